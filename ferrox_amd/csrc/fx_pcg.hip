/*
 * Preconditioned conjugate gradients for the Poisson operator with a
 * Helmholtz term on a laterally periodic, z-layered stack: the solve the
 * semiconductor outer loop (ComputePhi_Rho, ElectrostaticSolver.cpp:678-747)
 * hands to MLMG with the Newton coefficient of ComputePoissonRHS_Newton
 * (:85-132) as the A-coefficient.  That term varies across a plane (it
 * follows the carrier densities), so the direct solve of fx_fft.hip does
 * not apply to the operator itself; it does apply to the operator with each
 * plane's term replaced by its plane mean, which is symmetric positive
 * definite like the operator (the Newton coefficient is <= 0, the row
 * operator applies -a) and differs from it only by the lateral variation of
 * the carriers.  That layered operator, solved directly (its z-pivots
 * tabulated per solve), is the preconditioner.
 *
 * The kernels below are one pass each over the level: A p with the
 * multigrid's row operator (homogeneous faces) and the partial sums of p.q;
 * the updates of phi and r with the max-norm of r; r.z; the new search
 * direction.  Every reduction leaves PB partials, which the next kernel sums
 * itself (PB loads per workgroup from L2), so the scalars alpha and beta
 * never travel to the host; only max|r| does, for the stopping rule.
 */
#include "fx_kernels.h"

namespace {

using fxd::DBC;
using fxd::DLev;

constexpr int PB = fxk::PCG_PARTS;   /* workgroups of the reducing kernels = partials per reduction */
constexpr int PT = 256;

__device__ __forceinline__ double partial_sum(const double* __restrict__ part, double* sh) {
    double v = 0.0;
    for (int q = threadIdx.x; q < PB; q += PT) v += part[q];
    return fxd::block_sum<PT>(v, sh);
}

/* the A-coefficient the row operator applies at cell c (mg_coef) */
__device__ __forceinline__ double acoef(const DLev& l, int i, int j, int k, long long c) {
    if (l.read_a == 0) return 0.0;
    if (l.read_a == 2) return l.a[c];
    const int g = (i + 1) + l.msx * (j + 1) + l.msy * (k + 1);
    return ((l.mat[g] & 3) >= 2) ? l.a[c] : 0.0;
}

/* abar[k] = mean over plane k of the A-coefficient; one workgroup per plane */
__global__ __launch_bounds__(PT) void k_pcg_abar(DLev l, double* __restrict__ abar) {
    __shared__ double sh[PT / 64 + 1];
    const int k = blockIdx.x, s = l.n0 * l.n1;
    double v = 0.0;
    for (int q = threadIdx.x; q < s; q += PT) v += acoef(l, q % l.n0, q / l.n0, k, (long long)s * k + q);
    v = fxd::block_sum<PT>(v, sh);
    if (threadIdx.x == 0) abar[k] = v / (double)s;
}

/* q = A p (homogeneous faces) and the partial sums of p.q */
__global__ __launch_bounds__(PT) void k_pcg_matvec(DLev l, DBC bc, const double* __restrict__ p, double* __restrict__ q,
                                                  double* __restrict__ part) {
    __shared__ double sh[PT / 64 + 1];
    const int n0 = l.n0, n1 = l.n1;
    const long long N = (long long)n0 * n1 * l.n2;
    double acc = 0.0;
    for (long long c = blockIdx.x * (long long)PT + threadIdx.x; c < N; c += (long long)PB * PT) {
        const int i = (int)(c % n0);
        const long long rest = c / n0;
        const int j = (int)(rest % n1), k = (int)(rest / n1);
        double D, S;
        fxd::mg_row(l, p, i, j, k, bc, false, D, S);
        const double pc = p[c];
        const double v = D * pc - S;
        q[c] = v;
        acc += pc * v;
    }
    acc = fxd::block_sum<PT>(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

/* alpha = rz / pq (rz = sc[it & 1]); phi += alpha p; r -= alpha q; partial max|r| */
__global__ __launch_bounds__(PT) void k_pcg_update(double* __restrict__ phi, double* __restrict__ r,
                                                  const double* __restrict__ p, const double* __restrict__ q, long long N,
                                                  const double* __restrict__ part_pq, const double* __restrict__ sc,
                                                  int it, double* __restrict__ part_r) {
    __shared__ double sh[PT / 64 + 1];
    const double alpha = sc[it & 1] / partial_sum(part_pq, sh);
    double m = 0.0;
    for (long long c = blockIdx.x * (long long)PT + threadIdx.x; c < N; c += (long long)PB * PT) {
        phi[c] += alpha * p[c];
        const double rv = r[c] - alpha * q[c];
        r[c] = rv;
        m = fxd::nmax(m, fabs(rv));
    }
    m = fxd::block_max<PT>(m, sh);
    if (threadIdx.x == 0) part_r[blockIdx.x] = m;
}

/* partial sums of a.b */
__global__ __launch_bounds__(PT) void k_pcg_dot(const double* __restrict__ a, const double* __restrict__ b, long long N,
                                               double* __restrict__ part) {
    __shared__ double sh[PT / 64 + 1];
    double acc = 0.0;
    for (long long c = blockIdx.x * (long long)PT + threadIdx.x; c < N; c += (long long)PB * PT) acc += a[c] * b[c];
    acc = fxd::block_sum<PT>(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

/* rz_new = r.z from the partials; p = z + (rz_new / sc[it & 1]) p, or p = z
 * for it < 0; sc[(it + 1) & 1] = rz_new (the slot the next update reads;
 * no workgroup of this launch reads it) */
__global__ __launch_bounds__(PT) void k_pcg_dir(double* __restrict__ p, const double* __restrict__ z, long long N,
                                               const double* __restrict__ part_rz, double* __restrict__ sc, int it) {
    __shared__ double sh[PT / 64 + 1];
    const double rz = partial_sum(part_rz, sh);
    if (blockIdx.x == 0 && threadIdx.x == 0) sc[(it + 1) & 1] = rz;
    if (it < 0) {
        for (long long c = blockIdx.x * (long long)PT + threadIdx.x; c < N; c += (long long)PB * PT) p[c] = z[c];
    } else {
        const double beta = rz / sc[it & 1];
        for (long long c = blockIdx.x * (long long)PT + threadIdx.x; c < N; c += (long long)PB * PT)
            p[c] = z[c] + beta * p[c];
    }
}

}  // namespace

namespace fxk {

void pcg_abar(hipStream_t s, const DLev& l, double* abar) {
    hipLaunchKernelGGL(k_pcg_abar, dim3(l.n2), dim3(PT), 0, s, l, abar);
}
void pcg_matvec(hipStream_t s, const DLev& l, const DBC& bc, const double* p, double* q, double* part) {
    hipLaunchKernelGGL(k_pcg_matvec, dim3(PB), dim3(PT), 0, s, l, bc, p, q, part);
}
void pcg_update(hipStream_t s, double* phi, double* r, const double* p, const double* q, long long N,
                const double* part_pq, const double* sc, int it, double* part_r) {
    hipLaunchKernelGGL(k_pcg_update, dim3(PB), dim3(PT), 0, s, phi, r, p, q, N, part_pq, sc, it, part_r);
}
void pcg_dot(hipStream_t s, const double* a, const double* b, long long N, double* part) {
    hipLaunchKernelGGL(k_pcg_dot, dim3(PB), dim3(PT), 0, s, a, b, N, part);
}
void pcg_dir(hipStream_t s, double* p, const double* z, long long N, const double* part_rz, double* sc, int it) {
    hipLaunchKernelGGL(k_pcg_dir, dim3(PB), dim3(PT), 0, s, p, z, N, part_rz, sc, it);
}

}  // namespace fxk
