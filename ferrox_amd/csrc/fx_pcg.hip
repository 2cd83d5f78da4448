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
 * An iteration is three passes over the level besides the preconditioner:
 * the updates of phi and r with the max-norm of r; its publication; the new
 * search direction p and q = A p formed together from z and A z.  r.z comes
 * from the preconditioner's last transform (fft_solve_dot).  Every
 * reduction leaves partials, which the next kernel sums itself (a few
 * hundred loads per workgroup from L2), so the scalars alpha and beta never
 * travel to the host; only max|r| does, for the stopping rule.
 */
#include "fx_kernels.h"

namespace {

using fxd::DBC;
using fxd::DLev;

constexpr int PB = fxk::PCG_PARTS;   /* workgroups of the reducing kernels = partials per reduction */
constexpr int PT = 256;
constexpr int AB = 16;               /* workgroups per plane of the plane-mean kernel */

__device__ __forceinline__ double partial_sum(const double* __restrict__ part, double* sh) {
    double v = 0.0;
    for (int q = threadIdx.x; q < PB; q += PT) v += part[q];
    return fxd::block_sum<PT>(v, sh);
}

/* the A-coefficient the row operator applies at cell c (mg_coef) */
__device__ __forceinline__ double acoef(const DLev& l, int i, int j, int k, int c) {
    if (l.read_a == 0) return 0.0;
    if (l.read_a == 2) return l.a[c];
    const int g = (i + 1) + l.msx * (j + 1) + l.msy * (k + 1);
    return ((l.mat[g] & 3) >= 2) ? l.a[c] : 0.0;
}

/* partial sums of the A-coefficient over plane k = blockIdx.x, AB workgroups
 * per plane: part[k AB + blockIdx.y]; lateral extents 2^l0 x 2^l1 */
__global__ __launch_bounds__(PT) void k_pcg_abar(DLev l, int l0, double* __restrict__ part) {
    __shared__ double sh[PT / 64 + 1];
    const int k = blockIdx.x, s = l.n0 * l.n1;
    double v = 0.0;
    for (int q = blockIdx.y * PT + threadIdx.x; q < s; q += AB * PT)
        v += acoef(l, q & (l.n0 - 1), q >> l0, k, s * k + q);
    v = fxd::block_sum<PT>(v, sh);
    if (threadIdx.x == 0) part[k * AB + blockIdx.y] = v;
}

/* abar[k] = the plane mean from the partials */
__global__ __launch_bounds__(64) void k_pcg_abar_fin(const double* __restrict__ part, int n2, double inv_s,
                                                    double* __restrict__ abar) {
    const int k = blockIdx.x * 64 + threadIdx.x;
    if (k >= n2) return;
    double v = 0.0;
    for (int b = 0; b < AB; ++b) v += part[k * AB + b];
    abar[k] = v * inv_s;
}

/* The search direction and its image in one pass (fx_kernels.h,
 * pcg_dirmat): A z by the row operator of a layered level periodic in x
 * and y (mg_coef's layered branch and mg_acc, homogeneous faces, the same
 * accumulation order), then p = z + beta p and q = A z + beta q, which is
 * A p without a second stencil pass.  Bricks of 64 x 4 columns x MZ planes
 * (the planes unrolled, so their loads are in flight together), strided
 * over the PB workgroups. */
constexpr int MZ = 4;
__global__ __launch_bounds__(PT) void k_pcg_dirmat(DLev l, DBC bc, const double* __restrict__ z, double* __restrict__ p,
                                                  double* __restrict__ q, const double* __restrict__ part_rz, int nrz,
                                                  double* __restrict__ sc, int it, double* __restrict__ part_pq) {
    __shared__ double sh[PT / 64 + 1];
    if (sc[2] != 0.0) return;
    double v = 0.0;
    for (int t = threadIdx.x; t < nrz; t += PT) v += part_rz[t];
    const double rz = fxd::block_sum<PT>(v, sh);
    const bool first = it < 0;
    const double beta = first ? 0.0 : rz / sc[it & 1];
    if (blockIdx.x == 0 && threadIdx.x == 0) sc[(it + 1) & 1] = rz;   /* the slot the next update reads */
    const int n0 = l.n0, n1 = l.n1, n2 = l.n2, s = n0 * n1;
    const int bx = (n0 + 63) / 64, by = (n1 + 3) / 4, bz = (n2 + MZ - 1) / MZ;
    const int nbrick = bx * by * bz;
    const bool zlo_per = bc.lo[2] == fxd::BC_PER, zhi_per = bc.hi[2] == fxd::BC_PER;
    const double dlo = bc.lo[2] == fxd::BC_DIR ? 2.0 : 0.0, dhi = bc.hi[2] == fxd::BC_DIR ? 2.0 : 0.0;
    double acc = 0.0;
    for (int t = blockIdx.x; t < nbrick; t += PB) {
        const int tx = t % bx, rest = t / bx, ty = rest % by, tz = rest / by;
        const int i = tx * 64 + (threadIdx.x & 63), j = ty * 4 + (threadIdx.x >> 6), k0 = tz * MZ;
        if (i < n0 && j < n1) {
            const int im = i > 0 ? -1 : n0 - 1, ip = i < n0 - 1 ? 1 : -(n0 - 1);
            const int jm = j > 0 ? -n0 : n0 * (n1 - 1), jp = j < n1 - 1 ? n0 : -n0 * (n1 - 1);
#pragma unroll
            for (int dz = 0; dz < MZ; ++dz) {
                const int k = k0 + dz;
                if (k < n2) {
                    const int c = i + n0 * (j + n1 * k);
                    const double wx = l.bxk[k] * l.ih2[0], wy = l.byk[k] * l.ih2[1];
                    const double wzl = l.bzk[k] * l.ih2[2], wzh = l.bzk[k + 1] * l.ih2[2];
                    const double zc = z[c];
                    double D = -acoef(l, i, j, k, c), S = 0.0;
                    D += wx; S += wx * z[c + im];
                    D += wx; S += wx * z[c + ip];
                    D += wy; S += wy * z[c + jm];
                    D += wy; S += wy * z[c + jp];
                    if (k > 0 || zlo_per) { D += wzl; S += wzl * z[k > 0 ? c - s : c + s * (n2 - 1)]; }
                    else D += dlo * wzl;
                    if (k < n2 - 1 || zhi_per) { D += wzh; S += wzh * z[k < n2 - 1 ? c + s : c - s * (n2 - 1)]; }
                    else D += dhi * wzh;
                    const double az = D * zc - S;
                    const double pn = first ? zc : zc + beta * p[c];
                    const double qn = first ? az : az + beta * q[c];
                    p[c] = pn;
                    q[c] = qn;
                    acc += pn * qn;
                }
            }
        }
    }
    acc = fxd::block_sum<PT>(acc, sh);
    if (threadIdx.x == 0) part_pq[blockIdx.x] = acc;
}

/* alpha = rz / pq (rz = sc[it & 1]); phi += alpha p; r -= alpha q; partial max|r| */
__global__ __launch_bounds__(PT) void k_pcg_update(double* __restrict__ phi, double* __restrict__ r,
                                                  const double* __restrict__ p, const double* __restrict__ q, int N,
                                                  const double* __restrict__ part_pq, const double* __restrict__ sc,
                                                  int it, double* __restrict__ part_r) {
    __shared__ double sh[PT / 64 + 1];
    if (sc[2] != 0.0) return;
    const double alpha = sc[it & 1] / partial_sum(part_pq, sh);
    double m = 0.0;
    for (int c = blockIdx.x * PT + threadIdx.x; c < N; c += PB * PT) {
        phi[c] += alpha * p[c];
        const double rv = r[c] - alpha * q[c];
        r[c] = rv;
        m = fxd::nmax(m, fabs(rv));
    }
    m = fxd::block_max<PT>(m, sh);
    if (threadIdx.x == 0) part_r[blockIdx.x] = m;
}

/* the max of the PB partials to pinned host memory, then the sequence number
 * v (system-scope release), as k_publish; sc[2] = 1 once the max meets the
 * target sc[3] (every later pass of the solve then does nothing) */
__global__ __launch_bounds__(PT) void k_pcg_publish_max(const double* __restrict__ part, double* host, unsigned* seq,
                                                       unsigned v, double* __restrict__ sc) {
    __shared__ double sh[PT / 64 + 1];
    double m = 0.0;
    for (int q = threadIdx.x; q < PB; q += PT) m = fxd::nmax(m, part[q]);
    m = fxd::block_max<PT>(m, sh);
    if (threadIdx.x == 0) {
        if (m <= sc[3]) sc[2] = 1.0;
        host[0] = m;
        __threadfence_system();
        __hip_atomic_store(seq, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

inline int log2i(int n) {
    int l = 0;
    while ((1 << l) < n) ++l;
    return l;
}

}  // namespace

namespace fxk {

void pcg_abar(hipStream_t s, const DLev& l, double* part, double* abar) {
    hipLaunchKernelGGL(k_pcg_abar, dim3(l.n2, AB), dim3(PT), 0, s, l, log2i(l.n0), part);
    hipLaunchKernelGGL(k_pcg_abar_fin, dim3((l.n2 + 63) / 64), dim3(64), 0, s, part, l.n2, 1.0 / ((double)l.n0 * l.n1),
                       abar);
}
void pcg_dirmat(hipStream_t s, const DLev& l, const DBC& bc, const double* z, double* p, double* q,
                const double* part_rz, int nrz, double* sc, int it, double* part_pq) {
    hipLaunchKernelGGL(k_pcg_dirmat, dim3(PB), dim3(PT), 0, s, l, bc, z, p, q, part_rz, nrz, sc, it, part_pq);
}
void pcg_update(hipStream_t s, double* phi, double* r, const double* p, const double* q, long long N,
                const double* part_pq, const double* sc, int it, double* part_r) {
    hipLaunchKernelGGL(k_pcg_update, dim3(PB), dim3(PT), 0, s, phi, r, p, q, (int)N, part_pq, sc, it, part_r);
}
void pcg_publish_max(hipStream_t s, const double* part, double* host, unsigned* seq, unsigned v, double* sc) {
    hipLaunchKernelGGL(k_pcg_publish_max, dim3(1), dim3(PT), 0, s, part, host, seq, v, sc);
}

}  // namespace fxk
