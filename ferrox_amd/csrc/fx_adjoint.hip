/*
 * Discrete adjoint of the forward-Euler TDGL step with its Poisson solve
 * (scalar polarization P = P_z, no rotation, no semiconductor): the kernels
 * of one backward step.  The forward step (main.cpp:279-348 with
 * ComputePhi_Rho, ElectrostaticSolver.cpp:678-747) is
 *
 *   phi   = A^-1 (B P + bvec V)                 B P = -DPDz(P) on FE cells
 *   GL    = chi [ -G (alpha P + beta P^3 + gamma P^5)
 *                 -G (-g11 Hzz P - g44' (Hxx + Hyy) P)  - G fz(phi, V) ]
 *   P'    = P + dt GL
 *
 * with chi the ferroelectric (non t-phase) cells, G = BigGamma, H the masked
 * second-difference stencils (DoubleDPDx/y/z) and fz the z-derivative of phi
 * used for E (E_z = -fz).  With nu = dt G chi lambda' the adjoint step is
 *
 *   psi    = A^-1 (Fz^T nu)                      (A symmetric: same multigrid)
 *   lambda = lambda' - (alpha + 3 beta P^2 + 5 gamma P^4) nu
 *            + g11 Hzz^T nu + g44' (Hxx + Hyy)^T nu - B^T psi   (+ loss term)
 *
 * and the parameter gradients are the reductions listed at k_adj_step.  Every
 * transpose is exact: the coefficient that cell n's stencil applies to cell
 * c's value is obtained by evaluating the forward stencil at n on a unit
 * input at c, so the adjoint follows every boundary-flag branch of the
 * forward stencils without a hand-transposed copy of them.
 */
#include "fx_device.h"
#include "fx_kernels.h"

using namespace fxd;

namespace {

/* 1 at cell (ci, cj, ck), with P's ghost semantics (periodic image, 0 across
 * a non-periodic face). */
struct UnitPView {
    int ci, cj, ck;
    int nx, ny, nz, px, py, pz;
    __device__ __forceinline__ double operator()(int i, int j, int k) const {
        if (i < 0) { if (!px) return 0.0; do i += nx; while (i < 0); }
        else if (i >= nx) { if (!px) return 0.0; do i -= nx; while (i >= nx); }
        if (j < 0) { if (!py) return 0.0; do j += ny; while (j < 0); }
        else if (j >= ny) { if (!py) return 0.0; do j -= ny; while (j >= ny); }
        if (k < 0) { if (!pz) return 0.0; do k += nz; while (k < 0); }
        else if (k >= nz) { if (!pz) return 0.0; do k -= nz; while (k >= nz); }
        return (i == ci && j == cj && k == ck) ? 1.0 : 0.0;
    }
};

/* phi = unit at (ci, cj, ck) (unit != 0) or 0 everywhere, with the Dirichlet
 * values v and the ghost rule of PhiView. */
struct UnitPhiView {
    int ci, cj, ck, unit;
    int n[3];
    int lo[3], hi[3];
    double v[6];
    int ghost_face;
    __device__ __forceinline__ double pv(int i, int j, int k) const {
        return (unit && i == ci && j == cj && k == ck) ? 1.0 : 0.0;
    }
    __device__ __forceinline__ double at(int i, int j, int k) const {
        const int nx = n[0], ny = n[1], nz = n[2];
        if (i < 0 || i >= nx) {
            const bool l = i < 0; const int t = l ? lo[0] : hi[0];
            if (t == BC_PER) i = l ? i + nx : i - nx;
            else { const double p0 = pv(l ? 0 : nx - 1, j, k);
                   if (t == BC_NEU) return p0; const double pb = v[l ? 0 : 3];
                   return ghost_face ? pb : 2.0 * pb - p0; }
        }
        if (j < 0 || j >= ny) {
            const bool l = j < 0; const int t = l ? lo[1] : hi[1];
            if (t == BC_PER) j = l ? j + ny : j - ny;
            else { const double p0 = pv(i, l ? 0 : ny - 1, k);
                   if (t == BC_NEU) return p0; const double pb = v[l ? 1 : 4];
                   return ghost_face ? pb : 2.0 * pb - p0; }
        }
        if (k < 0 || k >= nz) {
            const bool l = k < 0; const int t = l ? lo[2] : hi[2];
            if (t == BC_PER) k = l ? k + nz : k - nz;
            else { const double p0 = pv(i, j, l ? 0 : nz - 1);
                   if (t == BC_NEU) return p0; const double pb = v[l ? 2 : 5];
                   return ghost_face ? pb : 2.0 * pb - p0; }
        }
        return pv(i, j, k);
    }
};

__device__ __forceinline__ bool is_fe(const MView& M, int i, int j, int k) {
    const int r = M.raw(i, j, k);
    return (r & 3) == 0 && !(r & 4);
}

/* neighbour (i,j,k) + o e_d inside the domain (periodic wrap) */
__device__ __forceinline__ bool shift(const Geom& g, int d, int o, int& i, int& j, int& k) {
    int q[3] = {i, j, k};
    q[d] += o;
    if (q[d] < 0 || q[d] >= g.n[d]) {
        if (!g.per[d]) return false;
        q[d] = (q[d] % g.n[d] + g.n[d]) % g.n[d];
    }
    i = q[0]; j = q[1]; k = q[2];
    return true;
}

/* offsets o in [-r, r] along d whose wrapped cells are distinct (a periodic
 * extent shorter than 2r+1 would reach one cell twice; each stencil row must
 * be visited once) */
__device__ __forceinline__ bool first_visit(const Geom& g, int d, int o, int base) {
    if (!g.per[d]) return true;
    const int n = g.n[d];
    const int w = ((base + o) % n + n) % n;
    for (int p = -2; p < o; ++p)
        if (((base + p) % n + n) % n == w) return false;
    return true;
}

__device__ __forceinline__ double fz_of(const UnitPhiView& U, const Geom& g, int i, int j, int k) {
    double fx, fy, fz;
    grad_phi(U, g, i, j, k, fx, fy, fz);
    return fz;
}

}  // namespace

/* s = Fz^T nu with nu = dtG chi lambda: s_c = sum over the cells n whose fz
 * reads phi_c of nu_n * (d fz_n / d phi_c). */
__global__ __launch_bounds__(256) void k_adj_fzT(fxk::AdjArgs a) {
    const Geom& g = a.g;
    const int i = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int k = blockIdx.z;
    if (i >= g.n[0] || j >= g.n[1]) return;
    const int nx = g.n[0], ny = g.n[1];
    UnitPhiView U;
    U.ci = i; U.cj = j; U.ck = k; U.unit = 1;
    for (int d = 0; d < 3; ++d) { U.n[d] = g.n[d]; U.lo[d] = a.bclo[d]; U.hi[d] = a.bchi[d]; }
    for (int q = 0; q < 6; ++q) U.v[q] = 0.0;
    U.ghost_face = a.ghost_face;
    double s = 0.0;
    for (int o = -1; o <= 1; ++o) {
        int ni = i, nj = j, nk = k;
        if (!shift(g, 2, o, ni, nj, nk)) continue;
        if (!first_visit(g, 2, o, k)) continue;
        if (!is_fe(a.M, ni, nj, nk)) continue;
        const int n = ni + nx * (nj + ny * nk);
        const double nu = a.dtG * a.lam[n];
        if (nu != 0.0) s += nu * fz_of(U, g, ni, nj, nk);
    }
    a.s[i + nx * (j + ny * k)] = s;
}

/* lambda_n from lambda_{n+1} and psi, with the per-workgroup partial sums of
 *   0 dJ/dalpha = -sum nu P        3 dJ/dg11  =  sum nu Hzz P
 *   1 dJ/dbeta  = -sum nu P^3      4 dJ/dg44  =  sum nu (Hxx + Hyy) P
 *   2 dJ/dgamma = -sum nu P^5      5 dJ/deps_fe = sum psi [(dA/deps) phi - (dbvec/deps) V]
 *   6 dJ/dV_n   = -sum nu dfz/dV - sum psi bvec_hi
 * (eps_fe: relative permittivity of the FE cells; eps_cell = eps_fe * eps_0). */
__global__ __launch_bounds__(256) void k_adj_step(fxk::AdjArgs a) {
    __shared__ double sh[256 / 64 + 1];
    const Geom& g = a.g;
    const int i = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int k = blockIdx.z;
    const int nx = g.n[0], ny = g.n[1];
    double part[7] = {0, 0, 0, 0, 0, 0, 0};
    if (i < g.n[0] && j < g.n[1]) {
        const int c = i + nx * (j + ny * k);
        const bool fe = is_fe(a.M, i, j, k);
        const double P = a.P[c];
        const double nu = fe ? a.dtG * a.lam[c] : 0.0;
        UnitPView U{i, j, k, g.n[0], g.n[1], g.n[2], g.per[0], g.per[1], g.per[2]};
        /* transposed gradient terms and B^T psi */
        double HT = 0.0, BT = 0.0;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const double gw = d == 2 ? a.g11 : a.g44m;
            const int base = d == 0 ? i : d == 1 ? j : k;
            for (int o = -2; o <= 2; ++o) {
                int ni = i, nj = j, nk = k;
                if (!shift(g, d, o, ni, nj, nk)) continue;
                if (!first_visit(g, d, o, base)) continue;
                if (!is_fe(a.M, ni, nj, nk)) continue;
                const int n = ni + nx * (nj + ny * nk);
                const double nun = a.dtG * a.lam[n];
                if (nun != 0.0) {
                    const double coef = d == 0 ? ddpd<0>(U, a.M, ni, nj, nk, g.dx, a.pbc)
                                      : d == 1 ? ddpd<1>(U, a.M, ni, nj, nk, g.dx, a.pbc)
                                               : ddpd<2>(U, a.M, ni, nj, nk, g.dx, a.pbc);
                    HT += gw * nun * coef;
                }
                if (d == 2 && o >= -1 && o <= 1) {
                    const double ps = a.psi[n];
                    if (ps != 0.0) BT += ps * (-dpd<2>(U, a.M, ni, nj, nk, g.dx, a.pbc));
                }
            }
        }
        double lam = a.lam[c] - (a.alpha + 3.0 * a.beta * P * P + 5.0 * a.gamma * P * P * P * P) * nu + HT - BT;
        if (fe) lam += a.loss_w;
        a.lam_out[c] = lam;

        if (fe && nu != 0.0) {
            const PView F{a.P, g.n[0], g.n[1], g.n[2], g.per[0], g.per[1], g.per[2]};
            const double P2 = P * P;
            part[0] = -nu * P;
            part[1] = -nu * P2 * P;
            part[2] = -nu * P2 * P2 * P;
            part[3] = nu * ddpd<2>(F, a.M, i, j, k, g.dx, a.pbc);
            part[4] = nu * (ddpd<0>(F, a.M, i, j, k, g.dx, a.pbc) + ddpd<1>(F, a.M, i, j, k, g.dx, a.pbc));
            /* direct dependence of fz on the top Dirichlet value */
            UnitPhiView Z;
            Z.ci = i; Z.cj = j; Z.ck = k; Z.unit = 0;
            for (int d = 0; d < 3; ++d) { Z.n[d] = g.n[d]; Z.lo[d] = a.bclo[d]; Z.hi[d] = a.bchi[d]; }
            for (int q = 0; q < 6; ++q) Z.v[q] = 0.0;
            Z.v[5] = 1.0;
            Z.ghost_face = a.ghost_face;
            part[6] = -nu * fz_of(Z, g, i, j, k);
        }
        /* permittivity and gate terms through the Poisson operator */
        const double ps = a.psi[c];
        if (ps != 0.0) {
            const double phic = a.phi[c];
            const double ce = fe ? a.eps0 : 0.0;      /* d eps_c / d eps_fe */
            double de = 0.0;
            for (int d = 0; d < 3; ++d)
                for (int side = 0; side < 2; ++side) {
                    const int o = side ? 1 : -1;
                    int ni = i, nj = j, nk = k;
                    const int idx[3] = {i, j, k};
                    const bool inside = side ? idx[d] < g.n[d] - 1 : idx[d] > 0;
                    const int t = side ? a.bchi[d] : a.bclo[d];
                    if (inside || t == BC_PER) {
                        shift(g, d, o, ni, nj, nk);
                        const double cn = is_fe(a.M, ni, nj, nk) ? a.eps0 : 0.0;
                        const double dw = 0.5 * (ce + cn) * a.ih2[d];
                        de += dw * (phic - a.phi[ni + nx * (nj + ny * nk)]);
                    } else if (t == BC_DIR) {
                        const double vb = a.bcv[side ? 3 + d : d];
                        de += 2.0 * ce * a.ih2[d] * (phic - vb);
                    }
                }
            part[5] = ps * de;
            if (k == g.n[2] - 1 && a.bchi[2] == BC_DIR) part[6] -= ps * 2.0 * a.eps[c] * a.ih2[2];
        }
    }
    const int nb = gridDim.x * gridDim.y * gridDim.z;
    const int bid = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
#pragma unroll
    for (int q = 0; q < 7; ++q) {
        const double v = block_sum<256>(part[q], sh);
        if (threadIdx.x == 0) a.part[q * nb + bid] = v;
    }
}

/* sum over the FE cells of P (the curve value is this over their count) */
__global__ __launch_bounds__(256) void k_fe_sum(const double* __restrict__ P, MView M, int nx, int ny, double* __restrict__ part) {
    __shared__ double sh[256 / 64 + 1];
    const int i = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int k = blockIdx.z;
    double v = 0.0;
    if (i < nx && j < ny && is_fe(M, i, j, k)) v = P[i + nx * (j + ny * k)];
    v = block_sum<256>(v, sh);
    if (threadIdx.x == 0) part[blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z)] = v;
}

/* lambda = w * chi (the loss term at the last state) */
__global__ void k_adj_init(double* __restrict__ lam, MView M, int nx, int ny, int nz, double w) {
    const long long N = (long long)nx * ny * nz;
    for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < N; c += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(c % nx), j = (int)((c / nx) % ny), k = (int)(c / ((long long)nx * ny));
        lam[c] = is_fe(M, i, j, k) ? w : 0.0;
    }
}

/* acc[0..5] += column sums of part (q = 0..5); gV[step] = column sum q = 6.
 * One workgroup; the sum order is fixed, so gradients are reproducible. */
__global__ __launch_bounds__(1024) void k_adj_accum(const double* __restrict__ part, int nb, double* __restrict__ acc,
                                                    double* __restrict__ gV, int step) {
    __shared__ double sh[1024 / 64 + 1];
    for (int q = 0; q < 7; ++q) {
        double v = 0.0;
        for (int b = threadIdx.x; b < nb; b += 1024) v += part[q * nb + b];
        v = block_sum<1024>(v, sh);
        if (threadIdx.x == 0) {
            if (q < 6) acc[q] += v;
            else gV[step] = v;
        }
    }
}

namespace fxk {
static inline dim3 grid_adj(const Geom& g) { return dim3((g.n[0] + 63) / 64, (g.n[1] + 3) / 4, g.n[2]); }
int adj_blocks(const Geom& g) { return ((g.n[0] + 63) / 64) * ((g.n[1] + 3) / 4) * g.n[2]; }
void adj_fzT(hipStream_t s, const AdjArgs& a) { hipLaunchKernelGGL(k_adj_fzT, grid_adj(a.g), dim3(256), 0, s, a); }
void adj_step(hipStream_t s, const AdjArgs& a) { hipLaunchKernelGGL(k_adj_step, grid_adj(a.g), dim3(256), 0, s, a); }
void adj_accum(hipStream_t s, const double* part, int nb, double* acc, double* gV, int step) {
    hipLaunchKernelGGL(k_adj_accum, dim3(1), dim3(1024), 0, s, part, nb, acc, gV, step);
}
void fe_sum(hipStream_t s, const double* P, const MView& M, const Geom& g, double* part, double* out) {
    hipLaunchKernelGGL(k_fe_sum, grid_adj(g), dim3(256), 0, s, P, M, g.n[0], g.n[1], part);
    final1024(s, part, adj_blocks(g), 1, 0, out);
}
void adj_init(hipStream_t s, double* lam, const MView& M, const Geom& g, double w) {
    const long long N = (long long)g.n[0] * g.n[1] * g.n[2];
    unsigned b = (unsigned)((N + 255) / 256); if (b > 4096) b = 4096; if (b == 0) b = 1;
    hipLaunchKernelGGL(k_adj_init, dim3(b), dim3(256), 0, s, lam, M, g.n[0], g.n[1], g.n[2], w);
}
}  // namespace fxk
