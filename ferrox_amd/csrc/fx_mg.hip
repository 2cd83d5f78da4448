/*
 * Geometric multigrid for  -a*phi - div(eps grad phi) = rhs  on gfx950: the
 * from-scratch replacement of AMReX MLABecLaplacian + MLMG
 * (ElectrostaticSolver.cpp:562-611, 713-720).
 *
 *   k_mg_gsrb_fused  a full red-black Gauss-Seidel sweep in one pass: each
 *                    workgroup owns a 64x4 column tile and a z-chunk, keeps four
 *                    z-planes of phi (with a 2-cell in-plane halo) in LDS and
 *                    streams upward, updating red on plane t and black on plane
 *                    t-1 -- the same values as a red sweep followed by a black
 *                    sweep, with phi, rhs and the material id read once and phi
 *                    written once.
 *   k_mg_smooth      one colour of the sweep (small levels, periodic z)
 *   k_mg_resid_restrict   residual averaged into the coarse rhs; coarse phi = 0
 *   k_mg_prolong     cell-centred linear prolongation added to the fine phi
 *   k_mg_resid_norm  max|r| and max|b| per workgroup
 *   k_bottom_cg      Jacobi-PCG on the coarsest level in one workgroup
 * Launches use 3-D grids (x tiles of 64, y tiles of 4, z) so no thread does
 * 64-bit index division.
 */
#include "fx_device.h"
#include "fx_kernels.h"

#include <algorithm>

using namespace fxd;

/* ------------------------- one colour of RB-GS ---------------------------- */
/* grid: x covers half-rows in chunks of 64 threads, y = j/4, z = k.
 * zero_init: the other colour is known to be zero (first red pass of a
 * correction solve), so phi = b/D without reading neighbours -- the value the
 * plain update gives when every neighbour is 0. */
template <class T>
__global__ __launch_bounds__(256) void k_mg_smooth(DLevT<T> l, DBC bc, int inh, int color, int zero_init) {
    const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int k = blockIdx.z;
    const int i = 2 * (blockIdx.x * 64 + (threadIdx.x & 63)) + ((j + k + color) & 1);
    if (i >= l.n0 || j >= l.n1) return;
    const int c = i + l.n0 * (j + l.n1 * k);
    T D, S;
    if (zero_init) {
        T a, w[6];
        const T nb[6] = {T(0), T(0), T(0), T(0), T(0), T(0)};
        mg_coef(l, bc, i, j, k, c, a, w);
        mg_acc(l, bc, inh != 0, i, j, k, a, w, nb, D, S);
    } else {
        mg_row(l, l.phi, i, j, k, bc, inh != 0, D, S);
    }
    l.phi[c] = (l.b[c] + S) / D;
}

/* ---------------------- fused red + black sweep --------------------------- */
namespace {
constexpr int FTX = 64, FTY = 4;              /* interior tile */
constexpr int HX = FTX + 4, HY = FTY + 4;     /* with the 2-cell halo */
constexpr int PL = HX * HY;                   /* cells per LDS plane */
}

/* Requires: z not periodic; n0 >= 2, n1 >= 2 in periodic directions (even,
 * as every smoothed level is). */
__global__ __launch_bounds__(256) void k_mg_gsrb_fused(DLev l, DBC bc, int inh, int kchunk) {
    __shared__ double sp[4][PL];
    const int n0 = l.n0, n1 = l.n1, n2 = l.n2;
    const int tx0 = blockIdx.x * FTX, ty0 = blockIdx.y * FTY;
    const int k0 = blockIdx.z * kchunk;
    const int k1 = min(k0 + kchunk, n2);
    const int tid = threadIdx.x;
    const bool perx = bc.lo[0] == BC_PER, pery = bc.lo[1] == BC_PER;
    const int s = n0 * n1;

    /* global (i, j) of an LDS position (hx, hy), or -1 if outside a
     * non-periodic face */
    auto gi_of = [&](int hx) -> int {
        int i = tx0 + hx - 2;
        if (i < 0) return perx ? i + n0 : -1;
        if (i >= n0) return perx ? i - n0 : -1;
        return i;
    };
    auto gj_of = [&](int hy) -> int {
        int j = ty0 + hy - 2;
        if (j < 0) return pery ? j + n1 : -1;
        if (j >= n1) return pery ? j - n1 : -1;
        return j;
    };
    auto load_plane = [&](int k) {
        double* dst = sp[(k + 8) & 3];
        for (int p = tid; p < PL; p += 256) {
            const int hx = p % HX, hy = p / HX;
            const int i = gi_of(hx), j = gj_of(hy);
            dst[p] = (i >= 0 && j >= 0) ? l.phi[i + n0 * (j + n1 * k)] : 0.0;
        }
    };
    /* update one cell from LDS neighbours; (hx, hy) in LDS coordinates */
    auto update = [&](int hx, int hy, int k) {
        const int i = tx0 + hx - 2, j = ty0 + hy - 2;      /* unwrapped */
        int iw = i, jw = j;
        if (iw < 0) iw += n0; else if (iw >= n0) iw -= n0;
        if (jw < 0) jw += n1; else if (jw >= n1) jw -= n1;
        const int c = iw + n0 * (jw + n1 * k);
        double a, w[6], nb[6];
        mg_coef(l, bc, iw, jw, k, c, a, w);
        const int p = hx + HX * hy;
        const double* cur = sp[(k + 8) & 3];
        nb[0] = cur[p - 1];
        nb[1] = cur[p + 1];
        nb[2] = cur[p - HX];
        nb[3] = cur[p + HX];
        nb[4] = (k > 0) ? sp[(k + 7) & 3][p] : 0.0;
        nb[5] = (k < n2 - 1) ? sp[(k + 9) & 3][p] : 0.0;
        double D, S;
        mg_acc(l, bc, inh != 0, iw, jw, k, a, w, nb, D, S);
        sp[(k + 8) & 3][p] = (l.b[c] + S) / D;
    };

    const int kr0 = max(k0 - 1, 0);
    for (int k = max(kr0 - 1, 0); k <= min(kr0 + 1, n2 - 1); ++k) load_plane(k);
    __syncthreads();
    for (int t = kr0; t <= min(k1, n2 - 1) + 1; ++t) {
        if (t <= min(k1, n2 - 1)) {
            if (t + 1 < n2 && t + 1 > kr0 + 1) load_plane(t + 1);
            __syncthreads();
            /* red on plane t over the tile and a 1-cell ring */
            for (int p = tid; p < (FTX + 2) * (FTY + 2); p += 256) {
                const int hx = p % (FTX + 2) + 1, hy = p / (FTX + 2) + 1;
                const int i = tx0 + hx - 2, j = ty0 + hy - 2;
                if (((i + j + t) & 1) != 0) continue;
                if ((!perx && (i < 0 || i >= n0)) || (!pery && (j < 0 || j >= n1))) continue;
                if (i >= n0 + 1 || j >= n1 + 1) continue;
                update(hx, hy, t);
            }
            __syncthreads();
        }
        /* black on plane t-1 over the tile interior, then write it back */
        const int kb = t - 1;
        if (kb >= k0 && kb < k1) {
            {
                const int hx = (tid & 63) + 2, hy = (tid >> 6) + 2;
                const int i = tx0 + hx - 2, j = ty0 + hy - 2;
                if (i < n0 && j < n1 && ((i + j + kb) & 1) == 1) update(hx, hy, kb);
            }
            __syncthreads();
            {
                const int hx = (tid & 63) + 2, hy = (tid >> 6) + 2;
                const int i = tx0 + hx - 2, j = ty0 + hy - 2;
                if (i < n0 && j < n1) l.phi[i + n0 * (j + n1 * kb)] = sp[(kb + 8) & 3][hx + HX * hy];
            }
        }
        __syncthreads();
    }
    (void)s;
}

/* --------------------- residual + restriction ----------------------------- */
/* grid over coarse cells: x tiles of 64, y tiles of 4, z. */
template <class T>
__global__ __launch_bounds__(256) void k_mg_resid_restrict(DLevT<T> f, DLevT<T> cl, DBC bc, int inh) {
    const int I = blockIdx.x * 64 + (threadIdx.x & 63);
    const int J = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int K = blockIdx.z;
    if (I >= cl.n0 || J >= cl.n1) return;
    const int cx = f.crs[0], cy = f.crs[1], cz = f.crs[2];
    T s = T(0);
    for (int kk = 0; kk <= cz; ++kk)
        for (int jj = 0; jj <= cy; ++jj)
            for (int ii = 0; ii <= cx; ++ii) {
                const int i = cx ? 2 * I + ii : I, j = cy ? 2 * J + jj : J, k = cz ? 2 * K + kk : K;
                T D, S;
                mg_row(f, f.phi, i, j, k, bc, inh != 0, D, S);
                const int c = i + f.n0 * (j + f.n1 * k);
                s += f.b[c] - (D * f.phi[c] - S);
            }
    const int cnt = (1 + cx) * (1 + cy) * (1 + cz);
    const int t = I + cl.n0 * (J + cl.n1 * K);
    cl.b[t] = s / T(cnt);
    cl.phi[t] = T(0);
}

/* ------------------------------ prolongation ------------------------------ */
/* In each coarsened direction the fine cell takes 3/4 of its parent and 1/4 of
 * the parent's neighbour on its side; across a domain face that neighbour is
 * the homogeneous ghost (-parent Dirichlet, +parent Neumann, periodic image).
 * Sum order: z outer, x inner, parent first. */
template <class T>
__global__ __launch_bounds__(256) void k_mg_prolong(DLevT<T> f, DLevT<T> cl, DBC bc) {
    const int i = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int k = blockIdx.z;
    if (i >= f.n0 || j >= f.n1) return;
    const int fi[3] = {i, j, k};
    const int cn[3] = {cl.n0, cl.n1, cl.n2};
    int id[3][2];
    T w[3][2];
    int cnt[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        if (!f.crs[d]) { id[d][0] = fi[d]; id[d][1] = fi[d]; w[d][0] = T(1); w[d][1] = T(0); cnt[d] = 1; continue; }
        const int I = fi[d] >> 1;
        int J = (fi[d] & 1) ? I + 1 : I - 1;
        T sg = T(1);
        if (J < 0 || J >= cn[d]) {
            const int bct = J < 0 ? bc.lo[d] : bc.hi[d];
            if (bct == BC_PER) J = J < 0 ? J + cn[d] : J - cn[d];
            else { sg = (bct == BC_DIR) ? T(-1) : T(1); J = I; }
        }
        id[d][0] = I; w[d][0] = T(0.75);
        id[d][1] = J; w[d][1] = T(0.25) * sg;
        cnt[d] = 2;
    }
    T e = T(0);
    for (int a = 0; a < cnt[2]; ++a)
        for (int b = 0; b < cnt[1]; ++b)
            for (int q = 0; q < cnt[0]; ++q)
                e += (w[2][a] * w[1][b]) * w[0][q] * cl.phi[id[0][q] + cl.n0 * (id[1][b] + cl.n1 * id[2][a])];
    f.phi[i + f.n0 * (j + f.n1 * k)] += e;
}

/* Cell average of a fine array into the coarse level (A-coefficient). */
template <class T>
__global__ __launch_bounds__(256) void k_mg_restrict_cells(DLevT<T> f, DLevT<T> cl, const T* __restrict__ src, T* __restrict__ dst) {
    const int I = blockIdx.x * 64 + (threadIdx.x & 63);
    const int J = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int K = blockIdx.z;
    if (I >= cl.n0 || J >= cl.n1) return;
    const int cx = f.crs[0], cy = f.crs[1], cz = f.crs[2];
    T s = T(0);
    for (int kk = 0; kk <= cz; ++kk)
        for (int jj = 0; jj <= cy; ++jj)
            for (int ii = 0; ii <= cx; ++ii)
                s += src[(cx ? 2 * I + ii : I) + f.n0 * ((cy ? 2 * J + jj : J) + f.n1 * (cz ? 2 * K + kk : K))];
    dst[I + cl.n0 * (J + cl.n1 * K)] = s / T((1 + cx) * (1 + cy) * (1 + cz));
}

/* Residual (optionally stored) and per-workgroup max|r| and max|b|; each
 * workgroup walks k with stride gridDim.z.  Partials: [max|r| of every
 * block][max|b| of every block]. */
__global__ __launch_bounds__(256) void k_mg_resid_norm(DLev l, DBC bc, int inh, double* __restrict__ r, double* __restrict__ part) {
    __shared__ double sh[256 / 64 + 1];
    const int i = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
    double mr = 0.0, mb = 0.0;
    if (i < l.n0 && j < l.n1) {
        for (int k = blockIdx.z; k < l.n2; k += gridDim.z) {
            double D, S;
            mg_row(l, l.phi, i, j, k, bc, inh != 0, D, S);
            const int c = i + l.n0 * (j + l.n1 * k);
            const double rr = l.b[c] - (D * l.phi[c] - S);
            if (r) r[c] = rr;
            mr = nmax(mr, fabs(rr));
            mb = nmax(mb, fabs(l.b[c]));
        }
    }
    mr = block_max<256>(mr, sh);
    mb = block_max<256>(mb, sh);
    const int nb = gridDim.x * gridDim.y * gridDim.z;
    const int bid = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
    if (threadIdx.x == 0) { part[bid] = mr; part[nb + bid] = mb; }
}

/* out[q] = max (or sum) over b of part[q*nb + b], one workgroup of 1024. */
__global__ __launch_bounds__(1024) void k_final1024(const double* __restrict__ part, int nb, int nq, int is_max, double* __restrict__ out) {
    __shared__ double sh[1024 / 64 + 1];
    for (int q = 0; q < nq; ++q) {
        double v = 0.0;
        for (int b = threadIdx.x; b < nb; b += 1024) v = is_max ? nmax(v, part[q * nb + b]) : v + part[q * nb + b];
        v = is_max ? block_max<1024>(v, sh) : block_sum<1024>(v, sh);
        if (threadIdx.x == 0) out[q] = v;
    }
}

/* Jacobi-preconditioned CG for A e = rhs (homogeneous BCs) in one workgroup.
 * work holds 5*N doubles (r, z, p, q, diag). */
template <class T>
__global__ __launch_bounds__(1024) void k_bottom_cg(DLevT<T> l, DBC bc, const T* __restrict__ rhs, T* __restrict__ e,
                                                    T* __restrict__ work, double tol, int maxit, int* __restrict__ iters) {
    constexpr int BS = 1024;
    __shared__ T sh[BS / 64 + 1];
    const int N = l.n0 * l.n1 * l.n2;
    T* r = work;
    T* z = work + N;
    T* p = work + 2 * N;
    T* q = work + 3 * N;
    T* Dg = work + 4 * N;
    const int tid = threadIdx.x;
    T rr = T(0);
    for (int c = tid; c < N; c += BS) {
        const int i = c % l.n0, j = (c / l.n0) % l.n1, k = c / (l.n0 * l.n1);
        T a, w[6], nb[6] = {T(0), T(0), T(0), T(0), T(0), T(0)}, D, S;
        mg_coef(l, bc, i, j, k, c, a, w);
        mg_acc(l, bc, false, i, j, k, a, w, nb, D, S);
        Dg[c] = D;
        r[c] = rhs[c];
        e[c] = T(0);
        rr += rhs[c] * rhs[c];
    }
    const T rr0 = block_sum<BS>(rr, sh);
    if (rr0 == T(0)) return;
    T rz = T(0);
    for (int c = tid; c < N; c += BS) { z[c] = r[c] / Dg[c]; p[c] = z[c]; rz += r[c] * z[c]; }
    rz = block_sum<BS>(rz, sh);
    const T tol2 = T(tol * tol) * rr0;
    int it = 0;
    for (; it < maxit; ++it) {
        T pq = T(0);
        for (int c = tid; c < N; c += BS) {
            const int i = c % l.n0, j = (c / l.n0) % l.n1, k = c / (l.n0 * l.n1);
            T D, S;
            mg_row(l, p, i, j, k, bc, false, D, S);
            const T qq = D * p[c] - S;
            q[c] = qq;
            pq += p[c] * qq;
        }
        pq = block_sum<BS>(pq, sh);
        const T al = rz / pq;
        rr = T(0);
        for (int c = tid; c < N; c += BS) {
            e[c] += al * p[c];
            r[c] -= al * q[c];
            rr += r[c] * r[c];
        }
        rr = block_sum<BS>(rr, sh);
        if (rr <= tol2) { ++it; break; }
        T rzn = T(0);
        for (int c = tid; c < N; c += BS) { z[c] = r[c] / Dg[c]; rzn += r[c] * z[c]; }
        rzn = block_sum<BS>(rzn, sh);
        const T be = rzn / rz;
        rz = rzn;
        for (int c = tid; c < N; c += BS) p[c] = z[c] + be * p[c];
        __syncthreads();
    }
    if (tid == 0) *iters += it;
}

/* --------------------- FP64 refinement of an FP32 correction -------------- */
/* phi_new = phi + e (e may be null), r = b - A(phi_new) in FP64 with the
 * inhomogeneous BCs, stored as the FP32 right-hand side of the next
 * correction solve; per-workgroup max|r| and max|b|.  phi_new is a separate
 * buffer: neighbours are read as phi + e from the old arrays. */
__global__ __launch_bounds__(256) void k_refine_resid(DLev l, DBC bc, const float* __restrict__ e, double* __restrict__ phi_new,
                                                      float* __restrict__ r32, double* __restrict__ part) {
    __shared__ double sh[256 / 64 + 1];
    const int i = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
    double mr = 0.0, mb = 0.0;
    if (i < l.n0 && j < l.n1) {
        const int n0 = l.n0, n1 = l.n1, n2 = l.n2, s = n0 * n1;
        for (int k = blockIdx.z; k < n2; k += gridDim.z) {
            const int c = i + n0 * (j + n1 * k);
            const int nbi[6] = {i > 0 ? c - 1 : c + (n0 - 1), i < n0 - 1 ? c + 1 : c - (n0 - 1),
                                j > 0 ? c - n0 : c + n0 * (n1 - 1), j < n1 - 1 ? c + n0 : c - n0 * (n1 - 1),
                                k > 0 ? c - s : c + s * (n2 - 1), k < n2 - 1 ? c + s : c - s * (n2 - 1)};
            double nb[6];
#pragma unroll
            for (int q = 0; q < 6; ++q) nb[q] = e ? l.phi[nbi[q]] + (double)e[nbi[q]] : l.phi[nbi[q]];
            const double pc = e ? l.phi[c] + (double)e[c] : l.phi[c];
            double a, w[6], D, S;
            mg_coef(l, bc, i, j, k, c, a, w);
            mg_acc(l, bc, true, i, j, k, a, w, nb, D, S);
            const double rr = l.b[c] - (D * pc - S);
            phi_new[c] = pc;
            r32[c] = (float)rr;
            mr = nmax(mr, fabs(rr));
            mb = nmax(mb, fabs(l.b[c]));
        }
    }
    mr = block_max<256>(mr, sh);
    mb = block_max<256>(mb, sh);
    const int nb = gridDim.x * gridDim.y * gridDim.z;
    const int bid = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
    if (threadIdx.x == 0) { part[bid] = mr; part[nb + bid] = mb; }
}

__global__ void k_d2f(float* __restrict__ d, const double* __restrict__ s, long long n) {
    for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < n; c += (long long)gridDim.x * blockDim.x) d[c] = (float)s[c];
}

/* ------------------------------- launchers -------------------------------- */
namespace fxk {

static inline dim3 grid3(int n0, int n1, int n2) { return dim3((n0 + 63) / 64, (n1 + 3) / 4, n2); }

int resid_blocks(const DLev& l) { return ((l.n0 + 63) / 64) * ((l.n1 + 3) / 4) * l.n2; }

bool fused_ok(const DLev& l, const DBC& bc) {
    return bc.lo[2] != BC_PER && l.n0 >= 64 && l.n1 >= 4 && l.n2 >= 2;
}

void mg_sweep(hipStream_t s, const DLev& l, const DBC& bc, int inh, int nsweeps, int allow_fused) {
    if (nsweeps <= 0) return;
    if (allow_fused && fused_ok(l, bc)) {
        const int tiles = ((l.n0 + FTX - 1) / FTX) * ((l.n1 + FTY - 1) / FTY);
        int kc = l.n2;
        while (kc > 4 && tiles * ((l.n2 + kc - 1) / kc) < 2048) kc = (kc + 1) / 2;
        const dim3 g((l.n0 + FTX - 1) / FTX, (l.n1 + FTY - 1) / FTY, (l.n2 + kc - 1) / kc);
        for (int it = 0; it < nsweeps; ++it)
            hipLaunchKernelGGL(k_mg_gsrb_fused, g, dim3(256), 0, s, l, bc, inh, kc);
        return;
    }
    const dim3 g(((l.n0 + 1) / 2 + 63) / 64, (l.n1 + 3) / 4, l.n2);
    for (int it = 0; it < nsweeps; ++it) {
        hipLaunchKernelGGL(k_mg_smooth<double>, g, dim3(256), 0, s, l, bc, inh, 0, 0);
        hipLaunchKernelGGL(k_mg_smooth<double>, g, dim3(256), 0, s, l, bc, inh, 1, 0);
    }
}
void mg_sweep(hipStream_t s, const DLevF& l, const DBC& bc, int nsweeps, int zero_first) {
    const dim3 g(((l.n0 + 1) / 2 + 63) / 64, (l.n1 + 3) / 4, l.n2);
    for (int it = 0; it < nsweeps; ++it) {
        hipLaunchKernelGGL(k_mg_smooth<float>, g, dim3(256), 0, s, l, bc, 0, 0, (zero_first && it == 0) ? 1 : 0);
        hipLaunchKernelGGL(k_mg_smooth<float>, g, dim3(256), 0, s, l, bc, 0, 1, 0);
    }
}
void mg_smooth(hipStream_t s, const DLev& l, const DBC& bc, int inh, int color) {
    const dim3 g(((l.n0 + 1) / 2 + 63) / 64, (l.n1 + 3) / 4, l.n2);
    hipLaunchKernelGGL(k_mg_smooth<double>, g, dim3(256), 0, s, l, bc, inh, color, 0);
}
void mg_resid_restrict(hipStream_t s, const DLev& f, const DLev& c, const DBC& bc, int inh) {
    hipLaunchKernelGGL(k_mg_resid_restrict<double>, grid3(c.n0, c.n1, c.n2), dim3(256), 0, s, f, c, bc, inh);
}
void mg_resid_restrict(hipStream_t s, const DLevF& f, const DLevF& c, const DBC& bc) {
    hipLaunchKernelGGL(k_mg_resid_restrict<float>, grid3(c.n0, c.n1, c.n2), dim3(256), 0, s, f, c, bc, 0);
}
void mg_prolong(hipStream_t s, const DLev& f, const DLev& c, const DBC& bc) {
    hipLaunchKernelGGL(k_mg_prolong<double>, grid3(f.n0, f.n1, f.n2), dim3(256), 0, s, f, c, bc);
}
void mg_prolong(hipStream_t s, const DLevF& f, const DLevF& c, const DBC& bc) {
    hipLaunchKernelGGL(k_mg_prolong<float>, grid3(f.n0, f.n1, f.n2), dim3(256), 0, s, f, c, bc);
}
void mg_restrict_cells(hipStream_t s, const DLev& f, const DLev& c, const double* src, double* dst) {
    hipLaunchKernelGGL(k_mg_restrict_cells<double>, grid3(c.n0, c.n1, c.n2), dim3(256), 0, s, f, c, src, dst);
}
void mg_restrict_cells(hipStream_t s, const DLevF& f, const DLevF& c, const float* src, float* dst) {
    hipLaunchKernelGGL(k_mg_restrict_cells<float>, grid3(c.n0, c.n1, c.n2), dim3(256), 0, s, f, c, src, dst);
}
int resid_norm_blocks(const DLev& l) { return ((l.n0 + 63) / 64) * ((l.n1 + 3) / 4) * std::min(l.n2, 8); }

void mg_resid_norm(hipStream_t s, const DLev& l, const DBC& bc, int inh, double* r, double* part, double* out) {
    hipLaunchKernelGGL(k_mg_resid_norm, dim3((l.n0 + 63) / 64, (l.n1 + 3) / 4, std::min(l.n2, 8)), dim3(256), 0, s,
                       l, bc, inh, r, part);
    hipLaunchKernelGGL(k_final1024, dim3(1), dim3(1024), 0, s, part, resid_norm_blocks(l), 2, 1, out);
}
void refine_resid(hipStream_t s, const DLev& l, const DBC& bc, const float* e, double* phi_new, float* r32, double* part, double* out) {
    hipLaunchKernelGGL(k_refine_resid, dim3((l.n0 + 63) / 64, (l.n1 + 3) / 4, std::min(l.n2, 8)), dim3(256), 0, s,
                       l, bc, e, phi_new, r32, part);
    hipLaunchKernelGGL(k_final1024, dim3(1), dim3(1024), 0, s, part, resid_norm_blocks(l), 2, 1, out);
}
void d2f(hipStream_t s, float* d, const double* src, long long n) {
    unsigned b = (unsigned)((n + 255) / 256); if (b > 4096) b = 4096; if (b == 0) b = 1;
    hipLaunchKernelGGL(k_d2f, dim3(b), dim3(256), 0, s, d, src, n);
}
void final1024(hipStream_t s, const double* part, int nb, int nq, int is_max, double* out) {
    hipLaunchKernelGGL(k_final1024, dim3(1), dim3(1024), 0, s, part, nb, nq, is_max, out);
}
void bottom_cg(hipStream_t s, const DLev& l, const DBC& bc, const double* rhs, double* e, double* work, double tol, int maxit, int* iters) {
    hipLaunchKernelGGL(k_bottom_cg<double>, dim3(1), dim3(1024), 0, s, l, bc, rhs, e, work, tol, maxit, iters);
}
void bottom_cg(hipStream_t s, const DLevF& l, const DBC& bc, const float* rhs, float* e, float* work, double tol, int maxit, int* iters) {
    hipLaunchKernelGGL(k_bottom_cg<float>, dim3(1), dim3(1024), 0, s, l, bc, rhs, e, work, tol, maxit, iters);
}

}  // namespace fxk
