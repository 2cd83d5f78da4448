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
constexpr int FTX = 64, FTY = 4, FKZ = 8;          /* interior block: 64 x 4 columns, 8 planes */
constexpr int HX = FTX + 4, HY = FTY + 4, HZ = FKZ + 4;
}

/* One full red-black sweep in one pass, from phi_in to phi_out (two buffers,
 * so no workgroup can read a neighbour's already-updated values).  A
 * workgroup loads its 64x4x8 block of phi with a 2-cell halo in x, y and z
 * into LDS (every load in flight at once), updates red over the block and a
 * 1-cell ring (planes k0-1..k1) from the old black values, then black over
 * the interior from the new red values, and stores the interior.  Each cell
 * is computed from the same values as by a red launch followed by a black
 * launch; the ring's red values are recomputed, identically, by the
 * neighbouring block.  Requires z not periodic, n0 >= 64, n1 >= 8. */
template <class T>
__global__ __launch_bounds__(256) void k_mg_gsrb_fused(DLevT<T> l, DBC bc, int inh, int zero_init,
                                                       const T* __restrict__ phi_in, T* __restrict__ phi_out) {
    __shared__ T sp[HZ * HY * HX];
    const int n0 = l.n0, n1 = l.n1, n2 = l.n2;
    const int tx0 = blockIdx.x * FTX, ty0 = blockIdx.y * FTY, k0 = blockIdx.z * FKZ;
    const int k1 = min(k0 + FKZ, n2);
    const int tid = threadIdx.x;
    const bool perx = bc.lo[0] == BC_PER, pery = bc.lo[1] == BC_PER;
    auto wrapx = [&](int i) { return i < 0 ? i + n0 : (i >= n0 ? i - n0 : i); };
    auto wrapy = [&](int j) { return j < 0 ? j + n1 : (j >= n1 ? j - n1 : j); };
    /* 1. load phi (zero-initial: nothing to load, every value is 0) */
    for (int p = tid; p < HZ * HY * HX; p += 256) {
        const int hx = p % HX, hy = (p / HX) % HY, hz = p / (HX * HY);
        const int i = tx0 + hx - 2, j = ty0 + hy - 2, k = k0 + hz - 2;
        T v = T(0);
        if (!zero_init && k >= 0 && k < n2 && (perx || (i >= 0 && i < n0)) && (pery || (j >= 0 && j < n1)))
            v = phi_in[wrapx(i) + n0 * (wrapy(j) + n1 * k)];
        sp[p] = v;
    }
    __syncthreads();
    auto update = [&](int hx, int hy, int hz) {
        const int i = tx0 + hx - 2, j = ty0 + hy - 2, k = k0 + hz - 2;
        const int iw = wrapx(i), jw = wrapy(j);
        const int c = iw + n0 * (jw + n1 * k);
        const int p = hx + HX * (hy + HY * hz);
        T a, w[6], nb[6];
        mg_coef(l, bc, iw, jw, k, c, a, w);
        nb[0] = sp[p - 1];
        nb[1] = sp[p + 1];
        nb[2] = sp[p - HX];
        nb[3] = sp[p + HX];
        nb[4] = sp[p - HX * HY];
        nb[5] = sp[p + HX * HY];
        T D, S;
        mg_acc(l, bc, inh != 0, iw, jw, k, a, w, nb, D, S);
        sp[p] = (l.b[c] + S) / D;
    };
    /* 2. red over planes k0-1..k1 and the 1-cell ring: 33 red cells per ring row */
    {
        const int rows = (FTY + 2) * (FKZ + 2);
        for (int q = tid; q < rows * 33; q += 256) {
            const int r = q / 33, x2 = q % 33;
            const int hy = 1 + r % (FTY + 2), hz = 1 + r / (FTY + 2);
            const int j = ty0 + hy - 2, k = k0 + hz - 2;
            const int i0 = tx0 - 1;                           /* hx = 1 */
            const int hx = 1 + 2 * x2 + ((i0 + j + k) & 1);
            const int i = tx0 + hx - 2;
            if (hx > FTX + 2 || k < 0 || k >= n2) continue;
            if (!perx && (i < 0 || i >= n0)) continue;
            if (!pery && (j < 0 || j >= n1)) continue;
            if (i > n0 || j > n1) continue;
            update(hx, hy, hz);
        }
    }
    __syncthreads();
    /* 3. black over the interior planes k0..k1-1: 32 black cells per row */
    {
        const int rows = FTY * FKZ;
        for (int q = tid; q < rows * 32; q += 256) {
            const int r = q / 32, x2 = q % 32;
            const int hy = 2 + r % FTY, hz = 2 + r / FTY;
            const int j = ty0 + hy - 2, k = k0 + hz - 2;
            const int hx = 2 + 2 * x2 + ((tx0 + j + k + 1) & 1);
            const int i = tx0 + hx - 2;
            if (k >= k1 || i >= n0 || j >= n1) continue;
            update(hx, hy, hz);
        }
    }
    __syncthreads();
    /* 4. store the interior */
    for (int q = tid; q < FTX * FTY * FKZ; q += 256) {
        const int hx = 2 + q % FTX, hy = 2 + (q / FTX) % FTY, hz = 2 + q / (FTX * FTY);
        const int i = tx0 + hx - 2, j = ty0 + hy - 2, k = k0 + hz - 2;
        if (i < n0 && j < n1 && k < k1) phi_out[i + n0 * (j + n1 * k)] = sp[hx + HX * (hy + HY * hz)];
    }
}

/* --------------------- residual + restriction ----------------------------- */
/* grid over coarse cells: x tiles of 64, y tiles of 4, z.  The coarsening
 * pattern is a template parameter so the children's loads are unrolled. */
template <class T, int CX, int CY, int CZ>
__global__ __launch_bounds__(256) void k_mg_resid_restrict(DLevT<T> f, DLevT<T> cl, DBC bc, int inh) {
    const int I = blockIdx.x * 64 + (threadIdx.x & 63);
    const int J = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int K = blockIdx.z;
    if (I >= cl.n0 || J >= cl.n1) return;
    T s = T(0);
#pragma unroll
    for (int kk = 0; kk <= CZ; ++kk)
#pragma unroll
        for (int jj = 0; jj <= CY; ++jj)
#pragma unroll
            for (int ii = 0; ii <= CX; ++ii) {
                const int i = CX ? 2 * I + ii : I, j = CY ? 2 * J + jj : J, k = CZ ? 2 * K + kk : K;
                T D, S;
                mg_row(f, f.phi, i, j, k, bc, inh != 0, D, S);
                const int c = i + f.n0 * (j + f.n1 * k);
                s += f.b[c] - (D * f.phi[c] - S);
            }
    constexpr int cnt = (1 + CX) * (1 + CY) * (1 + CZ);
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

/* Residual (optionally stored) and per-workgroup max|r| and max|b|, one
 * cell per thread.  Partials: [max|r| of every block][max|b| of every block]. */
__global__ __launch_bounds__(256) void k_mg_resid_norm(DLev l, DBC bc, int inh, double* __restrict__ r, double* __restrict__ part) {
    __shared__ double sh[256 / 64 + 1];
    const int i = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
    double mr = 0.0, mb = 0.0;
    const int k = blockIdx.z;
    if (i < l.n0 && j < l.n1) {
        double D, S;
        mg_row(l, l.phi, i, j, k, bc, inh != 0, D, S);
        const int c = i + l.n0 * (j + l.n1 * k);
        const double rr = l.b[c] - (D * l.phi[c] - S);
        if (r) r[c] = rr;
        mr = fabs(rr);
        mb = fabs(l.b[c]);
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

/* ------------------- coarse tail of the V-cycle in one workgroup ----------- */
/* Levels with few cells are launch-latency bound: the whole sub-V-cycle below
 * a given level (smoothing, residual + restriction, Jacobi-PCG bottom,
 * prolongation) runs here in one 1024-thread workgroup, with the same per-cell
 * arithmetic as the separate kernels.  levs[0] is the first tail level (its b
 * already restricted, its phi zero); levs[nl-1] is the bottom. */
template <class T>
__global__ __launch_bounds__(1024) void k_mg_tail(const DLevT<T>* __restrict__ levs, int nl, DBC bc, int read_a,
                                                  int nu1, int nu2, T* __restrict__ work, double tol, int maxit,
                                                  int* __restrict__ iters) {
    constexpr int BS = 1024;
    __shared__ T sh[BS / 64 + 1];
    const int tid = threadIdx.x;
    auto lev = [&](int t) { DLevT<T> l = levs[t]; l.read_a = read_a ? 2 : 0; return l; };
    auto half_sweep = [&](const DLevT<T>& l, int color) {
        const int half = (l.n0 + 1) >> 1;
        const int tot = half * l.n1 * l.n2;
        for (int q = tid; q < tot; q += BS) {
            const int ih = q % half, jk = q / half;
            const int j = jk % l.n1, k = jk / l.n1;
            const int i = 2 * ih + ((j + k + color) & 1);
            if (i >= l.n0) continue;
            T D, S;
            mg_row(l, l.phi, i, j, k, bc, false, D, S);
            const int c = i + l.n0 * (j + l.n1 * k);
            l.phi[c] = (l.b[c] + S) / D;
        }
        __syncthreads();
    };
    auto restrict_to = [&](const DLevT<T>& f, const DLevT<T>& cl) {
        const int tot = cl.n0 * cl.n1 * cl.n2;
        const int cx = f.crs[0], cy = f.crs[1], cz = f.crs[2];
        for (int t = tid; t < tot; t += BS) {
            const int I = t % cl.n0, J = (t / cl.n0) % cl.n1, K = t / (cl.n0 * cl.n1);
            T s = T(0);
            for (int kk = 0; kk <= cz; ++kk)
                for (int jj = 0; jj <= cy; ++jj)
                    for (int ii = 0; ii <= cx; ++ii) {
                        const int i = cx ? 2 * I + ii : I, j = cy ? 2 * J + jj : J, k = cz ? 2 * K + kk : K;
                        T D, S;
                        mg_row(f, f.phi, i, j, k, bc, false, D, S);
                        const int c = i + f.n0 * (j + f.n1 * k);
                        s += f.b[c] - (D * f.phi[c] - S);
                    }
            cl.b[t] = s / T((1 + cx) * (1 + cy) * (1 + cz));
            cl.phi[t] = T(0);
        }
        __syncthreads();
    };
    auto prolong_from = [&](const DLevT<T>& f, const DLevT<T>& cl) {
        const int tot = f.n0 * f.n1 * f.n2;
        const int cn[3] = {cl.n0, cl.n1, cl.n2};
        for (int t = tid; t < tot; t += BS) {
            const int fi[3] = {t % f.n0, (t / f.n0) % f.n1, t / (f.n0 * f.n1)};
            int id[3][2];
            T w[3][2];
            int cnt[3];
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
            f.phi[t] += e;
        }
        __syncthreads();
    };
    for (int t = 0; t + 1 < nl; ++t) {
        const DLevT<T> l = lev(t);
        for (int s = 0; s < nu1; ++s) { half_sweep(l, 0); half_sweep(l, 1); }
        restrict_to(l, lev(t + 1));
    }
    {   /* Jacobi-PCG bottom, as k_bottom_cg */
        const DLevT<T> l = lev(nl - 1);
        const int N = l.n0 * l.n1 * l.n2;
        T* r = work; T* z = work + N; T* p = work + 2 * N; T* q = work + 3 * N; T* Dg = work + 4 * N;
        const T* rhs = l.b;
        T* e = l.phi;
        T rr = T(0);
        for (int c = tid; c < N; c += BS) {
            const int i = c % l.n0, j = (c / l.n0) % l.n1, k = c / (l.n0 * l.n1);
            T a, w[6], nb[6] = {T(0), T(0), T(0), T(0), T(0), T(0)}, D, S;
            mg_coef(l, bc, i, j, k, c, a, w);
            mg_acc(l, bc, false, i, j, k, a, w, nb, D, S);
            Dg[c] = D; r[c] = rhs[c]; e[c] = T(0);
            rr += rhs[c] * rhs[c];
        }
        const T rr0 = block_sum<BS>(rr, sh);
        if (rr0 != T(0)) {
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
                for (int c = tid; c < N; c += BS) { e[c] += al * p[c]; r[c] -= al * q[c]; rr += r[c] * r[c]; }
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
        __syncthreads();
    }
    for (int t = nl - 2; t >= 0; --t) {
        const DLevT<T> l = lev(t);
        prolong_from(l, lev(t + 1));
        for (int s = 0; s < nu2; ++s) { half_sweep(l, 0); half_sweep(l, 1); }
    }
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
    const int k = blockIdx.z;
    if (i < l.n0 && j < l.n1) {
        const int n0 = l.n0, n1 = l.n1, n2 = l.n2, s = n0 * n1;
        {
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

/* phi += e (FP32 correction added into the FP64 solution). */
__global__ void k_add_f2d(double* __restrict__ phi, const float* __restrict__ e, long long n) {
    for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < n; c += (long long)gridDim.x * blockDim.x)
        phi[c] += (double)e[c];
}

/* FP64 residual of the current solution, stored as the FP32 right-hand side
 * of the next correction solve, with per-workgroup max|r| and max|b|. */
__global__ __launch_bounds__(256) void k_resid_to32(DLev l, DBC bc, float* __restrict__ r32, double* __restrict__ part) {
    __shared__ double sh[256 / 64 + 1];
    const int i = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int k = blockIdx.z;
    double mr = 0.0, mb = 0.0;
    if (i < l.n0 && j < l.n1) {
        double D, S;
        mg_row(l, l.phi, i, j, k, bc, true, D, S);
        const int c = i + l.n0 * (j + l.n1 * k);
        const double rr = l.b[c] - (D * l.phi[c] - S);
        r32[c] = (float)rr;
        mr = fabs(rr);
        mb = fabs(l.b[c]);
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
    return bc.lo[2] != BC_PER && l.n0 >= 64 && l.n1 >= 8 && l.phi2 != nullptr;
}

template <class T>
__global__ void k_copyT(T* __restrict__ d, const T* __restrict__ s, long long n) {
    for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < n; c += (long long)gridDim.x * blockDim.x) d[c] = s[c];
}

template <class T>
static void sweep_impl(hipStream_t s, const DLevT<T>& l, const DBC& bc, int inh, int nsweeps, int allow_fused, int zero_first) {
    if (nsweeps <= 0) return;
    if (allow_fused && l.phi2 && bc.lo[2] != BC_PER && l.n0 >= 64 && l.n1 >= 8) {
        const dim3 g((l.n0 + FTX - 1) / FTX, (l.n1 + FTY - 1) / FTY, (l.n2 + FKZ - 1) / FKZ);
        T* cur = l.phi;
        T* nxt = l.phi2;
        for (int it = 0; it < nsweeps; ++it) {
            hipLaunchKernelGGL(k_mg_gsrb_fused<T>, g, dim3(256), 0, s, l, bc, inh, (zero_first && it == 0) ? 1 : 0, cur, nxt);
            std::swap(cur, nxt);
        }
        if (cur != l.phi) {   /* odd number of sweeps: the result goes back to phi */
            const long long n = (long long)l.n0 * l.n1 * l.n2;
            hipLaunchKernelGGL(k_copyT<T>, dim3((unsigned)std::min<long long>((n + 255) / 256, 4096)), dim3(256), 0, s, l.phi, cur, n);
        }
        return;
    }
    const dim3 g(((l.n0 + 1) / 2 + 63) / 64, (l.n1 + 3) / 4, l.n2);
    for (int it = 0; it < nsweeps; ++it) {
        hipLaunchKernelGGL(k_mg_smooth<T>, g, dim3(256), 0, s, l, bc, inh, 0, (zero_first && it == 0) ? 1 : 0);
        hipLaunchKernelGGL(k_mg_smooth<T>, g, dim3(256), 0, s, l, bc, inh, 1, 0);
    }
}
void mg_sweep(hipStream_t s, const DLev& l, const DBC& bc, int inh, int nsweeps, int allow_fused) {
    sweep_impl<double>(s, l, bc, inh, nsweeps, allow_fused && fused_ok(l, bc), 0);
}
void mg_sweep(hipStream_t s, const DLevF& l, const DBC& bc, int nsweeps, int zero_first, int allow_fused) {
    sweep_impl<float>(s, l, bc, 0, nsweeps, allow_fused, zero_first);
}
void mg_smooth(hipStream_t s, const DLev& l, const DBC& bc, int inh, int color) {
    const dim3 g(((l.n0 + 1) / 2 + 63) / 64, (l.n1 + 3) / 4, l.n2);
    hipLaunchKernelGGL(k_mg_smooth<double>, g, dim3(256), 0, s, l, bc, inh, color, 0);
}
template <class T>
static void resid_restrict_impl(hipStream_t s, const DLevT<T>& f, const DLevT<T>& c, const DBC& bc, int inh) {
    const dim3 g = grid3(c.n0, c.n1, c.n2);
    const int code = f.crs[0] | (f.crs[1] << 1) | (f.crs[2] << 2);
#define RR(cx, cy, cz) hipLaunchKernelGGL((k_mg_resid_restrict<T, cx, cy, cz>), g, dim3(256), 0, s, f, c, bc, inh)
    switch (code) {
    case 0: RR(0, 0, 0); break;
    case 1: RR(1, 0, 0); break;
    case 2: RR(0, 1, 0); break;
    case 3: RR(1, 1, 0); break;
    case 4: RR(0, 0, 1); break;
    case 5: RR(1, 0, 1); break;
    case 6: RR(0, 1, 1); break;
    default: RR(1, 1, 1); break;
    }
#undef RR
}
void mg_resid_restrict(hipStream_t s, const DLev& f, const DLev& c, const DBC& bc, int inh) {
    resid_restrict_impl<double>(s, f, c, bc, inh);
}
void mg_resid_restrict(hipStream_t s, const DLevF& f, const DLevF& c, const DBC& bc) {
    resid_restrict_impl<float>(s, f, c, bc, 0);
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
int resid_norm_blocks(const DLev& l) { return ((l.n0 + 63) / 64) * ((l.n1 + 3) / 4) * l.n2; }

void mg_resid_norm(hipStream_t s, const DLev& l, const DBC& bc, int inh, double* r, double* part, double* out) {
    hipLaunchKernelGGL(k_mg_resid_norm, dim3((l.n0 + 63) / 64, (l.n1 + 3) / 4, l.n2), dim3(256), 0, s,
                       l, bc, inh, r, part);
    hipLaunchKernelGGL(k_final1024, dim3(1), dim3(1024), 0, s, part, resid_norm_blocks(l), 2, 1, out);
}
void refine_resid(hipStream_t s, const DLev& l, const DBC& bc, const float* e, double* phi_new, float* r32, double* part, double* out) {
    hipLaunchKernelGGL(k_refine_resid, dim3((l.n0 + 63) / 64, (l.n1 + 3) / 4, l.n2), dim3(256), 0, s,
                       l, bc, e, phi_new, r32, part);
    hipLaunchKernelGGL(k_final1024, dim3(1), dim3(1024), 0, s, part, resid_norm_blocks(l), 2, 1, out);
}
void add_f2d(hipStream_t s, double* phi, const float* e, long long n) {
    unsigned b = (unsigned)((n + 255) / 256); if (b > 8192) b = 8192; if (b == 0) b = 1;
    hipLaunchKernelGGL(k_add_f2d, dim3(b), dim3(256), 0, s, phi, e, n);
}
void resid_to32(hipStream_t s, const DLev& l, const DBC& bc, float* r32, double* part, double* out) {
    hipLaunchKernelGGL(k_resid_to32, dim3((l.n0 + 63) / 64, (l.n1 + 3) / 4, l.n2), dim3(256), 0, s, l, bc, r32, part);
    hipLaunchKernelGGL(k_final1024, dim3(1), dim3(1024), 0, s, part, resid_norm_blocks(l), 2, 1, out);
}
void mg_tail(hipStream_t s, const DLev* levs, int nl, const DBC& bc, int read_a, int nu1, int nu2, double* work,
             double tol, int maxit, int* iters) {
    hipLaunchKernelGGL(k_mg_tail<double>, dim3(1), dim3(1024), 0, s, levs, nl, bc, read_a, nu1, nu2, work, tol, maxit, iters);
}
void mg_tail(hipStream_t s, const DLevF* levs, int nl, const DBC& bc, int read_a, int nu1, int nu2, float* work,
             double tol, int maxit, int* iters) {
    hipLaunchKernelGGL(k_mg_tail<float>, dim3(1), dim3(1024), 0, s, levs, nl, bc, read_a, nu1, nu2, work, tol, maxit, iters);
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
