/*
 * Geometric multigrid for  -a*phi - div(eps grad phi) = rhs  on gfx950: the
 * from-scratch replacement of AMReX MLABecLaplacian + MLMG
 * (ElectrostaticSolver.cpp:562-611, 713-720).
 *
 *   k_mg_gsrb_stream a full red-black Gauss-Seidel sweep in one pass: each
 *                    workgroup owns a 64x16 column tile and a z-chunk, keeps five
 *                    z-planes of phi (with a 2-cell in-plane halo) in LDS and
 *                    streams upward, updating red on plane p and black on plane
 *                    p-1 -- the same values as a red sweep followed by a black
 *                    sweep, with phi and rhs read once and phi written once;
 *                    variants fold in the z-prolongation (PROLONG) or the
 *                    convergence check's residual norm of the input (NORM).
 *   k_mg_resid_stream the residual streamed the same way: z-restriction, norm,
 *                    FP32 residual of the refinement
 *   k_mg_smooth      one colour of the sweep (small levels, periodic z)
 *   k_mg_block_down/up  one leg of a small layered level in one launch
 *   k_mg_resid_restrict   residual averaged into the coarse rhs; coarse phi = 0
 *   k_mg_prolong_blk cell-centred linear prolongation added to the fine phi
 *   k_mg_resid_max   residual with max|r| and max|b| per workgroup (fixed grid)
 *   k_bottom         exact (dense) or Jacobi-PCG bottom solve in one workgroup
 *   k_mg_tail        the coarse end of the V-cycle in one workgroup, staged in LDS
 *   k_spec_fwd/tri/inv  exact solve of a laterally periodic layered level
 *                    (plane transforms on the matrix cores, a tridiagonal
 *                    solve per lateral mode)
 * Per-level launches use 3-D grids (x tiles of 64, y tiles of 4, z) so no
 * thread does 64-bit index division.
 */
#include "fx_device.h"
#include "fx_kernels.h"

#include <algorithm>
#include <type_traits>

using namespace fxd;

namespace {
/* c / n for 0 <= c < 2^22 from a float reciprocal r = 1/n, corrected by one step */
__device__ __forceinline__ int qdiv(int c, int n, float r) {
    const int q = (int)((float)c * r);
    const int m = c - q * n;
    return q + (m >= n ? 1 : 0) - (m < 0 ? 1 : 0);
}
}  // namespace

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

/* ------------------- one-pass red + black sweep (streaming) --------------- */
namespace {
constexpr int SNS = 5;                            /* plane slots in LDS */
constexpr int SCZ_MAX = 64;                       /* largest z-chunk */
constexpr int NPW = 6;                            /* per-plane table: wx wy wzl wzh D 1/D */
}
/* Tile of the streaming sweep: 64x16 columns, one thread per black cell of
 * the tile (64x32 tiles for FP32, twice the cells per plane step in the
 * same LDS, measured slower: 165 against 169 refined steps/s). */
template <class T>
struct SGeo {
    static constexpr int TX = 64, TY = 16;
    static constexpr int THR = TX * TY / 2;
    static constexpr int HX = TX + 4, HY = TY + 4;     /* with the 2-cell in-plane halo */
    static constexpr int PL = HX * HY;                 /* LDS values per plane */
    static constexpr int LD = (PL + THR - 1) / THR;    /* plane loads per thread */
    static constexpr int NRED = (TY + 2) * (TX / 2 + 1);   /* red cells of the tile and its ring */
    static_assert(NRED <= 2 * THR, "two red cells per thread");
};

template <class T> struct Vec2;
template <> struct Vec2<double> { using type = double2; };
template <> struct Vec2<float> { using type = float2; };

/* Arguments of the streaming sweep: only what it reads, so the kernel
 * argument block stays small. */
template <class T>
struct StreamArgs {
    int n0, n1, n2, msx, msy;
    const T* __restrict__ b;
    const T* __restrict__ a;
    const uint8_t* __restrict__ mat;
    const T* __restrict__ bxk;
    const T* __restrict__ byk;
    const T* __restrict__ bzk;
    T ih2[3];
    DBC bc;
    int inh, zero_init;
    const T* __restrict__ phi_in;
    T* __restrict__ phi_out;
    int ntx, nty, cz, remap;
    const T* __restrict__ ec;   /* PROLONG: coarse correction (z-coarsened level, same n0 x n1) */
    int cn2;                    /* its number of planes */
    double* part;               /* NORM: max|r| and max|b| of the input phi, per workgroup */
};

namespace {
template <class T>
struct StreamRegs {
    T phi[SGeo<T>::LD];   /* next phi plane (raw loads; out-of-domain lanes masked when stored to LDS) */
    T cn[SGeo<T>::LD];    /* PROLONG, odd planes 2K+1: coarse plane K+1, which enters the window */
    T b[3];          /* rhs of red cells u = 0, 1 and of the black cell */
    T a[3];          /* Helmholtz term of the same cells */
    uint8_t m[3];    /* their material id */
    typename Vec2<T>::type bp;   /* NORM: rhs of the thread's pair of cells on the step's plane */
};
/* geometry of one of this thread's cells, fixed for the whole sweep */
struct StreamCell {
    int q;           /* LDS index in a plane */
    int c2;          /* in-plane global index (wrapped) */
    int g2;          /* in-plane index on the grown material box */
    int i, j;        /* wrapped coordinates */
    bool ok;         /* a cell of the domain this thread updates */
    bool lat;        /* away from every non-periodic lateral face */
};
}

/* One full red-black Gauss-Seidel sweep, phi_in -> phi_out (two buffers, so a
 * workgroup never reads values another workgroup has already updated).
 *
 * A workgroup owns a 64x16 column tile and the planes [z0, z1) (z0 even), and
 * marches upward: at step p it updates red on plane p over the tile and a
 * 1-cell ring (from old black values), then black on plane p-1 over the tile
 * (from the new red values of planes p-2, p-1, p), and stores plane p-1.  The
 * ring's red values are the ones the neighbouring workgroup computes, from
 * the same old values, so every cell gets the value of a red launch followed
 * by a black launch (to rounding: interior cells multiply by a per-plane 1/D).
 * Old phi is staged in LDS one plane at a time with a 2-cell in-plane halo
 * (five plane slots, so two barriers per plane suffice).  Every global load
 * of step p+1 (the next phi plane, rhs and the Helmholtz term of this
 * thread's three cells) is issued during step p into the other half of a
 * register ping-pong; the loop is unrolled by two so the plane parity, and
 * with it each thread's red and black cells, is known at compile time.
 * Black values are never read again by this workgroup, so they go straight
 * to HBM with the red value of the same pair.
 *
 * AMODE: 0 no Helmholtz term; 1 a[c] in semiconductor cells (material id);
 * 2 a[c] everywhere.  Coefficients must be layered (per-plane face values).
 * Tiles are numbered x fastest, then y, then z-chunk; with remap the
 * workgroup id is permuted so that each of the eight XCDs (which receive
 * workgroups round-robin) works on one contiguous range of tiles and the
 * in-plane halos are shared in that XCD's L2.  Requires: z not periodic, n0
 * even, n0 >= 64, n1 >= 8, periodic extents even, cz even. */
template <class T, int AMODE, int PROLONG, int NORM>
__global__ __launch_bounds__(SGeo<T>::THR) void k_mg_gsrb_stream(StreamArgs<T> A) {
    using G = SGeo<T>;
    constexpr int STX = G::TX, STY = G::TY, STHR = G::THR, SHX = G::HX, SPL = G::PL, SLD = G::LD, NRED = G::NRED;
    static_assert(!NORM || (AMODE == 0 && !PROLONG), "the residual norm is folded into plain sweeps only");
    __shared__ T sp[SNS * SPL];
    __shared__ T scw[NPW * (SCZ_MAX + 2)];
    const int n0 = A.n0, n1 = A.n1, n2 = A.n2;
    int w = blockIdx.x;
    if (A.remap) w = (w & 7) * (gridDim.x >> 3) + (w >> 3);
    const int tx = w % A.ntx, rest = w / A.ntx, ty = rest % A.nty, tz = rest / A.nty;
    const int tx0 = tx * STX, ty0 = ty * STY, z0 = tz * A.cz, z1 = min(z0 + A.cz, n2);
    const int tid = threadIdx.x;
    const bool perx = A.bc.lo[0] == BC_PER, pery = A.bc.lo[1] == BC_PER;
    const int plane = n0 * n1;
    const int gplane = A.msy;

    /* per-plane weights beta/h^2 (the products of mg_coef's layered branch),
     * the interior diagonal accumulated in mg_acc's order, and its inverse */
    for (int e = tid; e < z1 - z0 + 2; e += STHR) {
        const int k = z0 - 1 + e;
        T wx = T(0), wy = T(0), wzl = T(0), wzh = T(0);
        if (k >= 0 && k < n2) {
            wx = A.bxk[k] * A.ih2[0];
            wy = A.byk[k] * A.ih2[1];
            wzl = A.bzk[k] * A.ih2[2];
            wzh = A.bzk[k + 1] * A.ih2[2];
        }
        T D = T(0);
        D += wx; D += wx; D += wy; D += wy; D += wzl; D += wzh;
        T* t = scw + NPW * e;
        t[0] = wx; t[1] = wy; t[2] = wzl; t[3] = wzh; t[4] = D;
        t[5] = D != T(0) ? T(1) / D : T(0);
    }

    /* plane loads: LDS index tid + STHR s */
    int goff[SLD];
#pragma unroll
    for (int s = 0; s < SLD; ++s) {
        const int q = tid + STHR * s;
        goff[s] = -1;
        if (q < SPL) {
            int i = tx0 + q % SHX - 2, j = ty0 + q / SHX - 2;
            bool ok = true;
            if (i < 0) { if (perx) i += n0; else ok = false; }
            else if (i >= n0) { if (perx) i -= n0; else ok = false; }
            if (j < 0) { if (pery) j += n1; else ok = false; }
            else if (j >= n1) { if (pery) j -= n1; else ok = false; }
            if (ok) goff[s] = i + n0 * j;
        }
    }
    auto make_cell = [&](int hx, int hy, bool valid) {
        StreamCell c;
        int i = tx0 + hx - 2, j = ty0 + hy - 2;
        bool ok = valid && i <= n0 && j <= n1;
        if (i < 0) { if (perx) i += n0; else ok = false; }
        else if (i >= n0) { if (perx) i -= n0; else ok = false; }
        if (j < 0) { if (pery) j += n1; else ok = false; }
        else if (j >= n1) { if (pery) j -= n1; else ok = false; }
        c.q = hx + SHX * hy;
        c.i = i; c.j = j;
        c.c2 = ok ? i + n0 * j : 0;
        c.g2 = ok ? (i + 1) + A.msx * (j + 1) : 0;
        c.ok = ok;
        c.lat = (perx || (i > 0 && i < n0 - 1)) && (pery || (j > 0 && j < n1 - 1));
        return c;
    };
    /* red cells u = 0, 1 (positions tid and tid + STHR of the rows x 33 red
     * cells of the tile and its ring) for planes of parity 0 and 1; the black
     * cell (one of the pair 2*x2, 2*x2+1 of tile row tid/32) for planes of
     * parity 0 and 1 */
    StreamCell red[2][2], blk[2];
#pragma unroll
    for (int par = 0; par < 2; ++par) {
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const int qn = tid + STHR * u;
            const bool valid = qn < NRED;
            const int qq = valid ? qn : 0;
            const int hy = 1 + qq / (STX / 2 + 1), x2 = qq % (STX / 2 + 1);
            const int hx = 1 + 2 * x2 + ((tx0 - 1 + ty0 + hy - 2 + par) & 1);
            red[par][u] = make_cell(hx, hy, valid && hx <= STX + 2);
        }
        const int bhy = 2 + (tid >> 5), bx2 = tid & 31;
        const int off = (ty0 + bhy - 2 + par + 1) & 1;
        blk[par] = make_cell(2 + 2 * bx2 + off, bhy, ty0 + bhy - 2 < n1 && tx0 + 2 * bx2 < n0);
    }
    const int bhy = 2 + (tid >> 5), bx2 = tid & 31;
    const int bj = ty0 + bhy - 2, bi0 = tx0 + 2 * bx2;
    const bool bstore = bj < n1 && bi0 < n0;

    const T* ap = AMODE ? A.a : A.b;
    const uint8_t* mp = AMODE == 1 ? A.mat : reinterpret_cast<const uint8_t*>(A.b);
    /* the loads of step p (plane parity PAR): phi plane p+1, red cells of
     * plane p, black cell of plane p-1.  Every load is unconditional (clamped
     * to a valid address, masked where used). */
    auto coarse_plane = [&](T* dst, int K) {
        const int kc = plane * min(max(K, 0), A.cn2 - 1);
#pragma unroll
        for (int s = 0; s < SLD; ++s) dst[s] = A.ec[(goff[s] >= 0 ? goff[s] : 0) + kc];
    };
    auto issue = [&](StreamRegs<T>& R, int p, auto par_tag) {
        constexpr int PAR = decltype(par_tag)::value;
        const int kf = min(max(p + 1, 0), n2 - 1);
        const int kof = plane * kf;
        if (!A.zero_init) {   /* zero_init: the old phi is taken as 0 and never read */
#pragma unroll
            for (int s = 0; s < SLD; ++s) R.phi[s] = A.phi_in[(goff[s] >= 0 ? goff[s] : 0) + kof];
        }
        /* plane p+1 odd (= 2K+1) needs coarse planes K and K+1: K+1 is new */
        if (PROLONG && PAR == 0) coarse_plane(R.cn, ((p + 1) >> 1) + 1);
        const int kr = min(max(p, 0), n2 - 1), kb = min(max(p - 1, 0), n2 - 1);
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const int c = red[PAR][u].c2 + plane * kr;
            R.b[u] = A.b[c];
            if (AMODE) R.a[u] = ap[c];
            if (AMODE == 1) R.m[u] = mp[red[PAR][u].g2 + gplane * (kr + 1)];
        }
        const int c = blk[1 - PAR].c2 + plane * kb;
        R.b[2] = A.b[c];
        if (AMODE) R.a[2] = ap[c];
        if (AMODE == 1) R.m[2] = mp[blk[1 - PAR].g2 + gplane * (kb + 1)];
        if constexpr (NORM != 0) R.bp = *reinterpret_cast<const typename Vec2<T>::type*>(A.b + (bstore ? bi0 + n0 * bj : 0) + plane * kr);
    };
    auto a_use = [&](const StreamRegs<T>& R, int u) -> T {
        if (AMODE == 0) return T(0);
        if (AMODE == 2) return R.a[u];
        return ((R.m[u] & 3) >= 2) ? R.a[u] : T(0);
    };
    /* PROLONG: the two coarse planes the next fine plane interpolates from,
     * {K-1, K} before an even plane 2K and {K, K+1} from the odd plane 2K+1 on
     * (clamped planes stand for the ghost, weighted below); each coarse plane
     * is loaded once, at the odd plane where it enters. */
    T cw0[PROLONG ? SLD : 1], cw1[PROLONG ? SLD : 1];
    auto put_plane = [&](const StreamRegs<T>& R, int k, auto odd_tag) {
        constexpr int KODD = decltype(odd_tag)::value;
        T* d = sp + ((k - z0 + 2) % SNS) * SPL;
        const bool kin = !A.zero_init && k >= 0 && k < n2;
        if (PROLONG && KODD) {
#pragma unroll
            for (int s = 0; s < SLD; ++s) { cw0[s] = cw1[s]; cw1[s] = R.cn[s]; }
        }
        /* PROLONG: plane k of phi + P e, with the weights and sum order of
         * k_mg_prolong_blk for a z-only coarsening (parent 3/4, neighbour 1/4;
         * across a z face the neighbour is the homogeneous ghost) */
        T wj = T(0);
        if (PROLONG) {
            const int K = k >> 1;
            const int J = (k & 1) ? K + 1 : K - 1;
            T sg = T(1);
            if (J < 0 || J >= A.cn2) sg = ((J < 0 ? A.bc.lo[2] : A.bc.hi[2]) == BC_DIR) ? T(-1) : T(1);
            wj = T(0.25) * sg;
        }
#pragma unroll
        for (int s = 0; s < SLD; ++s) {
            const int q = tid + STHR * s;
            T v = R.phi[s];
            if (PROLONG) {
                T e = T(0);
                e += T(0.75) * (KODD ? cw0[s] : cw1[s]);
                e += wj * (KODD ? cw1[s] : cw0[s]);
                v += e;
            }
            if (q < SPL) d[q] = (kin && goff[s] >= 0) ? v : T(0);
        }
    };
    auto slot = [&](int k) { return sp + ((k - z0 + 2) % SNS) * SPL; };
    /* relaxation of one cell of plane k at LDS index q */
    auto relax = [&](const StreamCell& cl, int k, const T* pl, const T* pm, const T* pu, T bv, T av) -> T {
        const int q = cl.q;
        const T* t = scw + NPW * (k - z0 + 1);
        const T wx = t[0], wy = t[1], wzl = t[2], wzh = t[3];
        const T n0v = pm[q - 1], n1v = pm[q + 1], n2v = pm[q - SHX], n3v = pm[q + SHX], n4v = pl[q], n5v = pu[q];
        if (cl.lat && k > 0 && k < n2 - 1) {
            T S = T(0);
            S += wx * n0v; S += wx * n1v; S += wy * n2v; S += wy * n3v; S += wzl * n4v; S += wzh * n5v;
            if (AMODE == 0) return (bv + S) * t[5];
            T D = -av;
            D += wx; D += wx; D += wy; D += wy; D += wzl; D += wzh;
            return (bv + S) / D;
        }
        DLevT<T> lv;
        lv.n0 = n0; lv.n1 = n1; lv.n2 = n2;
        const T wc[6] = {wx, wx, wy, wy, wzl, wzh};
        const T nb[6] = {n0v, n1v, n2v, n3v, n4v, n5v};
        T D, S;
        mg_acc(lv, A.bc, A.inh != 0, cl.i, cl.j, k, av, wc, nb, D, S);
        return (bv + S) / D;
    };
    /* NORM: r = b - A phi of the input phi on the thread's pair of cells of
     * plane p, from the old values in LDS before the red phase of step p
     * replaces them -- except the red value below the black cell, which the
     * red phase of step p-1 has replaced: the pair's old red value is kept in
     * a register from one step to the next (the cell below a black cell is
     * the red cell of the same pair one plane down).  The arithmetic of the
     * row operator as in k_mg_resid_stream. */
    T red_below = T(0);
    double mr = 0.0, mb = 0.0;
    auto norm_pair = [&](const StreamRegs<T>& R, int p) {
        if (!(p >= 0 && p < n2 && bstore)) return;
        const int ob = (bj + p + 1) & 1;          /* 1: the pair's first cell is red */
        const int q0 = (2 + 2 * bx2) + SHX * bhy;
        const int qr = q0 + (ob ? 0 : 1), qb = q0 + (ob ? 1 : 0);
        const T* pl = slot(p - 1);
        const T* pm = slot(p);
        const T* pu = slot(p + 1);
        const T rb = red_below;
        red_below = pm[qr];
        if (p < z0 || p >= z1) return;
        const T* t = scw + NPW * (p - z0 + 1);
        const T wx = t[0], wy = t[1], wzl = t[2], wzh = t[3];
#pragma unroll
        for (int e = 0; e < 2; ++e) {
            const int q = e == 0 ? qr : qb;
            const int ci = bi0 + (q - q0);
            const T n0v = pm[q - 1], n1v = pm[q + 1], n2v = pm[q - SHX], n3v = pm[q + SHX];
            const T n4v = e == 0 ? pl[q] : rb, n5v = pu[q];
            const T pc = pm[q];
            const bool lat = (perx || (ci > 0 && ci < n0 - 1)) && (pery || (bj > 0 && bj < n1 - 1));
            T D, S;
            if (lat && p > 0 && p < n2 - 1) {
                S = T(0);
                S += wx * n0v; S += wx * n1v; S += wy * n2v; S += wy * n3v; S += wzl * n4v; S += wzh * n5v;
                D = t[4];
            } else {
                DLevT<T> lv;
                lv.n0 = n0; lv.n1 = n1; lv.n2 = n2;
                const T wc[6] = {wx, wx, wy, wy, wzl, wzh};
                const T nb[6] = {n0v, n1v, n2v, n3v, n4v, n5v};
                mg_acc(lv, A.bc, A.inh != 0, ci, bj, p, T(0), wc, nb, D, S);
            }
            const T bv = (q - q0) ? R.bp.y : R.bp.x;
            const T r = bv - (D * pc - S);
            mr = nmax(mr, fabs((double)r));
            mb = nmax(mb, fabs((double)bv));
        }
    };
    /* the black plane computed by the previous step, stored one step late so
     * that its completion is never waited for right after it is issued */
    using T2 = typename Vec2<T>::type;
    T2 pv;
    T* pptr = nullptr;
    bool pend = false;
    auto step = [&](int p, StreamRegs<T>& cur, StreamRegs<T>& nxt, auto par_tag) {
        constexpr int PAR = decltype(par_tag)::value;
        /* cur's loads and the store were issued a step ago: wait for them
         * here, before this step issues its own loads.  (s_waitcnt immediate:
         * vmcnt [3:0] = 0, expcnt [6:4] = 7 and lgkmcnt [11:8] = 15 leave those
         * counters unconstrained.) */
        __builtin_amdgcn_s_waitcnt(0x0F70);
        put_plane(cur, p + 1, std::integral_constant<int, 1 - PAR>{});
        issue(nxt, p + 1, std::integral_constant<int, 1 - PAR>{});
        if (pend) *reinterpret_cast<T2*>(pptr) = pv;
        __syncthreads();
        if constexpr (NORM != 0) {   /* the old values are read before any red cell of plane p changes */
            norm_pair(cur, p);
            __syncthreads();
        }
        if (p >= 0 && p < n2) {       /* red on plane p */
            const T* pl = slot(p - 1);
            T* pm = slot(p);
            const T* pu = slot(p + 1);
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                const StreamCell& cl = red[PAR][u];
                if (cl.ok) pm[cl.q] = relax(cl, p, pl, pm, pu, cur.b[u], a_use(cur, u));
            }
        }
        __syncthreads();
        const int k = p - 1;          /* black on plane p-1 (parity 1 - PAR) */
        pend = k >= z0 && bstore;
        if (pend) {
            const StreamCell& cl = blk[1 - PAR];
            const T* pm = slot(k);
            const T vb = relax(cl, k, slot(k - 1), pm, slot(k + 1), cur.b[2], a_use(cur, 2));
            const int ob = (bj + k + 1) & 1;
            const T vr = pm[cl.q + (ob ? -1 : 1)];
            pv.x = ob ? vr : vb;
            pv.y = ob ? vb : vr;
            pptr = A.phi_out + bi0 + n0 * bj + (long long)plane * k;
        }
    };

    StreamRegs<T> R0, R1;
    /* planes z0-2 and z0-1 straight into LDS; step z0-1's loads into R0 */
    using P0 = std::integral_constant<int, 0>;
    using P1 = std::integral_constant<int, 1>;
    if (PROLONG) {               /* the window before the even plane z0-2 = 2K: {K-1, K} */
        coarse_plane(cw0, z0 / 2 - 2);
        coarse_plane(cw1, z0 / 2 - 1);
    }
    issue(R0, z0 - 3, P1{});
    put_plane(R0, z0 - 2, P0{});
    issue(R0, z0 - 2, P0{});
    put_plane(R0, z0 - 1, P1{});
    issue(R0, z0 - 1, P1{});
    for (int p = z0 - 1; p <= z1; p += 2) {    /* z0 even: p odd, p + 1 even */
        step(p, R0, R1, P1{});
        if (p + 1 <= z1) step(p + 1, R1, R0, P0{});
    }
    if (pend) *reinterpret_cast<T2*>(pptr) = pv;
    if constexpr (NORM != 0) {
        __shared__ double shn[STHR / 64 + 1];
        mr = block_max<STHR>(mr, shn);
        mb = block_max<STHR>(mb, shn);
        if (tid == 0) { A.part[blockIdx.x] = mr; A.part[gridDim.x + blockIdx.x] = mb; }
    }
}

/* -------------- two red-black sweeps in one streaming pass ---------------- */
namespace {
constexpr int W2X = 64, W2Y = 8, W2H = 4;                  /* output tile; halo of two sweeps */
constexpr int W2SX = W2X + 2 * W2H, W2SY = W2Y + 2 * W2H;  /* 72 x 16 */
constexpr int W2P = W2SX * W2SY;                           /* LDS values per plane */
constexpr int W2LD = (W2P + 255) / 256;                    /* plane loads per thread */
constexpr int W2NS = 7;                                    /* plane slots */
constexpr int W2C = 7;                                     /* cells per thread per step */
constexpr int W2ZX = 8;                                    /* planes beyond the chunk in the table */
}
/* dynamic LDS bytes of k_mg_gsrb_stream2 */
template <class T>
constexpr size_t stream2_lds() { return sizeof(T) * (size_t)(W2NS * W2P + NPW * (SCZ_MAX + W2ZX)); }

namespace {
template <class T>
struct Stream2Regs {
    T phi[W2LD];     /* the plane two steps ahead of the red phase */
    T cn[W2LD];      /* PROLONG: coarse plane entering the window */
    T b[W2C];        /* rhs of this thread's cells of the next step */
    T a[W2C];
    uint8_t m[W2C];
    int q[W2C];      /* LDS index of the cell in its plane, -1: none */
    int ij[W2C];     /* its wrapped coordinates, i | j << 16 */
};
}

/* Two full red-black Gauss-Seidel sweeps, phi_in -> phi_out, in one pass.
 * A workgroup owns a 64x8 column tile and the planes [z0, z1) and holds old
 * phi with a 4-cell in-plane halo in LDS, one plane per slot.  At step p it
 * updates red of the first sweep on plane p (tile and a ring of 3), black of
 * the first sweep on plane p-1 (ring 2), red of the second sweep on plane
 * p-2 (ring 1) and black of the second sweep on plane p-3 (the tile), and
 * stores plane p-3.  Each phase reads its neighbours after the phases it
 * depends on and before the phases that overwrite them, so every value equals
 * that of four separate colour passes, and the ring values are the ones the
 * neighbouring workgroups compute; the chunk's z-halo (planes z0-4 .. z0-1
 * and z1 .. z1+3) is recomputed the same way.  Cell geometry and the loads of
 * rhs and Helmholtz term are prepared one step ahead, the next phi plane two
 * steps ahead of its red phase.  Arithmetic per cell as in k_mg_gsrb_stream.
 * Requires what k_mg_gsrb_stream requires; the chunk cz is even. */
template <class T, int AMODE, int PROLONG>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 2)))
void k_mg_gsrb_stream2(StreamArgs<T> A) {
    extern __shared__ __align__(16) unsigned char s2_lds[];
    T* const sp = reinterpret_cast<T*>(s2_lds);
    T* const scw = sp + W2NS * W2P;
    const int n0 = A.n0, n1 = A.n1, n2 = A.n2;
    int w = blockIdx.x;
    if (A.remap) w = (w & 7) * (gridDim.x >> 3) + (w >> 3);
    const int tx = w % A.ntx, rest = w / A.ntx, ty = rest % A.nty, tz = rest / A.nty;
    const int tx0 = tx * W2X, ty0 = ty * W2Y, z0 = tz * A.cz, z1 = min(z0 + A.cz, n2);
    const int tid = threadIdx.x;
    const bool perx = A.bc.lo[0] == BC_PER, pery = A.bc.lo[1] == BC_PER;
    const int plane = n0 * n1;
    const int kt0 = z0 - W2H;                         /* first plane of the table */

    for (int e = tid; e < z1 - z0 + W2ZX; e += 256) {
        const int k = kt0 + e;
        T wx = T(0), wy = T(0), wzl = T(0), wzh = T(0);
        if (k >= 0 && k < n2) {
            wx = A.bxk[k] * A.ih2[0];
            wy = A.byk[k] * A.ih2[1];
            wzl = A.bzk[k] * A.ih2[2];
            wzh = A.bzk[k + 1] * A.ih2[2];
        }
        T D = T(0);
        D += wx; D += wx; D += wy; D += wy; D += wzl; D += wzh;
        T* t = scw + NPW * e;
        t[0] = wx; t[1] = wy; t[2] = wzl; t[3] = wzh; t[4] = D;
        t[5] = D != T(0) ? T(1) / D : T(0);
    }
    int goff[W2LD];
#pragma unroll
    for (int s = 0; s < W2LD; ++s) {
        const int q = tid + 256 * s;
        goff[s] = -1;
        if (q < W2P) {
            int i = tx0 + q % W2SX - W2H, j = ty0 + q / W2SX - W2H;
            bool ok = true;
            if (i < 0) { if (perx) i += n0; else ok = false; }
            else if (i >= n0) { if (perx) i -= n0; else ok = false; }
            if (j < 0) { if (pery) j += n1; else ok = false; }
            else if (j >= n1) { if (pery) j -= n1; else ok = false; }
            if (ok) goff[s] = i + n0 * j;
        }
    }
    /* this thread's slots: phases 0..3 (red ring 3, black ring 2, red ring 1,
     * black ring 0) take cells 0-1, 2-3, 4-5 and 6; slot = row << 8 | x2 */
    int slot2[W2C];
#pragma unroll
    for (int cidx = 0; cidx < W2C; ++cidx) {
        const int ph = cidx >> 1, u = cidx & 1, r = 3 - ph;
        const int per_row = W2X / 2 + r, rows = W2Y + 2 * r;
        const int s = tid + 256 * u;
        slot2[cidx] = (s < per_row * rows) ? (((W2H - r + s / per_row) << 8) | (s % per_row)) : -1;
    }
    const T* ap = AMODE ? A.a : A.b;
    const uint8_t* mp = AMODE == 1 ? A.mat : reinterpret_cast<const uint8_t*>(A.b);
    auto slot = [&](int k) { return sp + ((k - kt0) % W2NS) * W2P; };
    auto coarse_plane = [&](T* dst, int K) {
        const int kc = plane * min(max(K, 0), A.cn2 - 1);
#pragma unroll
        for (int s = 0; s < W2LD; ++s) dst[s] = A.ec[(goff[s] >= 0 ? goff[s] : 0) + kc];
    };
    /* geometry and loads of the cells of step pn (phase ph on plane pn - ph) and
     * the phi plane pn + 1 */
    auto issue = [&](Stream2Regs<T>& R, int pn, auto par_tag) {
        constexpr int PPAR = decltype(par_tag)::value;   /* parity of pn */
        const int kf = min(max(pn + 1, 0), n2 - 1);
#pragma unroll
        for (int s = 0; s < W2LD; ++s) R.phi[s] = A.phi_in[(goff[s] >= 0 ? goff[s] : 0) + plane * kf];
        /* plane pn+1 odd (= 2K+1) brings coarse plane K+1 into the window */
        if (PROLONG && PPAR == 0) coarse_plane(R.cn, ((pn + 1) >> 1) + 1);
#pragma unroll
        for (int cidx = 0; cidx < W2C; ++cidx) {
            const int ph = cidx >> 1, r = 3 - ph, col = ph & 1;
            const int k = pn - ph;
            const int sl = slot2[cidx];
            int q = -1, ij = 0, c = 0, g = 0;
            if (sl >= 0 && k >= 0 && k < n2 && k >= z0 - 3 + ph && k <= z1 + 2 - ph) {
                const int hy = sl >> 8, x2 = sl & 255;
                int jg = ty0 - W2H + hy;
                const int hx = W2H - r + 2 * x2 + ((col + r + jg + k) & 1);
                int ig = tx0 - W2H + hx;
                bool ok = true;
                if (ig < 0) { if (perx) ig += n0; else ok = false; }
                else if (ig >= n0) { if (perx) ig -= n0; else ok = false; }
                if (jg < 0) { if (pery) jg += n1; else ok = false; }
                else if (jg >= n1) { if (pery) jg -= n1; else ok = false; }
                if (ok) {
                    q = hx + W2SX * hy;
                    ij = ig | (jg << 16);
                    c = ig + n0 * jg + plane * k;
                    g = (ig + 1) + A.msx * (jg + 1) + A.msy * (k + 1);
                }
            }
            R.q[cidx] = q;
            R.ij[cidx] = ij;
            R.b[cidx] = A.b[c];
            if (AMODE) R.a[cidx] = ap[c];
            if (AMODE == 1) R.m[cidx] = mp[g];
        }
    };
    T cw0[PROLONG ? W2LD : 1], cw1[PROLONG ? W2LD : 1];
    auto put_plane = [&](const Stream2Regs<T>& R, int k, auto odd_tag) {
        constexpr int KODD = decltype(odd_tag)::value;
        T* d = slot(k);
        const bool kin = !A.zero_init && k >= 0 && k < n2;
        if (PROLONG && KODD) {
#pragma unroll
            for (int s = 0; s < W2LD; ++s) { cw0[s] = cw1[s]; cw1[s] = R.cn[s]; }
        }
        T wj = T(0);
        if (PROLONG) {
            const int K = k >> 1;
            const int J = (k & 1) ? K + 1 : K - 1;
            T sg = T(1);
            if (J < 0 || J >= A.cn2) sg = ((J < 0 ? A.bc.lo[2] : A.bc.hi[2]) == BC_DIR) ? T(-1) : T(1);
            wj = T(0.25) * sg;
        }
#pragma unroll
        for (int s = 0; s < W2LD; ++s) {
            const int q = tid + 256 * s;
            T v = R.phi[s];
            if (PROLONG) {
                T e = T(0);
                e += T(0.75) * (KODD ? cw0[s] : cw1[s]);
                e += wj * (KODD ? cw1[s] : cw0[s]);
                v += e;
            }
            if (q < W2P) d[q] = (kin && goff[s] >= 0) ? v : T(0);
        }
    };
    auto a_use = [&](const Stream2Regs<T>& R, int cidx) -> T {
        if (AMODE == 0) return T(0);
        if (AMODE == 2) return R.a[cidx];
        return ((R.m[cidx] & 3) >= 2) ? R.a[cidx] : T(0);
    };
    auto relax = [&](const Stream2Regs<T>& R, int cidx, int k) -> T {
        const int q = R.q[cidx];
        const T* pl = slot(k - 1);
        const T* pm = slot(k);
        const T* pu = slot(k + 1);
        const T* t = scw + NPW * (k - kt0);
        const T wx = t[0], wy = t[1], wzl = t[2], wzh = t[3];
        const T n0v = pm[q - 1], n1v = pm[q + 1], n2v = pm[q - W2SX], n3v = pm[q + W2SX], n4v = pl[q], n5v = pu[q];
        const int i = R.ij[cidx] & 0xFFFF, j = R.ij[cidx] >> 16;
        const T bv = R.b[cidx], av = a_use(R, cidx);
        const bool lat = (perx || (i > 0 && i < n0 - 1)) && (pery || (j > 0 && j < n1 - 1));
        if (lat && k > 0 && k < n2 - 1) {
            T S = T(0);
            S += wx * n0v; S += wx * n1v; S += wy * n2v; S += wy * n3v; S += wzl * n4v; S += wzh * n5v;
            if (AMODE == 0) return (bv + S) * t[5];
            T D = -av;
            D += wx; D += wx; D += wy; D += wy; D += wzl; D += wzh;
            return (bv + S) / D;
        }
        DLevT<T> lv;
        lv.n0 = n0; lv.n1 = n1; lv.n2 = n2;
        const T wc[6] = {wx, wx, wy, wy, wzl, wzh};
        const T nb[6] = {n0v, n1v, n2v, n3v, n4v, n5v};
        T D, S;
        mg_acc(lv, A.bc, A.inh != 0, i, j, k, av, wc, nb, D, S);
        return (bv + S) / D;
    };
    /* the output pair of this thread (black of ring 0 with its red neighbour) */
    const int orow = tid >> 5, ox2 = tid & 31;
    const int oj = ty0 + orow, oi0 = tx0 + 2 * ox2;
    const bool ostore = oj < n1 && oi0 < n0;
    using T2 = typename Vec2<T>::type;
    T2 pv;
    T* pptr = nullptr;
    bool pend = false;
    auto step = [&](int p, Stream2Regs<T>& cur, Stream2Regs<T>& nxt, auto par_tag) {
        constexpr int PAR = decltype(par_tag)::value;    /* parity of p */
        /* cur's loads and the store were issued a step ago (s_waitcnt: vmcnt 0,
         * expcnt and lgkmcnt unconstrained) */
        __builtin_amdgcn_s_waitcnt(0x0F70);
        put_plane(cur, p + 1, std::integral_constant<int, 1 - PAR>{});
        issue(nxt, p + 1, std::integral_constant<int, 1 - PAR>{});
        if (pend) *reinterpret_cast<T2*>(pptr) = pv;
        pend = false;
#pragma unroll
        for (int ph = 0; ph < 4; ++ph) {
            __syncthreads();
            const int k = p - ph;
#pragma unroll
            for (int u = 0; u < (ph == 3 ? 1 : 2); ++u) {
                const int cidx = 2 * ph + u;
                if (cur.q[cidx] < 0) continue;
                const T v = relax(cur, cidx, k);
                if (ph < 3) {
                    slot(k)[cur.q[cidx]] = v;
                } else {
                    /* black of the second sweep on plane p-3: store it with its red neighbour */
                    const int q = cur.q[cidx];
                    const T* pm = slot(k);
                    const int ob = (oj + k + 1) & 1;      /* black sits at the odd position of the pair */
                    const T vr = pm[q + (ob ? -1 : 1)];
                    pv.x = ob ? vr : v;
                    pv.y = ob ? v : vr;
                    pptr = A.phi_out + oi0 + (long long)n0 * oj + (long long)plane * k;
                    pend = ostore;
                }
            }
        }
    };
    Stream2Regs<T> R0, R1;
    using P0 = std::integral_constant<int, 0>;
    using P1 = std::integral_constant<int, 1>;
    if (PROLONG) {                 /* the window before the even plane z0-4 = 2K: {K-1, K} */
        coarse_plane(cw0, z0 / 2 - 3);
        coarse_plane(cw1, z0 / 2 - 2);
    }
    /* planes z0-4 and z0-3 into LDS, then step z0-3's loads */
    issue(R0, z0 - 5, P1{});
    put_plane(R0, z0 - 4, P0{});
    issue(R0, z0 - 4, P0{});
    put_plane(R0, z0 - 3, P1{});
    issue(R0, z0 - 3, P1{});
    for (int p = z0 - 3; p <= z1 + 2; p += 2) {   /* z0 even: p odd, p + 1 even */
        step(p, R0, R1, P1{});
        if (p + 1 <= z1 + 2) step(p + 1, R1, R0, P0{});
    }
    if (pend) *reinterpret_cast<T2*>(pptr) = pv;
}

/* ------------------ streaming residual (layered levels) ------------------- */
namespace {
constexpr int RTX = 64, RTY = 8;                  /* tile of columns */
constexpr int RHX = RTX + 2, RHY = RTY + 2;       /* with the 1-cell in-plane halo */
constexpr int RPL = RHX * RHY;                    /* LDS values per plane */
constexpr int RLD = (RPL + 255) / 256;            /* plane loads per thread */
constexpr int RNS = 4;                            /* plane slots */
}

template <class T>
struct ResidArgs {
    int n0, n1, n2, msx, msy;
    const T* __restrict__ phi;
    const T* __restrict__ b;
    const T* __restrict__ a;
    const uint8_t* __restrict__ mat;
    const T* __restrict__ bxk;
    const T* __restrict__ byk;
    const T* __restrict__ bzk;
    T ih2[3];
    DBC bc;
    int inh;
    int ntx, nty, cz, remap;
    T* __restrict__ cb;          /* MODE 1: rhs of the z-coarsened level */
    T* __restrict__ cphi;        /* MODE 1: its phi, set to 0 (null: left alone, the level's first sweep takes it as 0) */
    float* __restrict__ r32;     /* MODE 3: the residual in FP32 */
    int zero_phi;                /* MODE 3: phi is taken as 0 and not read */
    double* __restrict__ part;   /* MODE 2, 3: max|r| and max|b| per workgroup */
};

namespace {
template <class T>
struct ResidRegs {
    T phi[RLD];      /* next phi plane */
    T b[2];          /* rhs of this thread's two cells of the current plane */
    T a[2];
    uint8_t m[2];
};
}

/* r = b - A phi over a layered level, streamed upward plane by plane: a
 * workgroup owns a 64x8 column tile and the planes [z0, z1); phi is staged in
 * LDS one plane at a time with a 1-cell in-plane halo (four plane slots, one
 * barrier per plane), each thread takes two cells of the tile per plane, and
 * the loads of plane p+1 are in flight while plane p is evaluated.  Per cell
 * the arithmetic of the row operator (interior cells: the per-plane weights
 * and diagonal accumulated in its order; elsewhere the row operator itself).
 * MODE 1: the rhs of the z-coarsened level, the mean of the residuals of
 *         planes 2K and 2K+1 summed in that order, and its phi set to 0
 *         (the values of k_mg_resid_restrict for that coarsening);
 * MODE 2: max|r| and max|b| per workgroup into part[blockIdx.x] and
 *         part[gridDim.x + blockIdx.x];
 * MODE 3: as MODE 2, with r stored in FP32.
 * AMODE as in k_mg_gsrb_stream.  Requires: z not periodic, n2 >= 2; MODE 1:
 * n2 and cz even. */
template <class T, int AMODE, int MODE>
__global__ __launch_bounds__(256) void k_mg_resid_stream(ResidArgs<T> A) {
    __shared__ T sp[RNS * RPL];
    __shared__ T scw[5 * SCZ_MAX];
    __shared__ double sh[256 / 64 + 1];
    const int n0 = A.n0, n1 = A.n1, n2 = A.n2;
    int w = blockIdx.x;
    if (A.remap) w = (w & 7) * (gridDim.x >> 3) + (w >> 3);
    const int tx = w % A.ntx, rest = w / A.ntx, ty = rest % A.nty, tz = rest / A.nty;
    const int tx0 = tx * RTX, ty0 = ty * RTY, z0 = tz * A.cz, z1 = min(z0 + A.cz, n2);
    const int tid = threadIdx.x;
    const bool perx = A.bc.lo[0] == BC_PER, pery = A.bc.lo[1] == BC_PER;
    const int plane = n0 * n1;

    /* per-plane weights beta/h^2 and the interior diagonal in the row operator's order */
    for (int e = tid; e < z1 - z0; e += 256) {
        const int k = z0 + e;
        const T wx = A.bxk[k] * A.ih2[0], wy = A.byk[k] * A.ih2[1];
        const T wzl = A.bzk[k] * A.ih2[2], wzh = A.bzk[k + 1] * A.ih2[2];
        T D = T(0);
        D += wx; D += wx; D += wy; D += wy; D += wzl; D += wzh;
        T* t = scw + 5 * e;
        t[0] = wx; t[1] = wy; t[2] = wzl; t[3] = wzh; t[4] = D;
    }
    int goff[RLD];
#pragma unroll
    for (int s = 0; s < RLD; ++s) {
        const int q = tid + 256 * s;
        goff[s] = -1;
        if (q < RPL) {
            int i = tx0 + q % RHX - 1, j = ty0 + q / RHX - 1;
            bool ok = true;
            if (i < 0) { if (perx) i += n0; else ok = false; }
            else if (i >= n0) { if (perx) i -= n0; else ok = false; }
            if (j < 0) { if (pery) j += n1; else ok = false; }
            else if (j >= n1) { if (pery) j -= n1; else ok = false; }
            if (ok) goff[s] = i + n0 * j;
        }
    }
    /* this thread's cells: column i, rows j and j + 4 of the tile */
    const int ci = tx0 + (tid & 63);
    int cj[2], cq[2], cc[2], cg[2];
    bool cok[2], clat[2];
#pragma unroll
    for (int u = 0; u < 2; ++u) {
        const int j = ty0 + (tid >> 6) + 4 * u;
        cj[u] = j;
        cok[u] = ci < n0 && j < n1;
        cq[u] = (tid & 63) + 1 + RHX * ((tid >> 6) + 4 * u + 1);
        cc[u] = cok[u] ? ci + n0 * j : 0;
        cg[u] = cok[u] ? (ci + 1) + A.msx * (j + 1) : 0;
        clat[u] = (perx || (ci > 0 && ci < n0 - 1)) && (pery || (j > 0 && j < n1 - 1));
    }
    const T* ap = AMODE ? A.a : A.b;
    const uint8_t* mp = AMODE == 1 ? A.mat : reinterpret_cast<const uint8_t*>(A.b);
    auto slot = [&](int k) { return sp + ((k - z0 + 1) & (RNS - 1)) * RPL; };
    auto load_plane = [&](ResidRegs<T>& R, int k) {
        const int kof = plane * min(max(k, 0), n2 - 1);
#pragma unroll
        for (int s = 0; s < RLD; ++s) R.phi[s] = (MODE == 3 && A.zero_phi) ? T(0) : A.phi[(goff[s] >= 0 ? goff[s] : 0) + kof];
    };
    auto load_cells = [&](ResidRegs<T>& R, int k) {
        const int kk = min(max(k, 0), n2 - 1);
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const int c = cc[u] + plane * kk;
            R.b[u] = A.b[c];
            if (AMODE) R.a[u] = ap[c];
            if (AMODE == 1) R.m[u] = mp[cg[u] + A.msy * (kk + 1)];
        }
    };
    auto put_plane = [&](const ResidRegs<T>& R, int k) {
        T* d = slot(k);
        const bool kin = k >= 0 && k < n2;
#pragma unroll
        for (int s = 0; s < RLD; ++s) {
            const int q = tid + 256 * s;
            if (q < RPL) d[q] = (kin && goff[s] >= 0) ? R.phi[s] : T(0);
        }
    };
    double mr = 0.0, mb = 0.0;
    T racc[2] = {T(0), T(0)};
    auto eval = [&](const ResidRegs<T>& R, int k) {
        const T* pl = slot(k - 1);
        const T* pm = slot(k);
        const T* pu = slot(k + 1);
        const T* t = scw + 5 * (k - z0);
        const T wx = t[0], wy = t[1], wzl = t[2], wzh = t[3];
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const int q = cq[u];
            const T n0v = pm[q - 1], n1v = pm[q + 1], n2v = pm[q - RHX], n3v = pm[q + RHX], n4v = pl[q], n5v = pu[q];
            const T pc = pm[q];
            const T av = AMODE == 0 ? T(0) : (AMODE == 2 ? R.a[u] : (((R.m[u] & 3) >= 2) ? R.a[u] : T(0)));
            T D, S;
            if (clat[u] && k > 0 && k < n2 - 1) {
                S = T(0);
                S += wx * n0v; S += wx * n1v; S += wy * n2v; S += wy * n3v; S += wzl * n4v; S += wzh * n5v;
                if (AMODE == 0) D = t[4];
                else { D = -av; D += wx; D += wx; D += wy; D += wy; D += wzl; D += wzh; }
            } else {
                DLevT<T> lv;
                lv.n0 = n0; lv.n1 = n1; lv.n2 = n2;
                const T wc[6] = {wx, wx, wy, wy, wzl, wzh};
                const T nb[6] = {n0v, n1v, n2v, n3v, n4v, n5v};
                mg_acc(lv, A.bc, A.inh != 0, ci, cj[u], k, av, wc, nb, D, S);
            }
            const T r = R.b[u] - (D * pc - S);
            if (!cok[u]) continue;
            const long long c = cc[u] + (long long)plane * k;
            if (MODE == 1) {
                if ((k & 1) == 0) racc[u] = r;
                else {
                    T s = T(0);
                    s += racc[u];
                    s += r;
                    const long long cc2 = cc[u] + (long long)plane * (k >> 1);
                    A.cb[cc2] = s / T(2);
                    if (A.cphi) A.cphi[cc2] = T(0);
                }
            } else {
                if (MODE == 3) A.r32[c] = (float)r;
                mr = nmax(mr, fabs((double)r));
                mb = nmax(mb, fabs((double)R.b[u]));
            }
        }
    };
    ResidRegs<T> R0, R1;
    load_plane(R0, z0 - 1);
    put_plane(R0, z0 - 1);
    load_plane(R0, z0);
    put_plane(R0, z0);
    load_plane(R0, z0 + 1);
    load_cells(R0, z0);
    auto step = [&](int p, ResidRegs<T>& cur, ResidRegs<T>& nxt) {
        /* cur's loads were issued a step ago (s_waitcnt: vmcnt 0, expcnt and
         * lgkmcnt unconstrained) */
        __builtin_amdgcn_s_waitcnt(0x0F70);
        put_plane(cur, p + 1);
        load_plane(nxt, p + 2);
        load_cells(nxt, p + 1);
        __syncthreads();
        eval(cur, p);
    };
    for (int p = z0; p < z1; p += 2) {
        step(p, R0, R1);
        if (p + 1 < z1) step(p + 1, R1, R0);
    }
    if (MODE >= 2) {
        mr = block_max<256>(mr, sh);
        mb = block_max<256>(mb, sh);
        if (tid == 0) { A.part[blockIdx.x] = mr; A.part[gridDim.x + blockIdx.x] = mb; }
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

/* The same residual + restriction for a layered level: the children of the
 * coarse cell and their face neighbours are loaded once (a 4x4x4 block
 * without its edges and corners when every direction is coarsened), and
 * each child's row uses the per-plane weights, with the interior diagonal
 * accumulated in the row operator's order (the row operator itself at
 * domain faces).  Children summed in the order of k_mg_resid_restrict. */
template <class T, int CX, int CY, int CZ>
__global__ __launch_bounds__(256) void k_mg_resid_restrict_lay(DLevT<T> f, DLevT<T> cl, DBC bc, int inh) {
    const int I = blockIdx.x * 64 + (threadIdx.x & 63);
    const int J = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int K = blockIdx.z;
    if (I >= cl.n0 || J >= cl.n1) return;
    const int n0 = f.n0, n1 = f.n1, n2 = f.n2;
    constexpr int NX = CX ? 2 : 1, NY = CY ? 2 : 1, NZ = CZ ? 2 : 1;
    const int i0 = CX ? 2 * I : I, j0 = CY ? 2 * J : J, k0 = CZ ? 2 * K : K;
    /* fine index of offset o (-1 .. N) along a direction, wrapped where periodic, clamped elsewhere */
    auto wrap = [&](int v, int n, int d) {
        if (v < 0) return bc.lo[d] == BC_PER ? v + n : 0;
        if (v >= n) return bc.hi[d] == BC_PER ? v - n : n - 1;
        return v;
    };
    int xs[NX + 2], ys[NY + 2], zs[NZ + 2];
#pragma unroll
    for (int o = 0; o < NX + 2; ++o) xs[o] = wrap(i0 + o - 1, n0, 0);
#pragma unroll
    for (int o = 0; o < NY + 2; ++o) ys[o] = wrap(j0 + o - 1, n1, 1);
#pragma unroll
    for (int o = 0; o < NZ + 2; ++o) zs[o] = wrap(k0 + o - 1, n2, 2);
    /* phi on the children and their face neighbours: v[z][y][x], offsets -1 .. N */
    T v[NZ + 2][NY + 2][NX + 2];
#pragma unroll
    for (int z = 0; z < NZ + 2; ++z)
#pragma unroll
        for (int y = 0; y < NY + 2; ++y)
#pragma unroll
            for (int x = 0; x < NX + 2; ++x) {
                const int out = (x == 0 || x == NX + 1) + (y == 0 || y == NY + 1) + (z == 0 || z == NZ + 1);
                v[z][y][x] = out <= 1 ? f.phi[xs[x] + n0 * (ys[y] + (long long)n1 * zs[z])] : T(0);
            }
    T s = T(0);
#pragma unroll
    for (int kk = 0; kk < NZ; ++kk) {
        const int k = k0 + kk;
        const T wx = f.bxk[k] * f.ih2[0], wy = f.byk[k] * f.ih2[1];
        const T wzl = f.bzk[k] * f.ih2[2], wzh = f.bzk[k + 1] * f.ih2[2];
#pragma unroll
        for (int jj = 0; jj < NY; ++jj)
#pragma unroll
            for (int ii = 0; ii < NX; ++ii) {
                const int i = i0 + ii, j = j0 + jj;
                const long long c = i + n0 * (j + (long long)n1 * k);
                const T pc = v[kk + 1][jj + 1][ii + 1];
                const T nb[6] = {v[kk + 1][jj + 1][ii], v[kk + 1][jj + 1][ii + 2], v[kk + 1][jj][ii + 1],
                                 v[kk + 1][jj + 2][ii + 1], v[kk][jj + 1][ii + 1], v[kk + 2][jj + 1][ii + 1]};
                T a;
                if (f.read_a == 0) a = T(0);
                else if (f.read_a == 2) a = f.a[c];
                else a = ((f.mat[(i + 1) + f.msx * (j + 1) + f.msy * (k + 1)] & 3) >= 2) ? f.a[c] : T(0);
                const bool in = (bc.lo[0] == BC_PER || (i > 0 && i < n0 - 1)) &&
                                (bc.lo[1] == BC_PER || (j > 0 && j < n1 - 1)) &&
                                (bc.lo[2] == BC_PER || (k > 0 && k < n2 - 1));
                T D, S;
                if (in) {
                    S = T(0);
                    S += wx * nb[0]; S += wx * nb[1]; S += wy * nb[2]; S += wy * nb[3]; S += wzl * nb[4]; S += wzh * nb[5];
                    D = -a; D += wx; D += wx; D += wy; D += wy; D += wzl; D += wzh;
                } else {
                    const T wc[6] = {wx, wx, wy, wy, wzl, wzh};
                    mg_acc(f, bc, inh != 0, i, j, k, a, wc, nb, D, S);
                }
                s += f.b[c] - (D * pc - S);
            }
    }
    constexpr int cnt = NX * NY * NZ;
    const int t = I + cl.n0 * (J + cl.n1 * K);
    cl.b[t] = s / T(cnt);
    cl.phi[t] = T(0);
}

/* ------------------------------ prolongation ------------------------------ */
/* Cell-centred linear prolongation added to the fine phi.  In each coarsened
 * direction the fine cell takes 3/4 of its parent and 1/4 of the parent's
 * neighbour on its side; across a domain face that neighbour is the
 * homogeneous ghost (-parent Dirichlet, +parent Neumann, periodic image).
 * Per fine cell: e = sum over (z, then y, then x; parent first) of
 * (w_z * w_y) * w_x * value.  One thread per coarse cell: the parent's
 * coarse neighbourhood (-1..1 in each coarsened direction) is loaded once
 * and every child updated from it; children along x are adjacent, so they
 * are read and written as pairs when x is coarsened. */
template <class T, int CX, int CY, int CZ>
__global__ __launch_bounds__(256) void k_mg_prolong_blk(DLevT<T> f, DLevT<T> cl, DBC bc) {
    const int I = blockIdx.x * 64 + (threadIdx.x & 63);
    const int J = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int K = blockIdx.z;
    if (I >= cl.n0 || J >= cl.n1) return;
    const int P[3] = {I, J, K};
    const int cn[3] = {cl.n0, cl.n1, cl.n2};
    constexpr int CR[3] = {CX, CY, CZ};
    /* per direction: index of the neighbour below (0) and above (1), and its ghost sign */
    int nb[3][2];
    T sg[3][2];
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            int Jn = s ? P[d] + 1 : P[d] - 1;
            T g = T(1);
            if (Jn < 0 || Jn >= cn[d]) {
                const int bct = Jn < 0 ? bc.lo[d] : bc.hi[d];
                if (bct == BC_PER) Jn = Jn < 0 ? Jn + cn[d] : Jn - cn[d];
                else { g = (bct == BC_DIR) ? T(-1) : T(1); Jn = P[d]; }
            }
            nb[d][s] = Jn;
            sg[d][s] = g;
        }
    /* coarse values v[z][y][x]: index 0 parent, 1 neighbour below, 2 neighbour above */
    T v[CZ ? 3 : 1][CY ? 3 : 1][CX ? 3 : 1];
#pragma unroll
    for (int a = 0; a < (CZ ? 3 : 1); ++a)
#pragma unroll
        for (int b = 0; b < (CY ? 3 : 1); ++b)
#pragma unroll
            for (int q = 0; q < (CX ? 3 : 1); ++q) {
                const int kz = a == 0 ? K : nb[2][a - 1];
                const int jy = b == 0 ? J : nb[1][b - 1];
                const int ix = q == 0 ? I : nb[0][q - 1];
                v[a][b][q] = cl.phi[ix + cl.n0 * (jy + cl.n1 * kz)];
            }
    (void)CR;
#pragma unroll
    for (int c = 0; c < (CZ ? 2 : 1); ++c)
#pragma unroll
        for (int bb = 0; bb < (CY ? 2 : 1); ++bb) {
            const int k = CZ ? 2 * K + c : K, j = CY ? 2 * J + bb : J;
            T e[CX ? 2 : 1];
#pragma unroll
            for (int aa = 0; aa < (CX ? 2 : 1); ++aa) {
                /* child (aa, bb, c): its side's neighbour is 1 + side in v */
                T wz[2], wy[2], wx[2];
                int iz[2], iy[2], ix[2];
                wz[0] = CZ ? T(0.75) : T(1); wz[1] = T(0.25) * sg[2][c]; iz[0] = 0; iz[1] = 1 + c;
                wy[0] = CY ? T(0.75) : T(1); wy[1] = T(0.25) * sg[1][bb]; iy[0] = 0; iy[1] = 1 + bb;
                wx[0] = CX ? T(0.75) : T(1); wx[1] = T(0.25) * sg[0][aa]; ix[0] = 0; ix[1] = 1 + aa;
                T s = T(0);
#pragma unroll
                for (int za = 0; za < (CZ ? 2 : 1); ++za)
#pragma unroll
                    for (int yb = 0; yb < (CY ? 2 : 1); ++yb)
#pragma unroll
                        for (int xq = 0; xq < (CX ? 2 : 1); ++xq)
                            s += (wz[za] * wy[yb]) * wx[xq] * v[CZ ? iz[za] : 0][CY ? iy[yb] : 0][CX ? ix[xq] : 0];
                e[aa] = s;
            }
            const long long row = (long long)f.n0 * (j + (long long)f.n1 * k);
            if (CX) {
                using T2 = typename Vec2<T>::type;
                T2* p2 = reinterpret_cast<T2*>(f.phi + row + 2 * I);
                T2 o = *p2;
                o.x += e[0];
                o.y += e[CX ? 1 : 0];
                *p2 = o;
            } else {
                f.phi[row + I] += e[0];
            }
        }
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

/* r = b - A phi (inhomogeneous BCs) over the finest level with max|r| and
 * max|b| per workgroup.  A fixed grid of workgroups strides over bricks of
 * 64 x 4 columns x ZC planes (the ZC planes of a column unrolled, so their
 * loads are in flight together); the few partials are finished by
 * k_final1024.  (A last-workgroup-done atomic would serialise thousands of
 * same-address atomics at the L2.)  STORE: 0 nothing, 1 r in FP64, 2 r in
 * FP32 (the right-hand side of an FP32 correction solve). */
template <int STORE, int ZC>
__global__ __launch_bounds__(256) void k_mg_resid_max(DLev l, DBC bc, int inh, double* __restrict__ r,
                                                      float* __restrict__ r32, double* __restrict__ part) {
    __shared__ double sh[256 / 64 + 1];
    const int bx = (l.n0 + 63) / 64, by = (l.n1 + 3) / 4, bz = (l.n2 + ZC - 1) / ZC;
    const int nbrick = bx * by * bz;
    double mr = 0.0, mb = 0.0;
    for (int t = blockIdx.x; t < nbrick; t += gridDim.x) {
        const int tx = t % bx, rest = t / bx, ty = rest % by, tz = rest / by;
        const int i = tx * 64 + (threadIdx.x & 63);
        const int j = ty * 4 + (threadIdx.x >> 6);
        const int k0 = tz * ZC;
        if (i < l.n0 && j < l.n1) {
#pragma unroll
            for (int kk = 0; kk < ZC; ++kk) {
                const int k = k0 + kk;
                if (k < l.n2) {
                    double D, S;
                    mg_row(l, l.phi, i, j, k, bc, inh != 0, D, S);
                    const int c = i + l.n0 * (j + l.n1 * k);
                    const double bb = l.b[c];
                    const double rr = bb - (D * l.phi[c] - S);
                    if (STORE == 1) r[c] = rr;
                    if (STORE == 2) r32[c] = (float)rr;
                    mr = nmax(mr, fabs(rr));
                    mb = nmax(mb, fabs(bb));
                }
            }
        }
    }
    mr = block_max<256>(mr, sh);
    mb = block_max<256>(mb, sh);
    if (threadIdx.x == 0) { part[blockIdx.x] = mr; part[gridDim.x + blockIdx.x] = mb; }
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

/* ------------------------------- bottom solve ----------------------------- */
/* Bottom solve of A e = rhs (homogeneous BCs) by one workgroup of BS threads,
 * the device form of the oracle's bottom_solve: a bottom of at most
 * BOTTOM_DIRECT_MAX cells is solved exactly -- the dense matrix assembled
 * column by column from the row operator, Gaussian elimination with partial
 * pivoting (first largest |pivot|; f = M[i][k]/M[k][k]; M[i][j] -= f*M[k][j]
 * for j > k; y[i] -= f*y[k]) and back substitution by columns (e_k =
 * y_k/U_kk, then y_i -= U_ik e_k for i < k) -- and anything larger, or a singular operator (a pivot below
 * a small fraction of the largest diagonal entry: no Dirichlet face), by
 * Jacobi-PCG.  work: N*N + N values for the direct solve, 5*N for CG. */
constexpr int BOTTOM_DIRECT_MAX = 64;
constexpr int TAIL_MAXL = 16;

template <class T, int BS>
__device__ bool bottom_direct(const DLevT<T>& l, const DBC& bc, const T* rhs, T* e, T* work) {
    __shared__ int piv_sh, ok_sh;
    const int n0 = l.n0, n1 = l.n1, n2 = l.n2, s = n0 * n1;
    const int N = s * n2;
    T* M = work;
    T* y = work + N * N;
    const int tid = threadIdx.x;
    for (int idx = tid; idx < N * N; idx += BS) {
        const int r = idx / N, col = idx % N;
        const int i = r % n0, j = (r / n0) % n1, k = r / s;
        const int nbi[6] = {i > 0 ? r - 1 : r + (n0 - 1), i < n0 - 1 ? r + 1 : r - (n0 - 1),
                            j > 0 ? r - n0 : r + n0 * (n1 - 1), j < n1 - 1 ? r + n0 : r - n0 * (n1 - 1),
                            k > 0 ? r - s : r + s * (n2 - 1), k < n2 - 1 ? r + s : r - s * (n2 - 1)};
        T a, w[6], nb[6], D, S;
        mg_coef(l, bc, i, j, k, r, a, w);
#pragma unroll
        for (int q = 0; q < 6; ++q) nb[q] = nbi[q] == col ? T(1) : T(0);
        mg_acc(l, bc, false, i, j, k, a, w, nb, D, S);
        M[idx] = (r == col ? D : T(0)) - S;
    }
    for (int r = tid; r < N; r += BS) y[r] = rhs[r];
    __syncthreads();
    T dmax = T(0);
    if (tid == 0)
        for (int r = 0; r < N; ++r) dmax = fmax(dmax, fabs(M[r * N + r]));
    const T small = sizeof(T) == 8 ? T(1e-14) : T(1e-6);
    for (int kk = 0; kk < N; ++kk) {
        if (tid == 0) {
            int p = kk;
            T best = fabs(M[kk * N + kk]);
            for (int i = kk + 1; i < N; ++i) {
                const T v = fabs(M[i * N + kk]);
                if (v > best) { best = v; p = i; }
            }
            piv_sh = p;
            ok_sh = best > small * dmax;
        }
        __syncthreads();
        if (!ok_sh) return false;
        const int p = piv_sh;
        if (p != kk) {
            for (int jj = kk + tid; jj < N; jj += BS) {
                const T t = M[kk * N + jj];
                M[kk * N + jj] = M[p * N + jj];
                M[p * N + jj] = t;
            }
            if (tid == 0) { const T t = y[kk]; y[kk] = y[p]; y[p] = t; }
        }
        __syncthreads();
        const int rows = N - kk - 1, cols = N - kk;    /* column N stands for y */
        for (int idx = tid; idx < rows * cols; idx += BS) {
            const int i = kk + 1 + idx / cols, jj = kk + 1 + idx % cols;
            const T f = M[i * N + kk] / M[kk * N + kk];
            if (jj < N) M[i * N + jj] = M[i * N + jj] - f * M[kk * N + jj];
            else y[i] = y[i] - f * y[kk];
        }
        __syncthreads();
    }
    if (tid == 0) {
        for (int jj = N - 1; jj >= 0; --jj) {
            const T ej = y[jj] / M[jj * N + jj];
            e[jj] = ej;
            for (int i = 0; i < jj; ++i) y[i] = y[i] - M[i * N + jj] * ej;
        }
    }
    __syncthreads();
    return true;
}

template <class T, int BS>
__device__ void bottom_cg_dev(const DLevT<T>& l, const DBC& bc, const T* rhs, T* e, T* work, double tol, int maxit,
                              int* iters) {
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
    if (tid == 0 && iters) *iters += it;
    __syncthreads();
}

template <class T, int BS>
__device__ void bottom_solve_dev(const DLevT<T>& l, const DBC& bc, const T* rhs, T* e, T* work, double tol, int maxit,
                                 int* iters) {
    if (l.n0 * l.n1 * l.n2 <= BOTTOM_DIRECT_MAX && bottom_direct<T, BS>(l, bc, rhs, e, work)) return;
    bottom_cg_dev<T, BS>(l, bc, rhs, e, work, tol, maxit, iters);
}

/* Bottom solve as a kernel (hierarchies without a tail). */
template <class T>
__global__ __launch_bounds__(1024) void k_bottom(DLevT<T> l, DBC bc, const T* __restrict__ rhs, T* __restrict__ e,
                                                 T* __restrict__ work, double tol, int maxit, int* __restrict__ iters) {
    bottom_solve_dev<T, 1024>(l, bc, rhs, e, work, tol, maxit, iters);
}

/* ------------------- coarse tail of the V-cycle in one workgroup ----------- */
/* LDS values a staged tail level needs: phi, b, its Helmholtz term and its
 * face coefficients (per-plane values when layered, face arrays otherwise). */
template <class T>
__host__ __device__ inline size_t fxd_tail_level_values(const DLevT<T>& l) {
    const size_t n = (size_t)l.n0 * l.n1 * l.n2;
    const size_t faces = l.layered ? (size_t)(3 * l.n2 + 1)
                                   : (size_t)(l.n0 + 1) * l.n1 * l.n2 + (size_t)l.n0 * (l.n1 + 1) * l.n2 +
                                         (size_t)l.n0 * l.n1 * (l.n2 + 1);
    return 3 * n + faces;
}

/* Levels with few cells are launch-latency bound: the whole sub-V-cycle below
 * a given level (smoothing, residual + restriction, bottom solve,
 * prolongation) runs here in one 1024-thread workgroup, with the same per-cell
 * arithmetic as the separate kernels.  levs[0] is the first tail level (its b
 * already restricted, its phi zero); levs[nl-1] is the bottom.  The level
 * descriptors live in LDS, and levels ts..nl-1 are staged there entirely
 * (phi, b, Helmholtz term, face coefficients; level 0's b copied in and its
 * phi copied back out when ts == 0), as are the bottom solve's work arrays,
 * so that the phases between barriers cost LDS rather than memory latency.
 * Levels below ts stay in global memory. */
template <class T>
__global__ __launch_bounds__(1024) void k_mg_tail(const DLevT<T>* __restrict__ levs, int nl, DBC bc, int read_a,
                                                  int nu1, int nu2, T* __restrict__ gwork, double tol, int maxit,
                                                  int* __restrict__ iters, int ts) {
    constexpr int BS = 1024;
    extern __shared__ __align__(16) unsigned char tail_lds[];
    __shared__ DLevT<T> sl[TAIL_MAXL];
    __shared__ T* swork;
    const int tid = threadIdx.x;
    if (tid == 0) {
        T* q = reinterpret_cast<T*>(tail_lds);
        for (int t = 0; t < nl; ++t) {
            DLevT<T> l = levs[t];
            l.read_a = read_a ? 2 : 0;
            if (t >= ts) {
                const int n = l.n0 * l.n1 * l.n2;
                l.phi = q; q += n;
                l.b = q; q += n;
                l.a = q; q += n;
                if (l.layered) {
                    l.bxk = q; q += l.n2;
                    l.byk = q; q += l.n2;
                    l.bzk = q; q += l.n2 + 1;
                } else {
                    l.bx = q; q += (l.n0 + 1) * l.n1 * l.n2;
                    l.by = q; q += l.n0 * (l.n1 + 1) * l.n2;
                    l.bz = q; q += l.n0 * l.n1 * (l.n2 + 1);
                }
            }
            sl[t] = l;
        }
        swork = ts < nl ? q : gwork;
    }
    __syncthreads();
    /* copy the staged levels' coefficients in (and level 0's rhs) */
    for (int t = ts; t < nl; ++t) {
        const DLevT<T>& g = levs[t];
        const DLevT<T>& l = sl[t];
        const int n = l.n0 * l.n1 * l.n2;
        for (int c = tid; c < n; c += BS) {
            if (read_a) l.a[c] = g.a[c];
            if (t == 0) { l.b[c] = g.b[c]; l.phi[c] = T(0); }
        }
        if (l.layered) {
            /* the per-plane arrays are read-only everywhere else; here they are this kernel's LDS copies */
            T* bxk = const_cast<T*>(l.bxk);
            T* byk = const_cast<T*>(l.byk);
            T* bzk = const_cast<T*>(l.bzk);
            for (int c = tid; c < l.n2; c += BS) { bxk[c] = g.bxk[c]; byk[c] = g.byk[c]; }
            for (int c = tid; c <= l.n2; c += BS) bzk[c] = g.bzk[c];
        } else {
            const int nx = (l.n0 + 1) * l.n1 * l.n2, ny = l.n0 * (l.n1 + 1) * l.n2, nz = l.n0 * l.n1 * (l.n2 + 1);
            for (int c = tid; c < nx; c += BS) l.bx[c] = g.bx[c];
            for (int c = tid; c < ny; c += BS) l.by[c] = g.by[c];
            for (int c = tid; c < nz; c += BS) l.bz[c] = g.bz[c];
        }
    }
    __syncthreads();
    auto lev = [&](int t) -> const DLevT<T>& { return sl[t]; };
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
        const DLevT<T>& l = lev(t);
        for (int s = 0; s < nu1; ++s) { half_sweep(l, 0); half_sweep(l, 1); }
        restrict_to(l, lev(t + 1));
    }
    {
        const DLevT<T>& l = lev(nl - 1);
        bottom_solve_dev<T, BS>(l, bc, l.b, l.phi, swork, tol, maxit, iters);
    }
    for (int t = nl - 2; t >= 0; --t) {
        const DLevT<T>& l = lev(t);
        prolong_from(l, lev(t + 1));
        for (int s = 0; s < nu2; ++s) { half_sweep(l, 0); half_sweep(l, 1); }
    }
    if (ts == 0) {
        const int n = levs[0].n0 * levs[0].n1 * levs[0].n2;
        for (int c = tid; c < n; c += BS) levs[0].phi[c] = sl[0].phi[c];
    }
}

/* --------- small layered levels: a whole leg per launch, blocks in LDS ------- */
/* A level too small to fill the chip with streaming planes but too large for
 * the single-workgroup tail runs each leg of the V-cycle as one launch: every
 * workgroup owns a bx x by column block over all planes (z not periodic) and
 * holds it in LDS with an in-plane halo wide enough that its colour phases
 * need nothing from other workgroups.  A red phase on the block and a ring of
 * width r, then a black phase on ring r-1, and so on: the values on the
 * shrinking rings are exactly those the owning workgroups compute, so the
 * result is the result of separate launches.
 *   down: phi = 0 (the level's phi after restriction), nu sweeps on rings
 *         2nu .. 1 (halo 2nu+1), the residual on the block restricted into
 *         the coarse rhs (coarse phi = 0), and the block's phi stored;
 *   up:   phi on the block and halo 2nu plus the prolongated correction from
 *         the coarse level (staged in LDS with its own halo), nu sweeps on
 *         rings 2nu-1 .. 0, the block stored.
 * Per cell the arithmetic of the tail: the per-plane weights, the interior
 * diagonal in the row operator's order (times 1/D without a Helmholtz
 * term), the row operator's face terms elsewhere. */
namespace {
constexpr int BLK_THREADS = 1024;
constexpr int BLK_MAXZ = 64;
__device__ __forceinline__ bool blk_wrap(int& v, int n, bool per) {
    if (v < 0) { if (!per) return false; v += n; }
    else if (v >= n) { if (!per) return false; v -= n; }
    return true;
}
}  // namespace

template <class T>
struct BlkArgs {
    DLevT<T> l;        /* the level (layered) */
    DLevT<T> c;        /* the next coarser level */
    DBC bc;
    int nu, bx, by;
    T* out;            /* where the leg writes the block's phi */
};

/* (A phi)_c = D phi_c - S of the LDS cell q (global cell (i, j, k)) */
template <class T>
__device__ __forceinline__ bool blk_row(const T* V, int q, int sx, int sxy, int i, int j, int k, int n0, int n1,
                                        int n2, const DBC& bc, bool perx, bool pery, const T* tb, T av, bool ra,
                                        T& D, T& S) {
    const T wx = tb[0], wy = tb[1], wzl = tb[2], wzh = tb[3];
    const T v0 = V[q - 1], v1 = V[q + 1], v2 = V[q - sx], v3 = V[q + sx];
    const T v4 = V[k > 0 ? q - sxy : q], v5 = V[k < n2 - 1 ? q + sxy : q];
    const bool in = (perx || (i > 0 && i < n0 - 1)) && (pery || (j > 0 && j < n1 - 1)) && k > 0 && k < n2 - 1;
    if (in) {
        S = T(0);
        S += wx * v0; S += wx * v1; S += wy * v2; S += wy * v3; S += wzl * v4; S += wzh * v5;
        if (ra) { D = -av; D += wx; D += wx; D += wy; D += wy; D += wzl; D += wzh; }
        else D = tb[4];
        return true;
    }
    D = -av;
    S = T(0);
    auto face = [&](bool open, int kind, T wv, T nv) {
        D += open ? wv : (kind == BC_DIR ? T(2) * wv : T(0));
        S += open ? wv * nv : T(0);
    };
    face(i > 0 || perx, bc.lo[0], wx, v0);
    face(i < n0 - 1 || perx, bc.hi[0], wx, v1);
    face(j > 0 || pery, bc.lo[1], wy, v2);
    face(j < n1 - 1 || pery, bc.hi[1], wy, v3);
    face(k > 0, bc.lo[2], wzl, v4);
    face(k < n2 - 1, bc.hi[2], wzh, v5);
    return false;
}

/* Shared part of the block kernels: the per-plane table, the rhs (and
 * Helmholtz term) of the cells the phases update staged in LDS, and the
 * colour phases on rings of the block held in V (extent (bx+2H) x (by+2H) x
 * n2; the rhs on rings up to H-1). */
template <class T>
struct BlkLevel {
    const BlkArgs<T>& A;
    T* V;
    T* tab;
    T* Bv;
    T* Av;
    int H, LX, LY, BLX, BLY, x0, y0, tid;
    bool perx, pery, ra;
    __device__ BlkLevel(const BlkArgs<T>& a, T* v, T* t, int h) : A(a), V(v), tab(t), H(h) {
        LX = A.bx + 2 * H;
        LY = A.by + 2 * H;
        BLX = LX - 2;
        BLY = LY - 2;
        Bv = nullptr;
        Av = nullptr;
        const int nbx = A.l.n0 / A.bx;
        x0 = (blockIdx.x % nbx) * A.bx;
        y0 = (blockIdx.x / nbx) * A.by;
        tid = threadIdx.x;
        perx = A.bc.lo[0] == BC_PER;
        pery = A.bc.lo[1] == BC_PER;
        ra = A.l.read_a != 0;
    }
    __device__ void table() {
        const DLevT<T>& l = A.l;
        for (int k = tid; k < l.n2; k += BLK_THREADS) {
            const T wx = l.bxk[k] * l.ih2[0], wy = l.byk[k] * l.ih2[1];
            const T wzl = l.bzk[k] * l.ih2[2], wzh = l.bzk[k + 1] * l.ih2[2];
            T D = T(0);
            D += wx; D += wx; D += wy; D += wy; D += wzl; D += wzh;
            T* t = tab + 6 * k;
            t[0] = wx; t[1] = wy; t[2] = wzl; t[3] = wzh; t[4] = D; t[5] = T(1) / D;
        }
    }
    /* the rhs (and Helmholtz term) on rings 0 .. H-1, from b0 on (one pass of loads) */
    __device__ void load_coef(T* b0) {
        const DLevT<T>& l = A.l;
        Bv = b0;
        const int tot = BLX * BLY * l.n2;
        Av = ra ? b0 + tot : nullptr;
        const float rx = 1.0f / (float)BLX, rxy = 1.0f / (float)(BLX * BLY);
        for (int t = tid; t < tot; t += BLK_THREADS) {
            const int k = qdiv(t, BLX * BLY, rxy);
            const int rr = t - k * BLX * BLY;
            const int yl = qdiv(rr, BLX, rx);
            const int xl = rr - yl * BLX;
            int i = x0 - (H - 1) + xl, j = y0 - (H - 1) + yl;
            const bool ok = blk_wrap(i, l.n0, perx) && blk_wrap(j, l.n1, pery);
            const long long c = i + (long long)l.n0 * (j + (long long)l.n1 * k);
            Bv[t] = ok ? l.b[c] : T(0);
            if (ra) Av[t] = ok ? l.a[c] : T(0);
        }
    }
    __device__ int qb(int xl, int jl, int k) const { return xl + BLX * (jl + BLY * k); }
    /* one colour (0 red: i+j+k even) on the block and its ring of width r */
    __device__ void phase(int r, int col) {
        const DLevT<T>& l = A.l;
        const int X = A.bx + 2 * r, Y = A.by + 2 * r;
        const int half = (X + 1) >> 1;
        const int tot = half * Y * l.n2;
        const float rh = 1.0f / (float)half, ry = 1.0f / (float)Y;
        for (int t = tid; t < tot; t += BLK_THREADS) {
            const int jk = qdiv(t, half, rh);
            const int ih = t - jk * half;
            const int k = qdiv(jk, Y, ry);
            const int jl = jk - k * Y;
            const int ju = y0 - r + jl;
            /* x0 is even and periodic extents are even, so the parity of the
             * unwrapped coordinates is that of the cell */
            const int xl = 2 * ih + ((ju + k + col + r) & 1);
            if (xl >= X) continue;
            int i = x0 - r + xl, j = ju;
            if (!blk_wrap(i, l.n0, perx) || !blk_wrap(j, l.n1, pery)) continue;
            const int q = (H - r + xl) + LX * ((H - r + jl) + LY * k);
            const int qq = qb(H - 1 - r + xl, H - 1 - r + jl, k);
            const T* tb = tab + 6 * k;
            T D, S;
            const bool in = blk_row(V, q, LX, LX * LY, i, j, k, l.n0, l.n1, l.n2, A.bc, perx, pery, tb,
                                    ra ? Av[qq] : T(0), ra, D, S);
            const T bv = Bv[qq];
            V[q] = (in && !ra) ? (bv + S) * tb[5] : (bv + S) / D;
        }
    }
    /* the block (ring 0) to the leg's output */
    __device__ void store() {
        const DLevT<T>& l = A.l;
        const int tot = A.bx * A.by * l.n2;
        const float rx = 1.0f / (float)A.bx, rxy = 1.0f / (float)(A.bx * A.by);
        for (int t = tid; t < tot; t += BLK_THREADS) {
            const int k = qdiv(t, A.bx * A.by, rxy);
            const int rr = t - k * A.bx * A.by;
            const int jl = qdiv(rr, A.bx, rx);
            const int xl = rr - jl * A.bx;
            A.out[(x0 + xl) + (long long)l.n0 * ((y0 + jl) + (long long)l.n1 * k)] = V[(H + xl) + LX * ((H + jl) + LY * k)];
        }
    }
};

template <class T, int CX, int CY, int CZ>
__global__ __launch_bounds__(BLK_THREADS) void k_mg_block_down(BlkArgs<T> A) {
    extern __shared__ __align__(16) unsigned char blk_lds[];
    __shared__ T tab[6 * BLK_MAXZ];
    BlkLevel<T> B(A, reinterpret_cast<T*>(blk_lds), tab, 2 * A.nu + 1);
    const DLevT<T>& l = A.l;
    B.table();
    const int nv = B.LX * B.LY * l.n2;
    B.load_coef(B.V + nv);
    for (int q = B.tid; q < nv; q += BLK_THREADS) B.V[q] = T(0);
    __syncthreads();
    for (int s = 0; s < A.nu; ++s)
        for (int col = 0; col < 2; ++col) {
            B.phase(B.H - 1 - (2 * s + col), col);
            __syncthreads();
        }
    /* residual on the block, restricted: children summed z outer, x inner */
    const int cbx = A.bx >> CX, cby = A.by >> CY;
    const int ctot = cbx * cby * (l.n2 >> CZ);
    const float rcx = 1.0f / (float)cbx, rcxy = 1.0f / (float)(cbx * cby);
    for (int t = B.tid; t < ctot; t += BLK_THREADS) {
        const int K = qdiv(t, cbx * cby, rcxy);
        const int rr = t - K * cbx * cby;
        const int Jl = qdiv(rr, cbx, rcx);
        const int Il = rr - Jl * cbx;
        T s = T(0);
#pragma unroll
        for (int kk = 0; kk <= CZ; ++kk)
#pragma unroll
            for (int jj = 0; jj <= CY; ++jj)
#pragma unroll
                for (int ii = 0; ii <= CX; ++ii) {
                    const int xl = (Il << CX) + ii, jl = (Jl << CY) + jj, k = (K << CZ) + kk;
                    const int i = B.x0 + xl, j = B.y0 + jl;
                    const int q = (B.H + xl) + B.LX * ((B.H + jl) + B.LY * k);
                    const int qq = B.qb(B.H - 1 + xl, B.H - 1 + jl, k);
                    T D, S;
                    blk_row(B.V, q, B.LX, B.LX * B.LY, i, j, k, l.n0, l.n1, l.n2, A.bc, B.perx, B.pery, tab + 6 * k,
                            B.ra ? B.Av[qq] : T(0), B.ra, D, S);
                    s += B.Bv[qq] - (D * B.V[q] - S);
                }
        const long long ct = ((B.x0 >> CX) + Il) + (long long)A.c.n0 * (((B.y0 >> CY) + Jl) + (long long)A.c.n1 * K);
        A.c.b[ct] = s / T((1 + CX) * (1 + CY) * (1 + CZ));
        A.c.phi[ct] = T(0);
    }
    B.store();
}

template <class T, int CX, int CY, int CZ>
__global__ __launch_bounds__(BLK_THREADS) void k_mg_block_up(BlkArgs<T> A) {
    extern __shared__ __align__(16) unsigned char blk_lds[];
    __shared__ T tab[6 * BLK_MAXZ];
    BlkLevel<T> B(A, reinterpret_cast<T*>(blk_lds), tab, 2 * A.nu);
    const DLevT<T>& l = A.l;
    const DLevT<T>& c = A.c;
    const int H = B.H, LX = B.LX, LY = B.LY, n2 = l.n2;
    /* the coarse correction on the parents of the block and halo, with their neighbours */
    const int HC = H / 2 + 1;
    const int cx0 = CX ? (B.x0 >> 1) - HC : B.x0 - H, cy0 = CY ? (B.y0 >> 1) - HC : B.y0 - H;
    const int CLX = CX ? (A.bx >> 1) + 2 * HC : LX, CLY = CY ? (A.by >> 1) + 2 * HC : LY;
    T* const V = B.V;
    T* const C = V + LX * LY * n2;
    B.table();
    B.load_coef(C + CLX * CLY * c.n2);
    {
        const int tot = CLX * CLY * c.n2;
        const float rx = 1.0f / (float)CLX, rxy = 1.0f / (float)(CLX * CLY);
        for (int t = B.tid; t < tot; t += BLK_THREADS) {
            const int K = qdiv(t, CLX * CLY, rxy);
            const int rr = t - K * CLX * CLY;
            const int yl = qdiv(rr, CLX, rx);
            const int xl = rr - yl * CLX;
            int I = cx0 + xl, J = cy0 + yl;
            const bool ok = blk_wrap(I, c.n0, B.perx) && blk_wrap(J, c.n1, B.pery);
            C[t] = ok ? c.phi[I + (long long)c.n0 * (J + (long long)c.n1 * K)] : T(0);
        }
        const int nv = LX * LY * n2;
        const float fx = 1.0f / (float)LX, fxy = 1.0f / (float)(LX * LY);
        for (int t = B.tid; t < nv; t += BLK_THREADS) {
            const int k = qdiv(t, LX * LY, fxy);
            const int rr = t - k * LX * LY;
            const int yl = qdiv(rr, LX, fx);
            const int xl = rr - yl * LX;
            int i = B.x0 - H + xl, j = B.y0 - H + yl;
            const bool ok = blk_wrap(i, l.n0, B.perx) && blk_wrap(j, l.n1, B.pery);
            V[t] = ok ? l.phi[i + (long long)l.n0 * (j + (long long)l.n1 * k)] : T(0);
        }
    }
    __syncthreads();
    /* phi += P e on the block and halo (weights, products and sum order of k_mg_prolong_blk) */
    {
        const int nv = LX * LY * n2;
        const float fx = 1.0f / (float)LX, fxy = 1.0f / (float)(LX * LY);
        for (int t = B.tid; t < nv; t += BLK_THREADS) {
            const int k = qdiv(t, LX * LY, fxy);
            const int rr = t - k * LX * LY;
            const int yl = qdiv(rr, LX, fx);
            const int xl = rr - yl * LX;
            const int iu = B.x0 - H + xl, ju = B.y0 - H + yl;
            int i = iu, j = ju;
            if (!blk_wrap(i, l.n0, B.perx) || !blk_wrap(j, l.n1, B.pery)) continue;
            /* per direction: LDS coordinate of the parent and of its neighbour on the side, weights */
            int px[2], py[2], pz[2];
            T wx[2], wy[2], wz[2];
            if (CX) {
                const int P = iu >> 1, s = iu & 1;
                int Nb = s ? P + 1 : P - 1;
                T g = T(1);
                const int Nw = s ? (i >> 1) + 1 : (i >> 1) - 1;   /* the neighbour's wrapped index */
                if (!B.perx && (Nw < 0 || Nw >= c.n0)) { g = ((Nw < 0 ? A.bc.lo[0] : A.bc.hi[0]) == BC_DIR) ? T(-1) : T(1); Nb = P; }
                px[0] = P - cx0; px[1] = Nb - cx0; wx[0] = T(0.75); wx[1] = T(0.25) * g;
            } else { px[0] = px[1] = xl; wx[0] = T(1); wx[1] = T(0); }
            if (CY) {
                const int P = ju >> 1, s = ju & 1;
                int Nb = s ? P + 1 : P - 1;
                T g = T(1);
                const int Nw = s ? (j >> 1) + 1 : (j >> 1) - 1;
                if (!B.pery && (Nw < 0 || Nw >= c.n1)) { g = ((Nw < 0 ? A.bc.lo[1] : A.bc.hi[1]) == BC_DIR) ? T(-1) : T(1); Nb = P; }
                py[0] = P - cy0; py[1] = Nb - cy0; wy[0] = T(0.75); wy[1] = T(0.25) * g;
            } else { py[0] = py[1] = yl; wy[0] = T(1); wy[1] = T(0); }
            if (CZ) {
                const int P = k >> 1, s = k & 1;
                int Nb = s ? P + 1 : P - 1;
                T g = T(1);
                if (Nb < 0 || Nb >= c.n2) { g = ((Nb < 0 ? A.bc.lo[2] : A.bc.hi[2]) == BC_DIR) ? T(-1) : T(1); Nb = P; }
                pz[0] = P; pz[1] = Nb; wz[0] = T(0.75); wz[1] = T(0.25) * g;
            } else { pz[0] = pz[1] = k; wz[0] = T(1); wz[1] = T(0); }
            T e = T(0);
#pragma unroll
            for (int za = 0; za < (CZ ? 2 : 1); ++za)
#pragma unroll
                for (int yb = 0; yb < (CY ? 2 : 1); ++yb)
#pragma unroll
                    for (int xq = 0; xq < (CX ? 2 : 1); ++xq)
                        e += (wz[za] * wy[yb]) * wx[xq] * C[px[xq] + CLX * (py[yb] + CLY * pz[za])];
            V[t] += e;
        }
    }
    __syncthreads();
    for (int s = 0; s < A.nu; ++s)
        for (int col = 0; col < 2; ++col) {
            B.phase(H - 1 - (2 * s + col), col);
            __syncthreads();
        }
    B.store();
}

/* ------------------- bottom LU factors, once per operator ----------------- */
/* The elimination of bottom_direct split into factors and solve: the dense
 * matrix of the bottom level (N <= BOTTOM_DIRECT_MAX cells) assembled from
 * the row operator, Gaussian elimination with partial pivoting (the first
 * largest |pivot|), the multipliers f = M[i][k]/M[k][k] kept below the
 * diagonal and whole rows swapped, so that a solve applies the same
 * operations to any right-hand side: y = rhs permuted, y_i -= f_ik y_k in
 * order of k, then the column-oriented back substitution.  That solve is
 * applied to every unit vector at once, which gives the explicit inverse the
 * tail applies as one matrix-vector product.  One workgroup; the pivot search
 * is one wavefront's reduction.  Output: inv = A^-1 column-major
 * (inv[col*N + row]); perm[r] = the original row now at r; perm[N] = 1, or 0
 * when a pivot falls below a small fraction of the largest diagonal entry
 * (singular operator: the solve then uses Jacobi-PCG). */
template <class T>
__global__ __launch_bounds__(1024) void k_bottom_factor(DLevT<T> l, DBC bc, T* __restrict__ inv, int* __restrict__ perm) {
    constexpr int BS = 1024;
    constexpr int NM = BOTTOM_DIRECT_MAX;
    __shared__ T M[NM * NM];
    __shared__ T X[NM * NM];
    __shared__ int sp[NM];
    __shared__ int piv_sh, ok_sh;
    __shared__ T dmax_sh;
    const int n0 = l.n0, n1 = l.n1, n2 = l.n2, s = n0 * n1;
    const int N = s * n2;
    const int tid = threadIdx.x;
    for (int idx = tid; idx < N * N; idx += BS) {
        const int r = idx / N, col = idx % N;
        const int i = r % n0, j = (r / n0) % n1, k = r / s;
        const int nbi[6] = {i > 0 ? r - 1 : r + (n0 - 1), i < n0 - 1 ? r + 1 : r - (n0 - 1),
                            j > 0 ? r - n0 : r + n0 * (n1 - 1), j < n1 - 1 ? r + n0 : r - n0 * (n1 - 1),
                            k > 0 ? r - s : r + s * (n2 - 1), k < n2 - 1 ? r + s : r - s * (n2 - 1)};
        T a, w[6], nb[6], D, S;
        mg_coef(l, bc, i, j, k, r, a, w);
#pragma unroll
        for (int q = 0; q < 6; ++q) nb[q] = nbi[q] == col ? T(1) : T(0);
        mg_acc(l, bc, false, i, j, k, a, w, nb, D, S);
        M[idx] = (r == col ? D : T(0)) - S;
    }
    for (int r = tid; r < N; r += BS) sp[r] = r;
    __syncthreads();
    if (tid < 64) {
        T v = tid < N ? fabs(M[tid * N + tid]) : T(0);
        for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
        if (tid == 0) dmax_sh = v;
    }
    __syncthreads();
    const T small = sizeof(T) == 8 ? T(1e-14) : T(1e-6);
    int ok = 1;
    for (int kk = 0; kk < N; ++kk) {
        if (tid < 64) {
            T v = (tid >= kk && tid < N) ? fabs(M[tid * N + kk]) : T(-1);
            int p = tid;
            for (int o = 32; o > 0; o >>= 1) {
                const T v2 = __shfl_xor(v, o, 64);
                const int p2 = __shfl_xor(p, o, 64);
                if (v2 > v || (v2 == v && p2 < p)) { v = v2; p = p2; }
            }
            if (tid == 0) { piv_sh = p; ok_sh = v > small * dmax_sh; }
        }
        __syncthreads();
        if (!ok_sh) { ok = 0; break; }
        const int p = piv_sh;
        if (p != kk) {
            for (int jj = tid; jj < N; jj += BS) {
                const T t = M[kk * N + jj];
                M[kk * N + jj] = M[p * N + jj];
                M[p * N + jj] = t;
            }
            if (tid == 0) { const int t = sp[kk]; sp[kk] = sp[p]; sp[p] = t; }
        }
        __syncthreads();
        for (int i = kk + 1 + tid; i < N; i += BS) M[i * N + kk] = M[i * N + kk] / M[kk * N + kk];
        __syncthreads();
        const int cols = N - kk - 1;
        for (int idx = tid; idx < cols * cols; idx += BS) {
            const int i = kk + 1 + idx / cols, jj = kk + 1 + idx % cols;
            M[i * N + jj] = M[i * N + jj] - M[i * N + kk] * M[kk * N + jj];
        }
        __syncthreads();
    }
    if (ok) {
        /* the columns of A^-1: every unit vector through the solve above at once */
        for (int idx = tid; idx < N * N; idx += BS) X[idx] = sp[idx / N] == idx % N ? T(1) : T(0);
        __syncthreads();
        for (int kk = 0; kk + 1 < N; ++kk) {
            for (int idx = tid; idx < (N - kk - 1) * N; idx += BS) {
                const int i = kk + 1 + idx / N, c = idx % N;
                X[i * N + c] = X[i * N + c] - M[i * N + kk] * X[kk * N + c];
            }
            __syncthreads();
        }
        for (int jj = N - 1; jj >= 0; --jj) {
            for (int c = tid; c < N; c += BS) X[jj * N + c] = X[jj * N + c] / M[jj * N + jj];
            __syncthreads();
            for (int idx = tid; idx < jj * N; idx += BS) {
                const int i = idx / N, c = idx % N;
                X[i * N + c] = X[i * N + c] - M[i * N + jj] * X[jj * N + c];
            }
            __syncthreads();
        }
    }
    for (int q = tid; q < N * N; q += BS) inv[q] = X[(q % N) * N + q / N];
    for (int r = tid; r < N; r += BS) perm[r] = sp[r];
    if (tid == 0) perm[N] = ok;
}

/* ---------------- coarse tail of a layered hierarchy (lean) ---------------- */
namespace {
/* per-plane table: face weights beta/h^2 (wx, wy), the z weights applied to
 * the neighbours (0 across a closed z face) and to the diagonal (2w across a
 * Dirichlet face, 0 across a Neumann one), the diagonal D accumulated in the
 * row operator's order, and 1/D */
constexpr int TAIL_NPW = 8;
enum { TW_X = 0, TW_Y, TW_ZLN, TW_ZHN, TW_ZLD, TW_ZHD, TW_D, TW_ID };
constexpr int TAIL_CPT = 4;   /* cells per thread and pass of a phase (loads of all of them first) */
struct TLev {
    int n0, n1, n2, s, n, half;
    int cx, cy, cz;
    int ophi, ob, oa, otab;   /* offsets (in values) into the dynamic LDS */
    float r0, rs, rh, r1;     /* 1/n0, 1/(n0 n1), 1/half, 1/n1 for index splits */
};
__device__ __forceinline__ TLev tail_level(const int* n, const int* crs, int read_a, int& o) {
    TLev L;
    L.n0 = n[0]; L.n1 = n[1]; L.n2 = n[2];
    L.s = L.n0 * L.n1; L.n = L.s * L.n2; L.half = (L.n0 + 1) >> 1;
    L.cx = crs[0]; L.cy = crs[1]; L.cz = crs[2];
    L.ophi = o; o += L.n;
    L.ob = o; o += L.n;
    L.oa = o; if (read_a) o += L.n;
    L.otab = o; o += TAIL_NPW * L.n2;
    L.r0 = 1.0f / (float)L.n0; L.rs = 1.0f / (float)L.s;
    L.rh = 1.0f / (float)L.half; L.r1 = 1.0f / (float)L.n1;
    return L;
}
}  // namespace

/* Values of dynamic LDS the lean tail uses for levels of the given extents. */
__host__ inline size_t tail_lean_values(const int (*n)[3], int nl, int read_a) {
    size_t v = 0;
    for (int t = 0; t < nl; ++t) {
        const size_t c = (size_t)n[t][0] * n[t][1] * n[t][2];
        v += (read_a ? 3 : 2) * c + TAIL_NPW * (size_t)n[t][2];
    }
    const size_t N = (size_t)n[nl - 1][0] * n[nl - 1][1] * n[nl - 1][2];
    if (N <= (size_t)BOTTOM_DIRECT_MAX) v += N * N;
    return v;
}

/* The sub-V-cycle below a level of a layered hierarchy in one 1024-thread
 * workgroup, every level held in LDS: phi, b, the Helmholtz term (read_a) and
 * the per-plane table.  Buffers are addressed as offsets from the dynamic LDS
 * base, so every access is an LDS instruction.
 *  - Staging is one round of global loads (every thread issues all of its
 *    loads before its first LDS store).
 *  - A cell away from every non-periodic lateral face takes the per-plane
 *    table (z faces included: the weights there are the row operator's) and,
 *    without a Helmholtz term, multiplies by the per-plane 1/D; other cells
 *    run the row operator.
 *  - Each phase loads the neighbours of up to TAIL_CPT cells per thread
 *    before it updates any of them (the cells a phase writes are never read
 *    in that phase).
 *  - Bottom: the explicit inverse from k_bottom_factor, one lane per row, or
 *    Jacobi-PCG when the bottom is singular or too large.
 * Restriction, prolongation and the sweep order are those of k_mg_tail. */
template <class T>
__global__ __launch_bounds__(1024) void k_mg_tail_lay(TailPlan<T> P, DBC bc, int read_a, int nu1, int nu2,
                                                      T* __restrict__ gwork, double tol, int maxit,
                                                      int* __restrict__ iters) {
    constexpr int BS = 1024;
    extern __shared__ __align__(16) unsigned char tail_lds[];
    T* const V = reinterpret_cast<T*>(tail_lds);
    __shared__ TLev sl[TAIL_LEAN_MAXL];
    const int tid = threadIdx.x;
    const int nl = P.nl;
    const bool perx = bc.lo[0] == BC_PER, pery = bc.lo[1] == BC_PER, perz = bc.lo[2] == BC_PER;

    /* ---- staging: the layout is computed by every thread, then one round of loads ---- */
    int o = 0, ntab = 0;
    TLev L0, Lb;
    for (int t = 0; t < nl; ++t) {
        const TLev L = tail_level(P.n[t], P.crs[t], read_a, o);
        if (tid == 0) sl[t] = L;
        if (t == 0) L0 = L;
        Lb = L;
        ntab += L.n2;
    }
    const int oinv = o;
    const int NB = Lb.n;
    const bool direct = NB <= BOTTOM_DIRECT_MAX;
    /* the table entry of this thread (level tt, plane te) */
    int tt = -1, te = 0;
    for (int t = 0, acc = 0; t < nl; ++t) {
        if (tid >= acc && tid < acc + P.n[t][2]) { tt = t; te = tid - acc; }
        acc += P.n[t][2];
    }
    T gx = T(0), gy = T(0), gzl = T(0), gzh = T(0);
    if (tt >= 0) {
        gx = P.bxk[tt][te]; gy = P.byk[tt][te]; gzl = P.bzk[tt][te]; gzh = P.bzk[tt][te + 1];
    }
    T rb[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) { const int c = tid + u * BS; rb[u] = c < L0.n ? P.b0[c] : T(0); }
    T ri[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) { const int q = tid + u * BS; ri[u] = (direct && q < NB * NB) ? P.binv[q] : T(0); }
    const int okb = direct ? P.perm[NB] : 0;
    /* stores */
    auto put_table = [&](int tl, int e, T bx, T by, T bzl, T bzh) {
        const T wx = bx * P.ih2[tl][0], wy = by * P.ih2[tl][1];
        const T wzl = bzl * P.ih2[tl][2], wzh = bzh * P.ih2[tl][2];
        const int n2 = P.n[tl][2];
        const bool lo_open = e > 0 || perz, hi_open = e < n2 - 1 || perz;
        const T zln = lo_open ? wzl : T(0), zhn = hi_open ? wzh : T(0);
        const T zld = lo_open ? wzl : (bc.lo[2] == BC_DIR ? T(2) * wzl : T(0));
        const T zhd = hi_open ? wzh : (bc.hi[2] == BC_DIR ? T(2) * wzh : T(0));
        T D = T(0);
        D += wx; D += wx; D += wy; D += wy; D += zld; D += zhd;
        int ot = 0;
        for (int q = 0, oo = 0; q <= tl; ++q) ot = tail_level(P.n[q], P.crs[q], read_a, oo).otab;
        T* tb = V + ot + TAIL_NPW * e;
        tb[TW_X] = wx; tb[TW_Y] = wy; tb[TW_ZLN] = zln; tb[TW_ZHN] = zhn; tb[TW_ZLD] = zld; tb[TW_ZHD] = zhd;
        tb[TW_D] = D; tb[TW_ID] = D != T(0) ? T(1) / D : T(0);
    };
    if (tt >= 0) put_table(tt, te, gx, gy, gzl, gzh);
#pragma unroll
    for (int u = 0; u < 8; ++u) {
        const int c = tid + u * BS;
        if (c < L0.n) { V[L0.ob + c] = rb[u]; V[L0.ophi + c] = T(0); }
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) { const int q = tid + u * BS; if (direct && q < NB * NB) V[oinv + q] = ri[u]; }
    for (int c = tid + 8 * BS; c < L0.n; c += BS) { V[L0.ob + c] = P.b0[c]; V[L0.ophi + c] = T(0); }
    for (int q = tid + BS; q < ntab; q += BS) {   /* more planes than threads (never in practice) */
        int t2 = 0, e2 = q;
        while (e2 >= P.n[t2][2]) { e2 -= P.n[t2][2]; ++t2; }
        put_table(t2, e2, P.bxk[t2][e2], P.byk[t2][e2], P.bzk[t2][e2], P.bzk[t2][e2 + 1]);
    }
    if (read_a) {
        for (int t = 0, oo = 0; t < nl; ++t) {
            const TLev L = tail_level(P.n[t], P.crs[t], read_a, oo);
            for (int c = tid; c < L.n; c += BS) V[L.oa + c] = P.a[t][c];
        }
    }
    __syncthreads();

    /* (A phi)_c = D phi_c - S of cell (i,j,k) of level L; true on the table path */
    auto row = [&](const TLev& L, int i, int j, int k, int c, T& D, T& S) -> bool {
        const T* phi = V + L.ophi;
        const T* tb = V + L.otab + TAIL_NPW * k;
        const T v0 = phi[i > 0 ? c - 1 : c + (L.n0 - 1)];
        const T v1 = phi[i < L.n0 - 1 ? c + 1 : c - (L.n0 - 1)];
        const T v2 = phi[j > 0 ? c - L.n0 : c + L.n0 * (L.n1 - 1)];
        const T v3 = phi[j < L.n1 - 1 ? c + L.n0 : c - L.n0 * (L.n1 - 1)];
        const T v4 = phi[k > 0 ? c - L.s : c + L.s * (L.n2 - 1)];
        const T v5 = phi[k < L.n2 - 1 ? c + L.s : c - L.s * (L.n2 - 1)];
        const T av = read_a ? V[L.oa + c] : T(0);
        const T wx = tb[TW_X], wy = tb[TW_Y];
        const bool lat = (perx || (i > 0 && i < L.n0 - 1)) && (pery || (j > 0 && j < L.n1 - 1));
        if (lat) {
            S = T(0);
            S += wx * v0; S += wx * v1; S += wy * v2; S += wy * v3; S += tb[TW_ZLN] * v4; S += tb[TW_ZHN] * v5;
            if (read_a) { D = -av; D += wx; D += wx; D += wy; D += wy; D += tb[TW_ZLD]; D += tb[TW_ZHD]; }
            else D = tb[TW_D];
            return true;
        }
        /* the row operator's accumulation (homogeneous faces), face by face
         * with selects (adding 0 where the row operator adds nothing) */
        D = -av;
        S = T(0);
        auto face = [&](bool open, int kind, T wv, T nv) {
            D += open ? wv : (kind == BC_DIR ? T(2) * wv : T(0));
            S += open ? wv * nv : T(0);
        };
        face(i > 0 || perx, bc.lo[0], wx, v0);
        face(i < L.n0 - 1 || perx, bc.hi[0], wx, v1);
        face(j > 0 || pery, bc.lo[1], wy, v2);
        face(j < L.n1 - 1 || pery, bc.hi[1], wy, v3);
        /* z faces: the table holds the row operator's terms */
        D += tb[TW_ZLD]; S += tb[TW_ZLN] * v4;
        D += tb[TW_ZHD]; S += tb[TW_ZHN] * v5;
        return false;
    };
    auto half_sweep = [&](const TLev& L, int color) {
        const int tot = L.half * L.n1 * L.n2;
        for (int q0 = tid; q0 < tot; q0 += TAIL_CPT * BS) {
            T nv[TAIL_CPT];
            int cc[TAIL_CPT];
#pragma unroll
            for (int u = 0; u < TAIL_CPT; ++u) {
                const int q = q0 + u * BS;
                cc[u] = -1;
                nv[u] = T(0);
                if (q >= tot) continue;
                const int jk = qdiv(q, L.half, L.rh);
                const int ih = q - jk * L.half;
                const int k = qdiv(jk, L.n1, L.r1);
                const int j = jk - k * L.n1;
                const int i = 2 * ih + ((j + k + color) & 1);
                if (i >= L.n0) continue;
                const int c = i + L.n0 * (j + L.n1 * k);
                T D, S;
                const bool in = row(L, i, j, k, c, D, S);
                const T bv = V[L.ob + c];
                nv[u] = (in && !read_a) ? (bv + S) * V[L.otab + TAIL_NPW * k + TW_ID] : (bv + S) / D;
                cc[u] = c;
            }
#pragma unroll
            for (int u = 0; u < TAIL_CPT; ++u)
                if (cc[u] >= 0) V[L.ophi + cc[u]] = nv[u];
        }
        __syncthreads();
    };
    /* residual restricted: one thread per coarse cell, children summed z outer, x inner */
    auto restrict_to = [&](const TLev& F, const TLev& C) {
        for (int t = tid; t < C.n; t += BS) {
            const int K = qdiv(t, C.s, C.rs);
            const int rr = t - K * C.s;
            const int J = qdiv(rr, C.n0, C.r0);
            const int I = rr - J * C.n0;
            T sum = T(0);
            for (int kk = 0; kk <= F.cz; ++kk)
                for (int jj = 0; jj <= F.cy; ++jj)
                    for (int ii = 0; ii <= F.cx; ++ii) {
                        const int i = F.cx ? 2 * I + ii : I, j = F.cy ? 2 * J + jj : J, k = F.cz ? 2 * K + kk : K;
                        const int c = i + F.n0 * (j + F.n1 * k);
                        T D, S;
                        row(F, i, j, k, c, D, S);
                        sum += V[F.ob + c] - (D * V[F.ophi + c] - S);
                    }
            V[C.ob + t] = sum / T((1 + F.cx) * (1 + F.cy) * (1 + F.cz));
            V[C.ophi + t] = T(0);
        }
        __syncthreads();
    };
    auto prolong_from = [&](const TLev& F, const TLev& C) {
        const int crs[3] = {F.cx, F.cy, F.cz};
        const int cn[3] = {C.n0, C.n1, C.n2};
        for (int t0 = tid; t0 < F.n; t0 += TAIL_CPT * BS) {
            T ev[TAIL_CPT];
#pragma unroll
            for (int u = 0; u < TAIL_CPT; ++u) {
                const int t = t0 + u * BS;
                ev[u] = T(0);
                if (t >= F.n) continue;
                const int k = qdiv(t, F.s, F.rs);
                const int rr = t - k * F.s;
                const int j = qdiv(rr, F.n0, F.r0);
                const int fi[3] = {rr - j * F.n0, j, k};
                int id[3][2];
                T w[3][2];
                int cnt[3];
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    if (!crs[d]) { id[d][0] = fi[d]; id[d][1] = fi[d]; w[d][0] = T(1); w[d][1] = T(0); cnt[d] = 1; continue; }
                    const int I = fi[d] >> 1;
                    int Jn = (fi[d] & 1) ? I + 1 : I - 1;
                    T sg = T(1);
                    if (Jn < 0 || Jn >= cn[d]) {
                        const int bct = Jn < 0 ? bc.lo[d] : bc.hi[d];
                        if (bct == BC_PER) Jn = Jn < 0 ? Jn + cn[d] : Jn - cn[d];
                        else { sg = (bct == BC_DIR) ? T(-1) : T(1); Jn = I; }
                    }
                    id[d][0] = I; w[d][0] = T(0.75);
                    id[d][1] = Jn; w[d][1] = T(0.25) * sg;
                    cnt[d] = 2;
                }
                T e = T(0);
#pragma unroll
                for (int a = 0; a < 2; ++a)
#pragma unroll
                    for (int b = 0; b < 2; ++b)
#pragma unroll
                        for (int q = 0; q < 2; ++q)
                            if (a < cnt[2] && b < cnt[1] && q < cnt[0])
                                e += (w[2][a] * w[1][b]) * w[0][q] * V[C.ophi + id[0][q] + C.n0 * (id[1][b] + C.n1 * id[2][a])];
                ev[u] = e;
            }
#pragma unroll
            for (int u = 0; u < TAIL_CPT; ++u) {
                const int t = t0 + u * BS;
                if (t < F.n) V[F.ophi + t] += ev[u];
            }
        }
        __syncthreads();
    };
    for (int t = 0; t + 1 < nl; ++t) {
        for (int s = 0; s < nu1; ++s) { half_sweep(sl[t], 0); half_sweep(sl[t], 1); }
        restrict_to(sl[t], sl[t + 1]);
    }
    {
        const TLev& L = sl[nl - 1];
        const int N = L.n;
        if (direct && okb) {
            /* e = A^-1 b: lane i sums row i of the inverse (stored column-major) in column order */
            if (tid < N) {
                const T* inv = V + oinv;
                const T* b = V + L.ob;
                T e = T(0);
                for (int jj = 0; jj < N; ++jj) e += inv[jj * N + tid] * b[jj];
                V[L.ophi + tid] = e;
            }
            __syncthreads();
        } else {
            DLevT<T> bl;
            bl.n0 = L.n0; bl.n1 = L.n1; bl.n2 = L.n2;
            for (int d = 0; d < 3; ++d) { bl.ih2[d] = P.ih2[nl - 1][d]; bl.crs[d] = 0; }
            bl.phi = V + L.ophi; bl.b = V + L.ob; bl.a = read_a ? V + L.oa : nullptr;
            bl.bx = bl.by = bl.bz = nullptr; bl.mat = nullptr; bl.msx = bl.msy = 0; bl.matcoef = 0;
            bl.read_a = read_a ? 2 : 0;
            bl.layered = 1; bl.bxk = P.bxk[nl - 1]; bl.byk = P.byk[nl - 1]; bl.bzk = P.bzk[nl - 1];
            bl.phi2 = nullptr;
            bottom_cg_dev<T, BS>(bl, bc, bl.b, bl.phi, gwork, tol, maxit, iters);
        }
    }
    for (int t = nl - 2; t >= 0; --t) {
        prolong_from(sl[t], sl[t + 1]);
        for (int s = 0; s < nu2; ++s) { half_sweep(sl[t], 0); half_sweep(sl[t], 1); }
    }
    {
        const TLev& L = sl[0];
        for (int c = tid; c < L.n; c += BS) P.phi0[c] = V[L.ophi + c];
    }
}

/* ------------- exact spectral solve of a periodic layered level ------------ */
/* A e = b on one level of a laterally periodic layered hierarchy, exactly (to
 * rounding), in three launches: every plane transformed along x and y by one
 * workgroup on the matrix cores (with the plane as an n1 x n0 matrix F,
 * S = Qy^T F Qx), one Thomas solve in z per mode (mx, my) with diagonal
 * bx_k mu_x/dx^2 + by_k mu_y/dy^2 plus the z couplings of the row operator
 * (2w across a Dirichlet face, 0 across a Neumann one), and the inverse
 * transforms (Q is orthogonal: F = Qy S Qx^T). */
namespace {
/* C = X Y for an R x K and a K x N operand, R, K, N <= 64, on the matrix
 * cores: X[r][k] = x[r*xr + k*xk], Y[k][c] = y[k*yk + c*yc], each entry of C
 * handed to put(r, c, value).  The 16x16 output tiles are shared out over the
 * workgroup's waves; rows and columns past R, K, N read as zero. */
template <class T, class Put>
__device__ __forceinline__ void spec_mm(int R, int K, int N, const T* x, int xr, int xk, const T* y, int yk, int yc, Put put) {
    using MF = Mfma16<T>;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const int tr = (R + 15) >> 4, tc = (N + 15) >> 4;
    for (int t = wave; t < tr * tc; t += nw) {
        const int r0 = (t / tc) * 16, c0 = (t - (t / tc) * tc) * 16;
        const int r = r0 + li, c = c0 + li;
        typename MF::v4 acc = {T(0), T(0), T(0), T(0)};
        for (int k0 = 0; k0 < K; k0 += 4) {
            const int k = k0 + lk;
            const T a = (r < R && k < K) ? x[r * xr + k * xk] : T(0);
            const T b = (k < K && c < N) ? y[k * yk + c * yc] : T(0);
            acc = MF::mma(a, b, acc);
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int rr = r0 + MF::row(lane, v);
            if (rr < R && c < N) put(rr, c, acc[v]);
        }
    }
}

}  // namespace

/* The plane kernels split each plane by blocks of 16 output columns, one
 * workgroup per (column block, plane): column block c of
 * S = Qy^T (F Qx) needs only columns c of Qx, and column block c of
 * F = Qy (S Qx^T) only rows c of Qx.  LDS: the plane and Qy with rows padded
 * to n + 1 values, the 16 columns of the block (padded to 17) for Qx and for
 * the intermediate product. */
constexpr int SPEC_CB = 16, SPEC_CBL = SPEC_CB + 1, SPEC_THR = 1024;   /* 16 waves to stage, 4 to multiply */

/* plane k = blockIdx.y, output columns m0.. = 16 blockIdx.x: b -> S */
template <class T>
__global__ __launch_bounds__(SPEC_THR) void k_spec_fwd(SpecTail<T> P) {
    extern __shared__ __align__(16) unsigned char spec_lds[];
    const int n0 = P.n0, n1 = P.n1, s = n0 * n1, k = blockIdx.y, lx = n0 + 1, ly = n1 + 1;
    const int m0 = blockIdx.x * SPEC_CB, nb = min(SPEC_CB, n0 - m0);
    T* const A = reinterpret_cast<T*>(spec_lds);
    T* const Qy = A + n1 * lx;
    T* const Qc = Qy + n1 * ly;            /* Qc[i][mm] = Qx[i][m0 + mm] */
    T* const Bc = Qc + n0 * SPEC_CBL;      /* Bc[j][mm] = (F Qx)[j][m0 + mm] */
    for (int c = threadIdx.x; c < s; c += SPEC_THR) {
        const int j = c / n0;
        A[c + j] = P.b[(size_t)k * s + c];
    }
    for (int c = threadIdx.x; c < n1 * n1; c += SPEC_THR) {
        const int j = c / n1;
        Qy[c + j] = P.Qy[c];
    }
    for (int c = threadIdx.x; c < n0 * SPEC_CB; c += SPEC_THR) {
        const int i = c / SPEC_CB, mm = c - i * SPEC_CB;
        Qc[i * SPEC_CBL + mm] = mm < nb ? P.Qx[i * n0 + m0 + mm] : T(0);
    }
    __syncthreads();
    /* Bc[j][mm] = sum_i A[j][i] Qx[i][m0 + mm] */
    spec_mm<T>(n1, n0, nb, A, lx, 1, Qc, SPEC_CBL, 1, [&](int r, int c, T v) { Bc[r * SPEC_CBL + c] = v; });
    __syncthreads();
    /* S[q][m0 + mm] = sum_j Qy[j][q] Bc[j][mm] */
    T* const S = P.S + (size_t)k * s + m0;
    spec_mm<T>(n1, n1, nb, Qy, 1, ly, Bc, SPEC_CBL, 1, [&](int r, int c, T v) { S[r * n0 + c] = v; });
}

/* one thread per mode: the tridiagonal system in z, its column of S held in
 * registers (n2 <= NZ; NZ = 0: any n2, through global memory with C as
 * scratch) */
template <class T, int NZ>
__global__ __launch_bounds__(64) void k_spec_tri(SpecTail<T> P) {
    const int n0 = P.n0, s = n0 * P.n1, n2 = P.n2;
    const int mode = blockIdx.x * 64 + threadIdx.x;
    if (mode >= s) return;
    const int q = mode / n0, m = mode - q * n0;
    const T mx = P.mux[m], my = P.muy[q];
    /* row k of the system: a_k x_k-1 + d_k x_k + c_k x_k+1 = S_k */
    auto coef = [&](int k, T& a, T& d, T& c) {
        const T wzl = P.bzk[k] * P.ih2[2], wzh = P.bzk[k + 1] * P.ih2[2];
        T diag = (P.bxk[k] * P.ih2[0]) * mx;
        diag += (P.byk[k] * P.ih2[1]) * my;
        diag += k > 0 ? wzl : (P.dlo ? T(2) * wzl : T(0));
        diag += k < n2 - 1 ? wzh : (P.dhi ? T(2) * wzh : T(0));
        a = k > 0 ? -wzl : T(0);
        c = k < n2 - 1 ? -wzh : T(0);
        d = diag;
    };
    if constexpr (NZ > 0) {
        T sv[NZ], cv[NZ];
#pragma unroll
        for (int k = 0; k < NZ; ++k)
            if (k < n2) sv[k] = P.S[mode + (size_t)k * s];
        T cprev = T(0), dprev = T(0);
#pragma unroll
        for (int k = 0; k < NZ; ++k) {
            if (k < n2) {
                T a, d, c;
                coef(k, a, d, c);
                const T rden = T(1) / (d - a * cprev);
                cv[k] = c * rden;
                sv[k] = (sv[k] - a * dprev) * rden;
                cprev = cv[k];
                dprev = sv[k];
            }
        }
        T xn = T(0);
#pragma unroll
        for (int k = NZ - 1; k >= 0; --k) {
            if (k < n2) {
                xn = sv[k] - cv[k] * xn;
                P.S[mode + (size_t)k * s] = xn;
            }
        }
    } else {
        T cprev = T(0), dprev = T(0);
        for (int k = 0; k < n2; ++k) {
            T a, d, c;
            coef(k, a, d, c);
            const T rden = T(1) / (d - a * cprev);
            const T cq = c * rden;
            const T dq = (P.S[mode + (size_t)k * s] - a * dprev) * rden;
            P.C[mode + (size_t)k * s] = cq;
            P.S[mode + (size_t)k * s] = dq;
            cprev = cq;
            dprev = dq;
        }
        T xn = T(0);
        for (int k = n2 - 1; k >= 0; --k) {
            xn = P.S[mode + (size_t)k * s] - P.C[mode + (size_t)k * s] * xn;
            P.S[mode + (size_t)k * s] = xn;
        }
    }
}

/* plane k = blockIdx.y, output columns i0.. = 16 blockIdx.x: S -> phi */
template <class T>
__global__ __launch_bounds__(SPEC_THR) void k_spec_inv(SpecTail<T> P) {
    extern __shared__ __align__(16) unsigned char spec_lds[];
    const int n0 = P.n0, n1 = P.n1, s = n0 * n1, k = blockIdx.y, lx = n0 + 1, ly = n1 + 1;
    const int i0 = blockIdx.x * SPEC_CB, nb = min(SPEC_CB, n0 - i0);
    T* const A = reinterpret_cast<T*>(spec_lds);
    T* const Qy = A + n1 * lx;
    T* const Qr = Qy + n1 * ly;            /* Qr[m][ii] = Qx[i0 + ii][m] */
    T* const Tc = Qr + n0 * SPEC_CBL;      /* Tc[q][ii] = (S Qx^T)[q][i0 + ii] */
    for (int c = threadIdx.x; c < s; c += SPEC_THR) {
        const int j = c / n0;
        A[c + j] = P.S[(size_t)k * s + c];
    }
    for (int c = threadIdx.x; c < n1 * n1; c += SPEC_THR) {
        const int j = c / n1;
        Qy[c + j] = P.Qy[c];
    }
    for (int c = threadIdx.x; c < SPEC_CB * n0; c += SPEC_THR) {
        const int ii = c / n0, m = c - ii * n0;
        Qr[m * SPEC_CBL + ii] = ii < nb ? P.Qx[(i0 + ii) * n0 + m] : T(0);
    }
    __syncthreads();
    /* Tc[q][ii] = sum_m A[q][m] Qx[i0 + ii][m] */
    spec_mm<T>(n1, n0, nb, A, lx, 1, Qr, SPEC_CBL, 1, [&](int r, int c, T v) { Tc[r * SPEC_CBL + c] = v; });
    __syncthreads();
    /* phi[j][i0 + ii] = sum_q Qy[j][q] Tc[q][ii] */
    T* const phi = P.phi + (size_t)k * s + i0;
    spec_mm<T>(n1, n1, nb, Qy, ly, 1, Tc, SPEC_CBL, 1, [&](int r, int c, T v) { phi[r * n0 + c] = v; });
}

/* --------------------- FP64 refinement of an FP32 correction -------------- */
/* phi += e (FP32 correction added into the FP64 solution). */
__global__ void k_add_f2d(double* __restrict__ phi, const float* __restrict__ e, long long n, int set) {
    for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < n; c += (long long)gridDim.x * blockDim.x)
        phi[c] = set ? (double)e[c] : phi[c] + (double)e[c];   /* set: phi was 0 (0 + e is e exactly) */
}

__global__ void k_d2f(float* __restrict__ d, const double* __restrict__ s, long long n) {
    for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < n; c += (long long)gridDim.x * blockDim.x) d[c] = (float)s[c];
}

/* ------------------------------- launchers -------------------------------- */
namespace fxk {

static inline dim3 grid3(int n0, int n1, int n2) { return dim3((n0 + 63) / 64, (n1 + 3) / 4, n2); }

int resid_blocks(const DLev& l) { return ((l.n0 + 63) / 64) * ((l.n1 + 3) / 4) * l.n2; }

template <class T>
static bool stream_ok(const DLevT<T>& l, const DBC& bc) {
    return l.phi2 != nullptr && l.layered && bc.lo[2] != BC_PER && l.n0 >= 64 && l.n1 >= 8 && (l.n0 % 2) == 0 &&
           (bc.lo[1] != BC_PER || (l.n1 % 2) == 0) && l.n2 >= 2 && (l.read_a != 1 || l.mat != nullptr);
}
bool fused_ok(const DLev& l, const DBC& bc) { return stream_ok(l, bc); }
bool fused_ok32(const DLevF& l, const DBC& bc) { return stream_ok(l, bc); }

template <class T>
__global__ void k_copyT(T* __restrict__ d, const T* __restrict__ s, long long n) {
    for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < n; c += (long long)gridDim.x * blockDim.x) d[c] = s[c];
}

/* z-chunk of a streaming kernel: even, about nwg workgroups, 8..64 planes */
static int stream_chunk(int tiles, int n2, int nwg = 1024) {
    int cz = (int)(((long long)n2 * tiles + nwg - 1) / nwg);
    cz = std::max(cz, 8);
    cz = std::min(std::min(cz, n2), SCZ_MAX);
    return (cz + 1) & ~1;
}

/* Levels are streamed when large enough (DLevT::stream, from the
 * mg_stream_min_cells input, default 2M cells): below that a streaming
 * workgroup's planes are too few to cover the memory latency of each plane
 * step, and the per-level kernels or the block legs are faster. */
template <class T>
static bool stream_use(const DLevT<T>& l, const DBC& bc, int allow_fused) {
    return allow_fused && stream_ok(l, bc) && l.stream;
}

/* z-chunk of the two-sweep pass: even, about two workgroups per compute unit
 * (512 on this chip), at least 8 planes (the chunk recomputes 6 planes of
 * z-halo) and at most 64 */
static int stream2_chunk(int tiles, int n2) {
    int cz = (int)(((long long)n2 * tiles + 511) / 512);
    cz = std::max(cz, 8);
    cz = std::min(std::min(cz, n2), SCZ_MAX);
    return (cz + 1) & ~1;
}
template <class T, int AMODE, int PRO>
static void stream2_launch(hipStream_t s, const StreamArgs<T>& A, int nblk) {
    static bool set = false;
    if (!set) {
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_mg_gsrb_stream2<T, AMODE, PRO>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)stream2_lds<T>());
        set = true;
    }
    hipLaunchKernelGGL((k_mg_gsrb_stream2<T, AMODE, PRO>), dim3(nblk), dim3(256), stream2_lds<T>(), s, A);
}

/* Prolongation of the correction on c into f (one thread per coarse cell). */
template <class T>
static void prolong_impl(hipStream_t s, const DLevT<T>& f, const DLevT<T>& c, const DBC& bc) {
    const dim3 g = grid3(c.n0, c.n1, c.n2);
    const int code = f.crs[0] | (f.crs[1] << 1) | (f.crs[2] << 2);
#define PB(cx, cy, cz) hipLaunchKernelGGL((k_mg_prolong_blk<T, cx, cy, cz>), g, dim3(256), 0, s, f, c, bc)
    switch (code) {
    case 0: PB(0, 0, 0); break;
    case 1: PB(1, 0, 0); break;
    case 2: PB(0, 1, 0); break;
    case 3: PB(1, 1, 0); break;
    case 4: PB(0, 0, 1); break;
    case 5: PB(1, 0, 1); break;
    case 6: PB(0, 1, 1); break;
    default: PB(1, 1, 1); break;
    }
#undef PB
}

/* nsweeps RB-GS sweeps of level l.  With coarse != null the coarse
 * correction is first prolongated into l: folded into the first streaming
 * sweep's plane loads when l is z-coarsened (the sweep reads phi + P e), else
 * by k_mg_prolong_blk. */
template <class T>
static void sweep_impl(hipStream_t s, const DLevT<T>& l, const DBC& bc, int inh, int nsweeps, int allow_fused, int zero_first,
                       const DLevT<T>* coarse = nullptr, bool start_alt = false, bool keep_alt = false,
                       double* norm_part = nullptr, double* norm_out = nullptr) {
    const bool zonly = coarse && l.crs[0] == 0 && l.crs[1] == 0 && l.crs[2] == 1 && coarse->n0 == l.n0 && coarse->n1 == l.n1;
    if (coarse && !(zonly && nsweeps > 0 && stream_use(l, bc, allow_fused))) {
        prolong_impl<T>(s, l, *coarse, bc);
        coarse = nullptr;
    }
    if (nsweeps <= 0) return;
    if (stream_use(l, bc, allow_fused)) {
        StreamArgs<T> A;
        A.n0 = l.n0; A.n1 = l.n1; A.n2 = l.n2; A.msx = l.msx; A.msy = l.msy;
        A.b = l.b; A.a = l.a; A.mat = l.mat;
        A.bxk = l.bxk; A.byk = l.byk; A.bzk = l.bzk;
        for (int d = 0; d < 3; ++d) A.ih2[d] = l.ih2[d];
        A.bc = bc;
        A.inh = inh;
        A.ntx = (l.n0 + SGeo<T>::TX - 1) / SGeo<T>::TX;
        A.nty = (l.n1 + SGeo<T>::TY - 1) / SGeo<T>::TY;
        A.cz = stream_chunk(A.ntx * A.nty, l.n2, 1024 * 256 / SGeo<T>::THR);   /* about four waves per SIMD */
        const int nz = (l.n2 + A.cz - 1) / A.cz;
        const int nblk = A.ntx * A.nty * nz;
        A.remap = (nblk % 8) == 0 ? 1 : 0;
        /* start_alt: the input is in phi2; keep_alt: leave an odd result in phi2 */
        T* cur = start_alt ? l.phi2 : l.phi;
        T* nxt = start_alt ? l.phi : l.phi2;
        int it = 0;
        if (allow_fused >= 2 && nsweeps >= 2) {
            /* pairs of sweeps in one pass each */
            StreamArgs<T> B = A;
            B.ntx = (l.n0 + W2X - 1) / W2X;
            B.nty = (l.n1 + W2Y - 1) / W2Y;
            B.cz = stream2_chunk(B.ntx * B.nty, l.n2);
            const int nb2 = B.ntx * B.nty * ((l.n2 + B.cz - 1) / B.cz);
            B.remap = (nb2 % 8) == 0 ? 1 : 0;
            for (; it + 1 < nsweeps; it += 2) {
                const bool pro = coarse && it == 0;
                B.zero_init = (zero_first && it == 0 && !pro) ? 1 : 0;
                B.phi_in = cur;
                B.phi_out = nxt;
                B.ec = pro ? coarse->phi : nullptr;
                B.cn2 = pro ? coarse->n2 : 0;
#define SWEEP2(am, pr) stream2_launch<T, am, pr>(s, B, nb2)
                if (pro) {
                    switch (l.read_a) { case 0: SWEEP2(0, 1); break; case 1: SWEEP2(1, 1); break; default: SWEEP2(2, 1); break; }
                } else {
                    switch (l.read_a) { case 0: SWEEP2(0, 0); break; case 1: SWEEP2(1, 0); break; default: SWEEP2(2, 0); break; }
                }
#undef SWEEP2
                std::swap(cur, nxt);
            }
        }
        for (; it < nsweeps; ++it) {
            const bool pro = coarse && it == 0;
            A.zero_init = (zero_first && it == 0 && !pro) ? 1 : 0;
            A.phi_in = cur;
            A.phi_out = nxt;
            A.ec = pro ? coarse->phi : nullptr;
            A.cn2 = pro ? coarse->n2 : 0;
#define SWEEP(am, pr) hipLaunchKernelGGL((k_mg_gsrb_stream<T, am, pr, 0>), dim3(nblk), dim3(SGeo<T>::THR), 0, s, A)
            if (norm_part && it == 0 && !pro && l.read_a == 0) {   /* zero_init as set above */
                /* the first sweep also reduces max|r| and max|b| of its input */
                A.part = norm_part;
                hipLaunchKernelGGL((k_mg_gsrb_stream<T, 0, 0, 1>), dim3(nblk), dim3(SGeo<T>::THR), 0, s, A);
                hipLaunchKernelGGL(k_final1024, dim3(1), dim3(1024), 0, s, norm_part, nblk, 2, 1, norm_out);
            } else if (pro) {
                switch (l.read_a) { case 0: SWEEP(0, 1); break; case 1: SWEEP(1, 1); break; default: SWEEP(2, 1); break; }
            } else {
                switch (l.read_a) { case 0: SWEEP(0, 0); break; case 1: SWEEP(1, 0); break; default: SWEEP(2, 0); break; }
            }
#undef SWEEP
            std::swap(cur, nxt);
        }
        if (cur != l.phi && !keep_alt) {   /* the result goes back to phi */
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
void mg_sweep(hipStream_t s, const DLev& l, const DBC& bc, int inh, int nsweeps, int allow_fused, int zero_first) {
    sweep_impl<double>(s, l, bc, inh, nsweeps, allow_fused, zero_first);
}
bool mg_sweep_first(hipStream_t s, const DLev& l, const DBC& bc, int inh, int allow_fused) {
    if (allow_fused != 1 || !stream_use(l, bc, allow_fused)) return false;
    sweep_impl<double>(s, l, bc, inh, 1, allow_fused, 0, nullptr, false, true);
    return true;
}
bool mg_sweep_first_norm(hipStream_t s, const DLev& l, const DBC& bc, int inh, int allow_fused, double* part, double* out,
                         bool zero) {
    if (allow_fused != 1 || !stream_use(l, bc, allow_fused) || l.read_a != 0) return false;
    sweep_impl<double>(s, l, bc, inh, 1, allow_fused, zero ? 1 : 0, nullptr, false, true, part, out);
    return true;
}
void mg_sweep_rest(hipStream_t s, const DLev& l, const DBC& bc, int inh, int nsweeps, int allow_fused) {
    if (nsweeps <= 0) {
        const long long n = (long long)l.n0 * l.n1 * l.n2;
        hipLaunchKernelGGL(k_copyT<double>, dim3((unsigned)std::min<long long>((n + 255) / 256, 4096)), dim3(256), 0, s, l.phi, l.phi2, n);
        return;
    }
    sweep_impl<double>(s, l, bc, inh, nsweeps, allow_fused, 0, nullptr, true, false);
}
void mg_sweep(hipStream_t s, const DLevF& l, const DBC& bc, int nsweeps, int zero_first, int allow_fused) {
    sweep_impl<float>(s, l, bc, 0, nsweeps, allow_fused, zero_first);
}
/* The FP32 first sweep starts from e = 0 and does not read its input, so it
 * writes phi directly (the input named is phi2, never read): the correction
 * is never the answer a convergence check might keep. */
bool mg_sweep_first(hipStream_t s, const DLevF& l, const DBC& bc, int allow_fused) {
    if (allow_fused != 1 || !stream_use(l, bc, allow_fused)) return false;
    sweep_impl<float>(s, l, bc, 0, 1, allow_fused, 1, nullptr, true, false);
    return true;
}
void mg_sweep_rest(hipStream_t s, const DLevF& l, const DBC& bc, int nsweeps, int allow_fused) {
    if (nsweeps <= 0) return;   /* the first sweep's result is in phi */
    sweep_impl<float>(s, l, bc, 0, nsweeps, allow_fused, 0, nullptr, false, false);
}
void mg_prolong_sweep(hipStream_t s, const DLev& l, const DLev& c, const DBC& bc, int inh, int nsweeps, int allow_fused) {
    sweep_impl<double>(s, l, bc, inh, nsweeps, allow_fused, 0, &c);
}
void mg_prolong_sweep(hipStream_t s, const DLevF& l, const DLevF& c, const DBC& bc, int nsweeps, int allow_fused) {
    sweep_impl<float>(s, l, bc, 0, nsweeps, allow_fused, 0, &c);
}
void mg_smooth(hipStream_t s, const DLev& l, const DBC& bc, int inh, int color) {
    const dim3 g(((l.n0 + 1) / 2 + 63) / 64, (l.n1 + 3) / 4, l.n2);
    hipLaunchKernelGGL(k_mg_smooth<double>, g, dim3(256), 0, s, l, bc, inh, color, 0);
}
/* The streaming residual serves layered levels large enough to fill the
 * chip with 64x8 tiles; the fixed-grid kernels take the rest. */
template <class T>
static bool resid_stream_ok(const DLevT<T>& l, const DBC& bc) {
    return l.layered && bc.lo[2] != BC_PER && l.n2 >= 2 && l.n0 >= 64 && l.n1 >= 8 &&
           (l.read_a != 1 || l.mat != nullptr) && l.stream;
}
template <class T>
static ResidArgs<T> resid_args(const DLevT<T>& l, const DBC& bc, int inh, int* nblk) {
    ResidArgs<T> A;
    A.n0 = l.n0; A.n1 = l.n1; A.n2 = l.n2; A.msx = l.msx; A.msy = l.msy;
    A.phi = l.phi; A.b = l.b; A.a = l.a; A.mat = l.mat;
    A.bxk = l.bxk; A.byk = l.byk; A.bzk = l.bzk;
    for (int d = 0; d < 3; ++d) A.ih2[d] = l.ih2[d];
    A.bc = bc;
    A.inh = inh;
    A.ntx = (l.n0 + RTX - 1) / RTX;
    A.nty = (l.n1 + RTY - 1) / RTY;
    A.cz = stream_chunk(A.ntx * A.nty, l.n2);
    *nblk = A.ntx * A.nty * ((l.n2 + A.cz - 1) / A.cz);
    A.remap = (*nblk % 8) == 0 ? 1 : 0;
    A.cb = nullptr; A.cphi = nullptr; A.r32 = nullptr; A.part = nullptr;
    A.zero_phi = 0;
    return A;
}
template <class T, int MODE>
static void resid_stream_launch(hipStream_t s, const DLevT<T>& l, const ResidArgs<T>& A, int nblk) {
#define RS(am) hipLaunchKernelGGL((k_mg_resid_stream<T, am, MODE>), dim3(nblk), dim3(256), 0, s, A)
    switch (l.read_a) { case 0: RS(0); break; case 1: RS(1); break; default: RS(2); break; }
#undef RS
}

template <class T>
static void resid_restrict_impl(hipStream_t s, const DLevT<T>& f, const DLevT<T>& c, const DBC& bc, int inh,
                                bool zero_cphi) {
    const int zonly = f.crs[0] == 0 && f.crs[1] == 0 && f.crs[2] == 1 && c.n0 == f.n0 && c.n1 == f.n1 && (f.n2 % 2) == 0;
    if (zonly && resid_stream_ok(f, bc)) {
        int nblk;
        ResidArgs<T> A = resid_args(f, bc, inh, &nblk);
        A.cb = c.b;
        A.cphi = zero_cphi ? c.phi : nullptr;
        resid_stream_launch<T, 1>(s, f, A, nblk);
        return;
    }
    const dim3 g = grid3(c.n0, c.n1, c.n2);
    const int code = f.crs[0] | (f.crs[1] << 1) | (f.crs[2] << 2);
    if (f.layered) {
#define RL(cx, cy, cz) hipLaunchKernelGGL((k_mg_resid_restrict_lay<T, cx, cy, cz>), g, dim3(256), 0, s, f, c, bc, inh)
        switch (code) {
        case 0: RL(0, 0, 0); break;
        case 1: RL(1, 0, 0); break;
        case 2: RL(0, 1, 0); break;
        case 3: RL(1, 1, 0); break;
        case 4: RL(0, 0, 1); break;
        case 5: RL(1, 0, 1); break;
        case 6: RL(0, 1, 1); break;
        default: RL(1, 1, 1); break;
        }
#undef RL
        return;
    }
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
void mg_resid_restrict(hipStream_t s, const DLev& f, const DLev& c, const DBC& bc, int inh, bool zero_cphi) {
    resid_restrict_impl<double>(s, f, c, bc, inh, zero_cphi);
}
void mg_resid_restrict(hipStream_t s, const DLevF& f, const DLevF& c, const DBC& bc, bool zero_cphi) {
    resid_restrict_impl<float>(s, f, c, bc, 0, zero_cphi);
}
/* Block shape of one leg of level l (coarser level c): among 16x16, 16x8,
 * 8x8, 8x4 and 4x4 blocks that tile the level and whose LDS fits (down:
 * phi with halo 2nu+1 and the rhs [and Helmholtz term] on rings up to 2nu;
 * up: phi with halo 2nu, the coarse parents and the rhs on rings up to
 * 2nu-1), the one giving the most workgroups without exceeding the compute
 * units (these levels are latency bound: smaller blocks mean fewer cells per
 * workgroup and phase), else the fewest workgroups. */
template <class T>
static bool block_shape(const DLevT<T>& l, const DLevT<T>& c, const DBC& bc, int nu, bool down, int* bx, int* by,
                        size_t* lds) {
    if (!l.layered || l.read_a == 1 || bc.lo[2] == BC_PER || l.n2 < 1 || l.n2 > BLK_MAXZ || nu < 1) return false;
    const size_t lmax = tail_lds_max();
    const size_t cap = lmax > 8192 ? lmax - 8192 : 0;
    static int ncu = 0;
    if (!ncu) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu <= 0)
            ncu = 256;
    }
    static const int cand[5][2] = {{16, 16}, {16, 8}, {8, 8}, {8, 4}, {4, 4}};
    const int ncoef = l.read_a ? 2 : 1;
    int best = -1;
    long long best_n = 0;
    size_t best_lds = 0;
    for (int q = 0; q < 5; ++q) {
        const int X = cand[q][0], Y = cand[q][1];
        if (l.n0 % X || l.n1 % Y) continue;
        const int H = down ? 2 * nu + 1 : 2 * nu;
        size_t v = (size_t)(X + 2 * H) * (Y + 2 * H) * l.n2 + (size_t)ncoef * (X + 2 * H - 2) * (Y + 2 * H - 2) * l.n2;
        if (!down) {
            const int HC = H / 2 + 1;
            const size_t clx = l.crs[0] ? X / 2 + 2 * HC : X + 2 * H, cly = l.crs[1] ? Y / 2 + 2 * HC : Y + 2 * H;
            v += clx * cly * c.n2;
        }
        if (v * sizeof(T) > cap) continue;
        const long long n = (long long)(l.n0 / X) * (l.n1 / Y);
        const bool better = best < 0 || (n <= ncu && (best_n > ncu || n > best_n)) || (n > ncu && best_n > ncu && n < best_n);
        if (better) { best = q; best_n = n; best_lds = v * sizeof(T); }
    }
    if (best < 0) return false;
    *bx = cand[best][0]; *by = cand[best][1]; *lds = best_lds;
    return true;
}
template <class T, int CX, int CY, int CZ>
static void block_launch(hipStream_t s, const BlkArgs<T>& A, bool down, size_t lds) {
    const int nblk = (A.l.n0 / A.bx) * (A.l.n1 / A.by);
    if (down) {
        static size_t set = 0;
        if (lds > set) {
            (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_mg_block_down<T, CX, CY, CZ>),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
            set = lds;
        }
        hipLaunchKernelGGL((k_mg_block_down<T, CX, CY, CZ>), dim3(nblk), dim3(BLK_THREADS), lds, s, A);
    } else {
        static size_t set = 0;
        if (lds > set) {
            (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_mg_block_up<T, CX, CY, CZ>),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
            set = lds;
        }
        hipLaunchKernelGGL((k_mg_block_up<T, CX, CY, CZ>), dim3(nblk), dim3(BLK_THREADS), lds, s, A);
    }
}
template <class T>
static bool block_impl(hipStream_t s, const DLevT<T>& l, const DLevT<T>& c, const DBC& bc, int nu1, int nu2, int leg,
                       T* out = nullptr) {
    BlkArgs<T> A;
    A.out = leg == 0 ? l.phi : out;
    size_t lds;
    if (leg < 0) {
        int bx, by;
        return block_shape(l, c, bc, nu1, true, &bx, &by, &lds) && block_shape(l, c, bc, nu2, false, &bx, &by, &lds);
    }
    const bool down = leg == 0;
    A.nu = down ? nu1 : nu2;
    if (!block_shape(l, c, bc, A.nu, down, &A.bx, &A.by, &lds)) return false;
    A.l = l; A.c = c; A.bc = bc;
    const int code = l.crs[0] | (l.crs[1] << 1) | (l.crs[2] << 2);
    switch (code) {
    case 0: block_launch<T, 0, 0, 0>(s, A, down, lds); break;
    case 1: block_launch<T, 1, 0, 0>(s, A, down, lds); break;
    case 2: block_launch<T, 0, 1, 0>(s, A, down, lds); break;
    case 3: block_launch<T, 1, 1, 0>(s, A, down, lds); break;
    case 4: block_launch<T, 0, 0, 1>(s, A, down, lds); break;
    case 5: block_launch<T, 1, 0, 1>(s, A, down, lds); break;
    case 6: block_launch<T, 0, 1, 1>(s, A, down, lds); break;
    default: block_launch<T, 1, 1, 1>(s, A, down, lds); break;
    }
    return true;
}
bool mg_block_ok(const DLev& l, const DLev& c, const DBC& bc, int nu1, int nu2) {
    return block_impl<double>(nullptr, l, c, bc, nu1, nu2, -1);
}
bool mg_block_ok(const DLevF& l, const DLevF& c, const DBC& bc, int nu1, int nu2) {
    return block_impl<float>(nullptr, l, c, bc, nu1, nu2, -1);
}
void mg_block_down(hipStream_t s, const DLev& l, const DLev& c, const DBC& bc, int nu1, int nu2) {
    block_impl<double>(s, l, c, bc, nu1, nu2, 0);
}
void mg_block_down(hipStream_t s, const DLevF& l, const DLevF& c, const DBC& bc, int nu1, int nu2) {
    block_impl<float>(s, l, c, bc, nu1, nu2, 0);
}
void mg_block_up(hipStream_t s, const DLev& l, const DLev& c, const DBC& bc, int nu1, int nu2, double* out) {
    block_impl<double>(s, l, c, bc, nu1, nu2, 1, out);
}
void mg_block_up(hipStream_t s, const DLevF& l, const DLevF& c, const DBC& bc, int nu1, int nu2, float* out) {
    block_impl<float>(s, l, c, bc, nu1, nu2, 1, out);
}
void mg_prolong(hipStream_t s, const DLev& f, const DLev& c, const DBC& bc) { prolong_impl<double>(s, f, c, bc); }
void mg_prolong(hipStream_t s, const DLevF& f, const DLevF& c, const DBC& bc) { prolong_impl<float>(s, f, c, bc); }
void mg_restrict_cells(hipStream_t s, const DLev& f, const DLev& c, const double* src, double* dst) {
    hipLaunchKernelGGL(k_mg_restrict_cells<double>, grid3(c.n0, c.n1, c.n2), dim3(256), 0, s, f, c, src, dst);
}
void mg_restrict_cells(hipStream_t s, const DLevF& f, const DLevF& c, const float* src, float* dst) {
    hipLaunchKernelGGL(k_mg_restrict_cells<float>, grid3(c.n0, c.n1, c.n2), dim3(256), 0, s, f, c, src, dst);
}

static constexpr int RESID_ZC = 4;
static constexpr int RESID_NB = 1024;   /* workgroups of the residual kernels */
int resid_norm_blocks(const DLev&) { return RESID_NB; }
void mg_resid_norm(hipStream_t s, const DLev& l, const DBC& bc, int inh, double* r, double* part, double* out) {
    if (!r && resid_stream_ok(l, bc)) {
        int nblk;
        ResidArgs<double> A = resid_args(l, bc, inh, &nblk);
        if (nblk <= RESID_NB) {
            A.part = part;
            resid_stream_launch<double, 2>(s, l, A, nblk);
            hipLaunchKernelGGL(k_final1024, dim3(1), dim3(1024), 0, s, part, nblk, 2, 1, out);
            return;
        }
    }
    if (r) hipLaunchKernelGGL((k_mg_resid_max<1, RESID_ZC>), dim3(RESID_NB), dim3(256), 0, s, l, bc, inh, r, nullptr, part);
    else hipLaunchKernelGGL((k_mg_resid_max<0, RESID_ZC>), dim3(RESID_NB), dim3(256), 0, s, l, bc, inh, nullptr, nullptr, part);
    hipLaunchKernelGGL(k_final1024, dim3(1), dim3(1024), 0, s, part, RESID_NB, 2, 1, out);
}
void add_f2d(hipStream_t s, double* phi, const float* e, long long n, bool set) {
    unsigned b = (unsigned)((n + 255) / 256); if (b > 8192) b = 8192; if (b == 0) b = 1;
    hipLaunchKernelGGL(k_add_f2d, dim3(b), dim3(256), 0, s, phi, e, n, set ? 1 : 0);
}
void resid_to32(hipStream_t s, const DLev& l, const DBC& bc, float* r32, double* part, double* out, bool zero) {
    if (resid_stream_ok(l, bc)) {
        int nblk;
        ResidArgs<double> A = resid_args(l, bc, 1, &nblk);
        if (nblk <= RESID_NB) {
            A.part = part;
            A.r32 = r32;
            A.zero_phi = zero ? 1 : 0;
            resid_stream_launch<double, 3>(s, l, A, nblk);
            hipLaunchKernelGGL(k_final1024, dim3(1), dim3(1024), 0, s, part, nblk, 2, 1, out);
            return;
        }
    }
    if (zero) {
        const long long n = (long long)l.n0 * l.n1 * l.n2;
        fill(s, l.phi, n, 0.0);
    }
    hipLaunchKernelGGL((k_mg_resid_max<2, RESID_ZC>), dim3(RESID_NB), dim3(256), 0, s, l, bc, 1, nullptr, r32, part);
    hipLaunchKernelGGL(k_final1024, dim3(1), dim3(1024), 0, s, part, RESID_NB, 2, 1, out);
}
size_t tail_level_values(const DLev& l) { return fxd_tail_level_values(l); }
/* largest dynamic LDS a workgroup may use on the current device */
size_t tail_lds_max() {
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 65536;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || v <= 0) return 65536;
    return (size_t)v;
}
template <class T>
static void tail_impl(hipStream_t s, const DLevT<T>* levs, int nl, const DBC& bc, int read_a, int nu1, int nu2, T* work,
                      double tol, int maxit, int* iters, int ts, size_t lds_bytes) {
    static size_t attr_set = 0;      /* largest dynamic LDS requested so far for this instantiation */
    if (lds_bytes > attr_set) {
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_mg_tail<T>), hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)lds_bytes);
        attr_set = lds_bytes;
    }
    hipLaunchKernelGGL(k_mg_tail<T>, dim3(1), dim3(1024), lds_bytes, s, levs, nl, bc, read_a, nu1, nu2, work, tol, maxit,
                       iters, ts);
}
void mg_tail(hipStream_t s, const DLev* levs, int nl, const DBC& bc, int read_a, int nu1, int nu2, double* work,
             double tol, int maxit, int* iters, int ts, size_t lds_bytes) {
    tail_impl<double>(s, levs, nl, bc, read_a, nu1, nu2, work, tol, maxit, iters, ts, lds_bytes);
}
void mg_tail(hipStream_t s, const DLevF* levs, int nl, const DBC& bc, int read_a, int nu1, int nu2, float* work,
             double tol, int maxit, int* iters, int ts, size_t lds_bytes) {
    tail_impl<float>(s, levs, nl, bc, read_a, nu1, nu2, work, tol, maxit, iters, ts, lds_bytes);
}
template <class T>
static void tail_lay_impl(hipStream_t s, const TailPlan<T>& P, const DBC& bc, int read_a, int nu1, int nu2, T* work,
                          double tol, int maxit, int* iters) {
    const size_t lds = tail_lean_values(P.n, P.nl, read_a) * sizeof(T);
    static size_t attr_set = 0;
    if (lds > attr_set) {
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_mg_tail_lay<T>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        attr_set = lds;
    }
    hipLaunchKernelGGL(k_mg_tail_lay<T>, dim3(1), dim3(1024), lds, s, P, bc, read_a, nu1, nu2, work, tol, maxit, iters);
}
template <class T>
static void spec_solve_impl(hipStream_t s, const SpecTail<T>& P) {
    const size_t lds = spec_plane_values(P.n0, P.n1) * sizeof(T);
    static size_t attr_set = 0;
    if (lds > attr_set) {
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_spec_fwd<T>), hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_spec_inv<T>), hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        attr_set = lds;
    }
    const int np = P.n0 * P.n1;
    const dim3 g((P.n0 + SPEC_CB - 1) / SPEC_CB, P.n2);
    hipLaunchKernelGGL(k_spec_fwd<T>, g, dim3(SPEC_THR), lds, s, P);
    if (P.n2 <= 16) hipLaunchKernelGGL((k_spec_tri<T, 16>), dim3((np + 63) / 64), dim3(64), 0, s, P);
    else hipLaunchKernelGGL((k_spec_tri<T, 0>), dim3((np + 63) / 64), dim3(64), 0, s, P);
    hipLaunchKernelGGL(k_spec_inv<T>, g, dim3(SPEC_THR), lds, s, P);
}
void mg_spec_solve(hipStream_t s, const SpecTail<double>& P) { spec_solve_impl<double>(s, P); }
void mg_spec_solve(hipStream_t s, const SpecTail<float>& P) { spec_solve_impl<float>(s, P); }
void mg_tail_lay(hipStream_t s, const TailPlan<double>& P, const DBC& bc, int read_a, int nu1, int nu2, double* work,
                 double tol, int maxit, int* iters) {
    tail_lay_impl<double>(s, P, bc, read_a, nu1, nu2, work, tol, maxit, iters);
}
void mg_tail_lay(hipStream_t s, const TailPlan<float>& P, const DBC& bc, int read_a, int nu1, int nu2, float* work,
                 double tol, int maxit, int* iters) {
    tail_lay_impl<float>(s, P, bc, read_a, nu1, nu2, work, tol, maxit, iters);
}
size_t tail_lay_lds(const int (*n)[3], int nl, int read_a, size_t elem) { return tail_lean_values(n, nl, read_a) * elem; }
void bottom_factor(hipStream_t s, const DLev& l, const DBC& bc, double* lu, int* perm) {
    hipLaunchKernelGGL(k_bottom_factor<double>, dim3(1), dim3(1024), 0, s, l, bc, lu, perm);
}
void bottom_factor(hipStream_t s, const DLevF& l, const DBC& bc, float* lu, int* perm) {
    hipLaunchKernelGGL(k_bottom_factor<float>, dim3(1), dim3(1024), 0, s, l, bc, lu, perm);
}
void d2f(hipStream_t s, float* d, const double* src, long long n) {
    unsigned b = (unsigned)((n + 255) / 256); if (b > 4096) b = 4096; if (b == 0) b = 1;
    hipLaunchKernelGGL(k_d2f, dim3(b), dim3(256), 0, s, d, src, n);
}
void final1024(hipStream_t s, const double* part, int nb, int nq, int is_max, double* out) {
    hipLaunchKernelGGL(k_final1024, dim3(1), dim3(1024), 0, s, part, nb, nq, is_max, out);
}
void bottom_cg(hipStream_t s, const DLev& l, const DBC& bc, const double* rhs, double* e, double* work, double tol, int maxit, int* iters) {
    hipLaunchKernelGGL(k_bottom<double>, dim3(1), dim3(1024), 0, s, l, bc, rhs, e, work, tol, maxit, iters);
}
void bottom_cg(hipStream_t s, const DLevF& l, const DBC& bc, const float* rhs, float* e, float* work, double tol, int maxit, int* iters) {
    hipLaunchKernelGGL(k_bottom<float>, dim3(1), dim3(1024), 0, s, l, bc, rhs, e, work, tol, maxit, iters);
}

}  // namespace fxk
