/*
 * Batching form: many independent MFIM devices of uneven sizes advanced
 * together.  Each stage of a time step (Poisson right-hand side, the Poisson
 * solve, the TDGL update) is ONE persistent launch that walks a tile list
 * covering every device; a whole step is captured in a hipGraph and replayed;
 * the tile shapes and grid sizes of the stages are chosen by measurement
 * (fx_batch_autotune).  There is no counterpart in the reference, which runs
 * one device per process (main.cpp:48-623).
 *
 * The devices are laterally periodic layered stacks (an insulator under a
 * ferroelectric, or any stack whose permittivity depends on z only).  For
 * them the discrete Poisson operator of the hot path -- face-averaged
 * permittivity, 7-point stencil, Dirichlet faces in z with the ghost
 * 2 phi_b - phi_0 -- is diagonalised exactly by the real Fourier basis of
 * each periodic lateral direction: mode (mx, my) leaves a tridiagonal system
 * in z with diagonal eps_k (mu_x/dx^2 + mu_y/dy^2) + z couplings.  The batch
 * solves it directly (lateral transforms as small dense products, a Thomas
 * solve per mode), so every step is the same fixed sequence of kernels with
 * no convergence test: graph-capturable, and exact to rounding.  The TDGL and
 * right-hand-side stages use the same device stencil functions as the
 * single-device path.  FP32 mode: the solve runs in FP32 and one FP64
 * residual correction (r = b - A phi in FP64, a second FP32 solve) follows.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/ferrox_mi.h"
#include "fx_device.h"
#include "fx_kernels.h"

using namespace fxd;

int fxi_export(fx_ctx* x, std::vector<uint8_t>& mat, int& msx, int& msy, fx_inputs& in, double ih2[3],
               std::vector<double>& epsk, std::vector<double>& bzk, double bv[6]);

namespace {

/* per-device descriptor, in device memory */
struct BDev {
    Geom g;
    PBC pbc;
    const uint8_t* mat;
    int msx, msy;
    long long off;           /* first cell of this device in the batch arrays */
    int bclo[3], bchi[3];
    double bv[6];            /* Dirichlet values x,y,z lo then hi */
    int ghost_face;
    double alpha, beta, gamma, g11, g44m, G, dt;
    const double* epsk;      /* lateral face permittivity of plane k (nz) */
    const double* bzk;       /* z-face permittivity (nz + 1) */
    double ih2[3];
    int bx, by;              /* basis index of n_x and n_y */
};

/* orthonormal real Fourier basis of an n-point periodic direction */
struct BBasis {
    int n;
    const double* Q;         /* n x n, column m = basis vector m */
    const double* QT;        /* its transpose */
    const float* Q32;
    const float* QT32;
    const double* mu;        /* eigenvalue of (2 phi_i - phi_i-1 - phi_i+1) for column m */
};

}  // namespace

/* The real Fourier basis of an n-point periodic direction (library-internal
 * and exported for tests): column 0 constant, then cos/sin pairs of frequency
 * m, then the alternating vector when n is even; mu[m] = 2 - 2 cos(2 pi f/n). */
extern "C" int fx_batch_basis(int n, double* Q, double* mu) {
    if (n < 1 || !Q || !mu) return FX_EINVAL;
    const double pi = 3.14159265358979323846;
    int col = 0;
    auto put = [&](int f, int kind) {      /* kind 0 constant, 1 cos, 2 sin, 3 alternating */
        const double sc = (kind == 1 || kind == 2) ? std::sqrt(2.0 / n) : 1.0 / std::sqrt((double)n);
        for (int i = 0; i < n; ++i) {
            double v;
            if (kind == 0) v = sc;
            else if (kind == 3) v = (i & 1) ? -sc : sc;
            else {
                const double a = 2.0 * pi * (double)f * (double)i / n;
                v = sc * (kind == 1 ? std::cos(a) : std::sin(a));
            }
            Q[(size_t)i * n + col] = v;
        }
        mu[col] = 2.0 - 2.0 * std::cos(2.0 * pi * f / n);
        ++col;
    };
    put(0, 0);
    for (int f = 1; 2 * f < n; ++f) { put(f, 1); put(f, 2); }
    if (n % 2 == 0) put(n / 2, 3);
    return col == n ? FX_OK : FX_EINVAL;
}

/* ------------------------------- kernels ---------------------------------- */
namespace {

__device__ __forceinline__ int4 tile_at(const int4* tiles, int t) { return tiles[t]; }

/* Poisson right-hand side of every device, with the Dirichlet terms of the
 * z faces folded in (the solve is then homogeneous): FE cells -DPDz(P), as
 * k_rhs_newton for an unrotated scalar P. */
template <int TX>
__global__ __launch_bounds__(256) void k_b_rhs(const BDev* __restrict__ devs, const int4* __restrict__ tiles, int ntiles,
                                               const double* __restrict__ P, double* __restrict__ rhs) {
    constexpr int TY = 256 / TX;
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const int4 tl = tile_at(tiles, t);
        const BDev& D = devs[tl.x];
        const int i = tl.y * TX + (threadIdx.x % TX), j = tl.z * TY + (threadIdx.x / TX), k = tl.w;
        const int nx = D.g.n[0], ny = D.g.n[1], nz = D.g.n[2];
        if (i >= nx || j >= ny) continue;
        const MView M{D.mat, D.msx, D.msy};
        const PView F{P + D.off, nx, ny, nz, D.g.per[0], D.g.per[1], D.g.per[2]};
        double r = 0.0;
        if (M.code(i, j, k) == 0) {
            double x = 0.0;
            x += 1.0 * dpd<2>(F, M, i, j, k, D.g.dx, D.pbc);
            r = -x;
        }
        if (k == 0 && D.bclo[2] == BC_DIR) r += 2.0 * (D.bzk[0] * D.ih2[2]) * D.bv[2];
        if (k == nz - 1 && D.bchi[2] == BC_DIR) r += 2.0 * (D.bzk[nz] * D.ih2[2]) * D.bv[5];
        rhs[D.off + i + (long long)nx * (j + (long long)ny * k)] = r;
    }
}

/* out = in transformed along x (axis 0) or y (axis 1): for each vector v of
 * the direction, out_m = sum_q in_q M[q][m] with M = Q (forward: coefficients
 * in the basis) or Q^T (inverse).  A tile holds R vectors: R rows (x) or R
 * consecutive i of one plane (y), staged in LDS. */
template <class T, int R, int AXIS, int INV>
__global__ __launch_bounds__(256) void k_b_axis(const BDev* __restrict__ devs, const BBasis* __restrict__ bases,
                                                const int4* __restrict__ tiles, int ntiles, const T* __restrict__ in,
                                                T* __restrict__ out) {
    __shared__ T sv[R * 128];
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const int4 tl = tile_at(tiles, t);
        const BDev& D = devs[tl.x];
        const int nx = D.g.n[0], ny = D.g.n[1];
        const BBasis& B = bases[AXIS == 0 ? D.bx : D.by];
        const int n = B.n;
        const T* M;
        if (sizeof(T) == 8) M = reinterpret_cast<const T*>(INV ? B.QT : B.Q);
        else M = reinterpret_cast<const T*>(INV ? B.QT32 : B.Q32);
        /* vector v of the tile, element q: global index; vectors past the
         * end of the device (a partial last tile) are skipped */
        const int nvec = AXIS == 0 ? ny * D.g.n[2] - tl.y : nx - tl.y;
        auto gidx = [&](int v, int q) -> long long {
            if (AXIS == 0) { const long long row = (long long)tl.y + v; return D.off + q + (long long)nx * row; }
            const int k = tl.z, i = tl.y + v;
            return D.off + i + (long long)nx * (q + (long long)ny * k);
        };
        __syncthreads();
        for (int e = threadIdx.x; e < R * n; e += 256) {
            const int v = AXIS == 0 ? e / n : e % R, q = AXIS == 0 ? e % n : e / R;
            sv[v * n + q] = v < nvec ? in[gidx(v, q)] : T(0);
        }
        __syncthreads();
        for (int e = threadIdx.x; e < R * n; e += 256) {
            const int v = AXIS == 0 ? e / n : e % R, m = AXIS == 0 ? e % n : e / R;
            if (v >= nvec) continue;
            T s = T(0);
            const T* x = sv + v * n;
            for (int q = 0; q < n; ++q) s += x[q] * M[q * n + m];
            out[gidx(v, m)] = s;
        }
    }
}

/* Thomas solve in z of mode (mx, my) of every device: tile = (device, first
 * mode, -, -), 256 modes.  h holds the transformed right-hand side on entry
 * and the transformed solution on exit; cp is scratch of the same layout. */
template <class T>
__global__ __launch_bounds__(256) void k_b_tridiag(const BDev* __restrict__ devs, const BBasis* __restrict__ bases,
                                                   const int4* __restrict__ tiles, int ntiles, T* __restrict__ h,
                                                   T* __restrict__ cp) {
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const int4 tl = tile_at(tiles, t);
        const BDev& D = devs[tl.x];
        const int nx = D.g.n[0], ny = D.g.n[1], nz = D.g.n[2];
        const int mode = tl.y + threadIdx.x;
        if (mode >= nx * ny) continue;
        const int mx = mode % nx, my = mode / nx;
        const T lam = T(bases[D.bx].mu[mx] * D.ih2[0] + bases[D.by].mu[my] * D.ih2[1]);
        const long long plane = (long long)nx * ny;
        T* hc = h + D.off + mode;
        T* cc = cp + D.off + mode;
        T cprev = T(0), dprev = T(0);
        for (int k = 0; k < nz; ++k) {
            const T wl = T(D.bzk[k] * D.ih2[2]), wh = T(D.bzk[k + 1] * D.ih2[2]);
            T diag = T(D.epsk[k]) * lam;
            T a = T(0), c = T(0);
            if (k > 0) { diag += wl; a = -wl; }
            else if (D.bclo[2] == BC_DIR) diag += T(2) * wl;
            if (k < nz - 1) { diag += wh; c = -wh; }
            else if (D.bchi[2] == BC_DIR) diag += T(2) * wh;
            const T rden = T(1) / (diag - a * cprev);
            const T cq = c * rden;
            const T dq = (hc[k * plane] - a * dprev) * rden;
            cc[k * plane] = cq;
            hc[k * plane] = dq;
            cprev = cq;
            dprev = dq;
        }
        T xnext = T(0);
        for (int k = nz - 1; k >= 0; --k) {
            const T xv = hc[k * plane] - cc[k * plane] * xnext;
            hc[k * plane] = xv;
            xnext = xv;
        }
    }
}

/* TDGL right-hand side and forward-Euler update of every device (the scalar
 * unrotated branch of k_tdgl, same stencils and summation order). */
template <int TX>
__global__ __launch_bounds__(256) void k_b_tdgl(const BDev* __restrict__ devs, const int4* __restrict__ tiles, int ntiles,
                                                const double* __restrict__ P, const double* __restrict__ phi,
                                                double* __restrict__ Pout) {
    constexpr int TY = 256 / TX;
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const int4 tl = tile_at(tiles, t);
        const BDev& D = devs[tl.x];
        const int i = tl.y * TX + (threadIdx.x % TX), j = tl.z * TY + (threadIdx.x / TX), k = tl.w;
        const int nx = D.g.n[0], ny = D.g.n[1], nz = D.g.n[2];
        if (i >= nx || j >= ny) continue;
        const MView M{D.mat, D.msx, D.msy};
        const int raw = M.raw(i, j, k);
        if ((raw & 3) != 0) continue;
        const long long c = D.off + i + (long long)nx * (j + (long long)ny * k);
        const PView Fr{P + D.off, nx, ny, nz, D.g.per[0], D.g.per[1], D.g.per[2]};
        PhiView pv;
        pv.p = phi + D.off;
        for (int d = 0; d < 3; ++d) { pv.n[d] = D.g.n[d]; pv.lo[d] = D.bclo[d]; pv.hi[d] = D.bchi[d]; }
        for (int q = 0; q < 6; ++q) pv.v[q] = D.bv[q];
        pv.ghost_face = D.ghost_face;
        const double Gam = D.G;
        /* FE cells whose stencils are all the centred ones (MAT_INTERIOR_FE)
         * read their neighbours directly, as k_tdgl does */
        const bool inner = (raw & MAT_INTERIOR_FE) != 0;
        const long long pl = (long long)nx * ny;
        double fx = 0.0, fy = 0.0, fz;
        if (inner) fz = (phi[c + pl] - phi[c - pl]) / (2. * D.g.dx[2]);
        else grad_phi(pv, D.g, i, j, k, fx, fy, fz);
        double sE = 0.0;
        sE += 1.0 * fz;
        const double E2 = -sE;
        const double r = P[c];
        const double r2 = r * r;
        const double dF = D.alpha * r + D.beta * (r2 * r) + D.gamma * (r2 * r2 * r);
        double L = -1.0 * Gam * dF;
        double Hxx, Hyy, Hzz;
        if (inner) {
            const NbrOff o(i, j, nx, ny);
            const double h0 = D.g.dx[0], h1 = D.g.dx[1], h2 = D.g.dx[2];
            Hxx = (P[c + o.xp] - 2. * r + P[c + o.xm]) / (h0 * h0);
            Hyy = (P[c + o.yp] - 2. * r + P[c + o.ym]) / (h1 * h1);
            Hzz = (P[c + o.zp] - 2. * r + P[c + o.zm]) / (h2 * h2);
        } else {
            Hxx = ddpd<0>(Fr, M, i, j, k, D.g.dx, D.pbc);
            Hyy = ddpd<1>(Fr, M, i, j, k, D.g.dx, D.pbc);
            Hzz = ddpd<2>(Fr, M, i, j, k, D.g.dx, D.pbc);
        }
        const double gr = -D.g11 * (0.0 * Hxx + 0.0 * Hyy + 1.0 * Hzz)
                          - D.g44m * (1.0 * Hxx + 0.0 * Hyy + 0.0 * Hzz)
                          - D.g44m * (0.0 * Hxx + 1.0 * Hyy + 0.0 * Hzz);
        double G = -1.0 * Gam * gr;
        double El = -1.0 * Gam * (-E2);
        if (raw & 4) { L = 0.0; G = 0.0; El = 0.0; }
        double g = 0.0;
        g += L; g += G; g += El;
        Pout[c] = r + D.dt * g;
    }
}

/* FP64 residual r = b - A phi of every device (homogeneous operator; b has
 * the Dirichlet terms), written as FP32 for the correction solve. */
template <int TX>
__global__ __launch_bounds__(256) void k_b_resid(const BDev* __restrict__ devs, const int4* __restrict__ tiles, int ntiles,
                                                 const double* __restrict__ b, const double* __restrict__ phi,
                                                 float* __restrict__ r32) {
    constexpr int TY = 256 / TX;
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const int4 tl = tile_at(tiles, t);
        const BDev& D = devs[tl.x];
        const int i = tl.y * TX + (threadIdx.x % TX), j = tl.z * TY + (threadIdx.x / TX), k = tl.w;
        const int nx = D.g.n[0], ny = D.g.n[1], nz = D.g.n[2];
        if (i >= nx || j >= ny) continue;
        const double* p = phi + D.off;
        const long long c = i + (long long)nx * (j + (long long)ny * k);
        const double wx = D.epsk[k] * D.ih2[0], wy = D.epsk[k] * D.ih2[1];
        const double wl = D.bzk[k] * D.ih2[2], wh = D.bzk[k + 1] * D.ih2[2];
        const int im = i > 0 ? i - 1 : nx - 1, ip = i < nx - 1 ? i + 1 : 0;
        const int jm = j > 0 ? j - 1 : ny - 1, jp = j < ny - 1 ? j + 1 : 0;
        const long long rowj = (long long)nx * (j + (long long)ny * k);
        double Dg = 2.0 * wx + 2.0 * wy;
        double S = wx * (p[im + rowj] + p[ip + rowj]) +
                   wy * (p[i + (long long)nx * (jm + (long long)ny * k)] + p[i + (long long)nx * (jp + (long long)ny * k)]);
        if (k > 0) { Dg += wl; S += wl * p[c - (long long)nx * ny]; } else if (D.bclo[2] == BC_DIR) Dg += 2.0 * wl;
        if (k < nz - 1) { Dg += wh; S += wh * p[c + (long long)nx * ny]; } else if (D.bchi[2] == BC_DIR) Dg += 2.0 * wh;
        r32[D.off + c] = (float)(b[D.off + c] - (Dg * p[c] - S));
    }
}

__global__ void k_b_d2f(float* __restrict__ d, const double* __restrict__ s, long long n) {
    for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < n; c += (long long)gridDim.x * blockDim.x) d[c] = (float)s[c];
}
__global__ void k_b_f2d(double* __restrict__ d, const float* __restrict__ s, long long n, int add) {
    for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < n; c += (long long)gridDim.x * blockDim.x)
        d[c] = add ? d[c] + (double)s[c] : (double)s[c];
}

constexpr int AXR = 16;        /* vectors per transform tile */

/* The lateral transforms on the matrix cores: for the vectors of a 16-wide
 * tile (16 rows j..j+15 of a device for AXIS 0; 16 consecutive i of plane k
 * for AXIS 1) out_m = sum_q in_q M[q][m], as the 16x16 blocks
 * C[m][j] = sum_q M[q][m] in_j[q], K = n in steps of 4.  Tile = (device,
 * first vector, group of four 16-wide m blocks, plane); the tile's vectors
 * are staged in LDS once with coalesced loads (as [q][j], padded to 17 so
 * that the staging writes and the fragment reads spread over the banks), and
 * each of the four wavefronts takes one m block.  AXIS 0 writes its result
 * back through the same LDS so that whole rows are stored.  Every lateral
 * extent is a multiple of 16 (fx_batch_create checks it; otherwise k_b_axis
 * runs). */
constexpr int MF_LD = 17;
template <class T, int AXIS, int INV>
__global__ __launch_bounds__(256) void k_b_axis_mfma(const BDev* __restrict__ devs, const BBasis* __restrict__ bases,
                                                     const int4* __restrict__ tiles, int ntiles, const T* __restrict__ in,
                                                     T* __restrict__ out) {
    using MF = Mfma16<T>;
    __shared__ T sv[128 * MF_LD];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int li = lane & 15, lk = lane >> 4;
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const int4 tl = tile_at(tiles, t);
        const BDev& D = devs[tl.x];
        const int nx = D.g.n[0], ny = D.g.n[1];
        const BBasis& B = bases[AXIS == 0 ? D.bx : D.by];
        const int n = B.n;
        const int m0 = (tl.z * 4 + wave) * 16;
        const T* M;
        if (sizeof(T) == 8) M = reinterpret_cast<const T*>(INV ? B.QT : B.Q);
        else M = reinterpret_cast<const T*>(INV ? B.QT32 : B.Q32);
        /* element q of vector j of the tile */
        auto gidx = [&](int j, int q) -> long long {
            return AXIS == 0 ? D.off + (long long)nx * (tl.y + j) + q : D.off + (tl.y + j) + (long long)nx * (q + (long long)ny * tl.w);
        };
        __syncthreads();                             /* the previous tile is done with sv */
        for (int e = threadIdx.x; e < 16 * n; e += 256) {
            int j, q;
            if (AXIS == 0) { j = e / n; q = e - j * n; }   /* rows are contiguous */
            else { j = e & 15; q = e >> 4; }              /* the 16 i of a row are contiguous */
            sv[q * MF_LD + j] = in[gidx(j, q)];
        }
        __syncthreads();
        typename MF::v4 acc = {T(0), T(0), T(0), T(0)};
        if (m0 < n) {                                /* uniform in the wavefront */
#pragma unroll 4
            for (int k0 = 0; k0 < n; k0 += 4) {
                const int q = k0 + lk;
                acc = MF::mma(M[(long long)q * n + m0 + li], sv[q * MF_LD + li], acc);
            }
        }
        if (AXIS == 1) {
            if (m0 < n) {
#pragma unroll
                for (int v = 0; v < 4; ++v) out[gidx(li, m0 + MF::row(lane, v))] = acc[v];
            }
        } else {
            __syncthreads();                         /* every wavefront has read sv */
            if (m0 < n) {
#pragma unroll
                for (int v = 0; v < 4; ++v) sv[(m0 + MF::row(lane, v)) * MF_LD + li] = acc[v];
            }
            __syncthreads();
            const int mlo = tl.z * 64, w = min(n, mlo + 64) - mlo;
            for (int e = threadIdx.x; e < 16 * w; e += 256) {
                const int j = e / w, m = mlo + (e - j * w);
                out[gidx(j, m)] = sv[m * MF_LD + j];
            }
        }
    }
}

}  // namespace

/* ------------------------------- host side -------------------------------- */
struct fx_batch {
    int nd = 0;
    hipStream_t s = nullptr;
    std::vector<fx_inputs> in;
    std::vector<BDev> hdev;
    std::vector<long long> off;
    long long Ntot = 0;
    BDev* ddev = nullptr;
    BBasis* dbas = nullptr;
    std::vector<void*> allocs;
    double *P[2] = {nullptr, nullptr}, *phi = nullptr, *rhs = nullptr, *h1 = nullptr, *h2 = nullptr, *cp = nullptr;
    float *f1 = nullptr, *f2 = nullptr, *fc = nullptr, *r32 = nullptr;
    int cur = 0;
    int precision = FX_PREC_FP64;
    /* tile lists */
    int4 *t_cell = nullptr, *t_x = nullptr, *t_y = nullptr, *t_mode = nullptr;
    int n_cell = 0, n_x = 0, n_y = 0, n_mode = 0;
    bool mfma = false;                   /* lateral transforms on the matrix cores (extents multiples of 16) */
    /* tuned launch shapes */
    int tx = 64;                         /* stencil tile width (height 256/tx) */
    int g_cell = 0, g_axis = 0, g_mode = 0;
    int ncu = 256;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    long long steps = 0;
    std::string err, tune_report;
};

static int bfail(fx_batch* b, int code, const std::string& m) { b->err = m; return code; }
#define BCHK(b, call) do { hipError_t e_ = (call); if (e_ != hipSuccess) return bfail(b, FX_EHIP, std::string(#call) + ": " + hipGetErrorString(e_)); } while (0)

static int balloc(fx_batch* b, void** p, size_t bytes) {
    BCHK(b, hipMalloc(p, bytes ? bytes : 8));
    b->allocs.push_back(*p);
    BCHK(b, hipMemsetAsync(*p, 0, bytes ? bytes : 8, b->s));
    return FX_OK;
}

static void free_graph(fx_batch* b) {
    if (b->gexec) { (void)hipGraphExecDestroy(b->gexec); b->gexec = nullptr; }
    if (b->graph) { (void)hipGraphDestroy(b->graph); b->graph = nullptr; }
}

/* cell tiles (device, x tile, y tile, plane) for a stencil tile width tx */
static int build_cell_tiles(fx_batch* b, int tx) {
    const int ty = 256 / tx;
    std::vector<int4> t;
    for (int d = 0; d < b->nd; ++d) {
        const Geom& g = b->hdev[d].g;
        for (int k = 0; k < g.n[2]; ++k)
            for (int y = 0; y < (g.n[1] + ty - 1) / ty; ++y)
                for (int x = 0; x < (g.n[0] + tx - 1) / tx; ++x) t.push_back(make_int4(d, x, y, k));
    }
    if (b->t_cell) { (void)hipFree(b->t_cell); b->allocs.erase(std::find(b->allocs.begin(), b->allocs.end(), (void*)b->t_cell)); }
    int rc = balloc(b, (void**)&b->t_cell, t.size() * sizeof(int4));
    if (rc) return rc;
    BCHK(b, hipMemcpy(b->t_cell, t.data(), t.size() * sizeof(int4), hipMemcpyHostToDevice));
    b->n_cell = (int)t.size();
    b->tx = tx;
    return FX_OK;
}

static int build_fixed_tiles(fx_batch* b) {
    std::vector<int4> tx, ty, tm;
    for (int d = 0; d < b->nd; ++d) {
        const Geom& g = b->hdev[d].g;
        const int nx = g.n[0], ny = g.n[1], nz = g.n[2];
        if (b->mfma) {
            /* (device, first vector, group of four 16-wide m blocks, plane) */
            for (int r = 0; r < ny * nz; r += 16)
                for (int g = 0; g < (nx + 63) / 64; ++g) tx.push_back(make_int4(d, r, g, 0));
            for (int k = 0; k < nz; ++k)
                for (int i = 0; i < nx; i += 16)
                    for (int g = 0; g < (ny + 63) / 64; ++g) ty.push_back(make_int4(d, i, g, k));
        } else {
            for (int r = 0; r < ny * nz; r += AXR) tx.push_back(make_int4(d, r, 0, 0));
            for (int k = 0; k < nz; ++k)
                for (int i = 0; i < nx; i += AXR) ty.push_back(make_int4(d, i, k, 0));
        }
        for (int m = 0; m < nx * ny; m += 256) tm.push_back(make_int4(d, m, 0, 0));
    }
    int rc;
    if ((rc = balloc(b, (void**)&b->t_x, tx.size() * sizeof(int4))) || (rc = balloc(b, (void**)&b->t_y, ty.size() * sizeof(int4))) ||
        (rc = balloc(b, (void**)&b->t_mode, tm.size() * sizeof(int4))))
        return rc;
    BCHK(b, hipMemcpy(b->t_x, tx.data(), tx.size() * sizeof(int4), hipMemcpyHostToDevice));
    BCHK(b, hipMemcpy(b->t_y, ty.data(), ty.size() * sizeof(int4), hipMemcpyHostToDevice));
    BCHK(b, hipMemcpy(b->t_mode, tm.data(), tm.size() * sizeof(int4), hipMemcpyHostToDevice));
    b->n_x = (int)tx.size(); b->n_y = (int)ty.size(); b->n_mode = (int)tm.size();
    return FX_OK;
}

static inline unsigned gsize(int want, int ntiles) { return (unsigned)std::max(1, std::min(want, ntiles)); }

static void launch_rhs(fx_batch* b, const double* P) {
    const unsigned g = gsize(b->g_cell, b->n_cell);
    switch (b->tx) {
    case 32: hipLaunchKernelGGL(k_b_rhs<32>, dim3(g), dim3(256), 0, b->s, b->ddev, b->t_cell, b->n_cell, P, b->rhs); break;
    case 128: hipLaunchKernelGGL(k_b_rhs<128>, dim3(g), dim3(256), 0, b->s, b->ddev, b->t_cell, b->n_cell, P, b->rhs); break;
    default: hipLaunchKernelGGL(k_b_rhs<64>, dim3(g), dim3(256), 0, b->s, b->ddev, b->t_cell, b->n_cell, P, b->rhs); break;
    }
}
static void launch_tdgl(fx_batch* b, const double* P, double* Pout) {
    const unsigned g = gsize(b->g_cell, b->n_cell);
    switch (b->tx) {
    case 32: hipLaunchKernelGGL(k_b_tdgl<32>, dim3(g), dim3(256), 0, b->s, b->ddev, b->t_cell, b->n_cell, P, b->phi, Pout); break;
    case 128: hipLaunchKernelGGL(k_b_tdgl<128>, dim3(g), dim3(256), 0, b->s, b->ddev, b->t_cell, b->n_cell, P, b->phi, Pout); break;
    default: hipLaunchKernelGGL(k_b_tdgl<64>, dim3(g), dim3(256), 0, b->s, b->ddev, b->t_cell, b->n_cell, P, b->phi, Pout); break;
    }
}
static void launch_resid(fx_batch* b) {
    const unsigned g = gsize(b->g_cell, b->n_cell);
    switch (b->tx) {
    case 32: hipLaunchKernelGGL(k_b_resid<32>, dim3(g), dim3(256), 0, b->s, b->ddev, b->t_cell, b->n_cell, b->rhs, b->phi, b->r32); break;
    case 128: hipLaunchKernelGGL(k_b_resid<128>, dim3(g), dim3(256), 0, b->s, b->ddev, b->t_cell, b->n_cell, b->rhs, b->phi, b->r32); break;
    default: hipLaunchKernelGGL(k_b_resid<64>, dim3(g), dim3(256), 0, b->s, b->ddev, b->t_cell, b->n_cell, b->rhs, b->phi, b->r32); break;
    }
}

/* A^-1 src -> dst, with scratch t1, t2, scr (all of type T, batch layout);
 * dst must not be t1: a transform never writes its own input (each
 * wavefront reads whole vectors that other wavefronts are writing) */
template <class T>
static void spectral_solve(fx_batch* b, const T* src, T* t1, T* t2, T* scr, T* dst) {
    const unsigned gx = gsize(b->g_axis, b->n_x), gy = gsize(b->g_axis, b->n_y), gm = gsize(b->g_mode, b->n_mode);
    if (b->mfma) {
        hipLaunchKernelGGL((k_b_axis_mfma<T, 0, 0>), dim3(gx), dim3(256), 0, b->s, b->ddev, b->dbas, b->t_x, b->n_x, src, t1);
        hipLaunchKernelGGL((k_b_axis_mfma<T, 1, 0>), dim3(gy), dim3(256), 0, b->s, b->ddev, b->dbas, b->t_y, b->n_y, t1, t2);
        hipLaunchKernelGGL(k_b_tridiag<T>, dim3(gm), dim3(256), 0, b->s, b->ddev, b->dbas, b->t_mode, b->n_mode, t2, scr);
        hipLaunchKernelGGL((k_b_axis_mfma<T, 1, 1>), dim3(gy), dim3(256), 0, b->s, b->ddev, b->dbas, b->t_y, b->n_y, t2, t1);
        hipLaunchKernelGGL((k_b_axis_mfma<T, 0, 1>), dim3(gx), dim3(256), 0, b->s, b->ddev, b->dbas, b->t_x, b->n_x, t1, dst);
        return;
    }
    hipLaunchKernelGGL((k_b_axis<T, AXR, 0, 0>), dim3(gx), dim3(256), 0, b->s, b->ddev, b->dbas, b->t_x, b->n_x, src, t1);
    hipLaunchKernelGGL((k_b_axis<T, AXR, 1, 0>), dim3(gy), dim3(256), 0, b->s, b->ddev, b->dbas, b->t_y, b->n_y, t1, t2);
    hipLaunchKernelGGL(k_b_tridiag<T>, dim3(gm), dim3(256), 0, b->s, b->ddev, b->dbas, b->t_mode, b->n_mode, t2, scr);
    hipLaunchKernelGGL((k_b_axis<T, AXR, 1, 1>), dim3(gy), dim3(256), 0, b->s, b->ddev, b->dbas, b->t_y, b->n_y, t2, t1);
    hipLaunchKernelGGL((k_b_axis<T, AXR, 0, 1>), dim3(gx), dim3(256), 0, b->s, b->ddev, b->dbas, b->t_x, b->n_x, t1, dst);
}

static unsigned flat_blocks(long long n) { return (unsigned)std::max<long long>(1, std::min<long long>((n + 255) / 256, 8192)); }

/* one time step: P[from] -> P[1 - from] */
static void enqueue_step(fx_batch* b, int from) {
    const double* P = b->P[from];
    double* Pout = b->P[1 - from];
    launch_rhs(b, P);
    if (b->precision == FX_PREC_FP32_REFINE) {
        hipLaunchKernelGGL(k_b_d2f, dim3(flat_blocks(b->Ntot)), dim3(256), 0, b->s, b->r32, b->rhs, b->Ntot);
        spectral_solve<float>(b, b->r32, b->f1, b->f2, b->fc, b->f2);
        hipLaunchKernelGGL(k_b_f2d, dim3(flat_blocks(b->Ntot)), dim3(256), 0, b->s, b->phi, b->f2, b->Ntot, 0);
        launch_resid(b);
        spectral_solve<float>(b, b->r32, b->f1, b->f2, b->fc, b->f2);
        hipLaunchKernelGGL(k_b_f2d, dim3(flat_blocks(b->Ntot)), dim3(256), 0, b->s, b->phi, b->f2, b->Ntot, 1);
    } else {
        spectral_solve<double>(b, b->rhs, b->h1, b->h2, b->cp, b->phi);
    }
    launch_tdgl(b, P, Pout);
}

/* capture two steps (P0 -> P1 -> P0) as one graph */
static int capture(fx_batch* b) {
    free_graph(b);
    BCHK(b, hipStreamSynchronize(b->s));
    BCHK(b, hipStreamBeginCapture(b->s, hipStreamCaptureModeThreadLocal));
    enqueue_step(b, 0);
    enqueue_step(b, 1);
    hipGraph_t g = nullptr;
    BCHK(b, hipStreamEndCapture(b->s, &g));
    b->graph = g;
    BCHK(b, hipGraphInstantiate(&b->gexec, b->graph, nullptr, nullptr, 0));
    return FX_OK;
}

extern "C" {

int fx_batch_create(int n, const fx_inputs* const* inputs, fx_batch** out) {
    if (n < 1 || !inputs || !out) return FX_EINVAL;
    *out = nullptr;
    fx_batch* b = new fx_batch();
    b->nd = n;
    int rc = FX_OK;
    auto bail = [&](int code) { std::fprintf(stderr, "fx_batch_create: %s\n", b->err.c_str()); fx_batch_destroy(b); return code; };
    if (hipStreamCreateWithFlags(&b->s, hipStreamNonBlocking) != hipSuccess) { delete b; return FX_EHIP; }
    int dev = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&b->ncu, hipDeviceAttributeMultiprocessorCount, dev);
    if (b->ncu <= 0) b->ncu = 256;
    b->precision = inputs[0]->precision;

    /* per-device set-up through a single-device context: material layout,
     * permittivity, initial P, BC values */
    std::vector<std::vector<uint8_t>> mats(n);
    std::vector<std::vector<double>> epsk(n), bzk(n), P0(n);
    std::map<int, int> basis_of;
    b->hdev.resize(n);
    b->in.resize(n);
    for (int d = 0; d < n; ++d) {
        const fx_inputs& in = *inputs[d];
        if (!(in.is_periodic[0] && in.is_periodic[1]) || in.is_periodic[2])
            { b->err = "device " + std::to_string(d) + ": batched devices are periodic in x and y and bounded in z"; return bail(FX_EINVAL); }
        if (in.bc_lo[2] != FX_BC_DIRICHLET || in.bc_hi[2] != FX_BC_DIRICHLET)
            { b->err = "device " + std::to_string(d) + ": batched devices have Dirichlet z faces"; return bail(FX_EINVAL); }
        if (in.TimeIntegratorOrder != 1 || in.is_polarization_scalar != 1 || in.Coordinate_Transformation != 0)
            { b->err = "device " + std::to_string(d) + ": batched devices use forward Euler, scalar P, no rotation"; return bail(FX_EINVAL); }
        if (in.include_Landau != 1 || in.include_Grad != 1 || in.include_Elec != 1)
            { b->err = "device " + std::to_string(d) + ": batched devices include all three TDGL terms"; return bail(FX_EINVAL); }
        if (in.n_cell[0] > 128 || in.n_cell[1] > 128)
            { b->err = "device " + std::to_string(d) + ": lateral extent above 128 cells"; return bail(FX_EINVAL); }
        fx_ctx* c = nullptr;
        if ((rc = fx_create(&in, &c))) { b->err = "device " + std::to_string(d) + ": fx_create failed"; return bail(rc); }
        fx_inputs ein;
        double ih2[3], bv[6];
        BDev& D = b->hdev[d];
        rc = fxi_export(c, mats[d], D.msx, D.msy, ein, ih2, epsk[d], bzk[d], bv);
        if (!rc) {
            P0[d].resize((size_t)in.n_cell[0] * in.n_cell[1] * in.n_cell[2]);
            rc = fx_get_field(c, FX_FIELD_PZ, P0[d].data());
            double px = 0.0;
            std::vector<double> tmp(P0[d].size());
            for (int f : {FX_FIELD_PX, FX_FIELD_PY}) {
                if (!rc) rc = fx_get_field(c, f, tmp.data());
                for (double v : tmp) px = std::max(px, std::fabs(v));
            }
            if (!rc) {
                std::vector<double> sc(P0[d].size());
                rc = fx_get_field(c, FX_FIELD_MASK, sc.data());
                for (double v : sc) if (v >= 2.0) { b->err = "device " + std::to_string(d) + ": semiconductor regions are not batched"; rc = FX_EINVAL; break; }
            }
            if (!rc && px != 0.0) { b->err = "device " + std::to_string(d) + ": Px/Py must be zero"; rc = FX_EINVAL; }
        } else {
            b->err = "device " + std::to_string(d) + ": " + fx_last_error(c);
        }
        fx_destroy(c);
        if (rc) return bail(rc);
        b->in[d] = in;
        for (int q = 0; q < 3; ++q) {
            D.g.n[q] = in.n_cell[q]; D.g.per[q] = in.is_periodic[q];
            D.g.dx[q] = (in.prob_hi[q] - in.prob_lo[q]) / in.n_cell[q];
            D.g.prob_lo[q] = in.prob_lo[q]; D.g.prob_hi[q] = in.prob_hi[q];
            D.pbc.lo[q] = in.P_BC_flag_lo[q]; D.pbc.hi[q] = in.P_BC_flag_hi[q];
            D.bclo[q] = in.bc_lo[q]; D.bchi[q] = in.bc_hi[q];
            D.ih2[q] = ih2[q];
        }
        D.pbc.lambda = in.lambda;
        std::memcpy(D.bv, bv, sizeof(bv));
        D.ghost_face = in.phi_ghost_mode == FX_GHOST_FACE;
        D.alpha = in.alpha; D.beta = in.beta; D.gamma = in.gamma; D.g11 = in.g11; D.g44m = in.g44 - in.g44_p;
        D.G = in.BigGamma; D.dt = in.dt;
        for (int q = 0; q < 2; ++q) {
            const int nn = in.n_cell[q];
            if (!basis_of.count(nn)) { const int id = (int)basis_of.size(); basis_of[nn] = id; }
        }
        D.bx = basis_of[in.n_cell[0]]; D.by = basis_of[in.n_cell[1]];
        b->off.push_back(b->Ntot);
        D.off = b->Ntot;
        b->Ntot += (long long)P0[d].size();
    }
    /* device arrays */
    const size_t B = (size_t)b->Ntot * sizeof(double);
    if ((rc = balloc(b, (void**)&b->P[0], B)) || (rc = balloc(b, (void**)&b->P[1], B)) || (rc = balloc(b, (void**)&b->phi, B)) ||
        (rc = balloc(b, (void**)&b->rhs, B)))
        return bail(rc);
    if (b->precision == FX_PREC_FP32_REFINE) {
        const size_t Bf = (size_t)b->Ntot * sizeof(float);
        if ((rc = balloc(b, (void**)&b->f1, Bf)) || (rc = balloc(b, (void**)&b->f2, Bf)) || (rc = balloc(b, (void**)&b->fc, Bf)) ||
            (rc = balloc(b, (void**)&b->r32, Bf)))
            return bail(rc);
    } else {
        if ((rc = balloc(b, (void**)&b->h1, B)) || (rc = balloc(b, (void**)&b->h2, B)) || (rc = balloc(b, (void**)&b->cp, B)))
            return bail(rc);
    }
    for (int d = 0; d < n; ++d) {
        BDev& D = b->hdev[d];
        uint8_t* m = nullptr;
        double *ek = nullptr, *bk = nullptr;
        if ((rc = balloc(b, (void**)&m, mats[d].size())) || (rc = balloc(b, (void**)&ek, epsk[d].size() * sizeof(double))) ||
            (rc = balloc(b, (void**)&bk, bzk[d].size() * sizeof(double))))
            return bail(rc);
        BCHK(b, hipMemcpy(m, mats[d].data(), mats[d].size(), hipMemcpyHostToDevice));
        BCHK(b, hipMemcpy(ek, epsk[d].data(), epsk[d].size() * sizeof(double), hipMemcpyHostToDevice));
        BCHK(b, hipMemcpy(bk, bzk[d].data(), bzk[d].size() * sizeof(double), hipMemcpyHostToDevice));
        D.mat = m; D.epsk = ek; D.bzk = bk;
        for (int q = 0; q < 2; ++q)
            BCHK(b, hipMemcpy(b->P[q] + D.off, P0[d].data(), P0[d].size() * sizeof(double), hipMemcpyHostToDevice));
    }
    /* bases */
    std::vector<BBasis> hb(basis_of.size());
    for (auto& kv : basis_of) {
        const int nn = kv.first;
        std::vector<double> Q((size_t)nn * nn), QT((size_t)nn * nn), mu(nn);
        fx_batch_basis(nn, Q.data(), mu.data());
        for (int i = 0; i < nn; ++i)
            for (int j = 0; j < nn; ++j) QT[(size_t)j * nn + i] = Q[(size_t)i * nn + j];
        std::vector<float> Q32(Q.begin(), Q.end()), QT32(QT.begin(), QT.end());
        BBasis& bb = hb[kv.second];
        bb.n = nn;
        double *q = nullptr, *qt = nullptr, *m = nullptr;
        float *q32 = nullptr, *qt32 = nullptr;
        const size_t nb = Q.size();
        if ((rc = balloc(b, (void**)&q, nb * 8)) || (rc = balloc(b, (void**)&qt, nb * 8)) || (rc = balloc(b, (void**)&m, nn * 8)) ||
            (rc = balloc(b, (void**)&q32, nb * 4)) || (rc = balloc(b, (void**)&qt32, nb * 4)))
            return bail(rc);
        BCHK(b, hipMemcpy(q, Q.data(), nb * 8, hipMemcpyHostToDevice));
        BCHK(b, hipMemcpy(qt, QT.data(), nb * 8, hipMemcpyHostToDevice));
        BCHK(b, hipMemcpy(m, mu.data(), nn * 8, hipMemcpyHostToDevice));
        BCHK(b, hipMemcpy(q32, Q32.data(), nb * 4, hipMemcpyHostToDevice));
        BCHK(b, hipMemcpy(qt32, QT32.data(), nb * 4, hipMemcpyHostToDevice));
        bb.Q = q; bb.QT = qt; bb.mu = m; bb.Q32 = q32; bb.QT32 = qt32;
    }
    if ((rc = balloc(b, (void**)&b->dbas, hb.size() * sizeof(BBasis))) || (rc = balloc(b, (void**)&b->ddev, n * sizeof(BDev))))
        return bail(rc);
    BCHK(b, hipMemcpy(b->dbas, hb.data(), hb.size() * sizeof(BBasis), hipMemcpyHostToDevice));
    BCHK(b, hipMemcpy(b->ddev, b->hdev.data(), n * sizeof(BDev), hipMemcpyHostToDevice));
    b->mfma = true;
    for (int d = 0; d < n; ++d)
        if (b->hdev[d].g.n[0] % 16 || b->hdev[d].g.n[1] % 16) b->mfma = false;
    if ((rc = build_fixed_tiles(b)) || (rc = build_cell_tiles(b, 64))) return bail(rc);
    b->g_cell = b->g_axis = b->g_mode = 4 * b->ncu;
    if ((rc = capture(b))) return bail(rc);
    *out = b;
    return FX_OK;
}

int fx_batch_destroy(fx_batch* b) {
    if (!b) return FX_OK;
    if (b->s) (void)hipStreamSynchronize(b->s);
    free_graph(b);
    for (void* p : b->allocs) (void)hipFree(p);
    if (b->s) (void)hipStreamDestroy(b->s);
    delete b;
    return FX_OK;
}

const char* fx_batch_last_error(fx_batch* b) { return b ? b->err.c_str() : "null batch"; }

int fx_batch_num_devices(fx_batch* b) { return b ? b->nd : 0; }

long long fx_batch_cells(fx_batch* b) { return b ? b->Ntot : 0; }

int fx_batch_dims(fx_batch* b, int dev, int dims[3]) {
    if (!b || dev < 0 || dev >= b->nd || !dims) return FX_EINVAL;
    for (int q = 0; q < 3; ++q) dims[q] = b->hdev[dev].g.n[q];
    return FX_OK;
}

int fx_batch_step(fx_batch* b, int nsteps, int use_graph, double* wall_seconds) {
    if (!b || nsteps < 0) return FX_EINVAL;
    const auto t0 = std::chrono::steady_clock::now();
    int done = 0;
    if (use_graph && b->gexec) {
        /* the graph advances P[0] -> P[0] by two steps */
        if (b->cur == 1 && nsteps > 0) { enqueue_step(b, 1); b->cur = 0; ++done; }
        while (nsteps - done >= 2) { BCHK(b, hipGraphLaunch(b->gexec, b->s)); done += 2; }
    }
    while (done < nsteps) { enqueue_step(b, b->cur); b->cur = 1 - b->cur; ++done; }
    BCHK(b, hipGetLastError());
    BCHK(b, hipStreamSynchronize(b->s));
    b->steps += nsteps;
    if (wall_seconds) *wall_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return FX_OK;
}

int fx_batch_get_field(fx_batch* b, int dev, int field, double* host) {
    if (!b || dev < 0 || dev >= b->nd || !host) return FX_EINVAL;
    const BDev& D = b->hdev[dev];
    const size_t n = (size_t)D.g.n[0] * D.g.n[1] * D.g.n[2];
    const double* src = field == FX_FIELD_PZ ? b->P[b->cur] : field == FX_FIELD_PHI ? b->phi : nullptr;
    if (!src) return bfail(b, FX_EINVAL, "fx_batch_get_field: Pz and Phi only");
    BCHK(b, hipStreamSynchronize(b->s));
    BCHK(b, hipMemcpy(host, src + D.off, n * sizeof(double), hipMemcpyDeviceToHost));
    return FX_OK;
}

int fx_batch_set_field(fx_batch* b, int dev, int field, const double* host) {
    if (!b || dev < 0 || dev >= b->nd || !host) return FX_EINVAL;
    if (field != FX_FIELD_PZ) return bfail(b, FX_EINVAL, "fx_batch_set_field: Pz only");
    const BDev& D = b->hdev[dev];
    const size_t n = (size_t)D.g.n[0] * D.g.n[1] * D.g.n[2];
    BCHK(b, hipStreamSynchronize(b->s));
    for (int q = 0; q < 2; ++q) BCHK(b, hipMemcpy(b->P[q] + D.off, host, n * sizeof(double), hipMemcpyHostToDevice));
    return FX_OK;
}

/* Chooses, by timing on this batch, the stencil tile width (32, 64 or 128
 * cells by 256/width rows) and the persistent grid size of the stencil,
 * transform and tridiagonal stages (1, 2, 4 or 8 workgroups per CU), then
 * re-captures the step graph.  report (optional) receives a JSON summary. */
int fx_batch_autotune(fx_batch* b, char* report, int report_len) {
    if (!b) return FX_EINVAL;
    hipEvent_t e0, e1;
    BCHK(b, hipEventCreate(&e0));
    BCHK(b, hipEventCreate(&e1));
    auto time_ms = [&](auto&& fn) {
        fn();                                /* warm */
        (void)hipEventRecord(e0, b->s);
        for (int r = 0; r < 5; ++r) fn();
        (void)hipEventRecord(e1, b->s);
        (void)hipEventSynchronize(e1);
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        return ms / 5.0;
    };
    std::string js = "{\"stencil\": [";
    int best_tx = b->tx, best_gc = b->g_cell;
    double best = 1e30;
    const int factors[4] = {1, 2, 4, 8};
    bool first = true;
    for (int tx : {32, 64, 128}) {
        int rc = build_cell_tiles(b, tx);
        if (rc) { (void)hipEventDestroy(e0); (void)hipEventDestroy(e1); return rc; }
        for (int f : factors) {
            b->g_cell = f * b->ncu;
            const double ms = time_ms([&] { launch_tdgl(b, b->P[b->cur], b->P[1 - b->cur]); launch_rhs(b, b->P[b->cur]); });
            char buf[128];
            std::snprintf(buf, sizeof buf, "%s{\"tile\": [%d, %d], \"grid\": %d, \"ms\": %.5f}", first ? "" : ", ", tx, 256 / tx,
                          b->g_cell, ms);
            js += buf;
            first = false;
            if (ms < best) { best = ms; best_tx = tx; best_gc = b->g_cell; }
        }
    }
    int rc = build_cell_tiles(b, best_tx);
    if (rc) return rc;
    b->g_cell = best_gc;
    js += "], \"spectral\": [";
    double bestS = 1e30;
    int best_ga = b->g_axis, best_gm = b->g_mode;
    first = true;
    for (int fa : factors)
        for (int fm : {1, 2, 4}) {
            b->g_axis = fa * b->ncu;
            b->g_mode = fm * b->ncu;
            const double ms = time_ms([&] {
                if (b->precision == FX_PREC_FP32_REFINE) spectral_solve<float>(b, b->r32, b->f1, b->f2, b->fc, b->f2);
                else spectral_solve<double>(b, b->rhs, b->h1, b->h2, b->cp, b->h2);
            });
            char buf[128];
            std::snprintf(buf, sizeof buf, "%s{\"axis_grid\": %d, \"mode_grid\": %d, \"ms\": %.5f}", first ? "" : ", ", b->g_axis,
                          b->g_mode, ms);
            js += buf;
            first = false;
            if (ms < bestS) { bestS = ms; best_ga = b->g_axis; best_gm = b->g_mode; }
        }
    b->g_axis = best_ga;
    b->g_mode = best_gm;
    char buf[256];
    std::snprintf(buf, sizeof buf, "], \"chosen\": {\"tile\": [%d, %d], \"cell_grid\": %d, \"axis_grid\": %d, \"mode_grid\": %d}}",
                  b->tx, 256 / b->tx, b->g_cell, b->g_axis, b->g_mode);
    js += buf;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    /* the tuning runs wrote scratch only (rhs, phi and the transform buffers
     * are recomputed by every step; P[1 - cur] is the next step's output) */
    b->tune_report = js;
    if (report && report_len > 0) { std::strncpy(report, js.c_str(), report_len - 1); report[report_len - 1] = 0; }
    return capture(b);
}

}  // extern "C"
