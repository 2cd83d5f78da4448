/*
 * Direct Poisson solve of a laterally periodic, z-layered operator.
 *
 *   -div(beta grad phi) = b   on n0 x n1 x n2 cells, periodic in x and y,
 *   bounded in z (Dirichlet value on the face, ghost 2 phi_b - phi_0, or
 *   homogeneous Neumann; at least one Dirichlet face), with face
 *   coefficients that depend on the plane k only.
 *
 * This is the operator the reference hands to AMReX MLMG
 * (ElectrostaticSolver.cpp:562-611, 713-720) whenever the stack has no
 * semiconductor Newton term and its permittivity varies in z only, which is
 * the case for every MFIM stack of the reference's inputs.  Such an operator
 * is diagonalised exactly by the separable discrete Hartley transform in x
 * and y: the periodic second difference is a symmetric circulant, whose
 * eigenvectors cos(2 pi m n / N) and sin(2 pi m n / N) share the eigenvalue
 * mu_m = 2 - 2 cos(2 pi m / N), so cas_m = cos + sin is one too, and the cas
 * vectors are orthogonal with H H = N I.  What is left is one tridiagonal
 * system in z per lateral mode (mx, my):
 *
 *   (bx_k mu_x / hx^2 + by_k mu_y / hy^2 + z couplings) x_k
 *       - wz_k x_k-1 - wz_k+1 x_k+1 = b^_k.
 *
 * Five launches, each a pass over the field:
 *   k_fft_rows  (forward)  b (+ the Dirichlet face terms) -> S, Hartley along x
 *   k_fft_cols  (forward)  S -> S, Hartley along y
 *   k_fft_tri              S -> S, Thomas solve in z per mode, scaled by 1/(n0 n1)
 *   k_fft_cols  (inverse)  S -> S
 *   k_fft_rows  (inverse)  S -> phi
 * The forward transforms skip planes whose right-hand side is known to be 0
 * (the non-FE planes of a stack without semiconductor cells).  The pivots
 * of the z systems depend on the operator only; they are tabulated once
 * (k_fft_pivots), one system per mirror class of modes.
 *
 * The Hartley transforms of two real rows come from one complex FFT of the
 * row pair packed as re + i im (radix-2, in LDS): with Z = FFT(re + i im),
 * A + iB = Z[m] and C + iD = Z[-m],
 *   H(re)[m] = (A - B + C + D) / 2,   H(im)[m] = (A + B - C + D) / 2.
 * n0 and n1 are powers of two (4 .. 1024).  The result is exact to rounding:
 * no iteration and no stopping test.
 */
#include <type_traits>

#include "fx_kernels.h"

namespace {

using fxd::FftPlan;

/* Threads of a transform workgroup (R sequences x max(N1, N2)) and the y
 * transform's LDS layout: rows of the four-step buffer padded by one value
 * (PAD), twiddles staged in LDS (TWL).  Measured on 256 x 256 planes: the x
 * transform is fastest with 256 threads (23.7 against 25.4 us per pass); the
 * y transform with 128 threads, padded rows and twiddles in LDS (31.0 us;
 * without padding and with twiddles read from the cache five workgroups fit
 * a CU, but 36.9 us; 64 threads, 44.9 us). */
#ifndef FX_FFT_ROWS_THR
#define FX_FFT_ROWS_THR 256
#endif
#ifndef FX_FFT_COLS_THR
#define FX_FFT_COLS_THR 128
#endif
#ifndef FX_FFT_COLS_PAD
#define FX_FFT_COLS_PAD 1
#endif
#ifndef FX_FFT_COLS_TWL
#define FX_FFT_COLS_TWL 1
#endif
#ifndef FX_FFT_REG_HARTLEY
#define FX_FFT_REG_HARTLEY 1
#endif
#ifndef FX_DPP_ROR
#define FX_DPP_ROR 1
#endif
/* the four-step buffer carries the real parts across, then the imaginary
 * ones (two barriers more, half the LDS): on 80 planes of 256 x 256 the y
 * transform's 1280 workgroups then fit the CUs at once (eight a CU instead of
 * four) and the x transform's 640 (three instead of two); 170.8 -> 160.5 us
 * per time step with both */
#ifndef FX_FFT_COLS_SPLIT
#define FX_FFT_COLS_SPLIT 1
#endif
#ifndef FX_FFT_ROWS_SPLIT
#define FX_FFT_ROWS_SPLIT 1
#endif
/* one value of padding after each sequence of the split buffer: the y
 * transform's first-step stores (thread = sequence s, row b; its 16 values
 * along c) otherwise put the eight sequences of a row on the same banks
 * (159.9 -> 157.1 us per time step) */
#ifndef FX_FFT_COLS_SPAD
#define FX_FFT_COLS_SPAD 1
#endif
#ifndef FX_FFT_ROWS_SPAD
#define FX_FFT_ROWS_SPAD 0
#endif
constexpr int ROWS_THR = FX_FFT_ROWS_THR, COLS_THR = FX_FFT_COLS_THR;
constexpr int TRI_THR = 256;
constexpr int TRI_BLK = 8;    /* planes per load block of the tridiagonal solve */
#ifndef FX_TRI_REG
#define FX_TRI_REG 64
#endif
constexpr int TRI_REG = FX_TRI_REG;   /* planes of a 128-plane column whose y stays in registers (the rest in LDS) */

template <int N>
struct Lg { static constexpr int v = N <= 1 ? 0 : 1 + Lg<N / 2>::v; };

template <int N>
__host__ __device__ constexpr int brev_c(int v) {
    int r = 0;
    for (int i = 0; i < Lg<N>::v; ++i) r |= ((v >> i) & 1) << (Lg<N>::v - 1 - i);
    return r;
}

/* omega_n^q = exp(-2 pi i q / n) for 0 <= q < n from the half table
 * tw[2q'] = cos, tw[2q'+1] = -sin, q' < n/2 (omega^(q + n/2) = -omega^q) */
template <class T>
__device__ __forceinline__ void omega(const T* tw, int q, int half, T& c, T& sn) {
    const int r = q >= half ? q - half : q;
    c = tw[2 * r];
    sn = tw[2 * r + 1];
    if (q >= half) { c = -c; sn = -sn; }
}

/* In-register DFT of length N (decimation in frequency, radix 2): on
 * return v[i] holds X[bitrev(i)].  tw is the n-point half table and
 * ts = n / N, so that omega_N^j = tw[j ts]. */
template <class T, int N>
__device__ __forceinline__ void reg_dft(T (&vr)[N], T (&vi)[N], const T* tw, int ts) {
#pragma unroll
    for (int m = N / 2; m >= 1; m >>= 1) {
#pragma unroll
        for (int g = 0; g < N; g += 2 * m) {
#pragma unroll
            for (int j = 0; j < m; ++j) {
                const int i0 = g + j, i1 = i0 + m;
                const T ur = vr[i0], ui = vi[i0], wr = vr[i1], wi = vi[i1];
                vr[i0] = ur + wr;
                vi[i0] = ui + wi;
                const T dr = ur - wr, di = ui - wi;
                if (j == 0) {
                    vr[i1] = dr;
                    vi[i1] = di;
                } else {
                    const int q = j * (N / (2 * m)) * ts;   /* omega_2m^j */
                    const T c = tw[2 * q], sn = tw[2 * q + 1];
                    vr[i1] = dr * c - di * sn;
                    vi[i1] = dr * sn + di * c;
                }
            }
        }
    }
}

/* Four-step FFT of R packed sequences of length n = N1 N2 (x = N2 a + b,
 * X = c + N1 d):
 *   step 1, thread (s, b): DFT over a of x[N2 a + b], times omega_n^(b c),
 *           into Y[s][b][c] in LDS (rows padded to N1 + 1);
 *   step 3, thread (s, c): DFT over b of Y[s][b][c] gives X[c + N1 d],
 *           stored in LDS as X[s][k] over the Y buffer.
 * load(s, a, b, re, im) reads element N2 a + b of sequence s; after the
 * call, X is in Xr/Xi (Xr[s n + k]) for the Hartley step. */
template <class T, int N1, int N2, int PAD, class Load>
__device__ __forceinline__ void four_step_1(int R, const T* tw, T* Yr, T* Yi, Load load, int step1_s_fast) {
    constexpr int n = N1 * N2, YS = N1 + PAD;
    const int t = threadIdx.x;
    if (t < R * N2) {
        const int s = step1_s_fast ? t % R : t / N2;
        const int b = step1_s_fast ? t / R : t % N2;
        T vr[N1], vi[N1];
#pragma unroll
        for (int a = 0; a < N1; ++a) load(s, a, b, vr[a], vi[a]);
        reg_dft<T, N1>(vr, vi, tw, N2);
#pragma unroll
        for (int c = 0; c < N1; ++c) {
            T wc, ws;
            omega<T>(tw, b * c, n / 2, wc, ws);
            const T xr = vr[brev_c<N1>(c)], xi = vi[brev_c<N1>(c)];
            Yr[(s * N2 + b) * YS + c] = xr * wc - xi * ws;
            Yi[(s * N2 + b) * YS + c] = xr * ws + xi * wc;
        }
    }
    __syncthreads();
}

template <class T, int N1, int N2, int PAD, class Load>
__device__ __forceinline__ void four_step(int R, const T* tw, T* Yr, T* Yi, Load load, int step1_s_fast) {
    constexpr int n = N1 * N2, YS = N1 + PAD;
    const int t = threadIdx.x;
    four_step_1<T, N1, N2, PAD>(R, tw, Yr, Yi, load, step1_s_fast);
    T ur[N2], ui[N2];
    const int s3 = t / N1, c3 = t % N1;
    const bool act3 = t < R * N1;
    if (act3) {
#pragma unroll
        for (int b = 0; b < N2; ++b) {
            ur[b] = Yr[(s3 * N2 + b) * YS + c3];
            ui[b] = Yi[(s3 * N2 + b) * YS + c3];
        }
        reg_dft<T, N2>(ur, ui, tw, N1);
    }
    __syncthreads();
    if (act3) {
#pragma unroll
        for (int d = 0; d < N2; ++d) {
            Yr[s3 * n + c3 + N1 * d] = ur[brev_c<N2>(d)];
            Yi[s3 * n + c3 + N1 * d] = ui[brev_c<N2>(d)];
        }
    }
    __syncthreads();
}

/* Hartley values of both packed parts at index k of sequence s. */
template <class T>
__device__ __forceinline__ void hartley2(const T* Xr, const T* Xi, int s, int n, int k, T& h0, T& h1) {
    const int mk = (n - k) & (n - 1);
    const T A = Xr[s * n + k], B = Xi[s * n + k], C = Xr[s * n + mk], D = Xi[s * n + mk];
    h0 = T(0.5) * ((A - B) + (C + D));
    h1 = T(0.5) * ((A + B) + (D - C));
}

/* Lane c of each row of 16 lanes receives v of lane (16 - c) mod 16:
 * row_mirror (lane c <- lane 15 - c), then row_ror:1 (lane c <- lane c - 1),
 * one 32-bit word at a time. */
template <class T>
__device__ __forceinline__ T lane_partner16(T v) {
    constexpr int W = sizeof(T) / 4;
    union U { T v; int w[W]; };
    U a, b;
    a.v = v;
#pragma unroll
    for (int i = 0; i < W; ++i) {
        const int m = __builtin_amdgcn_update_dpp(0, a.w[i], 0x140, 0xF, 0xF, false);
        b.w[i] = __builtin_amdgcn_update_dpp(0, m, 0x120 + FX_DPP_ROR, 0xF, 0xF, false);
    }
    return b.v;
}

/* The four-step FFT with the Hartley step in registers, for N1 = 16: after
 * step 3, thread (s, c) holds X[c + 16 d], d < N2, of sequence s, and
 * X[n - c - 16 d] is X[(16 - c) + 16 (N2 - 1 - d)] of the thread 16 - c of the
 * same sequence (the lane (16 - c) mod 16 of the same 16-lane row; for c = 0,
 * X[16 ((N2 - d) mod N2)] of the thread itself).  The pair is exchanged by
 * DPP, so X is never written back to LDS: one barrier fewer.  The Hartley
 * values are formed as hartley2 forms them (bitwise the same);
 * store(s, k, h0, h1) receives index k of sequence s. */
template <class T, int N1, int N2, int PAD, bool SPLIT, int SP, class Load, class Store>
__device__ __forceinline__ void four_step_hartley16(int R, const T* tw, T* Yr, T* Yi, Load load, int step1_s_fast,
                                                    Store store) {
    static_assert(N1 == 16, "the lane exchange works on rows of 16 lanes");
    constexpr int n = N1 * N2, YS = N1 + PAD;
    const int t = threadIdx.x;
    const int s3 = t / N1, c3 = t % N1;
    T ur[N2], ui[N2];
    if (!SPLIT) {
        four_step_1<T, N1, N2, PAD>(R, tw, Yr, Yi, load, step1_s_fast);
        if (t < R * N1) {
#pragma unroll
            for (int b = 0; b < N2; ++b) {
                ur[b] = Yr[(s3 * N2 + b) * YS + c3];
                ui[b] = Yi[(s3 * N2 + b) * YS + c3];
            }
        }
    } else {
        /* one buffer of real values (Yr): the real parts of step 1 go across,
         * then the imaginary parts, so the workgroup needs half the LDS */
        T xr[N1], xi[N1];
        const bool act1 = t < R * N2;
        /* sequence s at s (N2 YS + SP): SP = 1 moves consecutive sequences to
         * other LDS banks */
        constexpr int SS = N2 * YS + SP;
        const int s1 = step1_s_fast ? t % R : t / N2, b1 = step1_s_fast ? t / R : t % N2;
        if (act1) {
            T vr[N1], vi[N1];
#pragma unroll
            for (int a = 0; a < N1; ++a) load(s1, a, b1, vr[a], vi[a]);
            reg_dft<T, N1>(vr, vi, tw, N2);
#pragma unroll
            for (int c = 0; c < N1; ++c) {
                T wc, ws;
                omega<T>(tw, b1 * c, n / 2, wc, ws);
                const T ar = vr[brev_c<N1>(c)], ai = vi[brev_c<N1>(c)];
                xr[c] = ar * wc - ai * ws;
                xi[c] = ar * ws + ai * wc;
            }
#pragma unroll
            for (int c = 0; c < N1; ++c) Yr[s1 * SS + b1 * YS + c] = xr[c];
        }
        __syncthreads();
        if (t < R * N1) {
#pragma unroll
            for (int b = 0; b < N2; ++b) ur[b] = Yr[s3 * SS + b * YS + c3];
        }
        __syncthreads();
        if (act1) {
#pragma unroll
            for (int c = 0; c < N1; ++c) Yr[s1 * SS + b1 * YS + c] = xi[c];
        }
        __syncthreads();
        if (t < R * N1) {
#pragma unroll
            for (int b = 0; b < N2; ++b) ui[b] = Yr[s3 * SS + b * YS + c3];
        }
    }
    if (t < R * N1) {
        T pr[N2], pi[N2];
        reg_dft<T, N2>(ur, ui, tw, N1);
#pragma unroll
        for (int r = 0; r < N2; ++r) {
            pr[r] = lane_partner16<T>(ur[r]);
            pi[r] = lane_partner16<T>(ui[r]);
        }
#pragma unroll
        for (int d = 0; d < N2; ++d) {
            const T A = ur[brev_c<N2>(d)], B = ui[brev_c<N2>(d)];
            const T C = c3 ? pr[brev_c<N2>(N2 - 1 - d)] : ur[brev_c<N2>((N2 - d) & (N2 - 1))];
            const T D = c3 ? pi[brev_c<N2>(N2 - 1 - d)] : ui[brev_c<N2>((N2 - d) & (N2 - 1))];
            store(s3, c3 + N1 * d, T(0.5) * ((A - B) + (C + D)), T(0.5) * ((A + B) + (D - C)));
        }
    }
}

template <class T>
__device__ __forceinline__ void load_tw(T* tw, const T* g, int n) {
    for (int q = threadIdx.x; q < n; q += blockDim.x) tw[q] = g[q];
}

/* Hartley transform along x of rows j0 .. j0 + 2R - 1 of plane k =
 * blockIdx.y: rows 2s and 2s + 1 of the block are sequence s.  FWD: input b
 * (TIN) plus the Dirichlet face terms on the first and last plane, output
 * S.  !FWD: input S, output phi (TOUT), stored or, with add, added; with
 * dotp, also the workgroup's partial sum of phi . dotv (the stored values)
 * to dotp[blockIdx.y gridDim.x + blockIdx.x]. */
template <class T, class TIN, class TOUT, bool FWD, int N1, int N2>
__global__ __launch_bounds__(ROWS_THR) void k_fft_rows(FftPlan<T> P, const TIN* __restrict__ in, TOUT* __restrict__ out,
                                                       int R, double blo, double bhi, int add, int kz0,
                                                       const double* __restrict__ dotv, double* __restrict__ dotp) {
    extern __shared__ __align__(16) unsigned char fft_lds[];
    __shared__ double dsh[ROWS_THR / 64 + 1];
    if (P.skip && *P.skip != 0.0) return;
    constexpr int n = N1 * N2;
    const int k = kz0 + blockIdx.y, j0 = blockIdx.x * 2 * R;
    T* const tw = reinterpret_cast<T*>(fft_lds);
    T* const Yr = tw + n;
    T* const Yi = Yr + R * N2 * (N1 + 1);
    load_tw<T>(tw, P.tw0, n);
    __syncthreads();
    const size_t row0 = ((size_t)k * P.n1 + j0) * n;
    double bterm = 0.0;
    if (FWD) bterm = (k == 0 ? blo : 0.0) + (k == P.n2 - 1 ? bhi : 0.0);
    auto load = [&](int s, int a, int b, T& re, T& im) {
        const size_t o = row0 + (size_t)(2 * s) * n + N2 * a + b;
        if (FWD) {
            re = (T)((double)in[o] + bterm);
            im = (T)((double)in[o + n] + bterm);
        } else {
            re = (T)in[o];
            im = (T)in[o + n];
        }
    };
    if constexpr (N1 == 16 && FX_FFT_REG_HARTLEY) {
        if (FWD || !dotp) {
            four_step_hartley16<T, N1, N2, 1, FX_FFT_ROWS_SPLIT != 0, FX_FFT_ROWS_SPAD>(R, tw, Yr, Yi, load, 0, [&](int s, int kk, T h0, T h1) {
                const size_t o = row0 + (size_t)(2 * s) * n + kk;
                if (FWD || !add) {
                    out[o] = (TOUT)h0;
                    out[o + n] = (TOUT)h1;
                } else {
                    out[o] = out[o] + (TOUT)h0;
                    out[o + n] = out[o + n] + (TOUT)h1;
                }
            });
            return;
        }
    }
    four_step<T, N1, N2, 1>(R, tw, Yr, Yi, load, 0);
    double dacc = 0.0;
    for (int t = threadIdx.x; t < R * n; t += blockDim.x) {
        const int s = t / n, kk = t % n;
        T h0, h1;
        hartley2<T>(Yr, Yi, s, n, kk, h0, h1);
        const size_t o = row0 + (size_t)(2 * s) * n + kk;
        if (FWD || !add) {
            out[o] = (TOUT)h0;
            out[o + n] = (TOUT)h1;
            if (!FWD && dotp) dacc += (double)(TOUT)h0 * dotv[o] + (double)(TOUT)h1 * dotv[o + n];
        } else {
            out[o] = out[o] + (TOUT)h0;
            out[o + n] = out[o + n] + (TOUT)h1;
        }
    }
    if (!FWD && dotp) {
        dacc = fxd::block_sum<ROWS_THR>(dacc, dsh);
        if (threadIdx.x == 0) dotp[blockIdx.y * gridDim.x + blockIdx.x] = dacc;
    }
}

template <class T> struct Vec2;
template <> struct Vec2<double> { using t = double2; };
template <> struct Vec2<float> { using t = float2; };

/* Hartley transform along y, in place, of the column pairs (i0 + 2s,
 * i0 + 2s + 1), s < R, of plane k = kz0 + blockIdx.y: each row segment of
 * 2R values is read and written whole (one vector of two per thread). */
template <class T, int N1, int N2>
__global__ __launch_bounds__(COLS_THR) void k_fft_cols(FftPlan<T> P, int R, int kz0) {
    extern __shared__ __align__(16) unsigned char fft_lds[];
    if (P.skip && *P.skip != 0.0) return;
    using V2 = typename Vec2<T>::t;
    constexpr int n = N1 * N2, PAD = FX_FFT_COLS_PAD;
    const int n0 = P.n0, k = kz0 + blockIdx.y, i0 = blockIdx.x * 2 * R;
    T* const twl = reinterpret_cast<T*>(fft_lds);
    T* const Yr = twl + (FX_FFT_COLS_TWL ? n : 0);
    T* const Yi = Yr + R * N2 * (N1 + PAD);
    const T* tw = P.tw1;
    if (FX_FFT_COLS_TWL) {
        load_tw<T>(twl, P.tw1, n);
        __syncthreads();
        tw = twl;
    }
    T* const S = P.S + (size_t)k * n * n0 + i0;
    auto load = [&](int s, int a, int b, T& re, T& im) {
        const V2 v = *reinterpret_cast<const V2*>(S + (size_t)(N2 * a + b) * n0 + 2 * s);
        re = v.x;
        im = v.y;
    };
    if constexpr (N1 == 16 && FX_FFT_REG_HARTLEY) {
        four_step_hartley16<T, N1, N2, PAD, FX_FFT_COLS_SPLIT != 0, FX_FFT_COLS_SPAD>(R, tw, Yr, Yi, load, 1, [&](int s, int j, T h0, T h1) {
            V2 v;
            v.x = h0;
            v.y = h1;
            *reinterpret_cast<V2*>(S + (size_t)j * n0 + 2 * s) = v;
        });
        return;
    }
    four_step<T, N1, N2, PAD>(R, tw, Yr, Yi, load, 1);
    for (int t = threadIdx.x; t < R * n; t += blockDim.x) {
        const int s = t % R, j = t / R;
        T h0, h1;
        hartley2<T>(Yr, Yi, s, n, j, h0, h1);
        V2 v;
        v.x = h0;
        v.y = h1;
        *reinterpret_cast<V2*>(S + (size_t)j * n0 + 2 * s) = v;
    }
}

/* Reciprocal pivots of the z systems (FftPlan::piv), once per operator: one
 * thread per mode class (mxc, myc), the Thomas recurrence's pivots
 * t_k = d_k - a_k c'_k-1, c'_k = c_k / t_k, with
 *   d_k = bx_k mu_x / hx^2 + by_k mu_y / hy^2 + (z couplings),
 *   a_k = -bz_k / hz^2 (0 on the first plane), c_k = -bz_k+1 / hz^2 (0 on the last),
 * the z couplings 2 bz / hz^2 across a Dirichlet face and 0 across a Neumann
 * one, and -abar_k where a Helmholtz term per plane is given; formed in FP64 (correctly rounded divisions) from the FP64 plan's
 * coefficients and stored as T. */
/* the per-plane coefficients of the pivot recurrences (bx, by, bz with its
 * last face, and the Helmholtz term) staged in LDS: each step of a
 * recurrence then waits for LDS, not for the cache */
__device__ __forceinline__ void stage_plane_coefs(const FftPlan<double>& P, const double* abar, double* sx, double* sy,
                                                  double* sz, double* sa) {
    const int n2 = P.n2;
    for (int k = threadIdx.x; k <= n2; k += blockDim.x) {
        sz[k] = P.bzk[k];
        if (k < n2) {
            sx[k] = P.bxk[k];
            sy[k] = P.byk[k];
            sa[k] = abar ? abar[k] : 0.0;
        }
    }
    __syncthreads();
}

template <class T>
__global__ __launch_bounds__(256) void k_fft_pivots(FftPlan<double> P, T* __restrict__ piv, const double* __restrict__ abar) {
    extern __shared__ __align__(16) unsigned char piv_lds[];
    const int n0c = P.n0 / 2 + 1, n1c = P.n1 / 2 + 1, n2 = P.n2;
    double* const sx = reinterpret_cast<double*>(piv_lds);
    double* const sy = sx + n2;
    double* const sa = sy + n2;
    double* const sz = sa + n2;
    stage_plane_coefs(P, abar, sx, sy, sz, sa);
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n0c * n1c) return;
    const int mxc = q % n0c, myc = q / n0c;
    const double lx = P.mux[mxc] * P.ih2[0], ly = P.muy[myc] * P.ih2[1], iz = P.ih2[2];
    const size_t ps = (size_t)n0c * n1c;
    double cp = 0.0;
    for (int k = 0; k < n2; ++k) {
        const double wzl = sz[k] * iz, wzh = sz[k + 1] * iz;
        double d = sx[k] * lx;
        d += sy[k] * ly;
        d += (k > 0 ? wzl : (P.dlo ? 2.0 * wzl : 0.0)) + (k < n2 - 1 ? wzh : (P.dhi ? 2.0 * wzh : 0.0));
        if (abar) d -= sa[k];
        const double a = k > 0 ? -wzl : 0.0, c = k < n2 - 1 ? -wzh : 0.0;
        const double r = 1.0 / (d - a * cp);
        cp = c * r;
        piv[k * ps + q] = (T)r;
    }
}

/* Upward pivots of the same systems (FftPlan::pivu): eliminating from the
 * last plane, g_k = d_k - c_k a'_k+1 with a'_k = a_k / g_k, stored as 1 / g_k
 * (FP64, correctly rounded divisions; no Helmholtz term). */
__global__ __launch_bounds__(256) void k_fft_pivots_up(FftPlan<double> P, double* __restrict__ pivu) {
    extern __shared__ __align__(16) unsigned char piv_lds[];
    const int n0c = P.n0 / 2 + 1, n1c = P.n1 / 2 + 1, n2 = P.n2;
    double* const sx = reinterpret_cast<double*>(piv_lds);
    double* const sy = sx + n2;
    double* const sa = sy + n2;
    double* const sz = sa + n2;
    stage_plane_coefs(P, nullptr, sx, sy, sz, sa);
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n0c * n1c) return;
    const int mxc = q % n0c, myc = q / n0c;
    const double lx = P.mux[mxc] * P.ih2[0], ly = P.muy[myc] * P.ih2[1], iz = P.ih2[2];
    const size_t ps = (size_t)n0c * n1c;
    double ap = 0.0;   /* a'_k+1 */
    for (int k = n2 - 1; k >= 0; --k) {
        const double wzl = sz[k] * iz, wzh = sz[k + 1] * iz;
        double d = sx[k] * lx;
        d += sy[k] * ly;
        d += (k > 0 ? wzl : (P.dlo ? 2.0 * wzl : 0.0)) + (k < n2 - 1 ? wzh : (P.dhi ? 2.0 * wzh : 0.0));
        const double a = k > 0 ? -wzl : 0.0, c = k < n2 - 1 ? -wzh : 0.0;
        const double r = 1.0 / (d - c * ap);
        ap = a * r;
        pivu[k * ps + q] = r;
    }
}

/* the two-ended z solve (k_fft_tri2): planes per load block, and the share
 * of a half's values kept in LDS */
#ifndef FX_TRI2_BLK
#define FX_TRI2_BLK 8
#endif
#ifndef FX_TRI2_LDS_EIGHTHS
#define FX_TRI2_LDS_EIGHTHS 4   /* eighths of a half's H values that wait in LDS */
#endif
/* One half of the two-ended z solve (UP false: downward from plane kb, true:
 * upward from plane kb), nf steps, step j at plane k = UP ? kb - j : kb + j,
 * from y, the value of the plane before the first; the first LB steps'
 * values wait in LDS (lv[j 256]), the rest in registers (yr).  Loads of
 * steps past nf read a plane clamped into [0, 2H) and are not used. */
template <int H, int LB, bool UP>
__device__ __forceinline__ void tri2_sweep(const double* __restrict__ col, const double* __restrict__ rp, size_t s,
                                           size_t ps, const double* tab, double* lv, double (&yr)[H - LB], int kb,
                                           int nf, int zlo, int zhi, bool any, int z0, int z1, bool c0, double scale,
                                           double flo, double fhi, double& y, double& rlast) {
    constexpr int TB = FX_TRI2_BLK, NB = H / TB, N2 = 2 * H;
    double sb[2][TB], rb[2][TB];
    auto load = [&](int b, double (&sv)[TB], double (&rv)[TB]) {
#pragma unroll
        for (int q = 0; q < TB; ++q) {
            const int j = b * TB + q, k0 = UP ? kb - j : kb + j;
            const int k = k0 < 0 ? 0 : k0 > N2 - 1 ? N2 - 1 : k0;
            const int kc = k < zlo ? zlo : k > zhi ? zhi : k;
            sv[q] = col[(size_t)kc * s];
            rv[q] = rp[(size_t)k * ps];
        }
    };
    load(0, sb[0], rb[0]);
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        if (b * TB >= nf) break;
        if ((b + 1) * TB < nf) load(b + 1, sb[(b + 1) & 1], rb[(b + 1) & 1]);
#pragma unroll
        for (int q = 0; q < TB; ++q) {
            const int j = b * TB + q, k = UP ? kb - j : kb + j;
            if (j < nf) {
                double bk = (any && k >= z0 && k <= z1) ? sb[b & 1][q] * scale : 0.0;
                if (c0) {
                    if (k == 0) bk += flo;
                    if (k == N2 - 1) bk += fhi;
                }
                /* downward: the coupling to the plane above (a_k); upward: below (c_k) */
                y = (bk - tab[2 * k + (UP ? 1 : 0)] * y) * rb[b & 1][q];
                if (j < LB) lv[(size_t)j * 256] = y;
                else yr[j - LB] = y;
                rlast = rb[b & 1][q];
            }
        }
    }
}

/* the back substitution of that half, from its last step (the plane next to
 * the meeting) back to step 0 and on outward to plane oe: downward x_k = y_k -
 * c'_k x_k+1, upward x_k = u_k - a'_k x_k-1, from x of the neighbour across
 * the middle (xn; the downward half's last step is the meeting's x_M-1
 * itself).  Past step 0 the right-hand side is 0: y_k = 0, or for mode (0, 0)
 * fv gr[k], the face term's share (FftPlan::green).  x is stored on the planes
 * from the meeting out to oe. */
template <int H, int LB, bool UP>
__device__ __forceinline__ void tri2_back(double* col, const double* __restrict__ rp, size_t s, size_t ps,
                                          const double* tab, const double* lv, const double (&yr)[H - LB], int kb,
                                          int nf, int oe, bool c0, double fv, const double* __restrict__ gr,
                                          double xn) {
    constexpr int TB = FX_TRI2_BLK, NB = H / TB, N2 = 2 * H;
    double rb[2][TB];
    auto loadr = [&](int b, double (&rv)[TB]) {
#pragma unroll
        for (int q = 0; q < TB; ++q) {
            const int j = b * TB + q, k0 = UP ? kb - j : kb + j;
            const int k = k0 < 0 ? 0 : k0 > N2 - 1 ? N2 - 1 : k0;
            rv[q] = rp[(size_t)k * ps];
        }
    };
    const int bl = (nf - 1) / TB;   /* the block of the last step */
#pragma unroll
    for (int b = NB - 1; b >= 0; --b) {
        if (b > bl) continue;
        if (b == bl) loadr(b, rb[b & 1]);
        if (b > 0) loadr(b - 1, rb[(b - 1) & 1]);
#pragma unroll
        for (int q = TB - 1; q >= 0; --q) {
            const int j = b * TB + q, k = UP ? kb - j : kb + j;
            if (j < nf) {
                if (UP || j != nf - 1) {
                    const double yj = j < LB ? lv[(size_t)j * 256] : yr[j - LB];
                    xn = yj - (tab[2 * k + (UP ? 0 : 1)] * rb[b & 1][q]) * xn;
                }
                if (UP ? k <= oe : k >= oe) col[(size_t)k * s] = xn;
            }
        }
    }
    for (int k = UP ? kb + 1 : kb - 1; UP ? k <= oe : k >= oe; k += UP ? 1 : -1) {
        const double yk = c0 ? fv * gr[k] : 0.0;
        xn = yk - (tab[2 * k + (UP ? 0 : 1)] * rp[(size_t)k * ps]) * xn;
        col[(size_t)k * s] = xn;
    }
}

/* The z systems of 2H planes solved from both ends (FP64, FftPlan::pivu
 * set): in a workgroup of 256 threads, threads 0..127 eliminate planes
 * kd0..M-1 of 128 modes downward (y_k = (b_k - a_k y_k-1) r_k, the Thomas
 * recurrence of k_fft_tri) and threads 128..255 planes ku1..M of the same
 * modes upward (u_k = (b_k - c_k u_k+1) s_k); they meet at planes M-1 and M,
 *   x_M-1 = (y_M-1 - c'_M-1 u_M) / (1 - c'_M-1 a'_M),  x_M = u_M - a'_M x_M-1,
 * (c'_k = c_k r_k, a'_k = a_k s_k) through LDS, and substitute back outward,
 * storing x on planes o0..o1.  The whole system is kd0 = 0, ku1 = 2H-1,
 * M = H, o0 = 0, o1 = 2H-1.  With a right-hand side on planes z0..z1 only,
 * kd0 = z0, ku1 = z1 and M their middle: below z0 and above z1 y and u are 0
 * (mode (0, 0): the face terms' Green's values times flo, fhi), so the sweeps
 * start there, and the back substitution stops at o0 and o1 (the planes the
 * inverse transforms read; k_fft_tri2_cont continues it when the rest is
 * needed).  Each thread keeps at most H values (half in registers, half in
 * LDS: 64 KB per workgroup for H = 64), so two workgroups fit a CU.  The
 * halves are whole waves, so each runs code with its direction fixed at
 * compile time.  Right-hand side, skipped planes and face terms as
 * k_fft_tri. */
template <int H>
__global__ __launch_bounds__(256, 2) void k_fft_tri2(FftPlan<double> P, double scale, int z0, int z1, double flo,
                                                     double fhi, int kd0, int ku1, int M, int o0, int o1) {
    __shared__ double tab[4 * H];   /* a_k, c_k */
    __shared__ double xch[2][128];
    constexpr int LB = H * FX_TRI2_LDS_EIGHTHS / 8, N2 = 2 * H;
    extern __shared__ __align__(16) unsigned char tri2_lds[];
    double* const lv = reinterpret_cast<double*>(tri2_lds) + threadIdx.x;
    if (P.skip && *P.skip != 0.0) return;
    for (int k = threadIdx.x; k < N2; k += 256) {
        const double iz = P.ih2[2];
        tab[2 * k] = k > 0 ? -(P.bzk[k] * iz) : 0.0;
        tab[2 * k + 1] = k < N2 - 1 ? -(P.bzk[k + 1] * iz) : 0.0;
    }
    __syncthreads();
    const bool up = threadIdx.x >= 128;
    const int lane = threadIdx.x & 127;
    const int c = blockIdx.x * 128 + lane;   /* n0 n1 is a multiple of 128 (the launch checks) */
    const size_t s = (size_t)P.n0 * P.n1;
    const int h0 = P.n0 >> 1, h1 = P.n1 >> 1;
    const int mx = c & (P.n0 - 1), my = c >> P.l0;
    const int mxc = mx <= h0 ? mx : P.n0 - mx, myc = my <= h1 ? my : P.n1 - my;
    const size_t ps = (size_t)(h0 + 1) * (h1 + 1);
    const double* __restrict__ rp = (up ? P.pivu : P.piv) + (size_t)myc * (h0 + 1) + mxc;
    double* const col = P.S + c;
    const int zlo = z0 <= z1 ? z0 : 0, zhi = z0 <= z1 ? z1 : 0;
    const bool any = z0 <= z1, c0 = c == 0;
    double yr[H - LB];
    double y, rlast = 0.0;
    if (up) {
        y = (c0 && ku1 < N2 - 1) ? fhi * P.green[N2 + ku1 + 1] : 0.0;
        tri2_sweep<H, LB, true>(col, rp, s, ps, tab, lv, yr, ku1, ku1 - M + 1, zlo, zhi, any, z0, z1, c0, scale, flo,
                                fhi, y, rlast);
    } else {
        y = (c0 && kd0 > 0) ? flo * P.green[kd0 - 1] : 0.0;
        tri2_sweep<H, LB, false>(col, rp, s, ps, tab, lv, yr, kd0, M - kd0, zlo, zhi, any, z0, z1, c0, scale, flo,
                                 fhi, y, rlast);
    }
    /* the meeting: the upward half hands over u_M and a'_M */
    if (up) {
        xch[0][lane] = y;
        xch[1][lane] = tab[2 * M] * rlast;
    }
    __syncthreads();
    if (!up) {
        const double cp = tab[2 * (M - 1) + 1] * rlast;
        xch[0][lane] = (y - cp * xch[0][lane]) / (1.0 - cp * xch[1][lane]);   /* x_M-1 */
    }
    __syncthreads();
    const double xm = xch[0][lane];   /* x_M-1 */
    if (up) tri2_back<H, LB, true>(col, rp, s, ps, tab, lv, yr, ku1, ku1 - M + 1, o1, c0, fhi, P.green + N2, xm);
    else tri2_back<H, LB, false>(col, rp, s, ps, tab, lv, yr, kd0, M - kd0, o0, c0, flo, P.green, xm);
}

/* The rest of a z solution that k_fft_tri2 stored on planes o0..o1 only
 * (still untransformed there), where the right-hand side was 0 outside
 * them: its back substitution continued from plane o0 down to 0 and from o1
 * up to the last plane, one thread per mode (y_k, u_k as in k_fft_tri2 past
 * its sweeps). */
__global__ __launch_bounds__(256) void k_fft_tri2_cont(FftPlan<double> P, int o0, int o1, double flo, double fhi) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= P.n0 * P.n1) return;
    const int n2 = P.n2;
    const size_t s = (size_t)P.n0 * P.n1;
    const int h0 = P.n0 >> 1, h1 = P.n1 >> 1;
    const int mx = c & (P.n0 - 1), my = c >> P.l0;
    const int mxc = mx <= h0 ? mx : P.n0 - mx, myc = my <= h1 ? my : P.n1 - my;
    const size_t ps = (size_t)(h0 + 1) * (h1 + 1), cls = (size_t)myc * (h0 + 1) + mxc;
    const double iz = P.ih2[2];
    const bool c0 = c == 0;
    double* const col = P.S + c;
    if (o0 > 0) {
        double x = col[(size_t)o0 * s];
        for (int k = o0 - 1; k >= 0; --k) {
            const double ck = -(P.bzk[k + 1] * iz);
            const double yk = c0 ? flo * P.green[k] : 0.0;
            x = yk - (ck * P.piv[k * ps + cls]) * x;
            col[(size_t)k * s] = x;
        }
    }
    if (o1 < n2 - 1) {
        double x = col[(size_t)o1 * s];
        for (int k = o1 + 1; k < n2; ++k) {
            const double ak = -(P.bzk[k] * iz);
            const double uk = c0 ? fhi * P.green[n2 + k] : 0.0;
            x = uk - (ak * P.pivu[k * ps + cls]) * x;
            col[(size_t)k * s] = x;
        }
    }
}

/* FftPlan::green: the Thomas values of mode (0, 0) for a unit right-hand
 * side on the first plane (downward, green[k]) and on the last (upward,
 * green[n2 + k]), with the pivots of the tables; one thread */
__global__ void k_fft_face_green(FftPlan<double> P, double* g) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const int n2 = P.n2;
    const size_t ps = (size_t)(P.n0 / 2 + 1) * (P.n1 / 2 + 1);
    const double iz = P.ih2[2];
    double y = 0.0;
    for (int k = 0; k < n2; ++k) {
        const double a = k > 0 ? -(P.bzk[k] * iz) : 0.0;
        y = ((k == 0 ? 1.0 : 0.0) - a * y) * P.piv[k * ps];
        g[k] = y;
    }
    double u = 0.0;
    for (int k = n2 - 1; k >= 0; --k) {
        const double cc = k < n2 - 1 ? -(P.bzk[k + 1] * iz) : 0.0;
        u = ((k == n2 - 1 ? 1.0 : 0.0) - cc * u) * P.pivu[k * ps];
        g[n2 + k] = u;
    }
}

/* One thread per lateral mode: the tridiagonal system in z by the Thomas
 * algorithm with the pivots of the table, in place on S, the right-hand side
 * scaled by `scale` (the 1/(n0 n1) of the two inverse transforms, exact for
 * powers of two):
 *   y_k = (b_k - a_k y_k-1) r_k,    x_k = y_k - c_k r_k x_k+1.
 * Planes of S outside [z0, z1] were not transformed and count as 0; the
 * Dirichlet face terms of a skipped first or last plane (flo, fhi, already
 * scaled) enter mode (0, 0), the only mode of a constant plane.  y of
 * planes [0, NZ - RB) waits in LDS for the backward sweep, the rest in
 * registers; S and r stream in blocks of TRI_BLK planes, the next block's
 * loads in flight while one is eliminated.  Every load is unconditional
 * (clamped plane, the value selected afterwards).  NZ = 0: any n2, y
 * written back to S between the sweeps. */
template <class T, int NZ, int RB, int THR>
__global__ __launch_bounds__(THR) void k_fft_tri(FftPlan<T> P, T scale, int z0, int z1, T flo, T fhi) {
    extern __shared__ __align__(16) unsigned char tri_lds[];
    if (P.skip && *P.skip != 0.0) return;
    constexpr int LB = NZ > RB ? NZ - RB : 0;
    constexpr int TB = TRI_BLK;
    const int s = P.n0 * P.n1, n2 = P.n2;
    T* const lv = reinterpret_cast<T*>(tri_lds) + threadIdx.x;                 /* y of plane k: lv[k THR] */
    T* const tab = reinterpret_cast<T*>(tri_lds) + (size_t)LB * THR;             /* a_k, c_k */
    for (int k = threadIdx.x; k < n2; k += THR) {
        const T iz = P.ih2[2];
        tab[2 * k] = k > 0 ? -(P.bzk[k] * iz) : T(0);
        tab[2 * k + 1] = k < n2 - 1 ? -(P.bzk[k + 1] * iz) : T(0);
    }
    __syncthreads();
    const int c = blockIdx.x * THR + threadIdx.x;
    if (c >= s) return;
    const int h0 = P.n0 >> 1, h1 = P.n1 >> 1;
    const int mx = c & (P.n0 - 1), my = c >> P.l0;
    const int mxc = mx <= h0 ? mx : P.n0 - mx, myc = my <= h1 ? my : P.n1 - my;
    const size_t ps = (size_t)(h0 + 1) * (h1 + 1);
    const T* __restrict__ rp = P.piv + (size_t)myc * (h0 + 1) + mxc;
    T* const col = P.S + c;
    const int zlo = z0 <= z1 ? z0 : 0, zhi = z0 <= z1 ? z1 : 0;
    const bool any = z0 <= z1;
    /* right-hand side of plane k from the loaded value v */
    auto rhs = [&](int k, T v) -> T {
        T b = (any && k >= z0 && k <= z1) ? v * scale : T(0);
        if (c == 0) {
            if (k == 0) b += flo;
            if (k == n2 - 1) b += fhi;
        }
        return b;
    };
    if constexpr (NZ > 0) {
        constexpr int NB = NZ / TB;
        T sb[2][TB], rb[2][TB], yr[RB > 0 ? RB : 1];
        auto load = [&](int b, T (&sv)[TB], T (&rv)[TB]) {
#pragma unroll
            for (int q = 0; q < TB; ++q) {
                const int k = b * TB + q;
                const int kc = k < zlo ? zlo : k > zhi ? zhi : k;
                sv[q] = col[(size_t)kc * s];
                rv[q] = rp[(size_t)(k < n2 ? k : n2 - 1) * ps];
            }
        };
        T y = T(0);
        load(0, sb[0], rb[0]);
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            if (b + 1 < NB) load(b + 1, sb[(b + 1) & 1], rb[(b + 1) & 1]);
#pragma unroll
            for (int q = 0; q < TB; ++q) {
                const int k = b * TB + q;
                if (k < n2) {
                    y = (rhs(k, sb[b & 1][q]) - tab[2 * k] * y) * rb[b & 1][q];
                    if (k < LB) lv[(size_t)k * THR] = y;
                    else yr[k - LB] = y;
                }
            }
        }
        auto loadr = [&](int b, T (&rv)[TB]) {
#pragma unroll
            for (int q = 0; q < TB; ++q) {
                const int k = b * TB + q;
                rv[q] = rp[(size_t)(k < n2 ? k : n2 - 1) * ps];
            }
        };
        T xn = T(0);
        loadr(NB - 1, rb[(NB - 1) & 1]);
#pragma unroll
        for (int b = NB - 1; b >= 0; --b) {
            if (b > 0) loadr(b - 1, rb[(b - 1) & 1]);
#pragma unroll
            for (int q = TB - 1; q >= 0; --q) {
                const int k = b * TB + q;
                if (k < n2) {
                    const T yk = k < LB ? lv[(size_t)k * THR] : yr[k - LB];
                    xn = yk - (tab[2 * k + 1] * rb[b & 1][q]) * xn;
                    col[(size_t)k * s] = xn;
                }
            }
        }
    } else {
        T y = T(0);
        for (int k = 0; k < n2; ++k) {
            const int kc = k < zlo ? zlo : k > zhi ? zhi : k;
            y = (rhs(k, col[(size_t)kc * s]) - tab[2 * k] * y) * rp[(size_t)k * ps];
            col[(size_t)k * s] = y;
        }
        T xn = T(0);
        for (int k = n2 - 1; k >= 0; --k) {
            xn = col[(size_t)k * s] - (tab[2 * k + 1] * rp[(size_t)k * ps]) * xn;
            col[(size_t)k * s] = xn;
        }
    }
}

/* sequences per workgroup: thr / max(N1, N2), and 2R no more than the
 * other lateral extent */
inline int seqs_per_block(int thr, int n1f, int n2f, int other) {
    int R = thr / (n1f > n2f ? n1f : n2f);
    if (2 * R > other) R = other / 2;
    return R;
}

template <class T>
size_t fft_lds_bytes(int N1, int N2, int R, int pad, int twl, bool half = false) {
    return ((twl ? (size_t)N1 * N2 : 0) + (half ? 1 : 2) * (size_t)R * N2 * (N1 + pad)) * sizeof(T);
}

/* raise the dynamic LDS limit of a kernel to what a launch needs, where that
 * is more than 64 KB (the limit plus the kernel's static LDS must stay
 * within the 160 KB of a CU, so it is raised to the need, not to 160 KB);
 * set: the largest limit set so far */
template <class K>
void lds_limit(K* kernel, size_t lds, size_t& set) {
    if (lds > 64 * 1024 && lds > set) {
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)lds);
        set = lds;
    }
}

template <class T, class TIN, class TOUT, int N1, int N2>
void rows_launch(hipStream_t s, const FftPlan<T>& P, bool fwd, const void* in, void* out, double blo, double bhi, int add,
                 int kz0, int nk, const double* dotv, double* dotp) {
    const int R = seqs_per_block(ROWS_THR, N1, N2, P.n1);
    /* the split four-step path (real, then imaginary parts across) needs half the buffer */
    const bool half = FX_FFT_ROWS_SPLIT && FX_FFT_REG_HARTLEY && N1 == 16 && (fwd || !dotp);
    const size_t lds = fft_lds_bytes<T>(N1, N2, R, 1, 1, half) + (half ? (size_t)R * FX_FFT_ROWS_SPAD * sizeof(T) : 0);
    const dim3 g(P.n1 / (2 * R), nk);
    static size_t lf = 0, li = 0;
    if (fwd) lds_limit(&k_fft_rows<T, TIN, T, true, N1, N2>, lds, lf);
    else lds_limit(&k_fft_rows<T, T, TOUT, false, N1, N2>, lds, li);
    if (fwd)
        hipLaunchKernelGGL((k_fft_rows<T, TIN, T, true, N1, N2>), g, dim3(ROWS_THR), lds, s, P,
                           static_cast<const TIN*>(in), static_cast<T*>(out), R, blo, bhi, 0, kz0, nullptr, nullptr);
    else
        hipLaunchKernelGGL((k_fft_rows<T, T, TOUT, false, N1, N2>), g, dim3(ROWS_THR), lds, s, P,
                           static_cast<const T*>(in), static_cast<TOUT*>(out), R, 0.0, 0.0, add, kz0, dotv, dotp);
}

/* workgroups of a rows pass over nk planes (the partials of its dot product) */
inline int rows_blocks(int l0, int n1, int nk) {
    const int n1f = l0 <= 2 ? 2 : l0 <= 4 ? 4 : l0 <= 6 ? 8 : l0 <= 8 ? 16 : 32;
    const int n2f = (l0 <= 3 ? 2 : l0 <= 5 ? 4 : l0 <= 7 ? 8 : l0 <= 9 ? 16 : 32);
    return n1 / (2 * seqs_per_block(ROWS_THR, n1f, n2f, n1)) * nk;
}

template <class T, int N1, int N2>
void cols_launch(hipStream_t s, const FftPlan<T>& P, int kz0, int nk) {
    const int R = seqs_per_block(COLS_THR, N1, N2, P.n0);
    const bool half = FX_FFT_COLS_SPLIT && FX_FFT_REG_HARTLEY && N1 == 16;
    const size_t lds = fft_lds_bytes<T>(N1, N2, R, FX_FFT_COLS_PAD, FX_FFT_COLS_TWL, half) +
                       (half ? (size_t)R * FX_FFT_COLS_SPAD * sizeof(T) : 0);
    static size_t lc = 0;
    lds_limit(&k_fft_cols<T, N1, N2>, lds, lc);
    hipLaunchKernelGGL((k_fft_cols<T, N1, N2>), dim3(P.n0 / (2 * R), nk), dim3(COLS_THR), lds, s, P, R, kz0);
}

/* n = N1 N2 with N1 >= N2, both powers of two up to 32; planes kz0 .. kz0 + nk - 1 */
template <class T, class TIN, class TOUT>
void rows(hipStream_t s, const FftPlan<T>& P, bool fwd, const void* in, void* out, double blo, double bhi, int add,
          int kz0, int nk, const double* dotv = nullptr, double* dotp = nullptr) {
    if (nk <= 0) return;
    switch (P.l0) {
    case 2: rows_launch<T, TIN, TOUT, 2, 2>(s, P, fwd, in, out, blo, bhi, add, kz0, nk, dotv, dotp); break;
    case 3: rows_launch<T, TIN, TOUT, 4, 2>(s, P, fwd, in, out, blo, bhi, add, kz0, nk, dotv, dotp); break;
    case 4: rows_launch<T, TIN, TOUT, 4, 4>(s, P, fwd, in, out, blo, bhi, add, kz0, nk, dotv, dotp); break;
    case 5: rows_launch<T, TIN, TOUT, 8, 4>(s, P, fwd, in, out, blo, bhi, add, kz0, nk, dotv, dotp); break;
    case 6: rows_launch<T, TIN, TOUT, 8, 8>(s, P, fwd, in, out, blo, bhi, add, kz0, nk, dotv, dotp); break;
    case 7: rows_launch<T, TIN, TOUT, 16, 8>(s, P, fwd, in, out, blo, bhi, add, kz0, nk, dotv, dotp); break;
    case 8: rows_launch<T, TIN, TOUT, 16, 16>(s, P, fwd, in, out, blo, bhi, add, kz0, nk, dotv, dotp); break;
    case 9: rows_launch<T, TIN, TOUT, 32, 16>(s, P, fwd, in, out, blo, bhi, add, kz0, nk, dotv, dotp); break;
    default: rows_launch<T, TIN, TOUT, 32, 32>(s, P, fwd, in, out, blo, bhi, add, kz0, nk, dotv, dotp); break;
    }
}

template <class T>
void cols(hipStream_t s, const FftPlan<T>& P, int kz0, int nk) {
    if (nk <= 0) return;
    switch (P.l1) {
    case 2: cols_launch<T, 2, 2>(s, P, kz0, nk); break;
    case 3: cols_launch<T, 4, 2>(s, P, kz0, nk); break;
    case 4: cols_launch<T, 4, 4>(s, P, kz0, nk); break;
    case 5: cols_launch<T, 8, 4>(s, P, kz0, nk); break;
    case 6: cols_launch<T, 8, 8>(s, P, kz0, nk); break;
    case 7: cols_launch<T, 16, 8>(s, P, kz0, nk); break;
    case 8: cols_launch<T, 16, 16>(s, P, kz0, nk); break;
    case 9: cols_launch<T, 32, 16>(s, P, kz0, nk); break;
    default: cols_launch<T, 32, 32>(s, P, kz0, nk); break;
    }
}

template <class T, int NZ, int RB, int THR>
void tri_launch_thr(hipStream_t s, const FftPlan<T>& P, T scale, int z0, int z1, T flo, T fhi) {
    constexpr int LB = NZ > RB ? NZ - RB : 0;
    const size_t lds = ((size_t)LB * THR + 2 * (size_t)P.n2) * sizeof(T);
    static bool attr = false;   /* the LDS limit is raised once per kernel */
    if (lds > 64 * 1024 && !attr) {
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_fft_tri<T, NZ, RB, THR>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
        attr = true;
    }
    hipLaunchKernelGGL((k_fft_tri<T, NZ, RB, THR>), dim3((P.n0 * P.n1 + THR - 1) / THR), dim3(THR), lds, s, P, scale,
                       z0, z1, flo, fhi);
}

/* one thread per mode: workgroups of TRI_THR threads, or of 64 where that
 * would leave CUs idle (fewer modes than 256 workgroups of TRI_THR) */
template <class T, int NZ, int RB>
void tri_launch(hipStream_t s, const FftPlan<T>& P, T scale, int z0, int z1, T flo, T fhi) {
    if (P.n0 * P.n1 < 256 * TRI_THR) tri_launch_thr<T, NZ, RB, 64>(s, P, scale, z0, z1, flo, fhi);
    else tri_launch_thr<T, NZ, RB, TRI_THR>(s, P, scale, z0, z1, flo, fhi);
}

/* returns true where the two-ended solve stored x on planes o0..o1 only (the
 * rest by fft_finish's continuation): an FP64 plan with upward pivots, a
 * right-hand side on z0..z1 only and o0 <= z0, z1 <= o1, not every plane */
template <class T>
bool tri(hipStream_t s, const FftPlan<T>& P, T scale, int z0, int z1, T flo, T fhi, int o0 = 0, int o1 = -1) {
    if constexpr (std::is_same<T, double>::value) {
        if (P.pivu && P.n2 == 128 && (P.n0 * P.n1) % 128 == 0) {
            const int N2 = P.n2;
            const bool part = z0 <= z1 && P.green && o0 <= o1 && o0 <= z0 && z1 <= o1 && (o0 > 0 || o1 < N2 - 1);
            int kd0 = 0, ku1 = N2 - 1;
            if (part) {
                kd0 = z0;
                ku1 = z1;
                if (ku1 == kd0) { if (ku1 < N2 - 1) ++ku1; else --kd0; }   /* one plane a side at least */
                /* one plane more each side: the continuation starts there, since
                 * the inverse transforms overwrite planes o0..o1 in place */
                o0 = o0 > 0 ? o0 - 1 : 0;
                o1 = o1 < N2 - 1 ? o1 + 1 : N2 - 1;
            } else {
                o0 = 0;
                o1 = N2 - 1;
            }
            const int M = (kd0 + ku1 + 1) / 2;   /* kd0 < M <= ku1, at most N2/2 planes a side */
            const size_t lds = (size_t)(64 * FX_TRI2_LDS_EIGHTHS / 8) * 256 * sizeof(double);
            static bool attr = false;   /* 64 KB of dynamic LDS: the limit raised once */
            if (!attr) {
                (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_fft_tri2<64>),
                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
                attr = true;
            }
            hipLaunchKernelGGL(k_fft_tri2<64>, dim3(P.n0 * P.n1 / 128), dim3(256), lds, s, P, scale, z0, z1, flo, fhi,
                               kd0, ku1, M, o0, o1);
            return part;
        }
    }
    if (P.n2 <= 16) tri_launch<T, 16, 16>(s, P, scale, z0, z1, flo, fhi);
    else if (P.n2 <= 32) tri_launch<T, 32, 32>(s, P, scale, z0, z1, flo, fhi);
    else if (P.n2 <= 64) tri_launch<T, 64, 64>(s, P, scale, z0, z1, flo, fhi);
    else if (P.n2 <= 128) tri_launch<T, 128, TRI_REG>(s, P, scale, z0, z1, flo, fhi);
    else tri_launch<T, 0, 0>(s, P, scale, z0, z1, flo, fhi);
    return false;
}

/* planes of b outside [z0, z1] hold zeros: their forward transforms are
 * skipped (z1 < z0: none is, every plane is transformed); phi is formed on
 * planes [o0, o1] only (o1 < o0: all), the other planes' solution left
 * transformed in S for fft_finish */
template <class T, class TIN, class TOUT>
bool fft_solve_impl(hipStream_t s, const FftPlan<T>& P, const TIN* b, TOUT* phi, double blo, double bhi, int add,
                    int z0, int z1, int o0 = 0, int o1 = -1, const double* dotv = nullptr, double* dotp = nullptr,
                    double* face = nullptr) {
    const T scale = T(1) / T(P.n0 * P.n1);
    if (z1 < z0) { z0 = 0; z1 = P.n2 - 1; }
    if (o1 < o0) { o0 = 0; o1 = P.n2 - 1; }
    const int nk = z1 - z0 + 1;
    /* the Dirichlet face terms of a skipped face plane go to the z solve */
    const T flo = z0 > 0 ? T(blo) : T(0), fhi = z1 < P.n2 - 1 ? T(bhi) : T(0);
    rows<T, TIN, TOUT>(s, P, true, b, P.S, blo, bhi, 0, z0, nk);
    cols<T>(s, P, z0, nk);
    const bool part = tri<T>(s, P, scale, z0, z1, flo, fhi, o0, o1);
    if (face) { face[0] = (double)flo; face[1] = (double)fhi; }
    cols<T>(s, P, o0, o1 - o0 + 1);
    rows<T, TIN, TOUT>(s, P, false, P.S, phi, 0.0, 0.0, add, o0, o1 - o0 + 1, dotv, dotp);
    return part;
}

}  // namespace

namespace fxk {

bool fft_size_ok(int n) { return n >= 4 && n <= 1024 && (n & (n - 1)) == 0; }

size_t fft_pivot_count(int n0, int n1, int n2) { return (size_t)(n0 / 2 + 1) * (n1 / 2 + 1) * n2; }

/* the recurrences are sequential in the plane: one thread per mode class in
 * workgroups of 64, so that the classes spread over the CUs */
void fft_pivots_up(hipStream_t s, const fxd::FftPlan<double>& P, double* pivu) {
    const int n = (P.n0 / 2 + 1) * (P.n1 / 2 + 1);
    const size_t lds = (4 * (size_t)P.n2 + 1) * sizeof(double);
    hipLaunchKernelGGL(k_fft_pivots_up, dim3((n + 63) / 64), dim3(64), lds, s, P, pivu);
}
void fft_face_green(hipStream_t s, const fxd::FftPlan<double>& P, double* g) {
    hipLaunchKernelGGL(k_fft_face_green, dim3(1), dim3(64), 0, s, P, g);
}
void fft_pivots(hipStream_t s, const fxd::FftPlan<double>& P, double* piv64, float* piv32, const double* abar) {
    const int n = (P.n0 / 2 + 1) * (P.n1 / 2 + 1);
    const size_t lds = (4 * (size_t)P.n2 + 1) * sizeof(double);
    hipLaunchKernelGGL(k_fft_pivots<double>, dim3((n + 63) / 64), dim3(64), lds, s, P, piv64, abar);
    if (piv32) hipLaunchKernelGGL(k_fft_pivots<float>, dim3((n + 63) / 64), dim3(64), lds, s, P, piv32, abar);
}

bool fft_solve(hipStream_t s, const fxd::FftPlan<double>& P, const double* b, double* phi, double blo, double bhi, int z0,
               int z1, int o0, int o1, double* face) {
    return fft_solve_impl<double, double, double>(s, P, b, phi, blo, bhi, 0, z0, z1, o0, o1, nullptr, nullptr, face);
}
void fft_solve_dot(hipStream_t s, const fxd::FftPlan<double>& P, const double* b, double* phi, double* dotp) {
    fft_solve_impl<double, double, double>(s, P, b, phi, 0.0, 0.0, 0, 0, -1, 0, -1, b, dotp);
}
int fft_dot_parts(const fxd::FftPlan<double>& P) { return rows_blocks(P.l0, P.n1, P.n2); }
void fft_finish(hipStream_t s, const fxd::FftPlan<double>& P, double* phi, int o0, int o1, const double* face) {
    if (face && (o0 > 1 || o1 < P.n2 - 2)) {   /* the solve stored planes o0 - 1 and o1 + 1 too */
        const int n = P.n0 * P.n1;
        hipLaunchKernelGGL(k_fft_tri2_cont, dim3((n + 255) / 256), dim3(256), 0, s, P, o0 > 0 ? o0 - 1 : 0,
                           o1 < P.n2 - 1 ? o1 + 1 : P.n2 - 1, face[0], face[1]);
    }
    if (o0 > 0) {
        cols<double>(s, P, 0, o0);
        rows<double, double, double>(s, P, false, P.S, phi, 0.0, 0.0, 0, 0, o0);
    }
    if (o1 < P.n2 - 1) {
        cols<double>(s, P, o1 + 1, P.n2 - 1 - o1);
        rows<double, double, double>(s, P, false, P.S, phi, 0.0, 0.0, 0, o1 + 1, P.n2 - 1 - o1);
    }
}
void fft_solve(hipStream_t s, const fxd::FftPlan<float>& P, const double* b, double* phi, double blo, double bhi, int z0,
               int z1) {
    fft_solve_impl<float, double, double>(s, P, b, phi, blo, bhi, 0, z0, z1);
}
void fft_solve_add(hipStream_t s, const fxd::FftPlan<float>& P, const float* r, double* phi) {
    fft_solve_impl<float, float, double>(s, P, r, phi, 0.0, 0.0, 1, 0, -1);
}

void fft_pass(hipStream_t s, const fxd::FftPlan<double>& P, int pass, const double* b, double* phi, double blo,
              double bhi) {
    switch (pass) {
    case 0: rows<double, double, double>(s, P, true, b, P.S, blo, bhi, 0, 0, P.n2); break;
    case 1: cols<double>(s, P, 0, P.n2); break;
    case 2: tri<double>(s, P, 1.0 / (double)(P.n0 * P.n1), 0, P.n2 - 1, 0.0, 0.0); break;
    default: rows<double, double, double>(s, P, false, P.S, phi, 0.0, 0.0, 0, 0, P.n2); break;
    }
}

}  // namespace fxk
