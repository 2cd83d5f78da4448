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
 *
 * The Hartley transforms of two real rows come from one complex FFT of the
 * row pair packed as re + i im (radix-2, in LDS): with Z = FFT(re + i im),
 * A + iB = Z[m] and C + iD = Z[-m],
 *   H(re)[m] = (A - B + C + D) / 2,   H(im)[m] = (A + B - C + D) / 2.
 * n0 and n1 are powers of two (4 .. 1024).  The result is exact to rounding:
 * no iteration and no stopping test.
 */
#include "fx_kernels.h"

namespace {

using fxd::FftPlan;

constexpr int FFT_THR = 128;  /* threads of a transform workgroup: R sequences x max(N1, N2) */
constexpr int TRI_THR = 256;
constexpr int TRI_BLK = 16;   /* planes per register block of the tridiagonal solve */
constexpr int TRI_REG = 64;   /* planes of a column held in registers (the rest in LDS) */

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
template <class T, int N1, int N2, class Load>
__device__ __forceinline__ void four_step(int R, const T* tw, T* Yr, T* Yi, Load load, int step1_s_fast) {
    constexpr int n = N1 * N2, YS = N1 + 1;
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

template <class T>
__device__ __forceinline__ void load_tw(T* tw, const T* g, int n) {
    for (int q = threadIdx.x; q < n; q += blockDim.x) tw[q] = g[q];
}

/* Hartley transform along x of rows j0 .. j0 + 2R - 1 of plane k =
 * blockIdx.y: rows 2s and 2s + 1 of the block are sequence s.  FWD: input b
 * (TIN) plus the Dirichlet face terms on the first and last plane, output
 * S.  !FWD: input S, output phi (TOUT), stored or, with add, added. */
template <class T, class TIN, class TOUT, bool FWD, int N1, int N2>
__global__ __launch_bounds__(FFT_THR) void k_fft_rows(FftPlan<T> P, const TIN* __restrict__ in, TOUT* __restrict__ out,
                                                      int R, double blo, double bhi, int add) {
    extern __shared__ __align__(16) unsigned char fft_lds[];
    constexpr int n = N1 * N2;
    const int k = blockIdx.y, j0 = blockIdx.x * 2 * R;
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
    four_step<T, N1, N2>(R, tw, Yr, Yi, load, 0);
    for (int t = threadIdx.x; t < R * n; t += blockDim.x) {
        const int s = t / n, kk = t % n;
        T h0, h1;
        hartley2<T>(Yr, Yi, s, n, kk, h0, h1);
        const size_t o = row0 + (size_t)(2 * s) * n + kk;
        if (FWD || !add) {
            out[o] = (TOUT)h0;
            out[o + n] = (TOUT)h1;
        } else {
            out[o] = out[o] + (TOUT)h0;
            out[o + n] = out[o + n] + (TOUT)h1;
        }
    }
}

template <class T> struct Vec2;
template <> struct Vec2<double> { using t = double2; };
template <> struct Vec2<float> { using t = float2; };

/* Hartley transform along y, in place, of the column pairs (i0 + 2s,
 * i0 + 2s + 1), s < R, of plane k = blockIdx.y: each row segment of 2R
 * values is read and written whole (one vector of two per thread). */
template <class T, int N1, int N2>
__global__ __launch_bounds__(FFT_THR) void k_fft_cols(FftPlan<T> P, int R) {
    extern __shared__ __align__(16) unsigned char fft_lds[];
    using V2 = typename Vec2<T>::t;
    constexpr int n = N1 * N2;
    const int n0 = P.n0, k = blockIdx.y, i0 = blockIdx.x * 2 * R;
    T* const tw = reinterpret_cast<T*>(fft_lds);
    T* const Yr = tw + n;
    T* const Yi = Yr + R * N2 * (N1 + 1);
    load_tw<T>(tw, P.tw1, n);
    __syncthreads();
    T* const S = P.S + (size_t)k * n * n0 + i0;
    auto load = [&](int s, int a, int b, T& re, T& im) {
        const V2 v = *reinterpret_cast<const V2*>(S + (size_t)(N2 * a + b) * n0 + 2 * s);
        re = v.x;
        im = v.y;
    };
    four_step<T, N1, N2>(R, tw, Yr, Yi, load, 1);
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

/* One thread per lateral mode: the tridiagonal system in z (Thomas
 * algorithm), in place on S, the right-hand side scaled by `scale` (the
 * 1/(n0 n1) of the two inverse transforms, exact for powers of two).
 * NZ > 0 (n2 <= NZ): the whole column is loaded into registers at once,
 * eliminated there and stored once; the c' of the recurrence are kept every
 * TRI_BLK planes and recomputed block by block for the back substitution
 * with the same operations.  NZ = 0: any n2, block by block through S, with
 * the checkpoints in P.ck. */
/* 1/t for the positive pivots of the tridiagonal recurrence: the hardware
 * reciprocal refined by Newton steps e = 1 - t r, r += r e (two for FP64,
 * one for FP32), shorter than the correctly rounded division sequence and
 * accurate to an ulp or two; the forward and the backward sweeps compute
 * the same c' from the same operations. */
__device__ __forceinline__ double tri_recip(double t) {
    double r = __builtin_amdgcn_rcp(t);
    double e = __builtin_fma(-t, r, 1.0);
    r = __builtin_fma(r, e, r);
    e = __builtin_fma(-t, r, 1.0);
    return __builtin_fma(r, e, r);
}
__device__ __forceinline__ float tri_recip(float t) {
    const float r = __builtin_amdgcn_rcpf(t);
    const float e = __builtin_fmaf(-t, r, 1.0f);
    return __builtin_fmaf(r, e, r);
}

/* Row k of a mode's tridiagonal system, a x_k-1 + d x_k + cc x_k+1, from the
 * per-plane table staged in LDS: tab[5k..5k+4] = bx_k, by_k, the z part of
 * the diagonal, the sub- and the super-diagonal. */
template <class T>
struct TriCoef {
    const T* tab;
    T lx, ly;
    __device__ __forceinline__ void operator()(int k, T& a, T& d, T& cc) const {
        const T* r = tab + 5 * k;
        T diag = r[0] * lx;
        diag += r[1] * ly;
        diag += r[2];
        a = r[3];
        cc = r[4];
        d = diag;
    }
};

template <class T, int NZ>
__global__ __launch_bounds__(TRI_THR, 1) void k_fft_tri(FftPlan<T> P, T scale) {
    extern __shared__ __align__(16) unsigned char tri_lds[];
    const int s = P.n0 * P.n1, n2 = P.n2;
    const int c = blockIdx.x * TRI_THR + threadIdx.x;
    constexpr int LBV = NZ > TRI_REG ? NZ - TRI_REG : 0;
    T* const tab = reinterpret_cast<T*>(tri_lds) + (size_t)LBV * TRI_THR;
    for (int k = threadIdx.x; k < n2; k += TRI_THR) {
        const T iz = P.ih2[2];
        const T wzl = P.bzk[k] * iz, wzh = P.bzk[k + 1] * iz;
        tab[5 * k + 0] = P.bxk[k];
        tab[5 * k + 1] = P.byk[k];
        tab[5 * k + 2] = (k > 0 ? wzl : (P.dlo ? T(2) * wzl : T(0))) + (k < n2 - 1 ? wzh : (P.dhi ? T(2) * wzh : T(0)));
        tab[5 * k + 3] = k > 0 ? -wzl : T(0);
        tab[5 * k + 4] = k < n2 - 1 ? -wzh : T(0);
    }
    __syncthreads();
    if (c >= s) return;
    TriCoef<T> coef{tab, P.mux[c & (P.n0 - 1)] * P.ih2[0], P.muy[c >> P.l0] * P.ih2[1]};
    T* const col = P.S + c;
    if constexpr (NZ > 0) {
        /* planes [0, LB) go through LDS (one value per thread and plane),
         * planes [LB, NZ) stay in registers */
        constexpr int RB = NZ < TRI_REG ? NZ : TRI_REG, LB = NZ - RB;
        constexpr int NB = (NZ + TRI_BLK - 1) / TRI_BLK;
        T* const lv = reinterpret_cast<T*>(tri_lds) + threadIdx.x;
        T v[RB], ck[NB];
#pragma unroll
        for (int q = 0; q < RB; ++q)
            if (LB + q < n2) v[q] = col[(size_t)(LB + q) * s];
        T cprev = T(0), dprev = T(0);
        auto elim = [&](int k, T x) -> T {
            T a, d, cc;
            coef(k, a, d, cc);
            const T r = tri_recip(d - a * cprev);
            cprev = cc * r;
            dprev = (x * scale - a * dprev) * r;
            return dprev;
        };
#pragma unroll
        for (int b0 = 0; b0 < LB; b0 += TRI_BLK) {
            T w[TRI_BLK];
#pragma unroll
            for (int q = 0; q < TRI_BLK; ++q)
                if (b0 + q < n2) w[q] = col[(size_t)(b0 + q) * s];
            ck[b0 / TRI_BLK] = cprev;
#pragma unroll
            for (int q = 0; q < TRI_BLK; ++q)
                if (b0 + q < n2) lv[(b0 + q) * TRI_THR] = elim(b0 + q, w[q]);
        }
#pragma unroll
        for (int q = 0; q < RB; ++q) {
            const int k = LB + q;
            if (k % TRI_BLK == 0) ck[k / TRI_BLK] = cprev;
            if (k < n2) v[q] = elim(k, v[q]);
        }
        T xn = T(0);
#pragma unroll
        for (int b = NB - 1; b >= 0; --b) {
            T cv[TRI_BLK];
            T cp = ck[b];
#pragma unroll
            for (int q = 0; q < TRI_BLK; ++q) {
                const int k = b * TRI_BLK + q;
                if (k < n2) {
                    T a, d, cc;
                    coef(k, a, d, cc);
                    const T r = tri_recip(d - a * cp);
                    cp = cc * r;
                    cv[q] = cp;
                }
            }
#pragma unroll
            for (int q = TRI_BLK - 1; q >= 0; --q) {
                const int k = b * TRI_BLK + q;
                if (k < n2) {
                    if (k >= LB) {
                        xn = v[k - LB] - cv[q] * xn;
                        v[k - LB] = xn;
                    } else {
                        xn = lv[k * TRI_THR] - cv[q] * xn;
                        col[(size_t)k * s] = xn;
                    }
                }
            }
        }
#pragma unroll
        for (int q = 0; q < RB; ++q)
            if (LB + q < n2) col[(size_t)(LB + q) * s] = v[q];
    } else {
        const int nb = (n2 + TRI_BLK - 1) / TRI_BLK;
        T cprev = T(0), dprev = T(0);
        for (int b = 0; b < nb; ++b) {
            const int k0 = b * TRI_BLK;
            P.ck[(size_t)b * s + c] = cprev;
            T v[TRI_BLK];
#pragma unroll
            for (int q = 0; q < TRI_BLK; ++q)
                if (k0 + q < n2) v[q] = col[(size_t)(k0 + q) * s];
#pragma unroll
            for (int q = 0; q < TRI_BLK; ++q) {
                if (k0 + q < n2) {
                    T a, d, cc;
                    coef(k0 + q, a, d, cc);
                    const T r = tri_recip(d - a * cprev);
                    cprev = cc * r;
                    dprev = (v[q] * scale - a * dprev) * r;
                    v[q] = dprev;
                }
            }
#pragma unroll
            for (int q = 0; q < TRI_BLK; ++q)
                if (k0 + q < n2) col[(size_t)(k0 + q) * s] = v[q];
        }
        T xn = T(0);
        for (int b = nb - 1; b >= 0; --b) {
            const int k0 = b * TRI_BLK;
            T cp = P.ck[(size_t)b * s + c];
            T cv[TRI_BLK], v[TRI_BLK];
#pragma unroll
            for (int q = 0; q < TRI_BLK; ++q)
                if (k0 + q < n2) v[q] = col[(size_t)(k0 + q) * s];
#pragma unroll
            for (int q = 0; q < TRI_BLK; ++q) {
                if (k0 + q < n2) {
                    T a, d, cc;
                    coef(k0 + q, a, d, cc);
                    const T r = tri_recip(d - a * cp);
                    cp = cc * r;
                    cv[q] = cp;
                }
            }
#pragma unroll
            for (int q = TRI_BLK - 1; q >= 0; --q) {
                if (k0 + q < n2) {
                    xn = v[q] - cv[q] * xn;
                    v[q] = xn;
                }
            }
#pragma unroll
            for (int q = 0; q < TRI_BLK; ++q)
                if (k0 + q < n2) col[(size_t)(k0 + q) * s] = v[q];
        }
    }
}

/* sequences per workgroup: FFT_THR / max(N1, N2), and 2R no more than the
 * other lateral extent */
inline int seqs_per_block(int n1f, int n2f, int other) {
    int R = FFT_THR / (n1f > n2f ? n1f : n2f);
    if (2 * R > other) R = other / 2;
    return R;
}

template <class T>
size_t fft_lds_bytes(int N1, int N2, int R) {
    return ((size_t)N1 * N2 + 2 * (size_t)R * N2 * (N1 + 1)) * sizeof(T);
}

template <class T, class TIN, class TOUT, int N1, int N2>
void rows_launch(hipStream_t s, const FftPlan<T>& P, bool fwd, const void* in, void* out, double blo, double bhi, int add) {
    const int R = seqs_per_block(N1, N2, P.n1);
    const size_t lds = fft_lds_bytes<T>(N1, N2, R);
    const dim3 g(P.n1 / (2 * R), P.n2);
    if (fwd)
        hipLaunchKernelGGL((k_fft_rows<T, TIN, T, true, N1, N2>), g, dim3(FFT_THR), lds, s, P,
                           static_cast<const TIN*>(in), static_cast<T*>(out), R, blo, bhi, 0);
    else
        hipLaunchKernelGGL((k_fft_rows<T, T, TOUT, false, N1, N2>), g, dim3(FFT_THR), lds, s, P,
                           static_cast<const T*>(in), static_cast<TOUT*>(out), R, 0.0, 0.0, add);
}

template <class T, int N1, int N2>
void cols_launch(hipStream_t s, const FftPlan<T>& P) {
    const int R = seqs_per_block(N1, N2, P.n0);
    const size_t lds = fft_lds_bytes<T>(N1, N2, R);
    hipLaunchKernelGGL((k_fft_cols<T, N1, N2>), dim3(P.n0 / (2 * R), P.n2), dim3(FFT_THR), lds, s, P, R);
}

/* n = N1 N2 with N1 >= N2, both powers of two up to 32 */
template <class T, class TIN, class TOUT>
void rows(hipStream_t s, const FftPlan<T>& P, bool fwd, const void* in, void* out, double blo, double bhi, int add) {
    switch (P.l0) {
    case 2: rows_launch<T, TIN, TOUT, 2, 2>(s, P, fwd, in, out, blo, bhi, add); break;
    case 3: rows_launch<T, TIN, TOUT, 4, 2>(s, P, fwd, in, out, blo, bhi, add); break;
    case 4: rows_launch<T, TIN, TOUT, 4, 4>(s, P, fwd, in, out, blo, bhi, add); break;
    case 5: rows_launch<T, TIN, TOUT, 8, 4>(s, P, fwd, in, out, blo, bhi, add); break;
    case 6: rows_launch<T, TIN, TOUT, 8, 8>(s, P, fwd, in, out, blo, bhi, add); break;
    case 7: rows_launch<T, TIN, TOUT, 16, 8>(s, P, fwd, in, out, blo, bhi, add); break;
    case 8: rows_launch<T, TIN, TOUT, 16, 16>(s, P, fwd, in, out, blo, bhi, add); break;
    case 9: rows_launch<T, TIN, TOUT, 32, 16>(s, P, fwd, in, out, blo, bhi, add); break;
    default: rows_launch<T, TIN, TOUT, 32, 32>(s, P, fwd, in, out, blo, bhi, add); break;
    }
}

template <class T>
void cols(hipStream_t s, const FftPlan<T>& P) {
    switch (P.l1) {
    case 2: cols_launch<T, 2, 2>(s, P); break;
    case 3: cols_launch<T, 4, 2>(s, P); break;
    case 4: cols_launch<T, 4, 4>(s, P); break;
    case 5: cols_launch<T, 8, 4>(s, P); break;
    case 6: cols_launch<T, 8, 8>(s, P); break;
    case 7: cols_launch<T, 16, 8>(s, P); break;
    case 8: cols_launch<T, 16, 16>(s, P); break;
    case 9: cols_launch<T, 32, 16>(s, P); break;
    default: cols_launch<T, 32, 32>(s, P); break;
    }
}

template <class T, int NZ>
void tri_launch(hipStream_t s, const FftPlan<T>& P, T scale) {
    constexpr int LB = NZ > TRI_REG ? NZ - TRI_REG : 0;
    const size_t lds = ((size_t)LB * TRI_THR + 5 * (size_t)P.n2) * sizeof(T);
    static bool attr = false;   /* the LDS limit is raised once per kernel */
    if (lds > 64 * 1024 && !attr) {
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_fft_tri<T, NZ>), hipFuncAttributeMaxDynamicSharedMemorySize,
                                  160 * 1024);
        attr = true;
    }
    hipLaunchKernelGGL((k_fft_tri<T, NZ>), dim3((P.n0 * P.n1 + TRI_THR - 1) / TRI_THR), dim3(TRI_THR), lds, s, P, scale);
}

template <class T>
void tri(hipStream_t s, const FftPlan<T>& P, T scale) {
    if (P.n2 <= 16) tri_launch<T, 16>(s, P, scale);
    else if (P.n2 <= 32) tri_launch<T, 32>(s, P, scale);
    else if (P.n2 <= 64) tri_launch<T, 64>(s, P, scale);
    else if (P.n2 <= 128) tri_launch<T, 128>(s, P, scale);
    else tri_launch<T, 0>(s, P, scale);
}

template <class T, class TIN, class TOUT>
void fft_solve_impl(hipStream_t s, const FftPlan<T>& P, const TIN* b, TOUT* phi, double blo, double bhi, int add) {
    const T scale = T(1) / T(P.n0 * P.n1);
    rows<T, TIN, TOUT>(s, P, true, b, P.S, blo, bhi, 0);
    cols<T>(s, P);
    tri<T>(s, P, scale);
    cols<T>(s, P);
    rows<T, TIN, TOUT>(s, P, false, P.S, phi, 0.0, 0.0, add);
}

}  // namespace

namespace fxk {

bool fft_size_ok(int n) { return n >= 4 && n <= 1024 && (n & (n - 1)) == 0; }

void fft_solve(hipStream_t s, const fxd::FftPlan<double>& P, const double* b, double* phi, double blo, double bhi) {
    fft_solve_impl<double, double, double>(s, P, b, phi, blo, bhi, 0);
}
void fft_solve(hipStream_t s, const fxd::FftPlan<float>& P, const double* b, double* phi, double blo, double bhi) {
    fft_solve_impl<float, double, double>(s, P, b, phi, blo, bhi, 0);
}
void fft_solve_add(hipStream_t s, const fxd::FftPlan<float>& P, const float* r, double* phi) {
    fft_solve_impl<float, float, double>(s, P, r, phi, 0.0, 0.0, 1);
}

void fft_pass(hipStream_t s, const fxd::FftPlan<double>& P, int pass, const double* b, double* phi, double blo,
              double bhi) {
    switch (pass) {
    case 0: rows<double, double, double>(s, P, true, b, P.S, blo, bhi, 0); break;
    case 1: cols<double>(s, P); break;
    case 2: tri<double>(s, P, 1.0 / (double)(P.n0 * P.n1)); break;
    default: rows<double, double, double>(s, P, false, P.S, phi, 0.0, 0.0, 0); break;
    }
}

}  // namespace fxk
