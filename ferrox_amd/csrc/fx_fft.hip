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

constexpr int FFT_THR = 256;
constexpr int TRI_BLK = 16;   /* planes per register block of the tridiagonal solve */

__device__ __forceinline__ int bitrev(int v, int logn) { return (int)(__brev((unsigned)v) >> (32 - logn)); }

/* nseq complex sequences of length n = 2^logn, in LDS as re[r*n + i],
 * im[r*n + i], loaded in bit-reversed order; tw[2q], tw[2q+1] =
 * cos(2 pi q / n), -sin(2 pi q / n), q < n/2.  In-place radix-2
 * decimation in time: after it, re + i im holds the DFT in natural order. */
template <class T>
__device__ __forceinline__ void lds_fft(T* re, T* im, int nseq, int n, int logn, const T* tw) {
    const int half = n >> 1;
    const int nb = nseq * half;
    for (int m = 1, ls = 0; m < n; m <<= 1, ++ls) {
        const int tsh = logn - 1 - ls;            /* twiddle index j * n / (2m) = j << tsh */
        for (int t = threadIdx.x; t < nb; t += blockDim.x) {
            const int r = t >> (logn - 1);
            const int b = t & (half - 1);
            const int j = b & (m - 1);
            const int i0 = r * n + ((b - j) << 1) + j;
            const int i1 = i0 + m;
            const T wr = tw[2 * (j << tsh)], wi = tw[2 * (j << tsh) + 1];
            const T xr = re[i1], xi = im[i1];
            const T tr = wr * xr - wi * xi;
            const T ti = wr * xi + wi * xr;
            const T ur = re[i0], ui = im[i0];
            re[i0] = ur + tr; im[i0] = ui + ti;
            re[i1] = ur - tr; im[i1] = ui - ti;
        }
        __syncthreads();
    }
}

/* Hartley value of the real part (part 0) or of the imaginary part (part 1)
 * at index m of a transformed packed sequence. */
template <class T>
__device__ __forceinline__ T hartley(const T* re, const T* im, int base, int n, int m, int part) {
    const int mm = (n - m) & (n - 1);
    const T A = re[base + m], B = im[base + m], C = re[base + mm], D = im[base + mm];
    return part ? T(0.5) * ((A + B) + (D - C)) : T(0.5) * ((A - B) + (C + D));
}

/* Hartley transform along x of rows j0 .. j0 + 2R - 1 of plane k = blockIdx.y
 * (R = rows_pairs).  FWD: input b (TIN) plus the Dirichlet face terms on the
 * first and last plane, output S.  !FWD: input S, output phi (TOUT), stored
 * or, with add, added. */
template <class T, class TIN, class TOUT, bool FWD>
__global__ __launch_bounds__(FFT_THR) void k_fft_rows(FftPlan<T> P, const TIN* __restrict__ in, TOUT* __restrict__ out,
                                                      int R, double blo, double bhi, int add) {
    extern __shared__ __align__(16) unsigned char fft_lds[];
    const int n = P.n0, logn = P.l0, k = blockIdx.y, j0 = blockIdx.x * 2 * R;
    T* const tw = reinterpret_cast<T*>(fft_lds);
    T* const re = tw + n;
    T* const im = re + R * n;
    for (int q = threadIdx.x; q < n; q += blockDim.x) tw[q] = P.tw0[q];
    const size_t row0 = ((size_t)k * P.n1 + j0) * n;
    double bterm = 0.0;
    if (FWD) bterm = (k == 0 ? blo : 0.0) + (k == P.n2 - 1 ? bhi : 0.0);
    for (int t = threadIdx.x; t < 2 * R * n; t += blockDim.x) {
        const int q = t >> logn, c = t & (n - 1);
        T v;
        if (FWD) v = (T)((double)in[row0 + t] + bterm);
        else v = (T)in[row0 + t];
        T* dst = (q & 1) ? im : re;
        dst[(q >> 1) * n + bitrev(c, logn)] = v;
    }
    __syncthreads();
    lds_fft<T>(re, im, R, n, logn, tw);
    for (int t = threadIdx.x; t < 2 * R * n; t += blockDim.x) {
        const int q = t >> logn, c = t & (n - 1);
        const T h = hartley<T>(re, im, (q >> 1) * n, n, c, q & 1);
        if (FWD || !add) out[row0 + t] = (TOUT)h;
        else out[row0 + t] = out[row0 + t] + (TOUT)h;
    }
}

/* Hartley transform along y, in place, of columns i0 .. i0 + 2R - 1 of
 * plane k = blockIdx.y (rows of 2R values are read and written whole). */
template <class T>
__global__ __launch_bounds__(FFT_THR) void k_fft_cols(FftPlan<T> P, int R) {
    extern __shared__ __align__(16) unsigned char fft_lds[];
    const int n = P.n1, logn = P.l1, n0 = P.n0, k = blockIdx.y, i0 = blockIdx.x * 2 * R, w = 2 * R;
    T* const tw = reinterpret_cast<T*>(fft_lds);
    T* const re = tw + n;
    T* const im = re + R * n;
    for (int q = threadIdx.x; q < n; q += blockDim.x) tw[q] = P.tw1[q];
    T* const S = P.S + (size_t)k * P.n1 * n0 + i0;
    for (int t = threadIdx.x; t < w * n; t += blockDim.x) {
        const int j = t / w, q = t - j * w;
        T* dst = (q & 1) ? im : re;
        dst[(q >> 1) * n + bitrev(j, logn)] = S[(size_t)j * n0 + q];
    }
    __syncthreads();
    lds_fft<T>(re, im, R, n, logn, tw);
    for (int t = threadIdx.x; t < w * n; t += blockDim.x) {
        const int j = t / w, q = t - j * w;
        S[(size_t)j * n0 + q] = hartley<T>(re, im, (q >> 1) * n, n, j, q & 1);
    }
}

/* One thread per lateral mode: the tridiagonal system in z (Thomas
 * algorithm), in place on S, the right-hand side scaled by `scale` (the
 * 1/(n0 n1) of the two inverse transforms, exact for powers of two).  The
 * forward sweep stores the eliminated right-hand side in S and, every
 * TRI_BLK planes, the recurrence value c'_k-1 in P.ck; the back substitution
 * recomputes each block's c' from its checkpoint with the same operations. */
template <class T>
__global__ __launch_bounds__(FFT_THR) void k_fft_tri(FftPlan<T> P, T scale) {
    const int s = P.n0 * P.n1, n2 = P.n2;
    const int c = blockIdx.x * FFT_THR + threadIdx.x;
    if (c >= s) return;
    const T lx = P.mux[c & (P.n0 - 1)] * P.ih2[0];
    const T ly = P.muy[c >> P.l0] * P.ih2[1];
    const T iz = P.ih2[2];
    auto coef = [&](int k, T& a, T& d, T& cc) {
        const T wzl = P.bzk[k] * iz, wzh = P.bzk[k + 1] * iz;
        T diag = P.bxk[k] * lx;
        diag += P.byk[k] * ly;
        diag += k > 0 ? wzl : (P.dlo ? T(2) * wzl : T(0));
        diag += k < n2 - 1 ? wzh : (P.dhi ? T(2) * wzh : T(0));
        a = k > 0 ? -wzl : T(0);
        cc = k < n2 - 1 ? -wzh : T(0);
        d = diag;
    };
    T* const col = P.S + c;
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
                const T r = T(1) / (d - a * cprev);
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
                const T r = T(1) / (d - a * cp);
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

/* sequences per workgroup: 2R rows (columns) of n values, R n <= 2048
 * complex values in LDS for FP64 (4096 for FP32), 2R <= the other extent */
template <class T>
int pairs_per_block(int n, int other) {
    const int cap = (sizeof(T) == 8 ? 2048 : 4096) / n;
    int R = cap < 1 ? 1 : cap;
    if (R > 8 * (int)(8 / sizeof(T))) R = 8 * (int)(8 / sizeof(T));
    if (2 * R > other) R = other / 2;
    return R;
}

template <class T>
size_t fft_lds_bytes(int n, int R) { return ((size_t)n + 2 * (size_t)R * n) * sizeof(T); }

template <class T, class TIN, class TOUT>
void fft_solve_impl(hipStream_t s, const FftPlan<T>& P, const TIN* b, TOUT* phi, double blo, double bhi, int add) {
    const int Rx = pairs_per_block<T>(P.n0, P.n1), Ry = pairs_per_block<T>(P.n1, P.n0);
    const size_t lx = fft_lds_bytes<T>(P.n0, Rx), ly = fft_lds_bytes<T>(P.n1, Ry);
    static bool attr = false;
    if (!attr) {
        attr = true;
        const int mx = 160 * 1024;
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_fft_rows<T, TIN, T, true>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, mx);
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_fft_rows<T, T, TOUT, false>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, mx);
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_fft_cols<T>), hipFuncAttributeMaxDynamicSharedMemorySize, mx);
    }
    const dim3 gx(P.n1 / (2 * Rx), P.n2), gy(P.n0 / (2 * Ry), P.n2);
    const int s0 = P.n0 * P.n1;
    const T scale = T(1) / T(s0);
    hipLaunchKernelGGL((k_fft_rows<T, TIN, T, true>), gx, dim3(FFT_THR), lx, s, P, b, P.S, Rx, blo, bhi, 0);
    hipLaunchKernelGGL(k_fft_cols<T>, gy, dim3(FFT_THR), ly, s, P, Ry);
    hipLaunchKernelGGL(k_fft_tri<T>, dim3((s0 + FFT_THR - 1) / FFT_THR), dim3(FFT_THR), 0, s, P, scale);
    hipLaunchKernelGGL(k_fft_cols<T>, gy, dim3(FFT_THR), ly, s, P, Ry);
    hipLaunchKernelGGL((k_fft_rows<T, T, TOUT, false>), gx, dim3(FFT_THR), lx, s, P, (const T*)P.S, phi, Rx, 0.0, 0.0, add);
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

}  // namespace fxk
