/*
 * Device-side building blocks of the FerroX hot path for gfx950: field views
 * with the reference's ghost semantics, the masked stencils of
 * Source/Solver/DerivativeAlgorithm.H, the rotation matrix of
 * TotalEnergyDensity.cpp:167-195, and the multigrid row operator.
 *
 * Layout in HBM: every cell field is a dense FP64 array over the valid box,
 * i fastest (c = i + nx*(j + ny*k)), no ghost cells; periodic neighbours are
 * found by index wrap, and a read across a non-periodic face returns the
 * ghost value the reference would hold there (0 for P, see main.cpp:157).
 * The material id is one byte per cell on a box grown by one cell, so the
 * stencils can test the neighbour's material exactly as the reference tests
 * MaterialMask's ghost layer (Initialization.cpp:177-205): bits 0-1 hold the
 * mask code (FE 0, DE 1, SC 2, Channel 3), bit 2 the t-phase flag, and in the
 * device copy bit 3 marks FE cells whose stencils are all centred.
 */
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fxd {

enum { BC_PER = 0, BC_DIR = 1, BC_NEU = 2 };

struct Geom {
    int n[3];
    int per[3];
    double dx[3];
    double prob_lo[3], prob_hi[3];
};

/* P with ghost semantics: periodic image or 0 across a non-periodic face. */
struct PView {
    const double* __restrict__ p;
    int nx, ny, nz, px, py, pz;
    __device__ __forceinline__ double operator()(int i, int j, int k) const {
        if (i < 0) { if (!px) return 0.0; do i += nx; while (i < 0); }
        else if (i >= nx) { if (!px) return 0.0; do i -= nx; while (i >= nx); }
        if (j < 0) { if (!py) return 0.0; do j += ny; while (j < 0); }
        else if (j >= ny) { if (!py) return 0.0; do j -= ny; while (j >= ny); }
        if (k < 0) { if (!pz) return 0.0; do k += nz; while (k < 0); }
        else if (k >= nz) { if (!pz) return 0.0; do k -= nz; while (k >= nz); }
        return p[i + nx * (j + ny * k)];
    }
};

/* Material id bit 3 (device copy only): an FE cell whose P stencils and
 * grad(phi) are all centred (see upload_mat). */
constexpr int MAT_INTERIOR_FE = 8;

/* Material id bits 4 and 5 (device copy only): an FE cell whose stencils are
 * the centred ones except in z, where the cells below it (MAT_ZFACE_LO) or
 * above it (MAT_ZFACE_HI) and below or above its four lateral neighbours
 * are not FE, and the P_BC_flag of that z side is 0..3 (see upload_mat).
 * Its masked z stencils of DerivativeAlgorithm.H are then linear in P at
 * planes k - 1, k, k + 1 with coefficients of the flag alone (zface_coef). */
constexpr int MAT_ZFACE_LO = 16;
constexpr int MAT_ZFACE_HI = 32;

/* The masked z stencils of a MAT_ZFACE_LO (hi = false) or MAT_ZFACE_HI cell
 * as linear forms (h = dz, F the P values at planes k - 1, k, k + 1):
 *   h^2 DoubleDPDz = z[0] F_k + z[1] F_k+1 + z[2] F_k-1   (DerivativeAlgorithm.H:257-467),
 *   2 h DPDz(q)    = z[3] F_k + z[4] F_k+1 + z[5] F_k-1   (:45-252),
 * the latter at the cell and at its lateral neighbours q, which take the
 * same branch; flag 3 gives the centred forms.  The products reassociate
 * the reference's expressions (rounding-level differences). */
__host__ __device__ inline void zface_coef(int flag, bool hi, double h, double lambda, double z[6]) {
    const double s = hi ? h / (1 - h / 2 / lambda) / lambda : h / (1 + h / 2 / lambda) / lambda;
    switch (flag) {
    case 0:
        z[0] = -4.0; z[1] = hi ? 0.0 : 4.0 / 3.0; z[2] = hi ? 4.0 / 3.0 : 0.0;
        z[3] = hi ? -2.0 : 2.0; z[4] = hi ? 0.0 : 2.0 / 3.0; z[5] = hi ? -2.0 / 3.0 : 0.0;
        break;
    case 1:
        z[0] = hi ? s - 1.0 : -s - 1.0; z[1] = hi ? 0.0 : 1.0; z[2] = hi ? 1.0 : 0.0;
        z[3] = hi ? s + 1.0 : s - 1.0; z[4] = hi ? 0.0 : 1.0; z[5] = hi ? -1.0 : 0.0;
        break;
    case 2:
        z[0] = -1.0; z[1] = hi ? 0.0 : 1.0; z[2] = hi ? 1.0 : 0.0;
        z[3] = hi ? 1.0 : -1.0; z[4] = hi ? 0.0 : 1.0; z[5] = hi ? -1.0 : 0.0;
        break;
    default:
        z[0] = -2.0; z[1] = 1.0; z[2] = 1.0;
        z[3] = 0.0; z[4] = 1.0; z[5] = -1.0;
        break;
    }
}

/* Material id on the grown box. */
struct MView {
    const uint8_t* __restrict__ m;
    int sx, sy;
    __device__ __forceinline__ int code(int i, int j, int k) const { return m[(i + 1) + sx * (j + 1) + sy * (k + 1)] & 3; }
    __device__ __forceinline__ int raw(int i, int j, int k) const { return m[(i + 1) + sx * (j + 1) + sy * (k + 1)]; }
};

struct PBC {
    int lo[3], hi[3];
    double lambda;
};

/* Neighbour offsets of an interior FE cell (periodic wrap laterally; such
 * cells never sit on a non-periodic face or the first or last plane). */
struct NbrOff {
    int xm, xp, ym, yp, zm, zp;
    __device__ __forceinline__ NbrOff(int i, int j, int nx, int ny) {
        xm = i > 0 ? -1 : nx - 1;
        xp = i < nx - 1 ? 1 : -(nx - 1);
        ym = j > 0 ? -nx : nx * (ny - 1);
        yp = j < ny - 1 ? nx : -nx * (ny - 1);
        zp = nx * ny;
        zm = -zp;
    }
};

/* DPDx/DPDy/DPDz, DerivativeAlgorithm.H:45-252.  Flags outside 0..4 are
 * rejected before any launch. */
template <int d, class V = PView>
__device__ __forceinline__ double dpd(const V& F, const MView& M, int i, int j, int k,
                                      const double* dx, const PBC& bc) {
    constexpr int di = (d == 0), dj = (d == 1), dk = (d == 2);
    const double h = dx[d];
    const double lambda = bc.lambda;
    const int m0 = M.code(i, j, k);
    if (m0 != 0) return 0.0;
    if (M.code(i - di, j - dj, k - dk) != 0) {
        switch (bc.lo[d]) {
        case 0: { const double F_lo = 0.0; return (-4. * F_lo + 3. * F(i, j, k) + F(i + di, j + dj, k + dk)) / (3. * h); }
        case 1: { const double F_lo = F(i, j, k) / (1 + h / 2 / lambda);
                  return (h * F_lo / lambda - F(i, j, k) + F(i + di, j + dj, k + dk)) / (2. * h); }
        case 2: return (-F(i, j, k) + F(i + di, j + dj, k + dk)) / (2. * h);
        case 3: return (-F(i - di, j - dj, k - dk) + F(i + di, j + dj, k + dk)) / (2. * h);
        default: return (-F(i, j, k) + F(i + di, j + dj, k + dk)) / (h);
        }
    } else if (M.code(i + di, j + dj, k + dk) != 0) {
        switch (bc.hi[d]) {
        case 0: { const double F_hi = 0.0; return (4. * F_hi - 3. * F(i, j, k) - F(i - di, j - dj, k - dk)) / (3. * h); }
        case 1: { const double F_hi = F(i, j, k) / (1 - h / 2 / lambda);
                  return (h * F_hi / lambda + F(i, j, k) - F(i - di, j - dj, k - dk)) / (2. * h); }
        case 2: return (F(i, j, k) - F(i - di, j - dj, k - dk)) / (2. * h);
        case 3: return (F(i + di, j + dj, k + dk) - F(i - di, j - dj, k - dk)) / (2. * h);
        default: return (F(i, j, k) - F(i - di, j - dj, k - dk)) / (h);
        }
    }
    return (F(i + di, j + dj, k + dk) - F(i - di, j - dj, k - dk)) / (2. * h);
}

/* DoubleDPDx/y/z, DerivativeAlgorithm.H:257-467. */
template <int d, class V = PView>
__device__ __forceinline__ double ddpd(const V& F, const MView& M, int i, int j, int k,
                                       const double* dx, const PBC& bc) {
    constexpr int di = (d == 0), dj = (d == 1), dk = (d == 2);
    const double h = dx[d];
    const double lambda = bc.lambda;
    const int m0 = M.code(i, j, k);
    if (m0 != 0) return 0.0;
    if (M.code(i - di, j - dj, k - dk) != 0) {
        switch (bc.lo[d]) {
        case 0: { const double F_lo = 0.0; return 4. * (2. * F_lo - 3. * F(i, j, k) + F(i + di, j + dj, k + dk)) / 3. / h / h; }
        case 1: { const double F_lo = F(i, j, k) / (1 + h / 2 / lambda);
                  return (-h * F_lo / lambda - F(i, j, k) + F(i + di, j + dj, k + dk)) / h / h; }
        case 2: return (-F(i, j, k) + F(i + di, j + dj, k + dk)) / h / h;
        case 3: return (F(i + di, j + dj, k + dk) - 2. * F(i, j, k) + F(i - di, j - dj, k - dk)) / h / h;
        default: return (F(i + 2 * di, j + 2 * dj, k + 2 * dk) - 2. * F(i + di, j + dj, k + dk) + F(i, j, k)) / h / h;
        }
    } else if (M.code(i + di, j + dj, k + dk) != 0) {
        switch (bc.hi[d]) {
        case 0: { const double F_hi = 0.0; return 4. * (2. * F_hi - 3. * F(i, j, k) + F(i - di, j - dj, k - dk)) / 3. / h / h; }
        case 1: { const double F_hi = F(i, j, k) / (1 - h / 2 / lambda);
                  return (h * F_hi / lambda - F(i, j, k) + F(i - di, j - dj, k - dk)) / h / h; }
        case 2: return (-F(i, j, k) + F(i - di, j - dj, k - dk)) / h / h;
        case 3: return (F(i + di, j + dj, k + dk) - 2. * F(i, j, k) + F(i - di, j - dj, k - dk)) / h / h;
        default: return (F(i, j, k) - 2. * F(i - di, j - dj, k - dk) + F(i - 2 * di, j - 2 * dj, k - 2 * dk)) / h / h;
        }
    }
    return (F(i + di, j + dj, k + dk) - 2. * F(i, j, k) + F(i - di, j - dj, k - dk)) / (h * h);
}

/* DoubleDPDxDy / DxDz / DyDz, DerivativeAlgorithm.H:471-497. */
template <int a, int b, class V = PView>
__device__ __forceinline__ double ddpd_mixed(const V& F, const MView& M, int i, int j, int k,
                                             const double* dx, const PBC& bc) {
    constexpr int ai = (a == 0), aj = (a == 1), ak = (a == 2);
    return (dpd<b>(F, M, i + ai, j + aj, k + ak, dx, bc) - dpd<b>(F, M, i - ai, j - aj, k - ak, dx, bc)) / 2. / dx[a];
}

/* Rotation from Euler angles in degrees (TotalEnergyDensity.cpp:167-195). */
__host__ __device__ inline void rotation(double a_deg, double b_deg, double t_deg, int use_Euler_angles, double R[9]) {
    const double Pi = 3.14159265358979323846;
    const double a = Pi / 180. * a_deg, b = Pi / 180. * b_deg, t = Pi / 180. * t_deg;
    if (use_Euler_angles) {
        R[0] = cos(a) * cos(t) - cos(b) * sin(a) * sin(t);
        R[1] = sin(a) * cos(t) + cos(b) * cos(a) * sin(t);
        R[2] = sin(b) * sin(t);
        R[3] = -cos(b) * cos(t) * sin(a) - cos(a) * sin(t);
        R[4] = cos(b) * cos(a) * cos(t) - sin(a) * sin(t);
        R[5] = sin(b) * cos(t);
        R[6] = sin(a) * sin(b);
        R[7] = -cos(a) * sin(b);
        R[8] = cos(b);
    } else {
        R[0] = cos(b) * cos(t);
        R[1] = sin(a) * sin(b) * cos(t) - cos(a) * sin(t);
        R[2] = cos(a) * sin(b) * cos(t) + sin(a) * sin(t);
        R[3] = cos(b) * sin(t);
        R[4] = sin(b) * sin(a) * sin(t) + cos(a) * cos(t);
        R[5] = cos(a) * sin(b) * sin(t) - sin(a) * cos(t);
        R[6] = -sin(b);
        R[7] = sin(a) * cos(b);
        R[8] = cos(a) * cos(b);
    }
}

/* phi across faces for E (ElectrostaticSolver.cpp:134-203).  The Dirichlet
 * ghost is 2*phi_b - phi_0 (what MLMG leaves in the aliased solution) or phi_b
 * (face value), per ghost_mode. */
struct PhiView {
    const double* __restrict__ p;
    int n[3];
    int lo[3], hi[3];          /* BC kinds */
    double v[6];               /* Dirichlet values x,y,z lo then hi */
    int ghost_face;
    __device__ __forceinline__ double at(int i, int j, int k) const {
        const int nx = n[0], ny = n[1], nz = n[2];
        if (i < 0 || i >= nx) {
            const bool l = i < 0; const int t = l ? lo[0] : hi[0];
            if (t == BC_PER) i = l ? i + nx : i - nx;
            else { const double p0 = p[(l ? 0 : nx - 1) + nx * (j + ny * k)];
                   if (t == BC_NEU) return p0; const double pb = v[l ? 0 : 3];
                   return ghost_face ? pb : 2.0 * pb - p0; }
        }
        if (j < 0 || j >= ny) {
            const bool l = j < 0; const int t = l ? lo[1] : hi[1];
            if (t == BC_PER) j = l ? j + ny : j - ny;
            else { const double p0 = p[i + nx * ((l ? 0 : ny - 1) + ny * k)];
                   if (t == BC_NEU) return p0; const double pb = v[l ? 1 : 4];
                   return ghost_face ? pb : 2.0 * pb - p0; }
        }
        if (k < 0 || k >= nz) {
            const bool l = k < 0; const int t = l ? lo[2] : hi[2];
            if (t == BC_PER) k = l ? k + nz : k - nz;
            else { const double p0 = p[i + nx * (j + ny * (l ? 0 : nz - 1))];
                   if (t == BC_NEU) return p0; const double pb = v[l ? 2 : 5];
                   return ghost_face ? pb : 2.0 * pb - p0; }
        }
        return p[i + nx * (j + ny * k)];
    }
};

/* grad(phi) with the reference's one-sided z formula at the metal faces
 * (DphiDz, DFDx, DFDy: DerivativeAlgorithm.H:8-40). */
template <class PV = PhiView>
__device__ __forceinline__ void grad_phi(const PV& P, const Geom& g, int i, int j, int k,
                                         double& dfdx, double& dfdy, double& dfdz) {
    const double* dx = g.dx;
    dfdx = (P.at(i + 1, j, k) - P.at(i - 1, j, k)) / (2. * dx[0]);
    dfdy = (P.at(i, j + 1, k) - P.at(i, j - 1, k)) / (2. * dx[1]);
    const double z_hi = g.prob_lo[2] + (k + 1.5) * dx[2];
    const double z_lo = g.prob_lo[2] + (k - 0.5) * dx[2];
    if (z_lo < g.prob_lo[2]) dfdz = (-4. * P.at(i, j, k - 1) + 3. * P.at(i, j, k) + P.at(i, j, k + 1)) / (3. * dx[2]);
    else if (z_hi > g.prob_hi[2]) dfdz = (4. * P.at(i, j, k + 1) - 3. * P.at(i, j, k) - P.at(i, j, k - 1)) / (3. * dx[2]);
    else dfdz = (P.at(i, j, k + 1) - P.at(i, j, k - 1)) / (2. * dx[2]);
}

/* ------------------------------ multigrid -------------------------------- */
/* One level of the hierarchy, in FP64 or FP32.  Face coefficients come from
 * per-k arrays (layered level: the stack is laterally uniform), from the
 * material id on the finest level (beta_face = 0.5*(eps_c + eps_nb), eps_nb =
 * eps_c across a non-periodic face, as the reference copies the adjacent cell
 * into the ghost, ElectrostaticSolver.cpp:257-298), or from face arrays.  The
 * three hold the same values. */
template <class T>
struct DLevT {
    int n0, n1, n2;
    T ih2[3];
    T* phi;
    T* b;
    T* bx;   /* (n0+1)*n1*n2 */
    T* by;   /* n0*(n1+1)*n2 */
    T* bz;   /* n0*n1*(n2+1) */
    T* a;
    int crs[3];
    const uint8_t* mat;   /* grown box, finest level */
    int msx, msy;
    int matcoef;          /* faces from mat + eps table */
    int read_a;           /* 0: a == 0; 1: a nonzero only in SC cells; 2: a anywhere */
    T eps[8];
    int layered;
    const T* bxk;
    const T* byk;
    const T* bzk;
    T* phi2;   /* second solution buffer of the one-pass sweep (null if not allocated) */
    int stream;   /* large enough for the streaming kernels (the mg_stream_min_cells input) */
};
using DLev = DLevT<double>;
using DLevF = DLevT<float>;

struct DBC {
    int lo[3], hi[3];
    double v[6];
};

/* The coarse tail of a layered hierarchy as one workgroup (k_mg_tail_lay):
 * each level's extents, coarsening pattern, 1/h^2, per-plane face
 * coefficients and Helmholtz term (global), the first level's rhs and
 * solution, and the bottom level's LU factors (column-major) with its row
 * permutation; perm[N] is 1 when the factors exist (non-singular bottom). */
constexpr int TAIL_LEAN_MAXL = 10;
template <class T>
struct TailPlan {
    int nl;
    int n[TAIL_LEAN_MAXL][3];
    int crs[TAIL_LEAN_MAXL][3];
    T ih2[TAIL_LEAN_MAXL][3];
    const T* bxk[TAIL_LEAN_MAXL];
    const T* byk[TAIL_LEAN_MAXL];
    const T* bzk[TAIL_LEAN_MAXL];
    const T* a[TAIL_LEAN_MAXL];
    const T* b0;
    T* phi0;
    const T* binv;   /* the bottom operator's inverse, column-major (bottom_factor) */
    const int* perm;
};

/* One 16x16x4 matrix-core step per element type, C += A B with A 16x4 and B
 * 4x16: lane l holds A[l % 16][l / 16] and B[l / 16][l % 16]; of the 16x16
 * result it holds four values, C[row(l, v)][l % 16] for v = 0..3 (the FP64
 * instruction interleaves the rows, the FP32 one keeps four consecutive rows
 * per lane group; both layouts were checked on the device by multiplying
 * known matrices). */
template <class T> struct Mfma16;
template <> struct Mfma16<double> {
    typedef double v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 mma(double a, double b, v4 c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
    static __device__ __forceinline__ int row(int lane, int v) { return 4 * v + (lane >> 4); }
};
template <> struct Mfma16<float> {
    typedef float v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 mma(float a, float b, v4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
    static __device__ __forceinline__ int row(int lane, int v) { return 4 * (lane >> 4) + v; }
};

/* The exact solve that ends the V-cycle of a laterally periodic layered
 * hierarchy (mg_spec_solve): one level, its operator block-diagonalised by
 * the real Fourier basis of each periodic lateral direction (column m of the
 * n x n matrix Q is basis vector m, stored Q[i*n + m]; mu[m] its eigenvalue of
 * 2 f_i - f_i-1 - f_i+1), then one tridiagonal system in z per mode.
 * Homogeneous faces; dlo/dhi = 1 for a Dirichlet z face, 0 for Neumann.
 * S and C are scratch of the level's size (the transformed values and the
 * eliminated upper diagonal of the z systems). */
template <class T>
struct SpecTail {
    int n0, n1, n2;
    const T* b;
    T* phi;
    T* S;
    T* C;
    const T* Qx;
    const T* Qy;
    const T* mux;
    const T* muy;
    const T* bxk;
    const T* byk;
    const T* bzk;
    T ih2[3];
    int dlo, dhi;
};
/* Direct solve of a laterally periodic, z-layered level (fx_fft.hip): the
 * Hartley twiddles and eigenvalues of each periodic direction (n0, n1
 * powers of two), the per-plane face coefficients, the spectrum buffer S
 * (n0 n1 n2 values) and the tridiagonal solve's checkpoints ck (n0 n1
 * ceil(n2 / 16) values). */
template <class T>
struct FftPlan {
    int n0, n1, n2;
    int l0, l1;            /* log2 n0, log2 n1 */
    const T* tw0;          /* cos(2 pi q / n0), -sin(2 pi q / n0), q < n0/2, interleaved */
    const T* tw1;
    const T* mux;          /* 2 - 2 cos(2 pi m / n0) */
    const T* muy;
    const T* bxk;
    const T* byk;
    const T* bzk;
    T ih2[3];
    int dlo, dhi;
    T* S;
    /* reciprocal pivots of the z systems, 1 / (d_k - a_k c'_k-1) of the Thomas
     * recurrence, per plane k and mode class (min(mx, n0-mx), min(my, n1-my)):
     * piv[k (n1/2+1) (n0/2+1) + myc (n0/2+1) + mxc] (mu_m = mu_n-m, so the
     * four mirror modes share one system) */
    const T* piv;
    /* non-null: the pivots of the same systems eliminated upward from the last
     * plane, 1 / (d_k - c_k a'_k+1) with a'_k = a_k pivu_k, laid out as piv;
     * the z solve then runs from both ends and meets in the middle */
    const T* pivu;
    /* with pivu: the Thomas values of mode (0, 0) for a unit right-hand side
     * on the first plane (downward, green[k]) and on the last (upward,
     * green[n2 + k]); null: the z solve always runs over every plane */
    const T* green;
    /* non-null: every pass of the solve does nothing while *skip != 0 (a
     * device flag set when an iterative solve around it has converged) */
    const double* skip;
};
/* Plane kernels of the exact solve (one workgroup per plane and block of 16
 * columns): dynamic LDS values, the plane and the y basis with every row
 * padded by one value so that the matrix-core operand loads, 16 rows at a
 * time, fall in distinct banks, and two n0 x 17 blocks (16 columns of the x
 * basis, 16 columns of the intermediate product). */
__host__ __device__ constexpr size_t spec_plane_values(int n0, int n1) {
    return (size_t)n1 * (n0 + 1) + (size_t)n1 * (n1 + 1) + 2 * (size_t)(n0 > n1 ? n0 : n1) * 17;
}
/* A level is solved exactly where n0*n1 <= SPEC_MAX_PLANE, n0, n1 <= SPEC_MAX_N
 * and it has at most SPEC_MAX_CELLS cells (the same rule as the oracle's). */
constexpr int SPEC_MAX_PLANE = 4096, SPEC_MAX_N = 64, SPEC_MAX_CELLS = 65536;

/* Cell permittivity of a material id: FE (t-phase or not), DE, SC; chosen
 * with selects on kernel-argument scalars rather than an indexed table. */
template <class T>
__device__ __forceinline__ T eps_of(const DLevT<T>& l, int m) {
    const int code = m & 3;
    const T fe = (m & 4) ? l.eps[4] : l.eps[0];
    return code == 0 ? fe : (code == 1 ? l.eps[1] : l.eps[2]);
}

/* Coefficients of row (i,j,k): A-coefficient and the six face weights
 * beta/h^2 in the order x lo, x hi, y lo, y hi, z lo, z hi. */
template <class T>
__device__ __forceinline__ void mg_coef(const DLevT<T>& l, const DBC& bc, int i, int j, int k, int c,
                                        T& a, T w[6]) {
    const int n0 = l.n0, n1 = l.n1;
    const int g = (i + 1) + l.msx * (j + 1) + l.msy * (k + 1);   /* finest level only */
    if (l.read_a == 0) a = T(0);
    else if (l.read_a == 2) a = l.a[c];
    else a = ((l.mat[g] & 3) >= 2) ? l.a[c] : T(0);
    if (l.layered) {
        const T bxv = l.bxk[k], byv = l.byk[k];
        w[0] = bxv * l.ih2[0];
        w[1] = bxv * l.ih2[0];
        w[2] = byv * l.ih2[1];
        w[3] = byv * l.ih2[1];
        w[4] = l.bzk[k] * l.ih2[2];
        w[5] = l.bzk[k + 1] * l.ih2[2];
    } else if (l.matcoef) {
        const int mc = l.mat[g];
        const T ec = eps_of(l, mc);
        const T exl = (i > 0 || bc.lo[0] == BC_PER) ? eps_of(l, l.mat[g - 1]) : ec;
        const T exh = (i < n0 - 1 || bc.hi[0] == BC_PER) ? eps_of(l, l.mat[g + 1]) : ec;
        const T eyl = (j > 0 || bc.lo[1] == BC_PER) ? eps_of(l, l.mat[g - l.msx]) : ec;
        const T eyh = (j < n1 - 1 || bc.hi[1] == BC_PER) ? eps_of(l, l.mat[g + l.msx]) : ec;
        const T ezl = (k > 0 || bc.lo[2] == BC_PER) ? eps_of(l, l.mat[g - l.msy]) : ec;
        const T ezh = (k < l.n2 - 1 || bc.hi[2] == BC_PER) ? eps_of(l, l.mat[g + l.msy]) : ec;
        const T h = T(0.5);
        w[0] = (h * (ec + exl)) * l.ih2[0];
        w[1] = (h * (exh + ec)) * l.ih2[0];
        w[2] = (h * (ec + eyl)) * l.ih2[1];
        w[3] = (h * (eyh + ec)) * l.ih2[1];
        w[4] = (h * (ec + ezl)) * l.ih2[2];
        w[5] = (h * (ezh + ec)) * l.ih2[2];
    } else {
        const int fx = i + (n0 + 1) * (j + n1 * k);
        const int fy = i + n0 * (j + (n1 + 1) * k);
        w[0] = l.bx[fx] * l.ih2[0];
        w[1] = l.bx[fx + 1] * l.ih2[0];
        w[2] = l.by[fy] * l.ih2[1];
        w[3] = l.by[fy + n0] * l.ih2[1];
        w[4] = l.bz[c] * l.ih2[2];
        w[5] = l.bz[c + n0 * n1] * l.ih2[2];
    }
}

/* (A phi)_c = D*phi_c - S for  -a*phi - div(beta grad phi)  with periodic /
 * Dirichlet (ghost 2*phi_b - phi_0, value on the face) / homogeneous Neumann
 * faces.  nb[6] are the neighbour values (read only for interior or periodic
 * faces).  Accumulation order: -a, x lo, x hi, y lo, y hi, z lo, z hi. */
template <class T>
__device__ __forceinline__ void mg_acc(const DLevT<T>& l, const DBC& bc, bool inh, int i, int j, int k,
                                       T a, const T w[6], const T nb[6], T& D, T& S) {
    const int n[3] = {l.n0, l.n1, l.n2};
    const int ijk[3] = {i, j, k};
    D = -a;
    S = T(0);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const T wlo = w[2 * d], whi = w[2 * d + 1];
        if (ijk[d] > 0 || bc.lo[d] == BC_PER) { D += wlo; S += wlo * nb[2 * d]; }
        else if (bc.lo[d] == BC_DIR) { D += T(2) * wlo; if (inh) S += T(2) * wlo * T(bc.v[d]); }
        if (ijk[d] < n[d] - 1 || bc.hi[d] == BC_PER) { D += whi; S += whi * nb[2 * d + 1]; }
        else if (bc.hi[d] == BC_DIR) { D += T(2) * whi; if (inh) S += T(2) * whi * T(bc.v[3 + d]); }
    }
}

/* Neighbour values from a global array (periodic wrap). */
template <class T>
__device__ __forceinline__ void mg_nb_global(const DLevT<T>& l, const T* __restrict__ phi, int i, int j, int k,
                                             int c, T nb[6]) {
    const int n0 = l.n0, n1 = l.n1, n2 = l.n2, s = n0 * n1;
    nb[0] = phi[i > 0 ? c - 1 : c + (n0 - 1)];
    nb[1] = phi[i < n0 - 1 ? c + 1 : c - (n0 - 1)];
    nb[2] = phi[j > 0 ? c - n0 : c + n0 * (n1 - 1)];
    nb[3] = phi[j < n1 - 1 ? c + n0 : c - n0 * (n1 - 1)];
    nb[4] = phi[k > 0 ? c - s : c + s * (n2 - 1)];
    nb[5] = phi[k < n2 - 1 ? c + s : c - s * (n2 - 1)];
}

template <class T>
__device__ __forceinline__ void mg_row(const DLevT<T>& l, const T* __restrict__ phi, int i, int j, int k,
                                       const DBC& bc, bool inh, T& D, T& S) {
    const int c = i + l.n0 * (j + l.n1 * k);
    T a, w[6], nb[6];
    mg_coef(l, bc, i, j, k, c, a, w);
    mg_nb_global(l, phi, i, j, k, c, nb);
    mg_acc(l, bc, inh, i, j, k, a, w, nb, D, S);
}

/* Block-wide reductions (wavefront shuffle, then LDS across waves). */
template <int BS, class T>
__device__ __forceinline__ T block_sum(T v, T* sh) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) sh[w] = v;
    __syncthreads();
    T r = T(0);
    if (threadIdx.x == 0) { for (int i = 0; i < BS / 64; ++i) r += sh[i]; sh[BS / 64] = r; }
    __syncthreads();
    r = sh[BS / 64];
    __syncthreads();
    return r;
}
/* max that propagates NaN (fmax would drop it and hide a diverged solve) */
__device__ __forceinline__ double nmax(double a, double b) { return (a > b || a != a) ? a : b; }

template <int BS>
__device__ __forceinline__ double block_max(double v, double* sh) {
    for (int o = 32; o > 0; o >>= 1) v = nmax(v, __shfl_xor(v, o, 64));
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) sh[w] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0) { for (int i = 0; i < BS / 64; ++i) r = nmax(r, sh[i]); sh[BS / 64] = r; }
    __syncthreads();
    r = sh[BS / 64];
    __syncthreads();
    return r;
}

}  // namespace fxd
