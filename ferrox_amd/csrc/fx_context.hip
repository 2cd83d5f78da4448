/*
 * Host side of the MI355X FerroX hot path: context, set-up, the multigrid
 * driver, the outer Newton loop, the time loop and the C ABI declared in
 * include/ferrox_mi.h.
 *
 * Reference call stack being replaced (Source/main.cpp:261-623):
 *   ComputePhi_Rho            -> compute_phi_rho()   ElectrostaticSolver.cpp:678-747
 *   ComputeEfromPhi           -> fused into k_tdgl / k_compute_E
 *   Calculate_* + TDGL_RHS    -> k_tdgl
 *   LinComb / Copy / Fill     -> fused into k_tdgl (Euler) and k_heun
 *   CheckSteadyState          -> steady_check()      ElectrostaticSolver.cpp:526-560
 *   voltage sweep + stops     -> fx_step()           main.cpp:572-621
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <mutex>
#include <vector>
#include <unordered_map>

#include "../../include/ferrox_mi.h"
#include "fx_device.h"
#include "fx_kernels.h"

using namespace fxd;

namespace {

constexpr int RED_NB = 2048;         /* blocks of the fixed-grid reductions */
constexpr double OUTER_TOL = 1.e-5;  /* ElectrostaticSolver.cpp:697 */
constexpr double BOTTOM_TOL = 1.e-6;
constexpr int BOTTOM_MAXIT = 400;
constexpr size_t TAIL_MAX_CELLS = 8192;   /* levels this small run inside the single-workgroup tail kernel */

inline uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
}
/* Counter-based uniform in [0,1): the build's seeded initial-P generator. */
inline double uniform01(uint64_t seed, uint64_t idx) {
    uint64_t x = splitmix64((seed * 0xD1B54A32D192ED03ULL) ^ splitmix64(idx));
    return (double)(x >> 11) * 0x1.0p-53;
}

struct Level {
    DLev d;
    size_t N;
    double* up = nullptr;   /* block levels: the up leg's output, swapped with d.phi after it */
};

}  // namespace

struct fx_ctx {
    fx_inputs in;
    int n[3];
    size_t N;
    double dx[3];
    hipStream_t s = nullptr;
    /* second stream for the face-cell kernels beside the streamed passes,
     * and the two events that order it (default priority, like s) */
    hipStream_t s2 = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    /* device fields, dense FP64 over the valid box */
    double* P[3] = {nullptr, nullptr, nullptr};
    double* Pn[3] = {nullptr, nullptr, nullptr};
    double* phi = nullptr;
    double* phi_old = nullptr;
    double* phi_prev = nullptr;
    double* rhs = nullptr;
    double* rho = nullptr;
    double* e_den = nullptr;
    double* h_den = nullptr;
    double* alpha_cc = nullptr;
    /* rho (alpha_cc) holds 0 in every cell, written so by this library and
     * never handed out as a device pointer: without semiconductor cells the
     * per-step rewrite with zeros is skipped */
    bool rho_clean = false, alpha_clean = false;
    bool rho_exposed = false, alpha_exposed = false;
    /* the steady-state check's max|phi| and max|phi - phi_old| (device), and
     * whether the host has yet to read the last of them */
    double* steady_dev = nullptr;
    bool steady_pending = false;
    bool phi_old_stale = false;    /* steps skipped the check: phi_old is not the previous step's phi */
    /* the TDGL update of the last step also formed the Poisson RHS of the P
     * it produced (k_tdgl_rhs_stream): rhs holds RHS(P) for the current P and
     * the next solve needs no RHS pass.  Anything else that writes P, rhs,
     * the materials or the rotations clears it. */
    bool rhs_next = false;
    int fe_k0 = -1, fe_k1 = -1;    /* first and last plane holding an FE cell (-1: none) */
    int* face_list = nullptr;      /* FE cells without centred stencils (global indices), then the planes holding them */
    int nface = 0;
    int nface_planes = 0;
    /* the same without the cells next to a z face only (MAT_ZFACE_LO/HI),
     * which the streamed interior passes update */
    int* face2_list = nullptr;
    int nface2 = 0;
    int nface2_planes = 0;
    double* noalloc_scratch = nullptr;   /* device address of the static workspace on this context's GPU */
    /* phi_z0 <= phi_z1: the last direct solve formed phi on those planes
     * only (what the TDGL update reads); the other planes' solution waits,
     * transformed, in fft64.S until phi_sync() finishes it */
    int phi_z0 = 0, phi_z1 = -1;
    /* the z solve, too, stored only planes phi_z0 .. phi_z1 of fft64.S:
     * phi_sync continues it from there with the solve's face terms */
    bool phi_cont = false;
    double phi_face[2] = {0.0, 0.0};
    bool phi_exposed = false;      /* a device pointer to phi was handed out: always form all of it */
    /* the fused update writes the next RHS into rhs_nx, which the next solve
     * swaps in (rhs keeps the RHS of the last solve, the PoissonRHS the
     * reference writes at plot time); the fused kernel writes only planes
     * fe_k0..fe_k1, so each buffer records whether it holds 0 in every
     * non-FE cell */
    double* rhs_nx = nullptr;
    bool rhs_clean = false, rhs_nx_clean = true, rhs_exposed = false;
    double* E[3] = {nullptr, nullptr, nullptr};
    double* GL[3] = {nullptr, nullptr, nullptr};
    double* GL1[3] = {nullptr, nullptr, nullptr};
    double* phidiff = nullptr;
    double* eps = nullptr;         /* cell permittivity (output field) */
    uint8_t* mat = nullptr;        /* material id on the grown box */
    int msx = 0, msy = 0;
    int* rot_idx = nullptr;
    double* rot_tab = nullptr;
    double* rot_atab = nullptr;
    int nrot = 0;                  /* rows of rot_tab / rot_atab */
    double Ru[9];
    /* reductions */
    double* red = nullptr;
    int nred = 0;                  /* capacity of red, in partials per quantity */
    double* red_out = nullptr;
    double* h_red = nullptr;       /* pinned: 8 reduction values, then the publish sequence number; from
                                      h_red + 16, PCG_SLOTS (value, sequence number) pairs of the CG solve */
    double* d_red_host = nullptr;  /* device view of h_red */
    unsigned pub_seq = 0;
    int* d_iters = nullptr;
    double* cg_work = nullptr;
    double* tmp = nullptr;         /* N doubles: residual of a one-level solve */
    std::vector<Level> lev;
    /* levels tail0.. run as one single-workgroup kernel (-1: none) */
    int tail0 = -1;
    int tail_base = -1;            /* first level small enough for the tail */
    /* layered hierarchies: every tail level in LDS (k_mg_tail_lay) with the
     * bottom's LU factors, refactored when the bottom operator changes */
    int tail_lean = 0;
    fxd::TailPlan<double> plan64{};
    fxd::TailPlan<float> plan32{};
    double* lu64 = nullptr;
    float* lu32 = nullptr;
    int* perm64 = nullptr;
    int* perm32 = nullptr;
    bool lu_ok64 = false, lu_ok32 = false;
    DLev* d_tail64 = nullptr;
    DLevF* d_tail32 = nullptr;
    std::vector<float*> up32;     /* FP32 block levels: the up leg's output (Level::up) */
    /* exact spectral solve of level spec0 (-1: none), used by the solves without a
     * Helmholtz term in place of the sub-V-cycle below it */
    int spec0 = -1;
    fxd::SpecTail<double> spec64{};
    fxd::SpecTail<float> spec32{};
    double* spec_basis64 = nullptr;   /* Qx, Qy, mu_x, mu_y of level spec0, then two scratch arrays of its size */
    float* spec_basis32 = nullptr;
    /* direct solve of the finest level when it is laterally periodic and
     * z-layered (fx_fft.hip), used by the solves without a Helmholtz term */
    bool direct = false;
    fxd::FftPlan<double> fft64{};
    fxd::FftPlan<float> fft32{};
    double* fft_mem64 = nullptr;      /* twiddles, eigenvalues, S, the z-pivot table */
    /* preconditioned CG (semiconductor stacks on the direct path's geometry):
     * r, z, p, q (N each), the plane means of the A-coefficient (n2), the
     * preconditioner's z-pivots, three sets of partials, four scalars (r.z of
     * two iterations, the converged flag, the target), the plane-mean partials */
    double* pcg_mem = nullptr;
    float* fft_mem32 = nullptr;
    double fft_w2lo = 0.0, fft_w2hi = 0.0;   /* 2 w of the first / last plane's Dirichlet face (0: Neumann) */
    int tail_ts64 = 0, tail_ts32 = 0;          /* first tail level staged in LDS */
    size_t tail_lds64 = 0, tail_lds32 = 0;     /* dynamic LDS of the tail kernel */
    /* FP32 correction hierarchy of the mixed-precision solve (precision = FP32_REFINE) */
    std::vector<DLevF> lev32;
    float* cg_work32 = nullptr;
    double* phi_alt = nullptr;     /* second phi buffer of the refinement update */
    /* host shadows */
    std::vector<uint8_t> h_mat;
    std::vector<double> h_tphase, h_ang[3], h_lalpha;
    bool has_sc = false;
    bool pxy_zero = true;
    /* run state (main.cpp:238-245) */
    double Phi_Bc_lo = 0, Phi_Bc_hi = 0;
    double bc_solve[6] = {0, 0, 0, 0, 0, 0};
    /* differentiation form: checkpoints of the last fx_diff_forward */
    struct Diff {
        int nsteps = 0, K = 0, nck = 0, nfe = 0;
        bool valid = false;
        std::vector<double> V;
        double* ck = nullptr;        /* nck x N: P at steps 0, K, 2K, ... */
        double* segP = nullptr;      /* K x N: P_n of the segment being reversed */
        double* segPhi = nullptr;    /* K x N: phi_n of that segment */
        double* lam[2] = {nullptr, nullptr};
        double* s = nullptr;         /* Fz^T nu */
        double* fin = nullptr;       /* P after the last forward step */
        double* acc = nullptr;       /* 6 parameter gradients, then dL/dV per step */
        double* m = nullptr;         /* FE sums per step */
        double* part = nullptr;      /* 7 x adj_blocks partials */
        size_t cap_ck = 0, cap_segP = 0, cap_segPhi = 0, cap_acc = 0, cap_m = 0;   /* bytes allocated */
    } diff;
    std::string plot_dir = ".";   /* where plt%08d directories go */
    int step = 0, sign = 1, num_Vapp = 0, steady_state_step = 1000000, inc_step = -1;
    double time = 0.0;
    std::vector<double> waveform;
    SCParams sc{};
    std::string err;
};

namespace {

int fail(fx_ctx* x, int code, const std::string& m) {
    if (x) x->err = m;
    return code;
}

#define HIPCHK(x, call)                                                                     \
    do {                                                                                    \
        hipError_t e_ = (call);                                                             \
        if (e_ != hipSuccess) return fail((x), FX_EHIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

int dalloc(fx_ctx* x, void** p, size_t bytes) {
    HIPCHK(x, hipMalloc(p, bytes ? bytes : 8));
    HIPCHK(x, hipMemsetAsync(*p, 0, bytes ? bytes : 8, x->s));
    return FX_OK;
}

double phi_ref_of(const SCParams& s) { return s.Chi + 0.5 * s.Eg + 0.5 * s.kb * s.T * log(s.Nc / s.Nv) / s.q; }

/* Material ids on the grown box from cell-centre coordinates, FE > DE > SC
 * (> Channel), spacer = DE, then the periodic images (InitializeMaterialMask,
 * Initialization.cpp:169-206). */
void mask_from_boxes(fx_ctx* x, std::vector<uint8_t>& m) {
    const fx_inputs& in = x->in;
    const int nx = x->n[0], ny = x->n[1], nz = x->n[2];
    const int sx = nx + 2, sy = sx * (ny + 2);
    m.assign((size_t)sy * (nz + 2), 1);
    auto inbox = [](const double* lo, const double* hi, double X, double Y, double Z) {
        return X <= hi[0] && X >= lo[0] && Y <= hi[1] && Y >= lo[1] && Z <= hi[2] && Z >= lo[2];
    };
    for (int k = -1; k <= nz; ++k)
        for (int j = -1; j <= ny; ++j)
            for (int i = -1; i <= nx; ++i) {
                const double X = in.prob_lo[0] + (i + 0.5) * x->dx[0];
                const double Y = in.prob_lo[1] + (j + 0.5) * x->dx[1];
                const double Z = in.prob_lo[2] + (k + 0.5) * x->dx[2];
                uint8_t v;
                if (inbox(in.FE_lo, in.FE_hi, X, Y, Z)) v = 0;
                else if (inbox(in.DE_lo, in.DE_hi, X, Y, Z)) v = 1;
                else if (inbox(in.SC_lo, in.SC_hi, X, Y, Z)) v = inbox(in.Channel_lo, in.Channel_hi, X, Y, Z) ? 3 : 2;
                else v = 1;
                m[(size_t)(i + 1) + sx * (size_t)(j + 1) + sy * (size_t)(k + 1)] = v;
            }
}

/* Ghost cells whose periodic image is a valid cell take its value. */
void periodic_fill_u8(const fx_ctx* x, std::vector<uint8_t>& m) {
    const int n[3] = {x->n[0], x->n[1], x->n[2]};
    const int sx = n[0] + 2, sy = sx * (n[1] + 2);
    for (int k = -1; k <= n[2]; ++k)
        for (int j = -1; j <= n[1]; ++j)
            for (int i = -1; i <= n[0]; ++i) {
                int id[3] = {i, j, k};
                bool ghost = false, ok = true;
                for (int d = 0; d < 3; ++d)
                    if (id[d] < 0 || id[d] >= n[d]) {
                        ghost = true;
                        if (x->in.is_periodic[d]) id[d] = (id[d] + n[d]) % n[d]; else ok = false;
                    }
                if (ghost && ok)
                    m[(size_t)(i + 1) + sx * (size_t)(j + 1) + sy * (size_t)(k + 1)] =
                        m[(size_t)(id[0] + 1) + sx * (size_t)(id[1] + 1) + sy * (size_t)(id[2] + 1)];
            }
}

inline size_t gidx(const fx_ctx* x, int i, int j, int k) {
    return (size_t)(i + 1) + (size_t)(x->n[0] + 2) * ((size_t)(j + 1) + (size_t)(x->n[1] + 2) * (k + 1));
}

/* Coarsening rule shared with the oracle's specification: direction d is
 * coarsened when n_d is even and h_d < mg_coarsen_ratio*min(h) (default
 * 1.9); a level with an odd
 * periodic extent > 1, or with nothing left to coarsen, is the bottom. */
void plan_levels(fx_ctx* x, std::vector<std::array<int, 3>>& dims, std::vector<std::array<double, 3>>& hs,
                 std::vector<std::array<int, 3>>& crs) {
    std::array<int, 3> n = {x->n[0], x->n[1], x->n[2]};
    std::array<double, 3> h = {x->dx[0], x->dx[1], x->dx[2]};
    dims.push_back(n); hs.push_back(h);
    while (dims.size() < 32) {
        bool oddp = false;
        for (int d = 0; d < 3; ++d) if (x->in.bc_lo[d] == FX_BC_PERIODIC && n[d] > 1 && (n[d] & 1)) oddp = true;
        std::array<int, 3> c = {0, 0, 0};
        if (!oddp) {
            const double hmin = std::min(h[0], std::min(h[1], h[2]));
            for (int d = 0; d < 3; ++d) c[d] = (n[d] % 2 == 0 && n[d] >= 2 && h[d] < x->in.mg_coarsen_ratio * hmin) ? 1 : 0;
        }
        const size_t Nl = (size_t)n[0] * n[1] * n[2];
        if (oddp || !(c[0] | c[1] | c[2]) || Nl <= 1) { crs.push_back({0, 0, 0}); break; }
        crs.push_back(c);
        for (int d = 0; d < 3; ++d) if (c[d]) { n[d] /= 2; h[d] *= 2.0; }
        dims.push_back(n); hs.push_back(h);
    }
    if (crs.size() < dims.size()) crs.push_back({0, 0, 0});
}

int build_levels(fx_ctx* x) {
    std::vector<std::array<int, 3>> dims, crs;
    std::vector<std::array<double, 3>> hs;
    plan_levels(x, dims, hs, crs);
    x->lev.resize(dims.size());
    for (size_t l = 0; l < dims.size(); ++l) {
        Level& L = x->lev[l];
        DLev& d = L.d;
        d.n0 = dims[l][0]; d.n1 = dims[l][1]; d.n2 = dims[l][2];
        for (int q = 0; q < 3; ++q) { d.ih2[q] = 1.0 / (hs[l][q] * hs[l][q]); d.crs[q] = crs[l][q]; }
        L.N = (size_t)d.n0 * d.n1 * d.n2;
        int rc;
        if ((rc = dalloc(x, (void**)&d.bx, sizeof(double) * (size_t)(d.n0 + 1) * d.n1 * d.n2))) return rc;
        if ((rc = dalloc(x, (void**)&d.by, sizeof(double) * (size_t)d.n0 * (d.n1 + 1) * d.n2))) return rc;
        if ((rc = dalloc(x, (void**)&d.bz, sizeof(double) * (size_t)d.n0 * d.n1 * (d.n2 + 1)))) return rc;
        d.mat = nullptr; d.msx = 0; d.msy = 0; d.matcoef = 0; d.read_a = 0;
        d.layered = 0; d.bxk = d.byk = d.bzk = nullptr; d.phi2 = nullptr;
        d.stream = L.N >= (size_t)std::max(0, x->in.mg_stream_min_cells) ? 1 : 0;
        for (int q = 0; q < 8; ++q) d.eps[q] = 0.0;
        if (l == 0) {
            d.phi = x->phi; d.b = x->rhs; d.a = x->alpha_cc;
            const fx_inputs& in = x->in;
            d.mat = x->mat; d.msx = x->msx; d.msy = x->msy; d.matcoef = 1;
            for (int q = 0; q < 8; ++q) {
                const int m = q & 3, tph = (q >> 2) & 1;
                d.eps[q] = m == 0 ? (tph ? in.epsilonX_fe_tphase * in.epsilon_0 : in.epsilonX_fe * in.epsilon_0)
                         : m == 1 ? in.epsilon_de * in.epsilon_0 : in.epsilon_si * in.epsilon_0;
            }
        } else {
            if ((rc = dalloc(x, (void**)&d.phi, sizeof(double) * L.N))) return rc;
            if ((rc = dalloc(x, (void**)&d.b, sizeof(double) * L.N))) return rc;
            if ((rc = dalloc(x, (void**)&d.a, sizeof(double) * L.N))) return rc;
        }
    }
    int rc;
    const size_t Nb = x->lev.back().N;
    const size_t bottom_work = std::max<size_t>(5 * Nb, Nb <= (size_t)fxk::BOTTOM_DIRECT_CELLS ? Nb * Nb + Nb : 0);
    if ((rc = dalloc(x, (void**)&x->cg_work, sizeof(double) * bottom_work))) return rc;
    x->tail0 = -1;
    for (int l = (int)x->lev.size() - 1; l >= 1 && x->lev[l].N <= TAIL_MAX_CELLS; --l) x->tail0 = l;
    if (x->tail0 >= 0 && (int)x->lev.size() - x->tail0 < 2) x->tail0 = -1;
    x->tail_base = x->tail0;
    if (Nb <= (size_t)fxk::BOTTOM_DIRECT_CELLS) {
        if ((rc = dalloc(x, (void**)&x->lu64, sizeof(double) * Nb * Nb))) return rc;
        if ((rc = dalloc(x, (void**)&x->perm64, sizeof(int) * (Nb + 1)))) return rc;
        if (x->in.precision == FX_PREC_FP32_REFINE) {
            if ((rc = dalloc(x, (void**)&x->lu32, sizeof(float) * Nb * Nb))) return rc;
            if ((rc = dalloc(x, (void**)&x->perm32, sizeof(int) * (Nb + 1)))) return rc;
        }
    }
    if (x->tail0 >= 0) {
        if ((rc = dalloc(x, (void**)&x->d_tail64, sizeof(DLev) * x->lev.size()))) return rc;
        if (x->in.precision == FX_PREC_FP32_REFINE && (rc = dalloc(x, (void**)&x->d_tail32, sizeof(DLevF) * x->lev.size()))) return rc;
    }
    /* levels that may run block legs (below the finest, above the tail) */
    auto block_level = [&](size_t l) { return x->in.mg_block_levels && l >= 1 && (x->tail0 < 0 || (int)l < x->tail0); };
    for (size_t l = 0; l < x->lev.size(); ++l)
        if (block_level(l) && (rc = dalloc(x, (void**)&x->lev[l].up, sizeof(double) * x->lev[l].N))) return rc;
    if (x->in.mg_fused_sweep)
        for (size_t l = 0; l < x->lev.size(); ++l)
            if (x->lev[l].d.n0 >= 64 && x->lev[l].d.n1 >= 8 && (rc = dalloc(x, (void**)&x->lev[l].d.phi2, sizeof(double) * x->lev[l].N))) return rc;
    if (x->in.precision == FX_PREC_FP32_REFINE) {
        x->lev32.resize(x->lev.size());
        x->up32.assign(x->lev.size(), nullptr);
        for (size_t l = 0; l < x->lev.size(); ++l) {
            const DLev& d = x->lev[l].d;
            DLevF& f = x->lev32[l];
            std::memset(&f, 0, sizeof(f));
            f.n0 = d.n0; f.n1 = d.n1; f.n2 = d.n2;
            f.stream = d.stream;
            for (int q = 0; q < 3; ++q) { f.ih2[q] = (float)d.ih2[q]; f.crs[q] = d.crs[q]; }
            const size_t Nl = x->lev[l].N;
            if ((rc = dalloc(x, (void**)&f.phi, sizeof(float) * Nl))) return rc;
            if ((rc = dalloc(x, (void**)&f.b, sizeof(float) * Nl))) return rc;
            if ((rc = dalloc(x, (void**)&f.a, sizeof(float) * Nl))) return rc;
            if ((rc = dalloc(x, (void**)&f.bx, sizeof(float) * (size_t)(d.n0 + 1) * d.n1 * d.n2))) return rc;
            if ((rc = dalloc(x, (void**)&f.by, sizeof(float) * (size_t)d.n0 * (d.n1 + 1) * d.n2))) return rc;
            if ((rc = dalloc(x, (void**)&f.bz, sizeof(float) * (size_t)d.n0 * d.n1 * (d.n2 + 1)))) return rc;
            if (x->in.mg_fused_sweep && d.n0 >= 64 && d.n1 >= 8 && (rc = dalloc(x, (void**)&f.phi2, sizeof(float) * Nl))) return rc;
            if (block_level(l) && (rc = dalloc(x, (void**)&x->up32[l], sizeof(float) * Nl))) return rc;
        }
        if ((rc = dalloc(x, (void**)&x->cg_work32, sizeof(float) * bottom_work))) return rc;
        if ((rc = dalloc(x, (void**)&x->phi_alt, sizeof(double) * x->N))) return rc;
    }
    return FX_OK;
}

/* Plan of the lean tail from level t0 down; lev(l) gives level l's descriptor. */
template <class T, class F>
void fill_plan(fxd::TailPlan<T>& P, const fx_ctx* x, int t0, F lev, const T* lu, const int* perm) {
    std::memset(&P, 0, sizeof(P));
    P.nl = (int)x->lev.size() - t0;
    for (int t = 0; t < P.nl; ++t) {
        const DLevT<T>& d = lev(t0 + t);
        P.n[t][0] = d.n0; P.n[t][1] = d.n1; P.n[t][2] = d.n2;
        for (int q = 0; q < 3; ++q) { P.crs[t][q] = d.crs[q]; P.ih2[t][q] = d.ih2[q]; }
        P.bxk[t] = d.bxk; P.byk[t] = d.byk; P.bzk[t] = d.bzk;
        P.a[t] = d.a;
    }
    P.b0 = lev(t0).b;
    P.phi0 = lev(t0).phi;
    P.binv = lu;
    P.perm = perm;
}

/* Permittivity with its ghost rule (InitializePermittivity,
 * ElectrostaticSolver.cpp:205-299), face averages (eXstaticUtil.cpp:105-143)
 * and the face means of every coarser level; uploaded to the device. */
int build_direct(fx_ctx* x, double bz_lo, double bz_hi);

int build_coefficients(fx_ctx* x) {
    const fx_inputs& in = x->in;
    const int nx = x->n[0], ny = x->n[1], nz = x->n[2];
    const int sx = nx + 2, sy = sx * (ny + 2);
    std::vector<double> e((size_t)sy * (nz + 2), 0.0);
    auto E = [&](int i, int j, int k) -> double& { return e[(size_t)(i + 1) + sx * (size_t)(j + 1) + sy * (size_t)(k + 1)]; };
    std::vector<double> cell(x->N);
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                const int m = x->h_mat[gidx(x, i, j, k)] & 3;
                const size_t c = (size_t)i + (size_t)nx * (j + (size_t)ny * k);
                double v;
                if (m == 0) v = (x->h_tphase[c] == 1.0) ? in.epsilonX_fe_tphase * in.epsilon_0 : in.epsilonX_fe * in.epsilon_0;
                else if (m == 1) v = in.epsilon_de * in.epsilon_0;
                else v = in.epsilon_si * in.epsilon_0;
                E(i, j, k) = v;
                cell[c] = v;
            }
    /* periodic images, then copies across Dirichlet / Neumann faces */
    for (int k = -1; k <= nz; ++k)
        for (int j = -1; j <= ny; ++j)
            for (int i = -1; i <= nx; ++i) {
                int id[3] = {i, j, k};
                const int n[3] = {nx, ny, nz};
                bool ghost = false, ok = true;
                for (int d = 0; d < 3; ++d)
                    if (id[d] < 0 || id[d] >= n[d]) {
                        ghost = true;
                        if (in.is_periodic[d]) id[d] = (id[d] + n[d]) % n[d]; else ok = false;
                    }
                if (ghost && ok) E(i, j, k) = E(id[0], id[1], id[2]);
            }
    for (int k = -1; k <= nz; ++k)
        for (int j = -1; j <= ny; ++j)
            for (int i = -1; i <= nx; ++i) {
                if (in.bc_lo[0] != FX_BC_PERIODIC && i < 0) E(i, j, k) = E(i + 1, j, k);
                if (in.bc_hi[0] != FX_BC_PERIODIC && i > nx - 1) E(i, j, k) = E(i - 1, j, k);
                if (in.bc_lo[1] != FX_BC_PERIODIC && j < 0) E(i, j, k) = E(i, j + 1, k);
                if (in.bc_hi[1] != FX_BC_PERIODIC && j > ny - 1) E(i, j, k) = E(i, j - 1, k);
                if (in.bc_lo[2] != FX_BC_PERIODIC && k < 0) E(i, j, k) = E(i, j, k + 1);
                if (in.bc_hi[2] != FX_BC_PERIODIC && k > nz - 1) E(i, j, k) = E(i, j, k - 1);
            }
    /* finest faces */
    std::vector<std::vector<double>> fx(x->lev.size()), fy(x->lev.size()), fz(x->lev.size());
    fx[0].resize((size_t)(nx + 1) * ny * nz);
    fy[0].resize((size_t)nx * (ny + 1) * nz);
    fz[0].resize((size_t)nx * ny * (nz + 1));
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i <= nx; ++i) fx[0][(size_t)i + (size_t)(nx + 1) * (j + (size_t)ny * k)] = 0.5 * (E(i, j, k) + E(i - 1, j, k));
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j <= ny; ++j)
            for (int i = 0; i < nx; ++i) fy[0][(size_t)i + (size_t)nx * (j + (size_t)(ny + 1) * k)] = 0.5 * (E(i, j, k) + E(i, j - 1, k));
    for (int k = 0; k <= nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) fz[0][(size_t)i + (size_t)nx * (j + (size_t)ny * k)] = 0.5 * (E(i, j, k) + E(i, j, k - 1));
    /* coarse faces: mean of the fine faces tiling each coarse face */
    for (size_t l = 0; l + 1 < x->lev.size(); ++l) {
        const DLev& f = x->lev[l].d;
        const DLev& c = x->lev[l + 1].d;
        const int cx = f.crs[0], cy = f.crs[1], cz = f.crs[2];
        fx[l + 1].resize((size_t)(c.n0 + 1) * c.n1 * c.n2);
        fy[l + 1].resize((size_t)c.n0 * (c.n1 + 1) * c.n2);
        fz[l + 1].resize((size_t)c.n0 * c.n1 * (c.n2 + 1));
        for (int k = 0; k < c.n2; ++k)
            for (int j = 0; j < c.n1; ++j)
                for (int i = 0; i <= c.n0; ++i) {
                    double s = 0; int cnt = 0;
                    for (int kk = 0; kk <= cz; ++kk)
                        for (int jj = 0; jj <= cy; ++jj) {
                            const int fi = cx ? 2 * i : i, fj = cy ? 2 * j + jj : j, fk = cz ? 2 * k + kk : k;
                            s += fx[l][(size_t)fi + (size_t)(f.n0 + 1) * (fj + (size_t)f.n1 * fk)]; ++cnt;
                        }
                    fx[l + 1][(size_t)i + (size_t)(c.n0 + 1) * (j + (size_t)c.n1 * k)] = s / cnt;
                }
        for (int k = 0; k < c.n2; ++k)
            for (int j = 0; j <= c.n1; ++j)
                for (int i = 0; i < c.n0; ++i) {
                    double s = 0; int cnt = 0;
                    for (int kk = 0; kk <= cz; ++kk)
                        for (int ii = 0; ii <= cx; ++ii) {
                            const int fi = cx ? 2 * i + ii : i, fj = cy ? 2 * j : j, fk = cz ? 2 * k + kk : k;
                            s += fy[l][(size_t)fi + (size_t)f.n0 * (fj + (size_t)(f.n1 + 1) * fk)]; ++cnt;
                        }
                    fy[l + 1][(size_t)i + (size_t)c.n0 * (j + (size_t)(c.n1 + 1) * k)] = s / cnt;
                }
        for (int k = 0; k <= c.n2; ++k)
            for (int j = 0; j < c.n1; ++j)
                for (int i = 0; i < c.n0; ++i) {
                    double s = 0; int cnt = 0;
                    for (int jj = 0; jj <= cy; ++jj)
                        for (int ii = 0; ii <= cx; ++ii) {
                            const int fi = cx ? 2 * i + ii : i, fj = cy ? 2 * j + jj : j, fk = cz ? 2 * k : k;
                            s += fz[l][(size_t)fi + (size_t)f.n0 * (fj + (size_t)f.n1 * fk)]; ++cnt;
                        }
                    fz[l + 1][(size_t)i + (size_t)c.n0 * (j + (size_t)c.n1 * k)] = s / cnt;
                }
    }
    for (size_t l = 0; l < x->lev.size(); ++l) {
        DLev& d = x->lev[l].d;
        /* layered level: every face coefficient depends on k only */
        bool lay = true;
        for (int k = 0; k < d.n2 && lay; ++k)
            for (int j = 0; j < d.n1 && lay; ++j)
                for (int i = 0; i <= d.n0 && lay; ++i)
                    lay = fx[l][(size_t)i + (size_t)(d.n0 + 1) * (j + (size_t)d.n1 * k)] == fx[l][(size_t)(d.n0 + 1) * ((size_t)d.n1 * k)];
        for (int k = 0; k < d.n2 && lay; ++k)
            for (int j = 0; j <= d.n1 && lay; ++j)
                for (int i = 0; i < d.n0 && lay; ++i)
                    lay = fy[l][(size_t)i + (size_t)d.n0 * (j + (size_t)(d.n1 + 1) * k)] == fy[l][(size_t)d.n0 * ((size_t)(d.n1 + 1) * k)];
        for (int k = 0; k <= d.n2 && lay; ++k)
            for (int j = 0; j < d.n1 && lay; ++j)
                for (int i = 0; i < d.n0 && lay; ++i)
                    lay = fz[l][(size_t)i + (size_t)d.n0 * (j + (size_t)d.n1 * k)] == fz[l][(size_t)d.n0 * ((size_t)d.n1 * k)];
        if (d.bxk) { (void)hipFree((void*)d.bxk); (void)hipFree((void*)d.byk); (void)hipFree((void*)d.bzk); d.bxk = d.byk = d.bzk = nullptr; }
        d.layered = lay ? 1 : 0;
        if (lay) {
            std::vector<double> kx(d.n2), ky(d.n2), kz(d.n2 + 1);
            for (int k = 0; k < d.n2; ++k) {
                kx[k] = fx[l][(size_t)(d.n0 + 1) * ((size_t)d.n1 * k)];
                ky[k] = fy[l][(size_t)d.n0 * ((size_t)(d.n1 + 1) * k)];
            }
            for (int k = 0; k <= d.n2; ++k) kz[k] = fz[l][(size_t)d.n0 * ((size_t)d.n1 * k)];
            double *pbx, *pby, *pbz;
            HIPCHK(x, hipMalloc((void**)&pbx, kx.size() * sizeof(double)));
            HIPCHK(x, hipMalloc((void**)&pby, ky.size() * sizeof(double)));
            HIPCHK(x, hipMalloc((void**)&pbz, kz.size() * sizeof(double)));
            HIPCHK(x, hipMemcpyAsync(pbx, kx.data(), kx.size() * sizeof(double), hipMemcpyHostToDevice, x->s));
            HIPCHK(x, hipMemcpyAsync(pby, ky.data(), ky.size() * sizeof(double), hipMemcpyHostToDevice, x->s));
            HIPCHK(x, hipMemcpyAsync(pbz, kz.data(), kz.size() * sizeof(double), hipMemcpyHostToDevice, x->s));
            HIPCHK(x, hipStreamSynchronize(x->s));
            d.bxk = pbx; d.byk = pby; d.bzk = pbz;
        }
        HIPCHK(x, hipMemcpyAsync(d.bx, fx[l].data(), fx[l].size() * sizeof(double), hipMemcpyHostToDevice, x->s));
        HIPCHK(x, hipMemcpyAsync(d.by, fy[l].data(), fy[l].size() * sizeof(double), hipMemcpyHostToDevice, x->s));
        HIPCHK(x, hipMemcpyAsync(d.bz, fz[l].data(), fz[l].size() * sizeof(double), hipMemcpyHostToDevice, x->s));
    }
    HIPCHK(x, hipMemcpyAsync(x->eps, cell.data(), cell.size() * sizeof(double), hipMemcpyHostToDevice, x->s));
    HIPCHK(x, hipStreamSynchronize(x->s));
    /* FP32 copies of every level's coefficients (rounded from FP64) */
    for (size_t l = 0; l < x->lev32.size(); ++l) {
        DLevF& f = x->lev32[l];
        const DLev& d = x->lev[l].d;
        auto up = [&](float* dst, const std::vector<double>& src) -> int {
            std::vector<float> t(src.size());
            for (size_t q = 0; q < src.size(); ++q) t[q] = (float)src[q];
            HIPCHK(x, hipMemcpy(dst, t.data(), t.size() * sizeof(float), hipMemcpyHostToDevice));
            return FX_OK;
        };
        int rc;
        if ((rc = up(f.bx, fx[l])) || (rc = up(f.by, fy[l])) || (rc = up(f.bz, fz[l]))) return rc;
        if (f.bxk) { (void)hipFree((void*)f.bxk); (void)hipFree((void*)f.byk); (void)hipFree((void*)f.bzk); f.bxk = f.byk = f.bzk = nullptr; }
        f.layered = d.layered;
        f.matcoef = d.matcoef;
        f.mat = d.mat; f.msx = d.msx; f.msy = d.msy;
        for (int q = 0; q < 8; ++q) f.eps[q] = (float)d.eps[q];
        if (d.layered) {
            std::vector<double> kx(d.n2), ky(d.n2), kz(d.n2 + 1);
            HIPCHK(x, hipMemcpy(kx.data(), d.bxk, kx.size() * sizeof(double), hipMemcpyDeviceToHost));
            HIPCHK(x, hipMemcpy(ky.data(), d.byk, ky.size() * sizeof(double), hipMemcpyDeviceToHost));
            HIPCHK(x, hipMemcpy(kz.data(), d.bzk, kz.size() * sizeof(double), hipMemcpyDeviceToHost));
            float *pbx, *pby, *pbz;
            HIPCHK(x, hipMalloc((void**)&pbx, kx.size() * sizeof(float)));
            HIPCHK(x, hipMalloc((void**)&pby, ky.size() * sizeof(float)));
            HIPCHK(x, hipMalloc((void**)&pbz, kz.size() * sizeof(float)));
            if ((rc = up(pbx, kx)) || (rc = up(pby, ky)) || (rc = up(pbz, kz))) return rc;
            f.bxk = pbx; f.byk = pby; f.bzk = pbz;
        }
    }
    /* the lean tail: the first level from the tail's start on from which
     * every level is layered and the whole tail fits in LDS (with a
     * Helmholtz term, so that one plan serves every solve) */
    x->tail0 = x->tail_base;
    x->tail_lean = 0;
    x->lu_ok64 = x->lu_ok32 = false;
    if (x->tail_base >= 0) {
        const size_t lmax = fxk::tail_lds_max();
        const size_t cap = lmax > 4096 ? lmax - 4096 : 0;   /* room for the kernel's static LDS */
        const int L = (int)x->lev.size();
        bool lay = true;
        for (int t = x->tail_base; t < L; ++t) lay = lay && x->lev[t].d.layered;
        for (int t0 = x->tail_base; lay && t0 + 2 <= L; ++t0) {
            const int nl = L - t0;
            if (nl > fxd::TAIL_LEAN_MAXL) continue;
            int n[fxd::TAIL_LEAN_MAXL][3];
            for (int t = 0; t < nl; ++t) {
                const DLev& d = x->lev[t0 + t].d;
                n[t][0] = d.n0; n[t][1] = d.n1; n[t][2] = d.n2;
            }
            if (fxk::tail_lay_lds(n, nl, 1, sizeof(double)) > cap) continue;
            x->tail0 = t0;
            x->tail_lean = 1;
            fill_plan(x->plan64, x, t0, [&](int l) -> const DLevT<double>& { return x->lev[l].d; }, x->lu64, x->perm64);
            if (!x->lev32.empty())
                fill_plan(x->plan32, x, t0, [&](int l) -> const DLevT<float>& { return x->lev32[l]; }, x->lu32, x->perm32);
            if (x->in.verbose)
                std::printf("multigrid tail: levels %d.. in one workgroup, all in LDS (%zu B FP64 with a Helmholtz term)\n",
                            t0, fxk::tail_lay_lds(n, nl, 1, sizeof(double)));
            break;
        }
    }
    /* the exact spectral solve that ends the V-cycle: on a laterally periodic
     * hierarchy bounded in z with a Dirichlet face, the finest level below the
     * finest whose planes fit one workgroup (n0 n1 <= SPEC_MAX_PLANE, n0, n1 <=
     * SPEC_MAX_N), of at most SPEC_MAX_CELLS cells, with coefficients that
     * depend on z only (the choice the oracle makes from the same rule) */
    x->spec0 = -1;
    if (x->in.mg_exact_tail && x->in.is_periodic[0] && x->in.is_periodic[1] && !x->in.is_periodic[2] &&
        (x->in.bc_lo[2] == FX_BC_DIRICHLET || x->in.bc_hi[2] == FX_BC_DIRICHLET)) {
        int lsel = -1;
        for (size_t l = 1; l < x->lev.size() && lsel < 0; ++l) {
            const DLev& d = x->lev[l].d;
            if (d.n0 <= fxd::SPEC_MAX_N && d.n1 <= fxd::SPEC_MAX_N && d.n0 * d.n1 <= fxd::SPEC_MAX_PLANE &&
                x->lev[l].N <= (size_t)fxd::SPEC_MAX_CELLS)
                lsel = (int)l;
        }
        if (lsel >= 1 && x->lev[lsel].d.layered) {
            const DLev& d = x->lev[lsel].d;
            const size_t nb = (size_t)d.n0 * d.n0 + (size_t)d.n1 * d.n1 + d.n0 + d.n1;
            const size_t N = x->lev[lsel].N;
            std::vector<double> h(nb);
            double* Qx = h.data();
            double* Qy = Qx + (size_t)d.n0 * d.n0;
            double* mux = Qy + (size_t)d.n1 * d.n1;
            double* muy = mux + d.n0;
            int rc = fx_batch_basis(d.n0, Qx, mux);
            if (!rc) rc = fx_batch_basis(d.n1, Qy, muy);
            if (rc) return fail(x, rc, "spectral coarse solve: basis");
            /* basis, then the two scratch arrays */
            if (!x->spec_basis64 && (rc = dalloc(x, (void**)&x->spec_basis64, (nb + 2 * N) * sizeof(double)))) return rc;
            HIPCHK(x, hipMemcpy(x->spec_basis64, h.data(), nb * sizeof(double), hipMemcpyHostToDevice));
            if (!x->lev32.empty()) {
                std::vector<float> hf(h.begin(), h.end());
                if (!x->spec_basis32 && (rc = dalloc(x, (void**)&x->spec_basis32, (nb + 2 * N) * sizeof(float)))) return rc;
                HIPCHK(x, hipMemcpy(x->spec_basis32, hf.data(), nb * sizeof(float), hipMemcpyHostToDevice));
            }
            auto fill = [&](auto& P, auto* base) {
                P.n0 = d.n0; P.n1 = d.n1; P.n2 = d.n2;
                P.Qx = base;
                P.Qy = base + (size_t)d.n0 * d.n0;
                P.mux = P.Qy + (size_t)d.n1 * d.n1;
                P.muy = P.mux + d.n0;
                P.S = base + nb;
                P.C = base + nb + N;
                for (int q = 0; q < 3; ++q) P.ih2[q] = d.ih2[q];
                P.dlo = x->in.bc_lo[2] == FX_BC_DIRICHLET;
                P.dhi = x->in.bc_hi[2] == FX_BC_DIRICHLET;
            };
            fill(x->spec64, x->spec_basis64);
            if (x->spec_basis32) fill(x->spec32, x->spec_basis32);
            x->spec0 = lsel;
        }
    }
    /* descriptors of the tail levels, read by the single-workgroup kernel */
    if (x->tail0 >= 0 && !x->tail_lean) {
        /* stage as many of the tail's levels (the smallest ones) in LDS as
         * fit, with the bottom solve's work arrays */
        const size_t lmax = fxk::tail_lds_max();
        const int nl = (int)x->lev.size() - x->tail0;
        const size_t Nb = x->lev.back().N;
        const size_t bottom_work = std::max<size_t>(5 * Nb, Nb <= (size_t)fxk::BOTTOM_DIRECT_CELLS ? Nb * Nb + Nb : 0);
        auto values = [&](int ts) {
            size_t v = bottom_work;
            for (int t = ts; t < nl; ++t) v += fxk::tail_level_values(x->lev[x->tail0 + t].d);
            return v;
        };
        x->tail_ts64 = nl; x->tail_lds64 = 0;
        x->tail_ts32 = nl; x->tail_lds32 = 0;
        for (int ts = nl - 1; ts >= 0; --ts) {
            if (values(ts) * sizeof(double) <= lmax) { x->tail_ts64 = ts; x->tail_lds64 = values(ts) * sizeof(double); }
            if (values(ts) * sizeof(float) <= lmax) { x->tail_ts32 = ts; x->tail_lds32 = values(ts) * sizeof(float); }
        }
        if (x->in.verbose)
            std::printf("multigrid tail: levels %d.. in one workgroup, LDS limit %zu B, staged from tail level %d (%zu B FP64)\n",
                        x->tail0, lmax, x->tail_ts64, x->tail_lds64);
        const int nt = (int)x->lev.size() - x->tail0;
        std::vector<DLev> t64(nt);
        for (int q = 0; q < nt; ++q) t64[q] = x->lev[x->tail0 + q].d;
        HIPCHK(x, hipMemcpy(x->d_tail64, t64.data(), nt * sizeof(DLev), hipMemcpyHostToDevice));
        if (x->d_tail32) {
            std::vector<DLevF> t32(nt);
            for (int q = 0; q < nt; ++q) t32[q] = x->lev32[x->tail0 + q];
            HIPCHK(x, hipMemcpy(x->d_tail32, t32.data(), nt * sizeof(DLevF), hipMemcpyHostToDevice));
        }
    }
    return build_direct(x, fz[0][0], fz[0][(size_t)nx * ny * nz]);
}

/* The direct solve of the finest level (fx_fft.hip) where it applies: x and y
 * periodic, z bounded with a Dirichlet face, the face coefficients a
 * function of the plane only, lateral extents powers of two (4..1024), and
 * the poisson_direct input set.  bz_lo, bz_hi: the z-face coefficients of
 * the first and last plane's boundary faces. */
int build_direct(fx_ctx* x, double bz_lo, double bz_hi) {
    const fx_inputs& in = x->in;
    const DLev& d = x->lev[0].d;
    const bool dlo = in.bc_lo[2] == FX_BC_DIRICHLET, dhi = in.bc_hi[2] == FX_BC_DIRICHLET;
    x->direct = in.poisson_direct && in.is_periodic[0] && in.is_periodic[1] && !in.is_periodic[2] && (dlo || dhi) &&
                d.layered && fxk::fft_size_ok(d.n0) && fxk::fft_size_ok(d.n1);
    if (!x->direct) return FX_OK;
    const int n0 = d.n0, n1 = d.n1, n2 = d.n2;
    const size_t ntab = 2 * (size_t)n0 + 2 * (size_t)n1;
    const size_t npiv = fxk::fft_pivot_count(n0, n1, n2);
    const size_t total = ntab + x->N + npiv;
    const bool twoway = n2 == 128 && ((size_t)n0 * n1) % 128 == 0;   /* upward pivots too (FP64) */
    std::vector<double> tab(ntab);
    double* tw0 = tab.data();
    double* tw1 = tw0 + n0;
    double* mux = tw1 + n1;
    double* muy = mux + n0;
    const double pi = 3.14159265358979323846;
    auto fill_dir = [&](int n, double* tw, double* mu) {
        for (int q = 0; q < n / 2; ++q) { tw[2 * q] = std::cos(2.0 * pi * q / n); tw[2 * q + 1] = -std::sin(2.0 * pi * q / n); }
        /* 2 - 2 cos(2 pi m / n), equal for m and n - m (the pivot table is shared by the two) */
        for (int m = 0; m < n; ++m) { const double sn = std::sin(pi * std::min(m, n - m) / n); mu[m] = 4.0 * sn * sn; }
    };
    fill_dir(n0, tw0, mux);
    fill_dir(n1, tw1, muy);
    int rc;
    if (!x->fft_mem64 && (rc = dalloc(x, (void**)&x->fft_mem64, (total + (twoway ? npiv + 2 * (size_t)n2 : 0)) * sizeof(double))))
        return rc;
    HIPCHK(x, hipMemcpy(x->fft_mem64, tab.data(), ntab * sizeof(double), hipMemcpyHostToDevice));
    const bool fp32 = in.precision == FX_PREC_FP32_REFINE && !x->lev32.empty() && x->lev32[0].bxk;
    if (fp32) {
        std::vector<float> tf(tab.begin(), tab.end());
        if (!x->fft_mem32 && (rc = dalloc(x, (void**)&x->fft_mem32, total * sizeof(float)))) return rc;
        HIPCHK(x, hipMemcpy(x->fft_mem32, tf.data(), ntab * sizeof(float), hipMemcpyHostToDevice));
    }
    int l0 = 0, l1 = 0;
    while ((1 << l0) < n0) ++l0;
    while ((1 << l1) < n1) ++l1;
    auto fill = [&](auto& P, auto* base, auto* bxk, auto* byk, auto* bzk) {
        P.n0 = n0; P.n1 = n1; P.n2 = n2; P.l0 = l0; P.l1 = l1;
        P.tw0 = base; P.tw1 = base + n0; P.mux = base + n0 + n1; P.muy = base + 2 * n0 + n1;
        P.bxk = bxk; P.byk = byk; P.bzk = bzk;
        for (int q = 0; q < 3; ++q) P.ih2[q] = d.ih2[q];
        P.dlo = dlo; P.dhi = dhi;
        P.S = base + ntab;
        P.piv = base + ntab + x->N;
        P.pivu = nullptr;
        P.green = nullptr;
    };
    fill(x->fft64, x->fft_mem64, d.bxk, d.byk, d.bzk);
    if (fp32) fill(x->fft32, x->fft_mem32, x->lev32[0].bxk, x->lev32[0].byk, x->lev32[0].bzk);
    fxk::fft_pivots(x->s, x->fft64, x->fft_mem64 + ntab + x->N, fp32 ? x->fft_mem32 + ntab + x->N : nullptr);
    if (twoway) {
        fxk::fft_pivots_up(x->s, x->fft64, x->fft_mem64 + total);
        x->fft64.pivu = x->fft_mem64 + total;
        fxk::fft_face_green(x->s, x->fft64, x->fft_mem64 + total + npiv);
        x->fft64.green = x->fft_mem64 + total + npiv;
    }
    HIPCHK(x, hipStreamSynchronize(x->s));
    /* the inhomogeneous face term of the row operator is (2 w) v, w = beta_face / hz^2 */
    x->fft_w2lo = dlo ? 2.0 * (bz_lo * d.ih2[2]) : 0.0;
    x->fft_w2hi = dhi ? 2.0 * (bz_hi * d.ih2[2]) : 0.0;
    if (in.verbose) std::printf("direct Poisson solve: %d x %d lateral Hartley transforms, %d planes\n", n0, n1, n2);
    return FX_OK;
}

/* The device copy of the material ids carries bit 3 (MAT_INTERIOR_FE) on FE
 * cells whose polarization stencils and grad(phi) are all the centred ones:
 * the 18 face and edge neighbours are FE cells inside the domain or across a
 * periodic face, and k is neither the first nor the last plane (where the
 * z-derivative of phi is one-sided); bit 4 or 5 (MAT_ZFACE_LO / _HI) on the
 * FE cells of such planes and lateral neighbourhoods whose only masked
 * stencils are the z ones of one side.  The host shadow keeps bits 0-2. */
void mark_interior_fe(const fx_ctx* x, std::vector<uint8_t>& dm) {
    const int nx = x->n[0], ny = x->n[1], nz = x->n[2];
    const bool px = x->in.is_periodic[0], py = x->in.is_periodic[1];
    for (int k = 1; k + 1 < nz; ++k)
        for (int j = 0; j < ny; ++j) {
            if (!py && (j < 1 || j > ny - 2)) continue;
            for (int i = 0; i < nx; ++i) {
                if (!px && (i < 1 || i > nx - 2)) continue;
                const size_t g = gidx(x, i, j, k);
                if (x->h_mat[g] & 3) continue;
                bool ok = true;
                for (int dk = -1; dk <= 1 && ok; ++dk)
                    for (int dj = -1; dj <= 1 && ok; ++dj)
                        for (int di = -1; di <= 1 && ok; ++di) {
                            const int nn = (di != 0) + (dj != 0) + (dk != 0);
                            if (nn == 0 || nn == 3) continue;
                            ok = (x->h_mat[gidx(x, i + di, j + dj, k + dk)] & 3) == 0;
                        }
                if (ok) { dm[g] |= fxd::MAT_INTERIOR_FE; continue; }
                /* next to a z face only: the cell's in-plane neighbours FE, and
                 * of the five cells below it (its own and its four lateral
                 * neighbours' z neighbours) and the five above, one group all FE
                 * and the other none, on a side whose flag is 0..3 */
                bool plane_ok = true;
                int below = 0, above = 0;
                for (int dj = -1; dj <= 1; ++dj)
                    for (int di = -1; di <= 1; ++di) {
                        if ((di || dj) && (x->h_mat[gidx(x, i + di, j + dj, k)] & 3)) plane_ok = false;
                        if (di && dj) continue;
                        below += (x->h_mat[gidx(x, i + di, j + dj, k - 1)] & 3) == 0;
                        above += (x->h_mat[gidx(x, i + di, j + dj, k + 1)] & 3) == 0;
                    }
                if (!plane_ok) continue;
                if (below == 0 && above == 5 && x->in.P_BC_flag_lo[2] >= 0 && x->in.P_BC_flag_lo[2] <= 3)
                    dm[g] |= fxd::MAT_ZFACE_LO;
                else if (above == 0 && below == 5 && x->in.P_BC_flag_hi[2] >= 0 && x->in.P_BC_flag_hi[2] <= 3)
                    dm[g] |= fxd::MAT_ZFACE_HI;
            }
        }
}
int upload_mat(fx_ctx* x) {
    std::vector<uint8_t> dm(x->h_mat);
    mark_interior_fe(x, dm);
    HIPCHK(x, hipMemcpyAsync(x->mat, dm.data(), dm.size(), hipMemcpyHostToDevice, x->s));
    /* FE cells whose stencils are not all centred: the fused TDGL + RHS
     * update forms them on their own (k_tdgl_face_cells, k_rhs_face_cells) */
    std::vector<int> face, face2;
    for (int k = 0; k < x->n[2]; ++k)
        for (int j = 0; j < x->n[1]; ++j)
            for (int i = 0; i < x->n[0]; ++i) {
                const uint8_t r = dm[gidx(x, i, j, k)];
                if ((r & 3) == 0 && !(r & fxd::MAT_INTERIOR_FE)) face.push_back(i + x->n[0] * (j + x->n[1] * k));
                if ((r & 3) == 0 && !(r & (fxd::MAT_INTERIOR_FE | fxd::MAT_ZFACE_LO | fxd::MAT_ZFACE_HI)))
                    face2.push_back(i + x->n[0] * (j + x->n[1] * k));
            }
    const int plane = x->n[0] * x->n[1];
    /* a list of cells (in plane order) followed by the planes that hold them */
    auto upload = [&](std::vector<int>& f, int*& dev, int& ncells, int& nplanes) -> int {
        if (dev) { (void)hipFree(dev); dev = nullptr; }
        ncells = (int)f.size();
        nplanes = 0;
        for (int q = 0; q < ncells; ++q)
            if (q == 0 || f[q] / plane != f[q - 1] / plane) { f.push_back(f[q] / plane); ++nplanes; }
        if (!f.empty()) {
            HIPCHK(x, hipMalloc((void**)&dev, f.size() * sizeof(int)));
            HIPCHK(x, hipMemcpy(dev, f.data(), f.size() * sizeof(int), hipMemcpyHostToDevice));
        }
        return FX_OK;
    };
    int rc = upload(face, x->face_list, x->nface, x->nface_planes);
    if (!rc) rc = upload(face2, x->face2_list, x->nface2, x->nface2_planes);
    if (rc) return rc;
    x->has_sc = false;
    for (int k = 0; k < x->n[2]; ++k)
        for (int j = 0; j < x->n[1]; ++j)
            for (int i = 0; i < x->n[0]; ++i)
                if ((x->h_mat[gidx(x, i, j, k)] & 3) >= 2) x->has_sc = true;
    x->fe_k0 = x->fe_k1 = -1;
    for (int k = 0; k < x->n[2]; ++k)
        for (int j = 0; j < x->n[1]; ++j)
            for (int i = 0; i < x->n[0]; ++i)
                if ((x->h_mat[gidx(x, i, j, k)] & 3) == 0) {
                    if (x->fe_k0 < 0) x->fe_k0 = k;
                    x->fe_k1 = k;
                }
    x->rhs_next = false;
    x->rhs_clean = x->rhs_nx_clean = false;   /* the FE cells moved: the zeros outside them are not known */
    HIPCHK(x, hipStreamSynchronize(x->s));
    return FX_OK;
}

/* Rotation and Landau alpha: uniform, or a table of the distinct
 * (angle_alpha, angle_beta, angle_theta, landau_alpha) rows with a per-cell
 * index.  R is evaluated on the host with the reference formula. */
int refresh_rotation(fx_ctx* x) {
    const size_t N = x->N;
    bool uniform = true;
    for (size_t c = 1; c < N && uniform; ++c)
        uniform = x->h_ang[0][c] == x->h_ang[0][0] && x->h_ang[1][c] == x->h_ang[1][0] && x->h_ang[2][c] == x->h_ang[2][0] &&
                  x->h_lalpha[c] == x->h_lalpha[0];
    rotation(x->h_ang[0][0], x->h_ang[1][0], x->h_ang[2][0], x->in.use_Euler_angles, x->Ru);
    if (x->rot_idx) { (void)hipFree(x->rot_idx); x->rot_idx = nullptr; }
    if (x->rot_tab) { (void)hipFree(x->rot_tab); x->rot_tab = nullptr; }
    if (x->rot_atab) { (void)hipFree(x->rot_atab); x->rot_atab = nullptr; }
    x->nrot = 0;
    if (uniform && x->h_lalpha[0] == x->in.alpha) return FX_OK;
    /* rows in order of first appearance, keyed on the bit patterns */
    struct Key {
        uint64_t w[4];
        bool operator==(const Key& o) const { return w[0] == o.w[0] && w[1] == o.w[1] && w[2] == o.w[2] && w[3] == o.w[3]; }
    };
    struct KeyHash {
        size_t operator()(const Key& k) const {
            uint64_t h = 0x9e3779b97f4a7c15ull;
            for (int q = 0; q < 4; ++q) { h ^= k.w[q] + 0x9e3779b97f4a7c15ull + (h << 6) + (h >> 2); }
            return (size_t)h;
        }
    };
    std::unordered_map<Key, int, KeyHash> rows;
    std::vector<int> idx(N);
    std::vector<double> tab, atab;
    for (size_t c = 0; c < N; ++c) {
        const double v[4] = {x->h_ang[0][c], x->h_ang[1][c], x->h_ang[2][c], x->h_lalpha[c]};
        Key k;
        std::memcpy(k.w, v, sizeof v);
        auto it = rows.find(k);
        if (it == rows.end()) {
            it = rows.emplace(k, (int)atab.size()).first;
            double R[9];
            rotation(v[0], v[1], v[2], x->in.use_Euler_angles, R);
            tab.insert(tab.end(), R, R + 9);
            atab.push_back(v[3]);
        }
        idx[c] = it->second;
    }
    HIPCHK(x, hipMalloc((void**)&x->rot_idx, N * sizeof(int)));
    HIPCHK(x, hipMalloc((void**)&x->rot_tab, tab.size() * sizeof(double)));
    HIPCHK(x, hipMalloc((void**)&x->rot_atab, atab.size() * sizeof(double)));
    HIPCHK(x, hipMemcpyAsync(x->rot_idx, idx.data(), N * sizeof(int), hipMemcpyHostToDevice, x->s));
    HIPCHK(x, hipMemcpyAsync(x->rot_tab, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice, x->s));
    HIPCHK(x, hipMemcpyAsync(x->rot_atab, atab.data(), atab.size() * sizeof(double), hipMemcpyHostToDevice, x->s));
    HIPCHK(x, hipStreamSynchronize(x->s));
    x->nrot = (int)atab.size();
    return FX_OK;
}

/* ------------------------------ argument blocks --------------------------- */
Geom geom_of(const fx_ctx* x) {
    Geom g;
    for (int d = 0; d < 3; ++d) {
        g.n[d] = x->n[d]; g.per[d] = x->in.is_periodic[d]; g.dx[d] = x->dx[d];
        g.prob_lo[d] = x->in.prob_lo[d]; g.prob_hi[d] = x->in.prob_hi[d];
    }
    return g;
}
MView mview_of(const fx_ctx* x) { return MView{x->mat, x->msx, x->msy}; }
PBC pbc_of(const fx_ctx* x) {
    PBC b;
    for (int d = 0; d < 3; ++d) { b.lo[d] = x->in.P_BC_flag_lo[d]; b.hi[d] = x->in.P_BC_flag_hi[d]; }
    b.lambda = x->in.lambda;
    return b;
}
RotArgs rot_of(const fx_ctx* x) {
    RotArgs r;
    r.idx = x->rot_idx; r.tab = x->rot_tab; r.atab = x->rot_atab; r.nrows = x->nrot;
    std::memcpy(r.Ru, x->Ru, sizeof(r.Ru));
    return r;
}
/* Finish the planes of phi a partial direct solve left in its workspace. */
void phi_sync(fx_ctx* x) {
    if (x->phi_z0 > x->phi_z1) return;
    fxk::fft_finish(x->s, x->fft64, x->phi, x->phi_z0, x->phi_z1, x->phi_cont ? x->phi_face : nullptr);
    x->phi_z0 = 0;
    x->phi_z1 = -1;
    x->phi_cont = false;
}

PhiView phiview_of(const fx_ctx* x) {
    PhiView p;
    p.p = x->phi;
    for (int d = 0; d < 3; ++d) { p.n[d] = x->n[d]; p.lo[d] = x->in.bc_lo[d]; p.hi[d] = x->in.bc_hi[d]; }
    std::memcpy(p.v, x->bc_solve, sizeof(p.v));
    p.ghost_face = x->in.phi_ghost_mode == FX_GHOST_FACE;
    return p;
}
void bc_values(const fx_ctx* x, double v[6]) {
    const double phi_ref = x->sc.phi_ref;
    const double phi_m = x->in.use_work_function ? x->in.metal_work_function : phi_ref;
    v[0] = 0.0; v[1] = 0.0; v[2] = x->Phi_Bc_lo;
    v[3] = 0.0; v[4] = 0.0; v[5] = x->Phi_Bc_hi - (phi_m - phi_ref);   /* SetPhiBC_z, ElectrostaticSolver.cpp:512-520 */
}
DBC dbc_of(const fx_ctx* x, const double v[6]) {
    DBC b;
    for (int d = 0; d < 3; ++d) { b.lo[d] = x->in.bc_lo[d]; b.hi[d] = x->in.bc_hi[d]; }
    std::memcpy(b.v, v, sizeof(b.v));
    return b;
}

/* The first nq values of red_out, once the work queued before them is done.
 * A one-lane kernel writes them into coherent pinned host memory followed by
 * a sequence number, and the host spins on that number (a stream
 * synchronisation wakes the host much later than the GPU finishes); the
 * stream is polled now and then so that an error ends the wait, and a wait
 * of several seconds falls back to the synchronisation. */
unsigned post_red(fx_ctx* x, int nq, const double* src = nullptr) {
    const unsigned v = ++x->pub_seq;
    fxk::publish(x->s, src ? src : x->red_out, nq, x->d_red_host, reinterpret_cast<unsigned*>(x->d_red_host + 8), v);
    return v;
}
/* Waits for the values published under sequence number v (work queued after
 * the publishing kernel does not delay them). */
int wait_red(fx_ctx* x, unsigned v, int nq, double* out) {
    HIPCHK(x, hipGetLastError());
    volatile unsigned* seq = reinterpret_cast<volatile unsigned*>(x->h_red + 8);
    const auto t0 = std::chrono::steady_clock::now();
    for (unsigned spins = 1; *seq != v; ++spins) {
        if ((spins & 1023) == 0) {
            const hipError_t q = hipStreamQuery(x->s);
            if (q != hipSuccess && q != hipErrorNotReady) return fail(x, FX_EHIP, hipGetErrorString(q));
            if (q == hipSuccess || std::chrono::steady_clock::now() - t0 > std::chrono::seconds(5)) break;
        }
    }
    if (*seq != v) {
        HIPCHK(x, hipStreamSynchronize(x->s));
        if (*seq != v) return fail(x, FX_EHIP, "reduction results did not reach the host");
    }
    std::atomic_thread_fence(std::memory_order_acquire);
    for (int q = 0; q < nq; ++q) out[q] = x->h_red[q];
    return FX_OK;
}
int read_red(fx_ctx* x, int nq, double* out, const double* src = nullptr) { return wait_red(x, post_red(x, nq, src), nq, out); }

/* CG iterations in flight: each publishes max|r| into its own slot */
constexpr int PCG_SLOTS = 4;
double* pcg_slot_host(fx_ctx* x, int w) { return x->h_red + 16 + 2 * w; }
double* pcg_slot_dev(fx_ctx* x, int w) { return x->d_red_host + 16 + 2 * w; }
/* wait_red for a CG slot */
int wait_slot(fx_ctx* x, int w, unsigned v, double* out) {
    HIPCHK(x, hipGetLastError());
    volatile unsigned* seq = reinterpret_cast<volatile unsigned*>(pcg_slot_host(x, w) + 1);
    const auto t0 = std::chrono::steady_clock::now();
    for (unsigned spins = 1; *seq != v; ++spins) {
        if ((spins & 1023) == 0) {
            const hipError_t q = hipStreamQuery(x->s);
            if (q != hipSuccess && q != hipErrorNotReady) return fail(x, FX_EHIP, hipGetErrorString(q));
            if (q == hipSuccess || std::chrono::steady_clock::now() - t0 > std::chrono::seconds(5)) break;
        }
    }
    if (*seq != v) {
        HIPCHK(x, hipStreamSynchronize(x->s));
        if (*seq != v) return fail(x, FX_EHIP, "reduction results did not reach the host");
    }
    std::atomic_thread_fence(std::memory_order_acquire);
    *out = pcg_slot_host(x, w)[0];
    return FX_OK;
}

/* --------------------------------- multigrid ------------------------------ */
/* pre_done: the first pre-smoothing sweep of level 0 has already run
 * (phi -> phi2, queued speculatively with the previous convergence check) */
void vcycle(fx_ctx* x, size_t li, const DBC& bc, bool pre_done = false) {
    Level& L = x->lev[li];
    const int inh = li == 0;
    if ((int)li == x->spec0 && !L.d.read_a) {
        fxd::SpecTail<double> P = x->spec64;
        P.b = L.d.b; P.phi = L.d.phi; P.bxk = L.d.bxk; P.byk = L.d.byk; P.bzk = L.d.bzk;
        fxk::mg_spec_solve(x->s, P);
        return;
    }
    if ((int)li == x->tail0) {
        if (x->tail_lean) {
            if (!x->lu_ok64 && x->lu64) fxk::bottom_factor(x->s, x->lev.back().d, bc, x->lu64, x->perm64);
            x->lu_ok64 = true;
            fxk::mg_tail_lay(x->s, x->plan64, bc, L.d.read_a ? 1 : 0, x->in.mg_nu1, x->in.mg_nu2, x->cg_work,
                             BOTTOM_TOL, BOTTOM_MAXIT, x->d_iters);
            return;
        }
        fxk::mg_tail(x->s, x->d_tail64, (int)(x->lev.size() - li), bc, L.d.read_a, x->in.mg_nu1, x->in.mg_nu2,
                     x->cg_work, BOTTOM_TOL, BOTTOM_MAXIT, x->d_iters, x->tail_ts64, x->tail_lds64);
        return;
    }
    if (li + 1 == x->lev.size()) {
        if (li == 0) {
            fxk::mg_resid_norm(x->s, L.d, bc, 1, x->tmp, x->red, x->red_out + 4);
            fxk::bottom_cg(x->s, L.d, bc, x->tmp, x->tmp + L.N, x->cg_work, BOTTOM_TOL, BOTTOM_MAXIT, x->d_iters);
            fxk::add(x->s, L.d.phi, x->tmp + L.N, (long long)L.N);
        } else {
            fxk::bottom_cg(x->s, L.d, bc, L.d.b, L.d.phi, x->cg_work, BOTTOM_TOL, BOTTOM_MAXIT, x->d_iters);
        }
        return;
    }
    Level& C = x->lev[li + 1];
    if (li >= 1 && x->in.mg_block_levels && L.up && !(L.d.stream && x->in.mg_fused_sweep && fxk::fused_ok(L.d, bc)) &&
        fxk::mg_block_ok(L.d, C.d, bc, x->in.mg_nu1, x->in.mg_nu2)) {
        fxk::mg_block_down(x->s, L.d, C.d, bc, x->in.mg_nu1, x->in.mg_nu2);
        vcycle(x, li + 1, bc);
        fxk::mg_block_up(x->s, L.d, C.d, bc, x->in.mg_nu1, x->in.mg_nu2, L.up);
        std::swap(L.d.phi, L.up);
        return;
    }
    /* below the finest level phi starts at 0 (the restriction's), so the first
     * sweep takes it as 0 without reading it; the restriction then need not
     * clear the next level's phi unless the tail or the bottom solve reads it */
    if (pre_done) fxk::mg_sweep_rest(x->s, L.d, bc, inh, x->in.mg_nu1 - 1, x->in.mg_fused_sweep);
    else fxk::mg_sweep(x->s, L.d, bc, inh, x->in.mg_nu1, x->in.mg_fused_sweep, li > 0 ? 1 : 0);
    const bool plain_next = (int)(li + 1) != x->tail0 && li + 2 < x->lev.size();
    fxk::mg_resid_restrict(x->s, L.d, C.d, bc, inh, !plain_next);
    vcycle(x, li + 1, bc);
    fxk::mg_prolong_sweep(x->s, L.d, C.d, bc, inh, x->in.mg_nu2, x->in.mg_fused_sweep);
}

/* Which levels read the A-coefficient: 0 none, 1 the finest level only in
 * SC cells (it comes from the Newton linearisation), 2 anywhere. */
void set_read_a(fx_ctx* x, int mode) {
    /* the bottom operator changes with its Helmholtz term */
    if (!x->lev.empty() && x->lev.back().d.read_a != (x->lev.size() == 1 ? mode : (mode ? 2 : 0)))
        x->lu_ok64 = x->lu_ok32 = false;
    for (size_t l = 0; l < x->lev.size(); ++l) x->lev[l].d.read_a = (l == 0) ? mode : (mode ? 2 : 0);
    for (size_t l = 0; l < x->lev32.size(); ++l) x->lev32[l].read_a = (l == 0) ? mode : (mode ? 2 : 0);
}

/* FP32 V-cycle on the correction equation A e = r (homogeneous BCs).  zero0:
 * level 0 starts from e = 0 (the zero-initial first red pass); coarse levels
 * always do, their phi being cleared by the restriction. */
void vcycle32(fx_ctx* x, size_t li, const DBC& bc, bool zero0 = true, bool pre_done = false) {
    DLevF& L = x->lev32[li];
    if ((int)li == x->spec0 && !L.read_a && x->spec_basis32) {
        fxd::SpecTail<float> P = x->spec32;
        P.b = L.b; P.phi = L.phi; P.bxk = L.bxk; P.byk = L.byk; P.bzk = L.bzk;
        fxk::mg_spec_solve(x->s, P);
        return;
    }
    if ((int)li == x->tail0) {
        if (x->tail_lean) {
            if (!x->lu_ok32 && x->lu32) fxk::bottom_factor(x->s, x->lev32.back(), bc, x->lu32, x->perm32);
            x->lu_ok32 = true;
            fxk::mg_tail_lay(x->s, x->plan32, bc, L.read_a ? 1 : 0, x->in.mg_refine_nu1, x->in.mg_refine_nu2, x->cg_work32, 1e-5,
                             BOTTOM_MAXIT, x->d_iters);
            return;
        }
        fxk::mg_tail(x->s, x->d_tail32, (int)(x->lev32.size() - li), bc, L.read_a, x->in.mg_refine_nu1, x->in.mg_refine_nu2,
                     x->cg_work32, 1e-5, BOTTOM_MAXIT, x->d_iters, x->tail_ts32, x->tail_lds32);
        return;
    }
    if (li + 1 == x->lev32.size()) {
        fxk::bottom_cg(x->s, L, bc, L.b, L.phi, x->cg_work32, 1e-5, BOTTOM_MAXIT, x->d_iters);
        return;
    }
    DLevF& C = x->lev32[li + 1];
    if (li >= 1 && x->in.mg_block_levels && x->up32[li] && !(L.stream && x->in.mg_fused_sweep && fxk::fused_ok32(L, bc)) &&
        fxk::mg_block_ok(L, C, bc, x->in.mg_refine_nu1, x->in.mg_refine_nu2)) {
        fxk::mg_block_down(x->s, L, C, bc, x->in.mg_refine_nu1, x->in.mg_refine_nu2);
        vcycle32(x, li + 1, bc, true);
        fxk::mg_block_up(x->s, L, C, bc, x->in.mg_refine_nu1, x->in.mg_refine_nu2, x->up32[li]);
        std::swap(L.phi, x->up32[li]);
        return;
    }
    if (pre_done) fxk::mg_sweep_rest(x->s, L, bc, x->in.mg_refine_nu1 - 1, x->in.mg_fused_sweep);
    else fxk::mg_sweep(x->s, L, bc, x->in.mg_refine_nu1, (li > 0 || zero0) ? 1 : 0, x->in.mg_fused_sweep);
    const bool plain_next = (int)(li + 1) != x->tail0 && li + 2 < x->lev32.size();
    fxk::mg_resid_restrict(x->s, L, C, bc, !plain_next);
    vcycle32(x, li + 1, bc, true);
    fxk::mg_prolong_sweep(x->s, L, C, bc, x->in.mg_refine_nu2, x->in.mg_fused_sweep);
}

/* MLMG-style solve (stopping rule ||r||_inf <= tol*max(||b||_inf, ||r0||_inf),
 * tol_abs ignored as in MLMG::solve(..., 1e-10, -1)).  Level 0 holds phi
 * (initial guess), rhs and the A-coefficient. */
int mg_solve(fx_ctx* x, const double bv[6], double tol, int maxit, fx_solve_stats* st, bool phi_zero = false) {
    const DBC bc = dbc_of(x, bv);
    Level& L0 = x->lev[0];
    double nr[2];
    /* Each convergence check reads the residual of phi, and the first
     * pre-smoothing sweep of the next V-cycle only reads phi (it writes
     * phi2): it is queued with the check, so it runs while the host waits for
     * the norms and queues the rest of the cycle.  Where level 0 streams, the
     * check is folded into that sweep (its residual is of its input; 58 us
     * against 41 + 27 us for the sweep and a separate residual pass).  When
     * the check stops the solve, phi is the answer and phi2 is not used.
     * (Running the second pre-smoothing sweep speculatively as well, into a
     * third buffer, hid the host's reaction time but was slower overall: the
     * third 67 MB buffer of the 256x256x128 level pushes the sweeps' working
     * set out of the 256 MB last-level cache.) */
    const bool spec = x->lev.size() > 1 && x->tail0 != 0 && x->in.mg_nu1 >= 1;
    bool pre = false;   /* the next V-cycle's first pre-smoothing sweep is queued */
    /* zero: phi_zero's first check, which takes phi as 0 without reading it */
    auto check = [&](bool speculate, bool zero) -> unsigned {
        if (speculate && fxk::mg_sweep_first_norm(x->s, L0.d, bc, 1, x->in.mg_fused_sweep, x->red, x->red_out, zero)) {
            pre = true;
            return post_red(x, 2);
        }
        if (zero) fxk::fill(x->s, L0.d.phi, (long long)L0.N, 0.0);
        fxk::mg_resid_norm(x->s, L0.d, bc, 1, nullptr, x->red, x->red_out);
        const unsigned q = post_red(x, 2);
        pre = speculate && fxk::mg_sweep_first(x->s, L0.d, bc, 1, x->in.mg_fused_sweep);
        return q;
    };
    unsigned v = check(spec, phi_zero);
    const bool zero_left = phi_zero && pre;   /* phi itself was never written */
    int rc = wait_red(x, v, 2, nr);
    if (rc) return rc;
    const double r0 = nr[0], bn = nr[1];
    const double target = tol * std::max(bn, r0);
    st->mg_resid0 = r0; st->mg_bnorm = bn; st->mg_resid = r0; st->vcycles = 0;
    if (!(r0 == r0) || !(bn == bn)) return fail(x, FX_ENAN, "non-finite Poisson residual or right-hand side");
    if (r0 <= target) {
        if (zero_left) fxk::fill(x->s, L0.d.phi, (long long)L0.N, 0.0);
        return FX_OK;
    }
    for (int it = 1; it <= maxit; ++it) {
        vcycle(x, 0, bc, pre);
        v = check(spec && it < maxit, false);
        if ((rc = wait_red(x, v, 2, nr))) return rc;
        st->mg_resid = nr[0]; st->vcycles = it;
        if (!(nr[0] == nr[0])) return fail(x, FX_ENAN, "multigrid produced a non-finite residual");
        if (nr[0] <= target) return FX_OK;
    }
    return fail(x, FX_ENOCONV_MG, "multigrid did not converge in mg_max_iter V-cycles");
}

/* Mixed precision: FP64 iterative refinement around FP32 V-cycles.
 * r = b - A phi is formed in FP64 and stored in FP32; n_A FP32 V-cycles run
 * back to back on the correction equation A e = r; then phi += e in FP64, the
 * FP64 residual is formed again and the same stopping rule as mg_solve is
 * checked after every refinement.  n_A is the number of cycles that cannot
 * overshoot the target: the first refinement assumes a cycle cuts the
 * residual by at most 50x (floor(log50(r/target)) - 1); later ones use the
 * contraction per cycle measured on the refinement before, rounded down, and
 * stop at the FP32 limit (a correction solved to 1e-6 of its right-hand
 * side).  As in mg_solve, the first sweep of the next correction solve is
 * queued right behind the kernel that publishes each check. */
int mg_solve_mixed(fx_ctx* x, const double bv[6], double tol, int maxit, fx_solve_stats* st, bool phi_zero = false) {
    const DBC bc = dbc_of(x, bv);
    double nr[2];
    const bool spec = x->lev32.size() > 1 && x->tail0 != 0 && x->in.mg_refine_nu1 >= 1;
    /* phi_zero: the first residual takes phi as 0 without reading it, and the
     * first correction is stored rather than added */
    bool zero = phi_zero;
    fxk::resid_to32(x->s, x->lev[0].d, bc, x->lev32[0].b, x->red, x->red_out, zero);
    unsigned v = post_red(x, 2);
    bool pre = spec && fxk::mg_sweep_first(x->s, x->lev32[0], bc, x->in.mg_fused_sweep);
    int rc = wait_red(x, v, 2, nr);
    if (rc) return rc;
    const double r0 = nr[0], bn = nr[1];
    const double target = tol * std::max(bn, r0);
    st->mg_resid0 = r0; st->mg_bnorm = bn; st->mg_resid = r0; st->vcycles = 0;
    if (!(r0 == r0) || !(bn == bn)) return fail(x, FX_ENAN, "non-finite Poisson residual or right-hand side");
    if (r0 <= target) {
        if (zero) fxk::fill(x->s, x->phi, (long long)x->N, 0.0);
        return FX_OK;
    }
    int cycles = 0;
    double r = r0, rho = 0.0;   /* measured contraction per FP32 cycle (0: none yet) */
    while (cycles < maxit) {
        int nA = 1;
        if (r > 0.0 && target > 0.0) {
            if (rho > 0.0 && rho < 1.0) {
                const double need = std::log(r / target) / -std::log(rho);
                const double cap = std::log(1e6) / -std::log(rho);
                nA = std::max(1, (int)std::floor(std::min(need, cap)));
            } else {
                const double q = std::log(r / target) / std::log(50.0);
                nA = std::max(1, (int)std::floor(q) - 1);
            }
        }
        nA = std::min(nA, maxit - cycles);
        for (int c = 0; c < nA; ++c) vcycle32(x, 0, bc, c == 0, pre && c == 0);
        cycles += nA;
        fxk::add_f2d(x->s, x->phi, x->lev32[0].phi, (long long)x->N, zero);
        zero = false;
        fxk::resid_to32(x->s, x->lev[0].d, bc, x->lev32[0].b, x->red, x->red_out);
        v = post_red(x, 2);
        pre = spec && cycles < maxit && fxk::mg_sweep_first(x->s, x->lev32[0], bc, x->in.mg_fused_sweep);
        if ((rc = wait_red(x, v, 2, nr))) return rc;
        if (r > 0.0 && nr[0] > 0.0) rho = std::pow(nr[0] / r, 1.0 / nA);
        r = nr[0];
        st->mg_resid = r; st->vcycles = cycles;
        if (!(r == r)) return fail(x, FX_ENAN, "mixed-precision multigrid produced a non-finite residual");
        if (r <= target) return FX_OK;
    }
    return fail(x, FX_ENOCONV_MG, "mixed-precision multigrid did not converge in mg_max_iter V-cycles");
}

/* Direct solve (fx_fft.hip) of the finest level without a Helmholtz term.
 * check: form the FP64 residual afterwards and apply the multigrid stopping
 * rule to it, ||r|| <= tol max(||b||, ||r0||), with ||r0|| (the zero
 * guess's residual, b plus the Dirichlet face terms) replaced by its lower
 * bound max(||b||, |face term| - ||b||), so the test is never looser; without
 * it the stats carry no norms (0). */
int direct_solve(fx_ctx* x, const double bv[6], double tol, fx_solve_stats* st, bool check, int z0, int z1, int o0,
                 int o1) {
    const double blo = x->fft_w2lo * bv[2], bhi = x->fft_w2hi * bv[5];
    if (check || x->phi_exposed) o1 = o0 - 1;   /* the residual (or a caller) reads all of phi */
    x->phi_cont = fxk::fft_solve(x->s, x->fft64, x->rhs, x->phi, blo, bhi, z0, z1, o0, o1, x->phi_face);
    x->phi_z0 = o0 <= o1 ? o0 : 0;
    x->phi_z1 = o0 <= o1 ? o1 : -1;
    st->vcycles = 0; st->mg_resid0 = 0.0; st->mg_bnorm = 0.0; st->mg_resid = 0.0;
    if (!check) return FX_OK;
    fxk::mg_resid_norm(x->s, x->lev[0].d, dbc_of(x, bv), 1, nullptr, x->red, x->red_out);
    double nr[2];
    int rc = read_red(x, 2, nr);
    if (rc) return rc;
    const double r0 = std::max({nr[1], std::fabs(blo) - nr[1], std::fabs(bhi) - nr[1]});
    st->mg_resid = nr[0]; st->mg_bnorm = nr[1]; st->mg_resid0 = r0;
    if (!(nr[0] == nr[0]) || !(nr[1] == nr[1])) return fail(x, FX_ENAN, "non-finite Poisson residual or right-hand side");
    if (nr[0] > tol * r0) return fail(x, FX_ENOCONV_MG, "direct Poisson solve: residual above the stopping rule's bound");
    return FX_OK;
}

/* Precision form of the direct solve: lateral transforms and tridiagonal
 * solves in FP32, FP64 iterative refinement.  The first solve takes the FP64
 * right-hand side (with the face terms) and stores phi in FP64; then, until
 * the FP64 residual meets the stopping rule (with ||r0|| bounded below as in
 * direct_solve), r = b - A phi is formed in FP64, stored in FP32, and
 * phi += A^-1 r in FP32.  st->vcycles counts the FP32 solves. */
int direct_solve_mixed(fx_ctx* x, const double bv[6], double tol, int maxit, fx_solve_stats* st, int z0, int z1) {
    const DBC bc = dbc_of(x, bv);
    const double blo = x->fft_w2lo * bv[2], bhi = x->fft_w2hi * bv[5];
    fxk::fft_solve(x->s, x->fft32, x->rhs, x->phi, blo, bhi, z0, z1);
    st->vcycles = 1; st->mg_resid0 = 0.0; st->mg_bnorm = 0.0; st->mg_resid = 0.0;
    for (;;) {
        fxk::resid_to32(x->s, x->lev[0].d, bc, x->lev32[0].b, x->red, x->red_out);
        double nr[2];
        int rc = read_red(x, 2, nr);
        if (rc) return rc;
        const double r0 = std::max({nr[1], std::fabs(blo) - nr[1], std::fabs(bhi) - nr[1]});
        st->mg_resid = nr[0]; st->mg_bnorm = nr[1]; st->mg_resid0 = r0;
        if (!(nr[0] == nr[0]) || !(nr[1] == nr[1])) return fail(x, FX_ENAN, "non-finite Poisson residual or right-hand side");
        if (nr[0] <= tol * r0) return FX_OK;
        if (st->vcycles >= maxit) return fail(x, FX_ENOCONV_MG, "FP32 direct solve with refinement did not converge");
        fxk::fft_solve_add(x->s, x->fft32, x->lev32[0].b, x->phi);
        ++st->vcycles;
    }
}

/* Preconditioned conjugate gradients (fx_pcg.hip) from phi = 0 on the
 * direct path's geometry with a Helmholtz term.  The preconditioner is the
 * direct solve of the operator with each plane's A-coefficient replaced by
 * its plane mean (z-pivots tabulated for this operator).  The stopping rule
 * is the multigrid's, ||r|| <= tol max(||b||, ||r0||) in the max-norm, r0
 * the residual of phi = 0 with the Dirichlet face terms; it is tested on the
 * updated residual every iteration (the host reads max|r| while the next
 * preconditioner solve runs) and, once that meets it, on the residual
 * b - A phi formed anew.  st->vcycles counts the iterations. */
int pcg_solve(fx_ctx* x, const double bv[6], double tol, int maxit, fx_solve_stats* st) {
    const size_t N = x->N;
    const int n2 = x->n[2];
    const size_t npiv = fxk::fft_pivot_count(x->n[0], x->n[1], n2);
    const int PB = fxk::PCG_PARTS;
    int rc;
    if (N >= (size_t)1 << 31) return fail(x, FX_EINVAL, "conjugate gradients index cells with 32-bit integers");
    const DBC bc = dbc_of(x, bv);
    const DLev& L = x->lev[0].d;
    if (!L.layered || bc.lo[0] != fxd::BC_PER || bc.lo[1] != fxd::BC_PER)
        return fail(x, FX_EINVAL, "conjugate gradients need a layered level periodic in x and y");
    const size_t nab = (size_t)n2 * fxk::PCG_ABAR_PARTS;
    const int nrz = fxk::fft_dot_parts(x->fft64);
    const size_t nrzb = (size_t)std::max(nrz, PB);
    if (!x->pcg_mem) {
        if ((rc = dalloc(x, (void**)&x->pcg_mem,
                         (4 * N + (size_t)n2 + npiv + 2 * (size_t)PB + nrzb + 4 + nab) * sizeof(double))))
            return rc;
        fxk::fill(x->s, x->pcg_mem, (long long)(4 * N), 0.0);   /* p and q are read (times 0) by the first pass */
    }
    double* r = x->pcg_mem;
    double* z = r + N;
    double* p = z + N;
    double* q = p + N;
    double* abar = q + N;
    double* piv = abar + n2;
    double* part_pq = piv + npiv;
    double* part_r = part_pq + PB;
    double* part_rz = part_r + PB;
    double* sc = part_rz + nrzb;   /* r.z of two iterations, the converged flag, the target */
    double* part_ab = sc + 4;
    /* the preconditioner: the plane-averaged operator's pivots */
    fxk::pcg_abar(x->s, L, part_ab, abar);
    fxk::fft_pivots(x->s, x->fft64, piv, nullptr, abar);
    FftPlan<double> M = x->fft64;
    M.piv = piv;
    M.pivu = nullptr;   /* the upward pivots are of the operator without the plane means */
    M.green = nullptr;
    FftPlan<double> Mg = M;   /* the preconditioner as the iterations use it: idle once converged */
    Mg.skip = sc + 2;
    /* r0 = b - A 0 with the face terms; z0, p0 */
    fxk::fill(x->s, sc + 2, 1, 0.0);
    fxk::fill(x->s, x->phi, (long long)N, 0.0);
    fxk::mg_resid_norm(x->s, L, bc, 1, r, x->red, x->red_out);
    fxk::fft_solve_dot(x->s, M, r, z, part_rz);
    fxk::pcg_dirmat(x->s, L, bc, z, p, q, part_rz, nrz, sc, -1, part_pq);
    double nr[2];
    if ((rc = read_red(x, 2, nr))) return rc;
    const double r0 = nr[0], bn = nr[1];
    const double target = tol * std::max(bn, r0);
    st->mg_resid0 = r0; st->mg_bnorm = bn; st->mg_resid = r0; st->vcycles = 0;
    if (!(r0 == r0) || !(bn == bn)) return fail(x, FX_ENAN, "non-finite Poisson residual or right-hand side");
    if (r0 <= target) return FX_OK;
    fxk::fill(x->s, sc + 3, 1, target);
    /* PCG_SLOTS - 1 iterations in flight: the host reads max|r| of iteration
     * it while later ones run; the one that meets the rule sets the device
     * flag, after which every queued pass does nothing, so phi is that
     * iteration's, as if the host had stopped there */
    std::vector<unsigned> vseq(maxit > 0 ? maxit : 1);
    int queued = 0;
    auto queue_iter = [&](int it) {
        fxk::pcg_update(x->s, x->phi, r, p, q, (long long)N, part_pq, sc, it, part_r);
        vseq[it] = ++x->pub_seq;
        double* slot = pcg_slot_dev(x, it % PCG_SLOTS);
        fxk::pcg_publish_max(x->s, part_r, slot, reinterpret_cast<unsigned*>(slot + 1), vseq[it], sc);
        fxk::fft_solve_dot(x->s, Mg, r, z, part_rz);
        fxk::pcg_dirmat(x->s, L, bc, z, p, q, part_rz, nrz, sc, it, part_pq);
    };
    auto fill_window = [&](int from) {
        queued = std::max(queued, from);
        while (queued < maxit && queued < from + PCG_SLOTS - 1) queue_iter(queued++);
    };
    fill_window(0);
    for (int it = 0; it < maxit; ++it) {
        double rm;
        if ((rc = wait_slot(x, it % PCG_SLOTS, vseq[it], &rm))) return rc;
        st->vcycles = it + 1;
        st->mg_resid = rm;
        if (!(rm == rm)) return fail(x, FX_ENAN, "conjugate gradients produced a non-finite residual");
        if (rm <= target) {
            /* the rule on b - A phi formed anew; if it fails (rounding drift of
             * the updated residual), clear the flag and resume from here */
            fxk::mg_resid_norm(x->s, L, bc, 1, nullptr, x->red, x->red_out);
            if ((rc = read_red(x, 2, nr))) return rc;
            st->mg_resid = nr[0];
            if (nr[0] <= target) return FX_OK;
            fxk::fill(x->s, sc + 2, 1, 0.0);
            fxk::fill(x->s, sc + 3, 1, 0.0);    /* no further early stop: the host decides */
            fxk::fft_solve_dot(x->s, Mg, r, z, part_rz);
            fxk::pcg_dirmat(x->s, L, bc, z, p, q, part_rz, nrz, sc, it, part_pq);
            queued = it + 1;
        }
        fill_window(it + 1);
    }
    return fail(x, FX_ENOCONV_MG, "preconditioned conjugate gradients did not converge in mg_max_iter iterations");
}

/* phi_zero: the solve starts from phi = 0, whatever phi holds.  Planes of
 * the right-hand side outside [z0, z1] hold zeros (z1 < z0: not known).
 * The FP64 direct solve may form phi on planes [o0, o1] only (o1 < o0:
 * all); every other solve forms all of it. */
int solve_any(fx_ctx* x, const double bv[6], double tol, int maxit, fx_solve_stats* st, bool phi_zero = false,
              bool check = false, int z0 = 0, int z1 = -1, int o0 = 0, int o1 = -1) {
    if (x->direct && x->lev[0].d.read_a == 0) {
        if (x->in.precision == FX_PREC_FP32_REFINE && x->fft_mem32) {
            phi_sync(x);
            return direct_solve_mixed(x, bv, tol, maxit, st, z0, z1);
        }
        return direct_solve(x, bv, tol, st, check || x->in.poisson_direct_check, z0, z1, o0, o1);
    }
    phi_sync(x);
    if (x->direct && x->in.poisson_pcg && x->in.precision == FX_PREC_FP64) return pcg_solve(x, bv, tol, maxit, st);
    if (x->in.precision == FX_PREC_FP32_REFINE) {
        /* FP32 copy of the A-coefficient for the correction levels */
        if (x->lev32[0].read_a) {
            fxk::d2f(x->s, x->lev32[0].a, x->alpha_cc, (long long)x->N);
            for (size_t l = 0; l + 1 < x->lev32.size(); ++l)
                fxk::mg_restrict_cells(x->s, x->lev32[l], x->lev32[l + 1], x->lev32[l].a, x->lev32[l + 1].a);
            x->lu_ok32 = false;
        }
        return mg_solve_mixed(x, bv, tol, maxit, st, phi_zero);
    }
    return mg_solve(x, bv, tol, maxit, st, phi_zero);
}

void restrict_acoef(fx_ctx* x) {
    x->lu_ok64 = false;
    for (size_t l = 0; l + 1 < x->lev.size(); ++l)
        fxk::mg_restrict_cells(x->s, x->lev[l].d, x->lev[l + 1].d, x->lev[l].d.a, x->lev[l + 1].d.a);
}

RhsArgs rhs_args(fx_ctx* x, double* const* P) {
    RhsArgs a;
    a.g = geom_of(x); a.M = mview_of(x); a.pbc = pbc_of(x); a.rot = rot_of(x);
    for (int d = 0; d < 3; ++d) a.P[d] = P[d];
    a.phi = x->phi; a.rho = x->rho; a.rhs = x->rhs; a.alpha = x->alpha_cc;
    a.delta = x->in.delta; a.pxy_zero = x->pxy_zero; a.sc = x->sc;
    a.kz0 = 0; a.kz1 = x->n[2] - 1;
    return a;
}

fxk::FaceSet face_set(const fx_ctx* x) {
    fxk::FaceSet f;
    f.cells = x->face2_list; f.ncells = x->nface2;
    f.planes = x->face2_list ? x->face2_list + x->nface2 : nullptr; f.nplanes = x->nface2_planes;
    return f;
}

/* ComputePhi_Rho (ElectrostaticSolver.cpp:678-747).  partial_ok: only the
 * forward-Euler update reads phi before the next solve, so a direct solve
 * may form it on the FE planes and their neighbours only (phi_sync
 * finishes it for every other reader). */
int compute_phi_rho(fx_ctx* x, double* const* P, fx_solve_stats* st, bool partial_ok = false) {
    double err = 1.0;
    int iter = 0;
    double bv[6];
    bc_values(x, bv);
    std::memcpy(x->bc_solve, bv, sizeof(bv));
    st->outer_iters = 0; st->vcycles = 0; st->outer_err = 0.0;
    while (err > OUTER_TOL) {
        RhsArgs ra0 = rhs_args(x, P);
        if (!x->has_sc && x->alpha_clean) ra0.alpha = nullptr;   /* alpha is 0 outside SC cells: written once */
        int z0 = 0, z1 = -1;   /* planes of the RHS outside [z0, z1] hold zeros (z1 < z0: not known) */
        /* the RHS of this P was formed by the last TDGL update (no SC cells) */
        if (x->rhs_next && !x->has_sc && x->alpha_clean && P == x->P) {
            std::swap(x->rhs, x->rhs_nx);
            std::swap(x->rhs_clean, x->rhs_nx_clean);
            x->lev[0].d.b = x->rhs;
            if (x->fe_k0 >= 0) { z0 = x->fe_k0; z1 = x->fe_k1; }
        } else {
            ra0.rhs = x->rhs;
            /* without SC cells the RHS is 0 outside FE cells: once the buffer
             * holds those zeros (and alpha is not written), only the planes
             * with FE cells are formed */
            const bool narrow = !x->has_sc && x->rhs_clean && !x->rhs_exposed && !ra0.alpha;
            if (narrow) {
                ra0.kz0 = x->fe_k0 >= 0 ? x->fe_k0 : 1;
                ra0.kz1 = x->fe_k0 >= 0 ? x->fe_k1 : 0;
                if (x->fe_k0 >= 0) { z0 = x->fe_k0; z1 = x->fe_k1; }
            }
            if (narrow && x->pxy_zero && x->in.tdgl_stream && x->fe_k0 >= 0)
                fxk::rhs_stream(x->s, x->s2, x->ev0, x->ev1, ra0, x->fe_k0, x->fe_k1, face_set(x), x->in.tdgl_stream);
            else
                fxk::rhs_newton(x->s, ra0);
            x->rhs_clean = !x->has_sc;
        }
        x->rhs_next = false;
        x->alpha_clean = !x->has_sc && !x->alpha_exposed;
        set_read_a(x, x->has_sc ? 1 : 0);
        /* the coarse levels' A-coefficient, for the multigrid (the CG path
         * forms its preconditioner from the finest level's plane means) */
        if (x->has_sc && !(x->direct && x->in.poisson_pcg && x->in.precision == FX_PREC_FP64)) restrict_acoef(x);
        fx_solve_stats one{};
        int o0 = 0, o1 = -1;
        if (partial_ok && !x->has_sc && x->fe_k0 >= 0) {
            o0 = std::max(x->fe_k0 - 1, 0);
            o1 = std::min(x->fe_k1 + 1, x->n[2] - 1);
        }
        int rc = solve_any(x, bv, x->in.mg_tol_rel, x->in.mg_max_iter, &one, !x->in.mg_warm_start, false, z0, z1, o0, o1);
        st->vcycles += one.vcycles;
        st->mg_resid0 = one.mg_resid0; st->mg_resid = one.mg_resid; st->mg_bnorm = one.mg_bnorm;
        if (rc) { st->outer_iters = iter + 1; return rc; }
        RhoArgs ra;
        for (int d = 0; d < 3; ++d) ra.n[d] = x->n[d];
        ra.M = mview_of(x); ra.phi = x->phi; ra.rho = x->rho; ra.e_den = x->e_den; ra.h_den = x->h_den;
        ra.phi_prev = x->phi_prev; ra.part = x->red; ra.want_err = x->has_sc; ra.sc = x->sc;
        int nbr = 0;
        if (x->has_sc || !x->rho_clean) nbr = fxk::rho_update(x->s, ra);   /* rho = 0 everywhere without SC cells */
        x->rho_clean = !x->has_sc && !x->rho_exposed;
        if (!x->has_sc) {
            err = 0.;
            st->outer_iters = 1;
        } else {
            /* the error of the first iteration is never used (it is formed
             * from the second on), so the host does not wait for it */
            if (iter > 0) {
                fxk::final1024(x->s, x->red, nbr, 2, 0, x->red_out);
                double nd[2];
                if ((rc = read_red(x, 2, nd))) return rc;
                err = nd[0] / nd[1];
            }
            iter = iter + 1;
            st->outer_iters = iter;
            st->outer_err = err;
            if (x->in.verbose) std::printf("%d iterations :: err = %g\n", iter, err);
            /* a NaN error ends the loop, as while (err > tol) does in the reference */
            if (iter >= x->in.outer_max_iter && err > OUTER_TOL)
                return fail(x, FX_ENOCONV_OUTER, "semiconductor outer loop exceeded outer_max_iter");
        }
    }
    return FX_OK;
}

TdglArgs tdgl_args(fx_ctx* x, double* const* P) {
    TdglArgs a;
    const fx_inputs& in = x->in;
    a.g = geom_of(x); a.M = mview_of(x); a.pbc = pbc_of(x); a.phv = phiview_of(x); a.rot = rot_of(x);
    for (int d = 0; d < 3; ++d) { a.P[d] = P[d]; a.Pout[d] = nullptr; a.GLout[d] = nullptr; }
    a.alpha = in.alpha; a.beta = in.beta; a.gamma = in.gamma;
    a.a12 = in.alpha_12; a.a112 = in.alpha_112; a.a123 = in.alpha_123;
    a.BigGamma = in.BigGamma; a.g11 = in.g11; a.g44 = in.g44; a.g12 = in.g12; a.g44_p = in.g44_p; a.dt = in.dt;
    a.incL = in.include_Landau; a.incG = in.include_Grad; a.incE = in.include_Elec;
    a.scalar = in.is_polarization_scalar == 1;
    a.kz0 = 0; a.kz1 = x->n[2] - 1;
    return a;
}

bool scalar_fast(const fx_ctx* x) { return x->in.is_polarization_scalar == 1 && x->pxy_zero; }

/* One update of P (main.cpp:279-348). */
int update(fx_ctx* x, int order, fx_solve_stats* st2) {
    const bool fast = scalar_fast(x);
    TdglArgs a = tdgl_args(x, x->P);
    x->rhs_next = false;
    if (order == 1) {
        for (int d = 0; d < 3; ++d) a.Pout[d] = (fast && d < 2) ? nullptr : x->Pn[d];
        /* scalar P changes only in FE cells (Gamma = 0 elsewhere) */
        if (fast) {
            a.kz0 = x->fe_k0 >= 0 ? x->fe_k0 : 1;
            a.kz1 = x->fe_k0 >= 0 ? x->fe_k1 : 0;
        }
        if (fast && !x->has_sc && x->in.tdgl_fused_rhs && !x->rhs_exposed && x->fe_k0 >= 0 && x->n[0] >= 4 &&
            x->n[1] >= 4) {
            if (!x->rhs_nx_clean) {
                fxk::fill(x->s, x->rhs_nx, (long long)x->N, 0.0);
                x->rhs_nx_clean = true;
            }
            TdglRhsArgs f;
            f.t = a;
            f.rhs = x->rhs_nx;
            f.k0 = x->fe_k0;
            f.k1 = x->fe_k1;
            f.face = x->face_list;
            f.nface = x->nface;
            fxk::tdgl_rhs_stream(x->s, f);
            x->rhs_next = true;
        } else if (fast && x->in.tdgl_stream && x->fe_k0 >= 0) {
            TdglRhsArgs f;
            f.t = a;
            f.rhs = nullptr;
            f.k0 = x->fe_k0;
            f.k1 = x->fe_k1;
            f.face = x->face_list;
            f.nface = x->nface;
            fxk::tdgl_stream(x->s, x->s2, x->ev0, x->ev1, f, face_set(x), x->in.tdgl_stream);
        } else {
            fxk::tdgl(x->s, a, fast);
        }
        if (fast) std::swap(x->P[2], x->Pn[2]);
        else for (int d = 0; d < 3; ++d) std::swap(x->P[d], x->Pn[d]);
        return FX_OK;
    }
    /* Heun: f0 at P^n, P* = P^n + dt f0, solve at P*, f1 at P*, P^{n+1} = P^n + dt(f0+f1)/2 */
    for (int d = 0; d < 3; ++d) { a.GLout[d] = x->GL[d]; a.Pout[d] = x->Pn[d]; }
    if (fast) { a.Pout[0] = a.Pout[1] = nullptr; a.GLout[0] = a.GLout[1] = nullptr; }
    fxk::tdgl(x->s, a, fast);
    int rc = compute_phi_rho(x, x->Pn, st2);
    if (rc) return rc;
    TdglArgs b = tdgl_args(x, x->Pn);
    for (int d = 0; d < 3; ++d) b.GLout[d] = x->GL1[d];
    if (fast) b.GLout[0] = b.GLout[1] = nullptr;
    fxk::tdgl(x->s, b, fast);
    HeunArgs h;
    for (int d = 0; d < 3; ++d) { h.P[d] = (fast && d < 2) ? nullptr : x->P[d]; h.G0[d] = x->GL[d]; h.G1[d] = x->GL1[d]; }
    h.dt = x->in.dt; h.N = (long long)x->N;
    fxk::heun(x->s, h);
    /* keep the second buffer equal to P in the cells the next Euler stage skips */
    for (int d = 0; d < 3; ++d) if (!(fast && d < 2)) fxk::copy(x->s, x->Pn[d], x->P[d], (long long)x->N);
    return FX_OK;
}

/* CheckSteadyState, ElectrostaticSolver.cpp:526-560.  phidiff divides by
 * max|phi| read from device memory.  With phi_tolerance <= 0 the value can
 * never stop or advance the run (val < tol is false for val >= 0 and for
 * NaN), so unless it is printed the host does not wait for it: fx_step reads
 * the last one when it ends (steady_read). */
int steady_read(fx_ctx* x, double* metric) {
    double m[2];
    int rc = read_red(x, 2, m, x->steady_dev);
    if (rc) return rc;
    x->steady_pending = false;
    *metric = (m[0] == 0.0) ? NAN : m[1] / m[0];
    return FX_OK;
}
int steady_check(fx_ctx* x, int step, double* metric) {
    phi_sync(x);
    fxk::maxabs2(x->s, x->phi, x->phi_old, (long long)x->N, x->red, x->steady_dev);
    fxk::phidiff(x->s, x->phi, x->phi_old, x->phidiff, (long long)x->N, x->steady_dev);
    x->steady_pending = true;
    if (!(x->in.phi_tolerance > 0.0) && !x->in.verbose) return FX_OK;
    double val;
    int rc = steady_read(x, &val);
    if (rc) return rc;
    if (step > 1 && val < x->in.phi_tolerance) { x->steady_state_step = step; x->inc_step = step; }
    if (x->in.verbose) std::printf("Steady state check : (phi(t) - phi(t-1)).norm0() = %g\n", val);
    *metric = val;
    return FX_OK;
}

void free_all(fx_ctx* x) {
    auto F = [](void* p) { if (p) hipFree(p); };
    for (int d = 0; d < 3; ++d) { F(x->P[d]); F(x->Pn[d]); F(x->E[d]); F(x->GL[d]); F(x->GL1[d]); }
    F(x->phi); F(x->phi_old); F(x->phi_prev); F(x->rhs); F(x->rhs_nx); F(x->face_list); F(x->face2_list); F(x->rho); F(x->e_den); F(x->h_den);
    F(x->alpha_cc); F(x->phidiff); F(x->eps); F(x->mat); F(x->rot_idx); F(x->rot_tab); F(x->rot_atab);
    F(x->diff.ck); F(x->diff.segP); F(x->diff.segPhi); F(x->diff.lam[0]); F(x->diff.lam[1]); F(x->diff.s);
    F(x->diff.fin); F(x->diff.acc); F(x->diff.m); F(x->diff.part);
    F(x->red); F(x->red_out); F(x->steady_dev); F(x->d_iters); F(x->cg_work); F(x->tmp);
    for (size_t l = 0; l < x->lev.size(); ++l) {
        DLev& d = x->lev[l].d;
        F(d.bx); F(d.by); F(d.bz); F(x->lev[l].up);
        if (l > 0) { F(d.phi); F(d.b); F(d.a); }
    }
    for (size_t l = 0; l < x->lev32.size(); ++l) {
        DLevF& f = x->lev32[l];
        F(f.phi); F(f.b); F(f.a); F(f.bx); F(f.by); F(f.bz); F((void*)f.bxk); F((void*)f.byk); F((void*)f.bzk); F(f.phi2);
    }
    for (float* p : x->up32) F(p);
    F(x->spec_basis64); F(x->spec_basis32); F(x->fft_mem64); F(x->fft_mem32); F(x->pcg_mem);
    F(x->cg_work32); F(x->phi_alt); F(x->d_tail64); F(x->d_tail32);
    F(x->lu64); F(x->lu32); F(x->perm64); F(x->perm32);
    if (x->h_red) hipHostFree(x->h_red);
    for (size_t l = 0; l < x->lev.size(); ++l) { F((void*)x->lev[l].d.bxk); F((void*)x->lev[l].d.byk); F((void*)x->lev[l].d.bzk); F(x->lev[l].d.phi2); }
    if (x->ev0) (void)hipEventDestroy(x->ev0);
    if (x->ev1) (void)hipEventDestroy(x->ev1);
    if (x->s2) (void)hipStreamDestroy(x->s2);
    if (x->s) (void)hipStreamDestroy(x->s);
}

double* field_dev(fx_ctx* x, int f) {
    switch (f) {
    case FX_FIELD_PX: return x->P[0];
    case FX_FIELD_PY: return x->P[1];
    case FX_FIELD_PZ: return x->P[2];
    case FX_FIELD_PHI: return x->phi;
    case FX_FIELD_POISSON_RHS: return x->rhs;
    case FX_FIELD_EX: return x->E[0];
    case FX_FIELD_EY: return x->E[1];
    case FX_FIELD_EZ: return x->E[2];
    case FX_FIELD_HOLES: return x->h_den;
    case FX_FIELD_ELECTRONS: return x->e_den;
    case FX_FIELD_CHARGE: return x->rho;
    case FX_FIELD_EPSILON: return x->eps;
    case FX_FIELD_PHIDIFF: return x->phidiff;
    case FX_FIELD_ALPHA_CC: return x->alpha_cc;
    case FX_FIELD_GL_RHS_X: return x->GL[0];
    case FX_FIELD_GL_RHS_Y: return x->GL[1];
    case FX_FIELD_GL_RHS_Z: return x->GL[2];
    default: return nullptr;
    }
}

}  // namespace

/* ---------------------------------- C ABI --------------------------------- */
/* ---- workspace of fx_poisson_solve_noalloc -------------------------------
 * The direct solve's spectrum and tridiagonal checkpoints live in this
 * file-scope device array instead of library-allocated memory, for callers
 * that cannot allocate.  In host code the name is only a stand-in: its device
 * address on each GPU comes from hipGetSymbolAddress, taken once per device
 * at context set-up (fx_create) with that device current, and only that
 * address is handed to kernels. */
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)64 << 20;
static __device__ __align__(256) double scratch[NOALLOC_SCRATCH_BYTES / sizeof(double)];

namespace {
constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

/* device address of `scratch` on the current device, looked up once */
int scratch_address(fx_ctx* x) {
    int dev = 0;
    HIPCHK(x, hipGetDevice(&dev));
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return fail(x, FX_EINVAL, "device index beyond the workspace table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return fail(x, FX_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = static_cast<double*>(p);
    }
    x->noalloc_scratch = g_scratch_ptr[dev];
    return FX_OK;
}
}  // namespace

extern "C" {

int fx_create(const fx_inputs* in_, fx_ctx** out) {
    if (!in_ || !out) return FX_EINVAL;
    *out = nullptr;
    fx_ctx* x = new fx_ctx();
    x->in = *in_;
    char ebuf[256] = {0};
    int rc = fx_inputs_finalize(&x->in, ebuf, sizeof(ebuf));
    if (rc) { std::fprintf(stderr, "fx_create: %s\n", ebuf); delete x; return rc; }
    const fx_inputs& in = x->in;
    for (int d = 0; d < 3; ++d) { x->n[d] = in.n_cell[d]; x->dx[d] = (in.prob_hi[d] - in.prob_lo[d]) / in.n_cell[d]; }
    x->N = (size_t)x->n[0] * x->n[1] * x->n[2];
    if (x->N > (size_t)INT32_MAX / 2) { delete x; return FX_EINVAL; }
    x->msx = x->n[0] + 2; x->msy = x->msx * (x->n[1] + 2);
    /* semiconductor constants, FerroX.cpp:534-554 */
    SCParams& s = x->sc;
    s.Nc = 2.8e25; s.Nv = 1.83e25; s.Eg = 1.12; s.Chi = 4.05; s.q = 1.602e-19; s.kb = 1.38e-23; s.T = 300;
    s.Ea = 44.0e-3; s.Ed = 46.0e-3; s.gA = 4.0; s.gD = 2.0;
    s.NA = in.acceptor_doping; s.ND = in.donor_doping; s.fermi_dirac = in.use_Fermi_Dirac;
    s.phi_ref = phi_ref_of(s);
    x->Phi_Bc_lo = in.Phi_Bc_lo; x->Phi_Bc_hi = in.Phi_Bc_hi; x->inc_step = in.inc_step;
    if (in.bias_ramp_steps > 0) {
        x->waveform.resize(in.bias_ramp_steps);
        for (int t = 1; t <= in.bias_ramp_steps; ++t)
            x->waveform[t - 1] = in.bias_ramp_start + (in.bias_ramp_end - in.bias_ramp_start) * ((double)t / in.bias_ramp_steps);
    }
    if (hipStreamCreate(&x->s) != hipSuccess) { delete x; return FX_EHIP; }
    if (hipStreamCreate(&x->s2) != hipSuccess || hipEventCreateWithFlags(&x->ev0, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&x->ev1, hipEventDisableTiming) != hipSuccess) {
        free_all(x);
        delete x;
        return FX_EHIP;
    }
    if ((rc = scratch_address(x))) {
        std::fprintf(stderr, "fx_create: %s\n", x->err.c_str());
        free_all(x);
        delete x;
        return rc;
    }
    const size_t B = x->N * sizeof(double);
#define ALLOC(p, b) do { if ((rc = dalloc(x, (void**)&(p), (b)))) { free_all(x); std::fprintf(stderr, "fx_create: %s\n", x->err.c_str()); delete x; return rc; } } while (0)
    for (int d = 0; d < 3; ++d) { ALLOC(x->P[d], B); ALLOC(x->Pn[d], B); ALLOC(x->E[d], B); ALLOC(x->GL[d], B); ALLOC(x->GL1[d], B); }
    ALLOC(x->phi, B); ALLOC(x->phi_old, B); ALLOC(x->phi_prev, B); ALLOC(x->rhs, B); ALLOC(x->rhs_nx, B); ALLOC(x->rho, B);
    ALLOC(x->e_den, B); ALLOC(x->h_den, B); ALLOC(x->alpha_cc, B); ALLOC(x->phidiff, B); ALLOC(x->eps, B);
    ALLOC(x->tmp, 2 * B);
    ALLOC(x->mat, (size_t)x->msy * (x->n[2] + 2));
    x->nred = std::max(RED_NB, ((x->n[0] + 63) / 64) * ((x->n[1] + 3) / 4) * x->n[2]);
    ALLOC(x->red, sizeof(double) * 4 * (size_t)x->nred);
    ALLOC(x->red_out, sizeof(double) * 8);
    ALLOC(x->steady_dev, sizeof(double) * 2);
    ALLOC(x->d_iters, sizeof(int) * 4);
#undef ALLOC
    if (hipHostMalloc((void**)&x->h_red, 64 * sizeof(double), hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess ||
        hipHostGetDevicePointer((void**)&x->d_red_host, x->h_red, 0) != hipSuccess) {
        free_all(x); delete x; return FX_EHIP;
    }
    for (int q = 8; q < 64; ++q) x->h_red[q] = 0.0;   /* every sequence number 0 */
    if ((rc = build_levels(x))) { std::fprintf(stderr, "fx_create: %s\n", x->err.c_str()); free_all(x); delete x; return rc; }
    x->h_tphase.assign(x->N, 0.0);
    for (int d = 0; d < 3; ++d) x->h_ang[d].assign(x->N, 0.0);
    x->h_lalpha.assign(x->N, in.alpha);
    if (in.Coordinate_Transformation == 1) {
        /* Initialize_tphase_Mask / Initialize_Euler_angles (Initialization.cpp:253-383):
         * the functions of the cell centre; the reference dereferences an
         * absent parser, so a missing function is an error here */
        const char* fns[4] = {in.tphase_geom_function, in.angle_alpha_function, in.angle_beta_function, in.angle_theta_function};
        const char* keys[4] = {"tphase_geom.tphase_geom_function(x,y,z)", "angle_alpha.alpha_function(x,y,z)",
                               "angle_beta.beta_function(x,y,z)", "angle_theta.theta_function(x,y,z)"};
        std::vector<double>* dst[4] = {&x->h_tphase, &x->h_ang[0], &x->h_ang[1], &x->h_ang[2]};
        for (int q = 0; q < 4; ++q) {
            char eb[256] = {0};
            if (!fns[q][0]) rc = fail(x, FX_EINVAL, std::string("Coordinate_Transformation = 1 needs ") + keys[q]);
            else if ((rc = fx_eval_cell_function(&in, fns[q], dst[q]->data(), eb, sizeof eb))) rc = fail(x, rc, std::string(keys[q]) + ": " + eb);
            if (rc) { std::fprintf(stderr, "fx_create: %s\n", x->err.c_str()); free_all(x); delete x; return rc; }
        }
    }
    mask_from_boxes(x, x->h_mat);
    periodic_fill_u8(x, x->h_mat);
    for (int k = 0; k < x->n[2]; ++k)
        for (int j = 0; j < x->n[1]; ++j)
            for (int i = 0; i < x->n[0]; ++i)
                if (x->h_tphase[(size_t)i + (size_t)x->n[0] * (j + (size_t)x->n[1] * k)] == 1.0) x->h_mat[gidx(x, i, j, k)] |= 4;
    if ((rc = upload_mat(x)) || (rc = build_coefficients(x)) || (rc = refresh_rotation(x))) {
        std::fprintf(stderr, "fx_create: %s\n", x->err.c_str()); free_all(x); delete x; return rc;
    }
    /* InitializePandRho, Initialization.cpp:6-166 */
    std::vector<double> hp[3], hrho(x->N, 0.0), hn(x->N, 0.0), hh(x->N, 0.0);
    for (int d = 0; d < 3; ++d) hp[d].assign(x->N, 0.0);
    const uint64_t seed = (uint64_t)(int64_t)in.random_seed;
    const double ni = std::sqrt(s.Nc * s.Nv) * exp(-0.5 * s.q * s.Eg / (s.kb * s.T));
    for (int k = 0; k < x->n[2]; ++k)
        for (int j = 0; j < x->n[1]; ++j)
            for (int i = 0; i < x->n[0]; ++i) {
                const size_t c = (size_t)i + (size_t)x->n[0] * (j + (size_t)x->n[1] * k);
                const int m = x->h_mat[gidx(x, i, j, k)] & 3;
                const double X = in.prob_lo[0] + (i + 0.5) * x->dx[0];
                const double Y = in.prob_lo[1] + (j + 0.5) * x->dx[1];
                const double Z = in.prob_lo[2] + (k + 0.5) * x->dx[2];
                double p = 0, q = 0, r = 0;
                if (m == 0) {
                    if (in.prob_type == 1) {
                        const double u = uniform01(seed, (uint64_t)(i + k * x->n[2]));
                        p = (-1.0 + 2.0 * u) * in.Remnant_P[0];
                        q = (-1.0 + 2.0 * u) * in.Remnant_P[1];
                        r = (-1.0 + 2.0 * u) * in.Remnant_P[2];
                    } else if (in.prob_type == 2) {
                        p = (-1.0 + 2.0 * uniform01(seed, 3 * (uint64_t)c + 0)) * in.Remnant_P[0];
                        q = (-1.0 + 2.0 * uniform01(seed, 3 * (uint64_t)c + 1)) * in.Remnant_P[1];
                        r = (-1.0 + 2.0 * uniform01(seed, 3 * (uint64_t)c + 2)) * in.Remnant_P[2];
                    } else {
                        const double g = exp(-(X * X / (2.0 * 5.e-9 * 5.e-9) + Y * Y / (2.0 * 5.e-9 * 5.e-9)
                                               + (Z - 1.5 * in.DE_hi[2]) * (Z - 1.5 * in.DE_hi[2]) / (2.0 * 2.0e-9 * 2.0e-9)));
                        p = in.Remnant_P[0] * g; q = in.Remnant_P[1] * g; r = in.Remnant_P[2] * g;
                    }
                    if (x->h_tphase[c] == 1.0) r = 0.0;
                }
                if (in.is_polarization_scalar == 1) { p = 0.0; q = 0.0; }
                hp[0][c] = p; hp[1][c] = q; hp[2][c] = r;
                if (m >= 2) {
                    hh[c] = ni; hn[c] = ni;
                    hrho[c] = s.q * (ni - ni - in.acceptor_doping + in.donor_doping);
                }
            }
    x->pxy_zero = true;
    for (size_t c = 0; c < x->N && x->pxy_zero; ++c) x->pxy_zero = (hp[0][c] == 0.0 && hp[1][c] == 0.0);
    for (int d = 0; d < 3; ++d) {
        if (hipMemcpyAsync(x->P[d], hp[d].data(), B, hipMemcpyHostToDevice, x->s) != hipSuccess ||
            hipMemcpyAsync(x->Pn[d], hp[d].data(), B, hipMemcpyHostToDevice, x->s) != hipSuccess) { free_all(x); delete x; return FX_EHIP; }
    }
    if (hipMemcpyAsync(x->rho, hrho.data(), B, hipMemcpyHostToDevice, x->s) != hipSuccess ||
        hipMemcpyAsync(x->e_den, hn.data(), B, hipMemcpyHostToDevice, x->s) != hipSuccess ||
        hipMemcpyAsync(x->h_den, hh.data(), B, hipMemcpyHostToDevice, x->s) != hipSuccess ||
        hipStreamSynchronize(x->s) != hipSuccess) { free_all(x); delete x; return FX_EHIP; }
    *out = x;
    return FX_OK;
}

int fx_destroy(fx_ctx* x) {
    if (!x) return FX_EINVAL;
    hipStreamSynchronize(x->s);
    free_all(x);
    delete x;
    return FX_OK;
}

const char* fx_last_error(fx_ctx* x) { return x ? x->err.c_str() : "null context"; }

int fx_get_dims(fx_ctx* x, int dims[3]) {
    if (!x || !dims) return FX_EINVAL;
    for (int d = 0; d < 3; ++d) dims[d] = x->n[d];
    return FX_OK;
}

int fx_get_field(fx_ctx* x, int f, double* host) {
    if (!x || !host) return FX_EINVAL;
    const size_t N = x->N;
    if (f == FX_FIELD_MASK || f == FX_FIELD_TPHASE) {
        for (int k = 0; k < x->n[2]; ++k)
            for (int j = 0; j < x->n[1]; ++j)
                for (int i = 0; i < x->n[0]; ++i) {
                    const size_t c = (size_t)i + (size_t)x->n[0] * (j + (size_t)x->n[1] * k);
                    host[c] = (f == FX_FIELD_MASK) ? (double)(x->h_mat[gidx(x, i, j, k)] & 3) : x->h_tphase[c];
                }
        return FX_OK;
    }
    if (f == FX_FIELD_ANGLE_ALPHA || f == FX_FIELD_ANGLE_BETA || f == FX_FIELD_ANGLE_THETA) {
        std::memcpy(host, x->h_ang[f - FX_FIELD_ANGLE_ALPHA].data(), N * sizeof(double));
        return FX_OK;
    }
    if (f == FX_FIELD_LANDAU_ALPHA) {
        std::memcpy(host, x->h_lalpha.data(), N * sizeof(double));
        return FX_OK;
    }
    if (f == FX_FIELD_GAMMA) {
        for (int k = 0; k < x->n[2]; ++k)
            for (int j = 0; j < x->n[1]; ++j)
                for (int i = 0; i < x->n[0]; ++i)
                    host[(size_t)i + (size_t)x->n[0] * (j + (size_t)x->n[1] * k)] = ((x->h_mat[gidx(x, i, j, k)] & 3) == 0) ? x->in.BigGamma : 0.0;
        return FX_OK;
    }
    if (f == FX_FIELD_EX || f == FX_FIELD_EY || f == FX_FIELD_EZ) {
        int rc = fx_compute_E(x);
        if (rc) return rc;
    }
    if (f == FX_FIELD_PHI) phi_sync(x);
    double* d = field_dev(x, f);
    if (!d) return fail(x, FX_EINVAL, "unknown field id");
    HIPCHK(x, hipMemcpyAsync(host, d, N * sizeof(double), hipMemcpyDeviceToHost, x->s));
    HIPCHK(x, hipStreamSynchronize(x->s));
    return FX_OK;
}

int fx_set_field(fx_ctx* x, int f, const double* host) {
    if (!x || !host) return FX_EINVAL;
    x->rhs_next = false;
    const size_t N = x->N, B = N * sizeof(double);
    switch (f) {
    case FX_FIELD_PX: case FX_FIELD_PY: case FX_FIELD_PZ: {
        const int d = f - FX_FIELD_PX;
        HIPCHK(x, hipMemcpyAsync(x->P[d], host, B, hipMemcpyHostToDevice, x->s));
        HIPCHK(x, hipMemcpyAsync(x->Pn[d], host, B, hipMemcpyHostToDevice, x->s));
        HIPCHK(x, hipStreamSynchronize(x->s));
        if (d < 2) {
            bool z = true;
            for (size_t c = 0; c < N && z; ++c) z = host[c] == 0.0;
            if (!z) x->pxy_zero = false;
            else {
                /* re-check the other component */
                std::vector<double> o(N);
                HIPCHK(x, hipMemcpy(o.data(), x->P[1 - d], B, hipMemcpyDeviceToHost));
                bool z2 = true;
                for (size_t c = 0; c < N && z2; ++c) z2 = o[c] == 0.0;
                x->pxy_zero = z2;
            }
        }
        return FX_OK;
    }
    case FX_FIELD_PHI:
        x->phi_z0 = 0; x->phi_z1 = -1; x->phi_cont = false;   /* all of phi rewritten */
        HIPCHK(x, hipMemcpyAsync(x->phi, host, B, hipMemcpyHostToDevice, x->s));
        HIPCHK(x, hipStreamSynchronize(x->s));
        return FX_OK;
    case FX_FIELD_CHARGE:
        x->rho_clean = false;
        HIPCHK(x, hipMemcpyAsync(x->rho, host, B, hipMemcpyHostToDevice, x->s));
        HIPCHK(x, hipStreamSynchronize(x->s));
        return FX_OK;
    case FX_FIELD_TPHASE: {
        for (size_t c = 0; c < N; ++c) {
            if (host[c] != 0.0 && host[c] != 1.0) return fail(x, FX_EINVAL, "tphase values must be 0 or 1");
            x->h_tphase[c] = host[c];
        }
        for (int k = 0; k < x->n[2]; ++k)
            for (int j = 0; j < x->n[1]; ++j)
                for (int i = 0; i < x->n[0]; ++i) {
                    uint8_t& m = x->h_mat[gidx(x, i, j, k)];
                    m = (uint8_t)((m & 3) | (x->h_tphase[(size_t)i + (size_t)x->n[0] * (j + (size_t)x->n[1] * k)] == 1.0 ? 4 : 0));
                }
        int rc;
        if ((rc = upload_mat(x)) || (rc = build_coefficients(x))) return rc;
        return FX_OK;
    }
    case FX_FIELD_MASK: {
        std::vector<uint8_t> m;
        mask_from_boxes(x, m);
        for (int k = 0; k < x->n[2]; ++k)
            for (int j = 0; j < x->n[1]; ++j)
                for (int i = 0; i < x->n[0]; ++i) {
                    const size_t c = (size_t)i + (size_t)x->n[0] * (j + (size_t)x->n[1] * k);
                    const double v = host[c];
                    if (v != 0.0 && v != 1.0 && v != 2.0 && v != 3.0) return fail(x, FX_EINVAL, "mask values must be 0, 1, 2 or 3");
                    m[gidx(x, i, j, k)] = (uint8_t)v;
                }
        periodic_fill_u8(x, m);
        for (int k = 0; k < x->n[2]; ++k)
            for (int j = 0; j < x->n[1]; ++j)
                for (int i = 0; i < x->n[0]; ++i)
                    m[gidx(x, i, j, k)] |= (x->h_tphase[(size_t)i + (size_t)x->n[0] * (j + (size_t)x->n[1] * k)] == 1.0 ? 4 : 0);
        x->h_mat = m;
        int rc;
        if ((rc = upload_mat(x)) || (rc = build_coefficients(x))) return rc;
        return FX_OK;
    }
    case FX_FIELD_ANGLE_ALPHA: case FX_FIELD_ANGLE_BETA: case FX_FIELD_ANGLE_THETA:
        std::memcpy(x->h_ang[f - FX_FIELD_ANGLE_ALPHA].data(), host, B);
        return refresh_rotation(x);
    case FX_FIELD_LANDAU_ALPHA:
        std::memcpy(x->h_lalpha.data(), host, B);
        return refresh_rotation(x);
    default:
        return fail(x, FX_EINVAL, "field is not settable");
    }
}

int fx_set_bias(fx_ctx* x, double lo, double hi) {
    if (!x) return FX_EINVAL;
    x->Phi_Bc_lo = lo; x->Phi_Bc_hi = hi;
    return FX_OK;
}

int fx_set_bias_waveform(fx_ctx* x, const double* v, int n) {
    if (!x || (n > 0 && !v)) return FX_EINVAL;
    x->waveform.assign(v, v + (n > 0 ? n : 0));
    return FX_OK;
}

int fx_compute_phi_rho(fx_ctx* x, fx_solve_stats* st) {
    if (!x) return FX_EINVAL;
    fx_solve_stats tmp{};
    int rc = compute_phi_rho(x, x->P, st ? st : &tmp);
    if (!rc) { hipError_t e = hipStreamSynchronize(x->s); if (e != hipSuccess) return fail(x, FX_EHIP, hipGetErrorString(e)); }
    return rc;
}

int fx_compute_E(fx_ctx* x) {
    if (!x) return FX_EINVAL;
    phi_sync(x);
    EArgs a;
    a.g = geom_of(x); a.phv = phiview_of(x); a.rot = rot_of(x);
    for (int d = 0; d < 3; ++d) a.E[d] = x->E[d];
    fxk::compute_E(x->s, a);
    HIPCHK(x, hipGetLastError());
    HIPCHK(x, hipStreamSynchronize(x->s));
    return FX_OK;
}

int fx_tdgl_rhs(fx_ctx* x) {
    if (!x) return FX_EINVAL;
    TdglArgs a = tdgl_args(x, x->P);
    for (int d = 0; d < 3; ++d) a.GLout[d] = x->GL[d];
    const bool fast = scalar_fast(x);
    if (fast) {
        fxk::fill(x->s, x->GL[0], (long long)x->N, 0.0);
        fxk::fill(x->s, x->GL[1], (long long)x->N, 0.0);
        a.GLout[0] = a.GLout[1] = nullptr;
    }
    fxk::tdgl(x->s, a, fast);
    HIPCHK(x, hipGetLastError());
    HIPCHK(x, hipStreamSynchronize(x->s));
    return FX_OK;
}

int fx_update(fx_ctx* x, int order) {
    if (!x || (order != 1 && order != 2)) return FX_EINVAL;
    fx_solve_stats st{};
    int rc = update(x, order, &st);
    if (rc) return rc;
    HIPCHK(x, hipGetLastError());
    HIPCHK(x, hipStreamSynchronize(x->s));
    return FX_OK;
}

int fx_step(fx_ctx* x, int nsteps, fx_step_stats* out) {
    if (!x) return FX_EINVAL;
    fx_step_stats st{};
    const fx_inputs& in = x->in;
    if (nsteps <= 0) nsteps = in.nsteps;
    auto t0 = std::chrono::steady_clock::now();
    const double tiny = 1.e-6;
    int rc = FX_OK;
    auto plot = [&](int step) -> int {
        char name[32];
        std::snprintf(name, sizeof name, "plt%08d", step);
        return fx_write_plotfile(x, (x->plot_dir + "/" + name).c_str(), x->time, step);
    };
    /* plotfile of the initial data (main.cpp:228-234) */
    if (in.plot_int > 0 && x->step == 0 && (rc = plot(0))) return fail(x, rc, "writing plt00000000 failed");
    /* With phi_tolerance <= 0 and no voltage sweep, CheckSteadyState
     * (ElectrostaticSolver.cpp:526-560) can neither stop nor advance the run:
     * its value is seen only after the last step of this call and through
     * PhiDiff at plot steps.  It is then formed only at those steps, with
     * phi_old = phi copied at the start of such a step when the step before
     * skipped it (the check itself copies phi_old = phi). */
    const bool lazy = !(in.phi_tolerance > 0.0) && !in.verbose && in.voltage_sweep == 0;
    for (int it = 0; it < nsteps; ++it) {
        const int step = ++x->step;
        const bool check_now = !lazy || it == nsteps - 1 || (in.plot_int > 0 && step % in.plot_int == 0);
        if (check_now && x->phi_old_stale) {
            phi_sync(x);
            fxk::copy(x->s, x->phi_old, x->phi, (long long)x->N);
            x->phi_old_stale = false;
        }
        if (!x->waveform.empty()) x->Phi_Bc_hi = x->waveform[std::min((size_t)step, x->waveform.size()) - 1];
        fx_solve_stats ss{};
        rc = compute_phi_rho(x, x->P, &ss, in.TimeIntegratorOrder == 1);
        st.total_vcycles += ss.vcycles; st.total_outer_iters += ss.outer_iters; st.total_solves += ss.outer_iters;
        if (rc) break;
        fx_solve_stats ss2{};
        rc = update(x, in.TimeIntegratorOrder, &ss2);
        st.total_vcycles += ss2.vcycles; st.total_outer_iters += ss2.outer_iters; st.total_solves += ss2.outer_iters;
        if (rc) break;
        if (check_now) {
            if ((rc = steady_check(x, step, &st.last_phi_change))) break;
        } else {
            x->phi_old_stale = true;
        }
        x->time = x->time + in.dt;
        st.steps_done++;
        if (in.verbose) std::printf("Advanced step %d\n", step);
        /* main.cpp:563-569 */
        if (in.plot_int > 0 && (step % in.plot_int == 0 || step == x->steady_state_step) && (rc = plot(step))) {
            rc = fail(x, rc, "writing a plotfile failed");
            break;
        }
        if (in.voltage_sweep == 1 && x->inc_step > 0 && step == x->inc_step) {
            x->Phi_Bc_hi += x->sign * in.Phi_Bc_inc;
            x->num_Vapp += 1;
            if (std::fabs(std::fabs(x->Phi_Bc_hi) - in.Phi_Bc_hi_max) <= tiny) x->sign *= -1;
            fx_solve_stats ss3{};
            rc = compute_phi_rho(x, x->P, &ss3);
            st.total_vcycles += ss3.vcycles; st.total_outer_iters += ss3.outer_iters; st.total_solves += ss3.outer_iters;
            if (rc) break;
        }
        if (in.voltage_sweep == 0 && step == x->steady_state_step) { st.stopped_steady = 1; break; }
        if (in.voltage_sweep == 1 && x->Phi_Bc_hi > 0. && x->Phi_Bc_hi - in.Phi_Bc_hi_max > tiny) break;
        if (in.voltage_sweep == 1 && x->Phi_Bc_hi < 0. && -x->Phi_Bc_hi - in.Phi_Bc_hi_max > tiny) break;
        if (in.voltage_sweep == 1 && x->num_Vapp == in.num_Vapp_max) break;
    }
    if (x->steady_pending) {
        const int rc2 = steady_read(x, &st.last_phi_change);
        if (!rc) rc = rc2;
    }
    hipError_t e = hipStreamSynchronize(x->s);
    if (!rc && e != hipSuccess) rc = fail(x, FX_EHIP, hipGetErrorString(e));
    st.time = x->time; st.Phi_Bc_hi = x->Phi_Bc_hi; st.num_Vapp = x->num_Vapp;
    st.wall_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (out) *out = st;
    return rc;
}

int fx_set_plot_dir(fx_ctx* x, const char* dir) {
    if (!x || !dir) return FX_EINVAL;
    x->plot_dir = dir;
    return FX_OK;
}

int fx_poisson_solve(fx_ctx* x, const double* rhs, const double* acoef, double* phi, fx_solve_stats* st) {
    if (!x || !rhs || !phi) return FX_EINVAL;
    x->rhs_next = false;
    const size_t B = x->N * sizeof(double);
    HIPCHK(x, hipMemcpyAsync(x->rhs, rhs, B, hipMemcpyHostToDevice, x->s));
    x->rhs_clean = false;
    x->alpha_clean = false;
    if (acoef) HIPCHK(x, hipMemcpyAsync(x->alpha_cc, acoef, B, hipMemcpyHostToDevice, x->s));
    else fxk::fill(x->s, x->alpha_cc, (long long)x->N, 0.0);
    set_read_a(x, acoef ? 2 : 0);
    restrict_acoef(x);
    HIPCHK(x, hipMemcpyAsync(x->phi, phi, B, hipMemcpyHostToDevice, x->s));
    double bv[6];
    bc_values(x, bv);
    std::memcpy(x->bc_solve, bv, sizeof(bv));
    fx_solve_stats tmp{};
    int rc = solve_any(x, bv, x->in.mg_tol_rel, x->in.mg_max_iter, st ? st : &tmp, false, true);
    HIPCHK(x, hipMemcpyAsync(phi, x->phi, B, hipMemcpyDeviceToHost, x->s));
    HIPCHK(x, hipStreamSynchronize(x->s));
    return rc;
}

int fx_poisson_solve_noalloc(fx_ctx* x, const double* rhs_dev, double* phi_dev) {
    if (!x || !rhs_dev || !phi_dev) return FX_EINVAL;
    if (!x->direct) return fail(x, FX_EINVAL, "fx_poisson_solve_noalloc: the context's operator has no direct solve "
                                             "(needs x, y periodic, z-layered permittivity, a Dirichlet z face, "
                                             "lateral extents powers of two)");
    const FftPlan<double>& P0 = x->fft64;
    const size_t need = x->N * sizeof(double);
    if (need > NOALLOC_SCRATCH_BYTES)
        return fail(x, FX_EINVAL, "fx_poisson_solve_noalloc: " + std::to_string(x->N) + " cells need " + std::to_string(need) +
                                      " bytes of workspace, more than the " + std::to_string(NOALLOC_SCRATCH_BYTES) +
                                      " bytes of the static array");
    if (!x->noalloc_scratch) return fail(x, FX_EHIP, "fx_poisson_solve_noalloc: workspace address unknown");
    FftPlan<double> P = P0;
    P.S = x->noalloc_scratch;
    double bv[6];
    bc_values(x, bv);
    fxk::fft_solve(x->s, P, rhs_dev, phi_dev, x->fft_w2lo * bv[2], x->fft_w2hi * bv[5]);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(x, FX_EHIP, hipGetErrorString(e));
    return FX_OK;
}

int fx_field_device_ptr(fx_ctx* x, int f, void** dptr) {
    if (!x || !dptr) return FX_EINVAL;
    double* d = field_dev(x, f);
    if (!d) return fail(x, FX_EINVAL, "field has no device array");
    if (d == x->rho) x->rho_exposed = true, x->rho_clean = false;
    if (d == x->rhs) x->rhs_exposed = true;   /* the buffers are not swapped from then on */
    if (d == x->alpha_cc) x->alpha_exposed = true, x->alpha_clean = false;
    if (d == x->phi) phi_sync(x), x->phi_exposed = true;   /* every later solve forms all of phi */
    *dptr = d;
    return FX_OK;
}

int fx_sync(fx_ctx* x) {
    if (!x) return FX_EINVAL;
    HIPCHK(x, hipStreamSynchronize(x->s));
    return FX_OK;
}

void* fx_stream(fx_ctx* x) { return x ? (void*)x->s : nullptr; }

/* ------------------------- differentiation form --------------------------- */
/* Discrete adjoint of the forward-Euler step with its Poisson solve; the
 * kernels and the derivation are in fx_adjoint.hip.  The forward stores P
 * every K steps; the backward re-runs each K-step segment from its checkpoint
 * (the same kernels in the same order, so the same values), keeping P_n and
 * phi_n of the segment, and reverses it. */
static int diff_check(fx_ctx* x) {
    const fx_inputs& in = x->in;
    if (in.TimeIntegratorOrder != 1) return fail(x, FX_EINVAL, "differentiation form needs TimeIntegratorOrder = 1");
    if (x->has_sc) return fail(x, FX_EINVAL, "differentiation form does not cover semiconductor regions");
    if (!scalar_fast(x)) return fail(x, FX_EINVAL, "differentiation form needs is_polarization_scalar = 1 with Px = Py = 0");
    if (x->rot_idx) return fail(x, FX_EINVAL, "differentiation form needs uniform angles and Landau alpha");
    for (int q = 0; q < 9; ++q)
        if (x->Ru[q] != ((q % 4) == 0 ? 1.0 : 0.0)) return fail(x, FX_EINVAL, "differentiation form needs zero Euler angles");
    if (in.include_Landau != 1 || in.include_Grad != 1 || in.include_Elec != 1)
        return fail(x, FX_EINVAL, "differentiation form needs include_Landau = include_Grad = include_Elec = 1");
    return FX_OK;
}

static int diff_reserve(fx_ctx* x, int nsteps, int K) {
    fx_ctx::Diff& D = x->diff;
    const size_t N = x->N, B = N * sizeof(double);
    const int nck = nsteps / K + 1;
    int rc;
    auto grow = [&](double*& p, size_t& cap, size_t need) -> int {
        if (need <= cap) return FX_OK;
        if (p) { (void)hipFree(p); p = nullptr; }
        cap = 0;
        int r = dalloc(x, (void**)&p, need);
        if (!r) cap = need;
        return r;
    };
    if ((rc = grow(D.ck, D.cap_ck, (size_t)nck * B))) return rc;
    if ((rc = grow(D.segP, D.cap_segP, (size_t)K * B))) return rc;
    if ((rc = grow(D.segPhi, D.cap_segPhi, (size_t)K * B))) return rc;
    if (!D.lam[0]) {
        if ((rc = dalloc(x, (void**)&D.lam[0], B)) || (rc = dalloc(x, (void**)&D.lam[1], B)) ||
            (rc = dalloc(x, (void**)&D.s, B)) || (rc = dalloc(x, (void**)&D.fin, B)) ||
            (rc = dalloc(x, (void**)&D.part, sizeof(double) * 7 * (size_t)fxk::adj_blocks(geom_of(x)))))
            return rc;
    }
    if ((rc = grow(D.acc, D.cap_acc, sizeof(double) * (size_t)(6 + nsteps)))) return rc;
    if ((rc = grow(D.m, D.cap_m, sizeof(double) * (size_t)nsteps))) return rc;
    D.nck = nck;
    int nfe = 0;
    for (int k = 0; k < x->n[2]; ++k)
        for (int j = 0; j < x->n[1]; ++j)
            for (int i = 0; i < x->n[0]; ++i) {
                const int r = x->h_mat[(i + 1) + x->msx * (j + 1) + x->msy * (k + 1)];
                if ((r & 3) == 0 && !(r & 4)) ++nfe;
            }
    D.nfe = nfe;
    if (nfe == 0) return fail(x, FX_EINVAL, "differentiation form: no ferroelectric cells");
    return FX_OK;
}

/* restore P (both buffers: the scalar step leaves non-FE cells in place) */
static void diff_set_P(fx_ctx* x, const double* src) {
    x->rhs_next = false;
    fxk::copy(x->s, x->P[2], src, (long long)x->N);
    fxk::copy(x->s, x->Pn[2], src, (long long)x->N);
}

static int diff_step(fx_ctx* x, double V) {
    x->Phi_Bc_hi = V;
    fx_solve_stats ss{}, ss2{};
    int rc = compute_phi_rho(x, x->P, &ss, true);
    if (rc) return rc;
    return update(x, 1, &ss2);
}

int fx_diff_set_params(fx_ctx* x, const double* th) {
    if (!x || !th) return FX_EINVAL;
    x->rhs_next = false;
    fx_inputs& in = x->in;
    const double a_old = in.alpha;
    in.alpha = th[0]; in.beta = th[1]; in.gamma = th[2]; in.g11 = th[3]; in.g44 = th[4];
    for (double& v : x->h_lalpha) if (v == a_old) v = in.alpha;
    x->diff.valid = false;
    if (th[5] != in.epsilonX_fe) {
        if (in.epsilonX_fe_tphase == in.epsilonX_fe) in.epsilonX_fe_tphase = th[5];
        in.epsilonX_fe = th[5];
        return build_coefficients(x);
    }
    return FX_OK;
}

int fx_diff_get_params(fx_ctx* x, double* th) {
    if (!x || !th) return FX_EINVAL;
    const fx_inputs& in = x->in;
    th[0] = in.alpha; th[1] = in.beta; th[2] = in.gamma; th[3] = in.g11; th[4] = in.g44; th[5] = in.epsilonX_fe;
    return FX_OK;
}

int fx_diff_forward(fx_ctx* x, const double* V, int nsteps, int K, double* m_out) {
    if (!x || !V || nsteps < 1 || K < 1) return FX_EINVAL;
    int rc;
    if ((rc = diff_check(x)) || (rc = diff_reserve(x, nsteps, K))) return rc;
    fx_ctx::Diff& D = x->diff;
    D.valid = false;
    D.nsteps = nsteps; D.K = K;
    D.V.assign(V, V + nsteps);
    const size_t N = x->N;
    const Geom g = geom_of(x);
    const MView M = mview_of(x);
    fxk::copy(x->s, D.ck, x->P[2], (long long)N);
    for (int n = 0; n < nsteps; ++n) {
        if ((rc = diff_step(x, V[n]))) return rc;
        fxk::fe_sum(x->s, x->P[2], M, g, D.part, D.m + n);
        if ((n + 1) % K == 0) fxk::copy(x->s, D.ck + (size_t)((n + 1) / K) * N, x->P[2], (long long)N);
    }
    fxk::copy(x->s, D.fin, x->P[2], (long long)N);
    std::vector<double> m(nsteps);
    HIPCHK(x, hipMemcpyAsync(m.data(), D.m, sizeof(double) * nsteps, hipMemcpyDeviceToHost, x->s));
    HIPCHK(x, hipStreamSynchronize(x->s));
    for (int n = 0; n < nsteps; ++n) m_out[n] = m[n] / D.nfe;
    x->step += nsteps;
    x->time += nsteps * x->in.dt;
    D.valid = true;
    return FX_OK;
}

int fx_diff_backward(fx_ctx* x, const double* dL_dm, double* g_theta, double* g_V) {
    if (x) x->rhs_next = false;
    if (!x || !dL_dm) return FX_EINVAL;
    fx_ctx::Diff& D = x->diff;
    if (!D.valid) return fail(x, FX_EINVAL, "fx_diff_backward needs a preceding fx_diff_forward with the same parameters");
    const int N_ = D.nsteps, K = D.K;
    const size_t N = x->N;
    const Geom g = geom_of(x);
    const MView M = mview_of(x);
    const fx_inputs& in = x->in;
    int rc;
    fxk::fill(x->s, D.acc, 6 + N_, 0.0);
    fxk::adj_init(x->s, D.lam[0], M, g, dL_dm[N_ - 1] / D.nfe);
    int cur = 0;
    const int nb = fxk::adj_blocks(g);
    for (int seg = (N_ - 1) / K; seg >= 0; --seg) {
        const int n0 = seg * K, n1 = std::min(n0 + K, N_);
        diff_set_P(x, D.ck + (size_t)seg * N);
        for (int n = n0; n < n1; ++n) {
            fxk::copy(x->s, D.segP + (size_t)(n - n0) * N, x->P[2], (long long)N);
            x->Phi_Bc_hi = D.V[n];
            fx_solve_stats ss{}, ss2{};
            if ((rc = compute_phi_rho(x, x->P, &ss))) return rc;
            fxk::copy(x->s, D.segPhi + (size_t)(n - n0) * N, x->phi, (long long)N);
            if ((rc = update(x, 1, &ss2))) return rc;
        }
        for (int n = n1 - 1; n >= n0; --n) {
            x->Phi_Bc_hi = D.V[n];
            double bv[6];
            bc_values(x, bv);
            fxk::AdjArgs a;
            a.g = g; a.M = M; a.pbc = pbc_of(x);
            a.lam = D.lam[cur]; a.lam_out = D.lam[1 - cur];
            a.P = D.segP + (size_t)(n - n0) * N;
            a.phi = D.segPhi + (size_t)(n - n0) * N;
            a.psi = x->phi;
            a.s = D.s;
            a.dtG = in.dt * in.BigGamma;
            a.alpha = in.alpha; a.beta = in.beta; a.gamma = in.gamma; a.g11 = in.g11; a.g44m = in.g44 - in.g44_p;
            a.loss_w = n >= 1 ? dL_dm[n - 1] / D.nfe : 0.0;
            a.eps0 = in.epsilon_0;
            for (int d = 0; d < 3; ++d) { a.bclo[d] = in.bc_lo[d]; a.bchi[d] = in.bc_hi[d]; a.ih2[d] = x->lev[0].d.ih2[d]; }
            std::memcpy(a.bcv, bv, sizeof(bv));
            a.eps = x->eps;
            a.ghost_face = in.phi_ghost_mode == FX_GHOST_FACE;
            a.part = D.part;
            fxk::adj_fzT(x->s, a);
            /* psi = A^-1 (Fz^T nu): the symmetric operator, homogeneous BCs */
            fxk::copy(x->s, x->rhs, D.s, (long long)N);
            x->rhs_clean = false;
            fxk::fill(x->s, x->phi, (long long)N, 0.0);
            set_read_a(x, 0);
            const double bz[6] = {0, 0, 0, 0, 0, 0};
            fx_solve_stats st{};
            if ((rc = solve_any(x, bz, in.mg_tol_rel, in.mg_max_iter, &st))) return rc;
            fxk::adj_step(x->s, a);
            fxk::adj_accum(x->s, D.part, nb, D.acc, D.acc + 6, n);
            cur = 1 - cur;
        }
    }
    diff_set_P(x, D.fin);
    std::vector<double> acc(6 + N_);
    HIPCHK(x, hipMemcpyAsync(acc.data(), D.acc, sizeof(double) * acc.size(), hipMemcpyDeviceToHost, x->s));
    HIPCHK(x, hipStreamSynchronize(x->s));
    if (g_theta) for (int q = 0; q < 6; ++q) g_theta[q] = acc[q];
    if (g_V) for (int n = 0; n < N_; ++n) g_V[n] = acc[6 + n];
    return FX_OK;
}

/* Average duration of one launch of a hot-path kernel, timed with HIP events
 * on the context's stream (the stream the kernel runs on).  which:
 *   "smooth0"  red-black Gauss-Seidel on the finest level: one colour
 *              (k_mg_smooth), or one full sweep when mg_fused_sweep is set
 *   "resid0"   residual + restriction from the finest level
 *   "tdgl"     fused TDGL right-hand side + forward-Euler update (state unchanged)
 *   "rhs"      Poisson right-hand side + Newton coefficient
 *   "fft_x", "fft_y", "fft_tri", "fft_xinv", "fft_solve"
 *              passes of the direct solve (forward x transform, y transform,
 *              tridiagonal solves, inverse x transform) and the whole solve
 *   "copy"     plain copy of N doubles (bandwidth calibration)
 * Kernels that write state write it into scratch or re-smooth phi; use after
 * the measured run. */
int fx_bench_kernel(fx_ctx* x, const char* which, int reps, double* avg_ms) {
    if (!x || !which || reps < 1 || !avg_ms) return FX_EINVAL;
    phi_sync(x);   /* the transform passes reuse the direct solve's workspace */
    const std::string w(which);
    double bv[6];
    bc_values(x, bv);
    const DBC bc = dbc_of(x, bv);
    auto launch = [&]() -> bool {
        if (w == "smooth0") {   /* fused: two sweeps (the pair returns phi to its buffer); reported per sweep */
            if (x->in.mg_fused_sweep && fxk::fused_ok(x->lev[0].d, bc)) fxk::mg_sweep(x->s, x->lev[0].d, bc, 1, 2, 1);
            else fxk::mg_smooth(x->s, x->lev[0].d, bc, 1, 0);
        }
        else if (w == "resid0") {
            if (x->lev.size() < 2) return false;
            fxk::mg_resid_restrict(x->s, x->lev[0].d, x->lev[1].d, bc, 1);
        } else if (w == "smooth0x2") {   /* two sweeps in one streaming pass */
            if (!(x->in.mg_fused_sweep >= 2 && fxk::fused_ok(x->lev[0].d, bc))) return false;
            fxk::mg_sweep(x->s, x->lev[0].d, bc, 1, 2, 2);
        } else if (w == "prolong0") {   /* prolongation from level 1 folded into one sweep of level 0 */
            if (x->lev.size() < 2) return false;
            fxk::mg_prolong_sweep(x->s, x->lev[0].d, x->lev[1].d, bc, 1, 1, x->in.mg_fused_sweep);
        } else if (w == "residmax0") {
            fxk::mg_resid_norm(x->s, x->lev[0].d, bc, 1, nullptr, x->red, x->red_out);
        } else if (w == "tdgl") {
            /* the launches update() makes for an Euler step, into the second P buffer */
            const bool fast = scalar_fast(x);
            TdglArgs a = tdgl_args(x, x->P);
            for (int d = 0; d < 3; ++d) a.Pout[d] = (fast && d < 2) ? nullptr : x->Pn[d];
            if (fast && x->fe_k0 >= 0) { a.kz0 = x->fe_k0; a.kz1 = x->fe_k1; }
            if (fast && x->in.tdgl_stream && x->fe_k0 >= 0) {
                TdglRhsArgs f;
                f.t = a; f.rhs = nullptr; f.k0 = x->fe_k0; f.k1 = x->fe_k1; f.face = x->face_list; f.nface = x->nface;
                fxk::tdgl_stream(x->s, x->s2, x->ev0, x->ev1, f, face_set(x), x->in.tdgl_stream);
            } else {
                fxk::tdgl(x->s, a, fast);
            }
        } else if (w == "rhs") {
            /* the launches compute_phi_rho() makes, into scratch */
            RhsArgs a = rhs_args(x, x->P);
            const bool narrow = !x->has_sc && x->fe_k0 >= 0;
            a.rhs = x->tmp; a.alpha = narrow ? nullptr : x->tmp + x->N;
            if (narrow) { a.kz0 = x->fe_k0; a.kz1 = x->fe_k1; }
            if (narrow && x->pxy_zero && x->in.tdgl_stream)
                fxk::rhs_stream(x->s, x->s2, x->ev0, x->ev1, a, x->fe_k0, x->fe_k1, face_set(x), x->in.tdgl_stream);
            else
                fxk::rhs_newton(x->s, a);
        } else if (w.compare(0, 4, "fft_") == 0) {
            /* passes of the direct solve; the inverse x transform and the
             * whole solve write into scratch, the others into the solve's
             * own workspace */
            if (!x->direct) return false;
            const double blo = x->fft_w2lo * bv[2], bhi = x->fft_w2hi * bv[5];
            if (w == "fft_x") fxk::fft_pass(x->s, x->fft64, 0, x->rhs, nullptr, blo, bhi);
            else if (w == "fft_y") fxk::fft_pass(x->s, x->fft64, 1, nullptr, nullptr, 0.0, 0.0);
            else if (w == "fft_tri") fxk::fft_pass(x->s, x->fft64, 2, nullptr, nullptr, 0.0, 0.0);
            else if (w == "fft_xinv") fxk::fft_pass(x->s, x->fft64, 3, nullptr, x->tmp, 0.0, 0.0);
            else if (w == "fft_solve") fxk::fft_solve(x->s, x->fft64, x->rhs, x->tmp, blo, bhi);
            else return false;
        } else if (w == "copy") fxk::copy(x->s, x->tmp, x->phi, (long long)x->N);
        else return false;
        return true;
    };
    if (!launch()) return fail(x, FX_EINVAL, "unknown kernel name");
    hipEvent_t e0, e1;
    HIPCHK(x, hipEventCreate(&e0));
    HIPCHK(x, hipEventCreate(&e1));
    HIPCHK(x, hipEventRecord(e0, x->s));
    for (int r = 0; r < reps; ++r) launch();
    HIPCHK(x, hipEventRecord(e1, x->s));
    HIPCHK(x, hipEventSynchronize(e1));
    float ms = 0.f;
    HIPCHK(x, hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    HIPCHK(x, hipGetLastError());
    *avg_ms = (double)ms / reps;
    if (w == "smooth0" && x->in.mg_fused_sweep && fxk::fused_ok(x->lev[0].d, bc)) *avg_ms *= 0.5;
    return FX_OK;
}

int fx_num_levels(fx_ctx* x) { return x ? (int)x->lev.size() : 0; }

int fx_level_dims(fx_ctx* x, int l, int dims[3]) {
    if (!x || l < 0 || l >= (int)x->lev.size()) return FX_EINVAL;
    dims[0] = x->lev[l].d.n0; dims[1] = x->lev[l].d.n1; dims[2] = x->lev[l].d.n2;
    return FX_OK;
}

}  // extern "C"

/* Library-internal (fx_batch.hip): the material id on the grown box, its
 * strides, the inputs and the per-plane Poisson coefficients of level 0
 * (layered levels only). */
int fxi_export(fx_ctx* x, std::vector<uint8_t>& mat, int& msx, int& msy, fx_inputs& in, double ih2[3],
               std::vector<double>& epsk, std::vector<double>& bzk, double bv[6]) {
    if (!x) return FX_EINVAL;
    mat = x->h_mat; msx = x->msx; msy = x->msy; in = x->in;
    mark_interior_fe(x, mat);   /* the device copy's bit 3, as upload_mat */
    const DLev& d = x->lev[0].d;
    for (int q = 0; q < 3; ++q) ih2[q] = d.ih2[q];
    if (!d.layered) return fail(x, FX_EINVAL, "batched devices must be laterally uniform (layered) stacks");
    const int nz = x->n[2];
    epsk.resize(nz); bzk.resize(nz + 1);
    HIPCHK(x, hipMemcpy(epsk.data(), d.bxk, nz * sizeof(double), hipMemcpyDeviceToHost));
    HIPCHK(x, hipMemcpy(bzk.data(), d.bzk, (nz + 1) * sizeof(double), hipMemcpyDeviceToHost));
    bc_values(x, bv);
    return FX_OK;
}

const fx_inputs* fxi_inputs(fx_ctx* x) { return x ? &x->in : nullptr; }
