/*
 * Kernel argument blocks and host launchers (fx_kernels.hip).
 */
#pragma once
#include <hip/hip_runtime.h>
#include "fx_device.h"

/* Semiconductor constants (FerroX.cpp:534-554) and doping. */
struct SCParams {
    double q, kb, T, Nc, Nv, Eg, Chi, phi_ref, Ea, Ed, gA, gD, NA, ND;
    int fermi_dirac;
};

/* Rotation and Landau alpha: uniform, or a table of (R, alpha) rows indexed
 * per cell. */
struct RotArgs {
    const int* idx;        /* null: uniform */
    const double* tab;     /* 9 doubles of R per row */
    const double* atab;    /* alpha per row */
    double Ru[9];
    int nrows;             /* rows of tab and atab */
};

struct RhsArgs {
    fxd::Geom g;
    fxd::MView M;
    fxd::PBC pbc;
    RotArgs rot;
    const double* P[3];
    const double* phi;
    const double* rho;
    double* rhs;
    double* alpha;
    double delta;
    int pxy_zero;
    SCParams sc;
    int kz0, kz1;   /* planes formed (the caller narrows them to the FE planes when the rest is known to be 0) */
};

struct RhoArgs {
    int n[3];
    fxd::MView M;
    const double* phi;
    double* rho;
    double* e_den;
    double* h_den;
    double* phi_prev;
    double* part;
    int want_err;
    SCParams sc;
};

struct TdglArgs {
    fxd::Geom g;
    fxd::MView M;
    fxd::PBC pbc;
    fxd::PhiView phv;
    RotArgs rot;
    const double* P[3];
    double* Pout[3];
    double* GLout[3];
    double alpha, beta, gamma, a12, a112, a123, BigGamma, g11, g44, g12, g44_p, dt;
    int incL, incG, incE, scalar;
    int kz0, kz1;   /* planes updated (P changes only where Gamma != 0, i.e. in FE cells) */
};

/* TDGL update fused with the next step's Poisson RHS (scalar P, forward
 * Euler, no semiconductor cells): t's Pout[2] receives P_new, rhs the RHS of
 * P_new in planes k0..k1 (the planes holding FE cells); face lists the FE
 * cells next to a material or domain face (global cell indices) */
struct TdglRhsArgs {
    TdglArgs t;
    double* rhs;
    int k0, k1;
    const int* face;   /* FE cells whose stencils are not all centred (no MAT_INTERIOR_FE) */
    int nface;
};

struct HeunArgs {
    double* P[3];
    const double* G0[3];
    const double* G1[3];
    double dt;
    long long N;
};

struct EArgs {
    fxd::Geom g;
    fxd::PhiView phv;
    RotArgs rot;
    double* E[3];
};

namespace fxk {
void fill(hipStream_t s, double* p, long long n, double v);
/* direct solve of a laterally periodic z-layered operator (fx_fft.hip):
 * phi = A^-1 (b + the Dirichlet face terms blo, bhi of the first and last
 * plane); the FP32 plan transforms in FP32 and stores phi in FP64;
 * fft_solve_add: phi += A^-1 r (homogeneous faces) */
bool fft_size_ok(int n);
/* values of the pivot table (FftPlan::piv) and its set-up from the FP64
 * plan (piv32 may be null); abar (n2 values, may be null): a Helmholtz term
 * per plane, the row operator's -a on the diagonal */
size_t fft_pivot_count(int n0, int n1, int n2);
void fft_pivots(hipStream_t s, const fxd::FftPlan<double>& P, double* piv64, float* piv32, const double* abar = nullptr);
/* the upward pivots (FftPlan::pivu, fft_pivot_count values) of the same
 * operator without a Helmholtz term: with them an FP64 solve of 128 planes
 * eliminates from both ends */
void fft_pivots_up(hipStream_t s, const fxd::FftPlan<double>& P, double* pivu);
/* preconditioned conjugate gradients (fx_pcg.hip): reductions leave
 * partials (PCG_PARTS, or the preconditioner's fft_dot_parts for r.z),
 * summed by the kernel that uses them; sc[0..1] hold r.z of the last two
 * iterations, sc[2] the converged flag (set by pcg_publish_max when
 * max|r| <= sc[3]; every other pass does nothing while it is set) */
constexpr int PCG_PARTS = 1024;
constexpr int PCG_ABAR_PARTS = 16;   /* partials per plane of pcg_abar (part holds n2 of them) */
/* abar[k] = mean of the A-coefficient over plane k (lateral extents powers of two) */
void pcg_abar(hipStream_t s, const fxd::DLev& l, double* part, double* abar);
/* the max of the PCG_PARTS partials to pinned host memory, then seq = v */
void pcg_publish_max(hipStream_t s, const double* part, double* host, unsigned* seq, unsigned v, double* sc);
/* the new search direction and its image, from z and A z: rz = the sum of
 * the nrz partials part_rz (written to sc[(it + 1) & 1]); it < 0: p = z,
 * q = A z; else p = z + beta p, q = A z + beta q, beta = rz / sc[it & 1];
 * the partial sums of p.q.  A: the level's row operator with homogeneous
 * faces, for a layered level periodic in x and y */
void pcg_dirmat(hipStream_t s, const fxd::DLev& l, const fxd::DBC& bc, const double* z, double* p, double* q,
                const double* part_rz, int nrz, double* sc, int it, double* part_pq);
void pcg_update(hipStream_t s, double* phi, double* r, const double* p, const double* q, long long N,
                const double* part_pq, const double* sc, int it, double* part_r);
/* planes of b outside [z0, z1] hold zeros (z1 < z0: nothing is known);
 * phi is formed on planes [o0, o1] (o1 < o0: all), the others by a later
 * fft_finish(o0, o1) while P.S is untouched.  Returns true where the z
 * solve, too, stored only planes [o0, o1] of P.S: fft_finish then takes the
 * solve's face terms (face[0], face[1], written by the solve when face is
 * given) to continue it. */
bool fft_solve(hipStream_t s, const fxd::FftPlan<double>& P, const double* b, double* phi, double blo, double bhi,
               int z0 = 0, int z1 = -1, int o0 = 0, int o1 = -1, double* face = nullptr);
void fft_finish(hipStream_t s, const fxd::FftPlan<double>& P, double* phi, int o0, int o1, const double* face);
/* FftPlan::green of an FP64 plan with upward pivots (2 n2 values) */
void fft_face_green(hipStream_t s, const fxd::FftPlan<double>& P, double* g);
/* homogeneous solve phi = A^-1 b on every plane, with the partial sums of
 * phi . b (fft_dot_parts(P) of them) to dotp */
void fft_solve_dot(hipStream_t s, const fxd::FftPlan<double>& P, const double* b, double* phi, double* dotp);
int fft_dot_parts(const fxd::FftPlan<double>& P);
void fft_solve(hipStream_t s, const fxd::FftPlan<float>& P, const double* b, double* phi, double blo, double bhi,
               int z0 = 0, int z1 = -1);
void fft_solve_add(hipStream_t s, const fxd::FftPlan<float>& P, const float* r, double* phi);
/* one pass of the FP64 direct solve: 0 forward x transform (b -> S), 1 y
 * transform of S in place (forward and inverse are the same), 2 tridiagonal
 * solves in z, 3 inverse x transform (S -> phi) */
void fft_pass(hipStream_t s, const fxd::FftPlan<double>& P, int pass, const double* b, double* phi, double blo,
              double bhi);
/* host[0..nq) = src[0..nq), then *seq = v (host and seq: device views of
 * coherent pinned host memory) */
void publish(hipStream_t s, const double* src, int nq, double* host, unsigned* seq, unsigned v);
void copy(hipStream_t s, double* d, const double* src, long long n);
void add(hipStream_t s, double* d, const double* src, long long n);
void final_reduce(hipStream_t s, const double* part, int nb, int nq, int is_max, double* out);
/* out[0] = max|a|, out[1] = max|a-b| (b may be null); part holds 2*2048 doubles */
void maxabs2(hipStream_t s, const double* a, const double* b, long long n, double* part, double* out);
/* diff = |phi - phi_old| / *phi_max (device), then phi_old = phi */
void phidiff(hipStream_t s, const double* phi, double* phi_old, double* diff, long long n, const double* phi_max);
void rhs_newton(hipStream_t s, const RhsArgs& a);
/* returns the number of per-block partials written ([num...][den...]) */
int rho_update(hipStream_t s, const RhoArgs& a);
void tdgl(hipStream_t s, const TdglArgs& a, bool scalar_fast);
void tdgl_rhs_stream(hipStream_t s, const TdglRhsArgs& a);
/* FE cells whose stencils are not all centred: their global indices and the
 * planes that hold them (device arrays) */
struct FaceSet {
    const int* cells;
    int ncells;
    const int* planes;
    int nplanes;
};
/* k_tdgl<true> with the stencil operands streamed through LDS; the cells
 * next to a face beside it on the second stream s2 (events e0, e1 order it
 * after the work queued on s and s after it): mode 1, by the streamed kernel
 * over their planes when those are no more than a quarter of [k0, k1], else
 * by list; 2 always by list; 3 always over their planes.  a.rhs is not used */
void tdgl_stream(hipStream_t s, hipStream_t s2, hipEvent_t e0, hipEvent_t e1, const TdglRhsArgs& a,
                 const FaceSet& fs, int mode);
/* the Poisson RHS of scalar P in planes k0..k1 without SC cells (the RHS
 * buffer 0 in every non-FE cell), streamed through LDS, the cells next to a
 * face as in tdgl_stream */
void rhs_stream(hipStream_t s, hipStream_t s2, hipEvent_t e0, hipEvent_t e1, const RhsArgs& a, int k0, int k1,
                const FaceSet& fs, int mode);
void heun(hipStream_t s, const HeunArgs& a);
void compute_E(hipStream_t s, const EArgs& a);
void mg_smooth(hipStream_t s, const fxd::DLev& l, const fxd::DBC& bc, int inh, int color);
/* nsweeps full red-black sweeps (fused kernel where the level allows it) */
void mg_sweep(hipStream_t s, const fxd::DLev& l, const fxd::DBC& bc, int inh, int nsweeps, int allow_fused,
              int zero_first = 0);
/* prolongation of the coarse correction c into l followed by nsweeps sweeps
 * (the prolongation folded into the first one-pass sweep where it applies) */
void mg_prolong_sweep(hipStream_t s, const fxd::DLev& l, const fxd::DLev& c, const fxd::DBC& bc, int inh, int nsweeps,
                      int allow_fused);
void mg_prolong_sweep(hipStream_t s, const fxd::DLevF& l, const fxd::DLevF& c, const fxd::DBC& bc, int nsweeps,
                      int allow_fused);
bool fused_ok(const fxd::DLev& l, const fxd::DBC& bc);
/* the first streaming sweep of a level on its own, phi -> phi2 (phi is left
 * as it was; false when the level is not streamed one sweep per pass), and
 * the remaining sweeps from phi2, the result in phi */
bool mg_sweep_first(hipStream_t s, const fxd::DLev& l, const fxd::DBC& bc, int inh, int allow_fused);
/* mg_sweep_first that also leaves max|r| and max|b| of its input phi in
 * out[0], out[1] (mg_resid_norm's values, folded into the sweep; part holds
 * two partials per workgroup); zero: the input is phi = 0, not read; false,
 * with nothing queued, where the level has a Helmholtz term or is not
 * streamed */
bool mg_sweep_first_norm(hipStream_t s, const fxd::DLev& l, const fxd::DBC& bc, int inh, int allow_fused, double* part,
                         double* out, bool zero = false);
void mg_sweep_rest(hipStream_t s, const fxd::DLev& l, const fxd::DBC& bc, int inh, int nsweeps, int allow_fused);
bool fused_ok32(const fxd::DLevF& l, const fxd::DBC& bc);
/* residual of f restricted into c's rhs; c's phi is set to 0 unless zero_cphi
 * is false (the caller's next use of c takes its phi as 0 without reading it) */
void mg_resid_restrict(hipStream_t s, const fxd::DLev& f, const fxd::DLev& c, const fxd::DBC& bc, int inh,
                       bool zero_cphi = true);
void mg_prolong(hipStream_t s, const fxd::DLev& f, const fxd::DLev& c, const fxd::DBC& bc);
void mg_restrict_cells(hipStream_t s, const fxd::DLev& f, const fxd::DLev& c, const double* src, double* dst);
/* r = b - A phi (r may be null); out[0] = max|r|, out[1] = max|b|; part holds
 * 2*resid_norm_blocks(l) doubles */
void mg_resid_norm(hipStream_t s, const fxd::DLev& l, const fxd::DBC& bc, int inh, double* r, double* part, double* out);
int resid_norm_blocks(const fxd::DLev& l);
void final1024(hipStream_t s, const double* part, int nb, int nq, int is_max, double* out);
int resid_blocks(const fxd::DLev& l);
void bottom_cg(hipStream_t s, const fxd::DLev& l, const fxd::DBC& bc, const double* rhs, double* e, double* work,
               double tol, int maxit, int* iters);
/* FP32 correction hierarchy (homogeneous BCs) */
void mg_sweep(hipStream_t s, const fxd::DLevF& l, const fxd::DBC& bc, int nsweeps, int zero_first, int allow_fused);
/* the zero-initial first sweep of a correction solve on its own, into phi
 * (its input is not read), and the remaining sweeps from phi */
bool mg_sweep_first(hipStream_t s, const fxd::DLevF& l, const fxd::DBC& bc, int allow_fused);
void mg_sweep_rest(hipStream_t s, const fxd::DLevF& l, const fxd::DBC& bc, int nsweeps, int allow_fused);
void mg_resid_restrict(hipStream_t s, const fxd::DLevF& f, const fxd::DLevF& c, const fxd::DBC& bc, bool zero_cphi = true);
void mg_prolong(hipStream_t s, const fxd::DLevF& f, const fxd::DLevF& c, const fxd::DBC& bc);
void mg_restrict_cells(hipStream_t s, const fxd::DLevF& f, const fxd::DLevF& c, const float* src, float* dst);
void bottom_cg(hipStream_t s, const fxd::DLevF& l, const fxd::DBC& bc, const float* rhs, float* e, float* work,
               double tol, int maxit, int* iters);
void d2f(hipStream_t s, float* d, const double* src, long long n);
/* phi += e (set: phi = e, phi being 0) */
void add_f2d(hipStream_t s, double* phi, const float* e, long long n, bool set = false);
/* whole sub-V-cycle of the levels levs[0..nl-1] (device array) in one workgroup */
void mg_tail(hipStream_t s, const fxd::DLev* levs, int nl, const fxd::DBC& bc, int read_a, int nu1, int nu2, double* work,
             double tol, int maxit, int* iters, int ts, size_t lds_bytes);
void mg_tail(hipStream_t s, const fxd::DLevF* levs, int nl, const fxd::DBC& bc, int read_a, int nu1, int nu2, float* work,
             double tol, int maxit, int* iters, int ts, size_t lds_bytes);
/* largest dynamic LDS per workgroup on the current device */
size_t tail_lds_max();
/* LDS values (of the level's type) the tail kernel stages for a level */
size_t tail_level_values(const fxd::DLev& l);
/* bottom solves of more cells than this use Jacobi-PCG, below it the exact dense solve */
constexpr int BOTTOM_DIRECT_CELLS = 64;
/* the sub-V-cycle of a layered hierarchy's coarse levels in one workgroup,
 * every level in LDS; the bottom's inverse and perm from bottom_factor */
void mg_tail_lay(hipStream_t s, const fxd::TailPlan<double>& P, const fxd::DBC& bc, int read_a, int nu1, int nu2,
                 double* work, double tol, int maxit, int* iters);
void mg_tail_lay(hipStream_t s, const fxd::TailPlan<float>& P, const fxd::DBC& bc, int read_a, int nu1, int nu2,
                 float* work, double tol, int maxit, int* iters);
/* exact solve of one laterally periodic layered level (homogeneous faces):
 * per-plane transforms, a tridiagonal solve per mode, inverse transforms */
void mg_spec_solve(hipStream_t s, const fxd::SpecTail<double>& P);
void mg_spec_solve(hipStream_t s, const fxd::SpecTail<float>& P);
/* one launch per leg for a small layered level l above c: down = nu1 sweeps
 * from phi = 0 + residual restricted into c, written to l.phi; up =
 * prolongation from c + nu2 sweeps, read from l.phi and written to out, a
 * second buffer of the level's size (each workgroup reads the halo of its
 * neighbours' blocks, so the leg cannot update l.phi in place); the caller
 * then takes out as the level's phi (mg_block_ok: the level qualifies and
 * the blocks fit in LDS) */
bool mg_block_ok(const fxd::DLev& l, const fxd::DLev& c, const fxd::DBC& bc, int nu1, int nu2);
bool mg_block_ok(const fxd::DLevF& l, const fxd::DLevF& c, const fxd::DBC& bc, int nu1, int nu2);
void mg_block_down(hipStream_t s, const fxd::DLev& l, const fxd::DLev& c, const fxd::DBC& bc, int nu1, int nu2);
void mg_block_down(hipStream_t s, const fxd::DLevF& l, const fxd::DLevF& c, const fxd::DBC& bc, int nu1, int nu2);
void mg_block_up(hipStream_t s, const fxd::DLev& l, const fxd::DLev& c, const fxd::DBC& bc, int nu1, int nu2, double* out);
void mg_block_up(hipStream_t s, const fxd::DLevF& l, const fxd::DLevF& c, const fxd::DBC& bc, int nu1, int nu2, float* out);
/* dynamic LDS bytes of mg_tail_lay for levels of extents n[0..nl-1] */
size_t tail_lay_lds(const int (*n)[3], int nl, int read_a, size_t elem);
/* Inverse of the bottom operator (<= BOTTOM_DIRECT_CELLS cells) through its
 * LU factors with partial pivoting: inv N*N values column-major, perm N+1
 * ints (the pivot order; perm[N] = 1 when non-singular) */
void bottom_factor(hipStream_t s, const fxd::DLev& l, const fxd::DBC& bc, double* inv, int* perm);
void bottom_factor(hipStream_t s, const fxd::DLevF& l, const fxd::DBC& bc, float* inv, int* perm);
/* r32 = (float)(b - A phi) in FP64 (inhomogeneous BCs); out[0] = max|r|, out[1] = max|b| */
void resid_to32(hipStream_t s, const fxd::DLev& l, const fxd::DBC& bc, float* r32, double* part, double* out, bool zero = false);
/* ------------------------------- adjoint ---------------------------------- */
/* One backward step of the discrete adjoint (fx_adjoint.hip). */
struct AdjArgs {
    fxd::Geom g;
    fxd::MView M;
    fxd::PBC pbc;
    const double* lam;       /* lambda_{n+1} */
    double* lam_out;         /* lambda_n */
    const double* P;         /* P_n */
    const double* phi;       /* phi_n */
    const double* psi;       /* A^-1 Fz^T nu */
    double* s;               /* Fz^T nu */
    double dtG;              /* dt * BigGamma */
    double alpha, beta, gamma, g11, g44m;
    double loss_w;           /* loss weight added to lambda_n on FE cells */
    double eps0;
    int bclo[3], bchi[3];    /* potential BC kinds */
    double bcv[6];           /* Dirichlet values of step n */
    const double* eps;       /* cell permittivity */
    double ih2[3];
    int ghost_face;
    double* part;            /* 7 * adj_blocks partial sums */
};
int adj_blocks(const fxd::Geom& g);
void adj_fzT(hipStream_t s, const AdjArgs& a);
void adj_step(hipStream_t s, const AdjArgs& a);
void adj_accum(hipStream_t s, const double* part, int nb, double* acc, double* gV, int step);
void adj_init(hipStream_t s, double* lam, const fxd::MView& M, const fxd::Geom& g, double w);
/* out[0] = sum of P over the FE (non t-phase) cells; part: adj_blocks doubles */
void fe_sum(hipStream_t s, const double* P, const fxd::MView& M, const fxd::Geom& g, double* part, double* out);
}  // namespace fxk
