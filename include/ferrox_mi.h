/*
 * ferrox_mi.h -- C ABI of the MI355X-native FerroX TDGL + Poisson hot path.
 *
 * The reference has no plugin or FFI layer: its hot path is a set of free C++
 * functions over amrex::MultiFab, called only from main_main's time loop
 * (Source/main.cpp:261-623).  This header is the drop-in boundary that replaces
 * them.  Each entry point names the reference function(s) it stands for.
 *
 * Conventions
 *   - One context = one device (one rectangular box of cells), one HIP stream.
 *     Calls on one context are not thread-safe; different contexts are
 *     independent.
 *   - Host buffers are dense over the valid region, i fastest, then j, then k
 *     (the order of an AMReX plotfile FAB of a single box), FP64.
 *   - Every call returns an fx_status; fx_last_error() gives the message.
 *   - Results are deterministic: reductions run in a fixed order.
 */
#ifndef FERROX_MI_H
#define FERROX_MI_H

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes (the reference aborts or loops forever in these cases). */
typedef enum fx_status {
    FX_OK = 0,
    FX_EINVAL = 1,          /* bad key, flag or argument (amrex::Abort in the reference, e.g. DerivativeAlgorithm.H:72) */
    FX_ENOCONV_MG = 2,      /* multigrid did not reach its tolerance in mg_max_iter V-cycles (MLMG abort) */
    FX_ENOCONV_OUTER = 3,   /* semiconductor outer loop exceeded outer_max_iter (reference loops forever, ElectrostaticSolver.cpp:703,742) */
    FX_ENAN = 4,            /* non-finite value in a reduction */
    FX_EHIP = 5,            /* HIP runtime error */
    FX_ENOMEM = 6
} fx_status;

/* Boundary-condition kinds for the Poisson problem (boundary.lo / boundary.hi tokens,
 * BoundaryConditions.cpp:52-214): per, dir(...), neu. */
enum { FX_BC_PERIODIC = 0, FX_BC_DIRICHLET = 1, FX_BC_NEUMANN = 2 };

/* Where the ghost value of phi at a Dirichlet face sits when E is formed
 * (ElectrostaticSolver.cpp:134-203 reads the ghost after MLMG::solve). */
enum {
    FX_GHOST_AMREX = 0,   /* ghost = 2*phi_b - phi_0, the linear extrapolation the solver leaves behind */
    FX_GHOST_FACE = 1     /* ghost = phi_b, the face value */
};

/* Arithmetic of the Poisson solve. */
enum { FX_PREC_FP64 = 0, FX_PREC_FP32_REFINE = 1 };

/*
 * Parsed inputs.  Field names are the reference's inputs keys
 * (FerroX.cpp:273-572, GeometryProperties.cpp:107-124); the trailing block
 * holds the solver controls this build adds.  Fill with fx_inputs_default()
 * then fx_inputs_parse_file() / fx_inputs_set().
 */
typedef struct fx_inputs {
    /* domain.* */
    int    n_cell[3];
    double prob_lo[3];
    double prob_hi[3];
    int    is_periodic[3];
    int    max_grid_size[3];      /* accepted, not used: one box per context */
    /* boundary.lo / boundary.hi */
    int    bc_lo[3];
    int    bc_hi[3];
    /* control */
    int    TimeIntegratorOrder;
    int    prob_type;
    int    is_polarization_scalar;
    int    nsteps;
    int    plot_int;
    double dt;
    double delta;
    int    random_seed;
    int    voltage_sweep;
    int    inc_step;
    int    num_Vapp_max;
    double Phi_Bc_lo;
    double Phi_Bc_hi;
    double Phi_Bc_inc;
    double Phi_Bc_hi_max;
    double phi_tolerance;
    int    include_Landau;
    int    include_Grad;
    int    include_Elec;
    int    P_BC_flag_lo[3];
    int    P_BC_flag_hi[3];
    double lambda;
    /* materials */
    double epsilon_0;
    double epsilonX_fe;
    double epsilonX_fe_tphase;
    double epsilonZ_fe;           /* read but unused by the reference (FerroX.cpp:347) */
    double epsilon_de;
    double epsilon_si;
    double alpha;
    double beta;
    double gamma;
    double alpha_12;
    double alpha_112;
    double alpha_123;
    double BigGamma;
    double g11;
    double g44;
    double g12;
    double g44_p;
    /* stack geometry */
    double FE_lo[3], FE_hi[3];
    double DE_lo[3], DE_hi[3];
    double SC_lo[3], SC_hi[3];
    double Channel_lo[3], Channel_hi[3];
    double Remnant_P[3];
    /* semiconductor */
    double acceptor_doping;
    double donor_doping;
    int    use_Fermi_Dirac;
    int    use_work_function;
    double metal_work_function;
    /* rotated frame */
    int    Coordinate_Transformation;
    int    use_Euler_angles;
    /* ---- solver controls added by this build ---- */
    double mg_tol_rel;            /* 1e-10 = the reference's MLMG::solve(…, 1e-10, -1) */
    int    mg_max_iter;           /* 200 */
    int    mg_nu1;                /* pre-smoothing RB-GS sweeps (2) */
    int    mg_nu2;                /* post-smoothing RB-GS sweeps (2) */
    int    mg_warm_start;         /* 0: zero initial guess as the reference (ElectrostaticSolver.cpp:716) */
    int    outer_max_iter;        /* cap on the semiconductor outer loop */
    int    phi_ghost_mode;        /* FX_GHOST_AMREX (default) or FX_GHOST_FACE */
    int    precision;             /* FX_PREC_FP64 (default) or FX_PREC_FP32_REFINE */
    int    bias_ramp_steps;       /* >0: Phi_Bc_hi(step) = Phi_Bc_hi_start + step/bias_ramp_steps*(Phi_Bc_hi_end-start) */
    double bias_ramp_start;
    double bias_ramp_end;
    int    verbose;
    int    mg_fused_sweep;        /* streamed levels: 1 (default) one red+black sweep per pass, 2 two sweeps per pass, 0 two launches per sweep */
    int    mg_block_levels;       /* 1: small layered levels run their down and up legs as one launch each, column blocks in LDS (default 1) */
    int    mg_stream_min_cells;   /* levels of at least this many cells use the streaming kernels (default 2097152) */
    double mg_coarsen_ratio;      /* direction d is coarsened when h_d < ratio * min(h) (default 1.9) */
    int    mg_exact_tail;         /* 1 (default): laterally periodic layered hierarchies solve the tail's first level
                                     exactly (lateral Fourier transforms + tridiagonal solves) instead of a sub-V-cycle */
    int    mg_refine_nu1;         /* precision = 1: pre- and post-smoothing sweeps of the FP32 correction V-cycles */
    int    mg_refine_nu2;         /* (defaults 1 and 2; the FP64 V-cycles use mg_nu1, mg_nu2) */
    int    poisson_direct;        /* 1 (default): a laterally periodic stack whose permittivity depends on z only,
                                     with a Dirichlet z face, lateral extents powers of two (4..1024) and no
                                     Helmholtz term is solved directly (lateral Hartley transforms + one
                                     tridiagonal solve per mode) instead of by multigrid; 0: multigrid always */
    int    poisson_direct_check;  /* 1: after each direct solve, form the FP64 residual and apply the multigrid
                                     stopping rule to it (FX_ENOCONV_MG if it fails); 0 (default): no check */
    int    tdgl_fused_rhs;        /* 1: with scalar P, forward Euler and no semiconductor cells, the TDGL update
                                     also forms the next step's Poisson RHS (one pass over P per step, z-streamed
                                     through LDS); 0 (default, measured faster on MI355X): separate TDGL and RHS
                                     passes */
    int    tdgl_stream;           /* 1 (default): the scalar-P forward-Euler update and the Poisson RHS stream
                                     P and phi through LDS (k_tdgl_stream, k_rhs_stream); the cells next to a
                                     face beside them on a second stream, streamed over the planes that hold
                                     them when those are at most a quarter of the FE planes, else by list;
                                     2: always by list; 3: always over their planes; 0: one thread per cell
                                     with gathered operands (k_tdgl, k_rhs_newton) */
    int    poisson_pcg;           /* 1 (default): where the direct solve applies but for a Helmholtz term
                                     (semiconductor stacks), FP64 solves are preconditioned conjugate gradients
                                     with the direct solve of the plane-averaged operator as preconditioner,
                                     under the multigrid stopping rule; 0: multigrid */
    /* plotfile variable switches, default 1 (FerroX.cpp:300-325) */
    int    plot_Phi, plot_PoissonRHS, plot_E, plot_holes, plot_electrons, plot_charge, plot_epsilon;
    int    plot_mask, plot_tphase, plot_alpha, plot_beta, plot_theta, plot_PhiDiff;
    /* Coordinate_Transformation = 1: functions of the cell centre (x,y,z)
     * (Initialization.cpp:253-383), my_constants substituted; keys
     * tphase_geom.tphase_geom_function(x,y,z), angle_alpha.alpha_function(x,y,z),
     * angle_beta.beta_function(x,y,z), angle_theta.theta_function(x,y,z) */
    char   tphase_geom_function[4096];
    char   angle_alpha_function[2048];
    char   angle_beta_function[2048];
    char   angle_theta_function[2048];
} fx_inputs;

/* Evaluates a function of the cell centre (x,y,z) -- the AMReX parser
 * subset: + - * / ^ **, comparisons < > <= >= == != (1 or 0), and/or,
 * if(c,a,b), sqrt exp log log10 sin cos tan asin acos atan sinh cosh tanh abs
 * floor ceil, pow atan2 min max fmod, pi -- at every cell centre of the
 * inputs' domain, i fastest (ConvertParserIntoMultiFab_3vars,
 * eXstaticUtil.cpp). */
int fx_eval_cell_function(const fx_inputs* in, const char* expr, double* out, char* err, int errlen);

/* Field identifiers.  0..17 follow the plotfile variable order of the
 * reference (Plotfile.cpp:30-106). */
typedef enum fx_field {
    FX_FIELD_PX = 0, FX_FIELD_PY, FX_FIELD_PZ,
    FX_FIELD_PHI, FX_FIELD_POISSON_RHS,
    FX_FIELD_EX, FX_FIELD_EY, FX_FIELD_EZ,
    FX_FIELD_HOLES, FX_FIELD_ELECTRONS, FX_FIELD_CHARGE,
    FX_FIELD_EPSILON, FX_FIELD_MASK, FX_FIELD_TPHASE,
    FX_FIELD_ANGLE_ALPHA, FX_FIELD_ANGLE_BETA, FX_FIELD_ANGLE_THETA,
    FX_FIELD_PHIDIFF,
    FX_FIELD_ALPHA_CC,            /* Newton A-coefficient of the last solve */
    FX_FIELD_GAMMA,               /* BigGamma in FE, 0 elsewhere */
    FX_FIELD_GL_RHS_X, FX_FIELD_GL_RHS_Y, FX_FIELD_GL_RHS_Z,   /* TDGL right-hand side, last fx_tdgl_rhs */
    FX_FIELD_LANDAU_ALPHA,        /* per-cell Landau alpha (extension: the reference's alpha is global,
                                     FerroX.cpp:350); defaults to the inputs' alpha everywhere */
    FX_FIELD_COUNT
} fx_field;

typedef struct fx_solve_stats {
    int    outer_iters;           /* iterations of the semiconductor outer loop (1 without SC) */
    int    vcycles;               /* V-cycles summed over the outer loop */
    double outer_err;             /* last ||phi_k - phi_{k-1}||_1 / ||phi_k||_1 (0 without SC) */
    double mg_resid0;             /* ||r0||_inf of the last solve */
    double mg_resid;              /* ||r||_inf at exit of the last solve */
    double mg_bnorm;              /* ||b||_inf of the last solve */
} fx_solve_stats;

typedef struct fx_step_stats {
    int    steps_done;
    int    stopped_steady;        /* 1 if the steady-state test ended the loop (main.cpp:606-609) */
    int    total_vcycles;
    int    total_outer_iters;
    int    total_solves;
    double last_phi_change;       /* CheckSteadyState metric of the last step (ElectrostaticSolver.cpp:546) */
    double time;                  /* simulated time after the loop */
    double Phi_Bc_hi;             /* current top bias */
    int    num_Vapp;
    double wall_seconds;          /* host wall time of the call (device synchronized) */
} fx_step_stats;

typedef struct fx_ctx fx_ctx;

/* Library / inputs -------------------------------------------------------- */
const char* fx_version(void);
/* sizeof(fx_inputs), for bindings that mirror the struct. */
int fx_inputs_size(void);
/* Fills every key with the reference default (FerroX.cpp:273-572; Appendix of keys). */
int fx_inputs_default(fx_inputs* in);
/* ParmParse-compatible reader for a FerroX inputs file: key = v1 v2 ..., '#'
 * comments, last-wins duplicates, "key=value" overrides in argv style
 * (replaces amrex::ParmParse + InitializeFerroXNamespace, FerroX.cpp:273). */
int fx_inputs_parse_file(const char* path, fx_inputs* in, char* err, int errlen);
int fx_inputs_set(fx_inputs* in, const char* key_equals_value, char* err, int errlen);
/* Required-key check (pp.get aborts when a key is missing) and the defaults
 * that depend on other keys (DE/SC/Channel boxes outside the domain,
 * epsilonX_fe_tphase = epsilonX_fe, BCs from domain.is_periodic).  fx_create
 * calls it on its own copy; idempotent. */
int fx_inputs_finalize(fx_inputs* in, char* err, int errlen);

/* Context ----------------------------------------------------------------- */
/* Builds the material mask, permittivity, multigrid hierarchy and initial P
 * and rho on the device.  Replaces main_main setup, main.cpp:65-226:
 * InitializeMaterialMask (Initialization.cpp:169), InitializePermittivity
 * (ElectrostaticSolver.cpp:205), AverageCellCenteredMultiFabToCellFaces
 * (eXstaticUtil.cpp:105), SetupMLMG (ElectrostaticSolver.cpp:562),
 * InitializePandRho (Initialization.cpp:6). */
int fx_create(const fx_inputs* in, fx_ctx** out);
int fx_destroy(fx_ctx* ctx);
const char* fx_last_error(fx_ctx* ctx);
int fx_get_dims(fx_ctx* ctx, int dims[3]);

/* Host <-> device copies of one field (dense valid region, i fastest). */
int fx_get_field(fx_ctx* ctx, int field, double* host);
/* Settable: PX, PY, PZ, PHI, CHARGE, TPHASE, ANGLE_ALPHA/BETA/THETA, MASK.
 * Setting MASK/TPHASE/angles rebuilds permittivity and the multigrid
 * coefficients (Initialize_tphase_Mask / Initialize_Euler_angles,
 * Initialization.cpp:253-383). */
int fx_set_field(fx_ctx* ctx, int field, const double* host);

/* Top/bottom Dirichlet values of phi.  Replaces SetPhiBC_z + setLevelBC
 * (ElectrostaticSolver.cpp:502-524, main.cpp:585-592). */
int fx_set_bias(fx_ctx* ctx, double phi_bc_lo, double phi_bc_hi);

/* Self-consistent phi and rho for the current P: the outer loop of
 * ComputePhi_Rho (ElectrostaticSolver.cpp:678-747) with ComputePoissonRHS
 * (:8), dF_dPhi (:85), ComputePoissonRHS_Newton (:114), MLMG solve
 * (:713-720) and ComputeRho (ChargeDensity.cpp:15). */
int fx_compute_phi_rho(fx_ctx* ctx, fx_solve_stats* stats);

/* E = -R grad(phi) into the E fields.  ComputeEfromPhi,
 * ElectrostaticSolver.cpp:134-203. */
int fx_compute_E(fx_ctx* ctx);

/* GL right-hand side of the current P and phi into the GL_RHS fields.
 * Calculate_Landau + Calculate_Grad + Calculate_Elec + CalculateTDGL_RHS,
 * TotalEnergyDensity.cpp:6-294. */
int fx_tdgl_rhs(fx_ctx* ctx);

/* One forward-Euler (order 1) or Heun (order 2) update of P using the current
 * phi for the first stage: main.cpp:293-348. */
int fx_update(fx_ctx* ctx, int order);

/* nsteps iterations of the reference time loop, main.cpp:261-623: solve,
 * update, CheckSteadyState (ElectrostaticSolver.cpp:526), voltage sweep and
 * the stop tests.  nsteps <= 0 means the inputs' nsteps. */
int fx_step(fx_ctx* ctx, int nsteps, fx_step_stats* stats);

/* Multigrid solve of  -alpha_cc*phi - div(eps grad phi) = rhs  on the
 * context's geometry and BCs, from host arrays (test and tooling entry point;
 * replaces MLMG::solve, ElectrostaticSolver.cpp:719).  acoef may be NULL (0). */
int fx_poisson_solve(fx_ctx* ctx, const double* rhs, const double* acoef,
                     double* phi, fx_solve_stats* stats);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): phi_dev = A^-1 rhs_dev with the context's operator and
 * current bias, on device arrays of N doubles (dense, i fastest), queued on
 * the context's stream (fx_stream) with no host wait and no allocation.
 * Only where fx_poisson_solve takes the direct path (x and y periodic,
 * permittivity depending on z only, a Dirichlet z face, lateral extents
 * powers of two, no Helmholtz term).  The transform workspace, N doubles,
 * lives in a static device array of 64 MiB per GPU (8388608 cells at most,
 * e.g. 256 x 256 x 128), shared by every call on that GPU and left holding
 * the last call's values: calls that overlap on one GPU race.
 * FX_EINVAL off the direct path or beyond the limit. */
int fx_poisson_solve_noalloc(fx_ctx* ctx, const double* rhs_dev, double* phi_dev);

/* AMReX single-level plotfile of the current state into directory dir
 * (WritePlotfile, Plotfile.cpp:5-177; variables gated by the plot_* inputs
 * flags).  fx_step writes plt%08d itself when plot_int > 0, at step 0 and
 * every plot_int steps and at the steady-state step (main.cpp:228-234,
 * 563-569), into the directory set by fx_set_plot_dir (default "."). */
int fx_write_plotfile(fx_ctx* ctx, const char* dir, double time, int step);
int fx_set_plot_dir(fx_ctx* ctx, const char* dir);

/* Per-step top bias (extension: the reference changes Phi_Bc_hi only at
 * steady-state / inc_step events, main.cpp:572-604).  Step s (1-based, counted
 * over the context's lifetime) uses v[min(s, n) - 1]; n = 0 clears it. */
int fx_set_bias_waveform(fx_ctx* ctx, const double* v, int n);

/* Measurement: average duration in ms of one launch of a named hot-path
 * kernel ("smooth0", "smooth0x2", "prolong0", "resid0", "residmax0", "tdgl",
 * "rhs", "copy"),
 * timed with HIP events
 * on the context's stream.  No reference counterpart (instrumentation). */
int fx_bench_kernel(fx_ctx* ctx, const char* which, int reps, double* avg_ms);

/* ---- Differentiation form (extension; no counterpart in the reference) ----
 * Discrete adjoint of the forward-Euler TDGL step (main.cpp:293-306) and its
 * Poisson solve (ElectrostaticSolver.cpp:678-747) for the scalar, unrotated,
 * semiconductor-free case, with checkpoint-and-recompute.  theta holds
 * (alpha, beta, gamma, g11, g44, epsilonX_fe) in the inputs' units. */
int fx_diff_set_params(fx_ctx* ctx, const double theta[6]);
int fx_diff_get_params(fx_ctx* ctx, double theta[6]);
/* nsteps Euler steps from the current state, step n with top bias V[n];
 * P is kept every K steps.  m_out[n] = mean of Pz over the ferroelectric
 * cells after step n (the P-V curve). */
int fx_diff_forward(fx_ctx* ctx, const double* V, int nsteps, int K, double* m_out);
/* Gradients of a loss L(m) through the last fx_diff_forward: given dL/dm
 * (nsteps values), returns dL/dtheta (6) and dL/dV (nsteps; either output may
 * be NULL).  The state is left as after the forward. */
int fx_diff_backward(fx_ctx* ctx, const double* dL_dm, double* grad_theta, double* grad_V);

/* ---- Batching form (extension; no counterpart in the reference) ----------
 * n independent devices, each given by its own inputs (periodic in x and y,
 * Dirichlet z faces, laterally uniform layers, scalar P, forward Euler,
 * lateral extents <= 128), advanced together: every stage of a step is one
 * persistent launch over all devices, a step pair is a captured hipGraph,
 * and Poisson is solved directly (lateral real-Fourier transforms + a
 * tridiagonal solve per mode; FP32 with one FP64 residual correction when
 * inputs[0]->precision is FX_PREC_FP32_REFINE).  Set-up of each device
 * (materials, permittivity, initial P, bias) is that of fx_create. */
typedef struct fx_batch fx_batch;
int fx_batch_create(int n, const fx_inputs* const* inputs, fx_batch** out);
int fx_batch_destroy(fx_batch* batch);
const char* fx_batch_last_error(fx_batch* batch);
int fx_batch_num_devices(fx_batch* batch);
long long fx_batch_cells(fx_batch* batch);
int fx_batch_dims(fx_batch* batch, int dev, int dims[3]);
/* nsteps steps of every device (use_graph: replay the captured step pair) */
int fx_batch_step(fx_batch* batch, int nsteps, int use_graph, double* wall_seconds);
/* FX_FIELD_PZ or FX_FIELD_PHI of one device (host array, i fastest) */
int fx_batch_get_field(fx_batch* batch, int dev, int field, double* host);
int fx_batch_set_field(fx_batch* batch, int dev, int field, const double* host);
/* measure and choose tile shapes and grid sizes; JSON summary into report */
int fx_batch_autotune(fx_batch* batch, char* report, int report_len);
/* the real Fourier basis of an n-point periodic direction (n x n, column
 * vectors) and the eigenvalues of the 1-D periodic second difference */
int fx_batch_basis(int n, double* Q, double* mu);

/* Multigrid hierarchy (diagnostics). */
int fx_num_levels(fx_ctx* ctx);
int fx_level_dims(fx_ctx* ctx, int level, int dims[3]);

/* Device-resident field pointer (FP64, dense, i fastest) for zero-copy
 * callers such as the torch layer.  Valid until fx_destroy. */
int fx_field_device_ptr(fx_ctx* ctx, int field, void** dptr);
/* Device synchronisation of the context's stream. */
int fx_sync(fx_ctx* ctx);
/* HIP stream of the context (hipStream_t). */
void* fx_stream(fx_ctx* ctx);

#ifdef __cplusplus
}
#endif
#endif
