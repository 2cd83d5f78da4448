/*
 * HIP kernels (gfx950) of the FerroX per-timestep hot path.
 *
 *   k_rhs_newton   ComputePoissonRHS + dF_dPhi + ComputePoissonRHS_Newton
 *                  (ElectrostaticSolver.cpp:8-132) fused: one pass writes the
 *                  Newton right-hand side and the A-coefficient.
 *   k_rho_update   ComputeRho (ChargeDensity.cpp:15-102) fused with the L1
 *                  norms of the outer-loop error (ElectrostaticSolver.cpp:731-738).
 *   k_tdgl<...>    Calculate_Landau + Calculate_Grad + Calculate_Elec +
 *                  CalculateTDGL_RHS (TotalEnergyDensity.cpp) with E = -R grad(phi)
 *                  formed on the fly (ComputeEfromPhi, ElectrostaticSolver.cpp:134)
 *                  and the forward-Euler update (main.cpp:293-306) fused, so P and
 *                  phi cross HBM once per step.
 *   k_compute_E    ComputeEfromPhi for output.
 * The multigrid kernels live in fx_mg.hip.
 */
#include "fx_device.h"
#include "fx_kernels.h"

using namespace fxd;

/* --------------------------------- utils --------------------------------- */
__global__ void k_fill(double* __restrict__ p, long long n, double v) {
    for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < n; c += (long long)gridDim.x * blockDim.x) p[c] = v;
}
__global__ void k_copy(double* __restrict__ d, const double* __restrict__ s, long long n) {
    for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < n; c += (long long)gridDim.x * blockDim.x) d[c] = s[c];
}
__global__ void k_add(double* __restrict__ d, const double* __restrict__ s, long long n) {
    for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < n; c += (long long)gridDim.x * blockDim.x) d[c] += s[c];
}

/* The host reads reductions without a stream synchronisation: one lane
 * copies them into coherent pinned host memory, fences at system scope and
 * then publishes the sequence number the host is waiting for. */
__global__ void k_publish(const double* __restrict__ src, int nq, double* host, unsigned* seq, unsigned v) {
    if (threadIdx.x == 0) {
        for (int q = 0; q < nq; ++q) host[q] = src[q];
        __threadfence_system();
        __hip_atomic_store(seq, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

/* Final reduction of per-block partials: out[q] = reduce_b partial[q*nb + b]. */
__global__ __launch_bounds__(256) void k_final(const double* __restrict__ part, int nb, int nq, int is_max, double* __restrict__ out) {
    __shared__ double sh[256 / 64 + 1];
    for (int q = 0; q < nq; ++q) {
        double v = 0.0;
        for (int b = threadIdx.x; b < nb; b += 256) v = is_max ? nmax(v, part[q * nb + b]) : v + part[q * nb + b];
        v = is_max ? block_max<256>(v, sh) : block_sum<256>(v, sh);
        if (threadIdx.x == 0) out[q] = v;
    }
}

/* max|a| and max|a-b| (b may be null) per workgroup. */
__global__ __launch_bounds__(256) void k_maxabs2(const double* __restrict__ a, const double* __restrict__ b, long long n,
                                                 double* __restrict__ part) {
    __shared__ double sh[256 / 64 + 1];
    double m1 = 0.0, m2 = 0.0;
    for (long long c = blockIdx.x * 256LL + threadIdx.x; c < n; c += (long long)gridDim.x * 256) {
        m1 = nmax(m1, fabs(a[c]));
        if (b) m2 = nmax(m2, fabs(a[c] - b[c]));
    }
    m1 = block_max<256>(m1, sh);
    m2 = block_max<256>(m2, sh);
    if (threadIdx.x == 0) { part[blockIdx.x] = m1; part[gridDim.x + blockIdx.x] = m2; }
}

/* PhiDiff = |phi - phi_old| / phi_max, then phi_old = phi (CheckSteadyState,
 * ElectrostaticSolver.cpp:531-556). */
__global__ void k_phidiff(const double* __restrict__ phi, double* __restrict__ phi_old, double* __restrict__ diff,
                          long long n, const double* __restrict__ phi_max_p) {
    const double phi_max = *phi_max_p;
    for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < n; c += (long long)gridDim.x * blockDim.x) {
        diff[c] = fabs(phi[c] - phi_old[c]) / phi_max;
        phi_old[c] = phi[c];
    }
}

/* --------------------------- charge density ------------------------------- */
__device__ __forceinline__ double fd_half(double eta) {
    const double nu = pow(eta, 4.0) + 50.0 + 33.6 * eta * (1.0 - 0.68 * exp(-0.17 * pow((eta + 1.0), 2.0)));
    const double xi = 3.0 * sqrt(3.14) / (4.0 * pow(nu, 3. / 8.));
    return pow(exp(-eta) + xi, -1.0);
}

/* n, p and rho at potential phi in a semiconductor cell (ChargeDensity.cpp:46-93). */
__device__ __forceinline__ double rho_of(const SCParams& s, double phi, double& n, double& p) {
    const double q = s.q, kb = s.kb, T = s.T, Ef = 0.0, Eg = s.Eg, Chi = s.Chi;
    const double phi_ref = s.phi_ref;
    const double Ec_corr = -q * (phi - phi_ref) - Chi * q;
    const double Ev_corr = Ec_corr - q * Eg;
    if (s.fermi_dirac == 1) {
        const double eta_n = -(Ec_corr - q * Ef) / (kb * T);
        const double eta_p = -(q * Ef - Ev_corr) / (kb * T);
        n = s.Nc * fd_half(eta_n);
        p = s.Nv * fd_half(eta_p);
    } else {
        n = s.Nc * exp(-(Ec_corr - q * Ef) / (kb * T));
        p = s.Nv * exp(-(q * Ef - Ev_corr) / (kb * T));
    }
    const double NA = s.NA / (1.0 + s.gA * exp((-q * Ef + q * s.Ea + q * phi_ref - q * Chi - q * Eg - q * phi) / (kb * T)));
    const double ND = s.ND / (1.0 + s.gD * exp((q * Ef + q * s.Ed - q * phi_ref + q * Chi + q * phi) / (kb * T)));
    return q * (p - n - NA + ND);
}

__device__ __forceinline__ void rot_of(const RotArgs& r, int c, double R[9]) {
    if (r.idx) {
        const double* t = r.tab + 9 * r.idx[c];
#pragma unroll
        for (int q = 0; q < 9; ++q) R[q] = t[q];
    } else {
#pragma unroll
        for (int q = 0; q < 9; ++q) R[q] = r.Ru[q];
    }
}

/* R_31, R_32, R_33 into R[6..8] (the rest left unset) */
__device__ __forceinline__ void rot_row3(const RotArgs& r, int c, double R[9]) {
    if (r.idx) {
        const double* t = r.tab + 9 * r.idx[c];
        R[6] = t[6]; R[7] = t[7]; R[8] = t[8];
    } else {
        R[6] = r.Ru[6]; R[7] = r.Ru[7]; R[8] = r.Ru[8];
    }
}

/* Poisson RHS + Newton linearisation, one thread per cell. */
__global__ __launch_bounds__(256) void k_rhs_newton(RhsArgs a) {
    const int nx = a.g.n[0], ny = a.g.n[1], nz = a.g.n[2];
    const int i = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int k = a.kz0 + blockIdx.z;
    if (i >= nx || j >= ny || k >= nz) return;
    const int c = i + nx * (j + ny * k);
    const int code = a.M.code(i, j, k);
    double rhs, al;
    if (code >= 2) {
        const double rho = a.rho[c], ph = a.phi[c];
        double n, p;
        const double rho_d = rho_of(a.sc, ph + a.delta, n, p);
        al = (1. / a.delta) * rho_d + (-1. / a.delta) * rho;
        rhs = rho - al * ph;
    } else if (code == 1) {
        rhs = 0.;
        al = 0.;
    } else {
        double R[9];
        if (a.pxy_zero) rot_row3(a.rot, c, R);   /* scalar P: only R_3 enters */
        else rot_of(a.rot, c, R);
        const PView Fr{a.P[2], nx, ny, nz, a.g.per[0], a.g.per[1], a.g.per[2]};
        const double* dx = a.g.dx;
        double x = 0.0;
        if (a.pxy_zero && (a.M.raw(i, j, k) & MAT_INTERIOR_FE)) {   /* centred DPDx/y/z */
            const NbrOff o(i, j, nx, ny);
            const double* P = a.P[2];
            /* half reciprocal spacings multiplied in: the same differences to
             * rounding; all three formed before R is known, so their loads
             * issue beside the rotation's */
            const double gx = (P[c + o.xp] - P[c + o.xm]) * (0.5 / dx[0]);
            const double gy = (P[c + o.yp] - P[c + o.ym]) * (0.5 / dx[1]);
            const double gz = (P[c + o.zp] - P[c + o.zm]) * (0.5 / dx[2]);
            if (R[6] != 0.0) x += R[6] * gx;
            if (R[7] != 0.0) x += R[7] * gy;
            x += R[8] * gz;
            rhs = -x;
        } else if (a.pxy_zero) {
            if (R[6] != 0.0) x += R[6] * dpd<0>(Fr, a.M, i, j, k, dx, a.pbc);
            if (R[7] != 0.0) x += R[7] * dpd<1>(Fr, a.M, i, j, k, dx, a.pbc);
            x += R[8] * dpd<2>(Fr, a.M, i, j, k, dx, a.pbc);
            rhs = -x;
        } else {
            const PView Fp{a.P[0], nx, ny, nz, a.g.per[0], a.g.per[1], a.g.per[2]};
            const PView Fq{a.P[1], nx, ny, nz, a.g.per[0], a.g.per[1], a.g.per[2]};
            rhs = -(R[0] * dpd<0>(Fp, a.M, i, j, k, dx, a.pbc) + R[1] * dpd<1>(Fp, a.M, i, j, k, dx, a.pbc) + R[2] * dpd<2>(Fp, a.M, i, j, k, dx, a.pbc))
                  - (R[3] * dpd<0>(Fq, a.M, i, j, k, dx, a.pbc) + R[4] * dpd<1>(Fq, a.M, i, j, k, dx, a.pbc) + R[5] * dpd<2>(Fq, a.M, i, j, k, dx, a.pbc))
                  - (R[6] * dpd<0>(Fr, a.M, i, j, k, dx, a.pbc) + R[7] * dpd<1>(Fr, a.M, i, j, k, dx, a.pbc) + R[8] * dpd<2>(Fr, a.M, i, j, k, dx, a.pbc));
        }
        al = 0.0;
    }
    a.rhs[c] = rhs;
    if (a.alpha) a.alpha[c] = al;   /* null: alpha is known to be 0 everywhere (no SC cells) */
}

/* rho(phi) in SC cells (0 elsewhere), optional L1 partial sums of
 * |phi - phi_prev| and |phi| with phi_prev <- phi.  grid: 64x4 tiles, z. */
__global__ __launch_bounds__(256) void k_rho_update(RhoArgs a) {
    __shared__ double sh[256 / 64 + 1];
    const int nx = a.n[0], ny = a.n[1];
    const int i = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int k = blockIdx.z;
    double s_num = 0.0, s_den = 0.0;
    if (i < nx && j < ny) {
        const int c = i + nx * (j + ny * k);
        const double ph = a.phi[c];
        if (a.M.code(i, j, k) >= 2) {
            double n, p;
            a.rho[c] = rho_of(a.sc, ph, n, p);
            a.e_den[c] = n;
            a.h_den[c] = p;
        } else {
            a.rho[c] = 0.0;
        }
        if (a.want_err) {
            s_num = fabs(ph - a.phi_prev[c]);
            s_den = fabs(ph);
            a.phi_prev[c] = ph;
        }
    }
    if (a.want_err) {
        s_num = block_sum<256>(s_num, sh);
        s_den = block_sum<256>(s_den, sh);
        const int nb = gridDim.x * gridDim.y * gridDim.z;
        const int bid = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
        if (threadIdx.x == 0) { a.part[bid] = s_num; a.part[nb + bid] = s_den; }
    }
}

/* ------------------------------ TDGL right-hand side ---------------------- */
/* dF_grad/dP_r with the rotated tensor, TotalEnergyDensity.cpp:209-231, from
 * the Hessian of P_r; the mixed entries are needed only where R couples them. */
__device__ __forceinline__ void mixed_needed(const double R[9], bool& nxy, bool& nyz, bool& nxz) {
    const double R11 = R[0], R12 = R[1], R13 = R[2], R21 = R[3], R22 = R[4], R23 = R[5], R31 = R[6], R32 = R[7], R33 = R[8];
    nxy = (R31 * R32 != 0.0) || (R11 * R12 != 0.0) || (R21 * R22 != 0.0);
    nyz = (R32 * R33 != 0.0) || (R12 * R13 != 0.0) || (R22 * R23 != 0.0);
    nxz = (R33 * R31 != 0.0) || (R13 * R11 != 0.0) || (R23 * R21 != 0.0);
}
__device__ __forceinline__ double grad_r_of(const double R[9], double Hxx, double Hyy, double Hzz, double Hxy,
                                            double Hyz, double Hxz, double g11, double g44m) {
    const double R11 = R[0], R12 = R[1], R13 = R[2], R21 = R[3], R22 = R[4], R23 = R[5], R31 = R[6], R32 = R[7], R33 = R[8];
    return -g11 * (R31 * R31 * Hxx + R32 * R32 * Hyy + R33 * R33 * Hzz + 2. * R31 * R32 * Hxy + 2. * R32 * R33 * Hyz + 2. * R33 * R31 * Hxz)
           - g44m * (R11 * R11 * Hxx + R12 * R12 * Hyy + R13 * R13 * Hzz + 2. * R11 * R12 * Hxy + 2. * R12 * R13 * Hyz + 2. * R13 * R11 * Hxz)
           - g44m * (R21 * R21 * Hxx + R22 * R22 * Hyy + R23 * R23 * Hzz + 2. * R21 * R22 * Hxy + 2. * R22 * R23 * Hyz + 2. * R23 * R21 * Hxz);
}
__device__ __forceinline__ double grad_r(const PView& Fr, const MView& M, int i, int j, int k, const double* dx,
                                         const PBC& bc, const double R[9], double g11, double g44m) {
    const double Hxx = ddpd<0>(Fr, M, i, j, k, dx, bc);
    const double Hyy = ddpd<1>(Fr, M, i, j, k, dx, bc);
    const double Hzz = ddpd<2>(Fr, M, i, j, k, dx, bc);
    bool nxy, nyz, nxz;
    mixed_needed(R, nxy, nyz, nxz);
    const double Hxy = nxy ? ddpd_mixed<0, 1>(Fr, M, i, j, k, dx, bc) : 0.0;
    const double Hyz = nyz ? ddpd_mixed<1, 2>(Fr, M, i, j, k, dx, bc) : 0.0;
    const double Hxz = nxz ? ddpd_mixed<0, 2>(Fr, M, i, j, k, dx, bc) : 0.0;
    return grad_r_of(R, Hxx, Hyy, Hzz, Hxy, Hyz, Hxz, g11, g44m);
}

/* The same Hessian for an interior FE cell: every masked stencil is the
 * centred one (DoubleDPDx/y/z and DoubleDPDxDy/DxDz/DyDz interior branches),
 * so the neighbours are read directly.  The spacings enter as the reciprocals
 * ih = 1/dx, multiplied in (the masked stencils divide): the same differences
 * to rounding, three FP64 divisions per cell instead of fifteen. */
template <class Acc>
__device__ __forceinline__ double grad_r_interior_acc(const Acc& F, const double* ih, const double R[9], double g11,
                                                      double g44m) {
    const double pc = F(0, 0, 0);
    const double Hxx = (F(1, 0, 0) - 2. * pc + F(-1, 0, 0)) * (ih[0] * ih[0]);
    const double Hyy = (F(0, 1, 0) - 2. * pc + F(0, -1, 0)) * (ih[1] * ih[1]);
    const double Hzz = (F(0, 0, 1) - 2. * pc + F(0, 0, -1)) * (ih[2] * ih[2]);
    /* the mixed entries are formed whether or not R couples them (a zero
     * product of R entries then multiplies a finite value), so every load
     * can be issued before R is known */
    const double Hxy = ((F(1, 1, 0) - F(1, -1, 0)) - (F(-1, 1, 0) - F(-1, -1, 0))) * (0.25 * ih[0] * ih[1]);
    const double Hyz = ((F(0, 1, 1) - F(0, 1, -1)) - (F(0, -1, 1) - F(0, -1, -1))) * (0.25 * ih[1] * ih[2]);
    const double Hxz = ((F(1, 0, 1) - F(1, 0, -1)) - (F(-1, 0, 1) - F(-1, 0, -1))) * (0.25 * ih[0] * ih[2]);
    return grad_r_of(R, Hxx, Hyy, Hzz, Hxy, Hyz, Hxz, g11, g44m);
}

/* P of an interior cell's neighbourhood from global memory (periodic wrap
 * laterally through the offsets). */
struct GlobalNbr {
    const double* __restrict__ P;
    int c;
    NbrOff o;
    __device__ __forceinline__ double operator()(int di, int dj, int dk) const {
        return P[c + (di > 0 ? o.xp : di < 0 ? o.xm : 0) + (dj > 0 ? o.yp : dj < 0 ? o.ym : 0) +
                 (dk > 0 ? o.zp : dk < 0 ? o.zm : 0)];
    }
};

__device__ __forceinline__ double grad_r_interior(const double* __restrict__ P, int c, const NbrOff& o,
                                                  const double* ih, const double R[9], double g11, double g44m) {
    return grad_r_interior_acc(GlobalNbr{P, c, o}, ih, R, g11, g44m);
}

/* The scalar TDGL update of one FE cell (SCALAR_FAST: is_polarization_scalar
 * with P_p = P_q = 0): GL_r and the new P_r.  Interior cells take the Hessian
 * of P from F (neighbour offsets -1..1) and grad(phi) from the centred
 * differences; the others the masked stencils on global memory. */
template <class Acc>
__device__ __forceinline__ double tdgl_scalar_cell(const TdglArgs& a, int i, int j, int k, int c, int raw,
                                                   const Acc& F, double& g_out) {
    const int nx = a.g.n[0], ny = a.g.n[1], nz = a.g.n[2];
    const bool tph = (raw & 4) != 0;
    const double Gam = a.BigGamma;
    const double* dx = a.g.dx;
    double R[9];
    rot_of(a.rot, c, R);
    const bool inner = (raw & MAT_INTERIOR_FE) != 0;
    const NbrOff o(i, j, nx, ny);
    double ih[3] = {0.0, 0.0, 0.0};
    if (inner) { ih[0] = 1.0 / dx[0]; ih[1] = 1.0 / dx[1]; ih[2] = 1.0 / dx[2]; }
    double E2 = 0.0;
    if (a.incE == 1) {
        double fx, fy, fz;
        if (inner) {
            const double* ph = a.phv.p;
            fx = (ph[c + o.xp] - ph[c + o.xm]) * (0.5 * ih[0]);
            fy = (ph[c + o.yp] - ph[c + o.ym]) * (0.5 * ih[1]);
            fz = (ph[c + o.zp] - ph[c + o.zm]) * (0.5 * ih[2]);
        } else {
            grad_phi(a.phv, a.g, i, j, k, fx, fy, fz);
        }
        double s = 0.0;
        if (R[6] != 0.0) s += R[6] * fx;
        if (R[7] != 0.0) s += R[7] * fy;
        s += R[8] * fz;
        E2 = -s;
    }
    const double r = F(0, 0, 0);
    double L = 0.0, G = 0.0, El = 0.0;
    if (a.incL == 1) {
        const double r2 = r * r;
        const double alpha = a.rot.idx ? a.rot.atab[a.rot.idx[c]] : a.alpha;
        const double dF = alpha * r + a.beta * (r2 * r) + a.gamma * (r2 * r2 * r);
        L = -1.0 * Gam * dF;
    }
    if (a.incG == 1) {
        if (inner) {
            G = -1.0 * Gam * grad_r_interior_acc(F, ih, R, a.g11, a.g44 - a.g44_p);
        } else {
            const PView Fr{a.P[2], nx, ny, nz, a.g.per[0], a.g.per[1], a.g.per[2]};
            G = -1.0 * Gam * grad_r(Fr, a.M, i, j, k, dx, a.pbc, R, a.g11, a.g44 - a.g44_p);
        }
    }
    if (a.incE == 1) El = -1.0 * Gam * (-E2);
    if (tph) { L = 0.0; G = 0.0; El = 0.0; }
    double g = 0.0;
    if (a.incL == 1) g += L;
    if (a.incG == 1) g += G;
    if (a.incE == 1) g += El;
    g_out = g;
    return r + a.dt * g;
}


/* SCALAR_FAST: is_polarization_scalar with P_p = P_q = 0 everywhere (only P_r
 * is read or written).  Non-FE cells have Gamma = 0, so P never changes there;
 * both P buffers hold the same values in those cells and they are skipped. */
template <bool SCALAR_FAST>
__global__ __launch_bounds__(256) void k_tdgl(TdglArgs a) {
    const int nx = a.g.n[0], ny = a.g.n[1], nz = a.g.n[2];
    const int i = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int k = a.kz0 + blockIdx.z;
    if (i >= nx || j >= ny || k >= nz) return;
    const int c = i + nx * (j + ny * k);
    const int raw = a.M.raw(i, j, k);
    if ((raw & 3) != 0) {
        if (a.GLout[2]) {
            if (!SCALAR_FAST) { a.GLout[0][c] = 0.0; a.GLout[1][c] = 0.0; }
            a.GLout[2][c] = 0.0;
        }
        return;
    }
    if (SCALAR_FAST) {
        double g;
        const double rn = tdgl_scalar_cell(a, i, j, k, c, raw, GlobalNbr{a.P[2], c, NbrOff(i, j, nx, ny)}, g);
        if (a.GLout[2]) a.GLout[2][c] = g;
        if (a.Pout[2]) a.Pout[2][c] = rn;
        return;
    }
    const bool tph = (raw & 4) != 0;
    const double Gam = a.BigGamma;
    const double* dx = a.g.dx;
    double R[9];
    rot_of(a.rot, c, R);
    const PView Fr{a.P[2], nx, ny, nz, a.g.per[0], a.g.per[1], a.g.per[2]};
    double GL[3] = {0.0, 0.0, 0.0};
    /* electric field E = -R grad(phi) */
    double E[3] = {0.0, 0.0, 0.0};
    if (a.incE == 1) {
        double fx, fy, fz;
        grad_phi(a.phv, a.g, i, j, k, fx, fy, fz);
        E[0] = -(R[0] * fx + R[1] * fy + R[2] * fz);
        E[1] = -(R[3] * fx + R[4] * fy + R[5] * fz);
        E[2] = -(R[6] * fx + R[7] * fy + R[8] * fz);
    }

    const PView Fp{a.P[0], nx, ny, nz, a.g.per[0], a.g.per[1], a.g.per[2]};
    const PView Fq{a.P[1], nx, ny, nz, a.g.per[0], a.g.per[1], a.g.per[2]};
    const double p = a.P[0][c], q = a.P[1][c], r = a.P[2][c];
    double L[3] = {0, 0, 0}, G[3] = {0, 0, 0}, El[3] = {0, 0, 0};
    if (a.incL == 1) {
        const double p2 = p * p, q2 = q * q, r2 = r * r;
        const double a12 = a.a12, a112 = a.a112, a123 = a.a123;
        const double al = a.rot.idx ? a.rot.atab[a.rot.idx[c]] : a.alpha;
        const double dFp = al * p + a.beta * (p2 * p) + a.gamma * (p2 * p2 * p)
                         + 2. * a12 * p * q2 + 2. * a12 * p * r2
                         + 4. * a112 * (p2 * p) * (q2 + r2)
                         + 2. * a112 * p * (q2 * q2) + 2. * a112 * p * (r2 * r2)
                         + 2. * a123 * p * q2 * r2;
        const double dFq = al * q + a.beta * (q2 * q) + a.gamma * (q2 * q2 * q)
                         + 2. * a12 * q * p2 + 2. * a12 * q * r2
                         + 4. * a112 * (q2 * q) * (p2 + r2)
                         + 2. * a112 * q * (p2 * p2) + 2. * a112 * q * (r2 * r2)
                         + 2. * a123 * q * p2 * r2;
        const double dFr = al * r + a.beta * (r2 * r) + a.gamma * (r2 * r2 * r)
                         + 2. * a12 * r * p2 + 2. * a12 * r * q2
                         + 4. * a112 * (r2 * r) * (p2 + q2)
                         + 2. * a112 * r * (p2 * p2) + 2. * a112 * r * (q2 * q2)
                         + 2. * a123 * r * p2 * q2;
        L[0] = -1.0 * Gam * dFp; L[1] = -1.0 * Gam * dFq; L[2] = -1.0 * Gam * dFr;
        if (a.scalar) { L[0] = 0.0; L[1] = 0.0; }
        if (tph) { L[0] = L[1] = L[2] = 0.0; }
    }
    if (a.incG == 1) {
        const double g11 = a.g11, g44 = a.g44, g12 = a.g12, g44_p = a.g44_p;
        double dFp = 0.0, dFq = 0.0;
        if (!a.scalar) {
            dFp = -g11 * ddpd<0>(Fp, a.M, i, j, k, dx, a.pbc)
                  - (g44 + g44_p) * ddpd<1>(Fp, a.M, i, j, k, dx, a.pbc)
                  - (g44 + g44_p) * ddpd<2>(Fp, a.M, i, j, k, dx, a.pbc)
                  - (g12 + g44 - g44_p) * ddpd_mixed<0, 1>(Fq, a.M, i, j, k, dx, a.pbc)
                  - (g12 + g44 - g44_p) * ddpd_mixed<0, 2>(Fr, a.M, i, j, k, dx, a.pbc);
            dFq = -g11 * ddpd<1>(Fq, a.M, i, j, k, dx, a.pbc)
                  - (g44 - g44_p) * ddpd<0>(Fq, a.M, i, j, k, dx, a.pbc)
                  - (g44 - g44_p) * ddpd<2>(Fq, a.M, i, j, k, dx, a.pbc)
                  - (g12 + g44 + g44_p) * ddpd_mixed<0, 1>(Fp, a.M, i, j, k, dx, a.pbc)
                  - (g12 + g44 - g44_p) * ddpd_mixed<1, 2>(Fr, a.M, i, j, k, dx, a.pbc);
        }
        double dFr = grad_r(Fr, a.M, i, j, k, dx, a.pbc, R, g11, g44 - g44_p);
        dFr = dFr - (g44 + g44_p + g12) * ddpd_mixed<1, 2>(Fq, a.M, i, j, k, dx, a.pbc)
                  - (g44 + g44_p + g12) * ddpd_mixed<0, 2>(Fp, a.M, i, j, k, dx, a.pbc);
        G[0] = -1.0 * Gam * dFp; G[1] = -1.0 * Gam * dFq; G[2] = -1.0 * Gam * dFr;
        if (a.scalar) { G[0] = 0.0; G[1] = 0.0; }
        if (tph) { G[0] = G[1] = G[2] = 0.0; }
    }
    if (a.incE == 1) {
        El[0] = -1.0 * Gam * (-E[0]); El[1] = -1.0 * Gam * (-E[1]); El[2] = -1.0 * Gam * (-E[2]);
        if (a.scalar) { El[0] = 0.0; El[1] = 0.0; }
        if (tph) { El[0] = El[1] = El[2] = 0.0; }
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        double g = 0.0;
        if (a.incL == 1) g += L[d];
        if (a.incG == 1) g += G[d];
        if (a.incE == 1) g += El[d];
        GL[d] = g;
        if (a.GLout[d]) a.GLout[d][c] = g;
    }
    if (a.Pout[0]) {
        a.Pout[0][c] = p + a.dt * GL[0];
        a.Pout[1][c] = q + a.dt * GL[1];
        a.Pout[2][c] = r + a.dt * GL[2];
    }
}

/* ------------- TDGL update fused with the next step's Poisson RHS ---------- */
/* Scalar P (P_p = P_q = 0), forward Euler, no semiconductor cells: the
 * Poisson right-hand side of the next step (ComputePoissonRHS,
 * ElectrostaticSolver.cpp:8-83: -R grad P_new in FE cells, 0 elsewhere)
 * depends on P_new only, so it is formed by the workgroup that computes
 * P_new, from a window of P_new in LDS: P is read once and written once per
 * step, and the separate RHS pass (which reads P again) is gone.
 *
 * Three launches.  FE cells next to a material face or a domain face (a
 * list, built with the material ids) take the masked stencils of
 * k_tdgl<true> on global memory first (k_tdgl_face_cells).  Then
 * k_tdgl_rhs_stream: a workgroup owns a TRS_TX x TRS_TY column tile and a
 * chunk of TRS_Z planes of [k0, k1] (the planes holding FE cells; elsewhere P
 * never changes and the RHS stays 0) and marches upward.  Old P is staged
 * plane by plane in LDS with a 2-cell lateral halo (three planes, the next
 * one loaded into registers while the current one is computed); P_new is
 * formed on the tile and a 1-cell ring (one thread per cell of tile and
 * ring; the ring is also formed by the neighbouring workgroups) for the
 * interior FE cells, taken from the first launch's output for the face
 * cells and from P elsewhere, and kept for three planes; the RHS of plane
 * p-1 (interior FE cells, 0 in the others) is formed once P_new of planes
 * p-2 .. p is in LDS.  Last, the face cells' RHS with the masked stencils
 * on P_new (k_rhs_face_cells).  The arithmetic is that of k_tdgl<true> and
 * k_rhs_newton. */
constexpr int TRS_TX = 30, TRS_TY = 14, TRS_Z = 16;
constexpr int TRS_HX = TRS_TX + 4, TRS_HY = TRS_TY + 4;     /* old P: tile + 2-cell halo */
constexpr int TRS_RX = TRS_TX + 2, TRS_RY = TRS_TY + 2;     /* P_new: tile + 1-cell ring */
constexpr int TRS_THR = TRS_RX * TRS_RY;                     /* one thread per cell of tile and ring: 512 */
constexpr int TRS_PV = (TRS_HX * TRS_HY + TRS_THR - 1) / TRS_THR;   /* old-P values per thread and plane */

__device__ __forceinline__ int slot3(int q) { return ((q % 3) + 3) % 3; }

/* old P around window position o of plane slots (k-1, k, k+1) */
struct LdsNbr {
    const double* w0;
    const double* w1;
    const double* w2;
    int o;
    __device__ __forceinline__ double operator()(int di, int dj, int dk) const {
        const double* w = dk < 0 ? w0 : dk > 0 ? w2 : w1;
        return w[o + di + dj * TRS_HX];
    }
};

__global__ __launch_bounds__(256) void k_tdgl_face_cells(TdglRhsArgs A) {
    const TdglArgs& a = A.t;
    const int q = blockIdx.x * 256 + threadIdx.x;
    if (q >= A.nface) return;
    const int nx = a.g.n[0], ny = a.g.n[1];
    const int c = A.face[q];
    const int i = c % nx, j = (c / nx) % ny, k = c / (nx * ny);
    double g;
    a.Pout[2][c] = tdgl_scalar_cell(a, i, j, k, c, a.M.raw(i, j, k), GlobalNbr{a.P[2], c, NbrOff(i, j, nx, ny)}, g);
}

__global__ __launch_bounds__(256) void k_rhs_face_cells(TdglRhsArgs A) {
    const TdglArgs& a = A.t;
    const int q = blockIdx.x * 256 + threadIdx.x;
    if (q >= A.nface) return;
    const int nx = a.g.n[0], ny = a.g.n[1], nz = a.g.n[2];
    const int c = A.face[q];
    const int i = c % nx, j = (c / nx) % ny, k = c / (nx * ny);
    double R[9];
    rot_of(a.rot, c, R);
    const PView Fr{a.Pout[2], nx, ny, nz, a.g.per[0], a.g.per[1], a.g.per[2]};
    const double* dx = a.g.dx;
    double x = 0.0;
    if (R[6] != 0.0) x += R[6] * dpd<0>(Fr, a.M, i, j, k, dx, a.pbc);
    if (R[7] != 0.0) x += R[7] * dpd<1>(Fr, a.M, i, j, k, dx, a.pbc);
    x += R[8] * dpd<2>(Fr, a.M, i, j, k, dx, a.pbc);
    A.rhs[c] = -x;
}

__global__ __launch_bounds__(TRS_THR, 2) void k_tdgl_rhs_stream(TdglRhsArgs A) {
    __shared__ double Pw[3 * TRS_HX * TRS_HY];
    __shared__ double Nw[3 * TRS_RX * TRS_RY];
    const TdglArgs& a = A.t;
    const int nx = a.g.n[0], ny = a.g.n[1], nz = a.g.n[2];
    const int per0 = a.g.per[0], per1 = a.g.per[1];
    const int i0 = blockIdx.x * TRS_TX, j0 = blockIdx.y * TRS_TY;
    const int ks = A.k0 + blockIdx.z * TRS_Z;
    const int ke = min(ks + TRS_Z - 1, A.k1);
    const double* __restrict__ P = a.P[2];
    const double* __restrict__ Pn = a.Pout[2];
    /* global index of a window column/row, -1 across a non-periodic face */
    auto wrap = [](int v, int n, int per) -> int {
        if (v < 0) return per ? v + n : -1;
        if (v >= n) return per ? v - n : -1;
        return v;
    };
    /* this thread's share of a halo plane of old P */
    long long src[TRS_PV];
#pragma unroll
    for (int r = 0; r < TRS_PV; ++r) {
        const int t = threadIdx.x + r * TRS_THR;
        src[r] = -1;
        if (t < TRS_HX * TRS_HY) {
            const int v = t / TRS_HX, u = t - v * TRS_HX;
            const int i = wrap(i0 + u - 2, nx, per0), j = wrap(j0 + v - 2, ny, per1);
            if (i >= 0 && j >= 0) src[r] = i + (long long)nx * j;
        }
    }
    const long long plane = (long long)nx * ny;
    auto fetch = [&](int q, double (&vals)[TRS_PV]) {
#pragma unroll
        for (int r = 0; r < TRS_PV; ++r)
            vals[r] = (q >= 0 && q < nz && src[r] >= 0) ? P[src[r] + plane * q] : 0.0;
    };
    auto stash = [&](int q, const double (&vals)[TRS_PV]) {
        double* dst = Pw + slot3(q) * (TRS_HX * TRS_HY);
#pragma unroll
        for (int r = 0; r < TRS_PV; ++r) {
            const int t = threadIdx.x + r * TRS_THR;
            if (t < TRS_HX * TRS_HY) dst[t] = vals[r];
        }
    };
    const int tu = threadIdx.x % TRS_RX - 1, tv = threadIdx.x / TRS_RX - 1;
    const int ci = wrap(i0 + tu, nx, per0), cj = wrap(j0 + tv, ny, per1);
    const bool in_tile = tu >= 0 && tu < TRS_TX && tv >= 0 && tv < TRS_TY && i0 + tu < nx && j0 + tv < ny;
    const int wo = (tv + 2) * TRS_HX + (tu + 2);
    {
        double v0[TRS_PV], v1[TRS_PV], v2[TRS_PV];
        fetch(ks - 2, v0); fetch(ks - 1, v1); fetch(ks, v2);
        stash(ks - 2, v0); stash(ks - 1, v1); stash(ks, v2);
    }
    __syncthreads();
    for (int p = ks - 1; p <= ke + 1; ++p) {
        double nxt[TRS_PV];
        fetch(p + 2, nxt);                             /* in flight while plane p is formed */
        /* P_new of plane p at this thread's tile or ring position */
        double pn = 0.0;
        if (p >= 0 && p < nz && ci >= 0 && cj >= 0) {
            const int c = ci + nx * (cj + ny * p);
            const int raw = a.M.raw(ci, cj, p);
            const double* w1 = Pw + slot3(p) * (TRS_HX * TRS_HY);
            if ((raw & 3) != 0) {
                pn = w1[wo];                           /* Gamma = 0: P does not change */
            } else if (!(raw & MAT_INTERIOR_FE)) {
                pn = Pn[c];                            /* formed by k_tdgl_face_cells */
            } else {
                const LdsNbr F{Pw + slot3(p - 1) * (TRS_HX * TRS_HY), w1, Pw + slot3(p + 1) * (TRS_HX * TRS_HY), wo};
                double g;
                pn = tdgl_scalar_cell(a, ci, cj, p, c, raw, F, g);
                if (in_tile && p >= ks && p <= ke) a.Pout[2][c] = pn;
            }
        }
        Nw[slot3(p) * (TRS_RX * TRS_RY) + threadIdx.x] = pn;
        __syncthreads();
        stash(p + 2, nxt);                             /* over plane p-1, no longer read */
        /* Poisson RHS of plane p - 1 from P_new of planes p-2 .. p */
        const int q = p - 1;
        if (in_tile && q >= ks && q <= ke) {
            const int c = ci + nx * (cj + ny * q);
            const int raw = a.M.raw(ci, cj, q);
            if ((raw & 3) != 0) {
                A.rhs[c] = 0.0;
            } else if (raw & MAT_INTERIOR_FE) {
                double R[9];
                rot_of(a.rot, c, R);
                const double* dx = a.g.dx;
                const int o = (tv + 1) * TRS_RX + (tu + 1);
                const double* n0 = Nw + slot3(q - 1) * (TRS_RX * TRS_RY);
                const double* n1 = Nw + slot3(q) * (TRS_RX * TRS_RY);
                const double* n2 = Nw + slot3(q + 1) * (TRS_RX * TRS_RY);
                const double gx = (n1[o + 1] - n1[o - 1]) * (0.5 / dx[0]);
                const double gy = (n1[o + TRS_RX] - n1[o - TRS_RX]) * (0.5 / dx[1]);
                const double gz = (n2[o] - n0[o]) * (0.5 / dx[2]);
                double x = 0.0;
                if (R[6] != 0.0) x += R[6] * gx;
                if (R[7] != 0.0) x += R[7] * gy;
                x += R[8] * gz;
                A.rhs[c] = -x;
            }
        }
        __syncthreads();
    }
}

/* ----------------- TDGL update streamed through LDS (scalar P) ------------ */
/* k_tdgl<true> for the interior FE cells, with the stencil operands from LDS
 * instead of per-cell gathers: a workgroup owns a TDS_TX x TDS_TY column tile
 * and a chunk of TDS_Z planes of [k0, k1] and marches upward, P and phi of
 * three planes (tile + 1-cell halo) in LDS, the plane after next loaded into
 * registers while the current one is updated.  FE cells next to a face
 * (the list of k_tdgl_face_cells) are updated by that kernel; other cells
 * keep their P (Gamma = 0). */
constexpr int TDS_TX = 64, TDS_TY = 8, TDS_Z = 16;
constexpr int TDS_HX = TDS_TX + 2, TDS_HY = TDS_TY + 2;
constexpr int TDS_THR = TDS_TX * TDS_TY;                                  /* 512 */
constexpr int TDS_PV = (TDS_HX * TDS_HY + TDS_THR - 1) / TDS_THR;       /* halo-plane values per thread */
constexpr int TDS_MAXROWS = 512;   /* rotation / alpha tables up to this many rows are staged in LDS */

/* old P (or phi) around window position o of plane slots (k-1, k, k+1) */
struct WinNbr {
    const double* w0;
    const double* w1;
    const double* w2;
    int o;
    __device__ __forceinline__ double operator()(int di, int dj, int dk) const {
        const double* w = dk < 0 ? w0 : dk > 0 ? w2 : w1;
        return w[o + di + dj * TDS_HX];
    }
};

template <bool TAB>
__global__ __launch_bounds__(TDS_THR, 2) void k_tdgl_stream(TdglRhsArgs A) {
    __shared__ double Pw[3 * TDS_HX * TDS_HY];
    __shared__ double Fw[3 * TDS_HX * TDS_HY];
    extern __shared__ __align__(16) unsigned char tds_dyn[];
    double* const Rt = reinterpret_cast<double*>(tds_dyn);   /* TAB: 9 R values, then alpha, per row */
    const TdglArgs& a = A.t;
    if (TAB) {
        const int nr = a.rot.nrows;
        for (int t = threadIdx.x; t < 10 * nr; t += TDS_THR)
            Rt[t] = t < 9 * nr ? a.rot.tab[t] : a.rot.atab[t - 9 * nr];
    }
    const int nx = a.g.n[0], ny = a.g.n[1], nz = a.g.n[2];
    const int per0 = a.g.per[0], per1 = a.g.per[1];
    const int i0 = blockIdx.x * TDS_TX, j0 = blockIdx.y * TDS_TY;
    const int ks = A.k0 + blockIdx.z * TDS_Z;
    const int ke = min(ks + TDS_Z - 1, A.k1);
    const double* __restrict__ P = a.P[2];
    const double* __restrict__ ph = a.phv.p;
    auto wrap = [](int v, int n, int per) -> int {
        if (v < 0) return per ? v + n : -1;
        if (v >= n) return per ? v - n : -1;
        return v;
    };
    long long src[TDS_PV];
#pragma unroll
    for (int r = 0; r < TDS_PV; ++r) {
        const int t = threadIdx.x + r * TDS_THR;
        src[r] = -1;
        if (t < TDS_HX * TDS_HY) {
            const int v = t / TDS_HX, u = t - v * TDS_HX;
            const int i = wrap(i0 + u - 1, nx, per0), j = wrap(j0 + v - 1, ny, per1);
            if (i >= 0 && j >= 0) src[r] = i + (long long)nx * j;
        }
    }
    const long long plane = (long long)nx * ny;
    /* halo planes outside the domain (or across a non-periodic side) read 0;
     * interior FE cells never read them */
    auto fetch = [&](int q, double (&vp)[TDS_PV], double (&vf)[TDS_PV]) {
#pragma unroll
        for (int r = 0; r < TDS_PV; ++r) {
            const bool ok = q >= 0 && q < nz && src[r] >= 0;
            vp[r] = ok ? P[src[r] + plane * q] : 0.0;
            vf[r] = ok ? ph[src[r] + plane * q] : 0.0;
        }
    };
    auto stash = [&](int q, const double (&vp)[TDS_PV], const double (&vf)[TDS_PV]) {
        double* dp = Pw + slot3(q) * (TDS_HX * TDS_HY);
        double* df = Fw + slot3(q) * (TDS_HX * TDS_HY);
#pragma unroll
        for (int r = 0; r < TDS_PV; ++r) {
            const int t = threadIdx.x + r * TDS_THR;
            if (t < TDS_HX * TDS_HY) { dp[t] = vp[r]; df[t] = vf[r]; }
        }
    };
    const int tu = threadIdx.x % TDS_TX, tv = threadIdx.x / TDS_TX;
    const int ci = i0 + tu, cj = j0 + tv;
    const bool in_dom = ci < nx && cj < ny;
    const int wo = (tv + 1) * TDS_HX + (tu + 1);
    {
        double vp[TDS_PV], vf[TDS_PV];
        fetch(ks - 1, vp, vf); stash(ks - 1, vp, vf);
        fetch(ks, vp, vf); stash(ks, vp, vf);
        fetch(ks + 1, vp, vf); stash(ks + 1, vp, vf);
    }
    __syncthreads();
    const double Gam = a.BigGamma;
    const double* dx = a.g.dx;
    const double ih[3] = {1.0 / dx[0], 1.0 / dx[1], 1.0 / dx[2]};
    /* material id and table row of this thread's cell, one plane ahead */
    const int* __restrict__ ridx = a.rot.idx;
    int raw_c = 0, row_c = 0;
    if (in_dom) {
        raw_c = a.M.raw(ci, cj, ks);
        if (ridx) row_c = ridx[ci + nx * (cj + ny * ks)];
    }
    for (int p = ks; p <= ke; ++p) {
        double np_[TDS_PV], nf[TDS_PV];
        fetch(p + 2, np_, nf);                         /* in flight while plane p is updated */
        int raw_n = 0, row_n = 0;
        if (in_dom && p < ke) {
            raw_n = a.M.raw(ci, cj, p + 1);
            if (ridx) row_n = ridx[ci + nx * (cj + ny * (p + 1))];
        }
        if (in_dom) {
            const int c = ci + nx * (cj + ny * p);
            const int raw = raw_c;
            if ((raw & 3) == 0 && (raw & MAT_INTERIOR_FE)) {
                const double* p0 = Pw + slot3(p - 1) * (TDS_HX * TDS_HY);
                const double* p1 = Pw + slot3(p) * (TDS_HX * TDS_HY);
                const double* p2 = Pw + slot3(p + 1) * (TDS_HX * TDS_HY);
                const double* f0 = Fw + slot3(p - 1) * (TDS_HX * TDS_HY);
                const double* f1 = Fw + slot3(p) * (TDS_HX * TDS_HY);
                const double* f2 = Fw + slot3(p + 1) * (TDS_HX * TDS_HY);
                const WinNbr F{p0, p1, p2, wo};
                double R[9];
                double alpha_c = a.alpha;
                if (TAB) {
#pragma unroll
                    for (int q = 0; q < 9; ++q) R[q] = Rt[9 * row_c + q];
                    alpha_c = Rt[9 * a.rot.nrows + row_c];
                } else if (ridx) {
                    const double* t = a.rot.tab + 9 * row_c;
#pragma unroll
                    for (int q = 0; q < 9; ++q) R[q] = t[q];
                    alpha_c = a.rot.atab[row_c];
                } else {
#pragma unroll
                    for (int q = 0; q < 9; ++q) R[q] = a.rot.Ru[q];
                }
                const bool tph = (raw & 4) != 0;
                double E2 = 0.0;
                if (a.incE == 1) {
                    const double fx = (f1[wo + 1] - f1[wo - 1]) * (0.5 * ih[0]);
                    const double fy = (f1[wo + TDS_HX] - f1[wo - TDS_HX]) * (0.5 * ih[1]);
                    const double fz = (f2[wo] - f0[wo]) * (0.5 * ih[2]);
                    double sE = 0.0;
                    if (R[6] != 0.0) sE += R[6] * fx;
                    if (R[7] != 0.0) sE += R[7] * fy;
                    sE += R[8] * fz;
                    E2 = -sE;
                }
                const double r = p1[wo];
                double L = 0.0, G = 0.0, El = 0.0;
                if (a.incL == 1) {
                    const double r2 = r * r;
                    const double dF = alpha_c * r + a.beta * (r2 * r) + a.gamma * (r2 * r2 * r);
                    L = -1.0 * Gam * dF;
                }
                if (a.incG == 1) G = -1.0 * Gam * grad_r_interior_acc(F, ih, R, a.g11, a.g44 - a.g44_p);
                if (a.incE == 1) El = -1.0 * Gam * (-E2);
                if (tph) { L = 0.0; G = 0.0; El = 0.0; }
                double g = 0.0;
                if (a.incL == 1) g += L;
                if (a.incG == 1) g += G;
                if (a.incE == 1) g += El;
                a.Pout[2][c] = r + a.dt * g;
            }
        }
        raw_c = raw_n;
        row_c = row_n;
        __syncthreads();
        stash(p + 2, np_, nf);                         /* over plane p-1, no longer read */
        __syncthreads();
    }
}

/* ------------------- Poisson RHS streamed through LDS (scalar P) ----------- */
/* k_rhs_newton's centred branch for the interior FE cells of planes
 * [k0, k1] (no SC cells, the RHS buffer already 0 in every non-FE cell, the
 * Newton coefficient not written): P of three planes (tile + 1-cell halo)
 * in LDS, the plane after next in registers, R_3 from an LDS copy of the
 * rotation table; FE cells next to a face by list (k_rhs_face_list). */
template <bool TAB>
__global__ __launch_bounds__(TDS_THR, 2) void k_rhs_stream(RhsArgs a, int k0, int k1) {
    __shared__ double Pw[3 * TDS_HX * TDS_HY];
    extern __shared__ __align__(16) unsigned char rs_dyn[];
    double* const Rt = reinterpret_cast<double*>(rs_dyn);   /* TAB: R_31, R_32, R_33 per row */
    if (TAB) {
        const int nr = a.rot.nrows;
        for (int t = threadIdx.x; t < 3 * nr; t += TDS_THR) Rt[t] = a.rot.tab[9 * (t / 3) + 6 + t % 3];
    }
    const int nx = a.g.n[0], ny = a.g.n[1], nz = a.g.n[2];
    const int per0 = a.g.per[0], per1 = a.g.per[1];
    const int i0 = blockIdx.x * TDS_TX, j0 = blockIdx.y * TDS_TY;
    const int ks = k0 + blockIdx.z * TDS_Z;
    const int ke = min(ks + TDS_Z - 1, k1);
    const double* __restrict__ P = a.P[2];
    auto wrap = [](int v, int n, int per) -> int {
        if (v < 0) return per ? v + n : -1;
        if (v >= n) return per ? v - n : -1;
        return v;
    };
    long long src[TDS_PV];
#pragma unroll
    for (int r = 0; r < TDS_PV; ++r) {
        const int t = threadIdx.x + r * TDS_THR;
        src[r] = -1;
        if (t < TDS_HX * TDS_HY) {
            const int v = t / TDS_HX, u = t - v * TDS_HX;
            const int i = wrap(i0 + u - 1, nx, per0), j = wrap(j0 + v - 1, ny, per1);
            if (i >= 0 && j >= 0) src[r] = i + (long long)nx * j;
        }
    }
    const long long plane = (long long)nx * ny;
    auto fetch = [&](int q, double (&vp)[TDS_PV]) {
#pragma unroll
        for (int r = 0; r < TDS_PV; ++r) vp[r] = (q >= 0 && q < nz && src[r] >= 0) ? P[src[r] + plane * q] : 0.0;
    };
    auto stash = [&](int q, const double (&vp)[TDS_PV]) {
        double* dp = Pw + slot3(q) * (TDS_HX * TDS_HY);
#pragma unroll
        for (int r = 0; r < TDS_PV; ++r) {
            const int t = threadIdx.x + r * TDS_THR;
            if (t < TDS_HX * TDS_HY) dp[t] = vp[r];
        }
    };
    const int tu = threadIdx.x % TDS_TX, tv = threadIdx.x / TDS_TX;
    const int ci = i0 + tu, cj = j0 + tv;
    const bool in_dom = ci < nx && cj < ny;
    const int wo = (tv + 1) * TDS_HX + (tu + 1);
    {
        double vp[TDS_PV];
        fetch(ks - 1, vp); stash(ks - 1, vp);
        fetch(ks, vp); stash(ks, vp);
        fetch(ks + 1, vp); stash(ks + 1, vp);
    }
    __syncthreads();
    const double* dx = a.g.dx;
    const int* __restrict__ ridx = a.rot.idx;
    int raw_c = 0, row_c = 0;
    if (in_dom) {
        raw_c = a.M.raw(ci, cj, ks);
        if (ridx) row_c = ridx[ci + nx * (cj + ny * ks)];
    }
    for (int p = ks; p <= ke; ++p) {
        double np_[TDS_PV];
        fetch(p + 2, np_);
        int raw_n = 0, row_n = 0;
        if (in_dom && p < ke) {
            raw_n = a.M.raw(ci, cj, p + 1);
            if (ridx) row_n = ridx[ci + nx * (cj + ny * (p + 1))];
        }
        if (in_dom && (raw_c & 3) == 0 && (raw_c & MAT_INTERIOR_FE)) {
            const int c = ci + nx * (cj + ny * p);
            double R6, R7, R8;
            if (TAB) { R6 = Rt[3 * row_c]; R7 = Rt[3 * row_c + 1]; R8 = Rt[3 * row_c + 2]; }
            else if (ridx) { const double* t = a.rot.tab + 9 * row_c; R6 = t[6]; R7 = t[7]; R8 = t[8]; }
            else { R6 = a.rot.Ru[6]; R7 = a.rot.Ru[7]; R8 = a.rot.Ru[8]; }
            const double* p0 = Pw + slot3(p - 1) * (TDS_HX * TDS_HY);
            const double* p1 = Pw + slot3(p) * (TDS_HX * TDS_HY);
            const double* p2 = Pw + slot3(p + 1) * (TDS_HX * TDS_HY);
            const double gx = (p1[wo + 1] - p1[wo - 1]) * (0.5 / dx[0]);
            const double gy = (p1[wo + TDS_HX] - p1[wo - TDS_HX]) * (0.5 / dx[1]);
            const double gz = (p2[wo] - p0[wo]) * (0.5 / dx[2]);
            double x = 0.0;
            if (R6 != 0.0) x += R6 * gx;
            if (R7 != 0.0) x += R7 * gy;
            x += R8 * gz;
            a.rhs[c] = -x;
        }
        raw_c = raw_n;
        row_c = row_n;
        __syncthreads();
        stash(p + 2, np_);
        __syncthreads();
    }
}

/* the masked branch of k_rhs_newton (scalar P) for the FE cells of a list */
__global__ __launch_bounds__(256) void k_rhs_face_list(RhsArgs a, const int* __restrict__ face, int nface) {
    const int q = blockIdx.x * 256 + threadIdx.x;
    if (q >= nface) return;
    const int nx = a.g.n[0], ny = a.g.n[1], nz = a.g.n[2];
    const int c = face[q];
    const int i = c % nx, j = (c / nx) % ny, k = c / (nx * ny);
    double R[9];
    rot_row3(a.rot, c, R);
    const PView Fr{a.P[2], nx, ny, nz, a.g.per[0], a.g.per[1], a.g.per[2]};
    const double* dx = a.g.dx;
    double x = 0.0;
    if (R[6] != 0.0) x += R[6] * dpd<0>(Fr, a.M, i, j, k, dx, a.pbc);
    if (R[7] != 0.0) x += R[7] * dpd<1>(Fr, a.M, i, j, k, dx, a.pbc);
    x += R[8] * dpd<2>(Fr, a.M, i, j, k, dx, a.pbc);
    a.rhs[c] = -x;
}

/* Heun combination, main.cpp:336-341: P <- P + dt*(0.5*f0 + 0.5*f1) in FE cells. */
__global__ void k_heun(HeunArgs a) {
    for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < a.N; c += (long long)gridDim.x * blockDim.x) {
        for (int d = 0; d < 3; ++d) {
            if (!a.P[d]) continue;
            const double avg = 0.5 * a.G0[d][c] + 0.5 * a.G1[d][c];
            a.P[d][c] = a.P[d][c] + a.dt * avg;
        }
    }
}

/* ComputeEfromPhi for all cells (output field). */
__global__ __launch_bounds__(256) void k_compute_E(EArgs a) {
    const int nx = a.g.n[0], ny = a.g.n[1], nz = a.g.n[2];
    const int i = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int k = blockIdx.z;
    if (i >= nx || j >= ny || k >= nz) return;
    const int c = i + nx * (j + ny * k);
    double R[9];
    rot_of(a.rot, c, R);
    double fx, fy, fz;
    grad_phi(a.phv, a.g, i, j, k, fx, fy, fz);
    a.E[0][c] = -(R[0] * fx + R[1] * fy + R[2] * fz);
    a.E[1][c] = -(R[3] * fx + R[4] * fy + R[5] * fz);
    a.E[2][c] = -(R[6] * fx + R[7] * fy + R[8] * fz);
}

/* ------------------------------ launchers -------------------------------- */
namespace fxk {

static inline unsigned nblk(long long n, int bs) { return (unsigned)((n + bs - 1) / bs); }
static inline dim3 grid_cells(const int n[3]) { return dim3((n[0] + 63) / 64, (n[1] + 3) / 4, n[2]); }

void publish(hipStream_t s, const double* src, int nq, double* host, unsigned* seq, unsigned v) {
    hipLaunchKernelGGL(k_publish, dim3(1), dim3(64), 0, s, src, nq, host, seq, v);
}
void fill(hipStream_t s, double* p, long long n, double v) {
    unsigned b = nblk(n, 256); if (b > 4096) b = 4096; if (b == 0) b = 1;
    hipLaunchKernelGGL(k_fill, dim3(b), dim3(256), 0, s, p, n, v);
}
void copy(hipStream_t s, double* d, const double* src, long long n) {
    unsigned b = nblk(n, 256); if (b > 4096) b = 4096; if (b == 0) b = 1;
    hipLaunchKernelGGL(k_copy, dim3(b), dim3(256), 0, s, d, src, n);
}
void add(hipStream_t s, double* d, const double* src, long long n) {
    unsigned b = nblk(n, 256); if (b > 4096) b = 4096; if (b == 0) b = 1;
    hipLaunchKernelGGL(k_add, dim3(b), dim3(256), 0, s, d, src, n);
}
void final_reduce(hipStream_t s, const double* part, int nb, int nq, int is_max, double* out) {
    hipLaunchKernelGGL(k_final, dim3(1), dim3(256), 0, s, part, nb, nq, is_max, out);
}
void maxabs2(hipStream_t s, const double* a, const double* b, long long n, double* part, double* out) {
    hipLaunchKernelGGL(k_maxabs2, dim3(2048), dim3(256), 0, s, a, b, n, part);
    final1024(s, part, 2048, 2, 1, out);
}
void phidiff(hipStream_t s, const double* phi, double* phi_old, double* diff, long long n, const double* phi_max) {
    unsigned b = nblk(n, 256); if (b > 4096) b = 4096; if (b == 0) b = 1;
    hipLaunchKernelGGL(k_phidiff, dim3(b), dim3(256), 0, s, phi, phi_old, diff, n, phi_max);
}
void rhs_newton(hipStream_t s, const RhsArgs& a) {
    dim3 g = grid_cells(a.g.n);
    g.z = a.kz1 - a.kz0 + 1;
    if (a.kz1 >= a.kz0) hipLaunchKernelGGL(k_rhs_newton, g, dim3(256), 0, s, a);
}
int rho_update(hipStream_t s, const RhoArgs& a) {
    const dim3 g = grid_cells(a.n);
    hipLaunchKernelGGL(k_rho_update, g, dim3(256), 0, s, a);
    return (int)(g.x * g.y * g.z);
}
void tdgl(hipStream_t s, const TdglArgs& a, bool scalar_fast) {
    dim3 g = grid_cells(a.g.n);
    g.z = a.kz1 - a.kz0 + 1;
    if (a.kz1 < a.kz0) return;
    if (scalar_fast) hipLaunchKernelGGL(k_tdgl<true>, g, dim3(256), 0, s, a);
    else hipLaunchKernelGGL(k_tdgl<false>, g, dim3(256), 0, s, a);
}
void tdgl_rhs_stream(hipStream_t s, const TdglRhsArgs& a) {
    const unsigned nb = (unsigned)((a.nface + 255) / 256);
    if (nb) hipLaunchKernelGGL(k_tdgl_face_cells, dim3(nb), dim3(256), 0, s, a);
    const dim3 g((a.t.g.n[0] + TRS_TX - 1) / TRS_TX, (a.t.g.n[1] + TRS_TY - 1) / TRS_TY,
                 (a.k1 - a.k0 + TRS_Z) / TRS_Z);
    hipLaunchKernelGGL(k_tdgl_rhs_stream, g, dim3(TRS_THR), 0, s, a);
    if (nb) hipLaunchKernelGGL(k_rhs_face_cells, dim3(nb), dim3(256), 0, s, a);
}
void tdgl_stream(hipStream_t s, hipStream_t s2, hipEvent_t e0, hipEvent_t e1, const TdglRhsArgs& a) {
    const unsigned nb = (unsigned)((a.nface + 255) / 256);
    /* the face cells on the second stream, beside the streamed interior */
    if (nb) {
        (void)hipEventRecord(e0, s);
        (void)hipStreamWaitEvent(s2, e0, 0);
        hipLaunchKernelGGL(k_tdgl_face_cells, dim3(nb), dim3(256), 0, s2, a);
        (void)hipEventRecord(e1, s2);
    }
    const dim3 g((a.t.g.n[0] + TDS_TX - 1) / TDS_TX, (a.t.g.n[1] + TDS_TY - 1) / TDS_TY,
                 (a.k1 - a.k0 + TDS_Z) / TDS_Z);
    if (a.t.rot.idx && a.t.rot.nrows <= TDS_MAXROWS)
        hipLaunchKernelGGL(k_tdgl_stream<true>, g, dim3(TDS_THR), (size_t)10 * a.t.rot.nrows * sizeof(double), s, a);
    else
        hipLaunchKernelGGL(k_tdgl_stream<false>, g, dim3(TDS_THR), 0, s, a);
    if (nb) (void)hipStreamWaitEvent(s, e1, 0);
}
void rhs_stream(hipStream_t s, hipStream_t s2, hipEvent_t e0, hipEvent_t e1, const RhsArgs& a, int k0, int k1,
                const int* face, int nface) {
    const unsigned nb = (unsigned)((nface + 255) / 256);
    if (nb) {
        (void)hipEventRecord(e0, s);
        (void)hipStreamWaitEvent(s2, e0, 0);
        hipLaunchKernelGGL(k_rhs_face_list, dim3(nb), dim3(256), 0, s2, a, face, nface);
        (void)hipEventRecord(e1, s2);
    }
    const dim3 g((a.g.n[0] + TDS_TX - 1) / TDS_TX, (a.g.n[1] + TDS_TY - 1) / TDS_TY, (k1 - k0 + TDS_Z) / TDS_Z);
    if (a.rot.idx && a.rot.nrows <= 3 * TDS_MAXROWS)
        hipLaunchKernelGGL(k_rhs_stream<true>, g, dim3(TDS_THR), (size_t)3 * a.rot.nrows * sizeof(double), s, a, k0, k1);
    else
        hipLaunchKernelGGL(k_rhs_stream<false>, g, dim3(TDS_THR), 0, s, a, k0, k1);
    if (nb) (void)hipStreamWaitEvent(s, e1, 0);
}
void heun(hipStream_t s, const HeunArgs& a) {
    unsigned b = nblk(a.N, 256); if (b > 4096) b = 4096; if (b == 0) b = 1;
    hipLaunchKernelGGL(k_heun, dim3(b), dim3(256), 0, s, a);
}
void compute_E(hipStream_t s, const EArgs& a) {
    hipLaunchKernelGGL(k_compute_E, grid_cells(a.g.n), dim3(256), 0, s, a);
}
}  // namespace fxk
