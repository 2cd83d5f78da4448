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
template <class V = PView>
__device__ __forceinline__ double grad_r(const V& Fr, const MView& M, int i, int j, int k, const double* dx,
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

/* The gradient and field terms of the scalar update of an interior FE cell
 * as fixed weights of the raw differences, per rotation: with
 * D_xx = P(i+1) - 2P + P(i-1) (likewise y, z) and D_xy = (P(+1,+1) -
 * P(+1,-1)) - (P(-1,+1) - P(-1,-1)) (likewise yz, xz),
 *   grad_r_of(R, H) = -(w0 D_xx + w1 D_yy + w2 D_zz + w3 D_xy + w4 D_yz + w5 D_xz),
 *   w0 = (g11 R31^2 + g44m (R11^2 + R21^2)) / dx^2,
 *   w3 = 2 (g11 R31 R32 + g44m (R11 R12 + R21 R22)) / (4 dx dy), ...;
 * and the field term E2 = -(w6 d_x + w7 d_y + w8 d_z) with d_x = phi(i+1) -
 * phi(i-1), w6 = R31 / (2 dx), ...  The same sums as grad_r_interior_acc and
 * the centred E, reassociated (rounding-level differences). */
__device__ __forceinline__ void interior_weights(const double R[9], double g11, double g44m, const double ih[3],
                                                 double w[9]) {
    const double R11 = R[0], R12 = R[1], R13 = R[2], R21 = R[3], R22 = R[4], R23 = R[5], R31 = R[6], R32 = R[7], R33 = R[8];
    w[0] = (g11 * R31 * R31 + g44m * (R11 * R11 + R21 * R21)) * (ih[0] * ih[0]);
    w[1] = (g11 * R32 * R32 + g44m * (R12 * R12 + R22 * R22)) * (ih[1] * ih[1]);
    w[2] = (g11 * R33 * R33 + g44m * (R13 * R13 + R23 * R23)) * (ih[2] * ih[2]);
    w[3] = 2. * (g11 * R31 * R32 + g44m * (R11 * R12 + R21 * R22)) * (0.25 * ih[0] * ih[1]);
    w[4] = 2. * (g11 * R32 * R33 + g44m * (R12 * R13 + R22 * R23)) * (0.25 * ih[1] * ih[2]);
    w[5] = 2. * (g11 * R33 * R31 + g44m * (R13 * R11 + R23 * R21)) * (0.25 * ih[0] * ih[2]);
    w[6] = R31 * (0.5 * ih[0]);
    w[7] = R32 * (0.5 * ih[1]);
    w[8] = R33 * (0.5 * ih[2]);
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
    /* in-plane source of a window column/row: the cell, its periodic image
     * one cell past either side, or -1 (across a non-periodic side, or past
     * the domain where the tile is wider than it) */
    auto wrap = [](int v, int n, int per) -> int {
        if (v >= 0 && v < n) return v;
        if (v == -1) return per ? n - 1 : -1;
        if (v == n) return per ? 0 : -1;
        if (v == -2) return per && n >= 2 ? n - 2 : -1;
        if (v == n + 1) return per && n >= 2 ? 1 : -1;
        return -1;
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
 * and a chunk of zc planes of [k0, k1] and marches upward.  P and phi of
 * each plane (tile + 1-cell halo) go to a ring of four LDS slots: while
 * plane p is updated from slots p-1, p, p+1, plane p+2 (loaded into
 * registers at the top of the step) is stored into the fourth slot, so one
 * barrier per plane suffices.  Every global load of a step is issued
 * unconditionally (clamped addresses, the value selected afterwards), so
 * the update of plane p never waits for the loads in flight.  FE cells next
 * to a face (the list of k_tdgl_face_cells) are updated by that kernel;
 * other cells keep their P (Gamma = 0). */
constexpr int TDS_TX = 64, TDS_TY = 8;
constexpr int TDS_HX = TDS_TX + 2, TDS_HY = TDS_TY + 2;
constexpr int TDS_HP = TDS_HX * TDS_HY;                                   /* values of a halo plane */
constexpr int TDS_THR = TDS_TX * TDS_TY;                                  /* 512 */
constexpr int TDS_PV = (TDS_HP + TDS_THR - 1) / TDS_THR;                  /* halo-plane values per thread */
constexpr int TDS_SLOTS = 4;
constexpr int TDS_MAXROWS = 416;   /* rotation / alpha tables up to this many rows are staged in LDS (two
                                      workgroups per CU: 2 x (42 KB ring + 6 KB ids + 80 B per row) <= 160 KB) */

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

/* P at (i, j, k) with the ghost semantics of PView, from the LDS window of
 * the streamed passes around (ci, cj, p) (slots of planes p-1, p, p+1, tile
 * offset wo) where it reaches, else from global memory: the masked stencils
 * of the cells next to a face read at most one cell away, except the
 * two-cell one-sided second differences (P_BC_flag 4). */
struct WinPView {
    const double* w0;
    const double* w1;
    const double* w2;
    int ci, cj, p, wo;
    PView g;
    __device__ __forceinline__ double operator()(int i, int j, int k) const {
        const int di = i - ci, dj = j - cj, dk = k - p;
        if (di >= -1 && di <= 1 && dj >= -1 && dj <= 1 && dk >= -1 && dk <= 1) {
            const double* w = dk < 0 ? w0 : dk > 0 ? w2 : w1;
            return w[wo + di + dj * TDS_HX];
        }
        return g(i, j, k);
    }
};

/* phi at (i, j, k) with the ghost semantics of PhiView: from the window
 * where the cell is inside the domain (or across a periodic side, whose
 * image the window holds), else PhiView's Dirichlet / Neumann ghost */
struct WinPhi {
    const double* w0;
    const double* w1;
    const double* w2;
    int ci, cj, p, wo;
    PhiView g;
    __device__ __forceinline__ double at(int i, int j, int k) const {
        const int di = i - ci, dj = j - cj, dk = k - p;
        const bool in_i = (i >= 0 && i < g.n[0]) || g.lo[0] == BC_PER;
        const bool in_j = (j >= 0 && j < g.n[1]) || g.lo[1] == BC_PER;
        if (di >= -1 && di <= 1 && dj >= -1 && dj <= 1 && dk >= -1 && dk <= 1 && in_i && in_j && k >= 0 &&
            k < g.n[2]) {
            const double* w = dk < 0 ? w0 : dk > 0 ? w2 : w1;
            return w[wo + di + dj * TDS_HX];
        }
        return g.at(i, j, k);
    }
};

/* Tile of a streamed pass's workgroup.  The dispatcher hands workgroup L
 * (in x, y, z order) to XCD L mod 8; each XCD takes one contiguous share of
 * the tiles instead (x fastest), so that neighbouring tiles, whose halos
 * overlap, share an L2.  A bijection of the grid onto itself. */
#ifndef FX_STREAM_XCD
#define FX_STREAM_XCD 1
#endif
__device__ __forceinline__ void stream_tile(int& bx, int& by, int& bz) {
    if (!FX_STREAM_XCD) { bx = blockIdx.x; by = blockIdx.y; bz = blockIdx.z; return; }
    const int T = gridDim.x * gridDim.y * gridDim.z;
    const int L = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
    const int x = L & 7, q = T >> 3, r = T & 7;
    const int t = x * q + min(x, r) + (L >> 3);
    bx = t % gridDim.x;
    const int rest = t / gridDim.x;
    by = rest % gridDim.y;
    bz = rest / gridDim.y;
}

/* In-plane source of window column (or row) v of a domain of n cells:
 * cells 0 .. n-1 themselves, -1 and n their periodic images (or nothing
 * across a non-periodic side); positions further out, which arise when the
 * tile is wider than the domain, are read by no cell of the domain and have
 * no source.  Returns -1 for no source. */
__device__ __forceinline__ int halo_src(int v, int n, int per) {
    if (v >= 0 && v < n) return v;
    if (v == -1) return per ? n - 1 : -1;
    if (v == n) return per ? 0 : -1;
    return -1;
}

/* The halo plane a thread stages: for r < TDS_PV, window position t =
 * threadIdx.x + r TDS_THR (if below TDS_HP), its in-plane source index
 * (0, the plane's first cell, where it has none) and whether the source
 * exists.  Every offset lies inside the plane. */
struct HaloSrc {
    int off[TDS_PV];
    bool ok[TDS_PV];
    __device__ __forceinline__ HaloSrc(int i0, int j0, int nx, int ny, int per0, int per1) {
#pragma unroll
        for (int r = 0; r < TDS_PV; ++r) {
            const int t = threadIdx.x + r * TDS_THR;
            const int v = t / TDS_HX, u = t - v * TDS_HX;
            const int i = halo_src(i0 + u - 1, nx, per0), j = halo_src(j0 + v - 1, ny, per1);
            const bool good = t < TDS_HP && i >= 0 && j >= 0;
            ok[r] = good;
            off[r] = good ? i + nx * j : 0;
        }
    }
};

/* What a streamed pass loads for plane q: the halo plane of P (and of phi
 * for the update) and the material id and table row of this thread's tile
 * cell.  Every load is issued unconditionally (the plane index clamped to
 * [0, qmax], halo sources off the domain read the plane's first value), so
 * the compiler's wait counting never merges paths with different loads in
 * flight; what is off the domain is replaced by 0 when stored to LDS. */
template <bool PHI>
struct PlaneRegs {
    double p[TDS_PV];
    double f[TDS_PV];
    int raw, row;
};

template <bool PHI, int ROT>
__device__ __forceinline__ void plane_fetch(PlaneRegs<PHI>& R, int q, int qmax, const double* __restrict__ P,
                                            const double* __restrict__ ph, const HaloSrc& hs, long long plane,
                                            const MView& M, int cic, int cjc, const int* __restrict__ ridx,
                                            long long cc0) {
    const int qc = q < 0 ? 0 : q > qmax ? qmax : q;
    const long long base = plane * (long long)qc;
#pragma unroll
    for (int r = 0; r < TDS_PV; ++r) {
        R.p[r] = P[base + hs.off[r]];
        if (PHI) R.f[r] = ph[base + hs.off[r]];
    }
    R.raw = M.raw(cic, cjc, qc);
    R.row = ROT ? ridx[cc0 + base] : 0;
}

/* LDS ring of a streamed pass: four slots of the P (and phi) halo planes,
 * three of the tile's packed material id | table row << 8. */
template <bool PHI>
__device__ __forceinline__ void plane_stash(const PlaneRegs<PHI>& R, int q, int nz, const HaloSrc& hs, double* Pw,
                                            double* Fw, int* Rr) {
    const bool qok = q >= 0 && q < nz;
    double* dp = Pw + (q & 3) * TDS_HP;
    double* df = Fw + (q & 3) * TDS_HP;
#pragma unroll
    for (int r = 0; r < TDS_PV; ++r) {
        const int t = threadIdx.x + r * TDS_THR;
        if (t < TDS_HP) {
            const bool ok = qok && hs.ok[r];
            dp[t] = ok ? R.p[r] : 0.0;
            if (PHI) df[t] = ok ? R.f[r] : 0.0;
        }
    }
    Rr[((q + 3) % 3) * TDS_THR + threadIdx.x] = R.raw | (R.row << 8);
}

/* ROT: 0 uniform rotation (no table), 1 rotation/alpha table staged in LDS,
 * 2 table read from global memory.  FACES false: the interior FE cells of
 * chunks of zc planes of [k0, k1]; true: the FE cells next to a face (masked
 * stencils on the window) of plane zlist[blockIdx.z] (the planes that hold
 * such cells), a separate launch because the masked stencils need twice the
 * registers.
 *
 * The march: planes ks-1 .. ks+1 are staged first and ks+2 is loaded into
 * registers; step p loads plane p+3 into the other register set, updates
 * plane p from LDS, stores plane p+2 (loaded one step earlier) into the
 * ring slot of plane p-2 and ends with the plane's one barrier.  The loop
 * is unrolled by two so that the two register sets stay fixed; planes past
 * ke + 1 are loaded as ke + 1 again (cache hits) and not used. */
template <int ROT, bool FACES>
__global__ __launch_bounds__(TDS_THR, 2) void k_tdgl_stream(TdglRhsArgs A, int zc, const int* __restrict__ zlist) {
    __shared__ double Pw[TDS_SLOTS * TDS_HP];
    __shared__ double Fw[TDS_SLOTS * TDS_HP];
    __shared__ int Rr[3 * TDS_THR];
    extern __shared__ __align__(16) unsigned char tds_dyn[];
    /* ROT 1: per row, the interior weights w0..w8 and alpha (interior pass)
     * or the 9 R values and alpha (face pass) */
    double* const Rt = reinterpret_cast<double*>(tds_dyn);
    const TdglArgs& a = A.t;
    const double* dx = a.g.dx;
    const double ih[3] = {1.0 / dx[0], 1.0 / dx[1], 1.0 / dx[2]};
    const double g44m = a.g44 - a.g44_p;
    if (ROT == 1) {
        const int nr = a.rot.nrows;
        if (FACES) {
            for (int t = threadIdx.x; t < 10 * nr; t += TDS_THR)
                Rt[t] = t < 9 * nr ? a.rot.tab[t] : a.rot.atab[t - 9 * nr];
        } else {
            for (int q = threadIdx.x; q < nr; q += TDS_THR) {
                double R[9], w[9];
#pragma unroll
                for (int m = 0; m < 9; ++m) R[m] = a.rot.tab[9 * q + m];
                interior_weights(R, a.g11, g44m, ih, w);
#pragma unroll
                for (int m = 0; m < 9; ++m) Rt[10 * q + m] = w[m];
                Rt[10 * q + 9] = a.rot.atab[q];
            }
        }
    }
    double wu[9];   /* ROT 0: the uniform rotation's weights */
    if (ROT == 0 && !FACES) interior_weights(a.rot.Ru, a.g11, g44m, ih, wu);
    double zl[6], zh[6];   /* the masked z stencils of the cells next to a z face (MAT_ZFACE_LO/HI) */
    zface_coef(a.pbc.lo[2], false, dx[2], a.pbc.lambda, zl);
    zface_coef(a.pbc.hi[2], true, dx[2], a.pbc.lambda, zh);
    const int nx = a.g.n[0], ny = a.g.n[1], nz = a.g.n[2];
    int bx, by, bz;
    stream_tile(bx, by, bz);
    const int i0 = bx * TDS_TX, j0 = by * TDS_TY;
    const int ks = FACES ? zlist[bz] : A.k0 + bz * zc;
    const int ke = FACES ? ks : min(ks + zc - 1, A.k1);
    const int qmax = min(ke + 1, nz - 1);
    const double* __restrict__ P = a.P[2];
    const double* __restrict__ ph = a.phv.p;
    const HaloSrc hs(i0, j0, nx, ny, a.g.per[0], a.g.per[1]);
    const long long plane = (long long)nx * ny;
    const int tu = threadIdx.x % TDS_TX, tv = threadIdx.x / TDS_TX;
    const int ci = i0 + tu, cj = j0 + tv;
    const bool in_dom = ci < nx && cj < ny;
    const int cic = in_dom ? ci : 0, cjc = in_dom ? cj : 0;   /* clamped: every load is in bounds */
    const int wo = (tv + 1) * TDS_HX + (tu + 1);
    const int* __restrict__ ridx = a.rot.idx;
    const long long cc0 = cic + (long long)nx * cjc;
    auto fetch = [&](PlaneRegs<true>& R, int q) { plane_fetch<true, ROT>(R, q, qmax, P, ph, hs, plane, a.M, cic, cjc, ridx, cc0); };
    auto stash = [&](const PlaneRegs<true>& R, int q) { plane_stash<true>(R, q, nz, hs, Pw, Fw, Rr); };
    PlaneRegs<true> RA, RB;
    if (FACES) {
        /* one plane: its three planes' loads in flight together */
        PlaneRegs<true> RC;
        fetch(RA, ks - 1); fetch(RB, ks); fetch(RC, ks + 1);
        stash(RA, ks - 1); stash(RB, ks); stash(RC, ks + 1);
    } else {
        fetch(RA, ks - 1); stash(RA, ks - 1);
        fetch(RA, ks); stash(RA, ks);
        fetch(RA, ks + 1); stash(RA, ks + 1);
        fetch(RA, ks + 2);
    }
    __syncthreads();
    const double Gam = a.BigGamma;

    auto update = [&](int p) {
        const int rr = Rr[((p + 3) % 3) * TDS_THR + threadIdx.x];
        const int raw_c = rr & 255, row_c = rr >> 8;
        const int s0 = ((p - 1) & 3) * TDS_HP, s1 = (p & 3) * TDS_HP, s2 = ((p + 1) & 3) * TDS_HP;
        if (!FACES && in_dom && (raw_c & 3) == 0 && (raw_c & (MAT_INTERIOR_FE | MAT_ZFACE_LO | MAT_ZFACE_HI))) {
            const int c = ci + nx * (cj + ny * p);
            const double* f0 = Fw + s0;
            const double* f1 = Fw + s1;
            const double* f2 = Fw + s2;
            const WinNbr F{Pw + s0, Pw + s1, Pw + s2, wo};
            double w[9];
            double alpha_c = a.alpha;
            if (ROT == 1) {
#pragma unroll
                for (int q = 0; q < 9; ++q) w[q] = Rt[10 * row_c + q];
                alpha_c = Rt[10 * row_c + 9];
            } else if (ROT == 2) {
                double R[9];
                const double* t = a.rot.tab + 9 * row_c;
#pragma unroll
                for (int q = 0; q < 9; ++q) R[q] = t[q];
                interior_weights(R, a.g11, g44m, ih, w);
                alpha_c = a.rot.atab[row_c];
            } else {
#pragma unroll
                for (int q = 0; q < 9; ++q) w[q] = wu[q];
            }
            const bool tph = (raw_c & 4) != 0;
            const double r = Pw[s1 + wo];
            double L = 0.0, G = 0.0, El = 0.0;
            if (a.incL == 1) {
                const double r2 = r * r;
                const double dF = alpha_c * r + a.beta * (r2 * r) + a.gamma * (r2 * r2 * r);
                L = -1.0 * Gam * dF;
            }
            if (a.incG == 1) {
                const double Dxx = (F(1, 0, 0) - 2. * r) + F(-1, 0, 0);
                const double Dyy = (F(0, 1, 0) - 2. * r) + F(0, -1, 0);
                double Dzz = (F(0, 0, 1) - 2. * r) + F(0, 0, -1);
                const double Dxy = (F(1, 1, 0) - F(1, -1, 0)) - (F(-1, 1, 0) - F(-1, -1, 0));
                double Dyz = (F(0, 1, 1) - F(0, 1, -1)) - (F(0, -1, 1) - F(0, -1, -1));
                double Dxz = (F(1, 0, 1) - F(1, 0, -1)) - (F(-1, 0, 1) - F(-1, 0, -1));
                if (!(raw_c & MAT_INTERIOR_FE)) {
                    /* next to a z face: the masked z stencils as linear forms
                     * (h^2 DoubleDPDz, and 2h DPDz at the lateral neighbours of
                     * the mixed entries) */
                    const bool lo = (raw_c & MAT_ZFACE_LO) != 0;
                    double z[6];
#pragma unroll
                    for (int q = 0; q < 6; ++q) z[q] = lo ? zl[q] : zh[q];
                    Dzz = z[0] * r + z[1] * F(0, 0, 1) + z[2] * F(0, 0, -1);
                    auto D2 = [&](int di, int dj) { return z[3] * F(di, dj, 0) + z[4] * F(di, dj, 1) + z[5] * F(di, dj, -1); };
                    Dyz = D2(0, 1) - D2(0, -1);
                    Dxz = D2(1, 0) - D2(-1, 0);
                }
                const double gr = w[0] * Dxx + w[1] * Dyy + w[2] * Dzz + w[3] * Dxy + w[4] * Dyz + w[5] * Dxz;
                G = Gam * gr;                          /* -Gam * grad_r_of, grad_r_of = -gr */
            }
            if (a.incE == 1) {
                const double sE = w[6] * (f1[wo + 1] - f1[wo - 1]) + w[7] * (f1[wo + TDS_HX] - f1[wo - TDS_HX]) +
                                  w[8] * (f2[wo] - f0[wo]);
                El = -Gam * sE;                        /* -Gam * (-E2), E2 = -sE */
            }
            if (tph) { L = 0.0; G = 0.0; El = 0.0; }
            double g = 0.0;
            if (a.incL == 1) g += L;
            if (a.incG == 1) g += G;
            if (a.incE == 1) g += El;
            a.Pout[2][c] = r + a.dt * g;
        } else if (FACES && in_dom && (raw_c & 3) == 0 && !(raw_c & (MAT_INTERIOR_FE | MAT_ZFACE_LO | MAT_ZFACE_HI))) {
            /* an FE cell next to a face: the masked stencils (tdgl_scalar_cell),
             * operands from the window */
            const int c = ci + nx * (cj + ny * p);
            const WinPView Fr{Pw + s0, Pw + s1, Pw + s2, ci, cj, p, wo,
                              PView{P, nx, ny, nz, a.g.per[0], a.g.per[1], a.g.per[2]}};
            const WinPhi Fp{Fw + s0, Fw + s1, Fw + s2, ci, cj, p, wo, a.phv};
            double R[9];
            double alpha_c = a.alpha;
            if (ROT == 1) {
#pragma unroll
                for (int q = 0; q < 9; ++q) R[q] = Rt[9 * row_c + q];
                alpha_c = Rt[9 * a.rot.nrows + row_c];
            } else if (ROT == 2) {
                const double* t = a.rot.tab + 9 * row_c;
#pragma unroll
                for (int q = 0; q < 9; ++q) R[q] = t[q];
                alpha_c = a.rot.atab[row_c];
            } else {
#pragma unroll
                for (int q = 0; q < 9; ++q) R[q] = a.rot.Ru[q];
            }
            const bool tph = (raw_c & 4) != 0;
            double E2 = 0.0;
            if (a.incE == 1) {
                double fx, fy, fz;
                grad_phi(Fp, a.g, ci, cj, p, fx, fy, fz);
                double sE = 0.0;
                if (R[6] != 0.0) sE += R[6] * fx;
                if (R[7] != 0.0) sE += R[7] * fy;
                sE += R[8] * fz;
                E2 = -sE;
            }
            const double r = Pw[s1 + wo];
            double L = 0.0, G = 0.0, El = 0.0;
            if (a.incL == 1) {
                const double r2 = r * r;
                const double dF = alpha_c * r + a.beta * (r2 * r) + a.gamma * (r2 * r2 * r);
                L = -1.0 * Gam * dF;
            }
            if (a.incG == 1) G = -1.0 * Gam * grad_r(Fr, a.M, ci, cj, p, dx, a.pbc, R, a.g11, a.g44 - a.g44_p);
            if (a.incE == 1) El = -1.0 * Gam * (-E2);
            if (tph) { L = 0.0; G = 0.0; El = 0.0; }
            double g = 0.0;
            if (a.incL == 1) g += L;
            if (a.incG == 1) g += G;
            if (a.incE == 1) g += El;
            a.Pout[2][c] = r + a.dt * g;
        }
    };
    if (FACES) {
        update(ks);
        return;
    }
    for (int p = ks; p <= ke; p += 2) {
        fetch(RB, p + 3);
        update(p);
        stash(RA, p + 2);                              /* the slot of plane p-2, read by no step after p-1 */
        __syncthreads();
        fetch(RA, p + 4);
        if (p + 1 <= ke) update(p + 1);
        stash(RB, p + 3);
        __syncthreads();
    }
}

/* compute units of the current device (256 if it cannot be read) */
static int num_cus() {
    static int ncu = 0;
    if (!ncu) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu <= 0)
            ncu = 256;
    }
    return ncu;
}

/* Workgroups of the streamed passes resident per CU: the update's LDS (the
 * four-slot P and phi ring, the id ring and the rotation table, 80 KB)
 * holds it to two; the RHS would fit four, but its extra warm-up planes
 * cost more than the occupancy gains (measured). */
#ifndef FX_TDGL_WG_PER_CU
#define FX_TDGL_WG_PER_CU 2
#endif
#ifndef FX_RHS_WG_PER_CU
#define FX_RHS_WG_PER_CU 2
#endif

/* z-chunk length of the streamed passes over nk planes with nxy column
 * tiles: the fewest chunks that keep the grid within whole rounds of wpc
 * workgroups per CU, measured by (rounds) x (chunk + 2 warm-up planes) */
int stream_chunk(int nxy, int nk, int ncu, int wpc) {
    const int slots = wpc * (ncu > 0 ? ncu : 256);
    int best = nk, best_cost = 1 << 30;
    for (int nch = 1; nch <= nk; ++nch) {
        const int zc = (nk + nch - 1) / nch;
        const int rounds = (nxy * ((nk + zc - 1) / zc) + slots - 1) / slots;
        const int cost = rounds * (zc + 2);
        if (cost < best_cost) { best_cost = cost; best = zc; }
    }
    return best;
}

/* ------------------- Poisson RHS streamed through LDS (scalar P) ----------- */
/* k_rhs_newton's centred branch for the interior FE cells of planes
 * [k0, k1] (no SC cells, the RHS buffer already 0 in every non-FE cell, the
 * Newton coefficient not written): P of each plane (tile + 1-cell halo) in
 * the four-slot LDS ring of k_tdgl_stream, the plane after next in
 * registers, R_3 from an LDS copy of the rotation table; FE cells next to a
 * face by list (k_rhs_face_list). */
template <int ROT, bool FACES>
__global__ __launch_bounds__(TDS_THR, 2) void k_rhs_stream(RhsArgs a, int k0, int k1, int zc,
                                                           const int* __restrict__ zlist) {
    __shared__ double Pw[TDS_SLOTS * TDS_HP];
    __shared__ int Rr[3 * TDS_THR];
    extern __shared__ __align__(16) unsigned char rs_dyn[];
    double* const Rt = reinterpret_cast<double*>(rs_dyn);   /* ROT 1: R_31, R_32, R_33 per row */
    if (ROT == 1) {
        const int nr = a.rot.nrows;
        for (int t = threadIdx.x; t < 3 * nr; t += TDS_THR) Rt[t] = a.rot.tab[9 * (t / 3) + 6 + t % 3];
    }
    const int nx = a.g.n[0], ny = a.g.n[1], nz = a.g.n[2];
    int bx, by, bz;
    stream_tile(bx, by, bz);
    const int i0 = bx * TDS_TX, j0 = by * TDS_TY;
    const int ks = FACES ? zlist[bz] : k0 + bz * zc;
    const int ke = FACES ? ks : min(ks + zc - 1, k1);
    const int qmax = min(ke + 1, nz - 1);
    const double* __restrict__ P = a.P[2];
    const HaloSrc hs(i0, j0, nx, ny, a.g.per[0], a.g.per[1]);
    const long long plane = (long long)nx * ny;
    const int tu = threadIdx.x % TDS_TX, tv = threadIdx.x / TDS_TX;
    const int ci = i0 + tu, cj = j0 + tv;
    const bool in_dom = ci < nx && cj < ny;
    const int cic = in_dom ? ci : 0, cjc = in_dom ? cj : 0;
    const int wo = (tv + 1) * TDS_HX + (tu + 1);
    const int* __restrict__ ridx = a.rot.idx;
    const long long cc0 = cic + (long long)nx * cjc;
    auto fetch = [&](PlaneRegs<false>& R, int q) { plane_fetch<false, ROT>(R, q, qmax, P, nullptr, hs, plane, a.M, cic, cjc, ridx, cc0); };
    auto stash = [&](const PlaneRegs<false>& R, int q) { plane_stash<false>(R, q, nz, hs, Pw, nullptr, Rr); };
    PlaneRegs<false> RA, RB;
    if (FACES) {
        PlaneRegs<false> RC;
        fetch(RA, ks - 1); fetch(RB, ks); fetch(RC, ks + 1);
        stash(RA, ks - 1); stash(RB, ks); stash(RC, ks + 1);
    } else {
        fetch(RA, ks - 1); stash(RA, ks - 1);
        fetch(RA, ks); stash(RA, ks);
        fetch(RA, ks + 1); stash(RA, ks + 1);
        fetch(RA, ks + 2);
    }
    __syncthreads();
    const double* dx = a.g.dx;
    double zl[6], zh[6];   /* the masked z stencils of the cells next to a z face (MAT_ZFACE_LO/HI) */
    zface_coef(a.pbc.lo[2], false, dx[2], a.pbc.lambda, zl);
    zface_coef(a.pbc.hi[2], true, dx[2], a.pbc.lambda, zh);
    auto rhs_of = [&](int p) {
        const int rr = Rr[((p + 3) % 3) * TDS_THR + threadIdx.x];
        const int raw_c = rr & 255, row_c = rr >> 8;
        const double* p0 = Pw + ((p - 1) & 3) * TDS_HP;
        const double* p1 = Pw + (p & 3) * TDS_HP;
        const double* p2 = Pw + ((p + 1) & 3) * TDS_HP;
        if (!FACES && in_dom && (raw_c & 3) == 0 && (raw_c & (MAT_INTERIOR_FE | MAT_ZFACE_LO | MAT_ZFACE_HI))) {
            const int c = ci + nx * (cj + ny * p);
            double R6, R7, R8;
            if (ROT == 1) { R6 = Rt[3 * row_c]; R7 = Rt[3 * row_c + 1]; R8 = Rt[3 * row_c + 2]; }
            else if (ROT == 2) { const double* t = a.rot.tab + 9 * row_c; R6 = t[6]; R7 = t[7]; R8 = t[8]; }
            else { R6 = a.rot.Ru[6]; R7 = a.rot.Ru[7]; R8 = a.rot.Ru[8]; }
            const double gx = (p1[wo + 1] - p1[wo - 1]) * (0.5 / dx[0]);
            const double gy = (p1[wo + TDS_HX] - p1[wo - TDS_HX]) * (0.5 / dx[1]);
            double gz = (p2[wo] - p0[wo]) * (0.5 / dx[2]);
            if (!(raw_c & MAT_INTERIOR_FE)) {   /* next to a z face: 2h DPDz as a linear form */
                const bool lo = (raw_c & MAT_ZFACE_LO) != 0;
                gz = ((lo ? zl[3] : zh[3]) * p1[wo] + (lo ? zl[4] : zh[4]) * p2[wo] + (lo ? zl[5] : zh[5]) * p0[wo]) *
                     (0.5 / dx[2]);
            }
            double x = 0.0;
            if (R6 != 0.0) x += R6 * gx;
            if (R7 != 0.0) x += R7 * gy;
            x += R8 * gz;
            a.rhs[c] = -x;
        } else if (FACES && in_dom && (raw_c & 3) == 0 && !(raw_c & (MAT_INTERIOR_FE | MAT_ZFACE_LO | MAT_ZFACE_HI))) {
            /* an FE cell next to a face: the masked branch (k_rhs_face_list) on the window */
            const int c = ci + nx * (cj + ny * p);
            double R6, R7, R8;
            if (ROT == 1) { R6 = Rt[3 * row_c]; R7 = Rt[3 * row_c + 1]; R8 = Rt[3 * row_c + 2]; }
            else if (ROT == 2) { const double* t = a.rot.tab + 9 * row_c; R6 = t[6]; R7 = t[7]; R8 = t[8]; }
            else { R6 = a.rot.Ru[6]; R7 = a.rot.Ru[7]; R8 = a.rot.Ru[8]; }
            const WinPView Fr{p0, p1, p2, ci, cj, p, wo, PView{P, nx, ny, nz, a.g.per[0], a.g.per[1], a.g.per[2]}};
            double x = 0.0;
            if (R6 != 0.0) x += R6 * dpd<0>(Fr, a.M, ci, cj, p, dx, a.pbc);
            if (R7 != 0.0) x += R7 * dpd<1>(Fr, a.M, ci, cj, p, dx, a.pbc);
            x += R8 * dpd<2>(Fr, a.M, ci, cj, p, dx, a.pbc);
            a.rhs[c] = -x;
        }
    };
    if (FACES) {
        rhs_of(ks);
        return;
    }
    for (int p = ks; p <= ke; p += 2) {
        fetch(RB, p + 3);
        rhs_of(p);
        stash(RA, p + 2);
        __syncthreads();
        fetch(RA, p + 4);
        if (p + 1 <= ke) rhs_of(p + 1);
        stash(RB, p + 3);
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
/* The cells next to a face: with no more than a quarter of the planes
 * holding such cells, the streamed kernel over those planes, queued after
 * the interior pass on s (it needs twice the interior kernel's registers,
 * so running beside it would halve the interior pass's occupancy); else one
 * thread per listed cell beside the interior pass on the second stream s2
 * (events e0, e1 order it after the work queued on s, and s after it).
 * FX_FACE_ROT_GLOBAL 1: the face-plane kernels read the rotation rows from
 * the cache instead of staging the whole table in LDS (one plane per
 * workgroup leaves nothing to amortise the staging over; 212.6 against
 * 214.3 us per step on the 256x256x128 polycrystal, same box). */
#ifndef FX_FACE_ROT_GLOBAL
#define FX_FACE_ROT_GLOBAL 1
#endif
#ifndef FX_FACE_PLANES_BESIDE
#define FX_FACE_PLANES_BESIDE 0
#endif
static bool use_face_planes(int nplanes, int nk, int mode) { return nplanes > 0 && (mode == 3 || (mode == 1 && 4 * nplanes <= nk)); }

void tdgl_stream(hipStream_t s, hipStream_t s2, hipEvent_t e0, hipEvent_t e1, const TdglRhsArgs& a,
                 const FaceSet& fs, int mode) {
    const int gx = (a.t.g.n[0] + TDS_TX - 1) / TDS_TX, gy = (a.t.g.n[1] + TDS_TY - 1) / TDS_TY;
    const int nk = a.k1 - a.k0 + 1;
    const bool planes = use_face_planes(fs.nplanes, nk, mode);
    const size_t tab = (size_t)10 * a.t.rot.nrows * sizeof(double);
    const int rot = !a.t.rot.idx ? 0 : a.t.rot.nrows <= TDS_MAXROWS ? 1 : 2;
    const bool beside = fs.ncells && (!planes || FX_FACE_PLANES_BESIDE);
    auto faces = [&](hipStream_t q) {
        if (planes) {
            const dim3 gf(gx, gy, fs.nplanes);
            if (rot == 0) hipLaunchKernelGGL((k_tdgl_stream<0, true>), gf, dim3(TDS_THR), 0, q, a, 1, fs.planes);
            else if (rot == 1 && !FX_FACE_ROT_GLOBAL) hipLaunchKernelGGL((k_tdgl_stream<1, true>), gf, dim3(TDS_THR), tab, q, a, 1, fs.planes);
            else hipLaunchKernelGGL((k_tdgl_stream<2, true>), gf, dim3(TDS_THR), 0, q, a, 1, fs.planes);
        } else {
            TdglRhsArgs af = a;   /* the cells of fs (the update's own list may hold more) */
            af.face = fs.cells;
            af.nface = fs.ncells;
            hipLaunchKernelGGL(k_tdgl_face_cells, dim3((fs.ncells + 255) / 256), dim3(256), 0, q, af);
        }
    };
    if (beside) {
        (void)hipEventRecord(e0, s);
        (void)hipStreamWaitEvent(s2, e0, 0);
        faces(s2);
        (void)hipEventRecord(e1, s2);
    }
    const int zc = stream_chunk(gx * gy, nk, num_cus(), FX_TDGL_WG_PER_CU);
    const dim3 g(gx, gy, (nk + zc - 1) / zc);
    if (rot == 0) hipLaunchKernelGGL((k_tdgl_stream<0, false>), g, dim3(TDS_THR), 0, s, a, zc, nullptr);
    else if (rot == 1) hipLaunchKernelGGL((k_tdgl_stream<1, false>), g, dim3(TDS_THR), tab, s, a, zc, nullptr);
    else hipLaunchKernelGGL((k_tdgl_stream<2, false>), g, dim3(TDS_THR), 0, s, a, zc, nullptr);
    if (beside) (void)hipStreamWaitEvent(s, e1, 0);
    else if (fs.ncells) faces(s);
}
void rhs_stream(hipStream_t s, hipStream_t s2, hipEvent_t e0, hipEvent_t e1, const RhsArgs& a, int k0, int k1,
                const FaceSet& fs, int mode) {
    const int gx = (a.g.n[0] + TDS_TX - 1) / TDS_TX, gy = (a.g.n[1] + TDS_TY - 1) / TDS_TY;
    const int nk = k1 - k0 + 1;
    const bool planes = use_face_planes(fs.nplanes, nk, mode);
    const size_t tab = (size_t)3 * a.rot.nrows * sizeof(double);
    const int rot = !a.rot.idx ? 0 : a.rot.nrows <= 3 * TDS_MAXROWS ? 1 : 2;
    const bool beside = fs.ncells && (!planes || FX_FACE_PLANES_BESIDE);
    auto faces = [&](hipStream_t q) {
        if (planes) {
            const dim3 gf(gx, gy, fs.nplanes);
            if (rot == 0) hipLaunchKernelGGL((k_rhs_stream<0, true>), gf, dim3(TDS_THR), 0, q, a, k0, k1, 1, fs.planes);
            else if (rot == 1 && !FX_FACE_ROT_GLOBAL) hipLaunchKernelGGL((k_rhs_stream<1, true>), gf, dim3(TDS_THR), tab, q, a, k0, k1, 1, fs.planes);
            else hipLaunchKernelGGL((k_rhs_stream<2, true>), gf, dim3(TDS_THR), 0, q, a, k0, k1, 1, fs.planes);
        } else {
            hipLaunchKernelGGL(k_rhs_face_list, dim3((fs.ncells + 255) / 256), dim3(256), 0, q, a, fs.cells, fs.ncells);
        }
    };
    if (beside) {
        (void)hipEventRecord(e0, s);
        (void)hipStreamWaitEvent(s2, e0, 0);
        faces(s2);
        (void)hipEventRecord(e1, s2);
    }
    const int zc = stream_chunk(gx * gy, nk, num_cus(), FX_RHS_WG_PER_CU);
    const dim3 g(gx, gy, (nk + zc - 1) / zc);
    if (rot == 0) hipLaunchKernelGGL((k_rhs_stream<0, false>), g, dim3(TDS_THR), 0, s, a, k0, k1, zc, nullptr);
    else if (rot == 1) hipLaunchKernelGGL((k_rhs_stream<1, false>), g, dim3(TDS_THR), tab, s, a, k0, k1, zc, nullptr);
    else hipLaunchKernelGGL((k_rhs_stream<2, false>), g, dim3(TDS_THR), 0, s, a, k0, k1, zc, nullptr);
    if (beside) (void)hipStreamWaitEvent(s, e1, 0);
    else if (fs.ncells) faces(s);
}
void heun(hipStream_t s, const HeunArgs& a) {
    unsigned b = nblk(a.N, 256); if (b > 4096) b = 4096; if (b == 0) b = 1;
    hipLaunchKernelGGL(k_heun, dim3(b), dim3(256), 0, s, a);
}
void compute_E(hipStream_t s, const EArgs& a) {
    hipLaunchKernelGGL(k_compute_E, grid_cells(a.g.n), dim3(256), 0, s, a);
}
}  // namespace fxk
