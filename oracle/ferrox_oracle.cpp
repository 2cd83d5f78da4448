/*
 * ORACLE -- test infrastructure only (the checker, never the thing measured or
 * shipped).  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it.  The product (ferrox_amd/) never includes, links or calls
 * anything under oracle/.
 *
 * FP64 CPU restatement of FerroX's per-timestep hot path, written from the
 * reference sources read as text:
 *   Source/main.cpp:48-623                     time loop (Euler/Heun, steady state, sweep)
 *   Source/FerroX.cpp:273-572                  keys, defaults, hard-coded Si constants
 *   Source/Solver/DerivativeAlgorithm.H         stencils (fx_oracle_stencils.h)
 *   Source/Solver/TotalEnergyDensity.cpp        Landau / gradient / electric RHS
 *   Source/Solver/ElectrostaticSolver.cpp       Poisson RHS, dF_dPhi, Newton RHS, E, eps, BCs,
 *                                              CheckSteadyState, ComputePhi_Rho
 *   Source/Solver/ChargeDensity.cpp             rho(phi), FD_half
 *   Source/Solver/Initialization.cpp:6-206      material mask, initial P and rho
 *   Source/Utils/eXstaticUtils/eXstaticUtil.cpp:105-143   face averaging of eps
 *
 * The Poisson solve (AMReX MLMG in the reference, not available here) is a
 * geometric multigrid written to the same specification as the device solver
 * (semi-coarsening rule, RB Gauss-Seidel, face-averaged coefficients,
 * cell-centred linear prolongation, exact (dense) bottom solve, MLMG stopping rule
 * ||r||_inf <= tol*max(||b||_inf, ||r0||_inf)); the tests pin it against a
 * sparse direct solve of the assembled operator.
 *
 * AMReX behaviour restated as assumptions (parity unpinned at that boundary):
 *   - Dirichlet ghost for max order 2 is the linear extrapolation 2*phi_b - phi_0
 *     with the boundary value located at the face;
 *   - after MLMG::solve the solution's ghost cells hold that extrapolation
 *     (phi_ghost_mode FX_GHOST_AMREX); FX_GHOST_FACE uses phi_b instead;
 *   - the random initial P uses this build's own counter-based generator
 *     (SplitMix64), not AMReX's RNG stream.
 */
#include "../include/ferrox_mi.h"
#include "fx_oracle_stencils.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

using fxo::BCFlags;

namespace {

/* ---- counter-based RNG shared by oracle and device path (definition) ---- */
inline uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
}
inline double uniform01(uint64_t seed, uint64_t idx) {
    uint64_t x = splitmix64((seed * 0xD1B54A32D192ED03ULL) ^ splitmix64(idx));
    return (double)(x >> 11) * 0x1.0p-53;
}

/* Hard-coded silicon constants, FerroX.cpp:534-554. */
struct SiConst {
    double Nc = 2.8e25, Nv = 1.83e25, bandgap = 1.12, affinity = 4.05;
    double q = 1.602e-19, kb = 1.38e-23, T = 300;
    double Ea = 44.0e-3, Ed = 46.0e-3;
};

/* Array with a one-cell ghost layer (used for the mask and eps, whose ghost
 * values follow AMReX's fill rules). */
struct G3 {
    int nx = 0, ny = 0, nz = 0;
    size_t sx = 0, sy = 0;
    std::vector<double> v;
    void init(int a, int b, int c, double val = 0.0) {
        nx = a; ny = b; nz = c; sx = nx + 2; sy = sx * (ny + 2);
        v.assign(sy * (nz + 2), val);
    }
    double& operator()(int i, int j, int k) { return v[(i + 1) + sx * (j + 1) + sy * (k + 1)]; }
    double operator()(int i, int j, int k) const { return v[(i + 1) + sx * (j + 1) + sy * (k + 1)]; }
};

/* FillBoundary with periodicity: a ghost cell whose periodic image is a
 * valid cell takes that value; other ghost cells keep theirs. */
void fill_periodic(G3& a, const int per[3]) {
    const int n[3] = {a.nx, a.ny, a.nz};
    for (int k = -1; k <= a.nz; ++k)
        for (int j = -1; j <= a.ny; ++j)
            for (int i = -1; i <= a.nx; ++i) {
                int idx[3] = {i, j, k};
                bool ghost = false, ok = true;
                for (int d = 0; d < 3; ++d) {
                    if (idx[d] < 0 || idx[d] >= n[d]) {
                        ghost = true;
                        if (per[d]) idx[d] = (idx[d] + n[d]) % n[d];
                        else ok = false;
                    }
                }
                if (ghost && ok) a(i, j, k) = a(idx[0], idx[1], idx[2]);
            }
}

/* ------------------------------ multigrid ------------------------------- */
struct Lev {
    int n[3];
    double h[3], ih2[3];
    int crs[3];                 /* directions coarsened to reach the next level */
    size_t N;
    std::vector<double> bx, by, bz, a, phi, b, r;
    size_t c(int i, int j, int k) const { return (size_t)i + (size_t)n[0] * ((size_t)j + (size_t)n[1] * k); }
    size_t fx(int i, int j, int k) const { return (size_t)i + (size_t)(n[0] + 1) * ((size_t)j + (size_t)n[1] * k); }
    size_t fy(int i, int j, int k) const { return (size_t)i + (size_t)n[0] * ((size_t)j + (size_t)(n[1] + 1) * k); }
    size_t fz(int i, int j, int k) const { return (size_t)i + (size_t)n[0] * ((size_t)j + (size_t)n[1] * k); }
};

struct MG {
    std::vector<Lev> L;
    int bclo[3], bchi[3];
    double ratio = 1.9;
    int nu1 = 2, nu2 = 2;
    double bottom_tol = 1e-6;
    int bottom_maxit = 400;
    long bottom_iters_total = 0;

    /* Off-diagonal sum S and diagonal D of row (i,j,k): (A phi)_c = D*phi_c - S.
     * bv: Dirichlet values [lo x,y,z, hi x,y,z] or nullptr (homogeneous). */
    inline void row(const Lev& l, const double* phi, int i, int j, int k, const double* bv,
                    double& D, double& S) const {
        const int n0 = l.n[0], n1 = l.n[1], n2 = l.n[2];
        D = -l.a[l.c(i, j, k)];
        S = 0.0;
        const int ijk[3] = {i, j, k};
        for (int d = 0; d < 3; ++d) {
            const int nd = l.n[d];
            double wlo, whi;
            if (d == 0) { wlo = l.bx[l.fx(i, j, k)] * l.ih2[0]; whi = l.bx[l.fx(i + 1, j, k)] * l.ih2[0]; }
            else if (d == 1) { wlo = l.by[l.fy(i, j, k)] * l.ih2[1]; whi = l.by[l.fy(i, j + 1, k)] * l.ih2[1]; }
            else { wlo = l.bz[l.fz(i, j, k)] * l.ih2[2]; whi = l.bz[l.fz(i, j, k + 1)] * l.ih2[2]; }
            /* low face */
            if (ijk[d] > 0 || bclo[d] == FX_BC_PERIODIC) {
                int m[3] = {i, j, k};
                m[d] = ijk[d] > 0 ? ijk[d] - 1 : nd - 1;
                D += wlo; S += wlo * phi[(size_t)m[0] + (size_t)n0 * ((size_t)m[1] + (size_t)n1 * m[2])];
            } else if (bclo[d] == FX_BC_DIRICHLET) {
                D += 2.0 * wlo; if (bv) S += 2.0 * wlo * bv[d];
            }
            /* high face */
            if (ijk[d] < nd - 1 || bchi[d] == FX_BC_PERIODIC) {
                int m[3] = {i, j, k};
                m[d] = ijk[d] < nd - 1 ? ijk[d] + 1 : 0;
                D += whi; S += whi * phi[(size_t)m[0] + (size_t)n0 * ((size_t)m[1] + (size_t)n1 * m[2])];
            } else if (bchi[d] == FX_BC_DIRICHLET) {
                D += 2.0 * whi; if (bv) S += 2.0 * whi * bv[3 + d];
            }
        }
        (void)n2;
    }

    void build(int depth_cap) {
        /* Levels after the finest are made by the coarsening rule: direction d
         * is coarsened when n_d is even and h_d < ratio*min(h) (the
         * mg_coarsen_ratio input, default 1.9).  A level with an
         * odd periodic extent > 1 (red-black colouring would break across the
         * wrap) or with no coarsenable direction is the bottom. */
        while ((int)L.size() < depth_cap) {
            Lev& f = L.back();
            bool odd_periodic = false;
            for (int d = 0; d < 3; ++d)
                if (bclo[d] == FX_BC_PERIODIC && f.n[d] > 1 && (f.n[d] & 1)) odd_periodic = true;
            if (odd_periodic) break;
            double hmin = std::min(f.h[0], std::min(f.h[1], f.h[2]));
            int any = 0;
            for (int d = 0; d < 3; ++d) {
                f.crs[d] = (f.n[d] % 2 == 0 && f.n[d] >= 2 && f.h[d] < ratio * hmin) ? 1 : 0;
                any |= f.crs[d];
            }
            if (!any || f.N <= 1) { f.crs[0] = f.crs[1] = f.crs[2] = 0; break; }
            Lev c;
            for (int d = 0; d < 3; ++d) {
                c.n[d] = f.crs[d] ? f.n[d] / 2 : f.n[d];
                c.h[d] = f.crs[d] ? 2.0 * f.h[d] : f.h[d];
                c.ih2[d] = 1.0 / (c.h[d] * c.h[d]);
                c.crs[d] = 0;
            }
            c.N = (size_t)c.n[0] * c.n[1] * c.n[2];
            c.bx.assign((size_t)(c.n[0] + 1) * c.n[1] * c.n[2], 0.0);
            c.by.assign((size_t)c.n[0] * (c.n[1] + 1) * c.n[2], 0.0);
            c.bz.assign((size_t)c.n[0] * c.n[1] * (c.n[2] + 1), 0.0);
            c.a.assign(c.N, 0.0);
            c.phi.assign(c.N, 0.0);
            c.b.assign(c.N, 0.0);
            c.r.assign(c.N, 0.0);
            L.push_back(std::move(c));
        }
        L.back().crs[0] = L.back().crs[1] = L.back().crs[2] = 0;
        restrict_faces();
    }

    /* Coarse face coefficient = mean of the fine faces that tile it. */
    void restrict_faces() {
        for (size_t l = 0; l + 1 < L.size(); ++l) {
            const Lev& f = L[l];
            Lev& c = L[l + 1];
            const int cx = f.crs[0], cy = f.crs[1], cz = f.crs[2];
            for (int k = 0; k < c.n[2]; ++k)
                for (int j = 0; j < c.n[1]; ++j)
                    for (int i = 0; i <= c.n[0]; ++i) {
                        double s = 0; int cnt = 0;
                        for (int kk = 0; kk <= cz; ++kk)
                            for (int jj = 0; jj <= cy; ++jj) {
                                s += f.bx[f.fx(cx ? 2 * i : i, cy ? 2 * j + jj : j, cz ? 2 * k + kk : k)]; ++cnt;
                            }
                        c.bx[c.fx(i, j, k)] = s / cnt;
                    }
            for (int k = 0; k < c.n[2]; ++k)
                for (int j = 0; j <= c.n[1]; ++j)
                    for (int i = 0; i < c.n[0]; ++i) {
                        double s = 0; int cnt = 0;
                        for (int kk = 0; kk <= cz; ++kk)
                            for (int ii = 0; ii <= cx; ++ii) {
                                s += f.by[f.fy(cx ? 2 * i + ii : i, cy ? 2 * j : j, cz ? 2 * k + kk : k)]; ++cnt;
                            }
                        c.by[c.fy(i, j, k)] = s / cnt;
                    }
            for (int k = 0; k <= c.n[2]; ++k)
                for (int j = 0; j < c.n[1]; ++j)
                    for (int i = 0; i < c.n[0]; ++i) {
                        double s = 0; int cnt = 0;
                        for (int jj = 0; jj <= cy; ++jj)
                            for (int ii = 0; ii <= cx; ++ii) {
                                s += f.bz[f.fz(cx ? 2 * i + ii : i, cy ? 2 * j + jj : j, cz ? 2 * k : k)]; ++cnt;
                            }
                        c.bz[c.fz(i, j, k)] = s / cnt;
                    }
        }
    }

    /* Coarse A-coefficient = mean of the fine cells it covers. */
    void restrict_acoef() {
        for (size_t l = 0; l + 1 < L.size(); ++l) {
            const Lev& f = L[l];
            Lev& c = L[l + 1];
            restrict_cells(f, c, f.a.data(), c.a.data());
        }
    }

    static void restrict_cells(const Lev& f, const Lev& c, const double* fine, double* crse) {
        const int cx = f.crs[0], cy = f.crs[1], cz = f.crs[2];
        const int cnt = (1 + cx) * (1 + cy) * (1 + cz);
        #pragma omp parallel for schedule(static)
        for (int k = 0; k < c.n[2]; ++k)
            for (int j = 0; j < c.n[1]; ++j)
                for (int i = 0; i < c.n[0]; ++i) {
                    double s = 0;
                    for (int kk = 0; kk <= cz; ++kk)
                        for (int jj = 0; jj <= cy; ++jj)
                            for (int ii = 0; ii <= cx; ++ii)
                                s += fine[f.c(cx ? 2 * i + ii : i, cy ? 2 * j + jj : j, cz ? 2 * k + kk : k)];
                    crse[c.c(i, j, k)] = s / cnt;
                }
    }

    void smooth(Lev& l, int color, const double* bv) {
        double* phi = l.phi.data();
        #pragma omp parallel for schedule(static)
        for (int k = 0; k < l.n[2]; ++k)
            for (int j = 0; j < l.n[1]; ++j)
                for (int i = ((j + k + color) & 1); i < l.n[0]; i += 2) {
                    double D, S;
                    row(l, phi, i, j, k, bv, D, S);
                    phi[l.c(i, j, k)] = (l.b[l.c(i, j, k)] + S) / D;
                }
    }

    void residual(Lev& l, const double* bv) {
        #pragma omp parallel for schedule(static)
        for (int k = 0; k < l.n[2]; ++k)
            for (int j = 0; j < l.n[1]; ++j)
                for (int i = 0; i < l.n[0]; ++i) {
                    double D, S;
                    row(l, l.phi.data(), i, j, k, bv, D, S);
                    size_t c = l.c(i, j, k);
                    l.r[c] = l.b[c] - (D * l.phi[c] - S);
                }
    }

    double norm_inf(const std::vector<double>& v) const {
        double m = 0.0;
        for (double x : v) m = std::max(m, std::fabs(x));
        return m;
    }

    /* Bottom solve of A e = rhs (homogeneous BCs).  A bottom of at most
     * BOTTOM_DIRECT_MAX cells is solved exactly: the dense matrix is assembled
     * column by column from the row operator, then Gaussian elimination with
     * partial pivoting (first largest |pivot| in row order; factor
     * f = M[i][k]/M[k][k]; M[i][j] -= f*M[k][j] for j > k; y[i] -= f*y[k]) and
     * back substitution by columns (from the last column down).  A pivot
     * below 1e-14 of the largest diagonal entry (singular operator: no
     * Dirichlet face) falls back to Jacobi-PCG, as do larger bottoms. */
    static constexpr int BOTTOM_DIRECT_MAX = 64;
    void bottom_solve(Lev& l, const double* rhs, double* e) {
        const int N = (int)l.N;
        if (N <= BOTTOM_DIRECT_MAX && bottom_direct(l, rhs, e)) return;
        bottom_cg(l, rhs, e);
    }
    bool bottom_direct(Lev& l, const double* rhs, double* e) {
        const int N = (int)l.N;
        std::vector<double> M((size_t)N * N), y(rhs, rhs + N), unit(N, 0.0);
        double dmax = 0.0;
        for (int col = 0; col < N; ++col) {
            unit[col] = 1.0;
            for (int k = 0; k < l.n[2]; ++k)
                for (int j = 0; j < l.n[1]; ++j)
                    for (int i = 0; i < l.n[0]; ++i) {
                        double D, S;
                        row(l, unit.data(), i, j, k, nullptr, D, S);
                        const int r = (int)l.c(i, j, k);
                        M[(size_t)r * N + col] = (r == col ? D : 0.0) - S;
                    }
            unit[col] = 0.0;
        }
        for (int r = 0; r < N; ++r) dmax = std::max(dmax, std::fabs(M[(size_t)r * N + r]));
        for (int k = 0; k < N; ++k) {
            int p = k;
            double best = std::fabs(M[(size_t)k * N + k]);
            for (int i = k + 1; i < N; ++i)
                if (std::fabs(M[(size_t)i * N + k]) > best) { best = std::fabs(M[(size_t)i * N + k]); p = i; }
            if (!(best > 1e-14 * dmax)) return false;
            if (p != k) {
                for (int j = k; j < N; ++j) std::swap(M[(size_t)k * N + j], M[(size_t)p * N + j]);
                std::swap(y[k], y[p]);
            }
            for (int i = k + 1; i < N; ++i) {
                const double f = M[(size_t)i * N + k] / M[(size_t)k * N + k];
                for (int j = k + 1; j < N; ++j) M[(size_t)i * N + j] = M[(size_t)i * N + j] - f * M[(size_t)k * N + j];
                y[i] = y[i] - f * y[k];
            }
        }
        /* back substitution by columns: e_k = y_k / U_kk, then y_i -= U_ik e_k for i < k */
        for (int k = N - 1; k >= 0; --k) {
            e[k] = y[k] / M[(size_t)k * N + k];
            for (int i = 0; i < k; ++i) y[i] = y[i] - M[(size_t)i * N + k] * e[k];
        }
        return true;
    }

    /* Jacobi-preconditioned CG on level l for A e = rhs (homogeneous BCs). */
    void bottom_cg(Lev& l, const double* rhs, double* e) {
        const size_t N = l.N;
        std::vector<double> r(rhs, rhs + N), z(N), p(N), q(N), Dg(N);
        std::vector<double> zero(N, 0.0);
        for (int k = 0; k < l.n[2]; ++k)
            for (int j = 0; j < l.n[1]; ++j)
                for (int i = 0; i < l.n[0]; ++i) {
                    double D, S; row(l, zero.data(), i, j, k, nullptr, D, S);
                    Dg[l.c(i, j, k)] = D;
                }
        std::fill(e, e + N, 0.0);
        double rr0 = 0; for (size_t c = 0; c < N; ++c) rr0 += r[c] * r[c];
        if (rr0 == 0.0) return;
        for (size_t c = 0; c < N; ++c) z[c] = r[c] / Dg[c];
        p = z;
        double rz = 0; for (size_t c = 0; c < N; ++c) rz += r[c] * z[c];
        const double tol2 = bottom_tol * bottom_tol * rr0;
        for (int it = 0; it < bottom_maxit; ++it) {
            for (int k = 0; k < l.n[2]; ++k)
                for (int j = 0; j < l.n[1]; ++j)
                    for (int i = 0; i < l.n[0]; ++i) {
                        double D, S; row(l, p.data(), i, j, k, nullptr, D, S);
                        size_t c = l.c(i, j, k);
                        q[c] = D * p[c] - S;
                    }
            double pq = 0; for (size_t c = 0; c < N; ++c) pq += p[c] * q[c];
            double al = rz / pq;
            double rr = 0;
            for (size_t c = 0; c < N; ++c) { e[c] += al * p[c]; r[c] -= al * q[c]; rr += r[c] * r[c]; }
            ++bottom_iters_total;
            if (rr <= tol2) break;
            double rzn = 0;
            for (size_t c = 0; c < N; ++c) { z[c] = r[c] / Dg[c]; rzn += r[c] * z[c]; }
            double be = rzn / rz; rz = rzn;
            for (size_t c = 0; c < N; ++c) p[c] = z[c] + be * p[c];
        }
    }

    /* Cell-centred linear prolongation of the coarse correction: in each
     * coarsened direction the fine cell takes 3/4 of its parent and 1/4 of the
     * parent's neighbour on its side; across a domain face that neighbour is
     * the homogeneous ghost (-parent for Dirichlet, parent for Neumann,
     * periodic image).  Terms are summed z-outer, x-inner, parent first. */
    void prolong(Lev& l, const Lev& c) {
        #pragma omp parallel for schedule(static)
        for (int k = 0; k < l.n[2]; ++k)
            for (int j = 0; j < l.n[1]; ++j)
                for (int i = 0; i < l.n[0]; ++i) {
                    const int f[3] = {i, j, k};
                    int id[3][2];
                    double w[3][2];
                    int cnt[3];
                    for (int d = 0; d < 3; ++d) {
                        if (!l.crs[d]) { id[d][0] = f[d]; w[d][0] = 1.0; cnt[d] = 1; continue; }
                        const int I = f[d] >> 1;
                        int J = (f[d] & 1) ? I + 1 : I - 1;
                        double sg = 1.0;
                        if (J < 0 || J >= c.n[d]) {
                            const int bct = J < 0 ? bclo[d] : bchi[d];
                            if (bct == FX_BC_PERIODIC) J = (J + c.n[d]) % c.n[d];
                            else { sg = (bct == FX_BC_DIRICHLET) ? -1.0 : 1.0; J = I; }
                        }
                        id[d][0] = I; w[d][0] = 0.75;
                        id[d][1] = J; w[d][1] = 0.25 * sg;
                        cnt[d] = 2;
                    }
                    double e = 0.0;
                    for (int a = 0; a < cnt[2]; ++a)
                        for (int b = 0; b < cnt[1]; ++b)
                            for (int q = 0; q < cnt[0]; ++q)
                                e += (w[2][a] * w[1][b]) * w[0][q] * c.phi[c.c(id[0][q], id[1][b], id[2][a])];
                    l.phi[l.c(i, j, k)] += e;
                }
    }

    /* The level the V-cycle solves exactly instead of cycling below it (-1:
     * none): on a hierarchy periodic in x and y and bounded in z with a
     * Dirichlet face, the finest level below the finest with n0 n1 <= 4096,
     * n0, n1 <= 64 and at most 65536 cells, when its face coefficients depend
     * on z only and it has no Helmholtz term.  The device solves that level
     * with lateral Fourier transforms and a tridiagonal solve per mode; here
     * it is solved by repeated V-cycles of the levels below until the
     * residual stops falling, which is the same solution to rounding. */
    int exact_tail = 1;
    int exact_level() const {
        if (!exact_tail || bclo[0] != FX_BC_PERIODIC || bclo[1] != FX_BC_PERIODIC || bclo[2] == FX_BC_PERIODIC) return -1;
        if (bclo[2] != FX_BC_DIRICHLET && bchi[2] != FX_BC_DIRICHLET) return -1;
        int t = -1;
        for (size_t l = 1; l < L.size() && t < 0; ++l)
            if (L[l].n[0] <= 64 && L[l].n[1] <= 64 && L[l].n[0] * L[l].n[1] <= 4096 && L[l].N <= 65536) t = (int)l;
        if (t < 0) return -1;
        const Lev& v = L[t];
        for (size_t c = 0; c < v.N; ++c) if (v.a[c] != 0.0) return -1;
        for (int k = 0; k < v.n[2]; ++k) {
            const double bx0 = v.bx[v.fx(0, 0, k)], by0 = v.by[v.fy(0, 0, k)];
            for (int j = 0; j < v.n[1]; ++j)
                for (int i = 0; i <= v.n[0]; ++i) if (v.bx[v.fx(i, j, k)] != bx0) return -1;
            for (int j = 0; j <= v.n[1]; ++j)
                for (int i = 0; i < v.n[0]; ++i) if (v.by[v.fy(i, j, k)] != by0) return -1;
        }
        for (int k = 0; k <= v.n[2]; ++k) {
            const double bz0 = v.bz[v.fz(0, 0, k)];
            for (int j = 0; j < v.n[1]; ++j)
                for (int i = 0; i < v.n[0]; ++i) if (v.bz[v.fz(i, j, k)] != bz0) return -1;
        }
        return t;
    }
    int exact_lvl = -1;          /* exact_level() of the current solve */
    void exact_solve(size_t li) {
        Lev& l = L[li];
        std::fill(l.phi.begin(), l.phi.end(), 0.0);
        const double bn = norm_inf(l.b);
        if (bn == 0.0) return;
        double prev = bn;
        for (int it = 0; it < 100; ++it) {
            vcycle(li, nullptr, false);
            residual(l, nullptr);
            const double r = norm_inf(l.r);
            if (r <= 1e-15 * bn || r >= prev) break;
            prev = r;
        }
    }

    void vcycle(size_t li, const double* bv0, bool allow_exact = true) {
        Lev& l = L[li];
        const double* bv = (li == 0) ? bv0 : nullptr;
        if (allow_exact && li > 0 && (int)li == exact_lvl) {
            exact_solve(li);
            return;
        }
        if (li + 1 == L.size()) {
            if (li == 0) {     /* single-level hierarchy: correct the solution */
                residual(l, bv);
                std::vector<double> e(l.N);
                bottom_solve(l, l.r.data(), e.data());
                for (size_t c = 0; c < l.N; ++c) l.phi[c] += e[c];
            } else {
                bottom_solve(l, l.b.data(), l.phi.data());
            }
            return;
        }
        for (int s = 0; s < nu1; ++s) { smooth(l, 0, bv); smooth(l, 1, bv); }
        residual(l, bv);
        Lev& c = L[li + 1];
        restrict_cells(l, c, l.r.data(), c.b.data());
        std::fill(c.phi.begin(), c.phi.end(), 0.0);
        vcycle(li + 1, bv0, allow_exact);
        prolong(l, c);
        for (int s = 0; s < nu2; ++s) { smooth(l, 0, bv); smooth(l, 1, bv); }
    }

    /* MLMG-style solve on level 0: phi (in/out), b, Dirichlet values bv. */
    int solve(double tol_rel, int maxit, const double* bv, fx_solve_stats* st) {
        exact_lvl = exact_level();
        Lev& l = L[0];
        residual(l, bv);
        double r0 = norm_inf(l.r), bn = norm_inf(l.b);
        double target = tol_rel * std::max(bn, r0);
        st->mg_resid0 = r0; st->mg_bnorm = bn; st->mg_resid = r0; st->vcycles = 0;
        if (!(r0 == r0) || !(bn == bn)) return FX_ENAN;
        if (r0 <= target) return FX_OK;
        for (int it = 1; it <= maxit; ++it) {
            vcycle(0, bv);
            residual(l, bv);
            double rn = norm_inf(l.r);
            st->mg_resid = rn; st->vcycles = it;
            if (!(rn == rn)) return FX_ENAN;
            if (rn <= target) return FX_OK;
        }
        return FX_ENOCONV_MG;
    }
};

}  // namespace

/* ------------------------------ oracle state ------------------------------ */
struct fxo_state {
    fx_inputs in;
    SiConst si;
    int n[3];
    size_t N;
    double dx[3];
    G3 mask, eps;
    std::vector<double> tphase, ang[3], lalpha;
    std::vector<double> P[3], Gam, rho, e_den, h_den, phi, phi_old, phi_prev, alpha_cc, rhs, phidiff;
    std::vector<double> E[3], GL[3];
    MG mg;
    /* run state */
    double Phi_Bc_lo, Phi_Bc_hi;
    double bc_solve[6];            /* Dirichlet values the last solve used (x,y,z lo, x,y,z hi) */
    int step = 0, sign = 1, num_Vapp = 0, steady_state_step = 1000000, inc_step;
    double time = 0.0;
    std::vector<double> waveform;  /* per-step top bias (extension) */
    BCFlags pbc;
    std::string err;
    int rot_uniform = 1;
    double Ru[9];

    size_t c(int i, int j, int k) const { return (size_t)i + (size_t)n[0] * ((size_t)j + (size_t)n[1] * k); }

    double phi_ref() const {
        return si.affinity + 0.5 * si.bandgap + 0.5 * si.kb * si.T * log(si.Nc / si.Nv) / si.q;
    }
};

namespace {

/* P accessor with the reference's ghost semantics: periodic image in
 * periodic directions, 0 beyond non-periodic faces (P ghosts are set to 0 at
 * main.cpp:157 and never written). */
struct PAcc {
    const double* p; int n[3]; int per[3];
    double operator()(int i, int j, int k) const {
        int idx[3] = {i, j, k};
        for (int d = 0; d < 3; ++d) {
            if (idx[d] < 0 || idx[d] >= n[d]) {
                if (!per[d]) return 0.0;
                idx[d] = ((idx[d] % n[d]) + n[d]) % n[d];
            }
        }
        return p[(size_t)idx[0] + (size_t)n[0] * ((size_t)idx[1] + (size_t)n[1] * idx[2])];
    }
};
struct MAcc {
    const G3* m;
    double operator()(int i, int j, int k) const {
        if (i < -1 || j < -1 || k < -1 || i > m->nx || j > m->ny || k > m->nz) return 1.0;
        return (*m)(i, j, k);
    }
};

bool in_box(const double lo[3], const double hi[3], double x, double y, double z) {
    return x <= hi[0] && x >= lo[0] && y <= hi[1] && y >= lo[1] && z <= hi[2] && z >= lo[2];
}

/* InitializeMaterialMask, Initialization.cpp:169-206: codes on the grown box
 * from cell-centre coordinates (FE > DE > SC(> Channel), spacer = DE), then
 * the periodic fill. */
void build_mask(fxo_state* s) {
    const fx_inputs& in = s->in;
    s->mask.init(s->n[0], s->n[1], s->n[2], 1.0);
    for (int k = -1; k <= s->n[2]; ++k)
        for (int j = -1; j <= s->n[1]; ++j)
            for (int i = -1; i <= s->n[0]; ++i) {
                double x = in.prob_lo[0] + (i + 0.5) * s->dx[0];
                double y = in.prob_lo[1] + (j + 0.5) * s->dx[1];
                double z = in.prob_lo[2] + (k + 0.5) * s->dx[2];
                double m;
                if (in_box(in.FE_lo, in.FE_hi, x, y, z)) m = 0.;
                else if (in_box(in.DE_lo, in.DE_hi, x, y, z)) m = 1.;
                else if (in_box(in.SC_lo, in.SC_hi, x, y, z)) {
                    m = 2.;
                    if (in_box(in.Channel_lo, in.Channel_hi, x, y, z)) m = 3.;
                } else m = 1.;
                s->mask(i, j, k) = m;
            }
    fill_periodic(s->mask, in.is_periodic);
}

/* Ghost layer of a mask whose valid region was set by the caller: the ghost
 * cells keep the coordinate rule and the periodic images. */
void rebuild_mask_ghosts(fxo_state* s) {
    G3 valid = s->mask;
    build_mask(s);
    for (int k = 0; k < s->n[2]; ++k)
        for (int j = 0; j < s->n[1]; ++j)
            for (int i = 0; i < s->n[0]; ++i) s->mask(i, j, k) = valid(i, j, k);
    fill_periodic(s->mask, s->in.is_periodic);
}

/* InitializePermittivity (ElectrostaticSolver.cpp:205-299) and the face
 * averages (eXstaticUtil.cpp:105-143) into the finest multigrid level. */
void build_eps(fxo_state* s) {
    const fx_inputs& in = s->in;
    s->eps.init(s->n[0], s->n[1], s->n[2], 0.0);
    for (int k = 0; k < s->n[2]; ++k)
        for (int j = 0; j < s->n[1]; ++j)
            for (int i = 0; i < s->n[0]; ++i) {
                double m = s->mask(i, j, k), e;
                if (m == 0.0) {
                    e = in.epsilonX_fe * in.epsilon_0;
                    if (s->tphase[s->c(i, j, k)] == 1.0) e = in.epsilonX_fe_tphase * in.epsilon_0;
                } else if (m == 1.0) e = in.epsilon_de * in.epsilon_0;
                else if (m >= 2.0) e = in.epsilon_si * in.epsilon_0;
                else e = in.epsilon_de * in.epsilon_0;
                s->eps(i, j, k) = e;
            }
    fill_periodic(s->eps, in.is_periodic);
    const int nx = s->n[0], ny = s->n[1], nz = s->n[2];
    for (int k = -1; k <= nz; ++k)
        for (int j = -1; j <= ny; ++j)
            for (int i = -1; i <= nx; ++i) {
                if (in.bc_lo[0] != FX_BC_PERIODIC && i < 0) s->eps(i, j, k) = s->eps(i + 1, j, k);
                if (in.bc_hi[0] != FX_BC_PERIODIC && i > nx - 1) s->eps(i, j, k) = s->eps(i - 1, j, k);
                if (in.bc_lo[1] != FX_BC_PERIODIC && j < 0) s->eps(i, j, k) = s->eps(i, j + 1, k);
                if (in.bc_hi[1] != FX_BC_PERIODIC && j > ny - 1) s->eps(i, j, k) = s->eps(i, j - 1, k);
                if (in.bc_lo[2] != FX_BC_PERIODIC && k < 0) s->eps(i, j, k) = s->eps(i, j, k + 1);
                if (in.bc_hi[2] != FX_BC_PERIODIC && k > nz - 1) s->eps(i, j, k) = s->eps(i, j, k - 1);
            }
    Lev& l = s->mg.L[0];
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i <= nx; ++i) l.bx[l.fx(i, j, k)] = 0.5 * (s->eps(i, j, k) + s->eps(i - 1, j, k));
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j <= ny; ++j)
            for (int i = 0; i < nx; ++i) l.by[l.fy(i, j, k)] = 0.5 * (s->eps(i, j, k) + s->eps(i, j - 1, k));
    for (int k = 0; k <= nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) l.bz[l.fz(i, j, k)] = 0.5 * (s->eps(i, j, k) + s->eps(i, j, k - 1));
    s->mg.restrict_faces();
}

void rotation_at(const fxo_state* s, size_t c, double R[9]) {
    if (s->rot_uniform) { std::memcpy(R, s->Ru, sizeof(s->Ru)); return; }
    fxo::rotation(s->ang[0][c], s->ang[1][c], s->ang[2][c], s->in.use_Euler_angles, R);
}

void refresh_rotation(fxo_state* s) {
    s->rot_uniform = 1;
    for (size_t c = 0; c < s->N; ++c)
        if (s->ang[0][c] != s->ang[0][0] || s->ang[1][c] != s->ang[1][0] || s->ang[2][c] != s->ang[2][0]) {
            s->rot_uniform = 0; break;
        }
    fxo::rotation(s->ang[0][0], s->ang[1][0], s->ang[2][0], s->in.use_Euler_angles, s->Ru);
}

/* ComputeRho, ChargeDensity.cpp:15-102 (FD_half :4-11). */
double FD_half(double eta) {
    double nu = std::pow(eta, 4.0) + 50.0 + 33.6 * eta * (1.0 - 0.68 * exp(-0.17 * std::pow((eta + 1.0), 2.0)));
    double xi = 3.0 * sqrt(3.14) / (4.0 * std::pow(nu, 3. / 8.));
    return std::pow(exp(-eta) + xi, -1.0);
}

void compute_rho(fxo_state* s, const std::vector<double>& phi) {
    const fx_inputs& in = s->in;
    const SiConst& k = s->si;
    const double q = k.q, kb = k.kb, T = k.T, Nc = k.Nc, Nv = k.Nv;
    const double Ef = 0.0, Eg = k.bandgap, Chi = k.affinity;
    const double phi_ref = Chi + 0.5 * Eg + 0.5 * kb * T * log(Nc / Nv) / q;
    const double g_A = 4.0, g_D = 2.0, Ea = k.Ea, Ed = k.Ed;
    #pragma omp parallel for schedule(static)
    for (int kk = 0; kk < s->n[2]; ++kk)
        for (int j = 0; j < s->n[1]; ++j)
            for (int i = 0; i < s->n[0]; ++i) {
                size_t c = s->c(i, j, kk);
                if (s->mask(i, j, kk) >= 2.0) {
                    double Ec_corr = -q * (phi[c] - phi_ref) - Chi * q;
                    double Ev_corr = Ec_corr - q * Eg;
                    double n, p;
                    if (in.use_Fermi_Dirac == 1) {
                        double eta_n = -(Ec_corr - q * Ef) / (kb * T);
                        double eta_p = -(q * Ef - Ev_corr) / (kb * T);
                        n = Nc * FD_half(eta_n);
                        p = Nv * FD_half(eta_p);
                    } else {
                        n = Nc * exp(-(Ec_corr - q * Ef) / (kb * T));
                        p = Nv * exp(-(q * Ef - Ev_corr) / (kb * T));
                    }
                    double NA = in.acceptor_doping / (1.0 + g_A * exp((-q * Ef + q * Ea + q * phi_ref - q * Chi - q * Eg - q * phi[c]) / (kb * T)));
                    double ND = in.donor_doping / (1.0 + g_D * exp((q * Ef + q * Ed - q * phi_ref + q * Chi + q * phi[c]) / (kb * T)));
                    s->e_den[c] = n;
                    s->h_den[c] = p;
                    s->rho[c] = q * (p - n - NA + ND);
                } else {
                    s->rho[c] = 0.0;
                }
            }
}

/* ComputePoissonRHS, ElectrostaticSolver.cpp:8-83. */
void poisson_rhs(fxo_state* s, std::vector<double>* const P, const std::vector<double>& rho, std::vector<double>& out) {
    PAcc Fp{P[0].data(), {s->n[0], s->n[1], s->n[2]}, {s->in.is_periodic[0], s->in.is_periodic[1], s->in.is_periodic[2]}};
    PAcc Fq = Fp, Fr = Fp; Fq.p = P[1].data(); Fr.p = P[2].data();
    MAcc M{&s->mask};
    #pragma omp parallel for schedule(static)
    for (int k = 0; k < s->n[2]; ++k) {
        BCFlags bc = s->pbc;
        for (int j = 0; j < s->n[1]; ++j)
            for (int i = 0; i < s->n[0]; ++i) {
                size_t c = s->c(i, j, k);
                double m = s->mask(i, j, k);
                if (m >= 2.0) out[c] = rho[c];
                else if (m == 1.0) out[c] = 0.;
                else {
                    double R[9]; rotation_at(s, c, R);
                    out[c] = -(R[0] * fxo::DPD(Fp, M, i, j, k, 0, s->dx, bc) + R[1] * fxo::DPD(Fp, M, i, j, k, 1, s->dx, bc) + R[2] * fxo::DPD(Fp, M, i, j, k, 2, s->dx, bc))
                             - (R[3] * fxo::DPD(Fq, M, i, j, k, 0, s->dx, bc) + R[4] * fxo::DPD(Fq, M, i, j, k, 1, s->dx, bc) + R[5] * fxo::DPD(Fq, M, i, j, k, 2, s->dx, bc))
                             - (R[6] * fxo::DPD(Fr, M, i, j, k, 0, s->dx, bc) + R[7] * fxo::DPD(Fr, M, i, j, k, 1, s->dx, bc) + R[8] * fxo::DPD(Fr, M, i, j, k, 2, s->dx, bc));
                }
            }
        if (bc.bad) s->pbc.bad = true;
    }
}

void current_bc_values(const fxo_state* s, double bv[6]) {
    const double phi_ref = s->phi_ref();
    const double phi_m = s->in.use_work_function ? s->in.metal_work_function : phi_ref;
    for (int d = 0; d < 2; ++d) { bv[d] = 0.0; bv[3 + d] = 0.0; }
    bv[2] = s->Phi_Bc_lo;
    bv[5] = s->Phi_Bc_hi - (phi_m - phi_ref);
}

/* Ghost value of phi next to cell (i,j,k) across a domain face, as
 * ComputeEfromPhi reads it (ElectrostaticSolver.cpp:195-197). */
double phi_at(const fxo_state* s, int i, int j, int k) {
    int idx[3] = {i, j, k};
    for (int d = 0; d < 3; ++d) {
        if (idx[d] < 0 || idx[d] >= s->n[d]) {
            const bool lo = idx[d] < 0;
            const int bct = lo ? s->in.bc_lo[d] : s->in.bc_hi[d];
            if (bct == FX_BC_PERIODIC) { idx[d] = ((idx[d] % s->n[d]) + s->n[d]) % s->n[d]; continue; }
            int in_idx[3] = {idx[0], idx[1], idx[2]};
            in_idx[d] = lo ? 0 : s->n[d] - 1;
            for (int e = 0; e < 3; ++e)
                if (e != d && s->in.bc_lo[e] == FX_BC_PERIODIC) in_idx[e] = ((in_idx[e] % s->n[e]) + s->n[e]) % s->n[e];
            double p0 = s->phi[s->c(in_idx[0], in_idx[1], in_idx[2])];
            if (bct == FX_BC_NEUMANN) return p0;
            double pb = s->bc_solve[(lo ? 0 : 3) + d];
            return s->in.phi_ghost_mode == FX_GHOST_FACE ? pb : 2.0 * pb - p0;
        }
    }
    return s->phi[s->c(idx[0], idx[1], idx[2])];
}

/* ComputeEfromPhi, ElectrostaticSolver.cpp:134-203 (DphiDz, DFDx, DFDy at
 * DerivativeAlgorithm.H:8-40). */
void compute_E(fxo_state* s) {
    const double* dx = s->dx;
    const double plo = s->in.prob_lo[2], phi_hi = s->in.prob_hi[2];
    #pragma omp parallel for schedule(static)
    for (int k = 0; k < s->n[2]; ++k)
        for (int j = 0; j < s->n[1]; ++j)
            for (int i = 0; i < s->n[0]; ++i) {
                size_t c = s->c(i, j, k);
                double z_hi = plo + (k + 1.5) * dx[2];
                double z_lo = plo + (k - 0.5) * dx[2];
                double dfdx = (phi_at(s, i + 1, j, k) - phi_at(s, i - 1, j, k)) / (2. * dx[0]);
                double dfdy = (phi_at(s, i, j + 1, k) - phi_at(s, i, j - 1, k)) / (2. * dx[1]);
                double dfdz;
                if (z_lo < plo) dfdz = (-4. * phi_at(s, i, j, k - 1) + 3. * phi_at(s, i, j, k) + phi_at(s, i, j, k + 1)) / (3. * dx[2]);
                else if (z_hi > phi_hi) dfdz = (4. * phi_at(s, i, j, k + 1) - 3. * phi_at(s, i, j, k) - phi_at(s, i, j, k - 1)) / (3. * dx[2]);
                else dfdz = (phi_at(s, i, j, k + 1) - phi_at(s, i, j, k - 1)) / (2. * dx[2]);
                double R[9]; rotation_at(s, c, R);
                s->E[0][c] = -(R[0] * dfdx + R[1] * dfdy + R[2] * dfdz);
                s->E[1][c] = -(R[3] * dfdx + R[4] * dfdy + R[5] * dfdz);
                s->E[2][c] = -(R[6] * dfdx + R[7] * dfdy + R[8] * dfdz);
            }
}

/* Landau + gradient + electric terms and their sum (TotalEnergyDensity.cpp).
 * Requires E (compute_E) for the current phi. */
void tdgl_rhs(fxo_state* s, std::vector<double>* const P, std::vector<double>* GL) {
    const fx_inputs& in = s->in;
    PAcc Fp{P[0].data(), {s->n[0], s->n[1], s->n[2]}, {in.is_periodic[0], in.is_periodic[1], in.is_periodic[2]}};
    PAcc Fq = Fp, Fr = Fp; Fq.p = P[1].data(); Fr.p = P[2].data();
    MAcc M{&s->mask};
    const double beta = in.beta, gamma = in.gamma;
    const double a12 = in.alpha_12, a112 = in.alpha_112, a123 = in.alpha_123;
    const double g11 = in.g11, g44 = in.g44, g12 = in.g12, g44_p = in.g44_p;
    const double* dx = s->dx;
    #pragma omp parallel for schedule(static)
    for (int k = 0; k < s->n[2]; ++k) {
        BCFlags bc = s->pbc;
        for (int j = 0; j < s->n[1]; ++j)
            for (int i = 0; i < s->n[0]; ++i) {
                size_t c = s->c(i, j, k);
                const double Gam = s->Gam[c];
                const double tph = s->tphase[c];
                const double alpha = s->lalpha[c];   /* per-cell Landau alpha (extension; default: inputs' alpha) */
                const double p = P[0][c], q = P[1][c], r = P[2][c];
                double L[3] = {0, 0, 0}, G[3] = {0, 0, 0}, El[3] = {0, 0, 0};
                if (in.include_Landau == 1) {
                    /* Calculate_Landau, :86-128 (powers written as products) */
                    double p2 = p * p, q2 = q * q, r2 = r * r;
                    double dFp = alpha * p + beta * (p2 * p) + gamma * (p2 * p2 * p)
                               + 2. * a12 * p * q2 + 2. * a12 * p * r2
                               + 4. * a112 * (p2 * p) * (q2 + r2)
                               + 2. * a112 * p * (q2 * q2) + 2. * a112 * p * (r2 * r2)
                               + 2. * a123 * p * q2 * r2;
                    double dFq = alpha * q + beta * (q2 * q) + gamma * (q2 * q2 * q)
                               + 2. * a12 * q * p2 + 2. * a12 * q * r2
                               + 4. * a112 * (q2 * q) * (p2 + r2)
                               + 2. * a112 * q * (p2 * p2) + 2. * a112 * q * (r2 * r2)
                               + 2. * a123 * q * p2 * r2;
                    double dFr = alpha * r + beta * (r2 * r) + gamma * (r2 * r2 * r)
                               + 2. * a12 * r * p2 + 2. * a12 * r * q2
                               + 4. * a112 * (r2 * r) * (p2 + q2)
                               + 2. * a112 * r * (p2 * p2) + 2. * a112 * r * (q2 * q2)
                               + 2. * a123 * r * p2 * q2;
                    L[0] = -1.0 * Gam * dFp; L[1] = -1.0 * Gam * dFq; L[2] = -1.0 * Gam * dFr;
                    if (in.is_polarization_scalar == 1) { L[0] = 0.0; L[1] = 0.0; }
                    if (tph == 1.0) { L[0] = L[1] = L[2] = 0.0; }
                }
                if (in.include_Grad == 1) {
                    /* Calculate_Grad, :164-249 */
                    double R[9]; rotation_at(s, c, R);
                    const double R11 = R[0], R12 = R[1], R13 = R[2], R21 = R[3], R22 = R[4], R23 = R[5], R31 = R[6], R32 = R[7], R33 = R[8];
                    double dFp = 0, dFq = 0;
                    if (in.is_polarization_scalar != 1) {
                        dFp = -g11 * fxo::DDPD(Fp, M, i, j, k, 0, dx, bc)
                              - (g44 + g44_p) * fxo::DDPD(Fp, M, i, j, k, 1, dx, bc)
                              - (g44 + g44_p) * fxo::DDPD(Fp, M, i, j, k, 2, dx, bc)
                              - (g12 + g44 - g44_p) * fxo::DDPD_mixed(Fq, M, i, j, k, 0, 1, dx, bc)
                              - (g12 + g44 - g44_p) * fxo::DDPD_mixed(Fr, M, i, j, k, 0, 2, dx, bc);
                        dFq = -g11 * fxo::DDPD(Fq, M, i, j, k, 1, dx, bc)
                              - (g44 - g44_p) * fxo::DDPD(Fq, M, i, j, k, 0, dx, bc)
                              - (g44 - g44_p) * fxo::DDPD(Fq, M, i, j, k, 2, dx, bc)
                              - (g12 + g44 + g44_p) * fxo::DDPD_mixed(Fp, M, i, j, k, 0, 1, dx, bc)
                              - (g12 + g44 - g44_p) * fxo::DDPD_mixed(Fr, M, i, j, k, 1, 2, dx, bc);
                    }
                    double Hxx = fxo::DDPD(Fr, M, i, j, k, 0, dx, bc);
                    double Hyy = fxo::DDPD(Fr, M, i, j, k, 1, dx, bc);
                    double Hzz = fxo::DDPD(Fr, M, i, j, k, 2, dx, bc);
                    double Hxy = fxo::DDPD_mixed(Fr, M, i, j, k, 0, 1, dx, bc);
                    double Hyz = fxo::DDPD_mixed(Fr, M, i, j, k, 1, 2, dx, bc);
                    double Hxz = fxo::DDPD_mixed(Fr, M, i, j, k, 0, 2, dx, bc);
                    double dFr = -g11 * (R31 * R31 * Hxx + R32 * R32 * Hyy + R33 * R33 * Hzz
                                         + 2. * R31 * R32 * Hxy + 2. * R32 * R33 * Hyz + 2. * R33 * R31 * Hxz)
                                 - (g44 - g44_p) * (R11 * R11 * Hxx + R12 * R12 * Hyy + R13 * R13 * Hzz
                                                    + 2. * R11 * R12 * Hxy + 2. * R12 * R13 * Hyz + 2. * R13 * R11 * Hxz)
                                 - (g44 - g44_p) * (R21 * R21 * Hxx + R22 * R22 * Hyy + R23 * R23 * Hzz
                                                    + 2. * R21 * R22 * Hxy + 2. * R22 * R23 * Hyz + 2. * R23 * R21 * Hxz);
                    dFr = dFr - (g44 + g44_p + g12) * fxo::DDPD_mixed(Fq, M, i, j, k, 1, 2, dx, bc)
                              - (g44 + g44_p + g12) * fxo::DDPD_mixed(Fp, M, i, j, k, 0, 2, dx, bc);
                    G[0] = -1.0 * Gam * dFp; G[1] = -1.0 * Gam * dFq; G[2] = -1.0 * Gam * dFr;
                    if (in.is_polarization_scalar == 1) { G[0] = 0.0; G[1] = 0.0; }
                    if (tph == 1.0) { G[0] = G[1] = G[2] = 0.0; }
                }
                if (in.include_Elec == 1) {
                    /* Calculate_Elec, :273-292 */
                    El[0] = -1.0 * Gam * (-s->E[0][c]);
                    El[1] = -1.0 * Gam * (-s->E[1][c]);
                    El[2] = -1.0 * Gam * (-s->E[2][c]);
                    if (in.is_polarization_scalar == 1) { El[0] = 0.0; El[1] = 0.0; }
                    if (tph == 1.0) { El[0] = El[1] = El[2] = 0.0; }
                }
                /* CalculateTDGL_RHS, :44-62: GL = 0 (+ L) (+ G) (+ E) */
                for (int d = 0; d < 3; ++d) {
                    double g = 0.0;
                    if (in.include_Landau == 1) g += L[d];
                    if (in.include_Grad == 1) g += G[d];
                    if (in.include_Elec == 1) g += El[d];
                    GL[d][c] = g;
                }
            }
        if (bc.bad) s->pbc.bad = true;
    }
}

bool contains_sc(const fxo_state* s) {
    for (int k = 0; k < s->n[2]; ++k)
        for (int j = 0; j < s->n[1]; ++j)
            for (int i = 0; i < s->n[0]; ++i)
                if (s->mask(i, j, k) >= 2.0) return true;
    return false;
}

/* ComputePhi_Rho, ElectrostaticSolver.cpp:678-747 (with ComputePoissonRHS at
 * the head of each outer iteration). */
int compute_phi_rho(fxo_state* s, std::vector<double>* const P, fx_solve_stats* st) {
    const fx_inputs& in = s->in;
    const double tol = 1.e-5;
    double err = 1.0;
    int iter = 0;
    const bool sc = contains_sc(s);
    std::vector<double> rhs_d(s->N), phi_d(s->N);
    Lev& l0 = s->mg.L[0];
    double bv[6];
    current_bc_values(s, bv);
    std::memcpy(s->bc_solve, bv, sizeof(bv));
    fx_solve_stats one{};
    st->outer_iters = 0; st->vcycles = 0; st->outer_err = 0;
    while (err > tol) {
        poisson_rhs(s, P, s->rho, s->rhs);
        /* dF_dPhi, :85-113 */
        for (size_t c = 0; c < s->N; ++c) phi_d[c] = s->phi[c] + in.delta;
        compute_rho(s, phi_d);
        poisson_rhs(s, P, s->rho, rhs_d);
        for (size_t c = 0; c < s->N; ++c) s->alpha_cc[c] = (1. / in.delta) * rhs_d[c] + (-1. / in.delta) * s->rhs[c];
        /* ComputePoissonRHS_Newton, :114-132 */
        for (size_t c = 0; c < s->N; ++c) s->rhs[c] = s->rhs[c] - s->alpha_cc[c] * s->phi[c];
        /* setACoeffs + zero initial guess + solve, :713-719 */
        l0.a = s->alpha_cc;
        s->mg.restrict_acoef();
        if (!in.mg_warm_start) std::fill(s->phi.begin(), s->phi.end(), 0.0);
        l0.phi = s->phi;
        l0.b = s->rhs;
        int rc = s->mg.solve(in.mg_tol_rel, in.mg_max_iter, bv, &one);
        s->phi = l0.phi;
        st->vcycles += one.vcycles;
        st->mg_resid0 = one.mg_resid0; st->mg_resid = one.mg_resid; st->mg_bnorm = one.mg_bnorm;
        if (rc != FX_OK) { st->outer_iters = iter + 1; return rc; }
        compute_rho(s, s->phi);
        if (!sc) {
            err = 0.;
            st->outer_iters = 1;
        } else {
            if (iter > 0) {
                double num = 0, den = 0;
                for (size_t c = 0; c < s->N; ++c) { num += std::fabs(s->phi[c] - s->phi_prev[c]); den += std::fabs(s->phi[c]); }
                err = num / den;
            }
            s->phi_prev = s->phi;
            iter = iter + 1;
            st->outer_iters = iter;
            st->outer_err = err;
            if (iter >= in.outer_max_iter && err > tol) return FX_ENOCONV_OUTER;
        }
    }
    if (s->pbc.bad) return FX_EINVAL;
    return FX_OK;
}

/* CheckSteadyState, ElectrostaticSolver.cpp:526-560. */
double check_steady_state(fxo_state* s, int step) {
    double phi_max = 0.0;
    for (double v : s->phi) phi_max = std::max(phi_max, std::fabs(v));
    double m = 0.0;
    for (size_t c = 0; c < s->N; ++c) {
        s->phidiff[c] = std::fabs(s->phi[c] - s->phi_old[c]) / phi_max;
        m = std::max(m, s->phidiff[c]);
    }
    if (phi_max == 0.0) m = NAN;
    if (step > 1) {
        if (m < s->in.phi_tolerance) { s->steady_state_step = step; s->inc_step = step; }
    }
    s->phi_old = s->phi;
    return m;
}

void lincomb(std::vector<double>& dst, double a, const std::vector<double>& x, double b, const std::vector<double>& y) {
    for (size_t c = 0; c < dst.size(); ++c) dst[c] = a * x[c] + b * y[c];
}

/* One update of P (main.cpp:279-348). */
int update(fxo_state* s, int order, fx_solve_stats* st2) {
    compute_E(s);
    tdgl_rhs(s, s->P, s->GL);
    std::vector<double> Pn[3];
    for (int d = 0; d < 3; ++d) { Pn[d].resize(s->N); lincomb(Pn[d], 1.0, s->P[d], s->in.dt, s->GL[d]); }
    if (order == 1) {
        for (int d = 0; d < 3; ++d) s->P[d] = Pn[d];
        return s->pbc.bad ? FX_EINVAL : FX_OK;
    }
    int rc = compute_phi_rho(s, Pn, st2);
    if (rc) return rc;
    compute_E(s);
    std::vector<double> GLp[3];
    for (int d = 0; d < 3; ++d) GLp[d].resize(s->N);
    tdgl_rhs(s, Pn, GLp);
    for (int d = 0; d < 3; ++d) {
        std::vector<double> avg(s->N);
        lincomb(avg, 0.5, s->GL[d], 0.5, GLp[d]);
        lincomb(s->P[d], 1.0, s->P[d], s->in.dt, avg);
    }
    return s->pbc.bad ? FX_EINVAL : FX_OK;
}

}  // namespace

/* ------------------------------- C API ----------------------------------- */
extern "C" {

fxo_state* fxo_create(const fx_inputs* in) {
    fxo_state* s = new fxo_state();
    s->in = *in;
    for (int d = 0; d < 3; ++d) {
        s->n[d] = in->n_cell[d];
        s->dx[d] = (in->prob_hi[d] - in->prob_lo[d]) / in->n_cell[d];
    }
    s->N = (size_t)s->n[0] * s->n[1] * s->n[2];
    const size_t N = s->N;
    s->tphase.assign(N, 0.0);
    s->lalpha.assign(N, in->alpha);
    for (int d = 0; d < 3; ++d) {
        s->ang[d].assign(N, 0.0); s->P[d].assign(N, 0.0); s->E[d].assign(N, 0.0); s->GL[d].assign(N, 0.0);
        s->pbc.lo[d] = in->P_BC_flag_lo[d]; s->pbc.hi[d] = in->P_BC_flag_hi[d];
    }
    s->pbc.lambda = in->lambda; s->pbc.bad = false;
    s->Gam.assign(N, 0.0); s->rho.assign(N, 0.0); s->e_den.assign(N, 0.0); s->h_den.assign(N, 0.0);
    s->phi.assign(N, 0.0); s->phi_old.assign(N, 0.0); s->phi_prev.assign(N, 0.0);
    s->alpha_cc.assign(N, 0.0); s->rhs.assign(N, 0.0); s->phidiff.assign(N, 0.0);
    s->Phi_Bc_lo = in->Phi_Bc_lo; s->Phi_Bc_hi = in->Phi_Bc_hi;
    s->inc_step = in->inc_step;
    refresh_rotation(s);
    /* multigrid finest level */
    for (int d = 0; d < 3; ++d) { s->mg.bclo[d] = in->bc_lo[d]; s->mg.bchi[d] = in->bc_hi[d]; }
    s->mg.nu1 = in->mg_nu1; s->mg.nu2 = in->mg_nu2; s->mg.ratio = in->mg_coarsen_ratio;
    s->mg.exact_tail = in->mg_exact_tail;
    Lev l0;
    for (int d = 0; d < 3; ++d) { l0.n[d] = s->n[d]; l0.h[d] = s->dx[d]; l0.ih2[d] = 1.0 / (s->dx[d] * s->dx[d]); l0.crs[d] = 0; }
    l0.N = N;
    l0.bx.assign((size_t)(s->n[0] + 1) * s->n[1] * s->n[2], 0.0);
    l0.by.assign((size_t)s->n[0] * (s->n[1] + 1) * s->n[2], 0.0);
    l0.bz.assign((size_t)s->n[0] * s->n[1] * (s->n[2] + 1), 0.0);
    l0.a.assign(N, 0.0); l0.phi.assign(N, 0.0); l0.b.assign(N, 0.0); l0.r.assign(N, 0.0);
    s->mg.L.push_back(std::move(l0));
    s->mg.build(32);
    build_mask(s);
    build_eps(s);
    /* InitializePandRho, Initialization.cpp:6-166 */
    const fx_inputs& I = *in;
    const uint64_t seed = (uint64_t)(int64_t)I.random_seed;
    for (int k = 0; k < s->n[2]; ++k)
        for (int j = 0; j < s->n[1]; ++j)
            for (int i = 0; i < s->n[0]; ++i) {
                size_t c = s->c(i, j, k);
                double x = I.prob_lo[0] + (i + 0.5) * s->dx[0];
                double y = I.prob_lo[1] + (j + 0.5) * s->dx[1];
                double z = I.prob_lo[2] + (k + 0.5) * s->dx[2];
                double p = 0, q = 0, r = 0;
                if (s->mask(i, j, k) == 0.0) {
                    if (I.prob_type == 1) {
                        double u = uniform01(seed, (uint64_t)(i + k * s->n[2]));
                        p = (-1.0 + 2.0 * u) * I.Remnant_P[0];
                        q = (-1.0 + 2.0 * u) * I.Remnant_P[1];
                        r = (-1.0 + 2.0 * u) * I.Remnant_P[2];
                    } else if (I.prob_type == 2) {
                        p = (-1.0 + 2.0 * uniform01(seed, 3 * (uint64_t)c + 0)) * I.Remnant_P[0];
                        q = (-1.0 + 2.0 * uniform01(seed, 3 * (uint64_t)c + 1)) * I.Remnant_P[1];
                        r = (-1.0 + 2.0 * uniform01(seed, 3 * (uint64_t)c + 2)) * I.Remnant_P[2];
                    } else if (I.prob_type == 3) {
                        double g = exp(-(x * x / (2.0 * 5.e-9 * 5.e-9) + y * y / (2.0 * 5.e-9 * 5.e-9)
                                         + (z - 1.5 * I.DE_hi[2]) * (z - 1.5 * I.DE_hi[2]) / (2.0 * 2.0e-9 * 2.0e-9)));
                        p = I.Remnant_P[0] * g; q = I.Remnant_P[1] * g; r = I.Remnant_P[2] * g;
                    }
                    s->Gam[c] = I.BigGamma;
                    if (s->tphase[c] == 1.0) r = 0.0;
                }
                if (I.is_polarization_scalar == 1) { p = 0.0; q = 0.0; }
                s->P[0][c] = p; s->P[1][c] = q; s->P[2][c] = r;
                if (s->mask(i, j, k) >= 2.0) {
                    const SiConst& K = s->si;
                    double ni = std::sqrt(K.Nc * K.Nv) * exp(-0.5 * K.q * K.bandgap / (K.kb * K.T));
                    s->h_den[c] = ni; s->e_den[c] = ni;
                    s->rho[c] = K.q * (ni - ni - I.acceptor_doping + I.donor_doping);
                }
            }
    if (I.prob_type < 1 || I.prob_type > 3) s->err = "invalid prob_type";
    if (I.bias_ramp_steps > 0) {
        s->waveform.resize(I.bias_ramp_steps);
        for (int t = 1; t <= I.bias_ramp_steps; ++t)
            s->waveform[t - 1] = I.bias_ramp_start + (I.bias_ramp_end - I.bias_ramp_start) * ((double)t / I.bias_ramp_steps);
    }
    return s;
}

void fxo_destroy(fxo_state* s) { delete s; }
const char* fxo_error(fxo_state* s) { return s->err.c_str(); }

int fxo_num_levels(fxo_state* s) { return (int)s->mg.L.size(); }
int fxo_level_dims(fxo_state* s, int l, int dims[3]) {
    if (l < 0 || l >= (int)s->mg.L.size()) return FX_EINVAL;
    for (int d = 0; d < 3; ++d) dims[d] = s->mg.L[l].n[d];
    return FX_OK;
}

static std::vector<double>* field_ptr(fxo_state* s, int f) {
    switch (f) {
    case FX_FIELD_PX: return &s->P[0];
    case FX_FIELD_PY: return &s->P[1];
    case FX_FIELD_PZ: return &s->P[2];
    case FX_FIELD_PHI: return &s->phi;
    case FX_FIELD_POISSON_RHS: return &s->rhs;
    case FX_FIELD_EX: return &s->E[0];
    case FX_FIELD_EY: return &s->E[1];
    case FX_FIELD_EZ: return &s->E[2];
    case FX_FIELD_HOLES: return &s->h_den;
    case FX_FIELD_ELECTRONS: return &s->e_den;
    case FX_FIELD_CHARGE: return &s->rho;
    case FX_FIELD_TPHASE: return &s->tphase;
    case FX_FIELD_ANGLE_ALPHA: return &s->ang[0];
    case FX_FIELD_ANGLE_BETA: return &s->ang[1];
    case FX_FIELD_ANGLE_THETA: return &s->ang[2];
    case FX_FIELD_PHIDIFF: return &s->phidiff;
    case FX_FIELD_ALPHA_CC: return &s->alpha_cc;
    case FX_FIELD_GAMMA: return &s->Gam;
    case FX_FIELD_GL_RHS_X: return &s->GL[0];
    case FX_FIELD_GL_RHS_Y: return &s->GL[1];
    case FX_FIELD_GL_RHS_Z: return &s->GL[2];
    case FX_FIELD_LANDAU_ALPHA: return &s->lalpha;
    default: return nullptr;
    }
}

int fxo_get_field(fxo_state* s, int f, double* out) {
    if (f == FX_FIELD_EPSILON || f == FX_FIELD_MASK) {
        for (int k = 0; k < s->n[2]; ++k)
            for (int j = 0; j < s->n[1]; ++j)
                for (int i = 0; i < s->n[0]; ++i)
                    out[s->c(i, j, k)] = (f == FX_FIELD_MASK) ? s->mask(i, j, k) : s->eps(i, j, k);
        return FX_OK;
    }
    std::vector<double>* v = field_ptr(s, f);
    if (!v) return FX_EINVAL;
    std::memcpy(out, v->data(), s->N * sizeof(double));
    return FX_OK;
}

int fxo_set_field(fxo_state* s, int f, const double* in) {
    if (f == FX_FIELD_MASK) {
        for (int k = 0; k < s->n[2]; ++k)
            for (int j = 0; j < s->n[1]; ++j)
                for (int i = 0; i < s->n[0]; ++i) s->mask(i, j, k) = in[s->c(i, j, k)];
        rebuild_mask_ghosts(s);
        build_eps(s);
        for (size_t c = 0; c < s->N; ++c) {
            int i = (int)(c % s->n[0]), j = (int)((c / s->n[0]) % s->n[1]), k = (int)(c / ((size_t)s->n[0] * s->n[1]));
            s->Gam[c] = (s->mask(i, j, k) == 0.0) ? s->in.BigGamma : 0.0;
        }
        return FX_OK;
    }
    std::vector<double>* v = field_ptr(s, f);
    if (!v) return FX_EINVAL;
    std::memcpy(v->data(), in, s->N * sizeof(double));
    if (f == FX_FIELD_TPHASE) build_eps(s);
    if (f == FX_FIELD_ANGLE_ALPHA || f == FX_FIELD_ANGLE_BETA || f == FX_FIELD_ANGLE_THETA) refresh_rotation(s);
    return FX_OK;
}

int fxo_set_bias(fxo_state* s, double lo, double hi) { s->Phi_Bc_lo = lo; s->Phi_Bc_hi = hi; return FX_OK; }

int fxo_set_bias_waveform(fxo_state* s, const double* v, int n) {
    s->waveform.assign(v, v + (n > 0 ? n : 0));
    return FX_OK;
}

int fxo_compute_phi_rho(fxo_state* s, fx_solve_stats* st) {
    fx_solve_stats tmp{};
    return compute_phi_rho(s, s->P, st ? st : &tmp);
}

/* ComputeRho (ChargeDensity.cpp:15-102) at the current phi, without a solve. */
int fxo_compute_rho(fxo_state* s) { compute_rho(s, s->phi); return FX_OK; }

int fxo_compute_E(fxo_state* s) { compute_E(s); return FX_OK; }

int fxo_tdgl_rhs(fxo_state* s) {
    compute_E(s);
    tdgl_rhs(s, s->P, s->GL);
    return s->pbc.bad ? FX_EINVAL : FX_OK;
}

int fxo_update(fxo_state* s, int order) {
    fx_solve_stats st{};
    return update(s, order, &st);
}

/* The reference time loop, main.cpp:261-623. */
int fxo_step(fxo_state* s, int nsteps, fx_step_stats* out) {
    fx_step_stats st{};
    const fx_inputs& in = s->in;
    if (nsteps <= 0) nsteps = in.nsteps;
    auto t0 = std::chrono::steady_clock::now();
    const double tiny = 1.e-6;
    int rc = FX_OK;
    for (int it = 0; it < nsteps; ++it) {
        const int step = ++s->step;
        if (!s->waveform.empty()) {
            size_t w = std::min((size_t)step, s->waveform.size());
            s->Phi_Bc_hi = s->waveform[w - 1];
        }
        fx_solve_stats ss{};
        rc = compute_phi_rho(s, s->P, &ss);
        st.total_vcycles += ss.vcycles; st.total_outer_iters += ss.outer_iters; st.total_solves += ss.outer_iters;
        if (rc) break;
        fx_solve_stats ss2{};
        rc = update(s, in.TimeIntegratorOrder, &ss2);
        st.total_vcycles += ss2.vcycles; st.total_outer_iters += ss2.outer_iters; st.total_solves += ss2.outer_iters;
        if (rc) break;
        st.last_phi_change = check_steady_state(s, step);
        compute_E(s);
        s->time = s->time + in.dt;
        st.steps_done++;
        if (in.voltage_sweep == 1 && s->inc_step > 0 && step == s->inc_step) {
            s->Phi_Bc_hi += s->sign * in.Phi_Bc_inc;
            s->num_Vapp += 1;
            if (std::fabs(std::fabs(s->Phi_Bc_hi) - in.Phi_Bc_hi_max) <= tiny) s->sign *= -1;
            fx_solve_stats ss3{};
            rc = compute_phi_rho(s, s->P, &ss3);
            st.total_vcycles += ss3.vcycles; st.total_outer_iters += ss3.outer_iters; st.total_solves += ss3.outer_iters;
            if (rc) break;
        }
        if (in.voltage_sweep == 0 && step == s->steady_state_step) { st.stopped_steady = 1; break; }
        if (in.voltage_sweep == 1 && s->Phi_Bc_hi > 0. && s->Phi_Bc_hi - in.Phi_Bc_hi_max > tiny) break;
        if (in.voltage_sweep == 1 && s->Phi_Bc_hi < 0. && -s->Phi_Bc_hi - in.Phi_Bc_hi_max > tiny) break;
        if (in.voltage_sweep == 1 && s->num_Vapp == in.num_Vapp_max) break;
    }
    st.time = s->time; st.Phi_Bc_hi = s->Phi_Bc_hi; st.num_Vapp = s->num_Vapp;
    st.wall_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (out) *out = st;
    return rc;
}

/* Multigrid solve with explicit rhs / A-coefficient (tests).  Uses the
 * current bias as Dirichlet data; phi is the initial guess on entry. */
int fxo_poisson_solve(fxo_state* s, const double* rhs, const double* acoef, double* phi, double tol, fx_solve_stats* st) {
    Lev& l0 = s->mg.L[0];
    l0.b.assign(rhs, rhs + s->N);
    if (acoef) l0.a.assign(acoef, acoef + s->N); else std::fill(l0.a.begin(), l0.a.end(), 0.0);
    s->mg.restrict_acoef();
    l0.phi.assign(phi, phi + s->N);
    double bv[6]; current_bc_values(s, bv);
    fx_solve_stats tmp{};
    int rc = s->mg.solve(tol, s->in.mg_max_iter, bv, st ? st : &tmp);
    std::memcpy(phi, l0.phi.data(), s->N * sizeof(double));
    return rc;
}

/* y = A x on the finest level (inhomog != 0 adds the Dirichlet data). */
int fxo_apply_op(fxo_state* s, const double* acoef, const double* x, double* y, int inhomog) {
    Lev& l0 = s->mg.L[0];
    if (acoef) l0.a.assign(acoef, acoef + s->N); else std::fill(l0.a.begin(), l0.a.end(), 0.0);
    double bv[6]; current_bc_values(s, bv);
    for (int k = 0; k < s->n[2]; ++k)
        for (int j = 0; j < s->n[1]; ++j)
            for (int i = 0; i < s->n[0]; ++i) {
                double D, S;
                s->mg.row(l0, x, i, j, k, inhomog ? bv : nullptr, D, S);
                size_t c = s->c(i, j, k);
                y[c] = D * x[c] - S;
            }
    return FX_OK;
}

/* Face coefficients of the finest level (for tests): which = 0,1,2. */
int fxo_get_faces(fxo_state* s, int which, double* out) {
    const Lev& l0 = s->mg.L[0];
    const std::vector<double>& v = which == 0 ? l0.bx : which == 1 ? l0.by : l0.bz;
    std::memcpy(out, v.data(), v.size() * sizeof(double));
    return FX_OK;
}

int fxo_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* threads of the parallel loops from here on (timing samples); returns the count in effect */
int fxo_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

double fxo_uniform01(unsigned long long seed, unsigned long long idx) { return uniform01(seed, idx); }

}  // extern "C"
