/*
 * ORACLE -- test infrastructure only.  Nothing in the product (ferrox_amd/)
 * may include, link or call this code; only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg use it, as the checker.
 *
 * Masked finite-difference stencils of FerroX, restated in plain C++ from
 * Source/Solver/DerivativeAlgorithm.H (read as text).  The arithmetic of every
 * formula (operand order, "/dx/dx" versus "/(dx*dx)") follows the reference
 * line it cites, so that a faithful device implementation can match it to
 * rounding.
 *
 * Accessors: F(i,j,k) returns P with the reference's ghost semantics (periodic
 * image in periodic directions, 0 outside the domain in the others); M(i,j,k)
 * returns the FP64 material mask including its ghost layer.
 */
#ifndef FX_ORACLE_STENCILS_H
#define FX_ORACLE_STENCILS_H

#include <cmath>

namespace fxo {

struct BCFlags {
    int lo[3];
    int hi[3];
    double lambda;
    bool bad;   /* set when a flag outside 0..4 was hit (amrex::Abort in the reference) */
};

/* DPDx / DPDy / DPDz: DerivativeAlgorithm.H:45-252.  d = 0,1,2; (di,dj,dk) is
 * the unit offset along d. */
template <class FA, class MA>
inline double DPD(const FA& F, const MA& M, int i, int j, int k, int d,
                  const double dx[3], BCFlags& bc) {
    const int di = (d == 0), dj = (d == 1), dk = (d == 2);
    const double h = dx[d];
    const double lambda = bc.lambda;
    if (M(i - di, j - dj, k - dk) != 0.0 && M(i, j, k) == 0.0) {          /* FE lower boundary */
        switch (bc.lo[d]) {
        case 0: { double F_lo = 0.0;
                  return (-4. * F_lo + 3. * F(i, j, k) + F(i + di, j + dj, k + dk)) / (3. * h); }
        case 1: { double F_lo = F(i, j, k) / (1 + h / 2 / lambda);
                  return (h * F_lo / lambda - F(i, j, k) + F(i + di, j + dj, k + dk)) / (2. * h); }
        case 2:   return (-F(i, j, k) + F(i + di, j + dj, k + dk)) / (2. * h);
        case 3:   return (-F(i - di, j - dj, k - dk) + F(i + di, j + dj, k + dk)) / (2. * h);
        case 4:   return (-F(i, j, k) + F(i + di, j + dj, k + dk)) / (h);
        default:  bc.bad = true; return 0.0;
        }
    } else if (M(i + di, j + dj, k + dk) != 0.0 && M(i, j, k) == 0.0) {   /* FE higher boundary */
        switch (bc.hi[d]) {
        case 0: { double F_hi = 0.0;
                  return (4. * F_hi - 3. * F(i, j, k) - F(i - di, j - dj, k - dk)) / (3. * h); }
        case 1: { double F_hi = F(i, j, k) / (1 - h / 2 / lambda);
                  return (h * F_hi / lambda + F(i, j, k) - F(i - di, j - dj, k - dk)) / (2. * h); }
        case 2:   return (F(i, j, k) - F(i - di, j - dj, k - dk)) / (2. * h);
        case 3:   return (F(i + di, j + dj, k + dk) - F(i - di, j - dj, k - dk)) / (2. * h);
        case 4:   return (F(i, j, k) - F(i - di, j - dj, k - dk)) / (h);
        default:  bc.bad = true; return 0.0;
        }
    } else if (M(i, j, k) == 0.0) {                                        /* inside FE */
        return (F(i + di, j + dj, k + dk) - F(i - di, j - dj, k - dk)) / (2. * h);
    }
    return 0.0;
}

/* DoubleDPDx / DoubleDPDy / DoubleDPDz: DerivativeAlgorithm.H:257-467. */
template <class FA, class MA>
inline double DDPD(const FA& F, const MA& M, int i, int j, int k, int d,
                   const double dx[3], BCFlags& bc) {
    const int di = (d == 0), dj = (d == 1), dk = (d == 2);
    const double h = dx[d];
    const double lambda = bc.lambda;
    if (M(i - di, j - dj, k - dk) != 0.0 && M(i, j, k) == 0.0) {          /* FE lower boundary */
        switch (bc.lo[d]) {
        case 0: { double F_lo = 0.0;
                  return 4. * (2. * F_lo - 3. * F(i, j, k) + F(i + di, j + dj, k + dk)) / 3. / h / h; }
        case 1: { double F_lo = F(i, j, k) / (1 + h / 2 / lambda);
                  return (-h * F_lo / lambda - F(i, j, k) + F(i + di, j + dj, k + dk)) / h / h; }
        case 2:   return (-F(i, j, k) + F(i + di, j + dj, k + dk)) / h / h;
        case 3:   return (F(i + di, j + dj, k + dk) - 2. * F(i, j, k) + F(i - di, j - dj, k - dk)) / h / h;
        case 4:   return (F(i + 2 * di, j + 2 * dj, k + 2 * dk) - 2. * F(i + di, j + dj, k + dk) + F(i, j, k)) / h / h;
        default:  bc.bad = true; return 0.0;
        }
    } else if (M(i + di, j + dj, k + dk) != 0.0 && M(i, j, k) == 0.0) {   /* FE higher boundary */
        switch (bc.hi[d]) {
        case 0: { double F_hi = 0.0;
                  return 4. * (2. * F_hi - 3. * F(i, j, k) + F(i - di, j - dj, k - dk)) / 3. / h / h; }
        case 1: { double F_hi = F(i, j, k) / (1 - h / 2 / lambda);
                  return (h * F_hi / lambda - F(i, j, k) + F(i - di, j - dj, k - dk)) / h / h; }
        case 2:   return (-F(i, j, k) + F(i - di, j - dj, k - dk)) / h / h;
        case 3:   return (F(i + di, j + dj, k + dk) - 2. * F(i, j, k) + F(i - di, j - dj, k - dk)) / h / h;
        case 4:   return (F(i, j, k) - 2. * F(i - di, j - dj, k - dk) + F(i - 2 * di, j - 2 * dj, k - 2 * dk)) / h / h;
        default:  bc.bad = true; return 0.0;
        }
    } else if (M(i, j, k) == 0.0) {                                        /* inside FE */
        return (F(i + di, j + dj, k + dk) - 2. * F(i, j, k) + F(i - di, j - dj, k - dk)) / (h * h);
    }
    return 0.0;
}

/* Mixed second derivatives: DerivativeAlgorithm.H:471-497.
 * d2/(da db) = (DPD_b(i+e_a) - DPD_b(i-e_a)) / 2 / dx[a]. */
template <class FA, class MA>
inline double DDPD_mixed(const FA& F, const MA& M, int i, int j, int k, int a, int b,
                         const double dx[3], BCFlags& bc) {
    const int ai = (a == 0), aj = (a == 1), ak = (a == 2);
    return (DPD(F, M, i + ai, j + aj, k + ak, b, dx, bc) -
            DPD(F, M, i - ai, j - aj, k - ak, b, dx, bc)) / 2. / dx[a];
}

/* Rotation matrix from Euler angles in degrees, both conventions of the
 * reference (TotalEnergyDensity.cpp:167-195, ElectrostaticSolver.cpp:35-63). */
inline void rotation(double a_deg, double b_deg, double t_deg, int use_Euler_angles, double R[9]) {
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

}  // namespace fxo
#endif
