/*
 * AMReX plotfile writer for the hot path's fields: the drop-in for
 * WritePlotfile (Source/Plotfile.cpp:5-177), which copies the fields into one
 * MultiFab and calls amrex::WriteSingleLevelPlotfile.  Directory
 * plt%08d (amrex::Concatenate("plt", step, 8)), variables Px, Py, Pz then
 * Phi, PoissonRHS, Ex, Ey, Ez, holes, electrons, charge, epsilon, mask,
 * tphase, alpha, beta, theta, PhiDiff, each gated by its plot_* flag
 * (FerroX.cpp:300-325).
 *
 * The on-disk layout is AMReX's single-level plotfile, restated from its
 * documented format (no AMReX here to compare against; the tests read the
 * files back with an independent reader):
 *   <dir>/Header            "HyperCLaw-V1.1", variables, dimension, time,
 *                           finest level 0, prob_lo/hi, domain box, step,
 *                           cell size, coordinate system, box bounds,
 *                           "Level_0/Cell"
 *   <dir>/Level_0/Cell_H    VisMF header (version 1, NFiles), box array,
 *                           FabOnDisk entry, per-component minima/maxima
 *   <dir>/Level_0/Cell_D_00000   "FAB <real descriptor><box> <ncomp>\n" and
 *                           the components one after another, FP64
 *                           little-endian, i fastest (one box: the domain).
 */
#include <sys/stat.h>
#include <sys/types.h>

#include <cerrno>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <iomanip>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/ferrox_mi.h"

const fx_inputs* fxi_inputs(fx_ctx* x);

namespace {

int mkdir_p(const std::string& path) {
    std::string cur;
    std::stringstream ss(path);
    std::string part;
    if (!path.empty() && path[0] == '/') cur = "/";
    while (std::getline(ss, part, '/')) {
        if (part.empty()) continue;
        cur += part;
        if (::mkdir(cur.c_str(), 0755) != 0 && errno != EEXIST) return -1;
        cur += "/";
    }
    return 0;
}

std::string box_str(int nx, int ny, int nz) {
    std::ostringstream o;
    o << "((0,0,0) (" << nx - 1 << "," << ny - 1 << "," << nz - 1 << ") (0,0,0))";
    return o.str();
}

}  // namespace

extern "C" int fx_write_plotfile(fx_ctx* x, const char* dir, double time, int step) {
    if (!x || !dir) return FX_EINVAL;
    const fx_inputs* in = fxi_inputs(x);
    if (!in) return FX_EINVAL;
    int n[3];
    if (fx_get_dims(x, n)) return FX_EINVAL;
    const size_t N = (size_t)n[0] * n[1] * n[2];
    struct Var { const char* name; int field; int on; };
    const Var all[] = {
        {"Px", FX_FIELD_PX, 1}, {"Py", FX_FIELD_PY, 1}, {"Pz", FX_FIELD_PZ, 1},
        {"Phi", FX_FIELD_PHI, in->plot_Phi}, {"PoissonRHS", FX_FIELD_POISSON_RHS, in->plot_PoissonRHS},
        {"Ex", FX_FIELD_EX, in->plot_E}, {"Ey", FX_FIELD_EY, in->plot_E}, {"Ez", FX_FIELD_EZ, in->plot_E},
        {"holes", FX_FIELD_HOLES, in->plot_holes}, {"electrons", FX_FIELD_ELECTRONS, in->plot_electrons},
        {"charge", FX_FIELD_CHARGE, in->plot_charge}, {"epsilon", FX_FIELD_EPSILON, in->plot_epsilon},
        {"mask", FX_FIELD_MASK, in->plot_mask}, {"tphase", FX_FIELD_TPHASE, in->plot_tphase},
        {"alpha", FX_FIELD_ANGLE_ALPHA, in->plot_alpha}, {"beta", FX_FIELD_ANGLE_BETA, in->plot_beta},
        {"theta", FX_FIELD_ANGLE_THETA, in->plot_theta}, {"PhiDiff", FX_FIELD_PHIDIFF, in->plot_PhiDiff},
    };
    std::vector<const Var*> vars;
    for (const Var& v : all) if (v.on) vars.push_back(&v);
    /* E is derived from phi after every step in the reference (main.cpp:551) */
    if (in->plot_E) { int rc = fx_compute_E(x); if (rc) return rc; }
    std::vector<double> data(N * vars.size());
    for (size_t q = 0; q < vars.size(); ++q) {
        int rc = fx_get_field(x, vars[q]->field, data.data() + q * N);
        if (rc) return rc;
    }
    const std::string root(dir);
    if (mkdir_p(root + "/Level_0")) return FX_EINVAL;
    double dx[3];
    for (int d = 0; d < 3; ++d) dx[d] = (in->prob_hi[d] - in->prob_lo[d]) / n[d];
    {
        std::ofstream h(root + "/Header");
        if (!h) return FX_EINVAL;
        h.precision(17);
        h << "HyperCLaw-V1.1\n" << vars.size() << '\n';
        for (const Var* v : vars) h << v->name << '\n';
        h << 3 << '\n' << time << '\n' << 0 << '\n';
        for (int d = 0; d < 3; ++d) h << in->prob_lo[d] << ' ';
        h << '\n';
        for (int d = 0; d < 3; ++d) h << in->prob_hi[d] << ' ';
        h << '\n' << '\n';
        h << box_str(n[0], n[1], n[2]) << ' ' << '\n';
        h << step << ' ' << '\n';
        for (int d = 0; d < 3; ++d) h << dx[d] << ' ';
        h << '\n';
        h << 0 << '\n' << "0\n";
        h << 0 << ' ' << 1 << ' ' << time << '\n' << step << '\n';
        for (int d = 0; d < 3; ++d) h << in->prob_lo[d] << ' ' << in->prob_lo[d] + n[d] * dx[d] << '\n';
        h << "Level_0/Cell\n";
        if (!h) return FX_EINVAL;
    }
    {
        std::ofstream f(root + "/Level_0/Cell_D_00000", std::ios::binary);
        if (!f) return FX_EINVAL;
        f << "FAB ((8, (64 11 52 0 1 12 0 1023)),(8, (8 7 6 5 4 3 2 1)))" << box_str(n[0], n[1], n[2]) << ' '
          << vars.size() << '\n';
        f.write(reinterpret_cast<const char*>(data.data()), (std::streamsize)(data.size() * sizeof(double)));
        if (!f) return FX_EINVAL;
    }
    {
        std::ofstream c(root + "/Level_0/Cell_H");
        if (!c) return FX_EINVAL;
        c.precision(17);
        c << 1 << '\n' << 1 << '\n' << vars.size() << '\n' << 0 << '\n';
        c << '(' << 1 << ' ' << 0 << '\n' << box_str(n[0], n[1], n[2]) << '\n' << ')' << '\n';
        c << 1 << '\n' << "FabOnDisk: Cell_D_00000 0" << '\n' << '\n';
        c << 1 << ',' << vars.size() << '\n';
        for (size_t q = 0; q < vars.size(); ++q) {
            double mn = data[q * N];
            for (size_t i = 0; i < N; ++i) mn = std::fmin(mn, data[q * N + i]);
            c << mn << ',';
        }
        c << '\n' << '\n' << 1 << ',' << vars.size() << '\n';
        for (size_t q = 0; q < vars.size(); ++q) {
            double mx = data[q * N];
            for (size_t i = 0; i < N; ++i) mx = std::fmax(mx, data[q * N + i]);
            c << mx << ',';
        }
        c << '\n';
        if (!c) return FX_EINVAL;
    }
    return FX_OK;
}
