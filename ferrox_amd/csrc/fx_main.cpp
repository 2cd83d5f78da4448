/*
 * ferrox_mi -- command-line driver with the reference's calling convention:
 *
 *     ferrox_mi <inputs_file> [key=value ...]
 *
 * It reads a FerroX inputs file (plus command-line overrides, last wins, as
 * ParmParse does), builds one context and runs the time loop of
 * Source/main.cpp:261-623 through the C ABI, printing the per-step lines the
 * reference prints (main.cpp:556) and the total run time (:649).
 */
#include <chrono>
#include <cstdio>
#include <cstring>
#include <string>

#include "../../include/ferrox_mi.h"

int main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s <inputs_file> [key=value ...]\n", argv[0]);
        return 2;
    }
    fx_inputs in;
    fx_inputs_default(&in);
    in.verbose = 1;
    char err[512] = {0};
    if (fx_inputs_parse_file(argv[1], &in, err, sizeof(err))) {
        std::fprintf(stderr, "inputs: %s\n", err);
        return 1;
    }
    for (int a = 2; a < argc; ++a)
        if (fx_inputs_set(&in, argv[a], err, sizeof(err))) {
            std::fprintf(stderr, "override %s: %s\n", argv[a], err);
            return 1;
        }
    if (fx_inputs_finalize(&in, err, sizeof(err))) {
        std::fprintf(stderr, "inputs: %s\n", err);
        return 1;
    }
    fx_ctx* ctx = nullptr;
    int rc = fx_create(&in, &ctx);
    if (rc) {
        std::fprintf(stderr, "fx_create failed (%d)\n", rc);
        return 1;
    }
    auto t0 = std::chrono::steady_clock::now();
    fx_step_stats st;
    rc = fx_step(ctx, in.nsteps, &st);
    if (rc) std::fprintf(stderr, "time loop failed: %s\n", fx_last_error(ctx));
    const int done = st.steps_done;
    std::printf("V-cycles %d, outer iterations %d, Poisson solves %d\n", st.total_vcycles, st.total_outer_iters,
                st.total_solves);
    const double t = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::printf("Total run time %g seconds (%d steps)\n", t, done);
    fx_destroy(ctx);
    return rc ? 1 : 0;
}
