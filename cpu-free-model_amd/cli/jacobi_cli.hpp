// jacobi2D / jacobi3D driver: the reference's version table and flags
// (Stencil/jacobi2D/src/main.cu:24-69, Stencil/jacobi3D/src/main.cu:30-73,
// flags README.md:127-136) on top of the cfm C ABI.
//
// One GPU: the multi-GPU versions collapse onto the two single-GPU schemes.
// "Baseline *" run one launch per sweep (the reference's single_gpu loop),
// "Design 1/2" run the single persistent cooperative kernel one sweep per
// pass, "PERKS" the persistent kernel that keeps the intermediate sweeps on
// chip (two sweeps per HBM pass in 3D, four in 2D, where the shape allows
// it; -tblock 1|2|3|4 to force), and the "(No computation)" versions run the
// persistent kernel's barriers only.  -phase_timers prints the persistent
// kernel's in-place breakdown (time sweeping vs waiting at the device-wide
// barrier, per workgroup).
//
// -compare runs the launch-per-sweep scheme first (the reference's
// single_gpu, Stencil/jacobi2D/src/common.cu:154-273) and compares every
// interior point; unlike the reference, a mismatch fails the run (exit 1).
#pragma once

#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>

#include "cli_common.hpp"

namespace cli {

inline int run_jacobi(int argc, char **argv, int dims) {
    // line-buffered even into a file: the report survives a tool (a profiler)
    // that ends the process abnormally after main returns
    std::setvbuf(stdout, nullptr, _IOLBF, 0);
    const std::array<const char *, 11> names{
        "Baseline Copy", "Baseline Overlap", "Baseline P2P", "Design 1", "Design 2", "PERKS",
        "Baseline Copy (No computation)", "Baseline Overlap (No Computation)",
        "Baseline P2P (No Computation)", "Design 1 (No Computation)", "Design 2 (No Computation)"};
    const int variants[11] = {CFM_JACOBI_LAUNCH_PER_ITER, CFM_JACOBI_LAUNCH_PER_ITER,
                              CFM_JACOBI_LAUNCH_PER_ITER, CFM_JACOBI_PERSISTENT,
                              CFM_JACOBI_PERSISTENT,      CFM_JACOBI_PERSISTENT,
                              CFM_JACOBI_NO_COMPUTE,      CFM_JACOBI_NO_COMPUTE,
                              CFM_JACOBI_NO_COMPUTE,      CFM_JACOBI_NO_COMPUTE,
                              CFM_JACOBI_NO_COMPUTE};
    char **b = argv, **e = argv + argc;
    const int selection = get_argval<int>(b, e, "-v", 0);
    const bool silent = get_arg(b, e, "-s");
    if (selection < 0 || selection >= (int)names.size()) {
        std::fprintf(stderr, "ERROR: -v %d: no such version\n", selection);
        return 1;
    }
    if (!silent) print_versions(names, selection);

    cfm_jacobi_params p{};
    p.niter = get_argval<int>(b, e, "-niter", 1000);
    p.nx = get_argval<int>(b, e, "-nx", dims == 2 ? 16384 : 512);
    p.ny = get_argval<int>(b, e, "-ny", dims == 2 ? 16384 : 512);
    p.nz = dims == 2 ? 1 : get_argval<int>(b, e, "-nz", 512);
    const bool compare = get_arg(b, e, "-compare");
    const std::string prec = get_argval<std::string>(b, e, "-prec", "f64");
    const std::string bc = get_argval<std::string>(b, e, "-bc", "periodic");
    const std::string init = get_argval<std::string>(b, e, "-init", dims == 2 ? "sin" : "ref");
    p.norm_every = get_argval<int>(b, e, "-norm", 0);
    p.tile_rows = get_argval<int>(b, e, "-tile", 0);
    p.grid_blocks = get_argval<int>(b, e, "-grid", 0);
    p.chunks = get_argval<int>(b, e, "-chunks", 0);
    p.block_threads = get_argval<int>(b, e, "-block", 0);
    p.time_block = get_argval<int>(b, e, "-tblock", selection == 5 ? 0 : 1);
    const bool json = get_arg(b, e, "-json");
    // -phase_timers: the persistent kernel's in-place breakdown (sweeps / barrier wait per workgroup)
    p.phase_timers = get_arg(b, e, "-phase_timers") ? 1 : 0;
    p.prec = prec == "f32" ? CFM_F32 : CFM_F64;
    p.bc = bc == "dirichlet" ? CFM_BC_DIRICHLET : CFM_BC_PERIODIC;
    p.variant = variants[selection];
    const int kind = init == "ref" ? CFM_INIT_REF_QUIRK : CFM_INIT_SIN;
    const size_t elems = (size_t)p.nx * p.ny * p.nz;
    const size_t esz = p.prec == CFM_F64 ? 8 : 4;

    std::vector<unsigned char> field(elems * esz), out(elems * esz);
    int st = cfm_jacobi_init_field(p.nx, p.ny, p.nz, kind, p.prec, field.data());
    if (st != CFM_OK) return fail("cfm_jacobi_init_field", st);

    double runtime_single = 0.0;
    std::vector<unsigned char> ref;
    if (compare) {
        cfm_jacobi_params q = p;
        q.variant = CFM_JACOBI_LAUNCH_PER_ITER;
        q.norm_every = 0;
        q.time_block = 1;  // the reference scheme: one sweep per launch
        if (dims == 2)
            std::printf("Single GPU jacobi relaxation (non-persistent kernel): %d iterations on %d x %d mesh "
                        "with norm check every %d iterations\n",
                        q.niter, q.ny, q.nx, 0);
        else
            std::printf("Single GPU jacobi relaxation (non-persistent kernel): %d iterations on %d x %d x %d "
                        "mesh with norm check every %d iterations\n",
                        q.niter, q.nx, q.ny, q.nz, 0);
        std::fflush(stdout);
        ref.resize(elems * esz);
        cfm_timing t{};
        st = cfm_jacobi_run(&q, field.data(), ref.data(), nullptr, 0, nullptr, &t);
        if (st != CFM_OK) return fail("single-GPU reference run", st);
        runtime_single = t.wall_ms / 1e3;
    }

    const int ncap = p.norm_every > 0 ? (p.niter + p.norm_every - 1) / p.norm_every : 0;
    std::vector<double> norms(ncap > 0 ? ncap : 1);
    int nnorm = 0;
    cfm_timing t{};
    cfm_jacobi_plan *plan = nullptr;
    st = cfm_jacobi_plan_create(&p, &plan);
    if (st != CFM_OK) return fail("cfm_jacobi_plan_create", st);
    st = cfm_jacobi_plan_load(plan, field.data());
    if (st == CFM_OK) st = cfm_jacobi_plan_run(plan, norms.data(), ncap, &nnorm, &t);
    if (st == CFM_OK) st = cfm_jacobi_plan_download(plan, out.data());
    cfm_phase_times ph{};
    if (st == CFM_OK && p.phase_timers && p.variant != CFM_JACOBI_LAUNCH_PER_ITER)
        st = cfm_jacobi_plan_phase_times(plan, &ph);
    cfm_jacobi_plan_destroy(plan);
    if (st != CFM_OK) return fail("jacobi run", st);

    bool correct = true;
    if (compare && p.variant != CFM_JACOBI_NO_COMPUTE) {
        const int nz = p.nz, z0 = dims == 2 ? 0 : 1, z1 = dims == 2 ? 1 : nz - 1;
        long long bad = 0;
        for (int iz = z0; iz < z1; ++iz)
            for (int iy = 1; iy < p.ny - 1; ++iy)
                for (int ix = 1; ix < p.nx - 1; ++ix) {
                    const size_t i = ((size_t)iz * p.ny + iy) * p.nx + ix;
                    const double a = esz == 8 ? ((double *)out.data())[i] : ((float *)out.data())[i];
                    const double r = esz == 8 ? ((double *)ref.data())[i] : ((float *)ref.data())[i];
                    if (a != r) {
                        if (bad < 5)
                            std::fprintf(stderr, "ERROR: a[%d][%d][%d] = %.17g does not match %.17g (reference)\n",
                                         iz, iy, ix, a, r);
                        ++bad;
                    }
                }
        correct = bad == 0;
    }
    for (int i = 0; i < nnorm && i < ncap; ++i) std::printf("%5d, %0.6e\n", i * p.norm_every, norms[i]);
    if (!correct) return 1;
    const double secs = t.wall_ms / 1e3;
    std::printf("Execution time: %8.4f s\n", secs);
    if (compare) {
        if (dims == 2)
            std::printf("Non-persistent kernel - %dx%d: 1 GPU: %8.4f s, %d GPUs: %8.4f s, speedup: %8.2f, "
                        "efficiency: %8.2f \n",
                        p.ny, p.nx, runtime_single, 1, secs, runtime_single / secs, runtime_single / secs * 100);
        else
            std::printf("Non-persistent kernel - %dx%dx%d: 1 GPU: %8.4f s, %d GPUs: %8.4f s, speedup: %8.2f, "
                        "efficiency: %8.2f \n",
                        p.nx, p.ny, p.nz, runtime_single, 1, secs, runtime_single / secs,
                        runtime_single / secs * 100);
    }
    if (ph.workgroups > 0)
        std::printf("phases (ms, mean over the kernel's %d workgroups): sweeps %.4f (largest %.4f) barrier %.4f "
                    "(largest %.4f), total %.4f\n",
                    ph.workgroups, ph.ms[CFM_PHASE_SPMV], ph.max_ms[CFM_PHASE_SPMV], ph.ms[CFM_PHASE_BARRIER],
                    ph.max_ms[CFM_PHASE_BARRIER], ph.total_ms);
    if (json) {
        const double lups = (double)(p.nx - 2) * (p.ny - 2) * (dims == 2 ? 1 : p.nz - 2) * p.niter;
        std::printf("{\"version\": \"%s\", \"kernel_ms\": %.4f, \"glups\": %.3f, \"gbs\": %.1f, \"grid\": %d, "
                    "\"block\": %d, \"passes\": %d, \"cooperative\": %d",
                    names[selection], t.kernel_ms, lups / (t.kernel_ms * 1e6),
                    lups * 2 * esz / (t.kernel_ms * 1e6), t.grid_blocks, t.block_threads, t.outer_iterations,
                    t.cooperative);
        if (ph.workgroups > 0)
            std::printf(", \"phases\": {\"workgroups\": %d, \"sweeps_ms\": %.4f, \"sweeps_max_ms\": %.4f, "
                        "\"barrier_ms\": %.4f, \"barrier_max_ms\": %.4f, \"total_ms\": %.4f}",
                        ph.workgroups, ph.ms[CFM_PHASE_SPMV], ph.max_ms[CFM_PHASE_SPMV], ph.ms[CFM_PHASE_BARRIER],
                        ph.max_ms[CFM_PHASE_BARRIER], ph.total_ms);
        std::printf("}\n");
    }
    return 0;
}

}  // namespace cli
