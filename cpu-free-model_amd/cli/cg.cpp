// cg: Conjugate Gradient driver with the reference's version table and flags
// (CG/src/main.cu:23-278) on top of the cfm C ABI.
//
// Matrix: -nx/-ny/-nz build the 3D 7-point Poisson system (default 128^3);
// -tridiag N builds the reference's random SPD tridiagonal (CG/src/common.cu:46-76,
// glibc rand(), default seed); -matrix_path reads a Matrix Market file by the
// reference's rules (coordinate, real or integer, general / symmetric /
// skew-symmetric; 0- or 1-based indices detected as CG/src/main.cu:180-186).
// -json adds one JSON line per run (iterations, residual, times, launch kind).
// -phase_timers prints each run's in-place time breakdown (spmv / update /
// reduce / barrier / other); versions 6 and 7 always time their operations.
// Right-hand side: -rhs uniform (SplitMix64 seed -seed, default) or -rhs ones
// (the reference's b = 1).  -tol 0 runs exactly -niter iterations as the
// reference does; otherwise the solve stops at ||r|| <= tol ||b||.
//
// One GPU: "Baseline Discrete" / "Profiling Discrete" / "Single GPU Discrete"
// run the host-driven schedule (one launch per operation, host reads of the
// scalars); "Single Stream *" run the persistent cooperative kernel.
// -compare-single-gpu runs version 8 first and compares x within 1e-5 as the
// reference does (CG/src/common.cu:78-112).
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>

#include "cli_common.hpp"

namespace {

struct Csr {
    int n = 0;
    int base = 0;  // 0 or 1, as rowptr[0] (the reference's zero_indexed flag)
    std::vector<int> rowptr, col;
    std::vector<double> val;
};

// The reference's default matrix (genTridiag, CG/src/common.cu:46-76) and its
// Matrix Market path (loadMMSparseMatrix, CG/src/mmio_wrapper.cpp:170-369)
// live in the library: cfm_gen_tridiag and cfm_mm_read.
bool gen_tridiag(Csr &A, int n) {
    const int64_t nnz = cfm_gen_tridiag(n, 1u, nullptr, nullptr, nullptr);
    if (nnz < 0) return false;
    A.n = n;
    A.base = 0;
    A.rowptr.resize((size_t)n + 1);
    A.col.resize((size_t)nnz);
    A.val.resize((size_t)nnz);
    return cfm_gen_tridiag(n, 1u, A.rowptr.data(), A.col.data(), A.val.data()) == nnz;
}

int read_matrix_market(const std::string &path, Csr &A) {
    int n = 0, nnz = 0, base = 0, *rp = nullptr, *cp = nullptr;
    double *vp = nullptr;
    const int st = cfm_mm_read(path.c_str(), &n, &nnz, &base, &rp, &cp, &vp);
    if (st != CFM_OK) return st;
    A.n = n;
    A.base = base;
    A.rowptr.assign(rp, rp + n + 1);
    A.col.assign(cp, cp + nnz);
    A.val.assign(vp, vp + nnz);
    cfm_host_free(rp);
    cfm_host_free(cp);
    cfm_host_free(vp);
    return CFM_OK;
}

}  // namespace

int main(int argc, char *argv[]) {
    using namespace cli;
    std::setvbuf(stdout, nullptr, _IOLBF, 0);  // the report survives an abnormal end after main
    const std::array<const char *, 11> names{
        "Baseline Discrete Standard", "Baseline Discrete Pipelined", "Single Stream Standard",
        "Single Stream Pipelined", "Single Stream Pipelined Multi Overlap",
        "Single Stream Standard Saxpy Overlap", "Profiling Discrete Standard", "Profiling Discrete Pipelined",
        "Single GPU Discrete Standard", "Single Stream Pipelined Gather",
        "Single Stream Mixed-Precision Refinement"};
    const int methods[11] = {CFM_CG_STANDARD, CFM_CG_PIPELINED, CFM_CG_STANDARD, CFM_CG_PIPELINED,
                             CFM_CG_PIPELINED, CFM_CG_STANDARD, CFM_CG_STANDARD, CFM_CG_PIPELINED,
                             CFM_CG_STANDARD, CFM_CG_PIPELINED, CFM_CG_IR_F32};
    const int modes[11] = {CFM_CG_DISCRETE, CFM_CG_DISCRETE, CFM_CG_PERSISTENT, CFM_CG_PERSISTENT,
                           CFM_CG_PERSISTENT, CFM_CG_PERSISTENT, CFM_CG_DISCRETE, CFM_CG_DISCRETE,
                           CFM_CG_DISCRETE, CFM_CG_PERSISTENT, CFM_CG_PERSISTENT};
    char **b = argv, **e = argv + argc;
    const std::string vlist = get_argval<std::string>(b, e, "-v", "0");
    const bool silent = get_arg(b, e, "-s");
    const int iter_max = get_argval<int>(b, e, "-niter", 10000);
    const int num_runs = get_argval<int>(b, e, "-num_runs", 1);
    const std::string matrix_path = get_argval<std::string>(b, e, "-matrix_path", "");
    const bool compare_single = get_arg(b, e, "-compare-single-gpu") || get_arg(b, e, "-compare");
    if (get_arg(b, e, "-compare-cpu"))
        std::fprintf(stderr, "note: -compare-cpu is not supported (the reference's CPU CG never stores its SpMV "
                             "result); use -compare-single-gpu\n");
    const int nx = get_argval<int>(b, e, "-nx", 128), ny = get_argval<int>(b, e, "-ny", nx),
              nz = get_argval<int>(b, e, "-nz", nx);
    const int tridiag = get_argval<int>(b, e, "-tridiag", 0);
    const double tol = get_argval<double>(b, e, "-tol", 1e-8);
    const unsigned long long seed = get_argval<unsigned long long>(b, e, "-seed", 0ull);
    const std::string rhs = get_argval<std::string>(b, e, "-rhs", "uniform");
    const int quirk = get_argval<int>(b, e, "-quirk", 1);
    const int grid = get_argval<int>(b, e, "-grid", 0);
    const int block = get_argval<int>(b, e, "-block", 0);
    const bool json = get_arg(b, e, "-json");
    const bool phase_flag = get_arg(b, e, "-phase_timers");

    std::vector<int> versions;
    {
        std::stringstream ss(vlist);
        std::string tok;
        while (std::getline(ss, tok, ',')) versions.push_back(std::atoi(tok.c_str()));
    }
    Csr A;
    if (!matrix_path.empty()) {
        const int st = read_matrix_market(matrix_path, A);
        if (st != CFM_OK) return fail("reading the Matrix Market file", st);
    } else if (tridiag > 0) {
        if (!gen_tridiag(A, tridiag)) return fail("cfm_gen_tridiag", CFM_EINVAL);
    } else {
        const int64_t nnz = cfm_gen_poisson3d(nx, ny, nz, nullptr, nullptr, nullptr);
        if (nnz < 0) return fail("cfm_gen_poisson3d", CFM_EINVAL);
        A.n = nx * ny * nz;
        A.rowptr.resize(A.n + 1);
        A.col.resize(nnz);
        A.val.resize(nnz);
        cfm_gen_poisson3d(nx, ny, nz, A.rowptr.data(), A.col.data(), A.val.data());
    }
    const int nnz = A.rowptr[A.n] - A.base;
    std::vector<double> bvec(A.n, 1.0);
    if (rhs != "ones") cfm_gen_uniform(seed, 0, bvec.data(), A.n);

    // versions 6 and 7 are the reference's "Profiling Discrete" ones: every operation is timed
    // (stream synchronisation + a roctx range around it); -phase_timers asks any version for its
    // in-place breakdown (the persistent kernels' own per-workgroup phase clocks)
    auto timed = [&](int version) { return phase_flag || version == 6 || version == 7; };
    auto solve = [&](int version, std::vector<double> &x, cfm_timing &t, int &iters, std::vector<double> &hist,
                     cfm_phase_times *phases) {
        cfm_cg_params p{};
        p.method = methods[version];
        p.mode = modes[version];
        p.tol = tol;
        p.max_iter = iter_max;
        p.restart_quirk = quirk;
        p.grid_blocks = grid;
        p.block_threads = block;
        p.phase_timers = timed(version) ? 1 : 0;
        x.assign(A.n, 0.0);
        hist.assign(iter_max + 2, 0.0);
        cfm_cg_plan *plan = nullptr;
        int st = cfm_cg_plan_create(A.n, nnz, A.rowptr.data(), A.col.data(), A.val.data(), A.base, 0, &p, &plan);
        if (st != CFM_OK) return st;
        st = cfm_cg_plan_set_rhs(plan, bvec.data());
        if (st == CFM_OK) st = cfm_cg_plan_set_x0(plan, nullptr);
        int solve_st = CFM_OK;
        if (st == CFM_OK) {
            solve_st = cfm_cg_plan_solve(plan, hist.data(), (int)hist.size(), &iters, &t);
            if (solve_st == CFM_OK || solve_st == CFM_ENOTCONV || solve_st == CFM_EBREAKDOWN) {
                st = cfm_cg_plan_get_x(plan, x.data());
                if (st == CFM_OK && phases && p.phase_timers) st = cfm_cg_plan_phase_times(plan, phases);
            } else {
                st = solve_st;
            }
        }
        cfm_cg_plan_destroy(plan);
        return st != CFM_OK ? st : solve_st;
    };

    std::vector<double> x_ref;
    double single_runtime = -1.0;
    if (compare_single) {
        cfm_timing t{};
        int it = 0;
        std::vector<double> h;
        const int st = solve(8, x_ref, t, it, h, nullptr);
        if (st != CFM_OK && st != CFM_ENOTCONV) return fail("single-GPU reference solve", st);
        single_runtime = t.wall_ms / 1e3;
    }
    for (int v : versions) {
        if (v < 0 || v >= (int)names.size()) {
            std::fprintf(stderr, "ERROR: -v %d: no such version\n", v);
            return 1;
        }
        if (!silent) print_versions(names, v);
        bool cmp = compare_single;
        for (int run = 1; run <= num_runs; ++run) {
            std::vector<double> x, hist;
            cfm_timing t{};
            int iters = 0;
            cfm_phase_times ph{};
            const int st = solve(v, x, t, iters, hist, &ph);
            if (st != CFM_OK && st != CFM_ENOTCONV) return fail("CG solve", st);
            bool correct = true;
            if (cmp) {
                for (int i = 0; correct && i < A.n; ++i)
                    if (std::fabs(x_ref[i] - x[i]) > 1e-5 || std::isnan(x[i]) || std::isnan(x_ref[i])) {
                        std::fprintf(stderr, "ERROR: x[%d] = %.8f does not match %.8f (Single GPU reference)\n", i,
                                     x[i], x_ref[i]);
                        correct = false;
                    }
            }
            if (correct) {
                const double secs = t.wall_ms / 1e3;
                std::printf("Execution time: %8.4f s\n", secs);
                if (cmp)
                    std::printf("Non-persistent kernel - 1 GPU: %8.4f s, %d GPUs: %8.4f s, speedup: %8.2f, "
                                "efficiency: %8.2f \n",
                                single_runtime, 1, secs, single_runtime / secs, single_runtime / secs * 100);
                const int nh = methods[v] == CFM_CG_IR_F32 ? t.outer_iterations : iters;
                double bb = 0.0;
                for (double q : bvec) bb += q * q;
                std::printf("iterations: %d (outer %d), relative residual: %.3e, kernel: %.4f ms, grid: %d x %d, "
                            "status: %s\n",
                            iters, t.outer_iterations, std::sqrt(hist[nh] / bb), t.kernel_ms, t.grid_blocks,
                            t.block_threads, cfm_status_str(st));
                static const char *phase_names[CFM_NPHASES] = {"spmv", "update", "reduce", "barrier", "other"};
                if (timed(v)) {
                    std::printf("phases (ms, %s):", ph.host_timed ? "host-timed operations"
                                                                  : "mean over the kernel's workgroups");
                    for (int i = 0; i < CFM_NPHASES; ++i) std::printf(" %s %.4f", phase_names[i], ph.ms[i]);
                    std::printf(", total %.4f\n", ph.total_ms);
                }
                if (json) {
                    std::printf("{\"version\": \"%s\", \"iterations\": %d, \"outer_iterations\": %d, "
                                "\"rel_residual\": %.6e, \"kernel_ms\": %.4f, \"wall_ms\": %.4f, \"grid\": %d, "
                                "\"block\": %d, \"status\": \"%s\", \"cooperative\": %d, \"n\": %d, \"nnz\": %d",
                                names[v], iters, t.outer_iterations, std::sqrt(hist[nh] / bb), t.kernel_ms,
                                t.wall_ms, t.grid_blocks, t.block_threads, cfm_status_str(st), t.cooperative, A.n, nnz);
                    if (timed(v)) {
                        std::printf(", \"phases\": {\"host_timed\": %d, \"workgroups\": %d, \"total_ms\": %.4f",
                                    ph.host_timed, ph.workgroups, ph.total_ms);
                        for (int i = 0; i < CFM_NPHASES; ++i)
                            std::printf(", \"%s_ms\": %.4f, \"%s_max_ms\": %.4f", phase_names[i], ph.ms[i],
                                        phase_names[i], ph.max_ms[i]);
                        std::printf("}");
                    }
                    std::printf("}\n");
                }
            } else {
                return 1;
            }
            cmp = false;
        }
    }
    return 0;
}
