// cg: Conjugate Gradient driver with the reference's version table and flags
// (CG/src/main.cu:23-278) on top of the cfm C ABI.
//
// Matrix: -nx/-ny/-nz build the 3D 7-point Poisson system (default 128^3);
// -tridiag N builds the reference's random SPD tridiagonal (CG/src/common.cu:46-76,
// glibc rand(), default seed); -matrix_path reads a Matrix Market file
// (coordinate, real/integer/pattern, general or symmetric).
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
    std::vector<int> rowptr, col;
    std::vector<double> val;
};

void gen_tridiag(Csr &A, int n) {
    // CG/src/common.cu:46-76, same rand() call order
    const int nnz = (n - 2) * 3 + 4;
    A.n = n;
    A.rowptr.assign(n + 1, 0);
    A.col.assign(nnz, 0);
    A.val.assign(nnz, 0.0);
    int *I = A.rowptr.data(), *J = A.col.data();
    double *val = A.val.data();
    I[0] = 0, J[0] = 0, J[1] = 1;
    val[0] = (double)rand() / RAND_MAX + 10.0f;
    val[1] = (double)rand() / RAND_MAX;
    for (int i = 1; i < n; i++) {
        if (i > 1)
            I[i] = I[i - 1] + 3;
        else
            I[1] = 2;
        const int start = (i - 1) * 3 + 2;
        J[start] = i - 1;
        J[start + 1] = i;
        if (i < n - 1) J[start + 2] = i + 1;
        val[start] = val[start - 1];
        val[start + 1] = (double)rand() / RAND_MAX + 10.0f;
        if (i < n - 1) val[start + 2] = (double)rand() / RAND_MAX;
    }
    I[n] = nnz;
}

bool read_matrix_market(const std::string &path, Csr &A) {
    std::ifstream in(path);
    if (!in) return false;
    std::string line;
    std::getline(in, line);
    if (line.rfind("%%MatrixMarket", 0) != 0) return false;
    const bool pattern = line.find("pattern") != std::string::npos;
    const bool symmetric = line.find("symmetric") != std::string::npos;
    if (line.find("coordinate") == std::string::npos || line.find("complex") != std::string::npos) return false;
    while (std::getline(in, line))
        if (!line.empty() && line[0] != '%') break;
    long rows = 0, cols = 0, entries = 0;
    std::istringstream hdr(line);
    hdr >> rows >> cols >> entries;
    if (rows <= 0 || rows != cols || rows > 0x7ffffff0) return false;
    std::vector<std::map<int, double>> r(rows);
    for (long k = 0; k < entries; ++k) {
        long i, j;
        double v = 1.0;
        if (!(in >> i >> j)) return false;
        if (!pattern && !(in >> v)) return false;
        if (i < 1 || j < 1 || i > rows || j > rows) return false;
        r[i - 1][(int)(j - 1)] += v;
        if (symmetric && i != j) r[j - 1][(int)(i - 1)] += v;
    }
    A.n = (int)rows;
    A.rowptr.assign(rows + 1, 0);
    for (long i = 0; i < rows; ++i) {
        for (auto &kv : r[i]) {
            A.col.push_back(kv.first);
            A.val.push_back(kv.second);
        }
        A.rowptr[i + 1] = (int)A.col.size();
    }
    return true;
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

    std::vector<int> versions;
    {
        std::stringstream ss(vlist);
        std::string tok;
        while (std::getline(ss, tok, ',')) versions.push_back(std::atoi(tok.c_str()));
    }
    Csr A;
    if (!matrix_path.empty()) {
        if (!read_matrix_market(matrix_path, A)) {
            std::fprintf(stderr, "ERROR: cannot read Matrix Market file %s\n", matrix_path.c_str());
            return 1;
        }
    } else if (tridiag > 0) {
        gen_tridiag(A, tridiag);
    } else {
        const int64_t nnz = cfm_gen_poisson3d(nx, ny, nz, nullptr, nullptr, nullptr);
        if (nnz < 0) return fail("cfm_gen_poisson3d", CFM_EINVAL);
        A.n = nx * ny * nz;
        A.rowptr.resize(A.n + 1);
        A.col.resize(nnz);
        A.val.resize(nnz);
        cfm_gen_poisson3d(nx, ny, nz, A.rowptr.data(), A.col.data(), A.val.data());
    }
    const int nnz = A.rowptr[A.n];
    std::vector<double> bvec(A.n, 1.0);
    if (rhs != "ones") cfm_gen_uniform(seed, 0, bvec.data(), A.n);

    auto solve = [&](int version, std::vector<double> &x, cfm_timing &t, int &iters, std::vector<double> &hist) {
        cfm_cg_params p{};
        p.method = methods[version];
        p.mode = modes[version];
        p.tol = tol;
        p.max_iter = iter_max;
        p.restart_quirk = quirk;
        p.grid_blocks = grid;
        p.block_threads = block;
        x.assign(A.n, 0.0);
        hist.assign(iter_max + 2, 0.0);
        return cfm_cg_solve(A.n, nnz, A.rowptr.data(), A.col.data(), A.val.data(), 0, bvec.data(), x.data(), &p,
                            hist.data(), (int)hist.size(), &iters, &t);
    };

    std::vector<double> x_ref;
    double single_runtime = -1.0;
    if (compare_single) {
        cfm_timing t{};
        int it = 0;
        std::vector<double> h;
        const int st = solve(8, x_ref, t, it, h);
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
            const int st = solve(v, x, t, iters, hist);
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
            } else {
                return 1;
            }
            cmp = false;
        }
    }
    return 0;
}
