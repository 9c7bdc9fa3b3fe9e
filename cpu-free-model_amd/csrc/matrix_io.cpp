// Host-side inputs of the reference's CG program: its default random
// tridiagonal matrix and its Matrix Market reader.  Plain C++ (no device
// code); part of libcfm so the CLI, the tests and any C caller share them.
//
//   cfm_gen_tridiag  <- genTridiag            CG/src/common.cu:46-76
//                       (size and call        CG/src/main.cu:118-129)
//   cfm_mm_read      <- loadMMSparseMatrix    CG/src/mmio_wrapper.cpp:170-369
//                       mm_read_banner / _crd CG/src/mmio.c:91-160, :270-302
//                       index-base test       CG/src/main.cu:180-186
#include <stdlib.h>

#include <algorithm>
#include <cctype>
#include <cerrno>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/cfm.h"

namespace cfm {
void set_error(const char *fmt, ...) __attribute__((format(printf, 1, 2)));
}

// ------------------------------------------------------------ tridiagonal --
// The reference fills, in this order of rand() calls: a(0,0), a(0,1), then
// for every row i >= 1 the diagonal a(i,i) and, below the last row, a(i,i+1);
// a(i,i-1) repeats a(i-1,i).  Diagonal = rand()/RAND_MAX + 10, off-diagonal
// = rand()/RAND_MAX, in double (`real` is double in the CG program).  The
// generator is the C library's rand() stream (glibc: random() with its
// default 128-byte state), run on a private state so the caller's rand()
// sequence is untouched; seed 1 is what an unseeded program gets.
extern "C" int64_t cfm_gen_tridiag(int n, unsigned int seed, int *rowptr, int *col, double *val) {
    if (n < 2 || n > (0x7fffffff - 4) / 3) return -1;
    const int64_t nnz = 3LL * n - 2;
    if (!rowptr && !col && !val) return nnz;
    if (!rowptr || !col || !val) return -1;
    char state[128];
    struct random_data rd;
    std::memset(&rd, 0, sizeof(rd));
    if (initstate_r(seed, state, sizeof(state), &rd) != 0) return -1;
    auto next = [&rd]() {
        int32_t v = 0;
        random_r(&rd, &v);
        return (double)v / RAND_MAX;
    };
    rowptr[0] = 0;
    col[0] = 0;
    col[1] = 1;
    val[0] = next() + 10.0;
    val[1] = next();
    int e = 2;  // entries written so far; row i starts at 3i - 1
    for (int i = 1; i < n; ++i) {
        rowptr[i] = e;
        col[e] = i - 1;
        val[e] = val[e - 1];  // a(i,i-1) = a(i-1,i), the entry just written
        col[e + 1] = i;
        val[e + 1] = next() + 10.0;
        e += 2;
        if (i < n - 1) {
            col[e] = i + 1;
            val[e] = next();
            ++e;
        }
    }
    rowptr[n] = e;
    return nnz;
}

// --------------------------------------------------------- Matrix Market --
namespace {

struct Entry {
    int i, j;
    double v;
};

// Next line that does not start with '%' (the size line may follow comments).
bool next_data_line(FILE *f, std::string &line) {
    char buf[1030];
    for (;;) {
        if (!std::fgets(buf, sizeof(buf), f)) return false;
        if (buf[0] != '%') {
            line = buf;
            return true;
        }
    }
}

std::string lower(const char *s) {
    std::string r(s);
    for (auto &c : r) c = (char)std::tolower((unsigned char)c);
    return r;
}

}  // namespace

// Accepted as the reference accepts: "%%MatrixMarket matrix coordinate" with
// data type real or integer (pattern and complex are refused, as are dense
// "array" files) and storage general, symmetric or skew-symmetric (hermitian
// is not valid for real data); banner words are case-insensitive.  Symmetric
// storage is expanded by mirroring every off-diagonal entry (negated for
// skew-symmetric); entries are sorted by (row, column); a repeated (row,
// column) pair fails the reference's sorted-column check and is refused.
// Index base: 0 if any row or column index is 0, 1 if any equals the matrix
// dimension, refused if both, else 0 -- the reference's test; rowptr starts at
// the base and the column indices are kept as read.  Refused beyond the
// reference: non-square matrices (CG needs a square one) and indices outside
// [0, dimension], which the reference would write out of bounds.
extern "C" cfm_status cfm_mm_read(const char *path, int *n_out, int *nnz_out, int *base_out, int **rowptr_out,
                                  int **col_out, double **val_out) {
    if (!path || !n_out || !nnz_out || !base_out || !rowptr_out || !col_out || !val_out) return CFM_EINVAL;
    *rowptr_out = nullptr;
    *col_out = nullptr;
    *val_out = nullptr;
    FILE *f = std::fopen(path, "r");
    if (!f) {
        cfm::set_error("mm: cannot open %s: %s", path, std::strerror(errno));
        return CFM_EINVAL;
    }
    struct Closer {
        FILE *f;
        ~Closer() { std::fclose(f); }
    } closer{f};
    auto refuse = [&](const char *why) {
        cfm::set_error("mm: %s: %s", path, why);
        return CFM_EINVAL;
    };
    char line[1030];
    if (!std::fgets(line, sizeof(line), f)) return refuse("empty file");
    char banner[1030], object[1030], format[1030], field[1030], symmetry[1030];
    if (std::sscanf(line, "%1029s %1029s %1029s %1029s %1029s", banner, object, format, field, symmetry) != 5)
        return refuse("incomplete banner line");
    if (std::strncmp(banner, "%%MatrixMarket", 14) != 0) return refuse("no MatrixMarket banner");
    const std::string obj = lower(object), fmt = lower(format), typ = lower(field), sym = lower(symmetry);
    if (obj != "matrix") return refuse("object is not a matrix");
    if (fmt != "coordinate") return refuse("dense (array) files are not supported");
    if (typ == "pattern" || typ == "complex") return refuse("pattern and complex matrices are not supported");
    if (typ != "real" && typ != "integer") return refuse("unknown data type");
    const bool symmetric = sym == "symmetric", skew = sym == "skew-symmetric";
    if (!symmetric && !skew && sym != "general") return refuse("unsupported storage scheme");

    std::string size_line;
    if (!next_data_line(f, size_line)) return refuse("missing size line");
    long rows = 0, cols = 0, entries = 0;
    if (std::sscanf(size_line.c_str(), "%ld %ld %ld", &rows, &cols, &entries) != 3) {
        // a blank size line: the counts follow as the next three integers
        if (std::fscanf(f, "%ld %ld %ld", &rows, &cols, &entries) != 3) return refuse("missing size line");
    }
    if (rows < 1 || cols < 1 || entries < 0 || rows > 0x7ffffff0L || entries > 0x3ffffff0L)
        return refuse("bad size line");
    if (rows != cols) return refuse("matrix is not square");
    std::vector<Entry> e;
    e.reserve((size_t)entries * ((symmetric || skew) ? 2 : 1));
    for (long k = 0; k < entries; ++k) {
        int i, j;
        double v;
        if (std::fscanf(f, "%d %d %lg", &i, &j, &v) != 3) return refuse("premature end of entries");
        if (i < 0 || j < 0 || i > rows || j > cols) return refuse("index outside the matrix");
        e.push_back({i, j, v});
        if ((symmetric || skew) && i != j) e.push_back({j, i, skew ? -v : v});
    }
    if ((long long)e.size() > 0x7ffffff0LL) return refuse("too many entries");
    std::stable_sort(e.begin(), e.end(), [](const Entry &a, const Entry &b) {
        return a.i != b.i ? a.i < b.i : a.j < b.j;
    });
    for (size_t k = 1; k < e.size(); ++k)
        if (e[k].i == e[k - 1].i && e[k].j == e[k - 1].j)
            return refuse("repeated (row, column) entry (the reference's sorted-column check fails)");
    bool base0 = false, base1 = false;
    for (const Entry &x : e) {
        if (x.i == 0 || x.j == 0) base0 = true;
        if (x.i == rows || x.j == cols) base1 = true;
    }
    if (base0 && base1) return refuse("indices are both 0-based and 1-based");
    const int base = base1 ? 1 : 0;
    const int n = (int)rows, nnz = (int)e.size();
    int *rp = static_cast<int *>(std::malloc(sizeof(int) * ((size_t)n + 1)));
    int *cp = static_cast<int *>(std::malloc(sizeof(int) * std::max<size_t>(nnz, 1)));
    double *vp = static_cast<double *>(std::malloc(sizeof(double) * std::max<size_t>(nnz, 1)));
    if (!rp || !cp || !vp) {
        std::free(rp);
        std::free(cp);
        std::free(vp);
        cfm::set_error("mm: %s: host allocation failed", path);
        return CFM_ENOMEM;
    }
    // counts per row, then prefix sums starting at the base
    std::fill(rp, rp + n + 1, 0);
    for (const Entry &x : e) {
        const int r = x.i - base;
        if (r < 0 || r >= n) {  // guard only: the index and base tests above keep r in [0, n)
            std::free(rp);
            std::free(cp);
            std::free(vp);
            return refuse("row index outside the matrix for the detected base");
        }
        ++rp[r + 1];
    }
    rp[0] = base;
    for (int r = 0; r < n; ++r) rp[r + 1] += rp[r];
    for (int k = 0; k < nnz; ++k) {
        cp[k] = e[k].j;
        vp[k] = e[k].v;
    }
    *n_out = n;
    *nnz_out = nnz;
    *base_out = base;
    *rowptr_out = rp;
    *col_out = cp;
    *val_out = vp;
    return CFM_OK;
}

extern "C" void cfm_host_free(void *p) { std::free(p); }
