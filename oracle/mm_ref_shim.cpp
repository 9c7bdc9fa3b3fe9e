// C entry point into the reference's own Matrix Market reader, compiled from
// the reference sources where they lie (CG/src/mmio_wrapper.cpp, CG/src/mmio.c)
// into oracle/_ref/libmmref.so.  TEST INFRASTRUCTURE ONLY: the tests use it to
// check the library's cfm_mm_read against the reference's loadMMSparseMatrix
// on the same files; nothing in the product links it.
//
// The template is defined and explicitly instantiated for double in
// CG/src/mmio_wrapper.cpp:170-378; its declaration in CG/include/common.h sits
// next to CUDA headers, so the caller-side prototype is written here.
template <typename T_ELEM>
int loadMMSparseMatrix(char *filename, char elem_type, bool csrFormat, int *m, int *n, int *nnz, T_ELEM **aVal,
                       int **aRowInd, int **aColInd, int extendSymMatrix);

// The call CG/src/main.cu:131-133 makes: elem 'd', CSR, symmetric expansion on.
extern "C" int ref_load_mm(const char *path, int *m, int *n, int *nnz, double **val, int **rowptr, int **col) {
    return loadMMSparseMatrix<double>(const_cast<char *>(path), 'd', true, m, n, nnz, val, rowptr, col, 1);
}
