// Wavelet-compressed dense matrices for the time stepping (the reference's CompressedMatrix<1>,
// include/wavelets.hpp:402-448, with the Haar filters h_2 / g_2 of :15-17).
//
// A matrix A (n x n, zero-padded to np x np, np the power of two >= n) goes to Beylkin's nonstandard form:
// at each level of size sz = np/2, np/4, ..., 1 the 2x2 blocks of the current coarse matrix s give
//   alpha = sum g_k g_m s(2i+k, 2l+m)   at (off + i,      off + l)
//   beta  = sum g_k h_m s(...)          at (off + i,      off + sz + l)
//   gamma = sum h_k g_m s(...)          at (off + sz + i, off + l)
//   s'    = sum h_k h_m s(...)          (the next level's coarse matrix),   off += 2 sz,
// and the last s' lands at (2np - 3, 2np - 3) of a (2np - 2)^2 matrix (wavelets.hpp:154-219).  Of these np^2
// coefficients the num_elems largest in magnitude (and nonzero) are kept.  Vectors go to the matching
// redundant form of length 2np - 2 (decompose_vec, :221-251), products are formed in that space and
// mapped back by reconstruct (:253-281).
//
// Here the transform and the selection run on the GPU: one kernel per level, a radix select of the
// num_elems-th magnitude over the 64-bit patterns of |coefficient| (six histogram passes), and a
// per-row compaction to CSR in column order.  Coefficients tied with the threshold magnitude are kept
// in row-major order (the reference's heap keeps some tie set of its own choosing).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace cqbem {

// CSR matrix in device memory (rows = 2 np - 2; rowptr has rows + 1 entries).
struct DeviceCsr {
    int64_t n = 0;     // size of the matrix it compresses
    int64_t rows = 0;  // 2 np - 2
    int64_t nnz = 0;
    int64_t *rowptr = nullptr;
    int32_t *col = nullptr;
    double *val = nullptr;
    void release();
};

// np for an n x n matrix: the power of two >= n (std::round(std::pow(2, std::ceil(std::log2(n))))).
int64_t wavelet_padded(int64_t n);

// Compresses the row-major n x n device matrix A (n >= 2) keeping num_elems coefficients (num_elems <= 0: every
// nonzero coefficient, as decompose_mat does without a count, wavelets.hpp:180-184); `scale` multiplies
// the kept values (exact for -1, which gives the compression of -A).  Allocates out's arrays.  Runs on
// stream st and returns after it has finished.  Throws std::runtime_error on HIP errors.
void wavelet_compress(const double *A, int64_t n, int64_t num_elems, double scale, DeviceCsr &out, hipStream_t st);

// Same, into caller-provided arrays: rowptr (rows + 1), col and val (capacity num_elems, or np^2 when
// num_elems <= 0); returns nnz.
int64_t wavelet_compress_into(const double *A, int64_t n, int64_t num_elems, double scale, int64_t *rowptr, int32_t *col,
                              double *val, hipStream_t st);

// y = M x for one CSR matrix (device pointers).
hipError_t launch_csr_matvec(const int64_t *rowptr, const int32_t *col, const double *val, const double *x, int64_t rows,
                             double *y, hipStream_t st);

// y = sum_i M_i x_i over m CSR matrices of equal row count (device pointers; x_i of length rows).
struct CsrRef {
    const int64_t *rowptr;
    const int32_t *col;
    const double *val;
    const double *x;
};
hipError_t launch_csr_matvec_sum(const CsrRef *refs_device, int m, int64_t rows, double *y, hipStream_t st);

// Host-side vector transforms (O(n)).  decompose: x (n) -> 2 np - 2 coefficients of the zero-padded x;
// reconstruct: 2 np - 2 coefficients of a product -> the first n entries of the result.
std::vector<double> wavelet_decompose_vec(const double *x, int64_t n);
std::vector<double> wavelet_reconstruct(const double *c, int64_t n);

}  // namespace cqbem
