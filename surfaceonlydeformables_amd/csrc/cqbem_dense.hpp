// Dense FP64 factorisations used after the assembly, on the GPU: the partial-pivoting LU of H_0
// (the reference's Eigen::PartialPivLU<MatrixXs> H_inv, src/elastodynamics_cqbem_matrix_collection.cpp:184-215)
// through rocSOLVER, and a complete-orthogonal least-squares solve for the thin 3N x 3 B blocks
// (Eigen::CompleteOrthogonalDecomposition B_trans_inv / B_euler_inv, same lines) on the host.
//
// rocSOLVER and rocBLAS are opened with dlopen when the first factorisation is requested, so
// libcqbem.so itself keeps no link-time dependency on them.
#pragma once
#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

namespace cqbem {

// Opens rocBLAS and rocSOLVER and runs a small LU and solves, so that their first-use cost (library load,
// handle set-up, kernel code objects: seconds in a fresh process) is not paid inside a later DeviceLU::factor.
// Callable from any thread, beside other GPU work on non-blocking streams; failures are left to factor().
void warm_up_lu();

// LU factors of one n x n row-major matrix, resident in device memory.
class DeviceLU {
  public:
    DeviceLU() = default;
    DeviceLU(const DeviceLU &) = delete;
    DeviceLU &operator=(const DeviceLU &) = delete;
    ~DeviceLU();

    // Factorises the row-major matrix A (host, n x n).  Returns false (and sets err) on a library
    // error or an exactly singular matrix.
    bool factor(int64_t n, const double *A, std::string &err);
    // Solves A X = B in place for the row-major n x nrhs host matrix B.  Thread-safe.
    bool solve(double *B, int64_t nrhs, std::string &err) const;
    // Writes A^-1 (row-major n x n) to the device buffer dX, from the factors (for repeated one-vector
    // solves as a product: the triangular solves of one right-hand side are latency-bound).  Thread-safe.
    bool inverse_to_device(double *dX, std::string &err) const;
    int64_t size() const { return n_; }

  private:
    int64_t n_ = 0;
    double *lu_ = nullptr;  // column-major factors of A^T (the row-major A read as column-major)
    int32_t *ipiv_ = nullptr;
    void *handle_ = nullptr;
    mutable std::mutex mu_;  // the rocBLAS handle serves one call at a time
};

// Least-squares (minimum-norm for full column rank) solver for a thin m x k row-major matrix B,
// k <= 3: X = argmin |B X - Y| through Householder QR, for row-major right-hand sides Y (m x r).
class ThinLeastSquares {
  public:
    bool compute(const std::vector<double> &B, int64_t m, int k, std::string &err);
    // X is k x r row-major.
    std::vector<double> solve(const double *Y, int64_t r) const;

  private:
    int64_t m_ = 0;
    int k_ = 0;
    std::vector<double> qr_, tau_;  // Householder vectors below the diagonal of R (row-major m x k)
};

}  // namespace cqbem
