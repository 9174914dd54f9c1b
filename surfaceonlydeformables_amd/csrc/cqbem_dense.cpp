// Dense factorisations after the assembly: rocSOLVER LU on the GPU and a thin Householder least-squares
// solve on the host.  See cqbem_dense.hpp.
#include "cqbem_dense.hpp"

#include "host_parallel.hpp"

#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <dlfcn.h>

#include <algorithm>
#include <cmath>
#include <mutex>

namespace cqbem {

namespace {

struct RocLib {
    decltype(&rocblas_create_handle) create_handle = nullptr;
    decltype(&rocblas_destroy_handle) destroy_handle = nullptr;
    decltype(&rocsolver_dgetrf) dgetrf = nullptr;
    decltype(&rocsolver_dgetrs) dgetrs = nullptr;
    decltype(&rocblas_dgeam) dgeam = nullptr;  // optional: device transposes of many right-hand sides
    std::string error;
};

void *open_first(std::initializer_list<const char *> names) {
    for (const char *n : names)
        if (void *h = dlopen(n, RTLD_NOW | RTLD_GLOBAL)) return h;
    return nullptr;
}

// Opens rocBLAS and rocSOLVER once per process.  When a process already holds copies (a framework
// that bundles them), dlopen returns those, so the HIP runtime stays single.
const RocLib &roclib() {
    static RocLib lib;
    static std::once_flag once;
    std::call_once(once, [] {
        void *blas = open_first({"librocblas.so.5", "librocblas.so"});
        void *solver = open_first({"librocsolver.so.0", "librocsolver.so"});
        if (!blas || !solver) {
            lib.error = "rocBLAS/rocSOLVER not found";
            return;
        }
        lib.create_handle = reinterpret_cast<decltype(lib.create_handle)>(dlsym(blas, "rocblas_create_handle"));
        lib.destroy_handle = reinterpret_cast<decltype(lib.destroy_handle)>(dlsym(blas, "rocblas_destroy_handle"));
        lib.dgetrf = reinterpret_cast<decltype(lib.dgetrf)>(dlsym(solver, "rocsolver_dgetrf"));
        lib.dgetrs = reinterpret_cast<decltype(lib.dgetrs)>(dlsym(solver, "rocsolver_dgetrs"));
        lib.dgeam = reinterpret_cast<decltype(lib.dgeam)>(dlsym(blas, "rocblas_dgeam"));
        if (!lib.create_handle || !lib.destroy_handle || !lib.dgetrf || !lib.dgetrs) lib.error = "rocSOLVER symbols missing";
    });
    return lib;
}

bool hip_ok(hipError_t e, const char *what, std::string &err) {
    if (e == hipSuccess) return true;
    err = std::string(what) + ": " + hipGetErrorString(e);
    return false;
}

// Row-major (rows x cols) -> column-major, i.e. a transpose into `out`.
void to_colmajor(const double *in, int64_t rows, int64_t cols, double *out) {
    constexpr int64_t kB = 64;
    parallel_for((rows + kB - 1) / kB, [&](int64_t b0, int64_t b1) {
        for (int64_t r0 = b0 * kB; r0 < std::min(rows, b1 * kB); r0 += kB)
            for (int64_t c0 = 0; c0 < cols; c0 += kB)
                for (int64_t r = r0; r < std::min(rows, r0 + kB); r++)
                    for (int64_t c = c0; c < std::min(cols, c0 + kB); c++) out[c * rows + r] = in[r * cols + c];
    });
}

}  // namespace

void warm_up_lu() {
    const RocLib &lib = roclib();
    if (!lib.error.empty()) return;
    rocblas_handle h = nullptr;
    if (lib.create_handle(&h) != rocblas_status_success) return;
    constexpr int n = 2048;  // large enough for the blocked factorisation and its trailing-update products
    std::vector<double> eye((size_t)n * n, 0.0);
    for (int i = 0; i < n; i++) eye[(size_t)i * n + i] = 2.0;
    double *a = nullptr, *b = nullptr;
    int32_t *piv = nullptr;
    if (hipMalloc(&a, sizeof(double) * eye.size()) == hipSuccess && hipMalloc(&b, sizeof(double) * eye.size()) == hipSuccess &&
        hipMalloc(&piv, sizeof(int32_t) * (n + 1)) == hipSuccess &&
        hipMemcpy(a, eye.data(), sizeof(double) * eye.size(), hipMemcpyHostToDevice) == hipSuccess &&
        hipMemcpy(b, eye.data(), sizeof(double) * eye.size(), hipMemcpyHostToDevice) == hipSuccess &&
        lib.dgetrf(h, n, n, a, n, piv, piv + n) == rocblas_status_success) {
        (void)lib.dgetrs(h, rocblas_operation_transpose, n, 3, a, n, piv, b, n);
        (void)lib.dgetrs(h, rocblas_operation_transpose, n, n, a, n, piv, b, n);
        (void)hipStreamSynchronize(nullptr);  // the handle's (null) stream only
    }
    for (void *p : {(void *)a, (void *)b, (void *)piv})
        if (p) (void)hipFree(p);
    (void)hipGetLastError();
    lib.destroy_handle(h);
}

DeviceLU::~DeviceLU() {
    if (lu_) (void)hipFree(lu_);
    if (ipiv_) (void)hipFree(ipiv_);
    if (handle_) roclib().destroy_handle(static_cast<rocblas_handle>(handle_));
}

bool DeviceLU::factor(int64_t n, const double *A, std::string &err) {
    const RocLib &lib = roclib();
    if (!lib.error.empty()) return err = lib.error, false;
    if (n <= 0 || n > (int64_t)INT32_MAX / 8) return err = "LU size out of range", false;
    if (!handle_) {
        rocblas_handle h = nullptr;
        if (lib.create_handle(&h) != rocblas_status_success) return err = "rocblas_create_handle failed", false;
        handle_ = h;
    }
    if (lu_) (void)hipFree(lu_), lu_ = nullptr;
    if (ipiv_) (void)hipFree(ipiv_), ipiv_ = nullptr;
    n_ = n;
    int32_t *info = nullptr;
    if (!hip_ok(hipMalloc(&lu_, sizeof(double) * n * n), "hipMalloc LU", err)) return false;
    if (!hip_ok(hipMalloc(&ipiv_, sizeof(int32_t) * (n + 1)), "hipMalloc pivots", err)) return false;
    info = ipiv_ + n;
    // the row-major A read as column-major is A^T; factorising A^T and solving with the transpose
    // solves A x = b
    if (!hip_ok(hipMemcpy(lu_, A, sizeof(double) * n * n, hipMemcpyHostToDevice), "hipMemcpy A", err)) return false;
    if (lib.dgetrf(static_cast<rocblas_handle>(handle_), (rocblas_int)n, (rocblas_int)n, lu_, (rocblas_int)n, ipiv_, info) !=
        rocblas_status_success)
        return err = "rocsolver_dgetrf failed", false;
    int32_t h_info = 0;
    if (!hip_ok(hipMemcpy(&h_info, info, sizeof(int32_t), hipMemcpyDeviceToHost), "hipMemcpy info", err)) return false;
    if (h_info != 0) return err = "H_0 is singular (zero pivot " + std::to_string(h_info) + ")", false;
    return true;
}

bool DeviceLU::solve(double *B, int64_t nrhs, std::string &err) const {
    std::lock_guard<std::mutex> lk(mu_);
    const RocLib &lib = roclib();
    if (!lu_) return err = "LU not computed", false;
    if (nrhs > 1 && lib.dgeam) {
        // The row-major n x nrhs B is a column-major nrhs x n matrix: transposed on the device (rocblas_dgeam)
        // into the column-major layout getrs solves in, and back; the host only copies.
        rocblas_handle h = static_cast<rocblas_handle>(handle_);
        const double one = 1.0, zero = 0.0;
        double *d = nullptr, *t = nullptr;
        bool ok = hip_ok(hipMalloc(&d, sizeof(double) * n_ * nrhs), "hipMalloc rhs", err) &&
                  hip_ok(hipMalloc(&t, sizeof(double) * n_ * nrhs), "hipMalloc rhs", err) &&
                  hip_ok(hipMemcpy(t, B, sizeof(double) * n_ * nrhs, hipMemcpyHostToDevice), "hipMemcpy rhs", err) &&
                  // dgeam's B operand is d itself with factor 0: zeroed so that no stale NaN survives 0 * x
                  hip_ok(hipMemset(d, 0, sizeof(double) * n_ * nrhs), "hipMemset rhs", err);
        if (ok && lib.dgeam(h, rocblas_operation_transpose, rocblas_operation_none, (rocblas_int)n_, (rocblas_int)nrhs, &one, t,
                            (rocblas_int)nrhs, &zero, d, (rocblas_int)n_, d, (rocblas_int)n_) != rocblas_status_success)
            ok = false, err = "rocblas_dgeam failed";
        if (ok && lib.dgetrs(h, rocblas_operation_transpose, (rocblas_int)n_, (rocblas_int)nrhs, lu_, (rocblas_int)n_, ipiv_, d,
                             (rocblas_int)n_) != rocblas_status_success)
            ok = false, err = "rocsolver_dgetrs failed";
        if (ok && lib.dgeam(h, rocblas_operation_transpose, rocblas_operation_none, (rocblas_int)nrhs, (rocblas_int)n_, &one, d,
                            (rocblas_int)n_, &zero, t, (rocblas_int)nrhs, t, (rocblas_int)nrhs) != rocblas_status_success)
            ok = false, err = "rocblas_dgeam failed";
        if (ok) ok = hip_ok(hipMemcpy(B, t, sizeof(double) * n_ * nrhs, hipMemcpyDeviceToHost), "hipMemcpy x", err);
        if (d) (void)hipFree(d);
        if (t) (void)hipFree(t);
        return ok;
    }
    std::vector<double> cm;
    const double *src = B;
    if (nrhs > 1) {
        cm.resize((size_t)(n_ * nrhs));
        to_colmajor(B, n_, nrhs, cm.data());
        src = cm.data();
    }
    double *d = nullptr;
    if (!hip_ok(hipMalloc(&d, sizeof(double) * n_ * nrhs), "hipMalloc rhs", err)) return false;
    bool ok = hip_ok(hipMemcpy(d, src, sizeof(double) * n_ * nrhs, hipMemcpyHostToDevice), "hipMemcpy rhs", err);
    if (ok && lib.dgetrs(static_cast<rocblas_handle>(handle_), rocblas_operation_transpose, (rocblas_int)n_,
                         (rocblas_int)nrhs, lu_, (rocblas_int)n_, ipiv_, d, (rocblas_int)n_) != rocblas_status_success)
        ok = false, err = "rocsolver_dgetrs failed";
    if (ok) ok = hip_ok(hipMemcpy(nrhs > 1 ? cm.data() : B, d, sizeof(double) * n_ * nrhs, hipMemcpyDeviceToHost), "hipMemcpy x", err);
    (void)hipFree(d);
    if (ok && nrhs > 1) to_colmajor(cm.data(), nrhs, n_, B);  // column-major n x nrhs back to row-major
    return ok;
}

bool DeviceLU::inverse_to_device(double *dX, std::string &err) const {
    std::lock_guard<std::mutex> lk(mu_);
    const RocLib &lib = roclib();
    if (!lu_) return err = "LU not computed", false;
    std::vector<double> eye((size_t)(n_ * n_), 0.0);
    for (int64_t i = 0; i < n_; i++) eye[(size_t)(i * n_ + i)] = 1.0;
    if (!hip_ok(hipMemcpy(dX, eye.data(), sizeof(double) * n_ * n_, hipMemcpyHostToDevice), "hipMemcpy identity", err))
        return false;
    // lu_ factors A^T; A^T Y = I gives the column-major Y = A^-T, whose memory read row-major is A^-1
    if (lib.dgetrs(static_cast<rocblas_handle>(handle_), rocblas_operation_none, (rocblas_int)n_, (rocblas_int)n_, lu_,
                   (rocblas_int)n_, ipiv_, dX, (rocblas_int)n_) != rocblas_status_success)
        return err = "rocsolver_dgetrs failed", false;
    return hip_ok(hipDeviceSynchronize(), "inverse", err);
}

bool ThinLeastSquares::compute(const std::vector<double> &B, int64_t m, int k, std::string &err) {
    if (k < 1 || k > 3 || m < k || (int64_t)B.size() != m * k) return err = "least-squares shape", false;
    m_ = m, k_ = k;
    qr_ = B;
    tau_.assign(k, 0.0);
    double rmax = 0.0;
    for (int j = 0; j < k; j++) {
        // Householder reflector zeroing column j below the diagonal
        double norm2 = 0.0;
        for (int64_t i = j; i < m; i++) norm2 += qr_[i * k + j] * qr_[i * k + j];
        const double alpha = qr_[j * k + j], norm = std::sqrt(norm2);
        if (norm == 0.0) return err = "rank-deficient B block", false;
        const double beta = alpha >= 0 ? -norm : norm;
        tau_[j] = (beta - alpha) / beta;
        const double scale = 1.0 / (alpha - beta);
        for (int64_t i = j + 1; i < m; i++) qr_[i * k + j] *= scale;
        qr_[j * k + j] = beta;
        for (int c = j + 1; c < k; c++) {  // apply to the remaining columns
            double s = qr_[j * k + c];
            for (int64_t i = j + 1; i < m; i++) s += qr_[i * k + j] * qr_[i * k + c];
            s *= tau_[j];
            qr_[j * k + c] -= s;
            for (int64_t i = j + 1; i < m; i++) qr_[i * k + c] -= s * qr_[i * k + j];
        }
        rmax = std::max(rmax, std::fabs(beta));
    }
    for (int j = 0; j < k; j++)
        if (std::fabs(qr_[j * k + j]) <= 1e-13 * rmax) return err = "rank-deficient B block", false;
    return true;
}

std::vector<double> ThinLeastSquares::solve(const double *Y, int64_t r) const {
    const int64_t m = m_;
    const int k = k_;
    // thin Q (m x k) = H_0 ... H_{k-1} [I; 0]
    std::vector<double> Q((size_t)(m * k), 0.0);
    for (int j = 0; j < k; j++) Q[j * k + j] = 1.0;
    for (int j = k - 1; j >= 0; j--)
        for (int c = 0; c < k; c++) {
            double s = Q[j * k + c];
            for (int64_t i = j + 1; i < m; i++) s += qr_[i * k + j] * Q[i * k + c];
            s *= tau_[j];
            Q[j * k + c] -= s;
            for (int64_t i = j + 1; i < m; i++) Q[i * k + c] -= s * qr_[i * k + j];
        }
    // Z = Q^T Y (k x r), streamed over the rows of Y; then X = R^-1 Z.  Columns of Y are split over threads.
    std::vector<double> X((size_t)(k * r), 0.0);
    parallel_for(r, [&](int64_t t0, int64_t t1) {
        for (int64_t i = 0; i < m; i++) {
            const double *y = Y + i * r;
            for (int c = 0; c < k; c++) {
                const double q = Q[i * k + c];
                double *z = X.data() + c * r;
                for (int64_t t = t0; t < t1; t++) z[t] += q * y[t];
            }
        }
        for (int c = k - 1; c >= 0; c--) {
            double *x = X.data() + c * r;
            for (int d = c + 1; d < k; d++) {
                const double rcd = qr_[c * k + d];
                const double *xd = X.data() + d * r;
                for (int64_t t = t0; t < t1; t++) x[t] -= rcd * xd[t];
            }
            const double inv = 1.0 / qr_[c * k + c];
            for (int64_t t = t0; t < t1; t++) x[t] *= inv;
        }
    });
    return X;
}

}  // namespace cqbem
