// Haar nonstandard-form compression of dense matrices on the GPU, CSR products, and the vector
// transforms (include/wavelets.hpp:154-281,402-448 of the reference).  See cqbem_wavelet.hpp.
#include "cqbem_wavelet.hpp"

#include <algorithm>
#include <stdexcept>
#include <string>

namespace cqbem {

namespace {

constexpr double kHaar = 0.70710678118654752440;  // M_SQRT1_2: h_2 = (c, c), g_2 = (c, -c)

void check(hipError_t e, const char *what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// One level of the nonstandard form.  s is (2 sz) x (2 sz): src with leading dimension src_ld, entries
// at rows or columns >= n_valid read as zero (the padding of the first level).  Writes this level's
// candidates (d rows: alpha | beta, 2 sz each; then s rows: gamma, sz each) at `cand` and the next
// coarse matrix (sz x sz) at `next`.  The four products of each sum are rounded separately and added in
// the order k, m = (0,0), (0,1), (1,0), (1,1) (wavelets.hpp:189-201), with g_k g_m etc. = +-c^2.
__global__ void __launch_bounds__(256) k_haar_level(const double *__restrict__ src, int64_t src_ld, int64_t n_valid,
                                                    int64_t sz, double *__restrict__ cand, double *__restrict__ next) {
#pragma clang fp contract(off)
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= sz * sz) return;
    const int64_t i = t / sz, l = t - i * sz;
    double s[2][2];
#pragma unroll
    for (int k = 0; k < 2; k++)
#pragma unroll
        for (int m = 0; m < 2; m++) {
            const int64_t r = 2 * i + k, c = 2 * l + m;
            s[k][m] = (r < n_valid && c < n_valid) ? src[r * src_ld + c] : 0.0;
        }
    const double cc = kHaar * kHaar;
    const double p00 = cc * s[0][0], p01 = cc * s[0][1], p10 = cc * s[1][0], p11 = cc * s[1][1];
    const double alpha = (((0.0 + p00) - p01) - p10) + p11;
    const double beta = (((0.0 + p00) + p01) - p10) - p11;
    const double gamma = (((0.0 + p00) - p01) + p10) - p11;
    const double coarse = (((0.0 + p00) + p01) + p10) + p11;
    cand[i * 2 * sz + l] = alpha;
    cand[i * 2 * sz + sz + l] = beta;
    cand[2 * sz * sz + i * sz + l] = gamma;
    next[i * sz + l] = coarse;
}

__device__ inline uint64_t mag_key(double v) { return (uint64_t)__double_as_longlong(v) & 0x7fffffffffffffffull; }

// Histogram of the `width` bits at `shift` of |v| over the nonzero candidates whose higher bits equal prefix.
__global__ void __launch_bounds__(256) k_mag_hist(const double *__restrict__ cand, int64_t total, int shift, int width,
                                                  uint64_t prefix, unsigned int *__restrict__ hist) {
    __shared__ unsigned int h[4096];
    const int nb = 1 << width;
    for (int b = threadIdx.x; b < nb; b += 256) h[b] = 0;
    __syncthreads();
    for (int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x; j < total; j += (int64_t)gridDim.x * 256) {
        const uint64_t key = mag_key(cand[j]);
        if (key == 0 || (key >> (shift + width)) != prefix) continue;
        atomicAdd(&h[(key >> shift) & (uint64_t)(nb - 1)], 1u);
    }
    __syncthreads();
    for (int b = threadIdx.x; b < nb; b += 256)
        if (h[b]) atomicAdd(&hist[b], h[b]);
}

// Per row (one wavefront each): candidates above the threshold key T and equal to it.
__global__ void __launch_bounds__(256) k_row_count(const double *__restrict__ cand, const int64_t *__restrict__ rstart,
                                                   const int32_t *__restrict__ rlen, int64_t rows, uint64_t T,
                                                   int32_t *__restrict__ gt, int32_t *__restrict__ eq) {
    const int64_t r = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (r >= rows) return;
    const double *c = cand + rstart[r];
    const int len = rlen[r];
    int ng = 0, ne = 0;
    for (int j = lane; j < len; j += 64) {
        const uint64_t key = mag_key(c[j]);
        ng += key > T;
        ne += key == T && key != 0;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) ng += __shfl_xor(ng, o, 64), ne += __shfl_xor(ne, o, 64);
    if (lane == 0) gt[r] = ng, eq[r] = ne;
}

// Per row: the kept candidates in column order (all above T, then ties up to the row's quota in order).
__global__ void __launch_bounds__(256) k_row_fill(const double *__restrict__ cand, const int64_t *__restrict__ rstart,
                                                  const int32_t *__restrict__ rlen, const int64_t *__restrict__ rcol,
                                                  const int32_t *__restrict__ quota, const int64_t *__restrict__ rowptr,
                                                  int64_t rows, uint64_t T, double scale, int32_t *__restrict__ col,
                                                  double *__restrict__ val) {
    const int64_t r = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (r >= rows) return;
    const double *c = cand + rstart[r];
    const int len = rlen[r], q = quota[r];
    const int64_t base = rowptr[r];
    const uint64_t below = (1ull << lane) - 1;
    int written = 0, ties = 0;
    for (int j0 = 0; j0 < len; j0 += 64) {
        const int j = j0 + lane;
        const bool valid = j < len;
        const double v = valid ? c[j] : 0.0;
        const uint64_t key = mag_key(v);
        const bool is_gt = valid && key > T;
        const bool is_eq = valid && key == T && key != 0;
        const uint64_t eq_mask = __ballot(is_eq);
        const bool keep = is_gt || (is_eq && ties + __popcll(eq_mask & below) < q);
        const uint64_t keep_mask = __ballot(keep);
        if (keep) {
            const int64_t pos = base + written + __popcll(keep_mask & below);
            col[pos] = (int32_t)(rcol[r] + j);
            val[pos] = v * scale;
        }
        written += __popcll(keep_mask);
        ties += __popcll(eq_mask);
    }
}

// One lane's part of a CSR row, s += val[e] x[col[e]] for e = e0 + lane, e0 + lane + 64, ... < e1, in that order.
// Four (col, val) pairs and their gathers are in flight per lane: with one, a wavefront per row leaves the memory
// system mostly idle.  The matrix arrays are streamed once per product (non-temporal); x is gathered through L2.
__device__ inline double csr_row_part(const int32_t *__restrict__ col, const double *__restrict__ val,
                                     const double *__restrict__ x, int64_t e0, int64_t e1, int lane, double s) {
    int64_t e = e0 + lane;
    for (; e + 64 * 3 < e1; e += 256) {
        int32_t c[4];
        double v[4], xv[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            c[k] = __builtin_nontemporal_load(col + e + 64 * k);
            v[k] = __builtin_nontemporal_load(val + e + 64 * k);
        }
#pragma unroll
        for (int k = 0; k < 4; k++) xv[k] = x[c[k]];
#pragma unroll
        for (int k = 0; k < 4; k++) s = fma(v[k], xv[k], s);
    }
    for (; e < e1; e += 64) s = fma(val[e], x[col[e]], s);
    return s;
}

__global__ void __launch_bounds__(256) k_csr_matvec_sum(const CsrRef *__restrict__ refs, int m, int64_t rows,
                                                        double *__restrict__ y) {
    const int64_t r = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (r >= rows) return;
    double s = 0.0;
    for (int i = 0; i < m; i++) {
        const CsrRef R = refs[i];
        s = csr_row_part(R.col, R.val, R.x, R.rowptr[r], R.rowptr[r + 1], lane, s);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if (lane == 0) y[r] = s;
}

__global__ void __launch_bounds__(256) k_csr_matvec(const int64_t *__restrict__ rowptr, const int32_t *__restrict__ col,
                                                    const double *__restrict__ val, const double *__restrict__ x, int64_t rows,
                                                    double *__restrict__ y) {
    const int64_t r = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (r >= rows) return;
    double s = csr_row_part(col, val, x, rowptr[r], rowptr[r + 1], lane, 0.0);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if (lane == 0) y[r] = s;
}

template <typename T>
struct Dev {
    T *p = nullptr;
    explicit Dev(size_t n) { check(hipMalloc(&p, sizeof(T) * std::max<size_t>(n, 1)), "hipMalloc (wavelet)"); }
    ~Dev() {
        if (p) (void)hipFree(p);
    }
    Dev(const Dev &) = delete;
    Dev &operator=(const Dev &) = delete;
};

unsigned grid_for(int64_t items_per_block_256, int64_t n) { return (unsigned)std::max<int64_t>(1, (n + items_per_block_256 - 1) / items_per_block_256); }

}  // namespace

void DeviceCsr::release() {
    if (rowptr) (void)hipFree(rowptr);
    if (col) (void)hipFree(col);
    if (val) (void)hipFree(val);
    rowptr = nullptr, col = nullptr, val = nullptr;
    rows = nnz = n = 0;
}

int64_t wavelet_padded(int64_t n) {
    int64_t np = 1;
    while (np < n) np <<= 1;
    return np;
}

int64_t wavelet_compress_into(const double *A, int64_t n, int64_t num_elems, double scale, int64_t *rowptr, int32_t *col,
                              double *val, hipStream_t st) {
    if (n < 2) throw std::runtime_error("wavelet compression needs n >= 2");
    const int64_t np = wavelet_padded(n), rows = 2 * np - 2, total = np * np;
    // nonstandard form, level by level (coarse matrices ping-pong between two buffers)
    Dev<double> cand((size_t)total), sa((size_t)((np / 2) * (np / 2))), sb((size_t)std::max<int64_t>(1, (np / 4) * (np / 4)));
    {
        const double *src = A;
        int64_t src_ld = n, n_valid = n, base = 0;
        double *bufs[2] = {sa.p, sb.p};
        int which = 0;
        for (int64_t sz = np / 2; sz >= 1; sz /= 2) {
            double *next = sz == 1 ? cand.p + total - 1 : bufs[which];  // the last coarse value is the final coefficient
            hipLaunchKernelGGL(k_haar_level, dim3(grid_for(256, sz * sz)), dim3(256), 0, st, src, src_ld, n_valid, sz,
                               cand.p + base, next);
            check(hipGetLastError(), "k_haar_level");
            base += 3 * sz * sz;
            src = next, src_ld = sz, n_valid = sz;
            which ^= 1;
        }
    }
    // rows of the (2np - 2)^2 form: their candidate segments and first columns
    std::vector<int64_t> rstart((size_t)rows), rcol((size_t)rows);
    std::vector<int32_t> rlen((size_t)rows);
    {
        int64_t base = 0, off = 0;
        for (int64_t sz = np / 2; sz >= 1; sz /= 2) {
            for (int64_t i = 0; i < sz; i++) {
                rstart[off + i] = base + i * 2 * sz, rlen[off + i] = (int32_t)(2 * sz), rcol[off + i] = off;
                rstart[off + sz + i] = base + 2 * sz * sz + i * sz, rlen[off + sz + i] = (int32_t)sz, rcol[off + sz + i] = off;
            }
            base += 3 * sz * sz, off += 2 * sz;
        }
        rlen[rows - 1] += 1;  // the final coarse coefficient follows the last gamma, at column 2np - 3
    }
    // radix select of the num_elems-th largest |coefficient| over the 63-bit magnitude patterns
    Dev<unsigned int> hist(4096);
    const int shifts[6] = {51, 39, 27, 15, 3, 0}, widths[6] = {12, 12, 12, 12, 12, 3};
    const unsigned hgrid = (unsigned)std::min<int64_t>(2048, std::max<int64_t>(1, total / 1024));
    uint64_t prefix = 0, T = 0;
    int64_t remaining = 0;
    std::vector<unsigned int> h(4096);
    for (int p = 0; p < 6; p++) {
        const int nb = 1 << widths[p];
        check(hipMemsetAsync(hist.p, 0, sizeof(unsigned int) * nb, st), "hipMemsetAsync");
        hipLaunchKernelGGL(k_mag_hist, dim3(hgrid), dim3(256), 0, st, cand.p, total, shifts[p], widths[p], prefix, hist.p);
        check(hipGetLastError(), "k_mag_hist");
        check(hipMemcpyAsync(h.data(), hist.p, sizeof(unsigned int) * nb, hipMemcpyDeviceToHost, st), "hipMemcpyAsync");
        check(hipStreamSynchronize(st), "hipStreamSynchronize");
        if (p == 0) {
            int64_t nonzero = 0;
            for (int b = 0; b < nb; b++) nonzero += h[b];
            if (num_elems <= 0 || num_elems >= nonzero) {  // everything nonzero is kept (num_elems <= 0: wavelets.hpp:180-184)
                T = 0, remaining = 0;
                break;
            }
            remaining = num_elems;
        }
        int64_t above = 0;
        int d = nb - 1;
        for (; d > 0; d--) {
            if (above + (int64_t)h[d] >= remaining) break;
            above += h[d];
        }
        remaining -= above;
        prefix = (prefix << widths[p]) | (uint64_t)d;
        if (p == 5) T = prefix;
    }
    // per-row counts, the row pointers and the tie quota of each row (ties kept in row-major order)
    Dev<int64_t> d_rstart((size_t)rows), d_rcol((size_t)rows);
    Dev<int32_t> d_rlen((size_t)rows), d_gt((size_t)rows), d_eq((size_t)rows), d_quota((size_t)rows);
    check(hipMemcpyAsync(d_rstart.p, rstart.data(), sizeof(int64_t) * rows, hipMemcpyHostToDevice, st), "hipMemcpyAsync");
    check(hipMemcpyAsync(d_rcol.p, rcol.data(), sizeof(int64_t) * rows, hipMemcpyHostToDevice, st), "hipMemcpyAsync");
    check(hipMemcpyAsync(d_rlen.p, rlen.data(), sizeof(int32_t) * rows, hipMemcpyHostToDevice, st), "hipMemcpyAsync");
    const unsigned rgrid = grid_for(4, rows);
    hipLaunchKernelGGL(k_row_count, dim3(rgrid), dim3(256), 0, st, cand.p, d_rstart.p, d_rlen.p, rows, T, d_gt.p, d_eq.p);
    check(hipGetLastError(), "k_row_count");
    std::vector<int32_t> gt((size_t)rows), eq((size_t)rows), quota((size_t)rows);
    check(hipMemcpyAsync(gt.data(), d_gt.p, sizeof(int32_t) * rows, hipMemcpyDeviceToHost, st), "hipMemcpyAsync");
    check(hipMemcpyAsync(eq.data(), d_eq.p, sizeof(int32_t) * rows, hipMemcpyDeviceToHost, st), "hipMemcpyAsync");
    check(hipStreamSynchronize(st), "hipStreamSynchronize");
    std::vector<int64_t> rp((size_t)rows + 1, 0);
    int64_t ties_left = remaining;
    for (int64_t r = 0; r < rows; r++) {
        quota[r] = (int32_t)std::min<int64_t>(eq[r], ties_left);
        ties_left -= quota[r];
        rp[r + 1] = rp[r] + gt[r] + quota[r];
    }
    const int64_t nnz = rp[rows];
    if (nnz > (num_elems > 0 ? num_elems : total)) throw std::runtime_error("wavelet selection kept more than requested");
    check(hipMemcpyAsync(rowptr, rp.data(), sizeof(int64_t) * (rows + 1), hipMemcpyHostToDevice, st), "hipMemcpyAsync");
    check(hipMemcpyAsync(d_quota.p, quota.data(), sizeof(int32_t) * rows, hipMemcpyHostToDevice, st), "hipMemcpyAsync");
    hipLaunchKernelGGL(k_row_fill, dim3(rgrid), dim3(256), 0, st, cand.p, d_rstart.p, d_rlen.p, d_rcol.p, d_quota.p, rowptr,
                       rows, T, scale, col, val);
    check(hipGetLastError(), "k_row_fill");
    check(hipStreamSynchronize(st), "hipStreamSynchronize");
    return nnz;
}

void wavelet_compress(const double *A, int64_t n, int64_t num_elems, double scale, DeviceCsr &out, hipStream_t st) {
    out.release();
    const int64_t np = wavelet_padded(n), rows = 2 * np - 2;
    const int64_t cap = std::max<int64_t>(1, num_elems > 0 ? std::min<int64_t>(num_elems, np * np) : np * np);
    check(hipMalloc(&out.rowptr, sizeof(int64_t) * (rows + 1)), "hipMalloc rowptr");
    check(hipMalloc(&out.col, sizeof(int32_t) * cap), "hipMalloc col");
    check(hipMalloc(&out.val, sizeof(double) * cap), "hipMalloc val");
    out.n = n, out.rows = rows;
    out.nnz = wavelet_compress_into(A, n, num_elems > 0 ? std::min<int64_t>(num_elems, np * np) : 0, scale, out.rowptr, out.col,
                                    out.val, st);
}

hipError_t launch_csr_matvec(const int64_t *rowptr, const int32_t *col, const double *val, const double *x, int64_t rows,
                             double *y, hipStream_t st) {
    hipLaunchKernelGGL(k_csr_matvec, dim3(grid_for(4, rows)), dim3(256), 0, st, rowptr, col, val, x, rows, y);
    return hipGetLastError();
}

hipError_t launch_csr_matvec_sum(const CsrRef *refs_device, int m, int64_t rows, double *y, hipStream_t st) {
    hipLaunchKernelGGL(k_csr_matvec_sum, dim3(grid_for(4, rows)), dim3(256), 0, st, refs_device, m, rows, y);
    return hipGetLastError();
}

std::vector<double> wavelet_decompose_vec(const double *x, int64_t n) {
#pragma clang fp contract(off)
    const int64_t np = wavelet_padded(n);
    std::vector<double> out((size_t)(2 * np - 2), 0.0), pad((size_t)np, 0.0);
    std::copy(x, x + n, pad.begin());
    const double h[2] = {kHaar, kHaar}, g[2] = {kHaar, -kHaar};
    const double *s = pad.data();
    int64_t off = 0;
    for (int64_t sz = np / 2; sz >= 1; sz /= 2) {
        double *d = out.data() + off, *ns = out.data() + off + sz;
        for (int64_t j = 0; j < 2 * sz; j++) {
            const double sp = s[j];
            if (sp == 0.0) continue;
            const int nn = (int)(j % 2);
            ns[j / 2] += h[nn] * sp;
            d[j / 2] += g[nn] * sp;
        }
        off += 2 * sz;
        s = ns;
    }
    return out;
}

std::vector<double> wavelet_reconstruct(const double *c, int64_t n) {
#pragma clang fp contract(off)
    const int64_t np = wavelet_padded(n);
    const double h[2] = {kHaar, kHaar}, g[2] = {kHaar, -kHaar};
    std::vector<double> res((size_t)np, 0.0), s((size_t)std::max<int64_t>(1, np / 2));
    int64_t offset = -2;
    for (int64_t size = 1; size <= np / 2; size *= 2) {
        const double *d = c + 2 * np - 2 + offset;
        for (int64_t j = 0; j < size; j++) s[j] = c[2 * np - 2 + offset + size + j] + res[j];
        std::fill(res.begin(), res.begin() + 2 * size, 0.0);
        for (int64_t j = 0; j < size; j++) {
            res[2 * j] += h[0] * s[j] + g[0] * d[j];
            res[2 * j + 1] += h[1] * s[j] + g[1] * d[j];
        }
        offset -= 4 * size;
    }
    res.resize((size_t)n);
    return res;
}

}  // namespace cqbem
