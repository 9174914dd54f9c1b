// Device-side math shared by the CQBEM kernels: FP64 complex arithmetic and the
// Laplace-domain elastodynamic kernels in factored form.
//
// The reference evaluates each kernel entry (i,j) separately
// (G_collocation.cu:44-60, H_collocation.cu:203-234, B_trans_collocation.cu:41-58,
// B_angular_collocation.cu:185-204).  Here every kernel is written as a few
// complex scalar coefficients times frequency-independent geometric tensors, so
// one point/frequency evaluation costs two complex exponentials plus ~30 complex
// operations for all four matrices:
//   u_hat  = alpha rh rh^T + beta I
//   p_hat  = (1/4pi) [ h1 rh n^T + h2 (n rh^T + (rh.n) I) + h4 (rh.n) rh rh^T ]
//   b_hat  = t1 (rh n^T + n rh^T) + t2 (rh.n) I
//   b_hat_ijk = u1 (rh_i rh_k n_j + rh_j rh_k n_i) + (u1+u2)(d_jk n_i + d_ik n_j) + u3 d_ij n_k
#pragma once
#include <hip/hip_runtime.h>

namespace cqbem {

struct cd {
    double x, y;
};

__host__ __device__ inline cd mkc(double a, double b) { return cd{a, b}; }
__host__ __device__ inline cd operator+(cd a, cd b) { return cd{a.x + b.x, a.y + b.y}; }
__host__ __device__ inline cd operator-(cd a, cd b) { return cd{a.x - b.x, a.y - b.y}; }
__host__ __device__ inline cd operator-(cd a) { return cd{-a.x, -a.y}; }
__host__ __device__ inline cd operator*(cd a, cd b) { return cd{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
__host__ __device__ inline cd operator*(cd a, double s) { return cd{a.x * s, a.y * s}; }
__host__ __device__ inline cd operator*(double s, cd a) { return cd{a.x * s, a.y * s}; }
__host__ __device__ inline cd &operator+=(cd &a, cd b) {
    a.x += b.x, a.y += b.y;
    return a;
}
__host__ __device__ inline cd &operator-=(cd &a, cd b) {
    a.x -= b.x, a.y -= b.y;
    return a;
}
// a += b * s (complex times real)
__host__ __device__ inline void fma_cr(cd &a, cd b, double s) {
    a.x = fma(b.x, s, a.x);
    a.y = fma(b.y, s, a.y);
}

// exp(-z)
__device__ inline cd cexp_neg(cd z) {
    const double m = exp(-z.x);
    double sn, cs;
    sincos(z.y, &sn, &cs);
    return cd{m * cs, -m * sn};
}

// Per-frequency constants, uniform across a wavefront (loaded with scalar loads).
struct Freq {
    cd s;    // s_l
    cd sc1;  // s / c1
    cd sc2;  // s / c2
    cd is;   // 1 / s
    cd is2;  // 1 / s^2
};

// Material constants.
struct Mat {
    double c1, c2, rho;
    double ic1, ic2;    // 1/c1, 1/c2
    double ic1s, ic2s;  // 1/c1^2, 1/c2^2
    double kap;         // c2^2 / c1^2
    double six_c2s;     // 6 c2^2
    double g_scale;     // 1 / (4 pi rho)
};

// Shared exponentials of one (point pair, frequency): everything the four
// kernels need that depends on the frequency.
struct Expo {
    cd e1, e2;  // exp(-r s/c1), exp(-r s/c2)
    cd dA;      // A1 - A2, A_k = (s r/c_k + 1) exp(-r s/c_k) / s^2
    cd A2;
};

__device__ inline Expo expo(double r, const Freq &f) {
    Expo E;
    const cd z1 = f.sc1 * r, z2 = f.sc2 * r;
    E.e1 = cexp_neg(z1);
    E.e2 = cexp_neg(z2);
    const cd A1 = (z1 + cd{1.0, 0.0}) * E.e1 * f.is2;
    E.A2 = (z2 + cd{1.0, 0.0}) * E.e2 * f.is2;
    E.dA = A1 - E.A2;
    return E;
}

// u_hat = alpha rh rh^T + beta I  (times r if `times_r`: r*u_hat for polar rules)
__device__ inline void g_coef(const Expo &E, double ir, double ir2, const Mat &m, cd &alpha, cd &beta) {
    const cd t = E.dA * ir2;
    alpha = (3.0 * t + E.e1 * m.ic1s - E.e2 * m.ic2s) * m.g_scale;
    beta = (E.e2 * m.ic2s - t) * m.g_scale;
}

// p_hat coefficients without the 1/(4 pi) factor.
__device__ inline void h_coef(const Expo &E, double ir, double ir2, const Freq &f, const Mat &m, cd &h1, cd &h2, cd &h4) {
    const cd irs = f.is * ir;  // 1/(r s)
    const cd D = (((E.e1 * m.ic1) - (E.e2 * m.ic2)) * irs + (E.e1 - E.e2) * (irs * irs)) * (m.six_c2s * ir2);
    const cd T1 = E.e1 * ir2, T2 = E.e2 * ir2;
    const cd T3 = E.e1 * f.sc1 * ir, T4 = E.e2 * f.sc2 * ir;
    const double k = m.kap;
    h1 = D + T1 * (4.0 * k - 1.0) - T2 * 2.0 - T3 * (1.0 - 2.0 * k);
    h2 = D + T1 * (2.0 * k) - T2 * 3.0 - T4;
    h4 = D * (-5.0) - T1 * (12.0 * k) + T2 * 12.0 - T3 * (2.0 * k) + T4 * 2.0;
}

constexpr double kInv4Pi = 0.079577471545947667884;  // 1/(4 pi)
constexpr double kInv8Pi = 0.039788735772973833942;  // 1/(8 pi)

// b_hat (translational) and b_hat_ijk (angular) coefficients.
__device__ inline void b_coef(const Expo &E, double ir, double ir2, const Freq &f, const Mat &m, cd &t1, cd &t2, cd &u1,
                              cd &u12, cd &u3) {
    t1 = E.dA * (ir2 * kInv8Pi);
    t2 = (E.A2 - f.is2) * (ir2 * kInv4Pi);
    u1 = E.dA * (ir * kInv8Pi);
    const cd u2 = ((E.e2 * m.ic2) - (E.e1 * m.ic1)) * f.is * kInv8Pi;
    u12 = u1 + u2;
    u3 = E.e2 * f.is * (m.ic2 * kInv4Pi);
}

// symmetric 3x3 packed index: 00 11 22 01 02 12
__host__ __device__ inline int sidx(int a, int b) {
    if (a == b) return a;
    const int lo = a < b ? a : b, hi = a < b ? b : a;
    return lo == 0 ? (hi == 1 ? 3 : 4) : 5;
}

}  // namespace cqbem
