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

// ---------------------------------------------------------------------------
// FP64 elementary functions for the kernel arguments.  The CQ exponent
// z = s r / c has Re z >= 0 (the BDF symbols are A-stable, so Re s_l > 0) and
// a moderate imaginary part, which allows an exact Cody-Waite reduction
// (W. J. Cody, W. Waite, "Software Manual for the Elementary Functions", 1980)
// followed by truncated Taylor series (the first omitted terms are below
// 1e-17 relative) whose leading terms are summed with their rounding errors
// carried (Fast2Sum), so e^{-a} stays within 0.65 ulp and sin, cos within
// 0.4 ulp of 1, at about half the instruction count of the library calls.
// Contraction is off inside them: every FMA is written out, so the error
// analysis holds as written.
// ---------------------------------------------------------------------------
constexpr double kLog2e = 1.4426950408889634;         // 1/ln 2
constexpr double kLn2Hi = 0x1.62e42fefa3800p-1;       // ln 2 to 42 bits: k kLn2Hi exact for |k| < 2^11
constexpr double kLn2Lo = 5.497923018708371e-14;      // ln 2 - kLn2Hi
constexpr double kTwoOverPi = 0.6366197723675814;     // 2/pi
constexpr double kPio2A = 0x1.921fb54400000p+0;       // pi/2 to 33 bits: k kPio2A exact for |k| < 2^20
constexpr double kPio2B = 0x1.0b4611a600000p-34;      // next 33 bits of pi/2
constexpr double kPio2C = 2.0222662487959506e-21;     // pi/2 - kPio2A - kPio2B
// |Im z| bound of the exact reduction (|k| < 2^20).  The host rejects jobs whose largest argument
// max|s_l| diam / min(c1, c2) reaches it.
constexpr double kTrigArgMax = 1.0e6;

// e^{-a}, a >= -700.  k = rint(-a / ln 2); f + c = -a - k ln 2 with |f| <= ln2/2 (the k kLn2Hi part
// is exact by Sterbenz); e^(f+c) = (1 + f) + (f^2 p(f) + c (1 + f)) with p the Taylor series of
// (e^f - 1 - f) / f^2 through f^11 (next term < 5e-18); 1 + f is split exactly.  Scaled by 2^k;
// -a is clamped at -740, where the result has underflowed to 0.
__device__ inline double exp_neg(double a) {
#pragma clang fp contract(off)
    const double t = fmin(fmax(-a, -740.0), 700.0);
    const double k = rint(t * kLog2e);
    const double fh = fma(-k, kLn2Hi, t);
    const double fl = -k * kLn2Lo;
    const double f = fh + fl;
    const double c = (fh - f) + fl;
    double p = 1.0 / 6227020800.0;     // 1/13!
    p = fma(p, f, 1.0 / 479001600.0);  // 1/12!
    p = fma(p, f, 1.0 / 39916800.0);   // 1/11!
    p = fma(p, f, 1.0 / 3628800.0);
    p = fma(p, f, 1.0 / 362880.0);
    p = fma(p, f, 1.0 / 40320.0);
    p = fma(p, f, 1.0 / 5040.0);
    p = fma(p, f, 1.0 / 720.0);
    p = fma(p, f, 1.0 / 120.0);
    p = fma(p, f, 1.0 / 24.0);
    p = fma(p, f, 1.0 / 6.0);
    p = fma(p, f, 0.5);
    const double tail = fma(f * f, p, fma(c, f, c));
    const double s1 = 1.0 + f, e1 = (1.0 - s1) + f;
    return ldexp(s1 + (e1 + tail), (int)k);
}

// sin y, cos y for |y| < kTrigArgMax.  k = rint(2y/pi); r + rr = y - k pi/2 to ~1e-30 (three-part
// pi/2, the first two products exact, the subtraction split by TwoSum); sin r through r^17 and
// cos r through r^18 (next terms < 1e-19), first-order in rr; 1 - r^2/2 split exactly.  The
// quadrant k mod 4 then selects and signs them.
__device__ inline void cq_sincos(double y, double *s, double *c) {
#pragma clang fp contract(off)
    const double k = rint(y * kTwoOverPi);
    const double r1 = fma(-k, kPio2A, y);
    const double t = k * kPio2B;
    const double r = r1 - t;
    const double bb = r - r1;
    double rr = (r1 - (r - bb)) - (t + bb);
    rr = fma(-k, kPio2C, rr);
    const double x = r * r;
    double ps = 1.0 / 355687428096000.0;      // 1/17!
    ps = fma(ps, x, -1.0 / 1307674368000.0);  // -1/15!
    ps = fma(ps, x, 1.0 / 6227020800.0);
    ps = fma(ps, x, -1.0 / 39916800.0);
    ps = fma(ps, x, 1.0 / 362880.0);
    ps = fma(ps, x, -1.0 / 5040.0);
    ps = fma(ps, x, 1.0 / 120.0);
    ps = fma(ps, x, -1.0 / 6.0);
    double pc = -1.0 / 6402373705728000.0;    // -1/18!
    pc = fma(pc, x, 1.0 / 20922789888000.0);  // 1/16!
    pc = fma(pc, x, -1.0 / 87178291200.0);
    pc = fma(pc, x, 1.0 / 479001600.0);
    pc = fma(pc, x, -1.0 / 3628800.0);
    pc = fma(pc, x, 1.0 / 40320.0);
    pc = fma(pc, x, -1.0 / 720.0);
    pc = fma(pc, x, 1.0 / 24.0);
    const double sr = r + fma(r * x, ps, rr * fma(-0.5, x, 1.0));
    const double xl = fma(r, r, -x);
    const double hz = 0.5 * x;
    const double w = 1.0 - hz, ew = (1.0 - w) - hz;
    const double cr = w + (ew + fma(x * x, pc, -fma(0.5, xl, rr * r)));
    const int q = (int)k;
    const double so = (q & 1) ? cr : sr, co = (q & 1) ? sr : cr;
    *s = (q & 2) ? -so : so;
    *c = ((q + 1) & 2) ? -co : co;
}

// 1/sqrt(d), d > 0: hardware estimate refined by two Newton steps y += y (1 - d y^2) / 2.
__device__ inline double rsqrt_nr(double d) {
    double y = __builtin_amdgcn_rsq(d);
    double e = fma(-d * y, y, 1.0);
    y = fma(0.5 * y, e, y);
    e = fma(-d * y, y, 1.0);
    return fma(0.5 * y, e, y);
}

// exp(-z), Re z >= 0
__device__ inline cd cexp_neg(cd z) {
    const double m = exp_neg(z.x);
    double sn, cs;
    cq_sincos(z.y, &sn, &cs);
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
