// Device-side math shared by the CQBEM kernels: FP64 complex arithmetic and the
// Laplace-domain elastodynamic kernels in factored form.
//
// The reference evaluates each kernel entry (i,j) separately
// (G_collocation.cu:44-60, H_collocation.cu:45-76, B_trans_collocation.cu:41-58,
// B_angular_collocation.cu:44-63).  Here every kernel is written as a few
// complex scalar coefficients times frequency-independent geometric tensors, so
// one point/frequency evaluation costs two complex exponentials plus ~30 complex
// operations for all four matrices:
//   u_hat  = alpha rh rh^T + beta I
//   p_hat  = (1/4pi) [ h1 rh n^T + h2 (n rh^T + (rh.n) I) + h4 (rh.n) rh rh^T ]
//   b_hat  = t1 (rh n^T + n rh^T) + t2 (rh.n) I
//   b_hat_ijk = u1 (rh_i rh_k n_j + rh_j rh_k n_i) + (u1+u2)(d_jk n_i + d_ik n_j) + u3 d_ij n_k
#pragma once
#include <hip/hip_runtime.h>

#include "cqbem_math_tables.hpp"

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

// ---------------------------------------------------------------------------
// Table-driven variants (P. T. P. Tang, "Table-driven implementation of the
// exponential function in IEEE floating-point arithmetic", ACM TOMS 15, 1989;
// the same scheme for sin/cos): the reduction goes to |r| <= ln2/128 and
// |r| <= pi/64 against 64-entry tables of 2^(j/64) and sin, cos(j pi/32) held as
// double-double (hi, lo) pairs, so short series suffice and the table value is
// added last: e^{-a}, sin and cos stay within about 0.52 ulp (of 1 for sin, cos).
// The kernels that use them copy the tables to LDS first (load_math_tables).
// ---------------------------------------------------------------------------
struct MathTables {
    double2 e[64];  // 2^(j/64): hi, lo
    double4 t[64];  // sin(j pi/32) hi, lo, cos(j pi/32) hi, lo
};

static __constant__ double kExpTable[64][2] = CQBEM_EXP_TABLE;
static __constant__ double kTrigTable[64][4] = CQBEM_TRIG_TABLE;

// Copies the tables (constant memory) into `s`; every thread of the block calls it, a barrier must follow.
__device__ inline void load_math_tables(MathTables *s) {
    const auto &kExp = kExpTable;
    const auto &kTrig = kTrigTable;
    for (int k = threadIdx.x; k < 64; k += blockDim.x) {
        s->e[k] = make_double2(kExp[k][0], kExp[k][1]);
        s->t[k] = make_double4(kTrig[k][0], kTrig[k][1], kTrig[k][2], kTrig[k][3]);
    }
}

// e^{-a}, a >= -700, as exp_neg.  k = rint(t 64/ln2), t = -a clamped to [-740, 700];
// r = t - k ln2/64 (the k hi product and the first subtraction exact); e^r - 1 through r^6
// (next term < 3e-20); 2^(k/64) = 2^(k >> 6) T[k & 63], and T_hi + (T_hi p + T_lo) last.
__host__ __device__ inline double exp_neg_tab(double a, const double2 *T) {
#pragma clang fp contract(off)
    const double t = fmin(fmax(-a, -740.0), 700.0);
    const double k = rint(t * CQBEM_64_OVER_LN2);
    const double r1 = fma(-k, CQBEM_LN2_OVER_64_HI, t);
    const double r = fma(-k, CQBEM_LN2_OVER_64_LO, r1);
    double q = fma(r, 1.0 / 720.0, 1.0 / 120.0);
    q = fma(q, r, 1.0 / 24.0);
    q = fma(q, r, 1.0 / 6.0);
    q = fma(q, r, 0.5);
    const double p = fma(r * r, q, r);
    const int ki = (int)k;
    const double2 e = T[ki & 63];
    return ldexp(e.x + fma(e.x, p, e.y), ki >> 6);
}

// sin y, cos y for |y| < kTrigArgMax, as cq_sincos.  k = rint(32 y/pi); r + rr = y - k pi/32 to
// ~1e-28 (three-part pi/32, the first two products exact, TwoSum on the subtraction; |rr| reaches
// 1e-12 at |k| ~ 1e7, so rr enters to first order, sin(r + rr) = sin r + rr cos r and
// cos(r + rr) - 1 = cos r - 1 - rr sin r, the rr^2 terms below 1e-24); sin r through r^9 and
// cos r - 1 through r^8 (next terms < 1e-19); then sin(a + r) = S + (S (cos r - 1) + C sin r) and
// cos(a + r) = C + (C (cos r - 1) - S sin r) with a = (k mod 64) pi/32 and the table's lo parts
// added inside the brackets.
__host__ __device__ inline void sincos_tab(double y, const double4 *T, double *s, double *c) {
#pragma clang fp contract(off)
    const double k = rint(y * CQBEM_32_OVER_PI);
    const double r1 = fma(-k, CQBEM_PI_OVER_32_A, y);
    const double tb = k * CQBEM_PI_OVER_32_B;
    const double r = r1 - tb;
    const double bb = r - r1;
    double rr = (r1 - (r - bb)) - (tb + bb);
    rr = fma(-k, CQBEM_PI_OVER_32_C, rr);
    const double x = r * r;
    double ps = fma(x, 1.0 / 362880.0, -1.0 / 5040.0);
    ps = fma(ps, x, 1.0 / 120.0);
    ps = fma(ps, x, -1.0 / 6.0);
    const double s0 = (r * x) * ps;  // sin r - r
    double pc = fma(x, 1.0 / 40320.0, -1.0 / 720.0);
    pc = fma(pc, x, 1.0 / 24.0);
    pc = fma(pc, x, -0.5);
    const double c0 = x * pc;                        // cos r - 1
    const double sr = r + fma(rr, 1.0 + c0, s0);     // sin(r + rr)
    const double cm1 = fma(-rr, r + s0, c0);         // cos(r + rr) - 1
    const double4 e = T[(int)k & 63];
    *s = e.x + fma(e.z, sr, fma(e.x, cm1, e.y));
    *c = e.z + fma(-e.x, sr, fma(e.z, cm1, e.w));
}

// exp(-z), Re z >= 0, through the tables
__host__ __device__ inline cd cexp_neg_tab(cd z, const MathTables *M) {
    const double m = exp_neg_tab(z.x, M->e);
    double sn, cs;
    sincos_tab(z.y, M->t, &sn, &cs);
    return cd{m * cs, -m * sn};
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

// expo with the table-driven exponentials (tables in LDS)
__device__ inline Expo expo(double r, const Freq &f, const MathTables *M) {
    Expo E;
    const cd z1 = f.sc1 * r, z2 = f.sc2 * r;
    E.e1 = cexp_neg_tab(z1, M);
    E.e2 = cexp_neg_tab(z2, M);
    const cd A1 = (z1 + cd{1.0, 0.0}) * E.e1 * f.is2;
    E.A2 = (z2 + cd{1.0, 0.0}) * E.e2 * f.is2;
    E.dA = A1 - E.A2;
    return E;
}

// The exponentials G and H need: e1, e2 and dA, the last formed as ((z1 + 1) e1 - (z2 + 1) e2) / s^2 with one product
// by 1/s^2 (A2 is not formed: it is left zero).
__device__ inline Expo expo_gh(double r, const Freq &f, const MathTables *M) {
    Expo E;
    const cd z1 = f.sc1 * r, z2 = f.sc2 * r;
    E.e1 = cexp_neg_tab(z1, M);
    E.e2 = cexp_neg_tab(z2, M);
    E.dA = ((z1 + cd{1.0, 0.0}) * E.e1 - (z2 + cd{1.0, 0.0}) * E.e2) * f.is2;
    E.A2 = cd{0.0, 0.0};
    return E;
}

// u_hat = alpha rh rh^T + beta I  (times r if `times_r`: r*u_hat for polar rules).  SCALED = false leaves out the
// factor g_scale = 1/(4 pi rho), for callers that apply it to their sums.
template <bool SCALED = true>
__device__ inline void g_coef(const Expo &E, double ir, double ir2, const Mat &m, cd &alpha, cd &beta) {
    const cd t = E.dA * ir2;
    alpha = 3.0 * t + E.e1 * m.ic1s - E.e2 * m.ic2s;
    beta = E.e2 * m.ic2s - t;
    if (SCALED) alpha = alpha * m.g_scale, beta = beta * m.g_scale;
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

// b_coef times 8 pi: t1, u1 and u12 without their factor 1/(8 pi), t2 and u3 without 1/(4 pi) but doubled (exact), so
// that every sum of them takes the one factor 1/(8 pi) at its end.
__device__ inline void b_coef_8pi(const Expo &E, double ir, double ir2, const Freq &f, const Mat &m, cd &t1, cd &t2,
                                  cd &u1, cd &u12, cd &u3) {
    t1 = E.dA * ir2;
    t2 = (E.A2 - f.is2) * (2.0 * ir2);
    u1 = E.dA * ir;
    const cd u2 = ((E.e2 * m.ic2) - (E.e1 * m.ic1)) * f.is;
    u12 = u1 + u2;
    u3 = E.e2 * f.is * (2.0 * m.ic2);
}

// symmetric 3x3 packed index: 00 11 22 01 02 12
__host__ __device__ inline int sidx(int a, int b) {
    if (a == b) return a;
    const int lo = a < b ? a : b, hi = a < b ? b : a;
    return lo == 0 ? (hi == 1 ? 3 : 4) : 5;
}

}  // namespace cqbem
