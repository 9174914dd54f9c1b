// CQBEM precompute oracle -- TEST INFRASTRUCTURE ONLY.
//
// A CPU restatement of the reference's convolution-quadrature BEM precompute
// (rsugimoto/SurfaceOnlyDeformables): Laplace-domain G/H/B_trans/B_angular
// assembly in collocation and Galerkin form, the half-spectrum inverse CQ
// transform, the C (jump) matrices and the B_euler derivation.  It is used by
// tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the
// checker and the CPU baseline, never by the product path.
//
// Precision policies (template parameter P):
//   PolFP64      every scalar is double                         ("fp64")
//   PolRefMixed  the reference's own typedefs: float far-field collocation
//                G/B_trans/B_angular (G_collocation.cu:15,
//                B_trans_collocation.cu:14, B_angular_collocation.cu:14),
//                double collocation H (H_collocation.cu:15), float Galerkin
//                non-singular (elastodynamics_cqbem_galerkin_kernels.cuh:15),
//                double singular parts                          ("ref-mixed")
//   PolLD        every scalar is long double (x87, 64-bit mantissa); used only
//                to measure the rounding noise of the FP64 modes  ("ld")
//
// Three deliberate, documented differences from the reference's operation
// sequence (same mathematics, less rounding noise; the first two are needed because
// the R^-n factor of the inverse transform amplifies per-frequency rounding):
//   * the transform twiddles zeta^(-n l) are evaluated as cis(-2 pi ((n l) mod L)/L)
//     instead of std::pow(exp(2 pi i/L), -n l) (G.cu:172-174);
//   * zeta^l in the CQ frequencies is evaluated as cis(2 pi l/L) instead of
//     std::pow(exp(2 pi i/L), l) (G.cu:65);
//   * the polar rule's angle beta between the foot-of-perpendicular direction and the
//     first edge is atan2(|w x u1|, w . u1) instead of acos(clamp(w . u1))
//     (integrator_collocation.hpp:36): near a right angle at the second vertex acos
//     loses half the digits (1e-8 in beta, 1e-9 in the singular G rows of a cube).
// Parity pinning status of every function is listed at the end of this file.

#include <algorithm>
#include <array>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <map>
#include <utility>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "cqbem_oracle.h"

namespace {

constexpr long double kPiL = 3.141592653589793238462643383279502884L;

template <class T> using cx = std::complex<T>;

template <class T> inline T pi_t() { return (T)kPiL; }
template <class T> inline T inv4pi() { return (T)(1.0L / (4.0L * kPiL)); }

// ---------------------------------------------------------------- vectors --
template <class T> struct v3 {
    T a[3];
    T &operator[](int i) { return a[i]; }
    const T &operator[](int i) const { return a[i]; }
};
template <class T> inline v3<T> mk(T x, T y, T z) { return v3<T>{{x, y, z}}; }
template <class T, class U> inline v3<T> cast3(const U *p) { return mk<T>((T)p[0], (T)p[1], (T)p[2]); }
template <class T> inline v3<T> operator+(const v3<T> &u, const v3<T> &v) { return mk(u[0] + v[0], u[1] + v[1], u[2] + v[2]); }
template <class T> inline v3<T> operator-(const v3<T> &u, const v3<T> &v) { return mk(u[0] - v[0], u[1] - v[1], u[2] - v[2]); }
template <class T> inline v3<T> operator*(const v3<T> &u, T s) { return mk(u[0] * s, u[1] * s, u[2] * s); }
template <class T> inline v3<T> operator*(T s, const v3<T> &u) { return u * s; }
template <class T> inline v3<T> operator/(const v3<T> &u, T s) { return mk(u[0] / s, u[1] / s, u[2] / s); }
template <class T> inline T dot(const v3<T> &u, const v3<T> &v) { return u[0] * v[0] + u[1] * v[1] + u[2] * v[2]; }
template <class T> inline v3<T> cross(const v3<T> &u, const v3<T> &v) {
    return mk(u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0]);
}
template <class T> inline T norm(const v3<T> &u) { return std::sqrt(dot(u, u)); }
template <class T> inline v3<T> unit(const v3<T> &u) {
    T n = norm(u);
    return n > T(0) ? u / n : u;
}

// ------------------------------------------------------------- quadrature --
// Symmetric triangle rules of polynomial order 1..5 with 1/3/4/6/7 points
// (Cowper 1973; Strang-Fix).  Barycentric coordinates and weights summing to 1,
// as used by integrator.hpp:55-113.
struct TriRule {
    int n;
    long double c[7][3];
    long double w[7];
};

TriRule tri_rule(int order) {
    TriRule r{};
    auto put = [&](long double a, long double b, long double c, long double w) {
        r.c[r.n][0] = a, r.c[r.n][1] = b, r.c[r.n][2] = c, r.w[r.n] = w, r.n++;
    };
    auto orbit3 = [&](long double a, long double w) {  // (1-2a, a, a) and rotations
        put(1 - 2 * a, a, a, w), put(a, 1 - 2 * a, a, w), put(a, a, 1 - 2 * a, w);
    };
    const long double third = 1.0L / 3.0L;
    switch (order) {
    case 1: put(third, third, third, 1.0L); break;
    case 2: put(0, 0.5L, 0.5L, third), put(0.5L, 0, 0.5L, third), put(0.5L, 0.5L, 0, third); break;
    case 3:
        put(third, third, third, -27.0L / 48.0L);
        put(11.0L / 15, 2.0L / 15, 2.0L / 15, 25.0L / 48), put(2.0L / 15, 11.0L / 15, 2.0L / 15, 25.0L / 48),
            put(2.0L / 15, 2.0L / 15, 11.0L / 15, 25.0L / 48);
        break;
    case 4:
        orbit3(0.445948490915965L, 2 * 0.111690794839005L);
        orbit3(0.091576213509771L, 2 * 0.054975871827661L);
        break;
    case 5:
        put(third, third, third, 9.0L / 40.0L);
        orbit3(0.470142064105115L, 2 * 0.0661970763942530L);
        orbit3(0.101286507323456L, 2 * 0.0629695902724135L);
        break;
    default: r.n = 0; break;
    }
    return r;
}

// Gauss-Legendre rules with 1..5 nodes on [-1,1] (integrator.hpp:115-150),
// written from the closed forms of the Legendre roots.
struct LineRule {
    int n;
    long double x[5];
    long double w[5];
};

LineRule gauss_rule(int order) {
    LineRule g{};
    switch (order) {
    case 1: g.n = 1, g.x[0] = 0, g.w[0] = 2; break;
    case 2: {
        long double a = 1.0L / std::sqrt(3.0L);
        g.n = 2, g.x[0] = -a, g.x[1] = a, g.w[0] = g.w[1] = 1;
        break;
    }
    case 3: {
        long double a = std::sqrt(0.6L);
        g.n = 3, g.x[0] = -a, g.x[1] = 0, g.x[2] = a;
        g.w[0] = g.w[2] = 5.0L / 9.0L, g.w[1] = 8.0L / 9.0L;
        break;
    }
    case 4: {
        long double t = 2.0L / 7.0L * std::sqrt(1.2L);
        long double a = std::sqrt(3.0L / 7.0L - t), b = std::sqrt(3.0L / 7.0L + t);
        long double wa = (18.0L + std::sqrt(30.0L)) / 36.0L, wb = (18.0L - std::sqrt(30.0L)) / 36.0L;
        g.n = 4, g.x[0] = -b, g.x[1] = -a, g.x[2] = a, g.x[3] = b;
        g.w[0] = g.w[3] = wb, g.w[1] = g.w[2] = wa;
        break;
    }
    case 5: {
        long double t = 2.0L * std::sqrt(10.0L / 7.0L);
        long double a = std::sqrt(5.0L - t) / 3.0L, b = std::sqrt(5.0L + t) / 3.0L;
        long double wa = (322.0L + 13.0L * std::sqrt(70.0L)) / 900.0L, wb = (322.0L - 13.0L * std::sqrt(70.0L)) / 900.0L;
        g.n = 5, g.x[0] = -b, g.x[1] = -a, g.x[2] = 0, g.x[3] = a, g.x[4] = b;
        g.w[0] = g.w[4] = wb, g.w[1] = g.w[3] = wa, g.w[2] = 128.0L / 225.0L;
        break;
    }
    default: g.n = 0; break;
    }
    return g;
}

// Composite rule on [a,b]: `sub` equal pieces, each with the Gauss rule
// (integrate_subdivision_1d + gaussian_quadrature_1d, integrator.hpp:15-22,116-150).
template <class T> struct Line1D {
    std::vector<T> x, w;
};
template <class T> Line1D<T> composite_line(T a, T b, int order, int sub) {
    Line1D<T> out;
    LineRule g = gauss_rule(order);
    for (int k = 0; k < sub; k++) {
        T lo = a + (b - a) * T(k) / T(sub), hi = a + (b - a) * T(k + 1) / T(sub);
        for (int q = 0; q < g.n; q++) {
            out.x.push_back((hi - lo) * (T)g.x[q] / T(2) + (lo + hi) / T(2));
            out.w.push_back((T)g.w[q] * ((hi - lo) / T(2)));
        }
    }
    return out;
}

// Uniform m x m subdivision of a triangle into m^2 congruent sub-triangles
// (integrate_subdivision_triangle, integrator.hpp:24-51): m(m+1)/2 "upright"
// and m(m-1)/2 "inverted" children.  The child corners are returned in the
// parent's barycentric coordinates.
struct SubTri {
    long double b[3][3];  // corner k, barycentric coordinate j
};
std::vector<SubTri> subdivide(int m) {
    std::vector<SubTri> out;
    auto P = [&](int a, int c, long double *dst) {  // point x0 + (a v1 + c v2)/m
        dst[1] = (long double)a / m, dst[2] = (long double)c / m, dst[0] = 1 - dst[1] - dst[2];
    };
    for (int i = 0; i < m; i++)
        for (int j = 0; i + j < m; j++) {
            SubTri t;
            P(i, j, t.b[0]), P(i + 1, j, t.b[1]), P(i, j + 1, t.b[2]);
            out.push_back(t);
            if (i + j < m - 1) {
                SubTri u;
                P(i + 1, j + 1, u.b[0]), P(i + 1, j, u.b[1]), P(i, j + 1, u.b[2]);
                out.push_back(u);
            }
        }
    return out;
}

// Composite triangle rule on the reference triangle: points as barycentric
// coordinates of the parent, weights summing to 1 (so a physical integral is
// area * sum w f).
struct RefPts {
    std::vector<std::array<long double, 3>> b;
    std::vector<long double> w;
};
RefPts composite_tri(int order, int sub) {
    RefPts out;
    TriRule r = tri_rule(order);
    for (const SubTri &t : subdivide(sub))
        for (int q = 0; q < r.n; q++) {
            std::array<long double, 3> p{};
            for (int k = 0; k < 3; k++)
                for (int j = 0; j < 3; j++) p[j] += r.c[q][k] * t.b[k][j];
            out.b.push_back(p);
            out.w.push_back(r.w[q] / ((long double)sub * sub));
        }
    return out;
}

// ---------------------------------------------------- fundamental solutions --
// Laplace-domain elastodynamic kernels.  rh = unit vector from the
// collocation/test point x to the source point y, r = |y-x|, n = source normal.
template <class T> struct KParam {
    T c1, c2, rho;
    cx<T> s;
};

// Frequency- and distance-dependent factors shared by every entry of one point pair.  The reference's
// integrands evaluate them inside each entry (the same expressions on the same operands); computing them once
// per point pair gives bit-identical entries.
template <class T> struct KExp {
    cx<T> e1, e2, A1, A2;   // e_k = exp(-r s / c_k), A_k = (s r / c_k + 1) / s^2 e_k
    cx<T> q, e1r2, e2r2, e1s, e2s;  // traction-kernel factors (H_collocation.cu:45-76)
    cx<T> A2mA1, is2mA2, ec;        // B kernels: A2 - A1, 1/s^2 - A2, e2/c2 - e1/c1
};
template <class T> KExp<T> k_exp(T r, const KParam<T> &k) {
    KExp<T> x;
    const cx<T> s = k.s;
    const T c1 = k.c1, c2 = k.c2;
    x.e1 = std::exp(-r * s / c1), x.e2 = std::exp(-r * s / c2);
    x.A1 = ((s * r / c1 + T(1)) / (s * s)) * x.e1;
    x.A2 = ((s * r / c2 + T(1)) / (s * s)) * x.e2;
    const cx<T> rs = r * s;
    x.q = (x.e1 / rs) * (T(1) / c1 + T(1) / rs) - (x.e2 / rs) * (T(1) / c2 + T(1) / rs);
    x.e1r2 = x.e1 / (r * r), x.e2r2 = x.e2 / (r * r);
    x.e1s = x.e1 * s / (r * c1), x.e2s = x.e2 * s / (r * c2);
    x.A2mA1 = x.A2 - x.A1, x.is2mA2 = T(1) / (s * s) - x.A2, x.ec = x.e2 / c2 - x.e1 / c1;
    return x;
}

// r * u_hat_ij (G_collocation.cu:44-60)
template <class T> cx<T> k_u_r(const v3<T> &rh, T r, int i, int j, const KParam<T> &k, const KExp<T> &x) {
    const T d = (i == j) ? T(1) : T(0);
    return (inv4pi<T>() / k.rho) *
           (((T(3) * rh[i] * rh[j] - d) / (r * r)) * (x.A1 - x.A2) +
            (rh[i] * rh[j]) * (x.e1 / (k.c1 * k.c1) - x.e2 / (k.c2 * k.c2)) + (d / (k.c2 * k.c2)) * x.e2);
}
template <class T> cx<T> k_u_r(const v3<T> &rh, T r, int i, int j, const KParam<T> &k) {
    return k_u_r(rh, r, i, j, k, k_exp(r, k));
}

// p_hat_ij: Laplace-domain traction kernel (H_collocation.cu:45-76)
template <class T> cx<T> k_p(const v3<T> &rh, T r, const v3<T> &n, int i, int j, const KParam<T> &k, const KExp<T> &x) {
    const T d = (i == j) ? T(1) : T(0);
    const T c1 = k.c1, c2 = k.c2, kap = c2 * c2 / (c1 * c1);
    const T rn = dot(rh, n);
    const cx<T> t1 = ((T(6) * c2 * c2 / (r * r)) * (rh[i] * n[j] + rh[j] * n[i] + (d - T(5) * rh[i] * rh[j]) * rn)) * x.q;
    const cx<T> t2 = x.e1r2 * (T(2) * kap * (T(2) * rh[i] * n[j] + rh[j] * n[i] - (T(6) * rh[i] * rh[j] - d) * rn) - rh[i] * n[j]);
    const cx<T> t3 = x.e2r2 * (T(12) * rh[i] * rh[j] * rn - T(2) * rh[i] * n[j] - T(3) * rh[j] * n[i] - T(3) * d * rn);
    const cx<T> t4 = x.e1s * (rh[i] * n[j] + T(2) * kap * (rh[i] * rh[j] * rn - rh[i] * n[j]));
    const cx<T> t5 = x.e2s * (T(2) * rh[i] * rh[j] * rn - d * rn - rh[j] * n[i]);
    return inv4pi<T>() * (t1 + t2 + t3 - t4 + t5);
}
template <class T> cx<T> k_p(const v3<T> &rh, T r, const v3<T> &n, int i, int j, const KParam<T> &k) {
    return k_p(rh, r, n, i, j, k, k_exp(r, k));
}

// Static, strongly singular part of p_hat removed by the Guiggiani rule
// (H_collocation.cu:78-83).
template <class T> T k_p_sing(const v3<T> &rh, T r, const v3<T> &n, int i, int j, const KParam<T> &k) {
    return inv4pi<T>() * (k.c2 * k.c2 / (k.c1 * k.c1)) * (rh[i] * n[j] - rh[j] * n[i]) / (r * r);
}

// r * b_hat_ij, translational fictitious-force kernel (B_trans_collocation.cu:41-58)
template <class T> cx<T> k_bt_r(const v3<T> &rh, T r, const v3<T> &n, int i, int j, const KParam<T> &k, const KExp<T> &x) {
    const T d = (i == j) ? T(1) : T(0);
    return -inv4pi<T>() * (((rh[i] * n[j] + rh[j] * n[i]) / (T(2) * r)) * x.A2mA1 + (d * dot(rh, n) / r) * x.is2mA2);
}
template <class T> cx<T> k_bt_r(const v3<T> &rh, T r, const v3<T> &n, int i, int j, const KParam<T> &k) {
    return k_bt_r(rh, r, n, i, j, k, k_exp(r, k));
}

// r * b_hat_ijk, angular fictitious-force kernel (B_angular_collocation.cu:44-63)
template <class T>
cx<T> k_ba_r(const v3<T> &rh, T r, const v3<T> &n, int i, int j, int kk, const KParam<T> &k, const KExp<T> &x) {
    auto dl = [](int a, int b) { return a == b ? T(1) : T(0); };
    const cx<T> s = k.s;
    const T X = rh[i] * rh[kk] * n[j] + rh[j] * rh[kk] * n[i] + dl(j, kk) * n[i] + dl(i, kk) * n[j];
    const T Y = dl(j, kk) * n[i] + dl(i, kk) * n[j];
    return -inv4pi<T>() * ((X / T(2)) * x.A2mA1 - (r * Y) / (T(2) * s) * x.ec - (r * dl(i, j) * n[kk]) / (k.c2 * s) * x.e2);
}
template <class T> cx<T> k_ba_r(const v3<T> &rh, T r, const v3<T> &n, int i, int j, int kk, const KParam<T> &k) {
    return k_ba_r(rh, r, n, i, j, kk, k, k_exp(r, k));
}

// ------------------------------------------------------------------ mesh --
struct Mesh {
    int nv, nf;
    const double *V;  // nv x 3
    const int *F;     // nf x 3
    std::vector<double> N;  // nf x 3 unit normals
    const int *vmap;  // nullable vertex_map_inverse
};

void face_normals(const double *V, const int *F, int nf, double *N) {
    for (int f = 0; f < nf; f++) {
        v3<double> a = cast3<double>(V + 3 * F[3 * f]), b = cast3<double>(V + 3 * F[3 * f + 1]),
                   c = cast3<double>(V + 3 * F[3 * f + 2]);
        v3<double> n = unit(cross(b - a, c - a));
        for (int k = 0; k < 3; k++) N[3 * f + k] = n[k];
    }
}

inline bool same_vertex(const Mesh &m, int a, int b) {
    return a == b || (m.vmap != nullptr && m.vmap[a] == m.vmap[b]);
}

// ------------------------------------------------------ precision policies --
template <class CF, class CH, class GN, class SG, class AC> struct Policy {
    using ColFar = CF;   // collocation far field: G, B_trans, B_angular
    using ColHFar = CH;  // collocation far field: H
    using GalNS = GN;    // Galerkin non-singular: all four
    using Sing = SG;     // every singular / near-singular rule
    using Acc = AC;      // Laplace-domain storage and inverse transform
};
using PolFP64 = Policy<double, double, double, double, double>;
using PolRefMixed = Policy<float, double, float, double, double>;
using PolLD = Policy<long double, long double, long double, long double, long double>;

struct Which {
    bool G, H, Bt, Ba;
};

// Per-frequency Laplace-domain output for one job: row-major complex buffers.
template <class A> struct Laplace {
    int n3;                   // 3 nv
    std::vector<cx<A>> U, P;  // n3 x n3
    std::vector<cx<A>> Bt;    // n3 x 3
    std::vector<cx<A>> Ba;    // n3 x 9, index (3v+i)*9 + j*3 + k
    void alloc(int nv, const Which &w) {
        n3 = 3 * nv;
        if (w.G) U.assign((size_t)n3 * n3, cx<A>(0));
        if (w.H) P.assign((size_t)n3 * n3, cx<A>(0));
        if (w.Bt) Bt.assign((size_t)n3 * 3, cx<A>(0));
        if (w.Ba) Ba.assign((size_t)n3 * 9, cx<A>(0));
    }
    void zero() {
        std::fill(U.begin(), U.end(), cx<A>(0));
        std::fill(P.begin(), P.end(), cx<A>(0));
        std::fill(Bt.begin(), Bt.end(), cx<A>(0));
        std::fill(Ba.begin(), Ba.end(), cx<A>(0));
    }
};

struct QParams {
    int order, sub;
};

// ---------------------------------------------- collocation: far field rule --
// Integrate f over face (y1,y2,y3) with the uniform-subdivision triangle rule in
// physical coordinates; f receives the point and its clamped barycentric
// coordinates (integrate_collocation_nonsingular, integrator_collocation.hpp:6-17,
// with the projection of barycentric_coordinates.cuh:15-38 and the clamp of
// G_collocation.cu:96-97).
template <class T, class Fn>
void far_rule(const v3<T> &y1, const v3<T> &y2, const v3<T> &y3, const QParams &q, Fn &&f) {
    TriRule r = tri_rule(q.order);
    const int m = q.sub;
    const v3<T> e1 = y2 - y1, e2 = y3 - y1;
    auto corner = [&](int a, int c) { return y1 + (T(a) * e1 + T(c) * e2) / T(m); };
    auto do_tri = [&](const v3<T> &x0, const v3<T> &x1, const v3<T> &x2) {
        const T area = norm(cross(x1 - x0, x2 - x0)) / T(2);
        for (int k = 0; k < r.n; k++) {
            v3<T> y = x0 * (T)r.c[k][0] + x1 * (T)r.c[k][1] + x2 * (T)r.c[k][2];
            // projection barycentrics w.r.t. the parent face, then clamp to [0,1]
            v3<T> a = y2 - y1, b = y3 - y1, c = y - y1;
            T d00 = dot(a, a), d01 = dot(a, b), d11 = dot(b, b), d20 = dot(c, a), d21 = dot(c, b);
            T den = d00 * d11 - d01 * d01;
            T l1 = (d11 * d20 - d01 * d21) / den, l2 = (d00 * d21 - d01 * d20) / den;
            T l0 = T(1) - (l1 + l2);
            auto cl = [](T v) { return std::min(T(1), std::max(T(0), v)); };
            f(y, mk(cl(l0), cl(l1), cl(l2)), (T)r.w[k] * area);
        }
    };
    for (int i = 0; i < m; i++)
        for (int j = 0; i + j < m; j++) {
            do_tri(corner(i, j), corner(i + 1, j), corner(i, j + 1));
            if (i + j < m - 1) do_tri(corner(i + 1, j + 1), corner(i + 1, j), corner(i, j + 1));
        }
}

// ----------------------------------------- collocation: weakly singular rule --
// Polar rule around the singular vertex y1 (integrate_collocation_weakly_singular,
// integrator_collocation.hpp:19-63): theta in [0, alpha], radius in
// [0, a / cos(theta - beta)], composite Gauss in both.  f(y, rhat, r, w) is
// called with the polar weight only; the caller supplies r*kernel.
template <class T, class Fn>
void polar_rule(const v3<T> &y1, const v3<T> &y2, const v3<T> &y3, const QParams &q, Fn &&f) {
    const v3<T> v1 = y2 - y1, v2 = y3 - y1;
    const v3<T> u1 = unit(v1), u2 = unit(v2);
    const v3<T> xh = u1;
    const v3<T> yh = unit(cross(cross(xh, u2), xh));
    const v3<T> v12 = v1 + dot(v1, unit(v1 - v2)) * unit(v2 - v1);
    const T a = norm(v12);
    const T alpha = std::acos(dot(u1, u2));
    const T beta = std::atan2(norm(cross(unit(v12), u1)), dot(unit(v12), u1));  // deliberate difference 3 (header)
    Line1D<T> th = composite_line<T>(T(0), alpha, q.order, q.sub);
    for (size_t it = 0; it < th.x.size(); it++) {
        const T t = th.x[it];
        const v3<T> rh = xh * std::cos(t) + yh * std::sin(t);
        Line1D<T> rr = composite_line<T>(T(0), a / std::cos(t - beta), q.order, q.sub);
        for (size_t ir = 0; ir < rr.x.size(); ir++) {
            const T r = rr.x[ir];
            f(y1 + r * rh, rh, r, th.w[it] * rr.w[ir]);
        }
    }
}

// Projection barycentrics clamped to [0,1] (used by the singular G integrand,
// G_collocation.cu:133-135).
template <class T> v3<T> bary_clamped(const v3<T> &y, const v3<T> &y1, const v3<T> &y2, const v3<T> &y3) {
    v3<T> a = y2 - y1, b = y3 - y1, c = y - y1;
    T d00 = dot(a, a), d01 = dot(a, b), d11 = dot(b, b), d20 = dot(c, a), d21 = dot(c, b);
    T den = d00 * d11 - d01 * d01;
    T l1 = (d11 * d20 - d01 * d21) / den, l2 = (d00 * d21 - d01 * d20) / den;
    T l0 = T(1) - (l1 + l2);
    auto cl = [](T v) { return std::min(T(1), std::max(T(0), v)); };
    return mk(cl(l0), cl(l1), cl(l2));
}

// ---------------------------------------- collocation: strongly singular rule --
// Guiggiani-Gigante rule on the unit right triangle in (rho, theta) around the
// singular vertex y1 (integrate_collocation_strongly_singular,
// integrator_collocation.hpp:65-124).  Returns, for every (p,q) and vertex k,
// the finite-part integral of p_hat_pq * phi_k into out[(p*3+q)*3+k].
template <class T> void guiggiani_rule(const v3<T> &y1, const v3<T> &y2, const v3<T> &y3, const v3<T> &n, const KParam<T> &kp,
                                       const QParams &q, cx<T> out[27]) {
    const v3<T> v1 = y2 - y1, v2 = y3 - y1;
    const T J = norm(cross(v1, v2));
    for (int k = 0; k < 27; k++) out[k] = cx<T>(0);
    Line1D<T> th = composite_line<T>(T(0), pi_t<T>() / T(2), q.order, q.sub);
    for (size_t it = 0; it < th.x.size(); it++) {
        const T t = th.x[it], ct = std::cos(t), st = std::sin(t);
        const v3<T> Av = ct * v1 + st * v2;
        const v3<T> Ah = unit(Av);
        const T A = norm(Av);
        const T rho_bar = T(1) / (std::cos(t - pi_t<T>() / T(4)) * std::sqrt(T(2)));
        Line1D<T> rr = composite_line<T>(T(0), rho_bar, q.order, q.sub);
        const T logterm = std::log(rho_bar * A);
        // cancelling term f_pq = p_sing(Ahat, A) * J (basis (1,0,0))
        T fpq[9];
        for (int p = 0; p < 3; p++)
            for (int qq = 0; qq < 3; qq++) fpq[p * 3 + qq] = k_p_sing(Ah, A, n, p, qq, kp) * J;
        cx<T> part1[27];
        for (int k = 0; k < 27; k++) part1[k] = cx<T>(0);
        for (size_t ir = 0; ir < rr.x.size(); ir++) {
            const T rho = rr.x[ir], w = rr.w[ir];
            const T L[3] = {T(1) - rho * (ct + st), rho * ct, rho * st};
            const v3<T> rv = rho * Av;
            const T r = norm(rv);
            const v3<T> rh = unit(rv);
            const KExp<T> e = k_exp(r, kp);
            for (int p = 0; p < 3; p++)
                for (int qq = 0; qq < 3; qq++) {
                    const cx<T> F = k_p(rh, r, n, p, qq, kp, e);
                    for (int k = 0; k < 3; k++) {
                        cx<T> val = F * (L[k] * J * rho);
                        if (k == 0) val -= fpq[p * 3 + qq] / rho;
                        part1[(p * 3 + qq) * 3 + k] += w * val;
                    }
                }
        }
        for (int p = 0; p < 9; p++) {
            for (int k = 0; k < 3; k++) {
                cx<T> v = part1[p * 3 + k];
                if (k == 0) v += fpq[p] * logterm;
                out[p * 3 + k] += th.w[it] * v;
            }
        }
    }
}

// Singular-vertex test shared by every collocation kernel (G_collocation.cu:78-88).
inline int singular_corner(const Mesh &m, int i, int f) {
    for (int c = 0; c < 3; c++)
        if (same_vertex(m, i, m.F[3 * f + c])) return c;
    return -1;
}

// ------------------------------------------------ collocation: row kernels --
// Row vertex i of U_hat (compute_elastodynamic_G_kernel_collocation, G_collocation.cu:18-142).
template <class Pol>
void colloc_G_row(const Mesh &m, int i, double c1, double c2, double rho, cx<double> s, const QParams &q, cx<typename Pol::Acc> *U) {
    using T = typename Pol::ColFar;
    using S = typename Pol::Sing;
    using A = typename Pol::Acc;
    const int n3 = 3 * m.nv;
    for (int p = 0; p < 3; p++)
        for (int c = 0; c < n3; c++) U[(size_t)(3 * i + p) * n3 + c] = cx<A>(0);
    const KParam<T> kt{(T)c1, (T)c2, (T)rho, cx<T>((T)s.real(), (T)s.imag())};
    const KParam<S> ks{(S)c1, (S)c2, (S)rho, cx<S>((S)s.real(), (S)s.imag())};
    const v3<T> x = cast3<T>(m.V + 3 * i);
    for (int f = 0; f < m.nf; f++) {
        const int sc = singular_corner(m, i, f);
        if (sc < 0) {
            const int j[3] = {m.F[3 * f], m.F[3 * f + 1], m.F[3 * f + 2]};
            const v3<T> y1 = cast3<T>(m.V + 3 * j[0]), y2 = cast3<T>(m.V + 3 * j[1]), y3 = cast3<T>(m.V + 3 * j[2]);
            cx<T> acc[3][9];
            for (auto &a : acc)
                for (auto &b : a) b = cx<T>(0);
            far_rule<T>(y1, y2, y3, q, [&](const v3<T> &y, const v3<T> &L, T w) {
                const v3<T> rv = y - x;
                const T r = norm(rv);
                const v3<T> rh = unit(rv);
                const KExp<T> e = k_exp(r, kt);
                cx<T> u[9];
                for (int a = 0; a < 3; a++)
                    for (int b = 0; b <= a; b++) {
                        u[a * 3 + b] = k_u_r(rh, r, a, b, kt, e) / r;
                        u[b * 3 + a] = u[a * 3 + b];
                    }
                for (int k = 0; k < 3; k++)
                    for (int e = 0; e < 9; e++) acc[k][e] += (u[e] * L[k]) * w;
            });
            for (int k = 0; k < 3; k++)
                for (int a = 0; a < 3; a++)
                    for (int b = 0; b < 3; b++)
                        U[(size_t)(3 * i + a) * n3 + 3 * j[k] + b] +=
                            cx<A>((A)acc[k][a * 3 + b].real(), (A)acc[k][a * 3 + b].imag());
        } else {
            const int j[3] = {m.F[3 * f + sc], m.F[3 * f + (sc + 1) % 3], m.F[3 * f + (sc + 2) % 3]};
            const v3<S> y1 = cast3<S>(m.V + 3 * j[0]), y2 = cast3<S>(m.V + 3 * j[1]), y3 = cast3<S>(m.V + 3 * j[2]);
            cx<S> acc[3][9];
            for (auto &a : acc)
                for (auto &b : a) b = cx<S>(0);
            polar_rule<S>(y1, y2, y3, q, [&](const v3<S> &y, const v3<S> &rh, S r, S w) {
                const v3<S> L = bary_clamped(y, y1, y2, y3);
                const KExp<S> e = k_exp(r, ks);
                for (int a = 0; a < 3; a++)
                    for (int b = 0; b <= a; b++) {
                        cx<S> v = k_u_r(rh, r, a, b, ks, e);
                        for (int k = 0; k < 3; k++) acc[k][a * 3 + b] += w * (v * L[k]);
                    }
            });
            for (int a = 0; a < 3; a++)
                for (int b = 0; b <= a; b++)
                    for (int k = 0; k < 3; k++) {
                        cx<A> v((A)acc[k][a * 3 + b].real(), (A)acc[k][a * 3 + b].imag());
                        U[(size_t)(3 * i + a) * n3 + 3 * j[k] + b] += v;
                        if (a != b) U[(size_t)(3 * i + b) * n3 + 3 * j[k] + a] += v;
                    }
        }
    }
}

// Row vertex i of P_hat (compute_elastodynamic_H_kernel_collocation, H_collocation.cu:18-158).
template <class Pol>
void colloc_H_row(const Mesh &m, int i, double c1, double c2, cx<double> s, const QParams &q, cx<typename Pol::Acc> *P) {
    using T = typename Pol::ColHFar;
    using S = typename Pol::Sing;
    using A = typename Pol::Acc;
    const int n3 = 3 * m.nv;
    for (int p = 0; p < 3; p++)
        for (int c = 0; c < n3; c++) P[(size_t)(3 * i + p) * n3 + c] = cx<A>(0);
    const KParam<T> kt{(T)c1, (T)c2, T(0), cx<T>((T)s.real(), (T)s.imag())};
    const KParam<S> ks{(S)c1, (S)c2, S(0), cx<S>((S)s.real(), (S)s.imag())};
    const v3<T> x = cast3<T>(m.V + 3 * i);
    for (int f = 0; f < m.nf; f++) {
        const int sc = singular_corner(m, i, f);
        if (sc < 0) {
            const int j[3] = {m.F[3 * f], m.F[3 * f + 1], m.F[3 * f + 2]};
            const v3<T> n = cast3<T>(m.N.data() + 3 * f);
            const v3<T> y1 = cast3<T>(m.V + 3 * j[0]), y2 = cast3<T>(m.V + 3 * j[1]), y3 = cast3<T>(m.V + 3 * j[2]);
            cx<T> acc[3][9];
            for (auto &a : acc)
                for (auto &b : a) b = cx<T>(0);
            far_rule<T>(y1, y2, y3, q, [&](const v3<T> &y, const v3<T> &L, T w) {
                const v3<T> rv = y - x;
                const T r = norm(rv);
                const v3<T> rh = unit(rv);
                const KExp<T> e = k_exp(r, kt);
                cx<T> pv[9];
                for (int a = 0; a < 3; a++)
                    for (int b = 0; b < 3; b++) pv[a * 3 + b] = k_p(rh, r, n, a, b, kt, e);
                for (int k = 0; k < 3; k++)
                    for (int e = 0; e < 9; e++) acc[k][e] += (pv[e] * L[k]) * w;
            });
            for (int k = 0; k < 3; k++)
                for (int a = 0; a < 3; a++)
                    for (int b = 0; b < 3; b++)
                        P[(size_t)(3 * i + a) * n3 + 3 * j[k] + b] +=
                            cx<A>((A)acc[k][a * 3 + b].real(), (A)acc[k][a * 3 + b].imag());
        } else {
            const int j[3] = {m.F[3 * f + sc], m.F[3 * f + (sc + 1) % 3], m.F[3 * f + (sc + 2) % 3]};
            const v3<S> n = cast3<S>(m.N.data() + 3 * f);
            const v3<S> y1 = cast3<S>(m.V + 3 * j[0]), y2 = cast3<S>(m.V + 3 * j[1]), y3 = cast3<S>(m.V + 3 * j[2]);
            cx<S> out[27];
            guiggiani_rule<S>(y1, y2, y3, n, ks, q, out);
            for (int a = 0; a < 3; a++)
                for (int b = 0; b < 3; b++)
                    for (int k = 0; k < 3; k++)
                        P[(size_t)(3 * i + a) * n3 + 3 * j[k] + b] +=
                            cx<A>((A)out[(a * 3 + b) * 3 + k].real(), (A)out[(a * 3 + b) * 3 + k].imag());
        }
    }
}

// Row block of B_trans_hat (B_trans_collocation.cu:17-124).
template <class Pol>
void colloc_Bt_row(const Mesh &m, int i, double c1, double c2, cx<double> s, const QParams &q, cx<typename Pol::Acc> *Bt) {
    using T = typename Pol::ColFar;
    using S = typename Pol::Sing;
    using A = typename Pol::Acc;
    const KParam<T> kt{(T)c1, (T)c2, T(0), cx<T>((T)s.real(), (T)s.imag())};
    const KParam<S> ks{(S)c1, (S)c2, S(0), cx<S>((S)s.real(), (S)s.imag())};
    const v3<T> x = cast3<T>(m.V + 3 * i);
    cx<T> Bi[9];
    for (auto &b : Bi) b = cx<T>(0);
    for (int f = 0; f < m.nf; f++) {
        const int sc = singular_corner(m, i, f);
        if (sc < 0) {
            const int j[3] = {m.F[3 * f], m.F[3 * f + 1], m.F[3 * f + 2]};
            const v3<T> n = cast3<T>(m.N.data() + 3 * f);
            const v3<T> y1 = cast3<T>(m.V + 3 * j[0]), y2 = cast3<T>(m.V + 3 * j[1]), y3 = cast3<T>(m.V + 3 * j[2]);
            cx<T> acc[9];
            for (auto &b : acc) b = cx<T>(0);
            far_rule<T>(y1, y2, y3, q, [&](const v3<T> &y, const v3<T> &, T w) {
                const v3<T> rv = y - x;
                const T r = norm(rv);
                const v3<T> rh = unit(rv);
                const KExp<T> e = k_exp(r, kt);
                for (int a = 0; a < 3; a++)
                    for (int b = 0; b <= a; b++) {
                        cx<T> v = (k_bt_r(rh, r, n, a, b, kt, e) / r) * w;
                        acc[a * 3 + b] += v;
                        if (a != b) acc[b * 3 + a] += v;
                    }
            });
            for (int e = 0; e < 9; e++) Bi[e] += acc[e];
        } else {
            const int j[3] = {m.F[3 * f + sc], m.F[3 * f + (sc + 1) % 3], m.F[3 * f + (sc + 2) % 3]};
            const v3<S> n = cast3<S>(m.N.data() + 3 * f);
            const v3<S> y1 = cast3<S>(m.V + 3 * j[0]), y2 = cast3<S>(m.V + 3 * j[1]), y3 = cast3<S>(m.V + 3 * j[2]);
            cx<S> acc[9];
            for (auto &b : acc) b = cx<S>(0);
            polar_rule<S>(y1, y2, y3, q, [&](const v3<S> &, const v3<S> &rh, S r, S w) {
                const KExp<S> e = k_exp(r, ks);
                for (int a = 0; a < 3; a++)
                    for (int b = 0; b <= a; b++) acc[a * 3 + b] += w * k_bt_r(rh, r, n, a, b, ks, e);
            });
            for (int a = 0; a < 3; a++)
                for (int b = 0; b <= a; b++) {
                    cx<T> v((T)acc[a * 3 + b].real(), (T)acc[a * 3 + b].imag());
                    Bi[a * 3 + b] += v;
                    if (a != b) Bi[b * 3 + a] += v;
                }
        }
    }
    for (int e = 0; e < 9; e++) Bt[(size_t)(3 * i + e / 3) * 3 + e % 3] = cx<A>((A)Bi[e].real(), (A)Bi[e].imag());
}

// Row block of B_angular_hat (B_angular_collocation.cu:17-143).
template <class Pol>
void colloc_Ba_row(const Mesh &m, int i, double c1, double c2, cx<double> s, const QParams &q, cx<typename Pol::Acc> *Ba) {
    using T = typename Pol::ColFar;
    using S = typename Pol::Sing;
    using A = typename Pol::Acc;
    const KParam<T> kt{(T)c1, (T)c2, T(0), cx<T>((T)s.real(), (T)s.imag())};
    const KParam<S> ks{(S)c1, (S)c2, S(0), cx<S>((S)s.real(), (S)s.imag())};
    const v3<T> x = cast3<T>(m.V + 3 * i);
    cx<T> Bi[27];
    for (auto &b : Bi) b = cx<T>(0);
    for (int f = 0; f < m.nf; f++) {
        const int sc = singular_corner(m, i, f);
        if (sc < 0) {
            const int j[3] = {m.F[3 * f], m.F[3 * f + 1], m.F[3 * f + 2]};
            const v3<T> n = cast3<T>(m.N.data() + 3 * f);
            const v3<T> y1 = cast3<T>(m.V + 3 * j[0]), y2 = cast3<T>(m.V + 3 * j[1]), y3 = cast3<T>(m.V + 3 * j[2]);
            for (int kk = 0; kk < 3; kk++) {  // three separate integrations (B_angular_collocation.cu:97-105)
                cx<T> acc[9];
                for (auto &b : acc) b = cx<T>(0);
                far_rule<T>(y1, y2, y3, q, [&](const v3<T> &y, const v3<T> &, T w) {
                    const v3<T> rv = y - x;
                    const T r = norm(rv);
                    const v3<T> rh = unit(rv);
                    const KExp<T> e = k_exp(r, kt);
                    for (int a = 0; a < 3; a++)
                        for (int b = 0; b <= a; b++) {
                            cx<T> v = (k_ba_r(rh, r, n, a, b, kk, kt, e) / r) * w;
                            acc[a * 3 + b] += v;
                            if (a != b) acc[b * 3 + a] += v;
                        }
                });
                for (int a = 0; a < 3; a++)
                    for (int b = 0; b < 3; b++) Bi[a * 9 + b * 3 + kk] += acc[a * 3 + b];
            }
        } else {
            const int j[3] = {m.F[3 * f + sc], m.F[3 * f + (sc + 1) % 3], m.F[3 * f + (sc + 2) % 3]};
            const v3<S> n = cast3<S>(m.N.data() + 3 * f);
            const v3<S> y1 = cast3<S>(m.V + 3 * j[0]), y2 = cast3<S>(m.V + 3 * j[1]), y3 = cast3<S>(m.V + 3 * j[2]);
            cx<S> acc[27];
            for (auto &b : acc) b = cx<S>(0);
            polar_rule<S>(y1, y2, y3, q, [&](const v3<S> &, const v3<S> &rh, S r, S w) {
                const KExp<S> e = k_exp(r, ks);
                for (int a = 0; a < 3; a++)
                    for (int b = 0; b <= a; b++)
                        for (int kk = 0; kk < 3; kk++) acc[a * 9 + b * 3 + kk] += w * k_ba_r(rh, r, n, a, b, kk, ks, e);
            });
            for (int a = 0; a < 3; a++)
                for (int b = 0; b <= a; b++)
                    for (int kk = 0; kk < 3; kk++) {
                        cx<T> v((T)acc[a * 9 + b * 3 + kk].real(), (T)acc[a * 9 + b * 3 + kk].imag());
                        Bi[a * 9 + b * 3 + kk] += v;
                        if (a != b) Bi[b * 9 + a * 3 + kk] += v;
                    }
        }
    }
    for (int e = 0; e < 27; e++) Ba[(size_t)(3 * i + e / 9) * 9 + e % 9] = cx<A>((A)Bi[e].real(), (A)Bi[e].imag());
}

// ------------------------------------------------------ Galerkin integrands --
// Kernel blocks shared by the non-singular and singular Galerkin integrals
// (G/H/B_trans/B_angular _galerkin_non_singular.cu and galerkin_singular.cpp).
// Output layout per point pair: U: 3x3, P: 3x3, Bt: 3x3, Ba: 3x9 (column 3k+j).
template <class T> struct GalKernels {
    Which w;
    KParam<T> k;
    void eval(const v3<T> &x, const v3<T> &y, const v3<T> &n, cx<T> *u, cx<T> *p, cx<T> *bt, cx<T> *ba) const {
        const v3<T> rv = y - x;
        const T r = norm(rv);
        const v3<T> rh = unit(rv);
        const KExp<T> e = k_exp(r, k);
        if (w.G)
            for (int a = 0; a < 3; a++)
                for (int b = 0; b <= a; b++) u[a * 3 + b] = u[b * 3 + a] = k_u_r(rh, r, a, b, k, e) / r;
        if (w.H)
            for (int a = 0; a < 3; a++)
                for (int b = 0; b < 3; b++) p[a * 3 + b] = k_p(rh, r, n, a, b, k, e);
        if (w.Bt)
            for (int a = 0; a < 3; a++)
                for (int b = 0; b <= a; b++) bt[a * 3 + b] = bt[b * 3 + a] = k_bt_r(rh, r, n, a, b, k, e) / r;
        if (w.Ba)
            for (int kk = 0; kk < 3; kk++)
                for (int a = 0; a < 3; a++)
                    for (int b = 0; b <= a; b++)
                        ba[a * 9 + 3 * kk + b] = ba[b * 9 + 3 * kk + a] = k_ba_r(rh, r, n, a, b, kk, k, e) / r;
    }
};

// Per face pair accumulator: U/P 9x9 (row 3i+p, col 3j+q), Bt 9x3, Ba 9x9 (row 3i+p, col 3k+q).
template <class T> struct PairAcc {
    cx<T> U[81], P[81], Bt[27], Ba[81];
    void zero() {
        for (auto &v : U) v = cx<T>(0);
        for (auto &v : P) v = cx<T>(0);
        for (auto &v : Bt) v = cx<T>(0);
        for (auto &v : Ba) v = cx<T>(0);
    }
    // add wt * Lx_i * Ly_j * kernel
    void add(const Which &w, const T Lx[3], const T Ly[3], T wt, const cx<T> *u, const cx<T> *p, const cx<T> *bt, const cx<T> *ba) {
        for (int i = 0; i < 3; i++) {
            const T wi = wt * Lx[i];
            if (w.Bt)
                for (int a = 0; a < 3; a++)
                    for (int b = 0; b < 3; b++) Bt[(3 * i + a) * 3 + b] += bt[a * 3 + b] * wi;
            if (w.Ba)
                for (int a = 0; a < 3; a++)
                    for (int c = 0; c < 9; c++) Ba[(3 * i + a) * 9 + c] += ba[a * 9 + c] * wi;
            for (int j = 0; j < 3; j++) {
                const T wij = wi * Ly[j];
                if (w.G)
                    for (int a = 0; a < 3; a++)
                        for (int b = 0; b < 3; b++) U[(3 * i + a) * 9 + 3 * j + b] += u[a * 3 + b] * wij;
                if (w.H)
                    for (int a = 0; a < 3; a++)
                        for (int b = 0; b < 3; b++) P[(3 * i + a) * 9 + 3 * j + b] += p[a * 3 + b] * wij;
            }
        }
    }
    void scale(T s) {
        for (auto &v : U) v *= s;
        for (auto &v : P) v *= s;
        for (auto &v : Bt) v *= s;
        for (auto &v : Ba) v *= s;
    }
};

// Scatter a face-pair block with test vertices j1[] and source vertices j2[].
template <class T, class A>
void scatter_pair(const Which &w, const PairAcc<T> &acc, const int j1[3], const int j2[3], Laplace<A> &out) {
    const int n3 = out.n3;
    for (int i = 0; i < 3; i++)
        for (int a = 0; a < 3; a++) {
            const size_t row = (size_t)(3 * j1[i] + a);
            if (w.Bt)
                for (int b = 0; b < 3; b++) {
                    cx<T> v = acc.Bt[(3 * i + a) * 3 + b];
                    out.Bt[row * 3 + b] += cx<A>((A)v.real(), (A)v.imag());
                }
            if (w.Ba)
                for (int kk = 0; kk < 3; kk++)
                    for (int b = 0; b < 3; b++) {
                        cx<T> v = acc.Ba[(3 * i + a) * 9 + 3 * kk + b];
                        out.Ba[row * 9 + b * 3 + kk] += cx<A>((A)v.real(), (A)v.imag());
                    }
            for (int j = 0; j < 3; j++)
                for (int b = 0; b < 3; b++) {
                    if (w.G) {
                        cx<T> v = acc.U[(3 * i + a) * 9 + 3 * j + b];
                        out.U[row * n3 + 3 * j2[j] + b] += cx<A>((A)v.real(), (A)v.imag());
                    }
                    if (w.H) {
                        cx<T> v = acc.P[(3 * i + a) * 9 + 3 * j + b];
                        out.P[row * n3 + 3 * j2[j] + b] += cx<A>((A)v.real(), (A)v.imag());
                    }
                }
        }
}

// Non-singular face pair (integrate_galerkin_nonsingular, integrator_galerkin.hpp:214-260):
// subdivided triangle rule on both reference triangles, barycentrics (1-a-b, a, b).
template <class T>
bool galerkin_nonsingular_pair(const Mesh &m, int f1, int f2, const GalKernels<T> &K, const QParams &q, const RefPts &ref,
                               PairAcc<T> &acc) {
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++)
            if (same_vertex(m, m.F[3 * f1 + a], m.F[3 * f2 + b])) return false;
    const int *j1 = m.F + 3 * f1, *j2 = m.F + 3 * f2;
    const v3<T> x0 = cast3<T>(m.V + 3 * j1[0]), x1 = cast3<T>(m.V + 3 * j1[1]), x2 = cast3<T>(m.V + 3 * j1[2]);
    const v3<T> y0 = cast3<T>(m.V + 3 * j2[0]), y1 = cast3<T>(m.V + 3 * j2[1]), y2 = cast3<T>(m.V + 3 * j2[2]);
    const v3<T> n = cast3<T>(m.N.data() + 3 * f2);
    const v3<T> ex0 = x1 - x0, ex1 = x2 - x0, ey0 = y1 - y0, ey1 = y2 - y0;
    acc.zero();
    cx<T> u[9], p[9], bt[9], ba[27];
    const size_t np = ref.w.size();
    for (size_t ia = 0; ia < np; ia++) {
        const T xa = (T)ref.b[ia][1], xb = (T)ref.b[ia][2];
        const T Lx[3] = {T(1) - xa - xb, xa, xb};
        const v3<T> x = x0 + xa * ex0 + xb * ex1;
        for (size_t ib = 0; ib < np; ib++) {
            const T ya = (T)ref.b[ib][1], yb = (T)ref.b[ib][2];
            const T Ly[3] = {T(1) - ya - yb, ya, yb};
            const v3<T> y = y0 + ya * ey0 + yb * ey1;
            K.eval(x, y, n, u, p, bt, ba);
            // reference-triangle rule weights carry the reference area 1/2
            acc.add(K.w, Lx, Ly, (T)(ref.w[ia] * 0.5L) * (T)(ref.w[ib] * 0.5L), u, p, bt, ba);
        }
    }
    acc.scale(norm(cross(ex0, ex1)) * norm(cross(ey0, ey1)));
    return true;
}

// Sauter-Schwab rules for coincident / edge-adjacent / vertex-adjacent face pairs
// (integrator_galerkin.hpp:42-212), classification as in integrate_galerkin_singular
// (integrator_galerkin.hpp:262-334).  Returns false for non-touching pairs.
template <class T>
bool galerkin_singular_pair(const Mesh &m, int f1, int f2, const GalKernels<T> &K, const QParams &q, PairAcc<T> &acc,
                            int j1[3], int j2[3]) {
    const int *F = m.F;
    int kind = 0;  // 1 coincident, 2 edge, 3 vertex
    if (f1 == f2) {
        kind = 1;
        for (int k = 0; k < 3; k++) j1[k] = j2[k] = F[3 * f1 + k];
    }
    if (!kind) {
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                const bool exact = F[3 * f1 + i] == F[3 * f2 + j] && F[3 * f1 + (i + 1) % 3] == F[3 * f2 + (j + 2) % 3];
                const bool mapped = m.vmap != nullptr && m.vmap[F[3 * f1 + i]] == m.vmap[F[3 * f2 + j]] &&
                                    m.vmap[F[3 * f1 + (i + 1) % 3]] == m.vmap[F[3 * f2 + (j + 2) % 3]];
                if (exact || mapped) {
                    kind = 2;
                    j1[0] = F[3 * f1 + i], j1[1] = F[3 * f1 + (i + 1) % 3], j1[2] = F[3 * f1 + (i + 2) % 3];
                    j2[0] = F[3 * f2 + j], j2[1] = F[3 * f2 + (j + 2) % 3], j2[2] = F[3 * f2 + (j + 1) % 3];
                    break;  // inner loop only, as in the reference
                }
            }
    }
    if (!kind) {
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                if (same_vertex(m, F[3 * f1 + i], F[3 * f2 + j])) {
                    kind = 3;
                    j1[0] = F[3 * f1 + i], j1[1] = F[3 * f1 + (i + 1) % 3], j1[2] = F[3 * f1 + (i + 2) % 3];
                    j2[0] = F[3 * f2 + j], j2[1] = F[3 * f2 + (j + 1) % 3], j2[2] = F[3 * f2 + (j + 2) % 3];
                    break;
                }
            }
    }
    if (!kind) return false;

    const v3<T> n = cast3<T>(m.N.data() + 3 * f2);
    Line1D<T> l = composite_line<T>(T(0), T(1), q.order, q.sub);
    const size_t nl = l.x.size();
    cx<T> u[9], p[9], bt[9], ba[27];
    acc.zero();
    PairAcc<T> tmp;

    // x(xh) = p0 + d0 xh0 + d1 xh1 and y likewise; barycentrics (1-h0, h0-h1, h1)
    v3<T> px0, dx0, dx1, py0, dy0, dy1;
    T jac;
    if (kind == 1) {
        const v3<T> x0 = cast3<T>(m.V + 3 * j1[0]), x1 = cast3<T>(m.V + 3 * j1[1]), x2 = cast3<T>(m.V + 3 * j1[2]);
        px0 = py0 = x0, dx0 = dy0 = x1 - x0, dx1 = dy1 = x2 - x1;
        const T J = norm(cross(dx0, dx1));
        jac = J * J;
    } else if (kind == 2) {
        const v3<T> xy0 = cast3<T>(m.V + 3 * j1[0]), xy1 = cast3<T>(m.V + 3 * j1[1]), x2 = cast3<T>(m.V + 3 * j1[2]),
                    y2 = cast3<T>(m.V + 3 * j2[2]);
        px0 = py0 = xy0, dx0 = dy0 = xy1 - xy0, dx1 = x2 - xy1, dy1 = y2 - xy1;
        jac = norm(cross(dx0, dx1)) * norm(cross(dy0, dy1));
    } else {
        const v3<T> xy0 = cast3<T>(m.V + 3 * j1[0]), x1 = cast3<T>(m.V + 3 * j1[1]), x2 = cast3<T>(m.V + 3 * j1[2]),
                    y1 = cast3<T>(m.V + 3 * j2[1]), y2 = cast3<T>(m.V + 3 * j2[2]);
        px0 = py0 = xy0, dx0 = x1 - xy0, dx1 = x2 - x1, dy0 = y1 - xy0, dy1 = y2 - y1;
        jac = norm(cross(dx0, dx1)) * norm(cross(dy0, dy1));
    }
    auto eval = [&](T xh0, T xh1, T yh0, T yh1, T wt) {
        const v3<T> x = px0 + dx0 * xh0 + dx1 * xh1, y = py0 + dy0 * yh0 + dy1 * yh1;
        const T Lx[3] = {T(1) - xh0, xh0 - xh1, xh1}, Ly[3] = {T(1) - yh0, yh0 - yh1, yh1};
        K.eval(x, y, n, u, p, bt, ba);
        tmp.add(K.w, Lx, Ly, wt, u, p, bt, ba);
    };
    for (size_t a = 0; a < nl; a++)
        for (size_t b = 0; b < nl; b++)
            for (size_t c = 0; c < nl; c++)
                for (size_t d = 0; d < nl; d++) {
                    const T e1 = l.x[a], e2 = l.x[b], e3 = l.x[c], xi = l.x[d];
                    const T w = l.w[a] * l.w[b] * l.w[c] * l.w[d];
                    tmp.zero();
                    if (kind == 1) {
                        const T W = w * xi * xi * xi * e1 * e1 * e2;
                        eval(xi, xi * (T(1) - e1 + e1 * e2), xi * (T(1) - e1 * e2 * e3), xi * (T(1) - e1), W);
                        eval(xi * (T(1) - e1 * e2 * e3), xi * (T(1) - e1), xi, xi * (T(1) - e1 + e1 * e2), W);
                        eval(xi, xi * e1 * (T(1) - e2 + e2 * e3), xi * (T(1) - e1 * e2), xi * e1 * (T(1) - e2), W);
                        eval(xi * (T(1) - e1 * e2), xi * e1 * (T(1) - e2), xi, xi * e1 * (T(1) - e2 + e2 * e3), W);
                        eval(xi * (T(1) - e1 * e2 * e3), xi * e1 * (T(1) - e2 * e3), xi, xi * e1 * (T(1) - e2), W);
                        eval(xi, xi * e1 * (T(1) - e2), xi * (T(1) - e1 * e2 * e3), xi * e1 * (T(1) - e2 * e3), W);
                    } else if (kind == 2) {
                        const T W = w * xi * xi * xi * e1 * e1;
                        eval(xi, xi * e1 * e3, xi * (T(1) - e1 * e2), xi * e1 * (T(1) - e2), W);
                        eval(xi, xi * e1, xi * (T(1) - e1 * e2 * e3), xi * e1 * e2 * (T(1) - e3), W * e2);
                        eval(xi * (T(1) - e1 * e2), xi * e1 * (T(1) - e2), xi, xi * e1 * e2 * e3, W * e2);
                        eval(xi * (T(1) - e1 * e2 * e3), xi * e1 * e2 * (T(1) - e3), xi, xi * e1, W * e2);
                        eval(xi * (T(1) - e1 * e2 * e3), xi * e1 * (T(1) - e2 * e3), xi, xi * e1 * e2, W * e2);
                    } else {
                        const T W = w * xi * xi * xi * e2;
                        eval(xi, xi * e1, xi * e2, xi * e2 * e3, W);
                        eval(xi * e2, xi * e2 * e3, xi, xi * e1, W);
                    }
                    for (int k = 0; k < 81; k++) acc.U[k] += tmp.U[k], acc.P[k] += tmp.P[k], acc.Ba[k] += tmp.Ba[k];
                    for (int k = 0; k < 27; k++) acc.Bt[k] += tmp.Bt[k];
                }
    acc.scale(jac);
    return true;
}

// ------------------------------------------------------------ C matrices --
// Collocation jump term (compute_C_matrix_collocation, elastics_bem_matrix_collection.cpp:387-445):
// 1/2 I + dihedral-angle term per edge - 1/(8 pi (1-nu)) ((r2-r1) x n) n^T per face corner.
void C_collocation(const double *V, int nv, const int *F, int nf, double nu, double *C /* 3nv x 3 */) {
    std::vector<double> N(3 * (size_t)nf);
    face_normals(V, F, nf, N.data());
    std::fill(C, C + 9 * (size_t)nv, 0.0);
    for (int i = 0; i < nv; i++)
        for (int a = 0; a < 3; a++) C[(3 * i + a) * 3 + a] += 0.5;
    // half-edges a->b of face f, listed in the order c*nf + f (edge opposite corner c)
    std::map<std::pair<int, int>, std::vector<int>> edges;
    for (int c = 0; c < 3; c++)
        for (int f = 0; f < nf; f++) {
            int a = F[3 * f + (c + 1) % 3], b = F[3 * f + (c + 2) % 3];
            edges[{std::min(a, b), std::max(a, b)}].push_back(c * nf + f);
        }
    auto sgn = [](double x) { return std::abs(x) < 1e-12 ? 0.0 : (x > 0 ? 1.0 : -1.0); };
    for (auto &kv : edges) {
        std::vector<int> he = kv.second;
        std::sort(he.begin(), he.end());
        if (he.size() < 2) continue;  // boundary edge (closed meshes have none)
        const int e1 = he[0], e2 = he[1];
        const int f1 = e1 % nf, f2 = e2 % nf, c1 = e1 / nf;
        const int a = F[3 * f1 + (c1 + 1) % 3], b = F[3 * f1 + (c1 + 2) % 3];
        const v3<double> n1 = cast3<double>(&N[3 * f1]), n2 = cast3<double>(&N[3 * f2]);
        const v3<double> r = unit(cast3<double>(V + 3 * a) - cast3<double>(V + 3 * b));
        const double wgt = (1.0 / (4.0 * (double)kPiL)) * sgn(dot(cross(n1, n2), r)) *
                           std::acos(std::min(1.0, std::max(-1.0, dot(n1, n2))));
        for (int k = 0; k < 3; k++) C[(3 * a + k) * 3 + k] += wgt, C[(3 * b + k) * 3 + k] += wgt;
    }
    for (int f = 0; f < nf; f++) {
        const v3<double> n = cast3<double>(&N[3 * f]);
        for (int i = 0; i < 3; i++) {
            const v3<double> p = cast3<double>(V + 3 * F[3 * f + i]);
            const v3<double> r1 = unit(cast3<double>(V + 3 * F[3 * f + (i + 1) % 3]) - p);
            const v3<double> r2 = unit(cast3<double>(V + 3 * F[3 * f + (i + 2) % 3]) - p);
            const v3<double> t = cross(r2 - r1, n);
            const double k = 1.0 / (8.0 * (double)kPiL * (1.0 - nu));
            for (int a = 0; a < 3; a++)
                for (int b = 0; b < 3; b++) C[(3 * F[3 * f + i] + a) * 3 + b] -= k * t[a] * n[b];
        }
    }
}

// Vertex-split mesh: the jump term is evaluated on the original (unsplit) mesh and copied to every copy
// of a vertex (elastics_bem_matrix_collection.cpp:392-398,436-442).  The original mesh is recovered from
// the map: original vertex vmap[v] sits at V[v] and face f reads vmap[F[f]].
void C_collocation_split(const double *V, int nv, const int *F, int nf, const int *vmap, double nu, double *C) {
    int no = 0;
    for (int v = 0; v < nv; v++) no = std::max(no, vmap[v] + 1);
    std::vector<double> Vo(3 * (size_t)no, 0.0);
    for (int v = 0; v < nv; v++)
        for (int k = 0; k < 3; k++) Vo[3 * (size_t)vmap[v] + k] = V[3 * v + k];
    std::vector<int> Fo(3 * (size_t)nf);
    for (size_t k = 0; k < Fo.size(); k++) Fo[k] = vmap[F[k]];
    std::vector<double> Co(9 * (size_t)no);
    C_collocation(Vo.data(), no, Fo.data(), nf, nu, Co.data());
    for (int v = 0; v < nv; v++)
        for (int e = 0; e < 9; e++) C[9 * (size_t)v + e] = Co[9 * (size_t)vmap[v] + e];
}

// Galerkin jump term: 1/2 of the P1 mass matrix per displacement component
// (compute_C_matrix_galerkin, elastics_bem_matrix_collection.cpp:447-467).
void C_galerkin_add(const double *V, int nv, const int *F, int nf, double *H0 /* dense 3nv x 3nv */) {
    const size_t n3 = 3 * (size_t)nv;
    for (int f = 0; f < nf; f++) {
        const v3<double> a = cast3<double>(V + 3 * F[3 * f]), b = cast3<double>(V + 3 * F[3 * f + 1]),
                         c = cast3<double>(V + 3 * F[3 * f + 2]);
        const double A2 = norm(cross(b - a, c - a));  // double area
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++)
                for (int k = 0; k < 3; k++)
                    H0[(3 * (size_t)F[3 * f + i] + k) * n3 + 3 * F[3 * f + j] + k] += (i == j) ? A2 / 24.0 : A2 / 48.0;
    }
}

// ------------------------------------------------------ CQ frequencies --
// s_l = gamma(R zeta^l)/dt, l = 0..floor(L/2) (G.cu:19-24,65)
std::vector<cx<double>> cq_frequencies(int L, double R, double dt, int bdf) {
    std::vector<cx<double>> s;
    for (int l = 0; l <= L / 2; l++) {
        const double ang = 2.0 * (double)kPiL * (double)l / (double)L;
        const cx<double> z = R * cx<double>(std::cos(ang), std::sin(ang));
        const cx<double> g = (bdf == 1) ? (1.0 - z) : (1.5 - 2.0 * z + z * z / 2.0);
        s.push_back(g / dt);
    }
    return s;
}

// Transform coefficient of frequency l for weight n:
// W_n = R^-n / L * sum_l Re(X_l zeta^{-nl}) (+ conjugate partner), G.cu:166-179.
template <class A> void transform_accumulate(int L, int l, const cx<A> *X, size_t M, A *W /* L x M */) {
    const bool pair = (l != 0) && (l != L / 2 || L % 2 == 1);
    for (int n = 0; n < L; n++) {
        const long long k1 = ((long long)n * l) % L;
        const A ang = -(A)2 * (A)kPiL * (A)k1 / (A)L;
        cx<A> tw(std::cos(ang), std::sin(ang));
        cx<A> tw2 = pair ? std::conj(tw) : cx<A>(0);  // zeta^{-n(L-l)} = conj(zeta^{-nl})
        A *dst = W + (size_t)n * M;
        for (size_t e = 0; e < M; e++) {
            A v = (X[e] * tw).real();
            if (pair) v += (std::conj(X[e]) * tw2).real();
            dst[e] += v;
        }
    }
}
template <class A> void transform_scale(int L, double R, size_t M, A *W) {
    for (int n = 0; n < L; n++) {
        const A sc = std::pow((A)R, -(A)n) / (A)L;
        for (size_t e = 0; e < M; e++) W[(size_t)n * M + e] *= sc;
    }
}

// ----------------------------------------------------- Laplace assembly --
template <class Pol>
void laplace_collocation(const Mesh &m, const orc_params &p, cx<double> s, const Which &w, int row0, int row1,
                         Laplace<typename Pol::Acc> &out) {
    const QParams q{p.order, p.subdiv};
#pragma omp parallel for schedule(dynamic, 1)
    for (int i = row0; i < row1; i++) {
        if (w.H) colloc_H_row<Pol>(m, i, p.c1, p.c2, s, q, out.P.data());
        if (w.G) colloc_G_row<Pol>(m, i, p.c1, p.c2, p.rho, s, q, out.U.data());
        if (w.Bt) colloc_Bt_row<Pol>(m, i, p.c1, p.c2, s, q, out.Bt.data());
        if (w.Ba) colloc_Ba_row<Pol>(m, i, p.c1, p.c2, s, q, out.Ba.data());
    }
}

// Rows [row0, row1) (vertex range): only the test faces with a corner in the range are integrated, so the rows
// of those vertices are complete and every other row holds partial sums (row-sample checks at full mesh sizes).
template <class Pol>
void laplace_galerkin(const Mesh &m, const orc_params &p, cx<double> s, const Which &w, int row0, int row1,
                      Laplace<typename Pol::Acc> &out) {
    using T = typename Pol::GalNS;
    using S = typename Pol::Sing;
    using A = typename Pol::Acc;
    const QParams q{p.order, p.subdiv};
    const RefPts ref = composite_tri(p.order, p.subdiv);
    GalKernels<T> KT{w, KParam<T>{(T)p.c1, (T)p.c2, (T)p.rho, cx<T>((T)s.real(), (T)s.imag())}};
    GalKernels<S> KS{w, KParam<S>{(S)p.c1, (S)p.c2, (S)p.rho, cx<S>((S)s.real(), (S)s.imag())}};
    out.zero();
    // One test face f1 at a time: its source faces are split into kChunks cyclic chunks (f2 = ch, ch + kChunks,
    // ...), each summed in face order into a row block of its own (the 9 rows of f1's corners) by whichever thread
    // takes the chunk; the blocks are then added into the output rows in chunk order.  The sums therefore do not
    // depend on the number of threads.  The reference serializes the same scatter with omp critical
    // (G_galerkin_non_singular.cu:84).
    constexpr int kChunks = 16;
    const int n3 = out.n3;
    struct RowBlock {
        std::vector<cx<A>> U, P, Bt, Ba;  // 9 x n3, 9 x n3, 9 x 3, 9 x 9 (row 3c + a for corner c of f1)
    };
    std::vector<RowBlock> blk(kChunks);
    for (auto &b : blk) {
        if (w.G) b.U.resize(9 * (size_t)n3);
        if (w.H) b.P.resize(9 * (size_t)n3);
        if (w.Bt) b.Bt.resize(27);
        if (w.Ba) b.Ba.resize(81);
    }
    for (int f1 = 0; f1 < m.nf; f1++) {
        bool in_range = false;
        for (int c = 0; c < 3; c++) in_range |= m.F[3 * f1 + c] >= row0 && m.F[3 * f1 + c] < row1;
        if (!in_range) continue;
#pragma omp parallel for schedule(dynamic, 1)
        for (int ch = 0; ch < kChunks; ch++) {
            RowBlock &mine = blk[ch];
            for (auto *v : {&mine.U, &mine.P, &mine.Bt, &mine.Ba}) std::fill(v->begin(), v->end(), cx<A>(0));
            PairAcc<T> accT;
            PairAcc<S> accS;
            // the corners of f1 in the order a rule reports them (singular rules permute them) -> local block rows
            auto add = [&](const auto &acc, const int j1[3], const int j2[3]) {
                for (int i = 0; i < 3; i++) {
                    int c = 0;
                    while (m.F[3 * f1 + c] != j1[i]) c++;
                    for (int a = 0; a < 3; a++) {
                        const size_t row = (size_t)(3 * c + a);
                        if (w.Bt)
                            for (int b = 0; b < 3; b++) {
                                const auto v = acc.Bt[(3 * i + a) * 3 + b];
                                mine.Bt[row * 3 + b] += cx<A>((A)v.real(), (A)v.imag());
                            }
                        if (w.Ba)
                            for (int kk = 0; kk < 3; kk++)
                                for (int b = 0; b < 3; b++) {
                                    const auto v = acc.Ba[(3 * i + a) * 9 + 3 * kk + b];
                                    mine.Ba[row * 9 + b * 3 + kk] += cx<A>((A)v.real(), (A)v.imag());
                                }
                        for (int j = 0; j < 3; j++)
                            for (int b = 0; b < 3; b++) {
                                if (w.G) {
                                    const auto v = acc.U[(3 * i + a) * 9 + 3 * j + b];
                                    mine.U[row * n3 + 3 * j2[j] + b] += cx<A>((A)v.real(), (A)v.imag());
                                }
                                if (w.H) {
                                    const auto v = acc.P[(3 * i + a) * 9 + 3 * j + b];
                                    mine.P[row * n3 + 3 * j2[j] + b] += cx<A>((A)v.real(), (A)v.imag());
                                }
                            }
                    }
                }
            };
            for (int f2 = ch; f2 < m.nf; f2 += kChunks) {  // cyclic: the expensive touching pairs spread out
                if (galerkin_nonsingular_pair<T>(m, f1, f2, KT, q, ref, accT)) {
                    add(accT, m.F + 3 * f1, m.F + 3 * f2);
                } else {
                    int j1[3], j2[3];
                    if (galerkin_singular_pair<S>(m, f1, f2, KS, q, accS, j1, j2)) add(accS, j1, j2);
                }
            }
        }
        for (const auto &b : blk)
            for (int c = 0; c < 3; c++)
                for (int a = 0; a < 3; a++) {
                    const size_t lr = (size_t)(3 * c + a), gr = (size_t)(3 * m.F[3 * f1 + c] + a);
                    if (w.G)
                        for (int k = 0; k < n3; k++) out.U[gr * n3 + k] += b.U[lr * n3 + k];
                    if (w.H)
                        for (int k = 0; k < n3; k++) out.P[gr * n3 + k] += b.P[lr * n3 + k];
                    if (w.Bt)
                        for (int k = 0; k < 3; k++) out.Bt[gr * 3 + k] += b.Bt[lr * 3 + k];
                    if (w.Ba)
                        for (int k = 0; k < 9; k++) out.Ba[gr * 9 + k] += b.Ba[lr * 9 + k];
                }
    }
}

template <class Pol>
void laplace_any(const Mesh &m, const orc_params &p, cx<double> s, const Which &w, int row0, int row1,
                 Laplace<typename Pol::Acc> &out) {
    if (p.galerkin)
        laplace_galerkin<Pol>(m, p, s, w, row0, row1, out);
    else
        laplace_collocation<Pol>(m, p, s, w, row0, row1, out);
}

Which which_from(int mask) { return Which{(mask & 1) != 0, (mask & 2) != 0, (mask & 4) != 0, (mask & 8) != 0}; }

template <class Pol>
int assemble_impl(const Mesh &m, const orc_params &p, int mask, double *G, double *H, double *Bt, double *Ba, double *B6,
                  double *spec_U, double *spec_P, double *spec_Bt, double *spec_Ba) {
    using A = typename Pol::Acc;
    Which w = which_from(mask);
    if (B6) w.Bt = w.Ba = true;
    if (Ba) w.Bt = true;  // B_angular's correction term needs B_trans
    const int L = p.L;
    const size_t n3 = 3 * (size_t)m.nv;
    const auto s = cq_frequencies(L, p.R, p.dt, p.bdf);
    std::vector<A> WG, WH, WBt, WBa;
    if (w.G) WG.assign(L * n3 * n3, A(0));
    if (w.H) WH.assign(L * n3 * n3, A(0));
    if (w.Bt) WBt.assign(L * n3 * 3, A(0));
    if (w.Ba) WBa.assign(L * n3 * 9, A(0));
    Laplace<A> lap;
    lap.alloc(m.nv, w);
    for (size_t l = 0; l < s.size(); l++) {
        laplace_any<Pol>(m, p, s[l], w, 0, m.nv, lap);
        auto dump = [&](double *dst, const std::vector<cx<A>> &src) {
            if (!dst) return;
            for (size_t e = 0; e < src.size(); e++) {
                dst[2 * (l * src.size() + e)] = (double)src[e].real();
                dst[2 * (l * src.size() + e) + 1] = (double)src[e].imag();
            }
        };
        dump(spec_U, lap.U), dump(spec_P, lap.P), dump(spec_Bt, lap.Bt), dump(spec_Ba, lap.Ba);
        if (w.G) transform_accumulate<A>(L, (int)l, lap.U.data(), n3 * n3, WG.data());
        if (w.H) transform_accumulate<A>(L, (int)l, lap.P.data(), n3 * n3, WH.data());
        if (w.Bt) transform_accumulate<A>(L, (int)l, lap.Bt.data(), n3 * 3, WBt.data());
        if (w.Ba) transform_accumulate<A>(L, (int)l, lap.Ba.data(), n3 * 9, WBa.data());
    }
    if (w.G) transform_scale<A>(L, p.R, n3 * n3, WG.data());
    if (w.H) transform_scale<A>(L, p.R, n3 * n3, WH.data());
    if (w.Bt) transform_scale<A>(L, p.R, n3 * 3, WBt.data());
    if (w.Ba) transform_scale<A>(L, p.R, n3 * 9, WBa.data());
    // H_0 += C (H.cu:184-193)
    if (w.H) {
        if (p.galerkin) {
            std::vector<double> H0(n3 * n3, 0.0);
            C_galerkin_add(m.V, m.nv, m.F, m.nf, H0.data());
            for (size_t e = 0; e < n3 * n3; e++) WH[e] += (A)H0[e];
        } else {
            std::vector<double> C(n3 * 3);
            if (m.vmap)
                C_collocation_split(m.V, m.nv, m.F, m.nf, m.vmap, p.nu, C.data());
            else
                C_collocation(m.V, m.nv, m.F, m.nf, p.nu, C.data());
            for (int v = 0; v < m.nv; v++)
                for (int a = 0; a < 3; a++)
                    for (int b = 0; b < 3; b++) WH[(size_t)(3 * v + a) * n3 + 3 * v + b] += (A)C[(3 * v + a) * 3 + b];
        }
    }
    // B_angular += B_trans (x) (x_v - cm) with cm = 0 after centering (B_angular.cu:167-175)
    if (w.Ba)
        for (int n = 0; n < L; n++)
            for (int v = 0; v < m.nv; v++)
                for (int a = 0; a < 3; a++)
                    for (int j = 0; j < 3; j++)
                        for (int k = 0; k < 3; k++)
                            WBa[(size_t)n * n3 * 9 + (3 * v + a) * 9 + j * 3 + k] +=
                                WBt[(size_t)n * n3 * 3 + (3 * v + a) * 3 + j] * (A)m.V[3 * v + k];
    auto out = [&](double *dst, const std::vector<A> &src) {
        if (dst)
            for (size_t e = 0; e < src.size(); e++) dst[e] = (double)src[e];
    };
    out(G, WG), out(H, WH), out(Bt, WBt), out(Ba, WBa);
    // B_n = [B_trans | B_euler], B_euler(i,j) = Ba(i,(j+2)%3,(j+1)%3) - Ba(i,(j+1)%3,(j+2)%3)  (cpp:280-305,168-172)
    if (B6)
        for (int n = 0; n < L; n++)
            for (size_t r = 0; r < n3; r++) {
                for (int j = 0; j < 3; j++) B6[((size_t)n * n3 + r) * 6 + j] = (double)WBt[(size_t)n * n3 * 3 + r * 3 + j];
                for (int j = 0; j < 3; j++) {
                    const A *b = &WBa[(size_t)n * n3 * 9 + r * 9];
                    B6[((size_t)n * n3 + r) * 6 + 3 + j] = (double)(b[((j + 2) % 3) * 3 + (j + 1) % 3] - b[((j + 1) % 3) * 3 + (j + 2) % 3]);
                }
            }
    return 0;
}

Mesh make_mesh(const double *V, int nv, const int *F, int nf, const int *vmap) {
    Mesh m{nv, nf, V, F, std::vector<double>(3 * (size_t)nf), vmap};
    face_normals(V, F, nf, m.N.data());
    return m;
}

int check_params(const orc_params *p, int nv, int nf) {
    if (!p || nv <= 0 || nf <= 0) return -1;
    if (p->order < 1 || p->order > 5 || p->subdiv < 1 || p->L < 2 || p->R <= 0 || p->dt <= 0) return -2;
    if (p->bdf != 1 && p->bdf != 2) return -3;
    if (p->mode < 0 || p->mode > 2) return -4;
    return 0;
}

void set_threads(const orc_params *p) {
#ifdef _OPENMP
    if (p->nthreads > 0) omp_set_num_threads(p->nthreads);
#endif
}

}  // namespace

// ================================================================ C API ====
extern "C" {

int orc_frequencies(int L, double R, double dt, int bdf, double *s_out) {
    auto s = cq_frequencies(L, R, dt, bdf);
    for (size_t l = 0; l < s.size(); l++) s_out[2 * l] = s[l].real(), s_out[2 * l + 1] = s[l].imag();
    return (int)s.size();
}

int orc_rule_tri(int order, double *bary, double *w) {
    TriRule r = tri_rule(order);
    for (int k = 0; k < r.n; k++) {
        for (int j = 0; j < 3; j++) bary[3 * k + j] = (double)r.c[k][j];
        w[k] = (double)r.w[k];
    }
    return r.n;
}

int orc_rule_tri_composite(int order, int sub, double *bary, double *w) {
    RefPts r = composite_tri(order, sub);
    for (size_t k = 0; k < r.w.size(); k++) {
        for (int j = 0; j < 3; j++) bary[3 * k + j] = (double)r.b[k][j];
        w[k] = (double)r.w[k];
    }
    return (int)r.w.size();
}

int orc_rule_gauss(int order, double *x, double *w) {
    LineRule g = gauss_rule(order);
    for (int k = 0; k < g.n; k++) x[k] = (double)g.x[k], w[k] = (double)g.w[k];
    return g.n;
}

// Kernel values at one point pair.  which: 0 u_hat (9), 1 p_hat (9), 2 b_hat (9), 3 b_hat_ijk (27, index i*9+j*3+k)
// mode: 0 double, 2 long double.  out: complex interleaved.
int orc_kernel(int which, int mode, const double *x, const double *y, const double *n, double c1, double c2, double rho,
               double s_re, double s_im, double *out) {
    auto run = [&](auto tag) {
        using T = decltype(tag);
        const v3<T> rv = cast3<T>(y) - cast3<T>(x);
        const T r = norm(rv);
        const v3<T> rh = unit(rv), nn = cast3<T>(n);
        KParam<T> k{(T)c1, (T)c2, (T)rho, cx<T>((T)s_re, (T)s_im)};
        auto put = [&](int e, cx<T> v) { out[2 * e] = (double)v.real(), out[2 * e + 1] = (double)v.imag(); };
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) {
                if (which == 0) put(a * 3 + b, k_u_r(rh, r, a, b, k) / r);
                if (which == 1) put(a * 3 + b, k_p(rh, r, nn, a, b, k));
                if (which == 2) put(a * 3 + b, k_bt_r(rh, r, nn, a, b, k) / r);
                if (which == 3)
                    for (int c = 0; c < 3; c++) put(a * 9 + b * 3 + c, k_ba_r(rh, r, nn, a, b, c, k) / r);
            }
    };
    if (mode == 2)
        run((long double)0);
    else if (mode == 1)
        run(0.0f);
    else
        run(0.0);
    return 0;
}

// Integral over triangle (y1,y2,y3) of 1/|y - y1| with the weakly singular polar rule (test hook).
double orc_polar_inv_r(const double *y1, const double *y2, const double *y3, int order, int sub) {
    double acc = 0.0;
    polar_rule<double>(cast3<double>(y1), cast3<double>(y2), cast3<double>(y3), QParams{order, sub},
                       [&](const v3<double> &, const v3<double> &, double, double w) { acc += w; });
    return acc;
}

// Integral of f(y) = 1 + a.y + (b.y)^2 over a triangle with the far-field rule (test hook).
double orc_far_poly(const double *y1, const double *y2, const double *y3, const double *a, const double *b, int order, int sub) {
    double acc = 0.0;
    far_rule<double>(cast3<double>(y1), cast3<double>(y2), cast3<double>(y3), QParams{order, sub},
                     [&](const v3<double> &y, const v3<double> &, double w) {
                         double bb = dot(cast3<double>(b), y);
                         acc += w * (1.0 + dot(cast3<double>(a), y) + bb * bb);
                     });
    return acc;
}

int orc_C_collocation(const double *V, int nv, const int *F, int nf, double nu, double *C) {
    C_collocation(V, nv, F, nf, nu, C);
    return 0;
}

int orc_C_collocation_split(const double *V, int nv, const int *F, int nf, const int *vmap, double nu, double *C) {
    C_collocation_split(V, nv, F, nf, vmap, nu, C);
    return 0;
}

int orc_C_galerkin(const double *V, int nv, const int *F, int nf, double *H0) {
    C_galerkin_add(V, nv, F, nf, H0);
    return 0;
}

int orc_transform(const double *spec, long long M, int L, double R, double *W) {
    const int Lf = L / 2 + 1;
    std::vector<double> w((size_t)L * M, 0.0);
    for (int l = 0; l < Lf; l++)
        transform_accumulate<double>(L, l, reinterpret_cast<const cx<double> *>(spec) + (size_t)l * M, (size_t)M, w.data());
    transform_scale<double>(L, R, (size_t)M, w.data());
    std::copy(w.begin(), w.end(), W);
    return 0;
}

// Laplace-domain matrices for one complex frequency s, complete in the rows of vertices [row0,row1)
// (collocation: those rows only, the others left untouched; Galerkin: the test faces touching the range,
// the other rows then hold partial sums).  Outputs complex interleaved, full size (3nv x 3nv, 3nv x 3, 3nv x 9).
int orc_laplace(const double *V, int nv, const int *F, int nf, const int *vmap, const orc_params *p, double s_re, double s_im,
                int mask, int row0, int row1, double *U, double *P, double *Bt, double *Ba) {
    int rc = check_params(p, nv, nf);
    if (rc) return rc;
    set_threads(p);
    Mesh m = make_mesh(V, nv, F, nf, vmap);
    Which w = which_from(mask);
    if (row1 <= row0 || row0 < 0 || row1 > nv) row0 = 0, row1 = nv;
    auto run = [&](auto pol) {
        using Pol = decltype(pol);
        using A = typename Pol::Acc;
        Laplace<A> lap;
        lap.alloc(nv, w);
        laplace_any<Pol>(m, *p, cx<double>(s_re, s_im), w, row0, row1, lap);
        auto dump = [&](double *dst, const std::vector<cx<A>> &src) {
            if (!dst) return;
            for (size_t e = 0; e < src.size(); e++) dst[2 * e] = (double)src[e].real(), dst[2 * e + 1] = (double)src[e].imag();
        };
        dump(U, lap.U), dump(P, lap.P), dump(Bt, lap.Bt), dump(Ba, lap.Ba);
    };
    if (p->mode == 1)
        run(PolRefMixed{});
    else if (p->mode == 2)
        run(PolLD{});
    else
        run(PolFP64{});
    return 0;
}

// Full precompute: time-domain weights for n = 0..L-1.
//   G, H: L x 3nv x 3nv (H_0 includes C);  Bt: L x 3nv x 3;  Ba: L x 3nv x 9 (with the
//   B_trans (x) x correction);  B6: L x 3nv x 6 ([B_trans | B_euler]).
//   spec_*: optional Laplace-domain dumps, (L/2+1) x size, complex interleaved.
int orc_assemble(const double *V, int nv, const int *F, int nf, const int *vmap, const orc_params *p, int mask, double *G,
                 double *H, double *Bt, double *Ba, double *B6, double *spec_U, double *spec_P, double *spec_Bt,
                 double *spec_Ba) {
    int rc = check_params(p, nv, nf);
    if (rc) return rc;
    set_threads(p);
    Mesh m = make_mesh(V, nv, F, nf, vmap);
    if (G) mask |= 1;
    if (H) mask |= 2;
    if (Bt) mask |= 4;
    if (Ba) mask |= 8;
    if (p->mode == 1) return assemble_impl<PolRefMixed>(m, *p, mask, G, H, Bt, Ba, B6, spec_U, spec_P, spec_Bt, spec_Ba);
    if (p->mode == 2) return assemble_impl<PolLD>(m, *p, mask, G, H, Bt, Ba, B6, spec_U, spec_P, spec_Bt, spec_Ba);
    return assemble_impl<PolFP64>(m, *p, mask, G, H, Bt, Ba, B6, spec_U, spec_P, spec_Bt, spec_Ba);
}

}  // extern "C"

// Parity pinning (what checks each function, see tests/test_oracle_*.py):
//   tri_rule / composite_tri      polynomial exactness up to the rule order (test_oracle_quadrature)
//   gauss_rule                    numpy.polynomial.legendre.leggauss
//   far_rule                      exact integral of quadratic polynomials over a triangle
//   polar_rule                    closed-form vertex potential of a triangle (int 1/r)
//   guiggiani_rule                static rigid-body identity sum_j P(s->0)_ij + C_i = 0
//   k_u_r, k_p, k_bt_r, k_ba_r    static limits (Kelvin solution), symmetry, conjugation,
//                                 Laplace-domain Navier residual, traction = sigma(u).n
//   galerkin_* rules              static-limit Galerkin identities and symmetry on tiny meshes
//   cq_frequencies, transform     known CQ weights of exp(-s T), 1/s, 1/s^2 (BDF1, BDF2)
//   C_collocation                 sum of jump terms; 1/2 I on flat interior vertices
//   C_collocation_split           rigid-body identity on the vertex-split cube (test_oracle_identities)
//   C_galerkin_add                total mass = surface area per component
