// Quadrature tables, CQ frequency set, transform table and jump (C) matrices,
// built on the host in FP64 and uploaded once per job.
#pragma once
#include <algorithm>
#include <array>
#include <cmath>
#include <complex>
#include <map>
#include <utility>
#include <vector>

namespace cqbem {

constexpr double kPi = 3.14159265358979323846;

// Symmetric triangle rule of polynomial order 1..5 (Cowper 1973), as barycentric
// coordinates and weights summing to 1 (integrator.hpp:55-113).
struct TriPoint {
    double b[3];
    double w;
};

inline std::vector<TriPoint> tri_rule(int order) {
    std::vector<TriPoint> r;
    auto orbit = [&](double a, double w) {
        const double c = 1.0 - 2.0 * a;
        r.push_back({{c, a, a}, w});
        r.push_back({{a, c, a}, w});
        r.push_back({{a, a, c}, w});
    };
    const double t = 1.0 / 3.0;
    switch (order) {
    case 1: r.push_back({{t, t, t}, 1.0}); break;
    case 2: orbit(0.5, t); break;
    case 3:
        r.push_back({{t, t, t}, -27.0 / 48.0});
        orbit(2.0 / 15.0, 25.0 / 48.0);
        break;
    case 4:
        orbit(0.445948490915965, 2 * 0.111690794839005);
        orbit(0.091576213509771, 2 * 0.054975871827661);
        break;
    case 5:
        r.push_back({{t, t, t}, 9.0 / 40.0});
        orbit(0.470142064105115, 2 * 0.0661970763942530);
        orbit(0.101286507323456, 2 * 0.0629695902724135);
        break;
    default: break;
    }
    return r;
}

// Uniform m x m subdivision composed with the triangle rule: points in the
// parent's barycentric coordinates, weights summing to 1
// (integrate_subdivision_triangle, integrator.hpp:24-51).
inline std::vector<TriPoint> composite_tri_rule(int order, int m) {
    std::vector<TriPoint> base = tri_rule(order), out;
    auto corner = [&](int a, int c, double *b) {
        b[1] = (double)a / m, b[2] = (double)c / m, b[0] = 1.0 - b[1] - b[2];
    };
    auto emit = [&](const double B[3][3]) {
        for (const TriPoint &p : base) {
            TriPoint q{{0, 0, 0}, p.w / ((double)m * m)};
            for (int k = 0; k < 3; k++)
                for (int j = 0; j < 3; j++) q.b[j] += p.b[k] * B[k][j];
            out.push_back(q);
        }
    };
    for (int i = 0; i < m; i++)
        for (int j = 0; i + j < m; j++) {
            double B[3][3];
            corner(i, j, B[0]), corner(i + 1, j, B[1]), corner(i, j + 1, B[2]);
            emit(B);
            if (i + j < m - 1) {
                corner(i + 1, j + 1, B[0]), corner(i + 1, j, B[1]), corner(i, j + 1, B[2]);
                emit(B);
            }
        }
    return out;
}

// Composite Gauss-Legendre rule on [0,1]: `m` pieces of `order` nodes
// (integrate_subdivision_1d + gaussian_quadrature_1d, integrator.hpp:15-22,116-150).
inline std::vector<std::pair<double, double>> composite_line01(int order, int m) {
    std::vector<double> x, w;
    switch (order) {
    case 1: x = {0.0}, w = {2.0}; break;
    case 2: {
        const double a = 1.0 / std::sqrt(3.0);
        x = {-a, a}, w = {1.0, 1.0};
        break;
    }
    case 3: {
        const double a = std::sqrt(0.6);
        x = {-a, 0.0, a}, w = {5.0 / 9.0, 8.0 / 9.0, 5.0 / 9.0};
        break;
    }
    case 4: {
        const double t = 2.0 / 7.0 * std::sqrt(1.2);
        const double a = std::sqrt(3.0 / 7.0 - t), b = std::sqrt(3.0 / 7.0 + t);
        const double wa = (18.0 + std::sqrt(30.0)) / 36.0, wb = (18.0 - std::sqrt(30.0)) / 36.0;
        x = {-b, -a, a, b}, w = {wb, wa, wa, wb};
        break;
    }
    case 5: {
        const double t = 2.0 * std::sqrt(10.0 / 7.0);
        const double a = std::sqrt(5.0 - t) / 3.0, b = std::sqrt(5.0 + t) / 3.0;
        const double wa = (322.0 + 13.0 * std::sqrt(70.0)) / 900.0, wb = (322.0 - 13.0 * std::sqrt(70.0)) / 900.0;
        x = {-b, -a, 0.0, a, b}, w = {wb, wa, 128.0 / 225.0, wa, wb};
        break;
    }
    default: break;
    }
    std::vector<std::pair<double, double>> out;
    for (int k = 0; k < m; k++) {
        const double lo = (double)k / m, hi = (double)(k + 1) / m;
        for (size_t q = 0; q < x.size(); q++)
            out.push_back({(hi - lo) * x[q] / 2.0 + (lo + hi) / 2.0, w[q] * (hi - lo) / 2.0});
    }
    return out;
}

// CQ frequencies s_l = gamma(R zeta^l)/dt, l = 0..L/2, zeta = exp(2 pi i / L)
// (G.cu:19-24,65); zeta^l evaluated as cis(2 pi l / L).
inline std::vector<std::complex<double>> cq_frequencies(int L, double R, double dt, int bdf) {
#pragma clang fp contract(off)
    std::vector<std::complex<double>> s;
    for (int l = 0; l <= L / 2; l++) {
        const double ang = 2.0 * kPi * (double)l / (double)L;
        const std::complex<double> z = R * std::complex<double>(std::cos(ang), std::sin(ang));
        const std::complex<double> g = (bdf == 1) ? (1.0 - z) : (1.5 - 2.0 * z + z * z / 2.0);
        s.push_back(g / dt);
    }
    return s;
}

// Inverse-transform table T[n][l] = R^-n / L * mult_l * zeta^(-n l), mult_l = 1 for
// l = 0 and for l = L/2 with L even, 2 otherwise (the conjugate partner of
// G.cu:173-174 folded in); (n l) reduced mod L before the twiddle is formed.
inline std::vector<std::complex<double>> transform_table(int L, double R) {
#pragma clang fp contract(off)
    const int Lf = L / 2 + 1;
    std::vector<std::complex<double>> t((size_t)L * Lf);
    for (int n = 0; n < L; n++) {
        const double sc = std::pow(R, -(double)n) / (double)L;
        for (int l = 0; l < Lf; l++) {
            const bool pair = (l != 0) && (l != L / 2 || L % 2 == 1);
            const long long k = ((long long)n * l) % L;
            const double ang = -2.0 * kPi * (double)k / (double)L;
            t[(size_t)n * Lf + l] = (pair ? 2.0 : 1.0) * sc * std::complex<double>(std::cos(ang), std::sin(ang));
        }
    }
    return t;
}

struct V3 {
    double x, y, z;
};
inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 crs(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline double dt3(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline double nrm(V3 a) { return std::sqrt(dt3(a, a)); }
inline V3 unit(V3 a) {
    const double n = nrm(a);
    return n > 0 ? V3{a.x / n, a.y / n, a.z / n} : a;
}
inline V3 vtx(const double *V, int i) { return {V[3 * i], V[3 * i + 1], V[3 * i + 2]}; }

// Collocation jump term C_i (3x3 per vertex), compute_C_matrix_collocation
// (elastics_bem_matrix_collection.cpp:387-445): 1/2 I, plus per interior edge
// (1/4pi) sgn((n1 x n2).r) acos(n1.n2) I on both end vertices, minus per face corner
// 1/(8 pi (1-nu)) ((r2 - r1) x n) n^T.
inline std::vector<double> c_matrix_collocation(const double *V, int nv, const int *F, int nf, const double *N, double nu) {
    std::vector<double> C(9 * (size_t)nv, 0.0);
    for (int i = 0; i < nv; i++)
        for (int a = 0; a < 3; a++) C[9 * i + 4 * a] += 0.5;
    std::map<std::pair<int, int>, std::vector<int>> edges;  // half-edges c*nf+f, edge opposite corner c
    for (int c = 0; c < 3; c++)
        for (int f = 0; f < nf; f++) {
            const int a = F[3 * f + (c + 1) % 3], b = F[3 * f + (c + 2) % 3];
            edges[{std::min(a, b), std::max(a, b)}].push_back(c * nf + f);
        }
    for (auto &kv : edges) {
        if (kv.second.size() < 2) continue;
        const int e1 = kv.second[0], e2 = kv.second[1];  // ascending by construction
        const int f1 = e1 % nf, f2 = e2 % nf, c1 = e1 / nf;
        const int a = F[3 * f1 + (c1 + 1) % 3], b = F[3 * f1 + (c1 + 2) % 3];
        const V3 n1 = vtx(N, f1), n2 = vtx(N, f2);
        const V3 r = unit(sub(vtx(V, a), vtx(V, b)));
        const double x = dt3(crs(n1, n2), r);
        const double sg = std::abs(x) < 1e-12 ? 0.0 : (x > 0 ? 1.0 : -1.0);
        const double w = sg * std::acos(std::min(1.0, std::max(-1.0, dt3(n1, n2)))) / (4.0 * kPi);
        for (int k = 0; k < 3; k++) C[9 * a + 4 * k] += w, C[9 * b + 4 * k] += w;
    }
    const double kf = 1.0 / (8.0 * kPi * (1.0 - nu));
    for (int f = 0; f < nf; f++) {
        const V3 n = vtx(N, f);
        const double nn[3] = {n.x, n.y, n.z};
        for (int i = 0; i < 3; i++) {
            const V3 p = vtx(V, F[3 * f + i]);
            const V3 r1 = unit(sub(vtx(V, F[3 * f + (i + 1) % 3]), p)), r2 = unit(sub(vtx(V, F[3 * f + (i + 2) % 3]), p));
            const V3 t = crs(sub(r2, r1), n);
            const double tt[3] = {t.x, t.y, t.z};
            for (int a = 0; a < 3; a++)
                for (int b = 0; b < 3; b++) C[9 * F[3 * f + i] + 3 * a + b] -= kf * tt[a] * nn[b];
        }
    }
    return C;
}

// Jump term of a vertex-split mesh: evaluated on the original (unsplit) mesh and copied to every copy of a
// vertex (elastics_bem_matrix_collection.cpp:392-398,436-442).  The original mesh is recovered from the
// map: original vertex vmap[v] sits at V[v] (copies share the position) and face f reads vmap[F[f]].
inline std::vector<double> c_matrix_collocation_split(const double *V, int nv, const int *F, int nf, const double *N,
                                                      const int *vmap, double nu) {
    int no = 0;
    for (int v = 0; v < nv; v++) no = std::max(no, vmap[v] + 1);
    std::vector<double> Vo(3 * (size_t)no, 0.0);
    for (int v = 0; v < nv; v++)
        for (int k = 0; k < 3; k++) Vo[3 * (size_t)vmap[v] + k] = V[3 * v + k];
    std::vector<int> Fo(3 * (size_t)nf);
    for (size_t k = 0; k < Fo.size(); k++) Fo[k] = vmap[F[k]];
    const std::vector<double> Co = c_matrix_collocation(Vo.data(), no, Fo.data(), nf, N, nu);
    std::vector<double> C(9 * (size_t)nv);
    for (int v = 0; v < nv; v++)
        for (int e = 0; e < 9; e++) C[9 * (size_t)v + e] = Co[9 * (size_t)vmap[v] + e];
    return C;
}

}  // namespace cqbem
