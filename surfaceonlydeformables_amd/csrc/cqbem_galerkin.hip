// Galerkin assembly on gfx950.
//
// Regular (non-touching) face pairs, G and H: one lane per (test face f1, source
// vertex j, frequency l).  The lane walks f1's test points and, for each, j's
// vertex star with j's basis function (integrate_galerkin_nonsingular,
// integrator_galerkin.hpp:214-260, restricted to one source corner), then
// projects onto f1's three test basis functions.  Test faces are launched by
// vertex-disjoint colour classes, so the three row blocks a lane adds into are
// owned by that lane alone within a launch: plain read-modify-write, no atomics,
// a fixed summation order.
//
// Regular pairs, B_trans / B_angular: one 256-lane block per (test face, l); lanes
// stride over source faces per test point, a block reduction per test point,
// test-basis projection kept in LDS.
//
// Touching pairs (Sauter-Schwab, integrator_galerkin.hpp:42-212): one 64-lane
// block per (pair, l); lanes split the 4D tensor rule, reduce, and write the
// pair's corner blocks to a per-pair scratch.  A gather kernel then owns each
// output block and adds its pairs' contributions in a fixed order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "cqbem_galerkin.hpp"
#include "cqbem_pool.hpp"

namespace cqbem {

namespace {

__device__ inline bool same_v(const Geom &g, int a, int b) { return a == b || (g.vmap != nullptr && g.vmap[a] == g.vmap[b]); }

__device__ inline bool faces_touch(const Geom &g, const int *f1v, int f2) {
    const int b0 = g.F[3 * f2], b1 = g.F[3 * f2 + 1], b2 = g.F[3 * f2 + 2];
    for (int k = 0; k < 3; k++)
        if (same_v(g, f1v[k], b0) || same_v(g, f1v[k], b1) || same_v(g, f1v[k], b2)) return true;
    return false;
}

// ---------------------------------------------------------------------------
// Regular G/H: grid (ceil(nv/64), nl, faces in colour)
// ---------------------------------------------------------------------------
template <bool DG, bool DH>
__global__ void __launch_bounds__(64) k_gal_gh(Geom g, const double4 *__restrict__ quad, int npts,
                                                const Freq *__restrict__ freqs, Mat m, const int *__restrict__ faces,
                                                cd *__restrict__ UG, cd *__restrict__ UH, long long ld, long long lstride) {
    const int j = blockIdx.x * 64 + threadIdx.x;
    const int l = blockIdx.y;
    const int f1 = faces[blockIdx.z];
    if (j >= g.nv) return;
    const Freq f = freqs[l];
    const int tv[3] = {g.F[3 * f1], g.F[3 * f1 + 1], g.F[3 * f1 + 2]};
    const double *X0 = g.V + 3 * tv[0], *X1 = g.V + 3 * tv[1], *X2 = g.V + 3 * tv[2];
    const double tarea = g.area[f1];
    cd ga[3][6], gb[3], hh[3][9];
    for (int k = 0; k < 3; k++) {
        gb[k] = cd{0, 0};
        for (int e = 0; e < 6; e++) ga[k][e] = cd{0, 0};
        for (int e = 0; e < 9; e++) hh[k][e] = cd{0, 0};
    }
    const int t0 = g.vs_ptr[j], t1 = g.vs_ptr[j + 1];
    for (int a = 0; a < npts; ++a) {
        const double4 Qa = quad[a];
        const double x0 = Qa.x * X0[0] + Qa.y * X1[0] + Qa.z * X2[0];
        const double x1 = Qa.x * X0[1] + Qa.y * X1[1] + Qa.z * X2[1];
        const double x2 = Qa.x * X0[2] + Qa.y * X1[2] + Qa.z * X2[2];
        const double wa = Qa.w * tarea;
        cd pa[6], pb = {0, 0}, ph[9];
        for (int e = 0; e < 6; e++) pa[e] = cd{0, 0};
        for (int e = 0; e < 9; e++) ph[e] = cd{0, 0};
        for (int t = t0; t < t1; ++t) {
            const int pk = g.vs_face[t];
            const int fc = pk >> 2, c = pk & 3;
            if (faces_touch(g, tv, fc)) continue;
            const int a0 = g.F[3 * fc], a1 = g.F[3 * fc + 1], a2 = g.F[3 * fc + 2];
            const double *p0 = g.V + 3 * a0, *p1 = g.V + 3 * a1, *p2 = g.V + 3 * a2;
            const double n0 = g.N[3 * fc], n1 = g.N[3 * fc + 1], n2 = g.N[3 * fc + 2];
            const double ar = g.area[fc];
            cd v1[3], v2[3], s2 = {0, 0}, m4[6];
            if (DH) {
                for (int k = 0; k < 3; k++) v1[k] = v2[k] = cd{0, 0};
                for (int k = 0; k < 6; k++) m4[k] = cd{0, 0};
            }
            for (int q = 0; q < npts; ++q) {
                const double4 Q = quad[q];
                const double r0 = Q.x * p0[0] + Q.y * p1[0] + Q.z * p2[0] - x0;
                const double r1 = Q.x * p0[1] + Q.y * p1[1] + Q.z * p2[1] - x1;
                const double r2 = Q.x * p0[2] + Q.y * p1[2] + Q.z * p2[2] - x2;
                const double d2 = r0 * r0 + r1 * r1 + r2 * r2;
                const double ir = rsqrt_nr(d2), r = d2 * ir, ir2 = ir * ir;
                const double h0 = r0 * ir, h1 = r1 * ir, h2 = r2 * ir;
                const double bc = c == 0 ? Q.x : (c == 1 ? Q.y : Q.z);
                const double wL = Q.w * ar * bc;
                const Expo E = expo(r, f);
                if (DG) {
                    cd al, be;
                    g_coef(E, ir, ir2, m, al, be);
                    const double wr = wL * ir;
                    al = al * wr;
                    fma_cr(pb, be, wr);
                    fma_cr(pa[0], al, h0 * h0);
                    fma_cr(pa[1], al, h1 * h1);
                    fma_cr(pa[2], al, h2 * h2);
                    fma_cr(pa[3], al, h0 * h1);
                    fma_cr(pa[4], al, h0 * h2);
                    fma_cr(pa[5], al, h1 * h2);
                }
                if (DH) {
                    cd c1, c2, c4;
                    h_coef(E, ir, ir2, f, m, c1, c2, c4);
                    const double rn = h0 * n0 + h1 * n1 + h2 * n2;
                    c1 = c1 * wL;
                    c2 = c2 * wL;
                    c4 = c4 * (wL * rn);
                    fma_cr(v1[0], c1, h0), fma_cr(v1[1], c1, h1), fma_cr(v1[2], c1, h2);
                    fma_cr(v2[0], c2, h0), fma_cr(v2[1], c2, h1), fma_cr(v2[2], c2, h2);
                    fma_cr(s2, c2, rn);
                    fma_cr(m4[0], c4, h0 * h0);
                    fma_cr(m4[1], c4, h1 * h1);
                    fma_cr(m4[2], c4, h2 * h2);
                    fma_cr(m4[3], c4, h0 * h1);
                    fma_cr(m4[4], c4, h0 * h2);
                    fma_cr(m4[5], c4, h1 * h2);
                }
            }
            if (DH) {
                const double n[3] = {n0, n1, n2};
                for (int p = 0; p < 3; p++)
                    for (int b = 0; b < 3; b++) {
                        cd v = m4[sidx(p, b)];
                        fma_cr(v, v1[p], n[b]);
                        fma_cr(v, v2[b], n[p]);
                        if (p == b) v += s2;
                        ph[p * 3 + b] += v;
                    }
            }
        }
        const double wk[3] = {wa * Qa.x, wa * Qa.y, wa * Qa.z};
        for (int k = 0; k < 3; k++) {
            if (DG) {
                fma_cr(gb[k], pb, wk[k]);
                for (int e = 0; e < 6; e++) fma_cr(ga[k][e], pa[e], wk[k]);
            }
            if (DH)
                for (int e = 0; e < 9; e++) fma_cr(hh[k][e], ph[e], wk[k]);
        }
    }
    for (int k = 0; k < 3; k++) {
        const long long base = (long long)l * lstride + (long long)(3 * tv[k]) * ld + 3 * (long long)j;
        for (int p = 0; p < 3; p++)
            for (int b = 0; b < 3; b++) {
                if (DG) {
                    cd v = ga[k][sidx(p, b)];
                    if (p == b) v += gb[k];
                    UG[base + p * ld + b] += v;
                }
                if (DH) UH[base + p * ld + b] += hh[k][p * 3 + b] * kInv4Pi;
            }
    }
}

__device__ inline double block_sum256(double v, double *red) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[w] = v;
    __syncthreads();
    const double s = red[0] + red[1] + red[2] + red[3];
    __syncthreads();
    return s;
}

// ---------------------------------------------------------------------------
// Regular B_trans / B_angular: grid (nl, faces in colour), 256 lanes.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_gal_b(Geom g, const double4 *__restrict__ quad, int npts,
                                                const Freq *__restrict__ freqs, Mat m, const int *__restrict__ faces,
                                                cd *__restrict__ Bt, cd *__restrict__ Ba, long long nv3) {
    __shared__ double red[4];
    __shared__ double corner[3][24][2];
    const int l = blockIdx.x;
    const int f1 = faces[blockIdx.y];
    const Freq f = freqs[l];
    const int tv[3] = {g.F[3 * f1], g.F[3 * f1 + 1], g.F[3 * f1 + 2]};
    const double *X0 = g.V + 3 * tv[0], *X1 = g.V + 3 * tv[1], *X2 = g.V + 3 * tv[2];
    const double tarea = g.area[f1];
    if (threadIdx.x < 3 * 24 * 2) (&corner[0][0][0])[threadIdx.x] = 0.0;
    __syncthreads();
    for (int a = 0; a < npts; ++a) {
        const double4 Qa = quad[a];
        const double x0 = Qa.x * X0[0] + Qa.y * X1[0] + Qa.z * X2[0];
        const double x1 = Qa.x * X0[1] + Qa.y * X1[1] + Qa.z * X2[1];
        const double x2 = Qa.x * X0[2] + Qa.y * X1[2] + Qa.z * X2[2];
        cd bt[6], ba[18];
        for (int k = 0; k < 6; k++) bt[k] = cd{0, 0};
        for (int k = 0; k < 18; k++) ba[k] = cd{0, 0};
        for (int fc = threadIdx.x; fc < g.nf; fc += 256) {
            if (faces_touch(g, tv, fc)) continue;
            const int a0 = g.F[3 * fc], a1 = g.F[3 * fc + 1], a2 = g.F[3 * fc + 2];
            const double *p0 = g.V + 3 * a0, *p1 = g.V + 3 * a1, *p2 = g.V + 3 * a2;
            const double n[3] = {g.N[3 * fc], g.N[3 * fc + 1], g.N[3 * fc + 2]};
            const double ar = g.area[fc];
            cd mt[3] = {{0, 0}, {0, 0}, {0, 0}}, m2 = {0, 0}, M[6], S = {0, 0}, S3 = {0, 0};
            for (int k = 0; k < 6; k++) M[k] = cd{0, 0};
            for (int q = 0; q < npts; ++q) {
                const double4 Q = quad[q];
                const double r0 = Q.x * p0[0] + Q.y * p1[0] + Q.z * p2[0] - x0;
                const double r1 = Q.x * p0[1] + Q.y * p1[1] + Q.z * p2[1] - x1;
                const double r2 = Q.x * p0[2] + Q.y * p1[2] + Q.z * p2[2] - x2;
                const double d2 = r0 * r0 + r1 * r1 + r2 * r2;
                const double ir = rsqrt_nr(d2), r = d2 * ir, ir2 = ir * ir;
                const double h[3] = {r0 * ir, r1 * ir, r2 * ir};
                const double w = Q.w * ar;
                const Expo E = expo(r, f);
                cd t1, t2, u1, u12, u3;
                b_coef(E, ir, ir2, f, m, t1, t2, u1, u12, u3);
                const double rn = h[0] * n[0] + h[1] * n[1] + h[2] * n[2];
                t1 = t1 * w;
                u1 = u1 * w;
                for (int p = 0; p < 3; p++) fma_cr(mt[p], t1, h[p]);
                fma_cr(m2, t2, w * rn);
                for (int p = 0; p < 3; p++)
                    for (int b = p; b < 3; b++) fma_cr(M[sidx(p, b)], u1, h[p] * h[b]);
                fma_cr(S, u12, w);
                fma_cr(S3, u3, w);
            }
            for (int p = 0; p < 3; p++)
                for (int b = p; b < 3; b++) {
                    cd v = {0, 0};
                    fma_cr(v, mt[p], n[b]);
                    fma_cr(v, mt[b], n[p]);
                    if (p == b) v += m2;
                    bt[sidx(p, b)] += v;
                    for (int k = 0; k < 3; k++) {
                        cd u = {0, 0};
                        fma_cr(u, M[sidx(p, k)], n[b]);
                        fma_cr(u, M[sidx(b, k)], n[p]);
                        fma_cr(u, S, (b == k ? n[p] : 0.0) + (p == k ? n[b] : 0.0));
                        if (p == b) fma_cr(u, S3, n[k]);
                        ba[sidx(p, b) * 3 + k] += u;
                    }
                }
        }
        const double wa = Qa.w * tarea;
        const double wk[3] = {wa * Qa.x, wa * Qa.y, wa * Qa.z};
#pragma unroll
        for (int e = 0; e < 24; e++) {
            const cd v = e < 6 ? bt[e] : ba[e - 6];
            const double sx = block_sum256(v.x, red), sy = block_sum256(v.y, red);
            if (threadIdx.x == 0)
                for (int k = 0; k < 3; k++) corner[k][e][0] += wk[k] * sx, corner[k][e][1] += wk[k] * sy;
        }
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        const int k = threadIdx.x;
        cd *bt_out = Bt + (long long)l * nv3 * 3 + (long long)(3 * tv[k]) * 3;
        cd *ba_out = Ba + (long long)l * nv3 * 9 + (long long)(3 * tv[k]) * 9;
        for (int p = 0; p < 3; p++)
            for (int b = 0; b < 3; b++) {
                const int e = sidx(p, b);
                bt_out[p * 3 + b] += cd{corner[k][e][0], corner[k][e][1]};
                for (int q = 0; q < 3; q++) ba_out[p * 9 + b * 3 + q] += cd{corner[k][6 + e * 3 + q][0], corner[k][6 + e * 3 + q][1]};
            }
    }
}

// ---------------------------------------------------------------------------
// Touching pairs, Sauter-Schwab.  Pair record: kind (1 coincident, 2 edge, 3 vertex),
// f2 (source normal), j1[3], j2[3] (re-ordered corner vertices, integrator_galerkin.hpp:262-334).
// ---------------------------------------------------------------------------
struct Pair {
    int kind, f2;
    int j1[3], j2[3];
    int f1;
};

// Output per pair and frequency (complex): G 81 (row 3i+p, col 3j+q), H 81, Bt 27 (row 3i+p, col q), Ba 81 (row 3i+p, col q*3+k)
constexpr int kPairG = 0, kPairH = 81, kPairBt = 162, kPairBa = 189, kPairStride = 270;

// Point-pair kernel accumulation for MODE 0 (G), 1 (H), 2 (B).
template <int MODE> struct SingAcc {
    cd v[MODE == 0 ? 63 : (MODE == 1 ? 81 : 36)];
};

// MIRROR (B of a coincident pair integrated over regions 1, 3, 5): the point also stands for its exchanged image,
// whose test weight is Ly; the odd terms (t1 rh, t2 rh.n) change sign there and the even ones (u1 rh rh, u12, u3) do
// not, so they take the weights wt (Lx - Ly) and wt (Lx + Ly).
template <int MODE, bool MIRROR = false>
__device__ __forceinline__ void sing_eval(const double x[3], const double y[3], const double Lx[3], const double Ly[3], double wt,
                                 const double n[3], const Freq &f, const Mat &m, const MathTables *M, cd *acc) {
    const double r0 = y[0] - x[0], r1 = y[1] - x[1], r2 = y[2] - x[2];
    const double d2 = r0 * r0 + r1 * r1 + r2 * r2;
    const double ir = rsqrt_nr(d2), r = d2 * ir, ir2 = ir * ir;
    const double h[3] = {r0 * ir, r1 * ir, r2 * ir};
    const Expo E = expo(r, f, M);
    if (MODE == 0) {
        cd al, be;
        g_coef(E, ir, ir2, m, al, be);
        al = al * ir;
        be = be * ir;
        const double hh[6] = {h[0] * h[0], h[1] * h[1], h[2] * h[2], h[0] * h[1], h[0] * h[2], h[1] * h[2]};
        #pragma unroll
        for (int i = 0; i < 3; i++)
            #pragma unroll
            for (int j = 0; j < 3; j++) {
                const double w = wt * Lx[i] * Ly[j];
                const cd aw = al * w;
                cd *a = acc + (i * 3 + j) * 7;
                #pragma unroll
                for (int e = 0; e < 6; e++) fma_cr(a[e], aw, hh[e]);
                fma_cr(a[6], be, w);
            }
    } else if (MODE == 1) {
        cd c1, c2, c4;
        h_coef(E, ir, ir2, f, m, c1, c2, c4);
        const double rn = h[0] * n[0] + h[1] * n[1] + h[2] * n[2];
        cd pv[9];
        #pragma unroll
        for (int p = 0; p < 3; p++)
            #pragma unroll
            for (int b = 0; b < 3; b++) {
                cd v = c1 * (h[p] * n[b]);
                fma_cr(v, c2, n[p] * h[b] + (p == b ? rn : 0.0));
                fma_cr(v, c4, rn * h[p] * h[b]);
                pv[p * 3 + b] = v;
            }
        #pragma unroll
        for (int i = 0; i < 3; i++)
            #pragma unroll
            for (int j = 0; j < 3; j++) {
                const double w = wt * Lx[i] * Ly[j];
                cd *a = acc + (i * 3 + j) * 9;
                #pragma unroll
                for (int e = 0; e < 9; e++) fma_cr(a[e], pv[e], w);
            }
    } else {
        cd t1, t2, u1, u12, u3;
        b_coef(E, ir, ir2, f, m, t1, t2, u1, u12, u3);
        const double rn = h[0] * n[0] + h[1] * n[1] + h[2] * n[2];
        #pragma unroll
        for (int i = 0; i < 3; i++) {
            const double wo = MIRROR ? wt * (Lx[i] - Ly[i]) : wt * Lx[i];  // odd terms
            const double we = MIRROR ? wt * (Lx[i] + Ly[i]) : wt * Lx[i];  // even terms
            cd *a = acc + i * 12;
            const cd tw = t1 * wo, uw = u1 * we;
            #pragma unroll
            for (int p = 0; p < 3; p++) fma_cr(a[p], tw, h[p]);
            fma_cr(a[3], t2, wo * rn);
            #pragma unroll
            for (int p = 0; p < 3; p++)
                #pragma unroll
                for (int b = p; b < 3; b++) fma_cr(a[4 + sidx(p, b)], uw, h[p] * h[b]);
            fma_cr(a[10], u12, we);
            fma_cr(a[11], u3, we);
        }
    }
}

// Point r of the Sauter-Schwab region maps (integrator_galerkin.hpp:42-212) at the 4D rule point
// (e1, e2, e3, xi) with rule weight w: test barycentric coordinates (a0, a1), source (b0, b1), weight wt.
template <int KIND>
__device__ __forceinline__ void region_point(int r, double e1, double e2, double e3, double xi, double w, double &a0,
                                             double &a1, double &b0, double &b1, double &wt) {
    if (KIND == 1) {
        wt = w * xi * xi * xi * e1 * e1 * e2;
        switch (r) {
        case 0: a0 = xi, a1 = xi * (1.0 - e1 + e1 * e2), b0 = xi * (1.0 - e1 * e2 * e3), b1 = xi * (1.0 - e1); break;
        case 1: a0 = xi * (1.0 - e1 * e2 * e3), a1 = xi * (1.0 - e1), b0 = xi, b1 = xi * (1.0 - e1 + e1 * e2); break;
        case 2: a0 = xi, a1 = xi * e1 * (1.0 - e2 + e2 * e3), b0 = xi * (1.0 - e1 * e2), b1 = xi * e1 * (1.0 - e2); break;
        case 3: a0 = xi * (1.0 - e1 * e2), a1 = xi * e1 * (1.0 - e2), b0 = xi, b1 = xi * e1 * (1.0 - e2 + e2 * e3); break;
        case 4: a0 = xi * (1.0 - e1 * e2 * e3), a1 = xi * e1 * (1.0 - e2 * e3), b0 = xi, b1 = xi * e1 * (1.0 - e2); break;
        default: a0 = xi, a1 = xi * e1 * (1.0 - e2), b0 = xi * (1.0 - e1 * e2 * e3), b1 = xi * e1 * (1.0 - e2 * e3); break;
        }
    } else if (KIND == 2) {
        const double W = w * xi * xi * xi * e1 * e1;
        wt = r == 0 ? W : W * e2;
        switch (r) {
        case 0: a0 = xi, a1 = xi * e1 * e3, b0 = xi * (1.0 - e1 * e2), b1 = xi * e1 * (1.0 - e2); break;
        case 1: a0 = xi, a1 = xi * e1, b0 = xi * (1.0 - e1 * e2 * e3), b1 = xi * e1 * e2 * (1.0 - e3); break;
        case 2: a0 = xi * (1.0 - e1 * e2), a1 = xi * e1 * (1.0 - e2), b0 = xi, b1 = xi * e1 * e2 * e3; break;
        case 3: a0 = xi * (1.0 - e1 * e2 * e3), a1 = xi * e1 * e2 * (1.0 - e3), b0 = xi, b1 = xi * e1; break;
        default: a0 = xi * (1.0 - e1 * e2 * e3), a1 = xi * e1 * (1.0 - e2 * e3), b0 = xi, b1 = xi * e1 * e2; break;
        }
    } else {
        wt = w * xi * xi * xi * e2;
        if (r == 0) a0 = xi, a1 = xi * e1, b0 = xi * e2, b1 = xi * e2 * e3;
        else a0 = xi * e2, a1 = xi * e2 * e3, b0 = xi, b1 = xi * e1;
    }
}

// HALF (B of coincident pairs): regions 1, 3, 5 only, each point standing for its exchanged image as well.
// B (MODE 2) is held to 2 waves per SIMD: its 36 accumulators and working set take 316 registers at 1 wave; at 2 it
// spills 46-59 of them to scratch and still runs 15% faster.  G and H (the lane variant) keep 1 wave.
template <int MODE, int KIND, bool HALF = false>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(MODE == 2 ? 2 : 1, 2))) k_gal_sing(Geom g, const Pair *__restrict__ pairs, int pair0,
                                                  const double2 *__restrict__ line, int nline,
                                                  const Freq *__restrict__ freqs, Mat m, cd *__restrict__ out) {
    constexpr int NA = MODE == 0 ? 63 : (MODE == 1 ? 81 : 36);
    const int pi = pair0 + blockIdx.x, l = blockIdx.y;
    const Pair P = pairs[pi];
    const Freq f = freqs[l];
    __shared__ MathTables mtab;  // tables of the exponentials
    load_math_tables(&mtab);
    __syncthreads();
    const double n[3] = {g.N[3 * P.f2], g.N[3 * P.f2 + 1], g.N[3 * P.f2 + 2]};
    // x(xh) = px + dx0 xh0 + dx1 xh1 and y likewise; barycentrics (1-h0, h0-h1, h1)
    double px[3], dx0[3], dx1[3], dy0[3], dy1[3], jac;
    {
        const double *A = g.V + 3 * P.j1[0], *B = g.V + 3 * P.j1[1], *Cc = g.V + 3 * P.j1[2];
        #pragma unroll
        for (int k = 0; k < 3; k++) px[k] = A[k];
        if (KIND == 1) {
            #pragma unroll
            for (int k = 0; k < 3; k++) dx0[k] = dy0[k] = B[k] - A[k], dx1[k] = dy1[k] = Cc[k] - B[k];
        } else if (KIND == 2) {
            const double *Y2 = g.V + 3 * P.j2[2];
            #pragma unroll
            for (int k = 0; k < 3; k++) dx0[k] = dy0[k] = B[k] - A[k], dx1[k] = Cc[k] - B[k], dy1[k] = Y2[k] - B[k];
        } else {
            const double *Y1 = g.V + 3 * P.j2[1], *Y2 = g.V + 3 * P.j2[2];
            #pragma unroll
            for (int k = 0; k < 3; k++) dx0[k] = B[k] - A[k], dx1[k] = Cc[k] - B[k], dy0[k] = Y1[k] - A[k], dy1[k] = Y2[k] - Y1[k];
        }
        const double cx[3] = {dx0[1] * dx1[2] - dx0[2] * dx1[1], dx0[2] * dx1[0] - dx0[0] * dx1[2], dx0[0] * dx1[1] - dx0[1] * dx1[0]};
        const double cy[3] = {dy0[1] * dy1[2] - dy0[2] * dy1[1], dy0[2] * dy1[0] - dy0[0] * dy1[2], dy0[0] * dy1[1] - dy0[1] * dy1[0]};
        jac = sqrt(cx[0] * cx[0] + cx[1] * cx[1] + cx[2] * cx[2]) * sqrt(cy[0] * cy[0] + cy[1] * cy[1] + cy[2] * cy[2]);
    }
    cd acc[NA];
    #pragma unroll
    for (int k = 0; k < NA; k++) acc[k] = cd{0, 0};
    const int n4 = nline * nline * nline * nline;
    for (int p = threadIdx.x; p < n4; p += 64) {
        const int ia = p / (nline * nline * nline), ib = (p / (nline * nline)) % nline, ic = (p / nline) % nline, id = p % nline;
        const double e1 = line[ia].x, e2 = line[ib].x, e3 = line[ic].x, xi = line[id].x;
        const double w = line[ia].y * line[ib].y * line[ic].y * line[id].y;
        static_assert(!HALF || (MODE == 2 && KIND == 1), "the exchanged half is folded in for B of coincident pairs");
        constexpr int NR = KIND == 1 ? (HALF ? 3 : 6) : (KIND == 2 ? 5 : 2);
#pragma unroll 1
        for (int r = 0; r < NR; r++) {
            double a0, a1, b0, b1, wt;
            region_point<KIND>(HALF ? 2 * r : r, e1, e2, e3, xi, w, a0, a1, b0, b1, wt);
            double x[3], y[3];
            #pragma unroll
            for (int k = 0; k < 3; k++) x[k] = px[k] + dx0[k] * a0 + dx1[k] * a1, y[k] = px[k] + dy0[k] * b0 + dy1[k] * b1;
            const double Lx[3] = {1.0 - a0, a0 - a1, a1}, Ly[3] = {1.0 - b0, b0 - b1, b1};
            sing_eval<MODE, HALF>(x, y, Lx, Ly, wt, n, f, m, &mtab, acc);
        }
    }
    // wave reduction; lane 0 expands and writes
    #pragma unroll
    for (int k = 0; k < NA; k++) {
        double vx = acc[k].x, vy = acc[k].y;
        #pragma unroll
        for (int o = 32; o > 0; o >>= 1) vx += __shfl_xor(vx, o, 64), vy += __shfl_xor(vy, o, 64);
        acc[k] = cd{vx * jac, vy * jac};
    }
    if (threadIdx.x != 0) return;
    cd *o = out + ((long long)pi * gridDim.y + l) * kPairStride;
    if (MODE == 0) {
        #pragma unroll
        for (int i = 0; i < 3; i++)
            #pragma unroll
            for (int j = 0; j < 3; j++)
                #pragma unroll
                for (int p = 0; p < 3; p++)
                    #pragma unroll
                    for (int b = 0; b < 3; b++) {
                        cd v = acc[(i * 3 + j) * 7 + sidx(p, b)];
                        if (p == b) v += acc[(i * 3 + j) * 7 + 6];
                        o[kPairG + (3 * i + p) * 9 + 3 * j + b] = v;
                    }
    } else if (MODE == 1) {
        #pragma unroll
        for (int i = 0; i < 3; i++)
            #pragma unroll
            for (int j = 0; j < 3; j++)
                #pragma unroll
                for (int p = 0; p < 3; p++)
                    #pragma unroll
                    for (int b = 0; b < 3; b++) o[kPairH + (3 * i + p) * 9 + 3 * j + b] = acc[(i * 3 + j) * 9 + p * 3 + b] * kInv4Pi;
    } else {
        #pragma unroll
        for (int i = 0; i < 3; i++) {
            const cd *a = acc + i * 12;
            #pragma unroll
            for (int p = 0; p < 3; p++)
                #pragma unroll
                for (int b = 0; b < 3; b++) {
                    cd v = {0, 0};
                    fma_cr(v, a[p], n[b]);
                    fma_cr(v, a[b], n[p]);
                    if (p == b) v += a[3];
                    o[kPairBt + (3 * i + p) * 3 + b] = v;
                    #pragma unroll
                    for (int k = 0; k < 3; k++) {
                        cd u = {0, 0};
                        fma_cr(u, a[4 + sidx(p, k)], n[b]);
                        fma_cr(u, a[4 + sidx(b, k)], n[p]);
                        fma_cr(u, a[10], (b == k ? n[p] : 0.0) + (p == k ? n[b] : 0.0));
                        if (p == b) fma_cr(u, a[11], n[k]);
                        o[kPairBa + (3 * i + p) * 9 + b * 3 + k] = u;
                    }
                }
        }
    }
}

// ---------------------------------------------------------------------------
// Touching pairs, G and H, quad form.  The 9 (test corner i, source corner j) basis pairs of a point
// pair need 63 (G) or 81 (H) complex accumulators when one lane takes all of them, which spills.  Here
// the four lanes of a quad each evaluate their own point, pass the point's kernel data around the quad
// (three DPP rotations), and each lane accumulates only its basis pairs c = q, q + 4, q + 8 over all four
// points.  The quads' partial sums are added across the wavefront at the end.
// ---------------------------------------------------------------------------
template <int D> __device__ inline double quad_rot(double v) {  // lane q receives lane (q + D) mod 4's value
    constexpr int sel = ((0 + D) & 3) | (((1 + D) & 3) << 2) | (((2 + D) & 3) << 4) | (((3 + D) & 3) << 6);
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)b, sel, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), sel, 0xF, 0xF, false);
    return __longlong_as_double(((long long)(unsigned)hi << 32) | (unsigned)lo);
}

// Kernel data of one point pair.  G: alpha/r, beta/r, the 6 products rh_a rh_b; H: the 9 entries of p_hat.
// Then Lx[3], Ly[3] and the weight.
template <int MODE> struct QuadK {
    static constexpr int NK = MODE == 0 ? 10 : 18;
    double v[NK + 7];
};

template <int MODE>
__device__ __forceinline__ void quad_point(const double x[3], const double y[3], const double Lx[3], const double Ly[3], double wt,
                                           const double n[3], const Freq &f, const Mat &m, const MathTables *M,
                                           QuadK<MODE> &K) {
    const double r0 = y[0] - x[0], r1 = y[1] - x[1], r2 = y[2] - x[2];
    const double d2 = r0 * r0 + r1 * r1 + r2 * r2;
    const double ir = rsqrt_nr(d2), r = d2 * ir, ir2 = ir * ir;
    const double h[3] = {r0 * ir, r1 * ir, r2 * ir};
    const Expo E = expo(r, f, M);
    if (MODE == 0) {
        cd al, be;
        g_coef(E, ir, ir2, m, al, be);
        al = al * ir, be = be * ir;
        K.v[0] = al.x, K.v[1] = al.y, K.v[2] = be.x, K.v[3] = be.y;
        K.v[4] = h[0] * h[0], K.v[5] = h[1] * h[1], K.v[6] = h[2] * h[2];
        K.v[7] = h[0] * h[1], K.v[8] = h[0] * h[2], K.v[9] = h[1] * h[2];
    } else {
        cd c1, c2, c4;
        h_coef(E, ir, ir2, f, m, c1, c2, c4);
        const double rn = h[0] * n[0] + h[1] * n[1] + h[2] * n[2];
#pragma unroll
        for (int p = 0; p < 3; p++)
#pragma unroll
            for (int b = 0; b < 3; b++) {
                cd v = c1 * (h[p] * n[b]);
                fma_cr(v, c2, n[p] * h[b] + (p == b ? rn : 0.0));
                fma_cr(v, c4, rn * h[p] * h[b]);
                K.v[2 * (p * 3 + b)] = v.x, K.v[2 * (p * 3 + b) + 1] = v.y;
            }
    }
    constexpr int o = QuadK<MODE>::NK;
#pragma unroll
    for (int k = 0; k < 3; k++) K.v[o + k] = Lx[k], K.v[o + 3 + k] = Ly[k];
    K.v[o + 6] = wt;
}

__device__ __forceinline__ double pick3(const double *a, int i) { return i == 0 ? a[0] : (i == 1 ? a[1] : a[2]); }

template <int MODE>
__device__ __forceinline__ void quad_acc(const QuadK<MODE> &K, const int ci[3], const int cj[3], const double cv[3],
                                         cd (*acc)[MODE == 0 ? 7 : 9]) {
    constexpr int o = QuadK<MODE>::NK;
#pragma unroll
    for (int s = 0; s < 3; s++) {
        const double w = K.v[o + 6] * pick3(K.v + o, ci[s]) * pick3(K.v + o + 3, cj[s]) * cv[s];
        if (MODE == 0) {
            const cd aw = cd{K.v[0], K.v[1]} * w;
#pragma unroll
            for (int e = 0; e < 6; e++) fma_cr(acc[s][e], aw, K.v[4 + e]);
            fma_cr(acc[s][6], cd{K.v[2], K.v[3]}, w);
        } else {
#pragma unroll
            for (int e = 0; e < 9; e++) fma_cr(acc[s][e], cd{K.v[2 * e], K.v[2 * e + 1]}, w);
        }
    }
}

// HALF (coincident pairs only): regions 1, 3, 5 of the six, whose exchanged images are the other three; the
// exchanged half is added afterwards by k_gal_coincident_sym.
template <int MODE, int KIND, bool HALF = false>
__global__ void __launch_bounds__(64) k_gal_sing_quad(Geom g, const Pair *__restrict__ pairs, int pair0,
                                                      const double2 *__restrict__ line, int nline,
                                                      const Freq *__restrict__ freqs, Mat m, cd *__restrict__ out) {
    constexpr int NE = MODE == 0 ? 7 : 9;
    const int pi = pair0 + blockIdx.x, l = blockIdx.y;
    const Pair P = pairs[pi];
    const Freq f = freqs[l];
    __shared__ MathTables mtab;  // tables of the exponentials
    load_math_tables(&mtab);
    __syncthreads();
    const double n[3] = {g.N[3 * P.f2], g.N[3 * P.f2 + 1], g.N[3 * P.f2 + 2]};
    double px[3], dx0[3], dx1[3], dy0[3], dy1[3], jac;
    {
        const double *A = g.V + 3 * P.j1[0], *B = g.V + 3 * P.j1[1], *Cc = g.V + 3 * P.j1[2];
        for (int k = 0; k < 3; k++) px[k] = A[k];
        if (KIND == 1) {
            for (int k = 0; k < 3; k++) dx0[k] = dy0[k] = B[k] - A[k], dx1[k] = dy1[k] = Cc[k] - B[k];
        } else if (KIND == 2) {
            const double *Y2 = g.V + 3 * P.j2[2];
            for (int k = 0; k < 3; k++) dx0[k] = dy0[k] = B[k] - A[k], dx1[k] = Cc[k] - B[k], dy1[k] = Y2[k] - B[k];
        } else {
            const double *Y1 = g.V + 3 * P.j2[1], *Y2 = g.V + 3 * P.j2[2];
            for (int k = 0; k < 3; k++) dx0[k] = B[k] - A[k], dx1[k] = Cc[k] - B[k], dy0[k] = Y1[k] - A[k], dy1[k] = Y2[k] - Y1[k];
        }
        const double cx[3] = {dx0[1] * dx1[2] - dx0[2] * dx1[1], dx0[2] * dx1[0] - dx0[0] * dx1[2], dx0[0] * dx1[1] - dx0[1] * dx1[0]};
        const double cy[3] = {dy0[1] * dy1[2] - dy0[2] * dy1[1], dy0[2] * dy1[0] - dy0[0] * dy1[2], dy0[0] * dy1[1] - dy0[1] * dy1[0]};
        jac = sqrt(cx[0] * cx[0] + cx[1] * cx[1] + cx[2] * cx[2]) * sqrt(cy[0] * cy[0] + cy[1] * cy[1] + cy[2] * cy[2]);
    }
    // this lane's basis pairs c = q + 4 s (c = 9..11 are empty slots)
    const int q = threadIdx.x & 3;
    int ci[3], cj[3];
    double cv[3];
    for (int s = 0; s < 3; s++) {
        const int c = q + 4 * s;
        cv[s] = c < 9 ? 1.0 : 0.0;
        ci[s] = (c < 9 ? c : 0) / 3, cj[s] = (c < 9 ? c : 0) % 3;
    }
    cd acc[3][NE];
    for (int s = 0; s < 3; s++)
        for (int e = 0; e < NE; e++) acc[s][e] = cd{0, 0};
    const int n4 = nline * nline * nline * nline;
    const int iters = (n4 + 63) / 64;  // every lane runs the same count: the quad exchanges need all four
    for (int itr = 0; itr < iters; ++itr) {
        int p = threadIdx.x + 64 * itr;
        const double pw = p < n4 ? 1.0 : 0.0;  // past the end: a repeated point with weight 0
        if (p >= n4) p = n4 - 1;
        const int ia = p / (nline * nline * nline), ib = (p / (nline * nline)) % nline, ic = (p / nline) % nline, id = p % nline;
        const double e1 = line[ia].x, e2 = line[ib].x, e3 = line[ic].x, xi = line[id].x;
        const double w = line[ia].y * line[ib].y * line[ic].y * line[id].y * pw;
        constexpr int NR = KIND == 1 ? (HALF ? 3 : 6) : (KIND == 2 ? 5 : 2);
#pragma unroll 1
        for (int r = 0; r < NR; r++) {
            double a0, a1, b0, b1, wt;
            region_point<KIND>(HALF ? 2 * r : r, e1, e2, e3, xi, w, a0, a1, b0, b1, wt);
            double x[3], y[3];
            for (int k = 0; k < 3; k++) x[k] = px[k] + dx0[k] * a0 + dx1[k] * a1, y[k] = px[k] + dy0[k] * b0 + dy1[k] * b1;
            const double Lx[3] = {1.0 - a0, a0 - a1, a1}, Ly[3] = {1.0 - b0, b0 - b1, b1};
            QuadK<MODE> K;
            quad_point<MODE>(x, y, Lx, Ly, wt, n, f, m, &mtab, K);
            quad_acc<MODE>(K, ci, cj, cv, acc);
#pragma unroll 1
            for (int rot = 0; rot < 3; rot++) {  // the next lane's point, then the one after, then the last
#pragma unroll
                for (int k = 0; k < QuadK<MODE>::NK + 7; k++) K.v[k] = quad_rot<1>(K.v[k]);
                quad_acc<MODE>(K, ci, cj, cv, acc);
            }
        }
    }
    // add the 16 quads' partial sums (lanes with the same q)
    for (int s = 0; s < 3; s++)
        for (int e = 0; e < NE; e++) {
            double vx = acc[s][e].x, vy = acc[s][e].y;
            for (int o = 4; o < 64; o <<= 1) vx += __shfl_xor(vx, o, 64), vy += __shfl_xor(vy, o, 64);
            acc[s][e] = cd{vx * jac, vy * jac};
        }
    if (threadIdx.x >= 4) return;
    cd *ob = out + ((long long)pi * gridDim.y + l) * kPairStride;
    for (int s = 0; s < 3; s++) {
        const int c = q + 4 * s;
        if (c >= 9) continue;
        const int i = c / 3, j = c % 3;
        for (int p = 0; p < 3; p++)
            for (int b = 0; b < 3; b++) {
                if (MODE == 0) {
                    cd v = acc[s][sidx(p, b)];
                    if (p == b) v += acc[s][6];
                    ob[kPairG + (3 * i + p) * 9 + 3 * j + b] = v;
                } else {
                    ob[kPairH + (3 * i + p) * 9 + 3 * j + b] = acc[s][p * 3 + b] * kInv4Pi;
                }
            }
    }
}

// Vertex-adjacent mirror pairs.  The rule of (f2, f1) is that of (f1, f2) with the two faces' points exchanged
// (integrator_galerkin.hpp:190-197: region 2 is region 1 with x and y swapped; the pair list orders both
// faces from the shared vertex), and the G kernel is even in y - x, so the G block of the mirror pair at
// (test corner i, row p; source corner j, column b) is the primary pair's at (j, p; i, b).  One thread per
// (mirror, frequency, entry); mirrors[k] = (mirror pair, primary pair).
__global__ void k_gal_mirror_g(const int2 *__restrict__ mirrors, int nm, int nl, cd *__restrict__ out) {
    const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
    if (t >= (long long)nm * nl * 81) return;
    const int e = (int)(t % 81);
    const long long kl = t / 81;
    const int l = (int)(kl % nl), k = (int)(kl / nl);
    const int2 mp = mirrors[k];
    const int i = e / 27, p = (e / 9) % 3, j = (e % 9) / 3, b = e % 3;
    out[((long long)mp.x * nl + l) * kPairStride + kPairG + e] =
        out[((long long)mp.y * nl + l) * kPairStride + kPairG + (3 * j + p) * 9 + 3 * i + b];
}

// Coincident pairs integrated over regions 1, 3, 5 only (sums S).  Regions 2, 4, 6 are those with x and y
// exchanged (integrator_galerkin.hpp:64-92), which swaps the test and source corners, keeps the G kernel
// (even in y - x) and negates the H kernel (odd in y - x; both points lie on the one face, so n is shared):
//   G(i p, j b) = S(i p, j b) + S(j p, i b),   H(i p, j b) = S(i p, j b) - S(j p, i b).
// One thread per (pair, frequency, matrix, entry pair {(i p, j b), (j p, i b)} with i <= j).
__global__ void k_gal_coincident_sym(int np, int nl, int do_g, int do_h, cd *__restrict__ out) {
    const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
    if (t >= (long long)np * nl * 2 * 54) return;
    const int u = (int)(t % 54);
    const int mat = (int)((t / 54) % 2);
    const long long kl = t / 108;
    if (mat == 0 ? !do_g : !do_h) return;
    // u = 9 c + 3 p + b over the corner pairs c: (0,0) (1,1) (2,2) (0,1) (0,2) (1,2)
    const int c = u / 9, p = (u / 3) % 3, b = u % 3;
    const int ci[6] = {0, 1, 2, 0, 0, 1}, cj[6] = {0, 1, 2, 1, 2, 2};
    const int i = ci[c], j = cj[c];
    cd *o = out + kl * kPairStride + (mat == 0 ? kPairG : kPairH);
    const int ea = (3 * i + p) * 9 + 3 * j + b, eb = (3 * j + p) * 9 + 3 * i + b;
    const cd sa = o[ea], sb = o[eb];
    if (mat == 0) {
        o[ea] = sa + sb;
        if (eb != ea) o[eb] = sb + sa;
    } else {
        o[ea] = sa - sb;
        if (eb != ea) o[eb] = sb - sa;
    }
}

// Gather touching-pair blocks into the spectra.  One lane per (output block, l);
// items[ptr[b]..ptr[b+1]) = pair*9 + i*3 + j in a fixed order.
__global__ void k_gal_gather_gh(const int2 *__restrict__ blocks, const int *__restrict__ ptr, const int *__restrict__ items,
                                int nblocks, int nl, const cd *__restrict__ scratch, cd *__restrict__ UG, cd *__restrict__ UH,
                                long long ld, long long lstride) {
    const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
    if (t >= (long long)nblocks * nl) return;
    const int b = (int)(t / nl), l = (int)(t % nl);
    const int2 uv = blocks[b];
    cd g[9], h[9];
    for (int e = 0; e < 9; e++) g[e] = h[e] = cd{0, 0};
    for (int k = ptr[b]; k < ptr[b + 1]; k++) {
        const int it = items[k];
        const int pi = it / 9, i = (it % 9) / 3, j = it % 3;
        const cd *o = scratch + ((long long)pi * nl + l) * kPairStride;
        for (int p = 0; p < 3; p++)
            for (int q = 0; q < 3; q++) {
                g[p * 3 + q] += o[kPairG + (3 * i + p) * 9 + 3 * j + q];
                h[p * 3 + q] += o[kPairH + (3 * i + p) * 9 + 3 * j + q];
            }
    }
    const long long base = (long long)l * lstride + (long long)(3 * uv.x) * ld + 3LL * uv.y;
    for (int p = 0; p < 3; p++)
        for (int q = 0; q < 3; q++) {
            if (UG) UG[base + p * ld + q] += g[p * 3 + q];
            if (UH) UH[base + p * ld + q] += h[p * 3 + q];
        }
}

__global__ void k_gal_gather_b(const int *__restrict__ vptr, const int *__restrict__ items, int nv, int nl,
                               const cd *__restrict__ scratch, cd *__restrict__ Bt, cd *__restrict__ Ba, long long nv3) {
    const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
    if (t >= (long long)nv * nl) return;
    const int u = (int)(t / nl), l = (int)(t % nl);
    cd bt[9], ba[27];
    for (int e = 0; e < 9; e++) bt[e] = cd{0, 0};
    for (int e = 0; e < 27; e++) ba[e] = cd{0, 0};
    for (int k = vptr[u]; k < vptr[u + 1]; k++) {
        const int it = items[k];
        const int pi = it / 3, i = it % 3;
        const cd *o = scratch + ((long long)pi * nl + l) * kPairStride;
        for (int p = 0; p < 3; p++) {
            for (int q = 0; q < 3; q++) bt[p * 3 + q] += o[kPairBt + (3 * i + p) * 3 + q];
            for (int q = 0; q < 9; q++) ba[p * 9 + q] += o[kPairBa + (3 * i + p) * 9 + q];
        }
    }
    cd *bo = Bt + (long long)l * nv3 * 3 + (long long)(3 * u) * 3;
    cd *ao = Ba + (long long)l * nv3 * 9 + (long long)(3 * u) * 9;
    for (int e = 0; e < 9; e++) bo[e] += bt[e];
    for (int e = 0; e < 27; e++) ao[e] += ba[e];
}

// Test points of the faces faces[0..nfl) in that order: tpts[fi * npts + a] = (x_a, face) with x_a the a-th rule
// point of the face, and psi[3 (fi * npts + a) + k] = w_a area phi_k(x_a), the weight of x_a for test corner k.
__global__ void k_gal_points(Geom g, const double4 *__restrict__ quad, int npts, const int *__restrict__ faces, int nfl,
                             double4 *__restrict__ tpts, double *__restrict__ psi) {
    const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
    if (t >= (long long)nfl * npts) return;
    const int fi = (int)(t / npts), a = (int)(t % npts);
    const int f1 = faces[fi];
    const double *X0 = g.V + 3 * g.F[3 * f1], *X1 = g.V + 3 * g.F[3 * f1 + 1], *X2 = g.V + 3 * g.F[3 * f1 + 2];
    const double4 Qa = quad[a];
    const double x0 = Qa.x * X0[0] + Qa.y * X1[0] + Qa.z * X2[0];
    const double x1 = Qa.x * X0[1] + Qa.y * X1[1] + Qa.z * X2[1];
    const double x2 = Qa.x * X0[2] + Qa.y * X1[2] + Qa.z * X2[2];
    tpts[t] = make_double4(x0, x1, x2, (double)f1);
    const double wa = Qa.w * g.area[f1];
    psi[3 * t] = wa * Qa.x, psi[3 * t + 1] = wa * Qa.y, psi[3 * t + 2] = wa * Qa.z;
}

// Projection of per-point rows onto the test basis.  src holds, per frequency (stride src_ls), rows 3p..3p+2 of
// ew complex entries for the local test points p = fl * npts + a of the chunk faces faces[0..nfc); the faces of a
// chunk share no vertex, so dst rows 3 F[f][k] + r are owned by one thread each:
//   dst[l][3 F[f][k] + r][c] += sum_a psi[k of point row0 + p] src[l][3 p + r][c].
__global__ void k_gal_project(const cd *__restrict__ src, long long ew, long long src_ls, const double *__restrict__ psi,
                              const int *__restrict__ faces, int nfc, int npts, long long row0, const int *__restrict__ F,
                              cd *__restrict__ dst, long long dst_ls) {
    const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
    if (t >= (long long)nfc * 3 * ew) return;
    const int l = blockIdx.y;
    const int fl = (int)(t / (3 * ew));
    const long long rem = t - (long long)fl * 3 * ew;
    const int r = (int)(rem / ew);
    const long long c = rem - r * ew;
    const cd *s = src + (long long)l * src_ls + (3LL * fl * npts + r) * ew + c;
    const double *w = psi + 3 * (row0 + (long long)fl * npts);
    cd a0 = {0, 0}, a1 = {0, 0}, a2 = {0, 0};
    for (int a = 0; a < npts; a++) {
        const cd v = s[3LL * a * ew];
        fma_cr(a0, v, w[3 * a]);
        fma_cr(a1, v, w[3 * a + 1]);
        fma_cr(a2, v, w[3 * a + 2]);
    }
    const int f = faces[fl];
    cd *d = dst + (long long)l * dst_ls + r * ew + c;
    d[3LL * F[3 * f] * ew] += a0;
    d[3LL * F[3 * f + 1] * ew] += a1;
    d[3LL * F[3 * f + 2] * ew] += a2;
}

__global__ void k_fill_zero(double2 *p, long long n) {
    const long long stride = (long long)gridDim.x * 256;
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += stride) p[i] = make_double2(0.0, 0.0);
}

void check(hipError_t e, const char *what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

struct DevVec {
    void *p = nullptr;
    bool pooled = false;
    ~DevVec() { dev_free(p, pooled); }
    void alloc(size_t n, const char *what) { check(dev_alloc(&p, std::max<size_t>(n, 1), &pooled), what); }
    template <class T> T *up(const std::vector<T> &v, hipStream_t st) {
        alloc(v.size() * sizeof(T), "device allocation");
        if (!v.empty()) check(hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, st), "hipMemcpyAsync");
        return (T *)p;
    }
};

void zero(cd *p, long long n, int num_cus, hipStream_t st) {
    if (!p || n <= 0) return;
    long long blocks = std::min<long long>((n + 255) / 256, (long long)num_cus * 8);
    hipLaunchKernelGGL(k_fill_zero, dim3((unsigned)blocks), dim3(256), 0, st, (double2 *)p, n);
    check(hipGetLastError(), "k_fill_zero");
}

struct EvTimer {
    hipEvent_t a = nullptr, b = nullptr;
    hipStream_t st;
    double *acc;
    bool on;
    EvTimer(hipStream_t s, double *d) : st(s), acc(d), on(phase_timers_on()) {
        if (on) (void)hipEventCreate(&a), (void)hipEventCreate(&b), (void)hipEventRecord(a, st);
    }
    ~EvTimer() {
        if (!on) return;
        float ms = 0;
        if (hipEventRecord(b, st) == hipSuccess && hipEventSynchronize(b) == hipSuccess && hipEventElapsedTime(&ms, a, b) == hipSuccess)
            *acc += ms;
        (void)hipEventDestroy(a), (void)hipEventDestroy(b);
    }
};

}  // namespace

int galerkin_chunk_faces(int nv, int nf, int npts, int nl, bool G, bool H, bool B, size_t budget) {
    const size_t per_face = (size_t)npts * 3 * nl * ((G ? 3ULL * nv : 0) + (H ? 3ULL * nv : 0) + (B ? 12 : 0)) * sizeof(cd);
    size_t cap = per_face ? budget / per_face : (size_t)nf;
    cap = std::max<size_t>(cap, 1);
    cap = std::min<size_t>(cap, (size_t)(65535 / std::max(npts, 1)));
    return (int)std::min<size_t>(cap, (size_t)std::max(nf, 1));
}

size_t galerkin_scratch_bytes(int nv, int nf, int npts, int nl, bool G, bool H, bool B, size_t budget) {
    const size_t cap = (size_t)galerkin_chunk_faces(nv, nf, npts, nl, G, H, B, budget);
    const size_t rows = 3 * cap * npts;
    return (size_t)nf * npts * (sizeof(double4) + 3 * sizeof(double)) +
           (size_t)nl * rows * ((G ? 3ULL * nv : 0) + (H ? 3ULL * nv : 0) + (B ? 12 : 0)) * sizeof(cd);
}

std::vector<std::vector<int>> touching_candidates(const int *F, int nf, int nv, const int *vmap) {
    // faces around each vertex, or around each original vertex when the mesh is split
    std::vector<std::vector<int>> around(nv);
    auto key = [&](int v) { return vmap ? vmap[v] : v; };
    for (int f = 0; f < nf; f++)
        for (int k = 0; k < 3; k++) around[key(F[3 * f + k])].push_back(f);
    std::vector<std::vector<int>> cand(nf);
    for (int f1 = 0; f1 < nf; f1++) {
        auto &c = cand[f1];
        for (int k = 0; k < 3; k++) {
            const auto &a = around[key(F[3 * f1 + k])];
            c.insert(c.end(), a.begin(), a.end());
        }
        std::sort(c.begin(), c.end());
        c.erase(std::unique(c.begin(), c.end()), c.end());
    }
    return cand;
}

void run_galerkin(const GalerkinJob &job, const int *F, const int *vmap, cd *U, cd *P, cd *Bt, cd *Ba, double *tm,
                  hipStream_t st) {
    const Geom &g = job.g;
    const int nv = g.nv, nf = g.nf, nl = job.nl;
    const long long n3 = 3LL * nv;
    zero(U, U ? (long long)nl * n3 * n3 : 0, job.num_cus, st);
    zero(P, P ? (long long)nl * n3 * n3 : 0, job.num_cus, st);
    zero(Bt, Bt ? (long long)nl * n3 * 3 : 0, job.num_cus, st);
    zero(Ba, Ba ? (long long)nl * n3 * 9 : 0, job.num_cus, st);

    // Greedy colouring of test faces: faces sharing a vertex get different colours.
    std::vector<int> color(nf, -1), vmark;
    std::vector<std::vector<int>> classes;
    {
        std::vector<std::vector<char>> used;  // used[c][v]
        for (int f = 0; f < nf; f++) {
            int c = 0;
            for (;; c++) {
                if (c == (int)used.size()) used.emplace_back(nv, 0), classes.emplace_back();
                if (!used[c][F[3 * f]] && !used[c][F[3 * f + 1]] && !used[c][F[3 * f + 2]]) break;
            }
            color[f] = c;
            for (int k = 0; k < 3; k++) used[c][F[3 * f + k]] = 1;
            classes[c].push_back(f);
        }
    }
    std::vector<int> flat, cls_off{0};
    for (auto &c : classes) flat.insert(flat.end(), c.begin(), c.end()), cls_off.push_back((int)flat.size());
    DevVec dfaces;
    const int *faces = dfaces.up(flat, st);
    // faces sharing an (original) vertex with each face: its touching pairs, which the regular kernels exclude
    const std::vector<std::vector<int>> cands = touching_candidates(F, nf, nv, vmap);

    const char *variant = std::getenv("CQBEM_GALERKIN_KERNEL");
    if ((U || P || Bt || Ba) && !(variant && !std::strcmp(variant, "star"))) {
        // Test-point rows: the collocation kernels evaluate, for every rule point x_a of the test faces of a
        // chunk, the row sum over the source faces not touching x_a's face; k_gal_project then weights the
        // rows by the test basis.  Chunks are runs of one colour class, so their faces share no vertex.
        EvTimer t(st, &tm[1]);
        const int npts = job.npts;
        const long long ntp = (long long)nf * npts;
        DevVec dtp, dpsi, dsG, dsH, dsBt, dsBa;
        dtp.alloc((size_t)ntp * sizeof(double4), "device allocation: test points");
        dpsi.alloc((size_t)ntp * 3 * sizeof(double), "device allocation: test weights");
        double4 *tpts = (double4 *)dtp.p;
        double *psi = (double *)dpsi.p;
        hipLaunchKernelGGL(k_gal_points, dim3((unsigned)((ntp + 255) / 256)), dim3(256), 0, st, g, job.quad, npts, faces, nf,
                           tpts, psi);
        check(hipGetLastError(), "k_gal_points");
        const int cap = galerkin_chunk_faces(nv, nf, npts, nl, U != nullptr, P != nullptr, Bt && Ba, job.scratch_bytes);
        const long long cap_rows = 3LL * cap * npts, ls = cap_rows * n3;
        if (U) dsG.alloc((size_t)nl * ls * sizeof(cd), "device allocation: point rows G");
        if (P) dsH.alloc((size_t)nl * ls * sizeof(cd), "device allocation: point rows H");
        if (Bt && Ba) {
            dsBt.alloc((size_t)nl * cap_rows * 3 * sizeof(cd), "device allocation: point rows Bt");
            dsBa.alloc((size_t)nl * cap_rows * 9 * sizeof(cd), "device allocation: point rows Ba");
        }
        cd *sG = (cd *)dsG.p, *sH = (cd *)dsH.p, *sBt = (cd *)dsBt.p, *sBa = (cd *)dsBa.p;
        auto project = [&](const cd *src, long long ew, long long src_ls, const int *cf, int nfc, long long row0, cd *dst,
                           long long dst_ls) {
            const long long n = (long long)nfc * 3 * ew;
            hipLaunchKernelGGL(k_gal_project, dim3((unsigned)((n + 255) / 256), nl), dim3(256), 0, st, src, ew, src_ls, psi, cf,
                               nfc, npts, row0, g.F, dst, dst_ls);
            check(hipGetLastError(), "k_gal_project");
        };
        for (size_t c = 0; c < classes.size(); c++) {
            const int cnt = cls_off[c + 1] - cls_off[c];
            for (int b = 0; b < cnt; b += cap) {
                const int nfc = std::min(cap, cnt - b), fi0 = cls_off[c] + b;
                const int row0 = fi0 * npts, nrows = nfc * npts;
                if (U || P) {
                    {
                        // the pair kernel at the test points: timed on its own, with its launches and its unique
                        // evaluations (test point, regular source face, source point, frequency)
                        EvTimer tp(st, &tm[2]);
                        check(launch_colloc_gh_pair(g, job.pair_pts, job.npts_even, job.freqs, 0, nl, job.mat, row0, nrows,
                                                    tpts, job.patch_pass, job.pass_info, job.pass_faces, job.patch_verts,
                                                    job.npatches, sG, sH, n3, ls, st),
                              "k_colloc_gh_pair (test points)");
                    }
                    tm[6] += 1;
                    for (int fi = fi0; fi < fi0 + nfc; fi++)
                        tm[7] += (double)npts * (double)(nf - (int)cands[flat[fi]].size()) * npts * nl;
                    if (U) project(sG, n3, ls, faces + fi0, nfc, row0, U, n3 * n3);
                    if (P) project(sH, n3, ls, faces + fi0, nfc, row0, P, n3 * n3);
                }
                if (Bt && Ba) {
                    if (job.freqs32)
                        check(launch_points_b32(g, job.quad, npts, job.freqs32, 0, nl, job.mat32, tpts, row0, nrows, sBt, sBa,
                                                cap_rows, st),
                              "k_colloc_b32 (test points)");
                    else
                        check(launch_points_b(g, job.quad, npts, job.freqs, 0, nl, job.mat, tpts, row0, nrows, sBt, sBa,
                                              cap_rows, st),
                              "k_colloc_b (test points)");
                    project(sBt, 3, cap_rows * 3, faces + fi0, nfc, row0, Bt, n3 * 3);
                    project(sBa, 9, cap_rows * 9, faces + fi0, nfc, row0, Ba, n3 * 9);
                }
            }
        }
    } else if (U || P || Bt || Ba) {
        EvTimer t(st, &tm[1]);
        for (size_t c = 0; c < classes.size(); c++) {
            const int cnt = cls_off[c + 1] - cls_off[c];
            const int *fc = faces + cls_off[c];
            if (U || P) {
                dim3 grid((nv + 63) / 64, nl, cnt);
                if (U) hipLaunchKernelGGL((k_gal_gh<true, false>), grid, dim3(64), 0, st, g, job.quad, job.npts, job.freqs, job.mat, fc, U, (cd *)nullptr, n3, n3 * n3);
                if (P) hipLaunchKernelGGL((k_gal_gh<false, true>), grid, dim3(64), 0, st, g, job.quad, job.npts, job.freqs, job.mat, fc, (cd *)nullptr, P, n3, n3 * n3);
                check(hipGetLastError(), "k_gal_gh");
            }
            if (Bt && Ba) {
                hipLaunchKernelGGL(k_gal_b, dim3(nl, cnt), dim3(256), 0, st, g, job.quad, job.npts, job.freqs, job.mat, fc, Bt, Ba, n3);
                check(hipGetLastError(), "k_gal_b");
            }
        }
    }

    // Touching pairs, classified exactly as integrate_galerkin_singular (integrator_galerkin.hpp:262-334).
    std::vector<Pair> pairs;
    {
        auto same = [&](int a, int b) { return a == b || (vmap && vmap[a] == vmap[b]); };
        for (int f1 = 0; f1 < nf; f1++) {
            for (int f2 : cands[f1]) {
                Pair p{};
                p.f1 = f1, p.f2 = f2;
                const int *A = F + 3 * f1, *B = F + 3 * f2;
                if (f1 == f2) {
                    p.kind = 1;
                    for (int k = 0; k < 3; k++) p.j1[k] = p.j2[k] = A[k];
                } else {
                    for (int i = 0; i < 3; i++)
                        for (int j = 0; j < 3; j++)
                            if ((A[i] == B[j] && A[(i + 1) % 3] == B[(j + 2) % 3]) ||
                                (vmap && vmap[A[i]] == vmap[B[j]] && vmap[A[(i + 1) % 3]] == vmap[B[(j + 2) % 3]])) {
                                p.kind = 2;
                                p.j1[0] = A[i], p.j1[1] = A[(i + 1) % 3], p.j1[2] = A[(i + 2) % 3];
                                p.j2[0] = B[j], p.j2[1] = B[(j + 2) % 3], p.j2[2] = B[(j + 1) % 3];
                                break;
                            }
                    if (!p.kind)
                        for (int i = 0; i < 3; i++)
                            for (int j = 0; j < 3; j++)
                                if (same(A[i], B[j])) {
                                    p.kind = 3;
                                    p.j1[0] = A[i], p.j1[1] = A[(i + 1) % 3], p.j1[2] = A[(i + 2) % 3];
                                    p.j2[0] = B[j], p.j2[1] = B[(j + 1) % 3], p.j2[2] = B[(j + 2) % 3];
                                    break;
                                }
                }
                if (p.kind) pairs.push_back(p);
            }
        }
    }
    if (pairs.empty()) return;
    std::stable_sort(pairs.begin(), pairs.end(), [](const Pair &a, const Pair &b) { return a.kind < b.kind; });
    int kind_off[4] = {0, 0, 0, 0};
    for (const Pair &p : pairs) kind_off[p.kind]++;
    kind_off[3] += kind_off[2] += kind_off[1];  // exclusive ends per kind: [0,k1) coincident, [k1,k2) edge, [k2,k3) vertex
    // Vertex-adjacent mirrors: a pair (f1, f2) with f1 > f2 whose pair (f2, f1) is listed with both faces ordered from
    // the same shared vertex (as the classification orders them on a consistently oriented mesh).  Its G block is
    // the transpose of that primary's and is not integrated (quad kernels); the mirrors go after the other
    // vertex-adjacent pairs, and a pair without such a primary is integrated in full.
    std::vector<int2> mirrors;  // (mirror pair, primary pair)
    int vprim = kind_off[3];    // end of the vertex-adjacent pairs that are integrated for G
    {
        auto key = [](const Pair &p) { return std::make_pair(p.f1, p.f2); };
        std::map<std::pair<int, int>, int> at;
        for (int k = kind_off[2]; k < kind_off[3]; k++) at[key(pairs[k])] = k;
        auto mirrored = [&](const Pair &p) {
            if (p.f1 < p.f2) return false;
            auto it = at.find({p.f2, p.f1});
            if (it == at.end()) return false;
            const Pair &q = pairs[it->second];
            for (int c = 0; c < 3; c++)
                if (q.j1[c] != p.j2[c] || q.j2[c] != p.j1[c]) return false;
            return true;
        };
        std::vector<char> mir(pairs.size(), 0);
        for (int k = kind_off[2]; k < kind_off[3]; k++) mir[k] = mirrored(pairs[k]) ? 1 : 0;
        std::vector<Pair> prim, rest;
        for (int k = kind_off[2]; k < kind_off[3]; k++) (mir[k] ? rest : prim).push_back(pairs[k]);
        std::copy(prim.begin(), prim.end(), pairs.begin() + kind_off[2]);
        std::copy(rest.begin(), rest.end(), pairs.begin() + kind_off[2] + (int)prim.size());
        vprim = kind_off[2] + (int)prim.size();
        at.clear();
        for (int k = kind_off[2]; k < vprim; k++) at[key(pairs[k])] = k;
        for (int k = vprim; k < kind_off[3]; k++) mirrors.push_back(make_int2(k, at.at({pairs[k].f2, pairs[k].f1})));
    }
    const long long npairs = (long long)pairs.size();
    DevVec dpairs, dscratch, dblocks, dptr, ditems, dvptr, dvitems, dmirrors;
    const Pair *dp = dpairs.up(pairs, st);
    dscratch.alloc((size_t)npairs * nl * kPairStride * sizeof(cd), "device allocation: touching-pair scratch");
    cd *scratch = (cd *)dscratch.p;
    {
        EvTimer t(st, &tm[3]);
        const int beg[3] = {0, kind_off[1], kind_off[2]}, end[3] = {kind_off[1], kind_off[2], kind_off[3]};
#define CQBEM_SING(MODE, KIND)                                                                                        \
    if (end[KIND - 1] > beg[KIND - 1])                                                                                \
        hipLaunchKernelGGL((k_gal_sing<MODE, KIND>), dim3((unsigned)(end[KIND - 1] - beg[KIND - 1]), nl), dim3(64), 0, st, g, \
                           dp, beg[KIND - 1], job.line, job.nline, job.freqs, job.mat, scratch);
        const char *sv = std::getenv("CQBEM_GALERKIN_SINGULAR");
        const bool lane = sv && !std::strcmp(sv, "lane");  // every region of every ordered pair integrated
        if (!lane) {
#define CQBEM_SINGQ(MODE, KIND, HALF, E)                                                                              \
    if ((E) > beg[KIND - 1])                                                                                          \
        hipLaunchKernelGGL((k_gal_sing_quad<MODE, KIND, HALF>), dim3((unsigned)((E) - beg[KIND - 1]), nl), dim3(64), 0, \
                           st, g, dp, beg[KIND - 1], job.line, job.nline, job.freqs, job.mat, scratch);
            // coincident: half the regions, then the exchanged half; vertex-adjacent G: the primaries, then
            // their mirrors' transposes
            if (U) { CQBEM_SINGQ(0, 1, true, end[0]) CQBEM_SINGQ(0, 2, false, end[1]) CQBEM_SINGQ(0, 3, false, vprim) }
            if (P) { CQBEM_SINGQ(1, 1, true, end[0]) CQBEM_SINGQ(1, 2, false, end[1]) CQBEM_SINGQ(1, 3, false, end[2]) }
#undef CQBEM_SINGQ
            if ((U || P) && end[0] > 0) {
                const long long n = (long long)end[0] * nl * 108;
                hipLaunchKernelGGL(k_gal_coincident_sym, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, end[0], nl,
                                   U ? 1 : 0, P ? 1 : 0, scratch);
            }
            if (U && !mirrors.empty()) {
                const int2 *dm = dmirrors.up(mirrors, st);
                const long long n = (long long)mirrors.size() * nl * 81;
                hipLaunchKernelGGL(k_gal_mirror_g, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, dm, (int)mirrors.size(),
                                   nl, scratch);
            }
        } else {
            if (U) { CQBEM_SING(0, 1) CQBEM_SING(0, 2) CQBEM_SING(0, 3) }
            if (P) { CQBEM_SING(1, 1) CQBEM_SING(1, 2) CQBEM_SING(1, 3) }
        }
        if (Bt && Ba) {
            if (lane) {
                CQBEM_SING(2, 1)
            } else if (end[0] > beg[0]) {  // coincident B over half the regions, the exchanged half folded in
                hipLaunchKernelGGL((k_gal_sing<2, 1, true>), dim3((unsigned)(end[0] - beg[0]), nl), dim3(64), 0, st, g, dp,
                                   beg[0], job.line, job.nline, job.freqs, job.mat, scratch);
            }
            CQBEM_SING(2, 2) CQBEM_SING(2, 3)
        }
#undef CQBEM_SING
        check(hipGetLastError(), "k_gal_sing");
    }
    // gather lists: output block (u, v) <- (pair, i, j); vertex u <- (pair, i)
    std::map<std::pair<int, int>, std::vector<int>> bl;
    std::vector<std::vector<int>> vl(nv);
    for (long long k = 0; k < npairs; k++) {
        for (int i = 0; i < 3; i++) {
            vl[pairs[k].j1[i]].push_back((int)(k * 3 + i));
            for (int j = 0; j < 3; j++) bl[{pairs[k].j1[i], pairs[k].j2[j]}].push_back((int)(k * 9 + i * 3 + j));
        }
    }
    std::vector<int2> blocks;
    std::vector<int> ptr{0}, items, vptr{0}, vitems;
    for (auto &kv : bl) {
        blocks.push_back(make_int2(kv.first.first, kv.first.second));
        items.insert(items.end(), kv.second.begin(), kv.second.end());
        ptr.push_back((int)items.size());
    }
    for (int v = 0; v < nv; v++) vitems.insert(vitems.end(), vl[v].begin(), vl[v].end()), vptr.push_back((int)vitems.size());
    const int2 *db = dblocks.up(blocks, st);
    const int *dpt = dptr.up(ptr, st), *dit = ditems.up(items, st), *dvp = dvptr.up(vptr, st), *dvi = dvitems.up(vitems, st);
    {
        EvTimer t(st, &tm[3]);
        if (U || P) {
            const long long n = (long long)blocks.size() * nl;
            hipLaunchKernelGGL(k_gal_gather_gh, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, db, dpt, dit, (int)blocks.size(), nl,
                               scratch, U, P, n3, n3 * n3);
            check(hipGetLastError(), "k_gal_gather_gh");
        }
        if (Bt && Ba) {
            const long long n = (long long)nv * nl;
            hipLaunchKernelGGL(k_gal_gather_b, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, dvp, dvi, nv, nl, scratch, Bt, Ba, n3);
            check(hipGetLastError(), "k_gal_gather_b");
        }
    }
    check(hipStreamSynchronize(st), "galerkin sync");
}

}  // namespace cqbem
