// Mixed-precision (CQBEM_MIXED) far-field kernels: the reference's precision policy, where the far-field
// collocation G, B_trans and B_angular integrals are evaluated in float (G_collocation.cu:14-15,
// B_trans_collocation.cu:14, B_angular_collocation.cu:14) while H, the singular faces and the inverse
// transform stay in double.  Everything that depends on the point pair and the frequency is float here:
// coordinates, r and r-hat, the exponent z = r s / c, e^{-z} (hardware v_exp_f32 and v_sin_f32 / v_cos_f32
// after reduction to one revolution), the kernel coefficients and the quadrature sums.  Each face's sum over
// its quadrature points is formed in float and added to the running entry with Kahan compensation
// (W. Kahan, "Further remarks on reducing truncation errors", CACM 8, 1965); the compensated float pair is
// widened to double when the entry is stored.
//
// Work decomposition, as the double kernels' simplest forms: G takes a lane per (row vertex, column vertex,
// frequency) that integrates the faces of the column vertex's star with that corner's basis function (each
// point is evaluated for each of its face's three corners); B takes a 256-lane block per (row, frequency) whose
// lanes stride the faces, reduced over the wave in double.
#include <hip/hip_runtime.h>

#include "cqbem_kernels.hpp"
#include "cqbem_mixed.hpp"

namespace cqbem {
namespace {

struct cf {
    float x, y;
};
__device__ inline cf operator+(cf a, cf b) { return cf{a.x + b.x, a.y + b.y}; }
__device__ inline cf operator-(cf a, cf b) { return cf{a.x - b.x, a.y - b.y}; }
__device__ inline cf operator*(cf a, cf b) { return cf{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
__device__ inline cf operator*(cf a, float s) { return cf{a.x * s, a.y * s}; }
__device__ inline void fma_cf(cf &a, cf b, float s) {
    a.x = fmaf(b.x, s, a.x);
    a.y = fmaf(b.y, s, a.y);
}

// Kahan-compensated running sum of complex float addends (one per face).
struct KahanC {
    cf s{0.f, 0.f}, c{0.f, 0.f};
    __device__ inline void add(cf v) {
#pragma clang fp contract(off)
        const float yx = v.x - c.x, yy = v.y - c.y;
        const float tx = s.x + yx, ty = s.y + yy;
        c.x = (tx - s.x) - yx, c.y = (ty - s.y) - yy;
        s.x = tx, s.y = ty;
    }
    __device__ inline cd value() const { return cd{(double)s.x - (double)c.x, (double)s.y - (double)c.y}; }
};

// e^{-z} in float, Re z >= 0: 2^(-Re z log2 e) by v_exp_f32, and the phase reduced to [-1/2, 1/2] revolution
// for v_sin_f32 / v_cos_f32 (which take their argument in revolutions).
__device__ inline cf cexp_neg_f(cf z) {
    const float m = __builtin_amdgcn_exp2f(-z.x * 1.44269504f);
    float t = z.y * 0.159154943f;
    t = t - rintf(t);
    return cf{m * __builtin_amdgcn_cosf(t), -m * __builtin_amdgcn_sinf(t)};
}

struct ExpoF {
    cf e1, e2, dA, A2;
};

__device__ inline ExpoF expo_f(float r, const FreqF &f) {
    ExpoF E;
    const cf sc1{f.sc1x, f.sc1y}, sc2{f.sc2x, f.sc2y}, is2{f.is2x, f.is2y};
    const cf z1 = sc1 * r, z2 = sc2 * r;
    E.e1 = cexp_neg_f(z1);
    E.e2 = cexp_neg_f(z2);
    const cf A1 = (z1 + cf{1.f, 0.f}) * E.e1 * is2;
    E.A2 = (z2 + cf{1.f, 0.f}) * E.e2 * is2;
    E.dA = A1 - E.A2;
    return E;
}

__device__ inline bool same_vertex_g(const Geom &g, int a, int b) {
    return a == b || (g.vmap != nullptr && g.vmap[a] == g.vmap[b]);
}

// -------------------------------------------------------------------- G --
// u_hat = alpha rh rh^T + beta I with alpha = (3 dA / r^2 + e1/c1^2 - e2/c2^2) / (4 pi rho r),
// beta = (e2/c2^2 - dA / r^2) / (4 pi rho r)  (G_collocation.cu:44-60 in factored form).
__global__ void __launch_bounds__(64) k_colloc_g32(Geom g, const double4 *__restrict__ quad, int npts,
                                                    const FreqF *__restrict__ freqs, int l0, MatF m, int row0,
                                                    cd *__restrict__ UG, long long ld, long long lstride) {
    const int j = blockIdx.x * 64 + threadIdx.x;
    const int il = blockIdx.z;
    const int i = row0 + il;
    const int l = blockIdx.y;
    if (j >= g.nv) return;
    const FreqF f = freqs[l0 + l];
    const float x0 = (float)g.V[3 * i], x1 = (float)g.V[3 * i + 1], x2 = (float)g.V[3 * i + 2];
    KahanC ga[6], gb;
    const int t_end = g.vs_ptr[j + 1];
    for (int t = g.vs_ptr[j]; t < t_end; ++t) {
        const int pk = g.vs_face[t];
        const int fc = pk >> 2, c = pk & 3;
        const int a0 = g.F[3 * fc], a1 = g.F[3 * fc + 1], a2 = g.F[3 * fc + 2];
        if (same_vertex_g(g, i, a0) || same_vertex_g(g, i, a1) || same_vertex_g(g, i, a2)) continue;
        const double *p0 = g.V + 3 * a0, *p1 = g.V + 3 * a1, *p2 = g.V + 3 * a2;
        const float y00 = (float)p0[0], y01 = (float)p0[1], y02 = (float)p0[2];
        const float y10 = (float)p1[0], y11 = (float)p1[1], y12 = (float)p1[2];
        const float y20 = (float)p2[0], y21 = (float)p2[1], y22 = (float)p2[2];
        const float ar = (float)g.area[fc];
        cf fa[6] = {{0.f, 0.f}, {0.f, 0.f}, {0.f, 0.f}, {0.f, 0.f}, {0.f, 0.f}, {0.f, 0.f}}, fb = {0.f, 0.f};
        for (int q = 0; q < npts; ++q) {
            const double4 Q = quad[q];
            const float b0 = (float)Q.x, b1 = (float)Q.y, b2 = (float)Q.z;
            const float r0 = b0 * y00 + b1 * y10 + b2 * y20 - x0;
            const float r1 = b0 * y01 + b1 * y11 + b2 * y21 - x1;
            const float r2 = b0 * y02 + b1 * y12 + b2 * y22 - x2;
            const float d2 = r0 * r0 + r1 * r1 + r2 * r2;
            const float ir = __builtin_amdgcn_rsqf(d2), r = d2 * ir, ir2 = ir * ir;
            const float h0 = r0 * ir, h1 = r1 * ir, h2 = r2 * ir;
            const float bc = c == 0 ? b0 : (c == 1 ? b1 : b2);
            const float wr = (float)Q.w * ar * bc * ir;
            const ExpoF E = expo_f(r, f);
            const cf tt = E.dA * ir2;
            const cf al = (tt * 3.f + E.e1 * m.ic1s - E.e2 * m.ic2s) * (m.g_scale * wr);
            const cf be = (E.e2 * m.ic2s - tt) * (m.g_scale * wr);
            fb = fb + be;
            fma_cf(fa[0], al, h0 * h0);
            fma_cf(fa[1], al, h1 * h1);
            fma_cf(fa[2], al, h2 * h2);
            fma_cf(fa[3], al, h0 * h1);
            fma_cf(fa[4], al, h0 * h2);
            fma_cf(fa[5], al, h1 * h2);
        }
        for (int k = 0; k < 6; k++) ga[k].add(fa[k]);
        gb.add(fb);
    }
    const long long base = (long long)l * lstride + (long long)(3 * il) * ld + 3 * (long long)j;
    const cd beta = gb.value();
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) {
            cd v = ga[sidx(a, b)].value();
            if (a == b) v += beta;
            UG[base + a * ld + b] = v;
        }
}

// -------------------------------------------------------------------- B --
constexpr int kBWords32 = 48;

// GAL: the rows are test points tpts[row0 + r] = (x, y, z, test face), each excluding the faces that share a vertex
// with its test face, and row r of the launch goes to rows 3r..3r+2 of Bt/Ba (as k_colloc_b<true>).
template <bool GAL>
__global__ void __launch_bounds__(256) k_colloc_b32(Geom g, const double4 *__restrict__ quad, int npts,
                                                     const FreqF *__restrict__ freqs, int l0, MatF m, int row0,
                                                     const double4 *__restrict__ tpts, cd *__restrict__ Bt,
                                                     cd *__restrict__ Ba, long long nv3, FaceBands fb) {
    __shared__ double red[4][kBWords32];
    const int l = blockIdx.x;
    const int i = row0 + blockIdx.y;
    const FreqF f = freqs[l0 + l];
    const cf is{f.isx, f.isy}, is2{f.is2x, f.is2y};
    double xd[3];  // the row point (double: the face bands decide as the host plan does)
    int t[3] = {i, i, i};
    if (GAL) {
        const double4 T = tpts[i];
        xd[0] = T.x, xd[1] = T.y, xd[2] = T.z;
        const int f1 = (int)T.w;
        t[0] = g.F[3 * f1], t[1] = g.F[3 * f1 + 1], t[2] = g.F[3 * f1 + 2];
    } else {
        xd[0] = g.V[3 * i], xd[1] = g.V[3 * i + 1], xd[2] = g.V[3 * i + 2];
    }
    const float x0 = (float)xd[0], x1 = (float)xd[1], x2 = (float)xd[2];
    const int io = GAL ? (int)blockIdx.y : i;
    constexpr float kI8 = 0.0397887358f, kI4 = 0.0795774715f;  // 1/(8 pi), 1/(4 pi)
    KahanC bt[6], ba[18];
    for (int fc = threadIdx.x; fc < g.nf; fc += 256) {
        const int a0 = g.F[3 * fc], a1 = g.F[3 * fc + 1], a2 = g.F[3 * fc + 2];
        bool touch = false;
        for (int k = 0; k < (GAL ? 3 : 1); k++)
            touch = touch || same_vertex_g(g, t[k], a0) || same_vertex_g(g, t[k], a1) || same_vertex_g(g, t[k], a2);
        if (touch) continue;
        const double *p0 = g.V + 3 * a0, *p1 = g.V + 3 * a1, *p2 = g.V + 3 * a2;
        const float n[3] = {(float)g.N[3 * fc], (float)g.N[3 * fc + 1], (float)g.N[3 * fc + 2]};
        const float ar = (float)g.area[fc];
        const float y0[3] = {(float)p0[0], (float)p0[1], (float)p0[2]}, y1[3] = {(float)p1[0], (float)p1[1], (float)p1[2]},
                    y2[3] = {(float)p2[0], (float)p2[1], (float)p2[2]};
        const double4 *qf = quad;  // the face's rule
        int nq = npts;
        if (fb.sphere && beyond_sphere(xd[0], xd[1], xd[2], fb.sphere[fc])) qf = fb.far_quad, nq = fb.far_npts;
        cf mt[3] = {{0.f, 0.f}, {0.f, 0.f}, {0.f, 0.f}}, m2 = {0.f, 0.f}, M[6], S = {0.f, 0.f}, S3 = {0.f, 0.f};
#pragma unroll
        for (int k = 0; k < 6; k++) M[k] = cf{0.f, 0.f};
        for (int q = 0; q < nq; ++q) {
            const double4 Q = qf[q];
            const float b0 = (float)Q.x, b1 = (float)Q.y, b2 = (float)Q.z;
            const float r0 = b0 * y0[0] + b1 * y1[0] + b2 * y2[0] - x0;
            const float r1 = b0 * y0[1] + b1 * y1[1] + b2 * y2[1] - x1;
            const float r2 = b0 * y0[2] + b1 * y1[2] + b2 * y2[2] - x2;
            const float d2 = r0 * r0 + r1 * r1 + r2 * r2;
            const float ir = __builtin_amdgcn_rsqf(d2), r = d2 * ir, ir2 = ir * ir;
            const float h[3] = {r0 * ir, r1 * ir, r2 * ir};
            const float w = (float)Q.w * ar;
            const ExpoF E = expo_f(r, f);
            // b_hat = t1 (rh n^T + n rh^T) + t2 (rh.n) I; b_hat_ijk = u1 (...) + (u1 + u2)(...) + u3 d_ij n_k
            // (B_trans_collocation.cu:41-58, B_angular_collocation.cu:44-63 in factored form)
            const cf t1 = E.dA * (ir2 * kI8 * w);
            const cf t2 = (E.A2 - is2) * (ir2 * kI4);
            const cf u1 = E.dA * (ir * kI8);
            const cf u2 = (E.e2 * m.ic2 - E.e1 * m.ic1) * is * kI8;
            const cf u12 = u1 + u2;
            const cf u1w = u1 * w;
            const cf u3 = E.e2 * is * (m.ic2 * kI4);
            const float rn = h[0] * n[0] + h[1] * n[1] + h[2] * n[2];
#pragma unroll
            for (int a = 0; a < 3; a++) fma_cf(mt[a], t1, h[a]);
            fma_cf(m2, t2, w * rn);
#pragma unroll
            for (int a = 0; a < 3; a++)
#pragma unroll
                for (int b = a; b < 3; b++) fma_cf(M[sidx(a, b)], u1w, h[a] * h[b]);
            fma_cf(S, u12, w);
            fma_cf(S3, u3, w);
        }
#pragma unroll
        for (int a = 0; a < 3; a++)
#pragma unroll
            for (int b = a; b < 3; b++) {
                cf v = {0.f, 0.f};
                fma_cf(v, mt[a], n[b]);
                fma_cf(v, mt[b], n[a]);
                if (a == b) v = v + m2;
                bt[sidx(a, b)].add(v);
#pragma unroll
                for (int k = 0; k < 3; k++) {
                    cf u = {0.f, 0.f};
                    fma_cf(u, M[sidx(a, k)], n[b]);
                    fma_cf(u, M[sidx(b, k)], n[a]);
                    fma_cf(u, S, (b == k ? n[a] : 0.f) + (a == k ? n[b] : 0.f));
                    if (a == b) fma_cf(u, S3, n[k]);
                    ba[sidx(a, b) * 3 + k].add(u);
                }
            }
    }
    // the compensated lane sums, widened to double, reduced over the wave and then the 4 waves (fixed order)
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
    for (int k = 0; k < kBWords32; k++) {
        const cd e = k < 12 ? bt[k >> 1].value() : ba[(k - 12) >> 1].value();
        double v = (k & 1) ? e.y : e.x;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
        if (lane == 0) red[wave][k] = v;
    }
    __syncthreads();
    if (threadIdx.x < kBWords32) {
        const int k = threadIdx.x;
        const double v = (red[0][k] + red[1][k]) + (red[2][k] + red[3][k]);
        const int c = k & 1;
        if (k < 12) {
            const int e = k >> 1;
            double *bt_out = reinterpret_cast<double *>(Bt + (long long)l * nv3 * 3 + (long long)(3 * io) * 3);
            for (int a = 0; a < 3; a++)
                for (int b = 0; b < 3; b++)
                    if (sidx(a, b) == e) bt_out[2 * (a * 3 + b) + c] = v;
        } else {
            const int e = (k - 12) >> 1, sab = e / 3, kk = e % 3;
            double *ba_out = reinterpret_cast<double *>(Ba + (long long)l * nv3 * 9 + (long long)(3 * io) * 9);
            for (int a = 0; a < 3; a++)
                for (int b = 0; b < 3; b++)
                    if (sidx(a, b) == sab) ba_out[2 * (a * 9 + b * 3 + kk) + c] = v;
        }
    }
}

}  // namespace

hipError_t launch_colloc_g32(const Geom &g, const double4 *quad, int npts, const FreqF *freqs, int l0, int nl, const MatF &m,
                             int row0, int nrows, cd *UG, long long ld, long long lstride, hipStream_t st) {
    if (nrows <= 0 || nl <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_colloc_g32, dim3((g.nv + 63) / 64, nl, nrows), dim3(64), 0, st, g, quad, npts, freqs, l0, m, row0,
                       UG, ld, lstride);
    return hipGetLastError();
}

hipError_t launch_colloc_b32(const Geom &g, const double4 *quad, int npts, const FreqF *freqs, int l0, int nl, const MatF &m,
                             int row0, int nrows, cd *Bt, cd *Ba, hipStream_t st, const FaceBands &fb) {
    if (nrows <= 0 || nl <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_colloc_b32<false>, dim3(nl, nrows), dim3(256), 0, st, g, quad, npts, freqs, l0, m, row0,
                       (const double4 *)nullptr, Bt, Ba, (long long)3 * g.nv, fb);
    return hipGetLastError();
}

hipError_t launch_points_b32(const Geom &g, const double4 *quad, int npts, const FreqF *freqs, int l0, int nl, const MatF &m,
                             const double4 *tpts, int row0, int nrows, cd *Bt, cd *Ba, long long out_rows, hipStream_t st) {
    if (nrows <= 0 || nrows > 65535 || 3LL * nrows > out_rows) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_colloc_b32<true>, dim3(nl, nrows), dim3(256), 0, st, g, quad, npts, freqs, l0, m, row0, tpts, Bt,
                       Ba, out_rows, FaceBands{});
    return hipGetLastError();
}

std::vector<FreqF> freqs_f32(const std::vector<Freq> &f) {
    std::vector<FreqF> out(f.size());
    for (size_t l = 0; l < f.size(); l++)
        out[l] = FreqF{(float)f[l].sc1.x, (float)f[l].sc1.y, (float)f[l].sc2.x, (float)f[l].sc2.y,
                       (float)f[l].is.x,  (float)f[l].is.y,  (float)f[l].is2.x, (float)f[l].is2.y};
    return out;
}

MatF mat_f32(const Mat &m) { return MatF{(float)m.ic1, (float)m.ic2, (float)m.ic1s, (float)m.ic2s, (float)m.g_scale}; }

}  // namespace cqbem
