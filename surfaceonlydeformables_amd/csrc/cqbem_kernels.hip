// gfx950 kernels of the CQBEM precompute (collocation path, inverse CQ
// transform, jump/B post-processing).
//
// Work decomposition (collocation, regular pairs): one lane owns one output
// block (row vertex i, column vertex j, frequency l) and integrates the faces
// of j's vertex star with the basis function of j's corner.  Each block is
// therefore written exactly once, with no atomics and no read-modify-write;
// lanes of a wavefront share (i, l), so x_i and the frequency constants live
// in scalar registers.  Faces touching i (the singular ones) are skipped and
// integrated by the singular kernels, which own rows (i, l) and add into them.
//
// Reference counterparts: compute_elastodynamic_{G,H,B_trans,B_angular}_kernel_collocation
// (src/elastodynamics_cqbem_*_collocation.cu) and the inverse transform loops
// of src/elastodynamics_cqbem_{G,H,B_trans,B_angular}.cu.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <cstring>

#include "cqbem_device.hpp"
#include "cqbem_kernels.hpp"

namespace cqbem {

__device__ inline bool same_vertex(const Geom &g, int a, int b) {
    return a == b || (g.vmap != nullptr && g.vmap[a] == g.vmap[b]);
}

// ------------------------------------------------------------------------
// G and H, regular (non-touching) faces.  Grid: (ceil(nv/64), nl, rows).
// ------------------------------------------------------------------------
template <bool DG, bool DH>
__global__ void __launch_bounds__(64) k_colloc_gh(Geom g, const double4 *__restrict__ quad, int npts,
                                                   const Freq *__restrict__ freqs, int l0, Mat m, int row0,
                                                   cd *__restrict__ UG, cd *__restrict__ UH, long long ld,
                                                   long long lstride) {
    const int j = blockIdx.x * 64 + threadIdx.x;
    const int il = blockIdx.z;
    const int i = row0 + il;
    const int l = blockIdx.y;
    if (j >= g.nv) return;
    const Freq f = freqs[l0 + l];
    const double x0 = g.V[3 * i], x1 = g.V[3 * i + 1], x2 = g.V[3 * i + 2];
    cd ga[6], gb = {0, 0}, hh[9];
    for (int k = 0; k < 6; k++) ga[k] = cd{0, 0};
    for (int k = 0; k < 9; k++) hh[k] = cd{0, 0};

    const int t_end = g.vs_ptr[j + 1];
    for (int t = g.vs_ptr[j]; t < t_end; ++t) {
        const int pk = g.vs_face[t];
        const int fc = pk >> 2, c = pk & 3;
        const int a0 = g.F[3 * fc], a1 = g.F[3 * fc + 1], a2 = g.F[3 * fc + 2];
        if (same_vertex(g, i, a0) || same_vertex(g, i, a1) || same_vertex(g, i, a2)) continue;
        const double *p0 = g.V + 3 * a0, *p1 = g.V + 3 * a1, *p2 = g.V + 3 * a2;
        const double y00 = p0[0], y01 = p0[1], y02 = p0[2];
        const double y10 = p1[0], y11 = p1[1], y12 = p1[2];
        const double y20 = p2[0], y21 = p2[1], y22 = p2[2];
        const double n0 = g.N[3 * fc], n1 = g.N[3 * fc + 1], n2 = g.N[3 * fc + 2];
        const double ar = g.area[fc];
        cd v1[3], v2[3], s2 = {0, 0}, m4[6];
        if (DH) {
            for (int k = 0; k < 3; k++) v1[k] = v2[k] = cd{0, 0};
            for (int k = 0; k < 6; k++) m4[k] = cd{0, 0};
        }
        for (int q = 0; q < npts; ++q) {
            const double4 Q = quad[q];
            const double r0 = Q.x * y00 + Q.y * y10 + Q.z * y20 - x0;
            const double r1 = Q.x * y01 + Q.y * y11 + Q.z * y21 - x1;
            const double r2 = Q.x * y02 + Q.y * y12 + Q.z * y22 - x2;
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
                fma_cr(gb, be, wr);
                fma_cr(ga[0], al, h0 * h0);
                fma_cr(ga[1], al, h1 * h1);
                fma_cr(ga[2], al, h2 * h2);
                fma_cr(ga[3], al, h0 * h1);
                fma_cr(ga[4], al, h0 * h2);
                fma_cr(ga[5], al, h1 * h2);
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
            for (int a = 0; a < 3; a++)
                for (int b = 0; b < 3; b++) {
                    cd v = m4[sidx(a, b)];
                    fma_cr(v, v1[a], n[b]);
                    fma_cr(v, v2[b], n[a]);
                    if (a == b) v += s2;
                    hh[a * 3 + b] += v;
                }
        }
    }
    const long long base = (long long)l * lstride + (long long)(3 * il) * ld + 3 * (long long)j;
    for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) {
            if (DG) {
                cd v = ga[sidx(a, b)];
                if (a == b) v += gb;
                UG[base + a * ld + b] = v;
            }
            if (DH) UH[base + a * ld + b] = hh[a * 3 + b] * kInv4Pi;
        }
}


// ------------------------------------------------------------------------
// G and H, regular faces, pair form.  Column vertices are grouped into
// compact patches of at most kPairPV vertices; a block owns one (patch,
// frequency, row vertex) and evaluates every face touching the patch once.
// Two adjacent lanes share a face: per round each evaluates one quadrature
// point, forms its 3x3 G and H blocks K (7 + 9 complex), and the two swap K
// with a DPP quad permutation, so both see both points.  Lane A (even)
// accumulates corner 0 and lane B (odd) corner 2 over both points; corner 1
// is summed over each lane's own points and the two halves are added after
// the face (lane A keeps its G part, lane B its H part) -- each point is
// evaluated once for all three corner basis functions.  The face sums are added into
// per-vertex LDS accumulators in sub-steps that the host schedules so that
// no vertex is updated twice in one sub-step, in face order: deterministic,
// no atomics.  The block then writes each (row, column) block once.
// ------------------------------------------------------------------------
constexpr int kPairPV = 256;                   // vertex slots per patch (64 KB of LDS accumulators)
constexpr int kPairThreads = 256;
constexpr int kPairFaces = kPairThreads / 2;  // faces per pass
constexpr int kPairWords = 32;                // doubles per slot: G (6 sym + beta) then H (9), complex
constexpr int kFaceWords = 13;                // per face of a pass: 3 corners, normal, area (16 KB per block)

// Swap a double with the neighbouring lane of the pair (DPP quad_perm [1,0,3,2]).
__device__ inline double pair_swap(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)b, 0xB1, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), 0xB1, 0xF, 0xF, false);
    return __longlong_as_double(((long long)(unsigned)hi << 32) | (unsigned)lo);
}

// One (patch pc, frequency l, row il) tile of the pair form, for a block of kPairThreads threads.  acc, fgeo and
// mtab are the block's LDS (mtab loaded and visible); the tile zeroes acc itself and ends with its write-out
// (callers that run another tile on the same LDS synchronise first).
template <bool DG, bool DH, bool GAL>
__device__ __attribute__((always_inline)) inline void pair_tile(const Geom &g, const double4 *__restrict__ pts, int nrounds,
                                 const Bands &bands, const Freq *__restrict__ freqs,
                                 int l0, const Mat &m, int row0, const double4 *__restrict__ tpts,
                                 const int *__restrict__ patch_pass, const int2 *__restrict__ pass_info,
                                 const int4 *__restrict__ pass_faces, const int *__restrict__ patch_verts,
                                 cd *__restrict__ UG, cd *__restrict__ UH, long long ld, long long lstride, int pc, int l,
                                 int il, double *acc, double *fgeo, const MathTables &mtab) {
    const int i = row0 + il;
    const int tid = threadIdx.x, pr = tid >> 1, side = tid & 1;
    const Freq f = freqs[l0 + l];
    // the row point: vertex i (collocation), or test point i with its face's corners t[] excluded (Galerkin)
    double x0, x1, x2;
    int t[3] = {i, i, i};
    if (GAL) {
        const double4 T = tpts[i];
        x0 = T.x, x1 = T.y, x2 = T.z;
        const int f1 = (int)T.w;
        t[0] = g.F[3 * f1], t[1] = g.F[3 * f1 + 1], t[2] = g.F[3 * f1 + 2];
    } else {
        x0 = g.V[3 * i], x1 = g.V[3 * i + 1], x2 = g.V[3 * i + 2];
    }
    if (bands.far_pts) {  // the tile's rule: block-uniform (patch and row are)
        if (__builtin_amdgcn_readfirstlane((int)beyond_sphere(x0, x1, x2, bands.sphere[pc])))
            pts = bands.far_pts, nrounds = bands.far_rounds;
    }
    for (int k = tid; k < kPairWords * kPairPV; k += kPairThreads) acc[k] = 0.0;
    __syncthreads();

    for (int pass = patch_pass[pc]; pass < patch_pass[pc + 1]; ++pass) {
        const int2 info = pass_info[pass];
        const int nfaces = info.y & 0xffff, nsub = info.y >> 16;
        // pair-uniform: both lanes of a pair take the same branches around the swaps
        int4 fe = make_int4(-1, -1, -1, -1);
        bool active = pr < nfaces;
        if (active) {
            fe = pass_faces[info.x + pr];
            const int a0 = g.F[3 * fe.x], a1 = g.F[3 * fe.x + 1], a2 = g.F[3 * fe.x + 2];
            for (int k = 0; k < (GAL ? 3 : 1); k++)
                active = active && !(same_vertex(g, t[k], a0) || same_vertex(g, t[k], a1) || same_vertex(g, t[k], a2));
        }
        cd accF[16], accX[16];
#pragma unroll
        for (int k = 0; k < 16; k++) accF[k] = cd{0, 0}, accX[k] = cd{0, 0};
        if (active) {
            const int fc = fe.x;
            if (side == 0) {
                const int a0 = g.F[3 * fc], a1 = g.F[3 * fc + 1], a2 = g.F[3 * fc + 2];
                double *dst = fgeo + pr;
                for (int k = 0; k < 3; k++) {
                    dst[k * kPairFaces] = g.V[3 * a0 + k];
                    dst[(3 + k) * kPairFaces] = g.V[3 * a1 + k];
                    dst[(6 + k) * kPairFaces] = g.V[3 * a2 + k];
                    dst[(9 + k) * kPairFaces] = g.N[3 * fc + k];
                }
                dst[12 * kPairFaces] = g.area[fc];
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            double4 qcur = pts[2 * side], wcur = pts[2 * side + 1];  // next round's point, prefetched
            for (int rd = 0; rd < nrounds; ++rd) {
                // the face slot is made opaque per round so that the loads below stay in the loop (LDS reads
                // instead of 13 loop-invariant values held in VGPRs)
                int slot = pr;
                asm volatile("" : "+v"(slot));
                const double *fg = fgeo + slot;
                const double y00 = fg[0], y01 = fg[kPairFaces], y02 = fg[2 * kPairFaces];
                const double y10 = fg[3 * kPairFaces], y11 = fg[4 * kPairFaces], y12 = fg[5 * kPairFaces];
                const double y20 = fg[6 * kPairFaces], y21 = fg[7 * kPairFaces], y22 = fg[8 * kPairFaces];
                const double ar = fg[12 * kPairFaces];
                const double4 Qo = qcur, Wt = wcur;  // this lane's point and its corner weights
                if (rd + 1 < nrounds) qcur = pts[4 * (rd + 1) + 2 * side], wcur = pts[4 * (rd + 1) + 2 * side + 1];
                const double r0 = Qo.x * y00 + Qo.y * y10 + Qo.z * y20 - x0;
                const double r1 = Qo.x * y01 + Qo.y * y11 + Qo.z * y21 - x1;
                const double r2 = Qo.x * y02 + Qo.y * y12 + Qo.z * y22 - x2;
                const double d2 = r0 * r0 + r1 * r1 + r2 * r2;
                const double ir = rsqrt_nr(d2), r = d2 * ir, ir2 = ir * ir;
                const double h[3] = {r0 * ir, r1 * ir, r2 * ir};
                const double wL = Qo.w * ar;
                const Expo E = expo_gh(r, f, &mtab);
                cd al, be, c1, c2, c4;
                if (DG) g_coef<false>(E, ir, ir2, m, al, be);  // 1/(4 pi rho) at the write-out
                if (DH) h_coef(E, ir, ir2, f, m, c1, c2, c4);
                // corner weights: the full corner (0 on lane A, 2 on lane B) takes the own and the partner point;
                // corner 1 takes the own point only, and the pair adds its two halves after the face
                const double fo = Wt.x, fp = Wt.y, xo = Qo.y;
                auto consume = [&](const cd &Kk, cd &aF, cd &aX) {
                    const cd Kq{pair_swap(Kk.x), pair_swap(Kk.y)};
                    fma_cr(aF, Kk, fo);
                    fma_cr(aF, Kq, fp);
                    fma_cr(aX, Kk, xo);
                };
                if (DG) {  // G block: alpha rh rh^T + beta I, packed 00 11 22 01 02 12 then beta
                    const double wr = wL * ir;
                    al = al * wr;
#pragma unroll
                    for (int a = 0; a < 3; a++)
#pragma unroll
                        for (int b = a; b < 3; b++) consume(al * (h[a] * h[b]), accF[sidx(a, b)], accX[sidx(a, b)]);
                    consume(be * wr, accF[6], accX[6]);
                }
                if (DH) {  // H block (without 1/4pi): rh_a u_b + n_a v_b + d I with u = c4 rn rh + c1 n, v = c2 rh, d = c2 rn
                    const double n[3] = {fg[9 * kPairFaces], fg[10 * kPairFaces], fg[11 * kPairFaces]};
                    const double rn = h[0] * n[0] + h[1] * n[1] + h[2] * n[2];
                    c1 = c1 * wL;
                    c2 = c2 * wL;
                    c4 = c4 * (wL * rn);
                    const cd d = c2 * rn;
                    cd u[3], v[3];
#pragma unroll
                    for (int b = 0; b < 3; b++) {
                        u[b] = c4 * h[b];
                        fma_cr(u[b], c1, n[b]);
                        v[b] = c2 * h[b];
                    }
#pragma unroll
                    for (int a = 0; a < 3; a++)
#pragma unroll
                        for (int b = 0; b < 3; b++) {
                            cd w = u[b] * h[a];
                            fma_cr(w, v[b], n[a]);
                            if (a == b) w += d;
                            consume(w, accF[7 + a * 3 + b], accX[7 + a * 3 + b]);
                        }
                }
            }
#pragma unroll
            for (int k = 0; k < 16; k++) accX[k] += cd{pair_swap(accX[k].x), pair_swap(accX[k].y)};
        }
        // scheduled updates of the vertex accumulators; corner code: slot | sub-step << 10, -1 = outside the patch
        const int cF = side ? fe.w : fe.y, cX = fe.z;
        // read-modify-write of n consecutive complex words of slot sl: all reads issued before the adds
        auto rmw = [&](int sl, int w0, const cd *v, int n) {
            double t[32];
#pragma unroll
            for (int k = 0; k < 16; k++)
                if (k < n) t[2 * k] = acc[(w0 + 2 * k) * kPairPV + sl], t[2 * k + 1] = acc[(w0 + 2 * k + 1) * kPairPV + sl];
#pragma unroll
            for (int k = 0; k < 16; k++)
                if (k < n) acc[(w0 + 2 * k) * kPairPV + sl] = t[2 * k] + v[k].x, acc[(w0 + 2 * k + 1) * kPairPV + sl] = t[2 * k + 1] + v[k].y;
        };
        for (int sub = 0; sub < nsub; ++sub) {
            __syncthreads();
            if (!active) continue;
            if (cF >= 0 && (cF >> 10) == sub) {
                if (DG && DH) rmw(cF & 0x3ff, 0, accF, 16);
                else if (DG) rmw(cF & 0x3ff, 0, accF, 7);
                else rmw(cF & 0x3ff, 14, accF + 7, 9);
            }
            if (cX >= 0 && (cX >> 10) == sub) {
                if (side == 0) {
                    if (DG) rmw(cX & 0x3ff, 0, accX, 7);
                } else if (DH) {
                    rmw(cX & 0x3ff, 14, accX + 7, 9);
                }
            }
        }
    }
    __syncthreads();
    // one thread per vertex slot writes its (row, column) G and H blocks
    for (int sl = tid; sl < kPairPV; sl += kPairThreads) {
        const int j = patch_verts[pc * kPairPV + sl];
        if (j < 0) continue;
        auto a = [&](int e) { return acc[e * kPairPV + sl]; };
        const long long obase = (long long)l * lstride + (long long)(3 * il) * ld + 3 * (long long)j;
#pragma unroll
        for (int p = 0; p < 3; p++)
#pragma unroll
            for (int b = 0; b < 3; b++) {
                const int e = sidx(p, b);
                if (DG) {
                    cd v{a(2 * e), a(2 * e + 1)};
                    if (p == b) v += cd{a(12), a(13)};
                    UG[obase + p * ld + b] = v * m.g_scale;
                }
                if (DH) UH[obase + p * ld + b] = cd{a(14 + 2 * (p * 3 + b)), a(15 + 2 * (p * 3 + b))} * kInv4Pi;
            }
    }
}

// BANDED: the distance-banded rule (CQBEM_QUAD_BANDED) is compiled in; without it the tile rule check is not.
template <bool DG, bool DH, bool GAL, bool BANDED>
__global__ void __launch_bounds__(kPairThreads, 2) k_colloc_gh_pair(Geom g, const double4 *__restrict__ pts, int nrounds,
                                                                  const Freq *__restrict__ freqs, int l0, Mat m, int row0,
                                                                  const double4 *__restrict__ tpts,
                                                                  const int *__restrict__ patch_pass,
                                                                  const int2 *__restrict__ pass_info,
                                                                  const int4 *__restrict__ pass_faces,
                                                                  const int *__restrict__ patch_verts, cd *__restrict__ UG,
                                                                  cd *__restrict__ UH, long long ld, long long lstride,
                                                                  Bands bands) {
    __shared__ double acc[kPairWords * kPairPV];      // [word][slot]
    __shared__ double fgeo[kFaceWords * kPairFaces];  // [word][face of pass]: re-read every round, not held in VGPRs
    __shared__ MathTables mtab;                       // tables of the exponentials
    load_math_tables(&mtab);
    pair_tile<DG, DH, GAL>(g, pts, nrounds, BANDED ? bands : Bands{}, freqs, l0, m, row0, tpts, patch_pass, pass_info,
                           pass_faces, patch_verts, UG, UH, ld, lstride, blockIdx.x, blockIdx.y, blockIdx.z, acc, fgeo, mtab);
}

// ------------------------------------------------------------------------
// Several objects' G/H regular-face work in one launch: each piece (an object's row slab) contributes
// npatches x nl x nrows tiles, numbered from its `base`, and each tile runs exactly the computation of
// k_colloc_gh_pair's block for it, so every object's spectra are bit-identical to its own launch.  Two forms:
//   k_scene_gh_tiles: one block per tile; the hardware's workgroup dispatcher hands tiles to CUs in tile order as
//     they free up (pieces are ordered largest first by the host).  Same registers as the single-object kernel
//     plus the piece lookup: 2 waves per SIMD.
//   k_scene_gh_pair: persistent blocks (2 per CU) that take tile numbers from a device counter with one atomic add
//     per tile.  The loop-carried state costs 8 registers beyond 256, i.e. 1 wave per SIMD (CQBEM_SCENE_KERNEL=
//     persistent selects it, for comparison).
// ------------------------------------------------------------------------
template <bool DG, bool DH>
__device__ __attribute__((always_inline)) inline void scene_tile(const ScenePiece *__restrict__ pieces, int npieces, unsigned t, double *acc, double *fgeo,
                                  const MathTables &mtab) {
    int pi = 0;
    while (pi + 1 < npieces && pieces[pi + 1].base <= t) ++pi;
    const ScenePiece *P = pieces + pi;
    const unsigned local = t - P->base, np = (unsigned)P->npatches, nl = (unsigned)P->nl;
    const unsigned rest = local / np;
    pair_tile<DG, DH, false>(P->g, P->pts, P->nrounds, P->bands, P->freqs, 0, P->m, P->row0, nullptr, P->patch_pass, P->pass_info,
                             P->pass_faces, P->patch_verts, P->UG, P->UH, P->ld, P->lstride, (int)(local - rest * np),
                             (int)(rest % nl), (int)(rest / nl), acc, fgeo, mtab);
}

template <bool DG, bool DH>
__global__ void __launch_bounds__(kPairThreads, 2) k_scene_gh_tiles(const ScenePiece *__restrict__ pieces, int npieces,
                                                                  unsigned t0) {
    __shared__ double acc[kPairWords * kPairPV];
    __shared__ double fgeo[kFaceWords * kPairFaces];
    __shared__ MathTables mtab;
    load_math_tables(&mtab);
    scene_tile<DG, DH>(pieces, npieces, t0 + blockIdx.x, acc, fgeo, mtab);
}

template <bool DG, bool DH>
__global__ void __launch_bounds__(kPairThreads, 2) k_scene_gh_pair(const ScenePiece *__restrict__ pieces, int npieces,
                                                                 unsigned ntiles, unsigned *__restrict__ counter) {
    __shared__ double acc[kPairWords * kPairPV];
    __shared__ double fgeo[kFaceWords * kPairFaces];
    __shared__ MathTables mtab;
    __shared__ unsigned next_tile;
    load_math_tables(&mtab);
    for (;;) {
        __syncthreads();  // the previous tile's write-out has read acc; next_tile is free
        if (threadIdx.x == 0) next_tile = atomicAdd(counter, 1u);
        __syncthreads();
        const unsigned t = __builtin_amdgcn_readfirstlane(next_tile);
        if (t >= ntiles) break;  // uniform across the block: every wave leaves
        scene_tile<DG, DH>(pieces, npieces, t, acc, fgeo, mtab);
    }
}

// Zeroes the tile counter of a scene launch (a kernel, not a memset: it is captured in the scene's graph).
__global__ void k_zero_counter(unsigned *counter) {
    if (threadIdx.x == 0) *counter = 0u;
}

// ------------------------------------------------------------------------
// B_trans and B_angular, regular faces.  Grid: (nl, rows), 256 threads; lanes
// stride over the faces, then a wave shuffle reduction and one LDS pass.
// ------------------------------------------------------------------------
constexpr int kBWords = 48;  // B_trans (6 sym complex) + B_angular (18 complex), as doubles

// GAL: the rows are test points tpts[row0 + y] (x, y, z, face) whose face's corners are excluded, written to
// rows 3y..3y+2 of an nv3-row buffer; otherwise collocation vertex i = row0 + y, written to rows 3i..3i+2.
template <bool GAL>
__global__ void __launch_bounds__(256) k_colloc_b(Geom g, const double4 *__restrict__ quad, int npts,
                                                   const Freq *__restrict__ freqs, int l0, Mat m, int row0,
                                                   const double4 *__restrict__ tpts, cd *__restrict__ Bt,
                                                   cd *__restrict__ Ba, long long nv3, FaceBands fb) {
    __shared__ double red[4][kBWords];
    __shared__ MathTables mtab;  // tables of the exponentials
    load_math_tables(&mtab);
    __syncthreads();
    const int l = blockIdx.x;
    const int i = row0 + blockIdx.y;
    const Freq f = freqs[l0 + l];
    double x0, x1, x2;
    int t[3] = {i, i, i};
    if (GAL) {
        const double4 T = tpts[i];
        x0 = T.x, x1 = T.y, x2 = T.z;
        const int f1 = (int)T.w;
        t[0] = g.F[3 * f1], t[1] = g.F[3 * f1 + 1], t[2] = g.F[3 * f1 + 2];
    } else {
        x0 = g.V[3 * i], x1 = g.V[3 * i + 1], x2 = g.V[3 * i + 2];
    }
    const int io = GAL ? (int)blockIdx.y : i;
    cd bt[6], ba[18];
#pragma unroll
    for (int k = 0; k < 6; k++) bt[k] = cd{0, 0};
#pragma unroll
    for (int k = 0; k < 18; k++) ba[k] = cd{0, 0};
    for (int fc = threadIdx.x; fc < g.nf; fc += 256) {
        const int a0 = g.F[3 * fc], a1 = g.F[3 * fc + 1], a2 = g.F[3 * fc + 2];
        bool touch = false;
        for (int k = 0; k < (GAL ? 3 : 1); k++)
            touch = touch || same_vertex(g, t[k], a0) || same_vertex(g, t[k], a1) || same_vertex(g, t[k], a2);
        if (touch) continue;
        const double *p0 = g.V + 3 * a0, *p1 = g.V + 3 * a1, *p2 = g.V + 3 * a2;
        const double n[3] = {g.N[3 * fc], g.N[3 * fc + 1], g.N[3 * fc + 2]};
        const double ar = g.area[fc];
        const double4 *qf = quad;  // the face's rule
        int nq = npts;
        if (fb.sphere && beyond_sphere(x0, x1, x2, fb.sphere[fc])) qf = fb.far_quad, nq = fb.far_npts;
        cd mt[3] = {{0, 0}, {0, 0}, {0, 0}}, m2 = {0, 0}, M[6], S = {0, 0}, S3 = {0, 0};
#pragma unroll
        for (int k = 0; k < 6; k++) M[k] = cd{0, 0};
        for (int q = 0; q < nq; ++q) {
            const double4 Q = qf[q];
            const double r0 = Q.x * p0[0] + Q.y * p1[0] + Q.z * p2[0] - x0;
            const double r1 = Q.x * p0[1] + Q.y * p1[1] + Q.z * p2[1] - x1;
            const double r2 = Q.x * p0[2] + Q.y * p1[2] + Q.z * p2[2] - x2;
            const double d2 = r0 * r0 + r1 * r1 + r2 * r2;
            const double ir = rsqrt_nr(d2), r = d2 * ir, ir2 = ir * ir;
            const double h[3] = {r0 * ir, r1 * ir, r2 * ir};
            const double w = Q.w * ar;
            const Expo E = expo(r, f, &mtab);
            cd t1, t2, u1, u12, u3;
            b_coef_8pi(E, ir, ir2, f, m, t1, t2, u1, u12, u3);  // 1/(8 pi) at the write-out
            const double rn = h[0] * n[0] + h[1] * n[1] + h[2] * n[2];
            t1 = t1 * w;
            u1 = u1 * w;
#pragma unroll
            for (int a = 0; a < 3; a++) fma_cr(mt[a], t1, h[a]);
            fma_cr(m2, t2, w * rn);
#pragma unroll
            for (int a = 0; a < 3; a++)
#pragma unroll
                for (int b = a; b < 3; b++) fma_cr(M[sidx(a, b)], u1, h[a] * h[b]);
            fma_cr(S, u12, w);
            fma_cr(S3, u3, w);
        }
#pragma unroll
        for (int a = 0; a < 3; a++)
#pragma unroll
            for (int b = a; b < 3; b++) {
                cd v = {0, 0};
                fma_cr(v, mt[a], n[b]);
                fma_cr(v, mt[b], n[a]);
                if (a == b) v += m2;
                bt[sidx(a, b)] += v;
#pragma unroll
                for (int k = 0; k < 3; k++) {
                    cd u = {0, 0};
                    fma_cr(u, M[sidx(a, k)], n[b]);
                    fma_cr(u, M[sidx(b, k)], n[a]);
                    fma_cr(u, S, (b == k ? n[a] : 0.0) + (a == k ? n[b] : 0.0));
                    if (a == b) fma_cr(u, S3, n[k]);
                    ba[sidx(a, b) * 3 + k] += u;
                }
            }
    }
    // wave reduction of the 48 partial sums, then the 4 waves through LDS (fixed order)
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
    for (int k = 0; k < kBWords; k++) {
        double v = k < 12 ? (k & 1 ? bt[k >> 1].y : bt[k >> 1].x) : (k & 1 ? ba[(k - 12) >> 1].y : ba[(k - 12) >> 1].x);
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
        if (lane == 0) red[wave][k] = v;
    }
    __syncthreads();
    if (threadIdx.x < kBWords) {
        const int k = threadIdx.x;
        const double v = ((red[0][k] + red[1][k]) + (red[2][k] + red[3][k])) * kInv8Pi;
        // scatter the packed symmetric entries to their full positions
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

// ------------------------------------------------------------------------
// Singular faces (faces touching the collocation vertex).  One lane per
// (row vertex, frequency); the lane owns those rows and adds into them.
//   KIND 0: G, weakly singular polar rule (integrator_collocation.hpp:19-63)
//   KIND 1: B_trans + B_angular, same polar rule
//   KIND 2: H, Guiggiani finite-part rule (integrator_collocation.hpp:65-124)
// ------------------------------------------------------------------------
constexpr int kSingMaxFaces = 16;  // singular faces per batch of one (row, frequency) wavefront

// Sum of v over lanes lane .. lane + gs - 1, in lane order (the group leader's value is the group's sum).
__device__ inline cd group_sum(cd v, int lane, int gs) {
    double sx = v.x, sy = v.y;
    for (int q = 1; q < gs; q++) sx += __shfl(v.x, lane + q), sy += __shfl(v.y, lane + q);
    return cd{sx, sy};
}

__device__ inline void unit3(double v[3]) {
    const double n = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    if (n > 0) v[0] /= n, v[1] /= n, v[2] /= n;
}

template <int KIND>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) k_colloc_singular(Geom g, const double2 *__restrict__ line, int nline,
                                                         const Freq *__restrict__ freqs, int l0, int nl, Mat m, int row0,
                                                         int nrows, cd *__restrict__ U, long long ld, long long lstride,
                                                         cd *__restrict__ Bt, cd *__restrict__ Ba, long long nv3) {
    constexpr int NE = KIND == 1 ? 36 : 27;  // per face: B 9 + 27 entries; G/H 3 corners x 3 x 3
    __shared__ cd fsum[kSingMaxFaces][NE];
    __shared__ int fj[kSingMaxFaces][3], umap[kSingMaxFaces][3], ucol[3 * kSingMaxFaces + 1];
    const int il = blockIdx.x / nl, l = blockIdx.x % nl;
    const int i = row0 + il;
    const Freq f = freqs[l0 + l];
    // GS lanes per singular face, lane t of a group takes the angular points t, t + GS, ...
    const int GS = nline < 64 / kSingMaxFaces ? 64 / kSingMaxFaces : (nline > 64 ? 64 : nline);
    const int FB = 64 / GS;  // faces per batch
    const int lane = threadIdx.x, slot = lane / GS, t = lane % GS;
    const int fbeg = g.sg_ptr[i], fend = g.sg_ptr[i + 1];
    for (int fb = fbeg; fb < fend; fb += FB) {
        const int nb = min(FB, fend - fb);
        const bool active = slot < nb;
        const int pk = g.sg_face[active ? fb + slot : fbeg];
        const int fc = pk >> 2, c = pk & 3;
        const int j[3] = {g.F[3 * fc + c], g.F[3 * fc + (c + 1) % 3], g.F[3 * fc + (c + 2) % 3]};
        double y1[3], y2[3], y3[3], n[3];
        for (int k = 0; k < 3; k++)
            y1[k] = g.V[3 * j[0] + k], y2[k] = g.V[3 * j[1] + k], y3[k] = g.V[3 * j[2] + k], n[k] = g.N[3 * fc + k];
        double v1[3], v2[3];
        for (int k = 0; k < 3; k++) v1[k] = y2[k] - y1[k], v2[k] = y3[k] - y1[k];
        const bool leader = active && t == 0;  // holds the face's sums after group_sum
        if (KIND == 0 || KIND == 1) {
            // polar frame around y1
            double u1[3] = {v1[0], v1[1], v1[2]}, u2[3] = {v2[0], v2[1], v2[2]};
            unit3(u1), unit3(u2);
            const double xh[3] = {u1[0], u1[1], u1[2]};
            double cxu[3] = {xh[1] * u2[2] - xh[2] * u2[1], xh[2] * u2[0] - xh[0] * u2[2], xh[0] * u2[1] - xh[1] * u2[0]};
            double yh[3] = {cxu[1] * xh[2] - cxu[2] * xh[1], cxu[2] * xh[0] - cxu[0] * xh[2], cxu[0] * xh[1] - cxu[1] * xh[0]};
            unit3(yh);
            double e[3] = {v1[0] - v2[0], v1[1] - v2[1], v1[2] - v2[2]};
            unit3(e);
            const double d = v1[0] * e[0] + v1[1] * e[1] + v1[2] * e[2];
            double v12[3] = {v1[0] - d * e[0], v1[1] - d * e[1], v1[2] - d * e[2]};
            const double a = sqrt(v12[0] * v12[0] + v12[1] * v12[1] + v12[2] * v12[2]);
            unit3(v12);
            const double alpha = acos(u1[0] * u2[0] + u1[1] * u2[1] + u1[2] * u2[2]);
            // beta = angle(v12, u1), which the reference takes as acos of the clamped dot product
            // (integrator_collocation.hpp:36); atan2(|v12 x u1|, v12 . u1) is the same angle without acos's
            // loss of half the digits near 0 (a right angle at y2, common on cube meshes)
            const double bx = v12[1] * u1[2] - v12[2] * u1[1], by = v12[2] * u1[0] - v12[0] * u1[2],
                         bz = v12[0] * u1[1] - v12[1] * u1[0];
            const double beta = atan2(sqrt(bx * bx + by * by + bz * bz), v12[0] * u1[0] + v12[1] * u1[1] + v12[2] * u1[2]);
            // barycentric projection setup (barycentric_coordinates.cuh:15-38)
            const double d00 = v1[0] * v1[0] + v1[1] * v1[1] + v1[2] * v1[2];
            const double d01 = v1[0] * v2[0] + v1[1] * v2[1] + v1[2] * v2[2];
            const double d11 = v2[0] * v2[0] + v2[1] * v2[1] + v2[2] * v2[2];
            const double iden = 1.0 / (d00 * d11 - d01 * d01);
            cd ga[3][6], gb[3];
            cd mt[3], m2, M[6], S, S3;
            if (KIND == 0) {
                for (int k = 0; k < 3; k++) {
                    gb[k] = cd{0, 0};
                    for (int e2 = 0; e2 < 6; e2++) ga[k][e2] = cd{0, 0};
                }
            } else {
                for (int k = 0; k < 3; k++) mt[k] = cd{0, 0};
                for (int k = 0; k < 6; k++) M[k] = cd{0, 0};
                m2 = S = S3 = cd{0, 0};
            }
            for (int it = active ? t : nline; it < nline; it += GS) {
                const double th = alpha * line[it].x, wth = alpha * line[it].y;
                double sth, cth;
                sincos(th, &sth, &cth);
                const double h[3] = {xh[0] * cth + yh[0] * sth, xh[1] * cth + yh[1] * sth, xh[2] * cth + yh[2] * sth};
                const double rmax = a / cos(th - beta);
                const double rn = h[0] * n[0] + h[1] * n[1] + h[2] * n[2];
                for (int ir_ = 0; ir_ < nline; ++ir_) {
                    const double r = rmax * line[ir_].x, w = wth * rmax * line[ir_].y;
                    const double ir = 1.0 / r, ir2 = ir * ir;
                    const Expo E = expo(r, f);
                    if (KIND == 0) {
                        // clamped projection barycentrics of y = y1 + r h
                        const double c0 = r * h[0], c1 = r * h[1], c2 = r * h[2];
                        const double d20 = c0 * v1[0] + c1 * v1[1] + c2 * v1[2];
                        const double d21 = c0 * v2[0] + c1 * v2[1] + c2 * v2[2];
                        double L1 = (d11 * d20 - d01 * d21) * iden, L2 = (d00 * d21 - d01 * d20) * iden;
                        double L0 = 1.0 - (L1 + L2);
                        L0 = fmin(1.0, fmax(0.0, L0)), L1 = fmin(1.0, fmax(0.0, L1)), L2 = fmin(1.0, fmax(0.0, L2));
                        const double Lb[3] = {L0, L1, L2};
                        cd al, be;
                        g_coef(E, ir, ir2, m, al, be);  // r * u_hat coefficients
                        for (int k = 0; k < 3; k++) {
                            const double wk = w * Lb[k];
                            const cd ak = al * wk;
                            fma_cr(gb[k], be, wk);
                            fma_cr(ga[k][0], ak, h[0] * h[0]);
                            fma_cr(ga[k][1], ak, h[1] * h[1]);
                            fma_cr(ga[k][2], ak, h[2] * h[2]);
                            fma_cr(ga[k][3], ak, h[0] * h[1]);
                            fma_cr(ga[k][4], ak, h[0] * h[2]);
                            fma_cr(ga[k][5], ak, h[1] * h[2]);
                        }
                    } else {
                        cd t1, t2, u1c, u12, u3;
                        b_coef(E, ir, ir2, f, m, t1, t2, u1c, u12, u3);
                        const double wr = w * r;  // polar rules integrate r * kernel
                        t1 = t1 * wr;
                        u1c = u1c * wr;
                        for (int a2 = 0; a2 < 3; a2++) fma_cr(mt[a2], t1, h[a2]);
                        fma_cr(m2, t2, wr * rn);
                        for (int a2 = 0; a2 < 3; a2++)
                            for (int b = a2; b < 3; b++) fma_cr(M[sidx(a2, b)], u1c, h[a2] * h[b]);
                        fma_cr(S, u12, wr);
                        fma_cr(S3, u3, wr);
                    }
                }
            }
            if (KIND == 0) {
                for (int k = 0; k < 3; k++) {
                    gb[k] = group_sum(gb[k], lane, GS);
                    for (int e2 = 0; e2 < 6; e2++) ga[k][e2] = group_sum(ga[k][e2], lane, GS);
                }
                if (leader)
                    for (int k = 0; k < 3; k++)
                        for (int a2 = 0; a2 < 3; a2++)
                            for (int b = 0; b < 3; b++) {
                                cd v = ga[k][sidx(a2, b)];
                                if (a2 == b) v += gb[k];
                                fsum[slot][k * 9 + a2 * 3 + b] = v;
                            }
            } else {
                for (int k = 0; k < 3; k++) mt[k] = group_sum(mt[k], lane, GS);
                for (int k = 0; k < 6; k++) M[k] = group_sum(M[k], lane, GS);
                m2 = group_sum(m2, lane, GS), S = group_sum(S, lane, GS), S3 = group_sum(S3, lane, GS);
                if (leader)
                for (int a2 = 0; a2 < 3; a2++)
                    for (int b = 0; b < 3; b++) {
                        cd v = {0, 0};
                        fma_cr(v, mt[a2], n[b]);
                        fma_cr(v, mt[b], n[a2]);
                        if (a2 == b) v += m2;
                        fsum[slot][a2 * 3 + b] = v;
                        for (int k = 0; k < 3; k++) {
                            cd u = {0, 0};
                            fma_cr(u, M[sidx(a2, k)], n[b]);
                            fma_cr(u, M[sidx(b, k)], n[a2]);
                            fma_cr(u, S, (b == k ? n[a2] : 0.0) + (a2 == k ? n[b] : 0.0));
                            if (a2 == b) fma_cr(u, S3, n[k]);
                            fsum[slot][9 + a2 * 9 + b * 3 + k] = u;
                        }
                    }
            }
        } else {
            // Guiggiani: rho-theta on the unit right triangle around y1
            double cr[3] = {v1[1] * v2[2] - v1[2] * v2[1], v1[2] * v2[0] - v1[0] * v2[2], v1[0] * v2[1] - v1[1] * v2[0]};
            const double J = sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
            cd out[27];
            for (int k = 0; k < 27; k++) out[k] = cd{0, 0};
            const double half_pi = 1.5707963267948966192;
            for (int it = active ? t : nline; it < nline; it += GS) {
                const double th = half_pi * line[it].x, wth = half_pi * line[it].y;
                double st, ct;
                sincos(th, &st, &ct);
                double Av[3] = {ct * v1[0] + st * v2[0], ct * v1[1] + st * v2[1], ct * v1[2] + st * v2[2]};
                const double A = sqrt(Av[0] * Av[0] + Av[1] * Av[1] + Av[2] * Av[2]);
                const double Ah[3] = {Av[0] / A, Av[1] / A, Av[2] / A};
                const double rho_bar = 1.0 / (cos(th - 0.78539816339744830962) * 1.4142135623730950488);
                const double rn = Ah[0] * n[0] + Ah[1] * n[1] + Ah[2] * n[2];
                // cancelling term f_pq = (1/4pi) kap (Ah_p n_q - Ah_q n_p) / A^2 * J, basis (1,0,0)
                const double fs = kInv4Pi * m.kap * J / (A * A);
                cd S1[3], S2[3], S4[3];
                for (int k = 0; k < 3; k++) S1[k] = S2[k] = S4[k] = cd{0, 0};
                double sinv = 0.0;
                for (int ir_ = 0; ir_ < nline; ++ir_) {
                    const double rho = rho_bar * line[ir_].x, w = rho_bar * line[ir_].y;
                    const double r = rho * A;
                    const double ir = 1.0 / r, ir2 = ir * ir;
                    const Expo E = expo(r, f);
                    cd c1, c2, c4;
                    h_coef(E, ir, ir2, f, m, c1, c2, c4);
                    const double Lb[3] = {1.0 - rho * (ct + st), rho * ct, rho * st};
                    for (int k = 0; k < 3; k++) {
                        const double wk = w * Lb[k] * J * rho;
                        fma_cr(S1[k], c1, wk);
                        fma_cr(S2[k], c2, wk);
                        fma_cr(S4[k], c4, wk);
                    }
                    sinv += w / rho;
                }
                const double lg = log(rho_bar * A);
                for (int p = 0; p < 3; p++)
                    for (int qq = 0; qq < 3; qq++) {
                        const double fpq = fs * (Ah[p] * n[qq] - Ah[qq] * n[p]);
                        for (int k = 0; k < 3; k++) {
                            cd v = {0, 0};
                            fma_cr(v, S1[k], Ah[p] * n[qq]);
                            fma_cr(v, S2[k], n[p] * Ah[qq] + (p == qq ? rn : 0.0));
                            fma_cr(v, S4[k], rn * Ah[p] * Ah[qq]);
                            v = v * kInv4Pi;
                            if (k == 0) v.x += fpq * (lg - sinv);
                            fma_cr(out[(p * 3 + qq) * 3 + k], v, wth);
                        }
                    }
            }
            for (int e = 0; e < 27; e++) out[e] = group_sum(out[e], lane, GS);
            if (leader)
                for (int p = 0; p < 3; p++)
                    for (int qq = 0; qq < 3; qq++)
                        for (int k = 0; k < 3; k++) fsum[slot][k * 9 + p * 3 + qq] = out[(p * 3 + qq) * 3 + k];
        }
        if (leader)
            for (int k = 0; k < 3; k++) fj[slot][k] = j[k];
        __syncthreads();
        // add the batch's faces into the row: each output entry sums its faces in face order, then one
        // read-modify-write per entry (no dependent chains through memory)
        if (KIND == 1) {
            if (lane < 36) {
                cd sum = fsum[0][lane];
                for (int s2 = 1; s2 < nb; s2++) sum += fsum[s2][lane];
                cd *o = lane < 9 ? Bt + (long long)l * nv3 * 3 + (long long)(3 * i) * 3 + lane
                                 : Ba + (long long)l * nv3 * 9 + (long long)(3 * i) * 9 + (lane - 9);
                *o += sum;
            }
        } else {
            // distinct columns of the batch in order of first appearance
            if (lane == 0) {
                int nu = 0;
                for (int s2 = 0; s2 < nb; s2++)
                    for (int k = 0; k < 3; k++) {
                        int u = 0;
                        while (u < nu && ucol[u] != fj[s2][k]) u++;
                        if (u == nu) ucol[nu++] = fj[s2][k];
                        umap[s2][k] = u;
                    }
                ucol[3 * kSingMaxFaces] = nu;
            }
            __syncthreads();
            const int nu = ucol[3 * kSingMaxFaces];
            for (int item = lane; item < 9 * nu; item += 64) {
                const int u = item / 9, e = item % 9, a2 = e / 3, b = e % 3;
                cd sum = {0, 0};
                for (int s2 = 0; s2 < nb; s2++)
                    for (int k = 0; k < 3; k++)
                        if (umap[s2][k] == u) sum += fsum[s2][k * 9 + e];
                U[(long long)l * lstride + (long long)(3 * il + a2) * ld + 3 * (long long)ucol[u] + b] += sum;
            }
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------------------
// Inverse CQ transform: W_n[e] = sum_l Re(X_l[e] * T[n][l]), with the table
// T[n][l] = R^-n / L * mult_l * zeta^(-n l) (G.cu:166-179).  One lane per
// entry, NC weights per pass; the table index is uniform (scalar loads).
// ------------------------------------------------------------------------
// Inverse CQ transform of one chunk of NC weights n0 .. n0+NC-1: out[n][e] = sum_l Re(X_l(e) T(n, l)), with
// T(n, l) = (2 or 1) R^-n / L exp(-2 pi i n l / L) (the half-spectrum table).  One lane per entry e; the
// chunk's table lives in LDS (read as broadcasts) and the spectrum is read LC frequencies at a time so
// that LC loads per lane are in flight.
constexpr int kTransformNC = 32, kTransformLC = 8, kTransformMaxLf = 64;

// OutT: double, or float for resident weights stored in single precision (the sum is formed in double and
// rounded once).
template <int NC, int LC, class OutT>
__global__ void __launch_bounds__(256) k_transform(const cd *__restrict__ spec, long long E, long long spec_ls, int Lf,
                                                    const double2 *__restrict__ tab, int L, int n0,
                                                    OutT *__restrict__ out, long long out_nstride) {
    __shared__ double2 T[NC * kTransformMaxLf];  // [l][k]
    for (int t = threadIdx.x; t < NC * Lf; t += 256) {
        const int l = t / NC, k = t % NC;
        T[t] = n0 + k < L ? tab[(long long)(n0 + k) * Lf + l] : make_double2(0.0, 0.0);
    }
    __syncthreads();
    const long long stride = (long long)gridDim.x * 256;
    for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < E; e += stride) {
        double acc[NC];
#pragma unroll
        for (int k = 0; k < NC; k++) acc[k] = 0.0;
        for (int l0 = 0; l0 < Lf; l0 += LC) {
            cd X[LC];
#pragma unroll
            for (int c = 0; c < LC; c++) {
                // each spectrum value is read once: streaming (non-temporal) loads
                const double *src = reinterpret_cast<const double *>(spec + (long long)(l0 + c) * spec_ls + e);
                X[c] = l0 + c < Lf ? cd{__builtin_nontemporal_load(src), __builtin_nontemporal_load(src + 1)} : cd{0.0, 0.0};
            }
#pragma unroll
            for (int c = 0; c < LC; c++) {
                if (l0 + c >= Lf) break;
                const double2 *Tl = T + (l0 + c) * NC;
#pragma unroll
                for (int k = 0; k < NC; k++) acc[k] = fma(X[c].x, Tl[k].x, fma(-X[c].y, Tl[k].y, acc[k]));
            }
        }
#pragma unroll
        for (int k = 0; k < NC; k++) {
            const int n = n0 + k;
            if (n < L) __builtin_nontemporal_store((OutT)acc[k], out + (long long)n * out_nstride + e);  // written once
        }
    }
}

// The same sums for two adjacent entries per lane (E, spec_ls and out_nstride even, out 16-byte aligned): each
// frequency's two complex values in one 32-byte read, each weight's two values in one 16-byte store (the store
// width HBM writes best), and a grid that covers every pair at once.  Each weight is k_transform's sum, term for
// term.  LC = 2 frequencies per batch of loads keeps the kernel within 168 VGPRs: 3 waves per SIMD (with 4 it takes
// 181 VGPRs and 2 waves, 2.6% slower).
typedef double d2v __attribute__((ext_vector_type(2)));
typedef float f2v __attribute__((ext_vector_type(2)));
template <class OutT> struct Pair2;
template <> struct Pair2<double> {
    typedef d2v type;
};
template <> struct Pair2<float> {
    typedef f2v type;
};
template <int LC, class OutT>
__global__ void __launch_bounds__(256, 3) k_transform_pair(const cd *__restrict__ spec, long long E, long long spec_ls, int Lf,
                                                        const double2 *__restrict__ tab, int L, int n0,
                                                        OutT *__restrict__ out, long long out_nstride) {
    constexpr int NC = kTransformNC;
    __shared__ double2 T[NC * kTransformMaxLf];  // [l][k]
    for (int t = threadIdx.x; t < NC * Lf; t += 256) {
        const int l = t / NC, k = t % NC;
        T[t] = n0 + k < L ? tab[(long long)(n0 + k) * Lf + l] : make_double2(0.0, 0.0);
    }
    __syncthreads();
    const long long e = ((long long)blockIdx.x * 256 + threadIdx.x) * 2;
    if (e >= E) return;
    double a0[NC], a1[NC];
#pragma unroll
    for (int k = 0; k < NC; k++) a0[k] = a1[k] = 0.0;
    for (int l0 = 0; l0 < Lf; l0 += LC) {
        d2v X0[LC], X1[LC];
#pragma unroll
        for (int c = 0; c < LC; c++) {
            // each spectrum value is read once: streaming (non-temporal) loads
            const d2v *src = reinterpret_cast<const d2v *>(spec + (long long)(l0 + c) * spec_ls + e);
            if (l0 + c < Lf) {
                X0[c] = __builtin_nontemporal_load(src);
                X1[c] = __builtin_nontemporal_load(src + 1);
            } else {
                X0[c] = d2v{0.0, 0.0}, X1[c] = d2v{0.0, 0.0};
            }
        }
#pragma unroll
        for (int c = 0; c < LC; c++) {
            if (l0 + c >= Lf) break;
            const double2 *Tl = T + (l0 + c) * NC;
#pragma unroll
            for (int k = 0; k < NC; k++) {
                a0[k] = fma(X0[c].x, Tl[k].x, fma(-X0[c].y, Tl[k].y, a0[k]));
                a1[k] = fma(X1[c].x, Tl[k].x, fma(-X1[c].y, Tl[k].y, a1[k]));
            }
        }
    }
    typedef typename Pair2<OutT>::type V2;
#pragma unroll
    for (int k = 0; k < NC; k++) {
        const int n = n0 + k;
        if (n < L) {
            const V2 w = {(OutT)a0[k], (OutT)a1[k]};
            __builtin_nontemporal_store(w, reinterpret_cast<V2 *>(out + (long long)n * out_nstride + e));  // written once
        }
    }
}

// Same for Lf > kTransformMaxLf, with the table read from global memory.
template <int NC, class OutT>
__global__ void __launch_bounds__(256) k_transform_wide(const cd *__restrict__ spec, long long E, long long spec_ls, int Lf,
                                                         const double2 *__restrict__ tab, int L, int n0,
                                                         OutT *__restrict__ out, long long out_nstride) {
    const long long stride = (long long)gridDim.x * 256;
    for (long long e = (long long)blockIdx.x * 256 + threadIdx.x; e < E; e += stride) {
        double acc[NC];
#pragma unroll
        for (int k = 0; k < NC; k++) acc[k] = 0.0;
        for (int l = 0; l < Lf; l++) {
            const cd X = spec[(long long)l * spec_ls + e];
#pragma unroll
            for (int k = 0; k < NC; k++) {
                const int n = n0 + k;
                if (n < L) {
                    const double2 Tv = tab[(long long)n * Lf + l];
                    acc[k] = fma(X.x, Tv.x, fma(-X.y, Tv.y, acc[k]));
                }
            }
        }
#pragma unroll
        for (int k = 0; k < NC; k++) {
            const int n = n0 + k;
            if (n < L) out[(long long)n * out_nstride + e] = (OutT)acc[k];
        }
    }
}

// ------------------------------------------------------------------------
// History term of one CQ time step on the resident matrices:
//   b = sum_{i=1..m} (G_i p[nh-i] - H_i u[nh-i] + B_i acc[nh-i]).
// One wavefront per row; lanes stride the columns with kHistUnroll independent loads per matrix in
// flight (the 2 (L-1) row streams are read once, HBM-bound); fixed-order wave reduction.
// ------------------------------------------------------------------------
constexpr int kHistUnroll = 4;

template <bool HG, bool HH>
__device__ inline double history_row(const double *__restrict__ gr, const double *__restrict__ hr, const double *__restrict__ pv,
                                     const double *__restrict__ uv, long long n3, int lane, double s) {
    long long c = lane;
    for (; c + 64 * (kHistUnroll - 1) < n3; c += 64 * kHistUnroll) {
        double g[kHistUnroll], h[kHistUnroll], x[kHistUnroll], y[kHistUnroll];
#pragma unroll
        for (int k = 0; k < kHistUnroll; k++) {
            if (HG) g[k] = gr[c + 64 * k], x[k] = pv[c + 64 * k];
            if (HH) h[k] = hr[c + 64 * k], y[k] = uv[c + 64 * k];
        }
#pragma unroll
        for (int k = 0; k < kHistUnroll; k++) {
            if (HG) s = fma(g[k], x[k], s);
            if (HH) s = fma(-h[k], y[k], s);
        }
    }
    for (; c < n3; c += 64) {
        if (HG) s = fma(gr[c], pv[c], s);
        if (HH) s = fma(-hr[c], uv[c], s);
    }
    return s;
}

template <bool HG, bool HH>
__global__ void __launch_bounds__(256) k_history(const double *__restrict__ G, const double *__restrict__ H,
                                                  const double *__restrict__ B, const double *__restrict__ p,
                                                  const double *__restrict__ u, const double *__restrict__ acc,
                                                  long long n3, int m, int nh, double *__restrict__ b) {
    const long long r = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (r >= n3) return;
    double s = 0.0;
    for (int i = 1; i <= m; i++) {
        const long long ro = (long long)i * n3 * n3 + r * n3;
        s = history_row<HG, HH>(HG ? G + ro : nullptr, HH ? H + ro : nullptr, HG ? p + (long long)(nh - i) * n3 : nullptr,
                                HH ? u + (long long)(nh - i) * n3 : nullptr, n3, lane, s);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if (lane == 0) {
        if (B)
            for (int i = 1; i <= m; i++) {
                const double *br = B + ((long long)i * n3 + r) * 6, *av = acc + (long long)(nh - i) * 6;
#pragma unroll
                for (int k = 0; k < 6; k++) s = fma(br[k], av[k], s);
            }
        b[r] = s;
    }
}

// The same history sum for an even n3 with 16-byte aligned matrices and vectors (every row then starts on a 16-byte
// boundary).  Each wavefront takes R consecutive rows, so a chunk of p and u loaded once serves R matrix rows
// (the vectors are re-read from L2 by every wavefront: with one row per wavefront that traffic equals the matrix
// stream).  The matrix rows are streamed with 16-byte non-temporal loads: each weight is read once per step.  The
// B terms are spread over the lanes before the wave reduction.  Matrix i starts at G + i * mat_stride (n3^2 for the
// history; 0 with m = nh = 1 makes it y = A x).
typedef double hist_d2 __attribute__((ext_vector_type(2)));
typedef float hist_f2 __attribute__((ext_vector_type(2)));
template <class S> struct HistPair;
template <> struct HistPair<double> { typedef hist_d2 type; };
template <> struct HistPair<float> { typedef hist_f2 type; };

// S: the matrices' element type (double, or float for weights stored in single precision; the products and
// sums are formed in double either way).
template <class S, bool HG, bool HH, int R>
__global__ void __launch_bounds__(256) k_history_v2(const S *__restrict__ G, const S *__restrict__ H,
                                                     const double *__restrict__ B, const double *__restrict__ p,
                                                     const double *__restrict__ u, const double *__restrict__ acc,
                                                     long long n3, long long mat_stride, int m, int nh,
                                                     double *__restrict__ b) {
    const long long r0 = ((long long)blockIdx.x * 4 + (threadIdx.x >> 6)) * R;
    const int lane = threadIdx.x & 63;
    if (r0 >= n3) return;
    const long long nc = n3 >> 1;  // 16-byte columns
    long long rr[R];
#pragma unroll
    for (int k = 0; k < R; k++) rr[k] = r0 + k < n3 ? r0 + k : n3 - 1;  // rows past the end repeat the last (unused)
    double s[R];
#pragma unroll
    for (int k = 0; k < R; k++) s[k] = 0.0;
    for (int i = 1; i <= m; i++) {
        const long long mo = (long long)i * mat_stride;
        const hist_d2 *xp = HG ? reinterpret_cast<const hist_d2 *>(p + (long long)(nh - i) * n3) : nullptr;
        const hist_d2 *xu = HH ? reinterpret_cast<const hist_d2 *>(u + (long long)(nh - i) * n3) : nullptr;
        typedef typename HistPair<S>::type s2;
        const s2 *gr[R], *hr[R];
#pragma unroll
        for (int k = 0; k < R; k++) {
            gr[k] = HG ? reinterpret_cast<const s2 *>(G + mo + rr[k] * n3) : nullptr;
            hr[k] = HH ? reinterpret_cast<const s2 *>(H + mo + rr[k] * n3) : nullptr;
        }
        long long c = lane;
        for (; c + 64 * (kHistUnroll - 1) < nc; c += 64 * kHistUnroll) {
            hist_d2 x[kHistUnroll], y[kHistUnroll];
            s2 g[R][kHistUnroll], h[R][kHistUnroll];
#pragma unroll
            for (int j = 0; j < kHistUnroll; j++) {
                if (HG) x[j] = xp[c + 64 * j];
                if (HH) y[j] = xu[c + 64 * j];
#pragma unroll
                for (int k = 0; k < R; k++) {
                    if (HG) g[k][j] = __builtin_nontemporal_load(gr[k] + c + 64 * j);
                    if (HH) h[k][j] = __builtin_nontemporal_load(hr[k] + c + 64 * j);
                }
            }
#pragma unroll
            for (int j = 0; j < kHistUnroll; j++)
#pragma unroll
                for (int k = 0; k < R; k++) {
                    if (HG) s[k] = fma((double)g[k][j].y, x[j].y, fma((double)g[k][j].x, x[j].x, s[k]));
                    if (HH) s[k] = fma(-(double)h[k][j].y, y[j].y, fma(-(double)h[k][j].x, y[j].x, s[k]));
                }
        }
        for (; c < nc; c += 64) {
            const hist_d2 x = HG ? xp[c] : hist_d2{0.0, 0.0}, y = HH ? xu[c] : hist_d2{0.0, 0.0};
#pragma unroll
            for (int k = 0; k < R; k++) {
                if (HG) {
                    const s2 g = __builtin_nontemporal_load(gr[k] + c);
                    s[k] = fma((double)g.y, x.y, fma((double)g.x, x.x, s[k]));
                }
                if (HH) {
                    const s2 h = __builtin_nontemporal_load(hr[k] + c);
                    s[k] = fma(-(double)h.y, y.y, fma(-(double)h.x, y.x, s[k]));
                }
            }
        }
    }
    if (B)
        for (int j = lane; j < 6 * m; j += 64) {
            const int i = j / 6 + 1, q = j - 6 * (i - 1);
            const double a = acc[(long long)(nh - i) * 6 + q];
#pragma unroll
            for (int k = 0; k < R; k++) s[k] = fma(B[((long long)i * n3 + rr[k]) * 6 + q], a, s[k]);
        }
#pragma unroll
    for (int k = 0; k < R; k++)
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) s[k] += __shfl_xor(s[k], o, 64);
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < R; k++)
            if (r0 + k < n3) b[r0 + k] = s[k];
}

// y = A x for a row-major n x n matrix (u_p p of a constrained time step); one wavefront per row, as k_history.
__global__ void __launch_bounds__(256) k_matvec(const double *__restrict__ A, const double *__restrict__ x, long long n,
                                                double *__restrict__ y) {
    const long long r = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (r >= n) return;
    double s = history_row<true, false>(A + r * n, nullptr, x, nullptr, n, lane, 0.0);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if (lane == 0) y[r] = s;
}

// H_0 += C on the 3x3 diagonal blocks (collocation, H.cu:186-191).
template <class T>
__global__ void k_add_c_colloc(T *H0, const double *C, int nv) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= 9 * nv) return;
    const int v = t / 9, a = (t % 9) / 3, b = t % 3;
    const long long n3 = 3LL * nv;
    T &h = H0[(3LL * v + a) * n3 + 3 * v + b];
    h = (T)((double)h + C[t]);
}

// H_0 += C for Galerkin: sparse (row, col, value) triplets, one lane each, rows unique per launch.
template <class T>
__global__ void k_add_triplets(T *H0, const long long *idx, const double *val, long long n) {
    const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
    if (t >= n) return;
    H0[idx[t]] = (T)((double)H0[idx[t]] + val[t]);
}

// B_angular += B_trans (x) x_v and B = [B_trans | B_euler] (B_angular.cu:167-175, matrix_collection.cpp:280-305).
__global__ void k_b_post(double *Bt, double *Ba, double *B6, const double *V, int nv, int L) {
    const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
    const long long nr = 3LL * nv;
    if (t >= nr * L) return;
    const long long n = t / nr, r = t % nr;
    const double *bt = Bt + (n * nr + r) * 3;
    double *ba = Ba + (n * nr + r) * 9;
    const double *x = V + 3 * (r / 3);
    for (int j = 0; j < 3; j++)
        for (int k = 0; k < 3; k++) ba[j * 3 + k] += bt[j] * x[k];
    if (B6) {
        double *b6 = B6 + (n * nr + r) * 6;
        for (int j = 0; j < 3; j++) b6[j] = bt[j];
        for (int j = 0; j < 3; j++) b6[3 + j] = ba[((j + 2) % 3) * 3 + (j + 1) % 3] - ba[((j + 1) % 3) * 3 + (j + 2) % 3];
    }
}

// ------------------------------------------------------------- launchers --
hipError_t launch_colloc_gh(const Geom &g, const double4 *quad, int npts, const Freq *freqs, int l0, int nl, const Mat &m,
                            int row0, int nrows, cd *UG, cd *UH, long long ld, long long lstride, hipStream_t st) {
    dim3 grid((g.nv + 63) / 64, nl, nrows);
    if (UG && UH)
        hipLaunchKernelGGL((k_colloc_gh<true, true>), grid, dim3(64), 0, st, g, quad, npts, freqs, l0, m, row0, UG, UH, ld,
                           lstride);
    else if (UG)
        hipLaunchKernelGGL((k_colloc_gh<true, false>), grid, dim3(64), 0, st, g, quad, npts, freqs, l0, m, row0, UG, UH, ld,
                           lstride);
    else if (UH)
        hipLaunchKernelGGL((k_colloc_gh<false, true>), grid, dim3(64), 0, st, g, quad, npts, freqs, l0, m, row0, UG, UH, ld,
                           lstride);
    return hipGetLastError();
}


hipError_t launch_colloc_gh_pair(const Geom &g, const double4 *pts, int npts_even, const Freq *freqs, int l0, int nl, const Mat &m,
                                 int row0, int nrows, const double4 *tpts, const int *patch_pass, const int2 *pass_info,
                                 const int4 *pass_faces, const int *patch_verts, int npatches, cd *UG, cd *UH, long long ld,
                                 long long lstride, hipStream_t st, const Bands &bands) {
    if ((!UG && !UH) || (npts_even & 1) || nrows <= 0 || nrows > 65535) return hipErrorInvalidValue;
    if (bands.far_pts && (tpts || bands.far_rounds <= 0 || !bands.sphere)) return hipErrorInvalidValue;
    dim3 grid(npatches, nl, nrows);
#define CQBEM_PAIR(DG, DH, GAL, BANDED)                                                                                    \
    hipLaunchKernelGGL((k_colloc_gh_pair<DG, DH, GAL, BANDED>), grid, dim3(kPairThreads), 0, st, g, pts, npts_even / 2, freqs, \
                       l0, m, row0, tpts, patch_pass, pass_info, pass_faces, patch_verts, UG, UH, ld, lstride, bands)
    if (tpts) {  // test points: never banded (refused above)
        if (UG && UH) CQBEM_PAIR(true, true, true, false);
        else if (UG) CQBEM_PAIR(true, false, true, false);
        else CQBEM_PAIR(false, true, true, false);
    } else if (bands.far_pts) {
        if (UG && UH) CQBEM_PAIR(true, true, false, true);
        else if (UG) CQBEM_PAIR(true, false, false, true);
        else CQBEM_PAIR(false, true, false, true);
    } else {
        if (UG && UH) CQBEM_PAIR(true, true, false, false);
        else if (UG) CQBEM_PAIR(true, false, false, false);
        else CQBEM_PAIR(false, true, false, false);
    }
#undef CQBEM_PAIR
    return hipGetLastError();
}
hipError_t launch_scene_gh_pair(const ScenePiece *pieces, int npieces, long long ntiles, unsigned *counter, bool G, bool H,
                                int num_cus, hipStream_t st) {
    if ((!G && !H) || npieces <= 0 || ntiles <= 0 || ntiles >= (1LL << 31)) return hipErrorInvalidValue;
    const char *env = std::getenv("CQBEM_SCENE_KERNEL");  // read per launch (tests switch it)
    const bool persistent = env && !std::strcmp(env, "persistent");
    if (!persistent) {
        // a dispatch holds fewer than 2^32 work-items: grids of at most 2^22 blocks of 256, in tile order
        constexpr long long kChunk = 1LL << 22;
        for (long long t0 = 0; t0 < ntiles; t0 += kChunk) {
            const dim3 grid((unsigned)std::min(kChunk, ntiles - t0));
            if (G && H)
                hipLaunchKernelGGL((k_scene_gh_tiles<true, true>), grid, dim3(kPairThreads), 0, st, pieces, npieces, (unsigned)t0);
            else if (G)
                hipLaunchKernelGGL((k_scene_gh_tiles<true, false>), grid, dim3(kPairThreads), 0, st, pieces, npieces, (unsigned)t0);
            else
                hipLaunchKernelGGL((k_scene_gh_tiles<false, true>), grid, dim3(kPairThreads), 0, st, pieces, npieces, (unsigned)t0);
        }
        return hipGetLastError();
    }
    hipLaunchKernelGGL(k_zero_counter, dim3(1), dim3(64), 0, st, counter);
    const unsigned blocks = (unsigned)std::min<long long>(ntiles, 2LL * num_cus);
    if (G && H)
        hipLaunchKernelGGL((k_scene_gh_pair<true, true>), dim3(blocks), dim3(kPairThreads), 0, st, pieces, npieces,
                           (unsigned)ntiles, counter);
    else if (G)
        hipLaunchKernelGGL((k_scene_gh_pair<true, false>), dim3(blocks), dim3(kPairThreads), 0, st, pieces, npieces,
                           (unsigned)ntiles, counter);
    else
        hipLaunchKernelGGL((k_scene_gh_pair<false, true>), dim3(blocks), dim3(kPairThreads), 0, st, pieces, npieces,
                           (unsigned)ntiles, counter);
    return hipGetLastError();
}

int pair_patch_size() { return kPairPV; }
int pair_pass_faces() { return kPairFaces; }

hipError_t launch_colloc_b(const Geom &g, const double4 *quad, int npts, const Freq *freqs, int l0, int nl, const Mat &m,
                           int row0, int nrows, cd *Bt, cd *Ba, hipStream_t st, const FaceBands &fb) {
    if (nrows <= 0 || nrows > 65535) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_colloc_b<false>, dim3(nl, nrows), dim3(256), 0, st, g, quad, npts, freqs, l0, m, row0,
                       (const double4 *)nullptr, Bt, Ba, 3LL * g.nv, fb);
    return hipGetLastError();
}

hipError_t launch_points_b(const Geom &g, const double4 *quad, int npts, const Freq *freqs, int l0, int nl, const Mat &m,
                           const double4 *tpts, int row0, int nrows, cd *Bt, cd *Ba, long long out_rows, hipStream_t st) {
    if (nrows <= 0 || nrows > 65535 || 3LL * nrows > out_rows) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_colloc_b<true>, dim3(nl, nrows), dim3(256), 0, st, g, quad, npts, freqs, l0, m, row0, tpts, Bt, Ba,
                       out_rows, FaceBands{});
    return hipGetLastError();
}

hipError_t launch_colloc_singular(int kind, const Geom &g, const double2 *line, int nline, const Freq *freqs, int l0, int nl,
                                  const Mat &m, int row0, int nrows, cd *U, long long ld, long long lstride, cd *Bt,
                                  cd *Ba, hipStream_t st) {
    const long long n = (long long)nrows * nl;
    if (n <= 0 || n > 0x7fffffffLL) return hipErrorInvalidValue;
    dim3 grid((unsigned)n);  // one wavefront per (row, frequency)
    if (kind == 0)
        hipLaunchKernelGGL((k_colloc_singular<0>), grid, dim3(64), 0, st, g, line, nline, freqs, l0, nl, m, row0, nrows, U,
                           ld, lstride, Bt, Ba, 3LL * g.nv);
    else if (kind == 1)
        hipLaunchKernelGGL((k_colloc_singular<1>), grid, dim3(64), 0, st, g, line, nline, freqs, l0, nl, m, row0, nrows, U,
                           ld, lstride, Bt, Ba, 3LL * g.nv);
    else
        hipLaunchKernelGGL((k_colloc_singular<2>), grid, dim3(64), 0, st, g, line, nline, freqs, l0, nl, m, row0, nrows, U,
                           ld, lstride, Bt, Ba, 3LL * g.nv);
    return hipGetLastError();
}

template <class OutT>
static hipError_t launch_transform_t(const cd *spec, long long E, long long spec_ls, int Lf, const double2 *tab, int L,
                                     OutT *out, long long out_nstride, int num_cus, hipStream_t st) {
    // two entries per lane when the pairs stay aligned: even sizes and strides, a 2-value-aligned output
    if (Lf <= kTransformMaxLf && E % 2 == 0 && spec_ls % 2 == 0 && out_nstride % 2 == 0 &&
        reinterpret_cast<uintptr_t>(out) % (2 * sizeof(OutT)) == 0 && E / 2 <= 256LL * 0x7fffffffLL) {
        const long long pblocks = (E / 2 + 255) / 256;
        for (int n0 = 0; n0 < L; n0 += kTransformNC) {
            hipLaunchKernelGGL((k_transform_pair<2, OutT>), dim3((unsigned)pblocks), dim3(256), 0, st, spec, E, spec_ls, Lf,
                               tab, L, n0, out, out_nstride);
            hipError_t e = hipGetLastError();
            if (e != hipSuccess) return e;
        }
        return hipSuccess;
    }
    long long blocks = (E + 255) / 256;
    const long long cap = (long long)num_cus * 64;
    if (blocks > cap) blocks = cap;
    if (blocks < 1) blocks = 1;
    for (int n0 = 0; n0 < L; n0 += kTransformNC) {
        if (Lf <= kTransformMaxLf)
            hipLaunchKernelGGL((k_transform<kTransformNC, kTransformLC, OutT>), dim3((unsigned)blocks), dim3(256), 0, st, spec,
                               E, spec_ls, Lf, tab, L, n0, out, out_nstride);
        else
            hipLaunchKernelGGL((k_transform_wide<kTransformNC, OutT>), dim3((unsigned)blocks), dim3(256), 0, st, spec, E,
                               spec_ls, Lf, tab, L, n0, out, out_nstride);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}
hipError_t launch_transform(const cd *spec, long long E, long long spec_ls, int Lf, const double2 *tab, int L, double *out,
                            long long out_nstride, int num_cus, hipStream_t st) {
    return launch_transform_t(spec, E, spec_ls, Lf, tab, L, out, out_nstride, num_cus, st);
}
hipError_t launch_transform(const cd *spec, long long E, long long spec_ls, int Lf, const double2 *tab, int L, float *out,
                            long long out_nstride, int num_cus, hipStream_t st) {
    return launch_transform_t(spec, E, spec_ls, Lf, tab, L, out, out_nstride, num_cus, st);
}

template <int R, class S>
static void launch_history_v2(const S *G, const S *H, const double *B, const double *p, const double *u,
                              const double *acc, long long n3, long long mat_stride, int m, int nh, double *b,
                              hipStream_t st) {
    const long long waves = (n3 + R - 1) / R;
    const dim3 grid((unsigned)((waves + 3) / 4));
    if (G && H)
        hipLaunchKernelGGL((k_history_v2<S, true, true, R>), grid, dim3(256), 0, st, G, H, B, p, u, acc, n3, mat_stride, m,
                           nh, b);
    else if (G)
        hipLaunchKernelGGL((k_history_v2<S, true, false, R>), grid, dim3(256), 0, st, G, H, B, p, u, acc, n3, mat_stride, m,
                           nh, b);
    else if (H)
        hipLaunchKernelGGL((k_history_v2<S, false, true, R>), grid, dim3(256), 0, st, G, H, B, p, u, acc, n3, mat_stride, m,
                           nh, b);
    else
        hipLaunchKernelGGL((k_history_v2<S, false, false, R>), grid, dim3(256), 0, st, G, H, B, p, u, acc, n3, mat_stride,
                           m, nh, b);
}

// Rows per wavefront of the 16-byte history and matvec kernel: CQBEM_HISTORY_ROWS = 1 selects the 8-byte one-row
// kernel, 2 or 4 (default) the vectorized one.  Read at every call (a few hundred ns against a launch).
static int history_rows() {
    const char *e = std::getenv("CQBEM_HISTORY_ROWS");
    const int v = e ? std::atoi(e) : 4;
    return v == 1 || v == 2 ? v : 4;
}

static bool aligned16(const double *q) { return (reinterpret_cast<uintptr_t>(q) & 15) == 0; }
static bool aligned8(const void *q) { return (reinterpret_cast<uintptr_t>(q) & 7) == 0; }

hipError_t launch_history_f32(const float *G, const float *H, const double *B, const double *p, const double *u,
                              const double *acc, long long n3, int m, int nh, double *b, hipStream_t st) {
    // single-precision matrices: 8-byte pairs of weights against 16-byte pairs of history entries
    if (n3 % 2 != 0 || (G && (!aligned8(G) || !aligned16(p))) || (H && (!aligned8(H) || !aligned16(u))))
        return hipErrorInvalidValue;
    launch_history_v2<4>(G, H, B, p, u, acc, n3, n3 * n3, m, nh, b, st);
    return hipGetLastError();
}

hipError_t launch_history(const double *G, const double *H, const double *B, const double *p, const double *u,
                          const double *acc, long long n3, int m, int nh, double *b, hipStream_t st) {
    const int rows = history_rows();
    if (rows > 1 && n3 % 2 == 0 && (!G || (aligned16(G) && aligned16(p))) && (!H || (aligned16(H) && aligned16(u)))) {
        if (rows == 4)
            launch_history_v2<4>(G, H, B, p, u, acc, n3, n3 * n3, m, nh, b, st);
        else
            launch_history_v2<2>(G, H, B, p, u, acc, n3, n3 * n3, m, nh, b, st);
        return hipGetLastError();
    }
    const dim3 grid((unsigned)((n3 + 3) / 4));
    if (G && H)
        hipLaunchKernelGGL((k_history<true, true>), grid, dim3(256), 0, st, G, H, B, p, u, acc, n3, m, nh, b);
    else if (G)
        hipLaunchKernelGGL((k_history<true, false>), grid, dim3(256), 0, st, G, H, B, p, u, acc, n3, m, nh, b);
    else if (H)
        hipLaunchKernelGGL((k_history<false, true>), grid, dim3(256), 0, st, G, H, B, p, u, acc, n3, m, nh, b);
    else
        hipLaunchKernelGGL((k_history<false, false>), grid, dim3(256), 0, st, G, H, B, p, u, acc, n3, m, nh, b);
    return hipGetLastError();
}

hipError_t launch_matvec(const double *A, const double *x, long long n, double *y, hipStream_t st) {
    const int rows = history_rows();
    if (rows > 1 && n % 2 == 0 && aligned16(A) && aligned16(x)) {
        if (rows == 4)
            launch_history_v2<4, double>(A, nullptr, nullptr, x, nullptr, nullptr, n, 0, 1, 1, y, st);
        else
            launch_history_v2<2, double>(A, nullptr, nullptr, x, nullptr, nullptr, n, 0, 1, 1, y, st);
        return hipGetLastError();
    }
    hipLaunchKernelGGL(k_matvec, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, st, A, x, n, y);
    return hipGetLastError();
}

hipError_t launch_add_c_colloc(double *H0, const double *C, int nv, hipStream_t st) {
    hipLaunchKernelGGL(k_add_c_colloc<double>, dim3((9 * nv + 255) / 256), dim3(256), 0, st, H0, C, nv);
    return hipGetLastError();
}
hipError_t launch_add_c_colloc(float *H0, const double *C, int nv, hipStream_t st) {
    hipLaunchKernelGGL(k_add_c_colloc<float>, dim3((9 * nv + 255) / 256), dim3(256), 0, st, H0, C, nv);
    return hipGetLastError();
}

hipError_t launch_add_triplets(double *H0, const long long *idx, const double *val, long long n, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_add_triplets<double>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, H0, idx, val, n);
    return hipGetLastError();
}
hipError_t launch_add_triplets(float *H0, const long long *idx, const double *val, long long n, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_add_triplets<float>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, H0, idx, val, n);
    return hipGetLastError();
}

hipError_t launch_b_post(double *Bt, double *Ba, double *B6, const double *V, int nv, int L, hipStream_t st) {
    const long long n = 3LL * nv * L;
    hipLaunchKernelGGL(k_b_post, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, Bt, Ba, B6, V, nv, L);
    return hipGetLastError();
}

}  // namespace cqbem
