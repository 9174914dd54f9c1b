// Kernel-side data structures and launchers shared by the host driver.
#pragma once
#include <hip/hip_runtime.h>

#include "cqbem_device.hpp"

namespace cqbem {

// Device-resident mesh of one job.
struct Geom {
    const double *V;     // nv x 3
    const int *F;        // nf x 3
    const double *N;     // nf x 3 unit normals
    const double *area;  // nf
    const int *vs_ptr;   // nv+1: vertex star CSR (faces incident to each vertex)
    const int *vs_face;  // (face << 2) | corner, ascending face index
    const int *sg_ptr;   // nv+1: faces singular for each collocation vertex
    const int *sg_face;  // (face << 2) | first matching corner
    const int *vmap;     // nullable vertex_map_inverse
    int nv, nf;
};

hipError_t launch_colloc_gh(const Geom &g, const double4 *quad, int npts, const Freq *freqs, int l0, int nl, const Mat &m,
                            int row0, int nrows, cd *UG, cd *UH, long long ld, long long lstride, hipStream_t st);

// Distance bands of the pair kernel's regular-face rule (CQBEM_QUAD_BANDED).  A tile (patch, row point x) takes the
// far rule at every frequency when |x - centre| >= radius + kappa h (h: the patch's largest face diameter);
// otherwise the reference rule.  far_pts == nullptr: the reference rule for every tile.
struct Bands {
    const double4 *far_pts = nullptr;  // pair point table of the far rule (layout of the reference rule's)
    int far_rounds = 0;                // its padded points / 2
    const double4 *sphere = nullptr;   // per patch: centre, radius + kappa h
};
// The bands' rule choice: whether row point x lies at or beyond distance S.w from (S.x, S.y, S.z).  Compiled without
// contraction, so that the host plans and the kernels decide every tile and face alike.
__host__ __device__ inline bool beyond_sphere(double x0, double x1, double x2, const double4 &S) {
#pragma clang fp contract(off)
    const double dx = x0 - S.x, dy = x1 - S.y, dz = x2 - S.z;
    return dx * dx + dy * dy + dz * dz >= S.w * S.w;
}
// Distance bands of the B kernels' regular-face rule (CQBEM_QUAD_BANDED).  Face f takes the far rule for row point x
// at every frequency when |x - centre_f| >= radius_f + kappa h_f (h_f: its longest edge); otherwise the reference
// rule.  sphere == nullptr: the reference rule for every face.
struct FaceBands {
    const double4 *far_quad = nullptr;  // the far rule, (b0, b1, b2, w) per point
    int far_npts = 0;
    const double4 *sphere = nullptr;    // per face: centre, radius + kappa h
};
// Pair form of the G/H regular-face kernel (both matrices): per patch a range of passes
// [patch_pass[p], patch_pass[p+1]); per pass (first face entry, nfaces | nsub << 16); per face entry
// (face, corner 0, corner 1, corner 2) with corner = slot | sub-step << 10 or -1 outside the patch.
// `pts` holds, for each point q of the (even-padded) rule, (b0, b1, b2, w) and (phi_F(q), phi_F(partner), 0, 0)
// with phi_F the barycentric of corner 0 for even q (lane A) and of corner 2 for odd q (lane B), partner = q ^ 1.
// Rows are collocation vertices row0..row0+nrows-1 when tpts is null.  Otherwise they are the test points
// tpts[row0..row0+nrows-1] = (x, y, z, test face), each excluding every face that shares a vertex with its
// test face (the Galerkin non-singular set), and row r of the launch goes to rows 3r..3r+2 of UG/UH.
hipError_t launch_colloc_gh_pair(const Geom &g, const double4 *pts, int npts_even, const Freq *freqs, int l0, int nl, const Mat &m,
                                 int row0, int nrows, const double4 *tpts, const int *patch_pass, const int2 *pass_info,
                                 const int4 *pass_faces, const int *patch_verts, int npatches, cd *UG, cd *UH, long long ld,
                                 long long lstride, hipStream_t st, const Bands &bands = Bands{});
int pair_patch_size();
int pair_pass_faces();

// One object's row slab in a multi-object launch of the pair kernel (collocation rows row0..row0+nrows-1 at all
// nl frequencies of `freqs`): its npatches x nl x nrows tiles are numbered base .. base + npatches nl nrows - 1.
struct ScenePiece {
    Geom g;
    const double4 *pts;
    int nrounds;  // padded rule points / 2
    const Freq *freqs;
    Mat m;
    int row0, nrows, nl, npatches;
    const int *patch_pass;
    const int2 *pass_info;
    const int4 *pass_faces;
    const int *patch_verts;
    cd *UG, *UH;
    long long ld, lstride;
    unsigned base;
    Bands bands;
};
// The pieces (device array, ascending base, all with the same G / H selection) in one persistent launch of
// min(ntiles, 2 blocks per CU) blocks that take tiles from *counter (zeroed by the launcher with a kernel).
hipError_t launch_scene_gh_pair(const ScenePiece *pieces, int npieces, long long ntiles, unsigned *counter, bool G, bool H,
                                int num_cus, hipStream_t st);
hipError_t launch_colloc_b(const Geom &g, const double4 *quad, int npts, const Freq *freqs, int l0, int nl, const Mat &m,
                           int row0, int nrows, cd *Bt, cd *Ba, hipStream_t st, const FaceBands &fb = FaceBands{});
// B_trans / B_angular rows of the test points tpts[row0..row0+nrows-1] (Galerkin non-singular set, as above), into
// rows 3r..3r+2 of Bt (out_rows x 3 per frequency) and Ba (out_rows x 9).
hipError_t launch_points_b(const Geom &g, const double4 *quad, int npts, const Freq *freqs, int l0, int nl, const Mat &m,
                           const double4 *tpts, int row0, int nrows, cd *Bt, cd *Ba, long long out_rows, hipStream_t st);
hipError_t launch_colloc_singular(int kind, const Geom &g, const double2 *line, int nline, const Freq *freqs, int l0, int nl,
                                  const Mat &m, int row0, int nrows, cd *U, long long ld, long long lstride, cd *Bt,
                                  cd *Ba, hipStream_t st);
// out[n * out_nstride + e] = sum_l Re(spec[l * spec_ls + e] T(n, l)) for e < E (spec_ls >= E: a slab buffer may be
// sized for more rows than the slab holds).
hipError_t launch_transform(const cd *spec, long long E, long long spec_ls, int Lf, const double2 *tab, int L, double *out,
                            long long out_nstride,
                            int num_cus, hipStream_t st);
// The same with the weights rounded to float (resident storage in single precision).
hipError_t launch_transform(const cd *spec, long long E, long long spec_ls, int Lf, const double2 *tab, int L, float *out,
                            long long out_nstride, int num_cus, hipStream_t st);
hipError_t launch_add_c_colloc(double *H0, const double *C, int nv, hipStream_t st);
hipError_t launch_add_c_colloc(float *H0, const double *C, int nv, hipStream_t st);
// b = sum_{i=1..m} (G_i p[nh-i] - H_i u[nh-i] + B_i acc[nh-i]); any of G, H, B may be null.
hipError_t launch_history(const double *G, const double *H, const double *B, const double *p, const double *u,
                          const double *acc, long long n3, int m, int nh, double *b, hipStream_t st);
hipError_t launch_matvec(const double *A, const double *x, long long n, double *y, hipStream_t st);
hipError_t launch_add_triplets(double *H0, const long long *idx, const double *val, long long n, hipStream_t st);
hipError_t launch_add_triplets(float *H0, const long long *idx, const double *val, long long n, hipStream_t st);
// The history sum with G_i, H_i stored in float (n3 even, G/H 8-byte and p/u 16-byte aligned).
hipError_t launch_history_f32(const float *G, const float *H, const double *B, const double *p, const double *u,
                              const double *acc, long long n3, int m, int nh, double *b, hipStream_t st);
hipError_t launch_b_post(double *Bt, double *Ba, double *B6, const double *V, int nv, int L, hipStream_t st);

}  // namespace cqbem
