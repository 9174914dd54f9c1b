/*
 * cqbem.h -- C-ABI of the MI355X (gfx950) CQBEM precompute.
 *
 * Assembles the convolution-quadrature boundary-element weight matrices G_n, H_n,
 * B_trans_n and B_angular_n (n = 0..L-1) of a closed triangle surface mesh, in
 * collocation or Galerkin form, on one GPU.  Plain pointers and sizes only.
 *
 * Reference interfaces replaced (rsugimoto/SurfaceOnlyDeformables):
 *   ElastodynamicsCQBEMMatrixCollection::compute_H_matrices   src/elastodynamics_cqbem_H.cu:18-200
 *   ElastodynamicsCQBEMMatrixCollection::compute_G_matrices   src/elastodynamics_cqbem_G.cu:18-183
 *   ElastodynamicsCQBEMMatrixCollection::compute_B_trans_matrices   src/elastodynamics_cqbem_B_trans.cu:14-164
 *   ElastodynamicsCQBEMMatrixCollection::compute_B_angular_tensors  src/elastodynamics_cqbem_B_angular.cu:14-179
 *   ElastodynamicsCQBEMMatrixCollection::compute_B_euler_matrices   src/elastodynamics_cqbem_matrix_collection.cpp:280-305
 *   the kernel-level seam of include/elastodynamics_cqbem_collocation_kernels.cuh:15-77 and
 *   include/elastodynamics_cqbem_galerkin_kernels.cuh:18-108 (one call per frequency there; one call per
 *   job here, all frequencies and the inverse transform included).
 *
 * Layouts follow include/type_declaration.hpp:5,16-48 of the reference: double, row-major, 64-bit sizes.
 *   G, H        L x (3 nv) x (3 nv)       H_0 includes the C (jump) matrix (H.cu:184-193); not negated
 *   B_trans     L x (3 nv) x 3
 *   B_angular   L x (3 nv) x 3 x 3        includes + B_trans (x) (x_v - cm), cm = 0 (B_angular.cu:167-175)
 *   B           L x (3 nv) x 6            [B_trans | B_euler]  (matrix_collection.cpp:168-172)
 * Every function returns CQBEM_OK (0) or a negative status; cqbem_last_error() describes the last failure.
 * Calls on one context are serialized by the context; distinct contexts may be used from distinct threads.
 */
#ifndef CQBEM_ABI_H_
#define CQBEM_ABI_H_
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    CQBEM_OK = 0,
    CQBEM_ERR_ARG = -1,       /* invalid argument */
    CQBEM_ERR_HIP = -2,       /* HIP runtime failure (no device, launch failure, ...) */
    CQBEM_ERR_NOMEM = -3,     /* job does not fit the context's HBM budget or device memory */
    CQBEM_ERR_UNSUPPORTED = -4,
    CQBEM_ERR_BUSY = -5,      /* serving: no free slot (retry later) */
    CQBEM_ERR_NOTFOUND = -6,
    CQBEM_ERR_CANCELLED = -7  /* serving: the job was cancelled (its outputs are partial) */
};

enum { CQBEM_BDF1 = 1, CQBEM_BDF2 = 2 };
/* CQBEM_MIXED (collocation): far-field B_trans and B_angular integrated in float, as the reference's typedefs do
 * (B_trans_collocation.cu:14, B_angular_collocation.cu:14), with Kahan-compensated face sums; H, the singular faces
 * and the inverse transform in double.  Far-field G stays in double by default: it shares H's exponentials in the
 * pair kernel, which costs less than a float kernel of its own; the environment variable CQBEM_MIXED_G=float
 * selects the reference's float G (G_collocation.cu:14-15) instead.  Galerkin: the far-field B at the test points
 * in float, G and H in double (CQBEM_MIXED_G=float with Galerkin returns CQBEM_ERR_UNSUPPORTED). */
enum { CQBEM_FP64 = 0, CQBEM_MIXED = 1 };
/* Element type of the G and H outputs: double, or float (weights rounded once from the double transform; device
 * outputs only).  B_trans, B_angular and B are always double.  No counterpart in the reference: a storage policy for
 * jobs whose double weights exceed the memory they may hold resident (an icosphere-L5 job at L = 16 needs 242 GB
 * in double, 121 GB in float). */
enum { CQBEM_STORE_F64 = 0, CQBEM_STORE_F32 = 1 };
/* Regular-face quadrature of the collocation G, H, B_trans and B_angular (no counterpart in the reference, which
 * applies one uniform rule to every non-touching pair: integrator_collocation.hpp:6-17).  CQBEM_QUAD_REFERENCE
 * (default): that rule everywhere.  CQBEM_QUAD_BANDED: the rule (far_order, far_subdivision) replaces it, for G and
 * H per tile (column patch, row point) when the row point lies at least band_kappa x h beyond the patch's bounding
 * sphere (h: the patch's largest face diameter), and, only when band_kappa_b > 0, for B per (row point, face) when
 * it lies band_kappa_b x h_f beyond the face's own sphere (h_f: its longest edge); at every frequency alike (a rule that changed along the spectrum would make each entry's
 * quadrature error jump between frequencies, which the inverse transform's R^-n factor amplifies).  An
 * approximation with a measured error, not bit-compatible with the reference.  Zeros select the tuned defaults
 * (far rule: order, subdivision 1). */
enum { CQBEM_QUAD_REFERENCE = 0, CQBEM_QUAD_BANDED = 1 };
enum { CQBEM_G = 1, CQBEM_H = 2, CQBEM_B_TRANS = 4, CQBEM_B_ANGULAR = 8, CQBEM_ALL = 15 };

typedef struct cqbem_context cqbem_context;

/* Surface mesh (reference: V, F after load_mesh and centering, physics_base_matrix_collection.cpp:32-57). */
typedef struct {
    const double *V;                  /* nv x 3, centred on the volume centroid */
    int64_t nv;
    const int32_t *F;                 /* nf x 3, outward counter-clockwise */
    int64_t nf;
    const int32_t *vertex_map_inverse; /* nullable; traction-discontinuity vertex map (G_collocation.cu:81-88) */
} cqbem_mesh;

/* Material, CQ and quadrature parameters (reference: matrix_collection.cpp:59-69, G.cu:19-24,
 * elastics_bem_matrix_collection.cpp:19-38).  L and R are explicit: the reference derives
 * L = max_time_history + 1 and R = 1e-10^(1/(2 max_time_history)). */
typedef struct {
    double c1, c2, rho, nu;
    double dt, R;
    int32_t L;
    int32_t bdf;          /* CQBEM_BDF1 | CQBEM_BDF2 */
    int32_t order;        /* gaussian_quadrature_order, 1..5 */
    int32_t subdivision;  /* quadrature_subdivision, >= 1 */
    int32_t galerkin;     /* 0 collocation, 1 Galerkin */
    int32_t precision;    /* CQBEM_FP64 | CQBEM_MIXED */
    int32_t which;        /* bitmask of CQBEM_G | CQBEM_H | CQBEM_B_TRANS | CQBEM_B_ANGULAR */
    int32_t storage;      /* CQBEM_STORE_F64 | CQBEM_STORE_F32: element type of the G and H outputs */
    int32_t quad_policy;  /* CQBEM_QUAD_REFERENCE | CQBEM_QUAD_BANDED (collocation; FP64 or CQBEM_MIXED) */
    int32_t far_order;    /* banded: far rule order (0: order) */
    int32_t far_subdivision; /* banded: far rule subdivision (0: 1) */
    int32_t reserved;     /* 0 */
    double band_kappa;    /* banded G/H: distance threshold in patch face diameters (0: tuned default) */
    double band_kappa_b;  /* banded B: distance threshold in the face's longest edges (0: B keeps the reference rule) */
} cqbem_params;

/* Output buffers; any may be NULL.  on_device != 0: device pointers on the context's GPU;
 * otherwise host pointers (the library stages through device memory). */
typedef struct {
    double *G, *H, *B_trans, *B_angular, *B;
    int32_t on_device;
} cqbem_outputs;

/* Creates a context on HIP device `device`; hbm_budget_bytes = 0 means 90% of free device memory.
 * A context works on streams of its own.  Device buffers passed to it must hold their data when the call is
 * made: work the caller queued on another stream to produce them (or to fill an output buffer) must have
 * completed, e.g. by synchronizing that stream first.  Calls that return results complete before returning. */
int cqbem_create(cqbem_context **ctx, int device, uint64_t hbm_budget_bytes);
int cqbem_destroy(cqbem_context *ctx);
const char *cqbem_last_error(const cqbem_context *ctx);

/* Blocking assembly of one job (replaces compute_H/G/B_trans/B_angular_matrices + compute_B_euler_matrices). */
int cqbem_assemble(cqbem_context *ctx, const cqbem_mesh *mesh, const cqbem_params *params, cqbem_outputs *out);

/* cqbem_assemble without any device allocation.  No counterpart in the reference: for callers that cannot
 * allocate device memory while they run (serving loops that forbid it, code under HIP graph capture in the
 * caller's own design, harnesses that ban hipMalloc).  The outputs must be caller-provided device buffers
 * (out->on_device != 0); every other device buffer of the job (mesh copies, rule tables, Laplace-domain spectra
 * in row slabs, B spectra) is carved from a fixed scratch array of CQBEM_NOALLOC_SCRATCH_MIB MiB that the
 * library holds on each GPU it is loaded on until the process ends.  No hipMalloc, hipFree or HIP event is
 * created.  A job whose buffers exceed the array (even with one-row slabs) returns CQBEM_ERR_NOMEM with the size
 * it needed.  Results are bit-identical to cqbem_assemble's.  Calls on one GPU share its array and run one at a
 * time within a process (the library serializes them); the array keeps the previous call's contents, and every
 * kernel writes what it later reads.  Collocation and Galerkin are both accepted. */
#define CQBEM_NOALLOC_SCRATCH_MIB 256
int cqbem_assemble_noalloc(cqbem_context *ctx, const cqbem_mesh *mesh, const cqbem_params *params,
                           const cqbem_outputs *out);

/* All objects of a scene in one call (device outputs only).  No counterpart in the reference, whose
 * Simulator::init assembles one object group after another (simulator.cpp:101-125).  The G/H regular-face tiles
 * of every collocation object -- (column patch, frequency, row) tiles of all objects, largest objects first --
 * run in shared launches of the pair kernel, in waves whose Laplace-domain spectra fit the HBM budget; each
 * wave (the shared launch, the float G kernel of CQBEM_MIXED objects, the singular faces and the transforms of
 * every object slab) is captured in one HIP graph and launched once, and B, C and the B post-processing of all
 * objects follow in one more graph.  Galerkin objects run after them, one by one, as cqbem_assemble runs them.
 * Every object's outputs are bit-identical to its own cqbem_assemble.  cqbem_last_timings: [1] the waves' wall
 * time, [2] the B graph, [6] the waves, [7] the far-field evaluations. */
int cqbem_assemble_scene(cqbem_context *ctx, int32_t n, const cqbem_mesh *meshes, const cqbem_params *params,
                         const cqbem_outputs *outs);

/* Device bytes cqbem_assemble needs for a job (outputs on device included when outputs_on_device != 0). */
int cqbem_job_footprint(const cqbem_mesh *mesh, const cqbem_params *params, int outputs_on_device, uint64_t *bytes);

/* The tile plan of CQBEM_QUAD_BANDED (host only, no context): *npatches receives the column patch count;
 * patch_of_vertex (nv, may be NULL) the patch holding each vertex's column; far (nv x npatches, row-major, may be
 * NULL) whether tile (row vertex, patch) takes the far rule, at every frequency.  Entry (row i, column j) of G and
 * H is integrated entirely in tile (i, patch_of_vertex[j]).  No counterpart in the reference: for tests and tuning. */
int cqbem_band_plan(const cqbem_mesh *mesh, const cqbem_params *params, int32_t *patch_of_vertex, uint8_t *far,
                    int32_t *npatches);

/* The face plan of CQBEM_QUAD_BANDED for the far-field B_trans / B_angular (host only, no context): far (nv x nf,
 * row-major) whether face f takes the far rule in row vertex i's sums, at every frequency (the faces at vertex i are
 * integrated by the singular rule whatever this says).  A face is far when the row point lies at least
 * radius + band_kappa_b h from the mean of its corners (radius: its farthest corner; h: its longest edge).  No
 * counterpart in the reference: for tests and tuning.  With band_kappa_b = 0 no face is far. */
int cqbem_band_plan_faces(const cqbem_mesh *mesh, const cqbem_params *params, uint8_t *far);

/* Laplace-domain matrices of one CQ frequency index l (0..L/2), complex interleaved, row-major:
 * U, P (3nv x 3nv), Bt (3nv x 3), Ba (3nv x 9).  Host pointers; any may be NULL.  Used for parity checks
 * of the per-frequency assembly (the U_l / P_l matrices of G.cu:150 / H.cu:155). */
int cqbem_laplace(cqbem_context *ctx, const cqbem_mesh *mesh, const cqbem_params *params, int32_t l, double *U,
                  double *P, double *Bt, double *Ba);

/* Traction-discontinuity vertex split, the mesh step of PhysicsBaseMatrixCollection::load_mesh with
 * enable_traction_discontinuity (src/physics_base_matrix_collection.cpp:203-365, vertex reordering off):
 * vertices on sharp edges (face normals more than threshold_deg apart; the reference's default is 20)
 * get one copy per smooth sheet, placed right after the original.  In: V (nv x 3), F (nf x 3) of a
 * closed, consistently oriented manifold.  Out: V_out (nv_out x 3), F_out (nf x 3), vertex_map_inverse
 * (nv_out; output vertex -> input vertex), ready for cqbem_mesh.  With V_out, F_out and
 * vertex_map_inverse all NULL only *nv_out is set; otherwise *nv_out must hold the capacity on entry. */
int cqbem_split_traction_discontinuity(const double *V, int64_t nv, const int32_t *F, int64_t nf, double threshold_deg,
                                       double *V_out, int32_t *F_out, int32_t *vertex_map_inverse, int64_t *nv_out);

/* Subdivision schemes of cqbem_subdivide_mesh (the reference's mesh.subdivision_scheme values "Upsample", "Loop",
 * "FalseBarycentric"; an unknown name selects Upsample there). */
enum cqbem_subdivision_scheme {
    CQBEM_SUBDIV_UPSAMPLE = 0,          /* igl::upsample: four faces per face, edge midpoints */
    CQBEM_SUBDIV_LOOP = 1,              /* igl::loop: four faces per face, Loop's stencils with Warren's weights */
    CQBEM_SUBDIV_FALSE_BARYCENTRIC = 2  /* igl::false_barycentric_subdivision: three faces per face */
};

/* Mesh subdivision, the mesh step of PhysicsBaseMatrixCollection::load_mesh with subdivision_level > 0
 * (src/physics_base_matrix_collection.cpp:109-140; libigl v2.4.0's upsample, loop and
 * false_barycentric_subdivision): `levels` rounds of `scheme`.  Old vertices keep their indices, new ones follow,
 * and the positions are bit-identical to the reference's sparse products.  In: V (nv x 3), F (nf x 3) of an
 * edge-manifold mesh without repeated face vertices.  Out: V_out (nv_out x 3), F_out (nf_out x 3).  With V_out and
 * F_out both NULL only the sizes are set; otherwise *nv_out and *nf_out must hold the capacities on entry. */
int cqbem_subdivide_mesh(const double *V, int64_t nv, const int32_t *F, int64_t nf, int32_t scheme, int32_t levels,
                         double *V_out, int32_t *F_out, int64_t *nv_out, int64_t *nf_out);

/* Mesh decimation, the mesh step of PhysicsBaseMatrixCollection::load_mesh with decimation_scale < 1
 * (src/physics_base_matrix_collection.cpp:143-151; libigl v2.4.0's igl::decimate with its default shortest-edge
 * cost and midpoint placement, called there with max_faces = (int64)(nf * decimation_scale)): edges are collapsed,
 * shortest first, keeping the mesh a closed 2-manifold of the same topology, until at most max_faces faces remain
 * (at least one collapse is made) or no edge can be collapsed.  In: a closed, consistently oriented, edge-manifold
 * mesh.  Sizes and capacities as cqbem_subdivide_mesh. */
int cqbem_decimate_mesh(const double *V, int64_t nv, const int32_t *F, int64_t nf, int64_t max_faces, double *V_out,
                        int32_t *F_out, int64_t *nv_out, int64_t *nf_out);

/* CQ frequencies s_l = gamma(R zeta^l)/dt, l = 0..L/2, complex interleaved (G.cu:19-24,65). Returns L/2+1. */
int cqbem_frequencies(int32_t L, double R, double dt, int32_t bdf, double *s_out);

/* Statistics of the last cqbem_assemble on this context (device times from HIP events on the context's
 * stream, milliseconds): [0] total wall, [1] regular G/H quadrature (the dominant kernel; Galerkin: all
 * non-singular quadrature, G, H and B, with the test-basis projection), [2] regular B_trans/B_angular
 * quadrature (collocation; Galerkin: the G/H pair kernel at the test points, a part of [1]), [3] singular
 * kernels (collocation: faces at the point; Galerkin: touching pairs and their gather), [4] inverse transform
 * and post-processing, [5] host preparation and staging, [6] launches of the dominant G/H kernel (the pair
 * kernel; Galerkin: at the test points), [7] far-field kernel evaluations (collocation point or test point,
 * source point, frequency) those launches performed.
 * Copies min(n, 8) values; returns that count. */
int cqbem_last_timings(const cqbem_context *ctx, double *out, int32_t n);

/* ---- serving: continuous batching of assembly jobs on one context ----
 * cqbem_submit admits a job when its device footprint fits the budget next to the jobs in flight,
 * copies the mesh, and enqueues it on a free stream slot; results land in `out` (device pointers
 * required: matrices stay resident).  The caller keeps `out` buffers alive until the job is done. */
int cqbem_submit(cqbem_context *ctx, const cqbem_mesh *mesh, const cqbem_params *params, const cqbem_outputs *out,
                 uint64_t *job_id);
/* *done = 1 once the job finished (its scratch is then released); status of the job in the return value. */
int cqbem_poll(cqbem_context *ctx, uint64_t job_id, int32_t *done);
int cqbem_wait(cqbem_context *ctx, uint64_t job_id);
/* Bytes currently reserved: the scratch of in-flight jobs and the outputs of resident jobs. */
int cqbem_reserved_bytes(const cqbem_context *ctx, uint64_t *bytes);
/* cqbem_submit for a job whose outputs stay resident after it finishes (the matrices a simulated object steps on):
 * admission counts the job's output bytes (G and H at their storage size, B_trans, B_angular and B) next to its
 * scratch, and those bytes stay reserved -- so later jobs are admitted against them -- until cqbem_release.
 * cqbem_poll and cqbem_wait work as for cqbem_submit; the job's record is kept until cqbem_release.  No
 * counterpart in the reference (its precompute runs once per object group at start-up). */
int cqbem_submit_resident(cqbem_context *ctx, const cqbem_mesh *mesh, const cqbem_params *params,
                          const cqbem_outputs *out, uint64_t *job_id);
/* Ends the residency of a finished job submitted with cqbem_submit_resident (the caller may then free its outputs):
 * its output bytes leave the reservation.  CQBEM_ERR_BUSY while the job still runs, CQBEM_ERR_NOTFOUND for an
 * unknown id or a job that was not resident. */
int cqbem_release(cqbem_context *ctx, uint64_t job_id);
/* Asks a submitted job to stop: it ends at its next row-slab or phase boundary (the device work already launched
 * completes), its scratch reservation is released, and cqbem_poll / cqbem_wait then return CQBEM_ERR_CANCELLED; its
 * outputs are partial.  A job that has already finished is unaffected; a resident job keeps its output reservation
 * until cqbem_release.  CQBEM_ERR_NOTFOUND for an unknown id.  No counterpart in the reference: a serving control. */
int cqbem_cancel(cqbem_context *ctx, uint64_t job_id);

/* ---- time stepping on the resident matrices ----
 * History term of one CQ time step (reference: ElastodynamicsCQBEMObject::estimate_next_state,
 * src/elastodynamics_cqbem_object.cpp:42-49, whose H_matrices are stored negated):
 *   b = sum_{i=1..m} ( G_i p[nhist-i] - H_i u[nhist-i] + B_i acc[nhist-i] ),   m = min(nhist, L-1),
 * with G, H (L x n3 x n3) and B (L x n3 x 6) as cqbem_assemble writes them, and the histories
 * p_hist, u_hist (nhist x n3) and acc_hist (nhist x 6: translational then rotational acceleration),
 * oldest first.  All pointers are device pointers on the context's GPU; any of G, H, B may be NULL
 * (its term is skipped).  b (n3) is complete when the call returns. */
int cqbem_history(cqbem_context *ctx, int64_t n3, int32_t L, const double *G, const double *H, const double *B,
                  const double *p_hist, const double *u_hist, const double *acc_hist, int32_t nhist, double *b);
/* The same history term on G and H stored in float (CQBEM_STORE_F32), products and sums in double.  n3 must be
 * even, G and H 8-byte and the histories 16-byte aligned.  Same reference as cqbem_history. */
int cqbem_history_f32(cqbem_context *ctx, int64_t n3, int32_t L, const float *G, const float *H, const double *B,
                      const double *p_hist, const double *u_hist, const double *acc_hist, int32_t nhist, double *b);
/* y = A x for a row-major n x n device matrix A and device vectors x, y (the u_p * p product of a
 * constrained step, elastodynamics_cqbem_object.cpp:136).  y is complete when the call returns. */
int cqbem_matvec(cqbem_context *ctx, int64_t n, const double *A, const double *x, double *y);

/* ---- wavelet-compressed matrices for the time stepping ----
 * Haar compression of a dense matrix, the reference's CompressedMatrix<1>(mat, num_elems)
 * (include/wavelets.hpp:154-219,433-445): the n x n row-major device matrix A (n >= 2, zero-padded to the
 * power of two np >= n) in Beylkin's nonstandard form, (2np - 2) x (2np - 2), of which the num_elems
 * largest-magnitude nonzero coefficients are kept, as CSR with columns ascending in each row, into
 * caller-provided device arrays rowptr (2np - 1 entries), col and val (num_elems entries).  *nnz (host)
 * receives the count kept (num_elems, or fewer when fewer coefficients are nonzero).  Coefficients tied
 * with the smallest kept magnitude are kept in row-major order.  The reference keeps
 * (int)(n * n * compression_ratio) (wavelets.hpp:443-445); when that count is <= 0 it keeps every nonzero
 * coefficient (decompose_mat, wavelets.hpp:168-184), and so does this call: col and val then need np * np
 * entries. */
int cqbem_wavelet_compress(cqbem_context *ctx, int64_t n, const double *A, int64_t num_elems, int64_t *rowptr, int32_t *col,
                           double *val, int64_t *nnz);
/* y = M x for a CSR device matrix of `rows` rows (the product in the compressed space,
 * CompressedMatrixVectorProduct, wavelets.hpp:352-356). */
int cqbem_csr_matvec(cqbem_context *ctx, int64_t rows, const int64_t *rowptr, const int32_t *col, const double *val,
                     const double *x, double *y);

#ifdef __cplusplus
}
#endif
#endif
