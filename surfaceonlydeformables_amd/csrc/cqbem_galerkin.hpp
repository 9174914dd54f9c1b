// Galerkin assembly (non-singular and Sauter-Schwab singular face pairs).
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "cqbem_kernels.hpp"
#include "cqbem_mixed.hpp"

namespace cqbem {

struct GalerkinJob {
    Geom g;
    const double4 *quad;  // composite triangle rule (barycentric b0,b1,b2, weight summing to 1)
    int npts;
    const double2 *line;  // composite Gauss rule on [0,1]
    int nline;
    const Freq *freqs;    // first frequency of the job's range
    int nl;               // number of frequencies
    Mat mat;
    int num_cus;
    // pair-form tables of the collocation G/H kernel (column patches; see launch_colloc_gh_pair)
    const double4 *pair_pts;
    int npts_even;
    const int *patch_pass;
    const int2 *pass_info;
    const int4 *pass_faces;
    const int *patch_verts;
    int npatches;
    size_t scratch_bytes;  // bound of the per-test-point row scratch (at least one face is always processed)
    // CQBEM_MIXED: the far-field B at the test points in float (first frequency of the job's range), else null
    const FreqF *freqs32 = nullptr;
    MatF mat32{};
};

// Test faces per chunk of the per-test-point row scratch, and the device bytes the scratch takes
// (test points, their basis weights, and one chunk of G/H/B rows for nl frequencies).
int galerkin_chunk_faces(int nv, int nf, int npts, int nl, bool G, bool H, bool B, size_t budget);
size_t galerkin_scratch_bytes(int nv, int nf, int npts, int nl, bool G, bool H, bool B, size_t budget);

// Accumulates the Laplace-domain spectra of all frequencies of `job` into
// U, P (nl x 3nv x 3nv), Bt (nl x 3nv x 3), Ba (nl x 3nv x 9); any may be null.
// The buffers are zero-filled first.  tm[1] += non-singular ms (G, H and B), tm[3] += singular ms.
// F and vmap (nullable vertex_map_inverse) are the host copies of the job's mesh.
void run_galerkin(const GalerkinJob &job, const int *F, const int *vmap, cd *U, cd *P, cd *Bt, cd *Ba, double *tm,
                  hipStream_t st);

// Per face f1, the faces that share a vertex with it (or a copy of one, with vmap): the candidates of the
// touching-pair classification, f1 itself included, ascending.
std::vector<std::vector<int>> touching_candidates(const int *F, int nf, int nv, const int *vmap);

}  // namespace cqbem
