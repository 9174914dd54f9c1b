// Float far-field kernels of the mixed-precision mode (CQBEM_MIXED) and their float constants.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "cqbem_kernels.hpp"

namespace cqbem {

// Per-frequency constants rounded to float: s/c1, s/c2, 1/s, 1/s^2 (real, imaginary).
struct FreqF {
    float sc1x, sc1y, sc2x, sc2y, isx, isy, is2x, is2y;
};
// Material constants rounded to float: 1/c1, 1/c2, 1/c1^2, 1/c2^2, 1/(4 pi rho).
struct MatF {
    float ic1, ic2, ic1s, ic2s, g_scale;
};

std::vector<FreqF> freqs_f32(const std::vector<Freq> &f);
MatF mat_f32(const Mat &m);

// G far field of collocation rows row0..row0+nrows-1 at frequencies l0..l0+nl-1, written (as double) like
// launch_colloc_gh's G part: every (row, column) block of the regular faces, the singular faces excluded.
hipError_t launch_colloc_g32(const Geom &g, const double4 *quad, int npts, const FreqF *freqs, int l0, int nl, const MatF &m,
                             int row0, int nrows, cd *UG, long long ld, long long lstride, hipStream_t st);
// B_trans / B_angular far field of collocation rows, written (as double) like launch_colloc_b.
hipError_t launch_colloc_b32(const Geom &g, const double4 *quad, int npts, const FreqF *freqs, int l0, int nl, const MatF &m,
                             int row0, int nrows, cd *Bt, cd *Ba, hipStream_t st, const FaceBands &fb = FaceBands{});
// The same at Galerkin test points (as launch_points_b): rows tpts[row0 .. row0 + nrows - 1], out_rows rows per
// frequency in Bt/Ba.
hipError_t launch_points_b32(const Geom &g, const double4 *quad, int npts, const FreqF *freqs, int l0, int nl, const MatF &m,
                             const double4 *tpts, int row0, int nrows, cd *Bt, cd *Ba, long long out_rows, hipStream_t st);

}  // namespace cqbem
