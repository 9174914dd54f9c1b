#include "cqbem_galerkin.hpp"

namespace cqbem {
void run_galerkin(const GalerkinJob &, const std::vector<int> &, const std::vector<int> &, const int *, cd *, cd *, cd *, cd *,
                  double *, hipStream_t) {}
}  // namespace cqbem
