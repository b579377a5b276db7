// Chebyshev launchers and kernels of the single propagation, FP64 and compact-term instances
// (the code is in qdx_cheby_impl.hpp).
#include "qdx_cheby_impl.hpp"

namespace qdx {
QDX_INST_CHEBY(double)
}  // namespace qdx
