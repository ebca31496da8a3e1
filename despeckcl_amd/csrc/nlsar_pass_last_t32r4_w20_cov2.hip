// Instantiation of the NL-SAR scale pass for the last pass on dimension-2 covariance
// input: tiles of 32 x 32 with 4 rows per thread, search windows of half width up to 20.
#include "dkx_nlsar_pass.h"

namespace dkx {
DKX_NLSAR_PASS_INSTANTIATE(true, 32, 4, 20, 2)
}  // namespace dkx
