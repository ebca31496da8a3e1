// Instantiation of the NL-SAR scale pass for the last pass: tiles of 32 x 32 with 2
// rows per thread, search windows of half width up to 10.
#include "dkx_nlsar_pass.h"

namespace dkx {
DKX_NLSAR_PASS_INSTANTIATE(true, 32, 2, 10, 0)
}  // namespace dkx
