// Instantiation of the NL-SAR scale pass for the statistics-only passes: tiles of 32 x 32 with 4
// rows per thread, search windows of half width up to 20.
#include "dkx_nlsar_pass.h"

namespace dkx {
DKX_NLSAR_PASS_INSTANTIATE(false, 32, 4, 20, 0)
}  // namespace dkx
