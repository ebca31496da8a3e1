// Instantiation of the NL-SAR scale pass for the statistics-only passes: tiles of 32 x 32 with 2
// rows per thread, search windows of half width up to 10.
#include "dkx_nlsar_pass.h"

namespace dkx {
DKX_NLSAR_PASS_INSTANTIATE(false, 32, 2, 10, 0)
}  // namespace dkx
