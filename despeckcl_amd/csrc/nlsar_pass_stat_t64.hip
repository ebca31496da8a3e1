// Instantiation of the NL-SAR scale pass for the statistics-only passes, tiles of 32 x 64.
#include "dkx_nlsar_pass.h"

namespace dkx {
DKX_NLSAR_PASS_INSTANTIATE(false, 64)
}  // namespace dkx
