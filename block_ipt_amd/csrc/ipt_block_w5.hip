// Block-cluster update kernels for cluster widths 26..29 (see
// ipt_block_width.h); split from the other widths so that the unrolled
// instantiations compile in parallel.
#include "ipt_block_width.h"

namespace ipt {

IPT_BLOCK_WIDTH_INSTANTIATE(26)
IPT_BLOCK_WIDTH_INSTANTIATE(27)
IPT_BLOCK_WIDTH_INSTANTIATE(28)
IPT_BLOCK_WIDTH_INSTANTIATE(29)

}  // namespace ipt
