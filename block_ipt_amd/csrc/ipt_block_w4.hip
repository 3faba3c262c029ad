// Block-cluster update kernels for cluster widths 22..25 (see
// ipt_block_width.h); split from the other widths so that the unrolled
// instantiations compile in parallel.
#include "ipt_block_width.h"

namespace ipt {

IPT_BLOCK_WIDTH_INSTANTIATE(22)
IPT_BLOCK_WIDTH_INSTANTIATE(23)
IPT_BLOCK_WIDTH_INSTANTIATE(24)
IPT_BLOCK_WIDTH_INSTANTIATE(25)

}  // namespace ipt
