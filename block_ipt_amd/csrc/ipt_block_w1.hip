// Block-cluster update kernels for cluster widths 2..12 (see
// ipt_block_width.h); split from the other widths so that the unrolled
// instantiations compile in parallel.
#include "ipt_block_width.h"

namespace ipt {

IPT_BLOCK_WIDTH_INSTANTIATE(2)
IPT_BLOCK_WIDTH_INSTANTIATE(3)
IPT_BLOCK_WIDTH_INSTANTIATE(4)
IPT_BLOCK_WIDTH_INSTANTIATE(5)
IPT_BLOCK_WIDTH_INSTANTIATE(6)
IPT_BLOCK_WIDTH_INSTANTIATE(7)
IPT_BLOCK_WIDTH_INSTANTIATE(8)
IPT_BLOCK_WIDTH_INSTANTIATE(9)
IPT_BLOCK_WIDTH_INSTANTIATE(10)
IPT_BLOCK_WIDTH_INSTANTIATE(11)
IPT_BLOCK_WIDTH_INSTANTIATE(12)

}  // namespace ipt
