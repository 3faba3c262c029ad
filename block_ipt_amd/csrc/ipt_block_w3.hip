// Block-cluster update kernels for cluster widths 18..21 (see
// ipt_block_width.h); split from the other widths so that the unrolled
// instantiations compile in parallel.
#include "ipt_block_width.h"

namespace ipt {

IPT_BLOCK_WIDTH_INSTANTIATE(18)
IPT_BLOCK_WIDTH_INSTANTIATE(19)
IPT_BLOCK_WIDTH_INSTANTIATE(20)
IPT_BLOCK_WIDTH_INSTANTIATE(21)

}  // namespace ipt
