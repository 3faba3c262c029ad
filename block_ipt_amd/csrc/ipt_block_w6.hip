// Block-cluster update kernels for cluster widths 30..32 (see
// ipt_block_width.h); split from the other widths so that the unrolled
// instantiations compile in parallel.
#include "ipt_block_width.h"

namespace ipt {

IPT_BLOCK_WIDTH_INSTANTIATE(30)
IPT_BLOCK_WIDTH_INSTANTIATE(31)
IPT_BLOCK_WIDTH_INSTANTIATE(32)

}  // namespace ipt
