// Block-cluster update kernels for cluster widths 13..17 (see
// ipt_block_width.h); split from the other widths so that the unrolled
// instantiations compile in parallel.
#include "ipt_block_width.h"

namespace ipt {

IPT_BLOCK_WIDTH_INSTANTIATE(13)
IPT_BLOCK_WIDTH_INSTANTIATE(14)
IPT_BLOCK_WIDTH_INSTANTIATE(15)
IPT_BLOCK_WIDTH_INSTANTIATE(16)
IPT_BLOCK_WIDTH_INSTANTIATE(17)

}  // namespace ipt
