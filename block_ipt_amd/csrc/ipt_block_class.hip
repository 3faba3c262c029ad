// Block-cluster update of several widths in one launch: the widths of one
// lane class (3..4, 5..8, 9..16: the same lanes per row, so the same rows
// per workgroup and grid) dispatched per workgroup on the cluster's width.
// Each workgroup runs the width-M code of ipt_block_width.h unchanged, so
// the results are those of the per-width launches; one launch instead of
// one per width present saves the dependent-launch gaps between them.
// The registers are those of the class's widest width.
#include "ipt_block_width.h"

namespace ipt {

namespace {

template <int LO, int HI>
__device__ __forceinline__ void class_rows(int m, int cid, int lblk, const BlockUpdateArgs &g)
{
    if constexpr (LO <= HI) {
        if (m == LO) {
            block_update_rows_m<LO>(cid, lblk, g);
            return;
        }
        class_rows<LO + 1, HI>(m, cid, lblk, g);
    }
}

// blockIdx.y = cluster list[blockIdx.y] (widths LO..HI), blockIdx.x = near
// or far block (the widths of a class have the same near-block count).
template <int LO, int HI>
__global__ __launch_bounds__(256) void k_block_update_class(BlockUpdateArgs g, const int *__restrict__ list)
{
    const int cid = list[blockIdx.y];
    class_rows<LO, HI>(g.clusters[cid].m, cid, blockIdx.x, g);
}

template <int LO, int HI>
int launch_class(const int *list, int count, const BlockUpdateArgs &g, hipStream_t stream)
{
    static_assert(LO > HI / 2, "one lane class per launch");
    static_assert((LO >= kSeriesWidth) == (HI >= kSeriesWidth) && near_blocks_for(LO) == near_blocks_for(HI),
                  "one lane class and one row split per launch");
    const dim3 grid((unsigned)update_blocks_for(HI, g.n), (unsigned)count);
    hipLaunchKernelGGL((k_block_update_class<LO, HI>), grid, dim3(256), 0, stream, g, list);
    return hipGetLastError() == hipSuccess ? IPT_CUDA_SUCCESS : IPT_CUDA_CUDA_ERROR;
}

}  // namespace

int block_update_class(int lo, const int *list, int count, const BlockUpdateArgs &g, hipStream_t stream)
{
    switch (lo) {
    case 3: return launch_class<3, 4>(list, count, g, stream);
    case 5: return launch_class<5, 8>(list, count, g, stream);
    case 9: return launch_class<9, 16>(list, count, g, stream);
    default: return IPT_CUDA_INVALID_ARGUMENT;
    }
}

}  // namespace ipt
