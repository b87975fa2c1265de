// xf_sched.h -- block-to-work mapping shared by the per-colour launches
// (tets and vertex rings).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace xf {

constexpr uint32_t kNoSlot = 0xFFFFFFFFu;
constexpr uint32_t kNoNode = 0xFFFFFFFFu;  // padding of a sweep block's node list

// Schedule position of this lane's item in a colour [begin, begin + count),
// or kNoSlot when the lane has none; `group` consecutive lanes share one item
// (a workgroup holds blockDim.x / group items). slabs != 0: workgroup b works
// on eighth (b mod 8) of the colour, chunk b / 8 (grid 8 * ceil(ceil(count /
// 8) / items per workgroup)), so under round-robin dispatch each XCD sweeps
// the same eighth in every colour.
__device__ __forceinline__ uint32_t colour_slot(uint32_t begin, uint32_t count, uint32_t slabs, uint32_t group = 1) {
    uint32_t per = blockDim.x / group, local = threadIdx.x / group;
    uint32_t k;
    if (slabs) {
        uint32_t r = blockIdx.x & 7u, j = blockIdx.x >> 3;
        uint32_t lo = (uint32_t)(((uint64_t)count * r) >> 3), hi = (uint32_t)(((uint64_t)count * (r + 1u)) >> 3);
        k = lo + j * per + local;
        if (k >= hi) return kNoSlot;
    } else {
        k = blockIdx.x * per + local;
        if (k >= count) return kNoSlot;
    }
    return begin + k;
}

// Grid of a per-colour launch of `block` items per workgroup (see colour_slot).
inline uint32_t colour_grid(uint32_t count, uint32_t block, uint32_t slabs) {
    if (!slabs) return (count + block - 1) / block;
    uint32_t eighth = (count + 7) / 8;
    return 8 * ((eighth + block - 1) / block);
}

}  // namespace xf
