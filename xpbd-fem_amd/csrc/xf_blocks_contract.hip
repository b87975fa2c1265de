// xf_blocks_contract.hip -- the sweep-block projection with multiply-add
// contraction (xf_scene_opts.contract).
//
// The kernels of xf_kernels.hip compiled a second time, with
// -ffp-contract=fast, inside namespace xf_contract (so that their symbols
// differ from the exact ones), keeping only the sweep-block launcher,
// exported as xf_launch_blocks_contract with the arguments of
// xf_launch_blocks. The element chain then uses fused multiply-adds: about a
// quarter fewer VALU instructions per tet, and results that no longer equal
// the reference's serial FP32 sweep bit for bit (each product-sum rounds
// once instead of twice).
#include <hip/hip_runtime.h>

#include "xf_device.h"
#include "xf_element.h"
#include "xf_sched.h"

#define XF_BLOCKS_ONLY 1
#define XF_BLOCKS_LAUNCH xf_launch_blocks_contract

namespace xf_contract {
#include "xf_kernels.hip"
}  // namespace xf_contract
