// D3Q19 TRT step kernels: batched fused step and the split step for the LES pass.
#include "step_inst.h"

namespace aero {

AERO_INST_BATCH(D3Q19, COLL_TRT, d3q19_trt)
AERO_INST_SPLIT(D3Q19, COLL_TRT, d3q19_trt)

}  // namespace aero
