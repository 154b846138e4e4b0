// D3Q19 MRT step kernels: batched fused step.
#include "step_inst.h"

namespace aero {

AERO_INST_BATCH(D3Q19, COLL_MRT, d3q19_mrt)

}  // namespace aero
