// D3Q19 BGK step kernels: batched fused step and the split step for the LES pass.
#include "step_inst.h"

namespace aero {

AERO_INST_BATCH(D3Q19, COLL_BGK, d3q19_bgk)
AERO_INST_SPLIT(D3Q19, COLL_BGK, d3q19_bgk)

}  // namespace aero
