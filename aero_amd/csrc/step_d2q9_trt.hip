// D2Q9 TRT step kernels: batched fused step and the split step for the LES pass.
#include "step_inst.h"

namespace aero {

AERO_INST_BATCH(D2Q9, COLL_TRT, d2q9_trt)
AERO_INST_SPLIT(D2Q9, COLL_TRT, d2q9_trt)

}  // namespace aero
