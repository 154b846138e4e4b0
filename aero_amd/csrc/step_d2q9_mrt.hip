// D2Q9 MRT step kernels: batched fused step.
#include "step_inst.h"

namespace aero {

AERO_INST_BATCH(D2Q9, COLL_MRT, d2q9_mrt)

}  // namespace aero
