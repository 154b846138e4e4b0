// D2Q9 BGK step kernels: batched fused step and the split step for the LES pass.
#include "step_inst.h"

namespace aero {

AERO_INST_BATCH(D2Q9, COLL_BGK, d2q9_bgk)
AERO_INST_SPLIT(D2Q9, COLL_BGK, d2q9_bgk)

}  // namespace aero
