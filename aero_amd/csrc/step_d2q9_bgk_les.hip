// D2Q9 BGK batched step kernels with the local Smagorinsky rates.
#include "step_inst.h"

namespace aero {

AERO_INST_BATCH_LES(D2Q9, COLL_BGK, d2q9_bgk)

}  // namespace aero
