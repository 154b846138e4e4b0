// D2Q9 TRT batched step kernels with the local Smagorinsky rates.
#include "step_inst.h"

namespace aero {

AERO_INST_BATCH_LES(D2Q9, COLL_TRT, d2q9_trt)

}  // namespace aero
