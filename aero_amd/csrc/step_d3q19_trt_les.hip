// D3Q19 TRT batched step kernels with the local Smagorinsky rates.
#include "step_inst.h"

namespace aero {

AERO_INST_BATCH_LES(D3Q19, COLL_TRT, d3q19_trt)

}  // namespace aero
