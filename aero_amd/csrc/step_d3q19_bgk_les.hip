// D3Q19 BGK batched step kernels with the local Smagorinsky rates.
#include "step_inst.h"

namespace aero {

AERO_INST_BATCH_LES(D3Q19, COLL_BGK, d3q19_bgk)

}  // namespace aero
