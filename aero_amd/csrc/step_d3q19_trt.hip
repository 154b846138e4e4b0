// D3Q19 TRT instantiations of the fused step.
#include "launch.h"
#include "step_impl.h"

namespace aero {

hipError_t launch_step_d3q19_trt(int st, bool thermal, int kmode, const StepParams& p, hipStream_t s) {
  const dim3 g((unsigned)p.nblocks_grid);
#define AERO_CASE(ST)                                                                    \
  case ST:                                                                               \
    if (thermal) {                                                                       \
      if (kmode == 0) step_kernel<D3Q19, COLL_TRT, ST, true, 0><<<g, BLOCK, 0, s>>>(p);  \
      else step_kernel<D3Q19, COLL_TRT, ST, true, 1><<<g, BLOCK, 0, s>>>(p);             \
    } else {                                                                             \
      if (kmode == 0) step_kernel<D3Q19, COLL_TRT, ST, false, 0><<<g, BLOCK, 0, s>>>(p); \
      else step_kernel<D3Q19, COLL_TRT, ST, false, 1><<<g, BLOCK, 0, s>>>(p);            \
    }                                                                                    \
    return hipGetLastError();
  switch (st) {
    AERO_CASE(ST_F64)
    AERO_CASE(ST_F32)
    AERO_CASE(ST_F16S)
    AERO_CASE(ST_BF16S)
  }
#undef AERO_CASE
  return hipErrorInvalidValue;
}

}  // namespace aero
