// D3Q19 BGK instantiations of the step kernels (one fused launch per step).
#include "launch.h"
#include "step_impl.h"

namespace aero {

namespace {
template <int ST, bool TH, int KM>
hipError_t pair(const StepParams& p, hipStream_t s) {
  return launch_step<D3Q19, COLL_BGK, ST, TH, KM>(p, s);
}

template <int ST>
hipError_t step_t(bool thermal, int kmode, const StepParams& p, hipStream_t s) {
  return thermal ? (kmode == 0 ? pair<ST, true, 0>(p, s) : pair<ST, true, 1>(p, s))
                 : (kmode == 0 ? pair<ST, false, 0>(p, s) : pair<ST, false, 1>(p, s));
}
}  // namespace

hipError_t launch_step_d3q19_bgk(int st, bool thermal, int kmode, const StepParams& p, hipStream_t s) {
  switch (st) {
    case ST_F64: return step_t<ST_F64>(thermal, kmode, p, s);
    case ST_F32: return step_t<ST_F32>(thermal, kmode, p, s);
    case ST_F16S: return step_t<ST_F16S>(thermal, kmode, p, s);
    case ST_BF16S: return step_t<ST_BF16S>(thermal, kmode, p, s);
  }
  return hipErrorInvalidValue;
}

}  // namespace aero
