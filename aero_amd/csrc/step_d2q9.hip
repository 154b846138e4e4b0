// D2Q9 instantiations of the step kernels (one fused launch per step).
#include "launch.h"
#include "step_impl.h"

namespace aero {

namespace {
template <int COLL, int ST, bool TH, int KM>
hipError_t pair(const StepParams& p, hipStream_t s) {
  return launch_step<D2Q9, COLL, ST, TH, KM>(p, s);
}

template <int COLL, int ST>
hipError_t step_t(bool thermal, int kmode, const StepParams& p, hipStream_t s) {
  if (thermal) return kmode == 0 ? pair<COLL, ST, true, 0>(p, s) : pair<COLL, ST, true, 1>(p, s);
  return kmode == 0 ? pair<COLL, ST, false, 0>(p, s) : pair<COLL, ST, false, 1>(p, s);
}

template <int COLL>
hipError_t step_c(int st, bool thermal, int kmode, const StepParams& p, hipStream_t s) {
  switch (st) {
    case ST_F64: return step_t<COLL, ST_F64>(thermal, kmode, p, s);
    case ST_F32: return step_t<COLL, ST_F32>(thermal, kmode, p, s);
    case ST_F16S: return step_t<COLL, ST_F16S>(thermal, kmode, p, s);
    case ST_BF16S: return step_t<COLL, ST_BF16S>(thermal, kmode, p, s);
  }
  return hipErrorInvalidValue;
}
}  // namespace

hipError_t launch_step_d2q9(int coll, int st, bool thermal, int kmode, const StepParams& p, hipStream_t s) {
  switch (coll) {
    case COLL_BGK: return step_c<COLL_BGK>(st, thermal, kmode, p, s);
    case COLL_TRT: return step_c<COLL_TRT>(st, thermal, kmode, p, s);
    case COLL_MRT: return step_c<COLL_MRT>(st, thermal, 0, p, s);  // MRT never takes the LES field
  }
  return hipErrorInvalidValue;
}

}  // namespace aero
