// D2Q9 instantiations of the fused step, the materialisation pass and the
// field collision.
#include "launch.h"
#include "step_impl.h"

namespace aero {

namespace {
dim3 grid_of(const StepParams& p) { return dim3((unsigned)p.nblocks_grid); }

template <int COLL, int ST>
hipError_t step_t(bool thermal, int kmode, const StepParams& p, hipStream_t s) {
  if (thermal) {
    if (kmode == 0) step_kernel<D2Q9, COLL, ST, true, 0><<<grid_of(p), BLOCK, 0, s>>>(p);
    else step_kernel<D2Q9, COLL, ST, true, 1><<<grid_of(p), BLOCK, 0, s>>>(p);
  } else {
    if (kmode == 0) step_kernel<D2Q9, COLL, ST, false, 0><<<grid_of(p), BLOCK, 0, s>>>(p);
    else step_kernel<D2Q9, COLL, ST, false, 1><<<grid_of(p), BLOCK, 0, s>>>(p);
  }
  return hipGetLastError();
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
    case COLL_MRT: return step_c<COLL_MRT>(st, thermal, kmode, p, s);
  }
  return hipErrorInvalidValue;
}

}  // namespace aero
