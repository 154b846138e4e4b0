// Host-side launch entry points for the templated kernels.  Each kernel
// family is instantiated in its own translation unit so the library builds in
// parallel; these functions pick the instantiation at run time.
#pragma once

#include "lbm_device.h"

namespace aero {

struct CollideParams;

// kmode 0: fused step, 1: split step (LES finite-difference pass follows)
hipError_t launch_step_d2q9(int coll, int st, bool thermal, int kmode, const StepParams& p, hipStream_t s);
hipError_t launch_step_d3q19(int coll, int st, bool thermal, int kmode, const StepParams& p, hipStream_t s);
hipError_t launch_materialize(int q, int st, bool thermal, const StepParams& p, hipStream_t s);
hipError_t launch_collide(int q, int coll, int st, const CollideParams& p, hipStream_t s);
hipError_t launch_recycling_plane(int st, const StepParams& p, double* out, hipStream_t s);
hipError_t launch_materialize_g(int q, int st, const StepParams& p, double* out, hipStream_t s);
hipError_t launch_store_from_f64(int st, int q, long long n, const double* src, void* dst, double shift_scale,
                                 hipStream_t s);
hipError_t launch_compute_from_f64(int st, long long count, const double* src, void* dst, hipStream_t s);
hipError_t launch_macro_from_compute(int st, int q, long long n, const void* F, void* U, hipStream_t s);
hipError_t launch_macro4_f64(int q, long long n, const double* F, double* out, hipStream_t s);
hipError_t launch_check_f64(int q, int nx, int ny, int nz, int zc, const double* F, int nblocks, double* part,
                            hipStream_t s);
hipError_t launch_reduce_forces(const ForceReduce& r, hipStream_t s);
hipError_t launch_reduce_check(const double* part, int nblocks, double* out, hipStream_t s);
hipError_t launch_reduce_recycling(double* recyc, int nrow, double target_u0, hipStream_t s);

}  // namespace aero
