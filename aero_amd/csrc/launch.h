// Host-side launch entry points for the templated kernels.  Each kernel
// family is instantiated in its own translation unit so the library builds in
// parallel; these functions pick the instantiation at run time.
#pragma once

#include "lbm_device.h"

namespace aero {

struct CollideParams;

// Per (lattice, collision) instantiations (step_<lattice>_<collision>.hip).
#define AERO_DECL_BATCH(NAME)                                                                   \
  hipError_t launch_batch_##NAME(int st, bool thermal, const BatchCase* cases, const int2* tiles, \
                                 int ntiles, int k, hipStream_t s, void* e0, void* e1);
#define AERO_DECL_SPLIT(NAME) \
  hipError_t launch_split_##NAME(int st, bool thermal, const StepParams& p, hipStream_t s);

AERO_DECL_BATCH(d2q9_bgk)
AERO_DECL_BATCH(d2q9_trt)
AERO_DECL_BATCH(d2q9_mrt)
AERO_DECL_BATCH(d3q19_bgk)
AERO_DECL_BATCH(d3q19_trt)
AERO_DECL_BATCH(d3q19_mrt)
AERO_DECL_BATCH(les_d2q9_bgk)
AERO_DECL_BATCH(les_d2q9_trt)
AERO_DECL_BATCH(les_d3q19_bgk)
AERO_DECL_BATCH(les_d3q19_trt)
AERO_DECL_SPLIT(d2q9_bgk)
AERO_DECL_SPLIT(d2q9_trt)
AERO_DECL_SPLIT(d3q19_bgk)
AERO_DECL_SPLIT(d3q19_trt)
#undef AERO_DECL_BATCH
#undef AERO_DECL_SPLIT

// Synthetic-eddy-method inlet planes on the device (sem.hip): a generator for
// one ny x nz inlet holding up to max_planes planes per call.
struct SemGen;
SemGen* sem_create(int ny, int nz, int n_eddies, double sigma, double Tu, double u0, int max_planes);
void sem_destroy(SemGen* g);
// nk planes (nk x 3 x nz*ny doubles, device) from nk host eddy snapshots
// (nk x 6 x n_eddies doubles: x, y, z, then the three signs)
void sem_planes(SemGen* g, const double* eddies_host, int nk, double* planes_dev, hipStream_t s);

// Step k of a chunk for a group of cases sharing lattice, collision, storage
// and scalar transport: one launch over `ntiles` packed blocks.  Events, when
// given, are recorded at the kernel's start and end.
hipError_t launch_batch_step(int q, int coll, int st, bool thermal, const BatchCase* cases, const int2* tiles,
                             int ntiles, int k, hipStream_t s, void* ev_start, void* ev_stop,
                             bool les_local = false);
// Split step of one case (LES finite-difference pass follows; BGK / TRT).
hipError_t launch_split_step(int q, int coll, int st, bool thermal, const StepParams& p, hipStream_t s);
// Last-step force reduction of many cases of one lattice: block b reduces rs[b].
// max_terms > 0: first form the terms of the cases whose record has a terms
// buffer (at most max_terms link cells each)
hipError_t launch_reduce_forces_many(int q, const ForceReduce* rs, int n, int max_terms, hipStream_t s);
// Coefficient rows of a chunk to pinned host memory: for each of n cases the
// first nk rows (NCOEF doubles each) of its block of ch rows, same layout.  A
// kernel storing into the pinned buffer rather than hipMemcpyAsync, whose
// first device-to-host call in a process blocked the host ~6 ms with work
// queued on the stream.
hipError_t launch_rows_to_host(const double* rows, double* host, int n, int ch, int nk, hipStream_t s);
// The 5 D3Q19 populations with e_z = dir of z plane `z` (row-interleaved
// storage, esz-byte elements) to (pack) or from (unpack) a [5][ny][nx] buffer.
hipError_t launch_halo(bool pack, int esz, void* base, void* buf, int nx, int ny, int xp, int z, int dir,
                       hipStream_t s);
hipError_t launch_materialize(int q, int st, bool thermal, const StepParams& p, hipStream_t s);
hipError_t launch_collide(int q, int coll, int st, const CollideParams& p, hipStream_t s);
hipError_t launch_recycling_plane(int st, const StepParams& p, double* out, hipStream_t s);
hipError_t launch_materialize_g(int q, int st, const StepParams& p, double* out, hipStream_t s);
hipError_t launch_store_from_f64(int st, int q, long long n, int nx, int xp, const double* src, void* dst,
                                 double shift_scale, hipStream_t s);
hipError_t launch_compute_from_f64(int st, long long count, const double* src, void* dst, hipStream_t s);
hipError_t launch_ibm_field(int q, int st, long long n, void* F, const uint8_t* flags, double coeff, hipStream_t s);
// Convergence statistics (stats.hip).  A case's drag and lift histories,
// as FP64 ring buffers (newest sample at index (count - 1) mod cap) or plain
// arrays (cap 0, count samples from index 0); lz null in 2D.
struct StatSeries {
  const double* cd;
  const double* ly;
  const double* lz;
  long long count;
  long long cap;
  double D, u0;
};
struct StatParams {
  int conv_window;
  double conv_tol;
  int st_window;
  double st_tol;
};
// Per case NSTAT doubles: stop, Cd converged, ratio, mean |Cd|, std Cd,
// St (NaN when none), peak frequency, peak amplitude, window used, St drift
// between the half windows (NaN when not formed), stationary, lift series
// (1: Cl / Cly, 2: Clz), then the three signals' peak amplitudes.
constexpr int NSTAT = 15;
// The largest Strouhal window the device test takes (its DFT twiddle table
// sits in LDS: n + 2 doubles for a window of n samples).
constexpr int STAT_MAX_WINDOW = 4096;
hipError_t launch_stationarity(const StatSeries* series, int n, const StatParams& sp, double* out, double* work,
                               int* siglen, int* peak, hipStream_t s);
// Append nk coefficient rows (NCOEF doubles per step: Cd, Cl_y, Cl_z first)
// to a case's ring of capacity cap (three series of cap doubles) from
// position start.
struct RingAppend {
  const double* rows;
  double* ring;
  long long start;
  long long cap;
  int nk;
};
hipError_t launch_ring_append(const RingAppend* ra, int n, int max_nk, hipStream_t s);

// STL geometry (mesh.hip).
size_t mesh_inside_scratch_bytes(long long ntri);
hipError_t launch_mesh_inside(const double* tris, long long ntri, void* scratch, const double* pts, long long npts,
                              int nx, int ny, const double ray[3], uint8_t* out, hipStream_t s);
hipError_t launch_mesh_sqdist(const double* tris, long long ntri, const long long* cells, long long ncells, int nx,
                              int ny, double* out, hipStream_t s);
hipError_t launch_feq_fill(int q, long long n, double* f, const double* rho, double rho0, const double* ux,
                           double ux0, const double* uy, const double* uz, hipStream_t s);
hipError_t launch_macro_from_compute(int st, int q, long long n, const void* F, void* U, hipStream_t s);
hipError_t launch_macro4_f64(int q, long long n, const double* F, double* out, hipStream_t s);
hipError_t launch_check_f64(int q, int nx, int ny, int nz, int zc, const double* F, int nblocks, double* part,
                            hipStream_t s);
hipError_t launch_reduce_forces(const ForceReduce& r, hipStream_t s);
hipError_t launch_reduce_check(const double* part, int nblocks, double* out, hipStream_t s);
// Batched form over n cases whose parameters (out_mode 2) sit in device memory:
// the block partials of every tile (case, block) of one (lattice, storage,
// scalar) group, then one fixed-order reduction per case into out[NCHECK * i].
hipError_t launch_check_batch(int q, int st, bool thermal, const StepParams* ps, const int2* tiles, int ntiles,
                              hipStream_t s);
hipError_t launch_reduce_check_batch(const StepParams* ps, int n, double* out, hipStream_t s);
hipError_t launch_reduce_recycling(double* recyc, int nrow, double target_u0, hipStream_t s);

}  // namespace aero
