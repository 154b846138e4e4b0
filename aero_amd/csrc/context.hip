// Host side of the C ABI: context state, geometry records, the step loop and
// state materialisation.  See include/aero_hip.h for the contract.
#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "aero_hip.h"
#include "launch.h"
#include "step_impl.h"

using namespace aero;

namespace {

thread_local std::string g_err;

struct HipError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

void ck(hipError_t e, const char* what) {
  if (e != hipSuccess) {
    throw HipError(std::string(what) + ": " + hipGetErrorString(e));
  }
}

template <class F>
int guarded(F&& fn) {
  try {
    fn();
    return 0;
  } catch (const std::exception& ex) {
    g_err = ex.what();
    return 1;
  }
}

// Steps per chunk of aero_step (device history rows and events are reused per chunk).
constexpr int STEP_CHUNK = 256;

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  void alloc(size_t b) {
    if (bytes >= b && p) return;
    release();
    if (b == 0) return;
    ck(hipMalloc(&p, b), "hipMalloc");
    bytes = b;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  template <class T>
  T* as() const { return (T*)p; }
};

// Antisymmetric TRT rate.  Variant 0 is the reference's formula
// (trt2d.py:8-18: tau+ = lam tau + 1 - lam, tau- = (tau+ tau - 1/2)/(tau+ - 1/2));
// variant 1 is the standard magic-parameter TRT, tau- = lam/(tau - 1/2) + 1/2.
double trt_s_minus(double omega, double lam, int variant) {
  const double tau = 1.0 / omega;
  if (variant == 1) return 1.0 / (lam / (tau - 0.5) + 0.5);
  const double tp = lam * tau + (1.0 - lam);
  const double tm = (tp * tau - 0.5) / (tp - 0.5);
  return 1.0 / tm;
}

constexpr size_t NT_STORE_BYTES = (size_t)192 << 20;

// MI355X hands the workgroups of a launch to its 8 XCDs round-robin (block b
// to XCD b mod 8), and each XCD has its own L2.  Launch slots are grouped in
// rounds of XCDS * run (xcd_block): within a round XCD x takes `run`
// consecutive tiles of the natural order, so the rows a block pulls from its
// neighbours (x +- 1, y +- 1) were read by the same XCD a moment earlier.
// One round over the whole list gives every XCD one contiguous eighth of the
// cells (best when the z +- 1 planes of all eight eighths fit the last-level
// cache together: 2D grids, small 3D ones); short runs keep the whole device
// sweeping the grid together, so large 3D grids still find the planes at
// z +- 1 in that cache.
constexpr int XCD_RUN_LARGE = 8;
constexpr size_t ZREUSE_BYTES = (size_t)64 << 20;
static void xcd_order(const std::vector<int2>& nat, int run, std::vector<int2>& out) {
  const int n = (int)nat.size();
  for (int pos = 0; pos < n; ++pos) out.push_back(nat[xcd_block(pos, run, n)]);
}

size_t storage_bytes(int st) {
  switch (st) {
    case ST_F64: return 8;
    case ST_F32: return 4;
    default: return 2;
  }
}

}  // namespace

// Launch tuning of the batched step, set by aero_tune_set (the autotuner's
// table): per kernel group -- lattice, collision, storage, scalar transport --
// and grid class (floor(log2) of the group's tile count), how the group's
// tiles are ordered and launched.  The choices change only the schedule:
// every cell computes the same bits whatever they are.
struct TuneVal {
  int xcd_run = 0;  // 0: built-in rule; -1: natural order; k > 0: runs of k tiles per XCD
  int nt = -1;      // -1: built-in rule (non-temporal stores past NT_STORE_BYTES); 0 / 1: off / on
  int graph = 0;    // 1: a chunk's step launches replay from a captured HIP graph
  int rowgroups = -1;  // -1: grouped row waves where the kernel lists its waves; 0: one row per wave
};
namespace {
std::mutex g_tune_mu;
std::map<std::array<int, 5>, TuneVal> g_tune;
std::atomic<unsigned> g_tune_gen{1};

int grid_class(long long ntiles) {
  int c = 0;
  while ((2ll << c) <= ntiles) ++c;
  return c;
}

TuneVal tune_lookup(int q, int coll, int st, int th, int cls) {
  std::lock_guard<std::mutex> lock(g_tune_mu);
  auto it = g_tune.find({q, coll, st, th, cls});
  return it == g_tune.end() ? TuneVal{} : it->second;
}
}  // namespace

// Buffers for stepping a set of cases in chunks of at most STEP_CHUNK steps:
// the batched kernel's case table (pinned host copy + device copy, double
// buffered) and packed tile list, double-buffered coefficient rows with
// pinned staging, and the last-step force reductions.  They only grow, so
// stepping a set of cases no larger than before allocates nothing.
struct StepRunner {
  DevBuf cases[2], tiles, rows[2], reds[2], rapp[2];
  BatchCase* cases_host[2] = {nullptr, nullptr};
  ForceReduce* reds_host[2] = {nullptr, nullptr};
  RingAppend* rapp_host[2] = {nullptr, nullptr};
  double* stage[2] = {nullptr, nullptr};
  size_t ncase_cap = 0, rows_cap = 0;
  hipEvent_t ev[2] = {nullptr, nullptr};
  // creation serials of the case list the tile table was built for: a serial
  // is never reused, so a context created at a freed context's address never
  // matches a stale table
  std::vector<unsigned long long> tile_key;
  unsigned tune_gen = 0;  // tuning table generation the tile table was built under
  // kernel groups of that case list: lattice, collision, storage, scalar
  // transport, the group's slice of the tile table and its launch tuning
  struct Group {
    int q, coll, st, th, les, toff, ntiles;  // les: local Smagorinsky kernels
    TuneVal tune;
    bool holds(const aero_ctx* c) const;
  };
  std::vector<Group> groups;
  std::vector<char> listed;  // per case of that list: its step walks the case's wave list
  // captured chunks (graph launch tuning): by buffer parity and chunk length,
  // valid for the tile table and case-table buffers they were captured with
  struct Captured {
    hipGraphExec_t exec = nullptr;
    const void* cases = nullptr;
    const void* tiles = nullptr;
  };
  std::map<std::pair<int, int>, Captured> graphs;
  // Streams beside the runner's own for kernel groups stepped concurrently:
  // groups hold disjoint cases, so each group's steps need only their own
  // order, and one group's launches fill the device while another's drain.
  // Forked from and joined back to the runner's stream by events per chunk.
  static constexpr int NSIDE = 3;
  hipStream_t side[NSIDE] = {};
  hipEvent_t fork_ev = nullptr, join_ev[NSIDE] = {};
  void ensure_side(int m) {
    if (!fork_ev) ck(hipEventCreateWithFlags(&fork_ev, hipEventDisableTiming), "hipEventCreate");
    for (int i = 0; i < m && i < NSIDE; ++i) {
      if (!side[i]) ck(hipStreamCreateWithFlags(&side[i], hipStreamNonBlocking), "hipStreamCreate");
      if (!join_ev[i]) ck(hipEventCreateWithFlags(&join_ev[i], hipEventDisableTiming), "hipEventCreate");
    }
  }
  void drop_graphs() {
    for (auto& kv : graphs)
      if (kv.second.exec) (void)hipGraphExecDestroy(kv.second.exec);
    graphs.clear();
  }

  template <class T>
  static void host_grow(T*& p, size_t n) {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    ck(hipHostMalloc((void**)&p, sizeof(T) * std::max<size_t>(1, n), hipHostMallocDefault), "hipHostMalloc");
  }
  void reserve(int ncase, int ch) {
    const size_t nrows = (size_t)ncase * ch * NCOEF;
    for (int b = 0; b < 2; ++b) {
      if (!ev[b]) ck(hipEventCreateWithFlags(&ev[b], hipEventDisableTiming), "hipEventCreate");
      cases[b].alloc(sizeof(BatchCase) * ncase);
      reds[b].alloc(sizeof(ForceReduce) * ncase);
      rapp[b].alloc(sizeof(RingAppend) * ncase);
      rows[b].alloc(sizeof(double) * nrows);
    }
    if ((size_t)ncase > ncase_cap) {
      for (int b = 0; b < 2; ++b) {
        host_grow(cases_host[b], ncase);
        host_grow(reds_host[b], ncase);
        host_grow(rapp_host[b], ncase);
      }
      ncase_cap = ncase;
    }
    if (nrows > rows_cap) {
      for (int b = 0; b < 2; ++b) host_grow(stage[b], nrows);
      rows_cap = nrows;
    }
  }
  void release() {
    for (int b = 0; b < 2; ++b) {
      cases[b].release();
      rows[b].release();
      reds[b].release();
      rapp[b].release();
      if (cases_host[b]) (void)hipHostFree(cases_host[b]);
      if (reds_host[b]) (void)hipHostFree(reds_host[b]);
      if (rapp_host[b]) (void)hipHostFree(rapp_host[b]);
      rapp_host[b] = nullptr;
      if (stage[b]) (void)hipHostFree(stage[b]);
      if (ev[b]) (void)hipEventDestroy(ev[b]);
      cases_host[b] = nullptr;
      reds_host[b] = nullptr;
      stage[b] = nullptr;
      ev[b] = nullptr;
    }
    drop_graphs();
    for (int i = 0; i < NSIDE; ++i) {
      if (side[i]) {
        (void)hipStreamSynchronize(side[i]);
        (void)hipStreamDestroy(side[i]);
      }
      if (join_ev[i]) (void)hipEventDestroy(join_ev[i]);
      side[i] = nullptr;
      join_ev[i] = nullptr;
    }
    if (fork_ev) (void)hipEventDestroy(fork_ev);
    fork_ev = nullptr;
    tiles.release();
    ncase_cap = rows_cap = 0;
    tile_key.clear();
    groups.clear();
  }
};

static std::atomic<unsigned long long> g_ctx_serial{0};

// Streams of destroyed contexts, per device, handed to the next contexts
// created there: creating and destroying a HIP stream costs about a
// millisecond, which a serving loop that retires and admits cases every
// chunk would pay per case.  A pooled stream is idle (synchronised before it
// is returned) and keeps the default priority.
namespace {
std::mutex g_stream_mu;
std::map<int, std::vector<hipStream_t>> g_stream_pool;

hipStream_t take_stream(int device) {
  {
    std::lock_guard<std::mutex> lock(g_stream_mu);
    std::vector<hipStream_t>& v = g_stream_pool[device];
    if (!v.empty()) {
      hipStream_t s = v.back();
      v.pop_back();
      return s;
    }
  }
  hipStream_t s = nullptr;
  ck(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
  return s;
}

void return_stream(int device, hipStream_t s) {
  if (!s) return;
  (void)hipStreamSynchronize(s);
  std::lock_guard<std::mutex> lock(g_stream_mu);
  g_stream_pool[device].push_back(s);
}
}  // namespace

struct aero_ctx {
  aero_desc d;
  unsigned long long serial = ++g_ctx_serial;  // unique per context for the process lifetime
  int device = 0;
  hipStream_t stream = nullptr;
  int Q = 0;
  long long N = 0;
  int nrow = 0;
  int nblocks = 0;  // blocks over the N cells (per-cell kernels)
  int xp = 0;       // padded row length (storage rows and the step's thread grid)
  int sblocks = 0;       // blocks over the step's waves: the listed waves, else nrow * xp threads
  int sblocks_rows = 0;  // blocks over nrow * xp threads (one row per wave)
  int nitems = 0;        // grouped kernels: listed waves (witems); 0 for the others
  unsigned long long layout_gen = 0;  // bumped when the step's wave list changes
  size_t esz = 8;  // storage element size
  size_t csz = 8;  // compute element size
  // state: ping-pong sides, `side` is the one the next step reads
  DevBuf B[2], G[2], prev[2];
  int side = 0;
  int prev_side = 0;
  // geometry
  DevBuf witems;  // grouped kernels: the step's wave list
  DevBuf flags, wave_any, lrec, rec_of, qdir, fpart[2], sigma, minv_d, minv_f, damp;
  DevBuf les_c2d, les_c2f;  // local Smagorinsky: (C_s d)^2 per cell with Van Driest damping d
  DevBuf halo_buf;          // aero_halo_exchange staging (this slab as the lower one)
  DevBuf fterms;            // NFORCE doubles per link cell for the chunk-end force reduction
  int fpart_side = 0;  // buffer holding the last completed step's per-link-cell force sums
  int nrec = 0;
  std::vector<int> rec_cells;
  std::vector<uint32_t> rec_masks;
  // per-step inputs
  // per-step inputs and outputs, in chunks of at most STEP_CHUNK steps:
  // double-buffered device coefficient rows and pinned staging, so the
  // device never waits on the host between chunks and memory stays bounded
  DevBuf planes[2], recyc, hist_dev[2], last_plane_buf;
  SemGen* sem = nullptr;  // device SEM inlet (aero_set_sem): `planes` arguments carry eddy snapshots
  double* hist_stage[2] = {nullptr, nullptr};
  hipEvent_t chunk_ev[2] = {nullptr, nullptr};
  const double* last_plane = nullptr;
  long long last_step_no = 0;
  int last_g_skip = 0;
  int g_skip_pending = 0;
  bool has_state = false;  // set_f has defined the state
  int ibm_on = 0;          // immersed-boundary mode (aero_set_ibm)
  int sponge_x0 = 0;       // first sponge column (nx: none)
  double ibm_coeff = 0.0;
  // FP64 rings of the drag and lift histories (aero_set_history_ring): three
  // series of ring_cap doubles, sample s of the stream at index s % ring_cap
  DevBuf ring, ring_rec;  // ring_rec: the LES path's two append records
  long long ring_cap = 0;
  std::atomic<long long> ring_count{0};  // read by aero_stationarity while a step call appends
  RingAppend ring_append(const double* rows, int nk) const {
    return RingAppend{rows, ring.as<double>(), ring_count.load(), ring_cap, nk};
  }
  // LES split buffers
  DevBuf fsplit, usplit;
  // caches of the reference state in f64
  DevBuf fcache, gcache, tmp;
  bool fcache_valid = false, gcache_valid = false;
  DevBuf checkpart, checkout;
  // optional event timing of the step kernel
  bool timing = false;
  StepRunner runner;  // single-case stepping (a batch of one)
  std::vector<hipEvent_t> ev_pool;
  double timed_ms = 0.0;
  long long timed_launches = 0;

  // bytes of one population buffer in the storage layout
  size_t state_bytes() const { return esz * Q * (size_t)nrow * xp; }
  // the reference's finite-difference LES: a split step (MRT ignores LES)
  bool les_split() const {
    return (d.les == AERO_LES_SMAGORINSKY || d.les == AERO_LES_WALE) && d.collision != AERO_MRT;
  }
  // Smagorinsky from the local non-equilibrium stress, fused into the step
  bool les_local() const { return d.les == AERO_LES_SMAGORINSKY_LOCAL && d.collision != AERO_MRT; }
  bool need_planes() const { return d.inlet == AERO_IN_FIELD || d.inlet == AERO_IN_SEM; }
  // Inlet planes of steps s0 .. s0+n-1 of a call into planes[b]: uploaded
  // from the host planes, or, with the device SEM, generated from the host's
  // per-step eddy snapshots.
  void fill_planes(int b, const double* host, int s0, int n, hipStream_t s) {
    if (sem) {
      sem_planes(sem, host + (size_t)6 * sem_n * s0, n, planes[b].as<double>(), s);
      return;
    }
    const size_t plane_cnt = 3 * (size_t)nrow;
    ck(hipMemcpyAsync(planes[b].p, host + plane_cnt * s0, sizeof(double) * plane_cnt * n, hipMemcpyHostToDevice, s),
       "upload planes");
  }
  int sem_n = 0;
  // inlet planes are 3 x nrow doubles per step: smaller chunks bound their buffers
  int chunk_steps() const { return need_planes() ? STEP_CHUNK / 4 : STEP_CHUNK; }
  bool convective() const { return d.outlet == AERO_OUT_CONVECTIVE; }

  void set_device() const { ck(hipSetDevice(device), "hipSetDevice"); }

  StepParams base_params() const {
    StepParams p;
    std::memset(&p, 0, sizeof(p));
    p.nx = d.nx;
    p.ny = d.ny;
    p.nz = d.nz;
    p.nrow = nrow;
    p.xp = xp;
    p.qxp = Q * xp;
    p.div_xp = make_div_magic((unsigned)xp);
    p.sblocks = sblocks;
    p.nitems = nitems;
    p.witems = (decltype(p.witems))(witems.as<uint32_t>());
    p.ibm_on = ibm_on;
    p.ibm_coeff = ibm_coeff;
    p.sponge_x0 = sponge_x0;
    p.div_nx = make_div_magic((unsigned)d.nx);
    p.div_ny = make_div_magic((unsigned)d.ny);
    p.ncell = N;
    p.nblocks = nblocks;
    p.flags = (decltype(p.flags))(flags.as<uint8_t>());
    p.wave_any = (decltype(p.wave_any))(wave_any.as<uint32_t>());
    p.lrec = (decltype(p.lrec))(lrec.as<LinkCell>());
    p.rec_of = (decltype(p.rec_of))(rec_of.as<int>());
    p.qdir = (decltype(p.qdir))(qdir.as<float>());
    p.fpart = (decltype(p.fpart))(fpart[fpart_side].as<double>());
    p.nrec = nrec;
    p.fdenom = force_denom();
    p.mdenom = moment_denom();
    p.sigma = (decltype(p.sigma))(d.sponge_thickness > 0 ? sigma.as<double>() : nullptr);
    p.minv_d = (decltype(p.minv_d))(minv_d.as<double>());
    p.minv_f = (decltype(p.minv_f))(minv_f.as<float>());
    for (int a = 0; a < 19; ++a) p.mrt_s[a] = d.mrt_s[a];
    for (int a = 0; a < Q; ++a) {
      if (Q == 19) p.fsh[fsh_class<D3Q19>(a)] = (float)fshift<D3Q19, double>(a, d.u0);
      else p.fsh[fsh_class<D2Q9>(a)] = (float)fshift<D2Q9, double>(a, d.u0);
    }
    p.omega = d.omega;
    p.trt_lambda = d.trt_lambda;
    p.trt_sm = trt_s_minus(d.omega, d.trt_lambda, d.trt_variant);
    p.trt_variant = d.trt_variant;
    p.les_tau0 = 1.0 / d.omega;
    p.les_c2 = d.les_cs * d.les_cs;
    p.les_c2d = (decltype(p.les_c2d))(les_c2d.as<double>());
    p.les_c2f = (decltype(p.les_c2f))(les_c2f.as<float>());
    p.u0 = d.u0;
    p.u_conv = d.u0;
    p.inlet_amp = d.inlet_perturbation > 0.0 ? d.inlet_perturbation : 0.0;
    p.rho_in = d.rho_in;
    p.rho_out = d.rho_out;
    p.u_top = d.u_wall_top;
    p.u_bot = d.u_wall_bottom;
    p.bf[0] = d.body_force[0];
    p.bf[1] = d.body_force[1];
    p.bf[2] = d.body_force[2];
    p.bf_coeff = 1.0 - 0.5 / (1.0 / d.omega);
    p.omega_T = 1.0 / (3.0 * d.alpha_T + 0.5);
    p.gbeta = d.g_gravity * d.beta;
    p.T_ref = d.T_ref;
    p.T_hot = d.T_hot;
    p.T_cold = d.T_cold;
    p.gshift = (d.storage == ST_F16S || d.storage == ST_BF16S) ? d.T_ref : 0.0;
    p.center[0] = d.center[0];
    p.center[1] = d.center[1];
    p.center[2] = d.center[2];
    p.inlet_mode = d.inlet;
    p.outlet_mode = d.outlet;
    p.wall_mode = d.wall;
    p.body_force_on = (Q == 19 && d.body_force_on) ? 1 : 0;
    p.buoyancy = d.buoyancy;
    p.bouzidi = d.bouzidi;
    p.check_z = -1;
    p.recyc = (decltype(p.recyc))(recyc.as<double>());
    return p;
  }

  double force_denom() const {
    return Q == 19 ? 0.5 * d.rho0 * d.u0 * d.u0 * d.D * d.D : 0.5 * d.rho0 * d.u0 * d.u0 * d.D;
  }
  double moment_denom() const {
    return Q == 19 ? 0.5 * d.rho0 * d.u0 * d.u0 * d.D * d.D * d.D : 0.5 * d.rho0 * d.u0 * d.u0 * d.D * d.D;
  }

  void launch_split_kernel(const StepParams& p) {
    ck(launch_split_step(Q, d.collision, d.storage, d.thermal != 0, p, stream), "split step launch");
  }

  // Parameters of the batched kernel for the next steps from the current
  // state: p[0] for even steps of the chunk, p[1] for odd ones.
  BatchCase make_case(long long step0, const double* planes_dev, double* hist_rows) const {
    BatchCase bc;
    std::memset(&bc, 0, sizeof(bc));
    for (int par = 0; par < 2; ++par) {
      StepParams p = base_params();
      const int rd = (side + par) & 1;
      p.fin = (decltype(p.fin))(B[rd].p);
      p.fout = (decltype(p.fout))(B[rd ^ 1].p);
      p.gin = (decltype(p.gin))(G[rd].p);
      p.gout = (decltype(p.gout))(G[rd ^ 1].p);
      const int ps = convective() ? ((prev_side + par) & 1) : prev_side;
      p.prev_in = (decltype(p.prev_in))(prev[ps].as<double>());
      p.prev_out = (decltype(p.prev_out))(prev[ps ^ 1].as<double>());
      // force sums of a step go to the other buffer; block 0 of the launch
      // reduces the previous step's
      p.fpart = (decltype(p.fpart))(fpart[(fpart_side + par + 1) & 1].as<double>());
      p.fpart_prev = (decltype(p.fpart_prev))(fpart[(fpart_side + par) & 1].as<double>());
      p.commit = 1;
      bc.p[par] = p;
    }
    bc.step0 = step0;
    bc.planes = (decltype(bc.planes))planes_dev;
    bc.hist = (decltype(bc.hist))hist_rows;
    bc.first_g_skip = g_skip_pending;
    return bc;
  }

  // State bookkeeping after n steps ending at step number `last`.
  void advance_state(int n, long long last, const double* last_plane_dev) {
    const int flip = n & 1;
    side ^= flip;
    if (convective()) prev_side ^= flip;
    fpart_side ^= flip;
    last_step_no = last;
    last_g_skip = n == 1 ? g_skip_pending : 0;
    g_skip_pending = 0;
    last_plane = last_plane_dev;
  }

  // B[dst] = collide(F) with the optional LES finite-difference omega.
  void collide_field(const void* F, const void* U, int dst) {
    CollideParams cp;
    std::memset(&cp, 0, sizeof(cp));
    cp.nx = d.nx;
    cp.ny = d.ny;
    cp.nz = d.nz;
    cp.dim = Q == 19 ? 3 : 2;
    cp.xp = xp;
    cp.qxp = Q * xp;
    cp.ncell = N;
    cp.nblocks = nblocks;
    cp.F = F;
    cp.U = U;
    cp.fout = B[dst].p;
    cp.flags = flags.as<uint8_t>();
    cp.damp = damp.as<double>();
    cp.minv_d = minv_d.as<double>();
    cp.minv_f = minv_f.as<float>();
    for (int a = 0; a < 19; ++a) cp.mrt_s[a] = d.mrt_s[a];
    StepParams bp = base_params();
    cp.omega = d.omega;
    cp.trt_sm = bp.trt_sm;
    cp.trt_lambda = d.trt_lambda;
    cp.trt_variant = d.trt_variant;
    cp.base_nu = (1.0 / d.omega - 0.5) / 3.0;  // physics.py:27-29
    cp.les_cs = d.les_cs;
    cp.u0 = d.u0;
    for (int a = 0; a < NSHIFT; ++a) cp.fsh[a] = bp.fsh[a];
    cp.les = (U || les_local()) ? d.les : AERO_LES_OFF;
    ck(launch_collide(Q, d.collision, d.storage, cp, stream), "collide launch");
  }

  void recycling_prepass(int read_side, hipStream_t s = nullptr) {
    if (!s) s = stream;
    StepParams p = base_params();
    p.fin = (decltype(p.fin))(B[read_side].p);
    ck(launch_recycling_plane(d.storage, p, recyc.as<double>(), s), "recycling plane");
    ck(launch_reduce_recycling(recyc.as<double>(), nrow, d.u0, s), "recycling reduce");
  }

  // Parameters that re-derive the post-boundary state of the last completed
  // step (materialize_kernel), queued on stream s (the recycling inlet's
  // plane of that step first).
  StepParams materialize_params(hipStream_t s) {
    if (!has_state) throw std::runtime_error("no state: call aero_set_f first");
    const int old = side ^ 1;
    if (d.inlet == AERO_IN_RECYCLING) recycling_prepass(old, s);
    StepParams p = base_params();
    p.fin = (decltype(p.fin))(B[old].p);
    p.gin = (decltype(p.gin))(G[old].p);
    p.prev_in = (decltype(p.prev_in))(prev[convective() ? (prev_side ^ 1) : prev_side].as<double>());
    p.plane = (decltype(p.plane))(last_plane);
    p.phase = 0.17 * (double)last_step_no;
    p.g_bc_skip = last_g_skip;
    p.commit = 0;
    return p;
  }

  // f64 post-boundary state of the last completed step into fcache.
  void materialize_f() {
    if (fcache_valid) return;
    if (!has_state) throw std::runtime_error("no state: call aero_set_f first");
    fcache.alloc(sizeof(double) * Q * N);
    StepParams p = materialize_params(stream);
    p.out_mode = 0;
    p.fmat = (decltype(p.fmat))(fcache.as<double>());
    ck(launch_materialize(Q, d.storage, d.thermal != 0, p, stream), "materialize launch");
    fcache_valid = true;
  }

  // check_stability's statistics (diagnostics.py:32-71) of the post-boundary
  // state of the last completed step, reduced on the device into out (NCHECK
  // doubles), queued on stream s: per-block partials as the state is
  // re-derived (3D: the z mid-plane only, solver3d.py:608-615), then one
  // block sums them in block order.
  // Parameters of the stability check's pass over the last completed step
  // (materialize, out_mode 2): NCHECK block partials per block into part.
  StepParams check_params(hipStream_t s, double* part) {
    StepParams p = materialize_params(s);
    p.out_mode = 2;
    p.check_z = Q == 19 ? d.nz / 2 : -1;
    p.check_part = (decltype(p.check_part))part;
    return p;
  }

  void check_stats(hipStream_t s, double* out) {
    checkpart.alloc(sizeof(double) * NCHECK * nblocks);
    const StepParams p = check_params(s, checkpart.as<double>());
    ck(launch_materialize(Q, d.storage, d.thermal != 0, p, s), "check launch");
    ck(launch_reduce_check(checkpart.as<double>(), nblocks, out, s), "check reduce");
  }

  void materialize_g() {
    if (gcache_valid) return;
    gcache.alloc(sizeof(double) * Q * N);
    StepParams p = base_params();
    p.gin = (decltype(p.gin))(G[side].p);
    p.g_bc_skip = 0;
    ck(launch_materialize_g(Q, d.storage, p, gcache.as<double>(), stream), "materialize g");
    gcache_valid = true;
  }

  void invalidate() {
    fcache_valid = false;
    gcache_valid = false;
  }
};

// Wait for the work queued on a context's own stream.  An idle stream (the
// usual case between batched chunks) costs a status query, not a round trip
// through the device as a synchronisation does.
static void drain(aero_ctx* c) {
  const hipError_t q = hipStreamQuery(c->stream);
  if (q == hipSuccess) return;
  if (q != hipErrorNotReady) ck(q, "stream query");
  ck(hipStreamSynchronize(c->stream), "context sync");
}

bool StepRunner::Group::holds(const aero_ctx* c) const {
  return q == c->d.q && coll == c->d.collision && st == c->d.storage && th == (c->d.thermal != 0) &&
         les == (c->les_local() ? 1 : 0);
}

namespace {

// Inverse periodic stream on the host: G[i](x) = g[i](x + e_i), so that the
// pull performed by the next step returns g unchanged.
void inverse_stream(const aero_ctx* c, const double* g, std::vector<double>& out) {
  const int nx = c->d.nx, ny = c->d.ny, nz = c->d.nz;
  const int* EX = c->Q == 9 ? D2Q9::EX : D3Q19::EX;
  const int* EY = c->Q == 9 ? D2Q9::EY : D3Q19::EY;
  const int* EZ = c->Q == 9 ? D2Q9::EZ : D3Q19::EZ;
  out.resize((size_t)c->Q * c->N);
  for (int i = 0; i < c->Q; ++i) {
    const size_t base = (size_t)i * c->N;
    for (int z = 0; z < nz; ++z) {
      const int zs = ((z + EZ[i]) % nz + nz) % nz;
      for (int y = 0; y < ny; ++y) {
        const int ys = ((y + EY[i]) % ny + ny) % ny;
        for (int x = 0; x < nx; ++x) {
          const int xs = ((x + EX[i]) % nx + nx) % nx;
          out[base + ((size_t)z * ny + y) * nx + x] = g[base + ((size_t)zs * ny + ys) * nx + xs];
        }
      }
    }
  }
}

}  // namespace

static void set_geometry_impl(aero_ctx* c, const uint8_t* solid, int64_t nlinks, const int32_t* links,
                              const float* q);

// Advance every case in `cs` by nsteps steps on `stream`: per step one
// batched-kernel launch for each group of cases that share lattice,
// collision, storage and scalar transport (plus the recycling prepass of
// cases that have one).  Coefficient rows go to hists[i] (or nowhere when
// null); per-step inlet planes come from planes[i].
static void run_cases(const std::vector<aero_ctx*>& cs, int nsteps, const long long* step0s,
                      const double* const* planes, double* const* hists, hipStream_t stream, StepRunner& R) {
  const int n = (int)cs.size();
  if (n == 0 || nsteps <= 0) return;
  int CH = STEP_CHUNK;
  for (aero_ctx* c : cs) CH = std::min(CH, c->chunk_steps());
  R.reserve(n, CH);
  // groups and the packed tile list, rebuilt only when the case list changes
  std::vector<unsigned long long> key;
  for (aero_ctx* c : cs) {
    key.push_back(c->serial);
    key.push_back(c->layout_gen);  // a new wave list changes the case's block count
  }
  const unsigned tgen = g_tune_gen.load();
  if (key != R.tile_key || tgen != R.tune_gen) {
    R.drop_graphs();
    std::vector<StepRunner::Group>& groups = R.groups;
    groups.clear();
    std::vector<std::vector<int>> members;
    for (int i = 0; i < n; ++i) {
      const aero_desc& d = cs[i]->d;
      int g = 0;
      for (; g < (int)groups.size(); ++g)
        if (groups[g].holds(cs[i])) break;
      if (g == (int)groups.size()) {
        groups.push_back(StepRunner::Group{d.q, d.collision, d.storage, d.thermal != 0, cs[i]->les_local() ? 1 : 0,
                                           0, 0, TuneVal{}});
        members.emplace_back();
      }
      members[g].push_back(i);
    }
    std::vector<int2> tiles, nat;
    R.listed.assign(n, 0);
    for (size_t g = 0; g < groups.size(); ++g) {
      groups[g].toff = (int)tiles.size();
      nat.clear();
      size_t zreuse = 0;  // planes z-1, z, z+1 of every population, on all XCDs at once
      long long row_tiles = 0;  // the group's tiles with one row per wave: its grid class
      for (int i : members[g]) {
        row_tiles += cs[i]->sblocks_rows + 1;
        const aero_ctx* c = cs[i];
        if (c->d.nz > 1)
          zreuse = std::max(zreuse, (size_t)XCDS * 3 * c->Q * (size_t)c->xp * c->d.ny * c->esz);
      }
      const StepRunner::Group& G = groups[g];
      const TuneVal tv = tune_lookup(G.q, G.coll, G.st, G.th, grid_class(row_tiles));
      groups[g].tune = tv;
      for (int i : members[g]) {
        R.listed[i] = tv.rowgroups != 0 && cs[i]->nitems > 0;
        const int nb = R.listed[i] ? cs[i]->sblocks : cs[i]->sblocks_rows;
        for (int b = 0; b <= nb; ++b) nat.push_back(make_int2(i, b));
      }
      int run = zreuse > ZREUSE_BYTES ? XCD_RUN_LARGE : ((int)nat.size() + XCDS - 1) / XCDS;
      if (tv.xcd_run > 0) run = tv.xcd_run;
      if (tv.xcd_run < 0) tiles.insert(tiles.end(), nat.begin(), nat.end());
      else xcd_order(nat, std::max(run, 1), tiles);
      groups[g].ntiles = (int)tiles.size() - groups[g].toff;
    }
    R.tiles.alloc(sizeof(int2) * tiles.size());
    ck(hipMemcpyAsync(R.tiles.p, tiles.data(), sizeof(int2) * tiles.size(), hipMemcpyHostToDevice, stream),
       "upload tiles");
    ck(hipStreamSynchronize(stream), "tiles sync");
    R.tile_key = key;
    R.tune_gen = tgen;
  }
  const std::vector<StepRunner::Group>& groups = R.groups;
  const bool timed = n == 1 && cs[0]->timing;
  int pend_off[2] = {-1, -1}, pend_n[2] = {0, 0};
  auto drain = [&](int b) {
    if (pend_off[b] < 0) return;
    ck(hipEventSynchronize(R.ev[b]), "chunk sync");
    for (int i = 0; i < n; ++i)
      if (hists[i])
        std::memcpy(hists[i] + (size_t)NCOEF * pend_off[b], R.stage[b] + (size_t)i * CH * NCOEF,
                    sizeof(double) * NCOEF * pend_n[b]);
    pend_off[b] = -1;
  };
  // Population footprint of the launch (both buffers of every case): beyond
  // most of MI355X's 256 MB last-level cache a stored line is evicted before
  // the next step reads it, so the stores bypass the caches.
  size_t ws = 0;
  for (aero_ctx* c : cs) ws += 2 * c->state_bytes() * (c->d.thermal ? 2 : 1);
  const bool nt_rule = ws > NT_STORE_BYTES;
  std::vector<char> nt(n, nt_rule);  // per case, the tuning of its group may override the rule
  for (int i = 0; i < n; ++i)
    for (const StepRunner::Group& g : groups)
      if (g.holds(cs[i]) && g.tune.nt >= 0) nt[i] = g.tune.nt == 1;
  // the chunk's step launches replay from a captured graph when every
  // group's tuning asks for it (not with per-step host work in between:
  // the recycling prepass, or the timing events)
  bool use_graph = !timed && !groups.empty();
  for (const StepRunner::Group& g : groups) use_graph = use_graph && g.tune.graph == 1;
  for (aero_ctx* c : cs) use_graph = use_graph && c->d.inlet != AERO_IN_RECYCLING;
  // several kernel groups step on streams of their own (not under capture,
  // timing events or the recycling prepass, which run on the one stream)
  bool fork = !use_graph && !timed && groups.size() > 1;
  for (aero_ctx* c : cs) fork = fork && c->d.inlet != AERO_IN_RECYCLING;
  const int nside = fork ? std::min<int>((int)groups.size() - 1, StepRunner::NSIDE) : 0;
  if (nside) R.ensure_side(nside);
  bool any_hist = false;
  for (int i = 0; i < n; ++i) any_hist |= hists[i] != nullptr;
  for (int s0 = 0, ci = 0; s0 < nsteps; s0 += CH, ++ci) {
    const int nk = std::min(CH, nsteps - s0);
    const int b = ci & 1;
    drain(b);  // also frees this buffer's pinned case table and reductions
    double* rows = R.rows[b].as<double>();
    ck(hipMemsetAsync(rows, 0, sizeof(double) * NCOEF * CH * (size_t)n, stream), "zero rows");
    std::vector<const double*> last_planes(n, nullptr);
    for (int i = 0; i < n; ++i) {
      aero_ctx* c = cs[i];
      const double* pdev = nullptr;
      if (c->need_planes()) {
        const size_t plane_cnt = 3 * (size_t)c->nrow;
        c->fill_planes(b, planes[i], s0, nk, stream);
        pdev = c->planes[b].as<double>();
        last_planes[i] = pdev + plane_cnt * (nk - 1);
      }
      R.cases_host[b][i] = c->make_case(step0s[i] + s0, pdev, rows + (size_t)i * CH * NCOEF);
    }
    for (int i = 0; i < n; ++i) {
      if (nt[i]) R.cases_host[b][i].p[0].nt_store = R.cases_host[b][i].p[1].nt_store = 1;
      if (!R.listed[i]) {  // one row per wave over the whole thread grid
        for (StepParams& p : R.cases_host[b][i].p) {
          p.nitems = 0;
          p.sblocks = cs[i]->sblocks_rows;
        }
      }
    }
    ck(hipMemcpyAsync(R.cases[b].p, R.cases_host[b], sizeof(BatchCase) * n, hipMemcpyHostToDevice, stream),
       "upload cases");
    const BatchCase* cases = R.cases[b].as<BatchCase>();
    auto launch_steps = [&]() {
      if (nside) {
        ck(hipEventRecord(R.fork_ev, stream), "fork event");
        for (int j = 0; j < nside; ++j) ck(hipStreamWaitEvent(R.side[j], R.fork_ev, 0), "fork wait");
      }
      for (int k = 0; k < nk; ++k) {
        for (aero_ctx* c : cs)
          if (c->d.inlet == AERO_IN_RECYCLING) c->recycling_prepass((c->side + k) & 1, stream);
        for (size_t gi = 0; gi < groups.size(); ++gi) {
          const StepRunner::Group& g = groups[gi];
          const int j = nside ? (int)(gi % (size_t)(nside + 1)) : 0;  // 0: the runner's stream
          ck(launch_batch_step(g.q, g.coll, g.st, g.th, cases, R.tiles.as<int2>() + g.toff, g.ntiles, k,
                               j ? R.side[j - 1] : stream, timed ? cs[0]->ev_pool[2 * k] : nullptr,
                               timed ? cs[0]->ev_pool[2 * k + 1] : nullptr, g.les != 0),
             "step launch");
        }
      }
      for (int j = 0; j < nside; ++j) {
        ck(hipEventRecord(R.join_ev[j], R.side[j]), "join event");
        ck(hipStreamWaitEvent(stream, R.join_ev[j], 0), "join wait");
      }
    };
    if (use_graph) {
      StepRunner::Captured& cap = R.graphs[{b, nk}];
      if (!cap.exec || cap.cases != R.cases[b].p || cap.tiles != R.tiles.p) {
        if (cap.exec) (void)hipGraphExecDestroy(cap.exec);
        cap.exec = nullptr;
        hipGraph_t graph = nullptr;
        ck(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal), "begin capture");
        try {
          launch_steps();
        } catch (...) {
          (void)hipStreamEndCapture(stream, &graph);
          if (graph) (void)hipGraphDestroy(graph);
          throw;
        }
        ck(hipStreamEndCapture(stream, &graph), "end capture");
        const hipError_t e = hipGraphInstantiate(&cap.exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        ck(e, "instantiate graph");
        cap.cases = R.cases[b].p;
        cap.tiles = R.tiles.p;
      }
      ck(hipGraphLaunch(cap.exec, stream), "graph launch");
    } else {
      launch_steps();
    }
    // the chunk's last step is reduced here, so every chunk's rows are
    // complete when they are copied out
    int m = 0;
    for (int q : {9, 19}) {
      const int m0 = m;
      int max_terms = 0;
      for (int i = 0; i < n; ++i) {
        aero_ctx* c = cs[i];
        if (c->Q != q || c->nrec == 0) continue;
        R.reds_host[b][m++] = make_force_reduce(c->Q, c->fpart[(c->fpart_side + nk) & 1].as<double>(),
                                          c->lrec.as<LinkCell>(), c->nrec, c->d.nx, c->d.ny, c->d.center[0],
                                          c->d.center[1], c->d.center[2], c->force_denom(), c->moment_denom(),
                                          rows + (size_t)i * CH * NCOEF + (size_t)NCOEF * (nk - 1));
        c->fterms.alloc(sizeof(double) * NFORCE * (size_t)c->nrec);
        R.reds_host[b][m - 1].terms = (decltype(R.reds_host[b][m - 1].terms))(c->fterms.as<double>());
        max_terms = std::max(max_terms, c->nrec);
      }
      if (m > m0) {
        ck(hipMemcpyAsync(R.reds[b].as<ForceReduce>() + m0, R.reds_host[b] + m0, sizeof(ForceReduce) * (m - m0),
                          hipMemcpyHostToDevice, stream),
           "upload reductions");
        ck(launch_reduce_forces_many(q, R.reds[b].as<ForceReduce>() + m0, m - m0, max_terms, stream),
           "force reduction");
      }
    }
    int nr = 0;
    for (int i = 0; i < n; ++i)
      if (cs[i]->ring_cap > 0) R.rapp_host[b][nr++] = cs[i]->ring_append(rows + (size_t)i * CH * NCOEF, nk);
    if (nr > 0) {
      ck(hipMemcpyAsync(R.rapp[b].p, R.rapp_host[b], sizeof(RingAppend) * nr, hipMemcpyHostToDevice, stream),
         "upload ring appends");
      ck(launch_ring_append(R.rapp[b].as<RingAppend>(), nr, nk, stream), "ring append");
      for (aero_ctx* c : cs)
        if (c->ring_cap > 0) c->ring_count += nk;
    }
    if (any_hist)  // the chunk's rows of every case, stored by a kernel into the pinned staging buffer
      ck(launch_rows_to_host(rows, R.stage[b], n, CH, nk, stream), "download rows");
    ck(hipEventRecord(R.ev[b], stream), "chunk event");
    pend_off[b] = s0;
    pend_n[b] = nk;
    for (int i = 0; i < n; ++i) cs[i]->advance_state(nk, step0s[i] + s0 + nk, last_planes[i]);
    if (timed) {
      aero_ctx* c = cs[0];
      ck(hipEventSynchronize(c->ev_pool[2 * nk - 1]), "timing sync");
      for (int k = 0; k < nk; ++k) {
        float ms = 0.f;
        ck(hipEventElapsedTime(&ms, c->ev_pool[2 * k], c->ev_pool[2 * k + 1]), "event time");
        c->timed_ms += ms;
      }
      c->timed_launches += nk;
    }
  }
  bool tail = false;  // work queued after the last chunk event
  for (aero_ctx* c : cs) {
    // keep the last step's inlet plane for re-deriving its state later
    if (c->need_planes()) {
      ck(hipMemcpyAsync(c->last_plane_buf.p, c->last_plane, sizeof(double) * 3 * (size_t)c->nrow,
                        hipMemcpyDeviceToDevice, stream),
         "keep last plane");
      c->last_plane = c->last_plane_buf.as<double>();
      tail = true;
    }
    c->invalidate();
  }
  drain(0);
  drain(1);
  // the chunk events close everything else on the stream
  if (tail) ck(hipStreamSynchronize(stream), "step sync");
}

extern "C" {

const char* aero_last_error(void) { return g_err.c_str(); }

int aero_version(void) { return AERO_ABI_VERSION; }

int aero_device_count(int32_t* n) {
  return guarded([&] {
    int k = 0;
    ck(hipGetDeviceCount(&k), "hipGetDeviceCount");
    *n = k;
  });
}

int aero_create(const aero_desc* desc, aero_ctx** out) {
  *out = nullptr;
  aero_ctx* c = new aero_ctx();
  int rc = guarded([&] {
    c->d = *desc;
    const aero_desc& d = c->d;
    if (d.q != 9 && d.q != 19) throw std::invalid_argument("q must be 9 or 19");
    if (d.q == 9 && d.nz != 1) throw std::invalid_argument("D2Q9 requires nz == 1");
    if (d.nx < 3 || d.ny < 1 || d.nz < 1) throw std::invalid_argument("grid must have nx >= 3, ny >= 1, nz >= 1");
    if (d.collision < 0 || d.collision > 2) throw std::invalid_argument("unknown collision");
    if (d.les < AERO_LES_OFF || d.les > AERO_LES_SMAGORINSKY_LOCAL) throw std::invalid_argument("unknown LES model");
    if (d.storage < 0 || d.storage > 3) throw std::invalid_argument("unknown storage");
    if (d.q == 9 && (d.inlet == AERO_IN_SEM || d.inlet == AERO_IN_RECYCLING))
      throw std::invalid_argument("SEM and recycling inlets are 3D only");
    if (d.q == 19 && (d.inlet == AERO_IN_PRESSURE || d.outlet == AERO_OUT_PRESSURE))
      throw std::invalid_argument("pressure inlet/outlet are 2D only");
    if (!(d.omega > 0.0)) throw std::invalid_argument("omega must be positive");
    const long long N = (long long)d.nx * d.ny * d.nz;
    if (N * d.q >= (1ll << 31)) throw std::invalid_argument("grid too large for 32-bit cell indexing");
    const long long nthread = (long long)padded_row(d.nx) * d.ny * d.nz;
    if (nthread >= (1ll << 31)) throw std::invalid_argument("grid too large for 32-bit step thread indexing");
    c->Q = d.q;
    c->N = N;
    c->nrow = d.nz * d.ny;
    c->xp = padded_row(d.nx);
    c->nblocks = (int)((N + BLOCK - 1) / BLOCK);
    c->sblocks_rows = (int)((nthread + BLOCK - 1) / BLOCK);
    c->sblocks = c->sblocks_rows;  // grouped kernels: set with the flags
    c->esz = storage_bytes(d.storage);
    c->csz = d.storage == ST_F64 ? 8 : 4;
    ck(hipGetDevice(&c->device), "hipGetDevice");
    c->stream = take_stream(c->device);
    const size_t fbytes = c->state_bytes();
    c->B[0].alloc(fbytes);
    c->B[1].alloc(fbytes);
    if (d.thermal) {
      c->G[0].alloc(fbytes);
      c->G[1].alloc(fbytes);
    }
    const size_t pbytes = sizeof(double) * d.q * (size_t)c->nrow;
    c->prev[0].alloc(pbytes);
    c->prev[1].alloc(pbytes);
    ck(hipMemset(c->prev[0].p, 0, pbytes), "memset");
    ck(hipMemset(c->prev[1].p, 0, pbytes), "memset");
    // sponge sigma(x), sponge.py:8-21 (same arithmetic as the reference)
    std::vector<double> sig(d.nx, 0.0);
    if (d.sponge_thickness > 0 && d.sponge_strength > 0.0) {
      const int start = std::max(d.nx - d.sponge_thickness, 0);
      for (int x = start; x < d.nx; ++x) {
        const double frac = (double)(x - start + 1) / (double)std::max(d.sponge_thickness, 1);
        sig[x] = d.sponge_strength * frac * frac;
      }
    }
    c->sigma.alloc(sizeof(double) * d.nx);
    ck(hipMemcpy(c->sigma.p, sig.data(), sizeof(double) * d.nx, hipMemcpyHostToDevice), "copy sigma");
    c->sponge_x0 = d.nx;
    if (d.sponge_thickness > 0)
      for (int x = d.nx - 1; x >= 0; --x)
        if (sig[x] > 0.0) c->sponge_x0 = x;
    c->minv_d.alloc(sizeof(double) * 361);
    c->minv_f.alloc(sizeof(float) * 361);
    std::vector<float> mf(361);
    for (int k = 0; k < 361; ++k) mf[k] = (float)d.mrt_minv[k];
    ck(hipMemcpy(c->minv_d.p, d.mrt_minv, sizeof(double) * 361, hipMemcpyHostToDevice), "copy minv");
    ck(hipMemcpy(c->minv_f.p, mf.data(), sizeof(float) * 361, hipMemcpyHostToDevice), "copy minv");
    c->recyc.alloc(sizeof(double) * (4 * (size_t)c->nrow + 1));
    if (c->les_split()) {
      for (int b = 0; b < 2; ++b) {
        c->hist_dev[b].alloc(sizeof(double) * NCOEF * STEP_CHUNK);
        ck(hipHostMalloc((void**)&c->hist_stage[b], sizeof(double) * NCOEF * STEP_CHUNK, hipHostMallocDefault),
           "hipHostMalloc");
        ck(hipEventCreateWithFlags(&c->chunk_ev[b], hipEventDisableTiming), "hipEventCreate");
      }
    }
    if (d.inlet == AERO_IN_FIELD || d.inlet == AERO_IN_SEM) {
      for (int b = 0; b < 2; ++b) c->planes[b].alloc(sizeof(double) * 3 * (size_t)c->nrow * c->chunk_steps());
      c->last_plane_buf.alloc(sizeof(double) * 3 * (size_t)c->nrow);
    }
    if (c->les_split()) {
      c->fsplit.alloc(c->csz * d.q * (size_t)N);
      c->usplit.alloc(c->csz * 3 * (size_t)N);
    }
    // default geometry: no solid, no links
    std::vector<uint8_t> none((size_t)N, 0);
    set_geometry_impl(c, none.data(), 0, nullptr, nullptr);
  });
  if (rc) {
    aero_destroy(c);
    return rc;
  }
  *out = c;
  return 0;
}

int aero_destroy(aero_ctx* c) {
  if (!c) return 0;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  sem_destroy(c->sem);
  c->sem = nullptr;
  DevBuf* bufs[] = {&c->B[0], &c->B[1], &c->G[0], &c->G[1], &c->prev[0], &c->prev[1], &c->flags, &c->wave_any,
                    &c->witems, &c->lrec, &c->rec_of, &c->qdir, &c->fpart[0], &c->fpart[1], &c->sigma, &c->minv_d, &c->minv_f,
                    &c->damp, &c->les_c2d, &c->les_c2f, &c->halo_buf, &c->fterms, &c->planes[0], &c->planes[1], &c->recyc, &c->hist_dev[0], &c->hist_dev[1], &c->last_plane_buf, &c->fsplit, &c->usplit, &c->fcache,
                    &c->gcache, &c->tmp, &c->checkpart, &c->checkout, &c->ring, &c->ring_rec};
  for (DevBuf* b : bufs) b->release();
  c->runner.release();
  for (hipEvent_t e : c->ev_pool) (void)hipEventDestroy(e);
  for (int b = 0; b < 2; ++b) {
    if (c->hist_stage[b]) (void)hipHostFree(c->hist_stage[b]);
    if (c->chunk_ev[b]) (void)hipEventDestroy(c->chunk_ev[b]);
  }
  return_stream(c->device, c->stream);
  delete c;
  return 0;
}

int aero_mem_info(int64_t* free_bytes, int64_t* total_bytes) {
  return guarded([&] {
    size_t f = 0, t = 0;
    ck(hipMemGetInfo(&f, &t), "hipMemGetInfo");
    *free_bytes = (int64_t)f;
    *total_bytes = (int64_t)t;
  });
}

// ------------------------------------------------------------------ batch
struct aero_batch {
  int device = 0;
  hipStream_t stream = nullptr;
  StepRunner runner;
  DevBuf checkout;  // aero_batch_check: NCHECK doubles per case
  DevBuf checkpart;  // aero_batch_check: block partials of every case
  DevBuf checkpar;   // aero_batch_check: the cases' parameters, then the tile list
  std::vector<std::pair<const aero_ctx*, int>> check_sig;  // the cases of the tile list held in checkpar
};

int aero_batch_create(aero_batch** out) {
  *out = nullptr;
  aero_batch* b = new aero_batch();
  const int rc = guarded([&] {
    ck(hipGetDevice(&b->device), "hipGetDevice");
    ck(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking), "hipStreamCreate");
  });
  if (rc) {
    aero_batch_destroy(b);
    return rc;
  }
  *out = b;
  return 0;
}

int aero_batch_destroy(aero_batch* b) {
  if (!b) return 0;
  (void)hipSetDevice(b->device);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  b->runner.release();
  b->checkout.release();
  b->checkpart.release();
  b->checkpar.release();
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return 0;
}

int aero_batch_step(aero_batch* b, int32_t n, aero_ctx* const* ctxs, int32_t nsteps, const int64_t* step0s,
                    const double* const* planes, double* const* hists) {
  return guarded([&] {
    if (n <= 0 || nsteps <= 0) return;
    ck(hipSetDevice(b->device), "hipSetDevice");
    std::vector<aero_ctx*> cs(ctxs, ctxs + n);
    std::vector<long long> s0(n);
    std::vector<const double*> pl(n, nullptr);
    std::vector<double*> hs(n, nullptr);
    for (int i = 0; i < n; ++i) {
      aero_ctx* c = cs[i];
      if (!c) throw std::invalid_argument("null context in batch");
      if (c->device != b->device) throw std::invalid_argument("every context of a batch must be on its device");
      if (!c->has_state) throw std::runtime_error("no state: call aero_set_f first");
      if (c->les_split()) throw std::invalid_argument("LES contexts step on their own (aero_step)");
      for (int j = 0; j < i; ++j)
        if (cs[j] == c) throw std::invalid_argument("a context appears twice in a batch");
      if (c->need_planes() && !(planes && planes[i])) throw std::invalid_argument("this inlet needs per-step planes");
      s0[i] = step0s ? step0s[i] : 0;
      pl[i] = planes ? planes[i] : nullptr;
      hs[i] = hists ? hists[i] : nullptr;
      drain(c);  // work queued on the context's own stream
    }
    run_cases(cs, nsteps, s0.data(), pl.data(), hs.data(), b->stream, b->runner);
  });
}

}  // extern "C"

// Upload the cell flags and their per-wave summary: bit w is set when any of
// the 64 cells [64 w, 64 w + 64) is flagged, so a step wave can tell from one
// scalar load that every cell it covers is plain fluid.
static void upload_flags(aero_ctx* c, const std::vector<uint8_t>& fl) {
  ck(hipMemcpy(c->flags.p, fl.data(), (size_t)c->N, hipMemcpyHostToDevice), "copy flags");
  // waves of the step's thread grid: cell (row, x) is step thread row * xp + x
  const size_t nwave = ((size_t)c->nrow * c->xp + 63) / 64;
  std::vector<uint32_t> words((nwave + 31) / 32, 0u);
  const int nx = c->d.nx;
  for (size_t k = 0; k < (size_t)c->N; ++k)
    if (fl[k]) {
      const size_t w = ((k / nx) * c->xp + k % nx) >> 6;
      words[w >> 5] |= 1u << (w & 31);
    }
  c->wave_any.alloc(sizeof(uint32_t) * words.size());
  ck(hipMemcpy(c->wave_any.p, words.data(), sizeof(uint32_t) * words.size(), hipMemcpyHostToDevice),
     "copy wave flags");
  // grouped kernels: the wave list, row groups of RPW_GROUP rows by 64-column segments;
  // a group item where no cell of the group (nor the outlet's x = nx - 2 source) is flagged
  const aero_desc& d = c->d;
  if (rows_per_wave(d.q, d.collision, d.storage, d.thermal != 0) > 1) {
    const int RG = RPW_GROUP, segs = c->xp / 64;
    std::vector<uint32_t> items;
    items.reserve(nwave);
    for (int r0 = 0; r0 < c->nrow; r0 += RG) {
      const bool rows_ok = r0 + RG <= c->nrow;
      for (int sgm = 0; sgm < segs; ++sgm) {
        const int x1 = std::min(sgm * 64 + 63, nx - 1);
        // a convective outlet row keeps per-row state: its waves step one row each
        bool bulk = rows_ok && !(x1 == nx - 1 && d.outlet == AERO_OUT_CONVECTIVE);
        for (int r = r0; bulk && r < r0 + RG; ++r) {
          const size_t w = (size_t)r * segs + sgm;
          bulk = !((words[w >> 5] >> (w & 31)) & 1u);
          // the outlet lane pulls x = nx - 2, in the previous wave when x = nx - 1 is lane 0
          if (bulk && sgm * 64 == nx - 1) bulk = !((words[(w - 1) >> 5] >> ((w - 1) & 31)) & 1u);
        }
        if (bulk) {
          items.push_back(WITEM_GROUP | (uint32_t)((size_t)r0 * segs + sgm));
        } else {
          for (int r = r0; r < std::min(r0 + RG, c->nrow); ++r) items.push_back((uint32_t)((size_t)r * segs + sgm));
        }
      }
    }
    c->nitems = (int)items.size();
    c->witems.alloc(sizeof(uint32_t) * std::max<size_t>(1, items.size()));
    ck(hipMemcpy(c->witems.p, items.data(), sizeof(uint32_t) * items.size(), hipMemcpyHostToDevice),
       "copy wave list");
    c->sblocks = (int)(((long long)c->nitems * 64 + BLOCK - 1) / BLOCK);
  }
  ++c->layout_gen;
}

// Cell flags and link-cell records for one geometry.
static void set_geometry_impl(aero_ctx* c, const uint8_t* solid, int64_t nlinks, const int32_t* links,
                              const float* q) {
    c->set_device();
    const int nx = c->d.nx, ny = c->d.ny, nz = c->d.nz;
    c->flags.alloc((size_t)c->N);
    std::vector<uint8_t> fl((size_t)c->N);
    for (long long k = 0; k < c->N; ++k) fl[k] = solid[k] ? FLAG_SOLID : 0;
    // group links by cell: direction mask + q per direction
    const int Q1 = c->Q - 1;
    std::map<int, std::pair<uint32_t, std::vector<float>>> cells;
    for (int64_t k = 0; k < nlinks; ++k) {
      const int i = links[4 * k], z = links[4 * k + 1], y = links[4 * k + 2], x = links[4 * k + 3];
      if (i < 1 || i >= c->Q || x < 0 || x >= nx || y < 0 || y >= ny || z < 0 || z >= nz)
        throw std::invalid_argument("link out of range");
      const int cell = (z * ny + y) * nx + x;
      if (solid[cell]) throw std::invalid_argument("link origin must be a fluid cell");
      auto& e = cells[cell];
      if (e.first & (1u << i)) throw std::invalid_argument("duplicate link");
      if (e.second.empty()) e.second.assign(Q1, 0.0f);
      e.first |= 1u << i;
      e.second[i - 1] = q ? q[k] : 0.5f;
    }
    std::vector<LinkCell> recs;
    std::vector<float> qv;
    std::vector<int> rof((size_t)c->N, -1);
    recs.reserve(cells.size());
    qv.reserve(cells.size() * Q1);
    for (auto& kv : cells) {
      rof[kv.first] = (int)recs.size();
      recs.push_back(LinkCell{kv.first, kv.second.first});
      qv.insert(qv.end(), kv.second.second.begin(), kv.second.second.end());
      fl[kv.first] |= FLAG_LINKS;
    }
    c->nrec = (int)recs.size();
    // new flags carry no IBM band: immersed-boundary mode ends until
    // aero_set_ibm marks the band again
    c->ibm_on = 0;
    c->ibm_coeff = 0.0;
    c->rec_cells.clear();
    c->rec_masks.clear();
    for (auto& r : recs) {
      c->rec_cells.push_back(r.cell);
      c->rec_masks.push_back(r.mask);
    }
    upload_flags(c, fl);
    c->lrec.alloc(sizeof(LinkCell) * std::max<size_t>(1, recs.size()));
    c->qdir.alloc(sizeof(float) * std::max<size_t>(1, qv.size()));
    c->rec_of.alloc(sizeof(int) * (size_t)c->N);
    for (int k = 0; k < 2; ++k) c->fpart[k].alloc(sizeof(double) * NPART * std::max<size_t>(1, recs.size()));
    c->fpart_side = 0;
    if (!recs.empty()) {
      ck(hipMemcpy(c->lrec.p, recs.data(), sizeof(LinkCell) * recs.size(), hipMemcpyHostToDevice), "copy recs");
      ck(hipMemcpy(c->qdir.p, qv.data(), sizeof(float) * qv.size(), hipMemcpyHostToDevice), "copy q");
    }
    ck(hipMemcpy(c->rec_of.p, rof.data(), sizeof(int) * (size_t)c->N, hipMemcpyHostToDevice), "copy record index");
    for (int k = 0; k < 2; ++k) ck(hipMemset(c->fpart[k].p, 0, c->fpart[k].bytes), "memset fpart");
    c->invalidate();
}

extern "C" {

int aero_set_geometry(aero_ctx* c, const uint8_t* solid, int64_t nlinks, const int32_t* links, const float* q) {
  return guarded([&] { set_geometry_impl(c, solid, nlinks, links, q); });
}

int aero_set_damping(aero_ctx* c, const double* damp) {
  return guarded([&] {
    c->set_device();
    if (!damp) {
      c->damp.release();
      c->les_c2d.release();
      c->les_c2f.release();
      return;
    }
    c->damp.alloc(sizeof(double) * c->N);
    ck(hipMemcpy(c->damp.p, damp, sizeof(double) * c->N, hipMemcpyHostToDevice), "copy damping");
    if (c->d.les == AERO_LES_SMAGORINSKY_LOCAL) {
      // the local step's C^2 per cell, in both arithmetic types
      std::vector<double> c2d((size_t)c->N);
      std::vector<float> c2f((size_t)c->N);
      for (size_t i = 0; i < c2d.size(); ++i) {
        const double cs = c->d.les_cs * damp[i];
        c2d[i] = cs * cs;
        c2f[i] = (float)c2d[i];
      }
      c->les_c2d.alloc(sizeof(double) * c->N);
      c->les_c2f.alloc(sizeof(float) * c->N);
      ck(hipMemcpy(c->les_c2d.p, c2d.data(), sizeof(double) * c->N, hipMemcpyHostToDevice), "copy C^2");
      ck(hipMemcpy(c->les_c2f.p, c2f.data(), sizeof(float) * c->N, hipMemcpyHostToDevice), "copy C^2");
    }
  });
}

// The state from the FP64 field in fcache (the reference's self.f).
static void state_from_fcache(aero_ctx* c) {
  const size_t cnt = (size_t)c->Q * c->N;
  {
    // F in compute precision, then B[side] = collide(F) (with LES omega from F)
    const void* F = c->fcache.p;
    const bool own = c->d.storage != ST_F64 || c->ibm_on;  // fcache itself stays the unforced self.f
    if (own) {
      c->tmp.alloc((c->d.storage == ST_F64 ? sizeof(double) : sizeof(float)) * cnt);
      ck(launch_compute_from_f64(c->d.storage, (long long)cnt, c->fcache.as<double>(), c->tmp.p, c->stream),
         "convert f");
      F = c->tmp.p;
    }
    if (c->ibm_on)  // the first step's forcing acts on this state (solver.py:280-285)
      ck(launch_ibm_field(c->Q, c->d.storage, c->N, c->tmp.p, c->flags.as<uint8_t>(), c->ibm_coeff, c->stream),
         "ibm forcing");
    const void* U = nullptr;
    if (c->les_split()) {
      ck(launch_macro_from_compute(c->d.storage, c->Q, c->N, F, c->usplit.p, c->stream), "macro");
      U = c->usplit.p;
    }
    c->collide_field(F, U, c->side);
    ck(hipStreamSynchronize(c->stream), "sync");
    if (own) c->tmp.release();
    c->fcache_valid = true;
    c->has_state = true;  // state defined; fcache holds it until the next step
  }
}

int aero_set_ibm(aero_ctx* c, const double* phi) {
  return guarded([&] {
    c->set_device();
    if (!c->flags.p) throw std::runtime_error("aero_set_ibm: call aero_set_geometry first");
    std::vector<uint8_t> fl((size_t)c->N);
    ck(hipMemcpy(fl.data(), c->flags.p, (size_t)c->N, hipMemcpyDeviceToHost), "read flags");
    for (long long k = 0; k < c->N; ++k) {
      fl[k] &= (uint8_t)~FLAG_IBM;
      // ibm_guo.py:133: skipped where solid, phi <= 0 or phi > 1.5 (a NaN phi is not skipped)
      if (phi && !(fl[k] & FLAG_SOLID) && !(phi[k] <= 0.0) && !(phi[k] > 1.5)) fl[k] |= FLAG_IBM;
    }
    upload_flags(c, fl);
    c->ibm_on = 1;
    const double tau = 1.0 / c->d.omega;  // solver.py:184
    c->ibm_coeff = 1.0 - 0.5 / tau;
  });
}

int aero_set_f(aero_ctx* c, const double* f) {
  return guarded([&] {
    c->set_device();
    const size_t cnt = (size_t)c->Q * c->N;
    c->fcache.alloc(sizeof(double) * cnt);
    ck(hipMemcpyAsync(c->fcache.p, f, sizeof(double) * cnt, hipMemcpyHostToDevice, c->stream), "upload f");
    state_from_fcache(c);
  });
}

// Device buffer freed when it goes out of scope (hipFree waits for its use).
struct ScopedBuf : DevBuf {
  ScopedBuf() = default;
  ScopedBuf(const ScopedBuf&) = delete;
  ScopedBuf& operator=(const ScopedBuf&) = delete;
  ~ScopedBuf() { release(); }
};

struct ScopedStream {
  hipStream_t s = nullptr;
  ScopedStream() { ck(hipStreamCreate(&s), "hipStreamCreate"); }
  ~ScopedStream() { (void)hipStreamDestroy(s); }
};

int aero_set_sem(aero_ctx* c, int32_t n_eddies, double sigma, double Tu) {
  return guarded([&] {
    if (c->d.q != 19 || c->d.inlet != AERO_IN_SEM) throw std::invalid_argument("aero_set_sem: not a 3D SEM inlet");
    if (n_eddies < 0) throw std::invalid_argument("aero_set_sem: negative eddy count");
    c->set_device();
    ck(hipStreamSynchronize(c->stream), "SEM sync");
    sem_destroy(c->sem);
    c->sem = nullptr;
    c->sem = sem_create(c->d.ny, c->d.nz, n_eddies, sigma, Tu, c->d.u0, c->chunk_steps());
    c->sem_n = n_eddies;
  });
}

int aero_sem_planes(aero_ctx* c, int32_t n, const double* eddies, double* planes_out) {
  return guarded([&] {
    if (!c->sem) throw std::invalid_argument("aero_sem_planes: call aero_set_sem first");
    c->set_device();
    const size_t plane_cnt = 3 * (size_t)c->nrow;
    const int CH = c->chunk_steps();
    ScopedBuf out;
    out.alloc(sizeof(double) * plane_cnt * (size_t)std::min<int>(std::max<int>(n, 1), CH));
    for (int s0 = 0; s0 < n; s0 += CH) {
      const int k = std::min(CH, n - s0);
      sem_planes(c->sem, eddies + (size_t)6 * c->sem_n * s0, k, out.as<double>(), c->stream);
      ck(hipMemcpyAsync(planes_out + plane_cnt * s0, out.p, sizeof(double) * plane_cnt * k, hipMemcpyDeviceToHost,
                        c->stream),
         "download SEM planes");
      ck(hipStreamSynchronize(c->stream), "SEM planes sync");
    }
  });
}

int aero_mesh_inside(const double* tris, int64_t ntri, const double* ray, const double* points, int64_t npts,
                     int32_t nx, int32_t ny, int32_t nz, uint8_t* inside) {
  return guarded([&] {
    if (ntri < 0 || (ntri > 0 && !tris) || !ray || !inside) throw std::runtime_error("aero_mesh_inside: bad arguments");
    const bool grid = points == nullptr;
    if (grid && (nx <= 0 || ny <= 0 || nz <= 0)) throw std::runtime_error("aero_mesh_inside: bad grid");
    if (!grid && npts < 0) throw std::runtime_error("aero_mesh_inside: bad point count");
    const long long n = grid ? (long long)nx * ny * nz : (long long)npts;
    if (n == 0) return;
    ScopedStream st;
    ScopedBuf dt, scr, dp, dout;
    if (ntri > 0) {
      dt.alloc(sizeof(double) * 9 * (size_t)ntri);
      ck(hipMemcpyAsync(dt.p, tris, sizeof(double) * 9 * (size_t)ntri, hipMemcpyHostToDevice, st.s), "upload mesh");
    }
    scr.alloc(mesh_inside_scratch_bytes(ntri));
    if (!grid) {
      dp.alloc(sizeof(double) * 3 * (size_t)n);
      ck(hipMemcpyAsync(dp.p, points, sizeof(double) * 3 * (size_t)n, hipMemcpyHostToDevice, st.s), "upload points");
    }
    dout.alloc((size_t)n);
    ck(launch_mesh_inside(dt.as<double>(), (long long)ntri, scr.p, dp.as<double>(), n, grid ? nx : 1,
                          grid ? ny : 1, ray, dout.as<uint8_t>(), st.s),
       "mesh inside");
    ck(hipMemcpyAsync(inside, dout.p, (size_t)n, hipMemcpyDeviceToHost, st.s), "download inside");
    ck(hipStreamSynchronize(st.s), "sync");
  });
}

int aero_mesh_sqdist(const double* tris, int64_t ntri, const int64_t* cells, int64_t ncells, int32_t nx, int32_t ny,
                     int32_t nz, double* sqdist) {
  return guarded([&] {
    if (ntri < 0 || (ntri > 0 && !tris) || ncells < 0 || (ncells > 0 && (!cells || !sqdist)))
      throw std::runtime_error("aero_mesh_sqdist: bad arguments");
    if (nx <= 0 || ny <= 0 || nz <= 0) throw std::runtime_error("aero_mesh_sqdist: bad grid");
    const long long N = (long long)nx * ny * nz;
    for (int64_t k = 0; k < ncells; ++k)
      if (cells[k] < 0 || cells[k] >= N) throw std::runtime_error("aero_mesh_sqdist: cell index outside the grid");
    if (ncells == 0) return;
    ScopedStream st;
    ScopedBuf dt, dc, dout;
    dt.alloc(sizeof(double) * 9 * (size_t)(ntri > 0 ? ntri : 1));
    if (ntri > 0)
      ck(hipMemcpyAsync(dt.p, tris, sizeof(double) * 9 * (size_t)ntri, hipMemcpyHostToDevice, st.s), "upload mesh");
    dc.alloc(sizeof(long long) * (size_t)ncells);
    ck(hipMemcpyAsync(dc.p, cells, sizeof(long long) * (size_t)ncells, hipMemcpyHostToDevice, st.s), "upload cells");
    dout.alloc(sizeof(double) * (size_t)ncells);
    ck(launch_mesh_sqdist(dt.as<double>(), (long long)ntri, dc.as<long long>(), (long long)ncells, nx, ny,
                          dout.as<double>(), st.s),
       "mesh distance");
    ck(hipMemcpyAsync(sqdist, dout.p, sizeof(double) * (size_t)ncells, hipMemcpyDeviceToHost, st.s), "download");
    ck(hipStreamSynchronize(st.s), "sync");
  });
}

int aero_set_f_equilibrium(aero_ctx* c, const double* rho, const double* ux, const double* uy, const double* uz) {
  return guarded([&] {
    c->set_device();
    const size_t N = (size_t)c->N;
    const double* src[4] = {rho, ux, uy, uz};
    int nf = 0;
    for (const double* q : src) nf += q != nullptr;
    if (c->Q == 9 && uz) throw std::runtime_error("aero_set_f_equilibrium: uz given for a 2D context");
    c->fcache.alloc(sizeof(double) * (size_t)c->Q * N);
    DevBuf fields;
    struct Release {
      DevBuf& b;
      ~Release() { b.release(); }  // hipFree waits for the stream's use of it
    } release_fields{fields};
    const double* dev[4] = {nullptr, nullptr, nullptr, nullptr};
    if (nf) {
      fields.alloc(sizeof(double) * N * nf);
      int k = 0;
      for (int a = 0; a < 4; ++a)
        if (src[a]) {
          double* d = fields.as<double>() + N * k++;
          ck(hipMemcpyAsync(d, src[a], sizeof(double) * N, hipMemcpyHostToDevice, c->stream), "upload field");
          dev[a] = d;
        }
    }
    ck(launch_feq_fill(c->Q, (long long)N, c->fcache.as<double>(), dev[0], c->d.rho0, dev[1], c->d.u0, dev[2],
                       dev[3], c->stream),
       "equilibrium");
    state_from_fcache(c);
  });
}

int aero_get_f(aero_ctx* c, double* f) {
  return guarded([&] {
    c->set_device();
    c->materialize_f();
    ck(hipMemcpyAsync(f, c->fcache.p, sizeof(double) * c->Q * c->N, hipMemcpyDeviceToHost, c->stream), "download f");
    ck(hipStreamSynchronize(c->stream), "sync");
  });
}

int aero_set_g(aero_ctx* c, const double* g) {
  return guarded([&] {
    if (!c->d.thermal) throw std::runtime_error("scalar transport is off for this context");
    c->set_device();
    std::vector<double> inv;
    inverse_stream(c, g, inv);
    const size_t cnt = (size_t)c->Q * c->N;
    c->gcache.alloc(sizeof(double) * cnt);
    ck(hipMemcpy(c->gcache.p, g, sizeof(double) * cnt, hipMemcpyHostToDevice), "upload g");
    c->tmp.alloc(sizeof(double) * cnt);
    ck(hipMemcpy(c->tmp.p, inv.data(), sizeof(double) * cnt, hipMemcpyHostToDevice), "upload g");
    const double shift = (c->d.storage == ST_F16S || c->d.storage == ST_BF16S) ? c->d.T_ref : 0.0;
    ck(launch_store_from_f64(c->d.storage, c->Q, c->N, c->d.nx, c->xp, c->tmp.as<double>(), c->G[c->side].p, shift,
                             c->stream),
       "store g");
    ck(hipStreamSynchronize(c->stream), "sync");
    c->tmp.release();
    c->gcache_valid = true;
    c->g_skip_pending = 1;  // the given field may not carry the wall values yet
  });
}

int aero_get_g(aero_ctx* c, double* g) {
  return guarded([&] {
    if (!c->d.thermal) throw std::runtime_error("scalar transport is off for this context");
    c->set_device();
    c->materialize_g();
    ck(hipMemcpyAsync(g, c->gcache.p, sizeof(double) * c->Q * c->N, hipMemcpyDeviceToHost, c->stream), "download g");
    ck(hipStreamSynchronize(c->stream), "sync");
  });
}

int aero_set_outlet_prev(aero_ctx* c, const double* prev) {
  return guarded([&] {
    c->set_device();
    ck(hipMemcpy(c->prev[c->prev_side].p, prev, sizeof(double) * c->Q * c->nrow, hipMemcpyHostToDevice),
       "upload prev");
  });
}

int aero_get_outlet_prev(aero_ctx* c, double* prev) {
  return guarded([&] {
    c->set_device();
    ck(hipStreamSynchronize(c->stream), "sync");
    ck(hipMemcpy(prev, c->prev[c->prev_side].p, sizeof(double) * c->Q * c->nrow, hipMemcpyDeviceToHost),
       "download prev");
  });
}

int aero_step(aero_ctx* c, int32_t nsteps, int64_t step0, const double* planes, double* hist) {
  return guarded([&] {
    if (nsteps <= 0) return;
    if (!c->has_state) throw std::runtime_error("no state: call aero_set_f first");
    c->set_device();
    const bool need_planes = c->need_planes();
    if (need_planes && !planes) throw std::invalid_argument("this inlet needs per-step planes");
    if (!c->les_split()) {
      // the batched kernel with a batch of one: the same code path as in a batch
      const std::vector<aero_ctx*> cs{c};
      const long long s0 = step0;
      run_cases(cs, nsteps, &s0, &planes, &hist, c->stream, c->runner);
      return;
    }
    // LES: split step (post-boundary state + velocity), finite-difference
    // relaxation field, then the collision, every step
    const size_t plane_cnt = 3 * (size_t)c->nrow;
    const double fd = c->force_denom(), md = c->moment_denom();
    const int CH = c->chunk_steps();
    // rows of chunk b wait in hist_stage[b] until that buffer is reused
    int pend_off[2] = {-1, -1}, pend_n[2] = {0, 0};
    auto drain = [&](int b) {
      if (pend_off[b] < 0) return;
      ck(hipEventSynchronize(c->chunk_ev[b]), "chunk sync");
      if (hist) std::memcpy(hist + (size_t)NCOEF * pend_off[b], c->hist_stage[b], sizeof(double) * NCOEF * pend_n[b]);
      pend_off[b] = -1;
    };
    for (int s0 = 0, ci = 0; s0 < nsteps; s0 += CH, ++ci) {
      const int n = std::min(CH, nsteps - s0);
      const int b = ci & 1;
      drain(b);
      double* rows = c->hist_dev[b].as<double>();
      if (need_planes) c->fill_planes(b, planes, s0, n, c->stream);
      for (int k = 0; k < n; ++k) {
        const long long step_no = step0 + s0 + k + 1;
        const int rd = c->side, wr = c->side ^ 1;
        if (c->d.inlet == AERO_IN_RECYCLING) c->recycling_prepass(rd);
        StepParams p = c->base_params();
        p.fin = (decltype(p.fin))(c->B[rd].p);
        p.fout = (decltype(p.fout))(c->B[wr].p);
        p.gin = (decltype(p.gin))(c->G[rd].p);
        p.gout = (decltype(p.gout))(c->G[wr].p);
        p.prev_in = (decltype(p.prev_in))(c->prev[c->prev_side].as<double>());
        p.prev_out = (decltype(p.prev_out))(c->prev[c->prev_side ^ 1].as<double>());
        p.plane = (decltype(p.plane))(need_planes ? c->planes[b].as<double>() + plane_cnt * k : nullptr);
        p.phase = 0.17 * (double)step_no;
        p.g_bc_skip = c->g_skip_pending;
        p.commit = 1;
        p.fsplit = (decltype(p.fsplit))(c->fsplit.p);
        p.usplit = (decltype(p.usplit))(c->usplit.p);
        // force sums of this step go to the other buffer; the launch's block 0
        // reduces the previous step's into that step's coefficient row
        const int fs = c->fpart_side ^ 1;
        p.fpart = (decltype(p.fpart))(c->fpart[fs].as<double>());
        p.fpart_prev = (decltype(p.fpart_prev))(c->fpart[c->fpart_side].as<double>());
        p.reduce_prev = (k > 0 && c->nrec > 0) ? 1 : 0;
        p.hist_prev = (decltype(p.hist_prev))(k > 0 ? rows + (size_t)NCOEF * (k - 1) : nullptr);
        c->launch_split_kernel(p);
        c->collide_field(c->fsplit.p, c->usplit.p, wr);
        c->fpart_side = fs;
        c->last_plane = p.plane;
        c->last_step_no = step_no;
        c->last_g_skip = c->g_skip_pending;
        c->g_skip_pending = 0;
        c->side = wr;
        if (c->convective()) c->prev_side ^= 1;
      }
      // the chunk's last step is reduced here, so every chunk's rows are
      // complete when its copy is issued
      if (c->nrec > 0) {
        const ForceReduce r = make_force_reduce(c->Q, c->fpart[c->fpart_side].as<double>(), c->lrec.as<LinkCell>(),
                                                c->nrec, c->d.nx, c->d.ny, c->d.center[0], c->d.center[1],
                                                c->d.center[2], fd, md, rows + (size_t)NCOEF * (n - 1));
        ck(launch_reduce_forces(r, c->stream), "force reduction");
      } else {
        ck(hipMemsetAsync(rows, 0, sizeof(double) * NCOEF * (size_t)n, c->stream), "zero history");
      }
      if (c->ring_cap > 0) {
        const RingAppend ra = c->ring_append(rows, n);
        RingAppend* rec = c->ring_rec.as<RingAppend>() + b;
        ck(hipMemcpyAsync(rec, &ra, sizeof(RingAppend), hipMemcpyHostToDevice, c->stream), "upload ring append");
        ck(launch_ring_append(rec, 1, n, c->stream), "ring append");
        c->ring_count += n;
      }
      if (hist) {
        ck(launch_rows_to_host(rows, c->hist_stage[b], 1, n, n, c->stream), "download history");
      }
      ck(hipEventRecord(c->chunk_ev[b], c->stream), "chunk event");
      pend_off[b] = s0;
      pend_n[b] = n;
    }
    // keep the last step's inlet plane for re-deriving its state later
    if (need_planes) {
      ck(hipMemcpyAsync(c->last_plane_buf.p, c->last_plane, sizeof(double) * plane_cnt, hipMemcpyDeviceToDevice,
                        c->stream),
         "keep last plane");
      c->last_plane = c->last_plane_buf.as<double>();
    }
    c->invalidate();
    drain(0);
    drain(1);
    if (need_planes) ck(hipStreamSynchronize(c->stream), "step sync");  // the copy above follows the last event
  });
}

// ------------------------------------------------ allocation-free stepping
// aero_step_noalloc keeps every buffer the fused step needs beyond the
// context's own state -- the batched kernel's case table, its tile list, the
// chunk's coefficient rows and the last step's reduction record -- in this
// file-scope device array instead of buffers the library allocates.  Each
// device holds its own copy for the life of the process.  In host code the
// array's name is not a device address: it is reached only through the
// per-device address hipGetSymbolAddress returns (noalloc_scratch()).
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)4 << 20;
static __device__ __align__(256) unsigned char g_noalloc_scratch[NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[MAX_DEVICES] = {};

// Device address of the scratch array on `device` (current), looked up once
// per device.
void* noalloc_scratch(int device) {
  if (device < 0 || device >= MAX_DEVICES) throw std::runtime_error("aero_step_noalloc: device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* ptr = nullptr;
    ck(hipGetSymbolAddress(&ptr, HIP_SYMBOL(g_noalloc_scratch)), "hipGetSymbolAddress(scratch)");
    if (!ptr) throw std::runtime_error("hipGetSymbolAddress(scratch) returned null");
    g_scratch_ptr[device] = ptr;
  }
  return g_scratch_ptr[device];
}

// Scratch layout: case tables for the two step parities' chunk, reduction
// record, coefficient rows of one chunk, then the tile list.
struct NoallocLayout {
  size_t cases, red, rapp, rows, tiles, end;
  explicit NoallocLayout(int ntiles) {
    auto up = [](size_t v) { return (v + 255) / 256 * 256; };
    cases = 0;
    red = up(cases + sizeof(BatchCase));
    rapp = up(red + sizeof(ForceReduce));
    rows = up(rapp + sizeof(RingAppend));
    tiles = up(rows + sizeof(double) * NCOEF * STEP_CHUNK);
    end = tiles + sizeof(int2) * (size_t)ntiles;
  }
};

}  // namespace

// The tile list of one case (case index 0, blocks 0..n-1) in the XCD order
// xcd_order() forms on the host for `run`-tile runs.
__global__ __launch_bounds__(BLOCK) void noalloc_tiles_kernel(int2* tiles, int n, int run) {
  const int b = blockIdx.x * BLOCK + threadIdx.x;
  if (b >= n) return;
  tiles[b] = make_int2(0, xcd_block(b, run, n));
}

__global__ __launch_bounds__(BLOCK) void noalloc_zero_kernel(double* p, int n) {
  const int k = blockIdx.x * BLOCK + threadIdx.x;
  if (k < n) p[k] = 0.0;
}

extern "C" {

int aero_step_noalloc(aero_ctx* c, int32_t nsteps, int64_t step0, double* hist) {
  return guarded([&] {
    if (nsteps <= 0) return;
    if (!c->has_state) throw std::runtime_error("no state: call aero_set_f first");
    if (c->les_split()) throw std::invalid_argument("aero_step_noalloc: LES contexts step with aero_step");
    if (c->need_planes())
      throw std::invalid_argument("aero_step_noalloc: inlets with per-step planes step with aero_step");
    c->set_device();
    const int ntiles = c->sblocks + 1;
    const NoallocLayout L(ntiles);
    if (L.end > NOALLOC_SCRATCH_BYTES)
      throw std::invalid_argument("aero_step_noalloc: grid too large for the " +
                                  std::to_string(NOALLOC_SCRATCH_BYTES >> 20) + " MiB scratch (" +
                                  std::to_string(ntiles) + " tiles; at most " +
                                  std::to_string((NOALLOC_SCRATCH_BYTES - L.tiles) / sizeof(int2)) + ")");
    unsigned char* base = (unsigned char*)noalloc_scratch(c->device);
    BatchCase* d_cases = (BatchCase*)(base + L.cases);
    ForceReduce* d_red = (ForceReduce*)(base + L.red);
    double* d_rows = (double*)(base + L.rows);
    int2* d_tiles = (int2*)(base + L.tiles);
    hipStream_t st = c->stream;
    // tiles as the allocating path orders them
    const size_t zreuse = c->d.nz > 1 ? (size_t)XCDS * 3 * c->Q * (size_t)c->xp * c->d.ny * c->esz : 0;
    const int run = zreuse > ZREUSE_BYTES ? XCD_RUN_LARGE : std::max(1, (ntiles + XCDS - 1) / XCDS);
    noalloc_tiles_kernel<<<(ntiles + BLOCK - 1) / BLOCK, BLOCK, 0, st>>>(d_tiles, ntiles, run);
    ck(hipGetLastError(), "noalloc tiles");
    const bool nt = 2 * c->state_bytes() * (c->d.thermal ? 2 : 1) > NT_STORE_BYTES;
    for (int s0 = 0; s0 < nsteps; s0 += STEP_CHUNK) {
      const int nk = std::min(STEP_CHUNK, nsteps - s0);
      noalloc_zero_kernel<<<(NCOEF * STEP_CHUNK + BLOCK - 1) / BLOCK, BLOCK, 0, st>>>(d_rows, NCOEF * STEP_CHUNK);
      ck(hipGetLastError(), "noalloc zero rows");
      BatchCase bc = c->make_case(step0 + s0, nullptr, d_rows);
      bc.p[0].nt_store = bc.p[1].nt_store = nt ? 1 : 0;
      ck(hipMemcpyAsync(d_cases, &bc, sizeof(BatchCase), hipMemcpyHostToDevice, st), "noalloc case table");
      for (int k = 0; k < nk; ++k) {
        if (c->d.inlet == AERO_IN_RECYCLING) c->recycling_prepass((c->side + k) & 1, st);
        ck(launch_batch_step(c->Q, c->d.collision, c->d.storage, c->d.thermal != 0, d_cases, d_tiles, ntiles, k, st,
                             nullptr, nullptr, c->les_local()),
           "noalloc step launch");
      }
      if (c->nrec > 0) {
        const ForceReduce r = make_force_reduce(c->Q, c->fpart[(c->fpart_side + nk) & 1].as<double>(),
                                                c->lrec.as<LinkCell>(), c->nrec, c->d.nx, c->d.ny, c->d.center[0],
                                                c->d.center[1], c->d.center[2], c->force_denom(), c->moment_denom(),
                                                d_rows + (size_t)NCOEF * (nk - 1));
        ck(hipMemcpyAsync(d_red, &r, sizeof(ForceReduce), hipMemcpyHostToDevice, st), "noalloc reduction");
        ck(launch_reduce_forces_many(c->Q, d_red, 1, 0, st), "noalloc force reduction");
      }
      if (c->ring_cap > 0) {
        const RingAppend ra = c->ring_append(d_rows, nk);
        RingAppend* d_rapp = (RingAppend*)(base + L.rapp);
        ck(hipMemcpyAsync(d_rapp, &ra, sizeof(RingAppend), hipMemcpyHostToDevice, st), "noalloc ring append");
        ck(launch_ring_append(d_rapp, 1, nk, st), "noalloc ring append");
        c->ring_count += nk;
      }
      if (hist)
        ck(hipMemcpyAsync(hist + (size_t)NCOEF * s0, d_rows, sizeof(double) * NCOEF * nk, hipMemcpyDeviceToHost, st),
           "noalloc rows");
      // the host copies of the case table and reduction record, and the
      // scratch, are reused by the next chunk
      ck(hipStreamSynchronize(st), "noalloc chunk sync");
      c->advance_state(nk, step0 + s0 + nk, nullptr);
    }
    c->invalidate();
  });
}

}  // extern "C"

int aero_tune_set(int32_t q, int32_t collision, int32_t storage, int32_t thermal, int32_t grid_class_,
                  int32_t xcd_run, int32_t nt, int32_t graph, int32_t rowgroups) {
  return guarded([&] {
    if ((q != 9 && q != 19) || collision < 0 || collision > 2 || storage < 0 || storage > 3 || grid_class_ < 0 ||
        grid_class_ > 40 || nt < -1 || nt > 1 || graph < 0 || graph > 1 || xcd_run < -1 || rowgroups < -1 ||
        rowgroups > 0)
      throw std::invalid_argument("aero_tune_set: argument out of range");
    std::lock_guard<std::mutex> lock(g_tune_mu);
    g_tune[{q, collision, storage, thermal != 0, grid_class_}] = TuneVal{xcd_run, nt, graph, rowgroups};
    ++g_tune_gen;
  });
}

int aero_tune_clear(void) {
  return guarded([&] {
    std::lock_guard<std::mutex> lock(g_tune_mu);
    g_tune.clear();
    ++g_tune_gen;
  });
}

int aero_tune_grid_class(int64_t ntiles, int32_t* cls) {
  return guarded([&] {
    if (ntiles < 1) throw std::invalid_argument("aero_tune_grid_class: ntiles must be positive");
    *cls = grid_class(ntiles);
  });
}

// Halo faces of a z-slab decomposition (D3Q19): the populations with
// e_z = dir of plane z of the state the next step reads, to or from a
// device buffer of aero_halo_bytes bytes: f, then g when the scalar is on.
static void halo_call(aero_ctx* c, bool pack, int32_t z, int32_t dir, void* buf) {
  if (c->Q != 19) throw std::invalid_argument("halo faces are D3Q19 (z slabs)");
  if (z < 0 || z >= c->d.nz || (dir != 1 && dir != -1)) throw std::invalid_argument("halo: bad plane or direction");
  if (!buf) throw std::invalid_argument("halo: null buffer");
  c->set_device();
  const size_t fb = 5 * (size_t)c->d.nx * c->d.ny * c->esz;
  ck(launch_halo(pack, (int)c->esz, c->B[c->side].p, buf, c->d.nx, c->d.ny, c->xp, z, dir, c->stream), "halo f");
  if (c->d.thermal)
    ck(launch_halo(pack, (int)c->esz, c->G[c->side].p, (char*)buf + fb, c->d.nx, c->d.ny, c->xp, z, dir, c->stream),
       "halo g");
  if (!pack) c->invalidate();
}

int aero_halo_bytes(aero_ctx* c, int64_t* bytes) {
  return guarded([&] { *bytes = (int64_t)(5 * (size_t)c->d.nx * c->d.ny * c->esz * (c->d.thermal ? 2 : 1)); });
}

int aero_halo_pack(aero_ctx* c, int32_t z, int32_t dir, void* buf) {
  return guarded([&] { halo_call(c, true, z, dir, buf); });
}

int aero_halo_unpack(aero_ctx* c, int32_t z, int32_t dir, const void* buf) {
  return guarded([&] { halo_call(c, false, z, dir, (void*)buf); });
}

int aero_halo_exchange(aero_ctx* lower, aero_ctx* upper) {
  return guarded([&] {
    if (!lower || !upper || lower == upper) throw std::invalid_argument("halo exchange: two distinct slabs");
    if (lower->d.nx != upper->d.nx || lower->d.ny != upper->d.ny || lower->d.storage != upper->d.storage ||
        lower->d.thermal != upper->d.thermal || lower->device != upper->device)
      throw std::invalid_argument("halo exchange: slabs differ in plane shape, storage, scalar or device");
    const size_t hb = 5 * (size_t)lower->d.nx * lower->d.ny * lower->esz * (lower->d.thermal ? 2 : 1);
    lower->set_device();
    lower->halo_buf.alloc(2 * hb);
    ck(hipStreamSynchronize(upper->stream), "halo sync");
    // lower's top plane -> upper's bottom halo; upper's bottom plane -> lower's top halo
    halo_call(lower, true, lower->d.nz - 2, +1, lower->halo_buf.p);
    halo_call(upper, true, 1, -1, (char*)lower->halo_buf.p + hb);
    ck(hipStreamSynchronize(lower->stream), "halo sync");
    ck(hipStreamSynchronize(upper->stream), "halo sync");
    halo_call(upper, false, 0, +1, lower->halo_buf.p);
    halo_call(lower, false, lower->d.nz - 1, -1, (char*)lower->halo_buf.p + hb);
    ck(hipStreamSynchronize(upper->stream), "halo sync");
    ck(hipStreamSynchronize(lower->stream), "halo sync");
  });
}

int aero_step_tiles(aero_ctx* c, int64_t* ntiles) {
  return guarded([&] { *ntiles = (int64_t)c->sblocks_rows + 1; });
}

int aero_set_history_ring(aero_ctx* c, int64_t cap) {
  return guarded([&] {
    if (cap < 0) throw std::invalid_argument("aero_set_history_ring: negative capacity");
    c->set_device();
    ck(hipStreamSynchronize(c->stream), "ring sync");
    c->ring.release();
    c->ring_rec.release();
    c->ring_cap = 0;
    c->ring_count = 0;
    if (cap == 0) return;
    c->ring.alloc(sizeof(double) * 3 * (size_t)cap);
    c->ring_rec.alloc(sizeof(RingAppend) * 2);
    c->ring_cap = cap;
  });
}

}  // extern "C"

namespace {
// Buffers of the stationarity test, grown as needed, with their own stream.
struct StatScratch {
  hipStream_t stream = nullptr;
  DevBuf series, out, work, siglen, peak, data;
};
std::mutex g_stat_mu;
StatScratch g_stat[64];

void run_stationarity(std::vector<StatSeries>& ser, const StatParams& sp, double* out_host,
                      const std::vector<std::vector<double>>* upload) {
  int dev = 0;
  ck(hipGetDevice(&dev), "hipGetDevice");
  if (dev < 0 || dev >= 64) throw std::runtime_error("aero_stationarity: device index out of range");
  std::lock_guard<std::mutex> lock(g_stat_mu);
  StatScratch& S = g_stat[dev];
  if (!S.stream) ck(hipStreamCreateWithFlags(&S.stream, hipStreamNonBlocking), "hipStreamCreate");
  const int n = (int)ser.size();
  if (upload) {  // host histories: plain arrays in one device buffer
    size_t total = 0;
    for (const auto& v : *upload) total += v.size();
    S.data.alloc(sizeof(double) * std::max<size_t>(total, 1));
    size_t off = 0;
    std::vector<const double*> base(upload->size());
    for (size_t k = 0; k < upload->size(); ++k) {
      const auto& v = (*upload)[k];
      if (!v.empty())
        ck(hipMemcpyAsync(S.data.as<double>() + off, v.data(), sizeof(double) * v.size(), hipMemcpyHostToDevice,
                          S.stream),
           "upload histories");
      base[k] = S.data.as<double>() + off;
      off += v.size();
    }
    for (int i = 0; i < n; ++i) {
      ser[i].cd = base[3 * i];
      ser[i].ly = base[3 * i + 1];
      ser[i].lz = (*upload)[3 * i + 2].empty() ? nullptr : base[3 * i + 2];
    }
  }
  S.series.alloc(sizeof(StatSeries) * n);
  S.out.alloc(sizeof(double) * NSTAT * n);
  S.work.alloc(sizeof(double) * 3 * (size_t)n * sp.st_window);
  S.siglen.alloc(sizeof(int) * 3 * n);
  S.peak.alloc(sizeof(int) * 3 * n);
  ck(hipMemcpyAsync(S.series.p, ser.data(), sizeof(StatSeries) * n, hipMemcpyHostToDevice, S.stream), "upload series");
  ck(launch_stationarity(S.series.as<StatSeries>(), n, sp, S.out.as<double>(), S.work.as<double>(),
                         S.siglen.as<int>(), S.peak.as<int>(), S.stream),
     "stationarity");
  ck(hipMemcpyAsync(out_host, S.out.p, sizeof(double) * NSTAT * n, hipMemcpyDeviceToHost, S.stream), "download");
  ck(hipStreamSynchronize(S.stream), "stationarity sync");
}

StatParams stat_params(int32_t conv_window, double conv_tol, int32_t st_window, double st_tol) {
  if (conv_window < 1 || st_window < 1 || st_window > STAT_MAX_WINDOW)
    throw std::invalid_argument("aero_stationarity: windows must be positive and the Strouhal window at most " +
                                std::to_string(STAT_MAX_WINDOW));
  return StatParams{conv_window, conv_tol, st_window, st_tol};
}
}  // namespace

extern "C" {

int aero_stationarity(int32_t n, aero_ctx* const* ctxs, const int64_t* counts, int32_t conv_window, double conv_tol,
                      int32_t st_window, double st_tol, double* out) {
  return guarded([&] {
    if (n <= 0) return;
    const StatParams sp = stat_params(conv_window, conv_tol, st_window, st_tol);
    std::vector<StatSeries> ser(n);
    int dev = -1;
    for (int i = 0; i < n; ++i) {
      aero_ctx* c = ctxs[i];
      if (!c || c->ring_cap <= 0) throw std::invalid_argument("aero_stationarity: context without a history ring");
      if (dev >= 0 && c->device != dev) throw std::invalid_argument("aero_stationarity: contexts on different devices");
      dev = c->device;
      if (c->ring_cap < std::max<long long>(conv_window, st_window))
        throw std::invalid_argument("aero_stationarity: ring smaller than the windows");
      if (counts[i] < 0 || counts[i] > c->ring_count)
        throw std::invalid_argument("aero_stationarity: count beyond the samples the ring has received");
      const double* r = c->ring.as<double>();
      ser[i] = StatSeries{r, r + c->ring_cap, c->Q == 19 ? r + 2 * c->ring_cap : nullptr, counts[i], c->ring_cap,
                          c->d.D, c->d.u0};
    }
    ck(hipSetDevice(dev), "hipSetDevice");
    run_stationarity(ser, sp, out, nullptr);
  });
}

int aero_stationarity_arrays(int32_t n, const double* const* cd, const double* const* ly, const double* const* lz,
                             const int64_t* len, const double* D, const double* u0, int32_t conv_window,
                             double conv_tol, int32_t st_window, double st_tol, double* out) {
  return guarded([&] {
    if (n <= 0) return;
    const StatParams sp = stat_params(conv_window, conv_tol, st_window, st_tol);
    const long long keep = std::max<long long>(conv_window, st_window);
    std::vector<std::vector<double>> up(3 * (size_t)n);
    std::vector<StatSeries> ser(n);
    for (int i = 0; i < n; ++i) {
      if (len[i] < 0) throw std::invalid_argument("aero_stationarity_arrays: negative length");
      const long long m = std::min<long long>(len[i], keep);
      const long long from = len[i] - m;
      up[3 * i].assign(cd[i] + from, cd[i] + len[i]);
      up[3 * i + 1].assign(ly[i] + from, ly[i] + len[i]);
      if (lz && lz[i]) up[3 * i + 2].assign(lz[i] + from, lz[i] + len[i]);
      ser[i] = StatSeries{nullptr, nullptr, nullptr, m, 0, D[i], u0[i]};
    }
    run_stationarity(ser, sp, out, &up);
  });
}

int aero_set_timing(aero_ctx* c, int32_t enabled) {
  return guarded([&] {
    c->set_device();
    c->timing = enabled != 0;
    // the events are created here, outside any timed stepping
    while (c->timing && c->ev_pool.size() < 2 * (size_t)STEP_CHUNK) {
      hipEvent_t e;
      ck(hipEventCreate(&e), "hipEventCreate");
      c->ev_pool.push_back(e);
    }
    c->timed_ms = 0.0;
    c->timed_launches = 0;
  });
}

int aero_cell_counts(aero_ctx* c, int64_t* total, int64_t* link_cells) {
  return guarded([&] {
    *total = c->N;
    *link_cells = c->nrec;
  });
}

int aero_get_timing(aero_ctx* c, double* kernel_ms, int64_t* launches) {
  return guarded([&] {
    *kernel_ms = c->timed_ms;
    *launches = c->timed_launches;
  });
}

int aero_macroscopic(aero_ctx* c, double* rho, double* ux, double* uy, double* uz) {
  return guarded([&] {
    c->set_device();
    c->materialize_f();
    c->tmp.alloc(sizeof(double) * 4 * c->N);
    ck(launch_macro4_f64(c->Q, c->N, c->fcache.as<double>(), c->tmp.as<double>(), c->stream), "macro");
    double* outs[4] = {rho, ux, uy, uz};
    for (int k = 0; k < 4; ++k)
      if (outs[k])
        ck(hipMemcpyAsync(outs[k], c->tmp.as<double>() + (size_t)k * c->N, sizeof(double) * c->N,
                          hipMemcpyDeviceToHost, c->stream),
           "download macro");
    ck(hipStreamSynchronize(c->stream), "sync");
  });
}

int aero_check(aero_ctx* c, double* stats) {
  return guarded([&] {
    c->set_device();
    c->checkout.alloc(sizeof(double) * NCHECK);
    c->check_stats(c->stream, c->checkout.as<double>());
    ck(hipMemcpyAsync(stats, c->checkout.p, sizeof(double) * NCHECK, hipMemcpyDeviceToHost, c->stream), "download");
    ck(hipStreamSynchronize(c->stream), "sync");
  });
}

int aero_batch_check(aero_batch* b, int32_t n, aero_ctx* const* ctxs, double* stats) {
  return guarded([&] {
    if (n <= 0) return;
    ck(hipSetDevice(b->device), "hipSetDevice");
    // One pass per (lattice, storage, scalar) group over the tiles (case,
    // block) of its cases, then one reduction block per case: the blocks and
    // the reduction order of aero_check, so each row equals that case's own.
    long long nparts = 0;
    std::vector<long long> part0((size_t)n);
    for (int i = 0; i < n; ++i) {
      aero_ctx* c = ctxs[i];
      if (!c) throw std::invalid_argument("null context in batch");
      if (c->device != b->device) throw std::invalid_argument("every context of a batch must be on its device");
      part0[i] = nparts;
      nparts += c->nblocks;
    }
    b->checkout.alloc(sizeof(double) * NCHECK * (size_t)n);
    b->checkpart.alloc(sizeof(double) * NCHECK * (size_t)nparts);
    std::vector<StepParams> ps((size_t)n);
    std::vector<int> order((size_t)n);
    for (int i = 0; i < n; ++i) {
      aero_ctx* c = ctxs[i];
      drain(c);  // work queued on the context's own stream
      ps[i] = c->check_params(b->stream, b->checkpart.as<double>() + (size_t)NCHECK * part0[i]);
      order[i] = i;
    }
    auto key = [&](int i) {
      const aero_ctx* c = ctxs[i];
      return std::make_tuple(c->Q, c->d.storage, c->d.thermal != 0);
    };
    std::stable_sort(order.begin(), order.end(), [&](int a, int c) { return key(a) < key(c); });
    struct Group {
      int rep;        // a case of the group
      size_t t0, t1;  // its tiles
    };
    std::vector<Group> groups;
    for (size_t j = 0; j < order.size(); ++j) {
      const int i = order[j];
      if (j == 0 || key(i) != key(order[j - 1])) groups.push_back({i, 0, 0});
    }
    // the tile list depends only on the cases (their order and sizes): kept
    // on the device while the same cases are checked again
    std::vector<std::pair<const aero_ctx*, int>> sig((size_t)n);
    for (int i = 0; i < n; ++i) sig[i] = {ctxs[i], ctxs[i]->nblocks};
    const size_t pbytes = sizeof(StepParams) * (size_t)n;
    const size_t toff = (pbytes + 255) / 256 * 256;
    std::vector<int2> tiles;
    {
      size_t nt = 0, g = 0;
      const bool fresh = sig != b->check_sig;
      if (fresh) tiles.resize((size_t)nparts);
      for (size_t j = 0; j < order.size(); ++j) {
        const int i = order[j];
        if (j > 0 && key(i) != key(order[j - 1])) groups[++g].t0 = nt;
        if (fresh)
          for (int blk = 0; blk < ctxs[i]->nblocks; ++blk) tiles[nt + blk] = make_int2(i, blk);
        nt += ctxs[i]->nblocks;
        groups[g].t1 = nt;
      }
      if (fresh) {
        b->check_sig.clear();
        b->checkpar.alloc(toff + sizeof(int2) * tiles.size());
      }
    }
    const StepParams* dps = (const StepParams*)b->checkpar.p;
    const int2* dtiles = (const int2*)((char*)b->checkpar.p + toff);
    ck(hipMemcpyAsync(b->checkpar.p, ps.data(), pbytes, hipMemcpyHostToDevice, b->stream), "upload check params");
    if (!tiles.empty())
      ck(hipMemcpyAsync((void*)dtiles, tiles.data(), sizeof(int2) * tiles.size(), hipMemcpyHostToDevice, b->stream),
         "upload check tiles");
    for (const Group& g : groups) {
      const aero_ctx* c = ctxs[g.rep];
      ck(launch_check_batch(c->Q, c->d.storage, c->d.thermal != 0, dps, dtiles + g.t0, (int)(g.t1 - g.t0), b->stream),
         "check launch");
    }
    ck(launch_reduce_check_batch(dps, n, b->checkout.as<double>(), b->stream), "check reduce");
    ck(hipMemcpyAsync(stats, b->checkout.p, sizeof(double) * NCHECK * (size_t)n, hipMemcpyDeviceToHost, b->stream),
       "download");
    ck(hipStreamSynchronize(b->stream), "sync");  // the host vectors above stay alive until here
    if (!tiles.empty()) b->check_sig = std::move(sig);
  });
}

int aero_link_forces(aero_ctx* c, int32_t* cells, double* forces, int32_t* nrec) {
  return guarded([&] {
    c->set_device();
    if (nrec) *nrec = c->nrec;
    if (cells) std::memcpy(cells, c->rec_cells.data(), sizeof(int) * c->nrec);
    if (forces && c->nrec) {
      ck(hipStreamSynchronize(c->stream), "sync");
      std::vector<double> part((size_t)NPART * c->nrec);
      ck(hipMemcpy(part.data(), c->fpart[c->fpart_side].p, sizeof(double) * part.size(), hipMemcpyDeviceToHost),
         "download");
      for (int k = 0; k < c->nrec; ++k) {
        double* out = forces + (size_t)NFORCE * k;
        const double* pk = part.data() + (size_t)NPART * k;
        if (c->Q == 9)
          cell_force_terms<D2Q9>(pk, c->rec_cells[k], c->rec_masks[k], c->d.nx, c->d.ny, c->d.center[0],
                                 c->d.center[1], c->d.center[2], out);
        else
          cell_force_terms<D3Q19>(pk, c->rec_cells[k], c->rec_masks[k], c->d.nx, c->d.ny, c->d.center[0],
                                  c->d.center[1], c->d.center[2], out);
      }
    }
  });
}

int aero_synchronize(aero_ctx* c) {
  return guarded([&] {
    c->set_device();
    ck(hipStreamSynchronize(c->stream), "sync");
  });
}

int aero_device_bytes(aero_ctx* c, int64_t* bytes) {
  return guarded([&] {
    const DevBuf* bufs[] = {&c->B[0], &c->B[1], &c->G[0], &c->G[1], &c->prev[0], &c->prev[1], &c->flags, &c->wave_any,
                            &c->witems, &c->lrec, &c->rec_of, &c->qdir, &c->fpart[0], &c->fpart[1], &c->sigma, &c->minv_d, &c->minv_f,
                            &c->damp, &c->les_c2d, &c->les_c2f, &c->halo_buf, &c->fterms, &c->planes[0], &c->planes[1], &c->recyc, &c->hist_dev[0], &c->hist_dev[1], &c->last_plane_buf, &c->fsplit, &c->usplit, &c->fcache,
                            &c->gcache, &c->tmp, &c->checkpart, &c->checkout, &c->ring, &c->ring_rec};
    int64_t t = 0;
    for (const DevBuf* b : bufs) t += (int64_t)b->bytes;
    *bytes = t;
  });
}

}  // extern "C"
