/*
 * aero_hip.h -- C ABI of the MI355X (gfx950) lattice-Boltzmann step library.
 *
 * The reference drives its step loop from Python (aero/lbm/solver.py,
 * aero/lbm/solver3d.py); its "native" layer is a set of Numba kernels behind
 * plain NumPy arrays.  This header is the boundary a Python host binds with
 * ctypes in place of those kernels: plain C types, caller-owned host buffers
 * in the reference's own float64 layouts, no torch or HIP types.
 *
 * Conventions
 *  - Every function returns 0 on success and a nonzero code on failure; the
 *    message is available from aero_last_error() on the same thread.
 *  - Host buffers are borrowed for the duration of the call only.
 *  - Device memory is owned by the context.  A context is not thread-safe;
 *    calls on one context must be serialised by the caller.
 *  - Field layouts are the reference's C-order arrays:
 *      f, g           (Q, Nz, Ny, Nx) float64  (2D: Nz = 1)
 *      f_outlet_prev  (Q, Nz, Ny)     float64
 *      solid          (Nz, Ny, Nx)    uint8 (0/1)
 *      links          (L, 4) int32 rows (i, z, y, x)   (2D: z = 0)
 *      q_vals         (L,)   float32
 */
#ifndef AERO_HIP_H
#define AERO_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AERO_ABI_VERSION 1

/* collision (aero/lbm/solver.py:186-187) */
#define AERO_BGK 0
#define AERO_TRT 1
#define AERO_MRT 2
/* population storage */
#define AERO_STORE_F64 0
#define AERO_STORE_F32 1
#define AERO_STORE_F16S 2  /* float16 holding f - w_i, float32 arithmetic */
#define AERO_STORE_BF16S 3 /* bfloat16 holding f - w_i, float32 arithmetic */
/* inlet at x = 0 */
#define AERO_IN_NONE 0
#define AERO_IN_ZOUHE 1     /* velocity Zou-He (+ travelling sine)   boundary.py:15-45, boundary3d.py:24-68 */
#define AERO_IN_PRESSURE 2  /* 2D Zou-He pressure                    boundary.py:116-136 */
#define AERO_IN_FIELD 3     /* target velocity planes                boundary.py:48-65, boundary3d.py:71-93 */
#define AERO_IN_SEM 4       /* SEM fluctuation planes (3D)           boundary3d.py:96-137 */
#define AERO_IN_RECYCLING 5 /* recycling-rescaling (3D)              boundary3d.py:172-198 */
/* outlet at x = Nx-1 */
#define AERO_OUT_NONE 0
#define AERO_OUT_CONVECTIVE 1 /* boundary.py:78-113 */
#define AERO_OUT_ZEROGRAD 2   /* boundary.py:68-75 */
#define AERO_OUT_PRESSURE 3   /* 2D, boundary.py:139-159 */
/* walls at y = 0 and y = Ny-1 */
#define AERO_WALL_NONE 0
#define AERO_WALL_SLIP 1
#define AERO_WALL_NOSLIP 2
#define AERO_WALL_MOVING 3
/* LES (aero/lbm/les.py:273-344), finite differences on the pre-collision velocity */
#define AERO_LES_OFF 0
#define AERO_LES_SMAGORINSKY 1
#define AERO_LES_WALE 2
/* Smagorinsky eddy viscosity from the local non-equilibrium stress of each
 * cell, fused into the step (BGK / TRT; MRT ignores LES as in the reference).
 * No counterpart in the reference: tau = (tau0 + sqrt(tau0^2 + 18 C^2 |Pi^neq| / rho)) / 2
 * with |Pi^neq| = sqrt(2 Pi^neq:Pi^neq), C = les_cs times the Van Driest
 * factor of aero_set_damping.  Such contexts batch with aero_batch_step. */
#define AERO_LES_SMAGORINSKY_LOCAL 3

/* Number of doubles per step in the coefficient history written by aero_step:
 * Cd, Cl_y, Cl_z, Cm_x, Cm_y, Cm_z, Cd_p, Cd_v, then the raw FP64 sums
 * Fx Fy Fz Mx My Mz Fpx Fpy Fpz Fvx Fvy Fvz.  2D uses Cl_y as Cl and Cm_z as Cm
 * (aero/forces.py:78-139); 3D normalises by D^2 and D^3 (aero/forces3d.py:364-439). */
#define AERO_NCOEF 20
/* aero_check output: any NaN, any rho <= 0, max |(ux, uy)|, sum rho, cell count
 * (aero/diagnostics.py:32-71; 3D checks the z-midplane, solver3d.py:608-615). */
#define AERO_NCHECK 5

typedef struct aero_ctx aero_ctx;

/* Everything the constructors take (aero/lbm/solver.py:75-116,
 * aero/lbm/solver3d.py:71-119), resolved to numbers by the host. */
typedef struct aero_desc {
  int32_t q;  /* 9 or 19 */
  int32_t nx, ny, nz;
  int32_t collision;
  int32_t storage;
  int32_t inlet;
  int32_t outlet;
  int32_t wall;
  int32_t bouzidi;     /* 0: mid-link bounce-back; 1: the reference's Bouzidi rule
                          (boundary.py:303-352); 2: linear Bouzidi-Firdaouss-Lallemand */
  int32_t sponge_thickness;
  int32_t les;
  int32_t thermal;
  int32_t buoyancy;
  int32_t body_force_on;
  int32_t trt_variant; /* 0: reference TRT rates (trt2d.py:8-18); 1: standard Lambda-TRT,
                          tau+ = tau, tau- = Lambda/(tau+ - 1/2) + 1/2 with Lambda = trt_lambda */
  double omega, u0, rho0, D, trt_lambda;
  double rho_in, rho_out, inlet_perturbation;
  double sponge_strength;
  double les_cs;
  double u_wall_top, u_wall_bottom;
  double body_force[3];
  double alpha_T, T_hot, T_cold, T_ref, g_gravity, beta;
  double center[3];    /* moment reference point (x, y, z): mean of solid cells */
  double mrt_s[19];    /* MRT relaxation vector (mrt2d.py:51-77, mrt3d.py:94-135) */
  double mrt_minv[361];/* inv(M9) or pinv(M19), row-major Q x Q */
} aero_desc;

/* Library / device */
const char* aero_last_error(void);
int aero_version(void);
int aero_device_count(int32_t* n);

/* STL geometry on the device (no context needed).  tris: ntri triangles of
 * three (x, y, z) vertices, FP64, in grid coordinates.
 *
 * aero_mesh_inside: parity of the forward hits of the ray (unit direction
 * ray[3]) from each point over all triangles, with the Moller-Trumbore test
 * and eps = 1e-9 of points_inside_mesh (aero/geometry3d/mesh_mask.py:14-45).
 * points: npts (x, y, z) triples, or NULL for the nz*ny*nx cell centres
 * (x + 0.5, y + 0.5, z + 0.5), x fastest (MeshMask.mark_solid,
 * mesh_mask.py:114-124).  inside: one byte per point, 1 = inside.
 *
 * aero_mesh_sqdist: squared Euclidean distance from the centre of each listed
 * cell (linear index z*ny*nx + y*nx + x) to the nearest triangle, as
 * _sq_dist_points_to_triangle + the minimum over triangles of
 * compute_phi_field (signed_distance.py:38-137, 222-231). */
int aero_mesh_inside(const double* tris, int64_t ntri, const double* ray, const double* points, int64_t npts,
                     int32_t nx, int32_t ny, int32_t nz, uint8_t* inside);
int aero_mesh_sqdist(const double* tris, int64_t ntri, const int64_t* cells, int64_t ncells, int32_t nx, int32_t ny,
                     int32_t nz, double* sqdist);

/* Context lifecycle.  Replaces Solver.__init__ / Solver3D.__init__ state
 * allocation (solver.py:143-257, solver3d.py:136-295). */
int aero_create(const aero_desc* desc, aero_ctx** out);
int aero_destroy(aero_ctx* ctx);

/* Geometry: solid mask and the surface links + q values built by
 * build_surface_links(_3d) (boundary.py:232-280, boundary3d.py:279-330).
 * Replaces the whole geometry, so it also leaves immersed-boundary mode:
 * call aero_set_ibm again afterwards to keep it. */
int aero_set_geometry(aero_ctx* ctx, const uint8_t* solid, int64_t nlinks, const int32_t* links,
                      const float* q_vals);
/* Van Driest factor (1 - exp(-|phi|/A))^2 per cell (les.py:512-514), or NULL. */
int aero_set_damping(aero_ctx* ctx, const double* damp);
/* Immersed-boundary mode (ibm_enabled=True, solver.py:280-285, 379-383,
 * solver3d.py:328-333, 426): the surface links stop bouncing back (they
 * still carry the momentum-exchange sums) and, when phi (N doubles, the
 * SDF) is given, the fluid cells with 0 < phi <= 1.5 receive Guo forcing
 * toward a resting wall before every collision (ibm_guo.py:110-182).  Call
 * after aero_set_geometry and before the state is set. */
int aero_set_ibm(aero_ctx* ctx, const double* phi);

/* Synthetic-eddy-method inlet planes formed on the device: replaces the
 * per-step SEMInlet.fluctuation() (aero/lbm/sem_inlet.py:71-136) of an
 * n_eddies-eddy SEM with eddy size sigma (L_int) and intensity Tu.  The host
 * keeps SEMInlet.step() (sem_inlet.py:58-69), whose random regeneration is
 * the reference's.  Once enabled, the `planes` argument of aero_step and
 * aero_batch_step carries, per step, the eddy state after that step's
 * advance -- 6 x n_eddies doubles: x, y, z, then the u, v, w signs --
 * instead of 3 x nz x ny plane values.  aero_sem_planes returns the planes
 * that n such snapshots give (n x 3 x nz x ny doubles), for checking. */
int aero_set_sem(aero_ctx* ctx, int32_t n_eddies, double sigma, double Tu);
int aero_sem_planes(aero_ctx* ctx, int32_t n, const double* eddies, double* planes);

/* State: the reference's self.f (post-boundary, pre-collision), self.g and
 * self._f_outlet_prev, read and written in the reference layout. */
int aero_set_f(aero_ctx* ctx, const double* f);
int aero_get_f(aero_ctx* ctx, double* f);
/* self.f = compute_feq(rho, ux, uy[, uz]) (d2q9.py:66-89, d3q19.py:105-129)
 * evaluated on the device, bit-identical to the host formula: the
 * initial states of solver.py:218-223 / solver3d.py:239-247 and of the
 * benchmark cases without an N x Q host field and its upload.  Each field
 * is N doubles in the reference layout or NULL for a constant: rho0 and u0
 * of the descriptor, 0 for uy and uz (uz must be NULL in 2D). */
int aero_set_f_equilibrium(aero_ctx* ctx, const double* rho, const double* ux, const double* uy,
                           const double* uz);
int aero_set_g(aero_ctx* ctx, const double* g);
int aero_get_g(aero_ctx* ctx, double* g);
int aero_set_outlet_prev(aero_ctx* ctx, const double* prev);
int aero_get_outlet_prev(aero_ctx* ctx, double* prev);

/* Advance nsteps reference steps (Solver._step, solver.py:276-464;
 * Solver3D._step, solver3d.py:325-541).  step0 is the step count before the
 * call (the travelling inlet sine uses step0 + k + 1).  planes: for
 * AERO_IN_FIELD / AERO_IN_SEM, nsteps x 3 x (Nz*Ny) doubles (2D: ux, uy, 0 per
 * row; 3D: the target or fluctuation planes (u, v, w)), else NULL.
 * hist: nsteps x AERO_NCOEF doubles, or NULL. */
int aero_step(aero_ctx* ctx, int32_t nsteps, int64_t step0, const double* planes, double* hist);

/* aero_step without any device-memory allocation: no counterpart in the
 * reference.  For callers that cannot allocate (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same arguments and
 * results as aero_step for a context whose inlet takes no per-step planes and
 * that has no LES split step (errors otherwise).  The case table, tile list,
 * coefficient rows and reduction record live in one file-scope device array
 * of AERO_NOALLOC_SCRATCH_BYTES per GPU, shared by every call on that GPU:
 * calls that overlap on one GPU race.  A grid needs at most
 * (AERO_NOALLOC_SCRATCH_BYTES - 64 KiB) / 8 tiles of 256 padded cells (about
 * 130 M cells); larger grids are rejected. */
#define AERO_NOALLOC_SCRATCH_BYTES (4u << 20)
int aero_step_noalloc(aero_ctx* ctx, int32_t nsteps, int64_t step0, double* hist);

/* Launch tuning of the batched step (no counterpart in the reference: the
 * autotuner's table, aero_amd/autotune.py).  One entry per kernel group --
 * lattice q, collision, storage, scalar transport -- and grid class (floor
 * of log2 of the group's tile count with one row per wave,
 * aero_tune_grid_class): xcd_run 0 keeps the built-in tile order, -1 the
 * natural order, k > 0 runs of k tiles per XCD; nt -1 keeps the built-in
 * non-temporal-store rule, 0 / 1 turn it off / on; graph 1 replays each
 * chunk's step launches from a captured HIP graph; rowgroups -1 keeps the
 * built-in wave shape (the 2-byte D2Q9 kernels step bulk rows four to a
 * wave), 0 steps one row per wave.  Entries change the schedule only, never
 * the results.  aero_step_tiles returns the tile count of one context's step
 * with one row per wave (the grid-class basis). */
int aero_tune_set(int32_t q, int32_t collision, int32_t storage, int32_t thermal, int32_t grid_class,
                  int32_t xcd_run, int32_t nt, int32_t graph, int32_t rowgroups);
int aero_tune_clear(void);
int aero_tune_grid_class(int64_t ntiles, int32_t* grid_class);
int aero_step_tiles(aero_ctx* ctx, int64_t* ntiles);

/* Halo faces of a z-slab decomposition.  Replaces the ghost-plane exchange
 * of MPISolver3D._halo_exchange (aero/lbm/mpi_solver3d.py:116-150), which
 * sends whole boundary planes with MPI Sendrecv on the host; here only the 5
 * populations that cross the face move, on the device.
 * A slab context holds its planes plus one halo plane below (z = 0) and one
 * above (z = nz-1); each step, before aero_step, it sends the populations that
 * cross its faces and receives its neighbours': aero_halo_pack copies the 5
 * D3Q19 populations with e_z = dir of plane z of the state the next step
 * reads into buf (device memory, aero_halo_bytes bytes: [5][ny][nx] storage
 * elements of f, then of g with the scalar on); aero_halo_unpack writes them
 * into plane z.  Bottom face: pack (1, -1), the neighbour below unpacks into
 * its (nz-1, -1); top face: pack (nz-2, +1), the neighbour above unpacks
 * into its (0, +1).  Calls are ordered on the context's stream. */
int aero_halo_bytes(aero_ctx* ctx, int64_t* bytes);
int aero_halo_pack(aero_ctx* ctx, int32_t z, int32_t dir, void* buf);
int aero_halo_unpack(aero_ctx* ctx, int32_t z, int32_t dir, const void* buf);
/* Both faces between two slab contexts on one device (lower below upper):
 * the pack / unpack pairs above through a staging buffer of the lower
 * context, synchronised with both contexts' streams. */
int aero_halo_exchange(aero_ctx* lower, aero_ctx* upper);

/* Convergence statistics on the device, in FP64 (the auto-stop test of
 * aero/diagnostics.py:87-120, 153-228 and the 3D lift choice of
 * aero/lbm/solver3d.py:634-638).  aero_set_history_ring gives a context
 * FP64 rings of `cap` samples of its drag and lift histories (Cd, Cl_y,
 * Cl_z), appended on the device by every step call from then on (cap 0
 * removes them).  aero_stationarity evaluates n contexts at once over the
 * first counts[i] samples their rings received; aero_stationarity_arrays
 * does the same for host histories (lz NULL or lz[i] NULL in 2D).  Per case
 * AERO_NSTAT doubles: stop, Cd converged, std/mean ratio, mean |Cd|, std Cd,
 * St (NaN when none), peak frequency, peak amplitude, samples in the
 * Strouhal window, St drift between its halves (NaN when not formed),
 * stationary, lift series (1: Cl / Cl_y, 2: Cl_z), then the peak amplitudes
 * of the window and of its two halves.  The Strouhal window is at most
 * AERO_STAT_MAX_WINDOW samples; a ring read while its context steps needs
 * cap >= the larger window + the steps of that call. */
#define AERO_NSTAT 15
#define AERO_STAT_MAX_WINDOW 4096
int aero_set_history_ring(aero_ctx* ctx, int64_t cap);
int aero_stationarity(int32_t n, aero_ctx* const* ctxs, const int64_t* counts, int32_t conv_window,
                      double conv_tol, int32_t st_window, double st_tol, double* out);
int aero_stationarity_arrays(int32_t n, const double* const* cd, const double* const* ly, const double* const* lz,
                             const int64_t* len, const double* D, const double* u0, int32_t conv_window,
                             double conv_tol, int32_t st_window, double st_tol, double* out);

/* Observables of the current state (compute_macroscopic, d2q9.py:44-63). */
int aero_macroscopic(aero_ctx* ctx, double* rho, double* ux, double* uy, double* uz);
/* check_stability's statistics (diagnostics.py:32-71) of the current state:
 * NCHECK doubles -- any NaN, any rho <= 0, max |u|, sum rho, cell count;
 * 3D on the z mid-plane (solver3d.py:608-615).  aero_batch_check (below)
 * computes them for many contexts at once. */
int aero_check(aero_ctx* ctx, double* stats);
/* Per-link-cell FP64 momentum-exchange sums of the last step (nrec x 12) and
 * the cells they belong to; *nrec returns the count (pass NULL arrays to query). */
int aero_link_forces(aero_ctx* ctx, int32_t* cells, double* forces, int32_t* nrec);
int aero_synchronize(aero_ctx* ctx);
/* Bytes of device memory held by the context. */
int aero_device_bytes(aero_ctx* ctx, int64_t* bytes);
/* Kernel timing with HIP events on the context's stream: when enabled,
 * aero_step brackets every step kernel launch (one per step) with events
 * recorded at the kernel's start and end; aero_get_timing returns the summed
 * kernel milliseconds and the launch count since enabling.  No reference
 * counterpart (the reference times whole steps from the host). */
int aero_set_timing(aero_ctx* ctx, int32_t enabled);
int aero_get_timing(aero_ctx* ctx, double* kernel_ms, int64_t* launches);
/* Cell counts: total cells and link cells (fluid cells with a surface link). */
int aero_cell_counts(aero_ctx* ctx, int64_t* total, int64_t* link_cells);

/* Batched stepping.  aero_batch_step advances n contexts of one device by
 * nsteps steps each with one kernel launch per step for every group of
 * contexts sharing lattice, collision, storage and scalar transport (a packed
 * tile list over all of them), on the batch's own stream.  A context computes
 * the same bits in a batch as through aero_step, which runs the same kernel
 * with a batch of one.  step0s[i], planes[i] and hists[i] mean what they do
 * for aero_step (planes / hists may be NULL, or hold NULL entries).  LES
 * contexts (finite-difference pass every step) are rejected: step them with
 * aero_step.  No counterpart in the reference, which runs one case per
 * process; this is the device side of the batched / serving path. */
typedef struct aero_batch aero_batch;
/* Free and total device memory of the current device (admission control). */
int aero_mem_info(int64_t* free_bytes, int64_t* total_bytes);
int aero_batch_create(aero_batch** out);
int aero_batch_destroy(aero_batch* batch);
int aero_batch_step(aero_batch* batch, int32_t n, aero_ctx* const* ctxs, int32_t nsteps, const int64_t* step0s,
                    const double* const* planes, double* const* hists);
/* aero_check for n contexts at once on the batch's stream (stats: n x
 * NCHECK), one host synchronisation for all: the serving loop checks every
 * live case after each chunk, as the reference's run loop does every
 * check_every steps (solver.py:517-570). */
int aero_batch_check(aero_batch* batch, int32_t n, aero_ctx* const* ctxs, double* stats);

#ifdef __cplusplus
}
#endif

#endif /* AERO_HIP_H */
