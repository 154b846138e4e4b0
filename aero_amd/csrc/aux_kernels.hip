// Support kernels around the fused step: state materialisation, field
// collision, recycling prepass, conversions and the fixed-order FP64
// reductions (forces, stability check, recycling mean).
#include "launch.h"
#include "step_impl.h"

namespace aero {

hipError_t launch_batch_step(int q, int coll, int st, bool thermal, const BatchCase* cases, const int2* tiles,
                             int ntiles, int k, hipStream_t s, void* e0, void* e1, bool les_local) {
  if (les_local) {  // local Smagorinsky kernels (BGK and TRT; MRT ignores LES)
    if (q == 9 && coll == COLL_BGK) return launch_batch_les_d2q9_bgk(st, thermal, cases, tiles, ntiles, k, s, e0, e1);
    if (q == 9 && coll == COLL_TRT) return launch_batch_les_d2q9_trt(st, thermal, cases, tiles, ntiles, k, s, e0, e1);
    if (q == 19 && coll == COLL_BGK)
      return launch_batch_les_d3q19_bgk(st, thermal, cases, tiles, ntiles, k, s, e0, e1);
    if (q == 19 && coll == COLL_TRT)
      return launch_batch_les_d3q19_trt(st, thermal, cases, tiles, ntiles, k, s, e0, e1);
    return hipErrorInvalidValue;
  }
  if (q == 9) {
    switch (coll) {
      case COLL_BGK: return launch_batch_d2q9_bgk(st, thermal, cases, tiles, ntiles, k, s, e0, e1);
      case COLL_TRT: return launch_batch_d2q9_trt(st, thermal, cases, tiles, ntiles, k, s, e0, e1);
      case COLL_MRT: return launch_batch_d2q9_mrt(st, thermal, cases, tiles, ntiles, k, s, e0, e1);
    }
  } else {
    switch (coll) {
      case COLL_BGK: return launch_batch_d3q19_bgk(st, thermal, cases, tiles, ntiles, k, s, e0, e1);
      case COLL_TRT: return launch_batch_d3q19_trt(st, thermal, cases, tiles, ntiles, k, s, e0, e1);
      case COLL_MRT: return launch_batch_d3q19_mrt(st, thermal, cases, tiles, ntiles, k, s, e0, e1);
    }
  }
  return hipErrorInvalidValue;
}

hipError_t launch_split_step(int q, int coll, int st, bool thermal, const StepParams& p, hipStream_t s) {
  if (q == 9) {
    if (coll == COLL_BGK) return launch_split_d2q9_bgk(st, thermal, p, s);
    if (coll == COLL_TRT) return launch_split_d2q9_trt(st, thermal, p, s);
  } else {
    if (coll == COLL_BGK) return launch_split_d3q19_bgk(st, thermal, p, s);
    if (coll == COLL_TRT) return launch_split_d3q19_trt(st, thermal, p, s);
  }
  return hipErrorInvalidValue;
}

// ------------------------------------------------------------ materialize
template <class L>
static hipError_t mat_l(int st, bool thermal, const StepParams& p, hipStream_t s) {
  const dim3 g((unsigned)p.nblocks);
#define AERO_CASE(ST)                                                                  \
  case ST:                                                                             \
    if (thermal) materialize_kernel<L, ST, true><<<g, BLOCK, 0, s>>>(p);               \
    else materialize_kernel<L, ST, false><<<g, BLOCK, 0, s>>>(p);                      \
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

hipError_t launch_materialize(int q, int st, bool thermal, const StepParams& p, hipStream_t s) {
  return q == 9 ? mat_l<D2Q9>(st, thermal, p, s) : mat_l<D3Q19>(st, thermal, p, s);
}

template <class L>
static hipError_t chk_l(int st, bool thermal, const StepParams* ps, const int2* tiles, int ntiles, hipStream_t s) {
  const dim3 g((unsigned)ntiles);
#define AERO_CASE(ST)                                                                  \
  case ST:                                                                             \
    if (thermal) check_batch_kernel<L, ST, true><<<g, BLOCK, 0, s>>>(ps, tiles);       \
    else check_batch_kernel<L, ST, false><<<g, BLOCK, 0, s>>>(ps, tiles);              \
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

hipError_t launch_check_batch(int q, int st, bool thermal, const StepParams* ps, const int2* tiles, int ntiles,
                              hipStream_t s) {
  if (ntiles <= 0) return hipSuccess;
  return q == 9 ? chk_l<D2Q9>(st, thermal, ps, tiles, ntiles, s) : chk_l<D3Q19>(st, thermal, ps, tiles, ntiles, s);
}

// ---------------------------------------------------------------- collide
template <class L, int COLL>
static hipError_t col_lc(int st, const CollideParams& p, hipStream_t s) {
  const dim3 g((unsigned)p.nblocks);
  switch (st) {
    case ST_F64: collide_kernel<L, COLL, ST_F64><<<g, BLOCK, 0, s>>>(p); break;
    case ST_F32: collide_kernel<L, COLL, ST_F32><<<g, BLOCK, 0, s>>>(p); break;
    case ST_F16S: collide_kernel<L, COLL, ST_F16S><<<g, BLOCK, 0, s>>>(p); break;
    case ST_BF16S: collide_kernel<L, COLL, ST_BF16S><<<g, BLOCK, 0, s>>>(p); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <class L>
static hipError_t col_l(int coll, int st, const CollideParams& p, hipStream_t s) {
  switch (coll) {
    case COLL_BGK: return col_lc<L, COLL_BGK>(st, p, s);
    case COLL_TRT: return col_lc<L, COLL_TRT>(st, p, s);
    case COLL_MRT: return col_lc<L, COLL_MRT>(st, p, s);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_collide(int q, int coll, int st, const CollideParams& p, hipStream_t s) {
  return q == 9 ? col_l<D2Q9>(coll, st, p, s) : col_l<D3Q19>(coll, st, p, s);
}

// ------------------------------------------------------ recycling prepass
// Macroscopic state of plane x = nx-2 after streaming, body force and links
// (the plane apply_recycling_rescaling_inlet_3d reads, boundary3d.py:185-190).
template <int ST>
__global__ __launch_bounds__(BLOCK) void recycling_plane_kernel(StepParams p, double* out) {
  using S = Storage<ST>;
  using C = typename S::C;
  const int row = blockIdx.x * BLOCK + threadIdx.x;
  if (row >= p.nrow) return;
  const int y = row % p.ny, z = row / p.ny;
  const int x = p.nx - 2;
  const int cell = row * p.nx + x;
  C fn[19];
  pre_outlet<D3Q19, ST>(p, dyn_of(p), (const typename S::T*)p.fin, cell, x, y, z, p.flags[cell], fn);
  double rho = 0, mx = 0, my = 0, mz = 0;
#pragma unroll
  for (int i = 0; i < 19; ++i) {
    const double v = (double)fn[i];
    rho += v;
    mx += D3Q19::EX[i] * v;
    my += D3Q19::EY[i] * v;
    mz += D3Q19::EZ[i] * v;
  }
  const double inv = rho > 0.0 ? 1.0 / rho : 0.0;
  out[row] = rho;
  out[p.nrow + row] = inv * mx;
  out[2 * p.nrow + row] = inv * my;
  out[3 * p.nrow + row] = inv * mz;
}

hipError_t launch_recycling_plane(int st, const StepParams& p, double* out, hipStream_t s) {
  const dim3 g((unsigned)((p.nrow + BLOCK - 1) / BLOCK));
  switch (st) {
    case ST_F64: recycling_plane_kernel<ST_F64><<<g, BLOCK, 0, s>>>(p, out); break;
    case ST_F32: recycling_plane_kernel<ST_F32><<<g, BLOCK, 0, s>>>(p, out); break;
    case ST_F16S: recycling_plane_kernel<ST_F16S><<<g, BLOCK, 0, s>>>(p, out); break;
    case ST_BF16S: recycling_plane_kernel<ST_BF16S><<<g, BLOCK, 0, s>>>(p, out); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// Mean ux over the plane (fixed order) and the rescale factor
// target/mean (1 when |mean| < 1e-12), stored after the 4 plane arrays.
__global__ __launch_bounds__(BLOCK) void reduce_recycling_kernel(double* recyc, int nrow, double target) {
  __shared__ double red[BLOCK];
  double acc = 0.0;
  for (int r = threadIdx.x; r < nrow; r += BLOCK) acc += recyc[nrow + r];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = BLOCK / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double mean = red[0] / (double)nrow;
    recyc[4 * nrow] = fabs(mean) < 1e-12 ? 1.0 : target / mean;
  }
}

hipError_t launch_reduce_recycling(double* recyc, int nrow, double target_u0, hipStream_t s) {
  reduce_recycling_kernel<<<1, BLOCK, 0, s>>>(recyc, nrow, target_u0);
  return hipGetLastError();
}

// ------------------------------------------------------- scalar state
// Pre-collision scalar state (the reference's self.g): pull of the stored
// post-collision g plus the y-wall Dirichlet values (thermal.py:85-100).
template <class L, int ST>
__global__ __launch_bounds__(BLOCK) void materialize_g_kernel(StepParams p, double* out) {
  using S = Storage<ST>;
  using C = typename S::C;
  const long long c64 = (long long)blockIdx.x * BLOCK + threadIdx.x;
  if (c64 >= p.ncell) return;
  const int cell = (int)c64;
  const int x = cell % p.nx, yz = cell / p.nx, y = yz % p.ny, z = yz / p.ny;
  C g[L::Q];
  gather<L, ST, 1>(p, (const typename S::T*)p.gin, x, y, z, (C)p.gshift, g);
  const size_t N = (size_t)p.ncell;
#pragma unroll
  for (int i = 0; i < L::Q; ++i) {
    double v = (double)g[i];
    if (!p.g_bc_skip) {
      if (y == 0) v = L::W[i] * p.T_hot;
      if (y == p.ny - 1) v = L::W[i] * p.T_cold;
    }
    out[(size_t)i * N + cell] = v;
  }
}

hipError_t launch_materialize_g(int q, int st, const StepParams& p, double* out, hipStream_t s) {
  const dim3 g((unsigned)((p.ncell + BLOCK - 1) / BLOCK));
#define AERO_CASE(ST)                                                                      \
  case ST:                                                                                 \
    if (q == 9) materialize_g_kernel<D2Q9, ST><<<g, BLOCK, 0, s>>>(p, out);                \
    else materialize_g_kernel<D3Q19, ST><<<g, BLOCK, 0, s>>>(p, out);                      \
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

// ------------------------------------------------------------ conversions
struct Weights {
  double w[19];
};

// A reference-layout FP64 field (q x n, n = nrow * nx) into the storage
// layout (row-interleaved, rows of q runs of xp values).
template <int ST>
__global__ __launch_bounds__(BLOCK) void store_from_f64_kernel(long long n, int q, int nx, int xp, const double* src,
                                                               typename Storage<ST>::T* dst, Weights w,
                                                               double shift_scale) {
  using S = Storage<ST>;
  using C = typename S::C;
  const long long k = (long long)blockIdx.x * BLOCK + threadIdx.x;
  if (k >= n * q) return;
  const int i = (int)(k / n);
  const long long c = k - (long long)i * n;
  const long long row = c / nx;
  const long long x = c - row * nx;
  S::store(dst, (size_t)((row * q + i) * xp + x), (C)src[k], (C)(w.w[i] * shift_scale));
}

hipError_t launch_store_from_f64(int st, int q, long long n, int nx, int xp, const double* src, void* dst,
                                 double shift_scale, hipStream_t s) {
  const dim3 g((unsigned)((n * q + BLOCK - 1) / BLOCK));
  Weights w{};
  for (int i = 0; i < q; ++i) w.w[i] = q == 9 ? D2Q9::W[i] : D3Q19::W[i];
  switch (st) {
    case ST_F64:
      store_from_f64_kernel<ST_F64><<<g, BLOCK, 0, s>>>(n, q, nx, xp, src, (double*)dst, w, shift_scale);
      break;
    case ST_F32:
      store_from_f64_kernel<ST_F32><<<g, BLOCK, 0, s>>>(n, q, nx, xp, src, (float*)dst, w, shift_scale);
      break;
    case ST_F16S:
      store_from_f64_kernel<ST_F16S><<<g, BLOCK, 0, s>>>(n, q, nx, xp, src, (__half*)dst, w, shift_scale);
      break;
    case ST_BF16S:
      store_from_f64_kernel<ST_BF16S><<<g, BLOCK, 0, s>>>(n, q, nx, xp, src, (__hip_bfloat16*)dst, w, shift_scale);
      break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

__global__ __launch_bounds__(BLOCK) void f64_to_f32_kernel(long long count, const double* src, float* dst) {
  const long long k = (long long)blockIdx.x * BLOCK + threadIdx.x;
  if (k < count) dst[k] = (float)src[k];
}

hipError_t launch_ibm_field(int q, int st, long long n, void* F, const uint8_t* flags, double coeff, hipStream_t s) {
  const dim3 g((unsigned)((n + BLOCK - 1) / BLOCK));
  if (st == ST_F64) {
    if (q == 19) ibm_field_kernel<D3Q19, double><<<g, BLOCK, 0, s>>>(n, (double*)F, flags, coeff);
    else ibm_field_kernel<D2Q9, double><<<g, BLOCK, 0, s>>>(n, (double*)F, flags, coeff);
  } else {
    if (q == 19) ibm_field_kernel<D3Q19, float><<<g, BLOCK, 0, s>>>(n, (float*)F, flags, coeff);
    else ibm_field_kernel<D2Q9, float><<<g, BLOCK, 0, s>>>(n, (float*)F, flags, coeff);
  }
  return hipGetLastError();
}

hipError_t launch_compute_from_f64(int st, long long count, const double* src, void* dst, hipStream_t s) {
  if (st == ST_F64) {
    return hipMemcpyAsync(dst, src, sizeof(double) * count, hipMemcpyDeviceToDevice, s);
  }
  f64_to_f32_kernel<<<dim3((unsigned)((count + BLOCK - 1) / BLOCK)), BLOCK, 0, s>>>(count, src, (float*)dst);
  return hipGetLastError();
}

// f = feq(rho, u) in FP64 over the reference layout (Q x N), as
// compute_feq / compute_feq_3d (aero/lbm/d2q9.py:66-89, d3q19.py:105-129):
// every product and sum is rounded in the reference's order and FMA
// contraction is off, so the field is bit-identical to the host's numpy
// version.  Null field pointers stand for the constants rho0, ux0, 0, 0.
template <class L>
__global__ __launch_bounds__(BLOCK) void feq_fill_kernel(long long n, double* __restrict__ f,
                                                         const double* __restrict__ rho, double rho0,
                                                         const double* __restrict__ ux, double ux0,
                                                         const double* __restrict__ uy,
                                                         const double* __restrict__ uz) {
#pragma clang fp contract(off)
  const long long c = (long long)blockIdx.x * BLOCK + threadIdx.x;
  if (c >= n) return;
  const double r = rho ? rho[c] : rho0;
  const double u = ux ? ux[c] : ux0;
  const double v = uy ? uy[c] : 0.0;
  const double w = uz ? uz[c] : 0.0;
  double usq = u * u + v * v;
  if (L::DIM == 3) usq = usq + w * w;
#pragma unroll
  for (int i = 0; i < L::Q; ++i) {
    double eu = (double)L::EX[i] * u + (double)L::EY[i] * v;
    if (L::DIM == 3) eu = eu + (double)L::EZ[i] * w;
    f[(size_t)i * n + c] = (L::W[i] * r) * (1.0 + 3.0 * eu + 4.5 * eu * eu - 1.5 * usq);
  }
}

hipError_t launch_feq_fill(int q, long long n, double* f, const double* rho, double rho0, const double* ux,
                           double ux0, const double* uy, const double* uz, hipStream_t s) {
  const dim3 g((unsigned)((n + BLOCK - 1) / BLOCK));
  if (q == 19)
    feq_fill_kernel<D3Q19><<<g, BLOCK, 0, s>>>(n, f, rho, rho0, ux, ux0, uy, uz);
  else
    feq_fill_kernel<D2Q9><<<g, BLOCK, 0, s>>>(n, f, rho, rho0, ux, ux0, uy, uz);
  return hipGetLastError();
}

// Velocity of an explicit field (rho > 0 guard, solids not zeroed: les.py:510).
template <class C, int Q>
__global__ __launch_bounds__(BLOCK) void macro_kernel(long long n, const C* F, C* U) {
  const long long c = (long long)blockIdx.x * BLOCK + threadIdx.x;
  if (c >= n) return;
  using L = typename std::conditional<Q == 9, D2Q9, D3Q19>::type;
  C rho = 0, mx = 0, my = 0, mz = 0;
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    const C v = F[(size_t)i * n + c];
    rho += v;
    mx += C(L::EX[i]) * v;
    my += C(L::EY[i]) * v;
    mz += C(L::EZ[i]) * v;
  }
  const C inv = rho > C(0) ? C(1) / rho : C(0);
  U[c] = inv * mx;
  U[n + c] = inv * my;
  U[2 * n + c] = inv * mz;
}

hipError_t launch_macro_from_compute(int st, int q, long long n, const void* F, void* U, hipStream_t s) {
  const dim3 g((unsigned)((n + BLOCK - 1) / BLOCK));
  if (st == ST_F64) {
    if (q == 9) macro_kernel<double, 9><<<g, BLOCK, 0, s>>>(n, (const double*)F, (double*)U);
    else macro_kernel<double, 19><<<g, BLOCK, 0, s>>>(n, (const double*)F, (double*)U);
  } else {
    if (q == 9) macro_kernel<float, 9><<<g, BLOCK, 0, s>>>(n, (const float*)F, (float*)U);
    else macro_kernel<float, 19><<<g, BLOCK, 0, s>>>(n, (const float*)F, (float*)U);
  }
  return hipGetLastError();
}

// rho, ux, uy, uz of an f64 field (d2q9.py:44-63: rho > 0 guard).
template <int Q>
__global__ __launch_bounds__(BLOCK) void macro4_kernel(long long n, const double* F, double* out) {
  const long long c = (long long)blockIdx.x * BLOCK + threadIdx.x;
  if (c >= n) return;
  using L = typename std::conditional<Q == 9, D2Q9, D3Q19>::type;
  double rho = 0, mx = 0, my = 0, mz = 0;
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    const double v = F[(size_t)i * n + c];
    rho += v;
    mx += L::EX[i] * v;
    my += L::EY[i] * v;
    mz += L::EZ[i] * v;
  }
  const double inv = rho > 0.0 ? 1.0 / rho : 0.0;
  out[c] = rho;
  out[n + c] = inv * mx;
  out[2 * n + c] = inv * my;
  out[3 * n + c] = inv * mz;
}

hipError_t launch_macro4_f64(int q, long long n, const double* F, double* out, hipStream_t s) {
  const dim3 g((unsigned)((n + BLOCK - 1) / BLOCK));
  if (q == 9) macro4_kernel<9><<<g, BLOCK, 0, s>>>(n, F, out);
  else macro4_kernel<19><<<g, BLOCK, 0, s>>>(n, F, out);
  return hipGetLastError();
}

// Per-block partials of check_stability (diagnostics.py:32-71) on an f64
// field, optionally restricted to one z plane (3D midplane, solver3d.py:608-615).
template <int Q>
__global__ __launch_bounds__(BLOCK) void check_f64_kernel(int nx, int ny, int nz, int zc, const double* F,
                                                          double* part) {
  using L = typename std::conditional<Q == 9, D2Q9, D3Q19>::type;
  __shared__ double red[BLOCK / 64][NCHECK];
  const long long n = (long long)nx * ny * nz;
  const long long c = (long long)blockIdx.x * BLOCK + threadIdx.x;
  double v[NCHECK] = {0, 0, 0, 0, 0};
  if (c < n && (zc < 0 || (int)(c / ((long long)nx * ny)) == zc)) {
    double rho = 0, mx = 0, my = 0;
    bool nan = false;
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      const double f = F[(size_t)i * n + c];
      nan |= isnan(f);
      rho += f;
      mx += L::EX[i] * f;
      my += L::EY[i] * f;
    }
    const double inv = rho > 0.0 ? 1.0 / rho : 0.0;
    const double ux = inv * mx, uy = inv * my;
    const double um = sqrt(ux * ux + uy * uy);
    v[0] = nan ? 1.0 : 0.0;
    v[1] = rho <= 0.0 ? 1.0 : 0.0;
    v[2] = isnan(um) ? 0.0 : um;
    v[3] = rho;
    v[4] = 1.0;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    v[0] = fmax(v[0], __shfl_xor(v[0], off));
    v[1] = fmax(v[1], __shfl_xor(v[1], off));
    v[2] = fmax(v[2], __shfl_xor(v[2], off));
    v[3] += __shfl_xor(v[3], off);
    v[4] += __shfl_xor(v[4], off);
  }
  if ((threadIdx.x & 63) == 0) {
#pragma unroll
    for (int k = 0; k < NCHECK; ++k) red[threadIdx.x >> 6][k] = v[k];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double r[NCHECK] = {0, 0, 0, 0, 0};
    for (int w = 0; w < BLOCK / 64; ++w) {
      r[0] = fmax(r[0], red[w][0]);
      r[1] = fmax(r[1], red[w][1]);
      r[2] = fmax(r[2], red[w][2]);
      r[3] += red[w][3];
      r[4] += red[w][4];
    }
#pragma unroll
    for (int k = 0; k < NCHECK; ++k) part[(size_t)blockIdx.x * NCHECK + k] = r[k];
  }
}

hipError_t launch_check_f64(int q, int nx, int ny, int nz, int zc, const double* F, int nblocks, double* part,
                            hipStream_t s) {
  if (q == 9) check_f64_kernel<9><<<nblocks, BLOCK, 0, s>>>(nx, ny, nz, zc, F, part);
  else check_f64_kernel<19><<<nblocks, BLOCK, 0, s>>>(nx, ny, nz, zc, F, part);
  return hipGetLastError();
}

// ------------------------------------------------------------- reductions
// Momentum-exchange totals of one step in the same fixed order the step
// kernel's block 0 uses (reduce_forces_block), for the last step of a chunk.
template <class L>
__global__ __launch_bounds__(BLOCK) void reduce_forces_kernel(ForceReduce r) { reduce_forces_block<L, 8>(r); }

// The NFORCE terms of every link cell of the cases with a terms buffer, one
// thread per cell (case = blockIdx.y), ahead of the one-block reduction.
template <class L>
__global__ __launch_bounds__(BLOCK) void force_terms_kernel(const ForceReduce* rs) {
  const ForceReduce& r = rs[blockIdx.y];
  const int k = blockIdx.x * BLOCK + threadIdx.x;
  if (!r.terms || k >= r.nrec) return;
  double t[NFORCE];
  const LinkCell lc = r.lrec[k];
  cell_force_terms<L>(r.part + (size_t)k * NPART, lc.cell, lc.mask, r.nx, r.ny, r.cx, r.cy, r.cz, t);
#pragma unroll
  for (int j = 0; j < NFORCE; ++j) r.terms[(size_t)k * NFORCE + j] = t[j];
}

template <class L>
__global__ __launch_bounds__(BLOCK) void reduce_forces_many_kernel(const ForceReduce* rs) {
  reduce_forces_block<L, 8, true>(rs[blockIdx.x]);
}

hipError_t launch_reduce_forces_many(int q, const ForceReduce* rs, int n, int max_terms, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  if (max_terms > 0) {
    const dim3 g((unsigned)((max_terms + BLOCK - 1) / BLOCK), (unsigned)n);
    if (q == 9) force_terms_kernel<D2Q9><<<g, BLOCK, 0, s>>>(rs);
    else force_terms_kernel<D3Q19><<<g, BLOCK, 0, s>>>(rs);
  }
  if (q == 9) reduce_forces_many_kernel<D2Q9><<<n, BLOCK, 0, s>>>(rs);
  else reduce_forces_many_kernel<D3Q19><<<n, BLOCK, 0, s>>>(rs);
  return hipGetLastError();
}

__global__ __launch_bounds__(BLOCK) void rows_to_host_kernel(const double* __restrict__ rows, double* host, int n,
                                                             int ch, int nk) {
  const long long per = (long long)nk * NCOEF;
  const long long e = (long long)blockIdx.x * BLOCK + threadIdx.x;
  if (e >= per * n) return;
  const long long i = e / per, r = e - i * per;
  host[i * ch * NCOEF + r] = rows[i * ch * NCOEF + r];
}

hipError_t launch_rows_to_host(const double* rows, double* host, int n, int ch, int nk, hipStream_t s) {
  const long long total = (long long)n * nk * NCOEF;
  if (total <= 0) return hipSuccess;
  rows_to_host_kernel<<<(unsigned)((total + BLOCK - 1) / BLOCK), BLOCK, 0, s>>>(rows, host, n, ch, nk);
  return hipGetLastError();
}

// ------------------------------------------------------------ halo faces
// The D3Q19 populations that cross a z face (e_z = dir), of one z plane of
// the row-interleaved storage, to or from a compact buffer
// [5][ny][nx] of raw storage elements: what one z-slab sends its neighbour
// each step (the neighbour's pull at its first / last plane reads them).
template <class E, bool PACK>
__global__ __launch_bounds__(BLOCK) void halo_kernel(E* __restrict__ base, E* __restrict__ buf, int nx, int ny,
                                                     int xp, int z, int dir) {
  constexpr int UP[5] = {5, 11, 12, 15, 16};    // e_z = +1
  constexpr int DOWN[5] = {6, 13, 14, 17, 18};  // e_z = -1
  const long long plane = (long long)nx * ny;
  const long long e = (long long)blockIdx.x * BLOCK + threadIdx.x;
  if (e >= 5 * plane) return;
  const int k = (int)(e / plane);
  const long long r = e - (long long)k * plane;
  const int y = (int)(r / nx), x = (int)(r - (long long)y * nx);
  int i = 0;
#pragma unroll
  for (int j = 0; j < 5; ++j)
    if (j == k) i = dir > 0 ? UP[j] : DOWN[j];
  E* cell = base + (((size_t)z * ny + y) * 19 + i) * (size_t)xp + x;
  if (PACK) buf[e] = *cell;
  else *cell = buf[e];
}

template <bool PACK>
static hipError_t halo_n(int esz, void* base, void* buf, int nx, int ny, int xp, int z, int dir, hipStream_t s) {
  const long long n = 5ll * nx * ny;
  const dim3 g((unsigned)((n + BLOCK - 1) / BLOCK));
  if (esz == 8) halo_kernel<uint64_t, PACK><<<g, BLOCK, 0, s>>>((uint64_t*)base, (uint64_t*)buf, nx, ny, xp, z, dir);
  else if (esz == 4) halo_kernel<uint32_t, PACK><<<g, BLOCK, 0, s>>>((uint32_t*)base, (uint32_t*)buf, nx, ny, xp, z, dir);
  else halo_kernel<uint16_t, PACK><<<g, BLOCK, 0, s>>>((uint16_t*)base, (uint16_t*)buf, nx, ny, xp, z, dir);
  return hipGetLastError();
}

hipError_t launch_halo(bool pack, int esz, void* base, void* buf, int nx, int ny, int xp, int z, int dir,
                       hipStream_t s) {
  return pack ? halo_n<true>(esz, base, buf, nx, ny, xp, z, dir, s) : halo_n<false>(esz, base, buf, nx, ny, xp, z, dir, s);
}

hipError_t launch_reduce_forces(const ForceReduce& r, hipStream_t s) {
  if (r.q == 9) reduce_forces_kernel<D2Q9><<<1, BLOCK, 0, s>>>(r);
  else reduce_forces_kernel<D3Q19><<<1, BLOCK, 0, s>>>(r);
  return hipGetLastError();
}

__device__ __forceinline__ void reduce_check_block(const double* part, int nblocks, double* out,
                                                   double (*red)[BLOCK]) {
  double a[NCHECK] = {0, 0, 0, 0, 0};
  for (int b = threadIdx.x; b < nblocks; b += BLOCK) {
    a[0] = fmax(a[0], part[(size_t)b * NCHECK + 0]);
    a[1] = fmax(a[1], part[(size_t)b * NCHECK + 1]);
    a[2] = fmax(a[2], part[(size_t)b * NCHECK + 2]);
    a[3] += part[(size_t)b * NCHECK + 3];
    a[4] += part[(size_t)b * NCHECK + 4];
  }
#pragma unroll
  for (int k = 0; k < NCHECK; ++k) red[k][threadIdx.x] = a[k];
  __syncthreads();
  for (int w = BLOCK / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      red[0][threadIdx.x] = fmax(red[0][threadIdx.x], red[0][threadIdx.x + w]);
      red[1][threadIdx.x] = fmax(red[1][threadIdx.x], red[1][threadIdx.x + w]);
      red[2][threadIdx.x] = fmax(red[2][threadIdx.x], red[2][threadIdx.x + w]);
      red[3][threadIdx.x] += red[3][threadIdx.x + w];
      red[4][threadIdx.x] += red[4][threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NCHECK; ++k) out[k] = red[k][0];
  }
}

__global__ __launch_bounds__(BLOCK) void reduce_check_kernel(const double* part, int nblocks, double* out) {
  __shared__ double red[NCHECK][BLOCK];
  reduce_check_block(part, nblocks, out, red);
}

// One block per case: the same fixed-order reduction of the case's partials.
__global__ __launch_bounds__(BLOCK) void reduce_check_batch_kernel(const StepParams* ps, double* out) {
  __shared__ double red[NCHECK][BLOCK];
  const StepParams& p = ps[blockIdx.x];
  reduce_check_block(p.check_part, p.nblocks, out + (size_t)NCHECK * blockIdx.x, red);
}

hipError_t launch_reduce_check(const double* part, int nblocks, double* out, hipStream_t s) {
  reduce_check_kernel<<<1, BLOCK, 0, s>>>(part, nblocks, out);
  return hipGetLastError();
}

hipError_t launch_reduce_check_batch(const StepParams* ps, int n, double* out, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  reduce_check_batch_kernel<<<(unsigned)n, BLOCK, 0, s>>>(ps, out);
  return hipGetLastError();
}

}  // namespace aero
