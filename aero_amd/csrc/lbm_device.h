// Device-side building blocks of the MI355X lattice-Boltzmann step.
//
// Populations live in HBM row-interleaved, f[z][y][q][xp] with x fastest and
// rows padded to xp (a multiple of 64; see padded_row below), so one
// wavefront of 64 lanes reads 64 consecutive cells of one direction of one
// row and every access is coalesced.  The host converts from and to the
// reference's (Q, Nz, Ny, Nx) layout (aero/lbm/solver3d.py:239-247).
#pragma once

#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace aero {

enum Collision { COLL_BGK = 0, COLL_TRT = 1, COLL_MRT = 2 };
enum StorageKind { ST_F64 = 0, ST_F32 = 1, ST_F16S = 2, ST_BF16S = 3 };
enum InletMode {
  IN_NONE = 0,
  IN_ZOUHE = 1,      // velocity Zou-He, optional travelling sine
  IN_PRESSURE = 2,   // 2D Zou-He pressure
  IN_FIELD = 3,      // per-row / per-cell target velocity planes
  IN_SEM = 4,        // 3D SEM fluctuation planes
  IN_RECYCLING = 5,  // 3D recycling-rescaling from plane x = nx-2
};
enum OutletMode { OUT_NONE = 0, OUT_CONVECTIVE = 1, OUT_ZEROGRAD = 2, OUT_PRESSURE = 3 };
enum WallMode { WALL_NONE = 0, WALL_SLIP = 1, WALL_NOSLIP = 2, WALL_MOVING = 3 };
// LES_FD_*: the reference's finite-difference models (split step); LES_LOCAL_SMAG:
// Smagorinsky from the local non-equilibrium stress, fused into the step.
enum LesMode { LES_OFF = 0, LES_FD_SMAG = 1, LES_FD_WALE = 2, LES_LOCAL_SMAG = 3 };
// Per-cell flags: solid, or a fluid cell with at least one surface link.
enum CellFlag { FLAG_SOLID = 1, FLAG_LINKS = 2, FLAG_IBM = 4 };  // FLAG_IBM: fluid cell with 0 < phi <= 1.5

// ------------------------------------------------------------------ lattices
// D2Q9: aero/lbm/d2q9.py:18-41.  D3Q19: aero/lbm/d3q19.py:36-73.
struct D2Q9 {
  static constexpr int Q = 9;
  static constexpr int DIM = 2;
  static constexpr int EX[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
  static constexpr int EY[9] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
  static constexpr int EZ[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  static constexpr int OPP[9] = {0, 3, 4, 1, 2, 7, 8, 5, 6};
  static constexpr double W[9] = {4.0 / 9, 1.0 / 9, 1.0 / 9, 1.0 / 9, 1.0 / 9,
                                  1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36};
  // MRT basis rows (aero/lbm/mrt2d.py:36-46), exact small integers
  static constexpr int M[9][9] = {
      {1, 1, 1, 1, 1, 1, 1, 1, 1},     {-4, -1, -1, -1, -1, 2, 2, 2, 2},
      {4, -2, -2, -2, -2, 1, 1, 1, 1}, {0, 1, 0, -1, 0, 1, -1, -1, 1},
      {0, -2, 0, 2, 0, 1, -1, -1, 1},  {0, 0, 1, 0, -1, 1, 1, -1, -1},
      {0, 0, -2, 0, 2, 1, 1, -1, -1},  {0, 1, -1, 1, -1, 0, 0, 0, 0},
      {0, 0, 0, 0, 0, 1, -1, 1, -1}};
  // wall pair tables (dst <- src); boundary.py:162-226
  static constexpr int NWALL = 3;
  static constexpr int SLIP_BOT[3][2] = {{2, 4}, {5, 8}, {6, 7}};
  static constexpr int SLIP_TOP[3][2] = {{4, 2}, {8, 5}, {7, 6}};
  static constexpr int NOSLIP_BOT[3][2] = {{2, 4}, {5, 7}, {6, 8}};
  static constexpr int NOSLIP_TOP[3][2] = {{4, 2}, {7, 5}, {8, 6}};
  static constexpr int MOVE_SGN_BOT[3] = {0, 1, -1};
  static constexpr int MOVE_SGN_TOP[3] = {0, -1, 1};
};

struct D3Q19 {
  static constexpr int Q = 19;
  static constexpr int DIM = 3;
  static constexpr int EX[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
  static constexpr int EY[19] = {0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1};
  static constexpr int EZ[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1};
  static constexpr int OPP[19] = {0, 2, 1, 4, 3, 6, 5, 10, 9, 8, 7, 14, 13, 12, 11, 18, 17, 16, 15};
  static constexpr double W[19] = {1.0 / 3,  1.0 / 18, 1.0 / 18, 1.0 / 18, 1.0 / 18,
                                   1.0 / 18, 1.0 / 18, 1.0 / 36, 1.0 / 36, 1.0 / 36,
                                   1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36,
                                   1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36};
  // d'Humieres-style basis as the reference writes it (aero/lbm/mrt3d.py:46-87);
  // rank 16, so the inverse used with it is the host-computed pseudo-inverse.
  static constexpr int M[19][19] = {
      {1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1},
      {-30, -11, -11, -11, -11, -11, -11, 8, 8, 8, 8, 8, 8, 8, 8, 8, 8, 8, 8},
      {12, -4, -4, -4, -4, -4, -4, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1},
      {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0},
      {0, -4, 4, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0},
      {0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1},
      {0, 0, 0, -4, 4, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1},
      {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1},
      {0, 0, 0, 0, 0, -4, 4, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1},
      {0, 2, 2, -1, -1, -1, -1, 1, 1, 1, 1, 1, 1, 1, 1, -2, -2, -2, -2},
      {0, 0, 0, 1, 1, -1, -1, 1, 1, 1, 1, -1, -1, -1, -1, 0, 0, 0, 0},
      {0, 0, 0, 0, 0, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 0, 0, 0, 0},
      {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, -1, -1, 1},
      {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0},
      {0, -2, 2, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0},
      {0, 0, 0, -2, 2, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1},
      {0, 0, 0, 0, 0, -2, 2, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1},
      {0, 0, 0, 0, 0, 0, 0, 1, -1, 1, -1, -1, 1, -1, 1, 0, 0, 0, 0},
      {0, 0, 0, 0, 0, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, -1, 1, -1, 1}};
  // Entries of the pseudo-inverse of M that are nonzero beyond rounding
  // (|pinv(M)| >= 3.3e-3; the other 145 are below 6.5e-16 in the float64
  // pseudo-inverse): the inverse transform skips the rest.
  static constexpr bool PINV_NZ[19][19] = {
      {1, 1, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0},
      {1, 1, 1, 1, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0},
      {1, 1, 1, 1, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0},
      {1, 1, 1, 0, 0, 1, 1, 0, 0, 1, 1, 0, 0, 0, 0, 1, 0, 0, 0},
      {1, 1, 1, 0, 0, 1, 1, 0, 0, 1, 1, 0, 0, 0, 0, 1, 0, 0, 0},
      {1, 1, 1, 0, 0, 0, 0, 1, 1, 1, 1, 0, 0, 0, 0, 0, 1, 0, 0},
      {1, 1, 1, 0, 0, 0, 0, 1, 1, 1, 1, 0, 0, 0, 0, 0, 1, 0, 0},
      {1, 1, 1, 1, 1, 1, 1, 0, 0, 1, 1, 1, 0, 0, 1, 1, 0, 1, 1},
      {1, 1, 1, 1, 1, 1, 1, 0, 0, 1, 1, 1, 0, 0, 1, 1, 0, 1, 1},
      {1, 1, 1, 1, 1, 1, 1, 0, 0, 1, 1, 1, 0, 0, 1, 1, 0, 1, 1},
      {1, 1, 1, 1, 1, 1, 1, 0, 0, 1, 1, 1, 0, 0, 1, 1, 0, 1, 1},
      {1, 1, 1, 1, 1, 0, 0, 1, 1, 1, 1, 0, 0, 1, 1, 0, 1, 1, 0},
      {1, 1, 1, 1, 1, 0, 0, 1, 1, 1, 1, 0, 0, 1, 1, 0, 1, 1, 0},
      {1, 1, 1, 1, 1, 0, 0, 1, 1, 1, 1, 0, 0, 1, 1, 0, 1, 1, 0},
      {1, 1, 1, 1, 1, 0, 0, 1, 1, 1, 1, 0, 0, 1, 1, 0, 1, 1, 0},
      {1, 1, 1, 0, 0, 1, 1, 1, 1, 1, 0, 0, 1, 0, 0, 1, 1, 0, 1},
      {1, 1, 1, 0, 0, 1, 1, 1, 1, 1, 0, 0, 1, 0, 0, 1, 1, 0, 1},
      {1, 1, 1, 0, 0, 1, 1, 1, 1, 1, 0, 0, 1, 0, 0, 1, 1, 0, 1},
      {1, 1, 1, 0, 0, 1, 1, 1, 1, 1, 0, 0, 1, 0, 0, 1, 1, 0, 1}};
  // boundary3d.py:205-272
  static constexpr int NWALL = 5;
  static constexpr int SLIP_BOT[5][2] = {{3, 4}, {7, 9}, {8, 10}, {15, 16}, {17, 18}};
  static constexpr int SLIP_TOP[5][2] = {{4, 3}, {9, 7}, {10, 8}, {16, 15}, {18, 17}};
  static constexpr int NOSLIP_BOT[5][2] = {{3, 4}, {7, 10}, {8, 9}, {15, 18}, {17, 16}};
  static constexpr int NOSLIP_TOP[5][2] = {{4, 3}, {10, 7}, {9, 8}, {18, 15}, {16, 17}};
  static constexpr int MOVE_SGN_BOT[5] = {0, 1, -1, 0, 0};
  static constexpr int MOVE_SGN_TOP[5] = {0, -1, 1, 0, 0};
};

// acc + e*v for a lattice component e in {-1, 0, 1}: with e a constant of an
// unrolled loop the zero terms vanish at compile time (0*v cannot be folded
// in IEEE arithmetic, so writing e*v would keep a multiply-add per zero).
template <class T>
__host__ __device__ __forceinline__ T madd(T acc, int e, T v) {
  return e == 0 ? acc : (e == 1 ? acc + v : (e == -1 ? acc - v : acc + T(e) * v));
}
// e*v for a lattice component: the first term of a sum, so no 0 + x add is
// left for the compiler (it may not drop one in IEEE arithmetic).
template <class T>
__host__ __device__ __forceinline__ T mfirst(int e, T v) {
  return e == 1 ? v : (e == -1 ? -v : T(e) * v);
}
template <class L, class T>
__host__ __device__ __forceinline__ T edot(int i, T ux, T uy, T uz) {
  if (L::EX[i] != 0) return madd(madd(mfirst(L::EX[i], ux), L::EY[i], uy), L::EZ[i], uz);
  if (L::EY[i] != 0) return madd(mfirst(L::EY[i], uy), L::EZ[i], uz);
  return L::EZ[i] != 0 ? mfirst(L::EZ[i], uz) : T(0);
}
// Density and momentum of one cell's populations; every sum starts from its
// first nonzero term.
template <class L, class T>
__host__ __device__ __forceinline__ void moments(const T f[L::Q], T& rho, T& mx, T& my, T& mz) {
  rho = f[0];
  bool sx = false, sy = false, sz = false;
  mx = my = mz = T(0);
#pragma unroll
  for (int i = 1; i < L::Q; ++i) {
    rho += f[i];
    if (L::EX[i] != 0) { mx = sx ? madd(mx, L::EX[i], f[i]) : mfirst(L::EX[i], f[i]); sx = true; }
    if (L::EY[i] != 0) { my = sy ? madd(my, L::EY[i], f[i]) : mfirst(L::EY[i], f[i]); sy = true; }
    if (L::EZ[i] != 0) { mz = sz ? madd(mz, L::EZ[i], f[i]) : mfirst(L::EZ[i], f[i]); sz = true; }
  }
}

// Division of a cell index (below 2^31) by a grid extent d >= 1 as a
// multiply-high and a shift (Granlund & Montgomery, "Division by invariant
// integers using multiplication", 1994): with l = ceil(log2 d) and
// m = ceil(2^(31+l) / d) < 2^32, floor(n/d) = floor(n m / 2^(31+l)) for all
// n < 2^31.  d = 1 is flagged by m = 0.
struct DivMagic {
  unsigned m;
  unsigned s;  // l - 1
};
inline DivMagic make_div_magic(unsigned d) {
  if (d <= 1) return DivMagic{0u, 0u};
  unsigned l = 0;
  while ((1ull << l) < d) ++l;
  const unsigned long long m = ((1ull << (31 + l)) + d - 1) / d;
  return DivMagic{(unsigned)m, l - 1};
}
__device__ __forceinline__ unsigned udiv(unsigned n, DivMagic dm) {
  return dm.m ? (__umulhi(n, dm.m) >> dm.s) : n;
}

// Direction components for a runtime index (loops over the set bits of a
// link mask): two bits per direction, value + 1.
template <class L>
struct DirPack {
  static constexpr unsigned long long pack(const int* e) {
    unsigned long long v = 0;
    for (int i = 0; i < L::Q; ++i) v |= (unsigned long long)(e[i] + 1) << (2 * i);
    return v;
  }
  static constexpr unsigned long long PX = pack(L::EX), PY = pack(L::EY), PZ = pack(L::EZ);
  static __host__ __device__ __forceinline__ int ex(int i) { return (int)((PX >> (2 * i)) & 3ull) - 1; }
  static __host__ __device__ __forceinline__ int ey(int i) { return (int)((PY >> (2 * i)) & 3ull) - 1; }
  static __host__ __device__ __forceinline__ int ez(int i) { return (int)((PZ >> (2 * i)) & 3ull) - 1; }
  // weights by shell: rest, axis (|e| = 1), diagonal
  static __host__ __device__ __forceinline__ double w(int i) {
    return i == 0 ? L::W[0] : (i <= 2 * L::DIM ? L::W[1] : L::W[L::Q - 1]);
  }
};

// Storage shift of the half-precision forms for population i of f: the
// free-stream equilibrium feq_i(rho = 1, u = (u0, 0, 0)), so the stored
// 10- or 7-bit-mantissa value is the (small) deviation from the free stream
// wherever the flow is near it.  The scalar g shifts by w_i * T_ref.
// Its value depends only on the weight shell (rest, axis, diagonal) and e_x
// of the direction, so the kernels keep one value per class: 9 classes.
constexpr int NSHIFT = 9;
template <class L>
__host__ __device__ constexpr int fsh_class(int i) {
  return (i == 0 ? 0 : (L::W[i] == L::W[1] ? 1 : 2)) * 3 + L::EX[i] + 1;
}
template <class L, class C>
__host__ __device__ __forceinline__ C fshift(int i, C u0) {
  const C a = C(1) - C(1.5) * u0 * u0;
  const int ex = L::EX[i];
  return C(L::W[i]) * (ex == 0 ? a : a + C(4.5) * u0 * u0 + (ex > 0 ? C(3) * u0 : C(-3) * u0));
}

// ------------------------------------------------------------------ storage
// Populations are stored as T and computed in C.  The half forms store the
// shifted value f - w_i*s (s = 1 for f, T_ref for the scalar g) so the 10/7
// mantissa bits resolve the deviation from rest equilibrium, not its bulk.
template <int ST> struct Storage;
// ld/st address one population run of a storage row with a 32-bit byte
// offset from the run's base (an x offset, below the padded row length), so
// the compiler forms scalar-base + 32-bit-offset addresses: one VGPR per
// access instead of a 64-bit pair.
#define AERO_PLANE_ACCESS                                                                    \
  static __device__ __forceinline__ C ld(const T* plane, unsigned k, C s) {                 \
    return load((const T*)((const char*)plane + k * (unsigned)sizeof(T)), 0, s);            \
  }                                                                                          \
  /* the stored value, unconverted; cvt(raw(plane, k), s) == ld(plane, k, s) */              \
  static __device__ __forceinline__ T raw(const T* plane, unsigned k) {                     \
    return *(const T*)((const char*)plane + k * (unsigned)sizeof(T));                       \
  }                                                                                          \
  static __device__ __forceinline__ void st(T* plane, unsigned k, C v, C s) {               \
    store((T*)((char*)plane + k * (unsigned)sizeof(T)), 0, v, s);                           \
  }                                                                                          \
  static __device__ __forceinline__ void st_nt(T* plane, unsigned k, C v, C s) {            \
    store_nt((T*)((char*)plane + k * (unsigned)sizeof(T)), v, s);                           \
  }
template <> struct Storage<ST_F64> {
  using T = double;
  using C = double;
  static __device__ __forceinline__ C load(const T* p, size_t k, C) { return p[k]; }
  static __device__ __forceinline__ C cvt(T v, C) { return v; }
  static __device__ __forceinline__ void store(T* p, size_t k, C v, C) { p[k] = v; }
  static __device__ __forceinline__ void store_nt(T* q, C v, C) { __builtin_nontemporal_store(v, q); }
  AERO_PLANE_ACCESS
};
template <> struct Storage<ST_F32> {
  using T = float;
  using C = float;
  static __device__ __forceinline__ C load(const T* p, size_t k, C) { return p[k]; }
  static __device__ __forceinline__ C cvt(T v, C) { return v; }
  static __device__ __forceinline__ void store(T* p, size_t k, C v, C) { p[k] = v; }
  static __device__ __forceinline__ void store_nt(T* q, C v, C) { __builtin_nontemporal_store(v, q); }
  AERO_PLANE_ACCESS
};
// FP16: the conversion and the shift are one mixed-precision FMA each way
// (v_fma_mix_f32 on load, v_fma_mixlo_f16 on store).  The unit multiplier
// is held opaque in a scalar register: with a literal 1.0 the compiler folds
// the FMA back into a separate add and convert.  The load is exact-then-
// one-rounding as before; the store rounds v - s once, straight to FP16.
template <> struct Storage<ST_F16S> {
  using T = __half;
  using C = float;
  static __device__ __forceinline__ float one() {
    float o = 1.0f;
    asm("" : "+s"(o));
    return o;
  }
  static __device__ __forceinline__ C load(const T* p, size_t k, C s) {
    return __builtin_fmaf(__half2float(p[k]), one(), s);
  }
  static __device__ __forceinline__ C cvt(T v, C s) { return __builtin_fmaf(__half2float(v), one(), s); }
  static __device__ __forceinline__ void store(T* p, size_t k, C v, C s) {
    p[k] = __float2half(__builtin_fmaf(v, one(), -s));
  }
  static __device__ __forceinline__ void store_nt(T* q, C v, C s) {
    __builtin_nontemporal_store(__half_as_ushort(__float2half(__builtin_fmaf(v, one(), -s))), (unsigned short*)q);
  }
  AERO_PLANE_ACCESS
};
// Stochastic rounding to BF16.  With an 8-bit significand, round-to-nearest
// stalls: a relaxation step smaller than half a unit in the last place leaves
// a population unchanged, step after step, and the errors pile up in one
// direction (the 128 x 64 x 64 sphere's drag +1.07% after 5000 steps).  Adding
// 16 uniform random bits below the BF16 cut and truncating rounds up with
// probability equal to the discarded fraction, so the error has zero mean
// (the same drag then within 0.002%).  The bits come from an integer hash of
// (cell, step number, population), so a case computes the same bits alone, in
// a batch or chunked differently: a hash of the cell and step seeds a 32-bit
// linear congruential sequence (Numerical Recipes' constants) whose high 16
// bits serve the cell's populations in turn.  mix32 is "lowbias32" from
// C. Wellons' hash prospector.
__host__ __device__ inline unsigned mix32(unsigned x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}
__host__ __device__ inline unsigned sr_seed(int cell, unsigned step) {
  return mix32((unsigned)cell * 0x9e3779b1u ^ mix32(step));
}
// the next state of the sequence; its high 16 bits are the random bits
__host__ __device__ inline unsigned sr_next(unsigned h) { return h * 1664525u + 1013904223u; }
// v rounded to BF16 with the random bits in the high half of r: gfx950's
// v_cvt_sr_bf16_f32, which (checked bit for bit over normal, subnormal,
// near-overflow, infinite and NaN inputs) returns (bits(v) + (r >> 16)) >> 16,
// NaN kept NaN.
__device__ inline unsigned short bf16_round_sr(float v, unsigned r) {
  typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
  bf16x2_t o = {};
  o = __builtin_amdgcn_cvt_sr_bf16_f32(o, v, r, false);
  return __builtin_bit_cast(unsigned short, o[0]);
}

template <> struct Storage<ST_BF16S> {
  using T = __hip_bfloat16;
  using C = float;
  static __device__ __forceinline__ C load(const T* p, size_t k, C s) { return __bfloat162float(p[k]) + s; }
  static __device__ __forceinline__ C cvt(T v, C s) { return __bfloat162float(v) + s; }
  static __device__ __forceinline__ void store(T* p, size_t k, C v, C s) { p[k] = __float2bfloat16(v - s); }
  static __device__ __forceinline__ void store_nt(T* q, C v, C s) {
    __builtin_nontemporal_store(__bfloat16_as_ushort(__float2bfloat16(v - s)), (unsigned short*)q);
  }
  AERO_PLANE_ACCESS
};
#undef AERO_PLANE_ACCESS

// Link-cell record: one per fluid cell that has at least one surface link.
// Its q values live in a separate (Q-1)-per-record table (direction i at
// index i-1, zero where i is not a link).
struct LinkCell {
  int cell;       // linear cell index z*ny*nx + y*nx + x
  uint32_t mask;  // bit i set: direction i is a link (neighbour along e_i solid)
};

constexpr int NCOEF = 20;  // per-step history row: 8 coefficients + 12 raw sums
constexpr int NFORCE = 12; // Fx Fy Fz Mx My Mz Fpx Fpy Fpz Fvx Fvy Fvz
constexpr int NPART = 10;  // per link cell: rho, u xyz of f_pre; F_pre xyz; F_new xyz (see link_pre_terms)
constexpr int NCHECK = 5;  // nan, nonpositive rho, max|u|, sum rho, count
constexpr int BLOCK = 256;

// Pointers the step kernel streams through are declared in the global
// address space on the device side (plain pointers in host code; same
// layout), so accesses compile to global rather than flat instructions also
// when the parameters are read from a table in memory.
#if defined(__HIP_DEVICE_COMPILE__)
#define AERO_GLOBAL __attribute__((address_space(1)))
#else
#define AERO_GLOBAL
#endif

// Population storage (f and the scalar's g, in the storage type): grid row
// r = z ny + y holds its cells' Q populations as Q runs of xp values, with
// xp = nx rounded up to 64, population i of cell (x, y, z) at element
// (r Q + i) xp + x.  The fused step runs one thread per padded cell
// (nrow * xp threads), so a 64-lane wavefront always covers 64 aligned
// cells of one row: its y and z, every neighbour row's base address and the
// wall / inlet row tests are wave-uniform (scalar registers), and each lane
// adds only its 32-bit x offset.  A wave touches Q contiguous runs of one
// row instead of Q planes spread over the whole buffer, which keeps far
// fewer HBM pages open at once than plane-major [q][z][y][x] storage.
// Per-cell tables (flags, link records, output fields) stay compact,
// indexed by cell = r nx + x.
constexpr int ROW_ALIGN = 64;
__host__ __device__ constexpr int padded_row(int nx) { return (nx + ROW_ALIGN - 1) / ROW_ALIGN * ROW_ALIGN; }

// Row groups of the 2-byte D2Q9 kernels (no scalar).  Their step waves are
// listed per case (StepParams::witems): a plain item is one row's 64
// aligned columns; a group item (bit 31) covers the same columns in
// RPW_GROUP consecutive rows, none of them holding a flagged cell (the inlet
// column, a zero-gradient or pressure outlet, sponge columns and the wall
// rows may be among them).  A group wave issues all RPW_GROUP x
// 9 two-byte loads before it collides its first row: a one-row wave keeps
// only 9 x 128 bytes in flight, too few to cover HBM latency at the
// occupancy the kernel reaches (a pull + BGK probe on 8192 x 2048 ran at
// 48% of 8 TB/s with one row per wave, 71% with four).  Every other kernel
// runs one row per wave over the whole nrow * xp thread grid.
constexpr int RPW_GROUP = 4;

// MI355X hands a launch's workgroups to its 8 XCDs round-robin (workgroup i
// to XCD i mod 8).  Blocks in XCD order with runs of `run`: within each
// round of 8 run positions, XCD x walks a contiguous run of the round's
// blocks, so each XCD's L2 sees neighbouring rows.  xcd_block maps a launch
// position to its block; the host's tile lists use the same order.
constexpr int XCDS = 8;
__host__ __device__ inline int xcd_block(int pos, int run, int n) {
  if (run <= 0) return pos;
  const int round = XCDS * run;
  const int r0 = pos / round * round;
  const int m = n - r0 < round ? n - r0 : round;
  const int b = pos - r0, x = b % XCDS;
  int off = 0;
  for (int i = 0; i < x; ++i) off += m > i ? (m - 1 - i) / XCDS + 1 : 0;
  return r0 + off + b / XCDS;
}
constexpr uint32_t WITEM_GROUP = 0x80000000u;
__host__ __device__ constexpr int rows_per_wave(int q, int collision, int storage, bool thermal) {
  return (q == 9 && (storage == ST_F16S || storage == ST_BF16S) && !thermal) ? RPW_GROUP
                                                                                                      : 1;
}

// Everything one step launch needs; passed by value as the kernel argument.
struct StepParams {
  int nx, ny, nz;
  int nrow;  // nz * ny
  int xp;    // padded row length of the population storage and of the step's thread grid
  int qxp;   // Q * xp: elements of one storage row
  DivMagic div_nx, div_ny, div_xp;  // cell -> (x, y, z) and step thread -> row without integer division
  long long ncell;
  int nblocks;  // blocks covering ncell (per-cell kernels)
  int sblocks;  // blocks covering the step's waves (the step launch adds one reduction block)
  int nitems;   // grouped kernels (rows_per_wave > 1): waves listed in witems
  const AERO_GLOBAL uint32_t* witems;  // per wave: its first thread in the one-row grid / 64, | WITEM_GROUP
  int nt_store;  // population stores bypass the caches (launch working set beyond the last-level cache)
  int ibm_on;      // immersed-boundary mode: no link bounce-back; Guo forcing of FLAG_IBM cells
  int sponge_x0;   // first column with sponge sigma > 0 (nx when there is none)
  double ibm_coeff;  // 1 - 1/(2 tau)
  const AERO_GLOBAL void* fin;
  AERO_GLOBAL void* fout;
  const AERO_GLOBAL void* gin;
  AERO_GLOBAL void* gout;
  AERO_GLOBAL void* fsplit;  // split mode: post-BC populations, compute type, Q x N
  AERO_GLOBAL void* usplit;  // split mode: ux, uy, uz, compute type, 3 x N
  const AERO_GLOBAL uint8_t* flags;
  const AERO_GLOBAL uint32_t* wave_any;  // bit w: some cell of step threads [64 w, 64 w + 64) is flagged
  const AERO_GLOBAL LinkCell* lrec;
  const AERO_GLOBAL int* rec_of;   // per cell: its link record (read only where FLAG_LINKS is set)
  const AERO_GLOBAL float* qdir;   // Q-1 wall fractions per link record
  AERO_GLOBAL double* fpart;             // NPART doubles per link cell, this step
  const AERO_GLOBAL double* fpart_prev;  // the previous step's sums, reduced by the extra block
  AERO_GLOBAL double* hist_prev;         // coefficient row the previous step's sums go to
  int reduce_prev;
  int nrec;  // link cells
  double fdenom, mdenom;
  const AERO_GLOBAL double* prev_in;
  AERO_GLOBAL double* prev_out;
  const AERO_GLOBAL double* sigma;  // sponge sigma(x), nx entries, or null
  const AERO_GLOBAL double* plane;  // inlet planes, 3 x nrow
  const AERO_GLOBAL double* recyc;  // recycling: rho, ux, uy, uz per row-plane cell (4 x nrow), then scale
  const AERO_GLOBAL double* minv_d;
  const AERO_GLOBAL float* minv_f;
  double mrt_s[19];
  float fsh[NSHIFT];  // storage shift of f by fsh_class (half forms): feq_i(1, (u0, 0, 0))
  double omega, trt_sm, trt_lambda, u0, u_conv;
  // local Smagorinsky (LES_LOCAL_SMAG kernels): tau0 = 1/omega and C_s^2, or
  // per cell (C_s d)^2 with the Van Driest factor d (les_c2d for FP64
  // arithmetic, les_c2f for FP32; null without damping)
  double les_tau0, les_c2;
  const AERO_GLOBAL double* les_c2d;
  const AERO_GLOBAL float* les_c2f;
  int trt_variant;
  double inlet_amp, phase;
  double rho_in, rho_out, u_top, u_bot;
  double bf[3], bf_coeff;
  double omega_T, gbeta, T_ref, T_hot, T_cold, gshift;
  double center[3];
  int inlet_mode, outlet_mode, wall_mode, body_force_on, buoyancy, g_bc_skip, bouzidi;
  int commit;    // write side effects (outlet column state, scalar, forces)
  int out_mode;  // materialize: 0 f64 field, 1 f64 macroscopic, 2 check partials
  int check_z;   // materialize/check: restrict to this z plane (-1: all)
  void* ev_start;  // host side only: events around the step launch (timing)
  void* ev_stop;
  AERO_GLOBAL double* fmat;
  AERO_GLOBAL double* check_part;
};

// One case of a batched step launch: its step parameters for even and odd
// steps of the chunk (fin/fout, g, outlet state and force buffers swapped),
// plus what changes every step.
struct BatchCase {
  StepParams p[2];
  long long step0;        // step count before the chunk
  const AERO_GLOBAL double* planes;   // per-step inlet planes (3 x nrow doubles each) or null
  AERO_GLOBAL double* hist;           // NCOEF doubles per step of the chunk
  int first_g_skip;       // skip the scalar wall values on the chunk's first step
  int pad;
};

// Inputs of the per-step force reduction (one block, fixed order).
struct ForceReduce {
  int q;               // lattice (9 or 19)
  const AERO_GLOBAL double* part;  // NPART doubles per link cell
  const AERO_GLOBAL LinkCell* lrec;
  int nrec, nx, ny;
  double cx, cy, cz;  // moment reference centre
  double fd, md;      // force and moment denominators
  AERO_GLOBAL double* row;        // NCOEF doubles out
  // NFORCE doubles per link cell, formed beforehand by force_terms_kernel, or
  // null: the reduction forms each cell's terms itself (the same bits)
  AERO_GLOBAL double* terms;
};

// Host-side construction (casts plain device pointers to the struct's types).
inline ForceReduce make_force_reduce(int q, const double* part, const LinkCell* lrec, int nrec, int nx, int ny,
                                     double cx, double cy, double cz, double fd, double md, double* row) {
  ForceReduce r;
  r.terms = nullptr;
  r.q = q;
  r.part = (decltype(r.part))part;
  r.lrec = (decltype(r.lrec))lrec;
  r.nrec = nrec;
  r.nx = nx;
  r.ny = ny;
  r.cx = cx;
  r.cy = cy;
  r.cz = cz;
  r.fd = fd;
  r.md = md;
  r.row = (decltype(r.row))row;
  return r;
}

}  // namespace aero
