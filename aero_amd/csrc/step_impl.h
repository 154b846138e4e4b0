// The fused pull-scheme step for D2Q9 / D3Q19 (templated on lattice,
// collision, storage and the scalar transport).
//
// State convention: the buffer B read by step k holds the post-collision
// populations of step k (the reference's f_pre, aero/lbm/solver.py:348).  One
// launch completes reference step k and starts step k+1:
//   pull-stream   f_new[i](x) = B[i](x - e_i), periodic in every direction
//                 (the inverse of the reference's push, kernels.py:94-115)
//   [3D body force (solver3d.py:411-423)]
//   bounce-back / Bouzidi on the cell's own links   (solver.py:379-383)
//   inlet (x = 0) -> outlet (x = nx-1) -> sponge -> walls (y = 0, ny-1)
//   scalar + Boussinesq force                        (solver.py:427-443)
//   momentum-exchange partial sums (FP64) at link cells (solver.py:449-463)
//   collision of step k+1 -> B'                     (kernels.py:42-92, ...)
// Every boundary condition in the reference reads only the cell it writes,
// the reference's f_pre at that cell, or (outlet) the cell at x = nx-2 in the
// same row; that cell's pre-outlet state is re-derived by the thread at
// x = nx-1 from B, so one launch per step is enough and no grid-wide barrier
// is needed.
#pragma once
#include <hip/hip_ext.h>

#include "lbm_device.h"

namespace aero {

// Tell the compiler a pointer addresses global memory (not LDS, not
// scratch).  Pointers read from a parameter table in memory are otherwise
// generic to it and their accesses compile to flat instructions.
template <class T>
__device__ __forceinline__ void assume_global(const T* p) {
#if defined(__HIP_DEVICE_COMPILE__)
  __builtin_assume(!__builtin_amdgcn_is_shared((const void*)p) && !__builtin_amdgcn_is_private((const void*)p));
#endif
}

// The parts of a step's parameters that change every step (kept apart so
// the batched kernel can read the rest straight from the case table).
struct StepDyn {
  double phase;                     // travelling inlet sine phase, 0.17 * step number
  const AERO_GLOBAL double* plane;  // this step's inlet planes or null
  AERO_GLOBAL double* hist_prev;    // coefficient row of the previous step
  int g_bc_skip;         // skip the scalar wall values (state just set from a raw field)
  int reduce_prev;       // block 0 reduces the previous step's force sums
  unsigned rstep;        // step number (mod 2^32): seeds the BF16 stochastic rounding
};
__device__ __forceinline__ StepDyn dyn_of(const StepParams& p) {
  return StepDyn{p.phase, p.plane, p.hist_prev, p.g_bc_skip, p.reduce_prev, 0u};
}

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// `false` as a per-lane value the optimiser cannot see through.  The paths
// that know their cells are fluid (bulk waves, grouped rows) pass it as the
// collision's solid flag, so the collision compiles to the same arithmetic
// as in the full path, where the flag is loaded: with a constant the
// optimiser would fuse u = m / rho into the equilibrium's sums, and a cell's
// last bit would depend on the path that stepped it.
__device__ __forceinline__ bool opaque_false() {
  int z = 0;
  asm volatile("" : "+v"(z));
  return z != 0;
}

// Grid coordinates of a linear cell index (x fastest).
__device__ __forceinline__ void cell_xyz(const StepParams& p, unsigned cell, int& x, int& y, int& z) {
  const unsigned yz = udiv(cell, p.div_nx);
  const unsigned zz = udiv(yz, p.div_ny);
  x = (int)(cell - yz * (unsigned)p.nx);
  y = (int)(yz - zz * (unsigned)p.ny);
  z = (int)zz;
}

// 1/rho for the collision: the hardware reciprocal (1 ulp) when computing in
// FP32, an IEEE division in FP64.
template <class C>
__device__ __forceinline__ C recip(C v) {
  if constexpr (sizeof(C) == 4) return __builtin_amdgcn_rcpf(v);
  else return C(1) / v;
}

// A plane base the optimiser may not fold into per-lane addresses: every
// store of a cell then forms scalar base + the shared 32-bit lane offset
// instead of a 64-bit vector address per plane.
template <class T>
__device__ __forceinline__ AERO_GLOBAL T* scalar_base(AERO_GLOBAL T* q) {
  asm("" : "+s"(q));
  return q;
}

// Guo immersed-boundary forcing toward a resting wall (ibm_guo.py:110-182):
// F = coeff (0 - u) from the cell's own populations (none where rho <= 0),
// then f_i += w_i e_i.F.
template <class L, class C>
__device__ __forceinline__ void ibm_force(C coeff, C fn[L::Q]) {
  C rho, mx, my, mz;
  moments<L>(fn, rho, mx, my, mz);
  if (!(rho > C(0))) return;
  const C Fx = coeff * (C(0) - mx / rho), Fy = coeff * (C(0) - my / rho), Fz = coeff * (C(0) - mz / rho);
#pragma unroll
  for (int i = 0; i < L::Q; ++i) fn[i] += C(L::W[i]) * edot<L>(i, Fx, Fy, Fz);
}

template <class C>
__device__ __forceinline__ C feq_i(C w, C rho, C eu, C usq) {
  return w * rho * (C(1) + C(3) * eu + C(4.5) * eu * eu - C(1.5) * usq);
}

// Storage row of grid row (y, z) in buffer B: its Q runs of xp values
// (population i of the row's cell x at element i * xp + x).  Wave-uniform in
// the fused step, where a wave covers one row.
template <class T>
__device__ __forceinline__ T* srow(const StepParams& p, T* B, int y, int z) {
  return B + (size_t)((unsigned)z * (unsigned)p.ny + (unsigned)y) * (unsigned)p.qxp;
}

// Pull gather of all Q populations of cell (x, y, z) from buffer B: the
// populations f (KIND 0, shifted by the free-stream equilibrium in the half
// forms) or the scalar's g (KIND 1, shifted by w_i * T_ref).
template <class L, int ST, int KIND = 0>
__device__ __forceinline__ void gather(const StepParams& p, const typename Storage<ST>::T* __restrict__ B,
                                       int x, int y, int z, typename Storage<ST>::C shift_scale /* u0 or T_ref */,
                                       typename Storage<ST>::C fn[L::Q]) {
  using S = Storage<ST>;
  using C = typename S::C;
  const int xm = x == 0 ? p.nx - 1 : x - 1, xp = x == p.nx - 1 ? 0 : x + 1;
  const int ym = y == 0 ? p.ny - 1 : y - 1, yp = y == p.ny - 1 ? 0 : y + 1;
  int zm = z, zp = z;
  if (L::DIM == 3) {
    zm = z == 0 ? p.nz - 1 : z - 1;
    zp = z == p.nz - 1 ? 0 : z + 1;
  }
#pragma unroll
  for (int i = 0; i < L::Q; ++i) {
    const int xs = L::EX[i] == 0 ? x : (L::EX[i] > 0 ? xm : xp);
    const int ys = L::EY[i] == 0 ? y : (L::EY[i] > 0 ? ym : yp);
    const int zs = L::EZ[i] == 0 ? z : (L::EZ[i] > 0 ? zm : zp);
    fn[i] = S::ld(srow(p, B, ys, zs) + (unsigned)(i * p.xp), (unsigned)xs,
                  KIND == 0 ? (C)p.fsh[fsh_class<L>(i)] : C(L::W[i]) * shift_scale);
  }
}

// v if bit is 1, else +0.0, by masking the bits: no per-lane condition (a
// 64-bit scalar register pair each) is kept live across the unrolled loops.
__device__ __forceinline__ double keep_if(double v, unsigned bit) {
  return __longlong_as_double(__double_as_longlong(v) & -(long long)bit);
}
__device__ __forceinline__ float keep_if(float v, unsigned bit) {
  return __int_as_float(__float_as_int(v) & -(int)bit);
}

// Link-cell record of a cell flagged FLAG_LINKS (a per-cell index only link
// cells read, so finding it costs one load).
__device__ __forceinline__ int link_record(const StepParams& p, int cell) { return p.rec_of[cell]; }

#ifndef LINK_GROUP
#define LINK_GROUP 9
#endif

// Bounce-back (boundary.py:283-300, boundary3d.py:333-350) or Bouzidi
// (boundary.py:303-352, boundary3d.py:353-404) on the links of record `rec`.
// For link i the pulled fn[OPP[i]] is replaced by the cell's own outgoing
// f_pre[i] (plain bounce-back), or interpolated with q: bouzidi 1 follows the
// reference (q < 1/2 uses f_pre[OPP[i]] at the cell, q >= 1/2 f_pre[i] at the
// clipped x - e_i), bouzidi 2 is the linear rule (q < 1/2 takes f_pre[i] at
// x - e_i, q >= 1/2 f_pre[OPP[i]] at the cell).  Every direction's operands
// are loaded whether or not it is a link (addresses stay in the domain) and
// the result is selected by the mask, so all loads issue together: one
// memory round trip instead of one per link.
template <class L, int ST>
__device__ __forceinline__ void apply_links_light(const StepParams& p, const typename Storage<ST>::T* __restrict__ B,
                                                  int rec, int x, int y, int z, typename Storage<ST>::C fn[L::Q]) {
  using S = Storage<ST>;
  using C = typename S::C;
  using T = typename S::T;
  constexpr int Q = L::Q;
  const uint32_t mask = p.lrec[rec].mask;
  const T* own = srow(p, B, y, z);
  if (!p.bouzidi) {
#pragma unroll
    for (int i = 1; i < Q; ++i) {
      const C fi = S::ld(own + (unsigned)(i * p.xp), (unsigned)x, (C)p.fsh[fsh_class<L>(i)]);
      fn[L::OPP[i]] = ((mask >> i) & 1u) ? fi : fn[L::OPP[i]];
    }
    return;
  }
  const float* qr = p.qdir + (size_t)rec * (Q - 1);
#pragma unroll
  for (int i = 1; i < Q; ++i) {
    // directions go in groups of LINK_GROUP: each group's loads issue
    // together, and the barrier keeps the scheduler from hoisting every
    // group's loads at once (which would need ~2Q extra registers)
    if (i > 1 && (i - 1) % LINK_GROUP == 0) __builtin_amdgcn_sched_barrier(0);
    const int o = L::OPP[i];
    const C q = (C)qr[i - 1];
    const bool lo = q < C(0.5);
    const bool upstream = p.bouzidi == 1 ? !lo : lo;
    const int xn = clampi(x + L::EX[o], 0, p.nx - 1);
    const int yn = clampi(y + L::EY[o], 0, p.ny - 1);
    const int zn = clampi(z + L::EZ[o], 0, p.nz - 1);
    const C fi = S::ld(own + (unsigned)(i * p.xp), (unsigned)x, (C)p.fsh[fsh_class<L>(i)]);
    const C f_up = S::ld(srow(p, B, yn, zn) + (unsigned)(i * p.xp), (unsigned)xn,
                         (C)p.fsh[fsh_class<L>(i)]);                                        // f_pre[i](x - e_i)
    const C f_op = S::ld(own + (unsigned)(o * p.xp), (unsigned)x, (C)p.fsh[fsh_class<L>(o)]);  // f_pre[OPP[i]](x)
    const C partner = upstream ? f_up : f_op;
    const C a = lo ? C(2) * q : C(1) / (C(2) * q);
    const C v = a * fi + (C(1) - a) * partner;
    fn[o] = ((mask >> i) & 1u) ? v : fn[o];
  }
}

// Momentum exchange at one link cell (forces.py:15-75 and 106-139,
// forces3d.py:8-66 and 398-425): per link i, f_pre[i](x) + f_new[OPP[i]](x)
// times e_i; pressure part from feq of the cell's f_pre (u not zeroed), the
// viscous part the remainder; moments about the reference centre.  The sums
// split into terms of f_pre alone (link_pre_terms, taken before the cell's
// populations are pulled) and
//   S = sum over links of e_i f_new[OPP[i]]
// (link_post_sum, just before the collision), each written to the cell's
// NPART-double record slot as soon as it is known, so no force term is held
// in registers across the step.  The reduction combines them per cell.
template <class L, int ST>
__device__ __forceinline__ void link_pre_terms(const StepParams& p, const typename Storage<ST>::T* __restrict__ B,
                                               int rec, int x, int y, int z) {
  using St = Storage<ST>;
  using C = typename St::C;
  constexpr int Q = L::Q;
  const uint32_t mask = p.lrec[rec].mask;
  const typename St::T* own = srow(p, B, y, z);
  C fp[Q];
#pragma unroll
  for (int i = 0; i < Q; ++i) fp[i] = St::ld(own + (unsigned)(i * p.xp), (unsigned)x, (C)p.fsh[fsh_class<L>(i)]);
  double rho = 0, mx = 0, my = 0, mz = 0, fx = 0, fy = 0, fz = 0;
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    const double v = (double)fp[i];
    rho += v;
    mx = madd(mx, L::EX[i], v);
    my = madd(my, L::EY[i], v);
    mz = madd(mz, L::EZ[i], v);
    const double lv = keep_if(v, (mask >> i) & 1u);
    fx = madd(fx, L::EX[i], lv);
    fy = madd(fy, L::EY[i], lv);
    fz = madd(fz, L::EZ[i], lv);
  }
  const double inv = rho > 0.0 ? 1.0 / rho : 0.0;
  double* out = p.fpart + (size_t)rec * NPART;
  out[0] = rho;
  out[1] = inv * mx;
  out[2] = inv * my;
  out[3] = inv * mz;
  out[4] = fx;
  out[5] = fy;
  out[6] = fz;
}

template <class L, class C>
__device__ __forceinline__ void link_post_sum(const StepParams& p, int rec, const C fn[L::Q]) {
  const uint32_t mask = p.lrec[rec].mask;
  double sx = 0, sy = 0, sz = 0;
#pragma unroll
  for (int i = 1; i < L::Q; ++i) {
    const double v = keep_if((double)fn[L::OPP[i]], (mask >> i) & 1u);
    sx = madd(sx, L::EX[i], v);
    sy = madd(sy, L::EY[i], v);
    sz = madd(sz, L::EZ[i], v);
  }
  double* out = p.fpart + (size_t)rec * NPART;
  out[7] = sx;
  out[8] = sy;
  out[9] = sz;
}

// The twelve per-cell sums (Fx Fy Fz, M = r x F, Fp xyz, Fv = F - Fp) of one
// link cell from its NPART record slot; r = cell - centre.  The pressure part
// uses feq_i + feq_OPP[i] = 2 w_i rho (1 + 4.5 (e_i.u)^2 - 1.5 u.u) (the
// linear terms cancel, w_OPP[i] = w_i), summed over the set bits in order.
template <class L>
__device__ __host__ __forceinline__ void cell_force_terms(const double* part, int cell, uint32_t mask, int nx, int ny,
                                                          double cx, double cy, double cz, double out[NFORCE]) {
  // no contraction into fused multiply-adds: the terms must come out the
  // same in every kernel that forms them (the step's reduction block and
  // force_terms_kernel)
#pragma clang fp contract(off)
  using D = DirPack<L>;
  const double rho = part[0], ux = part[1], uy = part[2], uz = part[3];
  const double base = 1.0 - 1.5 * (ux * ux + uy * uy + uz * uz);
  double px = 0, py = 0, pz = 0;
  for (uint32_t m = mask & ~1u; m; m &= m - 1) {
    const int i = __builtin_ctz(m);
    const int ex = D::ex(i), ey = D::ey(i), ez = D::ez(i);
    const double eu = ex * ux + ey * uy + ez * uz;
    const double mp = 2.0 * D::w(i) * rho * (base + 4.5 * eu * eu);
    px += ex * mp;
    py += ey * mp;
    pz += ez * mp;
  }
  const int x = cell % nx, yz = cell / nx, y = yz % ny, z = yz / ny;
  const double Fx = part[4] + part[7], Fy = part[5] + part[8], Fz = part[6] + part[9];
  const double rx = (double)x - cx, ry = (double)y - cy, rz = (double)z - cz;
  out[0] = Fx;
  out[1] = Fy;
  out[2] = Fz;
  out[3] = ry * Fz - rz * Fy;
  out[4] = rz * Fx - rx * Fz;
  out[5] = rx * Fy - ry * Fx;
  out[6] = px;
  out[7] = py;
  out[8] = pz;
  out[9] = Fx - px;
  out[10] = Fy - py;
  out[11] = Fz - pz;
}

// Uniform Guo body force on a fluid cell (ibm_guo.py:264-301).
template <class L, class C>
__device__ __forceinline__ void body_force(const StepParams& p, C fn[L::Q]) {
  C rho = 0, mx = 0, my = 0, mz = 0;
#pragma unroll
  for (int i = 0; i < L::Q; ++i) {
    rho += fn[i];
    mx += C(L::EX[i]) * fn[i];
    my += C(L::EY[i]) * fn[i];
    mz += C(L::EZ[i]) * fn[i];
  }
  if (!(rho > C(0))) return;
  const C ux = mx / rho, uy = my / rho, uz = mz / rho;
  const C fx = (C)p.bf[0], fy = (C)p.bf[1], fz = (C)p.bf[2];
  const C cs2 = C(1.0 / 3.0);
  const C uf = ux * fx + uy * fy + uz * fz;
  const C coeff = (C)p.bf_coeff;
#pragma unroll
  for (int i = 0; i < L::Q; ++i) {
    const C eu = C(L::EX[i]) * ux + C(L::EY[i]) * uy + C(L::EZ[i]) * uz;
    const C ef = C(L::EX[i]) * fx + C(L::EY[i]) * fy + C(L::EZ[i]) * fz;
    const C term = (ef - uf) / cs2 + (eu * ef) / (cs2 * cs2);
    fn[i] += coeff * C(L::W[i]) * rho * term;
  }
}

// ---------------------------------------------------------------- inlets
// The 2D boundary conditions below (inlet, pressure outlet, sponge, walls)
// compute with floating-point contraction off: a cell gets the same bits
// whichever step path (one row per wave, grouped rows) inlines them.
template <class C>
__device__ __forceinline__ void inlet_2d(const StepParams& p, const StepDyn& dyn, int y, C f[9]) {
#pragma clang fp contract(off)
  if (p.inlet_mode == IN_ZOUHE) {  // boundary.py:15-45
    const C u0 = (C)p.u0;
    const C rho = ((f[0] + f[2] + f[4]) + C(2) * (f[3] + f[6] + f[7])) / (C(1) - u0);
    f[1] = f[3] + C(2.0 / 3.0) * rho * u0;
    f[5] = f[7] - C(0.5) * (f[2] - f[4]) + C(1.0 / 6.0) * rho * u0;
    f[8] = f[6] + C(0.5) * (f[2] - f[4]) + C(1.0 / 6.0) * rho * u0;
    if (p.inlet_amp > 0.0) {
      const double ny = p.ny > 1 ? p.ny : 1;
      const C uy = (C)(p.inlet_amp * p.u0 * sin(2.0 * M_PI * (double)y / ny + dyn.phase));
      f[2] += uy * C(0.25);
      f[4] -= uy * C(0.25);
      f[5] += uy * C(0.15);
      f[8] -= uy * C(0.15);
    }
  } else if (p.inlet_mode == IN_FIELD) {  // boundary.py:48-65
    const C ux_t = (C)dyn.plane[y], uy_t = (C)dyn.plane[p.nrow + y];
    C den = C(1) - ux_t;
    if (den < C(1e-8)) den = C(1e-8);
    const C rho = ((f[0] + f[2] + f[4]) + C(2) * (f[3] + f[6] + f[7])) / den;
    const C cy = rho * uy_t;
    f[1] = f[3] + C(2.0 / 3.0) * rho * ux_t;
    f[5] = f[7] - C(0.5) * cy + C(1.0 / 6.0) * rho * ux_t;
    f[8] = f[6] + C(0.5) * cy + C(1.0 / 6.0) * rho * ux_t;
  } else if (p.inlet_mode == IN_PRESSURE) {  // boundary.py:116-136
    const C rin = (C)p.rho_in;
    const C ux = C(1) - (f[0] + f[2] + f[4] + C(2) * (f[3] + f[6] + f[7])) / rin;
    f[1] = f[3] + C(2.0 / 3.0) * rin * ux;
    f[5] = f[7] - C(0.5) * (f[2] - f[4]) + C(1.0 / 6.0) * rin * ux;
    f[8] = f[6] + C(0.5) * (f[2] - f[4]) + C(1.0 / 6.0) * rin * ux;
  }
}

// 2D Zou-He pressure outlet at x = nx-1 (boundary.py:139-159).
template <class C>
__device__ __forceinline__ void pressure_outlet_2d(const StepParams& p, C fn[9]) {
#pragma clang fp contract(off)
  const C rout = (C)p.rho_out;
  const C ux = C(-1) + (fn[0] + fn[2] + fn[4] + C(2) * (fn[1] + fn[5] + fn[8])) / rout;
  fn[3] = fn[1] - C(2.0 / 3.0) * rout * ux;
  fn[7] = fn[5] + C(0.5) * (fn[2] - fn[4]) - C(1.0 / 6.0) * rout * ux;
  fn[6] = fn[8] - C(0.5) * (fn[2] - fn[4]) - C(1.0 / 6.0) * rout * ux;
}

template <class C>
__device__ __forceinline__ void inlet_3d(const StepParams& p, const StepDyn& dyn, int y, int z, C f[19]) {
  const int row = z * p.ny + y;
  if (p.inlet_mode == IN_RECYCLING) {  // boundary3d.py:172-198
    const double scale = p.recyc[4 * p.nrow];
    const C rho = (C)p.recyc[row];
    const C ux = (C)(p.recyc[p.nrow + row] * scale);
    const C uy = (C)(p.recyc[2 * p.nrow + row] * scale);
    const C uz = (C)(p.recyc[3 * p.nrow + row] * scale);
    const C usq = ux * ux + uy * uy + uz * uz;
#pragma unroll
    for (int i = 0; i < 19; ++i) {
      const C eu = C(D3Q19::EX[i]) * ux + C(D3Q19::EY[i]) * uy + C(D3Q19::EZ[i]) * uz;
      f[i] = feq_i(C(D3Q19::W[i]), rho, eu, usq);
    }
    return;
  }
  const C Fm = f[2] + f[8] + f[10] + f[12] + f[14];
  const C Fn = f[0] + f[3] + f[4] + f[5] + f[6] + f[15] + f[16] + f[17] + f[18];
  const C cyb = f[3] - f[4] + f[15] - f[16] + f[17] - f[18];
  const C czb = f[5] - f[6] + f[15] + f[16] - f[17] - f[18];
  C ux, rho, cy, cz;
  if (p.inlet_mode == IN_ZOUHE) {  // boundary3d.py:24-68 (uz sine only, solver3d.py:447-452)
    ux = (C)p.u0;
    rho = (C(2) * Fm + Fn) / (C(1) - ux);
    cy = cyb;
    cz = czb;
    if (p.inlet_amp > 0.0) {
      const double nz = p.nz > 1 ? p.nz : 1;
      const C uz = (C)(p.inlet_amp * p.u0 * sin(2.0 * M_PI * (double)z / nz + dyn.phase));
      cz = cz + rho * uz;
    }
  } else if (p.inlet_mode == IN_FIELD) {  // boundary3d.py:71-93
    ux = (C)dyn.plane[row];
    C den = C(1) - ux;
    if (den < C(1e-8)) den = C(1e-8);
    rho = (C(2) * Fm + Fn) / den;
    cy = rho * (C)dyn.plane[p.nrow + row];
    cz = rho * (C)dyn.plane[2 * p.nrow + row];
  } else if (p.inlet_mode == IN_SEM) {  // boundary3d.py:96-137
    ux = (C)p.u0 + (C)dyn.plane[row];
    C den = C(1) - ux;
    if (den < C(1e-10)) den = C(1e-10);
    rho = (C(2) * Fm + Fn) / den;
    cy = cyb - rho * (C)dyn.plane[p.nrow + row];
    cz = czb - rho * (C)dyn.plane[2 * p.nrow + row];
  } else {
    return;
  }
  f[1] = f[2] + C(1.0 / 3.0) * rho * ux;
  f[7] = f[10] + C(1.0 / 6.0) * rho * ux - C(0.5) * cy;
  f[9] = f[8] + C(1.0 / 6.0) * rho * ux + C(0.5) * cy;
  f[11] = f[14] + C(1.0 / 6.0) * rho * ux - C(0.5) * cz;
  f[13] = f[12] + C(1.0 / 6.0) * rho * ux + C(0.5) * cz;
}

template <class L, class C>
__device__ __forceinline__ void apply_inlet(const StepParams& p, const StepDyn& dyn, int y, int z, C fn[L::Q]) {
  if constexpr (L::DIM == 2) inlet_2d<C>(p, dyn, y, fn);
  else inlet_3d<C>(p, dyn, y, z, fn);
}

// Sponge relaxation (sponge.py:24-70) and the y-wall conditions
// (boundary.py:162-226), in the reference's order; both read only the cell.
template <class L, class C>
__device__ __forceinline__ void sponge_and_walls(const StepParams& p, int x, int y, C fn[L::Q]) {
#pragma clang fp contract(off)
  constexpr int Q = L::Q;
  if (p.sigma != nullptr && x >= p.sponge_x0) {  // sponge.py:24-70; sigma is 0 before sponge_x0
    const double s = p.sigma[x];
    if (s > 0.0) {
      const C cs = (C)s;
      C rho = 0;
#pragma unroll
      for (int i = 0; i < Q; ++i) rho += fn[i];
      const C ut = (C)p.u0;
      const C usq = ut * ut;
#pragma unroll
      for (int i = 0; i < Q; ++i) {
        const C eu = C(L::EX[i]) * ut;
        fn[i] = (C(1) - cs) * fn[i] + cs * feq_i(C(L::W[i]), rho, eu, usq);
      }
    }
  }
  if (p.wall_mode != WALL_NONE && (y == 0 || y == p.ny - 1)) {  // boundary.py:162-226
    // Every branch indexes fn with compile-time constants only, so fn stays
    // in registers (a runtime-selected index would move it to scratch).
    if (p.wall_mode == WALL_SLIP) {
      if (y == 0) {
#pragma unroll
        for (int k = 0; k < L::NWALL; ++k) fn[L::SLIP_BOT[k][0]] = fn[L::SLIP_BOT[k][1]];
      }
      if (y == p.ny - 1) {
#pragma unroll
        for (int k = 0; k < L::NWALL; ++k) fn[L::SLIP_TOP[k][0]] = fn[L::SLIP_TOP[k][1]];
      }
    } else {
      const bool moving = p.wall_mode == WALL_MOVING;
      C rsum = 0;
      if (moving) {
#pragma unroll
        for (int i = 0; i < Q; ++i) rsum += fn[i];
        rsum = rsum > C(1e-12) ? rsum : C(1e-12);
      }
      if (y == 0) {
        const C add = moving ? C(6.0 * (1.0 / 36.0)) * rsum * (C)p.u_bot : C(0);
#pragma unroll
        for (int k = 0; k < L::NWALL; ++k)
          fn[L::NOSLIP_BOT[k][0]] = fn[L::NOSLIP_BOT[k][1]] + C(L::MOVE_SGN_BOT[k]) * add;
      }
      if (y == p.ny - 1) {
        const C add = moving ? C(6.0 * (1.0 / 36.0)) * rsum * (C)p.u_top : C(0);
#pragma unroll
        for (int k = 0; k < L::NWALL; ++k)
          fn[L::NOSLIP_TOP[k][0]] = fn[L::NOSLIP_TOP[k][1]] + C(L::MOVE_SGN_TOP[k]) * add;
      }
    }
  }
}

// State of a cell after streaming, body force, links and inlet: what the
// outlet and the recycling inlet read at x = nx-2 (boundary.py:68-113,
// boundary3d.py:185-190).  Returns the cell's link record or -1.
// Value of lane `i` of a float held in a VGPR, as a wave-uniform value.
__device__ __forceinline__ float lane_value(float v, int i) {
  return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), i));
}

// The previous convective-outlet state of the one outlet row in a wave,
// spread over lanes 0..Q-1 (lane i: population i), loaded with one coalesced
// instruction before the pull so its latency overlaps the pull's.  Only for
// waves holding exactly one x = nx-1 lane (always so when nx > 64) whose
// lanes 0..Q-1 are active (the grid's last wave can be partial), and for the
// 2-byte storage forms, whose outlet waves are latency-bound (the FP32 form
// measured slower with the extra live register); otherwise `coop` is false
// and the outlet lane loads its own state.
template <class L, int ST>
struct OutletRow {
  using C = typename Storage<ST>::C;
  C pv = 0;
  bool coop = false;
  int src = 0;  // the outlet lane
  __device__ __forceinline__ OutletRow(const StepParams& p, int x, int y, int z) {
    if constexpr (sizeof(C) == 4 && sizeof(typename Storage<ST>::T) == 2) {
      if (p.outlet_mode == OUT_CONVECTIVE) {
        const uint64_t m = __builtin_amdgcn_ballot_w64(x == p.nx - 1);
        const uint64_t lowq = (1ull << L::Q) - 1;
        if (m != 0 && (m & (m - 1)) == 0 && (__builtin_amdgcn_ballot_w64(true) & lowq) == lowq) {
          src = __builtin_ctzll(m);
          const int row = __builtin_amdgcn_readlane(z * p.ny + y, src);
          const unsigned lane = __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
          if (lane < (unsigned)L::Q) pv = (C)p.prev_in[(size_t)lane * p.nrow + row];
          coop = true;
        }
      }
    }
  }
};

template <class L, int ST>
__device__ __forceinline__ int pre_outlet(const StepParams& p, const StepDyn& dyn,
                                          const typename Storage<ST>::T* __restrict__ B, int cell,
                                          int x, int y, int z, uint8_t flag, typename Storage<ST>::C fn[L::Q]) {
  using C = typename Storage<ST>::C;
  gather<L, ST>(p, B, x, y, z, (C)p.u0, fn);
  if (L::DIM == 3 && p.body_force_on && !(flag & FLAG_SOLID)) body_force<L, C>(p, fn);
  int rec = -1;
  if (flag & FLAG_LINKS) {
    rec = link_record(p, cell);
    if (!p.ibm_on) apply_links_light<L, ST>(p, B, rec, x, y, z, fn);  // IBM: solver.py:379-383
  }
  if (x == 0) apply_inlet<L, C>(p, dyn, y, z, fn);
  return rec;
}

// Post-boundary state of one cell (the reference's self.f before the scalar
// step): pull, [3D body force], links, inlet, outlet, sponge, walls.  The
// outlet overwrites every population of the x = nx-1 column with the state of
// x = nx-2 (boundary.py:68-113), so that column pulls the x = nx-2 state
// instead of its own.  With commit set the convective outlet's state is
// advanced.  Returns the cell's own link record or -1.
//
// `nbr_flags` false: the x = nx-2 source of an outlet lane is known to be
// unflagged (no flag load).  `coop`: the wave holds exactly one outlet lane
// and lane i of `pv` already holds population i of that row's previous
// outlet state (OutletRow), so the outlet lane reads it across lanes instead
// of issuing its own dependent loads after the pull.
template <class L, int ST>
__device__ __forceinline__ int post_bc(const StepParams& p, const StepDyn& dyn,
                                       const typename Storage<ST>::T* __restrict__ B, int cell,
                                       int x, int y, int z, uint8_t flag, bool commit,
                                       typename Storage<ST>::C fn[L::Q], bool nbr_flags = true,
                                       typename Storage<ST>::C pv = 0, bool coop = false, int src = 0) {
  using C = typename Storage<ST>::C;
  constexpr int Q = L::Q;
  const int row = z * p.ny + y;
  const bool redirect = x == p.nx - 1 && (p.outlet_mode == OUT_ZEROGRAD || p.outlet_mode == OUT_CONVECTIVE);
  const int xg = redirect ? p.nx - 2 : x;
  const int cg = redirect ? cell - 1 : cell;
  const uint8_t fg = redirect ? (nbr_flags ? p.flags[cg] : (uint8_t)0) : flag;
  int rec = pre_outlet<L, ST>(p, dyn, B, cg, xg, y, z, fg, fn);
  C pvs[Q];  // wave-uniform copies of the outlet row's previous state (coop only)
  if constexpr (sizeof(C) == 4) {
    if (coop) {
#pragma unroll
      for (int i = 0; i < Q; ++i) pvs[i] = lane_value(pv, i);
    }
  }
  if (redirect) {
    if (p.outlet_mode == OUT_CONVECTIVE) {  // boundary.py:78-113
      const double u = p.u_conv < 0.0 ? 0.0 : (p.u_conv > 1.0 ? 1.0 : p.u_conv);
      const C cu = (C)u;
#pragma unroll
      for (int i = 0; i < Q; ++i) {
        const size_t k = (size_t)i * p.nrow + row;
        C prev_i;
        if constexpr (sizeof(C) == 4) prev_i = coop ? pvs[i] : (C)p.prev_in[k];
        else prev_i = (C)p.prev_in[k];
        const C v = (C(1) - cu) * prev_i + cu * fn[i];
        fn[i] = v;
        if (commit && !coop) p.prev_out[k] = (double)v;
      }
    }  // zero gradient (boundary.py:68-75): fn already is the x = nx-2 state
    rec = (flag & FLAG_LINKS) ? link_record(p, cell) : -1;
  }
  if constexpr (sizeof(C) == 4) {
    if (coop && commit) {  // lane i stores population i of the outlet lane's new state
      const unsigned lane = __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
      const int orow = __builtin_amdgcn_readlane(row, src);
      C mine = 0;
#pragma unroll
      for (int i = 0; i < Q; ++i) {
        const C vi = lane_value(fn[i], src);
        mine = lane == (unsigned)i ? vi : mine;
      }
      if (lane < (unsigned)Q) p.prev_out[(size_t)lane * p.nrow + orow] = (double)mine;
    }
  }
  if constexpr (L::DIM == 2) {
    if (x == p.nx - 1 && p.outlet_mode == OUT_PRESSURE) pressure_outlet_2d<C>(p, fn);
  }
  sponge_and_walls<L, C>(p, x, y, fn);
  return rec;
}

// Scalar transport + Boussinesq force (thermal.py:47-124, 149-222).  Pulls
// the previous scalar post-collision state from G (scalar_pull), applies the
// y-wall Dirichlet values (unless the state was just set from a raw field),
// adds the buoyancy force to fn and, with commit, writes the new
// post-collision g.  The fused step issues the scalar's pull together with
// the populations', before the flow's boundary conditions, so the two sets
// of loads share one memory round trip.
template <class L, int ST>
__device__ __forceinline__ void scalar_pull(const StepParams& p, int x, int y, int z,
                                            typename Storage<ST>::C g[L::Q]) {
  using S = Storage<ST>;
  gather<L, ST, 1>(p, (const typename S::T*)p.gin, x, y, z, (typename S::C)p.gshift, g);
}

template <class L, int ST>
__device__ __forceinline__ void scalar_apply(const StepParams& p, const StepDyn& dyn, int cell, int x, int y, int z,
                                             bool solid, typename Storage<ST>::C fn[L::Q],
                                             typename Storage<ST>::C g[L::Q], bool write_g) {
  using S = Storage<ST>;
  using C = typename S::C;
  constexpr int Q = L::Q;
  if (!dyn.g_bc_skip) {
    if (y == 0) {
#pragma unroll
      for (int i = 0; i < Q; ++i) g[i] = C(L::W[i]) * (C)p.T_hot;
    }
    if (y == p.ny - 1) {
#pragma unroll
      for (int i = 0; i < Q; ++i) g[i] = C(L::W[i]) * (C)p.T_cold;
    }
  }
  C T = g[0];
#pragma unroll
  for (int i = 1; i < Q; ++i) T += g[i];
  C rho, mx, my, mz;
  moments<L>(fn, rho, mx, my, mz);
  const C inv = rho > C(0) ? recip(rho) : C(0);
  const C ux = inv * mx, uy = inv * my, uz = inv * mz;
  if (p.buoyancy) {
    // directions with e_y = 0 receive exactly zero: skipped
    const C buoy = (C)p.gbeta * (T - (C)p.T_ref);
#pragma unroll
    for (int i = 0; i < Q; ++i)
      if (L::EY[i] != 0) fn[i] += C(L::W[i]) * C(3) * rho * buoy * C(L::EY[i]);
  }
  if (write_g) {
    const C om = (C)p.omega_T;
    AERO_GLOBAL typename S::T* grow = srow(p, (AERO_GLOBAL typename S::T*)p.gout, y, z);
    // BF16: stochastic rounding as the populations', with bits of its own
    // (seeded from the complemented step number, hashed per population)
    const unsigned gseed = ST == ST_BF16S ? sr_seed(cell, ~dyn.rstep) : 0u;
    auto put = [&](int i, C gp) {
      if constexpr (ST == ST_BF16S) {
        ((unsigned short*)(grow + (unsigned)(i * p.xp)))[x] =
            bf16_round_sr(gp - C(L::W[i]) * (C)p.gshift, mix32(gseed + (unsigned)i));
      } else {
        S::st((typename S::T*)(grow + (unsigned)(i * p.xp)), (unsigned)x, gp, C(L::W[i]) * (C)p.gshift);
      }
    };
    // geq_i = w T (1 + 3 e_i.u): a pair (i, opp i) shares w T and +-3 w T e_i.u
    put(0, solid ? C(L::W[0]) * T : g[0] - om * (g[0] - C(L::W[0]) * T));
#pragma unroll
    for (int i = 1; i < Q; ++i) {
      const int j = L::OPP[i];
      if (i > j) continue;
      const C wt = C(L::W[i]) * T;
      const C eu = edot<L>(i, ux, uy, uz);
      const C geq_i = wt * (C(1) + C(3) * eu);
      const C geq_j = wt * (C(1) - C(3) * eu);
      put(i, solid ? wt : g[i] - om * (g[i] - geq_i));
      put(j, solid ? wt : g[j] - om * (g[j] - geq_j));
    }
  }
}

template <class L, int ST>
__device__ __forceinline__ void scalar_step(const StepParams& p, const StepDyn& dyn, int cell, int x, int y, int z,
                                            bool solid, typename Storage<ST>::C fn[L::Q], bool write_g) {
  typename Storage<ST>::C g[L::Q];
  scalar_pull<L, ST>(p, x, y, z, g);
  scalar_apply<L, ST>(p, dyn, cell, x, y, z, solid, fn, g, write_g);
}

// Equilibrium moments of the MRT basis in closed form.  With
// feq_b = w_b rho (1 + 3 e_b.u + 4.5 (e_b.u)^2 - 1.5 u^2), row a of M feq is
// rho (A_a + 3 B_a.u + sum_d Cq_a,d q_d / 72) / 36, where 36 w_b is an integer
// for both lattices, A_a = sum_b M_ab 36 w_b, B_a = sum_b M_ab 36 w_b e_b and
// Cq holds the coefficients of q = (ux^2, uy^2, uz^2, ux uy, ux uz, uy uz)
// times 72: all exact integers, so a term whose coefficient vanishes is
// dropped at compile time.  D_a = sum_b M_ab^2 (the row norms).
template <class L>
struct MrtEq {
  struct Tab {
    long A[L::Q];
    long B[L::Q][3];
    long Cq[L::Q][6];
    long D[L::Q];
    bool even[L::Q];  // row a is even under i <-> opp(i) (else odd)
  };
  static constexpr Tab make() {
    Tab t{};
    for (int a = 0; a < L::Q; ++a) {
      for (int b = 0; b < L::Q; ++b) {
        const long w = (long)(L::W[b] * 36.0 + 0.5);
        const long m = (long)L::M[a][b] * w;
        const long e[3] = {L::EX[b], L::EY[b], L::EZ[b]};
        t.A[a] += m;
        for (int d = 0; d < 3; ++d) t.B[a][d] += m * e[d];
        // 72 * (4.5 (e.u)^2): 324 e_d e_f over ordered pairs; diagonal 324 e_d^2,
        // each cross pair 2 * 324 e_d e_f
        t.Cq[a][0] += 324 * m * e[0] * e[0];
        t.Cq[a][1] += 324 * m * e[1] * e[1];
        t.Cq[a][2] += 324 * m * e[2] * e[2];
        t.Cq[a][3] += 648 * m * e[0] * e[1];
        t.Cq[a][4] += 648 * m * e[0] * e[2];
        t.Cq[a][5] += 648 * m * e[1] * e[2];
        t.D[a] += (long)L::M[a][b] * L::M[a][b];
      }
      // 72 * (-1.5 u^2) per unit of 36 w: -108 A on each square
      for (int d = 0; d < 3; ++d) t.Cq[a][d] -= 108 * t.A[a];
      t.even[a] = true;
      for (int b = 0; b < L::Q; ++b) t.even[a] = t.even[a] && L::M[a][b] == L::M[a][L::OPP[b]];
    }
    return t;
  }
  static constexpr Tab T = make();
};

// Inputs of the local Smagorinsky relaxation time: C^2 = (C_s d)^2 (d the Van
// Driest factor, 1 without damping), tau0 = 1/omega, and the TRT rule for
// the antisymmetric rate.
template <class C>
struct LesLocal {
  C c2, tau0, lam;
  int variant;
};

// Antisymmetric TRT rate from a cell's relaxation time (kernels_trt.py:14-18
// for variant 0; variant 1 is the magic-parameter form).
template <class C>
__device__ __forceinline__ C trt_sm_of(C tau, C lam, int variant) {
  if (variant == 1) return recip(lam * recip(tau - C(0.5)) + C(0.5));
  const C tp = lam * tau + (C(1) - lam);
  return recip((tp * tau - C(0.5)) * recip(tp - C(0.5)));
}

// Smagorinsky closure on the local non-equilibrium stress.  With the
// Chapman-Enskog estimate Pi^neq = -2 rho c_s^2 tau S of the pre-collision
// populations (c_s^2 = 1/3), |S| = sqrt(2 S:S) = 3 Q / (2 rho tau) with
// Q = sqrt(2 Pi^neq:Pi^neq), so nu = nu0 + C^2 |S| and tau = 3 nu + 1/2 give
// tau^2 - tau0 tau - 9 C^2 Q / (2 rho) = 0, whose positive root is
//   tau = (tau0 + sqrt(tau0^2 + 18 C^2 Q / rho)) / 2
// (Hou, Sterling, Chen & Doolen, Fields Inst. Commun. 6 (1996) 151).
// Pi_ab = sum_i e_ia e_ib f_i is even under i <-> opp(i), so it sums the pair
// sums sp over one direction of each pair; Pi^eq = rho (delta/3 + u u).
template <class L, class C>
__device__ __forceinline__ C les_local_tau(const LesLocal<C>& la, C rho, C inv, C ux, C uy, C uz, const C sp[L::Q]) {
  C pxx = C(0), pyy = C(0), pzz = C(0), pxy = C(0), pxz = C(0), pyz = C(0);
#pragma unroll
  for (int i = 1; i < L::Q; ++i) {
    if (i > L::OPP[i]) continue;
    const int ex = L::EX[i], ey = L::EY[i], ez = L::EZ[i];
    if (ex != 0) pxx += sp[i];
    if (ey != 0) pyy += sp[i];
    if (ez != 0) pzz += sp[i];
    if (ex * ey > 0) pxy += sp[i];
    if (ex * ey < 0) pxy -= sp[i];
    if (ex * ez > 0) pxz += sp[i];
    if (ex * ez < 0) pxz -= sp[i];
    if (ey * ez > 0) pyz += sp[i];
    if (ey * ez < 0) pyz -= sp[i];
  }
  const C r3 = rho * C(1.0 / 3.0);
  const C nxx = pxx - (r3 + rho * ux * ux), nyy = pyy - (r3 + rho * uy * uy);
  const C nxy = pxy - rho * ux * uy;
  C pp = nxx * nxx + nyy * nyy + C(2) * nxy * nxy;
  if constexpr (L::DIM == 3) {
    const C nzz = pzz - (r3 + rho * uz * uz);
    const C nxz = pxz - rho * ux * uz, nyz = pyz - rho * uy * uz;
    pp += nzz * nzz + C(2) * (nxz * nxz + nyz * nyz);
  }
  const C qn = sqrt(C(2) * pp);
  return C(0.5) * (la.tau0 + sqrt(la.tau0 * la.tau0 + C(18) * la.c2 * qn * inv));
}

// Collision (BGK kernels.py:42-92, TRT kernels_trt.py:228-277,
// MRT kernels_mrt.py:30-82 / kernels3d_mrt.py:158-217).  u = 0 in solids.
// LESL: BGK / TRT rates from the local Smagorinsky relaxation time of each
// fluid cell (solids keep om, sm); MRT ignores it, as MRT ignores the
// reference's LES field.
template <class L, int COLL, class C, bool LESL = false>
__device__ __forceinline__ void collide(const double* mrt_s, const double* minv_d, const float* minv_f,
                                        bool solid, C om, C sm, const C fn[L::Q], C out[L::Q],
                                        const LesLocal<C>& la = LesLocal<C>{}) {
  constexpr int Q = L::Q;
  if constexpr (COLL == COLL_BGK || COLL == COLL_TRT) {
    // Opposite pairs (i, j = OPP[i], i < j): s = f_i + f_j and d = f_i - f_j
    // give the density (f_0 + sum s), the momentum (sum e_i d, as e_j = -e_i)
    // and the even / odd parts the relaxation acts on.
    C sp[Q], sd[Q];
    C rho = fn[0];
#pragma unroll
    for (int i = 1; i < Q; ++i) {
      const int j = L::OPP[i];
      if (i > j) continue;
      sp[i] = fn[i] + fn[j];
      sd[i] = fn[i] - fn[j];
      rho += sp[i];
    }
    C mx = C(0), my = C(0), mz = C(0);
    bool sx = false, sy = false, sz = false;
#pragma unroll
    for (int i = 1; i < Q; ++i) {
      if (i > L::OPP[i]) continue;
      if (L::EX[i] != 0) { mx = sx ? madd(mx, L::EX[i], sd[i]) : mfirst(L::EX[i], sd[i]); sx = true; }
      if (L::EY[i] != 0) { my = sy ? madd(my, L::EY[i], sd[i]) : mfirst(L::EY[i], sd[i]); sy = true; }
      if (L::EZ[i] != 0) { mz = sz ? madd(mz, L::EZ[i], sd[i]) : mfirst(L::EZ[i], sd[i]); sz = true; }
    }
    const C inv = rho > C(0) ? recip(rho) : C(0);
    C ux = inv * mx, uy = inv * my, uz = inv * mz;
    if (solid) {
      ux = C(0);
      uy = C(0);
      uz = C(0);
    }
    if constexpr (LESL) {
      if (!solid) {
        const C tau = les_local_tau<L, C>(la, rho, inv, ux, uy, uz, sp);
        om = recip(tau);
        if constexpr (COLL == COLL_TRT) sm = trt_sm_of<C>(tau, la.lam, la.variant);
      }
    }
    const C base = C(1) - C(1.5) * (ux * ux + uy * uy + uz * uz);
    // feq_i = even + odd with even = w rho (base + 4.5 eu^2), odd = 3 w rho eu
    // (eu = e_i.u) and feq_j = even - odd.  The new pair is
    //   a = om even + (1 - om) s/2,  b = sm odd + (1 - sm) d/2,
    //   f_i = a + b,  f_j = a - b,
    // which for BGK (sm = om) is (1 - om) f + om feq.  om w rho base,
    // 4.5 om w rho and 3 sm w rho are common to a weight shell.
    const C smv = COLL == COLL_TRT ? sm : om;
    const C h1 = C(0.5) * (C(1) - om);
    const C h2 = C(0.5) * (C(1) - smv);
    out[0] = (C(1) - om) * fn[0] + om * (C(L::W[0]) * rho) * base;
#pragma unroll
    for (int i = 1; i < Q; ++i) {
      const int j = L::OPP[i];
      if (i > j) continue;
      const C wr = C(L::W[i]) * rho;
      const C eu = edot<L>(i, ux, uy, uz);
      const C even = (C(4.5) * om * wr) * eu * eu + om * wr * base;
      const C a = h1 * sp[i] + even;
      const C b = h2 * sd[i] + (C(3) * smv * wr) * eu;
      out[i] = a + b;
      out[j] = a - b;
    }
  } else {
    C rho, mx, my, mz;
    moments<L>(fn, rho, mx, my, mz);
    const C inv = rho > C(0) ? recip(rho) : C(0);
    C ux = inv * mx, uy = inv * my, uz = inv * mz;
    if (solid) {
      ux = C(0);
      uy = C(0);
      uz = C(0);
    }
    // Moments m = M f; equilibrium moments M feq in closed form (MrtEq:
    // rho * (A + B.u + u^T C u), the same sums the reference's M @ feq forms);
    // relaxed moments mp = m - s (m - m_eq).
    using E_ = MrtEq<L>;
    const C uq[6] = {ux * ux, uy * uy, uz * uz, ux * uy, ux * uz, uy * uz};
    const C uv[3] = {ux, uy, uz};
    // Every row of M is even or odd under i <-> opp(i), so each row is a sum
    // over direction pairs of f_i + f_opp(i) (even) or f_i - f_opp(i) (odd):
    // half the terms.
    C ps[Q], pd[Q];
#pragma unroll
    for (int i = 1; i < Q; ++i) {
      if (i < L::OPP[i]) {
        ps[i] = fn[i] + fn[L::OPP[i]];
        pd[i] = fn[i] - fn[L::OPP[i]];
      }
    }
    C mp[Q];
#pragma unroll
    for (int a = 0; a < Q; ++a) {
      C m = 0;
      bool mfirst_ = true;
      if (E_::T.even[a] && L::M[a][0] != 0) {
        m = L::M[a][0] == 1 ? fn[0] : C(L::M[a][0]) * fn[0];
        mfirst_ = false;
      }
#pragma unroll
      for (int b = 1; b < Q; ++b) {
        if (b < L::OPP[b] && L::M[a][b] != 0) {
          const C v = E_::T.even[a] ? ps[b] : pd[b];
          const C term = L::M[a][b] == 1 ? v : (L::M[a][b] == -1 ? -v : C(L::M[a][b]) * v);
          m = mfirst_ ? term : m + term;
          mfirst_ = false;
        }
      }
      C t = 0;
      bool tfirst_ = true;
      if (E_::T.A[a] != 0) {
        t = C(E_::T.A[a] / 36.0);
        tfirst_ = false;
      }
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        if (E_::T.B[a][d] != 0) {
          const C term = C(3.0 * E_::T.B[a][d] / 36.0) * uv[d];
          t = tfirst_ ? term : t + term;
          tfirst_ = false;
        }
      }
#pragma unroll
      for (int d = 0; d < 6; ++d) {
        if (E_::T.Cq[a][d] != 0) {
          const C term = C(E_::T.Cq[a][d] / 2592.0) * uq[d];
          t = tfirst_ ? term : t + term;
          tfirst_ = false;
        }
      }
      const C me = rho * t;
      mp[a] = m - C(mrt_s[a]) * (m - me);
    }
    if constexpr (L::Q == 9) {
      // D2Q9's rows are orthogonal: M^-1 = M^T D^-1 with D = diag(M M^T),
      // exact in closed form (the reference inverts M9 numerically; the two
      // agree to 2e-16)
      C r[Q];
#pragma unroll
      for (int a = 0; a < Q; ++a) r[a] = mp[a] * C(1.0 / E_::T.D[a]);
      // out_i = E_i + O_i and out_opp(i) = E_i - O_i over the even / odd rows
#pragma unroll
      for (int i = 0; i < Q; ++i) {
        if (i != 0 && i > L::OPP[i]) continue;
        C ev = 0, od = 0;
        bool efirst_ = true, ofirst_ = true;
#pragma unroll
        for (int a = 0; a < Q; ++a) {
          if (L::M[a][i] != 0) {
            const C term = L::M[a][i] == 1 ? r[a] : (L::M[a][i] == -1 ? -r[a] : C(L::M[a][i]) * r[a]);
            if (E_::T.even[a]) {
              ev = efirst_ ? term : ev + term;
              efirst_ = false;
            } else {
              od = ofirst_ ? term : od + term;
              ofirst_ = false;
            }
          }
        }
        if (i == 0) {
          out[0] = ev;
        } else {
          out[i] = ev + od;
          out[L::OPP[i]] = ev - od;
        }
      }
    } else {
      // The host's pseudo-inverse P, skipping its entries that are zero up
      // to rounding (D3Q19::PINV_NZ).  P keeps the rows' symmetry
      // (P[opp i][a] = +-P[i][a] for even / odd rows, to 2e-15 in the
      // float64 pseudo-inverse), so row i gives both out_i = E_i + O_i and
      // out_opp(i) = E_i - O_i.
      const C* minv;
      if constexpr (sizeof(C) == 8) minv = (const C*)minv_d;
      else minv = (const C*)minv_f;
#pragma unroll
      for (int i = 0; i < Q; ++i) {
        if (i != 0 && i > L::OPP[i]) continue;
        C ev = 0, od = 0;
        bool efirst_ = true, ofirst_ = true;
#pragma unroll
        for (int a = 0; a < Q; ++a) {
          if (L::PINV_NZ[i][a]) {
            const C term = minv[i * Q + a] * mp[a];
            if (E_::T.even[a]) {
              ev = efirst_ ? term : ev + term;
              efirst_ = false;
            } else {
              od = ofirst_ ? term : od + term;
              ofirst_ = false;
            }
          }
        }
        if (i == 0) {
          out[0] = ev;
        } else {
          out[i] = ev + od;
          out[L::OPP[i]] = ev - od;
        }
      }
    }
  }
}

// ---------------------------------------------------------------- kernels
// Tail of a step for one cell: collide into fout (KMODE 0; KMODE 2 with the
// local Smagorinsky rates), or -- split step for the LES finite-difference
// pass -- write the post-boundary populations and velocity (KMODE 1).
// `sb` is the storage element of the wave's first cell: its row's base plus
// the wave's first x (wave-uniform), so a lane's cell sits at sb + lane.

template <class L, int COLL, int ST, int KMODE>
__device__ __forceinline__ void finish_cell(const StepParams& p, int cell, size_t sb, bool solid,
                                            const typename Storage<ST>::C fn[L::Q], unsigned rstep = 0) {
  using S = Storage<ST>;
  using C = typename S::C;
  using T = typename S::T;
  constexpr int Q = L::Q;
  const size_t N = (size_t)p.ncell;
  if constexpr (KMODE == 0 || KMODE == 2) {
    C out[Q];
    if constexpr (KMODE == 2) {
      C c2 = (C)p.les_c2;
      if constexpr (sizeof(C) == 8) {
        if (p.les_c2d) c2 = (C)p.les_c2d[cell];
      } else {
        if (p.les_c2f) c2 = (C)p.les_c2f[cell];
      }
      const LesLocal<C> la{c2, (C)p.les_tau0, (C)p.trt_lambda, p.trt_variant};
      collide<L, COLL, C, true>(p.mrt_s, p.minv_d, p.minv_f, solid, (C)p.omega, (C)p.trt_sm, fn, out, la);
    } else {
      collide<L, COLL, C>(p.mrt_s, p.minv_d, p.minv_f, solid, (C)p.omega, (C)p.trt_sm, fn, out);
    }
    // the store offset is the lane index, rebuilt here rather than held in a
    // register across the whole step; the run bases are scalar
    const unsigned lane = __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
    AERO_GLOBAL T* wbase = (AERO_GLOBAL T*)p.fout + sb;
    if constexpr (ST == ST_BF16S) {
      // BF16 rounds stochastically (bf16_round_sr), bits from (cell, step, population)
      unsigned h = sr_seed(cell, rstep);
#pragma unroll
      for (int i = 0; i < Q; ++i) {
        h = sr_next(h);
        const unsigned short b = bf16_round_sr(out[i] - (float)p.fsh[fsh_class<L>(i)], h);
        unsigned short* a = (unsigned short*)scalar_base(wbase + (unsigned)(i * p.xp)) + lane;
        if (p.nt_store) __builtin_nontemporal_store(b, a);
        else *a = b;
      }
      return;
    }
    if (p.nt_store) {
#pragma unroll
      for (int i = 0; i < Q; ++i)
        S::st_nt((T*)scalar_base(wbase + (unsigned)(i * p.xp)), lane, out[i], (C)p.fsh[fsh_class<L>(i)]);
    } else {
#pragma unroll
      for (int i = 0; i < Q; ++i)
        S::st((T*)scalar_base(wbase + (unsigned)(i * p.xp)), lane, out[i], (C)p.fsh[fsh_class<L>(i)]);
    }
  } else {
    C* F = (C*)p.fsplit;
    C* U = (C*)p.usplit;
    C rho = 0, mx = 0, my = 0, mz = 0;
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      F[(size_t)i * N + cell] = fn[i];
      rho += fn[i];
      mx += C(L::EX[i]) * fn[i];
      my += C(L::EY[i]) * fn[i];
      mz += C(L::EZ[i]) * fn[i];
    }
    const C inv = rho > C(0) ? C(1) / rho : C(0);
    U[cell] = inv * mx;
    U[N + cell] = inv * my;
    U[2 * N + cell] = inv * mz;
  }
}

__device__ __forceinline__ void write_coefficients(const double F[NFORCE], double fd, double md, double* row) {
  // forces.py:78-139 / forces3d.py:364-439
  const bool ok = fd != 0.0;
  row[0] = ok ? F[0] / fd : 0.0;
  row[1] = ok ? F[1] / fd : 0.0;
  row[2] = ok ? F[2] / fd : 0.0;
  row[3] = md != 0.0 ? F[3] / md : 0.0;
  row[4] = md != 0.0 ? F[4] / md : 0.0;
  row[5] = md != 0.0 ? F[5] / md : 0.0;
  row[6] = ok ? F[6] / fd : 0.0;
  row[7] = ok ? F[9] / fd : 0.0;
#pragma unroll
  for (int k = 0; k < NFORCE; ++k) row[8 + k] = F[k];
}

// Fixed-order FP64 reduction of the per-link-cell momentum-exchange sums by
// one block of BLOCK threads: thread t sums link cells t, t+256, ...
// sequentially; the four waves then add their partials into a 64-lane LDS
// row one wave after another, and a fixed tree of wavefront shuffles
// reduces that row.  The order
// never depends on timing or on what else runs, so a case reproduces its own
// bits alone or in a batch, whichever launch performs the reduction.  6 KB of
// LDS, so the step kernel that hosts it keeps its occupancy.
// PF link cells of a thread have their inputs loaded before the first is
// summed (same order of additions for every PF): the standalone reduction
// launch, which sits on the critical path at the end of a chunk, uses PF = 8
// to overlap the memory latency of successive cells; inside the step kernel,
// where it overlaps the cell blocks and shares their register budget, PF = 1.
// TERMS: the record may carry terms formed beforehand (the standalone
// launch); the step kernel's reduction block forms them itself.
template <class L, int PF = 1, bool TERMS = false>
__device__ __forceinline__ void reduce_forces_block(const ForceReduce& r) {
  __shared__ double red[NFORCE][64];
  double acc[NFORCE];
#pragma unroll
  for (int k = 0; k < NFORCE; ++k) acc[k] = 0.0;
  if (TERMS && r.terms) {
    // terms formed beforehand: a plain sum in the same order, PF cells ahead
    for (int k0 = threadIdx.x; k0 < r.nrec; k0 += PF * BLOCK) {
      double tt[PF][NFORCE];
#pragma unroll
      for (int j = 0; j < PF; ++j) {
        const int k = k0 + j * BLOCK;
        if (k < r.nrec) {
#pragma unroll
          for (int q = 0; q < NFORCE; ++q) tt[j][q] = r.terms[(size_t)k * NFORCE + q];
        }
      }
#pragma unroll
      for (int j = 0; j < PF; ++j) {
        if (k0 + j * BLOCK < r.nrec) {
#pragma unroll
          for (int q = 0; q < NFORCE; ++q) acc[q] += tt[j][q];
        }
      }
    }
  } else if constexpr (PF == 1) {
    // operands read as they are used: fewer registers live in the step kernel
    for (int k = threadIdx.x; k < r.nrec; k += BLOCK) {
      double t[NFORCE];
      const LinkCell lc = r.lrec[k];
      cell_force_terms<L>(r.part + (size_t)k * NPART, lc.cell, lc.mask, r.nx, r.ny, r.cx, r.cy, r.cz, t);
#pragma unroll
      for (int j = 0; j < NFORCE; ++j) acc[j] += t[j];
    }
  } else {
    for (int k0 = threadIdx.x; k0 < r.nrec; k0 += PF * BLOCK) {
      LinkCell lc[PF];
      double pp[PF][NPART];
#pragma unroll
      for (int j = 0; j < PF; ++j) {
        const int k = k0 + j * BLOCK;
        if (k < r.nrec) {
          lc[j] = r.lrec[k];
#pragma unroll
          for (int q = 0; q < NPART; ++q) pp[j][q] = r.part[(size_t)k * NPART + q];
        }
      }
#pragma unroll
      for (int j = 0; j < PF; ++j) {
        if (k0 + j * BLOCK < r.nrec) {
          double t[NFORCE];
          cell_force_terms<L>(pp[j], lc[j].cell, lc[j].mask, r.nx, r.ny, r.cx, r.cy, r.cz, t);
#pragma unroll
          for (int q = 0; q < NFORCE; ++q) acc[q] += t[q];
        }
      }
    }
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int w = 0; w < BLOCK / 64; ++w) {
    if (wave == w) {
#pragma unroll
      for (int k = 0; k < NFORCE; ++k) red[k][lane] = (w == 0 ? 0.0 : red[k][lane]) + acc[k];
    }
    __syncthreads();
  }
  // the 64-lane row by wavefront shuffles in wave 0: at each halving, lane
  // i adds lane i + h (the fixed tree; lanes past the live half carry
  // values nobody reads)
  if (wave == 0) {
    double F[NFORCE];
#pragma unroll
    for (int k = 0; k < NFORCE; ++k) F[k] = red[k][lane];
#pragma unroll
    for (int h = 32; h > 0; h >>= 1) {
#pragma unroll
      for (int k = 0; k < NFORCE; ++k) F[k] += __shfl_down(F[k], h, 64);
    }
    if (lane == 0) write_coefficients(F, r.fd, r.md, r.row);
  }
}

// One reference step for every cell in a single launch.  Each thread pulls
// its cell's populations and applies, in the reference's order, the
// conditions that concern it: links, inlet, outlet (the outlet column pulls
// the x = nx-2 state instead of its own, since the outlet overwrites all of
// it), sponge, walls, scalar + buoyancy; link cells then write their FP64
// momentum-exchange sums, and the cell is collided for the next step.
// Rare work (links, outlet) sits behind flag / column branches and streams
// its operands from memory, so the register budget -- and the occupancy that
// hides HBM latency -- is set by the common pull + collide path.  One extra
// block (index 0, dispatched first so it overlaps the cell blocks) reduces the
// previous step's force sums, written by the previous launch, into that
// step's coefficient row.
// Occupancy target of the step kernel: the 4- and 2-byte BGK/TRT forms ask
// the register allocator for 7 waves per SIMD (at most 72 VGPRs; the common
// pull + collide path needs ~50, the rare link / outlet / reduction paths
// would otherwise raise the whole kernel to ~85).  The 2-byte D2Q9 forms ask
// for 8 (64 VGPRs, a 3-VGPR spill on a rare path): a D2Q9 wave keeps only 9
// two-byte loads in flight, so the extra wave per SIMD raised their batches
// 5%; likewise D2Q9 MRT (+5%).  D3Q19 MRT in the 2-byte forms
// targets 6 (the compiler's own choice, 5, measured 10% slower), or 5 with
// the scalar (the compiler's 4: 1.441 against 1.411 ms per step on the
// 512 x 256 x 256 MRT + scalar configuration; 6 spills, 2.11 ms); FP64,
// FP32 MRT and D2Q9 MRT with the scalar keep the compiler's choice.  The D3Q19
// local-Smagorinsky kernels (KMODE 2) without the scalar target 6: at 512 x
// 256 x 256, 0.852 vs 0.888 ms (FP32) and 0.513 vs 0.566 ms (FP16) at 7,
// while the plain kernels stay fastest at 7 (0.832 vs 0.845, 0.463 vs 0.485).
template <class L, int COLL, int ST, bool THERMAL, int KMODE = 0>
struct StepWaves {
  static constexpr int value = (KMODE == 2 && L::Q == 19 && ST != ST_F64 && !THERMAL) ? 6
                              : (ST != ST_F64 && COLL != COLL_MRT)
                                  ? ((L::Q == 9 && (ST == ST_F16S || ST == ST_BF16S)) ? 8 : 7)
                              : (COLL == COLL_MRT && (ST == ST_F16S || ST == ST_BF16S)) ? (THERMAL ? (L::Q == 19 ? 5 : 1)
                                                                                                   : (L::Q == 9 ? 8 : 6))
                                                                                        : 1;
};

// The D3Q19 FP32 MRT kernel with the scalar issues the scalar's pull with the
// populations' (scalar_pull): 1.885 against 2.005 ms per step on the
// 512 x 256 x 256 MRT + scalar configuration, at the 4 waves per SIMD the
// compiler then picks (6 without it).  Forcing 5 or 6 waves spills, and in
// the 2-byte forms the early pull drops the kernel to 3 waves and measured
// slower (1.521 against 1.411 ms at 5 waves without it); the BGK / TRT
// kernels with the scalar spill heavily at their 7-wave target with it.
template <class L, int COLL, int ST>
struct ScalarPrefetch {
  static constexpr bool value = COLL == COLL_MRT && L::Q == 19 && ST == ST_F32;
};

// The rows of a group item (grouped kernels, D2Q9): rows y0 .. y0+RG-1 of 64
// aligned columns x, none of them flagged (the inlet, a zero-gradient or
// pressure outlet, sponge columns and the wall rows may be among them).
// Every pull of the RG rows is issued, into the raw storage type, before the
// first row is converted: the wave keeps RG x Q loads in flight.  Each row
// then runs what the one-row path runs for such a cell, in the same order and
// with the same arithmetic: the pull (the outlet column pulls x = nx-2, as
// post_bc), the inlet at x = 0, the pressure outlet, the sponge and the
// walls, the collision (cvt(raw) == ld).
template <class L, int COLL, int ST, int KMODE, int RG>
__device__ __forceinline__ void step_rows_group(const StepParams& p, const StepDyn& dyn,
                                                const typename Storage<ST>::T* __restrict__ B, int x, int y0,
                                                unsigned xw) {
  using S = Storage<ST>;
  using C = typename S::C;
  using T = typename S::T;
  constexpr int Q = L::Q;
  static_assert(L::DIM == 2 && sizeof(T) == 2 && RG % 2 == 0, "grouped rows: D2Q9, 2-byte storage");
  const bool redirect = x == p.nx - 1 && (p.outlet_mode == OUT_ZEROGRAD || p.outlet_mode == OUT_CONVECTIVE);
  const int xg = redirect ? p.nx - 2 : x;
  const int xm = xg == 0 ? p.nx - 1 : xg - 1, xq = xg == p.nx - 1 ? 0 : xg + 1;
  // rows r and r + 1 of a population share one register (16-bit halves)
  typedef unsigned short pair_t __attribute__((ext_vector_type(2)));
  pair_t raw[RG / 2][Q];
#pragma unroll
  for (int r = 0; r < RG; ++r) {
    const int y = y0 + r;
    const int ym = y == 0 ? p.ny - 1 : y - 1, yp = y == p.ny - 1 ? 0 : y + 1;
#pragma unroll
    for (int i = 0; i < Q; ++i) {
      const int xs = L::EX[i] == 0 ? xg : (L::EX[i] > 0 ? xm : xq);
      const int ys = L::EY[i] == 0 ? y : (L::EY[i] > 0 ? ym : yp);
      const T v = S::raw(srow(p, B, ys, 0) + (unsigned)(i * p.xp), (unsigned)xs);
      raw[r / 2][i][r % 2] = __builtin_bit_cast(unsigned short, v);
    }
  }
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int r = 0; r < RG; ++r) {
    const int y = y0 + r;
    C fn[Q];
#pragma unroll
    for (int i = 0; i < Q; ++i)
      fn[i] = S::cvt(__builtin_bit_cast(T, (unsigned short)raw[r / 2][i][r % 2]), (C)p.fsh[fsh_class<L>(i)]);
    if (xg == 0) apply_inlet<L, C>(p, dyn, y, 0, fn);
    if (x == p.nx - 1 && p.outlet_mode == OUT_PRESSURE) pressure_outlet_2d<C>(p, fn);
    sponge_and_walls<L, C>(p, x, y, fn);
    finish_cell<L, COLL, ST, KMODE>(p, y * p.nx + x, (size_t)(unsigned)y * (unsigned)p.qxp + xw, opaque_false(),
                                    fn, dyn.rstep);
  }
}

template <class L, int COLL, int ST, bool THERMAL>
struct RowsPerWave {
  static constexpr int value = rows_per_wave(L::Q, COLL, ST, THERMAL);
};

// The step of the cells of one wave (or, in block 0, the previous step's
// force reduction), expanded textually into each step kernel with `p` (the
// case's StepParams), `dyn` (this step's StepDyn) and `blk` (the case-local
// block index) in scope.  A macro rather than an inlined function: with the
// body in a callee the register allocator's outcome for the 2-byte storage
// forms degrades sharply (heavy spilling at the same occupancy target).
// A wave covers 64 aligned columns (xp is a multiple of 64) of one row, or,
// for a group item of the grouped kernels' wave list, of RPW_GROUP rows
// (step_rows_group).  Without a list (nitems 0: other kernels, or the launch
// tuning's one-row setting) the step's thread grid is nrow * xp.
#define AERO_STEP_BODY(L, COLL, ST, THERMAL, KMODE)                                                        \
  do {                                                                                                     \
    using S_ = Storage<ST>;                                                                                \
    using C_ = typename S_::C;                                                                             \
    using T_ = typename S_::T;                                                                             \
    if (blk == 0) {                                                                                        \
      if (dyn.reduce_prev) {                                                                               \
        const ForceReduce r_{L::Q, p.fpart_prev, p.lrec, p.nrec, p.nx, p.ny, p.center[0], p.center[1],     \
                             p.center[2], p.fdenom, p.mdenom, dyn.hist_prev};                              \
        reduce_forces_block<L>(r_);                                                                        \
      }                                                                                                    \
      break;                                                                                               \
    }                                                                                                      \
    /* T0_: the wave's first thread in the one-row grid; thread indices fit in 31 bits (checked */        \
    /* when the context is created).  T0_, the row, xw_ (the wave's first column) and every row */         \
    /* base are wave-uniform (scalar registers). */                                                        \
    constexpr bool LISTED_ = RowsPerWave<L, COLL, ST, THERMAL>::value > 1;                                 \
    unsigned T0_;                                                                                          \
    bool group_ = false;                                                                                   \
    if (LISTED_ && p.nitems > 0) {                                                                         \
      const unsigned wid_ = (unsigned)(blk - 1) * (unsigned)(BLOCK / 64) +                                 \
                            ((unsigned)__builtin_amdgcn_readfirstlane(threadIdx.x) >> 6);                  \
      if (wid_ >= (unsigned)p.nitems) break;                                                               \
      const unsigned it_ = __builtin_amdgcn_readfirstlane(p.witems[wid_]);                                 \
      group_ = (it_ & WITEM_GROUP) != 0u;                                                                  \
      T0_ = (it_ & ~WITEM_GROUP) << 6;                                                                     \
    } else {                                                                                               \
      T0_ = (unsigned)(blk - 1) * (unsigned)BLOCK + __builtin_amdgcn_readfirstlane(threadIdx.x);           \
    }                                                                                                      \
    const unsigned row_ = udiv(T0_, p.div_xp);                                                             \
    if (row_ >= (unsigned)p.nrow) break;                                                                   \
    const unsigned xw_ = T0_ - row_ * (unsigned)p.xp;                                                      \
    const int x = (int)(xw_ + __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)));          \
    if (x >= p.nx) break; /* row padding */                                                                \
    const T_* B = (const T_*)p.fin;                                                                        \
    if constexpr (LISTED_) {                                                                               \
      if (group_) {                                                                                        \
        step_rows_group<L, COLL, ST, KMODE, RPW_GROUP>(p, dyn, B, x, (int)row_, xw_);                            \
        break;                                                                                             \
      }                                                                                                    \
    }                                                                                                      \
    AERO_STEP_ROW(L, COLL, ST, THERMAL, KMODE);                                                            \
  } while (0)

// The step of the wave's cells in grid row row_ (columns x = xw_ + lane),
// inside AERO_STEP_BODY; a `break` ends the wave.
#define AERO_STEP_ROW(L, COLL, ST, THERMAL, KMODE)                                                         \
  do {                                                                                                     \
    const unsigned t0_ = T0_; /* wave index t0_ / 64 indexes the wave flag words */                       \
    const int z = (int)udiv(row_, p.div_ny);                                                               \
    const int y = (int)row_ - z * p.ny;                                                                    \
    const int cell = (int)(row_ * (unsigned)p.nx) + x;                                                     \
    const size_t sb_ = (size_t)row_ * (unsigned)p.qxp + xw_;                                               \
    /* Bulk waves: no cell of the wave flagged (solid, link, IBM; one bit per 64 cells, a scalar */        \
    /* load) and no lane at a domain edge or in the sponge -> pull, [body force,] collide, store, */      \
    /* straight-line.  Every boundary condition below is a no-op for such a cell, so this is the */       \
    /* same arithmetic (the solid flag passed opaque, as the full sequence's loaded one).  Only */         \
    /* D2Q9 BGK/TRT keep it (+6-8% in FP16 batches); in D3Q19 the full sequence below, which skips */     \
    /* the flag load in unflagged waves, measured as fast (and faster on small grids), and MRT */         \
    /* kernels are register-bound. */                                                                     \
    if constexpr (!THERMAL && COLL != COLL_MRT && L::DIM == 2) {                                           \
      const unsigned w_ = t0_ >> 6;                                                                        \
      const bool edge_ = x == 0 || x == p.nx - 1 || x >= p.sponge_x0 ||                                    \
                         (p.wall_mode != WALL_NONE && (y == 0 || y == p.ny - 1));                          \
      if (!((p.wave_any[w_ >> 5] >> (w_ & 31)) & 1u) && __builtin_amdgcn_ballot_w64(edge_) == 0) {        \
        C_ fq_[L::Q];                                                                                      \
        gather<L, ST>(p, B, x, y, z, (C_)p.u0, fq_);                                                       \
        if (L::DIM == 3 && p.body_force_on) body_force<L, C_>(p, fq_);                                     \
        finish_cell<L, COLL, ST, KMODE>(p, cell, sb_, opaque_false(), fq_, dyn.rstep);                     \
        break;                                                                                             \
      }                                                                                                    \
    }                                                                                                      \
    /* Remaining waves: the cell flags are loaded only where the wave (or, for the outlet's */            \
    /* x = nx-2 source, the wave before) holds a flagged cell, and a convective outlet row's */            \
    /* previous state is fetched before the pull, so neither adds a dependent memory round trip. */       \
    const OutletRow<L, ST> orow_(p, x, y, z);                                                              \
    const unsigned wv_ = t0_ >> 6;                                                                         \
    const unsigned wp_ = wv_ > 0 ? wv_ - 1 : 0;                                                            \
    const bool wflag_ = (p.wave_any[wv_ >> 5] >> (wv_ & 31)) & 1u;                                         \
    const bool pflag_ = wflag_ || ((p.wave_any[wp_ >> 5] >> (wp_ & 31)) & 1u);                             \
    const uint8_t flag = wflag_ ? p.flags[cell] : (uint8_t)0;                                              \
    const bool solid = flag & FLAG_SOLID;                                                                  \
    if (flag & FLAG_LINKS) link_pre_terms<L, ST>(p, B, link_record(p, cell), x, y, z);                    \
    __builtin_amdgcn_sched_barrier(0);                                                                     \
    C_ g_[THERMAL ? L::Q : 1];                                                                             \
    if constexpr (THERMAL && ScalarPrefetch<L, COLL, ST>::value) scalar_pull<L, ST>(p, x, y, z, g_);       \
    C_ fn[L::Q];                                                                                           \
    const int rec = post_bc<L, ST>(p, dyn, B, cell, x, y, z, flag, true, fn, pflag_, orow_.pv, orow_.coop,  \
                                  orow_.src);                                                              \
    if constexpr (THERMAL && !ScalarPrefetch<L, COLL, ST>::value) scalar_pull<L, ST>(p, x, y, z, g_);      \
    if constexpr (THERMAL) scalar_apply<L, ST>(p, dyn, cell, x, y, z, solid, fn, g_, true);                \
    if (rec >= 0) link_post_sum<L, C_>(p, rec, fn);                                                        \
    /* the next step's IBM forcing acts on this post-boundary state (solver.py:280-285) */                 \
    if (p.ibm_on && (flag & FLAG_IBM)) ibm_force<L, C_>((C_)p.ibm_coeff, fn);                             \
    finish_cell<L, COLL, ST, KMODE>(p, cell, sb_, solid, fn, dyn.rstep);                                   \
  } while (0)

// One case, parameters by value (split LES step, recycling inlet).
template <class L, int COLL, int ST, bool THERMAL, int KMODE>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(StepWaves<L, COLL, ST, THERMAL>::value)))
void step_kernel(StepParams p) {
  const StepDyn dyn = dyn_of(p);
  const int blk = blockIdx.x;
  AERO_STEP_BODY(L, COLL, ST, THERMAL, KMODE);
}

// Step k of a chunk for every case of a packed tile list: tile b is block
// `y` of case `x` (block 0 of a case reduces its previous step's forces).
// One workgroup per tile: a loop over several tiles per workgroup around this
// body made the register allocator spill (96-108 bytes per lane).
// Each case's parameters sit in device memory for both parities of k (the
// ping-pong buffers swap every step); the per-step fields are derived here.
// A single case runs through this same kernel as a batch of one, so a case
// computes the same bits alone or among others.
// KMODE 0: the plain collision; 2: with the local Smagorinsky rates.
template <class L, int COLL, int ST, bool THERMAL, int KMODE = 0>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(StepWaves<L, COLL, ST, THERMAL, KMODE>::value)))
void batch_step_kernel(const BatchCase* __restrict__ cases, const int2* __restrict__ tiles, int k) {
  const int2 t = tiles[blockIdx.x];
  const BatchCase& bc = cases[t.x];
  const StepParams& p = bc.p[k & 1];
  const StepDyn dyn{0.17 * (double)(bc.step0 + k + 1), bc.planes ? bc.planes + (size_t)3 * p.nrow * k : nullptr,
                    bc.hist + (size_t)NCOEF * (k > 0 ? k - 1 : 0), k == 0 ? bc.first_g_skip : 0,
                    (k > 0 && p.nrec > 0) ? 1 : 0, (unsigned)(bc.step0 + k + 1)};
  const int blk = t.y;
  AERO_STEP_BODY(L, COLL, ST, THERMAL, KMODE);
}

template <class L, int COLL, int ST, bool THERMAL, int KMODE = 0>
hipError_t launch_batch(const BatchCase* cases, const int2* tiles, int ntiles, int k, hipStream_t s, void* ev_start,
                        void* ev_stop) {
  const dim3 grid((unsigned)ntiles);
  if (ev_start)
    hipExtLaunchKernelGGL(batch_step_kernel<L, COLL, ST, THERMAL, KMODE>, grid, dim3(BLOCK), 0, s,
                          (hipEvent_t)ev_start, (hipEvent_t)ev_stop, 0, cases, tiles, k);
  else
    batch_step_kernel<L, COLL, ST, THERMAL, KMODE><<<grid, BLOCK, 0, s>>>(cases, tiles, k);
  return hipGetLastError();
}

// One step = one launch of nblocks + 1 blocks; with timing events set they
// are recorded by the packet processor at the kernel's start and end.
template <class L, int COLL, int ST, bool THERMAL, int KMODE>
hipError_t launch_step(const StepParams& p, hipStream_t s) {
  const dim3 grid((unsigned)(p.sblocks + 1));
  if (p.ev_start)
    hipExtLaunchKernelGGL(step_kernel<L, COLL, ST, THERMAL, KMODE>, grid, dim3(BLOCK), 0, s,
                          (hipEvent_t)p.ev_start, (hipEvent_t)p.ev_stop, 0, p);
  else
    step_kernel<L, COLL, ST, THERMAL, KMODE><<<grid, BLOCK, 0, s>>>(p);
  return hipGetLastError();
}

// Side-effect-free re-derivation of the reference's post-boundary state of
// the last completed step (reads the previous buffer side).  out_mode 0
// writes the f64 field, 1 writes rho/ux/uy/uz, 2 reduces the stability check.
template <class L, int ST, bool THERMAL>
__device__ __forceinline__ void materialize_block(const StepParams& p, int lb, double (*red)[NCHECK]) {
  using S = Storage<ST>;
  using C = typename S::C;
  using T = typename S::T;
  constexpr int Q = L::Q;
  const long long c64 = (long long)lb * BLOCK + threadIdx.x;
  double v_nan = 0, v_neg = 0, v_umax = 0, v_rho = 0, v_cnt = 0;
  if (c64 < p.ncell) {
    const int cell = (int)c64;
    int x, y, z;
    cell_xyz(p, (unsigned)cell, x, y, z);
    const bool in_plane = p.check_z < 0 || z == p.check_z;
    if (in_plane) {
      const uint8_t flag = p.flags[cell];
      C fn[Q];
      const StepDyn dyn = dyn_of(p);
      post_bc<L, ST>(p, dyn, (const T*)p.fin, cell, x, y, z, flag, false, fn);
      if constexpr (THERMAL) scalar_step<L, ST>(p, dyn, cell, x, y, z, flag & FLAG_SOLID, fn, false);
      const size_t N = (size_t)p.ncell;
      if (p.out_mode == 0) {
#pragma unroll
        for (int i = 0; i < Q; ++i) p.fmat[(size_t)i * N + cell] = (double)fn[i];
      } else {
        double rho = 0, mx = 0, my = 0, mz = 0;
        bool nan = false;
#pragma unroll
        for (int i = 0; i < Q; ++i) {
          const double v = (double)fn[i];
          nan |= isnan(v);
          rho += v;
          mx += L::EX[i] * v;
          my += L::EY[i] * v;
          mz += L::EZ[i] * v;
        }
        const double inv = rho > 0.0 ? 1.0 / rho : 0.0;
        if (p.out_mode == 1) {
          p.fmat[cell] = rho;
          p.fmat[N + cell] = inv * mx;
          p.fmat[2 * N + cell] = inv * my;
          p.fmat[3 * N + cell] = inv * mz;
        } else {
          // check_stability (diagnostics.py:32-71): |u| uses ux, uy only
          v_nan = nan ? 1.0 : 0.0;
          v_neg = rho <= 0.0 ? 1.0 : 0.0;
          const double ux = inv * mx, uy = inv * my;
          v_umax = sqrt(ux * ux + uy * uy);
          if (isnan(v_umax)) v_umax = 0.0;
          v_rho = rho;
          v_cnt = 1.0;
        }
      }
    }
  }
  if (p.out_mode == 2) {
    double v[NCHECK] = {v_nan, v_neg, v_umax, v_rho, v_cnt};
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      v[0] = fmax(v[0], __shfl_xor(v[0], off));
      v[1] = fmax(v[1], __shfl_xor(v[1], off));
      v[2] = fmax(v[2], __shfl_xor(v[2], off));
      v[3] += __shfl_xor(v[3], off);
      v[4] += __shfl_xor(v[4], off);
    }
    const int wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
      for (int k = 0; k < NCHECK; ++k) red[wv][k] = v[k];
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
      for (int k = 0; k < NCHECK; ++k) p.check_part[(size_t)lb * NCHECK + k] = r[k];
    }
  }
}

template <class L, int ST, bool THERMAL>
__global__ __launch_bounds__(BLOCK) void materialize_kernel(StepParams p) {
  __shared__ double red[BLOCK / 64][NCHECK];
  materialize_block<L, ST, THERMAL>(p, blockIdx.x, red);
}

// The stability check's block partials (out_mode 2) of many cases in one
// launch: tile b is block `y` of case `x`, whose parameters sit in device
// memory.  Each block computes what materialize_kernel's block does.
template <class L, int ST, bool THERMAL>
__global__ __launch_bounds__(BLOCK) void check_batch_kernel(const StepParams* __restrict__ ps,
                                                            const int2* __restrict__ tiles) {
  __shared__ double red[BLOCK / 64][NCHECK];
  const int2 t = tiles[blockIdx.x];
  materialize_block<L, ST, THERMAL>(ps[t.x], t.y, red);
}

// Collision from an explicit post-boundary field F (compute type): used after
// set_f and by the LES split step.  les: LES_OFF or a finite-difference
// model on U (les.py:273-344).
struct CollideParams {
  int nx, ny, nz, dim;
  int xp, qxp;  // storage row layout of fout (as StepParams)
  long long ncell;
  int nblocks;
  const void* F;
  const void* U;
  void* fout;
  const uint8_t* flags;
  const double* damp;
  const double* minv_d;
  const float* minv_f;
  double mrt_s[19];
  double omega, trt_sm, trt_lambda, base_nu, les_cs, u0;
  float fsh[NSHIFT];  // storage shift of f by fsh_class (half forms), as in StepParams
  int les;
  int trt_variant;
};

template <class C>
__device__ __forceinline__ C les_omega_fd(const CollideParams& p, const C* U, int cell, int x, int y, int z,
                                          bool solid) {
  const size_t N = (size_t)p.ncell;
  C nu_sgs = C(0);
  bool interior = x >= 1 && x <= p.nx - 2 && y >= 1 && y <= p.ny - 2;
  if (p.dim == 3) interior = interior && z >= 1 && z <= p.nz - 2;
  if (interior && !solid) {
    const C csl = p.damp ? (C)(p.les_cs * p.damp[cell]) : (C)p.les_cs;
    C g[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    const size_t sy = (size_t)p.nx, sz = (size_t)p.nx * p.ny;
    for (int a = 0; a < p.dim; ++a) {
      const C* u = U + (size_t)a * N;
      g[a][0] = C(0.5) * (u[cell + 1] - u[cell - 1]);
      g[a][1] = C(0.5) * (u[cell + sy] - u[cell - sy]);
      if (p.dim == 3) g[a][2] = C(0.5) * (u[cell + sz] - u[cell - sz]);
    }
    if (p.les == LES_FD_SMAG) {
      C smag;
      if (p.dim == 2) {
        const C s11 = g[0][0], s22 = g[1][1], s12 = C(0.5) * (g[0][1] + g[1][0]);
        smag = sqrt(C(2) * (s11 * s11 + s22 * s22 + C(2) * s12 * s12));
      } else {
        const C s11 = g[0][0], s22 = g[1][1], s33 = g[2][2];
        const C s12 = C(0.5) * (g[0][1] + g[1][0]);
        const C s13 = C(0.5) * (g[0][2] + g[2][0]);
        const C s23 = C(0.5) * (g[1][2] + g[2][1]);
        smag = sqrt(C(2) * (s11 * s11 + s22 * s22 + s33 * s33) + C(4) * (s12 * s12 + s13 * s13 + s23 * s23));
      }
      nu_sgs = (csl * csl) * smag;
    } else {
      C s_sq, sd_sq;
      if (p.dim == 2) {
        const C g11 = g[0][0], g12 = g[0][1], g21 = g[1][0], g22 = g[1][1];
        const C s12 = C(0.5) * (g12 + g21);
        s_sq = g11 * g11 + g22 * g22 + C(2) * s12 * s12;
        const C q11 = g11 * g11 + g12 * g21, q12 = g11 * g12 + g12 * g22;
        const C q21 = g21 * g11 + g22 * g21, q22 = g21 * g12 + g22 * g22;
        const C tr = C(0.5) * (q11 + q22);
        const C d11 = q11 - tr, d22 = q22 - tr, d12 = C(0.5) * (q12 + q21);
        sd_sq = d11 * d11 + d22 * d22 + C(2) * d12 * d12;
      } else {
        C sq = 0;
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b) {
            const C sab = C(0.5) * (g[a][b] + g[b][a]);
            sq += sab * sab;
          }
        s_sq = sq;
        C g2[3][3];
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b) g2[a][b] = g[a][0] * g[0][b] + g[a][1] * g[1][b] + g[a][2] * g[2][b];
        const C tr = (g2[0][0] + g2[1][1] + g2[2][2]) / C(3);
        C sd = 0;
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b) {
            const C v = C(0.5) * (g2[a][b] + g2[b][a]) - (a == b ? tr : C(0));
            sd += v * v;
          }
        sd_sq = sd;
      }
      const C num = pow(sd_sq, C(1.5));
      const C den = pow(s_sq, C(2.5)) + pow(sd_sq, C(1.25)) + C(1e-30);
      nu_sgs = (csl * csl) * num / den;
    }
  }
  const C nu_eff = solid ? (C)p.base_nu : (C)p.base_nu + nu_sgs;
  return C(1) / (C(3) * nu_eff + C(0.5));
}

template <class L, int COLL, int ST>
__global__ __launch_bounds__(BLOCK) void collide_kernel(CollideParams p) {
  using S = Storage<ST>;
  using C = typename S::C;
  using T = typename S::T;
  constexpr int Q = L::Q;
  const long long c64 = (long long)blockIdx.x * BLOCK + threadIdx.x;
  if (c64 >= p.ncell) return;
  const int cell = (int)c64;
  const size_t N = (size_t)p.ncell;
  const C* F = (const C*)p.F;
  C fn[Q], out[Q];
#pragma unroll
  for (int i = 0; i < Q; ++i) fn[i] = F[(size_t)i * N + cell];
  const bool solid = p.flags[cell] & FLAG_SOLID;
  C om = (C)p.omega, sm = (C)p.trt_sm;
  if (p.les == LES_LOCAL_SMAG && COLL != COLL_MRT) {
    // the first collision of a state set from the host, as the fused local step does it
    const double cs = p.damp ? p.les_cs * p.damp[cell] : p.les_cs;  // C^2 rounded as the step's table
    const LesLocal<C> la{(C)(cs * cs), (C)(1.0 / p.omega), (C)p.trt_lambda, p.trt_variant};
    collide<L, COLL, C, true>(p.mrt_s, p.minv_d, p.minv_f, solid, om, sm, fn, out, la);
  } else {
    if (p.les != LES_OFF && COLL != COLL_MRT) {
      const int x = cell % p.nx, yz = cell / p.nx, y = yz % p.ny, z = yz / p.ny;
      om = les_omega_fd<C>(p, (const C*)p.U, cell, x, y, z, solid);
      if (COLL == COLL_TRT) {  // per-cell s- from the local omega (kernels_trt.py:14-18)
        const C tau = C(1) / om;
        const C lam = (C)p.trt_lambda;
        if (p.trt_variant == 1) {
          sm = C(1) / (lam / (tau - C(0.5)) + C(0.5));
        } else {
          const C tp = lam * tau + (C(1) - lam);
          const C tm = (tp * tau - C(0.5)) / (tp - C(0.5));
          sm = C(1) / tm;
        }
      }
    }
    collide<L, COLL, C>(p.mrt_s, p.minv_d, p.minv_f, solid, om, sm, fn, out);
  }
  const unsigned row = (unsigned)cell / (unsigned)p.nx;
  T* Bo = (T*)p.fout + (size_t)row * (unsigned)p.qxp;
  const unsigned x = (unsigned)cell - row * (unsigned)p.nx;
#pragma unroll
  for (int i = 0; i < Q; ++i) S::st(Bo + (unsigned)(i * p.xp), x, out[i], (C)p.fsh[fsh_class<L>(i)]);
}

// IBM forcing applied in place to a compute-precision field (Q x N), for
// states set from the host before their first collision.
template <class L, class C>
__global__ __launch_bounds__(BLOCK) void ibm_field_kernel(long long n, C* __restrict__ F,
                                                          const uint8_t* __restrict__ flags, double coeff) {
  const long long c = (long long)blockIdx.x * BLOCK + threadIdx.x;
  if (c >= n || !(flags[c] & FLAG_IBM)) return;
  C fn[L::Q];
#pragma unroll
  for (int i = 0; i < L::Q; ++i) fn[i] = F[(size_t)i * n + c];
  ibm_force<L, C>((C)coeff, fn);
#pragma unroll
  for (int i = 0; i < L::Q; ++i) F[(size_t)i * n + c] = fn[i];
}

}  // namespace aero
