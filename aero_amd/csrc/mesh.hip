// STL geometry on the device: the parity ray-cast inside test and the exact
// point-to-triangle distance of the narrow-band signed-distance field.
//
// Both are brute force over every (point, triangle) pair in FP64, as the
// reference's numpy loops are (aero/geometry3d/mesh_mask.py:14-45,
// aero/geometry3d/signed_distance.py:38-137); here one thread owns one point
// and walks the triangle list, whose per-triangle constants every lane of a
// wave reads at the same address (scalar loads).  FMA contraction is off so
// each product and sum rounds as the reference's separate numpy operations.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "launch.h"

namespace aero {

constexpr int MESH_BLOCK = 256;
constexpr double MT_EPS = 1e-9;  // mesh_mask.py:25

// Moller-Trumbore constants of one triangle for the fixed ray direction.
struct TriMT {
  double v0[3], e1[3], e2[3], pv[3];
  double inv_det;
  int valid;  // |det| >= eps (a degenerate or ray-parallel triangle is never hit)
  int pad;
};

__global__ __launch_bounds__(MESH_BLOCK) void mt_prep_kernel(const double* __restrict__ tris, long long n, double rx,
                                                             double ry, double rz, TriMT* __restrict__ out) {
#pragma clang fp contract(off)
  const long long t = (long long)blockIdx.x * MESH_BLOCK + threadIdx.x;
  if (t >= n) return;
  const double* v = tris + 9 * t;
  TriMT m;
  for (int k = 0; k < 3; ++k) {
    m.v0[k] = v[k];
    m.e1[k] = v[3 + k] - v[k];
    m.e2[k] = v[6 + k] - v[k];
  }
  // pvec = ray x edge2
  m.pv[0] = ry * m.e2[2] - rz * m.e2[1];
  m.pv[1] = rz * m.e2[0] - rx * m.e2[2];
  m.pv[2] = rx * m.e2[1] - ry * m.e2[0];
  const double det = m.e1[0] * m.pv[0] + m.e1[1] * m.pv[1] + m.e1[2] * m.pv[2];
  m.valid = !(fabs(det) < MT_EPS);
  m.inv_det = 1.0 / det;
  m.pad = 0;
  out[t] = m;
}

// Point i of the query: an explicit (x, y, z) triple, or the centre
// (x + 0.5, y + 0.5, z + 0.5) of grid cell i (x fastest).
__device__ __forceinline__ void query_point(const double* pts, long long i, int nx, int ny, double p[3]) {
  if (pts) {
    p[0] = pts[3 * i];
    p[1] = pts[3 * i + 1];
    p[2] = pts[3 * i + 2];
  } else {
    const long long yz = i / nx;
    p[0] = (double)(i - yz * nx) + 0.5;
    p[1] = (double)(yz % ny) + 0.5;
    p[2] = (double)(yz / ny) + 0.5;
  }
}

// Parity of forward hits of the ray from each point (mesh_mask.py:40-45).
__global__ __launch_bounds__(MESH_BLOCK) void mesh_inside_kernel(const TriMT* __restrict__ T, long long ntri,
                                                                 const double* __restrict__ pts, long long i0,
                                                                 long long npts, int nx, int ny, double rx, double ry,
                                                                 double rz, uint8_t* __restrict__ out) {
#pragma clang fp contract(off)
  const long long i = i0 + (long long)blockIdx.x * MESH_BLOCK + threadIdx.x;
  if (i >= npts) return;
  double p[3];
  query_point(pts, i, nx, ny, p);
  unsigned parity = 0;
  for (long long t = 0; t < ntri; ++t) {
    const TriMT& m = T[t];
    if (!m.valid) continue;
    const double tx = p[0] - m.v0[0], ty = p[1] - m.v0[1], tz = p[2] - m.v0[2];
    const double u = (tx * m.pv[0] + ty * m.pv[1] + tz * m.pv[2]) * m.inv_det;
    // qvec = tvec x edge1
    const double qx = ty * m.e1[2] - tz * m.e1[1];
    const double qy = tz * m.e1[0] - tx * m.e1[2];
    const double qz = tx * m.e1[1] - ty * m.e1[0];
    const double v = (qx * rx + qy * ry + qz * rz) * m.inv_det;
    const double d = (qx * m.e2[0] + qy * m.e2[1] + qz * m.e2[2]) * m.inv_det;
    parity ^= (unsigned)((u >= 0.0) & (v >= 0.0) & (u + v <= 1.0) & (d > MT_EPS));
  }
  out[i] = (uint8_t)parity;
}

// Squared distance from each listed cell centre to the nearest triangle:
// the closest point of triangle (A, B, C) by the region of the minimiser of
// |A + s e0 + t e1 - P|^2 (Eberly, "Distance Between Point and Triangle in
// 3D", 1999), with the reference's degenerate-triangle rule
// (signed_distance.py:38-137).  Triangles are read raw (9 doubles each).
__device__ __forceinline__ double clamp01(double v) { return v < 0.0 ? 0.0 : (v > 1.0 ? 1.0 : v); }

__global__ __launch_bounds__(MESH_BLOCK) void mesh_sqdist_kernel(const double* __restrict__ tris, long long ntri,
                                                                 const long long* __restrict__ cells,
                                                                 long long ncells, int nx, int ny,
                                                                 double* __restrict__ out) {
#pragma clang fp contract(off)
  const long long k = (long long)blockIdx.x * MESH_BLOCK + threadIdx.x;
  if (k >= ncells) return;
  double P[3];
  query_point(nullptr, cells[k], nx, ny, P);
  double best = __builtin_inf();
  for (long long t = 0; t < ntri; ++t) {
    const double* v = tris + 9 * t;
    const double A[3] = {v[0], v[1], v[2]};
    const double e0[3] = {v[3] - v[0], v[4] - v[1], v[5] - v[2]};
    const double e1[3] = {v[6] - v[0], v[7] - v[1], v[8] - v[2]};
    const double df[3] = {A[0] - P[0], A[1] - P[1], A[2] - P[2]};
    const double a = e0[0] * e0[0] + e0[1] * e0[1] + e0[2] * e0[2];
    const double b = e0[0] * e1[0] + e0[1] * e1[1] + e0[2] * e1[2];
    const double c = e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2];
    const double d = df[0] * e0[0] + df[1] * e0[1] + df[2] * e0[2];
    const double e = df[0] * e1[0] + df[1] * e1[1] + df[2] * e1[2];
    const double det = a * c - b * b;
    double Q[3];
    if (det < 1e-14) {  // degenerate: the longer of the two edges from A
      const bool use0 = a >= c;
      const double s = clamp01(-(use0 ? d : e) / fmax(use0 ? a : c, 1e-30));
      const double* ed = use0 ? e0 : e1;
      for (int j = 0; j < 3; ++j) Q[j] = A[j] + s * ed[j];
    } else {
      const double s = b * e - c * d;
      const double tt = b * d - a * e;
      const bool beyond = (s + tt) > det, neg_s = s < 0.0, neg_t = tt < 0.0;
      if (neg_s && neg_t && !beyond) {  // vertex A
        for (int j = 0; j < 3; ++j) Q[j] = A[j];
      } else if (neg_s) {  // edge A-C
        const double r = clamp01(-e / fmax(c, 1e-30));
        for (int j = 0; j < 3; ++j) Q[j] = A[j] + r * e1[j];
      } else if (neg_t) {  // edge A-B
        const double r = clamp01(-d / fmax(a, 1e-30));
        for (int j = 0; j < 3; ++j) Q[j] = A[j] + r * e0[j];
      } else if (beyond) {  // edge B-C
        const double B[3] = {v[3], v[4], v[5]};
        const double bc[3] = {v[6] - v[3], v[7] - v[4], v[8] - v[5]};
        const double bc2 = bc[0] * bc[0] + bc[1] * bc[1] + bc[2] * bc[2];
        if (bc2 > 1e-14) {
          const double pb[3] = {P[0] - B[0], P[1] - B[1], P[2] - B[2]};
          const double r = clamp01((pb[0] * bc[0] + pb[1] * bc[1] + pb[2] * bc[2]) / bc2);
          for (int j = 0; j < 3; ++j) Q[j] = B[j] + r * bc[j];
        } else {
          for (int j = 0; j < 3; ++j) Q[j] = B[j];
        }
      } else {  // interior
        const double inv = 1.0 / det;
        const double ss = s * inv, ts = tt * inv;
        for (int j = 0; j < 3; ++j) Q[j] = A[j] + ss * e0[j] + ts * e1[j];
      }
    }
    const double dx = P[0] - Q[0], dy = P[1] - Q[1], dz = P[2] - Q[2];
    const double sq = dx * dx + dy * dy + dz * dz;
    best = sq < best ? sq : best;
  }
  out[k] = best;
}

// Launches cover at most this many points each, so no single kernel runs
// for long on large grids.
constexpr long long MESH_CHUNK = 1ll << 22;

hipError_t launch_mesh_inside(const double* tris, long long ntri, void* scratch, const double* pts, long long npts,
                              int nx, int ny, const double ray[3], uint8_t* out, hipStream_t s) {
  TriMT* T = (TriMT*)scratch;
  if (ntri > 0)
    mt_prep_kernel<<<dim3((unsigned)((ntri + MESH_BLOCK - 1) / MESH_BLOCK)), MESH_BLOCK, 0, s>>>(
        tris, ntri, ray[0], ray[1], ray[2], T);
  for (long long i0 = 0; i0 < npts; i0 += MESH_CHUNK) {
    const long long m = npts - i0 < MESH_CHUNK ? npts - i0 : MESH_CHUNK;
    mesh_inside_kernel<<<dim3((unsigned)((m + MESH_BLOCK - 1) / MESH_BLOCK)), MESH_BLOCK, 0, s>>>(
        T, ntri, pts, i0, i0 + m, nx, ny, ray[0], ray[1], ray[2], out);
  }
  return hipGetLastError();
}

size_t mesh_inside_scratch_bytes(long long ntri) { return sizeof(TriMT) * (size_t)(ntri > 0 ? ntri : 1); }

hipError_t launch_mesh_sqdist(const double* tris, long long ntri, const long long* cells, long long ncells, int nx,
                              int ny, double* out, hipStream_t s) {
  for (long long k0 = 0; k0 < ncells; k0 += MESH_CHUNK) {
    const long long m = ncells - k0 < MESH_CHUNK ? ncells - k0 : MESH_CHUNK;
    mesh_sqdist_kernel<<<dim3((unsigned)((m + MESH_BLOCK - 1) / MESH_BLOCK)), MESH_BLOCK, 0, s>>>(
        tris, ntri, cells + k0, m, nx, ny, out + k0);
  }
  return hipGetLastError();
}

}  // namespace aero
