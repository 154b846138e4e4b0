// Synthetic-eddy-method inlet planes on the device (the fluctuation() of
// aero/lbm/sem_inlet.py:60-136).  The host advances the eddies (their random
// regeneration must follow the reference's generator draw for draw) and hands
// over one snapshot per step; the device then forms, for a whole chunk of
// steps at once:
//   1. the raw fields u, v, w = (1/sqrt N) sum_k eps_k max(0, 1 - r_k / sigma)
//      over the plane, eddies summed in index order as numpy's reduction
//      over the first axis does (sem_sum_kernel);
//   2. the divergence-free projection of (v, w) in Fourier space, with the
//      Nyquist row / column zeroed for even sizes (batched 2D FFTs, hipFFT,
//      around sem_project_kernel);
//   3. the scaling of u by Tu u0 / std(u) and of (v, w) by
//      Tu u0 / rms(v, w) (sem_scale_kernel, one block per plane).
// The arithmetic is FP64 without contraction, so step 1 matches the host
// formula bit for bit; the FFTs and the reductions agree with numpy's to
// rounding.
#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>

#include <cmath>
#include <stdexcept>
#include <string>

#include "launch.h"

#pragma clang fp contract(off)

namespace aero {

namespace {

constexpr int SEM_THREADS = 256;
constexpr int SEM_TILE = 512;  // eddies staged in LDS at a time

// raw fields of plane blockIdx.y: u into uraw, v and w (imaginary part 0)
// into the FFT buffer
__global__ __launch_bounds__(SEM_THREADS) void sem_sum_kernel(int nrow, int ny, int n, const double* __restrict__ ed,
                                                              double sigma, double inv_sqrt_n,
                                                              double* __restrict__ uraw, double2* __restrict__ cv) {
  __shared__ double s[6 * SEM_TILE];
  const int k = blockIdx.y;
  const double* e = ed + (size_t)k * 6 * n;
  const int cell = blockIdx.x * SEM_THREADS + threadIdx.x;
  const int y = cell % ny, z = cell / ny;
  const double yc = (double)y + 0.5, zc = (double)z + 0.5;
  double u = 0.0, v = 0.0, w = 0.0;
  for (int c0 = 0; c0 < n; c0 += SEM_TILE) {
    const int m = n - c0 < SEM_TILE ? n - c0 : SEM_TILE;
    __syncthreads();
    for (int t = threadIdx.x; t < 6 * m; t += SEM_THREADS) {
      const int comp = t / m, j = t - comp * m;
      s[comp * SEM_TILE + j] = e[(size_t)comp * n + c0 + j];
    }
    __syncthreads();
    if (cell < nrow) {
      for (int j = 0; j < m; ++j) {
        const double xk = s[j];
        const double dy = yc - s[SEM_TILE + j];
        const double dz = zc - s[2 * SEM_TILE + j];
        const double r = sqrt(xk * xk + dy * dy + dz * dz);
        const double fk = fmax(0.0, 1.0 - r / sigma);
        u += s[3 * SEM_TILE + j] * fk;
        v += s[4 * SEM_TILE + j] * fk;
        w += s[5 * SEM_TILE + j] * fk;
      }
    }
  }
  if (cell < nrow) {
    uraw[(size_t)k * nrow + cell] = inv_sqrt_n * u;
    cv[(size_t)(2 * k) * nrow + cell] = make_double2(inv_sqrt_n * v, 0.0);
    cv[(size_t)(2 * k + 1) * nrow + cell] = make_double2(inv_sqrt_n * w, 0.0);
  }
}

// 2*pi * numpy.fft.fftfreq(n)[i]
__device__ __forceinline__ double wavenumber(int i, int n) {
  const int s = i <= (n - 1) / 2 ? i : i - n;
  return (double)s * (1.0 / (double)n) * 2.0 * M_PI;
}

// (v, w) spectra of every plane: zero the Nyquist row / column of even sizes,
// then remove the divergent part, v -= ky / k^2 (ky v + kz w), likewise w
__global__ __launch_bounds__(SEM_THREADS) void sem_project_kernel(long long total, int nrow, int ny, int nz,
                                                                  double2* __restrict__ cv) {
  const long long i = (long long)blockIdx.x * SEM_THREADS + threadIdx.x;
  if (i >= total) return;
  const int k = (int)(i / nrow), cell = (int)(i - (long long)k * nrow);
  const int iy = cell % ny, iz = cell / ny;
  double2* V = cv + (size_t)(2 * k) * nrow + cell;
  double2* W = cv + (size_t)(2 * k + 1) * nrow + cell;
  double2 v = *V, w = *W;
  if ((nz % 2 == 0 && iz == nz / 2) || (ny % 2 == 0 && iy == ny / 2)) {
    v = make_double2(0.0, 0.0);
    w = make_double2(0.0, 0.0);
  }
  const double kz = wavenumber(iz, nz), ky = wavenumber(iy, ny);
  double k2 = kz * kz + ky * ky;
  if (iz == 0 && iy == 0) k2 = 1.0;
  const double dre = ky * v.x + kz * w.x, dim = ky * v.y + kz * w.y;
  const double ay = ky / k2, az = kz / k2;
  *V = make_double2(v.x - ay * dre, v.y - ay * dim);
  *W = make_double2(w.x - az * dre, w.y - az * dim);
}

__device__ __forceinline__ double block_sum(double v, double* red) {
  red[threadIdx.x] = v;
  __syncthreads();
  for (int s = SEM_THREADS / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  const double out = red[0];
  __syncthreads();
  return out;
}

// plane blockIdx.x: u scaled to the target intensity by its standard
// deviation, (v, w) = real part of the inverse transform (1/n of the
// unnormalised FFT) scaled by their joint rms
__global__ __launch_bounds__(SEM_THREADS) void sem_scale_kernel(int nrow, double target,
                                                                const double* __restrict__ uraw,
                                                                const double2* __restrict__ cv,
                                                                double* __restrict__ planes) {
  __shared__ double red[SEM_THREADS];
  const int k = blockIdx.x;
  const double* u = uraw + (size_t)k * nrow;
  const double2* v = cv + (size_t)(2 * k) * nrow;
  const double2* w = cv + (size_t)(2 * k + 1) * nrow;
  const double inv_n = 1.0 / (double)nrow;
  double a = 0.0;
  for (int c = threadIdx.x; c < nrow; c += SEM_THREADS) a += u[c];
  const double mean = block_sum(a, red) / (double)nrow;
  double b = 0.0, q = 0.0;
  for (int c = threadIdx.x; c < nrow; c += SEM_THREADS) {
    const double d = u[c] - mean;
    b += d * d;
    const double vv = v[c].x * inv_n, ww = w[c].x * inv_n;
    q += vv * vv + ww * ww;
  }
  const double u_std = sqrt(block_sum(b, red) / (double)nrow);
  const double vw_rms = sqrt(block_sum(q, red) / (double)nrow / 2.0);
  const double su = u_std > 1e-14 ? target / u_std : 0.0;
  const double sv = vw_rms > 1e-14 ? target / vw_rms : 0.0;
  double* P = planes + (size_t)k * 3 * nrow;
  for (int c = threadIdx.x; c < nrow; c += SEM_THREADS) {
    P[c] = u_std > 1e-14 ? u[c] * su : 0.0;
    P[nrow + c] = vw_rms > 1e-14 ? (v[c].x * inv_n) * sv : 0.0;
    P[2 * nrow + c] = vw_rms > 1e-14 ? (w[c].x * inv_n) * sv : 0.0;
  }
}

void hk(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}
void fk(hipfftResult r, const char* what) {
  if (r != HIPFFT_SUCCESS) throw std::runtime_error(std::string(what) + ": hipFFT error " + std::to_string((int)r));
}

}  // namespace

struct SemGen {
  int ny = 0, nz = 0, nrow = 0, n = 0, max_planes = 0;
  double sigma = 0.0, target = 0.0, inv_sqrt_n = 0.0;
  bool active = false;  // Tu > 0, sigma > 0 and n > 0 (otherwise every plane is zero)
  double* eddies = nullptr;  // max_planes x 6 x n
  double* uraw = nullptr;    // max_planes x nrow
  double2* cv = nullptr;     // max_planes x 2 x nrow
  hipfftHandle plan = 0;
  bool has_plan = false;
};

SemGen* sem_create(int ny, int nz, int n_eddies, double sigma, double Tu, double u0, int max_planes) {
  SemGen* g = new SemGen;
  try {
    g->ny = ny;
    g->nz = nz;
    g->nrow = ny * nz;
    g->n = n_eddies;
    g->max_planes = max_planes;
    g->sigma = sigma;
    g->target = Tu * u0;
    g->inv_sqrt_n = n_eddies > 0 ? 1.0 / std::sqrt((double)n_eddies) : 0.0;
    g->active = Tu > 0.0 && sigma > 0.0 && n_eddies > 0;
    if (g->active) {
      hk(hipMalloc(&g->eddies, sizeof(double) * 6 * (size_t)n_eddies * max_planes), "SEM eddies");
      hk(hipMalloc(&g->uraw, sizeof(double) * (size_t)g->nrow * max_planes), "SEM u field");
      hk(hipMalloc(&g->cv, sizeof(double2) * 2 * (size_t)g->nrow * max_planes), "SEM spectra");
      int dims[2] = {nz, ny};
      fk(hipfftPlanMany(&g->plan, 2, dims, nullptr, 1, g->nrow, nullptr, 1, g->nrow, HIPFFT_Z2Z, 2 * max_planes),
         "SEM FFT plan");
      g->has_plan = true;
    }
  } catch (...) {
    sem_destroy(g);
    throw;
  }
  return g;
}

void sem_destroy(SemGen* g) {
  if (!g) return;
  if (g->has_plan) (void)hipfftDestroy(g->plan);
  if (g->eddies) (void)hipFree(g->eddies);
  if (g->uraw) (void)hipFree(g->uraw);
  if (g->cv) (void)hipFree(g->cv);
  delete g;
}

void sem_planes(SemGen* g, const double* eddies_host, int nk, double* planes_dev, hipStream_t s) {
  if (nk <= 0) return;
  if (nk > g->max_planes) throw std::invalid_argument("SEM: more planes than the generator holds");
  if (!g->active) {
    hk(hipMemsetAsync(planes_dev, 0, sizeof(double) * 3 * (size_t)g->nrow * nk, s), "SEM zero planes");
    return;
  }
  hk(hipMemcpyAsync(g->eddies, eddies_host, sizeof(double) * 6 * (size_t)g->n * nk, hipMemcpyHostToDevice, s),
     "SEM upload eddies");
  const dim3 sg((unsigned)((g->nrow + SEM_THREADS - 1) / SEM_THREADS), (unsigned)nk);
  sem_sum_kernel<<<sg, SEM_THREADS, 0, s>>>(g->nrow, g->ny, g->n, g->eddies, g->sigma, g->inv_sqrt_n, g->uraw,
                                            g->cv);
  hk(hipGetLastError(), "SEM eddy sum");
  // the plan covers max_planes planes; unused ones transform stale data that
  // nothing reads
  fk(hipfftSetStream(g->plan, s), "SEM FFT stream");
  fk(hipfftExecZ2Z(g->plan, (hipfftDoubleComplex*)g->cv, (hipfftDoubleComplex*)g->cv, HIPFFT_FORWARD), "SEM FFT");
  const long long total = (long long)nk * g->nrow;
  sem_project_kernel<<<(unsigned)((total + SEM_THREADS - 1) / SEM_THREADS), SEM_THREADS, 0, s>>>(total, g->nrow, g->ny,
                                                                                               g->nz, g->cv);
  hk(hipGetLastError(), "SEM projection");
  fk(hipfftExecZ2Z(g->plan, (hipfftDoubleComplex*)g->cv, (hipfftDoubleComplex*)g->cv, HIPFFT_BACKWARD),
     "SEM inverse FFT");
  sem_scale_kernel<<<(unsigned)nk, SEM_THREADS, 0, s>>>(g->nrow, g->target, g->uraw, g->cv, planes_dev);
  hk(hipGetLastError(), "SEM scaling");
}

}  // namespace aero
