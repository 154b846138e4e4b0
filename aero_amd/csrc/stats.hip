// Convergence statistics of coefficient histories on the device, in FP64:
// the reference's auto-stop test (aero/diagnostics.py:87-120, 153-228) over
// the recent drag and lift samples of many cases at once.
//
// Per case, one workgroup forms the drag window's mean |Cd| and standard
// deviation, picks the lift series (3D: Clz when its standard deviation over
// the Strouhal window exceeds Cly's, aero/lbm/solver3d.py:634-638), and
// writes three mean-free signals: the Strouhal window and its two halves.
// One workgroup per (case, signal) then finds the peak of |rfft| over the
// positive frequencies with a direct DFT (twiddles from an LDS table of
// sincospi(2 j / n) for j <= n / 2, products indexed by (k j) mod n, so every
// twiddle is one correctly rounded value), keeping the first index of the
// maximum as np.argmax does.  A last pass combines them into the reference's decision.
#include <cmath>

#include "launch.h"

namespace aero {

namespace {

constexpr int STB = 256;

__device__ __forceinline__ double sample(const StatSeries& s, const double* ring, long long len, long long j) {
  // sample j (0 = oldest) of the newest `len` of the series
  const long long idx = s.count - len + j;
  return ring[s.cap > 0 ? idx % s.cap : idx];
}

template <class F>
__device__ double block_sum(F&& f, long long n, double* red) {
  double acc = 0.0;
  for (long long j = threadIdx.x; j < n; j += STB) acc += f(j);
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int h = STB / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  const double v = red[0];
  __syncthreads();
  return v;
}

__device__ double block_max(double v, double* red) {
  red[threadIdx.x] = v;
  __syncthreads();
  for (int h = STB / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + h]);
    __syncthreads();
  }
  const double r = red[0];
  __syncthreads();
  return r;
}

// Per case: the drag window's statistics, the lift choice, and the signals.
__global__ __launch_bounds__(STB) void stat_prep_kernel(const StatSeries* ss, StatParams sp, double* out,
                                                        double* work, int* siglen) {
  __shared__ double red[STB];
  const StatSeries s = ss[blockIdx.x];
  double* o = out + (size_t)blockIdx.x * NSTAT;
  const long long W = sp.conv_window;
  // analyze_convergence
  if (s.count < W) {
    if (threadIdx.x == 0) {
      o[1] = 0.0;
      o[2] = INFINITY;
      o[3] = 0.0;
      o[4] = 0.0;
    }
  } else {
    const double sum = block_sum([&](long long j) { return sample(s, s.cd, W, j); }, W, red);
    const double sabs = block_sum([&](long long j) { return fabs(sample(s, s.cd, W, j)); }, W, red);
    const double mean = sum / (double)W, mean_abs = sabs / (double)W;
    const double var = block_sum([&](long long j) {
                         const double d = sample(s, s.cd, W, j) - mean;
                         return d * d;
                       }, W, red) / (double)W;
    const double sd = sqrt(var);
    if (threadIdx.x == 0) {
      const bool tiny = mean_abs < 1e-12;
      const double ratio = tiny ? 0.0 : sd / mean_abs;
      o[1] = (tiny || ratio < sp.conv_tol) ? 1.0 : 0.0;
      o[2] = ratio;
      o[3] = mean_abs;
      o[4] = sd;
    }
  }
  // the lift series (3D: Cly, or Clz when its std over the window is larger)
  const long long SW = sp.st_window;
  const long long used = s.count < SW ? s.count : SW;
  const double* lift = s.ly;
  int which = 1;
  if (s.lz && used > 0) {
    auto sd_of = [&](const double* r) {
      const double m = block_sum([&](long long j) { return sample(s, r, used, j); }, used, red) / (double)used;
      return sqrt(block_sum([&](long long j) {
                    const double d = sample(s, r, used, j) - m;
                    return d * d;
                  }, used, red) / (double)used);
    };
    if (sd_of(s.lz) > sd_of(s.ly)) {
      lift = s.lz;
      which = 2;
    }
  }
  // signals: 0 the window, 1 its first half, 2 its second half (each minus its own mean)
  const long long half = used / 2;
  const long long off[3] = {0, 0, used - half}, len[3] = {used, used - half, half};
  for (int g = 0; g < 3; ++g) {
    double* w = work + ((size_t)blockIdx.x * 3 + g) * (size_t)SW;
    const long long n = len[g];
    double amp = 0.0;
    if (n > 0) {
      const double m = block_sum([&](long long j) { return sample(s, lift, used, off[g] + j); }, n, red) / (double)n;
      double mx = 0.0;
      for (long long j = threadIdx.x; j < n; j += STB) {
        const double v = sample(s, lift, used, off[g] + j) - m;
        w[j] = v;
        mx = fmax(mx, fabs(v));
      }
      amp = block_max(mx, red);
    }
    if (threadIdx.x == 0) {
      siglen[blockIdx.x * 3 + g] = (int)n;
      if (g == 0) {
        o[7] = amp;
        o[8] = (double)used;
      }
      o[12 + g] = amp;
    }
  }
  if (threadIdx.x == 0) o[11] = (double)which;
}

// Peak bin of |rfft(signal)| over k >= 1 for one (case, signal).
__global__ __launch_bounds__(STB) void stat_dft_kernel(const double* work, const int* siglen, int SW, int* peak) {
  // cos and sin of 2 pi j / n for j <= n / 2; the rest by symmetry
  // (cos(2 pi (n - j) / n) = cos, sin = -sin)
  extern __shared__ double tw[];
  __shared__ double best_v[STB];
  __shared__ int best_k[STB];
  const int sig = blockIdx.x;
  const int n = siglen[sig];
  const int h = n / 2;
  const double* x = work + (size_t)sig * SW;
  for (int j = threadIdx.x; j <= h; j += STB) {
    double sv, cv;
    sincospi(2.0 * (double)j / (double)n, &sv, &cv);
    tw[j] = cv;
    tw[h + 1 + j] = sv;
  }
  __syncthreads();
  double bv = -1.0;
  int bk = 0;
  const int nb = n / 2 + 1;
  for (int k = 1 + threadIdx.x; k < nb; k += STB) {
    double re = 0.0, im = 0.0;
    int idx = 0;
    for (int j = 0; j < n; ++j) {
      const bool lo = idx <= h;
      const int t = lo ? idx : n - idx;
      const double sv = tw[h + 1 + t];
      re = fma(x[j], tw[t], re);
      im = fma(-x[j], lo ? sv : -sv, im);
      idx += k;
      if (idx >= n) idx -= n;
    }
    const double m2 = re * re + im * im;
    if (m2 > bv) {  // k increases along the loop: keeps the first maximum
      bv = m2;
      bk = k;
    }
  }
  best_v[threadIdx.x] = bv;
  best_k[threadIdx.x] = bk;
  __syncthreads();
  for (int h = STB / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) {
      const double v2 = best_v[threadIdx.x + h];
      const int k2 = best_k[threadIdx.x + h];
      if (v2 > best_v[threadIdx.x] || (v2 == best_v[threadIdx.x] && k2 != 0 && (best_k[threadIdx.x] == 0 || k2 < best_k[threadIdx.x]))) {
        best_v[threadIdx.x] = v2;
        best_k[threadIdx.x] = k2;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) peak[sig] = n >= 2 ? best_k[0] : 0;
}

// The reference's decision (detect_statistical_stationarity).
__global__ void stat_final_kernel(const StatSeries* ss, int ncase, StatParams sp, const int* siglen, const int* peak,
                                  double* out) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncase) return;
  const StatSeries s = ss[c];
  double* o = out + (size_t)c * NSTAT;
  const bool conv = o[1] != 0.0;
  // analyze_strouhal of signal g with min_samples ms: St or NaN
  auto st_of = [&](int g, int ms) {
    const int n = siglen[c * 3 + g];
    if (s.u0 <= 0.0 || s.D <= 0.0 || n < (ms > 2 ? ms : 2)) return (double)NAN;
    if (o[12 + g] < 1e-6 || n / 2 + 1 <= 1) return (double)NAN;
    const double f = (double)peak[c * 3 + g] / (double)n;
    return f * s.D / s.u0;
  };
  const double st = st_of(0, 512);
  o[5] = st;
  o[6] = std::isnan(st) ? NAN : (double)peak[c * 3] / (double)siglen[c * 3];
  o[9] = NAN;
  o[10] = 0.0;
  double stop = conv ? 1.0 : 0.0;
  const long long need = sp.st_window > 1024 ? sp.st_window : 1024;
  if (!std::isnan(st) && (long long)siglen[c * 3] >= need) {
    const double a = st_of(1, 128), b = st_of(2, 128);
    if (!std::isnan(a) && !std::isnan(b)) {
      const double drift = fabs(b - a) / fmax(fabs(a), 1e-12);
      o[9] = drift;
      o[10] = drift < sp.st_tol ? 1.0 : 0.0;
      stop = (conv && drift < sp.st_tol) ? 1.0 : 0.0;
    }
  }
  o[0] = stop;
}

__global__ void ring_append_kernel(const RingAppend* ra, int n) {
  const RingAppend r = ra[blockIdx.y];
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= r.nk) return;
  const double* row = r.rows + (size_t)NCOEF * k;
  const long long pos = (r.start + k) % r.cap;
  r.ring[pos] = row[0];
  r.ring[r.cap + pos] = row[1];
  r.ring[2 * (size_t)r.cap + pos] = row[2];
}

}  // namespace

hipError_t launch_stationarity(const StatSeries* series, int n, const StatParams& sp, double* out, double* work,
                               int* siglen, int* peak, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  stat_prep_kernel<<<n, STB, 0, s>>>(series, sp, out, work, siglen);
  const size_t lds = sizeof(double) * 2 * (size_t)(sp.st_window / 2 + 1);
  stat_dft_kernel<<<3 * n, STB, lds, s>>>(work, siglen, sp.st_window, peak);
  stat_final_kernel<<<(n + 63) / 64, 64, 0, s>>>(series, n, sp, siglen, peak, out);
  return hipGetLastError();
}

hipError_t launch_ring_append(const RingAppend* ra, int n, int max_nk, hipStream_t s) {
  if (n <= 0 || max_nk <= 0) return hipSuccess;
  ring_append_kernel<<<dim3((max_nk + 255) / 256, n), 256, 0, s>>>(ra, n);
  return hipGetLastError();
}

}  // namespace aero
