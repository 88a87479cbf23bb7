// HIP kernels of the waveform + likelihood path, written for gfx950.
//
// Data layout of one analysis in HBM ("band planes"): band bins
// j = k - imin, k in [imin, imax), padded to nbp; FP64 planes
//   0: f_k = k/Tobs   1: f_k^(1/6)   2: f_k^(-1/6)   3: log f_k
//   4+3d: D_re/S   5+3d: D_im/S   6+3d: 1/S          (detector slot d)
// and, for the mixed-precision kernel, FP32 planes scaled by w_ref (per
// detector, so 1/S fits in FP32):
//   3d: D_re/(S sqrt(w_ref))   3d+1: D_im/(S sqrt(w_ref))   3d+2: 1/(S w_ref)
#include "qcbc_device.h"
#include "qcbc_internal.h"

namespace qcbc {

// ------------------------------------------------------------ utilities
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// ---------------------------------------------- fused FP64 likelihood
// Grid: x = group of CPB points (one wave per point), y = bin split.  Each
// block walks the LDS tiles of its bin range; the tile (frequency tables and
// whitened data of every detector) is loaded once and read by all CPB waves.
// Per bin: amplitude and phase from table powers, projection, and the two
// inner products per detector accumulated in FP64 registers; one wave
// reduction at the end, one partial per (point, split).
template <int ND>
__global__ __launch_bounds__(512) void k_ll_f64(const double* __restrict__ planes, int nbp, int imin, int nb,
                                                const Point* __restrict__ pts, int B, int tile,
                                                double* __restrict__ partial) {
  extern __shared__ double lds[];
  constexpr int NP = 4 + 3 * ND;
  __shared__ int s_lo, s_hi;
  const int cpb = blockDim.x >> 6;
  const int wave = uniform(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int b = blockIdx.x * cpb + wave;
  const int splits = gridDim.y, split = blockIdx.y;
  const int per = ((nb + splits - 1) / splits + 63) & ~63;
  const int j0 = split * per, j1 = min(nb, j0 + per);
  const bool active = b < B && pts[b].status == ST_OK;
  int ks = 0, ke = 0;
  if (active) {
    ks = max(pts[b].kstart - imin, j0);
    ke = min(pts[b].kend - imin, j1);
  }
  if (threadIdx.x == 0) { s_lo = j1; s_hi = j0; }
  __syncthreads();
  if (lane == 0 && active && ke > ks) {
    atomicMin(&s_lo, ks);
    atomicMax(&s_hi, ke);
  }
  __syncthreads();
  const int blo = s_lo, bhi = s_hi;
  double acc[2 * ND];
#pragma unroll
  for (int q = 0; q < 2 * ND; q++) acc[q] = 0.0;

  for (int t0 = blo; t0 < bhi; t0 += tile) {
    const int tn = min(tile, bhi - t0);
    __syncthreads();
    for (int i = threadIdx.x; i < tn; i += blockDim.x) {
#pragma unroll
      for (int p = 0; p < NP; p++) lds[p * tile + i] = planes[(size_t)p * nbp + t0 + i];
    }
    __syncthreads();
    const int lo = max(ks, t0), hi = min(ke, t0 + tn);
    if (active && lo < hi) {
      const Point& c = pts[b];
      for (int j = lo + lane; j < hi; j += 64) {
        const int i = j - t0;
        const double f = lds[i];
        Pow pw = pow_table(c, f, lds[tile + i], lds[2 * tile + i], lds[3 * tile + i]);
        const double a = amplitude(c, pw);
        const double ph = phase(c, pw);
        const double tf = TPI * f;
#pragma unroll
        for (int d = 0; d < ND; d++) {
          double s, co;
          sincos(tf * c.tshift[d] + c.lambda[d] - ph, &s, &co);
          const double A = c.Ascale[d] * a;
          const double hr = A * co, hi2 = A * s;
          const double w = lds[(6 + 3 * d) * tile + i];
          acc[2 * d] += (hr * hr + hi2 * hi2) * w;
          acc[2 * d + 1] += lds[(4 + 3 * d) * tile + i] * hr + lds[(5 + 3 * d) * tile + i] * hi2;
        }
      }
    }
  }
#pragma unroll
  for (int q = 0; q < 2 * ND; q++) acc[q] = wave_sum(acc[q]);
  if (lane == 0 && b < B) {
#pragma unroll
    for (int q = 0; q < 2 * ND; q++) partial[((size_t)b * splits + split) * (2 * ND) + q] = acc[q];
  }
}

// ------------------------------------------- fused mixed-precision likelihood
// Same walk as k_ll_f64.  FP64: powers, amplitude and phase, and the full
// argument 2 pi f (ts - td) + lambda - Psi reduced to a fraction of a turn
// (v_fract_f64), which keeps 2 pi f t_c + Psi ~ 1e5 rad exact to ~1e-11.
// FP32: sin/cos of the reduced turn (hardware v_sin/v_cos, which take turns),
// data products, and per-lane Kahan-compensated sums; FP64 wave/split combine.
__device__ __forceinline__ void kahan_add(float& s, float& comp, float x) {
  float y = x - comp;
  float t = s + y;
  comp = (t - s) - y;
  s = t;
}

template <int ND>
__global__ __launch_bounds__(512) void k_ll_mixed(const double* __restrict__ planes, const float* __restrict__ planes32,
                                                  int nbp, int imin, int nb, const Point* __restrict__ pts, int B,
                                                  int tile, double* __restrict__ partial) {
  extern __shared__ double lds[];
  float* lds32 = reinterpret_cast<float*>(lds + 4 * tile);
  __shared__ int s_lo, s_hi;
  const int cpb = blockDim.x >> 6;
  const int wave = uniform(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int b = blockIdx.x * cpb + wave;
  const int splits = gridDim.y, split = blockIdx.y;
  const int per = ((nb + splits - 1) / splits + 63) & ~63;
  const int j0 = split * per, j1 = min(nb, j0 + per);
  const bool active = b < B && pts[b].status == ST_OK;
  int ks = 0, ke = 0;
  if (active) {
    ks = max(pts[b].kstart - imin, j0);
    ke = min(pts[b].kend - imin, j1);
  }
  if (threadIdx.x == 0) { s_lo = j1; s_hi = j0; }
  __syncthreads();
  if (lane == 0 && active && ke > ks) {
    atomicMin(&s_lo, ks);
    atomicMax(&s_hi, ke);
  }
  __syncthreads();
  const int blo = s_lo, bhi = s_hi;
  float sum[2 * ND], comp[2 * ND];
#pragma unroll
  for (int q = 0; q < 2 * ND; q++) sum[q] = comp[q] = 0.0f;

  for (int t0 = blo; t0 < bhi; t0 += tile) {
    const int tn = min(tile, bhi - t0);
    __syncthreads();
    for (int i = threadIdx.x; i < tn; i += blockDim.x) {
#pragma unroll
      for (int p = 0; p < 4; p++) lds[p * tile + i] = planes[(size_t)p * nbp + t0 + i];
#pragma unroll
      for (int p = 0; p < 3 * ND; p++) lds32[p * tile + i] = planes32[(size_t)p * nbp + t0 + i];
    }
    __syncthreads();
    const int lo = max(ks, t0), hi = min(ke, t0 + tn);
    if (active && lo < hi) {
      const Point& c = pts[b];
      for (int j = lo + lane; j < hi; j += 64) {
        const int i = j - t0;
        const double f = lds[i];
        Pow pw = pow_table(c, f, lds[tile + i], lds[2 * tile + i], lds[3 * tile + i]);
        const float a = (float)amplitude(c, pw);
        const double phr = phase(c, pw) * INV_TPI;  // phase in turns
        const float a2 = a * a;
#pragma unroll
        for (int d = 0; d < ND; d++) {
          // turns: f (ts - td) + lambda/2pi - Psi/2pi, reduced to [0, 1)
          double q = fma(f, c.tshift[d], c.lambda[d] * INV_TPI - phr);
          float r = (float)__builtin_amdgcn_fract(q);
          float sn = __builtin_amdgcn_sinf(r), cs = __builtin_amdgcn_cosf(r);
          kahan_add(sum[2 * d], comp[2 * d], a2 * lds32[(3 * d + 2) * tile + i]);
          kahan_add(sum[2 * d + 1], comp[2 * d + 1], a * (lds32[(3 * d) * tile + i] * cs + lds32[(3 * d + 1) * tile + i] * sn));
        }
      }
    }
  }
  double acc[2 * ND];
#pragma unroll
  for (int q = 0; q < 2 * ND; q++) acc[q] = wave_sum((double)sum[q] - (double)comp[q]);
  if (lane == 0 && b < B) {
#pragma unroll
    for (int q = 0; q < 2 * ND; q++) partial[((size_t)b * splits + split) * (2 * ND) + q] = acc[q];
  }
}

// ------------------------------------------------------------ finalize
// logL = sum_d [HD_d - HH_d/2], rho_d = HD_d / sqrt(HH_d)
// (QuickCBC.c:4259-4266).  Splits are summed in index order, so a point's
// result does not depend on what else shares the launch.
__global__ void k_finalize(const double* __restrict__ partial, int splits, int nd, const Point* __restrict__ pts,
                           int B, int mixed, const double* __restrict__ wref, double* __restrict__ logL,
                           double* __restrict__ rho) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  const Point& c = pts[b];
  if (c.status != ST_OK) {
    logL[b] = 0.0;
    for (int d = 0; d < nd; d++) rho[(size_t)b * nd + d] = 0.0;
    return;
  }
  double L = 0.0;
  for (int d = 0; d < nd; d++) {
    double HH = 0.0, HD = 0.0;
    for (int s = 0; s < splits; s++) {
      HH += partial[((size_t)b * splits + s) * (2 * nd) + 2 * d];
      HD += partial[((size_t)b * splits + s) * (2 * nd) + 2 * d + 1];
    }
    if (mixed) {
      const double A = c.Ascale[d];
      HH *= (A * A) * wref[d];
      HD *= A * sqrt(wref[d]);
    }
    rho[(size_t)b * nd + d] = HD / sqrt(HH);
    L += HD - 0.5 * HH;
  }
  logL[b] = L;
}

// ------------------------------------------------------------ launchers
size_t ll_lds_bytes(int precision, int nd, int tile) {
  if (precision == 0) return (size_t)(4 + 3 * nd) * tile * sizeof(double);
  return (size_t)4 * tile * sizeof(double) + (size_t)3 * nd * tile * sizeof(float);
}

hipError_t launch_ll(int precision, int nd, const double* p64, const float* p32, int nbp, int imin, int nb,
                     const Point* pts, int B, int cpb, int tile, int splits, double* partial, hipStream_t s) {
  if (B <= 0) return hipSuccess;
  dim3 grid((B + cpb - 1) / cpb, splits), block(64 * cpb);
  size_t lds = ll_lds_bytes(precision, nd, tile);
  if (precision == 0) {
    switch (nd) {
      case 1: hipLaunchKernelGGL(k_ll_f64<1>, grid, block, lds, s, p64, nbp, imin, nb, pts, B, tile, partial); break;
      case 2: hipLaunchKernelGGL(k_ll_f64<2>, grid, block, lds, s, p64, nbp, imin, nb, pts, B, tile, partial); break;
      default: hipLaunchKernelGGL(k_ll_f64<3>, grid, block, lds, s, p64, nbp, imin, nb, pts, B, tile, partial); break;
    }
  } else {
    switch (nd) {
      case 1: hipLaunchKernelGGL(k_ll_mixed<1>, grid, block, lds, s, p64, p32, nbp, imin, nb, pts, B, tile, partial); break;
      case 2: hipLaunchKernelGGL(k_ll_mixed<2>, grid, block, lds, s, p64, p32, nbp, imin, nb, pts, B, tile, partial); break;
      default: hipLaunchKernelGGL(k_ll_mixed<3>, grid, block, lds, s, p64, p32, nbp, imin, nb, pts, B, tile, partial); break;
    }
  }
  return hipGetLastError();
}

hipError_t launch_finalize(const double* partial, int splits, int nd, const Point* pts, int B, int mixed,
                           const double* wref, double* logL, double* rho, hipStream_t s) {
  if (B <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_finalize, dim3((B + 127) / 128), dim3(128), 0, s, partial, splits, nd, pts, B, mixed, wref, logL, rho);
  return hipGetLastError();
}

// ------------------------------------------------------ PT ensemble step
// Counter-based random numbers: SplitMix64 finaliser of (seed, step, point,
// draw), so a step's draws depend only on its counters (graph replays and
// reruns give the same chain).

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z += 0x9e3779b97f4a7c15ULL;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}
__device__ __forceinline__ double u01(uint64_t seed, uint64_t a, uint64_t b) {
  uint64_t r = mix64(seed ^ mix64(a * 0x2545F4914F6CDD1DULL + b));
  return ((r >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

// Gaussian jiggle proposal (QuickCBC.c:2290-2291 shape): x' = x + sqrt(T) s z.
__global__ void k_pt_propose(const double* __restrict__ x, double* __restrict__ y, const double* __restrict__ sigma,
                             int n_chains, int P, double ladder, uint64_t seed, const unsigned long long* step) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  const uint64_t st = *step;
  const double T = pow(ladder, (double)(p / n_chains));
  const double sT = sqrt(T);
  for (int i = 0; i < 11; i += 2) {
    double u1 = u01(seed, st * (uint64_t)P + p, 2 * i), u2 = u01(seed, st * (uint64_t)P + p, 2 * i + 1);
    double r = sqrt(-2.0 * log(u1));
    double s, c;
    sincospi(2.0 * u2, &s, &c);
    y[11 * p + i] = x[11 * p + i] + sT * sigma[i] * r * c;
    if (i + 1 < 11) y[11 * p + i + 1] = x[11 * p + i + 1] + sT * sigma[i + 1] * r * s;
  }
}

__device__ __forceinline__ bool in_prior(const double* q) {
  return q[2] > -1.0 && q[2] < 1.0 && q[3] > -1.0 && q[3] < 1.0 && q[8] > -1.0 && q[8] < 1.0 && q[10] > -1.0 &&
         q[10] < 1.0 && q[0] < q[1] + (-0.8317766166719343);  // eta <= 1/4: ln Mc <= ln Mt + 0.6 ln(1/4)
}

// Metropolis-Hastings at inverse temperature 1/T; flat priors inside bounds.
__global__ void k_pt_accept(double* __restrict__ x, double* __restrict__ L, const double* __restrict__ y,
                            const double* __restrict__ Ly, const int* __restrict__ st_y, int n_chains, int P,
                            double ladder, uint64_t seed, const unsigned long long* step,
                            unsigned long long* __restrict__ acc) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  const double beta = 1.0 / pow(ladder, (double)(p / n_chains));
  const double u = u01(seed ^ 0x5bd1e995ULL, *step * (uint64_t)P + p, 99);
  const double* q = y + 11 * p;
  if (st_y[p] == 0 && in_prior(q) && log(u) < beta * (Ly[p] - L[p])) {
    for (int i = 0; i < 11; i++) x[11 * p + i] = q[i];
    L[p] = Ly[p];
    acc[p] += 1;
  }
}

// Adjacent-temperature swaps for every chain column; even pairs on even
// steps, odd pairs on odd steps.  Advances the step counter at the end.
__global__ void k_pt_swap(double* __restrict__ x, double* __restrict__ L, int n_temps, int n_chains, double ladder,
                          uint64_t seed, unsigned long long* step, unsigned long long* __restrict__ swaps) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  const unsigned long long s = *step;
  if (c < n_chains) {
    for (int t = (int)(s & 1); t + 1 < n_temps; t += 2) {
      int a = t * n_chains + c, b = (t + 1) * n_chains + c;
      double ba = 1.0 / pow(ladder, (double)t), bb = 1.0 / pow(ladder, (double)(t + 1));
      double u = u01(seed ^ 0x27d4eb2fULL, s * (uint64_t)(n_temps * n_chains) + a, 7);
      if (log(u) < (ba - bb) * (L[b] - L[a])) {
        for (int i = 0; i < 11; i++) {
          double tmp = x[11 * a + i];
          x[11 * a + i] = x[11 * b + i];
          x[11 * b + i] = tmp;
        }
        double tl = L[a];
        L[a] = L[b];
        L[b] = tl;
        swaps[c] += 1;
      }
    }
  }
  __syncthreads();
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    // every block has read *step before block 0 can finish only if the grid
    // is one block; launch_pt_swap guarantees that (n_chains <= 1024)
    *step = s + 1;
  }
}

hipError_t launch_pt_propose(const double* x, double* y, const double* sigma, int n_chains, int P, double ladder,
                             uint64_t seed, const unsigned long long* step, hipStream_t s) {
  hipLaunchKernelGGL(k_pt_propose, dim3((P + 127) / 128), dim3(128), 0, s, x, y, sigma, n_chains, P, ladder, seed, step);
  return hipGetLastError();
}

hipError_t launch_pt_accept(double* x, double* L, const double* y, const double* Ly, const int* st_y, int n_chains,
                            int P, double ladder, uint64_t seed, const unsigned long long* step,
                            unsigned long long* acc, hipStream_t s) {
  hipLaunchKernelGGL(k_pt_accept, dim3((P + 127) / 128), dim3(128), 0, s, x, L, y, Ly, st_y, n_chains, P, ladder, seed,
                     step, acc);
  return hipGetLastError();
}

hipError_t launch_pt_swap(double* x, double* L, int n_temps, int n_chains, double ladder, uint64_t seed,
                          unsigned long long* step, unsigned long long* swaps, hipStream_t s) {
  if (n_chains > 1024) return hipErrorInvalidValue;
  int threads = ((n_chains + 63) / 64) * 64;
  hipLaunchKernelGGL(k_pt_swap, dim3(1), dim3(threads), 0, s, x, L, n_temps, n_chains, ladder, seed, step, swaps);
  return hipGetLastError();
}

__global__ void k_fill_u64(unsigned long long* p, int n, unsigned long long v) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

hipError_t launch_fill_u64(unsigned long long* p, int n, unsigned long long v, hipStream_t s) {
  hipLaunchKernelGGL(k_fill_u64, dim3((n + 255) / 256), dim3(256), 0, s, p, n, v);
  return hipGetLastError();
}

}  // namespace qcbc
