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

// sin(2 pi r) and cos(2 pi r) for a phase r given in turns.  v_fract_f64
// drops the whole turns (exactly), rounding to the nearest quarter turn leaves
// |t| <= 1/8 (also exact), and the Taylor series of sin and cos through
// x^15 and x^16 on |x| = |2 pi t| <= pi/4 are below 5e-17 truncation error.
// This replaces the general sincos, whose large-argument reduction the
// ~1e5 rad projected phase would otherwise pull in.
__device__ __forceinline__ void sincos_turns(double r, double* s, double* c) {
  const double u = __builtin_amdgcn_fract(r);
  const double k = rint(4.0 * u);
  const double x = TPI * (u - 0.25 * k);
  const double x2 = x * x;
  const double ps =
      x2 * (-1.0 / 6.0 +
            x2 * (1.0 / 120.0 +
                  x2 * (-1.0 / 5040.0 +
                        x2 * (1.0 / 362880.0 +
                              x2 * (-1.0 / 39916800.0 + x2 * (1.0 / 6227020800.0 + x2 * (-1.0 / 1307674368000.0)))))));
  const double sn = x + x * ps;
  const double cs =
      1.0 + x2 * (-0.5 +
                  x2 * (1.0 / 24.0 +
                        x2 * (-1.0 / 720.0 +
                              x2 * (1.0 / 40320.0 +
                                    x2 * (-1.0 / 3628800.0 +
                                          x2 * (1.0 / 479001600.0 +
                                                x2 * (-1.0 / 87178291200.0 + x2 * (1.0 / 20922789888000.0))))))));
  const int q = ((int)k) & 3;
  // rotate (cos, sin) by q quarter turns
  const double s1 = (q & 1) ? cs : sn;
  const double c1 = (q & 1) ? sn : cs;
  *s = (q & 2) ? -s1 : s1;
  *c = ((q + 1) & 2) ? -c1 : c1;
}

__device__ __forceinline__ void kahan_add(float& s, float& comp, float x) {
  float y = x - comp;
  float t = s + y;
  comp = (t - s) - y;
  s = t;
}

// ------------------------------------------------- fused likelihood blocks
// A block works on one group of cpb points of one analysis (one wave per
// point, so every per-point constant is wave-uniform) and one split of the
// band.  It walks the LDS tiles of the split that any of its points needs:
// each tile (frequency tables and whitened data of every detector) is read
// from HBM once and used by all cpb waves.  Tiles sit on a fixed grid of the
// band and a lane's bins within a tile start at the point's own first bin,
// so a point's sums do not depend on the other points of its block.
// Per-lane sums, one wave reduction, one partial per (point, split).

// Bins of the split that the block's points need: [lo, hi), with lo rounded
// down to the tile grid.
__device__ __forceinline__ void block_range(const LLInst& in, int b, bool active, int j0, int j1, int tile,
                                            int* s_lohi, int* ks, int* ke, int* lo, int* hi) {
  const int lane = threadIdx.x & 63;
  *ks = 0;
  *ke = 0;
  if (active) {
    *ks = max(in.pts[b].kstart - in.imin, j0);
    *ke = min(in.pts[b].kend - in.imin, j1);
  }
  __syncthreads();  // s_lohi may still be read by the previous work item
  if (threadIdx.x == 0) { s_lohi[0] = j1; s_lohi[1] = j0; }
  __syncthreads();
  if (lane == 0 && active && *ke > *ks) {
    atomicMin(&s_lohi[0], *ks);
    atomicMax(&s_lohi[1], *ke);
  }
  __syncthreads();
  *lo = j0 + ((s_lohi[0] - j0) / tile) * tile;
  *hi = s_lohi[1];
}

template <int ND>
__device__ void ll_f64_block(const LLInst& in, int group, int split, int tile, double* lds, int* s_lohi) {
  constexpr int NP = 4 + 3 * ND;
  const int cpb = blockDim.x >> 6;
  const int wave = uniform(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int b = group * cpb + wave;
  const int j0 = split * in.split_len, j1 = min(in.nb, j0 + in.split_len);
  const bool active = b < in.B && in.pts[b].status == ST_OK;
  int ks, ke, blo, bhi;
  block_range(in, b, active, j0, j1, tile, s_lohi, &ks, &ke, &blo, &bhi);
  double acc[2 * ND];
#pragma unroll
  for (int q = 0; q < 2 * ND; q++) acc[q] = 0.0;
  // Detector d > 0 differs from detector 0 by the phase
  // theta_d(f) = 2 pi f (tshift_d - tshift_0) + lambda_d - lambda_0, linear in
  // f.  A lane steps 64 bins at a time, so exp(i theta_d) advances by the
  // constant rotation w_d = exp(i 2 pi 64 (tshift_d - tshift_0) / Tobs); the
  // recurrence is re-seeded directly at a lane's first bin of a run and
  // every 32 steps (drift below ~1e-14 relative).
  double rc[ND], rs[ND], wc[ND], ws[ND];
#pragma unroll
  for (int d = 0; d < ND; d++) { rc[d] = 1.0; rs[d] = 0.0; wc[d] = 1.0; ws[d] = 0.0; }
  if (active) {
    const Point& c = in.pts[b];
#pragma unroll
    for (int d = 1; d < ND; d++) sincos_turns(64.0 * in.inv_T * (c.tshift[d] - c.tshift[0]), &ws[d], &wc[d]);
  }
  int jnext = -1, age = 0;

  for (int t0 = blo; t0 < bhi; t0 += tile) {
    const int tn = min(tile, bhi - t0);
    __syncthreads();
    for (int i = threadIdx.x; i < tn; i += blockDim.x) {
#pragma unroll
      for (int p = 0; p < NP; p++) lds[p * tile + i] = in.p64[(size_t)p * in.nbp + t0 + i];
    }
    __syncthreads();
    const int lo = max(ks, t0), hi = min(ke, t0 + tn);
    if (active && lo < hi) {
      const Point& c = in.pts[b];
      for (int j = lo + lane; j < hi; j += 64) {
        const int i = j - t0;
        const double f = lds[i];
        if (j != jnext || age >= 32) {
#pragma unroll
          for (int d = 1; d < ND; d++)
            sincos_turns(f * (c.tshift[d] - c.tshift[0]) + (c.lambda[d] - c.lambda[0]) * INV_TPI, &rs[d], &rc[d]);
          age = 0;
        }
        Pow pw = pow_table(c, f, lds[tile + i], lds[2 * tile + i], lds[3 * tile + i]);
        const double a = amplitude(c, pw);
        const double ph = phase(c, pw);
        double s0, c0;
        // projected phase in turns: f (ts - td_0) + (lambda_0 - Psi) / 2 pi
        sincos_turns(fma(f, c.tshift[0], (c.lambda[0] - ph) * INV_TPI), &s0, &c0);
#pragma unroll
        for (int d = 0; d < ND; d++) {
          const double co = (d == 0) ? c0 : c0 * rc[d] - s0 * rs[d];
          const double sn = (d == 0) ? s0 : s0 * rc[d] + c0 * rs[d];
          const double A = c.Ascale[d] * a;
          const double hr = A * co, hi2 = A * sn;
          const double w = lds[(6 + 3 * d) * tile + i];
          acc[2 * d] += (hr * hr + hi2 * hi2) * w;
          acc[2 * d + 1] += lds[(4 + 3 * d) * tile + i] * hr + lds[(5 + 3 * d) * tile + i] * hi2;
          if (d > 0) {
            const double t = rc[d] * wc[d] - rs[d] * ws[d];
            rs[d] = rs[d] * wc[d] + rc[d] * ws[d];
            rc[d] = t;
          }
        }
        jnext = j + 64;
        age++;
      }
    }
  }
#pragma unroll
  for (int q = 0; q < 2 * ND; q++) acc[q] = wave_sum(acc[q]);
  if (lane == 0 && b < in.B) {
#pragma unroll
    for (int q = 0; q < 2 * ND; q++) in.partial[((size_t)b * in.splits + split) * (2 * ND) + q] = acc[q];
  }
}

// Mixed precision.  FP64: powers, amplitude and phase (gfx950 issues FP64
// FMAs at the FP32 rate, and the phase needs FP64), and the full argument
// 2 pi f (ts - td) + lambda - Psi reduced to a fraction of a turn
// (v_fract_f64), which keeps 2 pi f t_c + Psi ~ 1e5 rad exact to ~1e-11.
// FP32: sin/cos of the reduced turn (hardware v_sin/v_cos take turns), the
// products with the scaled data, and per-lane Kahan-compensated sums; the
// wave and split combine is FP64.
template <int ND>
__device__ void ll_mixed_block(const LLInst& in, int group, int split, int tile, double* lds, int* s_lohi) {
  float* lds32 = reinterpret_cast<float*>(lds + 4 * tile);
  const int cpb = blockDim.x >> 6;
  const int wave = uniform(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int b = group * cpb + wave;
  const int j0 = split * in.split_len, j1 = min(in.nb, j0 + in.split_len);
  const bool active = b < in.B && in.pts[b].status == ST_OK;
  int ks, ke, blo, bhi;
  block_range(in, b, active, j0, j1, tile, s_lohi, &ks, &ke, &blo, &bhi);
  float sum[2 * ND], comp[2 * ND];
#pragma unroll
  for (int q = 0; q < 2 * ND; q++) sum[q] = comp[q] = 0.0f;

  for (int t0 = blo; t0 < bhi; t0 += tile) {
    const int tn = min(tile, bhi - t0);
    __syncthreads();
    for (int i = threadIdx.x; i < tn; i += blockDim.x) {
#pragma unroll
      for (int p = 0; p < 4; p++) lds[p * tile + i] = in.p64[(size_t)p * in.nbp + t0 + i];
#pragma unroll
      for (int p = 0; p < 3 * ND; p++) lds32[p * tile + i] = in.p32[(size_t)p * in.nbp + t0 + i];
    }
    __syncthreads();
    const int lo = max(ks, t0), hi = min(ke, t0 + tn);
    if (active && lo < hi) {
      const Point& c = in.pts[b];
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
          kahan_add(sum[2 * d + 1], comp[2 * d + 1],
                    a * (lds32[(3 * d) * tile + i] * cs + lds32[(3 * d + 1) * tile + i] * sn));
        }
      }
    }
  }
  double acc[2 * ND];
#pragma unroll
  for (int q = 0; q < 2 * ND; q++) acc[q] = wave_sum((double)sum[q] - (double)comp[q]);
  if (lane == 0 && b < in.B) {
#pragma unroll
    for (int q = 0; q < 2 * ND; q++) in.partial[((size_t)b * in.splits + split) * (2 * ND) + q] = acc[q];
  }
}

template <int PREC>
__device__ __forceinline__ void ll_block(const LLInst& in, int group, int split, int tile, double* lds,
                                         int* s_lohi) {
  if (PREC == 0) {
    if (in.nd == 1) ll_f64_block<1>(in, group, split, tile, lds, s_lohi);
    else if (in.nd == 2) ll_f64_block<2>(in, group, split, tile, lds, s_lohi);
    else ll_f64_block<3>(in, group, split, tile, lds, s_lohi);
  } else {
    if (in.nd == 1) ll_mixed_block<1>(in, group, split, tile, lds, s_lohi);
    else if (in.nd == 2) ll_mixed_block<2>(in, group, split, tile, lds, s_lohi);
    else ll_mixed_block<3>(in, group, split, tile, lds, s_lohi);
  }
}

// One analysis: grid x = point groups, y = splits.
template <int ND, int PREC>
__global__ __launch_bounds__(512) void k_ll(LLInst in, int tile) {
  extern __shared__ double lds[];
  __shared__ int s_lohi[2];
  if (PREC == 0)
    ll_f64_block<ND>(in, blockIdx.x, blockIdx.y, tile, lds, s_lohi);
  else
    ll_mixed_block<ND>(in, blockIdx.x, blockIdx.y, tile, lds, s_lohi);
}

// Many analyses in one persistent launch: every block takes (analysis, point
// group, split) items from a queue until it is empty.  The exit test reads a
// block-uniform item index, so every wave of every block reaches it.
template <int PREC>
__global__ __launch_bounds__(512) void k_ll_queue(const LLInst* __restrict__ insts, const LLItem* __restrict__ items,
                                                  int n_items, int* __restrict__ counter, int tile) {
  extern __shared__ double lds[];
  __shared__ int s_lohi[2];
  __shared__ int s_item;
  for (;;) {
    __syncthreads();
    if (threadIdx.x == 0) s_item = atomicAdd(counter, 1);
    __syncthreads();
    const int it = s_item;
    if (it >= n_items) break;
    const LLItem w = items[it];
    const LLInst in = insts[w.inst];
    ll_block<PREC>(in, w.group, w.split, tile, lds, s_lohi);
  }
}

// ------------------------------------------------------------ finalize
// logL = sum_d [HD_d - HH_d/2], rho_d = HD_d / sqrt(HH_d) with HH = 4 sum |h|^2/S,
// HD = 4 sum Re(d h*)/S (QuickCBC.c:4259-4266, Utilities.c:716-736).  Splits are summed in index order, so a point's
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
    HH *= 4.0;  // fourier_nwip returns 4 sum (Utilities.c:734)
    HD *= 4.0;
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

hipError_t launch_ll(int precision, const LLInst& in, int cpb, int tile, hipStream_t s) {
  if (in.B <= 0) return hipSuccess;
  dim3 grid((in.B + cpb - 1) / cpb, in.splits), block(64 * cpb);
  size_t lds = ll_lds_bytes(precision, in.nd, tile);
  if (precision == 0) {
    switch (in.nd) {
      case 1: hipLaunchKernelGGL((k_ll<1, 0>), grid, block, lds, s, in, tile); break;
      case 2: hipLaunchKernelGGL((k_ll<2, 0>), grid, block, lds, s, in, tile); break;
      default: hipLaunchKernelGGL((k_ll<3, 0>), grid, block, lds, s, in, tile); break;
    }
  } else {
    switch (in.nd) {
      case 1: hipLaunchKernelGGL((k_ll<1, 1>), grid, block, lds, s, in, tile); break;
      case 2: hipLaunchKernelGGL((k_ll<2, 1>), grid, block, lds, s, in, tile); break;
      default: hipLaunchKernelGGL((k_ll<3, 1>), grid, block, lds, s, in, tile); break;
    }
  }
  return hipGetLastError();
}

hipError_t launch_ll_queue(int precision, const LLInst* d_insts, const LLItem* d_items, int n_items, int* d_counter,
                           int max_nd, int cpb, int tile, int blocks, hipStream_t s) {
  size_t lds = ll_lds_bytes(precision, max_nd, tile);
  if (precision == 0)
    hipLaunchKernelGGL((k_ll_queue<0>), dim3(blocks), dim3(64 * cpb), lds, s, d_insts, d_items, n_items, d_counter, tile);
  else
    hipLaunchKernelGGL((k_ll_queue<1>), dim3(blocks), dim3(64 * cpb), lds, s, d_insts, d_items, n_items, d_counter, tile);
  return hipGetLastError();
}

hipError_t launch_finalize(const double* partial, int splits, int nd, const Point* pts, int B, int mixed,
                           const double* wref, double* logL, double* rho, hipStream_t s) {
  if (B <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_finalize, dim3((B + 127) / 128), dim3(128), 0, s, partial, splits, nd, pts, B, mixed, wref, logL, rho);
  return hipGetLastError();
}

__global__ void k_fill_i32(int* p, int n, int v) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

hipError_t launch_fill_i32(int* p, int n, int v, hipStream_t s) {
  hipLaunchKernelGGL(k_fill_i32, dim3((n + 255) / 256), dim3(256), 0, s, p, n, v);
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

// Adjacent-temperature swaps: one thread per (chain column, temperature
// pair); even pairs on even steps, odd pairs on odd steps, so no two threads
// touch the same point.
__global__ void k_pt_swap(double* __restrict__ x, double* __restrict__ L, int n_temps, int n_chains, double ladder,
                          uint64_t seed, const unsigned long long* step, unsigned long long* __restrict__ swaps) {
  const unsigned long long s = *step;
  const int npairs = (n_temps - (int)(s & 1)) / 2;
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= npairs * n_chains) return;
  const int c = q % n_chains;
  const int t = (int)(s & 1) + 2 * (q / n_chains);
  const int a = t * n_chains + c, b = (t + 1) * n_chains + c;
  const double ba = pow(ladder, -(double)t), bb = pow(ladder, -(double)(t + 1));
  const double u = u01(seed ^ 0x27d4eb2fULL, s * (uint64_t)(n_temps * n_chains) + a, 7);
  if (log(u) < (ba - bb) * (L[b] - L[a])) {
    for (int i = 0; i < 11; i++) {
      double tmp = x[11 * a + i];
      x[11 * a + i] = x[11 * b + i];
      x[11 * b + i] = tmp;
    }
    double tl = L[a];
    L[a] = L[b];
    L[b] = tl;
    atomicAdd(&swaps[c], 1ULL);
  }
}

// Advance the step counter once every kernel of the step has read it.
__global__ void k_pt_tick(unsigned long long* step) {
  if (threadIdx.x == 0) *step += 1;
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
  int n = ((n_temps + 1) / 2) * n_chains;
  hipLaunchKernelGGL(k_pt_swap, dim3((n + 255) / 256), dim3(256), 0, s, x, L, n_temps, n_chains, ladder, seed, step,
                     swaps);
  hipLaunchKernelGGL(k_pt_tick, dim3(1), dim3(64), 0, s, step);
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

namespace qcbc {
// Plain device copy, 16 bytes per lane: the calibration reference for the HBM
// byte counters (a copy of n bytes reads n and writes n).
__global__ __launch_bounds__(256) void k_copy(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n16) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n16; i += (size_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

hipError_t launch_copy(const void* src, void* dst, size_t bytes, hipStream_t s) {
  size_t n16 = bytes / 16;
  hipLaunchKernelGGL(k_copy, dim3(2048), dim3(256), 0, s, (const uint4*)src, (uint4*)dst, n16);
  return hipGetLastError();
}
}  // namespace qcbc
