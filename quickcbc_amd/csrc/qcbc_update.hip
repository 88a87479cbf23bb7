// QuickCBC's own PT step on the GPU, compiled with -ffp-contract=off: the
// proposal mix of update() (QuickCBC.c:2238-2405) and the step of MCMC_all
// (QuickCBC.c:2029-2092) for an ensemble of independent chain columns, each
// column one of the reference's temperature ladders.
//
// Per column and step s, as MCMC_all's iteration: with probability 0.2 (more
// than one temperature) one swap of a random adjacent pair of temperatures,
// otherwise update() at every temperature: the proposal y made at the end of
// step s - 1 (its heterodyned logL already evaluated) is accepted when
//   H = (logL_y - logL_x)/T + pMcM(y) + 3 lnDL_y - 3 lnDL_x - pMcM(x) > log u
// and y lies in the prior (bounds, eta <= 1/4, eta >= etamin), with
// pMcM = 2 ln Mt + ln eta - ln(1 - 4 eta)/2 the Jacobian that makes the prior
// flat in m1, m2 (:2345-2358).  On update steps with s % hist_every == 0 the
// column's states enter its history (:2071-2080).  Then the proposals of
// step s + 1, by update()'s thresholds on one uniform alpha:
//   alpha > a       Fisher jump: y = x + sqrt(T) ej[i] N(0,1) ev[i], i uniform in 0..10
//   alpha > b       differential evolution (de_jump, :2407-2425): y = x + g (h_i - h_j),
//                   i != j from the temperature's history, g = N(0, 0.5) with
//                   probability 0.9, else 1
//   alpha > c       jiggle: y = x + 0.01 10^-floor(3u) sqrt(T) N(0,1) per parameter
//   otherwise       the sky ring (Ring_all / skymap_all, :2298-2330), not built: c
//                   must be 0 (the reference's alternative setting, :2256-2260)
// then the angles wrapped into range (:2336-2342).
//
// Random numbers are counter-based (SplitMix64 of (seed, step, point,
// draw)), the point number p = t n_chains + c whatever the storage order,
// so every number of column groups and every graph replay gives the same
// chains, and the oracle replays them from the same counters.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "qcbc_update_internal.h"

namespace qcbc {

namespace {

constexpr double TPI_U = 6.283185307179586476925286766559;
constexpr double PI_U = 3.141592653589793238462643383279;

__device__ __forceinline__ uint64_t mix(uint64_t z) {
  z += 0x9e3779b97f4a7c15ULL;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}
// uniform in (0, 1) from the counters: 53 random bits, centred in their cell
__device__ __forceinline__ double uni(uint64_t seed, uint64_t salt, uint64_t a, uint64_t k) {
  const uint64_t r = mix((seed ^ salt) ^ mix(a * 0x2545F4914F6CDD1DULL + k));
  return ((double)(r >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}
// standard normal by Box-Muller from draws k and k + 1
__device__ __forceinline__ double gauss(uint64_t seed, uint64_t salt, uint64_t a, uint64_t k) {
  const double u1 = uni(seed, salt, a, k), u2 = uni(seed, salt, a, k + 1);
  return sqrt(-2.0 * log(u1)) * cos(TPI_U * u2);
}

// update()'s prior checks (:2345-2363) and the mass Jacobian; returns true
// when y is outside the prior
__device__ bool prior_flag(const double* y, const UpdateConst& u, double* pMcM) {
  bool flag = false;
  for (int i = 0; i < 11; i++)  // (a NaN fails too)
    if (!(y[i] <= u.max[i] && y[i] >= u.min[i])) flag = true;
  const double leta = (5.0 / 3.0) * (y[0] - y[1]);
  if (leta > log(0.25)) flag = true;
  *pMcM = 0.0;
  if (!flag) {
    const double eta = exp(leta);
    *pMcM = 2.0 * y[1] + leta - 0.5 * log(1.0 - 4.0 * eta);
    if (eta < u.etamin) flag = true;
  }
  return flag;
}

// the Jacobian term of a state already in the prior (eta <= 1/4)
__device__ double mass_jacobian(const double* x) {
  const double leta = (5.0 / 3.0) * (x[0] - x[1]);
  const double eta = exp(leta);
  return 2.0 * x[1] + leta - 0.5 * log(1.0 - 4.0 * eta);
}

// Component j of update()'s proposal for the state x at temperature T
// (sT = sqrt(T)) from the counters (seed, ctr); *typ gets the proposal type
// (1 Fisher, 2 DE, 3 jiggle).  One lane per component: the draws a proposal
// shares (its type, eigendirection, history rows, scale) are the same
// counters on every lane, so each component is computed exactly as a
// sequential loop over the components would compute it.
__device__ double propose_component(const double* __restrict__ x, const double* __restrict__ ej,
                                    const double* __restrict__ ev, const double* __restrict__ hist, int n_hist,
                                    double sT, const UpdateConst& u, uint64_t seed, uint64_t ctr, int j, int* typ) {
  const double alpha = uni(seed, SALT_PROPOSE, ctr, D_ALPHA);
  double y;
  if (alpha > u.a) {  // Fisher jump along an eigendirection
    *typ = 1;
    const int i = (int)(uni(seed, SALT_PROPOSE, ctr, D_FISHER) * 11.0);
    const double beta = sT * ej[i] * gauss(seed, SALT_PROPOSE, ctr, D_FISHER + 1);
    y = x[j] + beta * ev[11 * i + j];
  } else if (alpha > u.b) {  // differential evolution from the temperature's history
    *typ = 2;
    const int i = (int)((double)n_hist * uni(seed, SALT_PROPOSE, ctr, D_DE));
    int k = i;
    for (int r = 0; r < DE_TRIES && k == i; r++) k = (int)((double)n_hist * uni(seed, SALT_PROPOSE, ctr, D_DE + 1 + r));
    if (k == i) k = (i + 1) % n_hist;
    double g = 1.0;
    if (uni(seed, SALT_PROPOSE, ctr, D_DE_SCALE) < 0.9) g = 0.5 * gauss(seed, SALT_PROPOSE, ctr, D_DE_SCALE + 1);
    y = x[j] + g * (hist[11 * i + j] - hist[11 * k + j]);
  } else {  // jiggle (alpha > c; the sky ring below c is not built and c is 0)
    *typ = 3;
    const double beta = 0.01 * pow(10.0, -floor(3.0 * uni(seed, SALT_PROPOSE, ctr, D_JIGGLE))) * sT;
    y = x[j] + beta * gauss(seed, SALT_PROPOSE, ctr, D_JIGGLE + 1 + 2 * j);
  }
  // angles back into range (:2336-2342).  Only an angle within +-ANGLE_WRAP_MAX
  // rad is wrapped, so each loop ends within ~1600 turns; one beyond (or not
  // finite) stays out of its prior range and the proposal is rejected.
  if ((j == 4 || j == 7) && fabs(y) <= ANGLE_WRAP_MAX) {
    while (y > TPI_U) y -= TPI_U;
    while (y < 0.0) y += TPI_U;
  } else if (j == 9 && fabs(y) <= ANGLE_WRAP_MAX) {
    while (y > PI_U) y -= PI_U;
    while (y < 0.0) y += PI_U;
  }
  return y;
}

// lanes per point of the proposal loops (11 used)
constexpr int PROP_LANES = 16;

// The proposals of step s for every temperature of the group's column b
// (storage lp = t ncols + b): thread l of the caller's nthreads is lane
// l % PROP_LANES of temperature l / PROP_LANES, then of every nthreads /
// PROP_LANES-th temperature after it.
__device__ void propose_column(const double* __restrict__ x, const double* __restrict__ ej,
                               const double* __restrict__ ev, const double* __restrict__ hist,
                               const double* __restrict__ tc, const UpdateConst& u, int n_temps, int n_chains,
                               int c0, int ncols, int b, uint64_t seed, uint64_t s, int l, int nthreads,
                               double* __restrict__ y, int* __restrict__ typ) {
  const uint64_t P = (uint64_t)n_temps * n_chains;
  const int per = nthreads / PROP_LANES, sub = l / PROP_LANES, j = l % PROP_LANES;
  if (j >= 11) return;
  for (int t = sub; t < n_temps; t += per) {
    const int lp = t * ncols + b;
    const uint64_t gp = (uint64_t)t * n_chains + c0 + b;
    const double* h = hist + ((size_t)b * n_temps + t) * (size_t)u.n_hist * 11;
    int ty;
    const double yj = propose_component(x + 11 * lp, ej + 11 * lp, ev + 121 * lp, h, u.n_hist, tc[3 * t], u, seed,
                                        s * P + gp, j, &ty);
    y[11 * lp + j] = yj;
    if (j == 0) typ[lp] = ty;
  }
}

// Proposals of step *step for the group's points (used at creation and after
// a Fisher refresh; later steps get theirs from k_pt_update).  One block per
// column.
__global__ __launch_bounds__(512) void k_update_propose(const double* __restrict__ x, const double* __restrict__ ej,
                                                        const double* __restrict__ ev, const double* __restrict__ hist,
                                                        const double* __restrict__ tc, UpdateConst u, int n_temps,
                                                        int n_chains, int c0, int ncols, uint64_t seed,
                                                        const unsigned long long* __restrict__ step,
                                                        double* __restrict__ y, int* __restrict__ typ) {
  propose_column(x, ej, ev, hist, tc, u, n_temps, n_chains, c0, ncols, blockIdx.x, seed, *step, threadIdx.x,
                 blockDim.x, y, typ);
}

// One PT step of a group of chain columns after the heterodyned logL of the
// proposals (Ly, status st_y): block b owns column c0 + b.
__global__ __launch_bounds__(512) void k_pt_update(
    double* __restrict__ x, double* __restrict__ L, double* __restrict__ ej, double* __restrict__ ev,
    double* __restrict__ hist, unsigned long long* __restrict__ hcount, double* __restrict__ y,
    const double* __restrict__ Ly, const int* __restrict__ st_y, int* __restrict__ typ, const double* __restrict__ tc,
    UpdateConst u, int n_temps, int n_chains, int c0, int ncols, uint64_t seed, unsigned long long* __restrict__ step,
    unsigned long long* __restrict__ stats, unsigned int* __restrict__ done) {
  const int b = blockIdx.x, c = c0 + b;
  const uint64_t P = (uint64_t)n_temps * n_chains, s = *step;
  const uint64_t cs = s * (uint64_t)n_chains + c;  // the column's counter for this step
  // stats per column: [0] update steps, [1] swap steps, [2] swaps accepted,
  // [3 + 2 (typ - 1)] proposals of type typ, [4 + 2 (typ - 1)] accepted
  unsigned long long* cst = stats + (size_t)c * UPD_STATS;
  const bool swap_step = n_temps > 1 && uni(seed, SALT_SWAP, cs, 0) < 0.2;
  if (!swap_step) {
    for (int t = threadIdx.x; t < n_temps; t += blockDim.x) {
      const int lp = t * ncols + b;
      const uint64_t gp = (uint64_t)t * n_chains + c;
      const double* q = y + 11 * lp;
      double pMcMy;
      const bool flag = prior_flag(q, u, &pMcMy) || st_y[lp] != 0;
      const int k = typ[lp];
      atomicAdd(&cst[3 + 2 * (k - 1)], 1ULL);
      if (!flag) {
        const double* xp = x + 11 * lp;
        const double H = (Ly[lp] - L[lp]) * tc[3 * t + 1] + pMcMy + 3.0 * q[6] - 3.0 * xp[6] - mass_jacobian(xp);
        if (H > log(uni(seed, SALT_ACCEPT, s * P + gp, 0))) {
          for (int i = 0; i < 11; i++) x[11 * lp + i] = q[i];
          L[lp] = Ly[lp];
          atomicAdd(&cst[4 + 2 * (k - 1)], 1ULL);
        }
      }
    }
    __syncthreads();
    if (s % (uint64_t)u.hist_every == 0) {  // the column's states enter its history
      const uint64_t m = hcount[c];
      for (int t = threadIdx.x; t < n_temps; t += blockDim.x) {
        double* h = hist + (((size_t)b * n_temps + t) * u.n_hist + (size_t)(m % (uint64_t)u.n_hist)) * 11;
        for (int i = 0; i < 11; i++) h[i] = x[11 * (t * ncols + b) + i];
      }
      __syncthreads();
      if (threadIdx.x == 0) hcount[c] = m + 1;
    }
    if (threadIdx.x == 0) cst[0] += 1;
  } else {
    // one swap of temperatures j, j + 1 (:2038-2051)
    __shared__ int sw;
    if (threadIdx.x == 0) {
      const int j = (int)((double)(n_temps - 1) * uni(seed, SALT_SWAP, cs, 1));
      const double dL = L[j * ncols + b] - L[(j + 1) * ncols + b];
      const double beta = exp(dL * tc[3 * (j + 1) + 1] - dL * tc[3 * j + 1]);
      sw = beta > uni(seed, SALT_SWAP, cs, 2) ? j : -1;
      cst[1] += 1;
      if (sw >= 0) cst[2] += 1;
    }
    __syncthreads();
    if (sw >= 0) {
      const int la = sw * ncols + b, lb = (sw + 1) * ncols + b;
      for (int i = threadIdx.x; i < 11 + 11 + 121 + 1; i += blockDim.x) {
        double* pa;
        double* pb;
        if (i < 11) {
          pa = x + 11 * la + i;
          pb = x + 11 * lb + i;
        } else if (i < 22) {
          pa = ej + 11 * la + (i - 11);
          pb = ej + 11 * lb + (i - 11);
        } else if (i < 143) {
          pa = ev + 121 * la + (i - 22);
          pb = ev + 121 * lb + (i - 22);
        } else {
          pa = L + la;
          pb = L + lb;
        }
        const double tmp = *pa;
        *pa = *pb;
        *pb = tmp;
      }
    }
  }
  __syncthreads();
  // proposals of step s + 1
  propose_column(x, ej, ev, hist, tc, u, n_temps, n_chains, c0, ncols, b, seed, s + 1, threadIdx.x, blockDim.x, y,
                 typ);
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    if (atomicAdd(done, 1u) == gridDim.x - 1) {
      *step = s + 1;
      *done = 0u;
      __threadfence();
    }
  }
}

// The initial history of each column (MCMC_all, :1946-1954): row i of
// temperature t holds the initial state of a temperature q drawn uniformly
// from the column's.
__global__ void k_update_hist_init(const double* __restrict__ x, int n_temps, int n_chains, int c0, int ncols,
                                   int n_hist, uint64_t seed, double* __restrict__ hist) {
  const size_t n = (size_t)ncols * n_temps * n_hist;
  const size_t r = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const int i = (int)(r % n_hist);
  const int t = (int)((r / n_hist) % n_temps);
  const int b = (int)(r / ((size_t)n_hist * n_temps));
  const int c = c0 + b;
  const int q = (int)((double)n_temps * uni(seed, SALT_HIST, ((uint64_t)c * n_temps + t), (uint64_t)i));
  for (int k = 0; k < 11; k++) hist[r * 11 + k] = x[11 * (q * ncols + b) + k];
}

}  // namespace

// threads per column block: PROP_LANES per temperature, whole waves, at most 512
static int update_threads(int n_temps) {
  const int n = n_temps * PROP_LANES;
  return n >= 512 ? 512 : ((n + 63) / 64) * 64;
}

hipError_t launch_update_hist_init(const double* x, int n_temps, int n_chains, int c0, int ncols, int n_hist,
                                   uint64_t seed, double* hist, hipStream_t s) {
  const size_t n = (size_t)ncols * n_temps * n_hist;
  hipLaunchKernelGGL(k_update_hist_init, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, x, n_temps, n_chains, c0,
                     ncols, n_hist, seed, hist);
  return hipGetLastError();
}

hipError_t launch_update_propose(const double* x, const double* ej, const double* ev, const double* hist,
                                 const double* tc, const UpdateConst& u, int n_temps, int n_chains, int c0, int ncols,
                                 uint64_t seed, const unsigned long long* step, double* y, int* typ, hipStream_t s) {
  hipLaunchKernelGGL(k_update_propose, dim3(ncols), dim3(update_threads(n_temps)), 0, s, x, ej, ev, hist, tc, u,
                     n_temps, n_chains, c0, ncols, seed, step, y, typ);
  return hipGetLastError();
}

hipError_t launch_pt_update(double* x, double* L, double* ej, double* ev, double* hist, unsigned long long* hcount,
                            double* y, const double* Ly, const int* st_y, int* typ, const double* tc,
                            const UpdateConst& u, int n_temps, int n_chains, int c0, int ncols, uint64_t seed,
                            unsigned long long* step, unsigned long long* stats, unsigned int* done, hipStream_t s) {
  hipLaunchKernelGGL(k_pt_update, dim3(ncols), dim3(update_threads(n_temps)), 0, s, x, L, ej, ev, hist, hcount, y, Ly, st_y, typ, tc,
                     u, n_temps, n_chains, c0, ncols, seed, step, stats, done);
  return hipGetLastError();
}

}  // namespace qcbc
