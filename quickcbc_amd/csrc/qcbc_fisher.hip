// HIP kernels of the heterodyned Fisher matrix (Fisher_Het, QuickCBC.c:5446-5617)
// and its eigen-decomposition into jump directions (FisherEvec,
// QuickCBC.c:2449-2546), compiled with -ffp-contract=off: every per-sample
// product and every rotation is the reference's double arithmetic, operation
// for operation; only the sum over a heterodyne's samples is regrouped (one
// weight per sample, as in the heterodyned likelihood).
//
// A Fisher matrix needs fullphaseamp at 23 parameter sets per point: the
// point itself, then params[k] + 1e-5 and params[k] - 1e-5 for k = 0..10
// (variant 2k+1 and 2k+2).  k_fisher_params writes them, the ordinary set-up
// kernel turns them into records, and k_het_fisher takes one block per point:
// per chunk of 64 grid samples, the 23 variants' h22 amplitude and phase
// (one variant per wave), the 22 central differences per detector, then the
// 66 (i <= j) sums over the chunk's samples.
#include "qcbc_device.h"
#include "qcbc_fisher_internal.h"
#include "qcbc_het_device.h"

namespace qcbc {

namespace {

constexpr int NP = FISHER_NP, NV = FISHER_NV;
constexpr int NPAIR = NP * (NP + 1) / 2;  // 66 entries of the upper triangle
constexpr int NQ = 4;                     // each pair's sum split over four sample quarters
constexpr double EPS = 1.0e-5;            // Fisher_Het's epsilon (QuickCBC.c:5486)

// The 23 parameter sets of each point: q[(b NV + v) 11 + k]
__global__ void k_fisher_params(const double* __restrict__ p, int B, double* __restrict__ q) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= B * NV) return;
  const int b = i / NV, v = i % NV;
  double x[NP];
#pragma unroll
  for (int k = 0; k < NP; k++) x[k] = p[(size_t)b * NP + k];
  if (v > 0) {
    const int k = (v - 1) >> 1;
#pragma unroll
    for (int j = 0; j < NP; j++)
      if (j == k) x[j] = (v & 1) ? x[j] + EPS : x[j] - EPS;
  }
#pragma unroll
  for (int k = 0; k < NP; k++) q[(size_t)i * NP + k] = x[k];
}

// (i, j) of upper-triangle entry n in row order
__device__ __forceinline__ void pair_of(int n, int* i, int* j) {
  int r = 0;
  while (n >= NP - r) {
    n -= NP - r;
    r++;
  }
  *i = r;
  *j = r + n;
}

// Fisher_Het for B points, one 256-thread block each.  pts holds NV records
// per point (k_fisher_params' order), ha the reference amplitude (nd x M), w
// the raw evaluation weights, whose first nd x M are wH: sample m of region
// i enters HH with sum_j IP[i][j][k] SL[i][d][j], less aa[d][i] at the
// region starts 1 .. NR-2 (log_likelihood_het's and Fisher_Het's overcount
// correction).  The slow term per sample is the reference's
//   cc = 4 (dA_i dA_j + A^2 dphi_i dphi_j) / (a_ref a_ref),
// so fish[i][j] = sum_d sum_m wH cc, evaluated as sum (wH / a_ref^2) 4 (...):
// one division per sample instead of one per sample and entry.  With
// |chi1| or |chi2| >= 0.9999 the
// matrix is 1e4 I and nothing is evaluated (:5467-5469); a record the set-up
// rejected leaves 1e4 I and its status (the first failing variant's, in
// Fisher_Het's order of evaluation).
//
// The grid may be split over gridDim.y blocks per point (Fisher_All's full
// band): block (b, s) sums the chunks of samples [s len, (s + 1) len), len a
// multiple of 64, into partial[(b splits + s) 66 + n] (upper-triangle entry
// n in row order), and k_fisher_reduce adds the splits in order.  With one
// split the block writes the matrix itself.  ha null: a_ref = 1 (w is then
// the per-sample weight itself, as 1/S for the full band).
template <int ND>
__global__ __launch_bounds__(256) void k_het_fisher(const Point* __restrict__ pts, const double* __restrict__ params,
                                                    int B, const double* __restrict__ freq,
                                                    const double* __restrict__ pl, const double* __restrict__ ha,
                                                    const double* __restrict__ w, int M, double* __restrict__ fish,
                                                    int* __restrict__ status, int len,
                                                    double* __restrict__ partial) {
  __shared__ double V[NV][3][64];       // h22 amplitude, phase, taper per variant and sample
  __shared__ double U[ND][2 * NP][64];  // dA_k (k < 11) and dphi_k per detector and sample
  __shared__ double S[ND][2][64];       // A^2 and wH / a_ref^2 per detector and sample
  __shared__ double red[NQ][NPAIR];
  const int b = blockIdx.x, t = threadIdx.x;
  const int wave = __builtin_amdgcn_readfirstlane(t >> 6), lane = t & 63;
  if (b >= B) return;
  const Point* P = pts + (size_t)b * NV;
  int st = ST_OK;
  for (int v = 0; v < NV && st == ST_OK; v++) st = P[v].status;
  const bool spins = fabs(params[(size_t)b * NP + 2]) < 0.9999 && fabs(params[(size_t)b * NP + 3]) < 0.9999;
  double* F = fish + (size_t)b * NP * NP;
  const int split = blockIdx.y, splits = gridDim.y;
  if (!spins || st != ST_OK) {  // block-uniform
    if (split == 0) {
      if (t < NP * NP) F[t] = (t % (NP + 1) == 0) ? 1.0e4 : 0.0;
      if (t == 0) status[b] = spins ? st : ST_OK;
    }
    return;
  }
  const int m_begin = split * len, m_end = min(M, m_begin + len);
  // the pair sums this thread owns: items t and t + 256 of NQ x NPAIR
  int pi0, pj0, pi1 = 0, pj1 = 0;
  const int q0 = t / NPAIR, q1 = (t + 256) / NPAIR;
  pair_of(t % NPAIR, &pi0, &pj0);
  const bool two = t + 256 < NQ * NPAIR;
  if (two) pair_of((t + 256) % NPAIR, &pi1, &pj1);
  double acc0 = 0.0, acc1 = 0.0;
  for (int m0 = m_begin; m0 < m_end; m0 += 64) {
    // the 23 variants' h22 amplitude and phase at the chunk's samples
    for (int v = wave; v < NV; v += 4) {
      const Point& c = P[v];
      const int m = m0 + lane;
      double a = 0.0, p = 0.0, y = 0.0;
      if (m < m_end) {
        const double f = freq[m];
        h22_planes(c, 1.0 / c.Msec, f, pl, M, m, &a, &p);
        y = taper(f, c.fs);
      }
      V[v][0][lane] = a;
      V[v][1][lane] = p;
      V[v][2][lane] = y;
    }
    __syncthreads();
    // central differences (:5497-5512): per detector, 22 rows of 64 samples
    for (int r = wave; r < ND * NP; r += 4) {
      const int d = r / NP, k = r % NP, m = m0 + lane;
      double dA = 0.0, dp = 0.0;
      if (m < m_end) {
        const double f = freq[m];
        double aP, xP, aM, xM;
        project(P[2 * k + 1], d, f, V[2 * k + 1][0][lane], V[2 * k + 1][1][lane], V[2 * k + 1][2][lane], &aP, &xP);
        project(P[2 * k + 2], d, f, V[2 * k + 2][0][lane], V[2 * k + 2][1][lane], V[2 * k + 2][2][lane], &aM, &xM);
        aP -= aM;
        xP -= xM;
        dA = aP / (2.0 * EPS);
        dp = xP / (2.0 * EPS);
      }
      U[d][k][lane] = dA;
      U[d][NP + k][lane] = dp;
    }
    for (int d = wave; d < ND; d += 4) {
      const int m = m0 + lane;
      double A2 = 0.0, wh = 0.0;
      if (m < m_end) {
        double a0, x0;
        project(P[0], d, freq[m], V[0][0][lane], V[0][1][lane], V[0][2][lane], &a0, &x0);
        A2 = a0 * a0;
        wh = w[(size_t)d * M + m];
        if (ha) {
          const double h = ha[(size_t)d * M + m];
          wh = wh / (h * h);
        }
      }
      S[d][0][lane] = A2;
      S[d][1][lane] = wh;
    }
    __syncthreads();
    // the chunk's share of each owned entry: 16 samples of every detector
    {
      const int s0 = q0 * 16;
#pragma unroll 1
      for (int d = 0; d < ND; d++)
#pragma unroll 4
        for (int s = s0; s < s0 + 16; s++) {
          const double cc = 4.0 * (U[d][pi0][s] * U[d][pj0][s] + S[d][0][s] * U[d][NP + pi0][s] * U[d][NP + pj0][s]);
          acc0 += S[d][1][s] * cc;
        }
    }
    if (two) {
      const int s0 = q1 * 16;
#pragma unroll 1
      for (int d = 0; d < ND; d++)
#pragma unroll 4
        for (int s = s0; s < s0 + 16; s++) {
          const double cc = 4.0 * (U[d][pi1][s] * U[d][pj1][s] + S[d][0][s] * U[d][NP + pi1][s] * U[d][NP + pj1][s]);
          acc1 += S[d][1][s] * cc;
        }
    }
    __syncthreads();
  }
  red[q0][t % NPAIR] = acc0;
  if (two) red[q1][(t + 256) % NPAIR] = acc1;
  __syncthreads();
  if (t < NPAIR) {
    int i, j;
    pair_of(t, &i, &j);
    double x = red[0][t];
    for (int q = 1; q < NQ; q++) x += red[q][t];
    if (splits == 1) {
      F[i * NP + j] = x;
      F[j * NP + i] = x;
    } else {
      partial[((size_t)b * splits + split) * NPAIR + t] = x;
    }
  }
  if (t == 0 && split == 0) status[b] = ST_OK;
}

// The sum over the splits of k_het_fisher's partial entries, in split order,
// for the points whose matrix k_het_fisher evaluated (the others already
// hold 1e4 I).  One thread per point and entry.
__global__ void k_fisher_reduce(const Point* __restrict__ pts, const double* __restrict__ params, int B, int splits,
                                const double* __restrict__ partial, double* __restrict__ fish) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= B * NPAIR) return;
  const int b = i / NPAIR, n = i % NPAIR;
  const Point* P = pts + (size_t)b * NV;
  int st = ST_OK;
  for (int v = 0; v < NV && st == ST_OK; v++) st = P[v].status;
  const bool spins = fabs(params[(size_t)b * NP + 2]) < 0.9999 && fabs(params[(size_t)b * NP + 3]) < 0.9999;
  if (!spins || st != ST_OK) return;
  double x = 0.0;
  for (int sp = 0; sp < splits; sp++) x += partial[((size_t)b * splits + sp) * NPAIR + n];
  int r, c;
  pair_of(n, &r, &c);
  fish[(size_t)b * NP * NP + r * NP + c] = x;
  fish[(size_t)b * NP * NP + c * NP + r] = x;
}


// ---------------------------------------------------------------- Fisher_Full
// Fisher_Full (QuickCBC.c:2551-2688), MCMC_intrinsic's Fisher matrix of the
// seven reference-frame parameters over the band [imin, imax) of every bin:
// AmpPhase (:3395-3444) at the point and at params[k] +- 1e-5, k < 7, gives
// the bare h22 A = h22fac amp and phase P = 2 pi f ts - phase (no taper, no
// projection, ts = Tobs - params[5]), and
//   fish[i][j] = sum_l Scale_l sum_k (dA_i dA_j + A^2 dP_i dP_j) / S_l[k]
// with Scale_0 = 4 and Scale_l = 4 r_l^2 (r_l the amplitude ratio
// params[7 + 3 (l-1) + 2]); diagonal entries below 1 become 1.  With |chi1|
// or |chi2| >= 0.9999 the matrix is 1e4 I.  Here the detector sum is one
// weight per bin, sum_l Scale_l / S_l[k] (a regrouping of the reference's
// sums).  A d < 7 matrix of the reference is the top-left block of this one.
constexpr int FF_NP = FISHER_FULL_NP, FF_NV = 2 * FF_NP + 1, FF_NPAIR = FF_NP * (FF_NP + 1) / 2;

__device__ __forceinline__ void ff_pair_of(int n, int* i, int* j) {
  int r = 0;
  while (n >= FF_NP - r) {
    n -= FF_NP - r;
    r++;
  }
  *i = r;
  *j = r + n;
}

// the FF_NV parameter sets (seven entries each) of B points of ips entries
__global__ void k_fisher_full_params(const double* __restrict__ p, int ips, int B, double* __restrict__ q) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= B * FF_NV) return;
  const int b = i / FF_NV, v = i % FF_NV;
  double x[FF_NP];
#pragma unroll
  for (int k = 0; k < FF_NP; k++) x[k] = p[(size_t)b * ips + k];
  if (v > 0) {
    const int k = (v - 1) >> 1;
#pragma unroll
    for (int j = 0; j < FF_NP; j++)
      if (j == k) x[j] = (v & 1) ? x[j] + EPS : x[j] - EPS;
  }
#pragma unroll
  for (int k = 0; k < FF_NP; k++) q[(size_t)i * FF_NP + k] = x[k];
}

__device__ __forceinline__ bool ff_evaluated(const Point* P, const double* prm, int* st) {
  *st = ST_OK;
  for (int v = 0; v < FF_NV && *st == ST_OK; v++) *st = P[v].status;
  const bool spins = fabs(prm[2]) < 0.9999 && fabs(prm[3]) < 0.9999;
  if (!spins) *st = ST_OK;
  return spins && *st == ST_OK;
}

// one 256-thread block per point and band slice [split len, (split+1) len):
// each thread owns samples t, t + 256, ..., evaluates the 15 variants there
// and accumulates the 28 upper-triangle sums; waves reduce by shuffles, the
// block in LDS; partial[(b splits + split) 28 + n]
template <int ND>
__global__ __launch_bounds__(256) void k_fisher_full(const Point* __restrict__ pts, const double* __restrict__ q,
                                                     const double* __restrict__ params, int ips, int B,
                                                     const double* __restrict__ SN, int n2, int imin, int nb,
                                                     double Tobs, int len, double* __restrict__ partial) {
  __shared__ double red[4][FF_NPAIR];
  __shared__ double sdA[FF_NP][256], sdP[FF_NP][256];  // each thread's own column
  const int b = blockIdx.x, split = blockIdx.y, splits = gridDim.y, t = threadIdx.x;
  if (b >= B) return;
  const Point* P = pts + (size_t)b * FF_NV;
  const double* prm = params + (size_t)b * ips;
  int st;
  if (!ff_evaluated(P, prm, &st)) return;  // block-uniform; the reduce writes 1e4 I
  double scale[ND];
  scale[0] = 4.0;
#pragma unroll
  for (int l = 1; l < ND; l++) {
    const double r = prm[(l - 1) * 3 + FF_NP + 2];
    scale[l] = 4.0 * r * r;
  }
  double acc[FF_NPAIR];
#pragma unroll
  for (int n = 0; n < FF_NPAIR; n++) acc[n] = 0.0;
  const int m_end = min(nb, (split + 1) * len);
  for (int m = split * len + t; m < m_end; m += 256) {
    const int k = imin + m;
    const double f = (double)k / Tobs;
    double A0 = 0.0;
    {
      double a, ph;
      h22_at(P[0], f, &a, &ph);
      A0 = H22FAC * a;
    }
#pragma unroll 1
    for (int i = 0; i < FF_NP; i++) {
      double aP, pP, aM, pM;
      h22_at(P[2 * i + 1], f, &aP, &pP);
      h22_at(P[2 * i + 2], f, &aM, &pM);
      const double tsP = Tobs - q[((size_t)b * FF_NV + 2 * i + 1) * FF_NP + 5];
      const double tsM = Tobs - q[((size_t)b * FF_NV + 2 * i + 2) * FF_NP + 5];
      double A = H22FAC * aP - H22FAC * aM;
      double X = (TPI * f * tsP - pP) - (TPI * f * tsM - pM);
      sdA[i][t] = A / (2.0 * EPS);
      sdP[i][t] = X / (2.0 * EPS);
    }
    double wk = 0.0;
#pragma unroll
    for (int l = 0; l < ND; l++) wk += scale[l] / SN[(size_t)l * n2 + k];
    const double A2 = A0 * A0;
    int n = 0;
#pragma unroll
    for (int i = 0; i < FF_NP; i++)
#pragma unroll
      for (int j = i; j < FF_NP; j++, n++) acc[n] += wk * (sdA[i][t] * sdA[j][t] + A2 * sdP[i][t] * sdP[j][t]);
  }
  const int wave = t >> 6, lane = t & 63;
#pragma unroll
  for (int n = 0; n < FF_NPAIR; n++) {
    const double v = het_wave_sum(acc[n]);
    if (lane == 0) red[wave][n] = v;
  }
  __syncthreads();
  if (t < FF_NPAIR) partial[((size_t)b * splits + split) * FF_NPAIR + t] = ((red[0][t] + red[1][t]) + red[2][t]) + red[3][t];
}

// the matrices: 1e4 I where nothing was evaluated, else the splits' sums in
// split order, mirrored, the diagonal at least 1; fish B x 7 x 7
__global__ void k_fisher_full_reduce(const Point* __restrict__ pts, const double* __restrict__ params, int ips, int B,
                                     int splits, const double* __restrict__ partial, double* __restrict__ fish,
                                     int* __restrict__ status) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= B * FF_NPAIR) return;
  const int b = i / FF_NPAIR, n = i % FF_NPAIR;
  int r, c;
  ff_pair_of(n, &r, &c);
  double* F = fish + (size_t)b * FF_NP * FF_NP;
  int st;
  double x;
  if (ff_evaluated(pts + (size_t)b * FF_NV, params + (size_t)b * ips, &st)) {
    x = 0.0;
    for (int sp = 0; sp < splits; sp++) x += partial[((size_t)b * splits + sp) * FF_NPAIR + n];
    if (r == c && x < 1.0) x = 1.0;
  } else {
    x = (r == c) ? 1.0e4 : 0.0;
  }
  F[r * FF_NP + c] = x;
  F[c * FF_NP + r] = x;
  if (n == 0) status[b] = st;
}

// FisherEvec (QuickCBC.c:2449-2546) of d x d matrices (d <= 11: the
// reference calls it with NP = 11, NX = 7 and NS = 7, ConstCBC.h:67-69): the
// cyclic Jacobi method (Golub & Van Loan, Alg. 8.5.3, row-cyclic order) in
// place of GSL's eigen_symmv, which is absent.  A rotation is skipped, and
// a_pq set to 0, once |a_pq| <= 2^-53 sqrt(|a_pp a_qq|) (the relative
// criterion of Demmel and Veselic); sweeps stop when one rotates nothing, at
// most 60.  Then the eigenpairs ordered by |lambda| ascending
// (GSL_EIGEN_SORT_ABS_ASC, stable), each eigenvector's largest-magnitude
// component made positive (the first on a tie), and the jump sizes
// 1/sqrt(max(lambda, 10)).  A diagonal entry below 1e-16 in magnitude gives
// 1e4 and the unit vectors.
//
// EV_G lanes per matrix, EV_PER matrices per 64-lane block, matrix and
// eigenvectors in LDS.  Every lane of a group computes the rotation (t, c, s)
// from the same three entries; lane k then updates row and column k of A and
// row k of V (entries no other lane of the rotation touches), lane 0 the
// diagonal pair: per element the same operations as a sequential sweep, so
// the same results.  The rotation sequence stays serial; the block steps all
// its groups through it together and a group that has converged idles.
constexpr int EV_D = NP, EV_G = 16, EV_PER = 64 / EV_G;

__global__ __launch_bounds__(64) void k_fisher_evec(const double* __restrict__ fish, int B, int d,
                                                    double* __restrict__ ej, double* __restrict__ ev,
                                                    int* __restrict__ sweeps) {
  __shared__ double sa[EV_PER][EV_D * EV_D];
  __shared__ double sv[EV_PER][EV_D * EV_D];
  const int g = threadIdx.x / EV_G, k = threadIdx.x % EV_G;
  const int b = blockIdx.x * EV_PER + g;
  const bool live = b < B;
  double* A = sa[g];
  double* V = sv[g];
  const int dd = d * d;
  for (int i = k; i < dd; i += EV_G) {
    A[i] = live ? fish[(size_t)b * dd + i] : 1.0;
    V[i] = (i / d == i % d) ? 1.0 : 0.0;
  }
  __syncthreads();
  bool ec = false;
  for (int i = 0; i < d; i++)
    if (fabs(A[i * d + i]) < 1.0e-16) ec = true;
  bool done = !live || ec;
  int nsweep = 0;
  for (int sweep = 0; sweep < 60; sweep++) {
    if (__syncthreads_and(done)) break;  // block-uniform
    int rot = 0;
    for (int p = 0; p < d - 1; p++)
      for (int q = p + 1; q < d; q++) {
        const double apq = A[p * d + q], app = A[p * d + p], aqq = A[q * d + q];
        const bool skip = fabs(apq) <= 0x1p-53 * sqrt(fabs(app * aqq));
        double t = 0.0, c = 1.0, s = 0.0;
        if (!done && !skip) {
          const double tau = (aqq - app) / (2.0 * apq);
          if (fabs(tau) > 1.0e150)
            t = 0.5 / tau;
          else
            t = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
          c = 1.0 / sqrt(1.0 + t * t);
          s = t * c;
        }
        __syncthreads();  // every lane has read a_pq, a_pp, a_qq
        if (!done) {
          if (skip) {
            if (k == 0) {
              A[p * d + q] = 0.0;
              A[q * d + p] = 0.0;
            }
          } else {
            rot++;
            if (k < d && k != p && k != q) {
              const double akp = A[k * d + p], akq = A[k * d + q];
              const double np = c * akp - s * akq, nq = s * akp + c * akq;
              A[k * d + p] = np;
              A[p * d + k] = np;
              A[k * d + q] = nq;
              A[q * d + k] = nq;
            }
            if (k == 0) {
              A[p * d + p] = app - t * apq;
              A[q * d + q] = aqq + t * apq;
              A[p * d + q] = 0.0;
              A[q * d + p] = 0.0;
            }
            if (k < d) {
              const double vkp = V[k * d + p], vkq = V[k * d + q];
              V[k * d + p] = c * vkp - s * vkq;
              V[k * d + q] = s * vkp + c * vkq;
            }
          }
        }
        __syncthreads();
      }
    if (!done) {
      nsweep = sweep + 1;
      if (rot == 0) done = true;
    }
  }
  if (!live || k != 0) return;
  double* e = ej + (size_t)b * d;
  double* E = ev + (size_t)b * dd;
  if (sweeps) sweeps[b] = ec ? 0 : nsweep;
  if (ec) {
    for (int i = 0; i < d; i++) {
      e[i] = 10000.0;
      for (int j = 0; j < d; j++) E[i * d + j] = (i == j) ? 1.0 : 0.0;
    }
    return;
  }
  int idx[EV_D];
  for (int i = 0; i < d; i++) idx[i] = i;
  for (int i = 1; i < d; i++) {
    const int t = idx[i];
    int j = i - 1;
    while (j >= 0 && fabs(A[idx[j] * d + idx[j]]) > fabs(A[t * d + t])) {
      idx[j + 1] = idx[j];
      j--;
    }
    idx[j + 1] = t;
  }
  for (int i = 0; i < d; i++) {
    const int col = idx[i];
    int kmax = 0;
    double vmax = -1.0;
    for (int r = 0; r < d; r++)
      if (fabs(V[r * d + col]) > vmax) {
        vmax = fabs(V[r * d + col]);
        kmax = r;
      }
    const double sg = V[kmax * d + col] < 0.0 ? -1.0 : 1.0;
    double l = A[col * d + col];
    if (l < 10.0) l = 10.0;
    e[i] = 1.0 / sqrt(l);
    for (int r = 0; r < d; r++) E[i * d + r] = sg * V[r * d + col];
  }
}

}  // namespace

hipError_t launch_fisher_params(const double* d_params, int B, double* d_q, hipStream_t s) {
  if (B <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fisher_params, dim3((B * NV + 255) / 256), dim3(256), 0, s, d_params, B, d_q);
  return hipGetLastError();
}

hipError_t launch_het_fisher(const Point* pts, const double* d_params, int B, int nd, const double* freq,
                             const double* pl, const double* ha, const double* w, int M, double* fish, int* status,
                             hipStream_t s, int splits, double* partial) {
  if (B <= 0) return hipSuccess;
  if (splits < 1 || (splits > 1 && !partial)) return hipErrorInvalidValue;
  int len = (M + splits - 1) / splits;
  len = (len + 63) / 64 * 64;
  const dim3 grid(B, splits);
  if (nd == 1)
    hipLaunchKernelGGL(k_het_fisher<1>, grid, dim3(256), 0, s, pts, d_params, B, freq, pl, ha, w, M, fish, status,
                       len, partial);
  else if (nd == 2)
    hipLaunchKernelGGL(k_het_fisher<2>, grid, dim3(256), 0, s, pts, d_params, B, freq, pl, ha, w, M, fish, status,
                       len, partial);
  else
    hipLaunchKernelGGL(k_het_fisher<3>, grid, dim3(256), 0, s, pts, d_params, B, freq, pl, ha, w, M, fish, status,
                       len, partial);
  if (splits > 1)
    hipLaunchKernelGGL(k_fisher_reduce, dim3((B * NPAIR + 255) / 256), dim3(256), 0, s, pts, d_params, B, splits,
                       partial, fish);
  return hipGetLastError();
}

hipError_t launch_fisher_full_params(const double* d_params, int ips, int B, double* d_q, hipStream_t s) {
  if (B <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fisher_full_params, dim3((B * FF_NV + 255) / 256), dim3(256), 0, s, d_params, ips, B, d_q);
  return hipGetLastError();
}

hipError_t launch_fisher_full(const Point* pts, const double* d_q, const double* d_params, int ips, int B, int nd,
                              const double* SN, int n2, int imin, int nb, double Tobs, int splits, double* partial,
                              double* fish, int* status, hipStream_t s) {
  if (B <= 0) return hipSuccess;
  if (splits < 1 || nb < 1 || !partial || nd < 1 || nd > 3) return hipErrorInvalidValue;
  const int len = (nb + splits - 1) / splits;
  const dim3 grid(B, splits);
  if (nd == 1)
    hipLaunchKernelGGL(k_fisher_full<1>, grid, dim3(256), 0, s, pts, d_q, d_params, ips, B, SN, n2, imin, nb, Tobs,
                       len, partial);
  else if (nd == 2)
    hipLaunchKernelGGL(k_fisher_full<2>, grid, dim3(256), 0, s, pts, d_q, d_params, ips, B, SN, n2, imin, nb, Tobs,
                       len, partial);
  else
    hipLaunchKernelGGL(k_fisher_full<3>, grid, dim3(256), 0, s, pts, d_q, d_params, ips, B, SN, n2, imin, nb, Tobs,
                       len, partial);
  hipLaunchKernelGGL(k_fisher_full_reduce, dim3((B * FF_NPAIR + 255) / 256), dim3(256), 0, s, pts, d_params, ips, B,
                     splits, partial, fish, status);
  return hipGetLastError();
}

hipError_t launch_fisher_evec(const double* fish, int B, int d, double* ej, double* ev, int* sweeps, hipStream_t s) {
  if (B <= 0) return hipSuccess;
  if (d < 1 || d > EV_D) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_fisher_evec, dim3((B + EV_PER - 1) / EV_PER), dim3(64), 0, s, fish, B, d, ej, ev, sweeps);
  return hipGetLastError();
}

}  // namespace qcbc
