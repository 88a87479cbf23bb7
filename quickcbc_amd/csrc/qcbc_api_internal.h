// Host-side state of libqcbc shared by its C++ files.
#pragma once

#include <hip/hip_runtime_api.h>

#include <mutex>
#include <utility>
#include <vector>

#include "../../include/qcbc.h"
#include "qcbc_types.h"

#define QCBC_MAX_SPLITS 64
#define QCBC_PLAIN_BLOCKS 64
#define QCBC_PT_MAX_GROUPS 4

// One resident analysis: band planes (FP64 and scaled FP32), the raw
// half-complex data and PSD, per-batch scratch, launch shape and timing.
struct qcbc_ctx {
  qcbc_desc desc{};
  qcbc::AnalysisConst an{};
  int nhalf = 0, nb = 0, nbp = 0;
  int device = 0, n_cu = 256;
  hipStream_t stream = nullptr;
  double* d_p64 = nullptr;
  float* d_p32 = nullptr;
  double* d_D = nullptr;
  double* d_SN = nullptr;
  double* d_wref = nullptr;
  double wref[3] = {1, 1, 1};
  bool bound = false;
  // batch scratch
  int cap_B = 0;
  double* d_params = nullptr;
  qcbc::Point* d_pts = nullptr;
  int* d_status = nullptr;
  double* d_logL = nullptr;
  double* d_rho = nullptr;
  double* d_partial = nullptr;
  // plain-path scratch (one point)
  double* d_amp = nullptr;
  double* d_phase = nullptr;
  double* d_hw = nullptr;
  double* d_plain_partial = nullptr;
  // data-conditioning scratch, kept for the next call of the same shape:
  // time series (n_det x N), its transform (n_det x (N/2 + 1) complex) and
  // the real-to-complex plan; makespec's knots, lines and output
  double* d_cond_x = nullptr;
  void* d_cond_X = nullptr;
  void* cond_plan = nullptr;  // hipfftHandle
  void* cond_iplan = nullptr;  // complex-to-real plan of the whitened series
  double* d_spec = nullptr;
  size_t spec_cap = 0;  // doubles
  // fused-kernel launch shape per precision
  int cpb[2] = {8, 8}, tile[2] = {512, 512}, splits[2] = {0, 0};
  // kernel timing
  bool timing = false;
  double kernel_ms = 0.0;
  long long launches = 0;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_pairs;
  std::mutex mu;
};

namespace qcbc {
extern thread_local hipError_t t_last_hip;
int ensure_device();
// the no-allocation entry point's device array: its address on device dev
// (called once per device, with dev current), and its size in bytes
int noalloc_init_device(int dev);
size_t noalloc_scratch_bytes();
// the analysis constants, half length, band bins and padded band bins of a
// description (QCBC_EINVAL for an invalid one)
int make_analysis(const qcbc_desc* d, AnalysisConst* an, int* nhalf, int* nb, int* nbp);
int effective_splits(const qcbc_ctx* c, int prec, int B);
LLInst make_inst(const qcbc_ctx* c, int tile, int B, int splits, const Point* pts, double* partial);
// the fused likelihood launch on stream s, bracketed by timing events when
// the context's kernel timing is on (caller holds mu)
int ll_launch_timed(qcbc_ctx* c, int prec, const LLInst& in, hipStream_t s);
// setup + fused likelihood partial sums on the context stream (caller holds
// mu); *out describes the launch (splits, partial buffer) for the finalize
int loglike_partials(qcbc_ctx* c, const double* d_params, int B, int prec, int* d_status, LLInst* out);
// setup + fused likelihood + finalize on the context stream (caller holds mu)
int loglike_core(qcbc_ctx* c, const double* d_params, int B, int prec, double* d_logL, double* d_rho,
                 int* d_status);
// the heterodyned likelihood of B points on stream s: their set-up into pts,
// status into d_status, logL into d_logL (the heterodyne's data are read only)
int het_enqueue(qcbc_het* h, const double* d_params, int B, Point* pts, int* d_status, double* d_logL, hipStream_t s);
int het_device(const qcbc_het* h);
// Fisher_Het and FisherEvec of B points (device params) into d_ej (B x 11)
// and d_ev (B x 11 x 11) on stream s, with the heterodyne's scratch; returns
// after the stream has finished them
int het_fisher_enqueue(qcbc_het* h, const double* d_params, int B, double* d_ej, double* d_ev, hipStream_t s);

// A heterodyne's grid as the grid-choice kernels read it: frequencies (M),
// sample powers (4 x M), folded weights (3 x nd x M, first wH / a_ref^2) and
// the reference amplitude (nd x M), all device memory
struct HetGridView {
  int device, M, J, NR, nd;
  AnalysisConst an;
  const double *freq, *pl, *wf, *amp, *phase;
  double logLb;
};
int het_grid_view(const qcbc_het* h, HetGridView* out);

// The heterodyne's grid choice (qcbc_gridchoice.cpp), on stream s; PSD rows
// SN on the host, d_SN (n_det x N/2) on the device.
// limits (QuickCBC.c:4275-4385): stst[2], ratio (N/2, may be null)
int limits_impl(qcbc_ctx* c, const double* const* SN, const double* params, int* stst, double* ratio, hipStream_t s);
// Fisher_All (QuickCBC.c:2695-2817) of B points (host memory)
int fisher_full_impl(qcbc_ctx* c, const double* d_SN, const double* params, int B, int d, double* fish, int* status,
                     hipStream_t s);
int fisher_all_impl(qcbc_ctx* c, const double* d_SN, const double* params, int B, double* fish, int* status,
                    hipStream_t s);
// diff (QuickCBC.c:4913-4964), or het_diff (:4966-5053) when h is given, of
// K displaced points px (K x 11) against start points params[of[k]]
int distances_impl(qcbc_ctx* c, const double* d_SN, qcbc_het* h, const double* params, int B, const double* px,
                   const int* of, int K, double* z, hipStream_t s);
// efix (QuickCBC.c:4387-4567) of B points: ej (B x 11) scaled in place
int efix_impl(qcbc_ctx* c, const double* d_SN, qcbc_het* h, const double* params, int B, double* ej, const double* ev,
              const double* mn, const double* mx, double etamin, double zs, int* trials, hipStream_t s);
// the same on device states, jump sizes and directions (the PT step's)
int het_efix_device(qcbc_het* h, const double* d_x, int B, double* d_ej, const double* d_ev, const double* mn,
                    const double* mx, double etamin, double zs, hipStream_t s);
// het_space (QuickCBC.c:4570-4910): fgrid has room for cap entries
int het_space_impl(qcbc_ctx* c, const double* const* SN, const double* d_SN, const double* params, const double* mn,
                   const double* mx, double etamin, int* MN, int* MM, int* M, int* NR, int* fgrid, int cap,
                   double* eval, double* evec, double* dparams, hipStream_t s);
// The search likelihoods (qcbc_search.cpp), params in place (7 + 3 (n_det - 1) each)
int search_apmax(qcbc_ctx* c, double* params, int B, double* logL, int* status);
int search_llmax(qcbc_ctx* c, double* params, int B, double tmin, double tmax, double net_tmax, const double* delays,
                 double* logL, int* status);
int search_penalized_one(const double* params, const double* D, const double* twave, const double* SN, int N,
                         double Tobs, double fmin, double fmax, double* logL, double* A, double* P, hipStream_t s);
int search_pdwave(const double* params, double tcmin, int N, double Tobs, double fref, double fmin0, double* wavef,
                  hipStream_t s);
int search_pbt_shift(double* corr, double* corrf, const double* data1, const double* data2, const double* Sn, int imin,
                     int imax, int N, hipStream_t s);

// the small conditioning drop-ins on the GPU (qcbc_ingest.cpp): tukey in
// place on N samples, tukey_scale into s12 = {s1, s2}, whiten in place on a
// half-complex row of N with Sn (N/2); synchronous on s
int ingest_tukey(double* data, double alpha, int N, hipStream_t s);
int ingest_tukey_scale(double alpha, int N, double* s12, hipStream_t s);
int ingest_whiten(double* data, const double* Sn, int N, hipStream_t s);
}  // namespace qcbc
