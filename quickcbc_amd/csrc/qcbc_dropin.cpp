// QuickCBC's own entry points, re-implemented on the batch API so that a
// QuickCBC build can link this library in place of the reference objects.
//
// QuickCBC calls these from up to 12 OpenMP threads at once, one chain per
// thread (QuickCBC.c:880, :1422, :2063).  Each calling thread therefore gets
// its own stream, scratch buffers and cached analyses (thread_local), and the
// analyses are keyed by the caller's D/SN row pointers so that the data are
// copied to the GPU once, not on every call.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/qcbc.h"
#include "qcbc_api_internal.h"
#include "qcbc_fisher_internal.h"
#include "qcbc_het_internal.h"
#include "qcbc_internal.h"

using namespace qcbc;

namespace {

// ConstCBC.h:31-34 (the reference's compile-time values)
std::atomic<double> g_fmin{16.0}, g_fmax{1024.0}, g_fref{100.0};
std::atomic<int> g_exit_on_error{1};
std::atomic<unsigned long long> g_generation{1};

struct CtxEntry {
  qcbc_desc desc{};
  qcbc_ctx* ctx = nullptr;
  std::vector<const double*> drows, srows;  // bound data (empty: none)
  unsigned long long gen = 0;
  unsigned long long last_use = 0;
};

struct DropState {
  hipStream_t stream = nullptr;
  double* d_buf = nullptr;
  size_t cap = 0;  // doubles
  Point* d_pt = nullptr;
  int* d_st = nullptr;
  std::vector<CtxEntry> entries;
  unsigned long long tick = 0;
  int init_status = -1;
};

thread_local DropState t_drop;

int drop_init() {
  DropState& s = t_drop;
  if (s.init_status == QCBC_OK) return QCBC_OK;
  int st = ensure_device();
  if (st) return st;
  if (hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking) != hipSuccess ||
      hipMalloc(&s.d_pt, sizeof(Point)) != hipSuccess || hipMalloc(&s.d_st, sizeof(int)) != hipSuccess)
    return QCBC_ENOMEM;
  s.init_status = QCBC_OK;
  return QCBC_OK;
}

int drop_buf(size_t n) {
  DropState& s = t_drop;
  if (n <= s.cap) return QCBC_OK;
  if (s.d_buf) hipFree(s.d_buf);
  s.d_buf = nullptr;
  s.cap = 0;
  if (hipMalloc(&s.d_buf, sizeof(double) * n) != hipSuccess) return QCBC_ENOMEM;
  s.cap = n;
  return QCBC_OK;
}

[[noreturn]] void fail(const char* where, int code) {
  std::fprintf(stderr, "libqcbc %s: %s\n", where, qcbc_strerror(code));
  std::exit(-1);
}

// Domain errors follow the reference's ERROR() (print, exit(-1)) unless
// qcbc_set_exit_on_error(0); device or runtime errors always exit.
bool domain_error(const char* where, int status) {
  if (status == QCBC_OK) return false;
  if (status != QCBC_EDOM || g_exit_on_error.load()) fail(where, status);
  return true;
}

bool same_desc(const qcbc_desc& a, const qcbc_desc& b) {
  if (a.n_det != b.n_det || a.N != b.N || a.Tobs != b.Tobs || a.gmst != b.gmst || a.fmin != b.fmin ||
      a.fmax != b.fmax || a.fref != b.fref)
    return false;
  for (int i = 0; i < a.n_det; i++)
    if (a.labels[i] != b.labels[i]) return false;
  return true;
}

// Find or create the thread's analysis for this shape; bind D/SN if given and
// not already bound (same row pointers, same generation), or always when
// rebind is set.
qcbc_ctx* get_ctx(const qcbc_desc& desc, double** D, double** SN, int* status, bool rebind = false) {
  DropState& s = t_drop;
  *status = drop_init();
  if (*status) return nullptr;
  CtxEntry* hit = nullptr;
  for (auto& e : s.entries)
    if (same_desc(e.desc, desc)) {
      hit = &e;
      break;
    }
  if (!hit) {
    if (s.entries.size() >= 8) {  // evict the least recently used analysis
      size_t lru = 0;
      for (size_t i = 1; i < s.entries.size(); i++)
        if (s.entries[i].last_use < s.entries[lru].last_use) lru = i;
      qcbc_ctx_destroy(s.entries[lru].ctx);
      s.entries.erase(s.entries.begin() + lru);
    }
    CtxEntry e;
    e.desc = desc;
    *status = qcbc_ctx_create(&desc, &e.ctx);
    if (*status) return nullptr;
    s.entries.push_back(e);
    hit = &s.entries.back();
  }
  hit->last_use = ++s.tick;
  if (D && SN) {
    bool same = !rebind && hit->gen == g_generation.load() && (int)hit->drows.size() == desc.n_det;
    for (int d = 0; same && d < desc.n_det; d++) same = hit->drows[d] == D[d] && hit->srows[d] == SN[d];
    if (!same) {
      *status = qcbc_bind_rows(hit->ctx, D, SN);
      if (*status) return nullptr;
      hit->drows.assign(D, D + desc.n_det);
      hit->srows.assign(SN, SN + desc.n_det);
      hit->gen = g_generation.load();
    }
  }
  return hit->ctx;
}

// The kernels evaluate the analysis grid f_i = i / Tobs (freq[0] = 1/Tobs),
// the grid QuickCBC builds (QuickCBC.c:417-419, :635-636).  A freq that is
// not that grid, to a few ulps, is rejected rather than ignored.  A list
// once checked is remembered by its pointer, length and Tobs until
// qcbc_invalidate().  NULL stands for the grid itself.
struct GridSeen {
  const double* data = nullptr;
  size_t length = 0;
  int nhalf = 0;
  double Tobs = 0.0;
  unsigned long long gen = 0;
};
thread_local GridSeen t_grid_seen[4];
thread_local unsigned t_grid_next = 0;

bool grid_ok(const RealVector* freq, int nhalf, double Tobs) {
  if (!freq) return true;
  if (!freq->data || freq->length < (size_t)nhalf || !(Tobs > 0)) return false;
  const unsigned long long gen = g_generation.load();
  for (const auto& g : t_grid_seen)
    if (g.data == freq->data && g.length == freq->length && g.nhalf == nhalf && g.Tobs == Tobs && g.gen == gen)
      return true;
  const double* f = freq->data;
  const double tol = 1e-15;  // about 4.5 ulp
  if (nhalf > 0 && !(std::fabs(f[0] - 1.0 / Tobs) <= tol / Tobs)) return false;
  for (int i = 1; i < nhalf; i++) {
    const double want = (double)i / Tobs;
    if (!(std::fabs(f[i] - want) <= tol * want)) return false;
  }
  GridSeen& g = t_grid_seen[t_grid_next++ & 3];
  g.data = freq->data;
  g.length = freq->length;
  g.nhalf = nhalf;
  g.Tobs = Tobs;
  g.gen = gen;
  return true;
}

qcbc_desc desc_from(const struct Net* net, int N, double Tobs) {
  qcbc_desc d{};
  d.n_det = net->Nifo;
  for (int i = 0; i < net->Nifo && i < 3; i++) d.labels[i] = net->labels[i];
  d.gmst = net->GMST;
  d.N = N;
  d.Tobs = Tobs;
  d.fmin = g_fmin.load();
  d.fmax = g_fmax.load();
  d.fref = g_fref.load();
  return d;
}

}  // namespace

extern "C" {

void qcbc_set_band(double fmin, double fmax) {
  g_fmin = fmin;
  g_fmax = fmax;
}
void qcbc_set_fref(double fref) { g_fref = fref; }
void qcbc_set_exit_on_error(int on) { g_exit_on_error = on ? 1 : 0; }
void qcbc_invalidate(void) { g_generation++; }

double log_likelihood_full(struct Net* net, double** D, double* params, RealVector* freq, double** SN, double* rho,
                           int N, double Tobs) {
  if (!net || net->Nifo < 1 || net->Nifo > 3 || !D || !SN || !params || !grid_ok(freq, N / 2, Tobs))
    fail("log_likelihood_full", QCBC_EINVAL);
  qcbc_desc desc = desc_from(net, N, Tobs);
  int st;
  qcbc_ctx* ctx = get_ctx(desc, D, SN, &st);
  if (!ctx) fail("log_likelihood_full", st);
  double logL = 0.0, r[3] = {0, 0, 0};
  int pst = 0;
  st = qcbc_loglike(ctx, params, 1, QCBC_FP64, &logL, r, &pst);
  if (st) fail("log_likelihood_full", st);
  if (domain_error("log_likelihood_full", pst)) return 0.0;
  for (int i = 0; i < net->Nifo; i++)
    if (rho) rho[i] = r[i];
  return logL;
}

void fulltemplates(struct Net* net, double** hwave, RealVector* freq, double* params, int N) {
  if (!net || net->Nifo < 1 || net->Nifo > 3 || !hwave || !freq || !freq->data || freq->length < 2 || !params ||
      !grid_ok(freq, N / 2, 1.0 / freq->data[1]))
    fail("fulltemplates", QCBC_EINVAL);
  qcbc_desc desc = desc_from(net, N, 1.0 / freq->data[1]);  // QuickCBC.c:3702
  int st;
  qcbc_ctx* ctx = get_ctx(desc, nullptr, nullptr, &st);
  if (!ctx) fail("fulltemplates", st);
  std::vector<double> tmp((size_t)net->Nifo * N);
  int pst = 0;
  st = qcbc_templates(ctx, params, 1, tmp.data(), &pst);
  if (st) fail("fulltemplates", st);
  if (domain_error("fulltemplates", pst)) std::fill(tmp.begin(), tmp.end(), 0.0);
  for (int d = 0; d < net->Nifo; d++) std::memcpy(hwave[d], tmp.data() + (size_t)d * N, sizeof(double) * N);
}

void templates(struct Net* net, double** hwave, RealVector* freq, double* params, int N) {
  if (!net || net->Nifo < 1 || net->Nifo > 3 || !hwave || !freq || !freq->data || freq->length < 2 || !params ||
      !grid_ok(freq, N / 2, 1.0 / freq->data[1]))
    fail("templates", QCBC_EINVAL);
  qcbc_desc desc = desc_from(net, N, 1.0 / freq->data[1]);  // QuickCBC.c:3539
  int st;
  qcbc_ctx* ctx = get_ctx(desc, nullptr, nullptr, &st);
  if (!ctx) fail("templates", st);
  std::vector<double> tmp((size_t)net->Nifo * N);
  int pst = 0;
  st = qcbc_templates_intrinsic(ctx, params, 1, tmp.data(), &pst);
  if (st) fail("templates", st);
  if (domain_error("templates", pst)) std::fill(tmp.begin(), tmp.end(), 0.0);
  for (int d = 0; d < net->Nifo; d++) std::memcpy(hwave[d], tmp.data() + (size_t)d * N, sizeof(double) * N);
}

double log_likelihood_intrinsic(struct Net* net, double** D, double* params, RealVector* freq, double** SN, int N,
                                double Tobs) {
  if (!net || net->Nifo < 1 || net->Nifo > 3 || !D || !SN || !params || !grid_ok(freq, N / 2, Tobs))
    fail("log_likelihood_intrinsic", QCBC_EINVAL);
  qcbc_desc desc = desc_from(net, N, Tobs);
  int st;
  qcbc_ctx* ctx = get_ctx(desc, D, SN, &st);
  if (!ctx) fail("log_likelihood_intrinsic", st);
  double logL = 0.0;
  int pst = 0;
  st = qcbc_loglike_intrinsic(ctx, params, 1, &logL, nullptr, &pst);
  if (st) fail("log_likelihood_intrinsic", st);
  if (domain_error("log_likelihood_intrinsic", pst)) return 0.0;
  return logL;
}

int IMRPhenomDGenerateh22FDAmpPhase(AmpPhaseFDWaveform** h22, RealVector* freq, double phi0, double fRef_in,
                                    double m1_SI, double m2_SI, double chi1, double chi2, double distance) {
  // IMRPhenomD.c:197-210: pointer checks
  if (!freq || !freq->data || freq->length == 0 || !h22) {
    if (g_exit_on_error.load()) fail("IMRPhenomDGenerateh22FDAmpPhase", QCBC_EFAULT);
    return QCBC_EFAULT;
  }
  if (*h22) {
    if (g_exit_on_error.load()) fail("IMRPhenomDGenerateh22FDAmpPhase", QCBC_EFAULT);
    return QCBC_EFAULT;
  }
  int st = drop_init();
  if (st) fail("IMRPhenomDGenerateh22FDAmpPhase", st);
  const size_t n = freq->length;
  DropState& s = t_drop;
  if ((st = drop_buf(3 * n))) fail("IMRPhenomDGenerateh22FDAmpPhase", st);
  double* d_f = s.d_buf;
  double* d_a = s.d_buf + n;
  double* d_p = s.d_buf + 2 * n;
  int pst = 0;
  // fRef = 0 means f_min (IMRPhenomD.c:227); the reference then passes
  // fRef_in, not fRef, to the internal routine (:242, :511).
  if (hipMemcpyAsync(d_f, freq->data, sizeof(double) * n, hipMemcpyHostToDevice, s.stream) != hipSuccess ||
      launch_setup_h22(m1_SI, m2_SI, chi1, chi2, distance, phi0, fRef_in, freq->data[0], s.d_pt, s.d_st, s.stream) !=
          hipSuccess ||
      hipMemcpyAsync(&pst, s.d_st, sizeof(int), hipMemcpyDeviceToHost, s.stream) != hipSuccess ||
      hipStreamSynchronize(s.stream) != hipSuccess)
    fail("IMRPhenomDGenerateh22FDAmpPhase", QCBC_EHIP);
  if (pst != QCBC_OK) {
    if (g_exit_on_error.load()) fail("IMRPhenomDGenerateh22FDAmpPhase", pst);
    return pst;
  }
  AmpPhaseFDWaveform* wf = CreateAmpPhaseFDWaveform(n);
  if (launch_waveform_freqs(s.d_pt, d_f, (int)n, d_a, d_p, s.stream) != hipSuccess ||
      hipMemcpyAsync(wf->amp, d_a, sizeof(double) * n, hipMemcpyDeviceToHost, s.stream) != hipSuccess ||
      hipMemcpyAsync(wf->phase, d_p, sizeof(double) * n, hipMemcpyDeviceToHost, s.stream) != hipSuccess ||
      hipStreamSynchronize(s.stream) != hipSuccess)
    fail("IMRPhenomDGenerateh22FDAmpPhase", QCBC_EHIP);
  std::memcpy(wf->freq, freq->data, sizeof(double) * n);
  *h22 = wf;
  return 0;
}

double fourier_nwip(double* a, double* b, double* Sn, int imin, int imax, int N) {
  if (imax <= imin) return 0.0;
  if (!a || !b || !Sn || imin < 0 || N < 2 * imax - 1) fail("fourier_nwip", QCBC_EINVAL);
  int st = drop_init();
  if (st) fail("fourier_nwip", st);
  DropState& s = t_drop;
  const int nblk = 64;
  if ((st = drop_buf(2 * (size_t)N + imax + nblk + 1))) fail("fourier_nwip", st);
  double* d_a = s.d_buf;
  double* d_b = d_a + N;
  double* d_s = d_b + N;
  double* d_part = d_s + imax;
  double* d_out = d_part + nblk;
  double out = 0.0;
  // bins start at 1: the reference's i = 0 term would read a[N], past the N samples
  if (hipMemcpyAsync(d_a, a, sizeof(double) * N, hipMemcpyHostToDevice, s.stream) != hipSuccess ||
      hipMemcpyAsync(d_b, b, sizeof(double) * N, hipMemcpyHostToDevice, s.stream) != hipSuccess ||
      hipMemcpyAsync(d_s, Sn, sizeof(double) * imax, hipMemcpyHostToDevice, s.stream) != hipSuccess ||
      launch_nwip_pair(d_a, d_b, d_s, imin < 1 ? 1 : imin, imax, N, nblk, d_part, d_out, s.stream) != hipSuccess ||
      hipMemcpyAsync(&out, d_out, sizeof(double), hipMemcpyDeviceToHost, s.stream) != hipSuccess ||
      hipStreamSynchronize(s.stream) != hipSuccess)
    fail("fourier_nwip", QCBC_EHIP);
  return out;
}

static void geometry(int what, const double* in, size_t nin, double* out, size_t nout, const char* where) {
  int st = drop_init();
  if (st) fail(where, st);
  DropState& s = t_drop;
  if ((st = drop_buf(nin + nout))) fail(where, st);
  if (hipMemcpyAsync(s.d_buf, in, sizeof(double) * nin, hipMemcpyHostToDevice, s.stream) != hipSuccess ||
      launch_geometry(what, s.d_buf, 1, s.d_buf + nin, s.stream) != hipSuccess ||
      hipMemcpyAsync(out, s.d_buf + nin, sizeof(double) * nout, hipMemcpyDeviceToHost, s.stream) != hipSuccess ||
      hipStreamSynchronize(s.stream) != hipSuccess)
    fail(where, QCBC_EHIP);
}

void F_ant(double psi, double alpha, double sindelta, double GMST, double* Fplus, double* Fcross, int id) {
  if (id < 0 || id > 2) return;  // Utilities.c:995-1000 leaves the outputs untouched
  double in[5] = {psi, alpha, sindelta, GMST, (double)id}, out[2];
  geometry(0, in, 5, out, 2, "F_ant");
  *Fplus = out[0];
  *Fcross = out[1];
}

void Times(double alpha, double sindelta, double GMST, double* dtimes) {
  double in[3] = {alpha, sindelta, GMST};
  geometry(1, in, 3, dtimes, 3, "Times");
}

double fbegin(double* param) {
  double out = 0.0;
  geometry(2, param, 11, &out, 1, "fbegin");
  return out;
}

void fullphaseamp(struct Net* net, double** amp, double** phase, RealVector* freq, double* params, int N) {
  const int n = N / 2;
  if (!net || net->Nifo < 1 || net->Nifo > 3 || !amp || !phase || !freq || !freq->data || !params || n < 1 ||
      freq->length < (size_t)n)
    fail("fullphaseamp", QCBC_EINVAL);
  for (int d = 0; d < net->Nifo; d++)
    if (!amp[d] || !phase[d]) fail("fullphaseamp", QCBC_EINVAL);
  // the analysis's sky and detector constants; Tobs is net->Tobs (QuickCBC.c:3799)
  qcbc_desc desc = desc_from(net, 2 * n < 4 ? 4 : 2 * n, net->Tobs);
  int st;
  qcbc_ctx* ctx = get_ctx(desc, nullptr, nullptr, &st);
  if (!ctx) fail("fullphaseamp", st);
  DropState& s = t_drop;
  const int nd = net->Nifo;
  if ((st = drop_buf(11 + (size_t)n * (1 + 2 * nd)))) fail("fullphaseamp", st);
  double* d_p = s.d_buf;
  double* d_f = d_p + 11;
  double* d_a = d_f + n;
  double* d_ph = d_a + (size_t)nd * n;
  // IMRPhenomDGenerateh22FDAmpPhase's f_min is the list's first frequency
  AnalysisConst an = ctx->an;
  an.fmin0 = freq->data[0];
  std::vector<double> a((size_t)nd * n), p((size_t)nd * n);
  int pst = 0;
  if (hipMemcpyAsync(d_p, params, sizeof(double) * 11, hipMemcpyHostToDevice, s.stream) != hipSuccess ||
      hipMemcpyAsync(d_f, freq->data, sizeof(double) * n, hipMemcpyHostToDevice, s.stream) != hipSuccess ||
      launch_setup(d_p, 1, an, s.d_pt, s.d_st, s.stream) != hipSuccess ||
      launch_phaseamp(s.d_pt, d_f, n, nd, net->Tobs, std::sqrt(net->Tobs), d_a, d_ph, s.stream) != hipSuccess ||
      hipMemcpyAsync(&pst, s.d_st, sizeof(int), hipMemcpyDeviceToHost, s.stream) != hipSuccess ||
      hipMemcpyAsync(a.data(), d_a, sizeof(double) * a.size(), hipMemcpyDeviceToHost, s.stream) != hipSuccess ||
      hipMemcpyAsync(p.data(), d_ph, sizeof(double) * p.size(), hipMemcpyDeviceToHost, s.stream) != hipSuccess ||
      hipStreamSynchronize(s.stream) != hipSuccess)
    fail("fullphaseamp", QCBC_EHIP);
  if (domain_error("fullphaseamp", pst)) {
    std::fill(a.begin(), a.end(), 0.0);
    std::fill(p.begin(), p.end(), 0.0);
  }
  for (int d = 0; d < nd; d++) {
    std::memcpy(amp[d], a.data() + (size_t)d * n, sizeof(double) * n);
    std::memcpy(phase[d], p.data() + (size_t)d * n, sizeof(double) * n);
  }
}

}  // extern "C"

namespace {

// The heterodynes a thread has on the GPU, keyed by the struct, its array
// pointers and the reference point it was made at.
struct HetEntry {
  const struct Het* het = nullptr;
  const void* arrays[4] = {};
  int M = 0, J = 0, NR = 0;
  double pref[11] = {};
  bool has_pref = false;
  unsigned long long gen = 0, last_use = 0;
  qcbc_het* h = nullptr;
};
thread_local std::vector<HetEntry> t_hets;

HetEntry key_of(const struct Het* het) {
  HetEntry e;
  e.het = het;
  e.arrays[0] = het->IP;
  e.arrays[1] = het->SL;
  e.arrays[2] = het->amp;
  e.arrays[3] = het->freq ? (const void*)het->freq->data : nullptr;
  e.M = het->M;
  e.J = het->J;
  e.NR = het->NR;
  e.has_pref = het->pref != nullptr;
  if (e.has_pref) std::memcpy(e.pref, het->pref, sizeof(e.pref));
  e.gen = g_generation.load();
  return e;
}

bool same_key(const HetEntry& a, const HetEntry& b) {
  if (a.het != b.het || a.M != b.M || a.J != b.J || a.NR != b.NR || a.has_pref != b.has_pref || a.gen != b.gen)
    return false;
  for (int i = 0; i < 4; i++)
    if (a.arrays[i] != b.arrays[i]) return false;
  return !a.has_pref || std::memcmp(a.pref, b.pref, sizeof(a.pref)) == 0;
}

void het_cache_put(HetEntry e) {
  for (auto it = t_hets.begin(); it != t_hets.end();)
    if (it->het == e.het) {  // a newer heterodyne of the same struct
      qcbc_het_destroy(it->h);
      it = t_hets.erase(it);
    } else {
      ++it;
    }
  if (t_hets.size() >= 4) {
    size_t lru = 0;
    for (size_t i = 1; i < t_hets.size(); i++)
      if (t_hets[i].last_use < t_hets[lru].last_use) lru = i;
    qcbc_het_destroy(t_hets[lru].h);
    t_hets.erase(t_hets.begin() + lru);
  }
  e.last_use = ++t_drop.tick;
  t_hets.push_back(e);
}

double*** tensor(int n, int m, int l) {  // double_tensor's shape (Utilities.c:3001-3016)
  double*** t = (double***)std::malloc(sizeof(double**) * (n + 1));
  if (!t) fail("heterodyne", QCBC_ENOMEM);
  for (int i = 0; i < n + 1; i++) {
    t[i] = (double**)std::malloc(sizeof(double*) * (m + 1));
    if (!t[i]) fail("heterodyne", QCBC_ENOMEM);
    for (int j = 0; j < m + 1; j++) {
      t[i][j] = (double*)std::malloc(sizeof(double) * (l + 1));
      if (!t[i][j]) fail("heterodyne", QCBC_ENOMEM);
    }
  }
  return t;
}

double** matrix(int n, int m) {  // double_matrix's shape (Utilities.c:2981-2992)
  double** a = (double**)std::malloc(sizeof(double*) * (n + 1));
  if (!a) fail("heterodyne", QCBC_ENOMEM);
  for (int i = 0; i < n + 1; i++) {
    a[i] = (double*)std::malloc(sizeof(double) * (m + 1));
    if (!a[i]) fail("heterodyne", QCBC_ENOMEM);
  }
  return a;
}

}  // namespace

extern "C" {

void heterodyne(struct Net* net, struct Het* het, double** D, double* params, RealVector* freq, double** SN, int N,
                double Tobs) {
  if (!net || net->Nifo < 1 || net->Nifo > 3 || !het || !D || !SN || !params || !het->fgrid ||
      !grid_ok(freq, N / 2, Tobs))
    fail("heterodyne", QCBC_EINVAL);
  const int M = het->M, J = het->J, NR = het->NR, nd = net->Nifo;
  if (J < 1 || M != NR * J + 1) fail("heterodyne", QCBC_EINVAL);
  qcbc_desc desc = desc_from(net, N, Tobs);
  int st;
  // heterodyne runs rarely and often right after SN was redrawn in place
  // (noise marginalisation, QuickCBC.c:2001-2006): always bind D/SN afresh
  qcbc_ctx* ctx = get_ctx(desc, D, SN, &st, true);
  if (!ctx) fail("heterodyne", st);
  qcbc_het* h = nullptr;
  st = qcbc_het_create(ctx, params, het->fgrid, M, J, het->MN, het->MM, &h);
  if (st == QCBC_EDOM) {
    domain_error("heterodyne", st);
    return;
  }
  if (st) fail("heterodyne", st);
  const size_t J1 = J + 1;
  std::vector<double> amp((size_t)nd * M), phase((size_t)nd * M), IP(NR * J1 * J1), SL(NR * nd * J1),
      lc(NR * nd * J1), ls(NR * nd * J1), rc((size_t)nd * NR), rs((size_t)nd * NR), aa((size_t)nd * NR);
  double logLb = 0.0;
  if ((st = qcbc_het_export(h, nullptr, amp.data(), phase.data(), IP.data(), SL.data(), lc.data(), ls.data(),
                            rc.data(), rs.data(), aa.data())) ||
      (st = qcbc_het_info(h, nullptr, nullptr, nullptr, nullptr, &logLb)))
    fail("heterodyne", st);
  // the struct's arrays, allocated as QuickCBC.c:5128-5138 does
  het->IP = tensor(NR, J + 1, J + 1);
  het->SL = tensor(NR, nd, J + 1);
  het->amp = matrix(nd, M);
  het->phase = matrix(nd, M);
  het->rc = matrix(nd, NR);
  het->rs = matrix(nd, NR);
  het->aa = matrix(nd, NR);
  het->lc = tensor(NR, nd, J + 1);
  het->ls = tensor(NR, nd, J + 1);
  het->pref = (double*)std::malloc(sizeof(double) * (11 + 1));
  if (!het->pref) fail("heterodyne", QCBC_ENOMEM);
  std::memcpy(het->pref, params, sizeof(double) * 11);
  for (int d = 0; d < nd; d++) {
    std::memcpy(het->amp[d], amp.data() + (size_t)d * M, sizeof(double) * M);
    std::memcpy(het->phase[d], phase.data() + (size_t)d * M, sizeof(double) * M);
    std::memcpy(het->rc[d], rc.data() + (size_t)d * NR, sizeof(double) * NR);
    std::memcpy(het->rs[d], rs.data() + (size_t)d * NR, sizeof(double) * NR);
    std::memcpy(het->aa[d], aa.data() + (size_t)d * NR, sizeof(double) * NR);
  }
  for (int i = 0; i < NR; i++) {
    for (size_t j = 0; j < J1; j++) std::memcpy(het->IP[i][j], IP.data() + (i * J1 + j) * J1, sizeof(double) * J1);
    for (int d = 0; d < nd; d++) {
      const size_t o = ((size_t)i * nd + d) * J1;
      std::memcpy(het->SL[i][d], SL.data() + o, sizeof(double) * J1);
      std::memcpy(het->lc[i][d], lc.data() + o, sizeof(double) * J1);
      std::memcpy(het->ls[i][d], ls.data() + o, sizeof(double) * J1);
    }
  }
  het->logLb = logLb;
  std::printf("ref logL = %e\n", logLb);  // QuickCBC.c:5180
  HetEntry e = key_of(het);
  e.h = h;
  het_cache_put(e);
}

// The GPU heterodyne of a struct Het: the cached one heterodyne() made, or
// one made by the reference's own set-up, flattened and imported on first use
// (samples beyond NR J + 1 are never read, QuickCBC.c:5358-5417)
static qcbc_het* het_for(struct Net* net, struct Het* het, const char* where) {
  if (!net || net->Nifo < 1 || net->Nifo > 3 || !het || !het->IP || !het->SL || !het->lc || !het->ls ||
      !het->amp || !het->phase || !het->rc || !het->rs || !het->aa || !het->freq || !het->freq->data)
    fail(where, QCBC_EINVAL);
  const int M = het->M, J = het->J, NR = het->NR, nd = net->Nifo;
  if (J < 1 || NR < 1 || M < NR * J + 1 || het->freq->length < (size_t)M) fail(where, QCBC_EINVAL);
  HetEntry key = key_of(het);
  for (auto& e : t_hets)
    if (same_key(e, key)) {
      e.last_use = ++t_drop.tick;
      return e.h;
    }
  const int Mu = NR * J + 1;
  const size_t J1 = J + 1;
  qcbc_desc desc = desc_from(net, 2 * Mu < 4 ? 4 : 2 * Mu, net->Tobs);
  int st;
  qcbc_ctx* ctx = get_ctx(desc, nullptr, nullptr, &st);
  if (!ctx) fail(where, st);
  std::vector<double> amp((size_t)nd * Mu), phase((size_t)nd * Mu), IP(NR * J1 * J1), SL(NR * nd * J1),
      lc(NR * nd * J1), ls(NR * nd * J1), rc((size_t)nd * NR), rs((size_t)nd * NR), aa((size_t)nd * NR);
  for (int d = 0; d < nd; d++) {
    std::memcpy(amp.data() + (size_t)d * Mu, het->amp[d], sizeof(double) * Mu);
    std::memcpy(phase.data() + (size_t)d * Mu, het->phase[d], sizeof(double) * Mu);
    std::memcpy(rc.data() + (size_t)d * NR, het->rc[d], sizeof(double) * NR);
    std::memcpy(rs.data() + (size_t)d * NR, het->rs[d], sizeof(double) * NR);
    std::memcpy(aa.data() + (size_t)d * NR, het->aa[d], sizeof(double) * NR);
  }
  for (int i = 0; i < NR; i++) {
    for (size_t j = 0; j < J1; j++) std::memcpy(IP.data() + (i * J1 + j) * J1, het->IP[i][j], sizeof(double) * J1);
    for (int d = 0; d < nd; d++) {
      const size_t o = ((size_t)i * nd + d) * J1;
      std::memcpy(SL.data() + o, het->SL[i][d], sizeof(double) * J1);
      std::memcpy(lc.data() + o, het->lc[i][d], sizeof(double) * J1);
      std::memcpy(ls.data() + o, het->ls[i][d], sizeof(double) * J1);
    }
  }
  qcbc_het* h = nullptr;
  st = qcbc_het_import(ctx, Mu, J, het->freq->data, amp.data(), phase.data(), IP.data(), SL.data(), lc.data(),
                       ls.data(), rc.data(), rs.data(), aa.data(), het->logLb, &h);
  if (st) fail(where, st);
  key.h = h;
  het_cache_put(key);
  return h;
}

double log_likelihood_het(struct Net* net, struct Het* het, double* params, double Tobs) {
  (void)Tobs;  // fullphaseamp takes Tobs from net (QuickCBC.c:3799)
  if (!params) fail("log_likelihood_het", QCBC_EINVAL);
  qcbc_het* h = het_for(net, het, "log_likelihood_het");
  double logL = 0.0;
  int pst = 0;
  int st = qcbc_het_loglike(h, params, 1, &logL, &pst);
  if (st) fail("log_likelihood_het", st);
  if (domain_error("log_likelihood_het", pst)) return 0.0;
  return logL;
}

void Fisher_Het(struct Net* net, struct Het* het, double** fish, double* params, double Tobs) {
  (void)Tobs;  // fullphaseamp takes Tobs from net (QuickCBC.c:3799)
  if (!fish || !params) fail("Fisher_Het", QCBC_EINVAL);
  qcbc_het* h = het_for(net, het, "Fisher_Het");
  double F[11 * 11];
  int pst = 0;
  int st = qcbc_het_fisher(h, params, 1, F, nullptr, nullptr, &pst);
  if (st) fail("Fisher_Het", st);
  domain_error("Fisher_Het", pst);  // the matrix is then 1e4 I
  for (int i = 0; i < 11; i++) std::memcpy(fish[i], F + 11 * i, sizeof(double) * 11);
}

void FisherEvec(double** fish, double* eval, double** evec, int d) {
  if (!fish || !eval || !evec || d < 1 || d > FISHER_EVEC_MAX_D) fail("FisherEvec", QCBC_EINVAL);
  int st = drop_init();
  if (st) fail("FisherEvec", st);
  DropState& s = t_drop;
  const size_t dd = (size_t)d * d;
  if ((st = drop_buf(2 * dd + d))) fail("FisherEvec", st);
  double F[FISHER_EVEC_MAX_D * FISHER_EVEC_MAX_D], E[FISHER_EVEC_MAX_D * FISHER_EVEC_MAX_D], J[FISHER_EVEC_MAX_D];
  for (int i = 0; i < d; i++) std::memcpy(F + (size_t)i * d, fish[i], sizeof(double) * d);
  double *d_f = s.d_buf, *d_ev = d_f + dd, *d_ej = d_ev + dd;
  if (hipMemcpyAsync(d_f, F, sizeof(double) * dd, hipMemcpyHostToDevice, s.stream) != hipSuccess ||
      launch_fisher_evec(d_f, 1, d, d_ej, d_ev, nullptr, s.stream) != hipSuccess ||
      hipMemcpyAsync(E, d_ev, sizeof(double) * dd, hipMemcpyDeviceToHost, s.stream) != hipSuccess ||
      hipMemcpyAsync(J, d_ej, sizeof(double) * d, hipMemcpyDeviceToHost, s.stream) != hipSuccess ||
      hipStreamSynchronize(s.stream) != hipSuccess)
    fail("FisherEvec", QCBC_EHIP);
  for (int i = 0; i < d; i++) {
    eval[i] = J[i];
    std::memcpy(evec[i], E + (size_t)i * d, sizeof(double) * d);
  }
}

// ------------------------------------------------- the heterodyne's grid choice
// The PSD rows of a call on the thread's device buffer (these entry points
// take SN without D, so the analysis's bound data are left alone).
static const double* psd_on_device(const struct Net* net, double** SN, int N, const char* where) {
  const int n2 = N / 2, nd = net->Nifo;
  for (int d = 0; d < nd; d++)
    if (!SN[d]) fail(where, QCBC_EINVAL);
  int st = drop_init();
  if (st) fail(where, st);
  DropState& s = t_drop;
  if ((st = drop_buf((size_t)nd * n2))) fail(where, st);
  for (int d = 0; d < nd; d++)
    if (hipMemcpyAsync(s.d_buf + (size_t)d * n2, SN[d], sizeof(double) * n2, hipMemcpyHostToDevice, s.stream) !=
        hipSuccess)
      fail(where, QCBC_EHIP);
  if (hipStreamSynchronize(s.stream) != hipSuccess) fail(where, QCBC_EHIP);
  return s.d_buf;
}

static qcbc_ctx* grid_ctx(struct Net* net, RealVector* freq, int N, double Tobs, const char* where) {
  if (!net || net->Nifo < 1 || net->Nifo > 3 || N < 4 || !grid_ok(freq, N / 2, Tobs)) fail(where, QCBC_EINVAL);
  int st;
  qcbc_ctx* ctx = get_ctx(desc_from(net, N, Tobs), nullptr, nullptr, &st);
  if (!ctx) fail(where, st);
  return ctx;
}

void limits(struct Net* net, double* params, RealVector* freq, double** SN, int* stst, double* ratio, int N,
            double Tobs) {
  if (!params || !SN || !stst) fail("limits", QCBC_EINVAL);
  qcbc_ctx* ctx = grid_ctx(net, freq, N, Tobs, "limits");
  std::lock_guard<std::mutex> lk(ctx->mu);
  int st = limits_impl(ctx, (const double* const*)SN, params, stst, ratio, ctx->stream);
  if (domain_error("limits", st)) stst[0] = stst[1] = 0;
}

void Fisher_All(struct Net* net, double** fish, double* params, RealVector* freq, double** SN, int N, double Tobs) {
  if (!fish || !params || !SN) fail("Fisher_All", QCBC_EINVAL);
  qcbc_ctx* ctx = grid_ctx(net, freq, N, Tobs, "Fisher_All");
  const double* d_SN = psd_on_device(net, SN, N, "Fisher_All");
  double F[11 * 11];
  int pst = 0;
  std::lock_guard<std::mutex> lk(ctx->mu);
  int st = fisher_all_impl(ctx, d_SN, params, 1, F, &pst, ctx->stream);
  if (st) fail("Fisher_All", st);
  domain_error("Fisher_All", pst);  // the matrix is then 1e4 I
  for (int i = 0; i < 11; i++) std::memcpy(fish[i], F + 11 * i, sizeof(double) * 11);
}

void Fisher_Full(struct Net* net, double** fish, int d, double* params, RealVector* freq, double** SN, int N,
                 double Tobs) {
  if (!fish || !params || !SN || d < 1 || d > 7) fail("Fisher_Full", QCBC_EINVAL);
  qcbc_ctx* ctx = grid_ctx(net, freq, N, Tobs, "Fisher_Full");
  const double* d_SN = psd_on_device(net, SN, N, "Fisher_Full");
  double F[7 * 7];
  int pst = 0;
  std::lock_guard<std::mutex> lk(ctx->mu);
  int st = fisher_full_impl(ctx, d_SN, params, 1, d, F, &pst, ctx->stream);
  if (st) fail("Fisher_Full", st);
  domain_error("Fisher_Full", pst);  // the matrix is then 1e4 I
  for (int i = 0; i < d; i++) std::memcpy(fish[i], F + d * i, sizeof(double) * d);
}

double diff(struct Net* net, double* params, double* dparams, RealVector* freq, double** SM, int N, double Tobs) {
  if (!params || !dparams || !SM) fail("diff", QCBC_EINVAL);
  qcbc_ctx* ctx = grid_ctx(net, freq, N, Tobs, "diff");
  const double* d_SN = psd_on_device(net, SM, N, "diff");
  const int of = 0;
  double z = 0.0;
  std::lock_guard<std::mutex> lk(ctx->mu);
  int st = distances_impl(ctx, d_SN, nullptr, params, 1, dparams, &of, 1, &z, ctx->stream);
  if (st) fail("diff", st);
  return z;
}

double het_diff(struct Net* net, struct Het* het, double* params, double* dparams) {
  if (!params || !dparams) fail("het_diff", QCBC_EINVAL);
  qcbc_het* h = het_for(net, het, "het_diff");
  const int of = 0;
  double z = 0.0;
  int st = drop_init();
  if (st) fail("het_diff", st);
  st = distances_impl(nullptr, nullptr, h, params, 1, dparams, &of, 1, &z, t_drop.stream);
  if (st) fail("het_diff", st);
  return z;
}

void efix(struct Net* net, struct Het* het, double* params, double* min, double* max, RealVector* freq, double** SM,
          int N, double Tobs, double* eval, double** evec, double zs, int hflag) {
  if (!params || !min || !max || !eval || !evec || !(zs > 0.0)) fail("efix", QCBC_EINVAL);
  double ev[11 * 11];
  for (int i = 0; i < 11; i++) {
    if (!evec[i]) fail("efix", QCBC_EINVAL);
    std::memcpy(ev + 11 * i, evec[i], sizeof(double) * 11);
  }
  const double etamin = 0.05;  // ConstCBC.h:57
  int st;
  if (hflag == 0) {
    if (!SM) fail("efix", QCBC_EINVAL);
    qcbc_ctx* ctx = grid_ctx(net, freq, N, Tobs, "efix");
    const double* d_SN = psd_on_device(net, SM, N, "efix");
    std::lock_guard<std::mutex> lk(ctx->mu);
    st = efix_impl(ctx, d_SN, nullptr, params, 1, eval, ev, min, max, etamin, zs, nullptr, ctx->stream);
  } else {
    qcbc_het* h = het_for(net, het, "efix");
    if ((st = drop_init())) fail("efix", st);
    st = efix_impl(nullptr, nullptr, h, params, 1, eval, ev, min, max, etamin, zs, nullptr, t_drop.stream);
  }
  if (st) fail("efix", st);
}

void het_space(struct Net* net, struct Het* het, double* params, double* min, double* max, RealVector* freq,
               double** SN, int N, double Tobs) {
  if (!het || !params || !min || !max || !SN) fail("het_space", QCBC_EINVAL);
  qcbc_ctx* ctx = grid_ctx(net, freq, N, Tobs, "het_space");
  const double* d_SN = psd_on_device(net, SN, N, "het_space");
  std::vector<int> g(N / 2);
  int MN = 0, MM = 0, M = 0, NR = 0;
  int st;
  {
    std::lock_guard<std::mutex> lk(ctx->mu);
    st = het_space_impl(ctx, (const double* const*)SN, d_SN, params, min, max, 0.05, &MN, &MM, &M, &NR, g.data(),
                        (int)g.size(), nullptr, nullptr, nullptr, ctx->stream);
  }
  if (st == QCBC_EDOM) {
    domain_error("het_space", st);
    return;
  }
  if (st) fail("het_space", st);
  het->MN = MN;
  het->MM = MM;
  het->J = 2;
  het->NR = NR;
  het->M = M;
  // QuickCBC.c:4887-4891
  het->freq = CreateRealVector(M);
  het->fgrid = (int*)std::malloc(sizeof(int) * (M + 1));
  if (!het->fgrid) fail("het_space", QCBC_ENOMEM);
  for (int i = 0; i < M; i++) {
    het->fgrid[i] = g[i];
    het->freq->data[i] = ((double)g[i]) / Tobs;
  }
}

// ------------------------------------------------- search likelihoods
double log_likelihood_APmax(struct Net* net, double** D, double* params, RealVector* freq, double** SN, int N,
                            double Tobs) {
  if (!net || net->Nifo < 1 || net->Nifo > 3 || !D || !SN || !params || !grid_ok(freq, N / 2, Tobs))
    fail("log_likelihood_APmax", QCBC_EINVAL);
  int st;
  qcbc_ctx* ctx = get_ctx(desc_from(net, N, Tobs), D, SN, &st);
  if (!ctx) fail("log_likelihood_APmax", st);
  double logL = 0.0;
  int pst = 0;
  {
    std::lock_guard<std::mutex> lk(ctx->mu);
    st = search_apmax(ctx, params, 1, &logL, &pst);
  }
  if (st) fail("log_likelihood_APmax", st);
  if (domain_error("log_likelihood_APmax", pst)) return 0.0;
  return logL;
}

double log_likelihood_max(struct Net* net, double** D, double* params, RealVector* freq, double** SN, int N,
                          double Tobs, double tmin, double tmax, int pflag) {
  (void)pflag;  // its prints are not made
  if (!net || net->Nifo < 1 || net->Nifo > 3 || !D || !SN || !params || !grid_ok(freq, N / 2, Tobs) ||
      (net->Nifo > 1 && !net->delays) || !(tmin < tmax))
    fail("log_likelihood_max", QCBC_EINVAL);
  const int nd = net->Nifo;
  double dl[9] = {0};
  for (int i = 0; i < nd && nd > 1; i++) {
    if (!net->delays[i]) fail("log_likelihood_max", QCBC_EINVAL);
    for (int j = 0; j < nd; j++) dl[i * nd + j] = net->delays[i][j];
  }
  int st;
  qcbc_ctx* ctx = get_ctx(desc_from(net, N, Tobs), D, SN, &st);
  if (!ctx) fail("log_likelihood_max", st);
  double logL = 0.0;
  int pst = 0;
  {
    std::lock_guard<std::mutex> lk(ctx->mu);
    st = search_llmax(ctx, params, 1, tmin, tmax, net->tmax, dl, &logL, &pst);
  }
  if (st) fail("log_likelihood_max", st);
  if (domain_error("log_likelihood_max", pst)) return 0.0;
  return logL;
}

double log_likelihood_penalized(double* params, double* D, double* twave, double* SN, double* Ax, double* Px, int N,
                                double Tobs, int pflag) {
  (void)pflag;  // its prints are not made
  if (!params || !D || !twave || !SN || !Ax || !Px || N < 4 || !(Tobs > 0)) fail("log_likelihood_penalized", QCBC_EINVAL);
  int st = drop_init();
  if (st) fail("log_likelihood_penalized", st);
  double logL = 0.0;
  st = search_penalized_one(params, D, twave, SN, N, Tobs, g_fmin.load(), g_fmax.load(), &logL, Ax, Px, t_drop.stream);
  if (st) fail("log_likelihood_penalized", st);
  return logL;
}

void PDwave(double* wavef, RealVector* freq, double tcmin, double* params, int N) {
  if (!wavef || !freq || !freq->data || freq->length < 2 || !params || N < 4) fail("PDwave", QCBC_EINVAL);
  const double Tobs = 1.0 / freq->data[1];  // QuickCBC.c:3459
  if (!grid_ok(freq, N / 2, Tobs)) fail("PDwave", QCBC_EINVAL);
  int st = drop_init();
  if (st) fail("PDwave", st);
  st = search_pdwave(params, tcmin, N, Tobs, g_fref.load(), freq->data[0], wavef, t_drop.stream);
  if (st == QCBC_EDOM) {
    if (domain_error("PDwave", st)) std::fill(wavef, wavef + N, 0.0);
    return;
  }
  if (st) fail("PDwave", st);
}

void tukey(double* data, double alpha, int N) {
  if (!data || N < 1) fail("tukey", QCBC_EINVAL);
  int st = drop_init();
  if (st) fail("tukey", st);
  st = ingest_tukey(data, alpha, N, t_drop.stream);
  if (st) fail("tukey", st);
}

void tukey_scale(double* s1, double* s2, double alpha, int N) {
  if (!s1 || !s2 || N < 1) fail("tukey_scale", QCBC_EINVAL);
  int st = drop_init();
  if (st) fail("tukey_scale", st);
  double s12[2];
  st = ingest_tukey_scale(alpha, N, s12, t_drop.stream);
  if (st) fail("tukey_scale", st);
  *s1 = s12[0];
  *s2 = s12[1];
}

void whiten(double* data, double* Sn, int N) {
  if (!data || !Sn || N < 2 || (N & 1)) fail("whiten", QCBC_EINVAL);
  int st = drop_init();
  if (st) fail("whiten", st);
  st = ingest_whiten(data, Sn, N, t_drop.stream);
  if (st) fail("whiten", st);
}

void pbt_shift(double* corr, double* corrf, double* data1, double* data2, double* Sn, int imin, int imax, int N) {
  if (!corr || !corrf || !data1 || !data2 || !Sn || N < 4 || (N & (N - 1)) || imin < 1 || imax > N / 2)
    fail("pbt_shift", QCBC_EINVAL);
  int st = drop_init();
  if (st) fail("pbt_shift", st);
  st = search_pbt_shift(corr, corrf, data1, data2, Sn, imin, imax, N, t_drop.stream);
  if (st) fail("pbt_shift", st);
}

// IMRPhenomD_internals.c:141-181
AmpPhaseFDWaveform* CreateAmpPhaseFDWaveform(size_t length) {
  AmpPhaseFDWaveform* wf = (AmpPhaseFDWaveform*)std::malloc(sizeof(AmpPhaseFDWaveform));
  if (!wf) fail("CreateAmpPhaseFDWaveform", QCBC_ENOMEM);
  wf->length = length;
  wf->freq = (double*)std::calloc(length ? length : 1, sizeof(double));
  wf->amp = (double*)std::calloc(length ? length : 1, sizeof(double));
  wf->phase = (double*)std::calloc(length ? length : 1, sizeof(double));
  if (!wf->freq || !wf->amp || !wf->phase) fail("CreateAmpPhaseFDWaveform", QCBC_ENOMEM);
  return wf;
}

void DestroyAmpPhaseFDWaveform(AmpPhaseFDWaveform* wf) {
  if (!wf) return;
  std::free(wf->freq);
  std::free(wf->amp);
  std::free(wf->phase);
  std::free(wf);
}

RealVector* CreateRealVector(size_t length) {
  RealVector* v = (RealVector*)std::malloc(sizeof(RealVector));
  if (!v) fail("CreateRealVector", QCBC_ENOMEM);
  v->length = length;
  v->data = (double*)std::calloc(length ? length : 1, sizeof(double));
  if (!v->data) fail("CreateRealVector", QCBC_ENOMEM);
  return v;
}

void DestroyRealVector(RealVector* v) {
  if (!v) return;
  std::free(v->data);
  std::free(v);
}

}  // extern "C"
