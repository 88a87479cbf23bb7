// The fused likelihood on device memory without any allocation in the call:
// the point records and the per-split partial sums live in a file-scope
// device array (one copy per GPU, held until the process ends) instead of
// the context's batch scratch, which qcbc_loglike_device grows with
// hipMalloc when a batch is larger than any before.  For callers that may
// not allocate: code under stream capture, tight serving loops.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/qcbc.h"
#include "qcbc_api_internal.h"
#include "qcbc_internal.h"

using namespace qcbc;

namespace {

constexpr size_t round256(size_t n) { return (n + 255) / 256 * 256; }

// layout of the array: QCBC_NOALLOC_MAX_POINTS point records, then the
// partial sums (2 n_det doubles per point and split, at most 3 detectors and
// QCBC_MAX_SPLITS splits)
constexpr size_t kPtsBytes = round256(sizeof(Point) * QCBC_NOALLOC_MAX_POINTS);
constexpr size_t kPartialBytes = sizeof(double) * 6 * (size_t)QCBC_MAX_SPLITS * QCBC_NOALLOC_MAX_POINTS;
constexpr size_t kScratchBytes = kPtsBytes + kPartialBytes;

}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
std::mutex g_mu;
void* g_scratch_ptr[64] = {};  // per device: the array's device address
}  // namespace

// Called once per device by the library's device set-up, with that device
// current.  The array's name is only a host-side handle; its address on this
// device comes from the runtime and is the one every launch uses.
int qcbc::noalloc_init_device(int dev) {
  if (dev < 0 || dev >= 64) return QCBC_ENODEV;
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_scratch_ptr[dev]) return QCBC_OK;
  void* p = nullptr;
  hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    t_last_hip = e;
    return QCBC_EHIP;
  }
  g_scratch_ptr[dev] = p;
  return QCBC_OK;
}

size_t qcbc::noalloc_scratch_bytes() { return kScratchBytes; }

extern "C" {

int qcbc_loglike_device_noalloc(qcbc_ctx* c, const double* d_params, int B, int precision, double* d_logL,
                                double* d_rho, int* d_status) {
  if (!c || B < 0 || (precision != QCBC_FP64 && precision != QCBC_MIXED)) return QCBC_EINVAL;
  if (B > QCBC_NOALLOC_MAX_POINTS) return QCBC_ERANGE;
  if (B == 0) return QCBC_OK;
  if (!d_params || !d_logL || !d_rho) return QCBC_EINVAL;
  std::lock_guard<std::mutex> lk(c->mu);
  if (!c->bound) return QCBC_ENOTBOUND;
  if (hipSetDevice(c->device) != hipSuccess) return QCBC_ENODEV;
  void* base = nullptr;
  {
    std::lock_guard<std::mutex> g(g_mu);
    base = g_scratch_ptr[c->device];
  }
  if (!base) return QCBC_ENODEV;
  Point* pts = static_cast<Point*>(base);
  double* partial = reinterpret_cast<double*>(static_cast<unsigned char*>(base) + kPtsBytes);
  hipError_t e = launch_setup(d_params, B, c->an, pts, d_status, c->stream);
  // every (point, split) partial is written by the likelihood kernel, so
  // nothing in the array needs clearing between calls
  const LLInst in = make_inst(c, c->tile[precision], B, effective_splits(c, precision, B), pts, partial);
  if (e == hipSuccess) e = launch_ll(precision, in, c->cpb[precision], c->tile[precision], c->stream);
  if (e == hipSuccess)
    e = launch_finalize(partial, in.splits, c->desc.n_det, pts, B, precision == QCBC_MIXED ? FIN_MIXED : FIN_FP64,
                        c->d_wref, d_logL, d_rho, c->stream);
  if (e != hipSuccess) {
    t_last_hip = e;
    return QCBC_EHIP;
  }
  return QCBC_OK;
}

}  // extern "C"
