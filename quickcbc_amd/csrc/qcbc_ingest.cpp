// QuickCBC's input and output text formats (QuickCBC.c:195-248, :267-306)
// and the GPU halves of the small conditioning drop-ins (tukey, tukey_scale,
// whiten; Utilities.c:294-309, :635-683).
//
// The readers parse the whole file in memory with strtod, which converts
// as scanf's %lf does (both correctly rounded), and reproduce the
// reference's count of a framed_* file: its loop `while (!feof(in)) {
// fscanf(in, "%lf%lf", ...); N++; }` from N = -1 counts the fscanf calls made
// before the end of the file is seen, and the end is seen by the call that
// finds no further number or whose last number runs up to the end.
#include <hip/hip_runtime_api.h>

#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "../../include/qcbc.h"
#include "qcbc_api_internal.h"
#include "qcbc_condition_internal.h"
#include "qcbc_internal.h"

#define HIPCHK(x)                \
  do {                           \
    hipError_t e_ = (x);         \
    if (e_ != hipSuccess) {      \
      qcbc::t_last_hip = e_;     \
      return QCBC_EHIP;          \
    }                            \
  } while (0)

namespace qcbc {
namespace {

int slurp(const char* path, std::string& out) {
  FILE* f = std::fopen(path, "rb");
  if (!f) return QCBC_EIO;
  char chunk[1 << 16];
  size_t n;
  while ((n = std::fread(chunk, 1, sizeof(chunk), f)) > 0) out.append(chunk, n);
  const bool bad = std::ferror(f) != 0;
  std::fclose(f);
  return bad ? QCBC_EIO : QCBC_OK;
}

// scanf's %lf over an in-memory file: 1 converted, 0 a token that is not a
// number, -1 the end of the file (sets eof, as the stream's flag)
struct Scanner {
  const char* p;
  const char* end;
  bool eof = false;
  int next(double* v) {
    while (p < end && std::isspace((unsigned char)*p)) p++;
    if (p == end) {
      eof = true;
      return -1;
    }
    char* q = nullptr;
    const double x = std::strtod(p, &q);
    if (q == p) return 0;
    p = q;
    if (p >= end) eof = true;  // the conversion looked past the last character
    *v = x;
    return 1;
  }
};

// device buffers of one drop-in call, freed on every return path
struct Temp {
  std::vector<void*> bufs;
  ~Temp() {
    for (void* b : bufs) (void)hipFree(b);
  }
  int alloc(size_t n, double** out) {
    void* b = nullptr;
    if (hipMalloc(&b, sizeof(double) * (n ? n : 1)) != hipSuccess) return QCBC_ENOMEM;
    bufs.push_back(b);
    *out = static_cast<double*>(b);
    return QCBC_OK;
  }
};

}  // namespace

int ingest_tukey(double* data, double alpha, int N, hipStream_t s) {
  Temp t;
  double* d;
  int st = t.alloc((size_t)N, &d);
  if (st) return st;
  HIPCHK(hipMemcpyAsync(d, data, sizeof(double) * N, hipMemcpyHostToDevice, s));
  HIPCHK(launch_tukey(d, 1, N, alpha, s));
  HIPCHK(hipMemcpyAsync(data, d, sizeof(double) * N, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  return QCBC_OK;
}

int ingest_tukey_scale(double alpha, int N, double* s12, hipStream_t s) {
  Temp t;
  double* d;
  int st = t.alloc(2, &d);
  if (st) return st;
  HIPCHK(launch_tukey_scale(N, alpha, d, s));
  HIPCHK(hipMemcpyAsync(s12, d, sizeof(double) * 2, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  return QCBC_OK;
}

int ingest_whiten(double* data, const double* Sn, int N, hipStream_t s) {
  Temp t;
  double *d, *S;
  int st;
  if ((st = t.alloc((size_t)N, &d)) || (st = t.alloc((size_t)N / 2, &S))) return st;
  HIPCHK(hipMemcpyAsync(d, data, sizeof(double) * N, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(S, Sn, sizeof(double) * (N / 2), hipMemcpyHostToDevice, s));
  HIPCHK(launch_whiten(d, S, N, s));
  HIPCHK(hipMemcpyAsync(data, d, sizeof(double) * N, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  return QCBC_OK;
}

}  // namespace qcbc

using namespace qcbc;

extern "C" {

int qcbc_data_file_name(char* buf, size_t len, const char* kind, double Tobs, double ttrig, int label) {
  if (!buf || !len || !kind) return QCBC_EINVAL;
  const int n = std::snprintf(buf, len, "%s_%d_%d_%d.dat", kind, (int)(Tobs), (int)ttrig, label);
  if (n < 0 || (size_t)n >= len) return QCBC_ERANGE;
  return QCBC_OK;
}

int qcbc_read_framed(const char* path, int cap, double* time, double* x, int* N) {
  if (!path || !N || cap < 0) return QCBC_EINVAL;
  std::string buf;
  int st = slurp(path, buf);
  if (st) return st;
  buf.push_back('\0');  // strtod stops at the terminator
  const char* b = buf.data();
  const char* e = b + buf.size() - 1;
  // the reference's count (QuickCBC.c:204-210)
  Scanner sc{b, e};
  int n = -1;
  while (!sc.eof) {
    double a, c;
    int r = sc.next(&a);
    if (r == 0) return QCBC_EINVAL;
    if (r == 1 && sc.next(&c) == 0) return QCBC_EINVAL;
    n++;
  }
  *N = n;
  if (!time && !x) return QCBC_OK;
  if (n > cap) return QCBC_ERANGE;
  // the first N pairs (QuickCBC.c:219-220)
  Scanner rd{b, e};
  for (int i = 0; i < n; i++) {
    double a, c;
    if (rd.next(&a) != 1 || rd.next(&c) != 1) return QCBC_EINVAL;
    if (time) time[i] = a;
    if (x) x[i] = c;
  }
  return QCBC_OK;
}

int qcbc_read_spec(const char* path, int n2, double* SN, double* SM) {
  if (!path || n2 < 0) return QCBC_EINVAL;
  std::string buf;
  int st = slurp(path, buf);
  if (st) return st;
  buf.push_back('\0');
  Scanner rd{buf.data(), buf.data() + buf.size() - 1};
  for (int i = 0; i < n2; i++) {
    double v[4];
    for (int k = 0; k < 4; k++)
      if (rd.next(&v[k]) != 1) return QCBC_EINVAL;
    if (SN) SN[i] = v[1];
    if (SM) SM[i] = v[2];
  }
  return QCBC_OK;
}

int qcbc_write_table(const char* path, int nrows, int ncols, const double* const* cols, int digits) {
  if (!path || nrows < 0 || ncols < 0 || (ncols > 0 && !cols) || digits < 0 || digits > 30) return QCBC_EINVAL;
  for (int c = 0; c < ncols; c++)
    if (!cols[c] && nrows > 0) return QCBC_EINVAL;
  FILE* f = std::fopen(path, "w");
  if (!f) return QCBC_EIO;
  for (int r = 0; r < nrows; r++) {
    for (int c = 0; c < ncols; c++) std::fprintf(f, "%.*e ", digits, cols[c][r]);
    std::fputc('\n', f);
  }
  const bool bad = std::ferror(f) != 0;
  if (std::fclose(f) != 0 || bad) return QCBC_EIO;
  return QCBC_OK;
}

}  // extern "C"
