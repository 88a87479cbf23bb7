// Launch wrappers shared by the kernels (qcbc_kernels.hip) and the host side
// of libqcbc (qcbc_api.cpp).  Host-safe: no device code in here.
#pragma once

#include <hip/hip_runtime_api.h>
#include <stddef.h>
#include <stdint.h>

#include "qcbc_types.h"

namespace qcbc {

hipError_t launch_qnm_init(hipStream_t s);
// k_setup for templates' reference-frame parameters, ips per point (7 + 3 (n_det - 1))
hipError_t launch_setup_intrinsic(const double* d_params, int ips, int B, const AnalysisConst& an, Point* d_pts,
                                  int* d_status, hipStream_t s);
// log_likelihood_intrinsic's logL and per-detector snrsq from FP64 partials
hipError_t launch_finalize_intrinsic(const double* partial, int splits, int nd, const Point* pts, int B,
                                     double* logL, double* snrsq, hipStream_t s);
hipError_t launch_setup(const double* d_params, int B, const AnalysisConst& an, Point* d_pts, int* d_status,
                        hipStream_t s);
hipError_t launch_setup_h22(double m1, double m2, double c1, double c2, double dist, double phi0, double fref,
                            double fmin0, Point* d_pts, int* d_status, hipStream_t s);
hipError_t launch_bind(const double* D, const double* SN, int nd, int N, double Tobs, int imin, int nb, int nbp,
                       const double* wref, double* p64, float* p32, hipStream_t s);
size_t ll_lds_bytes(int precision, int nd, int tile);
// LDS tile lengths the fused kernels are built for: 128, 256, 512 bins
bool ll_tile_ok(int tile);
hipError_t launch_ll(int precision, const LLInst& in, int cpb, int tile, hipStream_t s);
hipError_t launch_ll_queue(int precision, const LLInst* d_insts, const LLItem* d_items, int n_items, int* d_counter,
                           int max_nd, int cpb, int tile, int blocks, hipStream_t s);
hipError_t launch_fill_i32(int* p, int n, int v, hipStream_t s);
// partial sums as written by: the plain path (raw), k_ll / k_ll_queue FP64,
// mixed; FIN_DONE: logL is already final (the heterodyned likelihood)
enum { FIN_RAW = 0, FIN_FP64 = 1, FIN_MIXED = 2, FIN_DONE = 3 };
hipError_t launch_finalize(const double* partial, int splits, int nd, const Point* pts, int B, int mode,
                           const double* wref, double* logL, double* rho, hipStream_t s);
hipError_t launch_waveform(const Point* pts, int B, int nhalf, double Tobs, double* amp, double* phs, hipStream_t s);
hipError_t launch_waveform_freqs(const Point* pts, const double* freq, int n, double* amp, double* phs,
                                 hipStream_t s);
hipError_t launch_project(const Point* pts, int B, const double* amp, const double* phs, int nd, int N, double Tobs,
                          double sqT, double* hw, hipStream_t s);
hipError_t launch_nwip(const double* hw, const double* D, const double* SN, int B, int nd, int N, int imin, int imax,
                       int nblk, double* partial, hipStream_t s);
hipError_t launch_nwip_pair(const double* a, const double* b, const double* Sn, int imin, int imax, int N, int nblk,
                            double* partial, double* out, hipStream_t s);
hipError_t launch_geometry(int what, const double* in, int n, double* out, hipStream_t s);

}  // namespace qcbc

namespace qcbc {
// per-temperature constants tc[3 n_temps]: sqrt(T), 1/T, ladder^-t
hipError_t launch_pt_temps(double* tc, int n_temps, double ladder, hipStream_t s);
// proposals of the group of chain columns [c0, c0 + ncols) at step *step
hipError_t launch_pt_propose(const double* x, double* y, const double* sigma, const double* tc, int n_temps,
                             int n_chains, int c0, int ncols, uint64_t seed, const unsigned long long* step,
                             hipStream_t s);
// finalize + accept + swap + next proposals + step counter of one group, one
// launch (block per chain column)
hipError_t launch_pt_post(const LLInst& in, int mode, const double* wref, double* x, double* L, double* y, double* Ly,
                          double* rho, const int* st_y, const double* sigma, const double* tc, int n_temps,
                          int n_chains, int c0, int ncols, uint64_t seed, unsigned long long* step,
                          unsigned long long* acc, unsigned long long* swaps, unsigned int* done, hipStream_t s);
hipError_t launch_fill_u64(unsigned long long* p, int n, unsigned long long v, hipStream_t s);
hipError_t launch_copy(const void* src, void* dst, size_t bytes, hipStream_t s);
}  // namespace qcbc
