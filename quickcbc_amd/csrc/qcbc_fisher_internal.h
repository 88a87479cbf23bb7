// Launch wrappers of the Fisher-matrix and eigen-decomposition kernels
// (qcbc_fisher.hip).  Host-safe: no device code in here.
#pragma once

#include <hip/hip_runtime_api.h>

#include "qcbc_types.h"

namespace qcbc {

constexpr int FISHER_NP = 11;                 // parameters of a point (NP, ConstCBC.h:68)
constexpr int FISHER_NV = 2 * FISHER_NP + 1;  // parameter sets per point: itself, then +eps / -eps each
constexpr int FISHER_EVEC_MAX_D = FISHER_NP;  // largest matrix the eigen kernel takes

// the FISHER_NV parameter sets of each of B points: q is B x FISHER_NV x 11
hipError_t launch_fisher_params(const double* d_params, int B, double* d_q, hipStream_t s);
constexpr int FISHER_NPAIR = FISHER_NP * (FISHER_NP + 1) / 2;  // entries of the upper triangle

// Fisher_Het for B points from the set-up records of their FISHER_NV
// parameter sets (pts, B x FISHER_NV), on a heterodyne's grid: fish is
// B x 11 x 11, status B.  The same kernel gives Fisher_All on a grid of
// every band bin (ha null, w = 1/S per detector and bin), split over
// `splits` blocks per point with partial (B x splits x FISHER_NPAIR)
// scratch when splits > 1.
hipError_t launch_het_fisher(const Point* pts, const double* d_params, int B, int nd, const double* freq,
                             const double* pl, const double* ha, const double* w, int M, double* fish, int* status,
                             hipStream_t s, int splits = 1, double* partial = nullptr);
// Fisher_Full (QuickCBC.c:2551-2688): the parameters it varies (NX,
// ConstCBC.h:67), its variant parameter sets (B x (2 FISHER_FULL_NP + 1) x
// FISHER_FULL_NP from B points of ips entries), and the matrices (B x 7 x 7)
// from their set-up records, over band bins imin .. imin + nb - 1 of the PSD
// SN (nd x n2) split over `splits` blocks per point with partial
// (B x splits x 28) scratch
constexpr int FISHER_FULL_NP = 7;
constexpr int FISHER_FULL_NV = 2 * FISHER_FULL_NP + 1;
constexpr int FISHER_FULL_NPAIR = FISHER_FULL_NP * (FISHER_FULL_NP + 1) / 2;
hipError_t launch_fisher_full_params(const double* d_params, int ips, int B, double* d_q, hipStream_t s);
hipError_t launch_fisher_full(const Point* pts, const double* d_q, const double* d_params, int ips, int B, int nd,
                              const double* SN, int n2, int imin, int nb, double Tobs, int splits, double* partial,
                              double* fish, int* status, hipStream_t s);
// FisherEvec of B d x d matrices (d <= FISHER_EVEC_MAX_D): ej B x d, ev B x d x d,
// sweeps B (optional)
hipError_t launch_fisher_evec(const double* fish, int B, int d, double* ej, double* ev, int* sweeps, hipStream_t s);

}  // namespace qcbc
