// Launch wrappers of the grid-choice kernels (qcbc_gridchoice.hip).
// Host-safe: no device code in here.
#pragma once

#include <hip/hip_runtime_api.h>

#include "qcbc_types.h"

namespace qcbc {

// One distance of efix: set-up records a (the point the jumps start from) and
// b (the displaced point); for the full-data distance the bin range [lo, hi)
// of the analysis grid (diff's, from point a)
struct DiffPair {
  int a, b, lo, hi;
};

// splits per pair of the full-data distance for ranges up to max_bins bins
int diff_splits(int max_bins);
// diff (QuickCBC.c:4913-4964) of K pairs into z (K); SN nd rows of n2 on the
// device; partial K x splits scratch
hipError_t launch_diff_full(const Point* pts, const DiffPair* pairs, int K, int nd, const double* SN, int n2,
                            double Tobs, int max_bins, int splits, double* partial, double* z, hipStream_t s);
// het_diff (QuickCBC.c:4966-5053) of K pairs on a heterodyne's grid (freq M,
// sample powers pl 4 x M, folded weights wf whose first nd x M are wH / a_ref^2)
hipError_t launch_diff_het(const Point* pts, const DiffPair* pairs, int K, int nd, const double* freq,
                           const double* pl, const double* wf, int M, double* z, hipStream_t s);
// the band bins [imin, imin + nb) as a grid for Fisher_All: freq nb, w nd x nb (1/S)
hipError_t launch_band_grid(const double* SN, int nd, int n2, int imin, int nb, double Tobs, double* freq, double* w,
                            hipStream_t s);

}  // namespace qcbc
