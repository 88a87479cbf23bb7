// QuickCBC's PT step (update(), QuickCBC.c:2238-2405; MCMC_all's step,
// :2029-2092) on the GPU: constants, random-draw layout and launch wrappers
// (qcbc_update.hip).  Host-safe: no device code in here.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdint>

#include "qcbc_types.h"

namespace qcbc {

// Prior box and proposal mix of one ensemble.
struct UpdateConst {
  double min[11], max[11];  // prior bounds per parameter (update()'s min, max)
  double a, b, c;           // proposal thresholds on alpha (0.5, 0.3, 0.2 in the reference; c = 0 here)
  double etamin;            // lower limit of the symmetric mass ratio (ConstCBC.h:57)
  int n_hist;               // history rows per temperature (NH, ConstCBC.h:72)
  int hist_every;           // steps between history entries (20, QuickCBC.c:2071)
  // the analysis's detectors, for the sky ring
  double gmst;
  int n_det, labels[3];
};

// Random streams (xor-ed into the seed) and draw indices within a stream.
constexpr uint64_t SALT_PROPOSE = 0x6a09e667f3bcc908ULL;
constexpr uint64_t SALT_ACCEPT = 0xbb67ae8584caa73bULL;
constexpr uint64_t SALT_SWAP = 0x3c6ef372fe94f82bULL;
constexpr uint64_t SALT_HIST = 0xa54ff53a5f1d36f1ULL;
constexpr int D_ALPHA = 0;     // proposal type
constexpr int D_FISHER = 1;    // eigendirection, then a normal (2 draws)
constexpr int D_DE = 4;        // history row i, then row j: one draw per try
constexpr int DE_TRIES = 32;
constexpr int D_DE_SCALE = 40; // whether the scale is N(0, 0.5), then a normal
constexpr int D_JIGGLE = 50;   // the scale's exponent, then 11 normals (2 draws each)
constexpr int D_SKY = 80;      // first detector of the pair, then the second: one draw per try
constexpr int D_SKY_ROT = 120; // the rotation about the pair's baseline
constexpr int D_SKY_FLIP = 121;// the half-period shift of phase and polarisation

// angles wrapped into range after a jump only within this magnitude (rad)
constexpr double ANGLE_WRAP_MAX = 1.0e4;

// per column: update steps, swap steps, swaps accepted, then (proposed,
// accepted) for the Fisher, DE, jiggle and sky-ring proposals
constexpr int UPD_STATS = 11;

hipError_t launch_update_hist_init(const double* x, int n_temps, int n_chains, int c0, int ncols, int n_hist,
                                   uint64_t seed, double* hist, hipStream_t s);
hipError_t launch_update_propose(const double* x, const double* ej, const double* ev, const double* hist,
                                 const double* tc, const UpdateConst& u, int n_temps, int n_chains, int c0, int ncols,
                                 uint64_t seed, const unsigned long long* step, double* y, int* typ, double* jac,
                                 hipStream_t s);
// fills the fused step kernel's copy of the QNM spline slopes (once per device)
hipError_t launch_update_qnm_init(hipStream_t s);

// The heterodyned likelihood's arrays (qcbc_het.cpp's grid, nd x M rows).
struct HetArgs {
  const double *freq, *pl, *ha, *hp, *wf;
  int M, nd;
  double logLb;
};

// nsteps whole PT steps of a group of chain columns in one launch (block b
// runs column c0 + b through all of them): the set-up and heterodyned logL of
// the pending proposals y (skipped on a swap step, whose proposals are
// discarded), acceptance or the swap, the history and the next proposals.
hipError_t launch_pt_steps(double* x, double* L, double* ej, double* ev, double* hist, unsigned long long* hcount,
                           double* y, int* typ, double* jac, const double* tc, const UpdateConst& u,
                           const AnalysisConst& an, const HetArgs& h, int n_temps, int n_chains, int c0, int ncols,
                           uint64_t seed, int nsteps, unsigned long long* step, unsigned long long* stats,
                           unsigned int* done, hipStream_t s);


}  // namespace qcbc
