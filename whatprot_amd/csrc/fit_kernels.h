// Launch interface of the Baum-Welch E-step kernel (host side).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>

#include "hmm_device.h"

namespace wp {

// Arguments of the E-step kernel. Device pointers unless noted.
struct FitArgs {
    DevModel m;                   // model of this iteration, uncut
    int N;                        // reads
    const double* rads;           // [N][T][C], intensities / mu
    int NB;                       // prod_c (D_c + 1): count states per Edman row
    int D[WP_MAX_CHANNELS];       // dyes of each channel in the dye-seq
    const unsigned char* crd;     // [NB][C] per-channel counts of count index q
    const unsigned short* fbase;  // per channel c, NB / (D_c + 1) fiber base indices
    int nfib;                     // total fiber bases
    const signed char* seq;       // [T] channel of position e (-1 = none)
    const unsigned char* track;   // [T][C] dyes of channel c at positions >= e
    const double* binom;          // [2][C][tri] dud then bleach tables
    int tri;
    void* hist;                   // per workgroup: T x 2 x T*NB states (Real)
    size_t hist_stride;           // elements per workgroup
    double* partial;              // [grid][L + 1] per-workgroup totals
    int L;                        // WP_FIT_STATS_LEN(T, C)
    const int* weight;            // [rounds][N] multiplicity of each read (bootstrap resamples), or null = 1
    // Several E-steps in one launch (bootstrap rounds): nslots models,
    // binom holds nslots tables back to back, partial nslots x grid totals,
    // d_out nslots x (L + 1) results.
    int nslots;                   // 1 for a plain E-step
    const DevModel* models;       // [nslots] device copies, or null: a.m
    const int* slot_round;        // [nslots] weight row of each slot, or null: row 0
    int rows;                     // 1: row-lane kernel (fit_rows_supports), hist per wavefront
};

// Dynamic LDS of one workgroup ((size_t)-1 for an unsupported C).
size_t fit_lds_bytes(int precision, int T, int C, int NB, int tri, int nfib);
int fit_threads();

// E-step over all reads; d_out[0..L) receives the summed statistics and
// d_out[L] the sum of ln P. grid workgroups, each owning hist and partial
// slots.
hipError_t launch_fit_estep(const FitArgs& a, int precision, int grid, double* d_out, hipStream_t stream);

// Row-lane E-step (one wavefront per read, lane = tensor x Edman row, count
// states in registers) for T <= 32 and the count shapes it is compiled for.
bool fit_rows_supports(int T, int C, const int* D);
int fit_rows_threads();
size_t fit_rows_lds_bytes(int precision, int T, int C, int NB, int tri);
// history elements of one workgroup (T x NB x 64 per wavefront)
size_t fit_rows_hist_per_group(int T, int NB);
// workgroups of the row-lane kernel resident per CU (its register bound)
int fit_rows_groups_per_cu(int precision);
// the E-step launch alone (per-workgroup totals into a.partial)
hipError_t launch_fit_rows(const FitArgs& a, int precision, int grid, hipStream_t stream);

}  // namespace wp
