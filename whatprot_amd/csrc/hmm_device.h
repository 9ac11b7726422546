// Device-side model arithmetic shared by the HMM kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/whatprot_amd.h"

namespace wp {

// Everything a kernel needs from the sequencing model, passed by value.
struct DevModel {
    int T, C, M;          // timesteps, channels, max_num_dyes
    int uncut;            // dist_cutoff == DBL_MAX
    double cutoff;        // dist_cutoff
    double pf, p_ib, p_cb;
    double p_detach[WP_MAX_TIMESTEPS];  // DecayingRateModel values, index i
    double mu[WP_MAX_CHANNELS], sig[WP_MAX_CHANNELS], bg_sig[WP_MAX_CHANNELS];
    double inter[WP_MAX_CHANNELS][WP_MAX_CHANNELS];
    double flat[WP_MAX_CHANNELS][WP_MAX_CHANNELS];
};

// ChannelModel::adjusted_mu (channel-model.h:35-55). counts[] holds C values;
// C is a compile-time bound so the count arrays stay in registers.
template <int C>
__device__ inline double dev_adjusted_mu(const DevModel& m, int channel, const unsigned int* counts) {
#pragma clang fp contract(off)
    double amu = m.mu[channel];
    unsigned int own = 0;
#pragma unroll
    for (int c = 0; c < C; c++)
        if (c == channel) own = counts[c];
    amu *= (double)own;
    // A factor of exactly 1.0 (no interaction, the usual case) multiplies by
    // pow(1.0, n) == 1.0: skipping it leaves amu bit for bit unchanged.
#pragma unroll
    for (int c = 0; c < C; c++) {
        const double fl = m.flat[channel][c], in = m.inter[channel][c];
        if (c == channel) {
            if (counts[c] > 1) {
                if (fl != 1.0) amu *= fl;
                if (in != 1.0) amu *= pow(in, (double)(counts[c] - 1));
            }
        } else if (counts[c] > 0) {
            if (fl != 1.0) amu *= fl;
            if (in != 1.0) amu *= pow(in, (double)counts[c]);
        }
    }
    return amu;
}

// ChannelModel::sigma (channel-model.cc:47-49).
__device__ inline double dev_sigma(const DevModel& m, int channel, double amu) {
#pragma clang fp contract(off)
    return sqrt(m.bg_sig[channel] * m.bg_sig[channel] + amu * m.sig[channel] * m.sig[channel]);
}

// ChannelModel::pdf_helper (channel-model.h:62-70).
template <int C>
__device__ inline double dev_pdf(const DevModel& m, int channel, double observed, const unsigned int* counts) {
#pragma clang fp contract(off)
    const double pi = 3.141592653589793238;
    double amu = dev_adjusted_mu<C>(m, channel, counts);
    double offset = observed - amu;
    double s = dev_sigma(m, channel, amu);
    return (1.0 / (s * sqrt(2.0 * pi))) * exp(-offset * offset / (2.0 * s * s));
}

// PeptideEmission's own range per channel at one timestep
// (peptide-emission.cc:49-164). x[] holds the C intensities of that timestep.
// FMA contraction is off so every comparison sees the same doubles as the
// reference's x86-64 build.
template <int C>
__device__ inline void dev_emission_range(const DevModel& m, const double* x, unsigned int* lo,
                                          unsigned int* hi) {
#pragma clang fp contract(off)
    const unsigned int M = (unsigned int)m.M;
    if (m.uncut) {
#pragma unroll
        for (int c = 0; c < C; c++) {
            lo[c] = 0;
            hi[c] = M + 1;
        }
        return;
    }
    unsigned int counts_lo[C];
    unsigned int counts_hi[C];
#pragma unroll
    for (int c = 0; c < C; c++) {
        counts_lo[c] = 0;
        counts_hi[c] = 0;
    }
    bool did_change = true;
    int guard = 0;
    while (did_change && guard < 4096) {
        did_change = false;
        guard++;
#pragma unroll
        for (int c = 0; c < C; c++) {
            unsigned int old = counts_hi[c];
            counts_hi[c] = 1;
            double mu1 = dev_adjusted_mu<C>(m, c, counts_hi);
            counts_hi[c] = old;
            double mu = dev_adjusted_mu<C>(m, c, counts_hi);
            double s = m.cutoff * dev_sigma(m, c, mu);
            long rounded = lround((x[c] + s) / mu1);
            unsigned int new_cts = (unsigned int)rounded;
            if (new_cts > M + 1) new_cts = M + 1;
            if (new_cts != counts_hi[c]) {
                did_change = true;
                counts_hi[c] = new_cts;
            }
        }
    }
#pragma unroll
    for (int c = 0; c < C; c++) {
        unsigned int cmin = M + 1;
        for (unsigned int d = 0; d < M + 1; d++) {
            counts_lo[c] = d;
            double mu = dev_adjusted_mu<C>(m, c, counts_lo);
            double s = m.cutoff * dev_sigma(m, c, mu);
            if (mu + s > x[c]) {
                cmin = d;
                break;
            }
        }
        lo[c] = cmin;
        counts_lo[c] = 0;
        unsigned int cmax = M + 1;
        unsigned int old_hi = counts_hi[c];
        for (unsigned int d = cmin; d < M + 1; d++) {
            counts_hi[c] = d;
            double mu = dev_adjusted_mu<C>(m, c, counts_hi);
            double s = m.cutoff * dev_sigma(m, c, mu);
            if (mu - s > x[c]) {
                cmax = d;
                break;
            }
        }
        hi[c] = cmax;
        counts_hi[c] = old_hi;
    }
}

}  // namespace wp
