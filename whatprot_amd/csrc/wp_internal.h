// Internal declarations shared by the host-side translation units of
// libwhatprot_amd. Not part of the C ABI.
#pragma once

#include <cmath>
#include <cstdint>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/whatprot_amd.h"

namespace wp {

void set_error(const std::string& msg);

// ChannelModel::adjusted_mu (channel-model.h:35-55), host side.
inline double adjusted_mu(const wp_model& m, int channel, const unsigned int* counts) {
    const wp_channel_model& cm = m.channels[channel];
    double amu = cm.mu;
    amu *= (double)counts[channel];
    for (int c = 0; c < m.num_channels; c++) {
        if (c == channel) {
            if (counts[c] > 1) {
                amu *= cm.flat_interactions[c];
                amu *= std::pow(cm.interactions[c], (double)(counts[c] - 1));
            }
        } else if (counts[c] > 0) {
            amu *= cm.flat_interactions[c];
            amu *= std::pow(cm.interactions[c], (double)counts[c]);
        }
    }
    return amu;
}

// ChannelModel::sigma (channel-model.cc:47-49), host side.
inline double sigma(const wp_model& m, int channel, double amu) {
    const wp_channel_model& cm = m.channels[channel];
    return std::sqrt(cm.bg_sig * cm.bg_sig + amu * cm.sig * cm.sig);
}

// DyeSeq(C, string) (dye-seq.cc:23-40): returns the length after trimming
// trailing '.'; seq[i] = channel digit or -1.
int parse_dye_seq(const char* s, short* seq, int cap);

// Dye-seq index of each id of an HMM classifier's table; false (map
// unspecified) when two dye-seqs share an id.
bool dye_seq_index_of_ids(const wp_hmm_classifier* c, std::unordered_map<int, int>* out);

// True iff every parsed dye names a channel < C (else sets the error text).
bool dye_channels_ok(const short* seqs, int n, int cap, const int* lens, int C);

void generate_dye_seqs(int C, int n, int len_min, int len_max, int k_min, int k_max,
                       uint64_t seed, char* out, int stride);
void simulate_radiometries(const wp_model& m, int T, int C, int nds, const short* seqs, int cap,
                           const int* lens, const int* weights, int n, uint64_t seed,
                           double* out, int* out_idx, unsigned int* out_tracks);

// p_detach of each timestep t < T (decaying-rate-model.cc:47-50).
void detach_table(const wp_model& m, int T, double* pd);

struct DevModel;
// GPU simulator (simulate_gpu.hip): the host simulator's streams, on the device.
int gpu_simulate_dye_tracks(const wp_model& m, int T, int C, int nds, const short* seqs, int cap, const int* lens,
                            const int* counts, const int* ids, int per_peptide, uint64_t seed, int max_tracks,
                            int max_sources, int* out_num, int* out_num_sources, short* out_tracks, int* out_off,
                            int* out_id, int* out_count, int* out_hits);
int gpu_simulate_radiometries(const wp_model& m, const DevModel& dm, int T, int C, int nds, const short* seqs,
                              int cap, const int* lens, const int* weights, int n, uint64_t seed, double* out,
                              int* out_idx);

int simulate_dye_tracks(const wp_model& m, int T, int C, int nds, const short* seqs, int cap, const int* lens,
                        const int* counts, const int* ids, int per_peptide, uint64_t seed,
                        std::vector<short>* track_counts, std::vector<int>* src_off, std::vector<int>* src_id,
                        std::vector<int>* src_count, std::vector<int>* src_hits);

// least_squares_fit_of_offset_exponential (util/levmar-helper.cc:44-85):
// weighted fit of y_t ~ p[0] + p[1] exp(-p[2] t), t = 0..n-1, bounds [0, 1],
// hold[i] != 0 pins p[i]; p is updated in place. Own restatement of levmar
// 2.6's box-constrained method (levmar_bc.cpp).
void offset_exponential_fit(int n, const double* y, const double* w, const int* hold, double* p,
                            int itmax = 10);

}  // namespace wp
