// Seeded synthetic-input generator, host side: random dye-seqs, and dye tracks
// and radiometries drawn from the sequencing model's forward process.
// Restates the generative model of whatprot's simulator
// (simulation/generate-dye-track.cc:29-110, generate-radiometry.cc:28-62,
// generate-radiometries.cc:30-64) with a portable RNG (xoshiro256** seeded by
// splitmix64) instead of the reference's clock-seeded
// std::default_random_engine, so runs are reproducible across machines.
// Streams therefore differ from the reference; the distribution is the same.
//
// Every simulated peptide (dye tracks) or read attempt (radiometries) draws
// from its own stream g, seeded from seed + (g + 1) * 0xD1B54A32D192ED03, and
// the detach probability of each timestep comes from one table, so the GPU
// simulator (simulate_gpu.hip) produces the same tracks bit for bit.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <vector>

#include "sim_rng.h"
#include "wp_internal.h"

namespace wp {

namespace {



}  // namespace

// p_detach of each timestep: DecayingRateModel (decaying-rate-model.cc:47-50).
void detach_table(const wp_model& m, int T, double* pd) {
    for (int t = 0; t < T; t++) pd[t] = m.p_detach + m.p_initial_detach * std::exp(-(double)t * m.p_initial_detach_decay);
}

// Dye-seq generator: length ~ U{len_min..len_max}; per-channel dye counts
// k_c ~ U{k_min..k_max} capped so the total fits the length and redrawn while
// the total is zero; dye positions uniform without replacement.
void generate_dye_seqs(int C, int n, int len_min, int len_max, int k_min, int k_max,
                       uint64_t seed, char* out, int stride) {
    Rng rng(seed);
    for (int i = 0; i < n; i++) {
        int len = rng.range(len_min, len_max);
        std::vector<int> k(C);
        int total;
        do {
            total = 0;
            for (int c = 0; c < C; c++) {
                k[c] = rng.range(k_min, k_max);
                if (total + k[c] > len) k[c] = len - total;
                total += k[c];
            }
        } while (total == 0);
        std::vector<char> s(len, '.');
        std::vector<int> free_pos(len);
        for (int p = 0; p < len; p++) free_pos[p] = p;
        int nfree = len;
        for (int c = 0; c < C; c++) {
            for (int j = 0; j < k[c]; j++) {
                int pick = rng.range(0, nfree - 1);
                s[free_pos[pick]] = (char)('0' + c);
                free_pos[pick] = free_pos[--nfree];
            }
        }
        char* dst = out + (size_t)i * stride;
        std::memset(dst, 0, stride);
        std::memcpy(dst, s.data(), std::min(len, stride - 1));
    }
}

// One dye track following generate-dye-track.cc:29-110 (duds; initial block;
// per cycle: record, cyclic block, detach, Edman unless blocked, bleach).
static void dye_track(const wp_model& m, const double* pdt, const short* seq0, int len, int T, int C, Rng& rng,
                      unsigned int* track) {
    std::vector<short> ds(seq0, seq0 + len);
    for (int i = 0; i < len; i++)
        if (ds[i] != -1 && rng.bernoulli(m.channels[ds[i]].p_dud)) ds[i] = -1;
    int counts[WP_MAX_CHANNELS] = {0};
    for (int i = 0; i < len; i++)
        if (ds[i] != -1) counts[ds[i]]++;
    int e = 0;
    bool block = rng.bernoulli(m.p_initial_block);
    for (int t = 0; t < T; t++) {
        if (e < len) {
            for (int c = 0; c < C; c++) track[t * C + c] = (unsigned int)counts[c];
            if (rng.bernoulli(m.p_cyclic_block)) block = true;
            if (rng.bernoulli(pdt[t])) {
                e = len;
                continue;
            }
            if (!block) {
                if (!rng.bernoulli(m.p_edman_failure)) {
                    if (ds[e] != -1) counts[ds[e]]--;
                    e++;
                }
            }
            for (int i = e; i < len; i++) {
                if (ds[i] != -1 && rng.bernoulli(m.channels[ds[i]].p_bleach)) {
                    counts[ds[i]]--;
                    ds[i] = -1;
                }
            }
        } else {
            for (int c = 0; c < C; c++) track[t * C + c] = 0;
        }
    }
}

// generate-radiometries.cc:30-64 with the visibility filter of
// generate-radiometry.cc:41-53, repeated until exactly n visible reads exist.
// Values are rounded to float32 (the precision of the synthetic files).
void simulate_radiometries(const wp_model& m, int T, int C, int nds, const short* seqs, int cap,
                           const int* lens, const int* weights, int n, uint64_t seed,
                           double* out, int* out_idx, unsigned int* out_tracks) {
#pragma clang fp contract(off)
    std::vector<double> pdt(T);
    detach_table(m, T, pdt.data());
    std::vector<double> cum(nds);
    double acc = 0.0;
    for (int i = 0; i < nds; i++) {
        acc += weights ? (double)weights[i] : 1.0;
        cum[i] = acc;
    }
    std::vector<unsigned int> track((size_t)T * C);
    int made = 0;
    for (uint64_t attempt = 0; made < n; attempt++) {
        Rng rng = stream_rng(seed, attempt);
        double u = rng.uniform() * acc;
        int i = (int)(std::upper_bound(cum.begin(), cum.end(), u) - cum.begin());
        if (i >= nds) i = nds - 1;
        dye_track(m, pdt.data(), seqs + (size_t)i * cap, lens[i], T, C, rng, track.data());
        bool trivial = true;
        for (int c = 0; c < C; c++)
            if (track[c] != 0) trivial = false;
        if (trivial) continue;
        double* r = out + (size_t)made * T * C;
        for (int t = 0; t < T; t++) {
            for (int c = 0; c < C; c++) {
                unsigned int cnt[WP_MAX_CHANNELS];
                for (int k = 0; k < C; k++) cnt[k] = track[t * C + k];
                double mu = adjusted_mu(m, c, cnt);
                double sd = sigma(m, c, mu);
                r[t * C + c] = (double)(float)(mu + sd * rng.normal());
            }
        }
        if (out_idx) out_idx[made] = i;
        if (out_tracks) std::memcpy(out_tracks + (size_t)made * T * C, track.data(), sizeof(unsigned int) * T * C);
        made++;
    }
}

// generate_dye_tracks (simulation/generate-dye-tracks.cc:28-63): per_peptide
// tracks for every peptide (count) of every dye-seq, all-zero tracks dropped;
// then deduplicated as dedup-dye-tracks.cc does (one entry per distinct track
// with, per source dye-seq id, its count and the number of hits). Entries are
// ordered by their counts (lexicographically), sources by id.
int simulate_dye_tracks(const wp_model& m, int T, int C, int nds, const short* seqs, int cap, const int* lens,
                        const int* counts, const int* ids, int per_peptide, uint64_t seed,
                        std::vector<short>* track_counts, std::vector<int>* src_off, std::vector<int>* src_id,
                        std::vector<int>* src_count, std::vector<int>* src_hits) {
    std::vector<double> pdt(T);
    detach_table(m, T, pdt.data());
    std::map<std::vector<short>, std::map<int, std::pair<int, int>>> table;
    std::vector<unsigned int> track((size_t)T * C);
    std::vector<short> key((size_t)T * C);
    uint64_t g = 0;  // peptide index: dye-seq, then copy, then repeat
    for (int i = 0; i < nds; i++) {
        for (int copy = 0; copy < counts[i]; copy++) {
            for (int j = 0; j < per_peptide; j++, g++) {
                Rng rng = stream_rng(seed, g);
                dye_track(m, pdt.data(), seqs + (size_t)i * cap, lens[i], T, C, rng, track.data());
                bool trivial = true;
                for (int c = 0; c < C; c++)
                    if (track[c] != 0) trivial = false;
                if (trivial) continue;
                for (size_t q = 0; q < key.size(); q++) key[q] = (short)track[q];
                auto& src = table[key][ids[i]];
                src.first = counts[i];
                src.second++;
            }
        }
    }
    track_counts->clear();
    src_off->assign(1, 0);
    src_id->clear();
    src_count->clear();
    src_hits->clear();
    for (const auto& kv : table) {
        track_counts->insert(track_counts->end(), kv.first.begin(), kv.first.end());
        for (const auto& s : kv.second) {
            src_id->push_back(s.first);
            src_count->push_back(s.second.first);
            src_hits->push_back(s.second.second);
        }
        src_off->push_back((int)src_id->size());
    }
    return (int)table.size();
}

}  // namespace wp
