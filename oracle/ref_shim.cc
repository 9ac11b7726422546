// ORACLE — test infrastructure only. A thin extern "C" face over the
// reference's own C++ classes, compiled in place from the reference sources by
// oracle/ref.mk into oracle/_ref/libwpref.so. It constructs the reference's
// objects exactly as its drivers do (run-classify-hmm.cc:48-88,
// hmm-classifier.h:45-78, peptide-hmm.test.cc:152-189) and returns their
// results, so the tests can pin the C restatement and the HIP path against the
// reference itself.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <string>
#include <vector>

#include <map>

#include "classifiers/hmm-classifier.h"
#include "classifiers/hybrid-classifier.h"
#include "classifiers/nn-classifier.h"
#include "common/dye-track.h"
#include "common/dye-seq.h"
#include "common/radiometry.h"
#include "common/scored-classification.h"
#include "common/sourced-data.h"
#include "fitters/hmm-fitter.h"
#include "hmm/hmm/peptide-hmm.h"
#include "hmm/precomputations/dye-seq-precomputations.h"
#include "hmm/precomputations/radiometry-precomputations.h"
#include "hmm/precomputations/universal-precomputations.h"
#include "parameterization/fit/sequencing-model-fitter.h"
#include "parameterization/settings/fit-settings.h"
#include "util/levmar-helper.h"
#include "levmar.h"
#include "parameterization/model/sequencing-model.h"
#include "parameterization/settings/sequencing-settings.h"
#include "simulation/generate-dye-track.h"
#include "io/dye-tracks-io.h"
#include "io/radiometries-io.h"
#include "io/scored-classifications-io.h"
#include "whatprot_amd.h"

using namespace whatprot;

namespace {

void fill_model(const wp_model* m, unsigned int C, SequencingModel* sm) {
    sm->p_edman_failure = m->p_edman_failure;
    sm->p_detach.base = m->p_detach;
    sm->p_detach.initial = m->p_initial_detach;
    sm->p_detach.initial_decay = m->p_initial_detach_decay;
    sm->p_initial_block = m->p_initial_block;
    sm->p_cyclic_block = m->p_cyclic_block;
    for (unsigned int c = 0; c < C; c++) {
        ChannelModel* cm = new ChannelModel(c, C);
        cm->p_bleach = m->channels[c].p_bleach;
        cm->p_dud = m->channels[c].p_dud;
        cm->bg_sig = m->channels[c].bg_sig;
        cm->mu = m->channels[c].mu;
        cm->sig = m->channels[c].sig;
        for (unsigned int o = 0; o < C; o++) {
            cm->interactions[o] = m->channels[c].interactions[o];
            cm->flat_interactions[o] = m->channels[c].flat_interactions[o];
        }
        sm->channel_models.push_back(cm);
    }
}

// SequencingModelFitter fields into the flat layout of WP_FIT_STATS_LEN.
void add_stats(const SequencingModelFitter& f, unsigned int T, unsigned int C, double* st) {
    st[WP_FIT_P_EDMAN_FAILURE] += f.p_edman_failure_fit.numerator;
    st[WP_FIT_P_EDMAN_FAILURE + 1] += f.p_edman_failure_fit.denominator;
    st[WP_FIT_P_INITIAL_BLOCK] += f.p_initial_block_fit.numerator;
    st[WP_FIT_P_INITIAL_BLOCK + 1] += f.p_initial_block_fit.denominator;
    st[WP_FIT_P_CYCLIC_BLOCK] += f.p_cyclic_block_fit.numerator;
    st[WP_FIT_P_CYCLIC_BLOCK + 1] += f.p_cyclic_block_fit.denominator;
    for (unsigned int t = 0; t + 1 < T; t++) {
        st[WP_FIT_DETACH_X + t] += f.p_detach_fit.xvec[t];
        st[WP_FIT_DETACH_X + (T - 1) + t] += f.p_detach_fit.nvec[t];
    }
    for (unsigned int c = 0; c < C; c++) {
        st[WP_FIT_CHANNEL(T, c) + 0] += f.channel_fits[c]->p_bleach_fit.numerator;
        st[WP_FIT_CHANNEL(T, c) + 1] += f.channel_fits[c]->p_bleach_fit.denominator;
        st[WP_FIT_CHANNEL(T, c) + 2] += f.channel_fits[c]->p_dud_fit.numerator;
        st[WP_FIT_CHANNEL(T, c) + 3] += f.channel_fits[c]->p_dud_fit.denominator;
    }
}

void read_model(const SequencingModel& sm, unsigned int C, wp_model* m) {
    m->num_channels = (int)C;
    m->p_edman_failure = sm.p_edman_failure;
    m->p_detach = sm.p_detach.base;
    m->p_initial_detach = sm.p_detach.initial;
    m->p_initial_detach_decay = sm.p_detach.initial_decay;
    m->p_initial_block = sm.p_initial_block;
    m->p_cyclic_block = sm.p_cyclic_block;
    for (unsigned int c = 0; c < C; c++) {
        const ChannelModel& cm = *sm.channel_models[c];
        m->channels[c].p_bleach = cm.p_bleach;
        m->channels[c].p_dud = cm.p_dud;
        m->channels[c].bg_sig = cm.bg_sig;
        m->channels[c].mu = cm.mu;
        m->channels[c].sig = cm.sig;
        for (unsigned int o = 0; o < C; o++) {
            m->channels[c].interactions[o] = cm.interactions[o];
            m->channels[c].flat_interactions[o] = cm.flat_interactions[o];
        }
    }
}

// holds[0..5]: edman, detach, initial detach, initial detach decay, initial
// block, cyclic block; holds[6 + 2c], holds[7 + 2c]: bleach, dud of channel c.
void fill_fit_settings(const int* holds, unsigned int C, FitSettings* fs) {
    fs->hold_p_edman_failure = holds && holds[0];
    fs->hold_p_detach = holds && holds[1];
    fs->hold_p_initial_detach = holds && holds[2];
    fs->hold_p_initial_detach_decay = holds && holds[3];
    fs->hold_p_initial_block = holds && holds[4];
    fs->hold_p_cyclic_block = holds && holds[5];
    for (unsigned int c = 0; c < C; c++) {
        fs->channel_fit_settings[c]->hold_p_bleach = holds && holds[6 + 2 * c];
        fs->channel_fit_settings[c]->hold_p_dud = holds && holds[7 + 2 * c];
    }
}

std::vector<SourcedData<DyeTrack, SourceCountHitsList<int>>> make_tracks(
        unsigned int T, unsigned int C, int n, const short* counts, const int* off, const int* ids,
        const int* cnts, const int* hits) {
    std::vector<SourcedData<DyeTrack, SourceCountHitsList<int>>> v;
    v.reserve(n);
    for (int i = 0; i < n; i++) {
        DyeTrack dt(T, C, const_cast<short*>(counts + (size_t)i * T * C));
        int ns = off[i + 1] - off[i];
        SourceCountHits<int>** src = new SourceCountHits<int>*[ns];
        for (int j = 0; j < ns; j++)
            src[j] = new SourceCountHits<int>(ids[off[i] + j], cnts[off[i] + j], hits[off[i] + j]);
        v.push_back(SourcedData<DyeTrack, SourceCountHitsList<int>>(std::move(dt),
                                                                    SourceCountHitsList<int>(ns, src)));
    }
    return v;
}

std::vector<Radiometry> make_rads(unsigned int T, unsigned int C, int N, const double* rads) {
    std::vector<Radiometry> out;
    out.reserve(N);
    for (int r = 0; r < N; r++) {
        out.push_back(Radiometry(T, C));
        for (unsigned int i = 0; i < T * C; i++) out.back().intensities[i] = rads[(size_t)r * T * C + i];
    }
    return out;
}

}  // namespace

extern "C" {

// generate_dye_track (simulation/generate-dye-track.cc:29-110) run n times
// on one std::default_random_engine seeded with seed (the reference seeds it
// from the clock, util/time.cc:28-30): out[n][T][C].
void wpref_generate_dye_tracks(const wp_model* m, unsigned int T, unsigned int C, const char* dye_string, int n,
                               unsigned int seed, short* out) {
    SequencingModel sm;
    fill_model(m, C, &sm);
    DyeSeq ds(C, std::string(dye_string));
    std::default_random_engine gen(seed);
    for (int k = 0; k < n; k++) {
        DyeTrack dt(T, C);
        generate_dye_track(sm, ds, T, C, &gen, &dt);
        for (unsigned int t = 0; t < T; t++)
            for (unsigned int c = 0; c < C; c++) out[((size_t)k * T + t) * C + c] = dt(t, c);
    }
}

// write_radiometries (io/radiometries-io.cc:58-81) of values already divided by mu.
void wpref_write_radiometries(const char* path, const wp_model* m, unsigned int T, unsigned int C, int n,
                              const double* values) {
    SequencingModel sm;
    fill_model(m, C, &sm);
    std::vector<SourcedData<Radiometry, SourceCount<int>>> rads;
    for (int i = 0; i < n; i++) {
        Radiometry r(T, C);
        for (unsigned int k = 0; k < T * C; k++) r.intensities[k] = values[(size_t)i * T * C + k];
        rads.push_back(SourcedData<Radiometry, SourceCount<int>>(r, SourceCount<int>(0, 1)));
    }
    write_radiometries(std::string(path), sm, T, C, rads);
}

// read_radiometries (io/radiometries-io.cc:33-56): returns N; fills out[cap].
int wpref_read_radiometries(const char* path, const wp_model* m, double* out, int cap) {
    SequencingModel sm;
    fill_model(m, (unsigned int)m->num_channels, &sm);
    unsigned int T, C, N;
    std::vector<Radiometry> rads;
    read_radiometries(std::string(path), sm, &T, &C, &N, &rads);
    for (unsigned int i = 0; i < N; i++)
        for (unsigned int k = 0; k < T * C && (int)(i * T * C + k) < cap; k++) out[(size_t)i * T * C + k] = rads[i].intensities[k];
    return (int)N;
}

// write_dye_tracks (io/dye-tracks-io.cc:66-103).
void wpref_write_dye_tracks(const char* path, unsigned int T, unsigned int C, int n, const short* counts,
                            const int* off, const int* ids, const int* cnts, const int* hits) {
    std::vector<SourcedData<DyeTrack, SourceCountHitsList<int>>> tracks;
    for (int i = 0; i < n; i++) {
        DyeTrack dt(T, C);
        for (unsigned int k = 0; k < T * C; k++) dt.counts[k] = counts[(size_t)i * T * C + k];
        const int ns = off[i + 1] - off[i];
        SourceCountHits<int>** src = new SourceCountHits<int>*[ns];
        for (int j = 0; j < ns; j++) src[j] = new SourceCountHits<int>(ids[off[i] + j], cnts[off[i] + j], hits[off[i] + j]);
        tracks.push_back(SourcedData<DyeTrack, SourceCountHitsList<int>>(dt, SourceCountHitsList<int>(ns, src)));
    }
    write_dye_tracks(std::string(path), T, C, tracks);
}

// read_dye_tracks (io/dye-tracks-io.cc:32-64): returns n and the number of
// sources; fills counts[n*T*C], off[n+1] and the source triples (capacities
// cap_t tracks, cap_s sources).
int wpref_read_dye_tracks(const char* path, short* counts, int* off, int* ids, int* cnts, int* hits, int cap_t,
                          int cap_s, int* out_sources) {
    unsigned int T, C;
    std::vector<SourcedData<DyeTrack, SourceCountHitsList<int>>> tracks;
    read_dye_tracks(std::string(path), &T, &C, &tracks);
    int s = 0;
    if (cap_t >= 0) off[0] = 0;
    for (size_t i = 0; i < tracks.size(); i++) {
        if ((int)i < cap_t)
            for (unsigned int k = 0; k < T * C; k++) counts[i * T * C + k] = tracks[i].value.counts[k];
        for (int j = 0; j < tracks[i].source.num_sources; j++, s++) {
            if (s < cap_s) {
                ids[s] = tracks[i].source.sources[j]->source;
                cnts[s] = tracks[i].source.sources[j]->count;
                hits[s] = tracks[i].source.sources[j]->hits;
            }
        }
        if ((int)i < cap_t) off[i + 1] = s;
    }
    *out_sources = s;
    return (int)tracks.size();
}

// write_scored_classifications_raw (io/scored-classifications-io.cc:54-68).
void wpref_write_scored(const char* path, int n, const int* ids, const double* scores) {
    int* i2 = new int[n > 0 ? n : 1];
    double* s2 = new double[n > 0 ? n : 1];
    for (int i = 0; i < n; i++) {
        i2[i] = ids[i];
        s2[i] = scores[i];
    }
    write_scored_classifications_raw(std::string(path), n, i2, s2);  // frees i2, s2
}

double wpref_pdf(const wp_model* m, int channel, double observed, const unsigned int* counts) {
    SequencingModel sm;
    fill_model(m, (unsigned int)m->num_channels, &sm);
    return sm.channel_models[channel]->pdf(observed, counts);
}

double wpref_pair_probability(const wp_model* m, unsigned int T, unsigned int C,
                              const char* dye_string, const double* rad,
                              unsigned int max_num_dyes, double dist_cutoff) {
    SequencingModel sm;
    fill_model(m, C, &sm);
    SequencingSettings ss;
    ss.dist_cutoff = dist_cutoff;
    UniversalPrecomputations up(sm, T, C);
    up.set_max_num_dyes(max_num_dyes);
    DyeSeq ds(C, std::string(dye_string));
    DyeSeqPrecomputations dsp(ds, sm, T, C);
    Radiometry r(T, C);
    for (unsigned int i = 0; i < T * C; i++) r.intensities[i] = rad[i];
    RadiometryPrecomputations rp(r, sm, ss, max_num_dyes);
    PeptideHMM hmm(T, C, dsp, rp, up);
    return hmm.probability();
}

// HMMClassifier::classify(vector<Radiometry>) with OpenMP over reads
// (hmm-classifier.cc:84-93). Returns raw best score and total.
void wpref_classify(const wp_model* m, unsigned int T, unsigned int C,
                    int num_dye_seqs, const char* const* dye_strings,
                    const int* counts, const int* ids, int num_reads,
                    const double* rads, double dist_cutoff, int* out_id,
                    double* out_score, double* out_total) {
    SequencingModel sm;
    fill_model(m, C, &sm);
    SequencingSettings ss;
    ss.dist_cutoff = dist_cutoff;
    std::vector<SourcedData<DyeSeq, SourceCount<int>>> dye_seqs;
    dye_seqs.reserve(num_dye_seqs);
    for (int i = 0; i < num_dye_seqs; i++) {
        dye_seqs.push_back(SourcedData<DyeSeq, SourceCount<int>>(
                DyeSeq(C, std::string(dye_strings[i])),
                SourceCount<int>(ids[i], counts[i])));
    }
    std::vector<Radiometry> radiometries;
    radiometries.reserve(num_reads);
    for (int r = 0; r < num_reads; r++) {
        radiometries.push_back(Radiometry(T, C));
        for (unsigned int i = 0; i < T * C; i++)
            radiometries.back().intensities[i] = rads[(size_t)r * T * C + i];
    }
    HMMClassifier classifier(T, C, sm, ss, dye_seqs);
    std::vector<ScoredClassification> res = classifier.classify(radiometries);
    for (int r = 0; r < num_reads; r++) {
        out_id[r] = res[r].id;
        out_score[r] = res[r].score;
        out_total[r] = res[r].total;
    }
}

// GenericHMM::improve_fit of one pair with uncut ranges, as HMMFitter::fit
// runs it (hmm-fitter.h:63-78); adds the per-read fitter into stats.
double wpref_improve_fit(const wp_model* m, unsigned int T, unsigned int C,
                         const char* dye_string, const double* rad,
                         unsigned int max_num_dyes, double* stats) {
    SequencingModel sm;
    fill_model(m, C, &sm);
    SequencingSettings ss;
    ss.dist_cutoff = DBL_MAX;
    FitSettings fs(C);
    UniversalPrecomputations up(sm, T, C);
    up.set_max_num_dyes(max_num_dyes);
    DyeSeq ds(C, std::string(dye_string));
    DyeSeqPrecomputations dsp(ds, sm, T, C);
    Radiometry r(T, C);
    for (unsigned int i = 0; i < T * C; i++) r.intensities[i] = rad[i];
    RadiometryPrecomputations rp(r, sm, ss, max_num_dyes);
    PeptideHMM hmm(T, C, dsp, rp, up);
    SequencingModelFitter f(T, C, sm, fs);
    double P = hmm.improve_fit(&f);
    add_stats(f, T, C, stats);
    return P;
}

// least_squares_fit_of_offset_exponential (util/levmar-helper.cc:44-85).
void wpref_offset_exponential_fit(int n, const double* y, const double* w, const int* hold,
                                  double* p3) {
    std::vector<double> yv(y, y + n), wv(w, w + n), p(p3, p3 + 3);
    std::vector<bool> h(3);
    for (int i = 0; i < 3; i++) h[i] = hold[i] != 0;
    least_squares_fit_of_offset_exponential(yv, wv, h, &p);
    for (int i = 0; i < 3; i++) p3[i] = p[i];
}

// The M-step alone: SequencingModelFitter::get from flat statistics plus the
// dud correction and update_with_holds of HMMFitter::fit
// (hmm-fitter.h:80-104, hmm-fitter.cc:63-97). Returns the step size.
double wpref_mstep(const wp_model* m, unsigned int T, unsigned int C, const char* dye_string,
                   int num_reads, const double* stats, const int* holds, wp_model* out) {
    SequencingModel sm;
    fill_model(m, C, &sm);
    FitSettings fs(C);
    fill_fit_settings(holds, C, &fs);
    SequencingModelFitter f(T, C, sm, fs);
    f.p_edman_failure_fit.numerator = stats[WP_FIT_P_EDMAN_FAILURE];
    f.p_edman_failure_fit.denominator = stats[WP_FIT_P_EDMAN_FAILURE + 1];
    f.p_initial_block_fit.numerator = stats[WP_FIT_P_INITIAL_BLOCK];
    f.p_initial_block_fit.denominator = stats[WP_FIT_P_INITIAL_BLOCK + 1];
    f.p_cyclic_block_fit.numerator = stats[WP_FIT_P_CYCLIC_BLOCK];
    f.p_cyclic_block_fit.denominator = stats[WP_FIT_P_CYCLIC_BLOCK + 1];
    for (unsigned int t = 0; t + 1 < T; t++) {
        f.p_detach_fit.xvec[t] = stats[WP_FIT_DETACH_X + t];
        f.p_detach_fit.nvec[t] = stats[WP_FIT_DETACH_X + (T - 1) + t];
    }
    for (unsigned int c = 0; c < C; c++) {
        f.channel_fits[c]->p_bleach_fit.numerator = stats[WP_FIT_CHANNEL(T, c) + 0];
        f.channel_fits[c]->p_bleach_fit.denominator = stats[WP_FIT_CHANNEL(T, c) + 1];
        f.channel_fits[c]->p_dud_fit.numerator = stats[WP_FIT_CHANNEL(T, c) + 2];
        f.channel_fits[c]->p_dud_fit.denominator = stats[WP_FIT_CHANNEL(T, c) + 3];
    }
    DyeSeq ds(C, std::string(dye_string));
    double ratio_hidden = 1.0;
    for (unsigned int i = 0; i < ds.length; i++)
        if (ds[i] != -1) ratio_hidden *= sm.channel_models[ds[i]]->p_dud;
    double magic_ratio = 1.0 / (1.0 - ratio_hidden) - 1.0;
    double expected_hidden_count = magic_ratio * num_reads;
    for (unsigned int i = 0; i < ds.length; i++) {
        if (ds[i] != -1) {
            f.channel_fits[ds[i]]->p_dud_fit.numerator += expected_hidden_count;
            f.channel_fits[ds[i]]->p_dud_fit.denominator += expected_hidden_count;
        }
    }
    SequencingModel next = sm;
    SequencingSettings ss;
    ss.dist_cutoff = DBL_MAX;
    HMMFitter helper(T, C, 0.0, 0.0, sm, ss, fs, ds);
    helper.update_with_holds(f.get(), &next);
    read_model(next, C, out);
    return sm.distance(next);
}

// HMMFitter::fit (hmm-fitter.h:50-116) as run_fit calls it; returns log L
// under the fitted model.
double wpref_fit(const wp_model* m, unsigned int T, unsigned int C, const char* dye_string,
                 int num_reads, const double* rads, double stopping_threshold,
                 double max_runtime, const int* holds, wp_model* out, double* step_size) {
    SequencingModel sm;
    fill_model(m, C, &sm);
    SequencingSettings ss;
    ss.dist_cutoff = DBL_MAX;
    FitSettings fs(C);
    fill_fit_settings(holds, C, &fs);
    DyeSeq ds(C, std::string(dye_string));
    std::vector<Radiometry> radiometries;
    radiometries.reserve(num_reads);
    for (int r = 0; r < num_reads; r++) {
        radiometries.push_back(Radiometry(T, C));
        for (unsigned int i = 0; i < T * C; i++)
            radiometries.back().intensities[i] = rads[(size_t)r * T * C + i];
    }
    HMMFitter fitter(T, C, stopping_threshold, max_runtime, sm, ss, fs, ds);
    SequencingModel fitted;
    double log_l = fitter.fit(radiometries, &fitted, step_size);
    read_model(fitted, C, out);
    return log_l;
}

// dlevmar_bc_der exactly as least_squares_fit_of_offset_exponential calls
// it (util/levmar-helper.cc:44-85) but with a chosen iteration cap and the
// info vector returned, for step-by-step comparison.
void wpref_levmar_bc(int n, const double* y, const double* w, const int* hold, double* p3, int itmax,
                     double* info) {
    std::vector<double> lb(3, 0), ub(3, 1), sw(n), wy(n);
    for (int i = 0; i < 3; i++)
        if (hold[i]) lb[i] = ub[i] = p3[i];
    for (int i = 0; i < n; i++) sw[i] = std::sqrt(w[i]);
    for (int i = 0; i < n; i++) wy[i] = sw[i] * y[i];
    dlevmar_bc_der(&offset_exponential, &jacobian_of_offset_exponential, p3, wy.data(), 3, n, lb.data(),
                   ub.data(), NULL, itmax, NULL, info, NULL, NULL, sw.data());
}

// One E-step of HMMFitter::fit exactly as its loop runs it (hmm-fitter.h:57-79):
// serial over reads, each read's PeptideHMM::improve_fit into a fresh fitter
// added to the total. Returns the sum of ln P.
double wpref_fit_estep(const wp_model* m, unsigned int T, unsigned int C, const char* dye_string,
                       int num_reads, const double* rads, double* stats) {
    SequencingModel sm;
    fill_model(m, C, &sm);
    SequencingSettings ss;
    ss.dist_cutoff = DBL_MAX;
    FitSettings fs(C);
    DyeSeq ds(C, std::string(dye_string));
    unsigned int max_num_dyes = 0;
    for (unsigned int c = 0; c < C; c++) {
        unsigned int n = 0;
        for (unsigned int i = 0; i < ds.length; i++)
            if (ds[i] == (int)c) n++;
        max_num_dyes = std::max(max_num_dyes, n);
    }
    SequencingModelFitter total(T, C, sm, fs);
    DyeSeqPrecomputations dsp(ds, sm, T, C);
    UniversalPrecomputations up(sm, T, C);
    up.set_max_num_dyes(max_num_dyes);
    double log_p = 0.0;
    for (int r = 0; r < num_reads; r++) {
        Radiometry rad(T, C);
        for (unsigned int i = 0; i < T * C; i++) rad.intensities[i] = rads[(size_t)r * T * C + i];
        RadiometryPrecomputations rp(rad, sm, ss, max_num_dyes);
        PeptideHMM hmm(T, C, dsp, rp, up);
        SequencingModelFitter one(T, C, sm, fs);
        log_p += std::log(hmm.improve_fit(&one));
        total += one;
    }
    add_stats(total, T, C, stats);
    return log_p;
}

// KDTree::search of NNClassifier's tree for every read (kd-tree.h:40-46):
// neighbours farthest first, identified by the index of their (distinct)
// dye-track in the input. out arrays are [N][k].
void wpref_nn_search(const wp_model* m, unsigned int T, unsigned int C, int k, double sig, int n,
                     const short* counts, const int* off, const int* ids, const int* cnts, const int* hits,
                     int N, const double* rads, int* out_count, int* out_track, double* out_dist) {
    SequencingModel sm;
    fill_model(m, C, &sm);
    std::map<std::vector<short>, int> index;
    for (int i = 0; i < n; i++)
        index[std::vector<short>(counts + (size_t)i * T * C, counts + (size_t)(i + 1) * T * C)] = i;
    auto tracks = make_tracks(T, C, n, counts, off, ids, cnts, hits);
    NNClassifier nn(T, C, sm, k, sig, &tracks);
    auto radv = make_rads(T, C, N, rads);
    for (int r = 0; r < N; r++) {
        KDTQuery q(radv[r]);
        std::vector<KDTEntry*> kn;
        std::vector<double> ds;
        nn.kd_tree->search(q, &kn, &ds);
        out_count[r] = (int)kn.size();
        for (size_t i = 0; i < kn.size(); i++) {
            out_track[(size_t)r * k + i] = index[kn[i]->dye_track.value.counts];
            out_dist[(size_t)r * k + i] = ds[i];
        }
    }
}

// NNClassifier::classify(vector<Radiometry>) (nn-classifier.cc:206-215).
void wpref_nn_classify(const wp_model* m, unsigned int T, unsigned int C, int k, double sig, int n,
                       const short* counts, const int* off, const int* ids, const int* cnts, const int* hits,
                       int N, const double* rads, int* out_id, double* out_score, double* out_total) {
    SequencingModel sm;
    fill_model(m, C, &sm);
    auto tracks = make_tracks(T, C, n, counts, off, ids, cnts, hits);
    NNClassifier nn(T, C, sm, k, sig, &tracks);
    auto res = nn.classify(make_rads(T, C, N, rads));
    for (int r = 0; r < N; r++) {
        out_id[r] = res[r].id;
        out_score[r] = res[r].score;
        out_total[r] = res[r].total;
    }
}

// NNClassifier::classify(radiometry, h) (nn-classifier.cc:179-204).
void wpref_nn_classify_top(const wp_model* m, unsigned int T, unsigned int C, int k, double sig, int n,
                           const short* counts, const int* off, const int* ids, const int* cnts,
                           const int* hits, int N, const double* rads, int h, int* out_count, int* out_ids,
                           double* out_scores, double* out_total) {
    SequencingModel sm;
    fill_model(m, C, &sm);
    auto tracks = make_tracks(T, C, n, counts, off, ids, cnts, hits);
    NNClassifier nn(T, C, sm, k, sig, &tracks);
    auto radv = make_rads(T, C, N, rads);
    for (int r = 0; r < N; r++) {
        auto top = nn.classify(radv[r], (unsigned int)h);
        out_count[r] = (int)top.size();
        out_total[r] = top.empty() ? 0.0 : top[0].total;
        for (size_t i = 0; i < top.size(); i++) {
            out_ids[(size_t)r * h + i] = top[i].id;
            out_scores[(size_t)r * h + i] = top[i].score;
        }
    }
}

// HybridClassifier::classify(vector<Radiometry>) (hybrid-classifier.cc:84-95).
void wpref_hybrid_classify(const wp_model* m, unsigned int T, unsigned int C, double dist_cutoff, int k,
                           double sig, int n, const short* counts, const int* off, const int* ids,
                           const int* cnts, const int* hits, int h, int num_dye_seqs,
                           const char* const* dye_strings, const int* ds_counts, const int* ds_ids, int N,
                           const double* rads, int* out_id, double* out_score, double* out_total) {
    SequencingModel sm;
    fill_model(m, C, &sm);
    SequencingSettings ss;
    ss.dist_cutoff = dist_cutoff;
    auto tracks = make_tracks(T, C, n, counts, off, ids, cnts, hits);
    std::vector<SourcedData<DyeSeq, SourceCount<int>>> dye_seqs;
    for (int i = 0; i < num_dye_seqs; i++)
        dye_seqs.push_back(SourcedData<DyeSeq, SourceCount<int>>(DyeSeq(C, std::string(dye_strings[i])),
                                                                SourceCount<int>(ds_ids[i], ds_counts[i])));
    HybridClassifier hc(T, C, sm, ss, k, sig, &tracks, h, dye_seqs);
    auto res = hc.classify(make_rads(T, C, N, rads));
    for (int r = 0; r < N; r++) {
        out_id[r] = res[r].id;
        out_score[r] = res[r].score;
        out_total[r] = res[r].total;
    }
}

// The permuted point order of KDTree's construction (kd-tree.h:33-37): for
// each position, the index of the point given at that position in the input.
// Points are rows of X (n x d) with one hit each.
void wpref_kdtree_order(int k, int d, int n, const double* X, int* out_order) {
    struct P {
        std::vector<double> v;
        int hits;
        int orig;
        double operator[](int i) const { return v[i]; }
    };
    std::vector<P> pts;
    for (int i = 0; i < n; i++) pts.push_back(P{std::vector<double>(X + (size_t)i * d, X + (size_t)(i + 1) * d), 1, i});
    KDTree<P, std::vector<double>> tree(k, d, std::move(pts));
    for (int i = 0; i < n; i++) out_order[i] = tree.values[i].orig;
}

}  // extern "C"
