// ORACLE — test infrastructure only. A thin extern "C" face over the
// reference's own C++ classes, compiled in place from the reference sources by
// oracle/ref.mk into oracle/_ref/libwpref.so. It constructs the reference's
// objects exactly as its drivers do (run-classify-hmm.cc:48-88,
// hmm-classifier.h:45-78, peptide-hmm.test.cc:152-189) and returns their
// results, so the tests can pin the C restatement and the HIP path against the
// reference itself.
#include <cfloat>
#include <cmath>
#include <string>
#include <vector>

#include "classifiers/hmm-classifier.h"
#include "common/dye-seq.h"
#include "common/radiometry.h"
#include "common/scored-classification.h"
#include "common/sourced-data.h"
#include "hmm/hmm/peptide-hmm.h"
#include "hmm/precomputations/dye-seq-precomputations.h"
#include "hmm/precomputations/radiometry-precomputations.h"
#include "hmm/precomputations/universal-precomputations.h"
#include "parameterization/model/sequencing-model.h"
#include "parameterization/settings/sequencing-settings.h"
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

}  // namespace

extern "C" {

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

}  // extern "C"
