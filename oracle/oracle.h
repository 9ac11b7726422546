/*
 * ORACLE — test infrastructure only.
 *
 * A plain-C FP64 restatement of whatprot's PeptideHMM scoring path, written
 * from the reference's semantics (file:line citations at each function). It is
 * the checker for the HIP implementation: only tests/, the smoke entry and the
 * CPU-baseline leg of the benchmark may load it. Nothing in the product links
 * or calls it.
 *
 * Parity pins (tests/test_oracle.py, tests/test_boxes.py, tests/test_fit.py):
 *   - whole-HMM goldens from hmm/hmm/peptide-hmm.test.cc:188,227,266,305,344;
 *   - ChannelModel::pdf goldens from channel-model.test.cc;
 *   - binomial table vs the closed form C(i,j)(1-q)^j q^(i-j);
 *   - brute-force enumeration of hidden paths on tiny inputs;
 *   - mass conservation (emission == 1 => P == 1);
 *   - the compiled reference (oracle/_ref) on random inputs, where built.
 */
#ifndef WHATPROT_ORACLE_H
#define WHATPROT_ORACLE_H

#include "../include/whatprot_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

#define WPO_MAXD (WP_MAX_CHANNELS + 1)

/* ChannelModel::adjusted_mu (channel-model.h:35-55). */
double wpo_adjusted_mu(const wp_model* m, int channel, const unsigned int* counts);
/* ChannelModel::sigma (channel-model.cc:47-49). */
double wpo_sigma(const wp_model* m, int channel, double amu);
/* ChannelModel::pdf_helper (channel-model.h:62-70). */
double wpo_pdf(const wp_model* m, int channel, double observed, const unsigned int* counts);
/* DecayingRateModel::operator[] (decaying-rate-model.cc:47-50). */
double wpo_p_detach(const wp_model* m, unsigned int i);
/* BinomialTransition::reserve (binomial-transition.cc:51-67): lower-triangular
 * table P(i,j) at index i*(i+1)/2 + j, i in [0, max_n]. */
void wpo_binomial_table(double q, unsigned int max_n, double* out);

/* PeptideEmission's own pruned range for timestep t (peptide-emission.cc:39-164):
 * per channel [lo[c], hi[c]); the Edman axis is [0, t+1). */
void wpo_emission_range(const wp_model* m, const double* rad, unsigned int C,
                        unsigned int t, unsigned int max_num_dyes,
                        double dist_cutoff, unsigned int* lo, unsigned int* hi);

/* DyeSeq(C, string) (dye-seq.cc:23-40): writes up to cap entries of seq
 * (-1 = no dye), returns the trimmed length. */
int wpo_parse_dye_seq(const char* s, short* seq, int cap);

/* Per-emission boxes after PeptideHMM's forward and backward pruning passes
 * (peptide-hmm.cc:62-83), computed with the generic per-step rules.
 * Returns 1 if every range is non-empty, 0 otherwise. On success lo/hi hold,
 * for each timestep t, the channel bounds of the emission's final range at
 * [t*C + c]. */
int wpo_emission_boxes(const wp_model* m, unsigned int T, unsigned int C,
                       const short* seq, int len, const double* rad,
                       unsigned int max_num_dyes, double dist_cutoff,
                       unsigned int* lo, unsigned int* hi);

/* PeptideHMM(...).probability() (peptide-hmm.cc:33-100, generic-hmm.h:41-55). */
double wpo_pair_probability(const wp_model* m, unsigned int T, unsigned int C,
                            const short* seq, int len, const double* rad,
                            unsigned int max_num_dyes, double dist_cutoff);

/* HMMClassifier::classify over all reads (hmm-classifier.cc:84-93 with
 * classify_helper hmm-classifier.h:45-78). dye seqs are packed: seqs[i*seq_cap ..]
 * with lengths lens[i]. cand_offsets may be NULL (all dye-seqs). Outputs
 * id, raw best score, total (unadjusted). OpenMP over reads. */
void wpo_classify(const wp_model* m, unsigned int T, unsigned int C,
                  int num_dye_seqs, const short* seqs, int seq_cap,
                  const int* lens, const int* counts, const int* ids,
                  int num_reads, const double* rads, double dist_cutoff,
                  const int* cand_offsets, const int* cand_indices,
                  int* out_id, double* out_score, double* out_total);

/* P for every (read, dye-seq) pair, [read][dye_seq]. OpenMP over reads. */
void wpo_pair_probabilities(const wp_model* m, unsigned int T, unsigned int C,
                            int num_dye_seqs, const short* seqs, int seq_cap,
                            const int* lens, int num_reads, const double* rads,
                            double dist_cutoff, double* out_p);

/* Independent check: sums over all hidden paths of the generative process of
 * generate-dye-track.cc:29-110 (per-dye dud/bleach fates, block, detach,
 * Edman success per cycle) times the emission densities. No pruning. Only for
 * tiny inputs (total dyes <= 6, T <= 5). */
double wpo_brute_force_probability(const wp_model* m, unsigned int T,
                                   unsigned int C, const short* seq, int len,
                                   const double* rad);

/* GenericHMM::improve_fit for one (dye-seq, radiometry) pair
 * (generic-hmm.h:60-109 with every step's backward and improve_fit): adds the
 * pair's sufficient statistics into stats (layout WP_FIT_STATS_LEN in
 * whatprot_amd.h) and returns P. Nothing is added when P == 0. */
double wpo_improve_fit(const wp_model* m, unsigned int T, unsigned int C,
                       const short* seq, int len, const double* rad,
                       unsigned int M, double cutoff, double* stats);

/* One E-step of HMMFitter::fit (hmm-fitter.h:63-78): the sum over reads of
 * the per-read statistics (uncut ranges, as run-fit.cc:64-65 sets) and the
 * sum of ln P. OpenMP over reads, partial sums combined in thread order. */
void wpo_fit_estep(const wp_model* m, unsigned int T, unsigned int C,
                   const short* seq, int len, int num_reads, const double* rads,
                   unsigned int M, double* stats, double* sum_log_p);

int wpo_num_threads(void);

#ifdef __cplusplus
}
#endif

#endif
