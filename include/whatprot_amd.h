/*
 * whatprot_amd — C ABI of the MI355X (gfx950) implementation of whatprot's
 * tensor-state HMM scoring path.
 *
 * Plain C, caller-owned host buffers unless a name ends in _device, no
 * exceptions. Every entry point returns WP_OK (0) or a negative status; the
 * text of the last error on the calling thread is wp_last_error().
 *
 * Reference interfaces replaced (paths relative to whatprot's cc_code/src):
 *   SequencingModel / ChannelModel      parameterization/model/sequencing-model.h:22-49,
 *                                       parameterization/model/channel-model.h:20-83
 *   HMMClassifier(T, C, model, settings, dye_seqs)
 *                                       classifiers/hmm-classifier.h:32-38
 *   HMMClassifier::classify(vector<Radiometry>)
 *                                       classifiers/hmm-classifier.h:42-43, hmm-classifier.cc:84-93
 *   HMMClassifier::classify(Radiometry, vector<int> candidate_indices)
 *                                       classifiers/hmm-classifier.h:40-41, hmm-classifier.cc:79-82
 *   PeptideHMM(...).probability()       hmm/hmm/peptide-hmm.h:29-36, peptide-hmm.cc:94-100
 */
#ifndef WHATPROT_AMD_H
#define WHATPROT_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define WP_MAX_CHANNELS 8
#define WP_MAX_TIMESTEPS 64
#define WP_MAX_DYES 15

enum {
    WP_OK = 0,
    WP_ERR_ARG = -1,         /* invalid argument or shape */
    WP_ERR_HIP = -2,         /* HIP runtime error */
    WP_ERR_UNSUPPORTED = -3, /* shape beyond the limits above */
    WP_ERR_NOMEM = -4,       /* allocation failed */
    WP_ERR_NO_DEVICE = -5,   /* no gfx950 device visible */
    WP_ERR_INTERNAL = -6     /* a kernel reported a broken invariant */
};

/* Arithmetic of the forward recursion.
 *   WP_PRECISION_FP64: the reference's method, FP64 end to end, no rescaling.
 *   WP_PRECISION_FP32: FP32 recursion with exact power-of-two rescaling after
 *                      every timestep into an integer log2 accumulator; the
 *                      per-pair log-likelihood, the normaliser over dye-seqs
 *                      and the argmax are FP64. */
enum { WP_PRECISION_FP64 = 0, WP_PRECISION_FP32 = 1 };

/* One channel of the sequencing model; mirrors ChannelModel
 * (parameterization/model/channel-model.h:76-83). interactions[c] and
 * flat_interactions[c] hold the stored values, i.e. 1 - effect and
 * 1 - flat_effect as read from the JSON (sequencing-model.cc:75-88); 1.0 means
 * no interaction. */
typedef struct wp_channel_model {
    double p_bleach;
    double p_dud;
    double bg_sig;
    double mu;
    double sig;
    double interactions[WP_MAX_CHANNELS];
    double flat_interactions[WP_MAX_CHANNELS];
} wp_channel_model;

/* Mirrors SequencingModel (parameterization/model/sequencing-model.h:22-49)
 * with DecayingRateModel p_detach (decaying-rate-model.h): the detach rate at
 * step i is p_detach + p_initial_detach * exp(-i * p_initial_detach_decay). */
typedef struct wp_model {
    int num_channels;
    double p_edman_failure;
    double p_detach;
    double p_initial_detach;
    double p_initial_detach_decay;
    double p_initial_block;
    double p_cyclic_block;
    wp_channel_model channels[WP_MAX_CHANNELS];
} wp_model;

/* Text of the last error on this thread ("" if none). */
const char* wp_last_error(void);

/* Mirrors SequencingModel::with_mu_as_one (sequencing-model.cc:140-151,
 * channel-model.cc:35-45): bg_sig /= mu, sig /= mu, mu = 1. Host only. */
void wp_model_with_mu_as_one(const wp_model* in, wp_model* out);

/* Number of visible gfx950 devices (0 if none). */
int wp_device_count(void);

/* ---------------------------------------------------------------------------
 * HMM classifier. Replaces HMMClassifier (classifiers/hmm-classifier.h:30-78).
 *
 * num_timesteps, num_channels: T and C of the reads.
 * seq_model: the model with mu already set to one (run-classify-hmm.cc:60),
 *            radiometries already divided by mu (radiometries-io.cc:50-51).
 * dist_cutoff: SequencingSettings::dist_cutoff (the -p sigma cutoff);
 *              DBL_MAX means no pruning (peptide-emission.cc:49-53).
 * dye_strings[i]: the dye-seq text, e.g. "..0.1..0" (dye-seq.cc:23-40).
 * dye_seq_counts[i], dye_seq_ids[i]: SourceCount<int>{id, count}
 *              (dye-seqs-io.cc:96-103).
 * precision: WP_PRECISION_FP64 or WP_PRECISION_FP32.
 * The classifier copies everything it needs; the arrays may be freed after.
 * One classifier is used by one host thread at a time.
 * ------------------------------------------------------------------------- */
typedef struct wp_hmm_classifier wp_hmm_classifier;

int wp_hmm_classifier_create(unsigned int num_timesteps,
                             unsigned int num_channels,
                             const wp_model* seq_model,
                             double dist_cutoff,
                             int num_dye_seqs,
                             const char* const* dye_strings,
                             const int* dye_seq_counts,
                             const int* dye_seq_ids,
                             int precision,
                             wp_hmm_classifier** out);

void wp_hmm_classifier_destroy(wp_hmm_classifier* clf);

/* max_num_dyes over the classifier's dye-seqs (hmm-classifier.cc:47-65). */
int wp_hmm_classifier_max_num_dyes(const wp_hmm_classifier* clf);

/* Number of dye-seqs the classifier scores (no reference counterpart). */
int wp_hmm_classifier_num_dye_seqs(const wp_hmm_classifier* clf);

/* Classifies num_reads radiometries laid out [read][t][c] (t*C + c,
 * radiometry.cc:39-44) against every dye-seq. Replaces
 * HMMClassifier::classify(const vector<Radiometry>&)
 * (hmm-classifier.cc:84-93, classify_helper hmm-classifier.h:45-78).
 * Outputs per read (any pointer but out_id may be NULL):
 *   out_id[r]        dye_seq_ids of the best dye-seq (strict >, first wins)
 *   out_adjusted[r]  best / total (the value written to the results CSV by
 *                    scored-classifications-io.cc:41-46), 0 when NaN
 *                    (hmm-classifier.h:73-76)
 *   out_log_score[r] ln P(best)   (-inf when P = 0)
 *   out_log_total[r] ln sum_i P_i * count_i  (-inf when 0) */
int wp_hmm_classify(wp_hmm_classifier* clf,
                    int num_reads,
                    const double* radiometries,
                    int* out_id,
                    double* out_adjusted,
                    double* out_log_score,
                    double* out_log_total);

/* As wp_hmm_classify, restricted per read to candidate dye-seq indices
 * cand_indices[cand_offsets[r] .. cand_offsets[r+1]) in that order (ties go
 * to the earlier candidate). Replaces HMMClassifier::classify(rad,
 * candidate_indices) as called by HybridClassifier (hybrid-classifier.cc:67). */
int wp_hmm_classify_candidates(wp_hmm_classifier* clf,
                               int num_reads,
                               const double* radiometries,
                               const int* cand_offsets,
                               const int* cand_indices,
                               int* out_id,
                               double* out_adjusted,
                               double* out_log_score,
                               double* out_log_total);

/* Device-resident form of wp_hmm_classify: every pointer is device memory,
 * work is queued on `stream` (a hipStream_t, NULL = default stream) and the
 * call returns without waiting. */
int wp_hmm_classify_device(wp_hmm_classifier* clf,
                           int num_reads,
                           const double* d_radiometries,
                           int* d_out_id,
                           double* d_out_adjusted,
                           double* d_out_log_score,
                           double* d_out_log_total,
                           void* stream);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). Same arguments and outputs as wp_hmm_classify_device, and
 * the same kernels, but the call's workspace (read boxes, bin lists, counters,
 * near-best lists) lives in a 64 MiB device array of the library that each
 * GPU holds until the process ends, so no hipMalloc, hipFree, stream, event
 * or graph is created. Launches go eagerly onto `stream` in order (no
 * parallel branches, no graph replay; the quad-rows kernel variants are the
 * defaults, or the autotuned ones once another entry point tuned them). A
 * call whose workspace exceeds 64 MiB (about 85,000 reads at C = 3, T = 12
 * in FP32) returns WP_ERR_ARG before launching anything. Every call on one
 * GPU shares that GPU's array: calls that overlap on one GPU (another stream,
 * another classifier, another thread) race and must be serialised by the
 * caller. */
int wp_hmm_classify_device_noalloc(wp_hmm_classifier* clf,
                                   int num_reads,
                                   const double* d_radiometries,
                                   int* d_out_id,
                                   double* d_out_adjusted,
                                   double* d_out_log_score,
                                   double* d_out_log_total,
                                   void* stream);

/* Device-resident form of wp_hmm_classify_candidates: cand_offsets (N + 1)
 * and cand_indices in device memory; an index outside [0, num_dye_seqs)
 * scores as an empty pair (P = 0, id -1 if it is the pick). */
int wp_hmm_classify_candidates_device(wp_hmm_classifier* clf,
                                      int num_reads,
                                      const double* d_radiometries,
                                      const int* d_cand_offsets,
                                      const int* d_cand_indices,
                                      int* d_out_id,
                                      double* d_out_adjusted,
                                      double* d_out_log_score,
                                      double* d_out_log_total,
                                      void* stream);

/* ln PeptideHMM::probability() for every (read, dye-seq) pair, row-major
 * [read][dye_seq] (peptide-hmm.cc:94-100); -inf where P = 0. Validation and
 * accuracy reporting. */
int wp_hmm_pair_log_probability(wp_hmm_classifier* clf,
                                int num_reads,
                                const double* radiometries,
                                double* out_log_p);

/* Kernel timing (no reference counterpart; the reference prints wall times,
 * cmd-line-out.cc:30-114). When enabled, the classifier records HIP events on
 * the call's stream around its range/bounds kernels and its classify kernels;
 * wp_hmm_last_timing waits for the last call and returns both in ms. */
int wp_hmm_set_timing(wp_hmm_classifier* clf, int enable);
int wp_hmm_last_timing(wp_hmm_classifier* clf, double* prep_ms, double* classify_ms);

/* Kernel selection (no reference counterpart). FP32 reads with count boxes of
 * at most 16 states run the quad-rows kernel in one of two variants with
 * identical outputs: 0 computes every timestep at the bin's box capacity,
 * 1 at each timestep's own box class. mode -1 (the default) times both on the
 * first call with at least 20,000 reads and keeps the faster per bin (smaller
 * calls before it use 0 for bins of up to 4 states and 1 above); 0 or 1 forces
 * one. Setting the mode discards an earlier choice. */
int wp_hmm_set_variant(wp_hmm_classifier* clf, int mode);

/* The per-bin variant in use (0/1, or -1 before the first call) and, once a
 * call has timed them, each variant's best time in ms ([bin][2], -1 if
 * untimed). Writes up to cap bins; returns the number of bins. */
int wp_hmm_get_variants(const wp_hmm_classifier* clf, int* out_variant, double* out_ms, int cap);

/* Graph replay (no reference counterpart). The *_device entry points capture
 * their launch sequence as a HIP graph and replay it while the number of
 * reads, the device pointers and the kernel choice repeat (default on);
 * enable 0 launches every kernel eagerly. */
int wp_hmm_set_graph(wp_hmm_classifier* clf, int enable);

/* Parallel bins (no reference counterpart). Once its reads are binned, a call
 * runs its bins as three parallel branches (the caller's stream and two of
 * the classifier's, joined before the call returns); enable 0 runs them in
 * order on the caller's stream, so each kernel's duration is its own (for
 * profiling). Outputs are identical either way. */
int wp_hmm_set_branches(wp_hmm_classifier* clf, int enable);

/* FP64 decision of FP32 near-ties (no reference counterpart; it makes the FP32
 * path pick the candidate HMMClassifier::classify_helper picks,
 * classifiers/hmm-classifier.h:59-65). An FP32 classifier keeps, per read,
 * the candidates whose FP32 ln P lies within W = abs + rel |ln P_best| of the
 * best; when there are two or more, they are rescored in FP64 on the device
 * and the strict '>' with the earlier candidate winning decides the id and the
 * ln score (the normaliser stays the FP32 one). W must be at least twice the
 * FP32 error of ln P; the defaults are 2e-5 and 1e-5. enable 0 keeps the FP32
 * argmax. Returns WP_ERR_UNSUPPORTED for FP64 classifiers and FP32 ones whose
 * FP64 boxes do not fit in LDS. */
int wp_hmm_set_refine(wp_hmm_classifier* clf, int enable, double abs, double rel);

/* After an FP32 call with the FP64 decision on (synchronises the device):
 * out2[0] = reads whose near-best list was rescored, out2[1] = those among
 * them rescored over every candidate (their window held more than 128). */
int wp_hmm_refine_stats(wp_hmm_classifier* clf, int num_reads, int* out2);

/* Algorithmic work of classifying these reads against every dye-seq (no
 * reference counterpart; used for roofline accounting): out3[0] = non-empty
 * pairs, out3[1] = pruned state-timesteps sum_t (t+1) prod_c (H-L),
 * out3[2] = FLOPs by the per-state counts emission 2, block 3, detach 4,
 * Edman 8, binomial 2(D_c+2) per channel. Host buffers; synchronous. */
int wp_hmm_work(wp_hmm_classifier* clf, int num_reads, const double* radiometries,
                double* out3);

/* ---------------------------------------------------------------------------
 * kNN classifier. Replaces NNClassifier (classifiers/nn-classifier.h:44-82)
 * over the KDTree of kd-tree/kd-tree.h:22-63, and HybridClassifier
 * (classifiers/hybrid-classifier.h:25-52).
 *
 * Dye-tracks as read_dye_tracks gives them (io/dye-tracks-io.cc:32-62):
 *   track_counts[i][t][c]  dyes of channel c at timestep t of track i (short)
 *   sources of track i: [source_offsets[i], source_offsets[i+1]) into
 *   source_ids / source_counts / source_hits (SourceCountHits: dye-seq id,
 *   its count, simulated hits).
 * seq_model has mu = 1 (run-classify-nn.cc), radiometries are divided by mu.
 * ------------------------------------------------------------------------- */
typedef struct wp_nn_classifier wp_nn_classifier;
typedef struct wp_hybrid_classifier wp_hybrid_classifier;

/* The point order the kd-tree construction leaves (kd-tree.h:33-37 with
 * make-node.h:28-80): out_order[p] = input index of the point at position p;
 * out_depth (may be NULL) the tree depth. n points of dimension d, row-major.
 * Host only; the search breaks distance ties by this order. */
int wp_kdtree_order(int k, int d, int n, const double* points, int* out_order, int* out_depth);

/* NNClassifier(T, C, seq_model, k, sig, dye_tracks) (nn-classifier.cc:89-108):
 * builds the kd-tree on the host (make-node.h:28-80) and copies it to the
 * device. k >= 1. */
int wp_nn_classifier_create(unsigned int num_timesteps, unsigned int num_channels,
                            const wp_model* seq_model, int k, double sig, int num_dye_tracks,
                            const short* track_counts, const int* source_offsets,
                            const int* source_ids, const int* source_counts,
                            const int* source_hits, wp_nn_classifier** out);

void wp_nn_classifier_destroy(wp_nn_classifier* clf);

/* KDTree::search for every read (kd-tree.h:40-46): out_count[r] <= k
 * neighbours in KBest::fill order (farthest first, k-best.h:45-55),
 * out_track[r * k + i] the dye-track index, out_dist_sq[r * k + i] its squared
 * distance. */
int wp_nn_search(wp_nn_classifier* clf, int num_reads, const double* radiometries,
                 int* out_count, int* out_track, double* out_dist_sq);

/* NNClassifier::classify(vector<Radiometry>) (nn-classifier.cc:153-177,
 * 206-215): best id, its score and the total (adjusted = score / total; the
 * NaN guard sets score 0, total 1). */
int wp_nn_classify(wp_nn_classifier* clf, int num_reads, const double* radiometries,
                   int* out_id, double* out_score, double* out_total);

/* NNClassifier::classify(radiometry, h) (nn-classifier.cc:179-204): up to h
 * ids per read in ascending score order at out_ids[r * h + i], their scores,
 * and the read's total. */
int wp_nn_classify_top(wp_nn_classifier* clf, int num_reads, const double* radiometries, int h,
                       int* out_count, int* out_ids, double* out_scores, double* out_total);

/* HybridClassifier(T, C, seq_model, seq_settings, k, sig, dye_tracks, h,
 * dye_seqs) (hybrid-classifier.cc:25-41): an NN classifier and an HMM
 * classifier (precision as wp_hmm_classifier_create). */
int wp_hybrid_classifier_create(unsigned int num_timesteps, unsigned int num_channels,
                                const wp_model* seq_model, double dist_cutoff, int k, double sig,
                                int num_dye_tracks, const short* track_counts,
                                const int* source_offsets, const int* source_ids,
                                const int* source_counts, const int* source_hits, int h,
                                int num_dye_seqs, const char* const* dye_strings,
                                const int* dye_seq_counts, const int* dye_seq_ids, int precision,
                                wp_hybrid_classifier** out);

void wp_hybrid_classifier_destroy(wp_hybrid_classifier* clf);

/* HybridClassifier::classify (hybrid-classifier.cc:43-95): the top-h kNN ids
 * as candidates of the HMM, the HMM score scaled by the candidates' kNN
 * share. out_adjusted is the written score (score / total). */
int wp_hybrid_classify(wp_hybrid_classifier* clf, int num_reads, const double* radiometries,
                       int* out_id, double* out_adjusted);

/* ---------------------------------------------------------------------------
 * Learned shortlist (no reference counterpart: it proposes HMM candidates in
 * place of, or beside, the kNN stage of HybridClassifier). A ReLU MLP
 * in_dim -> hidden (x num_hidden) -> out_dim with BF16 weights and
 * activations, FP32 accumulation on the matrix cores, FP32 master weights
 * trained with Adam (beta 0.9 / 0.999, eps 1e-8) on the mean softmax
 * cross-entropy. Layer l's weights are [rows][cols] row-major with
 * rows = in_dim (l = 0) or hidden, cols = hidden or out_dim (last layer).
 * Inputs are radiometries divided by mu, [read][t][c].
 * ------------------------------------------------------------------------- */
typedef struct wp_mlp wp_mlp;

/* Random initialisation from seed (He-uniform for the ReLU layers,
 * LeCun-uniform for the logits, zero biases). */
int wp_mlp_create(int in_dim, int hidden, int num_hidden, int out_dim, uint64_t seed, wp_mlp** out);
void wp_mlp_destroy(wp_mlp* mlp);
int wp_mlp_layer_shape(const wp_mlp* mlp, int layer, int* rows, int* cols);
/* FP32 master weights and biases of one layer (host buffers). */
int wp_mlp_get_layer(wp_mlp* mlp, int layer, float* W, float* b);
int wp_mlp_set_layer(wp_mlp* mlp, int layer, const float* W, const float* b);
/* FP32 logits [B][out_dim] of B inputs [B][in_dim] (host buffers). */
int wp_mlp_logits(wp_mlp* mlp, int B, const float* x, float* out_logits);
/* Mean cross-entropy of B labelled inputs and its gradients: gW[l], gb[l]
 * host buffers shaped like the layers. No update. */
int wp_mlp_gradients(wp_mlp* mlp, int B, const float* x, const int* labels, double* out_loss, float* const* gW,
                     float* const* gb);
/* Adam training: `epochs` passes over N labelled inputs in seeded shuffled
 * order, batches of `batch` (the last partial batch of an epoch is dropped);
 * out_epoch_loss[e] (may be NULL) is the mean batch loss of epoch e. */
int wp_mlp_train(wp_mlp* mlp, int N, const float* x, const int* labels, int batch, int epochs, float lr,
                 uint64_t seed, float* out_epoch_loss);
/* The H highest-scoring outputs per read, in increasing index order
 * (out_index[r * H + j], -1 past the end), the largest softmax probability
 * per read and the softmax mass of the H (either may be NULL). */
int wp_mlp_shortlist(wp_mlp* mlp, int num_reads, const double* radiometries, int H, int* out_index,
                     float* out_max_posterior, float* out_mass);

/* Confidence a shortlist must reach to be trusted: the largest MLP
 * posterior, or the posterior mass of the H shortlisted dye-seqs. */
enum { WP_SHORTLIST_MAX_POSTERIOR = 0, WP_SHORTLIST_MASS = 1 };

/* Classification with the learned shortlist: reads whose confidence
 * (criterion) is >= min_posterior are scored by the HMM against their H
 * shortlisted dye-seqs (output index i = dye-seq i of clf), the others
 * against every dye-seq, so every output is an HMM score (outputs as
 * wp_hmm_classify's out_id / out_adjusted). */
int wp_shortlist_classify(wp_hmm_classifier* clf, wp_mlp* mlp, int num_reads, const double* radiometries, int H,
                          double min_posterior, int criterion, int* out_id, double* out_adjusted,
                          int* out_num_fallback);

/* Device-resident form of wp_shortlist_classify (no reference counterpart):
 * radiometries and outputs in device memory. The call first waits for
 * `stream`, and returns when the outputs are written; the MLP's and the
 * shortlist's buffers are kept by the MLP between calls. */
int wp_shortlist_classify_device(wp_hmm_classifier* clf, wp_mlp* mlp, int num_reads, const double* d_radiometries,
                                 int H, double min_posterior, int criterion, int* d_out_id, double* d_out_adjusted,
                                 int* out_num_fallback, void* stream);

/* ---------------------------------------------------------------------------
 * Baum-Welch fit. Replaces HMMFitter (fitters/hmm-fitter.h:30-151) as run by
 * run_fit (main/run-fit.cc:37-140): one dye-seq, uncut ranges (the reference
 * sets dist_cutoff = DBL_MAX for fitting, run-fit.cc:64-65).
 *
 * Sufficient statistics of one E-step, the fields of SequencingModelFitter
 * (parameterization/fit/sequencing-model-fitter.h:33-38) summed over reads,
 * as one FP64 array of WP_FIT_STATS_LEN(T, C):
 *   [WP_FIT_P_EDMAN_FAILURE + 0/1]  p_edman_failure numerator, denominator
 *   [WP_FIT_P_INITIAL_BLOCK + 0/1]  p_initial_block numerator, denominator
 *   [WP_FIT_P_CYCLIC_BLOCK + 0/1]   p_cyclic_block numerator, denominator
 *   [WP_FIT_DETACH_X + t]           p_detach x_t, t = 0..T-2
 *   [WP_FIT_DETACH_X + T-1 + t]     p_detach n_t
 *                                   (DecayingParameterFitter::add_timestep)
 *   [WP_FIT_CHANNEL(T, c) + 0/1]    channel c p_bleach numerator, denominator
 *   [WP_FIT_CHANNEL(T, c) + 2/3]    channel c p_dud numerator, denominator
 * ------------------------------------------------------------------------- */
#define WP_FIT_STATS_LEN(T, C) (6 + 2 * ((T) - 1) + 4 * (C))
#define WP_FIT_P_EDMAN_FAILURE 0
#define WP_FIT_P_INITIAL_BLOCK 2
#define WP_FIT_P_CYCLIC_BLOCK 4
#define WP_FIT_DETACH_X 6
#define WP_FIT_CHANNEL(T, c) (6 + 2 * ((T) - 1) + 4 * (c))

/* FitSettings (parameterization/settings/fit-settings.h:20-37,
 * channel-fit-settings.h:20-26): non-zero holds a parameter at its starting
 * value. bg_sig, mu and sig are never fitted (channel-model-fitter.cc:39-50). */
typedef struct wp_fit_settings {
    int hold_p_edman_failure;
    int hold_p_detach;
    int hold_p_initial_detach;
    int hold_p_initial_detach_decay;
    int hold_p_initial_block;
    int hold_p_cyclic_block;
    int hold_p_bleach[WP_MAX_CHANNELS];
    int hold_p_dud[WP_MAX_CHANNELS];
} wp_fit_settings;

typedef struct wp_hmm_fitter wp_hmm_fitter;

/* Replaces HMMFitter(T, C, ..., dye_seq) (hmm-fitter.h:35-43,
 * hmm-fitter.cc:30-61) with its reads: parses the dye-seq, copies
 * num_reads radiometries [read][t][c] (already divided by mu,
 * radiometries-io.cc:50-51) to the device and sizes the E-step. precision:
 * WP_PRECISION_FP64 (the reference's arithmetic) or WP_PRECISION_FP32 (FP32
 * forward/backward with power-of-two rescaling, FP64 statistics). */
int wp_hmm_fitter_create(unsigned int num_timesteps, unsigned int num_channels,
                         const char* dye_string, int num_reads, const double* radiometries,
                         int precision, wp_hmm_fitter** out);

void wp_hmm_fitter_destroy(wp_hmm_fitter* fitter);

/* One E-step under seq_model (mu = 1): the sufficient statistics of every
 * read's PeptideHMM::improve_fit summed over reads (hmm-fitter.h:63-78) and
 * sum_r ln P_r (HMMFitter::log_likelihood, hmm-fitter.h:120-140). Either
 * output may be NULL. A read with P = 0 adds no statistics and -inf to the
 * sum (generic-hmm.h:77-87). */
int wp_hmm_fitter_estep(wp_hmm_fitter* fitter, const wp_model* seq_model, double* out_stats,
                        double* out_sum_log_p);

/* The M-step of HMMFitter::fit (hmm-fitter.h:80-104): the dud correction for
 * reads hidden by all-dud dye-seqs, SequencingModelFitter::get
 * (sequencing-model-fitter.cc:68-78, decaying-parameter-fitter.cc:42-66) and
 * update_with_holds (hmm-fitter.cc:63-97). out_step_size receives
 * SequencingModel::distance (sequencing-model.cc:153-166). settings may be
 * NULL (nothing held). Host only. */
int wp_fit_mstep(unsigned int num_timesteps, unsigned int num_channels, const char* dye_string,
                 const wp_model* seq_model, int num_reads, const double* stats,
                 const wp_fit_settings* settings, wp_model* out_model, double* out_step_size);

/* HMMFitter::fit (hmm-fitter.h:50-116): EM until the step size falls below
 * stopping_threshold or max_runtime seconds have passed, returning in
 * out_log_l the log-likelihood under the fitted model. max_iterations > 0
 * also stops after that many iterations (no reference counterpart; the
 * reference stops on the threshold or the clock only). */
int wp_hmm_fit(wp_hmm_fitter* fitter, const wp_model* seq_model, const wp_fit_settings* settings,
               double stopping_threshold, double max_runtime, int max_iterations,
               wp_model* out_model, double* out_step_size, double* out_log_l, int* out_iterations);

/* Bootstrap rounds in one batch (no reference counterpart; default on): the
 * rounds' EM loops run in lockstep and every iteration's E-steps of all
 * running rounds are one launch, each round's statistics identical bit for
 * bit to a launch of its own; the wall-clock limit then counts from the start
 * of the batch. enable 0 runs the rounds one after another. */
int wp_hmm_fitter_set_batched(wp_hmm_fitter* fitter, int enable);

/* Bootstrap fit (fitters/bootstrap-fit.cc:37-173, as run by run-fit.cc:
 * 107-130): the fit of wp_hmm_fit on every read (out_best_model,
 * out_step_size, out_log_l), then num_rounds fits from seq_model on
 * resamples of the reads drawn with replacement (out_models[i],
 * out_log_ls[i]: the model of round i and the log-likelihood of its
 * resample), the largest final step size of a round, and per parameter the
 * order statistics at (1 - ci) / 2 and 1 - (1 - ci) / 2 of the rounds as the
 * lower and upper bounds (the other fields as out_best_model; index clamped
 * to the last round). A resample is evaluated on the device as per-read
 * multiplicities (wp_fit_bootstrap_counts), so the reads are not copied
 * again; round i draws from its own seeded stream in place of the
 * reference's clock-seeded engine. */
int wp_hmm_fit_bootstrap(wp_hmm_fitter* fitter, const wp_model* seq_model, const wp_fit_settings* settings,
                         double stopping_threshold, double max_runtime, int num_rounds,
                         double confidence_interval, uint64_t seed, wp_model* out_models, double* out_log_ls,
                         wp_model* out_best_model, double* out_step_size, double* out_log_l,
                         wp_model* out_ci_min, wp_model* out_ci_max, double* out_worst_step);

/* Multiplicity of each of num_reads reads in bootstrap round `round`:
 * num_reads draws of a uniform index from stream `round` of the seed (the
 * resampling of bootstrap-fit.cc:76-85). Host only. */
int wp_fit_bootstrap_counts(int num_reads, uint64_t seed, int round, int* out_counts);

/* least_squares_fit_of_offset_exponential (util/levmar-helper.cc:44-85):
 * weighted fit of y_t ~ p[0] + p[1] exp(-p[2] t) for t = 0..n-1 within
 * [0, 1]^3, hold[i] != 0 pinning p[i]; p is updated in place. Host only. */
int wp_fit_offset_exponential(int n, const double* y, const double* w, const int* hold, double* p);

/* ---------------------------------------------------------------------------
 * Synthetic inputs (host, and the same streams on the GPU below). Seeded and
 * portable; the streams differ from the reference's clock-seeded
 * std::default_random_engine, the distributions follow its simulator. Each
 * simulated peptide (dye tracks, in dye-seq, copy, repeat order) or read
 * attempt (radiometries) g draws from its own xoshiro256** stream seeded by
 * splitmix64 from seed + (g + 1) * 0xD1B54A32D192ED03.
 * ------------------------------------------------------------------------- */

/* Random dye-seq strings, num_dye_seqs rows of `stride` bytes (NUL padded):
 * length ~ U{len_min..len_max}, per-channel dye counts ~ U{dyes_min..dyes_max}
 * capped by the length and redrawn while all are zero, positions uniform
 * without replacement. No reference counterpart (the reference reads dye-seqs
 * made by its Python tools, README.md:310-325). */
int wp_generate_dye_seqs(int num_channels, int num_dye_seqs, int len_min, int len_max,
                         int dyes_min, int dyes_max, uint64_t seed, char* out_strings,
                         int stride);

/* Replaces generate_radiometries (simulation/generate-radiometries.cc:30-64,
 * generate-radiometry.cc:28-62, generate-dye-track.cc:29-110): draws a dye-seq
 * by weight (NULL = uniform), simulates its dye track, drops invisible tracks
 * (all zero at t = 0) and emits Normal(adjusted_mu, sigma) intensities rounded
 * to float32, until num_reads reads exist. out_radiometries is [read][t][c] in
 * the model's units (divide by mu for a model with mu = 1). */
int wp_simulate_radiometries(const wp_model* seq_model, unsigned int num_timesteps,
                             unsigned int num_channels, int num_dye_seqs,
                             const char* const* dye_strings, const int* weights,
                             int num_reads, uint64_t seed, double* out_radiometries,
                             int* out_dye_seq_index);

/* generate_dye_tracks (simulation/generate-dye-tracks.cc:28-63) followed by
 * the deduplication of dedup-dye-tracks.cc:35-148: per_peptide tracks for each
 * of dye_seq_counts[i] peptides of every dye-seq, all-zero tracks dropped,
 * identical tracks merged with per source dye-seq id its count and hits.
 * Output in the layout wp_nn_classifier_create takes, tracks ordered by their
 * counts, sources by id. out_track_counts holds max_tracks tracks and
 * out_source_offsets max_tracks + 1 entries; the three source arrays hold
 * max_sources entries (one per distinct (track, dye-seq id); at most the
 * number of simulated peptides). With more tracks or sources than that it
 * writes only *out_num_tracks and *out_num_sources and fails with
 * WP_ERR_ARG. */
int wp_simulate_dye_tracks(const wp_model* seq_model, unsigned int num_timesteps,
                           unsigned int num_channels, int num_dye_seqs,
                           const char* const* dye_strings, const int* dye_seq_counts,
                           const int* dye_seq_ids, int per_peptide, uint64_t seed, int max_tracks,
                           int max_sources, int* out_num_tracks, int* out_num_sources,
                           short* out_track_counts, int* out_source_offsets,
                           int* out_source_ids, int* out_source_counts, int* out_source_hits);

/* The two simulators above on the GPU (no reference counterpart for the
 * device itself; the survey's first "next" component, replacing the
 * reference's MR-MPI dedup-dye-tracks run): one thread per peptide or read
 * attempt, each drawing from its own stream exactly as the host functions
 * do, so the dye tracks (and the deduplicated output) are identical to
 * wp_simulate_dye_tracks / the tracks behind wp_simulate_radiometries; the
 * deduplication is a device sort of the packed tracks. Radiometries can
 * differ from the host's in the last bit of a float32 where the device's
 * log or pow rounds differently. Dye-seqs of at most 64 positions;
 * wp_simulate_radiometries_gpu takes at most 4 channels. */
int wp_simulate_radiometries_gpu(const wp_model* seq_model, unsigned int num_timesteps,
                                 unsigned int num_channels, int num_dye_seqs,
                                 const char* const* dye_strings, const int* weights, int num_reads,
                                 uint64_t seed, double* out_radiometries, int* out_dye_seq_index);

int wp_simulate_dye_tracks_gpu(const wp_model* seq_model, unsigned int num_timesteps,
                               unsigned int num_channels, int num_dye_seqs,
                               const char* const* dye_strings, const int* dye_seq_counts,
                               const int* dye_seq_ids, int per_peptide, uint64_t seed, int max_tracks,
                               int max_sources, int* out_num_tracks, int* out_num_sources,
                               short* out_track_counts, int* out_source_offsets,
                               int* out_source_ids, int* out_source_counts, int* out_source_hits);


/* ---------------------------------------------------------------------------
 * Text files. The reference's formats, read in parallel and written
 * byte-identically (17 significant digits where the reference writes
 * setprecision(17)).
 * ------------------------------------------------------------------------- */
typedef struct wp_radiometries {
    int num_timesteps, num_channels, num_reads;
    double* values; /* [read][t][c], divided by the model's mu_c; library-allocated */
} wp_radiometries;

/* read_radiometries (io/radiometries-io.cc:33-56): T, C, N, then the N*T*C
 * intensities, each divided by its channel's mu in seq_model (NULL: by 1).
 * Free with wp_free_radiometries. */
int wp_read_radiometries(const char* path, const wp_model* seq_model, wp_radiometries* out);
void wp_free_radiometries(wp_radiometries* r);

/* write_radiometries (io/radiometries-io.cc:58-81): each value times mu_c of
 * seq_model (NULL: 1), tab-separated, one read per line, setprecision(17). */
int wp_write_radiometries(const char* path, const wp_model* seq_model, int num_timesteps, int num_channels,
                          int num_reads, const double* values);

typedef struct wp_dye_tracks {
    int num_timesteps, num_channels, num_dye_tracks, num_sources;
    short* counts;       /* [track][t][c] */
    int* source_offsets; /* [num_dye_tracks + 1] */
    int* source_ids;     /* [num_sources]: SourceCountHits id, count, hits */
    int* source_counts;
    int* source_hits;
} wp_dye_tracks;

/* read_dye_tracks (io/dye-tracks-io.cc:32-64). Free with wp_free_dye_tracks. */
int wp_read_dye_tracks(const char* path, wp_dye_tracks* out);
void wp_free_dye_tracks(wp_dye_tracks* t);

/* write_dye_tracks (io/dye-tracks-io.cc:66-103). */
int wp_write_dye_tracks(const char* path, int num_timesteps, int num_channels, int num_dye_tracks,
                        const short* counts, const int* source_offsets, const int* source_ids,
                        const int* source_counts, const int* source_hits);

/* write_scored_classifications_raw (io/scored-classifications-io.cc:54-68):
 * header, then "i,id,score" with the score at setprecision(17). */
int wp_write_scored_classifications(const char* path, int num_reads, const int* ids, const double* scores);

#ifdef __cplusplus
}
#endif

#endif /* WHATPROT_AMD_H */
