// C ABI of the Baum-Welch fitter: the GPU E-step, the host M-step and the EM
// loop of HMMFitter::fit. Declarations and reference citations are in
// include/whatprot_amd.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "fit_kernels.h"
#include "model_tables.h"
#include "sim_rng.h"
#include "wp_internal.h"

struct wp_hmm_fitter {
    int T = 0, C = 0, N = 0, precision = 0, M = 0, NB = 1, tri = 1, nfib = 0, grid = 1, L = 0;
    int rows = 0;  // row-lane E-step kernel
    int D[WP_MAX_CHANNELS] = {0};
    std::vector<short> seq;  // parsed dye-seq
    int len = 0;
    double* d_rads = nullptr;
    unsigned char* d_crd = nullptr;
    unsigned short* d_fbase = nullptr;
    signed char* d_seq = nullptr;
    unsigned char* d_track = nullptr;
    double* d_binom = nullptr;
    void* d_hist = nullptr;
    size_t hist_stride = 0;
    double* d_partial = nullptr;
    double* d_out = nullptr;
    int* d_weight = nullptr;  // bootstrap multiplicities [rounds][N]
    int weight_rows = 0;
    // batched bootstrap: every active round's E-step in one launch
    int batched = 1;
    int cap_slots = 0;  // capacity of the arrays below
    wp::DevModel* d_models = nullptr;
    int* d_slot_round = nullptr;
    double* d_binom_b = nullptr;
    double* d_partial_b = nullptr;
    double* d_out_b = nullptr;
};

namespace {

void free_fitter(wp_hmm_fitter* f) {
    hipFree(f->d_rads);
    hipFree(f->d_crd);
    hipFree(f->d_fbase);
    hipFree(f->d_seq);
    hipFree(f->d_track);
    hipFree(f->d_binom);
    hipFree(f->d_hist);
    hipFree(f->d_partial);
    hipFree(f->d_out);
    hipFree(f->d_weight);
    hipFree(f->d_models);
    hipFree(f->d_slot_round);
    hipFree(f->d_binom_b);
    hipFree(f->d_partial_b);
    hipFree(f->d_out_b);
    delete f;
}

// Device copy of h[0..n) (h may be null: allocation only).
template <typename T>
int put(T** d, const T* h, size_t n) {
    WP_HIP(hipMalloc((void**)d, sizeof(T) * std::max<size_t>(n, 1)));
    if (n && h) WP_HIP(hipMemcpy(*d, h, sizeof(T) * n, hipMemcpyHostToDevice));
    return WP_OK;
}

// Parses a dye-seq for a fit with C channels; fills seq/len and D.
int parse_fit_dye_seq(const char* s, int C, std::vector<short>* seq, int* len, int* D) {
    if (!s) {
        wp::set_error("null dye-seq");
        return WP_ERR_ARG;
    }
    size_t n = std::strlen(s);
    seq->assign(n + 1, -1);
    *len = wp::parse_dye_seq(s, seq->data(), (int)n + 1);
    for (int c = 0; c < C; c++) D[c] = 0;
    for (int i = 0; i < *len; i++) {
        int c = (*seq)[i];
        if (c >= C) {
            wp::set_error("dye-seq names a channel >= num_channels");
            return WP_ERR_ARG;
        }
        if (c >= 0) D[c]++;
    }
    for (int c = 0; c < C; c++) {
        if (D[c] > WP_MAX_DYES) {
            wp::set_error("more than 15 dyes of one channel in the dye-seq");
            return WP_ERR_UNSUPPORTED;
        }
    }
    return WP_OK;
}

// SequencingModel::distance (sequencing-model.cc:153-166).
double model_distance(const wp_model& a, const wp_model& b, int C) {
    double d = 0.0;
    d = std::max(d, std::fabs(a.p_edman_failure - b.p_edman_failure));
    d = std::max(d, std::fabs(a.p_detach - b.p_detach));
    d = std::max(d, std::fabs(a.p_initial_detach - b.p_initial_detach));
    d = std::max(d, std::fabs(a.p_initial_detach_decay - b.p_initial_detach_decay));
    d = std::max(d, std::fabs(a.p_initial_block - b.p_initial_block));
    d = std::max(d, std::fabs(a.p_cyclic_block - b.p_cyclic_block));
    for (int c = 0; c < C; c++) {
        d = std::max(d, std::fabs(a.channels[c].p_bleach - b.channels[c].p_bleach));
        d = std::max(d, std::fabs(a.channels[c].p_dud - b.channels[c].p_dud));
    }
    return d;
}

int estep(wp_hmm_fitter* f, const wp_model* model, double* stats, double* sum_log_p, const int* d_weight = nullptr) {
    if (!model || model->num_channels != f->C) {
        wp::set_error("model channels != fitter channels");
        return WP_ERR_ARG;
    }
    wp::FitArgs a;
    std::memset(&a, 0, sizeof(a));
    wp::fill_dev_model(*model, f->T, f->C, f->M, DBL_MAX, &a.m);
    std::vector<double> binom((size_t)2 * f->C * f->tri);
    wp::binomial_tables(*model, f->C, f->M, binom.data());
    WP_HIP(hipMemcpy(f->d_binom, binom.data(), sizeof(double) * binom.size(), hipMemcpyHostToDevice));
    a.N = f->N;
    a.rads = f->d_rads;
    a.NB = f->NB;
    for (int c = 0; c < f->C; c++) a.D[c] = f->D[c];
    a.crd = f->d_crd;
    a.fbase = f->d_fbase;
    a.nfib = f->nfib;
    a.seq = f->d_seq;
    a.track = f->d_track;
    a.binom = f->d_binom;
    a.tri = f->tri;
    a.hist = f->d_hist;
    a.hist_stride = f->hist_stride;
    a.rows = f->rows;
    a.partial = f->d_partial;
    a.L = f->L;
    a.weight = d_weight;
    a.nslots = 1;
    a.models = nullptr;
    a.slot_round = nullptr;
    std::vector<double> out((size_t)f->L + 1, 0.0);
    if (f->N > 0) {
        WP_HIP(wp::launch_fit_estep(a, f->precision, f->grid, f->d_out, nullptr));
        WP_HIP(hipMemcpy(out.data(), f->d_out, sizeof(double) * out.size(), hipMemcpyDeviceToHost));
    }
    if (stats) std::memcpy(stats, out.data(), sizeof(double) * f->L);
    if (sum_log_p) *sum_log_p = out[f->L];
    return WP_OK;
}

// The E-steps of several bootstrap rounds in one launch: slot k runs under
// models[k] with weight row rounds[k] of f->d_weight; out[k][0..L] as
// estep's stats and sum of ln P. Each slot's result is estep's bit for bit.
int estep_batch(wp_hmm_fitter* f, const std::vector<wp_model>& models, const std::vector<int>& rounds,
                std::vector<double>* out) {
    const int K = (int)models.size();
    out->assign((size_t)K * (f->L + 1), 0.0);
    if (K == 0 || f->N == 0) return WP_OK;
    if (K > f->cap_slots) {
        hipFree(f->d_models);
        hipFree(f->d_slot_round);
        hipFree(f->d_binom_b);
        hipFree(f->d_partial_b);
        hipFree(f->d_out_b);
        f->d_models = nullptr;
        f->d_slot_round = nullptr;
        f->d_binom_b = f->d_partial_b = f->d_out_b = nullptr;
        f->cap_slots = 0;
        WP_HIP(hipMalloc(&f->d_models, sizeof(wp::DevModel) * K));
        WP_HIP(hipMalloc(&f->d_slot_round, sizeof(int) * K));
        WP_HIP(hipMalloc(&f->d_binom_b, sizeof(double) * (size_t)K * 2 * f->C * f->tri));
        WP_HIP(hipMalloc(&f->d_partial_b, sizeof(double) * (size_t)K * f->grid * (f->L + 1)));
        WP_HIP(hipMalloc(&f->d_out_b, sizeof(double) * (size_t)K * (f->L + 1)));
        f->cap_slots = K;
    }
    std::vector<wp::DevModel> dm(K);
    std::vector<double> binom((size_t)K * 2 * f->C * f->tri);
    for (int k = 0; k < K; k++) {
        if (models[k].num_channels != f->C) {
            wp::set_error("model channels != fitter channels");
            return WP_ERR_ARG;
        }
        wp::fill_dev_model(models[k], f->T, f->C, f->M, DBL_MAX, &dm[k]);
        wp::binomial_tables(models[k], f->C, f->M, binom.data() + (size_t)k * 2 * f->C * f->tri);
    }
    WP_HIP(hipMemcpy(f->d_models, dm.data(), sizeof(wp::DevModel) * K, hipMemcpyHostToDevice));
    WP_HIP(hipMemcpy(f->d_slot_round, rounds.data(), sizeof(int) * K, hipMemcpyHostToDevice));
    WP_HIP(hipMemcpy(f->d_binom_b, binom.data(), sizeof(double) * binom.size(), hipMemcpyHostToDevice));
    wp::FitArgs a;
    std::memset(&a, 0, sizeof(a));
    a.m = dm[0];  // T and C for the launch; each slot reads its own model
    a.N = f->N;
    a.rads = f->d_rads;
    a.NB = f->NB;
    for (int c = 0; c < f->C; c++) a.D[c] = f->D[c];
    a.crd = f->d_crd;
    a.fbase = f->d_fbase;
    a.nfib = f->nfib;
    a.seq = f->d_seq;
    a.track = f->d_track;
    a.binom = f->d_binom_b;
    a.tri = f->tri;
    a.hist = f->d_hist;
    a.hist_stride = f->hist_stride;
    a.rows = f->rows;
    a.partial = f->d_partial_b;
    a.L = f->L;
    a.weight = f->d_weight;
    a.nslots = K;
    a.models = f->d_models;
    a.slot_round = f->d_slot_round;
    WP_HIP(wp::launch_fit_estep(a, f->precision, f->grid, f->d_out_b, nullptr));
    WP_HIP(hipMemcpy(out->data(), f->d_out_b, sizeof(double) * out->size(), hipMemcpyDeviceToHost));
    return WP_OK;
}

}  // namespace

extern "C" {

int wp_fit_offset_exponential(int n, const double* y, const double* w, const int* hold, double* p) {
    if (n < 0 || (n > 0 && (!y || !w)) || !hold || !p) {
        wp::set_error("null argument");
        return WP_ERR_ARG;
    }
    wp::offset_exponential_fit(n, y, w, hold, p);
    return WP_OK;
}

int wp_fit_mstep(unsigned int num_timesteps, unsigned int num_channels, const char* dye_string,
                 const wp_model* model, int num_reads, const double* stats,
                 const wp_fit_settings* settings, wp_model* out_model, double* out_step_size) {
    const int T = (int)num_timesteps, C = (int)num_channels;
    if (!model || !stats || !out_model || T < 1 || T > WP_MAX_TIMESTEPS || C < 1 ||
        C > WP_MAX_CHANNELS || model->num_channels != C) {
        wp::set_error("bad argument to wp_fit_mstep");
        return WP_ERR_ARG;
    }
    std::vector<short> seq;
    int len = 0, D[WP_MAX_CHANNELS];
    int rc = parse_fit_dye_seq(dye_string, C, &seq, &len, D);
    if (rc) return rc;
    wp_fit_settings none;
    std::memset(&none, 0, sizeof(none));
    const wp_fit_settings& h = settings ? *settings : none;
    const int L = WP_FIT_STATS_LEN(T, C);
    std::vector<double> st(stats, stats + L);
    // Correction for reads never seen because every dye is a dud
    // (hmm-fitter.h:80-101), added once per dye.
    double ratio_hidden = 1.0;
    for (int i = 0; i < len; i++)
        if (seq[i] != -1) ratio_hidden *= model->channels[seq[i]].p_dud;
    double magic_ratio = 1.0 / (1.0 - ratio_hidden) - 1.0;
    double expected_hidden = magic_ratio * num_reads;
    for (int i = 0; i < len; i++) {
        if (seq[i] != -1) {
            st[WP_FIT_CHANNEL(T, seq[i]) + 2] += expected_hidden;
            st[WP_FIT_CHANNEL(T, seq[i]) + 3] += expected_hidden;
        }
    }
    // SequencingModelFitter::get (sequencing-model-fitter.cc:68-78): ratios,
    // and the decaying detach rate by a weighted fit of x_t / n_t
    // (decaying-parameter-fitter.cc:42-66).
    wp_model upd = *model;
    upd.p_edman_failure = st[WP_FIT_P_EDMAN_FAILURE] / st[WP_FIT_P_EDMAN_FAILURE + 1];
    upd.p_initial_block = st[WP_FIT_P_INITIAL_BLOCK] / st[WP_FIT_P_INITIAL_BLOCK + 1];
    upd.p_cyclic_block = st[WP_FIT_P_CYCLIC_BLOCK] / st[WP_FIT_P_CYCLIC_BLOCK + 1];
    {
        int n = T - 1;
        std::vector<double> y((size_t)std::max(n, 1)), w((size_t)std::max(n, 1));
        for (int k = 0; k < n; k++) {
            y[k] = st[WP_FIT_DETACH_X + k] / st[WP_FIT_DETACH_X + n + k];
            w[k] = st[WP_FIT_DETACH_X + n + k];
        }
        double p[3] = {model->p_detach, model->p_initial_detach, model->p_initial_detach_decay};
        int hold[3] = {h.hold_p_detach, h.hold_p_initial_detach, h.hold_p_initial_detach_decay};
        wp::offset_exponential_fit(n, y.data(), w.data(), hold, p);
        upd.p_detach = p[0];
        upd.p_initial_detach = p[1];
        upd.p_initial_detach_decay = p[2];
    }
    for (int c = 0; c < C; c++) {
        upd.channels[c].p_bleach = st[WP_FIT_CHANNEL(T, c)] / st[WP_FIT_CHANNEL(T, c) + 1];
        upd.channels[c].p_dud = st[WP_FIT_CHANNEL(T, c) + 2] / st[WP_FIT_CHANNEL(T, c) + 3];
    }
    // HMMFitter::update_with_holds (hmm-fitter.cc:63-97); bg_sig, mu and sig
    // are never fitted.
    wp_model next = *model;
    if (!h.hold_p_edman_failure) next.p_edman_failure = upd.p_edman_failure;
    if (!h.hold_p_detach) next.p_detach = upd.p_detach;
    if (!h.hold_p_initial_detach) next.p_initial_detach = upd.p_initial_detach;
    if (!h.hold_p_initial_detach_decay) next.p_initial_detach_decay = upd.p_initial_detach_decay;
    if (!h.hold_p_initial_block) next.p_initial_block = upd.p_initial_block;
    if (!h.hold_p_cyclic_block) next.p_cyclic_block = upd.p_cyclic_block;
    for (int c = 0; c < C; c++) {
        if (!h.hold_p_bleach[c]) next.channels[c].p_bleach = upd.channels[c].p_bleach;
        if (!h.hold_p_dud[c]) next.channels[c].p_dud = upd.channels[c].p_dud;
    }
    if (out_step_size) *out_step_size = model_distance(*model, next, C);
    *out_model = next;
    return WP_OK;
}

int wp_hmm_fitter_create(unsigned int num_timesteps, unsigned int num_channels, const char* dye_string,
                         int num_reads, const double* radiometries, int precision, wp_hmm_fitter** out) {
    if (!out || (num_reads > 0 && !radiometries)) {
        wp::set_error("null argument");
        return WP_ERR_ARG;
    }
    *out = nullptr;
    const int T = (int)num_timesteps, C = (int)num_channels;
    if (T < 1 || T > WP_MAX_TIMESTEPS || C < 1 || C > 4 || num_reads < 0) {
        wp::set_error("unsupported shape: need 1 <= T <= 64, 1 <= C <= 4, num_reads >= 0");
        return WP_ERR_UNSUPPORTED;
    }
    if (precision != WP_PRECISION_FP64 && precision != WP_PRECISION_FP32) {
        wp::set_error("unknown precision");
        return WP_ERR_ARG;
    }
    int rc = wp::require_device();
    if (rc) return rc;
    wp_hmm_fitter* f = new wp_hmm_fitter();
    f->T = T;
    f->C = C;
    f->N = num_reads;
    f->precision = precision;
    f->L = WP_FIT_STATS_LEN(T, C);
    rc = parse_fit_dye_seq(dye_string, C, &f->seq, &f->len, f->D);
    if (rc) {
        free_fitter(f);
        return rc;
    }
    // HMMFitter's max_num_dyes (hmm-fitter.cc:44-60)
    for (int c = 0; c < C; c++) f->M = std::max(f->M, f->D[c]);
    f->tri = (f->M + 1) * (f->M + 2) / 2;
    for (int c = 0; c < C; c++) f->NB *= f->D[c] + 1;
    // count coordinates of every q and, per channel, the fiber bases (q with
    // n_c = 0) in increasing order
    std::vector<unsigned char> crd((size_t)f->NB * C);
    std::vector<unsigned short> fbase;
    for (int q = 0; q < f->NB; q++) {
        int r = q;
        for (int c = C - 1; c >= 0; c--) {
            crd[(size_t)q * C + c] = (unsigned char)(r % (f->D[c] + 1));
            r /= f->D[c] + 1;
        }
    }
    for (int c = 0; c < C; c++)
        for (int q = 0; q < f->NB; q++)
            if (crd[(size_t)q * C + c] == 0) fbase.push_back((unsigned short)q);
    f->nfib = (int)fbase.size();
    std::vector<signed char> hseq(T);
    std::vector<unsigned char> track((size_t)T * C, 0);
    {
        std::vector<int> cs(C, 0);  // DyeTrack (dye-track.cc:26-43)
        for (int e = f->len - 1; e >= 0; e--) {
            if (f->seq[e] >= 0) cs[f->seq[e]]++;
            if (e < T)
                for (int c = 0; c < C; c++) track[(size_t)e * C + c] = (unsigned char)cs[c];
        }
        for (int e = 0; e < T; e++) hseq[e] = (signed char)(e < f->len ? f->seq[e] : -1);
    }
    // the row-lane kernel where the count shape has one (WP_FIT_KERNEL=generic
    // keeps the workgroup-per-read kernel, for comparisons)
    const char* kern = std::getenv("WP_FIT_KERNEL");
    f->rows = wp::fit_rows_supports(T, C, f->D) && !(kern && std::strcmp(kern, "generic") == 0);
    const size_t lds = f->rows ? wp::fit_rows_lds_bytes(precision, T, C, f->NB, f->tri)
                               : wp::fit_lds_bytes(precision, T, C, f->NB, f->tri, f->nfib);
    hipDeviceProp_t prop;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) {
        wp::set_error("cannot query the device");
        free_fitter(f);
        return WP_ERR_HIP;
    }
    if (f->NB > 65535 || lds > prop.sharedMemPerBlock) {
        wp::set_error("dye-seq state tensor too large for LDS at this precision");
        free_fitter(f);
        return WP_ERR_UNSUPPORTED;
    }
    // workgroups: as many as fit on the device at once, at most one per read
    const size_t real = precision == WP_PRECISION_FP32 ? 4 : 8;
    if (f->rows) {
        // as many workgroups of four wavefronts per CU as its register bound
        // keeps resident, one read per wavefront at a time
        const int waves = wp::fit_rows_threads() / 64;
        f->grid = std::max(1, std::min((num_reads + waves - 1) / waves,
                                       wp::fit_rows_groups_per_cu(precision) * prop.multiProcessorCount));
        f->hist_stride = wp::fit_rows_hist_per_group(T, f->NB);
    } else {
        int per_cu =
            std::max(1, std::min(8, (int)(prop.maxSharedMemoryPerMultiProcessor / std::max<size_t>(lds, 1))));
        f->grid = std::max(1, std::min(num_reads, per_cu * prop.multiProcessorCount));
        f->hist_stride = (size_t)T * 2 * T * f->NB;
    }
    rc = put(&f->d_rads, radiometries, (size_t)num_reads * T * C);
    if (!rc) rc = put(&f->d_crd, crd.data(), crd.size());
    if (!rc) rc = put(&f->d_fbase, fbase.data(), fbase.size());
    if (!rc) rc = put(&f->d_seq, hseq.data(), hseq.size());
    if (!rc) rc = put(&f->d_track, track.data(), track.size());
    if (!rc) rc = put<double>(&f->d_binom, nullptr, (size_t)2 * C * f->tri);
    if (!rc) rc = put<double>(&f->d_partial, nullptr, (size_t)f->grid * (f->L + 1));
    if (!rc) rc = put<double>(&f->d_out, nullptr, (size_t)f->L + 1);
    if (!rc && hipMalloc(&f->d_hist, real * f->hist_stride * f->grid) != hipSuccess) {
        wp::set_error("cannot allocate the forward history");
        rc = WP_ERR_NOMEM;
    }
    if (rc) {
        free_fitter(f);
        return rc;
    }
    *out = f;
    return WP_OK;
}

void wp_hmm_fitter_destroy(wp_hmm_fitter* f) {
    if (f) free_fitter(f);
}

int wp_hmm_fitter_estep(wp_hmm_fitter* f, const wp_model* model, double* out_stats, double* out_sum_log_p) {
    if (!f) {
        wp::set_error("null fitter");
        return WP_ERR_ARG;
    }
    return estep(f, model, out_stats, out_sum_log_p);
}

}  // extern "C"

namespace {

// The EM loop of HMMFitter::fit (hmm-fitter.h:55-113) over the fitter's reads,
// each counted d_weight[r] times (null: once).
int run_em(wp_hmm_fitter* f, const wp_model* model, const wp_fit_settings* settings, double stopping_threshold,
           double max_runtime, int max_iterations, const int* d_weight, wp_model* out_model, double* out_step_size,
           double* out_log_l, int* out_iterations) {
    std::string ds;
    for (int i = 0; i < f->len; i++) ds += f->seq[i] < 0 ? '.' : (char)('0' + f->seq[i]);
    std::vector<double> st((size_t)f->L);
    wp_model sm = *model, next;
    double step = 0.0;
    int it = 0;
    auto start = std::chrono::steady_clock::now();
    while (true) {
        int rc = estep(f, &sm, st.data(), nullptr, d_weight);
        if (rc) return rc;
        rc = wp_fit_mstep(f->T, f->C, ds.c_str(), &sm, f->N, st.data(), settings, &next, &step);
        if (rc) return rc;
        it++;
        double elapsed = std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
        if (step < stopping_threshold || elapsed > max_runtime || (max_iterations > 0 && it >= max_iterations))
            break;
        sm = next;
    }
    *out_model = next;
    if (out_step_size) *out_step_size = step;
    if (out_iterations) *out_iterations = it;
    double log_l = 0.0;  // HMMFitter::log_likelihood under the fitted model
    int rc = estep(f, &next, nullptr, &log_l, d_weight);
    if (rc) return rc;
    if (out_log_l) *out_log_l = log_l;
    return WP_OK;
}

// The EM loops of R bootstrap rounds in lockstep (each round as run_em with
// weight row i): one batched E-step launch per iteration over the rounds
// still running, then each round's M-step. Every round follows the same
// sequence of models as its own run_em, bit for bit; the wall-clock limit
// counts from the start of the batch.
int run_em_rounds(wp_hmm_fitter* f, const wp_model* model, const wp_fit_settings* settings,
                  double stopping_threshold, double max_runtime, int R, wp_model* out_models, double* out_steps,
                  double* out_log_ls) {
    std::string ds;
    for (int i = 0; i < f->len; i++) ds += f->seq[i] < 0 ? '.' : (char)('0' + f->seq[i]);
    std::vector<wp_model> sm(R, *model), next(R, *model);
    std::vector<double> step(R, 0.0), out;
    std::vector<char> done(R, 0);
    auto start = std::chrono::steady_clock::now();
    int left = R;
    while (left > 0) {
        std::vector<wp_model> ms;
        std::vector<int> rs;
        for (int i = 0; i < R; i++)
            if (!done[i]) {
                ms.push_back(sm[i]);
                rs.push_back(i);
            }
        int rc = estep_batch(f, ms, rs, &out);
        if (rc) return rc;
        const double elapsed = std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
        for (size_t k = 0; k < rs.size(); k++) {
            const int i = rs[k];
            rc = wp_fit_mstep(f->T, f->C, ds.c_str(), &sm[i], f->N, out.data() + k * (f->L + 1), settings, &next[i],
                              &step[i]);
            if (rc) return rc;
            if (step[i] < stopping_threshold || elapsed > max_runtime) {
                done[i] = 1;
                left--;
            } else {
                sm[i] = next[i];
            }
        }
    }
    std::vector<int> all(R);
    for (int i = 0; i < R; i++) all[i] = i;
    int rc = estep_batch(f, next, all, &out);  // HMMFitter::log_likelihood under each fitted model
    if (rc) return rc;
    for (int i = 0; i < R; i++) {
        out_models[i] = next[i];
        out_steps[i] = step[i];
        out_log_ls[i] = out[(size_t)i * (f->L + 1) + f->L];
    }
    return WP_OK;
}

// Order statistic i of one parameter over the rounds' models.
double order_stat(std::vector<double> v, size_t i) {
    std::sort(v.begin(), v.end());
    return v[std::min(i, v.size() - 1)];
}

}  // namespace

extern "C" {

int wp_hmm_fit(wp_hmm_fitter* f, const wp_model* model, const wp_fit_settings* settings,
               double stopping_threshold, double max_runtime, int max_iterations, wp_model* out_model,
               double* out_step_size, double* out_log_l, int* out_iterations) {
    if (!f || !model || !out_model) {
        wp::set_error("null argument");
        return WP_ERR_ARG;
    }
    return run_em(f, model, settings, stopping_threshold, max_runtime, max_iterations, nullptr, out_model,
                  out_step_size, out_log_l, out_iterations);
}

int wp_hmm_fitter_set_batched(wp_hmm_fitter* f, int enable) {
    if (!f) {
        wp::set_error("null fitter");
        return WP_ERR_ARG;
    }
    f->batched = enable ? 1 : 0;
    return WP_OK;
}

int wp_fit_bootstrap_counts(int num_reads, uint64_t seed, int round, int* out_counts) {
    if (num_reads < 0 || round < 0 || (num_reads > 0 && !out_counts)) {
        wp::set_error("bad bootstrap arguments");
        return WP_ERR_ARG;
    }
    std::fill(out_counts, out_counts + num_reads, 0);
    wp::Rng rng = wp::stream_rng(seed, (uint64_t)round);
    for (int j = 0; j < num_reads; j++) out_counts[rng.next() % (uint64_t)num_reads]++;
    return WP_OK;
}

int wp_hmm_fit_bootstrap(wp_hmm_fitter* f, const wp_model* model, const wp_fit_settings* settings,
                         double stopping_threshold, double max_runtime, int num_rounds,
                         double confidence_interval, uint64_t seed, wp_model* out_models, double* out_log_ls,
                         wp_model* out_best_model, double* out_step_size, double* out_log_l, wp_model* out_ci_min,
                         wp_model* out_ci_max, double* out_worst_step) {
    if (!f || !model || !out_best_model || num_rounds < 0 || (num_rounds > 0 && (!out_models || !out_log_ls)) ||
        !(confidence_interval >= 0.0 && confidence_interval <= 1.0)) {
        wp::set_error("bad bootstrap arguments");
        return WP_ERR_ARG;
    }
    // the fit on every read (bootstrap-fit.cc:72-75)
    int rc = run_em(f, model, settings, stopping_threshold, max_runtime, 0, nullptr, out_best_model, out_step_size,
                    out_log_l, nullptr);
    if (rc) return rc;
    if (num_rounds == 0) return WP_OK;
    // bootstrap-fit.cc:76-93: each round a resample of N reads with
    // replacement, as per-read multiplicities, all rounds resident
    if (num_rounds > f->weight_rows) {
        hipFree(f->d_weight);
        f->d_weight = nullptr;
        f->weight_rows = 0;
        WP_HIP(hipMalloc(&f->d_weight, sizeof(int) * (size_t)num_rounds * std::max(f->N, 1)));
        f->weight_rows = num_rounds;
    }
    std::vector<int> w((size_t)num_rounds * std::max(f->N, 1));
    for (int i = 0; i < num_rounds; i++) wp_fit_bootstrap_counts(f->N, seed, i, w.data() + (size_t)i * f->N);
    WP_HIP(hipMemcpy(f->d_weight, w.data(), sizeof(int) * (size_t)num_rounds * f->N, hipMemcpyHostToDevice));
    double worst = 0.0;
    std::vector<double> steps(num_rounds, 0.0);
    if (f->batched) {
        rc = run_em_rounds(f, model, settings, stopping_threshold, max_runtime, num_rounds, out_models, steps.data(),
                           out_log_ls);
        if (rc) return rc;
    } else {
        for (int i = 0; i < num_rounds; i++) {
            rc = run_em(f, model, settings, stopping_threshold, max_runtime, 0, f->d_weight + (size_t)i * f->N,
                        &out_models[i], &steps[i], &out_log_ls[i], nullptr);
            if (rc) return rc;
        }
    }
    for (double st : steps) worst = std::max(worst, st);
    if (out_worst_step) *out_worst_step = worst;
    // confidence bounds: per parameter, the order statistics at (1 - ci) / 2
    // and 1 - (1 - ci) / 2 of the rounds (bootstrap-fit.cc:98-167)
    const size_t lo = (size_t)((1.0 - confidence_interval) / 2.0 * (double)num_rounds);
    const size_t hi = (size_t)((1.0 - (1.0 - confidence_interval) / 2.0) * (double)num_rounds);
    wp_model mn = *out_best_model, mx = *out_best_model;
    auto bound = [&](auto get, auto set) {
        std::vector<double> v(num_rounds);
        for (int i = 0; i < num_rounds; i++) v[i] = get(out_models[i]);
        set(mn, order_stat(v, lo));
        set(mx, order_stat(v, hi));
    };
    bound([](const wp_model& m) { return m.p_edman_failure; }, [](wp_model& m, double x) { m.p_edman_failure = x; });
    bound([](const wp_model& m) { return m.p_detach; }, [](wp_model& m, double x) { m.p_detach = x; });
    bound([](const wp_model& m) { return m.p_initial_detach; }, [](wp_model& m, double x) { m.p_initial_detach = x; });
    bound([](const wp_model& m) { return m.p_initial_detach_decay; },
          [](wp_model& m, double x) { m.p_initial_detach_decay = x; });
    bound([](const wp_model& m) { return m.p_initial_block; }, [](wp_model& m, double x) { m.p_initial_block = x; });
    bound([](const wp_model& m) { return m.p_cyclic_block; }, [](wp_model& m, double x) { m.p_cyclic_block = x; });
    for (int c = 0; c < f->C; c++) {
        bound([c](const wp_model& m) { return m.channels[c].p_bleach; },
              [c](wp_model& m, double x) { m.channels[c].p_bleach = x; });
        bound([c](const wp_model& m) { return m.channels[c].p_dud; },
              [c](wp_model& m, double x) { m.channels[c].p_dud = x; });
    }
    if (out_ci_min) *out_ci_min = mn;
    if (out_ci_max) *out_ci_max = mx;
    return WP_OK;
}

}  // extern "C"
