/*
 * ORACLE — test infrastructure only (see oracle.h). FP64 CPU restatement of
 * whatprot's PeptideHMM forward scoring path. Every function cites the
 * reference file:line whose semantics it restates (paths relative to
 * whatprot's cc_code/src). Deliberately simple: dense tensors over the full
 * dye-seq shape, explicit per-step boxes, row-major loops in the reference's
 * iteration order.
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

int wpo_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ---- model ---------------------------------------------------------------- */

/* channel-model.h:35-55 */
double wpo_adjusted_mu(const wp_model* m, int channel, const unsigned int* counts) {
    const wp_channel_model* cm = &m->channels[channel];
    double amu = cm->mu;
    amu *= (double)counts[channel];
    for (int c = 0; c < m->num_channels; c++) {
        if (c == channel) {
            if (counts[c] > 1) {
                amu *= cm->flat_interactions[c];
                amu *= pow(cm->interactions[c], (double)(counts[c] - 1));
            }
        } else {
            if (counts[c] > 0) {
                amu *= cm->flat_interactions[c];
                amu *= pow(cm->interactions[c], (double)counts[c]);
            }
        }
    }
    return amu;
}

/* channel-model.cc:47-49 */
double wpo_sigma(const wp_model* m, int channel, double amu) {
    const wp_channel_model* cm = &m->channels[channel];
    return sqrt(cm->bg_sig * cm->bg_sig + amu * cm->sig * cm->sig);
}

/* channel-model.h:62-70 (pi literal as written there) */
double wpo_pdf(const wp_model* m, int channel, double observed, const unsigned int* counts) {
    static const double pi = 3.141592653589793238;
    double amu = wpo_adjusted_mu(m, channel, counts);
    double offset = observed - amu;
    double s = wpo_sigma(m, channel, amu);
    return (1.0 / (s * sqrt(2.0 * pi))) * exp(-offset * offset / (2.0 * s * s));
}

/* decaying-rate-model.cc:47-50 */
double wpo_p_detach(const wp_model* m, unsigned int i) {
    return m->p_detach + m->p_initial_detach * exp(-(double)i * m->p_initial_detach_decay);
}

/* binomial-transition.cc:51-75 */
void wpo_binomial_table(double q, unsigned int max_n, double* v) {
    v[0] = 1.0;
    double p = 1.0 - q;
    for (unsigned int i = 1; i <= max_n; i++) {
        unsigned int row = i * (i + 1) / 2, prev = (i - 1) * i / 2;
        v[row + 0] = v[prev + 0] * q;
        for (unsigned int j = 1; j < i; j++) {
            v[row + j] = v[prev + j] * q + v[prev + j - 1] * p;
        }
        v[row + i] = v[prev + i - 1] * p;
    }
}

/* peptide-emission.cc:49-164: the per-channel pruned range at timestep t. */
void wpo_emission_range(const wp_model* m, const double* rad, unsigned int C,
                        unsigned int t, unsigned int M, double cutoff,
                        unsigned int* lo, unsigned int* hi) {
    if (cutoff == DBL_MAX) {
        for (unsigned int c = 0; c < C; c++) {
            lo[c] = 0;
            hi[c] = M + 1;
        }
        return;
    }
    unsigned int counts_lo[WP_MAX_CHANNELS] = {0};
    unsigned int counts_hi[WP_MAX_CHANNELS] = {0};
    int did_change = 1;
    while (did_change) { /* fixed point, peptide-emission.cc:76-111 */
        did_change = 0;
        for (unsigned int c = 0; c < C; c++) {
            unsigned int old = counts_hi[c];
            counts_hi[c] = 1;
            double mu1 = wpo_adjusted_mu(m, (int)c, counts_hi);
            counts_hi[c] = old;
            double mu = wpo_adjusted_mu(m, (int)c, counts_hi);
            double s = cutoff * wpo_sigma(m, (int)c, mu);
            unsigned int new_cts = (unsigned int)lround((rad[t * C + c] + s) / mu1);
            if (new_cts > M + 1) new_cts = M + 1;
            if (new_cts != counts_hi[c]) {
                did_change = 1;
                counts_hi[c] = new_cts;
            }
        }
    }
    for (unsigned int c = 0; c < C; c++) { /* peptide-emission.cc:113-163 */
        double x = rad[t * C + c];
        unsigned int cmin = M + 1;
        unsigned int old_lo = counts_lo[c];
        for (unsigned int d = 0; d < M + 1; d++) {
            counts_lo[c] = d;
            double mu = wpo_adjusted_mu(m, (int)c, counts_lo);
            double s = cutoff * wpo_sigma(m, (int)c, mu);
            if (mu + s > x) {
                cmin = d;
                break;
            }
        }
        lo[c] = cmin;
        counts_lo[c] = old_lo;
        unsigned int cmax = M + 1;
        unsigned int old_hi = counts_hi[c];
        for (unsigned int d = cmin; d < M + 1; d++) {
            counts_hi[c] = d;
            double mu = wpo_adjusted_mu(m, (int)c, counts_hi);
            double s = cutoff * wpo_sigma(m, (int)c, mu);
            if (mu - s > x) {
                cmax = d;
                break;
            }
        }
        hi[c] = cmax;
        counts_hi[c] = old_hi;
    }
}

/* dye-seq.cc:23-40 */
int wpo_parse_dye_seq(const char* s, short* seq, int cap) {
    int len = (int)strlen(s);
    while (len > 0 && s[len - 1] == '.') len--;
    for (int i = 0; i < len && i < cap; i++) {
        char ch = s[i];
        if (ch >= '0' && ch <= '9') seq[i] = (short)(ch - '0');
        else seq[i] = -1;
    }
    return len;
}

/* dye-track.cc:26-43: track[t*C + c] = number of channel-c dyes at positions >= t. */
static void dye_track(unsigned int T, unsigned int C, const short* seq, int len, unsigned int* track) {
    unsigned int cs[WP_MAX_CHANNELS] = {0};
    memset(track, 0, sizeof(unsigned int) * T * C);
    for (int t = len - 1; t >= 0; t--) {
        if (seq[t] != -1) cs[seq[t]]++;
        if (t < (int)T) memcpy(&track[t * C], cs, sizeof(unsigned int) * C);
    }
}

/* ---- boxes (util/kd-range.cc:18-48) ---------------------------------------- */

typedef struct {
    unsigned int lo[WPO_MAXD], hi[WPO_MAXD];
} obox;

static obox box_intersect(const obox* a, const obox* b, int nd) {
    obox r;
    for (int i = 0; i < nd; i++) {
        r.lo[i] = a->lo[i] > b->lo[i] ? a->lo[i] : b->lo[i];
        r.hi[i] = a->hi[i] < b->hi[i] ? a->hi[i] : b->hi[i];
    }
    return r;
}

static int box_empty(const obox* b, int nd) {
    for (int i = 0; i < nd; i++)
        if (b->lo[i] >= b->hi[i]) return 1;
    return 0;
}

static int box_includes_zero(const obox* b, int nd) {
    for (int i = 0; i < nd; i++) {
        if (b->lo[i] > 0) return 0;
        if (b->hi[i] <= 0) return 0;
    }
    return 1;
}

/* Row-major iteration, last dimension fastest (tensor/base-tensor-iterator.h:46-92). */
static int box_next(const obox* b, int nd, unsigned int* loc) {
    for (int o = nd - 1; o >= 0; o--) {
        loc[o]++;
        if (loc[o] < b->hi[o]) return 1;
        loc[o] = b->lo[o];
    }
    return 0;
}

/* ---- steps (hmm/hmm/peptide-hmm.cc:40-61) --------------------------------- */

enum { ST_BLOCK_INITIAL, ST_DUD, ST_EMIT, ST_BLOCK_CYCLIC, ST_DETACH, ST_BLEACH, ST_EDMAN };

typedef struct {
    int kind, channel, t;
    obox a;  /* block/detach/emit: pruned range; binomial: forward_range; edman: true_forward */
    obox b;  /* binomial: backward_range; edman: safe_backward */
    obox tb; /* edman: true_backward */
    int allow;   /* emission allow_detached */
    int det_fwd; /* detach detached_forward */
} ostep;

typedef struct {
    unsigned int T, C, nd;
    unsigned int shape[WPO_MAXD];
    size_t strides[WPO_MAXD];
    size_t size;
    int nsteps;
    ostep* steps;
    obox own[WP_MAX_TIMESTEPS]; /* emission own ranges */
    obox forward_range, backward_range;
    int empty;
} oplan;

static void edman_set_true_forward(ostep* s, const obox* r, int nd) { /* edman-transition.cc:28-37 */
    s->a = *r;
    s->b = *r;
    s->b.hi[0]++;
    for (int c = 1; c < nd; c++)
        if (s->b.lo[c] != 0) s->b.lo[c]--;
}

/* Builds the step list and runs prune_forward / prune_backward
 * (peptide-hmm.cc:62-83 with each step's prune_* rules). */
static void plan_build(oplan* p, const wp_model* m, unsigned int T, unsigned int C,
                       const unsigned int* track, const double* rad,
                       unsigned int M, double cutoff) {
    int nd = (int)C + 1;
    p->T = T;
    p->C = C;
    p->nd = (unsigned int)nd;
    p->shape[0] = T;
    for (unsigned int c = 0; c < C; c++) p->shape[1 + c] = 1 + track[c];
    p->size = 1;
    for (int o = nd - 1; o >= 0; o--) {
        p->strides[o] = p->size;
        p->size *= p->shape[o];
    }
    p->nsteps = (int)((4 + C) * (T - 1) + 2 + C);
    p->steps = (ostep*)calloc((size_t)p->nsteps, sizeof(ostep));
    int k = 0;
    p->steps[k++].kind = ST_BLOCK_INITIAL;
    for (unsigned int c = 0; c < C; c++) {
        p->steps[k].kind = ST_DUD;
        p->steps[k++].channel = (int)c;
    }
    p->steps[k].kind = ST_EMIT;
    p->steps[k++].t = 0;
    for (unsigned int t = 1; t < T; t++) {
        p->steps[k++].kind = ST_BLOCK_CYCLIC;
        p->steps[k].kind = ST_DETACH;
        p->steps[k++].t = (int)t - 1;
        for (unsigned int c = 0; c < C; c++) {
            p->steps[k].kind = ST_BLEACH;
            p->steps[k++].channel = (int)c;
        }
        p->steps[k++].kind = ST_EDMAN;
        p->steps[k].kind = ST_EMIT;
        p->steps[k++].t = (int)t;
    }
    for (unsigned int t = 0; t < T; t++) {
        unsigned int lo[WP_MAX_CHANNELS], hi[WP_MAX_CHANNELS];
        wpo_emission_range(m, rad, C, t, M, cutoff, lo, hi);
        p->own[t].lo[0] = 0;
        p->own[t].hi[0] = t + 1;
        for (unsigned int c = 0; c < C; c++) {
            p->own[t].lo[1 + c] = lo[c];
            p->own[t].hi[1 + c] = hi[c];
        }
    }
    obox r;
    for (int o = 0; o < nd; o++) {
        r.lo[o] = 0;
        r.hi[o] = p->shape[o];
    }
    int allow = 0; /* uninitialized in the reference (peptide-hmm.cc:67); first written by Emit(0) */
    p->empty = 0;
    for (int i = 0; i < p->nsteps; i++) {
        ostep* s = &p->steps[i];
        switch (s->kind) {
        case ST_BLOCK_INITIAL:
        case ST_BLOCK_CYCLIC: /* block-transition.cc:22-24 */
            s->a = r;
            break;
        case ST_DUD:
        case ST_BLEACH: /* binomial-transition.cc:77-81 */
            s->a = r;
            r.lo[1 + s->channel] = 0;
            s->b = r;
            break;
        case ST_DETACH: /* detach-transition.cc:24-28 */
            s->a = r;
            s->det_fwd = allow;
            allow = 1;
            break;
        case ST_EDMAN: /* edman-transition.cc:50-53 */
            edman_set_true_forward(s, &r, nd);
            r = s->b;
            break;
        case ST_EMIT: /* peptide-emission.cc:198-203 */
            s->a = box_intersect(&p->own[s->t], &r, nd);
            r = s->a;
            s->allow = box_includes_zero(&r, nd);
            allow = s->allow;
            break;
        }
        if (box_empty(&r, nd)) {
            p->empty = 1;
            return;
        }
    }
    p->backward_range = r;
    for (int i = p->nsteps - 1; i >= 0; i--) {
        ostep* s = &p->steps[i];
        switch (s->kind) {
        case ST_BLOCK_INITIAL:
        case ST_BLOCK_CYCLIC: /* block-transition.cc:26-29 */
        case ST_DETACH:       /* detach-transition.cc:30-34 */
            s->a = box_intersect(&s->a, &r, nd);
            r = s->a;
            break;
        case ST_DUD:
        case ST_BLEACH: /* binomial-transition.cc:83-89 */
            s->b = box_intersect(&s->b, &r, nd);
            r = s->b;
            r.hi[1 + s->channel] = 0xffffffffu;
            s->a = box_intersect(&s->a, &r, nd);
            r = s->a;
            break;
        case ST_EDMAN: { /* edman-transition.cc:39-60 */
            r = box_intersect(&s->b, &r, nd);
            s->tb = r;
            obox sf = r;
            if (sf.lo[0] != 0) sf.lo[0]--;
            for (int c = 1; c < nd; c++) sf.hi[c]++;
            r = box_intersect(&sf, &s->a, nd);
            edman_set_true_forward(s, &r, nd);
            break;
        }
        case ST_EMIT: /* peptide-emission.cc:205-210 */
            s->a = box_intersect(&s->a, &r, nd);
            r = s->a;
            s->allow = box_includes_zero(&r, nd);
            allow = s->allow;
            break;
        }
        if (box_empty(&r, nd)) {
            p->empty = 1;
            return;
        }
    }
    p->forward_range = r;
}

int wpo_emission_boxes(const wp_model* m, unsigned int T, unsigned int C,
                       const short* seq, int len, const double* rad,
                       unsigned int M, double cutoff, unsigned int* lo, unsigned int* hi) {
    unsigned int track[WP_MAX_TIMESTEPS * WP_MAX_CHANNELS];
    dye_track(T, C, seq, len, track);
    oplan p;
    plan_build(&p, m, T, C, track, rad, M, cutoff);
    int ok = !p.empty;
    if (ok) {
        for (int i = 0; i < p.nsteps; i++) {
            const ostep* s = &p.steps[i];
            if (s->kind != ST_EMIT) continue;
            for (unsigned int c = 0; c < C; c++) {
                lo[s->t * C + c] = s->a.lo[1 + c];
                hi[s->t * C + c] = s->a.hi[1 + c];
            }
        }
    }
    free(p.steps);
    return ok;
}

/* ---- forward (generic-hmm.h:41-55 over the steps' forward operators) ------- */

static size_t flat(const oplan* p, const unsigned int* loc) {
    size_t i = 0;
    for (unsigned int o = 0; o < p->nd; o++) i += p->strides[o] * loc[o];
    return i;
}

double wpo_pair_probability(const wp_model* m, unsigned int T, unsigned int C,
                            const short* seq, int len, const double* rad,
                            unsigned int M, double cutoff) {
    unsigned int track[WP_MAX_TIMESTEPS * WP_MAX_CHANNELS];
    dye_track(T, C, seq, len, track);
    oplan p;
    plan_build(&p, m, T, C, track, rad, M, cutoff);
    if (p.empty) { /* peptide-hmm.cc:94-97 */
        free(p.steps);
        return 0.0;
    }
    int nd = (int)p.nd;
    size_t n = p.size;
    double* A = (double*)calloc(n, sizeof(double));
    double* B = (double*)calloc(n, sizeof(double));
    double* A2 = (double*)calloc(n, sizeof(double));
    double* B2 = (double*)calloc(n, sizeof(double));
    /* emission density table over [0, M+1)^C, filled on the own range only */
    size_t esz = 1;
    for (unsigned int c = 0; c < C; c++) esz *= (M + 1);
    double* E = (double*)malloc(esz * sizeof(double));
    double delta = 0.0;
    int allow = 1; /* PeptideStateVector ctor (peptide-state-vector.cc:33-36) */
    double qdud[WP_MAX_CHANNELS][(WP_MAX_DYES + 1) * (WP_MAX_DYES + 2) / 2];
    double qble[WP_MAX_CHANNELS][(WP_MAX_DYES + 1) * (WP_MAX_DYES + 2) / 2];
    for (unsigned int c = 0; c < C; c++) {
        wpo_binomial_table(m->channels[c].p_dud, M, qdud[c]);
        wpo_binomial_table(m->channels[c].p_bleach, M, qble[c]);
    }
    /* initialize_from_start: last element of the first Edman slice of the
     * forward_range tensor (peptide-state-vector.cc:38-40). */
    {
        unsigned int loc[WPO_MAXD];
        loc[0] = p.forward_range.lo[0];
        for (int o = 1; o < nd; o++) loc[o] = p.forward_range.hi[o] - 1;
        A[flat(&p, loc)] = 1.0;
    }
    unsigned int loc[WPO_MAXD];
    for (int si = 0; si < p.nsteps; si++) {
        const ostep* s = &p.steps[si];
        memset(A2, 0, n * sizeof(double));
        memset(B2, 0, n * sizeof(double));
        switch (s->kind) {
        case ST_BLOCK_INITIAL:
        case ST_BLOCK_CYCLIC: { /* block-transition.cc:31-60 */
            double pb = s->kind == ST_BLOCK_INITIAL ? m->p_initial_block : m->p_cyclic_block;
            memcpy(loc, s->a.lo, sizeof(loc));
            do {
                size_t i = flat(&p, loc);
                A2[i] = (1 - pb) * A[i];
                B2[i] = B[i] + pb * A[i];
            } while (box_next(&s->a, nd, loc));
            if (!allow) delta = 0.0;
            break;
        }
        case ST_DUD:
        case ST_BLEACH: { /* binomial-transition.cc:91-136 */
            int c = s->channel;
            const double* P = s->kind == ST_DUD ? qdud[c] : qble[c];
            obox fib = s->b; /* iterate fibers: fix dim 1+c at its low end */
            fib.hi[1 + c] = fib.lo[1 + c] + 1;
            memcpy(loc, fib.lo, sizeof(loc));
            do {
                unsigned int l2[WPO_MAXD];
                memcpy(l2, loc, sizeof(l2));
                for (unsigned int to = s->b.lo[1 + c]; to < s->b.hi[1 + c]; to++) {
                    double va = 0.0, vb = 0.0;
                    unsigned int from0 = to > s->a.lo[1 + c] ? to : s->a.lo[1 + c];
                    for (unsigned int from = from0; from < s->a.hi[1 + c]; from++) {
                        l2[1 + c] = from;
                        size_t i = flat(&p, l2);
                        double pr = P[from * (from + 1) / 2 + to];
                        va += pr * A[i];
                        vb += pr * B[i];
                    }
                    l2[1 + c] = to;
                    size_t o = flat(&p, l2);
                    A2[o] = va;
                    B2[o] = vb;
                }
            } while (box_next(&fib, nd, loc));
            if (!allow) delta = 0.0;
            break;
        }
        case ST_DETACH: { /* detach-transition.cc:36-68 */
            double pd = wpo_p_detach(m, (unsigned int)s->t);
            double sa = 0.0, sb = 0.0;
            memcpy(loc, s->a.lo, sizeof(loc));
            do {
                size_t i = flat(&p, loc);
                A2[i] = A[i] * (1 - pd);
                sa += A[i];
            } while (box_next(&s->a, nd, loc));
            memcpy(loc, s->a.lo, sizeof(loc));
            do {
                size_t i = flat(&p, loc);
                B2[i] = B[i] * (1 - pd);
                sb += B[i];
            } while (box_next(&s->a, nd, loc));
            double sum = sa + sb;
            /* detached_backward is always true (detach-transition.cc:33) */
            delta = s->det_fwd ? delta + pd * sum : pd * sum;
            allow = 1;
            break;
        }
        case ST_EDMAN: { /* edman-transition.cc:62-162 */
            double pf = m->p_edman_failure;
            size_t ts = p.strides[0];
            memcpy(loc, s->a.lo, sizeof(loc));
            do {
                size_t i = flat(&p, loc);
                double f = A[i];
                unsigned int e = loc[0];
                int c = (int)e < len ? seq[e] : -1;
                A2[i] += pf * f;
                if (c == -1) {
                    A2[i + ts] += (1 - pf) * f;
                } else {
                    unsigned int k = loc[1 + c];
                    unsigned int K = track[e * C + (unsigned int)c];
                    double ratio = (double)k / (double)K;
                    size_t cs = p.strides[1 + c];
                    if (k < K) A2[i + ts] += (1 - pf) * (1 - ratio) * f;
                    if (k > 0) A2[i + ts - cs] += (1 - pf) * ratio * f;
                }
            } while (box_next(&s->a, nd, loc));
            memcpy(loc, s->a.lo, sizeof(loc));
            do {
                size_t i = flat(&p, loc);
                B2[i] = B[i];
            } while (box_next(&s->a, nd, loc));
            if (!allow) delta = 0.0;
            break;
        }
        case ST_EMIT: { /* peptide-emission.cc:39-181, 212-243 */
            unsigned int t = (unsigned int)s->t;
            const obox* own = &p.own[t];
            unsigned int n[WP_MAX_CHANNELS];
            obox tr; /* own range without the Edman axis */
            for (unsigned int c = 0; c < C; c++) {
                tr.lo[c] = own->lo[1 + c];
                tr.hi[c] = own->hi[1 + c];
            }
            memcpy(n, tr.lo, sizeof(n));
            do {
                size_t ei = 0;
                for (unsigned int c = 0; c < C; c++) ei = ei * (M + 1) + n[c];
                double v = 1.0;
                for (unsigned int c = 0; c < C; c++) v *= wpo_pdf(m, (int)c, rad[t * C + c], n);
                E[ei] = v;
            } while (box_next(&tr, (int)C, n));
            memcpy(loc, s->a.lo, sizeof(loc));
            do {
                size_t i = flat(&p, loc);
                size_t ei = 0;
                for (unsigned int c = 0; c < C; c++) ei = ei * (M + 1) + loc[1 + c];
                A2[i] = A[i] * E[ei];
                B2[i] = B[i] * E[ei];
            } while (box_next(&s->a, nd, loc));
            if (s->allow) delta = delta * E[0];
            else delta = 0.0;
            allow = s->allow;
            break;
        }
        }
        double* tmp = A;
        A = A2;
        A2 = tmp;
        tmp = B;
        B = B2;
        B2 = tmp;
    }
    /* PeptideStateVector::sum over the last range (peptide-state-vector.cc:52-54) */
    const obox* fin = &p.steps[p.nsteps - 1].a;
    double sa = 0.0, sb = 0.0;
    memcpy(loc, fin->lo, sizeof(loc));
    do {
        sa += A[flat(&p, loc)];
    } while (box_next(fin, nd, loc));
    memcpy(loc, fin->lo, sizeof(loc));
    do {
        sb += B[flat(&p, loc)];
    } while (box_next(fin, nd, loc));
    double result = sa + sb + delta;
    free(A);
    free(B);
    free(A2);
    free(B2);
    free(E);
    free(p.steps);
    return result;
}

/* ---- classifier (classifiers/hmm-classifier.h:45-78) ---------------------- */

static unsigned int max_num_dyes(unsigned int C, int nds, const short* seqs, int cap, const int* lens) {
    unsigned int M = 0; /* hmm-classifier.cc:47-65 */
    for (int i = 0; i < nds; i++) {
        unsigned int cnt[WP_MAX_CHANNELS] = {0};
        for (int k = 0; k < lens[i]; k++)
            if (seqs[(size_t)i * cap + k] >= 0) cnt[seqs[(size_t)i * cap + k]]++;
        for (unsigned int c = 0; c < C; c++)
            if (cnt[c] > M) M = cnt[c];
    }
    return M;
}

void wpo_classify(const wp_model* m, unsigned int T, unsigned int C,
                  int nds, const short* seqs, int cap, const int* lens,
                  const int* counts, const int* ids, int nr, const double* rads,
                  double cutoff, const int* cand_offsets, const int* cand_indices,
                  int* out_id, double* out_score, double* out_total) {
    unsigned int M = max_num_dyes(C, nds, seqs, cap, lens);
#pragma omp parallel for schedule(dynamic, 1)
    for (int r = 0; r < nr; r++) {
        const double* rad = rads + (size_t)r * T * C;
        int best_i = -1;
        double best = -1.0, total = 0.0;
        int n = cand_offsets ? cand_offsets[r + 1] - cand_offsets[r] : nds;
        for (int k = 0; k < n; k++) {
            int i = cand_offsets ? cand_indices[cand_offsets[r] + k] : k;
            double s = wpo_pair_probability(m, T, C, seqs + (size_t)i * cap, lens[i], rad, M, cutoff);
            total += s * counts[i];
            if (s > best) {
                best = s;
                best_i = i;
            }
        }
        out_id[r] = best_i >= 0 ? ids[best_i] : -1;
        out_score[r] = best;
        out_total[r] = total;
        if (isnan(best / total)) { /* hmm-classifier.h:73-76 */
            out_score[r] = 0.0;
            out_total[r] = 1.0;
        }
    }
}

void wpo_pair_probabilities(const wp_model* m, unsigned int T, unsigned int C,
                            int nds, const short* seqs, int cap, const int* lens,
                            int nr, const double* rads, double cutoff, double* out_p) {
    unsigned int M = max_num_dyes(C, nds, seqs, cap, lens);
#pragma omp parallel for schedule(dynamic, 1)
    for (int r = 0; r < nr; r++) {
        for (int i = 0; i < nds; i++) {
            out_p[(size_t)r * nds + i] = wpo_pair_probability(
                    m, T, C, seqs + (size_t)i * cap, lens[i], rads + (size_t)r * T * C, M, cutoff);
        }
    }
}

/* ---- brute force over explicit dye fates (generate-dye-track.cc:29-110) ---- */

#define BF_MAXDYES 8

typedef struct {
    double w[1 << BF_MAXDYES][WP_MAX_TIMESTEPS + 1][2]; /* [live mask][e][blocked] */
} bf_state;

static double bf_emit(const wp_model* m, unsigned int C, const double* rad, unsigned int t,
                      unsigned int mask, const int* dye_ch, int ndyes) {
    unsigned int counts[WP_MAX_CHANNELS] = {0};
    for (int d = 0; d < ndyes; d++)
        if (mask & (1u << d)) counts[dye_ch[d]]++;
    double v = 1.0;
    for (unsigned int c = 0; c < C; c++) v *= wpo_pdf(m, (int)c, rad[t * C + c], counts);
    return v;
}

double wpo_brute_force_probability(const wp_model* m, unsigned int T, unsigned int C,
                                   const short* seq, int len, const double* rad) {
    int dye_pos[BF_MAXDYES], dye_ch[BF_MAXDYES], ndyes = 0;
    for (int i = 0; i < len; i++) {
        if (seq[i] >= 0) {
            if (ndyes == BF_MAXDYES) return NAN;
            dye_pos[ndyes] = i;
            dye_ch[ndyes] = seq[i];
            ndyes++;
        }
    }
    if (T > WP_MAX_TIMESTEPS) return NAN;
    unsigned int nm = 1u << ndyes;
    bf_state* cur = (bf_state*)calloc(1, sizeof(bf_state));
    bf_state* nxt = (bf_state*)calloc(1, sizeof(bf_state));
    /* duds, then the initial block */
    for (unsigned int mask = 0; mask < nm; mask++) {
        double pr = 1.0;
        for (int d = 0; d < ndyes; d++) {
            double q = m->channels[dye_ch[d]].p_dud;
            pr *= (mask & (1u << d)) ? (1 - q) : q;
        }
        cur->w[mask][0][0] += pr * (1 - m->p_initial_block);
        cur->w[mask][0][1] += pr * m->p_initial_block;
    }
    double detached = 0.0;
    /* record t = 0 */
    for (unsigned int mask = 0; mask < nm; mask++) {
        double em = bf_emit(m, C, rad, 0, mask, dye_ch, ndyes);
        cur->w[mask][0][0] *= em;
        cur->w[mask][0][1] *= em;
    }
    unsigned int zero_mask = 0;
    for (unsigned int t = 0; t + 1 < T; t++) {
        memset(nxt, 0, sizeof(bf_state));
        double pdet = wpo_p_detach(m, t);
        double det_new = detached;
        for (unsigned int mask = 0; mask < nm; mask++) {
            for (unsigned int e = 0; e <= t; e++) {
                for (int blk = 0; blk < 2; blk++) {
                    double w = cur->w[mask][e][blk];
                    if (w == 0.0) continue;
                    if ((int)e >= len) { /* fully cleaved: records zeros, no events */
                        nxt->w[mask][e][blk] += w;
                        continue;
                    }
                    for (int cb = 0; cb < 2; cb++) {
                        double pcb = cb ? m->p_cyclic_block : 1 - m->p_cyclic_block;
                        int b2 = blk || cb;
                        double w1 = w * pcb;
                        det_new += w1 * pdet; /* detach: zeros from here on */
                        w1 *= (1 - pdet);
                        for (int ed = 0; ed < 2; ed++) {
                            double ped;
                            unsigned int e2 = e, m2 = mask;
                            if (b2) {
                                if (ed) continue;
                                ped = 1.0;
                            } else if (ed) {
                                ped = 1 - m->p_edman_failure;
                                e2 = e + 1;
                                for (int d = 0; d < ndyes; d++)
                                    if (dye_pos[d] == (int)e) m2 &= ~(1u << d);
                            } else {
                                ped = m->p_edman_failure;
                            }
                            /* bleach every live dye (all live dyes sit at positions >= e2) */
                            for (unsigned int keep = 0; keep < nm; keep++) {
                                if ((keep & ~m2) != 0) continue;
                                double pb = 1.0;
                                for (int d = 0; d < ndyes; d++) {
                                    if (!(m2 & (1u << d))) continue;
                                    double q = m->channels[dye_ch[d]].p_bleach;
                                    pb *= (keep & (1u << d)) ? (1 - q) : q;
                                }
                                nxt->w[keep][e2][b2] += w1 * ped * pb;
                            }
                        }
                    }
                }
            }
        }
        detached = det_new * bf_emit(m, C, rad, t + 1, zero_mask, dye_ch, ndyes);
        for (unsigned int mask = 0; mask < nm; mask++) {
            double em = bf_emit(m, C, rad, t + 1, mask, dye_ch, ndyes);
            for (unsigned int e = 0; e <= t + 1; e++) {
                nxt->w[mask][e][0] *= em;
                nxt->w[mask][e][1] *= em;
            }
        }
        bf_state* tmp = cur;
        cur = nxt;
        nxt = tmp;
    }
    double total = detached;
    for (unsigned int mask = 0; mask < nm; mask++)
        for (unsigned int e = 0; e < T; e++) total += cur->w[mask][e][0] + cur->w[mask][e][1];
    free(cur);
    free(nxt);
    return total;
}
