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
    int det_bwd; /* detach detached_backward */
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
            s->a = box_intersect(&s->a, &r, nd);
            r = s->a;
            break;
        case ST_DETACH: /* detach-transition.cc:30-34 */
            s->a = box_intersect(&s->a, &r, nd);
            r = s->a;
            s->det_bwd = allow;
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

/* ---- step operators over dense tensors ------------------------------------ */

static size_t flat(const oplan* p, const unsigned int* loc) {
    size_t i = 0;
    for (unsigned int o = 0; o < p->nd; o++) i += p->strides[o] * loc[o];
    return i;
}

/* A PeptideStateVector (peptide-state-vector.cc:20-58): A = tensor,
 * B = broken_n_tensor, delta = p_detached, allow = allow_detached. */
typedef struct {
    double *A, *B;
    double delta;
    int allow;
} ostate;

#define WPO_TRI ((WP_MAX_DYES + 1) * (WP_MAX_DYES + 2) / 2)

/* Everything one (dye-seq, radiometry) pair needs besides its plan. */
typedef struct {
    const wp_model* m;
    const oplan* p;
    const short* seq;
    int len;
    const unsigned int* track;
    unsigned int C, M;
    size_t esz;  /* (M+1)^C */
    double* E;   /* [T][esz] emission densities on each timestep's own range */
    double qdud[WP_MAX_CHANNELS][WPO_TRI];
    double qble[WP_MAX_CHANNELS][WPO_TRI];
} octx;

static void ctx_init(octx* x, const wp_model* m, const oplan* p, const short* seq, int len,
                     const unsigned int* track, const double* rad, unsigned int M) {
    unsigned int C = p->C;
    x->m = m;
    x->p = p;
    x->seq = seq;
    x->len = len;
    x->track = track;
    x->C = C;
    x->M = M;
    x->esz = 1;
    for (unsigned int c = 0; c < C; c++) x->esz *= (M + 1);
    x->E = (double*)calloc(x->esz * p->T, sizeof(double));
    for (unsigned int c = 0; c < C; c++) { /* universal-precomputations.cc:23-59 */
        wpo_binomial_table(m->channels[c].p_dud, M, x->qdud[c]);
        wpo_binomial_table(m->channels[c].p_bleach, M, x->qble[c]);
    }
    /* PeptideEmission's tensor over its own range (peptide-emission.cc:165-181) */
    for (unsigned int t = 0; t < p->T; t++) {
        const obox* own = &p->own[t];
        unsigned int n[WP_MAX_CHANNELS];
        obox tr;
        for (unsigned int c = 0; c < C; c++) {
            tr.lo[c] = own->lo[1 + c];
            tr.hi[c] = own->hi[1 + c];
        }
        if (box_empty(&tr, (int)C)) continue;
        memcpy(n, tr.lo, sizeof(n));
        do {
            size_t ei = 0;
            for (unsigned int c = 0; c < C; c++) ei = ei * (M + 1) + n[c];
            double v = 1.0;
            for (unsigned int c = 0; c < C; c++) v *= wpo_pdf(m, (int)c, rad[t * C + c], n);
            x->E[t * x->esz + ei] = v;
        } while (box_next(&tr, (int)C, n));
    }
}

static double emis(const octx* x, unsigned int t, const unsigned int* loc) {
    size_t ei = 0;
    for (unsigned int c = 0; c < x->C; c++) ei = ei * (x->M + 1) + loc[1 + c];
    return x->E[t * x->esz + ei];
}

static int dye_at(const octx* x, unsigned int e) { /* dye-seq.cc:52-58 */
    return (int)e < x->len ? x->seq[e] : -1;
}

/* One step forward: out = step(in). out's tensors arrive zeroed. */
static void fwd_step(const octx* x, const ostep* s, const ostate* in, ostate* out) {
    const oplan* p = x->p;
    const wp_model* m = x->m;
    int nd = (int)p->nd;
    unsigned int loc[WPO_MAXD];
    out->allow = in->allow;
    out->delta = in->allow ? in->delta : 0.0;
    switch (s->kind) {
    case ST_BLOCK_INITIAL:
    case ST_BLOCK_CYCLIC: { /* block-transition.cc:31-60 */
        double pb = s->kind == ST_BLOCK_INITIAL ? m->p_initial_block : m->p_cyclic_block;
        memcpy(loc, s->a.lo, sizeof(loc));
        do {
            size_t i = flat(p, loc);
            out->A[i] = (1 - pb) * in->A[i];
            out->B[i] = in->B[i] + pb * in->A[i];
        } while (box_next(&s->a, nd, loc));
        break;
    }
    case ST_DUD:
    case ST_BLEACH: { /* binomial-transition.cc:91-136 */
        int c = s->channel;
        const double* P = s->kind == ST_DUD ? x->qdud[c] : x->qble[c];
        obox fib = s->b; /* fibers: dim 1+c fixed at its low end */
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
                    size_t i = flat(p, l2);
                    double pr = P[from * (from + 1) / 2 + to];
                    va += pr * in->A[i];
                    vb += pr * in->B[i];
                }
                l2[1 + c] = to;
                size_t o = flat(p, l2);
                out->A[o] = va;
                out->B[o] = vb;
            }
        } while (box_next(&fib, nd, loc));
        break;
    }
    case ST_DETACH: { /* detach-transition.cc:36-68 */
        double pd = wpo_p_detach(m, (unsigned int)s->t);
        double sa = 0.0, sb = 0.0;
        memcpy(loc, s->a.lo, sizeof(loc));
        do {
            size_t i = flat(p, loc);
            out->A[i] = in->A[i] * (1 - pd);
            sa += in->A[i];
        } while (box_next(&s->a, nd, loc));
        memcpy(loc, s->a.lo, sizeof(loc));
        do {
            size_t i = flat(p, loc);
            out->B[i] = in->B[i] * (1 - pd);
            sb += in->B[i];
        } while (box_next(&s->a, nd, loc));
        double sum = sa + sb;
        out->delta = 0.0;
        if (s->det_bwd) out->delta = s->det_fwd ? in->delta + pd * sum : pd * sum;
        out->allow = s->det_bwd;
        break;
    }
    case ST_EDMAN: { /* edman-transition.cc:62-162 */
        double pf = m->p_edman_failure;
        size_t ts = p->strides[0];
        memcpy(loc, s->a.lo, sizeof(loc));
        do {
            size_t i = flat(p, loc);
            double f = in->A[i];
            unsigned int e = loc[0];
            int c = dye_at(x, e);
            out->A[i] += pf * f;
            if (c == -1) {
                out->A[i + ts] += (1 - pf) * f;
            } else {
                unsigned int k = loc[1 + c];
                unsigned int K = x->track[e * x->C + (unsigned int)c];
                double ratio = (double)k / (double)K;
                size_t cs = p->strides[1 + c];
                if (k < K) out->A[i + ts] += (1 - pf) * (1 - ratio) * f;
                if (k > 0) out->A[i + ts - cs] += (1 - pf) * ratio * f;
            }
        } while (box_next(&s->a, nd, loc));
        memcpy(loc, s->a.lo, sizeof(loc));
        do {
            size_t i = flat(p, loc);
            out->B[i] = in->B[i];
        } while (box_next(&s->a, nd, loc));
        break;
    }
    case ST_EMIT: { /* peptide-emission.cc:212-243 */
        unsigned int t = (unsigned int)s->t;
        memcpy(loc, s->a.lo, sizeof(loc));
        do {
            size_t i = flat(p, loc);
            double e = emis(x, t, loc);
            out->A[i] = in->A[i] * e;
            out->B[i] = in->B[i] * e;
        } while (box_next(&s->a, nd, loc));
        unsigned int zero[WPO_MAXD] = {0};
        out->delta = s->allow ? in->delta * emis(x, t, zero) : 0.0;
        out->allow = s->allow;
        break;
    }
    }
}

/* One step backward: out = step.backward(in). out's tensors arrive zeroed. */
static void bwd_step(const octx* x, const ostep* s, const ostate* in, ostate* out) {
    const oplan* p = x->p;
    const wp_model* m = x->m;
    int nd = (int)p->nd;
    unsigned int loc[WPO_MAXD];
    out->allow = in->allow;
    out->delta = in->allow ? in->delta : 0.0;
    switch (s->kind) {
    case ST_BLOCK_INITIAL:
    case ST_BLOCK_CYCLIC: { /* block-transition.cc:62-91 */
        double pb = s->kind == ST_BLOCK_INITIAL ? m->p_initial_block : m->p_cyclic_block;
        memcpy(loc, s->a.lo, sizeof(loc));
        do {
            size_t i = flat(p, loc);
            out->A[i] = pb * in->B[i] + (1 - pb) * in->A[i];
            out->B[i] = in->B[i];
        } while (box_next(&s->a, nd, loc));
        break;
    }
    case ST_DUD:
    case ST_BLEACH: { /* binomial-transition.cc:138-183 */
        int c = s->channel;
        const double* P = s->kind == ST_DUD ? x->qdud[c] : x->qble[c];
        obox fib = s->a;
        fib.hi[1 + c] = fib.lo[1 + c] + 1;
        memcpy(loc, fib.lo, sizeof(loc));
        do {
            unsigned int l2[WPO_MAXD];
            memcpy(l2, loc, sizeof(l2));
            for (int from = (int)s->a.hi[1 + c] - 1; from >= (int)s->a.lo[1 + c]; from--) {
                double va = 0.0, vb = 0.0;
                int to_min = (int)s->b.lo[1 + c];
                int to_max = from < (int)s->b.hi[1 + c] - 1 ? from : (int)s->b.hi[1 + c] - 1;
                for (int to = to_min; to <= to_max; to++) {
                    l2[1 + c] = (unsigned int)to;
                    size_t i = flat(p, l2);
                    double pr = P[from * (from + 1) / 2 + to];
                    va += pr * in->A[i];
                    vb += pr * in->B[i];
                }
                l2[1 + c] = (unsigned int)from;
                size_t o = flat(p, l2);
                out->A[o] = va;
                out->B[o] = vb;
            }
        } while (box_next(&fib, nd, loc));
        break;
    }
    case ST_DETACH: { /* detach-transition.cc:70-103 */
        double pd = wpo_p_detach(m, (unsigned int)s->t);
        memcpy(loc, s->a.lo, sizeof(loc));
        do {
            size_t i = flat(p, loc);
            out->A[i] = (1 - pd) * in->A[i];
            if (s->det_bwd) out->A[i] += pd * in->delta;
        } while (box_next(&s->a, nd, loc));
        memcpy(loc, s->a.lo, sizeof(loc));
        do {
            size_t i = flat(p, loc);
            out->B[i] = (1 - pd) * in->B[i];
            if (s->det_bwd) out->B[i] += pd * in->delta;
        } while (box_next(&s->a, nd, loc));
        out->delta = (s->det_fwd && s->det_bwd) ? in->delta : 0.0;
        out->allow = s->det_fwd;
        break;
    }
    case ST_EDMAN: { /* edman-transition.cc:164-277 */
        double pf = m->p_edman_failure;
        size_t ts = p->strides[0];
        memcpy(loc, s->tb.lo, sizeof(loc));
        do {
            size_t i = flat(p, loc);
            double f = in->A[i];
            unsigned int e = loc[0];
            out->A[i] += pf * f;
            if (e > 0) {
                int c = dye_at(x, e - 1);
                if (c == -1) {
                    out->A[i - ts] += (1 - pf) * f;
                } else {
                    unsigned int K = x->track[(e - 1) * x->C + (unsigned int)c];
                    size_t cs = p->strides[1 + c];
                    unsigned int k = loc[1 + c];
                    if (k < K) out->A[i - ts] += (1 - pf) * (1 - (double)k / (double)K) * f;
                    /* the reference's tensor here spans safe_forward_range, one
                     * wider than the dye-seq's shape; entries past the shape
                     * are never read again, so they are not kept */
                    if (k + 1 < p->shape[1 + c])
                        out->A[i - ts + cs] += (1 - pf) * ((double)(k + 1) / (double)K) * f;
                }
            }
        } while (box_next(&s->tb, nd, loc));
        memcpy(loc, s->tb.lo, sizeof(loc));
        do {
            size_t i = flat(p, loc);
            out->B[i] = in->B[i];
        } while (box_next(&s->tb, nd, loc));
        break;
    }
    case ST_EMIT: { /* peptide-emission.cc:212-248 */
        unsigned int t = (unsigned int)s->t;
        memcpy(loc, s->a.lo, sizeof(loc));
        do {
            size_t i = flat(p, loc);
            double e = emis(x, t, loc);
            out->A[i] = in->A[i] * e;
            out->B[i] = in->B[i] * e;
        } while (box_next(&s->a, nd, loc));
        unsigned int zero[WPO_MAXD] = {0};
        out->delta = s->allow ? in->delta * emis(x, t, zero) : 0.0;
        out->allow = s->allow;
        break;
    }
    }
}

/* Source entry: the last element of the first Edman slice of a tensor over
 * forward_range (peptide-state-vector.cc:38-40, 63-65). */
static size_t source_index(const oplan* p) {
    unsigned int loc[WPO_MAXD];
    loc[0] = p->forward_range.lo[0];
    for (unsigned int o = 1; o < p->nd; o++) loc[o] = p->forward_range.hi[o] - 1;
    return flat(p, loc);
}

static void state_alloc(ostate* s, size_t n) {
    s->A = (double*)calloc(n, sizeof(double));
    s->B = (double*)calloc(n, sizeof(double));
    s->delta = 0.0;
    s->allow = 1; /* PeptideStateVector ctor (peptide-state-vector.cc:33-36) */
}

static void state_clear(ostate* s, size_t n) {
    memset(s->A, 0, n * sizeof(double));
    memset(s->B, 0, n * sizeof(double));
}

static void state_free(ostate* s) {
    free(s->A);
    free(s->B);
}

static int pair_setup(oplan* p, octx* x, const wp_model* m, unsigned int T, unsigned int C,
                      const short* seq, int len, unsigned int* track, const double* rad,
                      unsigned int M, double cutoff) {
    dye_track(T, C, seq, len, track);
    plan_build(p, m, T, C, track, rad, M, cutoff);
    if (p->empty) return 0;
    ctx_init(x, m, p, seq, len, track, rad, M);
    return 1;
}

/* PeptideHMM::probability (peptide-hmm.cc:94-100, generic-hmm.h:41-55). */
double wpo_pair_probability(const wp_model* m, unsigned int T, unsigned int C,
                            const short* seq, int len, const double* rad,
                            unsigned int M, double cutoff) {
    unsigned int track[WP_MAX_TIMESTEPS * WP_MAX_CHANNELS];
    oplan p;
    octx x;
    if (!pair_setup(&p, &x, m, T, C, seq, len, track, rad, M, cutoff)) {
        free(p.steps);
        return 0.0;
    }
    ostate a, b;
    state_alloc(&a, p.size);
    state_alloc(&b, p.size);
    a.A[source_index(&p)] = 1.0;
    for (int si = 0; si < p.nsteps; si++) {
        state_clear(&b, p.size);
        fwd_step(&x, &p.steps[si], &a, &b);
        ostate tmp = a;
        a = b;
        b = tmp;
    }
    /* PeptideStateVector::sum over the last range (peptide-state-vector.cc:52-54) */
    const obox* fin = &p.steps[p.nsteps - 1].a;
    unsigned int loc[WPO_MAXD];
    double sa = 0.0, sb = 0.0;
    memcpy(loc, fin->lo, sizeof(loc));
    do {
        sa += a.A[flat(&p, loc)];
    } while (box_next(fin, (int)p.nd, loc));
    memcpy(loc, fin->lo, sizeof(loc));
    do {
        sb += a.B[flat(&p, loc)];
    } while (box_next(fin, (int)p.nd, loc));
    double result = sa + sb + a.delta;
    state_free(&a);
    state_free(&b);
    free(x.E);
    free(p.steps);
    return result;
}

/* ---- Baum-Welch E-step (generic-hmm.h:60-109) ----------------------------- */

/* Per-step improve_fit into the flat statistics array (layout of
 * WP_FIT_STATS_LEN in whatprot_amd.h). f = forward state entering the step,
 * b = backward state entering it, nb = backward state leaving it. */
static void fit_step(const octx* x, const ostep* s, const ostate* f, const ostate* b,
                     const ostate* nb, double P, double* st) {
    const oplan* p = x->p;
    const wp_model* m = x->m;
    unsigned int T = p->T;
    int nd = (int)p->nd;
    unsigned int loc[WPO_MAXD];
    switch (s->kind) {
    case ST_BLOCK_INITIAL:
    case ST_BLOCK_CYCLIC: { /* block-transition.cc:93-116 */
        double pb = s->kind == ST_BLOCK_INITIAL ? m->p_initial_block : m->p_cyclic_block;
        double* acc = s->kind == ST_BLOCK_INITIAL ? &st[WP_FIT_P_INITIAL_BLOCK]
                                                  : &st[WP_FIT_P_CYCLIC_BLOCK];
        memcpy(loc, s->a.lo, sizeof(loc));
        do {
            size_t i = flat(p, loc);
            acc[0] += f->A[i] * pb * nb->B[i] / P;
            acc[1] += f->A[i] * b->A[i] / P;
        } while (box_next(&s->a, nd, loc));
        break;
    }
    case ST_DUD:
    case ST_BLEACH: { /* binomial-transition.cc:185-256, dud/bleach-transition.cc:25-37 */
        int c = s->channel;
        const double* Pq = s->kind == ST_DUD ? x->qdud[c] : x->qble[c];
        double* acc = &st[WP_FIT_CHANNEL(T, c) + (s->kind == ST_DUD ? 2 : 0)];
        for (int tensor = 0; tensor < 2; tensor++) {
            const double* fv = tensor ? f->B : f->A;
            const double* bv = tensor ? b->B : b->A;
            const double* nbv = tensor ? nb->B : nb->A;
            obox fib = s->a;
            fib.hi[1 + c] = fib.lo[1 + c] + 1;
            memcpy(loc, fib.lo, sizeof(loc));
            do {
                unsigned int l2[WPO_MAXD];
                memcpy(l2, loc, sizeof(l2));
                int from_min = (int)s->a.lo[1 + c];
                if (from_min < 1) from_min = 1;
                for (int from = (int)s->a.hi[1 + c] - 1; from >= from_min; from--) {
                    l2[1 + c] = (unsigned int)from;
                    size_t fi = flat(p, l2);
                    double p_state = fv[fi] * bv[fi] / P;
                    acc[1] += p_state * (double)from;
                    int to_max = (int)s->b.hi[1 + c] < from ? (int)s->b.hi[1 + c] : from;
                    for (int to = (int)s->b.lo[1 + c]; to < to_max; to++) {
                        l2[1 + c] = (unsigned int)to;
                        size_t ti = flat(p, l2);
                        double p_tr = fv[fi] * Pq[from * (from + 1) / 2 + to] * nbv[ti] / P;
                        acc[0] += p_tr * (double)(from - to);
                    }
                }
            } while (box_next(&fib, nd, loc));
        }
        break;
    }
    case ST_DETACH: { /* detach-transition.cc:105-166 */
        double pd = wpo_p_detach(m, (unsigned int)s->t);
        int skip = box_includes_zero(&s->a, nd);
        for (int tensor = 0; tensor < 2; tensor++) {
            const double* fv = tensor ? f->B : f->A;
            const double* bv = tensor ? b->B : b->A;
            double fs = 0.0, fbs = 0.0;
            unsigned int old_t = 0xffffffffu;
            memcpy(loc, s->a.lo, sizeof(loc));
            do {
                if (skip && loc[0] != old_t) { /* first entry of each Edman slice */
                    old_t = loc[0];
                    continue;
                }
                size_t i = flat(p, loc);
                fs += fv[i];
                fbs += fv[i] * bv[i];
            } while (box_next(&s->a, nd, loc));
            st[WP_FIT_DETACH_X + s->t] += fs * pd * nb->delta / P;
            st[WP_FIT_DETACH_X + (T - 1) + s->t] += fbs / P;
        }
        break;
    }
    case ST_EDMAN: { /* edman-transition.cc:279-325 */
        double pf = m->p_edman_failure;
        int skip = box_includes_zero(&s->a, nd);
        unsigned int old_t = 0xffffffffu;
        memcpy(loc, s->a.lo, sizeof(loc));
        do {
            if (skip && loc[0] != old_t) {
                old_t = loc[0];
                continue;
            }
            size_t i = flat(p, loc);
            st[WP_FIT_P_EDMAN_FAILURE] += f->A[i] * pf * nb->A[i] / P;
            st[WP_FIT_P_EDMAN_FAILURE + 1] += f->A[i] * b->A[i] / P;
        } while (box_next(&s->a, nd, loc));
        break;
    }
    case ST_EMIT: /* peptide-emission.cc:255-260: no statistics */
        break;
    }
}

double wpo_improve_fit(const wp_model* m, unsigned int T, unsigned int C,
                       const short* seq, int len, const double* rad,
                       unsigned int M, double cutoff, double* stats) {
    unsigned int track[WP_MAX_TIMESTEPS * WP_MAX_CHANNELS];
    oplan p;
    octx x;
    if (!pair_setup(&p, &x, m, T, C, seq, len, track, rad, M, cutoff)) {
        free(p.steps);
        return 0.0;
    }
    int K = p.nsteps;
    /* backward states: bs[k] enters step k from the right side's view, bs[K]
     * is the finish state (peptide-state-vector.cc:42-50) */
    ostate* bs = (ostate*)malloc(sizeof(ostate) * (size_t)(K + 1));
    for (int k = 0; k <= K; k++) state_alloc(&bs[k], p.size);
    {
        const obox* r = &p.backward_range;
        unsigned int loc[WPO_MAXD];
        memcpy(loc, r->lo, sizeof(loc));
        do {
            size_t i = flat(&p, loc);
            bs[K].A[i] = 1.0;
            bs[K].B[i] = 1.0;
        } while (box_next(r, (int)p.nd, loc));
        bs[K].delta = 1.0;
        bs[K].allow = 1;
    }
    for (int k = K - 1; k >= 0; k--) bwd_step(&x, &p.steps[k], &bs[k + 1], &bs[k]);
    double P = bs[0].A[source_index(&p)];
    if (P != 0.0) { /* generic-hmm.h:77-87 */
        ostate a, b;
        state_alloc(&a, p.size);
        state_alloc(&b, p.size);
        a.A[source_index(&p)] = 1.0;
        for (int k = 0; k < K; k++) {
            fit_step(&x, &p.steps[k], &a, &bs[k], &bs[k + 1], P, stats);
            state_clear(&b, p.size);
            fwd_step(&x, &p.steps[k], &a, &b);
            ostate tmp = a;
            a = b;
            b = tmp;
        }
        state_free(&a);
        state_free(&b);
    }
    for (int k = 0; k <= K; k++) state_free(&bs[k]);
    free(bs);
    free(x.E);
    free(p.steps);
    return P;
}

void wpo_fit_estep(const wp_model* m, unsigned int T, unsigned int C,
                   const short* seq, int len, int num_reads, const double* rads,
                   unsigned int M, double* stats, double* sum_log_p) {
    int L = WP_FIT_STATS_LEN(T, C);
    int nt = wpo_num_threads();
    double* part = (double*)calloc((size_t)nt * (L + 1), sizeof(double));
#pragma omp parallel
    {
#ifdef _OPENMP
        int tid = omp_get_thread_num();
#else
        int tid = 0;
#endif
        double* mine = &part[(size_t)tid * (L + 1)];
        double* one = (double*)malloc(sizeof(double) * (size_t)L);
#pragma omp for schedule(static)
        for (int r = 0; r < num_reads; r++) {
            /* each read fills its own fitter, then fitter += it (hmm-fitter.h:63-78) */
            memset(one, 0, sizeof(double) * (size_t)L);
            double P = wpo_improve_fit(m, T, C, seq, len, &rads[(size_t)r * T * C], M, DBL_MAX, one);
            for (int i = 0; i < L; i++) mine[i] += one[i];
            mine[L] += log(P);
        }
        free(one);
    }
    for (int i = 0; i <= L; i++) {
        double s = 0.0;
        for (int t = 0; t < nt; t++) s += part[(size_t)t * (L + 1) + i];
        if (i < L) stats[i] = s;
        else *sum_log_p = s;
    }
    free(part);
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
