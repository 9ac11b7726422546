// Per-read tables shared by the FP32 candidate-loop classify kernels (the
// quad-rows kernel, hmm_v5.hip, and the lane-group kernel, hmm_v6.hip): a
// workgroup owns one read, builds its tables in LDS, then scores the read's
// candidates against them.
//
// Read-level boxes (see hmm_kernels.hip): every candidate of a read is scored
// on the read-level box [L, Hr), so box shapes, composite Bleach matrices and
// emission densities are computed once per read; only the dud start D and the
// dye position table differ per candidate.
//
// Scaling. Emission densities are stored per timestep relative to the
// largest density of that timestep's box (E = Ehat 2^X_t); a read whose box
// densities span more than 2^kMaxDensitySpan is handed to the wide pass
// (kernels that keep one exponent per density).
//
// Blocked and detached mass. The N-terminally blocked tensor B and the
// detached share are never Edman-shifted: from one step to the next they see
// Detach, Bleach and Emit only, which are the read's, not the candidate's. So
// a per-read backward pass gives, for every timestep t, what a unit of A at
// (t - 1, q) will add to P through CyclicBlock and Detach(t - 1) at step t,
//   g_t[q] = pd_t phi_t + pcb (1 - pd_t) sum_n Wt[q][n] E_t[n] beta_t[n],
// with beta_t[n] the final worth of a unit of blocked mass at (t, n) after
// Emit(t) and phi_t that of a unit of detached mass entering at step t
// (the product of E_u(0) over u >= t where box(u) holds the zero count, else
// 0); and the initial blocked share p_ib Dud(D -> n) E_0[n] is worth
// beta_0[n]. Each candidate then runs A alone and adds row . g_t at every
// step: P = sum A_(T-1) + those sums, which is the reference's
// sum A + sum B + detached share (peptide-state-vector.cc:52-54) with the same
// step order, regrouped.
//
// Per-timestep box classes. A read's box shrinks as its dyes leave (most
// timesteps of most reads have a one-state box), so timestep t can run at
// the class K_t = the power of two >= max(|box(t-1)|, |box(t)|) rather than
// at the bin's capacity NB: the composite matrices of t are then stored
// K_t x K_t and the step body is instantiated per class, chosen uniformly
// per read.
#pragma once

#include <float.h>
#include <math.h>

#include <algorithm>

#include "classify_common.h"
#include "hmm_kernels.h"

namespace wp {
namespace rt {

using namespace cc;

constexpr int kReadThreads = 256;
constexpr int kMaxDensitySpan = 90;  // largest binary-exponent spread of one box's densities
constexpr int kMaxBackSpread = 800;  // largest spread of one level of the blocked/detached backward pass

struct ReadLayout {
    size_t rad, L, H, nbt, kcl, woff, bstr, coord, kcf, ev, etab, xt, W, binom, list, scan, gres, ties, gtab, gx,
        b0, bd, total;
};

// W block per timestep: (1 + C) matrices (Wt, then W1[c]) of K_t x K_t with a
// stride of K_t * K_t + 4 entries, so the shifted matrices of different
// channels start in different LDS banks; sized for K_t = NB at every t.
// ngroups: candidate groups of the workgroup (accumulators and tie slots);
// wcap: if nonzero, the W block's capacity in floats (per-timestep classes
// only: a read whose classes need more goes to the wide pass).
__host__ __device__ inline ReadLayout read_layout(int T, int C, int tri, int NB, int ngroups,
                                              size_t wcap = 0) {
    ReadLayout l;
    const size_t ms = (size_t)NB * NB + 4;
    const size_t wfull = (size_t)T * (1 + C) * ms;
    const size_t wsize = wcap && wcap < wfull ? wcap : wfull;
    size_t o = 0;
    l.rad = o;   o = align16(o + sizeof(double) * T * C);
    l.L = o;     o = align16(o + sizeof(int) * T * C);
    l.H = o;     o = align16(o + sizeof(int) * T * C);
    l.nbt = o;   o = align16(o + sizeof(int) * T);
    l.kcl = o;   o = align16(o + sizeof(int) * T);
    l.woff = o;  o = align16(o + sizeof(int) * T);
    l.bstr = o;  o = align16(o + sizeof(int) * T * C);
    l.coord = o; o = align16(o + sizeof(int) * (size_t)T * NB * C);
    l.kcf = o;   o = align16(o + sizeof(float) * (size_t)T * NB * C);
    l.ev = o;    o = align16(o + sizeof(double) * (size_t)T * NB);
    l.etab = o;  o = align16(o + sizeof(float) * (size_t)T * NB);
    l.xt = o;    o = align16(o + sizeof(int) * T);
    l.W = o;     o = align16(o + sizeof(float) * wsize);
    l.binom = o; o = align16(o + sizeof(double) * 2 * C * tri);
    l.list = o;  o = align16(o + sizeof(int) * kReadThreads * 4);
    l.scan = o;  o = align16(o + sizeof(int) * (kReadThreads / 64 + 2));
    l.gres = o;  o = align16(o + sizeof(double) * 4 * ngroups);
    l.ties = o;  o = align16(o + sizeof(TieSlot) * ngroups);
    l.gtab = o;  o = align16(o + sizeof(double) * (size_t)T * NB);
    l.gx = o;    o = align16(o + sizeof(int) * (T + 1));
    l.b0 = o;    o = align16(o + sizeof(double) * NB);
    l.bd = o;    o = align16(o + sizeof(double) * NB);
    l.total = o;
    return l;
}

// Lane p of each quad receives the value of lane (p + 3) % 4.
__device__ inline int quad_rot_i(int v) {
    return __builtin_amdgcn_update_dpp(0, v, 3 | (0 << 2) | (1 << 4) | (2 << 6), 0xf, 0xf, false);
}
__device__ inline float quad_rot(float v) { return __int_as_float(quad_rot_i(__float_as_int(v))); }

// s 2^e += v 2^ev, s kept in [0.5, 1) (s = 0 holds any e). FP64: a row's
// share of g_t can lie far below the row's own scale.
__device__ __forceinline__ void side_add(double& s, int& e, double v, int ev) {
    if (v > 0.0) {
        const int mx = max(e, ev);
        int k;
        s = frexp(ldexp(s, e - mx) + ldexp(v, ev - mx), &k);
        e = mx + k;
    }
}

// LDS tables of one read (offsets of ReadLayout).
struct ReadTables {
    double* s_rad;
    int *sL, *sH, *nbt, *kcl, *woff, *bstr, *coord;
    float* kcf;
    double* ev;
    float* etab;
    int* xt;
    float* Wall;
    double* bin;
    int *list, *scan;
    double* gres;
    TieSlot* ties;
    double* gtab;
    int* gx;  // [T]: exponents of g_t (gx[0]: of beta_0); [T]: spread flag
    double *b0d, *bd;
};

__device__ inline ReadTables read_views(unsigned char* smem, const ReadLayout& lay) {
    ReadTables s;
    s.s_rad = (double*)(smem + lay.rad);
    s.sL = (int*)(smem + lay.L);
    s.sH = (int*)(smem + lay.H);
    s.nbt = (int*)(smem + lay.nbt);
    s.kcl = (int*)(smem + lay.kcl);
    s.woff = (int*)(smem + lay.woff);
    s.bstr = (int*)(smem + lay.bstr);
    s.coord = (int*)(smem + lay.coord);
    s.kcf = (float*)(smem + lay.kcf);
    s.ev = (double*)(smem + lay.ev);
    s.etab = (float*)(smem + lay.etab);
    s.xt = (int*)(smem + lay.xt);
    s.Wall = (float*)(smem + lay.W);
    s.bin = (double*)(smem + lay.binom);
    s.list = (int*)(smem + lay.list);
    s.scan = (int*)(smem + lay.scan);
    s.gres = (double*)(smem + lay.gres);
    s.ties = (TieSlot*)(smem + lay.ties);
    s.gtab = (double*)(smem + lay.gtab);
    s.gx = (int*)(smem + lay.gx);
    s.b0d = (double*)(smem + lay.b0);
    s.bd = (double*)(smem + lay.bd);
    return s;
}

// Whether read r belongs to this launch's bin.
__device__ inline bool in_bin(const ClassifyArgs& a, int r) {
    if (r >= a.N) return false;
    const int nbr = a.nbmax[r];
    return (nbr > a.nb_lo && nbr <= a.nb_hi) || (nbr == 0 && a.nb_lo == 0);
}

// The read-level prologue shared by the candidate loops: boxes, coordinates,
// emission densities relative to each timestep's largest, the composite
// Bleach matrices (per box class with kClass, else at NB) and the blocked /
// detached backward tables. Returns false, uniformly over the workgroup, when
// the read is handed to the wide pass instead. Block-wide.
// Timestep t's (1 + C) matrices of class K, built by the whole workgroup.
// Entry (q, q2) of matrix mat (0: the composite Bleach Wt between the boxes
// of t - 1 and t; 1 + c: its Edman-shifted form W1[c], which maps onto
// n'(q2) + 1_c and is zero at or above U = min(Hr[t] + 1, Hr[t-1]),
// edman-transition.cc:28-60) is the product over channels of the bleach
// table entries from the read's box coordinates. Each thread takes entries
// (q, q2) of the K x K block (K a power of two: shifts, no divisions), loads
// both states' coordinates once and writes the entry of every matrix;
// matrices are K * K + 4 entries apart.
template <int C, int NB>
__device__ inline void fill_w_block(const ReadTables& s, int tri, int t, int K, float* dst) {
    const int ms = K * K + 4;
    const int lk = 31 - __builtin_clz(K);
    const int n0 = s.nbt[t - 1], n1 = s.nbt[t];
    const double* ble = s.bin + C * tri;
    int lim[C];  // shifted channel c: target counts below U
#pragma unroll
    for (int c = 0; c < C; c++) lim[c] = min(s.sH[t * C + c] + 1, s.sH[(t - 1) * C + c]);
    for (int e = threadIdx.x; e < K * K; e += kReadThreads) {
        const int q = e >> lk, q2 = e & (K - 1);
        const bool in = q < n0 && q2 < n1;
        int from[C], to[C];
#pragma unroll
        for (int c = 0; c < C; c++) {
            from[c] = in ? s.coord[((t - 1) * NB + q) * C + c] : 0;
            to[c] = in ? s.coord[(t * NB + q2) * C + c] : 0;
        }
#pragma unroll
        for (int mat = 0; mat <= C; mat++) {
            double v = in ? 1.0 : 0.0;
#pragma unroll
            for (int c = 0; c < C; c++) {
                const int tc = to[c] + (c == mat - 1 ? 1 : 0);
                if (c == mat - 1 && tc >= lim[c]) v = 0.0;
                v = from[c] < tc ? 0.0 : v * ble[c * tri + from[c] * (from[c] + 1) / 2 + tc];
            }
            dst[mat * ms + e] = (float)v;
        }
    }
}

// kFillW false: the W block is left to the caller (built per timestep), and
// no W capacity applies.
template <int C, int NB, bool kClass, bool kFillW = true>
__device__ __forceinline__ bool read_prologue(const ClassifyArgs& a, const int r, const ReadTables& s, int wcap) {
    constexpr int MSNB = NB * NB + 4;
    const int T = a.m.T;
    const int tri = a.tri;
    const int tid = threadIdx.x;
    double* s_rad = s.s_rad;
    int* sL = s.sL;
    int* sH = s.sH;
    int* nbt = s.nbt;
    int* kcl = s.kcl;
    int* woff = s.woff;
    int* bstr = s.bstr;
    int* coord = s.coord;
    float* kcf = s.kcf;
    double* ev = s.ev;
    float* etab = s.etab;
    int* xt = s.xt;
    float* Wall = s.Wall;
    double* bin = s.bin;
    int* scan = s.scan;
    double* gtab = s.gtab;
    int* gx = s.gx;
    double* b0d = s.b0d;
    double* bd = s.bd;
    const bool read_empty = a.nbmax[r] == 0;

    // ---- read-level boxes, coordinates, emission densities
    for (int k = tid; k < T * C; k += kReadThreads) {
        s_rad[k] = a.rads[(size_t)r * T * C + k];
        sL[k] = a.Lr[(size_t)r * T * C + k];
        sH[k] = a.Hr[(size_t)r * T * C + k];
    }
    for (int k = tid; k < 2 * C * tri; k += kReadThreads) bin[k] = a.binom[k];
    if (tid == 0) scan[0] = scan[1] = 0;
    __syncthreads();
    if (tid < T) {
        const int t = tid;
        int s = 1;
        for (int c = C - 1; c >= 0; c--) {
            bstr[t * C + c] = s;
            const int w = sH[t * C + c] - sL[t * C + c];
            s *= w > 0 ? w : 0;
        }
        nbt[t] = read_empty ? 0 : s;
    }
    __syncthreads();
    if (kClass && tid == 0) {  // per-timestep classes and the offsets of their W blocks
        int o = 0;
        for (int t = 0; t < T; t++) {
            const int need = t == 0 ? nbt[0] : max(nbt[t - 1], nbt[t]);
            int k = 1;
            while (k < need) k *= 2;
            kcl[t] = k;
            woff[t] = o;
            if (t >= 1) o += (1 + C) * (k * k + 4);
        }
        if (kFillW && o > wcap) scan[1] = 1;  // the classes' matrices exceed the W block
    }
    __syncthreads();
    // peptide-emission.cc:165-180
    for (int k = tid; k < T * NB; k += kReadThreads) {
        const int t = k / NB, q = k % NB;
        unsigned int n[WP_MAX_CHANNELS];
        double v = 0.0;
        if (q < nbt[t]) {
            int rem = q;
#pragma unroll
            for (int c = 0; c < C; c++) {
                const int d = rem / bstr[t * C + c];
                rem -= d * bstr[t * C + c];
                n[c] = (unsigned int)(sL[t * C + c] + d);
            }
            v = 1.0;
#pragma unroll
            for (int c = 0; c < C; c++) v *= dev_pdf<C>(a.m, c, s_rad[t * C + c], n);
        } else {
#pragma unroll
            for (int c = 0; c < C; c++) n[c] = 0;
        }
#pragma unroll
        for (int c = 0; c < C; c++) {
            coord[k * C + c] = (int)n[c];
            kcf[k * C + c] = (float)n[c];
        }
        ev[k] = v;
    }
    __syncthreads();
    if (tid < T) {  // X_t: the largest binary exponent of the box's densities
        int x = kNoExp, lo = -kNoExp;
        for (int q = 0; q < NB; q++) {
            const double v = ev[tid * NB + q];
            if (v > 0.0) {
                int ex;
                frexp(v, &ex);
                x = max(x, ex);
                lo = min(lo, ex);
            }
        }
        xt[tid] = x == kNoExp ? 0 : x;
        if (x != kNoExp && x - lo > kMaxDensitySpan) scan[0] = 1;  // cleared before the barrier above
    }
    __syncthreads();
    // Densities spanning more than 2^kMaxDensitySpan in one box could fall below
    // FP32's range relative to the box maximum: such a read is flagged and
    // left to the kernels that keep one exponent per density. A read whose
    // box classes outgrow a capped W block goes to the launch's over list
    // (the uncapped pass, which builds each timestep's matrices before its
    // step) or, without one, to the wide pass as well.
    const bool span = scan[0] != 0, over = scan[1] != 0;
    if (tid == 0) {
        const bool to_over = over && !span && a.over_list != nullptr;
        a.wide[r] = (span || over) && !to_over ? 1 : 0;
        if (to_over)
            a.over_list[atomicAdd(a.over_count, 1)] = r;
        else if (span || over)
            a.wide_list[atomicAdd(a.wide_count, 1)] = r;
    }
    if (span || over) return false;
    for (int k = tid; k < T * NB; k += kReadThreads) etab[k] = (float)ldexp(ev[k], -xt[k / NB]);
    // Composite Bleach matrices of every timestep (unless the caller builds
    // each timestep's just before its step): timestep t's matrices are
    // K_t x K_t, K_t * K_t + 4 entries apart (different LDS banks for the
    // shifted matrices of different channels).
    // entries of the W blocks of t >= 1: (1 + C) matrices of K_t x K_t each
    const int wtotal = !kFillW ? 0
                       : kClass ? woff[T - 1] + (T > 1 ? (1 + C) * (kcl[T - 1] * kcl[T - 1] + 4) : 0)
                                : (T - 1) * (1 + C) * MSNB;
    if (wtotal > 0)
        for (int t = 1; t < T; t++)
            fill_w_block<C, NB>(s, tri, t, kClass ? kcl[t] : NB, Wall + (kClass ? woff[t] : t * (1 + C) * MSNB));
    // Blocked and detached mass (one wave, lane q = box entry q): from level
    // t + 1 = u, beta_u = bq 2^eb and phi_(u+1) = ph 2^eb, to level t, both
    // relative to 2^(eb + X_u) (densities E = Ehat 2^X), in FP64:
    //   phi_u = Ehat_u(0) ph (box(u) holds the zero count) or 0,
    //   v[q] = sum_m Bleach(q -> m) Ehat_u[m] bq[m]  (Bleach_c, Detach and Emit of step u),
    //   beta_t[q] = pd phi_u + (1 - pd) v[q],  g_u[q] = pd phi_u + pcb (1 - pd) v[q],
    // then level t is renormalised by a power of two. A level whose nonzero
    // entries spread over more than 2^kMaxBackSpread hands the read to the
    // wide pass: below that every product of the next level (Bleach >= 2^-65,
    // relative densities >= 2^-kMaxDensitySpan) is a normal double, so no share
    // underflows unseen.
    const double* ble = bin + C * tri;
    if (tid < 64) {
        const int q = tid;
        double bq = q < nbt[T - 1] ? 1.0 : 0.0;
        double ph = 1.0;
        int eb = 0;
        bool spread = false;
        const double pcb_d = a.m.p_cb;
        for (int t = T - 2; t >= 0; t--) {
            const int u = t + 1;
            if (q < NB) bd[q] = bq;
            wave_sync();
            bool zero_in = true;
#pragma unroll
            for (int c = 0; c < C; c++) zero_in = zero_in && sL[u * C + c] == 0;
            ph = zero_in ? ldexp(ev[u * NB], -xt[u]) * ph : 0.0;
            double v = 0.0;
            if (q < nbt[t]) {
                for (int m2 = 0; m2 < nbt[u]; m2++) {
                    double wv = ldexp(ev[u * NB + m2], -xt[u]) * bd[m2];
#pragma unroll
                    for (int c = 0; c < C; c++) {
                        const int from = coord[(t * NB + q) * C + c];
                        const int to = coord[(u * NB + m2) * C + c];
                        wv = from < to ? 0.0 : wv * ble[c * tri + from * (from + 1) / 2 + to];
                    }
                    v += wv;
                }
            }
            const double pd = a.m.p_detach[u - 1];
            const double g = q < nbt[t] ? pd * ph + pcb_d * (1.0 - pd) * v : 0.0;
            bq = q < nbt[t] ? pd * ph + (1.0 - pd) * v : 0.0;
            double gm = g, bm = fmax(bq, ph);
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) {
                gm = fmax(gm, __shfl_xor(gm, off, 64));
                bm = fmax(bm, __shfl_xor(bm, off, 64));
            }
            int ge = 0, be = 0;
            if (gm > 0.0) frexp(gm, &ge);
            if (bm > 0.0) frexp(bm, &be);
            // smallest nonzero exponents of the level
            int gl = 100000, bl = 100000, k2;
            if (g > 0.0) {
                frexp(g, &k2);
                gl = k2;
            }
            if (bq > 0.0) {
                frexp(bq, &k2);
                bl = k2;
            }
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) {
                gl = min(gl, __shfl_xor(gl, off, 64));
                bl = min(bl, __shfl_xor(bl, off, 64));
            }
            spread = spread || (gm > 0.0 && ge - gl > kMaxBackSpread) || (bm > 0.0 && be - bl > kMaxBackSpread);
            if (q < NB) gtab[u * NB + q] = ldexp(g, -ge);
            if (q == 0) gx[u] = eb + xt[u] + ge;
            bq = ldexp(bq, -be);
            ph = ldexp(ph, -be);
            eb += xt[u] + be;
            wave_sync();
        }
        if (q < NB) b0d[q] = bq;
        if (q == 0) {
            gx[0] = eb;  // beta_0 = b0d 2^gx[0]
            gx[T] = spread ? 1 : 0;
        }
    }
    __syncthreads();
    if (gx[T]) {
        if (tid == 0) {
            a.wide[r] = 1;
            a.wide_list[atomicAdd(a.wide_count, 1)] = r;
        }
        return false;
    }
    return true;
}

}  // namespace rt
}  // namespace wp
