// Classify kernel (FP32 rescaled) for reads whose count box holds at most 16
// states: a quad of lanes per candidate, several Edman rows per lane.
//
// Same recursion as the row-per-lane kernel (read-level boxes, composite
// Bleach matrices Wt and Edman-shifted W1[c] shared by every candidate of
// the read, Edman success handed to the next row), laid out so that no lane
// works on rows that cannot hold mass yet: lane p of a candidate's quad keeps
// rows e = p + 4k for k < R in registers (row-slot k). Row e is zero before
// t = e, so slot k is skipped, uniformly over the wave, while t < 4k. The
// Edman success of row e - 1 reaches row e through one quad DPP rotation:
// lanes 1..3 take slot k of the lane below, lane 0 takes slot k - 1 of lane 3.
//
// Scaling. Emission densities are stored per timestep relative to the
// largest density of that timestep's box (E = Ehat 2^X_t), and every row
// carries one power-of-two exponent, so the per-entry work of a step is the
// Edman coefficients, the two updates, the emission product and a rescale;
// row exponents meet only where the Edman success crosses rows and in the
// final sum of a pair.
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
// timesteps of most reads have a one-state box), so each timestep t runs at
// the class K_t = the power of two >= max(|box(t-1)|, |box(t)|), not at the
// bin's capacity NB: the composite matrices of t are stored K_t x K_t and the
// step body is instantiated per class, chosen uniformly per read.
#include <float.h>
#include <math.h>

#include <algorithm>

#include "classify_common.h"
#include "hmm_kernels.h"

namespace wp {
namespace {

using namespace cc;

constexpr int kV5Threads = 256;
constexpr int kV5Lanes = 4;  // lanes per candidate: one DPP quad
constexpr int kV5MaxSpan = 90;  // largest binary-exponent spread of one box's densities
constexpr int kV5MaxBackSpread = 800;  // largest spread of one level of the blocked/detached backward pass

struct V5Layout {
    size_t rad, L, H, nbt, kcl, woff, bstr, coord, kcf, ev, etab, xt, W, binom, list, scan, gres, ties, gtab, gx,
        b0, bd, total;
};

// W block per timestep: (1 + C) matrices (Wt, then W1[c]) of K_t x K_t with a
// stride of K_t * K_t + 4 entries, so the shifted matrices of different
// channels start in different LDS banks; sized for K_t = NB at every t.
__host__ __device__ inline V5Layout v5_layout(int T, int C, int tri, int NB) {
    V5Layout l;
    const size_t ms = (size_t)NB * NB + 4;
    const int ngroups = kV5Threads / kV5Lanes;
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
    l.W = o;     o = align16(o + sizeof(float) * (size_t)T * (1 + C) * ms);
    l.binom = o; o = align16(o + sizeof(double) * 2 * C * tri);
    l.list = o;  o = align16(o + sizeof(int) * kV5Threads * 4);
    l.scan = o;  o = align16(o + sizeof(int) * (kV5Threads / 64 + 2));
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

// Per-timestep constants of one step t >= 1, shared by every candidate.
struct V5Step {
    const float* Wt;   // K x K composite Bleach, then the C Edman-shifted forms, K * K + 4 apart
    const float* Et;   // densities of box(t) relative to 2^Xt
    const float* kct;  // [NB][C] counts of box(t) as floats
    const double* gt;  // g_t over box(t - 1), relative to 2^Gx
    int nbp, t, Xt, Gx, p;
    float cAA, pf;
};

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

// One step t >= 1 of a candidate's recursion (CyclicBlock, Detach(t-1),
// Bleach_c, Edman, Emit(t)) on a box of class K (both box(t-1) and box(t) fit
// in K entries), A only: the shares CyclicBlock and Detach take from each row
// go to the row's side sum (sm, se) through g_t. Entries >= K of A are stale
// and never read: inputs stop at |box(t-1)| and outputs at K.
template <int C, int NB, int R, int K>
__device__ __forceinline__ void v5_step(const V5Step& st, const int (&cix)[R], const float (&aK)[R],
                                        const float (&cst)[R], const float (&Kf)[R], float (&A)[R][NB],
                                        int (&kx)[R], double (&sm)[R], int (&se)[R]) {
    constexpr int G = kV5Lanes;
    constexpr int MS = K * K + 4;
    const float* Wt = st.Wt;
    const int p = st.p;
    float succ_prev[K];  // this lane's slot k - 1 Edman success (old scale)
    int kprev = kNoExp;  // and slot k - 1's exponent before this step
#pragma unroll
    for (int q = 0; q < K; q++) succ_prev[q] = 0.0f;
#pragma unroll
    for (int k = 0; k < R; k++) {
        if (st.t < G * k) continue;  // every row of slot k starts after t (uniform)
        // Every slot reads the same Wt; without this the compiler keeps the
        // previous slot's Wt loads live across slots, which costs registers
        // for no LDS traffic saved worth having: the reads are broadcasts.
        if constexpr (K > 1) asm volatile("" ::: "memory");
        const float* W1 = Wt + (1 + cix[k]) * MS;
        float BA[K], BA1[K];
#pragma unroll
        for (int q2 = 0; q2 < K; q2++) {
            BA[q2] = 0.0f;
            BA1[q2] = 0.0f;
        }
        double sv = 0.0;  // row . g_t: the row's blocked and detached shares of this step
#pragma unroll
        for (int q = 0; q < K; q++) {
            if (q >= st.nbp) continue;  // uniform: the previous box's size
            const float av = A[k][q];
            sv = fma((double)av, st.gt[q], sv);
#pragma unroll
            for (int q2 = 0; q2 < K; q2++) {
                BA[q2] += Wt[q * K + q2] * av;
                BA1[q2] += W1[q * K + q2] * av;
            }
        }
        // exponents of this row and of the row below it (e - 1)
        const int ko = kx[k];
        if (ko != kNoExp) side_add(sm[k], se[k], sv, ko + st.Gx);
        const int kfrom = quad_rot_i(p == G - 1 ? kprev : ko);
        int kn = max(ko, kfrom);
        if (kn == kNoExp) kn = 0;
        const float sown = ko == kNoExp ? 0.0f : ldexpf(1.0f, ko - kn);
        const float sfrom = kfrom == kNoExp ? 0.0f : ldexpf(1.0f, kfrom - kn);
        const float fAo = st.cAA * st.pf * sown, fAf = st.cAA * sfrom;
        float m = 0.0f;
#pragma unroll
        for (int q2 = 0; q2 < K; q2++) {
            const float kc = st.kct[q2 * C + cix[k]];
            const float cba = fmaf(fmaxf(Kf[k] - kc, 0.0f), aK[k], cst[k]);
            const float cba1 = fmaf(kc, aK[k], aK[k]);
            const float succ = fmaf(cba1, BA1[q2], cba * BA[q2]);
            const float from = quad_rot(p == G - 1 ? succ_prev[q2] : succ);
            succ_prev[q2] = succ;
            // Emit(t) (peptide-emission.cc:212-243)
            const float na = fmaf(fAf, from, fAo * BA[q2]) * st.Et[q2];
            m = fmaxf(m, na);
            A[k][q2] = na;
        }
        kprev = ko;
        const int km = exp_of(m);
        const int kf = km == kNoExp ? 0 : km;
#pragma unroll
        for (int q2 = 0; q2 < K; q2++) A[k][q2] = ldexpf(A[k][q2], -kf);
        kx[k] = km == kNoExp ? kNoExp : kn + st.Xt + kf;
    }
}

// One read: prologue (boxes, densities, composite matrices) then its
// candidates. kClass: each timestep at its own box class; otherwise every
// timestep at the bin's capacity NB. Block-uniform returns only.
template <int C, int NB, int R, bool kClass>
__device__ __forceinline__ void v5_read(const ClassifyArgs& a, const int r, unsigned char* smem) {
    constexpr int G = kV5Lanes;
    constexpr int NGR = kV5Threads / G;
    constexpr int MSNB = NB * NB + 4;
    if (r >= a.N) return;
    const int nbr = a.nbmax[r];
    if (!((nbr > a.nb_lo && nbr <= a.nb_hi) || (nbr == 0 && a.nb_lo == 0))) return;

    const int T = a.m.T;
    const int tri = a.tri;
    const int tid = threadIdx.x;
    const int group = tid / G;
    const int p = tid % G;
    const V5Layout lay = v5_layout(T, C, tri, NB);
    double* s_rad = (double*)(smem + lay.rad);
    int* sL = (int*)(smem + lay.L);
    int* sH = (int*)(smem + lay.H);
    int* nbt = (int*)(smem + lay.nbt);
    int* kcl = (int*)(smem + lay.kcl);
    int* woff = (int*)(smem + lay.woff);
    int* bstr = (int*)(smem + lay.bstr);
    int* coord = (int*)(smem + lay.coord);
    float* kcf = (float*)(smem + lay.kcf);
    double* ev = (double*)(smem + lay.ev);
    float* etab = (float*)(smem + lay.etab);
    int* xt = (int*)(smem + lay.xt);
    float* Wall = (float*)(smem + lay.W);
    double* bin = (double*)(smem + lay.binom);
    int* list = (int*)(smem + lay.list);
    int* scan = (int*)(smem + lay.scan);
    double* gres = (double*)(smem + lay.gres);
    TieSlot* ties = (TieSlot*)(smem + lay.ties);
    double* gtab = (double*)(smem + lay.gtab);
    int* gx = (int*)(smem + lay.gx);  // [T]: exponents; [T]: spread flag
    double* b0d = (double*)(smem + lay.b0);
    double* bd = (double*)(smem + lay.bd);
    int ncand;
    const int* cidx;
    read_candidates(a, r, &ncand, &cidx);
    const bool read_empty = nbr == 0;

    // ---- prologue: read-level boxes, coordinates, emission densities
    for (int k = tid; k < T * C; k += kV5Threads) {
        s_rad[k] = a.rads[(size_t)r * T * C + k];
        sL[k] = a.Lr[(size_t)r * T * C + k];
        sH[k] = a.Hr[(size_t)r * T * C + k];
    }
    for (int k = tid; k < 2 * C * tri; k += kV5Threads) bin[k] = a.binom[k];
    if (tid == 0) scan[0] = 0;
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
    }
    __syncthreads();
    // peptide-emission.cc:165-180
    for (int k = tid; k < T * NB; k += kV5Threads) {
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
        if (x != kNoExp && x - lo > kV5MaxSpan) scan[0] = 1;  // cleared before the barrier above
    }
    __syncthreads();
    // Densities spanning more than 2^kV5MaxSpan in one box could fall below
    // FP32's range relative to the box maximum: such a read is flagged and
    // left to the kernels that keep one exponent per density.
    const bool wide = scan[0] != 0;
    if (tid == 0) {
        a.wide[r] = wide ? 1 : 0;
        if (wide) a.wide_list[atomicAdd(a.wide_count, 1)] = r;
    }
    if (wide) return;
    for (int k = tid; k < T * NB; k += kV5Threads) etab[k] = (float)ldexp(ev[k], -xt[k / NB]);
    // Composite Bleach between the boxes of t-1 and t and its Edman-shifted
    // forms: W1[c][q][q2] maps onto n'(q2) + 1_c and is zero at or above
    // U = min(Hr[t] + 1, Hr[t-1]) (edman-transition.cc:28-60). Timestep t's
    // matrices are K_t x K_t, K_t + 4 entries apart (different LDS banks for
    // the shifted matrices of different channels).
    const double* ble = bin + C * tri;
    // entries of the W blocks of t >= 1: (1 + C) matrices of K_t x K_t each
    const int wtotal = kClass ? woff[T - 1] + (T > 1 ? (1 + C) * (kcl[T - 1] * kcl[T - 1] + 4) : 0)
                              : (T - 1) * (1 + C) * MSNB;
    for (int k = tid; k < wtotal; k += kV5Threads) {
        int t = 1, K = NB, rem;
        if constexpr (kClass) {
            while (t + 1 < T && woff[t + 1] <= k) t++;
            K = kcl[t];
            rem = k - woff[t];
        } else {
            t = 1 + k / ((1 + C) * MSNB);
            rem = k - (t - 1) * (1 + C) * MSNB;
        }
        const int ms = K * K + 4;
        const int mat = rem / ms;
        rem -= mat * ms;
        if (rem >= K * K) continue;  // the 4-entry gap between matrices
        const int q = rem / K, q2 = rem % K;
        double v = 0.0;
        if (t >= 1 && q < nbt[t - 1] && q2 < nbt[t]) {
            v = 1.0;
            const int sc = mat - 1;  // shifted channel, -1 for Wt
            for (int c = 0; c < C; c++) {
                const int from = coord[((t - 1) * NB + q) * C + c];
                const int to = coord[(t * NB + q2) * C + c] + (c == sc ? 1 : 0);
                if (c == sc && to >= min(sH[t * C + c] + 1, sH[(t - 1) * C + c])) v = 0.0;
                if (from < to) v = 0.0;
                else v *= ble[c * tri + from * (from + 1) / 2 + to];
            }
        }
        Wall[(kClass ? woff[t] : t * (1 + C) * MSNB) + mat * ms + rem] = (float)v;
    }
    // Blocked and detached mass (one wave, lane q = box entry q): from level
    // t + 1 = u, beta_u = bq 2^eb and phi_(u+1) = ph 2^eb, to level t, both
    // relative to 2^(eb + X_u) (densities E = Ehat 2^X), in FP64:
    //   phi_u = Ehat_u(0) ph (box(u) holds the zero count) or 0,
    //   v[q] = sum_m Bleach(q -> m) Ehat_u[m] bq[m]  (Bleach_c, Detach and Emit of step u),
    //   beta_t[q] = pd phi_u + (1 - pd) v[q],  g_u[q] = pd phi_u + pcb (1 - pd) v[q],
    // then level t is renormalised by a power of two. A level whose nonzero
    // entries spread over more than 2^kV5MaxBackSpread hands the read to the
    // wide pass: below that every product of the next level (Bleach >= 2^-65,
    // relative densities >= 2^-kV5MaxSpan) is a normal double, so no share
    // underflows unseen.
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
            spread = spread || (gm > 0.0 && ge - gl > kV5MaxBackSpread) || (bm > 0.0 && be - bl > kV5MaxBackSpread);
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
        return;
    }

    CandAcc<true> acc;
    const bool collect = a.tie_mode == 1;
    if (collect && p == 0) tie_init(&ties[group]);
    const float pf = (float)a.m.pf;
    const float opf = (float)(1.0 - a.m.pf);
    const double* dud = bin;

    for (int base = 0; base < ncand; base += kV5Threads * 4) {
        const int nlive = compact_chunk<C>(a, r, ncand, cidx, base, read_empty, sL, list, scan);
        for (int q0 = group; q0 < nlive; q0 += NGR) {
            const int kk = list[q0];
            const int i = cidx ? cidx[kk] : kk;
            int D[C];
#pragma unroll
            for (int c = 0; c < C; c++) D[c] = a.dsD[(size_t)i * C + c];
            // Edman coefficients of each row-slot: keep = aK max(K - kc, 0) + cst,
            // success = aK (kc + 1), with aK = (1 - pf) / K for a dye of channel
            // c_e at position e, and aK = 0, cst = 1 - pf for no dye there.
            int cix[R];
            float aK[R], cst[R], Kf[R];
#pragma unroll
            for (int k = 0; k < R; k++) {
                const int e = p + G * k;
                const int ce = e < T ? (int)a.dsSeq[(size_t)i * T + e] : -1;
                const int K = ce >= 0 ? (int)a.dsTrack[((size_t)i * T + e) * C + ce] : 1;
                cix[k] = ce < 0 ? 0 : ce;
                aK[k] = ce >= 0 ? opf / (float)K : 0.0f;
                cst[k] = ce >= 0 ? 0.0f : opf;
                Kf[k] = (float)K;
            }
            float A[R][NB];
            int kx[R];     // slot k holds values x 2^-kx[k]
            double sm[R];  // side sum of slot k's row (blocked and detached shares), sm 2^se
            int se[R];
#pragma unroll
            for (int k = 0; k < R; k++) {
#pragma unroll
                for (int q = 0; q < NB; q++) A[k][q] = 0.0f;
                kx[k] = kNoExp;
                sm[k] = 0.0;
                se[k] = kNoExp;
            }
            // t = 0: InitialBlock, Dud_c, Emit(0): row 0 (lane 0, slot 0); the
            // blocked share p_ib Dud(D -> n) E_0[n] is worth beta_0[n]
            if (p == 0) {
                const float pib = (float)a.m.p_ib;
                float m = 0.0f;
                double b0 = 0.0;
#pragma unroll
                for (int q = 0; q < NB; q++) {
                    bool ok = true;
                    float xa = 1.0f;
#pragma unroll
                    for (int c = 0; c < C; c++) {
                        const int n = coord[q * C + c];
                        ok = ok && n <= D[c];
                        if (ok) xa *= (float)dud[c * tri + D[c] * (D[c] + 1) / 2 + n];
                    }
                    const float em = ok ? xa * etab[q] : 0.0f;
                    A[0][q] = (1.0f - pib) * em;
                    b0 = fma((double)(pib * em), b0d[q], b0);
                    m = fmaxf(m, A[0][q]);
                }
                side_add(sm[0], se[0], b0, xt[0] + gx[0]);
                const int km = exp_of(m);
                const int kf = km == kNoExp ? 0 : km;
#pragma unroll
                for (int q = 0; q < NB; q++) A[0][q] = ldexpf(A[0][q], -kf);
                kx[0] = km == kNoExp ? kNoExp : xt[0] + kf;
            }

            for (int t = 1; t < T; t++) {
                V5Step st;
                st.Wt = Wall + (kClass ? woff[t] : t * (1 + C) * MSNB);
                st.nbp = nbt[t - 1];
                st.t = t;
                const float pd = (float)a.m.p_detach[t - 1];
                st.cAA = (1.0f - pd) * (1.0f - (float)a.m.p_cb);
                st.pf = pf;
                st.Et = etab + (size_t)t * NB;
                st.Xt = xt[t];
                st.gt = gtab + (size_t)t * NB;
                st.Gx = gx[t];
                st.kct = kcf + (size_t)t * NB * C;
                st.p = p;
                if constexpr (!kClass) {
                    v5_step<C, NB, R, NB>(st, cix, aK, cst, Kf, A, kx, sm, se);
                    continue;
                }
                const int Kt = kcl[t];  // uniform over the workgroup
                if (Kt <= 1) v5_step<C, NB, R, 1>(st, cix, aK, cst, Kf, A, kx, sm, se);
                else if (Kt == 2) v5_step<C, NB, R, 2>(st, cix, aK, cst, Kf, A, kx, sm, se);
                else if (Kt == 4 || NB == 4) v5_step<C, NB, R, (NB < 4 ? NB : 4)>(st, cix, aK, cst, Kf, A, kx, sm, se);
                else if (Kt == 8 || NB == 8) v5_step<C, NB, R, (NB < 8 ? NB : 8)>(st, cix, aK, cst, Kf, A, kx, sm, se);
                else v5_step<C, NB, R, NB>(st, cix, aK, cst, Kf, A, kx, sm, se);
            }

            // P = sum A over the last box + the rows' side sums
            const int nlast = nbt[T - 1];
            float rs[R];
            int kl = kNoExp;
#pragma unroll
            for (int k = 0; k < R; k++) {
                rs[k] = 0.0f;
#pragma unroll
                for (int q = 0; q < NB; q++)
                    if (q < nlast) rs[k] += A[k][q];
                if (rs[k] > 0.0f && kx[k] != kNoExp) kl = max(kl, kx[k]);
                if (sm[k] > 0.0) kl = max(kl, se[k]);
            }
            double part = 0.0;
#pragma unroll
            for (int k = 0; k < R; k++) {
                if (rs[k] > 0.0f && kx[k] != kNoExp) part += ldexp((double)rs[k], kx[k] - kl);
                if (sm[k] > 0.0) part += ldexp(sm[k], se[k] - kl);
            }
            const int K = group_max_int<G>(kl);
            part = (K == kNoExp || kl == kNoExp) ? 0.0 : ldexp(part, kl - K);
            const double s = group_sum<double, G>(part);
            const double logp = s > 0.0 ? log(s) + (double)K * 0.69314718055994530942 : -INFINITY;
            if (a.out_pair_logp && !cidx && p == 0) a.out_pair_logp[(size_t)r * a.S + i] = logp;
            acc.add(logp, exp(logp), kk, a.dsCount[i]);
            if (collect && p == 0) tie_note(&ties[group], acc.best_v, logp, kk, a.tie_abs, a.tie_rel);
        }
        __syncthreads();
    }
    merge_and_store<true>(a, r, ncand, cidx, NGR, group, p, acc, gres, ties, list);
}

// One workgroup per read of the bin's compacted list: a persistent loop over
// a queue costs this kernel registers (the compiler keeps read-invariant
// values live across reads), which lowers its occupancy.
// Waves per SIMD. The step is latency-bound at one wave per SIMD, so up to 3
// row slots (T <= 12) the bounds trade register spills for occupancy, as
// measured on 10^5 reads x 1,000 dye-seqs (C=3, T=12): boxes of up to 16
// states at 2 waves (unbounded, 390 registers and 1 wave; bounded, ~100
// spilled: 19.3 -> 11.3 ms), up to 8 states at 3 waves (253 registers and 2
// waves; bounded, 86 spilled: 8.4 -> 7.3 ms), up to 4 states at 2. More row
// slots would spill their state itself, so they keep 2 waves up to 8 states
// and 1 above.
template <int NB, int R>
constexpr int v5_min_waves() {
    if (R <= 3) return NB <= 4 ? 2 : NB <= 8 ? 3 : 2;
    return NB <= 8 ? 2 : 1;
}

template <int C, int NB, int R, bool kClass>
__global__ void __launch_bounds__(kV5Threads, (v5_min_waves<NB, R>())) classify_v5_kernel(const ClassifyArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    if ((int)blockIdx.x >= *a.list_count) return;
    v5_read<C, NB, R, kClass>(a, a.list[blockIdx.x], smem);
}

template <int C, int NB, int R, bool kClass>
hipError_t launch_v5_v(const ClassifyArgs& a, hipStream_t stream) {
    const V5Layout l = v5_layout(a.m.T, C, a.tri, NB);
    if (l.total > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute((const void*)classify_v5_kernel<C, NB, R, kClass>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)l.total);
        if (e != hipSuccess) return e;
    }
    // one workgroup per read of the bin's list (workgroups past its length exit)
    hipLaunchKernelGGL((classify_v5_kernel<C, NB, R, kClass>), dim3(a.N), dim3(kV5Threads), l.total, stream, a);
    return hipGetLastError();
}

template <int C, int NB, int R>
hipError_t launch_v5_t(const ClassifyArgs& a, hipStream_t stream) {
    return a.variant ? launch_v5_v<C, NB, R, true>(a, stream) : launch_v5_v<C, NB, R, false>(a, stream);
}

// Row-slots per lane: R = ceil(T / 4) rounded up to an instantiated count.
template <int C, int NB>
hipError_t launch_v5_nb(const ClassifyArgs& a, hipStream_t stream) {
    const int T = a.m.T;
    if (T <= 4) return launch_v5_t<C, NB, 1>(a, stream);
    if (T <= 8) return launch_v5_t<C, NB, 2>(a, stream);
    if (T <= 12) return launch_v5_t<C, NB, 3>(a, stream);
    if (T <= 16) return launch_v5_t<C, NB, 4>(a, stream);
    if constexpr (NB <= 8) {
        if (T <= 32) return launch_v5_t<C, NB, 8>(a, stream);
    }
    return hipErrorInvalidValue;
}

template <int C>
hipError_t launch_v5_c(const ClassifyArgs& a, hipStream_t stream) {
    switch (a.NB) {
    case 4: return launch_v5_nb<C, 4>(a, stream);
    case 8: return launch_v5_nb<C, 8>(a, stream);
    case 16: return launch_v5_nb<C, 16>(a, stream);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace

bool v5_supports(int precision, int T, int NB) {
    if (precision != WP_PRECISION_FP32) return false;
    if (NB == 4 || NB == 8) return T <= 32;
    return NB == 16 && T <= 16;
}

size_t v5_lds_bytes(int T, int C, int tri, int NB) { return v5_layout(T, C, tri, NB).total; }

hipError_t launch_classify_v5(const ClassifyArgs& a, hipStream_t stream) {
    if (a.N <= 0) return hipSuccess;
    switch (a.m.C) {
    case 1: return launch_v5_c<1>(a, stream);
    case 2: return launch_v5_c<2>(a, stream);
    case 3: return launch_v5_c<3>(a, stream);
    case 4: return launch_v5_c<4>(a, stream);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace wp
