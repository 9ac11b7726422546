// Fused PeptideHMM forward scoring for gfx950.
//
// Reference path (whatprot cc_code/src): HMMClassifier::classify_helper
// (classifiers/hmm-classifier.h:45-78) builds a PeptideHMM per (read,
// dye-seq) pair (hmm/hmm/peptide-hmm.cc:33-84) and runs the step chain
// InitialBlock, Dud_c, Emit(0), then per timestep t>=1: CyclicBlock,
// Detach(t-1), Bleach_c, Edman, Emit(t) (generic-hmm.h:41-55).
//
// Pruning in closed form. Running the per-step prune rules of
// peptide-hmm.cc:62-83 (block-transition.cc:22-29, binomial-transition.cc:77-89,
// detach-transition.cc:24-34, edman-transition.cc:28-60,
// peptide-emission.cc:198-210) symbolically gives, for the emission at
// timestep t, the box [0, t+1) x prod_c [L[t][c], H[t][c]) with
//   L[t][c] = max over t' >= t of lo_t'(c)          (read only)
//   H[t][c] = min(D_c + 1, min over t' <= t of hi_t'(c))
// where [lo_t(c), hi_t(c)) is PeptideEmission's own range. Bleach_c at
// timestep t maps channel c from [L[t-1][c], H[t-1][c]) onto
// [L[t][c], U[t][c]), U = min(H[t] + 1, H[t-1]); Edman reads
// [0,t) x [L[t], U[t]) and the emission keeps [0,t+1) x [L[t], H[t]). The pair
// is empty (P = 0) iff some L >= H. tests/test_boxes.py checks this against
// the oracle's literal step-by-step box algebra.
//
// One gather pass per timestep. Block, Detach, Bleach and Edman are linear, so
// for t >= 1 every output state (e', n') of the emission box X_t is computed
// directly from the previous state X_{t-1}:
//   Bl(X)[e, m] = sum_{n >= m, n in X_{t-1} box} prod_c Pbleach_c(n_c, m_c) X[e, n]
//   A_t[e', n'] = (1-pd)(1-pcb) [ pf Bl(A)[e', n']
//                 + (1-pf)(1 - n'_c/K) Bl(A)[e'-1, n']            (c = dye at e'-1)
//                 + (1-pf)((n'_c+1)/K) Bl(A)[e'-1, n'+1_c] ]       (if n'_c+1 < U_c)
//   B_t[e', n'] = (1-pd) (Bl(B)[e', n'] + pcb Bl(A)[e', n'])      (e' < t)
//   then times the emission density E_t(n'); the detached mass gets
//   pd * sum(A + B) of X_{t-1} (the previous output total) and E_t(0).
// No intermediate tensor is materialised, so a pair needs only its previous
// and current state, sized by the emission boxes.
//
// Mapping. A workgroup owns one read; its emission table, binomial tables and
// pruned bounds sit in LDS. Pairs are scored by groups of G = 16 lanes (four
// per wave); the lanes of a group stride over the flattened output states of
// each timestep and read the previous state from the group's LDS buffer. One
// wave-scope fence per timestep orders the group's writes before the next
// timestep's reads. Non-empty candidates are compacted first (73% of pairs are
// empty at a 5-sigma cutoff) so the groups of a wave work on live pairs.
//
// Precision. FP64 keeps the reference's arithmetic (doubles, no scaling).
// FP32 keeps the state in float; the emission table holds mantissa and
// exponent separately, and at every timestep the products (mass x density) are
// normalised by 2^-K with K the largest exponent among them, K accumulated in
// an integer. ln P, the argmax and the normaliser over dye-seqs are FP64.
#include <float.h>
#include <math.h>

#include <algorithm>

#include "classify_common.h"
#include "hmm_kernels.h"

// Register budget of the gather kernel: 3 waves per SIMD unless overridden.
// The strided C = 3 forms spill 78-93 dwords per lane at this bound; at 2
// waves per SIMD (FP64, 21 dwords) the near-tie decision was no faster.
#ifndef WP_WAVES_PER_EU
#define WP_WAVES_PER_EU 3
#endif
#if WP_WAVES_PER_EU > 0
#define WP_CLASSIFY_OCCUPANCY __attribute__((amdgpu_waves_per_eu(WP_WAVES_PER_EU, 8)))
#else
#define WP_CLASSIFY_OCCUPANCY
#endif

namespace wp {

namespace {

using namespace cc;

// LDS carve-up shared by host sizing and the kernel.
// Entries between the state buffers of consecutive pair groups: 4 * cap
// rounded so that group g starts 16 4-byte words (a quarter of the 64 banks)
// after group g - 1, and lanes of different groups reading the same offset of
// their buffers do not collide on one bank.
template <typename Real>
__host__ __device__ inline int group_stride(int cap) {
    constexpr int per = 64 * 4 / (int)sizeof(Real);  // entries per 64 banks
    return (4 * cap + per - 1) / per * per + per / 4;
}

struct Layout {
    size_t rad, L, H, eoff, estr, etab, eexp, binom, list, scan, gres, ties, gseq, state, total;
};

template <typename Real>
__host__ __device__ inline Layout make_layout(int T, int C, int tri, int NB, int cap, int ngroups, int G,
                                              int nthreads) {
    Layout l;
    size_t o = 0;
    l.rad = o;   o = align16(o + sizeof(double) * T * C);
    l.L = o;     o = align16(o + sizeof(int) * T * C);
    l.H = o;     o = align16(o + sizeof(int) * T * C);
    l.eoff = o;  o = align16(o + sizeof(int) * (T + 1));
    l.estr = o;  o = align16(o + sizeof(int) * T * C);
    l.etab = o;  o = align16(o + sizeof(Real) * (size_t)T * NB);
    l.eexp = o;  o = align16(o + (sizeof(Real) == 4 ? sizeof(short) * (size_t)T * NB : 0));
    l.binom = o; o = align16(o + sizeof(Real) * 2 * C * tri);
    l.list = o;  o = align16(o + sizeof(int) * nthreads * 4);
    l.scan = o;  o = align16(o + sizeof(int) * (nthreads / 64 + 2));
    l.gres = o;  o = align16(o + sizeof(double) * 4 * ngroups);
    l.ties = o;  o = align16(o + (sizeof(Real) == 4 ? sizeof(TieSlot) * ngroups : 0));
    l.gseq = o;  o = align16(o + sizeof(int) * 3 * T * ngroups);
    l.state = o; o = align16(o + sizeof(Real) * (size_t)group_stride<Real>(cap) * ngroups);
    l.total = o;
    return l;
}

// Ranges of one (read, timestep).
template <int C>
__global__ void range_kernel(DevModel m, int N, const double* __restrict__ rads,
                             uint8_t* __restrict__ lo_out, uint8_t* __restrict__ hi_out) {
    int idx = blockIdx.x * blockDim.x + threadIdx.x;
    const int T = m.T;
    if (idx >= N * T) return;
    double x[C];
#pragma unroll
    for (int c = 0; c < C; c++) x[c] = rads[(size_t)idx * C + c];
    unsigned int lo[C], hi[C];
    dev_emission_range<C>(m, x, lo, hi);
#pragma unroll
    for (int c = 0; c < C; c++) {
        lo_out[(size_t)idx * C + c] = (uint8_t)lo[c];
        hi_out[(size_t)idx * C + c] = (uint8_t)hi[c];
    }
}

// Turns per-timestep ranges into the pruned bounds in place (suffix max of
// lows, prefix min of highs capped at M+1) and the read's box bound.
__global__ void bounds_kernel(DevModel m, int N, uint8_t* __restrict__ L, uint8_t* __restrict__ H,
                              int* __restrict__ nbmax) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= N) return;
    const int T = m.T, C = m.C;
    uint8_t* l = L + (size_t)r * T * C;
    uint8_t* h = H + (size_t)r * T * C;
    for (int c = 0; c < C; c++) {
        for (int t = T - 2; t >= 0; t--) {
            uint8_t a = l[t * C + c], b = l[(t + 1) * C + c];
            l[t * C + c] = a > b ? a : b;
        }
        uint8_t cap = (uint8_t)(m.M + 1);
        if (h[c] > cap) h[c] = cap;
        for (int t = 1; t < T; t++) {
            uint8_t a = h[t * C + c], b = h[(t - 1) * C + c];
            h[t * C + c] = a < b ? a : b;
        }
    }
    // Capacity per pair: the largest emission box (the state of one timestep).
    int best = 0;
    bool empty = false;
    for (int t = 0; t < T; t++) {
        int nb = 1;
        for (int c = 0; c < C; c++) {
            int w = (int)h[t * C + c] - (int)l[t * C + c];
            if (w <= 0) empty = true;
            nb *= w > 0 ? w : 0;
        }
        if (nb > best) best = nb;
    }
    nbmax[r] = empty ? 0 : best;
}

template <typename Real>
struct PairResult {
    double logp;  // ln P (-inf when P == 0)
    double p;     // P itself (FP64 path; FP32 fills exp(logp))
};

template <int C>
__device__ inline int box_strides(const int* w, int* str) {
    int s = 1;
#pragma unroll
    for (int c = C - 1; c >= 0; c--) {
        str[c] = s;
        s *= w[c];
    }
    return s;
}

// Emission density of table entry ei. FP64 keeps the reference's value; FP32
// stores mantissa and exponent separately and returns the density times
// 2^-kmax, kmax being the largest exponent over the pair's count box at that
// timestep, so no candidate's box underflows whatever the read's range.
template <typename Real, bool kRescale>
__device__ inline Real emission(const Real* etab, const short* eexp, int ei, int kmax) {
    if constexpr (kRescale) return ldexpf(etab[ei], (int)eexp[ei] - kmax);
    else return etab[ei];
}

// Composite Bleach over channels c..C-1 evaluated at counts m: accumulates
// coef * prod P_c(n_c, m_c) * X[e, n] over n >= m inside the previous box
// (binomial-transition.cc:124-136 applied channel after channel).
template <typename Real, int C, int c, bool kWithB>
struct CompositeBleach {
    __device__ static inline void eval(const Real* A, const Real* B, int e, int es, const int* m,
                                       const int* lo, const int* hi, const int* str, const Real* ble,
                                       int tri, int idx, Real coef, Real& sa, Real& sb) {
        const int mc = m[c];
        for (int n = max(mc, lo[c]); n < hi[c]; n++) {
            Real p = ble[c * tri + n * (n + 1) / 2 + mc] * coef;
            CompositeBleach<Real, C, c + 1, kWithB>::eval(A, B, e, es, m, lo, hi, str, ble, tri,
                                                          idx + (n - lo[c]) * str[c], p, sa, sb);
        }
    }
};
template <typename Real, int C, bool kWithB>
struct CompositeBleach<Real, C, C, kWithB> {
    __device__ static inline void eval(const Real* A, const Real* B, int e, int es, const int*,
                                       const int*, const int*, const int*, const Real*, int, int idx,
                                       Real coef, Real& sa, Real& sb) {
        sa += coef * A[idx * es + e];
        if (kWithB) sb += coef * B[idx * es + e];
    }
};

// Forward algorithm of one (read, dye-seq) pair on one group of G lanes.
// st holds four buffers of cap = T * NB entries (A, B in; A, B out); gseq/gK
// hold the dye-seq's channel at each position e < T and K = dyes of that
// channel at positions >= e.
template <typename Real, int C, int G, bool kRescale>
__device__ PairResult<Real> score_pair(const ClassifyArgs& a, int i, int lane, Real* st, int cap,
                                       const int* sL, const int* sH, const int* eoff,
                                       const int* estr, const Real* etab, const short* eexp,
                                       const Real* bin, int* gseq, int* gK) {
    const int T = a.m.T;
    const int tri = a.tri;
    int D[C];
#pragma unroll
    for (int c = 0; c < C; c++) D[c] = a.dsD[(size_t)i * C + c];
    for (int e = lane; e < T; e += G) {
        int ch = a.dsSeq[(size_t)i * T + e];
        gseq[e] = ch;
        int K = ch >= 0 ? (int)a.dsTrack[((size_t)i * T + e) * C + ch] : 1;
        gK[e] = K;
        ((float*)(gK + T))[e] = 1.0f / (float)K;
    }
    const float* ginvK = (const float*)(gK + T);  // 1/K per position (FP32 path)
    const Real pf = (Real)a.m.pf;
    const Real one_m_pf = (Real)(1.0 - a.m.pf);
    const Real* dud = bin;
    const Real* ble = bin + C * tri;
    Real* Ain = st;
    Real* Bin = st + cap;
    Real* Aout = st + 2 * cap;
    Real* Bout = st + 3 * cap;

    // --- t = 0: InitialBlock, Dud_c, Emit(0) over [0,1) x [L0, H0) ---------
    int lo[C], hi[C], w[C], str[C];
#pragma unroll
    for (int c = 0; c < C; c++) {
        lo[c] = sL[c];
        hi[c] = min(sH[c], D[c] + 1);
        w[c] = hi[c] - lo[c];
    }
    int nb = box_strides<C>(w, str);
    int logacc = 0;
    Real S = 0;  // sum of A + B over the current state
    {
        const Real pib = (Real)a.m.p_ib;
        int K0 = 0;
        if constexpr (kRescale) {
            int km = kNoExp;
            for (int q = lane; q < nb; q += G) {
                int rem = q, ei = eoff[0];
                Real v = 1;
#pragma unroll
                for (int c = 0; c < C; c++) {
                    int n = lo[c] + rem / str[c];
                    rem -= (n - lo[c]) * str[c];
                    v = dud[c * tri + D[c] * (D[c] + 1) / 2 + n] * v;
                    ei += (n - sL[c]) * estr[c];
                }
                km = max(km, exp_of(v) + (int)eexp[ei]);
            }
            km = group_max_int<G>(km);
            K0 = km == kNoExp ? 0 : km;
            logacc += K0;
        }
        for (int q = lane; q < nb; q += G) {
            int rem = q, ei = eoff[0];
            Real va = (Real)1 - pib, vb = pib;
#pragma unroll
            for (int c = 0; c < C; c++) {
                int n = lo[c] + rem / str[c];
                rem -= (n - lo[c]) * str[c];
                Real p = dud[c * tri + D[c] * (D[c] + 1) / 2 + n];
                va = p * va;
                vb = p * vb;
                ei += (n - sL[c]) * estr[c];
            }
            Real em = emission<Real, kRescale>(etab, eexp, ei, K0);
            va *= em;
            vb *= em;
            Ain[q] = va;
            Bin[q] = vb;
            S += va + vb;
        }
        S = group_sum<Real, G>(S);
    }
    Real delta = 0;
    wave_sync();

    for (int t = 1; t < T; t++) {
        const int* Lt = sL + t * C;
        const int* Ht = sH + t * C;
        int lo2[C], hi2[C], U[C], w2[C], str2[C];
#pragma unroll
        for (int c = 0; c < C; c++) {
            lo2[c] = Lt[c];
            hi2[c] = min(Ht[c], D[c] + 1);
            U[c] = min(hi2[c] + 1, hi[c]);
            w2[c] = hi2[c] - lo2[c];
        }
        const int nb2 = box_strides<C>(w2, str2);
        const int nout = (t + 1) * nb2;
        const Divider by_e(t + 1);
        Divider by_str[C];
#pragma unroll
        for (int c = 0; c < C; c++) by_str[c] = Divider(str2[c]);
        const int eo = eoff[t];
        const int* es = estr + t * C;
        // Detach(t-1) adds pd times the previous total (block keeps A + B).
        const Real pd = (Real)a.m.p_detach[t - 1];
        const Real pcb = (Real)a.m.p_cb;
        delta = delta + pd * S;
        const Real cAA = ((Real)1 - pd) * ((Real)1 - pcb);
        const Real cB = (Real)1 - pd;
        int km = kNoExp;
        Real Snew = 0;
        for (int q = lane; q < nout; q += G) {
            const int ni = by_e.div(q);
            const int e = q - ni * (t + 1);
            int n[C];
            int rem = ni, ei = eo;
#pragma unroll
            for (int c = 0; c < C; c++) {
                const int qc = by_str[c].div(rem);
                n[c] = lo2[c] + qc;
                rem -= qc * str2[c];
                ei += (n[c] - Lt[c]) * es[c];
            }
            Real blA = 0, blB = 0;
            if (e < t)
                CompositeBleach<Real, C, 0, true>::eval(Ain, Bin, e, t, n, lo, hi, str, ble, tri, 0,
                                                        (Real)1, blA, blB);
            Real acc = 0, dummy = 0;
            if (e >= 1) {
                const int ce = gseq[e - 1];
                Real p1 = 0;
                CompositeBleach<Real, C, 0, false>::eval(Ain, Bin, e - 1, t, n, lo, hi, str, ble, tri, 0,
                                                         (Real)1, p1, dummy);
                if (ce < 0) {
                    acc += one_m_pf * p1;
                } else {
                    const int Ke = gK[e - 1];
                    int kc = 0, Uc = 0;
#pragma unroll
                    for (int c = 0; c < C; c++)
                        if (c == ce) {
                            kc = n[c];
                            Uc = U[c];
                        }
                    Real ratio, ratio2;
                    if constexpr (kRescale) {
                        const float inv = ginvK[e - 1];
                        ratio = (Real)kc * inv;
                        ratio2 = (Real)(kc + 1) * inv;
                    } else {
                        ratio = (Real)kc / (Real)Ke;
                        ratio2 = (Real)(kc + 1) / (Real)Ke;
                    }
                    if (kc < Ke) acc += one_m_pf * ((Real)1 - ratio) * p1;
                    if (kc + 1 < Uc) {
                        int n1[C];
#pragma unroll
                        for (int c = 0; c < C; c++) n1[c] = n[c] + (c == ce ? 1 : 0);
                        Real p2 = 0;
                        CompositeBleach<Real, C, 0, false>::eval(Ain, Bin, e - 1, t, n1, lo, hi, str, ble,
                                                                 tri, 0, (Real)1, p2, dummy);
                        acc += one_m_pf * ratio2 * p2;
                    }
                }
            }
            acc += pf * blA;
            Real ao = cAA * acc;
            Real bo = e < t ? cB * (blB + pcb * blA) : (Real)0;
            if constexpr (kRescale) {
                Aout[q] = ao;
                Bout[q] = bo;
                km = max(km, max(exp_of(ao), exp_of(bo)) + (int)eexp[ei]);
            } else {
                Real em = etab[ei];
                ao *= em;
                bo *= em;
                Aout[q] = ao;
                Bout[q] = bo;
                Snew += ao + bo;
            }
        }
        bool allow = true;
#pragma unroll
        for (int c = 0; c < C; c++) allow = allow && lo2[c] == 0;
        int ei0 = eo;
#pragma unroll
        for (int c = 0; c < C; c++) ei0 -= Lt[c] * es[c];
        if constexpr (kRescale) {
            km = group_max_int<G>(km);
            if (allow) km = max(km, exp_of(delta) + (int)eexp[ei0]);
            const int K = km == kNoExp ? 0 : km;
            logacc += K;
            for (int q = lane; q < nout; q += G) {
                const int ni = by_e.div(q);
                int rem = ni, ei = eo;
#pragma unroll
                for (int c = 0; c < C; c++) {
                    const int nc = by_str[c].div(rem);
                    rem -= nc * str2[c];
                    ei += (nc + lo2[c] - Lt[c]) * es[c];
                }
                Real em = emission<Real, kRescale>(etab, eexp, ei, K);
                Real ao = Aout[q] * em, bo = Bout[q] * em;
                Aout[q] = ao;
                Bout[q] = bo;
                Snew += ao + bo;
            }
            delta = allow ? delta * emission<Real, kRescale>(etab, eexp, ei0, K) : (Real)0;
        } else {
            delta = allow ? delta * etab[ei0] : (Real)0;
        }
        S = group_sum<Real, G>(Snew);
        wave_sync();
        Real* tA = Ain; Ain = Aout; Aout = tA;
        Real* tB = Bin; Bin = Bout; Bout = tB;
#pragma unroll
        for (int c = 0; c < C; c++) {
            lo[c] = lo2[c];
            hi[c] = hi2[c];
            w[c] = w2[c];
            str[c] = str2[c];
        }
        nb = nb2;
    }

    // PeptideStateVector::sum (peptide-state-vector.cc:52-54): S + delta.
    PairResult<Real> res;
    if (kRescale) {
        double s = (double)S + (double)delta;
        res.logp = s > 0 ? log(s) + (double)logacc * 0.69314718055994530942 : -INFINITY;
        res.p = exp(res.logp);
    } else {
        double s = (double)S + (double)delta;
        res.p = s;
        res.logp = log(s);
    }
    wave_sync();
    return res;
}

// One channel's Bleach (binomial-transition.cc:91-136) as a pass over a
// group's state: out[.., j, ..] = sum_{i >= j} P_c(i, j) in[.., i, ..] along
// axis c, every other axis unchanged, rows e < es. Boxes are absolute count
// ranges; arrays are [box index][e].
template <typename Real, int C, int G, int c>
__device__ inline void bleach_pass(const Real* inA, const Real* inB, Real* outA, Real* outB, int es,
                                   const int* ilo, const int* ihi, const int* istr, const int* olo,
                                   const int* ohi, const Real* ble, int tri, int lane) {
    int ow[C], ostr[C];
#pragma unroll
    for (int k = 0; k < C; k++) ow[k] = ohi[k] - olo[k];
    const int nbox = box_strides<C>(ow, ostr);
    const int nout = nbox * es;
    const Divider by_e(es);
    Divider by_str[C];
#pragma unroll
    for (int k = 0; k < C; k++) by_str[k] = Divider(ostr[k]);
    for (int q = lane; q < nout; q += G) {
        const int ni = by_e.div(q);
        const int e = q - ni * es;
        int rem = ni, base = 0, j = 0;
#pragma unroll
        for (int k = 0; k < C; k++) {
            const int m = by_str[k].div(rem);
            rem -= m * ostr[k];
            if (k == c) j = olo[k] + m;
            else base += m * istr[k];  // same range on both sides
        }
        Real sa = 0, sb = 0;
        const Real* row = ble + c * tri;
        for (int n = max(j, ilo[c]); n < ihi[c]; n++) {
            const Real pw = row[n * (n + 1) / 2 + j];
            const int idx = (base + (n - ilo[c]) * istr[c]) * es + e;
            sa += pw * inA[idx];
            sb += pw * inB[idx];
        }
        outA[q] = sa;
        outB[q] = sb;
    }
}

template <typename Real, int C, int G, int c>
struct BleachPasses {
    // Applies channels c..C-1 in turn, ping-ponging between (pA, pB) and
    // (qA, qB); the box of channels < c is already the output box. Returns
    // true when the result sits in (qA, qB).
    __device__ static inline bool run(Real* pA, Real* pB, Real* qA, Real* qB, int es, int* lo, int* hi,
                                      const int* lo2, const int* U, const Real* ble, int tri, int lane) {
        int w[C], str[C], nlo[C], nhi[C];
#pragma unroll
        for (int k = 0; k < C; k++) {
            w[k] = hi[k] - lo[k];
            nlo[k] = k == c ? lo2[k] : lo[k];
            nhi[k] = k == c ? U[k] : hi[k];
        }
        box_strides<C>(w, str);
        bleach_pass<Real, C, G, c>(pA, pB, qA, qB, es, lo, hi, str, nlo, nhi, ble, tri, lane);
        wave_sync();
        lo[c] = lo2[c];
        hi[c] = U[c];
        return !BleachPasses<Real, C, G, c + 1>::run(qA, qB, pA, pB, es, lo, hi, lo2, U, ble, tri, lane);
    }
};
template <typename Real, int C, int G>
struct BleachPasses<Real, C, G, C> {
    __device__ static inline bool run(Real*, Real*, Real*, Real*, int, int*, int*, const int*, const int*,
                                      const Real*, int, int) {
        return false;
    }
};

// score_pair with the composite Bleach applied as C separable passes: per
// step, the state is bleached channel by channel onto the box
// X = prod_c [lo2_c, U_c) (the next box plus the Edman source row n + 1_c),
// then Edman and Emit read X. Same recursion and scaling as score_pair, with
// per-state work proportional to the box widths instead of the box sizes.
// Every array of a step must fit in cap entries.
template <typename Real, int C, int G, bool kRescale>
__device__ PairResult<Real> score_pair_sep(const ClassifyArgs& a, int i, int lane, Real* st, int cap,
                                           const int* sL, const int* sH, const int* eoff,
                                           const int* estr, const Real* etab, const short* eexp,
                                           const Real* bin, int* gseq, int* gK) {
    const int T = a.m.T;
    const int tri = a.tri;
    int D[C];
#pragma unroll
    for (int c = 0; c < C; c++) D[c] = a.dsD[(size_t)i * C + c];
    for (int e = lane; e < T; e += G) {
        int ch = a.dsSeq[(size_t)i * T + e];
        gseq[e] = ch;
        int K = ch >= 0 ? (int)a.dsTrack[((size_t)i * T + e) * C + ch] : 1;
        gK[e] = K;
        ((float*)(gK + T))[e] = 1.0f / (float)K;
    }
    const float* ginvK = (const float*)(gK + T);
    const Real pf = (Real)a.m.pf;
    const Real one_m_pf = (Real)(1.0 - a.m.pf);
    const Real* dud = bin;
    const Real* ble = bin + C * tri;
    Real* curA = st;
    Real* curB = st + cap;
    Real* altA = st + 2 * cap;
    Real* altB = st + 3 * cap;

    // --- t = 0: InitialBlock, Dud_c, Emit(0) over [0,1) x [L0, H0) ---------
    int lo[C], hi[C], w[C], str[C];
#pragma unroll
    for (int c = 0; c < C; c++) {
        lo[c] = sL[c];
        hi[c] = min(sH[c], D[c] + 1);
        w[c] = hi[c] - lo[c];
    }
    int nb = box_strides<C>(w, str);
    int logacc = 0;
    Real S = 0;
    {
        const Real pib = (Real)a.m.p_ib;
        int K0 = 0;
        if constexpr (kRescale) {
            int km = kNoExp;
            for (int q = lane; q < nb; q += G) {
                int rem = q, ei = eoff[0];
                Real v = 1;
#pragma unroll
                for (int c = 0; c < C; c++) {
                    int n = lo[c] + rem / str[c];
                    rem -= (n - lo[c]) * str[c];
                    v = dud[c * tri + D[c] * (D[c] + 1) / 2 + n] * v;
                    ei += (n - sL[c]) * estr[c];
                }
                km = max(km, exp_of(v) + (int)eexp[ei]);
            }
            km = group_max_int<G>(km);
            K0 = km == kNoExp ? 0 : km;
            logacc += K0;
        }
        for (int q = lane; q < nb; q += G) {
            int rem = q, ei = eoff[0];
            Real va = (Real)1 - pib, vb = pib;
#pragma unroll
            for (int c = 0; c < C; c++) {
                int n = lo[c] + rem / str[c];
                rem -= (n - lo[c]) * str[c];
                Real p = dud[c * tri + D[c] * (D[c] + 1) / 2 + n];
                va = p * va;
                vb = p * vb;
                ei += (n - sL[c]) * estr[c];
            }
            Real em = emission<Real, kRescale>(etab, eexp, ei, K0);
            va *= em;
            vb *= em;
            curA[q] = va;
            curB[q] = vb;
            S += va + vb;
        }
        S = group_sum<Real, G>(S);
    }
    Real delta = 0;
    wave_sync();

    for (int t = 1; t < T; t++) {
        const int* Lt = sL + t * C;
        const int* Ht = sH + t * C;
        int lo2[C], hi2[C], U[C], w2[C], str2[C];
#pragma unroll
        for (int c = 0; c < C; c++) {
            lo2[c] = Lt[c];
            hi2[c] = min(Ht[c], D[c] + 1);
            U[c] = min(hi2[c] + 1, hi[c]);
            w2[c] = hi2[c] - lo2[c];
        }
        // Bleach_c for every channel: (cur) -> X, rows e < t
        int xlo[C], xhi[C], xw[C], xstr[C];
#pragma unroll
        for (int c = 0; c < C; c++) {
            xlo[c] = lo[c];
            xhi[c] = hi[c];
        }
        const bool in_alt = BleachPasses<Real, C, G, 0>::run(curA, curB, altA, altB, t, xlo, xhi, lo2, U, ble,
                                                             tri, lane);
        const Real* XA = in_alt ? altA : curA;
        const Real* XB = in_alt ? altB : curB;
        Real* Aout = in_alt ? curA : altA;
        Real* Bout = in_alt ? curB : altB;
#pragma unroll
        for (int c = 0; c < C; c++) xw[c] = xhi[c] - xlo[c];
        box_strides<C>(xw, xstr);

        const int nb2 = box_strides<C>(w2, str2);
        const int nout = (t + 1) * nb2;
        const Divider by_e(t + 1);
        Divider by_str[C];
#pragma unroll
        for (int c = 0; c < C; c++) by_str[c] = Divider(str2[c]);
        const int eo = eoff[t];
        const int* es = estr + t * C;
        const Real pd = (Real)a.m.p_detach[t - 1];
        const Real pcb = (Real)a.m.p_cb;
        delta = delta + pd * S;
        const Real cAA = ((Real)1 - pd) * ((Real)1 - pcb);
        const Real cB = (Real)1 - pd;
        int km = kNoExp;
        Real Snew = 0;
        // Edman (edman-transition.cc:62-162) and Emit(t), reading X
        for (int q = lane; q < nout; q += G) {
            const int ni = by_e.div(q);
            const int e = q - ni * (t + 1);
            int n[C];
            int rem = ni, ei = eo, xi = 0;
#pragma unroll
            for (int c = 0; c < C; c++) {
                const int qc = by_str[c].div(rem);
                n[c] = lo2[c] + qc;
                rem -= qc * str2[c];
                ei += (n[c] - Lt[c]) * es[c];
                xi += qc * xstr[c];
            }
            const Real blA = e < t ? XA[xi * t + e] : (Real)0;
            const Real blB = e < t ? XB[xi * t + e] : (Real)0;
            Real acc = 0;
            if (e >= 1) {
                const int ce = gseq[e - 1];
                const Real p1 = XA[xi * t + e - 1];
                if (ce < 0) {
                    acc += one_m_pf * p1;
                } else {
                    const int Ke = gK[e - 1];
                    int kc = 0, Uc = 0, sc = 0;
#pragma unroll
                    for (int c = 0; c < C; c++)
                        if (c == ce) {
                            kc = n[c];
                            Uc = U[c];
                            sc = xstr[c];
                        }
                    Real ratio, ratio2;
                    if constexpr (kRescale) {
                        const float inv = ginvK[e - 1];
                        ratio = (Real)kc * inv;
                        ratio2 = (Real)(kc + 1) * inv;
                    } else {
                        ratio = (Real)kc / (Real)Ke;
                        ratio2 = (Real)(kc + 1) / (Real)Ke;
                    }
                    if (kc < Ke) acc += one_m_pf * ((Real)1 - ratio) * p1;
                    if (kc + 1 < Uc) acc += one_m_pf * ratio2 * XA[(xi + sc) * t + e - 1];
                }
            }
            acc += pf * blA;
            Real ao = cAA * acc;
            Real bo = e < t ? cB * (blB + pcb * blA) : (Real)0;
            if constexpr (kRescale) {
                Aout[q] = ao;
                Bout[q] = bo;
                km = max(km, max(exp_of(ao), exp_of(bo)) + (int)eexp[ei]);
            } else {
                Real em = etab[ei];
                ao *= em;
                bo *= em;
                Aout[q] = ao;
                Bout[q] = bo;
                Snew += ao + bo;
            }
        }
        bool allow = true;
#pragma unroll
        for (int c = 0; c < C; c++) allow = allow && lo2[c] == 0;
        int ei0 = eo;
#pragma unroll
        for (int c = 0; c < C; c++) ei0 -= Lt[c] * es[c];
        if constexpr (kRescale) {
            km = group_max_int<G>(km);
            if (allow) km = max(km, exp_of(delta) + (int)eexp[ei0]);
            const int K = km == kNoExp ? 0 : km;
            logacc += K;
            for (int q = lane; q < nout; q += G) {
                const int ni = by_e.div(q);
                int rem = ni, ei = eo;
#pragma unroll
                for (int c = 0; c < C; c++) {
                    const int nc = by_str[c].div(rem);
                    rem -= nc * str2[c];
                    ei += (nc + lo2[c] - Lt[c]) * es[c];
                }
                Real em = emission<Real, kRescale>(etab, eexp, ei, K);
                Real ao = Aout[q] * em, bo = Bout[q] * em;
                Aout[q] = ao;
                Bout[q] = bo;
                Snew += ao + bo;
            }
            delta = allow ? delta * emission<Real, kRescale>(etab, eexp, ei0, K) : (Real)0;
        } else {
            delta = allow ? delta * etab[ei0] : (Real)0;
        }
        S = group_sum<Real, G>(Snew);
        wave_sync();
        // the new state is in (Aout, Bout); the other pair is free
        Real* fa = in_alt ? altA : curA;
        Real* fb = in_alt ? altB : curB;
        curA = Aout;
        curB = Bout;
        altA = fa;
        altB = fb;
#pragma unroll
        for (int c = 0; c < C; c++) {
            lo[c] = lo2[c];
            hi[c] = hi2[c];
            w[c] = w2[c];
            str[c] = str2[c];
        }
        nb = nb2;
    }

    PairResult<Real> res;
    if (kRescale) {
        double s = (double)S + (double)delta;
        res.logp = s > 0 ? log(s) + (double)logacc * 0.69314718055994530942 : -INFINITY;
        res.p = exp(res.logp);
    } else {
        double s = (double)S + (double)delta;
        res.p = s;
        res.logp = log(s);
    }
    wave_sync();
    return res;
}

// Entries a pair's forward pass needs in each state buffer, with the box
// bounds capped at D_c + 1 (cap[c]; pass a large cap for the read-level
// bound of every candidate): (t + 1) * |box(t)| for the state, and with
// kSepPasses also t * |box| for each intermediate box of the separable
// passes of step t.
template <int C, bool kSepPasses>
__device__ inline int state_need(int T, const int* sL, const int* sH, const int* cap) {
    int need = 0;
    for (int t = 0; t < T; t++) {
        int s = t + 1;
#pragma unroll
        for (int c = 0; c < C; c++) s *= max(min(sH[t * C + c], cap[c]) - sL[t * C + c], 0);
        need = max(need, s);
        if (!kSepPasses || t == 0) continue;
#pragma unroll
        for (int k = 0; k < C; k++) {  // after the pass of channel k
            int v = t;
#pragma unroll
            for (int c = 0; c < C; c++) {
                const int h1 = min(sH[(t - 1) * C + c], cap[c]), l1 = sL[(t - 1) * C + c];
                const int u = min(min(sH[t * C + c], cap[c]) + 1, h1);
                v *= c <= k ? max(u - sL[t * C + c], 0) : max(h1 - l1, 0);
            }
            need = max(need, v);
        }
    }
    return need;
}

// One read of the gather kernel. Block-uniform returns only.
template <typename Real, int C, int G, bool kSep>
__device__ __forceinline__ void gather_read(const ClassifyArgs& a, const int r, unsigned char* smem) {
    constexpr bool kRescale = sizeof(Real) == 4;
    if (r >= a.N) return;
    const int nbr = a.nbmax[r];
    // Empty reads go with the first bin (nb_lo == 0 there).
    if (!((nbr > a.nb_lo && nbr <= a.nb_hi) || (nbr == 0 && a.nb_lo == 0))) return;
    if (a.wide_pass && !a.wide[r]) return;

    const int T = a.m.T;
    const int tid = threadIdx.x;
    const int nthreads = blockDim.x;
    const int ngroups = a.ngroups;
    const int group = tid / G;
    const int lane = tid % G;
    const Layout lay = make_layout<Real>(T, C, a.tri, a.NB, a.cap, ngroups, G, nthreads);
    double* s_rad = (double*)(smem + lay.rad);
    int* sL = (int*)(smem + lay.L);
    int* sH = (int*)(smem + lay.H);
    int* eoff = (int*)(smem + lay.eoff);
    int* estr = (int*)(smem + lay.estr);
    Real* etab = (Real*)(smem + lay.etab);
    short* eexp = (short*)(smem + lay.eexp);
    Real* bin = (Real*)(smem + lay.binom);
    int* list = (int*)(smem + lay.list);
    int* scan = (int*)(smem + lay.scan);
    double* gres = (double*)(smem + lay.gres);
    TieSlot* ties = (TieSlot*)(smem + lay.ties);
    int* gseq_all = (int*)(smem + lay.gseq);  // per group: seq[T], K[T], 1/K[T]
    Real* state = (Real*)(smem + lay.state);

    int ncand;
    const int* cidx;
    read_candidates(a, r, &ncand, &cidx);

    // ---- prologue: read data, bounds, emission table, binomial tables ----
    for (int k = tid; k < T * C; k += nthreads) {
        s_rad[k] = a.rads[(size_t)r * T * C + k];
        sL[k] = a.Lr[(size_t)r * T * C + k];
        sH[k] = a.Hr[(size_t)r * T * C + k];
    }
    for (int k = tid; k < 2 * C * a.tri; k += nthreads) bin[k] = (Real)a.binom[k];
    __syncthreads();
    if (tid == 0) {
        int acc = 0;
        for (int t = 0; t < T; t++) {
            eoff[t] = acc;
            int s = 1;
            for (int c = C - 1; c >= 0; c--) {
                estr[t * C + c] = s;
                int wdt = sH[t * C + c] - sL[t * C + c];
                s *= wdt > 0 ? wdt : 0;
            }
            acc += s;
        }
        eoff[T] = acc;
    }
    __syncthreads();
    const int nent = eoff[T];
    // emission densities over each timestep's box (peptide-emission.cc:165-180)
    for (int k = tid; k < nent; k += nthreads) {
        int t = 0;
        while (eoff[t + 1] <= k) t++;
        int rem = k - eoff[t];
        unsigned int n[WP_MAX_CHANNELS];
#pragma unroll
        for (int c = 0; c < C; c++) {
            int s = estr[t * C + c];
            n[c] = (unsigned int)(sL[t * C + c] + rem / s);
            rem %= s;
        }
        double v = 1.0;
#pragma unroll
        for (int c = 0; c < C; c++) v *= dev_pdf<C>(a.m, c, s_rad[t * C + c], n);
        if constexpr (kRescale) {
            int ex = 0;
            double mant = frexp(v, &ex);
            etab[k] = (Real)mant;
            eexp[k] = v > 0.0 ? (short)ex : (short)-30000;
        } else {
            etab[k] = (Real)v;
        }
    }
    __syncthreads();
    bool read_empty = (nbr == 0);
    // The pair is non-empty iff D_c >= L[0][c] for every channel.

    CandAcc<kRescale> acc;
    const bool collect = kRescale && a.tie_mode == 1;
    if (collect && group < ngroups && lane == 0) tie_init(&ties[group]);
    const int cap = a.cap;
    if constexpr (kSep) {
        // a read one of whose candidates could outgrow the state buffers even
        // in the composite form is handed to this bin's wide pass; each pair
        // of the others runs separable if its own passes fit, composite if not
        if (tid == 0) {
            int nocap[C];
#pragma unroll
            for (int c = 0; c < C; c++) nocap[c] = 1 << 20;
            scan[0] = state_need<C, false>(T, sL, sH, nocap) > cap ? 1 : 0;
        }
        __syncthreads();
        const bool over = scan[0] != 0;
        __syncthreads();
        if (over) {
            if (tid == 0) {
                a.wide[r] = 1;
                a.wide_list[atomicAdd(a.wide_count, 1)] = r;
            }
            return;
        }
    }
    Real* st = state + (size_t)group * group_stride<Real>(cap);
    int* gseq = gseq_all + group * 3 * T;
    int* gK = gseq + T;
    for (int base = 0; base < ncand; base += nthreads * 4) {
        const int nlive = compact_chunk<C>(a, r, ncand, cidx, base, read_empty, sL, list, scan);
        if (group < ngroups) {
            for (int q = group; q < nlive; q += ngroups) {
                const int kk = list[q];
                const int i = cidx ? cidx[kk] : kk;
                PairResult<Real> pr;
                bool sep = kSep;
                if constexpr (kSep) {
                    int dcap[C];
#pragma unroll
                    for (int c = 0; c < C; c++) dcap[c] = (int)a.dsD[(size_t)i * C + c] + 1;
                    sep = state_need<C, true>(T, sL, sH, dcap) <= cap;  // uniform over the group
                }
                if (sep)
                    pr = score_pair_sep<Real, C, G, kRescale>(a, i, lane, st, cap, sL, sH, eoff, estr, etab, eexp,
                                                              bin, gseq, gK);
                else
                    pr = score_pair<Real, C, G, kRescale>(a, i, lane, st, cap, sL, sH, eoff, estr, etab, eexp,
                                                          bin, gseq, gK);
                if (a.out_pair_logp && !cidx && lane == 0) a.out_pair_logp[(size_t)r * a.S + i] = pr.logp;
                acc.add(pr.logp, pr.p, kk, a.dsCount[i]);
                if (collect && lane == 0) tie_note(&ties[group], acc.best_v, pr.logp, kk, a.tie_abs, a.tie_rel);
            }
        }
        __syncthreads();
    }
    merge_and_store<kRescale>(a, r, ncand, cidx, ngroups, group, lane, acc, gres, ties, list);
}

// kStrided false: one workgroup per read of the bin's compacted list
// (workgroups past its length exit). kStrided: workgroups stride over the
// list, for lists that hold few of the call's reads; the barrier keeps a
// read's LDS tables from being overwritten while the previous read's outputs
// are written. Two instantiations: the loop holds read-invariant values
// across reads, which costs registers where every read of a large list has
// its own workgroup anyway.
template <typename Real, int C, int G, bool kSep, bool kStrided>
__global__ void __launch_bounds__(kBlockThreads) WP_CLASSIFY_OCCUPANCY
classify_kernel(const ClassifyArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int n = *a.list_count;
    if constexpr (!kStrided) {
        if ((int)blockIdx.x < n) gather_read<Real, C, G, kSep>(a, a.list[blockIdx.x], smem);
    } else {
        for (int k = blockIdx.x; k < n; k += gridDim.x) {
            if (k != (int)blockIdx.x) __syncthreads();
            gather_read<Real, C, G, kSep>(a, a.list[k], smem);
        }
    }
}

template <typename Real, int C, int G, bool kSep>
hipError_t launch_t(const ClassifyArgs& a, hipStream_t stream) {
    int nthreads = a.ngroups * G;
    if (nthreads < 64) nthreads = 64;
    nthreads = (nthreads + 63) / 64 * 64;
    Layout l = make_layout<Real>(a.m.T, C, a.tri, a.NB, a.cap, a.ngroups, G, nthreads);
    const bool strided = list_grid(a) < a.N;
    const void* fn = strided ? (const void*)classify_kernel<Real, C, G, kSep, true>
                             : (const void*)classify_kernel<Real, C, G, kSep, false>;
    if (l.total > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)l.total);
        if (e != hipSuccess) return e;
    }
    if (strided)
        hipLaunchKernelGGL((classify_kernel<Real, C, G, kSep, true>), dim3(list_grid(a)), dim3(nthreads), l.total, stream,
                           a);
    else
        hipLaunchKernelGGL((classify_kernel<Real, C, G, kSep, false>), dim3(a.N), dim3(nthreads), l.total, stream, a);
    return hipGetLastError();
}

template <typename Real, int C>
hipError_t launch_c(const ClassifyArgs& a, int G, hipStream_t stream) {
    if (G != kGroupLanes) return hipErrorInvalidValue;
    return a.variant ? launch_t<Real, C, kGroupLanes, true>(a, stream) : launch_t<Real, C, kGroupLanes, false>(a, stream);
}

template <typename Real>
hipError_t launch_r(const ClassifyArgs& a, int G, hipStream_t stream) {
    switch (a.m.C) {
    case 1: return launch_c<Real, 1>(a, G, stream);
    case 2: return launch_c<Real, 2>(a, G, stream);
    case 3: return launch_c<Real, 3>(a, G, stream);
    case 4: return launch_c<Real, 4>(a, G, stream);
    default: return hipErrorInvalidValue;
    }
}

// Algorithmic work of a batch: per read, the number of non-empty pairs, the
// pruned state-timesteps sum_t S_t with S_t = (t+1) prod_c (H - L), and the
// FLOP count F = S_0 (5 + sum_c 2(D_c+2)) + sum_{t>=1} S_t (17 + sum_c 2(D_c+2))
// (emission 2, block 3, detach 4, Edman 8, binomial 2(D_c+2) per state).
__global__ void work_kernel(DevModel m, int N, int S, const uint8_t* __restrict__ Lr,
                            const uint8_t* __restrict__ Hr, const int* __restrict__ nbmax,
                            const uint8_t* __restrict__ dsD, double* __restrict__ out) {
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= N) return;
    const int T = m.T, C = m.C;
    double pairs = 0, states = 0, flops = 0;
    if (nbmax[r] > 0) {
        const uint8_t* l = Lr + (size_t)r * T * C;
        const uint8_t* h = Hr + (size_t)r * T * C;
        for (int i = 0; i < S; i++) {
            bool ok = true;
            int bino = 0;
            for (int c = 0; c < C; c++) {
                int D = dsD[(size_t)i * C + c];
                ok = ok && D >= l[c];
                bino += 2 * (D + 2);
            }
            if (!ok) continue;
            pairs += 1;
            for (int t = 0; t < T; t++) {
                double nb = t + 1;
                for (int c = 0; c < C; c++) {
                    int D = dsD[(size_t)i * C + c];
                    int hh = min((int)h[t * C + c], D + 1);
                    nb *= (double)(hh - (int)l[t * C + c]);
                }
                states += nb;
                flops += nb * (double)((t == 0 ? 5 : 17) + bino);
            }
        }
    }
    out[(size_t)r * 3 + 0] = pairs;
    out[(size_t)r * 3 + 1] = states;
    out[(size_t)r * 3 + 2] = flops;
}

__global__ void zero_ints_kernel(int* d, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d[i] = 0;
}

// Appends read r to its bin's list. The order inside a list follows the
// atomics; every read's outputs are computed independently of that order.
// Appends read r (lane's b >= 0) to list b: one atomic per wave and bin
// (the lanes of a bin take consecutive slots in lane order). Every lane of
// the wave calls it; b < 0 appends nothing.
__device__ inline void append_wave(int b, int r, int N, int* __restrict__ lists, int* __restrict__ counts) {
    const int lane = (int)__lane_id();
    unsigned long long pending = __ballot(b >= 0);
    while (pending) {
        const int leader = __ffsll((long long)pending) - 1;
        const int lb = __shfl(b, leader);
        const unsigned long long mask = __ballot(b == lb);
        int base = 0;
        if (lane == leader) base = atomicAdd(&counts[lb], (int)__popcll(mask));
        base = __shfl(base, leader);
        if (b == lb) lists[(size_t)lb * N + base + (int)__popcll(mask & ((1ull << lane) - 1ull))] = r;
        pending &= ~mask;
    }
}

__global__ void bin_reads_kernel(int N, const int* __restrict__ nbmax, BinEdges e, int* __restrict__ lists,
                                 int* __restrict__ counts) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    int b = -1;  // outside every bin (or past N): no launch takes it
    if (r < N) {
        const int nb = nbmax[r];
        b = 0;
        if (nb > 0) {
            b = -1;
            for (int k = 0; k < e.nbins; k++)
                if (nb > e.edge[k] && nb <= e.edge[k + 1]) b = k;
        }
    }
    append_wave(b, r, N, lists, counts);
}
__global__ void bin_refine_kernel(int N, const int* __restrict__ nbmax, const int* __restrict__ tie_n, BinEdges e,
                                  int* __restrict__ lists, int* __restrict__ counts) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    int b = -1;
    if (r < N && tie_n[r] != 0) {
        const int nb = nbmax[r];
        for (int k = 0; k < e.nbins; k++)
            if (nb > e.edge[k] && nb <= e.edge[k + 1]) b = k;
    }
    append_wave(b, r, N, lists, counts);
}
__global__ void refine_fix_kernel(int N, const int* __restrict__ tie_n, const int* __restrict__ rid,
                                  const double* __restrict__ rls, const double* __restrict__ lt, int* __restrict__ id,
                                  double* __restrict__ adj, double* __restrict__ ls) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= N || tie_n[r] == 0) return;
    double s = rls[r];
    double t = lt[r];
    double v = exp(s - t);
    if (isnan(v)) {
        v = 0.0;
        s = -INFINITY;
    }
    id[r] = rid[r];
    if (adj) adj[r] = v;
    if (ls) ls[r] = s;
}

}  // namespace

void launch_bin_refine(int N, const int* d_nbmax, const int* d_tie_n, const BinEdges& e, int* d_lists,
                       int* d_counts, hipStream_t stream) {
    if (N > 0)
        hipLaunchKernelGGL(bin_refine_kernel, dim3((N + 255) / 256), dim3(256), 0, stream, N, d_nbmax, d_tie_n, e,
                           d_lists, d_counts);
}

void launch_refine_fix(int N, const int* d_tie_n, const int* d_rid, const double* d_rls, const double* d_lt,
                       int* d_id, double* d_adj, double* d_ls, hipStream_t stream) {
    if (N > 0)
        hipLaunchKernelGGL(refine_fix_kernel, dim3((N + 255) / 256), dim3(256), 0, stream, N, d_tie_n, d_rid, d_rls,
                           d_lt, d_id, d_adj, d_ls);
}

void launch_zero_ints(int* d, int n, hipStream_t stream) {
    if (n > 0) hipLaunchKernelGGL(zero_ints_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, d, n);
}

void launch_bin_reads(int N, const int* d_nbmax, const BinEdges& e, int* d_lists, int* d_counts,
                      hipStream_t stream) {
    if (N > 0)
        hipLaunchKernelGGL(bin_reads_kernel, dim3((N + 255) / 256), dim3(256), 0, stream, N, d_nbmax, e, d_lists,
                           d_counts);
}

void launch_read_prep(const DevModel& m, int N, const double* d_rads, uint8_t* d_Lr, uint8_t* d_Hr,
                      int* d_nbmax, hipStream_t stream) {
    int total = N * m.T;
    if (total > 0) {
        const dim3 grid((total + 255) / 256), block(256);
        switch (m.C) {
        case 1: hipLaunchKernelGGL(range_kernel<1>, grid, block, 0, stream, m, N, d_rads, d_Lr, d_Hr); break;
        case 2: hipLaunchKernelGGL(range_kernel<2>, grid, block, 0, stream, m, N, d_rads, d_Lr, d_Hr); break;
        case 3: hipLaunchKernelGGL(range_kernel<3>, grid, block, 0, stream, m, N, d_rads, d_Lr, d_Hr); break;
        case 4: hipLaunchKernelGGL(range_kernel<4>, grid, block, 0, stream, m, N, d_rads, d_Lr, d_Hr); break;
        default: break;
        }
        hipLaunchKernelGGL(bounds_kernel, dim3((N + 255) / 256), dim3(256), 0, stream, m, N, d_Lr, d_Hr,
                           d_nbmax);
    }
}

hipError_t launch_classify(const ClassifyArgs& a, int precision, int group_lanes, hipStream_t stream) {
    if (a.N <= 0) return hipSuccess;
    if (precision == WP_PRECISION_FP32) return launch_r<float>(a, group_lanes, stream);
    return launch_r<double>(a, group_lanes, stream);
}

void launch_work(const DevModel& m, int N, int S, const uint8_t* d_Lr, const uint8_t* d_Hr,
                 const int* d_nbmax, const uint8_t* d_D, double* d_out, hipStream_t stream) {
    if (N > 0)
        hipLaunchKernelGGL(work_kernel, dim3((N + 127) / 128), dim3(128), 0, stream, m, N, S, d_Lr, d_Hr,
                           d_nbmax, d_D, d_out);
}

size_t classify_lds_bytes(int precision, int T, int C, int tri, int NB, int cap, int ngroups, int G) {
    int nthreads = ngroups * G;
    if (nthreads < 64) nthreads = 64;
    nthreads = (nthreads + 63) / 64 * 64;
    if (precision == WP_PRECISION_FP32) return make_layout<float>(T, C, tri, NB, cap, ngroups, G, nthreads).total;
    return make_layout<double>(T, C, tri, NB, cap, ngroups, G, nthreads).total;
}

}  // namespace wp
