// Classify kernel (FP32 rescaled) for reads whose count box holds at most 16
// states: a quad of lanes per candidate, several Edman rows per lane, one
// power-of-two exponent per row.
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
// Every row carries its own exponent, so the per-entry work of a step is the
// Edman coefficients, the two updates, the emission product and a rescale;
// row exponents meet only where the Edman success crosses rows and in the
// final sum of a pair. The read's tables (read_tables.h) are shared with the
// lane-group kernel, which serves T <= 12; this kernel serves longer runs.
#include <float.h>
#include <math.h>

#include <algorithm>

#include "read_tables.h"

namespace wp {
namespace {

using namespace cc;
using namespace rt;

constexpr int kV5Lanes = 4;  // lanes per candidate: one DPP quad

// Per-timestep constants of one step t >= 1, shared by every candidate.
struct V5Step {
    const float* Wt;   // K x K composite Bleach, then the C Edman-shifted forms, K * K + 4 apart
    const float* Et;   // densities of box(t) relative to 2^Xt
    const float* kct;  // [NB][C] counts of box(t) as floats
    const double* gt;  // g_t over box(t - 1), relative to 2^Gx
    int nbp, t, Xt, Gx, p;
    float cAA, pf;
};

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
// candidates, a quad of lanes each. kClass: each timestep at its own box
// class; otherwise every timestep at the bin's capacity NB. Block-uniform
// returns only.
template <int C, int NB, int R, bool kClass>
__device__ __forceinline__ void v5_read(const ClassifyArgs& a, const int r, unsigned char* smem) {
    constexpr int G = kV5Lanes;
    constexpr int NGR = kReadThreads / G;
    constexpr int MSNB = NB * NB + 4;
    if (!in_bin(a, r)) return;
    const int T = a.m.T;
    const int tri = a.tri;
    const int tid = threadIdx.x;
    const int group = tid / G;
    const int p = tid % G;
    const ReadLayout lay = read_layout(T, C, tri, NB, NGR);
    const ReadTables s = read_views(smem, lay);
    if (!read_prologue<C, NB, kClass>(a, r, s, (int)((lay.binom - lay.W) / sizeof(float)))) return;
    const int* sL = s.sL;
    const int* nbt = s.nbt;
    const int* kcl = s.kcl;
    const int* woff = s.woff;
    const int* coord = s.coord;
    const float* kcf = s.kcf;
    const float* etab = s.etab;
    const int* xt = s.xt;
    const float* Wall = s.Wall;
    const double* gtab = s.gtab;
    const int* gx = s.gx;
    const double* b0d = s.b0d;
    int ncand;
    const int* cidx;
    read_candidates(a, r, &ncand, &cidx);
    const bool read_empty = a.nbmax[r] == 0;

    CandAcc<true> acc;
    const bool collect = a.tie_mode == 1;
    if (collect && p == 0) tie_init(&s.ties[group]);
    const float pf = (float)a.m.pf;
    const float opf = (float)(1.0 - a.m.pf);
    const double* dud = s.bin;

    for (int base = 0; base < ncand; base += kReadThreads * 4) {
        const int nlive = compact_chunk<C>(a, r, ncand, cidx, base, read_empty, sL, s.list, s.scan);
        for (int q0 = group; q0 < nlive; q0 += NGR) {
            const int kk = s.list[q0];
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
            const double sum = group_sum<double, G>(part);
            const double logp = sum > 0.0 ? log(sum) + (double)K * 0.69314718055994530942 : -INFINITY;
            if (a.out_pair_logp && !cidx && p == 0) a.out_pair_logp[(size_t)r * a.S + i] = logp;
            acc.add(logp, exp(logp), kk, a.dsCount[i]);
            if (collect && p == 0) tie_note(&s.ties[group], acc.best_v, logp, kk, a.tie_abs, a.tie_rel);
        }
        __syncthreads();
    }
    merge_and_store<true>(a, r, ncand, cidx, NGR, group, p, acc, s.gres, s.ties, s.list);
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
__global__ void __launch_bounds__(kReadThreads, (v5_min_waves<NB, R>())) classify_v5_kernel(const ClassifyArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    if ((int)blockIdx.x >= *a.list_count) return;
    v5_read<C, NB, R, kClass>(a, a.list[blockIdx.x], smem);
}

template <int C, int NB, int R, bool kClass>
hipError_t launch_v5_v(const ClassifyArgs& a, hipStream_t stream) {
    const ReadLayout l = read_layout(a.m.T, C, a.tri, NB, kReadThreads / kV5Lanes);
    if (l.total > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute((const void*)classify_v5_kernel<C, NB, R, kClass>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)l.total);
        if (e != hipSuccess) return e;
    }
    // one workgroup per read of the bin's list (workgroups past its length exit)
    hipLaunchKernelGGL((classify_v5_kernel<C, NB, R, kClass>), dim3(a.N), dim3(kReadThreads), l.total, stream, a);
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

size_t v5_lds_bytes(int T, int C, int tri, int NB) { return read_layout(T, C, tri, NB, kReadThreads / kV5Lanes).total; }

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
