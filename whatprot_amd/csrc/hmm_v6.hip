// Classify kernel (FP32 rescaled) for reads whose count box holds at most 16
// states and T <= 12: L lanes per candidate (L = 1, 2 or 4 for boxes of up
// to 4, 8 and 16 states), one scale per candidate.
//
// Lane p of a candidate's group keeps rows e = p + L k (k < R = T / L) in
// registers; with L = 1 one lane holds the whole candidate and the Edman
// success of row e - 1 reaches row e in registers. Rows above t are zero at
// step t, so a slot is skipped while L k > t (uniform) and no lane computes
// rows that cannot hold mass yet.
//
// Scaling. All rows of a candidate share one power-of-two exponent, taken
// from the group's largest entry after every step: each entry costs a max
// and an exact ldexp, with no exponent exchange between rows. Rows of one
// candidate meet the same densities at the same count state; what separates
// them is Edman failure (pf per step and row) and bleaching, a few dozen
// binary orders at the lengths scored, far inside FP32's range. Densities
// are relative to each box's largest, and a read whose box densities span
// more than 2^kMaxDensitySpan goes to the wide pass as in the quad-rows kernel.
// The blocked and detached shares go through g_t into an FP64 side sum per
// lane, with its own exponent. The read's tables are built as for the
// quad-rows kernel (read_tables.h).
#include <float.h>
#include <math.h>

#include <algorithm>

#include "read_tables.h"

namespace wp {
namespace {

using namespace cc;
using namespace rt;

// Lane p of each group of L receives lane (p - 1) mod L's value.
template <int L>
__device__ inline float grp_rot(float v) {
    if constexpr (L == 2) {
        return __int_as_float(
            __builtin_amdgcn_update_dpp(0, __float_as_int(v), 1 | (0 << 2) | (3 << 4) | (2 << 6), 0xf, 0xf, false));
    } else if constexpr (L == 4) {
        return quad_rot(v);
    } else {
        return v;
    }
}

// Maximum over a group of L lanes (every lane gets it).
template <int L>
__device__ inline float grp_max(float v) {
    if constexpr (L >= 2)
        v = fmaxf(v, __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 1 | (0 << 2) | (3 << 4) | (2 << 6),
                                                                 0xf, 0xf, false)));
    if constexpr (L >= 4)
        v = fmaxf(v, __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 2 | (3 << 2) | (0 << 4) | (1 << 6),
                                                                 0xf, 0xf, false)));
    return v;
}

// W block of the lane-group kernel, in floats: boxes of up to 16 states keep
// 32 KB (the classes of 10^5 simulated reads of that bin need 39 KB at most
// and 11 KB at the median; the reserve for K_t = 16 at every timestep was
// 50 KB, which held the kernel to 2 workgroups per CU).
// Staged launches (boxes of up to 32 states, or variant 1 of the 16-state
// bin) build each timestep's matrices just before its step: two timesteps'
// blocks at class NB.
__host__ __device__ constexpr size_t v6_wcap(int NB, int C, bool staged) {
    return staged ? 2 * (1 + C) * ((size_t)NB * NB + 4) : NB >= 16 ? 8192 : 0;
}

// Per-timestep constants of one step t >= 1 (lane-group kernel).
struct V6Step {
    const float* Wt;   // K x K composite Bleach, then the C Edman-shifted forms, K * K + 4 apart
    const float* Et;   // densities of box(t) relative to 2^Xt
    const float* kct;  // [NB][C] counts of box(t) as floats
    const double* gt;  // g_t over box(t - 1), relative to 2^Gx
    int nbp, t, Xt, Gx, p;
    float cA, cApf;    // (1 - pd)(1 - pcb), times pf for the row's own (failed Edman) share
};

// One step t >= 1 (CyclicBlock, Detach(t-1), Bleach_c, Edman, Emit(t)) of a
// candidate's rows on a box of class K, then the candidate's rescale. Per
// row: BA = A Wt, BA1 = A W1[c_e]; the Edman success
//   succ = (1-pf)[(1 - n_c/K_e) BA + ((n_c+1)/K_e) BA1]
// moves to row e + 1, and A'_e = (1-pd)(1-pcb)(pf BA_e + succ_{e-1}) E_t; the
// row's CyclicBlock and Detach shares (A . g_t) go to the side sum. ce: the
// candidate's exponent (A holds values x 2^-ce), kNoExp once all rows are 0.
template <int C, int NB, int L, int R, int K>
__device__ __forceinline__ void v6_step(const V6Step& st, const int (&cix)[R], const float (&aK)[R],
                                        const float (&cst)[R], const float (&Kf)[R], float (&A)[R][NB], int& ce,
                                        double& sm, int& se) {
    constexpr int MS = K * K + 4;
    const float* Wt = st.Wt;
    const int p = st.p;
    float succ_prev[K];  // L = 1: row e - 1's success; else this lane's slot k - 1 success
#pragma unroll
    for (int q = 0; q < K; q++) succ_prev[q] = 0.0f;
    double sv = 0.0;  // A . g_t over this lane's rows (before the step)
    float m = 0.0f;
#pragma unroll
    for (int k = 0; k < R; k++) {
        if (st.t < L * k) continue;  // every row of slot k starts after t (uniform)
        if constexpr (K > 1) asm volatile("" ::: "memory");
        const float* W1 = Wt + (1 + cix[k]) * MS;
        float BA[K], BA1[K];
#pragma unroll
        for (int q2 = 0; q2 < K; q2++) {
            BA[q2] = 0.0f;
            BA1[q2] = 0.0f;
        }
#pragma unroll
        for (int q = 0; q < K; q++) {
            if (q >= st.nbp) continue;  // uniform: the previous box's size
            const float av = A[k][q];
            sv = fma((double)av, st.gt[q], sv);
#pragma unroll
            for (int q2 = 0; q2 < K; q2++) {
                BA[q2] = fmaf(Wt[q * K + q2], av, BA[q2]);
                BA1[q2] = fmaf(W1[q * K + q2], av, BA1[q2]);
            }
        }
#pragma unroll
        for (int q2 = 0; q2 < K; q2++) {
            const float kc = st.kct[q2 * C + cix[k]];
            const float cba = fmaf(fmaxf(Kf[k] - kc, 0.0f), aK[k], cst[k]);
            const float cba1 = fmaf(kc, aK[k], aK[k]);
            const float succ = fmaf(cba1, BA1[q2], cba * BA[q2]);
            float from;
            if constexpr (L == 1) {
                from = succ_prev[q2];
            } else {
                from = grp_rot<L>(p == L - 1 ? succ_prev[q2] : succ);
            }
            succ_prev[q2] = succ;
            // Emit(t) (peptide-emission.cc:212-243)
            const float na = fmaf(st.cA, from, st.cApf * BA[q2]) * st.Et[q2];
            m = fmaxf(m, na);
            A[k][q2] = na;
        }
    }
    if (ce != kNoExp) side_add(sm, se, sv, ce + st.Gx);
    const int km = exp_of(grp_max<L>(m));
    if (km == kNoExp) {
        ce = kNoExp;
        return;
    }
#pragma unroll
    for (int k = 0; k < R; k++) {
        if (st.t < L * k) continue;
#pragma unroll
        for (int q2 = 0; q2 < K; q2++) A[k][q2] = ldexpf(A[k][q2], -km);
    }
    ce += st.Xt + km;
}

// Timestep t's (1 + C) matrices of class K, built by the whole workgroup.
template <int C, int NB>
__device__ inline void fill_step_w(const ReadTables& s, int tri, int t, int K, float* dst) {
    fill_w_block<C, NB>(s, tri, t, K, dst);
}

// One read: the shared prologue, then its candidates, L lanes each, in
// rounds of one candidate per group. kStaged: the composite matrices of each
// timestep are built just before its step into one of two LDS buffers (the
// groups step together, a barrier per timestep), instead of all timesteps'
// in the prologue; for boxes whose classes' matrices outgrow the LDS.
template <int C, int NB, int L, bool kStaged>
__device__ __forceinline__ void v6_read(const ClassifyArgs& a, const int r, unsigned char* smem) {
    constexpr int NGR = kReadThreads / L;
    constexpr int R = 12 / L;  // T <= 12
    constexpr int kWStep = (1 + C) * (NB * NB + 4);  // one timestep's block at class NB
    if (!in_bin(a, r)) return;
    const int T = a.m.T;
    const int tri = a.tri;
    const int tid = threadIdx.x;
    const int group = tid / L;
    const int p = tid % L;
    const ReadLayout lay = read_layout(T, C, tri, NB, NGR, v6_wcap(NB, C, kStaged));
    const ReadTables s = read_views(smem, lay);
    if (!read_prologue<C, NB, true, !kStaged>(a, r, s, (int)((lay.binom - lay.W) / sizeof(float)))) return;
    const int* sL = s.sL;
    const int* nbt = s.nbt;
    const int* kcl = s.kcl;
    const int* woff = s.woff;
    const int* coord = s.coord;
    const float* kcf = s.kcf;
    const float* etab = s.etab;
    const int* xt = s.xt;
    float* Wall = s.Wall;
    const double* gtab = s.gtab;
    const int* gx = s.gx;
    const double* b0d = s.b0d;
    int ncand;
    const int* cidx;
    read_candidates(a, r, &ncand, &cidx);
    const bool read_empty = a.nbmax[r] == 0;

    // The group's running argmax and normaliser live in its gres slot (four
    // doubles: best ln P, best candidate, log-sum-exp pair), read and written
    // once per candidate, so nothing of the accumulator is held in registers
    // across a candidate's timestep loop.
    double* gacc = s.gres + 4 * group;
    if (p == 0) {
        gacc[0] = -INFINITY;
        gacc[1] = 0.0;
        gacc[2] = -INFINITY;
        gacc[3] = 0.0;
    }
    const bool collect = a.tie_mode == 1;
    if (collect && p == 0) tie_init(&s.ties[group]);
    const float opf = (float)(1.0 - a.m.pf);
    const float pib = (float)a.m.p_ib;
    const double* dud = s.bin;

    for (int base = 0; base < ncand; base += kReadThreads * 4) {
        const int nlive = compact_chunk<C>(a, r, ncand, cidx, base, read_empty, sL, s.list, s.scan);
        for (int r0 = 0; r0 < nlive; r0 += NGR) {
            // staged: the previous round's last step may still read the
            // buffer this round's first step fills
            if (kStaged && r0 > 0) __syncthreads();
            const int q0 = r0 + group;
            const bool live = q0 < nlive;
            const int kk = live ? s.list[q0] : 0;
            const int i = live ? (cidx ? cidx[kk] : kk) : 0;
            int D[C];
#pragma unroll
            for (int c = 0; c < C; c++) D[c] = a.dsD[(size_t)i * C + c];
            // Edman coefficients of each row: keep = aK max(K - kc, 0) + cst,
            // success = aK (kc + 1), with aK = (1 - pf) / K for a dye of
            // channel c_e at position e, and aK = 0, cst = 1 - pf for none.
            int cix[R];
            float aK[R], cst[R], Kf[R];
#pragma unroll
            for (int k = 0; k < R; k++) {
                const int e = p + L * k;
                const int ce = e < T ? (int)a.dsSeq[(size_t)i * T + e] : -1;
                const int K = ce >= 0 ? (int)a.dsTrack[((size_t)i * T + e) * C + ce] : 1;
                cix[k] = ce < 0 ? 0 : ce;
                aK[k] = ce >= 0 ? opf / (float)K : 0.0f;
                cst[k] = ce >= 0 ? 0.0f : opf;
                Kf[k] = (float)K;
            }
            float A[R][NB];
#pragma unroll
            for (int k = 0; k < R; k++)
#pragma unroll
                for (int q = 0; q < NB; q++) A[k][q] = 0.0f;
            double sm = 0.0;  // side sum of this lane's rows, sm 2^se
            int se = kNoExp;
            // t = 0: InitialBlock, Dud_c, Emit(0) on row 0 (lane 0, slot 0);
            // the blocked share p_ib Dud(D -> n) E_0[n] is worth beta_0[n]
            float m = 0.0f;
            if (p == 0) {
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
                side_add(sm, se, b0, xt[0] + gx[0]);
            }
            int ce;
            {
                const int km = exp_of(grp_max<L>(m));
                ce = km == kNoExp ? kNoExp : xt[0] + km;
                if (km != kNoExp) {
#pragma unroll
                    for (int q = 0; q < NB; q++) A[0][q] = ldexpf(A[0][q], -km);
                }
            }

            for (int t = 1; t < T; t++) {
                const int Kt = kcl[t];  // uniform over the workgroup
                float* Wt_t;
                if constexpr (kStaged) {
                    Wt_t = Wall + (t & 1) * kWStep;
                    fill_step_w<C, NB>(s, tri, t, Kt, Wt_t);
                    __syncthreads();
                } else {
                    Wt_t = Wall + woff[t];
                }
                if (!live || ce == kNoExp) continue;  // ce: uniform over the group (every row zero)
                V6Step st;
                st.Wt = Wt_t;
                st.nbp = nbt[t - 1];
                st.t = t;
                const float pd = (float)a.m.p_detach[t - 1];
                st.cA = (1.0f - pd) * (1.0f - (float)a.m.p_cb);
                st.cApf = st.cA * (float)a.m.pf;
                st.Et = etab + (size_t)t * NB;
                st.Xt = xt[t];
                st.gt = gtab + (size_t)t * NB;
                st.Gx = gx[t];
                st.kct = kcf + (size_t)t * NB * C;
                st.p = p;
                if (Kt <= 1) v6_step<C, NB, L, R, 1>(st, cix, aK, cst, Kf, A, ce, sm, se);
                else if (Kt == 2) v6_step<C, NB, L, R, 2>(st, cix, aK, cst, Kf, A, ce, sm, se);
                else if (Kt == 4 || NB == 4) v6_step<C, NB, L, R, (NB < 4 ? NB : 4)>(st, cix, aK, cst, Kf, A, ce, sm, se);
                else if (Kt == 8 || NB == 8) v6_step<C, NB, L, R, (NB < 8 ? NB : 8)>(st, cix, aK, cst, Kf, A, ce, sm, se);
                else if (Kt == 16 || NB == 16) v6_step<C, NB, L, R, (NB < 16 ? NB : 16)>(st, cix, aK, cst, Kf, A, ce, sm, se);
                else v6_step<C, NB, L, R, NB>(st, cix, aK, cst, Kf, A, ce, sm, se);
            }

            // P = sum A over the last box (x 2^ce) + the side sums
            const int nlast = nbt[T - 1];
            float rs = 0.0f;
            if (ce != kNoExp) {
#pragma unroll
                for (int k = 0; k < R; k++)
#pragma unroll
                    for (int q = 0; q < NB; q++)
                        if (q < nlast) rs += A[k][q];
            }
            int kl = kNoExp;
            if (rs > 0.0f) kl = ce;
            if (sm > 0.0) kl = max(kl, se);
            double part = 0.0;
            if (rs > 0.0f) part += ldexp((double)rs, ce - kl);
            if (sm > 0.0) part += ldexp(sm, se - kl);
            double sum;
            int K;
            if constexpr (L == 1) {
                K = kl;
                sum = part;
            } else {
                K = group_max_int<L>(kl);
                part = (K == kNoExp || kl == kNoExp) ? 0.0 : ldexp(part, kl - K);
                sum = group_sum<double, L>(part);
            }
            const double logp = (sum > 0.0 && K != kNoExp) ? log(sum) + (double)K * 0.69314718055994530942 : -INFINITY;
            if (!live || p != 0) continue;
            // the candidate again from LDS, not held across the timestep loop
            asm volatile("" ::: "memory");
            const int kc = s.list[q0];
            const int ic = cidx ? cidx[kc] : kc;
            if (a.out_pair_logp && !cidx) a.out_pair_logp[(size_t)r * a.S + ic] = logp;
            CandAcc<true> acc;
            acc.best_v = gacc[0];
            acc.best_k = (int)gacc[1];
            acc.tot_a = gacc[2];
            acc.tot_b = gacc[3];
            acc.add(logp, 0.0, kc, a.dsCount[ic]);
            gacc[0] = acc.best_v;
            gacc[1] = (double)acc.best_k;
            gacc[2] = acc.tot_a;
            gacc[3] = acc.tot_b;
            if (collect) tie_note(&s.ties[group], acc.best_v, logp, kc, a.tie_abs, a.tie_rel);
        }
        __syncthreads();
    }
    CandAcc<true> acc;
    if (p == 0) {
        acc.best_v = gacc[0];
        acc.best_k = (int)gacc[1];
        acc.tot_a = gacc[2];
        acc.tot_b = gacc[3];
    }
    __syncthreads();  // the merge below reuses gres
    merge_lanes_and_store<L>(a, r, ncand, cidx, NGR, group, p, acc, s.gres, s.ties, s.list, (int*)(s.gres + 24));
}

// Waves per SIMD: 3 where a lane's rows fit in 168 registers (L * 4 >= NB),
// with some per-candidate values spilled around the timestep loop; at 2
// waves nothing spills but every bin runs slower (10^5 reads x 1,000
// dye-seqs, bins alone: 4.0 / 6.2 / 7.8 / 8.4 ms at 3 waves against
// 4.6 / 7.1 / 9.2 / 8.8 ms at 2).
template <int NB, int L>
constexpr int v6_min_waves() {
    return L * 4 >= NB ? 3 : 2;
}

// Lanes per candidate by box capacity, as measured alone on 10^5 reads x
// 1,000 dye-seqs (C = 3, T = 12): boxes <= 4 at 1 lane (3.6 ms against 4.0 at
// 2), <= 8 at 2 (5.6 against 7.3 at 1), <= 16 at 4 (7.2 against 9.3 at 2);
// boxes <= 32 at 4 (three rows of 32 entries per lane).
constexpr int v6_lanes(int NB) { return NB <= 4 ? 1 : NB <= 8 ? 2 : 4; }

template <int C, int NB, int L, bool kStaged>
__global__ void __launch_bounds__(kReadThreads, (v6_min_waves<NB, L>())) classify_v6_kernel(const ClassifyArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    if ((int)blockIdx.x >= *a.list_count) return;
    v6_read<C, NB, L, kStaged>(a, a.list[blockIdx.x], smem);
}

template <int C, int NB, bool kStaged>
hipError_t launch_v6_s(const ClassifyArgs& a, hipStream_t stream) {
    constexpr int L = v6_lanes(NB);
    const ReadLayout l = read_layout(a.m.T, C, a.tri, NB, kReadThreads / L, v6_wcap(NB, C, kStaged));
    if (l.total > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute((const void*)classify_v6_kernel<C, NB, L, kStaged>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)l.total);
        if (e != hipSuccess) return e;
    }
    // one workgroup per read of the bin's list (workgroups past its length exit)
    hipLaunchKernelGGL((classify_v6_kernel<C, NB, L, kStaged>), dim3(a.N), dim3(kReadThreads), l.total, stream, a);
    return hipGetLastError();
}

template <int C, int NB>
hipError_t launch_v6_nb(const ClassifyArgs& a, hipStream_t stream) {
    if constexpr (NB >= 32) {
        return launch_v6_s<C, NB, true>(a, stream);
    } else if constexpr (NB >= 16) {
        return a.variant ? launch_v6_s<C, NB, true>(a, stream) : launch_v6_s<C, NB, false>(a, stream);
    } else {
        return launch_v6_s<C, NB, false>(a, stream);
    }
}

template <int C>
hipError_t launch_v6_c(const ClassifyArgs& a, hipStream_t stream) {
    switch (a.NB) {
    case 4: return launch_v6_nb<C, 4>(a, stream);
    case 8: return launch_v6_nb<C, 8>(a, stream);
    case 16: return launch_v6_nb<C, 16>(a, stream);
    case 32: return launch_v6_nb<C, 32>(a, stream);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace

bool v6_supports(int precision, int T, int NB) {
    return precision == WP_PRECISION_FP32 && T <= 12 && (NB == 4 || NB == 8 || NB == 16 || NB == 32);
}

size_t v6_lds_bytes(int T, int C, int tri, int NB) {
    const int L = v6_lanes(NB);
    return std::max(read_layout(T, C, tri, NB, kReadThreads / L, v6_wcap(NB, C, NB >= 32)).total,
                    read_layout(T, C, tri, NB, kReadThreads / L, v6_wcap(NB, C, NB >= 16)).total);
}

hipError_t launch_classify_v6(const ClassifyArgs& a, hipStream_t stream) {
    if (a.N <= 0) return hipSuccess;
    switch (a.m.C) {
    case 1: return launch_v6_c<1>(a, stream);
    case 2: return launch_v6_c<2>(a, stream);
    case 3: return launch_v6_c<3>(a, stream);
    case 4: return launch_v6_c<4>(a, stream);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace wp
