// Row-lane Baum-Welch E-step: the same forward/backward passes and
// statistics as fit_estep_kernel (generic-hmm.h:60-109 with the steps of
// peptide-hmm.cc:40-61), mapped so that one wavefront scores one read with no
// workgroup barrier.
//
// Mapping: lane = 32 * tensor + e, tensor 0 = A (normal), 1 = B (blocked
// N-terminus), e = Edman row (amino acids removed). A lane holds the count
// states of its row in registers, x[q] with q the row-major index of the
// per-channel counts, so the count shape (D_c per channel) is a compile-time
// parameter. Every step is then lane-local except three exchanges:
//   - cyclic/initial block: row e of B gains p * row e of A (lane e -> lane
//     32 + e);
//   - Edman forward: row e of A gains from row e - 1 (lane e - 1 -> lane e);
//   - Edman backward: row e of A reads row e + 1.
// The binomial (dud, bleach) steps act on the fibers of one channel inside a
// row, all in registers. Rows past the current timestep are zero in the
// forward pass and excluded from every sum and maximum of the backward pass
// (their values are never read again).
//
// History: alpha_t (the state after emission t) goes to a per-wavefront
// global buffer, [t][q][lane]; the backward pass reloads alpha_{t-1} and
// rebuilds the intermediate states of timestep t from it. Those rebuilt
// states (after detach, then after each bleach channel) are recomputed from
// the first one as each backward step needs them, so only three row arrays
// (beta, the first rebuilt state and one working state) stay in registers.
//
// Scaling is fit_estep_kernel's: a power of two after every emission, the
// forward exponents K_t and the backward exponent J, statistics of timestep
// t scaled by 2^(K_{t-1} + J - log2 P). A lane sums the terms of one step of
// one timestep (at most NB of them) in the state type and adds that sum to
// its FP64 totals, which run over timesteps and reads.
#include <hip/hip_runtime.h>
#include <math.h>

#include "fit_kernels.h"

namespace wp {
namespace {

constexpr int kNoExp = -100000;
constexpr int kRowThreads = 256;
constexpr int kRowWaves = kRowThreads / 64;
// FP64 statistics of each lane live in LDS, [stat][lane] per wavefront:
// edman (2), initial block (2), cyclic block (2), per channel bleach (2) and
// dud (2), then detach x_t and n_t (lane t holds timestep t).
constexpr int kAccStride = 64;
constexpr int kAccEdman = 0, kAccInitial = 2, kAccCyclic = 4, kAccChannel = 6;
__host__ __device__ constexpr int acc_stats(int C) { return 8 + 4 * C; }
// dye counts 0..2 of one channel: the Edman coefficients of a row
constexpr int kCoefK = 3;

// Count shape: C channels with D0..D2 dyes (C <= 3).
template <int C_, int D0, int D1, int D2>
struct Shape {
    static constexpr int C = C_;
    static constexpr int d(int c) { return c == 0 ? D0 : c == 1 ? D1 : D2; }
    static constexpr int NB = (D0 + 1) * (C_ > 1 ? D1 + 1 : 1) * (C_ > 2 ? D2 + 1 : 1);
    static constexpr int st(int c) {
        int s = 1;
        for (int k = C_ - 1; k > c; k--) s *= d(k) + 1;
        return s;
    }
    static constexpr int coord(int q, int c) { return (q / st(c)) % (d(c) + 1); }
    static constexpr int dmax() {
        int m = 0;
        for (int c = 0; c < C_; c++) m = d(c) > m ? d(c) : m;
        return m;
    }
};

__device__ inline int exp_of(float x) {
    int e;
    frexpf(x, &e);
    return x > 0.0f ? e : kNoExp;
}
__device__ inline int exp_of(double x) {
    int e;
    frexp(x, &e);
    return x > 0.0 ? e : kNoExp;
}
__device__ inline float scale2(float x, int k) { return ldexpf(x, k); }
__device__ inline double scale2(double x, int k) { return ldexp(x, k); }

__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}
__device__ inline int wave_max(int v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off, 64));
    return v;
}

// LDS of one workgroup: shared tables, then per wavefront the emission table
// of its current read.
template <typename Real>
struct RowSmem {
    Real* bin;            // [2][C][tri] dud then bleach
    signed char* seq;     // [T]
    unsigned char* track; // [T][C]
    Real* Em;             // [waves][T][NB] emission (mantissa in FP32)
    int* Ee;              // [waves][T][NB] exponent (FP32 only)
    double* red;          // [waves][64] final reductions
    double* acc;          // [waves][acc_stats(C)][64] per-lane statistics
    Real* ecoef;          // [T][2][kCoefK] Edman coefficients of each row: stay[k], then up[k]
};

template <typename Real>
__host__ __device__ inline size_t row_smem(int T, int C, int NB, int tri, RowSmem<Real>* s, unsigned char* base) {
    size_t o = 0;
    auto take = [&](size_t bytes) {
        size_t at = o;
        o = (o + bytes + 15) & ~(size_t)15;
        return base ? base + at : nullptr;
    };
    RowSmem<Real> l;
    l.bin = (Real*)take(sizeof(Real) * 2 * C * tri);
    l.seq = (signed char*)take(T);
    l.track = (unsigned char*)take((size_t)T * C);
    l.Em = (Real*)take(sizeof(Real) * kRowWaves * T * NB);
    l.Ee = (int*)take(sizeof(Real) == 4 ? sizeof(int) * kRowWaves * T * NB : 0);
    l.red = (double*)take(sizeof(double) * kRowWaves * 64);
    l.acc = (double*)take(sizeof(double) * kRowWaves * acc_stats(C) * kAccStride);
    l.ecoef = (Real*)take(sizeof(Real) * T * 2 * kCoefK);
    if (s) *s = l;
    return o;
}

// out[to] = sum_{from >= to} P[from][to] in[from] along the fibers of channel
// c (binomial-transition.cc:91-136), in place (ascending to reads only
// entries not yet overwritten).
template <class S, int c, typename Real>
__device__ __attribute__((always_inline)) inline void bin_fwd(Real (&x)[S::NB], const Real* P) {
    constexpr int len = S::d(c) + 1, s = S::st(c);
    Real p[len * (len + 1) / 2];
#pragma unroll
    for (int i = 0; i < len * (len + 1) / 2; i++) p[i] = P[i];
#pragma unroll
    for (int q0 = 0; q0 < S::NB; q0++) {
        if (S::coord(q0, c) != 0) continue;
#pragma unroll
        for (int to = 0; to < len; to++) {
            Real v = 0;
#pragma unroll
            for (int from = to; from < len; from++) v += p[from * (from + 1) / 2 + to] * x[q0 + from * s];
            x[q0 + to * s] = v;
        }
    }
}

// Backward of the same step in place on b (binomial-transition.cc:138-183)
// with its statistics against the forward state f (:185-256); returns the
// lane's numerator and denominator sums (in the state type: one timestep's
// terms of one lane, added to the FP64 totals by the caller).
template <class S, int c, typename Real>
__device__ __attribute__((always_inline)) inline void bin_bwd(Real (&b)[S::NB], const Real (&f)[S::NB],
                                                              const Real* P, Real& num, Real& den) {
    constexpr int len = S::d(c) + 1, s = S::st(c);
    Real p[len * (len + 1) / 2];
#pragma unroll
    for (int i = 0; i < len * (len + 1) / 2; i++) p[i] = P[i];
    Real n = 0, d = 0;
#pragma unroll
    for (int q0 = 0; q0 < S::NB; q0++) {
        if (S::coord(q0, c) != 0) continue;
#pragma unroll
        for (int from = len - 1; from >= 0; from--) {
            Real v = 0;
#pragma unroll
            for (int to = 0; to <= from; to++) v += p[from * (from + 1) / 2 + to] * b[q0 + to * s];
            if (from >= 1) {
                const Real af = f[q0 + from * s];
                d += (af * v) * (Real)from;
#pragma unroll
                for (int to = 0; to < from; to++)
                    n += (af * p[from * (from + 1) / 2 + to] * b[q0 + to * s]) * (Real)(from - to);
            }
            b[q0 + from * s] = v;
        }
    }
    num = n;
    den = d;
}

// Bleach channels 0 .. n-1 applied to x in order.
template <class S, int n, typename Real>
__device__ __attribute__((always_inline)) inline void bleach_upto(Real (&x)[S::NB], const Real* Pb, int tri) {
    if constexpr (n > 0) {
        bleach_upto<S, n - 1, Real>(x, Pb, tri);
        bin_fwd<S, n - 1, Real>(x, Pb + (size_t)(n - 1) * tri);
    }
}

// Bleach backward of channels c = hi-1 down to 0 (peptide-hmm.cc step order
// reversed), each against the rebuilt forward state before that channel's
// bleach: Y0 bleached by channels 0 .. c-1.
template <class S, int hi, typename Real>
__device__ __attribute__((always_inline)) inline void bleach_bwd_all(Real (&b)[S::NB], const Real (&Y0)[S::NB],
                                                                     Real (&Y)[S::NB], const Real* Pb, int tri,
                                                                     bool live, double F, double* acc) {
    if constexpr (hi > 0) {
        constexpr int c = hi - 1;
#pragma unroll
        for (int q = 0; q < S::NB; q++) Y[q] = Y0[q];
        bleach_upto<S, c, Real>(Y, Pb, tri);
        Real n, d;
        bin_bwd<S, c, Real>(b, Y, Pb + (size_t)c * tri, n, d);
        if (live) {
            acc[kAccStride * 4 * c] += (double)n * F;
            acc[kAccStride * (4 * c + 1)] += (double)d * F;
        }
        bleach_bwd_all<S, c, Real>(b, Y0, Y, Pb, tri, live, F, acc);
    }
}

// Dud forward over every channel, then dud backward from the last channel
// with statistics (t = 0; rebuilt states recomputed from Y0 like the bleach).
template <class S, int n, typename Real>
__device__ __attribute__((always_inline)) inline void dud_upto(Real (&x)[S::NB], const Real* Pd, int tri) {
    if constexpr (n > 0) {
        dud_upto<S, n - 1, Real>(x, Pd, tri);
        bin_fwd<S, n - 1, Real>(x, Pd + (size_t)(n - 1) * tri);
    }
}
template <class S, int hi, typename Real>
__device__ __attribute__((always_inline)) inline void dud_bwd_all(Real (&b)[S::NB], const Real (&Y0)[S::NB],
                                                                  Real (&Y)[S::NB], const Real* Pd, int tri,
                                                                  bool live, double F, double* acc) {
    if constexpr (hi > 0) {
        constexpr int c = hi - 1;
#pragma unroll
        for (int q = 0; q < S::NB; q++) Y[q] = Y0[q];
        dud_upto<S, c, Real>(Y, Pd, tri);
        Real n, d;
        bin_bwd<S, c, Real>(b, Y, Pd + (size_t)c * tri, n, d);
        if (live) {
            acc[kAccStride * 4 * c] += (double)n * F;
            acc[kAccStride * (4 * c + 1)] += (double)d * F;
        }
        dud_bwd_all<S, c, Real>(b, Y0, Y, Pd, tri, live, F, acc);
    }
}

// Emission of one timestep on the lane's row and the detached scalar x0,
// renormalised by the largest exponent over the live rows of the wavefront;
// returns that exponent (kNoExp when every live entry is zero).
template <class S, typename Real>
__device__ __attribute__((always_inline)) inline int emit_row(Real (&x)[S::NB], Real& x0, const Real* Em,
                                                              const int* Ee, bool live) {
    constexpr bool split = sizeof(Real) == 4;
    int kx = kNoExp;
#pragma unroll
    for (int q = 0; q < S::NB; q++) {
        const Real va = x[q] * Em[q];
        x[q] = va;
        kx = max(kx, exp_of(va) + (split ? Ee[q] : 0));
    }
    if (!live) kx = kNoExp;
    const Real v0 = x0 * Em[0];
    const int e0 = split ? Ee[0] : 0;
    kx = max(kx, exp_of(v0) + e0);
    const int k = wave_max(kx);
    if (k <= kNoExp / 2) return kNoExp;
#pragma unroll
    for (int q = 0; q < S::NB; q++) x[q] = scale2(x[q], (split ? Ee[q] : 0) - k);
    x0 = scale2(v0, e0 - k);
    return k;
}

// Edman coefficients of a lane whose source row has a dye of channel c with
// Kc dyes of that channel remaining: stay[k] for count k keeping its dye
// count, up[k] for count k + 1 losing one (edman-transition.cc:43-110).
template <class S, typename Real>
struct EdmanCoef {
    Real stay[S::dmax() + 1], up[S::dmax() + 1];
};

template <typename Real>
__device__ inline Real coef_stay(Real pf, int k, int Kc) {
    return k < Kc ? (1 - pf) * (1 - (Real)k / (Real)Kc) : (Real)0;
}
template <typename Real>
__device__ inline Real coef_up(Real pf, int k, int Kc) {
    return (1 - pf) * ((Real)(k + 1) / (Real)Kc);
}

// The coefficients of row e from the workgroup's table (built once per model).
template <class S, typename Real>
__device__ inline void edman_coef(EdmanCoef<S, Real>& ec, const Real* table, int e) {
#pragma unroll
    for (int k = 0; k <= S::dmax(); k++) {
        ec.stay[k] = table[(size_t)e * 2 * kCoefK + k];
        ec.up[k] = table[(size_t)e * 2 * kCoefK + kCoefK + k];
    }
}

// count of channel c (run time) at compile-time q
template <class S, int q>
__device__ inline int coord_of(int c) {
    int k = S::coord(q, 0);
    if constexpr (S::C > 1)
        if (c == 1) k = S::coord(q, 1);
    if constexpr (S::C > 2)
        if (c == 2) k = S::coord(q, 2);
    return k;
}

template <class S, typename Real, int q>
__device__ __attribute__((always_inline)) inline void edman_fwd_q(Real (&x)[S::NB], const Real (&src)[S::NB], int c,
                                                                  const EdmanCoef<S, Real>& ec, Real pf,
                                                                  bool has_src, bool keep) {
    Real v = 0;
    if (has_src) {
        if (c < 0) {
            v += (1 - pf) * src[q];
        } else {
            const int k = coord_of<S, q>(c);
            Real cs = ec.stay[0], cu = ec.up[0];
#pragma unroll
            for (int kk = 1; kk <= S::dmax(); kk++)
                if (k == kk) {
                    cs = ec.stay[kk];
                    cu = ec.up[kk];
                }
            // src[q + st_c] exists when count k + 1 <= D_c
            Real upv = 0;
            bool has_up = false;
            if constexpr (S::coord(q, 0) + 1 <= S::d(0))
                if (c == 0) {
                    upv = src[q + S::st(0)];
                    has_up = true;
                }
            if constexpr (S::C > 1 && S::coord(q, 1) + 1 <= S::d(1))
                if (c == 1) {
                    upv = src[q + S::st(1)];
                    has_up = true;
                }
            if constexpr (S::C > 2 && S::coord(q, 2) + 1 <= S::d(2))
                if (c == 2) {
                    upv = src[q + S::st(2)];
                    has_up = true;
                }
            v += cs * src[q];
            if (has_up) v += cu * upv;
        }
    }
    if (keep) v += pf * x[q];
    x[q] = v;
}

template <class S, typename Real, int q = 0>
__device__ __attribute__((always_inline)) inline void edman_fwd_all(Real (&x)[S::NB], const Real (&src)[S::NB], int c,
                                                                    const EdmanCoef<S, Real>& ec, Real pf,
                                                                    bool has_src, bool keep) {
    if constexpr (q < S::NB) {
        edman_fwd_q<S, Real, q>(x, src, c, ec, pf, has_src, keep);
        edman_fwd_all<S, Real, q + 1>(x, src, c, ec, pf, has_src, keep);
    }
}

// Edman backward of one entry with its statistics (edman-transition.cc:112-
// 160, 279-325): v = pf b[q] + (1 - pf) terms from row e + 1.
template <class S, typename Real, int q>
__device__ __attribute__((always_inline)) inline void edman_bwd_q(Real (&b)[S::NB], const Real (&nxt)[S::NB],
                                                                  const Real (&f)[S::NB],
                                                                  int c, const EdmanCoef<S, Real>& ec, Real pf,
                                                                  Real& num, Real& den) {
    Real v = pf * b[q];
    if (c < 0) {
        v += (1 - pf) * nxt[q];
    } else {
        const int k = coord_of<S, q>(c);
        // k >= 1: dye lost from count k, coefficient up[k - 1]
        Real cd = 0, downv = 0;
        bool has_down = false;
        if constexpr (S::coord(q, 0) >= 1)
            if (c == 0) {
                downv = nxt[q - S::st(0)];
                has_down = true;
            }
        if constexpr (S::C > 1 && S::coord(q, 1) >= 1)
            if (c == 1) {
                downv = nxt[q - S::st(1)];
                has_down = true;
            }
        if constexpr (S::C > 2 && S::coord(q, 2) >= 1)
            if (c == 2) {
                downv = nxt[q - S::st(2)];
                has_down = true;
            }
        Real cs = ec.stay[0];
#pragma unroll
        for (int kk = 1; kk <= S::dmax(); kk++)
            if (k == kk) {
                cs = ec.stay[kk];
                cd = ec.up[kk - 1];
            }
        if (has_down) v += cd * downv;
        v += cs * nxt[q];
    }
    if constexpr (q != 0) {
        num += f[q] * pf * b[q];
        den += f[q] * v;
    }
    b[q] = v;
}

template <class S, typename Real, int q = 0>
__device__ __attribute__((always_inline)) inline void edman_bwd_all(Real (&b)[S::NB], const Real (&nxt)[S::NB],
                                                                    const Real (&f)[S::NB],
                                                                    int c, const EdmanCoef<S, Real>& ec, Real pf,
                                                                    Real& num, Real& den) {
    if constexpr (q < S::NB) {
        edman_bwd_q<S, Real, q>(b, nxt, f, c, ec, pf, num, den);
        edman_bwd_all<S, Real, q + 1>(b, nxt, f, c, ec, pf, num, den);
    }
}

template <class S, typename Real>
__device__ __attribute__((always_inline)) inline void xchg32(Real (&o)[S::NB], const Real (&x)[S::NB]) {
#pragma unroll
    for (int q = 0; q < S::NB; q++) o[q] = __shfl_xor(x[q], 32, 64);
}

// One workgroup's share of an E-step under model m (reads r = wavefront id +
// k * waves of the grid, weighted by weight, binomial tables binom); the
// workgroup's totals go to out[0..L].
template <class S, typename Real>
__device__ void rows_estep_one(const FitArgs& a, const DevModel& m, const double* binom, const int* weight,
                               double* out, unsigned char* smem) {
    constexpr int C = S::C, NB = S::NB;
    constexpr bool split = sizeof(Real) == 4;
    const int T = m.T, tri = a.tri, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int tensor = lane >> 5, e = lane & 31;
    RowSmem<Real> L;
    row_smem<Real>(T, C, NB, tri, &L, smem);
    for (int i = tid; i < 2 * C * tri; i += kRowThreads) L.bin[i] = (Real)binom[i];
    for (int i = tid; i < T; i += kRowThreads) L.seq[i] = a.seq[i];
    for (int i = tid; i < T * C; i += kRowThreads) L.track[i] = a.track[i];
    __syncthreads();
    for (int i = tid; i < T * kCoefK; i += kRowThreads) {
        const int e = i / kCoefK, k = i - e * kCoefK, c = L.seq[e];
        const int Kc = c >= 0 ? (int)L.track[e * C + c] : 1;
        L.ecoef[(size_t)e * 2 * kCoefK + k] = coef_stay<Real>((Real)m.pf, k, Kc);
        L.ecoef[(size_t)e * 2 * kCoefK + kCoefK + k] = coef_up<Real>((Real)m.pf, k, Kc);
    }
    __syncthreads();
    const Real* Pd = L.bin;
    const Real* Pb = L.bin + (size_t)C * tri;
    Real* Em = L.Em + (size_t)wave * T * NB;
    int* Ee = split ? L.Ee + (size_t)wave * T * NB : nullptr;

    const Real pf = (Real)m.pf, pib = (Real)m.p_ib, pcb = (Real)m.p_cb;
    // FP64 accumulators of this lane: edman, initial block (lane 0), cyclic
    // block, bleach and dud per channel; detach x_t and n_t in lane t
    double* acc = L.acc + (size_t)wave * acc_stats(C) * kAccStride + lane;
    auto A = [&](int k) -> double& { return acc[k * kAccStride]; };
    for (int k = 0; k < acc_stats(C); k++) A(k) = 0.0;
    double sum_logp = 0.0;

    const int wid = blockIdx.x * kRowWaves + wave, nw = gridDim.x * kRowWaves;
    Real* hist = (Real*)a.hist + (size_t)wid * ((size_t)T * NB * 64);
    auto hat = [&](int t, int q) { return hist + ((size_t)t * NB + q) * 64 + lane; };

    for (int r = wid; r < a.N; r += nw) {
        const int wr = weight ? weight[r] : 1;
        if (wr == 0) continue;
        // ---- emission table of this read (peptide-emission.cc:165-181),
        // written by the lanes of the wavefront after its reads of the last one
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        const double* rad = a.rads + (size_t)r * T * C;
        for (int i = lane; i < T * NB; i += 64) {
            const int t = i / NB, q = i - t * NB;
            unsigned int cnt[C];
            int rest = q;
#pragma unroll
            for (int c = C - 1; c >= 0; c--) {
                cnt[c] = (unsigned int)(rest % (S::d(c) + 1));
                rest /= S::d(c) + 1;
            }
            double v = 1.0;
#pragma unroll
            for (int c = 0; c < C; c++) v *= dev_pdf<C>(m, c, rad[t * C + c], cnt);
            if (split) {
                int ex;
                const double mnt = frexp(v, &ex);
                Em[i] = (Real)mnt;
                Ee[i] = v > 0.0 ? ex : kNoExp;
            } else {
                Em[i] = (Real)v;
            }
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");

        // ---- forward pass, alpha_t to the history after every emission
        Real x[NB], o[NB];
#pragma unroll
        for (int q = 0; q < NB; q++) x[q] = (tensor == 0 && e == 0 && q == NB - 1) ? (Real)1 : (Real)0;
        Real dl = 0;
        int K = 0, kt_lane = 0;
        bool zero = false;
        for (int t = 0; t < T; t++) {
            if (t == 0) {
                xchg32<S, Real>(o, x);  // initial block
#pragma unroll
                for (int q = 0; q < NB; q++) x[q] = tensor == 0 ? (1 - pib) * x[q] : x[q] + pib * o[q];
                dud_upto<S, C, Real>(x, Pd, tri);
            } else {
                const Real pd = (Real)m.p_detach[t - 1];
                xchg32<S, Real>(o, x);  // cyclic block + detach
                double sum = 0.0;
#pragma unroll
                for (int q = 0; q < NB; q++) {
                    const Real cv = tensor == 0 ? (1 - pcb) * x[q] : x[q] + pcb * o[q];
                    x[q] = cv * (1 - pd);
                    sum += (double)cv;
                }
                sum = wave_sum(e < t ? sum : 0.0);
                dl = (Real)((double)dl + (double)pd * sum);
                bleach_upto<S, C, Real>(x, Pb, tri);
                // Edman: row e takes row e - 1 of A (gathered)
#pragma unroll
                for (int q = 0; q < NB; q++) o[q] = __shfl_up(x[q], 1, 64);
                const int c = e >= 1 && e - 1 < T ? (int)L.seq[e - 1] : -1;
                EdmanCoef<S, Real> ec;
                edman_coef<S, Real>(ec, L.ecoef, c >= 0 ? e - 1 : 0);
                if (tensor == 0) {
                    edman_fwd_all<S, Real>(x, o, c, ec, pf, e >= 1 && e <= t, e < t);
                } else {
#pragma unroll
                    for (int q = 0; q < NB; q++) x[q] = e < t ? x[q] : (Real)0;
                }
            }
            const int k = emit_row<S, Real>(x, dl, Em + (size_t)t * NB, split ? Ee + (size_t)t * NB : nullptr, e <= t);
            if (k == kNoExp) {
                zero = true;
                break;
            }
            K += k;
            if (lane == t) kt_lane = K;
            if (e <= t) {
#pragma unroll
                for (int q = 0; q < NB; q++) *hat(t, q) = x[q];
            }
        }
        if (zero) {  // P == 0: no statistics (generic-hmm.h:77-87)
            if (lane == 0) sum_logp += -INFINITY;
            __builtin_amdgcn_wave_barrier();
            continue;
        }
        double tot = 0.0;
#pragma unroll
        for (int q = 0; q < NB; q++) tot += (double)x[q];
        tot = wave_sum(tot) + (double)dl;
        const double lp2 = (double)K + log2(tot);
        if (lane == 0) sum_logp += (double)wr * (lp2 * 0.69314718055994530942);

        // ---- backward pass with the statistics
        Real b[NB], Y0[NB], Y[NB];
#pragma unroll
        for (int q = 0; q < NB; q++) b[q] = 1;  // peptide-state-vector.cc:42-50
        Real bd = 1;
        int J = 0;
        for (int t = T - 1; t >= 0; t--) {
            const int k = emit_row<S, Real>(b, bd, Em + (size_t)t * NB, split ? Ee + (size_t)t * NB : nullptr, e <= t);
            J += k;
            if (t >= 1) {
                const int kprev = __shfl(kt_lane, t - 1, 64);
                const double F = (double)wr * exp2((double)(kprev + J) - lp2);
                const Real pd = (Real)m.p_detach[t - 1];
                const bool live = e < t;
                // Y0 = detach(cyclic block(alpha_{t-1}))
#pragma unroll
                for (int q = 0; q < NB; q++) Y[q] = live ? *hat(t - 1, q) : (Real)0;
                xchg32<S, Real>(Y0, Y);
#pragma unroll
                for (int q = 0; q < NB; q++)
                    Y0[q] = tensor == 0 ? ((1 - pcb) * Y[q]) * (1 - pd) : (Y[q] + pcb * Y0[q]) * (1 - pd);
                {  // Edman backward with its statistics (A only), against Y0 bleached by every channel
#pragma unroll
                    for (int q = 0; q < NB; q++) Y[q] = Y0[q];
                    bleach_upto<S, C, Real>(Y, Pb, tri);
                    Real nxt[NB];
#pragma unroll
                    for (int q = 0; q < NB; q++) nxt[q] = __shfl_down(b[q], 1, 64);
                    const int c = e < T ? (int)L.seq[e] : -1;
                    EdmanCoef<S, Real> ec;
                    edman_coef<S, Real>(ec, L.ecoef, c >= 0 ? e : 0);
                    Real num = 0, den = 0;
                    if (tensor == 0) {
                        edman_bwd_all<S, Real>(b, nxt, Y, c, ec, pf, num, den);
                        if (live) {
                            A(kAccEdman) += (double)num * F;
                            A(kAccEdman + 1) += (double)den * F;
                        }
                    }
                }
                bleach_bwd_all<S, C, Real>(b, Y0, Y, Pb, tri, live, F, acc + kAccChannel * kAccStride);
                {  // detach and cyclic block backward with their statistics
                    const Real d = bd;
#pragma unroll
                    for (int q = 0; q < NB; q++) Y[q] = live ? *hat(t - 1, q) : (Real)0;  // own tensor of alpha_{t-1}
                    xchg32<S, Real>(Y0, Y);  // the other tensor's alpha_{t-1}
                    Real ob[NB];
                    xchg32<S, Real>(ob, b);  // the other tensor's beta
                    // one formula for both tensors (no divergent branch): A keeps
                    // (1 - pcb) of its alpha and takes pcb of B's beta; B gains
                    // pcb of A's alpha and keeps its beta
                    const Real ca = tensor == 0 ? (1 - pcb) : (Real)1, co = tensor == 0 ? (Real)0 : pcb;
                    const Real u1 = tensor == 0 ? pcb : (Real)0, u2 = tensor == 0 ? (1 - pcb) : (Real)1;
                    Real fs = 0, fbs = 0, cn = 0, cdd = 0;
#pragma unroll
                    for (int q = 0; q < NB; q++) {
                        const Real x0 = Y[q];
                        const Real a1 = ca * x0 + co * Y0[q];
                        const Real l = (1 - pd) * b[q] + pd * d;
                        const Real lo = (1 - pd) * ob[q] + pd * d;
                        if (q != 0) {
                            fs += a1;
                            fbs += a1 * l;
                        }
                        const Real nb = u1 * lo + u2 * l;
                        cn += x0 * pcb * lo;
                        cdd += x0 * nb;
                        b[q] = nb;
                    }
                    if (live && tensor == 0) {
                        A(kAccCyclic) += (double)cn * F;
                        A(kAccCyclic + 1) += (double)cdd * F;
                    }
                    const double wfs = wave_sum(live ? (double)fs : 0.0);
                    const double wfbs = wave_sum(live ? (double)fbs : 0.0);
                    if (lane == t - 1) {
                        A(acc_stats(C) - 2) += wfs * (double)pd * (double)d * F;
                        A(acc_stats(C) - 1) += wfbs * F;
                    }
                }
            } else {
                const double F = (double)wr * exp2((double)J - lp2);
                const bool live = e == 0;
                // Y0 = initial block(start state)
#pragma unroll
                for (int q = 0; q < NB; q++)
                    Y0[q] = (live && q == NB - 1) ? (tensor == 0 ? (1 - pib) : pib) : (Real)0;
                dud_bwd_all<S, C, Real>(b, Y0, Y, Pd, tri, live, F, acc + (kAccChannel + 2) * kAccStride);
                // initial block: the start state is the only non-zero input
                const Real vA = __shfl(b[NB - 1], 0, 64), vB = __shfl(b[NB - 1], 32, 64);
                if (lane == 0) {
                    A(kAccInitial) += (double)(pib * vB) * F;
                    A(kAccInitial + 1) += (double)(pib * vB + (1 - pib) * vA) * F;
                }
            }
        }
    }

    // ---- this workgroup's totals (lanes, then wavefronts, summed in order)
    double* red = L.red + (size_t)wave * 64;
    __syncthreads();
    for (int k = 0; k < acc_stats(C) - 2; k++) {
        const double v = wave_sum(A(k));
        if (lane == 0) red[k] = v;
    }
    {
        const double v = wave_sum(lane == 0 ? sum_logp : 0.0);
        if (lane == 0) red[acc_stats(C) - 2] = v;
    }
    __syncthreads();
    if (tid < acc_stats(C) - 1) {
        double s = 0.0;
        for (int w = 0; w < kRowWaves; w++) s += L.red[(size_t)w * 64 + tid];
        int idx;
        if (tid < kAccChannel) idx = tid;  // edman, initial block, cyclic block: same order in the output
        else if (tid < acc_stats(C) - 2) idx = WP_FIT_CHANNEL(T, (tid - kAccChannel) / 4) + (tid - kAccChannel) % 4;
        else idx = a.L;  // sum of ln P
        out[idx] = s;
    }
    if (tid < T - 1) {
        double sx = 0.0, sn = 0.0;
        for (int w = 0; w < kRowWaves; w++) {
            const double* aw = L.acc + (size_t)w * acc_stats(C) * kAccStride + tid;
            sx += aw[(acc_stats(C) - 2) * kAccStride];
            sn += aw[(acc_stats(C) - 1) * kAccStride];
        }
        out[WP_FIT_DETACH_X + tid] = sx;
        out[WP_FIT_DETACH_X + (T - 1) + tid] = sn;
    }
    __syncthreads();
}

// The E-steps of a.nslots models in one launch, each slot exactly as a
// launch of that slot alone (fit_estep_kernel's slot contract).
// FP32: two wavefronts per SIMD (256 VGPRs, a few spilled); FP64: one,
// with the state arrays in 512 registers instead of scratch
template <typename Real>
constexpr int rows_waves_per_simd() { return sizeof(Real) == 4 ? 2 : 1; }

template <class S, typename Real>
__global__ void __launch_bounds__(kRowThreads, rows_waves_per_simd<Real>()) fit_rows_kernel(const FitArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    for (int k = 0; k < a.nslots; k++) {
        const int row = a.slot_round ? a.slot_round[k] : 0;
        rows_estep_one<S, Real>(a, a.models ? a.models[k] : a.m, a.binom + (size_t)k * 2 * S::C * a.tri,
                                a.weight ? a.weight + (size_t)row * a.N : nullptr,
                                a.partial + ((size_t)k * gridDim.x + blockIdx.x) * (a.L + 1), smem);
        __syncthreads();
    }
}

template <class S, typename Real>
hipError_t launch_s(const FitArgs& a, int grid, hipStream_t stream) {
    static_assert(S::dmax() < kCoefK, "Edman coefficient table holds counts 0..kCoefK-1");
    const size_t lds = row_smem<Real>(a.m.T, S::C, S::NB, a.tri, nullptr, nullptr);
    if (lds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute((const void*)fit_rows_kernel<S, Real>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL((fit_rows_kernel<S, Real>), dim3(grid), dim3(kRowThreads), lds, stream, a);
    return hipGetLastError();
}

// Count shapes with a row-lane kernel: every shape of C <= 3 channels with
// 1 or 2 dyes in each (each is its own code object).
template <typename Real, int C, int D0, int D1 = 1, int D2 = 1>
hipError_t launch_d(const FitArgs& a, int grid, hipStream_t stream) {
    if constexpr (C == 1) {
        return launch_s<Shape<1, D0, 0, 0>, Real>(a, grid, stream);
    } else if constexpr (C == 2) {
        return a.D[1] == 1 ? launch_s<Shape<2, D0, 1, 0>, Real>(a, grid, stream)
                           : launch_s<Shape<2, D0, 2, 0>, Real>(a, grid, stream);
    } else {
        if (a.D[1] == 1)
            return a.D[2] == 1 ? launch_s<Shape<3, D0, 1, 1>, Real>(a, grid, stream)
                               : launch_s<Shape<3, D0, 1, 2>, Real>(a, grid, stream);
        return a.D[2] == 1 ? launch_s<Shape<3, D0, 2, 1>, Real>(a, grid, stream)
                           : launch_s<Shape<3, D0, 2, 2>, Real>(a, grid, stream);
    }
}

template <typename Real, int C>
hipError_t launch_c(const FitArgs& a, int grid, hipStream_t stream) {
    return a.D[0] == 1 ? launch_d<Real, C, 1>(a, grid, stream) : launch_d<Real, C, 2>(a, grid, stream);
}

template <typename Real>
hipError_t launch_r(const FitArgs& a, int grid, hipStream_t stream) {
    switch (a.m.C) {
    case 1: return launch_c<Real, 1>(a, grid, stream);
    case 2: return launch_c<Real, 2>(a, grid, stream);
    case 3: return launch_c<Real, 3>(a, grid, stream);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace

bool fit_rows_supports(int T, int C, const int* D) {
    if (T < 1 || T > 32 || C < 1 || C > 3) return false;
    for (int c = 0; c < C; c++)
        if (D[c] < 1 || D[c] > 2) return false;
    return true;
}

int fit_rows_threads() { return kRowThreads; }

size_t fit_rows_lds_bytes(int precision, int T, int C, int NB, int tri) {
    return precision == WP_PRECISION_FP32 ? row_smem<float>(T, C, NB, tri, nullptr, nullptr)
                                          : row_smem<double>(T, C, NB, tri, nullptr, nullptr);
}

size_t fit_rows_hist_per_group(int T, int NB) { return (size_t)kRowWaves * T * NB * 64; }

int fit_rows_groups_per_cu(int precision) {
    return precision == WP_PRECISION_FP32 ? rows_waves_per_simd<float>() : rows_waves_per_simd<double>();
}

hipError_t launch_fit_rows(const FitArgs& a, int precision, int grid, hipStream_t stream) {
    return precision == WP_PRECISION_FP32 ? launch_r<float>(a, grid, stream) : launch_r<double>(a, grid, stream);
}

}  // namespace wp
