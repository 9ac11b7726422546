// Baum-Welch E-step of HMMFitter::fit on the GPU: for every read, the
// forward and backward passes of the PeptideHMM of the fitted dye-seq and the
// sufficient statistics of every step's improve_fit (generic-hmm.h:60-109),
// summed over reads.
//
// Fitting runs with uncut ranges (run-fit.cc:64-65), so every read's box at
// timestep t is the full [0, t+1) x prod_c [0, D_c + 1) and every range
// includes zero; the emission steps keep the detached state, and the Edman
// and detach statistics skip the all-zero entry of each Edman row
// (edman-transition.cc:295-312, detach-transition.cc:128-144).
//
// Mapping: one workgroup per read at a time (a grid-stride loop over reads);
// the state tensors A (normal) and B (blocked N-terminus) of the current
// timestep live in LDS, row-major [e][q] with q the row-major index of the
// per-channel counts. A forward pass stores the state after every emission
// (alpha_t) in a per-workgroup global history; the backward pass walks the
// timesteps from the end, rebuilds each timestep's intermediate forward
// states from alpha_{t-1}, and applies each step's backward operator while
// accumulating its statistics (block-transition.cc:93-116,
// binomial-transition.cc:185-256, detach-transition.cc:105-166,
// edman-transition.cc:279-325).
//
// Scaling: after every emission the state is renormalised by an exact power
// of two, tracked as an integer log2 (K_t forward, J backward), so FP32 keeps
// its range over any number of timesteps and FP64 reproduces the unscaled
// reference values. A statistic of the steps of timestep t is
// sum(alpha * x * beta) / P = sum(alpha_s * x * beta_s) * 2^(K_{t-1} + J - log2 P),
// computed in the state type and accumulated in FP64.
#include <hip/hip_runtime.h>
#include <math.h>

#include <algorithm>

#include "fit_kernels.h"

namespace wp {
namespace {

constexpr int kNoExp = -100000;
constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;

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

// Block-wide sums of two values; every thread gets the result.
__device__ inline void block_sum2(double& a, double& b, double* red) {
    a = wave_sum(a);
    b = wave_sum(b);
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) {
        red[2 * w] = a;
        red[2 * w + 1] = b;
    }
    __syncthreads();
    a = 0.0;
    b = 0.0;
    for (int i = 0; i < kWaves; i++) {
        a += red[2 * i];
        b += red[2 * i + 1];
    }
}

__device__ inline int block_max(int v, int* redi) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off, 64));
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) redi[w] = v;
    __syncthreads();
    int m = redi[0];
    for (int i = 1; i < kWaves; i++) m = max(m, redi[i]);
    return m;
}

template <typename Real>
struct Smem {
    Real* Em;         // [T][NB] emission mantissa (FP64: the density itself)
    int* Ee;          // [T][NB] emission exponent (FP32 only)
    Real* Y;          // [C+1][2][T*NB] rebuilt forward states of one timestep
    Real *bA, *bB, *bA2;
    Real* bin;        // [2][C][tri] dud, bleach
    unsigned char* crd;  // [NB][C] counts of q
    unsigned short* fbase;  // per channel: row-relative base q of each fiber
    signed char* seq;       // [T]
    unsigned char* track;   // [T][C]
    int* Kt;          // [T] cumulative forward exponent after emission t
    double* det;      // [2][T] detach x, n of this workgroup's reads
    double* red;      // reduction scratch
    int* redi;
};

template <typename Real, int C>
__host__ __device__ __attribute__((always_inline)) inline size_t smem_layout(int T, int NB, int tri, int nfib,
                                                                             Smem<Real>* s, unsigned char* base) {
    size_t o = 0;
    auto take = [&](size_t bytes) {
        size_t at = o;
        o = (o + bytes + 15) & ~(size_t)15;
        return base ? base + at : nullptr;
    };
    const size_t S = (size_t)T * NB;
    Smem<Real> l;
    l.Em = (Real*)take(sizeof(Real) * S);
    l.Ee = (int*)take(sizeof(Real) == 4 ? sizeof(int) * S : 0);
    l.Y = (Real*)take(sizeof(Real) * (C + 1) * 2 * S);
    l.bA = (Real*)take(sizeof(Real) * S);
    l.bB = (Real*)take(sizeof(Real) * S);
    l.bA2 = (Real*)take(sizeof(Real) * S);
    l.bin = (Real*)take(sizeof(Real) * 2 * C * tri);
    l.crd = (unsigned char*)take((size_t)NB * C);
    l.fbase = (unsigned short*)take(sizeof(unsigned short) * nfib);
    l.seq = (signed char*)take(T);
    l.track = (unsigned char*)take((size_t)T * C);
    l.Kt = (int*)take(sizeof(int) * T);
    l.det = (double*)take(sizeof(double) * 2 * T);
    l.red = (double*)take(sizeof(double) * 2 * kWaves);
    l.redi = (int*)take(sizeof(int) * kWaves);
    if (s) *s = l;
    return o;
}

// arr[c] for a run-time c without indexing a register array (which would
// move it to scratch memory).
template <int C>
__device__ inline int pick(const int (&arr)[C], int c) {
    int v = arr[0];
#pragma unroll
    for (int k = 1; k < C; k++)
        if (k == c) v = arr[k];
    return v;
}

// Fibers of channel c over rows [0, R) of both tensors: fiber f -> tensor,
// row, base index. Fibers of one row are numbered by fbase.
struct Fiber {
    int tensor, off;
};
__device__ inline Fiber fiber_at(int f, int R, int G, int NB, const unsigned short* fb) {
    Fiber r;
    r.tensor = f >= R * G;
    int ff = f - r.tensor * R * G;
    int e = ff / G;
    r.off = e * NB + fb[ff - e * G];
    return r;
}

// One workgroup's share of an E-step under model m: the reads r = blockIdx.x
// + k gridDim.x, weighted by weight (null: 1), with the binomial tables binom;
// the workgroup's totals go to out[0..L].
template <typename Real, int C>
__device__ void fit_estep_one(const FitArgs& a, const DevModel& m, const double* binom, const int* weight,
                              double* out, unsigned char* smem) {
    const int T = m.T, NB = a.NB, tid = threadIdx.x;
    const size_t S = (size_t)T * NB;
    Smem<Real> L;
    smem_layout<Real, C>(T, NB, a.tri, a.nfib, &L, smem);
    Real* bA = L.bA;  // backward state, swapped with bA2 around the Edman step
    Real* bB = L.bB;
    Real* bA2 = L.bA2;
    const bool split = sizeof(Real) == 4;

    // per-fitter constants into LDS
    for (int i = tid; i < 2 * C * a.tri; i += kThreads) L.bin[i] = (Real)binom[i];
    for (int i = tid; i < NB * C; i += kThreads) L.crd[i] = a.crd[i];
    for (int i = tid; i < a.nfib; i += kThreads) L.fbase[i] = a.fbase[i];
    for (int i = tid; i < T; i += kThreads) L.seq[i] = a.seq[i];
    for (int i = tid; i < T * C; i += kThreads) L.track[i] = a.track[i];
    for (int i = tid; i < 2 * T; i += kThreads) L.det[i] = 0.0;
    __syncthreads();

    const Real pf = (Real)m.pf, pib = (Real)m.p_ib, pcb = (Real)m.p_cb;
    int D[C], st[C], G[C], fo[C];
    {
        int acc = 1, off = 0;
        for (int c = C - 1; c >= 0; c--) {
            D[c] = a.D[c];
            st[c] = acc;
            acc *= D[c] + 1;
        }
        for (int c = 0; c < C; c++) {
            G[c] = NB / (D[c] + 1);
            fo[c] = off;
            off += G[c];
        }
    }
    // FP64 accumulators of this thread: edman, initial block, cyclic block,
    // then per channel bleach and dud (numerator, denominator each)
    double acc_e[2] = {0, 0}, acc_ib[2] = {0, 0}, acc_cb[2] = {0, 0};
    double acc_bl[C][2], acc_du[C][2];
    for (int c = 0; c < C; c++) acc_bl[c][0] = acc_bl[c][1] = acc_du[c][0] = acc_du[c][1] = 0.0;
    double sum_logp = 0.0;

    Real* hist_base = (Real*)a.hist + (size_t)blockIdx.x * a.hist_stride;
    auto hist_at = [&](int t, int tensor) { return hist_base + ((size_t)t * 2 + tensor) * S; };
    Real* Y = L.Y;
    auto YA = [&](int i) { return Y + (size_t)i * 2 * S; };
    auto YB = [&](int i) { return Y + (size_t)i * 2 * S + S; };

    // Binomial step on channel c over rows [0, R): out = step(in) (forward,
    // binomial-transition.cc:91-136).
    auto binom_fwd = [&](int kind, int c, int R, const Real* inA, const Real* inB, Real* outA,
                         Real* outB) __attribute__((always_inline)) {
        const Real* P = L.bin + (size_t)(kind * C + c) * a.tri;
        const int len = pick<C>(D, c) + 1, s = pick<C>(st, c), g = pick<C>(G, c);
        const unsigned short* fbc = L.fbase + pick<C>(fo, c);
        for (int f = tid; f < 2 * R * g; f += kThreads) {
            Fiber fb = fiber_at(f, R, g, NB, fbc);
            const Real* in = (fb.tensor ? inB : inA) + fb.off;
            Real* out = (fb.tensor ? outB : outA) + fb.off;
            for (int to = 0; to < len; to++) {
                Real v = 0;
                for (int from = to; from < len; from++) v += P[from * (from + 1) / 2 + to] * in[from * s];
                out[to * s] = v;
            }
        }
    };
    // Backward of the same step in place on b (binomial-transition.cc:138-183)
    // with its statistics against forward states fA, fB.
    auto binom_bwd = [&](int kind, int c, int R, const Real* fA, const Real* fB, double F, double& acc_num,
                         double& acc_den) __attribute__((always_inline)) {
        const Real* P = L.bin + (size_t)(kind * C + c) * a.tri;
        const int len = pick<C>(D, c) + 1, s = pick<C>(st, c), g = pick<C>(G, c);
        const unsigned short* fbc = L.fbase + pick<C>(fo, c);
        double num = 0.0, den = 0.0;
        for (int f = tid; f < 2 * R * g; f += kThreads) {
            Fiber fb = fiber_at(f, R, g, NB, fbc);
            const Real* fv = (fb.tensor ? fB : fA) + fb.off;
            Real* b = (fb.tensor ? bB : bA) + fb.off;
            for (int from = len - 1; from >= 0; from--) {
                Real v = 0;
                for (int to = 0; to <= from; to++) v += P[from * (from + 1) / 2 + to] * b[to * s];
                if (from >= 1) {
                    Real af = fv[from * s];
                    den += (double)(af * v) * from;
                    for (int to = 0; to < from; to++)
                        num += (double)(af * P[from * (from + 1) / 2 + to] * b[to * s]) * (from - to);
                }
                b[from * s] = v;
            }
        }
        acc_num += num * F;
        acc_den += den * F;
    };
    // Emission of timestep t on rows [0, R) of (xA, xB) and the scalar x0,
    // renormalised by a power of two; returns the exponent taken out
    // (kNoExp when everything is zero).
    auto emit = [&](int t, int R, Real* xA, Real* xB, Real& x0) __attribute__((always_inline)) -> int {
        const Real* Em = L.Em + (size_t)t * NB;
        const int* Ee = split ? L.Ee + (size_t)t * NB : nullptr;
        int kx = kNoExp;
        for (int i = tid; i < R * NB; i += kThreads) {
            int q = i % NB;
            int ee = split ? Ee[q] : 0;
            Real va = xA[i] * Em[q], vb = xB[i] * Em[q];
            xA[i] = va;
            xB[i] = vb;
            kx = max(kx, max(exp_of(va), exp_of(vb)) + ee);
        }
        Real v0 = x0 * Em[0];
        int e0 = split ? Ee[0] : 0;
        kx = max(kx, exp_of(v0) + e0);
        int k = block_max(kx, L.redi);
        if (k <= kNoExp / 2) return kNoExp;
        for (int i = tid; i < R * NB; i += kThreads) {
            int q = i % NB;
            int sh = (split ? Ee[q] : 0) - k;
            xA[i] = scale2(xA[i], sh);
            xB[i] = scale2(xB[i], sh);
        }
        x0 = scale2(v0, e0 - k);
        __syncthreads();
        return k;
    };

    for (int r = blockIdx.x; r < a.N; r += gridDim.x) {
        // A read drawn w times adds w times its statistics and ln P; reads
        // not drawn are skipped (uniform over the workgroup).
        const int wr = weight ? weight[r] : 1;
        if (wr == 0) continue;
        // ---- emission table of this read (peptide-emission.cc:165-181)
        const double* rad = a.rads + (size_t)r * T * C;
        for (int i = tid; i < (int)S; i += kThreads) {
            int t = i / NB, q = i - t * NB;
            unsigned int cnt[C];
            for (int c = 0; c < C; c++) cnt[c] = L.crd[q * C + c];
            double v = 1.0;
            for (int c = 0; c < C; c++) v *= dev_pdf<C>(m, c, rad[t * C + c], cnt);
            if (split) {
                int e;
                double mnt = frexp(v, &e);
                L.Em[i] = (Real)mnt;
                L.Ee[i] = v > 0.0 ? e : kNoExp;
            } else {
                L.Em[i] = (Real)v;
            }
        }
        __syncthreads();

        // ---- forward pass, storing alpha_t after every emission
        Real* cA = YA(0);
        Real* cB = YB(0);
        Real* nA = YA(1);
        Real* nB = YB(1);
        auto swap = [&]() {
            Real* t0 = cA;
            cA = nA;
            nA = t0;
            t0 = cB;
            cB = nB;
            nB = t0;
        };
        for (int i = tid; i < NB; i += kThreads) {
            cA[i] = i == NB - 1 ? (Real)1 : (Real)0;  // peptide-state-vector.cc:38-40
            cB[i] = 0;
        }
        __syncthreads();
        Real dl = 0;  // detached state, same scale as A and B
        int K = 0;
        bool zero = false;
        for (int t = 0; t < T && !zero; t++) {
            if (t == 0) {
                for (int i = tid; i < NB; i += kThreads) {  // initial block
                    Real x = cA[i];
                    nA[i] = (1 - pib) * x;
                    nB[i] = cB[i] + pib * x;
                }
                __syncthreads();
                swap();
                #pragma unroll
                for (int c = 0; c < C; c++) {  // duds
                    binom_fwd(0, c, 1, cA, cB, nA, nB);
                    __syncthreads();
                    swap();
                }
            } else {
                const Real pd = (Real)m.p_detach[t - 1];
                double sum = 0.0, unused = 0.0;
                for (int i = tid; i < t * NB; i += kThreads) {  // cyclic block + detach
                    Real x = cA[i];
                    Real ca = (1 - pcb) * x, cb = cB[i] + pcb * x;
                    nA[i] = ca * (1 - pd);
                    nB[i] = cb * (1 - pd);
                    sum += (double)ca + (double)cb;
                }
                block_sum2(sum, unused, L.red);
                dl = (Real)((double)dl + (double)pd * sum);
                swap();
                #pragma unroll
                for (int c = 0; c < C; c++) {  // bleach
                    binom_fwd(1, c, t, cA, cB, nA, nB);
                    __syncthreads();
                    swap();
                }
                for (int i = tid; i < (t + 1) * NB; i += kThreads) {  // Edman, gathered
                    int e = i / NB, q = i - e * NB;
                    Real v = 0;
                    if (e >= 1) {
                        const Real* src = cA + (size_t)(e - 1) * NB;
                        int c = L.seq[e - 1];
                        if (c < 0) {
                            v += (1 - pf) * src[q];
                        } else {
                            int k = L.crd[q * C + c], Kc = L.track[(e - 1) * C + c];
                            if (k < Kc) v += ((1 - pf) * (1 - (Real)k / (Real)Kc)) * src[q];
                            if (k + 1 <= pick<C>(D, c)) v += ((1 - pf) * ((Real)(k + 1) / (Real)Kc)) * src[q + pick<C>(st, c)];
                        }
                    }
                    if (e < t) v += pf * cA[i];
                    nA[i] = v;
                    nB[i] = e < t ? cB[i] : (Real)0;
                }
                __syncthreads();
                swap();
            }
            int k = emit(t, t + 1, cA, cB, dl);
            if (k == kNoExp) {
                zero = true;
                break;
            }
            K += k;
            if (tid == 0) L.Kt[t] = K;
            Real* hA = hist_at(t, 0);
            Real* hB = hist_at(t, 1);
            for (int i = tid; i < (t + 1) * NB; i += kThreads) {
                hA[i] = cA[i];
                hB[i] = cB[i];
            }
        }
        if (zero) {  // P == 0: no statistics (generic-hmm.h:77-87)
            if (tid == 0) sum_logp += -INFINITY;  // any positive multiplicity
            __syncthreads();
            continue;
        }
        double tot = 0.0, unused = 0.0;
        for (int i = tid; i < (int)S; i += kThreads) tot += (double)cA[i] + (double)cB[i];
        block_sum2(tot, unused, L.red);
        tot += (double)dl;
        const double lp2 = (double)K + log2(tot);
        if (tid == 0) sum_logp += (double)wr * (lp2 * 0.69314718055994530942);

        // ---- backward pass with the statistics
        for (int i = tid; i < (int)S; i += kThreads) {
            bA[i] = 1;  // peptide-state-vector.cc:42-50
            bB[i] = 1;
        }
        Real bd = 1;
        int J = 0;
        __syncthreads();
        for (int t = T - 1; t >= 0; t--) {
            int k = emit(t, t + 1, bA, bB, bd);
            J += k;
            if (t >= 1) {
                const double F = (double)wr * exp2((double)(L.Kt[t - 1] + J) - lp2);
                const Real pd = (Real)m.p_detach[t - 1];
                const Real* hA = hist_at(t - 1, 0);
                const Real* hB = hist_at(t - 1, 1);
                for (int i = tid; i < t * NB; i += kThreads) {  // Y0 = detach(cyclic block(alpha_{t-1}))
                    Real x = hA[i];
                    YA(0)[i] = ((1 - pcb) * x) * (1 - pd);
                    YB(0)[i] = (hB[i] + pcb * x) * (1 - pd);
                }
                __syncthreads();
                #pragma unroll
                for (int c = 0; c < C; c++) {
                    binom_fwd(1, c, t, YA(c), YB(c), YA(c + 1), YB(c + 1));
                    __syncthreads();
                }
                {  // Edman backward with its statistics (A only)
                    const Real* fA = YA(C);
                    double num = 0.0, den = 0.0;
                    for (int i = tid; i < t * NB; i += kThreads) {
                        int e = i / NB, q = i - e * NB;
                        Real b = bA[i];
                        Real v = pf * b;
                        const Real* nxt = bA + (size_t)(e + 1) * NB;
                        int c = L.seq[e];
                        if (c < 0) {
                            v += (1 - pf) * nxt[q];
                        } else {
                            int kk = L.crd[q * C + c], Kc = L.track[e * C + c];
                            if (kk >= 1) v += ((1 - pf) * ((Real)kk / (Real)Kc)) * nxt[q - pick<C>(st, c)];
                            if (kk < Kc) v += ((1 - pf) * (1 - (Real)kk / (Real)Kc)) * nxt[q];
                        }
                        if (q != 0) {
                            Real f = fA[i];
                            num += (double)(f * pf * b);
                            den += (double)(f * v);
                        }
                        bA2[i] = v;
                    }
                    acc_e[0] += num * F;
                    acc_e[1] += den * F;
                    __syncthreads();
                    Real* tmp = bA;
                    bA = bA2;
                    bA2 = tmp;
                }
                #pragma unroll
                for (int c = C - 1; c >= 0; c--) {  // bleach backward
                    binom_bwd(1, c, t, YA(c), YB(c), F, acc_bl[c][0], acc_bl[c][1]);
                    __syncthreads();
                }
                {  // detach and cyclic block backward with their statistics
                    double fs = 0.0, fbs = 0.0, cn = 0.0, cd = 0.0;
                    const Real d = bd;
                    for (int i = tid; i < t * NB; i += kThreads) {
                        Real x0 = hA[i];
                        Real a1A = (1 - pcb) * x0, a1B = hB[i] + pcb * x0;
                        Real lA = (1 - pd) * bA[i] + pd * d;
                        Real lB = (1 - pd) * bB[i] + pd * d;
                        if (i % NB != 0) {
                            fs += (double)a1A + (double)a1B;
                            fbs += (double)(a1A * lA) + (double)(a1B * lB);
                        }
                        Real cA2 = pcb * lB + (1 - pcb) * lA;
                        cn += (double)(x0 * pcb * lB);
                        cd += (double)(x0 * cA2);
                        bA[i] = cA2;
                        bB[i] = lB;
                    }
                    acc_cb[0] += cn * F;
                    acc_cb[1] += cd * F;
                    block_sum2(fs, fbs, L.red);
                    if (tid == 0) {
                        L.det[t - 1] += fs * (double)pd * (double)d * F;
                        L.det[T + t - 1] += fbs * F;
                    }
                    __syncthreads();
                }
            } else {
                const double F = (double)wr * exp2((double)J - lp2);
                for (int i = tid; i < NB; i += kThreads) {  // Y0 = initial block(start)
                    YA(0)[i] = i == NB - 1 ? (1 - pib) : (Real)0;
                    YB(0)[i] = i == NB - 1 ? pib : (Real)0;
                }
                __syncthreads();
                #pragma unroll
                for (int c = 0; c < C; c++) {
                    binom_fwd(0, c, 1, YA(c), YB(c), YA(c + 1), YB(c + 1));
                    __syncthreads();
                }
                #pragma unroll
                for (int c = C - 1; c >= 0; c--) {  // dud backward
                    binom_bwd(0, c, 1, YA(c), YB(c), F, acc_du[c][0], acc_du[c][1]);
                    __syncthreads();
                }
                if (tid == 0) {  // initial block: the start state is the only non-zero input
                    Real vB = bB[NB - 1], vA = bA[NB - 1];
                    acc_ib[0] += (double)(pib * vB) * F;
                    acc_ib[1] += (double)(pib * vB + (1 - pib) * vA) * F;
                }
                __syncthreads();
            }
        }
    }

    // ---- this workgroup's totals
    auto put = [&](int idx, double v0, double v1) {
        block_sum2(v0, v1, L.red);
        if (tid == 0) {
            out[idx] = v0;
            out[idx + 1] = v1;
        }
    };
    put(WP_FIT_P_EDMAN_FAILURE, acc_e[0], acc_e[1]);
    put(WP_FIT_P_INITIAL_BLOCK, acc_ib[0], acc_ib[1]);
    put(WP_FIT_P_CYCLIC_BLOCK, acc_cb[0], acc_cb[1]);
    for (int c = 0; c < C; c++) {
        put(WP_FIT_CHANNEL(T, c), acc_bl[c][0], acc_bl[c][1]);
        put(WP_FIT_CHANNEL(T, c) + 2, acc_du[c][0], acc_du[c][1]);
    }
    __syncthreads();
    for (int i = tid; i < T - 1; i += kThreads) {
        out[WP_FIT_DETACH_X + i] = L.det[i];
        out[WP_FIT_DETACH_X + (T - 1) + i] = L.det[T + i];
    }
    if (tid == 0) out[a.L] = sum_logp;
}

// The E-steps of a.nslots models in one launch. Each workgroup runs its share
// of every slot in turn, exactly as a launch of that slot alone would (same
// reads, same order, own totals), so every slot's result is the single
// launch's bit for bit. Slot k uses models[k] (a.m when there is one slot and
// no array), binom + k * 2C * tri and weight row slot_round[k] (row 0 when
// there is no map).
template <typename Real, int C>
__global__ void __launch_bounds__(kThreads) fit_estep_kernel(const FitArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    for (int k = 0; k < a.nslots; k++) {
        const int row = a.slot_round ? a.slot_round[k] : 0;
        fit_estep_one<Real, C>(a, a.models ? a.models[k] : a.m, a.binom + (size_t)k * 2 * C * a.tri,
                               a.weight ? a.weight + (size_t)row * a.N : nullptr,
                               a.partial + ((size_t)k * gridDim.x + blockIdx.x) * (a.L + 1), smem);
        __syncthreads();
    }
}

// Sums the per-workgroup totals of slot blockIdx.y in workgroup order
// (deterministic).
__global__ void fit_reduce_kernel(const double* partial, int nblocks, int L, double* out) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > L) return;
    const size_t k = blockIdx.y;
    double s = 0.0;
    for (int b = 0; b < nblocks; b++) s += partial[(k * nblocks + b) * (L + 1) + i];
    out[k * (L + 1) + i] = s;
}

template <typename Real, int C>
hipError_t launch_c(const FitArgs& a, int grid, hipStream_t stream) {
    size_t lds = smem_layout<Real, C>(a.m.T, a.NB, a.tri, a.nfib, nullptr, nullptr);
    hipError_t e = hipFuncSetAttribute((const void*)fit_estep_kernel<Real, C>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((fit_estep_kernel<Real, C>), dim3(grid), dim3(kThreads), lds, stream, a);
    return hipGetLastError();
}

}  // namespace

size_t fit_lds_bytes(int precision, int T, int C, int NB, int tri, int nfib) {
    auto f = [&](auto real) -> size_t {
        using Real = decltype(real);
        switch (C) {
        case 1: return smem_layout<Real, 1>(T, NB, tri, nfib, nullptr, nullptr);
        case 2: return smem_layout<Real, 2>(T, NB, tri, nfib, nullptr, nullptr);
        case 3: return smem_layout<Real, 3>(T, NB, tri, nfib, nullptr, nullptr);
        case 4: return smem_layout<Real, 4>(T, NB, tri, nfib, nullptr, nullptr);
        default: return (size_t)-1;
        }
    };
    return precision == WP_PRECISION_FP32 ? f(0.0f) : f(0.0);
}

int fit_threads() { return kThreads; }

hipError_t launch_fit_estep(const FitArgs& a, int precision, int grid, double* d_out, hipStream_t stream) {
    hipError_t e = hipErrorInvalidValue;
    const int C = a.m.C;
    if (a.rows) {
        e = launch_fit_rows(a, precision, grid, stream);
    } else if (precision == WP_PRECISION_FP32) {
        if (C == 1) e = launch_c<float, 1>(a, grid, stream);
        else if (C == 2) e = launch_c<float, 2>(a, grid, stream);
        else if (C == 3) e = launch_c<float, 3>(a, grid, stream);
        else if (C == 4) e = launch_c<float, 4>(a, grid, stream);
    } else {
        if (C == 1) e = launch_c<double, 1>(a, grid, stream);
        else if (C == 2) e = launch_c<double, 2>(a, grid, stream);
        else if (C == 3) e = launch_c<double, 3>(a, grid, stream);
        else if (C == 4) e = launch_c<double, 4>(a, grid, stream);
    }
    if (e != hipSuccess) return e;
    int n = a.L + 1;
    hipLaunchKernelGGL(fit_reduce_kernel, dim3((n + 255) / 256, std::max(a.nslots, 1)), dim3(256), 0, stream,
                       a.partial, grid, a.L, d_out);
    return hipGetLastError();
}

}  // namespace wp
