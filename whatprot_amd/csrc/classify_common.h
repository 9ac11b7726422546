// Device helpers shared by the classify kernels: wave-scope fences, group
// reductions, exponent handling, the order-preserving compaction of a read's
// non-empty candidates, and the per-read argmax / normaliser
// (classifiers/hmm-classifier.h:45-78).
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>

#include "hmm_kernels.h"

namespace wp {
namespace cc {

__device__ inline void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <typename Real, int G>
__device__ inline Real group_sum(Real v) {
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, G);
    return v;
}

template <int G>
__device__ inline int group_max_int(int v) {
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off, G));
    return v;
}

// Exponent e with x = m 2^e, m in [0.5, 1); kNoExp for zero.
constexpr int kNoExp = -100000;
__device__ inline int exp_of(float x) {
    int e;
    frexpf(x, &e);
    return x > 0.0f ? e : kNoExp;
}

__host__ __device__ inline size_t align16(size_t x) { return (x + 15) & ~(size_t)15; }

// q / d for 0 <= q < 2^16 and 1 <= d <= 64 with a per-divisor multiplier.
struct Divider {
    unsigned int m;
    __device__ Divider() : m(0) {}
    __device__ explicit Divider(int d) : m(((1u << 22) + (unsigned int)d - 1u) / (unsigned int)d) {}
    __device__ int div(int q) const { return (int)(((unsigned int)q * m) >> 22); }
};

// Running argmax and normaliser of one group over the candidates it scored.
// FP64 (kRescale false): v = P, total = sum P * count, strict > with the
// earlier candidate winning ties (hmm-classifier.h:59-65). FP32: v = ln P and
// the total is an online log-sum-exp (tot_a = running max, tot_b = scaled sum).
template <bool kRescale>
struct CandAcc {
    double best_v;
    int best_k;
    double tot_a, tot_b;
    __device__ CandAcc() : best_v(kRescale ? -INFINITY : 0.0), best_k(0), tot_a(kRescale ? -INFINITY : 0.0), tot_b(0.0) {}
    __device__ void add(double logp, double p, int kk, int count) {
        if (kRescale) {
            if (logp > best_v || (logp == best_v && kk < best_k)) {
                best_v = logp;
                best_k = kk;
            }
            if (count > 0 && logp > -INFINITY) {
                double x = logp + log((double)count);
                if (x > tot_a) {
                    tot_b = tot_b * exp(tot_a - x) + 1.0;
                    tot_a = x;
                } else {
                    tot_b += exp(x - tot_a);
                }
            }
        } else {
            tot_a += p * count;
            if (p > best_v || (p == best_v && kk < best_k)) {
                best_v = p;
                best_k = kk;
            }
        }
    }
};

// Near-best candidates of the FP32 path. FP32 ln P carries an error of at most
// eps(x) = a + b |x| against FP64 (measured, with margin), so the FP64 best of
// a read lies within W = 2 eps of the FP32 best. Every pair group keeps the
// candidates it scored within W of its own running best (the group's best only
// rises, and the window's floor best - W rises with it, so what a group drops
// is out of the final window too); at the end of the read the groups' lists
// are merged against the read's best. A read with two or more candidates in
// the window gets them, in candidate order, for an FP64 pass that applies the
// reference's strict '>' with the earlier candidate winning
// (hmm-classifier.h:59-65); a read whose window holds more than kTieMax, or
// whose group list overflowed inside it, is rescored over all its candidates.
constexpr int kTieSlots = 4;   // entries per pair group (kTieMax per read: hmm_kernels.h)
// A read whose FP32 best lies below this ln P is rescored over every
// candidate: there the FP64 recursion runs into the subnormal range (ln DBL_MIN
// = -708, and intermediate states can sit e^33 below the final P when later
// densities exceed 1), so the FP64 decision is no longer an ordering of the
// FP32 values (an FP64 P of 0 everywhere picks the first candidate).
constexpr double kTieFloorAll = -650.0;

struct TieSlot {
    double v[kTieSlots];
    int kk[kTieSlots];
    int n;
    int pad;
    double drop;  // largest value evicted from a full list (-inf if none)
};

__device__ inline double tie_floor(double best, double wa, double wr) { return best - (wa + wr * fabs(best)); }

__device__ inline void tie_init(TieSlot* ts) {
    ts->n = 0;
    ts->drop = -INFINITY;
}

// One lane per group, after the group's accumulator took (logp, kk); best is
// the group's best so far, logp included.
__device__ inline void tie_note(TieSlot* ts, double best, double logp, int kk, double wa, double wr) {
    if (!(logp > -INFINITY)) return;
    const double lim = tie_floor(best, wa, wr);
    if (logp < lim) return;
    int n = 0;
    const int n0 = ts->n;
    for (int j = 0; j < n0; j++) {
        const double v = ts->v[j];
        if (v >= lim) {
            ts->v[n] = v;
            ts->kk[n] = ts->kk[j];
            n++;
        }
    }
    if (n < kTieSlots) {
        ts->v[n] = logp;
        ts->kk[n] = kk;
        n++;
    } else {
        int jm = 0;
        for (int j = 1; j < kTieSlots; j++)
            if (ts->v[j] < ts->v[jm]) jm = j;
        if (logp > ts->v[jm]) {
            ts->drop = fmax(ts->drop, ts->v[jm]);
            ts->v[jm] = logp;
            ts->kk[jm] = kk;
        } else {
            ts->drop = fmax(ts->drop, logp);
        }
    }
    ts->n = n;
}

// The candidate list of read r: its own list (or every dye-seq), or in the
// FP64 rescoring pass (tie_mode 2) the near-best list the FP32 pass left
// (tie_n > 0; -1 keeps the whole list).
__device__ inline void read_candidates(const ClassifyArgs& a, int r, int* ncand, const int** cidx) {
    *ncand = a.cand_off ? a.cand_off[r + 1] - a.cand_off[r] : a.S;
    *cidx = a.cand_off ? a.cand_idx + a.cand_off[r] : nullptr;
    if (a.tie_mode == 2) {
        const int tn = a.tie_n[r];
        if (tn > 0) {
            *ncand = tn;
            *cidx = a.tie_idx + (size_t)r * kTieMax;
        }
    }
}

// Compacts, in candidate order, the non-empty candidates among
// [base, base + 4 * blockDim.x) of read r into list[]; returns how many. A
// pair is non-empty iff the read is and D_c >= L[0][c] for every channel.
// Empty pairs get ln P = -inf in the optional pair output. Block-wide; every
// thread must call it.
template <int C>
__device__ int compact_chunk(const ClassifyArgs& a, int r, int ncand, const int* cidx, int base,
                             bool read_empty, const int* sL, int* list, int* scan) {
    const int tid = threadIdx.x, nthreads = blockDim.x;
    int cnt = 0, flags = 0;
    const int kk0 = base + tid * 4;
    for (int q = 0; q < 4; q++) {
        int kk = kk0 + q;
        if (kk >= ncand) break;
        int i = cidx ? cidx[kk] : kk;
        bool ok = !read_empty && i >= 0 && i < a.S;  // an index outside the table scores as empty
        if (ok) {
#pragma unroll
            for (int c = 0; c < C; c++) ok = ok && (int)a.dsD[(size_t)i * C + c] >= sL[c];
        }
        if (ok) {
            flags |= 1 << q;
            cnt++;
        } else if (a.out_pair_logp && !cidx) {
            a.out_pair_logp[(size_t)r * a.S + i] = -INFINITY;
        }
    }
    int incl = cnt;
    const int wl = tid & 63, wv = tid >> 6;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        int y = __shfl_up(incl, off, 64);
        if (wl >= off) incl += y;
    }
    if (wl == 63) scan[wv] = incl;
    __syncthreads();
    if (tid == 0) {
        int acc = 0;
        for (int q = 0; q < nthreads / 64; q++) {
            int v = scan[q];
            scan[q] = acc;
            acc += v;
        }
        scan[nthreads / 64] = acc;
    }
    __syncthreads();
    int pos = scan[wv] + incl - cnt;
    for (int q = 0; q < 4; q++)
        if (flags & (1 << q)) list[pos++] = kk0 + q;
    const int nlive = scan[nthreads / 64];
    __syncthreads();
    return nlive;
}

// Merges the groups' accumulators (gres: 4 doubles per group) and writes the
// read's outputs, with the NaN guard of hmm-classifier.h:73-76. In the FP32
// collecting pass (tie_mode 1) it also writes the read's near-best list from
// the groups' tie slots (tmp: kTieMax ints of free LDS). Block-wide.
template <bool kRescale>
__device__ void merge_and_store(const ClassifyArgs& a, int r, int ncand, const int* cidx, int ngroups,
                                int group, int lane, const CandAcc<kRescale>& acc, double* gres,
                                const TieSlot* ties = nullptr, int* tmp = nullptr) {
    if (group < ngroups && lane == 0) {
        gres[group * 4 + 0] = acc.best_v;
        gres[group * 4 + 1] = (double)acc.best_k;
        gres[group * 4 + 2] = acc.tot_a;
        gres[group * 4 + 3] = acc.tot_b;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    double bv = gres[0];
    int bk = (int)gres[1];
    double ta = gres[2], tb = gres[3];
    for (int g = 1; g < ngroups; g++) {
        double v = gres[g * 4 + 0];
        int k = (int)gres[g * 4 + 1];
        if (v > bv || (v == bv && k < bk)) {
            bv = v;
            bk = k;
        }
        if (kRescale) {
            double xa = gres[g * 4 + 2], xb = gres[g * 4 + 3];
            if (xa > ta) {
                tb = tb * exp(ta - xa) + xb;
                ta = xa;
            } else if (xa > -INFINITY) {
                tb += xb * exp(xa - ta);
            }
        } else {
            ta += gres[g * 4 + 2];
        }
    }
    double log_score, log_total, adjusted;
    if (kRescale) {
        log_score = bv;
        log_total = tb > 0 ? ta + log(tb) : -INFINITY;
        adjusted = exp(log_score - log_total);
    } else {
        log_score = log(bv);
        log_total = log(ta);
        adjusted = bv / ta;
    }
    if (isnan(adjusted)) {
        adjusted = 0.0;
        log_score = -INFINITY;
        log_total = 0.0;
    }
    if (kRescale && a.tie_mode == 1 && ties) {
        int tn = 0;
        if (bv > -INFINITY) {
            const double lim = tie_floor(bv, a.tie_abs, a.tie_rel);
            int cnt = 0;
            bool over = false;
            for (int g = 0; g < ngroups; g++) {
                const TieSlot& s = ties[g];
                if (s.drop >= lim) over = true;
                for (int j = 0; j < s.n; j++) {
                    if (s.v[j] >= lim) {
                        if (cnt < kTieMax) tmp[cnt] = s.kk[j];
                        cnt++;
                    }
                }
            }
            if (over || cnt > kTieMax || bv < kTieFloorAll) {
                tn = -1;
            } else if (cnt >= 2) {
                for (int j = 1; j < cnt; j++) {  // candidate order
                    const int v = tmp[j];
                    int q = j - 1;
                    while (q >= 0 && tmp[q] > v) {
                        tmp[q + 1] = tmp[q];
                        q--;
                    }
                    tmp[q + 1] = v;
                }
                int* out = a.tie_idx + (size_t)r * kTieMax;
                for (int j = 0; j < cnt; j++) out[j] = cidx ? cidx[tmp[j]] : tmp[j];
                tn = cnt;
            }
        }
        a.tie_n[r] = tn;
    }
    int id = -1;
    if (ncand > 0) {
        int i = cidx ? cidx[bk] : bk;
        id = (i >= 0 && i < a.S) ? a.dsId[i] : -1;
    }
    a.out_id[r] = id;
    if (a.out_adjusted) a.out_adjusted[r] = adjusted;
    if (a.out_log_score) a.out_log_score[r] = log_score;
    if (a.out_log_total) a.out_log_total[r] = log_total;
}

// x takes in another FP32 accumulator (best ln P and candidate, log-sum-exp
// pair): the larger ln P wins, ties to the earlier candidate.
__device__ inline void acc_merge(CandAcc<true>& x, double bv, int bk, double ta, double tb) {
    if (bv > x.best_v || (bv == x.best_v && bk < x.best_k)) {
        x.best_v = bv;
        x.best_k = bk;
    }
    if (ta > x.tot_a) {
        x.tot_b = x.tot_b * exp(x.tot_a - ta) + tb;
        x.tot_a = ta;
    } else if (ta > -INFINITY) {
        x.tot_b += tb * exp(ta - x.tot_a);
    }
}

// merge_and_store for many small groups (L lanes each, lane 0 holding the
// group's accumulator and tie slot): each wave first reduces its groups'
// accumulators by shuffles, then the waves are merged; the near-best list is
// gathered from every group's tie slot in parallel (a block prefix sum over
// the groups' counts). red: 16 ints of free LDS; gres: at least 20 doubles;
// tmp: kTieMax ints of free LDS. Block-wide.
template <int L>
__device__ void merge_lanes_and_store(const ClassifyArgs& a, int r, int ncand, const int* cidx, int ngroups,
                                      int group, int p, CandAcc<true> acc, double* gres, const TieSlot* ties,
                                      int* tmp, int* red) {
    const int tid = threadIdx.x, wl = tid & 63, wv = tid >> 6, nw = blockDim.x >> 6;
    const bool lead = p == 0 && group < ngroups;
    double bv = lead ? acc.best_v : -INFINITY;
    int bk = lead ? acc.best_k : 0x7fffffff;
    double ta = lead ? acc.tot_a : -INFINITY;
    double tb = lead ? acc.tot_b : 0.0;
    CandAcc<true> x;
    x.best_v = bv;
    x.best_k = bk;
    x.tot_a = ta;
    x.tot_b = tb;
#pragma unroll
    for (int off = 32; off >= L; off >>= 1) {
        acc_merge(x, __shfl_xor(x.best_v, off, 64), __shfl_xor(x.best_k, off, 64), __shfl_xor(x.tot_a, off, 64),
                  __shfl_xor(x.tot_b, off, 64));
    }
    if (wl == 0) {
        gres[wv * 4 + 0] = x.best_v;
        gres[wv * 4 + 1] = (double)x.best_k;
        gres[wv * 4 + 2] = x.tot_a;
        gres[wv * 4 + 3] = x.tot_b;
    }
    __syncthreads();
    if (tid == 0) {
        CandAcc<true> y;
        y.best_v = gres[0];
        y.best_k = (int)gres[1];
        y.tot_a = gres[2];
        y.tot_b = gres[3];
        for (int w = 1; w < nw; w++) acc_merge(y, gres[w * 4], (int)gres[w * 4 + 1], gres[w * 4 + 2], gres[w * 4 + 3]);
        gres[16] = y.best_v;
        gres[17] = (double)y.best_k;
        gres[18] = y.tot_a;
        gres[19] = y.tot_b;
        red[8] = 0;  // near-best overflow flag
    }
    __syncthreads();
    const double best = gres[16];
    const bool collect = a.tie_mode == 1 && ties && best > -INFINITY;
    if (collect) {
        const double lim = tie_floor(best, a.tie_abs, a.tie_rel);
        int cnt = 0;
        if (lead) {
            const TieSlot& s = ties[group];
            if (s.drop >= lim) atomicOr(&red[8], 1);
            for (int j = 0; j < s.n; j++) cnt += s.v[j] >= lim ? 1 : 0;
        }
        int incl = cnt;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const int y = __shfl_up(incl, off, 64);
            if (wl >= off) incl += y;
        }
        if (wl == 63) red[wv] = incl;
        __syncthreads();
        int pos = incl - cnt;
        for (int w = 0; w < wv; w++) pos += red[w];
        if (lead) {
            const TieSlot& s = ties[group];
            for (int j = 0; j < s.n; j++)
                if (s.v[j] >= lim) {
                    if (pos < kTieMax) tmp[pos] = s.kk[j];
                    pos++;
                }
        }
        __syncthreads();
    }
    if (tid != 0) return;
    bv = gres[16];
    bk = (int)gres[17];
    ta = gres[18];
    tb = gres[19];
    double log_score = bv;
    double log_total = tb > 0 ? ta + log(tb) : -INFINITY;
    double adjusted = exp(log_score - log_total);
    if (isnan(adjusted)) {
        adjusted = 0.0;
        log_score = -INFINITY;
        log_total = 0.0;
    }
    if (a.tie_mode == 1 && ties) {
        int tn = 0;
        if (collect) {
            int cnt = 0;
            for (int w = 0; w < nw; w++) cnt += red[w];
            if (red[8] || cnt > kTieMax || bv < kTieFloorAll) {
                tn = -1;
            } else if (cnt >= 2) {
                for (int j = 1; j < cnt; j++) {  // candidate order
                    const int v = tmp[j];
                    int q = j - 1;
                    while (q >= 0 && tmp[q] > v) {
                        tmp[q + 1] = tmp[q];
                        q--;
                    }
                    tmp[q + 1] = v;
                }
                int* out = a.tie_idx + (size_t)r * kTieMax;
                for (int j = 0; j < cnt; j++) out[j] = cidx ? cidx[tmp[j]] : tmp[j];
                tn = cnt;
            }
        }
        a.tie_n[r] = tn;
    }
    int id = -1;
    if (ncand > 0) {
        const int i = cidx ? cidx[bk] : bk;
        id = (i >= 0 && i < a.S) ? a.dsId[i] : -1;
    }
    a.out_id[r] = id;
    if (a.out_adjusted) a.out_adjusted[r] = adjusted;
    if (a.out_log_score) a.out_log_score[r] = log_score;
    if (a.out_log_total) a.out_log_total[r] = log_total;
}

}  // namespace cc
}  // namespace wp
