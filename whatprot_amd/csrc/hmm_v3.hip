// Classify kernel for reads with small pruned boxes (count box <= NB states
// per timestep): row-per-lane registers, precomputed bleach matrices.
//
// Read-level boxes. For a pair the reference's pruned box at timestep t is
// [0, t+1) x prod_c [L[t][c], min(Hr[t][c], D_c + 1)) (see hmm_kernels.hip).
// Computing every candidate of a read on the read-level box [L, Hr) gives
// identical results: a state with more live dyes than the dye-seq carries
// (n_c > D_c) never receives mass (duds, bleaching and Edman only lower
// counts), and the Edman input bound U = min(H[t] + 1, H[t-1]) only admits
// such zero states when taken at read level. So the box shapes, the composite
// Bleach matrices and the emission densities are shared by all candidates of
// a read; only D (the dud start) and the dye position table (Edman) differ.
//
// Mapping. A workgroup owns one read. Slots of G lanes (G >= T) score one
// candidate each; lane e keeps row e of the state tensors A and B (the NB
// count-box entries) in registers. One timestep t >= 1 is
//   BA = A Wt, BB = B Wt, BA1 = A W1t[c_e]     (composite Bleach, Wt in LDS)
//   succ_e = (1-pf)[(1 - n_c/K) BA + ((n_c+1)/K) BA1]   (Edman success of
//            row e towards row e+1, c = dye channel at position e)
//   A'_e = (1-pd)(1-pcb) (pf BA_e + succ_{e-1}),  B'_e = (1-pd)(BB_e + pcb BA_e)
// then the emission. succ moves to the next lane with one DPP row shift; no
// other cross-lane traffic, no fences. The detached mass is kept per lane
// (it is linear in the row sums), and in FP32 every lane carries its own
// power-of-two exponent, so the only reductions are at the end of a pair.
#include <float.h>
#include <math.h>

#include <algorithm>

#include "classify_common.h"
#include "hmm_kernels.h"

namespace wp {
namespace {

using namespace cc;

template <typename Real>
struct V3Layout {
    size_t rad, L, H, nbt, bstr, coord, etab, eexp, W, binom, list, scan, gres, ties, total;
};

// W block per timestep: (1 + C) matrices of NB x NB: Wt then W1t[c].
template <typename Real>
__host__ __device__ inline V3Layout<Real> v3_layout(int T, int C, int tri, int NB, int ngroups,
                                                    int nthreads) {
    V3Layout<Real> l;
    size_t o = 0;
    l.rad = o;   o = align16(o + sizeof(double) * T * C);
    l.L = o;     o = align16(o + sizeof(int) * T * C);
    l.H = o;     o = align16(o + sizeof(int) * T * C);
    l.nbt = o;   o = align16(o + sizeof(int) * T);
    l.bstr = o;  o = align16(o + sizeof(int) * T * C);
    l.coord = o; o = align16(o + sizeof(int) * (size_t)T * NB * C);
    l.etab = o;  o = align16(o + sizeof(Real) * (size_t)T * NB);
    l.eexp = o;  o = align16(o + sizeof(int) * (size_t)T * NB);
    l.W = o;     o = align16(o + sizeof(Real) * (size_t)T * (1 + C) * NB * NB);
    l.binom = o; o = align16(o + sizeof(double) * 2 * C * tri);
    l.list = o;  o = align16(o + sizeof(int) * nthreads * 4);
    l.scan = o;  o = align16(o + sizeof(int) * (nthreads / 64 + 2));
    l.gres = o;  o = align16(o + sizeof(double) * 4 * ngroups);
    l.ties = o;  o = align16(o + (sizeof(Real) == 4 ? sizeof(TieSlot) * ngroups : 0));
    l.total = o;
    return l;
}

// Value of the previous lane of the same G-lane slot (0 on the slot's first
// lane). G == 16 uses a DPP row shift.
template <int G>
__device__ inline int shift_up_i(int v) {
    if constexpr (G == 16) {
        return __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, false);
    } else {
        int r = __shfl_up(v, 1, G);
        return (threadIdx.x % G) == 0 ? 0 : r;
    }
}
template <int G>
__device__ inline float shift_up(float v) {
    return __int_as_float(shift_up_i<G>(__float_as_int(v)));
}
template <int G>
__device__ inline double shift_up(double v) {
    long long b = __double_as_longlong(v);
    int lo = shift_up_i<G>((int)(b & 0xffffffffll));
    int hi = shift_up_i<G>((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// One read of the row-per-lane kernel. Block-uniform returns only.
template <typename Real, int C, int NB, int G>
__device__ __forceinline__ void v3_read(const ClassifyArgs& a, const int r, unsigned char* smem) {
    constexpr bool kRescale = sizeof(Real) == 4;
    if (r >= a.N) return;
    const int nbr = a.nbmax[r];
    if (!((nbr > a.nb_lo && nbr <= a.nb_hi) || (nbr == 0 && a.nb_lo == 0))) return;
    if (a.wide_pass && !a.wide[r]) return;

    const int T = a.m.T;
    const int tri = a.tri;
    const int tid = threadIdx.x;
    const int nthreads = blockDim.x;
    const int ngroups = nthreads / G;
    const int group = tid / G;
    const int lane = tid % G;
    const V3Layout<Real> lay = v3_layout<Real>(T, C, tri, NB, ngroups, nthreads);
    double* s_rad = (double*)(smem + lay.rad);
    int* sL = (int*)(smem + lay.L);
    int* sH = (int*)(smem + lay.H);
    int* nbt = (int*)(smem + lay.nbt);
    int* bstr = (int*)(smem + lay.bstr);
    int* coord = (int*)(smem + lay.coord);
    Real* etab = (Real*)(smem + lay.etab);
    int* eexp = (int*)(smem + lay.eexp);
    Real* Wall = (Real*)(smem + lay.W);
    double* bin = (double*)(smem + lay.binom);
    int* list = (int*)(smem + lay.list);
    int* scan = (int*)(smem + lay.scan);
    double* gres = (double*)(smem + lay.gres);
    TieSlot* ties = (TieSlot*)(smem + lay.ties);
    int ncand;
    const int* cidx;
    read_candidates(a, r, &ncand, &cidx);
    const bool read_empty = nbr == 0;

    // ---- prologue: read-level boxes, coordinates, emissions, bleach matrices
    for (int k = tid; k < T * C; k += nthreads) {
        s_rad[k] = a.rads[(size_t)r * T * C + k];
        sL[k] = a.Lr[(size_t)r * T * C + k];
        sH[k] = a.Hr[(size_t)r * T * C + k];
    }
    for (int k = tid; k < 2 * C * tri; k += nthreads) bin[k] = a.binom[k];
    __syncthreads();
    if (tid < T) {
        const int t = tid;
        int s = 1;
        for (int c = C - 1; c >= 0; c--) {
            bstr[t * C + c] = s;
            int w = sH[t * C + c] - sL[t * C + c];
            s *= w > 0 ? w : 0;
        }
        nbt[t] = read_empty ? 0 : s;
    }
    __syncthreads();
    // coordinates and emission densities (peptide-emission.cc:165-180)
    for (int k = tid; k < T * NB; k += nthreads) {
        const int t = k / NB, q = k % NB;
        unsigned int n[WP_MAX_CHANNELS];
        if (q < nbt[t]) {
            int rem = q;
#pragma unroll
            for (int c = 0; c < C; c++) {
                int d = rem / bstr[t * C + c];
                rem -= d * bstr[t * C + c];
                n[c] = (unsigned int)(sL[t * C + c] + d);
                coord[(t * NB + q) * C + c] = (int)n[c];
            }
            double v = 1.0;
#pragma unroll
            for (int c = 0; c < C; c++) v *= dev_pdf<C>(a.m, c, s_rad[t * C + c], n);
            if constexpr (kRescale) {
                int ex = 0;
                double mant = frexp(v, &ex);
                etab[k] = (Real)mant;
                eexp[k] = v > 0.0 ? ex : kNoExp;
            } else {
                etab[k] = (Real)v;
            }
        } else {
#pragma unroll
            for (int c = 0; c < C; c++) coord[(t * NB + q) * C + c] = 0;
            etab[k] = 0;
            if constexpr (kRescale) eexp[k] = kNoExp;
        }
    }
    __syncthreads();
    // Composite Bleach between the boxes of t-1 and t, and its Edman-shifted
    // forms: W1t[c][q][q2] maps onto n'(q2) + 1_c and is zero where that lies
    // at or above U = min(Hr[t] + 1, Hr[t-1]) (edman-transition.cc:28-60).
    const double* ble = bin + C * tri;
    for (int k = tid; k < T * (1 + C) * NB * NB; k += nthreads) {
        const int t = k / ((1 + C) * NB * NB);
        int rem = k - t * (1 + C) * NB * NB;
        const int mat = rem / (NB * NB);
        rem -= mat * NB * NB;
        const int q = rem / NB, q2 = rem % NB;
        double v = 0.0;
        if (t >= 1 && q < nbt[t - 1] && q2 < nbt[t]) {
            v = 1.0;
            const int sc = mat - 1;  // shifted channel, -1 for Wt
            for (int c = 0; c < C; c++) {
                int from = coord[((t - 1) * NB + q) * C + c];
                int to = coord[(t * NB + q2) * C + c] + (c == sc ? 1 : 0);
                if (c == sc) {
                    int U = min(sH[t * C + c] + 1, sH[(t - 1) * C + c]);
                    if (to >= U) v = 0.0;
                }
                if (from < to) v = 0.0;
                else v *= ble[c * tri + from * (from + 1) / 2 + to];
            }
        }
        Wall[k] = (Real)v;
    }
    __syncthreads();

    CandAcc<kRescale> acc;
    const bool collect = kRescale && a.tie_mode == 1;
    if (collect && lane == 0) tie_init(&ties[group]);
    const Real pf = (Real)a.m.pf;
    const Real one_m_pf = (Real)(1.0 - a.m.pf);
    const double* dud = bin;
    const int e = lane;

    for (int base = 0; base < ncand; base += nthreads * 4) {
        const int nlive = compact_chunk<C>(a, r, ncand, cidx, base, read_empty, sL, list, scan);
        // Per-candidate inputs, fetched one candidate ahead so the global
        // loads overlap the previous candidate's recursion.
        auto fetch = [&](int q0, int& kk, int& i, int* D, int& ce, int& Ke, int& cnt) {
            kk = list[q0];
            i = cidx ? cidx[kk] : kk;
#pragma unroll
            for (int c = 0; c < C; c++) D[c] = a.dsD[(size_t)i * C + c];
            ce = e < T ? (int)a.dsSeq[(size_t)i * T + e] : -1;
            Ke = ce >= 0 ? (int)a.dsTrack[((size_t)i * T + e) * C + ce] : 1;
            cnt = a.dsCount[i];
        };
        int nkk = 0, ni = 0, nD[C], nce = -1, nKe = 1, ncnt = 0;
        if (group < nlive) fetch(group, nkk, ni, nD, nce, nKe, ncnt);
        for (int q0 = group; q0 < nlive; q0 += ngroups) {
            const int kk = nkk, i = ni, ce = nce, Ke = nKe, cntw = ncnt;
            int D[C];
#pragma unroll
            for (int c = 0; c < C; c++) D[c] = nD[c];
            if (q0 + ngroups < nlive) fetch(q0 + ngroups, nkk, ni, nD, nce, nKe, ncnt);
            const int cix = ce < 0 ? 0 : ce;
            Real invK = (Real)1 / (Real)Ke;
            Real A[NB], B[NB];
            int kexp = 0;       // FP32: the row holds values x 2^-kexp
            Real dl = 0;        // this row's share of the detached mass
            Real Srow = 0;      // sum of the row after the last emission

            // t = 0: InitialBlock, Dud_c, Emit(0) (row 0 only)
            {
                const Real pib = (Real)a.m.p_ib;
                int km = kNoExp;
#pragma unroll
                for (int q = 0; q < NB; q++) {
                    Real va = 0, vb = 0;
                    if (e == 0) {
                        bool ok = true;
                        Real p = 1;
                        Real xa = (Real)1 - pib, xb = pib;
#pragma unroll
                        for (int c = 0; c < C; c++) {
                            int n = coord[q * C + c];
                            ok = ok && n <= D[c];
                            if (ok) {
                                p = (Real)dud[c * tri + D[c] * (D[c] + 1) / 2 + n];
                                xa = p * xa;
                                xb = p * xb;
                            }
                        }
                        if (ok) {
                            va = xa * etab[q];
                            vb = xb * etab[q];
                            if constexpr (kRescale) km = max(km, max(exp_of(va), exp_of(vb)) + eexp[q]);
                        }
                    }
                    A[q] = va;
                    B[q] = vb;
                }
                if constexpr (kRescale) {
                    const int kf = km == kNoExp ? 0 : km;
#pragma unroll
                    for (int q = 0; q < NB; q++) {
                        int sh = eexp[q] - kf;
                        A[q] = ldexpf(A[q], sh);
                        B[q] = ldexpf(B[q], sh);
                    }
                    kexp = km == kNoExp ? kNoExp : kf;
                }
#pragma unroll
                for (int q = 0; q < NB; q++) Srow += A[q] + B[q];
            }

            for (int t = 1; t < T; t++) {
                // Boxes below NB states skip their zero padding with uniform
                // branches for NB > 4; at NB = 4 straight-line code is faster.
                constexpr bool kGuard = NB > 4;
                const int nbp = kGuard ? nbt[t - 1] : NB;
                const Real* Wt = Wall + (size_t)t * (1 + C) * NB * NB;
                const Real* W1 = Wt + (size_t)(1 + cix) * NB * NB;
                Real BA[NB], BB[NB], BA1[NB];
#pragma unroll
                for (int q2 = 0; q2 < NB; q2++) {
                    BA[q2] = 0;
                    BB[q2] = 0;
                    BA1[q2] = 0;
                }
#pragma unroll
                for (int q = 0; q < NB; q++) {
                    if (kGuard && q >= nbp) break;
                    const Real av = A[q], bv = B[q];
#pragma unroll
                    for (int q2 = 0; q2 < NB; q2++) {
                        const Real w = Wt[q * NB + q2];
                        BA[q2] += w * av;
                        BB[q2] += w * bv;
                        BA1[q2] += W1[q * NB + q2] * av;
                    }
                }
                // Detach(t-1): this row's share, pd times its previous sum.
                const Real pd = (Real)a.m.p_detach[t - 1];
                const Real pcb = (Real)a.m.p_cb;
                const Real cAA = ((Real)1 - pd) * ((Real)1 - pcb);
                const Real cB = (Real)1 - pd;
                Real dlp = dl + pd * Srow;
                // Edman success of this row towards row e+1.
                const int* ct = coord + (size_t)t * NB * C;
                const int kfrom = kRescale ? shift_up_i<G>(kexp) : 0;
                Real nA[NB], nB[NB];
                int kn = 0, sh_own = 0, sh_from = 0;
                if constexpr (kRescale) {
                    const int ko = kexp, kf = (lane == 0) ? kNoExp : kfrom;
                    kn = max(ko, kf);
                    if (kn == kNoExp) kn = 0;
                    sh_own = ko == kNoExp ? -200 : ko - kn;
                    sh_from = kf == kNoExp ? -200 : kf - kn;
                }
#pragma unroll
                for (int q2 = 0; q2 < NB; q2++) {
                    const int kc = ct[q2 * C + cix];
                    Real ratio, ratio2;
                    if constexpr (kRescale) {
                        ratio = (Real)kc * invK;
                        ratio2 = (Real)(kc + 1) * invK;
                    } else {
                        ratio = (Real)kc / (Real)Ke;
                        ratio2 = (Real)(kc + 1) / (Real)Ke;
                    }
                    const Real keep = kc < Ke ? one_m_pf * ((Real)1 - ratio) * BA[q2] : (Real)0;
                    const Real succ = ce < 0 ? one_m_pf * BA[q2] : keep + one_m_pf * ratio2 * BA1[q2];
                    const Real from = shift_up<G>(succ);
                    Real own = pf * BA[q2], bown = cB * (BB[q2] + pcb * BA[q2]);
                    if constexpr (kRescale) {
                        nA[q2] = cAA * (ldexpf(own, sh_own) + ldexpf(from, sh_from));
                        nB[q2] = ldexpf(bown, sh_own);
                    } else {
                        nA[q2] = cAA * (own + from);
                        nB[q2] = bown;
                    }
                }
                if constexpr (kRescale) dlp = ldexpf(dlp, sh_own);
                // Emit(t) (peptide-emission.cc:212-243); the detached share
                // keeps density E_t(0) only if the box includes zero.
                bool allow = true;
#pragma unroll
                for (int c = 0; c < C; c++) allow = allow && sL[t * C + c] == 0;
                const Real* Et = etab + (size_t)t * NB;
                if constexpr (kRescale) {
                    const int* Xt = eexp + (size_t)t * NB;
                    int km = kNoExp;
#pragma unroll
                    for (int q2 = 0; q2 < NB; q2++) {
                        nA[q2] *= Et[q2];
                        nB[q2] *= Et[q2];
                        km = max(km, max(exp_of(nA[q2]), exp_of(nB[q2])) + Xt[q2]);
                    }
                    if (allow) {
                        dlp *= Et[0];
                        km = max(km, exp_of(dlp) + Xt[0]);
                    } else {
                        dlp = 0;
                    }
                    const int kf = km == kNoExp ? 0 : km;
                    Srow = 0;
#pragma unroll
                    for (int q2 = 0; q2 < NB; q2++) {
                        const int sh = Xt[q2] - kf;
                        A[q2] = ldexpf(nA[q2], sh);
                        B[q2] = ldexpf(nB[q2], sh);
                        Srow += A[q2] + B[q2];
                    }
                    dl = allow ? ldexpf(dlp, Xt[0] - kf) : (Real)0;
                    kexp = km == kNoExp ? kNoExp : kn + kf;
                } else {
                    Srow = 0;
#pragma unroll
                    for (int q2 = 0; q2 < NB; q2++) {
                        const Real em = Et[q2];
                        A[q2] = nA[q2] * em;
                        B[q2] = nB[q2] * em;
                        Srow += A[q2] + B[q2];
                    }
                    dl = allow ? dlp * Et[0] : (Real)0;
                }
            }

            // P = sum over rows of (row sum + detached share) (peptide-state-vector.cc:52-54)
            double logp, p;
            if constexpr (kRescale) {
                const Real own = Srow + dl;
                const int kme = own > 0 ? kexp : kNoExp;
                const int K = group_max_int<G>(kme);
                float part = (K == kNoExp || kme == kNoExp) ? 0.0f : ldexpf(own, kme - K);
                float s = group_sum<float, G>(part);
                logp = s > 0 ? log((double)s) + (double)K * 0.69314718055994530942 : -INFINITY;
                p = exp(logp);
            } else {
                p = group_sum<double, G>(Srow + dl);
                logp = log(p);
            }
            if (a.out_pair_logp && !cidx && lane == 0) a.out_pair_logp[(size_t)r * a.S + i] = logp;
            acc.add(logp, p, kk, cntw);
            if (collect && lane == 0) tie_note(&ties[group], acc.best_v, logp, kk, a.tie_abs, a.tie_rel);
        }
        __syncthreads();
    }
    merge_and_store<kRescale>(a, r, ncand, cidx, ngroups, group, lane, acc, gres, ties, list);
}

// kStrided false: one workgroup per read of the bin's compacted list
// (workgroups past its length exit). kStrided: workgroups stride over the
// list, for lists that hold few of the call's reads; the barrier keeps a
// read's LDS tables from being overwritten while the previous read's outputs
// are written. Two instantiations, because the loop holds read-invariant
// values across reads (164 -> 250 registers in FP64, 3 -> 2 waves per SIMD).
template <typename Real, int C, int NB, int G, bool kStrided>
__global__ void __launch_bounds__(kBlockThreads) classify_v3_kernel(const ClassifyArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int n = *a.list_count;
    if constexpr (!kStrided) {
        if ((int)blockIdx.x < n) v3_read<Real, C, NB, G>(a, a.list[blockIdx.x], smem);
    } else {
        for (int k = blockIdx.x; k < n; k += gridDim.x) {
            if (k != (int)blockIdx.x) __syncthreads();
            v3_read<Real, C, NB, G>(a, a.list[k], smem);
        }
    }
}

template <typename Real, int C, int NB, int G>
hipError_t launch_v3_t(const ClassifyArgs& a, hipStream_t stream) {
    V3Layout<Real> l = v3_layout<Real>(a.m.T, C, a.tri, NB, kBlockThreads / G, kBlockThreads);
    const bool strided = list_grid(a) < a.N;
    const void* fn = strided ? (const void*)classify_v3_kernel<Real, C, NB, G, true>
                             : (const void*)classify_v3_kernel<Real, C, NB, G, false>;
    if (l.total > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)l.total);
        if (e != hipSuccess) return e;
    }
    if (strided)
        hipLaunchKernelGGL((classify_v3_kernel<Real, C, NB, G, true>), dim3(list_grid(a)), dim3(kBlockThreads), l.total,
                           stream, a);
    else
        hipLaunchKernelGGL((classify_v3_kernel<Real, C, NB, G, false>), dim3(a.N), dim3(kBlockThreads), l.total, stream,
                           a);
    return hipGetLastError();
}

template <typename Real, int C, int NB>
hipError_t launch_v3_g(const ClassifyArgs& a, hipStream_t stream) {
    if (a.m.T <= 16) return launch_v3_t<Real, C, NB, 16>(a, stream);
    return launch_v3_t<Real, C, NB, 32>(a, stream);
}

template <typename Real, int C>
hipError_t launch_v3_nb(const ClassifyArgs& a, hipStream_t stream) {
    switch (a.NB) {
    case 4: return launch_v3_g<Real, C, 4>(a, stream);
    case 8: return launch_v3_g<Real, C, 8>(a, stream);
    default: return hipErrorInvalidValue;
    }
}

template <typename Real>
hipError_t launch_v3_r(const ClassifyArgs& a, hipStream_t stream) {
    switch (a.m.C) {
    case 1: return launch_v3_nb<Real, 1>(a, stream);
    case 2: return launch_v3_nb<Real, 2>(a, stream);
    case 3: return launch_v3_nb<Real, 3>(a, stream);
    case 4: return launch_v3_nb<Real, 4>(a, stream);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace

bool v3_supports(int precision, int T, int NB) {
    (void)precision;
    return T <= 32 && (NB == 4 || NB == 8);
}

size_t v3_lds_bytes(int precision, int T, int C, int tri, int NB) {
    const int G = T <= 16 ? 16 : 32;
    if (precision == WP_PRECISION_FP32)
        return v3_layout<float>(T, C, tri, NB, kBlockThreads / G, kBlockThreads).total;
    return v3_layout<double>(T, C, tri, NB, kBlockThreads / G, kBlockThreads).total;
}

hipError_t launch_classify_v3(const ClassifyArgs& a, int precision, hipStream_t stream) {
    if (a.N <= 0) return hipSuccess;
    if (precision == WP_PRECISION_FP32) return launch_v3_r<float>(a, stream);
    return launch_v3_r<double>(a, stream);
}

}  // namespace wp
