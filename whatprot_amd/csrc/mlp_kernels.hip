// Kernels of the learned shortlist: a BF16 MLP on the matrix cores.
//
// GEMM: C(m, n) = sum_k A(m, k) B(k, n) with bf16 operands read through
// arbitrary (row, column) strides -- the forward product X W, the backward
// products dY W^T and X^T dY all use it -- FP32 accumulation on
// v_mfma_f32_16x16x32_bf16, and a fused epilogue (bias, ReLU, ReLU-derivative
// mask, FP32 and/or BF16 stores). A 256-thread workgroup owns a 128 x 128
// tile of C; its four wavefronts own 64 x 64 quarters as 4 x 4 MFMA tiles.
// K advances 32 at a time through LDS tiles stored k-contiguous for both
// operands, so every MFMA operand is one 16-byte LDS read.
//
// MFMA 16x16x32 lane layout (lane l, g = l / 16, r = l % 16): A operand row r,
// k = 8g .. 8g+7; B operand column r, k = 8g .. 8g+7; accumulator rows
// 4g .. 4g+3 of column r.
#include <hip/hip_runtime.h>
#include <float.h>
#include <math.h>
#include <stdint.h>

#include <algorithm>

#include "mlp_kernels.h"

namespace wp {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int BM = 128, BN = 128, BK = 32, PAD = 8, LDK = BK + PAD;

__device__ inline __bf16 to_bf16(float x) { return (__bf16)x; }
__device__ inline float from_bf16(__bf16 x) { return (float)x; }

// Order-preserving unsigned image of a float (larger float, larger key; -0 < +0).
__device__ inline unsigned int fkey(float f) {
    unsigned int u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

// The logit an order-preserving key stands for (the inverse of fkey).
__device__ inline float key_value(unsigned int k) {
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}

// Tile (tx, ty) of workgroup w of an nx x ny grid, w in dispatch order.
// Workgroups go to the 8 XCDs round-robin (XCD w % 8), so XCD x is given the
// consecutive tiles [x per, (x + 1) per) of a band-major order -- bands of
// ceil(nx / 8) column tiles, each swept row tile by row tile -- and its L2
// holds one band of B and the few A row tiles in flight. A permutation of
// the tiles: every tile is computed once, as before.
__device__ inline void tile_of(int w, int nx, int ny, int& tx, int& ty) {
    const int per = nx * ny / 8;
    const int t = w < 8 * per ? (w % 8) * per + w / 8 : w;
    const int bw = (nx + 7) / 8, full = nx / bw, tb = ny * bw;
    if (t < full * tb) {
        const int rem = t % tb;
        ty = rem / bw;
        tx = (t / tb) * bw + rem % bw;
    } else {  // the narrower last band
        const int wl = nx - full * bw, rem = t - full * tb;
        ty = rem / wl;
        tx = full * bw + rem % wl;
    }
}

// Exchange with the partner lane of step S of the 16-lane transposing
// reductions below: lane l ^ 8 (row_ror:8), l ^ 7 (row_half_mirror), l ^ 2,
// l ^ 1 (quad permutations) -- four steps that together reach all 16 lanes
// of a row.
template <int S>
__device__ inline unsigned int xchg(unsigned int v) {
    constexpr int ctrl = S == 0 ? 0x128 : S == 1 ? 0x141 : S == 2 ? (2 | (3 << 2) | (0 << 4) | (1 << 6))
                                                                  : (1 | (0 << 2) | (3 << 4) | (2 << 6));
    return (unsigned int)__builtin_amdgcn_update_dpp(0, (int)v, ctrl, 0xf, 0xf, false);
}
template <int S>
__device__ inline float xchg(float v) { return __uint_as_float(xchg<S>(__float_as_uint(v))); }

// Lane bit that step S tells apart (the partners of the step differ in it).
template <int S>
__device__ inline bool step_bit(int lr) { return (lr >> (3 - S)) & 1; }

// Step S of a transposing reduction: of its 2n entries (entry e + n in the
// upper half) a lane keeps the half whose index bit equals its lane bit and
// combines it with the partner's copy of that half, which the partner sends.
template <int S, int n, typename T, typename Op>
__device__ inline void treduce(T* v, int lr, Op op) {
    const bool b = step_bit<S>(lr);
#pragma unroll
    for (int e = 0; e < n; e++) {
        const T keep = b ? v[e + n] : v[e];
        const T send = b ? v[e] : v[e + n];
        v[e] = op(keep, xchg<S>(send));
    }
}

// Step S of the inverse (an all-gather): n entries become 2n, the lane's own
// in the half its lane bit names and the partner's in the other.
template <int S, int n>
__device__ inline void tgather(float* v, int lr) {
    const bool b = step_bit<S>(lr);
#pragma unroll
    for (int e = n - 1; e >= 0; e--) {
        const float own = v[e], other = xchg<S>(v[e]);
        v[e + n] = b ? own : other;
        v[e] = b ? other : own;
    }
}

// Row statistics of a wavefront's 64 x 64 quarter of C for the top-k (the
// GemmArgs gkey / lsp outputs). In the accumulator layout the 16 lanes of a
// row of lanes (l / 16 = lg) hold the 16 columns of one group for 16
// (i, r) combinations q = 4 i + r. Each half of the combinations (i / 2 = h)
// goes through a transposing reduction over lanes l ^ 7, l ^ 2, l ^ 1, which
// leaves lane lr with combination q = 8 h + lr % 8 reduced over 8 lanes, and
// a last exchange with l ^ 8 completes it on both lanes; the lanes with
// lr / 8 = h store it: the row's four group keys as one 16-byte word, its
// slab partial as one 8-byte word.
__device__ inline void row_stats(const GemmArgs& g, const f32x4 (&acc)[4][4], int m0, int n0, int lr, int lg) {
    const long G = stat_groups(g.N), NS = stat_slabs(g.N);
    auto kmax = [](unsigned int a, unsigned int b) { return max(a, b); };
    auto fadd = [](float a, float b) { return a + b; };
    float mx[16];  // slab maxima of every combination, on every lane of the row
#pragma unroll
    for (int h = 0; h < 2; h++) {
        unsigned int key[32];  // [q % 8][j]
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int j = 0; j < 4; j++)
#pragma unroll
                for (int r = 0; r < 4; r++)
                    key[(4 * i + r) * 4 + j] = n0 + 16 * j + lr < g.N ? fkey(acc[2 * h + i][j][r]) : 0u;
        treduce<1, 16>(key, lr, kmax);
        treduce<2, 8>(key, lr, kmax);
        treduce<3, 4>(key, lr, kmax);
#pragma unroll
        for (int j = 0; j < 4; j++) key[j] = max(key[j], xchg<0>(key[j]));
        const int m = m0 + 16 * (2 * h + ((lr >> 2) & 1)) + 4 * lg + (lr & 3);
        if ((lr >> 3) == h && m < g.M)
            *(uint4*)&g.gkey[(long)m * G + n0 / 16] = uint4{key[0], key[1], key[2], key[3]};
        float* x = mx + 8 * h;
        x[0] = key_value(max(max(key[0], key[1]), max(key[2], key[3])));
        tgather<3, 1>(x, lr);
        tgather<2, 2>(x, lr);
        tgather<1, 4>(x, lr);
    }
#pragma unroll
    for (int h = 0; h < 2; h++) {
        float sum[8];
#pragma unroll
        for (int i = 0; i < 2; i++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int q = 4 * i + r;
                float t = 0.0f;
#pragma unroll
                for (int j = 0; j < 4; j++)
                    if (n0 + 16 * j + lr < g.N) t += __expf(acc[2 * h + i][j][r] - mx[8 * h + q]);
                sum[q] = t;
            }
        treduce<1, 4>(sum, lr, fadd);
        treduce<2, 2>(sum, lr, fadd);
        treduce<3, 1>(sum, lr, fadd);
        sum[0] += xchg<0>(sum[0]);
        const int m = m0 + 16 * (2 * h + ((lr >> 2) & 1)) + 4 * lg + (lr & 3);
        if ((lr >> 3) == h && m < g.M) g.lsp[(long)m * NS + n0 / 64] = float2{mx[8 * h + (lr & 7)], sum[0]};
    }
}

// kKC: both operands k-contiguous with rows of a multiple of 8 (the forward
// products; the host checks) -- only that load path is compiled, which keeps
// the register count of the prefetching main loop down. kStats: the row
// statistics epilogue (GemmArgs gkey, lsp).
template <bool kKC, bool kStats>
__device__ inline void gemm_body(const GemmArgs& g) {
    __shared__ __align__(16) __bf16 As[BM * LDK];
    __shared__ __align__(16) __bf16 Bs[BN * LDK];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    int tx, ty;
    tile_of(blockIdx.y * gridDim.x + blockIdx.x, gridDim.x, gridDim.y, tx, ty);
    const int m0 = ty * BM, n0 = tx * BN;
    const int wm = (wv >> 1) * 64, wn = (wv & 1) * 64;
    const int lr = lane & 15, lg = lane >> 4;
    f32x4 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    // 16-byte loads: along k when k is contiguous (stored as is), else along
    // the row index (m or n) when that is contiguous (scattered into LDS);
    // a vector that would cross the matrix edge is loaded element-wise
    const int a_mode = kKC ? 1 : (g.sak == 1 && g.sam % 8 == 0) ? 1 : (g.sam == 1 && g.sak % 8 == 0) ? 2 : 0;
    const int b_mode = kKC ? 1 : (g.sbk == 1 && g.sbn % 8 == 0) ? 1 : (g.sbn == 1 && g.sbk % 8 == 0) ? 2 : 0;
    auto mfma_tile = [&]() {
        bf16x8 af[4], bfr[4];
#pragma unroll
        for (int i = 0; i < 4; i++) af[i] = *(const bf16x8*)&As[(wm + 16 * i + lr) * LDK + 8 * lg];
#pragma unroll
        for (int j = 0; j < 4; j++) bfr[j] = *(const bf16x8*)&Bs[(wn + 16 * j + lr) * LDK + 8 * lg];
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int j = 0; j < 4; j++)
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i], bfr[j], acc[i][j], 0, 0, 0);
    };
    if (kKC) {
        // both operands k-contiguous (the forward product): the next K tile's
        // 16-byte loads are in flight while the current tile's MFMAs run
        bf16x8 ra[2], rb[2];
        auto fetch = [&](int k0) {
#pragma unroll
            for (int it = 0; it < 2; it++) {
                const int e = tid + it * 256, row = e >> 2, kv = (e & 3) * 8, gk = k0 + kv;
                const int gm = m0 + row, gn = n0 + row;
                if (gm < g.M && gk + 8 <= g.K) {
                    ra[it] = *(const bf16x8*)&g.A[(long)gm * g.sam + gk];
                } else {
#pragma unroll
                    for (int q = 0; q < 8; q++)
                        ra[it][q] = (gm < g.M && gk + q < g.K) ? g.A[(long)gm * g.sam + gk + q] : (__bf16)0.0f;
                }
                if (gn < g.N && gk + 8 <= g.K) {
                    rb[it] = *(const bf16x8*)&g.B[(long)gn * g.sbn + gk];
                } else {
#pragma unroll
                    for (int q = 0; q < 8; q++)
                        rb[it][q] = (gn < g.N && gk + q < g.K) ? g.B[(long)gn * g.sbn + gk + q] : (__bf16)0.0f;
                }
            }
        };
        fetch(0);
        for (int k0 = 0; k0 < g.K; k0 += BK) {
#pragma unroll
            for (int it = 0; it < 2; it++) {
                const int e = tid + it * 256, row = e >> 2, kv = (e & 3) * 8;
                *(bf16x8*)&As[row * LDK + kv] = ra[it];
                *(bf16x8*)&Bs[row * LDK + kv] = rb[it];
            }
            __syncthreads();
            if (k0 + BK < g.K) fetch(k0 + BK);
            mfma_tile();
            __syncthreads();
        }
    }
    for (int k0 = 0; k0 < g.K && !kKC; k0 += BK) {
        if (a_mode == 1) {
#pragma unroll
            for (int it = 0; it < 2; it++) {
                int e = tid + it * 256;
                int m = e >> 2, kv = (e & 3) * 8;
                int gm = m0 + m, gk = k0 + kv;
                bf16x8 v;
                if (gm < g.M && gk + 8 <= g.K) {
                    v = *(const bf16x8*)&g.A[(long)gm * g.sam + gk];
                } else {
#pragma unroll
                    for (int q = 0; q < 8; q++)
                        v[q] = (gm < g.M && gk + q < g.K) ? g.A[(long)gm * g.sam + gk + q] : (__bf16)0.0f;
                }
                *(bf16x8*)&As[m * LDK + kv] = v;
            }
        } else if (a_mode == 2) {
#pragma unroll
            for (int it = 0; it < 2; it++) {
                int e = tid + it * 256;
                int k = e >> 4, mv = (e & 15) * 8;
                int gm = m0 + mv, gk = k0 + k;
                bf16x8 v;
                if (gk < g.K && gm + 8 <= g.M) {
                    v = *(const bf16x8*)&g.A[(long)gk * g.sak + gm];
                } else {
#pragma unroll
                    for (int q = 0; q < 8; q++)
                        v[q] = (gk < g.K && gm + q < g.M) ? g.A[(long)gk * g.sak + gm + q] : (__bf16)0.0f;
                }
#pragma unroll
                for (int q = 0; q < 8; q++) As[(mv + q) * LDK + k] = v[q];
            }
        } else {
            for (int it = 0; it < BM * BK / 256; it++) {
                int e = tid + it * 256;
                int m = e / BK, k = e % BK;
                int gm = m0 + m, gk = k0 + k;
                __bf16 v = (__bf16)0.0f;
                if (gm < g.M && gk < g.K) v = g.A[(long)gm * g.sam + (long)gk * g.sak];
                As[m * LDK + k] = v;
            }
        }
        if (b_mode == 1) {
#pragma unroll
            for (int it = 0; it < 2; it++) {
                int e = tid + it * 256;
                int n = e >> 2, kv = (e & 3) * 8;
                int gn = n0 + n, gk = k0 + kv;
                bf16x8 v;
                if (gn < g.N && gk + 8 <= g.K) {
                    v = *(const bf16x8*)&g.B[(long)gn * g.sbn + gk];
                } else {
#pragma unroll
                    for (int q = 0; q < 8; q++)
                        v[q] = (gn < g.N && gk + q < g.K) ? g.B[(long)gn * g.sbn + gk + q] : (__bf16)0.0f;
                }
                *(bf16x8*)&Bs[n * LDK + kv] = v;
            }
        } else if (b_mode == 2) {
#pragma unroll
            for (int it = 0; it < 2; it++) {
                int e = tid + it * 256;
                int k = e >> 4, nv = (e & 15) * 8;
                int gn = n0 + nv, gk = k0 + k;
                bf16x8 v;
                if (gk < g.K && gn + 8 <= g.N) {
                    v = *(const bf16x8*)&g.B[(long)gk * g.sbk + gn];
                } else {
#pragma unroll
                    for (int q = 0; q < 8; q++)
                        v[q] = (gk < g.K && gn + q < g.N) ? g.B[(long)gk * g.sbk + gn + q] : (__bf16)0.0f;
                }
#pragma unroll
                for (int q = 0; q < 8; q++) Bs[(nv + q) * LDK + k] = v[q];
            }
        } else {
            for (int it = 0; it < BN * BK / 256; it++) {
                int e = tid + it * 256;
                int n = e / BK, k = e % BK;
                int gn = n0 + n, gk = k0 + k;
                __bf16 v = (__bf16)0.0f;
                if (gn < g.N && gk < g.K) v = g.B[(long)gk * g.sbk + (long)gn * g.sbn];
                Bs[n * LDK + k] = v;
            }
        }
        __syncthreads();
        mfma_tile();
        __syncthreads();
    }
    // epilogue: the stored values, kept in acc for the row statistics
#pragma unroll
    for (int i = 0; i < 4; i++) {
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int n = n0 + wn + 16 * j + lr;
            if (n >= g.N) continue;
            const float bias = g.bias ? g.bias[n] : 0.0f;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int m = m0 + wm + 16 * i + 4 * lg + r;
                if (m >= g.M) continue;
                float v = acc[i][j][r] + bias;
                if (g.relu) v = fmaxf(v, 0.0f);
                if (g.mask && !(from_bf16(g.mask[(long)m * g.ldm + n]) > 0.0f)) v = 0.0f;
                acc[i][j][r] = v;
                if (g.Cf) {
                    // the logits of the top-k path: streamed out (the top-k
                    // reads back only the few groups that reach its threshold)
                    if (kStats)
                        __builtin_nontemporal_store(v, &g.Cf[(long)m * g.ldc + n]);
                    else
                        g.Cf[(long)m * g.ldc + n] = v;
                }
                if (g.Cb) g.Cb[(long)m * g.ldc + n] = to_bf16(v);
            }
        }
    }
    if (kStats && n0 + wn < g.N) row_stats(g, acc, m0 + wm, n0 + wn, lr, lg);
}

// The forward products run at 3 waves per SIMD (the statistics epilogue
// included); the general strided form, for the backward products, at 2.
__global__ void __launch_bounds__(256, 3) gemm_kc_kernel(const GemmArgs g) { gemm_body<true, false>(g); }
__global__ void __launch_bounds__(256, 3) gemm_kc_stats_kernel(const GemmArgs g) { gemm_body<true, true>(g); }
__global__ void __launch_bounds__(256) gemm_kernel(const GemmArgs g) { gemm_body<false, false>(g); }

// Row-wise softmax cross-entropy over N logits (one workgroup per row):
// loss_r = logsumexp(z_r) - z_r[y_r]; dZ = (softmax - onehot) * scale as
// bf16; max posterior = exp(max - logsumexp).
__global__ void __launch_bounds__(256) softmax_ce_kernel(int N, const float* __restrict__ Z, long ldz,
                                                          const int* __restrict__ labels, float scale,
                                                          __bf16* __restrict__ dZ, float* __restrict__ loss,
                                                          float* __restrict__ max_post, float* __restrict__ lse_out) {
    __shared__ float red[8];
    const int r = blockIdx.x, tid = threadIdx.x;
    const float* z = Z + (long)r * ldz;
    float mx = -FLT_MAX;
    for (int n = tid; n < N; n += 256) mx = fmaxf(mx, z[n]);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off, 64));
    if ((tid & 63) == 0) red[tid >> 6] = mx;
    __syncthreads();
    mx = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
    __syncthreads();
    float s = 0.0f;
    for (int n = tid; n < N; n += 256) s += expf(z[n] - mx);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if ((tid & 63) == 0) red[tid >> 6] = s;
    __syncthreads();
    s = red[0] + red[1] + red[2] + red[3];
    const float lse = mx + logf(s);
    const int y = labels ? labels[r] : -1;
    if (dZ) {
        for (int n = tid; n < N; n += 256) {
            float p = expf(z[n] - lse);
            dZ[(long)r * ldz + n] = to_bf16((p - (n == y ? 1.0f : 0.0f)) * scale);
        }
    }
    if (tid == 0) {
        if (loss) loss[r] = (y >= 0 && y < N) ? lse - z[y] : 0.0f;
        if (max_post) max_post[r] = expf(mx - lse);
        if (lse_out) lse_out[r] = lse;
    }
}

// db[n] = sum_m dY(m, n) (bf16 input, FP32 sums) in two deterministic
// passes: workgroup (x, y) sums rows [y * kColsumRows, (y + 1) * kColsumRows)
// of 64 columns (4 row lanes of 64 coalesced columns, then an LDS sum) into
// part[y][n]; the second pass adds the row chunks in order.
constexpr int kColsumRows = 128;
__global__ void __launch_bounds__(256) colsum_part_kernel(int M, int N, const __bf16* __restrict__ dY, long ld,
                                                          float* __restrict__ part) {
    __shared__ float red[4][64];
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    const int n = blockIdx.x * 64 + tx;
    const int m0 = blockIdx.y * kColsumRows, m1 = min(M, m0 + kColsumRows);
    float s = 0.0f;
    if (n < N)
        for (int m = m0 + ty; m < m1; m += 4) s += from_bf16(dY[(long)m * ld + n]);
    red[ty][tx] = s;
    __syncthreads();
    if (ty == 0 && n < N) part[(long)blockIdx.y * N + n] = (red[0][tx] + red[1][tx]) + (red[2][tx] + red[3][tx]);
}

__global__ void colsum_final_kernel(int chunks, int N, const float* __restrict__ part, float* __restrict__ db) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    float s = 0.0f;
    for (int c = 0; c < chunks; c++) s += part[(long)c * N + n];
    db[n] = s;
}

// Adam on FP32 master weights; refreshes the bf16 copy when given.
__global__ void adam_kernel(long n, float* __restrict__ p, const float* __restrict__ grad, float* __restrict__ m1,
                            float* __restrict__ m2, float lr, float b1, float b2, float eps, float c1, float c2,
                            __bf16* __restrict__ pb) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float gr = grad[i];
    float a = b1 * m1[i] + (1.0f - b1) * gr;
    float b = b2 * m2[i] + (1.0f - b2) * gr * gr;
    m1[i] = a;
    m2[i] = b;
    float v = p[i] - lr * (a * c1) / (sqrtf(b * c2) + eps);
    p[i] = v;
    if (pb) pb[i] = to_bf16(v);
}

__global__ void to_bf16_kernel(long n, const float* __restrict__ x, __bf16* __restrict__ y) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = to_bf16(x[i]);
}

// Rows of x [rows][din] (FP64 or FP32) to bf16 [rows][dpad], zero padded.
template <typename T>
__global__ void pack_input_kernel(int rows, int din, int dpad, const T* __restrict__ x, const int* __restrict__ idx,
                                  __bf16* __restrict__ y) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)rows * dpad) return;
    int r = (int)(i / dpad), c = (int)(i % dpad);
    int src = idx ? idx[r] : r;
    y[i] = c < din ? to_bf16((float)x[(long)src * din + c]) : to_bf16(0.0f);
}

// Top-H logits of each row (one workgroup per row): radix selection of the
// H-th largest key over the order-preserving unsigned image of the floats,
// then every index above it and the lowest-indexed ties at it, in
// increasing index order.


__global__ void __launch_bounds__(256) topk_kernel(int N, int H, const float* __restrict__ Z, long ldz,
                                                    int* __restrict__ out, const float* __restrict__ lse,
                                                    float* __restrict__ mass) {
    __shared__ unsigned int hist[256];
    __shared__ unsigned int sel[2];
    const int r = blockIdx.x, tid = threadIdx.x;
    const float* z = Z + (long)r * ldz;
    if (N <= H) {  // every index, then padding
        for (int o = tid; o < H; o += 256) out[(long)r * H + o] = o < N ? o : -1;
        if (mass && tid == 0) mass[r] = 1.0f;
        return;
    }
    unsigned int prefix = 0, mask = 0;
    int need = H;  // how many of the remaining top keys are still to place
    for (int shift = 24; shift >= 0; shift -= 8) {
        hist[tid] = 0;
        __syncthreads();
        for (int n = tid; n < N; n += 256) {
            unsigned int k = fkey(z[n]);
            if ((k & mask) == prefix) atomicAdd(&hist[(k >> shift) & 255u], 1u);
        }
        __syncthreads();
        if (tid == 0) {
            int acc = 0, b = 255;
            for (; b > 0; b--) {
                if (acc + (int)hist[b] >= need) break;
                acc += (int)hist[b];
            }
            sel[0] = (unsigned int)b;
            sel[1] = (unsigned int)(need - acc);
        }
        __syncthreads();
        prefix |= sel[0] << shift;
        mask |= 255u << shift;
        need = (int)sel[1];
        __syncthreads();
    }
    // prefix is the H-th largest key; the first `need` keys equal to it (in
    // index order) are taken with every larger key. Each thread owns a
    // contiguous slice of the row; slice counts are scanned so the output
    // stays in index order.
    __shared__ int scan_hi[256], scan_eq[256];
    const int per = (N + 255) / 256;
    const int b = tid * per, e = min(N, b + per);
    int hi = 0, eq = 0;
    for (int n = b; n < e; n++) {
        unsigned int k = fkey(z[n]);
        hi += k > prefix;
        eq += k == prefix;
    }
    scan_hi[tid] = hi;
    scan_eq[tid] = eq;
    __syncthreads();
    if (tid == 0) {
        int ah = 0, ae = 0;
        for (int t = 0; t < 256; t++) {
            int h = scan_hi[t], q = scan_eq[t];
            scan_hi[t] = ah;
            scan_eq[t] = ae;
            ah += h;
            ae += q;
        }
    }
    __syncthreads();
    int eq_before = scan_eq[tid];
    int took_eq_before = min(eq_before, need);
    int o = scan_hi[tid] + took_eq_before;
    int* dst = out + (long)r * H;
    int eq_seen = eq_before;
    const float l = lse ? lse[r] : 0.0f;
    float ms = 0.0f;
    for (int n = b; n < e && o < H; n++) {
        unsigned int k = fkey(z[n]);
        if (k > prefix || (k == prefix && eq_seen < need)) {
            dst[o++] = n;
            ms += expf(z[n] - l);
        }
        if (k == prefix) eq_seen++;
    }
    if (mass) {  // softmax mass of the shortlist
        __shared__ float red[4];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) ms += __shfl_xor(ms, off, 64);
        if ((tid & 63) == 0) red[tid >> 6] = ms;
        __syncthreads();
        if (tid == 0) mass[r] = red[0] + red[1] + red[2] + red[3];
    }
}

// Top-H for rows of at most 256 * PER logits, held in registers (element
// n = tid + 256 j in key[j]): the H-th largest key K is found two bits at a
// time, MSB first, as the largest K with #(key >= K) >= H; each round counts
// the keys at or above the three candidate extensions of K in registers and
// reduces them over the wave and the block (no atomics, one barrier per
// round). Output positions in increasing index order come from per-(j, wave)
// ballot counts of keys > K and == K, scanned in (j, wave) order, which is
// index order.
template <int PER, int NT>
__device__ void topk_reg_row(int r, int N, int H, const unsigned int (&key)[PER], int* __restrict__ out, float l,
                             float* __restrict__ mass) {
    constexpr int NW = NT / 64;
    __shared__ int red[2][3][NW];
    __shared__ int cnt[2][PER][NW];
    __shared__ int tot[2];
    __shared__ float fred[NW];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    unsigned int K = 0;
#pragma unroll 1
    for (int round = 0; round < 16; round++) {
        const int sh = 30 - 2 * round;  // bits sh+1, sh of K decided this round
        int c1 = 0, c2 = 0, c3 = 0;  // wave counts from ballots (scalar, no lane exchange)
#pragma unroll
        for (int j = 0; j < PER; j++) {
            const unsigned int d = key[j] >= K ? (key[j] - K) >> sh : 0u;  // the key's next two bits, above K
            c1 += __popcll(__ballot(d >= 1u));
            c2 += __popcll(__ballot(d >= 2u));
            c3 += __popcll(__ballot(d >= 3u));
        }
        const int buf = round & 1;
        if (lane == 0) {
            red[buf][0][w] = c1;
            red[buf][1][w] = c2;
            red[buf][2][w] = c3;
        }
        __syncthreads();  // double-buffered: one barrier per round
        int t1 = 0, t2 = 0, t3 = 0;
#pragma unroll
        for (int v = 0; v < NW; v++) {
            t1 += red[buf][0][v];
            t2 += red[buf][1][v];
            t3 += red[buf][2][v];
        }
        const unsigned int add = t3 >= H ? 3u : t2 >= H ? 2u : t1 >= H ? 1u : 0u;
        K |= add << sh;
    }
#pragma unroll
    for (int j = 0; j < PER; j++) {
        const int n = tid + NT * j;
        const unsigned long long bg = __ballot(key[j] > K), be = __ballot(key[j] == K && n < N);
        if (lane == 0) {
            cnt[0][j][w] = __popcll(bg);
            cnt[1][j][w] = __popcll(be);
        }
        __builtin_amdgcn_sched_barrier(0);  // keeps the ballot masks of one j live at a time
    }
    __syncthreads();
    if (w < 2) {  // wave 0 scans the > K counts, wave 1 the == K counts
        int* a = &cnt[w][0][0];
        constexpr int M = NW * PER, PL = (M + 63) / 64;
        const int b = lane * PL, e = min(M, b + PL);
        int sum = 0;
        for (int i = b; i < e; i++) sum += a[i];
        int incl = sum;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const int t = __shfl_up(incl, off, 64);
            if (lane >= off) incl += t;
        }
        int run = incl - sum;
        for (int i = b; i < e; i++) {
            const int v = a[i];
            a[i] = run;
            run += v;
        }
        if (lane == 63) tot[w] = incl;
    }
    __syncthreads();
    const int need = H - tot[0];  // ties at K taken, lowest index first
    const unsigned long long below = (1ull << lane) - 1ull;
    int* dst = out + (long)r * H;
    float ms = 0.0f;
#pragma unroll
    for (int j = 0; j < PER; j++) {
        const int n = tid + NT * j;
        const bool g = key[j] > K, e = key[j] == K && n < N;
        const unsigned long long bg = __ballot(g), be = __ballot(e);
        const int gpos = cnt[0][j][w] + __popcll(bg & below);
        const int epos = cnt[1][j][w] + __popcll(be & below);
        const int o = gpos + min(epos, need);
        if ((g || (e && epos < need)) && o < H) {
            dst[o] = n;
            const unsigned int k = key[j];
            ms += expf(__uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k) - l);
        }
        __builtin_amdgcn_sched_barrier(0);
    }
    if (mass) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) ms += __shfl_xor(ms, off, 64);
        if (lane == 0) fred[w] = ms;
        __syncthreads();
        if (tid == 0) {
            float m = 0.0f;
#pragma unroll
            for (int v = 0; v < NW; v++) m += fred[v];
            mass[r] = m;
        }
    }
}

template <int PER, int NT>
__global__ void __launch_bounds__(NT) topk_reg_kernel(int N, int H, const float* __restrict__ Z, long ldz,
                                                       int* __restrict__ out, const float* __restrict__ lse,
                                                       float* __restrict__ mass) {
    const float* z = Z + (long)blockIdx.x * ldz;
    unsigned int key[PER];
#pragma unroll
    for (int j = 0; j < PER; j++) {
        const int n = threadIdx.x + NT * j;
        key[j] = n < N ? fkey(z[n]) : 0u;  // padding: below every finite logit
    }
    topk_reg_row<PER, NT>(blockIdx.x, N, H, key, out, lse ? lse[blockIdx.x] : 0.0f, mass);
}

// Wt[n][k] = bf16(W[k][n]) for W [rows][cols] FP32.
__global__ void transpose_bf16_kernel(int rows, int cols, const float* __restrict__ W, __bf16* __restrict__ Wt) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)rows * cols) return;
    int n = (int)(i / rows), k = (int)(i % rows);
    Wt[i] = to_bf16(W[(long)k * cols + n]);
}

__global__ void gather_int_kernel(int n, const int* __restrict__ src, const int* __restrict__ idx,
                                  int* __restrict__ dst) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[idx[i]];
}

// *out = mean of x[0..n) (one workgroup).
__global__ void __launch_bounds__(256) mean_kernel(int n, const float* __restrict__ x, float* __restrict__ out) {
    __shared__ float red[4];
    float s = 0.0f;
    for (int i = threadIdx.x; i < n; i += 256) s += x[i];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) *out = (red[0] + red[1] + red[2] + red[3]) / (float)n;
}

// dst[i][0..w) = src[idx[i]][0..w)
template <typename T>
__global__ void gather_rows_kernel(int rows, int w, const T* __restrict__ src, const int* __restrict__ idx,
                                   T* __restrict__ dst) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)rows * w) return;
    int r = (int)(i / w), c = (int)(i % w);
    dst[i] = src[(long)idx[r] * w + c];
}

template <typename T>
__global__ void scatter_kernel(int n, const T* __restrict__ src, const int* __restrict__ idx, T* __restrict__ dst) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[idx[i]] = src[i];
}

}  // namespace

hipError_t scatter(int n, const int* src, const int* idx, int* dst, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(scatter_kernel<int>, dim3((n + 255) / 256), dim3(256), 0, s, n, src, idx, dst);
    return hipGetLastError();
}

hipError_t scatter(int n, const double* src, const int* idx, double* dst, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(scatter_kernel<double>, dim3((n + 255) / 256), dim3(256), 0, s, n, src, idx, dst);
    return hipGetLastError();
}

hipError_t gather_rows(int rows, int w, const double* src, const int* idx, double* dst, hipStream_t s) {
    long n = (long)rows * w;
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(gather_rows_kernel<double>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, rows, w, src,
                       idx, dst);
    return hipGetLastError();
}

hipError_t gather_rows(int rows, int w, const int* src, const int* idx, int* dst, hipStream_t s) {
    long n = (long)rows * w;
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(gather_rows_kernel<int>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, rows, w, src, idx,
                       dst);
    return hipGetLastError();
}

hipError_t gather_int(int n, const int* src, const int* idx, int* dst, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(gather_int_kernel, dim3((n + 255) / 256), dim3(256), 0, s, n, src, idx, dst);
    return hipGetLastError();
}

hipError_t mean(int n, const float* x, float* out, hipStream_t s) {
    hipLaunchKernelGGL(mean_kernel, dim3(1), dim3(256), 0, s, n, x, out);
    return hipGetLastError();
}

hipError_t gemm(const GemmArgs& g, hipStream_t s) {
    if (g.M <= 0 || g.N <= 0) return hipSuccess;
    dim3 grid((g.N + BN - 1) / BN, (g.M + BM - 1) / BM);
    const bool kc = g.sak == 1 && g.sam % 8 == 0 && g.sbk == 1 && g.sbn % 8 == 0;
    if (g.gkey && !(kc && g.lsp)) return hipErrorInvalidValue;  // statistics only on the forward path
    if (g.gkey)
        hipLaunchKernelGGL(gemm_kc_stats_kernel, grid, dim3(256), 0, s, g);
    else if (kc)
        hipLaunchKernelGGL(gemm_kc_kernel, grid, dim3(256), 0, s, g);
    else
        hipLaunchKernelGGL(gemm_kernel, grid, dim3(256), 0, s, g);
    return hipGetLastError();
}

hipError_t softmax_ce(int rows, int N, const float* Z, long ldz, const int* labels, float scale, __bf16* dZ,
                      float* loss, float* max_post, float* lse, hipStream_t s) {
    if (rows <= 0) return hipSuccess;
    hipLaunchKernelGGL(softmax_ce_kernel, dim3(rows), dim3(256), 0, s, N, Z, ldz, labels, scale, dZ, loss, max_post,
                       lse);
    return hipGetLastError();
}

size_t colsum_work_floats(int M, int N) { return (size_t)((M + kColsumRows - 1) / kColsumRows) * N; }

hipError_t colsum(int M, int N, const __bf16* dY, long ld, float* db, float* work, hipStream_t s) {
    if (N <= 0) return hipSuccess;
    const int chunks = (M + kColsumRows - 1) / kColsumRows;
    if (chunks <= 0) {
        hipLaunchKernelGGL(colsum_final_kernel, dim3((N + 255) / 256), dim3(256), 0, s, 0, N, work, db);
        return hipGetLastError();
    }
    hipLaunchKernelGGL(colsum_part_kernel, dim3((N + 63) / 64, chunks), dim3(256), 0, s, M, N, dY, ld, work);
    hipLaunchKernelGGL(colsum_final_kernel, dim3((N + 255) / 256), dim3(256), 0, s, chunks, N, work, db);
    return hipGetLastError();
}

hipError_t adam(long n, float* p, const float* g, float* m1, float* m2, float lr, float b1, float b2, float eps,
                float c1, float c2, __bf16* pb, hipStream_t s) {
    hipLaunchKernelGGL(adam_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, p, g, m1, m2, lr, b1, b2,
                       eps, c1, c2, pb);
    return hipGetLastError();
}

hipError_t to_bf16(long n, const float* x, __bf16* y, hipStream_t s) {
    hipLaunchKernelGGL(to_bf16_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, x, y);
    return hipGetLastError();
}

hipError_t pack_input(int rows, int din, int dpad, const float* x, const int* idx, __bf16* y, hipStream_t s) {
    long n = (long)rows * dpad;
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(pack_input_kernel<float>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, rows, din, dpad,
                       x, idx, y);
    return hipGetLastError();
}

hipError_t pack_input(int rows, int din, int dpad, const double* x, const int* idx, __bf16* y, hipStream_t s) {
    long n = (long)rows * dpad;
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(pack_input_kernel<double>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, rows, din,
                       dpad, x, idx, y);
    return hipGetLastError();
}

namespace {

constexpr int kWaveRows = 4;     // rows per workgroup, one wavefront each
constexpr int kWaveGroup = 16;   // logits per group maximum
constexpr int kWaveCap = 1024;   // candidates a wavefront keeps in LDS

__device__ inline void wsync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// H-th largest of the keys the wavefront's lanes produce through get(i) for
// i < n (0 = absent), two bits per round from ballot counts: no barrier.
template <typename Get>
__device__ inline unsigned int wave_kth(int n, int H, Get get) {
    unsigned int K = 0;
    for (int round = 0; round < 16; round++) {
        const int sh = 30 - 2 * round;
        int c1 = 0, c2 = 0, c3 = 0;
        for (int b = 0; b < n; b += 64) {
            const int i = b + (int)(threadIdx.x & 63);
            const unsigned int k = i < n ? get(i) : 0u;
            const unsigned int d = k >= K ? (k - K) >> sh : 0u;
            c1 += __popcll(__ballot(d >= 1u));
            c2 += __popcll(__ballot(d >= 2u));
            c3 += __popcll(__ballot(d >= 3u));
        }
        K |= (c3 >= H ? 3u : c2 >= H ? 2u : c1 >= H ? 1u : 0u) << sh;
    }
    return K;
}

// Online log-sum-exp state of one lane: the running maximum m and the sum of
// exp(v - m) over the values seen.
__device__ inline void lse_add(float v, float& m, float& sacc) {
    if (v > m) {
        sacc = sacc * expf(m - v) + 1.0f;
        m = v;
    } else {
        sacc += expf(v - m);
    }
}

// Top-H with one wavefront per row. Pass 1 streams the row once, storing the
// maxima of groups of 16 consecutive logits in LDS and the row's
// log-sum-exp (online, per lane, then over the wave). tau = the H-th largest
// group maximum bounds the H-th largest logit from below (the H group
// maxima >= tau are H logits), so only the groups whose maximum is >= tau
// (at least H of them, usually few more) can hold a top-H logit: pass 2
// reads just those groups and compacts their logits >= tau in index order
// into LDS. The exact H-th largest key K is selected among them (thresholds
// <= tau count every candidate, >= H, as over the whole row), or, when more
// than kWaveCap logits are >= tau, by the same rounds over the whole row.
// Output: the logits > K and the first ties at K, in index order. vec: the
// rows are 16-byte aligned (pass 1 loads four logits per lane).
__global__ void __launch_bounds__(64 * kWaveRows) topk_wave_kernel(int rows, int N, int H, const float* __restrict__ Z,
                                                                   long ldz, int vec, int* __restrict__ out,
                                                                   const float* __restrict__ lse_in,
                                                                   float* __restrict__ mass, float* __restrict__ lse_out,
                                                                   float* __restrict__ post_out) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int r = blockIdx.x * kWaveRows + wv;
    if (r >= rows) return;  // whole wavefront
    const int G = (N + kWaveGroup - 1) / kWaveGroup;
    const size_t per_wave = (size_t)G * 4 + (size_t)kWaveCap * 8;
    unsigned int* gm = (unsigned int*)(smem + wv * per_wave);
    unsigned int* ckey = gm + G;
    int* cidx = (int*)(ckey + kWaveCap);
    const float* z = Z + (long)r * ldz;
    const unsigned long long below = (1ull << lane) - 1ull;

    // pass 1: group maxima, and the log-sum-exp
    float m = -INFINITY, sacc = 0.0f;
    if (vec) {
        // lane l holds logits n..n+3 of a span of 256: a group is 4 lanes
        constexpr int kU = 8;
        for (int b0 = 0; b0 < N; b0 += 256 * kU) {
            float4 vv[kU];
#pragma unroll
            for (int u = 0; u < kU; u++) {
                const int n = b0 + 256 * u + 4 * lane;
                if (n + 3 < N) {
                    vv[u] = *(const float4*)(z + n);
                } else {
                    vv[u].x = n < N ? z[n] : -INFINITY;
                    vv[u].y = n + 1 < N ? z[n + 1] : -INFINITY;
                    vv[u].z = n + 2 < N ? z[n + 2] : -INFINITY;
                    vv[u].w = -INFINITY;
                }
            }
#pragma unroll
            for (int u = 0; u < kU; u++) {
                const int n = b0 + 256 * u + 4 * lane;
                const float v[4] = {vv[u].x, vv[u].y, vv[u].z, vv[u].w};
                unsigned int k = 0u;
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    if (n + j < N) {
                        k = max(k, fkey(v[j]));
                        lse_add(v[j], m, sacc);
                    }
                }
                k = max(k, (unsigned int)__shfl_xor((int)k, 1, 64));
                k = max(k, (unsigned int)__shfl_xor((int)k, 2, 64));
                if ((lane & 3) == 0 && n < N) gm[n / kWaveGroup] = k;
            }
        }
    } else {
        constexpr int kUnroll = 16;  // loads in flight per lane
        for (int b0 = 0; b0 < N; b0 += 64 * kUnroll) {
            float vv[kUnroll];
#pragma unroll
            for (int u = 0; u < kUnroll; u++) {
                const int n = b0 + 64 * u + lane;
                vv[u] = n < N ? z[n] : -INFINITY;
            }
#pragma unroll
            for (int u = 0; u < kUnroll; u++) {
                const int n = b0 + 64 * u + lane;
                unsigned int k = n < N ? fkey(vv[u]) : 0u;
#pragma unroll
                for (int off = 1; off < kWaveGroup; off <<= 1) k = max(k, (unsigned int)__shfl_xor((int)k, off, 64));
                if ((lane & (kWaveGroup - 1)) == 0 && n < N) gm[n / kWaveGroup] = k;
                if (n < N) lse_add(vv[u], m, sacc);
            }
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        const float m2 = __shfl_xor(m, off, 64), s2 = __shfl_xor(sacc, off, 64);
        const float mm = fmaxf(m, m2);
        sacc = (m == -INFINITY ? 0.0f : sacc * expf(m - mm)) + (m2 == -INFINITY ? 0.0f : s2 * expf(m2 - mm));
        m = mm;
    }
    float l = lse_in ? lse_in[r] : 0.0f;
    if (lse_out) {
        l = m + logf(sacc);
        if (lane == 0) {
            lse_out[r] = l;
            if (post_out) post_out[r] = expf(m - l);
        }
    }
    wsync();
    const unsigned int tau = G >= H ? wave_kth(G, H, [&](int i) { return gm[i]; }) : 0u;

    // the groups whose maximum is >= tau, in index order, compacted in place
    // (every lane reads its entry before any lane writes, and writes land at
    // or below the entries read)
    int NG = 0;
    for (int b = 0; b < G; b += 64) {
        const int i = b + lane;
        const bool f = i < G && gm[i] >= tau;
        const unsigned long long bal = __ballot(f);
        if (f) gm[NG + __popcll(bal & below)] = (unsigned int)i;
        NG += __popcll(bal);
    }
    wsync();

    // pass 2: the logits >= tau of those groups, four groups per load
    // (16 lanes each), in index order
    int M = 0;
    constexpr int kU2 = 16;
    for (int b = 0; b < NG; b += 4 * kU2) {
        float vv[kU2];
        int nn[kU2];
#pragma unroll
        for (int u = 0; u < kU2; u++) {
            const int sl = b + 4 * u + (lane >> 4);
            nn[u] = sl < NG ? (int)gm[sl] * kWaveGroup + (lane & (kWaveGroup - 1)) : N;
        }
#pragma unroll
        for (int u = 0; u < kU2; u++) vv[u] = nn[u] < N ? z[nn[u]] : 0.0f;
#pragma unroll
        for (int u = 0; u < kU2; u++) {
            const int n = nn[u];
            const unsigned int k = n < N ? fkey(vv[u]) : 0u;
            const bool f = n < N && k >= tau;
            const unsigned long long bal = __ballot(f);
            const int pos = M + __popcll(bal & below);
            if (f && pos < kWaveCap) {
                ckey[pos] = k;
                cidx[pos] = n;
            }
            M += __popcll(bal);
        }
    }
    wsync();
    const bool listed = M <= kWaveCap;
    unsigned int K;
    if (listed)
        K = wave_kth(M, H, [&](int i) { return ckey[i]; });
    else
        K = wave_kth(N, H, [&](int i) { return fkey(z[i]); });

    // output in index order: keys > K, then ties at K, lowest index first
    int ngt = 0;
    const int nsrc = listed ? M : N;
    for (int b = 0; b < nsrc; b += 64) {
        const int i = b + lane;
        const unsigned int k = i < nsrc ? (listed ? ckey[i] : fkey(z[i])) : 0u;
        ngt += __popcll(__ballot(i < nsrc && k > K));
    }
    const int need = H - ngt;
    int* dst = out + (long)r * H;
    int o = 0, eqs = 0;
    float ms = 0.0f;
    for (int b = 0; b < nsrc && o < H; b += 64) {
        const int i = b + lane;
        const bool v = i < nsrc;
        const unsigned int k = v ? (listed ? ckey[i] : fkey(z[i])) : 0u;
        const bool g = v && k > K, e = v && k == K;
        const unsigned long long be = __ballot(e);
        const bool sel = g || (e && eqs + __popcll(be & below) < need);
        const unsigned long long bs = __ballot(sel);
        const int pos = o + __popcll(bs & below);
        if (sel && pos < H) {
            dst[pos] = listed ? cidx[i] : i;
            ms += expf(key_value(k) - l);
        }
        o += __popcll(bs);
        eqs += __popcll(be);
    }
    if (mass) {
        for (int off = 32; off > 0; off >>= 1) ms += __shfl_xor(ms, off, 64);
        if (lane == 0) mass[r] = ms;
    }
}

constexpr int kStatWords = 8;    // 16-byte words of group keys per lane: 2048 groups, 32768 logits
constexpr int kStatCand = 1024;  // largest candidate list (LDS), 16 keys per lane
// Entries of a row's list of groups at or above tau' (LDS). tau' leaves at
// most H + 32 groups (cap >= 2H) unless group keys tie at it; a row with more
// is searched over all its logits instead.
__host__ __device__ inline int stat_group_list(int G, int cap) { return G < cap + 64 ? G : cap + 64; }

// Radix selection over keys held in registers, NR per lane (absent keys 0,
// never counted): n keys count, n >= H >= 1, all in [lo, hi]. Two bits a round,
// from the highest bit where lo and hi differ (the bits above it are common to
// every key), each round by three thresholds, K + d 2^sh for d = 1, 2, 3,
// counted with ballots. Returns K with count(keys >= K) = cnt >= H, stopping
// as soon as cnt <= H + slack; with slack 0 either cnt == H or K is the exact
// H-th largest key -- either way the keys > K and the first ties at K are the
// top H.
template <int NR>
__device__ inline unsigned int radix_select(const unsigned int (&k)[NR], int nr, int n, int H, int slack,
                                            unsigned int lo, unsigned int hi, int& cnt) {
    cnt = n;
    if (lo == hi) return hi;
    const int hb = 31 - __builtin_clz(lo ^ hi);
    int sh = hb & ~1;
    unsigned int K = hi & ~((4u << sh) - 1u);  // 4u << 30 wraps to 0: every bit cleared
    for (; sh >= 0 && cnt > H + slack; sh -= 2) {
        const unsigned int t1 = K + (1u << sh), t2 = K + (2u << sh), t3 = K + (3u << sh);  // no wrap: K's low bits are 0
        int c1 = 0, c2 = 0, c3 = 0;
#pragma unroll
        for (int t = 0; t < NR; t++) {
            if (t < nr) {  // wave-uniform
                c1 += __popcll(__ballot(k[t] >= t1));
                c2 += __popcll(__ballot(k[t] >= t2));
                c3 += __popcll(__ballot(k[t] >= t3));
            }
        }
        const unsigned int d = c3 >= H ? 3u : c2 >= H ? 2u : c1 >= H ? 1u : 0u;
        cnt = d == 3u ? c3 : d == 2u ? c2 : d == 1u ? c1 : cnt;
        K += d << sh;
    }
    return K;
}

__device__ inline unsigned int wave_max_u32(unsigned int v) {
    for (int off = 32; off > 0; off >>= 1) v = max(v, (unsigned int)__shfl_xor((int)v, off, 64));
    return v;
}
__device__ inline unsigned int wave_min_u32(unsigned int v) {
    for (int off = 32; off > 0; off >>= 1) v = min(v, (unsigned int)__shfl_xor((int)v, off, 64));
    return v;
}

// Writes the H selected indices of row r (keys > K, then the first ties at K,
// in index order) from the candidates in LDS (listed) or the whole row, and
// returns the lane's share of their softmax mass.
__device__ inline float emit_topk(bool listed, int M, int N, int H, unsigned int K, const unsigned int* ckey,
                                  const int* cidx, const float* z, float l, int* dst) {
    const int lane = threadIdx.x & 63;
    const unsigned long long below = (1ull << lane) - 1ull;
    int ngt = 0;
    const int nsrc = listed ? M : N;
    for (int b = 0; b < nsrc; b += 64) {
        const int i = b + lane;
        const unsigned int k = i < nsrc ? (listed ? ckey[i] : fkey(z[i])) : 0u;
        ngt += __popcll(__ballot(i < nsrc && k > K));
    }
    const int need = H - ngt;
    int o = 0, eqs = 0;
    float ms = 0.0f;
    for (int b = 0; b < nsrc && o < H; b += 64) {
        const int i = b + lane;
        const bool v = i < nsrc;
        const unsigned int k = v ? (listed ? ckey[i] : fkey(z[i])) : 0u;
        const bool g = v && k > K, e = v && k == K;
        const unsigned long long be = __ballot(e);
        const bool sel = g || (e && eqs + __popcll(be & below) < need);
        const unsigned long long bs = __ballot(sel);
        const int pos = o + __popcll(bs & below);
        if (sel && pos < H) {
            dst[pos] = listed ? cidx[i] : i;
            ms += expf(key_value(k) - l);
        }
        o += __popcll(bs);
        eqs += __popcll(be);
    }
    return ms;
}

// Top-H of logits whose row statistics the GEMM wrote (GemmArgs gkey, lsp),
// one wavefront per row: no pass over every logit. The row's group keys are
// held in registers (lane l: groups 4 (l + 64 t) .. +3); a radix selection
// over them finds a threshold tau' <= the H-th largest group key (its top 24
// bits; at least H group maxima, hence at least H logits, are >= tau'), the
// groups whose key reaches tau' are compacted in index order, and their
// logits >= tau' are gathered (16 lanes a group) into the candidate list, in
// index order. The exact H-th largest key among them is the row's (every key
// >= tau' is a candidate); when more than cap logits reach tau' the rounds run
// over the whole row. The log-sum-exp merges the 64-column slab partials.
__global__ void __launch_bounds__(64 * kWaveRows) topk_stats_kernel(int rows, int N, int H, const float* __restrict__ Z,
                                                                    long ldz, const unsigned int* __restrict__ gk,
                                                                    const float2* __restrict__ lsp, int cap,
                                                                    int* __restrict__ out, float* __restrict__ mass,
                                                                    float* __restrict__ lse_out,
                                                                    float* __restrict__ post_out) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int r = blockIdx.x * kWaveRows + wv;
    if (r >= rows) return;  // whole wavefront
    const int G = (N + kWaveGroup - 1) / kWaveGroup, G4 = (G + 3) / 4;
    const int GL = stat_group_list(G, cap);
    const size_t per_wave = (size_t)GL * 4 + (size_t)cap * 8;
    int* glist = (int*)(smem + wv * per_wave);
    unsigned int* ckey = (unsigned int*)(glist + GL);
    int* cidx = (int*)(ckey + cap);
    const float* z = Z + (long)r * ldz;
    const unsigned long long below = (1ull << lane) - 1ull;

    // group keys (the words past the row's last group hold keys 0) and slab partials
    uint4 gv[kStatWords];
    const uint4* g4 = (const uint4*)(gk + (long)r * stat_groups(N));
#pragma unroll
    for (int t = 0; t < kStatWords; t++) {
        const int w = lane + 64 * t;
        gv[t] = w < G4 ? g4[w] : uint4{0u, 0u, 0u, 0u};
    }
    float m = -INFINITY, sacc = 0.0f;
    const float2* p = lsp + (long)r * stat_slabs(N);
    for (int i = lane; i < G4; i += 64) {
        const float2 q = p[i];
        const float mm = fmaxf(m, q.x);
        sacc = (m == -INFINITY ? 0.0f : sacc * expf(m - mm)) + q.y * expf(q.x - mm);
        m = mm;
    }
    for (int off = 32; off > 0; off >>= 1) {
        const float m2 = __shfl_xor(m, off, 64), s2 = __shfl_xor(sacc, off, 64);
        const float mm = fmaxf(m, m2);
        sacc = (m == -INFINITY ? 0.0f : sacc * expf(m - mm)) + (m2 == -INFINITY ? 0.0f : s2 * expf(m2 - mm));
        m = mm;
    }
    const float l = m + logf(sacc);
    if (lane == 0) {
        lse_out[r] = l;
        post_out[r] = expf(m - l);
    }

    // tau': a radix selection over the register-held group keys that stops
    // once at most H + 32 groups reach it
    unsigned int kk[4 * kStatWords];
    unsigned int khi = 0u, klo = 0xffffffffu;
#pragma unroll
    for (int t = 0; t < kStatWords; t++) {
        const unsigned int v4[4] = {gv[t].x, gv[t].y, gv[t].z, gv[t].w};
#pragma unroll
        for (int c = 0; c < 4; c++) {
            kk[4 * t + c] = v4[c];
            const bool valid = 4 * (lane + 64 * t) + c < G;
            khi = valid ? max(khi, v4[c]) : khi;
            klo = valid ? min(klo, v4[c]) : klo;
        }
    }
    khi = wave_max_u32(khi);
    klo = wave_min_u32(klo);
    int ng_above = G;
    const unsigned int tau =
        G >= H ? radix_select(kk, 4 * ((G4 + 63) / 64), G, H, 32, klo, khi, ng_above) : 0u;

    // the groups whose key is >= tau', in index order
    int NG = 0;
#pragma unroll
    for (int t = 0; t < kStatWords; t++) {
        if (64 * t >= G4) continue;  // wave-uniform
        const int i0 = 4 * (lane + 64 * t);
        const unsigned int kk[4] = {gv[t].x, gv[t].y, gv[t].z, gv[t].w};
        bool f[4];
        int cnt = 0;
#pragma unroll
        for (int c = 0; c < 4; c++) {
            f[c] = i0 + c < G && kk[c] >= tau;
            cnt += f[c];
        }
        const unsigned long long b0 = __ballot(cnt & 1), b1 = __ballot(cnt & 2), b2 = __ballot(cnt & 4);
        int pos = NG + __popcll(b0 & below) + 2 * __popcll(b1 & below) + 4 * __popcll(b2 & below);
#pragma unroll
        for (int c = 0; c < 4; c++)
            if (f[c]) {
                if (pos < GL) glist[pos] = i0 + c;
                pos++;
            }
        NG += __popcll(b0) + 2 * __popcll(b1) + 4 * __popcll(b2);
    }
    wsync();
    // more groups than the list holds (ties at tau'): the search over the whole row
    const bool glisted = NG <= GL;
    if (!glisted) NG = 0;

    // the logits >= tau' of those groups, four groups per load (16 lanes each)
    int M = 0;
    constexpr int kU2 = 16;
    for (int b = 0; b < NG; b += 4 * kU2) {
        float vv[kU2];
        int nn[kU2];
#pragma unroll
        for (int u = 0; u < kU2; u++) {
            const int sl = b + 4 * u + (lane >> 4);
            nn[u] = sl < NG ? glist[sl] * kWaveGroup + (lane & (kWaveGroup - 1)) : N;
        }
#pragma unroll
        for (int u = 0; u < kU2; u++) vv[u] = nn[u] < N ? z[nn[u]] : 0.0f;
#pragma unroll
        for (int u = 0; u < kU2; u++) {
            const int n = nn[u];
            const unsigned int k = n < N ? fkey(vv[u]) : 0u;
            const bool f = n < N && k >= tau;
            const unsigned long long bal = __ballot(f);
            const int pos = M + __popcll(bal & below);
            if (f && pos < cap) {
                ckey[pos] = k;
                cidx[pos] = n;
            }
            M += __popcll(bal);
        }
    }
    wsync();
    const bool listed = glisted && M <= cap;
    unsigned int K = 0u;
    if (listed) {
        // the exact selection among the candidates (all in [tau', khi]), held in registers
        unsigned int cv[kStatCand / 64];
#pragma unroll
        for (int t = 0; t < kStatCand / 64; t++) {
            const int i = lane + 64 * t;
            cv[t] = i < M ? ckey[i] : 0u;
        }
        int c;
        K = radix_select(cv, (M + 63) / 64, M, H, 0, tau, khi, c);
    } else {
        K = wave_kth(N, H, [&](int i) { return fkey(z[i]); });
    }
    float ms = emit_topk(listed, M, N, H, K, ckey, cidx, z, l, out + (long)r * H);
    for (int off = 32; off > 0; off >>= 1) ms += __shfl_xor(ms, off, 64);
    if (lane == 0) mass[r] = ms;
}

}  // namespace

// One wavefront per row (rows of up to 32768 logits, H <= 1024).
static hipError_t topk_wave(int rows, int N, int H, const float* Z, long ldz, int* out, const float* lse_in,
                            float* mass, float* lse_out, float* post_out, hipStream_t s) {
    const int G = (N + kWaveGroup - 1) / kWaveGroup;
    const size_t lds = (size_t)kWaveRows * ((size_t)G * 4 + (size_t)kWaveCap * 8);
    if (lds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute((const void*)topk_wave_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)lds);
        if (e != hipSuccess) return e;
    }
    const int vec = ldz % 4 == 0 && reinterpret_cast<uintptr_t>(Z) % 16 == 0;
    hipLaunchKernelGGL(topk_wave_kernel, dim3((rows + kWaveRows - 1) / kWaveRows), dim3(64 * kWaveRows), lds, s, rows,
                       N, H, Z, ldz, vec, out, lse_in, mass, lse_out, post_out);
    return hipGetLastError();
}

bool stats_topk_supported(int N, int H) {
    return N > 8192 && N <= 64 * kStatWords * 4 * kWaveGroup && N > H && H >= 1 && H <= kStatCand / 2;
}

hipError_t topk_softmax_stats(int rows, int N, int H, const float* Z, long ldz, const unsigned int* gkey,
                              const float2* lsp, int* out, float* lse, float* max_post, float* mass, hipStream_t s) {
    if (!stats_topk_supported(N, H)) return hipErrorNotSupported;
    if (rows <= 0) return hipSuccess;
    // candidate lists of 2H (at least 512) logits
    const int cap = std::min(kStatCand, std::max(512, 2 * H));
    const int G = (N + kWaveGroup - 1) / kWaveGroup;
    const size_t lds = (size_t)kWaveRows * ((size_t)stat_group_list(G, cap) * 4 + (size_t)cap * 8);
    if (lds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute((const void*)topk_stats_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(topk_stats_kernel, dim3((rows + kWaveRows - 1) / kWaveRows), dim3(64 * kWaveRows), lds, s, rows,
                       N, H, Z, ldz, gkey, lsp, cap, out, mass, lse, max_post);
    return hipGetLastError();
}

hipError_t topk(int rows, int N, int H, const float* Z, long ldz, int* out, const float* lse, float* mass,
                hipStream_t s) {
    if (rows <= 0) return hipSuccess;
    if (N > 8192 && N <= 32768 && N > H && H <= 1024)  // one wavefront per row (larger rows: topk_kernel)
        return topk_wave(rows, N, H, Z, ldz, out, lse, mass, nullptr, nullptr, s);
    if (N > H) {  // register-resident rows
        // at most 32 keys per thread, 256 threads per row (up to 8192 logits)
#define WP_TOPK_REG(P, NT)                                                                                      \
    if (N <= (NT) * (P)) {                                                                                      \
        hipLaunchKernelGGL((topk_reg_kernel<P, NT>), dim3(rows), dim3(NT), 0, s, N, H, Z, ldz, out, lse, mass);  \
        return hipGetLastError();                                                                               \
    }
        WP_TOPK_REG(4, 256)
        WP_TOPK_REG(8, 256)
        WP_TOPK_REG(16, 256)
        WP_TOPK_REG(32, 256)
#undef WP_TOPK_REG
    }
    hipLaunchKernelGGL(topk_kernel, dim3(rows), dim3(256), 0, s, N, H, Z, ldz, out, lse, mass);
    return hipGetLastError();
}

hipError_t topk_softmax(int rows, int N, int H, const float* Z, long ldz, int* out, float* lse, float* max_post,
                        float* mass, hipStream_t s) {
    if (rows <= 0) return hipSuccess;
    if (N > 8192 && N <= 32768 && N > H && H <= 1024)  // one wavefront per row, softmax statistics included
        return topk_wave(rows, N, H, Z, ldz, out, nullptr, mass, lse, max_post, s);
    hipError_t e = softmax_ce(rows, N, Z, ldz, nullptr, 1.0f, nullptr, nullptr, max_post, lse, s);
    if (e != hipSuccess) return e;
    return topk(rows, N, H, Z, ldz, out, lse, mass, s);
}

hipError_t transpose_bf16(int rows, int cols, const float* W, __bf16* Wt, hipStream_t s) {
    long n = (long)rows * cols;
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(transpose_bf16_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, rows, cols, W, Wt);
    return hipGetLastError();
}

}  // namespace wp
