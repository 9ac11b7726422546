// Launch interface of the learned-shortlist MLP kernels (host side).
#pragma once

#include <hip/hip_runtime.h>

namespace wp {

// C(m, n) = epilogue(sum_k A(m, k) B(k, n)); A(m, k) = A[m * sam + k * sak],
// B(k, n) = B[k * sbk + n * sbn] (bf16). Epilogue: + bias[n], ReLU, zero where
// mask(m, n) <= 0, then FP32 (Cf) and/or bf16 (Cb) stores with row stride ldc.
// Optional row statistics of the stored values for the top-k (gkey non-null):
// gkey[m * stat_groups(N) + n / 16] = the largest order-preserving key (fkey)
// of columns 16 (n / 16) .. +15 (columns >= N excluded), and
// lsp[m * stat_slabs(N) + n / 64] = (max, sum of exp(value - max)) over
// columns 64 (n / 64) .. +63.
struct GemmArgs {
    int M, N, K;
    const __bf16* A;
    long sam, sak;
    const __bf16* B;
    long sbk, sbn;
    float* Cf;
    __bf16* Cb;
    long ldc;
    const float* bias;
    int relu;
    const __bf16* mask;
    long ldm;
    unsigned int* gkey;
    float2* lsp;
};

// Row strides of the GEMM's row statistics.
__host__ __device__ inline long stat_groups(int N) { return ((long)(N + 15) / 16 + 3) / 4 * 4; }
__host__ __device__ inline long stat_slabs(int N) { return (long)(N + 63) / 64; }

hipError_t gemm(const GemmArgs& g, hipStream_t s);
hipError_t softmax_ce(int rows, int N, const float* Z, long ldz, const int* labels, float scale, __bf16* dZ,
                      float* loss, float* max_post, float* lse, hipStream_t s);
// db[n] = sum over M rows of dY(m, n); work holds colsum_work_floats(M, N) floats.
size_t colsum_work_floats(int M, int N);
hipError_t colsum(int M, int N, const __bf16* dY, long ld, float* db, float* work, hipStream_t s);
hipError_t adam(long n, float* p, const float* g, float* m1, float* m2, float lr, float b1, float b2, float eps,
                float c1, float c2, __bf16* pb, hipStream_t s);
hipError_t to_bf16(long n, const float* x, __bf16* y, hipStream_t s);
hipError_t pack_input(int rows, int din, int dpad, const float* x, const int* idx, __bf16* y, hipStream_t s);
hipError_t pack_input(int rows, int din, int dpad, const double* x, const int* idx, __bf16* y, hipStream_t s);
hipError_t gather_rows(int rows, int w, const double* src, const int* idx, double* dst, hipStream_t s);
// dst[idx[i]] = src[i], i < n.
hipError_t scatter(int n, const int* src, const int* idx, int* dst, hipStream_t s);
hipError_t scatter(int n, const double* src, const int* idx, double* dst, hipStream_t s);
hipError_t gather_rows(int rows, int w, const int* src, const int* idx, int* dst, hipStream_t s);
hipError_t gather_int(int n, const int* src, const int* idx, int* dst, hipStream_t s);
hipError_t mean(int n, const float* x, float* out, hipStream_t s);
hipError_t topk(int rows, int N, int H, const float* Z, long ldz, int* out, const float* lse, float* mass,
                hipStream_t s);
// topk and the row statistics of softmax_ce (log-sum-exp, largest posterior)
// in one pass over the logits where the row fits the filtered top-k.
hipError_t topk_softmax(int rows, int N, int H, const float* Z, long ldz, int* out, float* lse, float* max_post,
                        float* mass, hipStream_t s);
// topk_softmax over logits Z whose row statistics the GEMM wrote (gkey, lsp):
// the same outputs without a pass over every logit. Returns
// hipErrorNotSupported for shapes the one-wavefront-per-row kernel does not take
// (stats_topk_supported).
bool stats_topk_supported(int N, int H);
hipError_t topk_softmax_stats(int rows, int N, int H, const float* Z, long ldz, const unsigned int* gkey,
                              const float2* lsp, int* out, float* lse, float* max_post, float* mass, hipStream_t s);
hipError_t transpose_bf16(int rows, int cols, const float* W, __bf16* Wt, hipStream_t s);

}  // namespace wp
