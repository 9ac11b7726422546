// C ABI of the learned shortlist: a ReLU MLP with BF16 weights and
// activations, FP32 accumulation and FP32 master weights (Adam), trained on
// the GPU with the kernels of mlp_kernels.hip, and the shortlist classifier
// that rescores its top-H candidates with the exact HMM (no reference
// counterpart; the reference's candidate stage is the kNN of
// classifiers/hybrid-classifier.cc). Declarations are in
// include/whatprot_amd.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <string>
#include <unordered_map>
#include <vector>

#include "mlp_kernels.h"
#include "model_tables.h"
#include "wp_internal.h"

struct wp_mlp {
    int in_dim = 0, in_pad = 0, hidden = 0, nh = 0, out_dim = 0, L = 0;
    std::vector<int> din, dout;  // per weight layer (din[0] = in_pad)
    std::vector<float*> W, b, gW, gb, mW, vW, mb, vb;
    std::vector<__bf16*> Wb, Wt;  // bf16 copies [rows][cols] and transposed [cols][rows]
    long step = 0;
    // batch work buffers
    int cap = 0;
    __bf16* X = nullptr;
    std::vector<__bf16*> H, dH;
    float* Z = nullptr;
    __bf16* dZ = nullptr;
    int* y = nullptr;
    float* loss = nullptr;
    float* post = nullptr;
    float* lse = nullptr;
    float* mass = nullptr;
    float* csum = nullptr;  // row-chunk partial sums of the bias gradients
    unsigned int* gkey = nullptr;  // the output layer's row statistics for the top-k
    float2* lsp = nullptr;
    // shortlist classification workspace (reads x H candidates), kept across calls
    int sl_cap = 0, sl_H = 0, sl_off_n = -1, sl_off_h = 0;
    double* sl_rads = nullptr;  // the host entry point's upload
    int* sl_cand = nullptr;
    int* sl_off = nullptr;
    float* sl_stat = nullptr;  // [3][cap]: largest posterior, mass, log-sum-exp
    int* sl_id = nullptr;
    double* sl_adj = nullptr;
    int* sl_sel = nullptr;
    double* sl_rads2 = nullptr;
    int* sl_cand2 = nullptr;
};

namespace {

// Logits per forward chunk of the shortlist calls (rows = kChunkBytes / 4 out_dim).
constexpr long kChunkBytes = 2048l << 20;

void free_mlp(wp_mlp* m) {
    for (auto* v : {&m->W, &m->b, &m->gW, &m->gb, &m->mW, &m->vW, &m->mb, &m->vb})
        for (float* p : *v) hipFree(p);
    for (__bf16* p : m->Wb) hipFree(p);
    for (__bf16* p : m->Wt) hipFree(p);
    for (__bf16* p : m->H) hipFree(p);
    for (__bf16* p : m->dH) hipFree(p);
    hipFree(m->X);
    hipFree(m->Z);
    hipFree(m->dZ);
    hipFree(m->y);
    hipFree(m->loss);
    hipFree(m->post);
    hipFree(m->lse);
    hipFree(m->mass);
    hipFree(m->csum);
    hipFree(m->gkey);
    hipFree(m->lsp);
    for (void* p : {(void*)m->sl_rads, (void*)m->sl_cand, (void*)m->sl_off, (void*)m->sl_stat, (void*)m->sl_id,
                    (void*)m->sl_adj, (void*)m->sl_sel, (void*)m->sl_rads2, (void*)m->sl_cand2})
        hipFree(p);
    delete m;
}

template <typename T>
int dalloc(T** p, size_t n) {
    WP_HIP(hipMalloc((void**)p, sizeof(T) * std::max<size_t>(n, 1)));
    WP_HIP(hipMemset(*p, 0, sizeof(T) * std::max<size_t>(n, 1)));
    return WP_OK;
}

int ensure_batch(wp_mlp* m, int B) {
    if (B <= m->cap) return WP_OK;
    hipFree(m->X);
    for (__bf16* p : m->H) hipFree(p);
    for (__bf16* p : m->dH) hipFree(p);
    hipFree(m->Z);
    hipFree(m->dZ);
    hipFree(m->y);
    hipFree(m->loss);
    hipFree(m->post);
    hipFree(m->lse);
    hipFree(m->mass);
    hipFree(m->csum);
    hipFree(m->gkey);
    hipFree(m->lsp);
    m->H.assign(m->nh, nullptr);
    m->dH.assign(m->nh, nullptr);
    m->cap = 0;
    int rc = dalloc(&m->X, (size_t)B * m->in_pad);
    for (int l = 0; l < m->nh && !rc; l++) rc = dalloc(&m->H[l], (size_t)B * m->hidden);
    for (int l = 0; l < m->nh && !rc; l++) rc = dalloc(&m->dH[l], (size_t)B * m->hidden);
    if (!rc) rc = dalloc(&m->Z, (size_t)B * m->out_dim);
    if (!rc) rc = dalloc(&m->dZ, (size_t)B * m->out_dim);
    if (!rc) rc = dalloc(&m->y, (size_t)B);
    if (!rc) rc = dalloc(&m->loss, (size_t)B);
    if (!rc) rc = dalloc(&m->post, (size_t)B);
    if (!rc) rc = dalloc(&m->lse, (size_t)B);
    if (!rc) rc = dalloc(&m->mass, (size_t)B);
    if (!rc) rc = dalloc(&m->csum, wp::colsum_work_floats(B, std::max(m->hidden, m->out_dim)));
    if (!rc) rc = dalloc(&m->gkey, (size_t)B * wp::stat_groups(m->out_dim));
    if (!rc) rc = dalloc(&m->lsp, (size_t)B * wp::stat_slabs(m->out_dim));
    if (!rc) m->cap = B;
    return rc;
}

// Z = MLP(X) for the first B rows of m->X; stats: the output layer also
// writes the row statistics of Z (m->gkey, m->lsp) for topk_softmax_stats.
int forward(wp_mlp* m, int B, bool stats = false) {
    const __bf16* in = m->X;
    for (int l = 0; l < m->L; l++) {
        wp::GemmArgs g;
        std::memset(&g, 0, sizeof(g));
        g.M = B;
        g.N = m->dout[l];
        g.K = m->din[l];
        g.A = in;
        g.sam = m->din[l];
        g.sak = 1;
        g.B = m->Wt[l];  // B(k, n) = Wt[n][k]: k-contiguous
        g.sbk = 1;
        g.sbn = m->din[l];
        g.bias = m->b[l];
        g.ldc = m->dout[l];
        if (l < m->nh) {
            g.Cb = m->H[l];
            g.relu = 1;
        } else {
            g.Cf = m->Z;
            if (stats) {
                g.gkey = m->gkey;
                g.lsp = m->lsp;
            }
        }
        WP_HIP(wp::gemm(g, nullptr));
        if (l < m->nh) in = m->H[l];
    }
    return WP_OK;
}

// Gradients of the mean cross-entropy over the first B rows (labels in
// m->y) into gW, gb; per-row losses in m->loss.
int backward(wp_mlp* m, int B) {
    WP_HIP(wp::softmax_ce(B, m->out_dim, m->Z, m->out_dim, m->y, 1.0f / (float)B, m->dZ, m->loss, m->post, nullptr,
                          nullptr));
    const __bf16* dY = m->dZ;
    for (int l = m->L - 1; l >= 0; l--) {
        const __bf16* Ain = l == 0 ? m->X : m->H[l - 1];
        // gW[l] = Ain^T dY
        wp::GemmArgs g;
        std::memset(&g, 0, sizeof(g));
        g.M = m->din[l];
        g.N = m->dout[l];
        g.K = B;
        g.A = Ain;
        g.sam = 1;
        g.sak = m->din[l];
        g.B = dY;
        g.sbk = m->dout[l];
        g.sbn = 1;
        g.Cf = m->gW[l];
        g.ldc = m->dout[l];
        WP_HIP(wp::gemm(g, nullptr));
        WP_HIP(wp::colsum(B, m->dout[l], dY, m->dout[l], m->gb[l], m->csum, nullptr));
        if (l == 0) break;
        // dH[l-1] = (dY W[l]^T) masked by ReLU'(H[l-1])
        std::memset(&g, 0, sizeof(g));
        g.M = B;
        g.N = m->din[l];
        g.K = m->dout[l];
        g.A = dY;
        g.sam = m->dout[l];
        g.sak = 1;
        g.B = m->Wb[l];
        g.sbk = 1;
        g.sbn = m->dout[l];
        g.Cb = m->dH[l - 1];
        g.ldc = m->din[l];
        g.mask = m->H[l - 1];
        g.ldm = m->din[l];
        WP_HIP(wp::gemm(g, nullptr));
        dY = m->dH[l - 1];
    }
    return WP_OK;
}

int adam_step(wp_mlp* m, float lr) {
    m->step++;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float c1 = 1.0f / (1.0f - std::pow(b1, (float)m->step));
    const float c2 = 1.0f / (1.0f - std::pow(b2, (float)m->step));
    for (int l = 0; l < m->L; l++) {
        long nw = (long)m->din[l] * m->dout[l];
        WP_HIP(wp::adam(nw, m->W[l], m->gW[l], m->mW[l], m->vW[l], lr, b1, b2, eps, c1, c2, m->Wb[l], nullptr));
        WP_HIP(wp::transpose_bf16(m->din[l], m->dout[l], m->W[l], m->Wt[l], nullptr));
        WP_HIP(wp::adam(m->dout[l], m->b[l], m->gb[l], m->mb[l], m->vb[l], lr, b1, b2, eps, c1, c2, nullptr,
                        nullptr));
    }
    return WP_OK;
}

uint64_t splitmix(uint64_t* s) {
    uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

double unit(uint64_t* s) { return (double)(splitmix(s) >> 11) * (1.0 / 9007199254740992.0); }

int upload_host_batch(wp_mlp* m, int B, const float* x) {
    float* dx = nullptr;
    WP_HIP(hipMalloc(&dx, sizeof(float) * (size_t)std::max(B, 1) * m->in_dim));
    hipError_t e = hipMemcpy(dx, x, sizeof(float) * (size_t)B * m->in_dim, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = wp::pack_input(B, m->in_dim, m->in_pad, dx, nullptr, m->X, nullptr);
    hipFree(dx);
    WP_HIP(e);
    return WP_OK;
}

bool bad_layer(const wp_mlp* m, int l) { return !m || l < 0 || l >= m->L; }

}  // namespace

extern "C" {

int wp_mlp_create(int in_dim, int hidden, int num_hidden, int out_dim, uint64_t seed, wp_mlp** out) {
    if (!out || in_dim < 1 || hidden < 1 || num_hidden < 1 || out_dim < 1) {
        wp::set_error("bad MLP shape");
        return WP_ERR_ARG;
    }
    *out = nullptr;
    int rc = wp::require_device();
    if (rc) return rc;
    wp_mlp* m = new wp_mlp();
    m->in_dim = in_dim;
    m->in_pad = (in_dim + 31) / 32 * 32;
    m->hidden = hidden;
    m->nh = num_hidden;
    m->out_dim = out_dim;
    m->L = num_hidden + 1;
    for (int l = 0; l < m->L; l++) {
        m->din.push_back(l == 0 ? m->in_pad : hidden);
        m->dout.push_back(l == m->L - 1 ? out_dim : hidden);
    }
    uint64_t st = seed;
    for (int l = 0; l < m->L && !rc; l++) {
        const int di = m->din[l], dout = m->dout[l], real_in = l == 0 ? in_dim : di;
        // He-uniform for the ReLU layers, LeCun-uniform for the logits
        const double lim = std::sqrt((l == m->L - 1 ? 3.0 : 6.0) / real_in);
        std::vector<float> w((size_t)di * dout, 0.0f);
        for (int i = 0; i < real_in; i++)
            for (int o = 0; o < dout; o++) w[(size_t)i * dout + o] = (float)((2.0 * unit(&st) - 1.0) * lim);
        float *W = nullptr, *bb = nullptr, *gW = nullptr, *gb = nullptr, *mW = nullptr, *vW = nullptr, *mb = nullptr,
              *vb = nullptr;
        __bf16 *Wb = nullptr, *Wt = nullptr;
        rc = dalloc(&W, w.size());
        if (!rc) rc = dalloc(&bb, dout);
        if (!rc) rc = dalloc(&gW, w.size());
        if (!rc) rc = dalloc(&gb, dout);
        if (!rc) rc = dalloc(&mW, w.size());
        if (!rc) rc = dalloc(&vW, w.size());
        if (!rc) rc = dalloc(&mb, dout);
        if (!rc) rc = dalloc(&vb, dout);
        if (!rc) rc = dalloc(&Wb, w.size());
        if (!rc) rc = dalloc(&Wt, w.size());
        m->W.push_back(W);
        m->b.push_back(bb);
        m->gW.push_back(gW);
        m->gb.push_back(gb);
        m->mW.push_back(mW);
        m->vW.push_back(vW);
        m->mb.push_back(mb);
        m->vb.push_back(vb);
        m->Wb.push_back(Wb);
        m->Wt.push_back(Wt);
        if (!rc && hipMemcpy(W, w.data(), sizeof(float) * w.size(), hipMemcpyHostToDevice) != hipSuccess) {
            wp::set_error("cannot upload the MLP weights");
            rc = WP_ERR_HIP;
        }
        if (!rc && (wp::to_bf16((long)w.size(), W, Wb, nullptr) != hipSuccess ||
                    wp::transpose_bf16(di, dout, W, Wt, nullptr) != hipSuccess))
            rc = WP_ERR_HIP;
    }
    if (rc) {
        free_mlp(m);
        return rc;
    }
    *out = m;
    return WP_OK;
}

void wp_mlp_destroy(wp_mlp* m) {
    if (m) free_mlp(m);
}

int wp_mlp_layer_shape(const wp_mlp* m, int layer, int* rows, int* cols) {
    if (bad_layer(m, layer) || !rows || !cols) {
        wp::set_error("bad layer");
        return WP_ERR_ARG;
    }
    *rows = layer == 0 ? m->in_dim : m->din[layer];
    *cols = m->dout[layer];
    return WP_OK;
}

int wp_mlp_get_layer(wp_mlp* m, int layer, float* W, float* b) {
    if (bad_layer(m, layer) || !W || !b) {
        wp::set_error("bad layer");
        return WP_ERR_ARG;
    }
    int rows = layer == 0 ? m->in_dim : m->din[layer];
    WP_HIP(hipMemcpy(W, m->W[layer], sizeof(float) * (size_t)rows * m->dout[layer], hipMemcpyDeviceToHost));
    WP_HIP(hipMemcpy(b, m->b[layer], sizeof(float) * m->dout[layer], hipMemcpyDeviceToHost));
    return WP_OK;
}

int wp_mlp_set_layer(wp_mlp* m, int layer, const float* W, const float* b) {
    if (bad_layer(m, layer) || !W || !b) {
        wp::set_error("bad layer");
        return WP_ERR_ARG;
    }
    int rows = layer == 0 ? m->in_dim : m->din[layer];
    WP_HIP(hipMemcpy(m->W[layer], W, sizeof(float) * (size_t)rows * m->dout[layer], hipMemcpyHostToDevice));
    WP_HIP(hipMemcpy(m->b[layer], b, sizeof(float) * m->dout[layer], hipMemcpyHostToDevice));
    WP_HIP(wp::to_bf16((long)m->din[layer] * m->dout[layer], m->W[layer], m->Wb[layer], nullptr));
    WP_HIP(wp::transpose_bf16(m->din[layer], m->dout[layer], m->W[layer], m->Wt[layer], nullptr));
    return WP_OK;
}

int wp_mlp_logits(wp_mlp* m, int B, const float* x, float* out_logits) {
    if (!m || B < 0 || (B > 0 && (!x || !out_logits))) {
        wp::set_error("null argument");
        return WP_ERR_ARG;
    }
    if (B == 0) return WP_OK;
    int rc = ensure_batch(m, B);
    if (!rc) rc = upload_host_batch(m, B, x);
    if (!rc) rc = forward(m, B);
    if (rc) return rc;
    WP_HIP(hipMemcpy(out_logits, m->Z, sizeof(float) * (size_t)B * m->out_dim, hipMemcpyDeviceToHost));
    return WP_OK;
}

int wp_mlp_gradients(wp_mlp* m, int B, const float* x, const int* labels, double* out_loss, float* const* gW,
                     float* const* gb) {
    if (!m || B < 1 || !x || !labels || !gW || !gb) {
        wp::set_error("null argument");
        return WP_ERR_ARG;
    }
    int rc = ensure_batch(m, B);
    if (!rc) rc = upload_host_batch(m, B, x);
    if (rc) return rc;
    WP_HIP(hipMemcpy(m->y, labels, sizeof(int) * B, hipMemcpyHostToDevice));
    rc = forward(m, B);
    if (!rc) rc = backward(m, B);
    if (rc) return rc;
    std::vector<float> loss(B);
    WP_HIP(hipMemcpy(loss.data(), m->loss, sizeof(float) * B, hipMemcpyDeviceToHost));
    if (out_loss) *out_loss = std::accumulate(loss.begin(), loss.end(), 0.0) / B;
    for (int l = 0; l < m->L; l++) {
        int rows = l == 0 ? m->in_dim : m->din[l];
        WP_HIP(hipMemcpy(gW[l], m->gW[l], sizeof(float) * (size_t)rows * m->dout[l], hipMemcpyDeviceToHost));
        WP_HIP(hipMemcpy(gb[l], m->gb[l], sizeof(float) * m->dout[l], hipMemcpyDeviceToHost));
    }
    return WP_OK;
}

int wp_mlp_train(wp_mlp* m, int N, const float* x, const int* labels, int batch, int epochs, float lr,
                 uint64_t seed, float* out_epoch_loss) {
    if (!m || N < 1 || !x || !labels || batch < 1 || epochs < 0) {
        wp::set_error("bad training arguments");
        return WP_ERR_ARG;
    }
    for (int i = 0; i < N; i++)
        if (labels[i] < 0 || labels[i] >= m->out_dim) {
            wp::set_error("label outside [0, out_dim)");
            return WP_ERR_ARG;
        }
    const int B = std::min(batch, N);
    int rc = ensure_batch(m, B);
    if (rc) return rc;
    float* dx = nullptr;
    int *dy = nullptr, *didx = nullptr;
    float* dl = nullptr;
    const int steps = N / B;
    rc = dalloc(&dx, (size_t)N * m->in_dim);
    if (!rc) rc = dalloc(&dy, (size_t)N);
    if (!rc) rc = dalloc(&didx, (size_t)N);
    if (!rc) rc = dalloc(&dl, (size_t)std::max(steps, 1));
    if (!rc && (hipMemcpy(dx, x, sizeof(float) * (size_t)N * m->in_dim, hipMemcpyHostToDevice) != hipSuccess ||
                hipMemcpy(dy, labels, sizeof(int) * N, hipMemcpyHostToDevice) != hipSuccess)) {
        wp::set_error("cannot upload the training set");
        rc = WP_ERR_HIP;
    }
    std::vector<int> perm(N);
    std::vector<float> sl(std::max(steps, 1));
    uint64_t st = seed;
    for (int ep = 0; ep < epochs && !rc; ep++) {
        std::iota(perm.begin(), perm.end(), 0);
        for (int i = N - 1; i > 0; i--) std::swap(perm[i], perm[(int)(splitmix(&st) % (uint64_t)(i + 1))]);
        if (hipMemcpy(didx, perm.data(), sizeof(int) * N, hipMemcpyHostToDevice) != hipSuccess) {
            wp::set_error("cannot upload the epoch order");
            rc = WP_ERR_HIP;
            break;
        }
        for (int s = 0; s < steps && !rc; s++) {
            const int* idx = didx + (size_t)s * B;
            if (wp::pack_input(B, m->in_dim, m->in_pad, dx, idx, m->X, nullptr) != hipSuccess ||
                wp::gather_int(B, dy, idx, m->y, nullptr) != hipSuccess) {
                wp::set_error("batch gather failed");
                rc = WP_ERR_HIP;
                break;
            }
            rc = forward(m, B);
            if (!rc) rc = backward(m, B);
            if (!rc) rc = adam_step(m, lr);
            if (!rc && wp::mean(B, m->loss, dl + s, nullptr) != hipSuccess) rc = WP_ERR_HIP;
        }
        if (!rc && steps > 0) {
            if (hipMemcpy(sl.data(), dl, sizeof(float) * steps, hipMemcpyDeviceToHost) != hipSuccess) {
                wp::set_error("cannot read the losses");
                rc = WP_ERR_HIP;
            } else if (out_epoch_loss) {
                out_epoch_loss[ep] = (float)(std::accumulate(sl.begin(), sl.begin() + steps, 0.0) / steps);
            }
        }
    }
    hipFree(dx);
    hipFree(dy);
    hipFree(didx);
    hipFree(dl);
    return rc;
}

int wp_mlp_shortlist(wp_mlp* m, int N, const double* radiometries, int H, int* out_index, float* out_max_posterior,
                     float* out_mass) {
    if (!m || N < 0 || H < 1 || H > m->out_dim || (N > 0 && (!radiometries || !out_index))) {
        wp::set_error("bad shortlist arguments");
        return WP_ERR_ARG;
    }
    const int chunk = std::max(1, std::min(N, (int)std::max<long>(1, kChunkBytes / (4l * m->out_dim))));
    int rc = ensure_batch(m, chunk);
    if (rc) return rc;
    double* dr = nullptr;
    int* didx = nullptr;
    rc = dalloc(&dr, (size_t)chunk * m->in_dim);
    if (!rc) rc = dalloc(&didx, (size_t)chunk * H);
    for (int r0 = 0; r0 < N && !rc; r0 += chunk) {
        const int B = std::min(chunk, N - r0);
        if (hipMemcpy(dr, radiometries + (size_t)r0 * m->in_dim, sizeof(double) * (size_t)B * m->in_dim,
                      hipMemcpyHostToDevice) != hipSuccess ||
            wp::pack_input(B, m->in_dim, m->in_pad, dr, nullptr, m->X, nullptr) != hipSuccess) {
            wp::set_error("cannot stage the radiometries");
            rc = WP_ERR_HIP;
            break;
        }
        const bool stats = wp::stats_topk_supported(m->out_dim, H);
        rc = forward(m, B, stats);
        if (rc) break;
        if ((stats ? wp::topk_softmax_stats(B, m->out_dim, H, m->Z, m->out_dim, m->gkey, m->lsp, didx, m->lse,
                                            m->post, m->mass, nullptr)
                   : wp::topk_softmax(B, m->out_dim, H, m->Z, m->out_dim, didx, m->lse, m->post, m->mass,
                                      nullptr)) != hipSuccess ||
            hipMemcpy(out_index + (size_t)r0 * H, didx, sizeof(int) * (size_t)B * H, hipMemcpyDeviceToHost) !=
                hipSuccess ||
            (out_max_posterior &&
             hipMemcpy(out_max_posterior + r0, m->post, sizeof(float) * B, hipMemcpyDeviceToHost) != hipSuccess) ||
            (out_mass && hipMemcpy(out_mass + r0, m->mass, sizeof(float) * B, hipMemcpyDeviceToHost) != hipSuccess)) {
            wp::set_error("shortlist kernels failed");
            rc = WP_ERR_HIP;
        }
    }
    hipFree(dr);
    hipFree(didx);
    return rc;
}

}  // extern "C"

namespace {

// Grows the shortlist workspace to n reads of H candidates.
int ensure_shortlist(wp_mlp* m, int n, int H) {
    if (n <= m->sl_cap && H <= m->sl_H) return WP_OK;
    const int cap = std::max(n, m->sl_cap), h = std::max(H, m->sl_H);
    for (void* p : {(void*)m->sl_rads, (void*)m->sl_cand, (void*)m->sl_off, (void*)m->sl_stat, (void*)m->sl_id,
                    (void*)m->sl_adj, (void*)m->sl_sel, (void*)m->sl_rads2, (void*)m->sl_cand2})
        hipFree(p);
    m->sl_rads = m->sl_rads2 = nullptr;
    m->sl_cand = m->sl_off = m->sl_id = m->sl_sel = m->sl_cand2 = nullptr;
    m->sl_stat = nullptr;
    m->sl_adj = nullptr;
    m->sl_cap = m->sl_H = 0;
    m->sl_off_n = -1;
    const size_t d = (size_t)m->in_dim;
    int rc = dalloc(&m->sl_rads, (size_t)cap * d);
    if (!rc) rc = dalloc(&m->sl_cand, (size_t)cap * h);
    if (!rc) rc = dalloc(&m->sl_off, (size_t)cap + 1);
    if (!rc) rc = dalloc(&m->sl_stat, (size_t)3 * cap);
    if (!rc) rc = dalloc(&m->sl_id, (size_t)cap);
    if (!rc) rc = dalloc(&m->sl_adj, (size_t)cap);
    if (!rc) rc = dalloc(&m->sl_sel, (size_t)cap);
    if (!rc) rc = dalloc(&m->sl_rads2, (size_t)cap * d);
    if (!rc) rc = dalloc(&m->sl_cand2, (size_t)cap * h);
    if (rc) return rc;
    m->sl_cap = cap;
    m->sl_H = h;
    return WP_OK;
}

// The shortlist call on device radiometries (N x in_dim) into device outputs;
// h_rads: the same radiometries on the host, or null (the fallback reads are
// then copied down). Runs on the null stream and returns when the outputs
// are written.
int shortlist_run(wp_hmm_classifier* clf, wp_mlp* m, int N, const double* d_rads, const double* h_rads, int H,
                  double min_posterior, int criterion, int* d_out_id, double* d_out_adj, int* out_num_fallback) {
    const int S = wp_hmm_classifier_num_dye_seqs(clf), d = m->in_dim;
    H = std::min(H, S);
    const int chunk = std::max(1, std::min(N, (int)std::max<long>(1, kChunkBytes / (4l * m->out_dim))));
    int rc = ensure_batch(m, chunk);
    if (!rc) rc = ensure_shortlist(m, N, H);
    if (rc) return rc;
    auto hip_ok = [&](hipError_t e, const char* what) {
        if (e != hipSuccess && !rc) {
            wp::set_error(std::string(what) + ": " + hipGetErrorString(e));
            rc = WP_ERR_HIP;
        }
        return rc == WP_OK;
    };
    if (m->sl_off_n != N || m->sl_off_h != H) {  // candidate offsets r * H, uploaded once per shape
        std::vector<int> offs(N + 1);
        for (int r = 0; r <= N; r++) offs[r] = r * H;
        if (!hip_ok(hipMemcpy(m->sl_off, offs.data(), sizeof(int) * (N + 1), hipMemcpyHostToDevice), "upload"))
            return rc;
        m->sl_off_n = N;
        m->sl_off_h = H;
    }
    float* d_post = m->sl_stat;
    float* d_mass = m->sl_stat + m->sl_cap;
    float* d_lse = m->sl_stat + 2 * (size_t)m->sl_cap;
    // the chunks' statistics go straight to per-read arrays: no copy, and no
    // host synchronisation, between chunks
    for (int r0 = 0; r0 < N && !rc; r0 += chunk) {
        const int B = std::min(chunk, N - r0);
        if (!hip_ok(wp::pack_input(B, d, m->in_pad, d_rads + (size_t)r0 * d, nullptr, m->X, nullptr), "pack"))
            break;
        const bool stats = wp::stats_topk_supported(m->out_dim, H);
        rc = forward(m, B, stats);
        if (rc) break;
        int* cand = m->sl_cand + (size_t)r0 * H;
        hip_ok(stats ? wp::topk_softmax_stats(B, m->out_dim, H, m->Z, m->out_dim, m->gkey, m->lsp, cand, d_lse + r0,
                                              d_post + r0, d_mass + r0, nullptr)
                     : wp::topk_softmax(B, m->out_dim, H, m->Z, m->out_dim, cand, d_lse + r0, d_post + r0,
                                        d_mass + r0, nullptr),
               "top-k");
    }
    const float* d_conf = criterion == WP_SHORTLIST_MASS ? d_mass : d_post;
    std::vector<float> conf(N);
    if (!rc) hip_ok(hipMemcpy(conf.data(), d_conf, sizeof(float) * N, hipMemcpyDeviceToHost), "download");
    std::vector<int> sure, unsure;
    for (int r = 0; r < N && !rc; r++) (conf[r] >= min_posterior ? sure : unsure).push_back(r);
    if (!rc && !sure.empty()) {  // HMM rescoring of the shortlists, on the device
        const int ns = (int)sure.size();
        if (ns == N) {
            rc = wp_hmm_classify_candidates_device(clf, N, d_rads, m->sl_off, m->sl_cand, d_out_id, d_out_adj, nullptr,
                                                   nullptr, nullptr);
        } else {
            if (hip_ok(hipMemcpy(m->sl_sel, sure.data(), sizeof(int) * ns, hipMemcpyHostToDevice), "upload") &&
                hip_ok(wp::gather_rows(ns, d, d_rads, m->sl_sel, m->sl_rads2, nullptr), "gather") &&
                hip_ok(wp::gather_rows(ns, H, m->sl_cand, m->sl_sel, m->sl_cand2, nullptr), "gather"))
                rc = wp_hmm_classify_candidates_device(clf, ns, m->sl_rads2, m->sl_off, m->sl_cand2, m->sl_id,
                                                       m->sl_adj, nullptr, nullptr, nullptr);
            if (!rc && hip_ok(wp::scatter(ns, m->sl_id, m->sl_sel, d_out_id, nullptr), "scatter"))
                hip_ok(wp::scatter(ns, m->sl_adj, m->sl_sel, d_out_adj, nullptr), "scatter");
        }
    }
    if (rc) return rc;
    if (out_num_fallback) *out_num_fallback = (int)unsure.size();
    if (unsure.empty()) {
        hip_ok(hipDeviceSynchronize(), "shortlist");
        return rc;
    }
    // the fallback reads' radiometries on the host
    std::vector<double> hx;
    const double* radiometries = h_rads;
    if (!radiometries) {
        hx.assign((size_t)N * d, 0.0);
        for (int r : unsure)
            if (!hip_ok(hipMemcpy(&hx[(size_t)r * d], d_rads + (size_t)r * d, sizeof(double) * d,
                                  hipMemcpyDeviceToHost),
                        "download"))
                return rc;
        radiometries = hx.data();
    }
    std::vector<int> fid(unsure.size());
    std::vector<double> fadj(unsure.size());
    int* out_id = fid.data();  // per fallback read
    double* out_adjusted = fadj.data();
    if (!unsure.empty()) {
        // Full HMM fallback. A few reads against every dye-seq would each be
        // one workgroup working through S candidates while the rest of the
        // GPU idles, so each read's dye-seqs are scored in chunks as separate
        // candidate lists and the chunks are merged in order: the largest
        // score with the earlier candidate winning ties (a later chunk must be
        // strictly larger), and the totals by log-sum-exp over the chunks
        // that hold a non-empty pair (hmm-classifier.h:45-78 over all S).
        const int chunk = 512, nch = (S + chunk - 1) / chunk;
        const size_t nv = unsure.size() * (size_t)nch;
        std::vector<double> x(nv * (size_t)d);
        std::vector<int> off(nv + 1), idx((size_t)S * unsure.size());
        size_t o = 0;
        for (size_t i = 0; i < unsure.size(); i++)
            for (int ch = 0; ch < nch; ch++) {
                const size_t v = i * nch + ch;
                std::memcpy(&x[v * d], radiometries + (size_t)unsure[i] * d, sizeof(double) * d);
                off[v] = (int)o;
                for (int k = ch * chunk; k < std::min(S, (ch + 1) * chunk); k++) idx[o++] = k;
            }
        off[nv] = (int)o;
        std::vector<int> id(nv);
        std::vector<double> adj(nv), ls(nv), lt(nv);
        rc = wp_hmm_classify_candidates(clf, (int)nv, x.data(), off.data(), idx.data(), id.data(), adj.data(),
                                        ls.data(), lt.data());
        if (rc) return rc;
        // The read's pick is the best of its chunks' picks. They are decided
        // by a second call over the chunk winners in chunk order (one read per
        // fallback read), so the kernels' own rule picks among them: strict
        // '>' with the earlier candidate winning, near-ties in FP64 on the
        // FP32 path. The winners' dye-seq indices come from their ids, which
        // needs unique ids; otherwise the chunks' ln scores decide directly.
        std::unordered_map<int, int> index_of;
        const bool by_index = wp::dye_seq_index_of_ids(clf, &index_of);
        std::vector<int> off2(unsure.size() + 1, 0), idx2;
        std::vector<double> tots(unsure.size(), -INFINITY);
        std::vector<int> best_ids(unsure.size());
        std::vector<double> bests(unsure.size());
        for (size_t i = 0; i < unsure.size(); i++) {
            const size_t v0 = i * nch;
            int best_id = id[v0];
            double best = ls[v0], tot = -INFINITY;
            for (int ch = 0; ch < nch; ch++) {
                const size_t v = v0 + ch;
                if (ls[v] > best) {
                    best = ls[v];
                    best_id = id[v];
                }
                if (ls[v] > -INFINITY) {  // an all-empty chunk holds no mass
                    const double hi = std::max(tot, lt[v]);
                    tot = hi + std::log(std::exp(tot - hi) + std::exp(lt[v] - hi));
                    if (by_index) idx2.push_back(index_of.at(id[v]));
                }
            }
            off2[i + 1] = (int)idx2.size();
            tots[i] = tot;
            best_ids[i] = best_id;
            bests[i] = best;
        }
        std::vector<int> id2(unsure.size());
        std::vector<double> adj2(unsure.size()), ls2(unsure.size()), lt2(unsure.size());
        if (by_index && !idx2.empty()) {
            std::vector<double> x2(unsure.size() * (size_t)d);
            for (size_t i = 0; i < unsure.size(); i++)
                std::memcpy(&x2[i * d], radiometries + (size_t)unsure[i] * d, sizeof(double) * d);
            rc = wp_hmm_classify_candidates(clf, (int)unsure.size(), x2.data(), off2.data(), idx2.data(), id2.data(),
                                            adj2.data(), ls2.data(), lt2.data());
            if (rc) return rc;
        }
        for (size_t i = 0; i < unsure.size(); i++) {
            int best_id = best_ids[i];
            double best = bests[i];
            if (by_index && off2[i + 1] > off2[i]) {
                best_id = id2[i];
                best = ls2[i];
            }
            const double tot = tots[i];
            double a = tot > -INFINITY ? std::exp(best - tot) : 0.0;
            if (std::isnan(a)) a = 0.0;
            out_id[i] = best_id;
            out_adjusted[i] = a;
        }
    }
    // into the outputs at the fallback reads' places
    const int nu = (int)unsure.size();
    if (hip_ok(hipMemcpy(m->sl_sel, unsure.data(), sizeof(int) * nu, hipMemcpyHostToDevice), "upload") &&
        hip_ok(hipMemcpy(m->sl_id, fid.data(), sizeof(int) * nu, hipMemcpyHostToDevice), "upload") &&
        hip_ok(hipMemcpy(m->sl_adj, fadj.data(), sizeof(double) * nu, hipMemcpyHostToDevice), "upload") &&
        hip_ok(wp::scatter(nu, m->sl_id, m->sl_sel, d_out_id, nullptr), "scatter") &&
        hip_ok(wp::scatter(nu, m->sl_adj, m->sl_sel, d_out_adj, nullptr), "scatter"))
        hip_ok(hipDeviceSynchronize(), "shortlist");
    return rc;
}

}  // namespace

extern "C" {

int wp_shortlist_classify(wp_hmm_classifier* clf, wp_mlp* m, int num_reads, const double* radiometries, int H,
                          double min_posterior, int criterion, int* out_id, double* out_adjusted,
                          int* out_num_fallback) {
    if (!clf || !m || num_reads < 0 || H < 1 || (num_reads > 0 && (!radiometries || !out_id || !out_adjusted))) {
        wp::set_error("null argument");
        return WP_ERR_ARG;
    }
    if (criterion != WP_SHORTLIST_MAX_POSTERIOR && criterion != WP_SHORTLIST_MASS) {
        wp::set_error("unknown shortlist criterion");
        return WP_ERR_ARG;
    }
    if (m->out_dim != wp_hmm_classifier_num_dye_seqs(clf)) {
        wp::set_error("the MLP's outputs must be the classifier's dye-seqs");
        return WP_ERR_ARG;
    }
    if (out_num_fallback) *out_num_fallback = 0;
    const int N = num_reads;
    if (N == 0) return WP_OK;
    int rc = ensure_shortlist(m, N, std::min(H, m->out_dim));
    if (rc) return rc;
    const size_t d = (size_t)m->in_dim;
    int* d_id = nullptr;
    double* d_adj = nullptr;
    hipError_t e = hipMemcpy(m->sl_rads, radiometries, sizeof(double) * (size_t)N * d, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMalloc(&d_id, sizeof(int) * (size_t)N);
    if (e == hipSuccess) e = hipMalloc(&d_adj, sizeof(double) * (size_t)N);
    if (e == hipSuccess) {
        rc = shortlist_run(clf, m, N, m->sl_rads, radiometries, H, min_posterior, criterion, d_id, d_adj,
                           out_num_fallback);
        if (!rc) e = hipMemcpy(out_id, d_id, sizeof(int) * (size_t)N, hipMemcpyDeviceToHost);
        if (!rc && e == hipSuccess)
            e = hipMemcpy(out_adjusted, d_adj, sizeof(double) * (size_t)N, hipMemcpyDeviceToHost);
    }
    hipFree(d_id);
    hipFree(d_adj);
    if (!rc && e != hipSuccess) {
        wp::set_error(std::string("shortlist: ") + hipGetErrorString(e));
        rc = WP_ERR_HIP;
    }
    return rc;
}

int wp_shortlist_classify_device(wp_hmm_classifier* clf, wp_mlp* m, int num_reads, const double* d_radiometries,
                                 int H, double min_posterior, int criterion, int* d_out_id, double* d_out_adjusted,
                                 int* out_num_fallback, void* stream) {
    if (!clf || !m || num_reads < 0 || H < 1 ||
        (num_reads > 0 && (!d_radiometries || !d_out_id || !d_out_adjusted))) {
        wp::set_error("null argument");
        return WP_ERR_ARG;
    }
    if (criterion != WP_SHORTLIST_MAX_POSTERIOR && criterion != WP_SHORTLIST_MASS) {
        wp::set_error("unknown shortlist criterion");
        return WP_ERR_ARG;
    }
    if (m->out_dim != wp_hmm_classifier_num_dye_seqs(clf)) {
        wp::set_error("the MLP's outputs must be the classifier's dye-seqs");
        return WP_ERR_ARG;
    }
    if (out_num_fallback) *out_num_fallback = 0;
    if (num_reads == 0) return WP_OK;
    const hipError_t e = hipStreamSynchronize((hipStream_t)stream);  // the caller's inputs are ready
    if (e != hipSuccess) {
        wp::set_error(std::string("shortlist: ") + hipGetErrorString(e));
        return WP_ERR_HIP;
    }
    return shortlist_run(clf, m, num_reads, d_radiometries, nullptr, H, min_posterior, criterion, d_out_id,
                         d_out_adjusted, out_num_fallback);
}

}  // extern "C"
