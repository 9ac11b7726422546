// C ABI of the kNN classifier (NNClassifier) and the hybrid classifier
// (HybridClassifier): kd-tree search on the GPU, kernel-weighted voting and
// the top-h candidates on the host, HMM rescoring of the candidates on the
// GPU. Declarations and reference citations are in include/whatprot_amd.h.
//
// The vote runs on the host with std::unordered_map<int, double> and the
// top-h selection with std::priority_queue, as nn-classifier.cc:123-204 does:
// ties in the best score and at the top-h boundary are broken by those
// containers' iteration and heap order, which the same standard library
// reproduces exactly.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <functional>
#include <queue>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "kdtree.h"
#include "model_tables.h"
#include "wp_internal.h"

struct wp_nn_classifier {
    int T = 0, C = 0, d = 0, k = 0, n = 0, depth = 0;
    double two_sig_sq = 0.0;
    std::vector<int> src_off, src_id, src_count, src_hits;  // per original track
    wp::KdNode* d_nodes = nullptr;
    double* d_points = nullptr;
    int* d_hits = nullptr;
    int* d_perm = nullptr;
    // per-call buffers
    int cap = 0;
    double* d_q = nullptr;
    int* d_cnt = nullptr;
    int* d_idx = nullptr;
    double* d_dist = nullptr;
};

struct wp_hybrid_classifier {
    wp_nn_classifier* nn = nullptr;
    wp_hmm_classifier* hmm = nullptr;
    int h = 0;
    std::unordered_map<int, int> id_index;  // hybrid-classifier.cc:37-40
    std::unordered_map<int, int> id_count;
};

namespace {

// ScoredClassification (common/scored-classification.h:14-25) for the
// candidate heap; operator> compares score only (scored-classification.cc:27-29).
struct Scored {
    int id;
    double score, total;
    double adjusted() const { return score / total; }
};
bool operator>(const Scored& x, const Scored& y) { return x.score > y.score; }

void free_nn(wp_nn_classifier* c) {
    hipFree(c->d_nodes);
    hipFree(c->d_points);
    hipFree(c->d_hits);
    hipFree(c->d_perm);
    hipFree(c->d_q);
    hipFree(c->d_cnt);
    hipFree(c->d_idx);
    hipFree(c->d_dist);
    delete c;
}

int ensure_cap(wp_nn_classifier* c, int N) {
    if (N <= c->cap) return WP_OK;
    hipFree(c->d_q);
    hipFree(c->d_cnt);
    hipFree(c->d_idx);
    hipFree(c->d_dist);
    c->d_q = nullptr;
    c->d_cnt = nullptr;
    c->d_idx = nullptr;
    c->d_dist = nullptr;
    c->cap = 0;
    WP_HIP(hipMalloc(&c->d_q, sizeof(double) * (size_t)N * c->d));
    WP_HIP(hipMalloc(&c->d_cnt, sizeof(int) * (size_t)N));
    WP_HIP(hipMalloc(&c->d_idx, sizeof(int) * (size_t)N * c->k));
    WP_HIP(hipMalloc(&c->d_dist, sizeof(double) * (size_t)N * c->k));
    c->cap = N;
    return WP_OK;
}

// KDTree::search for N queries: per query up to k (track, dist^2), farthest
// first (k-best.h:45-55).
int search(wp_nn_classifier* c, int N, const double* rads, std::vector<int>* cnt, std::vector<int>* idx,
           std::vector<double>* dist) {
    cnt->assign(N, 0);
    idx->assign((size_t)N * c->k, -1);
    dist->assign((size_t)N * c->k, 0.0);
    if (N == 0 || c->n == 0) return WP_OK;
    int rc = ensure_cap(c, N);
    if (rc) return rc;
    WP_HIP(hipMemcpy(c->d_q, rads, sizeof(double) * (size_t)N * c->d, hipMemcpyHostToDevice));
    wp::KnnArgs a;
    a.N = N;
    a.d = c->d;
    a.k = c->k;
    a.depth = c->depth;
    a.queries = c->d_q;
    a.nodes = c->d_nodes;
    a.points = c->d_points;
    a.hits = c->d_hits;
    a.perm = c->d_perm;
    a.out_count = c->d_cnt;
    a.out_index = c->d_idx;
    a.out_dist_sq = c->d_dist;
    WP_HIP(wp::launch_knn(a, nullptr));
    WP_HIP(hipMemcpy(cnt->data(), c->d_cnt, sizeof(int) * N, hipMemcpyDeviceToHost));
    WP_HIP(hipMemcpy(idx->data(), c->d_idx, sizeof(int) * (size_t)N * c->k, hipMemcpyDeviceToHost));
    WP_HIP(hipMemcpy(dist->data(), c->d_dist, sizeof(double) * (size_t)N * c->k, hipMemcpyDeviceToHost));
    for (int r = 0; r < N; r++) {
        if ((*cnt)[r] < 0) {
            wp::set_error("kNN search: k-best heap full (a point with no hits)");
            return WP_ERR_INTERNAL;
        }
    }
    return WP_OK;
}

// NNClassifier::classify_helper (nn-classifier.cc:123-151).
double vote(const wp_nn_classifier* c, int cnt, const int* idx, const double* dist,
            std::unordered_map<int, double>* scores) {
    double total = 0.0;
    for (int i = 0; i < cnt; i++) {
        int t = idx[i];
        double weight = exp(-dist[i] / c->two_sig_sq);
        for (int j = c->src_off[t]; j < c->src_off[t + 1]; j++) {
            double count = (double)c->src_count[j];
            double hits = (double)c->src_hits[j];
            total += weight * hits;
            (*scores)[c->src_id[j]] += weight * hits / count;
        }
    }
    return total;
}

// NNClassifier::classify(radiometry) (nn-classifier.cc:153-177).
Scored best_of(const wp_nn_classifier* c, int cnt, const int* idx, const double* dist) {
    std::unordered_map<int, double> scores;
    double total = vote(c, cnt, idx, dist, &scores);
    int best_id = -1;
    double best_score = -1.0;
    for (const auto& kv : scores) {
        double id = kv.first;
        double score = kv.second;
        if (score > best_score) {
            best_id = (int)id;
            best_score = score;
        }
    }
    Scored r{best_id, best_score, total};
    if (std::isnan(r.adjusted())) {
        r.score = 0.0;
        r.total = 1.0;
    }
    return r;
}

// NNClassifier::classify(radiometry, h) (nn-classifier.cc:179-204): the h
// best ids in ascending score order.
std::vector<Scored> top_of(const wp_nn_classifier* c, int cnt, const int* idx, const double* dist, unsigned int h) {
    std::unordered_map<int, double> scores;
    double total = vote(c, cnt, idx, dist, &scores);
    std::priority_queue<Scored, std::vector<Scored>, std::greater<Scored>> pq;
    for (const auto& kv : scores) {
        double id = kv.first;
        double score = kv.second;
        if (pq.size() < h) {
            pq.push(Scored{(int)id, score, total});
        } else if (score > pq.top().score) {
            pq.push(Scored{(int)id, score, total});
            pq.pop();
        }
    }
    std::vector<Scored> out;
    out.reserve(pq.size());
    while (!pq.empty()) {
        out.push_back(pq.top());
        pq.pop();
    }
    return out;
}

// Runs f(begin, end) over [0, n) on a few host threads.
template <typename F>
void parallel_for(int n, F f) {
    int nt = (int)std::min<unsigned>(std::max(1u, std::thread::hardware_concurrency()), 16u);
    nt = std::max(1, std::min(nt, n / 256 + 1));
    if (nt == 1) {
        f(0, n);
        return;
    }
    std::vector<std::thread> th;
    int per = (n + nt - 1) / nt;
    for (int t = 0; t < nt; t++) {
        int b = t * per, e = std::min(n, b + per);
        if (b < e) th.emplace_back(f, b, e);
    }
    for (auto& x : th) x.join();
}

bool valid_rads(int N, const double* rads) { return N >= 0 && (N == 0 || rads); }

}  // namespace

extern "C" {

int wp_kdtree_order(int k, int d, int n, const double* points, int* out_order, int* out_depth) {
    if (k < 1 || d < 1 || n < 0 || (n > 0 && (!points || !out_order))) {
        wp::set_error("bad argument");
        return WP_ERR_ARG;
    }
    std::vector<int> perm;
    std::vector<wp::KdNode> nodes;
    int depth = wp::build_kdtree(k, d, n, points, &perm, &nodes);
    if (n > 0) std::memcpy(out_order, perm.data(), sizeof(int) * n);
    if (out_depth) *out_depth = depth;
    return WP_OK;
}

int wp_nn_classifier_create(unsigned int num_timesteps, unsigned int num_channels, const wp_model* seq_model,
                            int k, double sig, int num_dye_tracks, const short* track_counts,
                            const int* source_offsets, const int* source_ids, const int* source_counts,
                            const int* source_hits, wp_nn_classifier** out) {
    if (!out || !seq_model || num_dye_tracks < 0 ||
        (num_dye_tracks > 0 && (!track_counts || !source_offsets || !source_ids || !source_counts || !source_hits))) {
        wp::set_error("null argument");
        return WP_ERR_ARG;
    }
    *out = nullptr;
    const int T = (int)num_timesteps, C = (int)num_channels;
    if (T < 1 || T > WP_MAX_TIMESTEPS || C < 1 || C > WP_MAX_CHANNELS || seq_model->num_channels != C || k < 1) {
        wp::set_error("unsupported shape: need 1 <= T <= 64, 1 <= C <= 8, model channels == C, k >= 1");
        return WP_ERR_UNSUPPORTED;
    }
    // KBest::insert (k-best.h:26-43) pops only while the remaining hits cover
    // k, so with a zero-hit point the heap could grow past k + 1 entries; the
    // search kernel's LDS heap is sized for k + 1, so such input is refused.
    for (int i = 0; i < num_dye_tracks; i++) {
        if (source_offsets[i + 1] <= source_offsets[i]) {
            wp::set_error("every dye-track needs at least one source");
            return WP_ERR_ARG;
        }
        for (int j = source_offsets[i]; j < source_offsets[i + 1]; j++) {
            if (source_hits[j] < 1) {
                wp::set_error("every dye-track source needs hits >= 1");
                return WP_ERR_ARG;
            }
        }
    }
    int rc = wp::require_device();
    if (rc) return rc;
    wp_nn_classifier* c = new wp_nn_classifier();
    c->T = T;
    c->C = C;
    c->d = T * C;
    c->k = k;
    c->n = num_dye_tracks;
    c->two_sig_sq = 2.0 * sig * sig;  // nn-classifier.cc:100
    const int n = num_dye_tracks, d = c->d;
    c->src_off.assign(source_offsets, source_offsets + n + 1);
    int ns = n ? source_offsets[n] : 0;
    c->src_id.assign(source_ids, source_ids + ns);
    c->src_count.assign(source_counts, source_counts + ns);
    c->src_hits.assign(source_hits, source_hits + ns);
    // KDTEntry coordinates: adjusted_mu of the counts at t, entry i -> (t = i / C,
    // c = i % C) (nn-classifier.cc:37-58); hits = total hits of the sources
    std::vector<double> X((size_t)n * d);
    std::vector<int> hits(n);
    for (int i = 0; i < n; i++) {
        for (int t = 0; t < T; t++) {
            unsigned int cnt[WP_MAX_CHANNELS];
            for (int ch = 0; ch < C; ch++) cnt[ch] = (unsigned int)track_counts[((size_t)i * T + t) * C + ch];
            for (int ch = 0; ch < C; ch++) X[(size_t)i * d + t * C + ch] = wp::adjusted_mu(*seq_model, ch, cnt);
        }
        int hsum = 0;
        for (int j = source_offsets[i]; j < source_offsets[i + 1]; j++) hsum += source_hits[j];
        hits[i] = hsum;
    }
    std::vector<int> perm;
    std::vector<wp::KdNode> nodes;
    c->depth = wp::build_kdtree(k, d, n, X.data(), &perm, &nodes);
    if (n > 0) {
        if (wp::knn_lds_bytes(d, k, c->depth) > 64 * 1024) {
            wp::set_error("k, dimension or tree depth too large for the search kernel's LDS");
            free_nn(c);
            return WP_ERR_UNSUPPORTED;
        }
        // points in the permuted order, each leaf's block dimension-major
        std::vector<double> P((size_t)n * d);
        std::vector<int> H(n);
        for (const wp::KdNode& nd : nodes) {
            if (nd.left >= 0) continue;
            const int len = nd.end - nd.begin;
            double* blk = &P[(size_t)nd.begin * d];
            for (int j = 0; j < len; j++)
                for (int i = 0; i < d; i++) blk[(size_t)i * len + j] = X[(size_t)perm[nd.begin + j] * d + i];
        }
        for (int p = 0; p < n; p++) H[p] = hits[perm[p]];
        if (hipMalloc(&c->d_nodes, sizeof(wp::KdNode) * nodes.size()) != hipSuccess ||
            hipMalloc(&c->d_points, sizeof(double) * P.size()) != hipSuccess ||
            hipMalloc(&c->d_hits, sizeof(int) * n) != hipSuccess || hipMalloc(&c->d_perm, sizeof(int) * n) != hipSuccess) {
            wp::set_error("cannot allocate the kd-tree on the device");
            free_nn(c);
            return WP_ERR_NOMEM;
        }
        if (hipMemcpy(c->d_nodes, nodes.data(), sizeof(wp::KdNode) * nodes.size(), hipMemcpyHostToDevice) ||
            hipMemcpy(c->d_points, P.data(), sizeof(double) * P.size(), hipMemcpyHostToDevice) ||
            hipMemcpy(c->d_hits, H.data(), sizeof(int) * n, hipMemcpyHostToDevice) ||
            hipMemcpy(c->d_perm, perm.data(), sizeof(int) * n, hipMemcpyHostToDevice)) {
            wp::set_error("cannot copy the kd-tree to the device");
            free_nn(c);
            return WP_ERR_HIP;
        }
    }
    *out = c;
    return WP_OK;
}

void wp_nn_classifier_destroy(wp_nn_classifier* c) {
    if (c) free_nn(c);
}

int wp_nn_search(wp_nn_classifier* c, int num_reads, const double* radiometries, int* out_count,
                 int* out_track, double* out_dist_sq) {
    if (!c || !valid_rads(num_reads, radiometries) || !out_count || !out_track || !out_dist_sq) {
        wp::set_error("null argument");
        return WP_ERR_ARG;
    }
    std::vector<int> cnt, idx;
    std::vector<double> dist;
    int rc = search(c, num_reads, radiometries, &cnt, &idx, &dist);
    if (rc) return rc;
    std::memcpy(out_count, cnt.data(), sizeof(int) * num_reads);
    std::memcpy(out_track, idx.data(), sizeof(int) * idx.size());
    std::memcpy(out_dist_sq, dist.data(), sizeof(double) * dist.size());
    return WP_OK;
}

int wp_nn_classify(wp_nn_classifier* c, int num_reads, const double* radiometries, int* out_id, double* out_score,
                   double* out_total) {
    if (!c || !valid_rads(num_reads, radiometries) || !out_id || !out_score || !out_total) {
        wp::set_error("null argument");
        return WP_ERR_ARG;
    }
    std::vector<int> cnt, idx;
    std::vector<double> dist;
    int rc = search(c, num_reads, radiometries, &cnt, &idx, &dist);
    if (rc) return rc;
    parallel_for(num_reads, [&](int b, int e) {
        for (int r = b; r < e; r++) {
            Scored s = best_of(c, cnt[r], &idx[(size_t)r * c->k], &dist[(size_t)r * c->k]);
            out_id[r] = s.id;
            out_score[r] = s.score;
            out_total[r] = s.total;
        }
    });
    return WP_OK;
}

int wp_nn_classify_top(wp_nn_classifier* c, int num_reads, const double* radiometries, int h, int* out_count,
                       int* out_ids, double* out_scores, double* out_total) {
    if (!c || !valid_rads(num_reads, radiometries) || h < 1 || !out_count || !out_ids || !out_scores ||
        !out_total) {
        wp::set_error("null argument");
        return WP_ERR_ARG;
    }
    std::vector<int> cnt, idx;
    std::vector<double> dist;
    int rc = search(c, num_reads, radiometries, &cnt, &idx, &dist);
    if (rc) return rc;
    parallel_for(num_reads, [&](int b, int e) {
        for (int r = b; r < e; r++) {
            std::vector<Scored> top = top_of(c, cnt[r], &idx[(size_t)r * c->k], &dist[(size_t)r * c->k], (unsigned)h);
            out_count[r] = (int)top.size();
            out_total[r] = top.empty() ? 0.0 : top[0].total;
            for (size_t i = 0; i < top.size(); i++) {
                out_ids[(size_t)r * h + i] = top[i].id;
                out_scores[(size_t)r * h + i] = top[i].score;
            }
        }
    });
    return WP_OK;
}

int wp_hybrid_classifier_create(unsigned int num_timesteps, unsigned int num_channels, const wp_model* seq_model,
                                double dist_cutoff, int k, double sig, int num_dye_tracks,
                                const short* track_counts, const int* source_offsets, const int* source_ids,
                                const int* source_counts, const int* source_hits, int h, int num_dye_seqs,
                                const char* const* dye_strings, const int* dye_seq_counts,
                                const int* dye_seq_ids, int precision, wp_hybrid_classifier** out) {
    if (!out || h < 1) {
        wp::set_error("bad argument");
        return WP_ERR_ARG;
    }
    *out = nullptr;
    wp_hybrid_classifier* c = new wp_hybrid_classifier();
    int rc = wp_hmm_classifier_create(num_timesteps, num_channels, seq_model, dist_cutoff, num_dye_seqs, dye_strings,
                                      dye_seq_counts, dye_seq_ids, precision, &c->hmm);
    if (!rc)
        rc = wp_nn_classifier_create(num_timesteps, num_channels, seq_model, k, sig, num_dye_tracks, track_counts,
                                     source_offsets, source_ids, source_counts, source_hits, &c->nn);
    if (rc) {
        wp_hmm_classifier_destroy(c->hmm);
        wp_nn_classifier_destroy(c->nn);
        delete c;
        return rc;
    }
    c->h = h;
    for (int i = 0; i < num_dye_seqs; i++) {
        c->id_index[dye_seq_ids[i]] = i;
        c->id_count[dye_seq_ids[i]] = dye_seq_counts[i];
    }
    *out = c;
    return WP_OK;
}

void wp_hybrid_classifier_destroy(wp_hybrid_classifier* c) {
    if (!c) return;
    wp_hmm_classifier_destroy(c->hmm);
    wp_nn_classifier_destroy(c->nn);
    delete c;
}

int wp_hybrid_classify(wp_hybrid_classifier* c, int num_reads, const double* radiometries, int* out_id,
                       double* out_adjusted) {
    if (!c || !valid_rads(num_reads, radiometries) || !out_id || !out_adjusted) {
        wp::set_error("null argument");
        return WP_ERR_ARG;
    }
    const int N = num_reads;
    if (N == 0) return WP_OK;  // (the per-read vectors below would be empty, their data() null)
    wp_nn_classifier* nn = c->nn;
    std::vector<int> cnt, idx;
    std::vector<double> dist;
    int rc = search(nn, N, radiometries, &cnt, &idx, &dist);
    if (rc) return rc;
    // candidates per read (hybrid-classifier.cc:55-66), in the nn order
    std::vector<std::vector<Scored>> cands(N);
    std::vector<std::vector<int>> cidx(N);  // candidates as dye-seq table indices
    std::vector<double> subfraction(N, 0.0);
    parallel_for(N, [&](int b, int e) {
        for (int r = b; r < e; r++) {
            cands[r] = top_of(nn, cnt[r], &idx[(size_t)r * nn->k], &dist[(size_t)r * nn->k], (unsigned)c->h);
            double s = 0.0;
            cidx[r].reserve(cands[r].size());
            for (const Scored& x : cands[r]) {
                auto it = c->id_count.find(x.id);
                s += x.adjusted() * (double)(it == c->id_count.end() ? 0 : it->second);
                auto jt = c->id_index.find(x.id);
                cidx[r].push_back(jt == c->id_index.end() ? 0 : jt->second);
            }
            subfraction[r] = s;
        }
    });
    std::vector<int> off(N + 1, 0);
    for (int r = 0; r < N; r++) off[r + 1] = off[r] + (int)cidx[r].size();
    std::vector<int> list(std::max(off[N], 1), 0);
    parallel_for(N, [&](int b, int e) {
        for (int r = b; r < e; r++) std::copy(cidx[r].begin(), cidx[r].end(), list.begin() + off[r]);
    });
    std::vector<int> hid(N);
    std::vector<double> hadj(N);
    rc = wp_hmm_classify_candidates(c->hmm, N, radiometries, off.data(), list.data(), hid.data(), hadj.data(),
                                    nullptr, nullptr);
    if (rc) return rc;
    for (int r = 0; r < N; r++) {  // hybrid-classifier.cc:67-82
        double adj;
        int id;
        if (hid[r] == -1) {
            if (cands[r].empty()) {
                id = -1;
                adj = 0.0;
            } else {
                id = cands[r].back().id;
                adj = cands[r].back().adjusted();
            }
        } else {
            id = hid[r];
            adj = hadj[r] * subfraction[r];
        }
        if (std::isnan(adj)) adj = 0.0;
        out_id[r] = id;
        out_adjusted[r] = adj;
    }
    return WP_OK;
}

}  // extern "C"
