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
#include <memory>
#include <memory_resource>
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
    double* d_boxes = nullptr;
    double* d_points = nullptr;
    int* d_hits = nullptr;
    int* d_perm = nullptr;
    // per-call buffers
    int cap = 0;
    double* d_q = nullptr;
    int* d_cnt = nullptr;
    int* d_idx = nullptr;
    double* d_dist = nullptr;
    // pinned host copies of the results (the hybrid's pipelined fetches)
    int* h_cnt = nullptr;
    int* h_idx = nullptr;
    double* h_dist = nullptr;
    int* d_next = nullptr;  // the searches' query counters, one per stream slot
    hipStream_t st[2] = {nullptr, nullptr};  // the hybrid's two search streams (created on first use)
};

struct wp_hybrid_classifier {
    wp_nn_classifier* nn = nullptr;
    wp_hmm_classifier* hmm = nullptr;
    int h = 0;
    // pinned host lists of candidates ([N][h] slots, and packed), kept
    // across calls
    size_t list_cap = 0;
    int* h_slots = nullptr;
    int* h_packed = nullptr;
    std::unordered_map<int, int> id_index;  // hybrid-classifier.cc:37-40
    std::unordered_map<int, int> id_count;
    // the same two maps as arrays over [id_lo, id_lo + flat_index.size()) when
    // the ids are dense enough (index -1: no such id)
    int id_lo = 0;
    std::vector<int> flat_index, flat_count;
    // table index and count of dye-seq id (0 and 0 for an unknown id, as the
    // reference's map lookups with defaults give)
    void lookup(int id, int* index, int* count) const {
        if (!flat_index.empty()) {
            const long o = (long)id - id_lo;
            if (o >= 0 && o < (long)flat_index.size() && flat_index[o] >= 0) {
                *index = flat_index[o];
                *count = flat_count[o];
            } else {
                *index = 0;
                *count = 0;
            }
            return;
        }
        auto it = id_count.find(id);
        *count = it == id_count.end() ? 0 : it->second;
        auto jt = id_index.find(id);
        *index = jt == id_index.end() ? 0 : jt->second;
    }
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
    hipFree(c->d_boxes);
    hipFree(c->d_points);
    hipFree(c->d_hits);
    hipFree(c->d_perm);
    hipFree(c->d_q);
    hipFree(c->d_cnt);
    hipFree(c->d_idx);
    hipFree(c->d_dist);
    hipHostFree(c->h_cnt);
    hipHostFree(c->h_idx);
    hipHostFree(c->h_dist);
    hipFree(c->d_next);
    for (hipStream_t s : c->st)
        if (s) hipStreamDestroy(s);
    delete c;
}

int ensure_cap(wp_nn_classifier* c, int N) {
    if (N <= c->cap) return WP_OK;
    hipFree(c->d_q);
    hipFree(c->d_cnt);
    hipFree(c->d_idx);
    hipFree(c->d_dist);
    hipHostFree(c->h_cnt);
    hipHostFree(c->h_idx);
    hipHostFree(c->h_dist);
    c->d_q = nullptr;
    c->d_cnt = nullptr;
    c->d_idx = nullptr;
    c->d_dist = nullptr;
    c->h_cnt = nullptr;
    c->h_idx = nullptr;
    c->h_dist = nullptr;
    c->cap = 0;
    WP_HIP(hipMalloc(&c->d_q, sizeof(double) * (size_t)N * c->d));
    WP_HIP(hipMalloc(&c->d_cnt, sizeof(int) * (size_t)N));
    WP_HIP(hipMalloc(&c->d_idx, sizeof(int) * (size_t)N * c->k));
    WP_HIP(hipMalloc(&c->d_dist, sizeof(double) * (size_t)N * c->k));
    if (!c->d_next) WP_HIP(hipMalloc(&c->d_next, 2 * sizeof(int)));
    WP_HIP(hipHostMalloc((void**)&c->h_cnt, sizeof(int) * (size_t)N, hipHostMallocDefault));
    WP_HIP(hipHostMalloc((void**)&c->h_idx, sizeof(int) * (size_t)N * c->k, hipHostMallocDefault));
    WP_HIP(hipHostMalloc((void**)&c->h_dist, sizeof(double) * (size_t)N * c->k, hipHostMallocDefault));
    c->cap = N;
    return WP_OK;
}

// KDTree::search (kd-tree.h:40-46) in pieces: upload the N queries, launch
// the search of reads [r0, r1) (asynchronous on the null stream), fetch its
// results (waits for it): per query up to k (track, dist^2), farthest first
// (k-best.h:45-55).
int search_upload(wp_nn_classifier* c, int N, const double* rads) {
    int rc = ensure_cap(c, N);
    if (rc) return rc;
    WP_HIP(hipMemcpy(c->d_q, rads, sizeof(double) * (size_t)N * c->d, hipMemcpyHostToDevice));
    return WP_OK;
}

int search_launch(wp_nn_classifier* c, int r0, int r1, hipStream_t stream = nullptr, int slot = 0) {
    wp::KnnArgs a;
    a.N = r1 - r0;
    a.d = c->d;
    a.k = c->k;
    a.depth = c->depth;
    a.queries = c->d_q + (size_t)r0 * c->d;
    a.nodes = c->d_nodes;
    a.boxes = c->d_boxes;
    a.points = c->d_points;
    a.hits = c->d_hits;
    a.perm = c->d_perm;
    a.out_count = c->d_cnt + r0;
    a.out_index = c->d_idx + (size_t)r0 * c->k;
    a.out_dist_sq = c->d_dist + (size_t)r0 * c->k;
    a.next = c->d_next + slot;
    WP_HIP(wp::launch_knn(a, stream));
    return WP_OK;
}

// Queues the copies of reads [r0, r1)'s results into the pinned buffers
// (valid once the stream has been synchronized).
int search_queue_fetch(wp_nn_classifier* c, int r0, int r1, hipStream_t stream) {
    const size_t n = (size_t)(r1 - r0), k = (size_t)c->k;
    WP_HIP(hipMemcpyAsync(c->h_cnt + r0, c->d_cnt + r0, sizeof(int) * n, hipMemcpyDeviceToHost, stream));
    WP_HIP(hipMemcpyAsync(c->h_idx + r0 * k, c->d_idx + r0 * k, sizeof(int) * n * k, hipMemcpyDeviceToHost, stream));
    WP_HIP(hipMemcpyAsync(c->h_dist + r0 * k, c->d_dist + r0 * k, sizeof(double) * n * k, hipMemcpyDeviceToHost,
                          stream));
    return WP_OK;
}

int search_fetch(wp_nn_classifier* c, int r0, int r1, int* cnt, int* idx, double* dist) {
    const size_t n = (size_t)(r1 - r0), k = (size_t)c->k;
    WP_HIP(hipMemcpy(cnt + r0, c->d_cnt + r0, sizeof(int) * n, hipMemcpyDeviceToHost));
    WP_HIP(hipMemcpy(idx + r0 * k, c->d_idx + r0 * k, sizeof(int) * n * k, hipMemcpyDeviceToHost));
    WP_HIP(hipMemcpy(dist + r0 * k, c->d_dist + r0 * k, sizeof(double) * n * k, hipMemcpyDeviceToHost));
    for (int r = r0; r < r1; r++) {
        if (cnt[r] < 0) {
            wp::set_error("kNN search: k-best heap full (a point with no hits)");
            return WP_ERR_INTERNAL;
        }
    }
    return WP_OK;
}

int search(wp_nn_classifier* c, int N, const double* rads, std::vector<int>* cnt, std::vector<int>* idx,
           std::vector<double>* dist) {
    cnt->assign(N, 0);
    idx->assign((size_t)N * c->k, -1);
    dist->assign((size_t)N * c->k, 0.0);
    if (N == 0 || c->n == 0) return WP_OK;
    int rc = search_upload(c, N, rads);
    if (!rc) rc = search_launch(c, 0, N);
    if (!rc) rc = search_fetch(c, 0, N, cnt->data(), idx->data(), dist->data());
    return rc;
}

// NNClassifier::classify_helper (nn-classifier.cc:123-151).
template <typename Map>
double vote(const wp_nn_classifier* c, int cnt, const int* idx, const double* dist, Map* scores) {
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
// The score map is the reference's std::unordered_map<int, double> (its
// iteration order decides which of equal scores the heap keeps), with its
// nodes and buckets carved from a per-thread arena rewound for every read in
// place of one heap allocation per dye-seq.
std::vector<Scored> top_of(const wp_nn_classifier* c, int cnt, const int* idx, const double* dist, unsigned int h) {
    thread_local std::vector<unsigned char> arena(8u << 20);
    std::pmr::monotonic_buffer_resource res(arena.data(), arena.size());  // beyond it: new/delete
    std::pmr::unordered_map<int, double> scores(&res);
    double total = vote(c, cnt, idx, dist, &scores);
    std::vector<Scored> heap;
    heap.reserve((size_t)h + 1);
    std::priority_queue<Scored, std::vector<Scored>, std::greater<Scored>> pq(std::greater<Scored>(), std::move(heap));
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
        std::vector<double> boxes;
        wp::kdtree_boxes(d, X.data(), perm, nodes, &boxes);
        if (hipMalloc(&c->d_nodes, sizeof(wp::KdNode) * nodes.size()) != hipSuccess ||
            hipMalloc(&c->d_boxes, sizeof(double) * boxes.size()) != hipSuccess ||
            hipMalloc(&c->d_points, sizeof(double) * P.size()) != hipSuccess ||
            hipMalloc(&c->d_hits, sizeof(int) * n) != hipSuccess || hipMalloc(&c->d_perm, sizeof(int) * n) != hipSuccess) {
            wp::set_error("cannot allocate the kd-tree on the device");
            free_nn(c);
            return WP_ERR_NOMEM;
        }
        if (hipMemcpy(c->d_nodes, nodes.data(), sizeof(wp::KdNode) * nodes.size(), hipMemcpyHostToDevice) ||
            hipMemcpy(c->d_boxes, boxes.data(), sizeof(double) * boxes.size(), hipMemcpyHostToDevice) ||
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
    if (num_dye_seqs > 0) {
        const int lo = *std::min_element(dye_seq_ids, dye_seq_ids + num_dye_seqs);
        const int hi = *std::max_element(dye_seq_ids, dye_seq_ids + num_dye_seqs);
        if ((long)hi - lo < 8l * num_dye_seqs + 1024) {
            c->id_lo = lo;
            c->flat_index.assign((size_t)hi - lo + 1, -1);
            c->flat_count.assign((size_t)hi - lo + 1, 0);
            for (const auto& kv : c->id_index) c->flat_index[kv.first - lo] = kv.second;
            for (const auto& kv : c->id_count) c->flat_count[kv.first - lo] = kv.second;
        }
    }
    *out = c;
    return WP_OK;
}

void wp_hybrid_classifier_destroy(wp_hybrid_classifier* c) {
    if (!c) return;
    hipHostFree(c->h_slots);
    hipHostFree(c->h_packed);
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
    const int h = c->h;
    // candidates per read (hybrid-classifier.cc:55-66), in the nn order:
    // their table indices in slot r of a [N][h] array, the subfraction, and
    // the last (best) candidate for reads the HMM leaves unscored
    if ((size_t)N * h > c->list_cap) {
        hipHostFree(c->h_slots);
        hipHostFree(c->h_packed);
        c->h_slots = c->h_packed = nullptr;
        c->list_cap = 0;
        WP_HIP(hipHostMalloc((void**)&c->h_slots, sizeof(int) * (size_t)N * h, hipHostMallocDefault));
        WP_HIP(hipHostMalloc((void**)&c->h_packed, sizeof(int) * (size_t)N * h, hipHostMallocDefault));
        c->list_cap = (size_t)N * h;
    }
    int* slots = c->h_slots;
    std::vector<int> ncand(N, 0), fb_id(N, -1);
    std::vector<double> subfraction(N, 0.0), fb_adj(N, 0.0);
    // the reads in pieces, searched (and their results copied to pinned
    // memory) on two streams in turn: piece p + 2 is queued as soon as piece
    // p is in, the host votes over piece p while the GPU searches the next
    // two, and a piece's slowest queries overlap the next piece's search
    const bool any = nn->n > 0;
    const int pieces = N >= 16384 ? 8 : 1, per = (N + pieces - 1) / pieces;
    int rc = any ? search_upload(nn, N, radiometries) : WP_OK;
    for (int i = 0; i < 2 && !rc && any; i++)
        if (!nn->st[i] && hipStreamCreate(&nn->st[i]) != hipSuccess) {
            wp::set_error("cannot create the search streams");
            rc = WP_ERR_HIP;
        }
    auto queue_piece = [&](int p) -> int {
        const int a0 = p * per, a1 = std::min(N, a0 + per);
        if (a0 >= a1) return WP_OK;
        int e = search_launch(nn, a0, a1, nn->st[p % 2], p % 2);
        return e ? e : search_queue_fetch(nn, a0, a1, nn->st[p % 2]);
    };
    for (int p = 0; p < 2 && !rc && any; p++) rc = queue_piece(p);
    std::vector<int> none(nn->k, -1);
    std::vector<double> zero(nn->k, 0.0);
    for (int p = 0; p < pieces && !rc; p++) {
        const int r0 = p * per, r1 = std::min(N, r0 + per);
        if (r0 >= r1) break;
        if (any) {
            if (hipStreamSynchronize(nn->st[p % 2]) != hipSuccess) {
                wp::set_error("kNN search failed");
                rc = WP_ERR_HIP;
                break;
            }
            if (p + 2 < pieces) rc = queue_piece(p + 2);
            for (int r = r0; r < r1 && !rc; r++)
                if (nn->h_cnt[r] < 0) {
                    wp::set_error("kNN search: k-best heap full (a point with no hits)");
                    rc = WP_ERR_INTERNAL;
                }
            if (rc) break;
        }
        const int k = nn->k;
        auto row_cnt = [&](int r) { return any ? nn->h_cnt[r] : 0; };
        auto row_idx = [&](int r) { return any ? nn->h_idx + (size_t)r * k : none.data(); };
        auto row_dist = [&](int r) { return any ? nn->h_dist + (size_t)r * k : zero.data(); };
        parallel_for(r1 - r0, [&](int b, int e) {
        for (int r = r0 + b; r < r0 + e; r++) {
            const std::vector<Scored> cands = top_of(nn, row_cnt(r), row_idx(r), row_dist(r), (unsigned)h);
            int* dst = &slots[(size_t)r * h];
            double s = 0.0;
            for (size_t i = 0; i < cands.size(); i++) {
                int index, count;
                c->lookup(cands[i].id, &index, &count);
                s += cands[i].adjusted() * (double)count;
                dst[i] = index;
            }
            ncand[r] = (int)cands.size();
            subfraction[r] = s;
            if (!cands.empty()) {
                fb_id[r] = cands.back().id;
                fb_adj[r] = cands.back().adjusted();
            }
        }
        });
    }
    if (rc) {
        hipDeviceSynchronize();  // a search still in flight writes into the classifier's buffers
        return rc;
    }
    std::vector<int> off(N + 1, 0);
    for (int r = 0; r < N; r++) off[r + 1] = off[r] + ncand[r];
    const int* list = slots;
    if ((size_t)off[N] != (size_t)N * h) {  // some reads hold fewer than h: pack the lists
        int* packed = c->h_packed;
        parallel_for(N, [&](int b, int e) {
            for (int r = b; r < e; r++) std::copy(&slots[(size_t)r * h], &slots[(size_t)r * h] + ncand[r], &packed[off[r]]);
        });
        list = packed;
    }
    std::vector<int> hid(N);
    std::vector<double> hadj(N);
    rc = wp_hmm_classify_candidates(c->hmm, N, radiometries, off.data(), list, hid.data(), hadj.data(), nullptr,
                                    nullptr);
    if (rc) return rc;
    for (int r = 0; r < N; r++) {  // hybrid-classifier.cc:67-82
        double adj;
        int id;
        if (hid[r] == -1) {  // no scored candidate: the best of the kNN's (-1, 0 when it has none)
            id = fb_id[r];
            adj = fb_adj[r];
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
