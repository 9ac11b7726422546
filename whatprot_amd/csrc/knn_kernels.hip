// kd-tree k-nearest-neighbour search on the GPU (KDTree::search,
// kd-tree/kd-tree.h:40-46) with the reference's exact semantics.
//
// One wavefront per query walks the tree depth first in the reference's
// order (internal-node.h:33-56: the near side of the split first, the far
// side afterwards only if the current k-th distance still exceeds the squared
// gap, strictly). At a leaf the 64 lanes compute the squared distances of up
// to 64 points at once, summed as leaf-node.h:38-60 sums them (groups of four
// dimensions, ((x1^2 + x2^2) + x3^2) + x4^2 added to the running total, the
// rest one by one, no contraction); the partial-sum early exit there only
// skips points whose full distance is >= the k-th distance, so evaluating
// every point and then admitting, in leaf order, those strictly below the
// current k-th distance admits the same points. Admission is KBest::insert
// (k-best.h:26-43): a max-heap on (distance, position) where position is the
// point's place in the permuted array (the reference's pointer order), one
// pop when the farthest entry is not needed to keep k hits, and the k-th
// distance set once the hits reach k. The heap lives in LDS and is updated
// by lane 0.
//
// Box test: before a node's subtree is entered, the squared distance from the
// query to the bounding box of its points is summed over the lanes; when it
// reaches the k-th distance (by a relative margin of 1e-12, far above the
// rounding of that sum or of any point's), every point below would be
// rejected by the distance test above, so the subtree is left at once and
// the heap, the k-th distance and the order of the far sides visited after
// it are those of the full walk.
#include <hip/hip_runtime.h>
#include <float.h>

#include <algorithm>

#include "kdtree.h"

namespace wp {
namespace {

constexpr int kWaves = 4;
constexpr int kKnnGrid = 2048;  // workgroups of a launch (8 per CU)

__device__ inline void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

struct WaveLds {
    double* q;       // [d]
    double* hd;      // heap distances [k]
    int* hp;         // heap positions [k]
    int* sn;         // stack nodes [depth]
    double* sg;      // stack squared gaps [depth]
    double* cd;      // chunk distances [64]
    int* st;         // hsize, hits
    double* kth;
};

__host__ __device__ inline size_t align8(size_t x) { return (x + 7) & ~(size_t)7; }

// The heap holds k + 1 entries between KBest::insert's push and its pop; the
// stack holds one entry per internal node on the current path (< depth).
__host__ __device__ inline size_t wave_bytes(int d, int k, int depth) {
    return align8(8 * (size_t)d) + align8(8 * (size_t)(k + 1)) + align8(4 * (size_t)(k + 1)) +
           align8(4 * (size_t)(depth + 1)) + align8(8 * (size_t)(depth + 1)) + 8 * 64 + 16 + 8;
}

__device__ inline WaveLds carve(unsigned char* base, int d, int k, int depth) {
    WaveLds w;
    size_t o = 0;
    w.q = (double*)(base + o);
    o += align8(8 * (size_t)d);
    w.hd = (double*)(base + o);
    o += align8(8 * (size_t)(k + 1));
    w.hp = (int*)(base + o);
    o += align8(4 * (size_t)(k + 1));
    w.sn = (int*)(base + o);
    o += align8(4 * (size_t)(depth + 1));
    w.sg = (double*)(base + o);
    o += align8(8 * (size_t)(depth + 1));
    w.cd = (double*)(base + o);
    o += 8 * 64;
    w.st = (int*)(base + o);
    o += 16;
    w.kth = (double*)(base + o);
    return w;
}

// Whether the box [lo, hi] (box[0..d), box[d..2d)) is at least the k-th
// distance from the query: no point inside can be admitted. Wave-uniform
// (the butterfly leaves the same sum on every lane).
__device__ inline bool box_rejects(const double* box, int d, const double* q, double kth) {
#pragma clang fp contract(off)
    if (kth == DBL_MAX) return false;
    double t = 0.0;
    for (int i = threadIdx.x & 63; i < d; i += 64) {
        const double lo = box[i], hi = box[d + i], v = q[i];
        const double g = v < lo ? lo - v : (v > hi ? v - hi : 0.0);
        t += g * g;
    }
    for (int off = 32; off > 0; off >>= 1) t += __shfl_xor(t, off, 64);
    return t >= kth * (1.0 + 1e-12);
}

// (d1, p1) < (d2, p2) lexicographically (std::pair ordering).
__device__ inline bool less_dp(double d1, int p1, double d2, int p2) { return d1 < d2 || (d1 == d2 && p1 < p2); }

__device__ void heap_push(double* hd, int* hp, int& n, double d, int p) {
    int i = n++;
    while (i > 0) {
        int parent = (i - 1) >> 1;
        if (!less_dp(hd[parent], hp[parent], d, p)) break;
        hd[i] = hd[parent];
        hp[i] = hp[parent];
        i = parent;
    }
    hd[i] = d;
    hp[i] = p;
}

__device__ void heap_pop(double* hd, int* hp, int& n) {
    n--;
    double d = hd[n];
    int p = hp[n];
    int i = 0;
    while (true) {
        int c = 2 * i + 1;
        if (c >= n) break;
        if (c + 1 < n && less_dp(hd[c], hp[c], hd[c + 1], hp[c + 1])) c++;
        if (!less_dp(d, p, hd[c], hp[c])) break;
        hd[i] = hd[c];
        hp[i] = hp[c];
        i = c;
    }
    if (n > 0) {
        hd[i] = d;
        hp[i] = p;
    }
}

// The search of query qi by one wavefront, in its LDS L.
__device__ void search_query(const KnnArgs& a, const WaveLds& L, int qi) {
#pragma clang fp contract(off)
    const int lane = threadIdx.x & 63;
    const int d = a.d, k = a.k;
    for (int i = lane; i < d; i += 64) L.q[i] = a.queries[(size_t)qi * d + i];
    if (lane == 0) {
        L.st[0] = 0;  // heap size
        L.st[1] = 0;  // hits
        L.st[2] = 0;  // heap overflow
        L.kth[0] = DBL_MAX;
    }
    wave_sync();
    int sp = 0;
    int node = 0;
    bool done = false;
    auto boxed_out = [&](int id) {
        return a.boxes != nullptr && box_rejects(a.boxes + (size_t)id * 2 * d, d, L.q, L.kth[0]);
    };
    while (!done) {
        // descend to a leaf, stacking the far sides; a subtree whose box
        // is beyond the k-th distance is left at once
        bool skip = boxed_out(node);
        KdNode nd = a.nodes[node];
        while (!skip && nd.left >= 0) {
            double qv = L.q[nd.s];
            int near, far;
            double gap;
            if (qv < nd.split) {
                near = nd.left;
                far = nd.right;
                gap = nd.min_right - qv;
            } else {
                near = nd.right;
                far = nd.left;
                gap = qv - nd.max_left;
            }
            if (lane == 0) {
                L.sn[sp] = far;
                L.sg[sp] = gap * gap;
            }
            sp++;
            node = near;
            skip = boxed_out(node);
            nd = a.nodes[node];
        }
        // the leaf's points, 64 at a time; a leaf's block of the point array
        // is dimension-major (point j, dimension i at i * len + j), so the
        // lanes' loads coalesce
        const int len = skip ? 0 : nd.end - nd.begin;
        const double* blk = a.points + (size_t)nd.begin * d;
        for (int c0 = 0; c0 < len; c0 += 64) {
            const int j = c0 + lane;
            const double kth0 = L.kth[0];  // never below the k-th distance at admission
            bool alive = j < len;
            double s = 0.0;
            int i = 0;
            while (i < d - 3) {
                if (alive) {
                    double x1 = L.q[i] - blk[(size_t)i * len + j];
                    double x2 = L.q[i + 1] - blk[(size_t)(i + 1) * len + j];
                    double x3 = L.q[i + 2] - blk[(size_t)(i + 2) * len + j];
                    double x4 = L.q[i + 3] - blk[(size_t)(i + 3) * len + j];
                    s += x1 * x1 + x2 * x2 + x3 * x3 + x4 * x4;
                    if (s >= kth0) alive = false;  // leaf-node.h:52-55: cannot be admitted
                }
                i += 4;
                if (!__any(alive)) break;
            }
            if (alive) {
                for (; i < d; i++) {
                    double x = L.q[i] - blk[(size_t)i * len + j];
                    s += x * x;
                }
                alive = s < kth0;
            }
            if (alive) L.cd[lane] = s;
            unsigned long long cand = __ballot(alive);
            wave_sync();
            if (lane == 0 && cand) {
                int hn = L.st[0], hits = L.st[1];
                double kth = L.kth[0];
                while (cand) {
                    int jj = __builtin_ctzll(cand);  // ascending point order
                    cand &= cand - 1;
                    double dist = L.cd[jj];
                    if (dist >= kth) continue;
                    if (hn > k) {  // no room for a push: only points with no hits get here
                        L.st[2] = 1;
                        break;
                    }
                    int pos = nd.begin + c0 + jj;
                    hits += a.hits[pos];
                    heap_push(L.hd, L.hp, hn, dist, pos);
                    int top_hits = a.hits[L.hp[0]];
                    if (hits - top_hits >= k) {
                        hits -= top_hits;
                        heap_pop(L.hd, L.hp, hn);
                    }
                    if (hits >= k) kth = L.hd[0];
                }
                L.st[0] = hn;
                L.st[1] = hits;
                L.kth[0] = kth;
            }
            wave_sync();
        }
        // next far side worth visiting
        done = true;
        while (sp > 0) {
            sp--;
            double g = L.sg[sp];
            int far = L.sn[sp];
            if (L.kth[0] > g) {
                node = far;
                done = false;
                break;
            }
        }
    }
    // KBest::fill: farthest first
    if (lane == 0) {
        int hn = L.st[0];
        int* oi = a.out_index + (size_t)qi * k;
        double* od = a.out_dist_sq + (size_t)qi * k;
        a.out_count[qi] = L.st[2] ? -1 : hn;
        if (hn > k) hn = k;  // (overflow only) the k nearest still fit the output row
        int j = 0;
        while (hn > 0) {
            od[j] = L.hd[0];
            oi[j] = a.perm[L.hp[0]];
            j++;
            heap_pop(L.hd, L.hp, hn);
        }
        for (; j < k; j++) {  // unused slots
            od[j] = 0.0;
            oi[j] = -1;
        }
    }
}

// Wavefronts take queries in index order from a counter (a.next, zero at
// launch) until none is left: a slow query holds one wavefront rather than
// a share of the grid, so the launch ends with the last query, not with the
// slowest workgroup of every round.
__global__ void __launch_bounds__(64 * kWaves) knn_kernel(const KnnArgs a) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const WaveLds L = carve(smem + (size_t)wv * wave_bytes(a.d, a.k, a.depth), a.d, a.k, a.depth);
    for (;;) {
        int qi = 0;
        if (lane == 0) qi = atomicAdd(a.next, 1);
        qi = __shfl(qi, 0, 64);
        if (qi >= a.N) break;  // whole wavefront
        search_query(a, L, qi);
        wave_sync();  // the next query's set-up rewrites what the last one read
    }
}

__global__ void knn_zero_kernel(int* p) { *p = 0; }

}  // namespace

size_t knn_lds_bytes(int d, int k, int depth) { return kWaves * wave_bytes(d, k, depth); }

hipError_t launch_knn(const KnnArgs& a, hipStream_t stream) {
    if (a.N <= 0) return hipSuccess;
    size_t lds = knn_lds_bytes(a.d, a.k, a.depth);
    hipError_t e = hipFuncSetAttribute((const void*)knn_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(knn_zero_kernel, dim3(1), dim3(1), 0, stream, a.next);
    const int grid = std::min((a.N + kWaves - 1) / kWaves, kKnnGrid);
    hipLaunchKernelGGL(knn_kernel, dim3(grid), dim3(64 * kWaves), lds, stream, a);
    return hipGetLastError();
}

}  // namespace wp
