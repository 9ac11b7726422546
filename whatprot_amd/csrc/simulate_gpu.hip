// GPU simulator: dye tracks with their deduplication, and radiometries, from
// the sequencing model's forward process (simulation/generate-dye-track.cc:
// 29-110, generate-dye-tracks.cc:28-63, dedup-dye-tracks.cc:35-148,
// generate-radiometry.cc:28-62, generate-radiometries.cc:30-64).
//
// One thread simulates one peptide (dye tracks) or one read attempt
// (radiometries) from its own random stream (sim_rng.h), exactly as the host
// simulator in simulate.cpp does, so the integer tracks agree bit for bit.
//
// Deduplication: the visible tracks are packed into 64-bit key words (the
// counts in timestep-major, channel-minor order, most significant first, so
// the numeric order of the words is the lexicographic order of the counts)
// and sorted with the source dye-seq id as the last key by stable LSD radix
// passes; flags where a word or the id changes, two scans and a scatter then
// give the distinct tracks in lexicographic order with, per source id in
// increasing order, its count and its number of hits.
#include <hip/hip_runtime.h>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <cstring>
#include <vector>

#include "hmm_device.h"
#include "sim_rng.h"
#include "wp_internal.h"

namespace wp {
namespace {

constexpr int kMaxSeqLen = 64;  // longest dye-seq the GPU simulator takes

struct SimArgs {
    int T, C, cap;
    const short* seqs;  // [nds][cap] channel per position, -1 = none
    const int* lens;
    double pd[WP_MAX_TIMESTEPS];
    double p_ib, p_cb, pf;
    double p_dud[WP_MAX_CHANNELS], p_bleach[WP_MAX_CHANNELS];
    uint64_t seed;
};

// generate-dye-track.cc:29-110 (as simulate.cpp's dye_track): duds; initial
// block; per cycle: record, cyclic block, detach, Edman unless blocked,
// bleach. Returns whether the track is visible (non-zero at t = 0).
__device__ bool sim_dye_track(const SimArgs& a, int i, Rng& rng, uint8_t* track) {
    const int len = a.lens[i];
    const short* seq = a.seqs + (size_t)i * a.cap;
    short ds[kMaxSeqLen];
    for (int p = 0; p < len; p++) {
        ds[p] = seq[p];
        if (ds[p] != -1 && rng.bernoulli(a.p_dud[ds[p]])) ds[p] = -1;
    }
    int counts[WP_MAX_CHANNELS] = {0, 0, 0, 0};
    for (int p = 0; p < len; p++)
        if (ds[p] != -1) counts[ds[p]]++;
    int e = 0;
    bool block = rng.bernoulli(a.p_ib);
    for (int t = 0; t < a.T; t++) {
        if (e < len) {
            for (int c = 0; c < a.C; c++) track[t * a.C + c] = (uint8_t)counts[c];
            if (rng.bernoulli(a.p_cb)) block = true;
            if (rng.bernoulli(a.pd[t])) {
                e = len;
                continue;
            }
            if (!block) {
                if (!rng.bernoulli(a.pf)) {
                    if (ds[e] != -1) counts[ds[e]]--;
                    e++;
                }
            }
            for (int p = e; p < len; p++) {
                if (ds[p] != -1 && rng.bernoulli(a.p_bleach[ds[p]])) {
                    counts[ds[p]]--;
                    ds[p] = -1;
                }
            }
        } else {
            for (int c = 0; c < a.C; c++) track[t * a.C + c] = 0;
        }
    }
    bool visible = false;
    for (int c = 0; c < a.C; c++) visible = visible || track[c] != 0;
    return visible;
}

// Peptide g of dye-seq i = the last i with first[i] <= g / per_peptide.
__global__ void sim_tracks_kernel(const SimArgs a, long total, int per_peptide, int nds,
                                  const long* __restrict__ first, uint8_t* __restrict__ tracks,
                                  int* __restrict__ dsidx, int* __restrict__ visible) {
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= total) return;
    const long pep = g / per_peptide;
    int lo = 0, hi = nds - 1;
    while (lo < hi) {  // largest i with first[i] <= pep
        const int mid = (lo + hi + 1) / 2;
        if (first[mid] <= pep) lo = mid;
        else hi = mid - 1;
    }
    Rng rng = stream_rng(a.seed, (uint64_t)g);
    visible[g] = sim_dye_track(a, lo, rng, tracks + (size_t)g * a.T * a.C) ? 1 : 0;
    dsidx[g] = lo;
}

// Compacts the visible peptides (pos = exclusive scan of the flags) and
// packs their keys: word w holds counts w*per .. w*per+per-1 of the flattened
// track, `bits` bits each, the first count in the highest bits.
__global__ void sim_pack_kernel(long total, int TC, int bits, int words, const uint8_t* __restrict__ tracks,
                                const int* __restrict__ visible, const int* __restrict__ pos,
                                const int* __restrict__ dsidx, const int* __restrict__ ids,
                                unsigned long long* __restrict__ keys, unsigned int* __restrict__ idkey,
                                int* __restrict__ src) {
    const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= total || !visible[g]) return;
    const int v = pos[g];
    const int per = 64 / bits;
    const uint8_t* tr = tracks + (size_t)g * TC;
    for (int w = 0; w < words; w++) {
        unsigned long long k = 0;
        for (int j = 0; j < per; j++) {
            const int q = w * per + j;
            k = (k << bits) | (unsigned long long)(q < TC ? tr[q] : 0);
        }
        keys[(size_t)v * words + w] = k;
    }
    idkey[v] = (unsigned int)ids[dsidx[g]] ^ 0x80000000u;  // order-preserving for signed ids
    src[v] = (int)g;
}

__global__ void iota_kernel(int n, int* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = i;
}

__global__ void gather_key_kernel(int n, int words, int w, const unsigned long long* __restrict__ keys,
                                  const int* __restrict__ perm, unsigned long long* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = keys[(size_t)perm[i] * words + w];
}

__global__ void gather_id_kernel(int n, const unsigned int* __restrict__ idkey, const int* __restrict__ perm,
                                 unsigned long long* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = idkey[perm[i]];
}

// Segment starts in sorted order: a new track where a key word differs from
// the previous element, a new (track, id) group where either differs.
__global__ void sim_flags_kernel(int n, int words, const unsigned long long* __restrict__ keys,
                                 const unsigned int* __restrict__ idkey, const int* __restrict__ perm,
                                 int* __restrict__ tflag, int* __restrict__ gflag) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int nt = i == 0, ng = i == 0;
    if (i > 0) {
        const int a = perm[i], b = perm[i - 1];
        for (int w = 0; w < words; w++) nt |= keys[(size_t)a * words + w] != keys[(size_t)b * words + w];
        ng = nt | (idkey[a] != idkey[b]);
    }
    tflag[i] = nt;
    gflag[i] = ng;
}

__global__ void sim_scatter_kernel(int n, int TC, const int* __restrict__ perm, const int* __restrict__ tflag,
                                   const int* __restrict__ gflag, const int* __restrict__ tidx,
                                   const int* __restrict__ gidx, const int* __restrict__ src,
                                   const uint8_t* __restrict__ tracks, const int* __restrict__ dsidx,
                                   const int* __restrict__ ids, const int* __restrict__ counts,
                                   short* __restrict__ out_tracks, int* __restrict__ out_off,
                                   int* __restrict__ gstart, int* __restrict__ out_id,
                                   int* __restrict__ out_count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const long g = src[perm[i]];
    if (gflag[i]) {
        const int k = gidx[i] - 1;
        gstart[k] = i;
        out_id[k] = ids[dsidx[g]];
        out_count[k] = counts[dsidx[g]];
    }
    if (tflag[i]) {
        const int t = tidx[i] - 1;
        out_off[t] = gidx[i] - 1;
        for (int q = 0; q < TC; q++) out_tracks[(size_t)t * TC + q] = (short)tracks[(size_t)g * TC + q];
    }
}

__global__ void sim_hits_kernel(int G, int n, const int* __restrict__ gstart, int* __restrict__ hits) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < G) hits[k] = (k + 1 < G ? gstart[k + 1] : n) - gstart[k];
}

// Read attempts base .. base + count - 1 (generate-radiometries.cc:30-64):
// a dye-seq by weight, its track, and for a visible track the intensities
// Normal(adjusted_mu, sigma) rounded to float32.
template <int C>
__global__ void sim_reads_kernel(const SimArgs a, const DevModel m, long base, int count, int nds,
                                 const double* __restrict__ cum, double acc, uint8_t* __restrict__ tracks,
                                 int* __restrict__ visible, int* __restrict__ dsidx, double* __restrict__ rads) {
#pragma clang fp contract(off)
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    Rng rng = stream_rng(a.seed, (uint64_t)(base + k));
    const double u = rng.uniform() * acc;
    int lo = 0, hi = nds;  // upper_bound(cum, u)
    while (lo < hi) {
        const int mid = (lo + hi) / 2;
        if (cum[mid] <= u) lo = mid + 1;
        else hi = mid;
    }
    const int i = lo >= nds ? nds - 1 : lo;
    uint8_t* tr = tracks + (size_t)k * a.T * C;
    const bool vis = sim_dye_track(a, i, rng, tr);
    visible[k] = vis ? 1 : 0;
    dsidx[k] = i;
    if (!vis) return;
    for (int t = 0; t < a.T; t++) {
        unsigned int cnt[C];
        for (int c = 0; c < C; c++) cnt[c] = tr[t * C + c];
        for (int c = 0; c < C; c++) {
            const double mu = dev_adjusted_mu<C>(m, c, cnt);
            const double sd = dev_sigma(m, c, mu);
            rads[((size_t)k * a.T + t) * C + c] = (double)(float)(mu + sd * rng.normal());
        }
    }
}

__global__ void sim_take_kernel(int count, int TC, int need, const int* __restrict__ visible,
                                const int* __restrict__ pos, const int* __restrict__ dsidx,
                                const double* __restrict__ rads, double* __restrict__ out, int* __restrict__ out_idx) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count || !visible[k]) return;
    const int v = pos[k];
    if (v >= need) return;
    for (int q = 0; q < TC; q++) out[(size_t)v * TC + q] = rads[(size_t)k * TC + q];
    out_idx[v] = dsidx[k];
}

struct DevBuf {
    std::vector<void*> ptrs;
    ~DevBuf() {
        for (void* p : ptrs) hipFree(p);
    }
    template <typename T>
    T* alloc(size_t n) {
        void* p = nullptr;
        if (hipMalloc(&p, sizeof(T) * std::max<size_t>(n, 1)) != hipSuccess) return nullptr;
        ptrs.push_back(p);
        return (T*)p;
    }
};

#define SIM_CHECK(x)                                                         \
    do {                                                                     \
        hipError_t e_ = (x);                                                 \
        if (e_ != hipSuccess) {                                              \
            set_error(std::string("GPU simulator: ") + hipGetErrorString(e_)); \
            return WP_ERR_HIP;                                               \
        }                                                                    \
    } while (0)
#define SIM_ALLOC(p, T, n)                                                   \
    do {                                                                     \
        p = bufs.alloc<T>(n);                                                \
        if (!p) {                                                            \
            set_error("GPU simulator: out of device memory");                \
            return WP_ERR_HIP;                                               \
        }                                                                    \
    } while (0)

template <typename F>
hipError_t with_temp(DevBuf& bufs, F f) {
    size_t bytes = 0;
    hipError_t e = f(nullptr, bytes);
    if (e != hipSuccess) return e;
    void* tmp = bufs.alloc<unsigned char>(bytes);
    if (!tmp) return hipErrorOutOfMemory;
    return f(tmp, bytes);
}

int fill_sim_args(const wp_model& m, int T, int C, int cap, uint64_t seed, SimArgs* a) {
    std::memset(a, 0, sizeof(*a));
    a->T = T;
    a->C = C;
    a->cap = cap;
    detach_table(m, T, a->pd);
    a->p_ib = m.p_initial_block;
    a->p_cb = m.p_cyclic_block;
    a->pf = m.p_edman_failure;
    for (int c = 0; c < C; c++) {
        a->p_dud[c] = m.channels[c].p_dud;
        a->p_bleach[c] = m.channels[c].p_bleach;
    }
    a->seed = seed;
    return WP_OK;
}

}  // namespace

int gpu_simulate_dye_tracks(const wp_model& m, int T, int C, int nds, const short* seqs, int cap, const int* lens,
                            const int* counts, const int* ids, int per_peptide, uint64_t seed, int max_tracks,
                            int max_sources, int* out_num, int* out_num_sources, short* out_tracks, int* out_off, int* out_id, int* out_count,
                            int* out_hits) {
    if (cap > kMaxSeqLen) {
        set_error("GPU simulator: dye-seqs longer than 64 positions");
        return WP_ERR_UNSUPPORTED;
    }
    const int TC = T * C;
    std::vector<long> first(nds);
    long peptides = 0;
    for (int i = 0; i < nds; i++) {
        first[i] = peptides;
        peptides += std::max(counts[i], 0);
    }
    const long total = peptides * per_peptide;
    if (total >= (1l << 31)) {
        set_error("GPU simulator: more than 2^31 simulated peptides");
        return WP_ERR_ARG;
    }
    *out_num = 0;
    *out_num_sources = 0;
    if (total == 0) {
        if (out_off) out_off[0] = 0;
        return WP_OK;
    }
    int maxc = 1;  // widest count: at most the dyes of one channel
    for (int i = 0; i < nds; i++) {
        int k[WP_MAX_CHANNELS] = {0, 0, 0, 0};
        for (int p = 0; p < lens[i]; p++)
            if (seqs[(size_t)i * cap + p] >= 0) k[seqs[(size_t)i * cap + p]]++;
        for (int c = 0; c < C; c++) maxc = std::max(maxc, k[c]);
    }
    int bits = 1;
    while ((1 << bits) <= maxc) bits++;
    const int per = 64 / bits, words = (TC + per - 1) / per;

    SimArgs a;
    fill_sim_args(m, T, C, cap, seed, &a);
    DevBuf bufs;
    short* d_seqs;
    int *d_lens, *d_ids, *d_counts, *d_dsidx, *d_vis, *d_pos;
    long* d_first;
    uint8_t* d_tracks;
    SIM_ALLOC(d_seqs, short, (size_t)nds * cap);
    SIM_ALLOC(d_lens, int, nds);
    SIM_ALLOC(d_ids, int, nds);
    SIM_ALLOC(d_counts, int, nds);
    SIM_ALLOC(d_first, long, nds);
    SIM_ALLOC(d_tracks, uint8_t, (size_t)total * TC);
    SIM_ALLOC(d_dsidx, int, total);
    SIM_ALLOC(d_vis, int, total);
    SIM_ALLOC(d_pos, int, total);
    SIM_CHECK(hipMemcpy(d_seqs, seqs, sizeof(short) * nds * cap, hipMemcpyHostToDevice));
    SIM_CHECK(hipMemcpy(d_lens, lens, sizeof(int) * nds, hipMemcpyHostToDevice));
    SIM_CHECK(hipMemcpy(d_ids, ids, sizeof(int) * nds, hipMemcpyHostToDevice));
    SIM_CHECK(hipMemcpy(d_counts, counts, sizeof(int) * nds, hipMemcpyHostToDevice));
    SIM_CHECK(hipMemcpy(d_first, first.data(), sizeof(long) * nds, hipMemcpyHostToDevice));
    a.seqs = d_seqs;
    a.lens = d_lens;
    const unsigned blocks = (unsigned)((total + 255) / 256);
    hipLaunchKernelGGL(sim_tracks_kernel, dim3(blocks), dim3(256), 0, 0, a, total, per_peptide, nds, d_first,
                       d_tracks, d_dsidx, d_vis);
    SIM_CHECK(hipGetLastError());
    SIM_CHECK(with_temp(bufs, [&](void* tmp, size_t& bytes) {
        return rocprim::exclusive_scan(tmp, bytes, d_vis, d_pos, 0, (size_t)total, rocprim::plus<int>());
    }));
    int last_pos = 0, last_vis = 0;
    SIM_CHECK(hipMemcpy(&last_pos, d_pos + total - 1, sizeof(int), hipMemcpyDeviceToHost));
    SIM_CHECK(hipMemcpy(&last_vis, d_vis + total - 1, sizeof(int), hipMemcpyDeviceToHost));
    const int n = last_pos + last_vis;  // visible peptides
    if (n == 0) {
        if (out_off) out_off[0] = 0;
        return WP_OK;
    }
    unsigned long long *d_keys, *d_k0, *d_k1;
    unsigned int* d_idkey;
    int *d_src, *d_perm0, *d_perm1;
    SIM_ALLOC(d_keys, unsigned long long, (size_t)n * words);
    SIM_ALLOC(d_k0, unsigned long long, n);
    SIM_ALLOC(d_k1, unsigned long long, n);
    SIM_ALLOC(d_idkey, unsigned int, n);
    SIM_ALLOC(d_src, int, n);
    SIM_ALLOC(d_perm0, int, n);
    SIM_ALLOC(d_perm1, int, n);
    hipLaunchKernelGGL(sim_pack_kernel, dim3(blocks), dim3(256), 0, 0, total, TC, bits, words, d_tracks, d_vis,
                       d_pos, d_dsidx, d_ids, d_keys, d_idkey, d_src);
    SIM_CHECK(hipGetLastError());
    const unsigned nb = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(iota_kernel, dim3(nb), dim3(256), 0, 0, n, d_perm0);
    // stable LSD passes: the id first, then the key words from the last
    for (int pass = -1; pass < words; pass++) {
        const int w = words - 1 - pass;  // pass -1: the id
        if (pass < 0)
            hipLaunchKernelGGL(gather_id_kernel, dim3(nb), dim3(256), 0, 0, n, d_idkey, d_perm0, d_k0);
        else
            hipLaunchKernelGGL(gather_key_kernel, dim3(nb), dim3(256), 0, 0, n, words, w, d_keys, d_perm0, d_k0);
        SIM_CHECK(hipGetLastError());
        const int end_bit = pass < 0 ? 32 : 64;
        SIM_CHECK(with_temp(bufs, [&](void* tmp, size_t& bytes) {
            return rocprim::radix_sort_pairs(tmp, bytes, d_k0, d_k1, d_perm0, d_perm1, (size_t)n, 0, end_bit);
        }));
        std::swap(d_perm0, d_perm1);
    }
    int *d_tflag, *d_gflag, *d_tidx, *d_gidx, *d_gstart, *d_off, *d_gid, *d_gcount, *d_hits;
    short* d_out_tracks;
    SIM_ALLOC(d_tflag, int, n);
    SIM_ALLOC(d_gflag, int, n);
    SIM_ALLOC(d_tidx, int, n);
    SIM_ALLOC(d_gidx, int, n);
    hipLaunchKernelGGL(sim_flags_kernel, dim3(nb), dim3(256), 0, 0, n, words, d_keys, d_idkey, d_perm0, d_tflag,
                       d_gflag);
    SIM_CHECK(hipGetLastError());
    SIM_CHECK(with_temp(bufs, [&](void* tmp, size_t& bytes) {
        return rocprim::inclusive_scan(tmp, bytes, d_tflag, d_tidx, (size_t)n, rocprim::plus<int>());
    }));
    SIM_CHECK(with_temp(bufs, [&](void* tmp, size_t& bytes) {
        return rocprim::inclusive_scan(tmp, bytes, d_gflag, d_gidx, (size_t)n, rocprim::plus<int>());
    }));
    int ntracks = 0, ngroups = 0;
    SIM_CHECK(hipMemcpy(&ntracks, d_tidx + n - 1, sizeof(int), hipMemcpyDeviceToHost));
    SIM_CHECK(hipMemcpy(&ngroups, d_gidx + n - 1, sizeof(int), hipMemcpyDeviceToHost));
    *out_num = ntracks;
    *out_num_sources = ngroups;
    if (ntracks > max_tracks || ngroups > max_sources) {
        set_error("more distinct dye-tracks or sources than max_tracks / max_sources");
        return WP_ERR_ARG;
    }
    if (!out_tracks || !out_off || !out_id || !out_count || !out_hits) {
        set_error("null output");
        return WP_ERR_ARG;
    }
    SIM_ALLOC(d_gstart, int, ngroups);
    SIM_ALLOC(d_off, int, ntracks + 1);
    SIM_ALLOC(d_gid, int, ngroups);
    SIM_ALLOC(d_gcount, int, ngroups);
    SIM_ALLOC(d_hits, int, ngroups);
    SIM_ALLOC(d_out_tracks, short, (size_t)ntracks * TC);
    hipLaunchKernelGGL(sim_scatter_kernel, dim3(nb), dim3(256), 0, 0, n, TC, d_perm0, d_tflag, d_gflag, d_tidx,
                       d_gidx, d_src, d_tracks, d_dsidx, d_ids, d_counts, d_out_tracks, d_off, d_gstart, d_gid,
                       d_gcount);
    SIM_CHECK(hipGetLastError());
    hipLaunchKernelGGL(sim_hits_kernel, dim3((ngroups + 255) / 256), dim3(256), 0, 0, ngroups, n, d_gstart, d_hits);
    SIM_CHECK(hipGetLastError());
    SIM_CHECK(hipMemcpy(out_tracks, d_out_tracks, sizeof(short) * (size_t)ntracks * TC, hipMemcpyDeviceToHost));
    SIM_CHECK(hipMemcpy(out_off, d_off, sizeof(int) * ntracks, hipMemcpyDeviceToHost));
    out_off[ntracks] = ngroups;
    SIM_CHECK(hipMemcpy(out_id, d_gid, sizeof(int) * ngroups, hipMemcpyDeviceToHost));
    SIM_CHECK(hipMemcpy(out_count, d_gcount, sizeof(int) * ngroups, hipMemcpyDeviceToHost));
    SIM_CHECK(hipMemcpy(out_hits, d_hits, sizeof(int) * ngroups, hipMemcpyDeviceToHost));
    return WP_OK;
}

int gpu_simulate_radiometries(const wp_model& m, const DevModel& dm, int T, int C, int nds, const short* seqs,
                              int cap, const int* lens, const int* weights, int n, uint64_t seed, double* out,
                              int* out_idx) {
    if (cap > kMaxSeqLen) {
        set_error("GPU simulator: dye-seqs longer than 64 positions");
        return WP_ERR_UNSUPPORTED;
    }
    if (n <= 0) return WP_OK;
    const int TC = T * C;
    std::vector<double> cum(nds);
    double acc = 0.0;
    for (int i = 0; i < nds; i++) {
        acc += weights ? (double)weights[i] : 1.0;
        cum[i] = acc;
    }
    SimArgs a;
    fill_sim_args(m, T, C, cap, seed, &a);
    DevBuf bufs;
    short* d_seqs;
    int *d_lens, *d_vis, *d_pos, *d_dsidx, *d_outidx;
    double *d_cum, *d_rads, *d_out;
    uint8_t* d_tracks;
    const int batch = std::max(65536, std::min(n + n / 4 + 1024, 1 << 22));
    SIM_ALLOC(d_seqs, short, (size_t)nds * cap);
    SIM_ALLOC(d_lens, int, nds);
    SIM_ALLOC(d_cum, double, nds);
    SIM_ALLOC(d_tracks, uint8_t, (size_t)batch * TC);
    SIM_ALLOC(d_vis, int, batch);
    SIM_ALLOC(d_pos, int, batch);
    SIM_ALLOC(d_dsidx, int, batch);
    SIM_ALLOC(d_rads, double, (size_t)batch * TC);
    SIM_ALLOC(d_out, double, (size_t)n * TC);
    SIM_ALLOC(d_outidx, int, n);
    SIM_CHECK(hipMemcpy(d_seqs, seqs, sizeof(short) * nds * cap, hipMemcpyHostToDevice));
    SIM_CHECK(hipMemcpy(d_lens, lens, sizeof(int) * nds, hipMemcpyHostToDevice));
    SIM_CHECK(hipMemcpy(d_cum, cum.data(), sizeof(double) * nds, hipMemcpyHostToDevice));
    a.seqs = d_seqs;
    a.lens = d_lens;
    long base = 0;
    int made = 0;
    while (made < n) {  // attempts in order until n visible reads exist
        const unsigned blocks = (unsigned)((batch + 255) / 256);
        switch (C) {
        case 1: hipLaunchKernelGGL(sim_reads_kernel<1>, dim3(blocks), dim3(256), 0, 0, a, dm, base, batch, nds, d_cum, acc, d_tracks, d_vis, d_dsidx, d_rads); break;
        case 2: hipLaunchKernelGGL(sim_reads_kernel<2>, dim3(blocks), dim3(256), 0, 0, a, dm, base, batch, nds, d_cum, acc, d_tracks, d_vis, d_dsidx, d_rads); break;
        case 3: hipLaunchKernelGGL(sim_reads_kernel<3>, dim3(blocks), dim3(256), 0, 0, a, dm, base, batch, nds, d_cum, acc, d_tracks, d_vis, d_dsidx, d_rads); break;
        default: hipLaunchKernelGGL(sim_reads_kernel<4>, dim3(blocks), dim3(256), 0, 0, a, dm, base, batch, nds, d_cum, acc, d_tracks, d_vis, d_dsidx, d_rads); break;
        }
        SIM_CHECK(hipGetLastError());
        SIM_CHECK(with_temp(bufs, [&](void* tmp, size_t& bytes) {
            return rocprim::exclusive_scan(tmp, bytes, d_vis, d_pos, made, (size_t)batch, rocprim::plus<int>());
        }));
        hipLaunchKernelGGL(sim_take_kernel, dim3(blocks), dim3(256), 0, 0, batch, TC, n, d_vis, d_pos, d_dsidx,
                           d_rads, d_out, d_outidx);
        SIM_CHECK(hipGetLastError());
        int last_pos = 0, last_vis = 0;
        SIM_CHECK(hipMemcpy(&last_pos, d_pos + batch - 1, sizeof(int), hipMemcpyDeviceToHost));
        SIM_CHECK(hipMemcpy(&last_vis, d_vis + batch - 1, sizeof(int), hipMemcpyDeviceToHost));
        made = std::min(n, last_pos + last_vis);
        base += batch;
        if (base > (1l << 40)) {
            set_error("GPU simulator: too few visible reads");
            return WP_ERR_ARG;
        }
    }
    SIM_CHECK(hipMemcpy(out, d_out, sizeof(double) * (size_t)n * TC, hipMemcpyDeviceToHost));
    if (out_idx) SIM_CHECK(hipMemcpy(out_idx, d_outidx, sizeof(int) * n, hipMemcpyDeviceToHost));
    return WP_OK;
}

}  // namespace wp
