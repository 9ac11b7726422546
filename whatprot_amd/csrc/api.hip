// C ABI of libwhatprot_amd: model helpers, the HMM classifier context and the
// synthetic-input generator. Declarations and reference citations are in
// include/whatprot_amd.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "hmm_kernels.h"
#include "model_tables.h"
#include "wp_internal.h"

namespace wp {

static thread_local std::string g_err;

void set_error(const std::string& msg) { g_err = msg; }

int parse_dye_seq(const char* s, short* seq, int cap) {
    int len = (int)std::strlen(s);
    while (len > 0 && s[len - 1] == '.') len--;
    for (int i = 0; i < len && i < cap; i++) {
        char ch = s[i];
        seq[i] = (ch >= '0' && ch <= '9') ? (short)(ch - '0') : (short)-1;
    }
    return len;
}

void fill_dev_model(const wp_model& m, int T, int C, int M, double cutoff, DevModel* out) {
    DevModel& dm = *out;
    std::memset(&dm, 0, sizeof(dm));
    dm.T = T;
    dm.C = C;
    dm.M = M;
    dm.uncut = cutoff == DBL_MAX;
    dm.cutoff = cutoff;
    dm.pf = m.p_edman_failure;
    dm.p_ib = m.p_initial_block;
    dm.p_cb = m.p_cyclic_block;
    for (int i = 0; i < T; i++)
        dm.p_detach[i] = m.p_detach + m.p_initial_detach * std::exp(-(double)i * m.p_initial_detach_decay);
    for (int k = 0; k < C; k++) {
        dm.mu[k] = m.channels[k].mu;
        dm.sig[k] = m.channels[k].sig;
        dm.bg_sig[k] = m.channels[k].bg_sig;
        for (int o = 0; o < C; o++) {
            dm.inter[k][o] = m.channels[k].interactions[o];
            dm.flat[k][o] = m.channels[k].flat_interactions[o];
        }
    }
}

void binomial_tables(const wp_model& m, int C, int M, double* out) {
    const int tri = (M + 1) * (M + 2) / 2;
    for (int kind = 0; kind < 2; kind++) {
        for (int k = 0; k < C; k++) {
            double q = kind == 0 ? m.channels[k].p_dud : m.channels[k].p_bleach;
            double p = 1.0 - q;
            double* v = &out[((size_t)kind * C + k) * tri];
            v[0] = 1.0;
            for (int i = 1; i <= M; i++) {
                int row = i * (i + 1) / 2, prev = (i - 1) * i / 2;
                v[row] = v[prev] * q;
                for (int j = 1; j < i; j++) v[row + j] = v[prev + j] * q + v[prev + j - 1] * p;
                v[row + i] = v[prev + i - 1] * p;
            }
        }
    }
}

bool dye_channels_ok(const short* seqs, int n, int cap, const int* lens, int C) {
    for (int i = 0; i < n; i++) {
        for (int k = 0; k < lens[i]; k++) {
            if (seqs[(size_t)i * cap + k] >= C) {
                set_error("dye-seq names a channel >= num_channels");
                return false;
            }
        }
    }
    return true;
}

int require_device() {
    if (wp_device_count() < 1) {
        set_error("no gfx950 device visible");
        return WP_ERR_NO_DEVICE;
    }
    return WP_OK;
}

}  // namespace wp

namespace wp {

// Device workspace of one call, carved from one block (arrays 256-byte
// aligned): per-read boxes and bin lists, and for the FP32 path the near-best
// lists and the FP64 rescoring pass's own lists and outputs.
struct Workspace {
    uint8_t* L = nullptr;        // [N][T][C] read-level box lows
    uint8_t* H = nullptr;        // [N][T][C] read-level box highs
    int* nb = nullptr;           // [N] box bound
    uint8_t* wide = nullptr;     // [N] per read: handed to its bin's wide pass
    int* lists = nullptr;        // [nbins][N] reads of each bin
    int* wide_lists = nullptr;   // [nbins][N] reads each bin's main launch hands to its wide pass
    int* wide2_lists = nullptr;  // [nbins][N] gather bins: reads the full-size separable pass hands on
    int* ctr = nullptr;          // counters: bin sizes, wide-list sizes
    // FP32 only
    int* tie_n = nullptr;        // [N] near-best list length (0 none, -1 every candidate)
    int* tie_idx = nullptr;      // [N][kTieMax]
    int* rid = nullptr;          // [N] FP64 pass: id
    double* rls = nullptr;       // [N] FP64 pass: ln score
    double* lt = nullptr;        // [N] ln total when the caller does not ask for it
    uint8_t* rwide = nullptr;    // FP64 pass: as wide, lists, wide_lists, wide2_lists, ctr
    int* rlists = nullptr;
    int* rwide_lists = nullptr;
    int* rwide2_lists = nullptr;
    int* rctr = nullptr;
};

enum BinKernel { kGather = 0, kRowPerLane = 3, kQuadRows = 5, kLaneGroup = 6 };

struct Bin {
    int lo, hi, ngroups;
    // BinKernel: lane-group (FP32, boxes <= 16, T <= 12), quad-rows (FP32,
    // boxes <= 16), row-per-lane (FP64, <= 8) or gather
    int kernel;
    int cap_sep, ngroups_sep;  // gather bins: state capacity and groups of the separable launch
    int ngroups_wide;          // groups of the wide passes (few reads: as many groups as LDS allows)
};

// The bins of a classifier, fixed at creation (no heap use per call).
struct BinSet {
    Bin b[kMaxBins];
    int n = 0;
    int size() const { return n; }
    const Bin& operator[](int i) const { return b[i]; }
    const Bin* begin() const { return b; }
    const Bin* end() const { return b + n; }
};

}  // namespace wp

// Device staging of the host-array entry points, kept across calls and grown
// as needed (no allocation on calls that fit).
struct HostStage {
    double* rads = nullptr;
    int* id = nullptr;
    double* out = nullptr;
    double* pair = nullptr;
    int* off = nullptr;
    int* idx = nullptr;
    size_t n_rads = 0, n_reads = 0, n_pair = 0, n_idx = 0;
};

struct wp_hmm_classifier {
    int T = 0, C = 0, S = 0, M = 0, precision = 0, G = 16, tri = 1;
    HostStage hs;
    wp::DevModel dm;
    std::vector<int> ids, counts;
    uint8_t* dD = nullptr;
    int8_t* dSeq = nullptr;
    uint8_t* dTrack = nullptr;
    int* dCount = nullptr;
    int* dId = nullptr;
    double* dBinom = nullptr;
    // 1: the device tables and the per-read boxes belong to the FP32
    // classifier this FP64 sibling rescoring pass serves
    int borrowed = 0;
    // per-read workspace of the device path (one block, wcap reads)
    wp::Workspace ws;
    void* wbuf = nullptr;
    int wcap = 0;
    int maxbox = 1;  // prod_c (M+1), bound on any read's box
    wp::BinSet bins;  // read bins by box bound, with their kernels and LDS sizing
    // near-best list lengths of the last call (its workspace) and its reads
    const int* last_tie_n = nullptr;
    int last_n = 0;
    // FP32: near-best candidates decided in FP64 by the sibling's kernels
    wp_hmm_classifier* r64 = nullptr;
    int refine = 0;
    double tie_abs = 0.0, tie_rel = 0.0;
    // quad-rows kernel variant per bin: chosen by timing both on the first
    // call (variant_mode -1) or forced (0 bin-wide boxes, 1 per-timestep classes)
    int variant_mode = -1;
    int tuned = 0;
    int variants_set = 0;  // variant[] holds the defaults or the tuned choice
    int variant[wp::kMaxBins] = {0};
    float tune_ms[wp::kMaxBins][2];  // autotuner's best time per variant (-1 untimed)
    wp_hmm_classifier() {
        for (int b = 0; b < wp::kMaxBins; b++) tune_ms[b][0] = tune_ms[b][1] = -1.f;
    }
    // the device entry points replay a captured graph of the launch sequence
    // while the call's shape and pointers repeat
    int use_graph = 1;
    // the bins of a call run as three parallel branches (0: in order on the
    // caller's stream, as the no-allocation entry point always does)
    int branches = 1;
    hipStream_t cap_stream = nullptr;
    hipGraphExec_t gexec = nullptr;
    std::vector<const void*> gkey;
    // optional kernel timing: events around prep and classify launches
    int timing = 0;
    hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
    int ev_recorded = 0;
    // the bins of a call run as parallel branches (the caller's stream and
    // these), so one bin's tail of heavy reads overlaps the others' work
    hipStream_t branch[2] = {nullptr, nullptr};
    hipEvent_t fork = nullptr, join[2] = {nullptr, nullptr};
};

namespace wp {

bool dye_seq_index_of_ids(const wp_hmm_classifier* c, std::unordered_map<int, int>* out) {
    out->clear();
    for (int i = 0; i < c->S; i++)
        if (!out->emplace(c->ids[i], i).second) return false;
    return true;
}

}  // namespace wp

namespace {

using wp::Bin;
using wp::BinSet;
using wp::kGather;
using wp::kLaneGroup;
using wp::kQuadRows;
using wp::kRowPerLane;

// The separable gather launch sizes its state buffers at 1/kSepCapDiv of the
// T x NB entries the bin could need, so more pairs fit in LDS; a read that
// needs more goes to the composite wide pass.
constexpr int kSepCapDiv = 2;

size_t bin_lds_bytes(const wp_hmm_classifier* c, const Bin& b);

constexpr size_t kLdsBudget = 48 * 1024;

// Workgroups of the row-per-lane and gather launches, which stride over their
// read list: most of their lists (wide passes, the FP64 decision, the bins of
// large boxes) hold a small part of the call's reads, and a grid of one
// workgroup per read spent about 1.5 ms per launch on 10^5 exits.
constexpr int kListGrid = 8192;

// Largest dynamic LDS one workgroup may take on the current device (160 KB
// on gfx950), queried once.
size_t lds_hard_limit() {
    static size_t lim = 0;
    if (!lim) {
        int dev = 0;
        hipDeviceProp_t p;
        lim = 64 * 1024;
        if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&p, dev) == hipSuccess &&
            p.sharedMemPerBlock > lim)
            lim = p.sharedMemPerBlock;
    }
    return lim;
}

// Bins of reads by box bound, each with its LDS sizing.
BinSet make_bins(const wp_hmm_classifier* c) {
    const int edges[] = {4, 8, 16, 32, 64, 128, c->maxbox};
    BinSet bins;
    int lo = 0;
    for (int hi : edges) {
        if (hi <= lo) continue;
        if (hi > c->maxbox) hi = c->maxbox;
        auto groups_for = [&](int cap, size_t budget) {
            size_t fixed = wp::classify_lds_bytes(c->precision, c->T, c->C, c->tri, hi, cap, 0, c->G);
            size_t per = wp::classify_lds_bytes(c->precision, c->T, c->C, c->tri, hi, cap, 1, c->G) - fixed;
            int g = (int)((budget > fixed ? budget - fixed : 0) / (per ? per : 1));
            g = std::max(1, std::min(g, wp::kBlockThreads / c->G));
            // the layout also grows with the thread count: check the exact size
            while (g > 1 && wp::classify_lds_bytes(c->precision, c->T, c->C, c->tri, hi, cap, g, c->G) > budget) g--;
            return g;
        };
        const int ng = groups_for(c->T * hi, kLdsBudget);
        const int cap_sep = std::max(1, c->T * hi / kSepCapDiv);
        const int ng_sep = groups_for(cap_sep, kLdsBudget);
        const int ng_wide = groups_for(c->T * hi, lds_hard_limit() - 1024);
        int kernel = kGather;
        if (wp::v6_supports(c->precision, c->T, hi) &&
            wp::v6_lds_bytes(c->T, c->C, c->tri, hi) <= lds_hard_limit())
            kernel = kLaneGroup;
        else if (wp::v5_supports(c->precision, c->T, hi) &&
            wp::v5_lds_bytes(c->T, c->C, c->tri, hi) <= lds_hard_limit())
            kernel = kQuadRows;
        else if (wp::v3_supports(c->precision, c->T, hi) &&
                 wp::v3_lds_bytes(c->precision, c->T, c->C, c->tri, hi) <= lds_hard_limit())
            kernel = kRowPerLane;
        bins.b[bins.n++] = Bin{lo, hi, ng, kernel, cap_sep, ng_sep, ng_wide};
        lo = hi;
        if (lo >= c->maxbox) break;
    }
    return bins;
}

size_t bin_lds_bytes(const wp_hmm_classifier* c, const Bin& b) {
    switch (b.kernel) {
    case kRowPerLane: return wp::v3_lds_bytes(c->precision, c->T, c->C, c->tri, b.hi);
    case kQuadRows: return wp::v5_lds_bytes(c->T, c->C, c->tri, b.hi);
    case kLaneGroup: return wp::v6_lds_bytes(c->T, c->C, c->tri, b.hi);
    default:
        return std::max(
            wp::classify_lds_bytes(c->precision, c->T, c->C, c->tri, b.hi, c->T * b.hi, b.ngroups_wide, c->G),
            wp::classify_lds_bytes(c->precision, c->T, c->C, c->tri, b.hi, b.cap_sep, b.ngroups_sep, c->G));
    }
}

// Counter block of one call: bin sizes, then the sizes of the two wide lists.
constexpr int kCtrCount = 0;
constexpr int kCtrWide = wp::kMaxBins;
constexpr int kCtrWide2 = 2 * wp::kMaxBins;
constexpr int kCtrLen = 3 * wp::kMaxBins;

void drop_graph(wp_hmm_classifier* c) {
    if (c->gexec) hipGraphExecDestroy(c->gexec);
    c->gexec = nullptr;
    c->gkey.clear();
}

// Lays out the arrays of an N-read workspace from base (null: sizes only);
// returns the bytes it spans.
size_t carve_workspace(const wp_hmm_classifier* c, int N, char* base, wp::Workspace* w) {
    size_t o = 0;
    auto take = [&](size_t bytes) -> char* {
        char* p = base ? base + o : nullptr;
        o += (bytes + 255) & ~(size_t)255;
        return p;
    };
    const size_t n = (size_t)std::max(N, 1);
    const size_t tc = n * c->T * c->C;
    const size_t nbins = (size_t)c->bins.size();
    w->L = (uint8_t*)take(tc);
    w->H = (uint8_t*)take(tc);
    w->nb = (int*)take(sizeof(int) * n);
    w->wide = (uint8_t*)take(n);
    w->lists = (int*)take(sizeof(int) * nbins * n);
    w->wide_lists = (int*)take(sizeof(int) * nbins * n);
    w->wide2_lists = (int*)take(sizeof(int) * nbins * n);
    w->ctr = (int*)take(sizeof(int) * kCtrLen);
    if (c->r64) {
        const size_t nb64 = (size_t)c->r64->bins.size();
        w->tie_n = (int*)take(sizeof(int) * n);
        w->tie_idx = (int*)take(sizeof(int) * n * wp::kTieMax);
        w->rid = (int*)take(sizeof(int) * n);
        w->rls = (double*)take(sizeof(double) * n);
        w->lt = (double*)take(sizeof(double) * n);
        w->rwide = (uint8_t*)take(n);
        w->rlists = (int*)take(sizeof(int) * nb64 * n);
        w->rwide_lists = (int*)take(sizeof(int) * nb64 * n);
        w->rwide2_lists = (int*)take(sizeof(int) * nb64 * n);
        w->rctr = (int*)take(sizeof(int) * kCtrLen);
    }
    return o;
}

int ensure_workspace(wp_hmm_classifier* c, int N) {
    if (N <= c->wcap) return WP_OK;
    drop_graph(c);  // the captured launches hold the old workspace
    hipFree(c->wbuf);
    c->wbuf = nullptr;
    c->ws = wp::Workspace();
    c->wcap = 0;
    c->last_tie_n = nullptr;
    c->last_n = 0;
    const size_t bytes = carve_workspace(c, N, nullptr, &c->ws);
    WP_HIP(hipMalloc(&c->wbuf, bytes));
    carve_workspace(c, N, (char*)c->wbuf, &c->ws);
    c->wcap = N;
    return WP_OK;
}

struct CallArgs {
    int N;
    const double* d_rads;
    const int* d_cand_off;
    const int* d_cand_idx;
    int* d_id;
    double *d_adj, *d_ls, *d_lt, *d_pair;
    int tie_mode = 0;  // ClassifyArgs::tie_mode of the launches
    int* tie_n = nullptr;
    int* tie_idx = nullptr;
    double tie_abs = 0.0, tie_rel = 0.0;
};

wp::ClassifyArgs base_args(const wp_hmm_classifier* c, const CallArgs& x) {
    wp::ClassifyArgs a;
    a.m = c->dm;
    a.N = x.N;
    a.S = c->S;
    a.rads = x.d_rads;
    a.Lr = c->ws.L;
    a.Hr = c->ws.H;
    a.nbmax = c->ws.nb;
    a.wide = c->ws.wide;
    a.wide_pass = 0;
    a.list = nullptr;
    a.list_count = nullptr;
    a.wide_list = nullptr;
    a.wide_count = nullptr;
    a.over_list = nullptr;
    a.over_count = nullptr;
    a.variant = 0;
    a.cap = 0;
    a.grid_cap = kListGrid;
    a.dsD = c->dD;
    a.dsSeq = c->dSeq;
    a.dsTrack = c->dTrack;
    a.dsCount = c->dCount;
    a.dsId = c->dId;
    a.binom = c->dBinom;
    a.tri = c->tri;
    a.cand_off = x.d_cand_off;
    a.cand_idx = x.d_cand_idx;
    a.out_id = x.d_id;
    a.out_adjusted = x.d_adj;
    a.out_log_score = x.d_ls;
    a.out_log_total = x.d_lt;
    a.out_pair_logp = x.d_pair;
    a.tie_mode = x.tie_mode;
    a.tie_n = x.tie_n;
    a.tie_idx = x.tie_idx;
    a.tie_abs = x.tie_abs;
    a.tie_rel = x.tie_rel;
    return a;
}

// Counters cleared, emission ranges and boxes, reads appended to their bins.
void enqueue_prep(wp_hmm_classifier* c, const BinSet& bins, const CallArgs& x, hipStream_t stream) {
    wp::launch_zero_ints(c->ws.ctr, kCtrLen, stream);
    wp::launch_read_prep(c->dm, x.N, x.d_rads, c->ws.L, c->ws.H, c->ws.nb, stream);
    wp::BinEdges e;
    e.nbins = (int)bins.size();
    e.edge[0] = 0;
    for (size_t b = 0; b < bins.size(); b++) e.edge[b + 1] = bins[b].hi;
    wp::launch_bin_reads(x.N, c->ws.nb, e, c->ws.lists, c->ws.ctr + kCtrCount, stream);
}

// The main launch of bin b (for a quad-rows bin, the given variant).
hipError_t launch_bin_main(wp_hmm_classifier* c, const BinSet& bins, int b, const CallArgs& x,
                           int variant, hipStream_t stream) {
    const Bin& bin = bins[b];
    wp::ClassifyArgs a = base_args(c, x);
    a.nb_lo = bin.lo;
    a.nb_hi = bin.hi;
    a.NB = bin.hi;
    a.cap = c->T * bin.hi;
    a.ngroups = bin.ngroups;
    a.list = c->ws.lists + (size_t)b * x.N;
    a.list_count = c->ws.ctr + kCtrCount + b;
    a.wide_list = c->ws.wide_lists + (size_t)b * x.N;
    a.wide_count = c->ws.ctr + kCtrWide + b;
    a.variant = variant;
    // the bins of up to 32 states hold most reads: one workgroup per read
    // (the FP64 decision's lists are short in every bin)
    if (bin.hi <= 32 && x.tie_mode != 2) a.grid_cap = 0;
    if (bin.kernel == kRowPerLane) return wp::launch_classify_v3(a, c->precision, stream);
    if (bin.kernel == kLaneGroup) {
        a.over_list = c->ws.wide2_lists + (size_t)b * x.N;  // (unused by the lane-group bins otherwise)
        a.over_count = c->ws.ctr + kCtrWide2 + b;
        return wp::launch_classify_v6(a, stream);
    }
    if (bin.kernel == kQuadRows) return wp::launch_classify_v5(a, stream);
    a.variant = 1;  // gather bins: separable Bleach passes; reads that outgrow them go to the wide pass
    a.cap = bin.cap_sep;
    a.ngroups = bin.ngroups_sep;
    return wp::launch_classify(a, c->precision, c->G, stream);
}

// Whether a lane-group bin's main launch caps its W block (its over list can
// hold reads): 16-state boxes in the first variant.
bool lane_group_capped(const Bin& bin, int variant) { return bin.kernel == kLaneGroup && bin.hi == 16 && !variant; }

// The uncapped pass of a capped lane-group bin: the reads whose box classes
// outgrew the main launch's W block, scored by the form that builds each
// timestep's matrices just before its step (reads whose densities are too
// wide still go on to the wide pass).
hipError_t launch_bin_over(wp_hmm_classifier* c, const BinSet& bins, int b, const CallArgs& x, hipStream_t stream) {
    const Bin& bin = bins[b];
    wp::ClassifyArgs a = base_args(c, x);
    a.nb_lo = bin.lo;
    a.nb_hi = bin.hi;
    a.NB = bin.hi;
    a.cap = c->T * bin.hi;
    a.ngroups = bin.ngroups;
    a.list = c->ws.wide2_lists + (size_t)b * x.N;
    a.list_count = c->ws.ctr + kCtrWide2 + b;
    a.wide_list = c->ws.wide_lists + (size_t)b * x.N;
    a.wide_count = c->ws.ctr + kCtrWide + b;
    a.variant = 1;
    if (x.tie_mode != 2) a.grid_cap = 0;
    return wp::launch_classify_v6(a, stream);
}

// The wide pass of a bin: the reads its main launch handed over. Quad-rows
// bins: densities too wide for one exponent per timestep, scored by the
// per-entry exponent kernels. Gather bins: boxes too large for the
// half-size separable launch; stage 1 runs the separable passes at the
// bin's full capacity, stage 2 the composite form on what still does not fit.
hipError_t launch_bin_wide(wp_hmm_classifier* c, const BinSet& bins, int b, const CallArgs& x,
                           int stage, hipStream_t stream) {
    const Bin& bin = bins[b];
    wp::ClassifyArgs a = base_args(c, x);
    a.nb_lo = bin.lo;
    a.nb_hi = bin.hi;
    a.NB = bin.hi;
    a.cap = c->T * bin.hi;
    a.ngroups = bin.ngroups;
    a.wide_pass = 1;
    a.ngroups = bin.ngroups_wide;  // few reads: every group the LDS holds works on each
    a.list = (stage == 2 ? c->ws.wide2_lists : c->ws.wide_lists) + (size_t)b * x.N;
    a.list_count = c->ws.ctr + (stage == 2 ? kCtrWide2 : kCtrWide) + b;
    if (bin.kernel == kGather && stage == 1) {
        a.variant = 1;
        a.wide_list = c->ws.wide2_lists + (size_t)b * x.N;
        a.wide_count = c->ws.ctr + kCtrWide2 + b;
        return wp::launch_classify(a, c->precision, c->G, stream);
    }
    if (wp::v3_supports(c->precision, c->T, bin.hi) &&
        wp::v3_lds_bytes(c->precision, c->T, c->C, c->tri, bin.hi) <= lds_hard_limit())
        return wp::launch_classify_v3(a, c->precision, stream);
    return wp::launch_classify(a, c->precision, c->G, stream);
}

// Every bin's launches of classifier s (its main launch, then its wide
// passes; s->variant for quad-rows bins) onto up to three branches: the
// caller's stream and two streams of c, forked from and joined back into
// `stream` by events (parallel branches of the graph under capture). Once the
// reads are binned, the bins are independent: every read's outputs, wide
// lists and counters belong to one bin. Bins are taken from the largest boxes
// down, branch = position mod 3: the bins of large boxes hold few reads whose
// workgroups run long, so they start first and overlap the dense bins.
hipError_t launch_bins_branched(wp_hmm_classifier* c, wp_hmm_classifier* s, const CallArgs& x, hipStream_t stream,
                                bool branches) {
    const BinSet& bins = s->bins;
    constexpr int kBranches = 3;
    const int nbr = branches ? std::min(kBranches, (int)bins.size()) : 1;
    hipError_t e = hipSuccess;
    if (nbr > 1) {
        if (!c->fork) e = hipEventCreateWithFlags(&c->fork, hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventRecord(c->fork, stream);
        for (int k = 0; k + 1 < nbr && e == hipSuccess; k++) {
            if (!c->branch[k]) e = hipStreamCreateWithFlags(&c->branch[k], hipStreamNonBlocking);
            if (e == hipSuccess && !c->join[k]) e = hipEventCreateWithFlags(&c->join[k], hipEventDisableTiming);
            if (e == hipSuccess) e = hipStreamWaitEvent(c->branch[k], c->fork, 0);
        }
    }
    for (int pos = 0; pos < (int)bins.size() && e == hipSuccess; pos++) {
        const int b = (int)bins.size() - 1 - pos;
        hipStream_t st = pos % nbr == 0 ? stream : c->branch[pos % nbr - 1];
        e = launch_bin_main(s, bins, b, x, s->variant[b], st);
        if (e == hipSuccess && lane_group_capped(bins[b], s->variant[b])) e = launch_bin_over(s, bins, b, x, st);
        if (e == hipSuccess && bins[b].kernel != kRowPerLane) e = launch_bin_wide(s, bins, b, x, 1, st);
        if (e == hipSuccess && bins[b].kernel == kGather) e = launch_bin_wide(s, bins, b, x, 2, st);
    }
    for (int k = 0; k + 1 < nbr && e == hipSuccess; k++) {
        e = hipEventRecord(c->join[k], c->branch[k]);
        if (e == hipSuccess) e = hipStreamWaitEvent(stream, c->join[k], 0);
    }
    return e;
}

// FP32 calls: the reads whose near-best list holds two or more candidates are
// rescored over that list by the FP64 sibling's kernels (on the boxes this
// call's prep computed), and their id and ln score are replaced by the FP64
// decision.
int enqueue_refine(wp_hmm_classifier* c, const CallArgs& x, hipStream_t stream, bool branches) {
    wp_hmm_classifier* s = c->r64;
    s->ws = wp::Workspace();
    s->ws.L = c->ws.L;
    s->ws.H = c->ws.H;
    s->ws.nb = c->ws.nb;
    s->ws.wide = c->ws.rwide;
    s->ws.lists = c->ws.rlists;
    s->ws.wide_lists = c->ws.rwide_lists;
    s->ws.wide2_lists = c->ws.rwide2_lists;
    s->ws.ctr = c->ws.rctr;
    const BinSet& bins = s->bins;
    wp::launch_zero_ints(s->ws.ctr, kCtrLen, stream);
    wp::BinEdges e;
    e.nbins = (int)bins.size();
    e.edge[0] = 0;
    for (size_t b = 0; b < bins.size(); b++) e.edge[b + 1] = bins[b].hi;
    wp::launch_bin_refine(x.N, c->ws.nb, c->ws.tie_n, e, s->ws.lists, s->ws.ctr + kCtrCount, stream);
    CallArgs y{x.N, x.d_rads, x.d_cand_off, x.d_cand_idx, c->ws.rid, nullptr, c->ws.rls, nullptr, nullptr};
    y.tie_mode = 2;
    y.tie_n = c->ws.tie_n;
    y.tie_idx = c->ws.tie_idx;
    WP_HIP(launch_bins_branched(c, s, y, stream, branches));
    wp::launch_refine_fix(x.N, c->ws.tie_n, c->ws.rid, c->ws.rls, x.d_lt, x.d_id, x.d_adj, x.d_ls, stream);
    WP_HIP(hipGetLastError());
    return WP_OK;
}

// The whole call: prep, binning, each bin's launch (and wide pass).
int enqueue_call(wp_hmm_classifier* c, const CallArgs& x, hipStream_t stream, bool events,
                 bool branches = true) {
    const BinSet& bins = c->bins;
    if (events) WP_HIP(hipEventRecord(c->ev[0], stream));
    enqueue_prep(c, bins, x, stream);
    WP_HIP(hipGetLastError());
    if (events) WP_HIP(hipEventRecord(c->ev[1], stream));
    WP_HIP(launch_bins_branched(c, c, x, stream, branches));
    if (x.tie_mode == 1) {
        const int rc = enqueue_refine(c, x, stream, branches);
        if (rc) return rc;
    }
    if (events) {
        WP_HIP(hipEventRecord(c->ev[2], stream));
        c->ev_recorded = 1;
    }
    return WP_OK;
}

// Calls with fewer reads use the default variants (bin-wide boxes for bins of
// up to 4 states, per-timestep classes above, as measured on the 10^5-read
// benchmark) and leave the tuning to the first call with at least this many.
constexpr int kTuneMinReads = 20000;

// Default near-best window of the FP32 path, W(best) = abs + rel |best| in
// ln P: twice the FP32 error bound eps(x) = 1e-5 + 5e-6 |x| (measured errors
// on the 10^5 x 1,000 and 20,000 dye-seq workloads stay below a quarter of it).
constexpr double kDefaultTieAbs = 2e-5;
constexpr double kDefaultTieRel = 1e-5;

// Picks each quad-rows bin's variant by timing both on this call's reads
// (two runs each, the faster second run wins; both give identical outputs).
int autotune(wp_hmm_classifier* c, const CallArgs& x, hipStream_t stream) {
    const BinSet& bins = c->bins;
    bool any = false;
    for (const Bin& b : bins) any = any || b.kernel == kQuadRows;
    if (!any || c->variant_mode >= 0) {
        for (int b = 0; b < wp::kMaxBins; b++) c->variant[b] = c->variant_mode > 0 ? 1 : 0;
        c->tuned = 1;
        return WP_OK;
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    WP_HIP(hipEventCreate(&e0));
    WP_HIP(hipEventCreate(&e1));
    int rc = WP_OK;
    auto fail = [&](hipError_t e) {
        wp::set_error(std::string("autotune: ") + hipGetErrorString(e));
        rc = WP_ERR_HIP;
    };
    enqueue_prep(c, bins, x, stream);
    for (int b = 0; b < (int)bins.size() && !rc; b++) {
        c->variant[b] = 0;
        if (bins[b].kernel != kQuadRows) continue;
        float best[2] = {1e30f, 1e30f};
        for (int trial = 0; trial < 4 && !rc; trial++) {
            const int v = trial & 1;
            // a fresh wide list for each trial
            wp::launch_zero_ints(c->ws.ctr + kCtrWide + b, 1, stream);
            hipError_t e = hipEventRecord(e0, stream);
            if (e == hipSuccess) e = launch_bin_main(c, bins, b, x, v, stream);
            if (e == hipSuccess) e = hipEventRecord(e1, stream);
            if (e == hipSuccess) e = hipEventSynchronize(e1);
            float ms = 0.f;
            if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
            if (e != hipSuccess) {
                fail(e);
                break;
            }
            if (trial >= 2) best[v] = std::min(best[v], ms);
        }
        c->tune_ms[b][0] = best[0];
        c->tune_ms[b][1] = best[1];
        c->variant[b] = best[1] < best[0] ? 1 : 0;
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (!rc) c->tuned = 1;
    return rc;
}

void default_variants(wp_hmm_classifier* c) {
    if (c->variant_mode >= 0) {
        for (int b = 0; b < wp::kMaxBins; b++) c->variant[b] = c->variant_mode;
    } else {
        for (int b = 0; b < c->bins.size(); b++) c->variant[b] = c->bins[b].hi >= 8 ? 1 : 0;
    }
    c->variants_set = 1;
}

// FP32 near-best settings of a call (the normaliser goes to the workspace
// when the caller does not ask for it: the FP64 decision's adjusted score
// needs it).
void set_tie_args(wp_hmm_classifier* c, CallArgs* x) {
    if (!(c->r64 && c->refine)) return;
    x->tie_mode = 1;
    x->tie_n = c->ws.tie_n;
    x->tie_idx = c->ws.tie_idx;
    x->tie_abs = c->tie_abs;
    x->tie_rel = c->tie_rel;
    if (!x->d_lt) x->d_lt = c->ws.lt;
}

int run_device(wp_hmm_classifier* c, int N, const double* d_rads, const int* d_cand_off,
               const int* d_cand_idx, int* d_id, double* d_adj, double* d_ls, double* d_lt,
               double* d_pair, hipStream_t stream, bool graph_ok = false) {
    if (N < 0) {
        wp::set_error("num_reads < 0");
        return WP_ERR_ARG;
    }
    if (N == 0) return WP_OK;
    int rc = ensure_workspace(c, N);
    if (rc) return rc;
    CallArgs x{N, d_rads, d_cand_off, d_cand_idx, d_id, d_adj, d_ls, d_lt, d_pair};
    set_tie_args(c, &x);
    c->last_tie_n = x.tie_n;
    c->last_n = N;
    if (!c->tuned) {
        if (c->variant_mode < 0 && N < kTuneMinReads) {
            // too few reads to time the variants reliably: the measured
            // defaults until a large call tunes them
            default_variants(c);
        } else {
            rc = autotune(c, x, stream);
            if (rc) return rc;
            c->variants_set = 1;
        }
    }
    if (!graph_ok || !c->use_graph || c->timing) return enqueue_call(c, x, stream, c->timing != 0, c->branches != 0);
    std::vector<const void*> key = {(const void*)(intptr_t)N, d_rads, d_cand_off, d_cand_idx, d_id, d_adj,
                                    d_ls, d_lt, d_pair, (const void*)(intptr_t)x.tie_mode};
    for (int b = 0; b < wp::kMaxBins; b++) key.push_back((const void*)(intptr_t)c->variant[b]);
    if (!c->gexec || key != c->gkey) {
        drop_graph(c);
        if (!c->cap_stream) WP_HIP(hipStreamCreateWithFlags(&c->cap_stream, hipStreamNonBlocking));
        WP_HIP(hipStreamBeginCapture(c->cap_stream, hipStreamCaptureModeRelaxed));
        rc = enqueue_call(c, x, c->cap_stream, false, c->branches != 0);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(c->cap_stream, &g);
        if (rc) {
            if (g) hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) {
            wp::set_error(std::string("graph capture: ") + hipGetErrorString(e));
            return WP_ERR_HIP;
        }
        e = hipGraphInstantiate(&c->gexec, g, nullptr, nullptr, 0);
        hipGraphDestroy(g);
        if (e != hipSuccess) {
            c->gexec = nullptr;
            wp::set_error(std::string("graph instantiate: ") + hipGetErrorString(e));
            return WP_ERR_HIP;
        }
        c->gkey = key;
    }
    WP_HIP(hipGraphLaunch(c->gexec, stream));
    return WP_OK;
}

template <typename T>
int upload(T** d, const T* h, size_t n) {
    WP_HIP(hipMalloc((void**)d, sizeof(T) * std::max<size_t>(n, 1)));
    if (n) WP_HIP(hipMemcpy(*d, h, sizeof(T) * n, hipMemcpyHostToDevice));
    return WP_OK;
}

}  // namespace

extern "C" {

const char* wp_last_error(void) { return wp::g_err.c_str(); }

void wp_model_with_mu_as_one(const wp_model* in, wp_model* out) {
    wp_model m = *in;
    for (int c = 0; c < in->num_channels && c < WP_MAX_CHANNELS; c++) {
        double mu = in->channels[c].mu;
        m.channels[c].bg_sig = in->channels[c].bg_sig / mu;
        m.channels[c].mu = 1.0;
        m.channels[c].sig = in->channels[c].sig / mu;
    }
    *out = m;
}

int wp_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    int good = 0;
    for (int d = 0; d < n; d++) {
        hipDeviceProp_t p;
        if (hipGetDeviceProperties(&p, d) == hipSuccess && std::strncmp(p.gcnArchName, "gfx950", 6) == 0)
            good++;
    }
    return good;
}

int wp_hmm_classifier_create(unsigned int num_timesteps, unsigned int num_channels,
                             const wp_model* seq_model, double dist_cutoff, int num_dye_seqs,
                             const char* const* dye_strings, const int* dye_seq_counts,
                             const int* dye_seq_ids, int precision, wp_hmm_classifier** out) {
    if (!out || !seq_model || (num_dye_seqs > 0 && (!dye_strings || !dye_seq_counts || !dye_seq_ids))) {
        wp::set_error("null argument");
        return WP_ERR_ARG;
    }
    *out = nullptr;
    const int T = (int)num_timesteps, C = (int)num_channels;
    if (T < 1 || T > WP_MAX_TIMESTEPS || C < 1 || C > 4 || seq_model->num_channels != C ||
        num_dye_seqs < 0) {
        wp::set_error("unsupported shape: need 1 <= T <= 64, 1 <= C <= 4, model channels == C");
        return WP_ERR_UNSUPPORTED;
    }
    if (precision != WP_PRECISION_FP64 && precision != WP_PRECISION_FP32) {
        wp::set_error("unknown precision");
        return WP_ERR_ARG;
    }
    if (wp_device_count() < 1) {
        wp::set_error("no gfx950 device visible");
        return WP_ERR_NO_DEVICE;
    }
    const int S = num_dye_seqs;
    std::vector<uint8_t> hD((size_t)S * C), hTrack((size_t)S * T * C);
    std::vector<int8_t> hSeq((size_t)S * T);
    int M = 0;
    std::vector<short> seq(4096);
    for (int i = 0; i < S; i++) {
        const char* s = dye_strings[i];
        int len = (int)std::strlen(s);
        if (len > 4096) {
            wp::set_error("dye-seq longer than 4096");
            return WP_ERR_UNSUPPORTED;
        }
        len = wp::parse_dye_seq(s, seq.data(), 4096);
        std::vector<int> cs(C, 0);
        for (int k = 0; k < len; k++) {
            if (seq[k] >= C) {
                wp::set_error("dye-seq names a channel >= num_channels");
                return WP_ERR_ARG;
            }
        }
        // DyeTrack (dye-track.cc:26-43): dyes of each channel at positions >= e.
        for (int e = len - 1; e >= 0; e--) {
            if (seq[e] >= 0) cs[seq[e]]++;
            if (e < T)
                for (int c = 0; c < C; c++) hTrack[((size_t)i * T + e) * C + c] = (uint8_t)cs[c];
        }
        for (int c = 0; c < C; c++) {
            if (cs[c] > WP_MAX_DYES) {
                wp::set_error("more than 15 dyes of one channel in a dye-seq");
                return WP_ERR_UNSUPPORTED;
            }
            hD[(size_t)i * C + c] = (uint8_t)cs[c];
            M = std::max(M, cs[c]);
        }
        for (int e = 0; e < T; e++) hSeq[(size_t)i * T + e] = (int8_t)(e < len ? seq[e] : -1);
    }
    wp_hmm_classifier* c = new wp_hmm_classifier();
    c->T = T;
    c->C = C;
    c->S = S;
    c->M = M;
    c->precision = precision;
    c->G = wp::kGroupLanes;
    c->tri = (M + 1) * (M + 2) / 2;
    c->ids.assign(dye_seq_ids, dye_seq_ids + S);
    c->counts.assign(dye_seq_counts, dye_seq_counts + S);
    c->maxbox = 1;
    for (int k = 0; k < C; k++) c->maxbox *= (M + 1);
    wp::fill_dev_model(*seq_model, T, C, M, dist_cutoff, &c->dm);
    std::vector<double> binom((size_t)2 * C * c->tri);
    wp::binomial_tables(*seq_model, C, M, binom.data());
    // every bin must fit at least one pair group
    c->bins = make_bins(c);
    for (const Bin& b : c->bins) {
        if (bin_lds_bytes(c, b) > lds_hard_limit()) {
            wp::set_error("state box too large for LDS at this precision (max_num_dyes/channels)");
            delete c;
            return WP_ERR_UNSUPPORTED;
        }
    }
    int rc = WP_OK;
    if (!rc) rc = upload(&c->dD, hD.data(), hD.size());
    if (!rc) rc = upload(&c->dSeq, hSeq.data(), hSeq.size());
    if (!rc) rc = upload(&c->dTrack, hTrack.data(), hTrack.size());
    if (!rc) rc = upload(&c->dCount, dye_seq_counts, (size_t)S);
    if (!rc) rc = upload(&c->dId, dye_seq_ids, (size_t)S);
    if (!rc) rc = upload(&c->dBinom, binom.data(), binom.size());
    if (rc) {
        wp_hmm_classifier_destroy(c);
        return rc;
    }
    if (precision == WP_PRECISION_FP32) {
        // the FP64 rescoring pass of near-best candidates: the same tables,
        // FP64 bins (left out if one of them does not fit in LDS)
        wp_hmm_classifier* s = new wp_hmm_classifier();
        s->T = T;
        s->C = C;
        s->S = S;
        s->M = M;
        s->precision = WP_PRECISION_FP64;
        s->G = c->G;
        s->tri = c->tri;
        s->dm = c->dm;
        s->maxbox = c->maxbox;
        s->dD = c->dD;
        s->dSeq = c->dSeq;
        s->dTrack = c->dTrack;
        s->dCount = c->dCount;
        s->dId = c->dId;
        s->dBinom = c->dBinom;
        s->borrowed = 1;
        bool fits = true;
        s->bins = make_bins(s);
        for (const Bin& b : s->bins) fits = fits && bin_lds_bytes(s, b) <= lds_hard_limit();
        if (fits) {
            c->r64 = s;
            c->refine = 1;
            c->tie_abs = kDefaultTieAbs;
            c->tie_rel = kDefaultTieRel;
        } else {
            delete s;
        }
    }
    *out = c;
    return WP_OK;
}

void wp_hmm_classifier_destroy(wp_hmm_classifier* c) {
    if (!c) return;
    for (int k = 0; k < 3; k++)
        if (c->ev[k]) hipEventDestroy(c->ev[k]);
    drop_graph(c);
    if (c->cap_stream) hipStreamDestroy(c->cap_stream);
    for (int k = 0; k < 2; k++) {
        if (c->branch[k]) hipStreamDestroy(c->branch[k]);
        if (c->join[k]) hipEventDestroy(c->join[k]);
    }
    if (c->fork) hipEventDestroy(c->fork);
    if (c->r64) wp_hmm_classifier_destroy(c->r64);
    hipFree(c->wbuf);
    for (void* p : {(void*)c->hs.rads, (void*)c->hs.id, (void*)c->hs.out, (void*)c->hs.pair, (void*)c->hs.off,
                    (void*)c->hs.idx})
        hipFree(p);
    if (!c->borrowed) {
        hipFree(c->dD);
        hipFree(c->dSeq);
        hipFree(c->dTrack);
        hipFree(c->dCount);
        hipFree(c->dId);
        hipFree(c->dBinom);
    }
    delete c;
}

int wp_hmm_classifier_max_num_dyes(const wp_hmm_classifier* c) { return c ? c->M : -1; }

int wp_hmm_classifier_num_dye_seqs(const wp_hmm_classifier* c) { return c ? c->S : -1; }

int wp_hmm_classify_candidates_device(wp_hmm_classifier* c, int N, const double* d_rads, const int* d_cand_off,
                                      const int* d_cand_idx, int* d_id, double* d_adj, double* d_ls, double* d_lt,
                                      void* stream) {
    if (!c || (N > 0 && (!d_rads || !d_id || !d_cand_off || !d_cand_idx))) {
        wp::set_error("null argument");
        return WP_ERR_ARG;
    }
    return run_device(c, N, d_rads, d_cand_off, d_cand_idx, d_id, d_adj, d_ls, d_lt, nullptr, (hipStream_t)stream,
                      true);
}

int wp_hmm_classify_device(wp_hmm_classifier* c, int N, const double* d_rads, int* d_id, double* d_adj,
                           double* d_ls, double* d_lt, void* stream) {
    if (!c || (N > 0 && (!d_rads || !d_id))) {
        wp::set_error("null argument");
        return WP_ERR_ARG;
    }
    return run_device(c, N, d_rads, nullptr, nullptr, d_id, d_adj, d_ls, d_lt, nullptr,
                      (hipStream_t)stream, true);
}

}  // extern "C"

// Workspace of wp_hmm_classify_device_noalloc: a file-scope device array, so
// that the entry point allocates nothing. Each device holds its own copy until
// the process ends. Host code addresses it only through the pointer
// hipGetSymbolAddress returns for the current device (the array's name is a
// host-side stand-in).
static constexpr size_t kScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxScratchDevices] = {nullptr};

// The current device's scratch address, looked up once per device.
int scratch_for_current_device(void** out) {
    int dev = 0;
    WP_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxScratchDevices) {
        wp::set_error("device index beyond the scratch table");
        return WP_ERR_UNSUPPORTED;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        WP_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return WP_OK;
}

}  // namespace

extern "C" {

int wp_hmm_classify_device_noalloc(wp_hmm_classifier* c, int N, const double* d_rads, int* d_id, double* d_adj,
                                   double* d_ls, double* d_lt, void* stream) {
    if (!c || N < 0 || (N > 0 && (!d_rads || !d_id))) {
        wp::set_error("null argument or num_reads < 0");
        return WP_ERR_ARG;
    }
    if (N == 0) return WP_OK;
    wp::Workspace ws;
    const size_t need = carve_workspace(c, N, nullptr, &ws);
    if (need > kScratchBytes) {
        wp::set_error("wp_hmm_classify_device_noalloc: " + std::to_string(N) + " reads need " +
                      std::to_string(need) + " bytes of workspace, more than its " +
                      std::to_string(kScratchBytes) + "-byte scratch");
        return WP_ERR_ARG;
    }
    void* base = nullptr;
    int rc = scratch_for_current_device(&base);
    if (rc) return rc;
    carve_workspace(c, N, (char*)base, &ws);
    if (!c->variants_set) default_variants(c);  // no autotuning here: it times launches with events
    // the call's launches read the classifier's workspace pointers when they
    // are enqueued; the classifier's own workspace is put back afterwards
    const wp::Workspace own = c->ws;
    c->ws = ws;
    CallArgs x{N, d_rads, nullptr, nullptr, d_id, d_adj, d_ls, d_lt, nullptr};
    set_tie_args(c, &x);
    rc = enqueue_call(c, x, (hipStream_t)stream, false, false);
    c->ws = own;
    c->last_tie_n = x.tie_n;
    c->last_n = N;
    return rc;
}

static int classify_host(wp_hmm_classifier* c, int N, const double* rads, const int* cand_off,
                         const int* cand_idx, int* out_id, double* out_adj, double* out_ls,
                         double* out_lt, double* out_pair) {
    if (!c || N < 0 || (N > 0 && !rads)) {
        wp::set_error("null argument");
        return WP_ERR_ARG;
    }
    if (N == 0) return WP_OK;
    if (cand_off) {
        for (int r = 0; r < N; r++) {
            if (cand_off[r + 1] < cand_off[r]) {
                wp::set_error("cand_offsets not monotone");
                return WP_ERR_ARG;
            }
        }
        for (int k = cand_off[0]; k < cand_off[N]; k++) {
            if (cand_idx[k] < 0 || cand_idx[k] >= c->S) {
                wp::set_error("candidate index out of range");
                return WP_ERR_ARG;
            }
        }
    }
    size_t tc = (size_t)N * c->T * c->C;
    HostStage& hs = c->hs;
    int rc = WP_OK;
    hipStream_t s = nullptr;
    auto cleanup = [&]() {};
#define WP_HIPC(expr)                                                              \
    do {                                                                           \
        hipError_t e_ = (expr);                                                    \
        if (e_ != hipSuccess) {                                                    \
            wp::set_error(std::string(#expr) + ": " + hipGetErrorString(e_));      \
            cleanup();                                                             \
            return WP_ERR_HIP;                                                     \
        }                                                                          \
    } while (0)
    // grow-only staging buffers (freed with the classifier)
    auto grow = [&](auto** p, size_t* have, size_t need, size_t elem) -> hipError_t {
        if (need <= *have) return hipSuccess;
        hipFree(*p);
        *p = nullptr;
        *have = 0;
        hipError_t e = hipMalloc((void**)p, elem * std::max<size_t>(need, 1));
        if (e == hipSuccess) *have = need;
        return e;
    };
    {
        size_t nr = hs.n_reads;  // id and out share the reads capacity
        WP_HIPC(grow(&hs.rads, &hs.n_rads, tc, sizeof(double)));
        if ((size_t)N > nr) {
            size_t n1 = nr, n2 = nr, n3 = nr;
            WP_HIPC(grow(&hs.id, &n1, (size_t)N, sizeof(int)));
            WP_HIPC(grow(&hs.out, &n2, 3 * (size_t)N, sizeof(double)));
            WP_HIPC(grow(&hs.off, &n3, (size_t)N + 1, sizeof(int)));
            hs.n_reads = (size_t)N;
        }
    }
    if (out_pair) WP_HIPC(grow(&hs.pair, &hs.n_pair, (size_t)N * std::max(c->S, 1), sizeof(double)));
    double* d_rads = hs.rads;
    int* d_id = hs.id;
    double* d_out = hs.out;
    double* d_pair = out_pair ? hs.pair : nullptr;
    int* d_off = nullptr;
    int* d_idx = nullptr;
    if (cand_off) {
        size_t nc = (size_t)(cand_off[N] - cand_off[0]);
        std::vector<int> off(N + 1);
        for (int r = 0; r <= N; r++) off[r] = cand_off[r] - cand_off[0];
        WP_HIPC(grow(&hs.idx, &hs.n_idx, std::max<size_t>(nc, 1), sizeof(int)));
        d_off = hs.off;
        d_idx = hs.idx;
        WP_HIPC(hipMemcpy(d_off, off.data(), sizeof(int) * (N + 1), hipMemcpyHostToDevice));
        if (nc) WP_HIPC(hipMemcpy(d_idx, cand_idx + cand_off[0], sizeof(int) * nc, hipMemcpyHostToDevice));
    }
    WP_HIPC(hipMemcpy(d_rads, rads, sizeof(double) * tc, hipMemcpyHostToDevice));
    rc = run_device(c, N, d_rads, d_off, d_idx, d_id, d_out, d_out + N, d_out + 2 * (size_t)N, d_pair, s);
    if (rc) {
        cleanup();
        return rc;
    }
    WP_HIPC(hipStreamSynchronize(s));
    std::vector<double> hout(3 * (size_t)N);
    if (out_id) WP_HIPC(hipMemcpy(out_id, d_id, sizeof(int) * N, hipMemcpyDeviceToHost));
    WP_HIPC(hipMemcpy(hout.data(), d_out, sizeof(double) * 3 * N, hipMemcpyDeviceToHost));
    if (out_pair) WP_HIPC(hipMemcpy(out_pair, d_pair, sizeof(double) * (size_t)N * c->S, hipMemcpyDeviceToHost));
#undef WP_HIPC
    for (int r = 0; r < N; r++) {
        if (out_adj) out_adj[r] = hout[r];
        if (out_ls) out_ls[r] = hout[N + r];
        if (out_lt) out_lt[r] = hout[2 * (size_t)N + r];
    }
    cleanup();
    return WP_OK;
}

int wp_hmm_classify(wp_hmm_classifier* c, int N, const double* rads, int* out_id, double* out_adj,
                    double* out_ls, double* out_lt) {
    if (!out_id) {
        wp::set_error("out_id is required");
        return WP_ERR_ARG;
    }
    return classify_host(c, N, rads, nullptr, nullptr, out_id, out_adj, out_ls, out_lt, nullptr);
}

int wp_hmm_classify_candidates(wp_hmm_classifier* c, int N, const double* rads, const int* cand_off,
                               const int* cand_idx, int* out_id, double* out_adj, double* out_ls,
                               double* out_lt) {
    if (!out_id || !cand_off || !cand_idx) {
        wp::set_error("null argument");
        return WP_ERR_ARG;
    }
    return classify_host(c, N, rads, cand_off, cand_idx, out_id, out_adj, out_ls, out_lt, nullptr);
}

int wp_hmm_pair_log_probability(wp_hmm_classifier* c, int N, const double* rads, double* out_log_p) {
    if (!out_log_p) {
        wp::set_error("null argument");
        return WP_ERR_ARG;
    }
    std::vector<int> ids(std::max(N, 1));
    return classify_host(c, N, rads, nullptr, nullptr, ids.data(), nullptr, nullptr, nullptr, out_log_p);
}

int wp_hmm_set_variant(wp_hmm_classifier* c, int mode) {
    if (!c || mode < -1 || mode > 1) {
        wp::set_error("wp_hmm_set_variant: mode must be -1, 0 or 1");
        return WP_ERR_ARG;
    }
    c->variant_mode = mode;
    c->tuned = 0;
    c->variants_set = 0;
    for (int b = 0; b < wp::kMaxBins; b++) c->variant[b] = 0;
    for (int b = 0; b < wp::kMaxBins; b++) c->tune_ms[b][0] = c->tune_ms[b][1] = -1.f;
    drop_graph(c);
    return WP_OK;
}

int wp_hmm_get_variants(const wp_hmm_classifier* c, int* out_variant, double* out_ms, int cap) {
    if (!c) {
        wp::set_error("null classifier");
        return WP_ERR_ARG;
    }
    const int nb = c->bins.size();
    for (int b = 0; b < std::min(nb, cap); b++) {
        if (out_variant) out_variant[b] = c->variants_set ? c->variant[b] : -1;
        if (out_ms) {
            out_ms[2 * b] = c->tune_ms[b][0];
            out_ms[2 * b + 1] = c->tune_ms[b][1];
        }
    }
    return nb;
}

int wp_hmm_set_graph(wp_hmm_classifier* c, int enable) {
    if (!c) {
        wp::set_error("null classifier");
        return WP_ERR_ARG;
    }
    c->use_graph = enable ? 1 : 0;
    drop_graph(c);
    return WP_OK;
}

int wp_hmm_set_branches(wp_hmm_classifier* c, int enable) {
    if (!c) {
        wp::set_error("null classifier");
        return WP_ERR_ARG;
    }
    c->branches = enable ? 1 : 0;
    drop_graph(c);
    return WP_OK;
}

int wp_hmm_set_refine(wp_hmm_classifier* c, int enable, double abs, double rel) {
    if (!c || !(abs >= 0.0) || !(rel >= 0.0)) {
        wp::set_error("wp_hmm_set_refine: null classifier or negative window");
        return WP_ERR_ARG;
    }
    if (!c->r64) {
        wp::set_error("no FP64 decision for this classifier (FP64 itself, or FP64 boxes too large for LDS)");
        return WP_ERR_UNSUPPORTED;
    }
    c->refine = enable ? 1 : 0;
    c->tie_abs = abs;
    c->tie_rel = rel;
    drop_graph(c);
    return WP_OK;
}

int wp_hmm_refine_stats(wp_hmm_classifier* c, int N, int* out2) {
    if (!c || !out2 || N < 0 || N > c->last_n || (N > 0 && !c->last_tie_n)) {
        wp::set_error("wp_hmm_refine_stats: null argument, no FP64 decision, or more reads than the last call");
        return WP_ERR_ARG;
    }
    out2[0] = out2[1] = 0;
    if (N == 0) return WP_OK;
    std::vector<int> h(N);
    WP_HIP(hipDeviceSynchronize());
    WP_HIP(hipMemcpy(h.data(), c->last_tie_n, sizeof(int) * N, hipMemcpyDeviceToHost));
    for (int r = 0; r < N; r++) {
        if (h[r] != 0) out2[0]++;
        if (h[r] < 0) out2[1]++;
    }
    return WP_OK;
}

int wp_hmm_set_timing(wp_hmm_classifier* c, int enable) {
    if (!c) {
        wp::set_error("null classifier");
        return WP_ERR_ARG;
    }
    if (enable && !c->ev[0]) {
        for (int k = 0; k < 3; k++) WP_HIP(hipEventCreate(&c->ev[k]));
    }
    c->timing = enable ? 1 : 0;
    c->ev_recorded = 0;
    return WP_OK;
}

int wp_hmm_last_timing(wp_hmm_classifier* c, double* prep_ms, double* classify_ms) {
    if (!c || !c->timing || !c->ev_recorded) {
        wp::set_error("timing not enabled or nothing recorded");
        return WP_ERR_ARG;
    }
    WP_HIP(hipEventSynchronize(c->ev[2]));
    float a = 0.f, b = 0.f;
    WP_HIP(hipEventElapsedTime(&a, c->ev[0], c->ev[1]));
    WP_HIP(hipEventElapsedTime(&b, c->ev[1], c->ev[2]));
    if (prep_ms) *prep_ms = a;
    if (classify_ms) *classify_ms = b;
    return WP_OK;
}

int wp_hmm_work(wp_hmm_classifier* c, int N, const double* rads, double* out3) {
    if (!c || N < 0 || (N > 0 && (!rads || !out3))) {
        wp::set_error("null argument");
        return WP_ERR_ARG;
    }
    out3[0] = out3[1] = out3[2] = 0.0;
    if (N == 0) return WP_OK;
    int rc = ensure_workspace(c, N);
    if (rc) return rc;
    double* d_rads = nullptr;
    double* d_out = nullptr;
    size_t tc = (size_t)N * c->T * c->C;
    hipError_t e = hipMalloc(&d_rads, sizeof(double) * tc);
    if (e == hipSuccess) e = hipMalloc(&d_out, sizeof(double) * 3 * (size_t)N);
    if (e == hipSuccess) e = hipMemcpy(d_rads, rads, sizeof(double) * tc, hipMemcpyHostToDevice);
    std::vector<double> h(3 * (size_t)N);
    if (e == hipSuccess) {
        wp::launch_read_prep(c->dm, N, d_rads, c->ws.L, c->ws.H, c->ws.nb, nullptr);
        wp::launch_work(c->dm, N, c->S, c->ws.L, c->ws.H, c->ws.nb, c->dD, d_out, nullptr);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(h.data(), d_out, sizeof(double) * 3 * N, hipMemcpyDeviceToHost);
    hipFree(d_rads);
    hipFree(d_out);
    if (e != hipSuccess) {
        wp::set_error(std::string("wp_hmm_work: ") + hipGetErrorString(e));
        return WP_ERR_HIP;
    }
    for (int r = 0; r < N; r++)
        for (int k = 0; k < 3; k++) out3[k] += h[(size_t)r * 3 + k];
    return WP_OK;
}

int wp_generate_dye_seqs(int num_channels, int num_dye_seqs, int len_min, int len_max, int dyes_min,
                         int dyes_max, uint64_t seed, char* out_strings, int stride) {
    if (num_channels < 1 || num_channels > 10 || num_dye_seqs < 0 || len_min < 1 || len_max < len_min ||
        dyes_min < 0 || dyes_max < dyes_min || !out_strings || stride < len_max + 1) {
        wp::set_error("bad dye-seq generator arguments");
        return WP_ERR_ARG;
    }
    wp::generate_dye_seqs(num_channels, num_dye_seqs, len_min, len_max, dyes_min, dyes_max, seed,
                          out_strings, stride);
    return WP_OK;
}

int wp_simulate_radiometries(const wp_model* seq_model, unsigned int num_timesteps,
                             unsigned int num_channels, int num_dye_seqs,
                             const char* const* dye_strings, const int* weights, int num_reads,
                             uint64_t seed, double* out_radiometries, int* out_dye_seq_index) {
    if (!seq_model || num_dye_seqs < 1 || !dye_strings || num_reads < 0 || !out_radiometries ||
        num_timesteps < 1 || num_timesteps > WP_MAX_TIMESTEPS || num_channels < 1 ||
        num_channels > WP_MAX_CHANNELS) {
        wp::set_error("bad simulator arguments");
        return WP_ERR_ARG;
    }
    int cap = 0;
    for (int i = 0; i < num_dye_seqs; i++) cap = std::max(cap, (int)std::strlen(dye_strings[i]));
    cap = std::max(cap, 1);
    std::vector<short> seqs((size_t)num_dye_seqs * cap, -1);
    std::vector<int> lens(num_dye_seqs);
    for (int i = 0; i < num_dye_seqs; i++) lens[i] = wp::parse_dye_seq(dye_strings[i], &seqs[(size_t)i * cap], cap);
    if (!wp::dye_channels_ok(seqs.data(), num_dye_seqs, cap, lens.data(), (int)num_channels)) return WP_ERR_ARG;
    wp::simulate_radiometries(*seq_model, (int)num_timesteps, (int)num_channels, num_dye_seqs, seqs.data(), cap,
                              lens.data(), weights, num_reads, seed, out_radiometries, out_dye_seq_index,
                              nullptr);
    return WP_OK;
}

int wp_simulate_radiometries_gpu(const wp_model* seq_model, unsigned int num_timesteps,
                                 unsigned int num_channels, int num_dye_seqs,
                                 const char* const* dye_strings, const int* weights, int num_reads,
                                 uint64_t seed, double* out_radiometries, int* out_dye_seq_index) {
    if (!seq_model || num_dye_seqs < 1 || !dye_strings || num_reads < 0 || !out_radiometries ||
        num_timesteps < 1 || num_timesteps > WP_MAX_TIMESTEPS || num_channels < 1 || num_channels > 4) {
        wp::set_error("bad simulator arguments");
        return WP_ERR_ARG;
    }
    int rc = wp::require_device();
    if (rc) return rc;
    int cap = 0;
    for (int i = 0; i < num_dye_seqs; i++) cap = std::max(cap, (int)std::strlen(dye_strings[i]));
    cap = std::max(cap, 1);
    std::vector<short> seqs((size_t)num_dye_seqs * cap, -1);
    std::vector<int> lens(num_dye_seqs);
    for (int i = 0; i < num_dye_seqs; i++) lens[i] = wp::parse_dye_seq(dye_strings[i], &seqs[(size_t)i * cap], cap);
    if (!wp::dye_channels_ok(seqs.data(), num_dye_seqs, cap, lens.data(), (int)num_channels)) return WP_ERR_ARG;
    wp::DevModel dm;
    wp::fill_dev_model(*seq_model, (int)num_timesteps, (int)num_channels, 1, DBL_MAX, &dm);
    return wp::gpu_simulate_radiometries(*seq_model, dm, (int)num_timesteps, (int)num_channels, num_dye_seqs,
                                         seqs.data(), cap, lens.data(), weights, num_reads, seed, out_radiometries,
                                         out_dye_seq_index);
}

int wp_simulate_dye_tracks_gpu(const wp_model* seq_model, unsigned int num_timesteps, unsigned int num_channels,
                               int num_dye_seqs, const char* const* dye_strings, const int* dye_seq_counts,
                               const int* dye_seq_ids, int per_peptide, uint64_t seed, int max_tracks,
                               int max_sources, int* out_num_tracks, int* out_num_sources,
                               short* out_track_counts, int* out_source_offsets,
                               int* out_source_ids, int* out_source_counts, int* out_source_hits) {
    if (!seq_model || num_dye_seqs < 1 || !dye_strings || !dye_seq_counts || !dye_seq_ids || per_peptide < 0 ||
        !out_num_tracks || !out_num_sources || num_timesteps < 1 || num_timesteps > WP_MAX_TIMESTEPS || num_channels < 1 ||
        num_channels > WP_MAX_CHANNELS) {
        wp::set_error("bad simulator arguments");
        return WP_ERR_ARG;
    }
    int rc = wp::require_device();
    if (rc) return rc;
    int cap = 0;
    for (int i = 0; i < num_dye_seqs; i++) cap = std::max(cap, (int)std::strlen(dye_strings[i]));
    cap = std::max(cap, 1);
    std::vector<short> seqs((size_t)num_dye_seqs * cap, -1);
    std::vector<int> lens(num_dye_seqs);
    for (int i = 0; i < num_dye_seqs; i++) lens[i] = wp::parse_dye_seq(dye_strings[i], &seqs[(size_t)i * cap], cap);
    if (!wp::dye_channels_ok(seqs.data(), num_dye_seqs, cap, lens.data(), (int)num_channels)) return WP_ERR_ARG;
    return wp::gpu_simulate_dye_tracks(*seq_model, (int)num_timesteps, (int)num_channels, num_dye_seqs, seqs.data(),
                                       cap, lens.data(), dye_seq_counts, dye_seq_ids, per_peptide, seed, max_tracks,
                                       max_sources, out_num_tracks, out_num_sources, out_track_counts,
                                       out_source_offsets, out_source_ids,
                                       out_source_counts, out_source_hits);
}

int wp_simulate_dye_tracks(const wp_model* seq_model, unsigned int num_timesteps, unsigned int num_channels,
                           int num_dye_seqs, const char* const* dye_strings, const int* dye_seq_counts,
                           const int* dye_seq_ids, int per_peptide, uint64_t seed, int max_tracks,
                           int max_sources, int* out_num_tracks, int* out_num_sources,
                           short* out_track_counts, int* out_source_offsets,
                           int* out_source_ids, int* out_source_counts, int* out_source_hits) {
    if (!seq_model || num_dye_seqs < 1 || !dye_strings || !dye_seq_counts || !dye_seq_ids || per_peptide < 0 ||
        !out_num_tracks || !out_num_sources || num_timesteps < 1 || num_timesteps > WP_MAX_TIMESTEPS || num_channels < 1 ||
        num_channels > WP_MAX_CHANNELS) {
        wp::set_error("bad simulator arguments");
        return WP_ERR_ARG;
    }
    int cap = 0;
    for (int i = 0; i < num_dye_seqs; i++) cap = std::max(cap, (int)std::strlen(dye_strings[i]));
    cap = std::max(cap, 1);
    std::vector<short> seqs((size_t)num_dye_seqs * cap, -1);
    std::vector<int> lens(num_dye_seqs);
    for (int i = 0; i < num_dye_seqs; i++) lens[i] = wp::parse_dye_seq(dye_strings[i], &seqs[(size_t)i * cap], cap);
    if (!wp::dye_channels_ok(seqs.data(), num_dye_seqs, cap, lens.data(), (int)num_channels)) return WP_ERR_ARG;
    std::vector<short> tc;
    std::vector<int> off, id, cnt, hits;
    int n = wp::simulate_dye_tracks(*seq_model, (int)num_timesteps, (int)num_channels, num_dye_seqs, seqs.data(),
                                    cap, lens.data(), dye_seq_counts, dye_seq_ids, per_peptide, seed, &tc, &off, &id,
                                    &cnt, &hits);
    *out_num_tracks = n;
    *out_num_sources = (int)id.size();
    if (n > max_tracks || (int)id.size() > max_sources) {
        wp::set_error("more distinct dye-tracks or sources than max_tracks / max_sources");
        return WP_ERR_ARG;
    }
    if (!out_track_counts || !out_source_offsets || !out_source_ids || !out_source_counts || !out_source_hits) {
        wp::set_error("null output");
        return WP_ERR_ARG;
    }
    std::memcpy(out_track_counts, tc.data(), sizeof(short) * tc.size());
    std::memcpy(out_source_offsets, off.data(), sizeof(int) * off.size());
    std::memcpy(out_source_ids, id.data(), sizeof(int) * id.size());
    std::memcpy(out_source_counts, cnt.data(), sizeof(int) * cnt.size());
    std::memcpy(out_source_hits, hits.data(), sizeof(int) * hits.size());
    return WP_OK;
}

}  // extern "C"
