// Launch interface of the HMM scoring kernels (host side).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "hmm_device.h"

namespace wp {

// Near-best candidates per read handed from the FP32 pass to the FP64 pass.
constexpr int kTieMax = 128;

// Arguments of the fused classify kernel. Device pointers unless noted.
struct ClassifyArgs {
    DevModel m;
    int N;                    // reads
    int S;                    // dye-seqs in the table
    const double* rads;       // [N][T][C], intensities / mu
    const uint8_t* Lr;        // [N][T][C] lower box bound (suffix max of range lows)
    const uint8_t* Hr;        // [N][T][C] upper box bound (prefix min of range highs)
    const int* nbmax;         // [N] max_t prod_c (Hr - Lr); 0 = every pair empty
    int nb_lo, nb_hi;         // this launch takes reads with nb_lo < nbmax <= nb_hi
    uint8_t* wide;            // [N] set by the quad-rows kernel: densities too wide for it
    int wide_pass;            // 1: score only the reads flagged in wide[]
    // Reads of this launch: list[k] for k < *list_count (the list is built on
    // the device); workgroup k scores list[k] (quad-rows and lane-group
    // kernels: later workgroups exit at once) or strides over it (grid_cap).
    const int* list;
    const int* list_count;
    int* wide_list;           // quad-rows: reads handed to the wide pass are appended here
    int* wide_count;
    // lane-group, capped W block: reads whose box classes outgrow the block
    // (and whose densities fit) are appended here for the uncapped pass;
    // null: they go to the wide list
    int* over_list;
    int* over_count;
    int variant;              // quad-rows: 1 = per-timestep box classes, 0 = bin-wide boxes
    int NB;                   // per-group state capacity in count-box entries
    int cap;                  // gather kernel: entries of each of a group's four state buffers
    int ngroups;              // pair groups per workgroup
    // Workgroups of the launch: 0 = one per read of the list (grid a.N);
    // k > 0 = min(a.N, k) workgroups striding over the list (for lists that
    // hold few of the N reads, so the grid does not idle through N exits).
    int grid_cap;
    const uint8_t* dsD;       // [S][C] dyes per channel
    const int8_t* dsSeq;      // [S][T] channel of position e (-1 = none)
    const uint8_t* dsTrack;   // [S][T][C] dyes of channel c at positions >= e
    const int* dsCount;       // [S]
    const int* dsId;          // [S]
    const double* binom;      // [2][C][tri]: dud tables then bleach tables
    int tri;                  // (M+1)(M+2)/2
    const int* cand_off;      // [N+1] or null (every dye-seq)
    const int* cand_idx;      // candidate dye-seq indices
    int* out_id;
    double* out_adjusted;
    double* out_log_score;
    double* out_log_total;
    double* out_pair_logp;    // [N][S] or null
    // FP32 near-best lists: tie_mode 1 (FP32 pass) writes tie_n[r] (0: the
    // FP32 best stands; k >= 2: tie_idx[r][0..k) holds the dye-seqs within
    // the window of the best, in candidate order; -1: rescore every
    // candidate); tie_mode 2 (FP64 pass) scores those lists.
    int tie_mode;
    int* tie_n;               // [N]
    int* tie_idx;             // [N][kTieMax]
    double tie_abs, tie_rel;  // window W(best) = tie_abs + tie_rel |best| in ln P
};

// Workgroups of a launch over a read list (ClassifyArgs::grid_cap).
inline int list_grid(const ClassifyArgs& a) { return a.grid_cap > 0 && a.grid_cap < a.N ? a.grid_cap : a.N; }

// Largest number of read bins (box-size classes) of one classify call.
constexpr int kMaxBins = 16;

// Bin edges of one call: bin b takes reads with edge[b] < nbmax <= edge[b + 1]
// (reads with nbmax == 0, every pair empty, go to bin 0).
struct BinEdges {
    int nbins;
    int edge[kMaxBins + 1];
};

// Clears n ints (a fill kernel, so the sequence stays graph-capturable).
void launch_zero_ints(int* d, int n, hipStream_t stream);

// Appends every read to its bin's list: lists[b * N + k], counts[b].
void launch_bin_reads(int N, const int* d_nbmax, const BinEdges& e, int* d_lists, int* d_counts,
                      hipStream_t stream);

// Appends every read with a near-best list (tie_n != 0) to its bin's list.
void launch_bin_refine(int N, const int* d_nbmax, const int* d_tie_n, const BinEdges& e, int* d_lists,
                       int* d_counts, hipStream_t stream);

// Writes the FP64 decision of every read with a near-best list over the FP32
// outputs: id and ln score from the FP64 pass, adjusted = exp(ln score - ln
// total) with the FP32 normaliser (NaN guard as hmm-classifier.h:73-76).
void launch_refine_fix(int N, const int* d_tie_n, const int* d_rid, const double* d_rls, const double* d_lt,
                       int* d_id, double* d_adj, double* d_ls, hipStream_t stream);

// Per-timestep emission ranges then pruned boxes for every read.
void launch_read_prep(const DevModel& m, int N, const double* d_rads, uint8_t* d_Lr, uint8_t* d_Hr,
                      int* d_nbmax, hipStream_t stream);

// Fused classify over the reads of one box-size bin. precision: WP_PRECISION_*.
// Returns hipSuccess or the launch error.
hipError_t launch_classify(const ClassifyArgs& a, int precision, int group_lanes, hipStream_t stream);

// Row-per-lane classify for reads whose count box is at most NB in {4, 8}
// states, T <= 32.
bool v3_supports(int precision, int T, int NB);
size_t v3_lds_bytes(int precision, int T, int C, int tri, int NB);
hipError_t launch_classify_v3(const ClassifyArgs& a, int precision, hipStream_t stream);

// Quad-rows classify (FP32) for reads whose count box is NB in {4, 8, 16}
// states; T <= 32 (NB <= 8) or T <= 16 (NB = 16).
bool v5_supports(int precision, int T, int NB);
size_t v5_lds_bytes(int T, int C, int tri, int NB);
hipError_t launch_classify_v5(const ClassifyArgs& a, hipStream_t stream);

// Lane-group classify (FP32): 1, 2 or 4 lanes per candidate, one scale per
// candidate, per-timestep box classes; NB in {4, 8, 16, 32}, T <= 12.
bool v6_supports(int precision, int T, int NB);
size_t v6_lds_bytes(int T, int C, int tri, int NB);
hipError_t launch_classify_v6(const ClassifyArgs& a, hipStream_t stream);

// Per-read algorithmic work (non-empty pairs, state-timesteps, FLOPs) -> [N][3].
void launch_work(const DevModel& m, int N, int S, const uint8_t* d_Lr, const uint8_t* d_Hr,
                 const int* d_nbmax, const uint8_t* d_D, double* d_out, hipStream_t stream);

// Bytes of dynamic LDS one gather-kernel workgroup needs (state buffers of cap entries).
size_t classify_lds_bytes(int precision, int T, int C, int tri, int NB, int cap, int ngroups, int group_lanes);

// Chunk of candidates compacted per pass.
constexpr int kCandChunk = 1024;
constexpr int kBlockThreads = 256;
constexpr int kGroupLanes = 16;  // lanes scoring one pair

}  // namespace wp
