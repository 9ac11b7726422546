// kd-tree of the kNN stage: host construction and the GPU search launch.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

namespace wp {

// One node of the flattened tree. Internal nodes have children (left,
// right >= 0), split dimension s and the bounds of InternalNode
// (internal-node.h:22-31); leaves cover points [begin, end) of the permuted
// order (leaf-node.h:21).
struct KdNode {
    int left, right, s, begin, end;
    double max_left, min_right, split;
};

// Builds the tree over n points of dimension d (X row-major by original
// index). perm receives the reference's final point order (position ->
// original index); returns the tree depth, -1 for n <= 0. Node 0 is the root.
int build_kdtree(int k, int d, int n, const double* X, std::vector<int>* perm, std::vector<KdNode>* nodes);

// Bounding boxes of the nodes: boxes[(node * 2) * d + i] = the smallest, and
// boxes[(node * 2 + 1) * d + i] the largest, coordinate i of the node's points.
void kdtree_boxes(int d, const double* X, const std::vector<int>& perm, const std::vector<KdNode>& nodes,
                  std::vector<double>* boxes);

struct KnnArgs {
    int N, d, k, depth;
    const double* queries;   // [N][d]
    const KdNode* nodes;
    const double* boxes;     // kdtree_boxes, or null (no box test)
    const double* points;    // permuted order, each leaf's block [begin*d, end*d) dimension-major
    const int* hits;         // [n] in permuted order
    const int* perm;         // position -> original index
    int* out_count;          // [N]
    int* out_index;          // [N][k] original indices, farthest first
    double* out_dist_sq;     // [N][k]
    int* next;               // device counter of the launch's query queue (launch_knn zeroes it)
};

size_t knn_lds_bytes(int d, int k, int depth);
hipError_t launch_knn(const KnnArgs& a, hipStream_t stream);

}  // namespace wp
