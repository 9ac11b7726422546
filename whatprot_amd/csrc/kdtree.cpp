// kd-tree construction of the kNN stage (kd-tree/make-node.h:28-80,
// kd-tree/max-min-nth.h:19-110), on the host.
//
// The search result depends on the tree shape and on the order of the points
// inside it (ties in the k-best heap are broken by a point's position, and
// the k-best set depends on the visiting order), so the build restates the
// reference's construction step for step: split on the widest dimension
// (maxima starting at DBL_MIN, minima at DBL_MAX), a Hoare-style partition
// around the middle element that counts its equal keys, the nth-element
// recursion that moves nth as little as duplicates allow, and a leaf whenever
// either side would hold fewer than k points. The points are permuted in
// place (here: a permutation of point indices) exactly as the reference's
// swaps permute its vector.
#include <algorithm>
#include <cfloat>
#include <utility>
#include <vector>

#include "kdtree.h"

namespace wp {
namespace {

struct Builder {
    int k, d;
    const double* X;  // [n][d], indexed by original point index
    std::vector<int>& perm;
    std::vector<KdNode>& nodes;

    double x(int pos, int s) const { return X[(size_t)perm[pos] * d + s]; }
    void swap(int a, int b) { std::swap(perm[a], perm[b]); }

    // Partition [b, e) around the key at pivot; returns the first position of
    // the upper part and counts keys equal to the pivot's
    // (max-min-nth.h:19-52).
    int partition(int b, int e, int s, int pivot, int* num_eq) {
        *num_eq = 0;
        int l = b, r = e - 1;
        double pv = x(pivot, s);
        swap(pivot, r);
        while (l < r) {
            while (l < r) {
                double lv = x(l, s);
                if (lv == pv) (*num_eq)++;
                if (lv > pv) {
                    swap(l, r);
                    r--;
                    break;
                }
                l++;
            }
            while (l < r) {
                double rv = x(r, s);
                if (rv == pv) (*num_eq)++;
                if (rv <= pv) {
                    swap(r, l);
                    l++;
                    break;
                }
                r--;
            }
        }
        return l;
    }

    // Moves keys equal to the last one to the upper end (max-min-nth.h:54-78).
    void partition_alternate(int b, int e, int s) {
        int l = b, r = e - 1;
        double pv = x(r, s);
        while (l < r) {
            while (l < r) {
                if (x(l, s) >= pv) {
                    swap(l, r);
                    r--;
                    break;
                }
                l++;
            }
            while (l < r) {
                if (x(r, s) < pv) {
                    swap(r, l);
                    l++;
                    break;
                }
                r--;
            }
        }
    }

    // Every key in [b, nth) ends strictly below every key in [nth, e); nth
    // moves as little as equal keys require (max-min-nth.h:80-108).
    int nth_element(int b, int e, int s, int nth) {
        int num_eq;
        int piv = partition(b, e, s, nth, &num_eq);
        int r = piv, l = piv - num_eq;
        if (l <= nth && nth <= r) {
            if (nth - l > r - nth) return r + 1;
            partition_alternate(b, r + 1, s);
            return l;
        } else if (nth < l) {
            partition_alternate(b, r + 1, s);
            return nth_element(b, l, s, nth);
        }
        return nth_element(r + 1, e, s, nth);
    }

    int make(int b, int e) {
        std::vector<double> mins(d, DBL_MAX), maxes(d, DBL_MIN);
        for (int p = b; p < e; p++) {
            for (int i = 0; i < d; i++) {
                double v = x(p, i);
                if (v < mins[i]) mins[i] = v;
                if (maxes[i] < v) maxes[i] = v;
            }
        }
        double best = -1.0;
        int s = -1;
        for (int i = 0; i < d; i++) {
            double range = maxes[i] - mins[i];
            if (range > best) {
                best = range;
                s = i;
            }
        }
        int nth = b + (e - b) / 2;
        nth = nth_element(b, e, s, nth);
        int id = (int)nodes.size();
        nodes.push_back(KdNode());
        if (nth - b < k || e - nth < k) {
            KdNode& n = nodes[id];
            n.left = -1;
            n.right = -1;
            n.s = -1;
            n.begin = b;
            n.end = e;
            n.max_left = n.min_right = n.split = 0.0;
            return id;
        }
        double max_left = DBL_MIN, min_right = DBL_MAX;
        for (int p = b; p < nth; p++) {
            double v = x(p, s);
            if (max_left < v) max_left = v;
        }
        for (int p = nth; p < e; p++) {
            double v = x(p, s);
            if (v < min_right) min_right = v;
        }
        int left = make(b, nth);
        int right = make(nth, e);
        KdNode& n = nodes[id];
        n.left = left;
        n.right = right;
        n.s = s;
        n.begin = b;
        n.end = e;
        n.max_left = max_left;
        n.min_right = min_right;
        n.split = (max_left + min_right) / 2;  // internal-node.h:30
        return id;
    }
};

}  // namespace

int build_kdtree(int k, int d, int n, const double* X, std::vector<int>* perm, std::vector<KdNode>* nodes) {
    perm->resize(n);
    for (int i = 0; i < n; i++) (*perm)[i] = i;
    nodes->clear();
    if (n <= 0) return -1;
    Builder b{k, d, X, *perm, *nodes};
    int root = b.make(0, n);
    int depth = 0;
    // depth of the tree, for the search stack
    std::vector<std::pair<int, int>> st{{root, 1}};
    while (!st.empty()) {
        auto [id, dep] = st.back();
        st.pop_back();
        if (dep > depth) depth = dep;
        const KdNode& nd = (*nodes)[id];
        if (nd.left >= 0) {
            st.push_back({nd.left, dep + 1});
            st.push_back({nd.right, dep + 1});
        }
    }
    return depth;
}

void kdtree_boxes(int d, const double* X, const std::vector<int>& perm, const std::vector<KdNode>& nodes,
                  std::vector<double>* boxes) {
    boxes->assign(nodes.size() * 2 * (size_t)d, 0.0);
    // children have larger ids than their parent: leaves from their points,
    // then every internal node from its children
    for (size_t id = nodes.size(); id-- > 0;) {
        const KdNode& nd = nodes[id];
        double* lo = &(*boxes)[id * 2 * d];
        double* hi = lo + d;
        if (nd.left < 0) {
            for (int i = 0; i < d; i++) {
                lo[i] = DBL_MAX;
                hi[i] = -DBL_MAX;
            }
            for (int p = nd.begin; p < nd.end; p++) {
                const double* x = X + (size_t)perm[p] * d;
                for (int i = 0; i < d; i++) {
                    lo[i] = std::min(lo[i], x[i]);
                    hi[i] = std::max(hi[i], x[i]);
                }
            }
        } else {
            const double* l = &(*boxes)[(size_t)nd.left * 2 * d];
            const double* r = &(*boxes)[(size_t)nd.right * 2 * d];
            for (int i = 0; i < d; i++) {
                lo[i] = std::min(l[i], r[i]);
                hi[i] = std::max(l[d + i], r[d + i]);
            }
        }
    }
}

}  // namespace wp
