// Host-side SAH builder for the 8-wide BVH the kernels traverse.
//
// It reproduces rustracer's BvhTree construction (src/bvh.rs:47-225 with
// QBVH_WIDTH 8): three levels of binary 12-bucket SAH splits make up to eight
// child groups per node, the bucket-sorted partition is stable, a one-sided
// split falls back to halving, nodes are numbered in pre-order and leaves in
// depth-first order.  Every f32 expression is evaluated in the reference's
// order (the file is compiled without FMA contraction), so the tree — and
// with it the traversal order of the kernels — is the reference's tree.
// Unlike the reference it works on index arrays, never on object lists.
#pragma once
#include <vector>
#include <cstdint>
#include <cmath>
#include <cstring>
#include <algorithm>
#include <limits>

namespace rtb {

struct Box {
    float mn[3], mx[3];
};

inline Box box_union(const Box& a, const Box& b) {
    Box r;
    for (int k = 0; k < 3; k++) { r.mn[k] = std::fmin(a.mn[k], b.mn[k]); r.mx[k] = std::fmax(a.mx[k], b.mx[k]); }
    return r;
}
inline float box_area(const Box& b) {
    float dx = b.mx[0] - b.mn[0], dy = b.mx[1] - b.mn[1], dz = b.mx[2] - b.mn[2];
    return 2.0f * (dx * dy + dy * dz + dz * dx);
}
inline Box box_pad(Box b) {
    const float D = 0.0001f;
    for (int k = 0; k < 3; k++) {
        float p = b.mx[k] - b.mn[k] < D ? D : 0.0f;
        b.mn[k] -= p; b.mx[k] += p;
    }
    return b;
}

// 256-byte node: minx[8] miny[8] minz[8] maxx[8] maxy[8] maxz[8] child[8] pad[8]
struct Node {
    float b[48];
    int32_t child[8];
    int32_t pad[8];
};
static_assert(sizeof(Node) == 256, "node must be 256 bytes");

class Builder {
public:
    static const int W = 8;
    static const int NB = 12;

    // Builds a tree over `boxes`; appends nodes to `nodes` (global indices) and
    // the item order of the leaves to `leaf_items`.  Returns the root node index
    // (a single item still gets a one-slot root node, as BvhTree::from_list does).
    static int32_t build_tree(const std::vector<Box>& boxes, std::vector<Node>& nodes, std::vector<uint32_t>& leaf_items, Box* out_bbox) {
        Builder b(boxes, nodes, leaf_items);
        std::vector<uint32_t> idx(boxes.size());
        for (size_t i = 0; i < idx.size(); i++) idx[i] = (uint32_t)i;
        size_t first_node = nodes.size();
        size_t first_leaf = leaf_items.size();
        Box bb;
        int32_t root = b.build(idx, bb);
        if (root < 0) {
            Node n; blank(n);
            n.child[0] = root;
            set_child_box(n, 0, boxes[0]);
            nodes.push_back(n);
            root = (int32_t)(nodes.size() - 1);
        }
        (void)first_node; (void)first_leaf;
        if (out_bbox) *out_bbox = bb;
        return root;
    }

    // Leaf encoding is relative to the tree's own leaf list: -(k+1).
    // Worst-case stack use of ordered traversal rooted at `root`.
    static uint32_t stack_bound(const std::vector<Node>& nodes, int32_t root, const std::vector<uint32_t>* leaf_extra = nullptr) {
        return bound(nodes, root, leaf_extra);
    }

private:
    const std::vector<Box>& boxes_;
    std::vector<Node>& nodes_;
    std::vector<uint32_t>& leaves_;
    size_t leaf_base_;

    Builder(const std::vector<Box>& b, std::vector<Node>& n, std::vector<uint32_t>& l)
        : boxes_(b), nodes_(n), leaves_(l), leaf_base_(l.size()) {}

    static void blank(Node& n) {
        memset(&n, 0, sizeof n);
        for (int c = 0; c < W; c++) n.child[c] = INT32_MIN;
    }
    static void set_child_box(Node& n, int c, const Box& b) {
        for (int a = 0; a < 3; a++) { n.b[a * 8 + c] = b.mn[a]; n.b[24 + a * 8 + c] = b.mx[a]; }
    }

    static uint32_t bound(const std::vector<Node>& nodes, int32_t e, const std::vector<uint32_t>* leaf_extra) {
        if (e < 0) return leaf_extra ? (*leaf_extra)[(size_t)(-e - 1)] : 0u;
        const Node& n = nodes[(size_t)e];
        uint32_t k = 0, worst = 0;
        for (int c = 0; c < W; c++) if (n.child[c] != INT32_MIN) k++;
        for (int c = 0; c < W; c++) if (n.child[c] != INT32_MIN) {
            uint32_t b = bound(nodes, n.child[c], leaf_extra);
            worst = std::max(worst, b);
        }
        // push k children, the first popped leaves k-1 below it
        return std::max(k, k - 1 + worst);
    }

    uint32_t bucket(const Box& b, int axis, float cmin, float extent) const {
        float c = (b.mn[axis] + b.mx[axis]) * 0.5f;
        float t = (c - cmin) / extent;
        float f = (float)NB * t;
        uint32_t bi = (f > 0.0f) ? (uint32_t)std::min(f, 4294967040.0f) : 0u;
        return std::min<uint32_t>(bi, NB - 1);
    }

    // sah_partition: stable split of idx into (left, right)
    void partition(const std::vector<uint32_t>& idx, std::vector<uint32_t>& L, std::vector<uint32_t>& R) const {
        const float INF = std::numeric_limits<float>::infinity();
        size_t span = idx.size();
        float cmin[3] = {INF, INF, INF}, cmax[3] = {-INF, -INF, -INF};
        for (uint32_t i : idx) {
            const Box& b = boxes_[i];
            for (int a = 0; a < 3; a++) {
                float c = (b.mn[a] + b.mx[a]) * 0.5f;
                cmin[a] = std::fmin(cmin[a], c);
                cmax[a] = std::fmax(cmax[a], c);
            }
        }
        Box nb = boxes_[idx[0]];
        for (size_t k = 1; k < span; k++) nb = box_union(nb, boxes_[idx[k]]);
        float node_sa = box_area(nb);
        float best_cost = INF;
        int best_axis = 0, best_split = 0;
        for (int axis = 0; axis < 3; axis++) {
            float extent = cmax[axis] - cmin[axis];
            if (extent < 1e-6f) continue;
            Box bb[NB]; bool have[NB] = {false}; uint32_t cnt[NB] = {0};
            for (uint32_t i : idx) {
                const Box& b = boxes_[i];
                uint32_t bi = bucket(b, axis, cmin[axis], extent);
                cnt[bi]++;
                bb[bi] = have[bi] ? box_union(bb[bi], b) : b;
                have[bi] = true;
            }
            float lsa[NB]; uint32_t lcnt[NB];
            Box acc; bool ah = false; uint32_t an = 0;
            for (int i = 0; i < NB; i++) {
                an += cnt[i];
                if (have[i]) { acc = ah ? box_union(acc, bb[i]) : bb[i]; ah = true; }
                lcnt[i] = an;
                lsa[i] = ah ? box_area(acc) : 0.0f;
            }
            ah = false; an = 0;
            for (int s = NB - 2; s >= 0; s--) {
                an += cnt[s + 1];
                if (have[s + 1]) { acc = ah ? box_union(acc, bb[s + 1]) : bb[s + 1]; ah = true; }
                uint32_t nl = lcnt[s], nr = an;
                if (nl == 0 || nr == 0) continue;
                float rsa = ah ? box_area(acc) : 0.0f;
                float cost = 1.0f + (lsa[s] * (float)nl + rsa * (float)nr) / node_sa;
                if (cost < best_cost) { best_cost = cost; best_axis = axis; best_split = s; }
            }
        }
        float extent = cmax[best_axis] - cmin[best_axis];
        L.clear(); R.clear();
        if (extent < 1e-6f) {
            // every item in bucket 0: the one-sided fallback halves the input order
            L.assign(idx.begin(), idx.begin() + span / 2);
            R.assign(idx.begin() + span / 2, idx.end());
            return;
        }
        // stable sort by bucket, split at bucket <= best_split
        std::vector<uint32_t> bid(span);
        for (size_t k = 0; k < span; k++) bid[k] = bucket(boxes_[idx[k]], best_axis, cmin[best_axis], extent);
        std::vector<uint32_t> order(span);
        for (size_t k = 0; k < span; k++) order[k] = (uint32_t)k;
        std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return bid[a] < bid[b]; });
        size_t mid_off = 0;
        while (mid_off < span && bid[order[mid_off]] <= (uint32_t)best_split) mid_off++;
        size_t mid = (mid_off == 0 || mid_off == span) ? span / 2 : mid_off;
        L.reserve(mid); R.reserve(span - mid);
        for (size_t k = 0; k < span; k++) (k < mid ? L : R).push_back(idx[order[k]]);
    }

    void split_recursive(const std::vector<uint32_t>& idx, int depth, std::vector<std::vector<uint32_t>>& out) const {
        if (depth == 3 || idx.size() <= 1) { out.push_back(idx); return; }
        std::vector<uint32_t> L, R;
        partition(idx, L, R);
        split_recursive(L, depth + 1, out);
        split_recursive(R, depth + 1, out);
    }

    int32_t build(const std::vector<uint32_t>& idx, Box& out) {
        if (idx.size() == 1) {
            int32_t k = (int32_t)(leaves_.size() - leaf_base_);
            leaves_.push_back(idx[0]);
            out = boxes_[idx[0]];
            return -(k + 1);
        }
        std::vector<std::vector<uint32_t>> groups;
        if (idx.size() <= (size_t)W) {
            for (uint32_t i : idx) groups.push_back(std::vector<uint32_t>(1, i));
        } else {
            split_recursive(idx, 0, groups);
        }
        size_t me = nodes_.size();
        Node n; blank(n);
        nodes_.push_back(n);
        Node built; blank(built);
        Box nb; bool have = false;
        for (size_t g = 0; g < groups.size(); g++) {
            Box cb;
            int32_t c = build(groups[g], cb);
            built.child[g] = c;
            set_child_box(built, (int)g, cb);
            nb = have ? box_union(nb, cb) : cb;
            have = true;
        }
        nodes_[me] = built;
        out = nb;
        return (int32_t)me;
    }
};

}  // namespace rtb
