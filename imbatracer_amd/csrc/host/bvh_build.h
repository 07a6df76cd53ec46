// Host BVH builders: binned SAH and SBVH, restated from the reference so that they emit the
// same sequence of nodes and leaves, with the same boxes and triangle order.
//
//   BinnedSahBuilder  <- FastBvhBuilder   core/fast_bvh_builder.h:24-278 (Wald 2007, 32 bins)
//   SplitBvhBuilder   <- SplitBvhBuilder  core/sbvh_builder.h:25-481   (Stich et al. 2009)
//   Fanout            <- MultiNode        core/bvh_helper.h:9-61
//
// Both builders walk the tree depth first with an explicit stack of 128 entries
// (core/stack.h:8-31); a node whose children do not fit on the stack gets leaves as
// children (fast_bvh_builder.h:141-150, sbvh_builder.h:146-155).  Output goes to two
// callbacks in emission order:
//     on_node(const BBox& parent, int count, const BBox* child_boxes)
//     on_leaf(const BBox& box, int count, const int* primitive_ids)
// A layout writer (layout.h) turns this stream into the traversal arrays.
#pragma once

#include <algorithm>
#include <cassert>
#include <memory>
#include <vector>

#include "vec.h"

namespace imba {

// Read-only triangle source: triangle(i) returns the i-th triangle.
struct TriSource {
    const float* vertices = nullptr;   // 4 floats per vertex
    const uint32_t* indices = nullptr; // 4 per triangle
    int tri_count = 0;
    Tri triangle(int i) const {
        const uint32_t* t = indices + 4 * (size_t)i;
        auto v = [&](uint32_t k) { return float3(vertices[4 * (size_t)k], vertices[4 * (size_t)k + 1], vertices[4 * (size_t)k + 2]); };
        return Tri(v(t[0]), v(t[1]), v(t[2]));
    }
};

// Up to N children of a node under construction.  A split replaces candidate i by its left
// half and appends its right half.
template <typename Cand, int N>
struct Fanout {
    Cand c[N];
    BBox bbox;
    int count;

    explicit Fanout(const Cand& first) : bbox(first.bbox), count(1) { c[0] = first; }
    bool full() const { return count == N; }
    bool is_leaf() const { return count == 1; }
    bool any_untested() const {
        for (int i = 0; i < count; i++)
            if (!c[i].tested) return true;
        return false;
    }
    // Next candidate to split: the cheapest untested one (the only one for N == 2).
    int pick() const {
        if (N == 2) return 0;
        float best = FLT_MAX;
        int idx = 0;
        for (int i = 0; i < count; i++)
            if (!c[i].tested && best > c[i].cost) { idx = i; best = c[i].cost; }
        return idx;
    }
    void split(int i, const Cand& l, const Cand& r) {
        assert(count < N);
        c[i] = l;
        c[count++] = r;
    }
    // Children by increasing primitive count; equal counts keep their order (the reference
    // calls std::sort on <= 4 elements, which libstdc++ sorts by stable insertion).
    void order_by_size() {
        for (int i = 1; i < count; i++) {
            Cand x = c[i];
            int j = i;
            while (j > 0 && x.size() < c[j - 1].size()) { c[j] = c[j - 1]; j--; }
            c[j] = x;
        }
    }
};

template <typename T, int Cap = 128>
struct BuildStack {
    T items[Cap];
    int top = -1;
    static constexpr int capacity() { return Cap; }
    void push(const T& v) { assert(top < Cap - 1); items[++top] = v; }
    T pop() { return items[top--]; }
    bool empty() const { return top < 0; }
    int size() const { return top + 1; }
};

// Two-index partition: the first index skips elements that satisfy pred, the second walks
// back over elements that do not, then the two are swapped (libstdc++'s std::partition for
// bidirectional iterators, which the reference calls at fast_bvh_builder.h:266).
template <typename Pred>
int* partition_two_index(int* first, int* last, Pred pred) {
    for (;;) {
        for (;;) {
            if (first == last) return first;
            if (pred(*first)) ++first; else break;
        }
        --last;
        for (;;) {
            if (first == last) return first;
            if (!pred(*last)) --last; else break;
        }
        std::swap(*first, *last);
        ++first;
    }
}

// ---------------------------------------------------------------------------------------
// Binned SAH builder over primitive boxes and centroids.
template <int N, typename Cost>
class BinnedSahBuilder {
public:
    static constexpr int kBins = 32;

    template <typename OnNode, typename OnLeaf>
    void build_mesh(const TriSource& mesh, OnNode&& on_node, OnLeaf&& on_leaf) {
        const int n = mesh.tri_count;
        std::vector<BBox> boxes(n);
        std::vector<float3> centres(n);
        for (int i = 0; i < n; i++) {
            const Tri t = mesh.triangle(i);
            t.compute_bbox(boxes[i]);
            centres[i] = (1.0f / 3.0f) * (t.v0 + t.v1 + t.v2);
        }
        build(boxes.data(), centres.data(), n, on_node, on_leaf);
    }

    template <typename OnNode, typename OnLeaf>
    void build(const BBox* boxes, const float3* centres, int n, OnNode&& on_node, OnLeaf&& on_leaf) {
        BBox global = BBox::empty();
        for (int i = 0; i < n; i++) global.extend(boxes[i]);
        std::vector<int> refs(n);
        for (int i = 0; i < n; i++) refs[i] = i;

        BuildStack<Cand> stack;
        stack.push(Cand::make(0, n, global));
        while (!stack.empty()) {
            Fanout<Cand, N> fan(stack.pop());
            while (!fan.full() && fan.any_untested()) {
                const int id = fan.pick();
                const Cand cand = fan.c[id];
                fan.c[id].tested = true;
                if (try_split(fan, id, cand, boxes, centres, refs.data())) continue;
            }
            fan.order_by_size();
            if (fan.is_leaf()) {
                emit_leaf(fan.c[0], refs.data(), on_leaf);
            } else {
                BBox child_boxes[N];
                for (int i = 0; i < fan.count; i++) child_boxes[i] = fan.c[i].bbox;
                on_node(fan.bbox, fan.count, child_boxes);
                if (stack.size() + fan.count < stack.capacity()) {
                    for (int i = fan.count - 1; i >= 0; i--) stack.push(fan.c[i]);
                } else {
                    for (int i = 0; i < fan.count; i++) emit_leaf(fan.c[i], refs.data(), on_leaf);
                }
            }
        }
    }

private:
    struct Cand {
        BBox bbox;
        int begin, end;
        float cost;
        bool tested;
        int size() const { return end - begin; }
        static Cand make(int b, int e, const BBox& bb) {
            Cand c;
            c.bbox = bb;
            c.begin = b;
            c.end = e;
            c.cost = Cost::leaf_cost(e - b, bb.half_area());
            c.tested = false;
            return c;
        }
    };
    struct Bin {
        int count;
        BBox bbox;
    };

    static int bin_of(float c, float lo, float inv) {
        return std::min(kBins - 1, std::max(0, cvt_trunc((float)kBins * (c - lo) * inv)));
    }

    template <typename OnLeaf>
    static void emit_leaf(const Cand& c, const int* refs, OnLeaf& on_leaf) {
        on_leaf(c.bbox, c.end - c.begin, refs + c.begin);
    }

    // Tries the three axes, longest extent first; the first axis with an accepted split wins.
    bool try_split(Fanout<Cand, N>& fan, int id, const Cand& cand, const BBox* boxes,
                   const float3* centres, int* refs) {
        const int begin = cand.begin, end = cand.end;
        const BBox& pbb = cand.bbox;
        const float3 ext = pbb.max - pbb.min;
        int axes[3] = {0, 1, 2};
        if (ext[axes[0]] < ext[axes[1]]) std::swap(axes[0], axes[1]);
        if (ext[axes[1]] < ext[axes[2]]) std::swap(axes[1], axes[2]);
        if (ext[axes[0]] < ext[axes[1]]) std::swap(axes[0], axes[1]);

        for (int j = 0; j < 3; j++) {
            const int axis = axes[j];
            float cmin = pbb.max[axis];
            float cmax = pbb.min[axis];
            for (int i = begin; i < end; i++) {
                const float c = centres[refs[i]][axis];
                cmin = smin(cmin, c);
                cmax = smax(cmax, c);
            }

            Bin bins[kBins];
            for (int i = 0; i < kBins; i++) { bins[i].count = 0; bins[i].bbox = BBox::empty(); }
            const float inv = 1.0f / (cmax - cmin);
            for (int i = begin; i < end; i++) {
                const int r = refs[i];
                const int b = bin_of(centres[r][axis], cmin, inv);
                bins[b].count++;
                bins[b].bbox.extend(boxes[r]);
            }

            const float parea = pbb.half_area();
            const int best = best_split(bins, Cost::leaf_cost(end - begin, parea) - Cost::traversal_cost(parea));
            if (best >= 0 && best < kBins - 1) {
                const float inv2 = 1.0f / (cmax - cmin);
                const int mid = (int)(partition_two_index(refs + begin, refs + end, [&](int r) {
                    return bin_of(centres[r][axis], cmin, inv2) < best;
                }) - refs);
                BBox lbb = BBox::empty(), rbb = BBox::empty();
                if (kBins < end - begin) {
                    for (int i = 0; i < best; i++) lbb.extend(bins[i].bbox);
                    for (int i = best; i < kBins; i++) rbb.extend(bins[i].bbox);
                } else {
                    for (int i = begin; i < mid; i++) lbb.extend(boxes[refs[i]]);
                    for (int i = mid; i < end; i++) rbb.extend(boxes[refs[i]]);
                }
                fan.split(id, Cand::make(begin, mid, lbb), Cand::make(mid, end, rbb));
                return true;
            }
        }
        return false;
    }

    // SAH sweep over the bin boundaries; returns the first bin of the right side or -1.
    static int best_split(const Bin* bins, float max_cost) {
        float left_cost[kBins];
        int lc = 0;
        BBox lbb = BBox::empty();
        for (int i = 0; i < kBins - 1; i++) {
            lbb.extend(bins[i].bbox);
            lc += bins[i].count;
            left_cost[i] = Cost::leaf_cost(lc, lbb.half_area());
        }
        int rc = 0;
        BBox rbb = BBox::empty();
        float best_cost = max_cost;
        int best = -1;
        for (int i = kBins - 1; i > 0; i--) {
            rbb.extend(bins[i].bbox);
            rc += bins[i].count;
            const float cost = Cost::leaf_cost(rc, rbb.half_area()) + left_cost[i - 1];
            if (cost < best_cost) { best = i; best_cost = cost; }
        }
        return best;
    }
};

// ---------------------------------------------------------------------------------------
// Spatial-split BVH builder over triangles.
template <int N, typename Cost>
class SplitBvhBuilder {
public:
    static constexpr int kSpatialBins = 64;
    static constexpr int kBinningPasses = 2;

    template <typename OnNode, typename OnLeaf>
    void build(const TriSource& mesh, OnNode&& on_node, OnLeaf&& on_leaf, int leaf_threshold, float alpha) {
        const int n = mesh.tri_count;
        Ref* initial = alloc_refs(n);
        right_bbs_.assign(std::max(kSpatialBins, n), BBox());
        BBox mesh_bb = BBox::empty();
        for (int i = 0; i < n; i++) {
            mesh.triangle(i).compute_bbox(initial[i].bb);
            mesh_bb.extend(initial[i].bb);
            initial[i].id = (uint32_t)i;
        }
        const float spatial_threshold = mesh_bb.half_area() * alpha;

        std::vector<int> ids;
        BuildStack<Cand> stack;
        stack.push(Cand::make(initial, n, mesh_bb));
        while (!stack.empty()) {
            Fanout<Cand, N> fan(stack.pop());
            while (!fan.full() && fan.any_untested()) {
                const int id = fan.pick();
                const Cand cand = fan.c[id];
                Ref* refs = cand.refs;
                const int rc = cand.ref_count;
                const BBox& pbb = cand.bbox;

                if (rc <= leaf_threshold) { fan.c[id].tested = true; continue; }

                ObjectSplit os;
                for (int axis = 0; axis < 3; axis++) find_object_split(os, axis, refs, rc);

                SpatialSplit ss;
                if (BBox(os.left_bb).overlap(os.right_bb).half_area() > spatial_threshold) {
                    for (int axis = 0; axis < 3; axis++) {
                        if (pbb.min[axis] == pbb.max[axis]) continue;
                        find_spatial_split(ss, pbb, mesh, axis, refs, rc);
                    }
                }

                const bool spatial = ss.cost < os.cost;
                const float split_cost = spatial ? ss.cost : os.cost;
                if (split_cost + Cost::traversal_cost(pbb.half_area()) >= cand.cost) {
                    fan.c[id].tested = true;
                    continue;
                }

                if (spatial) {
                    Ref *lrefs, *rrefs;
                    BBox lbb, rbb;
                    int lcount, rcount;
                    apply_spatial_split(ss, mesh, refs, rc, lrefs, lcount, lbb, rrefs, rcount, rbb);
                    fan.split(id, Cand::make(lrefs, lcount, lbb), Cand::make(rrefs, rcount, rbb));
                } else {
                    sort_refs(os.axis, refs, rc);
                    fan.split(id, Cand::make(refs, os.left_count, os.left_bb),
                              Cand::make(refs + os.left_count, rc - os.left_count, os.right_bb));
                }
            }

            fan.order_by_size();
            if (fan.is_leaf()) {
                emit_leaf(fan.c[0], ids, on_leaf);
            } else {
                BBox child_boxes[N];
                for (int i = 0; i < fan.count; i++) child_boxes[i] = fan.c[i].bbox;
                on_node(fan.bbox, fan.count, child_boxes);
                if (stack.size() + fan.count < stack.capacity()) {
                    for (int i = fan.count - 1; i >= 0; i--) stack.push(fan.c[i]);
                } else {
                    for (int i = 0; i < fan.count; i++) emit_leaf(fan.c[i], ids, on_leaf);
                }
            }
        }
        pool_.clear();
        right_bbs_.clear();
    }

private:
    struct Ref {
        uint32_t id;
        BBox bb;
    };
    struct Bin {
        BBox bb;
        int entry, exit;
    };
    struct ObjectSplit {
        int axis = 0;
        float cost = FLT_MAX;
        BBox left_bb = BBox::empty(), right_bb = BBox::empty();
        int left_count = 0;
    };
    struct SpatialSplit {
        int axis = 0;
        float cost = FLT_MAX;
        float position = 0.0f;
    };
    struct Cand {
        Ref* refs;
        int ref_count;
        BBox bbox;
        float cost;
        bool tested;
        int size() const { return ref_count; }
        static Cand make(Ref* r, int n, const BBox& bb) {
            Cand c;
            c.refs = r;
            c.ref_count = n;
            c.bbox = bb;
            c.cost = Cost::leaf_cost(n, bb.half_area());
            c.tested = false;
            return c;
        }
    };

    Ref* alloc_refs(int n) {
        pool_.emplace_back(new Ref[std::max(n, 1)]);
        return pool_.back().get();
    }

    template <typename OnLeaf>
    static void emit_leaf(const Cand& c, std::vector<int>& ids, OnLeaf& on_leaf) {
        ids.resize(c.ref_count);
        for (int i = 0; i < c.ref_count; i++) ids[i] = (int)c.refs[i].id;
        on_leaf(c.bbox, c.ref_count, ids.data());
    }

    // Orders references by box centre (min + max) along `axis`, then by triangle id.  The ids
    // of one node are distinct, so this is a total order and any sort gives the same result;
    // sorting (key, id, position) triples and permuting once is faster than moving Refs.
    struct SortKey {
        float c;
        uint32_t id;
        int pos;
    };
    std::vector<SortKey> keys_;
    std::vector<Ref> tmp_;
    void sort_refs(int axis, Ref* refs, int n) {
        keys_.resize(n);
        for (int i = 0; i < n; i++) keys_[i] = {refs[i].bb.min[axis] + refs[i].bb.max[axis], refs[i].id, i};
        std::sort(keys_.begin(), keys_.end(), [](const SortKey& a, const SortKey& b) {
            return (a.c < b.c) || (a.c == b.c && a.id < b.id);
        });
        tmp_.assign(refs, refs + n);
        for (int i = 0; i < n; i++) refs[i] = tmp_[keys_[i].pos];
    }

    void find_object_split(ObjectSplit& split, int axis, Ref* refs, int n) {
        sort_refs(axis, refs, n);
        BBox cur = BBox::empty();
        for (int i = n - 1; i > 0; i--) {
            cur.extend(refs[i].bb);
            right_bbs_[i - 1] = cur;
        }
        cur = BBox::empty();
        for (int i = 0; i < n - 1; i++) {
            cur.extend(refs[i].bb);
            const float cost = Cost::leaf_cost(i + 1, cur.half_area()) +
                               Cost::leaf_cost(n - i - 1, right_bbs_[i].half_area());
            if (cost < split.cost) {
                split.axis = axis;
                split.cost = cost;
                split.left_count = i + 1;
                split.left_bb = cur;
                split.right_bb = right_bbs_[i];
            }
        }
    }

    static int clamp_bin(int v) { return (v < 0) ? 0 : ((v > kSpatialBins - 1) ? kSpatialBins - 1 : v); }

    // One binning pass over [axis_min, axis_max]; improves `split` in place and returns the
    // boundary index of the improvement, or -1.
    int spatial_binning(Bin* bins, SpatialSplit& split, const TriSource& mesh, int axis,
                        const Ref* refs, int n, float axis_min, float axis_max) {
        for (int i = 0; i < kSpatialBins; i++) { bins[i].entry = 0; bins[i].exit = 0; bins[i].bb = BBox::empty(); }
        const float bin_size = (axis_max - axis_min) / kSpatialBins;
        const float inv_size = 1.0f / bin_size;
        for (int i = 0; i < n; i++) {
            const Ref& ref = refs[i];
            const int first = clamp_bin(cvt_trunc(inv_size * (ref.bb.min[axis] - axis_min)));
            const int last = clamp_bin(cvt_trunc(inv_size * (ref.bb.max[axis] - axis_min)));
            BBox cur = ref.bb;
            for (int j = first; j < last; j++) {
                BBox lb, rb;
                const float plane = j < kSpatialBins - 1 ? axis_min + (float)(j + 1) * bin_size : axis_max;
                mesh.triangle((int)ref.id).compute_split(lb, rb, axis, plane);
                bins[j].bb.extend(lb.overlap(cur));
                cur.overlap(rb);
            }
            bins[last].bb.extend(cur);
            bins[first].entry++;
            bins[last].exit++;
        }

        BBox cur = BBox::empty();
        for (int i = kSpatialBins - 1; i > 0; i--) {
            cur.extend(bins[i].bb);
            right_bbs_[i - 1] = cur;
        }
        int lc = 0, rc = n;
        cur = BBox::empty();
        int index = -1;
        for (int i = 0; i < kSpatialBins - 1; i++) {
            lc += bins[i].entry;
            rc -= bins[i].exit;
            cur.extend(bins[i].bb);
            const float cost = Cost::leaf_cost(lc, cur.half_area()) + Cost::leaf_cost(rc, right_bbs_[i].half_area());
            if (cost < split.cost) {
                split.axis = axis;
                split.cost = cost;
                split.position = axis_min + (float)(i + 1) * bin_size;
                index = i;
            }
        }
        return index;
    }

    void find_spatial_split(SpatialSplit& split, const BBox& pbb, const TriSource& mesh, int axis,
                            const Ref* refs, int n) {
        float lo = pbb.min[axis];
        float hi = pbb.max[axis];
        Bin bins[kSpatialBins];
        int pass = 0;
        do {
            if (hi <= lo) break;
            if (spatial_binning(bins, split, mesh, axis, refs, n, lo, hi) < 0) break;
            const float bin_size = (hi - lo) / kSpatialBins;
            lo = split.position - bin_size;
            hi = split.position + bin_size;
            pass++;
        } while (pass < kBinningPasses);
    }

    // Three-way partition (left / straddling / right), then each straddler is put left,
    // put right, or duplicated into both, whichever is cheapest.
    void apply_spatial_split(const SpatialSplit& split, const TriSource& mesh, Ref* refs, int n,
                             Ref*& lrefs, int& lcount, BBox& lbb, Ref*& rrefs, int& rcount, BBox& rbb) {
        const int axis = split.axis;
        const float pos = split.position;
        int first_right = n;
        int cur = 0;
        lcount = 0;
        lbb = BBox::empty();
        rbb = BBox::empty();
        while (cur < first_right) {
            if (refs[cur].bb.max[axis] <= pos) {
                lbb.extend(refs[cur].bb);
                std::swap(refs[cur], refs[lcount]);
                cur++;
                lcount++;
            } else if (refs[cur].bb.min[axis] >= pos) {
                rbb.extend(refs[cur].bb);
                --first_right;
                std::swap(refs[cur], refs[first_right]);
            } else {
                cur++;
            }
        }
        rcount = n - first_right;

        std::vector<Ref> dups;
        while (lcount < first_right) {
            const Ref ref = refs[lcount];
            BBox lsb, rsb;
            mesh.triangle((int)ref.id).compute_split(lsb, rsb, axis, pos);
            lsb.overlap(ref.bb);
            rsb.overlap(ref.bb);

            const BBox l_unsplit = BBox(ref.bb).extend(lbb);
            const BBox r_unsplit = BBox(ref.bb).extend(rbb);
            const BBox l_dup = BBox(lsb).extend(lbb);
            const BBox r_dup = BBox(rsb).extend(rbb);

            const float c_left = Cost::leaf_cost(lcount + 1, l_unsplit.half_area()) + Cost::leaf_cost(rcount, rbb.half_area());
            const float c_right = Cost::leaf_cost(lcount, lbb.half_area()) + Cost::leaf_cost(rcount + 1, r_unsplit.half_area());
            const float c_dup = Cost::leaf_cost(lcount + 1, l_dup.half_area()) + Cost::leaf_cost(rcount + 1, r_dup.half_area());
            const float c_min = smin(c_dup, smin(c_left, c_right));

            if (c_min == c_left) {
                lbb = l_unsplit;
                lcount++;
            } else if (c_min == c_right) {
                rbb = r_unsplit;
                --first_right;
                std::swap(refs[first_right], refs[lcount]);
                rcount++;
            } else {
                lbb = l_dup;
                rbb = r_dup;
                refs[lcount].bb = lsb;
                Ref d;
                d.id = ref.id;
                d.bb = rsb;
                dups.push_back(d);
                lcount++;
                rcount++;
            }
        }

        lrefs = refs;
        if (dups.empty()) {
            rrefs = refs + lcount;
        } else {
            rrefs = alloc_refs(rcount);
            std::copy(refs + first_right, refs + n, rrefs + dups.size());
            std::copy(dups.begin(), dups.end(), rrefs);
        }
    }

    std::vector<BBox> right_bbs_;
    std::vector<std::unique_ptr<Ref[]>> pool_;
};

}  // namespace imba
