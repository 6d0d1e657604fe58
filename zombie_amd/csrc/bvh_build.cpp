// bvh_build.cpp -- host construction of the packed wide BVHs (cold path, once per scene).
//
// Plays the role of FCPW's Bvh_SurfaceArea build invoked by
// FcpwDirichletBoundaryHandler / FcpwNeumannBoundaryHandler::buildAccelerationStructure
// (utils/fcpw_geometric_queries.h:537-679) and of computeSilhouettes (:631,667).
// Design: a binary binned-SAH tree is built first (all axes, `bins` buckets, leaves of at
// most `leafSize` primitives) and then collapsed into a W-wide tree by repeatedly opening
// the inner child of largest surface area until a node has W children, which is the usual
// construction of wide BVHs for SIMD/SIMT traversal.  Nodes and primitives are then packed
// in the layout of layout.h.
#include "bvh_build.h"

#include <algorithm>
#include <climits>
#include <cmath>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <map>
#include <stdexcept>

namespace zg {

namespace {

template <int DIM>
struct Box {
    float lo[DIM], hi[DIM];
    Box() { for (int i = 0; i < DIM; i++) { lo[i] = FLT_MAX; hi[i] = -FLT_MAX; } }
    void expand(const Box& b) { for (int i = 0; i < DIM; i++) { lo[i] = std::min(lo[i], b.lo[i]); hi[i] = std::max(hi[i], b.hi[i]); } }
    void expandPt(const float* p) { for (int i = 0; i < DIM; i++) { lo[i] = std::min(lo[i], p[i]); hi[i] = std::max(hi[i], p[i]); } }
    float area() const {
        float e[DIM];
        for (int i = 0; i < DIM; i++) { e[i] = hi[i] - lo[i]; if (e[i] < 0) return 0.0f; }
        if (DIM == 2) return 2.0f * (e[0] + e[1]);
        return 2.0f * (e[0] * e[1] + e[1] * e[DIM - 1] + e[0] * e[DIM - 1]);
    }
};

template <int DIM>
struct BinNode {
    Box<DIM> box;
    int left = -1, right = -1, start = 0, count = 0;
};

template <int DIM>
struct BinaryBuilder {
    const std::vector<Box<DIM>>& boxes;
    std::vector<float> cent;  // DIM per item
    std::vector<int> order;
    std::vector<BinNode<DIM>> nodes;
    int leafSize, bins;

    BinaryBuilder(const std::vector<Box<DIM>>& b, int leafSize_, int bins_) : boxes(b), leafSize(leafSize_), bins(bins_) {
        int n = (int)b.size();
        cent.resize((size_t)n * DIM);
        order.resize(n);
        for (int i = 0; i < n; i++) {
            order[i] = i;
            for (int a = 0; a < DIM; a++) cent[(size_t)i * DIM + a] = 0.5f * (b[i].lo[a] + b[i].hi[a]);
        }
        nodes.reserve(2 * (size_t)n / std::max(1, leafSize) + 8);
    }

    int bucketOf(int item, int axis, float lo, float inv) const {
        int k = (int)((cent[(size_t)item * DIM + axis] - lo) * inv);
        return std::min(std::max(k, 0), bins - 1);
    }

    void build() {
        struct Task { int node, start, end; };
        std::vector<Task> stack;
        nodes.push_back(BinNode<DIM>());
        stack.push_back({0, 0, (int)order.size()});
        std::vector<Box<DIM>> bb(bins);
        std::vector<int> cnt(bins);
        std::vector<Box<DIM>> rightBox(bins);
        std::vector<int> rightCnt(bins);
        while (!stack.empty()) {
            Task t = stack.back(); stack.pop_back();
            Box<DIM> box, cb;
            for (int i = t.start; i < t.end; i++) {
                box.expand(boxes[order[i]]);
                cb.expandPt(&cent[(size_t)order[i] * DIM]);
            }
            nodes[t.node].box = box;
            int n = t.end - t.start;
            if (n <= leafSize) { nodes[t.node].start = t.start; nodes[t.node].count = n; continue; }
            float bestCost = FLT_MAX; int bestAxis = -1, bestSplit = -1;
            for (int a = 0; a < DIM; a++) {
                float ext = cb.hi[a] - cb.lo[a];
                if (!(ext > 0.0f)) continue;
                float inv = bins / ext;
                for (int k = 0; k < bins; k++) { bb[k] = Box<DIM>(); cnt[k] = 0; }
                for (int i = t.start; i < t.end; i++) {
                    int k = bucketOf(order[i], a, cb.lo[a], inv);
                    cnt[k]++; bb[k].expand(boxes[order[i]]);
                }
                Box<DIM> acc; int accN = 0;
                for (int k = bins - 1; k >= 1; k--) {
                    acc.expand(bb[k]); accN += cnt[k];
                    rightBox[k] = acc; rightCnt[k] = accN;
                }
                Box<DIM> l; int nl = 0;
                for (int s = 1; s < bins; s++) {
                    l.expand(bb[s - 1]); nl += cnt[s - 1];
                    int nr = rightCnt[s];
                    if (nl == 0 || nr == 0) continue;
                    float cost = l.area() * nl + rightBox[s].area() * nr;
                    if (cost < bestCost) { bestCost = cost; bestAxis = a; bestSplit = s; }
                }
            }
            int mid = -1;
            if (bestAxis >= 0) {
                float ext = cb.hi[bestAxis] - cb.lo[bestAxis];
                float inv = bins / ext;
                float lo = cb.lo[bestAxis];
                auto it = std::partition(order.begin() + t.start, order.begin() + t.end,
                                         [&](int item) { return bucketOf(item, bestAxis, lo, inv) < bestSplit; });
                mid = (int)(it - order.begin());
            }
            if (mid <= t.start || mid >= t.end) {
                // all centroids coincide (or no valid split): median split along the widest axis
                int a = 0;
                for (int k = 1; k < DIM; k++) if (box.hi[k] - box.lo[k] > box.hi[a] - box.lo[a]) a = k;
                mid = t.start + n / 2;
                std::nth_element(order.begin() + t.start, order.begin() + mid, order.begin() + t.end,
                                 [&](int x, int y) { return cent[(size_t)x * DIM + a] < cent[(size_t)y * DIM + a]; });
            }
            int l = (int)nodes.size(); nodes.push_back(BinNode<DIM>());
            int r = (int)nodes.size(); nodes.push_back(BinNode<DIM>());
            nodes[t.node].left = l; nodes[t.node].right = r;
            stack.push_back({r, mid, t.end});
            stack.push_back({l, t.start, mid});
        }
    }
};

template <int DIM, int W>
struct WideCollapser {
    using L = NodeLayout<DIM, W>;
    const std::vector<BinNode<DIM>>& bin;
    std::vector<float>& out;
    int maxDepth = 0;

    WideCollapser(const std::vector<BinNode<DIM>>& b, std::vector<float>& o) : bin(b), out(o) {}

    int newNode() {
        int id = (int)(out.size() / L::kFloats);
        out.resize(out.size() + L::kFloats, 0.0f);
        float* nd = &out[(size_t)id * L::kFloats];
        for (int k = 0; k < L::kBoxFloats; k++) nd[k] = kEmptyBox;
        int32_t* ch = reinterpret_cast<int32_t*>(nd + L::kChildOffset);
        for (int k = 0; k < W; k++) ch[k] = 0;
        return id;
    }

    void setChild(int id, int slot, const Box<DIM>& box, int32_t code) {
        float* nd = &out[(size_t)id * L::kFloats];
        for (int a = 0; a < DIM; a++) { nd[a * W + slot] = box.lo[a]; nd[(DIM + a) * W + slot] = box.hi[a]; }
        reinterpret_cast<int32_t*>(nd + L::kChildOffset)[slot] = code;
    }

    // collapses the binary subtree rooted at b into wide node `id`
    void collapse(int id, int b, int depth) {
        maxDepth = std::max(maxDepth, depth);
        std::vector<int> kids;
        if (bin[b].left < 0) kids.push_back(b);
        else { kids.push_back(bin[b].left); kids.push_back(bin[b].right); }
        while ((int)kids.size() < W) {
            int best = -1; float bestArea = -1.0f;
            for (int k = 0; k < (int)kids.size(); k++) {
                const auto& c = bin[kids[k]];
                if (c.left >= 0 && c.box.area() > bestArea) { bestArea = c.box.area(); best = k; }
            }
            if (best < 0) break;
            int c = kids[best];
            kids[best] = bin[c].left;
            kids.push_back(bin[c].right);
        }
        for (int k = 0; k < (int)kids.size(); k++) {
            const auto& c = bin[kids[k]];
            if (c.left < 0) {
                setChild(id, k, c.box, encodeLeaf(c.start, c.count));
            } else {
                int cid = newNode();
                setChild(id, k, c.box, cid);
                collapse(cid, kids[k], depth + 1);
            }
        }
    }
};

template <int DIM, int W>
void buildWide(const std::vector<Box<DIM>>& boxes, int leafSize, int bins, std::vector<float>& nodes,
               std::vector<int>& order, int& nNodes, int& depth) {
    nodes.clear(); order.clear(); nNodes = 0; depth = 0;
    if (boxes.empty()) return;
    BinaryBuilder<DIM> bb(boxes, leafSize, bins);
    bb.build();
    order = bb.order;
    WideCollapser<DIM, W> wc(bb.nodes, nodes);
    int root = wc.newNode();
    wc.collapse(root, 0, 1);
    nNodes = (int)(nodes.size() / NodeLayout<DIM, W>::kFloats);
    depth = wc.maxDepth;
}

template <int DIM>
void buildDim(int width, int leafSize, int bins, const float* V, const int32_t* I,
              const std::vector<int>& sel, BuiltBoundary& out) {
    int n = (int)sel.size();
    std::vector<Box<DIM>> boxes(n);
    for (int i = 0; i < n; i++) {
        int p = sel[i];
        for (int k = 0; k < DIM; k++) boxes[i].expandPt(V + (size_t)I[(size_t)p * DIM + k] * DIM);
    }
    std::vector<int> order;
    if (width == 4) buildWide<DIM, 4>(boxes, leafSize, bins, out.nodes, order, out.nNodes, out.depth);
    else buildWide<DIM, 8>(boxes, leafSize, bins, out.nodes, order, out.nNodes, out.depth);
    out.nPrims = n;
    out.prims.assign((size_t)n * PrimLayout<DIM>::kFloats, 0.0f);
    out.primIds.assign(n, 0);
    for (int i = 0; i < n; i++) {
        int p = sel[order[i]];
        float* dst = &out.prims[(size_t)i * PrimLayout<DIM>::kFloats];
        for (int k = 0; k < DIM; k++)
            for (int a = 0; a < DIM; a++) dst[k * DIM + a] = V[(size_t)I[(size_t)p * DIM + k] * DIM + a];
        if (DIM == 3) { int32_t id = p; std::memcpy(&dst[9], &id, 4); }
        out.primIds[i] = p;
    }
}

}  // namespace

void buildBoundary(int dim, int width, int leafSize, int bins, const float* V, const int32_t* I,
                   const std::vector<int>& sel, BuiltBoundary& out) {
    if (width != 4 && width != 8) throw std::invalid_argument("bvh_width must be 4 or 8");
    if (leafSize < 1 || leafSize > 8) throw std::invalid_argument("leaf_size must be in 1..8");
    if (bins < 4 || bins > 64) throw std::invalid_argument("sah_bins must be in 4..64");
    out = BuiltBoundary();
    if (dim == 2) buildDim<2>(width, leafSize, bins, V, I, sel, out);
    else buildDim<3>(width, leafSize, bins, V, I, sel, out);
}

// Normal cones of the silhouette BVH (the role of the spatialized normal cone hierarchy behind
// fcpw's findClosestSilhouettePoint): for every child subtree, the smallest arc of directions
// containing the normals of all segments adjacent to its candidates, widened by the angular
// margin asin(precision) + 1e-4.  A subtree seen from x under directions whose angle to every
// such normal stays away from 90 degrees by more than the margin holds only vertices whose two
// segments both face x, or both face away, by more than the precision: no silhouette.
// Subtrees holding an open end (always a silhouette) get a full cone.  Stored per child as
// (axis.x, axis.y, sin(half + margin), cos(half + margin)); a cone of half-angle + margin
// >= 90 degrees is stored as (0, 0, 1, -1), which never prunes.
template <int W>
static void buildSilhouetteCones(BuiltBoundary& out, float precision) {
    using L = NodeLayout<2, W>;
    const int nn = out.nSilNodes;
    out.silCones.assign((size_t)nn * W * 4, 0.0f);
    if (nn == 0) return;
    const double margin = std::asin(std::min(1.0, (double)std::max(precision, 0.0f))) + 1e-4;
    // candidate range [lo, hi) of every node (children are leaves or later nodes)
    std::vector<int> rlo(nn, INT32_MAX), rhi(nn, -1);
    auto childRange = [&](int c, int& lo, int& hi) {
        if (c < 0) { lo = leafFirst(c); hi = lo + leafCount(c); } else { lo = rlo[c]; hi = rhi[c]; }
    };
    for (int i = nn - 1; i >= 0; i--) {
        const int32_t* ch = reinterpret_cast<const int32_t*>(&out.silNodes[(size_t)i * L::kFloats + L::kChildOffset]);
        for (int k = 0; k < W; k++) {
            const float* lo0 = &out.silNodes[(size_t)i * L::kFloats];
            if (lo0[k] >= 1e29f) continue;  // empty slot
            int lo, hi;
            childRange(ch[k], lo, hi);
            if (hi <= lo) continue;
            rlo[i] = std::min(rlo[i], lo); rhi[i] = std::max(rhi[i], hi);
        }
    }
    std::vector<double> ang;
    for (int i = 0; i < nn; i++) {
        const int32_t* ch = reinterpret_cast<const int32_t*>(&out.silNodes[(size_t)i * L::kFloats + L::kChildOffset]);
        const float* lo0 = &out.silNodes[(size_t)i * L::kFloats];
        for (int k = 0; k < W; k++) {
            float* cone = &out.silCones[((size_t)i * W + k) * 4];
            cone[0] = 0.0f; cone[1] = 0.0f; cone[2] = 1.0f; cone[3] = -1.0f;   // never prunes
            if (lo0[k] >= 1e29f) continue;
            int lo, hi;
            childRange(ch[k], lo, hi);
            if (hi <= lo) continue;
            ang.clear();
            bool full = false;
            for (int c = lo; c < hi && !full; c++) {
                const float* cd = &out.sil[(size_t)c * kSilFloats];
                if (cd[6] != 0.0f) { full = true; break; }
                ang.push_back(std::atan2((double)cd[3], (double)cd[2]));
                ang.push_back(std::atan2((double)cd[5], (double)cd[4]));
            }
            if (full || ang.empty()) continue;
            std::sort(ang.begin(), ang.end());
            // the smallest covering arc is the complement of the largest gap between neighbours
            double gap = ang.front() + 2.0 * M_PI - ang.back(), start = ang.front();
            for (size_t j = 1; j < ang.size(); j++)
                if (ang[j] - ang[j - 1] > gap) { gap = ang[j] - ang[j - 1]; start = ang[j]; }
            double arc = 2.0 * M_PI - gap;
            double half = 0.5 * arc + margin;
            if (half >= 0.5 * M_PI) continue;
            double axis = start + 0.5 * arc;
            cone[0] = (float)std::cos(axis); cone[1] = (float)std::sin(axis);
            // sin rounded up, cos rounded down: the stored cone contains the exact one
            cone[2] = std::nextafter((float)std::sin(half), 2.0f);
            cone[3] = std::nextafter((float)std::cos(half), -2.0f);
        }
    }
}

// 2D silhouette candidates of a reflecting polyline: every vertex with its adjacent
// segments; a vertex with a single neighbour (an open end) is always a silhouette.  With
// one-sided problems convex vertices are dropped, as getIgnoreCandidateSilhouetteCallback
// does (fcpw_geometric_queries.h:1029-1038): the dihedral angle, positive at reflex
// vertices of a boundary with outward normals (s.y, -s.x), is compared with the precision.
void buildSilhouettes2D(int width, const float* V, int64_t nV, const int32_t* I, const std::vector<int>& sel,
                        bool solveDoubleSided, float precision, BuiltBoundary& out) {
    std::vector<int> first(nV, -1), second(nV, -1), asStart(nV, -1), asEnd(nV, -1);
    for (int p : sel) {
        for (int k = 0; k < 2; k++) {
            int v = I[(size_t)p * 2 + k];
            if (first[v] < 0) first[v] = p; else if (second[v] < 0) second[v] = p;
            if (k == 0) asStart[v] = p; else asEnd[v] = p;
        }
    }
    auto normal = [&](int p, float* n) {
        const float* a = V + (size_t)I[(size_t)p * 2] * 2;
        const float* b = V + (size_t)I[(size_t)p * 2 + 1] * 2;
        float sx = b[0] - a[0], sy = b[1] - a[1];
        float l = std::sqrt(sx * sx + sy * sy);
        n[0] = sy / l; n[1] = -sx / l;
    };
    std::vector<float> cand;
    for (int64_t v = 0; v < nV; v++) {
        if (first[v] < 0) continue;
        float n0[2] = {0, 0}, n1[2] = {0, 0};
        float single = 0.0f;
        normal(first[v], n0);
        if (second[v] < 0) {
            single = 1.0f;
        } else {
            normal(second[v], n1);
            if (!solveDoubleSided && asEnd[v] >= 0 && asStart[v] >= 0) {
                float a0[2], a1[2];
                normal(asEnd[v], a0); normal(asStart[v], a1);
                float ang = std::atan2(a1[0] * a0[1] - a1[1] * a0[0], a0[0] * a1[0] + a0[1] * a1[1]);
                if (ang < precision) continue;
            }
        }
        const float* pv = V + (size_t)v * 2;
        float c[kSilFloats] = {pv[0], pv[1], n0[0], n0[1], n1[0], n1[1], single, 0.0f};
        cand.insert(cand.end(), c, c + kSilFloats);
    }
    int n = (int)(cand.size() / kSilFloats);
    std::vector<Box<2>> boxes(n);
    for (int i = 0; i < n; i++) boxes[i].expandPt(&cand[(size_t)i * kSilFloats]);
    std::vector<int> order;
    int depth = 0;
    if (width == 4) buildWide<2, 4>(boxes, 4, 16, out.silNodes, order, out.nSilNodes, depth);
    else buildWide<2, 8>(boxes, 4, 16, out.silNodes, order, out.nSilNodes, depth);
    out.nSil = n;
    out.sil.resize(cand.size());
    for (int i = 0; i < n; i++)
        std::memcpy(&out.sil[(size_t)i * kSilFloats], &cand[(size_t)order[i] * kSilFloats], kSilFloats * 4);
    if (width == 4) buildSilhouetteCones<4>(out, precision);
    else buildSilhouetteCones<8>(out, precision);
}

// Normal cones of the 3D silhouette-edge BVH: for every child subtree a cone around the
// normalised sum of its candidates' triangle normals whose half-angle is the largest angle to any
// of them plus the margin asin(precision) + 1e-4 (a conservative, not minimal, cone).  The
// pruning rule of the 2D cones then holds on the sphere of directions: seen from x under view
// directions whose angle to every normal stays away from 90 degrees by more than the margin, both
// triangles of every edge face x, or both face away, by more than the precision -- no silhouette.
// Subtrees with an open edge (always a silhouette) get a full cone.  Stored per child as two
// float4: (axis.xyz, sin(half + margin)), (cos(half + margin), plane offset, flat, tolerance);
// (0, 0, 0, 1), (-1, ...) never prunes.  A child whose triangles are coplanar (every normal
// within 1e-5 rad of the axis, every edge end point within 1e-6 (1 + |p|) of the plane
// axis . y = offset) is flagged flat: from x at distance h of that plane every view direction
// v to a point at distance r has |n . v| = h / r up to the tolerance, which lets the query skip
// the child when all its points lie beyond the precision band or all inside it.
template <int W>
static void buildSilhouetteCones3D(BuiltBoundary& out, float precision) {
    using L = NodeLayout<3, W>;
    const int nn = out.nSilNodes;
    out.silCones.assign((size_t)nn * W * 8, 0.0f);
    if (nn == 0) return;
    const double margin = std::asin(std::min(1.0, (double)std::max(precision, 0.0f))) + 1e-4;
    std::vector<int> rlo(nn, INT32_MAX), rhi(nn, -1);
    auto childRange = [&](int c, int& lo, int& hi) {
        if (c < 0) { lo = leafFirst(c); hi = lo + leafCount(c); } else { lo = rlo[c]; hi = rhi[c]; }
    };
    for (int i = nn - 1; i >= 0; i--) {
        const int32_t* ch = reinterpret_cast<const int32_t*>(&out.silNodes[(size_t)i * L::kFloats + L::kChildOffset]);
        const float* lo0 = &out.silNodes[(size_t)i * L::kFloats];
        for (int k = 0; k < W; k++) {
            if (lo0[k] >= 1e29f) continue;
            int lo, hi;
            childRange(ch[k], lo, hi);
            if (hi <= lo) continue;
            rlo[i] = std::min(rlo[i], lo); rhi[i] = std::max(rhi[i], hi);
        }
    }
    for (int i = 0; i < nn; i++) {
        const int32_t* ch = reinterpret_cast<const int32_t*>(&out.silNodes[(size_t)i * L::kFloats + L::kChildOffset]);
        const float* lo0 = &out.silNodes[(size_t)i * L::kFloats];
        for (int k = 0; k < W; k++) {
            float* cone = &out.silCones[((size_t)i * W + k) * 8];
            cone[0] = cone[1] = cone[2] = 0.0f; cone[3] = 1.0f; cone[4] = -1.0f;   // never prunes
            if (lo0[k] >= 1e29f) continue;
            int lo, hi;
            childRange(ch[k], lo, hi);
            if (hi <= lo) continue;
            bool full = false;
            double ax[3] = {0.0, 0.0, 0.0};
            for (int c = lo; c < hi && !full; c++) {
                const float* cd = &out.sil[(size_t)c * kSil3Floats];
                if (cd[3] != 0.0f) { full = true; break; }
                for (int a = 0; a < 3; a++) ax[a] += (double)cd[8 + a] + (double)cd[12 + a];
            }
            if (full) continue;
            double l = std::sqrt(ax[0] * ax[0] + ax[1] * ax[1] + ax[2] * ax[2]);
            if (!(l > 1e-6)) continue;
            for (int a = 0; a < 3; a++) ax[a] /= l;
            double theta = 0.0;
            for (int c = lo; c < hi; c++) {
                const float* cd = &out.sil[(size_t)c * kSil3Floats];
                for (int o : {8, 12}) {
                    double dt = ax[0] * cd[o] + ax[1] * cd[o + 1] + ax[2] * cd[o + 2];
                    double nl = std::sqrt((double)cd[o] * cd[o] + (double)cd[o + 1] * cd[o + 1] + (double)cd[o + 2] * cd[o + 2]);
                    theta = std::max(theta, std::acos(std::min(1.0, std::max(-1.0, dt / std::max(nl, 1e-30)))));
                }
            }
            double half = theta + margin + 1e-6;
            if (half >= 0.5 * M_PI) continue;
            cone[0] = (float)ax[0]; cone[1] = (float)ax[1]; cone[2] = (float)ax[2];
            cone[3] = std::nextafter((float)std::sin(half), 2.0f);
            cone[4] = std::nextafter((float)std::cos(half), -2.0f);
            // coplanar subtree: plane through the first end point, every end point within tolerance
            if (theta < 1e-5) {
                const float* c0 = &out.sil[(size_t)lo * kSil3Floats];
                double off = ax[0] * c0[0] + ax[1] * c0[1] + ax[2] * c0[2];
                double dev = 0.0, mag = 0.0;
                for (int c = lo; c < hi; c++) {
                    const float* cd = &out.sil[(size_t)c * kSil3Floats];
                    for (int o : {0, 4}) {
                        dev = std::max(dev, std::fabs(ax[0] * cd[o] + ax[1] * cd[o + 1] + ax[2] * cd[o + 2] - off));
                        mag = std::max(mag, std::fabs((double)cd[o]) + std::fabs((double)cd[o + 1]) + std::fabs((double)cd[o + 2]));
                    }
                }
                if (dev <= 1e-6 * (1.0 + mag)) {
                    cone[5] = (float)off;
                    cone[6] = 1.0f;
                    // plane-distance error (end points, float rounding of x . n) + normal spread
                    cone[7] = (float)(2.0 * dev + 4e-7 * (1.0 + mag) + 1e-7);
                }
            }
        }
    }
}

// 3D silhouette candidates of a reflecting triangle mesh: every edge with its (first two)
// adjacent triangles; an edge with one triangle (an open edge) is always a silhouette.  With
// one-sided problems an edge whose dihedral angle atan2(e . (n1 x n0), n0 . n1) -- e the edge
// direction in the winding of its first triangle, n0 and n1 the unit normals of its first and
// second triangles, as fcpw_geometric_queries.h:955-967 measures it -- is below the precision
// (flat or convex) is dropped, as getIgnoreCandidateSilhouetteCallback does (:1029-1038).
// Candidate layout: a.xyz, open; b.xyz, 0; n0.xyz, 0; n1.xyz, 0 (kSil3Floats).
void buildSilhouettes3D(int width, const float* V, int64_t nV, const int32_t* I, const std::vector<int>& sel,
                        bool solveDoubleSided, float precision, BuiltBoundary& out) {
    (void)nV;
    auto normal = [&](int p, float* n) {
        const float* a = V + (size_t)I[(size_t)p * 3] * 3;
        const float* b = V + (size_t)I[(size_t)p * 3 + 1] * 3;
        const float* c = V + (size_t)I[(size_t)p * 3 + 2] * 3;
        float u[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]}, v[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
        n[0] = u[1] * v[2] - u[2] * v[1]; n[1] = u[2] * v[0] - u[0] * v[2]; n[2] = u[0] * v[1] - u[1] * v[0];
        float l = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
        n[0] /= l; n[1] /= l; n[2] /= l;
    };
    struct EdgeFaces { int f0, f1, a, b; };
    std::map<std::pair<int, int>, EdgeFaces> edges;
    for (int p : sel) {
        for (int j = 0; j < 3; j++) {
            int a = I[(size_t)p * 3 + j], b = I[(size_t)p * 3 + (j + 1) % 3];
            auto key = std::make_pair(std::min(a, b), std::max(a, b));
            auto it = edges.find(key);
            if (it == edges.end()) edges[key] = {p, -1, a, b};
            else if (it->second.f1 < 0) it->second.f1 = p;
        }
    }
    std::vector<float> cand;
    for (const auto& kv : edges) {
        const EdgeFaces& e = kv.second;
        float n0[3], n1[3] = {0.0f, 0.0f, 0.0f};
        normal(e.f0, n0);
        if (e.f1 >= 0) {
            normal(e.f1, n1);
            if (!solveDoubleSided) {
                const float* pa = V + (size_t)e.a * 3;
                const float* pb = V + (size_t)e.b * 3;
                float d[3] = {pb[0] - pa[0], pb[1] - pa[1], pb[2] - pa[2]};
                float l = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
                d[0] /= l; d[1] /= l; d[2] /= l;
                float cr[3] = {n1[1] * n0[2] - n1[2] * n0[1], n1[2] * n0[0] - n1[0] * n0[2], n1[0] * n0[1] - n1[1] * n0[0]};
                float ang = std::atan2(d[0] * cr[0] + d[1] * cr[1] + d[2] * cr[2], n0[0] * n1[0] + n0[1] * n1[1] + n0[2] * n1[2]);
                if (ang < precision) continue;
            }
        }
        const float* pa = V + (size_t)e.a * 3;
        const float* pb = V + (size_t)e.b * 3;
        float c[kSil3Floats] = {pa[0], pa[1], pa[2], e.f1 < 0 ? 1.0f : 0.0f, pb[0], pb[1], pb[2], 0.0f,
                                n0[0], n0[1], n0[2], 0.0f, n1[0], n1[1], n1[2], 0.0f};
        cand.insert(cand.end(), c, c + kSil3Floats);
    }
    int n = (int)(cand.size() / kSil3Floats);
    std::vector<Box<3>> boxes(n);
    for (int i = 0; i < n; i++) {
        boxes[i].expandPt(&cand[(size_t)i * kSil3Floats]);
        boxes[i].expandPt(&cand[(size_t)i * kSil3Floats + 4]);
    }
    std::vector<int> order;
    int depth = 0;
    if (width == 4) buildWide<3, 4>(boxes, 4, 16, out.silNodes, order, out.nSilNodes, depth);
    else buildWide<3, 8>(boxes, 4, 16, out.silNodes, order, out.nSilNodes, depth);
    out.nSil = n;
    out.sil.resize(cand.size());
    for (int i = 0; i < n; i++)
        std::memcpy(&out.sil[(size_t)i * kSil3Floats], &cand[(size_t)order[i] * kSil3Floats], kSil3Floats * 4);
    if (width == 4) buildSilhouetteCones3D<4>(out, precision);
    else buildSilhouetteCones3D<8>(out, precision);
}

}  // namespace zg
