// spatial_sort.cpp -- spatially sorted point indices (host, cold path).
//
// getSpatiallySortedPointIndices (utils/nearest_neighbor_finder.h:112-128) returns the index
// order of a nanoflann KDTreeSingleIndexAdaptor built over the points with leaves of at most
// 10 points (:66-79).  nanoflann (deps/nanoflann) is not vendored, so its published build is
// restated: the root box is the points' bounding box; a node of more than 10 points splits
// the dimension of largest box extent (among the extents within 1e-5 of the largest, the one
// of largest point spread) at the box midpoint clamped to the points' range, partitions its
// index range into [< cut | == cut | > cut] with two in-place passes, and cuts the range at
// the first boundary above half of it, the second below half, or at half; each child's box is
// the parent's cut at that value.  The resulting index order keeps every leaf contiguous and
// the leaves in depth-first order.
#include <algorithm>
#include <cfloat>
#include <cstdint>
#include <vector>

namespace zg {

namespace {

struct KdOrder {
    const float* pts;
    int dim;
    std::vector<int32_t> idx;

    float at(int64_t i, int a) const { return pts[(size_t)idx[(size_t)i] * dim + a]; }

    // two in-place passes: [< cut) then [== cut), returning the ends of both
    void planeSplit(int64_t lo, int64_t count, int a, float cut, int64_t& lim1, int64_t& lim2) {
        int64_t l = 0, r = count - 1;
        for (;;) {
            while (l <= r && at(lo + l, a) < cut) ++l;
            while (r && l <= r && at(lo + r, a) >= cut) --r;
            if (l > r || !r) break;
            std::swap(idx[(size_t)(lo + l)], idx[(size_t)(lo + r)]);
            ++l; --r;
        }
        lim1 = l;
        r = count - 1;
        for (;;) {
            while (l <= r && at(lo + l, a) <= cut) ++l;
            while (r && l <= r && at(lo + r, a) > cut) --r;
            if (l > r || !r) break;
            std::swap(idx[(size_t)(lo + l)], idx[(size_t)(lo + r)]);
            ++l; --r;
        }
        lim2 = l;
    }

    // box: 2 * dim floats (low, high per dimension); on return the union of the children's
    // boxes (the leaves' point boxes), as the tree's build leaves it
    void divide(int64_t lo, int64_t hi, float* box) {
        const int64_t count = hi - lo;
        if (count <= 10) {
            for (int a = 0; a < dim; a++) {
                float mn = FLT_MAX, mx = -FLT_MAX;
                for (int64_t i = lo; i < hi; i++) { mn = std::min(mn, at(i, a)); mx = std::max(mx, at(i, a)); }
                box[2 * a] = mn; box[2 * a + 1] = mx;
            }
            return;
        }
        float maxSpan = box[1] - box[0];
        for (int a = 1; a < dim; a++) maxSpan = std::max(maxSpan, box[2 * a + 1] - box[2 * a]);
        int cutDim = 0;
        float maxSpread = -1.0f, minE = 0.0f, maxE = 0.0f;
        for (int a = 0; a < dim; a++) {
            if (box[2 * a + 1] - box[2 * a] >= (1.0f - 1e-5f) * maxSpan) {
                float mn = FLT_MAX, mx = -FLT_MAX;
                for (int64_t i = lo; i < hi; i++) { mn = std::min(mn, at(i, a)); mx = std::max(mx, at(i, a)); }
                if (mx - mn > maxSpread) { cutDim = a; maxSpread = mx - mn; minE = mn; maxE = mx; }
            }
        }
        float cut = (box[2 * cutDim] + box[2 * cutDim + 1]) / 2;
        cut = cut < minE ? minE : (cut > maxE ? maxE : cut);
        int64_t lim1, lim2;
        planeSplit(lo, count, cutDim, cut, lim1, lim2);
        int64_t mid = lim1 > count / 2 ? lim1 : (lim2 < count / 2 ? lim2 : count / 2);
        float lbox[6], rbox[6];
        std::copy(box, box + 2 * dim, lbox);
        std::copy(box, box + 2 * dim, rbox);
        lbox[2 * cutDim + 1] = cut;
        rbox[2 * cutDim] = cut;
        divide(lo, lo + mid, lbox);
        divide(lo + mid, hi, rbox);
        for (int a = 0; a < dim; a++) {
            box[2 * a] = std::min(lbox[2 * a], rbox[2 * a]);
            box[2 * a + 1] = std::max(lbox[2 * a + 1], rbox[2 * a + 1]);
        }
    }
};

}  // namespace

void spatiallySortedIndices(int dim, const float* pts, int64_t n, int32_t* out) {
    KdOrder k{pts, dim, std::vector<int32_t>((size_t)n)};
    for (int64_t i = 0; i < n; i++) k.idx[(size_t)i] = (int32_t)i;
    if (n > 0) {
        float box[6];
        for (int a = 0; a < dim; a++) {
            float mn = FLT_MAX, mx = -FLT_MAX;
            for (int64_t i = 0; i < n; i++) { mn = std::min(mn, pts[(size_t)i * dim + a]); mx = std::max(mx, pts[(size_t)i * dim + a]); }
            box[2 * a] = mn; box[2 * a + 1] = mx;
        }
        k.divide(0, n, box);
    }
    std::copy(k.idx.begin(), k.idx.end(), out);
}

}  // namespace zg
