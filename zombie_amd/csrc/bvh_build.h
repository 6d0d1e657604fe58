// bvh_build.h -- host construction of packed wide BVHs (see layout.h for the format).
#pragma once
#include <cstdint>
#include <vector>

#include "layout.h"

namespace zg {

struct BuiltBoundary {
    std::vector<float> nodes;      // nNodes * NodeLayout<DIM,W>::kFloats
    std::vector<float> prims;      // nPrims * PrimLayout<DIM>::kFloats, leaf order
    std::vector<int32_t> primIds;  // caller primitive index per packed primitive
    std::vector<float> sil;        // 2D silhouette candidates, leaf order of silNodes
    std::vector<float> silNodes;   // wide BVH over the candidates
    std::vector<float> silCones;   // normal cone of each child subtree: nSilNodes * W float4 (2D), 2 float4 (3D)
    int nNodes = 0, nPrims = 0, depth = 0, nSil = 0, nSilNodes = 0;
};

// sel: caller primitive indices belonging to this boundary part
void buildBoundary(int dim, int width, int leafSize, int bins, const float* V, const int32_t* I,
                   const std::vector<int>& sel, BuiltBoundary& out);

void buildSilhouettes2D(int width, const float* V, int64_t nV, const int32_t* I, const std::vector<int>& sel,
                        bool solveDoubleSided, float precision, BuiltBoundary& out);
void buildSilhouettes3D(int width, const float* V, int64_t nV, const int32_t* I, const std::vector<int>& sel,
                        bool solveDoubleSided, float precision, BuiltBoundary& out);

// kd-tree leaf order of n points (spatial_sort.cpp)
void spatiallySortedIndices(int dim, const float* pts, int64_t n, int32_t* out);

}  // namespace zg
