// layout.h -- packed scene layout shared by the host builder and the gfx950 kernels.
//
// Wide BVH node (W = 4 or 8 children), structure of arrays inside the node so that one
// lane streams a node with W*2*DIM/4 + W/4 dwordx4 loads:
//   float lo[DIM][W]; float hi[DIM][W]; int32 child[W]; padding to NODE_BYTES.
// child >= 0 : index of an inner node
// child <  0 : leaf, ~child = (first_prim << 4) | (prim_count - 1)
// empty slots hold a degenerate box at +1e30 whose squared distance overflows to +inf
// and whose slab interval lies beyond any ray, so traversal never enters them.
//
// Primitives are stored in leaf order:
//   3D triangle : 12 floats  a.xyz b.xyz c.xyz, caller index (int bits), 0, 0   (48 B)
//   2D segment  :  4 floats  a.xy b.xy                                          (16 B)
//                 + int32 caller index in a separate array
#pragma once
#include <cstdint>

namespace zg {

constexpr float kEmptyBox = 1e30f;

template <int DIM, int W>
struct NodeLayout {
    static constexpr int kBoxFloats = 2 * DIM * W;
    static constexpr int kRawFloats = kBoxFloats + W;
    // pad to a multiple of 64 bytes (16 floats)
    static constexpr int kFloats = ((kRawFloats + 15) / 16) * 16;
    static constexpr int kBytes = kFloats * 4;
    static constexpr int kChildOffset = kBoxFloats;
};

template <int DIM>
struct PrimLayout {
    static constexpr int kFloats = DIM == 3 ? 12 : 4;
    static constexpr int kBytes = kFloats * 4;
};

// silhouette candidate (2D reflecting boundary): v.xy, n0.xy, n1.xy, single-neighbour flag, pad
constexpr int kSilFloats = 8;    // 2D candidate: vertex.xy, n0.xy, n1.xy, open end, pad
constexpr int kSil3Floats = 16;  // 3D candidate edge: a.xyz, open edge; b.xyz, 0; n0.xyz, 0; n1.xyz, 0

inline int leafFirst(int child) { return (~child) >> 4; }
inline int leafCount(int child) { return ((~child) & 15) + 1; }
inline int encodeLeaf(int first, int count) { return ~((first << 4) | (count - 1)); }

// Device view of one boundary part (absorbing or reflecting)
struct BoundaryView {
    const float* nodes;    // nNodes * NodeLayout::kFloats
    const float* prims;    // nPrims * PrimLayout::kFloats (leaf order)
    const int32_t* primIds;// 2D: caller index per packed primitive
    const float* sil;      // reflecting: silhouette candidates (kSilFloats each in 2D, kSil3Floats in 3D)
    const float* silNodes; // reflecting: wide BVH over the silhouette candidates
    const float* silCones; // per silhouette-BVH child: normal cone (2D: axis.xy, sin, cos of half-angle + margin;
                           // 3D: axis.xyz, sin; cos, 0, 0, 0)
    float silConePrecision;// the silhouette precision the cones' margin covers
    const float* robinMax; // 2D reflecting: largest Robin coefficient per packed primitive (null: 0)
    int32_t nNodes;
    int32_t nPrims;
    int32_t nSil;
    int32_t nSilNodes;
};

struct SceneView {
    BoundaryView abs, refl;
    float dmin[3], dmax[3];
    int32_t dim, width;
};

}  // namespace zg
