// kernels.h -- kernel parameter blocks and host launchers (implemented in *.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/zombie_amd.h"
#include "layout.h"

namespace zg {

// counters[] slots of a solve
enum CounterSlot {
    kCntReached = 0, kCntRR = 1, kCntMaxLen = 2, kCntEscaped = 3,
    kCntSteps = 4, kCntWalks = 5, kCntNodes = 6, kCntPrims = 7, kCntQueries = 8,
    kCntClaims = 9, kCntError = 10,
    // walk on stars, visit counting: reflecting-boundary BVH (ray, visibility, boundary sample,
    // Robin bound) and silhouette BVH (star radius) nodes and primitives
    kCntReflNodes = 11, kCntReflPrims = 12, kCntSilNodes = 13, kCntSilPrims = 14,
    kCntSlots = 16
};

// one instance of a batched solve: scene part walked (absorbing boundary), bounding domain, PDE
struct InstanceDesc {
    const float* nodes;
    const float* prims;
    int32_t nNodes;
    int32_t nPrims;
    float dmin[3], dmax[3];
    zg_pde pde;
};

struct WosParams {
    SceneView scene;
    zg_pde pde;
    zg_walk_settings ws;
    const float* pt;
    const float* normal;
    const uint8_t* type;
    const uint8_t* aligned;
    const float* dist0;
    const int32_t* nWalks;
    const float* cvGrad;
    const float4* cvGrid;       // gradient field nodes (x fastest), null when absent
    int32_t cvRes[3];
    float cvLo[3], cvScale[3];  // node coordinate = (x - lo) * scale
    int64_t nPoints;
    int32_t chunksPerPoint;
    int32_t wpl;
    int32_t regroup;
    int32_t stepInterval;
    unsigned long long nClaims;
    uint64_t seed;
    uint32_t streamOffset;
    float relax2;        // (1 - distance_tolerance)^2
    float guard2;        // relaxation only when best^2 > guard2 (keeps the shell test exact)
    int32_t nInstances;  // batched solves: size of the instance table
    int32_t countMode;   // COUNT kernels: 1 = visits of the kernel's own queries, 2 = visits of an exact
                         // nearest-first traversal from the root at every query point
    double* accSum;
    double* accSq;
    unsigned long long* accN;
    unsigned long long* accLen;
    unsigned long long* claimCounter;
    unsigned long long* counters;
    const InstanceDesc* instances;    // batched solves: instance table
    const int32_t* instanceOf;        // batched solves: instance of each point
    const uint32_t* streamOffsetDev;  // when set, replaces streamOffset (graph replays advance it)
    const zg_pde* pdeDev;             // walk on stars: device copy of pde, read where used
    const float* fsr;                 // walk on stars: first sphere radius per point
    const int64_t* prevN;             // walk on stars: caller's estimate counts (first call when 0)
    unsigned long long* accSplits;    // walk splitting: split parts per point
    // per-point estimation quantity (null: all ZG_ESTIMATE_SOLUTION).  Gradient launches walk the
    // points asking for gradients; solution launches get walk counts with the others zeroed.
    const uint8_t* quantity;
    // gradient launches (estimateSolutionAndGradient)
    const float* dirDeriv;            // direction of the directional derivative (null: e_0)
    const float* strat;               // per-point Latin hypercube tables
    int64_t stratStride;              // floats per table
    int32_t gradW0, gradW1;           // walk indices [gradW0, gradW1) of this stage
    int32_t gradHalves;               // 2 with antithetic pairs (walk w is half w & 1 of pair w >> 1)
    const double* cvMean;             // caller's solution means (boundary-term control variate)
    const int64_t* cvN;               // caller's solution counts
    const double* cvFirst;            // caller's first-step source totals (source-term control variate)
    double* accG;                     // [n*dim] sums of the gradient estimates
    double* accG2;                    // [n*dim] sums of their squares
    double* accFirst;                 // [n] first-step source terms
    double* accDeriv;                 // [n] directional derivative contributions
    const float* distR;               // walk on stars: caller's distances to the reflecting boundary (nullable)
    zg_function sdf;                  // SDF-grid scenes: the interpolated signed-distance grid (grid != null)
};

// capacity of the per-lane queue of split walk parts
constexpr int kSplitQueue = 8;

struct MergeParams {
    int64_t n;
    const float* pt;
    const uint8_t* type;
    const uint8_t* aligned;
    const float* dist0;
    const double* accSum;
    const double* accSq;
    const unsigned long long* accN;
    const unsigned long long* accLen;
    int64_t* sN;
    double* sMean;
    double* sM2;
    int64_t* sLen;
    float* firstSphereRadius;
    zg_function dirichlet;
    float absorption;
    int32_t ignoreAbsorbing;
    int32_t solveDoubleSided;
    int32_t writeFirstSphereRadius;
    const InstanceDesc* instances;
    const int32_t* instanceOf;
    int32_t nInstances;
    uint32_t* streamOffsetDev;        // advanced by streamAdvance once the merge is done
    uint32_t streamAdvance;
    const float* fsrSrc;              // walk on stars: first sphere radius to record (else dist0)
    const unsigned long long* accSplits;
    int64_t* sSplits;                 // caller's total split counts (nullable)
    const uint8_t* quantity;          // per-point estimation quantity (null: all solution)
    int32_t want;                     // quantity of the points this merge folds in
    // gradient statistics (want == ZG_ESTIMATE_SOLUTION_AND_GRADIENT)
    const double* accG;
    const double* accG2;
    const double* accFirst;
    const double* accDeriv;
    int64_t* gN;
    double* gMean;
    double* gM2;
    double* gFirst;
    double* gDeriv;
};

hipError_t launchFillZero(uint64_t* p, int64_t nWords, hipStream_t s);
// solves without a report: their error flags are latched on the device and reported by zg_synchronize
hipError_t launchLatchErrors(const unsigned long long* counters, hipStream_t s);
hipError_t deviceErrorsTakeAndClear(unsigned int* flags);
hipError_t launchMaskWalks(const int32_t* nWalks, const uint8_t* quantity, int64_t n, int32_t* out, hipStream_t s);
hipError_t launchClosestPoint(const SceneView& sv, int boundary, int64_t n, const float* pts, float* dist,
                              float* closest, float* normal, int32_t* prim, hipStream_t s);
hipError_t launchInsideDomain(const SceneView& sv, int64_t n, const float* pts, uint8_t* inside, hipStream_t s);
hipError_t launchRay(const SceneView& sv, int boundary, int64_t n, const float* o, const float* d, const float* tmax,
                     uint8_t* hit, float* t, float* point, float* normal, hipStream_t s);
hipError_t launchWos(const WosParams& P, int dim, int width, bool count, bool screened, bool split, int gridBlocks,
                     hipStream_t s);
hipError_t launchWosBatch(const WosParams& P, int width, bool count, bool screened, bool split, int gridBlocks,
                          hipStream_t s);
hipError_t launchMergeStats(const MergeParams& M, int dim, hipStream_t s);
// gradient solves: walk launch of one stage, and the per-point stratified tables
hipError_t launchWosGrad(const WosParams& P, int dim, int width, bool screened, bool split, int gridBlocks,
                         hipStream_t s);
hipError_t launchStratified(int dims, float* out, int64_t nTables, int64_t stride, int32_t nSamples,
                            const int32_t* nWalks, const uint8_t* quantity, int32_t halves, uint64_t seed,
                            uint32_t streamOffset, hipStream_t s);
hipError_t launchEvalFunction(const zg_function& f, int dim, int64_t n, const float* pts, const float* normals,
                              const uint8_t* aligned, float absorption, float* out, hipStream_t s);
// SDF-grid scenes: walk on spheres and the distance / projection query
hipError_t launchWosSdf(const WosParams& P, int dim, bool screened, bool split, int gridBlocks, hipStream_t s);
hipError_t launchSdfQuery(const zg_function& sdf, int dim, int64_t n, const float* pts, float* dist, float* closest,
                          float* normal, int32_t* prim, hipStream_t s);
hipError_t launchWosSdfGrad(const WosParams& P, int dim, bool screened, bool split, int gridBlocks, hipStream_t s);
hipError_t launchSdfRay(const zg_function& sdf, int dim, int64_t n, const float* o, const float* d, const float* tmax,
                        uint8_t* hit, float* t, float* point, float* normal, hipStream_t s);
hipError_t launchSdfInside(const zg_function& sdf, int dim, int64_t n, const float* pts, uint8_t* inside, hipStream_t s);
// boundary value caching splat
hipError_t launchBvcSplat(int dim, float absorption, const zg_bvc_samples& S, const zg_bvc_eval_points& E,
                          const zg_bvc_statistics& out, float radiusClamp, float kreg, float robinCutoff, float cutoffA,
                          float cutoffR, hipStream_t s);
// read-bandwidth probe
hipError_t launchReadProbe(const void* buf, int64_t bytes, int reps, int blocks, float* sink, hipStream_t s);
// walk on stars (2D)
hipError_t launchWostFirstSphere(const WosParams& P, int dim, int width, const float* callerFsr, float* fsr, hipStream_t s);
hipError_t launchWost(const WosParams& P, int dim, int width, bool count, bool split, bool interleave, int gridBlocks,
                      hipStream_t s);
hipError_t launchWostGradFirstSphere(const WosParams& P, int dim, int width, float* fsr, hipStream_t s);
hipError_t launchWostGrad(const WosParams& P, int dim, int width, int gridBlocks, hipStream_t s);
hipError_t launchStarRadius(const SceneView& sv, int64_t n, const float* pts, const float* minR, const float* maxR,
                            float prec, const uint8_t* flip, float* radius, hipStream_t s);
hipError_t launchLineOfSight(const SceneView& sv, int64_t n, const float* p1, const float* p2, uint8_t* visible,
                             hipStream_t s);
hipError_t launchSampleReflecting(const SceneView& sv, int64_t n, const float* x, const float* radius, const float* u,
                                  uint8_t* ok, float* pt, float* normal, float* pdf, hipStream_t s);

// zg_wos_solve_noalloc: fixed device workspace (noalloc.hip), one copy per GPU for the life of
// the process.  noallocScratch() returns the current device's copy.
constexpr size_t kNoallocScratchBytes = size_t(32) << 20;
hipError_t noallocScratch(unsigned long long** out);

}  // namespace zg
