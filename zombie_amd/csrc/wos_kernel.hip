// wos_kernel.hip -- persistent Walk-on-Spheres kernel for gfx950, plus the batched query
// kernels and the statistics merge.
//
// Work decomposition.  The walks of a solve are cut into claims: claim c covers walks
// [k*wpl, (k+1)*wpl) of point p = c / chunksPerPoint (k = c % chunksPerPoint).  A grid of
// persistent waves runs one walk per lane.  Every `regroup` steps (or as soon as every lane
// of the wave is idle) the wave ballots its idle lanes and hands them the next claims from a
// global counter with one atomic per wave (mbcnt gives each idle lane its slot), so the
// lanes of a wave stay busy although walk lengths are heavy-tailed.  A lane accumulates the
// FP64 sum and sum of squares of its claim's walk estimates in registers and flushes them
// with FP64 atomics when the claim is done; a merge kernel then folds the per-point batch
// into the caller's statistics with Chan's formula.
//
// Per walk the lane restates WalkOnSpheres::walk (point_estimation/walk_on_spheres.h:228-291)
// with estimateSolution's per-point rules (:320-414) and getTerminalContribution (:293-318):
//   ball update, source term (norm * f(y), y drawn from the ball Green's function),
//   uniform direction, move, throughput (Yukawa dampening), escape test, Russian roulette,
//   walk length / max length, Tikhonov switch, closest-point query of the next radius.
// The closest point of the last query is kept: it bounds the next query (the new closest
// distance is at most |x_new - y_prev|) and it is the projection used for the terminal
// Dirichlet value when the walk enters the epsilon shell.
#include <hip/hip_runtime.h>

#include <cfloat>

#include "device.h"
#include "kernels.h"

namespace zg {

// ------------------------------------------------------------------------------------
// batched queries
// ------------------------------------------------------------------------------------
template <int D, int W>
__global__ void __launch_bounds__(256) closestPointKernel(BoundaryView b, int64_t n, const float* __restrict__ pts,
                                                          float* __restrict__ dist, float* __restrict__ closest,
                                                          float* __restrict__ normal, int32_t* __restrict__ prim) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    unsigned long long spill[kSpillStack];
    Stack st; st.init(spill);
    VisitCount vc{0, 0};
    Vec<D> x = vload<D>(pts + i * D);
    float best2 = FLT_MAX;
    Vec<D> cp = vzero<D>(); int pi = -1;
    bool found = closestPoint<D, W, false>(b, x, best2, cp, pi, st, vc);
    dist[i] = found ? __builtin_sqrtf(best2) : FLT_MAX;
    if (closest) vstore<D>(closest + i * D, found ? cp : vzero<D>());
    if (normal || prim) {
        Vec<D> nn = vzero<D>(); int id = -1;
        if (found) {
            Prim<D> pr; pr.load(b.prims, pi);
            nn = pr.normal();
            if (D == 3) id = __float_as_int(b.prims[(size_t)pi * PrimLayout<D>::kFloats + 9]);
            else id = b.primIds[pi];
        }
        if (normal) vstore<D>(normal + i * D, nn);
        if (prim) prim[i] = id;
    }
}

template <int D, int W>
__global__ void __launch_bounds__(256) insideDomainKernel(SceneView s, int64_t n, const float* __restrict__ pts,
                                                          uint8_t* __restrict__ inside) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    unsigned long long spill[kSpillStack];
    Stack st; st.init(spill);
    Vec<D> x = vload<D>(pts + i * D);
    bool in = true;
    for (int a = 0; a < D; a++) {
        int h = countAxisHits<D, W>(s.abs, x, a, st) + countAxisHits<D, W>(s.refl, x, a, st);
        in = in && (h & 1);
    }
    inside[i] = in ? 1 : 0;
}

template <int D, int W>
__global__ void __launch_bounds__(256) rayKernel(BoundaryView b, int64_t n, const float* __restrict__ o,
                                                 const float* __restrict__ d, const float* __restrict__ tmax,
                                                 uint8_t* __restrict__ hit, float* __restrict__ t, float* __restrict__ point,
                                                 float* __restrict__ normal) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    unsigned long long spill[kSpillStack];
    Stack st; st.init(spill);
    Vec<D> oo = vload<D>(o + i * D), dd = vload<D>(d + i * D);
    float th; int pi = -1;
    bool h = intersectRay<D, W>(b, oo, dd, tmax[i], th, pi, st);
    hit[i] = h ? 1 : 0;
    if (t) t[i] = h ? th : tmax[i];
    if (point) vstore<D>(point + i * D, h ? oo + dd * th : vzero<D>());
    if (normal) {
        Vec<D> nn = vzero<D>();
        if (h) { Prim<D> pr; pr.load(b.prims, pi); nn = pr.normal(); }
        vstore<D>(normal + i * D, nn);
    }
}

// ------------------------------------------------------------------------------------
// walk-on-spheres
// ------------------------------------------------------------------------------------
// the part of a scene a walk traverses (per lane in batched solves)
struct WalkView {
    const float* nodes;
    const float* prims;
    int nNodes;
};

// Occupancy target of the walk kernel: 6 waves per SIMD for 4-wide BVHs (at most 80 VGPRs,
// a few spilled; the 24 KB LDS stack of a 256-thread block allows 6 blocks per CU), 4 for
// 8-wide BVHs, whose traversal needs more registers.  Measured on the 1M-point 3D Poisson
// workload: 4 -> 6 waves raised completed walks/s by 15% for width 4; 7 and 8 waves spill
// too much and lose.
template <int W> struct WosOccupancy { static constexpr int kWaves = W == 4 ? 6 : 4; };

// One loop iteration = one BVH node visit for every lane whose distance query is in flight.
// Lanes whose query has finished take their walk step every `stepInterval` iterations (or at
// once when no lane of the wave is traversing), and idle lanes are refilled with new claims
// every `regroup` iterations, so a long traversal on one lane no longer stalls the others.
// control-variate gradient field at x: (bi/tri)linear interpolation of the grid nodes,
// clamped to the grid box
template <int D>
ZG_DEV Vec<D> cvField(const WosParams& P, const Vec<D>& x) {
    int i0[3], i1[3];
    float t[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        if (a < D) {
            float u = fminf(fmaxf((x[a < D ? a : 0] - P.cvLo[a]) * P.cvScale[a], 0.0f), (float)(P.cvRes[a] - 1));
            int k = min((int)u, P.cvRes[a] - 2);
            i0[a] = k; i1[a] = k + 1; t[a] = u - (float)k;
        } else {
            i0[a] = 0; i1[a] = 0; t[a] = 0.0f;
        }
    }
    float4 acc = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
#pragma unroll
    for (int c = 0; c < (D == 3 ? 8 : 4); c++) {
        int ix = (c & 1) ? i1[0] : i0[0], iy = (c & 2) ? i1[1] : i0[1], iz = (c & 4) ? i1[2] : i0[2];
        float wgt = ((c & 1) ? t[0] : 1.0f - t[0]) * ((c & 2) ? t[1] : 1.0f - t[1]);
        if (D == 3) wgt *= (c & 4) ? t[2] : 1.0f - t[2];
        float4 v = P.cvGrid[((size_t)iz * P.cvRes[1] + iy) * P.cvRes[0] + ix];
        acc.x += wgt * v.x; acc.y += wgt * v.y; acc.z += wgt * v.z;
    }
    Vec<D> g;
    g[0] = acc.x; g[1] = acc.y;
    if (D == 3) g[D - 1] = acc.z;
    return g;
}

// SCREENED = false compiles the Yukawa (screened Poisson) terms out of the walk.
// SPLIT = true adds walk splitting (applyWeightWindow, walk_on_spheres.h:192-226): a walk whose
// throughput exceeds splitting_threshold continues at the threshold and queues copies of its
// state (throughput = threshold, length 0, no source sum) in a per-lane private queue; the
// estimate of the original walk is the sum over its parts.  When the queue cannot take the
// copies the walk is not split (an unsplit walk is just as unbiased).
// GRAD = true runs estimateSolutionAndGradient (walk_on_spheres.h:416-593) for the points asking
// for it: walk w of a point is half (w & 1) of antithetic pair w >> 1 (or pair w without
// antithetic variates).  The first step is explicit: a ball of radius 0.99 * d(x0), a source
// sample and a boundary point from the point's stratified table (directions) and the pair's
// generator seedWalk(seed, p, stream_offset + pair) (source radius, cosine-sampling
// sign), mirrored through x0 for the second half, which therefore continues with exactly the
// random numbers of the first (the reference replays its saved generator).  The walk then goes
// on from the boundary point as a plain walk; its estimate feeds the solution statistics and the
// gradient estimate (boundary term - cv) * grad P / P + (first source - cv) * grad G / (pdf |G|).
template <int D, int W, bool COUNT, bool BATCH, bool SCREENED, bool SPLIT = false, bool GRAD = false>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WosOccupancy<W>::kWaves, 8)))
wosKernel(WosParams P) {
    unsigned long long spill[kSpillStack];
    Stack st; st.init(spill);
    VisitCount vc{0, 0};
    const zg_walk_settings& ws = P.ws;
    const float eps = ws.epsilon_shell_for_absorbing_boundary;
    const int lane = __lane_id();
    const uint32_t streamOffset = P.streamOffsetDev ? *P.streamOffsetDev : P.streamOffset;
    // scene / PDE of the current walk: uniform, or the lane's instance in batched solves
    const InstanceDesc* inst = nullptr;
    WalkView view{P.scene.abs.nodes, P.scene.abs.prims, P.scene.abs.nNodes};
    float sigma = P.pde.absorption_coeff;
    // the PDE comes from the kernel arguments: walks here are cheap enough per step that reading
    // it from device memory at every use (as the walk-on-stars kernel does) cost 3.7x on 2D scenes
    auto pde = [&]() -> const zg_pde& { return BATCH ? inst->pde : P.pde; };

    // lane state: 0 idle, 1 step pending (dist holds the radius at x), 2 distance query in flight,
    // 3 the claim's next walk is to be started
    int mode = 0;
    int pendingCode = -1;   // >= 0: the walk ends with this code once the query (projection) is done
    int p = 0;
    int w = 0, wEnd = 0;
    Vec<D> x = vzero<D>(), yc = vzero<D>();
    int ycPrim = -1;
    bool ycValid = false;
    float dist = 0.0f, T = 1.0f, S = 0.0f;
    int len = 0;
    Ball<D> ball; ball.yukawa = false; ball.sqrtLambda = __builtin_sqrtf(sigma);
    Pcg32 rng; rng.state = 0; rng.inc = 1;
    // splitting: parts of the current walk still to run, the running sum of its parts
    float4 splitQ[SPLIT ? kSplitQueue : 1];
    int nq = 0, nSplits = 0;
    float partSum = 0.0f;
    bool partOk = false;
    // gradient walks: boundary / source gradient directions, first-step source term, control variates
    Vec<D> gB = vzero<D>(), gS = vzero<D>();
    float fsc = 0.0f, cvB = 0.0f, cvS = 0.0f;
    ClosestQuery<D> q;
    q.begin(FLT_MAX, 1.0f, 0.0f);
    // block counters in LDS: walks per completion code, steps, walks started
    __shared__ unsigned long long blk[6];
    if (threadIdx.x < 6) blk[threadIdx.x] = 0;
    __syncthreads();
    uint32_t queries = 0;

    VisitCount cvc{0, 0};   // countMode 2: visits of the exact traversal from the root
    auto beginQuery = [&](float bound2, float relax2, int code) {
        if (COUNT && P.countMode == 2) {
            // the algorithmic work of this query: an exact nearest-first traversal of the packed
            // BVH from the root, counted, then discarded (the walk uses its own query below)
            ClosestQuery<D> cq;
            cq.begin(FLT_MAX, 1.0f, 0.0f);
            st.sp = 0; st.ovf = 0;
            while (!cq.template visit<W, true>(view, x, st, cvc)) {}
        }
        q.begin(bound2, relax2, P.guard2);
        q.cp = yc;
        st.sp = 0; st.ovf = 0;
        pendingCode = code;
        mode = 2;
    };

    // first step of a gradient walk (walk_on_spheres.h:453-531); ends with the exact distance
    // query at the boundary point
    auto startGradWalk = [&]() {
        const int pair = P.gradHalves == 2 ? (w >> 1) : w;
        const bool mirror = P.gradHalves == 2 && (w & 1);
        const Vec<D> c0 = vload<D>(P.pt + (size_t)p * D);
        seedWalk(rng, P.seed, (uint32_t)p, streamOffset + (uint32_t)pair);
        T = 1.0f; S = 0.0f; len = 0;
        ball.yukawa = SCREENED && sigma > 0.0f && ws.steps_before_applying_tikhonov == 0;
        ball.update(0.99f * P.dist0[p]);
        const float* tab = P.strat + (size_t)p * (size_t)P.stratStride + (size_t)(D - 1) * 2 * (size_t)pair;
        fsc = 0.0f; gS = vzero<D>();
        if (!ws.ignore_source_contribution) {
            float gNorm = ball.norm();
            if (pde().is_source_constant) {
                fsc = gNorm * evalField<D>(pde().source, c0, vzero<D>(), false, sigma);
            } else {
                float pdf;
                Vec<D> sd = sphereFromU<D>(tab);
                float r = ball.sampleRadiusPdf(rng, pdf);
                Vec<D> ys = c0 + sd * r;
                if (mirror) ys = c0 - (ys - c0);
                fsc = gNorm * evalField<D>(pde().source, ys, vzero<D>(), false, sigma);
                gS = (ys - c0) * (ball.gradientNorm(r) / (pdf * gNorm));
            }
            S += fsc;
        }
        Vec<D> bd;
        float bpdf;
        if (ws.use_cosine_sampling_for_derivatives) {
            bd = hemisphereCosineFromU<D>(tab + (D - 1));
            if (rng.nextFloat() < 0.5f) bd[D - 1] = -bd[D - 1];
            bpdf = 0.5f * pdfHemisphereCosine<D>(fabsf(bd[D - 1]));
            Vec<D> nd = vzero<D>();
            nd[0] = 1.0f;
            if (P.dirDeriv) nd = vload<D>(P.dirDeriv + (size_t)p * D);
            bd = fromFrame(nd, bd);
        } else {
            bd = sphereFromU<D>(tab + (D - 1));
            bpdf = D == 2 ? 1.0f / (2.0f * kPi) : 1.0f / (4.0f * kPi);
        }
        Vec<D> yb = c0 + bd * ball.R;
        if (mirror) yb = c0 - (yb - c0);
        x = yb;
        T *= ball.poissonKernel() / bpdf;
        gB = (yb - c0) * (ball.poissonKernelGradientScale() / (bpdf * T));
        cvB = 0.0f; cvS = 0.0f;
        if (ws.use_gradient_control_variates) {
            int64_t N = P.cvN[p];
            cvB = (float)P.cvMean[p];
            cvS = (float)(P.cvFirst[p] / (double)(N > 1 ? N : 1));
        }
        ycValid = false; ycPrim = -1;
        if (SPLIT) { nq = 0; nSplits = 0; partSum = 0.0f; partOk = false; }
        atomicAdd(&blk[5], 1ull);
        beginQuery(FLT_MAX, 1.0f, -1);
    };

    // FP64 per-point accumulation of one walk's estimate (the sum over its parts when split)
    // Walks that end in the step section (where the lanes of a wave take their steps together)
    // defer their statistics: the lanes that ended a walk of the same point combine them with a
    // wave reduction and one lane adds them to the point's sums.  A device-scope atomic costs
    // far more memory traffic than its 8 bytes, and the lanes of a wave mostly walk one point
    // (claims are consecutive walks).  Elsewhere (terminal projections, gradient walks) the
    // atomics are issued at once.  The deferred values live only until the end of the step
    // section, so they add no register pressure to the traversal.
    bool deferStats = false;
    bool dHas = false;
    float dVal = 0.0f;
    unsigned long long dLen = 0;
    auto flushDeferred = [&]() {   // all lanes of the wave active
        uint64_t todo = __ballot(dHas || dLen > 0);
        while (todo) {   // one round per distinct point among the deferring lanes
            const int leader = __ffsll((unsigned long long)todo) - 1;
            const int pl = __shfl(p, leader);
            const bool mine = (dHas || dLen > 0) && p == pl;
            todo &= ~__ballot(mine);
            double sm = mine && dHas ? (double)dVal : 0.0;
            double sq = sm * sm;
            unsigned long long nn = mine && dHas ? 1ull : 0ull;
            unsigned long long ll = mine ? dLen : 0ull;
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) {
                sm += __shfl_xor(sm, off);
                sq += __shfl_xor(sq, off);
                nn += __shfl_xor(nn, off);
                ll += __shfl_xor(ll, off);
            }
            if (lane == leader) {
                if (nn) {
                    atomicAdd(&P.accSum[pl], sm);
                    atomicAdd(&P.accSq[pl], sq);
                    atomicAdd(&P.accN[pl], nn);
                }
                if (ll) atomicAdd(&P.accLen[pl], ll);
            }
        }
        dHas = false;
        dLen = 0;
    };
    auto accumulate = [&](float total) {
        if (!GRAD && deferStats) { dHas = true; dVal = total; return; }
        double c = (double)total;
        atomicAdd(&P.accSum[p], c);
        atomicAdd(&P.accSq[p], c * c);
        atomicAdd(&P.accN[p], 1ull);
        if (GRAD) {
            // gradient estimate and directional derivative (walk_on_spheres.h:566-586)
            float bc = total - fsc;
            Vec<D> nd = vzero<D>();
            nd[0] = 1.0f;
            if (P.dirDeriv) nd = vload<D>(P.dirDeriv + (size_t)p * D);
            float dd = 0.0f;
#pragma unroll
            for (int i = 0; i < D; i++) {
                float bi = (bc - cvB) * gB[i];
                float si = (fsc - cvS) * gS[i];
                dd += bi * nd[i];
                dd += si * nd[i];
                double e = (double)(bi + si);
                atomicAdd(&P.accG[(size_t)p * D + i], e);
                atomicAdd(&P.accG2[(size_t)p * D + i], e * e);
            }
            atomicAdd(&P.accFirst[p], (double)fsc);
            atomicAdd(&P.accDeriv[p], (double)dd);
        }
    };

    auto startWalk = [&]() {
        if (GRAD) { startGradWalk(); return; }
        if (BATCH) {
            inst = P.instances + P.instanceOf[p];
            view.nodes = inst->nodes; view.prims = inst->prims; view.nNodes = inst->nNodes;
            sigma = inst->pde.absorption_coeff;
            ball.sqrtLambda = __builtin_sqrtf(sigma);
        }
        x = vload<D>(P.pt + (size_t)p * D);
        dist = P.dist0[p];
        T = 1.0f; S = 0.0f; len = 0;
        ball.yukawa = SCREENED && sigma > 0.0f && ws.steps_before_applying_tikhonov == 0;
        ycValid = false;
        seedWalk(rng, P.seed, (uint32_t)p, streamOffset + (uint32_t)w);
        if (SPLIT) { nq = 0; nSplits = 0; partSum = 0.0f; partOk = false; }
        atomicAdd(&blk[5], 1ull);
        if (dist > eps) { mode = 1; pendingCode = -1; }
        else beginQuery(FLT_MAX, 1.0f, ZG_REACHED_ABSORBING_BOUNDARY);  // walk ends at once: project
    };

    auto finishWalk = [&](int code) {
        atomicAdd(&blk[code], 1ull);
        if (code != ZG_ESCAPED_DOMAIN) {
            float g = 0.0f;
            if ((code == ZG_REACHED_ABSORBING_BOUNDARY || code == ZG_EXCEEDED_MAX_WALK_LENGTH) &&
                !ws.ignore_absorbing_boundary_contribution) {
                Vec<D> nrm = vzero<D>();
                bool aligned = false;
                if (ws.solve_double_sided && ycPrim >= 0) {
                    Prim<D> pr; pr.load(view.prims, ycPrim);
                    nrm = pr.normal();
                    aligned = dot(x - yc, nrm) > 0.0f;
                }
                g = evalField<D>(pde().dirichlet, yc, nrm, aligned, sigma);
            }
            if (SPLIT) {
                partSum += T * g + S;
                partOk = true;
            } else {
                accumulate(T * g + S);
            }
            if (!GRAD && deferStats) dLen += (unsigned long long)len;
            else atomicAdd(&P.accLen[p], (unsigned long long)len);
            atomicAdd(&blk[4], (unsigned long long)len);
        }
        if (SPLIT) {
            if (nq > 0) {
                // next part of the same walk: its state, an exact distance query at its point
                nq--;
                float4 e = splitQ[nq];
                x[0] = e.x; x[1] = e.y;
                if (D == 3) x[D - 1] = e.z;
                T = e.w; S = 0.0f; len = 0;
                ball.yukawa = SCREENED && sigma > 0.0f && ws.steps_before_applying_tikhonov == 0;
                ycValid = false; ycPrim = -1;
                nSplits++;
                beginQuery(FLT_MAX, 1.0f, -1);
                return;
            }
            if (partOk) {
                accumulate(partSum);
                if (nSplits) atomicAdd(&P.accSplits[p], (unsigned long long)nSplits);
            }
        }
        w++;
        mode = w < wEnd ? 3 : 0;   // 3: the claim's next walk starts at the top of the loop
    };

    // one walk step of a lane whose distance is known (mode 1)
    auto stepWalk = [&]() {
        if (dist <= eps) { finishWalk(ZG_REACHED_ABSORBING_BOUNDARY); return; }
        float R = dist;
        Vec<D> c = x;
        ball.update(R);
        if (!ws.ignore_source_contribution) {
            float f;
            if (pde().is_source_constant) {
                f = evalField<D>(pde().source, x, vzero<D>(), false, sigma);
            } else {
                Vec<D> sdir = sampleUnitSphere<D>(rng);
                float r = ball.sampleRadius(rng);
                f = evalField<D>(pde().source, c + sdir * r, vzero<D>(), false, sigma);
            }
            S += T * ball.norm() * f;
        }
        Vec<D> dir = sampleUnitSphere<D>(rng);
        x = c + dir * R;
        if (len == 0 && P.cvGrad) {
            // first-step control variate: E[dir] = 0, so -grad c(x0).(x1 - x0) has zero mean
            S -= dot(vload<D>(P.cvGrad + (size_t)p * D), dir) * R;
        } else if (P.cvGrid) {
            // later steps: -T G(x_k).(x_{k+1} - x_k) with G a fixed field, zero mean as well
            S -= T * dot(cvField<D>(P, c), dir) * R;
        }
        if (SCREENED && ball.yukawa) T *= ball.dampening(fmaxf(1e-4f, __builtin_sqrtf(norm2(x - c))));
        if (outsideDomain<D>(BATCH ? inst->dmin : P.scene.dmin, BATCH ? inst->dmax : P.scene.dmax, x)) {
            finishWalk(ZG_ESCAPED_DOMAIN);
            return;
        }
        if (SPLIT && T > ws.splitting_threshold) {  // applyWeightWindow :198-213
            const float thr = ws.splitting_threshold;
            float left = T - thr;
            int copies = 0;
            for (float l = left; l > thr && copies <= kSplitQueue; l -= thr) copies++;
            if (nq + copies + 1 <= kSplitQueue) {
                T = thr;
                float4 e = make_float4(x[0], x[1], D == 3 ? x[D - 1] : 0.0f, thr);
                while (left > thr) { left -= thr; splitQ[nq++] = e; }
                if (rng.nextFloat() < left / thr) splitQ[nq++] = e;
            }
        } else if (T < ws.russian_roulette_threshold) {  // applyWeightWindow :214-223
            float survival = T / ws.russian_roulette_threshold;
            if (survival < rng.nextFloat()) { T = 0.0f; finishWalk(ZG_TERMINATED_WITH_RUSSIAN_ROULETTE); return; }
            T = ws.russian_roulette_threshold;
        }
        len++;
        if (len >= ws.max_walk_length) {
            if (ws.ignore_absorbing_boundary_contribution) finishWalk(ZG_EXCEEDED_MAX_WALK_LENGTH);
            else beginQuery(FLT_MAX, 1.0f, ZG_EXCEEDED_MAX_WALK_LENGTH);  // terminal projection, exact
            return;
        }
        if (SCREENED && sigma > 0.0f && ws.steps_before_applying_tikhonov == len) ball.yukawa = true;
        // next radius: the new distance is at most |x - y_prev| (y_prev lies on the boundary)
        float bound2;
        if (ycValid) bound2 = norm2(x - yc) * 1.0001f + 1e-30f;
        else { float r2 = 2.0f * R; bound2 = r2 * r2 * 1.0001f + 1e-30f; }
        beginQuery(bound2, P.relax2, -1);
        if (ycValid && ycPrim >= 0) {
            // the previous closest triangle is usually close to the new closest one: its exact
            // distance seeds the bound, so the traversal from the root prunes early
            Prim<D> pr; pr.load(view.prims, ycPrim);
            Vec<D> qq;
            q.seed(qq, ycPrim, pr.closest2(x, qq));
            if (COUNT) vc.prims++;
        }
    };

    bool workLeft = true;
    int iter = 0;
    for (;;) {
        // -------- regroup: refill idle lanes with new claims --------
        if (iter % P.regroup == 0 || !__any(mode != 0)) {
            for (int round = 0; round < 4; round++) {
                uint64_t idle = __ballot(mode == 0);
                if (!workLeft || idle == 0) break;
                uint32_t nIdle = (uint32_t)__popcll(idle);
                int leader = __ffsll((unsigned long long)idle) - 1;
                unsigned long long base = 0;
                if (lane == leader) base = atomicAdd(P.claimCounter, (unsigned long long)nIdle);
                base = __shfl(base, leader);
                if (base + nIdle >= P.nClaims) workLeft = false;
                if (mode == 0) {
                    unsigned long long c = base + mbcnt64(idle);
                    if (c < P.nClaims) {
                        p = (int)(c / (unsigned long long)P.chunksPerPoint);
                        int k = (int)(c - (unsigned long long)p * P.chunksPerPoint);
                        int nw = P.nWalks[p];
                        if (GRAD) {
                            // walks of the stage: [gradW0, min(gradW1, halves * pairs))
                            int pairs = P.gradHalves == 2 ? max(1, nw / 2) : nw;
                            nw = P.quantity[p] == ZG_ESTIMATE_SOLUTION_AND_GRADIENT ? min(pairs * P.gradHalves, P.gradW1) : 0;
                            w = P.gradW0 + k * P.wpl;
                        } else {
                            uint8_t ty = P.type ? P.type[p] : (uint8_t)ZG_SAMPLE_IN_DOMAIN;
                            if (ty == ZG_SAMPLE_ON_ABSORBING_BOUNDARY) nw = 0;
                            else if (P.dist0[p] <= eps) nw = min(nw, 1);
                            w = k * P.wpl;
                        }
                        if (BATCH && (uint32_t)P.instanceOf[p] >= (uint32_t)P.nInstances) {
                            nw = 0;  // bad instance index: no walks, error flag for the host
                            if (k == 0) atomicOr(reinterpret_cast<unsigned int*>(&P.counters[kCntError]), 2u);
                        }
                        wEnd = min(w + P.wpl, nw);
                        if (w < wEnd) mode = 3;
                    }
                }
            }
            if (!workLeft && !__any(mode != 0)) break;
        }
        iter++;
        if (mode == 3) startWalk();

        // -------- one node of every distance query in flight --------
        if (mode == 2 && q.template visit<W, COUNT>(view, x, st, vc)) {
            queries++;
            if (q.found) { yc = q.cp; ycPrim = q.prim; ycValid = true; dist = __builtin_sqrtf(q.radius2()); }
            else { dist = FLT_MAX; ycValid = false; }
            if (pendingCode >= 0) finishWalk(pendingCode);  // terminal projection done
            else mode = 1;
        }

        // -------- walk steps (walk_on_spheres.h:237-287) --------
        if (iter % P.stepInterval != 0 && __any(mode == 2)) continue;
        if (GRAD) {
            if (mode == 1) stepWalk();
        } else {
            deferStats = true;
            if (mode == 1) stepWalk();
            deferStats = false;
            flushDeferred();
        }
    }

    // -------- counters --------
    __syncthreads();
    if (threadIdx.x < 6 && blk[threadIdx.x]) atomicAdd(&P.counters[threadIdx.x], blk[threadIdx.x]);
    if (COUNT) {
        atomicAdd(&P.counters[6], (unsigned long long)(P.countMode == 2 ? cvc.nodes : vc.nodes));
        atomicAdd(&P.counters[7], (unsigned long long)(P.countMode == 2 ? cvc.prims : vc.prims));
        atomicAdd(&P.counters[8], (unsigned long long)queries);
    }
    if (st.overflow) atomicOr(reinterpret_cast<unsigned int*>(&P.counters[kCntError]), 1u);
}

// per-point batch -> caller statistics (Chan merge, common.h:268-310); sample points on the
// absorbing boundary record the boundary value once (walk_on_spheres.h:329-346)
template <int D>
__global__ void __launch_bounds__(256) mergeStatsKernel(MergeParams M) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0 && M.streamOffsetDev) *M.streamOffsetDev += M.streamAdvance;  // next replay: fresh streams
    if (i >= M.n) return;
    if (M.instances && (uint32_t)M.instanceOf[i] >= (uint32_t)M.nInstances) return;  // flagged by the walk kernel
    if (M.quantity && M.quantity[i] != (uint8_t)M.want) return;
    int64_t N1 = M.sN[i];
    bool hasPrev = N1 > 0;
    uint8_t ty = M.type ? M.type[i] : (uint8_t)ZG_SAMPLE_IN_DOMAIN;
    double sum = M.accSum[i], sq = M.accSq[i];
    int64_t nb = (int64_t)M.accN[i];
    if (M.want == ZG_ESTIMATE_SOLUTION_AND_GRADIENT) {
        // gradient points: first sphere 0.99 d (walk_on_spheres.h:430; walk on stars records
        // 0.99 min(d_absorbing, d_reflecting), :680-681), no boundary shortcut; gradient
        // components merged like the solution (common.h:288-309)
        if (M.firstSphereRadius) M.firstSphereRadius[i] = M.fsrSrc ? M.fsrSrc[i] : 0.99f * M.dist0[i];
        if (nb > 0) {
            int64_t G1 = M.gN[i];
            double N = (double)(G1 + nb);
            for (int k = 0; k < D; k++) {
                double gs = M.accG[(size_t)i * D + k], gq = M.accG2[(size_t)i * D + k];
                double meanB = gs / (double)nb;
                double m2B = fmax(0.0, gq - gs * meanB);
                double& gm = M.gMean[(size_t)i * D + k];
                double& g2 = M.gM2[(size_t)i * D + k];
                if (G1 == 0) { gm = meanB; g2 = m2B; }
                else {
                    double delta = meanB - gm;
                    gm = gm + delta * ((double)nb / N);
                    g2 = g2 + m2B + delta * delta * ((double)G1 * (double)nb / N);
                }
            }
            M.gN[i] = G1 + nb;
        }
        M.gFirst[i] += M.accFirst[i];
        M.gDeriv[i] += M.accDeriv[i];
    } else if (ty == ZG_SAMPLE_ON_ABSORBING_BOUNDARY) {
        if (!hasPrev) {
            float g = 0.0f;
            if (!M.ignoreAbsorbing) {
                Vec<D> x = vload<D>(M.pt + i * D);
                const zg_function& gd = M.instances ? M.instances[M.instanceOf[i]].pde.dirichlet : M.dirichlet;
                float sig = M.instances ? M.instances[M.instanceOf[i]].pde.absorption_coeff : M.absorption;
                g = evalField<D>(gd, x, vzero<D>(), M.solveDoubleSided && M.aligned && M.aligned[i], sig);
            }
            sum = g; sq = (double)g * g; nb = 1;
        }
        if (M.firstSphereRadius && !hasPrev) M.firstSphereRadius[i] = 0.0f;
    } else if (M.firstSphereRadius && !hasPrev && M.writeFirstSphereRadius) {
        M.firstSphereRadius[i] = M.fsrSrc ? M.fsrSrc[i] : M.dist0[i];
    }
    if (nb > 0) {
        double meanB = sum / (double)nb;
        double m2B = fmax(0.0, sq - sum * meanB);
        if (N1 == 0) {
            M.sMean[i] = meanB; M.sM2[i] = m2B;
        } else {
            double N = (double)(N1 + nb);
            double delta = meanB - M.sMean[i];
            M.sMean[i] = M.sMean[i] + delta * ((double)nb / N);
            M.sM2[i] = M.sM2[i] + m2B + delta * delta * ((double)N1 * (double)nb / N);
        }
        M.sN[i] = N1 + nb;
    }
    if (M.sLen) M.sLen[i] += (int64_t)M.accLen[i];
    if (M.sSplits && M.accSplits) M.sSplits[i] += (int64_t)M.accSplits[i];
}

// batched evaluation of a PDE data descriptor
template <int D>
__global__ void __launch_bounds__(256) evalFunctionKernel(zg_function f, int64_t n, const float* __restrict__ pts,
                                                          const float* __restrict__ normals,
                                                          const uint8_t* __restrict__ aligned, float absorption,
                                                          float* __restrict__ out) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Vec<D> x = vload<D>(pts + i * D);
    Vec<D> nr = normals ? vload<D>(normals + i * D) : vzero<D>();
    out[i] = evalField<D>(f, x, nr, aligned && aligned[i], absorption);
}

// Latin hypercube tables (generateStratifiedSamples<DIMS>, core/sampling.h:435-458), one lane per
// table: table t is drawn from the generator seedWalk(~seed, t, streamOffset).  With nWalks given, table
// p holds 2 * pairs(p) rows (at most nSamples, the rows a table has room for) for the gradient
// points only; otherwise every table has nSamples rows.
template <int DIMS>
__global__ void __launch_bounds__(256) stratifiedKernel(float* __restrict__ out, int64_t nTables, int64_t stride,
                                                        int32_t nSamples, const int32_t* __restrict__ nWalks,
                                                        const uint8_t* __restrict__ quantity, int32_t halves,
                                                        uint64_t seed, uint32_t streamOffset) {
#pragma clang fp contract(off)
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nTables) return;
    int ns = nSamples;
    if (nWalks) {
        if (quantity[t] != ZG_ESTIMATE_SOLUTION_AND_GRADIENT) return;
        int nw = nWalks[t];
        ns = min(2 * (halves == 2 ? max(1, nw / 2) : nw), nSamples);
    }
    if (ns <= 0) return;
    float* s = out + (size_t)t * (size_t)stride;
    Pcg32 rng;
    seedWalk(rng, ~seed, (uint32_t)t, streamOffset);
    const float inv = 1.0f / (float)ns;
    const float oneMinusEps = 1.0f - FLT_EPSILON;
    for (int i = 0; i < ns; i++)
        for (int j = 0; j < DIMS; j++) s[(size_t)DIMS * i + j] = fminf(((float)i + rng.nextFloat()) * inv, oneMinusEps);
    for (int j = 0; j < DIMS; j++)
        for (int i = 0; i < ns; i++) {
            int other = i + (int)rng.nextBounded((uint32_t)(ns - i));
            float a = s[(size_t)DIMS * i + j];
            s[(size_t)DIMS * i + j] = s[(size_t)DIMS * other + j];
            s[(size_t)DIMS * other + j] = a;
        }
}

// walk counts of the solution points (others 0) for the solution launch of a call with
// per-point estimation quantities
__global__ void __launch_bounds__(256) maskWalksKernel(const int32_t* __restrict__ nWalks,
                                                       const uint8_t* __restrict__ quantity, int64_t n,
                                                       int32_t* __restrict__ out) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = quantity[i] == ZG_ESTIMATE_SOLUTION ? nWalks[i] : 0;
}

__global__ void fillZeroKernel(uint64_t* __restrict__ p, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = 0;
}

// error flags of solves that return without synchronising (no report requested): the solve's
// counters are folded into this per-device word, read and cleared by deviceErrorsTakeAndClear
// (zg_synchronize)
__device__ unsigned int zgDeviceErrors;

__global__ void latchErrorsKernel(const unsigned long long* __restrict__ counters) {
    unsigned int f = (unsigned int)counters[kCntError];
    if (f) atomicOr(&zgDeviceErrors, f);
}

// ------------------------------------------------------------------------------------
// host launchers
// ------------------------------------------------------------------------------------
static inline unsigned blocksFor(int64_t n, int bs = 256) { return (unsigned)((n + bs - 1) / bs); }

hipError_t launchLatchErrors(const unsigned long long* counters, hipStream_t s) {
    hipLaunchKernelGGL(latchErrorsKernel, dim3(1), dim3(1), 0, s, counters);
    return hipGetLastError();
}

hipError_t deviceErrorsTakeAndClear(unsigned int* flags) {
    unsigned int zero = 0;
    hipError_t e = hipMemcpyFromSymbol(flags, HIP_SYMBOL(zgDeviceErrors), sizeof(unsigned int), 0,
                                       hipMemcpyDeviceToHost);
    if (e == hipSuccess && *flags) e = hipMemcpyToSymbol(HIP_SYMBOL(zgDeviceErrors), &zero, sizeof(unsigned int), 0,
                                                          hipMemcpyHostToDevice);
    return e;
}

hipError_t launchFillZero(uint64_t* p, int64_t nWords, hipStream_t s) {
    if (nWords <= 0) return hipSuccess;
    unsigned g = (unsigned)std::min<int64_t>(4096, (nWords + 255) / 256);
    hipLaunchKernelGGL(fillZeroKernel, dim3(g), dim3(256), 0, s, p, nWords);
    return hipGetLastError();
}

hipError_t launchMaskWalks(const int32_t* nWalks, const uint8_t* quantity, int64_t n, int32_t* out, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(maskWalksKernel, dim3(blocksFor(n)), dim3(256), 0, s, nWalks, quantity, n, out);
    return hipGetLastError();
}

hipError_t launchClosestPoint(const SceneView& sv, int boundary, int64_t n, const float* pts, float* dist,
                              float* closest, float* normal, int32_t* prim, hipStream_t s) {
    const BoundaryView& b = boundary ? sv.refl : sv.abs;
    if (n <= 0) return hipSuccess;
    dim3 g(blocksFor(n)), bl(256);
    if (sv.dim == 2 && sv.width == 4) hipLaunchKernelGGL((closestPointKernel<2, 4>), g, bl, kStackLdsBytes, s, b, n, pts, dist, closest, normal, prim);
    else if (sv.dim == 2) hipLaunchKernelGGL((closestPointKernel<2, 8>), g, bl, kStackLdsBytes, s, b, n, pts, dist, closest, normal, prim);
    else if (sv.width == 4) hipLaunchKernelGGL((closestPointKernel<3, 4>), g, bl, kStackLdsBytes, s, b, n, pts, dist, closest, normal, prim);
    else hipLaunchKernelGGL((closestPointKernel<3, 8>), g, bl, kStackLdsBytes, s, b, n, pts, dist, closest, normal, prim);
    return hipGetLastError();
}

hipError_t launchInsideDomain(const SceneView& sv, int64_t n, const float* pts, uint8_t* inside, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    dim3 g(blocksFor(n)), bl(256);
    if (sv.dim == 2 && sv.width == 4) hipLaunchKernelGGL((insideDomainKernel<2, 4>), g, bl, kStackLdsBytes, s, sv, n, pts, inside);
    else if (sv.dim == 2) hipLaunchKernelGGL((insideDomainKernel<2, 8>), g, bl, kStackLdsBytes, s, sv, n, pts, inside);
    else if (sv.width == 4) hipLaunchKernelGGL((insideDomainKernel<3, 4>), g, bl, kStackLdsBytes, s, sv, n, pts, inside);
    else hipLaunchKernelGGL((insideDomainKernel<3, 8>), g, bl, kStackLdsBytes, s, sv, n, pts, inside);
    return hipGetLastError();
}

hipError_t launchRay(const SceneView& sv, int boundary, int64_t n, const float* o, const float* d, const float* tmax,
                     uint8_t* hit, float* t, float* point, float* normal, hipStream_t s) {
    const BoundaryView& b = boundary ? sv.refl : sv.abs;
    if (n <= 0) return hipSuccess;
    dim3 g(blocksFor(n)), bl(256);
    if (sv.dim == 2 && sv.width == 4) hipLaunchKernelGGL((rayKernel<2, 4>), g, bl, kStackLdsBytes, s, b, n, o, d, tmax, hit, t, point, normal);
    else if (sv.dim == 2) hipLaunchKernelGGL((rayKernel<2, 8>), g, bl, kStackLdsBytes, s, b, n, o, d, tmax, hit, t, point, normal);
    else if (sv.width == 4) hipLaunchKernelGGL((rayKernel<3, 4>), g, bl, kStackLdsBytes, s, b, n, o, d, tmax, hit, t, point, normal);
    else hipLaunchKernelGGL((rayKernel<3, 8>), g, bl, kStackLdsBytes, s, b, n, o, d, tmax, hit, t, point, normal);
    return hipGetLastError();
}

hipError_t launchWos(const WosParams& P, int dim, int width, bool count, bool screened, bool split, int gridBlocks,
                     hipStream_t s) {
    dim3 g(gridBlocks), bl(256);
#define ZG_WOS(D, Wd, C, S, SP) hipLaunchKernelGGL((wosKernel<D, Wd, C, false, S, SP>), g, bl, kStackLdsBytes, s, P)
    if (split) {   // no visit counting with splitting
        if (dim == 2) { if (width == 4) ZG_WOS(2, 4, false, false, true); else ZG_WOS(2, 8, false, false, true); }
        else if (screened) { if (width == 4) ZG_WOS(3, 4, false, true, true); else ZG_WOS(3, 8, false, true, true); }
        else { if (width == 4) ZG_WOS(3, 4, false, false, true); else ZG_WOS(3, 8, false, false, true); }
    } else if (dim == 2) {  // 2D walks are harmonic (screened Poisson is 3D only)
        if (width == 4) { if (count) ZG_WOS(2, 4, true, false, false); else ZG_WOS(2, 4, false, false, false); }
        else { if (count) ZG_WOS(2, 8, true, false, false); else ZG_WOS(2, 8, false, false, false); }
    } else if (screened) {
        if (width == 4) { if (count) ZG_WOS(3, 4, true, true, false); else ZG_WOS(3, 4, false, true, false); }
        else { if (count) ZG_WOS(3, 8, true, true, false); else ZG_WOS(3, 8, false, true, false); }
    } else {
        if (width == 4) { if (count) ZG_WOS(3, 4, true, false, false); else ZG_WOS(3, 4, false, false, false); }
        else { if (count) ZG_WOS(3, 8, true, false, false); else ZG_WOS(3, 8, false, false, false); }
    }
#undef ZG_WOS
    return hipGetLastError();
}

hipError_t launchWosBatch(const WosParams& P, int width, bool count, bool screened, bool split, int gridBlocks,
                          hipStream_t s) {
    dim3 g(gridBlocks), bl(256);
#define ZG_WOSB(Wd, C, S, SP) hipLaunchKernelGGL((wosKernel<3, Wd, C, true, S, SP>), g, bl, kStackLdsBytes, s, P)
    if (split) {
        if (screened) { if (width == 4) ZG_WOSB(4, false, true, true); else ZG_WOSB(8, false, true, true); }
        else { if (width == 4) ZG_WOSB(4, false, false, true); else ZG_WOSB(8, false, false, true); }
    } else if (screened) {
        if (width == 4) { if (count) ZG_WOSB(4, true, true, false); else ZG_WOSB(4, false, true, false); }
        else { if (count) ZG_WOSB(8, true, true, false); else ZG_WOSB(8, false, true, false); }
    } else {
        if (width == 4) { if (count) ZG_WOSB(4, true, false, false); else ZG_WOSB(4, false, false, false); }
        else { if (count) ZG_WOSB(8, true, false, false); else ZG_WOSB(8, false, false, false); }
    }
#undef ZG_WOSB
    return hipGetLastError();
}

hipError_t launchWosGrad(const WosParams& P, int dim, int width, bool screened, bool split, int gridBlocks,
                         hipStream_t s) {
    dim3 g(gridBlocks), bl(256);
#define ZG_WOSG(D, Wd, S, SP) hipLaunchKernelGGL((wosKernel<D, Wd, false, false, S, SP, true>), g, bl, kStackLdsBytes, s, P)
    if (dim == 2) {
        if (split) { if (width == 4) ZG_WOSG(2, 4, false, true); else ZG_WOSG(2, 8, false, true); }
        else { if (width == 4) ZG_WOSG(2, 4, false, false); else ZG_WOSG(2, 8, false, false); }
    } else if (screened) {
        if (split) { if (width == 4) ZG_WOSG(3, 4, true, true); else ZG_WOSG(3, 8, true, true); }
        else { if (width == 4) ZG_WOSG(3, 4, true, false); else ZG_WOSG(3, 8, true, false); }
    } else {
        if (split) { if (width == 4) ZG_WOSG(3, 4, false, true); else ZG_WOSG(3, 8, false, true); }
        else { if (width == 4) ZG_WOSG(3, 4, false, false); else ZG_WOSG(3, 8, false, false); }
    }
#undef ZG_WOSG
    return hipGetLastError();
}

hipError_t launchStratified(int dims, float* out, int64_t nTables, int64_t stride, int32_t nSamples,
                            const int32_t* nWalks, const uint8_t* quantity, int32_t halves, uint64_t seed,
                            uint32_t streamOffset, hipStream_t s) {
    if (nTables <= 0) return hipSuccess;
    dim3 g(blocksFor(nTables)), bl(256);
    if (dims == 1) hipLaunchKernelGGL(stratifiedKernel<1>, g, bl, 0, s, out, nTables, stride, nSamples, nWalks, quantity, halves, seed, streamOffset);
    else hipLaunchKernelGGL(stratifiedKernel<2>, g, bl, 0, s, out, nTables, stride, nSamples, nWalks, quantity, halves, seed, streamOffset);
    return hipGetLastError();
}

hipError_t launchEvalFunction(const zg_function& f, int dim, int64_t n, const float* pts, const float* normals,
                              const uint8_t* aligned, float absorption, float* out, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    dim3 g((unsigned)((n + 255) / 256)), bl(256);
    if (dim == 2) hipLaunchKernelGGL(evalFunctionKernel<2>, g, bl, 0, s, f, n, pts, normals, aligned, absorption, out);
    else hipLaunchKernelGGL(evalFunctionKernel<3>, g, bl, 0, s, f, n, pts, normals, aligned, absorption, out);
    return hipGetLastError();
}

hipError_t launchMergeStats(const MergeParams& M, int dim, hipStream_t s) {
    if (M.n <= 0) return hipSuccess;
    dim3 g(blocksFor(M.n)), bl(256);
    if (dim == 2) hipLaunchKernelGGL(mergeStatsKernel<2>, g, bl, 0, s, M);
    else hipLaunchKernelGGL(mergeStatsKernel<3>, g, bl, 0, s, M);
    return hipGetLastError();
}

}  // namespace zg
