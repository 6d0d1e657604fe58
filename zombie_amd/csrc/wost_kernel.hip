// wost_kernel.hip -- persistent Walk-on-Stars kernel (2D and 3D, Neumann or Robin reflecting
// boundaries) for gfx950, its per-point first-sphere prologue, and the batched star-radius,
// line-of-sight and reflecting-boundary sampling queries.
//
// Work decomposition is the one of the walk-on-spheres kernel: claims of `wpl` walks of one
// point, handed to the idle lanes of a wave by ballot + mbcnt every `regroup` iterations (or
// as soon as the whole wave is idle), FP64 per-point atomics, a merge kernel for the caller's
// statistics.  One iteration advances every active lane by one full walk-on-stars step; the
// queries of a step run to completion inside it:
//   star radius     closest silhouette vertex (2D) / edge (3D) of the reflecting boundary
//   ray             first reflecting-boundary hit within the star radius
//   boundary sample stochastic descent of the reflecting BVH + line-of-sight test
//   distance        closest point of the absorbing boundary for the next step
// Per walk the lane restates WalkOnStars::walk (point_estimation/walk_on_stars.h:345-490)
// with estimateSolution's per-point rules (:518-662), computeReflectingBoundaryContribution
// (:189-250), computeSourceContribution (:252-275) and computeWalkStepThroughput (:277-307)
// for the harmonic Green's function: reflectance 1 under Neumann conditions, 1 - kappa G / P
// under Robin conditions (the star radius then also respects the Robin bound of every segment
// or triangle and shrinks to the reflecting epsilon on the Robin boundary).
#include <hip/hip_runtime.h>

#include <cfloat>

#include "device.h"
#include "kernels.h"

namespace zg {

// star radius of a walk: closest silhouette within [epsR, distA], shrunk by 1 % for numerical
// robustness when the reflecting boundary is not empty (walk_on_stars.h:370-392,576-590)
template <int D, int W, bool ROBIN = true>
ZG_DEV float wostStarRadius(const SceneView& sc, const zg_walk_settings& ws, const Vec<D>& x, float distA, bool flip,
                            Stack& st) {
    float r = distA;
    if (sc.refl.nPrims > 0) {
        const float minR = ws.epsilon_shell_for_reflecting_boundary;
        if constexpr (D == 2) {
            r = starRadius2<W>(sc.refl, x, minR, distA, ws.silhouette_precision, flip, st);
        } else {
            r = starRadius3<W>(sc.refl, x, minR, distA, ws.silhouette_precision, flip, st);
        }
        // Robin primitives bound the radius too (fcpw_geometric_queries.h:1272-1289)
        if (ROBIN && sc.refl.robinMax && minR <= distA)
            r = fmaxf(fminf(r, __builtin_sqrtf(robinRadius2<D, W>(sc.refl, x, distA < FLT_MAX ? distA * distA : FLT_MAX, st))), minR);
        if (minR <= distA) r = fmaxf(0.99f * r, minR);
    }
    return r;
}

// PDE::hasNonZeroRobinCoeff (pde.h:62-78)
template <int D>
ZG_DEV bool hasNonZeroRobinCoeff(const zg_pde& pde, const Vec<D>& x) {
    if (pde.are_robin_conditions_pure_neumann) return false;
    Vec<D> n = vzero<D>();
    n[0] = 1.0f;
    return fabsf(evalField<D>(pde.robin_coeff, x, n, true, pde.absorption_coeff)) > 0.0f ||
           fabsf(evalField<D>(pde.robin_coeff, x, n, false, pde.absorption_coeff)) > 0.0f;
}

// first reflecting-boundary sample of a step: a point, its normal and density (DIM uniform
// numbers, as sampleReflectingBoundary draws them), then the visibility test
template <int D, int W>
ZG_DEV bool sampleReflectingD(const BoundaryView& b, const Vec<D>& x, float radius, Pcg32& rng, Vec<D>& pt, Vec<D>& nrm,
                              float& pdf, VisitCount* vc = nullptr) {
    if constexpr (D == 2) {
        float u0 = rng.nextFloat();
        float u1 = rng.nextFloat();
        return sampleReflecting2<W>(b, x, radius, u0, u1, pt, nrm, pdf, vc);
    } else {
        float u0 = rng.nextFloat();
        float u1 = rng.nextFloat();
        float u2 = rng.nextFloat();
        return sampleReflecting3<W>(b, x, radius, u0, u1, u2, pt, nrm, pdf, vc);
    }
}
template <int D, int W>
ZG_DEV bool lineOfSightD(const BoundaryView& b, const Vec<D>& p1, const Vec<D>& p2, Stack& st) {
    if constexpr (D == 2) return lineOfSight2<W>(b, p1, p2, st);
    else return lineOfSight3<W>(b, p1, p2, st);
}
// point where the ray o + t dir meets reflecting primitive i: the parametric point on the
// segment in 2D (as the reference reports hits), o + t dir on a triangle
template <int D>
ZG_DEV Vec<D> hitPointD(const float* prims, int i, const Vec<D>& o, const Vec<D>& dir, float t) {
    if constexpr (D == 2) return segmentHitPoint(prims, i, o, dir, t);
    else return o + dir * t;
}

// first sphere radius per point (estimateSolution :566-590); a point with earlier estimates
// keeps the caller's recorded radius
template <int D, int W>
__global__ void __launch_bounds__(256) wostFirstSphereKernel(WosParams P, const float* __restrict__ callerFsr,
                                                             float* __restrict__ fsr) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.nPoints) return;
    unsigned long long spill[kSpillStack];
    Stack st; st.init(spill);
    const zg_walk_settings& ws = P.ws;
    uint8_t ty = P.type ? P.type[i] : (uint8_t)ZG_SAMPLE_IN_DOMAIN;
    float d = P.dist0[i];
    float r;
    if (ty == ZG_SAMPLE_ON_ABSORBING_BOUNDARY) r = 0.0f;
    else if (callerFsr && P.prevN && P.prevN[i] > 0) r = callerFsr[i];
    else if (d <= ws.epsilon_shell_for_absorbing_boundary || ws.steps_before_using_maximal_spheres == 0) r = d;
    else if (ty == ZG_SAMPLE_ON_REFLECTING_BOUNDARY && hasNonZeroRobinCoeff<D>(P.pde, vload<D>(P.pt + D * i)))
        r = ws.epsilon_shell_for_reflecting_boundary;
    else {
        bool flip = ws.solve_double_sided && ty == ZG_SAMPLE_ON_REFLECTING_BOUNDARY && P.aligned && P.aligned[i];
        r = wostStarRadius<D, W>(P.scene, ws, vload<D>(P.pt + D * i), d, flip, st);
    }
    fsr[i] = r;
}

// first ball of the gradient walks of walk on stars (walk_on_stars.h:678-681): radius
// 0.99 min(d_absorbing, d_reflecting), the reflecting distance taken from the caller or queried
template <int D, int W>
__global__ void __launch_bounds__(256) wostGradFirstSphereKernel(WosParams P, float* __restrict__ fsr) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P.nPoints || P.quantity[i] != ZG_ESTIMATE_SOLUTION_AND_GRADIENT) return;
    unsigned long long spill[kSpillStack];
    Stack st; st.init(spill);
    float dR = FLT_MAX;
    if (P.distR) {
        dR = P.distR[i];
    } else if (P.scene.refl.nPrims > 0) {
        VisitCount vc{0, 0};
        float b2 = FLT_MAX;
        Vec<D> cp = vzero<D>();
        int prim = -1;
        if (closestPoint<D, W, false>(P.scene.refl, vload<D>(P.pt + D * i), b2, cp, prim, st, vc)) dR = __builtin_sqrtf(b2);
    }
    fsr[i] = 0.99f * fminf(P.dist0[i], dR);
}

// 2D: 5 waves per SIMD (93 VGPRs, no VGPR spill) with the PDE read from its device copy (the
// kernel-argument copy held 106 VGPRs, 4 waves, and ran 13% slower on the 2D mixed problem).
// GRAD = true: estimateSolutionAndGradient (walk_on_stars.h:664-848) for the points asking for
// it -- the explicit first step of the walk-on-spheres gradient walks inside the first ball (no
// boundary in it), then walk-on-stars steps from the boundary point.
// ROBIN = true compiles the Robin reflectance and star-radius bound in (non-Neumann PDEs).
// SPLIT = true adds walk splitting (applyWeightWindow, walk_on_stars.h:309-343): a walk whose
// throughput exceeds splitting_threshold continues at the threshold and queues copies of its
// state (position, normal, previous direction and distance, on-boundary flag; throughput = the
// threshold, length 0, no contributions yet) in a per-lane private queue; each part restarts
// with an exact absorbing distance and a fresh Green's function (estimateSolution :600-650), and
// the walk's estimate is the sum over its parts.  When the queue cannot take the copies the walk
// is not split (an unsplit walk is just as unbiased).
template <int D, int W, bool COUNT, bool GRAD = false, bool ROBIN = false, bool SPLIT = false>
__global__ void __launch_bounds__(256) wostKernel(WosParams P) {
    unsigned long long spill[kSpillStack];
    Stack st; st.init(spill);
    VisitCount vc{0, 0};
    const zg_walk_settings& ws = P.ws;
    const zg_pde& pde = *P.pdeDev;   // device copy, read where used
    const SceneView& sc = P.scene;
    const float epsA = ws.epsilon_shell_for_absorbing_boundary;
    const float prec = ws.silhouette_precision;
    const bool hasRefl = sc.refl.nPrims > 0;
    const bool robin = ROBIN && !pde.are_robin_conditions_pure_neumann;
    const float sigma = pde.absorption_coeff;
    const bool screened = D == 3 && sigma > 0.0f;   // Yukawa ball of the 3D screened equation
    const int lane = __lane_id();
    const uint32_t streamOffset = P.streamOffsetDev ? *P.streamOffsetDev : P.streamOffset;

    __shared__ unsigned long long blk[6];
    if (threadIdx.x < 6) blk[threadIdx.x] = 0;
    __syncthreads();

    // walk state (WalkState, common.h:93-130)
    bool active = false;
    int p = 0, w = 0, wEnd = 0;
    Vec<D> x = vzero<D>(), nrm = vzero<D>(), prevDir = vzero<D>();
    float prevDist = FLT_MAX, T = 1.0f, Sref = 0.0f, Ssrc = 0.0f, distA = 0.0f, fsr0 = 0.0f;
    int len = 0;
    int absPrim = -1;   // closest absorbing primitive of the last distance query (seeds the next)
    bool onRefl = false, flip = false, firstStep = true;
    Ball<D> ball; ball.yukawa = false; ball.sqrtLambda = __builtin_sqrtf(fmaxf(sigma, 0.0f)); ball.R = 0.0f;
    Pcg32 rng; rng.state = 0; rng.inc = 1;
    // gradient walks: boundary / source gradient directions, first-step source term, control variates
    Vec<D> gB = vzero<D>(), gS = vzero<D>();
    float fsc = 0.0f, cvB = 0.0f, cvS = 0.0f;
    // splitting: parts of the current walk still to run (x, normal, previous direction, previous
    // distance with the on-boundary flag in its sign bit), the running sum of its parts
    struct Part { float v[3 * D + 1]; };
    Part splitQ[SPLIT ? kSplitQueue : 1];
    int nq = 0, nSplits = 0;
    float partSum = 0.0f;
    bool partOk = false;

    // first step of a gradient walk (walk_on_stars.h:705-791)
    auto startGradWalk = [&]() {
        const int pair = P.gradHalves == 2 ? (w >> 1) : w;
        const bool mirror = P.gradHalves == 2 && (w & 1);
        const Vec<D> c0 = vload<D>(P.pt + D * (size_t)p);
        seedWalk(rng, P.seed, (uint32_t)p, streamOffset + (uint32_t)pair);
        T = 1.0f; Sref = 0.0f; Ssrc = 0.0f; len = 0;
        ball.yukawa = screened && ws.steps_before_applying_tikhonov == 0;
        ball.update(P.fsr[p]);
        const float* tab = P.strat + (size_t)p * (size_t)P.stratStride + (size_t)(D - 1) * 2 * (size_t)pair;
        fsc = 0.0f; gS = vzero<D>();
        if (!ws.ignore_source_contribution) {
            float gNorm = ball.norm();
            if (pde.is_source_constant) {
                fsc = gNorm * evalField<D>(pde.source, c0, vzero<D>(), false, sigma);
            } else {
                float pdf;
                Vec<D> sd = sphereFromU<D>(tab);
                float r = ball.sampleRadiusPdf(rng, pdf);
                Vec<D> ys = c0 + sd * r;
                if (mirror) ys = c0 - (ys - c0);
                fsc = gNorm * evalField<D>(pde.source, ys, vzero<D>(), false, sigma);
                gS = (ys - c0) * (ball.gradientNorm(r) / (pdf * gNorm));
            }
            Ssrc += fsc;
        }
        Vec<D> bd;
        float bpdf;
        if (ws.use_cosine_sampling_for_derivatives) {
            bd = hemisphereCosineFromU<D>(tab + (D - 1));
            if (rng.nextFloat() < 0.5f) bd[D - 1] = -bd[D - 1];
            bpdf = 0.5f * pdfHemisphereCosine<D>(fabsf(bd[D - 1]));
            Vec<D> nd = vzero<D>();
            nd[0] = 1.0f;
            if (P.dirDeriv) nd = vload<D>(P.dirDeriv + D * (size_t)p);
            bd = fromFrame(nd, bd);
        } else {
            bd = sphereFromU<D>(tab + (D - 1));
            bpdf = D == 2 ? 1.0f / (2.0f * kPi) : 1.0f / (4.0f * kPi);
        }
        Vec<D> yb = c0 + bd * ball.R;
        if (mirror) yb = c0 - (yb - c0);
        prevDist = ball.R;
        prevDir = (yb - c0) * (1.0f / ball.R);
        x = yb;
        nrm = vzero<D>();
        onRefl = false; flip = false;
        T *= ball.poissonKernel() / bpdf;
        gB = (yb - c0) * (ball.poissonKernelGradientScale() / (bpdf * T));
        cvB = 0.0f; cvS = 0.0f;
        if (ws.use_gradient_control_variates) {
            int64_t N = P.cvN[p];
            cvB = (float)P.cvMean[p];
            cvS = (float)(P.cvFirst[p] / (double)(N > 1 ? N : 1));
        }
        // the walk goes on from the boundary point with an exact absorbing distance and a star
        // radius computed at its first step
        float b2 = FLT_MAX;
        Vec<D> cp = x;
        int prim = -1;
        bool gotA = closestPoint<D, W, COUNT>(sc.abs, x, b2, cp, prim, st, vc);
        distA = gotA ? __builtin_sqrtf(b2) : FLT_MAX;
        absPrim = gotA ? prim : -1;
        fsr0 = 0.0f;
        firstStep = true;
        atomicAdd(&blk[5], 1ull);
        active = true;
    };

    auto startWalk = [&]() {
        if (GRAD) { startGradWalk(); return; }
        uint8_t ty = P.type ? P.type[p] : (uint8_t)ZG_SAMPLE_IN_DOMAIN;
        Vec<D> n = P.normal ? vload<D>(P.normal + D * (size_t)p) : vzero<D>();
        bool al = P.aligned ? P.aligned[p] != 0 : false;
        x = vload<D>(P.pt + D * (size_t)p);
        nrm = n; prevDir = n; prevDist = FLT_MAX; flip = false;
        if (ws.solve_double_sided && ty == ZG_SAMPLE_ON_REFLECTING_BOUNDARY && al) {
            nrm = n * -1.0f; prevDir = n * -1.0f; flip = true;
        }
        onRefl = ty == ZG_SAMPLE_ON_REFLECTING_BOUNDARY;
        distA = P.dist0[p];
        fsr0 = P.fsr[p];
        firstStep = true;
        absPrim = -1;
        T = 1.0f; Sref = 0.0f; Ssrc = 0.0f; len = 0;
        ball.yukawa = screened && ws.steps_before_applying_tikhonov == 0;
        seedWalk(rng, P.seed, (uint32_t)p, streamOffset + (uint32_t)w);
        if (SPLIT) { nq = 0; nSplits = 0; partSum = 0.0f; partOk = false; }
        atomicAdd(&blk[5], 1ull);
        active = true;
    };

    auto finishWalk = [&](int code) {
        atomicAdd(&blk[code], 1ull);
        if (code != ZG_ESCAPED_DOMAIN) {
            float term = 0.0f;
            if (code == ZG_REACHED_ABSORBING_BOUNDARY && !ws.ignore_absorbing_boundary_contribution) {
                // project to the absorbing boundary (estimateSolution :627-645)
                float b2 = FLT_MAX;
                Vec<D> cp = x;
                int prim = -1;
                Vec<D> xp = x, nn = nrm;
                bool al = false;
                if (closestPoint<D, W, COUNT>(sc.abs, x, b2, cp, prim, st, vc, 1.0f, 0.0f, nullptr, absPrim)) {
                    Prim<D> pr; pr.load(sc.abs.prims, prim);
                    Vec<D> n = pr.normal();
                    float sd = dot(x - cp, n);
                    xp = cp; nn = n;
                    al = ws.solve_double_sided && sd > 0.0f;
                }
                term = evalField<D>(pde.dirichlet, xp, nn, al, sigma);
            }
            float total = T * term + Sref + Ssrc;
            if (SPLIT) {
                partSum += total;
                partOk = true;
            } else {
            double c = (double)total;
            atomicAdd(&P.accSum[p], c);
            atomicAdd(&P.accSq[p], c * c);
            atomicAdd(&P.accN[p], 1ull);
            }
            if (GRAD) {
                // gradient estimate and directional derivative (walk_on_stars.h:818-838)
                float bc = total - fsc;
                Vec<D> nd = vzero<D>();
                nd[0] = 1.0f;
                if (P.dirDeriv) nd = vload<D>(P.dirDeriv + D * (size_t)p);
                float dd = 0.0f;
#pragma unroll
                for (int i = 0; i < D; i++) {
                    float bi = (bc - cvB) * gB[i];
                    float si = (fsc - cvS) * gS[i];
                    dd += bi * nd[i];
                    dd += si * nd[i];
                    double e = (double)(bi + si);
                    atomicAdd(&P.accG[D * (size_t)p + i], e);
                    atomicAdd(&P.accG2[D * (size_t)p + i], e * e);
                }
                atomicAdd(&P.accFirst[p], (double)fsc);
                atomicAdd(&P.accDeriv[p], (double)dd);
            }
            atomicAdd(&P.accLen[p], (unsigned long long)len);
            atomicAdd(&blk[4], (unsigned long long)len);
        }
        if (SPLIT) {
            if (nq > 0) {
                // next part of the same walk: its state, an exact absorbing distance, a star
                // radius computed at its first step
                nq--;
                const Part& e = splitQ[nq];
#pragma unroll
                for (int i = 0; i < D; i++) { x[i] = e.v[i]; nrm[i] = e.v[D + i]; prevDir[i] = e.v[2 * D + i]; }
                prevDist = fabsf(e.v[3 * D]);
                onRefl = __float_as_uint(e.v[3 * D]) >> 31;
                T = ws.splitting_threshold; Sref = 0.0f; Ssrc = 0.0f; len = 0;
                ball.yukawa = screened && ws.steps_before_applying_tikhonov == 0;
                flip = false; firstStep = true; fsr0 = 0.0f; absPrim = -1;
                float b2 = FLT_MAX;
                Vec<D> cp = x;
                int prim = -1;
                bool gotA = closestPoint<D, W, COUNT>(sc.abs, x, b2, cp, prim, st, vc);
                distA = gotA ? __builtin_sqrtf(b2) : FLT_MAX;
                absPrim = gotA ? prim : -1;
                nSplits++;
                return;
            }
            if (partOk) {
                double c = (double)partSum;
                atomicAdd(&P.accSum[p], c);
                atomicAdd(&P.accSq[p], c * c);
                atomicAdd(&P.accN[p], 1ull);
                if (nSplits) atomicAdd(&P.accSplits[p], (unsigned long long)nSplits);
            }
        }
        w++;
        if (w < wEnd) startWalk();
        else active = false;
    };

    bool workLeft = true;
    int iter = 0;
    for (;;) {
        // -------- regroup: refill idle lanes with new claims --------
        if (iter % P.regroup == 0 || !__any(active)) {
            for (int round = 0; round < 4; round++) {
                uint64_t idle = __ballot(!active);
                if (!workLeft || idle == 0) break;
                uint32_t nIdle = (uint32_t)__popcll(idle);
                int leader = __ffsll((unsigned long long)idle) - 1;
                unsigned long long base = 0;
                if (lane == leader) base = atomicAdd(P.claimCounter, (unsigned long long)nIdle);
                base = __shfl(base, leader);
                if (base + nIdle >= P.nClaims) workLeft = false;
                if (!active) {
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
                            else if (P.dist0[p] <= epsA) nw = min(nw, 1);
                            w = k * P.wpl;
                        }
                        wEnd = min(w + P.wpl, nw);
                        if (w < wEnd) startWalk();
                    }
                }
            }
            if (!workLeft && !__any(active)) break;
        }
        iter++;
        if (!active) continue;

        // -------- one walk-on-stars step (walk_on_stars.h:366-486) --------
        if (!(distA > epsA)) { finishWalk(ZG_REACHED_ABSORBING_BOUNDARY); continue; }
        float R;
        if (firstStep && fsr0 > 0.0f) {
            R = fsr0;
        } else {
            flip = false;
            if (ws.solve_double_sided && onRefl && prevDist > 0.0f && dot(prevDir, nrm) < 0.0f) {
                nrm = nrm * -1.0f;
                flip = true;
            }
            if (ws.steps_before_using_maximal_spheres <= len) R = distA;
            else if (robin && onRefl && hasNonZeroRobinCoeff<D>(pde, x)) R = ws.epsilon_shell_for_reflecting_boundary;
            else R = wostStarRadius<D, W, ROBIN>(sc, ws, x, distA, flip, st);
        }
        ball.update(R);
        Vec<D> dir = sampleUnitSphere<D>(rng);
        if (onRefl && dot(nrm, dir) > 0.0f) dir = dir * -1.0f;

        // first reflecting-boundary hit within the star (intersectReflectingBoundary)
        Vec<D> o = onRefl ? offsetPoint<D>(x, nrm * -1.0f) : x;
        float tHit = R;
        int hitPrim = -1;
        bool hit = hasRefl && intersectRay<D, W>(sc.refl, o, dir, R, tHit, hitPrim, st);
        Vec<D> ipt, inrm = vzero<D>();
        float idist;
        if (hit) {
            ipt = hitPointD<D>(sc.refl.prims, hitPrim, o, dir, tHit);
            Prim<D> pr; pr.load(sc.refl.prims, hitPrim);
            inrm = pr.normal();
            idist = tHit;
        } else {
            ipt = o + dir * R;
            idist = R;
        }

        // reflecting boundary contribution (computeReflectingBoundaryContribution :189-250)
        if (hasRefl && !ws.ignore_reflecting_boundary_contribution) {
            Vec<D> bpt, bn;
            float bpdf = 0.0f;
            if (sampleReflectingD<D, W>(sc.refl, x, R, rng, bpt, bn, bpdf)) {
                Vec<D> toS = bpt - x;
                float distS = __builtin_sqrtf(norm2(toS));
                float alpha = onRefl ? 2.0f : 1.0f;
                bool al = false;
                if (ws.solve_double_sided) {
                    toS = toS * (1.0f / distS);
                    if (flip) {
                        al = true;
                    } else if (dot(toS, bn) < -prec) {
                        bool flipS = alpha > 1.0f ? dot(toS, nrm) < -prec : true;
                        if (flipS) al = true;
                    }
                }
                Vec<D> bnS = bn * (al ? -1.0f : 1.0f);
                if (bpdf > 0.0f && distS < R) {
                    Vec<D> p1 = onRefl ? offsetPoint<D>(x, nrm * -1.0f) : x;
                    Vec<D> p2 = offsetPoint<D>(bpt, bnS * -1.0f);
                    if (lineOfSightD<D, W>(sc.refl, p1, p2, st)) {
                        float G = ball.evaluate(fmaxf(1e-4f, distS));
                        float h = evalField<D>(pde.robin, bpt, bn, ws.solve_double_sided && al, sigma);
                        Sref += T * alpha * G * h / bpdf;
                    }
                }
            }
        }

        // source contribution along the walk direction (computeSourceContribution :252-275)
        if (!ws.ignore_source_contribution) {
            float r = ball.sampleRadius(rng);
            if (r <= idist) Ssrc += T * ball.norm() * evalField<D>(pde.source, x + dir * r, vzero<D>(), false, sigma);
        }

        // move; throughput (computeWalkStepThroughput :277-307): 1 for Neumann conditions and
        // the harmonic kernel, the clamped Robin reflectance 1 - kappa G / P after a boundary hit
        prevDist = idist;
        prevDir = dir;
        x = ipt;
        nrm = inrm;
        onRefl = hit;
        const bool yukawa = D == 3 && ball.yukawa;
        if (onRefl && prevDist > FLT_EPSILON) {
            if (robin || yukawa) {
                // reflectance Q - kappa G / P (distributions.h:535-543, 643-652, 971-985): Q the
                // Yukawa dampening (1 for the harmonic ball), G and P the ball's kernels without
                // their common constant (2D log(R / r), cos / r; 3D 1 / r - 1 / R or its
                // Yukawa form, cos / r^2)
                float refl = yukawa ? ball.dampening(prevDist) : 1.0f;
                if (robin) {
                    bool fl = ws.solve_double_sided && dot(prevDir, nrm) < 0.0f;
                    Vec<D> nr = fl ? nrm * -1.0f : nrm;
                    float kappa = evalField<D>(pde.robin_coeff, x, nrm, fl, sigma);
                    if (fabsf(kappa) > 0.0f) {
                        float Pk = D == 2 ? dot(nr, prevDir) / prevDist : dot(nr, prevDir) / (prevDist * prevDist);
                        float G = (D == 2 ? 2.0f * kPi : 4.0f * kPi) * ball.evaluate(prevDist);
                        refl -= kappa * G / Pk;
                    }
                }
                T *= fminf(fmaxf(refl, 0.0f), pde.are_robin_coeffs_nonnegative ? 1.0f : 2.0f);
            }
        } else if (yukawa) {
            T *= ball.dampening(fmaxf(1e-4f, prevDist));   // directionSampledPoissonKernel
        }
        if (!onRefl && outsideDomain<D>(sc.dmin, sc.dmax, x)) { finishWalk(ZG_ESCAPED_DOMAIN); continue; }
        if (SPLIT && T > ws.splitting_threshold) {  // applyWeightWindow :313-328
            const float thr = ws.splitting_threshold;
            float left = T - thr;
            int copies = 0;
            for (float l = left; l > thr && copies <= kSplitQueue; l -= thr) copies++;
            if (nq + copies + 1 <= kSplitQueue) {
                T = thr;
                Part e;
#pragma unroll
                for (int i = 0; i < D; i++) { e.v[i] = x[i]; e.v[D + i] = nrm[i]; e.v[2 * D + i] = prevDir[i]; }
                e.v[3 * D] = __uint_as_float(__float_as_uint(prevDist) | (onRefl ? 0x80000000u : 0u));
                while (left > thr) { left -= thr; splitQ[nq++] = e; }
                if (rng.nextFloat() < left / thr) splitQ[nq++] = e;
            }
        } else if (T < ws.russian_roulette_threshold) {  // applyWeightWindow :329-338
            float survival = T / ws.russian_roulette_threshold;
            if (survival < rng.nextFloat()) { T = 0.0f; finishWalk(ZG_TERMINATED_WITH_RUSSIAN_ROULETTE); continue; }
            T = ws.russian_roulette_threshold;
        }
        len++;
        if (len >= ws.max_walk_length) { finishWalk(ZG_EXCEEDED_MAX_WALK_LENGTH); continue; }
        if (screened && ws.steps_before_applying_tikhonov == len) ball.yukawa = true;
        // distance to the absorbing boundary for the next step
        float b2 = FLT_MAX;
        Vec<D> cp = x;
        int prim = -1;
        // with distance_tolerance > 0 the query returns a certified lower bound of the distance
        // (exact within 2 epsilon of the boundary): a smaller maximal star radius, never a larger one
        float rad2 = FLT_MAX;
        bool gotA = closestPoint<D, W, COUNT>(sc.abs, x, b2, cp, prim, st, vc, P.relax2, P.guard2, &rad2, absPrim);
        distA = gotA ? __builtin_sqrtf(rad2) : FLT_MAX;
        absPrim = gotA ? prim : -1;
        firstStep = false;
    }

    __syncthreads();
    if (threadIdx.x < 6 && blk[threadIdx.x]) atomicAdd(&P.counters[threadIdx.x], blk[threadIdx.x]);
    if (COUNT) {
        atomicAdd(&P.counters[kCntNodes], (unsigned long long)vc.nodes);
        atomicAdd(&P.counters[kCntPrims], (unsigned long long)vc.prims);
    }
    if (st.overflow) atomicOr(reinterpret_cast<unsigned int*>(&P.counters[kCntError]), 1u);
}

// ------------------------------------------------------------------------------------
// interleaved walk-on-stars kernel
// ------------------------------------------------------------------------------------

// Resumable queries of one walk-on-stars step, and the points of the step where a lane waits
// for the step logic (the kernel below).
enum StarsQueryKind : int {
    kQDist = 0,    // closest absorbing point for the next step (certified relaxed radius)
    kQProj = 1,    // closest absorbing point of a walk that ended there (exact)
    kQStar = 2,    // closest silhouette within the absorbing distance
    kQRay = 3,     // first reflecting-boundary hit within the star radius
    kQLos = 4,     // visibility of the reflecting-boundary sample
    kQStep = 5,    // (no query) the next step starts: the absorbing distance is known
    kQRadius = 6,  // (no query) the star radius is final
};

// Squared distances from x to the W children of silhouette-BVH node idx; children beyond
// bound2, or whose normal cone rules out a silhouette seen from x (coneMaySeeSilhouette /
// coneMaySeeSilhouette3), get FLT_MAX.
template <int D, int W>
ZG_DEV void starChildKeys(const BoundaryView& b, int idx, const Vec<D>& x, float bound2, float precision, bool cones,
                          float (&key)[W], int (&ch)[W]) {
    WideNode<D, W> nd;
    nd.load(b.silNodes, idx);
#pragma unroll
    for (int k = 0; k < W; k++) {
        float d2 = nd.dist2(k, x);
        key[k] = d2 <= bound2 ? d2 : FLT_MAX;
        ch[k] = nd.ch[k];
    }
    if (!cones) return;
    if constexpr (D == 2) {
        const float4* cn = reinterpret_cast<const float4*>(b.silCones) + (size_t)idx * W;
#pragma unroll
        for (int k = 0; k < W; k++)
            if (key[k] < FLT_MAX && !coneMaySeeSilhouette<W>(cn[k], nd.lo[0][k], nd.lo[1][k], nd.hi[0][k], nd.hi[1][k], x))
                key[k] = FLT_MAX;
    } else {
        const float4* cn = reinterpret_cast<const float4*>(b.silCones) + (size_t)idx * W * 2;
#pragma unroll
        for (int k = 0; k < W; k++) {
            if (!(key[k] < FLT_MAX)) continue;
            const float lo[3] = {nd.lo[0][k], nd.lo[1][k], nd.lo[2][k]};
            const float hi[3] = {nd.hi[0][k], nd.hi[1][k], nd.hi[2][k]};
            if (!coneMaySeeSilhouette3(cn[2 * k], cn[2 * k + 1], lo, hi, x, precision)) key[k] = FLT_MAX;
        }
    }
}

// Entry parameters t0 of the ray o + t dir (inv = 1 / dir) into the W child boxes of node idx
// for t in [0, tMax]; FLT_MAX for the children it misses.
template <int D, int W>
ZG_DEV void slabChildKeys(const float* nodes, int idx, const Vec<D>& o, const Vec<D>& inv, float tMax, float (&key)[W],
                          int (&ch)[W]) {
    WideNode<D, W> nd;
    nd.load(nodes, idx);
#pragma unroll
    for (int k = 0; k < W; k++) {
        float t0 = 0.0f, t1 = tMax;
#pragma unroll
        for (int a = 0; a < D; a++) {
            float ta = (nd.lo[a][k] - o[a]) * inv[a], tb = (nd.hi[a][k] - o[a]) * inv[a];
            t0 = fmaxf(t0, fminf(ta, tb));
            t1 = fminf(t1, fmaxf(ta, tb));
        }
        key[k] = t0 <= t1 ? t0 : FLT_MAX;
        ch[k] = nd.ch[k];
    }
}

// Walk on stars with the queries of every step interleaved across the lanes of a wave.
//
// The run-to-completion kernel above makes a wave wait for its slowest lane in each of the
// five traversals of a step.  Here every traversal is resumable: one loop iteration advances
// each lane's query in flight by one work item -- an inner BVH node (its W child keys sorted,
// the nearest held, the others pushed) or one primitive of a leaf -- whatever kind the query is
// (closest absorbing point, closest silhouette, first ray hit, visibility).  Lanes whose query
// has finished run the step logic between queries every `stepInterval` iterations (or at once
// when no lane of the wave has a query in flight), so the walk arithmetic is shared by the
// lanes that reach it together.  The walk, its random numbers and its estimates are those of
// wostKernel<D, W, false, false, ROBIN> (WalkOnStars::walk, walk_on_stars.h:345-490, with
// estimateSolution's rules :518-662); only the order in which the BVH items are visited
// differs, which leaves every query's answer unchanged.  The boundary sample's stochastic
// descent (one node per level, no stack) and the Robin radius bound run inside the step logic.
//
// Opt-in (wost_interleave_queries): on the measured scenes it is slower than the
// run-to-completion kernel.  On the 2D mixed problem (4,096 segments, 32 slits; 50,000 points x
// 128 walks) it takes 4.6 s at a step interval of 8 against 2.7 s: the lanes of a wave sit in
// up to five query kinds, so each iteration runs the node and leaf code of every kind present
// (VALU lane utilisation 9% against 14%, 517 against 262 VALU instructions per step), which
// costs more than the wait for the slowest lane it removes.  Advancing only the most common
// kind per iteration was slower still (5.1 s).  Its visit-counting variant is the one that
// counts the reflecting and silhouette BVH reads of a step.
// COUNT = true counts the BVH nodes and primitives every query reads, per BVH: absorbing
// (distance and projection), reflecting (ray, visibility, boundary sample, Robin bound) and
// silhouette (star radius).
template <int D, int W, bool ROBIN, bool COUNT = false>
__global__ void __launch_bounds__(256) wostInterleavedKernel(WosParams P) {
    unsigned long long spill[kSpillStack];
    Stack st; st.init(spill);
    VisitCount vcAbs{0, 0}, vcRefl{0, 0}, vcSil{0, 0};
    const zg_walk_settings& ws = P.ws;
    const zg_pde& pde = *P.pdeDev;
    const SceneView& sc = P.scene;
    const float epsA = ws.epsilon_shell_for_absorbing_boundary;
    const float epsR = ws.epsilon_shell_for_reflecting_boundary;
    const float prec = ws.silhouette_precision;
    const bool hasRefl = sc.refl.nPrims > 0;
    const bool cones = sc.refl.silCones != nullptr && prec <= sc.refl.silConePrecision;
    const bool robin = ROBIN && !pde.are_robin_conditions_pure_neumann;
    const float sigma = pde.absorption_coeff;
    const bool screened = D == 3 && sigma > 0.0f;
    const int lane = __lane_id();
    const uint32_t streamOffset = P.streamOffsetDev ? *P.streamOffsetDev : P.streamOffset;

    __shared__ unsigned long long blk[6];
    if (threadIdx.x < 6) blk[threadIdx.x] = 0;
    __syncthreads();

    // lane stage: no walk, the claim's next walk to start, a query in flight, a result (or a
    // step point) waiting for the step logic
    enum { kIdle = 0, kStart = 1, kQuery = 2, kReady = 3 };
    int stage = kIdle;
    int qk = kQStep;
    int p = 0, w = 0, wEnd = 0;

    // walk state (WalkState, common.h:93-130)
    Vec<D> x = vzero<D>(), nrm = vzero<D>(), prevDir = vzero<D>();
    float prevDist = FLT_MAX, T = 1.0f, Sref = 0.0f, Ssrc = 0.0f, distA = 0.0f, fsr0 = 0.0f;
    int len = 0;
    int absPrim = -1;
    bool onRefl = false, flip = false, firstStep = true;
    Ball<D> ball; ball.yukawa = false; ball.sqrtLambda = __builtin_sqrtf(fmaxf(sigma, 0.0f)); ball.R = 0.0f;
    Pcg32 rng; rng.state = 0; rng.inc = 1;
    // state of the current step carried across its queries
    float R = 0.0f, pendS = 0.0f;
    Vec<D> dir = vzero<D>(), o = vzero<D>();

    // the query in flight: current item, pruning bound (closest / silhouette: squared distance;
    // ray / visibility: ray parameter), closest-point bookkeeping (see ClosestQuery)
    int qItem = 0;
    float qKey = 0.0f, qBound = FLT_MAX, qBest2 = FLT_MAX, qLower2 = FLT_MAX, qRelax2 = 1.0f, qGuard2 = 0.0f;
    bool qHas = false, qFound = false, qRetried = false;
    int qPrim = -1;
    Vec<D> qCp = vzero<D>(), qd = vzero<D>(), qInv = vzero<D>();

    auto beginClosest = [&](int kind, float relax2, float guard2) {
        qk = kind; stage = kQuery;
        qBest2 = FLT_MAX; qRelax2 = relax2; qGuard2 = guard2;
        qBound = qBest2 > qGuard2 ? qBest2 * qRelax2 : qBest2; qLower2 = FLT_MAX;
        qItem = 0; qKey = 0.0f; qHas = true; qPrim = -1; qFound = false; qRetried = false;
        qCp = x;
        st.sp = 0; st.ovf = 0;
        if (absPrim >= 0) {
            // the previous closest primitive seeds the bound (as closestPoint's seedPrim)
            if (COUNT) vcAbs.prims++;
            Prim<D> pr; pr.load(sc.abs.prims, absPrim);
            Vec<D> qq;
            float d2 = pr.closest2(x, qq);
            if (d2 < qBest2) {
                qBest2 = d2; qCp = qq; qPrim = absPrim; qFound = true;
                qBound = qBest2 > qGuard2 ? qBest2 * qRelax2 : qBest2;
            }
        }
    };
    auto beginRayLike = [&](int kind, const Vec<D>& d, float tMax) {
        qk = kind; stage = kQuery;
        qd = d;
#pragma unroll
        for (int a = 0; a < D; a++) qInv[a] = 1.0f / d[a];
        qBound = tMax;
        qItem = 0; qKey = 0.0f; qHas = true; qPrim = -1; qFound = false;
        st.sp = 0; st.ovf = 0;
    };

    // one work item of the query in flight; true once the query is complete
    auto advance = [&]() -> bool {
        const bool closest = qk <= kQProj;
        if (!qHas) {
            int n; float dn;
            if (st.pop(n, dn)) {
                if (dn > qBound) {
                    if (closest && dn <= qBest2) qLower2 = fminf(qLower2, dn);
                    return false;
                }
                qItem = n; qKey = dn; qHas = true;
            } else {
                if (!closest || qFound || qRetried) return true;
                // nothing within the seeded bound: search again without it
                qRetried = true;
                qBest2 = FLT_MAX; qBound = FLT_MAX; qLower2 = FLT_MAX;
                qItem = 0; qKey = 0.0f; qHas = true;
            }
        }
        if (!isLeaf(qItem)) {
            float key[W];
            int ch[W];
            if (closest) {
                if (sc.abs.nNodes == 0) return true;
                childDistances<D, W>(sc.abs.nodes, qItem, x, key, ch);
                if (COUNT) vcAbs.nodes++;
            } else if (qk == kQStar) {
                starChildKeys<D, W>(sc.refl, qItem, x, qBound, prec, cones, key, ch);
                if (COUNT) vcSil.nodes++;
            } else {
                if (sc.refl.nNodes == 0) return true;
                slabChildKeys<D, W>(sc.refl.nodes, qItem, o, qInv, qBound, key, ch);
                if (COUNT) vcRefl.nodes++;
            }
#pragma unroll
            for (int k = 0; k < W; k++) {
                if (key[k] > qBound) {
                    if (closest && key[k] <= qBest2) qLower2 = fminf(qLower2, key[k]);
                    key[k] = FLT_MAX;
                }
            }
            ClosestQuery<D>::template sortAscending<W>(key, ch);
            qHas = key[0] < FLT_MAX;
            qItem = ch[0]; qKey = key[0];
            st.pushChildren<W>(key, ch);
            if (!qHas) return false;
        }
        if (isLeaf(qItem)) {
            // one primitive of the leaf
            const int first = leafFirstD(qItem), cnt = leafCountD(qItem);
            bool done = false;
            if (COUNT) {
                if (closest) vcAbs.prims++;
                else if (qk == kQStar) vcSil.prims++;
                else vcRefl.prims++;
            }
            if (closest) {
                Prim<D> pr; pr.load(sc.abs.prims, first);
                Vec<D> qq;
                float dd = pr.closest2(x, qq);
                if (dd < qBest2) {
                    qBest2 = dd; qCp = qq; qPrim = first; qFound = true;
                    qBound = qBest2 > qGuard2 ? qBest2 * qRelax2 : qBest2;
                }
            } else if (qk == kQStar) {
                float dd;
                bool sil;
                if constexpr (D == 2) {
                    const float4* c = reinterpret_cast<const float4*>(sc.refl.sil) + 2 * first;
                    float4 c0 = c[0], c1 = c[1];
                    float dx = x[0] - c0.x, dy = x[1] - c0.y;
                    dd = dx * dx + dy * dy;
                    sil = dd < qBound && isSilhouette2(c0, c1, x, prec, flip);
                } else {
                    const float4* c = reinterpret_cast<const float4*>(sc.refl.sil) + 4 * first;
                    float4 c0 = c[0], c1 = c[1];
                    Vec<3> pa, pb, cp;
                    pa[0] = c0.x; pa[1] = c0.y; pa[2] = c0.z;
                    pb[0] = c1.x; pb[1] = c1.y; pb[2] = c1.z;
                    dd = closestOnSegment3(pa, pb, x, cp);
                    sil = dd < qBound && isSilhouette3(c0, c[2], c[3], x, cp, prec, flip);
                }
                if (sil) { qBound = dd; qFound = true; }
                // a silhouette closer than the reflecting epsilon fixes the radius at epsilon
                done = qFound && qBound <= epsR * epsR;
            } else {
                Prim<D> pr; pr.load(sc.refl.prims, first);
                float t;
                if (qk == kQRay) {
                    if (pr.ray(o, qd, qBound, t) && (!qFound || t < qBound)) { qBound = t; qPrim = first; qFound = true; }
                } else if (pr.ray(o, qd, qBound, t) && t > 0.0f && t < qBound) {
                    qFound = true;   // visibility blocked
                    done = true;
                }
            }
            if (done) return true;
            qHas = cnt > 1 && qKey <= qBound;
            if (closest && cnt > 1 && !qHas && qKey <= qBest2) qLower2 = fminf(qLower2, qKey);
            qItem = ~(((first + 1) << 4) | (cnt - 2));
        }
        return false;
    };

    auto startWalk = [&]() {
        uint8_t ty = P.type ? P.type[p] : (uint8_t)ZG_SAMPLE_IN_DOMAIN;
        Vec<D> n = P.normal ? vload<D>(P.normal + D * (size_t)p) : vzero<D>();
        bool al = P.aligned ? P.aligned[p] != 0 : false;
        x = vload<D>(P.pt + D * (size_t)p);
        nrm = n; prevDir = n; prevDist = FLT_MAX; flip = false;
        if (ws.solve_double_sided && ty == ZG_SAMPLE_ON_REFLECTING_BOUNDARY && al) {
            nrm = n * -1.0f; prevDir = n * -1.0f; flip = true;
        }
        onRefl = ty == ZG_SAMPLE_ON_REFLECTING_BOUNDARY;
        distA = P.dist0[p];
        fsr0 = P.fsr[p];
        firstStep = true;
        absPrim = -1;
        T = 1.0f; Sref = 0.0f; Ssrc = 0.0f; len = 0;
        ball.yukawa = screened && ws.steps_before_applying_tikhonov == 0;
        seedWalk(rng, P.seed, (uint32_t)p, streamOffset + (uint32_t)w);
        atomicAdd(&blk[5], 1ull);
        stage = kReady; qk = kQStep;
    };

    auto accumulate = [&](int code, float term) {
        atomicAdd(&blk[code], 1ull);
        if (code != ZG_ESCAPED_DOMAIN) {
            float total = T * term + Sref + Ssrc;
            double c = (double)total;
            atomicAdd(&P.accSum[p], c);
            atomicAdd(&P.accSq[p], c * c);
            atomicAdd(&P.accN[p], 1ull);
            atomicAdd(&P.accLen[p], (unsigned long long)len);
            atomicAdd(&blk[4], (unsigned long long)len);
        }
        w++;
        stage = w < wEnd ? kStart : kIdle;
    };
    auto finishWalk = [&](int code) {
        if (code == ZG_REACHED_ABSORBING_BOUNDARY && !ws.ignore_absorbing_boundary_contribution)
            beginClosest(kQProj, 1.0f, 0.0f);   // projection to the absorbing boundary (exact)
        else
            accumulate(code, 0.0f);
    };

    bool workLeft = true;
    int iter = 0;
    for (;;) {
        // -------- regroup: refill idle lanes with new claims --------
        if (iter % P.regroup == 0 || !__any(stage != kIdle)) {
            for (int round = 0; round < 4; round++) {
                uint64_t idle = __ballot(stage == kIdle);
                if (!workLeft || idle == 0) break;
                uint32_t nIdle = (uint32_t)__popcll(idle);
                int leader = __ffsll((unsigned long long)idle) - 1;
                unsigned long long base = 0;
                if (lane == leader) base = atomicAdd(P.claimCounter, (unsigned long long)nIdle);
                base = __shfl(base, leader);
                if (base + nIdle >= P.nClaims) workLeft = false;
                if (stage == kIdle) {
                    unsigned long long c = base + mbcnt64(idle);
                    if (c < P.nClaims) {
                        p = (int)(c / (unsigned long long)P.chunksPerPoint);
                        int k = (int)(c - (unsigned long long)p * P.chunksPerPoint);
                        int nw = P.nWalks[p];
                        uint8_t ty = P.type ? P.type[p] : (uint8_t)ZG_SAMPLE_IN_DOMAIN;
                        if (ty == ZG_SAMPLE_ON_ABSORBING_BOUNDARY) nw = 0;
                        else if (P.dist0[p] <= epsA) nw = min(nw, 1);
                        w = k * P.wpl;
                        wEnd = min(w + P.wpl, nw);
                        if (w < wEnd) stage = kStart;
                    }
                }
            }
            if (!workLeft && !__any(stage != kIdle)) break;
        }
        iter++;

        // -------- one BVH work item of every query in flight --------
        if (stage == kQuery && advance()) stage = kReady;

        // -------- step logic --------
        if (iter % P.stepInterval != 0 && __any(stage == kQuery)) continue;

        if (stage == kReady && qk == kQProj) {
            // terminal projection done (estimateSolution :627-645)
            float term = 0.0f;
            Vec<D> xp = x, nn = nrm;
            bool al = false;
            if (qFound) {
                Prim<D> pr; pr.load(sc.abs.prims, qPrim);
                Vec<D> n = pr.normal();
                float sd = dot(x - qCp, n);
                xp = qCp; nn = n;
                al = ws.solve_double_sided && sd > 0.0f;
            }
            term = evalField<D>(pde.dirichlet, xp, nn, al, sigma);
            accumulate(ZG_REACHED_ABSORBING_BOUNDARY, term);
        }
        if (stage == kStart) startWalk();
        if (stage == kReady && qk == kQDist) {
            // with distance_tolerance > 0 the radius is a certified lower bound of the distance
            distA = qFound ? __builtin_sqrtf(fminf(qBest2, qLower2)) : FLT_MAX;
            absPrim = qFound ? qPrim : -1;
            firstStep = false;
            qk = kQStep;
        }
        if (stage == kReady && qk == kQStep) {
            // -------- a walk-on-stars step begins (walk_on_stars.h:366-392) --------
            if (!(distA > epsA)) {
                finishWalk(ZG_REACHED_ABSORBING_BOUNDARY);
            } else if (firstStep && fsr0 > 0.0f) {
                R = fsr0; qk = kQRadius;
            } else {
                flip = false;
                if (ws.solve_double_sided && onRefl && prevDist > 0.0f && dot(prevDir, nrm) < 0.0f) {
                    nrm = nrm * -1.0f;
                    flip = true;
                }
                if (ws.steps_before_using_maximal_spheres <= len) {
                    R = distA; qk = kQRadius;
                } else if (robin && onRefl && hasNonZeroRobinCoeff<D>(pde, x)) {
                    R = epsR; qk = kQRadius;
                } else if (!hasRefl) {
                    R = distA; qk = kQRadius;
                } else if (epsR > distA || sc.refl.nSilNodes == 0) {
                    // wostStarRadius without a traversal
                    float r = epsR > distA ? distA : fmaxf(distA, epsR);
                    if (ROBIN && sc.refl.robinMax && epsR <= distA)
                        r = fmaxf(fminf(r, __builtin_sqrtf(robinRadius2<D, W>(sc.refl, x, distA < FLT_MAX ? distA * distA : FLT_MAX, st, COUNT ? &vcRefl : nullptr))), epsR);
                    if (epsR <= distA) r = fmaxf(0.99f * r, epsR);
                    R = r; qk = kQRadius;
                } else {
                    // closest silhouette within [epsR, distA] (starRadius2 / starRadius3)
                    qk = kQStar; stage = kQuery;
                    qBound = distA < FLT_MAX ? distA * distA : FLT_MAX;
                    qItem = 0; qKey = 0.0f; qHas = true; qFound = false;
                    st.sp = 0; st.ovf = 0;
                }
            }
        }
        if (stage == kReady && qk == kQStar) {
            float r = qFound ? fmaxf(__builtin_sqrtf(qBound), epsR) : fmaxf(distA, epsR);
            // Robin primitives bound the radius too (fcpw_geometric_queries.h:1272-1289)
            if (ROBIN && sc.refl.robinMax)
                r = fmaxf(fminf(r, __builtin_sqrtf(robinRadius2<D, W>(sc.refl, x, distA < FLT_MAX ? distA * distA : FLT_MAX, st, COUNT ? &vcRefl : nullptr))), epsR);
            R = fmaxf(0.99f * r, epsR);
            qk = kQRadius;
        }
        if (stage == kReady && qk == kQRadius) {
            ball.update(R);
            dir = sampleUnitSphere<D>(rng);
            if (onRefl && dot(nrm, dir) > 0.0f) dir = dir * -1.0f;
            o = onRefl ? offsetPoint<D>(x, nrm * -1.0f) : x;
            // first reflecting-boundary hit within the star (intersectReflectingBoundary)
            if (hasRefl) beginRayLike(kQRay, dir, R);
            else { qFound = false; qBound = R; qk = kQRay; }
        }
        if (stage == kReady && qk == kQRay) {
            // the step's end point: the hit, else the star's boundary
            Vec<D> ipt, inrm = vzero<D>();
            float idist;
            const bool hit = qFound;
            if (hit) {
                ipt = hitPointD<D>(sc.refl.prims, qPrim, o, dir, qBound);
                Prim<D> pr; pr.load(sc.refl.prims, qPrim);
                inrm = pr.normal();
                idist = qBound;
            } else {
                ipt = o + dir * R;
                idist = R;
            }
            // reflecting boundary contribution (computeReflectingBoundaryContribution :189-250):
            // the sample and its weight now, its visibility by the next query
            pendS = 0.0f;
            Vec<D> p2 = o;
            bool needLos = false;
            if (hasRefl && !ws.ignore_reflecting_boundary_contribution) {
                Vec<D> bpt, bn;
                float bpdf = 0.0f;
                if (sampleReflectingD<D, W>(sc.refl, x, R, rng, bpt, bn, bpdf, COUNT ? &vcRefl : nullptr)) {
                    Vec<D> toS = bpt - x;
                    float distS = __builtin_sqrtf(norm2(toS));
                    float alpha = onRefl ? 2.0f : 1.0f;
                    bool al = false;
                    if (ws.solve_double_sided) {
                        toS = toS * (1.0f / distS);
                        if (flip) {
                            al = true;
                        } else if (dot(toS, bn) < -prec) {
                            bool flipS = alpha > 1.0f ? dot(toS, nrm) < -prec : true;
                            if (flipS) al = true;
                        }
                    }
                    Vec<D> bnS = bn * (al ? -1.0f : 1.0f);
                    if (bpdf > 0.0f && distS < R) {
                        p2 = offsetPoint<D>(bpt, bnS * -1.0f);
                        float G = ball.evaluate(fmaxf(1e-4f, distS));
                        float h = evalField<D>(pde.robin, bpt, bn, ws.solve_double_sided && al, sigma);
                        pendS = T * alpha * G * h / bpdf;
                        needLos = true;
                    }
                }
            }
            // the step's end point and normal wait in the walk state until the move
            prevDist = idist;
            prevDir = dir;
            onRefl = hit;
            nrm = inrm;
            // x stays the step's start until the source sample below; the end point is kept in qCp
            qCp = ipt;
            if (needLos) {
                Vec<D> d = p2 - o;
                float L = __builtin_sqrtf(norm2(d));
                if (L <= 0.0f || sc.refl.nNodes == 0) { qFound = false; qk = kQLos; }
                else beginRayLike(kQLos, d * (1.0f / L), L);
            } else {
                pendS = 0.0f; qFound = true; qk = kQLos;
            }
        }
        if (stage == kReady && qk == kQLos) {
            if (!qFound) Sref += pendS;   // visible
            // source contribution along the walk direction (computeSourceContribution :252-275)
            if (!ws.ignore_source_contribution) {
                float r = ball.sampleRadius(rng);
                if (r <= prevDist) Ssrc += T * ball.norm() * evalField<D>(pde.source, x + dir * r, vzero<D>(), false, sigma);
            }
            // move; throughput (computeWalkStepThroughput :277-307)
            x = qCp;
            const bool yukawa = D == 3 && ball.yukawa;
            if (onRefl && prevDist > FLT_EPSILON) {
                if (robin || yukawa) {
                    float refl = yukawa ? ball.dampening(prevDist) : 1.0f;
                    if (robin) {
                        bool fl = ws.solve_double_sided && dot(prevDir, nrm) < 0.0f;
                        Vec<D> nr = fl ? nrm * -1.0f : nrm;
                        float kappa = evalField<D>(pde.robin_coeff, x, nrm, fl, sigma);
                        if (fabsf(kappa) > 0.0f) {
                            float Pk = D == 2 ? dot(nr, prevDir) / prevDist : dot(nr, prevDir) / (prevDist * prevDist);
                            float G = (D == 2 ? 2.0f * kPi : 4.0f * kPi) * ball.evaluate(prevDist);
                            refl -= kappa * G / Pk;
                        }
                    }
                    T *= fminf(fmaxf(refl, 0.0f), pde.are_robin_coeffs_nonnegative ? 1.0f : 2.0f);
                }
            } else if (yukawa) {
                T *= ball.dampening(fmaxf(1e-4f, prevDist));   // directionSampledPoissonKernel
            }
            if (!onRefl && outsideDomain<D>(sc.dmin, sc.dmax, x)) {
                accumulate(ZG_ESCAPED_DOMAIN, 0.0f);
            } else {
                bool alive = true;
                if (T < ws.russian_roulette_threshold) {  // applyWeightWindow :329-338
                    float survival = T / ws.russian_roulette_threshold;
                    if (survival < rng.nextFloat()) { T = 0.0f; accumulate(ZG_TERMINATED_WITH_RUSSIAN_ROULETTE, 0.0f); alive = false; }
                    else T = ws.russian_roulette_threshold;
                }
                if (alive) {
                    len++;
                    if (len >= ws.max_walk_length) {
                        accumulate(ZG_EXCEEDED_MAX_WALK_LENGTH, 0.0f);
                    } else {
                        if (screened && ws.steps_before_applying_tikhonov == len) ball.yukawa = true;
                        // distance to the absorbing boundary for the next step
                        beginClosest(kQDist, P.relax2, P.guard2);
                    }
                }
            }
        }
    }

    __syncthreads();
    if (threadIdx.x < 6 && blk[threadIdx.x]) atomicAdd(&P.counters[threadIdx.x], blk[threadIdx.x]);
    if (COUNT) {
        atomicAdd(&P.counters[kCntNodes], (unsigned long long)vcAbs.nodes);
        atomicAdd(&P.counters[kCntPrims], (unsigned long long)vcAbs.prims);
        atomicAdd(&P.counters[kCntReflNodes], (unsigned long long)vcRefl.nodes);
        atomicAdd(&P.counters[kCntReflPrims], (unsigned long long)vcRefl.prims);
        atomicAdd(&P.counters[kCntSilNodes], (unsigned long long)vcSil.nodes);
        atomicAdd(&P.counters[kCntSilPrims], (unsigned long long)vcSil.prims);
    }
    if (st.overflow) atomicOr(reinterpret_cast<unsigned int*>(&P.counters[kCntError]), 1u);
}

// ------------------------------------------------------------------------------------
// batched reflecting-boundary queries
// ------------------------------------------------------------------------------------
template <int D, int W>
__global__ void __launch_bounds__(256) starRadiusKernel(SceneView sv, int64_t n, const float* __restrict__ pts,
                                                        const float* __restrict__ minR, const float* __restrict__ maxR,
                                                        float prec, const uint8_t* __restrict__ flip,
                                                        float* __restrict__ radius) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    unsigned long long spill[kSpillStack];
    Stack st; st.init(spill);
    bool fl = flip && flip[i] != 0;
    Vec<D> x = vload<D>(pts + D * i);
    float r;
    if (sv.refl.nPrims == 0) {
        r = maxR[i];
    } else {
        if constexpr (D == 2) r = starRadius2<W>(sv.refl, x, minR[i], maxR[i], prec, fl, st);
        else r = starRadius3<W>(sv.refl, x, minR[i], maxR[i], prec, fl, st);
        if (sv.refl.robinMax && minR[i] <= maxR[i]) {
            float m = maxR[i];
            r = fmaxf(fminf(r, __builtin_sqrtf(robinRadius2<D, W>(sv.refl, x, m < FLT_MAX ? m * m : FLT_MAX, st))), minR[i]);
        }
    }
    radius[i] = r;
}

template <int D, int W>
__global__ void __launch_bounds__(256) lineOfSightKernel(SceneView sv, int64_t n, const float* __restrict__ p1,
                                                         const float* __restrict__ p2, uint8_t* __restrict__ visible) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    unsigned long long spill[kSpillStack];
    Stack st; st.init(spill);
    visible[i] = lineOfSightD<D, W>(sv.refl, vload<D>(p1 + D * i), vload<D>(p2 + D * i), st) ? 1 : 0;
}

// u: DIM uniform numbers per query (primitive choice, then the point on it)
template <int D, int W>
__global__ void __launch_bounds__(256) sampleReflectingKernel(SceneView sv, int64_t n, const float* __restrict__ x,
                                                              const float* __restrict__ radius, const float* __restrict__ u,
                                                              uint8_t* __restrict__ ok, float* __restrict__ pt,
                                                              float* __restrict__ normal, float* __restrict__ pdf) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Vec<D> q = vzero<D>(), nn = vzero<D>();
    float pd = 0.0f;
    bool got;
    if constexpr (D == 2) got = sampleReflecting2<W>(sv.refl, vload<2>(x + 2 * i), radius[i], u[2 * i], u[2 * i + 1], q, nn, pd);
    else got = sampleReflecting3<W>(sv.refl, vload<3>(x + 3 * i), radius[i], u[3 * i], u[3 * i + 1], u[3 * i + 2], q, nn, pd);
    ok[i] = got ? 1 : 0;
    if (pt) vstore<D>(pt + D * i, q);
    if (normal) vstore<D>(normal + D * i, nn);
    if (pdf) pdf[i] = got ? pd : 0.0f;
}

// ------------------------------------------------------------------------------------
// launchers
// ------------------------------------------------------------------------------------
static inline unsigned wostBlocksFor(int64_t n) { return (unsigned)((n + 255) / 256); }

// one launch of a kernel template over (dim, width)
#define ZG_DW(K, dim, width, ...)                                                   \
    do {                                                                            \
        if ((dim) == 2) {                                                           \
            if ((width) == 4) hipLaunchKernelGGL((K<2, 4>), __VA_ARGS__);           \
            else hipLaunchKernelGGL((K<2, 8>), __VA_ARGS__);                        \
        } else {                                                                    \
            if ((width) == 4) hipLaunchKernelGGL((K<3, 4>), __VA_ARGS__);           \
            else hipLaunchKernelGGL((K<3, 8>), __VA_ARGS__);                        \
        }                                                                           \
    } while (0)

hipError_t launchWostFirstSphere(const WosParams& P, int dim, int width, const float* callerFsr, float* fsr,
                                 hipStream_t s) {
    if (P.nPoints <= 0) return hipSuccess;
    dim3 g(wostBlocksFor(P.nPoints)), bl(256);
    ZG_DW(wostFirstSphereKernel, dim, width, g, bl, kStackLdsBytes, s, P, callerFsr, fsr);
    return hipGetLastError();
}

hipError_t launchWostGradFirstSphere(const WosParams& P, int dim, int width, float* fsr, hipStream_t s) {
    if (P.nPoints <= 0) return hipSuccess;
    dim3 g(wostBlocksFor(P.nPoints)), bl(256);
    ZG_DW(wostGradFirstSphereKernel, dim, width, g, bl, kStackLdsBytes, s, P, fsr);
    return hipGetLastError();
}

hipError_t launchWostGrad(const WosParams& P, int dim, int width, int gridBlocks, hipStream_t s) {
    dim3 g(gridBlocks), bl(256);
    const bool robin = !P.pde.are_robin_conditions_pure_neumann;
#define ZG_WOSTG(Dd, Wd, R) hipLaunchKernelGGL((wostKernel<Dd, Wd, false, true, R>), g, bl, kStackLdsBytes, s, P)
    if (dim == 3) {
        if (width == 4) { if (robin) ZG_WOSTG(3, 4, true); else ZG_WOSTG(3, 4, false); }
        else { if (robin) ZG_WOSTG(3, 8, true); else ZG_WOSTG(3, 8, false); }
    } else if (width == 4) {
        if (robin) ZG_WOSTG(2, 4, true); else ZG_WOSTG(2, 4, false);
    } else {
        if (robin) ZG_WOSTG(2, 8, true); else ZG_WOSTG(2, 8, false);
    }
#undef ZG_WOSTG
    return hipGetLastError();
}

hipError_t launchWost(const WosParams& P, int dim, int width, bool count, bool split, bool interleave, int gridBlocks,
                      hipStream_t s) {
    dim3 g(gridBlocks), bl(256);
    const bool robin = !P.pde.are_robin_conditions_pure_neumann;
    // visit counting runs in the interleaved kernel, which counts every BVH a step reads (the
    // walks are those of the run-to-completion kernel)
    if ((interleave || count) && !split) {
#define ZG_WOSTI(Dd, Wd, R)                                                                                   \
    do {                                                                                                      \
        if (count) hipLaunchKernelGGL((wostInterleavedKernel<Dd, Wd, R, true>), g, bl, kStackLdsBytes, s, P); \
        else hipLaunchKernelGGL((wostInterleavedKernel<Dd, Wd, R, false>), g, bl, kStackLdsBytes, s, P);      \
    } while (0)
        if (dim == 3) {
            if (width == 4) { if (robin) ZG_WOSTI(3, 4, true); else ZG_WOSTI(3, 4, false); }
            else { if (robin) ZG_WOSTI(3, 8, true); else ZG_WOSTI(3, 8, false); }
        } else if (width == 4) { if (robin) ZG_WOSTI(2, 4, true); else ZG_WOSTI(2, 4, false); }
        else { if (robin) ZG_WOSTI(2, 8, true); else ZG_WOSTI(2, 8, false); }
#undef ZG_WOSTI
        return hipGetLastError();
    }
    if (split) {   // no visit counting with splitting
#define ZG_WOSTS(Dd, Wd, R) hipLaunchKernelGGL((wostKernel<Dd, Wd, false, false, R, true>), g, bl, kStackLdsBytes, s, P)
        if (dim == 3) {
            if (width == 4) { if (robin) ZG_WOSTS(3, 4, true); else ZG_WOSTS(3, 4, false); }
            else { if (robin) ZG_WOSTS(3, 8, true); else ZG_WOSTS(3, 8, false); }
        } else if (width == 4) { if (robin) ZG_WOSTS(2, 4, true); else ZG_WOSTS(2, 4, false); }
        else { if (robin) ZG_WOSTS(2, 8, true); else ZG_WOSTS(2, 8, false); }
#undef ZG_WOSTS
        return hipGetLastError();
    }
#define ZG_WOST(Dd, Wd, R) hipLaunchKernelGGL((wostKernel<Dd, Wd, false, false, R>), g, bl, kStackLdsBytes, s, P)
    if (dim == 3) {
        if (width == 4) { if (robin) ZG_WOST(3, 4, true); else ZG_WOST(3, 4, false); }
        else { if (robin) ZG_WOST(3, 8, true); else ZG_WOST(3, 8, false); }
    } else if (width == 4) {
        if (robin) ZG_WOST(2, 4, true); else ZG_WOST(2, 4, false);
    } else {
        if (robin) ZG_WOST(2, 8, true); else ZG_WOST(2, 8, false);
    }
#undef ZG_WOST
    return hipGetLastError();
}

hipError_t launchStarRadius(const SceneView& sv, int64_t n, const float* pts, const float* minR, const float* maxR,
                            float prec, const uint8_t* flip, float* radius, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    dim3 g(wostBlocksFor(n)), bl(256);
    ZG_DW(starRadiusKernel, sv.dim, sv.width, g, bl, kStackLdsBytes, s, sv, n, pts, minR, maxR, prec, flip, radius);
    return hipGetLastError();
}

hipError_t launchLineOfSight(const SceneView& sv, int64_t n, const float* p1, const float* p2, uint8_t* visible,
                             hipStream_t s) {
    if (n <= 0) return hipSuccess;
    dim3 g(wostBlocksFor(n)), bl(256);
    ZG_DW(lineOfSightKernel, sv.dim, sv.width, g, bl, kStackLdsBytes, s, sv, n, p1, p2, visible);
    return hipGetLastError();
}

hipError_t launchSampleReflecting(const SceneView& sv, int64_t n, const float* x, const float* radius, const float* u,
                                  uint8_t* ok, float* pt, float* normal, float* pdf, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    dim3 g(wostBlocksFor(n)), bl(256);
    ZG_DW(sampleReflectingKernel, sv.dim, sv.width, g, bl, 0, s, sv, n, x, radius, u, ok, pt, normal, pdf);
    return hipGetLastError();
}

#undef ZG_DW

}  // namespace zg
