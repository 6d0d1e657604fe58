// sdf_kernel.hip -- walk on spheres over an SDF-grid Dirichlet boundary for gfx950
// (utils/sdf_grid_geometric_queries.h:247-272 populateGeometricQueriesForDirichletBoundary with an
// SdfGrid; point_estimation/walk_on_spheres.h:228-414 walk / estimateSolution / terminal value).
//
// The distance to the boundary is |sdf(x)| of the trilinearly interpolated grid, one gather of
// 2^DIM grid values per step, so every step costs the same and no traversal state is kept: each
// lane claims `wpl` walks of one point at a time from a global counter and runs them to the
// end.  A walk that enters the epsilon shell (or stops at the maximum length) projects to the
// zero level set by steepest descent and takes the Dirichlet value there, on the side given by
// the sign of the SDF.  Statistics as in the walk-on-spheres kernel: FP64 per-point atomics,
// merged into the caller's statistics by mergeStatsKernel.  SPLIT = true adds walk splitting
// with the walk-on-spheres kernel's per-lane queue of parts (a walk's estimate is the sum over
// its parts).
#include <hip/hip_runtime.h>

#include <cfloat>

#include "device.h"
#include "kernels.h"

namespace zg {

// GRAD = true: estimateSolutionAndGradient (walk_on_spheres.h:416-593) for the points asking for
// it, as the mesh walk kernel runs it: the stage's walks [gradW0, gradW1) of antithetic pairs,
// an explicit first step inside the ball of radius 0.99 |sdf| (Latin hypercube directions from
// the point's table, source sample and boundary point, mirrored for the pair's second walk),
// then SDF walk-on-spheres steps from the boundary point; gradient control variates from the
// statistics; FP64 gradient, first-source and derivative sums.
template <int D, bool SCREENED, bool GRAD = false, bool SPLIT = false>
__global__ void __launch_bounds__(256) wosSdfKernel(WosParams P) {
    const zg_walk_settings& ws = P.ws;
    const zg_pde& pde = P.pde;
    const zg_function& sdf = P.sdf;
    const float eps = ws.epsilon_shell_for_absorbing_boundary;
    const float sigma = pde.absorption_coeff;
    const uint32_t streamOffset = P.streamOffsetDev ? *P.streamOffsetDev : P.streamOffset;
    __shared__ unsigned long long blk[6];
    if (threadIdx.x < 6) blk[threadIdx.x] = 0;
    __syncthreads();
    Ball<D> ball;
    ball.sqrtLambda = __builtin_sqrtf(sigma);
    Pcg32 rng;

    for (;;) {
        const unsigned long long c = atomicAdd(P.claimCounter, 1ull);
        if (c >= P.nClaims) break;
        const int p = (int)(c / (unsigned long long)P.chunksPerPoint);
        const int k = (int)(c - (unsigned long long)p * P.chunksPerPoint);
        int nw = P.nWalks[p];
        int w0 = k * P.wpl;
        if (GRAD) {
            // walks of the stage: [gradW0, min(gradW1, halves * pairs))
            const int pairs = P.gradHalves == 2 ? max(1, nw / 2) : nw;
            nw = P.quantity[p] == ZG_ESTIMATE_SOLUTION_AND_GRADIENT ? min(pairs * P.gradHalves, P.gradW1) : 0;
            w0 = P.gradW0 + k * P.wpl;
        } else {
            const uint8_t ty = P.type ? P.type[p] : (uint8_t)ZG_SAMPLE_IN_DOMAIN;
            if (ty == ZG_SAMPLE_ON_ABSORBING_BOUNDARY) nw = 0;
            else if (P.dist0[p] <= eps) nw = min(nw, 1);
        }
        const int wEnd = min(w0 + P.wpl, nw);
        for (int w = w0; w < wEnd; w++) {
            Vec<D> x = vload<D>(P.pt + (size_t)p * D);
            float dist = P.dist0[p];
            float T = 1.0f, S = 0.0f;
            int len = 0;
            ball.yukawa = SCREENED && sigma > 0.0f && ws.steps_before_applying_tikhonov == 0;
            // gradient walks: first-step source term, source / boundary gradient directions, control variates
            float fsc = 0.0f, cvB = 0.0f, cvS = 0.0f;
            Vec<D> gB = vzero<D>(), gS = vzero<D>();
            if (GRAD) {
                const int pair = P.gradHalves == 2 ? (w >> 1) : w;
                const bool mirror = P.gradHalves == 2 && (w & 1);
                const Vec<D> c0 = x;
                seedWalk(rng, P.seed, (uint32_t)p, streamOffset + (uint32_t)pair);
                ball.update(0.99f * dist);
                const float* tab = P.strat + (size_t)p * (size_t)P.stratStride + (size_t)(D - 1) * 2 * (size_t)pair;
                if (!ws.ignore_source_contribution) {
                    const float gNorm = ball.norm();
                    if (pde.is_source_constant) {
                        fsc = gNorm * evalField<D>(pde.source, c0, vzero<D>(), false, sigma);
                    } else {
                        float pdf;
                        const Vec<D> sd = sphereFromU<D>(tab);
                        const float r = ball.sampleRadiusPdf(rng, pdf);
                        Vec<D> ys = c0 + sd * r;
                        if (mirror) ys = c0 - (ys - c0);
                        fsc = gNorm * evalField<D>(pde.source, ys, vzero<D>(), false, sigma);
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
                if (ws.use_gradient_control_variates) {
                    const int64_t N = P.cvN[p];
                    cvB = (float)P.cvMean[p];
                    cvS = (float)(P.cvFirst[p] / (double)(N > 1 ? N : 1));
                }
                dist = fabsf(gridValue<D>(sdf, x));
            } else {
                seedWalk(rng, P.seed, (uint32_t)p, streamOffset + (uint32_t)w);
            }
            atomicAdd(&blk[5], 1ull);
            // splitting: parts of the walk still to run (position, throughput) and the sum over its
            // finished parts; a part starts with no source sum and length 0
            float4 splitQ[SPLIT ? kSplitQueue : 1];
            int nq = 0, nSplits = 0;
            float partSum = 0.0f;
            bool partOk = false;
            for (;;) {
                int code = ZG_REACHED_ABSORBING_BOUNDARY;
                while (dist > eps) {
                    const Vec<D> cen = x;
                    ball.update(dist);
                    if (!ws.ignore_source_contribution) {
                        float f;
                        if (pde.is_source_constant) {
                            f = evalField<D>(pde.source, cen, vzero<D>(), false, sigma);
                        } else {
                            Vec<D> sdir = sampleUnitSphere<D>(rng);
                            float r = ball.sampleRadius(rng);
                            f = evalField<D>(pde.source, cen + sdir * r, vzero<D>(), false, sigma);
                        }
                        S += T * ball.norm() * f;
                    }
                    Vec<D> dir = sampleUnitSphere<D>(rng);
                    x = cen + dir * dist;
                    if (SCREENED && ball.yukawa) T *= ball.dampening(fmaxf(1e-4f, __builtin_sqrtf(norm2(x - cen))));
                    if (outsideDomain<D>(P.scene.dmin, P.scene.dmax, x)) { code = ZG_ESCAPED_DOMAIN; break; }
                    if (SPLIT && T > ws.splitting_threshold) {   // applyWeightWindow :198-213
                        const float thr = ws.splitting_threshold;
                        float left = T - thr;
                        int copies = 0;
                        for (float l = left; l > thr && copies <= kSplitQueue; l -= thr) copies++;
                        if (nq + copies + 1 <= kSplitQueue) {   // else unsplit (just as unbiased)
                            T = thr;
                            const float4 e = make_float4(x[0], x[1], D == 3 ? x[D - 1] : 0.0f, thr);
                            while (left > thr) { left -= thr; splitQ[nq++] = e; }
                            if (rng.nextFloat() < left / thr) splitQ[nq++] = e;
                        }
                    } else if (T < ws.russian_roulette_threshold) {   // applyWeightWindow :214-223
                        const float survival = T / ws.russian_roulette_threshold;
                        if (survival < rng.nextFloat()) { T = 0.0f; code = ZG_TERMINATED_WITH_RUSSIAN_ROULETTE; break; }
                        T = ws.russian_roulette_threshold;
                    }
                    len++;
                    if (len >= ws.max_walk_length) { code = ZG_EXCEEDED_MAX_WALK_LENGTH; break; }
                    if (SCREENED && sigma > 0.0f && ws.steps_before_applying_tikhonov == len) ball.yukawa = true;
                    dist = fabsf(gridValue<D>(sdf, x));
                }
                atomicAdd(&blk[code], 1ull);
                if (code != ZG_ESCAPED_DOMAIN) {
                    float g = 0.0f;
                    if ((code == ZG_REACHED_ABSORBING_BOUNDARY || code == ZG_EXCEEDED_MAX_WALK_LENGTH) &&
                        !ws.ignore_absorbing_boundary_contribution) {
                        // projectToAbsorbingBoundary for an SdfGrid (sdf_grid_geometric_queries.h:254-259)
                        const float sd = gridValue<D>(sdf, x);
                        Vec<D> n;
                        sdfProject<D>(sdf, x, n);
                        g = evalField<D>(pde.dirichlet, x, n, ws.solve_double_sided && sd > 0.0f, sigma);
                    }
                    partSum += T * g + S;
                    partOk = true;
                    atomicAdd(&P.accLen[p], (unsigned long long)len);
                    atomicAdd(&blk[4], (unsigned long long)len);
                }
                if (!SPLIT || nq == 0) break;
                nq--;   // the next part of the same walk
                const float4 e = splitQ[SPLIT ? nq : 0];
                x[0] = e.x; x[1] = e.y;
                if (D == 3) x[D - 1] = e.z;
                T = e.w; S = 0.0f; len = 0;
                ball.yukawa = SCREENED && sigma > 0.0f && ws.steps_before_applying_tikhonov == 0;
                dist = fabsf(gridValue<D>(sdf, x));
                nSplits++;
            }
            if (!partOk) continue;
            if (SPLIT && nSplits) atomicAdd(&P.accSplits[p], (unsigned long long)nSplits);
            const float total = partSum;
            const double v = (double)total;
            atomicAdd(&P.accSum[p], v);
            atomicAdd(&P.accSq[p], v * v);
            atomicAdd(&P.accN[p], 1ull);
            if (GRAD) {
                // gradient estimate and directional derivative (walk_on_spheres.h:566-586)
                const float bc = total - fsc;
                Vec<D> nd = vzero<D>();
                nd[0] = 1.0f;
                if (P.dirDeriv) nd = vload<D>(P.dirDeriv + (size_t)p * D);
                float dd = 0.0f;
#pragma unroll
                for (int i = 0; i < D; i++) {
                    const float bi = (bc - cvB) * gB[i];
                    const float si = (fsc - cvS) * gS[i];
                    dd += bi * nd[i];
                    dd += si * nd[i];
                    const double e = (double)(bi + si);
                    atomicAdd(&P.accG[(size_t)p * D + i], e);
                    atomicAdd(&P.accG2[(size_t)p * D + i], e * e);
                }
                atomicAdd(&P.accFirst[p], (double)fsc);
                atomicAdd(&P.accDeriv[p], (double)dd);
            }
        }
    }
    __syncthreads();
    if (threadIdx.x < 6 && blk[threadIdx.x]) atomicAdd(&P.counters[threadIdx.x], blk[threadIdx.x]);
}

// |sdf|, projection to the zero level set and its normal (zg_closest_point on an SDF scene)
template <int D>
__global__ void __launch_bounds__(256) sdfQueryKernel(zg_function sdf, int64_t n, const float* __restrict__ pts,
                                                      float* __restrict__ dist, float* __restrict__ closest,
                                                      float* __restrict__ normal, int32_t* __restrict__ prim) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Vec<D> x = vload<D>(pts + (size_t)i * D);
    dist[i] = fabsf(gridValue<D>(sdf, x));
    if (prim) prim[i] = -1;   // no primitives
    if (closest || normal) {
        Vec<D> nn;
        sdfProject<D>(sdf, x, nn);
        if (closest) vstore<D>(closest + (size_t)i * D, x);
        if (normal) vstore<D>(normal + (size_t)i * D, nn);
    }
}

// first crossing of the zero level set along each ray (zg_intersect_ray on an SDF scene)
template <int D>
__global__ void __launch_bounds__(256) sdfRayKernel(zg_function sdf, int64_t n, const float* __restrict__ o,
                                                    const float* __restrict__ d, const float* __restrict__ tmax,
                                                    uint8_t* __restrict__ hit, float* __restrict__ t,
                                                    float* __restrict__ point, float* __restrict__ normal) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Vec<D> p = vzero<D>(), nn = vzero<D>();
    float th = 0.0f;
    const bool h = sdfIntersect<D>(sdf, vload<D>(o + (size_t)i * D), vload<D>(d + (size_t)i * D), tmax[i], p, nn, th);
    hit[i] = h ? 1 : 0;
    if (t) t[i] = h ? th : tmax[i];
    if (point) vstore<D>(point + (size_t)i * D, p);
    if (normal) vstore<D>(normal + (size_t)i * D, nn);
}

template <int D>
__global__ void __launch_bounds__(256) sdfInsideKernel(zg_function sdf, int64_t n, const float* __restrict__ pts,
                                                       uint8_t* __restrict__ inside) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) inside[i] = gridValue<D>(sdf, vload<D>(pts + (size_t)i * D)) < 0.0f ? 1 : 0;
}

template <bool GRAD>
static hipError_t launchSdfWalks(const WosParams& P, int dim, bool screened, bool split, int gridBlocks, hipStream_t s) {
    dim3 g(gridBlocks), bl(256);
#define ZG_SDF(Dd, S, SP) hipLaunchKernelGGL((wosSdfKernel<Dd, S, GRAD, SP>), g, bl, 0, s, P)
    if (dim == 2) { if (split) ZG_SDF(2, false, true); else ZG_SDF(2, false, false); }
    else if (screened) { if (split) ZG_SDF(3, true, true); else ZG_SDF(3, true, false); }
    else { if (split) ZG_SDF(3, false, true); else ZG_SDF(3, false, false); }
#undef ZG_SDF
    return hipGetLastError();
}

hipError_t launchWosSdf(const WosParams& P, int dim, bool screened, bool split, int gridBlocks, hipStream_t s) {
    return launchSdfWalks<false>(P, dim, screened, split, gridBlocks, s);
}

hipError_t launchWosSdfGrad(const WosParams& P, int dim, bool screened, bool split, int gridBlocks, hipStream_t s) {
    return launchSdfWalks<true>(P, dim, screened, split, gridBlocks, s);
}

hipError_t launchSdfQuery(const zg_function& sdf, int dim, int64_t n, const float* pts, float* dist, float* closest,
                          float* normal, int32_t* prim, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    dim3 g((unsigned)((n + 255) / 256)), bl(256);
    if (dim == 2) hipLaunchKernelGGL(sdfQueryKernel<2>, g, bl, 0, s, sdf, n, pts, dist, closest, normal, prim);
    else hipLaunchKernelGGL(sdfQueryKernel<3>, g, bl, 0, s, sdf, n, pts, dist, closest, normal, prim);
    return hipGetLastError();
}

hipError_t launchSdfRay(const zg_function& sdf, int dim, int64_t n, const float* o, const float* d, const float* tmax,
                        uint8_t* hit, float* t, float* point, float* normal, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    dim3 g((unsigned)((n + 255) / 256)), bl(256);
    if (dim == 2) hipLaunchKernelGGL(sdfRayKernel<2>, g, bl, 0, s, sdf, n, o, d, tmax, hit, t, point, normal);
    else hipLaunchKernelGGL(sdfRayKernel<3>, g, bl, 0, s, sdf, n, o, d, tmax, hit, t, point, normal);
    return hipGetLastError();
}

hipError_t launchSdfInside(const zg_function& sdf, int dim, int64_t n, const float* pts, uint8_t* inside, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    dim3 g((unsigned)((n + 255) / 256)), bl(256);
    if (dim == 2) hipLaunchKernelGGL(sdfInsideKernel<2>, g, bl, 0, s, sdf, n, pts, inside);
    else hipLaunchKernelGGL(sdfInsideKernel<3>, g, bl, 0, s, sdf, n, pts, inside);
    return hipGetLastError();
}

}  // namespace zg
