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
// merged into the caller's statistics by mergeStatsKernel.
#include <hip/hip_runtime.h>

#include <cfloat>

#include "device.h"
#include "kernels.h"

namespace zg {

template <int D, bool SCREENED>
__global__ void __launch_bounds__(256) wosSdfKernel(WosParams P) {
    const zg_walk_settings& ws = P.ws;
    const zg_pde& pde = *P.pdeDev;   // device copy, read where used
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
        const uint8_t ty = P.type ? P.type[p] : (uint8_t)ZG_SAMPLE_IN_DOMAIN;
        if (ty == ZG_SAMPLE_ON_ABSORBING_BOUNDARY) nw = 0;
        else if (P.dist0[p] <= eps) nw = min(nw, 1);
        const int wEnd = min((k + 1) * P.wpl, nw);
        for (int w = k * P.wpl; w < wEnd; w++) {
            Vec<D> x = vload<D>(P.pt + (size_t)p * D);
            float dist = P.dist0[p];
            float T = 1.0f, S = 0.0f;
            int len = 0;
            int code = ZG_REACHED_ABSORBING_BOUNDARY;
            ball.yukawa = SCREENED && sigma > 0.0f && ws.steps_before_applying_tikhonov == 0;
            rng.seed(P.seed, ((uint64_t)p << 32) | (uint64_t)(uint32_t)(streamOffset + (uint32_t)w));
            atomicAdd(&blk[5], 1ull);
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
                if (T < ws.russian_roulette_threshold) {   // applyWeightWindow :214-223
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
            if (code == ZG_ESCAPED_DOMAIN) continue;
            float g = 0.0f;
            if ((code == ZG_REACHED_ABSORBING_BOUNDARY || code == ZG_EXCEEDED_MAX_WALK_LENGTH) &&
                !ws.ignore_absorbing_boundary_contribution) {
                // projectToAbsorbingBoundary for an SdfGrid (sdf_grid_geometric_queries.h:254-259)
                const float sd = gridValue<D>(sdf, x);
                Vec<D> n;
                sdfProject<D>(sdf, x, n);
                g = evalField<D>(pde.dirichlet, x, n, ws.solve_double_sided && sd > 0.0f, sigma);
            }
            const double v = (double)(T * g + S);
            atomicAdd(&P.accSum[p], v);
            atomicAdd(&P.accSq[p], v * v);
            atomicAdd(&P.accN[p], 1ull);
            atomicAdd(&P.accLen[p], (unsigned long long)len);
            atomicAdd(&blk[4], (unsigned long long)len);
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

hipError_t launchWosSdf(const WosParams& P, int dim, bool screened, int gridBlocks, hipStream_t s) {
    dim3 g(gridBlocks), bl(256);
    if (dim == 2) hipLaunchKernelGGL((wosSdfKernel<2, false>), g, bl, 0, s, P);
    else if (screened) hipLaunchKernelGGL((wosSdfKernel<3, true>), g, bl, 0, s, P);
    else hipLaunchKernelGGL((wosSdfKernel<3, false>), g, bl, 0, s, P);
    return hipGetLastError();
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
