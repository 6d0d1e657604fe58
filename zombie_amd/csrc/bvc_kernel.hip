// bvc_kernel.hip -- boundary value caching splat for gfx950
// (variance_reduction/boundary_value_caching.h:370-402 splat, :683-810 splatBoundaryData /
// splatSourceData, core/distributions.h:12-55 kernel regularization, :84-177 free-space harmonic
// and :271-321 3D Yukawa Green's functions and Poisson kernels with their gradients).
//
// Every evaluation point x receives one estimate per cached sample y: the boundary term of
// Green's representation formula (Dirichlet, Neumann or Robin form) or the source term, each
// divided by the sample's density.  The work is an all-pairs gather: one thread per evaluation
// point, the sample cache streamed through LDS in tiles of 256 (every thread of the block loads
// one sample, then all threads walk the tile in lockstep, so the per-sample branch on the sample's
// category is wave-uniform and the five per-category accumulators stay in registers).  Sums and
// sums of squares are FP64; each point's batch is Chan-merged into the caller's statistics.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>

#include "../../include/zombie_amd.h"
#include "device.h"

namespace zg {

constexpr int kBvcTile = 256;

struct BvcParams {
    zg_bvc_samples S;
    zg_bvc_eval_points E;
    zg_bvc_statistics out;
    float radiusClamp, kreg, robinCutoff, cutoffA, cutoffR;
    float mu;   // sqrt(absorption): 3D Yukawa kernels when > 0
};

template <int D>
__global__ void __launch_bounds__(256) bvcSplatKernel(BvcParams B) {
    // tile of samples in LDS: position, oriented normal, and the scalars of the estimate
    __shared__ float sPt[D][kBvcTile];
    __shared__ float sN[D][kBvcTile];
    __shared__ float sPdf[kBvcTile], sU[kBvcTile], sDn[kBvcTile], sH[kBvcTile], sK[kBvcTile];
    __shared__ int sCat[kBvcTile];   // 0..3 boundary categories, 4 source, -1 padding

    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool active = i < B.E.n;
    Vec<D> x = vzero<D>();
    float alpha = 1.0f, alphaG = 1.0f;
    if (active) {
        float dA = B.E.dist_to_absorbing_boundary[i];
        float dR = B.E.dist_to_reflecting_boundary ? B.E.dist_to_reflecting_boundary[i] : FLT_MAX;
        if (dA < B.cutoffA || dR < B.cutoffR) active = false;   // near the boundary: not splatted
        uint8_t ty = B.E.type ? B.E.type[i] : (uint8_t)ZG_SAMPLE_IN_DOMAIN;
        if (ty == ZG_SAMPLE_ON_ABSORBING_BOUNDARY || ty == ZG_SAMPLE_ON_REFLECTING_BOUNDARY) { alpha = 2.0f; alphaG = 0.0f; }
        x = vload<D>(B.E.pt + (size_t)i * D);
    }
    int cN[5];
    double cS[5], cQ[5], gS[5][D], gQ[5][D];
#pragma unroll
    for (int c = 0; c < 5; c++) {
        cN[c] = 0; cS[c] = 0.0; cQ[c] = 0.0;
#pragma unroll
        for (int a = 0; a < D; a++) { gS[c][a] = 0.0; gQ[c][a] = 0.0; }
    }

    for (int64_t base = 0; base < B.S.n; base += kBvcTile) {
        __syncthreads();
        {
            const int64_t j = base + threadIdx.x;
            int cat = -1;
            if (j < B.S.n) {
                uint8_t ty = B.S.type[j];
                bool al = B.S.normal_aligned && B.S.normal_aligned[j];
                float sgn = al ? -1.0f : 1.0f;
#pragma unroll
                for (int a = 0; a < D; a++) {
                    sPt[a][threadIdx.x] = B.S.pt[(size_t)j * D + a];
                    sN[a][threadIdx.x] = ty == ZG_SAMPLE_IN_DOMAIN ? 0.0f : sgn * B.S.normal[(size_t)j * D + a];
                }
                sPdf[threadIdx.x] = B.S.pdf[j];
                sU[threadIdx.x] = B.S.solution[j];
                sDn[threadIdx.x] = B.S.normal_derivative[j];
                sH[threadIdx.x] = B.S.contribution ? B.S.contribution[j] : 0.0f;
                sK[threadIdx.x] = B.S.robin_coeff ? B.S.robin_coeff[j] : 0.0f;
                cat = ty == ZG_SAMPLE_IN_DOMAIN ? 4 : (ty == ZG_SAMPLE_ON_ABSORBING_BOUNDARY ? 0 : 2) + (al ? 1 : 0);
            }
            sCat[threadIdx.x] = cat;
        }
        __syncthreads();
        const int cnt = (int)min<int64_t>(kBvcTile, B.S.n - base);
        if (!active) continue;
        for (int t = 0; t < cnt; t++) {
            const int cat = sCat[t];
            Vec<D> y, n, xy;
#pragma unroll
            for (int a = 0; a < D; a++) { y[a] = sPt[a][t]; n[a] = sN[a][t]; xy[a] = x[a] - y[a]; }
            const float r = fmaxf(B.radiusClamp, sqrtf(norm2(xy)));
            const float r2 = r * r;
            // free-space kernels, gradients in the pole x: harmonic (distributions.h:84-177), or in
            // 3D with absorption the Yukawa ones (:271-321) with q1 = e^{-mu r} (mu r + 1):
            // G = e^{-mu r} / (4 pi r), grad G = -(x - y) q1 / (4 pi r^3), P = n.(x - y) q1 / (4 pi r^3),
            // grad P = [n q1 - (x - y) (n.(x - y) / r^2) e^{-mu r} (mu^2 r^2 + 3 mu r + 3)] / (4 pi r^3)
            // (the derivative of P; mu = 0 gives the harmonic kernels)
            float G, P, kG, kP, kPn;
            if (D == 2) {
                G = -logf(r) / (2.0f * kPi);
                kG = -1.0f / (2.0f * kPi * r2);
                P = dot(n, xy) / (2.0f * kPi * r2);
                kP = 2.0f * (dot(n, xy) / r2);
                kPn = 1.0f / (2.0f * kPi * r2);
            } else {
                const float r3 = r2 * r;
                const float mr = B.mu * r;
                const float e = B.mu > 0.0f ? expf(-mr) : 1.0f;
                const float q1 = e * (mr + 1.0f);
                G = e / (4.0f * kPi * r);
                kG = -q1 / (4.0f * kPi * r3);
                P = dot(n, xy) * q1 / (4.0f * kPi * r3);
                kP = (dot(n, xy) / r2) * ((mr * mr + 3.0f * mr + 3.0f) / (mr + 1.0f));
                kPn = q1 / (4.0f * kPi * r3);
            }
            Vec<D> dG, dP;
#pragma unroll
            for (int a = 0; a < D; a++) { dG[a] = xy[a] * kG; dP[a] = (n[a] - xy[a] * kP) * kPn; }
            const float dGn = sqrtf(norm2(dG));
            const float pdf = sPdf[t];
            float est;
            Vec<D> grad;
            if (cat == 4) {
                if (!isfinite(G) || !isfinite(dGn)) continue;
                if (B.kreg > 0.0f) G *= D == 2 ? 1.0f : erff(r / B.kreg);
                const float src = sH[t];
                est = alpha * G * src / pdf;
#pragma unroll
                for (int a = 0; a < D; a++) grad[a] = alphaG * dG[a] * src / pdf;
            } else {
                const float dPn = sqrtf(norm2(dP));
                if (!isfinite(G) || !isfinite(P) || !isfinite(dGn) || !isfinite(dPn)) continue;
                if (B.kreg > 0.0f) {
                    const float rr = r / B.kreg;
                    if (D == 2) {
                        P *= 1.0f - expf(-rr * rr);
                    } else {
                        const float e = erff(rr);
                        G *= e;
                        P *= e - 2.0f * rr * expf(-rr * rr) / sqrtf(kPi);
                    }
                }
                const float u = sU[t], dn = sDn[t], h = sH[t], kappa = sK[t];
                if (fabsf(kappa) > B.robinCutoff) {
                    est = alpha * ((G + P / kappa) * dn - P * h / kappa) / pdf;
#pragma unroll
                    for (int a = 0; a < D; a++) grad[a] = alphaG * ((dG[a] + dP[a] / kappa) * dn - dP[a] * h / kappa) / pdf;
                } else if (fabsf(kappa) > 0.0f) {
                    est = alpha * (G * h - (P + kappa * G) * u) / pdf;
#pragma unroll
                    for (int a = 0; a < D; a++) grad[a] = alphaG * (dG[a] * h - (dP[a] + kappa * dG[a]) * u) / pdf;
                } else {
                    est = alpha * (G * dn - P * u) / pdf;
#pragma unroll
                    for (int a = 0; a < D; a++) grad[a] = alphaG * (dG[a] * dn - dP[a] * u) / pdf;
                }
            }
            // wave-uniform category: static register indices after unrolling
#pragma unroll
            for (int c = 0; c < 5; c++) {
                if (c != cat) continue;
                const double e = (double)est;
                cN[c]++; cS[c] += e; cQ[c] += e * e;
#pragma unroll
                for (int a = 0; a < D; a++) { const double g = (double)grad[a]; gS[c][a] += g; gQ[c][a] += g * g; }
            }
        }
    }
    if (!active) return;
    // Chan merge of this call's batch into the caller's statistics (common.h:268-310)
#pragma unroll
    for (int c = 0; c < 5; c++) {
        if (cN[c] == 0) continue;
        const size_t k = (size_t)i * 5 + c;
        const int64_t N1 = B.out.n[k], nb = cN[c], N = N1 + nb;
        const double f2 = (double)nb / (double)N, f12 = (double)N1 * (double)nb / (double)N;
        const double mB = cS[c] / (double)nb, m2B = fmax(0.0, cQ[c] - cS[c] * mB);
        if (N1 == 0) { B.out.solution_mean[k] = mB; B.out.solution_m2[k] = m2B; }
        else {
            const double d = mB - B.out.solution_mean[k];
            B.out.solution_mean[k] += d * f2;
            B.out.solution_m2[k] += m2B + d * d * f12;
        }
#pragma unroll
        for (int a = 0; a < D; a++) {
            const size_t kg = k * D + a;
            const double gm = gS[c][a] / (double)nb, g2 = fmax(0.0, gQ[c][a] - gS[c][a] * gm);
            if (N1 == 0) { B.out.gradient_mean[kg] = gm; B.out.gradient_m2[kg] = g2; }
            else {
                const double d = gm - B.out.gradient_mean[kg];
                B.out.gradient_mean[kg] += d * f2;
                B.out.gradient_m2[kg] += g2 + d * d * f12;
            }
        }
        B.out.n[k] = N;
    }
}

hipError_t launchBvcSplat(int dim, float absorption, const zg_bvc_samples& S, const zg_bvc_eval_points& E,
                          const zg_bvc_statistics& out, float radiusClamp, float kreg, float robinCutoff, float cutoffA,
                          float cutoffR, hipStream_t s) {
    if (E.n <= 0) return hipSuccess;
    BvcParams B{S, E, out, radiusClamp, kreg, robinCutoff, cutoffA, cutoffR, sqrtf(fmaxf(absorption, 0.0f))};
    dim3 g((unsigned)((E.n + 255) / 256)), bl(256);
    if (dim == 2) hipLaunchKernelGGL(bvcSplatKernel<2>, g, bl, 0, s, B);
    else hipLaunchKernelGGL(bvcSplatKernel<3>, g, bl, 0, s, B);
    return hipGetLastError();
}

}  // namespace zg
