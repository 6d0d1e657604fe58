// zombie_oracle.cpp -- TEST INFRASTRUCTURE ONLY.
//
// CPU restatement of the reference estimator (rohan-sawhney/zombie) used as the parity
// oracle for the GPU engine.  Only tests/, __graft_entry__.smoke() and bench.py's CPU
// baseline leg load this library; the product (zombie_amd/) never links or calls it.
//
// What it restates (reference file:line):
//   pcg32 (deps/pcg32, absent: restated from O'Neill's PCG-XSH-RR 64/32 and the
//     mantissa nextFloat used by core/sampling.h:108-119,171-183)
//   SphereSampler<2,3>::sampleUnitSphereUniform         core/sampling.h:104-111,164-174
//   HarmonicGreensFnBall<2>, <3>                         core/distributions.h:453-665
//   YukawaGreensFnBall<3> (+ rejection sampler)          core/distributions.h:419-440,835-1005
//   closest point on segment / triangle (Ericson)        core/geometry_helpers.h:15-161
//   offsetPointAlongDirection (RTG ch. 6)                core/geometric_queries.h:153-207
//   outsideBoundingDomain, insideDomain                  core/geometric_queries.h:357-379
//   FCPW Bvh_SurfaceArea (absent dep): binary binned-SAH BVH, leaf size 4, 8 buckets,
//     nearest-child-first closest-point traversal, first-hit ray, all-hits parity, as seen
//     through utils/reflectance_boundary_bvh/bvh.h:38-42,369-506 and the call sites
//     utils/fcpw_geometric_queries.h:1069-1242
//   silhouette candidates + closest silhouette point     fcpw_geometric_queries.h:1029-1038,1254-1268
//     (2D vertices; 3D edges with the dihedral angle of reflectance_boundary_bvh setup,
//     fcpw_geometric_queries.h:955-967, and FCPW's edge test restated from the WoSt paper)
//   branchTraversalWeight                                 fcpw_geometric_queries.h:1391-1398
//   SdfGrid value / gradient / projection / sphere tracing  utils/sdf_grid_geometric_queries.h:87-205
//     (walks and gradient walks over it: walk_on_spheres.h:228-593 with the SdfGrid queries)
//   SampleStatistics (FP64 here; float in the reference) point_estimation/common.h:200-395
//   WalkOnSpheres::estimateSolution / walk / weight window / terminal
//                                                        point_estimation/walk_on_spheres.h:170-414
//   WalkOnSpheres::estimateSolutionAndGradient (stratified first step, antithetic pairs,
//     gradient control variates, cosine-weighted directions)
//                                                        point_estimation/walk_on_spheres.h:416-593
//   generateStratifiedSamples, sampleUnitHemisphereCosine, transformCoordinates
//                                                        core/sampling.h:124-157,185-247,435-458
//   Green's function / Poisson kernel gradients          core/distributions.h:493-550,600-659,904-994
//   WalkOnStars::estimateSolution / walk / reflecting + source contributions / throughput
//                                                        point_estimation/walk_on_stars.h:189-662
//   tbb::parallel_for over points -> std::thread pool    walk_on_spheres.h:153-166
//
// Parity pins: see tests/test_oracle_*.py (PCG32 known answers, closed forms of the ball
// Green's functions, brute force against the BVH, manufactured solutions).  The stochastic
// sphere-sampling pdf of the reflecting boundary and FCPW's exact silhouette precision test
// cannot be pinned without the (absent) FCPW source: parity unpinned for those two, checked
// only through unbiasedness identities.

#include <algorithm>
#include <array>
#include <atomic>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <thread>
#include <type_traits>
#include <vector>

#include "../include/zombie_amd.h"

namespace orc {

// ------------------------------------------------------------------------------------
// PCG32 (PCG-XSH-RR 64/32), M. E. O'Neill, "PCG: A Family of Simple Fast Space-Efficient
// Statistically Good Algorithms for Random Number Generation" (2014).
// ------------------------------------------------------------------------------------
struct Pcg32 {
    uint64_t state = 0x853c49e6748fea9bULL;
    uint64_t inc = 0xda3e39cb94b95bdbULL;
    Pcg32() {}
    Pcg32(uint64_t initstate, uint64_t initseq) { seed(initstate, initseq); }
    void seed(uint64_t initstate, uint64_t initseq) {
        state = 0u;
        inc = (initseq << 1u) | 1u;
        nextUInt();
        state += initstate;
        nextUInt();
    }
    uint32_t nextUInt() {
        uint64_t old = state;
        state = old * 0x5851f42d4c957f2dULL + inc;
        uint32_t xorshifted = (uint32_t)(((old >> 18u) ^ old) >> 27u);
        uint32_t rot = (uint32_t)(old >> 59u);
        return (xorshifted >> rot) | (xorshifted << ((~rot + 1u) & 31u));
    }
    // uniform integer in [0, bound): draws below (2^32 mod bound) are rejected so that every
    // residue is equally likely (pcg32.h nextUInt(bound) of the pcg32 dependency, absent;
    // same rule as O'Neill's pcg32_boundedrand_r)
    uint32_t nextUInt(uint32_t bound) {
        uint32_t threshold = (~bound + 1u) % bound;
        for (;;) {
            uint32_t r = nextUInt();
            if (r >= threshold) return r % bound;
        }
    }
    // 23-bit mantissa construction: a float in [1,2) minus 1
    float nextFloat() {
        uint32_t bits = (nextUInt() >> 9) | 0x3f800000u;
        float f;
        std::memcpy(&f, &bits, 4);
        return f - 1.0f;
    }
};

// SplitMix64's output finaliser (Steele, Lea & Flood, OOPSLA 2014; constants of Vigna's
// splitmix64.c), a bijection of 64-bit words.
inline uint64_t splitmix64Mix(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}
// Generator of walk w (counted from the call's stream offset) of point p: PCG32 stream p from
// the state mix(seed, p, w).  Three bijective rounds: for a fixed seed and point, distinct
// walks start at distinct, unrelated positions of the stream's cycle.
inline uint64_t walkInitState(uint64_t seed, uint32_t p, uint32_t w) {
    uint64_t h = splitmix64Mix(seed ^ 0x9e3779b97f4a7c15ULL);
    h = splitmix64Mix(h ^ ((uint64_t)p * 0xd6e8feb86659fd93ULL));
    return splitmix64Mix(h ^ ((uint64_t)w * 0xa0761d6478bd642fULL));
}
inline Pcg32 walkRng(uint64_t seed, uint32_t p, uint32_t w) { return Pcg32(walkInitState(seed, p, w), (uint64_t)p); }

// ------------------------------------------------------------------------------------
// small fixed-size vectors
// ------------------------------------------------------------------------------------
template <int D>
struct V {
    float v[D];
    float& operator[](int i) { return v[i]; }
    float operator[](int i) const { return v[i]; }
    static V zero() { V r; for (int i = 0; i < D; i++) r.v[i] = 0.0f; return r; }
};
template <int D> inline V<D> operator+(V<D> a, const V<D>& b) { for (int i = 0; i < D; i++) a.v[i] += b.v[i]; return a; }
template <int D> inline V<D> operator-(V<D> a, const V<D>& b) { for (int i = 0; i < D; i++) a.v[i] -= b.v[i]; return a; }
template <int D> inline V<D> operator*(V<D> a, float s) { for (int i = 0; i < D; i++) a.v[i] *= s; return a; }
template <int D> inline V<D> operator*(float s, V<D> a) { return a * s; }
template <int D> inline float dot(const V<D>& a, const V<D>& b) { float s = 0.0f; for (int i = 0; i < D; i++) s += a.v[i] * b.v[i]; return s; }
template <int D> inline float norm2(const V<D>& a) { return dot(a, a); }
template <int D> inline float norm(const V<D>& a) { return std::sqrt(norm2(a)); }
inline float cross2(const V<2>& a, const V<2>& b) { return a[0] * b[1] - a[1] * b[0]; }
inline V<3> cross3(const V<3>& a, const V<3>& b) {
    V<3> r; r[0] = a[1] * b[2] - a[2] * b[1]; r[1] = a[2] * b[0] - a[0] * b[2]; r[2] = a[0] * b[1] - a[1] * b[0]; return r;
}

template <int D>
V<D> load(const float* p) { V<D> r; for (int i = 0; i < D; i++) r.v[i] = p[i]; return r; }
template <int D>
void store(float* p, const V<D>& a) { for (int i = 0; i < D; i++) p[i] = a.v[i]; }

static const float PI_F = 3.14159265358979323846f;

// ------------------------------------------------------------------------------------
// SphereSampler (core/sampling.h:104-111, 164-174)
// ------------------------------------------------------------------------------------
inline V<2> sampleUnitSphere2(Pcg32& rng) {
    float u0 = rng.nextFloat();
    float phi = (float)(2.0 * M_PI * u0);
    V<2> d; d[0] = std::cos(phi); d[1] = std::sin(phi);
    return d;
}
inline V<3> sampleUnitSphere3(Pcg32& rng) {
    float u0 = rng.nextFloat();
    float u1 = rng.nextFloat();
    float z = 1.0f - 2.0f * u0;
    float r = std::sqrt(std::max(0.0f, 1.0f - z * z));
    float phi = (float)(2.0 * M_PI * u1);
    V<3> d; d[0] = r * std::cos(phi); d[1] = r * std::sin(phi); d[2] = z;
    return d;
}
template <int D> V<D> sampleUnitSphere(Pcg32& rng);
template <> V<2> sampleUnitSphere<2>(Pcg32& rng) { return sampleUnitSphere2(rng); }
template <> V<3> sampleUnitSphere<3>(Pcg32& rng) { return sampleUnitSphere3(rng); }

// the same maps from given uniforms u[0..DIM-2] (core/sampling.h:104-107,164-170)
template <int D> V<D> sphereFromU(const float* u);
template <> V<2> sphereFromU<2>(const float* u) {
    float phi = (float)(2.0 * M_PI * u[0]);
    V<2> d; d[0] = std::cos(phi); d[1] = std::sin(phi);
    return d;
}
template <> V<3> sphereFromU<3>(const float* u) {
    float z = 1.0f - 2.0f * u[0];
    float r = std::sqrt(std::max(0.0f, 1.0f - z * z));
    float phi = (float)(2.0 * M_PI * u[1]);
    V<3> d; d[0] = r * std::cos(phi); d[1] = r * std::sin(phi); d[2] = z;
    return d;
}

// cosine-weighted hemisphere about +e_{DIM-1} (core/sampling.h:124-129 in 2D; 3D via Shirley's
// concentric disk map :185-217) and its pdf (:146-149, :232-235)
template <int D> V<D> hemisphereCosineFromU(const float* u);
template <> V<2> hemisphereCosineFromU<2>(const float* u) {
    float u1 = 2.0f * u[0] - 1.0f;
    V<2> d; d[0] = u1; d[1] = std::sqrt(std::max(0.0f, 1.0f - u1 * u1));
    return d;
}
template <> V<3> hemisphereCosineFromU<3>(const float* u) {
    float u1 = 2.0f * u[0] - 1.0f, u2 = 2.0f * u[1] - 1.0f;
    float dx = 0.0f, dy = 0.0f;
    if (!(u1 == 0.0f && u2 == 0.0f)) {
        float theta, r;
        if (std::fabs(u1) > std::fabs(u2)) { r = u1; theta = (float)(0.25 * M_PI * (u2 / u1)); }
        else { r = u2; theta = (float)(0.5 * M_PI * (1.0f - 0.5f * (u1 / u2))); }
        dx = r * std::cos(theta); dy = r * std::sin(theta);
    }
    V<3> d; d[0] = dx; d[1] = dy; d[2] = std::sqrt(std::max(0.0f, 1.0f - (dx * dx + dy * dy)));
    return d;
}
template <int D> float pdfHemisphereCosine(float cosTheta) {
    return D == 2 ? cosTheta / 2.0f : (float)(cosTheta / M_PI);
}

// rotate d from the frame whose last axis is n into world coordinates (core/sampling.h:151-157,
// 237-247; the 3D basis is Duff et al., "Building an Orthonormal Basis, Revisited", JCGT 2017)
inline void transformCoordinates(const V<2>& n, V<2>& d) {
    V<2> s; s[0] = n[1]; s[1] = -n[0];
    d = d[0] * s + d[1] * n;
}
inline void transformCoordinates(const V<3>& n, V<3>& d) {
    float sign = std::copysign(1.0f, n[2]);
    float a = -1.0f / (sign + n[2]);
    float b = n[0] * n[1] * a;
    V<3> b1; b1[0] = 1.0f + sign * n[0] * n[0] * a; b1[1] = sign * b; b1[2] = -sign * n[0];
    V<3> b2; b2[0] = b; b2[1] = sign + n[1] * n[1] * a; b2[2] = -n[1];
    d = d[0] * b1 + d[1] * b2 + d[2] * n;
}

// Latin hypercube samples (core/sampling.h:435-458): nSamples points of `dims` coordinates,
// row i starts on the diagonal stratum (i + U)/nSamples, then every coordinate column gets its
// own Fisher-Yates shuffle driven by bounded draws
inline void stratifiedSamples(int dims, int nSamples, Pcg32& rng, std::vector<float>& s) {
    const float oneMinusEps = 1.0f - FLT_EPSILON;
    float inv = 1.0f / (float)nSamples;
    s.assign((size_t)dims * nSamples, 0.0f);
    for (int i = 0; i < nSamples; i++)
        for (int j = 0; j < dims; j++) s[(size_t)dims * i + j] = std::min(((float)i + rng.nextFloat()) * inv, oneMinusEps);
    for (int i = 0; i < dims; i++)
        for (int j = 0; j < nSamples; j++) {
            int other = j + (int)rng.nextUInt((uint32_t)(nSamples - j));
            std::swap(s[(size_t)dims * j + i], s[(size_t)dims * other + i]);
        }
}

// ------------------------------------------------------------------------------------
// Green's functions of the ball (core/distributions.h:323-665, 835-1005)
// ------------------------------------------------------------------------------------
// sinh x and I(x) = cosh x - sinh x / x from e = e^{-x}, with Taylor series below the points
// where the exponential forms lose their digits to cancellation in float (the reference's
// updateBall, distributions.h:843-852, uses the exponential forms throughout: at x = 0.004
// cosh x - sinh x / x keeps no correct digit, which turns its gradient terms into noise)
inline float yukawaSinh(float x, float e) {
    if (x < 0.1f) { float x2 = x * x; return x * (1.0f + x2 / 6.0f * (1.0f + x2 / 20.0f * (1.0f + x2 / 42.0f))); }
    return (1.0f - e * e) / (2.0f * e);
}
inline float yukawaI(float x, float e) {
    if (x < 0.5f) { float x2 = x * x; return x2 * (1.0f / 3.0f + x2 * (1.0f / 30.0f + x2 * (1.0f / 840.0f + x2 / 45360.0f))); }
    return (1.0f + e * e) / (2.0f * e) - yukawaSinh(x, e) / x;
}

template <int D>
struct GreensBall {
    bool yukawa = false;
    float sqrtLambda = 0.0f;
    V<D> c;
    float R = 0.0f;
    float rClamp = 1e-4f;
    // Yukawa cached terms (distributions.h:843-852)
    float muR = 0, expmuR = 0, sinhmuR = 0, K32muR = 0, I32muR = 0;

    void makeHarmonic() { yukawa = false; }
    void makeYukawa(float lambda) { yukawa = true; sqrtLambda = std::sqrt(lambda); }

    void updateBall(const V<D>& c_, float R_) {
        c = c_; R = R_; rClamp = 1e-4f;
        if (yukawa) {
            muR = R * sqrtLambda;
            expmuR = std::exp(-muR);
            sinhmuR = yukawaSinh(muR, expmuR);
            K32muR = expmuR * (1.0f + 1.0f / muR);
            I32muR = yukawaI(muR, expmuR);
        }
    }

    float evaluate(float r) const {
        if (!yukawa) {
            if (D == 2) return std::log(R / r) / (2.0f * PI_F);
            return (1.0f / r - 1.0f / R) / (4.0f * PI_F);
        }
        float mur = r * sqrtLambda;
        float expmur = std::exp(-mur);
        float sinhmur = yukawaSinh(mur, expmur);
        return (expmur - expmuR * sinhmur / sinhmuR) / (4.0f * PI_F * r);
    }

    float poissonKernel() const {  // centred, |y-c| = R
        if (!yukawa) return D == 2 ? 1.0f / (2.0f * PI_F) : 1.0f / (4.0f * PI_F);
        return muR / (4.0f * PI_F * sinhmuR);
    }

    float norm() const {
        if (!yukawa) return D == 2 ? R * R / 4.0f : R * R / 6.0f;
        if (muR < 0.5f) {   // 1 - x / sinh x = x^2/6 - 7 x^4/360 + 31 x^6/15120 - 127 x^8/604800
            float x2 = muR * muR;
            return R * R * (1.0f / 6.0f - x2 * (7.0f / 360.0f - x2 * (31.0f / 15120.0f - x2 * (127.0f / 604800.0f))));
        }
        float lambda = sqrtLambda * sqrtLambda;
        return (1.0f - 4.0f * PI_F * poissonKernel()) / lambda;
    }

    float poissonKernelDampeningFactor(float r) const {
        float mur = r * sqrtLambda;
        float expmur = std::exp(-mur);
        float K32mur = expmur * (1.0f + 1.0f / mur);
        float I32mur = yukawaI(mur, expmur);
        float Q = K32mur + I32mur * expmuR / sinhmuR;
        return mur * Q;
    }

    float directionSampledPoissonKernel(const V<D>& y) const {
        if (!yukawa) return 1.0f;
        float r = std::max(rClamp, orc::norm(y - c));
        return poissonKernelDampeningFactor(r);
    }

    // reflectance (distributions.h:536-543 harmonic 2D, :635-645 harmonic 3D, :957-971 Yukawa 3D)
    // for a step of length r along dir onto a boundary with normal n and Robin coefficient kappa
    float reflectance(float r, const V<D>& dir, const V<D>& n, float kappa) const {
        if (!yukawa) {
            if (std::fabs(kappa) > 0.0f) {
                float P = D == 2 ? dot(n, dir) / r : dot(n, dir) / (r * r);
                float G = D == 2 ? std::log(R / r) : 1.0f / r - 1.0f / R;
                return 1.0f - kappa * G / P;
            }
            return 1.0f;
        }
        float Q = poissonKernelDampeningFactor(r);
        if (std::fabs(kappa) > 0.0f) {
            float P = dot(n, dir) / (r * r);
            float mur = r * sqrtLambda;
            float expmur = std::exp(-mur);
            float sinhmur = yukawaSinh(mur, expmur);
            float G = (expmur - expmuR * sinhmur / sinhmuR) / r;
            return Q - kappa * G / P;
        }
        return Q;
    }

    // gradient of G(x, y) in y for the centred ball: (y - c) * gradientNorm(r)
    // (distributions.h:493-502 harmonic 2D, :600-609 harmonic 3D, :904-922 Yukawa 3D)
    float gradientNorm(float r) const {
        if (!yukawa) {
            if (D == 2) return (1.0f / (r * r) - 1.0f / (R * R)) / (2.0f * PI_F);
            return (1.0f / (r * r * r) - 1.0f / (R * R * R)) / (4.0f * PI_F);
        }
        float mur = r * sqrtLambda;
        float expmur = std::exp(-mur);
        float K32mur = expmur * (1.0f + 1.0f / mur);
        float I32mur = yukawaI(mur, expmur);
        float Qr = sqrtLambda * (K32mur - I32mur * K32muR / I32muR);
        return Qr / (4.0f * PI_F * r * r);
    }
    V<D> gradient(float r, const V<D>& y) const { return (y - c) * gradientNorm(r); }

    // gradient of the Poisson kernel at y on the sphere (distributions.h:545-550, :654-659, :987-994)
    V<D> poissonKernelGradient(const V<D>& y) const {
        V<D> d = y - c;
        if (!yukawa) {
            if (D == 2) return 2.0f * d * (1.0f / (2.0f * PI_F * R * R));
            return 3.0f * d * (1.0f / (4.0f * PI_F * R * R));
        }
        float lambda = sqrtLambda * sqrtLambda;
        float QR = lambda / I32muR;
        return d * QR * (1.0f / (4.0f * PI_F));
    }

    float pdfSampleSphereUniform(float r) const {
        return D == 2 ? 1.0f / (2.0f * PI_F * r) : 1.0f / (4.0f * PI_F * r * r);
    }

    // sampleVolume(dir, rng, r, pdf)
    V<D> sampleVolumeDir(const V<D>& dir, Pcg32& rng, float& r, float& pdf) {
        if (!yukawa) {
            if (D == 2) {
                float u1 = rng.nextFloat();
                float u2 = rng.nextFloat();
                r = R * std::sqrt(u1 * u2);
            } else {
                float u1 = rng.nextFloat();
                float u2 = rng.nextFloat();
                float u3 = rng.nextFloat();
                float uMedian = std::max(std::min(u1, u2), std::min(std::max(u1, u2), u3));
                r = R * uMedian;
            }
            r = std::max(rClamp, r);
            if (r > R) r = R / 2.0f;
            pdf = evaluate(r) / norm();
            return c + r * dir;
        }
        // rejection sampling (distributions.h:855-864 -> 419-440)
        float lambda = sqrtLambda * sqrtLambda;
        float bound = R <= lambda ?
            std::max(std::max(2.0f / R, 2.0f / lambda), std::max(0.5f * std::sqrt(R), 0.5f * sqrtLambda)) :
            std::max(std::min(2.0f / R, 2.0f / lambda), std::min(0.5f * std::sqrt(R), 0.5f * sqrtLambda));
        int iter = 0;
        float fnNorm = norm();
        do {
            float u = rng.nextFloat();
            r = rng.nextFloat() * R;
            pdf = evaluate(r) / fnNorm;
            float pdfRadius = pdf / pdfSampleSphereUniform(r);
            iter++;
            if (u < pdfRadius / bound) break;
        } while (iter < 1000);
        r = std::max(rClamp, r);
        if (r > R) r = R / 2.0f;
        return c + r * dir;
    }

    V<D> sampleVolume(Pcg32& rng, float& r, float& pdf) {
        V<D> dir = sampleUnitSphere<D>(rng);
        return sampleVolumeDir(dir, rng, r, pdf);
    }
};

// ------------------------------------------------------------------------------------
// manufactured PDE data (zg_function descriptors), evaluated in float like the
// reference's T=float callbacks
// ------------------------------------------------------------------------------------
// DenseGrid<float,1,DIM>::operator() (utils/dense_grid.h:264-356): cell-centred values over
// [grid_min, grid_max], last axis fastest; nearest cell, or bi/trilinear interpolation between
// cell centres with the blend fractions clamped to [0, 1] (clamp to edge)
template <int D>
float gridValue(const zg_function& f, const V<D>& x) {
    float l[3]; int sh[3];
    for (int a = 0; a < D; a++) { l[a] = (x[a] - f.grid_min[a]) / (f.grid_max[a] - f.grid_min[a]); sh[a] = f.shape[a]; }
    if (!f.interpolate) {
        size_t flat = 0, stride = 1;
        for (int a = D - 1; a >= 0; a--) {
            int idx = std::clamp((int)std::floor(l[a] * sh[a]), 0, sh[a] - 1);
            flat += (size_t)idx * stride; stride *= (size_t)sh[a];
        }
        return f.grid[flat];
    }
    int i0[3], i1[3]; float t[3];
    for (int a = 0; a < D; a++) {
        float u = sh[a] == 1 ? 0.0f : l[a] * sh[a] - 0.5f;
        i0[a] = std::clamp((int)std::floor(u), 0, sh[a] - 1);
        i1[a] = std::clamp(i0[a] + 1, 0, sh[a] - 1);
        t[a] = std::clamp(u - (float)i0[a], 0.0f, 1.0f);
    }
    const float* g = f.grid;
    if constexpr (D == 2) {
        return g[i0[0] * sh[1] + i0[1]] * (1 - t[0]) * (1 - t[1]) + g[i1[0] * sh[1] + i0[1]] * t[0] * (1 - t[1]) +
               g[i0[0] * sh[1] + i1[1]] * (1 - t[0]) * t[1] + g[i1[0] * sh[1] + i1[1]] * t[0] * t[1];
    } else {
        float r = 0.0f;
        for (int c = 0; c < 8; c++) {
            int ci = (c >> 2) & 1, cj = (c >> 1) & 1, ck = c & 1;
            float w = (ci ? t[0] : 1 - t[0]) * (cj ? t[1] : 1 - t[1]) * (ck ? t[2] : 1 - t[2]);
            r += w * g[((ci ? i1[0] : i0[0]) * sh[1] + (cj ? i1[1] : i0[1])) * sh[2] + (ck ? i1[2] : i0[2])];
        }
        return r;
    }
}

template <int D>
float evalField(const zg_function& f, const V<D>& x, const V<D>& n, bool normalAligned, float absorption) {
    if (f.kind == ZG_FN_DENSE_GRID) return gridValue<D>(f, x);
    const float* p = f.params;
    float z = D == 3 ? x[2] : 0.0f;
    float u = 0.0f;
    V<3> g; g[0] = g[1] = g[2] = 0.0f;  // gradient
    float negLap = 0.0f;
    switch (f.kind) {
    case ZG_FN_ZERO: break;
    case ZG_FN_CONSTANT: u = p[0]; break;
    case ZG_FN_EXP_COS: {
        float e = std::exp(p[1] * x[0]);
        float cs = std::cos(p[1] * x[1]);
        float sn = std::sin(p[1] * x[1]);
        u = p[0] * e * cs;
        g[0] = p[0] * p[1] * e * cs;
        g[1] = -p[0] * p[1] * e * sn;
        negLap = 0.0f;
        break;
    }
    case ZG_FN_SIN_PRODUCT: {
        float s[3] = {1.0f, 1.0f, 1.0f}, c[3] = {0.0f, 0.0f, 0.0f};
        for (int i = 0; i < D; i++) { s[i] = std::sin(p[1] * x[i]); c[i] = std::cos(p[1] * x[i]); }
        u = p[0] * s[0] * s[1] * s[2];
        g[0] = p[0] * p[1] * c[0] * s[1] * s[2];
        g[1] = p[0] * p[1] * s[0] * c[1] * s[2];
        if (D == 3) g[2] = p[0] * p[1] * s[0] * s[1] * c[2];
        negLap = (float)D * p[1] * p[1] * u;
        break;
    }
    case ZG_FN_CUBIC_HARMONIC: {
        float xx = x[0], yy = x[1];
        u = p[0] * (xx * xx * xx - 3.0f * xx * yy * yy);
        g[0] = p[0] * (3.0f * xx * xx - 3.0f * yy * yy);
        g[1] = -p[0] * 6.0f * xx * yy;
        negLap = 0.0f;
        break;
    }
    case ZG_FN_EXP_LINEAR: {
        float a = p[1] * x[0] + p[2] * x[1] + (D == 3 ? p[3] * z : 0.0f);
        u = p[0] * std::exp(a);
        g[0] = p[1] * u; g[1] = p[2] * u; g[2] = D == 3 ? p[3] * u : 0.0f;
        float a2 = p[1] * p[1] + p[2] * p[2] + (D == 3 ? p[3] * p[3] : 0.0f);
        negLap = -a2 * u;
        break;
    }
    default: break;
    }
    if (f.eval == ZG_EVAL_VALUE) return u;
    if (f.eval == ZG_EVAL_NORMAL_DERIVATIVE || f.eval == ZG_EVAL_ROBIN) {
        float dn = 0.0f;
        for (int i = 0; i < D; i++) dn += g[i] * n[i];
        dn = normalAligned ? -dn : dn;
        return f.eval == ZG_EVAL_ROBIN ? dn + f.params[5] * u : dn;
    }
    return negLap + absorption * u;
}

// ------------------------------------------------------------------------------------
// primitives (core/geometry_helpers.h)
// ------------------------------------------------------------------------------------
inline float closestPointSegment(const V<2>& pa, const V<2>& pb, const V<2>& x, V<2>& pt) {
    V<2> u = pb - pa, v = x - pa;
    float c1 = dot(u, v);
    if (c1 <= 0.0f) { pt = pa; return norm(x - pa); }
    float c2 = dot(u, u);
    if (c2 <= c1) { pt = pb; return norm(x - pb); }
    float t = c1 / c2;
    pt = pa + u * t;
    return norm(x - pt);
}

template <int D>
inline float closestPointSegmentD(const V<D>& pa, const V<D>& pb, const V<D>& x, V<D>& pt) {
    V<D> u = pb - pa, v = x - pa;
    float c1 = dot(u, v);
    if (c1 <= 0.0f) { pt = pa; return norm(x - pa); }
    float c2 = dot(u, u);
    if (c2 <= c1) { pt = pb; return norm(x - pb); }
    pt = pa + u * (c1 / c2);
    return norm(x - pt);
}

inline float closestPointTriangle(const V<3>& pa, const V<3>& pb, const V<3>& pc, const V<3>& x, V<3>& pt) {
    V<3> ab = pb - pa, ac = pc - pa, ax = x - pa;
    float d1 = dot(ab, ax), d2 = dot(ac, ax);
    if (d1 <= 0.0f && d2 <= 0.0f) { pt = pa; return norm(x - pa); }
    V<3> bx = x - pb;
    float d3 = dot(ab, bx), d4 = dot(ac, bx);
    if (d3 >= 0.0f && d4 <= d3) { pt = pb; return norm(x - pb); }
    V<3> cx = x - pc;
    float d5 = dot(ab, cx), d6 = dot(ac, cx);
    if (d6 >= 0.0f && d5 <= d6) { pt = pc; return norm(x - pc); }
    float vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f) { float v = d1 / (d1 - d3); pt = pa + ab * v; return norm(x - pt); }
    float vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f) { float w = d2 / (d2 - d6); pt = pa + ac * w; return norm(x - pt); }
    float va = d3 * d6 - d5 * d4;
    if (va <= 0.0f && (d4 - d3) >= 0.0f && (d5 - d6) >= 0.0f) {
        float w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        pt = pb + (pc - pb) * w;
        return norm(x - pt);
    }
    float denom = 1.0f / (va + vb + vc);
    float v = vb * denom, w = vc * denom;
    pt = pa + ab * v + ac * w;
    return norm(x - pt);
}

inline V<2> segmentNormal(const V<2>& pa, const V<2>& pb) {
    V<2> s = pb - pa; V<2> n; n[0] = s[1]; n[1] = -s[0];
    float l = norm(n); return n * (1.0f / l);
}
inline V<3> triangleNormal(const V<3>& pa, const V<3>& pb, const V<3>& pc) {
    V<3> n = cross3(pb - pa, pc - pa); float l = norm(n); return n * (1.0f / l);
}

// ray vs segment: origin o, direction d, t in [0, tMax]
inline bool raySegment(const V<2>& o, const V<2>& d, float tMax, const V<2>& pa, const V<2>& pb, float& t, float& s) {
    V<2> u = pa - o, v = pb - pa;
    float dv = cross2(d, v);
    if (std::fabs(dv) <= 0.0f) return false;
    float ud = cross2(u, d);
    s = ud / dv;
    if (s < 0.0f || s > 1.0f) return false;
    float uv = cross2(u, v);
    t = uv / dv;
    return t >= 0.0f && t <= tMax;
}

// Moller-Trumbore
inline bool rayTriangle(const V<3>& o, const V<3>& d, float tMax, const V<3>& pa, const V<3>& pb, const V<3>& pc, float& t) {
    V<3> e1 = pb - pa, e2 = pc - pa;
    V<3> p = cross3(d, e2);
    float det = dot(e1, p);
    if (std::fabs(det) <= 0.0f) return false;
    float inv = 1.0f / det;
    V<3> s = o - pa;
    float u = dot(s, p) * inv;
    if (u < 0.0f || u > 1.0f) return false;
    V<3> q = cross3(s, e1);
    float v = dot(d, q) * inv;
    if (v < 0.0f || u + v > 1.0f) return false;
    t = dot(e2, q) * inv;
    return t >= 0.0f && t <= tMax;
}

// offsetPointAlongDirection (core/geometric_queries.h:175-207)
inline int floatAsInt(float a) { int b; std::memcpy(&b, &a, 4); return b; }
inline float intAsFloat(int a) { float b; std::memcpy(&b, &a, 4); return b; }
template <int D>
V<D> offsetPoint(const V<D>& p, const V<D>& n) {
    const float origin = 1.0f / 32.0f, floatScale = 1.0f / 65536.0f, intScale = 256.0f;
    V<D> r;
    for (int i = 0; i < D; i++) {
        int nOff = (int)(n[i] * intScale);
        float pOff = intAsFloat(floatAsInt(p[i]) + (p[i] < 0 ? -nOff : nOff));
        r[i] = std::fabs(p[i]) < origin ? p[i] + floatScale * n[i] : pOff;
    }
    return r;
}

// ------------------------------------------------------------------------------------
// binary binned-SAH BVH (FCPW Bvh_SurfaceArea defaults: leaf size 4, 8 buckets)
// ------------------------------------------------------------------------------------
template <int D>
struct Box {
    V<D> lo, hi;
    Box() { for (int i = 0; i < D; i++) { lo[i] = FLT_MAX; hi[i] = -FLT_MAX; } }
    void expand(const V<D>& p) { for (int i = 0; i < D; i++) { lo[i] = std::min(lo[i], p[i]); hi[i] = std::max(hi[i], p[i]); } }
    void expand(const Box& b) { for (int i = 0; i < D; i++) { lo[i] = std::min(lo[i], b.lo[i]); hi[i] = std::max(hi[i], b.hi[i]); } }
    float area() const {
        V<D> e = hi - lo;
        for (int i = 0; i < D; i++) if (e[i] < 0) return 0.0f;
        if (D == 2) return 2.0f * (e[0] + e[1]);
        return 2.0f * (e[0] * e[1] + e[1] * e[2] + e[0] * e[2]);
    }
    float dist2(const V<D>& x) const {
        float s = 0.0f;
        for (int i = 0; i < D; i++) {
            float d = std::max(std::max(lo[i] - x[i], x[i] - hi[i]), 0.0f);
            s += d * d;
        }
        return s;
    }
    V<D> centroid() const { return (lo + hi) * 0.5f; }
    // slab test
    bool rayOverlap(const V<D>& o, const V<D>& invD, float tMax) const {
        float t0 = 0.0f, t1 = tMax;
        for (int i = 0; i < D; i++) {
            float a = (lo[i] - o[i]) * invD[i], b = (hi[i] - o[i]) * invD[i];
            if (a > b) std::swap(a, b);
            t0 = std::max(t0, a); t1 = std::min(t1, b);
            if (std::isnan(a) || std::isnan(b)) { t0 = 0.0f; }
        }
        return t0 <= t1 * 1.00001f + 1e-7f;
    }
};

template <int D>
struct Bvh {
    struct Node { Box<D> box; int left = -1, right = -1, start = 0, count = 0; };
    std::vector<Node> nodes;
    std::vector<int> order;  // primitive indices in leaf order
    int depth = 0;

    void build(const std::vector<Box<D>>& boxes, int leafSize, int nBuckets) {
        int n = (int)boxes.size();
        order.resize(n);
        for (int i = 0; i < n; i++) order[i] = i;
        nodes.clear();
        if (n == 0) return;
        std::vector<V<D>> cent(n);
        for (int i = 0; i < n; i++) cent[i] = boxes[i].centroid();
        nodes.reserve(2 * n);
        depth = 0;
        buildRec(boxes, cent, 0, n, leafSize, nBuckets, 1);
    }

    int buildRec(const std::vector<Box<D>>& boxes, const std::vector<V<D>>& cent,
                 int start, int end, int leafSize, int nBuckets, int level) {
        depth = std::max(depth, level);
        int id = (int)nodes.size();
        nodes.push_back(Node());
        Box<D> bb, cb;
        for (int i = start; i < end; i++) { bb.expand(boxes[order[i]]); cb.expand(cent[order[i]]); }
        nodes[id].box = bb;
        int n = end - start;
        if (n <= leafSize) { nodes[id].start = start; nodes[id].count = n; return id; }
        // binned SAH over all axes
        float bestCost = FLT_MAX; int bestAxis = -1, bestSplit = -1;
        for (int a = 0; a < D; a++) {
            float ext = cb.hi[a] - cb.lo[a];
            if (ext <= 0.0f) continue;
            std::vector<Box<D>> bbox(nBuckets);
            std::vector<int> cnt(nBuckets, 0);
            for (int i = start; i < end; i++) {
                int b = (int)(nBuckets * (cent[order[i]][a] - cb.lo[a]) / ext);
                b = std::min(std::max(b, 0), nBuckets - 1);
                cnt[b]++; bbox[b].expand(boxes[order[i]]);
            }
            for (int s = 1; s < nBuckets; s++) {
                Box<D> l, r; int nl = 0, nr = 0;
                for (int b = 0; b < s; b++) { if (cnt[b]) { l.expand(bbox[b]); nl += cnt[b]; } }
                for (int b = s; b < nBuckets; b++) { if (cnt[b]) { r.expand(bbox[b]); nr += cnt[b]; } }
                if (nl == 0 || nr == 0) continue;
                float cost = l.area() * nl + r.area() * nr;
                if (cost < bestCost) { bestCost = cost; bestAxis = a; bestSplit = s; }
            }
        }
        int mid;
        if (bestAxis < 0) {
            mid = start + n / 2;
        } else {
            float ext = cb.hi[bestAxis] - cb.lo[bestAxis];
            auto it = std::partition(order.begin() + start, order.begin() + end, [&](int p) {
                int b = (int)(nBuckets * (cent[p][bestAxis] - cb.lo[bestAxis]) / ext);
                b = std::min(std::max(b, 0), nBuckets - 1);
                return b < bestSplit;
            });
            mid = (int)(it - order.begin());
            if (mid == start || mid == end) mid = start + n / 2;
        }
        int l = buildRec(boxes, cent, start, mid, leafSize, nBuckets, level + 1);
        int r = buildRec(boxes, cent, mid, end, leafSize, nBuckets, level + 1);
        nodes[id].left = l; nodes[id].right = r;
        return id;
    }
};

// ------------------------------------------------------------------------------------
// boundary meshes (absorbing / reflecting) with queries
// ------------------------------------------------------------------------------------
template <int D>
struct Mesh {
    std::vector<V<D>> verts;
    std::vector<int> idx;          // D indices per primitive
    std::vector<int> origId;       // primitive id in the caller's mesh
    std::vector<Box<D>> boxes;
    Bvh<D> bvh;
    // 2D silhouette candidates: vertex -> adjacent primitive(s)
    struct SilVertex { int v; int p0; int p1; };  // p1 = -1 when the vertex has one neighbour
    std::vector<SilVertex> sil;
    // 3D silhouette candidates: edge (va, vb) with its adjacent triangles (f1 = -1: boundary edge)
    struct SilEdge { int va; int vb; int f0; int f1; };
    std::vector<SilEdge> silE;
    Bvh<D> silBvh;                                 // over the silhouette candidates (point boxes)
    std::vector<std::array<float, 4>> silCone;     // per silBvh node: axis.xy, sin, cos of half-angle + margin
    float silConePrecision = 0.0f;
    // Robin coefficient bounds per primitive (FcpwRobinBoundaryHandler::buildAccelerationStructure /
    // updateCoefficientValues, fcpw_geometric_queries.h:726-831); empty = 0 (pure Neumann)
    std::vector<float> robinMin, robinMax;
    bool hasRobin() const { for (float v : robinMax) if (v > FLT_EPSILON) return true; return false; }

    int nPrims() const { return (int)origId.size(); }
    V<D> vtx(int p, int k) const { return verts[idx[p * D + k]]; }

    // uniform point on primitive p from D - 1 uniform numbers (the triangle by the square-root
    // warp: (1 - sqrt(u1)) a + sqrt(u1) (1 - u2) b + sqrt(u1) u2 c)
    V<D> primSample(int p, const float* u) const {
        if constexpr (D == 2) return vtx(p, 0) + (vtx(p, 1) - vtx(p, 0)) * u[0];
        else {
            float su = std::sqrt(u[0]);
            return vtx(p, 0) + (vtx(p, 1) - vtx(p, 0)) * (su * (1.0f - u[1])) + (vtx(p, 2) - vtx(p, 0)) * (su * u[1]);
        }
    }
    V<D> primNormal(int p) const {
        if constexpr (D == 2) return segmentNormal(vtx(p, 0), vtx(p, 1));
        else return triangleNormal(vtx(p, 0), vtx(p, 1), vtx(p, 2));
    }
    float primArea(int p) const {
        if constexpr (D == 2) return norm(vtx(p, 1) - vtx(p, 0));
        else return 0.5f * norm(cross3(vtx(p, 1) - vtx(p, 0), vtx(p, 2) - vtx(p, 0)));
    }
    float primClosest(int p, const V<D>& x, V<D>& pt) const {
        if constexpr (D == 2) return closestPointSegment(vtx(p, 0), vtx(p, 1), x, pt);
        else return closestPointTriangle(vtx(p, 0), vtx(p, 1), vtx(p, 2), x, pt);
    }
    bool primRay(int p, const V<D>& o, const V<D>& d, float tMax, float& t) const {
        if constexpr (D == 2) { float s; return raySegment(o, d, tMax, vtx(p, 0), vtx(p, 1), t, s); }
        else return rayTriangle(o, d, tMax, vtx(p, 0), vtx(p, 1), vtx(p, 2), t);
    }

    void build(int leafSize, int nBuckets) {
        int n = nPrims();
        boxes.resize(n);
        for (int p = 0; p < n; p++) { Box<D> b; for (int k = 0; k < D; k++) b.expand(vtx(p, k)); boxes[p] = b; }
        bvh.build(boxes, leafSize, nBuckets);
    }

    // silhouette candidates (2D): each vertex with its adjacent segments; with one-sided
    // problems, convex vertices are ignored (getIgnoreCandidateSilhouetteCallback,
    // fcpw_geometric_queries.h:1029-1038)
    void buildSilhouettes(bool solveDoubleSided, float precision) {
        sil.clear();
        if constexpr (D == 2) {
            int nv = (int)verts.size();
            std::vector<int> first(nv, -1), second(nv, -1), asEnd(nv, -1), asStart(nv, -1);
            for (int p = 0; p < nPrims(); p++) {
                for (int k = 0; k < 2; k++) {
                    int v = idx[p * 2 + k];
                    if (first[v] < 0) first[v] = p; else if (second[v] < 0) second[v] = p;
                    if (k == 0) asStart[v] = p; else asEnd[v] = p;
                }
            }
            for (int v = 0; v < nv; v++) {
                if (first[v] < 0) continue;
                if (second[v] < 0) { sil.push_back({v, first[v], -1}); continue; }
                // dihedral angle at v, positive at reflex vertices of a boundary with outward
                // normals; convex (and flat) candidates are ignored for one-sided problems
                if (!solveDoubleSided && asEnd[v] >= 0 && asStart[v] >= 0) {
                    V<2> n0 = primNormal(asEnd[v]), n1 = primNormal(asStart[v]);
                    float ang = std::atan2(cross2(n1, n0), dot(n0, n1));
                    if (ang < precision) continue;
                }
                sil.push_back({v, first[v], second[v]});
            }
            std::vector<Box<2>> pb(sil.size());
            for (size_t k = 0; k < sil.size(); k++) pb[k].expand(verts[sil[k].v]);
            silBvh.build(pb, 4, 8);
            buildSilhouetteCones(precision);
            silConePrecision = precision;
        } else {
            // every edge with its first two adjacent triangles (a third one of a non-manifold
            // edge is not recorded); f0's winding gives the edge direction a -> b
            std::map<std::pair<int, int>, std::array<int, 3>> edges;   // (min, max) -> f0, f1, j in f0
            for (int p = 0; p < nPrims(); p++) {
                for (int j = 0; j < 3; j++) {
                    int a = idx[p * 3 + j], b = idx[p * 3 + (j + 1) % 3];
                    auto key = std::make_pair(std::min(a, b), std::max(a, b));
                    auto it = edges.find(key);
                    if (it == edges.end()) edges[key] = {p, -1, j};
                    else if (it->second[1] < 0) it->second[1] = p;
                }
            }
            silE.clear();
            for (const auto& kv : edges) {
                int f0 = kv.second[0], f1 = kv.second[1], j = kv.second[2];
                int a = idx[f0 * 3 + j], b = idx[f0 * 3 + (j + 1) % 3];
                if (f1 >= 0 && !solveDoubleSided) {
                    // dihedral angle atan2(e . (n1 x n0), n0 . n1), negative at convex edges of a
                    // surface with outward normals: ignored below the precision for one-sided problems
                    V<3> e = verts[b] - verts[a];
                    e = e * (1.0f / norm(e));
                    V<3> n0 = primNormal(f0), n1 = primNormal(f1);
                    float ang = std::atan2(dot(e, cross3(n1, n0)), dot(n0, n1));
                    if (ang < precision) continue;
                }
                silE.push_back({a, b, f0, f1});
            }
            std::vector<Box<3>> pb(silE.size());
            for (size_t k = 0; k < silE.size(); k++) { pb[k].expand(verts[silE[k].va]); pb[k].expand(verts[silE[k].vb]); }
            silBvh.build(pb, 4, 8);
        }
    }

    // Silhouette vertex seen from x.  With d0, d1 the cosines between the adjacent segments'
    // normals and the unit direction v -> x: a vertex whose neighbours lie on the same side of
    // the line (x, v) is a silhouette, i.e. d0 * d1 < 0.  When x lies (within the precision)
    // on the line of one adjacent segment, the vertex is a silhouette only when the other
    // segment turns towards the far side, sign * d_other > precision: a concave corner seen
    // from the walker's side, whose grazing ray bounds a region the walker cannot see.  sign
    // = +1 for normals pointing away from the walker (outward normals of a domain walked from
    // inside), -1 when the walker's normal orientation is flipped.  A vertex with one
    // neighbour (an open end) and a query at the vertex itself are always silhouettes.
    // normal cone of every silhouette-BVH node (the spatialized normal cone hierarchy of fcpw's
    // findClosestSilhouettePoint): the smallest arc containing the normals of the segments
    // adjacent to the node's candidates, widened by asin(precision) + 1e-4; full for nodes
    // holding an open end.  Stored as (axis.x, axis.y, sin, cos) of the widened half-angle;
    // (0, 0, 1, -1) never prunes.
    void buildSilhouetteCones(float precision) {
        const auto& nodes = silBvh.nodes;
        int nn = (int)nodes.size();
        silCone.assign(nn, {0.0f, 0.0f, 1.0f, -1.0f});
        std::vector<int> lo(nn), hi(nn);
        for (int i = nn - 1; i >= 0; i--) {   // children follow their parent in build order
            if (nodes[i].left < 0) { lo[i] = nodes[i].start; hi[i] = nodes[i].start + nodes[i].count; }
            else { lo[i] = std::min(lo[nodes[i].left], lo[nodes[i].right]); hi[i] = std::max(hi[nodes[i].left], hi[nodes[i].right]); }
        }
        double margin = std::asin(std::min(1.0, (double)std::max(precision, 0.0f))) + 1e-4;
        std::vector<double> ang;
        for (int i = 0; i < nn; i++) {
            ang.clear();
            bool full = false;
            for (int j = lo[i]; j < hi[i]; j++) {
                const SilVertex& sv = sil[silBvh.order[j]];
                if (sv.p1 < 0) { full = true; break; }
                V<2> n0 = primNormal(sv.p0), n1 = primNormal(sv.p1);
                ang.push_back(std::atan2((double)n0[1], (double)n0[0]));
                ang.push_back(std::atan2((double)n1[1], (double)n1[0]));
            }
            if (full || ang.empty()) continue;
            std::sort(ang.begin(), ang.end());
            double gap = ang.front() + 2.0 * M_PI - ang.back(), start = ang.front();
            for (size_t j = 1; j < ang.size(); j++)
                if (ang[j] - ang[j - 1] > gap) { gap = ang[j] - ang[j - 1]; start = ang[j]; }
            double arc = 2.0 * M_PI - gap, half = 0.5 * arc + margin;
            if (half >= 0.5 * M_PI) continue;
            double axis = start + 0.5 * arc;
            silCone[i] = {(float)std::cos(axis), (float)std::sin(axis), std::nextafter((float)std::sin(half), 2.0f),
                          std::nextafter((float)std::cos(half), -2.0f)};
        }
    }
    // whether a silhouette-BVH node may hold a silhouette seen from x (see buildSilhouetteCones)
    bool coneMaySee(int node, const V<2>& x) const {
        const auto& c = silCone[node];
        if (c[3] < 0.0f) return true;
        const Box<2>& b = silBvh.nodes[node].box;
        V<2> ctr = b.centroid(), u = x - ctr, e = (b.hi - b.lo) * 0.5f;
        float d2 = norm2(u), r2 = norm2(e);
        if (d2 <= r2 * 1.0001f) return true;
        float invd = 1.0f / std::sqrt(d2), sphi = std::sqrt(r2) * invd * 1.0001f;
        float cphi = std::sqrt(std::max(0.0f, 1.0f - sphi * sphi));
        float sinSum = c[2] * cphi + c[3] * sphi, cosSum = c[3] * cphi - c[2] * sphi;
        if (cosSum <= 0.0f) return true;
        float ca = (c[0] * u[0] + c[1] * u[1]) * invd;
        return std::fabs(ca) <= sinSum * 1.0001f + 1e-6f;
    }

    // Silhouette edge seen from x: the vertex rule with the unit direction from the edge's
    // closest point cp to x (both triangle planes contain the edge line, so only the
    // normalisation depends on cp).  Boundary edges and a query on the edge are silhouettes.
    bool isSilhouetteEdge(const SilEdge& e, const V<3>& x, const V<3>& cp, float precision, bool flip) const {
        if (e.f1 < 0) return true;
        V<3> view = x - cp;
        float l = norm(view);
        if (l <= 0.0f) return true;
        view = view * (1.0f / l);
        float sign = flip ? -1.0f : 1.0f;
        float d0 = dot(primNormal(e.f0), view), d1 = dot(primNormal(e.f1), view);
        if (std::fabs(d0) <= precision) return sign * d1 > precision;
        if (std::fabs(d1) <= precision) return sign * d0 > precision;
        return d0 * d1 < 0.0f;
    }

    bool isSilhouette(const SilVertex& s, const V<2>& x, float precision, bool flip) const {
        if (s.p1 < 0) return true;
        V<2> v = verts[s.v];
        V<2> view = x - v;
        float l = norm(view);
        if (l <= 0.0f) return true;
        view = view * (1.0f / l);
        float sign = flip ? -1.0f : 1.0f;
        float d0 = dot(primNormal(s.p0), view), d1 = dot(primNormal(s.p1), view);
        if (std::fabs(d0) <= precision) return sign * d1 > precision;
        if (std::fabs(d1) <= precision) return sign * d0 > precision;
        return d0 * d1 < 0.0f;
    }
};

template <int D>
struct Scene {
    Mesh<D> abs, refl;
    V<D> dmin, dmax;
    bool hasAbs() const { return abs.nPrims() > 0; }
    bool hasRefl() const { return refl.nPrims() > 0; }
    bool outsideBoundingDomain(const V<D>& x) const {
        for (int i = 0; i < D; i++) if (x[i] < dmin[i] || x[i] > dmax[i]) return true;
        return false;
    }
};

struct QueryCount { int64_t nodes = 0, prims = 0; };

// closest point: nearest-child-first traversal with pruning (fcpw findClosestPoint with
// sphere r2 = maxFloat, fcpw_geometric_queries.h:1069-1096)
template <int D>
bool closestPoint(const Mesh<D>& m, const V<D>& x, float& d, V<D>& pt, int& prim, QueryCount* qc) {
    prim = -1; d = FLT_MAX;
    if (m.nPrims() == 0) return false;
    float best2 = FLT_MAX;
    struct E { int n; float d2; };
    E stack[128];
    int sp = 0;
    const auto& nodes = m.bvh.nodes;
    stack[sp++] = {0, nodes[0].box.dist2(x)};
    while (sp > 0) {
        E e = stack[--sp];
        if (e.d2 > best2) continue;
        const auto& nd = nodes[e.n];
        if (qc) qc->nodes++;
        if (nd.left < 0) {
            for (int i = nd.start; i < nd.start + nd.count; i++) {
                int p = m.bvh.order[i];
                V<D> q; float dd = m.primClosest(p, x, q);
                if (qc) qc->prims++;
                if (dd * dd < best2) { best2 = dd * dd; d = dd; pt = q; prim = p; }
            }
            continue;
        }
        float dl = nodes[nd.left].box.dist2(x), dr = nodes[nd.right].box.dist2(x);
        int nearN = nd.left, farN = nd.right; float dn = dl, df = dr;
        if (dr < dl) { nearN = nd.right; farN = nd.left; dn = dr; df = dl; }
        if (df <= best2) stack[sp++] = {farN, df};
        if (dn <= best2) stack[sp++] = {nearN, dn};
    }
    return prim >= 0;
}

template <int D>
bool closestPointBrute(const Mesh<D>& m, const V<D>& x, float& d, V<D>& pt, int& prim) {
    prim = -1; d = FLT_MAX;
    for (int p = 0; p < m.nPrims(); p++) {
        V<D> q; float dd = m.primClosest(p, x, q);
        if (dd < d) { d = dd; pt = q; prim = p; }
    }
    return prim >= 0;
}

// first hit along a ray
template <int D>
bool intersectRay(const Mesh<D>& m, const V<D>& o, const V<D>& dir, float tMax, float& tHit, int& prim, bool brute) {
    prim = -1; tHit = tMax;
    if (m.nPrims() == 0) return false;
    if (brute) {
        for (int p = 0; p < m.nPrims(); p++) { float t; if (m.primRay(p, o, dir, tHit, t)) { if (prim < 0 || t < tHit) { tHit = t; prim = p; } } }
        return prim >= 0;
    }
    V<D> inv; for (int i = 0; i < D; i++) inv[i] = 1.0f / dir[i];
    int stack[128]; int sp = 0; stack[sp++] = 0;
    const auto& nodes = m.bvh.nodes;
    while (sp > 0) {
        int ni = stack[--sp];
        const auto& nd = nodes[ni];
        if (!nd.box.rayOverlap(o, inv, tHit)) continue;
        if (nd.left < 0) {
            for (int i = nd.start; i < nd.start + nd.count; i++) {
                int p = m.bvh.order[i]; float t;
                if (m.primRay(p, o, dir, tHit, t)) { if (prim < 0 || t < tHit) { tHit = t; prim = p; } }
            }
            continue;
        }
        stack[sp++] = nd.left; stack[sp++] = nd.right;
    }
    return prim >= 0;
}

// number of hits along a ray (all hits), for insideDomain
template <int D>
int countHits(const Mesh<D>& m, const V<D>& o, const V<D>& dir) {
    int hits = 0;
    for (int p = 0; p < m.nPrims(); p++) { float t; if (m.primRay(p, o, dir, FLT_MAX, t)) hits++; }
    return hits;
}

// closest silhouette point (2D), fcpw_geometric_queries.h:1254-1268
inline float starRadius2(const Mesh<2>& m, const V<2>& x, float minRadius, float maxRadius, float precision,
                         bool flip) {
    if (minRadius > maxRadius) return maxRadius;
    float best = FLT_MAX; bool found = false;
    for (const auto& s : m.sil) {
        float d = norm(x - m.verts[s.v]);
        if (d > maxRadius) continue;
        if (d >= best) continue;
        if (m.isSilhouette(s, x, precision, flip)) { best = d; found = true; }
    }
    return found ? std::max(best, minRadius) : std::max(maxRadius, minRadius);
}

// the same query over the silhouette-candidate BVH (nearest-first, pruned by the best
// distance and max radius, stopped once a silhouette within min radius is found); equal to
// starRadius2 (checked by tests)
inline float starRadius2Bvh(const Mesh<2>& m, const V<2>& x, float minRadius, float maxRadius, float precision,
                            bool flip) {
    if (minRadius > maxRadius) return maxRadius;
    if (m.sil.empty()) return std::max(maxRadius, minRadius);
    const bool cones = precision <= m.silConePrecision;
    float best = FLT_MAX, best2 = maxRadius < FLT_MAX ? maxRadius * maxRadius : FLT_MAX;
    bool found = false;
    const auto& nodes = m.silBvh.nodes;
    struct E { int n; float d2; };
    E stack[128];
    int sp = 0;
    stack[sp++] = {0, nodes[0].box.dist2(x)};
    while (sp > 0) {
        E e = stack[--sp];
        if (e.d2 > best2) continue;
        const auto& nd = nodes[e.n];
        if (nd.left < 0) {
            for (int i = nd.start; i < nd.start + nd.count; i++) {
                const auto& sv = m.sil[m.silBvh.order[i]];
                float d = norm(x - m.verts[sv.v]);
                if (d > maxRadius || d >= best) continue;
                if (m.isSilhouette(sv, x, precision, flip)) { best = d; best2 = d * d; found = true; }
            }
            if (found && best <= minRadius) break;
            continue;
        }
        float dl = nodes[nd.left].box.dist2(x), dr = nodes[nd.right].box.dist2(x);
        if (cones && !m.coneMaySee(nd.left, x)) dl = FLT_MAX;
        if (cones && !m.coneMaySee(nd.right, x)) dr = FLT_MAX;
        int nearN = nd.left, farN = nd.right; float dn = dl, df = dr;
        if (dr < dl) { nearN = nd.right; farN = nd.left; dn = dr; df = dl; }
        if (df <= best2) stack[sp++] = {farN, df};
        if (dn <= best2) stack[sp++] = {nearN, dn};
    }
    return found ? std::max(best, minRadius) : std::max(maxRadius, minRadius);
}

// closest silhouette point (3D): the closest point of the nearest silhouette edge
inline float starRadius3(const Mesh<3>& m, const V<3>& x, float minRadius, float maxRadius, float precision,
                         bool flip) {
    if (minRadius > maxRadius) return maxRadius;
    float best = FLT_MAX; bool found = false;
    for (const auto& e : m.silE) {
        V<3> cp;
        float d = closestPointSegmentD<3>(m.verts[e.va], m.verts[e.vb], x, cp);
        if (d > maxRadius || d >= best) continue;
        if (m.isSilhouetteEdge(e, x, cp, precision, flip)) { best = d; found = true; }
    }
    return found ? std::max(best, minRadius) : std::max(maxRadius, minRadius);
}

// the same over the silhouette-edge BVH (nearest-first, pruned by the best distance and the
// max radius, stopped once a silhouette within min radius is found); equal to starRadius3
inline float starRadius3Bvh(const Mesh<3>& m, const V<3>& x, float minRadius, float maxRadius, float precision,
                            bool flip) {
    if (minRadius > maxRadius) return maxRadius;
    if (m.silE.empty()) return std::max(maxRadius, minRadius);
    float best = FLT_MAX, best2 = maxRadius < FLT_MAX ? maxRadius * maxRadius : FLT_MAX;
    bool found = false;
    const auto& nodes = m.silBvh.nodes;
    struct E { int n; float d2; };
    E stack[128];
    int sp = 0;
    stack[sp++] = {0, nodes[0].box.dist2(x)};
    while (sp > 0) {
        E e = stack[--sp];
        if (e.d2 > best2) continue;
        const auto& nd = nodes[e.n];
        if (nd.left < 0) {
            for (int i = nd.start; i < nd.start + nd.count; i++) {
                const auto& se = m.silE[m.silBvh.order[i]];
                V<3> cp;
                float d = closestPointSegmentD<3>(m.verts[se.va], m.verts[se.vb], x, cp);
                if (d > maxRadius || d >= best) continue;
                if (m.isSilhouetteEdge(se, x, cp, precision, flip)) { best = d; best2 = d * d; found = true; }
            }
            if (found && best <= minRadius) break;
            continue;
        }
        float dl = nodes[nd.left].box.dist2(x), dr = nodes[nd.right].box.dist2(x);
        int nearN = nd.left, farN = nd.right; float dn = dl, df = dr;
        if (dr < dl) { nearN = nd.right; farN = nd.left; dn = dr; df = dl; }
        if (df <= best2) stack[sp++] = {farN, df};
        if (dn <= best2) stack[sp++] = {nearN, dn};
    }
    return found ? std::max(best, minRadius) : std::max(maxRadius, minRadius);
}

// Robin part of the star radius of a segment (ReflectanceLineSegment::computeSquaredStarRadius,
// reflectance_boundary_bvh/geometry.h:86-140, with RobinLineSegmentBound, robin_bounds.h:12-35):
// inside the radius (h / c) exp(c^2 / (h kappa_max)), c = sqrt(h kappa_max)/sqrt(2) clamped to the
// view-angle cosines of the segment's closest and farthest points, the reflectance of every
// point of the segment seen from x stays non-negative.  FCPW's epsilon is taken as FLT_EPSILON.
inline void robinSegmentBound(const Mesh<2>& m, int p, const V<2>& x, float& r2) {
    float kmax = m.robinMax.empty() ? 0.0f : m.robinMax[p];
    if (!(kmax > FLT_EPSILON)) return;            // perfectly reflecting: silhouettes only
    V<2> pa = m.vtx(p, 0), pb = m.vtx(p, 1), cp;
    float cd = closestPointSegment(pa, pb, x, cp);
    float cd2 = cd * cd;
    if (cd2 > r2) return;
    if (cd < FLT_EPSILON) { r2 = std::min(r2, cd); return; }
    V<2> n0 = segmentNormal(pa, pb);
    float h = std::fabs(dot(n0, pa - x));
    if (!(h > 0.0f)) { r2 = std::min(r2, cd2); return; }
    float da = norm(x - pa), db = norm(x - pb);
    V<2> fp = da >= db ? pa : pb;
    float fd = std::max(da, db);
    float cosUpper = std::fabs(dot((x - cp) * (1.0f / cd), n0));
    float cosLower = std::fabs(dot((x - fp) * (1.0f / fd), n0));
    float cosLine = std::sqrt(h * kmax) / std::sqrt(2.0f);
    float c = std::min(std::max(cosLine, cosLower), cosUpper);
    float rad = (h / c) * std::exp(c * c / (h * kmax));
    r2 = std::min(r2, rad * rad);
}
// Robin part of the star radius of a triangle (ReflectanceTriangle::computeSquaredStarRadius with
// RobinTriangleBound, reflectance_boundary_bvh/geometry.h:170-236, robin_bounds.h:52-80): h the
// distance to the triangle's plane, cosUpper / cosLower the view-angle cosines of its closest
// point and farthest vertex; when sqrt(h kappa_max) >= cosLower, c = sqrt(h kappa_max) / sqrt(3)
// clamped to [cosLower, cosUpper] gives the radius h^2 kappa_max / (c (h kappa_max - c^2)).
inline void robinTriangleBound(const Mesh<3>& m, int p, const V<3>& x, float& r2) {
    float kmax = m.robinMax.empty() ? 0.0f : m.robinMax[p];
    if (!(kmax > FLT_EPSILON)) return;
    V<3> pa = m.vtx(p, 0), pb = m.vtx(p, 1), pc = m.vtx(p, 2), cp;
    float cd = closestPointTriangle(pa, pb, pc, x, cp);
    float cd2 = cd * cd;
    if (cd2 > r2) return;
    if (cd < FLT_EPSILON) { r2 = std::min(r2, cd); return; }
    V<3> n0 = triangleNormal(pa, pb, pc);
    float h = std::fabs(dot(n0, pa - x));
    if (!(h > 0.0f)) { r2 = std::min(r2, cd2); return; }
    float da = norm(x - pa), db = norm(x - pb), dc = norm(x - pc);
    V<3> fp = pa; float fd = da;
    if (db > fd) { fp = pb; fd = db; }
    if (dc > fd) { fp = pc; fd = dc; }
    float cosUpper = std::fabs(dot((x - cp) * (1.0f / cd), n0));
    float cosLower = std::fabs(dot((x - fp) * (1.0f / fd), n0));
    float maxCos = std::sqrt(h * kmax);
    if (maxCos < cosLower) return;
    float c = std::min(std::max(maxCos / std::sqrt(3.0f), cosLower), cosUpper);
    float rad = h * h * kmax / (c * (h * kmax - c * c));
    r2 = std::min(r2, rad * rad);
}
template <int D>
inline void robinPrimBound(const Mesh<D>& m, int p, const V<D>& x, float& r2) {
    if constexpr (D == 2) robinSegmentBound(m, p, x, r2);
    else robinTriangleBound(m, p, x, r2);
}
// minimum over the reflecting primitives within sqrt(r2) of x (BVH pruned by box distance)
template <int D>
float robinRadius2(const Mesh<D>& m, const V<D>& x, float r2, bool brute = false) {
    if (m.nPrims() == 0 || m.robinMax.empty()) return r2;
    if (brute) { for (int p = 0; p < m.nPrims(); p++) robinPrimBound<D>(m, p, x, r2); return r2; }
    const auto& nodes = m.bvh.nodes;
    int stack[128]; int sp = 0; stack[sp++] = 0;
    while (sp > 0) {
        const auto& nd = nodes[stack[--sp]];
        if (nd.box.dist2(x) > r2) continue;
        if (nd.left < 0) { for (int i = nd.start; i < nd.start + nd.count; i++) robinPrimBound<D>(m, m.bvh.order[i], x, r2); continue; }
        stack[sp++] = nd.right; stack[sp++] = nd.left;
    }
    return r2;
}

// line of sight over the BVH: any reflecting primitive crossing the open segment p1 -> p2
template <int D>
bool hasLineOfSightBvh(const Mesh<D>& m, const V<D>& p1, const V<D>& p2) {
    V<D> d = p2 - p1; float L = norm(d);
    if (L <= 0.0f || m.nPrims() == 0) return true;
    V<D> dir = d * (1.0f / L);
    V<D> inv; for (int i = 0; i < D; i++) inv[i] = 1.0f / dir[i];
    int stack[128]; int sp = 0; stack[sp++] = 0;
    const auto& nodes = m.bvh.nodes;
    while (sp > 0) {
        const auto& nd = nodes[stack[--sp]];
        if (!nd.box.rayOverlap(p1, inv, L)) continue;
        if (nd.left < 0) {
            for (int i = nd.start; i < nd.start + nd.count; i++) {
                int p = m.bvh.order[i];
                float t;
                if (m.primRay(p, p1, dir, L, t) && t > 0.0f && t < L) return false;
            }
            continue;
        }
        stack[sp++] = nd.left; stack[sp++] = nd.right;
    }
    return true;
}

// line of sight: no reflecting primitive crosses the open segment p1 -> p2
template <int D>
bool hasLineOfSight(const Mesh<D>& m, const V<D>& p1, const V<D>& p2) {
    V<D> d = p2 - p1; float L = norm(d);
    if (L <= 0.0f) return true;
    V<D> dir = d * (1.0f / L);
    for (int p = 0; p < m.nPrims(); p++) {
        float t;
        if (m.primRay(p, p1, dir, L, t)) {
            if (t > 0.0f && t < L) return false;
        }
    }
    return true;
}

// sampleReflectingBoundary (2D): stochastic selection of a reflecting segment whose box
// overlaps the ball, with probability proportional to branchTraversalWeight(d^2 to the
// segment's box centre) * segment length, then a uniform point on it.  The returned pdf is
// the exact density of the sample (the estimator alpha*G*h/pdf stays unbiased for any pdf
// that is positive on the boundary inside the ball).
inline float branchWeight(float r2) {
    float r = std::max(std::sqrt(r2), 1e-2f);
    return std::fabs(1.0f / (4.0f * PI_F * r));  // HarmonicGreensFnFreeSpace<3>::evaluate
}
template <int D>
bool sampleReflecting(const Mesh<D>& m, const V<D>& x, float radius, const float* u,
                      V<D>& pt, V<D>& nrm, float& pdf) {
    float R2 = radius * radius;
    double total = 0.0;
    std::vector<std::pair<int, double>> cand;
    for (int p = 0; p < m.nPrims(); p++) {
        if (m.boxes[p].dist2(x) > R2) continue;
        double w = (double)branchWeight(norm2(m.boxes[p].centroid() - x)) * m.primArea(p);
        if (w <= 0.0) continue;
        cand.push_back({p, w}); total += w;
    }
    if (cand.empty()) return false;
    double target = u[0] * total, acc = 0.0; int chosen = cand.back().first; double wch = cand.back().second;
    for (auto& c : cand) { acc += c.second; if (target < acc) { chosen = c.first; wch = c.second; break; } }
    pt = m.primSample(chosen, u + 1);
    nrm = m.primNormal(chosen);
    pdf = (float)(wch / total) / m.primArea(chosen);
    return true;
}

// sampleReflectingBoundary by one stochastic descent of the binary BVH: at an inner node a
// child whose box meets the ball is taken with probability proportional to
// branchWeight(|box centre - x|^2); in the leaf a segment whose box meets the ball with
// probability proportional to branchWeight x length; u[0] is rescaled at every level.  pdf is
// the exact density of the sample (the product of the choice probabilities / length).
template <int D>
bool sampleReflectingBvh(const Mesh<D>& m, const V<D>& x, float radius, const float* u,
                         V<D>& pt, V<D>& nrm, float& pdf) {
    if (m.nPrims() == 0) return false;
    const float R2 = radius * radius;
    const auto& nodes = m.bvh.nodes;
    if (nodes[0].box.dist2(x) > R2) return false;
    double prob = 1.0;
    double uu = u[0];
    int n = 0;
    for (;;) {
        const auto& nd = nodes[n];
        if (nd.left < 0) {
            double total = 0.0;
            double w[64]; int cnt = std::min(nd.count, 64);
            for (int j = 0; j < cnt; j++) {
                int p = m.bvh.order[nd.start + j];
                w[j] = m.boxes[p].dist2(x) <= R2 ? (double)branchWeight(norm2(m.boxes[p].centroid() - x)) * m.primArea(p) : 0.0;
                total += w[j];
            }
            if (!(total > 0.0)) return false;
            double t = uu * total, acc = 0.0; int sel = -1;
            for (int j = 0; j < cnt; j++) {
                if (w[j] > 0.0 && (sel < 0 || t >= acc)) sel = j;
                acc += w[j];
            }
            int p = m.bvh.order[nd.start + sel];
            pt = m.primSample(p, u + 1);
            nrm = m.primNormal(p);
            pdf = (float)(prob * (w[sel] / total) / m.primArea(p));
            return true;
        }
        const auto& L = nodes[nd.left].box; const auto& Rb = nodes[nd.right].box;
        double wl = L.dist2(x) <= R2 ? (double)branchWeight(norm2(L.centroid() - x)) : 0.0;
        double wr = Rb.dist2(x) <= R2 ? (double)branchWeight(norm2(Rb.centroid() - x)) : 0.0;
        double total = wl + wr;
        if (!(total > 0.0)) return false;
        double t = uu * total;
        if (wl > 0.0 && (wr <= 0.0 || t < wl)) { prob *= wl / total; uu = std::min(t / wl, 0.99999999); n = nd.left; }
        else { prob *= wr / total; uu = std::min(std::max((t - wl) / wr, 0.0), 0.99999999); n = nd.right; }
    }
}

// ------------------------------------------------------------------------------------
// statistics accumulation (FP64 Welford + Chan merge, common.h:225-310,381-386)
// ------------------------------------------------------------------------------------
struct Stats {
    int64_t n = 0; double mean = 0.0, m2 = 0.0; int64_t len = 0, splits = 0;
    void add(double x) { n++; double d = x - mean; mean += d / (double)n; double d2 = x - mean; m2 += d * d2; }
};
inline void mergeInto(int64_t& n, double& mean, double& m2, const Stats& o) {
    if (o.n <= 0) return;
    int64_t N1 = n, N2 = o.n, N = N1 + N2;
    if (N1 == 0) { mean = o.mean; m2 = o.m2; }
    else {
        double d = o.mean - mean;
        mean = mean + d * ((double)N2 / (double)N);
        m2 = m2 + o.m2 + d * d * ((double)N1 * (double)N2 / (double)N);
    }
    n = N;
}

// ------------------------------------------------------------------------------------
// walk state (common.h:93-130)
// ------------------------------------------------------------------------------------
template <int D>
struct WalkState {
    float totalReflectingBoundaryContribution = 0.0f;
    float totalSourceContribution = 0.0f;
    V<D> currentPt, currentNormal, prevDirection;
    float prevDistance = 0.0f, throughput = 1.0f;
    int walkLength = 0;
    bool onReflectingBoundary = false;
};

enum Code { Reached = 0, RR = 1, MaxLen = 2, Escaped = 3 };

struct SolveCtx {
    const zg_pde* pde;
    const zg_walk_settings* ws;
    uint64_t seed; uint32_t streamOffset; int rngMode;
};

// applyWeightWindow (walk_on_spheres.h:192-226, walk_on_stars.h:309-343)
template <int D>
bool applyWeightWindow(const zg_walk_settings& ws, Pcg32& rng, WalkState<D>& st, std::deque<WalkState<D>>& q) {
    if (st.throughput > ws.splitting_threshold) {
        float left = st.throughput - ws.splitting_threshold;
        st.throughput = ws.splitting_threshold;
        WalkState<D> s2; s2.currentPt = st.currentPt; s2.currentNormal = st.currentNormal;
        s2.prevDirection = st.prevDirection; s2.prevDistance = st.prevDistance;
        s2.throughput = st.throughput; s2.walkLength = 0; s2.onReflectingBoundary = st.onReflectingBoundary;
        while (left > ws.splitting_threshold) { left -= ws.splitting_threshold; q.push_back(s2); }
        if (rng.nextFloat() < left / ws.splitting_threshold) q.push_back(s2);
    } else if (st.throughput < ws.russian_roulette_threshold) {
        float survival = st.throughput / ws.russian_roulette_threshold;
        if (survival < rng.nextFloat()) { st.throughput = 0.0f; return true; }
        st.throughput = ws.russian_roulette_threshold;
    }
    return false;
}

// ----- Walk on Spheres: walk (walk_on_spheres.h:228-291) -----
template <int D>
Code wosWalk(const Scene<D>& sc, const SolveCtx& cx, float dist, GreensBall<D>& gf, Pcg32& rng,
             WalkState<D>& st, std::deque<WalkState<D>>& q, QueryCount* qc) {
    const zg_walk_settings& ws = *cx.ws; const zg_pde& pde = *cx.pde;
    while (dist > ws.epsilon_shell_for_absorbing_boundary) {
        gf.updateBall(st.currentPt, dist);
        if (!ws.ignore_source_contribution) {  // computeSourceContribution :170-190
            float sc_;
            V<D> nz = V<D>::zero();
            if (pde.is_source_constant) {
                sc_ = gf.norm() * evalField<D>(pde.source, st.currentPt, nz, false, pde.absorption_coeff);
            } else {
                float r, pdf;
                V<D> y = gf.sampleVolume(rng, r, pdf);
                sc_ = gf.norm() * evalField<D>(pde.source, y, nz, false, pde.absorption_coeff);
            }
            st.totalSourceContribution += st.throughput * sc_;
        }
        V<D> dir = sampleUnitSphere<D>(rng);
        st.currentPt = st.currentPt + dist * dir;
        st.throughput *= gf.directionSampledPoissonKernel(st.currentPt);
        if (sc.outsideBoundingDomain(st.currentPt)) return Escaped;
        if (applyWeightWindow(ws, rng, st, q)) return RR;
        st.walkLength++;
        if (st.walkLength >= ws.max_walk_length) return MaxLen;
        if (pde.absorption_coeff > 0.0f && ws.steps_before_applying_tikhonov == st.walkLength)
            gf.makeYukawa(pde.absorption_coeff);
        V<D> cp; int prim;
        closestPoint(sc.abs, st.currentPt, dist, cp, prim, qc);
    }
    return Reached;
}

// getTerminalContribution (walk_on_spheres.h:293-318)
template <int D>
float wosTerminal(const Scene<D>& sc, const SolveCtx& cx, Code code, WalkState<D>& st, QueryCount* qc) {
    const zg_walk_settings& ws = *cx.ws;
    if ((code == Reached || code == MaxLen) && !ws.ignore_absorbing_boundary_contribution) {
        float d; V<D> cp; int prim;
        bool found = closestPoint(sc.abs, st.currentPt, d, cp, prim, qc);
        bool aligned = false;
        if (found) {
            V<D> n = sc.abs.primNormal(prim);
            float sd = dot(st.currentPt - cp, n);
            st.currentPt = cp; st.currentNormal = n;
            aligned = ws.solve_double_sided && sd > 0.0f;
        }
        return evalField<D>(cx.pde->dirichlet, st.currentPt, st.currentNormal, aligned, cx.pde->absorption_coeff);
    }
    return 0.0f;
}

// rng_mode 1: walk w's own generator; rng_mode 0: the point's one generator (that of its walk 0).
// rng_mode 2 is a deliberately flawed per-walk scheme kept as a counter-example for the
// independence tests: one shared initstate on increments 2((p << 32) | w) + 1, under which the
// states of a point's walks lie on an arithmetic progression in w and their draws correlate.
inline Pcg32 makeRng(const SolveCtx& cx, int64_t p, int w) {
    if (cx.rngMode == 2) return Pcg32(cx.seed, ((uint64_t)p << 32) | (uint64_t)(uint32_t)(cx.streamOffset + (uint32_t)w));
    return walkRng(cx.seed, (uint32_t)p, cx.streamOffset + (cx.rngMode == 1 ? (uint32_t)w : 0u));
}

// estimateSolution (walk_on_spheres.h:320-414)
template <int D>
void wosEstimate(const Scene<D>& sc, const SolveCtx& cx, int64_t p, const V<D>& pt, uint8_t type,
                 float distA, int nWalks, Stats& out, int64_t codes[4], QueryCount* qc,
                 int64_t prevN, float* firstSphereRadius) {
    const zg_walk_settings& ws = *cx.ws; const zg_pde& pde = *cx.pde;
    bool hasPrev = prevN > 0;
    if (type == ZG_SAMPLE_ON_ABSORBING_BOUNDARY) {
        if (!hasPrev) {
            float tc = 0.0f;
            if (!ws.ignore_absorbing_boundary_contribution)
                tc = evalField<D>(pde.dirichlet, pt, V<D>::zero(), false, pde.absorption_coeff);
            out.add(tc);
            if (firstSphereRadius) *firstSphereRadius = 0.0f;
        }
        return;
    } else if (distA <= ws.epsilon_shell_for_absorbing_boundary) {
        nWalks = std::min(nWalks, 1);
    }
    if (!hasPrev && firstSphereRadius) *firstSphereRadius = distA;
    Pcg32 pointRng = makeRng(cx, p, 0);
    std::deque<WalkState<D>> q;
    for (int w = 0; w < nWalks; w++) {
        Pcg32 walkRng = cx.rngMode != 0 ? makeRng(cx, p, w) : Pcg32();
        Pcg32& rng = cx.rngMode != 0 ? walkRng : pointRng;
        WalkState<D> st0; st0.currentPt = pt; st0.currentNormal = V<D>::zero(); st0.prevDirection = V<D>::zero();
        q.push_back(st0);
        int splits = -1; float total = 0.0f; bool success = false;
        while (!q.empty()) {
            WalkState<D> st = q.front(); q.pop_front();
            splits++;
            GreensBall<D> gf;
            if (pde.absorption_coeff > 0.0f && ws.steps_before_applying_tikhonov == 0) gf.makeYukawa(pde.absorption_coeff);
            float dist = distA;
            if (splits > 0) { V<D> cp; int prim; closestPoint(sc.abs, st.currentPt, dist, cp, prim, qc); }
            Code code = wosWalk<D>(sc, cx, dist, gf, rng, st, q, qc);
            codes[code]++;
            if (code != Escaped) {
                float term = wosTerminal<D>(sc, cx, code, st, qc);
                total += st.throughput * term + st.totalSourceContribution;
                out.len += st.walkLength;
                success = true;
            }
        }
        if (success) { out.add((double)total); out.splits += splits; }
    }
}

template <int D>
Code wostWalk(const Scene<D>& sc, const SolveCtx& cx, float distA, float firstSphereRadius, bool flip,
              GreensBall<D>& gf, Pcg32& rng, WalkState<D>& st, std::deque<WalkState<D>>& q, QueryCount* qc);

// terminal value of a walk-on-stars walk (estimateSolution :627-645): projection to the
// absorbing boundary when it reached the boundary
template <int D>
float wostTerminal(const Scene<D>& sc, const SolveCtx& cx, Code code, WalkState<D>& st, QueryCount* qc) {
    const zg_walk_settings& ws = *cx.ws;
    if (code == Reached && !ws.ignore_absorbing_boundary_contribution) {
        float d; V<D> cp; int prim;
        bool aligned = false;
        if (closestPoint(sc.abs, st.currentPt, d, cp, prim, qc)) {
            V<D> n = sc.abs.primNormal(prim);
            float sd = dot(st.currentPt - cp, n);
            st.currentPt = cp; st.currentNormal = n;
            aligned = ws.solve_double_sided && sd > 0.0f;
        }
        return evalField<D>(cx.pde->dirichlet, st.currentPt, st.currentNormal, aligned, cx.pde->absorption_coeff);
    }
    return 0.0f;
}

// gradient statistics of one point (SampleStatistics gradient members, common.h:226-247,
// 288-309), FP64
template <int D>
struct GradStats {
    int64_t n = 0;
    double mean[D] = {}, m2[D] = {};
    double firstSource = 0.0, derivative = 0.0;
    void add(const double* g) {
        n++;
        for (int i = 0; i < D; i++) {
            double d = g[i] - mean[i];
            mean[i] += d / (double)n;
            m2[i] += d * (g[i] - mean[i]);
        }
    }
};

// control variates of the gradient estimate for the next antithetic pair
struct GradCv {
    int mode;                 // 0: the running means before every pair; 1: frozen per stage
    int64_t prevN; double prevMean, prevFirst;
    float b = 0.0f, s = 0.0f; // current values
    int stageEnd = 0;
    // the solution mean and the mean first-step source term over prevN + o.n estimates
    void refresh(const Stats& o, double first) {
        int64_t N = prevN + o.n;
        double mean = N > 0 ? (prevMean * (double)prevN + o.mean * (double)o.n) / (double)N : 0.0;
        b = (float)mean;
        s = (float)((prevFirst + first) / (double)std::max<int64_t>(1, N));
    }
    void atPair(int w, const Stats& o, double first) {
        if (mode == 0) { refresh(o, first); return; }
        // stages [0,1), [1,2), [2,4), [4,8), ...: the values are refreshed when a stage starts
        if (w == 0 || w >= stageEnd) { refresh(o, first); stageEnd = w == 0 ? 1 : 2 * w; }
    }
};

// SDF-grid geometry (defined with its walk below): the gradient estimator continues its walks
// over it when given one
template <int D> struct SdfGeom;
template <int D>
Code sdfWalk(const SdfGeom<D>& g, const SolveCtx& cx, float dist, GreensBall<D>& gf, Pcg32& rng, WalkState<D>& st,
             std::deque<WalkState<D>>& q);
template <int D>
float sdfTerminal(const SdfGeom<D>& g, const SolveCtx& cx, Code code, const WalkState<D>& st);

// estimateSolutionAndGradient (walk_on_spheres.h:416-593; walk on stars :664-848, whose first
// ball has radius 0.99 min(d_absorbing, d_reflecting) and whose walk goes on as a walk-on-stars
// walk from the boundary point).  rng_mode 0 consumes one generator
// per point in the reference's order (stratified table, then every walk; the antithetic half
// replays the generator saved after the first step).  rng_mode 1 is the GPU's scheme: the table
// comes from walkRng(~seed, p, stream_offset) and both halves of pair w start from
// walkRng(seed, p, stream_offset + w), the second half redrawing the first step before
// mirroring it (the same numbers, so the same walk continuation as the reference's replay).
template <int D>
void wosEstimateGrad(const Scene<D>& sc, const SolveCtx& cx, GradCv cv, int64_t p, const V<D>& pt, float distA,
                     int nWalks, const V<D>& dirDeriv, Stats& out, GradStats<D>& g, int64_t codes[4],
                     QueryCount* qc, float* firstSphereRadius, bool wost = false, float distR = FLT_MAX,
                     const SdfGeom<D>* sdf = nullptr) {
    const zg_walk_settings& ws = *cx.ws; const zg_pde& pde = *cx.pde;
    int nHalves = 1;
    if (ws.use_gradient_antithetic_variates) { nWalks = std::max(1, nWalks / 2); nHalves = 2; }
    const float R0 = 0.99f * (wost ? std::min(distA, distR) : distA);  // ZOMBIE_RADIUS_SHRINK_PERCENTAGE
    if (firstSphereRadius) *firstSphereRadius = R0;
    Pcg32 pointRng = makeRng(cx, p, 0);
    std::vector<float> strat;
    if (cx.rngMode != 0) { Pcg32 t = walkRng(~cx.seed, (uint32_t)p, cx.streamOffset); stratifiedSamples(D - 1, 2 * nWalks, t, strat); }
    else stratifiedSamples(D - 1, 2 * nWalks, pointRng, strat);
    std::deque<WalkState<D>> q;
    for (int w = 0; w < nWalks; w++) {
        float sourceRadius = 0.0f, sourcePdf = 1.0f, boundaryPdf = 1.0f;
        V<D> sourcePt = V<D>::zero(), boundaryPt = V<D>::zero();
        Pcg32 savedRng;
        float cvB = 0.0f, cvS = 0.0f;
        if (ws.use_gradient_control_variates) { cv.atPair(w, out, g.firstSource); cvB = cv.b; cvS = cv.s; }
        for (int half = 0; half < nHalves; half++) {
            Pcg32 pairRng = cx.rngMode != 0 ? makeRng(cx, p, w) : Pcg32();
            Pcg32& rng = cx.rngMode != 0 ? pairRng : pointRng;
            // the first step: GPU scheme redraws it in the second half, then mirrors
            bool draw = half == 0 || cx.rngMode != 0;
            WalkState<D> st; st.currentPt = pt; st.currentNormal = V<D>::zero(); st.prevDirection = V<D>::zero();
            GreensBall<D> gf;
            if (pde.absorption_coeff > 0.0f && ws.steps_before_applying_tikhonov == 0) gf.makeYukawa(pde.absorption_coeff);
            gf.updateBall(st.currentPt, R0);
            float firstSource = 0.0f;
            V<D> sourceDir = V<D>::zero();
            if (!ws.ignore_source_contribution) {
                float gNorm = gf.norm();
                V<D> nz = V<D>::zero();
                if (pde.is_source_constant) {
                    firstSource = gNorm * evalField<D>(pde.source, st.currentPt, nz, false, pde.absorption_coeff);
                } else {
                    if (draw) {
                        V<D> sd = sphereFromU<D>(&strat[(size_t)(D - 1) * (2 * w + 0)]);
                        sourcePt = gf.sampleVolumeDir(sd, rng, sourceRadius, sourcePdf);
                    }
                    if (half == 1) { V<D> sdv = sourcePt - st.currentPt; sourcePt = st.currentPt - sdv; }
                    firstSource = gNorm * evalField<D>(pde.source, sourcePt, nz, false, pde.absorption_coeff);
                    sourceDir = gf.gradient(sourceRadius, sourcePt) * (1.0f / (sourcePdf * gNorm));
                }
                st.totalSourceContribution += st.throughput * firstSource;
            }
            if (draw) {
                const float* u = &strat[(size_t)(D - 1) * (2 * w + 1)];
                V<D> bd;
                if (ws.use_cosine_sampling_for_derivatives) {
                    bd = hemisphereCosineFromU<D>(u);
                    if (rng.nextFloat() < 0.5f) bd[D - 1] *= -1.0f;
                    boundaryPdf = 0.5f * pdfHemisphereCosine<D>(std::fabs(bd[D - 1]));
                    transformCoordinates(dirDeriv, bd);
                } else {
                    bd = sphereFromU<D>(u);
                    boundaryPdf = gf.pdfSampleSphereUniform(1.0f);
                }
                boundaryPt = gf.c + gf.R * bd;
            }
            if (half == 1) { V<D> bdv = boundaryPt - st.currentPt; boundaryPt = st.currentPt - bdv; }
            st.prevDistance = gf.R;
            st.prevDirection = (boundaryPt - st.currentPt) * (1.0f / gf.R);
            st.currentPt = boundaryPt;
            st.throughput *= gf.poissonKernel() / boundaryPdf;
            V<D> boundaryDir = gf.poissonKernelGradient(boundaryPt) * (1.0f / (boundaryPdf * st.throughput));
            if (cx.rngMode == 0) { if (half == 0) savedRng = rng; else rng = savedRng; }

            q.push_back(st);
            int splits = -1; float total = 0.0f; bool success = false;
            while (!q.empty()) {
                WalkState<D> s2 = q.front(); q.pop_front();
                splits++;
                if (splits > 0) {
                    gf = GreensBall<D>();
                    if (pde.absorption_coeff > 0.0f && ws.steps_before_applying_tikhonov == 0) gf.makeYukawa(pde.absorption_coeff);
                }
                float dist; V<D> cp; int prim;
                if (sdf) dist = std::fabs(sdf->value(s2.currentPt));
                else closestPoint(sc.abs, s2.currentPt, dist, cp, prim, qc);
                Code code = sdf ? sdfWalk<D>(*sdf, cx, dist, gf, rng, s2, q)
                            : wost ? wostWalk<D>(sc, cx, dist, 0.0f, false, gf, rng, s2, q, qc)
                                   : wosWalk<D>(sc, cx, dist, gf, rng, s2, q, qc);
                codes[code]++;
                if (code != Escaped) {
                    float term = sdf ? sdfTerminal<D>(*sdf, cx, code, s2)
                                 : wost ? wostTerminal<D>(sc, cx, code, s2, qc) : wosTerminal<D>(sc, cx, code, s2, qc);
                    total += s2.throughput * term + s2.totalReflectingBoundaryContribution + s2.totalSourceContribution;
                    out.len += s2.walkLength;
                    success = true;
                }
            }
            if (success) {
                float boundaryContribution = total - firstSource;
                double est[D];
                float dd = 0.0f;
                for (int i = 0; i < D; i++) {
                    float bi = (boundaryContribution - cvB) * boundaryDir[i];
                    float si = (firstSource - cvS) * sourceDir[i];
                    dd += bi * dirDeriv[i];
                    dd += si * dirDeriv[i];
                    est[i] = (double)(bi + si);
                }
                out.add((double)total);
                g.firstSource += firstSource;
                g.add(est);
                g.derivative += dd;
                out.splits += splits;
            }
        }
    }
}

// ----- Walk on Stars (walk_on_stars.h:189-662), 2D and 3D -----
template <int D>
void wostReflectingContribution(const Scene<D>& sc, const SolveCtx& cx, GreensBall<D>& gf, float starRadius,
                                bool flip, Pcg32& rng, WalkState<D>& st) {
    const zg_walk_settings& ws = *cx.ws; const zg_pde& pde = *cx.pde;
    if (!(sc.hasRefl() && !ws.ignore_reflecting_boundary_contribution)) return;
    float u[D]; for (int i = 0; i < D; i++) u[i] = rng.nextFloat();
    V<D> bpt, bn; float bpdf;
    bool ok = false;
    ok = sampleReflectingBvh<D>(sc.refl, st.currentPt, starRadius, u, bpt, bn, bpdf);
    if (!ok) return;
    V<D> toS = bpt - st.currentPt;
    float distS = norm(toS);
    float alpha = st.onReflectingBoundary ? 2.0f : 1.0f;
    bool aligned = false;
    if (ws.solve_double_sided) {
        toS = toS * (1.0f / distS);
        if (flip) aligned = true;
        else if (dot(toS, bn) < -ws.silhouette_precision) {
            bool flipS = true;
            if (alpha > 1.0f) flipS = dot(toS, st.currentNormal) < -ws.silhouette_precision;
            if (flipS) aligned = true;
        }
    }
    V<D> bnS = bn * (aligned ? -1.0f : 1.0f);
    if (bpdf > 0.0f && distS < starRadius) {
        V<D> p1 = st.onReflectingBoundary ? offsetPoint<D>(st.currentPt, st.currentNormal * -1.0f) : st.currentPt;
        V<D> p2 = offsetPoint<D>(bpt, bnS * -1.0f);
        bool blocked = false;
        blocked = !hasLineOfSightBvh<D>(sc.refl, p1, p2);
        if (!blocked) {
            float r = std::max(gf.rClamp, distS);
            float G = gf.evaluate(r);
            bool retAligned = ws.solve_double_sided && aligned;
            float h = evalField<D>(pde.robin, bpt, bn, retAligned, pde.absorption_coeff);
            st.totalReflectingBoundaryContribution += st.throughput * alpha * G * h / bpdf;
        }
    }
}

template <int D>
float wostStarRadius(const Scene<D>& sc, const SolveCtx& cx, const V<D>& x, float distA, bool flip) {
    const zg_walk_settings& ws = *cx.ws;
    float r = distA;
    if constexpr (D == 2) {
        const float minR = ws.epsilon_shell_for_reflecting_boundary;
        if (sc.hasRefl()) {
            r = starRadius2Bvh(sc.refl, x, minR, distA, ws.silhouette_precision, flip);
            // Robin boundaries also bound the radius (populateStarRadiusQueryForRobinBoundary,
            // fcpw_geometric_queries.h:1272-1289)
            if (minR <= distA && sc.refl.hasRobin()) {
                float r2 = robinRadius2<2>(sc.refl, x, distA < FLT_MAX ? distA * distA : FLT_MAX);
                r = std::max(std::min(r, std::sqrt(r2)), minR);
            }
        }
    } else {
        const float minR = ws.epsilon_shell_for_reflecting_boundary;
        if (sc.hasRefl()) {
            r = starRadius3Bvh(sc.refl, x, minR, distA, ws.silhouette_precision, flip);
            if (minR <= distA && sc.refl.hasRobin()) {
                float r2 = robinRadius2<3>(sc.refl, x, distA < FLT_MAX ? distA * distA : FLT_MAX);
                r = std::max(std::min(r, std::sqrt(r2)), minR);
            }
        }
    }
    if (sc.hasRefl() && ws.epsilon_shell_for_reflecting_boundary <= distA)
        r = std::max(0.99f * r, ws.epsilon_shell_for_reflecting_boundary);
    return r;
}

// PDE::hasNonZeroRobinCoeff (pde.h:62-78)
template <int D>
bool hasNonZeroRobinCoeff(const zg_pde& pde, const V<D>& x) {
    if (pde.are_robin_conditions_pure_neumann) return false;
    V<D> n = V<D>::zero(); n[0] = 1.0f;
    return std::fabs(evalField<D>(pde.robin_coeff, x, n, true, pde.absorption_coeff)) > 0.0f ||
           std::fabs(evalField<D>(pde.robin_coeff, x, n, false, pde.absorption_coeff)) > 0.0f;
}

template <int D>
Code wostWalk(const Scene<D>& sc, const SolveCtx& cx, float distA, float firstSphereRadius, bool flip,
              GreensBall<D>& gf, Pcg32& rng, WalkState<D>& st, std::deque<WalkState<D>>& q, QueryCount* qc) {
    const zg_walk_settings& ws = *cx.ws; const zg_pde& pde = *cx.pde;
    bool firstStep = true;
    while (distA > ws.epsilon_shell_for_absorbing_boundary) {
        float starRadius;
        if (firstStep && firstSphereRadius > 0.0f) {
            starRadius = firstSphereRadius;
        } else {
            flip = false;
            if (ws.solve_double_sided && st.onReflectingBoundary) {
                if (st.prevDistance > 0.0f && dot(st.prevDirection, st.currentNormal) < 0.0f) {
                    st.currentNormal = st.currentNormal * -1.0f; flip = true;
                }
            }
            if (ws.steps_before_using_maximal_spheres <= st.walkLength) starRadius = distA;
            else if (st.onReflectingBoundary && hasNonZeroRobinCoeff<D>(pde, st.currentPt))
                starRadius = ws.epsilon_shell_for_reflecting_boundary;   // reflectance vanishes on the boundary
            else starRadius = wostStarRadius<D>(sc, cx, st.currentPt, distA, flip);
        }
        gf.updateBall(st.currentPt, starRadius);
        V<D> dir = sampleUnitSphere<D>(rng);
        if (st.onReflectingBoundary && dot(st.currentNormal, dir) > 0.0f) dir = dir * -1.0f;
        // intersectReflectingBoundary (fcpw_geometric_queries.h:1178-1196)
        V<D> o = st.onReflectingBoundary ? offsetPoint<D>(st.currentPt, st.currentNormal * -1.0f) : st.currentPt;
        float tHit; int prim = -1;
        bool hit = sc.hasRefl() && intersectRay(sc.refl, o, dir, starRadius, tHit, prim, false);
        V<D> ipt, inrm = V<D>::zero(); float idist;  // IntersectionPoint() default normal is zero
        if (hit) { ipt = o + tHit * dir; inrm = sc.refl.primNormal(prim); idist = tHit;
            // fcpw reports the hit point on the primitive
            if constexpr (D == 2) {
                V<2> pa = sc.refl.vtx(prim, 0), pb = sc.refl.vtx(prim, 1); float t2, s2;
                if (raySegment(o, dir, FLT_MAX, pa, pb, t2, s2)) ipt = pa + (pb - pa) * s2;
            }
        } else { ipt = o + starRadius * dir; idist = starRadius; }
        wostReflectingContribution<D>(sc, cx, gf, starRadius, flip, rng, st);
        if (!ws.ignore_source_contribution) {  // computeSourceContribution :252-275
            float r, pdf;
            V<D> y = gf.sampleVolumeDir(dir, rng, r, pdf);
            if (r <= idist) {
                float s_ = gf.norm() * evalField<D>(pde.source, y, V<D>::zero(), false, pde.absorption_coeff);
                st.totalSourceContribution += st.throughput * s_;
            }
        }
        st.prevDistance = idist;
        st.prevDirection = dir;
        st.currentPt = ipt;
        st.currentNormal = inrm;
        st.onReflectingBoundary = hit;
        // computeWalkStepThroughput (:277-307)
        float thr;
        if (st.onReflectingBoundary && st.prevDistance > FLT_EPSILON) {
            float kappa = 0.0f;
            V<D> nr = st.currentNormal;
            if (!pde.are_robin_conditions_pure_neumann) {
                bool fl = false;
                if (ws.solve_double_sided) { fl = dot(st.prevDirection, st.currentNormal) < 0.0f; if (fl) nr = nr * -1.0f; }
                kappa = evalField<D>(pde.robin_coeff, st.currentPt, st.currentNormal, ws.solve_double_sided && fl,
                                     pde.absorption_coeff);
            }
            thr = std::min(std::max(gf.reflectance(st.prevDistance, st.prevDirection, nr, kappa), 0.0f),
                           pde.are_robin_coeffs_nonnegative ? 1.0f : 2.0f);
        } else {
            thr = gf.directionSampledPoissonKernel(st.currentPt);
        }
        st.throughput *= thr;
        if (!st.onReflectingBoundary && sc.outsideBoundingDomain(st.currentPt)) return Escaped;
        if (applyWeightWindow(ws, rng, st, q)) return RR;
        st.walkLength++;
        if (st.walkLength >= ws.max_walk_length) return MaxLen;
        if (pde.absorption_coeff > 0.0f && ws.steps_before_applying_tikhonov == st.walkLength)
            gf.makeYukawa(pde.absorption_coeff);
        V<D> cp; int pr;
        closestPoint(sc.abs, st.currentPt, distA, cp, pr, qc);
        firstStep = false;
    }
    return Reached;
}

template <int D>
void wostEstimate(const Scene<D>& sc, const SolveCtx& cx, int64_t p, const V<D>& pt, const V<D>& normal,
                  uint8_t type, bool estAligned, float distA, int nWalks, Stats& out, int64_t codes[4],
                  QueryCount* qc, int64_t prevN, float* firstSphereRadiusIO) {
    const zg_walk_settings& ws = *cx.ws; const zg_pde& pde = *cx.pde;
    bool hasPrev = prevN > 0;
    if (type == ZG_SAMPLE_ON_ABSORBING_BOUNDARY) {
        if (!hasPrev) {
            float tc = 0.0f;
            if (!ws.ignore_absorbing_boundary_contribution)
                tc = evalField<D>(pde.dirichlet, pt, V<D>::zero(), ws.solve_double_sided && estAligned, pde.absorption_coeff);
            out.add(tc);
            if (firstSphereRadiusIO) *firstSphereRadiusIO = 0.0f;
        }
        return;
    } else if (distA <= ws.epsilon_shell_for_absorbing_boundary) {
        nWalks = std::min(nWalks, 1);
    }
    V<D> curN = normal, prevDir = normal; float prevDist = FLT_MAX; bool flip = false;
    if (ws.solve_double_sided && type == ZG_SAMPLE_ON_REFLECTING_BOUNDARY && estAligned) {
        curN = curN * -1.0f; prevDir = prevDir * -1.0f; flip = true;
    }
    float fsr = firstSphereRadiusIO ? *firstSphereRadiusIO : 0.0f;
    if (!hasPrev) {
        if (distA <= ws.epsilon_shell_for_absorbing_boundary || ws.steps_before_using_maximal_spheres == 0) fsr = distA;
        else if (type == ZG_SAMPLE_ON_REFLECTING_BOUNDARY && hasNonZeroRobinCoeff<D>(pde, pt))
            fsr = ws.epsilon_shell_for_reflecting_boundary;
        else fsr = wostStarRadius<D>(sc, cx, pt, distA, flip);
        if (firstSphereRadiusIO) *firstSphereRadiusIO = fsr;
    }
    Pcg32 pointRng = makeRng(cx, p, 0);
    std::deque<WalkState<D>> q;
    for (int w = 0; w < nWalks; w++) {
        Pcg32 walkRng = cx.rngMode != 0 ? makeRng(cx, p, w) : Pcg32();
        Pcg32& rng = cx.rngMode != 0 ? walkRng : pointRng;
        WalkState<D> st0; st0.currentPt = pt; st0.currentNormal = curN; st0.prevDirection = prevDir;
        st0.prevDistance = prevDist; st0.onReflectingBoundary = type == ZG_SAMPLE_ON_REFLECTING_BOUNDARY;
        q.push_back(st0);
        int splits = -1; float total = 0.0f; bool success = false;
        while (!q.empty()) {
            WalkState<D> st = q.front(); q.pop_front();
            splits++;
            GreensBall<D> gf;
            if (pde.absorption_coeff > 0.0f && ws.steps_before_applying_tikhonov == 0) gf.makeYukawa(pde.absorption_coeff);
            float dA = distA, fs = fsr;
            if (splits > 0) { V<D> cp; int pr; closestPoint(sc.abs, st.currentPt, dA, cp, pr, qc); fs = 0.0f; }
            Code code = wostWalk<D>(sc, cx, dA, fs, flip, gf, rng, st, q, qc);
            codes[code]++;
            if (code != Escaped) {
                float term = wostTerminal<D>(sc, cx, code, st, qc);
                total += st.throughput * term + st.totalReflectingBoundaryContribution + st.totalSourceContribution;
                out.len += st.walkLength;
                success = true;
            }
        }
        if (success) { out.add((double)total); out.splits += splits; }
    }
}

// ------------------------------------------------------------------------------------
// thread pool over points (tbb::parallel_for replacement)
// ------------------------------------------------------------------------------------
template <class F>
void parallelFor(int64_t n, int nThreads, F f) {
    if (nThreads <= 1 || n < 64) { for (int64_t i = 0; i < n; i++) f(i, 0); return; }
    std::atomic<int64_t> next{0};
    const int64_t chunk = 8;
    std::vector<std::thread> th;
    for (int t = 0; t < nThreads; t++) {
        th.emplace_back([&, t]() {
            for (;;) {
                int64_t b = next.fetch_add(chunk);
                if (b >= n) break;
                int64_t e = std::min(n, b + chunk);
                for (int64_t i = b; i < e; i++) f(i, t);
            }
        });
    }
    for (auto& x : th) x.join();
}

struct SceneHandle {
    int dim;
    std::unique_ptr<Scene<2>> s2;
    std::unique_ptr<Scene<3>> s3;
};

template <int D>
void buildScene(Scene<D>& sc, const float* V_, int64_t nV, const int32_t* I, int64_t nP, const uint8_t* refl,
                const float* dmin, const float* dmax, int leafSize, int nBuckets, bool dbl, float prec) {
    for (Mesh<D>* m : {&sc.abs, &sc.refl}) {
        m->verts.resize(nV);
        for (int64_t i = 0; i < nV; i++) m->verts[i] = load<D>(V_ + i * D);
    }
    for (int64_t p = 0; p < nP; p++) {
        Mesh<D>& m = (refl && refl[p]) ? sc.refl : sc.abs;
        for (int k = 0; k < D; k++) m.idx.push_back(I[p * D + k]);
        m.origId.push_back((int)p);
    }
    sc.abs.build(leafSize, nBuckets);
    sc.refl.build(leafSize, nBuckets);
    sc.refl.buildSilhouettes(dbl, prec);
    sc.dmin = load<D>(dmin); sc.dmax = load<D>(dmax);
}

// ------------------------------------------------------------------------------------
// SDF-grid Dirichlet geometry (utils/sdf_grid_geometric_queries.h:87-190, 247-272): the
// distance is |sdf(x)| of a trilinearly interpolated grid, the projection steepest descent on
// the SDF with central-difference gradients, clamped into the grid
// ------------------------------------------------------------------------------------
template <int D>
struct SdfGeom {
    zg_function f;          // the grid (ZG_FN_DENSE_GRID, interpolated)
    V<D> dmin, dmax;        // bounding domain of the walks
    float value(const V<D>& x) const { return gridValue<D>(f, x); }
    V<D> clampIn(V<D> x) const {
        for (int a = 0; a < D; a++) x[a] = std::clamp(x[a], f.grid_min[a] + 1e-6f, f.grid_max[a] - 1e-6f);
        return x;
    }
    V<D> gradient(const V<D>& x) const {
        V<D> g = V<D>::zero();
        for (int a = 0; a < D; a++) {
            float spacing = (f.grid_max[a] - f.grid_min[a]) / (float)f.shape[a];
            V<D> o = V<D>::zero(); o[a] = spacing;
            V<D> xm = clampIn(x - o), xp = clampIn(x + o);
            float denom = xp[a] - xm[a];
            g[a] = denom > 0.0f ? (value(xp) - value(xm)) / denom : 0.0f;
        }
        return g;
    }
    V<D> normal(const V<D>& x) const { V<D> g = gradient(x); return g * (1.0f / (norm(g) + 1e-6f)); }
    void project(V<D>& x, V<D>& n, int maxIter = 8, float eps = 1e-5f) const {
        for (int i = 0; i < maxIter; i++) {
            float s = value(x);
            if (std::fabs(s) < eps) break;
            V<D> g = gradient(x);
            x = clampIn(x - g * (s / (norm(g) + 1e-6f)));
        }
        n = normal(x);
    }
    bool outside(const V<D>& x) const {
        for (int a = 0; a < D; a++) if (x[a] < dmin[a] || x[a] > dmax[a]) return true;
        return false;
    }
    // intersectZeroLevelSet (sdf_grid_geometric_queries.h:147-205): clip the ray to the grid
    // box, sphere-trace with steps |sdf| (at most 128, hit when |sdf| < 1e-5); on a sign change
    // between two samples, one secant step and one gradient step give the hit point
    bool intersect(const V<D>& o, const V<D>& dir, float tMax, V<D>& pt, V<D>& n, float& tHit) const {
        float tEnter = 0.0f, tExit = tMax;
        for (int a = 0; a < D; a++) {
            float inv = 1.0f / dir[a];
            float t0 = (f.grid_min[a] - o[a]) * inv, t1 = (f.grid_max[a] - o[a]) * inv;
            tEnter = std::fmax(tEnter, std::fmin(t0, t1));
            tExit = std::fmin(tExit, std::fmax(t0, t1));
        }
        if (tEnter > tExit) return false;
        float tPrev = tEnter, sPrev = 0.0f, t = tEnter;
        for (int i = 0; i < 128 && t <= tExit; i++) {
            V<D> p = o + dir * t;
            float s = value(p);
            if (std::fabs(s) < 1e-5f) { pt = p; n = normal(p); tHit = t; return true; }
            if (sPrev * s < 0.0f) {
                float th = tPrev + (sPrev / (sPrev - s)) * (t - tPrev);
                V<D> ph = o + dir * th;
                float sh = value(ph);
                V<D> g = gradient(ph);
                ph = ph - g * (sh / (norm(g) + 1e-6f));
                pt = ph; n = normal(ph); tHit = dot(ph - o, dir);
                return true;
            }
            tPrev = t; sPrev = s;
            t += std::fabs(s);
        }
        return false;
    }
};

// walk (walk_on_spheres.h:228-291) over the SDF geometry from a point at distance dist
template <int D>
Code sdfWalk(const SdfGeom<D>& g, const SolveCtx& cx, float dist, GreensBall<D>& gf, Pcg32& rng, WalkState<D>& st,
             std::deque<WalkState<D>>& q) {
    const zg_walk_settings& ws = *cx.ws; const zg_pde& pde = *cx.pde;
    while (dist > ws.epsilon_shell_for_absorbing_boundary) {
        gf.updateBall(st.currentPt, dist);
        if (!ws.ignore_source_contribution) {
            float sc_;
            if (pde.is_source_constant) sc_ = gf.norm() * evalField<D>(pde.source, st.currentPt, V<D>::zero(), false, pde.absorption_coeff);
            else { float r, pdf; V<D> y = gf.sampleVolume(rng, r, pdf); sc_ = gf.norm() * evalField<D>(pde.source, y, V<D>::zero(), false, pde.absorption_coeff); }
            st.totalSourceContribution += st.throughput * sc_;
        }
        V<D> dir = sampleUnitSphere<D>(rng);
        st.currentPt = st.currentPt + dist * dir;
        st.throughput *= gf.directionSampledPoissonKernel(st.currentPt);
        if (g.outside(st.currentPt)) return Escaped;
        if (applyWeightWindow(ws, rng, st, q)) return RR;
        st.walkLength++;
        if (st.walkLength >= ws.max_walk_length) return MaxLen;
        if (pde.absorption_coeff > 0.0f && ws.steps_before_applying_tikhonov == st.walkLength) gf.makeYukawa(pde.absorption_coeff);
        dist = std::fabs(g.value(st.currentPt));
    }
    return Reached;
}

// terminal contribution: the Dirichlet value at the steepest-descent projection onto the zero
// level set, on the side given by the sign of the SDF (getTerminalContribution :293-318)
template <int D>
float sdfTerminal(const SdfGeom<D>& g, const SolveCtx& cx, Code code, const WalkState<D>& st) {
    const zg_walk_settings& ws = *cx.ws; const zg_pde& pde = *cx.pde;
    if (!((code == Reached || code == MaxLen) && !ws.ignore_absorbing_boundary_contribution)) return 0.0f;
    V<D> x = st.currentPt;
    float sd = g.value(x);
    V<D> n; g.project(x, n);
    return evalField<D>(pde.dirichlet, x, n, ws.solve_double_sided && sd > 0.0f, pde.absorption_coeff);
}

// estimateSolution (walk_on_spheres.h:320-414) over the SDF geometry
template <int D>
void wosEstimateSdf(const SdfGeom<D>& g, const SolveCtx& cx, int64_t p, const V<D>& pt, uint8_t type, float distA,
                    int nWalks, Stats& out, int64_t codes[4], int64_t prevN) {
    const zg_walk_settings& ws = *cx.ws; const zg_pde& pde = *cx.pde;
    if (type == ZG_SAMPLE_ON_ABSORBING_BOUNDARY) {
        if (prevN == 0) out.add(ws.ignore_absorbing_boundary_contribution ? 0.0 :
                                (double)evalField<D>(pde.dirichlet, pt, V<D>::zero(), false, pde.absorption_coeff));
        return;
    } else if (distA <= ws.epsilon_shell_for_absorbing_boundary) {
        nWalks = std::min(nWalks, 1);
    }
    Pcg32 pointRng = makeRng(cx, p, 0);
    for (int w = 0; w < nWalks; w++) {
        Pcg32 walkRng = cx.rngMode != 0 ? makeRng(cx, p, w) : Pcg32();
        Pcg32& rng = cx.rngMode != 0 ? walkRng : pointRng;
        // the walk and its split parts (applyWeightWindow queues them); the estimate is their sum
        std::deque<WalkState<D>> q;
        WalkState<D> st0; st0.currentPt = pt;
        q.push_back(st0);
        int splits = -1; float total = 0.0f; bool success = false;
        while (!q.empty()) {
            WalkState<D> st = q.front(); q.pop_front();
            splits++;
            GreensBall<D> gf;
            if (pde.absorption_coeff > 0.0f && ws.steps_before_applying_tikhonov == 0) gf.makeYukawa(pde.absorption_coeff);
            float dist = splits > 0 ? std::fabs(g.value(st.currentPt)) : distA;
            Code code = sdfWalk<D>(g, cx, dist, gf, rng, st, q);
            codes[code]++;
            if (code == Escaped) continue;
            float term = sdfTerminal<D>(g, cx, code, st);
            total += st.throughput * term + st.totalSourceContribution;
            out.len += st.walkLength;
            success = true;
        }
        if (success) { out.add((double)total); out.splits += splits; }
    }
}

// ------------------------------------------------------------------------------------
// boundary value caching: free-space harmonic kernels (distributions.h:84-177), kernel
// regularization (:12-55) and the splat (variance_reduction/boundary_value_caching.h:370-402,
// 683-810)
// ------------------------------------------------------------------------------------
template <int D>
struct FreeSpaceKernels {
    float G, P;
    V<D> dG, dP;
    // pole x, sample y with normal n, r = clamped distance; mu = sqrt(absorption) > 0 selects the
    // 3D Yukawa kernels (YukawaGreensFnFreeSpace<3>, distributions.h:271-321):
    //   G = e^{-mu r} / (4 pi r),  grad_x G = -(x - y) e^{-mu r} (mu r + 1) / (4 pi r^3),
    //   P = n.(x - y) e^{-mu r} (mu r + 1) / (4 pi r^3),
    //   grad_x P = [n e^{-mu r} (mu r + 1) - (x - y) (n.(x - y) / r^2) e^{-mu r} (mu^2 r^2 + 3 mu r + 3)]
    //              / (4 pi r^3),
    // the last one the derivative of P; the reference's :307-317 writes (2 mu r + 3) for the
    // factor (mu^2 r^2 + 3 mu r + 3), which is not the derivative of its own P (both agree at mu = 0)
    void eval(const V<D>& x, const V<D>& y, const V<D>& n, float r, float mu = 0.0f) {
        V<D> xy = x - y;
        float r2 = r * r;
        if (D == 2) {
            G = -std::log(r) / (2.0f * PI_F);
            dG = xy * (-1.0f / (2.0f * PI_F * r2));
            P = dot(n, xy) / (2.0f * PI_F * r2);
            dP = (n - xy * (2.0f * (dot(n, xy) / r2))) * (1.0f / (2.0f * PI_F * r2));
        } else {
            float r3 = r2 * r;
            float e = std::exp(-mu * r), mr = mu * r;
            float q1 = e * (mr + 1.0f);
            G = e / (4.0f * PI_F * r);
            dG = xy * (-q1 / (4.0f * PI_F * r3));
            P = dot(n, xy) * q1 / (4.0f * PI_F * r3);
            float q2 = e * (mr * mr + 3.0f * mr + 3.0f);
            dP = (n * q1 - xy * ((dot(n, xy) / r2) * q2)) * (1.0f / (4.0f * PI_F * r3));
        }
    }
};
template <int D> float regularizationG(float r) { return D == 2 ? 1.0f : std::erf(r); }
template <int D> float regularizationP(float r) {
    if (D == 2) return 1.0f - std::exp(-r * r);
    return std::erf(r) - 2.0f * r * std::exp(-r * r) / std::sqrt((float)M_PI);
}

struct BvcAcc {   // Welford per category
    int64_t n = 0; double mean = 0.0, m2 = 0.0; double gm[3] = {0, 0, 0}, g2[3] = {0, 0, 0};
    void add(double x, const double* g, int D) {
        n++;
        double d = x - mean; mean += d / (double)n; m2 += d * (x - mean);
        for (int i = 0; i < D; i++) { double e = g[i] - gm[i]; gm[i] += e / (double)n; g2[i] += e * (g[i] - gm[i]); }
    }
};

template <int D>
void bvcSplatPoint(const zg_bvc_samples& S, const V<D>& x, bool evalOnBoundary, float radiusClamp, float kreg,
                   float robinCutoff, float mu, BvcAcc acc[5]) {
    const float alpha = evalOnBoundary ? 2.0f : 1.0f;
    const float alphaG = evalOnBoundary ? 0.0f : alpha;
    for (int64_t j = 0; j < S.n; j++) {
        V<D> y = load<D>(S.pt + j * D);
        uint8_t ty = S.type[j];
        float pdf = S.pdf[j];
        bool aligned = S.normal_aligned && S.normal_aligned[j];
        float r = std::max(radiusClamp, norm(y - x));
        FreeSpaceKernels<D> k;
        V<D> n = V<D>::zero();
        if (ty != ZG_SAMPLE_IN_DOMAIN) n = load<D>(S.normal + j * D) * (aligned ? -1.0f : 1.0f);
        k.eval(x, y, n, r, mu);
        double est, grad[3];
        if (ty == ZG_SAMPLE_IN_DOMAIN) {
            float dGn = norm(k.dG);
            if (std::isinf(k.G) || std::isnan(k.G) || std::isinf(dGn) || std::isnan(dGn)) continue;
            float G = k.G;
            if (kreg > 0.0f) G *= regularizationG<D>(r / kreg);
            float src = S.contribution[j];
            est = alpha * G * src / pdf;
            for (int i = 0; i < D; i++) grad[i] = alphaG * k.dG[i] * src / pdf;
            acc[4].add(est, grad, D);
            continue;
        }
        float dGn = norm(k.dG), dPn = norm(k.dP);
        if (std::isinf(k.G) || std::isinf(k.P) || std::isinf(dGn) || std::isinf(dPn) ||
            std::isnan(k.G) || std::isnan(k.P) || std::isnan(dGn) || std::isnan(dPn)) continue;
        float G = k.G, P = k.P;
        if (kreg > 0.0f) { float rr = r / kreg; G *= regularizationG<D>(rr); P *= regularizationP<D>(rr); }
        float u = S.solution[j], dn = S.normal_derivative[j];
        float h = S.contribution ? S.contribution[j] : 0.0f;
        float kappa = S.robin_coeff ? S.robin_coeff[j] : 0.0f;
        if (std::fabs(kappa) > robinCutoff) {
            est = alpha * ((G + P / kappa) * dn - P * h / kappa) / pdf;
            for (int i = 0; i < D; i++) grad[i] = alphaG * ((k.dG[i] + k.dP[i] / kappa) * dn - k.dP[i] * h / kappa) / pdf;
        } else if (std::fabs(kappa) > 0.0f) {
            est = alpha * (G * h - (P + kappa * G) * u) / pdf;
            for (int i = 0; i < D; i++) grad[i] = alphaG * (k.dG[i] * h - (k.dP[i] + kappa * k.dG[i]) * u) / pdf;
        } else {
            est = alpha * (G * dn - P * u) / pdf;
            for (int i = 0; i < D; i++) grad[i] = alphaG * (k.dG[i] * dn - k.dP[i] * u) / pdf;
        }
        int cat = (ty == ZG_SAMPLE_ON_ABSORBING_BOUNDARY ? 0 : 2) + (aligned ? 1 : 0);
        acc[cat].add(est, grad, D);
    }
}

}  // namespace orc

using namespace orc;

template <int D>
static void closestImpl(Scene<D>& s, int boundary, int64_t n, const float* pts, float* dist, float* cp, float* nrm,
                        int32_t* prim, int brute, int64_t* visits, int nThreads) {
    Mesh<D>& m = boundary ? s.refl : s.abs;
    std::atomic<int64_t> nodes{0}, prims{0};
    parallelFor(n, nThreads, [&](int64_t i, int) {
        V<D> x = load<D>(pts + i * D), q; float d; int p; QueryCount qc;
        bool ok = brute ? closestPointBrute(m, x, d, q, p) : closestPoint(m, x, d, q, p, &qc);
        nodes += qc.nodes; prims += qc.prims;
        dist[i] = ok ? d : FLT_MAX;
        if (cp) store<D>(cp + i * D, ok ? q : V<D>::zero());
        if (nrm) store<D>(nrm + i * D, ok ? m.primNormal(p) : V<D>::zero());
        if (prim) prim[i] = ok ? m.origId[p] : -1;
    });
    if (visits) { visits[0] = nodes.load(); visits[1] = prims.load(); }
}

template <int D>
static void rayImpl(Scene<D>& s, int boundary, int64_t n, const float* o, const float* d, const float* tmax,
                    uint8_t* hit, float* t, int32_t* prim, int brute) {
    Mesh<D>& m = boundary ? s.refl : s.abs;
    for (int64_t i = 0; i < n; i++) {
        float th; int p;
        bool h_ = intersectRay(m, load<D>(o + i * D), load<D>(d + i * D), tmax[i], th, p, brute != 0);
        hit[i] = h_; t[i] = h_ ? th : tmax[i]; if (prim) prim[i] = h_ ? m.origId[p] : -1;
    }
}

template <int D>
static void insideImpl(Scene<D>& s, int64_t n, const float* pts, uint8_t* inside, int nThreads) {
    parallelFor(n, nThreads, [&](int64_t i, int) {
        V<D> x = load<D>(pts + i * D);
        bool in = true;
        for (int a = 0; a < D; a++) {
            V<D> dir = V<D>::zero(); dir[a] = 1.0f;
            int hits = countHits(s.abs, x, dir) + countHits(s.refl, x, dir);
            in = in && (hits % 2 == 1);
        }
        inside[i] = in;
    });
}

template <int D>
static int solveImpl(bool wost, Scene<D>& sc, const zg_pde* pde, const zg_walk_settings* ws, int64_t n,
                     const float* pts, const float* normals, const uint8_t* types, const uint8_t* aligned,
                     const float* distA, const int32_t* nWalks, uint64_t seed, uint32_t streamOffset, int rngMode,
                     int nThreads, int64_t* sN, double* sMean, double* sM2, int64_t* sLen, int64_t* sSplits,
                     float* firstSphereRadius, int64_t* report, const uint8_t* quantity = nullptr,
                     const float* dirDeriv = nullptr, int cvMode = 0, int64_t* gN = nullptr, double* gMean = nullptr,
                     double* gM2 = nullptr, double* gFirst = nullptr, double* gDeriv = nullptr,
                     const float* distR = nullptr) {
    if (pde->absorption_coeff > 0.0f && D == 2) return ZG_ERR_UNSUPPORTED;
    bool anyGrad = false;
    if (quantity) for (int64_t i = 0; i < n; i++) anyGrad = anyGrad || quantity[i] == ZG_ESTIMATE_SOLUTION_AND_GRADIENT;
    if (anyGrad && (!gN || !gMean || !gM2 || !gFirst || !gDeriv)) return ZG_ERR_INVALID_ARGUMENT;
    SolveCtx cx{pde, ws, seed, streamOffset, rngMode};
    int nt = std::max(1, nThreads);
    std::vector<std::array<int64_t, 8>> per(nt);
    for (auto& a : per) a.fill(0);
    parallelFor(n, nt, [&](int64_t i, int t) {
        V<D> x = load<D>(pts + i * D);
        V<D> nr = normals ? load<D>(normals + i * D) : V<D>::zero();
        uint8_t ty = types ? types[i] : 0;
        Stats st; int64_t codes[4] = {0, 0, 0, 0}; QueryCount qc;
        uint8_t quant = quantity ? quantity[i] : (uint8_t)ZG_ESTIMATE_SOLUTION;
        if (quant == ZG_ESTIMATE_SKIP) return;
        if (quant == ZG_ESTIMATE_SOLUTION_AND_GRADIENT) {
            V<D> dd = V<D>::zero(); dd[0] = 1.0f;
            if (dirDeriv) dd = load<D>(dirDeriv + i * D);
            GradCv cv{cvMode, sN[i], sMean[i], gFirst[i]};
            GradStats<D> g;
            float dR = FLT_MAX;
            if (wost) {
                if (distR) dR = distR[i];
                else if (sc.hasRefl()) { V<D> cp; int pr; closestPoint(sc.refl, x, dR, cp, pr, &qc); }
            }
            wosEstimateGrad<D>(sc, cx, cv, i, x, distA[i], nWalks[i], dd, st, g, codes, &qc,
                               firstSphereRadius ? firstSphereRadius + i : nullptr, wost, dR);
            mergeInto(sN[i], sMean[i], sM2[i], st);
            if (g.n > 0) {
                int64_t N1 = gN[i], N2 = g.n, N = N1 + N2;
                for (int k = 0; k < D; k++) {
                    double& m = gMean[i * D + k]; double& m2 = gM2[i * D + k];
                    if (N1 == 0) { m = g.mean[k]; m2 = g.m2[k]; }
                    else {
                        double d = g.mean[k] - m;
                        m += d * ((double)N2 / (double)N);
                        m2 += g.m2[k] + d * d * ((double)N1 * (double)N2 / (double)N);
                    }
                }
                gN[i] = N;
            }
            gFirst[i] += g.firstSource; gDeriv[i] += g.derivative;
            sLen[i] += st.len; sSplits[i] += st.splits;
            for (int k = 0; k < 4; k++) per[t][k] += codes[k];
            per[t][4] += st.len; per[t][5] += qc.nodes; per[t][6] += qc.prims;
            return;
        }
        if (wost) wostEstimate<D>(sc, cx, i, x, nr, ty, aligned ? aligned[i] != 0 : false, distA[i], nWalks[i], st, codes, &qc, sN[i], firstSphereRadius ? firstSphereRadius + i : nullptr);
        else wosEstimate<D>(sc, cx, i, x, ty, distA[i], nWalks[i], st, codes, &qc, sN[i], firstSphereRadius ? firstSphereRadius + i : nullptr);
        mergeInto(sN[i], sMean[i], sM2[i], st);
        sLen[i] += st.len; sSplits[i] += st.splits;
        for (int k = 0; k < 4; k++) per[t][k] += codes[k];
        per[t][4] += st.len; per[t][5] += qc.nodes; per[t][6] += qc.prims;
    });
    if (report) { for (int k = 0; k < 7; k++) { report[k] = 0; for (auto& a : per) report[k] += a[k]; } }
    return ZG_OK;
}


extern "C" {

void orc_pcg32_u32(uint64_t initstate, uint64_t initseq, int64_t n, uint32_t* out) {
    Pcg32 r(initstate, initseq);
    for (int64_t i = 0; i < n; i++) out[i] = r.nextUInt();
}
uint64_t orc_splitmix64_mix(uint64_t z) { return splitmix64Mix(z); }
uint64_t orc_walk_init_state(uint64_t seed, uint32_t p, uint32_t w) { return walkInitState(seed, p, w); }
// the first n_draws nextFloat() values of the generators of walks (p[i], w[i]), row i of out
void orc_walk_floats(uint64_t seed, const uint32_t* p, const uint32_t* w, int64_t n, int32_t nDraws, float* out) {
    for (int64_t i = 0; i < n; i++) {
        Pcg32 r = walkRng(seed, p[i], w[i]);
        for (int32_t k = 0; k < nDraws; k++) out[i * nDraws + k] = r.nextFloat();
    }
}
void orc_pcg32_f32(uint64_t initstate, uint64_t initseq, int64_t n, float* out) {
    Pcg32 r(initstate, initseq);
    for (int64_t i = 0; i < n; i++) out[i] = r.nextFloat();
}
void orc_pcg32_default_u32(int64_t n, uint32_t* out) {
    Pcg32 r;
    for (int64_t i = 0; i < n; i++) out[i] = r.nextUInt();
}

// sphere / ball samplers for moment and closed-form tests
void orc_sample_unit_sphere(int dim, uint64_t seed, int64_t n, float* out) {
    Pcg32 r(seed, 7);
    for (int64_t i = 0; i < n; i++) {
        if (dim == 2) store<2>(out + 2 * i, sampleUnitSphere2(r));
        else store<3>(out + 3 * i, sampleUnitSphere3(r));
    }
}
// radius samples of the harmonic / Yukawa ball source sampler (sampleVolume); lambda 0 = harmonic
void orc_sample_ball_radius(int dim, float R, float lambda, uint64_t seed, int64_t n, float* r_out, float* pdf_out) {
    Pcg32 rng(seed, 11);
    for (int64_t i = 0; i < n; i++) {
        float r = 0, pdf = 0;
        if (dim == 2) { GreensBall<2> g; if (lambda > 0) g.makeYukawa(lambda); g.updateBall(V<2>::zero(), R); g.sampleVolume(rng, r, pdf); }
        else { GreensBall<3> g; if (lambda > 0) g.makeYukawa(lambda); g.updateBall(V<3>::zero(), R); g.sampleVolume(rng, r, pdf); }
        r_out[i] = r; if (pdf_out) pdf_out[i] = pdf;
    }
}
// Green's function quantities of a ball of radius R: out = [norm, poissonKernel, G(r), dampening(R)]
void orc_greens_ball(int dim, float R, float lambda, float r, float* out) {
    if (dim == 2) {
        GreensBall<2> g; if (lambda > 0) g.makeYukawa(lambda); g.updateBall(V<2>::zero(), R);
        out[0] = g.norm(); out[1] = g.poissonKernel(); out[2] = g.evaluate(r); out[3] = lambda > 0 ? g.poissonKernelDampeningFactor(R) : 1.0f;
    } else {
        GreensBall<3> g; if (lambda > 0) g.makeYukawa(lambda); g.updateBall(V<3>::zero(), R);
        out[0] = g.norm(); out[1] = g.poissonKernel(); out[2] = g.evaluate(r); out[3] = lambda > 0 ? g.poissonKernelDampeningFactor(R) : 1.0f;
    }
}
void orc_offset_point(int dim, const float* p, const float* n, float* out) {
    if (dim == 2) store<2>(out, offsetPoint<2>(load<2>(p), load<2>(n)));
    else store<3>(out, offsetPoint<3>(load<3>(p), load<3>(n)));
}
float orc_eval_function(int dim, const zg_function* f, const float* x, const float* n, int aligned, float absorption) {
    if (dim == 2) return evalField<2>(*f, load<2>(x), n ? load<2>(n) : V<2>::zero(), aligned != 0, absorption);
    return evalField<3>(*f, load<3>(x), n ? load<3>(n) : V<3>::zero(), aligned != 0, absorption);
}

void* orc_scene_create(int dim, const float* V_, int64_t nV, const int32_t* I, int64_t nP, const uint8_t* refl,
                       const float* dmin, const float* dmax, int leafSize, int nBuckets, int dbl, float prec) {
    SceneHandle* h = new SceneHandle();
    h->dim = dim;
    if (dim == 2) { h->s2.reset(new Scene<2>()); buildScene<2>(*h->s2, V_, nV, I, nP, refl, dmin, dmax, leafSize, nBuckets, dbl, prec); }
    else { h->s3.reset(new Scene<3>()); buildScene<3>(*h->s3, V_, nV, I, nP, refl, dmin, dmax, leafSize, nBuckets, dbl, prec); }
    return h;
}
void orc_scene_destroy(void* h) { delete (SceneHandle*)h; }
// walk on spheres over an SDF grid (values with the DenseGrid layout, trilinear); rng_mode as orc_solve
int orc_sdf_solve(int dim, const float* grid, const int32_t* shape, const float* gmin, const float* gmax,
                  const float* dmin, const float* dmax, const zg_pde* pde, const zg_walk_settings* ws, int64_t n,
                  const float* pts, const uint8_t* types, const float* distA, const int32_t* nWalks, uint64_t seed,
                  uint32_t streamOffset, int rngMode, int nThreads, int64_t* sN, double* sMean, double* sM2,
                  int64_t* sLen, int64_t* report) {
    zg_function f{}; f.kind = ZG_FN_DENSE_GRID; f.grid = grid; f.interpolate = 1;
    for (int a = 0; a < 3; a++) { f.shape[a] = a < dim ? shape[a] : 1; f.grid_min[a] = a < dim ? gmin[a] : 0; f.grid_max[a] = a < dim ? gmax[a] : 1; }
    SolveCtx cx{pde, ws, seed, streamOffset, rngMode};
    int nt = std::max(1, nThreads);
    std::vector<std::array<int64_t, 5>> per(nt);
    for (auto& a : per) a.fill(0);
    auto run = [&](auto dimTag) {
        constexpr int D = decltype(dimTag)::value;
        SdfGeom<D> g; g.f = f; g.dmin = load<D>(dmin); g.dmax = load<D>(dmax);
        parallelFor(n, nt, [&](int64_t i, int t) {
            Stats st; int64_t codes[4] = {0, 0, 0, 0};
            wosEstimateSdf<D>(g, cx, i, load<D>(pts + i * D), types ? types[i] : 0, distA[i], nWalks[i], st, codes, sN[i]);
            mergeInto(sN[i], sMean[i], sM2[i], st);
            sLen[i] += st.len;
            for (int k = 0; k < 4; k++) per[t][k] += codes[k];
            per[t][4] += st.len;
        });
    };
    if (dim == 2) run(std::integral_constant<int, 2>{}); else run(std::integral_constant<int, 3>{});
    if (report) { for (int k = 0; k < 5; k++) { report[k] = 0; for (auto& a : per) report[k] += a[k]; } }
    return 0;
}
// gradient solves over the SDF geometry: every point runs estimateSolutionAndGradient
// (walk_on_spheres.h:416-593) with its walks continued over the SDF (cvMode as orc_solve_grad)
int orc_sdf_solve_grad(int dim, const float* grid, const int32_t* shape, const float* gmin, const float* gmax,
                       const float* dmin, const float* dmax, const zg_pde* pde, const zg_walk_settings* ws, int64_t n,
                       const float* pts, const float* distA, const int32_t* nWalks, const float* dirDeriv,
                       uint64_t seed, uint32_t streamOffset, int rngMode, int cvMode, int nThreads, int64_t* sN,
                       double* sMean, double* sM2, int64_t* sLen, int64_t* gN, double* gMean, double* gM2,
                       double* gFirst, double* gDeriv) {
    zg_function f{}; f.kind = ZG_FN_DENSE_GRID; f.grid = grid; f.interpolate = 1;
    for (int a = 0; a < 3; a++) { f.shape[a] = a < dim ? shape[a] : 1; f.grid_min[a] = a < dim ? gmin[a] : 0; f.grid_max[a] = a < dim ? gmax[a] : 1; }
    SolveCtx cx{pde, ws, seed, streamOffset, rngMode};
    int nt = std::max(1, nThreads);
    auto run = [&](auto dimTag) {
        constexpr int D = decltype(dimTag)::value;
        SdfGeom<D> g; g.f = f; g.dmin = load<D>(dmin); g.dmax = load<D>(dmax);
        Scene<D> none;   // the walks query the SDF, not a mesh
        parallelFor(n, nt, [&](int64_t i, int) {
            V<D> dd = V<D>::zero(); dd[0] = 1.0f;
            if (dirDeriv) dd = load<D>(dirDeriv + i * D);
            GradCv cv{cvMode, sN[i], sMean[i], gFirst[i]};
            Stats st; GradStats<D> gs; int64_t codes[4] = {0, 0, 0, 0}; QueryCount qc;
            wosEstimateGrad<D>(none, cx, cv, i, load<D>(pts + i * D), distA[i], nWalks[i], dd, st, gs, codes, &qc,
                               nullptr, false, FLT_MAX, &g);
            mergeInto(sN[i], sMean[i], sM2[i], st);
            if (gs.n > 0) {
                int64_t N1 = gN[i], N2 = gs.n, N = N1 + N2;
                for (int k = 0; k < D; k++) {
                    double& m = gMean[i * D + k]; double& m2 = gM2[i * D + k];
                    if (N1 == 0) { m = gs.mean[k]; m2 = gs.m2[k]; }
                    else {
                        double dlt = gs.mean[k] - m;
                        m += dlt * ((double)N2 / (double)N);
                        m2 += gs.m2[k] + dlt * dlt * ((double)N1 * (double)N2 / (double)N);
                    }
                }
                gN[i] = N;
            }
            gFirst[i] += gs.firstSource; gDeriv[i] += gs.derivative;
            sLen[i] += st.len;
        });
    };
    if (dim == 2) run(std::integral_constant<int, 2>{}); else run(std::integral_constant<int, 3>{});
    return 0;
}
// SDF queries: |sdf|, projected point and normal (projectToAbsorbingBoundary for an SdfGrid)
void orc_sdf_query(int dim, const float* grid, const int32_t* shape, const float* gmin, const float* gmax, int64_t n,
                   const float* pts, float* dist, float* proj, float* nrm) {
    zg_function f{}; f.kind = ZG_FN_DENSE_GRID; f.grid = grid; f.interpolate = 1;
    for (int a = 0; a < 3; a++) { f.shape[a] = a < dim ? shape[a] : 1; f.grid_min[a] = a < dim ? gmin[a] : 0; f.grid_max[a] = a < dim ? gmax[a] : 1; }
    for (int64_t i = 0; i < n; i++) {
        if (dim == 2) {
            SdfGeom<2> g; g.f = f; V<2> x = load<2>(pts + 2 * i), nn; dist[i] = std::fabs(g.value(x));
            g.project(x, nn); store<2>(proj + 2 * i, x); store<2>(nrm + 2 * i, nn);
        } else {
            SdfGeom<3> g; g.f = f; V<3> x = load<3>(pts + 3 * i), nn; dist[i] = std::fabs(g.value(x));
            g.project(x, nn); store<3>(proj + 3 * i, x); store<3>(nrm + 3 * i, nn);
        }
    }
}

// sphere tracing of n rays against the zero level set: hit, t, point, normal per ray
void orc_sdf_ray(int dim, const float* grid, const int32_t* shape, const float* gmin, const float* gmax, int64_t n,
                 const float* o, const float* d, const float* tmax, uint8_t* hit, float* t, float* pt, float* nrm) {
    zg_function f{}; f.kind = ZG_FN_DENSE_GRID; f.grid = grid; f.interpolate = 1;
    for (int a = 0; a < 3; a++) { f.shape[a] = a < dim ? shape[a] : 1; f.grid_min[a] = a < dim ? gmin[a] : 0; f.grid_max[a] = a < dim ? gmax[a] : 1; }
    for (int64_t i = 0; i < n; i++) {
        if (dim == 2) {
            SdfGeom<2> g; g.f = f; V<2> p = V<2>::zero(), nn = V<2>::zero(); float th = 0.0f;
            hit[i] = g.intersect(load<2>(o + 2 * i), load<2>(d + 2 * i), tmax[i], p, nn, th);
            t[i] = th; store<2>(pt + 2 * i, p); store<2>(nrm + 2 * i, nn);
        } else {
            SdfGeom<3> g; g.f = f; V<3> p = V<3>::zero(), nn = V<3>::zero(); float th = 0.0f;
            hit[i] = g.intersect(load<3>(o + 3 * i), load<3>(d + 3 * i), tmax[i], p, nn, th);
            t[i] = th; store<3>(pt + 3 * i, p); store<3>(nrm + 3 * i, nn);
        }
    }
}

// BoundaryValueCaching::splat of a sample cache into evaluation points (host arrays); stats are
// Chan-merged like zg_bvc_statistics
int orc_bvc_splat(int dim, float absorption, const zg_bvc_samples* S, const zg_bvc_eval_points* E,
                  float radiusClamp, float kreg, float robinCutoff, float cutoffA, float cutoffR, int64_t* sn,
                  double* smean, double* sm2, double* sgm, double* sg2, int nThreads) {
    if (absorption < 0.0f || (absorption > 0.0f && dim != 3)) return 2;   // screened kernels in 3D only
    const float mu = std::sqrt(absorption);
    parallelFor(E->n, std::max(1, nThreads), [&](int64_t i, int) {
        float dA = E->dist_to_absorbing_boundary[i];
        float dR = E->dist_to_reflecting_boundary ? E->dist_to_reflecting_boundary[i] : FLT_MAX;
        if (dA < cutoffA || dR < cutoffR) return;
        uint8_t ty = E->type ? E->type[i] : (uint8_t)ZG_SAMPLE_IN_DOMAIN;
        bool onB = ty == ZG_SAMPLE_ON_ABSORBING_BOUNDARY || ty == ZG_SAMPLE_ON_REFLECTING_BOUNDARY;
        BvcAcc acc[5];
        if (dim == 2) bvcSplatPoint<2>(*S, load<2>(E->pt + 2 * i), onB, radiusClamp, kreg, robinCutoff, 0.0f, acc);
        else bvcSplatPoint<3>(*S, load<3>(E->pt + 3 * i), onB, radiusClamp, kreg, robinCutoff, mu, acc);
        for (int c = 0; c < 5; c++) {
            const BvcAcc& a = acc[c];
            if (a.n == 0) continue;
            size_t k = (size_t)i * 5 + c;
            int64_t N1 = sn[k], N = N1 + a.n;
            double f2 = (double)a.n / (double)N, f12 = (double)N1 * (double)a.n / (double)N;
            if (N1 == 0) { smean[k] = a.mean; sm2[k] = a.m2; }
            else { double d = a.mean - smean[k]; smean[k] += d * f2; sm2[k] += a.m2 + d * d * f12; }
            for (int g = 0; g < dim; g++) {
                size_t kg = k * dim + g;
                if (N1 == 0) { sgm[kg] = a.gm[g]; sg2[kg] = a.g2[g]; }
                else { double d = a.gm[g] - sgm[kg]; sgm[kg] += d * f2; sg2[kg] += a.g2[g] + d * d * f12; }
            }
            sn[k] = N;
        }
    });
    return 0;
}
// the 3D Yukawa ball of radius R at distance r from its centre, for pins: [norm, dampening(r),
// gradientNorm(r), poissonKernelGradient scale, evaluate(r), poissonKernel]
void orc_yukawa_ball(float R, float lambda, float r, float* out) {
    GreensBall<3> g;
    g.makeYukawa(lambda);
    g.updateBall(V<3>::zero(), R);
    V<3> e = V<3>::zero();
    e[0] = 1.0f;
    out[0] = g.norm();
    out[1] = g.poissonKernelDampeningFactor(r);
    out[2] = g.gradientNorm(r);
    out[3] = g.poissonKernelGradient(e)[0];
    out[4] = g.evaluate(r);
    out[5] = g.poissonKernel();
}

// free-space kernels at (x, y, n, r) for pins: out = [G, P, dG..., dP...], regularized G and P in
// out[2 + 2 dim], out[3 + 2 dim] (kreg > 0)
void orc_free_space_kernels(int dim, const float* x, const float* y, const float* n, float r, float kreg,
                            float absorption, float* out) {
    if (dim == 2) {
        FreeSpaceKernels<2> k; k.eval(load<2>(x), load<2>(y), load<2>(n), r);
        out[0] = k.G; out[1] = k.P; store<2>(out + 2, k.dG); store<2>(out + 4, k.dP);
        out[6] = kreg > 0 ? k.G * regularizationG<2>(r / kreg) : k.G; out[7] = kreg > 0 ? k.P * regularizationP<2>(r / kreg) : k.P;
    } else {
        FreeSpaceKernels<3> k; k.eval(load<3>(x), load<3>(y), load<3>(n), r, std::sqrt(std::max(absorption, 0.0f)));
        out[0] = k.G; out[1] = k.P; store<3>(out + 2, k.dG); store<3>(out + 5, k.dP);
        out[8] = kreg > 0 ? k.G * regularizationG<3>(r / kreg) : k.G; out[9] = kreg > 0 ? k.P * regularizationP<3>(r / kreg) : k.P;
    }
}

// per-primitive Robin coefficient bounds over the caller's primitives (reflecting ones are read);
// NULL clears them (pure Neumann)
extern "C++" {
template <int D>
static void setRobin(Mesh<D>& m, const float* minC, const float* maxC) {
    m.robinMin.clear(); m.robinMax.clear();
    if (!minC || !maxC) return;
    for (int p = 0; p < m.nPrims(); p++) { m.robinMin.push_back(minC[m.origId[p]]); m.robinMax.push_back(maxC[m.origId[p]]); }
}
}
void orc_scene_set_robin(void* hv, const float* minC, const float* maxC) {
    SceneHandle* h = (SceneHandle*)hv;
    if (h->dim == 2) setRobin<2>(h->s2->refl, minC, maxC);
    else setRobin<3>(h->s3->refl, minC, maxC);
}
// Robin radius bound sqrt(min(maxR^2, Robin bounds)) at n points (brute: every segment)
void orc_robin_radius(void* hv, int64_t n, const float* pts, const float* maxR, int brute, float* out) {
    SceneHandle* h = (SceneHandle*)hv;
    for (int64_t i = 0; i < n; i++) {
        float r2 = maxR[i] < FLT_MAX ? maxR[i] * maxR[i] : FLT_MAX;
        out[i] = h->dim == 2 ? std::sqrt(robinRadius2<2>(h->s2->refl, load<2>(pts + 2 * i), r2, brute != 0))
                             : std::sqrt(robinRadius2<3>(h->s3->refl, load<3>(pts + 3 * i), r2, brute != 0));
    }
}
// out: [absNodes, absDepth, reflNodes, reflDepth, nSilhouetteCandidates]
void orc_scene_stats(void* hv, int64_t* out) {
    SceneHandle* h = (SceneHandle*)hv;
    if (h->dim == 2) { auto& s = *h->s2; out[0] = s.abs.bvh.nodes.size(); out[1] = s.abs.bvh.depth; out[2] = s.refl.bvh.nodes.size(); out[3] = s.refl.bvh.depth; out[4] = s.refl.sil.size(); }
    else { auto& s = *h->s3; out[0] = s.abs.bvh.nodes.size(); out[1] = s.abs.bvh.depth; out[2] = s.refl.bvh.nodes.size(); out[3] = s.refl.bvh.depth; out[4] = s.refl.silE.size(); }
}

void orc_closest_point(void* hv, int boundary, int64_t n, const float* pts, float* dist, float* cp, float* nrm,
                       int32_t* prim, int brute, int64_t* visits, int nThreads) {
    SceneHandle* h = (SceneHandle*)hv;
    if (h->dim == 2) closestImpl<2>(*h->s2, boundary, n, pts, dist, cp, nrm, prim, brute, visits, nThreads);
    else closestImpl<3>(*h->s3, boundary, n, pts, dist, cp, nrm, prim, brute, visits, nThreads);
}

void orc_intersect_ray(void* hv, int boundary, int64_t n, const float* o, const float* d, const float* tmax,
                       uint8_t* hit, float* t, int32_t* prim, int brute) {
    SceneHandle* h = (SceneHandle*)hv;
    if (h->dim == 2) rayImpl<2>(*h->s2, boundary, n, o, d, tmax, hit, t, prim, brute);
    else rayImpl<3>(*h->s3, boundary, n, o, d, tmax, hit, t, prim, brute);
}

void orc_inside_domain(void* hv, int64_t n, const float* pts, uint8_t* inside, int nThreads) {
    SceneHandle* h = (SceneHandle*)hv;
    if (h->dim == 2) insideImpl<2>(*h->s2, n, pts, inside, nThreads);
    else insideImpl<3>(*h->s3, n, pts, inside, nThreads);
}

void orc_star_radius(void* hv, int64_t n, const float* pts, const float* minR, const float* maxR, float prec,
                     const uint8_t* flip, float* out) {
    SceneHandle* h = (SceneHandle*)hv;
    for (int64_t i = 0; i < n; i++) {
        bool fl = flip && flip[i] != 0;
        if (h->dim == 2)
            out[i] = h->s2->hasRefl() ? starRadius2(h->s2->refl, load<2>(pts + 2 * i), minR[i], maxR[i], prec, fl) : maxR[i];
        else
            out[i] = h->s3->hasRefl() ? starRadius3(h->s3->refl, load<3>(pts + 3 * i), minR[i], maxR[i], prec, fl) : maxR[i];
    }
}

// silhouette test of every candidate seen from x; out[k] = 1 when candidate k is a silhouette.
// verts_out: the vertex (2D, 2 floats) or the edge's two end points (3D, 6 floats) per candidate
int64_t orc_silhouette_candidates(void* hv, const float* x, float prec, int flip, float* verts_out, uint8_t* is_sil_out) {
    SceneHandle* h = (SceneHandle*)hv;
    if (h->dim == 3) {
        const auto& m = h->s3->refl;
        for (size_t k = 0; k < m.silE.size(); k++) {
            const auto& e = m.silE[k];
            if (verts_out) { store<3>(verts_out + 6 * k, m.verts[e.va]); store<3>(verts_out + 6 * k + 3, m.verts[e.vb]); }
            if (is_sil_out) {
                V<3> cp;
                closestPointSegmentD<3>(m.verts[e.va], m.verts[e.vb], load<3>(x), cp);
                is_sil_out[k] = m.isSilhouetteEdge(e, load<3>(x), cp, prec, flip != 0);
            }
        }
        return (int64_t)m.silE.size();
    }
    const auto& m = h->s2->refl;
    for (size_t k = 0; k < m.sil.size(); k++) {
        if (verts_out) store<2>(verts_out + 2 * k, m.verts[m.sil[k].v]);
        if (is_sil_out) is_sil_out[k] = m.isSilhouette(m.sil[k], load<2>(x), prec, flip != 0);
    }
    return (int64_t)m.sil.size();
}

extern "C++" {
template <int D>
static int sampleImpl(const Mesh<D>& m, bool bvh, const float* x, float radius, const float* u, float* pt, float* nrm,
                      float* pdf) {
    V<D> p, nn; float pd;
    bool ok = bvh ? sampleReflectingBvh<D>(m, load<D>(x), radius, u, p, nn, pd)
                  : sampleReflecting<D>(m, load<D>(x), radius, u, p, nn, pd);
    if (ok) { store<D>(pt, p); store<D>(nrm, nn); *pdf = pd; }
    return ok;
}
}

// u holds DIM uniform numbers: one for the primitive choice, DIM - 1 for the point on it
int orc_sample_reflecting(void* hv, const float* x, float radius, const float* u, float* pt, float* nrm, float* pdf) {
    SceneHandle* h = (SceneHandle*)hv;
    return h->dim == 2 ? sampleImpl<2>(h->s2->refl, false, x, radius, u, pt, nrm, pdf)
                       : sampleImpl<3>(h->s3->refl, false, x, radius, u, pt, nrm, pdf);
}

int orc_line_of_sight(void* hv, const float* p1, const float* p2) {
    SceneHandle* h = (SceneHandle*)hv;
    if (h->dim == 2) return hasLineOfSight<2>(h->s2->refl, load<2>(p1), load<2>(p2));
    return hasLineOfSight<3>(h->s3->refl, load<3>(p1), load<3>(p2));
}

// BVH-accelerated forms used by the estimator (equal to the brute-force forms above for the
// star radius and line of sight; a different but exact density for the sampler)
void orc_star_radius_bvh(void* hv, int64_t n, const float* pts, const float* minR, const float* maxR, float prec,
                         const uint8_t* flip, float* out) {
    SceneHandle* h = (SceneHandle*)hv;
    for (int64_t i = 0; i < n; i++) {
        bool fl = flip && flip[i] != 0;
        if (h->dim == 2)
            out[i] = h->s2->hasRefl() ? starRadius2Bvh(h->s2->refl, load<2>(pts + 2 * i), minR[i], maxR[i], prec, fl) : maxR[i];
        else
            out[i] = h->s3->hasRefl() ? starRadius3Bvh(h->s3->refl, load<3>(pts + 3 * i), minR[i], maxR[i], prec, fl) : maxR[i];
    }
}

int orc_sample_reflecting_bvh(void* hv, const float* x, float radius, const float* u, float* pt, float* nrm, float* pdf) {
    SceneHandle* h = (SceneHandle*)hv;
    return h->dim == 2 ? sampleImpl<2>(h->s2->refl, true, x, radius, u, pt, nrm, pdf)
                       : sampleImpl<3>(h->s3->refl, true, x, radius, u, pt, nrm, pdf);
}

int orc_line_of_sight_bvh(void* hv, const float* p1, const float* p2) {
    SceneHandle* h = (SceneHandle*)hv;
    if (h->dim == 2) return hasLineOfSightBvh<2>(h->s2->refl, load<2>(p1), load<2>(p2));
    return hasLineOfSightBvh<3>(h->s3->refl, load<3>(p1), load<3>(p2));
}

// Solves.  rng_mode 0: one generator per point consumed across its walks (the reference's
// scheme, seeded deterministically: the generator of the point's walk 0); 1: one generator per
// walk, walkRng(seed, p, stream_offset + w) (the GPU's scheme).
// Stats arrays are accumulated with Chan's merge.  report: [codes0..3, totalSteps, nodeVisits, primVisits]

int orc_solve(int wost, void* hv, const zg_pde* pde, const zg_walk_settings* ws, int64_t n, const float* pts,
              const float* normals, const uint8_t* types, const uint8_t* aligned, const float* distA,
              const int32_t* nWalks, uint64_t seed, uint32_t streamOffset, int rngMode, int nThreads,
              int64_t* sN, double* sMean, double* sM2, int64_t* sLen, int64_t* sSplits, float* firstSphereRadius,
              int64_t* report) {
    SceneHandle* h = (SceneHandle*)hv;
    if (h->dim == 2) return solveImpl<2>(wost != 0, *h->s2, pde, ws, n, pts, normals, types, aligned, distA, nWalks, seed, streamOffset, rngMode, nThreads, sN, sMean, sM2, sLen, sSplits, firstSphereRadius, report);
    return solveImpl<3>(wost != 0, *h->s3, pde, ws, n, pts, normals, types, aligned, distA, nWalks, seed, streamOffset, rngMode, nThreads, sN, sMean, sM2, sLen, sSplits, firstSphereRadius, report);
}

// solve with per-point estimation quantities (EstimationQuantity, common.h:145-149) and the
// gradient statistics; cv_mode 0 refreshes the gradient control variates before every pair
// (the reference), 1 once per stage [0,1), [1,2), [2,4), ... (the GPU)
int orc_solve_grad(int wost, void* hv, const zg_pde* pde, const zg_walk_settings* ws, int64_t n, const float* pts,
                   const uint8_t* types, const float* distA, const float* distR, const int32_t* nWalks,
                   const uint8_t* quantity, const float* dirDeriv, uint64_t seed, uint32_t streamOffset, int rngMode,
                   int cvMode, int nThreads, int64_t* sN, double* sMean, double* sM2, int64_t* sLen, int64_t* sSplits,
                   int64_t* gN, double* gMean, double* gM2, double* gFirst, double* gDeriv, float* firstSphereRadius,
                   int64_t* report) {
    SceneHandle* h = (SceneHandle*)hv;
    if (h->dim == 2) return solveImpl<2>(wost != 0, *h->s2, pde, ws, n, pts, nullptr, types, nullptr, distA, nWalks, seed, streamOffset, rngMode, nThreads, sN, sMean, sM2, sLen, sSplits, firstSphereRadius, report, quantity, dirDeriv, cvMode, gN, gMean, gM2, gFirst, gDeriv, distR);
    return solveImpl<3>(wost != 0, *h->s3, pde, ws, n, pts, nullptr, types, nullptr, distA, nWalks, seed, streamOffset, rngMode, nThreads, sN, sMean, sM2, sLen, sSplits, firstSphereRadius, report, quantity, dirDeriv, cvMode, gN, gMean, gM2, gFirst, gDeriv, distR);
}

// the per-point Latin hypercube tables of a gradient solve: table t (2 * n_pairs rows of
// `dims` coordinates) is drawn from the generator of walk stream_offset of point t under ~seed
void orc_stratified_samples(int dims, int64_t nTables, int nSamples, uint64_t seed, uint32_t streamOffset, float* out) {
    std::vector<float> s;
    for (int64_t t = 0; t < nTables; t++) {
        Pcg32 r = walkRng(~seed, (uint32_t)t, streamOffset);
        stratifiedSamples(dims, nSamples, r, s);
        std::memcpy(out + (size_t)t * dims * nSamples, s.data(), s.size() * sizeof(float));
    }
}
void orc_pcg32_bounded(uint64_t initstate, uint64_t initseq, uint32_t bound, int64_t n, uint32_t* out) {
    Pcg32 r(initstate, initseq);
    for (int64_t i = 0; i < n; i++) out[i] = r.nextUInt(bound);
}
// Green's function gradient quantities of a centred ball of radius R at distance r along e_0:
// out = [gradientNorm(r), poissonKernelGradient(R e_0)[0]]
void orc_greens_ball_gradient(int dim, float R, float lambda, float r, float* out) {
    if (dim == 2) {
        GreensBall<2> g; if (lambda > 0) g.makeYukawa(lambda); g.updateBall(V<2>::zero(), R);
        V<2> y = V<2>::zero(); y[0] = R;
        out[0] = g.gradientNorm(r); out[1] = g.poissonKernelGradient(y)[0];
    } else {
        GreensBall<3> g; if (lambda > 0) g.makeYukawa(lambda); g.updateBall(V<3>::zero(), R);
        V<3> y = V<3>::zero(); y[0] = R;
        out[0] = g.gradientNorm(r); out[1] = g.poissonKernelGradient(y)[0];
    }
}
// boundary-direction samplers of the gradient path from given uniforms: mode 0 uniform sphere,
// 1 cosine hemisphere (+ the sign flip given in u[dims]) rotated to direction n; out = [dir, pdf]
void orc_gradient_direction(int dim, int mode, const float* u, const float* n, float* out) {
    if (dim == 2) {
        V<2> d; float pdf;
        if (mode == 0) { d = sphereFromU<2>(u); pdf = 1.0f / (2.0f * PI_F); }
        else { d = hemisphereCosineFromU<2>(u); if (u[1] < 0.5f) d[1] *= -1.0f; pdf = 0.5f * pdfHemisphereCosine<2>(std::fabs(d[1])); transformCoordinates(load<2>(n), d); }
        store<2>(out, d); out[2] = pdf;
    } else {
        V<3> d; float pdf;
        if (mode == 0) { d = sphereFromU<3>(u); pdf = 1.0f / (4.0f * PI_F); }
        else { d = hemisphereCosineFromU<3>(u); if (u[2] < 0.5f) d[2] *= -1.0f; pdf = 0.5f * pdfHemisphereCosine<3>(std::fabs(d[2])); transformCoordinates(load<3>(n), d); }
        store<3>(out, d); out[3] = pdf;
    }
}

}  // extern "C"
