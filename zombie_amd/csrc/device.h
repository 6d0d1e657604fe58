// device.h -- gfx950 device building blocks: PCG32, samplers, ball Green's functions,
// manufactured PDE data, and wide-BVH traversals with an LDS short stack per lane.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/zombie_amd.h"
#include "layout.h"

namespace zg {

#define ZG_DEV __device__ __forceinline__

constexpr float kPi = 3.14159265358979323846f;
constexpr float kPiF = kPi;

// ------------------------------------------------------------------------------------
// small vectors
// ------------------------------------------------------------------------------------
template <int D>
struct Vec {
    float v[D];
    ZG_DEV float& operator[](int i) { return v[i]; }
    ZG_DEV float operator[](int i) const { return v[i]; }
};
template <int D> ZG_DEV Vec<D> vzero() { Vec<D> r; for (int i = 0; i < D; i++) r.v[i] = 0.0f; return r; }
template <int D> ZG_DEV Vec<D> operator+(Vec<D> a, const Vec<D>& b) { for (int i = 0; i < D; i++) a.v[i] += b.v[i]; return a; }
template <int D> ZG_DEV Vec<D> operator-(Vec<D> a, const Vec<D>& b) { for (int i = 0; i < D; i++) a.v[i] -= b.v[i]; return a; }
template <int D> ZG_DEV Vec<D> operator*(Vec<D> a, float s) { for (int i = 0; i < D; i++) a.v[i] *= s; return a; }
template <int D> ZG_DEV float dot(const Vec<D>& a, const Vec<D>& b) { float s = 0.0f; for (int i = 0; i < D; i++) s += a.v[i] * b.v[i]; return s; }
template <int D> ZG_DEV float norm2(const Vec<D>& a) { return dot(a, a); }
template <int D> ZG_DEV Vec<D> vload(const float* p) { Vec<D> r; for (int i = 0; i < D; i++) r.v[i] = p[i]; return r; }
template <int D> ZG_DEV void vstore(float* p, const Vec<D>& a) { for (int i = 0; i < D; i++) p[i] = a.v[i]; }
ZG_DEV Vec<3> cross3(const Vec<3>& a, const Vec<3>& b) {
    Vec<3> r; r[0] = a[1] * b[2] - a[2] * b[1]; r[1] = a[2] * b[0] - a[0] * b[2]; r[2] = a[0] * b[1] - a[1] * b[0]; return r;
}
ZG_DEV float cross2(const Vec<2>& a, const Vec<2>& b) { return a[0] * b[1] - a[1] * b[0]; }

// ------------------------------------------------------------------------------------
// PCG32 (PCG-XSH-RR 64/32, O'Neill 2014); same stream semantics as the pcg32 dependency
// of core/sampling.h:19 (seed(initstate, initseq), nextUInt, mantissa nextFloat)
// ------------------------------------------------------------------------------------
struct Pcg32 {
    uint64_t state, inc;
    ZG_DEV uint32_t next() {
        uint64_t old = state;
        state = old * 0x5851f42d4c957f2dULL + inc;
        uint32_t xorshifted = (uint32_t)(((old >> 18u) ^ old) >> 27u);
        uint32_t rot = (uint32_t)(old >> 59u);
        return (xorshifted >> rot) | (xorshifted << ((32u - rot) & 31u));
    }
    ZG_DEV void seed(uint64_t initstate, uint64_t initseq) {
        state = 0u;
        inc = (initseq << 1u) | 1u;
        next();
        state += initstate;
        next();
    }
    ZG_DEV float nextFloat() { return __uint_as_float((next() >> 9) | 0x3f800000u) - 1.0f; }
    // uniform integer in [0, bound): draws below 2^32 mod bound are rejected (every residue
    // equally likely; pcg32 nextUInt(bound))
    ZG_DEV uint32_t nextBounded(uint32_t bound) {
        uint32_t threshold = (0u - bound) % bound;
        for (;;) {
            uint32_t r = next();
            if (r >= threshold) return r % bound;
        }
    }
};

// SplitMix64's output finaliser (Steele, Lea & Flood, "Fast splittable pseudorandom number
// generators", OOPSLA 2014; the constants of Vigna's splitmix64.c): a bijection of 64-bit words
// with full avalanche.
ZG_DEV uint64_t splitmix64Mix(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

// The generator of walk w (counted from the call's stream offset) of point p: PCG32 stream p,
// started at the state mix(seed, p, w).  Each mixing round is a bijection, so for a fixed seed
// and point the walks start at distinct, unrelated positions of the stream's 2^64-long cycle
// and never overlap in practice (a walk takes at most a few thousand draws).  Walks that
// shared one initstate on increments 2(p << 32 | w) + 1 would instead move in lockstep: their
// states at any draw index lie on an arithmetic progression in w, and their outputs correlate.
ZG_DEV uint64_t walkInitState(uint64_t seed, uint32_t p, uint32_t w) {
    uint64_t h = splitmix64Mix(seed ^ 0x9e3779b97f4a7c15ULL);
    h = splitmix64Mix(h ^ ((uint64_t)p * 0xd6e8feb86659fd93ULL));
    return splitmix64Mix(h ^ ((uint64_t)w * 0xa0761d6478bd642fULL));
}
ZG_DEV void seedWalk(Pcg32& rng, uint64_t seed, uint32_t p, uint32_t w) {
    rng.seed(walkInitState(seed, p, w), (uint64_t)p);
}

// ------------------------------------------------------------------------------------
// SphereSampler<DIM>::sampleUnitSphereUniform (core/sampling.h:104-111,164-174).
// v_sin_f32 / v_cos_f32 take their argument in revolutions: sin(2*pi*u) = v_sin(u).
// ------------------------------------------------------------------------------------
template <int D> ZG_DEV Vec<D> sampleUnitSphere(Pcg32& rng);
template <> ZG_DEV Vec<2> sampleUnitSphere<2>(Pcg32& rng) {
    float u = rng.nextFloat();
    Vec<2> d; d[0] = __builtin_amdgcn_cosf(u); d[1] = __builtin_amdgcn_sinf(u);
    return d;
}
template <> ZG_DEV Vec<3> sampleUnitSphere<3>(Pcg32& rng) {
    float u0 = rng.nextFloat();
    float u1 = rng.nextFloat();
    float z = 1.0f - 2.0f * u0;
    float r = __builtin_sqrtf(fmaxf(0.0f, 1.0f - z * z));
    Vec<3> d; d[0] = r * __builtin_amdgcn_cosf(u1); d[1] = r * __builtin_amdgcn_sinf(u1); d[2] = z;
    return d;
}

// the same maps from given uniforms (the stratified directions of the gradient path)
template <int D> ZG_DEV Vec<D> sphereFromU(const float* u);
template <> ZG_DEV Vec<2> sphereFromU<2>(const float* u) {
    Vec<2> d; d[0] = __builtin_amdgcn_cosf(u[0]); d[1] = __builtin_amdgcn_sinf(u[0]);
    return d;
}
template <> ZG_DEV Vec<3> sphereFromU<3>(const float* u) {
    float z = 1.0f - 2.0f * u[0];
    float r = __builtin_sqrtf(fmaxf(0.0f, 1.0f - z * z));
    Vec<3> d; d[0] = r * __builtin_amdgcn_cosf(u[1]); d[1] = r * __builtin_amdgcn_sinf(u[1]); d[2] = z;
    return d;
}

// cosine-weighted hemisphere about +e_{DIM-1} (core/sampling.h:124-129; 3D through Shirley's
// concentric disk map :185-217) and its density per solid angle (:146-149, :232-235)
template <int D> ZG_DEV Vec<D> hemisphereCosineFromU(const float* u);
template <> ZG_DEV Vec<2> hemisphereCosineFromU<2>(const float* u) {
    float u1 = 2.0f * u[0] - 1.0f;
    Vec<2> d; d[0] = u1; d[1] = __builtin_sqrtf(fmaxf(0.0f, 1.0f - u1 * u1));
    return d;
}
template <> ZG_DEV Vec<3> hemisphereCosineFromU<3>(const float* u) {
    float u1 = 2.0f * u[0] - 1.0f, u2 = 2.0f * u[1] - 1.0f;
    float dx = 0.0f, dy = 0.0f;
    if (!(u1 == 0.0f && u2 == 0.0f)) {
        // angle in revolutions: pi/4 (u2/u1) or pi/2 (1 - u1/(2 u2))
        float rev, r;
        if (fabsf(u1) > fabsf(u2)) { r = u1; rev = 0.125f * (u2 / u1); }
        else { r = u2; rev = 0.25f * (1.0f - 0.5f * (u1 / u2)); }
        dx = r * __builtin_amdgcn_cosf(rev); dy = r * __builtin_amdgcn_sinf(rev);
    }
    Vec<3> d; d[0] = dx; d[1] = dy; d[2] = __builtin_sqrtf(fmaxf(0.0f, 1.0f - (dx * dx + dy * dy)));
    return d;
}
template <int D> ZG_DEV float pdfHemisphereCosine(float cosTheta) { return D == 2 ? 0.5f * cosTheta : cosTheta * (1.0f / kPiF); }

// rotate d from the frame whose last axis is n to world coordinates (core/sampling.h:151-157;
// the 3D basis of :237-247 is Duff et al., "Building an Orthonormal Basis, Revisited", 2017)
ZG_DEV Vec<2> fromFrame(const Vec<2>& n, const Vec<2>& d) {
    Vec<2> r; r[0] = d[0] * n[1] + d[1] * n[0]; r[1] = -d[0] * n[0] + d[1] * n[1];
    return r;
}
ZG_DEV Vec<3> fromFrame(const Vec<3>& n, const Vec<3>& d) {
    float sign = copysignf(1.0f, n[2]);
    float a = -1.0f / (sign + n[2]);
    float b = n[0] * n[1] * a;
    Vec<3> r;
    r[0] = d[0] * (1.0f + sign * n[0] * n[0] * a) + d[1] * b + d[2] * n[0];
    r[1] = d[0] * (sign * b) + d[1] * (sign + n[1] * n[1] * a) + d[2] * n[1];
    r[2] = d[0] * (-sign * n[0]) + d[1] * (-n[1]) + d[2] * n[2];
    return r;
}

// ------------------------------------------------------------------------------------
// Green's function of the ball (core/distributions.h:453-665 harmonic, 835-1005 Yukawa 3D)
// ------------------------------------------------------------------------------------

// sinh x and I(x) = cosh x - sinh x / x (x^{3/2} I_{3/2} up to a constant) from e = e^{-x}:
// below the thresholds the exponential forms lose most digits to cancellation in float (at
// x = 0.004, 1 - e^2 keeps 4 correct digits and cosh x - sinh x / x none), so the Taylor series
// take over there (sinh: x (1 + x^2/6 (1 + x^2/20 (1 + x^2/42))); I: x^2/3 + x^4/30 + x^6/840 +
// x^8/45360, relative truncation error below 1e-8 at x = 0.5)
ZG_DEV float yukawaSinh(float x, float e) {
    if (x < 0.1f) { float x2 = x * x; return x * (1.0f + x2 * (1.0f / 6.0f) * (1.0f + x2 * (1.0f / 20.0f) * (1.0f + x2 * (1.0f / 42.0f)))); }
    return (1.0f - e * e) / (2.0f * e);
}
ZG_DEV float yukawaI(float x, float e) {
    if (x < 0.5f) {
        float x2 = x * x;
        return x2 * (1.0f / 3.0f + x2 * (1.0f / 30.0f + x2 * (1.0f / 840.0f + x2 * (1.0f / 45360.0f))));
    }
    return (1.0f + e * e) / (2.0f * e) - yukawaSinh(x, e) / x;
}

template <int D>
struct Ball {
    bool yukawa;
    float sqrtLambda;
    float R;
    float muR, expmuR, sinhmuR;  // Yukawa terms of updateBall (distributions.h:843-852)

    ZG_DEV void update(float R_) {
        R = R_;
        if (yukawa) {
            muR = R * sqrtLambda;
            expmuR = __expf(-muR);
            sinhmuR = yukawaSinh(muR, expmuR);
        }
    }
    ZG_DEV float evaluate(float r) const {
        if (!yukawa) {
            if (D == 2) return __logf(R / r) / (2.0f * kPi);
            return (1.0f / r - 1.0f / R) / (4.0f * kPi);
        }
        float mur = r * sqrtLambda;
        float expmur = __expf(-mur);
        float sinhmur = yukawaSinh(mur, expmur);
        return (expmur - expmuR * sinhmur / sinhmuR) / (4.0f * kPi * r);
    }
    ZG_DEV float norm() const {
        if (!yukawa) return D == 2 ? R * R * 0.25f : R * R * (1.0f / 6.0f);
        // (1 - 4 pi P)/lambda, P = muR/(4 pi sinh muR); for small muR the series of
        // 1 - x / sinh x = x^2/6 - 7 x^4/360 + 31 x^6/15120 - 127 x^8/604800 (Bernoulli numbers)
        if (muR < 0.5f) {
            float x2 = muR * muR;
            return R * R * (1.0f / 6.0f - x2 * (7.0f / 360.0f - x2 * (31.0f / 15120.0f - x2 * (127.0f / 604800.0f))));
        }
        float lambda = sqrtLambda * sqrtLambda;
        return (1.0f - muR / sinhmuR) / lambda;
    }
    // poissonKernelDampeningFactor(r) = mur (K32(mur) + I32(mur) e^{-muR}/sinh muR)
    ZG_DEV float dampening(float r) const {
        float mur = r * sqrtLambda;
        float expmur = __expf(-mur);
        float K32 = expmur * (1.0f + 1.0f / mur);
        float I32 = yukawaI(mur, expmur);
        return mur * (K32 + I32 * expmuR / sinhmuR);
    }
    ZG_DEV float pdfSphere(float r) const { return D == 2 ? 1.0f / (2.0f * kPi * r) : 1.0f / (4.0f * kPi * r * r); }

    // gradient of G(c, y) in the centre: (y - c) gradientNorm(r) (distributions.h:493-502,
    // 600-609, 904-922)
    ZG_DEV float gradientNorm(float r) const {
        if (!yukawa) {
            if (D == 2) return (1.0f / (r * r) - 1.0f / (R * R)) / (2.0f * kPi);
            return (1.0f / (r * r * r) - 1.0f / (R * R * R)) / (4.0f * kPi);
        }
        float mur = r * sqrtLambda;
        float expmur = __expf(-mur);
        float K32mur = expmur * (1.0f + 1.0f / mur);
        float I32mur = yukawaI(mur, expmur);
        float Qr = sqrtLambda * (K32mur - I32mur * K32R() / I32R());
        return Qr / (4.0f * kPi * r * r);
    }
    ZG_DEV float K32R() const { return expmuR * (1.0f + 1.0f / muR); }
    ZG_DEV float I32R() const { return yukawaI(muR, expmuR); }
    // Poisson kernel of the centred ball per unit solid angle, and its gradient in the centre
    // at y on the sphere (distributions.h:512-515,545-550, 628-631,654-659, 934-936,987-994)
    ZG_DEV float poissonKernel() const {
        if (!yukawa) return D == 2 ? 1.0f / (2.0f * kPi) : 1.0f / (4.0f * kPi);
        return muR / (4.0f * kPi * sinhmuR);
    }
    ZG_DEV float poissonKernelGradientScale() const {  // grad P = (y - c) * scale
        if (!yukawa) return D == 2 ? 2.0f / (2.0f * kPi * R * R) : 3.0f / (4.0f * kPi * R * R);
        return sqrtLambda * sqrtLambda / I32R() / (4.0f * kPi);
    }

    // sampleVolume(dir, rng, r, pdf) radius part; returns r (distributions.h:460-470,565-577,855-864,419-440)
    ZG_DEV float sampleRadius(Pcg32& rng) const { float pdf; return sampleRadiusPdf(rng, pdf); }
    // ... and the density of the sampled point (evaluate(r) / norm; before the clamp of r for
    // the Yukawa rejection sampler, after it for the harmonic ones, as the reference computes it)
    ZG_DEV float sampleRadiusPdf(Pcg32& rng, float& pdfOut) const {
        float r;
        if (!yukawa) {
            if (D == 2) {
                float u1 = rng.nextFloat();
                float u2 = rng.nextFloat();
                r = R * __builtin_sqrtf(u1 * u2);
            } else {
                float u1 = rng.nextFloat();
                float u2 = rng.nextFloat();
                float u3 = rng.nextFloat();
                float med = fmaxf(fminf(u1, u2), fminf(fmaxf(u1, u2), u3));
                r = R * med;
            }
        } else {
            float lambda = sqrtLambda * sqrtLambda;
            float bound = R <= lambda ?
                fmaxf(fmaxf(2.0f / R, 2.0f / lambda), fmaxf(0.5f * __builtin_sqrtf(R), 0.5f * sqrtLambda)) :
                fmaxf(fminf(2.0f / R, 2.0f / lambda), fminf(0.5f * __builtin_sqrtf(R), 0.5f * sqrtLambda));
            float fnNorm = norm();
            int iter = 0;
            do {
                float u = rng.nextFloat();
                r = rng.nextFloat() * R;
                pdfOut = evaluate(r) / fnNorm;
                float pdfRadius = pdfOut / pdfSphere(r);
                iter++;
                if (u < pdfRadius / bound) break;
            } while (iter < 1000);
        }
        r = fmaxf(1e-4f, r);
        if (r > R) r = R * 0.5f;
        if (!yukawa) pdfOut = evaluate(r) / norm();
        return r;
    }
};

// ------------------------------------------------------------------------------------
// manufactured PDE data: zg_function descriptors (stand-ins for the PDE<T,DIM> callbacks)
// ------------------------------------------------------------------------------------
// The manufactured fields use the hardware sin/cos/exp (v_sin_f32, v_cos_f32, v_exp_f32;
// absolute error near 1e-6 at the arguments these fields see, far below the Monte Carlo
// error): the accurate library versions cost 5% of the walk kernel's time.
ZG_DEV void fieldSinCos(float a, float* s, float* c) { __sincosf(a, s, c); }
ZG_DEV float fieldExp(float a) { return __expf(a); }
// DenseGrid<float,1,DIM>::operator() (utils/dense_grid.h:264-356): cell-centred values over
// [grid_min, grid_max], last axis fastest; nearest cell, or bi/trilinear interpolation between
// cell centres with the blend fractions clamped to [0, 1] (clamp to edge).
template <int D>
ZG_DEV float gridValue(const zg_function& f, const Vec<D>& x) {
    float l[3];
    int sh[3];
#pragma unroll
    for (int a = 0; a < D; a++) {
        float ext = f.grid_max[a] - f.grid_min[a];
        l[a] = (x[a] - f.grid_min[a]) / ext;
        sh[a] = f.shape[a];
    }
    if (!f.interpolate) {
        int flat = 0, stride = 1;
#pragma unroll
        for (int a = D - 1; a >= 0; a--) {
            int idx = min(max((int)floorf(l[a] * (float)sh[a]), 0), sh[a] - 1);
            flat += idx * stride;
            stride *= sh[a];
        }
        return f.grid[flat];
    }
    int i0[3], i1[3];
    float t[3];
#pragma unroll
    for (int a = 0; a < D; a++) {
        float u = sh[a] == 1 ? 0.0f : l[a] * (float)sh[a] - 0.5f;
        i0[a] = min(max((int)floorf(u), 0), sh[a] - 1);
        i1[a] = min(max(i0[a] + 1, 0), sh[a] - 1);
        t[a] = fminf(fmaxf(u - (float)i0[a], 0.0f), 1.0f);
    }
    if (D == 2) {
        const float* g = f.grid;
        float f00 = g[i0[0] * sh[1] + i0[1]], f01 = g[i0[0] * sh[1] + i1[1]];
        float f10 = g[i1[0] * sh[1] + i0[1]], f11 = g[i1[0] * sh[1] + i1[1]];
        return f00 * (1.0f - t[0]) * (1.0f - t[1]) + f10 * t[0] * (1.0f - t[1]) + f01 * (1.0f - t[0]) * t[1] +
               f11 * t[0] * t[1];
    }
    const float* g = f.grid;
    auto at = [&](int i, int j, int k) { return g[(i * sh[1] + j) * sh[2] + k]; };
    float r = 0.0f;
#pragma unroll
    for (int c = 0; c < 8; c++) {
        int ci = (c >> 2) & 1, cj = (c >> 1) & 1, ck = c & 1;
        float wgt = (ci ? t[0] : 1.0f - t[0]) * (cj ? t[1] : 1.0f - t[1]) * (ck ? t[2] : 1.0f - t[2]);
        r += wgt * at(ci ? i1[0] : i0[0], cj ? i1[1] : i0[1], ck ? i1[2] : i0[2]);
    }
    return r;
}

template <int D>
ZG_DEV float evalField(const zg_function& f, const Vec<D>& x, const Vec<D>& n, bool aligned, float absorption) {
    if (f.kind == ZG_FN_DENSE_GRID) return gridValue<D>(f, x);
    const float* p = f.params;
    float u = 0.0f, g0 = 0.0f, g1 = 0.0f, g2 = 0.0f, negLap = 0.0f;
    switch (f.kind) {
    case ZG_FN_CONSTANT: u = p[0]; break;
    case ZG_FN_EXP_COS: {
        float e = fieldExp(p[1] * x[0]);
        float s, c;
        fieldSinCos(p[1] * x[1], &s, &c);
        u = p[0] * e * c;
        g0 = p[0] * p[1] * e * c;
        g1 = -p[0] * p[1] * e * s;
        break;
    }
    case ZG_FN_SIN_PRODUCT: {
        float s[3] = {1.0f, 1.0f, 1.0f}, c[3] = {0.0f, 0.0f, 0.0f};
        for (int i = 0; i < D; i++) fieldSinCos(p[1] * x[i], &s[i], &c[i]);
        u = p[0] * s[0] * s[1] * s[2];
        g0 = p[0] * p[1] * c[0] * s[1] * s[2];
        g1 = p[0] * p[1] * s[0] * c[1] * s[2];
        g2 = D == 3 ? p[0] * p[1] * s[0] * s[1] * c[2] : 0.0f;
        negLap = (float)D * p[1] * p[1] * u;
        break;
    }
    case ZG_FN_CUBIC_HARMONIC: {
        float xx = x[0], yy = x[1];
        u = p[0] * (xx * xx * xx - 3.0f * xx * yy * yy);
        g0 = p[0] * (3.0f * xx * xx - 3.0f * yy * yy);
        g1 = -p[0] * 6.0f * xx * yy;
        break;
    }
    case ZG_FN_EXP_LINEAR: {
        float a = p[1] * x[0] + p[2] * x[1] + (D == 3 ? p[3] * x[D == 3 ? 2 : 0] : 0.0f);
        u = p[0] * fieldExp(a);
        g0 = p[1] * u; g1 = p[2] * u; g2 = D == 3 ? p[3] * u : 0.0f;
        float a2 = p[1] * p[1] + p[2] * p[2] + (D == 3 ? p[3] * p[3] : 0.0f);
        negLap = -a2 * u;
        break;
    }
    default: break;
    }
    if (f.eval == ZG_EVAL_VALUE) return u;
    if (f.eval == ZG_EVAL_NORMAL_DERIVATIVE || f.eval == ZG_EVAL_ROBIN) {
        float dn = g0 * n[0] + g1 * n[1] + (D == 3 ? g2 * n[D == 3 ? 2 : 0] : 0.0f);
        dn = aligned ? -dn : dn;
        return f.eval == ZG_EVAL_ROBIN ? dn + p[5] * u : dn;
    }
    return negLap + absorption * u;
}

// ------------------------------------------------------------------------------------
// SDF-grid geometry (utils/sdf_grid_geometric_queries.h:87-190): value by trilinear
// interpolation of the grid, central-difference gradient with the stencil clamped into the grid,
// steepest-descent projection to the zero level set (at most 8 steps, |sdf| < 1e-5 stops)
// ------------------------------------------------------------------------------------
template <int D>
ZG_DEV Vec<D> sdfClampIn(const zg_function& f, Vec<D> x) {
#pragma unroll
    for (int a = 0; a < D; a++) x[a] = fminf(fmaxf(x[a], f.grid_min[a] + 1e-6f), f.grid_max[a] - 1e-6f);
    return x;
}
template <int D>
ZG_DEV Vec<D> sdfGradient(const zg_function& f, const Vec<D>& x) {
    Vec<D> g;
#pragma unroll
    for (int a = 0; a < D; a++) {
        const float spacing = (f.grid_max[a] - f.grid_min[a]) / (float)f.shape[a];
        Vec<D> xm = x, xp = x;
        xm[a] -= spacing; xp[a] += spacing;
        xm = sdfClampIn<D>(f, xm); xp = sdfClampIn<D>(f, xp);
        const float denom = xp[a] - xm[a];
        g[a] = denom > 0.0f ? (gridValue<D>(f, xp) - gridValue<D>(f, xm)) / denom : 0.0f;
    }
    return g;
}
template <int D>
ZG_DEV Vec<D> sdfNormal(const zg_function& f, const Vec<D>& x) {
    Vec<D> g = sdfGradient<D>(f, x);
    return g * (1.0f / (__builtin_sqrtf(norm2(g)) + 1e-6f));
}
template <int D>
ZG_DEV void sdfProject(const zg_function& f, Vec<D>& x, Vec<D>& n) {
    for (int i = 0; i < 8; i++) {
        const float v = gridValue<D>(f, x);
        if (fabsf(v) < 1e-5f) break;
        Vec<D> g = sdfGradient<D>(f, x);
        x = sdfClampIn<D>(f, x - g * (v / (__builtin_sqrtf(norm2(g)) + 1e-6f)));
    }
    n = sdfNormal<D>(f, x);
}
// intersectZeroLevelSet (sdf_grid_geometric_queries.h:147-205): the ray clipped to the grid box,
// sphere-traced with steps |sdf| (at most 128; hit when |sdf| < 1e-5); a sign change between two
// samples is resolved by one secant step and one gradient step.
template <int D>
ZG_DEV bool sdfIntersect(const zg_function& f, const Vec<D>& o, const Vec<D>& dir, float tMax, Vec<D>& pt, Vec<D>& n,
                         float& tHit) {
    float tEnter = 0.0f, tExit = tMax;
#pragma unroll
    for (int a = 0; a < D; a++) {
        const float inv = 1.0f / dir[a];
        const float t0 = (f.grid_min[a] - o[a]) * inv, t1 = (f.grid_max[a] - o[a]) * inv;
        tEnter = fmaxf(tEnter, fminf(t0, t1));
        tExit = fminf(tExit, fmaxf(t0, t1));
    }
    if (tEnter > tExit) return false;
    float tPrev = tEnter, sPrev = 0.0f, t = tEnter;
    for (int i = 0; i < 128 && t <= tExit; i++) {
        const Vec<D> p = o + dir * t;
        const float s = gridValue<D>(f, p);
        if (fabsf(s) < 1e-5f) { pt = p; n = sdfNormal<D>(f, p); tHit = t; return true; }
        if (sPrev * s < 0.0f) {
            const float th = tPrev + (sPrev / (sPrev - s)) * (t - tPrev);
            Vec<D> ph = o + dir * th;
            const float sh = gridValue<D>(f, ph);
            const Vec<D> g = sdfGradient<D>(f, ph);
            ph = ph - g * (sh / (__builtin_sqrtf(norm2(g)) + 1e-6f));
            pt = ph; n = sdfNormal<D>(f, ph); tHit = dot(ph - o, dir);
            return true;
        }
        tPrev = t; sPrev = s;
        t += fabsf(s);
    }
    return false;
}

// ------------------------------------------------------------------------------------
// primitives (Ericson, "Real-Time Collision Detection" 5.1.2 / 5.1.5; the same closed forms
// as core/geometry_helpers.h:28-49,94-161)
// ------------------------------------------------------------------------------------
ZG_DEV float closestOnSegment2(const Vec<2>& pa, const Vec<2>& pb, const Vec<2>& x, Vec<2>& pt) {
    Vec<2> u = pb - pa, v = x - pa;
    float c1 = dot(u, v);
    float c2 = dot(u, u);
    float t = c1 <= 0.0f ? 0.0f : (c2 <= c1 ? 1.0f : c1 / c2);
    pt = c1 <= 0.0f ? pa : (c2 <= c1 ? pb : pa + u * t);
    return norm2(x - pt);
}

ZG_DEV float closestOnTriangle2(const Vec<3>& pa, const Vec<3>& pb, const Vec<3>& pc, const Vec<3>& x, Vec<3>& pt) {
    Vec<3> ab = pb - pa, ac = pc - pa, ax = x - pa;
    float d1 = dot(ab, ax), d2 = dot(ac, ax);
    Vec<3> bx = x - pb;
    float d3 = dot(ab, bx), d4 = dot(ac, bx);
    Vec<3> cx = x - pc;
    float d5 = dot(ab, cx), d6 = dot(ac, cx);
    float vc = d1 * d4 - d3 * d2;
    float vb = d5 * d2 - d1 * d6;
    float va = d3 * d6 - d5 * d4;
    if (d1 <= 0.0f && d2 <= 0.0f) {
        pt = pa;
    } else if (d3 >= 0.0f && d4 <= d3) {
        pt = pb;
    } else if (d6 >= 0.0f && d5 <= d6) {
        pt = pc;
    } else if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f) {
        pt = pa + ab * (d1 / (d1 - d3));
    } else if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f) {
        pt = pa + ac * (d2 / (d2 - d6));
    } else if (va <= 0.0f && (d4 - d3) >= 0.0f && (d5 - d6) >= 0.0f) {
        pt = pb + (pc - pb) * ((d4 - d3) / ((d4 - d3) + (d5 - d6)));
    } else {
        float denom = 1.0f / (va + vb + vc);
        pt = pa + ab * (vb * denom) + ac * (vc * denom);
    }
    return norm2(x - pt);
}

ZG_DEV Vec<2> segmentNormal(const Vec<2>& pa, const Vec<2>& pb) {
    Vec<2> s = pb - pa; Vec<2> n; n[0] = s[1]; n[1] = -s[0];
    return n * (1.0f / __builtin_sqrtf(norm2(n)));
}
ZG_DEV Vec<3> triangleNormal(const Vec<3>& pa, const Vec<3>& pb, const Vec<3>& pc) {
    Vec<3> n = cross3(pb - pa, pc - pa);
    return n * (1.0f / __builtin_sqrtf(norm2(n)));
}

template <int D> struct Prim;
template <> struct Prim<2> {
    Vec<2> a, b;
    ZG_DEV void load(const float* prims, int i) {
        float4 q = reinterpret_cast<const float4*>(prims)[i];
        a[0] = q.x; a[1] = q.y; b[0] = q.z; b[1] = q.w;
    }
    ZG_DEV float closest2(const Vec<2>& x, Vec<2>& pt) const { return closestOnSegment2(a, b, x, pt); }
    ZG_DEV Vec<2> normal() const { return segmentNormal(a, b); }
    // ray o + t d, t in [0, tMax]
    ZG_DEV bool ray(const Vec<2>& o, const Vec<2>& d, float tMax, float& t) const {
        Vec<2> u = a - o, v = b - a;
        float dv = cross2(d, v);
        if (dv == 0.0f) return false;
        float s = cross2(u, d) / dv;
        if (s < 0.0f || s > 1.0f) return false;
        t = cross2(u, v) / dv;
        return t >= 0.0f && t <= tMax;
    }
};
template <> struct Prim<3> {
    Vec<3> a, b, c;
    ZG_DEV void load(const float* prims, int i) {
        const float4* q = reinterpret_cast<const float4*>(prims) + 3 * i;
        float4 q0 = q[0], q1 = q[1], q2 = q[2];
        a[0] = q0.x; a[1] = q0.y; a[2] = q0.z; b[0] = q0.w;
        b[1] = q1.x; b[2] = q1.y; c[0] = q1.z; c[1] = q1.w; c[2] = q2.x;
    }
    ZG_DEV float closest2(const Vec<3>& x, Vec<3>& pt) const { return closestOnTriangle2(a, b, c, x, pt); }
    ZG_DEV Vec<3> normal() const { return triangleNormal(a, b, c); }
    // Moller-Trumbore
    ZG_DEV bool ray(const Vec<3>& o, const Vec<3>& d, float tMax, float& t) const {
        Vec<3> e1 = b - a, e2 = c - a;
        Vec<3> p = cross3(d, e2);
        float det = dot(e1, p);
        if (det == 0.0f) return false;
        float inv = 1.0f / det;
        Vec<3> s = o - a;
        float u = dot(s, p) * inv;
        if (u < 0.0f || u > 1.0f) return false;
        Vec<3> q = cross3(s, e1);
        float v = dot(d, q) * inv;
        if (v < 0.0f || u + v > 1.0f) return false;
        t = dot(e2, q) * inv;
        return t >= 0.0f && t <= tMax;
    }
};

// ------------------------------------------------------------------------------------
// wide node access
// ------------------------------------------------------------------------------------
template <int D, int W>
struct WideNode {
    float lo[D][W], hi[D][W];
    int ch[W];
    ZG_DEV void load(const float* nodes, int idx) {
        using L = NodeLayout<D, W>;
        const float4* q = reinterpret_cast<const float4*>(nodes + (size_t)idx * L::kFloats);
        constexpr int kBoxVec = (2 * D * W) / 4;
        float tmp[2 * D * W];
#pragma unroll
        for (int k = 0; k < kBoxVec; k++) {
            float4 v = q[k];
            tmp[4 * k] = v.x; tmp[4 * k + 1] = v.y; tmp[4 * k + 2] = v.z; tmp[4 * k + 3] = v.w;
        }
#pragma unroll
        for (int a = 0; a < D; a++)
#pragma unroll
            for (int k = 0; k < W; k++) { lo[a][k] = tmp[a * W + k]; hi[a][k] = tmp[(D + a) * W + k]; }
        const int4* c = reinterpret_cast<const int4*>(q + kBoxVec);
#pragma unroll
        for (int k = 0; k < W / 4; k++) {
            int4 v = c[k];
            ch[4 * k] = v.x; ch[4 * k + 1] = v.y; ch[4 * k + 2] = v.z; ch[4 * k + 3] = v.w;
        }
    }
    ZG_DEV float dist2(int k, const Vec<D>& x) const {
        float s = 0.0f;
#pragma unroll
        for (int a = 0; a < D; a++) {
            float d = fmaxf(fmaxf(lo[a][k] - x[a], x[a] - hi[a][k]), 0.0f);
            s += d * d;
        }
        return s;
    }
};

ZG_DEV bool isLeaf(int c) { return c < 0; }
ZG_DEV int leafFirstD(int c) { return (~c) >> 4; }
ZG_DEV int leafCountD(int c) { return ((~c) & 15) + 1; }

// Traversal stack: the top kLdsStack entries of a lane live in LDS (dynamic shared memory), as
// two arrays laid out [depth][blockDim] so that lanes at the same depth hit consecutive banks:
// the 32-bit node codes, then the upper 16 bits of the keys.  A non-negative float cut to its
// upper 16 bits rounds toward zero, so a popped key is a lower bound of the pushed one (relative
// error below 2^-8): pruning on pop stays conservative (at worst a node is visited and its
// children pruned), and a radius certified from pruned keys can only get smaller.  Six bytes per
// entry give 16 entries in the 24 KB per 256-thread block that 6 blocks per CU allow (eight-byte
// entries gave 12, and 10.6% of the pushes of the 3D walk went to the spill array).  Deeper
// entries spill, with their full keys, to a per-lane private array owned by the caller.
// Kernels that traverse launch with kStackLdsBytes of dynamic LDS and at most 256 threads.
constexpr int kLdsStack = 16;
constexpr int kSpillStack = 96;
constexpr int kStackLdsBytes = kLdsStack * 256 * 6;

extern __shared__ uint2 zgLdsStack[];

// The spill array is addressed through a private (address space 5) pointer: its loads and
// stores are scratch instructions that stay in the cache hierarchy, and the LDS load of pop()
// cannot be merged with the spill load into one flat load.  (A volatile generic pointer, the
// earlier form, compiled to system-scope write-through flat stores, each followed by a wait.)
typedef __attribute__((address_space(5))) unsigned long long PrivU64;

struct Stack {
    int col;          // threadIdx.x
    int stride;       // blockDim.x
    int sp;           // entries in LDS
    int ovf;          // entries in the spill array
    bool overflow;
    PrivU64* spill;   // caller's private array of kSpillStack entries
    ZG_DEV void init(unsigned long long* spillArray) {
        col = threadIdx.x; stride = blockDim.x; sp = 0; ovf = 0; overflow = false; spill = (PrivU64*)spillArray;
    }
    ZG_DEV int* codes() const { return reinterpret_cast<int*>(zgLdsStack); }
    ZG_DEV uint16_t* keys() const {
        return reinterpret_cast<uint16_t*>(reinterpret_cast<int*>(zgLdsStack) + kLdsStack * stride);
    }
    // d2 >= 0 (a squared distance or a ray parameter)
    ZG_DEV void push(int node, float d2) {
        const unsigned k = __float_as_uint(d2);
        if (sp < kLdsStack) {
            codes()[sp * stride + col] = node;
            keys()[sp * stride + col] = (uint16_t)(k >> 16);
            sp++;
        } else if (ovf < kSpillStack) {
            spill[ovf] = ((unsigned long long)k << 32) | (unsigned)node;
            ovf++;
        } else {
            overflow = true;
        }
    }
    // pushes the children k = W-1 .. 1 of a sorted child list whose keys are < FLT_MAX (a
    // prefix: invalid children sort last), the nearest on top, as that many push() calls would.
    // With room for W-1 entries in LDS every slot is written without a branch: the valid
    // children fill sp .. sp+nv-1 and the invalid ones land above the new top, where they are
    // never read.
    template <int W>
    ZG_DEV void pushChildren(const float (&d2)[W], const int (&ch)[W]) {
        int nv = 0;
#pragma unroll
        for (int k = 1; k < W; k++) nv += d2[k] < FLT_MAX ? 1 : 0;
        if (ovf == 0 && sp + (W - 1) <= kLdsStack) {
#pragma unroll
            for (int k = 1; k < W; k++) {
                const int slot = k <= nv ? sp + nv - k : sp + k - 1;
                codes()[slot * stride + col] = ch[k];
                keys()[slot * stride + col] = (uint16_t)(__float_as_uint(d2[k]) >> 16);
            }
            sp += nv;
        } else {
#pragma unroll
            for (int k = W - 1; k >= 1; k--)
                if (d2[k] < FLT_MAX) push(ch[k], d2[k]);
        }
    }
    ZG_DEV bool pop(int& node, float& d2) {
        if (ovf > 0) {
            ovf--;
            const unsigned long long v = spill[ovf];
            node = (int)(unsigned)v;
            d2 = __uint_as_float((unsigned)(v >> 32));
        } else if (sp > 0) {
            sp--;
            node = codes()[sp * stride + col];
            d2 = __uint_as_float((unsigned)keys()[sp * stride + col] << 16);
        } else {
            return false;
        }
        return true;
    }
};

struct VisitCount { uint32_t nodes, prims; };

// Squared distances from x to the W child boxes of node `idx`, streamed axis by axis
// (two dwordx4 loads per bound and axis for W = 8) so that only the running sums stay live.
template <int D, int W>
ZG_DEV void childDistances(const float* nodes, int idx, const Vec<D>& x, float (&d2)[W], int (&ch)[W]) {
    using L = NodeLayout<D, W>;
    const float4* q = reinterpret_cast<const float4*>(nodes + (size_t)idx * L::kFloats);
#pragma unroll
    for (int k = 0; k < W; k++) d2[k] = 0.0f;
#pragma unroll
    for (int a = 0; a < D; a++) {
#pragma unroll
        for (int v = 0; v < W / 4; v++) {
            float4 lo = q[a * (W / 4) + v];
            float4 hi = q[(D + a) * (W / 4) + v];
            float lov[4] = {lo.x, lo.y, lo.z, lo.w}, hiv[4] = {hi.x, hi.y, hi.z, hi.w};
#pragma unroll
            for (int j = 0; j < 4; j++) {
                float d = fmaxf(fmaxf(lov[j] - x[a], x[a] - hiv[j]), 0.0f);
                d2[4 * v + j] += d * d;
            }
        }
    }
    const int4* c = reinterpret_cast<const int4*>(q + (2 * D * W) / 4);
#pragma unroll
    for (int v = 0; v < W / 4; v++) {
        int4 cv = c[v];
        ch[4 * v] = cv.x; ch[4 * v + 1] = cv.y; ch[4 * v + 2] = cv.z; ch[4 * v + 3] = cv.w;
    }
}

// Resumable closest-point query on one boundary part.  A query advances by one work item per
// call of visit(), so a persistent kernel can interleave the traversals of its lanes with
// their walk steps instead of making every lane of a wave wait for the slowest traversal.
//
// Work items are either an inner node (the squared distances to its W child boxes) or one
// primitive of a leaf: leaves are not tested when their parent is visited but queued like
// inner children, and a leaf of n primitives takes n items.  Both kinds of item cost about
// the same number of instructions, so the lanes of a wave stay in step (testing all leaf
// children of a node at once made one lane's visit cost up to W * leaf-size triangle tests
// while its neighbours idled).
//
// Order: the surviving children of a node (inner or leaf, box distance within the pruning
// bound) are sorted by a compare-exchange network; the nearest becomes the next item, the
// others are pushed farthest-first so the stack pops them nearest-first.  (Role of fcpw
// Aggregate::findClosestPoint, fcpw_geometric_queries.h:1069-1096.)
//
// Bound: best2 starts at the caller's bound (FLT_MAX for an unbounded search); a query that
// finds nothing within it is restarted once without bound.
//
// Certified relaxation: relax2 = (1 - tau)^2 < 1 prunes boxes already within a factor
// (1 - tau) of the best distance once best2 > guard2.  Every box pruned that way lies at least
// (1 - tau) times the final best distance away, so radius2 = min(best2, smallest such box
// distance^2) is a certified lower bound of the true squared distance (exact when nothing was
// pruned that way).  Walk on spheres stays unbiased with any ball inside the domain, so the
// walk uses radius2 while cp stays the best boundary point found.
template <int D>
struct ClosestQuery {
    Vec<D> cp;       // best point found
    float best2;     // its squared distance (or the bound)
    float prune2;    // current pruning bound
    float lower2;    // smallest box distance^2 pruned by the relaxation
    float relax2, guard2;
    float itemD2;    // box distance^2 of the current item
    int item;        // current item: inner node index (>= 0) or leaf code (< 0)
    int prim;        // packed index of the best primitive
    bool has;        // an item is held (else pop one)
    bool found;
    bool retried;

    ZG_DEV void begin(float bound2, float relax2_, float guard2_) {
        best2 = bound2; relax2 = relax2_; guard2 = guard2_;
        prune2 = best2 > guard2 ? best2 * relax2 : best2;
        lower2 = FLT_MAX; item = 0; itemD2 = 0.0f; has = true; prim = -1; found = false; retried = false;
    }
    ZG_DEV float radius2() const { return fminf(best2, lower2); }
    ZG_DEV void prunedBox(float d2) { if (d2 <= best2) lower2 = fminf(lower2, d2); }
    // a known point of the boundary (squared distance d2 from the query) tightens the bound
    // before the traversal starts
    ZG_DEV void seed(const Vec<D>& p, int primIdx, float d2) {
        if (d2 < best2) {
            best2 = d2; cp = p; prim = primIdx; found = true;
            prune2 = best2 > guard2 ? best2 * relax2 : best2;
        }
    }

    template <int W>
    ZG_DEV static void cmpx(float (&k)[W], int (&v)[W], int i, int j) {
        bool sw = k[j] < k[i];
        float ki = k[i], kj = k[j];
        int vi = v[i], vj = v[j];
        k[i] = sw ? kj : ki; k[j] = sw ? ki : kj;
        v[i] = sw ? vj : vi; v[j] = sw ? vi : vj;
    }
    // ascending sort network (Batcher odd-even merge for 8, optimal for 4)
    template <int W>
    ZG_DEV static void sortAscending(float (&k)[W], int (&v)[W]) {
        if constexpr (W == 4) {
            cmpx<4>(k, v, 0, 1); cmpx<4>(k, v, 2, 3); cmpx<4>(k, v, 0, 2); cmpx<4>(k, v, 1, 3); cmpx<4>(k, v, 1, 2);
        } else {
            cmpx<8>(k, v, 0, 1); cmpx<8>(k, v, 2, 3); cmpx<8>(k, v, 4, 5); cmpx<8>(k, v, 6, 7);
            cmpx<8>(k, v, 0, 2); cmpx<8>(k, v, 1, 3); cmpx<8>(k, v, 4, 6); cmpx<8>(k, v, 5, 7);
            cmpx<8>(k, v, 1, 2); cmpx<8>(k, v, 5, 6);
            cmpx<8>(k, v, 0, 4); cmpx<8>(k, v, 1, 5); cmpx<8>(k, v, 2, 6); cmpx<8>(k, v, 3, 7);
            cmpx<8>(k, v, 2, 4); cmpx<8>(k, v, 3, 5);
            cmpx<8>(k, v, 1, 2); cmpx<8>(k, v, 3, 4); cmpx<8>(k, v, 5, 6);
        }
    }

    // advances the query on the view b (anything with nodes, prims, nNodes) by one item;
    // returns true when the query is complete
    template <int W, bool COUNT, class View>
    ZG_DEV bool visit(const View& b, const Vec<D>& x, Stack& st, VisitCount& vc) {
        if (b.nNodes == 0) return true;
        if (!has) {
            // one pop per call: a pruned entry ends the call, so a lane draining a stack of
            // pruned entries does not hold the wave in a loop of dependent LDS reads
            int n; float dn;
            bool got = false;
            if (st.pop(n, dn)) {
                if (dn > prune2) { prunedBox(dn); return false; }
                got = true;
            }
            if (!got) {
                if (found || retried) return true;
                // nothing within the initial bound: search again without it
                retried = true;
                best2 = FLT_MAX; prune2 = FLT_MAX; lower2 = FLT_MAX;
                n = 0; dn = 0.0f;
            }
            item = n; itemD2 = dn; has = true;
        }
        if (!isLeaf(item)) {
            // an inner node: its nearest surviving child becomes the next item
            float d2[W];
            int ch[W];
            childDistances<D, W>(b.nodes, item, x, d2, ch);
            if (COUNT) vc.nodes++;
            float low = lower2;
#pragma unroll
            for (int k = 0; k < W; k++) {
                const bool pr = d2[k] > prune2;
                low = fminf(low, pr && d2[k] <= best2 ? d2[k] : FLT_MAX);
                d2[k] = pr ? FLT_MAX : d2[k];
            }
            lower2 = low;
            sortAscending<W>(d2, ch);
            has = d2[0] < FLT_MAX;
            item = ch[0]; itemD2 = d2[0];
            st.pushChildren<W>(d2, ch);
            if (!has) return false;
        }
        if (isLeaf(item)) {
            // one primitive of a leaf (in the same call as the node that led to it, so a call
            // costs at most one node and one primitive)
            int first = leafFirstD(item), cnt = leafCountD(item);
            Prim<D> pr; pr.load(b.prims, first);
            Vec<D> qq;
            float dd = pr.closest2(x, qq);
            if (COUNT) vc.prims++;
            const bool better = dd < best2;
#pragma unroll
            for (int a = 0; a < D; a++) cp[a] = better ? qq[a] : cp[a];
            prim = better ? first : prim;
            found = found || better;
            best2 = better ? dd : best2;
            prune2 = better ? (best2 > guard2 ? best2 * relax2 : best2) : prune2;
            has = cnt > 1 && itemD2 <= prune2;
            if (cnt > 1 && !has) prunedBox(itemD2);
            item = ~(((first + 1) << 4) | (cnt - 2));
        }
        return false;
    }
};

// Complete closest-point query (loops ClosestQuery::visit); see ClosestQuery for the rules.
// seedPrim >= 0 names a primitive (packed index) likely to be close, tested first.
template <int D, int W, bool COUNT>
ZG_DEV bool closestPoint(const BoundaryView& b, const Vec<D>& x, float& best2, Vec<D>& cp, int& prim,
                         Stack& st, VisitCount& vc, float relax2 = 1.0f, float guard2 = 0.0f,
                         float* radius2 = nullptr, int seedPrim = -1) {
    ClosestQuery<D> q;
    q.begin(best2, relax2, guard2);
    q.cp = cp;
    if (seedPrim >= 0) {
        Prim<D> pr; pr.load(b.prims, seedPrim);
        Vec<D> qq;
        q.seed(qq, seedPrim, pr.closest2(x, qq));
        if (COUNT) vc.prims++;
    }
    st.sp = 0; st.ovf = 0;
    while (!q.template visit<W, COUNT>(b, x, st, vc)) {}
    best2 = q.best2; cp = q.cp; prim = q.prim;
    if (radius2) *radius2 = q.radius2();
    return q.found;
}

// First hit of a ray with one boundary part, 0 <= t <= tMax.
template <int D, int W>
ZG_DEV bool intersectRay(const BoundaryView& b, const Vec<D>& o, const Vec<D>& dir, float tMax, float& tHit,
                         int& prim, Stack& st) {
    bool found = false;
    tHit = tMax;
    if (b.nNodes == 0) return false;
    Vec<D> inv;
#pragma unroll
    for (int a = 0; a < D; a++) inv[a] = 1.0f / dir[a];
    st.sp = 0; st.ovf = 0;
    int node = 0;
    for (;;) {
        WideNode<D, W> nd;
        nd.load(b.nodes, node);
        float tn[W];
        bool hitk[W];
#pragma unroll
        for (int k = 0; k < W; k++) {
            float t0 = 0.0f, t1 = tHit;
#pragma unroll
            for (int a = 0; a < D; a++) {
                float ta = (nd.lo[a][k] - o[a]) * inv[a], tb = (nd.hi[a][k] - o[a]) * inv[a];
                t0 = fmaxf(t0, fminf(ta, tb));
                t1 = fminf(t1, fmaxf(ta, tb));
            }
            hitk[k] = t0 <= t1;
            tn[k] = t0;
        }
#pragma unroll
        for (int k = 0; k < W; k++) {
            if (hitk[k] && isLeaf(nd.ch[k]) && tn[k] <= tHit) {
                int first = leafFirstD(nd.ch[k]), cnt = leafCountD(nd.ch[k]);
                for (int i = first; i < first + cnt; i++) {
                    Prim<D> pr; pr.load(b.prims, i);
                    float t;
                    if (pr.ray(o, dir, tHit, t) && (!found || t < tHit)) { tHit = t; prim = i; found = true; }
                }
            }
        }
        int next = -1; float nextT = 0.0f; int nk = -1;
#pragma unroll
        for (int k = 0; k < W; k++)
            if (hitk[k] && !isLeaf(nd.ch[k]) && tn[k] <= tHit && (next < 0 || tn[k] < nextT)) { next = nd.ch[k]; nextT = tn[k]; nk = k; }
#pragma unroll
        for (int k = 0; k < W; k++)
            if (hitk[k] && !isLeaf(nd.ch[k]) && tn[k] <= tHit && k != nk) st.push(nd.ch[k], tn[k]);
        if (next < 0) {
            int n; float tnn;
            bool got = false;
            while (st.pop(n, tnn)) { if (tnn <= tHit) { got = true; break; } }
            if (!got) break;
            next = n;
        }
        node = next;
    }
    return found;
}

// Number of primitives crossed by the ray from x along +axis (all hits, t >= 0).
template <int D, int W>
ZG_DEV int countAxisHits(const BoundaryView& b, const Vec<D>& x, int axis, Stack& st) {
    int hits = 0;
    if (b.nNodes == 0) return 0;
    Vec<D> dir = vzero<D>(); dir[axis] = 1.0f;
    st.sp = 0; st.ovf = 0;
    int node = 0;
    for (;;) {
        WideNode<D, W> nd;
        nd.load(b.nodes, node);
#pragma unroll
        for (int k = 0; k < W; k++) {
            bool ov = nd.hi[axis][k] >= x[axis] && nd.lo[axis][k] < 1e29f;
#pragma unroll
            for (int a = 0; a < D; a++)
                if (a != axis) ov = ov && nd.lo[a][k] <= x[a] && x[a] <= nd.hi[a][k];
            if (!ov) continue;
            if (isLeaf(nd.ch[k])) {
                int first = leafFirstD(nd.ch[k]), cnt = leafCountD(nd.ch[k]);
                for (int i = first; i < first + cnt; i++) {
                    Prim<D> pr; pr.load(b.prims, i);
                    float t;
                    if (pr.ray(x, dir, 3.0e38f, t)) hits++;
                }
            } else {
                st.push(nd.ch[k], 0.0f);
            }
        }
        int n; float dn;
        if (!st.pop(n, dn)) break;
        node = n;
    }
    return hits;
}


// lanes of the wave below this one that are set in mask
ZG_DEV uint32_t mbcnt64(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// outsideBoundingDomain (core/geometric_queries.h:376-379)
template <int D>
ZG_DEV bool outsideDomain(const float* dmin, const float* dmax, const Vec<D>& x) {
    bool o = false;
#pragma unroll
    for (int a = 0; a < D; a++) o = o || x[a] < dmin[a] || x[a] > dmax[a];
    return o;
}

// ------------------------------------------------------------------------------------
// Walk-on-stars queries on the 2D reflecting boundary
// ------------------------------------------------------------------------------------

// offsetPointAlongDirection (core/geometric_queries.h:175-207): the Ray Tracing Gems ch. 6
// offset, a few float ulps along n in integer space away from the origin, a fixed step near it.
template <int D>
ZG_DEV Vec<D> offsetPoint(const Vec<D>& p, const Vec<D>& n) {
    const float origin = 1.0f / 32.0f, floatScale = 1.0f / 65536.0f, intScale = 256.0f;
    Vec<D> r;
#pragma unroll
    for (int i = 0; i < D; i++) {
        int nOff = (int)(n[i] * intScale);
        float pOff = __int_as_float(__float_as_int(p[i]) + (p[i] < 0.0f ? -nOff : nOff));
        r[i] = fabsf(p[i]) < origin ? p[i] + floatScale * n[i] : pOff;
    }
    return r;
}

// Point where the ray o + t dir crosses segment i (the parametric point on the segment, as the
// reference reports hits).
ZG_DEV Vec<2> segmentHitPoint(const float* prims, int i, const Vec<2>& o, const Vec<2>& dir, float t) {
    Prim<2> pr; pr.load(prims, i);
    Vec<2> u = pr.a - o, v = pr.b - pr.a;
    float dv = cross2(dir, v);
    if (dv == 0.0f) return o + dir * t;
    float s = cross2(u, dir) / dv;
    return pr.a + v * s;
}

// Silhouette test of a 2D candidate (v, n0, n1, single) seen from x.  With d0, d1 the cosines
// between the adjacent segments' normals and the unit direction v -> x, a vertex whose
// neighbours lie on the same side of the line (x, v) is a silhouette: d0 * d1 < 0.  When x
// lies (within the precision) on the line of one adjacent segment, the vertex is a
// silhouette only when the other segment turns to the far side, sign * d_other > precision
// (a concave corner seen from the walker's side); sign = +1 for normals pointing away from
// the walker, -1 when the walker's normal orientation is flipped.  Open ends (one neighbour)
// and a query at the vertex itself are always silhouettes.
ZG_DEV bool isSilhouette2(const float4& c0, const float4& c1, const Vec<2>& x, float precision, bool flip) {
    if (c1.z != 0.0f) return true;
    Vec<2> view; view[0] = x[0] - c0.x; view[1] = x[1] - c0.y;
    float l = __builtin_sqrtf(norm2(view));
    if (l <= 0.0f) return true;
    view = view * (1.0f / l);
    float sign = flip ? -1.0f : 1.0f;
    float d0 = c0.z * view[0] + c0.w * view[1];
    float d1 = c1.x * view[0] + c1.y * view[1];
    if (fabsf(d0) <= precision) return sign * d1 > precision;
    if (fabsf(d1) <= precision) return sign * d0 > precision;
    return d0 * d1 < 0.0f;
}

// computeStarRadiusForReflectingBoundary (fcpw_geometric_queries.h:1254-1268): distance to
// the closest silhouette vertex within maxRadius, clamped below by minRadius; maxRadius
// clamped below by minRadius when there is none.  Nearest-first traversal of the
// silhouette-candidate BVH, stopped once a silhouette closer than minRadius is found.
// Normal-cone pruning: a child whose box is seen from x under directions that make an angle
// with every normal of its subtree (its cone, widened by the build's precision margin) away
// from 90 degrees can hold no silhouette and is skipped.  With u the direction from the box
// centre to x, sin(phi) = box radius / distance the angular radius of the box, and (a, theta)
// the cone, every normal-to-view angle lies within alpha +- (theta + phi), alpha = angle(a, u):
// the child is skipped when |cos alpha| > sin(theta + phi), theta + phi < 90 degrees.
template <int W>
ZG_DEV bool coneMaySeeSilhouette(const float4& cone, float lox, float loy, float hix, float hiy, const Vec<2>& x) {
    float cx = 0.5f * (lox + hix), cy = 0.5f * (loy + hiy);
    float ux = x[0] - cx, uy = x[1] - cy;
    float d2 = ux * ux + uy * uy;
    float ex = 0.5f * (hix - lox), ey = 0.5f * (hiy - loy);
    float r2 = ex * ex + ey * ey;
    if (d2 <= r2 * 1.0001f || cone.w < 0.0f) return true;   // x near or inside the box, or a wide cone
    float invd = 1.0f / __builtin_sqrtf(d2);
    float sphi = __builtin_sqrtf(r2) * invd * 1.0001f;
    float cphi = __builtin_sqrtf(fmaxf(0.0f, 1.0f - sphi * sphi));
    float sinSum = cone.z * cphi + cone.w * sphi;            // sin(theta + phi)
    float cosSum = cone.w * cphi - cone.z * sphi;            // cos(theta + phi)
    if (cosSum <= 0.0f) return true;
    float calpha = (cone.x * ux + cone.y * uy) * invd;
    return fabsf(calpha) <= sinSum * 1.0001f + 1e-6f;
}

template <int W>
ZG_DEV float starRadius2(const BoundaryView& b, const Vec<2>& x, float minRadius, float maxRadius, float precision,
                         bool flip, Stack& st) {
    if (minRadius > maxRadius) return maxRadius;
    if (b.nSilNodes == 0) return fmaxf(maxRadius, minRadius);
    float best2 = maxRadius < FLT_MAX ? maxRadius * maxRadius : FLT_MAX;
    const float min2 = minRadius * minRadius;
    const bool cones = b.silCones != nullptr && precision <= b.silConePrecision;
    bool found = false;
    st.sp = 0; st.ovf = 0;
    int node = 0;
    for (;;) {
        WideNode<2, W> nd;
        nd.load(b.silNodes, node);
        const float4* cn = reinterpret_cast<const float4*>(b.silCones) + (size_t)node * W;
        bool keep[W];
#pragma unroll
        for (int k = 0; k < W; k++) {
            keep[k] = nd.dist2(k, x) <= best2;
            if (cones && keep[k])
                keep[k] = coneMaySeeSilhouette<W>(cn[k], nd.lo[0][k], nd.lo[1][k], nd.hi[0][k], nd.hi[1][k], x);
        }
#pragma unroll
        for (int k = 0; k < W; k++) {
            if (!keep[k] || !isLeaf(nd.ch[k])) continue;
            int first = leafFirstD(nd.ch[k]), cnt = leafCountD(nd.ch[k]);
            for (int i = first; i < first + cnt; i++) {
                const float4* c = reinterpret_cast<const float4*>(b.sil) + 2 * i;
                float4 c0 = c[0], c1 = c[1];
                float dx = x[0] - c0.x, dy = x[1] - c0.y;
                float dd = dx * dx + dy * dy;
                if (dd < best2 && isSilhouette2(c0, c1, x, precision, flip)) { best2 = dd; found = true; }
            }
        }
        if (found && best2 <= min2) break;
        int next = -1; float ndist = 0.0f; int nk = -1;
#pragma unroll
        for (int k = 0; k < W; k++) {
            float dk = nd.dist2(k, x);
            if (keep[k] && !isLeaf(nd.ch[k]) && dk <= best2 && (next < 0 || dk < ndist)) { next = nd.ch[k]; ndist = dk; nk = k; }
        }
#pragma unroll
        for (int k = 0; k < W; k++) {
            float dk = nd.dist2(k, x);
            if (keep[k] && !isLeaf(nd.ch[k]) && dk <= best2 && k != nk) st.push(nd.ch[k], dk);
        }
        if (next < 0) {
            int n; float dn;
            bool got = false;
            while (st.pop(n, dn)) { if (dn <= best2) { got = true; break; } }
            if (!got) break;
            next = n;
        }
        node = next;
    }
    return found ? fmaxf(__builtin_sqrtf(best2), minRadius) : fmaxf(maxRadius, minRadius);
}

// Robin bound of the star radius (ReflectanceLineSegment::computeSquaredStarRadius with
// RobinLineSegmentBound, reflectance_boundary_bvh/geometry.h:86-140, robin_bounds.h:12-35): for
// every segment with kappa_max > 0 within sqrt(r2) of x, r2 shrinks to the square of
// (h / c) exp(c^2 / (h kappa_max)), h the distance to the segment's line and c = sqrt(h kappa_max
// / 2) clamped to the view-angle cosines of its closest and farthest points; inside that radius
// the Robin reflectance of the segment seen from x stays non-negative.  Pruned by box distance.
ZG_DEV void robinSegmentBound(const Prim<2>& pr, float kmax, const Vec<2>& x, float& r2) {
    if (!(kmax > FLT_EPSILON)) return;
    Vec<2> cp;
    float cd2 = pr.closest2(x, cp);
    if (cd2 > r2) return;
    float cd = __builtin_sqrtf(cd2);
    if (cd < FLT_EPSILON) { r2 = fminf(r2, cd); return; }
    Vec<2> n0 = pr.normal();
    float h = fabsf(dot(n0, pr.a - x));
    if (!(h > 0.0f)) { r2 = fminf(r2, cd2); return; }
    float da = __builtin_sqrtf(norm2(x - pr.a)), db = __builtin_sqrtf(norm2(x - pr.b));
    Vec<2> fp = da >= db ? pr.a : pr.b;
    float fd = fmaxf(da, db);
    float cosUpper = fabsf(dot(x - cp, n0)) / cd;
    float cosLower = fabsf(dot(x - fp, n0)) / fd;
    float cosLine = __builtin_sqrtf(h * kmax) * 0.70710678f;
    float c = fminf(fmaxf(cosLine, cosLower), cosUpper);
    float rad = (h / c) * __expf(c * c / (h * kmax));
    r2 = fminf(r2, rad * rad);
}

// Robin bound of a triangle (ReflectanceTriangle::computeSquaredStarRadius with
// RobinTriangleBound, reflectance_boundary_bvh/geometry.h:170-236, robin_bounds.h:52-80): h the
// distance to the triangle's plane, cosUpper / cosLower the view-angle cosines of its closest
// point and farthest vertex; when sqrt(h kappa_max) >= cosLower, c = sqrt(h kappa_max / 3)
// clamped to [cosLower, cosUpper] gives the radius h^2 kappa_max / (c (h kappa_max - c^2)),
// inside which the 3D reflectance of every point of the triangle seen from x stays >= 0.
ZG_DEV void robinTriangleBound(const Prim<3>& pr, float kmax, const Vec<3>& x, float& r2) {
    if (!(kmax > FLT_EPSILON)) return;
    Vec<3> cp;
    float cd2 = pr.closest2(x, cp);
    if (cd2 > r2) return;
    float cd = __builtin_sqrtf(cd2);
    if (cd < FLT_EPSILON) { r2 = fminf(r2, cd); return; }
    Vec<3> n0 = pr.normal();
    float h = fabsf(dot(n0, pr.a - x));
    if (!(h > 0.0f)) { r2 = fminf(r2, cd2); return; }
    float da = __builtin_sqrtf(norm2(x - pr.a)), db = __builtin_sqrtf(norm2(x - pr.b)), dc = __builtin_sqrtf(norm2(x - pr.c));
    Vec<3> fp = pr.a;
    float fd = da;
    if (db > fd) { fp = pr.b; fd = db; }
    if (dc > fd) { fp = pr.c; fd = dc; }
    float cosUpper = fabsf(dot(x - cp, n0)) / cd;
    float cosLower = fabsf(dot(x - fp, n0)) / fd;
    float maxCos = __builtin_sqrtf(h * kmax);
    if (maxCos < cosLower) return;
    float c = fminf(fmaxf(maxCos * 0.57735027f, cosLower), cosUpper);
    float rad = h * h * kmax / (c * (h * kmax - c * c));
    r2 = fminf(r2, rad * rad);
}

// minimum of the Robin bounds of the reflecting primitives within sqrt(r2) of x
template <int D, int W>
ZG_DEV float robinRadius2(const BoundaryView& b, const Vec<D>& x, float r2, Stack& st, VisitCount* vc = nullptr) {
    if (b.nNodes == 0 || !b.robinMax) return r2;
    st.sp = 0; st.ovf = 0;
    int node = 0;
    for (;;) {
        WideNode<D, W> nd;
        nd.load(b.nodes, node);
        if (vc) vc->nodes++;
#pragma unroll
        for (int k = 0; k < W; k++) {
            if (nd.dist2(k, x) > r2) continue;
            int c = nd.ch[k];
            if (isLeaf(c)) {
                int first = leafFirstD(c), cnt = leafCountD(c);
                if (vc) vc->prims += cnt;
                for (int j = first; j < first + cnt; j++) {
                    Prim<D> pr; pr.load(b.prims, j);
                    if constexpr (D == 2) robinSegmentBound(pr, b.robinMax[j], x, r2);
                    else robinTriangleBound(pr, b.robinMax[j], x, r2);
                }
            } else {
                st.push(c, nd.dist2(k, x));
            }
        }
        int n; float dn;
        bool got = false;
        while (st.pop(n, dn)) { if (dn <= r2) { got = true; break; } }
        if (!got) break;
        node = n;
    }
    return r2;
}

// hasLineOfSight: true unless a segment of b crosses the open segment p1 -> p2.
template <int W>
ZG_DEV bool lineOfSight2(const BoundaryView& b, const Vec<2>& p1, const Vec<2>& p2, Stack& st) {
    if (b.nNodes == 0) return true;
    Vec<2> d = p2 - p1;
    float L = __builtin_sqrtf(norm2(d));
    if (L <= 0.0f) return true;
    Vec<2> dir = d * (1.0f / L);
    Vec<2> inv; inv[0] = 1.0f / dir[0]; inv[1] = 1.0f / dir[1];
    st.sp = 0; st.ovf = 0;
    int node = 0;
    for (;;) {
        WideNode<2, W> nd;
        nd.load(b.nodes, node);
#pragma unroll
        for (int k = 0; k < W; k++) {
            float t0 = 0.0f, t1 = L;
#pragma unroll
            for (int a = 0; a < 2; a++) {
                float ta = (nd.lo[a][k] - p1[a]) * inv[a], tb = (nd.hi[a][k] - p1[a]) * inv[a];
                t0 = fmaxf(t0, fminf(ta, tb));
                t1 = fminf(t1, fmaxf(ta, tb));
            }
            if (t0 > t1) continue;
            if (isLeaf(nd.ch[k])) {
                int first = leafFirstD(nd.ch[k]), cnt = leafCountD(nd.ch[k]);
                for (int i = first; i < first + cnt; i++) {
                    Prim<2> pr; pr.load(b.prims, i);
                    float t;
                    if (pr.ray(p1, dir, L, t) && t > 0.0f && t < L) return false;
                }
            } else {
                st.push(nd.ch[k], t0);
            }
        }
        int n; float tn;
        if (!st.pop(n, tn)) break;
        node = n;
    }
    return true;
}

// branchTraversalWeight (fcpw_geometric_queries.h:1391-1398): |G| of the 3D free-space
// harmonic Green's function at max(r, 1e-2).
ZG_DEV float branchWeight(float r2) {
    float r = fmaxf(__builtin_sqrtf(r2), 1e-2f);
    return 1.0f / (4.0f * kPi * r);
}

// sampleReflectingBoundary (fcpw_geometric_queries.h:1227-1242): one stochastic descent of
// the wide BVH.  At an inner node a child whose box overlaps the ball |y - x| <= radius is
// chosen with probability proportional to branchWeight(|box centre - x|^2); in the leaf, a
// segment whose box overlaps the ball with probability proportional to branchWeight x length;
// then a uniform point on it.  u0 is rescaled at every level; pdf is the exact density of the
// sample.  Every segment whose box overlaps the ball has positive density, so the estimator
// alpha G h / pdf is unbiased.
template <int W>
ZG_DEV bool sampleReflecting2(const BoundaryView& b, const Vec<2>& x, float radius, float u0, float u1,
                              Vec<2>& pt, Vec<2>& nrm, float& pdf, VisitCount* vc = nullptr) {
    if (b.nNodes == 0) return false;
    const float R2 = radius * radius;
    float prob = 1.0f;
    float u = u0;
    int node = 0;
    for (int depth = 0; depth < 64; depth++) {
        WideNode<2, W> nd;
        nd.load(b.nodes, node);
        if (vc) vc->nodes++;
        float w[W];
        float total = 0.0f;
#pragma unroll
        for (int k = 0; k < W; k++) {
            float cx = 0.5f * (nd.lo[0][k] + nd.hi[0][k]) - x[0];
            float cy = 0.5f * (nd.lo[1][k] + nd.hi[1][k]) - x[1];
            w[k] = nd.dist2(k, x) <= R2 ? branchWeight(cx * cx + cy * cy) : 0.0f;
            total += w[k];
        }
        if (!(total > 0.0f)) return false;
        // inverse CDF: the last overlapping child whose interval starts at or below t
        float t = u * total, acc = 0.0f;
        int sel = -1;
        float wSel = 0.0f, accSel = 0.0f;
#pragma unroll
        for (int k = 0; k < W; k++) {
            if (w[k] > 0.0f) {
                if (sel < 0 || t >= acc) { sel = k; wSel = w[k]; accSel = acc; }
                acc += w[k];
            }
        }
        prob *= wSel / total;
        u = fminf(fmaxf((t - accSel) / wSel, 0.0f), 0.99999994f);
        int c = nd.ch[sel];
        if (!isLeaf(c)) { node = c; continue; }
        int first = leafFirstD(c), cnt = leafCountD(c);
        if (vc) vc->prims += cnt;
        float lw[16];
        float ltotal = 0.0f;
        for (int j = 0; j < cnt; j++) {
            Prim<2> pr; pr.load(b.prims, first + j);
            float lox = fminf(pr.a[0], pr.b[0]), hix = fmaxf(pr.a[0], pr.b[0]);
            float loy = fminf(pr.a[1], pr.b[1]), hiy = fmaxf(pr.a[1], pr.b[1]);
            float dx = fmaxf(fmaxf(lox - x[0], x[0] - hix), 0.0f), dy = fmaxf(fmaxf(loy - x[1], x[1] - hiy), 0.0f);
            float mx = 0.5f * (lox + hix) - x[0], my = 0.5f * (loy + hiy) - x[1];
            float len = __builtin_sqrtf(norm2(pr.b - pr.a));
            lw[j] = dx * dx + dy * dy <= R2 ? branchWeight(mx * mx + my * my) * len : 0.0f;
            ltotal += lw[j];
        }
        if (!(ltotal > 0.0f)) return false;
        float lt = u * ltotal, lacc = 0.0f;
        int ps = -1;
        for (int j = 0; j < cnt; j++) {
            if (lw[j] > 0.0f && (ps < 0 || lt >= lacc)) ps = j;
            lacc += lw[j];
        }
        Prim<2> pr; pr.load(b.prims, first + ps);
        float len = __builtin_sqrtf(norm2(pr.b - pr.a));
        pt = pr.a + (pr.b - pr.a) * u1;
        nrm = pr.normal();
        pdf = prob * (lw[ps] / ltotal) / len;
        return true;
    }
    return false;
}

// ------------------------------------------------------------------------------------
// Walk-on-stars queries on the 3D reflecting boundary (triangle meshes)
// ------------------------------------------------------------------------------------

ZG_DEV float closestOnSegment3(const Vec<3>& pa, const Vec<3>& pb, const Vec<3>& x, Vec<3>& pt) {
    Vec<3> u = pb - pa, v = x - pa;
    float c1 = dot(u, v);
    float c2 = dot(u, u);
    pt = c1 <= 0.0f ? pa : (c2 <= c1 ? pb : pa + u * (c1 / c2));
    return norm2(x - pt);
}

// Silhouette test of a 3D candidate edge seen from x: the vertex rule of isSilhouette2 with
// the unit direction from the edge's closest point cp to x and the normals of the edge's two
// triangles (both planes contain the edge line).  Open edges (one triangle) and a query on
// the edge are always silhouettes.
ZG_DEV bool isSilhouette3(const float4& c0, const float4& c2, const float4& c3, const Vec<3>& x, const Vec<3>& cp,
                          float precision, bool flip) {
    if (c0.w != 0.0f) return true;
    Vec<3> view = x - cp;
    float l = __builtin_sqrtf(norm2(view));
    if (l <= 0.0f) return true;
    view = view * (1.0f / l);
    float sign = flip ? -1.0f : 1.0f;
    float d0 = c2.x * view[0] + c2.y * view[1] + c2.z * view[2];
    float d1 = c3.x * view[0] + c3.y * view[1] + c3.z * view[2];
    if (fabsf(d0) <= precision) return sign * d1 > precision;
    if (fabsf(d1) <= precision) return sign * d0 > precision;
    return d0 * d1 < 0.0f;
}

// whether a 3D silhouette-BVH child may hold a silhouette seen from x: the rule of
// coneMaySeeSilhouette on the sphere of directions (cone c0 = (axis.xyz, sin), c1.x = cos of
// the widened half-angle; a negative cos never prunes).  A coplanar child (c1.z = 1, plane
// axis . y = c1.y within tolerance c1.w, normals within 1e-5 rad of the axis) is also skipped
// when every point y of its box sees x with |n . v| = h / |x - y| beyond precision + 1e-4 (both
// triangles of each edge face x the same way) or within precision - 1e-4 (the tie rule then
// reports no silhouette), h the distance of x to the plane.
ZG_DEV bool coneMaySeeSilhouette3(const float4& c0, const float4& c1, const float* lo, const float* hi, const Vec<3>& x,
                                  float precision) {
    if (c1.x < 0.0f) return true;
    if (c1.z != 0.0f) {
        const float h = fabsf(c0.x * x[0] + c0.y * x[1] + c0.z * x[2] - c1.y);
        float n2 = 0.0f, f2 = 0.0f;
#pragma unroll
        for (int a = 0; a < 3; a++) {
            const float dn = fmaxf(fmaxf(lo[a] - x[a], x[a] - hi[a]), 0.0f);
            const float df = fmaxf(fabsf(x[a] - lo[a]), fabsf(x[a] - hi[a]));
            n2 += dn * dn; f2 += df * df;
        }
        const float nearD = __builtin_sqrtf(n2) * 0.9999f, farD = __builtin_sqrtf(f2) * 1.0001f;
        if (h - c1.w > (precision + 1e-4f) * farD) return false;
        if (nearD > 0.0f && h + c1.w < (precision - 1e-4f) * nearD) return false;
    }
    float ux = x[0] - 0.5f * (lo[0] + hi[0]), uy = x[1] - 0.5f * (lo[1] + hi[1]), uz = x[2] - 0.5f * (lo[2] + hi[2]);
    float ex = 0.5f * (hi[0] - lo[0]), ey = 0.5f * (hi[1] - lo[1]), ez = 0.5f * (hi[2] - lo[2]);
    float d2 = ux * ux + uy * uy + uz * uz;
    float r2 = ex * ex + ey * ey + ez * ez;
    if (d2 <= r2 * 1.0001f) return true;
    float invd = 1.0f / __builtin_sqrtf(d2);
    float sphi = __builtin_sqrtf(r2) * invd * 1.0001f;
    float cphi = __builtin_sqrtf(fmaxf(0.0f, 1.0f - sphi * sphi));
    float sinSum = c0.w * cphi + c1.x * sphi;
    float cosSum = c1.x * cphi - c0.w * sphi;
    if (cosSum <= 0.0f) return true;
    float calpha = (c0.x * ux + c0.y * uy + c0.z * uz) * invd;
    return fabsf(calpha) <= sinSum * 1.0001f + 1e-6f;
}

// computeStarRadiusForReflectingBoundary in 3D: distance to the closest point of the nearest
// silhouette edge within maxRadius, clamped below by minRadius (maxRadius clamped below by
// minRadius when there is none).  Nearest-first traversal of the candidate-edge BVH, pruned by
// box distance and by the children's normal cones, stopped once a silhouette closer than
// minRadius is found.
template <int W>
ZG_DEV float starRadius3(const BoundaryView& b, const Vec<3>& x, float minRadius, float maxRadius, float precision,
                         bool flip, Stack& st) {
    if (minRadius > maxRadius) return maxRadius;
    if (b.nSilNodes == 0) return fmaxf(maxRadius, minRadius);
    float best2 = maxRadius < FLT_MAX ? maxRadius * maxRadius : FLT_MAX;
    const float min2 = minRadius * minRadius;
    const bool cones = b.silCones != nullptr && precision <= b.silConePrecision;
    bool found = false;
    st.sp = 0; st.ovf = 0;
    int node = 0;
    for (;;) {
        WideNode<3, W> nd;
        nd.load(b.silNodes, node);
        const float4* cn = reinterpret_cast<const float4*>(b.silCones) + (size_t)node * W * 2;
        bool keep[W];
#pragma unroll
        for (int k = 0; k < W; k++) {
            keep[k] = nd.dist2(k, x) <= best2;
            if (cones && keep[k]) {
                const float lo[3] = {nd.lo[0][k], nd.lo[1][k], nd.lo[2][k]};
                const float hi[3] = {nd.hi[0][k], nd.hi[1][k], nd.hi[2][k]};
                keep[k] = coneMaySeeSilhouette3(cn[2 * k], cn[2 * k + 1], lo, hi, x, precision);
            }
        }
#pragma unroll
        for (int k = 0; k < W; k++) {
            if (!keep[k] || !isLeaf(nd.ch[k])) continue;
            int first = leafFirstD(nd.ch[k]), cnt = leafCountD(nd.ch[k]);
            for (int i = first; i < first + cnt; i++) {
                const float4* c = reinterpret_cast<const float4*>(b.sil) + 4 * i;
                float4 c0 = c[0], c1 = c[1];
                Vec<3> pa, pb, cp;
                pa[0] = c0.x; pa[1] = c0.y; pa[2] = c0.z;
                pb[0] = c1.x; pb[1] = c1.y; pb[2] = c1.z;
                float dd = closestOnSegment3(pa, pb, x, cp);
                if (dd < best2 && isSilhouette3(c0, c[2], c[3], x, cp, precision, flip)) { best2 = dd; found = true; }
            }
        }
        if (found && best2 <= min2) break;
        int next = -1; float ndist = 0.0f; int nk = -1;
#pragma unroll
        for (int k = 0; k < W; k++) {
            float dk = nd.dist2(k, x);
            if (keep[k] && !isLeaf(nd.ch[k]) && dk <= best2 && (next < 0 || dk < ndist)) { next = nd.ch[k]; ndist = dk; nk = k; }
        }
#pragma unroll
        for (int k = 0; k < W; k++) {
            float dk = nd.dist2(k, x);
            if (keep[k] && !isLeaf(nd.ch[k]) && dk <= best2 && k != nk) st.push(nd.ch[k], dk);
        }
        if (next < 0) {
            int n; float dn;
            bool got = false;
            while (st.pop(n, dn)) { if (dn <= best2) { got = true; break; } }
            if (!got) break;
            next = n;
        }
        node = next;
    }
    return found ? fmaxf(__builtin_sqrtf(best2), minRadius) : fmaxf(maxRadius, minRadius);
}

// hasLineOfSight in 3D: true unless a triangle of b crosses the open segment p1 -> p2.
template <int W>
ZG_DEV bool lineOfSight3(const BoundaryView& b, const Vec<3>& p1, const Vec<3>& p2, Stack& st) {
    if (b.nNodes == 0) return true;
    Vec<3> d = p2 - p1;
    float L = __builtin_sqrtf(norm2(d));
    if (L <= 0.0f) return true;
    Vec<3> dir = d * (1.0f / L);
    Vec<3> inv;
#pragma unroll
    for (int a = 0; a < 3; a++) inv[a] = 1.0f / dir[a];
    st.sp = 0; st.ovf = 0;
    int node = 0;
    for (;;) {
        WideNode<3, W> nd;
        nd.load(b.nodes, node);
#pragma unroll
        for (int k = 0; k < W; k++) {
            float t0 = 0.0f, t1 = L;
#pragma unroll
            for (int a = 0; a < 3; a++) {
                float ta = (nd.lo[a][k] - p1[a]) * inv[a], tb = (nd.hi[a][k] - p1[a]) * inv[a];
                t0 = fmaxf(t0, fminf(ta, tb));
                t1 = fminf(t1, fmaxf(ta, tb));
            }
            if (t0 > t1) continue;
            if (isLeaf(nd.ch[k])) {
                int first = leafFirstD(nd.ch[k]), cnt = leafCountD(nd.ch[k]);
                for (int i = first; i < first + cnt; i++) {
                    Prim<3> pr; pr.load(b.prims, i);
                    float t;
                    if (pr.ray(p1, dir, L, t) && t > 0.0f && t < L) return false;
                }
            } else {
                st.push(nd.ch[k], t0);
            }
        }
        int n; float tn;
        if (!st.pop(n, tn)) break;
        node = n;
    }
    return true;
}

// sampleReflectingBoundary in 3D: the stochastic descent of sampleReflecting2 over the
// triangle BVH (children and leaf triangles whose boxes overlap the ball, weighted by
// branchWeight of the box-centre distance, triangles also by their area), then a uniform
// point on the triangle from (u1, u2) by the square-root warp.  pdf is the exact density.
template <int W>
ZG_DEV bool sampleReflecting3(const BoundaryView& b, const Vec<3>& x, float radius, float u0, float u1, float u2,
                              Vec<3>& pt, Vec<3>& nrm, float& pdf, VisitCount* vc = nullptr) {
    if (b.nNodes == 0) return false;
    const float R2 = radius * radius;
    float prob = 1.0f;
    float u = u0;
    int node = 0;
    for (int depth = 0; depth < 64; depth++) {
        WideNode<3, W> nd;
        nd.load(b.nodes, node);
        if (vc) vc->nodes++;
        float w[W];
        float total = 0.0f;
#pragma unroll
        for (int k = 0; k < W; k++) {
            float cx = 0.5f * (nd.lo[0][k] + nd.hi[0][k]) - x[0];
            float cy = 0.5f * (nd.lo[1][k] + nd.hi[1][k]) - x[1];
            float cz = 0.5f * (nd.lo[2][k] + nd.hi[2][k]) - x[2];
            w[k] = nd.dist2(k, x) <= R2 ? branchWeight(cx * cx + cy * cy + cz * cz) : 0.0f;
            total += w[k];
        }
        if (!(total > 0.0f)) return false;
        float t = u * total, acc = 0.0f;
        int sel = -1;
        float wSel = 0.0f, accSel = 0.0f;
#pragma unroll
        for (int k = 0; k < W; k++) {
            if (w[k] > 0.0f) {
                if (sel < 0 || t >= acc) { sel = k; wSel = w[k]; accSel = acc; }
                acc += w[k];
            }
        }
        prob *= wSel / total;
        u = fminf(fmaxf((t - accSel) / wSel, 0.0f), 0.99999994f);
        int c = nd.ch[sel];
        if (!isLeaf(c)) { node = c; continue; }
        int first = leafFirstD(c), cnt = leafCountD(c);
        if (vc) vc->prims += cnt;
        float lw[16];
        float ltotal = 0.0f;
        for (int j = 0; j < cnt; j++) {
            Prim<3> pr; pr.load(b.prims, first + j);
            float d2 = 0.0f, m2 = 0.0f;
#pragma unroll
            for (int a = 0; a < 3; a++) {
                float lo = fminf(fminf(pr.a[a], pr.b[a]), pr.c[a]), hi = fmaxf(fmaxf(pr.a[a], pr.b[a]), pr.c[a]);
                float dd = fmaxf(fmaxf(lo - x[a], x[a] - hi), 0.0f);
                float mm = 0.5f * (lo + hi) - x[a];
                d2 += dd * dd; m2 += mm * mm;
            }
            float area = 0.5f * __builtin_sqrtf(norm2(cross3(pr.b - pr.a, pr.c - pr.a)));
            lw[j] = d2 <= R2 ? branchWeight(m2) * area : 0.0f;
            ltotal += lw[j];
        }
        if (!(ltotal > 0.0f)) return false;
        float lt = u * ltotal, lacc = 0.0f;
        int ps = -1;
        for (int j = 0; j < cnt; j++) {
            if (lw[j] > 0.0f && (ps < 0 || lt >= lacc)) ps = j;
            lacc += lw[j];
        }
        Prim<3> pr; pr.load(b.prims, first + ps);
        float area = 0.5f * __builtin_sqrtf(norm2(cross3(pr.b - pr.a, pr.c - pr.a)));
        float su = __builtin_sqrtf(u1);
        pt = pr.a + (pr.b - pr.a) * (su * (1.0f - u2)) + (pr.c - pr.a) * (su * u2);
        nrm = pr.normal();
        pdf = prob * (lw[ps] / ltotal) / area;
        return true;
    }
    return false;
}

}  // namespace zg
