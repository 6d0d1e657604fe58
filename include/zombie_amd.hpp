/* zombie_amd.hpp -- header-only C++17 adapter with the reference's solver signatures over the
 * C-ABI of zombie_amd.h.
 *
 * The reference's C++ callers construct WalkOnSpheres<T, DIM> / WalkOnStars<T, DIM> from a
 * GeometricQueries<DIM> and call
 *     solve(pde, walkSettings, nWalks, samplePts, sampleStatistics, runSingleThreaded = false,
 *           reportProgress = {})
 * (point_estimation/walk_on_spheres.h:31-44, walk_on_stars.h:33-46).  This header keeps those
 * class names, constructor arguments, member names and the solve() overloads in namespace
 * zombie_amd, so a caller switches backends by changing the namespace and the scene/PDE set-up:
 *   - GeometricQueries<DIM> owns the device scene (zg_scene) built from the boundary mesh, the
 *     reflecting flags and the bounding domain (what the reference's FCPW handlers and
 *     populateGeometricQueriesFor*Boundary produce);
 *   - PDE<T, DIM> holds a zg_pde: device-evaluable data (constants, manufactured fields, dense
 *     grids) instead of std::function closures, which cannot run on the GPU;
 *   - WalkSettings, SamplePoint and SampleStatistics keep the reference's fields
 *     (core/common.h:19-395); statistics accumulate across calls (Chan merge in FP64);
 *   - errors throw std::runtime_error with the library's message (the reference exits);
 *   - the solvers take an explicit seed (the reference seeds from the clock) and successive
 *     solve() calls use fresh walk streams; runSingleThreaded and reportProgress are accepted
 *     and ignored (the GPU runs every point at once; progress is reported once at the end).
 * Only T = float is provided (every benchmark configuration is scalar).
 * Links against libzombie_amd.so; needs no HIP headers (device memory through zg_malloc).
 */
#pragma once

#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <limits>
#include <stdexcept>
#include <string>
#include <vector>

#include "zombie_amd.h"

namespace zombie_amd {

template <size_t DIM>
using Vector = std::array<float, DIM>;
template <size_t DIM>
using Vectori = std::array<int32_t, DIM>;

inline void check(zg_status s) {
    if (s != ZG_OK) {
        char buf[512];
        zg_last_error(buf, (int32_t)sizeof(buf));
        throw std::runtime_error(std::string("zombie_amd: ") + buf);
    }
}

// core/common.h:132-143
enum class WalkCompletionCode { ReachedAbsorbingBoundary, TerminatedWithRussianRoulette, ExceededMaxWalkLength,
                                EscapedDomain };
enum class SampleType { InDomain = ZG_SAMPLE_IN_DOMAIN, OnAbsorbingBoundary = ZG_SAMPLE_ON_ABSORBING_BOUNDARY,
                        OnReflectingBoundary = ZG_SAMPLE_ON_REFLECTING_BOUNDARY };
enum class EstimationQuantity { Solution = ZG_ESTIMATE_SOLUTION,
                                SolutionAndGradient = ZG_ESTIMATE_SOLUTION_AND_GRADIENT, None = ZG_ESTIMATE_SKIP };

// core/common.h:19-90: the two constructors and the public fields
struct WalkSettings {
    WalkSettings(float epsilonShellForAbsorbingBoundary_, float epsilonShellForReflectingBoundary_,
                 int maxWalkLength_, bool solveDoubleSided_)
        : WalkSettings(epsilonShellForAbsorbingBoundary_, epsilonShellForReflectingBoundary_, 1e-3f, 0.0f,
                       std::numeric_limits<float>::max(), maxWalkLength_, 0, maxWalkLength_, solveDoubleSided_,
                       true, true, false, false, false, false, false) {}
    WalkSettings(float epsilonShellForAbsorbingBoundary_, float epsilonShellForReflectingBoundary_,
                 float silhouettePrecision_, float russianRouletteThreshold_, float splittingThreshold_,
                 int maxWalkLength_, int stepsBeforeApplyingTikhonov_, int stepsBeforeUsingMaximalSpheres_,
                 bool solveDoubleSided_, bool useGradientControlVariates_, bool useGradientAntitheticVariates_,
                 bool useCosineSamplingForDerivatives_, bool ignoreAbsorbingBoundaryContribution_,
                 bool ignoreReflectingBoundaryContribution_, bool ignoreSourceContribution_, bool printLogs_)
        : epsilonShellForAbsorbingBoundary(epsilonShellForAbsorbingBoundary_),
          epsilonShellForReflectingBoundary(epsilonShellForReflectingBoundary_),
          silhouettePrecision(silhouettePrecision_), russianRouletteThreshold(russianRouletteThreshold_),
          splittingThreshold(splittingThreshold_), maxWalkLength(maxWalkLength_),
          stepsBeforeApplyingTikhonov(stepsBeforeApplyingTikhonov_),
          stepsBeforeUsingMaximalSpheres(stepsBeforeUsingMaximalSpheres_), solveDoubleSided(solveDoubleSided_),
          useGradientControlVariates(useGradientControlVariates_),
          useGradientAntitheticVariates(useGradientAntitheticVariates_),
          useCosineSamplingForDerivatives(useCosineSamplingForDerivatives_),
          ignoreAbsorbingBoundaryContribution(ignoreAbsorbingBoundaryContribution_),
          ignoreReflectingBoundaryContribution(ignoreReflectingBoundaryContribution_),
          ignoreSourceContribution(ignoreSourceContribution_), printLogs(printLogs_) {}

    zg_walk_settings toC() const {
        zg_walk_settings s;
        s.epsilon_shell_for_absorbing_boundary = epsilonShellForAbsorbingBoundary;
        s.epsilon_shell_for_reflecting_boundary = epsilonShellForReflectingBoundary;
        s.silhouette_precision = silhouettePrecision;
        s.russian_roulette_threshold = russianRouletteThreshold;
        s.splitting_threshold = splittingThreshold;
        s.max_walk_length = maxWalkLength;
        s.steps_before_applying_tikhonov = stepsBeforeApplyingTikhonov;
        s.steps_before_using_maximal_spheres = stepsBeforeUsingMaximalSpheres;
        s.solve_double_sided = solveDoubleSided;
        s.use_gradient_control_variates = useGradientControlVariates;
        s.use_gradient_antithetic_variates = useGradientAntitheticVariates;
        s.use_cosine_sampling_for_derivatives = useCosineSamplingForDerivatives;
        s.ignore_absorbing_boundary_contribution = ignoreAbsorbingBoundaryContribution;
        s.ignore_reflecting_boundary_contribution = ignoreReflectingBoundaryContribution;
        s.ignore_source_contribution = ignoreSourceContribution;
        s.print_logs = printLogs;
        return s;
    }

    float epsilonShellForAbsorbingBoundary, epsilonShellForReflectingBoundary, silhouettePrecision;
    float russianRouletteThreshold, splittingThreshold;
    int maxWalkLength, stepsBeforeApplyingTikhonov, stepsBeforeUsingMaximalSpheres;
    bool solveDoubleSided, useGradientControlVariates, useGradientAntitheticVariates;
    bool useCosineSamplingForDerivatives, ignoreAbsorbingBoundaryContribution;
    bool ignoreReflectingBoundaryContribution, ignoreSourceContribution, printLogs;
};

// core/common.h:145-195
template <typename T, size_t DIM>
struct SamplePoint {
    SamplePoint(const Vector<DIM>& pt_, const Vector<DIM>& normal_, SampleType type_,
                EstimationQuantity estimationQuantity_, float pdf_, float distToAbsorbingBoundary_,
                float distToReflectingBoundary_)
        : pt(pt_), normal(normal_), type(type_), estimationQuantity(estimationQuantity_), pdf(pdf_),
          distToAbsorbingBoundary(distToAbsorbingBoundary_), distToReflectingBoundary(distToReflectingBoundary_) {
        directionForDerivative.fill(0.0f);
        directionForDerivative[0] = 1.0f;
    }
    void reset() { firstSphereRadius = 0.0f; }

    Vector<DIM> pt, normal, directionForDerivative;
    SampleType type;
    EstimationQuantity estimationQuantity;
    float pdf, distToAbsorbingBoundary, distToReflectingBoundary;
    float firstSphereRadius = 0.0f;
    bool estimateBoundaryNormalAligned = false;
};

// core/common.h:200-395 (FP64 running mean and M2)
template <typename T, size_t DIM>
class SampleStatistics {
public:
    T getEstimatedSolution() const { return (T)solutionMean; }
    T getEstimatedSolutionVariance() const { return nSolution > 1 ? (T)(solutionM2 / (double)(nSolution - 1)) : T(0); }
    Vector<DIM> getEstimatedGradient() const {
        Vector<DIM> g;
        for (size_t i = 0; i < DIM; i++) g[i] = (float)gradientMean[i];
        return g;
    }
    T getEstimatedDerivative() const { return nGradient > 0 ? (T)(derivativeSum / (double)nGradient) : T(0); }
    int getSolutionEstimateCount() const { return (int)nSolution; }
    int getGradientEstimateCount() const { return (int)nGradient; }
    float getMeanWalkLength() const { return nSolution > 0 ? (float)((double)walkLength / (double)nSolution) : 0.0f; }

    int64_t nSolution = 0, walkLength = 0, splits = 0, nGradient = 0;
    double solutionMean = 0.0, solutionM2 = 0.0, firstSourceSum = 0.0, derivativeSum = 0.0;
    std::array<double, DIM> gradientMean{}, gradientM2{};
};

// device buffer (RAII over zg_malloc / zg_free)
class DeviceBuffer {
public:
    DeviceBuffer() = default;
    explicit DeviceBuffer(size_t bytes) { resize(bytes); }
    ~DeviceBuffer() { if (p_) zg_free(p_); }
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    void resize(size_t bytes) {
        if (bytes <= bytes_) return;
        if (p_) { zg_free(p_); p_ = nullptr; bytes_ = 0; }
        check(zg_malloc(&p_, bytes));
        bytes_ = bytes;
    }
    template <typename U> U* as() const { return static_cast<U*>(p_); }
    template <typename U> void upload(const std::vector<U>& v) {
        resize(std::max<size_t>(v.size() * sizeof(U), 1));
        if (!v.empty()) check(zg_memcpy_htod(p_, v.data(), v.size() * sizeof(U)));
    }
    template <typename U> void download(std::vector<U>& v) const {
        if (!v.empty()) check(zg_memcpy_dtoh(v.data(), p_, v.size() * sizeof(U)));
    }

private:
    void* p_ = nullptr;
    size_t bytes_ = 0;
};

// GeometricQueries<DIM> (core/geometric_queries.h:59-125) backed by a device scene
template <size_t DIM>
class GeometricQueries {
public:
    GeometricQueries(const std::vector<Vector<DIM>>& positions, const std::vector<Vectori<DIM>>& indices,
                     const std::vector<uint8_t>& isReflecting, const Vector<DIM>& domainMin_,
                     const Vector<DIM>& domainMax_, bool solveDoubleSided = false, float silhouettePrecision = 1e-3f,
                     int bvhWidth = 4, int leafSize = 4)
        : domainMin(domainMin_), domainMax(domainMax_) {
        zg_build_options o{bvhWidth, leafSize, 16, 0};
        check(zg_scene_create((int32_t)DIM, positions.empty() ? nullptr : positions[0].data(),
                              (int64_t)positions.size(), indices.empty() ? nullptr : indices[0].data(),
                              (int64_t)indices.size(), isReflecting.empty() ? nullptr : isReflecting.data(),
                              domainMin.data(), domainMax.data(), solveDoubleSided ? 1 : 0, silhouettePrecision, &o,
                              &scene_));
    }
    ~GeometricQueries() { if (scene_) zg_scene_destroy(scene_); }
    GeometricQueries(const GeometricQueries&) = delete;
    GeometricQueries& operator=(const GeometricQueries&) = delete;
    const zg_scene* scene() const { return scene_; }

    // computeDistToAbsorbingBoundary / computeDistToReflectingBoundary for many points
    std::vector<float> computeDistToBoundary(const std::vector<Vector<DIM>>& pts, bool reflecting) const {
        std::vector<float> d(pts.size());
        if (pts.empty()) return d;
        DeviceBuffer p, out(pts.size() * sizeof(float));
        p.upload(pts);
        check(zg_closest_point(scene_, reflecting ? 1 : 0, (int64_t)pts.size(), p.as<float>(), out.as<float>(),
                               nullptr, nullptr, nullptr, nullptr));
        check(zg_synchronize());
        out.download(d);
        return d;
    }

    Vector<DIM> domainMin, domainMax;

private:
    zg_scene* scene_ = nullptr;
};

// PDE<T, DIM> (core/pde.h:19-79): device-evaluable descriptors
template <typename T, size_t DIM>
struct PDE {
    zg_pde desc{};
};

template <typename T, size_t DIM, bool WOST>
class SolverBase {
public:
    explicit SolverBase(const GeometricQueries<DIM>& queries_, uint64_t seed_ = 0x5EED)
        : queries(queries_), seed(seed_) {}

    // single point (walk_on_spheres.h:31-34)
    void solve(const PDE<T, DIM>& pde, const WalkSettings& walkSettings, int nWalks, SamplePoint<T, DIM>& samplePt,
               SampleStatistics<T, DIM>& statistics) const {
        std::vector<SamplePoint<T, DIM>> pts{samplePt};
        std::vector<SampleStatistics<T, DIM>> st{statistics};
        solve(pde, walkSettings, std::vector<int>{nWalks}, pts, st);
        samplePt = pts[0];
        statistics = st[0];
    }

    // batch (walk_on_spheres.h:38-44)
    void solve(const PDE<T, DIM>& pde, const WalkSettings& walkSettings, const std::vector<int>& nWalks,
               std::vector<SamplePoint<T, DIM>>& samplePts, std::vector<SampleStatistics<T, DIM>>& sampleStatistics,
               bool runSingleThreaded = false, std::function<void(int, int)> reportProgress = {}) const {
        (void)runSingleThreaded;
        const size_t n = samplePts.size();
        if (nWalks.size() != n || sampleStatistics.size() != n)
            throw std::runtime_error("zombie_amd: nWalks, samplePts and sampleStatistics differ in size");
        if (n == 0) return;
        std::vector<float> pt(n * DIM), normal(n * DIM), dirD(n * DIM), dA(n), dR(n), fsr(n);
        std::vector<uint8_t> type(n), aligned(n), quantity(n);
        std::vector<int32_t> nw(nWalks.begin(), nWalks.end());
        bool grad = false, skip = false;
        int32_t nwMax = 0;
        for (size_t i = 0; i < n; i++) {
            const auto& s = samplePts[i];
            for (size_t k = 0; k < DIM; k++) {
                pt[i * DIM + k] = s.pt[k];
                normal[i * DIM + k] = s.normal[k];
                dirD[i * DIM + k] = s.directionForDerivative[k];
            }
            dA[i] = s.distToAbsorbingBoundary;
            dR[i] = s.distToReflectingBoundary;
            fsr[i] = s.firstSphereRadius;
            type[i] = (uint8_t)s.type;
            aligned[i] = s.estimateBoundaryNormalAligned ? 1 : 0;
            quantity[i] = (uint8_t)s.estimationQuantity;
            grad = grad || s.estimationQuantity == EstimationQuantity::SolutionAndGradient;
            skip = skip || s.estimationQuantity == EstimationQuantity::None;
            nwMax = std::max(nwMax, nw[i]);
        }
        std::vector<int64_t> sN(n), sLen(n), sSplit(n), gN(n);
        std::vector<double> sMean(n), sM2(n), gMean(n * DIM), gM2(n * DIM), gFirst(n), gDeriv(n);
        for (size_t i = 0; i < n; i++) {
            const auto& st = sampleStatistics[i];
            sN[i] = st.nSolution; sLen[i] = st.walkLength; sSplit[i] = st.splits; gN[i] = st.nGradient;
            sMean[i] = st.solutionMean; sM2[i] = st.solutionM2; gFirst[i] = st.firstSourceSum;
            gDeriv[i] = st.derivativeSum;
            for (size_t k = 0; k < DIM; k++) { gMean[i * DIM + k] = st.gradientMean[k]; gM2[i * DIM + k] = st.gradientM2[k]; }
        }
        DeviceBuffer bPt, bNormal, bDirD, bDA, bDR, bFsr, bType, bAligned, bQ, bNw;
        DeviceBuffer bSN, bSMean, bSM2, bSLen, bSSplit, bGN, bGMean, bGM2, bGFirst, bGDeriv;
        bPt.upload(pt); bNormal.upload(normal); bDirD.upload(dirD); bDA.upload(dA); bDR.upload(dR);
        bFsr.upload(fsr); bType.upload(type); bAligned.upload(aligned); bQ.upload(quantity); bNw.upload(nw);
        bSN.upload(sN); bSMean.upload(sMean); bSM2.upload(sM2); bSLen.upload(sLen); bSSplit.upload(sSplit);
        bGN.upload(gN); bGMean.upload(gMean); bGM2.upload(gM2); bGFirst.upload(gFirst); bGDeriv.upload(gDeriv);

        zg_sample_points sp;
        std::memset(&sp, 0, sizeof(sp));
        sp.n = (int64_t)n;
        sp.pt = bPt.as<float>(); sp.normal = bNormal.as<float>(); sp.type = bType.as<uint8_t>();
        sp.dist_to_absorbing_boundary = bDA.as<float>(); sp.dist_to_reflecting_boundary = bDR.as<float>();
        sp.estimate_boundary_normal_aligned = bAligned.as<uint8_t>();
        sp.first_sphere_radius = bFsr.as<float>();
        sp.n_walks = bNw.as<int32_t>();
        sp.n_walks_max = nwMax;
        sp.estimation_quantity = (grad || skip) ? bQ.as<uint8_t>() : nullptr;
        sp.direction_for_derivative = grad ? bDirD.as<float>() : nullptr;
        zg_statistics st;
        std::memset(&st, 0, sizeof(st));
        st.n_solution_estimates = bSN.as<int64_t>(); st.solution_mean = bSMean.as<double>();
        st.solution_m2 = bSM2.as<double>(); st.total_walk_length = bSLen.as<int64_t>();
        st.total_splits = bSSplit.as<int64_t>();
        if (grad) {
            st.n_gradient_estimates = bGN.as<int64_t>(); st.gradient_mean = bGMean.as<double>();
            st.gradient_m2 = bGM2.as<double>(); st.total_first_source_contribution = bGFirst.as<double>();
            st.total_derivative_contribution = bGDeriv.as<double>();
        }
        zg_solve_options o;
        zg_solve_options_default(&o);
        o.seed = seed;
        o.stream_offset = streamOffset_;
        zg_walk_settings ws = walkSettings.toC();
        zg_solve_report rep;
        check(WOST ? zg_wost_solve(queries.scene(), &pde.desc, &ws, &sp, &st, &o, &rep, nullptr)
                   : zg_wos_solve(queries.scene(), &pde.desc, &ws, &sp, &st, &o, &rep, nullptr));
        streamOffset_ += (uint32_t)nwMax;
        bSN.download(sN); bSMean.download(sMean); bSM2.download(sM2); bSLen.download(sLen);
        bSSplit.download(sSplit); bFsr.download(fsr);
        if (grad) { bGN.download(gN); bGMean.download(gMean); bGM2.download(gM2); bGFirst.download(gFirst); bGDeriv.download(gDeriv); }
        for (size_t i = 0; i < n; i++) {
            auto& s = sampleStatistics[i];
            s.nSolution = sN[i]; s.solutionMean = sMean[i]; s.solutionM2 = sM2[i]; s.walkLength = sLen[i];
            s.splits = sSplit[i];
            if (grad) {
                s.nGradient = gN[i]; s.firstSourceSum = gFirst[i]; s.derivativeSum = gDeriv[i];
                for (size_t k = 0; k < DIM; k++) { s.gradientMean[k] = gMean[i * DIM + k]; s.gradientM2[k] = gM2[i * DIM + k]; }
            }
            samplePts[i].firstSphereRadius = fsr[i];
        }
        lastReport = rep;
        if (reportProgress) reportProgress((int)n, 0);
    }

    const GeometricQueries<DIM>& queries;   // referenced, as the reference's solvers do
    uint64_t seed;
    mutable zg_solve_report lastReport{};

private:
    mutable uint32_t streamOffset_ = 0;
};

// WalkOnSpheres<T, DIM> (walk_on_spheres.h:22-168) and WalkOnStars<T, DIM> (walk_on_stars.h:22-187)
template <typename T, size_t DIM>
class WalkOnSpheres : public SolverBase<T, DIM, false> {
public:
    using SolverBase<T, DIM, false>::SolverBase;
};

template <typename T, size_t DIM>
class WalkOnStars : public SolverBase<T, DIM, true> {
public:
    using SolverBase<T, DIM, true>::SolverBase;
};

}  // namespace zombie_amd
