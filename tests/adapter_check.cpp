// adapter_check.cpp -- drives the header-only C++ adapter (include/zombie_amd.hpp) the way a
// reference caller drives WalkOnSpheres / WalkOnStars::solve, for tests/test_gpu_adapter.py.
//
//   adapter_check <input.bin> <output.bin>
//
// input (little endian): int32 dim (2), wost, nV, nP, nPts, nWalks; uint64 seed;
//   float positions[nV*2]; int32 indices[nP*2]; uint8 isReflecting[nP]; float domainMin[2],
//   domainMax[2]; float pts[nPts*2], distAbsorbing[nPts], distReflecting[nPts]; zg_pde (raw);
//   float epsA, epsR, silhouettePrecision, rrThreshold, splittingThreshold;
//   int32 maxWalkLength, stepsBeforeTikhonov, stepsBeforeMaximalSpheres; uint8 solveDoubleSided.
// It builds GeometricQueries<2> and a PDE, constructs the 16-argument WalkSettings, runs two
// progressive batch solve() calls with nWalks walks per point, then one single-point solve() on
// point 0 with fresh statistics, and writes per point: int64 n, walkLength; double mean, M2;
// then the single-point call's int64 n and double mean.
#include <cstdio>
#include <fstream>
#include <iostream>
#include <vector>

#include "zombie_amd.hpp"

namespace {
template <typename T>
void readInto(std::ifstream& f, T* dst, size_t count) {
    f.read(reinterpret_cast<char*>(dst), (std::streamsize)(count * sizeof(T)));
    if (!f) throw std::runtime_error("truncated input");
}
template <typename T>
T read1(std::ifstream& f) { T v; readInto(f, &v, 1); return v; }
}  // namespace

int main(int argc, char** argv) {
    if (argc != 3) { std::cerr << "usage: adapter_check <input.bin> <output.bin>\n"; return 2; }
    try {
        std::ifstream in(argv[1], std::ios::binary);
        const int dim = read1<int32_t>(in), wost = read1<int32_t>(in), nV = read1<int32_t>(in);
        const int nP = read1<int32_t>(in), nPts = read1<int32_t>(in), nWalks = read1<int32_t>(in);
        const uint64_t seed = read1<uint64_t>(in);
        if (dim != 2) throw std::runtime_error("adapter_check runs 2D problems");
        std::vector<zombie_amd::Vector<2>> V(nV);
        std::vector<zombie_amd::Vectori<2>> I(nP);
        std::vector<uint8_t> refl(nP);
        readInto(in, V.data()->data(), (size_t)nV * 2);
        readInto(in, I.data()->data(), (size_t)nP * 2);
        readInto(in, refl.data(), (size_t)nP);
        zombie_amd::Vector<2> lo, hi;
        readInto(in, lo.data(), 2);
        readInto(in, hi.data(), 2);
        std::vector<float> pts((size_t)nPts * 2), dA(nPts), dR(nPts);
        readInto(in, pts.data(), pts.size());
        readInto(in, dA.data(), dA.size());
        readInto(in, dR.data(), dR.size());
        zombie_amd::PDE<float, 2> pde;
        readInto(in, reinterpret_cast<char*>(&pde.desc), sizeof(zg_pde));
        float f5[5];
        readInto(in, f5, 5);
        int32_t i3[3];
        readInto(in, i3, 3);
        const bool doubleSided = read1<uint8_t>(in) != 0;

        zombie_amd::GeometricQueries<2> queries(V, I, refl, lo, hi, doubleSided, f5[2]);
        zombie_amd::WalkSettings ws(f5[0], f5[1], f5[2], f5[3], f5[4], i3[0], i3[1], i3[2], doubleSided,
                                    true, true, false, false, false, false, false);
        std::vector<zombie_amd::SamplePoint<float, 2>> samplePts;
        for (int i = 0; i < nPts; i++)
            samplePts.emplace_back(zombie_amd::Vector<2>{pts[2 * i], pts[2 * i + 1]}, zombie_amd::Vector<2>{0.0f, 0.0f},
                                   zombie_amd::SampleType::InDomain, zombie_amd::EstimationQuantity::Solution, 1.0f,
                                   dA[i], dR[i]);
        std::vector<zombie_amd::SampleStatistics<float, 2>> stats(nPts), one(1);
        std::vector<int> nw(nPts, nWalks);
        int reported = 0;
        auto progress = [&](int done, int) { reported += done; };
        if (wost) {
            zombie_amd::WalkOnStars<float, 2> solver(queries, seed);
            solver.solve(pde, ws, nw, samplePts, stats, false, progress);
            solver.solve(pde, ws, nw, samplePts, stats, false, progress);
            zombie_amd::WalkOnStars<float, 2> single(queries, seed);
            zombie_amd::SamplePoint<float, 2> p0 = samplePts[0];
            p0.reset();
            single.solve(pde, ws, nWalks, p0, one[0]);
        } else {
            zombie_amd::WalkOnSpheres<float, 2> solver(queries, seed);
            solver.solve(pde, ws, nw, samplePts, stats, false, progress);
            solver.solve(pde, ws, nw, samplePts, stats, false, progress);
            zombie_amd::WalkOnSpheres<float, 2> single(queries, seed);
            zombie_amd::SamplePoint<float, 2> p0 = samplePts[0];
            single.solve(pde, ws, nWalks, p0, one[0]);
        }
        if (reported != 2 * nPts) throw std::runtime_error("progress callback not called");
        std::ofstream out(argv[2], std::ios::binary);
        for (const auto& s : stats) {
            int64_t n = s.getSolutionEstimateCount(), len = s.walkLength;
            double mean = s.solutionMean, m2 = s.solutionM2;
            out.write(reinterpret_cast<const char*>(&n), 8);
            out.write(reinterpret_cast<const char*>(&len), 8);
            out.write(reinterpret_cast<const char*>(&mean), 8);
            out.write(reinterpret_cast<const char*>(&m2), 8);
        }
        int64_t n1 = one[0].getSolutionEstimateCount();
        double m1 = one[0].solutionMean;
        out.write(reinterpret_cast<const char*>(&n1), 8);
        out.write(reinterpret_cast<const char*>(&m1), 8);
        std::printf("adapter_check: %d points x %d walks x 2 calls (%s)\n", nPts, nWalks, wost ? "stars" : "spheres");
    } catch (const std::exception& e) {
        std::cerr << "adapter_check: " << e.what() << "\n";
        return 1;
    }
    return 0;
}
