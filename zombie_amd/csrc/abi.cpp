// abi.cpp -- the extern "C" entry points declared in include/zombie_amd.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/zombie_amd.h"
#include "bvh_build.h"
#include "kernels.h"

using namespace zg;

namespace {

thread_local std::string g_lastError;

zg_status fail(zg_status code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_lastError = buf;
    return code;
}

zg_status ok() { g_lastError.clear(); return ZG_OK; }

#define ZG_HIP(call)                                                                         \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess)                                                                \
            return fail(e_ == hipErrorOutOfMemory ? ZG_ERR_OUT_OF_MEMORY : ZG_ERR_DEVICE,    \
                        "%s failed: %s", #call, hipGetErrorString(e_));                      \
    } while (0)

int cuCount(int device) {
    static std::mutex m;
    static std::vector<int> cache;
    std::lock_guard<std::mutex> g(m);
    if ((int)cache.size() <= device) cache.resize(device + 1, 0);
    if (cache[device] == 0) {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, device) == hipSuccess) cache[device] = prop.multiProcessorCount;
        else cache[device] = 256;
    }
    return cache[device];
}

}  // namespace

struct DeviceBoundary {
    float* nodes = nullptr;
    float* prims = nullptr;
    int32_t* primIds = nullptr;
    float* sil = nullptr;
    float* silNodes = nullptr;
    BuiltBoundary host;  // sizes only after upload (vectors cleared)
    int64_t bytes = 0;
};

struct zg_scene {
    int dim = 3, width = 8, device = 0;
    DeviceBoundary abs, refl;
    SceneView view{};
    zg_scene_info info{};
};

namespace {

template <class T>
hipError_t upload(T** dst, const std::vector<T>& src, int64_t& bytes) {
    *dst = nullptr;
    if (src.empty()) return hipSuccess;
    size_t b = src.size() * sizeof(T);
    hipError_t e = hipMalloc((void**)dst, b);
    if (e != hipSuccess) return e;
    bytes += (int64_t)b;
    return hipMemcpy(*dst, src.data(), b, hipMemcpyHostToDevice);
}

void freeBoundary(DeviceBoundary& b) {
    if (b.nodes) (void)hipFree(b.nodes);
    if (b.prims) (void)hipFree(b.prims);
    if (b.primIds) (void)hipFree(b.primIds);
    if (b.sil) (void)hipFree(b.sil);
    if (b.silNodes) (void)hipFree(b.silNodes);
    b = DeviceBoundary();
}

zg_status uploadBoundary(DeviceBoundary& d, BoundaryView& v) {
    d.bytes = 0;
    ZG_HIP(upload(&d.nodes, d.host.nodes, d.bytes));
    ZG_HIP(upload(&d.prims, d.host.prims, d.bytes));
    ZG_HIP(upload(&d.primIds, d.host.primIds, d.bytes));
    ZG_HIP(upload(&d.sil, d.host.sil, d.bytes));
    ZG_HIP(upload(&d.silNodes, d.host.silNodes, d.bytes));
    v.nodes = d.nodes; v.prims = d.prims; v.primIds = d.primIds; v.sil = d.sil; v.silNodes = d.silNodes;
    v.nNodes = d.host.nNodes; v.nPrims = d.host.nPrims; v.nSil = d.host.nSil; v.nSilNodes = d.host.nSilNodes;
    return ZG_OK;
}

int64_t nodeBytes(int dim, int width) {
    if (dim == 2) return width == 4 ? NodeLayout<2, 4>::kBytes : NodeLayout<2, 8>::kBytes;
    return width == 4 ? NodeLayout<3, 4>::kBytes : NodeLayout<3, 8>::kBytes;
}

zg_status checkSettings(const zg_walk_settings* ws) {
    if (!ws) return fail(ZG_ERR_INVALID_ARGUMENT, "walk settings are NULL");
    if (ws->max_walk_length < 1) return fail(ZG_ERR_INVALID_ARGUMENT, "max_walk_length must be >= 1");
    if (!(ws->epsilon_shell_for_absorbing_boundary >= 0.0f))
        return fail(ZG_ERR_INVALID_ARGUMENT, "epsilon_shell_for_absorbing_boundary must be >= 0");
    // walk throughput never exceeds max(1, rr threshold) here: splitting below that level
    // would need per-walk state queues, which the GPU path does not implement
    float tmax = std::max(1.0f, ws->russian_roulette_threshold);
    if (ws->splitting_threshold < tmax)
        return fail(ZG_ERR_UNSUPPORTED, "splitting_threshold %g < %g: walk splitting is not supported on the GPU",
                    ws->splitting_threshold, tmax);
    return ZG_OK;
}

}  // namespace

extern "C" {

void zg_walk_settings_default(zg_walk_settings* ws, float epsA, float epsR, int32_t maxLen, int32_t dbl) {
    if (!ws) return;
    std::memset(ws, 0, sizeof(*ws));
    ws->epsilon_shell_for_absorbing_boundary = epsA;
    ws->epsilon_shell_for_reflecting_boundary = epsR;
    ws->silhouette_precision = 1e-3f;
    ws->russian_roulette_threshold = 0.0f;
    ws->splitting_threshold = FLT_MAX;
    ws->max_walk_length = maxLen;
    ws->steps_before_applying_tikhonov = 0;
    ws->steps_before_using_maximal_spheres = maxLen;
    ws->solve_double_sided = dbl ? 1 : 0;
    ws->use_gradient_control_variates = 1;
    ws->use_gradient_antithetic_variates = 1;
}

void zg_solve_options_default(zg_solve_options* o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->seed = 0x5EEDull;
    o->walks_per_lane = 8;
    o->regroup_interval = 16;
    o->step_interval = 4;
    o->waves_per_cu = 0;
}

int32_t zg_last_error(char* buf, int32_t n) {
    int32_t len = (int32_t)g_lastError.size();
    if (buf && n > 0) {
        int32_t c = std::min(len, n - 1);
        std::memcpy(buf, g_lastError.data(), c);
        buf[c] = 0;
    }
    return len;
}

const char* zg_version(void) { return "0.1.0"; }

zg_status zg_device_count(int32_t* count) {
    int c = 0;
    hipError_t e = hipGetDeviceCount(&c);
    if (count) *count = e == hipSuccess ? c : 0;
    if (e != hipSuccess) return fail(ZG_ERR_NO_DEVICE, "hipGetDeviceCount: %s", hipGetErrorString(e));
    return ok();
}
zg_status zg_malloc(void** p, size_t b) { ZG_HIP(hipMalloc(p, b)); return ok(); }
zg_status zg_free(void* p) { ZG_HIP(hipFree(p)); return ok(); }
zg_status zg_memcpy_htod(void* d, const void* s, size_t b) { ZG_HIP(hipMemcpy(d, s, b, hipMemcpyHostToDevice)); return ok(); }
zg_status zg_memcpy_dtoh(void* d, const void* s, size_t b) { ZG_HIP(hipMemcpy(d, s, b, hipMemcpyDeviceToHost)); return ok(); }
zg_status zg_memset(void* d, int32_t v, size_t b) { ZG_HIP(hipMemset(d, v, b)); return ok(); }
zg_status zg_synchronize(void) { ZG_HIP(hipDeviceSynchronize()); return ok(); }

zg_status zg_scene_create(int32_t dim, const float* V, int64_t nV, const int32_t* I, int64_t nP,
                          const uint8_t* refl, const float* dmin, const float* dmax, int32_t dbl, float prec,
                          const zg_build_options* opts, zg_scene** out) {
    if (!out) return fail(ZG_ERR_INVALID_ARGUMENT, "out is NULL");
    *out = nullptr;
    if (dim != 2 && dim != 3) return fail(ZG_ERR_INVALID_ARGUMENT, "dim must be 2 or 3, got %d", dim);
    if (nV < 0 || nP < 0 || (nP > 0 && (!V || !I))) return fail(ZG_ERR_INVALID_ARGUMENT, "invalid mesh arrays");
    if (nP >= (1ll << 27)) return fail(ZG_ERR_INVALID_ARGUMENT, "at most 2^27 primitives per scene");
    for (int64_t k = 0; k < nP * dim; k++)
        if (I[k] < 0 || I[k] >= nV) return fail(ZG_ERR_INVALID_ARGUMENT, "index %d out of range at %lld", I[k], (long long)k);
    zg_build_options o{8, 4, 16, 0};
    if (opts) o = *opts;
    if (o.bvh_width != 4 && o.bvh_width != 8) return fail(ZG_ERR_INVALID_ARGUMENT, "bvh_width must be 4 or 8");
    if (o.leaf_size < 1 || o.leaf_size > 8) return fail(ZG_ERR_INVALID_ARGUMENT, "leaf_size must be 1..8");
    if (o.sah_bins < 4 || o.sah_bins > 64) return fail(ZG_ERR_INVALID_ARGUMENT, "sah_bins must be 4..64");
    int dev = 0;
    ZG_HIP(hipGetDevice(&dev));
    zg_scene* s = new zg_scene();
    s->dim = dim; s->width = o.bvh_width; s->device = dev;
    std::vector<int> selA, selR;
    for (int64_t p = 0; p < nP; p++) ((refl && refl[p]) ? selR : selA).push_back((int)p);
    try {
        buildBoundary(dim, o.bvh_width, o.leaf_size, o.sah_bins, V, I, selA, s->abs.host);
        buildBoundary(dim, o.bvh_width, o.leaf_size, o.sah_bins, V, I, selR, s->refl.host);
        if (dim == 2 && !selR.empty()) buildSilhouettes2D(o.bvh_width, V, nV, I, selR, dbl != 0, prec, s->refl.host);
    } catch (const std::exception& e) {
        delete s;
        return fail(ZG_ERR_INVALID_ARGUMENT, "scene build failed: %s", e.what());
    }
    zg_status st = uploadBoundary(s->abs, s->view.abs);
    if (st == ZG_OK) st = uploadBoundary(s->refl, s->view.refl);
    if (st != ZG_OK) { freeBoundary(s->abs); freeBoundary(s->refl); delete s; return st; }
    s->view.dim = dim; s->view.width = o.bvh_width;
    for (int a = 0; a < 3; a++) {
        s->view.dmin[a] = a < dim && dmin ? dmin[a] : -FLT_MAX;
        s->view.dmax[a] = a < dim && dmax ? dmax[a] : FLT_MAX;
    }
    zg_scene_info& in = s->info;
    in.dim = dim; in.bvh_width = o.bvh_width;
    in.n_absorbing_prims = s->abs.host.nPrims; in.n_reflecting_prims = s->refl.host.nPrims;
    in.n_absorbing_nodes = s->abs.host.nNodes; in.n_reflecting_nodes = s->refl.host.nNodes;
    in.absorbing_depth = s->abs.host.depth; in.reflecting_depth = s->refl.host.depth;
    in.node_bytes = nodeBytes(dim, o.bvh_width);
    in.prim_bytes = dim == 3 ? PrimLayout<3>::kBytes : PrimLayout<2>::kBytes;
    in.device_bytes = s->abs.bytes + s->refl.bytes;
    in.n_silhouette_vertices = s->refl.host.nSil;
    // host copies are no longer needed
    for (DeviceBoundary* b : {&s->abs, &s->refl}) {
        std::vector<float>().swap(b->host.nodes); std::vector<float>().swap(b->host.prims);
        std::vector<int32_t>().swap(b->host.primIds); std::vector<float>().swap(b->host.sil);
        std::vector<float>().swap(b->host.silNodes);
    }
    *out = s;
    return ok();
}

zg_status zg_scene_destroy(zg_scene* s) {
    if (!s) return ok();
    freeBoundary(s->abs);
    freeBoundary(s->refl);
    delete s;
    return ok();
}

zg_status zg_scene_get_info(const zg_scene* s, zg_scene_info* info) {
    if (!s || !info) return fail(ZG_ERR_INVALID_ARGUMENT, "NULL argument");
    *info = s->info;
    return ok();
}

zg_status zg_closest_point(const zg_scene* s, int32_t boundary, int64_t n, const float* pts, float* dist,
                           float* closest, float* normal, int32_t* prim, void* stream) {
    if (!s || n < 0 || (n > 0 && (!pts || !dist))) return fail(ZG_ERR_INVALID_ARGUMENT, "invalid arguments");
    if (boundary != 0 && boundary != 1) return fail(ZG_ERR_INVALID_ARGUMENT, "boundary must be 0 or 1");
    ZG_HIP(launchClosestPoint(s->view, boundary, n, pts, dist, closest, normal, prim, (hipStream_t)stream));
    return ok();
}

zg_status zg_inside_domain(const zg_scene* s, int64_t n, const float* pts, uint8_t* inside, void* stream) {
    if (!s || n < 0 || (n > 0 && (!pts || !inside))) return fail(ZG_ERR_INVALID_ARGUMENT, "invalid arguments");
    ZG_HIP(launchInsideDomain(s->view, n, pts, inside, (hipStream_t)stream));
    return ok();
}

zg_status zg_intersect_ray(const zg_scene* s, int32_t boundary, int64_t n, const float* o, const float* d,
                           const float* tmax, uint8_t* hit, float* t, float* point, float* normal, void* stream) {
    if (!s || n < 0 || (n > 0 && (!o || !d || !tmax || !hit))) return fail(ZG_ERR_INVALID_ARGUMENT, "invalid arguments");
    if (boundary != 0 && boundary != 1) return fail(ZG_ERR_INVALID_ARGUMENT, "boundary must be 0 or 1");
    ZG_HIP(launchRay(s->view, boundary, n, o, d, tmax, hit, t, point, normal, (hipStream_t)stream));
    return ok();
}

zg_status zg_star_radius(const zg_scene* s, int64_t n, const float* pts, const float* minR, const float* maxR,
                         float prec, const uint8_t* flip, float* radius, void* stream) {
    (void)s; (void)n; (void)pts; (void)minR; (void)maxR; (void)prec; (void)flip; (void)radius; (void)stream;
    return fail(ZG_ERR_UNSUPPORTED, "zg_star_radius is not implemented yet");
}

zg_status zg_wos_solve(const zg_scene* s, const zg_pde* pde, const zg_walk_settings* ws, const zg_sample_points* pts,
                       zg_statistics* stats, const zg_solve_options* opts, zg_solve_report* report, void* stream) {
    if (!s || !pde || !pts || !stats) return fail(ZG_ERR_INVALID_ARGUMENT, "NULL argument");
    zg_status st = checkSettings(ws);
    if (st != ZG_OK) return st;
    if (pde->absorption_coeff < 0.0f) return fail(ZG_ERR_INVALID_ARGUMENT, "absorption_coeff must be >= 0");
    if (pde->absorption_coeff > 0.0f && s->dim == 2)
        return fail(ZG_ERR_UNSUPPORTED, "screened Poisson (absorption > 0) is supported in 3D only");
    if (pts->n < 0) return fail(ZG_ERR_INVALID_ARGUMENT, "negative point count");
    if (pts->n > 0 && (!pts->pt || !pts->dist_to_absorbing_boundary || !pts->n_walks))
        return fail(ZG_ERR_INVALID_ARGUMENT, "pt, dist_to_absorbing_boundary and n_walks are required");
    if (pts->n > 0 && (!stats->n_solution_estimates || !stats->solution_mean || !stats->solution_m2))
        return fail(ZG_ERR_INVALID_ARGUMENT, "statistics arrays are required");
    if (pts->n_walks_max < 0) return fail(ZG_ERR_INVALID_ARGUMENT, "n_walks_max must be >= 0");
    if (pts->n >= (1ll << 31)) return fail(ZG_ERR_INVALID_ARGUMENT, "at most 2^31-1 points per call");
    zg_solve_options o;
    zg_solve_options_default(&o);
    if (opts) o = *opts;
    if (o.walks_per_lane < 1 || o.walks_per_lane > 64) return fail(ZG_ERR_INVALID_ARGUMENT, "walks_per_lane must be 1..64");
    if (o.regroup_interval < 1 || o.regroup_interval > 1024) return fail(ZG_ERR_INVALID_ARGUMENT, "regroup_interval must be 1..1024");
    if (o.step_interval < 1 || o.step_interval > 1024) return fail(ZG_ERR_INVALID_ARGUMENT, "step_interval must be 1..1024");
    if (!(o.distance_tolerance >= 0.0f && o.distance_tolerance < 0.5f))
        return fail(ZG_ERR_INVALID_ARGUMENT, "distance_tolerance must be in [0, 0.5)");
    if ((uint64_t)o.stream_offset + (uint64_t)pts->n_walks_max > 0xffffffffull)
        return fail(ZG_ERR_INVALID_ARGUMENT, "stream_offset + n_walks_max must fit in 32 bits");
    hipStream_t strm = (hipStream_t)stream;
    int dev = 0;
    ZG_HIP(hipGetDevice(&dev));
    if (dev != s->device) return fail(ZG_ERR_INVALID_ARGUMENT, "scene lives on device %d, current device is %d", s->device, dev);
    int64_t n = pts->n;
    if (n == 0) { if (report) std::memset(report, 0, sizeof(*report)); return ok(); }

    // workspace: per-point batch sums + counters (stream-ordered allocation)
    size_t words = (size_t)n * 4 + kCntSlots;
    uint64_t* wsp = nullptr;
    ZG_HIP(hipMallocAsync((void**)&wsp, words * 8, strm));
    hipError_t e = launchFillZero(wsp, (int64_t)words, strm);
    WosParams P;
    std::memset(&P, 0, sizeof(P));
    P.scene = s->view; P.pde = *pde; P.ws = *ws;
    P.pt = pts->pt; P.normal = pts->normal; P.type = pts->type; P.aligned = pts->estimate_boundary_normal_aligned;
    P.dist0 = pts->dist_to_absorbing_boundary; P.nWalks = pts->n_walks; P.nPoints = n;
    P.wpl = o.walks_per_lane; P.regroup = o.regroup_interval; P.stepInterval = o.step_interval;
    P.chunksPerPoint = std::max(1, (pts->n_walks_max + o.walks_per_lane - 1) / o.walks_per_lane);
    P.nClaims = (unsigned long long)n * (unsigned long long)P.chunksPerPoint;
    P.seed = o.seed; P.streamOffset = o.stream_offset;
    P.relax2 = (1.0f - o.distance_tolerance) * (1.0f - o.distance_tolerance);
    {
        float e2 = 2.0f * ws->epsilon_shell_for_absorbing_boundary;
        P.guard2 = e2 * e2 / P.relax2;
    }
    P.accSum = reinterpret_cast<double*>(wsp);
    P.accSq = reinterpret_cast<double*>(wsp + n);
    P.accN = reinterpret_cast<unsigned long long*>(wsp + 2 * n);
    P.accLen = reinterpret_cast<unsigned long long*>(wsp + 3 * n);
    P.counters = reinterpret_cast<unsigned long long*>(wsp + 4 * n);
    P.claimCounter = P.counters + kCntClaims;
    int wavesPerCu = o.waves_per_cu > 0 ? o.waves_per_cu : 32;
    int blocks = std::max(1, cuCount(dev) * wavesPerCu / 4);
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    if (e == hipSuccess && report) {
        e = hipEventCreate(&ev0);
        if (e == hipSuccess) e = hipEventCreate(&ev1);
        if (e == hipSuccess) e = hipEventRecord(ev0, strm);
    }
    if (e == hipSuccess && pts->n_walks_max > 0) e = launchWos(P, s->dim, s->width, o.count_visits != 0, blocks, strm);
    if (e == hipSuccess && report) e = hipEventRecord(ev1, strm);
    MergeParams M;
    std::memset(&M, 0, sizeof(M));
    M.n = n; M.pt = pts->pt; M.type = pts->type; M.aligned = pts->estimate_boundary_normal_aligned;
    M.dist0 = pts->dist_to_absorbing_boundary;
    M.accSum = P.accSum; M.accSq = P.accSq; M.accN = P.accN; M.accLen = P.accLen;
    M.sN = stats->n_solution_estimates; M.sMean = stats->solution_mean; M.sM2 = stats->solution_m2;
    M.sLen = stats->total_walk_length; M.firstSphereRadius = pts->first_sphere_radius;
    M.dirichlet = pde->dirichlet; M.absorption = pde->absorption_coeff;
    M.ignoreAbsorbing = ws->ignore_absorbing_boundary_contribution;
    M.solveDoubleSided = ws->solve_double_sided;
    M.writeFirstSphereRadius = 1;
    if (e == hipSuccess) e = launchMergeStats(M, s->dim, strm);
    unsigned long long cnt[kCntSlots] = {0};
    if (e == hipSuccess) e = hipMemcpyAsync(cnt, P.counters, sizeof(cnt), hipMemcpyDeviceToHost, strm);
    if (e == hipSuccess) e = hipStreamSynchronize(strm);
    float ms = 0.0f;
    if (e == hipSuccess && report) e = hipEventElapsedTime(&ms, ev0, ev1);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    hipError_t ef = hipFreeAsync(wsp, strm);
    if (e == hipSuccess) e = ef;
    if (e != hipSuccess) return fail(ZG_ERR_DEVICE, "walk-on-spheres launch failed: %s", hipGetErrorString(e));
    if (cnt[kCntError] & 1u)
        return fail(ZG_ERR_DEVICE, "BVH traversal stack overflow (depth %d, width %d)", s->info.absorbing_depth, s->width);
    if (report) {
        for (int k = 0; k < 4; k++) report->completion_codes[k] = (int64_t)cnt[k];
        report->total_steps = (int64_t)cnt[kCntSteps];
        report->total_walks = (int64_t)cnt[kCntWalks];
        report->node_visits = (int64_t)cnt[kCntNodes];
        report->prim_visits = (int64_t)cnt[kCntPrims];
        report->queries = (int64_t)cnt[kCntQueries];
        report->kernel_ms = ms;
        report->reserved = 0.0f;
    }
    return ok();
}

zg_status zg_wost_solve(const zg_scene* s, const zg_pde* pde, const zg_walk_settings* ws, const zg_sample_points* pts,
                        zg_statistics* stats, const zg_solve_options* opts, zg_solve_report* report, void* stream) {
    (void)s; (void)pde; (void)ws; (void)pts; (void)stats; (void)opts; (void)report; (void)stream;
    return fail(ZG_ERR_UNSUPPORTED, "zg_wost_solve is not implemented yet");
}

}  // extern "C"
