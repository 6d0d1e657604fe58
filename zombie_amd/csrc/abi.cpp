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
#include "mlp.h"

#include <random>

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
    constexpr int kMaxDevices = 64;
    static std::mutex m;
    static int cache[kMaxDevices] = {};
    if (device < 0 || device >= kMaxDevices) return 256;
    std::lock_guard<std::mutex> g(m);
    if (cache[device] == 0) {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, device) == hipSuccess) cache[device] = prop.multiProcessorCount;
        else cache[device] = 256;
    }
    return cache[device];
}

}  // namespace

struct zg_mlp {
    MlpDevice d{};
    zg_mlp_config cfg{};
    int device = 0;
    std::vector<void*> allocations;
};

struct DeviceBoundary {
    float* nodes = nullptr;
    float* prims = nullptr;
    int32_t* primIds = nullptr;
    float* sil = nullptr;
    float* silNodes = nullptr;
    float* silCones = nullptr;
    float silConePrecision = 0.0f;
    float* robinMax = nullptr;   // largest Robin coefficient per packed primitive
    BuiltBoundary host;  // sizes only after upload (vectors cleared; reflecting primIds kept)
    int64_t bytes = 0;
};

struct zg_scene {
    int dim = 3, width = 8, device = 0;
    int64_t nPrims = 0;
    zg_function sdf{};          // SDF-grid scenes: device grid in sdf.grid (kind ZG_FN_DENSE_GRID, interpolated)
    float* sdfGrid = nullptr;
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
    if (b.silCones) (void)hipFree(b.silCones);
    if (b.robinMax) (void)hipFree(b.robinMax);
    b = DeviceBoundary();
}

zg_status uploadBoundary(DeviceBoundary& d, BoundaryView& v) {
    d.bytes = 0;
    ZG_HIP(upload(&d.nodes, d.host.nodes, d.bytes));
    ZG_HIP(upload(&d.prims, d.host.prims, d.bytes));
    ZG_HIP(upload(&d.primIds, d.host.primIds, d.bytes));
    ZG_HIP(upload(&d.sil, d.host.sil, d.bytes));
    ZG_HIP(upload(&d.silNodes, d.host.silNodes, d.bytes));
    ZG_HIP(upload(&d.silCones, d.host.silCones, d.bytes));
    v.nodes = d.nodes; v.prims = d.prims; v.primIds = d.primIds; v.sil = d.sil; v.silNodes = d.silNodes;
    v.silCones = d.silCones; v.silConePrecision = d.silConePrecision;
    v.nNodes = d.host.nNodes; v.nPrims = d.host.nPrims; v.nSil = d.host.nSil; v.nSilNodes = d.host.nSilNodes;
    return ZG_OK;
}

int64_t nodeBytes(int dim, int width) {
    if (dim == 2) return width == 4 ? NodeLayout<2, 4>::kBytes : NodeLayout<2, 8>::kBytes;
    return width == 4 ? NodeLayout<3, 4>::kBytes : NodeLayout<3, 8>::kBytes;
}

zg_status checkFunction(const zg_function& f, int dim) {
    if (f.kind < ZG_FN_ZERO || f.kind > ZG_FN_DENSE_GRID) return fail(ZG_ERR_INVALID_ARGUMENT, "unknown function kind %d", f.kind);
    if (f.eval < ZG_EVAL_VALUE || f.eval > ZG_EVAL_ROBIN) return fail(ZG_ERR_INVALID_ARGUMENT, "unknown eval mode %d", f.eval);
    if (f.kind != ZG_FN_DENSE_GRID) return ZG_OK;
    if (!f.grid) return fail(ZG_ERR_INVALID_ARGUMENT, "dense grid without data");
    for (int a = 0; a < 3; a++) {
        if (f.shape[a] < 1 || (a >= dim && f.shape[a] != 1))
            return fail(ZG_ERR_INVALID_ARGUMENT, "dense grid shape[%d] = %d is invalid", a, f.shape[a]);
        if (a < dim && !(f.grid_max[a] > f.grid_min[a])) return fail(ZG_ERR_INVALID_ARGUMENT, "dense grid box is empty");
    }
    return ZG_OK;
}

zg_status checkSettings(const zg_walk_settings* ws) {
    if (!ws) return fail(ZG_ERR_INVALID_ARGUMENT, "walk settings are NULL");
    if (ws->max_walk_length < 1) return fail(ZG_ERR_INVALID_ARGUMENT, "max_walk_length must be >= 1");
    if (!(ws->epsilon_shell_for_absorbing_boundary >= 0.0f))
        return fail(ZG_ERR_INVALID_ARGUMENT, "epsilon_shell_for_absorbing_boundary must be >= 0");
    if (!(ws->splitting_threshold > 0.0f))
        return fail(ZG_ERR_INVALID_ARGUMENT, "splitting_threshold must be > 0");
    if (!(ws->russian_roulette_threshold >= 0.0f))
        return fail(ZG_ERR_INVALID_ARGUMENT, "russian_roulette_threshold must be >= 0");
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
zg_status zg_synchronize(void) {
    ZG_HIP(hipDeviceSynchronize());
    // error flags latched by solves that returned without synchronising (no report requested)
    unsigned int flags = 0;
    ZG_HIP(deviceErrorsTakeAndClear(&flags));
    if (flags & 2u)
        return fail(ZG_ERR_INVALID_ARGUMENT, "an asynchronous batched solve met an instance_of_point outside [0, n_instances)");
    if (flags & 1u) return fail(ZG_ERR_DEVICE, "an asynchronous solve overflowed its BVH traversal stack");
    return ok();
}

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
        if (dim == 2 && !selR.empty()) {
            buildSilhouettes2D(o.bvh_width, V, nV, I, selR, dbl != 0, prec, s->refl.host);
            s->refl.silConePrecision = prec;   // the cones' margin covers queries up to this precision
        } else if (dim == 3 && !selR.empty()) {
            buildSilhouettes3D(o.bvh_width, V, nV, I, selR, dbl != 0, prec, s->refl.host);
            s->refl.silConePrecision = prec;
        }
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
        std::vector<float>().swap(b->host.sil); std::vector<float>().swap(b->host.silNodes);
    }
    std::vector<int32_t>().swap(s->abs.host.primIds);   // the reflecting ones map Robin coefficients
    s->nPrims = nP;
    *out = s;
    return ok();
}

zg_status zg_scene_set_robin_coefficients(zg_scene* s, const float* minCoeff, const float* maxCoeff) {
    if (!s) return fail(ZG_ERR_INVALID_ARGUMENT, "NULL scene");
    if ((minCoeff == nullptr) != (maxCoeff == nullptr))
        return fail(ZG_ERR_INVALID_ARGUMENT, "give both coefficient arrays or neither");
    const std::vector<int32_t>& ids = s->refl.host.primIds;
    std::vector<float> mx(ids.size());
    if (minCoeff) {
        for (size_t i = 0; i < ids.size(); i++) {
            int32_t p = ids[i];
            if (p < 0 || p >= s->nPrims) return fail(ZG_ERR_DEVICE, "reflecting primitive map is corrupt");
            float lo = minCoeff[p], hi = maxCoeff[p];
            if (!(lo <= hi) || !std::isfinite(lo) || !std::isfinite(hi))
                return fail(ZG_ERR_INVALID_ARGUMENT, "Robin coefficient bounds of primitive %d are invalid", (int)p);
            mx[i] = hi;
        }
    }
    if (s->refl.robinMax) { ZG_HIP(hipFree(s->refl.robinMax)); s->refl.robinMax = nullptr; }
    s->view.refl.robinMax = nullptr;
    if (!minCoeff || ids.empty()) return ok();
    int64_t bytes = 0;
    ZG_HIP(upload(&s->refl.robinMax, mx, bytes));
    s->view.refl.robinMax = s->refl.robinMax;
    return ok();
}

zg_status zg_sdf_scene_create(int32_t dim, const float* values, const int32_t* shape, const float* gridMin,
                              const float* gridMax, const float* domainMin, const float* domainMax, zg_scene** out) {
    if (!out) return fail(ZG_ERR_INVALID_ARGUMENT, "out is NULL");
    *out = nullptr;
    if (dim != 2 && dim != 3) return fail(ZG_ERR_INVALID_ARGUMENT, "dim must be 2 or 3, got %d", dim);
    if (!values || !shape || !gridMin || !gridMax) return fail(ZG_ERR_INVALID_ARGUMENT, "NULL grid argument");
    size_t count = 1;
    for (int a = 0; a < dim; a++) {
        if (shape[a] < 2) return fail(ZG_ERR_INVALID_ARGUMENT, "SDF grid shape[%d] = %d must be >= 2", a, shape[a]);
        if (!(gridMax[a] > gridMin[a])) return fail(ZG_ERR_INVALID_ARGUMENT, "SDF grid box is empty");
        count *= (size_t)shape[a];
    }
    int dev = 0;
    ZG_HIP(hipGetDevice(&dev));
    zg_scene* s = new zg_scene();
    s->dim = dim; s->width = 4; s->device = dev;
    hipError_t e = hipMalloc((void**)&s->sdfGrid, count * sizeof(float));
    if (e == hipSuccess) e = hipMemcpy(s->sdfGrid, values, count * sizeof(float), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        if (s->sdfGrid) (void)hipFree(s->sdfGrid);
        delete s;
        return fail(ZG_ERR_DEVICE, "zg_sdf_scene_create: %s", hipGetErrorString(e));
    }
    zg_function& f = s->sdf;
    f.kind = ZG_FN_DENSE_GRID; f.eval = ZG_EVAL_VALUE; f.grid = s->sdfGrid; f.interpolate = 1;
    for (int a = 0; a < 3; a++) {
        f.shape[a] = a < dim ? shape[a] : 1;
        f.grid_min[a] = a < dim ? gridMin[a] : 0.0f;
        f.grid_max[a] = a < dim ? gridMax[a] : 1.0f;
        s->view.dmin[a] = a < dim ? (domainMin ? domainMin[a] : gridMin[a]) : -FLT_MAX;
        s->view.dmax[a] = a < dim ? (domainMax ? domainMax[a] : gridMax[a]) : FLT_MAX;
    }
    s->view.dim = dim; s->view.width = 4;
    s->info.dim = dim; s->info.bvh_width = 4;
    s->info.device_bytes = (int64_t)(count * sizeof(float));
    *out = s;
    return ok();
}

zg_status zg_scene_destroy(zg_scene* s) {
    if (!s) return ok();
    if (s->sdfGrid) (void)hipFree(s->sdfGrid);
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

zg_status zg_scene_copy_bvh(const zg_scene* s, int32_t boundary, float* nodes_host, int64_t node_floats,
                            float* prims_host, int64_t prim_floats) {
    if (!s) return fail(ZG_ERR_INVALID_ARGUMENT, "NULL scene");
    if (boundary != 0 && boundary != 1) return fail(ZG_ERR_INVALID_ARGUMENT, "boundary must be 0 or 1");
    if (s->sdfGrid) return fail(ZG_ERR_UNSUPPORTED, "SDF-grid scenes have no BVH");
    const BoundaryView& b = boundary == 0 ? s->view.abs : s->view.refl;
    const int64_t nf = (int64_t)b.nNodes * (s->info.node_bytes / 4);
    const int64_t pf = (int64_t)b.nPrims * (s->info.prim_bytes / 4);
    if ((nodes_host && node_floats < nf) || (prims_host && prim_floats < pf))
        return fail(ZG_ERR_INVALID_ARGUMENT, "buffers hold %lld / %lld floats, the BVH needs %lld / %lld",
                    (long long)node_floats, (long long)prim_floats, (long long)nf, (long long)pf);
    if (nodes_host && nf > 0) ZG_HIP(hipMemcpy(nodes_host, b.nodes, (size_t)nf * 4, hipMemcpyDeviceToHost));
    if (prims_host && pf > 0) ZG_HIP(hipMemcpy(prims_host, b.prims, (size_t)pf * 4, hipMemcpyDeviceToHost));
    return ok();
}

zg_status zg_closest_point(const zg_scene* s, int32_t boundary, int64_t n, const float* pts, float* dist,
                           float* closest, float* normal, int32_t* prim, void* stream) {
    if (!s || n < 0 || (n > 0 && (!pts || !dist))) return fail(ZG_ERR_INVALID_ARGUMENT, "invalid arguments");
    if (boundary != 0 && boundary != 1) return fail(ZG_ERR_INVALID_ARGUMENT, "boundary must be 0 or 1");
    if (s->sdfGrid) {
        if (boundary != 0) return fail(ZG_ERR_UNSUPPORTED, "SDF-grid scenes have an absorbing boundary only");
        ZG_HIP(launchSdfQuery(s->sdf, s->dim, n, pts, dist, closest, normal, prim, (hipStream_t)stream));
        return ok();
    }
    ZG_HIP(launchClosestPoint(s->view, boundary, n, pts, dist, closest, normal, prim, (hipStream_t)stream));
    return ok();
}

zg_status zg_inside_domain(const zg_scene* s, int64_t n, const float* pts, uint8_t* inside, void* stream) {
    if (!s || n < 0 || (n > 0 && (!pts || !inside))) return fail(ZG_ERR_INVALID_ARGUMENT, "invalid arguments");
    if (s->sdfGrid) { ZG_HIP(launchSdfInside(s->sdf, s->dim, n, pts, inside, (hipStream_t)stream)); return ok(); }
    ZG_HIP(launchInsideDomain(s->view, n, pts, inside, (hipStream_t)stream));
    return ok();
}

zg_status zg_intersect_ray(const zg_scene* s, int32_t boundary, int64_t n, const float* o, const float* d,
                           const float* tmax, uint8_t* hit, float* t, float* point, float* normal, void* stream) {
    if (!s || n < 0 || (n > 0 && (!o || !d || !tmax || !hit))) return fail(ZG_ERR_INVALID_ARGUMENT, "invalid arguments");
    if (boundary != 0 && boundary != 1) return fail(ZG_ERR_INVALID_ARGUMENT, "boundary must be 0 or 1");
    if (s->sdfGrid) {
        if (boundary != 0) return fail(ZG_ERR_UNSUPPORTED, "SDF-grid scenes have no reflecting boundary");
        ZG_HIP(launchSdfRay(s->sdf, s->dim, n, o, d, tmax, hit, t, point, normal, (hipStream_t)stream));
        return ok();
    }
    ZG_HIP(launchRay(s->view, boundary, n, o, d, tmax, hit, t, point, normal, (hipStream_t)stream));
    return ok();
}

zg_status zg_star_radius(const zg_scene* s, int64_t n, const float* pts, const float* minR, const float* maxR,
                         float prec, const uint8_t* flip, float* radius, void* stream) {
    if (!s || n < 0 || (n > 0 && (!pts || !minR || !maxR || !radius))) return fail(ZG_ERR_INVALID_ARGUMENT, "invalid arguments");
    if (s->sdfGrid) return fail(ZG_ERR_UNSUPPORTED, "star radius queries run on mesh scenes");
    ZG_HIP(launchStarRadius(s->view, n, pts, minR, maxR, prec, flip, radius, (hipStream_t)stream));
    return ok();
}

zg_status zg_evaluate_function(int32_t dim, const zg_function* f, int64_t n, const float* pts, const float* normals,
                               const uint8_t* aligned, float absorption, float* out, void* stream) {
    if (!f || (dim != 2 && dim != 3) || n < 0 || (n > 0 && (!pts || !out)))
        return fail(ZG_ERR_INVALID_ARGUMENT, "invalid arguments");
    zg_status st = checkFunction(*f, dim);
    if (st != ZG_OK) return st;
    ZG_HIP(launchEvalFunction(*f, dim, n, pts, normals, aligned, absorption, out, (hipStream_t)stream));
    return ok();
}

zg_status zg_probe_read_bandwidth(const void* buf, int64_t bytes, int32_t reps, float* ms_out, void* stream) {
    if (!buf || bytes < 16 || reps < 1 || !ms_out) return fail(ZG_ERR_INVALID_ARGUMENT, "invalid arguments");
    int dev = 0;
    ZG_HIP(hipGetDevice(&dev));
    hipStream_t s = (hipStream_t)stream;
    float* sink = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipMallocAsync((void**)&sink, sizeof(float), s);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = hipEventRecord(e0, s);
    if (e == hipSuccess) e = launchReadProbe(buf, bytes, reps, cuCount(dev) * 8, sink, s);
    if (e == hipSuccess) e = hipEventRecord(e1, s);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    if (e == hipSuccess) e = hipEventElapsedTime(ms_out, e0, e1);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (sink) (void)hipFreeAsync(sink, s);
    if (e != hipSuccess) return fail(ZG_ERR_DEVICE, "read probe failed: %s", hipGetErrorString(e));
    return ok();
}

zg_status zg_line_of_sight(const zg_scene* s, int64_t n, const float* p1, const float* p2, uint8_t* visible,
                           void* stream) {
    if (!s || n < 0 || (n > 0 && (!p1 || !p2 || !visible))) return fail(ZG_ERR_INVALID_ARGUMENT, "invalid arguments");
    if (s->sdfGrid) return fail(ZG_ERR_UNSUPPORTED, "line-of-sight queries run on mesh scenes");
    ZG_HIP(launchLineOfSight(s->view, n, p1, p2, visible, (hipStream_t)stream));
    return ok();
}

zg_status zg_sample_reflecting(const zg_scene* s, int64_t n, const float* x, const float* radius, const float* u,
                               uint8_t* ok_, float* pt, float* normal, float* pdf, void* stream) {
    if (!s || n < 0 || (n > 0 && (!x || !radius || !u || !ok_))) return fail(ZG_ERR_INVALID_ARGUMENT, "invalid arguments");
    if (s->sdfGrid) return fail(ZG_ERR_UNSUPPORTED, "reflecting-boundary sampling runs on mesh scenes");
    ZG_HIP(launchSampleReflecting(s->view, n, x, radius, u, ok_, pt, normal, pdf, (hipStream_t)stream));
    return ok();
}

}  // extern "C"

// A batch of instances walked by one persistent launch (the scenes stay owned by the caller).
struct zg_batch {
    int n = 0, dim = 3, width = 8, device = 0, maxDepth = 0;
    std::vector<const zg_scene*> scenes;
    std::vector<zg_pde> pdes;
    InstanceDesc* table = nullptr;
};

// A solve captured in a hipGraph: fill, walk and merge kernels over a library-owned workspace.
struct zg_graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    uint64_t* workspace = nullptr;
    uint32_t* streamOffset = nullptr;
    int64_t n = 0;
    int depth = 0, width = 0;
};

namespace {

struct SolveSetup {
    WosParams P;
    MergeParams M;
    int dim = 3, width = 8, blocks = 1, depth = 0;
    bool batch = false, count = false, wost = false, screened = false, split = false, interleave = false;
    int64_t n = 0;
    int nWalksMax = 0;
    size_t words = 0;
    const float* callerFsr = nullptr;   // walk on stars: caller's recorded first sphere radii
    zg_statistics stats{};              // caller's statistics (gradient members checked per call)
    bool sdf = false;                   // SDF-grid scene
};

zg_status checkPde(const zg_pde* pde, int dim) {
    const zg_function* fs[4] = {&pde->source, &pde->dirichlet, &pde->robin, &pde->robin_coeff};
    for (const zg_function* f : fs) { zg_status st = checkFunction(*f, dim); if (st != ZG_OK) return st; }
    if (pde->absorption_coeff < 0.0f) return fail(ZG_ERR_INVALID_ARGUMENT, "absorption_coeff must be >= 0");
    if (pde->absorption_coeff > 0.0f && dim == 2)
        return fail(ZG_ERR_UNSUPPORTED, "screened Poisson (absorption > 0) is supported in 3D only");
    return ZG_OK;
}

// validation and kernel parameters of a walk-on-spheres solve (no allocation, no launch)
zg_status prepareSolve(const zg_scene* s, const zg_batch* b, const zg_pde* pde, const zg_walk_settings* ws,
                       const zg_sample_points* pts, const int32_t* instanceOf, zg_statistics* stats,
                       const zg_solve_options* opts, SolveSetup& S, bool wost = false) {
    if ((!s && !b) || !pts || !stats || (s && !pde)) return fail(ZG_ERR_INVALID_ARGUMENT, "NULL argument");
    zg_status st = checkSettings(ws);
    if (st != ZG_OK) return st;
    int dim = s ? s->dim : b->dim;
    int sceneDevice = s ? s->device : b->device;
    if (s) { st = checkPde(pde, dim); if (st != ZG_OK) return st; }
    // walk throughputs start at 1 and never grow under the harmonic and Yukawa kernels and pure
    // Neumann reflectance, so those walks split only below a threshold of 1; walk on stars with
    // Robin coefficients that may be negative clamps the reflectance to [0, 2]
    // (walk_on_stars.h:298-303), so its throughput can grow past any finite threshold
    bool split = ws->splitting_threshold < 1.0f;
    if (wost) {
        if (!s) return fail(ZG_ERR_UNSUPPORTED, "walk on stars runs on single scenes");
        if (pde && !pde->are_robin_conditions_pure_neumann && !pde->are_robin_coeffs_nonnegative &&
            ws->splitting_threshold < FLT_MAX)
            split = true;
    }
    if (pts->n < 0) return fail(ZG_ERR_INVALID_ARGUMENT, "negative point count");
    if (pts->n > 0 && (!pts->pt || !pts->dist_to_absorbing_boundary || !pts->n_walks))
        return fail(ZG_ERR_INVALID_ARGUMENT, "pt, dist_to_absorbing_boundary and n_walks are required");
    if (pts->n > 0 && (!stats->n_solution_estimates || !stats->solution_mean || !stats->solution_m2))
        return fail(ZG_ERR_INVALID_ARGUMENT, "statistics arrays are required");
    if (b && pts->n > 0 && !instanceOf) return fail(ZG_ERR_INVALID_ARGUMENT, "instance_of_point is required");
    if (pts->n_walks_max < 0) return fail(ZG_ERR_INVALID_ARGUMENT, "n_walks_max must be >= 0");
    if (pts->n >= (1ll << 31)) return fail(ZG_ERR_INVALID_ARGUMENT, "at most 2^31-1 points per call");
    zg_solve_options o;
    zg_solve_options_default(&o);
    if (opts) o = *opts;
    if (o.walks_per_lane < 1 || o.walks_per_lane > 64) return fail(ZG_ERR_INVALID_ARGUMENT, "walks_per_lane must be 1..64");
    if (o.regroup_interval < 1 || o.regroup_interval > 1024) return fail(ZG_ERR_INVALID_ARGUMENT, "regroup_interval must be 1..1024");
    if (o.step_interval < 1 || o.step_interval > 1024) return fail(ZG_ERR_INVALID_ARGUMENT, "step_interval must be 1..1024");
    if (o.count_visits < 0 || o.count_visits > 2) return fail(ZG_ERR_INVALID_ARGUMENT, "count_visits must be 0, 1 or 2");
    if (o.count_visits == 2 && wost) return fail(ZG_ERR_UNSUPPORTED, "count_visits 2 runs in walk on spheres only");
    if (!(o.distance_tolerance >= 0.0f && o.distance_tolerance < 0.5f))
        return fail(ZG_ERR_INVALID_ARGUMENT, "distance_tolerance must be in [0, 0.5)");
    if (o.wost_interleave_queries != 0 && o.wost_interleave_queries != 1)
        return fail(ZG_ERR_INVALID_ARGUMENT, "wost_interleave_queries must be 0 or 1");
    if ((uint64_t)o.stream_offset + (uint64_t)pts->n_walks_max > 0xffffffffull)
        return fail(ZG_ERR_INVALID_ARGUMENT, "stream_offset + n_walks_max must fit in 32 bits");
    int dev = 0;
    ZG_HIP(hipGetDevice(&dev));
    if (dev != sceneDevice) return fail(ZG_ERR_INVALID_ARGUMENT, "scene lives on device %d, current device is %d", sceneDevice, dev);

    WosParams& P = S.P;
    std::memset(&P, 0, sizeof(P));
    // control-variate gradient grid
    if (pts->cv_gradient_grid && pts->cv_gradient_grid->data) {
        const zg_grid_field& gf = *pts->cv_gradient_grid;
        for (int a = 0; a < 3; a++) {
            int r = gf.res[a];
            if (r < 1 || (a < dim && r < 2) || (a >= dim && r != 1))
                return fail(ZG_ERR_INVALID_ARGUMENT, "cv_gradient_grid res[%d] = %d is invalid", a, r);
            if (a < dim && !(gf.hi[a] > gf.lo[a])) return fail(ZG_ERR_INVALID_ARGUMENT, "cv_gradient_grid box is empty");
            P.cvRes[a] = r;
            P.cvLo[a] = gf.lo[a];
            P.cvScale[a] = a < dim ? (float)(r - 1) / (gf.hi[a] - gf.lo[a]) : 0.0f;
        }
        P.cvGrid = reinterpret_cast<const float4*>(gf.data);
    }
    if (s) { P.scene = s->view; P.pde = *pde; }
    else { P.instances = b->table; P.instanceOf = instanceOf; P.nInstances = b->n; }
    if (s && s->sdfGrid) {
        if (wost) return fail(ZG_ERR_UNSUPPORTED, "walk on stars needs a mesh scene");
        if (pts->cv_gradient || pts->cv_gradient_grid)
            return fail(ZG_ERR_UNSUPPORTED, "control-variate gradients are not available on SDF-grid scenes");
        P.sdf = s->sdf;
        S.sdf = true;
    }
    P.ws = *ws;
    int64_t n = pts->n;
    P.pt = pts->pt; P.normal = pts->normal; P.type = pts->type; P.aligned = pts->estimate_boundary_normal_aligned;
    P.dist0 = pts->dist_to_absorbing_boundary; P.nWalks = pts->n_walks; P.nPoints = n;
    P.cvGrad = pts->cv_gradient;
    P.wpl = o.walks_per_lane; P.regroup = o.regroup_interval; P.stepInterval = o.step_interval;
    P.chunksPerPoint = std::max(1, (pts->n_walks_max + o.walks_per_lane - 1) / o.walks_per_lane);
    P.nClaims = (unsigned long long)n * (unsigned long long)P.chunksPerPoint;
    P.seed = o.seed; P.streamOffset = o.stream_offset;
    P.countMode = o.count_visits;
    P.relax2 = (1.0f - o.distance_tolerance) * (1.0f - o.distance_tolerance);
    float e2 = 2.0f * ws->epsilon_shell_for_absorbing_boundary;
    P.guard2 = e2 * e2 / P.relax2;

    MergeParams& M = S.M;
    std::memset(&M, 0, sizeof(M));
    M.n = n; M.pt = pts->pt; M.type = pts->type; M.aligned = pts->estimate_boundary_normal_aligned;
    M.dist0 = pts->dist_to_absorbing_boundary;
    M.sN = stats->n_solution_estimates; M.sMean = stats->solution_mean; M.sM2 = stats->solution_m2;
    M.sLen = stats->total_walk_length; M.firstSphereRadius = pts->first_sphere_radius;
    if (s) { M.dirichlet = pde->dirichlet; M.absorption = pde->absorption_coeff; }
    else { M.instances = b->table; M.instanceOf = instanceOf; M.nInstances = b->n; }
    M.ignoreAbsorbing = ws->ignore_absorbing_boundary_contribution;
    M.solveDoubleSided = ws->solve_double_sided;
    M.writeFirstSphereRadius = 1;
    if (pts->estimation_quantity && b)
        return fail(ZG_ERR_UNSUPPORTED, "estimation_quantity is supported by zg_wos_solve and zg_wost_solve only");
    P.quantity = pts->estimation_quantity; M.quantity = pts->estimation_quantity;
    M.want = ZG_ESTIMATE_SOLUTION;
    P.dirDeriv = pts->direction_for_derivative;
    P.distR = pts->dist_to_reflecting_boundary;
    S.stats = *stats;

    S.dim = dim; S.width = s ? s->width : b->width; S.depth = s ? s->info.absorbing_depth : b->maxDepth;
    S.batch = b != nullptr; S.count = o.count_visits != 0; S.n = n; S.nWalksMax = pts->n_walks_max;
    if (s) S.screened = pde->absorption_coeff > 0.0f;
    else for (const zg_pde& ip : b->pdes) S.screened = S.screened || ip.absorption_coeff > 0.0f;
    S.words = (size_t)n * 4 + kCntSlots + (wost ? (size_t)(n + 1) / 2 : 0) + (split ? (size_t)n : 0);
    if (P.quantity) S.words += (size_t)(n + 1) / 2;   // masked walk counts, last in the workspace
    S.wost = wost;
    S.split = split;
    S.interleave = o.wost_interleave_queries != 0;
    M.sSplits = stats->total_splits;
    if (wost) {
        S.callerFsr = pts->first_sphere_radius;
        P.prevN = stats->n_solution_estimates;
    }
    int wavesPerCu = o.waves_per_cu > 0 ? o.waves_per_cu : 32;
    S.blocks = std::max(1, cuCount(dev) * wavesPerCu / 4);
    return ZG_OK;
}

void bindWorkspace(SolveSetup& S, uint64_t* wsp) {
    int64_t n = S.n;
    S.P.accSum = reinterpret_cast<double*>(wsp);
    S.P.accSq = reinterpret_cast<double*>(wsp + n);
    S.P.accN = reinterpret_cast<unsigned long long*>(wsp + 2 * n);
    S.P.accLen = reinterpret_cast<unsigned long long*>(wsp + 3 * n);
    S.P.counters = reinterpret_cast<unsigned long long*>(wsp + 4 * n);
    S.P.claimCounter = S.P.counters + kCntClaims;
    S.M.accSum = S.P.accSum; S.M.accSq = S.P.accSq; S.M.accN = S.P.accN; S.M.accLen = S.P.accLen;
    if (S.split) {   // after the walk-on-stars radii
        S.P.accSplits = reinterpret_cast<unsigned long long*>(wsp + 4 * n + kCntSlots + (S.wost ? (n + 1) / 2 : 0));
        S.M.accSplits = S.P.accSplits;
    }
    if (S.wost) {
        float* fsr = reinterpret_cast<float*>(wsp + 4 * n + kCntSlots);
        S.P.fsr = fsr;
        S.M.fsrSrc = fsr;
    }
}

// fill (a kernel, so the sequence can be captured in a graph), walk, merge
hipError_t enqueueSolve(SolveSetup& S, uint64_t* wsp, hipStream_t strm, hipEvent_t ev0, hipEvent_t ev1) {
    bindWorkspace(S, wsp);
    hipError_t e = launchFillZero(wsp, (int64_t)S.words, strm);
    if (e == hipSuccess && ev0) e = hipEventRecord(ev0, strm);
    WosParams P = S.P;
    if (e == hipSuccess && P.quantity) {
        // the walk kernels see walk counts with the non-solution points zeroed
        int32_t* masked = reinterpret_cast<int32_t*>(wsp + S.words - (S.n + 1) / 2);
        e = launchMaskWalks(P.nWalks, P.quantity, S.n, masked, strm);
        P.nWalks = masked;
        P.quantity = nullptr;
    }
    if (e == hipSuccess && S.wost) e = launchWostFirstSphere(P, S.dim, S.width, S.callerFsr, const_cast<float*>(P.fsr), strm);
    if (e == hipSuccess && S.nWalksMax > 0)
        e = S.wost ? launchWost(P, S.dim, S.width, S.count, S.split, S.interleave, S.blocks, strm)
            : S.batch ? launchWosBatch(P, S.width, S.count, S.screened, S.split, S.blocks, strm)
            : S.sdf ? launchWosSdf(P, S.dim, S.screened, S.split, S.blocks, strm)
                    : launchWos(P, S.dim, S.width, S.count, S.screened, S.split, S.blocks, strm);
    if (e == hipSuccess && ev1) e = hipEventRecord(ev1, strm);
    if (e == hipSuccess) e = launchMergeStats(S.M, S.dim, strm);
    return e;
}

zg_status checkErrorFlags(unsigned long long flags, int depth, int width) {
    if (flags & 2u)
        return fail(ZG_ERR_INVALID_ARGUMENT, "instance_of_point holds an index outside [0, n_instances); those points were skipped");
    if (flags & 1u) return fail(ZG_ERR_DEVICE, "BVH traversal stack overflow (depth %d, width %d)", depth, width);
    return ZG_OK;
}

void fillReport(const unsigned long long* cnt, float ms, zg_solve_report* report) {
    for (int k = 0; k < 4; k++) report->completion_codes[k] = (int64_t)cnt[k];
    report->total_steps = (int64_t)cnt[kCntSteps];
    report->total_walks = (int64_t)cnt[kCntWalks];
    report->node_visits = (int64_t)cnt[kCntNodes];
    report->prim_visits = (int64_t)cnt[kCntPrims];
    report->queries = (int64_t)cnt[kCntQueries];
    report->kernel_ms = ms;
    report->reserved = 0.0f;
    report->reflecting_node_visits = (int64_t)cnt[kCntReflNodes];
    report->reflecting_prim_visits = (int64_t)cnt[kCntReflPrims];
    report->silhouette_node_visits = (int64_t)cnt[kCntSilNodes];
    report->silhouette_prim_visits = (int64_t)cnt[kCntSilPrims];
}

// gradient layout after the solve workspace: sums and squares of the gradient estimates
// (n*dim each), first-step source and derivative totals (n each), then the stratified tables
struct GradLayout {
    int halves = 1, pairsMax = 0;
    int64_t stride = 0;     // floats per stratified table
    size_t words = 0;
};

GradLayout gradLayout(const SolveSetup& S) {
    GradLayout L;
    L.halves = S.P.ws.use_gradient_antithetic_variates ? 2 : 1;
    L.pairsMax = L.halves == 2 ? std::max(1, S.nWalksMax / 2) : S.nWalksMax;
    L.stride = (int64_t)(S.dim - 1) * 2 * L.pairsMax;
    L.words = (size_t)S.n * (2 * S.dim + 2) + ((size_t)S.n * (size_t)L.stride + 1) / 2;
    return L;
}

// estimateSolutionAndGradient for the points asking for it: stratified tables, then stages of
// antithetic pairs [0,1), [1,2), [2,4), ... (one stage without control variates); every stage
// reads the control variates from the statistics the previous stage merged
hipError_t enqueueGradient(SolveSetup& S, uint64_t* wsp, bool fillCounters, hipStream_t strm) {
    const int64_t n = S.n;
    const int D = S.dim;
    GradLayout L = gradLayout(S);
    uint64_t* g = wsp + S.words;
    WosParams P = S.P;
    MergeParams M = S.M;
    P.cvGrad = nullptr; P.cvGrid = nullptr;
    P.accG = reinterpret_cast<double*>(g);
    P.accG2 = reinterpret_cast<double*>(g + (size_t)n * D);
    P.accFirst = reinterpret_cast<double*>(g + (size_t)n * 2 * D);
    P.accDeriv = reinterpret_cast<double*>(g + (size_t)n * (2 * D + 1));
    P.strat = reinterpret_cast<const float*>(g + (size_t)n * (2 * D + 2));
    P.stratStride = L.stride;
    P.gradHalves = L.halves;
    P.cvMean = S.stats.solution_mean; P.cvN = S.stats.n_solution_estimates;
    P.cvFirst = S.stats.total_first_source_contribution;
    M.want = ZG_ESTIMATE_SOLUTION_AND_GRADIENT;
    M.accG = P.accG; M.accG2 = P.accG2; M.accFirst = P.accFirst; M.accDeriv = P.accDeriv;
    M.gN = S.stats.n_gradient_estimates; M.gMean = S.stats.gradient_mean; M.gM2 = S.stats.gradient_m2;
    M.gFirst = S.stats.total_first_source_contribution; M.gDeriv = S.stats.total_derivative_contribution;
    M.streamOffsetDev = nullptr;
    hipError_t e = hipSuccess;
    if (fillCounters) e = launchFillZero(reinterpret_cast<uint64_t*>(P.counters), kCntSlots, strm);
    if (e == hipSuccess)
        e = launchStratified(D - 1, const_cast<float*>(P.strat), n, L.stride, 2 * L.pairsMax, P.nWalks, P.quantity,
                             L.halves, P.seed, P.streamOffset, strm);
    if (e == hipSuccess && S.wost) e = launchWostGradFirstSphere(P, S.dim, S.width, const_cast<float*>(P.fsr), strm);
    const bool cv = S.P.ws.use_gradient_control_variates != 0;
    for (int u0 = 0; e == hipSuccess && u0 < L.pairsMax;) {
        int u1 = cv ? (u0 == 0 ? 1 : std::min(2 * u0, L.pairsMax)) : L.pairsMax;
        e = launchFillZero(wsp, n * 4, strm);   // solution sums, counts, lengths
        if (e == hipSuccess) e = launchFillZero(reinterpret_cast<uint64_t*>(P.claimCounter), 1, strm);
        if (e == hipSuccess) e = launchFillZero(g, n * (2 * D + 2), strm);
        if (e == hipSuccess && S.split) e = launchFillZero(reinterpret_cast<uint64_t*>(P.accSplits), n, strm);
        P.gradW0 = u0 * L.halves;
        P.gradW1 = u1 * L.halves;
        P.chunksPerPoint = std::max(1, (P.gradW1 - P.gradW0 + P.wpl - 1) / P.wpl);
        P.nClaims = (unsigned long long)n * (unsigned long long)P.chunksPerPoint;
        if (e == hipSuccess)
            e = S.wost ? launchWostGrad(P, S.dim, S.width, S.blocks, strm)
                : S.sdf ? launchWosSdfGrad(P, D, S.screened, S.split, S.blocks, strm)
                        : launchWosGrad(P, D, S.width, S.screened, S.split, S.blocks, strm);
        if (e == hipSuccess) e = launchMergeStats(M, D, strm);
        u0 = u1;
    }
    return e;
}

// eager solve: stream-ordered workspace, launch, counters back, synchronise
zg_status runSolve(SolveSetup& S, zg_solve_report* report, hipStream_t strm) {
    if (S.n == 0) { if (report) std::memset(report, 0, sizeof(*report)); return ok(); }
    // estimation quantities: which launches the call needs
    int64_t nSolve = S.n, nGrad = 0;
    if (S.P.quantity) {
        std::vector<uint8_t> hq((size_t)S.n);
        ZG_HIP(hipMemcpyAsync(hq.data(), S.P.quantity, (size_t)S.n, hipMemcpyDeviceToHost, strm));
        ZG_HIP(hipStreamSynchronize(strm));
        nSolve = 0;
        for (uint8_t q : hq) {
            if (q > ZG_ESTIMATE_SKIP) return fail(ZG_ERR_INVALID_ARGUMENT, "estimation_quantity value %d is invalid", (int)q);
            nSolve += q == ZG_ESTIMATE_SOLUTION;
            nGrad += q == ZG_ESTIMATE_SOLUTION_AND_GRADIENT;
        }
    }
    size_t words = S.words;
    if (nGrad > 0) {
        const zg_statistics& st = S.stats;
        if (!st.n_gradient_estimates || !st.gradient_mean || !st.gradient_m2 || !st.total_first_source_contribution ||
            !st.total_derivative_contribution)
            return fail(ZG_ERR_INVALID_ARGUMENT, "gradient points need the gradient members of zg_statistics");
        if (S.wost && S.split)
            return fail(ZG_ERR_UNSUPPORTED, "walk splitting with gradient points runs in walk on spheres only");
        words += gradLayout(S).words;
    }
    // the walk-on-stars kernel reads the PDE from a device copy after the workspace
    const size_t pdeWords = (sizeof(zg_pde) + 7) / 8;
    uint64_t* wsp = nullptr;
    ZG_HIP(hipMallocAsync((void**)&wsp, (words + pdeWords) * 8, strm));
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipError_t e = hipSuccess;
    if (!S.batch) {
        S.P.pdeDev = reinterpret_cast<const zg_pde*>(wsp + words);
        e = hipMemcpyAsync(wsp + words, &S.P.pde, sizeof(zg_pde), hipMemcpyHostToDevice, strm);
    }
    if (report && e == hipSuccess) {
        e = hipEventCreate(&ev0);
        if (e == hipSuccess) e = hipEventCreate(&ev1);
    }
    if (e == hipSuccess && nSolve > 0) e = enqueueSolve(S, wsp, strm, ev0, nGrad > 0 ? nullptr : ev1);
    if (e == hipSuccess && nGrad > 0) {
        if (nSolve == 0) {
            bindWorkspace(S, wsp);
            if (ev0) e = hipEventRecord(ev0, strm);
        }
        if (e == hipSuccess) e = enqueueGradient(S, wsp, nSolve == 0, strm);
        if (e == hipSuccess && ev1) e = hipEventRecord(ev1, strm);
    }
    if (e == hipSuccess && nSolve == 0 && nGrad == 0) {   // every point skipped
        bindWorkspace(S, wsp);
        e = launchFillZero(reinterpret_cast<uint64_t*>(S.P.counters), kCntSlots, strm);
        if (e == hipSuccess && ev0) e = hipEventRecord(ev0, strm);
        if (e == hipSuccess && ev1) e = hipEventRecord(ev1, strm);
    }
    if (!report && nGrad == 0) {
        // asynchronous: nothing waits for the device; the workspace is freed in stream order and
        // the solve's error flags are latched for zg_synchronize
        if (e == hipSuccess) e = launchLatchErrors(S.P.counters, strm);
        hipError_t ef = hipFreeAsync(wsp, strm);
        if (e == hipSuccess) e = ef;
        if (e != hipSuccess) return fail(ZG_ERR_DEVICE, "walk-on-spheres launch failed: %s", hipGetErrorString(e));
        return ok();
    }
    unsigned long long cnt[kCntSlots] = {0};
    if (e == hipSuccess) e = hipMemcpyAsync(cnt, S.P.counters, sizeof(cnt), hipMemcpyDeviceToHost, strm);
    if (e == hipSuccess) e = hipStreamSynchronize(strm);
    float ms = 0.0f;
    if (e == hipSuccess && report) e = hipEventElapsedTime(&ms, ev0, ev1);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    hipError_t ef = hipFreeAsync(wsp, strm);
    if (e == hipSuccess) e = ef;
    if (e != hipSuccess) return fail(ZG_ERR_DEVICE, "walk-on-spheres launch failed: %s", hipGetErrorString(e));
    zg_status est = checkErrorFlags(cnt[kCntError], S.depth, S.width);
    if (est != ZG_OK) return est;
    if (report) fillReport(cnt, ms, report);
    return ok();
}

}  // namespace

extern "C" {

zg_status zg_wos_solve(const zg_scene* s, const zg_pde* pde, const zg_walk_settings* ws, const zg_sample_points* pts,
                       zg_statistics* stats, const zg_solve_options* opts, zg_solve_report* report, void* stream) {
    if (!s) return fail(ZG_ERR_INVALID_ARGUMENT, "NULL scene");
    SolveSetup S;
    zg_status st = prepareSolve(s, nullptr, pde, ws, pts, nullptr, stats, opts, S);
    if (st != ZG_OK) return st;
    return runSolve(S, report, (hipStream_t)stream);
}

zg_status zg_wos_solve_noalloc(const zg_scene* s, const zg_pde* pde, const zg_walk_settings* ws,
                               const zg_sample_points* pts, zg_statistics* stats, const zg_solve_options* opts,
                               zg_solve_report* report, void* stream) {
    if (!s) return fail(ZG_ERR_INVALID_ARGUMENT, "NULL scene");
    SolveSetup S;
    zg_status st = prepareSolve(s, nullptr, pde, ws, pts, nullptr, stats, opts, S);
    if (st != ZG_OK) return st;
    if (S.P.quantity)
        return fail(ZG_ERR_UNSUPPORTED, "zg_wos_solve_noalloc estimates solutions only (estimation_quantity must be NULL)");
    const size_t cap = kNoallocScratchBytes / 8;
    if (S.words > cap)
        return fail(ZG_ERR_INVALID_ARGUMENT,
                    "%lld points need %zu bytes of workspace; zg_wos_solve_noalloc holds %zu (at most %lld points%s)",
                    (long long)S.n, S.words * 8, kNoallocScratchBytes,
                    (long long)((cap - kCntSlots) / (S.split ? 5 : 4)), S.split ? " with walk splitting" : "");
    if (S.n == 0) { if (report) std::memset(report, 0, sizeof(*report)); return ok(); }
    unsigned long long* wsp = nullptr;
    hipStream_t strm = (hipStream_t)stream;
    hipError_t e = noallocScratch(&wsp);
    if (e != hipSuccess) return fail(ZG_ERR_DEVICE, "workspace lookup failed: %s", hipGetErrorString(e));
    // the walk-on-spheres kernels read the PDE from their arguments (no device copy needed)
    e = enqueueSolve(S, reinterpret_cast<uint64_t*>(wsp), strm, nullptr, nullptr);
    unsigned long long cnt[kCntSlots] = {0};
    if (e == hipSuccess) e = hipMemcpyAsync(cnt, S.P.counters, sizeof(cnt), hipMemcpyDeviceToHost, strm);
    if (e == hipSuccess) e = hipStreamSynchronize(strm);
    if (e != hipSuccess) return fail(ZG_ERR_DEVICE, "walk-on-spheres launch failed: %s", hipGetErrorString(e));
    st = checkErrorFlags(cnt[kCntError], S.depth, S.width);
    if (st != ZG_OK) return st;
    if (report) fillReport(cnt, 0.0f, report);
    return ok();
}

zg_status zg_bvc_splat(int32_t dim, float absorption, const zg_bvc_samples* S, const zg_bvc_eval_points* E,
                       float radiusClamp, float kreg, float robinCutoff, float cutoffA, float cutoffR,
                       zg_bvc_statistics* out, void* stream) {
    if (!S || !E || !out) return fail(ZG_ERR_INVALID_ARGUMENT, "NULL argument");
    if (dim != 2 && dim != 3) return fail(ZG_ERR_INVALID_ARGUMENT, "dim must be 2 or 3");
    if (!(absorption >= 0.0f)) return fail(ZG_ERR_INVALID_ARGUMENT, "absorption must be >= 0");
    if (absorption > 0.0f && dim == 2)
        return fail(ZG_ERR_UNSUPPORTED, "screened boundary value caching (absorption > 0) is supported in 3D only");
    if (S->n < 0 || E->n < 0) return fail(ZG_ERR_INVALID_ARGUMENT, "negative counts");
    if (E->n >= (1ll << 31)) return fail(ZG_ERR_INVALID_ARGUMENT, "at most 2^31-1 evaluation points");
    if (S->n > 0 && (!S->pt || !S->normal || !S->type || !S->pdf || !S->solution || !S->normal_derivative))
        return fail(ZG_ERR_INVALID_ARGUMENT, "sample pt, normal, type, pdf, solution and normal_derivative are required");
    if (E->n > 0 && (!E->pt || !E->dist_to_absorbing_boundary))
        return fail(ZG_ERR_INVALID_ARGUMENT, "evaluation pt and dist_to_absorbing_boundary are required");
    if (E->n > 0 && (!out->n || !out->solution_mean || !out->solution_m2 || !out->gradient_mean || !out->gradient_m2))
        return fail(ZG_ERR_INVALID_ARGUMENT, "statistics arrays are required");
    if (!(radiusClamp >= 0.0f) || !(kreg >= 0.0f)) return fail(ZG_ERR_INVALID_ARGUMENT, "radius clamp and kernel regularization must be >= 0");
    hipError_t e = launchBvcSplat(dim, absorption, *S, *E, *out, radiusClamp, kreg, robinCutoff, cutoffA, cutoffR,
                                  (hipStream_t)stream);
    if (e != hipSuccess) return fail(ZG_ERR_DEVICE, "zg_bvc_splat: %s", hipGetErrorString(e));
    return ok();
}

zg_status zg_spatially_sorted_indices(int32_t dim, const float* pts, int64_t n, int32_t* out) {
    if (dim != 2 && dim != 3) return fail(ZG_ERR_INVALID_ARGUMENT, "dim must be 2 or 3, got %d", dim);
    if (n < 0 || n >= (1ll << 31) || (n > 0 && (!pts || !out)))
        return fail(ZG_ERR_INVALID_ARGUMENT, "invalid point array (n = %lld)", (long long)n);
    for (int64_t k = 0; k < n * dim; k++)
        if (!std::isfinite(pts[k])) return fail(ZG_ERR_INVALID_ARGUMENT, "coordinate %lld is not finite", (long long)k);
    spatiallySortedIndices(dim, pts, n, out);
    return ok();
}

zg_status zg_stratified_samples(int32_t dims, int64_t nTables, int32_t nSamples, uint64_t seed, uint32_t streamOffset,
                                float* out, void* stream) {
    if (dims < 1 || dims > 2) return fail(ZG_ERR_INVALID_ARGUMENT, "dims must be 1 or 2");
    if (nTables < 0 || nSamples < 1) return fail(ZG_ERR_INVALID_ARGUMENT, "n_tables >= 0 and n_samples >= 1 required");
    if (nTables >= (1ll << 31)) return fail(ZG_ERR_INVALID_ARGUMENT, "at most 2^31-1 tables");
    if (nTables > 0 && !out) return fail(ZG_ERR_INVALID_ARGUMENT, "NULL output");
    hipError_t e = launchStratified(dims, out, nTables, (int64_t)dims * nSamples, nSamples, nullptr, nullptr, 1, seed,
                                    streamOffset, (hipStream_t)stream);
    if (e != hipSuccess) return fail(ZG_ERR_DEVICE, "zg_stratified_samples: %s", hipGetErrorString(e));
    return ok();
}

zg_status zg_batch_create(int32_t nInstances, const zg_scene* const* scenes, const zg_pde* pdes, zg_batch** out) {
    if (!out || !scenes || !pdes || nInstances < 1) return fail(ZG_ERR_INVALID_ARGUMENT, "invalid arguments");
    *out = nullptr;
    zg_batch* b = new zg_batch();
    b->n = nInstances;
    std::vector<InstanceDesc> host(nInstances);
    for (int i = 0; i < nInstances; i++) {
        const zg_scene* s = scenes[i];
        if (!s) { delete b; return fail(ZG_ERR_INVALID_ARGUMENT, "scene %d is NULL", i); }
        if (s->sdfGrid) { delete b; return fail(ZG_ERR_UNSUPPORTED, "batched solves take mesh scenes"); }
        if (s->dim != 3) { delete b; return fail(ZG_ERR_UNSUPPORTED, "batched solves are 3D"); }
        if (i > 0 && (s->width != b->width || s->device != b->device)) {
            delete b;
            return fail(ZG_ERR_INVALID_ARGUMENT, "all scenes of a batch need the same BVH width and device");
        }
        zg_status st = checkPde(&pdes[i], 3);
        if (st != ZG_OK) { delete b; return st; }
        b->width = s->width; b->device = s->device; b->maxDepth = std::max(b->maxDepth, s->info.absorbing_depth);
        InstanceDesc& d = host[i];
        std::memset(&d, 0, sizeof(d));
        d.nodes = s->view.abs.nodes; d.prims = s->view.abs.prims; d.nNodes = s->view.abs.nNodes; d.nPrims = s->view.abs.nPrims;
        for (int a = 0; a < 3; a++) { d.dmin[a] = s->view.dmin[a]; d.dmax[a] = s->view.dmax[a]; }
        d.pde = pdes[i];
        b->scenes.push_back(s);
        b->pdes.push_back(pdes[i]);
    }
    hipError_t e = hipMalloc((void**)&b->table, sizeof(InstanceDesc) * nInstances);
    if (e == hipSuccess) e = hipMemcpy(b->table, host.data(), sizeof(InstanceDesc) * nInstances, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        if (b->table) (void)hipFree(b->table);
        delete b;
        return fail(ZG_ERR_DEVICE, "zg_batch_create: %s", hipGetErrorString(e));
    }
    *out = b;
    return ok();
}

zg_status zg_batch_destroy(zg_batch* b) {
    if (!b) return ok();
    if (b->table) (void)hipFree(b->table);
    delete b;
    return ok();
}

zg_status zg_batch_wos_solve(const zg_batch* b, const zg_walk_settings* ws, const zg_sample_points* pts,
                             const int32_t* instanceOf, zg_statistics* stats, const zg_solve_options* opts,
                             zg_solve_report* report, void* stream) {
    if (!b) return fail(ZG_ERR_INVALID_ARGUMENT, "NULL batch");
    SolveSetup S;
    zg_status st = prepareSolve(nullptr, b, nullptr, ws, pts, instanceOf, stats, opts, S);
    if (st != ZG_OK) return st;
    return runSolve(S, report, (hipStream_t)stream);
}

zg_status zg_batch_wos_graph_create(const zg_batch* b, const zg_walk_settings* ws, const zg_sample_points* pts,
                                    const int32_t* instanceOf, zg_statistics* stats, const zg_solve_options* opts,
                                    zg_graph** out) {
    if (!b || !out) return fail(ZG_ERR_INVALID_ARGUMENT, "NULL argument");
    *out = nullptr;
    SolveSetup S;
    zg_status st = prepareSolve(nullptr, b, nullptr, ws, pts, instanceOf, stats, opts, S);
    if (st != ZG_OK) return st;
    if (S.n == 0) return fail(ZG_ERR_INVALID_ARGUMENT, "empty point set");
    zg_graph* g = new zg_graph();
    g->n = S.n; g->depth = S.depth; g->width = S.width;
    hipStream_t cs = nullptr;
    hipError_t e = hipMalloc((void**)&g->workspace, S.words * 8);
    if (e == hipSuccess) e = hipMalloc((void**)&g->streamOffset, sizeof(uint32_t));
    if (e == hipSuccess) e = hipMemcpy(g->streamOffset, &S.P.streamOffset, sizeof(uint32_t), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&cs, hipStreamNonBlocking);
    if (e == hipSuccess) {
        // every replay reads its PCG32 stream offset from device memory and the merge kernel
        // advances it by n_walks_max, so successive replays draw fresh streams
        S.P.streamOffsetDev = g->streamOffset;
        S.M.streamOffsetDev = g->streamOffset;
        S.M.streamAdvance = (uint32_t)S.nWalksMax;
        e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
        if (e == hipSuccess) {
            hipError_t ek = enqueueSolve(S, g->workspace, cs, nullptr, nullptr);
            e = hipStreamEndCapture(cs, &g->graph);
            if (e == hipSuccess) e = ek;
        }
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
    if (cs) (void)hipStreamDestroy(cs);
    if (e != hipSuccess) {
        if (g->exec) (void)hipGraphExecDestroy(g->exec);
        if (g->graph) (void)hipGraphDestroy(g->graph);
        if (g->workspace) (void)hipFree(g->workspace);
        if (g->streamOffset) (void)hipFree(g->streamOffset);
        delete g;
        return fail(ZG_ERR_DEVICE, "graph capture failed: %s", hipGetErrorString(e));
    }
    *out = g;
    return ok();
}

zg_status zg_graph_launch(zg_graph* g, void* stream) {
    if (!g) return fail(ZG_ERR_INVALID_ARGUMENT, "NULL graph");
    ZG_HIP(hipGraphLaunch(g->exec, (hipStream_t)stream));
    return ok();
}

zg_status zg_graph_read_report(zg_graph* g, zg_solve_report* report, void* stream) {
    if (!g || !report) return fail(ZG_ERR_INVALID_ARGUMENT, "NULL argument");
    unsigned long long cnt[kCntSlots] = {0};
    hipStream_t s = (hipStream_t)stream;
    ZG_HIP(hipMemcpyAsync(cnt, g->workspace + 4 * g->n, sizeof(cnt), hipMemcpyDeviceToHost, s));
    ZG_HIP(hipStreamSynchronize(s));
    zg_status est = checkErrorFlags(cnt[kCntError], g->depth, g->width);
    if (est != ZG_OK) return est;
    fillReport(cnt, 0.0f, report);
    return ok();
}

zg_status zg_graph_destroy(zg_graph* g) {
    if (!g) return ok();
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    if (g->workspace) (void)hipFree(g->workspace);
    if (g->streamOffset) (void)hipFree(g->streamOffset);
    delete g;
    return ok();
}


// ------------------------------------------------------------------------------------
// learned control variate (MLP)
// ------------------------------------------------------------------------------------
static void mlpFree(zg_mlp* m) {
    for (void* p : m->allocations) (void)hipFree(p);
    m->allocations.clear();
}

}  // extern "C"

template <class T>
static hipError_t mlpAlloc(zg_mlp* m, T** p, int64_t count) {
    *p = nullptr;
    hipError_t e = hipMalloc((void**)p, (size_t)std::max<int64_t>(count, 1) * sizeof(T));
    if (e == hipSuccess) m->allocations.push_back((void*)*p);
    return e;
}

extern "C" {

zg_status zg_mlp_create(const zg_mlp_config* cfg, int64_t maxBatch, zg_mlp** out) {
    if (!cfg || !out) return fail(ZG_ERR_INVALID_ARGUMENT, "NULL argument");
    *out = nullptr;
    if (cfg->in_dim != 2 && cfg->in_dim != 3) return fail(ZG_ERR_INVALID_ARGUMENT, "in_dim must be 2 or 3");
    if (cfg->n_frequencies < 1 || cfg->n_frequencies > 256) return fail(ZG_ERR_INVALID_ARGUMENT, "n_frequencies must be 1..256");
    if (cfg->hidden < 1 || cfg->hidden > 512) return fail(ZG_ERR_INVALID_ARGUMENT, "hidden must be 1..512");
    if (cfg->n_hidden_layers < 1 || cfg->n_hidden_layers > kMaxHidden) return fail(ZG_ERR_INVALID_ARGUMENT, "n_hidden_layers must be 1..8");
    if (maxBatch < 1 || maxBatch > (1ll << 26)) return fail(ZG_ERR_INVALID_ARGUMENT, "max_batch must be 1..2^26");
    zg_mlp* m = new zg_mlp();
    m->cfg = *cfg;
    MlpDevice& d = m->d;
    d.shape = MlpShape{cfg->in_dim, cfg->n_frequencies, cfg->hidden, cfg->n_hidden_layers};
    int L = cfg->n_hidden_layers, H = cfg->hidden, F2 = 2 * cfg->n_frequencies;
    int64_t off = 0;
    for (int l = 0; l <= L; l++) {
        int in = l == 0 ? F2 : H, outd = l == L ? 1 : H;
        d.wOffset[l] = off; off += (int64_t)in * outd;
        d.bOffset[l] = off; off += outd;
    }
    d.nParams = off;
    d.maxRows = maxBatch;
    // host initialisation: PyTorch-style U(-1/sqrt(in), 1/sqrt(in)) weights and biases,
    // Gaussian frequencies
    std::mt19937_64 gen(cfg->seed);
    std::vector<float> theta(off), freq((size_t)cfg->n_frequencies * cfg->in_dim);
    std::vector<int64_t> bfIndex(off, -1);
    for (int l = 0; l <= L; l++) {
        int in = l == 0 ? F2 : H, outd = l == L ? 1 : H;
        std::uniform_real_distribution<float> u(-1.0f / std::sqrt((float)in), 1.0f / std::sqrt((float)in));
        for (int64_t i = 0; i < (int64_t)in * outd; i++) {
            theta[d.wOffset[l] + i] = u(gen);
            if (l < L) bfIndex[d.wOffset[l] + i] = d.wOffset[l] + i;
        }
        for (int i = 0; i < outd; i++) theta[d.bOffset[l] + i] = u(gen);
    }
    std::normal_distribution<float> nd(0.0f, cfg->frequency_sigma);
    for (auto& f : freq) f = nd(gen);
    hipError_t e = hipGetDevice(&m->device);
    if (e == hipSuccess) e = mlpAlloc(m, &d.theta, off);
    if (e == hipSuccess) e = mlpAlloc(m, &d.grad, off);
    if (e == hipSuccess) e = mlpAlloc(m, &d.adamM, off);
    if (e == hipSuccess) e = mlpAlloc(m, &d.adamV, off);
    if (e == hipSuccess) e = mlpAlloc(m, &d.wBf, off);
    if (e == hipSuccess) e = mlpAlloc(m, &d.bfIndex, off);
    if (e == hipSuccess) e = mlpAlloc(m, &d.freq, (int64_t)freq.size());
    if (e == hipSuccess) e = mlpAlloc(m, &d.phi, maxBatch * F2);
    if (e == hipSuccess) e = mlpAlloc(m, &d.dphi, maxBatch * F2);
    for (int l = 0; l < L && e == hipSuccess; l++) {
        e = mlpAlloc(m, &d.pre[l], maxBatch * H);
        if (e == hipSuccess) e = mlpAlloc(m, &d.act[l], maxBatch * H);
    }
    if (e == hipSuccess) e = mlpAlloc(m, &d.dz[0], maxBatch * H);
    if (e == hipSuccess) e = mlpAlloc(m, &d.dz[1], maxBatch * H);
    if (e == hipSuccess) e = mlpAlloc(m, &d.loss, 1);
    if (e == hipSuccess) e = hipMemcpy(d.theta, theta.data(), off * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d.bfIndex, bfIndex.data(), off * 8, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d.freq, freq.data(), freq.size() * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(d.adamM, 0, off * 4);
    if (e == hipSuccess) e = hipMemset(d.adamV, 0, off * 4);
    if (e == hipSuccess) e = mlpCast(d, 0);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) {
        mlpFree(m);
        delete m;
        return fail(e == hipErrorOutOfMemory ? ZG_ERR_OUT_OF_MEMORY : ZG_ERR_DEVICE, "zg_mlp_create: %s", hipGetErrorString(e));
    }
    *out = m;
    return ok();
}

zg_status zg_mlp_destroy(zg_mlp* m) {
    if (!m) return ok();
    mlpFree(m);
    delete m;
    return ok();
}

zg_status zg_mlp_parameter_count(const zg_mlp* m, int64_t* count) {
    if (!m || !count) return fail(ZG_ERR_INVALID_ARGUMENT, "NULL argument");
    *count = m->d.nParams;
    return ok();
}

zg_status zg_mlp_get_parameters(const zg_mlp* m, float* host, int64_t count) {
    if (!m || !host || count != m->d.nParams) return fail(ZG_ERR_INVALID_ARGUMENT, "count must equal the parameter count");
    ZG_HIP(hipMemcpy(host, m->d.theta, count * 4, hipMemcpyDeviceToHost));
    return ok();
}

zg_status zg_mlp_set_parameters(zg_mlp* m, const float* host, int64_t count) {
    if (!m || !host || count != m->d.nParams) return fail(ZG_ERR_INVALID_ARGUMENT, "count must equal the parameter count");
    ZG_HIP(hipMemcpy(m->d.theta, host, count * 4, hipMemcpyHostToDevice));
    ZG_HIP(mlpCast(m->d, 0));
    ZG_HIP(hipDeviceSynchronize());
    return ok();
}

zg_status zg_mlp_get_frequencies(const zg_mlp* m, float* host, int64_t count) {
    if (!m || !host || count != (int64_t)m->cfg.n_frequencies * m->cfg.in_dim)
        return fail(ZG_ERR_INVALID_ARGUMENT, "count must equal n_frequencies * in_dim");
    ZG_HIP(hipMemcpy(host, m->d.freq, count * 4, hipMemcpyDeviceToHost));
    return ok();
}

static zg_status mlpGradientsImpl(zg_mlp* m, const float* x, const float* y, int64_t batch, hipStream_t s) {
    if (!m || !x || !y) return fail(ZG_ERR_INVALID_ARGUMENT, "NULL argument");
    if (batch < 1 || batch > m->d.maxRows) return fail(ZG_ERR_INVALID_ARGUMENT, "batch must be 1..max_batch");
    ZG_HIP(mlpZero(m->d.grad, m->d.nParams, s));
    ZG_HIP(mlpZero(m->d.loss, 1, s));
    ZG_HIP(mlpForward(m->d, x, (int)batch, nullptr, s));
    ZG_HIP(mlpBackward(m->d, y, (int)batch, 1, m->d.loss, nullptr, s));
    return ZG_OK;
}

zg_status zg_mlp_gradients(zg_mlp* m, const float* x, const float* y, int64_t batch, float* gradHost, float* lossHost,
                           void* stream) {
    hipStream_t s = (hipStream_t)stream;
    zg_status st = mlpGradientsImpl(m, x, y, batch, s);
    if (st != ZG_OK) return st;
    if (gradHost) { ZG_HIP(hipMemcpyAsync(gradHost, m->d.grad, m->d.nParams * 4, hipMemcpyDeviceToHost, s)); }
    if (lossHost) { ZG_HIP(hipMemcpyAsync(lossHost, m->d.loss, 4, hipMemcpyDeviceToHost, s)); }
    if (gradHost || lossHost) { ZG_HIP(hipStreamSynchronize(s)); }
    return ok();
}

zg_status zg_mlp_train_step(zg_mlp* m, const float* x, const float* y, int64_t batch, float lr, float* lossHost,
                            void* stream) {
    hipStream_t s = (hipStream_t)stream;
    zg_status st = mlpGradientsImpl(m, x, y, batch, s);
    if (st != ZG_OK) return st;
    ZG_HIP(mlpAdam(m->d, lr, s));
    if (lossHost) {
        ZG_HIP(hipMemcpyAsync(lossHost, m->d.loss, 4, hipMemcpyDeviceToHost, s));
        ZG_HIP(hipStreamSynchronize(s));
    }
    return ok();
}

zg_status zg_mlp_evaluate(zg_mlp* m, const float* x, int64_t n, float* value, float* gradient, void* stream) {
    if (!m || (n > 0 && !x)) return fail(ZG_ERR_INVALID_ARGUMENT, "NULL argument");
    hipStream_t s = (hipStream_t)stream;
    int d = m->cfg.in_dim;
    for (int64_t r0 = 0; r0 < n; r0 += m->d.maxRows) {
        int rows = (int)std::min<int64_t>(m->d.maxRows, n - r0);
        const float* xr = x + r0 * d;
        ZG_HIP(mlpForward(m->d, xr, rows, nullptr, s));
        // mode 2 writes c(x) and back-propagates dc/dc = 1 down to the Fourier features
        ZG_HIP(mlpBackward(m->d, nullptr, rows, gradient ? 2 : 0, nullptr, value ? value + r0 : nullptr, s));
        if (gradient) { ZG_HIP(mlpInputGrad(m->d, xr, rows, gradient + r0 * d, s)); }
    }
    return ok();
}

zg_status zg_wost_solve(const zg_scene* s, const zg_pde* pde, const zg_walk_settings* ws, const zg_sample_points* pts,
                        zg_statistics* stats, const zg_solve_options* opts, zg_solve_report* report, void* stream) {
    if (!s) return fail(ZG_ERR_INVALID_ARGUMENT, "NULL scene");
    SolveSetup S;
    zg_status st = prepareSolve(s, nullptr, pde, ws, pts, nullptr, stats, opts, S, true);
    if (st != ZG_OK) return st;
    return runSolve(S, report, (hipStream_t)stream);
}

}  // extern "C"
