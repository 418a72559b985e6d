// api.hip — the C ABI (include/helios_hip.h): scene packing, BVH builds,
// integrator frame protocol, surfaces, tonemap, batching, debug tracing.
//
// Host-side restatement of:
//   Scene::create_gpu_resources   src/engine/resource/scene.cpp:915-1311
//   Scene::update (light count)   src/engine/resource/scene.cpp:875-911
//   Mesh::Mesh (BLAS per mesh)    src/engine/resource/mesh.cpp:46-117
//   Renderer::render (TLAS, clear) src/engine/gfx/renderer.cpp:107-223
//   PathIntegrator::render / launch_rays / compute_tile_coords
//                                 src/engine/gfx/path_integrator.cpp:48-84,125-200,312-336
//   CameraNode::update            src/engine/resource/scene.cpp:632-646
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "../../include/helios_hip.h"
#include "gpu_scene.hpp"
#include "internal.hpp"

using namespace hh;

// ---------------------------------------------------------------------------
// Errors
// ---------------------------------------------------------------------------
static thread_local std::string g_err;

static hh_status fail(hh_status s, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return s;
}
// Lets the host-only translation units (image and scene files) set the
// thread's error string.
hh_status hh_internal_fail(hh_status s, const char* msg) { return fail(s, "%s", msg); }
static hh_status hip_fail(hipError_t e, const char* what) {
    (void)hipGetLastError();
    return fail(e == hipErrorOutOfMemory ? HH_ERR_OUT_OF_MEMORY : HH_ERR_DEVICE, "%s: %s", what, hipGetErrorString(e));
}
#define HH_HIP(expr)                                         \
    do {                                                     \
        hipError_t _e = (expr);                              \
        if (_e != hipSuccess) return hip_fail(_e, #expr);    \
    } while (0)
#define HH_ARG(cond, msg)                                    \
    do {                                                     \
        if (!(cond)) return fail(HH_ERR_INVALID_ARG, msg);   \
    } while (0)

// ---------------------------------------------------------------------------
// Objects
// ---------------------------------------------------------------------------
struct GraphEntry {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

struct hh_context_t {
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t side = nullptr;                       // shadow launches beside the next extension launch
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    // Pinned staging of a frame's scene table and view records (one DMA each,
    // no pageable bounce buffer) and of the ray counters read back after it;
    // ev_staged follows the last upload, so the host waits for it before it
    // rewrites the staging. last_scenes: the scene table on the device, whose
    // upload is skipped while it is unchanged.
    unsigned char* h_stage = nullptr;
    size_t stage_cap = 0;
    hipEvent_t ev_staged = nullptr;
    std::vector<SceneGPU> last_scenes;
    unsigned long long* h_stats = nullptr;
    int num_cu = 0;
    Tuning tune;
    Wavefront wf;
    BuildScratch bs;
    std::map<std::string, GraphEntry> graphs;
    SceneGPU* d_scene_table = nullptr;
    uint32_t scene_table_cap = 0;
    ViewGPU* d_views = nullptr;
    uint32_t views_cap = 0;
    void* dbg = nullptr;
    size_t dbg_cap = 0;
    float* d_coef = nullptr;
    bool use_graphs = true;
    Prof prof;

    void drop_graphs() {
        for (auto& kv : graphs) {
            if (kv.second.exec) (void)hipGraphExecDestroy(kv.second.exec);
            if (kv.second.graph) (void)hipGraphDestroy(kv.second.graph);
        }
        graphs.clear();
    }
};

struct hh_denoiser_t {
    hh_context ctx = nullptr;
    dn::Net net;
    float* d_feat = nullptr;     // staged raw records
    float* d_target = nullptr;   // staged targets
    size_t stage_px = 0;
    uint32_t loss_kind = HH_DENOISER_LOSS_LOG_MSE;
    ~hh_denoiser_t() {
        if (d_feat) (void)hipFree(d_feat);
        if (d_target) (void)hipFree(d_target);
    }
};

struct TexHost {
    uint32_t w = 0, h = 0, fmt = 0;
    std::vector<float> texels;
};
struct MeshHost {
    std::vector<hh_vertex> v;
    std::vector<uint32_t> idx;
    std::vector<hh_submesh> sm;
    std::vector<uint32_t> mats;
};

struct hh_scene_t {
    hh_context ctx = nullptr;
    std::vector<TexHost> textures;
    std::vector<hh_material_desc> materials;
    std::vector<MeshHost> meshes;
    std::vector<hh_instance_desc> instances;
    std::vector<hh_light_desc> lights;
    std::vector<float> ibl;
    uint32_t ibl_size = 0;
    hh_sky_params sky{4.0f, 0.1f, 1.15f, 512};
    uint64_t generation = 0;
    bool committed = false;
    bool dirty = true;
    std::vector<void*> allocs;
    SceneGPU gpu{};
    SceneGPU* d_gpu = nullptr;
    std::vector<LightData> light_data;
    float sky_coef[30] = {0};
    uint64_t bvh_stats[4] = {0, 0, 0, 0};

    void free_device() {
        for (void* p : allocs) (void)hipFree(p);
        allocs.clear();
        d_gpu = nullptr;
        committed = false;
    }
};

struct hh_surface_t {
    hh_context ctx = nullptr;
    uint32_t w = 0, h = 0;
    float4* d = nullptr;
    float4* aux = nullptr;  // two feature planes (hh_surface_enable_features)
};

struct hh_integrator_t {
    hh_context ctx = nullptr;
    hh_settings s{7, 5000, 0.0f, 0, 0, 0};
    uint32_t num_acc = 0;
    uint32_t tile_idx = 0;
    std::vector<std::pair<uint32_t, uint32_t>> tiles;
    uint32_t tiles_for_w = 0, tiles_for_h = 0, tiles_for_tiled = 2;
    uint64_t last_generation = ~0ull;
    const void* last_scene = nullptr;
    uint64_t ray_stats[3] = {0, 0, 0};
    ViewGPU* d_view = nullptr;
};

static uint64_t g_generation = 1;

// ---------------------------------------------------------------------------
// Host math (glm semantics evaluated in double)
// ---------------------------------------------------------------------------
namespace {
struct M4 {
    double m[4][4];  // [col][row]
};
M4 ident() {
    M4 r{};
    for (int i = 0; i < 4; i++) r.m[i][i] = 1.0;
    return r;
}
M4 mul(const M4& a, const M4& b) {
    M4 r{};
    for (int c = 0; c < 4; c++)
        for (int rr = 0; rr < 4; rr++) {
            double s = 0;
            for (int k = 0; k < 4; k++) s += a.m[k][rr] * b.m[c][k];
            r.m[c][rr] = s;
        }
    return r;
}
M4 inverse(const M4& a) {
    double m[16], inv[16];
    for (int c = 0; c < 4; c++)
        for (int r = 0; r < 4; r++) m[c * 4 + r] = a.m[c][r];
    inv[0] = m[5] * m[10] * m[15] - m[5] * m[11] * m[14] - m[9] * m[6] * m[15] + m[9] * m[7] * m[14] + m[13] * m[6] * m[11] - m[13] * m[7] * m[10];
    inv[4] = -m[4] * m[10] * m[15] + m[4] * m[11] * m[14] + m[8] * m[6] * m[15] - m[8] * m[7] * m[14] - m[12] * m[6] * m[11] + m[12] * m[7] * m[10];
    inv[8] = m[4] * m[9] * m[15] - m[4] * m[11] * m[13] - m[8] * m[5] * m[15] + m[8] * m[7] * m[13] + m[12] * m[5] * m[11] - m[12] * m[7] * m[9];
    inv[12] = -m[4] * m[9] * m[14] + m[4] * m[10] * m[13] + m[8] * m[5] * m[14] - m[8] * m[6] * m[13] - m[12] * m[5] * m[10] + m[12] * m[6] * m[9];
    inv[1] = -m[1] * m[10] * m[15] + m[1] * m[11] * m[14] + m[9] * m[2] * m[15] - m[9] * m[3] * m[14] - m[13] * m[2] * m[11] + m[13] * m[3] * m[10];
    inv[5] = m[0] * m[10] * m[15] - m[0] * m[11] * m[14] - m[8] * m[2] * m[15] + m[8] * m[3] * m[14] + m[12] * m[2] * m[11] - m[12] * m[3] * m[10];
    inv[9] = -m[0] * m[9] * m[15] + m[0] * m[11] * m[13] + m[8] * m[1] * m[15] - m[8] * m[3] * m[13] - m[12] * m[1] * m[11] + m[12] * m[3] * m[9];
    inv[13] = m[0] * m[9] * m[14] - m[0] * m[10] * m[13] - m[8] * m[1] * m[14] + m[8] * m[2] * m[13] + m[12] * m[1] * m[10] - m[12] * m[2] * m[9];
    inv[2] = m[1] * m[6] * m[15] - m[1] * m[7] * m[14] - m[5] * m[2] * m[15] + m[5] * m[3] * m[14] + m[13] * m[2] * m[7] - m[13] * m[3] * m[6];
    inv[6] = -m[0] * m[6] * m[15] + m[0] * m[7] * m[14] + m[4] * m[2] * m[15] - m[4] * m[3] * m[14] - m[12] * m[2] * m[7] + m[12] * m[3] * m[6];
    inv[10] = m[0] * m[5] * m[15] - m[0] * m[7] * m[13] - m[4] * m[1] * m[15] + m[4] * m[3] * m[13] + m[12] * m[1] * m[7] - m[12] * m[3] * m[5];
    inv[14] = -m[0] * m[5] * m[14] + m[0] * m[6] * m[13] + m[4] * m[1] * m[14] - m[4] * m[2] * m[13] - m[12] * m[1] * m[6] + m[12] * m[2] * m[5];
    inv[3] = -m[1] * m[6] * m[11] + m[1] * m[7] * m[10] + m[5] * m[2] * m[11] - m[5] * m[3] * m[10] - m[9] * m[2] * m[7] + m[9] * m[3] * m[6];
    inv[7] = m[0] * m[6] * m[11] - m[0] * m[7] * m[10] - m[4] * m[2] * m[11] + m[4] * m[3] * m[10] + m[8] * m[2] * m[7] - m[8] * m[3] * m[6];
    inv[11] = -m[0] * m[5] * m[11] + m[0] * m[7] * m[9] + m[4] * m[1] * m[11] - m[4] * m[3] * m[9] - m[8] * m[1] * m[7] + m[8] * m[3] * m[5];
    inv[15] = m[0] * m[5] * m[10] - m[0] * m[6] * m[9] - m[4] * m[1] * m[10] + m[4] * m[2] * m[9] + m[8] * m[1] * m[6] - m[8] * m[2] * m[5];
    double det = m[0] * inv[0] + m[1] * inv[4] + m[2] * inv[8] + m[3] * inv[12];
    M4 r{};
    for (int c = 0; c < 4; c++)
        for (int rr = 0; rr < 4; rr++) r.m[c][rr] = inv[c * 4 + rr] / det;
    return r;
}
M4 from_glm(const float f[16]) {
    M4 r{};
    for (int c = 0; c < 4; c++)
        for (int rr = 0; rr < 4; rr++) r.m[c][rr] = f[c * 4 + rr];
    return r;
}
M4 quat_m4(const float q[4]) {
    double x = q[0], y = q[1], z = q[2], w = q[3];
    M4 r = ident();
    r.m[0][0] = 1 - 2 * (y * y + z * z);
    r.m[0][1] = 2 * (x * y + w * z);
    r.m[0][2] = 2 * (x * z - w * y);
    r.m[1][0] = 2 * (x * y - w * z);
    r.m[1][1] = 1 - 2 * (x * x + z * z);
    r.m[1][2] = 2 * (y * z + w * x);
    r.m[2][0] = 2 * (x * z + w * y);
    r.m[2][1] = 2 * (y * z - w * x);
    r.m[2][2] = 1 - 2 * (x * x + y * y);
    return r;
}
void quat_rot(const float q[4], const double v[3], double o[3]) {
    double qx = q[0], qy = q[1], qz = q[2], qw = q[3];
    double uv[3] = {qy * v[2] - qz * v[1], qz * v[0] - qx * v[2], qx * v[1] - qy * v[0]};
    double uuv[3] = {qy * uv[2] - qz * uv[1], qz * uv[0] - qx * uv[2], qx * uv[1] - qy * uv[0]};
    for (int i = 0; i < 3; i++) o[i] = v[i] + ((uv[i] * qw) + uuv[i]) * 2.0;
}
M4 perspective(double fovy, double aspect, double zn, double zf) {
    M4 r{};
    double th = std::tan(fovy / 2.0);
    r.m[0][0] = 1.0 / (aspect * th);
    r.m[1][1] = 1.0 / th;
    r.m[2][2] = -(zf + zn) / (zf - zn);
    r.m[2][3] = -1.0;
    r.m[3][2] = -(2.0 * zf * zn) / (zf - zn);
    return r;
}

// PushConstants as launch_rays fills them (path_integrator.cpp:136-159).
void consts_from_camera(const hh_camera& cam, uint32_t w, uint32_t h, hh_launch_consts& o) {
    std::memset(&o, 0, sizeof(o));
    M4 R = quat_m4(cam.orientation), T = ident();
    for (int i = 0; i < 3; i++) T.m[3][i] = cam.position[i];
    M4 view = inverse(mul(T, R));
    M4 proj = perspective(cam.fov_y_deg * (3.14159265358979323846 / 180.0), double(w) / double(h), cam.near_plane,
                          cam.far_plane);
    M4 vpi = inverse(mul(proj, view));
    for (int c = 0; c < 4; c++)
        for (int r = 0; r < 4; r++) o.view_proj_inverse[c * 4 + r] = (float)vpi.m[c][r];
    double ex[3] = {1, 0, 0}, ey[3] = {0, 1, 0}, ez[3] = {0, 0, 1}, right[3], up[3], fwd[3];
    quat_rot(cam.orientation, ex, right);
    quat_rot(cam.orientation, ey, up);
    quat_rot(cam.orientation, ez, fwd);
    for (int i = 0; i < 3; i++) fwd[i] = -fwd[i];
    double fp[4], fpp[3];
    for (int i = 0; i < 3; i++) fpp[i] = cam.position[i] + fwd[i] * cam.focal_length;
    for (int i = 0; i < 3; i++) fp[i] = -fwd[i];
    fp[3] = -(fp[0] * fpp[0] + fp[1] * fpp[1] + fp[2] * fpp[2]);
    for (int i = 0; i < 3; i++) {
        o.camera_pos[i] = cam.position[i];
        o.up_direction[i] = (float)up[i];
        o.right_direction[i] = (float)right[i];
    }
    for (int i = 0; i < 4; i++) o.focal_plane[i] = (float)fp[i];
    o.launch_id_size[2] = w;
    o.launch_id_size[3] = h;
    o.focal_length = cam.focal_length;
    o.aperture_radius = cam.aperture_radius;
}

ViewGPU make_view(const hh_camera& cam, uint32_t w, uint32_t h) {
    hh_launch_consts c;
    consts_from_camera(cam, w, h, c);
    ViewGPU v{};
    std::memcpy(v.vpi, c.view_proj_inverse, 64);
    std::memcpy(v.cam, c.camera_pos, 16);
    std::memcpy(v.up, c.up_direction, 16);
    std::memcpy(v.right, c.right_direction, 16);
    std::memcpy(v.fplane, c.focal_plane, 16);
    v.aperture = cam.aperture_radius;
    v.width = w;
    v.height = h;
    v.tile_x = v.tile_y = 0;
    v.tile_w = w;
    v.tile_h = h;
    v.spp = 1;
    return v;
}

double srgb_to_linear(double c) { return c <= 0.04045 ? c / 12.92 : std::pow((c + 0.055) / 1.055, 2.4); }

bool material_is_emissive(const hh_material_desc& m) {  // material.cpp:71-77
    if (m.emissive_texture != -1) return true;
    return m.emissive_value[0] > 0.0f || m.emissive_value[1] > 0.0f || m.emissive_value[2] > 0.0f;
}
}  // namespace

template <typename T>
static hipError_t upload(hh_scene s, const std::vector<T>& v, T** out) {
    size_t bytes = std::max<size_t>(v.size() * sizeof(T), 16);
    void* p = nullptr;
    HH_TRY(hipMalloc(&p, bytes));
    s->allocs.push_back(p);
    if (!v.empty()) HH_TRY(hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s->ctx->stream));
    *out = (T*)p;
    return hipSuccess;
}

static hipError_t dalloc(hh_scene s, size_t bytes, void** out) {
    HH_TRY(hipMalloc(out, std::max<size_t>(bytes, 16)));
    s->allocs.push_back(*out);
    return hipSuccess;
}

// ---------------------------------------------------------------------------
// Context
// ---------------------------------------------------------------------------
extern "C" {

const char* hh_last_error(void) { return g_err.c_str(); }

// ---------------------------------------------------------------------------
// Scratch of hh_trace_rays_noalloc: one fixed device array per GPU instead of
// allocations. It holds the traversal-stack spill words of kNoAllocMaxRays
// lanes (kSpillStackDepth words each, strided by lane) followed by the raw hit
// records (float4 t,u,v,bits(instance) and two ids per ray). Every entry is
// written before it is read within one call, so nothing needs clearing; calls
// on one GPU share the array, so calls that overlap on one GPU race.
// ---------------------------------------------------------------------------
constexpr uint32_t kNoAllocMaxRays = HH_TRACE_NOALLOC_MAX_RAYS;
constexpr size_t kNoAllocSpillBytes = size_t(kNoAllocMaxRays) * kSpillStackDepth * sizeof(uint32_t);
constexpr size_t kNoAllocHitBytes = size_t(kNoAllocMaxRays) * sizeof(float4);
constexpr size_t kNoAllocIdBytes = size_t(kNoAllocMaxRays) * 2 * sizeof(uint32_t);
constexpr size_t kNoAllocScratchBytes = kNoAllocSpillBytes + kNoAllocHitBytes + kNoAllocIdBytes;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];
static_assert(kNoAllocMaxRays % 256 == 0, "a k_trace_rays grid of up to kNoAllocMaxRays rays fits the spill area");
static_assert(kNoAllocSpillBytes % 256 == 0 && kNoAllocHitBytes % 256 == 0, "scratch sections stay aligned");
static_assert(sizeof(hh_hit) == 24 && sizeof(hh_ray) == 32, "hh_hit / hh_ray layouts");

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on `device`, which must be current. Host
// code never passes the array's name itself: on the host it is a stand-in, not
// a device address.
static hh_status scratch_init(int device) {
    if (device < 0 || device >= kMaxDevices) return fail(HH_ERR_LIMIT_EXCEEDED, "device ordinal %d beyond %d", device, kMaxDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return HH_OK;
    void* ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) return hip_fail(e != hipSuccess ? e : hipErrorInvalidSymbol, "hipGetSymbolAddress(scratch)");
    g_scratch_ptr[device] = ptr;
    return HH_OK;
}

hh_status hh_context_create(int device_ordinal, hh_context* out) {
    HH_ARG(out, "out is null");
    int n = 0;
    HH_HIP(hipGetDeviceCount(&n));
    HH_ARG(device_ordinal >= 0 && device_ordinal < n, "device ordinal out of range");
    HH_HIP(hipSetDevice(device_ordinal));
    {
        hh_status st = scratch_init(device_ordinal);
        if (st != HH_OK) return st;
    }
    hh_context c = new (std::nothrow) hh_context_t();
    if (!c) return fail(HH_ERR_OUT_OF_MEMORY, "context");
    c->device = device_ordinal;
    // a failure releases what was created so far
    auto setup = [&]() -> hh_status {
        HH_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        HH_HIP(hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking));
        HH_HIP(hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
        HH_HIP(hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming));
        HH_HIP(hipEventCreateWithFlags(&c->ev_staged, hipEventDisableTiming));
        HH_HIP(hipHostMalloc((void**)&c->h_stats, 2 * sizeof(unsigned long long)));
        HH_HIP(hipDeviceGetAttribute(&c->num_cu, hipDeviceAttributeMultiprocessorCount, device_ordinal));
        HH_HIP(hipMalloc(&c->d_coef, 30 * sizeof(float)));
        return HH_OK;
    };
    const hh_status st = setup();
    if (st != HH_OK) {
        (void)hh_context_destroy(c);
        return st;
    }
    *out = c;
    return HH_OK;
}

hh_status hh_context_destroy(hh_context ctx) {
    if (!ctx) return HH_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    ctx->drop_graphs();
    ctx->wf.release();
    ctx->bs.release();
    if (ctx->d_scene_table) (void)hipFree(ctx->d_scene_table);
    if (ctx->d_views) (void)hipFree(ctx->d_views);
    if (ctx->dbg) (void)hipFree(ctx->dbg);
    if (ctx->d_coef) (void)hipFree(ctx->d_coef);
    if (ctx->prof.d_counts) (void)hipFree(ctx->prof.d_counts);
    if (ctx->prof.d_snaps) (void)hipFree(ctx->prof.d_snaps);
    for (hipEvent_t e : ctx->prof.pool) (void)hipEventDestroy(e);
    if (ctx->side) (void)hipStreamSynchronize(ctx->side);
    if (ctx->ev_fork) (void)hipEventDestroy(ctx->ev_fork);
    if (ctx->ev_join) (void)hipEventDestroy(ctx->ev_join);
    if (ctx->ev_staged) (void)hipEventDestroy(ctx->ev_staged);
    if (ctx->h_stage) (void)hipHostFree(ctx->h_stage);
    if (ctx->h_stats) (void)hipHostFree(ctx->h_stats);
    if (ctx->side) (void)hipStreamDestroy(ctx->side);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return HH_OK;
}

hh_status hh_sync(hh_context ctx) {
    HH_ARG(ctx, "null context");
    HH_HIP(hipStreamSynchronize(ctx->stream));
    return HH_OK;
}

hh_status hh_context_stream(hh_context ctx, void** out) {
    HH_ARG(ctx && out, "null argument");
    *out = (void*)ctx->stream;
    return HH_OK;
}

hh_status hh_context_set_tuning(hh_context ctx, const hh_tuning* t) {
    HH_ARG(ctx && t, "null argument");
    HH_ARG(t->bvh_width == 4 || t->bvh_width == 8, "bvh_width must be 4 or 8");
    HH_ARG(t->extend_block == 64 || t->extend_block == 128 || t->extend_block == 256, "extend_block must be 64/128/256");
    HH_ARG(t->persistent_blocks_per_cu >= 1 && t->persistent_blocks_per_cu <= 64, "blocks per CU must be 1..64");
    HH_ARG(t->frames_in_flight >= 1 && t->frames_in_flight <= 256, "frames_in_flight must be 1..256");
    HH_ARG(t->refill_lanes >= 1 && t->refill_lanes <= 64, "refill_lanes must be 1..64");
    HH_ARG(t->traversal <= 1, "traversal must be 0 (while-while) or 1 (if-if)");
    HH_ARG(t->bvh_builder <= 3, "bvh_builder must be 0 (LBVH), 1 (binned SAH on the host), 2 (PLOC) or 3 (binned SAH on the GPU)");
    HH_ARG(t->continuation_depth <= 64, "continuation_depth must be 0..64");
    ctx->tune.bvh_builder = t->bvh_builder;
    ctx->tune.continuation_depth = t->continuation_depth;
    ctx->tune.refill_lanes = t->refill_lanes;
    ctx->tune.traversal = (int)t->traversal;
    ctx->tune.frames_in_flight = t->frames_in_flight;
    ctx->tune.bvh_compressed = t->bvh_compressed != 0 && t->bvh_width == 4;
    ctx->tune.bvh_width = (int)t->bvh_width;
    ctx->tune.extend_block = (int)t->extend_block;
    ctx->tune.shade_block = (int)t->extend_block;
    ctx->tune.queue_capacity = t->queue_capacity;
    ctx->tune.blocks_per_cu = (int)t->persistent_blocks_per_cu;
    ctx->drop_graphs();
    return HH_OK;
}

hh_status hh_context_set_profiling(hh_context ctx, uint32_t flags) {
    HH_ARG(ctx, "null context");
    HH_HIP(hipStreamSynchronize(ctx->stream));
    ctx->prof.timing = (flags & 1u) != 0;
    ctx->prof.count = (flags & 2u) != 0;
    if (ctx->prof.timing && !ctx->prof.d_snaps)
        HH_HIP(hipMalloc(&ctx->prof.d_snaps, size_t(Prof::kSnapCap) * 4 * sizeof(unsigned long long)));
    if (ctx->prof.count && !ctx->prof.d_counts) {
        HH_HIP(hipMalloc(&ctx->prof.d_counts, 8 * sizeof(unsigned long long)));
        HH_HIP(hipMemset(ctx->prof.d_counts, 0, 8 * sizeof(unsigned long long)));
    }
    return HH_OK;
}

hh_status hh_context_read_profile(hh_context ctx, double ms[6], uint64_t launches[6], uint64_t counts[6]) {
    HH_ARG(ctx && ms && launches && counts, "null argument");
    HH_HIP(hipStreamSynchronize(ctx->stream));
    Prof& p = ctx->prof;
    std::vector<unsigned long long> snaps(size_t(p.snaps_used) * 4);
    if (p.snaps_used)
        HH_HIP(hipMemcpy(snaps.data(), p.d_snaps, snaps.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost));
    for (const Prof::Rec& r : p.recs) {
        float t = 0.f;
        HH_HIP(hipEventElapsedTime(&t, r.a, r.b));
        p.ms[r.kind] += t;
        p.launches[r.kind]++;
        if (r.snap >= 0) {
            const unsigned long long* q = &snaps[4 * size_t(r.snap)];
            p.rays[r.kind][0] += q[2] - q[0];
            p.rays[r.kind][1] += q[3] - q[1];
        }
    }
    p.recs.clear();
    p.used = 0;
    p.snaps_used = 0;
    for (int k = 0; k < PK_COUNT; k++) {
        ms[k] = p.ms[k];
        launches[k] = p.launches[k];
    }
    unsigned long long c[8] = {0};
    if (p.d_counts) HH_HIP(hipMemcpy(c, p.d_counts, sizeof(c), hipMemcpyDeviceToHost));
    for (int k = 0; k < 6; k++) counts[k] = c[k];
    return HH_OK;
}

hh_status hh_context_profile_rays(hh_context ctx, uint64_t extension[6], uint64_t shadow[6]) {
    HH_ARG(ctx && extension && shadow, "null argument");
    for (int k = 0; k < PK_COUNT; k++) {
        extension[k] = ctx->prof.rays[k][0];
        shadow[k] = ctx->prof.rays[k][1];
    }
    return HH_OK;
}

hh_status hh_context_reset_profile(hh_context ctx) {
    HH_ARG(ctx, "null context");
    HH_HIP(hipStreamSynchronize(ctx->stream));
    Prof& p = ctx->prof;
    p.recs.clear();
    p.used = 0;
    p.snaps_used = 0;
    for (int k = 0; k < PK_COUNT; k++) {
        p.ms[k] = 0;
        p.launches[k] = 0;
        p.rays[k][0] = p.rays[k][1] = 0;
    }
    if (p.d_counts) HH_HIP(hipMemset(p.d_counts, 0, 8 * sizeof(unsigned long long)));
    return HH_OK;
}

hh_status hh_bandwidth_probe(hh_context ctx, uint64_t bytes, uint32_t iters, double* out_gbps) {
    HH_ARG(ctx && out_gbps && bytes >= 16 && iters >= 1, "bad argument");
    HH_ARG(bytes <= (uint64_t)64 << 30, "probe size above 64 GiB");
    HH_HIP(hipSetDevice(ctx->device));
    size_t n = bytes / 16;
    float4 *src = nullptr, *dst = nullptr;
    HH_HIP(hipMalloc(&src, n * 16));
    hipError_t e = hipMalloc(&dst, n * 16);
    if (e != hipSuccess) {
        (void)hipFree(src);
        return hip_fail(e, "probe allocation");
    }
    hipEvent_t a, b;
    (void)hipEventCreate(&a);
    (void)hipEventCreate(&b);
    e = launch_fill_f4(ctx->stream, src, make_float4(1.f, 2.f, 3.f, 4.f), (uint32_t)std::min<size_t>(n, 0xFFFFFFFFu));
    if (e == hipSuccess) e = launch_copy_probe(ctx->stream, ctx->num_cu, src, dst, n);  // warm-up
    if (e == hipSuccess) e = hipEventRecord(a, ctx->stream);
    for (uint32_t i = 0; i < iters && e == hipSuccess; i++) e = launch_copy_probe(ctx->stream, ctx->num_cu, src, dst, n);
    if (e == hipSuccess) e = hipEventRecord(b, ctx->stream);
    if (e == hipSuccess) e = hipEventSynchronize(b);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    (void)hipFree(src);
    (void)hipFree(dst);
    if (e != hipSuccess) return hip_fail(e, "bandwidth probe");
    *out_gbps = 2.0 * (double)n * 16.0 * iters / (ms * 1e-3) / 1e9;  // read + write
    return HH_OK;
}

hh_status hh_context_get_tuning(hh_context ctx, hh_tuning* t) {
    HH_ARG(ctx && t, "null argument");
    t->bvh_width = (uint32_t)ctx->tune.bvh_width;
    t->extend_block = (uint32_t)ctx->tune.extend_block;
    t->shade_block = (uint32_t)ctx->tune.shade_block;
    t->queue_capacity = ctx->tune.queue_capacity;
    t->persistent_blocks_per_cu = (uint32_t)ctx->tune.blocks_per_cu;
    t->frames_in_flight = ctx->tune.frames_in_flight;
    t->bvh_compressed = ctx->tune.bvh_compressed ? 1u : 0u;
    t->refill_lanes = ctx->tune.refill_lanes;
    t->traversal = (uint32_t)ctx->tune.traversal;
    t->bvh_builder = ctx->tune.bvh_builder;
    t->continuation_depth = ctx->tune.continuation_depth;
    return HH_OK;
}

// ---------------------------------------------------------------------------
// Scene
// ---------------------------------------------------------------------------
hh_status hh_scene_create(hh_context ctx, hh_scene* out) {
    HH_ARG(ctx && out, "null argument");
    hh_scene s = new (std::nothrow) hh_scene_t();
    if (!s) return fail(HH_ERR_OUT_OF_MEMORY, "scene");
    s->ctx = ctx;
    *out = s;
    return HH_OK;
}

hh_status hh_scene_destroy(hh_scene s) {
    if (!s) return HH_OK;
    (void)hipSetDevice(s->ctx->device);
    (void)hipStreamSynchronize(s->ctx->stream);
    s->ctx->drop_graphs();
    s->free_device();
    delete s;
    return HH_OK;
}

hh_status hh_scene_add_texture(hh_scene s, uint32_t w, uint32_t h, uint32_t format, const void* texels, uint32_t* out_id) {
    HH_ARG(s && texels && out_id, "null argument");
    HH_ARG(w > 0 && h > 0 && w <= 16384 && h <= 16384, "texture size out of range");
    HH_ARG(format <= HH_TEXTURE_RGBA32F, "unknown texture format");
    HH_ARG(s->textures.size() < HH_MAX_SCENE_MATERIAL_TEXTURE_COUNT, "too many textures");
    TexHost t;
    t.w = w;
    t.h = h;
    t.fmt = format;
    size_t n = size_t(w) * h;
    t.texels.resize(n * 4);
    if (format == HH_TEXTURE_RGBA32F) {
        std::memcpy(t.texels.data(), texels, n * 16);
    } else {
        float lin[256], srgb[256];
        for (int i = 0; i < 256; i++) {
            lin[i] = (float)(i / 255.0);
            srgb[i] = (float)srgb_to_linear(i / 255.0);
        }
        const uint8_t* b = (const uint8_t*)texels;
        for (size_t i = 0; i < n; i++)
            for (int c = 0; c < 4; c++) {
                uint8_t v = b[i * 4 + c];
                t.texels[i * 4 + c] = (format == HH_TEXTURE_RGBA8_SRGB && c < 3) ? srgb[v] : lin[v];
            }
    }
    s->textures.push_back(std::move(t));
    *out_id = (uint32_t)s->textures.size() - 1;
    s->dirty = true;
    return HH_OK;
}

hh_status hh_scene_add_material(hh_scene s, const hh_material_desc* d, uint32_t* out_id) {
    HH_ARG(s && d && out_id, "null argument");
    HH_ARG(s->materials.size() < HH_MAX_SCENE_MATERIAL_COUNT, "too many materials");
    int32_t tex[5] = {d->albedo_texture, d->normal_texture, d->roughness_texture, d->metallic_texture, d->emissive_texture};
    for (int32_t t : tex) HH_ARG(t >= -1 && t < (int32_t)s->textures.size(), "material references an unknown texture");
    s->materials.push_back(*d);
    *out_id = (uint32_t)s->materials.size() - 1;
    s->dirty = true;
    return HH_OK;
}

hh_status hh_scene_add_mesh(hh_scene s, const hh_vertex* v, uint32_t nv, const uint32_t* idx, uint32_t ni,
                            const hh_submesh* sm, uint32_t ns, const uint32_t* mats, uint32_t nm, uint32_t* out_id) {
    HH_ARG(s && out_id, "null argument");
    HH_ARG((nv == 0 || v) && (ni == 0 || idx) && (ns == 0 || sm) && (nm == 0 || mats), "null array");
    for (uint32_t i = 0; i < ni; i++)
        if (idx[i] >= nv) return fail(HH_ERR_INVALID_ARG, "index %u out of range (%u vertices)", idx[i], nv);
    for (uint32_t i = 0; i < ns; i++) {
        if (sm[i].mat_idx >= nm) return fail(HH_ERR_INVALID_ARG, "submesh %u material index out of range", i);
        if ((uint64_t)sm[i].base_index + sm[i].index_count > ni || sm[i].index_count % 3 != 0)
            return fail(HH_ERR_INVALID_ARG, "submesh %u index range invalid", i);
    }
    for (uint32_t i = 0; i < nm; i++) HH_ARG(mats[i] < s->materials.size(), "mesh references an unknown material");
    // positions feed the BVH builders' bounds and bins, which need finite values
    for (uint32_t i = 0; i < nv; i++)
        if (!(std::isfinite(v[i].position[0]) && std::isfinite(v[i].position[1]) && std::isfinite(v[i].position[2])))
            return fail(HH_ERR_INVALID_ARG, "vertex %u has a non-finite position", i);
    MeshHost m;
    m.v.assign(v, v + nv);
    m.idx.assign(idx, idx + ni);
    m.sm.assign(sm, sm + ns);
    m.mats.assign(mats, mats + nm);
    s->meshes.push_back(std::move(m));
    *out_id = (uint32_t)s->meshes.size() - 1;
    s->dirty = true;
    return HH_OK;
}

hh_status hh_scene_add_instance(hh_scene s, const hh_instance_desc* d, uint32_t* out_id) {
    HH_ARG(s && d && out_id, "null argument");
    if (s->instances.size() >= HH_MAX_SCENE_MESH_INSTANCE_COUNT)
        return fail(HH_ERR_LIMIT_EXCEEDED, "more than %d mesh instances", HH_MAX_SCENE_MESH_INSTANCE_COUNT);
    HH_ARG(d->mesh < s->meshes.size(), "instance references an unknown mesh");
    HH_ARG(d->material_override >= -1 && d->material_override < (int32_t)s->materials.size(), "bad material override");
    for (int i = 0; i < 16; i++) HH_ARG(std::isfinite(d->model[i]) && std::isfinite(d->normal[i]), "non-finite transform");
    s->instances.push_back(*d);
    *out_id = (uint32_t)s->instances.size() - 1;
    s->dirty = true;
    return HH_OK;
}

hh_status hh_scene_set_instance_transform(hh_scene s, uint32_t i, const float model[16], const float normal[16]) {
    HH_ARG(s && model && normal, "null argument");
    HH_ARG(i < s->instances.size(), "instance out of range");
    for (int k = 0; k < 16; k++) HH_ARG(std::isfinite(model[k]) && std::isfinite(normal[k]), "non-finite transform");
    std::memcpy(s->instances[i].model, model, 64);
    std::memcpy(s->instances[i].normal, normal, 64);
    s->dirty = true;
    return HH_OK;
}

hh_status hh_scene_add_light(hh_scene s, const hh_light_desc* d, uint32_t* out_id) {
    HH_ARG(s && d && out_id, "null argument");
    HH_ARG(d->type <= HH_LIGHT_POINT, "unknown light type");
    if (s->lights.size() >= HH_MAX_SCENE_LIGHT_COUNT) return fail(HH_ERR_LIMIT_EXCEEDED, "too many lights");
    s->lights.push_back(*d);
    *out_id = (uint32_t)s->lights.size() - 1;
    s->dirty = true;
    return HH_OK;
}

hh_status hh_scene_set_ibl(hh_scene s, uint32_t size, const float* faces) {
    HH_ARG(s, "null scene");
    if (size == 0) {
        s->ibl.clear();
        s->ibl_size = 0;
    } else {
        HH_ARG(faces && size <= 8192, "bad IBL cube");
        s->ibl.assign(faces, faces + size_t(6) * size * size * 4);
        s->ibl_size = size;
    }
    s->dirty = true;
    return HH_OK;
}

hh_status hh_scene_set_sky(hh_scene s, const hh_sky_params* p) {
    HH_ARG(s && p, "null argument");
    HH_ARG(p->cube_size >= 1 && p->cube_size <= 4096, "sky cube size out of range");
    HH_ARG(p->turbidity >= 1.0f && p->turbidity <= 10.0f, "turbidity must be in [1, 10]");
    HH_ARG(p->albedo >= 0.0f && p->albedo <= 1.0f, "ground albedo must be in [0, 1]");
    HH_ARG(p->equirect_width == 0 || (p->equirect_width >= 16 && p->equirect_width <= 16384 && p->equirect_width % 2 == 0),
           "equirect_width must be 0 or even in [16, 16384]");
    s->sky = *p;
    s->dirty = true;
    return HH_OK;
}

hh_status hh_scene_commit(hh_scene s) {
    HH_ARG(s, "null scene");
    hh_context ctx = s->ctx;
    HH_HIP(hipSetDevice(ctx->device));
    HH_HIP(hipStreamSynchronize(ctx->stream));
    ctx->drop_graphs();
    s->free_device();
    const int W = ctx->tune.bvh_width;
    // ---- packing (scene.cpp:915-1311) -------------------------------------
    std::vector<MaterialData> mat_data;
    std::vector<int32_t> tex_slots;
    std::vector<LightData> area;
    std::vector<InstanceData> inst_data;
    std::vector<uint32_t> submesh_off;
    std::vector<uint2> submesh_info;
    std::vector<int> mesh_global(s->meshes.size(), -1);
    std::vector<uint32_t> global_meshes;
    std::vector<int> mat_global(s->materials.size(), -1);
    std::vector<char> tex_seen(s->textures.size(), 0);
    auto pack_tex = [&](int32_t tex_id, int32_t& slot) {
        if (!tex_seen[tex_id]) {
            tex_seen[tex_id] = 1;
            slot = (int32_t)tex_slots.size();
            tex_slots.push_back(tex_id);
        }
    };
    for (size_t ni = 0; ni < s->instances.size(); ni++) {
        const hh_instance_desc& inst = s->instances[ni];
        const MeshHost& me = s->meshes[inst.mesh];
        auto material_of = [&](const hh_submesh& sm) -> uint32_t {
            return inst.material_override >= 0 ? (uint32_t)inst.material_override : me.mats[sm.mat_idx];
        };
        if (mesh_global[inst.mesh] < 0) {
            mesh_global[inst.mesh] = (int)global_meshes.size();
            global_meshes.push_back(inst.mesh);
            for (const hh_submesh& sm : me.sm) {
                uint32_t mid = material_of(sm);
                const hh_material_desc& m = s->materials[mid];
                if (mat_global[mid] < 0) {
                    MaterialData md;
                    for (int k = 0; k < 4; k++) {
                        md.tex0[k] = md.tex1[k] = -1;
                        md.albedo[k] = md.emissive[k] = md.roughness_metallic[k] = 0.f;
                    }
                    if (m.albedo_texture >= 0) pack_tex(m.albedo_texture, md.tex0[0]);
                    else {
                        for (int k = 0; k < 3; k++) md.albedo[k] = std::pow(m.albedo_value[k], 2.2f);  // scene.cpp:1017
                        md.albedo[3] = m.albedo_value[3];
                    }
                    if (m.normal_texture >= 0) pack_tex(m.normal_texture, md.tex0[1]);
                    if (m.roughness_texture >= 0) {
                        int32_t slot = -1;
                        pack_tex(m.roughness_texture, slot);
                        if (slot >= 0) { md.tex0[2] = slot; md.tex1[2] = m.roughness_channel; }
                    } else md.roughness_metallic[0] = m.roughness_value;
                    if (m.metallic_texture >= 0) {
                        int32_t slot = -1;
                        pack_tex(m.metallic_texture, slot);
                        if (slot >= 0) { md.tex0[3] = slot; md.tex1[3] = m.metallic_channel; }
                    } else md.roughness_metallic[1] = m.metallic_value;
                    if (m.emissive_texture >= 0) pack_tex(m.emissive_texture, md.tex1[0]);
                    else for (int k = 0; k < 4; k++) md.emissive[k] = m.emissive_value[k];
                    mat_data.push_back(md);
                    mat_global[mid] = (int)mat_data.size() - 1;
                }
                if (material_is_emissive(m)) {  // scene.cpp:1111-1119
                    LightData l{};
                    l.d0[0] = 4.0f;
                    l.d0[1] = float(ni);
                    l.d0[2] = float(mat_global[mid]);
                    l.d0[3] = float(sm.base_index / 3);
                    l.d1[0] = float(sm.index_count / 3);
                    area.push_back(l);
                }
            }
        }
        submesh_off.push_back((uint32_t)submesh_info.size());
        // scene.cpp:934,1143: an unordered_map lookup, so a material that was
        // never packed (the mesh was first packed under another instance's
        // override) reads as material 0
        for (const hh_submesh& sm : me.sm) {
            const int gm = mat_global[material_of(sm)];
            submesh_info.push_back(make_uint2(sm.base_index / 3, gm < 0 ? 0u : (uint32_t)gm));
        }
        InstanceData id{};
        std::memcpy(id.model, inst.model, 64);
        std::memcpy(id.normal, inst.normal, 64);
        id.mesh_index = (uint32_t)mesh_global[inst.mesh];
        inst_data.push_back(id);
    }
    if (mat_data.size() > HH_MAX_SCENE_MATERIAL_COUNT) return fail(HH_ERR_LIMIT_EXCEEDED, "too many materials");
    // lights: area, env, directional, point, spot (scene.cpp:1259-1309)
    std::vector<LightData> lights = area;
    uint32_t n_dir = 0, n_point = 0, n_spot = 0;
    const hh_light_desc* sun = nullptr;
    for (auto& l : s->lights) {
        if (l.type == HH_LIGHT_DIRECTIONAL) { n_dir++; if (!sun) sun = &l; }
        n_point += l.type == HH_LIGHT_POINT;
        n_spot += l.type == HH_LIGHT_SPOT;
    }
    const bool has_ibl = s->ibl_size > 0;
    if (has_ibl || n_dir > 0) {
        LightData l{};
        l.d0[0] = 3.0f;
        lights.push_back(l);
    }
    for (auto& l : s->lights)
        if (l.type == HH_LIGHT_DIRECTIONAL) {
            LightData d{};
            d.d0[0] = 0.0f;
            for (int k = 0; k < 3; k++) { d.d0[1 + k] = l.color[k]; d.d1[k] = l.forward[k]; }
            d.d1[3] = l.intensity;
            d.d2[3] = l.radius;
            lights.push_back(d);
        }
    for (auto& l : s->lights)
        if (l.type == HH_LIGHT_POINT) {
            LightData d{};
            d.d0[0] = 2.0f;
            for (int k = 0; k < 3; k++) { d.d0[1 + k] = l.color[k]; d.d2[k] = l.position[k]; }
            d.d1[3] = l.intensity;
            d.d2[3] = l.radius;
            lights.push_back(d);
        }
    for (auto& l : s->lights)
        if (l.type == HH_LIGHT_SPOT) {
            LightData d{};
            d.d0[0] = 1.0f;
            for (int k = 0; k < 3; k++) { d.d0[1 + k] = l.color[k]; d.d1[k] = l.forward[k]; d.d2[k] = l.position[k]; }
            d.d1[3] = l.intensity;
            d.d2[3] = l.radius;
            d.d3[0] = std::cos(l.inner_cone_deg * (3.14159265358979323846f / 180.0f));
            d.d3[1] = std::cos(l.outer_cone_deg * (3.14159265358979323846f / 180.0f));
            lights.push_back(d);
        }
    if (lights.size() > HH_MAX_SCENE_LIGHT_COUNT) return fail(HH_ERR_LIMIT_EXCEEDED, "too many lights");
    s->light_data = lights;
    const uint32_t num_lights = (uint32_t)area.size() + n_dir + n_spot + n_point + ((has_ibl || n_dir > 0) ? 1u : 0u);

    // ---- geometry pools (global mesh order) ---------------------------------
    std::vector<uint32_t> vbase, ibase;
    size_t nv = 0, ni = 0, ntri = 0;
    std::vector<uint32_t> mesh_tris(global_meshes.size());
    for (size_t g = 0; g < global_meshes.size(); g++) {
        const MeshHost& me = s->meshes[global_meshes[g]];
        vbase.push_back((uint32_t)nv);
        ibase.push_back((uint32_t)ni);
        nv += me.v.size();
        ni += me.idx.size();
        uint32_t t = 0;
        for (auto& sm : me.sm) t += sm.index_count / 3;
        mesh_tris[g] = std::max<uint32_t>(t, 1);  // an empty mesh gets one inert triangle
        ntri += mesh_tris[g];
    }
    if (ntri >= kLeafStartMask) return fail(HH_ERR_LIMIT_EXCEEDED, "too many triangles (%zu)", ntri);
    float4* d_vtx = nullptr;
    uint32_t* d_idx = nullptr;
    HH_HIP(dalloc(s, nv * sizeof(hh_vertex), (void**)&d_vtx));
    HH_HIP(dalloc(s, ni * 4, (void**)&d_idx));
    for (size_t g = 0; g < global_meshes.size(); g++) {
        const MeshHost& me = s->meshes[global_meshes[g]];
        if (!me.v.empty())
            HH_HIP(hipMemcpyAsync((char*)d_vtx + size_t(vbase[g]) * sizeof(hh_vertex), me.v.data(), me.v.size() * sizeof(hh_vertex),
                                  hipMemcpyHostToDevice, ctx->stream));
        if (!me.idx.empty())
            HH_HIP(hipMemcpyAsync(d_idx + ibase[g], me.idx.data(), me.idx.size() * 4, hipMemcpyHostToDevice, ctx->stream));
    }
    // ---- BLAS per mesh (mesh.cpp:46-117) ------------------------------------
    const size_t node_bytes = W == 8 ? sizeof(WideNode<8>) : sizeof(WideNode<4>);
    size_t node_cap = s->instances.size() + 2;
    for (uint32_t t : mesh_tris) node_cap += t + 1;
    void* d_nodes = nullptr;
    float4* d_tris = nullptr;
    float4* d_shade = nullptr;
    HH_HIP(dalloc(s, node_cap * node_bytes, &d_nodes));
    HH_HIP(dalloc(s, ntri * 3 * sizeof(float4), (void**)&d_tris));
    HH_HIP(dalloc(s, ntri * 4 * sizeof(float4), (void**)&d_shade));
    std::vector<uint32_t> blas_root(global_meshes.size());
    std::vector<float> blas_box(global_meshes.size() * 6);
    uint32_t node_base = 0, tri_base = 0;
    // Per-triangle positions (3 float4: xyz + primitive index / geometry index
    // and opaque bit in w) and shading records (4 float4, leaf order like the
    // positions): (u0 v0 u1 v1) (u2 v2 n0.x n0.y) (n0.z n1.xyz) (n2.xyz 0), the
    // vertex attributes the shading and the alpha test interpolate, in one
    // line. Mesh g owns triangles tri_off[g]..; an empty mesh one inert
    // triangle. Filled by host threads, one mesh at a time per thread.
    const size_t G = global_meshes.size();
    std::vector<uint32_t> tri_off(G + 1, 0);
    for (size_t g = 0; g < G; g++) tri_off[g + 1] = tri_off[g] + mesh_tris[g];
    std::vector<float4> soup(ntri * 3), shade_soup(ntri * 4);
    auto fill_mesh = [&](size_t g) {
        const MeshHost& me = s->meshes[global_meshes[g]];
        float4* ps = &soup[3 * size_t(tri_off[g])];
        float4* ss = &shade_soup[4 * size_t(tri_off[g])];
        size_t t = 0;
        for (uint32_t gi = 0; gi < me.sm.size(); gi++) {
            const hh_submesh& sm = me.sm[gi];
            const hh_material_desc& m = s->materials[me.mats[sm.mat_idx]];
            uint32_t opaque = (m.type == HH_MATERIAL_OPAQUE && !m.alpha_test) ? 0x80000000u : 0u;  // mesh.cpp:75-76
            for (uint32_t p = 0; p < sm.index_count / 3; p++, t++) {
                const hh_vertex* v3[3];
                for (int k = 0; k < 3; k++) v3[k] = &me.v[me.idx[sm.base_index + 3 * p + k]];
                uint32_t w0 = p, w1 = gi | opaque;
                float fw0, fw1;
                std::memcpy(&fw0, &w0, 4);
                std::memcpy(&fw1, &w1, 4);
                ps[3 * t] = make_float4(v3[0]->position[0], v3[0]->position[1], v3[0]->position[2], fw0);
                ps[3 * t + 1] = make_float4(v3[1]->position[0], v3[1]->position[1], v3[1]->position[2], fw1);
                ps[3 * t + 2] = make_float4(v3[2]->position[0], v3[2]->position[1], v3[2]->position[2], 0.f);
                ss[4 * t] = make_float4(v3[0]->tex_coord[0], v3[0]->tex_coord[1], v3[1]->tex_coord[0], v3[1]->tex_coord[1]);
                ss[4 * t + 1] = make_float4(v3[2]->tex_coord[0], v3[2]->tex_coord[1], v3[0]->normal[0], v3[0]->normal[1]);
                ss[4 * t + 2] = make_float4(v3[0]->normal[2], v3[1]->normal[0], v3[1]->normal[1], v3[1]->normal[2]);
                ss[4 * t + 3] = make_float4(v3[2]->normal[0], v3[2]->normal[1], v3[2]->normal[2], 0.f);
            }
        }
        if (t == 0) {
            for (int k = 0; k < 3; k++) ps[k] = make_float4(0.f, 0.f, 0.f, 0.f);
            for (int k = 0; k < 4; k++) ss[k] = make_float4(0.f, 0.f, 0.f, 0.f);
        }
    };
    {
        const size_t nth = std::min<size_t>(G, std::max(1u, std::min(16u, std::thread::hardware_concurrency())));
        std::vector<std::thread> pool;
        for (size_t w = 1; w < nth; w++)
            pool.emplace_back([&, w] {
                for (size_t g = w; g < G; g += nth) fill_mesh(g);
            });
        for (size_t g = 0; g < G; g += nth) fill_mesh(g);
        for (auto& th : pool) th.join();
    }
    float4* d_soup = nullptr;
    float4* d_shade_soup = nullptr;
    HH_HIP(hipMalloc(&d_soup, soup.size() * sizeof(float4)));
    if (hipMalloc(&d_shade_soup, shade_soup.size() * sizeof(float4)) != hipSuccess) {
        (void)hipFree(d_soup);
        return fail(HH_ERR_OUT_OF_MEMORY, "shading records");
    }
    hipError_t e = hipSuccess;
    if (ntri > 0) {
        e = hipMemcpyAsync(d_soup, soup.data(), soup.size() * sizeof(float4), hipMemcpyHostToDevice, ctx->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(d_shade_soup, shade_soup.data(), shade_soup.size() * sizeof(float4), hipMemcpyHostToDevice,
                               ctx->stream);
    }
    hh_status st = HH_OK;
    if (e != hipSuccess) {
        st = hip_fail(e, "triangle upload");
    } else if (G == 0) {
        // no meshes: no BLAS
    } else if (ctx->tune.bvh_builder == 3) {
        // every mesh's tree on the GPU together
        std::vector<int2> ranges(G);
        std::vector<uint32_t> counts(G);
        for (size_t g = 0; g < G; g++) ranges[g] = make_int2((int)tri_off[g], (int)mesh_tris[g]);
        e = build_blas_forest(ctx->stream, W, ctx->tune.leaf_size, ranges, (uint32_t)ntri, d_soup, d_tris, d_shade_soup,
                              d_shade, d_nodes, 0, ctx->bs, blas_root.data(), counts.data(), blas_box.data());
        if (e != hipSuccess) st = hip_fail(e, "BLAS build");
        for (uint32_t c : counts) node_base += c;
        tri_base = (uint32_t)ntri;
    } else {
        for (size_t g = 0; g < G; g++) {
            const uint32_t n = mesh_tris[g], off = tri_off[g];
            uint32_t nn = 0;
            e = build_blas(ctx->stream, W, ctx->tune.leaf_size, n, d_soup + 3 * size_t(off), d_tris,
                           d_shade_soup + 4 * size_t(off), d_shade,
                           ctx->tune.bvh_builder == 1 ? soup.data() + 3 * size_t(off) : nullptr, tri_base, d_nodes, node_base,
                           ctx->bs, &nn, &blas_box[g * 6], ctx->tune.bvh_builder == 2, false);
            if (e != hipSuccess) {
                st = hip_fail(e, "BLAS build");
                break;
            }
            blas_root[g] = node_base;
            node_base += nn;
            tri_base += n;
        }
    }
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(d_soup);
    (void)hipFree(d_shade_soup);
    if (st != HH_OK) return st;
    s->bvh_stats[0] = node_base;
    s->bvh_stats[1] = tri_base;
    // ---- TLAS over instances (renderer.cpp:113-181) ---------------------------
    std::vector<InstTrav> trav(s->instances.size());
    std::vector<float> ilo(s->instances.size() * 3), ihi(s->instances.size() * 3);
    for (size_t i = 0; i < s->instances.size(); i++) {
        const hh_instance_desc& inst = s->instances[i];
        uint32_t g = (uint32_t)mesh_global[inst.mesh];
        M4 inv = inverse(from_glm(inst.model));
        for (int r = 0; r < 3; r++)
            for (int c = 0; c < 4; c++) trav[i].w2o[r * 4 + c] = (float)inv.m[c][r];
        trav[i].blas_root = blas_root[g];
        trav[i].mesh = g;
        // one alpha source for all non-opaque geometry of the instance, if any
        // (the geometry flag follows the mesh's material, the alpha the
        // instance's, as the any-hit test does)
        {
            const MeshHost& me = s->meshes[inst.mesh];
            uint32_t src = kAlphaConst, val = 0;
            bool first = true;
            for (uint32_t gi = 0; gi < me.sm.size(); gi++) {
                const hh_material_desc& own = s->materials[me.mats[me.sm[gi].mat_idx]];
                if (own.type == HH_MATERIAL_OPAQUE && !own.alpha_test) continue;
                const MaterialData& md = mat_data[submesh_info[submesh_off[i] + gi].y];
                uint32_t gsrc = kAlphaConst, gval = 0;
                if (md.tex0[0] != -1) gsrc = (uint32_t)md.tex0[0];
                else std::memcpy(&gval, &md.albedo[3], 4);
                if (first) { src = gsrc; val = gval; first = false; }
                else if (gsrc != src || gval != val) { src = kAlphaMixed; val = 0; break; }
            }
            trav[i].alpha_src = src;
            trav[i].alpha_value = val;
        }
        const float* bb = &blas_box[g * 6];
        double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
        for (int c = 0; c < 8; c++) {
            double p[3] = {(c & 1) ? bb[3] : bb[0], (c & 2) ? bb[4] : bb[1], (c & 4) ? bb[5] : bb[2]};
            for (int r = 0; r < 3; r++) {
                double v = (double)inst.model[0 * 4 + r] * p[0] + (double)inst.model[1 * 4 + r] * p[1] +
                           (double)inst.model[2 * 4 + r] * p[2] + (double)inst.model[3 * 4 + r];
                lo[r] = std::min(lo[r], v);
                hi[r] = std::max(hi[r], v);
            }
        }
        for (int r = 0; r < 3; r++) {
            ilo[i * 3 + r] = std::nextafter((float)lo[r], -INFINITY);
            ihi[i * 3 + r] = std::nextafter((float)hi[r], INFINITY);
        }
    }
    uint32_t tlas_root = node_base, tlas_nodes = 0;
    if (s->instances.empty()) {
        // a root without children: all-infinite child boxes, which every ray misses
        std::vector<char> empty(node_bytes, 0);
        auto fill_empty = [&](auto* nd, int w) {
            for (int k = 0; k < w; k++)
                nd->lox[k] = nd->loy[k] = nd->loz[k] = nd->hix[k] = nd->hiy[k] = nd->hiz[k] = INFINITY;
        };
        if (W == 8) fill_empty((WideNode<8>*)empty.data(), 8);
        else fill_empty((WideNode<4>*)empty.data(), 4);
        HH_HIP(hipMemcpyAsync((char*)d_nodes + size_t(node_base) * node_bytes, empty.data(), node_bytes, hipMemcpyHostToDevice,
                              ctx->stream));
        HH_HIP(hipStreamSynchronize(ctx->stream));   // `empty` goes out of scope
        tlas_nodes = 1;
    } else {
        hipError_t e = build_tlas(ctx->stream, W, (uint32_t)s->instances.size(), ilo.data(), ihi.data(), d_nodes, node_base, ctx->bs,
                                  &tlas_nodes, ctx->tune.bvh_builder == 1, ctx->tune.bvh_builder == 2,
                                  ctx->tune.bvh_builder == 3);
        if (e != hipSuccess) return hip_fail(e, "TLAS build");
    }
    s->bvh_stats[2] = tlas_nodes;
    s->bvh_stats[3] = global_meshes.size();
    // Optional 64-B quantized BVH4 nodes (same topology and indices).
    int layout = W;
    if (ctx->tune.bvh_compressed && W == 4) {
        void* q = nullptr;
        const uint32_t total_nodes = node_base + tlas_nodes;
        HH_HIP(dalloc(s, size_t(total_nodes) * sizeof(QNode4), &q));
        HH_HIP(quantize_bvh4(ctx->stream, d_nodes, q, total_nodes));
        HH_HIP(hipStreamSynchronize(ctx->stream));
        d_nodes = q;
        layout = kLayoutQ4;
    }
    // ---- textures --------------------------------------------------------------
    std::vector<TexDesc> tdesc;
    std::vector<float4> texels;
    for (int32_t tid : tex_slots) {
        const TexHost& t = s->textures[tid];
        TexDesc d{(uint32_t)texels.size(), t.w, t.h, 0};
        tdesc.push_back(d);
        for (size_t i = 0; i < size_t(t.w) * t.h; i++)
            texels.push_back(make_float4(t.texels[i * 4], t.texels[i * 4 + 1], t.texels[i * 4 + 2], t.texels[i * 4 + 3]));
    }
    // ---- power-proportional emitters (HH_FLAG_LIGHT_POWER): every emissive
    // triangle of every instance in world space; selection probability
    // proportional to luminance(emission) x area, accumulated in double
    std::vector<float4> emit_tri, emit_le;
    std::vector<float> emit_cdf;
    {
        std::vector<double> power;
        for (size_t n2 = 0; n2 < s->instances.size(); n2++) {
            const hh_instance_desc& inst = s->instances[n2];
            const MeshHost& me = s->meshes[inst.mesh];
            for (const hh_submesh& sm : me.sm) {
                const uint32_t mid = inst.material_override >= 0 ? (uint32_t)inst.material_override : me.mats[sm.mat_idx];
                const hh_material_desc& m = s->materials[mid];
                const double le[3] = {m.emissive_value[0], m.emissive_value[1], m.emissive_value[2]};
                const double lum = 0.2126 * le[0] + 0.7152 * le[1] + 0.0722 * le[2];
                if (!(lum > 0.0)) continue;
                for (uint32_t t = 0; t < sm.index_count / 3; t++) {
                    double w[3][3], ns[3] = {0.0, 0.0, 0.0};
                    for (int k = 0; k < 3; k++) {
                        const hh_vertex& vx = me.v[me.idx[sm.base_index + 3 * t + k]];
                        const float* p = vx.position;
                        for (int r = 0; r < 3; r++) {
                            w[k][r] = (double)inst.model[r] * p[0] + (double)inst.model[4 + r] * p[1] +
                                      (double)inst.model[8 + r] * p[2] + (double)inst.model[12 + r];
                            ns[r] += (double)inst.normal[r] * vx.normal[0] + (double)inst.normal[4 + r] * vx.normal[1] +
                                     (double)inst.normal[8 + r] * vx.normal[2];
                        }
                    }
                    double e1[3] = {w[1][0] - w[0][0], w[1][1] - w[0][1], w[1][2] - w[0][2]};
                    double e2[3] = {w[2][0] - w[0][0], w[2][1] - w[0][1], w[2][2] - w[0][2]};
                    double cx = e1[1] * e2[2] - e1[2] * e2[1], cy = e1[2] * e2[0] - e1[0] * e2[2],
                           cz = e1[0] * e2[1] - e1[1] * e2[0];
                    // the emitting side is the vertex normals' side: wind so that
                    // (v1 - v0) x (v2 - v0) points there
                    if (cx * ns[0] + cy * ns[1] + cz * ns[2] < 0.0) {
                        for (int r = 0; r < 3; r++) std::swap(w[1][r], w[2][r]);
                        cx = -cx;
                        cy = -cy;
                        cz = -cz;
                    }
                    const double area = 0.5 * std::sqrt(cx * cx + cy * cy + cz * cz);
                    if (!(area > 0.0)) continue;
                    emit_tri.push_back(make_float4((float)w[0][0], (float)w[0][1], (float)w[0][2], (float)area));
                    emit_tri.push_back(make_float4((float)w[1][0], (float)w[1][1], (float)w[1][2], 0.0f));
                    emit_tri.push_back(make_float4((float)w[2][0], (float)w[2][1], (float)w[2][2], 0.0f));
                    emit_le.push_back(make_float4((float)le[0], (float)le[1], (float)le[2], 0.0f));
                    power.push_back(lum * area);
                }
            }
        }
        double total = 0.0;
        for (double p : power) total += p;
        double acc = 0.0;
        for (size_t i = 0; i < power.size(); i++) {
            acc += power[i];
            emit_le[i].w = (float)(power[i] / total);
            emit_cdf.push_back(i + 1 == power.size() ? 1.0f : (float)(acc / total));
        }
    }
    // ---- environment: IBL, sky (directional light), or the black 2x2 default -----
    uint32_t env_size;
    float4* d_env = nullptr;
    uint16_t* d_env_eq = nullptr;
    const uint32_t eq_w = (!has_ibl && n_dir > 0) ? s->sky.equirect_width : 0u;
    if (eq_w > 0) {
        // extension: the sky as an equirectangular RGBA16F map (no cube)
        float sd[3] = {-sun->forward[0], -sun->forward[1], -sun->forward[2]};
        sky_coefficients(s->sky.turbidity, s->sky.albedo, s->sky.normalized_sun_y, sd, s->sky_coef);
        HH_HIP(hipMemcpyAsync(ctx->d_coef, s->sky_coef, sizeof(s->sky_coef), hipMemcpyHostToDevice, ctx->stream));
        HH_HIP(dalloc(s, size_t(eq_w) * (eq_w / 2) * 4 * sizeof(uint16_t), (void**)&d_env_eq));
        HH_HIP(launch_sky_bake_equirect(ctx->stream, d_env_eq, eq_w, ctx->d_coef, sd));
        env_size = 2;
        HH_HIP(dalloc(s, 6 * 4 * sizeof(float4), (void**)&d_env));
        HH_HIP(launch_fill_f4(ctx->stream, d_env, make_float4(0.f, 0.f, 0.f, 0.f), 24));
    } else if (has_ibl) {
        env_size = s->ibl_size;
        HH_HIP(dalloc(s, size_t(6) * env_size * env_size * sizeof(float4), (void**)&d_env));
        HH_HIP(hipMemcpyAsync(d_env, s->ibl.data(), size_t(6) * env_size * env_size * sizeof(float4), hipMemcpyHostToDevice,
                              ctx->stream));
    } else if (n_dir > 0) {
        env_size = s->sky.cube_size;
        float sd[3] = {-sun->forward[0], -sun->forward[1], -sun->forward[2]};
        sky_coefficients(s->sky.turbidity, s->sky.albedo, s->sky.normalized_sun_y, sd, s->sky_coef);
        HH_HIP(hipMemcpyAsync(ctx->d_coef, s->sky_coef, sizeof(s->sky_coef), hipMemcpyHostToDevice, ctx->stream));
        HH_HIP(dalloc(s, size_t(6) * env_size * env_size * sizeof(float4), (void**)&d_env));
        HH_HIP(launch_sky_bake(ctx->stream, d_env, env_size, ctx->d_coef, sd));
    } else {
        env_size = 2;
        HH_HIP(dalloc(s, 6 * 4 * sizeof(float4), (void**)&d_env));
        HH_HIP(launch_fill_f4(ctx->stream, d_env, make_float4(0.f, 0.f, 0.f, 0.f), 24));
    }
    // ---- upload tables --------------------------------------------------------
    SceneGPU g{};
    g.nodes = d_nodes;
    g.tris = d_tris;
    g.shade = d_shade;
    g.vtx = d_vtx;
    g.idx = d_idx;
    g.env = d_env;
    g.env_size = env_size;
    g.num_lights = num_lights;
    g.tlas_root = tlas_root;
    g.env_eq = d_env_eq;
    g.env_eq_w = eq_w;
    g.num_area = (uint32_t)area.size();
    g.num_emit = (uint32_t)emit_cdf.size();
    if (!emit_cdf.empty()) {
        HH_HIP(upload(s, emit_tri, (float4**)&g.emit_tri));
        HH_HIP(upload(s, emit_le, (float4**)&g.emit_le));
        HH_HIP(upload(s, emit_cdf, (float**)&g.emit_cdf));
    }
    g.bvh_width = (uint32_t)layout;
    {
        const std::vector<unsigned long long> zeros(kNumFaults, 0ull);
        HH_HIP(upload(s, zeros, &g.faults));
    }
    HH_HIP(upload(s, trav, (InstTrav**)&g.inst_trav));
    HH_HIP(upload(s, inst_data, (InstanceData**)&g.inst));
    HH_HIP(upload(s, mat_data, (MaterialData**)&g.mats));
    HH_HIP(upload(s, lights, (LightData**)&g.lights));
    HH_HIP(upload(s, submesh_off, (uint32_t**)&g.submesh_off));
    HH_HIP(upload(s, submesh_info, (uint2**)&g.submesh_info));
    HH_HIP(upload(s, vbase, (uint32_t**)&g.vbase));
    HH_HIP(upload(s, ibase, (uint32_t**)&g.ibase));
    HH_HIP(upload(s, tdesc, (TexDesc**)&g.tex));
    HH_HIP(upload(s, texels, (float4**)&g.texels));
    std::vector<SceneGPU> one{g};
    HH_HIP(upload(s, one, &s->d_gpu));
    HH_HIP(hipStreamSynchronize(ctx->stream));
    s->gpu = g;
    s->committed = true;
    s->dirty = false;
    s->generation = g_generation++;
    return HH_OK;
}

hh_status hh_scene_traversal_faults(hh_scene s, uint64_t out[3], int reset) {
    HH_ARG(s && out, "null argument");
    if (!s->committed || !s->gpu.faults) {
        out[0] = out[1] = out[2] = 0;
        return HH_OK;
    }
    HH_HIP(hipSetDevice(s->ctx->device));
    unsigned long long v[kNumFaults];
    HH_HIP(hipMemcpyAsync(v, s->gpu.faults, sizeof(v), hipMemcpyDeviceToHost, s->ctx->stream));
    HH_HIP(hipStreamSynchronize(s->ctx->stream));
    for (int k = 0; k < kNumFaults; k++) out[k] = v[k];
    if (reset) {
        const unsigned long long z[kNumFaults] = {0, 0, 0};
        HH_HIP(hipMemcpyAsync(s->gpu.faults, z, sizeof(z), hipMemcpyHostToDevice, s->ctx->stream));
        HH_HIP(hipStreamSynchronize(s->ctx->stream));
    }
    return HH_OK;
}

static hh_status require_committed(hh_scene s) {
    if (!s) return fail(HH_ERR_INVALID_ARG, "null scene");
    if (!s->committed || s->dirty) return fail(HH_ERR_NOT_READY, "scene has uncommitted changes; call hh_scene_commit");
    return HH_OK;
}

hh_status hh_scene_num_lights(hh_scene s, uint32_t* out) {
    HH_ARG(out, "null argument");
    hh_status st = require_committed(s);
    if (st != HH_OK) return st;
    *out = s->gpu.num_lights;
    return HH_OK;
}

hh_status hh_scene_read_lights(hh_scene s, float* out, uint32_t cap) {
    HH_ARG(out, "null argument");
    hh_status st = require_committed(s);
    if (st != HH_OK) return st;
    uint32_t n = std::min<uint32_t>(cap, (uint32_t)s->light_data.size());
    std::memcpy(out, s->light_data.data(), n * sizeof(LightData));
    return HH_OK;
}

hh_status hh_scene_read_env(hh_scene s, float* out, uint32_t* out_size) {
    HH_ARG(out_size, "null argument");
    hh_status st = require_committed(s);
    if (st != HH_OK) return st;
    *out_size = s->gpu.env_size;
    if (out) {
        HH_HIP(hipMemcpy(out, s->gpu.env, size_t(6) * s->gpu.env_size * s->gpu.env_size * sizeof(float4), hipMemcpyDeviceToHost));
    }
    return HH_OK;
}

hh_status hh_scene_read_env_equirect(hh_scene s, uint16_t* out, uint32_t* out_width) {
    HH_ARG(s && out_width, "null argument");
    hh_status st = require_committed(s);
    if (st != HH_OK) return st;
    *out_width = s->gpu.env_eq_w;
    if (out && s->gpu.env_eq_w > 0) {
        HH_HIP(hipStreamSynchronize(s->ctx->stream));
        HH_HIP(hipMemcpy(out, s->gpu.env_eq, size_t(s->gpu.env_eq_w) * (s->gpu.env_eq_w / 2) * 4 * sizeof(uint16_t),
                         hipMemcpyDeviceToHost));
    }
    return HH_OK;
}

hh_status hh_scene_sky_coefficients(hh_scene s, float out[30]) {
    HH_ARG(out, "null argument");
    hh_status st = require_committed(s);
    if (st != HH_OK) return st;
    std::memcpy(out, s->sky_coef, sizeof(s->sky_coef));
    return HH_OK;
}

hh_status hh_sky_fit(const hh_sky_params* p, const float sun_dir[3], float out[30]) {
    HH_ARG(p && sun_dir && out, "null argument");
    sky_coefficients(p->turbidity, p->albedo, p->normalized_sun_y, sun_dir, out);
    return HH_OK;
}

hh_status hh_scene_bvh_stats(hh_scene s, uint64_t out[4]) {
    HH_ARG(out, "null argument");
    hh_status st = require_committed(s);
    if (st != HH_OK) return st;
    std::memcpy(out, s->bvh_stats, sizeof(s->bvh_stats));
    return HH_OK;
}

// ---------------------------------------------------------------------------
// Surfaces and tonemap
// ---------------------------------------------------------------------------
hh_status hh_surface_create(hh_context ctx, uint32_t w, uint32_t h, hh_surface* out) {
    HH_ARG(ctx && out, "null argument");
    HH_ARG(w > 0 && h > 0 && w <= 16384 && h <= 16384, "surface size out of range");
    HH_HIP(hipSetDevice(ctx->device));
    hh_surface s = new (std::nothrow) hh_surface_t();
    if (!s) return fail(HH_ERR_OUT_OF_MEMORY, "surface");
    s->ctx = ctx;
    s->w = w;
    s->h = h;
    hipError_t e = hipMalloc(&s->d, size_t(w) * h * sizeof(float4));
    if (e != hipSuccess) {
        delete s;
        return hip_fail(e, "surface allocation");
    }
    *out = s;
    return hh_surface_clear(s);
}

hh_status hh_surface_destroy(hh_surface s) {
    if (!s) return HH_OK;
    (void)hipStreamSynchronize(s->ctx->stream);
    s->ctx->drop_graphs();
    (void)hipFree(s->d);
    if (s->aux) (void)hipFree(s->aux);
    delete s;
    return HH_OK;
}

hh_status hh_surface_clear(hh_surface s) {  // vkCmdClearColorImage (renderer.cpp:212-223)
    HH_ARG(s, "null surface");
    HH_HIP(launch_fill_f4(s->ctx->stream, s->d, make_float4(0.f, 0.f, 0.f, 1.f), s->w * s->h));
    if (s->aux) HH_HIP(launch_fill_f4(s->ctx->stream, s->aux, make_float4(0.f, 0.f, 0.f, 0.f), 2 * s->w * s->h));
    return HH_OK;
}

hh_status hh_surface_enable_features(hh_surface s) {
    HH_ARG(s, "null surface");
    if (s->aux) return HH_OK;
    HH_HIP(hipSetDevice(s->ctx->device));
    HH_HIP(hipStreamSynchronize(s->ctx->stream));
    s->ctx->drop_graphs();
    HH_HIP(hipMalloc(&s->aux, 2 * sizeof(float4) * s->w * s->h));
    HH_HIP(launch_fill_f4(s->ctx->stream, s->aux, make_float4(0.f, 0.f, 0.f, 0.f), 2 * s->w * s->h));
    return HH_OK;
}

hh_status hh_surface_read_features(hh_surface s, float* out) {
    HH_ARG(s && out, "null argument");
    HH_ARG(s->aux, "the surface has no feature planes (hh_surface_enable_features)");
    HH_HIP(hipMemcpyAsync(out, s->aux, 2 * sizeof(float4) * s->w * s->h, hipMemcpyDeviceToHost, s->ctx->stream));
    HH_HIP(hipStreamSynchronize(s->ctx->stream));
    return HH_OK;
}

hh_status hh_surface_read(hh_surface s, float* out) {
    HH_ARG(s && out, "null argument");
    HH_HIP(hipMemcpyAsync(out, s->d, size_t(s->w) * s->h * sizeof(float4), hipMemcpyDeviceToHost, s->ctx->stream));
    HH_HIP(hipStreamSynchronize(s->ctx->stream));
    return HH_OK;
}

hh_status hh_surface_write(hh_surface s, const float* in) {
    HH_ARG(s && in, "null argument");
    HH_HIP(hipMemcpyAsync(s->d, in, size_t(s->w) * s->h * sizeof(float4), hipMemcpyHostToDevice, s->ctx->stream));
    HH_HIP(hipStreamSynchronize(s->ctx->stream));
    return HH_OK;
}

hh_status hh_surface_validate(hh_surface s, uint64_t* nonfinite) {
    HH_ARG(s, "null surface");
    hh_context ctx = s->ctx;
    HH_HIP(hipSetDevice(ctx->device));
    unsigned long long* d = nullptr;
    HH_HIP(hipMalloc((void**)&d, sizeof(unsigned long long)));
    hipError_t e = hipMemsetAsync(d, 0, sizeof(unsigned long long), ctx->stream);
    if (e == hipSuccess) e = launch_count_nonfinite(ctx->stream, s->d, size_t(s->w) * s->h, d);
    unsigned long long n = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&n, d, sizeof(n), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    (void)hipFree(d);
    if (e != hipSuccess) return hip_fail(e, "surface validation");
    if (nonfinite) *nonfinite = n;
    if (n) return fail(HH_ERR_NAN_DETECTED, "%llu non-finite values in the surface", n);
    return HH_OK;
}

hh_status hh_surface_device_ptr(hh_surface s, void** out) {
    HH_ARG(s && out, "null argument");
    *out = s->d;
    return HH_OK;
}

hh_status hh_tonemap(hh_surface s, float exposure, uint32_t op, uint8_t* out) {
    HH_ARG(s && out, "null argument");
    hh_context ctx = s->ctx;
    size_t n = size_t(s->w) * s->h;
    if (ctx->dbg_cap < n * 4) {
        if (ctx->dbg) (void)hipFree(ctx->dbg);
        ctx->dbg = nullptr;
        ctx->dbg_cap = 0;
        HH_HIP(hipMalloc(&ctx->dbg, n * 4));
        ctx->dbg_cap = n * 4;
    }
    HH_HIP(launch_tonemap(ctx->stream, s->d, (uint32_t)n, exposure, op, (uchar4*)ctx->dbg));
    HH_HIP(hipMemcpyAsync(out, ctx->dbg, n * 4, hipMemcpyDeviceToHost, ctx->stream));
    HH_HIP(hipStreamSynchronize(ctx->stream));
    return HH_OK;
}

// ---------------------------------------------------------------------------
// Frame launches (direct or replayed from a captured hipGraph)
// ---------------------------------------------------------------------------
static hipError_t ensure_tables(hh_context ctx, uint32_t nscenes, uint32_t nviews) {
    if (nscenes > ctx->scene_table_cap) {
        if (ctx->d_scene_table) HH_TRY(hipFree(ctx->d_scene_table));
        ctx->d_scene_table = nullptr;
        ctx->last_scenes.clear();
        ctx->drop_graphs();
        uint32_t cap = std::max<uint32_t>(nscenes, 64);
        HH_TRY(hipMalloc(&ctx->d_scene_table, cap * sizeof(SceneGPU)));
        ctx->scene_table_cap = cap;
    }
    if (nviews > ctx->views_cap) {
        if (ctx->d_views) HH_TRY(hipFree(ctx->d_views));
        ctx->d_views = nullptr;
        ctx->drop_graphs();
        uint32_t cap = std::max<uint32_t>(nviews, 64);
        HH_TRY(hipMalloc(&ctx->d_views, cap * sizeof(ViewGPU)));
        ctx->views_cap = cap;
    }
    return hipSuccess;
}

// Path slots for `total` paths and queue segments for every block size a
// frame may be launched with (the profiled frames use 256-thread blocks).
static hipError_t ensure_wavefront(hh_context ctx, uint32_t total) {
    uint64_t qneed = 0;
    for (int bs : {64, 128, 256}) {
        const uint32_t g = frame_grid(total, ctx->num_cu, ctx->tune.blocks_per_cu, bs);
        qneed = std::max<uint64_t>(qneed, queue_segment(total, g, (uint32_t)bs) * kQShards);
    }
    return ctx->wf.ensure(total, qneed);
}

// The frame's scene table (skipped while the device copy is unchanged) and
// view records, through the context's pinned staging.
static hipError_t upload_tables(hh_context ctx, const std::vector<SceneGPU>& scenes, const std::vector<ViewGPU>& views) {
    const size_t sb = scenes.size() * sizeof(SceneGPU), vb = views.size() * sizeof(ViewGPU);
    const bool same = ctx->last_scenes.size() == scenes.size() &&
                      std::memcmp(ctx->last_scenes.data(), scenes.data(), sb) == 0;
    const size_t need = (same ? 0 : sb) + vb;
    HH_TRY(hipEventSynchronize(ctx->ev_staged));   // the previous upload has left the staging
    if (need > ctx->stage_cap) {
        if (ctx->h_stage) HH_TRY(hipHostFree(ctx->h_stage));
        ctx->h_stage = nullptr;
        ctx->stage_cap = 0;
        const size_t cap = std::max<size_t>(need, 16384);
        HH_TRY(hipHostMalloc((void**)&ctx->h_stage, cap));
        ctx->stage_cap = cap;
    }
    size_t off = 0;
    if (!same) {
        std::memcpy(ctx->h_stage, scenes.data(), sb);
        HH_TRY(hipMemcpyAsync(ctx->d_scene_table, ctx->h_stage, sb, hipMemcpyHostToDevice, ctx->stream));
        off = sb;
    }
    std::memcpy(ctx->h_stage + off, views.data(), vb);
    HH_TRY(hipMemcpyAsync(ctx->d_views, ctx->h_stage + off, vb, hipMemcpyHostToDevice, ctx->stream));
    HH_TRY(hipEventRecord(ctx->ev_staged, ctx->stream));
    if (!same) ctx->last_scenes = scenes;
    return hipSuccess;
}

static hh_status run_frame(hh_context ctx, const std::vector<SceneGPU>& scenes, const std::vector<ViewGPU>& views,
                           uint32_t total, uint32_t max_bounces, bool mega, const std::string& key_extra) {
    HH_HIP(ensure_tables(ctx, (uint32_t)scenes.size(), (uint32_t)views.size()));
    uint32_t cap_before = ctx->wf.capacity;
    uint32_t* spill_before = ctx->wf.spill;
    const uint64_t qcap_before = ctx->wf.qcap;
    HH_HIP(ensure_wavefront(ctx, total));
    // Spill area for the most threads a frame launches at once: twice the
    // traversal grid (the shadow launch on the side stream has its own half),
    // and the counting frames' 256-thread blocks. launch_frame runs blocks of
    // 64 or 128 threads when extend_block says so, else 256.
    {
        const int eb = ctx->tune.extend_block, bs = (eb == 64 || eb == 128) ? eb : 256;
        const size_t per_block = std::max<size_t>(2 * (size_t)bs, 256);
        HH_HIP(ctx->wf.ensure_spill((size_t)ctx->num_cu * ctx->tune.blocks_per_cu * per_block));
    }
    if (ctx->wf.capacity != cap_before || ctx->wf.qcap != qcap_before || ctx->wf.spill != spill_before) ctx->drop_graphs();
    HH_HIP(upload_tables(ctx, scenes, views));
    LaunchCfg cfg{ctx->stream, ctx->num_cu, ctx->tune};
    cfg.side = ctx->side;
    cfg.fork = ctx->ev_fork;
    cfg.join = ctx->ev_join;
    if (ctx->prof.timing || ctx->prof.count) {
        // profiled frames are launched directly so every kernel gets its events
        cfg.prof = &ctx->prof;
        HH_HIP(launch_frame(cfg, ctx->d_scene_table, scenes.data(), ctx->d_views, views.data(), (uint32_t)views.size(), total,
                            max_bounces, ctx->wf, mega));
        return HH_OK;
    }
    if (!ctx->use_graphs) {
        HH_HIP(launch_frame(cfg, ctx->d_scene_table, scenes.data(), ctx->d_views, views.data(), (uint32_t)views.size(), total,
                            max_bounces, ctx->wf, mega));
        return HH_OK;
    }
    // The launch sequence depends on the view geometry (tile sizes, spp) and
    // the bounce count, not on camera or frame index (read from d_views).
    std::string key = key_extra + "|" + std::to_string(total) + "|" + std::to_string(max_bounces) + "|" + std::to_string(mega) +
                      "|" + std::to_string(views.size());
    for (const ViewGPU& v : views)
        key += "|" + std::to_string(v.tile_w) + "x" + std::to_string(v.tile_h) + "s" + std::to_string(v.spp);
    auto it = ctx->graphs.find(key);
    if (it == ctx->graphs.end()) {
        GraphEntry ge;
        HH_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        hipError_t e = launch_frame(cfg, ctx->d_scene_table, scenes.data(), ctx->d_views, views.data(), (uint32_t)views.size(), total,
                                    max_bounces, ctx->wf, mega);
        hipError_t e2 = hipStreamEndCapture(ctx->stream, &ge.graph);
        if (e != hipSuccess) return hip_fail(e, "frame capture");
        if (e2 != hipSuccess) return hip_fail(e2, "hipStreamEndCapture");
        HH_HIP(hipGraphInstantiate(&ge.exec, ge.graph, nullptr, nullptr, 0));
        it = ctx->graphs.emplace(key, ge).first;
    }
    HH_HIP(hipGraphLaunch(it->second.exec, ctx->stream));
    return HH_OK;
}

// ---------------------------------------------------------------------------
// Integrator (PathIntegrator)
// ---------------------------------------------------------------------------
hh_status hh_integrator_create(hh_context ctx, hh_integrator* out) {
    HH_ARG(ctx && out, "null argument");
    hh_integrator it = new (std::nothrow) hh_integrator_t();
    if (!it) return fail(HH_ERR_OUT_OF_MEMORY, "integrator");
    it->ctx = ctx;
    *out = it;
    return HH_OK;
}

hh_status hh_integrator_destroy(hh_integrator it) {
    delete it;
    return HH_OK;
}

hh_status hh_integrator_set_settings(hh_integrator it, const hh_settings* s) {
    HH_ARG(it && s, "null argument");
    HH_ARG(s->max_ray_bounces >= 1 && s->max_ray_bounces <= 64, "max_ray_bounces must be in [1, 64]");
    HH_ARG(s->max_samples >= 1, "max_samples must be >= 1");
    HH_ARG(std::isfinite(s->shadow_ray_bias), "shadow_ray_bias must be finite");
    if ((s->tiled != 0) != (it->s.tiled != 0)) it->tiles_for_tiled = 2;  // set_tiled recomputes tiles
    it->s = *s;
    return HH_OK;
}

hh_status hh_integrator_get_settings(hh_integrator it, hh_settings* s) {
    HH_ARG(it && s, "null argument");
    *s = it->s;
    return HH_OK;
}

hh_status hh_integrator_restart_bake(hh_integrator it) {
    HH_ARG(it, "null integrator");
    it->num_acc = 0;
    it->tile_idx = 0;
    return HH_OK;
}

static void compute_tiles(hh_integrator it, uint32_t w, uint32_t h) {  // path_integrator.cpp:312-336
    it->tiles.clear();
    if (it->s.tiled) {
        const uint32_t T = 128;
        uint32_t nx = (w + T - 1) / T, ny = (h + T - 1) / T;
        for (uint32_t x = 0; x < nx; x++)
            for (uint32_t y = 0; y < ny; y++) it->tiles.push_back({x * T, y * T});
    } else {
        it->tiles.push_back({0, 0});
    }
    it->tiles_for_w = w;
    it->tiles_for_h = h;
    it->tiles_for_tiled = it->s.tiled ? 1 : 0;
}

hh_status hh_integrator_render_frames(hh_integrator it, hh_scene scene, const hh_camera* cam, hh_surface out,
                                      uint32_t frames) {
    HH_ARG(it && cam && out, "null argument");
    hh_status st = require_committed(scene);
    if (st != HH_OK) return st;
    HH_ARG(out->ctx == it->ctx && scene->ctx == it->ctx, "objects belong to different contexts");
    for (int i = 0; i < 3; i++) HH_ARG(std::isfinite(cam->position[i]), "non-finite camera position");
    HH_ARG(cam->fov_y_deg > 0.0f && cam->fov_y_deg < 180.0f, "fov_y must be in (0, 180)");
    hh_context ctx = it->ctx;
    HH_HIP(hipSetDevice(ctx->device));
    if (it->tiles_for_w != out->w || it->tiles_for_h != out->h || it->tiles_for_tiled != (it->s.tiled ? 1u : 0u)) {
        compute_tiles(it, out->w, out->h);  // on_window_resize / set_tiled
        it->num_acc = 0;
        it->tile_idx = 0;
    }
    if (scene->generation != it->last_generation || (const void*)scene != it->last_scene) {
        it->num_acc = 0;  // scene_state != READY (path_integrator.cpp:52-56)
        it->tile_idx = 0;
        it->last_generation = scene->generation;
        it->last_scene = scene;
    }
    // Frames per pass: up to frames_in_flight, and never more path slots than
    // kMaxPathSlots (a 16384^2 surface still takes 8 frames per pass), so
    // slot and queue indices stay within 32 bits.
    const uint64_t px = (uint64_t)out->w * out->h;
    const uint32_t fif = std::max<uint32_t>(1, std::min<uint64_t>(ctx->tune.frames_in_flight, kMaxPathSlots / px));
    // size the wavefront for the largest chunk first, so the statistics
    // counters are not reallocated after they are cleared
    const uint32_t max_chunk = std::max<uint32_t>(1, std::min<uint32_t>(frames, fif));
    HH_HIP(ensure_wavefront(ctx, (uint32_t)(px * max_chunk)));
    HH_HIP(hipMemsetAsync(ctx->wf.stats, 0, 2 * sizeof(unsigned long long), ctx->stream));
    const bool mega = (it->s.flags & HH_FLAG_MEGAKERNEL) != 0;
    std::vector<SceneGPU> scenes{scene->gpu};
    uint32_t launched = 0;
    // Frames are traced in chunks of up to frames_in_flight progressive
    // samples per wavefront pass (never crossing a tile's max_samples).
    uint32_t f = 0;
    while (f < frames) {
        if (it->num_acc == 0 && it->tile_idx == 0) {
            HH_HIP(launch_fill_f4(ctx->stream, out->d, make_float4(0.f, 0.f, 0.f, 1.f), out->w * out->h));
            if (out->aux) HH_HIP(launch_fill_f4(ctx->stream, out->aux, make_float4(0.f, 0.f, 0.f, 0.f), 2 * out->w * out->h));
        }
        if (it->tile_idx >= it->tiles.size()) {
            f++;
            continue;
        }
        uint32_t chunk = std::min(frames - f, fif);
        chunk = std::min(chunk, it->s.max_samples - it->num_acc);
        ViewGPU v = make_view(*cam, out->w, out->h);
        auto t = it->tiles[it->tile_idx];
        v.tile_x = t.first;
        v.tile_y = t.second;
        v.tile_w = it->s.tiled ? std::min<uint32_t>(128, out->w - t.first) : out->w;
        v.tile_h = it->s.tiled ? std::min<uint32_t>(128, out->h - t.second) : out->h;
        v.spp = chunk;
        v.first_frame = it->num_acc;
        v.max_bounces = it->s.max_ray_bounces;
        v.rr_start = it->s.rr_start_depth;
        v.shadow_bias = it->s.shadow_ray_bias;
        v.flags = it->s.flags;
        v.scene = 0;
        v.slot_base = 0;
        v.accum = out->d;
        v.aux = out->aux;
        std::vector<ViewGPU> views{v};
        st = run_frame(ctx, scenes, views, v.tile_w * v.tile_h * chunk, it->s.max_ray_bounces, mega,
                       std::to_string(scene->generation) + "|" + std::to_string((uintptr_t)out->d));
        if (st != HH_OK) return st;
        it->num_acc += chunk;
        launched += chunk;
        f += chunk;
        if (it->num_acc == it->s.max_samples) {
            it->num_acc = 0;
            it->tile_idx++;
        }
    }
    HH_HIP(hipMemcpyAsync(ctx->h_stats, ctx->wf.stats, 2 * sizeof(unsigned long long), hipMemcpyDeviceToHost, ctx->stream));
    HH_HIP(hipStreamSynchronize(ctx->stream));
    it->ray_stats[0] = ctx->h_stats[0];
    it->ray_stats[1] = ctx->h_stats[1];
    it->ray_stats[2] = launched;
    return HH_OK;
}

hh_status hh_integrator_render(hh_integrator it, hh_scene scene, const hh_camera* cam, hh_surface out) {
    return hh_integrator_render_frames(it, scene, cam, out, 1);
}

hh_status hh_integrator_counters(hh_integrator it, uint32_t out[4]) {
    HH_ARG(it && out, "null argument");
    out[0] = it->s.max_samples * it->tile_idx + it->num_acc;              // num_accumulated_samples()
    out[1] = it->s.max_samples * (uint32_t)std::max<size_t>(it->tiles.size(), 1);  // num_target_samples()
    out[2] = it->tile_idx;
    out[3] = it->num_acc;
    return HH_OK;
}

hh_status hh_integrator_ray_stats(hh_integrator it, uint64_t out[3]) {
    HH_ARG(it && out, "null argument");
    for (int i = 0; i < 3; i++) out[i] = it->ray_stats[i];
    return HH_OK;
}

hh_status hh_compute_launch_consts(const hh_camera* cam, uint32_t w, uint32_t h, uint32_t num_lights, uint32_t num_frames,
                                   const hh_settings* s, hh_launch_consts* out) {
    HH_ARG(cam && s && out && w > 0 && h > 0, "bad argument");
    consts_from_camera(*cam, w, h, *out);
    out->num_lights = num_lights;
    out->num_frames = num_frames;
    out->accumulation = float(num_frames) / float(num_frames + 1);
    out->max_ray_bounces = s->max_ray_bounces;
    out->shadow_ray_bias = s->shadow_ray_bias;
    return HH_OK;
}

// ---------------------------------------------------------------------------
// Batched rendering
// ---------------------------------------------------------------------------
hh_status hh_render_batch(hh_context ctx, const hh_view* views, uint32_t n) {
    HH_ARG(ctx && views && n > 0, "bad argument");
    HH_HIP(hipSetDevice(ctx->device));
    uint32_t max_spp = 0;
    for (uint32_t i = 0; i < n; i++) {
        hh_status st = require_committed(views[i].scene);
        if (st != HH_OK) return st;
        HH_ARG(views[i].surface && views[i].surface->ctx == ctx && views[i].scene->ctx == ctx, "view objects mismatch");
        HH_ARG(views[i].max_ray_bounces >= 1 && views[i].max_ray_bounces <= 64, "max_ray_bounces must be in [1, 64]");
        max_spp = std::max(max_spp, views[i].spp);
    }
    std::map<hh_scene, uint32_t> scene_index;
    std::vector<SceneGPU> scenes;
    std::string key = "batch";
    for (uint32_t i = 0; i < n; i++)
        if (!scene_index.count(views[i].scene)) {
            scene_index[views[i].scene] = (uint32_t)scenes.size();
            scenes.push_back(views[i].scene->gpu);
            key += "|" + std::to_string(views[i].scene->generation);
        }
    for (auto& sg : scenes) HH_ARG(sg.bvh_width == scenes[0].bvh_width, "scenes of one batch must share the BVH width");
    unsigned long long zero[2] = {0, 0};
    size_t total_px = 0;
    for (uint32_t i = 0; i < n; i++) total_px += (size_t)views[i].surface->w * views[i].surface->h;
    HH_ARG(total_px <= 0xFFFFFFFFull, "batch too large");
    // Samples per wavefront pass: up to frames_in_flight per view, and no
    // more path slots than queue_capacity (when set). Every view's samples
    // are still accumulated in frame order.
    uint32_t chunk = std::max<uint32_t>(1, ctx->tune.frames_in_flight);
    if (ctx->tune.queue_capacity > 0)
        chunk = std::max<uint32_t>(1, std::min<uint32_t>(chunk, (uint32_t)(ctx->tune.queue_capacity / std::max<size_t>(total_px, 1))));
    chunk = std::max<uint32_t>(1, std::min<uint64_t>(chunk, kMaxPathSlots / std::max<size_t>(total_px, 1)));
    HH_ARG(total_px * std::min(chunk, max_spp) <= kMaxPathSlots, "batch too large");
    HH_HIP(ensure_wavefront(ctx, (uint32_t)(total_px * std::min(chunk, max_spp))));
    HH_HIP(hipMemcpyAsync(ctx->wf.stats, zero, sizeof(zero), hipMemcpyHostToDevice, ctx->stream));
    for (uint32_t s = 0; s < max_spp; s += chunk) {
        std::vector<ViewGPU> vg;
        size_t total = 0;
        uint32_t maxb = 1;
        for (uint32_t i = 0; i < n; i++) {
            if (views[i].spp <= s) continue;
            const hh_surface sf = views[i].surface;
            ViewGPU v = make_view(views[i].camera, sf->w, sf->h);
            v.spp = std::min(chunk, views[i].spp - s);
            v.first_frame = views[i].first_frame + s;
            v.max_bounces = views[i].max_ray_bounces;
            v.shadow_bias = views[i].shadow_ray_bias;
            v.flags = views[i].flags & (HH_FLAG_NO_RADIANCE_CLAMP | HH_FLAG_DIRECT_ONLY | HH_FLAG_LIGHT_POWER);
            v.scene = scene_index[views[i].scene];
            v.slot_base = (uint32_t)total;
            v.accum = sf->d;
            v.aux = sf->aux;
            total += (size_t)sf->w * sf->h * v.spp;
            maxb = std::max(maxb, v.max_bounces);
            vg.push_back(v);
        }
        HH_ARG(total <= kMaxPathSlots, "batch too large");
        std::string k = key;
        for (uint32_t i = 0; i < n; i++) k += "|" + std::to_string((uintptr_t)views[i].surface->d);
        hh_status st = run_frame(ctx, scenes, vg, (uint32_t)total, maxb, false, k);
        if (st != HH_OK) return st;
    }
    return HH_OK;
}

// ---------------------------------------------------------------------------
// Autotuner
// ---------------------------------------------------------------------------
// Milliseconds per progressive frame of `views` (rendered into the scratch
// surfaces they name) under the context's current tuning.
static hh_status time_views(hh_context ctx, const std::vector<hh_view>& views, uint32_t frames, double* ms) {
    std::vector<hh_view> v = views;
    for (auto& x : v) x.spp = std::max<uint32_t>(1, std::min<uint32_t>(frames, ctx->tune.frames_in_flight));
    hh_status st = hh_render_batch(ctx, v.data(), (uint32_t)v.size());  // warm-up: graphs captured
    if (st != HH_OK) return st;
    for (auto& x : v) x.spp = frames;
    hipEvent_t a, b;
    HH_HIP(hipEventCreate(&a));
    HH_HIP(hipEventCreate(&b));
    HH_HIP(hipEventRecord(a, ctx->stream));
    st = hh_render_batch(ctx, v.data(), (uint32_t)v.size());
    hipError_t e = hipEventRecord(b, ctx->stream);
    if (e == hipSuccess) e = hipEventSynchronize(b);
    float t = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, a, b);
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (st != HH_OK) return st;
    if (e != hipSuccess) return hip_fail(e, "autotune timing");
    *ms = (double)t / frames;
    return HH_OK;
}

hh_status hh_context_autotune(hh_context ctx, const hh_view* views, uint32_t n, uint32_t frames, uint32_t flags,
                              hh_tuning* out, double* out_ms) {
    HH_ARG(ctx && views && n > 0, "bad argument");
    HH_ARG(frames >= 1 && frames <= 256, "frames must be 1..256");
    HH_HIP(hipSetDevice(ctx->device));
    // scratch surfaces: the caller's images are not touched
    std::vector<hh_surface> scratch;
    std::vector<hh_view> v(views, views + n);
    hh_status st = HH_OK;
    for (uint32_t i = 0; i < n && st == HH_OK; i++) {
        HH_ARG(views[i].surface && views[i].scene, "view without surface or scene");
        hh_surface sf = nullptr;
        st = hh_surface_create(ctx, views[i].surface->w, views[i].surface->h, &sf);
        if (st == HH_OK) {
            scratch.push_back(sf);
            v[i].surface = sf;
            v[i].first_frame = 0;
        }
    }
    auto cleanup = [&]() {
        for (hh_surface sf : scratch) hh_surface_destroy(sf);
    };
    if (st != HH_OK) {
        cleanup();
        return st;
    }
    Tuning best = ctx->tune;
    double best_ms = 0.0;
    auto trial = [&](const Tuning& t, double* ms) -> hh_status {
        ctx->tune = t;
        ctx->drop_graphs();
        return time_views(ctx, v, frames, ms);
    };
    st = trial(best, &best_ms);
    // coordinate descent, one knob at a time, twice over
    for (int pass = 0; pass < 2 && st == HH_OK; pass++) {
        const uint32_t fif[] = {1, 2, 4, 8, 16};
        const int blk[] = {64, 128, 256};
        const int bpc[] = {4, 8, 16, 32};
        const uint32_t rf[] = {16, 24, 32, 48, 64};
        const int trav[] = {0, 1};
        for (int knob = 0; knob < 5 && st == HH_OK; knob++) {
            const int count = knob == 0 ? 5 : knob == 1 ? 3 : knob == 2 ? 4 : knob == 3 ? 5 : 2;
            for (int c = 0; c < count && st == HH_OK; c++) {
                Tuning t = best;
                if (knob == 0) t.frames_in_flight = fif[c];
                if (knob == 1) t.extend_block = t.shade_block = blk[c];
                if (knob == 2) t.blocks_per_cu = bpc[c];
                if (knob == 3) t.refill_lanes = rf[c];
                if (knob == 4) t.traversal = trav[c];
                if (t.frames_in_flight == best.frames_in_flight && t.extend_block == best.extend_block &&
                    t.blocks_per_cu == best.blocks_per_cu && t.refill_lanes == best.refill_lanes && t.traversal == best.traversal)
                    continue;
                double ms = 0.0;
                st = trial(t, &ms);
                if (st == HH_OK && ms < best_ms) {
                    best = t;
                    best_ms = ms;
                }
            }
        }
    }
    if (st == HH_OK && (flags & HH_AUTOTUNE_BVH_WIDTH)) {
        // BVH width is a property of the committed scenes: re-commit them
        std::vector<hh_scene> scenes;
        for (uint32_t i = 0; i < n; i++)
            if (std::find(scenes.begin(), scenes.end(), views[i].scene) == scenes.end()) scenes.push_back(views[i].scene);
        const int cur = best.bvh_width;
        Tuning t = best;
        t.bvh_width = cur == 4 ? 8 : 4;
        t.bvh_compressed = false;
        ctx->tune = t;
        for (hh_scene sc : scenes)
            if (st == HH_OK) st = hh_scene_commit(sc);
        double ms = 0.0;
        if (st == HH_OK) st = trial(t, &ms);
        if (st == HH_OK && ms < best_ms) {
            best = t;
            best_ms = ms;
        } else if (st == HH_OK) {
            ctx->tune = best;
            for (hh_scene sc : scenes)
                if (st == HH_OK) st = hh_scene_commit(sc);
        }
    }
    ctx->tune = best;
    ctx->drop_graphs();
    cleanup();
    if (st != HH_OK) return st;
    if (out) hh_context_get_tuning(ctx, out);
    if (out_ms) *out_ms = best_ms;
    return HH_OK;
}

// ---------------------------------------------------------------------------
// Debug tracing
// ---------------------------------------------------------------------------
static hipError_t ensure_dbg(hh_context ctx, size_t bytes) {
    if (ctx->dbg_cap >= bytes) return hipSuccess;
    if (ctx->dbg) HH_TRY(hipFree(ctx->dbg));
    ctx->dbg = nullptr;
    ctx->dbg_cap = 0;
    HH_TRY(hipMalloc(&ctx->dbg, bytes));
    ctx->dbg_cap = bytes;
    return hipSuccess;
}

static void to_hits(const std::vector<float4>& hit, const std::vector<uint32_t>& ids, hh_hit* out, size_t n) {
    for (size_t i = 0; i < n; i++) {
        float4 h = hit[i];
        out[i].t = h.x;
        out[i].u = h.y;
        out[i].v = h.z;
        uint32_t inst;
        std::memcpy(&inst, &h.w, 4);
        out[i].instance = inst;
        out[i].geometry = ids[2 * i];
        out[i].primitive = ids[2 * i + 1];
        if (h.x < 0.0f) {
            out[i].u = out[i].v = 0.0f;
            out[i].instance = out[i].geometry = out[i].primitive = 0xFFFFFFFFu;
        }
    }
}

hh_status hh_trace_rays(hh_scene s, const hh_ray* rays, uint32_t n, uint32_t flags, hh_hit* out) {
    HH_ARG(rays && out, "null argument");
    hh_status st = require_committed(s);
    if (st != HH_OK) return st;
    hh_context ctx = s->ctx;
    HH_HIP(hipSetDevice(ctx->device));
    size_t need = size_t(n) * (32 + 16 + 8);
    HH_HIP(ensure_dbg(ctx, need));
    float4* d_rays = (float4*)ctx->dbg;
    float4* d_hit = d_rays + 2 * (size_t)n;
    uint32_t* d_ids = (uint32_t*)(d_hit + n);
    HH_HIP(hipMemcpyAsync(d_rays, rays, size_t(n) * 32, hipMemcpyHostToDevice, ctx->stream));
    LaunchCfg cfg{ctx->stream, ctx->num_cu, ctx->tune};
    HH_HIP(ctx->wf.ensure_spill(((size_t)n + 255) / 256 * 256));
    ctx->drop_graphs();
    HH_HIP(launch_trace_rays(cfg, s->d_gpu, s->gpu, d_rays, n, flags, d_hit, d_ids, ctx->wf.spill));
    std::vector<float4> hit(n);
    std::vector<uint32_t> ids(2 * size_t(n));
    HH_HIP(hipMemcpyAsync(hit.data(), d_hit, size_t(n) * 16, hipMemcpyDeviceToHost, ctx->stream));
    HH_HIP(hipMemcpyAsync(ids.data(), d_ids, size_t(n) * 8, hipMemcpyDeviceToHost, ctx->stream));
    HH_HIP(hipStreamSynchronize(ctx->stream));
    to_hits(hit, ids, out, n);
    return HH_OK;
}

hh_status hh_trace_rays_noalloc(hh_scene s, const hh_ray* d_rays, uint32_t n, uint32_t flags, hh_hit* d_hits) {
    HH_ARG(d_rays && d_hits, "null argument");
    HH_ARG(((uintptr_t)d_rays & 15) == 0 && ((uintptr_t)d_hits & 3) == 0, "rays need 16-byte, hits 4-byte alignment");
    if (n > kNoAllocMaxRays)
        return fail(HH_ERR_LIMIT_EXCEEDED, "hh_trace_rays_noalloc: %u rays exceed the fixed scratch of %u rays", n,
                    kNoAllocMaxRays);
    hh_status st = require_committed(s);
    if (st != HH_OK) return st;
    hh_context ctx = s->ctx;
    HH_HIP(hipSetDevice(ctx->device));
    void* base = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        base = g_scratch_ptr[ctx->device];
    }
    if (!base) return fail(HH_ERR_DEVICE, "hh_trace_rays_noalloc: no scratch address for device %d", ctx->device);
    if (n == 0) return HH_OK;
    uint32_t* spill = (uint32_t*)base;
    float4* d_hit = (float4*)((char*)base + kNoAllocSpillBytes);
    uint32_t* d_ids = (uint32_t*)((char*)base + kNoAllocSpillBytes + kNoAllocHitBytes);
    LaunchCfg cfg{ctx->stream, ctx->num_cu, ctx->tune};
    // The spill area of a k_trace_rays launch is strided by its thread count,
    // cdiv(n, 256) * 256 <= kNoAllocMaxRays.
    HH_HIP(launch_trace_rays(cfg, s->d_gpu, s->gpu, (const float4*)d_rays, n, flags, d_hit, d_ids, spill));
    HH_HIP(launch_pack_hits(ctx->stream, d_hit, d_ids, n, d_hits));
    return HH_OK;
}

hh_status hh_primary_hits(hh_scene s, const hh_camera* cam, uint32_t w, uint32_t h, uint32_t frame, hh_hit* out) {
    HH_ARG(cam && out && w > 0 && h > 0 && w <= 16384 && h <= 16384, "bad argument");
    hh_status st = require_committed(s);
    if (st != HH_OK) return st;
    hh_context ctx = s->ctx;
    HH_HIP(hipSetDevice(ctx->device));
    size_t n = size_t(w) * h;
    HH_HIP(ensure_dbg(ctx, n * 24 + sizeof(ViewGPU) + 256));
    ViewGPU* d_view = (ViewGPU*)ctx->dbg;
    float4* d_hit = (float4*)((char*)ctx->dbg + 256 * ((sizeof(ViewGPU) + 255) / 256));
    uint32_t* d_ids = (uint32_t*)(d_hit + n);
    ViewGPU v = make_view(*cam, w, h);
    HH_HIP(hipMemcpyAsync(d_view, &v, sizeof(v), hipMemcpyHostToDevice, ctx->stream));
    LaunchCfg cfg{ctx->stream, ctx->num_cu, ctx->tune};
    HH_HIP(ctx->wf.ensure_spill((n + 255) / 256 * 256));
    ctx->drop_graphs();
    HH_HIP(launch_primary_hits(cfg, s->d_gpu, s->gpu, d_view, w, h, frame, d_hit, d_ids, ctx->wf.spill));
    std::vector<float4> hit(n);
    std::vector<uint32_t> ids(2 * n);
    HH_HIP(hipMemcpyAsync(hit.data(), d_hit, n * 16, hipMemcpyDeviceToHost, ctx->stream));
    HH_HIP(hipMemcpyAsync(ids.data(), d_ids, n * 8, hipMemcpyDeviceToHost, ctx->stream));
    HH_HIP(hipStreamSynchronize(ctx->stream));
    to_hits(hit, ids, out, n);
    return HH_OK;
}

// ---------------------------------------------------------------------------
// Output-buffer views and the ray-debug variant
// ---------------------------------------------------------------------------
static hipError_t ensure_spill_keep_graphs(hh_context ctx, size_t lanes) {
    uint32_t* before = ctx->wf.spill;
    HH_TRY(ctx->wf.ensure_spill(lanes));
    if (ctx->wf.spill != before) ctx->drop_graphs();  // captured frames hold the old pointer
    return hipSuccess;
}

hh_status hh_render_output_buffer(hh_scene s, const hh_camera* cam, uint32_t which, hh_surface out) {
    HH_ARG(cam && out, "null argument");
    HH_ARG(which <= HH_OUTPUT_BUFFER_FINAL, "unknown output buffer");
    HH_ARG(which != HH_OUTPUT_BUFFER_FINAL, "HH_OUTPUT_BUFFER_FINAL is the path-traced surface, not a view");
    HH_ARG(cam->near_plane > 0.0f && cam->far_plane > cam->near_plane, "camera needs 0 < near_plane < far_plane");
    hh_status st = require_committed(s);
    if (st != HH_OK) return st;
    hh_context ctx = s->ctx;
    HH_ARG(out->ctx == ctx, "surface belongs to another context");
    HH_HIP(hipSetDevice(ctx->device));
    const uint32_t w = out->w, h = out->h;
    const size_t n = size_t(w) * h;
    HH_HIP(ensure_dbg(ctx, sizeof(ViewGPU)));
    ViewGPU* d_view = (ViewGPU*)ctx->dbg;
    ViewGPU v = make_view(*cam, w, h);
    HH_HIP(hipMemcpyAsync(d_view, &v, sizeof(v), hipMemcpyHostToDevice, ctx->stream));
    HH_HIP(ensure_spill_keep_graphs(ctx, (n + 255) / 256 * 256));
    const double fn = cam->far_plane, nn = cam->near_plane;
    const float near_d = (float)(2.0 * fn * nn / (fn + nn));
    LaunchCfg cfg{ctx->stream, ctx->num_cu, ctx->tune};
    HH_HIP(launch_output_buffer(cfg, s->d_gpu, s->gpu, d_view, w, h, near_d, cam->far_plane, which, out->d, ctx->wf.spill));
    // the view record was read from a stack variable
    HH_HIP(hipStreamSynchronize(ctx->stream));
    return HH_OK;
}

struct hh_debug_rays_t {
    hh_context ctx = nullptr;
    float4* vtx = nullptr;        // 2 float4 per vertex: position, colour
    uint32_t* args = nullptr;     // {vertex count, instance count, first vertex, first instance}
    uint32_t cap = 0;             // vertices
    uint32_t reset_args[4] = {0, 1, 0, 0};
    ~hh_debug_rays_t() {
        if (vtx) (void)hipFree(vtx);
        if (args) (void)hipFree(args);
    }
};

hh_status hh_debug_rays_create(hh_context ctx, uint32_t max_segments, hh_debug_rays* out) {
    HH_ARG(ctx && out && max_segments > 0 && max_segments <= (1u << 24), "bad argument");
    HH_HIP(hipSetDevice(ctx->device));
    hh_debug_rays r = new (std::nothrow) hh_debug_rays_t();
    if (!r) return fail(HH_ERR_OUT_OF_MEMORY, "debug ray buffer");
    r->ctx = ctx;
    r->cap = 2 * max_segments;
    hipError_t e = hipMalloc(&r->vtx, size_t(r->cap) * 32);
    if (e == hipSuccess) e = hipMalloc(&r->args, 16);
    if (e == hipSuccess) e = hipMemcpy(r->args, r->reset_args, 16, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        delete r;
        return hip_fail(e, "hh_debug_rays_create");
    }
    *out = r;
    return HH_OK;
}

hh_status hh_debug_rays_destroy(hh_debug_rays r) {
    if (!r) return HH_OK;
    (void)hipSetDevice(r->ctx->device);
    (void)hipStreamSynchronize(r->ctx->stream);
    delete r;
    return HH_OK;
}

hh_status hh_debug_rays_clear(hh_debug_rays r) {  // renderer.cpp:234-246
    HH_ARG(r, "null argument");
    HH_HIP(hipSetDevice(r->ctx->device));
    HH_HIP(hipMemcpyAsync(r->args, r->reset_args, 16, hipMemcpyHostToDevice, r->ctx->stream));
    return HH_OK;
}

hh_status hh_integrator_gather_debug_rays(hh_integrator it, hh_scene s, const hh_camera* cam, uint32_t width,
                                          uint32_t height, uint32_t pixel_x, uint32_t pixel_y, uint32_t num_debug_rays,
                                          hh_debug_rays out) {
    HH_ARG(it && cam && out && width > 0 && height > 0 && width <= 16384 && height <= 16384, "bad argument");
    HH_ARG(pixel_x < width && pixel_y <= height, "pixel outside the view");
    HH_ARG(num_debug_rays <= (1u << 20), "too many debug rays");
    hh_status st = require_committed(s);
    if (st != HH_OK) return st;
    hh_context ctx = s->ctx;
    HH_ARG(it->ctx == ctx && out->ctx == ctx, "objects belong to different contexts");
    HH_HIP(hipSetDevice(ctx->device));
    HH_HIP(ensure_dbg(ctx, sizeof(ViewGPU)));
    ViewGPU* d_view = (ViewGPU*)ctx->dbg;
    ViewGPU v = make_view(*cam, width, height);
    v.max_bounces = it->s.max_ray_bounces;
    v.shadow_bias = it->s.shadow_ray_bias;
    v.flags = it->s.flags & HH_FLAG_LIGHT_POWER;
    HH_HIP(hipMemcpyAsync(d_view, &v, sizeof(v), hipMemcpyHostToDevice, ctx->stream));
    HH_HIP(ensure_spill_keep_graphs(ctx, (size_t(num_debug_rays) + 63) / 64 * 64));
    LaunchCfg cfg{ctx->stream, ctx->num_cu, ctx->tune};
    // path_integrator.cpp:146: ray_debug_pixel_coord = (x, height - y, width, height)
    HH_HIP(launch_ray_debug(cfg, s->d_gpu, s->gpu, d_view, num_debug_rays, pixel_x, height - pixel_y, it->num_acc, out->vtx,
                            out->cap, out->args, ctx->wf.spill));
    HH_HIP(hipStreamSynchronize(ctx->stream));
    return HH_OK;
}

hh_status hh_debug_rays_read(hh_debug_rays r, hh_debug_ray_vertex* out, uint32_t max_vertices, uint32_t draw_args[4]) {
    HH_ARG(r && draw_args && (out || max_vertices == 0), "bad argument");
    HH_HIP(hipSetDevice(r->ctx->device));
    HH_HIP(hipMemcpyAsync(draw_args, r->args, 16, hipMemcpyDeviceToHost, r->ctx->stream));
    HH_HIP(hipStreamSynchronize(r->ctx->stream));
    const uint32_t n = std::min(std::min(draw_args[0], r->cap), max_vertices);
    if (n) {
        HH_HIP(hipMemcpyAsync(out, r->vtx, size_t(n) * 32, hipMemcpyDeviceToHost, r->ctx->stream));
        HH_HIP(hipStreamSynchronize(r->ctx->stream));
    }
    return HH_OK;
}

hh_status hh_debug_rays_device_ptr(hh_debug_rays r, void** vertices, void** draw_args) {
    HH_ARG(r && vertices && draw_args, "null argument");
    *vertices = r->vtx;
    *draw_args = r->args;
    return HH_OK;
}

hh_status hh_surface_save_png(hh_surface s, float exposure, uint32_t op, const char* path) {
    HH_ARG(s && path, "null argument");
    std::vector<uint8_t> px(size_t(s->w) * s->h * 4);
    hh_status st = hh_tonemap(s, exposure, op, px.data());
    if (st != HH_OK) return st;
    return hh_write_png(path, px.data(), s->w, s->h, 1u);
}

// ---------------------------------------------------------------------------
// Denoiser
// ---------------------------------------------------------------------------
hh_status hh_denoiser_num_params(uint64_t* out) {
    HH_ARG(out, "null argument");
    *out = dn::net_num_params();
    return HH_OK;
}

hh_status hh_denoiser_create(hh_context ctx, uint64_t seed, hh_denoiser* out) {
    HH_ARG(ctx && out, "null argument");
    HH_HIP(hipSetDevice(ctx->device));
    hh_denoiser d = new (std::nothrow) hh_denoiser_t();
    if (!d) return fail(HH_ERR_OUT_OF_MEMORY, "denoiser");
    d->ctx = ctx;
    hipError_t e = d->net.init();
    if (e != hipSuccess) {
        delete d;
        return hip_fail(e, "denoiser allocation");
    }
    std::vector<float> p(d->net.nparams, 0.0f);
    dn::net_he_init(seed, p.data());
    hh_status st = hh_denoiser_set_params(d, p.data(), p.size());
    if (st != HH_OK) {
        delete d;
        return st;
    }
    *out = d;
    return HH_OK;
}

hh_status hh_denoiser_destroy(hh_denoiser d) {
    if (!d) return HH_OK;
    (void)hipStreamSynchronize(d->ctx->stream);
    delete d;
    return HH_OK;
}

hh_status hh_denoiser_set_params(hh_denoiser d, const float* params, uint64_t n) {
    HH_ARG(d && params, "null argument");
    HH_ARG(n == d->net.nparams, "parameter count mismatch");
    hipStream_t st = d->ctx->stream;
    HH_HIP(hipMemcpyAsync(d->net.master, params, n * 4, hipMemcpyHostToDevice, st));
    HH_HIP(d->net.zero_state(st));
    HH_HIP(d->net.refresh_weights(st));
    HH_HIP(hipStreamSynchronize(st));
    return HH_OK;
}

static hh_status dn_read(hh_denoiser d, const float* src, float* out, uint64_t n) {
    HH_ARG(d && out, "null argument");
    HH_ARG(n == d->net.nparams, "parameter count mismatch");
    HH_HIP(hipMemcpyAsync(out, src, n * 4, hipMemcpyDeviceToHost, d->ctx->stream));
    HH_HIP(hipStreamSynchronize(d->ctx->stream));
    return HH_OK;
}
hh_status hh_denoiser_get_params(hh_denoiser d, float* out, uint64_t n) { return dn_read(d, d ? d->net.master : nullptr, out, n); }
hh_status hh_denoiser_get_grads(hh_denoiser d, float* out, uint64_t n) { return dn_read(d, d ? d->net.grad : nullptr, out, n); }

static uint32_t pad8(uint32_t v) { return (v + 7u) & ~7u; }

// Uploads raw records (and targets) and packs the network input.
static hh_status dn_stage(hh_denoiser d, const float* feats, const float* target, uint32_t n, uint32_t h, uint32_t w,
                          bool training) {
    HH_ARG(d && feats, "null argument");
    HH_ARG(n >= 1 && h >= 1 && w >= 1 && (size_t)n * pad8(h) * pad8(w) <= (size_t(1) << 28), "bad batch size");
    hipStream_t st = d->ctx->stream;
    HH_HIP(hipSetDevice(d->ctx->device));
    const size_t px = (size_t)n * h * w;
    if (px > d->stage_px) {
        if (d->d_feat) (void)hipFree(d->d_feat);
        if (d->d_target) (void)hipFree(d->d_target);
        d->d_feat = d->d_target = nullptr;
        d->stage_px = 0;
        HH_HIP(hipMalloc(&d->d_feat, px * 12 * 4));
        HH_HIP(hipMalloc(&d->d_target, px * 3 * 4));
        d->stage_px = px;
    }
    HH_HIP(d->net.ensure(n, pad8(h), pad8(w), training));
    HH_HIP(hipMemcpyAsync(d->d_feat, feats, px * 12 * 4, hipMemcpyHostToDevice, st));
    if (target) HH_HIP(hipMemcpyAsync(d->d_target, target, px * 3 * 4, hipMemcpyHostToDevice, st));
    HH_HIP(d->net.pack(st, d->d_feat, nullptr, nullptr, 0, h, w));
    return HH_OK;
}

hh_status hh_denoiser_forward(hh_denoiser d, const float* feats, uint32_t n, uint32_t h, uint32_t w, float* out_log) {
    HH_ARG(out_log, "null argument");
    hh_status s = dn_stage(d, feats, nullptr, n, h, w, false);
    if (s != HH_OK) return s;
    hipStream_t st = d->ctx->stream;
    HH_HIP(d->net.forward(st, nullptr, h, w));
    const uint32_t Hp = d->net.H, Wp = d->net.W;
    std::vector<float> o((size_t)n * Hp * Wp * 4);
    HH_HIP(hipMemcpyAsync(o.data(), d->net.out, o.size() * 4, hipMemcpyDeviceToHost, st));
    HH_HIP(hipStreamSynchronize(st));
    for (uint32_t i = 0; i < n; i++)
        for (uint32_t y = 0; y < h; y++)
            for (uint32_t x = 0; x < w; x++)
                for (int c = 0; c < 3; c++)
                    out_log[(((size_t)i * h + y) * w + x) * 3 + c] = o[(((size_t)i * Hp + y) * Wp + x) * 4 + c];
    return HH_OK;
}

hh_status hh_denoiser_loss_grad(hh_denoiser d, const float* feats, const float* target, uint32_t n, uint32_t h, uint32_t w,
                                float* loss) {
    HH_ARG(target, "null target");
    hh_status s = dn_stage(d, feats, target, n, h, w, true);
    if (s != HH_OK) return s;
    hipStream_t st = d->ctx->stream;
    HH_HIP(d->net.forward(st, nullptr, h, w));
    HH_HIP(d->net.loss_and_backward(st, d->d_target, h, w, d->loss_kind));
    float l = 0.0f;
    HH_HIP(hipMemcpyAsync(&l, d->net.loss, 4, hipMemcpyDeviceToHost, st));
    HH_HIP(hipStreamSynchronize(st));
    if (loss) *loss = l / (3.0f * (float)n * (float)h * (float)w);
    return HH_OK;
}

hh_status hh_denoiser_set_loss(hh_denoiser d, uint32_t kind) {
    HH_ARG(d, "null denoiser");
    HH_ARG(kind == HH_DENOISER_LOSS_LOG_MSE || kind == HH_DENOISER_LOSS_RELMSE, "unknown loss");
    d->loss_kind = kind;
    return HH_OK;
}

hh_status hh_denoiser_train_step(hh_denoiser d, const float* feats, const float* target, uint32_t n, uint32_t h, uint32_t w,
                                 float lr, float* loss) {
    HH_ARG(std::isfinite(lr) && lr >= 0.0f, "learning rate must be finite and >= 0");
    hh_status s = hh_denoiser_loss_grad(d, feats, target, n, h, w, loss);
    if (s != HH_OK) return s;
    HH_HIP(d->net.adam(d->ctx->stream, lr));
    return HH_OK;
}

hh_status hh_denoise(hh_denoiser d, hh_surface in, uint32_t num_frames, hh_surface out) {
    HH_ARG(d && in && out, "null argument");
    HH_ARG(in->aux, "the input surface has no feature planes (hh_surface_enable_features)");
    HH_ARG(in->w == out->w && in->h == out->h, "surface sizes differ");
    HH_ARG(in != out, "denoise into a different surface");
    HH_ARG(num_frames >= 1, "num_frames must be >= 1");
    HH_ARG(in->ctx == d->ctx && out->ctx == d->ctx, "objects belong to different contexts");
    hipStream_t st = d->ctx->stream;
    HH_HIP(hipSetDevice(d->ctx->device));
    HH_HIP(d->net.ensure(1, pad8(in->h), pad8(in->w), false));
    const uint32_t held = num_frames >= 2 ? num_frames - 1 : 1;
    HH_HIP(d->net.pack(st, nullptr, in->d, in->aux, held, in->h, in->w));
    HH_HIP(d->net.forward(st, out->d, in->h, in->w));
    return HH_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Wavefront buffers
// ---------------------------------------------------------------------------
hipError_t hh::Wavefront::ensure(uint32_t n, uint64_t queue_entries) {
    if (n <= capacity && queue_entries <= qcap && T) return hipSuccess;
    const size_t N = std::max<uint32_t>(std::max(n, capacity), 1024);
    const size_t Q = std::max<uint64_t>(std::max(queue_entries, qcap), N);
    release();
    HH_TRY(hipMalloc(&T, N * 16));
    HH_TRY(hipMalloc(&L, N * 16));
    HH_TRY(hipMalloc(&rng, N * 8));
    for (int k = 0; k < 2; k++) {
        HH_TRY(hipMalloc(&q_o[k], Q * 16));
        HH_TRY(hipMalloc(&q_d[k], Q * 16));
        HH_TRY(hipMalloc(&q_slot[k], Q * 4));
    }
    HH_TRY(hipMalloc(&hit, Q * 16));
    HH_TRY(hipMalloc(&hit_tri, Q * 4));
    HH_TRY(hipMalloc(&s_o, Q * 16));
    HH_TRY(hipMalloc(&s_d, Q * 16));
    HH_TRY(hipMalloc(&s_c, Q * 16));
    HH_TRY(hipMalloc(&s_slot, Q * 4));
    HH_TRY(hipMalloc(&feat, N * 32));
    HH_TRY(hipMalloc(&counters, (size_t)kCtrKinds * kQShards * kCtrStride * 4));
    HH_TRY(hipMalloc(&stats, 64));
    capacity = (uint32_t)N;
    qcap = Q;
    return hipSuccess;
}

hipError_t hh::Wavefront::ensure_spill(size_t threads) {
    if (threads <= spill_threads && spill) return hipSuccess;
    if (spill) HH_TRY(hipFree(spill));
    spill = nullptr;
    spill_threads = 0;
    HH_TRY(hipMalloc(&spill, threads * kSpillStackDepth * sizeof(uint32_t)));
    spill_threads = threads;
    return hipSuccess;
}

void hh::Wavefront::release() {
    if (spill) (void)hipFree(spill);
    spill = nullptr;
    spill_threads = 0;
    void* ps[] = {T, L, rng, q_o[0], q_o[1], q_d[0], q_d[1], q_slot[0], q_slot[1], hit, hit_tri, s_o, s_d, s_c, s_slot, feat, counters, stats};
    for (void* p : ps)
        if (p) (void)hipFree(p);
    T = L = nullptr;
    rng = nullptr;
    q_o[0] = q_o[1] = q_d[0] = q_d[1] = nullptr;
    q_slot[0] = q_slot[1] = nullptr;
    hit = nullptr;
    hit_tri = nullptr;
    s_o = s_d = s_c = nullptr;
    s_slot = nullptr;
    feat = nullptr;
    counters = nullptr;
    stats = nullptr;
    capacity = 0;
    qcap = 0;
}
