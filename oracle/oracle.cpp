// oracle.cpp — CPU restatement of Helios's path-tracing hot path (see oracle.hpp).
//
// TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg through ctypes (liboracle.so). The product
// library never links or calls it.
//
// Parity status per function (what pins it; see tests/test_oracle_*.py):
//   RNG (random.glsl / sampling.glsl)     pinned: published xoroshiro64* and
//       Wang-hash reference values, brute-force period/bit checks.
//   camera consts (path_integrator.cpp)   pinned: closed form (pixel-centre ray
//       through un-jittered NDC point, focal plane at focal_length).
//   watertight intersection               pinned: FP64 brute force, closed-form
//       hits on axis-aligned quads, shared-edge watertightness on fans.
//   BRDF set (brdf.glsl)                  pinned: quadrature (∫D cos = 1,
//       ∫cos-lobe pdf = 1), Lambert limit, reciprocity of the GGX lobe.
//   lights / NEE / RR / miss              pinned: point light 1/r² closed form,
//       furnace test (white env, clamp off), FP32-vs-FP64 agreement.
//   sky (hosek_wilkie_sky_model.cpp)      pinned: sun luminance normalization
//       to 1.15, symmetry, spline end-points equal to table entries.
//   ACES / Reinhard (tone_map.frag)       pinned: closed forms at known points.
//   accumulation (path_trace_rgen.glsl)   pinned: exact algebra of Q5.
#include "oracle.hpp"

#include <atomic>
#include <cstdio>
#include <thread>

namespace orc {

// ===========================================================================
// Host math: quaternion, perspective, inverse (glm semantics, in double).
// ===========================================================================
struct M4 {
    double m[4][4];  // m[col][row], glm layout
};
static M4 m4_identity() {
    M4 r{};
    for (int i = 0; i < 4; i++) r.m[i][i] = 1.0;
    return r;
}
static M4 m4_mul(const M4& a, const M4& b) {
    M4 r{};
    for (int c = 0; c < 4; c++)
        for (int rr = 0; rr < 4; rr++) {
            double s = 0;
            for (int k = 0; k < 4; k++) s += a.m[k][rr] * b.m[c][k];
            r.m[c][rr] = s;
        }
    return r;
}
// General inverse by cofactors (double).
static M4 m4_inverse(const M4& a) {
    double inv[16], m[16];
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
// glm::mat4_cast(quat) with q = (x, y, z, w).
static M4 quat_to_m4(const float q4[4]) {
    double x = q4[0], y = q4[1], z = q4[2], w = q4[3];
    M4 r = m4_identity();
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
// glm: q * v
static void quat_rotate(const float q4[4], const double v[3], double out[3]) {
    double qx = q4[0], qy = q4[1], qz = q4[2], qw = q4[3];
    double uv[3] = {qy * v[2] - qz * v[1], qz * v[0] - qx * v[2], qx * v[1] - qy * v[0]};
    double uuv[3] = {qy * uv[2] - qz * uv[1], qz * uv[0] - qx * uv[2], qx * uv[1] - qy * uv[0]};
    for (int i = 0; i < 3; i++) out[i] = v[i] + ((uv[i] * qw) + uuv[i]) * 2.0;
}
// glm::perspective (right-handed, depth -1..1).
static M4 perspective(double fovy, double aspect, double zn, double zf) {
    M4 r{};
    double th = std::tan(fovy / 2.0);
    r.m[0][0] = 1.0 / (aspect * th);
    r.m[1][1] = 1.0 / th;
    r.m[2][2] = -(zf + zn) / (zf - zn);
    r.m[2][3] = -1.0;
    r.m[3][2] = -(2.0 * zf * zn) / (zf - zn);
    return r;
}

struct ConstsD {  // double version of the 192-B block (only the fields the path uses)
    double vpi[4][4];
    double cam[3], up[3], right[3], fplane[4];
    double aperture;
};

static void launch_consts_d(const hh_camera& cam, uint32_t w, uint32_t h, ConstsD& c) {
    // CameraNode::update (scene.cpp:632-646) + launch_rays (path_integrator.cpp:136-159)
    M4 R = quat_to_m4(cam.orientation);
    M4 T = m4_identity();
    for (int i = 0; i < 3; i++) T.m[3][i] = cam.position[i];
    M4 view = m4_inverse(m4_mul(T, R));
    M4 proj = perspective(cam.fov_y_deg * (3.14159265358979323846 / 180.0), double(w) / double(h), cam.near_plane,
                          cam.far_plane);
    M4 vpi = m4_inverse(m4_mul(proj, view));
    std::memcpy(c.vpi, vpi.m, sizeof(c.vpi));
    double ex[3] = {1, 0, 0}, ey[3] = {0, 1, 0}, ez[3] = {0, 0, 1}, fwd[3];
    quat_rotate(cam.orientation, ex, c.right);
    quat_rotate(cam.orientation, ey, c.up);
    quat_rotate(cam.orientation, ez, fwd);
    for (int i = 0; i < 3; i++) fwd[i] = -fwd[i];  // forward = -camera->forward()
    for (int i = 0; i < 3; i++) c.cam[i] = cam.position[i];
    double fpp[3];
    for (int i = 0; i < 3; i++) fpp[i] = c.cam[i] + fwd[i] * cam.focal_length;
    for (int i = 0; i < 3; i++) c.fplane[i] = -fwd[i];
    c.fplane[3] = -(c.fplane[0] * fpp[0] + c.fplane[1] * fpp[1] + c.fplane[2] * fpp[2]);
    c.aperture = cam.aperture_radius;
}

void launch_consts(const hh_camera& cam, uint32_t w, uint32_t h, uint32_t num_lights, uint32_t frame,
                   const hh_settings& st, hh_launch_consts& o) {
    ConstsD c;
    launch_consts_d(cam, w, h, c);
    std::memset(&o, 0, sizeof(o));
    for (int col = 0; col < 4; col++)
        for (int r = 0; r < 4; r++) o.view_proj_inverse[col * 4 + r] = (float)c.vpi[col][r];
    for (int i = 0; i < 3; i++) {
        o.camera_pos[i] = (float)c.cam[i];
        o.up_direction[i] = (float)c.up[i];
        o.right_direction[i] = (float)c.right[i];
    }
    for (int i = 0; i < 4; i++) o.focal_plane[i] = (float)c.fplane[i];
    o.launch_id_size[0] = 0;
    o.launch_id_size[1] = 0;
    o.launch_id_size[2] = w;
    o.launch_id_size[3] = h;
    o.num_lights = num_lights;
    o.num_frames = frame;
    o.accumulation = float(frame) / float(frame + 1);
    o.max_ray_bounces = st.max_ray_bounces;
    o.shadow_ray_bias = st.shadow_ray_bias;
    o.focal_length = cam.focal_length;
    o.aperture_radius = cam.aperture_radius;
}

// ===========================================================================
// Textures: Vulkan bilinear at LOD 0, REPEAT addressing (vk.cpp:3557-3576).
// Formats: UNORM8 -> byte/255; SRGB8 rgb decoded per texel before filtering.
// ===========================================================================
static double srgb_to_linear(double c) {
    return c <= 0.04045 ? c / 12.92 : std::pow((c + 0.055) / 1.055, 2.4);
}
static void decode_texture(Texture& t, const void* data) {
    size_t n = size_t(t.w) * t.h;
    t.texels.resize(n * 4);
    if (t.format == HH_TEXTURE_RGBA32F) {
        std::memcpy(t.texels.data(), data, n * 16);
        return;
    }
    const uint8_t* b = (const uint8_t*)data;
    float lut_lin[256], lut_srgb[256];
    for (int i = 0; i < 256; i++) {
        lut_lin[i] = (float)(i / 255.0);
        lut_srgb[i] = (float)srgb_to_linear(i / 255.0);
    }
    for (size_t i = 0; i < n; i++)
        for (int c = 0; c < 4; c++) {
            uint8_t v = b[i * 4 + c];
            t.texels[i * 4 + c] = (t.format == HH_TEXTURE_RGBA8_SRGB && c < 3) ? lut_srgb[v] : lut_lin[v];
        }
}
template <typename R>
static inline R lerp_(R a, R b, R w) { return a + (b - a) * w; }

template <typename R>
static void sample_bilinear_repeat(const Texture& t, R s, R u_t, R out[4]) {
    R x = s * R(t.w) - R(0.5), y = u_t * R(t.h) - R(0.5);
    R fx = std::floor(x), fy = std::floor(y);
    R ax = x - fx, ay = y - fy;
    int64_t i0 = (int64_t)fx, j0 = (int64_t)fy;
    auto wrap = [](int64_t i, int64_t n) { int64_t r = i % n; return r < 0 ? r + n : r; };
    int64_t i1 = wrap(i0 + 1, t.w), j1 = wrap(j0 + 1, t.h);
    i0 = wrap(i0, t.w);
    j0 = wrap(j0, t.h);
    const float* p00 = &t.texels[(size_t(j0) * t.w + i0) * 4];
    const float* p10 = &t.texels[(size_t(j0) * t.w + i1) * 4];
    const float* p01 = &t.texels[(size_t(j1) * t.w + i0) * 4];
    const float* p11 = &t.texels[(size_t(j1) * t.w + i1) * 4];
    for (int c = 0; c < 4; c++) {
        R top = lerp_(R(p00[c]), R(p10[c]), ax);
        R bot = lerp_(R(p01[c]), R(p11[c]), ax);
        out[c] = lerp_(top, bot, ay);
    }
}

// Cube sampling with Vulkan's face selection and seamless bilinear filtering
// (the reference's env sampler is LINEAR, vk.cpp:3557-3587; Vulkan cube
// lookups filter across face edges, "Cube Map Edge Handling"). Face texels
// are row-major, t = 0 at row 0.
//
// A bilinear tap (i, j) outside a face (i or j = -1 or n) is the texel of the
// adjacent face that holds that texel's centre: the centre is extended on the
// face's plane, (2i+1-n, 2j+1-n) over n, and the face is re-selected by the
// major axis of that point (all in integers, exact). A tap outside on both axes
// (a corner) is the mean of the three texels that meet there: this face's
// corner texel and its two neighbours across the edges (Vulkan leaves the
// corner implementation-defined).
static void cube_face_point(int face, int64_t a, int64_t b, int64_t m, int64_t P[3]) {
    switch (face) {  // inverse of the face selection below: sc = a/m, tc = b/m
        case 0: P[0] = m; P[1] = -b; P[2] = -a; break;
        case 1: P[0] = -m; P[1] = -b; P[2] = a; break;
        case 2: P[0] = a; P[1] = m; P[2] = b; break;
        case 3: P[0] = a; P[1] = -m; P[2] = -b; break;
        case 4: P[0] = a; P[1] = -b; P[2] = m; break;
        default: P[0] = -a; P[1] = -b; P[2] = -m; break;
    }
}
static int cube_select(int64_t x, int64_t y, int64_t z, int64_t& sc, int64_t& tc, int64_t& ma) {
    const int64_t ax = x < 0 ? -x : x, ay = y < 0 ? -y : y, az = z < 0 ? -z : z;
    if (ax >= ay && ax >= az) {
        ma = ax;
        if (x >= 0) { sc = -z; tc = -y; return 0; }
        sc = z; tc = -y; return 1;
    }
    if (ay >= az) {
        ma = ay;
        if (y >= 0) { sc = x; tc = z; return 2; }
        sc = x; tc = -z; return 3;
    }
    ma = az;
    if (z >= 0) { sc = x; tc = -y; return 4; }
    sc = -x; tc = -y; return 5;
}
// (face, i, j) with one of i, j just outside [0, n): the adjacent face's texel.
static void cube_wrap(int face, int64_t i, int64_t j, int64_t n, int& f2, int64_t& i2, int64_t& j2) {
    int64_t P[3], sc, tc, ma;
    cube_face_point(face, 2 * i + 1 - n, 2 * j + 1 - n, n, P);
    f2 = cube_select(P[0], P[1], P[2], sc, tc, ma);
    i2 = (n * (sc + ma)) / (2 * ma);
    j2 = (n * (tc + ma)) / (2 * ma);
    i2 = i2 < 0 ? 0 : (i2 >= n ? n - 1 : i2);
    j2 = j2 < 0 ? 0 : (j2 >= n ? n - 1 : j2);
}
template <typename R>
static void cube_texel(const std::vector<float>& faces, uint32_t size, int face, int64_t i, int64_t j, R out[3]) {
    const int64_t n = size;
    auto at = [&](int f, int64_t ii, int64_t jj, int c) {
        return R(faces[((size_t(f) * size + size_t(jj)) * size + size_t(ii)) * 4 + c]);
    };
    const bool oi = i < 0 || i >= n, oj = j < 0 || j >= n;
    if (!oi && !oj) {
        for (int c = 0; c < 3; c++) out[c] = at(face, i, j, c);
        return;
    }
    const int64_t ci = i < 0 ? 0 : (i >= n ? n - 1 : i), cj = j < 0 ? 0 : (j >= n ? n - 1 : j);
    int f1;
    int64_t i1, j1;
    if (oi && oj) {
        int f2;
        int64_t i2, j2;
        cube_wrap(face, i, cj, n, f1, i1, j1);
        cube_wrap(face, ci, j, n, f2, i2, j2);
        for (int c = 0; c < 3; c++) out[c] = ((at(face, ci, cj, c) + at(f1, i1, j1, c)) + at(f2, i2, j2, c)) / R(3);
        return;
    }
    cube_wrap(face, i, j, n, f1, i1, j1);
    for (int c = 0; c < 3; c++) out[c] = at(f1, i1, j1, c);
}
template <typename R>
static V3<R> sample_cube(const std::vector<float>& faces, uint32_t size, V3<R> d) {
    R ax = std::fabs(d.x), ay = std::fabs(d.y), az = std::fabs(d.z);
    int face;
    R sc, tc, ma;
    if (ax >= ay && ax >= az) {
        ma = ax;
        if (d.x >= R(0)) { face = 0; sc = -d.z; tc = -d.y; } else { face = 1; sc = d.z; tc = -d.y; }
    } else if (ay >= az) {
        ma = ay;
        if (d.y >= R(0)) { face = 2; sc = d.x; tc = d.z; } else { face = 3; sc = d.x; tc = -d.z; }
    } else {
        ma = az;
        if (d.z >= R(0)) { face = 4; sc = d.x; tc = -d.y; } else { face = 5; sc = -d.x; tc = -d.y; }
    }
    if (!(ma > R(0))) return V3<R>(0, 0, 0);
    R s = R(0.5) * (sc / ma + R(1)), t = R(0.5) * (tc / ma + R(1));
    R x = s * R(size) - R(0.5), y = t * R(size) - R(0.5);
    R fx = std::floor(x), fy = std::floor(y);
    R wx = x - fx, wy = y - fy;
    const int64_t i0 = (int64_t)fx, j0 = (int64_t)fy;
    R p00[3], p10[3], p01[3], p11[3];
    cube_texel<R>(faces, size, face, i0, j0, p00);
    cube_texel<R>(faces, size, face, i0 + 1, j0, p10);
    cube_texel<R>(faces, size, face, i0, j0 + 1, p01);
    cube_texel<R>(faces, size, face, i0 + 1, j0 + 1, p11);
    R out[3];
    for (int c = 0; c < 3; c++) {
        R top = lerp_(p00[c], p10[c], wx);
        R bot = lerp_(p01[c], p11[c], wx);
        out[c] = lerp_(top, bot, wy);
    }
    return V3<R>(out[0], out[1], out[2]);
}

// IEEE binary16 conversions (round to nearest even; finite inputs).
static uint16_t float_to_half(float f) {
    uint32_t x;
    std::memcpy(&x, &f, 4);
    const uint32_t sign = (x >> 16) & 0x8000u;
    const int32_t exp = (int32_t)((x >> 23) & 0xFF) - 127 + 15;
    uint32_t mant = x & 0x7FFFFFu;
    if (exp >= 31) return (uint16_t)(sign | 0x7C00u);               // overflow -> inf
    if (exp <= 0) {                                                  // subnormal or zero
        if (exp < -10) return (uint16_t)sign;
        mant |= 0x800000u;
        const int shift = 14 - exp;
        uint32_t h = mant >> shift;
        const uint32_t rem = mant & ((1u << shift) - 1u), half = 1u << (shift - 1);
        if (rem > half || (rem == half && (h & 1u))) h++;
        return (uint16_t)(sign | h);
    }
    uint32_t h = ((uint32_t)exp << 10) | (mant >> 13);
    const uint32_t rem = mant & 0x1FFFu;
    if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) h++;          // carries into the exponent correctly
    return (uint16_t)(sign | h);
}
static float half_to_float(uint16_t h) {
    const uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
    uint32_t exp = (h >> 10) & 0x1Fu, mant = h & 0x3FFu;
    uint32_t x;
    if (exp == 0) {
        if (mant == 0) x = sign;
        else {
            int e = -1;
            do { mant <<= 1; e++; } while (!(mant & 0x400u));
            x = sign | ((uint32_t)(127 - 15 - e) << 23) | ((mant & 0x3FFu) << 13);
        }
    } else if (exp == 31) x = sign | 0x7F800000u | (mant << 13);
    else x = sign | ((exp - 15 + 127) << 23) | (mant << 13);
    float f;
    std::memcpy(&f, &x, 4);
    return f;
}

// Extension: equirectangular lookup (bilinear, repeat in x, clamp in y).
template <typename R>
static V3<R> sample_equirect(const std::vector<uint16_t>& tex, uint32_t w, V3<R> d) {
    const int64_t W = w, H = w / 2;
    R theta = std::acos(gclamp(d.y, R(-1), R(1)));
    R phi = std::atan2(d.x, -d.z);
    R x = (phi / (R(2) * kPi<R>()) + R(0.5)) * R(W) - R(0.5);
    R y = (theta / kPi<R>()) * R(H) - R(0.5);
    R fx = std::floor(x), fy = std::floor(y);
    R wx = x - fx, wy = y - fy;
    int64_t i0 = (((int64_t)fx % W) + W) % W, i1 = (i0 + 1) % W;
    auto cl = [H](int64_t j) { return j < 0 ? 0 : (j >= H ? H - 1 : j); };
    int64_t j0 = cl((int64_t)fy), j1 = cl((int64_t)fy + 1);
    auto px = [&](int64_t i, int64_t j, int c) { return R(half_to_float(tex[((size_t)j * W + i) * 4 + c])); };
    R out[3];
    for (int c = 0; c < 3; c++) out[c] = lerp_(lerp_(px(i0, j0, c), px(i1, j0, c), wx), lerp_(px(i0, j1, c), px(i1, j1, c), wx), wy);
    return V3<R>(out[0], out[1], out[2]);
}

// ===========================================================================
// Hosek-Wilkie sky (hosek_wilkie_sky_model.cpp:41-75, 658-701;
// procedural_sky.frag:48-75). Dataset: helios_amd/data/hosek_rgb.bin.
// ===========================================================================
static std::vector<double> g_hosek;  // 3600 doubles
static const double* hosek_rgb(int ch) { return g_hosek.data() + ch * 1080; }
static const double* hosek_rad(int ch) { return g_hosek.data() + 3 * 1080 + ch * 120; }

static double evaluate_spline(const double* spline, size_t stride, double value) {
    return 1 * std::pow(1 - value, 5) * spline[0 * stride] + 5 * std::pow(1 - value, 4) * std::pow(value, 1) * spline[1 * stride] +
           10 * std::pow(1 - value, 3) * std::pow(value, 2) * spline[2 * stride] +
           10 * std::pow(1 - value, 2) * std::pow(value, 3) * spline[3 * stride] +
           5 * std::pow(1 - value, 1) * std::pow(value, 4) * spline[4 * stride] + 1 * std::pow(value, 5) * spline[5 * stride];
}
static double evaluate(const double* dataset, size_t stride, float turbidity, float albedo, float sunTheta) {
    // std::max<float> in the reference rounds the argument to float, and the
    // float overload of pow is taken.
    float e = std::max<float>(0.f, (float)(1.f - sunTheta / (3.14159265358979323846 / 2.f)));
    double elevationK = std::pow(e, 1.f / 3.0f);
    int turbidity0 = std::min(std::max((int)turbidity, 1), 10);
    int turbidity1 = std::min(turbidity0 + 1, 10);
    float turbidityK = std::min(std::max(turbidity - turbidity0, 0.f), 1.f);
    const double* a0 = dataset;
    const double* a1 = dataset + stride * 6 * 10;
    double a0t0 = evaluate_spline(a0 + stride * 6 * (turbidity0 - 1), stride, elevationK);
    double a1t0 = evaluate_spline(a1 + stride * 6 * (turbidity0 - 1), stride, elevationK);
    double a0t1 = evaluate_spline(a0 + stride * 6 * (turbidity1 - 1), stride, elevationK);
    double a1t1 = evaluate_spline(a1 + stride * 6 * (turbidity1 - 1), stride, elevationK);
    return a0t0 * (1 - albedo) * (1 - turbidityK) + a1t0 * albedo * (1 - turbidityK) + a0t1 * (1 - albedo) * turbidityK +
           a1t1 * albedo * turbidityK;
}
// hosek_wilkie_sky_model.cpp:69-73 (host float version)
static void hosek_wilkie_host(float cos_theta, float gamma, float cos_gamma, const float* c /*A..I as 9x3*/, float out[3]) {
    for (int i = 0; i < 3; i++) {
        float A = c[0 * 3 + i], B = c[1 * 3 + i], C = c[2 * 3 + i], D = c[3 * 3 + i], E = c[4 * 3 + i], F = c[5 * 3 + i],
              G = c[6 * 3 + i], H = c[7 * 3 + i], I = c[8 * 3 + i];
        float chi = (1.f + cos_gamma * cos_gamma) / std::pow(1.f + H * H - 2.f * cos_gamma * H, 1.5f);
        out[i] = (1.f + A * std::exp(B / (cos_theta + 0.01f))) *
                 (C + D * std::exp(E * gamma) + F * (cos_gamma * cos_gamma) + G * chi + I * (float)std::sqrt(std::max(0.f, cos_theta)));
    }
}
static void sky_coefficients(const hh_sky_params& sp, const float sun_dir[3], float out[30]) {
    const float sunTheta = std::acos(std::min(std::max(sun_dir[1], 0.f), 1.f));
    float* A = out;  // out[k*3+i] = coefficient k (A..I, Z) channel i
    for (int i = 0; i < 3; i++) {
        const double* d = hosek_rgb(i);
        for (int k = 0; k < 7; k++) A[k * 3 + i] = (float)evaluate(d + k, 9, sp.turbidity, sp.albedo, sunTheta);
        A[7 * 3 + i] = (float)evaluate(d + 8, 9, sp.turbidity, sp.albedo, sunTheta);  // H, swapped in the dataset
        A[8 * 3 + i] = (float)evaluate(d + 7, 9, sp.turbidity, sp.albedo, sunTheta);  // I
        A[9 * 3 + i] = (float)evaluate(hosek_rad(i), 1, sp.turbidity, sp.albedo, sunTheta);
    }
    if (sp.normalized_sun_y != 0.f) {
        float S[3];
        hosek_wilkie_host(std::cos(sunTheta), 0, 1.f, out, S);
        for (int i = 0; i < 3; i++) S[i] *= out[27 + i];
        float lum = S[0] * 0.2126f + S[1] * 0.7152f + S[2] * 0.0722f;
        for (int i = 0; i < 3; i++) out[27 + i] = out[27 + i] / lum * sp.normalized_sun_y;
    }
}
// procedural_sky.frag:48-64, evaluated in precision R.
template <typename R>
static V3<R> hosek_sky_rgb(V3<R> v, V3<R> sun, const float* c) {
    R cos_theta = gclamp(v.y, R(0), R(1));
    R cos_gamma = gclamp(dot(v, sun), R(0), R(1));
    R gamma = std::acos(cos_gamma);
    R o[3];
    for (int i = 0; i < 3; i++) {
        R A = c[0 * 3 + i], B = c[1 * 3 + i], C = c[2 * 3 + i], D = c[3 * 3 + i], E = c[4 * 3 + i], F = c[5 * 3 + i],
          G = c[6 * 3 + i], H = c[7 * 3 + i], I = c[8 * 3 + i], Z = c[9 * 3 + i];
        R chi = (R(1) + cos_gamma * cos_gamma) / std::pow(R(1) + H * H - R(2) * cos_gamma * H, R(1.5));
        R hw = (R(1) + A * std::exp(B / (cos_theta + R(0.01)))) *
               (C + D * std::exp(E * gamma) + F * (cos_gamma * cos_gamma) + G * chi + I * std::sqrt(cos_theta));
        o[i] = Z * hw;
    }
    return V3<R>(o[0], o[1], o[2]);
}
// Texel-centre direction of cube face `face` (inverse of the Vulkan face table).
template <typename R>
static V3<R> cube_texel_dir(int face, uint32_t i, uint32_t j, uint32_t n) {
    R sc = R(2) * ((R(i) + R(0.5)) / R(n)) - R(1);
    R tc = R(2) * ((R(j) + R(0.5)) / R(n)) - R(1);
    V3<R> d;
    switch (face) {
        case 0: d = V3<R>(R(1), -tc, -sc); break;
        case 1: d = V3<R>(R(-1), -tc, sc); break;
        case 2: d = V3<R>(sc, R(1), tc); break;
        case 3: d = V3<R>(sc, R(-1), -tc); break;
        case 4: d = V3<R>(sc, -tc, R(1)); break;
        default: d = V3<R>(-sc, -tc, R(-1)); break;
    }
    return normalize(d);
}

// ===========================================================================
// Packing (Scene::create_gpu_resources, scene.cpp:915-1311).
// ===========================================================================
static bool material_is_emissive(const hh_material_desc& m) {  // material.cpp:71-77
    if (m.emissive_texture != -1) return true;
    return m.emissive_value[0] > 0.0f || m.emissive_value[1] > 0.0f || m.emissive_value[2] > 0.0f;
}
static void invert_affine(const float model[16], float w2o[12]) {
    M4 m{};
    for (int c = 0; c < 4; c++)
        for (int r = 0; r < 4; r++) m.m[c][r] = model[c * 4 + r];
    M4 inv = m4_inverse(m);
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 4; c++) w2o[r * 4 + c] = (float)inv.m[c][r];
}

static void build_blas(Mesh& me, const std::vector<hh_material_desc>& mats);

void commit(Scene& s) {
    s.mat_data.clear();
    s.tex_slots.clear();
    s.light_data.clear();
    s.inst_data.clear();
    s.submesh_info.clear();
    std::vector<int> mesh_seen(s.meshes.size(), -1);
    std::vector<int> mat_global(s.materials.size(), -1);
    std::vector<char> tex_seen(s.textures.size(), 0);
    uint32_t mesh_counter = 0;
    std::vector<LightData> area;
    auto pack_tex = [&](int32_t tex_id, int32_t& slot) {
        // First use gets a descriptor slot; a texture already used by another
        // material leaves the slot at -1 (scene.cpp:981-994 and siblings).
        if (!tex_seen[tex_id]) {
            tex_seen[tex_id] = 1;
            slot = (int32_t)s.tex_slots.size();
            s.tex_slots.push_back(tex_id);
        }
    };
    for (size_t ni = 0; ni < s.instances.size(); ni++) {
        const hh_instance_desc& inst = s.instances[ni];
        Mesh& me = s.meshes[inst.mesh];
        auto material_of = [&](const hh_submesh& sm) -> uint32_t {
            if (inst.material_override >= 0) return (uint32_t)inst.material_override;
            return me.materials[sm.mat_idx];
        };
        if (mesh_seen[inst.mesh] < 0) {
            mesh_seen[inst.mesh] = (int)mesh_counter++;
            for (const hh_submesh& sm : me.submeshes) {
                uint32_t mid = material_of(sm);
                const hh_material_desc& m = s.materials[mid];
                if (mat_global[mid] < 0) {
                    MaterialData md;
                    for (int k = 0; k < 4; k++) { md.tex0[k] = -1; md.tex1[k] = -1; md.albedo[k] = 0; md.emissive[k] = 0; md.roughness_metallic[k] = 0; }
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
                    s.mat_data.push_back(md);
                    mat_global[mid] = (int)s.mat_data.size() - 1;
                }
                if (material_is_emissive(m)) {  // scene.cpp:1111-1119
                    LightData l{};
                    l.d0[0] = 4.0f;  // LIGHT_AREA
                    l.d0[1] = float(ni);
                    l.d0[2] = float(mat_global[mid]);
                    l.d0[3] = float(sm.base_index / 3);
                    l.d1[0] = float(sm.index_count / 3);
                    area.push_back(l);
                }
            }
        }
        std::vector<uint32_t> info;
        for (const hh_submesh& sm : me.submeshes) {
            info.push_back(sm.base_index / 3);
            // scene.cpp:934,1143: global_material_indices is an unordered_map, so
            // a material that was never packed (its mesh was first packed under
            // another node's override) reads as material 0 (Q19)
            const int gm = mat_global[material_of(sm)];
            info.push_back(gm < 0 ? 0u : (uint32_t)gm);
        }
        s.submesh_info.push_back(info);
        InstanceData id{};
        std::memcpy(id.model, inst.model, 64);
        std::memcpy(id.normal, inst.normal, 64);
        id.mesh_index = (uint32_t)mesh_seen[inst.mesh];
        invert_affine(inst.model, id.w2o);
        s.inst_data.push_back(id);
    }
    // Light buffer order: area, env, directional, point, spot (scene.cpp:1259-1309).
    s.light_data = area;
    uint32_t n_dir = 0, n_point = 0, n_spot = 0;
    for (auto& l : s.lights) {
        n_dir += l.type == HH_LIGHT_DIRECTIONAL;
        n_point += l.type == HH_LIGHT_POINT;
        n_spot += l.type == HH_LIGHT_SPOT;
    }
    bool has_ibl = s.ibl_size > 0;
    if (has_ibl || n_dir > 0) {
        LightData l{};
        l.d0[0] = 3.0f;
        s.light_data.push_back(l);
    }
    for (auto& l : s.lights)
        if (l.type == HH_LIGHT_DIRECTIONAL) {
            LightData d{};
            d.d0[0] = 0.0f;
            for (int k = 0; k < 3; k++) d.d0[1 + k] = l.color[k];
            for (int k = 0; k < 3; k++) d.d1[k] = l.forward[k];
            d.d1[3] = l.intensity;
            d.d2[3] = l.radius;
            s.light_data.push_back(d);
        }
    for (auto& l : s.lights)
        if (l.type == HH_LIGHT_POINT) {
            LightData d{};
            d.d0[0] = 2.0f;
            for (int k = 0; k < 3; k++) d.d0[1 + k] = l.color[k];
            d.d1[3] = l.intensity;
            for (int k = 0; k < 3; k++) d.d2[k] = l.position[k];
            d.d2[3] = l.radius;
            s.light_data.push_back(d);
        }
    for (auto& l : s.lights)
        if (l.type == HH_LIGHT_SPOT) {
            LightData d{};
            d.d0[0] = 1.0f;
            for (int k = 0; k < 3; k++) d.d0[1 + k] = l.color[k];
            for (int k = 0; k < 3; k++) d.d1[k] = l.forward[k];
            d.d1[3] = l.intensity;
            for (int k = 0; k < 3; k++) d.d2[k] = l.position[k];
            d.d2[3] = l.radius;
            d.d3[0] = std::cos(l.inner_cone_deg * (3.14159265358979323846f / 180.0f));
            d.d3[1] = std::cos(l.outer_cone_deg * (3.14159265358979323846f / 180.0f));
            s.light_data.push_back(d);
        }
    // num_lights (scene.cpp:891-899)
    s.num_lights = (uint32_t)area.size() + n_dir + n_spot + n_point + ((has_ibl || n_dir > 0) ? 1u : 0u);
    s.num_area = (uint32_t)area.size();
    // Extension (HH_FLAG_LIGHT_POWER): every emissive triangle of every
    // instance, world vertices in double, P(tri) = luminance(emission) x area
    // / total, inclusive CDF (last entry exactly 1).
    s.emit_v.clear(); s.emit_area.clear(); s.emit_p.clear(); s.emit_cdf.clear(); s.emit_le.clear();
    {
        std::vector<double> power;
        for (const hh_instance_desc& inst : s.instances) {
            const Mesh& me = s.meshes[inst.mesh];
            for (const hh_submesh& sm : me.submeshes) {
                const uint32_t mid = inst.material_override >= 0 ? (uint32_t)inst.material_override : me.materials[sm.mat_idx];
                const hh_material_desc& m = s.materials[mid];
                const double le[3] = {m.emissive_value[0], m.emissive_value[1], m.emissive_value[2]};
                const double lum = 0.2126 * le[0] + 0.7152 * le[1] + 0.0722 * le[2];
                if (!(lum > 0.0)) continue;
                for (uint32_t t = 0; t < sm.index_count / 3; t++) {
                    double w[3][3], ns[3] = {0.0, 0.0, 0.0};
                    for (int k = 0; k < 3; k++) {
                        const hh_vertex& vx = me.vertices[me.indices[sm.base_index + 3 * t + k]];
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
                    const double a = 0.5 * std::sqrt(cx * cx + cy * cy + cz * cz);
                    if (!(a > 0.0)) continue;
                    for (int k = 0; k < 3; k++)
                        for (int r = 0; r < 3; r++) s.emit_v.push_back((float)w[k][r]);
                    s.emit_area.push_back((float)a);
                    for (int c = 0; c < 3; c++) s.emit_le.push_back((float)le[c]);
                    power.push_back(lum * a);
                }
            }
        }
        double total = 0.0, acc = 0.0;
        for (double p : power) total += p;
        for (size_t i = 0; i < power.size(); i++) {
            acc += power[i];
            s.emit_p.push_back((float)(power[i] / total));
            s.emit_cdf.push_back(i + 1 == power.size() ? 1.0f : (float)(acc / total));
        }
    }
    // Environment: IBL, else the sky when a directional light exists, else the
    // 2x2 black default cube (scene.cpp:1149-1161; vk.cpp:3589-3610).
    if (has_ibl) {
        s.env = s.ibl;
        s.env_size = s.ibl_size;
    } else if (n_dir > 0 && s.sky.equirect_width > 0) {
        // extension: equirectangular RGBA16F sky, texel centres as the GPU bake
        const hh_light_desc* sun = nullptr;
        for (auto& l : s.lights)
            if (l.type == HH_LIGHT_DIRECTIONAL) { sun = &l; break; }
        float sd[3] = {-sun->forward[0], -sun->forward[1], -sun->forward[2]};
        sky_coefficients(s.sky, sd, s.sky_coef);
        const uint32_t w = s.sky.equirect_width, h = w / 2;
        s.env_eq_w = w;
        s.env_eq.assign(size_t(w) * h * 4, 0);
        V3<float> sunv(sd[0], sd[1], sd[2]);
        for (uint32_t y = 0; y < h; y++)
            for (uint32_t x = 0; x < w; x++) {
                const float theta = kPi<float>() * (((float)y + 0.5f) / (float)h);
                const float phi = 2.0f * kPi<float>() * (((float)x + 0.5f) / (float)w - 0.5f);
                V3<float> d(std::sin(theta) * std::sin(phi), std::cos(theta), -std::sin(theta) * std::cos(phi));
                V3<float> c = hosek_sky_rgb<float>(d, sunv, s.sky_coef);
                uint16_t* px = &s.env_eq[(size_t(y) * w + x) * 4];
                px[0] = float_to_half(c.x); px[1] = float_to_half(c.y); px[2] = float_to_half(c.z); px[3] = float_to_half(1.0f);
            }
        s.env_size = 2;
        s.env.assign(6 * 2 * 2 * 4, 0.f);
    } else if (n_dir > 0) {
        const hh_light_desc* sun = nullptr;
        for (auto& l : s.lights)
            if (l.type == HH_LIGHT_DIRECTIONAL) { sun = &l; break; }
        float sd[3] = {-sun->forward[0], -sun->forward[1], -sun->forward[2]};
        sky_coefficients(s.sky, sd, s.sky_coef);
        uint32_t n = s.sky.cube_size;
        s.env_size = n;
        s.env.assign(size_t(6) * n * n * 4, 0.f);
        V3<float> sunv(sd[0], sd[1], sd[2]);
        for (int f = 0; f < 6; f++)
            for (uint32_t j = 0; j < n; j++)
                for (uint32_t i = 0; i < n; i++) {
                    V3<float> d = cube_texel_dir<float>(f, i, j, n);
                    V3<float> c = hosek_sky_rgb<float>(d, sunv, s.sky_coef);
                    float* px = &s.env[((size_t(f) * n + j) * n + i) * 4];
                    px[0] = c.x; px[1] = c.y; px[2] = c.z; px[3] = 1.0f;
                }
    } else {
        s.env_size = 2;
        s.env.assign(6 * 2 * 2 * 4, 0.f);
    }
    for (size_t mi = 0; mi < s.meshes.size(); mi++) build_blas(s.meshes[mi], s.materials);
    // World AABB of every instance (conservative: corners in double, rounded out).
    s.inst_world_lo.assign(s.instances.size() * 3, 0.f);
    s.inst_world_hi.assign(s.instances.size() * 3, 0.f);
    for (size_t ni = 0; ni < s.instances.size(); ni++) {
        const Mesh& me = s.meshes[s.instances[ni].mesh];
        double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
        if (!me.nodes.empty()) {
            const auto& root = me.nodes[0];
            for (int c = 0; c < 8; c++) {
                double p[3] = {(c & 1) ? root.hi[0] : root.lo[0], (c & 2) ? root.hi[1] : root.lo[1], (c & 4) ? root.hi[2] : root.lo[2]};
                const float* M = s.instances[ni].model;
                for (int r = 0; r < 3; r++) {
                    double v = M[0 * 4 + r] * p[0] + M[1 * 4 + r] * p[1] + M[2 * 4 + r] * p[2] + M[3 * 4 + r];
                    lo[r] = std::min(lo[r], v);
                    hi[r] = std::max(hi[r], v);
                }
            }
        }
        for (int r = 0; r < 3; r++) {
            float l = (float)lo[r], h = (float)hi[r];
            s.inst_world_lo[ni * 3 + r] = std::nextafter(l, -INFINITY);
            s.inst_world_hi[ni * 3 + r] = std::nextafter(h, INFINITY);
        }
    }
    s.committed = true;
}

// CPU BVH2 per mesh (median split on the widest centroid axis, leaves <= 4).
static void build_blas(Mesh& me, const std::vector<hh_material_desc>& mats) {
    me.tris.clear();
    me.nodes.clear();
    me.geom_opaque.clear();
    for (uint32_t g = 0; g < me.submeshes.size(); g++) {
        const hh_submesh& sm = me.submeshes[g];
        const hh_material_desc& m = mats[me.materials[sm.mat_idx]];
        me.geom_opaque.push_back(m.type == HH_MATERIAL_OPAQUE && !m.alpha_test);
        for (uint32_t p = 0; p < sm.index_count / 3; p++) {
            Mesh::Tri t;
            for (int k = 0; k < 3; k++) {
                const hh_vertex& v = me.vertices[me.indices[sm.base_index + 3 * p + k]];
                for (int c = 0; c < 3; c++) t.v[k][c] = v.position[c];
            }
            t.geom = g;
            t.prim = p;
            me.tris.push_back(t);
        }
    }
    if (me.tris.empty()) return;
    std::vector<uint32_t> order(me.tris.size());
    for (uint32_t i = 0; i < order.size(); i++) order[i] = i;
    std::vector<float> cen(me.tris.size() * 3);
    for (size_t i = 0; i < me.tris.size(); i++)
        for (int c = 0; c < 3; c++) cen[i * 3 + c] = (me.tris[i].v[0][c] + me.tris[i].v[1][c] + me.tris[i].v[2][c]) / 3.0f;
    struct Job { int32_t node; uint32_t b, e; };
    std::vector<Job> stack;
    me.nodes.push_back({});
    stack.push_back({0, 0, (uint32_t)order.size()});
    while (!stack.empty()) {
        Job j = stack.back();
        stack.pop_back();
        Mesh::Node nd;
        for (int c = 0; c < 3; c++) { nd.lo[c] = INFINITY; nd.hi[c] = -INFINITY; }
        float clo[3] = {INFINITY, INFINITY, INFINITY}, chi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (uint32_t i = j.b; i < j.e; i++) {
            const Mesh::Tri& t = me.tris[order[i]];
            for (int k = 0; k < 3; k++)
                for (int c = 0; c < 3; c++) { nd.lo[c] = std::min(nd.lo[c], t.v[k][c]); nd.hi[c] = std::max(nd.hi[c], t.v[k][c]); }
            for (int c = 0; c < 3; c++) { clo[c] = std::min(clo[c], cen[order[i] * 3 + c]); chi[c] = std::max(chi[c], cen[order[i] * 3 + c]); }
        }
        if (j.e - j.b <= 4) {
            nd.left = nd.right = -1;
            nd.start = j.b;
            nd.count = j.e - j.b;
            me.nodes[j.node] = nd;
            continue;
        }
        int ax = 0;
        for (int c = 1; c < 3; c++)
            if (chi[c] - clo[c] > chi[ax] - clo[ax]) ax = c;
        uint32_t mid = (j.b + j.e) / 2;
        std::nth_element(order.begin() + j.b, order.begin() + mid, order.begin() + j.e,
                         [&](uint32_t a, uint32_t b) { return cen[a * 3 + ax] < cen[b * 3 + ax]; });
        nd.left = (int32_t)me.nodes.size();
        nd.right = nd.left + 1;
        nd.start = nd.count = 0;
        me.nodes[j.node] = nd;
        me.nodes.push_back({});
        me.nodes.push_back({});
        stack.push_back({nd.left, j.b, mid});
        stack.push_back({nd.right, mid, j.e});
    }
    std::vector<Mesh::Tri> sorted(me.tris.size());
    for (size_t i = 0; i < order.size(); i++) sorted[i] = me.tris[order[i]];
    me.tris.swap(sorted);
}

// ===========================================================================
// Traversal: watertight ray/triangle test (Woop, Benthin, Wald 2013) with the
// reference's closest-hit / any-hit / terminate-on-first-hit semantics.
// Ties on t are broken by the smaller (instance, geometry, primitive).
// ===========================================================================
template <typename R>
struct RayPre {
    int kx, ky, kz;
    R Sx, Sy, Sz;
};
template <typename R>
static RayPre<R> ray_pre(V3<R> d) {
    RayPre<R> p;
    R ax = std::fabs(d.x), ay = std::fabs(d.y), az = std::fabs(d.z);
    p.kz = ax > ay ? (ax > az ? 0 : 2) : (ay > az ? 1 : 2);
    p.kx = p.kz + 1;
    if (p.kx == 3) p.kx = 0;
    p.ky = p.kx + 1;
    if (p.ky == 3) p.ky = 0;
    if (d[p.kz] < R(0)) std::swap(p.kx, p.ky);
    p.Sx = d[p.kx] / d[p.kz];
    p.Sy = d[p.ky] / d[p.kz];
    p.Sz = R(1) / d[p.kz];
    return p;
}
template <typename R>
static bool tri_hit(const RayPre<R>& pr, V3<R> o, const float v[3][3], R tmin, R tmax, R& t, R& u, R& vv) {
    V3<R> A = V3<R>(R(v[0][0]), R(v[0][1]), R(v[0][2])) - o;
    V3<R> B = V3<R>(R(v[1][0]), R(v[1][1]), R(v[1][2])) - o;
    V3<R> C = V3<R>(R(v[2][0]), R(v[2][1]), R(v[2][2])) - o;
    R Ax = A[pr.kx] - pr.Sx * A[pr.kz], Ay = A[pr.ky] - pr.Sy * A[pr.kz];
    R Bx = B[pr.kx] - pr.Sx * B[pr.kz], By = B[pr.ky] - pr.Sy * B[pr.kz];
    R Cx = C[pr.kx] - pr.Sx * C[pr.kz], Cy = C[pr.ky] - pr.Sy * C[pr.kz];
    R U = Cx * By - Cy * Bx;
    R V = Ax * Cy - Ay * Cx;
    R W = Bx * Ay - By * Ax;
    if (sizeof(R) == 4 && (U == R(0) || V == R(0) || W == R(0))) {
        double CxBy = (double)Cx * (double)By, CyBx = (double)Cy * (double)Bx;
        U = (R)(CxBy - CyBx);
        double AxCy = (double)Ax * (double)Cy, AyCx = (double)Ay * (double)Cx;
        V = (R)(AxCy - AyCx);
        double BxAy = (double)Bx * (double)Ay, ByAx = (double)By * (double)Ax;
        W = (R)(BxAy - ByAx);
    }
    if ((U < R(0) || V < R(0) || W < R(0)) && (U > R(0) || V > R(0) || W > R(0))) return false;
    R det = U + V + W;
    if (det == R(0)) return false;
    R Az = pr.Sz * A[pr.kz], Bz = pr.Sz * B[pr.kz], Cz = pr.Sz * C[pr.kz];
    R T = U * Az + V * Bz + W * Cz;
    t = T / det;
    if (!(t >= tmin && t <= tmax)) return false;
    u = V / det;
    vv = W / det;
    return true;
}

// World -> object transform with the fixed operation order the GPU uses.
template <typename R>
static void xform_ray(const float w2o[12], V3<R> o, V3<R> d, V3<R>& oo, V3<R>& od) {
    R r[3], s[3];
    for (int i = 0; i < 3; i++) {
        const float* m = w2o + i * 4;
        r[i] = ((R(m[0]) * o.x + R(m[1]) * o.y) + R(m[2]) * o.z) + R(m[3]);
        s[i] = (R(m[0]) * d.x + R(m[1]) * d.y) + R(m[2]) * d.z;
    }
    oo = V3<R>(r[0], r[1], r[2]);
    od = V3<R>(s[0], s[1], s[2]);
}

// Robust slab test (Ize 2013: tfar scaled by 1 + 2*gamma(3)).
template <typename R>
static bool box_hit(const float lo[3], const float hi[3], V3<R> o, V3<R> inv, R tmin, R tmax) {
    R t0x = (R(lo[0]) - o.x) * inv.x, t1x = (R(hi[0]) - o.x) * inv.x;
    R t0y = (R(lo[1]) - o.y) * inv.y, t1y = (R(hi[1]) - o.y) * inv.y;
    R t0z = (R(lo[2]) - o.z) * inv.z, t1z = (R(hi[2]) - o.z) * inv.z;
    R tn = gmax(gmax(gmin(t0x, t1x), gmin(t0y, t1y)), gmax(gmin(t0z, t1z), tmin));
    R tf = gmin(gmin(gmax(t0x, t1x), gmax(t0y, t1y)), gmin(gmax(t0z, t1z), tmax));
    return tn <= tf * R(1.0000004);
}

template <typename R>
struct Hit {
    R t, u, v;
    uint32_t inst, geom, prim;
    bool found;
};
static inline bool key_less(uint32_t i1, uint32_t g1, uint32_t p1, uint32_t i2, uint32_t g2, uint32_t p2) {
    if (i1 != i2) return i1 < i2;
    if (g1 != g2) return g1 < g2;
    return p1 < p2;
}

template <typename R>
static R fetch_alpha(const Scene& s, uint32_t inst, uint32_t geom, uint32_t prim, R u, R v);

// flags: HH_TRACE_OPAQUE; closest = false -> stop at the first accepted hit.
template <typename R>
static Hit<R> trace(const Scene& s, V3<R> o, V3<R> d, R tmin, R tmax, uint32_t flags, bool closest, bool brute) {
    Hit<R> h;
    h.found = false;
    h.t = tmax;
    h.u = h.v = 0;
    h.inst = h.geom = h.prim = 0xFFFFFFFFu;
    V3<R> winv(R(1) / d.x, R(1) / d.y, R(1) / d.z);
    for (uint32_t ni = 0; ni < s.inst_data.size(); ni++) {
        if (!brute && !box_hit(&s.inst_world_lo[ni * 3], &s.inst_world_hi[ni * 3], o, winv, tmin, h.t)) continue;
        const InstanceData& id = s.inst_data[ni];
        const Mesh& me = s.meshes[s.instances[ni].mesh];
        if (me.tris.empty()) continue;
        V3<R> oo, od;
        xform_ray(id.w2o, o, d, oo, od);
        RayPre<R> pr = ray_pre(od);
        V3<R> inv(R(1) / od.x, R(1) / od.y, R(1) / od.z);
        auto test_tri = [&](const Mesh::Tri& tri) -> bool {
            R t, u, v;
            if (!tri_hit(pr, oo, tri.v, tmin, h.t, t, u, v)) return false;
            if (h.found && t == h.t && !key_less(ni, tri.geom, tri.prim, h.inst, h.geom, h.prim)) return false;
            if (!(flags & HH_TRACE_OPAQUE) && !me.geom_opaque[tri.geom]) {
                if (fetch_alpha<R>(s, ni, tri.geom, tri.prim, u, v) < R(0.1)) return false;  // ignoreIntersectionEXT
            }
            h.found = true;
            h.t = t;
            h.u = u;
            h.v = v;
            h.inst = ni;
            h.geom = tri.geom;
            h.prim = tri.prim;
            return !closest;
        };
        if (brute) {
            for (const auto& tri : me.tris)
                if (test_tri(tri)) return h;
            continue;
        }
        int32_t stack[128];
        int sp = 0;
        stack[sp++] = 0;
        while (sp > 0) {
            const Mesh::Node& nd = me.nodes[stack[--sp]];
            if (!box_hit(nd.lo, nd.hi, oo, inv, tmin, h.t)) continue;
            if (nd.left < 0) {
                for (uint32_t i = 0; i < nd.count; i++)
                    if (test_tri(me.tris[nd.start + i])) return h;
            } else {
                stack[sp++] = nd.right;
                stack[sp++] = nd.left;
            }
        }
    }
    return h;
}

// ===========================================================================
// Shading (path_trace_rchit.glsl, brdf.glsl, sampling.glsl).
// ===========================================================================
template <typename R>
struct Vert {
    V3<R> pos, normal, tangent, bitangent;
    R uv[2];
};
template <typename R>
struct Surface {  // common.glsl:63-74
    Vert<R> vertex;
    R albedo[4];
    V3<R> emissive, normal, F0;
    R metallic, roughness, alpha, alpha2;
};

template <typename R>
static void sample_texture(const Scene& s, int32_t slot, R u, R v, R out[4]) {
    const Texture& t = s.textures[s.tex_slots[slot]];
    sample_bilinear_repeat<R>(t, u, v, out);
}

// path_trace_rahit.glsl:162-188
template <typename R>
static R fetch_alpha(const Scene& s, uint32_t inst, uint32_t geom, uint32_t prim, R u, R v) {
    const uint32_t* info = &s.submesh_info[inst][geom * 2];
    const MaterialData& m = s.mat_data[info[1]];
    const Mesh& me = s.meshes[s.instances[inst].mesh];
    uint32_t p = prim + info[0];
    const hh_vertex& v0 = me.vertices[me.indices[3 * p]];
    const hh_vertex& v1 = me.vertices[me.indices[3 * p + 1]];
    const hh_vertex& v2 = me.vertices[me.indices[3 * p + 2]];
    R bx = (R(1) - u) - v, by = u, bz = v;
    R tu = (R(v0.tex_coord[0]) * bx + R(v1.tex_coord[0]) * by) + R(v2.tex_coord[0]) * bz;
    R tv = (R(v0.tex_coord[1]) * bx + R(v1.tex_coord[1]) * by) + R(v2.tex_coord[1]) * bz;
    if (m.tex0[0] == -1) return R(m.albedo[3]);
    R c[4];
    sample_texture<R>(s, m.tex0[0], tu, tv, c);
    return c[3];
}

template <typename R>
static V3<R> mat3_mul(const float* m4 /*col-major 4x4*/, V3<R> v) {
    return V3<R>(R(m4[0]) * v.x + R(m4[4]) * v.y + R(m4[8]) * v.z, R(m4[1]) * v.x + R(m4[5]) * v.y + R(m4[9]) * v.z,
                 R(m4[2]) * v.x + R(m4[6]) * v.y + R(m4[10]) * v.z);
}
template <typename R>
static V3<R> mat4_mul_point(const float* m4, V3<R> v, R w) {
    return V3<R>(R(m4[0]) * v.x + R(m4[4]) * v.y + R(m4[8]) * v.z + R(m4[12]) * w,
                 R(m4[1]) * v.x + R(m4[5]) * v.y + R(m4[9]) * v.z + R(m4[13]) * w,
                 R(m4[2]) * v.x + R(m4[6]) * v.y + R(m4[10]) * v.z + R(m4[14]) * w);
}
template <typename R>
static V3<R> v4xyz(const float* p) { return V3<R>(R(p[0]), R(p[1]), R(p[2])); }

// populate_surface_properties (path_trace_rchit.glsl:256-280)
template <typename R>
static void populate(const Scene& s, const Hit<R>& h, Surface<R>& p) {
    const InstanceData& inst = s.inst_data[h.inst];
    const uint32_t* info = &s.submesh_info[h.inst][h.geom * 2];
    const Mesh& me = s.meshes[s.instances[h.inst].mesh];
    const MaterialData& m = s.mat_data[info[1]];
    uint32_t pid = h.prim + info[0];
    const hh_vertex* tv[3] = {&me.vertices[me.indices[3 * pid]], &me.vertices[me.indices[3 * pid + 1]],
                              &me.vertices[me.indices[3 * pid + 2]]};
    R b[3] = {R(1) - h.u - h.v, h.u, h.v};
    // interpolated_vertex (common.glsl:130-141)
    V3<R> pos = v4xyz<R>(tv[0]->position) * b[0] + v4xyz<R>(tv[1]->position) * b[1] + v4xyz<R>(tv[2]->position) * b[2];
    R uv0 = R(tv[0]->tex_coord[0]) * b[0] + R(tv[1]->tex_coord[0]) * b[1] + R(tv[2]->tex_coord[0]) * b[2];
    R uv1 = R(tv[0]->tex_coord[1]) * b[0] + R(tv[1]->tex_coord[1]) * b[1] + R(tv[2]->tex_coord[1]) * b[2];
    V3<R> n = normalize(v4xyz<R>(tv[0]->normal) * b[0] + v4xyz<R>(tv[1]->normal) * b[1] + v4xyz<R>(tv[2]->normal) * b[2]);
    V3<R> t = normalize(v4xyz<R>(tv[0]->tangent) * b[0] + v4xyz<R>(tv[1]->tangent) * b[1] + v4xyz<R>(tv[2]->tangent) * b[2]);
    V3<R> bt = normalize(v4xyz<R>(tv[0]->bitangent) * b[0] + v4xyz<R>(tv[1]->bitangent) * b[1] + v4xyz<R>(tv[2]->bitangent) * b[2]);
    // transform_vertex (path_trace_rchit.glsl:177-186)
    p.vertex.pos = mat4_mul_point<R>(inst.model, pos, R(1));
    p.vertex.normal = mat3_mul<R>(inst.normal, n);
    p.vertex.tangent = mat3_mul<R>(inst.normal, t);
    p.vertex.bitangent = mat3_mul<R>(inst.normal, bt);
    p.vertex.uv[0] = uv0;
    p.vertex.uv[1] = uv1;
    // fetch_albedo / normal / roughness / metallic / emissive (:206-252)
    if (m.tex0[0] == -1) for (int k = 0; k < 4; k++) p.albedo[k] = R(m.albedo[k]);
    else sample_texture<R>(s, m.tex0[0], uv0, uv1, p.albedo);
    if (m.tex0[1] == -1) p.normal = p.vertex.normal;
    else {
        R c[4];
        sample_texture<R>(s, m.tex0[1], uv0, uv1, c);
        V3<R> T = normalize(p.vertex.tangent), Bt = normalize(p.vertex.bitangent), N = normalize(p.vertex.normal);
        V3<R> tn = normalize(V3<R>(c[0] * R(2) - R(1), c[1] * R(2) - R(1), c[2] * R(2) - R(1)));
        p.normal = normalize(T * tn.x + Bt * tn.y + N * tn.z);
    }
    if (m.tex0[2] == -1) p.roughness = R(m.roughness_metallic[0]);
    else {
        R c[4];
        sample_texture<R>(s, m.tex0[2], uv0, uv1, c);
        p.roughness = c[m.tex1[2] & 3];
    }
    if (m.tex0[3] == -1) p.metallic = R(m.roughness_metallic[1]);
    else {
        R c[4];
        sample_texture<R>(s, m.tex0[3], uv0, uv1, c);
        p.metallic = c[m.tex1[3] & 3];
    }
    if (m.tex1[0] == -1) p.emissive = v4xyz<R>(m.emissive);
    else {
        R c[4];
        sample_texture<R>(s, m.tex1[0], uv0, uv1, c);
        p.emissive = V3<R>(c[0], c[1], c[2]);
    }
    p.roughness = gmax(p.roughness, kMinRoughness<R>());
    p.F0 = mix(V3<R>(R(0.03), R(0.03), R(0.03)), V3<R>(p.albedo[0], p.albedo[1], p.albedo[2]), p.metallic);
    p.alpha = p.roughness * p.roughness;
    p.alpha2 = p.alpha * p.alpha;
}

// sampling.glsl:28-36 make_rotation_matrix; returns M * t.
template <typename R>
static V3<R> rot_mul(V3<R> z, V3<R> t) {
    V3<R> ref = std::fabs(dot(z, V3<R>(0, 1, 0))) > R(0.99f) ? V3<R>(0, 0, 1) : V3<R>(0, 1, 0);
    V3<R> x = normalize(cross(ref, z));
    V3<R> y = cross(z, x);
    return x * t.x + y * t.y + z * t.z;
}
// brdf.glsl:6-18
template <typename R>
static V3<R> sample_cosine_lobe(V3<R> n, R r0, R r1) {
    R rx = gmax(R(0.00001f), r0), ry = gmax(R(0.00001f), r1);
    R phi = R(2) * kPi<R>() * ry;
    R ct = std::sqrt(rx), st = std::sqrt(R(1) - rx);
    V3<R> t(st * std::cos(phi), st * std::sin(phi), ct);
    return normalize(rot_mul(n, t));
}
template <typename R> static R pdf_cosine_lobe(R ndotl) { return ndotl / kPi<R>(); }
template <typename R>
static R D_ggx(R ndoth, R alpha) {  // brdf.glsl:58-64
    R a2 = alpha * alpha;
    R denom = (ndoth * ndoth) * (a2 - R(1)) + R(1);
    return a2 / gmax(kEps<R>(), kPi<R>() * denom * denom);
}
template <typename R>
static R G1_schlick_ggx(R roughness, R ndotv) {  // brdf.glsl:66-71
    R k = ((roughness + R(1)) * (roughness + R(1))) / R(8);
    return ndotv / gmax(kEps<R>(), ndotv * (R(1) - k) + k);
}
template <typename R>
static R G_schlick_ggx(R ndotl, R ndotv, R roughness) { return G1_schlick_ggx(roughness, ndotl) * G1_schlick_ggx(roughness, ndotv); }
template <typename R>
static V3<R> F_schlick(V3<R> f0, R vdoth) {
    R w = std::pow(R(1) - vdoth, R(5));
    return f0 + (V3<R>(1, 1, 1) - f0) * w;
}
template <typename R>
static V3<R> sample_ggx(V3<R> n, R alpha, R x0, R x1) {  // brdf.glsl:83-96
    R phi = R(2) * kPi<R>() * x0;
    R ct = std::sqrt((R(1) - x1) / (R(1) + (alpha * alpha - R(1)) * x1));
    R st = std::sqrt(R(1) - ct * ct);
    V3<R> d(st * std::cos(phi), st * std::sin(phi), ct);
    return normalize(rot_mul(n, d));
}
template <typename R>
static V3<R> evaluate_uber(const Surface<R>& p, V3<R> Wo, V3<R> Wh, V3<R> Wi) {  // brdf.glsl:109-121
    R NdotL = gmax(dot(p.normal, Wi), R(0));
    R NdotV = gmax(dot(p.normal, Wo), R(0));
    R NdotH = gmax(dot(p.normal, Wh), R(0));
    R VdotH = gmax(dot(Wi, Wh), R(0));
    V3<R> F = F_schlick(p.F0, VdotH);
    R a = p.roughness * p.roughness;
    V3<R> spec = ((F * D_ggx(NdotH, a)) * G_schlick_ggx(NdotL, NdotV, p.roughness)) / gmax(kEps<R>(), R(4) * NdotL * NdotV);
    V3<R> diffuse = V3<R>(p.albedo[0], p.albedo[1], p.albedo[2]) / kPi<R>();
    return (V3<R>(1, 1, 1) - F) * diffuse + spec;
}
template <typename R>
static R pdf_uber(const Surface<R>& p, V3<R> Wo, V3<R> Wh, V3<R> Wi) {  // brdf.glsl:123-134
    R NdotL = gmax(dot(p.normal, Wi), R(0));
    R NdotH = gmax(dot(p.normal, Wh), R(0));
    R VdotH = gmax(dot(Wi, Wh), R(0));
    (void)Wo;
    R pd = pdf_cosine_lobe(NdotL);
    R ps = D_ggx(NdotH, p.roughness * p.roughness) * NdotH / gmax(kEps<R>(), R(4) * VdotH);
    return mix(pd, ps, R(0.5));
}
// brdf.glsl:136-167; rng is taken BY VALUE (Q1).
template <typename R>
static V3<R> sample_uber(const Surface<R>& p, V3<R> Wo, RNG rng, V3<R>& Wi, R& pdf) {
    R alpha = p.roughness * p.roughness;
    R rx = next_float(rng), ry = next_float(rng), rz = next_float(rng);
    V3<R> Wh;
    bool is_specular = false;
    if (rx < R(0.5)) {
        Wh = sample_ggx(p.normal, alpha, ry, rz);
        Wi = reflect(-Wo, Wh);
        R NdotL = gmax(dot(p.normal, Wi), R(0));
        R NdotV = gmax(dot(p.normal, Wo), R(0));
        if (NdotL > R(0) && NdotV > R(0)) is_specular = true;
    }
    if (!is_specular) {
        Wi = sample_cosine_lobe(p.normal, ry, rz);
        Wh = normalize(Wo + Wi);
    }
    pdf = pdf_uber(p, Wo, Wh, Wi);
    return evaluate_uber(p, Wo, Wh, Wi);
}

struct Stats {
    uint64_t ext = 0, shadow = 0;
};

template <typename R>
struct Tracer {
    const Scene& s;
    R shadow_bias;
    uint32_t num_lights, max_bounces, flags;
    bool brute;
    Stats st;
    uint32_t rr_start = 0;  // settings.rr_start_depth (0 = the reference: RR at every bounce)
    R last_tmax = R(0);     // tmax of the last shadow ray (probe only)

    V3<R> env(V3<R> d) const {
        if (s.env_eq_w > 0) return sample_equirect<R>(s.env_eq, s.env_eq_w, d);
        return sample_cube<R>(s.env, s.env_size, d);
    }

    // sample_light (path_trace_rchit.glsl:284-451); returns Li * visibility.
    V3<R> sample_light(const Surface<R>& p, const LightData& light, uint32_t depth, RNG& rng, V3<R>& Wi, R& pdf) {
        uint32_t ray_flags = (depth == 0) ? 0u : (HH_TRACE_OPAQUE | HH_TRACE_TERMINATE_ON_FIRST);
        R tmin = R(0.0001), tmax = R(10000.0);
        V3<R> origin = p.vertex.pos + p.normal * shadow_bias;
        V3<R> Li(0, 0, 0);
        uint32_t type = (uint32_t)light.d0[0];
        if (type == 0) {  // directional
            R r0 = next_float(rng), r1 = next_float(rng);
            V3<R> ld = -v4xyz<R>(light.d1);
            V3<R> lt = normalize(cross(ld, V3<R>(0, 1, 0)));
            V3<R> lb = normalize(cross(lt, ld));
            R pr = R(light.d2[3]) * std::sqrt(r0);
            R pa = r1 * R(2) * kPi<R>();
            R dx = pr * std::cos(pa), dy = pr * std::sin(pa);
            Wi = normalize(ld + lt * dx + lb * dy);
            Li = v4xyz<R>(light.d0 + 1) * R(light.d1[3]);
            pdf = 0;
        } else if (type == 1 || type == 2) {  // spot / point
            R r0 = next_float(rng), r1 = next_float(rng);
            V3<R> to = v4xyz<R>(light.d2) - p.vertex.pos;
            V3<R> ld = normalize(to);
            R dist = length(to);
            R lr = R(light.d2[3]) / dist;
            R att = R(1);
            if (type == 1) att = smoothstep(R(light.d3[1]), R(light.d3[0]), dot(ld, -v4xyz<R>(light.d1)));
            V3<R> lt = normalize(cross(ld, V3<R>(0, 1, 0)));
            V3<R> lb = normalize(cross(lt, ld));
            R pr = lr * std::sqrt(r0);
            R pa = r1 * R(2) * kPi<R>();
            R dx = pr * std::cos(pa), dy = pr * std::sin(pa);
            Wi = normalize(ld + lt * dx + lb * dy);
            if (type == 1) Li = v4xyz<R>(light.d0 + 1) * R(light.d1[3]) * att / (dist * dist);
            else Li = v4xyz<R>(light.d0 + 1) * R(light.d1[3]) / (dist * dist);
            pdf = 0;
            tmax = dist;
        } else if (type == 3) {  // environment
            R r0 = next_float(rng), r1 = next_float(rng);
            Wi = sample_cosine_lobe(p.normal, r0, r1);
            Li = env(Wi);
            pdf = pdf_cosine_lobe(dot(p.normal, Wi));
        } else if (type == 4) {  // area (Q2, Q3)
            uint32_t mesh_id = (uint32_t)light.d0[1];
            uint32_t num_tris = (uint32_t)light.d1[2];
            uint32_t prim = next_uint(rng, num_tris);
            uint32_t mat_idx = (uint32_t)light.d0[2], off = (uint32_t)light.d0[3];
            const InstanceData& inst = s.inst_data[mesh_id];
            const Mesh& me = s.meshes[s.instances[mesh_id].mesh];
            const MaterialData& mat = s.mat_data[mat_idx];
            uint32_t pid = prim + off;
            const hh_vertex* tv[3] = {&me.vertices[me.indices[3 * pid]], &me.vertices[me.indices[3 * pid + 1]],
                                      &me.vertices[me.indices[3 * pid + 2]]};
            R b0 = next_float(rng), b1 = next_float(rng);
            R su0 = std::sqrt(b0);
            R bx = R(1) - su0, by = b1 * su0;  // uniform_sample_triangle
            V3<R> P[3];
            for (int k = 0; k < 3; k++) P[k] = mat4_mul_point<R>(inst.model, v4xyz<R>(tv[k]->position), R(tv[k]->position[3]));
            R w0 = R(1) - bx - by;
            V3<R> lp = P[0] * w0 + P[1] * bx + P[2] * by;
            V3<R> ln = normalize(mat3_mul<R>(inst.normal, v4xyz<R>(tv[0]->normal) * w0 + v4xyz<R>(tv[1]->normal) * bx +
                                                              v4xyz<R>(tv[2]->normal) * by));
            V3<R> ldir = p.vertex.pos - lp;
            R dsq = dot(ldir, ldir);
            R area = R(0.5) * length(cross(P[1] - P[0], P[2] - P[0]));
            if (area == R(0) || dsq == R(0)) { pdf = 0; return V3<R>(0, 0, 0); }
            R dist = std::sqrt(dsq);
            ldir = ldir / dist;
            tmax = gmax(R(0), dist - kEps<R>());
            R ct = dot(ln, ldir);
            if (ct == R(0)) { pdf = 0; return V3<R>(0, 0, 0); }
            Li = v4xyz<R>(mat.emissive);
            Wi = -ldir;
            pdf = dsq / gmax(kEps<R>(), ct * area);
        }
        st.shadow++;
        last_tmax = tmax;
        Hit<R> h = trace<R>(s, origin, Wi, tmin, tmax, ray_flags, false, brute);
        return h.found ? Li * R(0) : Li;
    }

    // Extension (HH_FLAG_LIGHT_POWER): an emissive triangle by power
    // (first CDF entry above u), a uniform point on it, front faces only;
    // pdf in solid angle = P(tri) / area * d^2 / cos.
    V3<R> sample_emitter(const Surface<R>& p, uint32_t depth, RNG& rng, V3<R>& Wi, R& pdf) {
        uint32_t ray_flags = (depth == 0) ? 0u : (HH_TRACE_OPAQUE | HH_TRACE_TERMINATE_ON_FIRST);
        R tmin = R(0.0001), tmax = R(10000.0);
        V3<R> origin = p.vertex.pos + p.normal * shadow_bias;
        const float u = next_float(rng);
        size_t t = (size_t)(std::upper_bound(s.emit_cdf.begin(), s.emit_cdf.end(), u) - s.emit_cdf.begin());
        if (t >= s.emit_cdf.size()) t = s.emit_cdf.size() - 1;
        R b0 = next_float(rng), b1 = next_float(rng);
        R su0 = std::sqrt(b0);
        R bx = R(1) - su0, by = b1 * su0;
        const float* v = &s.emit_v[9 * t];
        V3<R> P0(v[0], v[1], v[2]), P1(v[3], v[4], v[5]), P2(v[6], v[7], v[8]);
        R w0 = R(1) - bx - by;
        V3<R> lp = P0 * w0 + P1 * bx + P2 * by;
        V3<R> ln = normalize(cross(P1 - P0, P2 - P0));
        V3<R> ldir = p.vertex.pos - lp;
        R dsq = dot(ldir, ldir);
        if (dsq == R(0)) { pdf = 0; return V3<R>(0, 0, 0); }
        R dist = std::sqrt(dsq);
        ldir = ldir / dist;
        tmax = gmax(R(0), dist - kEps<R>());
        R ct = dot(ln, ldir);
        if (!(ct > R(0))) { pdf = 0; return V3<R>(0, 0, 0); }
        V3<R> Li(R(s.emit_le[3 * t]), R(s.emit_le[3 * t + 1]), R(s.emit_le[3 * t + 2]));
        Wi = -ldir;
        pdf = R(s.emit_p[t]) / R(s.emit_area[t]) * dsq / ct;
        st.shadow++;
        Hit<R> h = trace<R>(s, origin, Wi, tmin, tmax, ray_flags, false, brute);
        return h.found ? Li * R(0) : Li;
    }

    V3<R> direct_lighting(const Surface<R>& p, V3<R> ray_dir, V3<R> T, uint32_t depth, RNG& rng) {  // :455-483
        V3<R> L(0, 0, 0);
        // reference: uniform among the packed lights; power mode: uniform among
        // [all emitters as one entry] + the lights after the area records
        const bool power = (flags & HH_FLAG_LIGHT_POWER) != 0u;
        const uint32_t nemit = power && !s.emit_cdf.empty() ? 1u : 0u;
        const uint32_t nl = power ? num_lights - s.num_area + nemit : num_lights;
        uint32_t li = next_uint(rng, nl);
        if (nl == 0) return L;
        V3<R> Wo = -ray_dir, Wi(0, 0, 0);
        R pdf = 0;
        V3<R> Li;
        if (nemit && li == 0) Li = sample_emitter(p, depth, rng, Wi, pdf);
        else Li = sample_light(p, s.light_data[power ? li - nemit + s.num_area : li], depth, rng, Wi, pdf);
        V3<R> Wh = normalize(Wo + Wi);
        V3<R> brdf = evaluate_uber(p, Wo, Wh, Wi);
        R ct = gclamp(dot(p.normal, Wi), R(0), R(1));
        if (!is_black(Li)) {
            if (pdf == R(0)) L = T * brdf * ct * Li;
            else L = (T * brdf * ct * Li) / pdf;
        }
        return L * R(nl);
    }

    // RAY_DEBUG_VIEW variant (path_trace_rchit.glsl:500-514,548-567;
    // path_trace_rmiss.glsl:40-58): no Russian roulette and no RR draw; every
    // non-primary ray appends (origin, end, colour) to `segments`.
    bool debug = false;
    V3<R> debug_color;
    std::vector<float>* segments = nullptr;

    // payload: L, T, depth, rng. Returns L of this ray.
    V3<R> trace_path(V3<R> o, V3<R> d, R tmin, uint32_t rflags, V3<R> T, uint32_t depth, RNG rng) {
        st.ext++;
        Hit<R> h = trace<R>(s, o, d, tmin, R(10000.0), rflags, true, brute);
        if (debug && depth > 0) {
            const V3<R> e = o + d * (h.found ? h.t : R(10000.0));
            const R rec[9] = {o.x, o.y, o.z, e.x, e.y, e.z, debug_color.x, debug_color.y, debug_color.z};
            for (R x : rec) segments->push_back((float)x);
        }
        if (!h.found) {  // path_trace_rmiss.glsl:60-65
            V3<R> e = env(d);
            return depth == 0 ? e : T * e;
        }
        Surface<R> p;
        populate(s, h, p);
        V3<R> L(0, 0, 0);
        if (depth == 0 && !is_black(p.emissive)) L = L + p.emissive;
        L = L + direct_lighting(p, d, T, depth, rng);
        if ((flags & HH_FLAG_DIRECT_ONLY) == 0 && depth + 1 < max_bounces) L = L + indirect(p, d, T, depth, rng);
        return L;
    }

    V3<R> indirect(const Surface<R>& p, V3<R> ray_dir, V3<R> T, uint32_t depth, RNG& rng) {  // :487-536
        V3<R> Wo = -ray_dir, Wi;
        R pdf;
        V3<R> brdf = sample_uber(p, Wo, rng, Wi, pdf);
        R ct = gclamp(dot(p.normal, Wi), R(0), R(1));
        V3<R> Tn = (T * (brdf * ct)) / pdf;
        if (debug) return trace_path(p.vertex.pos, Wi, R(0.0001), HH_TRACE_OPAQUE, Tn, depth + 1, rng);
        R prob = gmax(Tn.x, gmax(Tn.y, Tn.z));
        // the RR draw is consumed at every bounce so the stream stays aligned;
        // bounces before rr_start neither terminate nor rescale
        const R u = R(next_float(rng));
        if (depth >= rr_start) {
            if (u > prob) return V3<R>(0, 0, 0);
            Tn = Tn * (R(1) / prob);
        }
        return trace_path(p.vertex.pos, Wi, R(0.0001), HH_TRACE_OPAQUE, Tn, depth + 1, rng);
    }
};

// generate_ray (path_trace_rgen.glsl:132-174) in precision R from consts.
template <typename R>
struct CamR {
    R vpi[4][4];
    V3<R> cam, up, right;
    R fplane[4];
    R aperture;
};
template <typename R>
static CamR<R> cam_from(const hh_launch_consts& c, const ConstsD& cd) {
    CamR<R> r;
    for (int col = 0; col < 4; col++)
        for (int row = 0; row < 4; row++)
            r.vpi[col][row] = sizeof(R) == 4 ? R(c.view_proj_inverse[col * 4 + row]) : R(cd.vpi[col][row]);
    if (sizeof(R) == 4) {
        r.cam = v4xyz<R>(c.camera_pos);
        r.up = v4xyz<R>(c.up_direction);
        r.right = v4xyz<R>(c.right_direction);
        for (int i = 0; i < 4; i++) r.fplane[i] = R(c.focal_plane[i]);
        r.aperture = R(c.aperture_radius);
    } else {
        r.cam = V3<R>(cd.cam[0], cd.cam[1], cd.cam[2]);
        r.up = V3<R>(cd.up[0], cd.up[1], cd.up[2]);
        r.right = V3<R>(cd.right[0], cd.right[1], cd.right[2]);
        for (int i = 0; i < 4; i++) r.fplane[i] = R(cd.fplane[i]);
        r.aperture = R(cd.aperture);
    }
    return r;
}
template <typename R>
static void generate_ray(const CamR<R>& c, uint32_t px, uint32_t py, uint32_t w, uint32_t h, RNG& rng, V3<R>& o, V3<R>& d) {
    R jx = next_float(rng), jy = next_float(rng);
    R tcx = ((R(px) + R(0.5)) + jx) / R(w);
    R tcy = ((R(py) + R(0.5)) + jy) / R(h);
    R nx = tcx * R(2) - R(1), ny = tcy * R(2) - R(1);
    // view_proj_inverse * vec4(nx, ny, 0, 1)
    R tg[4];
    for (int r = 0; r < 4; r++) tg[r] = ((c.vpi[0][r] * nx + c.vpi[1][r] * ny) + c.vpi[2][r] * R(0)) + c.vpi[3][r] * R(1);
    V3<R> target(tg[0] / tg[3], tg[1] / tg[3], tg[2] / tg[3]);
    R angle = R(next_float(rng)) * R(2) * kPi<R>();
    R radius = std::sqrt(R(next_float(rng)));
    R ox = std::cos(angle) * radius * c.aperture, oy = std::sin(angle) * radius * c.aperture;
    V3<R> ap = (c.cam + c.right * ox) + c.up * oy;
    V3<R> rdir = -normalize(target - c.cam);
    V3<R> fp(c.fplane[0], c.fplane[1], c.fplane[2]);
    R t = -(dot(c.cam, fp) + c.fplane[3]) / dot(rdir, fp);
    V3<R> focus = c.cam + rdir * t;
    o = ap;
    d = normalize(focus - ap);
}

template <typename R>
static void render_impl(const Scene& s, const hh_camera& cam, const hh_settings& st, uint32_t w, uint32_t h, uint32_t x0,
                        uint32_t y0, uint32_t x1, uint32_t y1, uint32_t first_frame, uint32_t frames, int threads, float* img,
                        uint64_t* stats) {
    hh_launch_consts c;
    launch_consts(cam, w, h, s.num_lights, 0, st, c);
    ConstsD cd;
    launch_consts_d(cam, w, h, cd);
    CamR<R> cr = cam_from<R>(c, cd);
    std::atomic<uint32_t> next_row(y0);
    std::atomic<uint64_t> n_ext(0), n_sh(0);
    auto worker = [&]() {
        Tracer<R> tr{s, R(st.shadow_ray_bias), s.num_lights, st.max_ray_bounces, st.flags, false, {}};
        tr.rr_start = st.rr_start_depth;
        for (;;) {
            uint32_t y = next_row.fetch_add(1);
            if (y >= y1) break;
            for (uint32_t x = x0; x < x1; x++) {
                float* px = img + (size_t(y) * w + x) * 4;
                for (uint32_t f = first_frame; f < first_frame + frames; f++) {
                    RNG rng = rng_init(x, y, f);
                    V3<R> o, d;
                    generate_ray(cr, x, y, w, h, rng, o, d);
                    V3<R> L = tr.trace_path(o, d, R(0.001), 0u, V3<R>(1, 1, 1), 0, rng);
                    V3<R> cl = (st.flags & HH_FLAG_NO_RADIANCE_CLAMP) ? L : vmin(L, R(1));  // Q4
                    if (f == 0) {  // path_trace_rgen.glsl:221-237 (Q5)
                        px[0] = (float)cl.x; px[1] = (float)cl.y; px[2] = (float)cl.z;
                    } else {
                        R pr[3] = {R(px[0]), R(px[1]), R(px[2])};
                        R cc[3] = {cl.x, cl.y, cl.z};
                        for (int k = 0; k < 3; k++) px[k] = (float)(pr[k] + (cc[k] - pr[k]) / R(f));
                    }
                    px[3] = 1.0f;
                }
            }
        }
        n_ext += tr.st.ext;
        n_sh += tr.st.shadow;
    };
    std::vector<std::thread> pool;
    for (int i = 0; i < std::max(1, threads); i++) pool.emplace_back(worker);
    for (auto& t : pool) t.join();
    if (stats) {
        stats[0] = n_ext.load();
        stats[1] = n_sh.load();
    }
}

// Output-buffer views: debug_visualization.vert:49-60 and .frag:60-163,
// rasterised by Renderer::render_debug_visualization (renderer.cpp:491-547;
// pipeline state :1178-1300). Perspective-correct interpolation at a fragment
// centre is barycentric interpolation at the point where the ray from the eye
// through that centre meets the triangle, so each pixel is one un-jittered
// ray: depth test LESS_OR_EQUAL = closest hit; VK_CULL_MODE_NONE; no discard,
// so alpha-tested geometry is drawn solid (an opaque trace); glm's GL-style
// projection under Vulkan's 0 <= z_clip <= w_clip keeps view depths in
// [2fn/(f+n), f].
template <typename R>
static void output_buffer_value(const Scene& s, const Hit<R>& h, uint32_t which, R out[4]) {
    const InstanceData& inst = s.inst_data[h.inst];
    const uint32_t* info = &s.submesh_info[h.inst][h.geom * 2];
    const Mesh& me = s.meshes[s.instances[h.inst].mesh];
    const MaterialData& m = s.mat_data[info[1]];
    const uint32_t pid = h.prim + info[0];
    const hh_vertex* tv[3] = {&me.vertices[me.indices[3 * pid]], &me.vertices[me.indices[3 * pid + 1]],
                              &me.vertices[me.indices[3 * pid + 2]]};
    const R b[3] = {R(1) - h.u - h.v, h.u, h.v};
    R uv0 = 0, uv1 = 0;
    for (int k = 0; k < 3; k++) {
        uv0 += R(tv[k]->tex_coord[0]) * b[k];
        uv1 += R(tv[k]->tex_coord[1]) * b[k];
    }
    R c[4];
    V3<R> v(0, 0, 0);
    switch (which) {
        case 0:  // fetch_albedo
            if (m.tex0[0] == -1) v = v4xyz<R>(m.albedo);
            else {
                sample_texture<R>(s, m.tex0[0], uv0, uv1, c);
                v = V3<R>(c[0], c[1], c[2]);
            }
            break;
        case 1: {  // vertex stage normalises each transformed normal; no renormalisation after interpolation
            V3<R> n(0, 0, 0), t(0, 0, 0), bt(0, 0, 0);
            for (int k = 0; k < 3; k++) {
                n = n + normalize(mat3_mul<R>(inst.normal, v4xyz<R>(tv[k]->normal))) * b[k];
                t = t + normalize(mat3_mul<R>(inst.normal, v4xyz<R>(tv[k]->tangent))) * b[k];
                bt = bt + normalize(mat3_mul<R>(inst.normal, v4xyz<R>(tv[k]->bitangent))) * b[k];
            }
            if (m.tex0[1] != -1) {  // get_normal_from_map
                sample_texture<R>(s, m.tex0[1], uv0, uv1, c);
                V3<R> T = normalize(t), B = normalize(bt), N = normalize(n);
                V3<R> tn = normalize(V3<R>(c[0] * R(2) - R(1), c[1] * R(2) - R(1), c[2] * R(2) - R(1)));
                n = normalize(T * tn.x + B * tn.y + N * tn.z);
            }
            v = n * R(0.5) + V3<R>(R(0.5), R(0.5), R(0.5));
            break;
        }
        case 2:
        case 3: {  // fetch_roughness / fetch_metallic (the raw value: no MIN_ROUGHNESS)
            const int k = int(which) - 2;
            R x;
            if (m.tex0[2 + k] == -1) x = R(m.roughness_metallic[k]);
            else {
                sample_texture<R>(s, m.tex0[2 + k], uv0, uv1, c);
                x = c[m.tex1[2 + k] & 3];
            }
            v = V3<R>(x, x, x);
            break;
        }
        default:  // fetch_emissive
            if (m.tex1[0] == -1) v = v4xyz<R>(m.emissive);
            else {
                sample_texture<R>(s, m.tex1[0], uv0, uv1, c);
                v = V3<R>(c[0], c[1], c[2]);
            }
    }
    out[0] = v.x;
    out[1] = v.y;
    out[2] = v.z;
    out[3] = R(1);
}

}  // namespace orc

// ===========================================================================
// C ABI for ctypes.
// ===========================================================================
using namespace orc;

extern "C" {

int orc_load_hosek(const double* data, uint32_t n) {
    if (n != 3600) return 1;
    g_hosek.assign(data, data + n);
    return 0;
}
void* orc_scene_create(void) { return new Scene(); }
void orc_scene_destroy(void* s) { delete (Scene*)s; }
int orc_scene_add_texture(void* sp, uint32_t w, uint32_t h, uint32_t fmt, const void* texels) {
    Scene& s = *(Scene*)sp;
    Texture t;
    t.w = w;
    t.h = h;
    t.format = fmt;
    decode_texture(t, texels);
    s.textures.push_back(std::move(t));
    return (int)s.textures.size() - 1;
}
int orc_scene_add_material(void* sp, const hh_material_desc* d) {
    Scene& s = *(Scene*)sp;
    s.materials.push_back(*d);
    return (int)s.materials.size() - 1;
}
int orc_scene_add_mesh(void* sp, const hh_vertex* v, uint32_t nv, const uint32_t* idx, uint32_t ni, const hh_submesh* sm,
                       uint32_t ns, const uint32_t* mats, uint32_t nm) {
    Scene& s = *(Scene*)sp;
    Mesh m;
    m.vertices.assign(v, v + nv);
    m.indices.assign(idx, idx + ni);
    m.submeshes.assign(sm, sm + ns);
    m.materials.assign(mats, mats + nm);
    s.meshes.push_back(std::move(m));
    return (int)s.meshes.size() - 1;
}
int orc_scene_add_instance(void* sp, const hh_instance_desc* d) {
    Scene& s = *(Scene*)sp;
    s.instances.push_back(*d);
    return (int)s.instances.size() - 1;
}
int orc_scene_add_light(void* sp, const hh_light_desc* d) {
    Scene& s = *(Scene*)sp;
    s.lights.push_back(*d);
    return (int)s.lights.size() - 1;
}
int orc_scene_set_ibl(void* sp, uint32_t size, const float* faces) {
    Scene& s = *(Scene*)sp;
    s.ibl_size = size;
    s.ibl.assign(faces, faces + size_t(6) * size * size * 4);
    return 0;
}
int orc_scene_set_sky(void* sp, const hh_sky_params* p) {
    ((Scene*)sp)->sky = *p;
    return 0;
}
int orc_scene_commit(void* sp) {
    Scene& s = *(Scene*)sp;
    if (g_hosek.empty()) {
        for (auto& l : s.lights)
            if (l.type == HH_LIGHT_DIRECTIONAL && s.ibl_size == 0) return 2;
    }
    commit(s);
    return 0;
}
uint32_t orc_scene_num_lights(void* sp) { return ((Scene*)sp)->num_lights; }
uint32_t orc_scene_read_lights(void* sp, float* out, uint32_t cap) {
    Scene& s = *(Scene*)sp;
    uint32_t n = std::min<uint32_t>(cap, (uint32_t)s.light_data.size());
    std::memcpy(out, s.light_data.data(), n * sizeof(LightData));
    return n;
}
uint32_t orc_scene_read_materials(void* sp, void* out, uint32_t cap) {
    Scene& s = *(Scene*)sp;
    uint32_t n = std::min<uint32_t>(cap, (uint32_t)s.mat_data.size());
    std::memcpy(out, s.mat_data.data(), n * sizeof(MaterialData));
    return n;
}
uint32_t orc_scene_env_size(void* sp) { return ((Scene*)sp)->env_size; }
uint32_t orc_scene_env_equirect(void* sp, uint16_t* out) {
    const Scene& s = *(Scene*)sp;
    if (out && s.env_eq_w) std::memcpy(out, s.env_eq.data(), s.env_eq.size() * 2);
    return s.env_eq_w;
}
void orc_scene_read_env(void* sp, float* out) {
    Scene& s = *(Scene*)sp;
    std::memcpy(out, s.env.data(), s.env.size() * 4);
}
void orc_scene_sky_coefficients(void* sp, float out[30]) { std::memcpy(out, ((Scene*)sp)->sky_coef, 120); }
void orc_sky_coefficients(const hh_sky_params* p, const float sun_dir[3], float out[30]) { sky_coefficients(*p, sun_dir, out); }
void orc_sky_eval(const float coef[30], const float sun[3], const float* dirs, uint32_t n, int fp64, float* out) {
    for (uint32_t i = 0; i < n; i++) {
        if (fp64) {
            V3<double> c = hosek_sky_rgb<double>(V3<double>(dirs[i * 3], dirs[i * 3 + 1], dirs[i * 3 + 2]), V3<double>(sun[0], sun[1], sun[2]), coef);
            out[i * 3] = (float)c.x; out[i * 3 + 1] = (float)c.y; out[i * 3 + 2] = (float)c.z;
        } else {
            V3<float> c = hosek_sky_rgb<float>(V3<float>(dirs[i * 3], dirs[i * 3 + 1], dirs[i * 3 + 2]), V3<float>(sun[0], sun[1], sun[2]), coef);
            out[i * 3] = c.x; out[i * 3 + 1] = c.y; out[i * 3 + 2] = c.z;
        }
    }
}
void orc_launch_consts(const hh_camera* cam, uint32_t w, uint32_t h, uint32_t nl, uint32_t frame, const hh_settings* st,
                       hh_launch_consts* out) {
    launch_consts(*cam, w, h, nl, frame, *st, *out);
}
int orc_render(void* sp, const hh_camera* cam, const hh_settings* st, uint32_t w, uint32_t h, uint32_t x0, uint32_t y0,
               uint32_t x1, uint32_t y1, uint32_t first_frame, uint32_t frames, int fp64, int threads, float* img,
               uint64_t* stats) {
    Scene& s = *(Scene*)sp;
    if (!s.committed) return 1;
    if (fp64) render_impl<double>(s, *cam, *st, w, h, x0, y0, x1, y1, first_frame, frames, threads, img, stats);
    else render_impl<float>(s, *cam, *st, w, h, x0, y0, x1, y1, first_frame, frames, threads, img, stats);
    return 0;
}
int orc_primary_hits(void* sp, const hh_camera* cam, uint32_t w, uint32_t h, uint32_t frame, int brute, hh_hit* out) {
    Scene& s = *(Scene*)sp;
    if (!s.committed) return 1;
    hh_settings st{};
    hh_launch_consts c;
    launch_consts(*cam, w, h, s.num_lights, frame, st, c);
    ConstsD cd;
    launch_consts_d(*cam, w, h, cd);
    CamR<float> cr = cam_from<float>(c, cd);
    for (uint32_t y = 0; y < h; y++)
        for (uint32_t x = 0; x < w; x++) {
            RNG rng = rng_init(x, y, frame);
            V3<float> o, d;
            generate_ray(cr, x, y, w, h, rng, o, d);
            Hit<float> hh = trace<float>(s, o, d, 0.001f, 10000.0f, 0u, true, brute != 0);
            hh_hit& r = out[size_t(y) * w + x];
            r.t = hh.found ? hh.t : -1.0f;
            r.u = hh.u;
            r.v = hh.v;
            r.instance = hh.found ? hh.inst : 0xFFFFFFFFu;
            r.geometry = hh.found ? hh.geom : 0xFFFFFFFFu;
            r.primitive = hh.found ? hh.prim : 0xFFFFFFFFu;
        }
    return 0;
}
// Output-buffer view `which` (0 albedo .. 4 emissive) into w*h RGBA floats,
// rows in the path tracer's order; background (0, 0, 0, 1).
int orc_output_buffer(void* sp, const hh_camera* cam, uint32_t w, uint32_t h, uint32_t which, int brute, float* out) {
    Scene& s = *(Scene*)sp;
    if (!s.committed || which > 4) return 1;
    hh_settings st{};
    hh_launch_consts c;
    launch_consts(*cam, w, h, s.num_lights, 0, st, c);
    const double fd = cam->far_plane, nd = cam->near_plane;
    const float near_d = (float)(2.0 * fd * nd / (fd + nd)), far_d = cam->far_plane;
    const V3<float> eye = v4xyz<float>(c.camera_pos);
    const V3<float> fwd(-c.focal_plane[0], -c.focal_plane[1], -c.focal_plane[2]);
    const float* m = c.view_proj_inverse;
    for (uint32_t y = 0; y < h; y++)
        for (uint32_t x = 0; x < w; x++) {
            const float nx = ((float(x) + 0.5f) / float(w)) * 2.0f - 1.0f;
            const float ny = ((float(y) + 0.5f) / float(h)) * 2.0f - 1.0f;
            float tg[4];
            for (int r = 0; r < 4; r++) tg[r] = (m[0 * 4 + r] * nx + m[1 * 4 + r] * ny) + m[3 * 4 + r];
            const V3<float> d = normalize(V3<float>(tg[0] / tg[3], tg[1] / tg[3], tg[2] / tg[3]) - eye);
            const float cs = dot(d, fwd);
            float* px = out + (size_t(y) * w + x) * 4;
            px[0] = px[1] = px[2] = 0.0f;
            px[3] = 1.0f;
            if (!(cs > 0.0f)) continue;
            Hit<float> hh = trace<float>(s, eye, d, near_d / cs, far_d / cs, HH_TRACE_OPAQUE, true, brute != 0);
            if (hh.found) output_buffer_value<float>(s, hh, which, px);
        }
    return 0;
}

// The ray-debug variant (path_trace_debug.rgen): num_rays paths through the
// window pixel (px, py) of a w x h view (shader pixel (px, h - py)), launch id
// (i, 0), seed `frame`. Writes up to max_segments records of 9 floats
// (origin xyz, end xyz, colour rgb), ray by ray in depth order; returns the
// number of segments produced.
uint32_t orc_ray_debug(void* sp, const hh_camera* cam, const hh_settings* st, uint32_t w, uint32_t h, uint32_t px,
                       uint32_t py, uint32_t num_rays, uint32_t frame, float* out, uint32_t max_segments) {
    Scene& s = *(Scene*)sp;
    if (!s.committed) return 0;
    hh_launch_consts c;
    launch_consts(*cam, w, h, s.num_lights, frame, *st, c);
    ConstsD cd;
    launch_consts_d(*cam, w, h, cd);
    CamR<float> cr = cam_from<float>(c, cd);
    std::vector<float> segs;
    Tracer<float> tr{s, st->shadow_ray_bias, s.num_lights, st->max_ray_bounces, st->flags & HH_FLAG_LIGHT_POWER, false, {}};
    tr.debug = true;
    tr.segments = &segs;
    for (uint32_t i = 0; i < num_rays; i++) {
        RNG rng = rng_init(i, 0, frame);
        const float r = next_float(rng) * 0.5f + 0.5f;
        const float g = next_float(rng) * 0.5f + 0.5f;
        const float b = next_float(rng) * 0.5f + 0.5f;
        tr.debug_color = V3<float>(r, g, b);
        V3<float> o, d;
        generate_ray(cr, px, h - py, w, h, rng, o, d);
        tr.trace_path(o, d, 0.001f, 0u, V3<float>(1, 1, 1), 0, rng);
    }
    const uint32_t n = (uint32_t)(segs.size() / 9);
    std::memcpy(out, segs.data(), sizeof(float) * 9 * std::min(n, max_segments));
    return n;
}

int orc_trace_rays(void* sp, const hh_ray* rays, uint32_t n, uint32_t flags, int brute, int fp64, hh_hit* out) {
    Scene& s = *(Scene*)sp;
    if (!s.committed) return 1;
    bool closest = !(flags & HH_TRACE_TERMINATE_ON_FIRST);
    for (uint32_t i = 0; i < n; i++) {
        const hh_ray& r = rays[i];
        hh_hit& o = out[i];
        bool found;
        if (fp64) {
            Hit<double> h = trace<double>(s, V3<double>(r.origin[0], r.origin[1], r.origin[2]),
                                          V3<double>(r.direction[0], r.direction[1], r.direction[2]), r.tmin, r.tmax, flags, closest, brute != 0);
            found = h.found;
            o.t = (float)h.t; o.u = (float)h.u; o.v = (float)h.v;
            o.instance = h.inst; o.geometry = h.geom; o.primitive = h.prim;
        } else {
            Hit<float> h = trace<float>(s, V3<float>(r.origin[0], r.origin[1], r.origin[2]),
                                        V3<float>(r.direction[0], r.direction[1], r.direction[2]), r.tmin, r.tmax, flags, closest, brute != 0);
            found = h.found;
            o.t = h.t; o.u = h.u; o.v = h.v;
            o.instance = h.inst; o.geometry = h.geom; o.primitive = h.prim;
        }
        if (!found) {
            o.t = -1.0f;
            o.instance = o.geometry = o.primitive = 0xFFFFFFFFu;
        }
    }
    return 0;
}
// Environment lookup (cube or equirect, as the miss shader and the env light
// see it) for n directions.
void orc_sample_env(void* sp, const float* dirs, uint32_t n, int fp64, float* out) {
    Scene& s = *(Scene*)sp;
    for (uint32_t i = 0; i < n; i++) {
        if (fp64) {
            Tracer<double> tr{s, 0.0, s.num_lights, 1, 0, false};
            V3<double> e = tr.env(V3<double>(dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]));
            out[3 * i] = (float)e.x; out[3 * i + 1] = (float)e.y; out[3 * i + 2] = (float)e.z;
        } else {
            Tracer<float> tr{s, 0.0f, s.num_lights, 1, 0, false};
            V3<float> e = tr.env(V3<float>(dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]));
            out[3 * i] = e.x; out[3 * i + 1] = e.y; out[3 * i + 2] = e.z;
        }
    }
}
// One sample_light call (FP64) on packed light record `light` from a surface
// point with geometric and shading normal `nrm`: out = Wi xyz, Li xyz (times
// visibility), pdf, shadow-ray tmax, RNG draws consumed. The RNG state is
// advanced in place.
int orc_sample_light_probe(void* sp, uint32_t light, const double* pos, const double* nrm, uint32_t depth,
                           uint32_t st[2], double* out) {
    Scene& s = *(Scene*)sp;
    if (!s.committed || light >= s.light_data.size()) return 1;
    Tracer<double> tr{s, 0.0, s.num_lights, 1, 0, false};
    Surface<double> p{};
    p.vertex.pos = V3<double>(pos[0], pos[1], pos[2]);
    p.normal = p.vertex.normal = V3<double>(nrm[0], nrm[1], nrm[2]);
    RNG r{st[0], st[1]};
    const RNG r0 = r;
    V3<double> Wi(0, 0, 0);
    double pdf = 0;
    V3<double> Li = tr.sample_light(p, s.light_data[light], depth, r, Wi, pdf);
    out[0] = Wi.x; out[1] = Wi.y; out[2] = Wi.z;
    out[3] = Li.x; out[4] = Li.y; out[5] = Li.z;
    out[6] = pdf;
    out[7] = tr.last_tmax;
    uint32_t draws = 0;
    for (RNG q = r0; (q.s0 != r.s0 || q.s1 != r.s1) && draws < 16; draws++) rng_next(q);
    out[8] = draws;
    st[0] = r.s0;
    st[1] = r.s1;
    return 0;
}
// Unit-level entry points for known-answer tests.
uint32_t orc_rng_hash(uint32_t x) { return rng_hash(x); }
void orc_rng_init(uint32_t x, uint32_t y, uint32_t frame, uint32_t st[2]) {
    RNG r = rng_init(x, y, frame);
    st[0] = r.s0;
    st[1] = r.s1;
}
uint32_t orc_rng_next(uint32_t st[2]) {
    RNG r{st[0], st[1]};
    uint32_t v = rng_next(r);
    st[0] = r.s0;
    st[1] = r.s1;
    return v;
}
float orc_next_float(uint32_t st[2]) {
    RNG r{st[0], st[1]};
    float f = next_float(r);
    st[0] = r.s0;
    st[1] = r.s1;
    return f;
}
uint32_t orc_next_uint(uint32_t st[2], uint32_t nmax) {
    RNG r{st[0], st[1]};
    uint32_t v = next_uint(r, nmax);
    st[0] = r.s0;
    st[1] = r.s1;
    return v;
}
// BRDF probes in FP64: surface (albedo rgb, roughness, metallic, normal).
static Surface<double> mk_surface(const double* a) {
    Surface<double> p{};
    p.albedo[0] = a[0]; p.albedo[1] = a[1]; p.albedo[2] = a[2]; p.albedo[3] = 1;
    p.roughness = std::max(a[3], 0.1);
    p.metallic = a[4];
    p.normal = V3<double>(a[5], a[6], a[7]);
    p.F0 = mix(V3<double>(0.03, 0.03, 0.03), V3<double>(a[0], a[1], a[2]), p.metallic);
    p.alpha = p.roughness * p.roughness;
    p.alpha2 = p.alpha * p.alpha;
    return p;
}
void orc_eval_uber(const double* surf, const double* wo, const double* wi, double* out_rgb_pdf) {
    Surface<double> p = mk_surface(surf);
    V3<double> Wo(wo[0], wo[1], wo[2]), Wi(wi[0], wi[1], wi[2]);
    V3<double> Wh = normalize(Wo + Wi);
    V3<double> f = evaluate_uber(p, Wo, Wh, Wi);
    out_rgb_pdf[0] = f.x; out_rgb_pdf[1] = f.y; out_rgb_pdf[2] = f.z;
    out_rgb_pdf[3] = pdf_uber(p, Wo, Wh, Wi);
}
void orc_sample_uber(const double* surf, const double* wo, uint32_t st[2], double* out_wi_pdf_f) {
    Surface<double> p = mk_surface(surf);
    RNG r{st[0], st[1]};
    V3<double> Wi;
    double pdf;
    V3<double> f = sample_uber(p, V3<double>(wo[0], wo[1], wo[2]), r, Wi, pdf);
    out_wi_pdf_f[0] = Wi.x; out_wi_pdf_f[1] = Wi.y; out_wi_pdf_f[2] = Wi.z;
    out_wi_pdf_f[3] = pdf;
    out_wi_pdf_f[4] = f.x; out_wi_pdf_f[5] = f.y; out_wi_pdf_f[6] = f.z;
}
double orc_D_ggx(double ndoth, double alpha) { return D_ggx(ndoth, alpha); }
void orc_sample_cosine_lobe(const double* n, double r0, double r1, double* out) {
    V3<double> d = sample_cosine_lobe(V3<double>(n[0], n[1], n[2]), r0, r1);
    out[0] = d.x; out[1] = d.y; out[2] = d.z;
}
void orc_sample_ggx(const double* n, double alpha, double x0, double x1, double* out) {
    V3<double> d = sample_ggx(V3<double>(n[0], n[1], n[2]), alpha, x0, x1);
    out[0] = d.x; out[1] = d.y; out[2] = d.z;
}
// tone_map.frag:20-51: exposure -> op -> gamma -> UNORM8 (round to nearest).
void orc_tonemap(const float* rgba, uint32_t n, float exposure, uint32_t op, uint8_t* out) {
    for (uint32_t i = 0; i < n; i++) {
        for (int c = 0; c < 3; c++) {
            float x = rgba[i * 4 + c] * exposure;
            if (op == 0) {
                const float a = 2.51f, b = 0.03f, cc = 2.43f, d = 0.59f, e = 0.14f;
                x = std::min(std::max((x * (a * x + b)) / (x * (cc * x + d) + e), 0.0f), 1.0f);
            } else if (op == 1) {
                x = x / (1.0f + x);
            }
            x = std::pow(x, 1.0f / 2.2f);
            x = std::min(std::max(x, 0.0f), 1.0f);
            out[i * 4 + c] = (uint8_t)std::floor(x * 255.0f + 0.5f);
        }
        out[i * 4 + 3] = 255;
    }
}
// Watertight test probe (FP32 transcription), for edge-case tests.
int orc_tri_hit(const float* o, const float* d, const float* v9, float tmin, float tmax, float* out_tuv) {
    float v[3][3];
    for (int k = 0; k < 3; k++)
        for (int c = 0; c < 3; c++) v[k][c] = v9[k * 3 + c];
    RayPre<float> pr = ray_pre(V3<float>(d[0], d[1], d[2]));
    float t, u, vv;
    bool h = tri_hit(pr, V3<float>(o[0], o[1], o[2]), v, tmin, tmax, t, u, vv);
    if (h) { out_tuv[0] = t; out_tuv[1] = u; out_tuv[2] = vv; }
    return h ? 1 : 0;
}

}  // extern "C"
