// glTF 2.0 scene ingestion.
//
// The reference loads scenes through ResourceManager on top of AssetCore
// (resource_manager.cpp:48-107 images, 196-222 scenes, 296-424 materials,
// 428-490 meshes, 494-672 scene nodes); AssetCore and its file formats are not
// part of the reference, so the formats cannot be read here. glTF 2.0 carries
// the same model: a node tree with TRS transforms, meshes of triangle
// primitives with one material each, metallic-roughness materials with albedo
// / normal / roughness / metallic / emissive textures and an alpha mask,
// perspective cameras and (KHR_lights_punctual) directional / point / spot
// lights. The loader plays AssetCore's part (file -> resources) and hands
// every resource to a sink of add_* callbacks, ResourceManager's part;
// hh_scene_load_gltf binds the sink to an hh_scene.
//
// Mapping (glTF -> helios):
//   mesh -> one mesh (first use), primitive -> submesh with its own material,
//     indices absolute, position.w = submesh index (resource_manager.cpp:467-473)
//   node with a mesh -> one instance: model = global transform, normal matrix
//     = global transform with the scale removed from its columns
//   baseColorFactor (linear) -> albedo_value = factor^(1/2.2), which the
//     packer's pow(x, 2.2) (scene.cpp:1017) turns back into the factor
//   metallicRoughnessTexture -> roughness from channel 1 (G), metallic from
//     channel 2 (B); every texture use gets its own texture object, because
//     the packer binds a texture only at its first use in the scene
//     (scene.cpp:970-1109)
//   factors other than 1 are multiplied into a float copy of the texture;
//     alphaMode MASK thresholds alpha at alphaCutoff into 0 / 1 so the tracer's
//     fixed 0.1 test (path_trace_rahit.glsl:186) reproduces the cutoff;
//     BLEND -> transparent material (alpha-tested geometry)
//   emissiveFactor x KHR_materials_emissive_strength -> emissive_value
//   camera -> position / orientation of its node, yfov, znear, zfar (1000
//     when absent); focal length 8 and aperture 0.1 are CameraNode's defaults
//     (the reference loader never sets them, resource_manager.cpp:552-564)
//   KHR_lights_punctual -> lights with radius 0 (punctual), forward = the
//     node's -Z, cone angles in degrees
// Not read: samplers (the tracer samples bilinear with repeat at level 0),
// texture transforms, second UV sets, skins, morph targets, animations.
// Images: PNG (image_io.cpp); other encodings are rejected.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/helios_hip.h"

hh_status hh_internal_fail(hh_status s, const char* msg);
bool hh_internal_png_decode(const uint8_t* data, size_t n, std::vector<uint8_t>& rgba, uint32_t& width,
                            uint32_t& height, std::string& err);

namespace {

constexpr double kPi = 3.14159265358979323846;

// ---------------------------------------------------------------------------
// JSON (RFC 8259) into a small DOM
// ---------------------------------------------------------------------------
struct Json {
    enum Type { Null, Bool, Number, String, Array, Object } type = Null;
    bool b = false;
    double num = 0.0;
    std::string str;
    std::vector<Json> arr;
    std::vector<std::pair<std::string, Json>> obj;

    const Json* get(const char* k) const {
        if (type != Object) return nullptr;
        for (const auto& kv : obj)
            if (kv.first == k) return &kv.second;
        return nullptr;
    }
    double number(const char* k, double def) const {
        const Json* v = get(k);
        return v && v->type == Number ? v->num : def;
    }
    long integer(const char* k, long def) const {
        const Json* v = get(k);
        return v && v->type == Number ? v->as_long() : def;
    }
    // The number as a long, or -1 when it is not an integer in
    // [-2^53, 2^53] (every valid index, count and offset is; converting
    // a double outside long's range would be undefined).
    long as_long() const {
        if (type != Number || !(num >= -9007199254740992.0 && num <= 9007199254740992.0) || num != std::floor(num))
            return -1;
        return long(num);
    }
    std::string string(const char* k, const std::string& def = "") const {
        const Json* v = get(k);
        return v && v->type == String ? v->str : def;
    }
    size_t size() const { return type == Array ? arr.size() : 0; }
};

struct JsonParser {
    const char* p;
    const char* end;
    std::string err;
    int depth = 0;

    bool fail(const char* m) {
        if (err.empty()) err = m;
        return false;
    }
    void ws() {
        while (p < end && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r')) p++;
    }
    bool literal(const char* s) {
        const size_t n = std::strlen(s);
        if (size_t(end - p) < n || std::memcmp(p, s, n) != 0) return fail("bad literal");
        p += n;
        return true;
    }
    static void utf8(std::string& out, uint32_t c) {
        if (c < 0x80) out += char(c);
        else if (c < 0x800) { out += char(0xC0 | (c >> 6)); out += char(0x80 | (c & 0x3F)); }
        else if (c < 0x10000) { out += char(0xE0 | (c >> 12)); out += char(0x80 | ((c >> 6) & 0x3F)); out += char(0x80 | (c & 0x3F)); }
        else { out += char(0xF0 | (c >> 18)); out += char(0x80 | ((c >> 12) & 0x3F)); out += char(0x80 | ((c >> 6) & 0x3F)); out += char(0x80 | (c & 0x3F)); }
    }
    bool hex4(uint32_t& v) {
        if (end - p < 4) return fail("bad \\u escape");
        v = 0;
        for (int i = 0; i < 4; i++) {
            const char c = *p++;
            v <<= 4;
            if (c >= '0' && c <= '9') v |= uint32_t(c - '0');
            else if (c >= 'a' && c <= 'f') v |= uint32_t(c - 'a' + 10);
            else if (c >= 'A' && c <= 'F') v |= uint32_t(c - 'A' + 10);
            else return fail("bad \\u escape");
        }
        return true;
    }
    bool string(std::string& out) {
        p++;  // opening quote
        while (p < end && *p != '"') {
            if (*p == '\\') {
                if (++p >= end) break;
                const char e = *p++;
                switch (e) {
                    case '"': out += '"'; break;
                    case '\\': out += '\\'; break;
                    case '/': out += '/'; break;
                    case 'b': out += '\b'; break;
                    case 'f': out += '\f'; break;
                    case 'n': out += '\n'; break;
                    case 'r': out += '\r'; break;
                    case 't': out += '\t'; break;
                    case 'u': {
                        uint32_t c;
                        if (!hex4(c)) return false;
                        if (c >= 0xD800 && c < 0xDC00 && end - p >= 6 && p[0] == '\\' && p[1] == 'u') {
                            p += 2;
                            uint32_t lo;
                            if (!hex4(lo)) return false;
                            if (lo >= 0xDC00 && lo < 0xE000) c = 0x10000 + ((c - 0xD800) << 10) + (lo - 0xDC00);
                        }
                        utf8(out, c);
                        break;
                    }
                    default: return fail("bad escape");
                }
            } else {
                out += *p++;
            }
        }
        if (p >= end) return fail("unterminated string");
        p++;
        return true;
    }
    bool number(Json& out) {
        const char* s = p;
        if (p < end && *p == '-') p++;
        while (p < end && ((*p >= '0' && *p <= '9') || *p == '.' || *p == 'e' || *p == 'E' || *p == '+' || *p == '-')) p++;
        if (p == s) return fail("unexpected character");
        const std::string tok(s, p);
        char* q = nullptr;
        out.num = std::strtod(tok.c_str(), &q);
        if (!q || *q != '\0') return fail("bad number");
        out.type = Json::Number;
        return true;
    }
    bool value(Json& out) {
        ws();
        if (p >= end) return fail("unexpected end of JSON");
        if (++depth > 512) return fail("JSON nested too deeply");
        bool ok = true;
        switch (*p) {
            case '{': {
                out.type = Json::Object;
                p++;
                ws();
                if (p < end && *p == '}') { p++; break; }
                for (;;) {
                    ws();
                    if (p >= end || *p != '"') { ok = fail("expected a key"); break; }
                    std::string key;
                    if (!string(key)) { ok = false; break; }
                    ws();
                    if (p >= end || *p != ':') { ok = fail("expected ':'"); break; }
                    p++;
                    out.obj.emplace_back(std::move(key), Json());
                    if (!value(out.obj.back().second)) { ok = false; break; }
                    ws();
                    if (p < end && *p == ',') { p++; continue; }
                    if (p < end && *p == '}') { p++; break; }
                    ok = fail("expected ',' or '}'");
                    break;
                }
                break;
            }
            case '[': {
                out.type = Json::Array;
                p++;
                ws();
                if (p < end && *p == ']') { p++; break; }
                for (;;) {
                    out.arr.emplace_back();
                    if (!value(out.arr.back())) { ok = false; break; }
                    ws();
                    if (p < end && *p == ',') { p++; continue; }
                    if (p < end && *p == ']') { p++; break; }
                    ok = fail("expected ',' or ']'");
                    break;
                }
                break;
            }
            case '"': out.type = Json::String; ok = string(out.str); break;
            case 't': out.type = Json::Bool; out.b = true; ok = literal("true"); break;
            case 'f': out.type = Json::Bool; out.b = false; ok = literal("false"); break;
            case 'n': out.type = Json::Null; ok = literal("null"); break;
            default: ok = number(out);
        }
        depth--;
        return ok;
    }
};

// ---------------------------------------------------------------------------
// files and URIs
// ---------------------------------------------------------------------------
bool read_file(const std::string& path, std::vector<uint8_t>& out) {
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) return false;
    out.clear();
    uint8_t buf[1 << 16];
    size_t n;
    while ((n = std::fread(buf, 1, sizeof(buf), f)) > 0) out.insert(out.end(), buf, buf + n);
    const bool ok = !std::ferror(f);
    std::fclose(f);
    return ok;
}

bool base64_decode(const std::string& s, size_t start, std::vector<uint8_t>& out) {
    auto val = [](char c) -> int {
        if (c >= 'A' && c <= 'Z') return c - 'A';
        if (c >= 'a' && c <= 'z') return c - 'a' + 26;
        if (c >= '0' && c <= '9') return c - '0' + 52;
        if (c == '+' || c == '-') return 62;
        if (c == '/' || c == '_') return 63;
        return -1;
    };
    out.clear();
    uint32_t acc = 0;
    int nbits = 0;
    for (size_t i = start; i < s.size(); i++) {
        const char c = s[i];
        if (c == '=') break;
        const int v = val(c);
        if (v < 0) return false;
        acc = (acc << 6) | uint32_t(v);
        nbits += 6;
        if (nbits >= 8) {
            nbits -= 8;
            out.push_back(uint8_t(acc >> nbits));
        }
    }
    return true;
}

std::string percent_decode(const std::string& s) {
    std::string out;
    for (size_t i = 0; i < s.size(); i++) {
        if (s[i] == '%' && i + 2 < s.size()) {
            const std::string h = s.substr(i + 1, 2);
            char* q = nullptr;
            const long v = std::strtol(h.c_str(), &q, 16);
            if (q && *q == '\0') {
                out += char(v);
                i += 2;
                continue;
            }
        }
        out += s[i];
    }
    return out;
}

// A URI's bytes: data: URIs (base64) or a path relative to the file.
bool load_uri(const std::string& uri, const std::string& dir, std::vector<uint8_t>& out, std::string& err) {
    if (uri.compare(0, 5, "data:") == 0) {
        const size_t comma = uri.find(',');
        if (comma == std::string::npos || uri.rfind(";base64", comma) == std::string::npos) {
            err = "only base64 data URIs are supported";
            return false;
        }
        if (!base64_decode(uri, comma + 1, out)) {
            err = "bad base64 in a data URI";
            return false;
        }
        return true;
    }
    const std::string path = dir + percent_decode(uri);
    if (!read_file(path, out)) {
        err = "cannot read " + path;
        return false;
    }
    return true;
}

// ---------------------------------------------------------------------------
// math (double, column-major 4x4 like glm)
// ---------------------------------------------------------------------------
struct M4 {
    double m[16];
    static M4 identity() {
        M4 r{};
        r.m[0] = r.m[5] = r.m[10] = r.m[15] = 1.0;
        return r;
    }
    double& at(int row, int col) { return m[col * 4 + row]; }
    double at(int row, int col) const { return m[col * 4 + row]; }
};
M4 mul(const M4& a, const M4& b) {
    M4 r{};
    for (int c = 0; c < 4; c++)
        for (int rr = 0; rr < 4; rr++) {
            double s = 0.0;
            for (int k = 0; k < 4; k++) s += a.at(rr, k) * b.at(k, c);
            r.at(rr, c) = s;
        }
    return r;
}
M4 from_trs(const double t[3], const double q[4], const double s[3]) {
    const double x = q[0], y = q[1], z = q[2], w = q[3];
    const double R[3][3] = {{1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)},
                            {2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)},
                            {2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)}};
    M4 r = M4::identity();
    for (int row = 0; row < 3; row++) {
        for (int col = 0; col < 3; col++) r.at(row, col) = R[row][col] * s[col];
        r.at(row, 3) = t[row];
    }
    return r;
}
// the global transform with each of the three axis columns normalised
M4 without_scale(const M4& g) {
    M4 r = g;
    for (int c = 0; c < 3; c++) {
        const double n = std::sqrt(g.at(0, c) * g.at(0, c) + g.at(1, c) * g.at(1, c) + g.at(2, c) * g.at(2, c));
        if (n > 0.0)
            for (int k = 0; k < 3; k++) r.at(k, c) = g.at(k, c) / n;
    }
    return r;
}
void quat_from_rotation(const M4& r, double q[4]) {  // (x, y, z, w)
    const double t = r.at(0, 0) + r.at(1, 1) + r.at(2, 2);
    if (t > 0.0) {
        const double s = std::sqrt(t + 1.0) * 2.0;
        q[3] = 0.25 * s;
        q[0] = (r.at(2, 1) - r.at(1, 2)) / s;
        q[1] = (r.at(0, 2) - r.at(2, 0)) / s;
        q[2] = (r.at(1, 0) - r.at(0, 1)) / s;
    } else if (r.at(0, 0) > r.at(1, 1) && r.at(0, 0) > r.at(2, 2)) {
        const double s = std::sqrt(1.0 + r.at(0, 0) - r.at(1, 1) - r.at(2, 2)) * 2.0;
        q[3] = (r.at(2, 1) - r.at(1, 2)) / s;
        q[0] = 0.25 * s;
        q[1] = (r.at(0, 1) + r.at(1, 0)) / s;
        q[2] = (r.at(0, 2) + r.at(2, 0)) / s;
    } else if (r.at(1, 1) > r.at(2, 2)) {
        const double s = std::sqrt(1.0 + r.at(1, 1) - r.at(0, 0) - r.at(2, 2)) * 2.0;
        q[3] = (r.at(0, 2) - r.at(2, 0)) / s;
        q[0] = (r.at(0, 1) + r.at(1, 0)) / s;
        q[1] = 0.25 * s;
        q[2] = (r.at(1, 2) + r.at(2, 1)) / s;
    } else {
        const double s = std::sqrt(1.0 + r.at(2, 2) - r.at(0, 0) - r.at(1, 1)) * 2.0;
        q[3] = (r.at(1, 0) - r.at(0, 1)) / s;
        q[0] = (r.at(0, 2) + r.at(2, 0)) / s;
        q[1] = (r.at(1, 2) + r.at(2, 1)) / s;
        q[2] = 0.25 * s;
    }
    const double n = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
    for (int k = 0; k < 4; k++) q[k] /= n;
}

double srgb_to_linear(double c) { return c <= 0.04045 ? c / 12.92 : std::pow((c + 0.055) / 1.055, 2.4); }

struct V3 {
    double x, y, z;
};
V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
V3 scale(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
V3 normalized(V3 a) {
    const double n = std::sqrt(dot(a, a));
    return n > 0.0 ? scale(a, 1.0 / n) : a;
}

// ---------------------------------------------------------------------------
// the loader
// ---------------------------------------------------------------------------
struct Loader {
    const hh_scene_sink* sink;
    std::string dir;
    Json doc;
    std::vector<std::vector<uint8_t>> buffers;
    std::string err;
    hh_gltf_info info{};
    hh_status status = HH_ERR_IO;
    std::map<long, uint32_t> mesh_ids;       // glTF mesh -> helios mesh
    std::map<long, uint32_t> material_ids;   // glTF material -> helios material
    std::map<long, std::vector<uint8_t>> images;
    std::map<long, std::pair<uint32_t, uint32_t>> image_dims;
    long default_material = -1;
    uint32_t default_material_id = 0;

    bool fail(const std::string& m, hh_status s = HH_ERR_IO) {
        if (err.empty()) {
            err = m;
            status = s;
        }
        return false;
    }
    bool sink_ok(hh_status s, const char* what) {
        if (s == HH_OK) return true;
        const char* why = hh_last_error();
        return fail(std::string(what) + " failed: " + (why ? why : ""), s);
    }
    const Json* array(const char* k) const {
        const Json* a = doc.get(k);
        return a && a->type == Json::Array ? a : nullptr;
    }
    const Json* element(const char* k, long i) const {
        const Json* a = array(k);
        if (!a || i < 0 || size_t(i) >= a->arr.size()) return nullptr;
        return &a->arr[size_t(i)];
    }

    bool parse(const std::string& path) {
        const size_t slash = path.find_last_of('/');
        dir = slash == std::string::npos ? "" : path.substr(0, slash + 1);
        std::vector<uint8_t> file;
        if (!read_file(path, file)) return fail("cannot read " + path);
        std::vector<uint8_t> bin;
        bool glb = false;
        std::string text;
        if (file.size() >= 12 && std::memcmp(file.data(), "glTF", 4) == 0) {
            glb = true;
            uint32_t version, length;
            std::memcpy(&version, file.data() + 4, 4);
            std::memcpy(&length, file.data() + 8, 4);
            if (version != 2) return fail("only glTF 2.0 binaries are supported");
            if (length > file.size()) return fail("truncated GLB");
            size_t pos = 12;
            while (pos + 8 <= length) {
                uint32_t clen, ctype;
                std::memcpy(&clen, file.data() + pos, 4);
                std::memcpy(&ctype, file.data() + pos + 4, 4);
                if (clen > length - pos - 8) return fail("truncated GLB chunk");
                const uint8_t* body = file.data() + pos + 8;
                if (ctype == 0x4E4F534Au && text.empty()) text.assign((const char*)body, clen);           // JSON
                else if (ctype == 0x004E4942u && bin.empty()) bin.assign(body, body + clen);                // BIN
                pos += 8 + ((size_t(clen) + 3) & ~size_t(3));
            }
            if (text.empty()) return fail("GLB without a JSON chunk");
        } else {
            text.assign((const char*)file.data(), file.size());
        }
        JsonParser jp{text.data(), text.data() + text.size()};
        if (!jp.value(doc)) return fail("JSON: " + jp.err);
        jp.ws();
        if (jp.p != jp.end) return fail("JSON: trailing characters");
        if (doc.type != Json::Object) return fail("JSON: the document is not an object");
        const Json* asset = doc.get("asset");
        if (!asset || asset->string("version").compare(0, 1, "2") != 0) return fail("not a glTF 2.0 asset");
        if (const Json* req = doc.get("extensionsRequired")) {
            for (const Json& e : req->arr) {
                if (e.str != "KHR_lights_punctual" && e.str != "KHR_materials_emissive_strength")
                    return fail("required extension not supported: " + e.str);
            }
        }
        if (const Json* bufs = array("buffers")) {
            for (size_t i = 0; i < bufs->arr.size(); i++) {
                const Json& b = bufs->arr[i];
                std::vector<uint8_t> data;
                if (b.get("uri")) {
                    std::string e;
                    if (!load_uri(b.string("uri"), dir, data, e)) return fail(e);
                } else if (glb && i == 0) {
                    data = bin;
                } else {
                    return fail("buffer without a uri");
                }
                const long len = b.integer("byteLength", -1);
                if (len < 0 || size_t(len) > data.size()) return fail("buffer shorter than its byteLength");
                data.resize(size_t(len));
                buffers.push_back(std::move(data));
            }
        }
        return true;
    }

    // bytes of a bufferView
    bool view(long bv, const uint8_t*& ptr, size_t& len, size_t& stride) {
        const Json* v = element("bufferViews", bv);
        if (!v) return fail("bad bufferView index");
        const long b = v->integer("buffer", -1);
        if (b < 0 || size_t(b) >= buffers.size()) return fail("bad buffer index");
        const long off = v->integer("byteOffset", 0), n = v->integer("byteLength", -1);
        const size_t have = buffers[size_t(b)].size();
        if (off < 0 || n < 0 || size_t(off) > have || size_t(n) > have - size_t(off)) return fail("bufferView out of range");
        // glTF 2.0: byteStride, when present, is in [4, 252] and a multiple of 4
        const long st = v->integer("byteStride", 0);
        if (v->get("byteStride") && (st < 4 || st > 252 || st % 4 != 0)) return fail("bad bufferView byteStride");
        ptr = buffers[size_t(b)].data() + off;
        len = size_t(n);
        stride = size_t(st);
        return true;
    }

    // an accessor as `comps` doubles per element (normalised integers scaled)
    bool read_accessor(long idx, int comps, std::vector<double>& out, bool as_index = false) {
        const Json* a = element("accessors", idx);
        if (!a) return fail("bad accessor index");
        if (a->get("sparse")) return fail("sparse accessors are not supported");
        static const std::map<std::string, int> kTypes = {{"SCALAR", 1}, {"VEC2", 2}, {"VEC3", 3}, {"VEC4", 4}};
        const auto ti = kTypes.find(a->string("type"));
        if (ti == kTypes.end() || ti->second != comps) return fail("unexpected accessor type");
        const long count = a->integer("count", -1);
        if (count < 0 || count > (1l << 30)) return fail("bad accessor count");
        const long ctype = a->integer("componentType", 0);
        const bool norm = a->get("normalized") && a->get("normalized")->b;
        int csize;
        switch (ctype) {
            case 5120: case 5121: csize = 1; break;
            case 5122: case 5123: csize = 2; break;
            case 5125: case 5126: csize = 4; break;
            default: return fail("bad componentType");
        }
        if (as_index && ctype != 5121 && ctype != 5123 && ctype != 5125) return fail("bad index componentType");
        out.assign(size_t(count) * comps, 0.0);
        if (!a->get("bufferView")) return true;  // all zeros
        const uint8_t* p;
        size_t len, stride;
        if (!view(a->integer("bufferView", -1), p, len, stride)) return false;
        const size_t elem = size_t(csize) * comps;
        if (stride == 0) stride = elem;
        if (stride < elem) return fail("bufferView byteStride below the element size");
        const long aoff = a->integer("byteOffset", 0);
        if (aoff < 0 || aoff % csize != 0) return fail("bad accessor byteOffset");
        const size_t off = size_t(aoff);
        // off + stride * (count - 1) + elem <= len, without overflow
        if (count > 0 && (off > len || elem > len - off || size_t(count - 1) > (len - off - elem) / stride))
            return fail("accessor out of range");
        for (long i = 0; i < count; i++) {
            const uint8_t* e = p + off + stride * size_t(i);
            for (int c = 0; c < comps; c++) {
                const uint8_t* q = e + size_t(c) * csize;
                double v = 0.0;
                switch (ctype) {
                    case 5120: { int8_t x; std::memcpy(&x, q, 1); v = norm ? std::max(x / 127.0, -1.0) : x; break; }
                    case 5121: { uint8_t x = *q; v = norm ? x / 255.0 : x; break; }
                    case 5122: { int16_t x; std::memcpy(&x, q, 2); v = norm ? std::max(x / 32767.0, -1.0) : x; break; }
                    case 5123: { uint16_t x; std::memcpy(&x, q, 2); v = norm ? x / 65535.0 : x; break; }
                    case 5125: { uint32_t x; std::memcpy(&x, q, 4); v = x; break; }
                    default: { float x; std::memcpy(&x, q, 4); v = x; break; }
                }
                out[size_t(i) * comps + c] = v;
            }
        }
        return true;
    }

    // an image as top-down RGBA8 (row 0 = v 0, glTF's convention)
    bool image(long idx, const std::vector<uint8_t>*& rgba, uint32_t& w, uint32_t& h) {
        auto it = images.find(idx);
        if (it == images.end()) {
            const Json* im = element("images", idx);
            if (!im) return fail("bad image index");
            std::vector<uint8_t> bytes;
            if (im->get("uri")) {
                std::string e;
                if (!load_uri(im->string("uri"), dir, bytes, e)) return fail(e);
            } else {
                const uint8_t* p;
                size_t len, stride;
                if (!view(im->integer("bufferView", -1), p, len, stride)) return false;
                bytes.assign(p, p + len);
            }
            std::vector<uint8_t> px;
            std::string e;
            uint32_t iw = 0, ih = 0;
            if (bytes.size() >= 2 && bytes[0] == 0xFF && bytes[1] == 0xD8)
                return fail("image " + std::to_string(idx) + " is JPEG; only PNG textures are read");
            if (!hh_internal_png_decode(bytes.data(), bytes.size(), px, iw, ih, e))
                return fail("image " + std::to_string(idx) + ": " + e);
            image_dims[idx] = {iw, ih};
            it = images.emplace(idx, std::move(px)).first;
        }
        rgba = &it->second;
        w = image_dims[idx].first;
        h = image_dims[idx].second;
        return true;
    }

    // One texture object for one use. srgb: rgb is sRGB-encoded. factor
    // multiplies the decoded channels; cutoff >= 0 thresholds alpha.
    bool texture(const Json& ref, bool srgb, const double factor[4], double cutoff, int32_t& out) {
        const long ti = ref.integer("index", -1);
        if (ref.integer("texCoord", 0) != 0) info.warnings |= HH_GLTF_WARN_TEXCOORD1;
        const Json* t = element("textures", ti);
        if (!t) return fail("bad texture index");
        const long src = t->integer("source", -1);
        if (src < 0) return fail("texture without a PNG source");
        const std::vector<uint8_t>* px;
        uint32_t w, h;
        if (!image(src, px, w, h)) return false;
        const bool plain = factor[0] == 1.0 && factor[1] == 1.0 && factor[2] == 1.0 && factor[3] == 1.0 && cutoff < 0.0;
        uint32_t id = 0;
        if (plain) {
            if (!sink_ok(sink->add_texture(sink->user, w, h, srgb ? HH_TEXTURE_RGBA8_SRGB : HH_TEXTURE_RGBA8_UNORM,
                                           px->data(), &id), "add_texture"))
                return false;
        } else {
            std::vector<float> f(size_t(w) * h * 4);
            for (size_t i = 0; i < size_t(w) * h; i++) {
                for (int c = 0; c < 4; c++) {
                    double v = (*px)[4 * i + c] / 255.0;
                    if (srgb && c < 3) v = srgb_to_linear(v);
                    v *= factor[c];
                    if (c == 3 && cutoff >= 0.0) v = v >= cutoff ? 1.0 : 0.0;
                    f[4 * i + c] = float(v);
                }
            }
            if (!sink_ok(sink->add_texture(sink->user, w, h, HH_TEXTURE_RGBA32F, f.data(), &id), "add_texture"))
                return false;
        }
        info.num_textures++;
        out = int32_t(id);
        return true;
    }

    static void factors(const Json* a, double out[], int n, double def) {
        for (int i = 0; i < n; i++) out[i] = def;
        if (a && a->type == Json::Array)
            for (int i = 0; i < n && size_t(i) < a->arr.size(); i++) out[i] = a->arr[size_t(i)].num;
    }

    bool material(long idx, uint32_t& out) {
        if (idx < 0) {
            if (default_material < 0) {  // glTF's default material: white, metallic 1, roughness 1
                hh_material_desc d{};
                for (int k = 0; k < 4; k++) d.albedo_value[k] = 1.0f;
                d.metallic_value = 1.0f;
                d.roughness_value = 1.0f;
                d.albedo_texture = d.normal_texture = d.roughness_texture = d.metallic_texture = d.emissive_texture = -1;
                d.roughness_channel = d.metallic_channel = -1;
                if (!sink_ok(sink->add_material(sink->user, &d, &default_material_id), "add_material")) return false;
                info.num_materials++;
                default_material = 1;
            }
            out = default_material_id;
            return true;
        }
        auto it = material_ids.find(idx);
        if (it != material_ids.end()) {
            out = it->second;
            return true;
        }
        const Json* m = element("materials", idx);
        if (!m) return fail("bad material index");
        static const Json kEmpty;
        const Json* pbr = m->get("pbrMetallicRoughness");
        if (!pbr) pbr = &kEmpty;
        hh_material_desc d{};
        d.albedo_texture = d.normal_texture = d.roughness_texture = d.metallic_texture = d.emissive_texture = -1;
        d.roughness_channel = d.metallic_channel = -1;
        double base[4];
        factors(pbr->get("baseColorFactor"), base, 4, 1.0);
        const std::string mode = m->string("alphaMode", "OPAQUE");
        const double cutoff = m->number("alphaCutoff", 0.5);
        const bool mask = mode == "MASK", blend = mode == "BLEND";
        d.type = blend ? HH_MATERIAL_TRANSPARENT : HH_MATERIAL_OPAQUE;
        d.alpha_test = (mask || blend) ? 1u : 0u;
        for (int k = 0; k < 3; k++) d.albedo_value[k] = float(std::pow(std::max(base[k], 0.0), 1.0 / 2.2));
        d.albedo_value[3] = float(mask ? (base[3] >= cutoff ? 1.0 : 0.0) : base[3]);
        if (const Json* bt = pbr->get("baseColorTexture"))
            if (!texture(*bt, true, base, mask ? cutoff : -1.0, d.albedo_texture)) return false;
        d.metallic_value = float(pbr->number("metallicFactor", 1.0));
        d.roughness_value = float(pbr->number("roughnessFactor", 1.0));
        if (const Json* mr = pbr->get("metallicRoughnessTexture")) {
            const double f[4] = {1.0, d.roughness_value, d.metallic_value, 1.0};
            if (!texture(*mr, false, f, -1.0, d.roughness_texture)) return false;
            if (!texture(*mr, false, f, -1.0, d.metallic_texture)) return false;
            d.roughness_channel = 1;
            d.metallic_channel = 2;
        }
        if (const Json* nt = m->get("normalTexture")) {
            const double one[4] = {1.0, 1.0, 1.0, 1.0};
            if (!texture(*nt, false, one, -1.0, d.normal_texture)) return false;
        }
        double em[3];
        factors(m->get("emissiveFactor"), em, 3, 0.0);
        double strength = 1.0;
        if (const Json* ext = m->get("extensions"))
            if (const Json* es = ext->get("KHR_materials_emissive_strength")) strength = es->number("emissiveStrength", 1.0);
        for (int k = 0; k < 3; k++) d.emissive_value[k] = float(em[k] * strength);
        d.emissive_value[3] = 0.0f;
        if (const Json* et = m->get("emissiveTexture")) {
            if (em[0] != 0.0 || em[1] != 0.0 || em[2] != 0.0) {
                const double f[4] = {em[0] * strength, em[1] * strength, em[2] * strength, 1.0};
                if (!texture(*et, true, f, -1.0, d.emissive_texture)) return false;
            }
        }
        uint32_t id;
        if (!sink_ok(sink->add_material(sink->user, &d, &id), "add_material")) return false;
        info.num_materials++;
        material_ids[idx] = id;
        out = id;
        return true;
    }

    // per-vertex tangent frames: given, from UV gradients, or any orthonormal pair
    static void tangents(std::vector<hh_vertex>& v, const std::vector<uint32_t>& idx, size_t base, bool have_uv) {
        const size_t n = v.size() - base;
        std::vector<V3> tacc(n, V3{0, 0, 0}), bacc(n, V3{0, 0, 0});
        if (have_uv) {
            for (size_t t = 0; t + 2 < idx.size(); t += 3) {
                const size_t a = idx[t] - base, b = idx[t + 1] - base, c = idx[t + 2] - base;
                const hh_vertex &A = v[base + a], &B = v[base + b], &Cv = v[base + c];
                const V3 e1 = {B.position[0] - A.position[0], B.position[1] - A.position[1], B.position[2] - A.position[2]};
                const V3 e2 = {Cv.position[0] - A.position[0], Cv.position[1] - A.position[1], Cv.position[2] - A.position[2]};
                const double du1 = B.tex_coord[0] - A.tex_coord[0], dv1 = B.tex_coord[1] - A.tex_coord[1];
                const double du2 = Cv.tex_coord[0] - A.tex_coord[0], dv2 = Cv.tex_coord[1] - A.tex_coord[1];
                const double det = du1 * dv2 - du2 * dv1;
                if (std::fabs(det) < 1e-20) continue;
                const V3 T = scale(sub(scale(e1, dv2), scale(e2, dv1)), 1.0 / det);
                const V3 Bt = scale(sub(scale(e2, du1), scale(e1, du2)), 1.0 / det);
                for (size_t k : {a, b, c}) {
                    tacc[k] = add(tacc[k], T);
                    bacc[k] = add(bacc[k], Bt);
                }
            }
        }
        for (size_t i = 0; i < n; i++) {
            hh_vertex& x = v[base + i];
            const V3 N = {x.normal[0], x.normal[1], x.normal[2]};
            V3 T = sub(tacc[i], scale(N, dot(N, tacc[i])));
            if (dot(T, T) < 1e-24) {
                const V3 a = std::fabs(N.y) < 0.99 ? V3{0, 1, 0} : V3{1, 0, 0};
                T = cross(a, N);
            }
            T = normalized(T);
            V3 B = cross(N, T);
            if (dot(B, bacc[i]) < 0.0) B = scale(B, -1.0);
            x.tangent[0] = float(T.x); x.tangent[1] = float(T.y); x.tangent[2] = float(T.z);
            x.bitangent[0] = float(B.x); x.bitangent[1] = float(B.y); x.bitangent[2] = float(B.z);
        }
    }

    bool mesh(long idx, uint32_t& out) {
        auto it = mesh_ids.find(idx);
        if (it != mesh_ids.end()) {
            out = it->second;
            return true;
        }
        const Json* m = element("meshes", idx);
        if (!m) return fail("bad mesh index");
        const Json* prims = m->get("primitives");
        if (!prims || prims->type != Json::Array) return fail("mesh without primitives");
        std::vector<hh_vertex> verts;
        std::vector<uint32_t> indices;
        std::vector<hh_submesh> subs;
        std::vector<uint32_t> mats;
        for (const Json& p : prims->arr) {
            if (p.integer("mode", 4) != 4) {
                info.warnings |= HH_GLTF_WARN_NON_TRIANGLES;
                continue;
            }
            const Json* attr = p.get("attributes");
            if (!attr || !attr->get("POSITION")) return fail("primitive without POSITION");
            std::vector<double> P, N, UV, TG, IX;
            if (!read_accessor(attr->integer("POSITION", -1), 3, P)) return false;
            const size_t nv = P.size() / 3;
            const bool have_n = attr->get("NORMAL") != nullptr;
            const bool have_uv = attr->get("TEXCOORD_0") != nullptr;
            const bool have_t = attr->get("TANGENT") != nullptr;
            if (have_n && !read_accessor(attr->integer("NORMAL", -1), 3, N)) return false;
            if (have_uv && !read_accessor(attr->integer("TEXCOORD_0", -1), 2, UV)) return false;
            if (have_t && !read_accessor(attr->integer("TANGENT", -1), 4, TG)) return false;
            if ((have_n && N.size() != 3 * nv) || (have_uv && UV.size() != 2 * nv) || (have_t && TG.size() != 4 * nv))
                return fail("attribute counts differ");
            std::vector<uint32_t> local;
            if (p.get("indices")) {
                if (!read_accessor(p.integer("indices", -1), 1, IX, true)) return false;
                local.resize(IX.size());
                for (size_t i = 0; i < IX.size(); i++) {
                    if (IX[i] >= double(nv)) return fail("index out of range");
                    local[i] = uint32_t(IX[i]);
                }
            } else {
                local.resize(nv);
                for (size_t i = 0; i < nv; i++) local[i] = uint32_t(i);
            }
            local.resize(local.size() / 3 * 3);
            if (local.empty()) continue;
            const uint32_t sub_i = uint32_t(subs.size());
            const size_t base = verts.size(), ibase = indices.size();
            auto make = [&](size_t i) {
                hh_vertex x{};
                for (int k = 0; k < 3; k++) x.position[k] = float(P[3 * i + k]);
                x.position[3] = float(sub_i);
                if (have_uv) { x.tex_coord[0] = float(UV[2 * i]); x.tex_coord[1] = float(UV[2 * i + 1]); }
                if (have_n) for (int k = 0; k < 3; k++) x.normal[k] = float(N[3 * i + k]);
                return x;
            };
            if (have_n) {
                for (size_t i = 0; i < nv; i++) verts.push_back(make(i));
                for (uint32_t i : local) indices.push_back(uint32_t(base) + i);
            } else {
                // no normals: flat shading (glTF 2.0, 3.7.2.1), one vertex per corner
                for (size_t t = 0; t < local.size(); t += 3) {
                    hh_vertex c[3] = {make(local[t]), make(local[t + 1]), make(local[t + 2])};
                    const V3 a = {c[0].position[0], c[0].position[1], c[0].position[2]};
                    const V3 b = {c[1].position[0], c[1].position[1], c[1].position[2]};
                    const V3 d = {c[2].position[0], c[2].position[1], c[2].position[2]};
                    const V3 fn = normalized(cross(sub(b, a), sub(d, a)));
                    for (int k = 0; k < 3; k++) {
                        c[k].normal[0] = float(fn.x); c[k].normal[1] = float(fn.y); c[k].normal[2] = float(fn.z);
                        indices.push_back(uint32_t(verts.size()));
                        verts.push_back(c[k]);
                    }
                }
            }
            const std::vector<uint32_t> sub_idx(indices.begin() + long(ibase), indices.end());
            if (have_t && have_n) {
                for (size_t i = 0; i < nv; i++) {
                    hh_vertex& x = verts[base + i];
                    const V3 Nn = {x.normal[0], x.normal[1], x.normal[2]};
                    const V3 T = {TG[4 * i], TG[4 * i + 1], TG[4 * i + 2]};
                    const V3 B = scale(cross(Nn, T), TG[4 * i + 3] < 0.0 ? -1.0 : 1.0);
                    x.tangent[0] = float(T.x); x.tangent[1] = float(T.y); x.tangent[2] = float(T.z);
                    x.bitangent[0] = float(B.x); x.bitangent[1] = float(B.y); x.bitangent[2] = float(B.z);
                }
            } else {
                tangents(verts, sub_idx, base, have_uv);
            }
            uint32_t mid;
            if (!material(p.integer("material", -1), mid)) return false;
            hh_submesh s{};
            s.mat_idx = uint32_t(mats.size());
            s.index_count = uint32_t(indices.size() - ibase);
            s.vertex_count = uint32_t(verts.size() - base);
            s.base_vertex = uint32_t(base);
            s.base_index = uint32_t(ibase);
            subs.push_back(s);
            mats.push_back(mid);
        }
        if (subs.empty()) return fail("mesh " + std::to_string(idx) + " has no triangle primitives");
        uint32_t id;
        if (!sink_ok(sink->add_mesh(sink->user, verts.data(), uint32_t(verts.size()), indices.data(), uint32_t(indices.size()),
                                    subs.data(), uint32_t(subs.size()), mats.data(), uint32_t(mats.size()), &id),
                     "add_mesh"))
            return false;
        info.num_meshes++;
        info.num_triangles += indices.size() / 3;
        mesh_ids[idx] = id;
        out = id;
        return true;
    }

    M4 local_transform(const Json& n) {
        if (const Json* mat = n.get("matrix")) {
            M4 r = M4::identity();
            for (int i = 0; i < 16 && size_t(i) < mat->size(); i++) r.m[i] = mat->arr[size_t(i)].num;
            return r;
        }
        double t[3], q[4], s[3];
        factors(n.get("translation"), t, 3, 0.0);
        factors(n.get("scale"), s, 3, 1.0);
        q[0] = q[1] = q[2] = 0.0;
        q[3] = 1.0;
        if (const Json* r = n.get("rotation"))
            for (int i = 0; i < 4 && size_t(i) < r->size(); i++) q[i] = r->arr[size_t(i)].num;
        return from_trs(t, q, s);
    }

    bool node(long idx, const M4& parent, int depth) {
        if (depth > 128) return fail("node hierarchy too deep (a cycle?)");
        const Json* n = element("nodes", idx);
        if (!n) return fail("bad node index");
        const M4 g = mul(parent, local_transform(*n));
        const M4 r = without_scale(g);
        if (n->get("mesh")) {
            uint32_t mid;
            if (!mesh(n->integer("mesh", -1), mid)) return false;
            hh_instance_desc d{};
            for (int i = 0; i < 16; i++) {
                d.model[i] = float(g.m[i]);
                d.normal[i] = float(r.m[i]);
            }
            d.mesh = mid;
            d.material_override = -1;
            uint32_t id;
            if (!sink_ok(sink->add_instance(sink->user, &d, &id), "add_instance")) return false;
            info.num_instances++;
        }
        if (n->get("camera") && !info.has_camera) {
            const Json* c = element("cameras", n->integer("camera", -1));
            if (!c) return fail("bad camera index");
            if (c->string("type") == "perspective" && c->get("perspective")) {
                const Json* p = c->get("perspective");
                hh_camera& cam = info.camera;
                for (int k = 0; k < 3; k++) cam.position[k] = float(g.at(k, 3));
                double q[4];
                quat_from_rotation(r, q);
                for (int k = 0; k < 4; k++) cam.orientation[k] = float(q[k]);
                cam.fov_y_deg = float(p->number("yfov", 0.8) * 180.0 / kPi);
                cam.near_plane = float(p->number("znear", 0.1));
                cam.far_plane = float(p->number("zfar", 1000.0));
                cam.focal_length = 8.0f;
                cam.aperture_radius = 0.1f;
                info.has_camera = 1;
            }
            info.num_cameras++;
        }
        if (const Json* ext = n->get("extensions")) {
            if (const Json* lp = ext->get("KHR_lights_punctual")) {
                const Json* lights = doc.get("extensions") ? doc.get("extensions")->get("KHR_lights_punctual") : nullptr;
                const Json* list = lights ? lights->get("lights") : nullptr;
                const long li = lp->integer("light", -1);
                if (!list || li < 0 || size_t(li) >= list->size()) return fail("bad light index");
                const Json& L = list->arr[size_t(li)];
                hh_light_desc d{};
                const std::string type = L.string("type");
                if (type == "directional") d.type = HH_LIGHT_DIRECTIONAL;
                else if (type == "point") d.type = HH_LIGHT_POINT;
                else if (type == "spot") d.type = HH_LIGHT_SPOT;
                else return fail("unknown light type " + type);
                double col[3];
                factors(L.get("color"), col, 3, 1.0);
                for (int k = 0; k < 3; k++) d.color[k] = float(col[k]);
                d.intensity = float(L.number("intensity", 1.0));
                d.radius = 0.0f;
                for (int k = 0; k < 3; k++) d.position[k] = float(g.at(k, 3));
                const V3 fwd = normalized(V3{-r.at(0, 2), -r.at(1, 2), -r.at(2, 2)});
                d.forward[0] = float(fwd.x);
                d.forward[1] = float(fwd.y);
                d.forward[2] = float(fwd.z);
                const Json* spot = L.get("spot");
                const double inner = spot ? spot->number("innerConeAngle", 0.0) : 0.0;
                const double outer = spot ? spot->number("outerConeAngle", kPi / 4.0) : kPi / 4.0;
                d.inner_cone_deg = float(inner * 180.0 / kPi);
                d.outer_cone_deg = float(outer * 180.0 / kPi);
                uint32_t id;
                if (!sink_ok(sink->add_light(sink->user, &d, &id), "add_light")) return false;
                info.num_lights++;
            }
        }
        if (const Json* ch = n->get("children"))
            for (const Json& c : ch->arr)
                if (!node(c.as_long(), g, depth + 1)) return false;
        return true;
    }

    bool run() {
        std::vector<long> roots;
        const Json* scenes = array("scenes");
        if (scenes && !scenes->arr.empty()) {
            const long si = doc.integer("scene", 0);
            const Json* sc = element("scenes", si);
            if (!sc) return fail("bad scene index");
            if (const Json* ns = sc->get("nodes"))
                for (const Json& x : ns->arr) roots.push_back(x.as_long());
        } else if (const Json* nodes = array("nodes")) {
            std::vector<char> child(nodes->arr.size(), 0);
            for (const Json& n : nodes->arr)
                if (const Json* ch = n.get("children"))
                    for (const Json& c : ch->arr)
                        if (c.as_long() >= 0 && size_t(c.as_long()) < child.size()) child[size_t(c.as_long())] = 1;
            for (size_t i = 0; i < child.size(); i++)
                if (!child[i]) roots.push_back(long(i));
        }
        for (long r : roots)
            if (!node(r, M4::identity(), 0)) return false;
        return true;
    }
};

struct SceneSink {
    hh_scene scene;
};
hh_status s_texture(void* u, uint32_t w, uint32_t h, uint32_t fmt, const void* texels, uint32_t* id) {
    return hh_scene_add_texture(((SceneSink*)u)->scene, w, h, fmt, texels, id);
}
hh_status s_material(void* u, const hh_material_desc* d, uint32_t* id) {
    return hh_scene_add_material(((SceneSink*)u)->scene, d, id);
}
hh_status s_mesh(void* u, const hh_vertex* v, uint32_t nv, const uint32_t* ix, uint32_t ni, const hh_submesh* sm,
                 uint32_t ns, const uint32_t* mats, uint32_t nm, uint32_t* id) {
    return hh_scene_add_mesh(((SceneSink*)u)->scene, v, nv, ix, ni, sm, ns, mats, nm, id);
}
hh_status s_instance(void* u, const hh_instance_desc* d, uint32_t* id) {
    return hh_scene_add_instance(((SceneSink*)u)->scene, d, id);
}
hh_status s_light(void* u, const hh_light_desc* d, uint32_t* id) {
    return hh_scene_add_light(((SceneSink*)u)->scene, d, id);
}

}  // namespace

extern "C" hh_status hh_gltf_load(const char* path, const hh_scene_sink* sink, hh_gltf_info* info) {
    if (!path || !sink || !sink->add_texture || !sink->add_material || !sink->add_mesh || !sink->add_instance ||
        !sink->add_light)
        return hh_internal_fail(HH_ERR_INVALID_ARG, "bad argument");
    Loader ld;
    ld.sink = sink;
    const bool ok = ld.parse(path) && ld.run();
    if (info) *info = ld.info;
    if (!ok) return hh_internal_fail(ld.status, ("glTF " + std::string(path) + ": " + ld.err).c_str());
    return HH_OK;
}

extern "C" hh_status hh_scene_load_gltf(hh_scene scene, const char* path, hh_gltf_info* info) {
    if (!scene || !path) return hh_internal_fail(HH_ERR_INVALID_ARG, "bad argument");
    SceneSink s{scene};
    const hh_scene_sink sink{&s, s_texture, s_material, s_mesh, s_instance, s_light};
    return hh_gltf_load(path, &sink, info);
}
