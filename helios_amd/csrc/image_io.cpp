// PNG output of tone-mapped surfaces (the role stb_image_write plays for
// Renderer::copy_and_save_tone_mapped_image, renderer.cpp:637-713).
//
// Host-only code, built with g++ and linked with the system zlib. The file is
// an 8-bit RGBA, non-interlaced PNG (ISO/IEC 15948): signature, IHDR, one
// IDAT holding the zlib stream of the filtered rows, IEND. Each row takes the
// filter (None, Sub, Up, Average, Paeth) whose output has the smallest sum of
// absolute values as signed bytes, the usual encoder heuristic; any choice
// decodes to the same pixels.
#include <zlib.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/helios_hip.h"

hh_status hh_internal_fail(hh_status s, const char* msg);

namespace {

void put_u32(std::vector<uint8_t>& out, uint32_t v) {
    out.push_back(uint8_t(v >> 24));
    out.push_back(uint8_t(v >> 16));
    out.push_back(uint8_t(v >> 8));
    out.push_back(uint8_t(v));
}

void put_chunk(std::vector<uint8_t>& out, const char type[4], const uint8_t* data, size_t n) {
    put_u32(out, uint32_t(n));
    const size_t start = out.size();
    out.insert(out.end(), type, type + 4);
    if (n) out.insert(out.end(), data, data + n);
    uLong crc = crc32(0L, Z_NULL, 0);
    crc = crc32(crc, out.data() + start, uInt(4 + n));
    put_u32(out, uint32_t(crc));
}

inline uint8_t paeth(int a, int b, int c) {
    const int p = a + b - c;
    const int pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
    if (pa <= pb && pa <= pc) return uint8_t(a);
    if (pb <= pc) return uint8_t(b);
    return uint8_t(c);
}

// Filters one row of `stride` bytes (4 bytes per pixel) against the previous
// row (nullptr for the first) into out[0..stride] (out[0] = filter type).
void filter_row(const uint8_t* cur, const uint8_t* prev, size_t stride, uint8_t* out, uint8_t* scratch) {
    uint64_t best_cost = ~0ull;
    for (int f = 0; f < 5; f++) {
        uint64_t cost = 0;
        for (size_t i = 0; i < stride; i++) {
            const int a = i >= 4 ? cur[i - 4] : 0;
            const int b = prev ? prev[i] : 0;
            const int c = (prev && i >= 4) ? prev[i - 4] : 0;
            int pred = 0;
            switch (f) {
                case 1: pred = a; break;
                case 2: pred = b; break;
                case 3: pred = (a + b) >> 1; break;
                case 4: pred = paeth(a, b, c); break;
                default: break;
            }
            const uint8_t v = uint8_t(cur[i] - pred);
            scratch[i] = v;
            cost += uint64_t(std::abs(int(int8_t(v))));
        }
        if (cost < best_cost) {
            best_cost = cost;
            out[0] = uint8_t(f);
            std::memcpy(out + 1, scratch, stride);
        }
    }
}

}  // namespace

extern "C" hh_status hh_write_png(const char* path, const uint8_t* rgba8, uint32_t width, uint32_t height,
                                  uint32_t flip_rows) {
    if (!path || !rgba8 || width == 0 || height == 0) return hh_internal_fail(HH_ERR_INVALID_ARG, "bad argument");
    if (width > (1u << 24) || height > (1u << 24) || uint64_t(width) * height > (1ull << 30))
        return hh_internal_fail(HH_ERR_LIMIT_EXCEEDED, "image too large for the PNG writer");
    const size_t stride = size_t(width) * 4;
    std::vector<uint8_t> raw((stride + 1) * height);
    std::vector<uint8_t> scratch(stride);
    for (uint32_t r = 0; r < height; r++) {
        const uint32_t src = flip_rows ? height - 1 - r : r;
        const uint8_t* cur = rgba8 + size_t(src) * stride;
        const uint8_t* prev = nullptr;
        if (r > 0) prev = rgba8 + size_t(flip_rows ? src + 1 : src - 1) * stride;
        filter_row(cur, prev, stride, raw.data() + size_t(r) * (stride + 1), scratch.data());
    }
    uLongf zlen = compressBound(uLong(raw.size()));
    std::vector<uint8_t> z(zlen);
    if (compress2(z.data(), &zlen, raw.data(), uLong(raw.size()), 6) != Z_OK)
        return hh_internal_fail(HH_ERR_OUT_OF_MEMORY, "zlib compress2 failed");
    std::vector<uint8_t> file = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1A, '\n'};
    uint8_t ihdr[13];
    const uint32_t dims[2] = {width, height};
    for (int k = 0; k < 2; k++)
        for (int b = 0; b < 4; b++) ihdr[4 * k + b] = uint8_t(dims[k] >> (24 - 8 * b));
    ihdr[8] = 8;   // bit depth
    ihdr[9] = 6;   // colour type: truecolour with alpha
    ihdr[10] = 0;  // deflate
    ihdr[11] = 0;  // adaptive filtering
    ihdr[12] = 0;  // no interlace
    put_chunk(file, "IHDR", ihdr, sizeof(ihdr));
    put_chunk(file, "IDAT", z.data(), zlen);
    put_chunk(file, "IEND", nullptr, 0);
    FILE* f = std::fopen(path, "wb");
    if (!f) return hh_internal_fail(HH_ERR_IO, (std::string("cannot open for writing: ") + path).c_str());
    const size_t wrote = std::fwrite(file.data(), 1, file.size(), f);
    const int closed = std::fclose(f);
    if (wrote != file.size() || closed != 0)
        return hh_internal_fail(HH_ERR_IO, (std::string("write failed: ") + path).c_str());
    return HH_OK;
}

// ---------------------------------------------------------------------------
// PNG input for scene textures: every colour type (grey, RGB, palette,
// grey+alpha, RGBA) at the bit depths the specification allows for it
// (1/2/4/8/16), tRNS transparency (palette alpha and colour keys),
// non-interlaced. The result is RGBA8: 16-bit samples keep their high byte,
// sub-byte grey is scaled to 0..255.
// ---------------------------------------------------------------------------
namespace {
uint32_t be32(const uint8_t* p) { return (uint32_t(p[0]) << 24) | (uint32_t(p[1]) << 16) | (uint32_t(p[2]) << 8) | p[3]; }
}  // namespace

bool hh_internal_png_decode(const uint8_t* data, size_t n, std::vector<uint8_t>& rgba, uint32_t& width,
                            uint32_t& height, std::string& err) {
    static const uint8_t sig[8] = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1A, '\n'};
    if (n < 8 || std::memcmp(data, sig, 8) != 0) { err = "not a PNG file"; return false; }
    uint32_t w = 0, h = 0, depth = 0, ctype = 0, interlace = 0;
    bool have_ihdr = false, have_end = false;
    std::vector<uint8_t> idat, plte, trns;
    size_t pos = 8;
    while (pos + 12 <= n) {
        const uint32_t len = be32(data + pos);
        if (len > n - pos - 12) { err = "truncated chunk"; return false; }
        const uint8_t* type = data + pos + 4;
        const uint8_t* body = data + pos + 8;
        uLong crc = crc32(0L, Z_NULL, 0);
        crc = crc32(crc, type, uInt(4 + len));
        if (uint32_t(crc) != be32(body + len)) { err = "chunk CRC mismatch"; return false; }
        if (!std::memcmp(type, "IHDR", 4)) {
            if (len != 13) { err = "bad IHDR"; return false; }
            w = be32(body);
            h = be32(body + 4);
            depth = body[8];
            ctype = body[9];
            interlace = body[12];
            if (body[10] != 0 || body[11] != 0) { err = "unknown compression or filter method"; return false; }
            have_ihdr = true;
        } else if (!std::memcmp(type, "PLTE", 4)) {
            plte.assign(body, body + len);
        } else if (!std::memcmp(type, "tRNS", 4)) {
            trns.assign(body, body + len);
        } else if (!std::memcmp(type, "IDAT", 4)) {
            idat.insert(idat.end(), body, body + len);
        } else if (!std::memcmp(type, "IEND", 4)) {
            have_end = true;
            break;
        }
        pos += 12 + size_t(len);
    }
    if (!have_ihdr || !have_end) { err = "missing IHDR or IEND"; return false; }
    if (w == 0 || h == 0 || w > (1u << 15) || h > (1u << 15)) { err = "unsupported PNG size"; return false; }
    if (interlace != 0) { err = "interlaced PNG is not supported"; return false; }
    int channels = 0;
    switch (ctype) {
        case 0: channels = 1; break;
        case 2: channels = 3; break;
        case 3: channels = 1; break;
        case 4: channels = 2; break;
        case 6: channels = 4; break;
        default: err = "bad colour type"; return false;
    }
    const bool depth_ok = (ctype == 0 && (depth == 1 || depth == 2 || depth == 4 || depth == 8 || depth == 16)) ||
                          (ctype == 3 && (depth == 1 || depth == 2 || depth == 4 || depth == 8)) ||
                          ((ctype == 2 || ctype == 4 || ctype == 6) && (depth == 8 || depth == 16));
    if (!depth_ok) { err = "bad bit depth for the colour type"; return false; }
    if (ctype == 3 && (plte.empty() || plte.size() % 3 != 0)) { err = "palette image without PLTE"; return false; }
    const size_t bits = size_t(channels) * depth;
    const size_t stride = (size_t(w) * bits + 7) / 8;
    const size_t bpp = std::max<size_t>(1, bits / 8);
    std::vector<uint8_t> raw((stride + 1) * h);
    uLongf raw_len = uLongf(raw.size());
    if (uncompress(raw.data(), &raw_len, idat.data(), uLong(idat.size())) != Z_OK || raw_len != raw.size()) {
        err = "corrupt image data";
        return false;
    }
    // undo the per-row filters in place
    std::vector<uint8_t> zero(stride, 0);
    for (uint32_t r = 0; r < h; r++) {
        uint8_t* line = raw.data() + size_t(r) * (stride + 1);
        const uint8_t f = line[0];
        uint8_t* cur = line + 1;
        const uint8_t* prev = r ? raw.data() + size_t(r - 1) * (stride + 1) + 1 : zero.data();
        for (size_t i = 0; i < stride; i++) {
            const int a = i >= bpp ? cur[i - bpp] : 0;
            const int b = prev[i];
            const int c = i >= bpp ? prev[i - bpp] : 0;
            int pred = 0;
            switch (f) {
                case 0: break;
                case 1: pred = a; break;
                case 2: pred = b; break;
                case 3: pred = (a + b) >> 1; break;
                case 4: pred = paeth(a, b, c); break;
                default: err = "bad filter type"; return false;
            }
            cur[i] = uint8_t(cur[i] + pred);
        }
    }
    rgba.assign(size_t(w) * h * 4, 255);
    const int maxv = (1 << std::min<uint32_t>(depth, 8)) - 1;
    for (uint32_t r = 0; r < h; r++) {
        const uint8_t* s = raw.data() + size_t(r) * (stride + 1) + 1;
        uint8_t* o = rgba.data() + size_t(r) * w * 4;
        for (uint32_t x = 0; x < w; x++) {
            // raw samples (full 16-bit values kept for colour-key comparison)
            uint32_t v[4] = {0, 0, 0, 0};
            for (int c = 0; c < channels; c++) {
                if (depth == 16) {
                    const uint8_t* p = s + (size_t(x) * channels + c) * 2;
                    v[c] = (uint32_t(p[0]) << 8) | p[1];
                } else if (depth == 8) {
                    v[c] = s[size_t(x) * channels + c];
                } else {
                    const size_t bit = size_t(x) * depth;
                    v[c] = (s[bit / 8] >> (8 - depth - bit % 8)) & uint32_t(maxv);
                }
            }
            auto to8 = [&](uint32_t x8) -> uint8_t {
                if (depth == 16) return uint8_t(x8 >> 8);
                if (depth == 8) return uint8_t(x8);
                return uint8_t((x8 * 255) / uint32_t(maxv));
            };
            uint8_t* px = o + size_t(x) * 4;
            if (ctype == 3) {
                const uint32_t k = v[0];
                if (3 * k + 2 >= plte.size()) { err = "palette index out of range"; return false; }
                px[0] = plte[3 * k];
                px[1] = plte[3 * k + 1];
                px[2] = plte[3 * k + 2];
                px[3] = k < trns.size() ? trns[k] : 255;
            } else if (ctype == 0 || ctype == 4) {
                px[0] = px[1] = px[2] = to8(v[0]);
                px[3] = ctype == 4 ? to8(v[1]) : 255;
                if (ctype == 0 && trns.size() >= 2 && v[0] == ((uint32_t(trns[0]) << 8) | trns[1])) px[3] = 0;
            } else {
                for (int c = 0; c < 3; c++) px[c] = to8(v[c]);
                px[3] = ctype == 6 ? to8(v[3]) : 255;
                if (ctype == 2 && trns.size() >= 6 && v[0] == ((uint32_t(trns[0]) << 8) | trns[1]) &&
                    v[1] == ((uint32_t(trns[2]) << 8) | trns[3]) && v[2] == ((uint32_t(trns[4]) << 8) | trns[5]))
                    px[3] = 0;
            }
        }
    }
    width = w;
    height = h;
    return true;
}

// ---------------------------------------------------------------------------
// Environment maps for the IBL node (IBLNode, scene.h; the reference loads
// cube maps through AssetCore): a Radiance RGBE (.hdr) equirectangular image,
// resampled to the cube layout hh_scene_set_ibl takes.
//   * RGBE (Ward, "Real Pixels", Graphics Gems II): value = (m + 0.5) *
//     2^(e - 136) per channel, 0 when e = 0; scanlines flat or in the
//     run-length form (2, 2, width) with each channel encoded separately;
//     header FORMAT=32-bit_rle_rgbe, resolution "-Y H +X W" (top row first)
//     or "+Y H +X W".
//   * Lat-long layout: texel (i, j) of a W x H image holds direction
//     (sin t sin p, cos t, -sin t cos p), t = pi (j + 0.5) / H,
//     p = 2 pi ((i + 0.5) / W - 0.5); bilinear, repeating in i, clamped in j.
//   * Cube faces in Vulkan layer order (+X, -X, +Y, -Y, +Z, -Z), texel
//     (i, j) of face f at the direction of the samplers' face table; each
//     texel averages k x k lat-long samples over its footprint, k =
//     clamp(ceil(W / (4 size)), 1, 8).
// ---------------------------------------------------------------------------
namespace {

bool read_whole(const char* path, std::vector<uint8_t>& out) {
    FILE* f = std::fopen(path, "rb");
    if (!f) return false;
    uint8_t buf[1 << 16];
    size_t n;
    while ((n = std::fread(buf, 1, sizeof(buf), f)) > 0) out.insert(out.end(), buf, buf + n);
    const bool ok = !std::ferror(f);
    std::fclose(f);
    return ok;
}

bool hdr_decode(const std::vector<uint8_t>& d, std::vector<float>& rgb, uint32_t& width, uint32_t& height,
                std::string& err) {
    size_t pos = 0;
    auto line = [&](std::string& s) {
        s.clear();
        while (pos < d.size() && d[pos] != '\n') s += char(d[pos++]);
        if (pos >= d.size()) return false;
        pos++;
        return true;
    };
    std::string s;
    if (!line(s) || s.compare(0, 2, "#?") != 0) { err = "not a Radiance HDR file"; return false; }
    bool format_ok = false;
    for (;;) {
        if (!line(s)) { err = "truncated header"; return false; }
        if (s.empty()) break;
        if (s.compare(0, 7, "FORMAT=") == 0) {
            if (s != "FORMAT=32-bit_rle_rgbe") { err = "unsupported " + s; return false; }
            format_ok = true;
        }
    }
    if (!format_ok) { err = "missing FORMAT=32-bit_rle_rgbe"; return false; }
    if (!line(s)) { err = "missing resolution line"; return false; }
    char ya[3] = {0}, xa[3] = {0};
    int h = 0, w = 0;
    if (std::sscanf(s.c_str(), "%2s %d %2s %d", ya, &h, xa, &w) != 4 || std::string(xa) != "+X" ||
        (std::string(ya) != "-Y" && std::string(ya) != "+Y"))
        { err = "unsupported resolution line: " + s; return false; }
    if (w <= 0 || h <= 0 || w > (1 << 15) || h > (1 << 15)) { err = "bad image size"; return false; }
    const bool bottom_up = std::string(ya) == "+Y";
    rgb.assign(size_t(w) * h * 3, 0.0f);
    std::vector<uint8_t> row(size_t(w) * 4);
    for (int r = 0; r < h; r++) {
        if (w >= 8 && w < 32768 && pos + 4 <= d.size() && d[pos] == 2 && d[pos + 1] == 2 &&
            ((int(d[pos + 2]) << 8) | d[pos + 3]) == w && !(d[pos + 2] & 0x80)) {
            pos += 4;
            for (int c = 0; c < 4; c++) {
                int x = 0;
                while (x < w) {
                    if (pos >= d.size()) { err = "truncated scanline"; return false; }
                    int count = d[pos++];
                    if (count > 128) {
                        count -= 128;
                        if (count > w - x || pos >= d.size()) { err = "bad run length"; return false; }
                        const uint8_t v = d[pos++];
                        for (int k = 0; k < count; k++) row[size_t(x++) * 4 + c] = v;
                    } else {
                        if (count == 0 || count > w - x || pos + count > d.size()) { err = "bad literal run"; return false; }
                        for (int k = 0; k < count; k++) row[size_t(x++) * 4 + c] = d[pos++];
                    }
                }
            }
        } else {
            if (pos + row.size() > d.size()) { err = "truncated scanline"; return false; }
            if (d[pos] == 1 && d[pos + 1] == 1 && d[pos + 2] == 1) { err = "old run-length scanlines are not supported"; return false; }
            std::memcpy(row.data(), d.data() + pos, row.size());
            pos += row.size();
        }
        const int dst = bottom_up ? h - 1 - r : r;
        for (int x = 0; x < w; x++) {
            const uint8_t* p = &row[size_t(x) * 4];
            float* o = &rgb[(size_t(dst) * w + x) * 3];
            if (p[3] == 0) continue;
            const double f = std::ldexp(1.0, int(p[3]) - (128 + 8));
            for (int c = 0; c < 3; c++) o[c] = float((p[c] + 0.5) * f);
        }
    }
    width = uint32_t(w);
    height = uint32_t(h);
    return true;
}

void latlong_sample(const std::vector<float>& img, uint32_t w, uint32_t h, double dx, double dy, double dz, double out[3]) {
    const double kPi = 3.14159265358979323846;
    const double t = std::acos(std::max(-1.0, std::min(1.0, dy)));
    const double p = std::atan2(dx, -dz);
    const double x = (p / (2.0 * kPi) + 0.5) * w - 0.5, y = t / kPi * h - 0.5;
    const double fx = std::floor(x), fy = std::floor(y);
    const double ax = x - fx, ay = y - fy;
    auto wrap = [w](long i) { long r = i % long(w); return uint32_t(r < 0 ? r + long(w) : r); };
    auto clampy = [h](long j) { return uint32_t(j < 0 ? 0 : (j >= long(h) ? long(h) - 1 : j)); };
    const uint32_t i0 = wrap(long(fx)), i1 = wrap(long(fx) + 1), j0 = clampy(long(fy)), j1 = clampy(long(fy) + 1);
    for (int c = 0; c < 3; c++) {
        const double a = img[(size_t(j0) * w + i0) * 3 + c], b = img[(size_t(j0) * w + i1) * 3 + c];
        const double e = img[(size_t(j1) * w + i0) * 3 + c], f = img[(size_t(j1) * w + i1) * 3 + c];
        out[c] = (a + (b - a) * ax) + ((e + (f - e) * ax) - (a + (b - a) * ax)) * ay;
    }
}

}  // namespace

extern "C" hh_status hh_ibl_from_hdr(const char* path, uint32_t size, float* out_faces) {
    if (!path || !out_faces || size == 0 || size > 4096) return hh_internal_fail(HH_ERR_INVALID_ARG, "bad argument");
    std::vector<uint8_t> bytes;
    if (!read_whole(path, bytes)) return hh_internal_fail(HH_ERR_IO, (std::string("cannot read ") + path).c_str());
    std::vector<float> img;
    uint32_t w = 0, h = 0;
    std::string err;
    if (!hdr_decode(bytes, img, w, h, err)) return hh_internal_fail(HH_ERR_IO, (std::string(path) + ": " + err).c_str());
    const int k = int(std::min<uint32_t>(8, std::max<uint32_t>(1, (w + 4 * size - 1) / (4 * size))));
    for (int face = 0; face < 6; face++)
        for (uint32_t j = 0; j < size; j++)
            for (uint32_t i = 0; i < size; i++) {
                double acc[3] = {0.0, 0.0, 0.0};
                for (int sj = 0; sj < k; sj++)
                    for (int si = 0; si < k; si++) {
                        const double sc = 2.0 * ((i + (si + 0.5) / k) / size) - 1.0;
                        const double tc = 2.0 * ((j + (sj + 0.5) / k) / size) - 1.0;
                        double d[3];
                        switch (face) {
                            case 0: d[0] = 1.0; d[1] = -tc; d[2] = -sc; break;
                            case 1: d[0] = -1.0; d[1] = -tc; d[2] = sc; break;
                            case 2: d[0] = sc; d[1] = 1.0; d[2] = tc; break;
                            case 3: d[0] = sc; d[1] = -1.0; d[2] = -tc; break;
                            case 4: d[0] = sc; d[1] = -tc; d[2] = 1.0; break;
                            default: d[0] = -sc; d[1] = -tc; d[2] = -1.0; break;
                        }
                        const double n = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
                        double v[3];
                        latlong_sample(img, w, h, d[0] / n, d[1] / n, d[2] / n, v);
                        for (int c = 0; c < 3; c++) acc[c] += v[c];
                    }
                float* o = out_faces + ((size_t(face) * size + j) * size + i) * 4;
                for (int c = 0; c < 3; c++) o[c] = float(acc[c] / (k * k));
                o[3] = 1.0f;
            }
    return HH_OK;
}

extern "C" hh_status hh_scene_load_ibl_hdr(hh_scene scene, const char* path, uint32_t size) {
    if (!scene || !path || size == 0 || size > 4096) return hh_internal_fail(HH_ERR_INVALID_ARG, "bad argument");
    std::vector<float> faces(size_t(6) * size * size * 4);
    const hh_status st = hh_ibl_from_hdr(path, size, faces.data());
    if (st != HH_OK) return st;
    return hh_scene_set_ibl(scene, size, faces.data());
}
