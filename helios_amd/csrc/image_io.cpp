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
