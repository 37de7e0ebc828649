// psy_film.cpp -- film output: the per-pixel XYZ film of psy_render as an
// 8-bit sRGB image and as a PNG or PFM file.  Host code only.
//
// Restates:
//   Film::scanline_image_8bbc          film/film.hpp:173-225
//   XYZ_to_sRGB_E, sRGB_gamma          color/color.hpp:384-386, :425-439
//   write_png_from_film                renderer/renderer.cpp:24-42
// The reference dithers with an RNG seeded from the clock
// (utils/rng.hpp:58-82), so its dither values are not reproducible; here
// the dither of channel c of pixel i is a hash of (seed, i, c) mapped to
// [-0.5, 0.5), and seed 0 turns it off.  The PNG is written without a
// compression library: zlib stream of stored deflate blocks.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "psy_internal.hpp"

namespace {

float srgb_gamma(float n) { return n < 0.0031308f ? (n * 12.92f) : ((1.055f * std::pow(n, 1.0f / 2.4f)) - 0.055f); }
float clamp01(float x) { return x < 0.0f ? 0.0f : (x > 1.0f ? 1.0f : x); }

// XYZ -> linear sRGB with white point E (unclamped)
void xyz_to_rgbe_linear(const float* xyz, float rgb[3]) {
    rgb[0] = ((xyz[0] * 2.6897f) + (xyz[1] * -1.2759f)) + (xyz[2] * -0.4138f);
    rgb[1] = ((xyz[0] * -1.0216f) + (xyz[1] * 1.9778f)) + (xyz[2] * 0.0438f);
    rgb[2] = ((xyz[0] * 0.0613f) + (xyz[1] * -0.2245f)) + (xyz[2] * 1.1632f);
}

// a uniform value in [-0.5, 0.5) from (seed, pixel, channel)
float dither(uint32_t seed, uint64_t pixel, uint32_t channel) {
    uint64_t h = ((uint64_t)seed << 32) ^ (pixel * 3u + channel);
    h ^= h >> 33;
    h *= 0xff51afd7ed558ccdull;
    h ^= h >> 33;
    h *= 0xc4ceb9fe1a85ec53ull;
    h ^= h >> 33;
    return (float)(uint32_t)(h >> 40) * (1.0f / 16777216.0f) - 0.5f;
}

struct CrcTable {
    uint32_t t[256];
    CrcTable() {
        for (uint32_t i = 0; i < 256; ++i) {
            uint32_t x = i;
            for (int k = 0; k < 8; ++k) x = (x & 1u) ? 0xedb88320u ^ (x >> 1) : x >> 1;
            t[i] = x;
        }
    }
};
// CRC-32 (ISO-HDLC polynomial), as PNG chunks carry it
uint32_t crc32(const uint8_t* p, size_t n, uint32_t c = 0xffffffffu) {
    static const CrcTable table;
    for (size_t i = 0; i < n; ++i) c = table.t[(c ^ p[i]) & 0xffu] ^ (c >> 8);
    return c;
}

void put_be32(std::vector<uint8_t>& v, uint32_t x) {
    v.push_back((uint8_t)(x >> 24));
    v.push_back((uint8_t)(x >> 16));
    v.push_back((uint8_t)(x >> 8));
    v.push_back((uint8_t)x);
}

void chunk(std::vector<uint8_t>& png, const char* type, const std::vector<uint8_t>& data) {
    put_be32(png, (uint32_t)data.size());
    const size_t at = png.size();
    png.insert(png.end(), type, type + 4);
    png.insert(png.end(), data.begin(), data.end());
    put_be32(png, crc32(png.data() + at, png.size() - at) ^ 0xffffffffu);
}

}  // namespace

extern "C" {

int psy_film_to_srgb8(const float* film_xyz, const int32_t* sample_counts, int32_t width, int32_t height,
                      uint32_t dither_seed, uint8_t* rgb8) {
    if (!film_xyz || !rgb8 || width <= 0 || height <= 0)
        return psy::fail(PSY_ERR_INVALID, "psy_film_to_srgb8: bad argument");
    for (int64_t y = 0; y < height; ++y) {
        for (int64_t x = 0; x < width; ++x) {
            const int64_t i = y * width + x;
            float c[3];
            if (sample_counts && sample_counts[i] == 0) {
                // the grey checkerboard of unsampled pixels
                c[0] = c[1] = c[2] = (((y % 32) < 16) ^ ((x % 32) < 16)) ? 0.5f : 0.35f;
            } else {
                xyz_to_rgbe_linear(film_xyz + 3 * i, c);
                for (int k = 0; k < 3; ++k) c[k] = srgb_gamma(clamp01(c[k]));
            }
            for (int k = 0; k < 3; ++k) {
                float v = c[k] * 255.0f;
                if (dither_seed) v += dither(dither_seed, (uint64_t)i, (uint32_t)k);
                v = std::fmin(255.0f, std::fmax(0.0f, v));
                rgb8[3 * i + k] = (uint8_t)v;
            }
        }
    }
    return PSY_OK;
}

int psy_write_png(const char* path, const uint8_t* rgb8, int32_t width, int32_t height) {
    if (!path || !rgb8 || width <= 0 || height <= 0) return psy::fail(PSY_ERR_INVALID, "psy_write_png: bad argument");
    const size_t row = (size_t)width * 3 + 1;
    std::vector<uint8_t> raw(row * (size_t)height);
    for (int64_t y = 0; y < height; ++y) {
        raw[y * row] = 0;  // filter: none
        std::memcpy(&raw[y * row + 1], rgb8 + (size_t)y * width * 3, (size_t)width * 3);
    }
    // zlib stream (no compression): header, stored blocks, Adler-32
    std::vector<uint8_t> z = {0x78, 0x01};
    size_t off = 0;
    do {
        const size_t n = std::min<size_t>(65535, raw.size() - off);
        const bool last = off + n == raw.size();
        z.push_back(last ? 1 : 0);
        z.push_back((uint8_t)n);
        z.push_back((uint8_t)(n >> 8));
        z.push_back((uint8_t)~n);
        z.push_back((uint8_t)(~n >> 8));
        z.insert(z.end(), raw.begin() + off, raw.begin() + off + n);
        off += n;
    } while (off < raw.size());
    uint32_t a = 1, b = 0;
    for (uint8_t v : raw) {
        a = (a + v) % 65521u;
        b = (b + a) % 65521u;
    }
    put_be32(z, (b << 16) | a);
    std::vector<uint8_t> png = {0x89, 'P', 'N', 'G', 0x0d, 0x0a, 0x1a, 0x0a};
    std::vector<uint8_t> ihdr;
    put_be32(ihdr, (uint32_t)width);
    put_be32(ihdr, (uint32_t)height);
    ihdr.insert(ihdr.end(), {8, 2, 0, 0, 0});  // 8 bits, RGB, deflate, filter 0, no interlace
    chunk(png, "IHDR", ihdr);
    chunk(png, "IDAT", z);
    chunk(png, "IEND", {});
    FILE* f = std::fopen(path, "wb");
    if (!f) return psy::fail(PSY_ERR_INVALID, std::string("psy_write_png: cannot open ") + path);
    const bool ok = std::fwrite(png.data(), 1, png.size(), f) == png.size();
    if (std::fclose(f) != 0 || !ok) return psy::fail(PSY_ERR_INVALID, std::string("psy_write_png: cannot write ") + path);
    return PSY_OK;
}

int psy_write_pfm(const char* path, const float* film_xyz, int32_t width, int32_t height) {
    if (!path || !film_xyz || width <= 0 || height <= 0)
        return psy::fail(PSY_ERR_INVALID, "psy_write_pfm: bad argument");
    FILE* f = std::fopen(path, "wb");
    if (!f) return psy::fail(PSY_ERR_INVALID, std::string("psy_write_pfm: cannot open ") + path);
    bool ok = std::fprintf(f, "PF\n%d %d\n-1.0\n", width, height) > 0;
    std::vector<float> row((size_t)width * 3);
    // PFM rows run bottom to top; little-endian floats (scale -1)
    for (int64_t y = height - 1; ok && y >= 0; --y) {
        for (int64_t x = 0; x < width; ++x) xyz_to_rgbe_linear(film_xyz + 3 * (y * width + x), &row[3 * x]);
        ok = std::fwrite(row.data(), sizeof(float), row.size(), f) == row.size();
    }
    if (std::fclose(f) != 0 || !ok) return psy::fail(PSY_ERR_INVALID, std::string("psy_write_pfm: cannot write ") + path);
    return PSY_OK;
}

}  // extern "C"
