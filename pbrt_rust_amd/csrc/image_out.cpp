// image_out.cpp -- Film::write_image's file output (film.rs:15-33, 316-354),
// host only.  The film's final rgb (xyz_to_rgb(xyz) / weight_sum, clamped at
// zero, as the renderer computes it) goes to a file chosen by extension:
//   .pfm  the floats themselves (Portable Float Map, little endian, rows
//         bottom to top as the format stores them);
//   .ppm  8-bit binary PPM (P6) and
//   .png  8-bit RGB PNG, both with write_img's per-channel encoding
//         (film.rs:22-24): (255 p^(1/2.2) + 0.5) clamped to [0, 255] and
//         truncated (Rust's `as u8`: NaN gives 0).  The PNG is written
//         without external libraries: one IDAT of stored (uncompressed)
//         deflate blocks, filter type 0 on every row.
#include <ctype.h>
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/pbrtx_gpu.h"

namespace px {
int fail(int code, const std::string& msg);
}

namespace {

// write_img's to_byte (film.rs:22-24) in f32 arithmetic
uint8_t to_byte(float p) {
    const float g = 1.0f / 2.2f;
    float v = 255.0f * powf(p, g) + 0.5f;
    if (v != v) return 0;  // `as u8` of NaN
    if (v < 0.0f) v = 0.0f;
    if (v > 255.0f) v = 255.0f;
    return (uint8_t)v;  // truncation, as `as u8`
}

struct CrcTable {
    uint32_t t[256];
    CrcTable() {
        for (uint32_t i = 0; i < 256; ++i) {
            uint32_t v = i;
            for (int k = 0; k < 8; ++k) v = (v & 1u) ? 0xEDB88320u ^ (v >> 1) : v >> 1;
            t[i] = v;
        }
    }
};
// CRC-32 of PNG chunks (ISO 3309 polynomial, reflected)
uint32_t crc32(const uint8_t* p, size_t n) {
    static const CrcTable tab;  // initialised once, thread-safely
    uint32_t c = 0xFFFFFFFFu;
    for (size_t i = 0; i < n; ++i) c = tab.t[(c ^ p[i]) & 0xFFu] ^ (c >> 8);
    return ~c;
}

void put_be32(std::vector<uint8_t>& v, uint32_t x) {
    v.push_back((uint8_t)(x >> 24)); v.push_back((uint8_t)(x >> 16));
    v.push_back((uint8_t)(x >> 8)); v.push_back((uint8_t)x);
}

void png_chunk(std::vector<uint8_t>& out, const char* type, const std::vector<uint8_t>& data) {
    put_be32(out, (uint32_t)data.size());
    const size_t at = out.size();
    out.insert(out.end(), type, type + 4);
    out.insert(out.end(), data.begin(), data.end());
    put_be32(out, crc32(out.data() + at, 4 + data.size()));
}

std::vector<uint8_t> encode_png(const std::vector<uint8_t>& rgb8, uint32_t w, uint32_t h) {
    // raw scanlines: filter byte 0, then the row
    std::vector<uint8_t> raw;
    raw.reserve((size_t)h * (1 + 3 * (size_t)w));
    for (uint32_t y = 0; y < h; ++y) {
        raw.push_back(0);
        raw.insert(raw.end(), rgb8.begin() + (size_t)y * w * 3, rgb8.begin() + (size_t)(y + 1) * w * 3);
    }
    // zlib stream of stored deflate blocks (RFC 1950/1951)
    std::vector<uint8_t> z;
    z.push_back(0x78); z.push_back(0x01);
    size_t pos = 0;
    do {
        const size_t len = std::min<size_t>(65535, raw.size() - pos);
        const bool last = pos + len == raw.size();
        z.push_back(last ? 1 : 0);
        z.push_back((uint8_t)len); z.push_back((uint8_t)(len >> 8));
        z.push_back((uint8_t)~len); z.push_back((uint8_t)(~len >> 8));
        z.insert(z.end(), raw.begin() + pos, raw.begin() + pos + len);
        pos += len;
    } while (pos < raw.size());
    uint32_t a = 1, b = 0;  // Adler-32
    for (uint8_t c : raw) {
        a = (a + c) % 65521u;
        b = (b + a) % 65521u;
    }
    put_be32(z, (b << 16) | a);
    std::vector<uint8_t> out = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1A, '\n'};
    std::vector<uint8_t> ihdr;
    put_be32(ihdr, w);
    put_be32(ihdr, h);
    ihdr.push_back(8);  // bit depth
    ihdr.push_back(2);  // RGB
    ihdr.push_back(0); ihdr.push_back(0); ihdr.push_back(0);
    png_chunk(out, "IHDR", ihdr);
    png_chunk(out, "IDAT", z);
    png_chunk(out, "IEND", {});
    return out;
}

bool ends_with(const std::string& s, const char* suf) {
    const size_t n = strlen(suf);
    if (s.size() < n) return false;
    for (size_t i = 0; i < n; ++i)
        if (tolower((unsigned char)s[s.size() - n + i]) != suf[i]) return false;
    return true;
}

}  // namespace

extern "C" {

int pbrtx_image_encode8(const float* rgb, uint64_t n, uint8_t* out) {
    if ((!rgb || !out) && n) return px::fail(PBRTX_ERR_ARG, "null buffer");
    for (uint64_t i = 0; i < n; ++i) out[i] = to_byte(rgb[i]);
    return PBRTX_OK;
}

int pbrtx_write_image(const char* filename, const float* rgb, uint32_t xres, uint32_t yres) {
    if (!filename || !rgb || xres == 0 || yres == 0) return px::fail(PBRTX_ERR_ARG, "null argument or empty image");
    const std::string name(filename);
    const size_t n = (size_t)xres * yres * 3;
    std::vector<uint8_t> bytes;
    if (ends_with(name, ".pfm")) {
        char head[64];
        const int k = snprintf(head, sizeof(head), "PF\n%u %u\n-1.0\n", xres, yres);
        bytes.assign(head, head + k);
        // rows bottom to top, little-endian floats (the host's order on x86)
        for (uint32_t y = yres; y-- > 0;) {
            const uint8_t* row = reinterpret_cast<const uint8_t*>(rgb + (size_t)y * xres * 3);
            bytes.insert(bytes.end(), row, row + (size_t)xres * 3 * sizeof(float));
        }
    } else if (ends_with(name, ".ppm") || ends_with(name, ".png")) {
        std::vector<uint8_t> rgb8(n);
        for (size_t i = 0; i < n; ++i) rgb8[i] = to_byte(rgb[i]);
        if (ends_with(name, ".ppm")) {
            char head[64];
            const int k = snprintf(head, sizeof(head), "P6\n%u %u\n255\n", xres, yres);
            bytes.assign(head, head + k);
            bytes.insert(bytes.end(), rgb8.begin(), rgb8.end());
        } else {
            bytes = encode_png(rgb8, xres, yres);
        }
    } else {
        return px::fail(PBRTX_ERR_ARG, "image file name must end in .pfm, .ppm or .png: " + name);
    }
    FILE* f = fopen(filename, "wb");
    if (!f) return px::fail(PBRTX_ERR_ARG, "cannot open " + name + " for writing");
    const size_t wrote = fwrite(bytes.data(), 1, bytes.size(), f);
    const int closed = fclose(f);
    if (wrote != bytes.size() || closed != 0) return px::fail(PBRTX_ERR_STATE, "short write to " + name);
    return PBRTX_OK;
}

}  // extern "C"
