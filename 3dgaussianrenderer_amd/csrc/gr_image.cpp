// gr_image.cpp -- image output of the headless host (include/gr_capi.h:
// gr_image_write), replacing the GL display of app/main.cpp:435-466 and the
// image helpers of core/utilities.h:110-153.
//
// The back buffer is RGBA8 with row 0 at the bottom (clip y = -1,
// app/main.cpp:361); files are written top row first, so rows are flipped
// unless the caller says the buffer is already top-down.
//   .ppm  binary P6, RGB (alpha dropped)
//   .png  RGBA, 8 bits per channel, filter 0 on every row, zlib stream of
//         stored (uncompressed) deflate blocks -- no compression library
//         needed; CRC-32 (ISO 3309, polynomial 0xEDB88320) per chunk and
//         Adler-32 over the raw scanlines (RFC 1950/1951, PNG spec 1.2).
#include <algorithm>
#include <cctype>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/gr_capi.h"

namespace gr_detail {
void set_last_error(const std::string& msg);
}

namespace {

struct CrcTable {
    uint32_t t[256];
    CrcTable() {
        for (uint32_t n = 0; n < 256; ++n) {
            uint32_t c = n;
            for (int k = 0; k < 8; ++k) c = (c & 1u) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
            t[n] = c;
        }
    }
};

uint32_t crc_update(uint32_t crc, const uint8_t* p, size_t n) {
    static const CrcTable table;   // initialised once, thread-safe
    for (size_t i = 0; i < n; ++i) crc = table.t[(crc ^ p[i]) & 0xffu] ^ (crc >> 8);
    return crc;
}

void put32(std::vector<uint8_t>& v, uint32_t x) {
    v.push_back((uint8_t)(x >> 24));
    v.push_back((uint8_t)(x >> 16));
    v.push_back((uint8_t)(x >> 8));
    v.push_back((uint8_t)x);
}

bool write_chunk(FILE* f, const char type[4], const std::vector<uint8_t>& data) {
    std::vector<uint8_t> head;
    put32(head, (uint32_t)data.size());
    head.insert(head.end(), type, type + 4);
    uint32_t crc = crc_update(0xffffffffu, reinterpret_cast<const uint8_t*>(type), 4);
    crc = crc_update(crc, data.data(), data.size()) ^ 0xffffffffu;
    std::vector<uint8_t> tail;
    put32(tail, crc);
    return std::fwrite(head.data(), 1, head.size(), f) == head.size() &&
           std::fwrite(data.data(), 1, data.size(), f) == data.size() &&
           std::fwrite(tail.data(), 1, tail.size(), f) == tail.size();
}

const uint8_t* src_row(const uint8_t* rgba, int w, int h, int y, bool flip) {
    return rgba + (size_t)(flip ? h - 1 - y : y) * w * 4;
}

bool write_png(FILE* f, const uint8_t* rgba, int w, int h, bool flip) {
    static const uint8_t sig[8] = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
    if (std::fwrite(sig, 1, 8, f) != 8) return false;
    std::vector<uint8_t> ihdr;
    put32(ihdr, (uint32_t)w);
    put32(ihdr, (uint32_t)h);
    ihdr.insert(ihdr.end(), {8, 6, 0, 0, 0});   // 8-bit RGBA, deflate, filter set 0, no interlace
    if (!write_chunk(f, "IHDR", ihdr)) return false;
    // raw scanlines: filter byte 0 + RGBA row
    const size_t line = (size_t)w * 4 + 1;
    const size_t raw_n = line * (size_t)h;
    std::vector<uint8_t> z;
    z.reserve(raw_n + raw_n / 65535 * 5 + 16);
    z.push_back(0x78);   // CMF: deflate, 32K window
    z.push_back(0x01);   // FLG: no dictionary, check bits make 0x7801 % 31 == 0
    uint32_t a = 1, b = 0;   // Adler-32
    std::vector<uint8_t> raw;
    raw.reserve(raw_n);
    for (int y = 0; y < h; ++y) {
        raw.push_back(0);
        const uint8_t* r = src_row(rgba, w, h, y, flip);
        raw.insert(raw.end(), r, r + (size_t)w * 4);
    }
    for (size_t i = 0; i < raw_n; ++i) {
        a = (a + raw[i]) % 65521u;
        b = (b + a) % 65521u;
    }
    size_t off = 0;
    do {
        const size_t n = std::min<size_t>(65535, raw_n - off);
        const bool last = off + n == raw_n;
        z.push_back(last ? 1 : 0);   // BFINAL, BTYPE = 00 (stored)
        z.push_back((uint8_t)(n & 0xff));
        z.push_back((uint8_t)(n >> 8));
        z.push_back((uint8_t)(~n & 0xff));
        z.push_back((uint8_t)((~n >> 8) & 0xff));
        z.insert(z.end(), raw.begin() + (ptrdiff_t)off, raw.begin() + (ptrdiff_t)(off + n));
        off += n;
    } while (off < raw_n);
    put32(z, (b << 16) | a);
    if (!write_chunk(f, "IDAT", z)) return false;
    return write_chunk(f, "IEND", {});
}

bool write_ppm(FILE* f, const uint8_t* rgba, int w, int h, bool flip) {
    std::fprintf(f, "P6\n%d %d\n255\n", w, h);
    std::vector<uint8_t> row((size_t)w * 3);
    for (int y = 0; y < h; ++y) {
        const uint8_t* r = src_row(rgba, w, h, y, flip);
        for (int x = 0; x < w; ++x)
            for (int k = 0; k < 3; ++k) row[(size_t)3 * x + k] = r[(size_t)4 * x + k];
        if (std::fwrite(row.data(), 1, row.size(), f) != row.size()) return false;
    }
    return true;
}

bool ends_with(const std::string& s, const char* suf) {
    const size_t n = std::strlen(suf);
    if (s.size() < n) return false;
    for (size_t i = 0; i < n; ++i)
        if (std::tolower((unsigned char)s[s.size() - n + i]) != suf[i]) return false;
    return true;
}

}  // namespace

extern "C" gr_status gr_image_write(const char* path, const uint8_t* rgba8, int32_t width, int32_t height,
                                    int32_t bottom_up) {
    if (!path || !rgba8 || width <= 0 || height <= 0) {
        gr_detail::set_last_error("gr_image_write: bad argument");
        return GR_ERR_INVALID_ARGUMENT;
    }
    const std::string p(path);
    const bool png = ends_with(p, ".png");
    if (!png && !ends_with(p, ".ppm")) {
        gr_detail::set_last_error("gr_image_write: the file name must end in .png or .ppm");
        return GR_ERR_INVALID_ARGUMENT;
    }
    FILE* f = std::fopen(path, "wb");
    if (!f) {
        gr_detail::set_last_error("cannot create " + p);
        return GR_ERR_NOT_FOUND;
    }
    bool ok = png ? write_png(f, rgba8, width, height, bottom_up != 0)
                  : write_ppm(f, rgba8, width, height, bottom_up != 0);
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) {
        gr_detail::set_last_error("write to " + p + " failed");
        return GR_ERR_INVALID_ARGUMENT;
    }
    return GR_OK;
}
