/* image_io.cpp - 8-bit PNG output of tone-mapped frames (the PNG format, ISO/IEC 15948: signature,
   IHDR, one zlib-compressed IDAT, IEND, CRC-32 per chunk). Every scanline uses the Sub filter, as
   the reference writer does (common/image_io.cpp:37-66), so the files compress alike. */
#include "image_io.h"

#include <zlib.h>

#include <cstdio>
#include <vector>

namespace bh {

namespace {

void be32(uint8_t* p, uint32_t v) {
    p[0] = (uint8_t)(v >> 24);
    p[1] = (uint8_t)(v >> 16);
    p[2] = (uint8_t)(v >> 8);
    p[3] = (uint8_t)v;
}

bool chunk(FILE* f, const char type[4], const uint8_t* data, uint32_t len) {
    uint8_t head[8];
    be32(head, len);
    for (int i = 0; i < 4; ++i) head[4 + i] = (uint8_t)type[i];
    uLong crc = crc32(0L, head + 4, 4);
    if (len) crc = crc32(crc, data, len);
    uint8_t tail[4];
    be32(tail, (uint32_t)crc);
    return std::fwrite(head, 1, 8, f) == 8 && (len == 0 || std::fwrite(data, 1, len, f) == len) &&
           std::fwrite(tail, 1, 4, f) == 4;
}

}  // namespace

bool write_png(const std::string& path, int w, int h, int channels, const uint8_t* data, int level) {
    if (w <= 0 || h <= 0 || !data || (channels != 1 && channels != 3 && channels != 4)) return false;
    const size_t stride = (size_t)w * channels;
    // filtered scanlines: filter byte 1 (Sub), then each byte minus the byte one pixel to the left
    std::vector<uint8_t> raw((stride + 1) * (size_t)h);
    for (int y = 0; y < h; ++y) {
        uint8_t* o = &raw[(size_t)y * (stride + 1)];
        const uint8_t* r = data + (size_t)y * stride;
        o[0] = 1;
        for (size_t i = 0; i < stride; ++i) o[1 + i] = (uint8_t)(r[i] - (i >= (size_t)channels ? r[i - channels] : 0));
    }
    uLongf clen = compressBound((uLong)raw.size());
    std::vector<uint8_t> comp(clen);
    if (compress2(comp.data(), &clen, raw.data(), (uLong)raw.size(), level) != Z_OK) return false;
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    static const uint8_t sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
    uint8_t ihdr[13];
    be32(ihdr, (uint32_t)w);
    be32(ihdr + 4, (uint32_t)h);
    ihdr[8] = 8;                                                  // bit depth
    ihdr[9] = (uint8_t)(channels == 1 ? 0 : (channels == 3 ? 2 : 6));   // grey / RGB / RGBA
    ihdr[10] = ihdr[11] = ihdr[12] = 0;                           // deflate, adaptive filters, no interlace
    bool ok = std::fwrite(sig, 1, 8, f) == 8 && chunk(f, "IHDR", ihdr, 13) &&
              chunk(f, "IDAT", comp.data(), (uint32_t)clen) && chunk(f, "IEND", nullptr, 0);
    return std::fclose(f) == 0 && ok;
}

}  // namespace bh
