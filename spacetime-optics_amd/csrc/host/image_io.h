/* image_io.h - 8-bit PNG writer for tone-mapped frames (common/image_io.hpp:6-7 of the reference). */
#pragma once
#include <cstdint>
#include <string>

namespace bh {

// channels = 1, 3 or 4; data row-major from the top left; zlib level 0..9 (default 4 as bhrt).
bool write_png(const std::string& path, int w, int h, int channels, const uint8_t* data, int level = 4);

}  // namespace bh
