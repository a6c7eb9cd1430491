// nr_image.h — 8-bit raster decoders shared by the readers in nr_io.cpp:
// JPEG (baseline, extended sequential and progressive Huffman), BMP and TGA.
//
// Each decoder returns what stb_image 2.30 (the reference's image library,
// texture.h:8-27) returns for the file with no requested channel count:
// interleaved 8-bit samples, `C` = the channel count stb reports, `Cout` =
// samples per pixel in `hwc`.  They return false with a message in `err` for
// files they reject.
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace nrio {

// C = channel count the decoder reports, Cout = channels per pixel in hwc
struct Raw8 {
    int C = 0, Cout = 0, H = 0, W = 0;
    std::vector<uint8_t> hwc;
};

// Largest width or height any image reader accepts (stb_image's STBI_MAX_DIMENSIONS).
constexpr uint32_t MAX_DIM = 1u << 24;

bool is_jpeg(const uint8_t* d, size_t n);
bool is_bmp(const uint8_t* d, size_t n);
bool is_tga(const uint8_t* d, size_t n);

bool decode_jpeg(const uint8_t* d, size_t n, Raw8& out, std::string& err);
bool decode_bmp(const uint8_t* d, size_t n, Raw8& out, std::string& err);
bool decode_tga(const uint8_t* d, size_t n, Raw8& out, std::string& err);

}  // namespace nrio
