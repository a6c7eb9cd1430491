// nr_io.h — scene file ingest for the _pathtracer module: Wavefront OBJ meshes,
// 8-bit images for textures and floating-point HDRIs.
//
// Output formats are the ones the C ABI takes (include/nectar_hip.h):
//   meshes  : 11 floats per vertex (position, normal, tangent, uv) + uint32 triangles
//   textures: planar CHW uint8
//   HDRIs   : planar CHW fp32
// All loaders throw std::runtime_error on failure; pybind11 turns that into
// Python's RuntimeError as the reference's loaders do (mesh.h:44-48,
// texture.h:13, 33).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace nrio {

struct MeshData {
    std::vector<float> vertices;      // n_vertices x 11
    std::vector<uint32_t> triangles;  // n_triangles x 3
    size_t n_vertices() const { return vertices.size() / 11; }
    size_t n_triangles() const { return triangles.size() / 3; }
};

// parse_obj_file (csrc/hittable/include/hittable/mesh.h:39-141): vertices are
// deduplicated per (v, vt, vn) index triple in first-use order; normals are
// area-independent face-normal sums when the file has none; tangents come from
// the uv gradients and are Gram-Schmidt orthogonalised against the normal.
MeshData load_obj(const std::string& path);
// Same vertex construction from in-memory OBJ text (tests, generated assets).
MeshData parse_obj(const std::string& text, const std::string& name = "<memory>");

struct Image8 {
    int C = 0, H = 0, W = 0;          // C = channels the file stores
    std::vector<uint8_t> chw;         // C x H x W
};
struct ImageF {
    int C = 0, H = 0, W = 0;
    std::vector<float> chw;
};

// load_image_uint8 (csrc/material/include/material/texture.h:8-27).
// PNG (every colour type and bit depth, Adam7 interlacing), JPEG (baseline,
// extended and progressive Huffman), BMP, TGA, binary PPM/PGM and Radiance .hdr.
Image8 load_image_u8(const std::string& path);
// load_image_float (texture.h:29-47): Radiance .hdr (RGBE), PFM, and 8-bit
// formats converted as pow(x/255, 2.2) for colour channels.
ImageF load_image_f32(const std::string& path);

// Interleaved-to-planar conversion with the reference's indexing: source
// element (i*3 + c) goes to plane c for every reported channel c < C
// (texture.h:20-22), which is the plain transpose for 3-channel data; hwc
// holds Cout channels per pixel.  Grey images (C < 3, which the reference
// reads past the end of) become three equal planes.
void to_planar_u8(const uint8_t* hwc, int C, int Cout, int H, int W, std::vector<uint8_t>& out);

}  // namespace nrio
