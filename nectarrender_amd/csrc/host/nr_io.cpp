// nr_io.cpp — OBJ, PNG, PNM and Radiance HDR readers; JPEG, BMP and TGA are in
// nr_image.cpp (see nr_io.h).
//
// The formats are decoded from their published specifications (Wavefront OBJ,
// PNG ISO/IEC 15948 with zlib inflate, Netpbm, Greg Ward's RGBE); the
// conversions applied after decoding follow what the reference's loaders
// return: tinyobjloader 2.x with triangulation (quads split along the shorter
// diagonal, larger faces by its ear clipping) feeding parse_obj_file
// (mesh.h:39-141), and stb_image 2.30 for
// images (8-bit output, 16-bit samples reduced to their high byte, HDR texels
// as mantissa * 2^(exponent-136)).
#include "nr_io.h"
#include "nr_image.h"

#include <zlib.h>

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <limits>
#include <sstream>
#include <stdexcept>
#include <unordered_map>

namespace nrio {
namespace {

std::string read_file(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("cannot open file: " + path);
    std::ostringstream ss;
    ss << f.rdbuf();
    return ss.str();
}

std::string lower_ext(const std::string& path) {
    size_t dot = path.find_last_of('.');
    if (dot == std::string::npos) return "";
    std::string e = path.substr(dot + 1);
    for (auto& ch : e) ch = (char)std::tolower((unsigned char)ch);
    return e;
}

// ---------------------------------------------------------------------------
// OBJ
// ---------------------------------------------------------------------------
struct V3 { float x, y, z; };
inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 mul(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }                    // vector.h:455-459
inline V3 cross(V3 u, V3 v) {                                                               // vector.h:461-466
    return {u.y * v.z - u.z * v.y, u.z * v.x - u.x * v.z, u.x * v.y - u.y * v.x};
}
inline float length(V3 a) { return std::sqrt(((a.x * a.x) + (a.y * a.y)) + (a.z * a.z)); }
inline V3 normalize(V3 a) { return mul(a, 1.0f / length(a)); }                              // v / |v| = (1/|v|) v
inline bool near_zero(V3 a) { return length(a) < 1e-8f; }

struct Idx { int v, vt, vn; };

// OBJ index: 1-based, negative = relative to the current count, 0 invalid.
bool fix_index(long raw, size_t count, int& out) {
    if (raw > 0) { out = (int)(raw - 1); return true; }
    if (raw == 0) return false;
    long r = (long)count + raw;
    if (r < 0) return false;
    out = (int)r;
    return true;
}

float parse_real(const char*& s) {
    while (*s == ' ' || *s == '\t') s++;
    const char* b = s;
    while (*s && *s != ' ' && *s != '\t' && *s != '\r' && *s != '\n') s++;
    std::string tok(b, s);
    if (tok.empty()) return 0.0f;
    char* end = nullptr;
    double d = std::strtod(tok.c_str(), &end);   // correctly rounded, then to float
    return (float)d;
}

// Five or more corners: the ear clipping tinyobjloader 2.x runs under
// triangulate = true (its built-in method, all arithmetic in float).  The face
// is projected onto two coordinate axes chosen from the first corner whose
// edges are not parallel: dropping x when the x component of their cross
// product is the largest, else using (x, z), or (x, y) when z is the largest.
// Starting at corner 0, the triangle (g, g+1, g+2) is cut off when the sign of
// its turn agrees with the sign of the projected term x_g y_{g+1} - y_g x_{g+1}
// and no other remaining corner lies inside it (even-odd crossing test);
// otherwise g advances.  The middle corner of a cut triangle is removed, and
// the search stops after as many failed tries as corners remain (the rest of a
// degenerate face is dropped).  The last three corners form the final triangle.
bool inside_triangle(const float* x, const float* y, float tx, float ty) {
    bool in = false;
    for (int i = 0, j = 2; i < 3; j = i++)
        if (((y[i] > ty) != (y[j] > ty)) && (tx < (x[j] - x[i]) * (ty - y[i]) / (y[j] - y[i]) + x[i])) in = !in;
    return in;
}

void triangulate_ngon(const std::vector<Idx>& face, const std::vector<float>& pos, std::vector<Idx>& corners) {
    const size_t n = face.size();
    auto valid = [&](int v, size_t c) { return v >= 0 && (size_t)v * 3 + c < pos.size(); };
    size_t ax0 = 1, ax1 = 2;
    for (size_t k = 0; k < n; k++) {
        const int a = face[k].v, b = face[(k + 1) % n].v, c = face[(k + 2) % n].v;
        if (!valid(a, 2) || !valid(b, 2) || !valid(c, 2)) continue;
        const float e0x = pos[3 * b] - pos[3 * a], e0y = pos[3 * b + 1] - pos[3 * a + 1],
                    e0z = pos[3 * b + 2] - pos[3 * a + 2];
        const float e1x = pos[3 * c] - pos[3 * b], e1y = pos[3 * c + 1] - pos[3 * b + 1],
                    e1z = pos[3 * c + 2] - pos[3 * b + 2];
        const float cx = std::fabs(e0y * e1z - e0z * e1y), cy = std::fabs(e0z * e1x - e0x * e1z),
                    cz = std::fabs(e0x * e1y - e0y * e1x);
        const float eps = std::numeric_limits<float>::epsilon();
        if (cx > eps || cy > eps || cz > eps) {
            if (!(cx > cy && cx > cz)) {
                ax0 = 0;
                if (cz > cx && cz > cy) ax1 = 1;
            }
            break;
        }
    }
    std::vector<Idx> rem = face;
    size_t g = 0, tries = n, last_count = n;
    while (rem.size() > 3 && tries > 0) {
        const size_t m = rem.size();
        if (g >= m) g -= m;
        if (last_count != m) {
            last_count = m;
            tries = m;
        } else {
            tries--;
        }
        Idx ind[3];
        float x[3], y[3];
        for (int k = 0; k < 3; k++) {
            ind[k] = rem[(g + k) % m];
            const int v = ind[k].v;
            const bool ok = valid(v, ax0) && valid(v, ax1);
            x[k] = ok ? pos[3 * (size_t)v + ax0] : 0.0f;
            y[k] = ok ? pos[3 * (size_t)v + ax1] : 0.0f;
        }
        const float turn = (x[1] - x[0]) * (y[2] - y[1]) - (y[1] - y[0]) * (x[2] - x[1]);
        const float term = (x[0] * y[1] - y[0] * x[1]) * 0.5f;
        if (turn * term < 0.0f) { g++; continue; }
        bool blocked = false;
        for (size_t o = 3; o < m && !blocked; o++) {
            const int v = rem[(g + o) % m].v;
            if (!valid(v, ax0) || !valid(v, ax1)) continue;
            blocked = inside_triangle(x, y, pos[3 * (size_t)v + ax0], pos[3 * (size_t)v + ax1]);
        }
        if (blocked) { g++; continue; }
        corners.insert(corners.end(), {ind[0], ind[1], ind[2]});
        rem.erase(rem.begin() + (long)((g + 1) % m));
    }
    if (rem.size() == 3) corners.insert(corners.end(), {rem[0], rem[1], rem[2]});
}

}  // namespace

MeshData parse_obj(const std::string& text, const std::string& name) {
    std::vector<float> pos, nrm, tex;
    std::vector<Idx> corners;          // triangulated face corners, 3 per triangle
    std::istringstream in(text);
    std::string line;
    size_t lineno = 0;
    std::string pending;
    while (std::getline(in, line)) {
        lineno++;
        if (!line.empty() && line.back() == '\r') line.pop_back();
        if (!line.empty() && line.back() == '\\') {          // line continuation
            pending += line.substr(0, line.size() - 1);
            continue;
        }
        if (!pending.empty()) { line = pending + line; pending.clear(); }
        const char* s = line.c_str();
        while (*s == ' ' || *s == '\t') s++;
        if (*s == '\0' || *s == '#') continue;
        if (s[0] == 'v' && (s[1] == ' ' || s[1] == '\t')) {
            s += 2;
            for (int k = 0; k < 3; k++) pos.push_back(parse_real(s));
        } else if (s[0] == 'v' && s[1] == 'n' && (s[2] == ' ' || s[2] == '\t')) {
            s += 3;
            for (int k = 0; k < 3; k++) nrm.push_back(parse_real(s));
        } else if (s[0] == 'v' && s[1] == 't' && (s[2] == ' ' || s[2] == '\t')) {
            s += 3;
            float u = parse_real(s), v = parse_real(s);
            tex.push_back(u);
            tex.push_back(v);
        } else if (s[0] == 'f' && (s[1] == ' ' || s[1] == '\t')) {
            s += 2;
            std::vector<Idx> face;
            std::istringstream fs(s);
            std::string tok;
            while (fs >> tok) {
                Idx id{-1, -1, -1};
                long a[3] = {0, 0, 0};
                bool has[3] = {false, false, false};
                size_t p = 0;
                for (int k = 0; k < 3 && p <= tok.size(); k++) {
                    size_t q = tok.find('/', p);
                    std::string part = tok.substr(p, q == std::string::npos ? std::string::npos : q - p);
                    if (!part.empty()) { a[k] = std::strtol(part.c_str(), nullptr, 10); has[k] = true; }
                    if (q == std::string::npos) break;
                    p = q + 1;
                }
                if (!has[0] || !fix_index(a[0], pos.size() / 3, id.v))
                    throw std::runtime_error("Failed to load OBJ '" + name + "': bad vertex index on line " +
                                             std::to_string(lineno));
                if (has[1] && !fix_index(a[1], tex.size() / 2, id.vt))
                    throw std::runtime_error("Failed to load OBJ '" + name + "': bad texcoord index on line " +
                                             std::to_string(lineno));
                if (has[2] && !fix_index(a[2], nrm.size() / 3, id.vn))
                    throw std::runtime_error("Failed to load OBJ '" + name + "': bad normal index on line " +
                                             std::to_string(lineno));
                face.push_back(id);
            }
            if (face.size() < 3) continue;                   // degenerate face: skipped with a warning
            if (face.size() == 3) {
                corners.insert(corners.end(), face.begin(), face.end());
            } else if (face.size() == 4) {
                // split along the shorter diagonal, computed in float
                auto P = [&](int i) { return V3{pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]}; };
                for (auto& c : face)
                    if ((size_t)c.v * 3 + 2 >= pos.size())
                        throw std::runtime_error("Failed to load OBJ '" + name + "': vertex index out of range");
                V3 e02 = sub(P(face[2].v), P(face[0].v));
                V3 e13 = sub(P(face[3].v), P(face[1].v));
                float s02 = e02.x * e02.x + e02.y * e02.y + e02.z * e02.z;
                float s13 = e13.x * e13.x + e13.y * e13.y + e13.z * e13.z;
                if (s02 < s13) {
                    corners.insert(corners.end(), {face[0], face[1], face[2], face[0], face[2], face[3]});
                } else {
                    corners.insert(corners.end(), {face[0], face[1], face[3], face[1], face[2], face[3]});
                }
            } else {
                triangulate_ngon(face, pos, corners);
            }
        }
        // o, g, s, usemtl, mtllib, l, p: no geometry for the path tracer
    }
    const bool has_normals = !nrm.empty();

    MeshData out;
    struct KeyHash {
        size_t operator()(const Idx& k) const {
            size_t h = std::hash<int>()(k.v);
            h ^= std::hash<int>()(k.vt) + 0x9e3779b9 + (h << 6) + (h >> 2);
            h ^= std::hash<int>()(k.vn) + 0x9e3779b9 + (h << 6) + (h >> 2);
            return h;
        }
    };
    struct KeyEq {
        bool operator()(const Idx& a, const Idx& b) const { return a.v == b.v && a.vt == b.vt && a.vn == b.vn; }
    };
    std::unordered_map<Idx, uint32_t, KeyHash, KeyEq> uniq;
    struct Vtx { V3 p, n, t; float u, v; };
    std::vector<Vtx> verts;
    out.triangles.reserve(corners.size());
    for (const Idx& c : corners) {
        auto it = uniq.find(c);
        if (it != uniq.end()) { out.triangles.push_back(it->second); continue; }
        if ((size_t)c.v * 3 + 2 >= pos.size())
            throw std::runtime_error("Failed to load OBJ '" + name + "': vertex index out of range");
        Vtx x{};
        x.p = {pos[3 * c.v], pos[3 * c.v + 1], pos[3 * c.v + 2]};
        if (c.vn >= 0) {
            if ((size_t)c.vn * 3 + 2 >= nrm.size())
                throw std::runtime_error("Failed to load OBJ '" + name + "': normal index out of range");
            x.n = {nrm[3 * c.vn], nrm[3 * c.vn + 1], nrm[3 * c.vn + 2]};
        }
        if (c.vt >= 0) {
            if ((size_t)c.vt * 2 + 1 >= tex.size())
                throw std::runtime_error("Failed to load OBJ '" + name + "': texcoord index out of range");
            x.u = tex[2 * c.vt];
            x.v = tex[2 * c.vt + 1];
        }
        uint32_t id = (uint32_t)verts.size();
        verts.push_back(x);
        uniq.emplace(c, id);
        out.triangles.push_back(id);
    }
    const size_t nt = out.triangles.size() / 3;
    const uint32_t* T = out.triangles.data();
    if (!has_normals) {                                                      // mesh.h:103-114
        for (size_t f = 0; f < nt; f++) {
            Vtx &a = verts[T[3 * f]], &b = verts[T[3 * f + 1]], &c = verts[T[3 * f + 2]];
            V3 fn = normalize(cross(sub(b.p, a.p), sub(c.p, a.p)));
            a.n = add(a.n, fn);
            b.n = add(b.n, fn);
            c.n = add(c.n, fn);
        }
        for (auto& v : verts) v.n = near_zero(v.n) ? V3{0.0f, 1.0f, 0.0f} : normalize(v.n);
    }
    for (size_t f = 0; f < nt; f++) {                                        // mesh.h:116-131
        Vtx &a = verts[T[3 * f]], &b = verts[T[3 * f + 1]], &c = verts[T[3 * f + 2]];
        V3 e1 = sub(b.p, a.p), e2 = sub(c.p, a.p);
        float du1 = b.u - a.u, dv1 = b.v - a.v, du2 = c.u - a.u, dv2 = c.v - a.v;
        float denom = du1 * dv2 - du2 * dv1;
        V3 t = std::fabs(denom) < 1e-8f ? V3{1.0f, 0.0f, 0.0f} : mul(sub(mul(e1, dv2), mul(e2, dv1)), 1.0f / denom);
        a.t = add(a.t, t);
        b.t = add(b.t, t);
        c.t = add(c.t, t);
    }
    for (auto& v : verts) {                                                  // mesh.h:132-135
        V3 t = sub(v.t, mul(v.n, dot(v.t, v.n)));
        v.t = near_zero(t) ? V3{1.0f, 0.0f, 0.0f} : normalize(t);
    }
    out.vertices.resize(verts.size() * 11);
    for (size_t i = 0; i < verts.size(); i++) {
        const Vtx& v = verts[i];
        float* o = &out.vertices[11 * i];
        o[0] = v.p.x; o[1] = v.p.y; o[2] = v.p.z;
        o[3] = v.n.x; o[4] = v.n.y; o[5] = v.n.z;
        o[6] = v.t.x; o[7] = v.t.y; o[8] = v.t.z;
        o[9] = v.u; o[10] = v.v;
    }
    return out;
}

MeshData load_obj(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("Failed to load OBJ '" + path + "': cannot open file");
    std::ostringstream ss;
    ss << f.rdbuf();
    return parse_obj(ss.str(), path);
}

// ---------------------------------------------------------------------------
// images
// ---------------------------------------------------------------------------
namespace {

struct RawF { int C = 0, H = 0, W = 0; std::vector<float> hwc; };

uint32_t be32(const uint8_t* p) { return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3]; }

int paeth(int a, int b, int c) {
    int p = a + b - c, pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
    if (pa <= pb && pa <= pc) return a;
    if (pb <= pc) return b;
    return c;
}

// PNG (ISO/IEC 15948): every colour type / bit depth pair the specification
// allows, Adam7 interlacing, tRNS transparency.  Pairs the specification
// forbids, a tRNS chunk on a type that already has alpha, a tRNS chunk of the
// wrong size and images larger than MAX_DIM on a side are rejected, as
// stb_image rejects them.
bool decode_png(const std::string& buf, Raw8& out, std::string& err) {
    const uint8_t* d = (const uint8_t*)buf.data();
    const size_t n = buf.size();
    static const uint8_t sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
    if (n < 8 || std::memcmp(d, sig, 8) != 0) { err = "not a PNG"; return false; }
    size_t p = 8;
    uint32_t W = 0, H = 0;
    int depth = 0, ctype = -1, interlace = 0;
    std::vector<uint8_t> idat, pal, trns;
    bool has_trns = false;
    while (p + 8 <= n) {
        uint32_t len = be32(d + p);
        std::string type((const char*)d + p + 4, 4);
        if (p + 12 + (size_t)len > n) { err = "truncated PNG chunk"; return false; }
        const uint8_t* c = d + p + 8;
        if (type == "IHDR") {
            if (len < 13) { err = "bad IHDR"; return false; }
            W = be32(c); H = be32(c + 4); depth = c[8]; ctype = c[9]; interlace = c[12];
            if (W == 0 || H == 0 || W > MAX_DIM || H > MAX_DIM) { err = "PNG size out of range"; return false; }
            if (interlace > 1) { err = "bad PNG interlace method"; return false; }
            bool ok;
            switch (ctype) {
                case 0: ok = depth == 1 || depth == 2 || depth == 4 || depth == 8 || depth == 16; break;
                case 3: ok = depth == 1 || depth == 2 || depth == 4 || depth == 8; break;
                case 2: case 4: case 6: ok = depth == 8 || depth == 16; break;
                default: err = "bad PNG colour type"; return false;
            }
            if (!ok) { err = "bad PNG bit depth for its colour type"; return false; }
        } else if (type == "PLTE") {
            if (len == 0 || len % 3 != 0 || len > 768) { err = "bad PLTE"; return false; }
            pal.assign(c, c + len);
        } else if (type == "tRNS") {
            if (ctype < 0) { err = "tRNS before IHDR"; return false; }
            if (ctype == 4 || ctype == 6) { err = "tRNS with alpha"; return false; }
            if (ctype == 3) {
                if (pal.empty()) { err = "tRNS before PLTE"; return false; }
                if (len > pal.size() / 3) { err = "bad tRNS length"; return false; }
            } else if (len != (ctype == 0 ? 2u : 6u)) {
                err = "bad tRNS length";
                return false;
            }
            trns.assign(c, c + len);
            has_trns = true;
        } else if (type == "IDAT") {
            idat.insert(idat.end(), c, c + len);
        } else if (type == "IEND") {
            break;
        }
        p += 12 + (size_t)len;
    }
    if (W == 0 || H == 0 || ctype < 0) { err = "PNG without IHDR"; return false; }
    if (ctype == 3 && pal.empty()) { err = "PNG palette image without PLTE"; return false; }
    const int spp = ctype == 2 ? 3 : (ctype == 4 ? 2 : (ctype == 6 ? 4 : 1));   // samples per pixel in the file
    const size_t bits_pp = (size_t)spp * depth;
    const size_t bpp = std::max<size_t>(1, bits_pp / 8);
    // sub-images: the whole image, or the seven Adam7 passes (x0, y0, dx, dy)
    static const uint32_t adam7[7][4] = {{0, 0, 8, 8}, {4, 0, 8, 8}, {0, 4, 4, 8}, {2, 0, 4, 4},
                                         {0, 2, 2, 4}, {1, 0, 2, 2}, {0, 1, 1, 2}};
    struct Pass { uint32_t x0, y0, dx, dy, w, h; size_t stride, off; };
    std::vector<Pass> passes;
    size_t total = 0;
    for (int k = 0; k < (interlace ? 7 : 1); k++) {
        Pass ps{0, 0, 1, 1, W, H, 0, 0};
        if (interlace) {
            ps.x0 = adam7[k][0]; ps.y0 = adam7[k][1]; ps.dx = adam7[k][2]; ps.dy = adam7[k][3];
            ps.w = W > ps.x0 ? (W - ps.x0 + ps.dx - 1) / ps.dx : 0;
            ps.h = H > ps.y0 ? (H - ps.y0 + ps.dy - 1) / ps.dy : 0;
        }
        if (ps.w == 0 || ps.h == 0) continue;
        ps.stride = ((size_t)ps.w * bits_pp + 7) / 8;
        ps.off = total;
        total += (ps.stride + 1) * ps.h;
        passes.push_back(ps);
    }
    const size_t outC0 = ctype == 3 ? (has_trns ? 4 : 3) : spp + (has_trns ? 1 : 0);
    if (total > ((size_t)1 << 32) || (size_t)W * H * outC0 > ((size_t)1 << 32)) { err = "PNG too large"; return false; }
    std::vector<uint8_t> raw(total);
    z_stream zs{};
    if (inflateInit(&zs) != Z_OK) { err = "zlib init failed"; return false; }
    zs.next_in = idat.data();
    zs.avail_in = (uInt)idat.size();
    size_t done = 0;
    int zr = Z_OK;
    while (done < total) {   // inflate in pieces: avail_out is 32 bits
        const size_t chunk = std::min<size_t>(total - done, (size_t)1 << 30);
        zs.next_out = raw.data() + done;
        zs.avail_out = (uInt)chunk;
        zr = inflate(&zs, Z_FINISH);
        done += chunk - zs.avail_out;
        if (zr == Z_STREAM_END || (zr != Z_OK && zr != Z_BUF_ERROR) || zs.avail_out != 0) break;
    }
    inflateEnd(&zs);
    if (zr != Z_STREAM_END && !(zr == Z_BUF_ERROR && done == total) && !(zr == Z_OK && done == total)) {
        err = "corrupt PNG data";
        return false;
    }
    if (done < total) { err = "PNG data too short"; return false; }
    // palette images expand to RGB(A); grey/RGB with tRNS gain an alpha channel
    // (0 where the pixel equals the key colour), which counts as a channel
    const int outC = (int)outC0;
    out.C = outC;
    out.Cout = outC;
    out.H = (int)H; out.W = (int)W;
    out.hwc.assign((size_t)W * H * outC, 0);
    static const int scale[9] = {0, 0xff, 0x55, 0, 0x11, 0, 0, 0, 0x01};
    uint16_t key[3] = {0, 0, 0};
    if (has_trns && ctype != 3)
        for (int k = 0; k < spp; k++) {
            uint16_t v = (uint16_t)(trns[2 * k] << 8 | trns[2 * k + 1]);
            key[k] = depth == 16 ? v : (uint16_t)((v & 255) * scale[depth]);
        }
    auto pixel = [&](const uint8_t* row, uint32_t x, uint8_t* o) {
        if (ctype == 3) {
            int idx;
            if (depth == 8) {
                idx = row[x];
            } else {
                size_t bit = (size_t)x * depth;
                idx = (row[bit / 8] >> (8 - depth - (int)(bit % 8))) & ((1 << depth) - 1);
            }
            for (int k = 0; k < 3; k++) o[k] = (size_t)idx * 3 + k < pal.size() ? pal[idx * 3 + k] : 0;
            if (outC == 4) o[3] = (size_t)idx < trns.size() ? trns[idx] : 255;
            return;
        }
        uint16_t v16[4];
        for (int k = 0; k < spp; k++) {
            if (depth == 16) {
                v16[k] = (uint16_t)(row[((size_t)x * spp + k) * 2] << 8 | row[((size_t)x * spp + k) * 2 + 1]);
                o[k] = row[((size_t)x * spp + k) * 2];                                  // high byte
            } else if (depth == 8) {
                v16[k] = row[(size_t)x * spp + k];
                o[k] = row[(size_t)x * spp + k];
            } else {   // 1, 2, 4-bit grey
                size_t bit = (size_t)x * depth;
                int v = (row[bit / 8] >> (8 - depth - (int)(bit % 8))) & ((1 << depth) - 1);
                v16[k] = (uint16_t)(v * scale[depth]);
                o[k] = (uint8_t)(v * scale[depth]);
            }
        }
        if (has_trns) {
            bool eq = true;
            for (int k = 0; k < spp; k++) eq = eq && v16[k] == key[k];
            o[spp] = eq ? 0 : 255;
        }
    };
    for (const Pass& ps : passes) {
        std::vector<uint8_t> cur(ps.stride), prev(ps.stride, 0);   // the filters' previous row restarts per pass
        for (uint32_t y = 0; y < ps.h; y++) {
            const uint8_t* src = &raw[ps.off + (size_t)y * (ps.stride + 1)];
            int f = src[0];
            src++;
            for (size_t i = 0; i < ps.stride; i++) {
                int a = i >= bpp ? cur[i - bpp] : 0, b = prev[i], cc = i >= bpp ? prev[i - bpp] : 0;
                int v = src[i];
                switch (f) {
                    case 0: break;
                    case 1: v += a; break;
                    case 2: v += b; break;
                    case 3: v += (a + b) >> 1; break;
                    case 4: v += paeth(a, b, cc); break;
                    default: err = "bad PNG filter"; return false;
                }
                cur[i] = (uint8_t)v;
            }
            const size_t oy = ps.y0 + (size_t)y * ps.dy;
            for (uint32_t x = 0; x < ps.w; x++)
                pixel(cur.data(), x, &out.hwc[(oy * W + ps.x0 + (size_t)x * ps.dx) * outC]);
            std::swap(prev, cur);
        }
    }
    return true;
}

bool pnm_token(const std::string& b, size_t& p, std::string& tok) {
    tok.clear();
    while (p < b.size()) {
        char ch = b[p];
        if (ch == '#') { while (p < b.size() && b[p] != '\n') p++; continue; }
        if (std::isspace((unsigned char)ch)) { p++; continue; }
        break;
    }
    while (p < b.size() && !std::isspace((unsigned char)b[p])) tok += b[p++];
    return !tok.empty();
}

bool decode_pnm(const std::string& b, Raw8& out, std::string& err) {
    if (b.size() < 2 || b[0] != 'P' || (b[1] != '5' && b[1] != '6')) { err = "not a binary PNM"; return false; }
    size_t p = 2;
    std::string t;
    int v[3];
    for (int k = 0; k < 3; k++) {
        if (!pnm_token(b, p, t)) { err = "bad PNM header"; return false; }
        v[k] = std::atoi(t.c_str());
    }
    p++;   // one whitespace byte after maxval
    int C = b[1] == '5' ? 1 : 3;
    if (v[2] <= 0 || v[2] > 255) { err = "only 8-bit PNM is supported"; return false; }
    size_t need = (size_t)v[0] * v[1] * C;
    if (v[0] <= 0 || v[1] <= 0 || (uint32_t)v[0] > MAX_DIM || (uint32_t)v[1] > MAX_DIM) { err = "PNM size out of range"; return false; }
    if (p + need > b.size()) { err = "truncated PNM"; return false; }
    out.C = C; out.Cout = C; out.W = v[0]; out.H = v[1];
    out.hwc.assign(b.begin() + p, b.begin() + p + need);
    return true;
}

// Radiance RGBE: header lines, "-Y H +X W", then flat or new-style RLE scanlines.
bool decode_hdr(const std::string& b, RawF& out, std::string& err) {
    size_t p = 0;
    auto line = [&](std::string& s) {
        s.clear();
        while (p < b.size() && b[p] != '\n') s += b[p++];
        if (p < b.size()) p++;
        return true;
    };
    std::string s;
    line(s);
    if (s != "#?RADIANCE" && s != "#?RGBE") { err = "not a Radiance HDR file"; return false; }
    bool fmt = false;
    for (;;) {
        if (p >= b.size()) { err = "truncated HDR header"; return false; }
        line(s);
        if (s.empty()) break;
        if (s == "FORMAT=32-bit_rle_rgbe") fmt = true;
    }
    if (!fmt) { err = "unsupported HDR format"; return false; }
    line(s);
    int H = 0, W = 0;
    if (std::sscanf(s.c_str(), "-Y %d +X %d", &H, &W) != 2 || H <= 0 || W <= 0 || (uint32_t)H > MAX_DIM ||
        (uint32_t)W > MAX_DIM) {
        err = "unsupported HDR orientation";
        return false;
    }
    out.C = 3; out.H = H; out.W = W;
    out.hwc.assign((size_t)H * W * 3, 0.0f);
    std::vector<uint8_t> scan((size_t)W * 4);
    auto conv = [&](const uint8_t* rgbe, float* o) {
        if (rgbe[3] != 0) {
            float f = (float)std::ldexp(1.0f, (int)rgbe[3] - (128 + 8));
            o[0] = rgbe[0] * f; o[1] = rgbe[1] * f; o[2] = rgbe[2] * f;
        } else {
            o[0] = o[1] = o[2] = 0.0f;
        }
    };
    const uint8_t* d = (const uint8_t*)b.data();
    const size_t n = b.size();
    for (int y = 0; y < H; y++) {
        bool rle = false;
        if (W >= 8 && W < 32768 && p + 4 <= n && d[p] == 2 && d[p + 1] == 2 && !(d[p + 2] & 0x80)) {
            if ((int)((d[p + 2] << 8) | d[p + 3]) != W) { err = "HDR scanline width mismatch"; return false; }
            rle = true;
            p += 4;
        }
        if (!rle) {
            if (p + (size_t)W * 4 > n) { err = "truncated HDR data"; return false; }
            std::memcpy(scan.data(), d + p, (size_t)W * 4);
            p += (size_t)W * 4;
        } else {
            for (int k = 0; k < 4; k++) {
                int x = 0;
                while (x < W) {
                    if (p >= n) { err = "truncated HDR data"; return false; }
                    int count = d[p++];
                    if (count > 128) {
                        count -= 128;
                        if (p >= n || x + count > W) { err = "corrupt HDR run"; return false; }
                        uint8_t v = d[p++];
                        for (int i = 0; i < count; i++) scan[(size_t)(x++) * 4 + k] = v;
                    } else {
                        if (count == 0 || p + count > n || x + count > W) { err = "corrupt HDR dump"; return false; }
                        for (int i = 0; i < count; i++) scan[(size_t)(x++) * 4 + k] = d[p++];
                    }
                }
            }
        }
        for (int x = 0; x < W; x++) conv(&scan[(size_t)x * 4], &out.hwc[((size_t)y * W + x) * 3]);
    }
    return true;
}

// The formats in the order stb_image tries them (stbi__load_main): those with a
// signature first, TGA (which has none) last.
bool decode_any8(const std::string& path, Raw8& r, std::string& err) {
    std::string b = read_file(path);
    const uint8_t* d = (const uint8_t*)b.data();
    if (b.size() >= 8 && (uint8_t)b[0] == 137 && b[1] == 'P' && b[2] == 'N' && b[3] == 'G') return decode_png(b, r, err);
    if (is_bmp(d, b.size())) return decode_bmp(d, b.size(), r, err);
    if (is_jpeg(d, b.size())) return decode_jpeg(d, b.size(), r, err);
    if (b.size() >= 2 && b[0] == 'P' && (b[1] == '5' || b[1] == '6')) return decode_pnm(b, r, err);
    if (is_tga(d, b.size())) return decode_tga(d, b.size(), r, err);
    err = "unsupported image format";
    return false;
}

}  // namespace

void to_planar_u8(const uint8_t* hwc, int C, int Cout, int H, int W, std::vector<uint8_t>& out) {
    const size_t hw = (size_t)H * W;
    if (C >= 3) {
        // the reference reads element i*3 + c for every channel c < C
        out.assign((size_t)C * hw, 0);
        const size_t limit = (size_t)Cout * hw;
        for (int c = 0; c < C; c++)
            for (size_t i = 0; i < hw; i++) {
                size_t s = i * 3 + c;
                out[c * hw + i] = s < limit ? hwc[s] : 0;
            }
    } else {
        // grey (+alpha): replicate grey into three planes
        out.assign(3 * hw, 0);
        for (int c = 0; c < 3; c++)
            for (size_t i = 0; i < hw; i++) out[c * hw + i] = hwc[i * Cout];
    }
}

Image8 load_image_u8(const std::string& path) {
    Image8 img;
    if (lower_ext(path) == "hdr") {
        // stb converts HDR texels to 8 bits with gamma 1/2.2
        RawF f;
        std::string err;
        if (!decode_hdr(read_file(path), f, err)) throw std::runtime_error("Failed to load image: " + err);
        std::vector<uint8_t> hwc(f.hwc.size());
        for (size_t i = 0; i < f.hwc.size(); i++) {
            float z = (float)std::pow((double)(f.hwc[i] * 1.0f), (double)(1.0f / 2.2f)) * 255 + 0.5f;
            if (z < 0) z = 0;
            if (z > 255) z = 255;
            hwc[i] = (uint8_t)(int)z;
        }
        img.C = 3; img.H = f.H; img.W = f.W;
        to_planar_u8(hwc.data(), 3, 3, f.H, f.W, img.chw);
        return img;
    }
    Raw8 r;
    std::string err;
    if (!decode_any8(path, r, err)) throw std::runtime_error("Failed to load image: " + err);
    img.C = r.C; img.H = r.H; img.W = r.W;
    to_planar_u8(r.hwc.data(), r.C, r.Cout, r.H, r.W, img.chw);
    if (img.C < 3) img.C = 3;
    return img;
}

ImageF load_image_f32(const std::string& path) {
    ImageF img;
    std::string err;
    std::string b = read_file(path);
    if (b.rfind("#?", 0) == 0) {
        RawF f;
        if (!decode_hdr(b, f, err)) throw std::runtime_error("Failed to load HDRI: " + err);
        img.C = 3; img.H = f.H; img.W = f.W;
        const size_t hw = (size_t)f.H * f.W;
        img.chw.resize(3 * hw);
        for (int c = 0; c < 3; c++)
            for (size_t i = 0; i < hw; i++) img.chw[c * hw + i] = f.hwc[i * 3 + c];
        return img;
    }
    Raw8 r;
    if (!decode_any8(path, r, err)) throw std::runtime_error("Failed to load HDRI: " + err);
    // 8-bit to linear: colour channels pow(x/255, 2.2), an alpha channel x/255
    const int C = r.C, Co = r.Cout;
    const int ncol = (Co & 1) ? Co : Co - 1;
    const size_t hw = (size_t)r.H * r.W;
    img.C = C; img.H = r.H; img.W = r.W;
    img.chw.resize((size_t)C * hw);
    for (int c = 0; c < C; c++)
        for (size_t i = 0; i < hw; i++) {
            uint8_t v = r.hwc[i * Co + c];
            img.chw[c * hw + i] = c < ncol ? (float)(std::pow((double)(v / 255.0f), (double)2.2f) * 1.0f) : v / 255.0f;
        }
    if (C < 3) throw std::runtime_error("Failed to load HDRI: an HDRI needs three colour channels");
    return img;
}

}  // namespace nrio
