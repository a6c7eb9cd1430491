"""Scene-file readers of `_pathtracer` (csrc/host/nr_io.cpp) against the reference's.

The reference loads meshes with tinyobjloader + parse_obj_file
(csrc/hittable/include/hittable/mesh.h:39-141) and images with stb_image +
load_image_uint8 / load_image_float (csrc/material/include/material/texture.h:8-47).
oracle/_ref/libnr_ioref.so is those vendored libraries compiled from the
reference tree; the files below are written by the tests themselves (OBJ text,
PNG via zlib with chosen filters, PNM, Radiance RGBE flat and run-length).
Outputs must be bit-identical.  The restated glue of the checker is pinned by
closed forms (face normals, uv tangents, quad splits) in the first tests.
"""
import os
import struct
import zlib

import numpy as np
import pytest

from oracle import ioref

pt = pytest.importorskip("nectarrender_amd._pathtracer")
needs_ref = pytest.mark.skipif(not ioref.available(), reason="reference readers not built (no reference tree)")


# ---------------------------------------------------------------------------
# writers
# ---------------------------------------------------------------------------
def write_png(path, arr, ctype, depth=8, palette=None, trns=None, filters=(0, 1, 2, 3, 4), interlace=False,
              ihdr=None):
    """arr: H x W x spp integer samples (or H x W palette indices).  interlace: Adam7
    (each pass filtered as an image of its own); ihdr: raw IHDR fields to write instead
    (W, H, depth, ctype), for malformed files."""
    arr = np.asarray(arr)
    H, W = arr.shape[:2]
    spp = {0: 1, 2: 3, 3: 1, 4: 2, 6: 4}[ctype]
    a = arr.reshape(H, W, spp).astype(np.int64)
    if interlace:
        raw = bytearray()
        for x0, y0, dx, dy in ((0, 0, 8, 8), (4, 0, 8, 8), (0, 4, 4, 8), (2, 0, 4, 4), (0, 2, 2, 4), (1, 0, 2, 2),
                               (0, 1, 1, 2)):
            sub = a[y0::dy, x0::dx]
            if sub.shape[0] and sub.shape[1]:
                raw += _filtered_rows(sub, depth, spp, filters)
    else:
        raw = _filtered_rows(a, depth, spp, filters)

    def chunk(t, d):
        return struct.pack(">I", len(d)) + t + d + struct.pack(">I", zlib.crc32(t + d) & 0xFFFFFFFF)

    fw, fh_, fd, fc = ihdr if ihdr is not None else (W, H, depth, ctype)
    png = b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", fw, fh_, fd, fc, 0, 0, 1 if interlace else 0))
    if palette is not None:
        png += chunk(b"PLTE", bytes(np.asarray(palette, np.uint8).reshape(-1)))
    if trns is not None:
        png += chunk(b"tRNS", bytes(trns))
    png += chunk(b"IDAT", zlib.compress(bytes(raw), 6)) + chunk(b"IEND", b"")
    with open(path, "wb") as fh:
        fh.write(png)


def _filtered_rows(a, depth, spp, filters):
    """PNG scanlines of one (sub-)image a (h x w x spp), each row with its filter byte."""
    H = a.shape[0]
    rows = []
    for y in range(H):
        if depth == 16:
            line = b"".join(struct.pack(">H", int(v)) for v in a[y].reshape(-1))
        elif depth == 8:
            line = bytes(int(v) for v in a[y].reshape(-1))
        else:
            bits = "".join(format(int(v), f"0{depth}b") for v in a[y].reshape(-1))
            bits += "0" * (-len(bits) % 8)
            line = bytes(int(bits[i:i + 8], 2) for i in range(0, len(bits), 8))
        rows.append(line)
    bpp = max(1, spp * depth // 8)
    raw = bytearray()
    prev = bytes(len(rows[0]))
    for y, line in enumerate(rows):
        f = filters[y % len(filters)]
        out = bytearray()
        for i, x in enumerate(line):
            a_ = out_a = line[i - bpp] if i >= bpp else 0
            b_ = prev[i]
            c_ = prev[i - bpp] if i >= bpp else 0
            if f == 0:
                v = x
            elif f == 1:
                v = x - a_
            elif f == 2:
                v = x - b_
            elif f == 3:
                v = x - ((a_ + b_) >> 1)
            else:
                p = a_ + b_ - c_
                pa, pb, pc = abs(p - a_), abs(p - b_), abs(p - c_)
                pr = a_ if (pa <= pb and pa <= pc) else (b_ if pb <= pc else c_)
                v = x - pr
            del out_a
            out.append(v & 255)
        raw.append(f)
        raw += out
        prev = line
    return raw


def write_rgbe(path, rgbe, rle):
    """rgbe: H x W x 4 uint8."""
    H, W = rgbe.shape[:2]
    out = bytearray(b"#?RADIANCE\n# test\nFORMAT=32-bit_rle_rgbe\nEXPOSURE=1.0\n\n" + f"-Y {H} +X {W}\n".encode())
    for y in range(H):
        if not rle:
            out += bytes(rgbe[y].reshape(-1))
            continue
        out += bytes([2, 2, W >> 8, W & 255])
        for k in range(4):
            comp = rgbe[y, :, k]
            x = 0
            while x < W:
                run = 1
                while x + run < W and run < 127 and comp[x + run] == comp[x]:
                    run += 1
                if run >= 3:
                    out += bytes([128 + run, int(comp[x])])
                    x += run
                else:
                    n = 0
                    while x + n < W and n < 128:
                        if x + n + 2 < W and comp[x + n] == comp[x + n + 1] == comp[x + n + 2]:
                            break
                        n += 1
                    n = max(n, 1)
                    out += bytes([n]) + bytes(int(v) for v in comp[x:x + n])
                    x += n
    with open(path, "wb") as fh:
        fh.write(bytes(out))


def random_rgbe(rng, H, W):
    e = rng.integers(120, 140, (H, W, 1))
    m = rng.integers(128, 256, (H, W, 3))
    a = np.concatenate([m, e], 2).astype(np.uint8)
    a[0, :5] = a[0, 0]          # runs for the RLE coder
    a[1, 2, 3] = 0              # a zero exponent texel
    return a


# ---------------------------------------------------------------------------
# the checker's own glue, pinned by closed forms
# ---------------------------------------------------------------------------
@needs_ref
def test_ref_obj_closed_forms(tmp_path):
    p = tmp_path / "tri.obj"
    p.write_text("v 0 0 0\nv 2 0 0\nv 0 3 0\nvt 0 0\nvt 1 0\nvt 0 1\nf 1/1 2/2 3/3\n")
    v, t = ioref.obj(str(p))
    assert v.shape == (3, 11) and t.tolist() == [[0, 1, 2]]
    np.testing.assert_array_equal(v[:, 3:6], np.tile([0, 0, 1], (3, 1)))          # unit face normal
    np.testing.assert_array_equal(v[:, 6:9], np.tile([1, 0, 0], (3, 1)))          # dP/du = (2,0,0) normalised
    np.testing.assert_array_equal(v[:, 9:11], [[0, 0], [1, 0], [0, 1]])
    # file normals are used as given (not normalised), tangent orthogonalised against them
    p.write_text("v 0 0 0\nv 1 0 0\nv 0 1 0\nvn 0 0 2\nvt 0 0\nvt 1 0\nvt 0 1\nf 1/1/1 2/2/1 3/3/1\n")
    v, _ = ioref.obj(str(p))
    np.testing.assert_array_equal(v[:, 3:6], np.tile([0, 0, 2], (3, 1)))
    # quad split along the shorter diagonal: 0-2 shorter -> (0,1,2),(0,2,3)
    p.write_text("v 0 0 0\nv 4 0 0\nv 1 1 0\nv 0 4 0\nf 1 2 3 4\n")
    _, t = ioref.obj(str(p))
    assert t.tolist() == [[0, 1, 2], [0, 2, 3]]
    p.write_text("v 0 0 0\nv 1 0 0\nv 4 4 0\nv 0 1 0\nf 1 2 3 4\n")
    v, t = ioref.obj(str(p))
    # (v0, v1, v3), (v1, v2, v3) with vertices numbered in first-use order: v3 -> 2, v2 -> 3
    assert t.tolist() == [[0, 1, 2], [1, 3, 2]]
    np.testing.assert_array_equal(v[:, :3], [[0, 0, 0], [1, 0, 0], [0, 1, 0], [4, 4, 0]])
    # shared corners are deduplicated per (v, vt, vn); a missing uv gives tangent (1,0,0)
    p.write_text("v 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nf 1 2 3\nf 1 3 4\n")
    v, t = ioref.obj(str(p))
    assert v.shape[0] == 4 and t.tolist() == [[0, 1, 2], [0, 2, 3]]
    np.testing.assert_array_equal(v[:, 6:9], np.tile([1, 0, 0], (4, 1)))


@needs_ref
def test_ref_image_closed_forms(tmp_path):
    img = np.arange(2 * 3 * 3, dtype=np.uint8).reshape(2, 3, 3) * 7
    write_png(str(tmp_path / "a.png"), img, 2)
    planes, c = ioref.image_u8(str(tmp_path / "a.png"))
    assert c == 3
    np.testing.assert_array_equal(planes, img.transpose(2, 0, 1))
    rgbe = np.array([[[128, 64, 32, 129]]], np.uint8)                      # 2^(129-136) = 1/128
    write_rgbe(str(tmp_path / "a.hdr"), rgbe, rle=False)
    f = ioref.image_f32(str(tmp_path / "a.hdr"))
    np.testing.assert_array_equal(f[:, 0, 0], [1.0, 0.5, 0.25])


# ---------------------------------------------------------------------------
# product readers vs the reference's
# ---------------------------------------------------------------------------
def _obj_cases(rng):
    cases = {}
    cases["tris_no_normals"] = "v 0 0 0\nv 1 0 0\nv 0 1 0\nv 0 0 1\nf 1 2 3\nf 1 3 4\nf 1 4 2\nf 2 4 3\n"
    cases["quads_uv_normals"] = (
        "o box\n" + "".join(f"v {x} {y} {z}\n" for x in (-1, 1) for y in (-1, 1) for z in (-1, 1))
        + "vt 0 0\nvt 1 0\nvt 1 1\nvt 0 1\n"
        + "vn 1 0 0\nvn -1 0 0\nvn 0 1 0\nvn 0 -1 0\nvn 0 0 1\nvn 0 0 -1\n"
        + "f 5/1/1 7/2/1 8/3/1 6/4/1\nf 1/1/2 2/2/2 4/3/2 3/4/2\ng top\nf 3/1/3 4/2/3 8/3/3 7/4/3\n"
        + "f 1/1/4 5/2/4 6/3/4 2/4/4\nf 2/1/5 6/2/5 8/3/5 4/4/5\nf 1/1/6 3/2/6 7/3/6 5/4/6\n")
    cases["negative_and_slashes"] = ("v 0 0 0\nv 1 0.5 0\nv 0 1 0.25\nvn 0 0 1\nf -3//-1 -2//-1 -1//-1\n"
                                     "v 2 2 2\nvt 0.5 0.5\nf 1/1 2/1 -1/1\n")
    # a noisy grid with uv, quads and pentagon-free, 1e-7-level coordinates
    n = 9
    g = np.stack(np.meshgrid(np.linspace(0, 1, n), np.linspace(0, 1, n), indexing="ij"), -1).reshape(-1, 2)
    z = rng.normal(size=g.shape[0]) * 0.05
    lines = [f"v {a:.9g} {b:.9g} {c:.9g}" for (a, b), c in zip(g, z)]
    lines += [f"vt {a:.7g} {b:.7g}" for a, b in g]
    for i in range(n - 1):
        for j in range(n - 1):
            q = [i * n + j + 1, (i + 1) * n + j + 1, (i + 1) * n + j + 2, i * n + j + 2]
            if (i + j) % 3 == 0:
                lines.append("f " + " ".join(f"{k}/{k}" for k in q))
            else:
                lines.append(f"f {q[0]}/{q[0]} {q[1]}/{q[1]} {q[2]}/{q[2]}")
                lines.append(f"f {q[0]}/{q[0]} {q[2]}/{q[2]} {q[3]}/{q[3]}")
    cases["noisy_grid"] = "\n".join(lines) + "\n"
    cases["crlf_comments"] = "# header\r\nv 0 0 0\r\nv 1 0 0\r\nv 0 1 0\r\n\r\nf 1 2 3\r\n"
    return cases


@needs_ref
def test_obj_matches_reference(tmp_path):
    rng = np.random.default_rng(7)
    for name, text in _obj_cases(rng).items():
        p = tmp_path / f"{name}.obj"
        p.write_text(text)
        v_ref, t_ref = ioref.obj(str(p))
        v, t = pt.io.load_obj(str(p))
        assert t.tolist() == t_ref.tolist(), name
        assert v.shape == v_ref.shape, name
        same = (v == v_ref) | (np.isnan(v) & np.isnan(v_ref))
        assert same.all(), f"{name}: {np.argwhere(~same)[:5]}"


def test_obj_errors(tmp_path):
    p = tmp_path / "bad.obj"
    p.write_text("v 0 0 0\nf 1 2 3\n")
    with pytest.raises(RuntimeError):
        pt.io.load_obj(str(p))
    with pytest.raises(RuntimeError):
        pt.io.load_obj(str(tmp_path / "missing.obj"))


def _png_cases(rng):
    H, W = 7, 11
    yield "rgb8", dict(arr=rng.integers(0, 256, (H, W, 3)), ctype=2)
    yield "rgba8", dict(arr=rng.integers(0, 256, (H, W, 4)), ctype=6)
    yield "rgb16", dict(arr=rng.integers(0, 65536, (H, W, 3)), ctype=2, depth=16)
    pal = rng.integers(0, 256, (16, 3))
    yield "pal8", dict(arr=rng.integers(0, 16, (H, W)), ctype=3, palette=pal)
    yield "pal4_trns", dict(arr=rng.integers(0, 16, (H, W)), ctype=3, depth=4, palette=pal,
                            trns=list(rng.integers(0, 256, 10)))
    rgb = rng.integers(0, 4, (H, W, 3)) * 60
    key = rgb[0, 0]
    yield "rgb8_trns", dict(arr=rgb, ctype=2, trns=b"".join(struct.pack(">H", int(k)) for k in key))
    # Adam7: sizes where some passes are empty (1x1, 3x2) and ragged ones
    for h, w in ((1, 1), (3, 2), (9, 13), (17, 10)):
        yield f"rgb8_adam7_{h}x{w}", dict(arr=rng.integers(0, 256, (h, w, 3)), ctype=2, interlace=True)
    yield "rgba16_adam7", dict(arr=rng.integers(0, 65536, (H, W, 4)), ctype=6, depth=16, interlace=True)
    yield "pal2_adam7", dict(arr=rng.integers(0, 4, (H, W)), ctype=3, depth=2, palette=pal[:4], interlace=True,
                             trns=[0, 128])
    yield "pal1_adam7", dict(arr=rng.integers(0, 2, (10, 19)), ctype=3, depth=1, palette=pal[:2], interlace=True)


@needs_ref
def test_png_matches_reference(tmp_path):
    rng = np.random.default_rng(3)
    for name, kw in _png_cases(rng):
        p = str(tmp_path / f"{name}.png")
        write_png(p, **kw)
        ref, c_ref = ioref.image_u8(p)
        got, c = pt.io.load_image_uint8(p)
        assert c == c_ref, name
        np.testing.assert_array_equal(got, ref, err_msg=name)


def _bad_png_cases(rng):
    """Files the PNG specification forbids.  stb_image rejects each of them except the
    colour-type / bit-depth pairs below 8 bits for RGB(A), which it decodes anyway; the
    product rejects those too (the specification's table of allowed pairs)."""
    rgba = rng.integers(0, 256, (4, 5, 4))
    yield "tRNS on RGBA", dict(arr=rgba, ctype=6, trns=bytes(8))
    yield "tRNS on grey+alpha", dict(arr=rng.integers(0, 256, (4, 5, 2)), ctype=4, trns=bytes(2))
    yield "tRNS of the wrong size", dict(arr=rng.integers(0, 256, (4, 5, 3)), ctype=2, trns=bytes(4))
    pal = rng.integers(0, 256, (4, 3))
    yield "palette at 16 bits", dict(arr=rng.integers(0, 4, (4, 5)), ctype=3, palette=pal, ihdr=(5, 4, 16, 3))
    yield "RGB at 4 bits (stb decodes it)", dict(arr=rng.integers(0, 16, (4, 5, 3)), ctype=2, depth=4)
    yield "RGBA at 2 bits (stb decodes it)", dict(arr=rng.integers(0, 4, (4, 5, 4)), ctype=6, depth=2)
    yield "palette without PLTE", dict(arr=rng.integers(0, 4, (4, 5)), ctype=3)
    yield "width beyond 2^24", dict(arr=rng.integers(0, 256, (1, 2, 3)), ctype=2, ihdr=(1 << 25, 1, 8, 2))
    yield "height 2^31", dict(arr=rng.integers(0, 256, (1, 2, 3)), ctype=2, ihdr=(2, 1 << 31, 8, 2))


def test_png_malformed_rejected(tmp_path):
    rng = np.random.default_rng(9)
    for name, kw in _bad_png_cases(rng):
        p = str(tmp_path / "bad.png")
        write_png(p, **kw)
        with pytest.raises(RuntimeError):
            pt.io.load_image_uint8(p)
        if ioref.available() and "stb decodes it" not in name:
            with pytest.raises(RuntimeError):
                ioref.image_u8(p)


def test_grey_png_replicated(tmp_path):
    """Grey files: the reference reads past its buffer; the product uses three equal planes."""
    rng = np.random.default_rng(4)
    g = rng.integers(0, 256, (5, 6))
    p = str(tmp_path / "g.png")
    write_png(p, g, 0)
    got, c = pt.io.load_image_uint8(p)
    assert c == 3 and got.shape == (3, 5, 6)
    for k in range(3):
        np.testing.assert_array_equal(got[k], g)


@needs_ref
def test_pnm_matches_reference(tmp_path):
    rng = np.random.default_rng(5)
    a = rng.integers(0, 256, (4, 9, 3)).astype(np.uint8)
    p = tmp_path / "a.ppm"
    p.write_bytes(b"P6\n# c\n9 4\n255\n" + a.tobytes())
    ref, c_ref = ioref.image_u8(str(p))
    got, c = pt.io.load_image_uint8(str(p))
    assert c == c_ref == 3
    np.testing.assert_array_equal(got, ref)


@needs_ref
@pytest.mark.parametrize("rle", [False, True])
def test_hdr_matches_reference(tmp_path, rle):
    rng = np.random.default_rng(6)
    rgbe = random_rgbe(rng, 6, 40)
    p = str(tmp_path / "sky.hdr")
    write_rgbe(p, rgbe, rle)
    ref = ioref.image_f32(p)
    got = pt.io.load_image_float(p)
    assert got.shape == ref.shape == (3, 6, 40)
    np.testing.assert_array_equal(got, ref)
    ref8, c8 = ioref.image_u8(p)
    got8, c = pt.io.load_image_uint8(p)
    assert c == c8
    np.testing.assert_array_equal(got8, ref8)


@needs_ref
def test_ldr_as_hdri_matches_reference(tmp_path):
    rng = np.random.default_rng(8)
    for name, ctype, ch in (("rgb", 2, 3), ("rgba", 6, 4)):
        p = str(tmp_path / f"{name}.png")
        write_png(p, rng.integers(0, 256, (5, 8, ch)), ctype)
        np.testing.assert_array_equal(pt.io.load_image_float(p), ioref.image_f32(p), err_msg=name)
