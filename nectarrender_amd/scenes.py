"""Seeded synthetic scenes of the benchmark configurations.

Each builder returns ``(SceneDesc, NrCameraDesc, dict(render params))``.
Camera poses are not given by the workload definitions, so they are fixed
here.  Meshes carry the attributes the reference's OBJ loader derives when a
file has no normals or texture coordinates (csrc/hittable/include/hittable/
mesh.h:103-138): area-unweighted face-normal sums, tangent (1,0,0) projected
off the normal.
"""
from __future__ import annotations

import hashlib

import numpy as np

from . import abi
from .abi import Material, SceneDesc, camera_desc

FMIN = np.float32(1e-8)


# ---------------------------------------------------------------------------
# mesh attributes (parse_obj_file, mesh.h:103-138)
# ---------------------------------------------------------------------------
def mesh_vertices(pos: np.ndarray, tris: np.ndarray, uv: np.ndarray | None = None) -> np.ndarray:
    pos = np.asarray(pos, np.float32)
    tris = np.asarray(tris, np.int64)
    nv = pos.shape[0]
    a, b, c = pos[tris[:, 0]], pos[tris[:, 1]], pos[tris[:, 2]]
    fn = np.cross(b - a, c - a).astype(np.float32)
    fl = np.sqrt((fn * fn).sum(1, keepdims=True)).astype(np.float32)
    fn = fn * (np.float32(1.0) / fl)
    nrm = np.zeros((nv, 3), np.float32)
    for k in range(3):
        np.add.at(nrm, tris[:, k], fn)
    nl = np.sqrt((nrm * nrm).sum(1, keepdims=True)).astype(np.float32)
    nrm = np.where(nl < 1e-8, np.array([0, 1, 0], np.float32), nrm * (np.float32(1.0) / np.maximum(nl, 1e-30)))
    if uv is None:
        uv = np.zeros((nv, 2), np.float32)
        tang_face = np.tile(np.array([1.0, 0.0, 0.0], np.float32), (tris.shape[0], 1))
    else:
        uv = np.asarray(uv, np.float32)
        e1, e2 = b - a, c - a
        d1, d2 = uv[tris[:, 1]] - uv[tris[:, 0]], uv[tris[:, 2]] - uv[tris[:, 0]]
        den = d1[:, 0] * d2[:, 1] - d2[:, 0] * d1[:, 1]
        t = (e1 * d2[:, 1:2] - e2 * d1[:, 1:2]) * (np.float32(1.0) / np.where(np.abs(den) < 1e-8, 1, den))[:, None]
        tang_face = np.where((np.abs(den) < 1e-8)[:, None], np.array([1.0, 0, 0], np.float32), t).astype(np.float32)
    tang = np.zeros((nv, 3), np.float32)
    for k in range(3):
        np.add.at(tang, tris[:, k], tang_face)
    t = tang - nrm * (tang * nrm).sum(1, keepdims=True)
    tl = np.sqrt((t * t).sum(1, keepdims=True)).astype(np.float32)
    tang = np.where(tl < 1e-8, np.array([1, 0, 0], np.float32), t * (np.float32(1.0) / np.maximum(tl, 1e-30)))
    return np.concatenate([pos, nrm, tang, uv], axis=1).astype(np.float32)


# ---------------------------------------------------------------------------
# geodesic icosphere (frequency f: 20 f^2 triangles, 10 f^2 + 2 vertices)
# ---------------------------------------------------------------------------
def _icosahedron():
    t = (1.0 + 5 ** 0.5) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2],
                  [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5], [2, 4, 11],
                  [6, 2, 10], [8, 6, 7], [9, 8, 1]], np.int64)
    return v, f


def geodesic_sphere(freq: int):
    """Unit geodesic sphere: every icosahedron face split into freq^2 triangles."""
    V, F = _icosahedron()
    f = freq
    # local grid of a face: points (i, j) with i + j <= f, barycentric (f-i-j, i, j)/f
    ii, jj = np.meshgrid(np.arange(f + 1), np.arange(f + 1), indexing="ij")
    m = (ii + jj) <= f
    gi, gj = ii[m], jj[m]
    lid = -np.ones((f + 1, f + 1), np.int64)
    lid[gi, gj] = np.arange(gi.size)
    # local triangles
    up = [(lid[i, j], lid[i + 1, j], lid[i, j + 1]) for i in range(f) for j in range(f - i)]
    dn = [(lid[i + 1, j], lid[i + 1, j + 1], lid[i, j + 1]) for i in range(f) for j in range(f - i - 1)]
    lt = np.array(up + dn, np.int64)
    pts = []
    tris = []
    for k, (a, b, c) in enumerate(F):
        w0 = (f - gi - gj) / f
        w1 = gi / f
        w2 = gj / f
        p = w0[:, None] * V[a] + w1[:, None] * V[b] + w2[:, None] * V[c]
        pts.append(p)
        tris.append(lt + k * gi.size)
    P = np.concatenate(pts)
    P /= np.linalg.norm(P, axis=1, keepdims=True)
    T = np.concatenate(tris)
    key = np.round(P * 1e6).astype(np.int64)
    _, first, inv = np.unique(key, axis=0, return_index=True, return_inverse=True)
    inv = inv.reshape(-1)
    P = P[first]
    T = inv[T]
    return P, T


def _hash3(ix, iy, iz, seed):
    h = (ix.astype(np.uint32) * np.uint32(73856093)) ^ (iy.astype(np.uint32) * np.uint32(19349663)) ^ \
        (iz.astype(np.uint32) * np.uint32(83492791)) ^ np.uint32(seed * 2654435761 & 0xFFFFFFFF)
    h ^= h >> np.uint32(13)
    h *= np.uint32(0x5bd1e995)
    h ^= h >> np.uint32(15)
    return (h & np.uint32(0xFFFFFF)).astype(np.float64) / float(0xFFFFFF) * 2.0 - 1.0


def value_noise(p: np.ndarray, seed: int) -> np.ndarray:
    """Trilinear value noise with smoothstep weights, lattice values in [-1, 1]."""
    f = np.floor(p)
    t = p - f
    s = t * t * (3 - 2 * t)
    i = f.astype(np.int64)
    out = np.zeros(p.shape[0])
    for dx in (0, 1):
        for dy in (0, 1):
            for dz in (0, 1):
                w = (s[:, 0] if dx else 1 - s[:, 0]) * (s[:, 1] if dy else 1 - s[:, 1]) * (s[:, 2] if dz else 1 - s[:, 2])
                out += w * _hash3(i[:, 0] + dx, i[:, 1] + dy, i[:, 2] + dz, seed)
    return out


def fbm(p: np.ndarray, seed=0, octaves=4, lacunarity=2.0, gain=0.5, base_freq=3.0) -> np.ndarray:
    out = np.zeros(p.shape[0])
    amp, freq = 1.0, base_freq
    for o in range(octaves):
        out += amp * value_noise(p * freq, seed + o)
        amp *= gain
        freq *= lacunarity
    return out


def displaced_icosphere(freq=224, amp=0.1, seed=0):
    """1,003,520-triangle icosphere (freq 224) radially displaced by 4-octave value-noise fBm."""
    P, T = geodesic_sphere(freq)
    r = 1.0 + amp * fbm(P, seed=seed)
    pos = (P * r[:, None]).astype(np.float32)
    return mesh_vertices(pos, T), T.astype(np.uint32)


def procedural_hdri(W=2048, H=1024, seed=0, sigma=0.5, sun_radiance=1e4, sun_diameter_deg=0.5,
                    sun_elevation_deg=45.0, sun_azimuth_deg=30.0) -> np.ndarray:
    """CHW fp32 equirect: i.i.d. log-normal background + a sun disc.

    Texel (x, y) looks along the direction the reference's lookup maps to it
    (skylight.h:84-88): phi = (u - 0.5) 2 pi, theta = v pi,
    d = (sin theta cos phi, cos theta, sin theta sin phi)."""
    rng = np.random.default_rng(seed)
    img = np.exp(rng.normal(0.0, sigma, size=(3, H, W))).astype(np.float32)
    u = (np.arange(W) + 0.5) / W
    v = (np.arange(H) + 0.5) / H
    phi = (u - 0.5) * 2 * np.pi
    th = v * np.pi
    d = np.stack([np.sin(th)[:, None] * np.cos(phi)[None, :], np.cos(th)[:, None] * np.ones((1, W)),
                  np.sin(th)[:, None] * np.sin(phi)[None, :]])
    el, az = np.radians(sun_elevation_deg), np.radians(sun_azimuth_deg)
    sd = np.array([np.cos(el) * np.cos(az), np.sin(el), np.cos(el) * np.sin(az)])
    cosang = np.tensordot(sd, d, axes=(0, 0))
    sun = cosang >= np.cos(np.radians(sun_diameter_deg / 2.0))
    if not sun.any():   # keep at least the nearest texel lit
        sun = cosang >= cosang.max()
    img[:, sun] = np.float32(sun_radiance)
    return img


def checksum(*arrays) -> str:
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()[:16]


# ---------------------------------------------------------------------------
# configurations
# ---------------------------------------------------------------------------
def config1(spp=16):
    """512x512, 16 spp, depth 8, seed 0: Lambertian ground quad + PBR / dielectric / Lambertian spheres,
    constant sky radiance 1."""
    s = SceneDesc()
    s.quad((0, 0, 0), (0, 0, 0), (10, 1, 10), Material.lambertian((0.5, 0.5, 0.5)))
    s.sphere((-1.1, 0.5, 0), 0.5, Material.pbr(albedo=(1.0, 0.6, 0.15), roughness=0.2, metallic=1.0))
    s.sphere((0, 0.5, 0), 0.5, Material.dielectric((1, 1, 1), 1.5))
    s.sphere((1.1, 0.5, 0), 0.5, Material.lambertian((0.2, 0.4, 0.8)))
    s.sky_simple((1, 1, 1), (1, 1, 1))
    cam = camera_desc((512, 512), (0, 1, 4), (-10, 0, 0), spp, 2.0, 2.0, 0.0, 2.0, 1.0)
    return s, cam, dict(depth=8, seed=0)


def config2(spp=64, freq=224, hdri_w=2048, hdri_h=1024):
    """1080p (H=1080, W=1920), 64 spp, depth 12: displaced 1M-triangle icosphere (PBR metal rough 0.4) on a
    Lambertian quad under a procedural 2048x1024 HDRI with a 1e4-radiance 0.5 degree sun."""
    verts, tris = displaced_icosphere(freq)
    s = SceneDesc()
    m = s.add_mesh(verts, tris)
    s.mesh(m, (0, 0, 0), (0, 0, 0), (1, 1, 1), Material.pbr(albedo=(0.9, 0.9, 0.9), roughness=0.4, metallic=1.0))
    s.quad((0, -1.15, 0), (0, 0, 0), (12, 1, 12), Material.lambertian((0.5, 0.5, 0.5)))
    s.sky_hdri(procedural_hdri(hdri_w, hdri_h))
    # The reference projects x with resolution[0] (camera.h:269-272), which puts the
    # principal point of a 1080x1920 frame at sensor offset (+0.78, +0.78) of the
    # focal plane; the pose below centres the sphere in that frame.
    cam = camera_desc((1080, 1920), (0.0, 0.5, 3.8), (-30.0, 20.15, 0.0), spp, 2.0, 3.8, 0.0, 2.0, 1.0)
    return s, cam, dict(depth=12, seed=0)


def cornell(spp=16, res=128):
    """The reference's config.py scene: five quads, two cubes, one ObjectLight quad, no sky,
    camera aperture 0.01 (defocus on)."""
    s = SceneDesc()
    W = (1.0, 1.0, 1.0)
    s.quad((0, -0.5, 0), (0, 0, 0), (1, 1, 1), Material.lambertian(W))
    s.quad((0, 0.5, 0), (0, 180, 0), (1, 1, 1), Material.lambertian(W))
    s.quad((0.5, 0, 0), (0, 0, 90), (1, 1, 1), Material.lambertian((1, 0, 0)))
    s.quad((-0.5, 0, 0), (0, 0, -90), (1, 1, 1), Material.lambertian((0, 1, 0)))
    s.quad((0, 0, -0.5), (-90, 0, 0), (1, 1, 1), Material.lambertian(W))
    s.cube((0.2, -0.35, 0.2), (0, 35, 0), (0.3, 0.3, 0.3), Material.lambertian(W))
    s.cube((-0.2, -0.2, -0.1), (0, 35, 0), (0.3, 0.6, 0.3), Material.lambertian(W))
    s.object_light(SceneDesc.bound(abi.HIT_QUAD, (0, 0.499, 0), (0, 180, 0), (0.2, 0.2, 0.2)), 35.0, W)
    cam = camera_desc((res, res), (0, 0, 2), (0, 0, 0), spp, 3.0, 10.0, 0.01, 2.0, 1.0)
    return s, cam, dict(depth=5, seed=42)


def mixed_scene(spp=8, res=96, seed=7):
    """Every hittable and material type: quad, cube, spheres, small mesh, constant media (sphere and cube
    bounds), a sphere ObjectLight, image textures, HDRI sky."""
    rng = np.random.default_rng(seed)
    s = SceneDesc()
    img = s.add_image((rng.random((3, 16, 16)) * 255).astype(np.uint8))
    s.quad((0, -0.6, 0), (0, 15, 0), (6, 1, 6), Material.lambertian(abi.Texture(image=img, scale=0.5)))
    s.cube((1.0, -0.2, -0.5), (10, 35, 5), (0.5, 0.7, 0.5), Material.pbr(albedo=(0.8, 0.3, 0.2), roughness=0.5,
                                                                         metallic=0.3))
    s.sphere((-0.9, -0.1, 0.3), 0.45, Material.dielectric((0.9, 1.0, 0.9), 1.45))
    s.sphere((0.2, 0.0, -1.0), 0.5, Material.pbr(albedo=abi.Texture(image=img), roughness=0.15, metallic=1.0,
                                                 emission=(0.05, 0.0, 0.0)))
    P, T = geodesic_sphere(3)
    mv = mesh_vertices((P * 0.35).astype(np.float32), T)
    mi = s.add_mesh(mv, T)
    s.mesh(mi, (-0.2, 0.2, 0.8), (20, 30, 0), (1.0, 0.8, 1.2), Material.lambertian((0.3, 0.8, 0.4)))
    s.constant_medium(SceneDesc.bound(abi.HIT_SPHERE, (0.6, 0.1, 0.9), radius=0.35), 2.0, (0.8, 0.8, 0.9))
    s.constant_medium(SceneDesc.bound(abi.HIT_CUBE, (-1.3, 0.3, -0.8), (0, 20, 0), (0.5, 0.5, 0.5)), 1.0,
                      (0.9, 0.7, 0.5))
    s.sphere((1.6, 0.6, 0.6), 0.2, Material.emissive((1.0, 0.9, 0.7), 8.0))
    s.object_light(SceneDesc.bound(abi.HIT_SPHERE, (-0.5, 1.6, 0.5), radius=0.25), 20.0, (1, 1, 1))
    s.sky_hdri(procedural_hdri(64, 32, seed=seed, sun_diameter_deg=8.0, sun_radiance=5.0), rotation=30.0,
               intensity=0.8)
    cam = camera_desc((res, res), (0, 0.6, 3.5), (-8, 0, 0), spp, 2.0, 3.5, 0.02, 2.0, 1.0)
    return s, cam, dict(depth=6, seed=seed)


def random_rays(n: int, seed: int, center=(0, 0, 0), spread=3.0) -> np.ndarray:
    """Rays from random points on a sphere of radius `spread` around `center` toward random interior points."""
    rng = np.random.default_rng(seed)
    o = rng.normal(size=(n, 3))
    o = o / np.linalg.norm(o, axis=1, keepdims=True) * spread + np.asarray(center)
    tgt = (rng.random((n, 3)) - 0.5) * spread + np.asarray(center)
    d = tgt - o
    return np.concatenate([o, d], axis=1).astype(np.float32)


def look_rotation(position, target):
    """Euler degrees (pitch, yaw, 0) whose R = Rz Ry Rx maps -z onto the direction position -> target."""
    d = np.asarray(target, np.float64) - np.asarray(position, np.float64)
    d /= np.linalg.norm(d)
    pitch = np.degrees(np.arcsin(np.clip(d[1], -1.0, 1.0)))
    yaw = np.degrees(np.arctan2(-d[0], -d[2]))
    return (float(pitch), float(yaw), 0.0)


def config3(n_spheres=10_000, seed=0, scale=1.0, n_views=16):
    """10,000 spheres, centres U[-10,10]^3, radius U[0.05,0.3]: 40% PBR (rough U[0,1], metallic
    Bernoulli 0.5, albedo U[0.2,1]^3), 20% dielectric 1.5, 20% Lambertian (albedo U[0.2,0.9]^3),
    10% emissive (radiance U[1,20]), 10% constant media (sphere bound, density U[0.1,2], albedo 0.8);
    simple sky.  16 views of 256x256 .. 1920x1080 (sum 10,142,016 pixels) looking at the origin from a
    ring of radius 25; `scale` shrinks every view (parity tests)."""
    rng = np.random.default_rng(seed)
    s = SceneDesc()
    c = rng.uniform(-10, 10, (n_spheres, 3))
    r = rng.uniform(0.05, 0.3, n_spheres)
    kind = rng.random(n_spheres)
    for i in range(n_spheres):
        k = kind[i]
        if k < 0.4:
            m = Material.pbr(albedo=tuple(rng.uniform(0.2, 1.0, 3)), roughness=float(rng.random()),
                             metallic=float(rng.random() < 0.5))
            s.sphere(c[i], r[i], m)
        elif k < 0.6:
            s.sphere(c[i], r[i], Material.dielectric((1, 1, 1), 1.5))
        elif k < 0.8:
            s.sphere(c[i], r[i], Material.lambertian(tuple(rng.uniform(0.2, 0.9, 3))))
        elif k < 0.9:
            s.sphere(c[i], r[i], Material.emissive((1, 1, 1), float(rng.uniform(1, 20))))
        else:
            s.constant_medium(SceneDesc.bound(abi.HIT_SPHERE, tuple(c[i]), radius=float(r[i])),
                              float(rng.uniform(0.1, 2.0)), (0.8, 0.8, 0.8))
    s.sky_simple((1, 1, 1), (0.5, 0.7, 1.0))
    sizes = [(256, 256), (320, 256), (384, 288), (448, 320), (512, 384), (576, 432), (640, 480), (720, 540),
             (800, 600), (960, 640), (1024, 768), (1152, 768), (1280, 800), (1440, 900), (1600, 900), (1920, 1080)]
    cams = []
    for v, (w, h) in enumerate(sizes[:n_views]):
        ang = 2 * np.pi * v / len(sizes)
        pos = (25 * np.cos(ang), 4.0 * np.sin(3 * ang), 25 * np.sin(ang))
        res = (max(8, int(round(h * scale))), max(8, int(round(w * scale))))
        cams.append(camera_desc(res, pos, look_rotation(pos, (0, 0, 0)), 32, 2.0, 25.0, 0.0, 2.0, 1.0))
    return s, cams, dict(depth=16, seed=0)


def config5(n_spheres=500, seed=0, res=(1080, 1920), spp=4096):
    """The 1080p three-sphere scene (config1's ground, spheres and constant unit sky) plus 500 spheres in
    the style of config3, half constant media (density U[0.1,2], albedo 0.8) and half dielectric 1.5,
    radius U[0.05,0.3], spread over the visible region; depth 8."""
    rng = np.random.default_rng(seed + 5)
    s = SceneDesc()
    s.quad((0, 0, 0), (0, 0, 0), (10, 1, 10), Material.lambertian((0.5, 0.5, 0.5)))
    s.sphere((-1.1, 0.5, 0), 0.5, Material.pbr(albedo=(1.0, 0.6, 0.15), roughness=0.2, metallic=1.0))
    s.sphere((0, 0.5, 0), 0.5, Material.dielectric((1, 1, 1), 1.5))
    s.sphere((1.1, 0.5, 0), 0.5, Material.lambertian((0.2, 0.4, 0.8)))
    for i in range(n_spheres):
        c = (float(rng.uniform(-3.5, 3.5)), float(rng.uniform(0.05, 2.5)), float(rng.uniform(-5.0, 1.5)))
        r = float(rng.uniform(0.05, 0.3))
        if i % 2 == 0:
            s.constant_medium(SceneDesc.bound(abi.HIT_SPHERE, c, radius=r), float(rng.uniform(0.1, 2.0)),
                              (0.8, 0.8, 0.8))
        else:
            s.sphere(c, r, Material.dielectric((1, 1, 1), 1.5))
    s.sky_simple((1, 1, 1), (1, 1, 1))
    # same framing correction as config2 for the reference's res[0]-based x projection
    cam = camera_desc(res, (0.0, 1.2, 4.8), (-30.0, 20.15, 0.0), spp, 2.0, 4.8, 0.0, 2.0, 1.0)
    return s, cam, dict(depth=8, seed=0)


def training_scene_hdri_mesh(seed: int, res=256, spp=1024):
    """Randomised scenes in the style of config2 for the denoiser: a displaced icosphere (frequency
    12-40, fBm amplitude 0.05-0.15) of PBR metal or plastic on a Lambertian ground under a procedural
    HDRI with a small, very bright sun (radiance 1e3-1e4, 0.5-2 degrees), a camera on a ring."""
    rng = np.random.default_rng(20_000 + seed)
    s = SceneDesc()
    v, t = displaced_icosphere(freq=int(rng.integers(12, 41)), amp=float(rng.uniform(0.05, 0.15)), seed=seed)
    m = s.add_mesh(v, t)
    s.mesh(m, (0, 0, 0), tuple(rng.uniform(0, 90, 3)), (1, 1, 1),
           Material.pbr(albedo=tuple(rng.uniform(0.3, 1.0, 3)), roughness=float(rng.uniform(0.1, 0.7)),
                        metallic=float(rng.random() < 0.7)))
    s.quad((0, -1.15, 0), (0, 0, 0), (12, 1, 12), Material.lambertian(tuple(rng.uniform(0.2, 0.7, 3))))
    s.sky_hdri(procedural_hdri(256, 128, seed=seed, sigma=float(rng.uniform(0.2, 0.6)),
                               sun_radiance=float(rng.uniform(1e3, 1e4)), sun_diameter_deg=float(rng.uniform(0.5, 2.0)),
                               sun_elevation_deg=float(rng.uniform(15, 70)), sun_azimuth_deg=float(rng.uniform(0, 360))))
    ang = rng.uniform(0, 2 * np.pi)
    dist = rng.uniform(2.8, 5.0)
    pos = (dist * np.cos(ang), float(rng.uniform(-0.3, 1.8)), dist * np.sin(ang))
    cam = camera_desc((res, res), pos, look_rotation(pos, (0.0, 0.0, 0.0)), spp, float(rng.uniform(1.5, 3.0)), dist,
                      0.0, 2.0, 1.0)
    return s, cam, dict(depth=12, seed=int(rng.integers(1, 1 << 30)))


def training_scene(seed: int, res=256, spp=1024):
    """Randomised scenes in the style of config1/config5 for the denoiser: a ground quad, 3-40 spheres of
    every material (some inside constant media), sometimes a small mesh, a simple or HDRI sky, a camera
    on a ring looking at the middle."""
    rng = np.random.default_rng(10_000 + seed)
    s = SceneDesc()
    g = tuple(rng.uniform(0.2, 0.8, 3))
    s.quad((0, 0, 0), (0, float(rng.uniform(0, 90)), 0), (12, 1, 12), Material.lambertian(g))
    n = int(rng.integers(3, 40))
    for _ in range(n):
        r = float(rng.uniform(0.08, 0.7))
        c = (float(rng.uniform(-2.5, 2.5)), r * float(rng.uniform(0.8, 3.0)), float(rng.uniform(-2.5, 2.5)))
        k = rng.random()
        if k < 0.3:
            s.sphere(c, r, Material.pbr(albedo=tuple(rng.uniform(0.2, 1.0, 3)), roughness=float(rng.random()),
                                        metallic=float(rng.random() < 0.5)))
        elif k < 0.5:
            s.sphere(c, r, Material.dielectric(tuple(rng.uniform(0.8, 1.0, 3)), float(rng.uniform(1.3, 1.8))))
        elif k < 0.75:
            s.sphere(c, r, Material.lambertian(tuple(rng.uniform(0.1, 0.9, 3))))
        elif k < 0.82:
            s.sphere(c, r * 0.4, Material.emissive(tuple(rng.uniform(0.5, 1.0, 3)), float(rng.uniform(2, 20))))
        else:
            s.constant_medium(SceneDesc.bound(abi.HIT_SPHERE, c, radius=r), float(rng.uniform(0.1, 3.0)),
                              tuple(rng.uniform(0.5, 0.95, 3)))
    if rng.random() < 0.3:
        P, T = geodesic_sphere(4)
        mi = s.add_mesh(mesh_vertices((P * 0.5).astype(np.float32), T), T)
        s.mesh(mi, (float(rng.uniform(-1, 1)), 0.6, float(rng.uniform(-1, 1))), tuple(rng.uniform(0, 90, 3)),
               tuple(rng.uniform(0.6, 1.4, 3)), Material.pbr(albedo=tuple(rng.uniform(0.3, 1, 3)),
                                                             roughness=float(rng.uniform(0.1, 0.8)), metallic=0.0))
    if rng.random() < 0.5:
        top = tuple(rng.uniform(0.3, 1.2, 3))
        s.sky_simple(top, top if rng.random() < 0.5 else tuple(rng.uniform(0.3, 1.2, 3)))
    else:
        s.sky_hdri(procedural_hdri(128, 64, seed=seed, sigma=0.3, sun_radiance=float(rng.uniform(20, 400)),
                                   sun_diameter_deg=6.0, sun_elevation_deg=float(rng.uniform(15, 70)),
                                   sun_azimuth_deg=float(rng.uniform(0, 360))),
                   rotation=float(rng.uniform(0, 360)), intensity=float(rng.uniform(0.3, 1.0)))
    ang = rng.uniform(0, 2 * np.pi)
    dist = rng.uniform(3.0, 6.0)
    pos = (dist * np.cos(ang), float(rng.uniform(0.4, 2.5)), dist * np.sin(ang))
    tgt = (float(rng.uniform(-0.5, 0.5)), float(rng.uniform(0.2, 0.8)), float(rng.uniform(-0.5, 0.5)))
    cam = camera_desc((res, res), pos, look_rotation(pos, tgt), spp, float(rng.uniform(1.5, 3.0)), dist,
                      float(rng.choice([0.0, 0.0, 0.05])), 2.0, 1.0)
    return s, cam, dict(depth=8, seed=int(rng.integers(1, 1 << 30)))
