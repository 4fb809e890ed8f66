"""Seeded synthetic scenes and HDRIs for the benchmark configurations.

Each generator returns a SceneDesc that both the renderer (through
integrator.upload_scene) and the oracle consume, so the two see identical
float32 inputs.  Conventions follow the reference: Y up, HDRI row 0 at the
zenith, u = 0.5 looking down +Z (hdridome.h:455-476).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from .integrator import (Camera, DiffuseMtl, DirectionalEmitter, GlassMtl, HDRIDome, ImageTexture2D, MetalMtl,
                         TriangleMesh, View)


@dataclass
class SceneDesc:
    name: str
    meshes: List[TriangleMesh]
    instance_mesh: List[int]
    instance_material: List[int]
    instance_xform: np.ndarray  # (n, 3, 4) float32 object -> world
    materials: list
    emitters: list
    camera: Camera
    width: int
    height: int
    spp: int
    max_depth: int
    rr_depth: int
    view: int = 0
    samples_per_batch: int = 0
    info: dict = field(default_factory=dict)

    @property
    def hdri(self) -> Optional[np.ndarray]:
        for e in self.emitters:
            if isinstance(e, HDRIDome):
                return e.texture.image
        return None

    def num_triangles(self) -> int:
        return int(sum(self.meshes[m].triangles.shape[0] for m in self.instance_mesh))


# ------------------------------------------------------------------ camera

def look_at(eye, at, up, fov_deg: float, width: int, height: int, reference_cosfovy: bool = False) -> Camera:
    """Pinhole camera basis.  The default maps the vertical field of view
    with 2 tan(fov/2); reference_cosfovy=True reproduces CLViewer's basis
    (src/colvillea/CLViewer.cpp:244-260, Camera.h:60), which scales by
    cos(fovy) instead."""
    eye = np.asarray(eye, np.float64)
    fwd = np.asarray(at, np.float64) - eye
    fwd /= np.linalg.norm(fwd)
    aspect = width / float(height)
    s = math.cos(math.radians(fov_deg)) if reference_cosfovy else 2.0 * math.tan(math.radians(fov_deg) * 0.5)
    right = np.cross(fwd, np.asarray(up, np.float64))
    right /= np.linalg.norm(right)
    ddu = s * aspect * right
    upv = np.cross(ddu, fwd)
    upv /= np.linalg.norm(upv)
    ddv = s * upv
    d00 = fwd - 0.5 * ddu - 0.5 * ddv
    f = lambda v: np.asarray(v, np.float32)
    return Camera(f(eye), f(d00), f(ddu), f(ddv))


# ------------------------------------------------------------------ meshes

def icosphere(subdiv: int):
    """Unit icosphere: 20 * 4**subdiv triangles, normals = positions."""
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = np.array([(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
                  (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)], np.float64)
    f = np.array([(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
                  (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5),
                  (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)], np.int64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    for _ in range(subdiv):
        e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
        e.sort(axis=1)
        uniq, inv = np.unique(e, axis=0, return_inverse=True)
        inv = inv.reshape(-1)
        mid = v[uniq[:, 0]] + v[uniq[:, 1]]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        base = v.shape[0]
        v = np.concatenate([v, mid])
        nf = f.shape[0]
        ab, bc, ca = inv[:nf] + base, inv[nf:2 * nf] + base, inv[2 * nf:] + base
        a, b, c = f[:, 0], f[:, 1], f[:, 2]
        f = np.concatenate([np.stack([a, ab, ca], 1), np.stack([b, bc, ab], 1), np.stack([c, ca, bc], 1),
                            np.stack([ab, bc, ca], 1)])
    v32 = v.astype(np.float32)
    return TriangleMesh(v32, f.astype(np.uint32), normals=v32.copy())


def quad_y0(half: float) -> TriangleMesh:
    v = np.array([(-half, 0, -half), (half, 0, -half), (half, 0, half), (-half, 0, half)], np.float32)
    n = np.tile(np.array([[0, 1, 0]], np.float32), (4, 1))
    return TriangleMesh(v, np.array([(0, 2, 1), (0, 3, 2)], np.uint32), normals=n)


def value_noise_fbm(x: np.ndarray, z: np.ndarray, rng: np.random.Generator, octaves: int = 4) -> np.ndarray:
    """Fractional Brownian motion of smoothed value noise (lattice values
    drawn from `rng`), amplitude halving and frequency doubling per octave."""
    out = np.zeros_like(x, dtype=np.float64)
    amp, freq, norm = 1.0, 1.0, 0.0
    for _ in range(octaves):
        xs, zs = x * freq, z * freq
        x0, z0 = np.floor(xs), np.floor(zs)
        ox, oz = x0.min(), z0.min()
        nx, nz = int(x0.max() - ox) + 2, int(z0.max() - oz) + 2
        lat = rng.uniform(-1.0, 1.0, size=(nz, nx))
        ix, iz = (x0 - ox).astype(np.int64), (z0 - oz).astype(np.int64)
        fx, fz = xs - x0, zs - z0
        sx, sz = fx * fx * (3 - 2 * fx), fz * fz * (3 - 2 * fz)
        a = lat[iz, ix] * (1 - sx) + lat[iz, ix + 1] * sx
        b = lat[iz + 1, ix] * (1 - sx) + lat[iz + 1, ix + 1] * sx
        out += amp * (a * (1 - sz) + b * sz)
        norm += amp
        amp *= 0.5
        freq *= 2.0
    return out / norm


def fbm_grid(nx: int, nz: int, half_x: float, half_z: float, height: float, scale: float,
             rng: np.random.Generator) -> TriangleMesh:
    xs = np.linspace(-half_x, half_x, nx)
    zs = np.linspace(-half_z, half_z, nz)
    X, Z = np.meshgrid(xs, zs)
    H = height * value_noise_fbm(X / scale, Z / scale, rng)
    dx, dz = xs[1] - xs[0], zs[1] - zs[0]
    gx = np.gradient(H, dx, axis=1)
    gz = np.gradient(H, dz, axis=0)
    n = np.stack([-gx, np.ones_like(H), -gz], -1)
    n /= np.linalg.norm(n, axis=-1, keepdims=True)
    v = np.stack([X, H, Z], -1).reshape(-1, 3).astype(np.float32)
    i = np.arange(nx * nz, dtype=np.int64).reshape(nz, nx)
    a, b, c, d = i[:-1, :-1], i[:-1, 1:], i[1:, 1:], i[1:, :-1]
    tris = np.concatenate([np.stack([a, c, b], -1).reshape(-1, 3), np.stack([a, d, c], -1).reshape(-1, 3)])
    return TriangleMesh(v, tris.astype(np.uint32), normals=n.reshape(-1, 3).astype(np.float32)), H


def translate_scale(center, radius) -> np.ndarray:
    m = np.zeros((3, 4), np.float32)
    m[0, 0] = m[1, 1] = m[2, 2] = radius
    m[:, 3] = center
    return m


IDENTITY = np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 1, 0]], np.float32)

# Complex IOR (eta, k) at roughly 650/550/450 nm for common conductors
# (published optical constants: Au, Ag, Cu, Al, Cr).
CONDUCTOR_TABLE = [
    ((0.143, 0.374, 1.442), (3.983, 2.385, 1.603)),
    ((0.155, 0.117, 0.138), (4.828, 3.122, 2.147)),
    ((0.200, 0.924, 1.102), (3.912, 2.452, 2.142)),
    ((1.657, 0.880, 0.521), (9.224, 6.270, 4.837)),
    ((3.105, 3.183, 2.312), (3.337, 3.331, 3.137)),
]


# ------------------------------------------------------------------ HDRIs

def hdri_directions(w: int, h: int, rows=None) -> np.ndarray:
    """Direction of every texel centre under the dome mapping
    (HDRIDome::sampleDirect, hdridome.h:207-222); `rows` = (j0, j1)
    restricts it to texel rows [j0, j1)."""
    j0, j1 = rows if rows is not None else (0, h)
    u = (np.arange(w) + 0.5) / w
    v = (np.arange(j0, j1) + 0.5) / h
    U, V = np.meshgrid(u, v)
    theta, phi = np.pi * V, -2.0 * np.pi * (U - 0.5)
    return np.stack([np.sin(theta) * np.sin(phi), np.cos(theta), np.sin(theta) * np.cos(phi)], -1)


def sky_hdri(w: int, h: int, sun_dir_to: np.ndarray, sun_peak: float = 40.0, sun_sigma_deg: float = 2.0) -> np.ndarray:
    """Sky gradient plus a Gaussian sun lobe (RGBA32F)."""
    d = hdri_directions(w, h)
    y = d[..., 1:2]
    zenith = np.array([0.25, 0.45, 1.0]) * 0.8
    horizon = np.array([1.0, 0.95, 0.85]) * 0.9
    ground = np.array([0.18, 0.16, 0.14]) * 0.6
    t = np.sqrt(np.clip(y, 0, 1))
    sky = horizon * (1 - t) + zenith * t
    col = np.where(y >= 0, sky, ground * (1 + y))
    s = np.asarray(sun_dir_to, np.float64)
    s = s / np.linalg.norm(s)
    ang = np.arccos(np.clip(d @ s, -1, 1))
    sig = math.radians(sun_sigma_deg)
    col = col + sun_peak * np.exp(-0.5 * (ang / sig) ** 2)[..., None] * np.array([1.0, 0.9, 0.75])
    out = np.ones((h, w, 4), np.float32)
    out[..., :3] = col
    return out


def lognormal_hdri(w: int, h: int, sigma: float, rng: np.random.Generator, hot_factor: float = 1e4) -> np.ndarray:
    """Smooth log-normal luminance (std `sigma` in log space) with a colour
    tint, plus one hot spot whose peak is `hot_factor` x the mean."""
    cw, ch = 32, 16
    g = rng.standard_normal((ch + 1, cw + 1))
    g[:, -1] = g[:, 0]  # wrap in u
    ys, xs = np.linspace(0, ch, h), np.linspace(0, cw, w)
    y0 = np.minimum(np.floor(ys).astype(int), ch - 1)
    x0 = np.minimum(np.floor(xs).astype(int), cw - 1)
    fy, fx = (ys - y0)[:, None], (xs - x0)[None, :]
    # evaluated in row blocks (the same per-texel operations) so that the
    # temporaries stay cache-sized at 8k x 4k
    B = 128
    f = np.empty((h, w))
    for r in range(0, h, B):
        y, ry = y0[r:r + B], fy[r:r + B]
        f[r:r + B] = (g[y][:, x0] * (1 - fx) * (1 - ry) + g[y][:, x0 + 1] * fx * (1 - ry)
                      + g[y + 1][:, x0] * (1 - fx) * ry + g[y + 1][:, x0 + 1] * fx * ry)
    f = (f - f.mean()) / f.std()
    lum = np.exp(sigma * f)
    tint = 0.75 + 0.25 * rng.uniform(size=(ch + 1, cw + 1, 3))
    col = np.empty((h, w, 3))
    for r in range(0, h, B):
        col[r:r + B] = lum[r:r + B, :, None] * tint[y0[r:r + B]][:, x0]
    mean = col.mean()
    th, ph = rng.uniform(0.25 * np.pi, 0.45 * np.pi), rng.uniform(-np.pi, np.pi)
    hot = np.array([np.sin(th) * np.sin(ph), np.cos(th), np.sin(th) * np.cos(ph)])
    sig = 2.0 * np.pi / w * 2.0
    # the lobe only in the texel rows within 12 sigma of the hot spot's
    # polar angle: beyond them it is below 1e4 * exp(-72) of the mean, which
    # leaves every float64 sum unchanged
    j0 = max(0, int(math.floor((th - 12 * sig) / np.pi * h)) - 1)
    j1 = min(h, int(math.ceil((th + 12 * sig) / np.pi * h)) + 1)
    d = hdri_directions(w, h, (j0, j1))
    ang = np.arccos(np.clip(d @ hot, -1, 1))
    col[j0:j1] = col[j0:j1] + hot_factor * mean * np.exp(-0.5 * (ang / sig) ** 2)[..., None]
    out = np.ones((h, w, 4), np.float32)
    for r in range(0, h, B):
        out[r:r + B, :, :3] = col[r:r + B]
    return out


# ------------------------------------------------------------------ configs

def make_c0(seed: int = 1, width: int = 512, height: int = 512, spp: int = 16, hdri_w: int = 512,
            hdri_h: int = 256) -> SceneDesc:
    """8x8 ground quad + 64 icospheres (subdivision 3, 1280 triangles each),
    diffuse albedo U[0.2, 0.8], sun of 0.5 deg radius and radiance 5, and a
    procedural sky HDRI.  512x512, 16 spp, depth 5, Russian roulette from 3."""
    rng = np.random.default_rng(seed)
    ground, sphere = quad_y0(4.0), icosphere(3)
    centres = np.stack([rng.uniform(-4, 4, 64), rng.uniform(0.5, 1.5, 64), rng.uniform(-4, 4, 64)], 1)
    radii = rng.uniform(0.2, 0.6, 64)
    albedo = rng.uniform(0.2, 0.8, (64, 3))
    materials = [DiffuseMtl((0.5, 0.5, 0.5))] + [DiffuseMtl(tuple(a)) for a in albedo]
    xf = np.stack([IDENTITY] + [translate_scale(c, r) for c, r in zip(centres, radii)])
    el, az = math.radians(50.0), math.radians(30.0)
    to_sun = np.array([math.cos(el) * math.sin(az), math.sin(el), math.cos(el) * math.cos(az)])
    sun = DirectionalEmitter((5.0, 5.0, 5.0), tuple(-to_sun), math.radians(0.5))
    dome = HDRIDome(ImageTexture2D(sky_hdri(hdri_w, hdri_h, to_sun)))
    cam = look_at((0.0, 5.5, 11.0), (0.0, 0.5, 0.0), (0, 1, 0), 45.0, width, height)
    return SceneDesc("C0", [ground, sphere], [0] + [1] * 64, list(range(65)), xf.astype(np.float32), materials,
                     [sun, dome], cam, width, height, spp, 5, 3,
                     info={"seed": seed, "triangles": 2 + 64 * 1280})


def make_c1(seed: int = 2, width: int = 1920, height: int = 1080, spp: int = 64, grid=(1024, 512),
            n_spheres: int = 200, sphere_subdiv: int = 4, hdri_w: int = 2048, hdri_h: int = 1024) -> SceneDesc:
    """fBm-displaced grid (4 octaves) + instanced icospheres (subdivision 4),
    materials 1/3 diffuse / GGX conductor (alpha U[0.05, 0.5], complex IOR
    from a seeded table) / GGX dielectric (IOR 1.5, alpha U[0.01, 0.3]);
    log-normal HDRI (sigma 1.5) with a hot spot at 1e4 x the mean.
    1920x1080, 64 spp, depth 8."""
    rng = np.random.default_rng(seed)
    grid_mesh, H = fbm_grid(grid[0], grid[1], 16.0, 8.0, 1.2, 4.0, rng)
    sphere = icosphere(sphere_subdiv)
    cx = rng.uniform(-14, 14, n_spheres)
    cz = rng.uniform(-7, 7, n_spheres)
    r = rng.uniform(0.25, 0.9, n_spheres)
    gi = np.clip(((cx + 16.0) / 32.0 * (grid[0] - 1)).round().astype(int), 0, grid[0] - 1)
    gj = np.clip(((cz + 8.0) / 16.0 * (grid[1] - 1)).round().astype(int), 0, grid[1] - 1)
    cy = H[gj, gi] + r * rng.uniform(0.3, 1.2, n_spheres)
    materials = [DiffuseMtl((0.45, 0.42, 0.38))]
    for i in range(n_spheres):
        kind = i % 3
        if kind == 0:
            materials.append(DiffuseMtl(tuple(rng.uniform(0.2, 0.8, 3))))
        elif kind == 1:
            eta, k = CONDUCTOR_TABLE[int(rng.integers(len(CONDUCTOR_TABLE)))]
            materials.append(MetalMtl((1.0, 1.0, 1.0), float(np.sqrt(rng.uniform(0.05, 0.5))), eta, k))
        else:
            materials.append(GlassMtl(float(np.sqrt(rng.uniform(0.01, 0.3))), 1.5))
    xf = np.stack([IDENTITY] + [translate_scale((x, y, z), rr) for x, y, z, rr in zip(cx, cy, cz, r)])
    dome = HDRIDome(ImageTexture2D(lognormal_hdri(hdri_w, hdri_h, 1.5, rng)))
    cam = look_at((0.0, 7.0, 20.0), (0.0, 0.5, 0.0), (0, 1, 0), 45.0, width, height)
    ntri = grid_mesh.triangles.shape[0] + n_spheres * sphere.triangles.shape[0]
    return SceneDesc("C1", [grid_mesh, sphere], [0] + [1] * n_spheres, list(range(n_spheres + 1)),
                     xf.astype(np.float32), materials, [dome], cam, width, height, spp, 8, 0,
                     info={"seed": seed, "triangles": int(ntri)})


def make_c4(seed: int = 5, width: int = 3840, height: int = 2160, spp: int = 128, grid=(4096, 2048),
            n_spheres: int = 2000, sphere_subdiv: int = 4, hdri_w: int = 8192, hdri_h: int = 4096,
            max_depth: int = 12) -> SceneDesc:
    """The stress configuration: a 4096 x 2048 fBm grid (4 octaves) over
    64 x 32 units and 2000 instanced icospheres (subdivision 4, radius
    U[0.15, 0.5]), half of them GGX dielectric (IOR 1.5, alpha U[0.01, 0.3]), a quarter diffuse and a
    quarter GGX conductor; an 8192 x 4096 log-normal HDRI (sigma 1.5, hot
    spot at 1e4 x the mean).  3840x2160, 128 spp, depth 12."""
    rng = np.random.default_rng(seed)
    hx, hz = 32.0, 16.0
    grid_mesh, H = fbm_grid(grid[0], grid[1], hx, hz, 2.0, 6.0, rng)
    sphere = icosphere(sphere_subdiv)
    cx = rng.uniform(-0.9 * hx, 0.9 * hx, n_spheres)
    cz = rng.uniform(-0.9 * hz, 0.9 * hz, n_spheres)
    r = rng.uniform(0.15, 0.5, n_spheres)
    gi = np.clip(((cx + hx) / (2 * hx) * (grid[0] - 1)).round().astype(int), 0, grid[0] - 1)
    gj = np.clip(((cz + hz) / (2 * hz) * (grid[1] - 1)).round().astype(int), 0, grid[1] - 1)
    cy = H[gj, gi] + r * rng.uniform(0.3, 1.2, n_spheres)
    materials = [DiffuseMtl((0.45, 0.42, 0.38))]
    for i in range(n_spheres):
        kind = i % 4
        if kind == 1:
            materials.append(DiffuseMtl(tuple(rng.uniform(0.2, 0.8, 3))))
        elif kind == 3:
            eta, k = CONDUCTOR_TABLE[int(rng.integers(len(CONDUCTOR_TABLE)))]
            materials.append(MetalMtl((1.0, 1.0, 1.0), float(np.sqrt(rng.uniform(0.05, 0.5))), eta, k))
        else:
            materials.append(GlassMtl(float(np.sqrt(rng.uniform(0.01, 0.3))), 1.5))
    xf = np.stack([IDENTITY] + [translate_scale((x, y, z), rr) for x, y, z, rr in zip(cx, cy, cz, r)])
    dome = HDRIDome(ImageTexture2D(lognormal_hdri(hdri_w, hdri_h, 1.5, rng)))
    cam = look_at((0.0, 18.0, 20.0), (0.0, 0.0, -2.0), (0, 1, 0), 55.0, width, height)
    ntri = grid_mesh.triangles.shape[0] + n_spheres * sphere.triangles.shape[0]
    return SceneDesc("C4", [grid_mesh, sphere], [0] + [1] * n_spheres, list(range(n_spheres + 1)),
                     xf.astype(np.float32), materials, [dome], cam, width, height, spp, max_depth, 0,
                     info={"seed": seed, "triangles": int(ntri)})


def checker_texture(n: int, a, b, cells: int = 8) -> ImageTexture2D:
    i, j = np.meshgrid(np.arange(n), np.arange(n))
    on = ((i * cells // n + j * cells // n) % 2 == 0)[..., None]
    return ImageTexture2D(np.where(on, np.asarray(a, np.float32), np.asarray(b, np.float32)).astype(np.float32))


def bump_normal_map(n: int, amp: float = 0.4, waves: int = 4) -> ImageTexture2D:
    """A tangent-space normal map of sinusoidal bumps, encoded (n + 1) / 2."""
    i, j = np.meshgrid(np.arange(n), np.arange(n))
    x = amp * np.sin(2 * np.pi * waves * i / n)
    y = amp * np.cos(2 * np.pi * waves * j / n)
    v = np.stack([x, y, np.ones_like(x)], -1)
    v /= np.linalg.norm(v, axis=-1, keepdims=True)
    return ImageTexture2D(((v + 1.0) * 0.5).astype(np.float32))


def make_textured(seed: int = 6, width: int = 320, height: int = 240, spp: int = 16, hdri_w: int = 256,
                  hdri_h: int = 128) -> SceneDesc:
    """Texture-driven materials: a uv-mapped, tangent-carrying ground with a
    checker reflectance and a bump normal map; icospheres (no uvs: the
    barycentric fallback; no tangents: the frame from the normal) with a
    checker diffuse, a metal with reflectance and roughness textures and a
    normal map, a normal-mapped diffuse and a normal-mapped rough glass;
    sky HDRI and a sun.  Depth 6, Russian roulette from 3."""
    rng = np.random.default_rng(seed)
    half = 6.0
    q = quad_y0(half)
    uv = ((q.vertices[:, [0, 2]] + half) / (2 * half) * 3.0).astype(np.float32)  # wraps 3 times
    ground = TriangleMesh(q.vertices, q.triangles, normals=q.normals,
                          tangents=np.tile(np.array([[1, 0, 0]], np.float32), (4, 1)), uvs=uv)
    sphere = icosphere(3)
    rough = rng.uniform(0.05, 0.7, (16, 16)).astype(np.float32)
    g = DiffuseMtl(checker_texture(64, (0.7, 0.7, 0.65), (0.15, 0.2, 0.3)))
    g.setNormalmap(bump_normal_map(64))
    eta, k = CONDUCTOR_TABLE[0]
    metal = MetalMtl(checker_texture(32, (0.95, 0.8, 0.5), (0.6, 0.6, 0.6), cells=4),
                     ImageTexture2D(np.repeat(rough[..., None], 3, 2)), eta, k)
    metal.setNormalmap(bump_normal_map(32, 0.3, 2))
    bumpy = DiffuseMtl((0.6, 0.3, 0.2))
    bumpy.setNormalmap(bump_normal_map(32, 0.6, 3))
    glass = GlassMtl(0.3, 1.5)
    glass.setNormalmap(bump_normal_map(32, 0.2, 2))
    materials = [g, DiffuseMtl(checker_texture(32, (0.8, 0.1, 0.1), (0.9, 0.9, 0.9))), metal, bumpy, glass]
    centres = [(-2.2, 0.9, 0.0), (-0.7, 0.9, 0.6), (0.8, 0.9, -0.3), (2.3, 0.9, 0.4)]
    xf = np.stack([IDENTITY] + [translate_scale(c, 0.85) for c in centres])
    el, az = math.radians(40.0), math.radians(-35.0)
    to_sun = np.array([math.cos(el) * math.sin(az), math.sin(el), math.cos(el) * math.cos(az)])
    sun = DirectionalEmitter((4.0, 3.8, 3.5), tuple(-to_sun), math.radians(0.6))
    dome = HDRIDome(ImageTexture2D(sky_hdri(hdri_w, hdri_h, to_sun, sun_peak=5.0)))
    cam = look_at((0.0, 3.2, 7.5), (0.0, 0.6, 0.0), (0, 1, 0), 42.0, width, height)
    return SceneDesc("textured", [ground, sphere], [0, 1, 1, 1, 1], [0, 1, 2, 3, 4], xf.astype(np.float32),
                     materials, [sun, dome], cam, width, height, spp, 6, 3, info={"seed": seed})


def crop(desc: SceneDesc, x0: int, y0: int, w: int, h: int, spp: Optional[int] = None) -> SceneDesc:
    """The same scene seen through the pixel rectangle [x0, x0 + w) x
    [y0, y0 + h) of desc's film, as a w x h film of its own (the camera's
    pixel plane restricted to the rectangle)."""
    from dataclasses import replace
    c = desc.camera
    d00, ddu, ddv = (np.asarray(v, np.float64) for v in (c.m_camera_d00, c.m_camera_ddu, c.m_camera_ddv))
    f = lambda v: np.asarray(v, np.float32)
    cam = Camera(f(c.m_camera_pos), f(d00 + ddu * (x0 / desc.width) + ddv * (y0 / desc.height)),
                 f(ddu * (w / desc.width)), f(ddv * (h / desc.height)))
    return replace(desc, camera=cam, width=w, height=h, spp=desc.spp if spp is None else spp,
                   info=dict(desc.info, crop=(x0, y0, w, h)))


def make_c2_views(seed: int = 3, n_views: int = 256, side=(64, 512), spp=(4, 32), depth=(3, 10),
                  rr_depth: int = 0) -> List[View]:
    """Camera views of the 1080p scene (make_c1) for a batched render:
    square films with side log-uniform in [64, 512], spp uniform in
    {4..32}, max depth uniform in {3..10}; eyes on a seeded orbit around
    the terrain looking at points near its centre.  View i uses RNG lane i."""
    rng = np.random.default_rng(seed)
    views = []
    for i in range(n_views):
        s = int(round(math.exp(rng.uniform(math.log(side[0]), math.log(side[1])))))
        n = int(rng.integers(spp[0], spp[1] + 1))
        d = int(rng.integers(depth[0], depth[1] + 1))
        az = rng.uniform(0.0, 2.0 * math.pi)
        el = rng.uniform(0.15, 0.7)
        dist = rng.uniform(8.0, 22.0)
        eye = (dist * math.cos(el) * math.sin(az), dist * math.sin(el) + 1.0, dist * math.cos(el) * math.cos(az))
        at = (rng.uniform(-6.0, 6.0), rng.uniform(0.0, 1.5), rng.uniform(-3.0, 3.0))
        fov = rng.uniform(30.0, 60.0)
        views.append(View(s, s, n, d, look_at(eye, at, (0, 1, 0), fov, s, s), rr_depth, i))
    return views


def grid_rect(origin, u, v, nu: int, nv: int, normal, hole=None) -> TriangleMesh:
    """Rectangle origin + s u + t v (s, t in [0, 1]) as nu x nv cells of two
    triangles each, wound and shaded with `normal`; cells whose centre lies
    in hole = (s0, s1, t0, t1) are left out."""
    origin, u, v = (np.asarray(a, np.float64) for a in (origin, u, v))
    s = np.linspace(0.0, 1.0, nu + 1)
    t = np.linspace(0.0, 1.0, nv + 1)
    S, T = np.meshgrid(s, t, indexing="ij")
    verts = (origin + S[..., None] * u + T[..., None] * v).reshape(-1, 3)
    idx = []
    n = np.asarray(normal, np.float64)
    flip = np.dot(np.cross(u, v), n) < 0
    for i in range(nu):
        for j in range(nv):
            if hole is not None:
                cs, ct = (i + 0.5) / nu, (j + 0.5) / nv
                if hole[0] <= cs <= hole[1] and hole[2] <= ct <= hole[3]:
                    continue
            a, b, c, d = i * (nv + 1) + j, (i + 1) * (nv + 1) + j, (i + 1) * (nv + 1) + j + 1, i * (nv + 1) + j + 1
            idx += [(a, c, b), (a, d, c)] if flip else [(a, b, c), (a, c, d)]
    nrm = np.tile(n.astype(np.float32), (verts.shape[0], 1))
    return TriangleMesh(verts.astype(np.float32), np.array(idx, np.uint32), normals=nrm)


def merge_meshes(meshes) -> TriangleMesh:
    vs, ids, ns, off = [], [], [], 0
    for m in meshes:
        vs.append(m.vertices)
        ids.append(m.triangles + off)
        ns.append(m.normals)
        off += m.vertices.shape[0]
    return TriangleMesh(np.concatenate(vs), np.concatenate(ids).astype(np.uint32), normals=np.concatenate(ns))


def rotate_hdri_to(img: np.ndarray, target_phi: float) -> np.ndarray:
    """Rotate an HDRI about the vertical axis (a whole-texel column shift)
    so that its brightest texel's azimuth moves to target_phi, where a
    direction is (sin th sin ph, cos th, sin th cos ph)."""
    h, w = img.shape[:2]
    lum = img[..., :3].sum(-1)
    y, x = np.unravel_index(np.argmax(lum), lum.shape)
    d = hdri_directions(w, h).reshape(h, w, 3)[y, x]
    phi = math.atan2(d[0], d[2])
    shift = int(round((target_phi - phi) / (2 * math.pi) * w))
    # the dome's u grows with -phi or +phi depending on the mapping: pick the
    # shift direction that lands the peak on target_phi
    for sgn in (1, -1):
        r = np.roll(img, sgn * shift, axis=1)
        yy, xx = np.unravel_index(np.argmax(r[..., :3].sum(-1)), lum.shape)
        dd = hdri_directions(w, h).reshape(h, w, 3)[yy, xx]
        err = abs((math.atan2(dd[0], dd[2]) - target_phi + math.pi) % (2 * math.pi) - math.pi)
        if err <= 2 * math.pi / w * 1.5:
            return np.ascontiguousarray(r)
    raise RuntimeError("could not rotate the HDRI peak to the requested azimuth")


def make_c3(seed: int = 4, width: int = 1920, height: int = 1080, spp: int = 256, cell: float = 1.0 / 16.0,
            n_props: int = 50, prop_subdiv: int = 3, hdri_w: int = 2048, hdri_h: int = 1024, hdri_seed: int = 2,
            max_depth: int = 10) -> SceneDesc:
    """A closed 10 x 10 x 4 room (walls, floor and ceiling as grids of
    `cell`-sized squares, normals facing in) with a 1 x 1 window in the wall
    z = -5, 50 icosphere props (1/3 diffuse, GGX conductor, GGX dielectric)
    and no light but the 1080p scene's HDRI (log-normal, hot spot at 1e4 x
    the mean), rotated about the vertical so its hot spot shines in through
    the window.  1920x1080, 256 spp, depth 10."""
    rng = np.random.default_rng(seed)
    nx, ny = int(round(10.0 / cell)), int(round(4.0 / cell))
    hole = (0.45, 0.55, 0.5, 0.75)  # x in [-0.5, 0.5], y in [2, 3]
    walls = [
        grid_rect((-5, 0, -5), (10, 0, 0), (0, 0, 10), nx, nx, (0, 1, 0)),      # floor
        grid_rect((-5, 4, -5), (10, 0, 0), (0, 0, 10), nx, nx, (0, -1, 0)),     # ceiling
        grid_rect((-5, 0, -5), (10, 0, 0), (0, 4, 0), nx, ny, (0, 0, 1), hole),  # window wall
        grid_rect((-5, 0, 5), (10, 0, 0), (0, 4, 0), nx, ny, (0, 0, -1)),
        grid_rect((-5, 0, -5), (0, 0, 10), (0, 4, 0), nx, ny, (1, 0, 0)),
        grid_rect((5, 0, -5), (0, 0, 10), (0, 4, 0), nx, ny, (-1, 0, 0)),
    ]
    room = merge_meshes(walls)
    sphere = icosphere(prop_subdiv)
    r = rng.uniform(0.2, 0.6, n_props)
    cx = rng.uniform(-4.3, 4.3, n_props)
    cz = rng.uniform(-4.3, 4.3, n_props)
    cy = r * rng.uniform(1.0, 3.0, n_props)
    materials = [DiffuseMtl((0.7, 0.68, 0.62))]
    for i in range(n_props):
        kind = i % 3
        if kind == 0:
            materials.append(DiffuseMtl(tuple(rng.uniform(0.2, 0.8, 3))))
        elif kind == 1:
            eta, k = CONDUCTOR_TABLE[int(rng.integers(len(CONDUCTOR_TABLE)))]
            materials.append(MetalMtl((1.0, 1.0, 1.0), float(np.sqrt(rng.uniform(0.05, 0.5))), eta, k))
        else:
            materials.append(GlassMtl(float(np.sqrt(rng.uniform(0.01, 0.3))), 1.5))
    xf = np.stack([IDENTITY] + [translate_scale((x, y, z), rr) for x, y, z, rr in zip(cx, cy, cz, r)])
    hrng = np.random.default_rng(hdri_seed)
    hdri = rotate_hdri_to(lognormal_hdri(hdri_w, hdri_h, 1.5, hrng), math.pi)  # peak towards -z
    dome = HDRIDome(ImageTexture2D(hdri))
    cam = look_at((0.0, 1.7, 4.6), (0.0, 1.0, -5.0), (0, 1, 0), 65.0, width, height)
    ntri = room.triangles.shape[0] + n_props * sphere.triangles.shape[0]
    return SceneDesc("C3", [room, sphere], [0] + [1] * n_props, list(range(n_props + 1)), xf.astype(np.float32),
                     materials, [dome], cam, width, height, spp, max_depth, 0,
                     info={"seed": seed, "triangles": int(ntri)})
