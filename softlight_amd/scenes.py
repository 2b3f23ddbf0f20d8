"""Seeded synthetic workloads for the draw path.

All randomness comes from a SplitMix64 stream turned into floats through the
top 24 bits, so every backend receives bit-identical buffers.  The builders
only use the SL_Context-shaped API (`SLContext`), so the same call sequence
runs on any library that implements include/sl_gpu.h.

Workloads (sizes are the defaults):
  * `build_random_triangles` -- 4096 independent non-indexed triangles,
    clip-space xyz ~ U[-1,1]^3, w = 1, RGBA8 vertex colours ~ U{0..255},
    pass-through vertex program + Gouraud fragment program, 640x480 RGBA8 +
    R32F targets cleared to 0 / 1.0, depth LESS, no blending, culling off.
  * `build_sphere_scene` -- a 256x256-segment UV sphere (130,560 indexed u32
    triangles, 66,049 vertices of {pos, normal, uv} f32) drawn as 64
    instances at random positions in a 40-unit cube, 60 deg perspective
    (near 0.1, far 100) from the origin looking down -z, a 1024x1024 RGBA8
    value-noise texture sampled bilinearly with REPEAT, Lambert lighting from
    one directional light, 1920x1080 RGBA8 + R32F, depth LESS, back-face
    culling.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from . import api as A

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(seed: int, n: int, stream: int = 0) -> np.ndarray:
    """n outputs of SplitMix64 started from (seed, stream)."""
    with np.errstate(over="ignore"):
        base = np.uint64((seed * 0x9E3779B97F4A7C15 + stream * 0xD1B54A32D192ED03) & 0xFFFFFFFFFFFFFFFF)
        z = base + (np.arange(1, n + 1, dtype=np.uint64) * np.uint64(0x9E3779B97F4A7C15))
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform01(seed: int, n: int, stream: int = 0) -> np.ndarray:
    """float32 in [0, 1) with 24 random mantissa bits (exactly representable)."""
    r = splitmix64(seed, n, stream) >> np.uint64(40)
    return (r.astype(np.float64) * (1.0 / 16777216.0)).astype(np.float32)


def uniform(seed: int, n: int, lo: float, hi: float, stream: int = 0) -> np.ndarray:
    u = uniform01(seed, n, stream)
    return (np.float32(lo) + u * np.float32(hi - lo)).astype(np.float32)


def uniform_u8(seed: int, n: int, stream: int = 0) -> np.ndarray:
    return (splitmix64(seed, n, stream) >> np.uint64(56)).astype(np.uint8)


# --------------------------------------------------------------------------
# matrices (column-major, like ls::math::mat4)
# --------------------------------------------------------------------------
def perspective(fovy_deg: float, aspect: float, near: float, far: float) -> np.ndarray:
    f = 1.0 / math.tan(math.radians(fovy_deg) * 0.5)
    m = np.zeros((4, 4), dtype=np.float64)        # m[col][row]
    m[0][0] = f / aspect
    m[1][1] = f
    m[2][2] = (far + near) / (near - far)
    m[2][3] = -1.0
    m[3][2] = 2.0 * far * near / (near - far)
    return m.astype(np.float32)


def translate_scale(tx, ty, tz, s) -> np.ndarray:
    m = np.zeros((4, 4), dtype=np.float32)
    m[0][0] = m[1][1] = m[2][2] = s
    m[3][0], m[3][1], m[3][2], m[3][3] = tx, ty, tz, 1.0
    return m


def identity() -> np.ndarray:
    return np.eye(4, dtype=np.float32)


# --------------------------------------------------------------------------
# geometry
# --------------------------------------------------------------------------
def uv_sphere(stacks: int, sectors: int, radius: float = 1.0):
    """UV sphere with counter-clockwise outward faces (as seen from outside).

    Returns (vertices float32[(stacks+1)*(sectors+1), 8] = pos, normal, uv;
    indices uint32[ntri*3]).  Pole bands contribute one triangle per sector,
    the other bands two: ntri = 2*stacks*sectors - 2*sectors.
    """
    i = np.arange(stacks + 1, dtype=np.float64)[:, None]
    j = np.arange(sectors + 1, dtype=np.float64)[None, :]
    phi = math.pi * 0.5 - math.pi * i / stacks          # +pi/2 (north) .. -pi/2
    theta = 2.0 * math.pi * j / sectors
    x = np.cos(phi) * np.cos(theta)
    y = np.sin(phi) * np.ones_like(theta)
    z = -np.cos(phi) * np.sin(theta)
    nrm = np.stack([x, y, z], axis=-1).reshape(-1, 3)
    pos = nrm * radius
    uv = np.stack([np.broadcast_to(j / sectors, x.shape), np.broadcast_to(i / stacks, x.shape)], axis=-1).reshape(-1, 2)
    verts = np.concatenate([pos, nrm, uv], axis=1).astype(np.float32)

    k1 = (np.arange(stacks)[:, None] * (sectors + 1) + np.arange(sectors)[None, :])
    k2 = k1 + (sectors + 1)
    tris = []
    top = np.stack([k1, k2, k1 + 1], axis=-1)        # skipped on the north pole band
    bot = np.stack([k1 + 1, k2, k2 + 1], axis=-1)    # skipped on the south pole band
    for s in range(stacks):
        if s != 0:
            tris.append(top[s])
        if s != stacks - 1:
            tris.append(bot[s])
    idx = np.concatenate(tris, axis=0).astype(np.uint32).reshape(-1)
    return verts, idx


def value_noise_rgba8(size: int, seed: int, cells: int = 16) -> np.ndarray:
    """Smooth value noise, one lattice of cells x cells per channel, RGBA8."""
    out = np.empty((size, size, 4), dtype=np.uint8)
    t = (np.arange(size, dtype=np.float64) + 0.5) * cells / size
    i0 = np.floor(t).astype(np.int64) % cells
    i1 = (i0 + 1) % cells
    f = t - np.floor(t)
    f = f * f * (3.0 - 2.0 * f)
    for ch in range(4):
        lat = uniform01(seed, cells * cells, stream=100 + ch).astype(np.float64).reshape(cells, cells)
        if ch == 3:
            lat = 0.5 + 0.5 * lat
        a = lat[i0[:, None], i0[None, :]]
        b = lat[i0[:, None], i1[None, :]]
        c = lat[i1[:, None], i0[None, :]]
        d = lat[i1[:, None], i1[None, :]]
        fx = f[None, :]
        fy = f[:, None]
        v = (a * (1 - fx) + b * fx) * (1 - fy) + (c * (1 - fx) + d * fx) * fy
        out[..., ch] = np.clip(v * 255.0, 0, 255).astype(np.uint8)
    return out


# --------------------------------------------------------------------------
# scenes
# --------------------------------------------------------------------------
@dataclass
class Scene:
    ctx: A.SLContext
    W: int
    H: int
    fbo: int
    color: int
    depth: int
    shader: int
    draws: list = field(default_factory=list)      # (kind, mesh, n) tuples
    clear_rgba: tuple = (0.0, 0.0, 0.0, 0.0)
    clear_depth: float = 1.0
    n_triangles: int = 0
    debug: tuple | None = None
    info: dict = field(default_factory=dict)

    def render(self, clear: bool = True):
        if clear:
            self.ctx.clear_framebuffer(self.fbo, 0, self.clear_rgba, self.clear_depth)
        for kind, mesh, n in self.draws:
            if kind == "draw_shader":
                self.ctx.draw(mesh, n, self.fbo)
            elif kind == "draw":
                self.ctx.draw(mesh, self.shader, self.fbo)
            elif kind == "instanced":
                self.ctx.draw_instanced(mesh, n, self.shader, self.fbo)
            elif kind == "multiple":
                self.ctx.draw_multiple(mesh, self.shader, self.fbo)

    def read(self):
        self.ctx.finish()
        col = self.ctx.texture_read(self.color)
        dep = self.ctx.texture_read(self.depth)
        out = {"color": col, "depth": dep}
        if self.debug is not None:
            out["last"] = self.ctx.texture_read(self.debug[0])
            out["count"] = self.ctx.texture_read(self.debug[1])
        return out


def make_target(ctx: A.SLContext, W: int, H: int, debug: bool, color_type=A.SL_COLOR_RGBA_8U,
                depth_type=A.SL_COLOR_R_FLOAT):
    color = ctx.create_texture(color_type, W, H)
    depth = ctx.create_texture(depth_type, W, H)
    fbo = ctx.create_framebuffer([color], depth)
    dbg = None
    if debug:
        last = ctx.create_texture(A.SL_COLOR_R_32U, W, H)
        cnt = ctx.create_texture(A.SL_COLOR_R_32U, W, H)
        ctx.attach_debug_buffers(fbo, last, cnt)
        dbg = (last, cnt)
    return color, depth, fbo, dbg


def gouraud_shader(ctx, cull=A.SL_CULL_OFF, depth_test=A.SL_DEPTH_TEST_LESS_THAN,
                   depth_mask=A.SL_DEPTH_MASK_ON, blend=A.SL_BLEND_OFF):
    vs = A.sl_vertex_shader(1, cull, A.SL_PROGRAM_PASSTHROUGH_GOURAUD)
    fs = A.sl_fragment_shader(1, 1, blend, depth_test, depth_mask, A.SL_PROGRAM_PASSTHROUGH_GOURAUD)
    return ctx.create_shader(vs, fs)


def random_triangle_vertices(seed: int, ntri: int) -> np.ndarray:
    """Vertex records {f32 x, y, z; u8 r, g, b, a} for ntri independent triangles."""
    nv = ntri * 3
    xyz = uniform(seed, nv * 3, -1.0, 1.0, stream=1).reshape(nv, 3)
    rgba = uniform_u8(seed, nv * 4, stream=2).reshape(nv, 4)
    rec = np.zeros((nv, 16), dtype=np.uint8)
    rec[:, 0:12] = xyz.view(np.uint8).reshape(nv, 12)
    rec[:, 12:16] = rgba
    return rec


def build_from_vertex_records(ctx, rec: np.ndarray, W: int, H: int, debug=False, shader_kw=None,
                              color_type=A.SL_COLOR_RGBA_8U, mode=A.RENDER_MODE_TRIANGLES,
                              depth_type=A.SL_COLOR_R_FLOAT) -> Scene:
    nv = rec.shape[0]
    vbo = ctx.create_vbo(rec)
    vao = ctx.create_vao(vbo, None, [(0, 16, A.VERTEX_DIMENSION_3, A.VERTEX_DATA_FLOAT),
                                     (12, 16, A.VERTEX_DIMENSION_4, A.VERTEX_DATA_BYTE)])
    color, depth, fbo, dbg = make_target(ctx, W, H, debug, color_type, depth_type)
    shader = gouraud_shader(ctx, **(shader_kw or {}))
    mesh = A.sl_mesh(vao, 0, nv, mode, 0)
    return Scene(ctx, W, H, fbo, color, depth, shader, [("draw", mesh, 1)], n_triangles=nv // 3, debug=dbg)


def build_random_triangles(ctx, seed=0, ntri=4096, W=640, H=480, debug=False, wire=False,
                           color_type=A.SL_COLOR_RGBA_8U, depth_type=A.SL_COLOR_R_FLOAT, blend=A.SL_BLEND_OFF) -> Scene:
    mode = A.RENDER_MODE_TRI_WIRE if wire else A.RENDER_MODE_TRIANGLES
    s = build_from_vertex_records(ctx, random_triangle_vertices(seed, ntri), W, H, debug, mode=mode,
                                  color_type=color_type, depth_type=depth_type, shader_kw=dict(blend=blend))
    s.info.update(workload="random_triangles", seed=seed, triangles=ntri, W=W, H=H, mode=mode)
    return s


def build_random_points(ctx, seed=0, npoints=20000, W=640, H=480, debug=False) -> Scene:
    """npoints Gouraud points uniform in NDC [-1.1, 1.1]^2 x [-1, 1] (some
    outside the clip volume), depth-tested."""
    xyz = uniform(seed, npoints * 3, -1.1, 1.1, stream=11).reshape(npoints, 3)
    xyz[:, 2] = uniform(seed, npoints, -1.0, 1.0, stream=12)
    rec = np.zeros((npoints, 16), dtype=np.uint8)
    rec[:, 0:12] = xyz.astype(np.float32).view(np.uint8).reshape(npoints, 12)
    rec[:, 12:16] = uniform_u8(seed, npoints * 4, stream=13).reshape(npoints, 4)
    s = build_from_vertex_records(ctx, rec, W, H, debug, mode=A.RENDER_MODE_POINTS)
    s.info.update(workload="random_points", seed=seed, points=npoints, W=W, H=H)
    return s


def build_random_lines(ctx, seed=0, nlines=3000, W=640, H=480, debug=False) -> Scene:
    """nlines Gouraud segments with ends uniform in NDC [-1.5, 1.5]^2 x
    [-1, 1] (many cross the viewport's edges and are clipped), depth-tested."""
    nv = 2 * nlines
    xyz = uniform(seed, nv * 3, -1.5, 1.5, stream=14).reshape(nv, 3)
    xyz[:, 2] = uniform(seed, nv, -1.0, 1.0, stream=15)
    rec = np.zeros((nv, 16), dtype=np.uint8)
    rec[:, 0:12] = xyz.astype(np.float32).view(np.uint8).reshape(nv, 12)
    rec[:, 12:16] = uniform_u8(seed, nv * 4, stream=16).reshape(nv, 4)
    s = build_from_vertex_records(ctx, rec, W, H, debug, mode=A.RENDER_MODE_LINES)
    s.info.update(workload="random_lines", seed=seed, lines=nlines, W=W, H=H)
    return s


def sphere_instances(seed: int, n: int, extent: float = 20.0, radius: float = 2.0) -> np.ndarray:
    pos = uniform(seed, n * 3, -extent, extent, stream=7).reshape(n, 3)
    mats = np.stack([translate_scale(p[0], p[1], p[2], radius) for p in pos]).astype(np.float32)
    return mats


def build_sphere_scene(ctx, seed=1, n_instances=64, stacks=256, sectors=256, W=1920, H=1080,
                       tex_size=1024, radius=2.0, extent=20.0, debug=False, cull=A.SL_CULL_BACK_FACE,
                       blend=A.SL_BLEND_OFF, alpha=1.0, wire=False, prim="triangles",
                       color_type=A.SL_COLOR_RGBA_8U, depth_type=A.SL_COLOR_R_FLOAT) -> Scene:
    verts, idx = uv_sphere(stacks, sectors)
    vbo = ctx.create_vbo(verts)
    ibo = ctx.create_ibo(idx)
    vao = ctx.create_vao(vbo, ibo, [(0, 32, A.VERTEX_DIMENSION_3, A.VERTEX_DATA_FLOAT),
                                    (12, 32, A.VERTEX_DIMENSION_3, A.VERTEX_DATA_FLOAT),
                                    (24, 32, A.VERTEX_DIMENSION_2, A.VERTEX_DATA_FLOAT)])
    tex = ctx.create_texture(A.SL_COLOR_RGBA_8U, tex_size, tex_size, 1, value_noise_rgba8(tex_size, seed))
    inst = sphere_instances(seed, n_instances, extent, radius)
    inst_vbo = ctx.create_vbo(inst)
    u = A.sl_std_uniforms()
    P = perspective(60.0, W / H, 0.1, 100.0)
    u.mvp[:] = P.reshape(-1).tolist()
    u.model[:] = identity().reshape(-1).tolist()
    L = np.array([-1.0, -1.0, -1.0])
    L = (L / np.linalg.norm(L)).astype(np.float32)
    u.light_dir[:] = [float(L[0]), float(L[1]), float(L[2]), 0.0]
    u.ambient[:] = [0.2, 0.2, 0.2, 0.0]
    u.diffuse[:] = [0.8, 0.8, 0.8, 0.0]
    u.tint[:] = [1.0, 1.0, 1.0, alpha]
    u.texture = tex
    u.instance_vbo = inst_vbo
    u.flags = 0
    ubo = ctx.create_ubo(u)
    vs = A.sl_vertex_shader(2, cull, A.SL_PROGRAM_LAMBERT_TEXTURED)
    fs = A.sl_fragment_shader(2, 1, blend, A.SL_DEPTH_TEST_LESS_THAN, A.SL_DEPTH_MASK_ON, A.SL_PROGRAM_LAMBERT_TEXTURED)
    shader = ctx.create_shader(vs, fs, ubo)
    color, depth, fbo, dbg = make_target(ctx, W, H, debug, color_type, depth_type)
    # the same index buffer read as triangles, wireframe triangles, lines or points
    mode = {"triangles": A.RENDER_MODE_INDEXED_TRI_WIRE if wire else A.RENDER_MODE_INDEXED_TRIANGLES,
            "lines": A.RENDER_MODE_INDEXED_LINES, "points": A.RENDER_MODE_INDEXED_POINTS}[prim]
    mesh = A.sl_mesh(vao, 0, idx.size, mode, 0)
    s = Scene(ctx, W, H, fbo, color, depth, shader, [("instanced", mesh, n_instances)],
              n_triangles=(idx.size // 3) * n_instances, debug=dbg)
    s.info.update(workload="textured_spheres", seed=seed, instances=n_instances, stacks=stacks, sectors=sectors,
                  W=W, H=H, tex_size=tex_size, radius=radius, extent=extent, mode=mode,
                  triangles_per_instance=idx.size // 3, vertices_per_instance=verts.shape[0])
    s.info["_keep"] = (verts, idx, inst)
    s.info["vbo"] = vbo
    return s


# --------------------------------------------------------------------------
# batched multi-material frame
# --------------------------------------------------------------------------
_ICO_FACES = np.array([(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4),
                       (11, 10, 2), (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8),
                       (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)], dtype=np.int64)
_ico_cache = {}


def icosphere(level: int):
    """Unit icosphere subdivided `level` times: (positions float64[V,3],
    faces int64[20*4^level, 3]) with counter-clockwise outward faces."""
    if level in _ico_cache:
        return _ico_cache[level]
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = np.array([(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
                  (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)], dtype=np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = _ICO_FACES.copy()
    for _ in range(level):
        e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]], axis=0)
        key = np.minimum(e[:, 0], e[:, 1]) * (1 << 32) + np.maximum(e[:, 0], e[:, 1])
        uniq, inv = np.unique(key, return_inverse=True)
        a, b = uniq >> 32, uniq & 0xFFFFFFFF
        mid = (v[a] + v[b]) * 0.5
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        base = v.shape[0]
        v = np.concatenate([v, mid], axis=0)
        n = f.shape[0]
        ab, bc, ca = base + inv[:n], base + inv[n:2 * n], base + inv[2 * n:]
        f = np.concatenate([np.stack([f[:, 0], ab, ca], 1), np.stack([f[:, 1], bc, ab], 1),
                            np.stack([f[:, 2], ca, bc], 1), np.stack([ab, bc, ca], 1)], axis=0)
    _ico_cache[level] = (v, f)
    return v, f


def rotation(axis, angle):
    axis = np.asarray(axis, dtype=np.float64)
    axis = axis / np.linalg.norm(axis)
    x, y, z = axis
    c, s, C = np.cos(angle), np.sin(angle), 1 - np.cos(angle)
    R = np.array([[c + x * x * C, x * y * C - z * s, x * z * C + y * s],
                  [y * x * C + z * s, c + y * y * C, y * z * C - x * s],
                  [z * x * C - y * s, z * y * C + x * s, c + z * z * C]])
    return R


def build_batched_scene(ctx, seed=3, n_meshes=1000, median_tris=2000, sigma=1.0, min_tris=12, max_tris=200_000,
                        W=1920, H=1080, n_textures=16, tex_size=512, blend_fraction=0.2, debug=False, wire_every=0,
                        color_type=A.SL_COLOR_RGBA_8U, depth_type=A.SL_COLOR_R_FLOAT):
    """1000 displaced icospheres (log-normal triangle counts), random affine
    model matrices, 80% opaque / 20% alpha-blended Lambert materials over
    sixteen RGBA8 noise textures; one draw call per mesh, in index order."""
    rng_u = uniform01(seed, n_meshes * 16, stream=30).astype(np.float64).reshape(n_meshes, 16)
    # log-normal counts via Box-Muller on the seeded stream
    u1, u2 = np.clip(rng_u[:, 0], 1e-7, 1.0), rng_u[:, 1]
    normal = np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)
    counts = np.clip(np.round(median_tris * np.exp(sigma * normal)), min_tris, max_tris).astype(np.int64)
    texs = [ctx.create_texture(A.SL_COLOR_RGBA_8U, tex_size, tex_size, 1, value_noise_rgba8(tex_size, seed * 97 + i, cells=8))
            for i in range(n_textures)]
    P = perspective(60.0, W / H, 0.1, 100.0)
    L = np.array([-0.4, -1.0, -0.6])
    L = (L / np.linalg.norm(L)).astype(np.float32)
    vs_op = A.sl_vertex_shader(2, A.SL_CULL_BACK_FACE, A.SL_PROGRAM_LAMBERT_TEXTURED)
    fs_op = A.sl_fragment_shader(2, 1, A.SL_BLEND_OFF, A.SL_DEPTH_TEST_LESS_THAN, A.SL_DEPTH_MASK_ON, A.SL_PROGRAM_LAMBERT_TEXTURED)
    fs_bl = A.sl_fragment_shader(2, 1, A.SL_BLEND_ALPHA, A.SL_DEPTH_TEST_LESS_THAN, A.SL_DEPTH_MASK_OFF, A.SL_PROGRAM_LAMBERT_TEXTURED)
    color, depth, fbo, dbg = make_target(ctx, W, H, debug, color_type, depth_type)
    draws = []
    shaders = []
    keep = []
    level_ibo = {}
    total = 0
    n_blend = 0
    for m in range(n_meshes):
        n = int(counts[m])
        level = 0
        while 20 * 4 ** level < n:
            level += 1
        base_v, faces = icosphere(level)
        if level not in level_ibo:
            level_ibo[level] = faces.astype(np.uint32).reshape(-1)
        disp = 1.0 + 0.15 * (uniform01(seed * 7919 + m, base_v.shape[0], stream=31).astype(np.float64) - 0.5)
        pos = base_v * disp[:, None]
        uv = np.stack([np.arctan2(base_v[:, 2], base_v[:, 0]) / (2 * np.pi) + 0.5,
                       np.arcsin(np.clip(base_v[:, 1], -1, 1)) / np.pi + 0.5], axis=1)
        verts = np.concatenate([pos, base_v, uv], axis=1).astype(np.float32)
        vbo = ctx.create_vbo(verts)
        ibo = ctx.create_ibo(level_ibo[level])
        vao = ctx.create_vao(vbo, ibo, [(0, 32, A.VERTEX_DIMENSION_3, A.VERTEX_DATA_FLOAT),
                                        (12, 32, A.VERTEX_DIMENSION_3, A.VERTEX_DATA_FLOAT),
                                        (24, 32, A.VERTEX_DIMENSION_2, A.VERTEX_DATA_FLOAT)])
        r = rng_u[m]
        R = rotation([r[2] - 0.5, r[3] - 0.5, r[4] - 0.5 + 1e-3], r[5] * 2 * np.pi)
        S = np.diag([0.4 + 1.6 * r[6], 0.4 + 1.6 * r[7], 0.4 + 1.6 * r[8]])
        M = np.eye(4)
        M[:3, :3] = R @ S
        M[:3, 3] = [(r[9] - 0.5) * 36.0, (r[10] - 0.5) * 20.0, -6.0 - 40.0 * r[11]]
        u = A.sl_std_uniforms()
        u.mvp[:] = P.reshape(-1).tolist()
        u.model[:] = M.T.astype(np.float32).reshape(-1).tolist()      # column-major
        u.light_dir[:] = [float(L[0]), float(L[1]), float(L[2]), 0.0]
        u.ambient[:] = [0.25, 0.25, 0.25, 0.0]
        u.diffuse[:] = [0.75, 0.75, 0.75, 0.0]
        blended = r[12] < blend_fraction
        alpha = 0.2 + 0.6 * r[13] if blended else 1.0
        u.tint[:] = [1.0, 1.0, 1.0, float(alpha)]
        u.texture = texs[m % n_textures]
        u.instance_vbo = A.SL_NO_RESOURCE
        ubo = ctx.create_ubo(u)
        sh = ctx.create_shader(vs_op, fs_bl if blended else fs_op, ubo)
        shaders.append(sh)
        n_blend += int(blended)
        wire = wire_every > 0 and m % wire_every == wire_every - 1
        mesh = A.sl_mesh(vao, 0, 3 * n, A.RENDER_MODE_INDEXED_TRI_WIRE if wire else A.RENDER_MODE_INDEXED_TRIANGLES, m)
        draws.append(("draw_shader", mesh, sh))
        total += n
        keep.append(verts)
    s = Scene(ctx, W, H, fbo, color, depth, shaders[0] if shaders else 0, draws, n_triangles=total, debug=dbg)
    s.info.update(workload="batched_meshes", seed=seed, meshes=n_meshes, blended=n_blend, W=W, H=H,
                  triangles=total, median_tris=median_tris, sigma=sigma)
    s.info["_keep"] = keep
    return s


# --------------------------------------------------------------------------
# differentiable views: a colour-and-texture icosphere seen by many cameras
# --------------------------------------------------------------------------
def gaussian(seed: int, n: int, stream: int = 0) -> np.ndarray:
    """Standard normal float64 samples (Box-Muller on two uniform streams)."""
    u1 = uniform01(seed, n, stream=stream).astype(np.float64)
    u2 = uniform01(seed, n, stream=stream + 1).astype(np.float64)
    u1 = np.maximum(u1, 2.0 ** -24)
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * math.pi * u2)


def look_at(eye, target, up) -> np.ndarray:
    """View matrix (row-major math convention), camera looking down -z."""
    eye = np.asarray(eye, dtype=np.float64)
    f = np.asarray(target, dtype=np.float64) - eye
    f /= np.linalg.norm(f)
    s = np.cross(f, np.asarray(up, dtype=np.float64))
    s /= np.linalg.norm(s)
    u = np.cross(s, f)
    V = np.eye(4)
    V[0, :3], V[1, :3], V[2, :3] = s, u, -f
    V[0, 3], V[1, 3], V[2, 3] = -s @ eye, -u @ eye, f @ eye
    return V


class DiffScene:
    """A unit icosphere with per-vertex RGB colours and an RGBA32F texture
    (SL_PROGRAM_COLOR_TEXTURED), rendered from `views` random cameras into
    RGBA32F targets with visibility attachments.  Views are packed
    `views_per_fbo` to a framebuffer as a grid of viewports, so each
    framebuffer is one batch.  The loss is sum((C - T)^2) over every view;
    gradients go to an RGBA32F texture and a float4-per-vertex VBO."""

    def __init__(self, ctx: A.SLContext, seed=4, level=5, views=32, size=512, views_per_fbo=8, tex_size=512,
                 distance=3.0, fov=60.0, edge_views=0, flags=0):
        self.ctx = ctx
        self.seed, self.level, self.views, self.size = seed, level, views, size
        self.tex_size = tex_size
        pos, faces = icosphere(level)
        nv = pos.shape[0]
        self.n_vertices, self.n_triangles = nv, faces.shape[0]
        lon = np.arctan2(pos[:, 2], pos[:, 0])
        lat = np.arccos(np.clip(pos[:, 1], -1.0, 1.0))
        uv = np.stack([lon / (2 * math.pi) + 0.5, lat / math.pi], axis=1)
        self.verts = np.zeros((nv, 8), dtype=np.float32)
        self.verts[:, 0:3] = pos
        self.verts[:, 3:6] = uniform01(seed, nv * 3, stream=1).reshape(nv, 3)
        self.verts[:, 6:8] = uv
        self.texels = uniform01(seed, tex_size * tex_size * 4, stream=2).reshape(tex_size, tex_size, 4).astype(np.float32)
        self.indices = faces.astype(np.uint32).reshape(-1)

        self.vbo = ctx.create_vbo(self.verts)
        self.ibo = ctx.create_ibo(self.indices)
        self.vao = ctx.create_vao(self.vbo, self.ibo, [(0, 32, A.VERTEX_DIMENSION_3, A.VERTEX_DATA_FLOAT),
                                                       (12, 32, A.VERTEX_DIMENSION_3, A.VERTEX_DATA_FLOAT),
                                                       (24, 32, A.VERTEX_DIMENSION_2, A.VERTEX_DATA_FLOAT)])
        self.tex = ctx.create_texture(A.SL_COLOR_RGBA_FLOAT, tex_size, tex_size, 1, self.texels)
        self.grad_vbo = ctx.create_vbo(np.zeros((nv, 4), dtype=np.float32))
        self.grad_tex = ctx.create_texture(A.SL_COLOR_RGBA_FLOAT, tex_size, tex_size, 1)
        ctx.vbo_attach_gradient(self.vbo, self.grad_vbo)
        ctx.texture_attach_gradient(self.tex, self.grad_tex)
        self.mesh = A.sl_mesh(self.vao, 0, self.indices.size, A.RENDER_MODE_INDEXED_TRIANGLES, 0)

        # cameras: random directions at `distance`, looking at the centre; the
        # first edge_views look past the sphere so that the frustum's side
        # planes cut it (clipped triangles)
        P = perspective(fov, 1.0, 0.1, 100.0).T.astype(np.float64)
        d = gaussian(seed, views * 3, stream=3).reshape(views, 3)
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        self.shaders = []
        self.ubos = []
        for v in range(views):
            up = (0.0, 1.0, 0.0) if abs(d[v, 1]) < 0.9 else (1.0, 0.0, 0.0)
            target = np.zeros(3)
            if v < edge_views:
                side = np.cross(d[v], up)
                target = 1.2 * side / np.linalg.norm(side)
            V = look_at(d[v] * distance, target, up)
            mvp = P @ V
            u = A.sl_std_uniforms()
            u.mvp[:] = mvp.T.astype(np.float32).reshape(-1).tolist()
            u.model[:] = identity().reshape(-1).tolist()
            u.tint[:] = [1.0, 1.0, 1.0, 1.0]
            u.texture = self.tex
            u.instance_vbo = A.SL_NO_RESOURCE
            u.flags = flags
            ubo = ctx.create_ubo(u)
            vs = A.sl_vertex_shader(2, A.SL_CULL_BACK_FACE, A.SL_PROGRAM_COLOR_TEXTURED)
            fs = A.sl_fragment_shader(2, 1, A.SL_BLEND_OFF, A.SL_DEPTH_TEST_LESS_THAN, A.SL_DEPTH_MASK_ON,
                                      A.SL_PROGRAM_COLOR_TEXTURED)
            self.ubos.append(ubo)
            self.shaders.append(ctx.create_shader(vs, fs, ubo))

        # framebuffers: a gx x gy grid of size x size viewports each
        self.targets = []
        per = max(1, min(views_per_fbo, views))
        gx = 1
        while gx * gx < per:
            gx *= 2
        gy = (per + gx - 1) // gx
        for f0 in range(0, views, per):
            vs_ = list(range(f0, min(views, f0 + per)))
            W, H = gx * size, gy * size
            color = ctx.create_texture(A.SL_COLOR_RGBA_FLOAT, W, H)
            depth = ctx.create_texture(A.SL_COLOR_R_FLOAT, W, H)
            vis = ctx.create_texture(A.SL_COLOR_RGBA_FLOAT, W, H)
            target = ctx.create_texture(A.SL_COLOR_RGBA_FLOAT, W, H)
            gimg = ctx.create_texture(A.SL_COLOR_RGBA_FLOAT, W, H)
            fbo = ctx.create_framebuffer([color], depth)
            ctx.attach_visibility(fbo, vis)
            cells = [((i % gx) * size, (i // gx) * size) for i in range(len(vs_))]
            self.targets.append(dict(fbo=fbo, color=color, depth=depth, vis=vis, target=target, grad_image=gimg,
                                     W=W, H=H, views=vs_, cells=cells))

    # ---- parameters ----
    def colors(self) -> np.ndarray:
        return self.verts[:, 3:6].copy()

    def set_colors(self, rgb: np.ndarray):
        self.verts[:, 3:6] = rgb
        self.ctx.vbo_assign(self.vbo, self.verts)

    def set_color_value(self, vertex: int, channel: int, value: float):
        self.verts[vertex, 3 + channel] = value
        self.ctx.vbo_assign(self.vbo, self.verts[vertex:vertex + 1, 3 + channel:4 + channel].copy(),
                            offset=vertex * 32 + 12 + 4 * channel)

    def set_texels(self, texels: np.ndarray):
        self.texels = np.ascontiguousarray(texels, dtype=np.float32)
        self.ctx.texture_set_texels(self.tex, self.texels)

    # ---- passes ----
    def forward(self):
        for t in self.targets:
            self.ctx.clear_framebuffer(t["fbo"], 0, (0.0, 0.0, 0.0, 0.0), 1.0)
            for v, (x, y) in zip(t["views"], t["cells"]):
                self.ctx.set_viewport(x, y, self.size, self.size)
                self.ctx.draw(self.mesh, self.shaders[v], t["fbo"])
        self.ctx.set_viewport(0, 0, 65535, 65535)

    def images(self):
        self.ctx.finish()
        return [self.ctx.texture_read(t["color"]) for t in self.targets]

    def visibility(self):
        self.ctx.finish()
        return [self.ctx.texture_read(t["vis"]) for t in self.targets]

    def set_targets(self, images):
        for t, img in zip(self.targets, images):
            self.ctx.texture_set_texels(t["target"], np.ascontiguousarray(img, dtype=np.float32))

    def make_targets(self, std=0.1):
        """Targets rendered from the colours and texels perturbed by N(0, std)."""
        c0, t0 = self.colors(), self.texels.copy()
        nc, nt = c0.size, t0.size
        self.set_colors((c0 + std * gaussian(self.seed + 1, nc, stream=5).reshape(c0.shape)).astype(np.float32))
        self.set_texels((t0 + std * gaussian(self.seed + 1, nt, stream=7).reshape(t0.shape)).astype(np.float32))
        self.forward()
        self.set_targets(self.images())
        self.set_colors(c0)
        self.set_texels(t0)

    def loss_grad_images(self, scale=1.0):
        for t in self.targets:
            self.ctx.loss_l2_grad(t["color"], t["target"], t["grad_image"], scale)

    def backward(self):
        self.ctx.gradients_zero()
        for t in self.targets:
            self.ctx.backward(t["fbo"], t["grad_image"])

    def gradients(self):
        self.ctx.finish()
        gc = self.ctx.vbo_read(self.grad_vbo, self.n_vertices * 16, dtype=np.float32).reshape(-1, 4)
        gt = self.ctx.texture_read(self.grad_tex).reshape(self.tex_size, self.tex_size, 4)
        return gc, gt

    def loss64(self, images=None):
        """sum((C - T)^2) in float64 over every view, from the read-back images."""
        if images is None:
            images = self.images()
        tot = 0.0
        for t, img in zip(self.targets, images):
            tgt = self.ctx.texture_read(t["target"])
            d = img.astype(np.float64) - tgt.astype(np.float64)
            tot += float(np.sum(d * d))
        return tot


class SmoothScene:
    """Multi-mesh views for the smoothed-coverage mode: `meshes` UV spheres with
    log-uniform triangle counts in [min_tris, max_tris], per-vertex RGB and one
    RGBA32F texture (SL_PROGRAM_COLOR_TEXTURED), placed around the origin and
    seen by `views` random cameras.  Views are packed `views_per_fbo` to a
    framebuffer as a grid of viewports (one batch each); every view draws
    every mesh.  The loss is sum((S - T)^2) over the smoothed images S;
    gradients go to the vertex colours, the texels and the vertex screen
    positions (float2 per saved draw and vertex).  With shared_vis every
    framebuffer uses one visibility attachment: the backward passes then draw
    each framebuffer's views again to recompute its visibility."""

    def __init__(self, ctx: A.SLContext, seed=6, meshes=8, min_tris=5000, max_tris=60000, views=64, size=512,
                 views_per_fbo=8, tex_size=512, distance=4.0, fov=60.0, shared_vis=False):
        self.ctx = ctx
        self.seed, self.views, self.size = seed, views, size
        lo, hi = math.log(min_tris), math.log(max_tris)
        u = uniform01(seed, meshes, stream=21).astype(np.float64)
        self.mesh_tris, self.n_vertices_max, self.n_triangles = [], 0, 0
        self.meshes, self.vbos, self.pos = [], [], []
        for m in range(meshes):
            target = math.exp(lo + (hi - lo) * u[m])
            st = max(3, int(round(math.sqrt(target / 4.0))))
            verts, idx = uv_sphere(st, 2 * st)
            nv = verts.shape[0]
            v = np.zeros((nv, 8), dtype=np.float32)
            v[:, 0:3] = verts[:, 0:3]
            v[:, 3:6] = uniform01(seed + m, nv * 3, stream=22).reshape(nv, 3)
            v[:, 6:8] = verts[:, 6:8]
            vbo = ctx.create_vbo(v)
            ibo = ctx.create_ibo(idx)
            vao = ctx.create_vao(vbo, ibo, [(0, 32, A.VERTEX_DIMENSION_3, A.VERTEX_DATA_FLOAT),
                                           (12, 32, A.VERTEX_DIMENSION_3, A.VERTEX_DATA_FLOAT),
                                           (24, 32, A.VERTEX_DIMENSION_2, A.VERTEX_DATA_FLOAT)])
            self.meshes.append(A.sl_mesh(vao, 0, idx.size, A.RENDER_MODE_INDEXED_TRIANGLES, 0))
            self.vbos.append(vbo)
            self.pos.append(v)
            self.mesh_tris.append(idx.size // 3)
            self.n_triangles += idx.size // 3
            self.n_vertices_max = max(self.n_vertices_max, nv)
        self.tex = ctx.create_texture(A.SL_COLOR_RGBA_FLOAT, tex_size, tex_size, 1,
                                      uniform01(seed, tex_size * tex_size * 4, stream=23).reshape(tex_size, tex_size, 4))
        self.grad_tex = ctx.create_texture(A.SL_COLOR_RGBA_FLOAT, tex_size, tex_size)
        ctx.texture_attach_gradient(self.tex, self.grad_tex)
        self.grad_vbos = []
        for m in range(meshes):
            g = ctx.create_vbo(np.zeros((self.pos[m].shape[0], 4), dtype=np.float32))
            ctx.vbo_attach_gradient(self.vbos[m], g)
            self.grad_vbos.append(g)

        # mesh placements (model matrices, column-major) and cameras
        place = uniform(seed, meshes * 3, -1.2, 1.2, stream=24).reshape(meshes, 3)
        models = [translate_scale(p[0], p[1], p[2], 0.55) for p in place]
        P = perspective(fov, 1.0, 0.1, 100.0).T.astype(np.float64)
        d = gaussian(seed, views * 3, stream=25).reshape(views, 3)
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        self.shaders = []
        for v in range(views):
            up = (0.0, 1.0, 0.0) if abs(d[v, 1]) < 0.9 else (1.0, 0.0, 0.0)
            mvp = P @ look_at(d[v] * distance, np.zeros(3), up)
            row = []
            for m in range(meshes):
                un = A.sl_std_uniforms()
                un.mvp[:] = mvp.T.astype(np.float32).reshape(-1).tolist()
                un.model[:] = models[m].reshape(-1).tolist()
                un.tint[:] = [1.0, 1.0, 1.0, 1.0]
                un.texture = self.tex
                un.instance_vbo = A.SL_NO_RESOURCE
                ubo = ctx.create_ubo(un)
                vs = A.sl_vertex_shader(2, A.SL_CULL_BACK_FACE, A.SL_PROGRAM_COLOR_TEXTURED)
                fs = A.sl_fragment_shader(2, 1, A.SL_BLEND_OFF, A.SL_DEPTH_TEST_LESS_THAN, A.SL_DEPTH_MASK_ON,
                                          A.SL_PROGRAM_COLOR_TEXTURED)
                row.append(ctx.create_shader(vs, fs, ubo))
            self.shaders.append(row)

        per = max(1, min(views_per_fbo, views))
        gx = 1
        while gx * gx < per:
            gx *= 2
        gy = (per + gx - 1) // gx
        self.targets = []
        shared = ctx.create_texture(A.SL_COLOR_RGBA_FLOAT, gx * size, gy * size) if shared_vis else None
        for f0 in range(0, views, per):
            vs_ = list(range(f0, min(views, f0 + per)))
            W, H = gx * size, gy * size
            mk = lambda: ctx.create_texture(A.SL_COLOR_RGBA_FLOAT, W, H)
            color, smooth, target, gsm, gimg = mk(), mk(), mk(), mk(), mk()
            vis = shared if shared is not None else mk()
            depth = ctx.create_texture(A.SL_COLOR_R_FLOAT, W, H)
            fbo = ctx.create_framebuffer([color], depth)
            ctx.attach_visibility(fbo, vis)
            ndraw = len(vs_) * meshes
            pg = ctx.create_vbo(np.zeros((ndraw, self.n_vertices_max, 2), dtype=np.float32))
            ctx.attach_positions(fbo, pg, A.SL_INVALID_ID, self.n_vertices_max)
            cells = [((i % gx) * size, (i // gx) * size) for i in range(len(vs_))]
            self.targets.append(dict(fbo=fbo, color=color, vis=vis, smooth=smooth, target=target, grad_smooth=gsm,
                                     grad_image=gimg, pos_grad=pg, W=W, H=H, views=vs_, cells=cells))

    def forward(self, sigma):
        for t in self.targets:
            self.ctx.clear_framebuffer(t["fbo"], 0, (0.0, 0.0, 0.0, 0.0), 1.0)
            for v, (x, y) in zip(t["views"], t["cells"]):
                self.ctx.set_viewport(x, y, self.size, self.size)
                for m, mesh in enumerate(self.meshes):
                    self.ctx.draw(mesh, self.shaders[v][m], t["fbo"])
            self.ctx.smooth_coverage(t["fbo"], t["smooth"], sigma)
        self.ctx.set_viewport(0, 0, 65535, 65535)

    def make_targets(self, sigma, std=0.02):
        """Targets: the smoothed images of the meshes with vertices displaced by N(0, std)."""
        for m, v in enumerate(self.pos):
            d = v.copy()
            d[:, 0:3] += (std * gaussian(self.seed + 7 + m, d.shape[0] * 3, stream=26).reshape(-1, 3)).astype(np.float32)
            self.ctx.vbo_assign(self.vbos[m], d)
        self.forward(sigma)
        self.ctx.finish()
        for t in self.targets:
            self.ctx.texture_set_texels(t["target"], self.ctx.texture_read(t["smooth"]))
        for m, v in enumerate(self.pos):
            self.ctx.vbo_assign(self.vbos[m], v)

    def backward(self, sigma):
        self.ctx.gradients_zero()
        for t in self.targets:
            self.ctx.loss_l2_grad(t["smooth"], t["target"], t["grad_smooth"], 1.0)
            self.ctx.smooth_coverage_backward(t["fbo"], t["grad_smooth"], sigma, t["grad_image"])
            self.ctx.backward(t["fbo"], t["grad_image"])


# --------------------------------------------------------------------------
# scene files: Wavefront OBJ + MTL + PPM written from the seeded streams
# --------------------------------------------------------------------------
def _fmt(x) -> str:
    """Shortest text that reads back as the same float32."""
    return np.format_float_positional(np.float32(x), unique=True, trim="-")


def write_ppm(path, rgba: np.ndarray):
    """Binary PPM (P6) of an RGBA8 image's RGB channels."""
    h, w = rgba.shape[:2]
    with open(path, "wb") as f:
        f.write(b"P6\n%d %d\n255\n" % (w, h))
        f.write(np.ascontiguousarray(rgba[:, :, :3]).tobytes())


def write_icosphere_scene(directory, seed=8, n_objects=6, level=2, n_materials=3, tex_size=64, extent=6.0,
                          name="spheres.obj"):
    """`n_objects` displaced icospheres as OBJ objects ('o sphere<i>'), each
    at its own position in the file's coordinates (a cube of side 2*extent
    around (0, 0, -2*extent)), with v/vt/vn per corner; materials cycle
    over `n_materials` entries of which the even ones carry a value-noise
    diffuse map.  Every other object's second half of faces switches to the
    next material, so objects hold one or two meshes.  Returns the OBJ path
    and per-object triangle counts."""
    import os
    base_v, faces = icosphere(level)
    r = uniform01(seed, n_objects * 8, stream=40).astype(np.float64).reshape(n_objects, 8)
    mtl = []
    for m in range(n_materials):
        mtl.append(f"newmtl mat{m}")
        mtl.append(f"Ka {_fmt(0.1 + 0.05 * m)} {_fmt(0.1)} {_fmt(0.15)}")
        mtl.append(f"Kd {_fmt(0.9 - 0.1 * m)} {_fmt(0.8)} {_fmt(0.7 + 0.1 * m)}")
        if m % 2 == 0:
            img = f"tex{m}.ppm"
            write_ppm(os.path.join(directory, img), value_noise_rgba8(tex_size, seed * 31 + m, cells=8))
            mtl.append(f"map_Kd {img}")
    with open(os.path.join(directory, "spheres.mtl"), "w") as f:
        f.write("\n".join(mtl) + "\n")
    lines = ["mtllib spheres.mtl"]
    counts = []
    nv = 0
    for o in range(n_objects):
        disp = 1.0 + 0.2 * (uniform01(seed * 7919 + o, base_v.shape[0], stream=41).astype(np.float64) - 0.5)
        rad = 0.6 + 0.8 * r[o, 3]
        c = np.array([(r[o, 0] - 0.5) * 2 * extent, (r[o, 1] - 0.5) * 2 * extent, -2 * extent + (r[o, 2] - 0.5) * extent])
        pos = (base_v * (disp[:, None] * rad) + c).astype(np.float32)
        uv = np.stack([np.arctan2(base_v[:, 2], base_v[:, 0]) / (2 * np.pi) + 0.5,
                       np.arcsin(np.clip(base_v[:, 1], -1, 1)) / np.pi + 0.5], axis=1).astype(np.float32)
        nrm = base_v.astype(np.float32)
        lines.append(f"o sphere{o}")
        lines += [f"v {_fmt(p[0])} {_fmt(p[1])} {_fmt(p[2])}" for p in pos]
        lines += [f"vt {_fmt(t[0])} {_fmt(t[1])}" for t in uv]
        lines += [f"vn {_fmt(n[0])} {_fmt(n[1])} {_fmt(n[2])}" for n in nrm]
        m0 = o % n_materials
        lines.append(f"usemtl mat{m0}")
        half = faces.shape[0] // 2
        for k, tri in enumerate(faces):
            if o % 2 == 1 and k == half:
                lines.append(f"usemtl mat{(m0 + 1) % n_materials}")
            a, b, cc = (int(i) + 1 + nv for i in tri)
            lines.append(f"f {a}/{a}/{a} {b}/{b}/{b} {cc}/{cc}/{cc}")
        nv += base_v.shape[0]
        counts.append(int(faces.shape[0]))
    path = os.path.join(directory, name)
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
    return path, counts
