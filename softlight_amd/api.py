"""ctypes binding of the C ABI declared in include/sl_gpu.h.

`SLContext` mirrors the reference's `SL_Context` (SL_Context.hpp:59-306): the
same resource-creation calls returning dense `size_t` handles, the same draw
and clear calls.  It binds any library that exports the ABI under a given
symbol prefix; the product library is `softlight_amd/_build/libsoftlight_amd.so`
with prefix ``sl_`` (see softlight_amd.load_library).
"""
from __future__ import annotations

import ctypes as C
import numpy as np

SL_INVALID_ID = (1 << 64) - 1
SL_NO_RESOURCE = 0xFFFFFFFF

# error codes
SL_OK = 0
SL_ERR_INVALID_HANDLE = -100
SL_ERR_INVALID_ARG = -101
SL_ERR_OUT_OF_MEMORY = -102
SL_ERR_DEVICE = -103
SL_ERR_UNSUPPORTED = -104
SL_ERR_CAPACITY = -105
SL_ERR_STATE = -106

# SL_PipelineState.hpp enums
SL_CULL_BACK_FACE, SL_CULL_FRONT_FACE, SL_CULL_OFF = 0, 1, 2
(SL_DEPTH_TEST_OFF, SL_DEPTH_TEST_LESS_THAN, SL_DEPTH_TEST_LESS_EQUAL, SL_DEPTH_TEST_GREATER_THAN,
 SL_DEPTH_TEST_GREATER_EQUAL, SL_DEPTH_TEST_EQUAL, SL_DEPTH_TEST_NOT_EQUAL) = range(7)
SL_DEPTH_MASK_OFF, SL_DEPTH_MASK_ON = 0, 1
SL_BLEND_OFF, SL_BLEND_ALPHA, SL_BLEND_PREMULTIPLED_ALPHA, SL_BLEND_ADDITIVE, SL_BLEND_SCREEN = range(5)
RENDER_MODE_POINTS = 0x01
RENDER_MODE_INDEXED_POINTS = 0x03
RENDER_MODE_LINES = 0x04
RENDER_MODE_INDEXED_LINES = 0x0C
RENDER_MODE_TRIANGLES = 0x10
RENDER_MODE_INDEXED_TRIANGLES = 0x30
RENDER_MODE_TRI_WIRE = 0x50
RENDER_MODE_INDEXED_TRI_WIRE = 0x70
VERTEX_DIMENSION_1, VERTEX_DIMENSION_2, VERTEX_DIMENSION_3, VERTEX_DIMENSION_4 = range(4)
VERTEX_DATA_BYTE, VERTEX_DATA_SHORT, VERTEX_DATA_INT, VERTEX_DATA_LONG, VERTEX_DATA_FLOAT = range(5)
SL_COLOR_R_8U, SL_COLOR_R_32U, SL_COLOR_R_HALF, SL_COLOR_R_FLOAT = 0, 2, 4, 5
SL_COLOR_RG_8U, SL_COLOR_RG_FLOAT = 7, 12
SL_COLOR_RGB_8U, SL_COLOR_RGB_FLOAT = 14, 19
SL_COLOR_RGBA_8U, SL_COLOR_RGBA_FLOAT = 21, 26
SL_COLOR_RGB_332, SL_COLOR_RGB_565, SL_COLOR_RGBA_5551, SL_COLOR_RGBA_4444, SL_COLOR_RGBA_1010102 = 28, 29, 30, 31, 32

SL_PROGRAM_PASSTHROUGH_GOURAUD = 1
SL_PROGRAM_LAMBERT_TEXTURED = 2
SL_PROGRAM_COLOR_TEXTURED = 3
SL_PROGRAM_LAMBERT_MRT = 4
SL_UNIFORM_SAMPLE_NEAREST = 0x1      # sl_std_uniforms.flags: nearest instead of bilinear sampling
SL_RASTER_SPLIT_DEFAULT = 0xFFFFFFFF  # sl_set_raster_split: the default split policy

_TEXEL_DTYPE = {SL_COLOR_R_8U: (np.uint8, 1), SL_COLOR_RG_8U: (np.uint8, 2), SL_COLOR_RGB_8U: (np.uint8, 3),
                SL_COLOR_RGBA_8U: (np.uint8, 4), SL_COLOR_R_32U: (np.uint32, 1), SL_COLOR_R_HALF: (np.float16, 1),
                SL_COLOR_R_FLOAT: (np.float32, 1), SL_COLOR_RG_FLOAT: (np.float32, 2),
                SL_COLOR_RGB_FLOAT: (np.float32, 3), SL_COLOR_RGBA_FLOAT: (np.float32, 4),
                SL_COLOR_RGB_332: (np.uint8, 1), SL_COLOR_RGB_565: (np.uint16, 1), SL_COLOR_RGBA_5551: (np.uint16, 1),
                SL_COLOR_RGBA_4444: (np.uint16, 1), SL_COLOR_RGBA_1010102: (np.uint32, 1)}
_BYTES_PER_TEXEL = {t: np.dtype(dt).itemsize * n for t, (dt, n) in _TEXEL_DTYPE.items()}


class sl_mesh(C.Structure):
    _fields_ = [("vaoId", C.c_size_t), ("elementBegin", C.c_size_t), ("elementEnd", C.c_size_t),
                ("mode", C.c_uint32), ("materialId", C.c_uint32)]


class sl_vertex_shader(C.Structure):
    _fields_ = [("numVaryings", C.c_uint8), ("cullMode", C.c_uint8), ("programId", C.c_uint32)]


class sl_fragment_shader(C.Structure):
    _fields_ = [("numVaryings", C.c_uint8), ("numOutputs", C.c_uint8), ("blend", C.c_uint8),
                ("depthTest", C.c_uint8), ("depthMask", C.c_uint8), ("programId", C.c_uint32)]


class sl_std_uniforms(C.Structure):
    _fields_ = [("mvp", C.c_float * 16), ("model", C.c_float * 16), ("light_dir", C.c_float * 4),
                ("ambient", C.c_float * 4), ("diffuse", C.c_float * 4), ("tint", C.c_float * 4),
                ("texture", C.c_uint32), ("instance_vbo", C.c_uint32), ("flags", C.c_uint32),
                ("reserved", C.c_uint32)]


class sl_frame_stats(C.Structure):
    _fields_ = [("triangles_submitted", C.c_uint64), ("triangles_binned", C.c_uint64),
                ("tile_pairs", C.c_uint64), ("fragments_tested", C.c_uint64),
                ("fragments_shaded", C.c_uint64), ("frames", C.c_uint64),
                ("ms_vertex", C.c_double), ("ms_bin", C.c_double), ("ms_raster", C.c_double),
                ("ms_total", C.c_double), ("ms_setup_kernel", C.c_double), ("ms_clip_kernel", C.c_double),
                ("setup_launches", C.c_uint64), ("raster_launches", C.c_uint64),
                ("ms_backward", C.c_double), ("backward_launches", C.c_uint64),
                ("visibility_rebuilds", C.c_uint64)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


_P = C.c_void_p
_SZ = C.c_size_t
_I = C.c_int
_U = C.c_uint
_SIGS = {
    "context_create": (_I, [C.POINTER(_P)]),
    "context_destroy": (None, [_P]),
    "last_error": (C.c_char_p, [_P]),
    "num_threads": (_U, [_P]),
    "set_num_threads": (_U, [_P, _U]),
    "set_raster_config": (_I, [_P, _I]),
    "get_raster_config": (_I, [_P, _I, C.POINTER(_I), C.POINTER(_I), C.POINTER(_I)]),
    "set_raster_split": (_I, [_P, C.c_uint32]),
    "set_raster_occupancy": (_I, [_P, C.c_uint32]),
    "set_raster_depth_reject": (_I, [_P, _I]),
    "set_viewport": (_I, [_P, C.c_int32, C.c_int32, C.c_int32, C.c_int32]),
    "set_scissor": (_I, [_P, C.c_int32, C.c_int32, C.c_int32, C.c_int32]),
    "create_vbo": (_SZ, [_P]),
    "vbo_init": (_I, [_P, _SZ, _SZ, _P]),
    "vbo_assign": (_I, [_P, _SZ, _P, _SZ, _SZ]),
    "vbo_read": (_I, [_P, _SZ, _P, _SZ, _SZ]),
    "create_ibo": (_SZ, [_P]),
    "ibo_init": (_I, [_P, _SZ, C.c_uint32, _I, _P]),
    "create_vao": (_SZ, [_P]),
    "vao_set_vertex_buffer": (_I, [_P, _SZ, _SZ]),
    "vao_set_index_buffer": (_I, [_P, _SZ, _SZ]),
    "vao_set_num_bindings": (_I, [_P, _SZ, _SZ]),
    "vao_set_binding": (_I, [_P, _SZ, _SZ, C.c_ssize_t, C.c_ssize_t, _I, _I]),
    "create_texture": (_SZ, [_P]),
    "texture_init": (_I, [_P, _SZ, _I, C.c_uint16, C.c_uint16, C.c_uint16]),
    "texture_set_texels": (_I, [_P, _SZ, _P, _SZ]),
    "texture_read": (_I, [_P, _SZ, _P, _SZ]),
    "create_framebuffer": (_SZ, [_P]),
    "fbo_reserve_color_buffers": (_I, [_P, _SZ, _U]),
    "fbo_attach_color_buffer": (_I, [_P, _SZ, _U, _SZ]),
    "fbo_attach_depth_buffer": (_I, [_P, _SZ, _SZ]),
    "fbo_attach_debug_buffers": (_I, [_P, _SZ, _SZ, _SZ]),
    "fbo_valid": (_I, [_P, _SZ]),
    "create_ubo": (_SZ, [_P]),
    "ubo_assign": (_I, [_P, _SZ, _P, _SZ, _SZ]),
    "create_shader": (_SZ, [_P, C.POINTER(sl_vertex_shader), C.POINTER(sl_fragment_shader), _SZ]),
    "draw": (_I, [_P, C.POINTER(sl_mesh), _SZ, _SZ]),
    "draw_multiple": (_I, [_P, C.POINTER(sl_mesh), _SZ, _SZ, _SZ]),
    "draw_instanced": (_I, [_P, C.POINTER(sl_mesh), _SZ, _SZ, _SZ]),
    "draw_static": (_I, [_P, C.POINTER(sl_mesh), _SZ, _SZ]),
    "clear_color_buffer": (_I, [_P, _SZ, _U, C.POINTER(C.c_double)]),
    "clear_depth_buffer": (_I, [_P, _SZ, C.c_double]),
    "clear_framebuffer": (_I, [_P, _SZ, _U, C.POINTER(C.c_double), C.c_double]),
    "finish": (_I, [_P]),
    "frame_graph_capture": (_I, [_P]),
    "frame_graph_launch": (_I, [_P]),
    "frame_graph_release": (_I, [_P]),
    "blit": (_I, [_P, _SZ, _SZ]),
    "blit_region": (_I, [_P, _SZ, _SZ] + [C.c_uint16] * 8),
    "framebuffer_attach_visibility": (_I, [_P, _SZ, _SZ]),
    "vbo_attach_gradient": (_I, [_P, _SZ, _SZ]),
    "texture_attach_gradient": (_I, [_P, _SZ, _SZ]),
    "gradients_zero": (_I, [_P]),
    "loss_l2_grad": (_I, [_P, _SZ, _SZ, _SZ, C.c_float]),
    "framebuffer_backward": (_I, [_P, _SZ, _SZ]),
    "framebuffer_attach_positions": (_I, [_P, _SZ, _SZ, _SZ, C.c_uint32]),
    "smooth_coverage": (_I, [_P, _SZ, _SZ, C.c_float]),
    "smooth_coverage_backward": (_I, [_P, _SZ, _SZ, C.c_float, _SZ]),
    "get_stats": (_I, [_P, C.POINTER(sl_frame_stats)]),
    "reset_stats": (_I, [_P]),
}


class SLError(RuntimeError):
    pass


class Library:
    """A loaded implementation of the ABI (functions looked up as prefix+name)."""

    def __init__(self, cdll: C.CDLL, prefix: str, extra_sigs: dict | None = None):
        self.cdll = cdll
        self.prefix = prefix
        self.path = getattr(cdll, "_name", "?")
        sigs = dict(_SIGS)
        if extra_sigs:
            sigs.update(extra_sigs)
        for name, (res, args) in sigs.items():
            try:
                f = getattr(cdll, prefix + name)
            except AttributeError:
                continue
            f.restype = res
            f.argtypes = args
            setattr(self, name, f)

    def missing_symbols(self, names=None):
        names = names or list(_SIGS)
        return [n for n in names if not hasattr(self, n)]


def _buf(arr):
    arr = np.ascontiguousarray(arr)
    return arr, arr.ctypes.data_as(_P)


class SLContext:
    """Python mirror of SL_Context over the C ABI (test and benchmark harness)."""

    def __init__(self, lib: Library):
        self.lib = lib
        h = _P()
        rc = lib.context_create(C.byref(h))
        if rc != SL_OK or not h.value:
            raise SLError(f"context_create failed ({rc})")
        self.h = h
        self._tex_info = {}

    def close(self):
        if self.h:
            self.lib.context_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc != SL_OK:
            msg = self.lib.last_error(self.h)
            raise SLError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")
        return rc

    # ---- threads / viewport ----
    def num_threads(self, n=None):
        if n is None:
            return self.lib.num_threads(self.h)
        return self.lib.set_num_threads(self.h, n)

    def set_raster_config(self, config):
        return self._check(self.lib.set_raster_config(self.h, config), "set_raster_config")

    def set_raster_split(self, part_records):
        """Records per part of a split heavy tile (0: no split;
        SL_RASTER_SPLIT_DEFAULT: the library's default policy)."""
        return self._check(self.lib.set_raster_split(self.h, part_records), "set_raster_split")

    def set_raster_occupancy(self, workgroups_per_cu):
        """Cap on resident raster workgroups per CU (0: none)."""
        return self._check(self.lib.set_raster_occupancy(self.h, workgroups_per_cu), "set_raster_occupancy")

    def set_raster_depth_reject(self, enable):
        """Raster variant with per-wave depth rejection on (True) or off."""
        return self._check(self.lib.set_raster_depth_reject(self.h, 1 if enable else 0), "set_raster_depth_reject")

    def raster_config(self, config=-1):
        w, h, s = _I(), _I(), _I()
        idx = self.lib.get_raster_config(self.h, config, C.byref(w), C.byref(h), C.byref(s))
        return idx, (w.value, h.value, s.value)

    def set_viewport(self, x, y, w, h):
        return self._check(self.lib.set_viewport(self.h, x, y, w, h), "set_viewport")

    def set_scissor(self, x, y, w, h):
        return self._check(self.lib.set_scissor(self.h, x, y, w, h), "set_scissor")

    # ---- buffers ----
    def create_vbo(self, data=None):
        vbo = self.lib.create_vbo(self.h)
        if data is not None:
            arr, p = _buf(data)
            self._check(self.lib.vbo_init(self.h, vbo, arr.nbytes, p), "vbo_init")
        return vbo

    def vbo_read(self, vbo, nbytes, offset=0, dtype=np.uint8):
        out = np.empty(nbytes, dtype=np.uint8)
        self._check(self.lib.vbo_read(self.h, vbo, out.ctypes.data_as(C.c_void_p), offset, nbytes), "vbo_read")
        return out.view(dtype)

    def vbo_assign(self, vbo, data, offset=0):
        arr, p = _buf(data)
        return self._check(self.lib.vbo_assign(self.h, vbo, p, offset, arr.nbytes), "vbo_assign")

    def create_ibo(self, indices=None):
        ibo = self.lib.create_ibo(self.h)
        if indices is not None:
            arr = np.ascontiguousarray(indices)
            t = {np.dtype(np.uint8): VERTEX_DATA_BYTE, np.dtype(np.uint16): VERTEX_DATA_SHORT,
                 np.dtype(np.uint32): VERTEX_DATA_INT}[arr.dtype]
            self._check(self.lib.ibo_init(self.h, ibo, arr.size, t, arr.ctypes.data_as(_P)), "ibo_init")
        return ibo

    def vao_set_index_buffer(self, vao, ibo):
        return self._check(self.lib.vao_set_index_buffer(self.h, vao, ibo), "vao_set_index_buffer")

    def create_vao(self, vbo=None, ibo=None, bindings=()):
        vao = self.lib.create_vao(self.h)
        if vbo is not None:
            self._check(self.lib.vao_set_vertex_buffer(self.h, vao, vbo), "vao_set_vertex_buffer")
        if ibo is not None:
            self._check(self.lib.vao_set_index_buffer(self.h, vao, ibo), "vao_set_index_buffer")
        if bindings:
            n = self.lib.vao_set_num_bindings(self.h, vao, len(bindings))
            if n != len(bindings):
                raise SLError(f"vao_set_num_bindings returned {n}")
            for i, (off, stride, dims, dtype) in enumerate(bindings):
                self._check(self.lib.vao_set_binding(self.h, vao, i, off, stride, dims, dtype), "vao_set_binding")
        return vao

    # ---- textures ----
    def create_texture(self, color_type, w, h, d=1, texels=None):
        tex = self.lib.create_texture(self.h)
        self._check(self.lib.texture_init(self.h, tex, color_type, w, h, d), "texture_init")
        self._tex_info[tex] = (color_type, w, h, d)
        if texels is not None:
            self.texture_set_texels(tex, texels)
        return tex

    def texture_set_texels(self, tex, texels):
        arr, p = _buf(texels)
        return self._check(self.lib.texture_set_texels(self.h, tex, p, arr.nbytes), "texture_set_texels")

    def texture_read(self, tex):
        color_type, w, h, d = self._tex_info[tex]
        dt, ch = _TEXEL_DTYPE[color_type]
        shape = (d, h, w, ch) if d > 1 else (h, w, ch)
        out = np.empty(shape, dtype=dt)
        self._check(self.lib.texture_read(self.h, tex, out.ctypes.data_as(_P), out.nbytes), "texture_read")
        return out[..., 0] if ch == 1 else out

    # ---- framebuffers ----
    def create_framebuffer(self, colors=(), depth=None):
        fbo = self.lib.create_framebuffer(self.h)
        self._check(self.lib.fbo_reserve_color_buffers(self.h, fbo, len(colors)), "fbo_reserve_color_buffers")
        for i, t in enumerate(colors):
            self._check(self.lib.fbo_attach_color_buffer(self.h, fbo, i, t), "fbo_attach_color_buffer")
        if depth is not None:
            self._check(self.lib.fbo_attach_depth_buffer(self.h, fbo, depth), "fbo_attach_depth_buffer")
        return fbo

    def fbo_valid(self, fbo):
        return self.lib.fbo_valid(self.h, fbo)

    def attach_debug_buffers(self, fbo, last_tex, count_tex):
        return self._check(self.lib.fbo_attach_debug_buffers(self.h, fbo, last_tex, count_tex),
                           "fbo_attach_debug_buffers")

    # ---- blit ----
    def blit(self, out_tex, in_tex, src=None, dst=None):
        """SL_Context::blit; src/dst are (x0, y0, x1, y1) rectangles or None for whole textures."""
        if src is None and dst is None:
            return self._check(self.lib.blit(self.h, out_tex, in_tex), "blit")
        return self._check(self.lib.blit_region(self.h, out_tex, in_tex, *src, *dst), "blit_region")

    # ---- differentiable draws ----
    def attach_visibility(self, fbo, tex):
        return self._check(self.lib.framebuffer_attach_visibility(self.h, fbo, tex), "framebuffer_attach_visibility")

    def vbo_attach_gradient(self, vbo, grad_vbo):
        return self._check(self.lib.vbo_attach_gradient(self.h, vbo, grad_vbo), "vbo_attach_gradient")

    def texture_attach_gradient(self, tex, grad_tex):
        return self._check(self.lib.texture_attach_gradient(self.h, tex, grad_tex), "texture_attach_gradient")

    def gradients_zero(self):
        return self._check(self.lib.gradients_zero(self.h), "gradients_zero")

    def loss_l2_grad(self, color_tex, target_tex, grad_tex, scale=1.0):
        return self._check(self.lib.loss_l2_grad(self.h, color_tex, target_tex, grad_tex, scale), "loss_l2_grad")

    def backward(self, fbo, grad_image_tex):
        return self._check(self.lib.framebuffer_backward(self.h, fbo, grad_image_tex), "framebuffer_backward")

    def attach_positions(self, fbo, grad_vbo=SL_INVALID_ID, offset_vbo=SL_INVALID_ID, num_vertices=0):
        """float2 per (saved draw, vertex id): screen-position gradients / offsets."""
        return self._check(self.lib.framebuffer_attach_positions(self.h, fbo, grad_vbo, offset_vbo, num_vertices),
                           "framebuffer_attach_positions")

    def smooth_coverage(self, fbo, out_tex, sigma):
        return self._check(self.lib.smooth_coverage(self.h, fbo, out_tex, sigma), "smooth_coverage")

    def smooth_coverage_backward(self, fbo, grad_smooth_tex, sigma, grad_image_tex):
        return self._check(self.lib.smooth_coverage_backward(self.h, fbo, grad_smooth_tex, sigma, grad_image_tex),
                           "smooth_coverage_backward")

    # ---- uniforms / shaders ----
    def create_ubo(self, data=None):
        ubo = self.lib.create_ubo(self.h)
        if data is not None:
            self.ubo_assign(ubo, data)
        return ubo

    def ubo_assign(self, ubo, data, offset=0):
        if isinstance(data, C.Structure):
            raw = bytes(data)
        else:
            raw = np.ascontiguousarray(data).tobytes()
        buf = C.create_string_buffer(raw, len(raw))
        return self._check(self.lib.ubo_assign(self.h, ubo, buf, offset, len(raw)), "ubo_assign")

    def create_shader(self, vs: sl_vertex_shader, fs: sl_fragment_shader, ubo=SL_INVALID_ID):
        return self.lib.create_shader(self.h, C.byref(vs), C.byref(fs), ubo)

    # ---- draw / clear ----
    def draw(self, mesh: sl_mesh, shader, fbo):
        return self._check(self.lib.draw(self.h, C.byref(mesh), shader, fbo), "draw")

    def draw_multiple(self, meshes, shader, fbo):
        arr = (sl_mesh * len(meshes))(*meshes)
        return self._check(self.lib.draw_multiple(self.h, arr, len(meshes), shader, fbo), "draw_multiple")

    def draw_instanced(self, mesh: sl_mesh, n, shader, fbo):
        return self._check(self.lib.draw_instanced(self.h, C.byref(mesh), n, shader, fbo), "draw_instanced")

    def draw_static(self, mesh: sl_mesh, shader, fbo):
        """sl_draw_static: draw and run the pending batch at once in the
        library's fixed scratch array (no device allocation)."""
        return self._check(self.lib.draw_static(self.h, C.byref(mesh), shader, fbo), "draw_static")

    def clear_color_buffer(self, fbo, att, rgba):
        c = (C.c_double * 4)(*rgba)
        return self._check(self.lib.clear_color_buffer(self.h, fbo, att, c), "clear_color_buffer")

    def clear_depth_buffer(self, fbo, depth):
        return self._check(self.lib.clear_depth_buffer(self.h, fbo, depth), "clear_depth_buffer")

    def clear_framebuffer(self, fbo, att, rgba, depth):
        c = (C.c_double * 4)(*rgba)
        return self._check(self.lib.clear_framebuffer(self.h, fbo, att, c, depth), "clear_framebuffer")

    def finish(self):
        return self._check(self.lib.finish(self.h), "finish")

    def frame_graph_capture(self):
        return self._check(self.lib.frame_graph_capture(self.h), "frame_graph_capture")

    def frame_graph_launch(self):
        return self._check(self.lib.frame_graph_launch(self.h), "frame_graph_launch")

    def frame_graph_release(self):
        return self._check(self.lib.frame_graph_release(self.h), "frame_graph_release")

    def stats(self):
        s = sl_frame_stats()
        self._check(self.lib.get_stats(self.h, C.byref(s)), "get_stats")
        return s.as_dict()

    def reset_stats(self):
        return self._check(self.lib.reset_stats(self.h), "reset_stats")
