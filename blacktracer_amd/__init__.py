"""blacktracer_amd -- MI355X-native implementation of BlackTracer's
data-parallel core, exposed through the C ABI of include/bt_abi.h.

The Python classes mirror the reference's host interface
(annemiekie/BlackTracer, BlackTracer2/BlackTracer/):

    BlackHole(a)                                     BlackHole.h:8-15
    Camera(theta, phi, r, br, btheta, bphi)          Camera.h:40-61
    Grid(max_level, start_level, angle, cam, bh)     Grid.h:605-624
    Grid.integration_wrapper(theta, phi)             Grid.h:514-556
    Viewer, StarProcessor, make_image                Viewer.h, StarProcessor.h,
                                                     MakeImageKernel.cu:1206

All work runs on the HIP device through libbtamd.so; there is no CPU path.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import (BT_ERR_BUDGET, BT_PREC_FP64, BT_PREC_MIXED, BtError, BtGridParams, BtIntegParams, BtPose,
                   BtRenderFlags, BtTraceStats, check)

__all__ = ["Context", "BlackHole", "Camera", "Grid", "IntegParams", "trace_rays", "BtError",
           "BT_PREC_FP64", "BT_PREC_MIXED", "default_context"]


class Context:
    """A bt_context: one HIP device, its stream, work buffers and timers."""

    def __init__(self, device=0):
        L = _lib.load()
        h = C.c_void_p()
        check(L.bt_context_create(device, C.byref(h)))
        self.handle = h
        self.device = device

    def sync(self):
        check(_lib.load().bt_context_sync(self.handle))

    def timing(self, enable=True):
        check(_lib.load().bt_timing_enable(self.handle, 1 if enable else 0))

    def timer(self, kernel):
        n = C.c_int64()
        ms = C.c_double()
        check(_lib.load().bt_timer_get(self.handle, kernel.encode(), C.byref(n), C.byref(ms)))
        return n.value, ms.value

    def reset_timers(self):
        check(_lib.load().bt_timer_reset(self.handle))

    def tune(self):
        """The persistent integrators' launch shape: {chunk, block, waves}."""
        t = _lib.BtTune()
        check(_lib.load().bt_tune_get(self.handle, C.byref(t)))
        return {"chunk": t.chunk, "block": t.block, "waves": t.waves}

    def set_tune(self, chunk=128, block=256, waves=0):
        t = _lib.BtTune(chunk, block, waves, 0.0)
        check(_lib.load().bt_tune_set(self.handle, C.byref(t)))

    def autotune(self, camera, params=None, n_rays=1 << 20):
        """Times about n_rays directions of the camera's equal-angle sky over
        launch shapes, keeps the fastest; returns {chunk, block, waves,
        rays_per_lane, ms}.  Calibrate at the size that will be traced."""
        t = _lib.BtTune()
        ms = C.c_double()
        params = params or IntegParams()
        check(_lib.load().bt_autotune(self.handle, C.byref(camera.pose()), C.byref(params), int(n_rays),
                                      C.byref(t), C.byref(ms)))
        return {"chunk": t.chunk, "block": t.block, "waves": t.waves,
                "rays_per_lane": t.rays_per_lane, "ms": ms.value}

    def alloc(self, nbytes):
        """A device buffer owned by the caller (DeviceBuffer)."""
        return DeviceBuffer(self, nbytes)

    def close(self):
        if getattr(self, "handle", None):
            _lib.load().bt_context_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceBuffer:
    """Device memory from bt_device_alloc, freed with the object."""

    def __init__(self, ctx, nbytes):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        check(_lib.load().bt_device_alloc(ctx.handle, self.nbytes, C.byref(p)))
        self.ptr = p

    def to_host(self, out):
        check(_lib.load().bt_memcpy(self.ctx.handle, out.ctypes.data_as(C.c_void_p), self.ptr,
                                    min(self.nbytes, out.nbytes), 0))
        return out

    def from_host(self, arr):
        check(_lib.load().bt_memcpy(self.ctx.handle, self.ptr, arr.ctypes.data_as(C.c_void_p),
                                    min(self.nbytes, arr.nbytes), 1))

    def close(self):
        if getattr(self, "ptr", None):
            _lib.load().bt_device_free(self.ctx.handle, self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default = None


def default_context():
    global _default
    if _default is None:
        _default = Context(0)
    return _default


class BlackHole:
    """BlackHole(afactor) (BlackHole.h:8-15).  The spin is carried by the
    pose of each call instead of a process global."""

    def __init__(self, afactor):
        self.a = float(afactor)


class Camera:
    """Camera(theta, phi, r, br, btheta, bphi) on the circular equatorial
    orbit (Camera.h:40-52), or Camera(theta, phi, r, speed=...) (Camera.h:27-38)."""

    def __init__(self, theta, phi, r, br=0.0, btheta=0.0, bphi=1.0, speed=None, black_hole=None):
        self.theta, self.phi, self.r = float(theta), float(phi), float(r)
        self.br, self.btheta, self.bphi = float(br), float(btheta), float(bphi)
        self.user_speed = speed is not None
        self._speed_in = float(speed) if speed is not None else 0.0
        self.a = black_hole.a if black_hole is not None else None
        if self.a is not None:
            self._init_forms()

    def pose(self, a=None):
        a = self.a if a is None else a
        if a is None:
            raise ValueError("Camera needs a BlackHole (spin) for its pose")
        return BtPose(a=a, theta=self.theta, phi=self.phi, r=self.r, br=self.br,
                      btheta=self.btheta, bphi=self.bphi, speed=self._speed_in,
                      user_speed=1 if self.user_speed else 0)

    def _init_forms(self):
        out = (C.c_double * 6)()
        cam7 = (C.c_float * 7)()
        check(_lib.load().bt_camera_init(C.byref(self.pose()), out, cam7))
        self.speed, self.alpha, self.w, self.wbar, self.Delta, self.ro = list(out)
        self.param_array = np.array(list(cam7), dtype=np.float32)

    def bind(self, black_hole):
        self.a = black_hole.a
        self._init_forms()
        return self

    def getParamArray(self):  # noqa: N802  (reference name, Camera.h:67)
        return self.param_array.copy()


def IntegParams(precision=BT_PREC_FP64, **kw):  # noqa: N802
    """bt_integ_params with the reference defaults (Metric.h:307-318), overridable."""
    p = BtIntegParams()
    _lib.load().bt_integ_params_default(C.byref(p))
    p.precision = precision
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def radius_stop_params(precision=BT_PREC_FP64):
    """The radius-stop variant of the full-sky build: rel tol 1e-8, escape at
    r > 1000M, capture at r < 1.01 r+ (no a-priori capture test)."""
    return IntegParams(precision, eps=1e-8, max_steps=10000, apriori_capture=0, r_escape=1000.0,
                       r_capture_factor=1.01)


def _stats_dict(s):
    return {n: getattr(s, n) for n, _ in BtTraceStats._fields_}


def trace_rays(camera, theta, phi, params=None, ctx=None, return_stats=False):
    """Grid::integration_wrapper on the device: returns (theta', phi', step)."""
    ctx = ctx or default_context()
    th = np.ascontiguousarray(theta, dtype=np.float64).copy()
    ph = np.ascontiguousarray(phi, dtype=np.float64).copy()
    st = np.zeros(th.shape, dtype=np.int32)
    stats = BtTraceStats()
    params = params or IntegParams()
    check(_lib.load().bt_trace_rays(ctx.handle, C.byref(camera.pose()), C.byref(params), th.size,
                                    th.ctypes.data_as(_lib.DP), ph.ctypes.data_as(_lib.DP),
                                    st.ctypes.data_as(_lib.I32P), C.byref(stats)))
    if return_stats:
        return th, ph, st, _stats_dict(stats)
    return th, ph, st


class Grid:
    """Grid(maxLevel, startLevel, angle, camera, blackhole) (Grid.h:605-624),
    built and kept resident on the device.  ``cel``/``steps``/``levels``
    are dense host copies made on first access."""

    def __init__(self, max_level, start_level, angle, camera, black_hole, params=None,
                 refine_thresh=0.015, fix_tvertices=True, ctx=None):
        self.ctx = ctx or default_context()
        L = _lib.load()
        gp = BtGridParams()
        L.bt_grid_params_default(C.byref(gp))
        gp.refine_thresh = refine_thresh
        gp.fix_tvertices = 1 if fix_tvertices else 0
        self.params = params or IntegParams()
        h = C.c_void_p()
        check(L.bt_grid_build(self.ctx.handle, C.byref(camera.pose(black_hole.a)),
                              C.byref(self.params), max_level, start_level, 1 if angle else 0,
                              C.byref(gp), C.byref(h)))
        self._adopt(h)

    def _adopt(self, h):
        L = _lib.load()
        self.handle = h
        vals = [C.c_int() for _ in range(5)]
        rays = C.c_int64()
        check(L.bt_grid_info(h, *[C.byref(v) for v in vals], C.byref(rays)))
        self.N, self.M, self.MAXLEVEL, self.STARTLVL, self.equafactor = [v.value for v in vals]
        self.rays_traced = rays.value
        self._cel = self._steps = self._levels = None

    @classmethod
    def build_batch(cls, cameras, black_holes, max_levels, start_levels, angle=True, params=None,
                    refine_thresh=0.015, fix_tvertices=True, ctx=None):
        """Several grids built together (bt_grid_build_batch): returns
        (list of Grid, batch stats dict)."""
        ctx = ctx or default_context()
        L = _lib.load()
        n = len(cameras)
        gp = BtGridParams()
        L.bt_grid_params_default(C.byref(gp))
        gp.refine_thresh = refine_thresh
        gp.fix_tvertices = 1 if fix_tvertices else 0
        params = params or IntegParams()
        poses = (BtPose * n)(*[cam.pose(bh.a) for cam, bh in zip(cameras, black_holes)])
        mls = (C.c_int * n)(*max_levels)
        sls = (C.c_int * n)(*start_levels)
        hs = (C.c_void_p * n)()
        st = BtTraceStats()
        check(L.bt_grid_build_batch(ctx.handle, n, poses, C.byref(params), mls, sls, 1 if angle else 0,
                                    C.byref(gp), hs, C.byref(st)))
        grids = []
        for k in range(n):
            g = cls.__new__(cls)
            g.ctx = ctx
            g.params = params
            g._adopt(C.c_void_p(hs[k]))
            grids.append(g)
        return grids, _stats_dict(st)

    def stats(self):
        s = BtTraceStats()
        check(_lib.load().bt_grid_stats(self.handle, C.byref(s)))
        return _stats_dict(s)

    def _export(self):
        N, M = self.N, self.M
        cel = np.empty((N, M, 2), dtype=np.float64)
        steps = np.empty((N, M), dtype=np.int32)
        lv = np.empty((N, M), dtype=np.int8)
        check(_lib.load().bt_grid_export(self.handle, cel.ctypes.data_as(_lib.DP),
                                         steps.ctypes.data_as(_lib.I32P),
                                         lv.ctypes.data_as(_lib.I8P)))
        self._cel, self._steps, self._levels = cel, steps, lv

    @property
    def cel(self):
        if self._cel is None:
            self._export()
        return self._cel

    @property
    def steps(self):
        if self._steps is None:
            self._export()
        return self._steps

    @property
    def levels(self):
        if self._levels is None:
            self._export()
        return self._levels

    def write(self, path):
        """A .grid file (Grid::serialize, Grid.h:33-39) of this grid."""
        check(_lib.load().bt_grid_file_write(str(path).encode(), self.handle))

    def psh(self):
        """saveAsGpuHash (Grid.h:642-654): (hash_table, offset_table, hash_pos_tag, hw, ow)."""
        L = _lib.load()
        hw, ow = C.c_int(), C.c_int()
        check(L.bt_grid_export_psh(self.handle, C.byref(hw), C.byref(ow), None, None, None))
        ht = np.zeros(2 * hw.value * hw.value, dtype=np.float32)
        ot = np.zeros(2 * ow.value * ow.value, dtype=np.int32)
        tag = np.zeros(2 * hw.value * hw.value, dtype=np.int32)
        check(L.bt_grid_export_psh(self.handle, C.byref(hw), C.byref(ow),
                                   ht.ctypes.data_as(_lib.FP), ot.ctypes.data_as(_lib.I32P),
                                   tag.ctypes.data_as(_lib.I32P)))
        return ht, ot, tag, hw.value, ow.value

    def close(self):
        if getattr(self, "handle", None):
            _lib.load().bt_grid_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------- render


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def RenderFlags(star=True, map=True, lr=True, hor=0.0, ver=0.0, offset=0.0):  # noqa: N802
    """bt_render_flags (MakeImageKernel.cu:955-960, :1413)."""
    return BtRenderFlags(star=int(star), map=int(map), lr=int(lr), hor=hor, ver=ver, offset=offset)


def viewer_matrix(view_angle, off_up, off_right, width, height, sphere=False):
    """Viewer(viewangle, offUp, offRight, w, h, sphere).viewMatrix, (h+1, w+1, 2) float32."""
    out = np.zeros((height + 1, width + 1, 2), dtype=np.float32)
    check(_lib.load().bt_viewer_matrix(view_angle, off_up, off_right, width, height,
                                       1 if sphere else 0, out.ctypes.data_as(_lib.FP)))
    return out


def star_tree(stars4, tree_level):
    """StarProcessor::makeTree: (tree, star_pos (S,2), star_mag (S,2))."""
    s4 = _f32(stars4)
    S = s4.shape[0]
    tree = np.zeros((1 << (tree_level + 1)) - 1, dtype=np.int32)
    pos = np.zeros((S, 2), dtype=np.float32)
    mag = np.zeros((S, 2), dtype=np.float32)
    check(_lib.load().bt_star_tree(s4.ctypes.data_as(_lib.FP), S, tree_level,
                                   tree.ctypes.data_as(_lib.I32P), pos.ctypes.data_as(_lib.FP),
                                   mag.ctypes.data_as(_lib.FP)))
    return tree, pos, mag


def sky_add_low_stars(sky_bgr, stars4, magnitude_cut):
    """addLowLightStarsToImage (StarProcessor.h:108-158): a copy of the BGR8
    sky with every star fainter than magnitude_cut painted in."""
    sky = np.ascontiguousarray(sky_bgr, dtype=np.uint8).copy()
    s4 = _f32(stars4).reshape(-1, 4)
    check(_lib.load().bt_sky_add_low_stars(sky.ctypes.data_as(_lib.U8P), sky.shape[0], sky.shape[1],
                                           s4.ctypes.data_as(_lib.FP), s4.shape[0], float(magnitude_cut)))
    return sky


class Scene:
    """Device-resident stars (star tree, StarProcessor.h) and summed-area sky
    table (MakeImageKernel.cu:1312-1331)."""

    def __init__(self, stars4, sky_bgr, tree_level=8, ctx=None):
        self.ctx = ctx or default_context()
        s4 = _f32(stars4)
        sky = np.ascontiguousarray(sky_bgr, dtype=np.uint8)
        self.rows, self.cols = sky.shape[:2]
        self.tree_level = tree_level
        h = C.c_void_p()
        check(_lib.load().bt_scene_create(self.ctx.handle, s4.ctypes.data_as(_lib.FP), s4.shape[0],
                                          tree_level, sky.ctypes.data_as(_lib.U8P), self.rows,
                                          self.cols, C.byref(h)))
        self.handle = h

    def table(self):
        t = np.zeros((self.rows, self.cols, 4), dtype=np.float32)
        check(_lib.load().bt_scene_export_table(self.handle, t.ctypes.data_as(_lib.FP)))
        return t

    @classmethod
    def from_tree(cls, tree, tree_level, star_pos, star_mag, sky_bgr, ctx=None):
        """A scene from a star tree already built (the arrays of a .star file)."""
        self = cls.__new__(cls)
        self.ctx = ctx or default_context()
        tr = np.ascontiguousarray(tree, dtype=np.int32)
        pos, mag = _f32(star_pos), _f32(star_mag)
        sky = np.ascontiguousarray(sky_bgr, dtype=np.uint8)
        self.rows, self.cols = sky.shape[:2]
        self.tree_level = tree_level
        h = C.c_void_p()
        check(_lib.load().bt_scene_create_tree(self.ctx.handle, tr.ctypes.data_as(_lib.I32P), tree_level,
                                               pos.ctypes.data_as(_lib.FP), mag.ctypes.data_as(_lib.FP),
                                               pos.size // 2, sky.ctypes.data_as(_lib.U8P), self.rows,
                                               self.cols, C.byref(h)))
        self.handle = h
        return self

    def close(self):
        if getattr(self, "handle", None):
            _lib.load().bt_scene_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class View:
    """Device-resident Viewer matrix of one resolution."""

    def __init__(self, width, height, view_angle=np.pi / 2, off_up=0.0, off_right=0.25 * np.pi / 4,
                 sphere=False, ctx=None):
        self.ctx = ctx or default_context()
        self.width, self.height = width, height
        h = C.c_void_p()
        check(_lib.load().bt_view_create(self.ctx.handle, view_angle, off_up, off_right, width,
                                         height, 1 if sphere else 0, C.byref(h)))
        self.handle = h

    def close(self):
        if getattr(self, "handle", None):
            _lib.load().bt_view_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def frame_bytes(width, height, step=1):
    """Device bytes one frame of this size works in."""
    return int(_lib.load().bt_frame_bytes(int(width), int(height), int(step)))


def server_footprint(width, height, max_level=10, step=1):
    """What a server counts against its HBM budget: {slot, view, grid, build}
    bytes for frames of width x height and grids of max_level."""
    L = _lib.load()
    cfg = _lib.BtServerConfig()
    L.bt_server_config_default(C.byref(cfg))
    cfg.max_level = max_level
    cfg.step = step
    b = _lib.BtServerBytes()
    check(L.bt_server_footprint(C.byref(cfg), int(width), int(height), C.byref(b)))
    return {n: int(getattr(b, n)) for n, _ in _lib.BtServerBytes._fields_}


def render_frame_device(grid, scene, view, out_buffer, flags=None, cam7=None, step=1, ctx=None):
    """One frame into a device buffer (asynchronous; no host copy)."""
    ctx = ctx or grid.ctx
    flags = flags or RenderFlags()
    c7 = None if cam7 is None else _f32(cam7).ctypes.data_as(_lib.FP)
    check(_lib.load().bt_render_frame(ctx.handle, grid.handle, scene.handle, view.handle, c7,
                                      C.byref(flags), step, None, out_buffer.ptr))


def render_frame(grid, scene, view, flags=None, cam7=None, step=1, stages=False, ctx=None):
    """One frame (BGRA uint8, (h, w, 4)) of a device-resident grid."""
    ctx = ctx or grid.ctx
    W, H = view.width, view.height
    out = np.zeros((H, W, 4), dtype=np.uint8)
    flags = flags or RenderFlags()
    c7 = None if cam7 is None else _f32(cam7).ctypes.data_as(_lib.FP)
    L = _lib.load()
    if not stages:
        check(L.bt_render_frame(ctx.handle, grid.handle, scene.handle, view.handle, c7,
                                C.byref(flags), step, out.ctypes.data_as(_lib.U8P), None))
        return out
    thphi = np.zeros((H + 1, W + 1, 2), dtype=np.float32)
    gap = np.zeros((H + 1, W + 1), dtype=np.int32)
    bh = np.zeros((H, W), dtype=np.uint8)
    area = np.zeros((H, W), dtype=np.float32)
    star = np.zeros((H, W, 3), dtype=np.float32)
    check(L.bt_render_frame_stages(ctx.handle, grid.handle, scene.handle, view.handle, c7,
                                   C.byref(flags), step, out.ctypes.data_as(_lib.U8P), None,
                                   thphi.ctypes.data_as(_lib.FP), gap.ctypes.data_as(_lib.I32P),
                                   bh.ctypes.data_as(_lib.U8P), area.ctypes.data_as(_lib.FP),
                                   star.ctypes.data_as(_lib.FP)))
    return out, dict(thphi=thphi, gap=gap, bh=bh, area=area, star=star)


def render_frame_blend(grid0, grid1, alpha, scene, view, flags=None, cam14=None, step=1,
                       stages=False, ctx=None, out_buffer=None):
    """A frame between the poses of two grids at fraction alpha (BGRA uint8,
    (h, w, 4)): the reference's G > 1 frame (MakeImageKernel.cu:386-510).
    With stages, also the pixel-corner celestial angles and the shadow
    border (2 + 2000 float2).  With out_buffer (a device buffer) the frame
    goes there, asynchronously."""
    ctx = ctx or grid0.ctx
    W, H = view.width, view.height
    flags = flags or RenderFlags()
    c14 = None if cam14 is None else _f32(cam14).ctypes.data_as(_lib.FP)
    if out_buffer is not None and not stages:
        check(_lib.load().bt_render_frame_blend(
            ctx.handle, grid0.handle, grid1.handle, float(alpha), scene.handle, view.handle, c14,
            C.byref(flags), step, None, out_buffer.ptr, None, None))
        return None
    out = np.zeros((H, W, 4), dtype=np.uint8)
    thphi = np.zeros((H + 1, W + 1, 2), dtype=np.float32) if stages else None
    border = np.zeros((2 + 2000, 2), dtype=np.float32) if stages else None
    check(_lib.load().bt_render_frame_blend(
        ctx.handle, grid0.handle, grid1.handle, float(alpha), scene.handle, view.handle, c14,
        C.byref(flags), step, out.ctypes.data_as(_lib.U8P), None,
        None if thphi is None else thphi.ctypes.data_as(_lib.FP),
        None if border is None else border.ctypes.data_as(_lib.FP)))
    if stages:
        return out, dict(thphi=thphi, border=border)
    return out


class Trails:
    """State of a frame sequence with star trails (and diffraction when a
    sprite of (width // 16, width // 16, 3) uint8 is given): each
    render() is the next frame (MakeImageKernel.cu:98-174,775-800)."""

    def __init__(self, scene, width, height, diffraction=None, ctx=None):
        self.ctx = ctx or scene.ctx
        self.width, self.height = width, height
        self._sprite = None
        sp = None
        if diffraction is not None:
            fs = width // 16
            if fs < 1:
                raise ValueError("a diffraction sprite needs a frame at least 16 pixels wide")
            self._sprite = np.ascontiguousarray(diffraction, dtype=np.uint8).reshape(fs, fs, 3)
            sp = self._sprite.ctypes.data_as(_lib.U8P)
        h = C.c_void_p()
        check(_lib.load().bt_trails_create(self.ctx.handle, scene.handle, width, height, sp, C.byref(h)))
        self.handle = h

    def render(self, grid, scene, view, flags=None, cam7=None, step=1, stages=False, out_buffer=None):
        """The next frame (BGRA uint8, (h, w, 4)); with stages also the summed
        star light and this frame's trails ((h, w, 3) float32 each).  With
        out_buffer (a device buffer) the frame goes there, asynchronously."""
        W, H = view.width, view.height
        flags = flags or RenderFlags()
        c7 = None if cam7 is None else _f32(cam7).ctypes.data_as(_lib.FP)
        if out_buffer is not None and not stages:
            check(_lib.load().bt_render_frame_trails(
                self.ctx.handle, grid.handle, scene.handle, view.handle, c7, C.byref(flags), step,
                self.handle, None, out_buffer.ptr, None, None))
            return None
        out = np.zeros((H, W, 4), dtype=np.uint8)
        star = np.zeros((H, W, 3), dtype=np.float32) if stages else None
        trail = np.zeros((H, W, 3), dtype=np.float32) if stages else None
        check(_lib.load().bt_render_frame_trails(
            self.ctx.handle, grid.handle, scene.handle, view.handle, c7, C.byref(flags), step,
            self.handle, out.ctypes.data_as(_lib.U8P), None,
            None if star is None else star.ctypes.data_as(_lib.FP),
            None if trail is None else trail.ctypes.data_as(_lib.FP)))
        if stages:
            return out, dict(star=star, trail=trail)
        return out

    def reset(self):
        _lib.load().bt_trails_reset(self.handle)

    def stats(self):
        f, d = C.c_int64(), C.c_int64()
        check(_lib.load().bt_trails_stats(self.handle, C.byref(f), C.byref(d)))
        return dict(frames=f.value, dropped=d.value)

    def close(self):
        if getattr(self, "handle", None):
            _lib.load().bt_trails_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def make_image(stars, star_tree_arr, cam_param, magnitude, tree_level, M, N, step, sky_bgr,
               viewthing, GM, GN, gridlvl, offset_table, hash_table, hash_pos_tag, hw, ow,
               flags=None, ctx=None):
    """makeImage (MakeImageKernel.cu:1206) with G = 1; returns the BGRA frame."""
    ctx = ctx or default_context()
    out = np.zeros((N, M, 4), dtype=np.uint8)
    sky = np.ascontiguousarray(sky_bgr, dtype=np.uint8)
    ts = np.array([hw, ow], dtype=np.int32)
    ot = np.ascontiguousarray(offset_table, dtype=np.int32)
    ht = _f32(hash_table)
    tag = np.ascontiguousarray(hash_pos_tag, dtype=np.int32)
    st = _f32(stars)
    mg = _f32(magnitude)
    tr = np.ascontiguousarray(star_tree_arr, dtype=np.int32)
    flags = flags or RenderFlags()
    check(_lib.load().bt_make_image(
        ctx.handle, st.ctypes.data_as(_lib.FP), tr.ctypes.data_as(_lib.I32P), st.size // 2,
        _f32(cam_param).ctypes.data_as(_lib.FP), mg.ctypes.data_as(_lib.FP), tree_level, M, N,
        step, sky.ctypes.data_as(_lib.U8P), sky.shape[0], sky.shape[1], 1, 0.0, 0.0, None,
        _f32(viewthing).ctypes.data_as(_lib.FP), 0.0, GM, GN, gridlvl,
        ot.ctypes.data_as(_lib.I32P), ht.ctypes.data_as(_lib.FP), tag.ctypes.data_as(_lib.I32P),
        ts.ctypes.data_as(_lib.I32P), ot.size // 2, ht.size // 2, C.byref(flags),
        out.ctypes.data_as(_lib.U8P)))
    return out


# ------------------------------------------- on-disk caches and catalogue


def read_star_catalog(text, magnitude_cut):
    """readStars2 + magnitude cut (StarProcessor.h:34-46,71-96): stars4
    (S, 4) = theta, phi, magnitude, B-V of the stars at or below the cut."""
    raw = text.encode() if isinstance(text, str) else bytes(text)
    L = _lib.load()
    n = C.c_int()
    check(L.bt_star_catalog_parse(raw, len(raw), magnitude_cut, None, 0, C.byref(n)))
    out = np.zeros((n.value, 4), dtype=np.float32)
    if n.value:
        check(L.bt_star_catalog_parse(raw, len(raw), magnitude_cut, out.ctypes.data_as(_lib.FP), n.value,
                                      C.byref(n)))
    return out


def write_star_file(path, tree, tree_level, star_pos, star_mag):
    """A .star file (StarProcessor::serialize) of star_tree()'s arrays."""
    tr = np.ascontiguousarray(tree, dtype=np.int32)
    pos, mag = _f32(star_pos), _f32(star_mag)
    check(_lib.load().bt_star_file_write(str(path).encode(), tr.ctypes.data_as(_lib.I32P), tree_level,
                                         pos.ctypes.data_as(_lib.FP), mag.ctypes.data_as(_lib.FP),
                                         pos.size // 2))


def read_star_file(path):
    """(tree, tree_level, star_pos (S,2), star_mag (S,2)) of a .star file."""
    L = _lib.load()
    S, lv = C.c_int(), C.c_int()
    p = str(path).encode()
    check(L.bt_star_file_read(p, None, None, None, C.byref(S), C.byref(lv)))
    tree = np.zeros((1 << (lv.value + 1)) - 1, dtype=np.int32)
    pos = np.zeros((S.value, 2), dtype=np.float32)
    mag = np.zeros((S.value, 2), dtype=np.float32)
    check(L.bt_star_file_read(p, tree.ctypes.data_as(_lib.I32P), pos.ctypes.data_as(_lib.FP),
                              mag.ctypes.data_as(_lib.FP), C.byref(S), C.byref(lv)))
    return tree, lv.value, pos, mag


def read_grid_file(path):
    """A .grid file (Grid::serialize): dict of max_level, N, M and the PSH
    tables (hash_table, offset_table, hash_pos_tag, hw, ow) that make_image takes."""
    L = _lib.load()
    v = [C.c_int() for _ in range(5)]
    p = str(path).encode()
    check(L.bt_grid_file_read(p, *[C.byref(x) for x in v], None, None, None))
    ml, N, M, hw, ow = (x.value for x in v)
    ht = np.zeros(2 * hw * hw, dtype=np.float32)
    ot = np.zeros(2 * ow * ow, dtype=np.int32)
    tag = np.zeros(2 * hw * hw, dtype=np.int32)
    check(L.bt_grid_file_read(p, *[C.byref(x) for x in v], ht.ctypes.data_as(_lib.FP),
                              ot.ctypes.data_as(_lib.I32P), tag.ctypes.data_as(_lib.I32P)))
    return {"max_level": ml, "N": N, "M": M, "hash_table": ht, "offset_table": ot, "hash_pos_tag": tag,
            "hw": hw, "ow": ow}


# ---------------------------------------------------------------- serving


class Server:
    """Continuous-batching render server (bt_server_*): requests of mixed
    resolution and pose, grids cached in HBM, admission by HBM footprint.

    ``submit`` is non-blocking; the frame lands in the given uint8 array
    (height x width x 4, BGRA) before its completion is returned by ``poll``
    or counted by ``drain``."""

    def __init__(self, scene, hbm_budget=16 << 30, slots=4, max_level=10, start_level=1, params=None,
                 step=1, ctx=None):
        self.ctx = ctx or default_context()
        L = _lib.load()
        cfg = _lib.BtServerConfig()
        L.bt_server_config_default(C.byref(cfg))
        cfg.hbm_budget = hbm_budget
        cfg.slots = slots
        cfg.max_level = max_level
        cfg.start_level = start_level
        if params is not None:
            cfg.integ = params
        cfg.step = step
        h = C.c_void_p()
        check(L.bt_server_create(self.ctx.handle, scene.handle, C.byref(cfg), C.byref(h)))
        self.handle = h
        self.scene = scene          # must outlive the server
        self._outs = {}

    def submit(self, a, r, theta, offset, out, tag, star=True, map=True, lr=True):
        """out: uint8 array (height, width, 4), kept alive until completion."""
        assert out.dtype == np.uint8 and out.ndim == 3 and out.shape[2] == 4 and out.flags["C_CONTIGUOUS"]
        req = _lib.BtRequest(a, r, theta, offset, out.shape[1], out.shape[0], int(star), int(map), int(lr),
                             out.ctypes.data, tag)
        check(_lib.load().bt_server_submit(self.handle, C.byref(req)))
        self._outs[tag] = out

    def poll(self, max_n=256):
        buf = (_lib.BtCompletion * max_n)()
        n = C.c_int()
        check(_lib.load().bt_server_poll(self.handle, buf, max_n, C.byref(n)))
        out = []
        for k in range(n.value):
            c = buf[k]
            self._outs.pop(c.tag, None)
            out.append({"tag": c.tag, "status": c.status, "grid_hit": bool(c.grid_hit),
                        "latency_ms": c.latency_ms, "queue_ms": c.queue_ms, "frame_ms": c.frame_ms})
        return out

    def drain(self, timeout_s=0.0):
        check(_lib.load().bt_server_drain(self.handle, timeout_s))

    def stats(self):
        st = _lib.BtServerStats()
        check(_lib.load().bt_server_stats_get(self.handle, C.byref(st)))
        return {n: getattr(st, n) for n, _ in _lib.BtServerStats._fields_}

    def close(self):
        if getattr(self, "handle", None):
            _lib.load().bt_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
