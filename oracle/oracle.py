"""ctypes bindings for the CPU oracle (test infrastructure only).

Loaded by tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline
leg -- never by the blacktracer_amd product.  The oracle is a plain-C
restatement of the reference (annemiekie/BlackTracer) built from
oracle/bt_oracle*.c; `load_ref()` loads the reference's own Metric.h /
Camera.h / BlackHole.h compiled into oracle/_ref/libbtref.so.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None

DP = C.POINTER(C.c_double)
FP = C.POINTER(C.c_float)
IP = C.POINTER(C.c_int)
I8P = C.POINTER(C.c_int8)
U8P = C.POINTER(C.c_uint8)


class Pose(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "a", "asq", "theta", "phi", "r", "speed", "br", "btheta", "bphi",
        "alpha", "w", "wbar", "Delta", "ro")]


class Params(C.Structure):
    _fields_ = [("eps", C.c_double), ("h1", C.c_double), ("z_end", C.c_double),
                ("max_steps", C.c_int), ("apriori_capture", C.c_int),
                ("r_escape", C.c_double), ("r_capture_factor", C.c_double),
                ("newton_cap", C.c_int)]


class Stats(C.Structure):
    _fields_ = [(n, C.c_int64) for n in ("attempts", "newton", "golden", "rays", "integrated")]


class Grid(C.Structure):
    _fields_ = [("N", C.c_int), ("M", C.c_int), ("max_level", C.c_int),
                ("start_level", C.c_int), ("equafactor", C.c_int),
                ("rays_traced", C.c_int64), ("cel", DP), ("steps", IP),
                ("level", I8P), ("present", U8P)]


def _ptr(a, t):
    return a.ctypes.data_as(t)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "libbtoracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library not built: {path} (run make -C oracle)")
        L = C.CDLL(path)
        L.bto_default_params.argtypes = [C.POINTER(Params)]
        L.bto_pose_init.argtypes = [C.POINTER(Pose)] + [C.c_double] * 7
        L.bto_pose_init_speed.argtypes = [C.POINTER(Pose)] + [C.c_double] * 5
        for n in ("bto_calc_speed",):
            getattr(L, n).argtypes = [C.c_double] * 3
            getattr(L, n).restype = C.c_double
        for n in ("bto_b0", "bto_q0"):
            getattr(L, n).argtypes = [C.c_double] * 2
            getattr(L, n).restype = C.c_double
        L.bto_check_bq.argtypes = [C.c_double, C.c_double, C.c_double, C.c_int, C.POINTER(C.c_int64)]
        L.bto_check_rup.argtypes = [C.c_double] * 5 + [C.POINTER(C.c_int64)]
        L.bto_check_celest.argtypes = [C.c_double] * 6 + [C.c_int, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
        L.bto_derivs.argtypes = [C.c_double, DP, DP, C.c_double, C.c_double]
        L.bto_rkck.argtypes = [C.c_double, DP, DP, C.c_double, DP, DP, C.c_double, C.c_double]
        L.bto_wrap_to_pi.argtypes = [DP, DP]
        L.bto_initial_conditions.argtypes = [C.POINTER(Pose), C.c_double, C.c_double, DP]
        L.bto_trace_ray.argtypes = [C.POINTER(Pose), C.POINTER(Params), DP, DP, IP, C.POINTER(Stats)]
        L.bto_trace_rays.argtypes = [C.POINTER(Pose), C.POINTER(Params), C.c_int, DP, DP, IP,
                                     C.c_int, C.POINTER(Stats)]
        L.bto_grid_dims.argtypes = [C.c_int, C.c_int, IP, IP]
        L.bto_grid_build.argtypes = [C.POINTER(Pose), C.POINTER(Params), C.c_int, C.c_int, C.c_int,
                                     C.c_double, C.c_int, C.c_int, C.POINTER(Grid), C.POINTER(Stats)]
        L.bto_grid_fix_tvertices.argtypes = [C.POINTER(Grid)]
        L.bto_grid_free.argtypes = [C.POINTER(Grid)]
        L.bto_hermite2.argtypes = [C.c_double, DP, DP, DP, DP, DP]
        # Path B
        L.btor_camera_view.argtypes = [C.c_double, C.c_double, C.c_double, C.c_int, C.c_int, FP]
        L.btor_sphere_view.argtypes = [C.c_int, C.c_int, FP]
        L.btor_make_tree.argtypes = [FP, C.c_int, C.c_int, IP, FP, FP]
        L.btor_sum_table.argtypes = [U8P, C.c_int, C.c_int, FP]
        L.btor_make_grid.argtypes = [FP, IP, IP, C.c_int, C.c_int, C.c_int, C.c_int, FP]
        L.btor_frame.argtypes = [FP, C.c_int, C.c_int, C.c_int, FP, C.c_int, C.c_int, FP, IP,
                                 C.c_int, FP, FP, C.c_int, FP, C.c_int, C.c_int, C.c_int, C.c_int,
                                 C.c_int, C.c_float, C.c_float, C.c_float, C.c_int, U8P, FP, IP,
                                 U8P, FP, FP]
        L.btor_blend_border.argtypes = [FP, FP, C.c_int, C.c_int, FP]
        L.btor_frame_blend.argtypes = [FP, FP, C.c_int, C.c_int, C.c_int, C.c_float, FP, C.c_int,
                                       C.c_int, FP, IP, C.c_int, FP, FP, C.c_int, FP, C.c_int,
                                       C.c_int, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float,
                                       C.c_float, C.c_int, U8P, FP, FP]
        L.btor_trails_new.argtypes = [C.c_int, C.c_int, C.c_int, U8P]
        L.btor_trails_new.restype = C.c_void_p
        L.btor_trails_free.argtypes = [C.c_void_p]
        L.btor_frame_trails.argtypes = [C.c_void_p, FP, C.c_int, C.c_int, C.c_int, FP, C.c_int, C.c_int,
                                        FP, IP, C.c_int, FP, FP, C.c_int, FP, C.c_int, C.c_int,
                                        C.c_int, C.c_int, C.c_int, C.c_float, C.c_float, C.c_float,
                                        C.c_int, U8P, FP, FP]
        L.btor_rgb_to_hsp.argtypes = [C.c_float] * 3 + [FP]
        L.btor_hsp_to_rgb.argtypes = [C.c_float] * 3 + [FP]
        for n in ("btor_calc_area4", "btor_calc_area3"):
            getattr(L, n).argtypes = [FP, FP]
            getattr(L, n).restype = C.c_float
        L.btor_redshift.argtypes = [C.c_float, C.c_float, FP]
        L.btor_redshift.restype = C.c_float
        L.btor_star_in_polygon.argtypes = [FP, FP, C.c_float, C.c_float, C.c_int]
        L.btor_search_tree.argtypes = [IP, FP, FP, C.c_int, IP, C.c_int]
        L.btor_smooth_area.argtypes = [FP, U8P, IP, C.c_int, C.c_int]
        L.btor_make_pix.argtypes = [C.c_float] * 3 + [U8P]
        L.btor_gaussian.argtypes = [C.c_float]
        L.btor_gaussian.restype = C.c_float
        L.btor_interpolate_star.argtypes = [FP, FP, C.c_float, C.c_float, C.c_int, FP]
        L.btor_pi_check.argtypes = [FP, C.c_float]
        L.btor_hermite.argtypes = [C.c_float, FP, FP]
        L.btor_intersection.argtypes = [FP, FP]
        L.btor_wrap_to_pi.argtypes = [FP, FP]
        _LIB = L
    return _LIB


def initial_conditions_batch(ps, theta, phi):
    """(pR, pTheta, b, q) of every ray, (n, 4) float64 (Grid.h:524-547)."""
    th = np.ascontiguousarray(theta, dtype=np.float64)
    ph = np.ascontiguousarray(phi, dtype=np.float64)
    out = np.zeros((th.size, 4))
    for k in range(th.size):
        lib().bto_initial_conditions(C.byref(ps), float(th[k]), float(ph[k]), _ptr(out[k], DP))
    return out


def ref_trace_batch(a, r, theta_cam, phi_cam, ic, nthreads=1):
    """The reference's capture test + rkckIntegrate1 (its own compiled code)
    over rays with initial conditions ic (n, 4): returns (theta', phi', step)."""
    R = load_ref()
    ic = np.ascontiguousarray(ic, dtype=np.float64)
    n = ic.shape[0]
    cols = [np.ascontiguousarray(ic[:, k]) for k in range(4)]
    th, ph = np.zeros(n), np.zeros(n)
    st = np.zeros(n, dtype=np.int32)
    R.btref_trace_batch(a, r, theta_cam, phi_cam, n, _ptr(cols[0], DP), _ptr(cols[1], DP), _ptr(cols[2], DP),
                        _ptr(cols[3], DP), _ptr(th, DP), _ptr(ph, DP), _ptr(st, IP), int(nthreads))
    return th, ph, st


def load_ref():
    """The reference's own Metric.h/Camera.h/BlackHole.h, or None."""
    global _REF
    if _REF is None:
        path = os.path.join(HERE, "_ref", "libbtref.so")
        if not os.path.exists(path):
            return None
        R = C.CDLL(path)
        R.btref_pose.argtypes = [C.c_double] * 7 + [DP]
        R.btref_check_celest.argtypes = [C.c_double] * 6
        R.btref_check_bq.argtypes = [C.c_double] * 3
        R.btref_check_rup.argtypes = [C.c_double] * 5
        for n in ("btref_b0", "btref_q0"):
            getattr(R, n).argtypes = [C.c_double] * 2
            getattr(R, n).restype = C.c_double
        R.btref_calc_speed.argtypes = [C.c_double] * 3
        R.btref_calc_speed.restype = C.c_double
        R.btref_derivs.argtypes = [C.c_double, DP, DP, C.c_double, C.c_double]
        R.btref_rkck.argtypes = [C.c_double, DP, DP, C.c_double, DP, DP, C.c_double, C.c_double]
        R.btref_integrate.argtypes = [C.c_double] * 8 + [DP, DP, IP]
        R.btref_wrap_to_pi.argtypes = [DP, DP]
        R.btref_wrap_to_pi_f.argtypes = [FP, FP]
        R.btref_trace_batch.argtypes = [C.c_double] * 4 + [C.c_int, DP, DP, DP, DP, DP, DP, IP, C.c_int]
        _REF = R
    return _REF


def params(eps=1e-5, h1=0.01, z_end=-1e7, max_steps=1000, apriori_capture=1,
           r_escape=0.0, r_capture_factor=0.0, newton_cap=100000):
    p = Params()
    p.eps, p.h1, p.z_end, p.max_steps = eps, h1, z_end, max_steps
    p.apriori_capture, p.r_escape, p.r_capture_factor = apriori_capture, r_escape, r_capture_factor
    p.newton_cap = newton_cap
    return p


def radius_stop_params():
    """The radius-stop variant: rel tol 1e-8, stop r>1000M or r<1.01 r+."""
    return params(eps=1e-8, max_steps=10000, apriori_capture=0, r_escape=1000.0,
                  r_capture_factor=1.01)


def pose(a, theta, phi, r, br=0.0, btheta=0.0, bphi=1.0, speed=None):
    p = Pose()
    if speed is None:
        lib().bto_pose_init(C.byref(p), a, theta, phi, r, br, btheta, bphi)
    else:
        lib().bto_pose_init_speed(C.byref(p), a, theta, phi, r, speed)
    return p


def trace_rays(ps, prm, theta, phi, nthreads=0):
    """integration_wrapper on copies of theta/phi; returns (theta', phi', step, stats)."""
    th = np.ascontiguousarray(theta, dtype=np.float64).copy()
    ph = np.ascontiguousarray(phi, dtype=np.float64).copy()
    st = np.zeros(th.shape, dtype=np.int32)
    s = Stats()
    lib().bto_trace_rays(C.byref(ps), C.byref(prm), th.size, _ptr(th, DP), _ptr(ph, DP),
                         _ptr(st, IP), nthreads, C.byref(s))
    return th, ph, st, s


def set_trig_ulp(k):
    """Test knob: move the integrator's sin/cos one ulp up (k = 1) or down
    (k = -1); 0 turns it off."""
    lib().bto_set_trig_ulp(int(k))


def set_pow_ulp(k):
    """Test knob: move the step-size control's pow one ulp up (k = 1) or down
    (k = -1); 0 turns it off."""
    lib().bto_set_pow_ulp(int(k))


# The one-ulp perturbations of the libm calls on the integration path:
# sin/cos of the equations of motion and pow of the step-size control, each
# one ulp up and one down.
_ULP_KNOBS = [(set_trig_ulp, 1), (set_trig_ulp, -1), (set_pow_ulp, 1), (set_pow_ulp, -1)]


def ulp_sensitive(ps, prm, theta, phi, tol=1e-10):
    """Rays whose traced result moves by more than tol rad, or whose step
    count changes, when the integrator's sin/cos or its step control's pow
    move by one ulp either way: the rays a last-ulp difference between two
    libms can move."""
    t0, p0, s0, _ = trace_rays(ps, prm, theta, phi)
    out = np.zeros(len(theta), bool)
    for knob, k in _ULP_KNOBS:
        knob(k)
        try:
            t1, p1, s1, _ = trace_rays(ps, prm, theta, phi)
        finally:
            knob(0)
        dp = np.abs(p1 - p0)
        dp = np.minimum(dp, 2 * np.pi - dp)
        out |= (np.abs(t1 - t0) > tol) | (dp > tol) | (s1 != s0)
    return out


def ulp_sensitive_grid(ps, prm, max_level, start_level, angle=True, tol=1e-10):
    """(N, M) mask of the grid points whose traced value moves by more than
    tol rad, or whose step count changes, under the same one-ulp
    perturbations as ulp_sensitive."""
    g0 = OracleGrid(ps, prm, max_level, start_level, angle)
    out = np.zeros(g0.steps.shape, bool)
    for knob, k in _ULP_KNOBS:
        knob(k)
        try:
            g1 = OracleGrid(ps, prm, max_level, start_level, angle)
        finally:
            knob(0)
        d = np.abs(g1.cel - g0.cel).max(-1)
        d = np.minimum(d, 2 * np.pi - d)
        out |= (d > tol) | (g1.steps != g0.steps)
    return out


def initial_conditions(ps, theta, phi):
    out = np.zeros(4)
    lib().bto_initial_conditions(C.byref(ps), theta, phi, _ptr(out, DP))
    return out


def grid_dims(max_level, angle=True):
    n, m = C.c_int(), C.c_int()
    lib().bto_grid_dims(max_level, 1 if angle else 0, C.byref(n), C.byref(m))
    return n.value, m.value


class OracleGrid:
    """Dense view of Grid(maxLevel, startLevel, angle, cam, bh)."""

    def __init__(self, ps, prm, max_level, start_level, angle=True, thresh=0.015,
                 fix_tvertices=True, nthreads=0):
        self.g = Grid()
        self.stats = Stats()
        lib().bto_grid_build(C.byref(ps), C.byref(prm), max_level, start_level,
                             1 if angle else 0, thresh, 1 if fix_tvertices else 0,
                             nthreads, C.byref(self.g), C.byref(self.stats))
        N, M = self.g.N, self.g.M
        self.N, self.M = N, M
        self.cel = np.ctypeslib.as_array(self.g.cel, shape=(N, M, 2)).copy()
        self.steps = np.ctypeslib.as_array(self.g.steps, shape=(N, M)).copy()
        self.level = np.ctypeslib.as_array(self.g.level, shape=(N, M)).copy()
        self.present = np.ctypeslib.as_array(self.g.present, shape=(N, M)).copy()
        self.rays_traced = self.g.rays_traced

    def fix_tvertices(self):
        lib().bto_grid_fix_tvertices(C.byref(self.g))
        N, M = self.N, self.M
        self.cel = np.ctypeslib.as_array(self.g.cel, shape=(N, M, 2)).copy()

    def __del__(self):
        try:
            lib().bto_grid_free(C.byref(self.g))
        except Exception:
            pass


# ---------------------------------------------------------------- Path B


def f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def camera_view(view_angle, off_up, off_right, w, h):
    """Viewer(viewangle, offUp, offRight, w, h, false).viewMatrix as (h+1, w+1, 2)."""
    v = np.zeros((h + 1, w + 1, 2), dtype=np.float32)
    lib().btor_camera_view(view_angle, off_up, off_right, w, h, _ptr(v, FP))
    return v


def sphere_view(w, h):
    v = np.zeros((h + 1, w + 1, 2), dtype=np.float32)
    lib().btor_sphere_view(w, h, _ptr(v, FP))
    return v


def make_tree(stars4, tree_level):
    """StarProcessor::makeTree: returns (tree, pos (S,2), mag (S,2))."""
    s4 = f32(stars4)
    S = s4.shape[0]
    tree = np.zeros((1 << (tree_level + 1)) - 1, dtype=np.int32)
    pos = np.zeros((S, 2), dtype=np.float32)
    mag = np.zeros((S, 2), dtype=np.float32)
    lib().btor_make_tree(_ptr(s4, FP), S, tree_level, _ptr(tree, IP), _ptr(pos, FP), _ptr(mag, FP))
    return tree, pos, mag


def sky_low_stars(sky_bgr, stars4):
    """addLowLightStarsToImage of every star in stars4 (the caller applies the
    magnitude cut): a painted copy of the BGR8 sky."""
    sky = np.ascontiguousarray(sky_bgr, dtype=np.uint8).copy()
    s4 = f32(stars4).reshape(-1, 4)
    lib().btor_sky_low_stars(_ptr(sky, U8P), sky.shape[0], sky.shape[1], _ptr(s4, FP), s4.shape[0])
    return sky


def sum_table(sky_bgr):
    sky = np.ascontiguousarray(sky_bgr, dtype=np.uint8)
    rows, cols = sky.shape[:2]
    t = np.zeros((rows, cols, 4), dtype=np.float32)
    lib().btor_sum_table(_ptr(sky, U8P), rows, cols, _ptr(t, FP))
    return t


def make_grid(hash_table, offset_table, hash_pos_tag, hw, ow, GN, GM):
    sym = GM / GN > 3
    GN1 = 2 * GN - 1 if sym else GN
    grid = np.zeros((GN1, GM, 2), dtype=np.float32)
    lib().btor_make_grid(_ptr(f32(hash_table), FP), _ptr(np.ascontiguousarray(offset_table, np.int32), IP),
                         _ptr(np.ascontiguousarray(hash_pos_tag, np.int32), IP), hw, ow, GN, GM,
                         _ptr(grid, FP))
    return grid


def frame(grid, gridlvl, view, cam7, tree, tree_level, star_pos, star_mag, table, star=True,
          map=True, lr=True, hor=0.0, ver=0.0, offset=0.0, step=1, stages=False):
    """One G == 1 frame; returns the BGRA image (h, w, 4) [and intermediates]."""
    grid = f32(grid)
    view = f32(view)
    h, w = view.shape[0] - 1, view.shape[1] - 1
    GN, GM = grid.shape[:2]
    out = np.zeros((h, w, 4), dtype=np.uint8)
    thphi = np.zeros((h + 1, w + 1, 2), dtype=np.float32)
    gap = np.zeros((h + 1, w + 1), dtype=np.int32)
    bh = np.zeros((h, w), dtype=np.uint8)
    area = np.zeros((h, w), dtype=np.float32)
    starl = np.zeros((h, w, 3), dtype=np.float32)
    pos, mag = f32(star_pos), f32(star_mag)
    table = f32(table)
    lib().btor_frame(_ptr(grid, FP), GN, GM, gridlvl, _ptr(view, FP), w, h, _ptr(f32(cam7), FP),
                     _ptr(np.ascontiguousarray(tree, np.int32), IP), tree_level, _ptr(pos, FP),
                     _ptr(mag, FP), pos.shape[0], _ptr(table, FP), table.shape[0], table.shape[1],
                     int(star), int(map), int(lr), hor, ver, offset, step, _ptr(out, U8P),
                     _ptr(thphi, FP), _ptr(gap, IP), _ptr(bh, U8P), _ptr(area, FP), _ptr(starl, FP))
    if stages:
        return out, dict(thphi=thphi, gap=gap, bh=bh, area=area, star=starl)
    return out


def blend_border(grid_a, grid_b):
    """findBhCenter + findBhBorders + smoothBorder x 4 of two dense grids:
    (2 + 2000, 2) float32 (MakeImageKernel.cu:391-456)."""
    ga, gb = f32(grid_a), f32(grid_b)
    assert ga.shape == gb.shape
    GN, GM = ga.shape[:2]
    border = np.zeros((2 + 2000, 2), dtype=np.float32)
    lib().btor_blend_border(_ptr(ga, FP), _ptr(gb, FP), GN, GM, _ptr(border, FP))
    return border


def frame_blend(grid_a, grid_b, alpha, gridlvl, view, cam14, tree, tree_level, star_pos, star_mag,
                table, star=True, map=True, lr=True, hor=0.0, ver=0.0, offset=0.0, step=1,
                stages=False):
    """One G > 1 frame between grids A and B at alpha (MakeImageKernel.cu:386-510)."""
    ga, gb = f32(grid_a), f32(grid_b)
    assert ga.shape == gb.shape
    view = f32(view)
    h, w = view.shape[0] - 1, view.shape[1] - 1
    GN, GM = ga.shape[:2]
    out = np.zeros((h, w, 4), dtype=np.uint8)
    thphi = np.zeros((h + 1, w + 1, 2), dtype=np.float32)
    border = np.zeros((2 + 2000, 2), dtype=np.float32)
    pos, mag = f32(star_pos), f32(star_mag)
    table = f32(table)
    lib().btor_frame_blend(_ptr(ga, FP), _ptr(gb, FP), GN, GM, gridlvl, float(alpha), _ptr(view, FP),
                           w, h, _ptr(f32(cam14), FP),
                           _ptr(np.ascontiguousarray(tree, np.int32), IP), tree_level, _ptr(pos, FP),
                           _ptr(mag, FP), pos.shape[0], _ptr(table, FP), table.shape[0],
                           table.shape[1], int(star), int(map), int(lr), hor, ver, offset, step,
                           _ptr(out, U8P), _ptr(thphi, FP), _ptr(border, FP))
    if stages:
        return out, dict(thphi=thphi, border=border)
    return out


class Trails:
    """Oracle state of a frame sequence with star trails (and diffraction
    when a (w // 16, w // 16, 3) uint8 sprite is given)."""

    def __init__(self, n_stars, w, h, sprite=None):
        self.w, self.h, self.S = w, h, n_stars
        sp = None
        if sprite is not None:
            self._sprite = np.ascontiguousarray(sprite, dtype=np.uint8)
            sp = _ptr(self._sprite, U8P)
        self.handle = lib().btor_trails_new(n_stars, w, h, sp)

    def frame(self, grid, gridlvl, view, cam7, tree, tree_level, star_pos, star_mag, table,
              star=True, map=True, lr=True, hor=0.0, ver=0.0, offset=0.0, step=1):
        """The next frame: (BGRA (h, w, 4), summed star light (h, w, 3), trails (h, w, 3))."""
        grid = f32(grid)
        view = f32(view)
        h, w = view.shape[0] - 1, view.shape[1] - 1
        assert (w, h) == (self.w, self.h)
        GN, GM = grid.shape[:2]
        out = np.zeros((h, w, 4), dtype=np.uint8)
        light = np.zeros((h, w, 3), dtype=np.float32)
        trail = np.zeros((h, w, 3), dtype=np.float32)
        pos, mag = f32(star_pos), f32(star_mag)
        table = f32(table)
        rc = lib().btor_frame_trails(self.handle, _ptr(grid, FP), GN, GM, gridlvl, _ptr(view, FP), w, h,
                                     _ptr(f32(cam7), FP), _ptr(np.ascontiguousarray(tree, np.int32), IP),
                                     tree_level, _ptr(pos, FP), _ptr(mag, FP), pos.shape[0],
                                     _ptr(table, FP), table.shape[0], table.shape[1], int(star),
                                     int(map), int(lr), hor, ver, offset, step, _ptr(out, U8P),
                                     _ptr(light, FP), _ptr(trail, FP))
        assert rc == 0
        return out, light, trail

    def close(self):
        if getattr(self, "handle", None):
            lib().btor_trails_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
