"""ctypes binding of oracle/libgeon_oracle.so (test infrastructure only)."""
import ctypes as C
import os
import subprocess

import numpy as np

from physsim_amd import _abi as A

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgeon_oracle.so")
_LIB = None
_DP = C.POINTER(C.c_double)
_IP = C.POINTER(C.c_int32)
FAR_STRIDE = 22


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = [C.POINTER(A.GeonParams), C.c_int32]
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_set_params.argtypes = [C.c_void_p, C.POINTER(A.GeonParams)]
        L.orc_upload.argtypes = [C.c_void_p, C.c_int32] + [_DP] * 7
        L.orc_count.argtypes = [C.c_void_p]
        L.orc_count.restype = C.c_int32
        L.orc_download.argtypes = [C.c_void_p] + [_DP] * 6 + [_IP]
        L.orc_compute_forces.argtypes = [C.c_void_p]
        L.orc_get_forces.argtypes = [C.c_void_p] + [_DP] * 11
        L.orc_update.argtypes = [C.c_void_p, C.c_double, _IP]
        L.orc_stats.argtypes = [C.c_void_p, C.POINTER(A.GeonStats)]
        L.orc_build_tree.argtypes = [C.c_void_p]
        L.orc_build_tree.restype = C.c_int64
        L.orc_get_tree.argtypes = [C.c_void_p, C.POINTER(A.GeonTreeNode), C.c_int64]
        L.orc_get_tree.restype = C.c_int64
        L.orc_counters.argtypes = [C.c_void_p, C.POINTER(C.c_int64)]
        L.orc_collision_pairs.argtypes = [C.c_void_p, _IP, _IP, C.c_int64]
        L.orc_collision_pairs.restype = C.c_int64
        L.orc_direct_forces.argtypes = [C.c_void_p, _IP, C.c_int32, _DP, _DP]
        L.orc_set_threads.argtypes = [C.c_int32]
        L.orc_far_samples.argtypes = [C.c_void_p, _IP, C.c_int32, _DP, C.c_int64]
        L.orc_far_samples.restype = C.c_int64
        L.orc_far_set_model.argtypes = [C.c_void_p, C.POINTER(C.c_float)]
        L.orc_far_training_samples.argtypes = [C.c_void_p, _IP, C.c_int32, C.c_double, C.c_double,
                                               C.POINTER(C.c_float), _IP, C.c_int64]
        L.orc_far_training_samples.restype = C.c_int64
        L.orc_far_loss_grad.argtypes = [C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_int64, C.c_double, _DP]
        L.orc_far_loss_grad.restype = C.c_double
        L.orc_get_history.argtypes = [C.c_void_p, C.c_int32, C.c_int32] + [_DP] * 6
        L.orc_get_history.restype = C.c_int32
        L.orc_set_history.argtypes = [C.c_void_p, C.c_int32, C.c_int32] + [_DP] * 6
        L.orc_set_time.argtypes = [C.c_void_p, C.c_double]
        L.orc_get_f1pn.argtypes = [C.c_void_p, _DP, _DP]
        L.orc_min_image.argtypes = [C.c_void_p] + [C.c_double] * 4 + [_DP]
        L.orc_dead_count.argtypes = [C.c_void_p]
        L.orc_dead_count.restype = C.c_int32
        L.orc_set_lifetime.argtypes = [C.c_void_p, C.c_int32, C.c_double, C.c_double]
        L.orc_delayed_state.argtypes = [C.c_void_p, C.c_int32, C.c_double, C.c_double, _DP]
        L.orc_delayed_state.restype = C.c_int
        L.orc_collision_pairs_ordered.argtypes = [C.c_void_p, _IP, _IP, C.c_int64]
        L.orc_collision_pairs_ordered.restype = C.c_int64
        L.orc_collision_events.argtypes = [C.c_void_p, _DP, C.c_int64]
        L.orc_collision_events.restype = C.c_int64
        L.orc_set_antimatter.argtypes = [C.c_void_p, _IP]
        L.orc_download_ex.argtypes = [C.c_void_p, _DP, _DP, _IP, _DP]
        L.orc_leaf_points.argtypes = [C.c_void_p, C.POINTER(C.c_uint64), _IP, C.c_int64]
        L.orc_leaf_points.restype = C.c_int64
        L.orc_get_jerk.argtypes = [C.c_void_p, _DP, _DP]
        _LIB = L
    return _LIB


def _d(a):
    return a.ctypes.data_as(_DP) if a is not None else None


def far_loss_grad(model, samples, wscale=1.0):
    """FP64 loss and gradient of the learned far-field network on samples."""
    m = np.ascontiguousarray(model, dtype=np.float32)
    sm = np.ascontiguousarray(samples, dtype=np.float32)
    g = np.zeros(A.FAR_PARAMS)
    L = lib().orc_far_loss_grad(m.ctypes.data_as(C.POINTER(C.c_float)), sm.ctypes.data_as(C.POINTER(C.c_float)),
                                len(sm), float(wscale), _d(g))
    return L, g


def default_params(**kw):
    """Reference defaults (src/config.js) without needing the HIP library."""
    p = A.GeonParams()
    p.toggles0 = A.GRAVITY | A.COULOMB | A.BARNES_HUT
    p.boundary_mode = A.BOUND_DESPAWN
    p.topology = A.TORUS
    p.collision_mode = A.COL_PASS
    p.periodic_mode = A.PERIODIC_MIN_IMAGE
    p.max_substeps = 32
    p.domain_w = p.domain_h = 64.0
    p.softening, p.bh_softening, p.theta = 8.0, 4.0, 0.5
    p.yukawa_coupling, p.yukawa_mu = 14.0, 0.15
    p.higgs_mass, p.axion_mass, p.hubble = 0.5, 0.05, 0.001
    p.ext_gravity_angle = np.pi / 2
    p.field_grid = 128
    p.kahan = 1
    p.bounce_friction = 0.4
    for k, v in kw.items():
        setattr(p, k, v)
    return p


class Oracle:
    def __init__(self, params, capacity=1024):
        self.L = lib()
        self.params = params
        self.s = self.L.orc_create(C.byref(params), int(capacity))

    def __del__(self):
        if getattr(self, "s", None):
            self.L.orc_destroy(self.s)
            self.s = None

    def set_params(self, p):
        self.params = p
        self.L.orc_set_params(self.s, C.byref(p))

    def upload(self, x, y, wx, wy, mass, charge, angw=None):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, wx, wy, mass, charge)]
        ang = np.ascontiguousarray(angw if angw is not None else np.zeros(len(x)), dtype=np.float64)
        self.L.orc_upload(self.s, len(arrs[0]), *[_d(a) for a in arrs], _d(ang))

    def count(self):
        return self.L.orc_count(self.s)

    def download(self):
        n = self.count()
        out = {k: np.zeros(n) for k in ("x", "y", "wx", "wy", "mass", "angw")}
        ids = np.zeros(n, dtype=np.int32)
        self.L.orc_download(self.s, *[_d(out[k]) for k in ("x", "y", "wx", "wy", "mass", "angw")],
                            ids.ctypes.data_as(_IP))
        out["id"] = ids
        return out

    def compute_forces(self):
        self.L.orc_compute_forces(self.s)

    def forces(self):
        n = self.count()
        names = ("fx", "fy", "bz", "bgz", "dbzdx", "dbzdy", "dbgzdx", "dbgzdy", "frame_drag", "tidal")
        out = {k: np.zeros(n) for k in names}
        pe = np.zeros(1)
        self.L.orc_get_forces(self.s, *[_d(out[k]) for k in names], _d(pe))
        out["pe_total"] = float(pe[0])
        return out

    def update(self, dt):
        n = C.c_int32()
        self.L.orc_update(self.s, float(dt), C.byref(n))
        return n.value

    def stats(self):
        st = A.GeonStats()
        self.L.orc_stats(self.s, C.byref(st))
        return {k: getattr(st, k) for k, _ in A.GeonStats._fields_}

    def tree(self):
        nn = self.L.orc_build_tree(self.s)
        arr = (A.GeonTreeNode * max(nn, 1))()
        self.L.orc_get_tree(self.s, arr, nn)
        return np.ctypeslib.as_array(arr)[:nn].copy()

    def counters(self):
        c = (C.c_int64 * 3)()
        self.L.orc_counters(self.s, c)
        return {"pair": c[0], "node": c[1], "visit": c[2]}

    def collision_pairs(self):
        n = self.L.orc_collision_pairs(self.s, None, None, 0)
        a = np.zeros(max(n, 1), dtype=np.int32)
        b = np.zeros(max(n, 1), dtype=np.int32)
        self.L.orc_collision_pairs(self.s, a.ctypes.data_as(_IP), b.ctypes.data_as(_IP), n)
        return a[:n], b[:n]

    def collision_pairs_ordered(self):
        n = self.L.orc_collision_pairs_ordered(self.s, None, None, 0)
        a = np.zeros(max(n, 1), dtype=np.int32)
        b = np.zeros(max(n, 1), dtype=np.int32)
        self.L.orc_collision_pairs_ordered(self.s, a.ctypes.data_as(_IP), b.ctypes.data_as(_IP), n)
        return a[:n], b[:n]

    def collision_events(self):
        """Events of the last update: rows {type (1 merge, 2 annihilation), x, y, energy, px, py}."""
        n = self.L.orc_collision_events(self.s, None, 0)
        out = np.zeros((max(n, 1), 6))
        self.L.orc_collision_events(self.s, _d(out), n)
        return out[:n]

    def set_antimatter(self, flags):
        f = np.ascontiguousarray(flags, dtype=np.int32)
        self.L.orc_set_antimatter(self.s, f.ctypes.data_as(_IP))

    def leaf_points(self):
        """(key, id) rows of every leaf's stored point order; key = level << 56 | path."""
        m = self.L.orc_leaf_points(self.s, None, None, 0)
        key = np.zeros(max(m, 1), dtype=np.uint64)
        ids = np.zeros(max(m, 1), dtype=np.int32)
        self.L.orc_leaf_points(self.s, key.ctypes.data_as(C.POINTER(C.c_uint64)), ids.ctypes.data_as(_IP), m)
        return key[:m], ids[:m]

    def jerk(self):
        n = self.count()
        jx, jy = np.zeros(n), np.zeros(n)
        self.L.orc_get_jerk(self.s, _d(jx), _d(jy))
        return jx, jy

    def download_ex(self):
        n = self.count()
        out = {k: np.zeros(n) for k in ("base_mass", "charge", "radius")}
        anti = np.zeros(n, dtype=np.int32)
        self.L.orc_download_ex(self.s, _d(out["base_mass"]), _d(out["charge"]), anti.ctypes.data_as(_IP),
                               _d(out["radius"]))
        out["antimatter"] = anti
        return out

    def far_samples(self, idx):
        """Accepted far nodes of the observers in idx with their exact residuals
        (one row of FAR_STRIDE values per sample; see orc_far_samples)."""
        ii = np.ascontiguousarray(idx, dtype=np.int32)
        m = self.L.orc_far_samples(self.s, ii.ctypes.data_as(_IP), len(ii), None, 0)
        out = np.zeros((max(m, 1), FAR_STRIDE))
        self.L.orc_far_samples(self.s, ii.ctypes.data_as(_IP), len(ii), _d(out), m)
        return out[:m]

    def far_training_samples(self, idx, f2ref=1.0, pe2ref=1.0):
        """Samples of the observers in idx in the 20-float training layout, and
        the observer of each."""
        ii = np.ascontiguousarray(idx, dtype=np.int32)
        m = self.far_samples(ii).shape[0]
        out = np.zeros((max(m, 1), A.FAR_SAMPLE), dtype=np.float32)
        obs = np.zeros(max(m, 1), dtype=np.int32)
        self.L.orc_far_training_samples(self.s, ii.ctypes.data_as(_IP), len(ii), float(f2ref), float(pe2ref),
                                        out.ctypes.data_as(C.POINTER(C.c_float)), obs.ctypes.data_as(_IP), m)
        return out[:m], obs[:m]

    def far_set_model(self, model):
        m = np.ascontiguousarray(model, dtype=np.float32)
        self.L.orc_far_set_model(self.s, m.ctypes.data_as(C.POINTER(C.c_float)))

    def direct_forces(self, idx=None, threads=None):
        if threads:
            self.L.orc_set_threads(int(threads))
        n = self.count() if idx is None else len(idx)
        fx = np.zeros(n)
        fy = np.zeros(n)
        if idx is None:
            self.L.orc_direct_forces(self.s, None, 0, _d(fx), _d(fy))
        else:
            ii = np.ascontiguousarray(idx, dtype=np.int32)
            self.L.orc_direct_forces(self.s, ii.ctypes.data_as(_IP), len(ii), _d(fx), _d(fy))
        return fx, fy


    # ---- signal delay ----
    HIST_KEYS = ("x", "y", "vx", "vy", "angw", "t")

    def history(self, i, dead=False):
        """Particle i's history (or the i-th dead record's), oldest first."""
        out = {k: np.zeros(A.HISTORY_SIZE) for k in self.HIST_KEYS}
        n = self.L.orc_get_history(self.s, int(i), int(dead), *[_d(out[k]) for k in self.HIST_KEYS])
        return {k: v[:n] for k, v in out.items()}

    def set_history(self, i, h):
        n = len(h["t"])
        arrs = [np.ascontiguousarray(h[k], dtype=np.float64) for k in self.HIST_KEYS]
        self.L.orc_set_history(self.s, int(i), n, *[_d(a) for a in arrs])

    def min_image(self, ox, oy, sx, sy):
        out = np.zeros(2)
        self.L.orc_min_image(self.s, float(ox), float(oy), float(sx), float(sy), _d(out))
        return out

    def f1pn(self):
        n = self.count()
        fx, fy = np.zeros(n), np.zeros(n)
        self.L.orc_get_f1pn(self.s, _d(fx), _d(fy))
        return fx, fy

    def set_time(self, t):
        self.L.orc_set_time(self.s, float(t))

    def dead_count(self):
        return self.L.orc_dead_count(self.s)

    def set_lifetime(self, i, creation=-np.inf, death=np.inf):
        self.L.orc_set_lifetime(self.s, int(i), float(creation), float(death))

    def delayed_state(self, i, ox, oy):
        """getDelayedState of source i seen from (ox, oy) now: (x, y, vx, vy, angw) or None."""
        out = np.zeros(5)
        ok = self.L.orc_delayed_state(self.s, int(i), float(ox), float(oy), _d(out))
        return out if ok == 1 else None


def set_threads(n):
    lib().orc_set_threads(int(n))


def _field_protos(L):
    if getattr(L, "_field_bound", False):
        return
    L.orc_field_set.argtypes = [C.c_void_p, C.c_int32, _DP, _DP]
    L.orc_field_get.argtypes = [C.c_void_p, C.c_int32, C.c_int32, _DP]
    L.orc_field_get.restype = C.c_int
    L.orc_field_selfgrav.argtypes = [C.c_void_p, C.c_int32]
    L.orc_field_selfgrav.restype = C.c_int
    L.orc_pqs_deposit.argtypes = [C.c_int32, C.c_double, C.c_double, C.c_int32, C.c_int32, _DP, _DP, _DP, _DP]
    L.orc_fft2d.argtypes = [_DP, _DP, C.c_int32, C.c_int32]
    L._field_bound = True


FQ = {"field": 0, "field_dot": 1, "grad_x": 2, "grad_y": 3, "sg_phi": 4, "sg_gx": 5, "sg_gy": 6,
      "source": 7, "thermal": 8, "rho": 9}


def fft2d(re, im, inverse=False):
    L = lib()
    _field_protos(L)
    re = np.ascontiguousarray(re, dtype=np.float64).copy()
    im = np.ascontiguousarray(im, dtype=np.float64).copy()
    L.orc_fft2d(_d(re), _d(im), re.shape[0], int(inverse))
    return re, im


def pqs_deposit(G, W, H, bc, x, y, v):
    L = lib()
    _field_protos(L)
    out = np.zeros((G, G))
    xs, ys, vs = (np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, v))
    L.orc_pqs_deposit(G, W, H, bc, len(xs), _d(xs), _d(ys), _d(vs), _d(out))
    return out


def _oracle_field_methods():
    def field_set(self, which, field=None, field_dot=None):
        _field_protos(self.L)
        f = np.ascontiguousarray(field, dtype=np.float64).ravel() if field is not None else None
        fd = np.ascontiguousarray(field_dot, dtype=np.float64).ravel() if field_dot is not None else None
        self.L.orc_field_set(self.s, which, _d(f), _d(fd))

    def field_get(self, which, what):
        _field_protos(self.L)
        G = self.params.field_grid
        out = np.zeros(G * G)
        self.L.orc_field_get(self.s, which, FQ[what], _d(out))
        return out.reshape(G, G)

    def field_selfgrav(self, which):
        _field_protos(self.L)
        return self.L.orc_field_selfgrav(self.s, which)

    Oracle.field_set = field_set
    Oracle.field_get = field_get
    Oracle.field_selfgrav = field_selfgrav


_oracle_field_methods()
