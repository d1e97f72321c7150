"""Python host side of the gfx950 Geon backend.

`HipPhysics` mirrors the method surface the reference app calls on its WebGPU
backend (`GPUPhysics`, src/gpu/gpu-physics.js) so a host written against that
contract can drive this backend; every call goes through the C ABI of
libgeon_hip.so (include/geon.h).  There is no CPU fallback: if the HIP library
is missing or no GPU is present, construction raises.
"""
import ctypes as C
import json
import os

import numpy as np

from . import _abi as A
from . import state_v1

_LIB = None
LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libgeon_hip.so")


def load_library(path=None):
    """Load libgeon_hip.so (built in-tree by physsim_amd/build.py)."""
    global _LIB
    if _LIB is None or path is not None:
        p = path or LIB_PATH
        if not os.path.exists(p):
            raise RuntimeError("libgeon_hip.so not built: run __graft_entry__.build() (missing %s)" % p)
        _LIB = A.bind(C.CDLL(p))
        if _LIB.geon_abi_version() != A.ABI_VERSION:
            raise RuntimeError("libgeon_hip.so ABI mismatch")
    return _LIB


class GeonError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("geon error %d: %s" % (code, msg))
        self.code = code


def default_params(**kw):
    lib = load_library()
    p = A.GeonParams()
    lib.geon_default_params(C.byref(p))
    for k, v in kw.items():
        if not hasattr(p, k):
            raise KeyError(k)
        setattr(p, k, v)
    return p


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double)) if a is not None else None


def _check_lengths(n, arrays, names):
    """Every given column of an upload holds exactly n entries (the C ABI reads n from each)."""
    for a, name in zip(arrays, names):
        if a is not None and len(a) != n:
            raise ValueError("upload column %s has %d entries, x has %d" % (name, len(a), n))


class HipPhysics:
    """Third physics backend beside the reference's CPUPhysics and GPUPhysics."""

    def __init__(self, params=None, max_particles=1 << 16):
        self.lib = load_library()
        self.params = params if params is not None else default_params()
        h = C.c_void_p()
        rc = self.lib.geon_create(C.byref(self.params), int(max_particles), C.byref(h))
        if rc != 0:
            raise GeonError(rc, self.lib.geon_last_error(None).decode())
        self.h = h
        self.max_particles = int(max_particles)
        self._stats = None
        self._selected = None

    # ---- lifecycle ----------------------------------------------------
    def destroy(self):                                      # gpu-physics.js:4291
        if self.h:
            self.lib.geon_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    def _check(self, rc):
        if rc < 0:
            raise GeonError(rc, self.lib.geon_last_error(self.h).decode())
        return rc

    def reset(self):                                        # gpu-physics.js:3973
        self._check(self.lib.geon_reset(self.h))

    # ---- configuration ----------------------------------------------------
    def set_params(self, params):
        self._check(self.lib.geon_set_params(self.h, C.byref(params)))
        self.params = params

    def setToggles(self, physics):                          # gpu-physics.js:1969-2056
        """Accepts an object or dict with the reference's *Enabled booleans."""
        get = physics.get if isinstance(physics, dict) else (lambda k, d=None: getattr(physics, k, d))
        t = 0
        for name, bit in A.TOGGLE_BITS.items():
            if get(name, False):
                t |= bit
        p = A.GeonParams.from_buffer_copy(self.params)
        p.toggles0 = t
        if get("yukawaMu") is not None:
            p.yukawa_mu = float(get("yukawaMu"))
        self.set_params(p)

    # ---- particles ----------------------------------------------------
    def addParticle(self, x, y, vx=0.0, vy=0.0, mass=1.0, charge=0.0, angw=0.0, antimatter=False):
        """gpu-physics.js:1857: vx/vy are proper velocity w; returns slot or -1."""
        pi = A.GeonParticleInit(x, y, vx, vy, mass, charge, angw, int(bool(antimatter)))
        rc = self.lib.geon_add_particle(self.h, C.byref(pi))
        if rc == A.EFULL:
            return -1
        return self._check(rc)

    def removeParticle(self, idx, deathMass=0.0, deathAngVel=0.0):   # gpu-physics.js:1936
        self._check(self.lib.geon_remove_particle(self.h, int(idx), float(deathMass), float(deathAngVel)))

    def patchMass(self, idx, mass, base_mass):              # gpu-physics.js:1955
        self._check(self.lib.geon_patch_mass(self.h, int(idx), float(mass), float(base_mass)))

    def upload(self, x, y, wx, wy, mass, charge, angw=None, base_mass=None, antimatter=None):
        arrs = [np.ascontiguousarray(a, dtype=np.float64).ravel() if a is not None else None
                for a in (x, y, wx, wy, mass, base_mass, charge, angw)]
        n = len(arrs[0])
        anti = np.ascontiguousarray(antimatter, dtype=np.int32).ravel() if antimatter is not None else None
        _check_lengths(n, arrs + [anti], ("x", "y", "wx", "wy", "mass", "base_mass", "charge", "angw", "antimatter"))
        v = A.GeonSoAView(*[_dptr(a) for a in arrs], None,
                          anti.ctypes.data_as(C.POINTER(C.c_int32)) if anti is not None else None)
        self._check(self.lib.geon_upload_soa(self.h, C.byref(v), n))
        self._keep = (arrs, anti)

    def download(self):
        ns = C.c_int32()
        cap = self.max_particles
        out = {k: np.zeros(cap) for k in ("x", "y", "wx", "wy", "mass", "base_mass", "charge", "angw")}
        alive = np.zeros(cap, dtype=np.int32)
        anti = np.zeros(cap, dtype=np.int32)
        ip = C.POINTER(C.c_int32)
        v = A.GeonSoAView(*[_dptr(out[k]) for k in ("x", "y", "wx", "wy", "mass", "base_mass", "charge", "angw")],
                          alive.ctypes.data_as(ip), anti.ctypes.data_as(ip))
        self._check(self.lib.geon_download_soa(self.h, C.byref(v), cap, C.byref(ns)))
        n = ns.value
        res = {k: a[:n] for k, a in out.items()}
        res["alive"] = alive[:n].astype(bool)
        res["antimatter"] = anti[:n].astype(bool)
        return res

    def getParticleCount(self):                             # gpu-physics.js:4287
        n = C.c_int32()
        self._check(self.lib.geon_particle_count(self.h, C.byref(n)))
        return n.value

    # ---- stepping --------------------------------------------------------
    def update(self, dt):
        """One Physics.update call: adaptive substeps, boundary once at the end."""
        n = C.c_int32()
        self._check(self.lib.geon_update(self.h, float(dt), C.byref(n)))
        return n.value

    def compute_forces(self):
        self._check(self.lib.geon_compute_forces(self.h))

    def forces(self, n_slots):
        names = ("fx", "fy", "bz", "dbzdx", "dbzdy", "bgz", "dbgzdx", "dbgzdy", "frame_drag_torque",
                 "tidal_torque", "pe", "f1pn_x", "f1pn_y", "jerk_x", "jerk_y")
        out = {k: np.zeros(n_slots) for k in names}
        v = A.GeonForceView(*[_dptr(out[k]) for k in names])
        self._check(self.lib.geon_get_forces(self.h, C.byref(v), int(n_slots)))
        return out

    # ---- stats -----------------------------------------------------------
    def requestStats(self, sel_idx=-1):                     # gpu-physics.js:4146
        """The stats now; sel_idx selects the particle whose forces by type
        the force passes record from then on (-1: none; None keeps the
        current selection)."""
        if sel_idx is not None:
            self._check(self.lib.geon_select(self.h, int(sel_idx)))
        st = A.GeonStats()
        self._check(self.lib.geon_get_stats(self.h, C.byref(st)))
        self._stats = st
        self._selected = self.selected()

    def selected(self):                                     # readStats().selected, gpu-physics.js:4238-4276
        s = A.GeonSelected()
        self._check(self.lib.geon_get_selected(self.h, C.byref(s)))
        if s.slot < 0:
            return None
        out = {"slot": s.slot, "forcesValid": bool(s.forces_valid), "antimatter": bool(s.antimatter),
               "flags": int(s.antimatter)}
        for k, name in (("x", "posX"), ("y", "posY"), ("wx", "velWX"), ("wy", "velWY"), ("mass", "mass"),
                        ("charge", "charge"), ("angw", "angW"), ("base_mass", "baseMass"), ("radius", "radius"),
                        ("vx", "velX"), ("vy", "velY"), ("ang_vel", "angVel"), ("mag_moment", "magMoment"),
                        ("ang_momentum", "angMomentum")):
            out[name] = getattr(s, k)
        for t, name in enumerate(A.SF_NAMES):
            out[name] = {"x": s.force[t][0], "y": s.force[t][1]}
        return out

    def readStats(self):                                    # gpu-physics.js:4215-4281
        st = self._stats
        if st is None:
            return None
        return {
            "linearKE": st.linear_ke, "spinKE": st.spin_ke, "px": st.px, "py": st.py,
            "orbitalAngMom": st.orbital_ang_mom, "spinAngMom": st.spin_ang_mom,
            "comX": st.com_x, "comY": st.com_y, "totalMass": st.total_mass,
            "aliveCount": st.alive_count, "pe": st.pe, "fieldEnergy": st.field_energy,
            "fieldPx": st.field_px, "fieldPy": st.field_py,
            "higgsFieldEnergy": st.higgs_field_energy, "axionFieldEnergy": st.axion_field_energy,
            "pfiEnergy": st.pfi_energy, "simTime": st.sim_time, "deadCount": st.dead_count,
            "totalRadiated": st.radiated_energy, "totalRadiatedPx": st.radiated_px, "totalRadiatedPy": st.radiated_py,
            "selected": self._selected,
        }

    def stats(self):
        self.requestStats(None)
        return self.readStats()

    def history(self, slot):
        """Signal-delay history of `slot`, oldest sample first: dict of x, y,
        vx, vy, angw, t arrays (particle.js:77-85)."""
        keys = ("x", "y", "vx", "vy", "angw", "t")
        out = {k: np.zeros(A.HISTORY_SIZE) for k in keys}
        v = A.GeonHistoryView(*[_dptr(out[k]) for k in keys])
        n = C.c_int32()
        self._check(self.lib.geon_get_history(self.h, int(slot), C.byref(v), C.byref(n)))
        return {k: a[:n.value] for k, a in out.items()}

    # ---- scalar fields ---------------------------------------------------
    def set_field(self, which, field=None, field_dot=None):
        G = self.params.field_grid
        f = np.ascontiguousarray(field, dtype=np.float64).ravel() if field is not None else None
        fd = np.ascontiguousarray(field_dot, dtype=np.float64).ravel() if field_dot is not None else None
        self._check(self.lib.geon_set_field(self.h, int(which), _dptr(f), _dptr(fd), G * G))

    def get_field(self, which, what="field"):              # readbackFieldData (gpu-physics.js:3844)
        G = self.params.field_grid
        out = np.zeros(G * G)
        self._check(self.lib.geon_get_field(self.h, int(which), A.FQ[what], _dptr(out), G * G))
        return out.reshape(G, G)

    def field_selfgravity(self, which):
        self._check(self.lib.geon_field_selfgravity(self.h, int(which)))

    def fft2d(self, re, im, inverse=False):
        re = np.ascontiguousarray(re, dtype=np.float64).copy()
        im = np.ascontiguousarray(im, dtype=np.float64).copy()
        self._check(self.lib.geon_fft2d(self.h, _dptr(re), _dptr(im), re.shape[0], int(bool(inverse))))
        return re, im

    # ---- tree / diagnostics ---------------------------------------------
    def build_tree(self):
        self._check(self.lib.geon_build_tree(self.h))

    def tree(self):
        nn = C.c_int64()
        self._check(self.lib.geon_get_tree(self.h, None, 0, C.byref(nn)))
        arr = (A.GeonTreeNode * max(nn.value, 1))()
        self._check(self.lib.geon_get_tree(self.h, arr, nn.value, C.byref(nn)))
        return np.ctypeslib.as_array(arr)[: nn.value].copy() if nn.value else np.zeros(0)

    def collision_pairs(self):
        """Overlapping pairs (slot a < slot b), sorted (collisions.js:35-60)."""
        n = C.c_int64()
        self._check(self.lib.geon_collision_pairs(self.h, None, None, 0, C.byref(n)))
        a = np.zeros(max(n.value, 1), dtype=np.int32)
        b = np.zeros(max(n.value, 1), dtype=np.int32)
        ip = C.POINTER(C.c_int32)
        self._check(self.lib.geon_collision_pairs(self.h, a.ctypes.data_as(ip), b.ctypes.data_as(ip), n.value,
                                                  C.byref(n)))
        return a[:n.value], b[:n.value]

    def collision_pairs_noalloc(self, cap=None):
        """The same list through the entry point that allocates no device memory
        (at most 2^20 pairs); cap limits the pairs returned."""
        n = C.c_int64()
        if cap is None:
            self._check(self.lib.geon_collision_pairs_noalloc(self.h, None, None, 0, C.byref(n)))
            cap = n.value
        a = np.zeros(max(cap, 1), dtype=np.int32)
        b = np.zeros(max(cap, 1), dtype=np.int32)
        ip = C.POINTER(C.c_int32)
        self._check(self.lib.geon_collision_pairs_noalloc(self.h, a.ctypes.data_as(ip), b.ctypes.data_as(ip), int(cap),
                                                          C.byref(n)))
        m = min(cap, n.value)
        return a[:m], b[:m], n.value

    def collision_events(self):
        """Events of the last update in merge mode: rows {type (1 merge, 2
        annihilation), x, y, energy, px, py}, merges first (collisions.js:62-92)."""
        n = C.c_int64()
        self._check(self.lib.geon_collision_events(self.h, None, 0, C.byref(n)))
        out = np.zeros((max(n.value, 1), 6))
        self._check(self.lib.geon_collision_events(self.h, out.ctypes.data_as(C.POINTER(C.c_double)), n.value,
                                                   C.byref(n)))
        return out[:n.value]

    # ---- learned far-field correction ------------------------------------
    def far_collect(self, group_stride=1):
        """Append training samples from the current state; returns the count held."""
        n = C.c_int64()
        self._check(self.lib.geon_far_collect(self.h, int(group_stride), C.byref(n)))
        self._far_n = n.value
        return n.value

    def far_clear(self):
        self._check(self.lib.geon_far_clear(self.h))
        self._far_n = 0

    def far_samples(self, first=0, n=None, observers=False):
        """Samples (n x 16 float32: 10 features, 4 targets, 2 weights) and
        optionally the observer slot of each."""
        if n is None:
            n = getattr(self, "_far_n", 0) - first
        out = np.zeros((max(n, 1), A.FAR_SAMPLE), dtype=np.float32)
        obs = np.zeros(max(n, 1), dtype=np.int32)
        self._check(self.lib.geon_far_get_samples(self.h, int(first), int(n), out.ctypes.data_as(C.POINTER(C.c_float)),
                                                  obs.ctypes.data_as(C.POINTER(C.c_int32))))
        return (out[:n], obs[:n]) if observers else out[:n]

    def far_train(self, steps, batch=8192, lr=2e-3, lr_final=2e-5, seed=1):
        loss = np.zeros(max(int(steps), 1), dtype=np.float32)
        self._check(self.lib.geon_far_train(self.h, int(steps), int(batch), float(lr), float(lr_final), int(seed),
                                            loss.ctypes.data_as(C.POINTER(C.c_float))))
        return loss[:steps]

    def far_grad(self, rows):
        r = np.ascontiguousarray(rows, dtype=np.int64)
        g = np.zeros(A.FAR_PARAMS, dtype=np.float32)
        L = C.c_float()
        self._check(self.lib.geon_far_grad(self.h, r.ctypes.data_as(C.POINTER(C.c_int64)), len(r),
                                           g.ctypes.data_as(C.POINTER(C.c_float)), C.byref(L)))
        return L.value, g

    def far_get_model(self):
        m = np.zeros(A.FAR_MODEL_FLOATS, dtype=np.float32)
        self._check(self.lib.geon_far_get_model(self.h, m.ctypes.data_as(C.POINTER(C.c_float)), len(m)))
        return m

    def far_set_model(self, model):
        m = np.ascontiguousarray(model, dtype=np.float32)
        self._check(self.lib.geon_far_set_model(self.h, m.ctypes.data_as(C.POINTER(C.c_float)), len(m)))

    def counters(self):
        c = A.GeonCounters()
        self._check(self.lib.geon_counters(self.h, C.byref(c)))
        return {k: getattr(c, k) for k, _ in A.GeonCounters._fields_}

    def timing(self, on=True):
        self._check(self.lib.geon_timing_enable(self.h, int(bool(on))))

    def kernel_timing(self, kernel):
        ms = C.c_double()
        n = C.c_int64()
        self._check(self.lib.geon_kernel_timing(self.h, int(kernel), C.byref(ms), C.byref(n)))
        return ms.value, n.value

    # ---- save/load: the reference's v1 state (save-load.js:89-262) ---------
    def serialize(self, sim=None):
        """GPUPhysics.serialize (gpu-physics.js:3749-3836): the v1 state of the
        live particles in slot order, toggles, parameters and modes.  `sim`
        may carry the app's speedScale and camera (settings.friction is the
        bounce friction parameter)."""
        d = self.download()
        particles = [{"x": float(d["x"][i]), "y": float(d["y"][i]), "wx": float(d["wx"][i]),
                      "wy": float(d["wy"][i]), "mass": float(d["mass"][i]), "baseMass": float(d["base_mass"][i]),
                      "charge": float(d["charge"][i]), "angw": float(d["angw"][i]),
                      "antimatter": bool(d["antimatter"][i])}
                     for i in np.nonzero(d["alive"])[0]]
        return state_v1.params_to_state(self.params, particles, sim)

    def deserialize(self, state):
        """GPUPhysics.deserialize (gpu-physics.js:3886-3960) with loadState's
        restore of toggles, parameters and modes (save-load.js:129-190):
        False for anything but version 1.  Particles are renumbered 0..n-1 and
        taken as always existing (no signal-delay history)."""
        if isinstance(state, (str, bytes)):
            state = json.loads(state)
        if not state or state.get("version") != 1:
            return False
        self.set_params(state_v1.apply_state(self.params, state))
        c = state_v1.particles_from_state(state, self.max_particles)
        arr = lambda k: np.asarray(c[k], dtype=np.float64)
        self.upload(arr("x"), arr("y"), arr("wx"), arr("wy"), arr("mass"), arr("charge"), arr("angw"),
                    arr("base_mass"), antimatter=np.asarray(c["antimatter"], dtype=np.int32))
        return True


class HipBatch:
    """Many independent systems sharing one GeonParams, stepped together
    (include/geon.h, geon_batch_*): each system keeps the semantics of one
    Physics.update per call; results of a system do not depend on the batch."""

    def __init__(self, params, max_particles):
        self.lib = load_library()
        self.params = params
        caps = np.ascontiguousarray(max_particles, dtype=np.int32)
        self.n_systems = len(caps)
        h = C.c_void_p()
        rc = self.lib.geon_batch_create(C.byref(params), self.n_systems, caps.ctypes.data_as(C.POINTER(C.c_int32)),
                                        C.byref(h))
        if rc != 0:
            raise GeonError(rc, self.lib.geon_batch_last_error(None).decode())
        self.h = h
        self.caps = caps

    def destroy(self):
        if self.h:
            self.lib.geon_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    def _check(self, rc):
        if rc < 0:
            raise GeonError(rc, self.lib.geon_batch_last_error(self.h).decode())
        return rc

    def upload(self, system, x, y, wx, wy, mass, charge, angw=None):
        arrs = [np.ascontiguousarray(a, dtype=np.float64).ravel() if a is not None else None
                for a in (x, y, wx, wy, mass, None, charge, angw)]
        _check_lengths(len(arrs[0]), arrs, ("x", "y", "wx", "wy", "mass", "base_mass", "charge", "angw"))
        v = A.GeonSoAView(*[_dptr(a) for a in arrs], None)
        self._check(self.lib.geon_batch_upload(self.h, int(system), C.byref(v), len(arrs[0])))
        self._keep = arrs

    def download(self, system):
        cap = int(self.caps[system])
        out = {k: np.zeros(cap) for k in ("x", "y", "wx", "wy", "mass", "charge", "angw")}
        alive = np.zeros(cap, dtype=np.int32)
        v = A.GeonSoAView(_dptr(out["x"]), _dptr(out["y"]), _dptr(out["wx"]), _dptr(out["wy"]), _dptr(out["mass"]),
                          None, _dptr(out["charge"]), _dptr(out["angw"]), alive.ctypes.data_as(C.POINTER(C.c_int32)))
        ns = C.c_int32()
        self._check(self.lib.geon_batch_download(self.h, int(system), C.byref(v), cap, C.byref(ns)))
        n = ns.value
        res = {k: a[:n] for k, a in out.items()}
        res["alive"] = alive[:n].astype(bool)
        return res

    def update(self, dt):
        """One Physics.update(dt) of every system; returns each system's substeps."""
        sub = np.zeros(self.n_systems, dtype=np.int32)
        self._check(self.lib.geon_batch_update(self.h, float(dt), sub.ctypes.data_as(C.POINTER(C.c_int32))))
        return sub

    def compute_forces(self):
        self._check(self.lib.geon_batch_compute_forces(self.h))

    def forces(self, system, n_slots):
        names = ("fx", "fy", "bz", "dbzdx", "dbzdy", "bgz", "dbgzdx", "dbgzdy", "frame_drag_torque",
                 "tidal_torque", "pe", "f1pn_x", "f1pn_y")
        out = {k: np.zeros(n_slots) for k in names}
        v = A.GeonForceView(*[_dptr(out[k]) for k in names])
        self._check(self.lib.geon_batch_get_forces(self.h, int(system), C.byref(v), int(n_slots)))
        return out

    def stats(self, system):
        st = A.GeonStats()
        self._check(self.lib.geon_batch_get_stats(self.h, int(system), C.byref(st)))
        return {k: getattr(st, k) for k, _ in A.GeonStats._fields_}

    def autotune(self):
        t = A.GeonBatchTune()
        self._check(self.lib.geon_batch_autotune(self.h, C.byref(t)))
        return {k: getattr(t, k) for k, _ in A.GeonBatchTune._fields_}

    def tune(self, **kw):
        t = A.GeonBatchTune()
        self._check(self.lib.geon_batch_get_tune(self.h, C.byref(t)))
        for k, v in kw.items():
            setattr(t, k, v)
        if kw:
            self._check(self.lib.geon_batch_set_tune(self.h, C.byref(t)))
        return {k: getattr(t, k) for k, _ in A.GeonBatchTune._fields_}
