"""The reference's save-state format, version 1, mapped onto GeonParams and the
SoA upload.  Host code only: no device.

save-load.js:89-127 writes the format and :129-262 reads it; the GPU backend
keeps its own copy (gpu-physics.js:3749-3836 serialize, 3886-3960 deserialize),
whose per-particle defaults are the ones used here, since this backend stands
where GPUPhysics does.  A state is a JSON object:

    {"version": 1,
     "particles": [{"x", "y", "wx", "wy", "mass", "baseMass", "charge", "angw", "antimatter"}, ...],
     "toggles": {"gravityEnabled": bool, ...},
     "settings": {"collision": "pass"|"merge"|"bounce", "boundary": "despawn"|"bounce"|"loop",
                  "topology": "torus"|"klein"|"rp2", "speed": number, "friction": number},
     "camera": {"x", "y", "zoom"},
     "yukawaMu", "axionMass", "hubbleParam", "higgsMass",
     "extGravity", "extGravityAngle", "extElectric", "extElectricAngle", "extBz"}
"""
import math

from . import _abi as A

# toggles0 bits (gpu-constants.js:196-213)
TOGGLE_BITS = A.TOGGLE_BITS
# the toggles a state carries, in the order _cpuSaveState writes them (save-load.js:107-113)
TOGGLE_KEYS = ["gravityEnabled", "bosonInterEnabled", "coulombEnabled", "magneticEnabled", "gravitomagEnabled",
               "relativityEnabled", "barnesHutEnabled", "radiationEnabled", "blackHoleEnabled",
               "disintegrationEnabled", "spinOrbitEnabled", "onePNEnabled", "yukawaEnabled", "axionEnabled",
               "expansionEnabled", "higgsEnabled"]

# config.js:12-25
COL_NAMES = ["pass", "merge", "bounce"]
BOUND_NAMES = ["despawn", "bounce", "loop"]
TOPO_NAMES = ["torus", "klein", "rp2"]


def col_from_string(s):
    return A.COL_MERGE if s == "merge" else A.COL_BOUNCE if s == "bounce" else A.COL_PASS


def bound_from_string(s):
    return A.BOUND_LOOP if s == "loop" else A.BOUND_BOUNCE if s == "bounce" else A.BOUND_DESPAWN


def topo_from_string(s):
    return A.KLEIN if s == "klein" else A.RP2 if s == "rp2" else A.TORUS


# (state key, GeonParams field, restored when): "value" = `!= null`
# (present and not null), "defined" = `!== undefined` (present; null reads as 0)
# (save-load.js:116-125, 221-230)
PARAM_KEYS = [
    ("yukawaMu", "yukawa_mu", "value"), ("axionMass", "axion_mass", "value"), ("hubbleParam", "hubble", "value"),
    ("higgsMass", "higgs_mass", "defined"), ("extGravity", "ext_gravity", "defined"),
    ("extGravityAngle", "ext_gravity_angle", "defined"), ("extElectric", "ext_electric", "defined"),
    ("extElectricAngle", "ext_electric_angle", "defined"), ("extBz", "ext_bz", "defined"),
]


def params_to_state(params, particles, sim=None):
    """A v1 state object from GeonParams and the live particles (a list of
    dicts with the particle keys).  `sim` may carry the app's UI settings
    (`speedScale`, `camera`), which the format also stores; settings.friction
    is the physics' bounceFriction (save-load.js:103)."""
    sim = sim or {}
    state = {
        "version": 1,
        "particles": particles,
        "toggles": {k: bool(params.toggles0 & TOGGLE_BITS[k]) for k in TOGGLE_KEYS},
        "settings": {
            "collision": COL_NAMES[params.collision_mode],
            "boundary": BOUND_NAMES[params.boundary_mode],
            "topology": TOPO_NAMES[params.topology],
            "speed": sim.get("speedScale", 16),
            "friction": params.bounce_friction,
        },
    }
    if "camera" in sim:
        state["camera"] = dict(sim["camera"])
    for key, field, _ in PARAM_KEYS:
        state[key] = getattr(params, field)
    return state


def apply_state(params, state):
    """GeonParams with a state's toggles, physics parameters and mode settings
    applied over `params` (a copy; what the state leaves out stays)."""
    p = A.GeonParams.from_buffer_copy(params)
    t = p.toggles0
    for key, val in (state.get("toggles") or {}).items():
        bit = TOGGLE_BITS.get(key)
        if bit is None:
            continue                                    # `if (key in ph)` (save-load.js:215)
        t = (t | bit) if val else (t & ~bit)
    p.toggles0 = t
    for key, field, when in PARAM_KEYS:
        if key not in state:
            continue
        v = state[key]
        if v is None:
            if when == "value":
                continue
            v = 0.0
        setattr(p, field, float(v))
    st = state.get("settings")
    if st:
        col = "bounce" if st.get("collision") == "repel" else (st.get("collision") or "pass")   # save-load.js:194
        p.collision_mode = col_from_string(col)
        p.boundary_mode = bound_from_string(st.get("boundary"))
        p.topology = topo_from_string(st.get("topology"))
        if st.get("friction") is not None:
            p.bounce_friction = float(st["friction"])                  # save-load.js:223
    return p


def _num(v, default):
    """JavaScript's `v || default` on a JSON number (undefined, null, 0 and NaN give the default)."""
    if v is None or v == 0 or (isinstance(v, float) and math.isnan(v)):
        return default
    return float(v)


def particles_from_state(state, cap):
    """Columns for the SoA upload, with the GPU backend's defaults
    (gpu-physics.js:3896-3912): mass || 1, baseMass ?? mass, charge || 0,
    angw || 0, wx || 0, wy || 0; at most `cap` particles."""
    ps = state["particles"][:cap]
    cols = {k: [] for k in ("x", "y", "wx", "wy", "mass", "base_mass", "charge", "angw", "antimatter")}
    for pd in ps:
        m = _num(pd.get("mass"), 1.0)
        bm = pd.get("baseMass")
        cols["x"].append(float(pd["x"]))
        cols["y"].append(float(pd["y"]))
        cols["wx"].append(_num(pd.get("wx"), 0.0))
        cols["wy"].append(_num(pd.get("wy"), 0.0))
        cols["mass"].append(m)
        cols["base_mass"].append(m if bm is None else float(bm))
        cols["charge"].append(_num(pd.get("charge"), 0.0))
        cols["angw"].append(_num(pd.get("angw"), 0.0))
        cols["antimatter"].append(1 if pd.get("antimatter") else 0)
    return cols
