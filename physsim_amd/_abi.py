"""ctypes mirror of include/geon.h (structs, constants, prototypes)."""
import ctypes as C

ABI_VERSION = 1

OK, EINVAL, EFULL, EDEVICE, ENOMEM, ESTATE, EUNSUPPORTED, EOVERFLOW = 0, -1, -2, -3, -4, -5, -6, -7

GRAVITY, COULOMB, MAGNETIC, GRAVITOMAG = 1 << 0, 1 << 1, 1 << 2, 1 << 3
ONE_PN, RELATIVITY, SPIN_ORBIT, RADIATION = 1 << 4, 1 << 5, 1 << 6, 1 << 7
BLACK_HOLE, DISINTEGRATION, EXPANSION, YUKAWA = 1 << 8, 1 << 9, 1 << 10, 1 << 11
HIGGS, AXION, BARNES_HUT, BOSON_INTER = 1 << 12, 1 << 13, 1 << 14, 1 << 15

COL_PASS, COL_MERGE, COL_BOUNCE = 0, 1, 2
BOUND_DESPAWN, BOUND_BOUNCE, BOUND_LOOP = 0, 1, 2
TORUS, KLEIN, RP2 = 0, 1, 2
PERIODIC_MIN_IMAGE, PERIODIC_GHOSTS, PERIODIC_WINDOWED = 0, 1, 2
FAR_NONE, FAR_MULTIPOLE, FAR_LEARNED = 0, 1, 2
DELAY_FOLLOW, DELAY_OFF, DELAY_ON = 0, 1, 2
HISTORY_SIZE = 256
FAR_PARAMS, FAR_MODEL_FLOATS, FAR_SAMPLE = 5574, 5600, 20
FIELD_HIGGS, FIELD_AXION = 0, 1
FQ = {"field": 0, "field_dot": 1, "grad_x": 2, "grad_y": 3, "sg_phi": 4, "sg_gx": 5, "sg_gy": 6,
      "source": 7, "thermal": 8}

TOGGLE_BITS = {
    "gravityEnabled": GRAVITY, "coulombEnabled": COULOMB, "magneticEnabled": MAGNETIC,
    "gravitomagEnabled": GRAVITOMAG, "onePNEnabled": ONE_PN, "relativityEnabled": RELATIVITY,
    "spinOrbitEnabled": SPIN_ORBIT, "radiationEnabled": RADIATION, "blackHoleEnabled": BLACK_HOLE,
    "disintegrationEnabled": DISINTEGRATION, "expansionEnabled": EXPANSION, "yukawaEnabled": YUKAWA,
    "higgsEnabled": HIGGS, "axionEnabled": AXION, "barnesHutEnabled": BARNES_HUT,
    "bosonInterEnabled": BOSON_INTER,
}


class GeonParams(C.Structure):
    _fields_ = [
        ("toggles0", C.c_uint32), ("boundary_mode", C.c_int32), ("topology", C.c_int32),
        ("collision_mode", C.c_int32), ("periodic_mode", C.c_int32), ("max_substeps", C.c_int32),
        ("domain_w", C.c_double), ("domain_h", C.c_double), ("softening", C.c_double),
        ("bh_softening", C.c_double), ("theta", C.c_double), ("yukawa_coupling", C.c_double),
        ("yukawa_mu", C.c_double), ("higgs_mass", C.c_double), ("axion_mass", C.c_double),
        ("hubble", C.c_double), ("ext_gravity", C.c_double), ("ext_gravity_angle", C.c_double),
        ("ext_electric", C.c_double), ("ext_electric_angle", C.c_double), ("ext_bz", C.c_double),
        ("field_grid", C.c_int32), ("kahan", C.c_int32), ("far_field", C.c_int32), ("far_list_cap", C.c_int32),
        ("signal_delay", C.c_int32), ("history_stride", C.c_int32), ("bounce_friction", C.c_double),
    ]


class GeonParticleInit(C.Structure):
    _fields_ = [("x", C.c_double), ("y", C.c_double), ("wx", C.c_double), ("wy", C.c_double),
                ("mass", C.c_double), ("charge", C.c_double), ("angw", C.c_double),
                ("antimatter", C.c_int32)]


_DP = C.POINTER(C.c_double)
_IP = C.POINTER(C.c_int32)


class GeonSoAView(C.Structure):
    _fields_ = [("x", _DP), ("y", _DP), ("wx", _DP), ("wy", _DP), ("mass", _DP),
                ("base_mass", _DP), ("charge", _DP), ("angw", _DP), ("alive", _IP), ("antimatter", _IP)]


class GeonForceView(C.Structure):
    _fields_ = [("fx", _DP), ("fy", _DP), ("bz", _DP), ("dbzdx", _DP), ("dbzdy", _DP),
                ("bgz", _DP), ("dbgzdx", _DP), ("dbgzdy", _DP), ("frame_drag_torque", _DP),
                ("tidal_torque", _DP), ("pe", _DP), ("f1pn_x", _DP), ("f1pn_y", _DP),
                ("jerk_x", _DP), ("jerk_y", _DP)]


class GeonStats(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "linear_ke", "spin_ke", "px", "py", "orbital_ang_mom", "spin_ang_mom", "com_x", "com_y",
        "total_mass", "pe", "field_energy", "higgs_field_energy", "axion_field_energy", "pfi_energy",
        "field_px", "field_py", "sim_time")] + [("alive_count", C.c_int32), ("dead_count", C.c_int32)] + [
        (n, C.c_double) for n in ("radiated_energy", "radiated_px", "radiated_py")]


SF_NAMES = ("forceGravity", "forceCoulomb", "forceMagnetic", "forceGravitomag", "force1PN", "forceSpinCurv",
            "forceRadiation", "forceYukawa", "forceExternal", "forceHiggs", "forceAxion", "forceFarField")


class GeonSelected(C.Structure):
    _fields_ = [("slot", C.c_int32), ("forces_valid", C.c_int32), ("antimatter", C.c_int32), ("pad", C.c_int32)] + \
               [(k, C.c_double) for k in ("x", "y", "wx", "wy", "mass", "charge", "angw", "base_mass", "radius",
                                           "vx", "vy", "ang_vel", "mag_moment", "ang_momentum")] + \
               [("force", (C.c_double * 2) * len(SF_NAMES))]


class GeonHistoryView(C.Structure):
    _fields_ = [("x", _DP), ("y", _DP), ("vx", _DP), ("vy", _DP), ("angw", _DP), ("t", _DP)]


class GeonCounters(C.Structure):
    _fields_ = [(n, C.c_int64) for n in (
        "pair_interactions", "node_interactions", "nodes_visited", "tree_nodes", "tree_divided",
        "depth_overflow", "substeps_total")]


class GeonTreeNode(C.Structure):
    _fields_ = [("path", C.c_uint64), ("level", C.c_int32), ("divided", C.c_int32),
                ("count", C.c_int32), ("charge_quanta", C.c_int32), ("mass", C.c_double),
                ("com_x", C.c_double), ("com_y", C.c_double)]


class GeonBatchTune(C.Structure):
    _fields_ = [("walk_threads", C.c_int32), ("waves_per_cu", C.c_int32), ("stack", C.c_int32),
                ("use_graph", C.c_int32), ("graph_bodies", C.c_int32), ("walk_ms", C.c_double)]


H = C.c_void_p

# name -> (restype, argtypes)
PROTOTYPES = {
    "geon_abi_version": (C.c_int, []),
    "geon_default_params": (None, [C.POINTER(GeonParams)]),
    "geon_create": (C.c_int, [C.POINTER(GeonParams), C.c_int32, C.POINTER(H)]),
    "geon_destroy": (None, [H]),
    "geon_reset": (C.c_int, [H]),
    "geon_set_params": (C.c_int, [H, C.POINTER(GeonParams)]),
    "geon_get_params": (C.c_int, [H, C.POINTER(GeonParams)]),
    "geon_add_particle": (C.c_int32, [H, C.POINTER(GeonParticleInit)]),
    "geon_remove_particle": (C.c_int, [H, C.c_int32, C.c_double, C.c_double]),
    "geon_patch_mass": (C.c_int, [H, C.c_int32, C.c_double, C.c_double]),
    "geon_upload_soa": (C.c_int, [H, C.POINTER(GeonSoAView), C.c_int32]),
    "geon_download_soa": (C.c_int, [H, C.POINTER(GeonSoAView), C.c_int32, C.POINTER(C.c_int32)]),
    "geon_update": (C.c_int, [H, C.c_double, C.POINTER(C.c_int32)]),
    "geon_compute_forces": (C.c_int, [H]),
    "geon_get_forces": (C.c_int, [H, C.POINTER(GeonForceView), C.c_int32]),
    "geon_get_stats": (C.c_int, [H, C.POINTER(GeonStats)]),
    "geon_select": (C.c_int, [H, C.c_int32]),
    "geon_get_selected": (C.c_int, [H, C.POINTER(GeonSelected)]),
    "geon_particle_count": (C.c_int, [H, C.POINTER(C.c_int32)]),
    "geon_get_history": (C.c_int, [H, C.c_int32, C.POINTER(GeonHistoryView), C.POINTER(C.c_int32)]),
    "geon_build_tree": (C.c_int, [H]),
    "geon_get_tree": (C.c_int, [H, C.POINTER(GeonTreeNode), C.c_int64, C.POINTER(C.c_int64)]),
    "geon_set_field": (C.c_int, [H, C.c_int32, C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int32]),
    "geon_get_field": (C.c_int, [H, C.c_int32, C.c_int32, C.POINTER(C.c_double), C.c_int32]),
    "geon_field_selfgravity": (C.c_int, [H, C.c_int32]),
    "geon_fft2d": (C.c_int, [H, C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int32, C.c_int32]),
    "geon_collision_pairs": (C.c_int, [H, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.c_int64,
                                        C.POINTER(C.c_int64)]),
    "geon_collision_pairs_noalloc": (C.c_int, [H, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.c_int64,
                                        C.POINTER(C.c_int64)]),
    "geon_collision_events": (C.c_int, [H, C.POINTER(C.c_double), C.c_int64, C.POINTER(C.c_int64)]),
    "geon_counters": (C.c_int, [H, C.POINTER(GeonCounters)]),
    "geon_timing_enable": (C.c_int, [H, C.c_int32]),
    "geon_kernel_timing": (C.c_int, [H, C.c_int32, C.POINTER(C.c_double), C.POINTER(C.c_int64)]),
    "geon_last_error": (C.c_char_p, [H]),
    "geon_far_collect": (C.c_int, [H, C.c_int32, C.POINTER(C.c_int64)]),
    "geon_far_clear": (C.c_int, [H]),
    "geon_far_get_samples": (C.c_int, [H, C.c_int64, C.c_int64, C.POINTER(C.c_float), C.POINTER(C.c_int32)]),
    "geon_far_train": (C.c_int, [H, C.c_int32, C.c_int32, C.c_double, C.c_double, C.c_uint64,
                                  C.POINTER(C.c_float)]),
    "geon_far_grad": (C.c_int, [H, C.POINTER(C.c_int64), C.c_int32, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    "geon_far_get_model": (C.c_int, [H, C.POINTER(C.c_float), C.c_int32]),
    "geon_far_set_model": (C.c_int, [H, C.POINTER(C.c_float), C.c_int32]),
    "geon_batch_create": (C.c_int, [C.POINTER(GeonParams), C.c_int32, C.POINTER(C.c_int32), C.POINTER(H)]),
    "geon_batch_destroy": (None, [H]),
    "geon_batch_upload": (C.c_int, [H, C.c_int32, C.POINTER(GeonSoAView), C.c_int32]),
    "geon_batch_download": (C.c_int, [H, C.c_int32, C.POINTER(GeonSoAView), C.c_int32, C.POINTER(C.c_int32)]),
    "geon_batch_update": (C.c_int, [H, C.c_double, C.POINTER(C.c_int32)]),
    "geon_batch_compute_forces": (C.c_int, [H]),
    "geon_batch_get_forces": (C.c_int, [H, C.c_int32, C.POINTER(GeonForceView), C.c_int32]),
    "geon_batch_get_stats": (C.c_int, [H, C.c_int32, C.POINTER(GeonStats)]),
    "geon_batch_autotune": (C.c_int, [H, C.POINTER(GeonBatchTune)]),
    "geon_batch_set_tune": (C.c_int, [H, C.POINTER(GeonBatchTune)]),
    "geon_batch_get_tune": (C.c_int, [H, C.POINTER(GeonBatchTune)]),
    "geon_batch_last_error": (C.c_char_p, [H]),
}


def bind(lib):
    for name, (res, args) in PROTOTYPES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib
