"""ctypes binding of the C ABI in include/pibiti_sph.h and a host-side mirror
of the reference's cSPH class (source/SPH/SPH.h:9-50).

The binding loads the in-tree ``libpibiti_sph.so``.  There is no fallback: if
the library is missing, every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

_PKG = Path(__file__).resolve().parent
# PIBITI_SPH_HOST_LIB names a build of the library's host-only C++ (the
# sanitizer build of `make -C oracle asan`) for the CPU tests of the scene
# constants, scene files and injection; its entry points are the host-only
# ones, the device entry points stay unbound.
_HOST_LIB = os.environ.get("PIBITI_SPH_HOST_LIB")
LIB_PATH = Path(_HOST_LIB) if _HOST_LIB else _PKG / "libpibiti_sph.so"

# status codes (sph_status)
SPH_OK = 0
SPH_ERR_OUT_OF_MEMORY = 4
SPH_ERR_NOT_INITIALISED = 6
SPH_ERR_ORDER = 9
# sph_debug_poke targets
SPH_DEBUG_MOVER_COUNT, SPH_DEBUG_CELL_END, SPH_DEBUG_ORDER_ERROR = 1, 2, 3
# enums
SPH_EOS_LINEAR, SPH_EOS_TAIT = 0, 1
SPH_PREC_REFERENCE, SPH_PREC_KAHAN, SPH_PREC_FAST = 0, 1, 2
(SPH_ARRAY_POSITION, SPH_ARRAY_VELOCITY, SPH_ARRAY_DENSITY, SPH_ARRAY_PRESSURE, SPH_ARRAY_DVEL,
 SPH_ARRAY_TARGET, SPH_ARRAY_NEIGHBOURS, SPH_ARRAY_COLOR, SPH_ARRAY_DYE) = range(9)
_SCALAR_ARRAYS = (SPH_ARRAY_DENSITY, SPH_ARRAY_PRESSURE, SPH_ARRAY_NEIGHBOURS, SPH_ARRAY_DYE)
(SPH_CLR_DENS, SPH_CLR_ACCEL, SPH_CLR_DENS_ACC, SPH_CLR_VEL, SPH_CLR_VEL_ACC, SPH_CLR_VEL_RGB,
 SPH_CLR_NONE) = range(7)
SPH_DYE_OFF, SPH_DYE_BOX, SPH_DYE_SPHERE = range(3)
MLP_NUM_PARAMS = 5388
SPH_FLAG_KEEP_DVEL = 1
EMPTY_CELL = 0xFFFFFFFF

f3 = C.c_float * 3
u3 = C.c_uint32 * 3

# boundary types, z effects and accelerator shapes (source/CUDA/Params.cuh:23-38)
(SPH_BND_BOX, SPH_BND_CYL_Y, SPH_BND_CYL_Z, SPH_BND_CYL_YZ, SPH_BND_SPHERE,
 SPH_BND_PUMP_Y) = range(6)
SPH_BND_EFF_NONE, SPH_BND_EFF_WRAP, SPH_BND_EFF_CYCLE, SPH_BND_EFF_WAVE = range(4)
SPH_ACC_OFF, SPH_ACC_BOX, SPH_ACC_CYL_Y, SPH_ACC_CYL_Y_SMOOTH = range(4)
SPH_NUM_ACCEL = 4
SPH_PHYS_COLLIDER, SPH_PHYS_ACCEL, SPH_PHYS_ROTOR, SPH_PHYS_HMAP = 1, 2, 4, 8


class Accel(C.Structure):
    """sph_accel: one accelerator (struct Accel, source/CUDA/Params.cuh:40-45)."""

    _fields_ = [("pos", f3), ("size", f3), ("acc", f3), ("type", C.c_uint32)]


_ROTOR_FIELDS = [
    ("rot_type", C.c_int32),
    ("rot_size", C.c_int32 * 3),
    ("rot_blades", C.c_int32),
    ("rot_r", C.c_float),
    ("rot_spc", C.c_float),
    ("rot2_dist", C.c_float),
    ("rot2_vel", C.c_float),
    ("rot2_twist", C.c_float),
    ("hmap_type", C.c_int32),
    ("clr_type", C.c_uint32),
    ("hue", C.c_int32),
    ("brightness", C.c_float),
    ("contrast", C.c_float),
    ("dye_type", C.c_int32),
    ("dye_clear", C.c_int32),
    ("dye_fade", C.c_float),
    ("dye_pos", f3),
    ("dye_size", f3),
]

_SHAPE_FIELDS = [
    ("bnd_type", C.c_uint32),
    ("bnd_eff_z", C.c_uint32),
    ("rot2_angle", C.c_float),
    ("rot_twist", C.c_float),
    ("rot_angle", C.c_float),
    ("rot_vel", C.c_float),
    ("exit_vel", C.c_float),
    ("exit_dist", C.c_float),
    ("pump_ang_out", C.c_float),
    ("h_close", C.c_float),
    ("pump_rad_in", C.c_float),
    ("shape_s", C.c_float * 6),
]


class Scene(C.Structure):
    """sph_scene: raw scene values before Scene::_UpdatePar/_UpdateGrid."""

    _fields_ = [
        ("num_particles", C.c_uint32),
        ("max_par_in_cell", C.c_uint32),
        ("time_step", C.c_float),
        ("gravity", f3),
        ("global_damping", C.c_float),
        ("world_min", f3),
        ("world_max", f3),
        ("cell_size", C.c_float),
        ("particle_r", C.c_float),
        ("h", C.c_float),
        ("rest_density", C.c_float),
        ("stiffness", C.c_float),
        ("viscosity", C.c_float),
        ("min_dens", C.c_float),
        ("min_dist", C.c_float),
        ("particle_mass", C.c_float),
        ("dist_bnd_soft", C.c_float),
        ("dist_bnd_hard", C.c_float),
        ("bnd_stiff", C.c_float),
        ("bnd_damp", C.c_float),
        ("bnd_damp_c", C.c_float),
        ("eos", C.c_uint32),
        ("precision", C.c_uint32),
        ("cfl", C.c_float),
        *_SHAPE_FIELDS,
        ("collider", C.c_uint32),
        ("coll_pos", f3),
        ("coll_r", C.c_float),
        ("coll_spring", C.c_float),
        ("coll_damping", C.c_float),
        ("coll_shear", C.c_float),
        ("acc", Accel * SPH_NUM_ACCEL),
        *_ROTOR_FIELDS,
    ]


class Emitter(C.Structure):
    """sph_emitter (class Emitter, source/SPH/Scene.h:7-17)."""

    _fields_ = [("size", C.c_int32), ("size2", C.c_int32), ("pos", f3), ("rot", f3), ("vel", C.c_float)]


SPH_ROTOR_NONE, SPH_ROTOR_PUMP_Z, SPH_ROTOR_PUMP_Y, SPH_ROTOR_PROPELLER_Z = range(4)
SPH_HMAP_NONE, SPH_HMAP_XZ, SPH_HMAP_YZ = range(3)


class SceneInfo(C.Structure):
    """sph_scene_info: a scene file's non-parameter values (init volume,
    drops, emitters, camera)."""

    _fields_ = [
        ("title", C.c_char * 64),
        ("chapter", C.c_uint32),
        ("is_default", C.c_uint32),
        ("init_min", f3),
        ("init_max", f3),
        ("init_type", C.c_int32),
        ("init_last", C.c_int32),
        ("spacing", C.c_float),
        ("drop_r", C.c_float),
        ("rain", C.c_int32),
        ("emit", Emitter * 4),
        ("cam_pos", f3),
        ("cam_rot", f3),
        ("reserved", C.c_uint32 * 4),
    ]


class Params(C.Structure):
    """sph_params: derived kernel constants (SimParams equivalent)."""

    _fields_ = [
        ("time_step", C.c_float),
        ("num_particles", C.c_uint32),
        ("max_par_in_cell", C.c_uint32),
        ("gravity", f3),
        ("global_damping", C.c_float),
        ("grid_size", u3),
        ("cell_size", f3),
        ("grid_size_yx", C.c_uint32),
        ("num_cells", C.c_uint32),
        ("world_min", f3),
        ("world_max", f3),
        ("particle_r", C.c_float),
        ("h", C.c_float),
        ("h2", C.c_float),
        ("spiky_kern", C.c_float),
        ("lap_kern", C.c_float),
        ("poly6_kern", C.c_float),
        ("particle_mass", C.c_float),
        ("rest_density", C.c_float),
        ("stiffness", C.c_float),
        ("viscosity", C.c_float),
        ("min_dens", C.c_float),
        ("min_dist", C.c_float),
        ("dist_bnd_hard", C.c_float),
        ("dist_bnd_soft", C.c_float),
        ("bnd_damp", C.c_float),
        ("bnd_stiff", C.c_float),
        ("bnd_damp_c", C.c_float),
        ("bnd_type", C.c_uint32),
        ("bnd_eff_z", C.c_uint32),
        ("eos", C.c_uint32),
        ("precision", C.c_uint32),
        ("cfl", C.c_float),
        ("sound_speed", C.c_float),
        ("key_bits", C.c_uint32),
        ("force_r2_max", C.c_float),
        ("tune", C.c_uint32 * 4),
        ("cell_base", C.c_uint32),
        ("t_end", C.c_float),
        *[f for f in _SHAPE_FIELDS if f[0] not in ("bnd_type", "bnd_eff_z")],
        ("wave_zmin_shift", C.c_float),
        ("pump_xs", C.c_float),
        ("pump_zs", C.c_float),
        ("world_size_d", f3),
        ("collider", C.c_uint32),
        ("coll_pos", f3),
        ("coll_r", C.c_float),
        ("coll_spring", C.c_float),
        ("coll_damping", C.c_float),
        ("coll_shear", C.c_float),
        ("acc", Accel * SPH_NUM_ACCEL),
        *_ROTOR_FIELDS,
        ("physics", C.c_uint32),
        ("reserved", C.c_uint32 * 2),
    ]

    def copy(self) -> "Params":
        out = Params()
        C.memmove(C.byref(out), C.byref(self), C.sizeof(Params))
        return out


class DeviceView(C.Structure):
    """sph_device_view: device pointers of the current state."""

    _fields_ = [("pos4", C.c_void_p), ("vel4", C.c_void_p), ("colour4", C.c_void_p), ("index", C.c_void_p),
                ("stream", C.c_void_p), ("n", C.c_uint32), ("device", C.c_int32)]


class Stats(C.Structure):
    _fields_ = [
        ("mass", C.c_double),
        ("kinetic_energy", C.c_double),
        ("potential_energy", C.c_double),
        ("max_speed", C.c_double),
        ("dt_cfl", C.c_double),
        ("sim_time", C.c_double),
        ("steps", C.c_uint64),
        ("cap_cells", C.c_uint32),
        ("out_of_grid", C.c_uint32),
        ("max_cell_count", C.c_uint32),
        ("sort_errors", C.c_uint32),
        ("full_sort", C.c_uint32),
        ("reserved", C.c_uint32 * 5),
    ]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


class MlpNorm(C.Structure):
    """sph_mlp_norm: feature and output units of the learned surrogate."""
    _fields_ = [
        ("inv_h", C.c_float),
        ("inv_v", C.c_float),
        ("inv_rho0", C.c_float),
        ("inv_p", C.c_float),
        ("inv_nnb", C.c_float),
        ("agg_scale", C.c_float),
        ("a_scale", C.c_float),
        ("rho_scale", C.c_float),
    ]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class LearnedReport(C.Structure):
    _fields_ = [
        ("coarse_steps", C.c_uint32),
        ("fallbacks", C.c_uint32),
        ("plain_steps", C.c_uint32),
        ("rebuilds", C.c_uint32),
        ("drift_max", C.c_double),
        ("drift_last", C.c_double),
        ("anchor", C.c_double),
        ("sim_time", C.c_double),
        ("density_ms", C.c_double),
        ("mlp_ms", C.c_double),
        ("coarse_ms", C.c_double),
        ("pairs", C.c_double),
        ("rebuild_ms", C.c_double),
        ("early_rebuilds", C.c_uint32),
        ("reserved0", C.c_uint32),
    ]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_ if not k.startswith("reserved")}


class SphError(RuntimeError):
    def __init__(self, status: int, what: str):
        super().__init__(f"{what}: {strerror(status)} (status {status})")
        self.status = status


_lib_handle = None

_P = C.c_void_p
_fp = C.POINTER(C.c_float)
_up = C.POINTER(C.c_uint32)

# exported symbol -> (restype, argtypes); every symbol declared in pibiti_sph.h
SIGNATURES = {
    "sph_abi_version": (C.c_uint32, []),
    "sph_strerror": (C.c_char_p, [C.c_int]),
    "sph_scene_default": (C.c_int, [C.POINTER(Scene)]),
    "sph_params_from_scene": (C.c_int, [C.POINTER(Scene), C.POINTER(Params)]),
    "sph_scene_lattice": (C.c_int, [C.c_uint32, C.c_uint32, C.c_uint32, _fp, C.c_float, C.c_float,
                                    C.c_float, _fp, C.c_uint64, C.c_uint64, _fp, _fp]),
    "sph_xml_scene_count": (C.c_int, [C.c_char_p, C.c_size_t, _up, _up]),
    "sph_xml_scene": (C.c_int, [C.c_char_p, C.c_size_t, C.c_uint32, C.POINTER(Scene),
                                C.POINTER(SceneInfo)]),
    "sph_scene_fill": (C.c_int, [C.POINTER(Scene), C.POINTER(SceneInfo), C.c_uint64, _fp, _fp]),
    "sph_scene_drop": (C.c_int, [C.POINTER(Scene), C.POINTER(SceneInfo), C.c_int, C.c_uint64, C.c_uint64,
                                 C.c_uint32, _fp, _fp, _up]),
    "sph_scene_emit": (C.c_int, [C.POINTER(SceneInfo), C.c_uint32, _fp, _fp, _up]),
    "sph_create": (C.c_int, [C.POINTER(Params), C.c_int, C.c_uint32, C.POINTER(_P)]),
    "sph_destroy": (C.c_int, [_P]),
    "sph_set_params": (C.c_int, [_P, C.POINTER(Params)]),
    "sph_set_array": (C.c_int, [_P, C.c_int, _fp, C.c_uint32, C.c_uint32]),
    "sph_get_array": (C.c_int, [_P, C.c_int, _fp]),
    "sph_step": (C.c_int, [_P, C.c_uint32]),
    "sph_get_grid": (C.c_int, [_P, _up, _up, _up, _up]),
    "sph_get_stats": (C.c_int, [_P, C.POINTER(Stats)]),
    "sph_synchronize": (C.c_int, [_P]),
    "sph_debug_poke": (C.c_int, [_P, C.c_int, C.c_uint32, C.c_uint32]),
    "sph_debug_fail_alloc": (C.c_int, [C.c_uint32]),
    "sph_radix_sort": (C.c_int, [_P, _P, C.c_uint32, C.c_uint32, _P]),
    "sph_radix_sort_noalloc": (C.c_int, [_P, _P, C.c_uint32, C.c_uint32, _P]),
    "sph_get_device_view": (C.c_int, [_P, C.POINTER(DeviceView)]),
    "sph_set_timing": (C.c_int, [_P, C.c_int]),
    "sph_get_timing": (C.c_int, [_P, _fp]),
    "sph_capture": (C.c_int, [_P, C.c_uint32]),
    "sph_replay": (C.c_int, [_P, C.c_uint32]),
    "sph_batch_create": (C.c_int, [C.POINTER(Params), C.c_uint32, C.c_int, C.c_uint32,
                                   C.POINTER(_P)]),
    "sph_batch_destroy": (C.c_int, [_P]),
    "sph_batch_get_params": (C.c_int, [_P, C.c_uint32, C.POINTER(Params)]),
    "sph_batch_set_params": (C.c_int, [_P, C.c_uint32, C.POINTER(Params)]),
    "sph_batch_set_array": (C.c_int, [_P, C.c_uint32, C.c_int, _fp, C.c_uint32, C.c_uint32]),
    "sph_batch_get_array": (C.c_int, [_P, C.c_uint32, C.c_int, _fp]),
    "sph_batch_get_grid": (C.c_int, [_P, C.c_uint32, _up, _up, _up, _up]),
    "sph_batch_step": (C.c_int, [_P, C.c_uint32]),
    "sph_batch_capture": (C.c_int, [_P, C.c_uint32]),
    "sph_batch_replay": (C.c_int, [_P, C.c_uint32]),
    "sph_batch_get_stats": (C.c_int, [_P, C.c_uint32, C.POINTER(Stats)]),
    "sph_batch_synchronize": (C.c_int, [_P]),
    "sph_batch_set_tune": (C.c_int, [_P, _up]),
    "sph_batch_autotune": (C.c_int, [_P, C.c_char_p, _up, _fp]),
    "sph_batch_info": (C.c_int, [_P, _up, _up, _up, _up]),
    "sph_mfma_selftest": (C.c_int, [C.c_int, C.POINTER(C.c_uint16), C.POINTER(C.c_uint16), _fp]),
    "sph_model_create": (C.c_int, [C.c_int, C.POINTER(MlpNorm), C.c_uint64, C.POINTER(_P)]),
    "sph_model_destroy": (C.c_int, [_P]),
    "sph_model_get_norm": (C.c_int, [_P, C.POINTER(MlpNorm)]),
    "sph_model_get_weights": (C.c_int, [_P, _fp]),
    "sph_model_set_weights": (C.c_int, [_P, _fp]),
    "sph_learn_begin": (C.c_int, [_P, _P]),
    "sph_learn_add": (C.c_int, [_P, _P]),
    "sph_learn_train": (C.c_int, [_P, _P, C.c_uint32, C.c_float, C.POINTER(C.c_double)]),
    "sph_learn_gradient": (C.c_int, [_P, _P, C.POINTER(C.c_double), _fp]),
    "sph_learn_predict": (C.c_int, [_P, _P, _fp]),
    "sph_learn_get_window": (C.c_int, [_P, C.c_int, _fp]),
    "sph_batch_learn_begin": (C.c_int, [_P, _P]),
    "sph_batch_learn_add": (C.c_int, [_P, _P]),
    "sph_batch_learn_train": (C.c_int, [_P, _P, C.c_uint32, C.c_float, C.POINTER(C.c_double)]),
    "sph_batch_learn_gradient": (C.c_int, [_P, _P, C.POINTER(C.c_double), _fp]),
    "sph_batch_learn_get_window": (C.c_int, [_P, C.c_uint32, C.c_int, _fp]),
    "sph_step_learned": (C.c_int, [_P, _P, C.c_uint32, C.c_uint32, C.c_float, C.c_float,
                                   C.POINTER(LearnedReport)]),
}


def lib() -> C.CDLL:
    """Load the in-tree HIP library (raises if it has not been built)."""
    global _lib_handle
    if _lib_handle is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m pibiti_amd.build`")
        h = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(h, name, None) if _HOST_LIB else getattr(h, name)
            if fn is None:
                continue
            fn.restype = res
            fn.argtypes = args
        _lib_handle = h
    return _lib_handle


def strerror(status: int) -> str:
    return lib().sph_strerror(status).decode()


def _check(status: int, what: str) -> None:
    if status != SPH_OK:
        raise SphError(status, what)


def _f(a: np.ndarray):
    return a.ctypes.data_as(_fp)


def _u(a: np.ndarray):
    return a.ctypes.data_as(_up)


def scene_default() -> Scene:
    s = Scene()
    _check(lib().sph_scene_default(C.byref(s)), "sph_scene_default")
    return s


def params_from_scene(s: Scene) -> Params:
    p = Params()
    _check(lib().sph_params_from_scene(C.byref(s), C.byref(p)), "sph_params_from_scene")
    return p


def scene_lattice(nx: int, ny: int, nz: int, region_min, spacing: float, jitter: float,
                  vel_sigma: float = 0.0, vel_offset=(0.0, 0.0, 0.0), seed: int = 0,
                  stream: int = 0) -> tuple[np.ndarray, np.ndarray]:
    """Seeded jittered lattice (float4 positions, float4 velocities)."""
    n = nx * ny * nz
    pos = np.zeros((n, 4), np.float32)
    vel = np.zeros((n, 4), np.float32)
    rmin = np.asarray(region_min, np.float32)
    voff = np.asarray(vel_offset, np.float32)
    _check(lib().sph_scene_lattice(nx, ny, nz, _f(rmin), spacing, jitter, vel_sigma, _f(voff),
                                   seed, stream, _f(pos), _f(vel)), "sph_scene_lattice")
    return pos, vel


def _xml_bytes(text) -> bytes:
    return text.encode("latin-1") if isinstance(text, str) else bytes(text)


def xml_scene_count(text) -> tuple[int, int]:
    """(number of scenes, startup scene index) of a scene file's text
    (cSPH::LoadScenes, source/SPH/SPH_Scenes.cpp:53-83)."""
    b = _xml_bytes(text)
    n, d = C.c_uint32(), C.c_uint32()
    _check(lib().sph_xml_scene_count(b, len(b), C.byref(n), C.byref(d)), "sph_xml_scene_count")
    return int(n.value), int(d.value)


def xml_scene(text, index: int) -> tuple[Scene, SceneInfo]:
    """Scene `index` of a scene file: Scene::InitDefault + Scene::_FromXML
    (source/SPH/Scene.cpp:7-66, source/SPH/Scene_Load.cpp:11-111)."""
    b = _xml_bytes(text)
    s, info = Scene(), SceneInfo()
    _check(lib().sph_xml_scene(b, len(b), index, C.byref(s), C.byref(info)), "sph_xml_scene")
    return s, info


def scene_fill(s: Scene, info: SceneInfo, seed: int = 0) -> tuple[np.ndarray, np.ndarray]:
    """cSPH::Reset(initType) (source/SPH/SPH_Init.cpp:23-76): positions and
    zero velocities of the scene's particles."""
    n = int(s.num_particles)
    pos = np.zeros((n, 4), np.float32)
    vel = np.zeros((n, 4), np.float32)
    _check(lib().sph_scene_fill(C.byref(s), C.byref(info), seed, _f(pos), _f(vel)), "sph_scene_fill")
    return pos, vel


SPH_MAX_EMIT = 65536


def scene_drop(s: Scene, info: SceneInfo, random: bool = False, seed: int = 0, draw: int = 0,
               first: int = 0) -> tuple[np.ndarray, np.ndarray]:
    """cSPH::Drop (source/SPH/SPH_Init.cpp:82-118): the records of one drop
    written from particle `first` (those that fit below the particle count)."""
    r = int(info.drop_r) + 1
    cap = min((2 * r + 1) ** 3, int(s.num_particles) - first)
    pos = np.zeros((max(cap, 1), 4), np.float32)
    vel = np.zeros((max(cap, 1), 4), np.float32)
    n = C.c_uint32()
    _check(lib().sph_scene_drop(C.byref(s), C.byref(info), int(random), seed, draw, first, _f(pos), _f(vel),
                                C.byref(n)), "sph_scene_drop")
    return pos[:n.value].copy(), vel[:n.value].copy()


def scene_emit(info: SceneInfo, emitter: int) -> tuple[np.ndarray, np.ndarray]:
    """One frame of an emitter's block (App::UpdateEmitter, source/App/Update.cpp:62-92)."""
    e = info.emit[emitter]
    cap = max(1, min(SPH_MAX_EMIT, max(e.size, 0) * max(e.size2 or e.size, 0)))
    pos = np.zeros((cap, 4), np.float32)
    vel = np.zeros((cap, 4), np.float32)
    n = C.c_uint32()
    _check(lib().sph_scene_emit(C.byref(info), emitter, _f(pos), _f(vel), C.byref(n)), "sph_scene_emit")
    return pos[:n.value].copy(), vel[:n.value].copy()


def debug_fail_alloc(after: int) -> None:
    """Testing hook: the after-th device allocation of the library from now
    on fails with SPH_ERR_OUT_OF_MEMORY (0 disarms; sph_debug_fail_alloc)."""
    _check(lib().sph_debug_fail_alloc(int(after)), "sph_debug_fail_alloc")


class SphSystem:
    """Mirror of cSPH (source/SPH/SPH.h:9-50) over the C ABI.

    init with params ~ cSPH::InitScene/_InitMem, set_array ~ setArray,
    get_array ~ getArray, update ~ Update.
    """

    def __init__(self, params: Params, device: int = 0, keep_dvel: bool = False):
        self.params = params.copy()
        self.n = int(params.num_particles)
        self._ctx = C.c_void_p()
        flags = SPH_FLAG_KEEP_DVEL if keep_dvel else 0
        _check(lib().sph_create(C.byref(self.params), device, flags, C.byref(self._ctx)), "sph_create")

    # context management -------------------------------------------------
    def close(self) -> None:
        if self._ctx:
            lib().sph_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # reference interface ------------------------------------------------
    def set_params(self, params: Params) -> None:
        _check(lib().sph_set_params(self._ctx, C.byref(params)), "sph_set_params")
        self.params = params.copy()

    def set_array(self, which: int, data: np.ndarray, first: int = 0) -> None:
        a = np.ascontiguousarray(data, np.float32).reshape(-1, 4)
        _check(lib().sph_set_array(self._ctx, which, _f(a), first, a.shape[0]), "sph_set_array")

    def reset(self, pos4: np.ndarray, vel4: np.ndarray) -> None:
        self.set_array(SPH_ARRAY_POSITION, pos4)
        self.set_array(SPH_ARRAY_VELOCITY, vel4)

    def get_array(self, which: int) -> np.ndarray:
        """One array in original particle order (NEIGHBOURS as uint32)."""
        shape = (self.n,) if which in _SCALAR_ARRAYS else (self.n, 4)
        out = np.empty(shape, np.float32)
        _check(lib().sph_get_array(self._ctx, which, _f(out)), "sph_get_array")
        return out.view(np.uint32) if which == SPH_ARRAY_NEIGHBOURS else out

    def update(self, nsteps: int = 1) -> None:
        _check(lib().sph_step(self._ctx, nsteps), "sph_step")

    def device_view(self) -> "DeviceView":
        v = DeviceView()
        _check(lib().sph_get_device_view(self._ctx, C.byref(v)), "sph_get_device_view")
        return v

    def synchronize(self) -> None:
        _check(lib().sph_synchronize(self._ctx), "sph_synchronize")

    def debug_poke(self, what: int, index: int = 0, value: int = 0) -> None:
        """Testing hook: perturb one word of the order state (sph_debug_poke)."""
        _check(lib().sph_debug_poke(self._ctx, what, index, value & 0xFFFFFFFF), "sph_debug_poke")

    def get_grid(self) -> dict:
        n, nc = self.n, int(self.params.num_cells)
        out = {
            "sorted_hash": np.empty(n, np.uint32),
            "perm": np.empty(n, np.uint32),
            "cell_start": np.empty(nc, np.uint32),
            "cell_end": np.empty(nc, np.uint32),
        }
        _check(lib().sph_get_grid(self._ctx, _u(out["sorted_hash"]), _u(out["perm"]),
                                  _u(out["cell_start"]), _u(out["cell_end"])), "sph_get_grid")
        return out

    def stats(self) -> dict:
        s = Stats()
        _check(lib().sph_get_stats(self._ctx, C.byref(s)), "sph_get_stats")
        return s.as_dict()

    def set_timing(self, enable: bool) -> None:
        _check(lib().sph_set_timing(self._ctx, int(enable)), "sph_set_timing")

    def timing(self) -> dict:
        ms = np.zeros(6, np.float32)
        _check(lib().sph_get_timing(self._ctx, _f(ms)), "sph_get_timing")
        keys = ["integrate_hash", "sort", "reorder", "density", "force", "step"]
        return dict(zip(keys, map(float, ms)))

    def capture(self, nsteps: int) -> None:
        """Record nsteps (even) steps into a hipGraph."""
        _check(lib().sph_capture(self._ctx, nsteps), "sph_capture")

    def replay(self, times: int = 1) -> None:
        _check(lib().sph_replay(self._ctx, times), "sph_replay")


class Model:
    """Learned pair-force surrogate on one device (sph_model)."""

    def __init__(self, norm: MlpNorm, device: int = 0, seed: int = 0):
        self._m = C.c_void_p()
        self.device = device
        _check(lib().sph_model_create(device, C.byref(norm), seed, C.byref(self._m)), "sph_model_create")

    def close(self) -> None:
        if self._m:
            lib().sph_model_destroy(self._m)
            self._m = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._m

    def norm(self) -> MlpNorm:
        n = MlpNorm()
        _check(lib().sph_model_get_norm(self._m, C.byref(n)), "sph_model_get_norm")
        return n

    def weights(self) -> np.ndarray:
        w = np.empty(MLP_NUM_PARAMS, np.float32)
        _check(lib().sph_model_get_weights(self._m, _f(w)), "sph_model_get_weights")
        return w

    def set_weights(self, w: np.ndarray) -> None:
        a = np.ascontiguousarray(w, np.float32).reshape(-1)
        if a.size != MLP_NUM_PARAMS:
            raise ValueError(f"expected {MLP_NUM_PARAMS} weights, got {a.size}")
        _check(lib().sph_model_set_weights(self._m, _f(a)), "sph_model_set_weights")


def _window_shape(which: int, n: int) -> tuple:
    return (n,) if which in (SPH_ARRAY_DENSITY, SPH_ARRAY_PRESSURE) else (n, 4)


def _learn_methods(prefix: str):
    """Training-window methods shared by SphSystem and SphBatch."""

    def learn_begin(self, model: Model) -> None:
        _check(getattr(lib(), prefix + "learn_begin")(self._h, model.handle), prefix + "learn_begin")

    def learn_add(self, model: Model) -> None:
        _check(getattr(lib(), prefix + "learn_add")(self._h, model.handle), prefix + "learn_add")

    def learn_train(self, model: Model, iters: int, lr: float) -> float:
        loss = C.c_double()
        _check(getattr(lib(), prefix + "learn_train")(self._h, model.handle, iters, lr, C.byref(loss)),
               prefix + "learn_train")
        return loss.value

    def learn_gradient(self, model: Model) -> tuple[float, np.ndarray]:
        loss = C.c_double()
        g = np.empty(MLP_NUM_PARAMS, np.float32)
        _check(getattr(lib(), prefix + "learn_gradient")(self._h, model.handle, C.byref(loss), _f(g)),
               prefix + "learn_gradient")
        return loss.value, g

    return learn_begin, learn_add, learn_train, learn_gradient


(SphSystem.learn_begin, SphSystem.learn_add, SphSystem.learn_train,
 SphSystem.learn_gradient) = _learn_methods("sph_")
SphSystem._h = property(lambda self: self._ctx)


def _sys_learn_predict(self, model: Model) -> np.ndarray:
    out = np.empty((self.n, 4), np.float32)
    _check(lib().sph_learn_predict(self._ctx, model.handle, _f(out)), "sph_learn_predict")
    return out


def _sys_learn_window(self, which: int) -> np.ndarray:
    out = np.empty(_window_shape(which, self.n), np.float32)
    _check(lib().sph_learn_get_window(self._ctx, which, _f(out)), "sph_learn_get_window")
    return out


def _sys_step_learned(self, model: Model, ncoarse: int, factor: int = 4, tol: float = 0.01,
                      skin: float = 0.0) -> dict:
    """``ncoarse`` surrogate steps of ``factor`` x dt.  skin = 0 sorts every
    step; skin > 0 sorts once per tune[2] steps and in between walks every
    cell within (1 + skin) h of the last sort's cell ranges."""
    r = LearnedReport()
    _check(lib().sph_step_learned(self._ctx, model.handle, ncoarse, factor, tol, skin, C.byref(r)),
           "sph_step_learned")
    return r.as_dict()


SphSystem.learn_predict = _sys_learn_predict
SphSystem.learn_window = _sys_learn_window
SphSystem.step_learned = _sys_step_learned


class SphBatch:
    """Many independent scenes advanced by the same launches (sph_batch_*)."""

    def __init__(self, scene_params: list, device: int = 0, keep_dvel: bool = False):
        arr = (Params * len(scene_params))(*[p.copy() for p in scene_params])
        self._b = C.c_void_p()
        flags = SPH_FLAG_KEEP_DVEL if keep_dvel else 0
        _check(lib().sph_batch_create(arr, len(scene_params), device, flags, C.byref(self._b)),
               "sph_batch_create")
        self.nscenes = len(scene_params)
        self.counts = [int(p.num_particles) for p in scene_params]

    def close(self) -> None:
        if self._b:
            lib().sph_batch_destroy(self._b)
            self._b = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self, scene: int) -> Params:
        p = Params()
        _check(lib().sph_batch_get_params(self._b, scene, C.byref(p)), "sph_batch_get_params")
        return p

    def set_params(self, scene: int, p: Params) -> None:
        _check(lib().sph_batch_set_params(self._b, scene, C.byref(p)), "sph_batch_set_params")

    def set_array(self, scene: int, which: int, data: np.ndarray, first: int = 0) -> None:
        a = np.ascontiguousarray(data, np.float32).reshape(-1, 4)
        _check(lib().sph_batch_set_array(self._b, scene, which, _f(a), first, a.shape[0]),
               "sph_batch_set_array")

    def reset(self, scene: int, pos4: np.ndarray, vel4: np.ndarray) -> None:
        self.set_array(scene, SPH_ARRAY_POSITION, pos4)
        self.set_array(scene, SPH_ARRAY_VELOCITY, vel4)

    def get_array(self, scene: int, which: int) -> np.ndarray:
        n = self.counts[scene]
        shape = (n,) if which in _SCALAR_ARRAYS else (n, 4)
        out = np.empty(shape, np.float32)
        _check(lib().sph_batch_get_array(self._b, scene, which, _f(out)), "sph_batch_get_array")
        return out

    def get_grid(self, scene: int) -> dict:
        """One scene's grid in scene-local keys, indices and slots."""
        p = self.params(scene)
        n, nc = int(p.num_particles), int(p.num_cells)
        out = {"sorted_hash": np.empty(n, np.uint32), "perm": np.empty(n, np.uint32),
               "cell_start": np.empty(nc, np.uint32), "cell_end": np.empty(nc, np.uint32)}
        _check(lib().sph_batch_get_grid(self._b, scene, _u(out["sorted_hash"]), _u(out["perm"]),
                                        _u(out["cell_start"]), _u(out["cell_end"])), "sph_batch_get_grid")
        return out

    def update(self, nsteps: int = 1) -> None:
        _check(lib().sph_batch_step(self._b, nsteps), "sph_batch_step")

    def capture(self, nsteps: int) -> None:
        _check(lib().sph_batch_capture(self._b, nsteps), "sph_batch_capture")

    def replay(self, times: int = 1) -> None:
        _check(lib().sph_batch_replay(self._b, times), "sph_batch_replay")

    def synchronize(self) -> None:
        _check(lib().sph_batch_synchronize(self._b), "sph_batch_synchronize")

    def stats(self, scene: int) -> dict:
        s = Stats()
        _check(lib().sph_batch_get_stats(self._b, scene, C.byref(s)), "sph_batch_get_stats")
        return s.as_dict()

    def set_tune(self, tune) -> None:
        t = np.asarray(list(tune) + [0] * (4 - len(tune)), np.uint32)
        _check(lib().sph_batch_set_tune(self._b, _u(t)), "sph_batch_set_tune")

    def autotune(self, cache_path: str | None = None) -> tuple[list[int], float]:
        t = np.zeros(4, np.uint32)
        ms = np.zeros(1, np.float32)
        path = cache_path.encode() if cache_path else None
        _check(lib().sph_batch_autotune(self._b, path, _u(t), _f(ms)), "sph_batch_autotune")
        return [int(x) for x in t], float(ms[0])

    def info(self) -> dict:
        v = np.zeros(4, np.uint32)
        _check(lib().sph_batch_info(self._b, _u(v[0:1]), _u(v[1:2]), _u(v[2:3]), _u(v[3:4])),
               "sph_batch_info")
        return dict(nscenes=int(v[0]), particles=int(v[1]), cells=int(v[2]), key_bits=int(v[3]))


(SphBatch.learn_begin, SphBatch.learn_add, SphBatch.learn_train,
 SphBatch.learn_gradient) = _learn_methods("sph_batch_")
SphBatch._h = property(lambda self: self._b)


def _batch_learn_window(self, scene: int, which: int) -> np.ndarray:
    n = int(self.params(scene).num_particles)
    out = np.empty(_window_shape(which, n), np.float32)
    _check(lib().sph_batch_learn_get_window(self._b, scene, which, _f(out)), "sph_batch_learn_get_window")
    return out


SphBatch.learn_window = _batch_learn_window
