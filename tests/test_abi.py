"""The C-ABI library loads, exports every entry point the public headers
declare, and its host-only functions and error paths behave (no device use)."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

from pibiti_amd import sph

ROOT = Path(__file__).resolve().parent.parent


def declared_functions():
    names = []
    for hdr in sorted((ROOT / "include").glob("*.h")):
        text = hdr.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\**\s+\**\s*(sph_[A-Za-z0-9_]+)\s*\(",
                             text, flags=re.M):
            names.append(m.group(1))
    return sorted(set(names))


def test_headers_declare_entry_points():
    names = declared_functions()
    assert "sph_create" in names and "sph_step" in names and "sph_get_array" in names
    assert len(names) >= 16


def test_library_exports_every_declared_symbol():
    h = C.CDLL(str(sph.LIB_PATH))
    missing = [n for n in declared_functions() if not hasattr(h, n)]
    assert not missing, missing


def test_binding_covers_every_declared_symbol():
    assert set(declared_functions()) <= set(sph.SIGNATURES)


def test_abi_version_and_strerror():
    assert sph.lib().sph_abi_version() == 2
    assert sph.strerror(0) == "ok"
    for code in range(1, 9):
        assert sph.strerror(code) not in ("ok", "unknown status")
    assert sph.strerror(999) == "unknown status"


def test_scene_default_matches_reference_defaults():
    s = sph.scene_default()
    # source/SPH/Scene.cpp:11-31
    assert s.num_particles == 7 * 8 * 1024 and s.max_par_in_cell == 16
    assert s.time_step == np.float32(0.0026) and s.h == np.float32(0.01)
    assert s.particle_r == np.float32(0.004) and s.rest_density == 1000
    assert s.stiffness == 3 and s.viscosity == np.float32(0.5)
    assert s.dist_bnd_soft == 8 and s.dist_bnd_hard == 1
    assert (s.bnd_stiff, s.bnd_damp, s.bnd_damp_c) == (30000, 256, 60)
    assert s.cell_size == np.float32(0.008)


@pytest.mark.parametrize("field,value", [
    ("num_particles", 0), ("h", 0.0), ("h", -1.0), ("cell_size", 0.0), ("time_step", 0.0),
    ("rest_density", 0.0), ("max_par_in_cell", 0), ("eos", 7), ("precision", 9), ("cfl", -1.0),
])
def test_params_from_scene_rejects_bad_input(field, value):
    s = sph.scene_default()
    setattr(s, field, value)
    p = sph.Params()
    assert sph.lib().sph_params_from_scene(C.byref(s), C.byref(p)) == 2  # SPH_ERR_INVALID_PARAMS


def test_params_from_scene_rejects_inverted_world():
    s = sph.scene_default()
    s.world_max[1] = s.world_min[1]
    p = sph.Params()
    assert sph.lib().sph_params_from_scene(C.byref(s), C.byref(p)) == 2


def test_null_arguments_are_errors_not_crashes():
    L = sph.lib()
    assert L.sph_scene_default(None) == 1
    assert L.sph_params_from_scene(None, None) == 1
    assert L.sph_create(None, 0, 0, None) == 1
    assert L.sph_destroy(None) == 0
    assert L.sph_step(None, 1) == 1
    assert L.sph_get_array(None, 0, None) == 1
    assert L.sph_set_array(None, 0, None, 0, 0) == 1
    assert L.sph_get_stats(None, None) == 1
    assert L.sph_get_grid(None, None, None, None, None) == 1
    assert L.sph_synchronize(None) == 1


def test_create_rejects_invalid_params_before_touching_a_device():
    p = sph.params_from_scene(sph.scene_default())
    p.num_particles = 0
    ctx = C.c_void_p()
    assert sph.lib().sph_create(C.byref(p), 0, 0, C.byref(ctx)) == 2
    p = sph.params_from_scene(sph.scene_default())
    p.key_bits = 2  # fewer bits than the cell count needs
    assert sph.lib().sph_create(C.byref(p), 0, 0, C.byref(ctx)) == 2
    for field, value in (("bnd_type", 6), ("bnd_eff_z", 4), ("collider", 2)):
        p = sph.params_from_scene(sph.scene_default())
        setattr(p, field, value)  # no such boundary type, effect or collider mode
        assert sph.lib().sph_create(C.byref(p), 0, 0, C.byref(ctx)) == 2
    p = sph.params_from_scene(sph.scene_default())
    p.acc[2].type = 4  # no such accelerator shape
    assert sph.lib().sph_create(C.byref(p), 0, 0, C.byref(ctx)) == 2


def test_batch_rejects_a_moving_wave():
    """A moving wave needs new parameters every step, which a batch's device
    table does not take: refused before any device use."""
    L = sph.lib()
    good = sph.params_from_scene(sph.scene_default())
    wave = good.copy()
    wave.bnd_eff_z = sph.SPH_BND_EFF_WAVE
    wave.rot_vel = 6.0
    arr = (sph.Params * 2)(good, wave)
    out = C.c_void_p()
    assert L.sph_batch_create(arr, 2, 0, 0, C.byref(out)) in (3, 7)  # no device here, or unsupported


C_TYPES = {C.c_float: "float", C.c_uint32: "uint32_t", C.c_double: "double", C.c_uint64: "uint64_t"}


def test_struct_layouts_match_the_header(tmp_path):
    """Every field of the ctypes mirrors sits at the offset the C compiler
    gives the header's struct, and the sizes agree."""
    import shutil
    import subprocess
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    structs = {"sph_scene": sph.Scene, "sph_params": sph.Params, "sph_accel": sph.Accel,
               "sph_emitter": sph.Emitter, "sph_scene_info": sph.SceneInfo, "sph_device_view": sph.DeviceView,
               "sph_stats": sph.Stats, "sph_mlp_norm": sph.MlpNorm,
               "sph_learned_report": sph.LearnedReport}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "pibiti_sph.h"', "int main(void) {"]
    expect = []
    for cname, py in structs.items():
        lines.append(f'printf("%zu\\n", sizeof({cname}));')
        expect.append(C.sizeof(py))
        for fname, _ in py._fields_:
            lines.append(f'printf("%zu\\n", offsetof({cname}, {fname}));')
            expect.append(getattr(py, fname).offset)
    lines += ["return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run([cc, "-I", str(ROOT / "include"), str(src), "-o", str(exe)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], check=True, capture_output=True,
                                          text=True).stdout.split()]
    assert got == expect


def test_batch_create_rejects_bad_scenes_and_null():
    L = sph.lib()
    good = sph.params_from_scene(sph.scene_default())
    bad = good.copy()
    bad.num_particles = 0
    arr = (sph.Params * 2)(good, bad)
    out = C.c_void_p()
    assert L.sph_batch_create(arr, 2, 0, 0, C.byref(out)) == 2
    assert L.sph_batch_create(None, 2, 0, 0, C.byref(out)) == 1
    assert L.sph_batch_create(arr, 0, 0, 0, C.byref(out)) == 1
    assert L.sph_batch_destroy(None) == 0
    assert L.sph_batch_step(None, 1) == 1
    assert L.sph_batch_autotune(None, None, None, None) == 1
    assert L.sph_capture(None, 2) == 1
    assert L.sph_replay(None, 1) == 1
