"""The C++ host class (include/pibiti_sph.hpp, the mirror of cSPH) through
its command-line driver pibiti_amd/_build/sph_run: the same results as the
Python binding and as the FP32 oracle, bit for bit."""
import subprocess

import numpy as np
import pytest

import oracle
from pibiti_amd import build, scenes, sph


def _run(args, **kw):
    return subprocess.run([str(build.RUN_BIN), *map(str, args)], capture_output=True, text=True, timeout=120,
                          **kw)


def test_driver_rejects_bad_usage():
    assert build.RUN_BIN.exists(), "build the library first (python -m pibiti_amd.build)"
    r = _run([])
    assert r.returncode == 2 and "usage" in r.stderr


@pytest.mark.gpu
def test_cpp_host_class_matches_binding_and_oracle(tmp_path):
    nx, ny, nz, spacing, h, dt, steps, seed = 12, 10, 12, 0.01, 0.02, 5e-5, 20, 4
    origin = scenes.fluid_origin(spacing)
    b = origin[0]
    wm = (2 * nx * spacing + 2 * b, 2 * ny * spacing + 2 * b, 2 * nz * spacing + 2 * b)
    s = scenes.scene(nx * ny * nz, h, spacing, (0.0, 0.0, 0.0), wm, dt, viscosity=0.1)
    (tmp_path / "scene.bin").write_bytes(bytes(s))
    out = tmp_path / "out.bin"
    print("case: C++ host class,", nx * ny * nz, "particles,", steps, "steps", flush=True)
    r = _run([tmp_path / "scene.bin", nx, ny, nz, *origin, spacing, 0.1, 0.05, seed, steps, out])
    assert r.returncode == 0, r.stderr
    raw = out.read_bytes()
    n = int(np.frombuffer(raw[:4], np.uint32)[0])
    off = 4
    pos = np.frombuffer(raw[off:off + 16 * n], np.float32).reshape(n, 4)
    off += 16 * n
    vel = np.frombuffer(raw[off:off + 16 * n], np.float32).reshape(n, 4)
    off += 16 * n
    dens = np.frombuffer(raw[off:off + 4 * n], np.float32)
    off += 4 * n
    mass, ke, vmax, t = np.frombuffer(raw[off:off + 32], np.float64)

    p = sph.params_from_scene(s)
    x0, v0 = sph.scene_lattice(nx, ny, nz, origin, spacing, 0.1, 0.05, (0.0, 0.0, 0.0), seed, 0)
    with sph.SphSystem(p) as c:
        c.reset(x0, v0)
        c.update(steps)
        assert np.array_equal(pos, c.get_array(sph.SPH_ARRAY_POSITION))
        assert np.array_equal(vel, c.get_array(sph.SPH_ARRAY_VELOCITY))
        assert np.array_equal(dens, c.get_array(sph.SPH_ARRAY_DENSITY))
        st = c.stats()
    assert mass == st["mass"] and ke == st["kinetic_energy"] and abs(t - steps * dt) < 1e-9
    o = oracle.run_f32(p, x0, v0, steps)
    assert np.array_equal(pos, o["pos"]) and np.array_equal(dens, o["dens"])
