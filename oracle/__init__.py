"""CPU oracle of the WCSPH step - TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may
import this package, and only as the checker (or the timed CPU baseline);
the product library never calls it.

ctypes binding of oracle/liboracle.so (built from oracle/sph_oracle.cpp by
oracle/Makefile).  See sph_oracle.cpp for the reference file:line each
function restates.

Parity pinning status (what checks each oracle function; all in tests/):
  integrate_f32 ... closed forms: free particle, one wall band at a time,
                    hard clamp (test_oracle_closed_forms.py)
  hash_f32 ........ integer closed form + brute-force cell bounds
  step_f32 sort ... std::stable_sort is the textbook routine; checked against
                    numpy's stable argsort and permutation/sortedness invariants
  cell ranges ..... brute force over the sorted keys, sum(end-start) == N
  density ......... O(N^2) brute force (FP64), one-pair closed form, lattice
                    sum closed form, EOS closed forms
  force ........... two-particle closed form, antisymmetry (sum of m*dv == 0),
                    Galilean invariance, clamp cases (min_dist, min_dens)
  step_f64 ........ FP64 twin, pinned by the same closed forms at float64
  neighbour_counts  brute-force pair counts of the same predicates
                    (test_oracle_closed_forms.py), and cap truncation on a
                    crowded cell counted by hand
  grid_of ......... numpy's stable lexsort of the hashes; ranges against a
                    brute-force scan of the sorted keys
The reference ships no golden vectors or fixtures for this path, so there
are none under tests/golden.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

_DIR = Path(__file__).resolve().parent
# PIBITI_ORACLE_LIB names another build of the same source (the sanitizer
# build of `make -C oracle asan`).
LIB_PATH = Path(os.environ.get("PIBITI_ORACLE_LIB", str(_DIR / "liboracle.so")))
_lib = None

_fp = C.POINTER(C.c_float)
_dp = C.POINTER(C.c_double)
_up = C.POINTER(C.c_uint32)


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(_DIR)], check=True)
    return LIB_PATH


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        h = C.CDLL(str(LIB_PATH))
        h.oracle_num_threads.restype = C.c_int
        h.oracle_set_num_threads.argtypes = [C.c_int]
        h.oracle_step_f32.restype = C.c_int
        h.oracle_step_f32.argtypes = [C.c_void_p, C.c_int, C.c_uint32, _fp, _fp, _fp, _fp, _fp,
                                      _up, _up, _up, _up, _up]
        h.oracle_run_f32.restype = C.c_int
        h.oracle_run_f32.argtypes = [C.c_void_p, C.c_int, C.c_uint32, C.c_uint32, _fp, _fp, _fp, _fp]
        h.oracle_run_f32_colour.restype = C.c_int
        h.oracle_run_f32_colour.argtypes = [C.c_void_p, C.c_int, C.c_uint32, C.c_uint32, _fp, _fp, _fp, _fp]
        h.oracle_step_f32_ex.restype = C.c_int
        h.oracle_step_f32_ex.argtypes = [C.c_void_p, C.c_int, C.c_uint32, _fp, _fp, _fp, _fp, _fp, _fp,
                                         _up, _up, _up, _up, _up]
        h.oracle_step_f64.restype = C.c_int
        h.oracle_step_f64.argtypes = [C.c_void_p, C.c_uint32, _fp, _fp, C.c_int, _dp, _dp, _dp, _dp,
                                      _dp, _dp, _dp, _dp]
        h.oracle_integrate_f32.argtypes = [C.c_void_p, C.c_uint32, _fp, _fp]
        h.oracle_hash_f32.restype = C.c_uint32
        h.oracle_hash_f32.argtypes = [C.c_void_p, C.c_uint32, _fp, _up]
        h.oracle_density_f32.argtypes = [C.c_void_p, C.c_int, C.c_uint32, _fp, _up, _up, _fp, _fp]
        h.oracle_density_bruteforce_f64.argtypes = [C.c_void_p, C.c_uint32, _fp, _dp, _up]
        h.oracle_neighbour_counts_f32.argtypes = [C.c_void_p, C.c_int, C.c_uint32, _fp, _up, _up, _up]
        _lib = h
    return _lib


def _f(a):
    return a.ctypes.data_as(_fp)


def _d(a):
    return a.ctypes.data_as(_dp)


def _u(a):
    return a.ctypes.data_as(_up)


def _pp(params):
    return C.cast(C.pointer(params), C.c_void_p)


def set_threads(n: int) -> None:
    lib().oracle_set_num_threads(int(n))


def num_threads() -> int:
    return int(lib().oracle_num_threads())


def step_f32(params, pos4, vel4, prec: int = 0) -> dict:
    """One FP32 step (reference order or Kahan).  Inputs are not modified."""
    n = int(params.num_particles)
    nc = int(params.num_cells)
    pos = np.array(pos4, np.float32, copy=True).reshape(n, 4)
    vel = np.array(vel4, np.float32, copy=True).reshape(n, 4)
    out = dict(
        dens=np.empty(n, np.float32), pres=np.empty(n, np.float32), dvel=np.empty((n, 4), np.float32),
        vel_int=np.empty((n, 4), np.float32),
        sorted_hash=np.empty(n, np.uint32), perm=np.empty(n, np.uint32),
        cell_start=np.empty(nc, np.uint32), cell_end=np.empty(nc, np.uint32),
        counters=np.zeros(4, np.uint32))
    st = lib().oracle_step_f32_ex(_pp(params), prec, n, _f(pos), _f(vel), _f(out["vel_int"]),
                                  _f(out["dens"]), _f(out["pres"]), _f(out["dvel"]),
                                  _u(out["sorted_hash"]), _u(out["perm"]), _u(out["cell_start"]),
                                  _u(out["cell_end"]), _u(out["counters"]))
    assert st == 0
    out["pos"] = pos
    out["vel"] = vel
    return out


def run_f32(params, pos4, vel4, nsteps: int, prec: int = 0) -> dict:
    n = int(params.num_particles)
    pos = np.array(pos4, np.float32, copy=True).reshape(n, 4)
    vel = np.array(vel4, np.float32, copy=True).reshape(n, 4)
    dens = np.empty(n, np.float32)
    pres = np.empty(n, np.float32)
    st = lib().oracle_run_f32(_pp(params), prec, n, nsteps, _f(pos), _f(vel), _f(dens), _f(pres))
    assert st == 0
    return dict(pos=pos, vel=vel, dens=dens, pres=pres)


def step_f64(params, pos4, vel4, skip_integrate: bool = False) -> dict:
    """One FP64 step from float inputs.  skip_integrate: the inputs already are
    the post-integrate state (compare density/force sums on identical inputs)."""
    n = int(params.num_particles)
    pin = np.ascontiguousarray(pos4, np.float32).reshape(n, 4)
    vin = np.ascontiguousarray(vel4, np.float32).reshape(n, 4)
    out = dict(pos=np.empty((n, 4)), vel=np.empty((n, 4)), dens=np.empty(n), pres=np.empty(n),
               dvel=np.empty((n, 4)), dvel_scale=np.empty(n), dens_scale=np.empty(n),
               pres_scale=np.empty(n))
    st = lib().oracle_step_f64(_pp(params), n, _f(pin), _f(vin), int(skip_integrate), _d(out["pos"]),
                               _d(out["vel"]), _d(out["dens"]), _d(out["pres"]), _d(out["dvel"]),
                               _d(out["dvel_scale"]), _d(out["dens_scale"]), _d(out["pres_scale"]))
    assert st == 0
    return out


def kernels_f64_on(step32: dict, params) -> dict:
    """Density/pressure/force in FP64 on the exact post-integrate FP32 state of
    an FP32 step (its positions and integrated velocities)."""
    return step_f64(params, step32["pos"], step32["vel_int"], skip_integrate=True)


def integrate_f32(params, pos4, vel4):
    n = len(pos4)
    pos = np.array(pos4, np.float32, copy=True).reshape(n, 4)
    vel = np.array(vel4, np.float32, copy=True).reshape(n, 4)
    lib().oracle_integrate_f32(_pp(params), n, _f(pos), _f(vel))
    return pos, vel


def hash_f32(params, pos4):
    pos = np.ascontiguousarray(pos4, np.float32).reshape(-1, 4)
    h = np.empty(len(pos), np.uint32)
    oob = lib().oracle_hash_f32(_pp(params), len(pos), _f(pos), _u(h))
    return h, int(oob)


def density_f32(params, spos4, cell_start, cell_end, prec: int = 0):
    sp = np.ascontiguousarray(spos4, np.float32).reshape(-1, 4)
    n = len(sp)
    cs = np.ascontiguousarray(cell_start, np.uint32)
    ce = np.ascontiguousarray(cell_end, np.uint32)
    dens = np.empty(n, np.float32)
    pres = np.empty(n, np.float32)
    lib().oracle_density_f32(_pp(params), prec, n, _f(sp), _u(cs), _u(ce), _f(dens), _f(pres))
    return dens, pres


def neighbour_counts_f32(params, spos4, cell_start, cell_end, pred: int = 0):
    """Per sorted slot: candidates of the capped 27-cell walk that pass the
    density predicate (pred 0: r2 < h2) or the force predicate (pred 1)."""
    sp = np.ascontiguousarray(spos4, np.float32).reshape(-1, 4)
    n = len(sp)
    out = np.empty(n, np.uint32)
    lib().oracle_neighbour_counts_f32(_pp(params), int(pred), n, _f(sp),
                                      _u(np.ascontiguousarray(cell_start, np.uint32)),
                                      _u(np.ascontiguousarray(cell_end, np.uint32)), _u(out))
    return out


def grid_of(params, pos4):
    """The stable (cell, index) order of positions given in original order:
    keys, permutation (sorted slot -> original index), sorted positions and
    the cell ranges (empty cells 0xFFFFFFFF)."""
    pos = np.ascontiguousarray(pos4, np.float32).reshape(-1, 4)
    h, _ = hash_f32(params, pos)
    perm = np.lexsort((np.arange(len(h)), h)).astype(np.uint32)
    sh = h[perm]
    nc = int(params.num_cells)
    counts = np.bincount(sh, minlength=nc)
    start = np.full(nc, 0xFFFFFFFF, np.uint32)
    end = np.full(nc, 0xFFFFFFFF, np.uint32)
    occ = np.flatnonzero(counts)
    first = np.searchsorted(sh, occ)
    start[occ] = first
    end[occ] = first + counts[occ]
    return dict(hash=sh, perm=perm, spos=pos[perm], cell_start=start, cell_end=end)


def density_bruteforce_f64(params, pos4):
    pos = np.ascontiguousarray(pos4, np.float32).reshape(-1, 4)
    n = len(pos)
    dens = np.empty(n)
    nnb = np.empty(n, np.uint32)
    lib().oracle_density_bruteforce_f64(_pp(params), n, _f(pos), _d(dens), _u(nnb))
    return dens, nnb


def run_f32_colour(params, pos4, vel4, nsteps: int, dye=None, prec: int = 0) -> dict:
    """nsteps FP32 steps with the colour output; dye (float per particle,
    original order) starts at `dye` (zeros by default)."""
    n = int(params.num_particles)
    pos = np.array(pos4, np.float32, copy=True).reshape(n, 4)
    vel = np.array(vel4, np.float32, copy=True).reshape(n, 4)
    d = np.zeros(n, np.float32) if dye is None else np.array(dye, np.float32, copy=True)
    col = np.zeros((n, 4), np.float32)
    st = lib().oracle_run_f32_colour(_pp(params), prec, n, nsteps, _f(pos), _f(vel), _f(d), _f(col))
    assert st == 0
    return dict(pos=pos, vel=vel, dye=d, colour=col)
