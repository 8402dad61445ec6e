"""MI355X-native WCSPH step: C ABI over hand-written HIP kernels for gfx950.

`pibiti_amd.sph` binds the in-tree library (libpibiti_sph.so) and mirrors the
reference's cSPH host class; `pibiti_amd.scenes` builds the seeded scenes.
"""
__all__ = ["sph", "scenes"]
