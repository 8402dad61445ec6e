"""Scene-file oracle.  TEST INFRASTRUCTURE ONLY: the parity tests compare
the library's scene-file front end with it; the product never imports it.

Two parts:
  * ``tinyxml_scenes(path)`` runs the reference's own XML reader (TinyXML as
    vendored in the reference, compiled by ``make -C oracle ref`` into
    ``oracle/_ref/xml_dump``) and returns the scenes it finds, walked as
    cSPH::LoadScenes walks them (source/SPH/SPH_Scenes.cpp:56-71);
  * ``scene_values(scene)`` restates Scene::InitDefault, Scene::_FromXML and
    the non-kernel parts of Scene::_UpdatePar/_UpdateGrid
    (source/SPH/Scene.cpp:7-66, :76-102; source/SPH/Scene_Load.cpp:11-111)
    in float32 arithmetic, statement by statement, over those attribute
    lists.
"""
from __future__ import annotations

import json
import math
import re
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
XML_DUMP = HERE / "_ref" / "xml_dump"
REFERENCE_SRC = Path("/root/reference/source/external")

f32 = np.float32
PI = f32(3.141592654)  # source/CUDA/Params.cuh:18


def build_reference_reader() -> Path | None:
    """Compile the reference's TinyXML with the scene lister (needs the
    reference tree); None when it is not available."""
    if XML_DUMP.exists():
        return XML_DUMP
    if not (REFERENCE_SRC / "tinyxml.cpp").exists():
        return None
    r = subprocess.run(["make", "-s", "-C", str(HERE), "ref"], capture_output=True, text=True)
    return XML_DUMP if r.returncode == 0 and XML_DUMP.exists() else None


def tinyxml_scenes(path) -> tuple[dict, list[dict]]:
    exe = build_reference_reader()
    if exe is None:
        raise FileNotFoundError("reference XML reader not built")
    lines = subprocess.run([str(exe), str(path)], check=True, capture_output=True,
                           text=True, encoding="latin-1").stdout.splitlines()
    head = json.loads(lines[0])
    return head, [json.loads(x) for x in lines[1:]]


# ---------------------------------------------------------------- converters
_FLOAT = re.compile(r"\s*[+-]?(?:\d+\.?\d*|\.\d+)(?:[eE][+-]?\d+)?")
_INT = re.compile(r"\s*([+-]?)(0[xX][0-9a-fA-F]+|0[0-7]*|[1-9]\d*)")


def to_float(a: str) -> np.float32:
    """toFloat: (float)atof(str) (source/pch/header.h:155)."""
    m = _FLOAT.match(a)
    return f32(float(m.group(0))) if m else f32(0.0)


def to_int(a: str) -> int:
    """toInt: (int)strtol(str, NULL, 0) (source/pch/header.h:156)."""
    m = _INT.match(a)
    if not m:
        return 0
    v = int(m.group(2), 0) if not re.fullmatch(r"0[0-7]+", m.group(2)) else int(m.group(2), 8)
    return -v if m.group(1) == "-" else v


def to_vec3(a: str) -> list:
    """toVec3: sscanf(str, "%f %f %f") (source/pch/header.h:148-153); the
    components it does not read are 0 here."""
    out = []
    rest = a
    for _ in range(3):
        m = _FLOAT.match(rest)
        if not m:
            break
        out.append(f32(float(m.group(0))))
        rest = rest[m.end():]
    return out + [f32(0.0)] * (3 - len(out))


# ------------------------------------------------------------------- scene
def scene_values(sc: dict, index: int) -> dict:
    """Every value the library's sph_xml_scene returns, for one scene of
    tinyxml_scenes(): {field: value} with sph_scene names for simulation
    values and sph_scene_info names (prefixed "info.") for the rest."""
    A = dict(sc["attrs"])
    # Scene::InitDefault (source/SPH/Scene.cpp:7-66)
    v = dict(num_particles=7 * 8 * 1024, max_par_in_cell=16, time_step=f32(0.0026),
             global_damping=f32(1.0), gravity=[f32(0), f32(-9.81), f32(0)],
             particle_r=f32(0.004), min_dist=f32(1.0), h=f32(0.01), rest_density=f32(1000),
             min_dens=f32(1.0), stiffness=f32(3.0), viscosity=f32(0.5),
             dist_bnd_soft=f32(8), dist_bnd_hard=f32(1), bnd_stiff=f32(30000), bnd_damp=f32(256),
             bnd_damp_c=f32(60), bnd_type=0, bnd_eff_z=0,
             coll_r=f32(0.05), coll_spring=f32(80), coll_damping=f32(0.02), coll_shear=f32(0.1),
             pump_ang_out=f32(0.7), pump_rad_in=f32(0.4), exit_vel=f32(0.03), exit_dist=f32(0.04))
    w = [f32(0.2), f32(0.25), f32(0.2)]
    v["world_min"] = [-x for x in w]
    v["world_max"] = list(w)
    init_min, init_max = [-x for x in w], list(w)
    cell = f32(v["particle_r"] * f32(2.0))
    spacing = f32(1.38)
    drop_r, rain, init_type, init_last = f32(5), 0, 0, 1
    h_close = f32(-0.03)
    s = [f32(0.90), f32(0.90), f32(1.198), f32(0.7782), f32(0.8696), f32(0.843)]
    rot_type, r_angle, r_twist, r2_angle, r_vel = 0, f32(0), f32(0.04), f32(0), f32(0)
    coll_pos = [f32(-77.0), f32(0.0), f32(0.0)]
    hmap = 0
    title, chapter = "No name", False
    cam_pos, cam_rot = [f32(0), f32(0.04), f32(-0.44)], [f32(5), f32(0), f32(0)]
    emit = [dict(size=0, size2=0, pos=[f32(0)] * 3, rot=[f32(0)] * 3, vel=f32(1.5)) for _ in range(4)]
    acc = [dict(pos=[f32(0)] * 3, size=[f32(0.005), f32(0.010), f32(0.005)],
                acc=[f32(0), f32(10), f32(0)], type=0) for _ in range(4)]

    # Scene::_FromXML, in statement order (source/SPH/Scene_Load.cpp:22-113)
    if "World" in A:
        sw = to_vec3(A["World"])
        v["world_min"] = [x * f32(-0.5) for x in sw]
        v["world_max"] = [x * f32(0.5) for x in sw]
        init_min, init_max = [x * f32(-0.5) for x in sw], [x * f32(0.5) for x in sw]
    if "Init" in A:
        si = to_vec3(A["Init"])
        init_min, init_max = [x * f32(-0.5) for x in si], [x * f32(0.5) for x in si]
    if "WorldMin" in A:
        v["world_min"] = to_vec3(A["WorldMin"])
    if "InitMin" in A:
        init_min = to_vec3(A["InitMin"])
    if "WorldMax" in A:
        v["world_max"] = to_vec3(A["WorldMax"])
    if "InitMax" in A:
        init_max = to_vec3(A["InitMax"])
    if "InitType" in A:
        init_type = to_int(A["InitType"])
    if "InitLast" in A:
        init_last = to_int(A["InitLast"])
    if "CellSize" in A:
        cell = to_float(A["CellSize"])
    if "name" in A:
        title = A["name"]
    if "chapter" in A:
        chapter = True
    if "CamPos" in A:
        cam_pos = to_vec3(A["CamPos"])
    if "CamRot" in A:
        cam_rot = to_vec3(A["CamRot"])
    if "dropR" in A:
        drop_r = to_float(A["dropR"])
    if "rain" in A:
        rain = to_int(A["rain"])
    e0 = emit[0]
    if "EmitSize" in A:
        e0["size"] = to_int(A["EmitSize"])
    if "EmitPos" in A:
        e0["pos"] = to_vec3(A["EmitPos"])
    if "EmitVel" in A:
        e0["vel"] = to_float(A["EmitVel"])
    if "EmitSize2" in A:
        e0["size2"] = to_int(A["EmitSize2"])
    if "EmitRot" in A:
        e0["rot"] = to_vec3(A["EmitRot"])
    for i, ea in enumerate(sc["Emitter"][:4]):
        E = dict(ea)
        e = emit[i]
        if "size" in E:
            e["size"] = to_int(E["size"])
        if "pos" in E:
            e["pos"] = to_vec3(E["pos"])
        if "vel" in E:
            e["vel"] = to_float(E["vel"])
        if "size2" in E:
            e["size2"] = to_int(E["size2"])
        if "rot" in E:
            e["rot"] = to_vec3(E["rot"])
    for i, aa in enumerate(sc["Accel"][:4]):
        E = dict(aa)
        a = acc[i]
        if "type" in E:
            a["type"] = to_int(E["type"])
        for k in ("pos", "size", "acc"):
            if k in E:
                a[k] = to_vec3(E[k])
    simple_f = [("ColliderR", "coll_r"), ("spring", "coll_spring"), ("damping", "coll_damping"),
                ("shear", "coll_shear")]
    if "ColliderR" in A:
        v["coll_r"] = to_float(A["ColliderR"])
    if "ColliderPos" in A:
        coll_pos = to_vec3(A["ColliderPos"])
    for key, name in simple_f[1:]:
        if key in A:
            v[name] = to_float(A[key])
    if "Particles" in A:
        v["num_particles"] = to_int(A["Particles"]) & 0xFFFFFFFF
    if "ParticlesK" in A:
        v["num_particles"] = (to_int(A["ParticlesK"]) * 1024) & 0xFFFFFFFF
    if "maxParInCell" in A:
        v["max_par_in_cell"] = to_int(A["maxParInCell"]) & 0xFFFFFFFF
    for key, name in [("TimeStep", "time_step"), ("globalDamping", "global_damping")]:
        if key in A:
            v[name] = to_float(A[key])
    if "Gravity" in A:
        v["gravity"] = to_vec3(A["Gravity"])
    for key, name in [("particleR", "particle_r"), ("minDist", "min_dist"), ("particleH", "h")]:
        if key in A:
            v[name] = to_float(A[key])
    if "spacing" in A:
        spacing = to_float(A["spacing"])
    for key, name in [("RestDensity", "rest_density"), ("minDens", "min_dens"),
                      ("Stiffness", "stiffness"), ("Viscosity", "viscosity"),
                      ("distBndSoft", "dist_bnd_soft"), ("bndStiff", "bnd_stiff"),
                      ("distBndHard", "dist_bnd_hard"), ("bndDamp", "bnd_damp"),
                      ("bndDampC", "bnd_damp_c")]:
        if key in A:
            v[name] = to_float(A[key])
    if "bndType" in A:
        v["bnd_type"] = to_int(A["bndType"]) & 0xFFFFFFFF
    if "bndEffZ" in A:
        v["bnd_eff_z"] = to_int(A["bndEffZ"]) & 0xFFFFFFFF
    if "HmapType" in A:
        hmap = to_int(A["HmapType"])
    if "PumpAngOut" in A:
        v["pump_ang_out"] = to_float(A["PumpAngOut"])
    if "PumpHClose" in A:
        h_close = to_float(A["PumpHClose"])
    if "PumpRadIn" in A:
        v["pump_rad_in"] = to_float(A["PumpRadIn"])
    if "ExitVel" in A:
        v["exit_vel"] = to_float(A["ExitVel"])
    if "ExitDist" in A:
        v["exit_dist"] = to_float(A["ExitDist"])
    for k in range(6):
        if f"s{k + 1}" in A:
            s[k] = to_float(A[f"s{k + 1}"])
    rot = dict(rot_size=[0, 4, 12], rot_blades=4, rot_r=f32(3), rot_spc=f32(0.7), rot2_dist=f32(0),
               rot2_vel=f32(1), rot2_twist=f32(1))
    if "RotorSizes" in A:
        rot["rot_size"] = [int(x) for x in to_vec3(A["RotorSizes"])]  # make_int3 of floats
    if "RotorType" in A:
        rot_type = to_int(A["RotorType"])
    if "Rotor2Dist" in A:
        rot["rot2_dist"] = to_float(A["Rotor2Dist"])
    if "RotorVel" in A:
        r_vel = to_float(A["RotorVel"])
    if "Rotor2Vel" in A:
        rot["rot2_vel"] = to_float(A["Rotor2Vel"])
    if "RotorAngle" in A:
        r_angle = f32(f32(to_float(A["RotorAngle"]) * PI) / f32(180.0))
    if "Rotor2Angle" in A:
        r2_angle = f32(f32(to_float(A["Rotor2Angle"]) * PI) / f32(180.0))
    if "colParR" in A:
        rot["rot_r"] = to_float(A["colParR"])
    if "colParSpc" in A:
        rot["rot_spc"] = to_float(A["colParSpc"])
    if "RotorBlades" in A:
        rot["rot_blades"] = to_int(A["RotorBlades"])
    if "RotorTwist" in A:
        r_twist = to_float(A["RotorTwist"])
    if "WaveSpeed" in A:
        r_vel = to_float(A["WaveSpeed"])
        r_angle = f32(-PI / f32(2.0))
    if "WaveAmpl" in A:
        r_twist = to_float(A["WaveAmpl"])
    if "HSlope" in A:
        r2_angle = to_float(A["HSlope"])
    if "Hheight" in A:
        r2_angle = to_float(A["Hheight"])
    if "Hscale" in A:
        h_close = to_float(A["Hscale"])
    for k, key in enumerate(["HxFq", "HxOfs", "HzFq", "HzOfs", "HholeR"]):
        if key in A:
            s[k] = to_float(A[key])

    # what the library reports besides the raw values
    v["cell_size"] = cell
    v["h_close"] = h_close
    v["shape_s"] = s
    v["rot_angle"], v["rot_twist"], v["rot2_angle"] = r_angle, r_twist, r2_angle
    v["rot_vel"] = r_vel
    v["collider"] = 1
    # colours and dye (Scene.cpp:61-64): no attributes set them
    v.update(clr_type=4, hue=0, brightness=f32(0.3), contrast=f32(0.3), dye_type=0, dye_clear=2,
             dye_fade=f32(1.0), dye_pos=[f32(0), f32(-0.21), f32(0)], dye_size=[f32(0.008), f32(0.012), f32(0.008)])
    v["rot_type"], v["hmap_type"] = rot_type, hmap
    v.update(rot)
    v["coll_pos"] = coll_pos
    v["acc"] = acc
    r = v["particle_r"]
    b = f32(f32(v["dist_bnd_soft"] * r) - r)  # _UpdatePar then _UpdateGrid (Scene.cpp:90,100)
    info = {
        "title": title[:63], "chapter": int(chapter or index == 0),
        "init_min": [f32(x + b) for x in init_min], "init_max": [f32(x - b) for x in init_max],
        "init_type": init_type, "init_last": init_last, "spacing": f32(spacing * r),
        "drop_r": drop_r, "rain": rain, "emit": emit, "cam_pos": cam_pos, "cam_rot": cam_rot,
    }
    v.update({"info." + k: x for k, x in info.items()})
    return v


def startup_index(scenes: list[dict]) -> int:
    """The last scene with a default or def attribute (SPH_Scenes.cpp:66)."""
    d = 0
    for i, sc in enumerate(scenes):
        A = dict(sc["attrs"])
        if "default" in A or "def" in A:
            d = i
    return d
