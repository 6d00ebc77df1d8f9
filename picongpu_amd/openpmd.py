"""openPMD output and checkpoint/restart of a simulation context.

The reference writes meshes and particle species through openPMD-api
(include/picongpu/plugins/openPMD/openPMDWriter.x.cpp:1150-1184, 1376-1560,
WriteSpecies.hpp:240-330, writer/ParticleAttribute.hpp:60-160,
plugins/common/openPMDWriteMeta.hpp:110-245) and restarts from the same files
(restart/LoadSpecies.hpp:54-80, restart/RestartFieldLoader.hpp:205-208,
plugins/Checkpoint.x.cpp:45). openPMD-api, HDF5 and ADIOS2 are not available
here, so the series is written with the layout of openPMD-api's JSON backend
(openPMD standard 1.1.0, file-based iteration encoding, one file per
iteration): every group is an object with an "attributes" object of
{"datatype", "value"} entries; a dataset carries "datatype", "extent" and
"data" (nested lists in C order); a constant record component carries the
standard's "value" and "shape" attributes instead of data.

Records, as the reference names and scales them:
- meshes E, B (and J when asked for): components x, y, z of shape (nz, ny, nx)
  with axisLabels ("z", "y", "x"), the Yee "position" of each component in
  (z, y, x) order, gridSpacing and gridUnitSI, unitSI eField / bField /
  charge/(time length^2);
- with the PML absorber, its convolution fields "Convolutional PML E" and
  "Convolutional PML B" (absorber/pml/Field.hpp:320-410, checkpointed by the
  reference): components xy, xz, yx, yz, zx, zy at E's and B's Yee positions
  in E's and B's units. The reference stores only the layer cells; here the
  whole grid (zero outside the layers);
- particles <species>: "position" (in-cell, float32, unitSI = cell size),
  "positionOffset" (global cell index, int32, unitSI = cell size),
  "momentum" (macro-particle momentum, unitSI = unit mass * unit speed,
  macroWeighted, weightingPower 1), "weighting", the constant records "mass"
  and "charge" (per unit weighting, in the unit mass/charge) and
  "particlePatches" of the one patch.

Particles are written supercell by supercell in slot order, and the loader
bins them back stably, so a restart restores every supercell's slot order;
floats are written as the shortest decimal that reads back to the same FP32
value, so the restart resumes bit for bit. The field step counter is the
iteration; the current's fixed-point scale is stored as the attribute
"jScaleExp" of the iteration (this library's own, outside the standard).
"""
from __future__ import annotations

import datetime
import json
import os

import numpy as np

from . import abi

OPENPMD_VERSION = "1.1.0"
FILE_PATTERN = "simData_%T.json"

# L, M, T, I, theta, N, J (openPMD unitDimension order)
_DIM = {
    "E": [1.0, 1.0, -3.0, -1.0, 0.0, 0.0, 0.0],
    "B": [0.0, 1.0, -2.0, -1.0, 0.0, 0.0, 0.0],
    "J": [-2.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0],
    "position": [1.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0],
    "momentum": [1.0, 1.0, -1.0, 0.0, 0.0, 0.0, 0.0],
    "weighting": [0.0] * 7,
    "mass": [0.0, 1.0, 0.0, 0.0, 0.0, 0.0, 0.0],
    "charge": [0.0, 0.0, 1.0, 1.0, 0.0, 0.0, 0.0],
}
# Yee positions of the components (fields/YeeCell.hpp:69-123), x, y, z order
_YEE = {"E": [(0.5, 0.0, 0.0), (0.0, 0.5, 0.0), (0.0, 0.0, 0.5)],
        "B": [(0.0, 0.5, 0.5), (0.5, 0.0, 0.5), (0.5, 0.5, 0.0)],
        "J": [(0.5, 0.0, 0.0), (0.0, 0.5, 0.0), (0.0, 0.0, 0.5)]}
_FIELD_ID = {"E": abi.FIELD_E, "B": abi.FIELD_B, "J": abi.FIELD_J}
PML_E, PML_B = "Convolutional PML E", "Convolutional PML B"
_PML_ID = {PML_E: abi.FIELD_PML_E, PML_B: abi.FIELD_PML_B}
_PML_OF = {PML_E: "E", PML_B: "B"}  # the field whose units and positions each takes
_SHAPE_NAME = {abi.SHAPE_CIC: "CIC", abi.SHAPE_TSC: "TSC", abi.SHAPE_PQS: "PQS"}
_PUSHER_NAME = {abi.PUSHER_BORIS: "Boris", abi.PUSHER_VAY: "Vay", abi.PUSHER_HIGUERA_CARY: "HigueraCary"}


def _attr(datatype, value):
    return {"datatype": datatype, "value": value}


def _f32_list(a):
    """Nested lists of Python floats; each FP32 value becomes the double of the
    same value, whose shortest repr reads back to it exactly."""
    return np.asarray(a, np.float32).astype(np.float64).tolist()


def file_name(directory: str, iteration: int) -> str:
    return os.path.join(directory, FILE_PATTERN.replace("%T", str(int(iteration))))


def units(k) -> dict:
    """SI units of the context's constants (unitless/simulation.unitless:400-470)."""
    length, time, mass, charge = k.unit_length, k.unit_time, k.unit_mass, k.unit_charge
    speed = length / time
    return {"length": length, "time": time, "mass": mass, "charge": charge, "speed": speed,
            "energy": mass * length * length / (time * time),
            "efield": 1.0 / (time * time / mass / length * charge), "bfield": mass / (time * charge),
            "current": charge / time / (length * length)}


def snapshot(sim: "abi.Simulation", species_names, with_current=False) -> dict:
    """Host copy of what an iteration file holds: fields, particles (supercell
    by supercell in slot order) and constants."""
    k = sim.constants()
    snap = {"grid": tuple(sim.grid), "k": k, "step": sim.get_step(), "jexp": sim.j_scale_exp(),
            "version": int(sim.lib.pic_abi_version()), "fields": {}, "species": []}
    cfg = sim.cfg
    sides = []
    for d in range(3):
        kind = "periodic" if int(cfg.boundary[d]) == abi.BOUNDARY_PERIODIC else "open"
        sides += [kind, kind]
    snap["config"] = {"solver": "Lehe" if int(cfg.solver) == abi.SOLVER_LEHE else "Yee", "field_boundary": sides,
                      "current_smoothing": "Binomial" if int(cfg.current_interpolation) == abi.CURRENT_BINOMIAL
                      else "none"}
    for nm in ("E", "B") + (("J",) if with_current else ()):
        snap["fields"][nm] = sim.read_field(_FIELD_ID[nm])
    if int(cfg.absorber) == abi.ABSORBER_PML and any(int(b) == abi.BOUNDARY_OPEN for b in cfg.boundary):
        for nm, which in _PML_ID.items():
            snap["fields"][nm] = sim.read_field(which)
    for sp, name in enumerate(species_names):
        if name is None:
            continue
        cell, pos, mom, w = sim.read_particles(sp)
        d = sim.cfg.species[sp]
        desc = {"shape": int(d.shape), "pusher": int(d.pusher),
                "current": "EmZ" if int(d.current_solver) == abi.CURRENT_SOLVER_EMZ else "Esirkepov",
                "mass": float(k.species_mass[sp]), "charge": float(k.species_charge[sp])}
        snap["species"].append({"name": name, "index": sp, "cell": cell, "pos": pos, "mom": mom, "w": w,
                                "desc": desc})
    return snap


def iteration_dict(snap: dict, config: dict) -> dict:
    """The iteration group of a snapshot in the JSON backend's layout.
    config: solver name, boundary names and the species' descriptions."""
    k = snap["k"]
    u = units(k)
    nx, ny, nz = snap["grid"]
    cell = [float(np.float32(k.cell[d])) for d in range(3)]
    step = int(snap["step"])
    dt = float(np.float32(k.dt))
    it_attrs = {
        "dt": _attr("FLOAT", dt),
        "time": _attr("FLOAT", float(np.float32(np.float32(step) * np.float32(k.dt)))),
        "timeUnitSI": _attr("DOUBLE", u["time"]),
        "iteration": _attr("UINT", step),
        "cell_width": _attr("FLOAT", cell[0]), "cell_height": _attr("FLOAT", cell[1]),
        "cell_depth": _attr("FLOAT", cell[2]),
        "unit_energy": _attr("DOUBLE", u["energy"]), "unit_length": _attr("DOUBLE", u["length"]),
        "unit_speed": _attr("DOUBLE", u["speed"]), "unit_time": _attr("DOUBLE", u["time"]),
        "unit_mass": _attr("DOUBLE", u["mass"]), "unit_charge": _attr("DOUBLE", u["charge"]),
        "unit_efield": _attr("DOUBLE", u["efield"]), "unit_bfield": _attr("DOUBLE", u["bfield"]),
        "mue0": _attr("FLOAT", float(np.float32(k.mue0))), "eps0": _attr("FLOAT", float(np.float32(k.eps0))),
        "jScaleExp": _attr("INT", int(snap["jexp"])),
    }
    meshes = {"attributes": {"fieldSolver": _attr("STRING", config.get("solver", "Yee")),
                             "fieldBoundary": _attr("VEC_STRING", config.get("field_boundary", ["periodic"] * 6)),
                             "currentSmoothing": _attr("STRING", config.get("current_smoothing", "none")),
                             "chargeCorrection": _attr("STRING", "none")}}
    unit_of = {"E": u["efield"], "B": u["bfield"], "J": u["current"]}
    for nm, data in snap["fields"].items():
        base = _PML_OF.get(nm, nm)
        mesh = {"attributes": {
            "axisLabels": _attr("VEC_STRING", ["z", "y", "x"]),
            "dataOrder": _attr("STRING", "C"),
            "geometry": _attr("STRING", "cartesian"),
            "gridGlobalOffset": _attr("VEC_DOUBLE", [0.0, 0.0, 0.0]),
            "gridSpacing": _attr("VEC_FLOAT", [cell[2], cell[1], cell[0]]),
            "gridUnitSI": _attr("DOUBLE", u["length"]),
            "timeOffset": _attr("FLOAT", 0.0),
            "unitDimension": _attr("ARR_DBL_7", _DIM[base]),
            "fieldSmoothing": _attr("STRING", "none")}}
        comps = abi.PML_COMPONENTS if nm in _PML_ID else "xyz"
        for c, comp in enumerate(comps):
            # a psi component takes the Yee position of its field's component along
            # its first letter (the component the convolution enters)
            px, py, pz = _YEE[base]["xyz".index(comp[0])]
            mesh[comp] = {"attributes": {"position": _attr("VEC_FLOAT", [pz, py, px]),
                                         "unitSI": _attr("DOUBLE", unit_of[base])},
                          "datatype": "FLOAT", "extent": [nz, ny, nx],
                          "data": _f32_list(np.asarray(data)[c].reshape(nz, ny, nx))}
        meshes[nm] = mesh
    particles = {}
    for s in snap["species"]:
        n = int(s["w"].shape[0])
        d = s["desc"]
        rec = {"attributes": {"particleShape": _attr("FLOAT", float({abi.SHAPE_CIC: 1.0, abi.SHAPE_TSC: 2.0,
                                                                    abi.SHAPE_PQS: 3.0}[d["shape"]])),
                              "currentDeposition": _attr("STRING", d["current"]),
                              "particlePush": _attr("STRING", _PUSHER_NAME.get(d["pusher"], "None")),
                              "particleInterpolation": _attr("STRING", "trilinear"),
                              "particleSmoothing": _attr("STRING", "none")}}

        def vector_record(name, arr, datatype, unit_si, macro, power):
            r = {"attributes": {"unitDimension": _attr("ARR_DBL_7", _DIM[name if name in _DIM else "position"]),
                                "macroWeighted": _attr("UINT", int(macro)),
                                "weightingPower": _attr("DOUBLE", float(power)), "timeOffset": _attr("FLOAT", 0.0)}}
            for c, comp in enumerate("xyz"):
                col = arr[:, c]
                data = _f32_list(col) if datatype == "FLOAT" else np.asarray(col, np.int64).tolist()
                r[comp] = {"attributes": {"unitSI": _attr("DOUBLE", unit_si[c])}, "datatype": datatype,
                           "extent": [n], "data": data}
            return r

        cell_si = [cell[c] * u["length"] for c in range(3)]
        rec["position"] = vector_record("position", s["pos"], "FLOAT", cell_si, False, 0.0)
        rec["positionOffset"] = vector_record("position", s["cell"], "INT", cell_si, False, 0.0)
        rec["momentum"] = vector_record("momentum", s["mom"], "FLOAT", [u["mass"] * u["speed"]] * 3, True, 1.0)
        rec["weighting"] = {"attributes": {"unitDimension": _attr("ARR_DBL_7", _DIM["weighting"]),
                                           "macroWeighted": _attr("UINT", 1), "weightingPower": _attr("DOUBLE", 1.0),
                                           "timeOffset": _attr("FLOAT", 0.0), "unitSI": _attr("DOUBLE", 1.0)},
                            "datatype": "FLOAT", "extent": [n], "data": _f32_list(s["w"])}
        for name, value, unit in (("mass", d["mass"], u["mass"]), ("charge", d["charge"], u["charge"])):
            rec[name] = {"attributes": {"unitDimension": _attr("ARR_DBL_7", _DIM[name]),
                                        "macroWeighted": _attr("UINT", 0), "weightingPower": _attr("DOUBLE", 1.0),
                                        "timeOffset": _attr("DOUBLE", 0.0), "unitSI": _attr("DOUBLE", unit),
                                        "value": _attr("DOUBLE", float(value)), "shape": _attr("VEC_ULONG", [n])}}
        rec["particlePatches"] = {
            "numParticles": {"datatype": "ULONG", "extent": [1], "data": [n]},
            "numParticlesOffset": {"datatype": "ULONG", "extent": [1], "data": [0]},
            "offset": {c: {"datatype": "ULONG", "extent": [1], "data": [0]} for c in "xyz"},
            "extent": {c: {"datatype": "ULONG", "extent": [1], "data": [g]} for c, g in zip("xyz", (nx, ny, nz))}}
        particles[s["name"]] = rec
    return {"attributes": it_attrs, "meshes": meshes, "particles": particles}


def series_dict(snap: dict, config: dict) -> dict:
    return {"attributes": {
        "basePath": _attr("STRING", "/data/%T/"),
        "date": _attr("STRING", datetime.datetime.now(datetime.timezone.utc).strftime("%Y-%m-%d %H:%M:%S %z")),
        "iterationEncoding": _attr("STRING", "fileBased"),
        "iterationFormat": _attr("STRING", FILE_PATTERN),
        "meshesPath": _attr("STRING", "meshes/"),
        "openPMD": _attr("STRING", OPENPMD_VERSION),
        "openPMDextension": _attr("UINT", 0),
        "particlesPath": _attr("STRING", "particles/"),
        "software": _attr("STRING", "picongpu_amd"),
        "softwareVersion": _attr("STRING", str(snap.get("version", 1))),
    }, "data": {str(int(snap["step"])): iteration_dict(snap, config)}}


def write_snapshot(snap: dict, directory: str) -> str:
    os.makedirs(directory, exist_ok=True)
    path = file_name(directory, snap["step"])
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(series_dict(snap, snap.get("config", {})), f, allow_nan=False)
    os.replace(tmp, path)
    return path


def read_iteration(path: str) -> tuple[int, dict]:
    """(iteration, iteration group) of a file written by write_snapshot."""
    with open(path) as f:
        series = json.load(f)
    attrs = series["attributes"]
    if attrs["openPMD"]["value"] != OPENPMD_VERSION or attrs["iterationEncoding"]["value"] != "fileBased":
        raise ValueError(f"{path}: not a file-based openPMD {OPENPMD_VERSION} series")
    (key, it), = series["data"].items()
    return int(key), it


def mesh_array(it: dict, name: str) -> np.ndarray:
    """(3, nx*ny*nz) float32 planes, x fastest, of mesh `name` (6 planes in
    NodeValues order for a PML mesh)."""
    m = it["meshes"][name]
    comps = abi.PML_COMPONENTS if name in _PML_ID else "xyz"
    return np.stack([np.asarray(m[c]["data"], np.float32).reshape(-1) for c in comps])


def species_arrays(it: dict, name: str):
    """(cell (n,3) int32, pos (n,3) float32, mom (n,3) float32, w (n,) float32) of a species."""
    r = it["particles"][name]
    cell = np.stack([np.asarray(r["positionOffset"][c]["data"], np.int32) for c in "xyz"], 1)
    pos = np.stack([np.asarray(r["position"][c]["data"], np.float32) for c in "xyz"], 1)
    mom = np.stack([np.asarray(r["momentum"][c]["data"], np.float32) for c in "xyz"], 1)
    w = np.asarray(r["weighting"]["data"], np.float32)
    return cell, pos, mom, w


def write(sim: "abi.Simulation", directory: str, species_names, with_current: bool = False) -> str:
    """Write the context's current state as iteration <field step counter>
    (file simData_<iteration>.json in `directory`); species_names[i] names
    species i (None: not written)."""
    return write_snapshot(snapshot(sim, species_names, with_current), directory)


def restart(sim: "abi.Simulation", path: str, species_names) -> int:
    """Load E, B (the PML fields too, when the file has them) and every named
    species of an iteration file into a context created from the same
    configuration (with j_scale_exp set from restart_config), and set the
    field step counter; returns the iteration."""
    step, it = read_iteration(path)
    for nm in ("E", "B"):
        sim.load_field(_FIELD_ID[nm], mesh_array(it, nm))
    if "J" in it["meshes"]:
        sim.load_field(_FIELD_ID["J"], mesh_array(it, "J"))
    for nm, which in _PML_ID.items():
        if nm in it["meshes"]:
            sim.load_field(which, mesh_array(it, nm))
    for sp, name in enumerate(species_names):
        if name is None:
            continue
        cell, pos, mom, w = species_arrays(it, name)
        sim.load_particles(sp, cell, pos, mom, w)
    sim.set_step(step)
    jexp = int(it["attributes"]["jScaleExp"]["value"])
    if sim.j_scale_exp() != jexp:
        raise ValueError(f"the context's current scale 2^{sim.j_scale_exp()} differs from the file's 2^{jexp}: "
                         "create it with restart_config()")
    return step


def restart_config(cfg, path: str):
    """A copy of pic_config `cfg` whose current fixed-point scale is the one
    stored in the iteration file, so the restarted steps deposit the same J."""
    _, it = read_iteration(path)
    cfg.j_scale_exp = int(it["attributes"]["jScaleExp"]["value"])
    return cfg
