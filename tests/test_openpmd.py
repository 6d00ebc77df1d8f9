"""openPMD output (JSON-backend layout of openPMD-api, file-based encoding):
the written records, attributes and units, and an exact round trip of every
value. Parity with openPMD-api itself is unpinned (the library is not
available here); the layout follows the openPMD standard 1.1.0 and the
records the reference writes (plugins/openPMD/openPMDWriter.x.cpp,
WriteSpecies.hpp, plugins/common/openPMDWriteMeta.hpp)."""
import json

import numpy as np
import pytest

from picongpu_amd import abi, openpmd, setups


def synthetic_snapshot(step=7):
    st = setups.thermal(grid=(16, 8, 4), ppc=2)
    k = abi.constants_from_config(st.config())
    rng = np.random.default_rng(1)
    n = 300
    fields = {nm: (rng.standard_normal((3, st.ncells)) * 10.0 ** rng.integers(-8, 8, (3, st.ncells))).astype(np.float32)
              for nm in ("E", "B", "J")}
    for nm in (openpmd.PML_E, openpmd.PML_B):
        fields[nm] = (rng.standard_normal((6, st.ncells)) * 1e-3).astype(np.float32)
    sp = {"name": "e", "index": 0,
          "cell": np.stack([rng.integers(0, g, n) for g in st.grid], 1).astype(np.int32),
          "pos": rng.random((n, 3), dtype=np.float32),
          "mom": (rng.standard_normal((n, 3)) * 1e-3).astype(np.float32),
          "w": np.full(n, np.float32(1250.0), np.float32),
          "desc": {"shape": abi.SHAPE_TSC, "pusher": abi.PUSHER_BORIS, "current": "Esirkepov",
                   "mass": float(k.species_mass[0]), "charge": float(k.species_charge[0])}}
    sp["pos"][0] = np.nextafter(np.float32(1), np.float32(0))  # the largest in-cell position
    sp["pos"][1] = 0.0
    return st, {"grid": st.grid, "k": k, "step": step, "jexp": 37, "version": 1, "fields": fields, "species": [sp],
                "config": {"solver": "Yee", "field_boundary": ["periodic"] * 6, "current_smoothing": "none"}}


def test_round_trip_is_exact(tmp_path):
    st, snap = synthetic_snapshot()
    path = openpmd.write_snapshot(snap, str(tmp_path))
    assert path.endswith("simData_7.json")
    step, it = openpmd.read_iteration(path)
    assert step == 7
    for nm in ("E", "B", "J", openpmd.PML_E, openpmd.PML_B):
        got = openpmd.mesh_array(it, nm)
        np.testing.assert_array_equal(got.view(np.uint32), snap["fields"][nm].view(np.uint32))
    cell, pos, mom, w = openpmd.species_arrays(it, "e")
    s = snap["species"][0]
    np.testing.assert_array_equal(cell, s["cell"])
    for a, b in ((pos, s["pos"]), (mom, s["mom"]), (w, s["w"])):
        np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32))
    assert it["attributes"]["jScaleExp"]["value"] == 37


def test_layout_and_units(tmp_path):
    st, snap = synthetic_snapshot(step=3)
    path = openpmd.write_snapshot(snap, str(tmp_path))
    series = json.load(open(path))
    a = series["attributes"]
    assert a["openPMD"]["value"] == "1.1.0" and a["basePath"]["value"] == "/data/%T/"
    assert a["meshesPath"]["value"] == "meshes/" and a["particlesPath"]["value"] == "particles/"
    assert a["iterationEncoding"]["value"] == "fileBased" and a["iterationFormat"]["value"] == "simData_%T.json"
    it = series["data"]["3"]
    k = snap["k"]
    u = openpmd.units(k)
    # iteration: dt and time in the time unit, which is dt in SI (simulation.unitless:435-438)
    assert it["attributes"]["timeUnitSI"]["value"] == pytest.approx(st.dt_si, rel=1e-15)
    assert it["attributes"]["dt"]["value"] == pytest.approx(1.0, rel=1e-7)
    assert it["attributes"]["time"]["value"] == pytest.approx(3.0, rel=1e-7)
    E = it["meshes"]["E"]
    assert E["attributes"]["axisLabels"]["value"] == ["z", "y", "x"]
    assert E["attributes"]["dataOrder"]["value"] == "C"
    assert E["attributes"]["unitDimension"]["value"] == [1, 1, -3, -1, 0, 0, 0]  # V/m = kg m / (A s^3)
    # grid spacing (z, y, x) times gridUnitSI gives the cell size in metres
    gs = np.array(E["attributes"]["gridSpacing"]["value"]) * E["attributes"]["gridUnitSI"]["value"]
    np.testing.assert_allclose(gs, st.cell_si[::-1], rtol=1e-6)
    assert E["x"]["attributes"]["position"]["value"] == [0.0, 0.0, 0.5]  # E_x at (1/2, 0, 0), listed z, y, x
    assert it["meshes"]["B"]["x"]["attributes"]["position"]["value"] == [0.5, 0.5, 0.0]
    # the PML's psi: NodeValues components, each at the Yee position and in the
    # unit of the field component it enters (its first letter)
    P = it["meshes"][openpmd.PML_E]
    assert [c for c in P if c != "attributes"] == ["xy", "xz", "yx", "yz", "zx", "zy"]
    assert P["yz"]["attributes"]["position"]["value"] == E["y"]["attributes"]["position"]["value"]
    assert P["zx"]["attributes"]["unitSI"]["value"] == E["z"]["attributes"]["unitSI"]["value"]
    assert it["meshes"][openpmd.PML_B]["xz"]["attributes"]["position"]["value"] == [0.5, 0.5, 0.0]
    assert P["attributes"]["unitDimension"]["value"] == E["attributes"]["unitDimension"]["value"]
    assert E["x"]["extent"] == [st.grid[2], st.grid[1], st.grid[0]]
    # E unit: V/m per PIC unit; the electron's frame charge times E times dt gives momentum in the momentum unit
    assert E["x"]["attributes"]["unitSI"]["value"] == pytest.approx(u["mass"] * u["length"] /
                                                                     (u["time"] ** 2 * u["charge"]), rel=1e-12)
    e = it["particles"]["e"]
    np.testing.assert_allclose(e["position"]["x"]["attributes"]["unitSI"]["value"], st.cell_si[0], rtol=1e-6)
    np.testing.assert_allclose(e["positionOffset"]["y"]["attributes"]["unitSI"]["value"], st.cell_si[1], rtol=1e-6)
    assert e["momentum"]["attributes"]["macroWeighted"]["value"] == 1
    assert e["momentum"]["attributes"]["weightingPower"]["value"] == 1.0
    # constant records: the electron charge and mass per unit weighting, in SI
    q_si = e["charge"]["attributes"]["value"]["value"] * e["charge"]["attributes"]["unitSI"]["value"]
    m_si = e["mass"]["attributes"]["value"]["value"] * e["mass"]["attributes"]["unitSI"]["value"]
    assert q_si == pytest.approx(setups.ELECTRON_CHARGE_SI, rel=1e-6)
    assert m_si == pytest.approx(setups.ELECTRON_MASS_SI, rel=1e-6)
    assert e["charge"]["attributes"]["shape"]["value"] == [300]
    assert e["particlePatches"]["numParticles"]["data"] == [300]
    assert e["attributes"]["particlePush"]["value"] == "Boris"
