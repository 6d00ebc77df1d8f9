"""C ABI checks that need no GPU: the library loads, exports every function
include/picongpu_amd.h declares, derives the same FP32 constants as the
oracle bit for bit, and rejects invalid configurations before touching a
device."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from oracle import oracle as O
from picongpu_amd import abi, setups

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "picongpu_amd.h")


def header_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pic_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = abi.load_library()
    names = header_functions()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # the Python binding covers exactly the declared set
    assert sorted(abi.SIGNATURES) == names


def test_abi_version():
    assert abi.load_library().pic_abi_version() == 1


SETUPS = [
    setups.thermal(),
    setups.kelvin_helmholtz(),
    setups.laser_plasma(),
    setups.thermal(grid=(24, 24, 24), shape=abi.SHAPE_PQS, pusher=abi.PUSHER_VAY),
]


@pytest.mark.parametrize("st", SETUPS, ids=[s.name for s in SETUPS])
def test_constants_match_oracle_bitwise(st):
    k = abi.constants_from_config(st.config())
    o = O.constants(st)
    for name in ("dt", "c", "eps0", "mue0", "mue0_eps0", "add_current_coeff", "inv_cell_volume_dt", "lehe_alpha",
                 "lehe_delta"):
        a, b = np.float32(getattr(k, name)), np.float32(getattr(o, name))
        assert a.view(np.uint32) == b.view(np.uint32), name
    for d in range(3):
        assert np.float32(k.cell[d]).view(np.uint32) == np.float32(o.cell[d]).view(np.uint32)
    for s in range(len(st.species)):
        assert np.float32(k.species_mass[s]) == np.float32(o.species_mass[s])
        assert np.float32(k.species_charge[s]) == np.float32(o.species_charge[s])
    for name in ("unit_length", "unit_time", "unit_mass", "unit_charge", "typical_weighting"):
        assert getattr(k, name) == getattr(o, name)


def _create_status(cfg):
    lib = abi.load_library()
    h = C.c_void_p()
    rc = lib.pic_create(C.byref(cfg), C.byref(h))
    if rc == 0:
        lib.pic_destroy(h)
    return rc, lib.pic_last_error().decode()


def test_rejects_grid_not_multiple_of_supercell():
    st = setups.thermal(grid=(20, 16, 16))
    rc, msg = _create_status(st.config())
    assert rc == -1 and "multiple" in msg


def test_rejects_cfl_violation():
    st = setups.thermal(grid=(16, 16, 16))
    st.dt_si *= 1.2
    rc, msg = _create_status(st.config())
    assert rc == -1 and "Courant" in msg


def test_rejects_bad_species():
    cfg = setups.thermal(grid=(16, 16, 16)).config()
    cfg.species[0].shape = 7
    rc, msg = _create_status(cfg)
    assert rc == -1 and "shape" in msg


def test_bin_frames_layout():
    st = setups.thermal(grid=(16, 16, 12), ppc=2)
    cells, pos, mom, w = setups.host_particles(st, 0)
    f = setups.bin_frames(st, 0, cells, pos, mom, w)
    assert f["counts"].sum() == cells.shape[0]
    assert np.all(f["counts"] == 2 * 256)
    cap = st.species[0].frames() * 256
    # all frames full except the last: occupancy is contiguous from slot 0
    m = f["mask"].reshape(-1, cap)
    for sc in range(m.shape[0]):
        n = f["counts"][sc]
        assert np.all(m[sc, :n] == 1) and np.all(m[sc, n:] == 0)


def _incident_setup(**prof_over):
    st = setups.laser_plasma(grid=(32, 64, 32), ppc=1, laser="incident")
    prof = dict(st.incident[(1, 0)])
    prof.update(prof_over)
    st.incident = {(1, 0): prof}
    return st


def test_incident_setup_is_accepted():
    rc, msg = _create_status(_incident_setup().config())
    # no device here: creation fails only after validation, at the first HIP call
    assert rc in (0, -2), msg


@pytest.mark.parametrize("over,fragment", [
    (dict(direction=(0, 1.1, 0)), "direction"),
    (dict(polarisation_direction=(0, 1, 0)), "orthogonal"),
    (dict(polarisation_direction=(0.5, 0, 0)), "polarisation direction"),
    (dict(laguerre_modes=(1.0, -1.0)), "laguerreModes"),
    (dict(time_delay_si=-1e-15), "time delay"),
    (dict(w0_si=0.0), "W0"),
    (dict(kind=7), "profile kind"),
])
def test_incident_rejects_bad_profile(over, fragment):
    rc, msg = _create_status(_incident_setup(**over).config())
    assert rc == -1 and fragment in msg, msg


def test_incident_rejects_surface_inside_absorber():
    st = _incident_setup()
    # Lehe along y: the B update's margin is 2, so the surface must keep 12 + 2 - 1 cells from the border
    st.incident_position = ((16, -16), (12, -16), (16, -16))
    rc, msg = _create_status(st.config())
    assert rc == -1 and "too close to the boundary" in msg and "13" in msg
    st.incident_position = ((16, -16), (13, -16), (16, -16))
    rc, msg = _create_status(st.config())
    assert rc in (0, -2), msg


def test_incident_rejected_by_ensemble_and_differentiable_forms():
    st = _incident_setup()
    cfgs = (abi.Config * 1)(st.config())
    lib = abi.load_library()
    h = C.c_void_p()
    rc = lib.pic_ensemble_create(cfgs, 1, C.byref(h))
    assert rc == -1 and "incident" in lib.pic_last_error().decode()
