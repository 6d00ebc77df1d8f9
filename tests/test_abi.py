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
