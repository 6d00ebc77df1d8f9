"""Checkpoint/restart through openPMD files (plugins/Checkpoint.x.cpp:45,
restart/LoadSpecies.hpp:54-80, restart/RestartFieldLoader.hpp:205-208): a run
written at step k, restarted in a fresh context and continued equals the
uninterrupted run bit for bit (frames, E, B, J), for a periodic two-species
box and for an open box with the exponential absorber, an incident laser
(timed by the restored field step counter) and an immobile species."""
import pytest

from picongpu_amd import abi, openpmd, setups
from tests import helpers as H

pytestmark = pytest.mark.gpu


def species_names(st):
    return [s.name for s in st.species]


def compare(a, b, st, what):
    for sp in range(len(st.species)):
        H.assert_frames_equal(a.read_frames(sp), b.read_frames(sp), f"{what} species {sp}")
    for f in (abi.FIELD_E, abi.FIELD_B, abi.FIELD_J):
        H.assert_field_equal(a.read_field(f), b.read_field(f).astype("float64"), f"{what} field {f}")


@pytest.mark.parametrize("case", ["kelvin_helmholtz", "laser_plasma"])
def test_restart_continues_bit_exact(case, tmp_path):
    if case == "kelvin_helmholtz":
        st = setups.kelvin_helmholtz(grid=(32, 32, 16), ppc=2)
    else:
        st = setups.laser_plasma(grid=(32, 96, 16), ppc=2, laser="incident")
    print(f"case openPMD restart {case} grid={st.grid}", flush=True)
    first, second = 5, 6
    a = abi.Simulation(st.config())
    setups.init_device(a, st)
    a.step(first)
    path = openpmd.write(a, str(tmp_path), species_names(st))
    a.step(second)

    b = abi.Simulation(openpmd.restart_config(st.config(), path))
    assert openpmd.restart(b, path, species_names(st)) == first
    assert b.get_step() == first
    b.step(second)
    compare(a, b, st, f"{case} after {first} + {second} steps")
    a.close()
    b.close()
