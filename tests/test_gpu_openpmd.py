"""Checkpoint/restart through openPMD files (plugins/Checkpoint.x.cpp:45,
restart/LoadSpecies.hpp:54-80, restart/RestartFieldLoader.hpp:205-208): a run
written at step k, restarted in a fresh context and continued equals the
uninterrupted run bit for bit (frames, E, B, J and the PML's convolution
fields), for a periodic two-species box and for an open box with an
incident laser (timed by the restored field step counter), an immobile
species and the PML (its psi fields travel in the file, as the reference
checkpoints them) or the exponential absorber."""
import numpy as np
import pytest

from picongpu_amd import abi, openpmd, setups
from tests import helpers as H

pytestmark = pytest.mark.gpu


def species_names(st):
    return [s.name for s in st.species]


def compare(a, b, st, what):
    for sp in range(len(st.species)):
        H.assert_frames_equal(a.read_frames(sp), b.read_frames(sp), f"{what} species {sp}")
    fields = [abi.FIELD_E, abi.FIELD_B, abi.FIELD_J]
    if st.absorber == abi.ABSORBER_PML:
        fields += [abi.FIELD_PML_E, abi.FIELD_PML_B]
    for f in fields:
        fa, fb = a.read_field(f), b.read_field(f)
        assert fa.view("uint32").tobytes() == fb.view("uint32").tobytes(), f"{what} field {f}"


@pytest.mark.parametrize("case", ["kelvin_helmholtz", "thermal_pml", "laser_plasma", "laser_plasma_exponential"])
def test_restart_continues_bit_exact(case, tmp_path):
    if case == "kelvin_helmholtz":
        st = setups.kelvin_helmholtz(grid=(32, 32, 16), ppc=2)
    elif case == "thermal_pml":
        # a warm plasma filling an open box: its currents drive the fields,
        # and so psi, inside the 6-cell layers from the first step
        st = setups.thermal(grid=(32, 32, 16), ppc=2)
        st.species[0].u_th = (0.05, 0.05, 0.05)
        st.boundary = (abi.BOUNDARY_OPEN, abi.BOUNDARY_OPEN, abi.BOUNDARY_PERIODIC)
        st.absorber = abi.ABSORBER_PML
        st.absorber_cells = ((6, 6), (6, 6), (0, 0))
    else:
        # the Huygens surface lies 16 cells inside the open borders: x spans 48
        absorber = "exponential" if case.endswith("exponential") else "pml"
        st = setups.laser_plasma(grid=(48, 96, 16), ppc=2, laser="incident", absorber=absorber)
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
    if case == "thermal_pml":
        for f in (abi.FIELD_PML_E, abi.FIELD_PML_B):
            assert np.any(b.read_field(f) != 0.0), f"the restart loaded no psi ({f})"
    b.step(second)
    compare(a, b, st, f"{case} after {first} + {second} steps")
    a.close()
    b.close()
