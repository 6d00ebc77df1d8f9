"""The oracle's PML (absorber/pml/Pml.kernel:60-590, the reference's default
field absorber), pinned by what the scheme must do rather than by its
formulas: with zero conductivity and frequency shift and unit stretching it
is the plain Yee/Lehe update bit for bit (b = 1, c = 0: no cell is in the
layer), and a pulse that runs into a 12-cell layer with the
fieldAbsorber.param defaults leaves almost no energy behind, where an open
border without absorber reflects it whole and a 32-cell exponential absorber
keeps a few percent."""
import numpy as np
import pytest

from oracle import oracle as O
from picongpu_amd import abi
from tests.test_oracle_boundary import field_only

NEUTRAL = dict(sigma_kappa_grading_order=4.0, sigma_max_si=(0.0, 0.0, 0.0), kappa_max=(1.0, 1.0, 1.0),
               alpha_grading_order=1.0, alpha_max_si=(0.0, 0.0, 0.0))


def pulse_run(absorber, solver=abi.SOLVER_YEE, cells=12, pml=None, steps=400):
    st = field_only((256, 8, 4), absorber=absorber, cells=((cells, cells), (0, 0), (0, 0)),
                    strength=((0.1, 0.1), (0.0, 0.0), (0.0, 0.0)))
    st.pml = pml
    if solver == abi.SOLVER_LEHE:
        st.solver = abi.SOLVER_LEHE
        st.lehe_dir = 0
    ora = O.Oracle(st, 32, 1)
    k = O.constants(st)
    x = np.arange(256)
    E = np.zeros((3, st.ncells), np.float32)
    B = np.zeros((3, st.ncells), np.float32)
    E[1] = np.tile(np.exp(-((x - 128) / 10.0) ** 2), 32).astype(np.float32)
    B[2] = (np.tile(np.exp(-((x + 0.5 - 128) / 10.0) ** 2), 32) / k.c).astype(np.float32)
    ora.set_field(abi.FIELD_E, E)
    ora.set_field(abi.FIELD_B, B)
    e0 = ora.field_energy()[0]
    for _ in range(steps):  # the pulse travels about 230 cells: into the border and back
        ora.update_before_current()
        ora.add_current()
        ora.update_after_current()
    e1 = ora.field_energy()[0]
    fields = (ora.get_field(abi.FIELD_E), ora.get_field(abi.FIELD_B))
    ora.close()
    return e1 / e0, fields


@pytest.mark.parametrize("solver", [abi.SOLVER_YEE, abi.SOLVER_LEHE])
def test_pml_without_conductivity_is_the_plain_update(solver):
    _, (e_pml, b_pml) = pulse_run(abi.ABSORBER_PML, solver, pml=NEUTRAL, steps=60)
    _, (e_none, b_none) = pulse_run(abi.ABSORBER_NONE, solver, steps=60)
    np.testing.assert_array_equal(e_pml, e_none)
    np.testing.assert_array_equal(b_pml, b_none)


@pytest.mark.parametrize("solver", [abi.SOLVER_YEE, abi.SOLVER_LEHE])
def test_pml_absorbs_a_pulse(solver):
    left_pml, _ = pulse_run(abi.ABSORBER_PML, solver)
    left_none, _ = pulse_run(abi.ABSORBER_NONE, solver)
    left_exp, _ = pulse_run(abi.ABSORBER_EXPONENTIAL, solver, cells=32)
    print(f"energy left after 400 steps: PML (12 cells) {left_pml:.3e}, exponential (32 cells) {left_exp:.3e}, "
          f"open border {left_none:.3f}")
    assert left_none > 0.99
    assert left_pml < 1e-4
    assert left_pml < left_exp / 100
