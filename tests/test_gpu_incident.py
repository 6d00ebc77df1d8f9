"""Parity of the incident field (Huygens surface) with the CPU oracle.

Bar: bit-exact E, B (and particle state where particles are present) against
the FP32 oracle after several steps. Both evaluate exp, cos and atan as the
FP64 function rounded to FP32; the FP32 rounding of FP64 results computed by
the device library and by the host's could differ only where an FP64 result
lies within its error of an FP32 rounding boundary, which none of these
inputs hits (the comparison would show it as a one-ulp difference).
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from picongpu_amd import abi, setups
from tests import helpers as H

pytestmark = pytest.mark.gpu

C_SI = 2.99792458e8


def vacuum(grid, cell, dt, incident, solver=abi.SOLVER_YEE, lehe_dir=1, boundary=(0, 0, 0), position=8,
           absorber_cells=4):
    st = setups.Setup("vacuum", tuple(grid), tuple(cell), dt, 1e25, 1, [], solver=solver, lehe_dir=lehe_dir,
                      boundary=boundary, incident=incident,
                      incident_position=((position, -position),) * 3)
    if any(boundary):
        st.absorber = abi.ABSORBER_EXPONENTIAL
        st.absorber_cells = ((absorber_cells, absorber_cells),) * 3
        st.absorber_strength = ((0.05, 0.05),) * 3
    return st


def gauss(direction, pol, focus, kind=abi.INCIDENT_GAUSSIAN_PULSE, polarisation=abi.POLARISATION_LINEAR, **kw):
    p = dict(kind=kind, wavelength_si=0.8e-6, amplitude_si=4e12, pulse_duration_si=2e-15, direction=direction,
             polarisation_direction=pol, focus_position_si=focus,
             focus_origin=(abi.ORIGIN_CENTER, abi.ORIGIN_CENTER, abi.ORIGIN_CENTER), w0_si=1.0e-6, pulse_init=4.0,
             polarisation=polarisation, ramp_init=4.0, laser_nofocus_constant_si=3e-15)
    p.update(kw)
    return p


CELL = (0.1e-6, 0.05e-6, 0.1e-6)
DT_LEHE_Y = 0.995 * CELL[1] / C_SI
CUBE = (0.05e-6, 0.05e-6, 0.05e-6)

VACUUM_CASES = [
    # plane wave along +y, Yee, contiguous surface across periodic x and z
    ("plane-ymin-yee", lambda: vacuum((8, 96, 4), CUBE, setups.yee_dt(CUBE, 0.95),
                                      {(1, 0): gauss((0, 1, 0), (1, 0, 0), (0, 0, 0), kind=abi.INCIDENT_PLANE_WAVE)},
                                      boundary=(0, 1, 0))),
    # Gaussian along +y, Lehe (Cherenkov-free along y: the B update's 2x2x2 coefficients), circular
    ("gauss-ymin-lehe-circular", lambda: vacuum((32, 64, 32), CELL, DT_LEHE_Y,
                                                {(1, 0): gauss((0, 1, 0), (1, 0, 0), (0, 1.5e-6, 0),
                                                               polarisation=abi.POLARISATION_CIRCULAR)},
                                                solver=abi.SOLVER_LEHE, boundary=(1, 1, 0))),
    # Gaussian along -y from the positive border, Lehe, Laguerre modes and a time delay
    ("gauss-ymax-lehe-modes", lambda: vacuum((32, 64, 32), CELL, DT_LEHE_Y,
                                             {(1, 1): gauss((0, -1, 0), (0, 0, 1), (0, 1.5e-6, 0),
                                                            laguerre_modes=(1.0, 0.4, -0.2),
                                                            laguerre_phases=(0.0, 0.3, 1.1), time_delay_si=1e-15)},
                                             solver=abi.SOLVER_LEHE, boundary=(0, 1, 0))),
    # Gaussian along -x from the positive x border, Lehe along y (x faces: coefficients along y only)
    ("gauss-xmax-lehe", lambda: vacuum((64, 32, 32), CELL, DT_LEHE_Y,
                                       {(0, 1): gauss((-1, 0, 0), (0, 1, 0), (0, 0, 0))},
                                       solver=abi.SOLVER_LEHE, boundary=(1, 0, 0))),
    # plane wave along +z (Lehe Cherenkov-free along z) and a Gaussian from XMin at once
    ("plane-zmin-gauss-xmin-lehe-z", lambda: vacuum(
        (32, 32, 64), CUBE, 0.95 * CUBE[2] / C_SI / math.sqrt(3),
        {(2, 0): gauss((0, 0, 1), (0, 1, 0), (0, 0, 0), kind=abi.INCIDENT_PLANE_WAVE,
                       polarisation=abi.POLARISATION_CIRCULAR),
         (0, 0): gauss((1, 0, 0), (0, 0, 1), (-0.5e-6, 0, 0))},
        solver=abi.SOLVER_LEHE, lehe_dir=2, boundary=(1, 1, 1))),
    # oblique Gaussian through the YMin face, Yee
    ("gauss-ymin-oblique-yee", lambda: vacuum((32, 64, 32), CUBE, setups.yee_dt(CUBE, 0.95),
                                              {(1, 0): gauss((0.6, 0.8, 0), (0.8, -0.6, 0), (0, 0, 0))},
                                              boundary=(0, 1, 0))),
]


@pytest.mark.parametrize("name,make", VACUUM_CASES, ids=[c[0] for c in VACUUM_CASES])
def test_incident_vacuum_bit_exact(name, make):
    st = make()
    nsteps = 40
    print(f"case incident vacuum {name} grid={st.grid} steps={nsteps}", flush=True)
    sim, ora, _ = H.make_pair(st, e_amp=0.0, b_amp=0.0)
    sim.step(nsteps)
    ora.step(nsteps)
    assert sim.get_step() == nsteps and ora.get_step() == nsteps
    E = sim.read_field(abi.FIELD_E)
    assert np.abs(E).max() > 1e-3, "the case injects no field"
    for f in (abi.FIELD_E, abi.FIELD_B):
        H.assert_field_equal(sim.read_field(f), ora.get_field(f), f"{name} field {f}")
    sim.close()
    ora.close()


def small_laser_plasma():
    st = setups.laser_plasma(grid=(32, 64, 32), ppc=2, laser="incident")
    st.absorber_cells = ((6, 6),) * 3
    st.incident_position = ((8, -8),) * 3
    prof = dict(st.incident[(1, 0)])
    prof.update(w0_si=1.0e-6, pulse_duration_si=2e-15, pulse_init=4.0, focus_position_si=(0.0, 1.2e-6, 0.0))
    st.incident = {(1, 0): prof}
    for s in st.species:
        s.ramp = (1, 0.5e-6, 1e-6)
    return st


def test_incident_laser_plasma_steps_bit_exact():
    st = small_laser_plasma()
    nsteps = 24
    print(f"case incident laser-plasma grid={st.grid} steps={nsteps}", flush=True)
    sim, ora, _ = H.make_pair(st, e_amp=0.0, b_amp=0.0)
    sim.step(nsteps)
    ora.step(nsteps)
    for sp in range(len(st.species)):
        H.assert_frames_equal(sim.read_frames(sp), ora.get_frames(sp), f"laser-plasma species {sp}")
    for f in (abi.FIELD_E, abi.FIELD_B, abi.FIELD_J):
        H.assert_field_equal(sim.read_field(f), ora.get_field(f), f"laser-plasma field {f}")
    # the pulse reached the plasma: the electrons carry current
    assert np.abs(sim.read_field(abi.FIELD_J)).max() > 0.0
    sim.close()
    ora.close()


def test_incident_stage_calls_equal_step_and_graph():
    """pic_step, the stage entry points one by one and pic_step_graph give the
    same state; the field step counter advances once per step in each."""
    st = small_laser_plasma()
    print("case incident stage calls / graph", flush=True)
    a, _, frames = H.make_pair(st, e_amp=0.0, b_amp=0.0)
    E, B = H.random_fields(st, 11, 0.0, 0.0)  # the fields make_pair loaded (zeros of either sign)
    sims = [a]
    for _ in range(2):
        b = abi.Simulation(st.config())
        for sp, f in enumerate(frames):
            b.write_frames(sp, f)
        b.load_field(abi.FIELD_E, E)
        b.load_field(abi.FIELD_B, B)
        sims.append(b)
    a_, stage, graph = sims
    for it in range(6):
        a_.step(1)
        stage.current_reset()
        stage.push(-1, True)
        stage.shift(-1)
        stage.update_before_current()
        stage.add_current()
        stage.update_after_current()
        graph.step_graph(1)
        graph.field_energy()  # eager work between replays
        for other, what in ((stage, "stage"), (graph, "graph")):
            assert other.get_step() == it + 1
            for f in (abi.FIELD_E, abi.FIELD_B, abi.FIELD_J):
                H.assert_field_equal(other.read_field(f), a_.read_field(f).astype(np.float64), f"{what} {it} {f}")
            for sp in range(len(st.species)):
                H.assert_frames_equal(other.read_frames(sp), a_.read_frames(sp), f"{what} {it} species {sp}")
    for s in sims:
        s.close()


def test_incident_set_step():
    st = VACUUM_CASES[1][1]()
    print("case incident set_step", flush=True)
    sim, ora, _ = H.make_pair(st, e_amp=0.0, b_amp=0.0)
    sim.set_step(30)
    ora.set_step(30)
    assert sim.get_step() == 30
    sim.step(5)
    ora.step(5)
    assert sim.get_step() == 35
    for f in (abi.FIELD_E, abi.FIELD_B):
        H.assert_field_equal(sim.read_field(f), ora.get_field(f), f"set_step field {f}")
    assert np.abs(sim.read_field(abi.FIELD_E)).max() > 1e-3
    sim.close()
    ora.close()
