"""Parity of the HIP library with the CPU oracle on identical seeded inputs.

Bar: bit-exact for particle state (positions, momenta, cell indices,
multiMask, supercell counts) and for E, B, J, against the oracle's device
summation order. Every case prints what it runs before the GPU work.
"""
from dataclasses import replace

import numpy as np
import pytest

from oracle import oracle as O
from picongpu_amd import abi, setups
from tests import helpers as H

pytestmark = pytest.mark.gpu


def small_thermal(shape=abi.SHAPE_CIC, pusher=abi.PUSHER_BORIS, grid=(16, 16, 12), ppc=4, u=0.05):
    st = setups.thermal(grid=grid, ppc=ppc, shape=shape, pusher=pusher)
    st.species[0].u_th = (u, u, u)
    return st


def emz(st, species=None):
    """The EmZ current solver for the given species (all by default)."""
    for i, s in enumerate(st.species):
        if species is None or i in species:
            s.current_solver = abi.CURRENT_SOLVER_EMZ
    return st


def open_box(st, boundary=(1, 0, 1), cells=((4, 6), (3, 5), (2, 4)), strength=((0.05, 0.1), (0.2, 0.03), (0.1, 0.1))):
    """Open borders on the given axes with the exponential absorber."""
    st.boundary = boundary
    st.absorber = abi.ABSORBER_EXPONENTIAL
    st.absorber_cells = cells
    st.absorber_strength = strength
    return st


def pml_box(st, boundary=(1, 0, 1), cells=((4, 6), (3, 5), (2, 4)), pml=None):
    """Open borders on the given axes with the PML (fieldAbsorber.param's
    defaults unless pml is given)."""
    st.boundary = boundary
    st.absorber = abi.ABSORBER_PML
    st.absorber_cells = cells
    st.pml = pml
    return st


def small_two_species(shape=abi.SHAPE_TSC, grid=(16, 16, 12), ppc=3):
    st = setups.kelvin_helmholtz(grid=grid, ppc=ppc, shape=shape)
    for s in st.species:
        s.u_th = (0.02, 0.02, 0.02)
    return st


CASES = [
    ("cic-boris", lambda: small_thermal(abi.SHAPE_CIC, abi.PUSHER_BORIS)),
    ("tsc-boris", lambda: small_thermal(abi.SHAPE_TSC, abi.PUSHER_BORIS)),
    ("pqs-boris", lambda: small_thermal(abi.SHAPE_PQS, abi.PUSHER_BORIS)),
    ("cic-vay", lambda: small_thermal(abi.SHAPE_CIC, abi.PUSHER_VAY)),
    ("pqs-vay", lambda: small_thermal(abi.SHAPE_PQS, abi.PUSHER_VAY)),
    ("cic-hc", lambda: small_thermal(abi.SHAPE_CIC, abi.PUSHER_HIGUERA_CARY, u=0.3)),
    ("tsc-hc", lambda: small_thermal(abi.SHAPE_TSC, abi.PUSHER_HIGUERA_CARY, u=0.3)),
    ("cic-emz", lambda: emz(small_thermal(abi.SHAPE_CIC, abi.PUSHER_BORIS, u=0.3))),
    ("tsc-emz", lambda: emz(small_thermal(abi.SHAPE_TSC, abi.PUSHER_HIGUERA_CARY, u=0.3))),
    ("pqs-emz", lambda: emz(small_thermal(abi.SHAPE_PQS, abi.PUSHER_VAY, u=0.3))),
]


@pytest.mark.parametrize("name,make", CASES, ids=[c[0] for c in CASES])
def test_push_fused_bit_exact(name, make):
    st = make()
    print(f"case push_fused {name} grid={st.grid}", flush=True)
    sim, ora, _ = H.make_pair(st)
    sim.current_reset()
    sim.push(-1, True)
    ora.current_reset()
    ora.push(-1, True)
    H.assert_frames_equal(sim.read_frames(0), ora.get_frames(0), f"{name} push")
    H.assert_field_equal(sim.read_field(abi.FIELD_J), ora.get_field(abi.FIELD_J), f"{name} J")
    sim.close()


DEPOSIT_CASES = CASES[:3] + CASES[-3:]


@pytest.mark.parametrize("name,make", DEPOSIT_CASES, ids=[c[0] for c in DEPOSIT_CASES])
def test_deposit_unfused_bit_exact(name, make):
    st = make()
    print(f"case deposit {name} grid={st.grid}", flush=True)
    sim, ora, _ = H.make_pair(st)
    sim.current_reset()
    sim.deposit(-1)
    ora.current_reset()
    ora.deposit(-1)
    H.assert_field_equal(sim.read_field(abi.FIELD_J), ora.get_field(abi.FIELD_J), f"{name} J")
    sim.close()


@pytest.mark.parametrize("open_", [False, True], ids=["periodic", "open"])
def test_shift_bit_exact(open_):
    st = small_thermal(abi.SHAPE_CIC, abi.PUSHER_BORIS, u=0.3)
    if open_:
        open_box(st, boundary=(1, 1, 0))
    print(f"case shift grid={st.grid} open={open_}", flush=True)
    sim, ora, _ = H.make_pair(st)
    for it in range(3):
        sim.push(-1, False)
        sim.shift(-1)
        ora.push(-1, False)
        ora.shift(-1)
        H.assert_frames_equal(sim.read_frames(0), ora.get_frames(0), f"shift step {it}")
    assert sim.error_flags() == 0
    sim.close()


# the stage kernels: per-cell kernels on 16x16x12, the z-marching B half on
# 128x128x32 (1024 tiles of 32x8 over chunks of 2 planes)
FIELD_STAGE_CASES = [(abi.SOLVER_YEE, False, False, (16, 16, 12)), (abi.SOLVER_LEHE, False, False, (16, 16, 12)),
                     (abi.SOLVER_YEE, True, False, (16, 16, 12)), (abi.SOLVER_LEHE, True, False, (16, 16, 12)),
                     (abi.SOLVER_YEE, False, True, (16, 16, 12)), (abi.SOLVER_YEE, False, False, (128, 128, 32)),
                     (abi.SOLVER_LEHE, False, False, (128, 128, 32)), (abi.SOLVER_YEE, True, False, (128, 128, 32)),
                     (abi.SOLVER_LEHE, True, False, (128, 128, 32))]


@pytest.mark.parametrize("solver,open_,binomial,grid", FIELD_STAGE_CASES,
                         ids=[f"{'yee' if s == abi.SOLVER_YEE else 'lehe'}{'-open' if o else ''}"
                              f"{'-binomial' if b else ''}-{'x'.join(map(str, g))}" for s, o, b, g in FIELD_STAGE_CASES])
def test_fields_bit_exact(solver, open_, binomial, grid):
    st = small_thermal(grid=grid, ppc=1)
    if open_:
        open_box(st, boundary=(1, 1, 1))
    if binomial:
        st.current_interpolation = abi.CURRENT_BINOMIAL
    st.solver = solver
    st.lehe_dir = 1
    if solver == abi.SOLVER_LEHE:
        st.dt_si = 0.99 * st.cell_si[1] / setups.SPEED_OF_LIGHT_SI / 1.8
    print(f"case fields solver={solver} open={open_} binomial={binomial} grid={grid}", flush=True)
    sim, ora, _ = H.make_pair(st, e_amp=1.0, b_amp=1.0)
    J = H.random_fields(st, 5, 0.1, 0.1)[0]
    sim.load_field(abi.FIELD_J, J)
    ora.set_field(abi.FIELD_J, J)
    sim.update_before_current()
    ora.update_before_current()
    sim.add_current()
    ora.add_current()
    sim.update_after_current()
    ora.update_after_current()
    for f in (abi.FIELD_E, abi.FIELD_B):
        H.assert_field_equal(sim.read_field(f), ora.get_field(f), f"field {f}")
    sim.close()


CUSTOM_PML = dict(sigma_kappa_grading_order=3.0, sigma_max_si=(2.0e6, 1.5e6, 3.0e6), kappa_max=(1.0, 4.0, 7.0),
                  alpha_grading_order=1.5, alpha_max_si=(0.3, 0.2, 0.15))
# 128x64x64 takes the z-marching B half step on the inner box
PML_CASES = [(abi.SOLVER_YEE, (1, 1, 1), None, (16, 16, 12)), (abi.SOLVER_LEHE, (1, 1, 1), None, (16, 16, 12)),
             (abi.SOLVER_YEE, (0, 1, 1), CUSTOM_PML, (16, 16, 12)),
             (abi.SOLVER_YEE, (1, 0, 1), None, (128, 64, 64)), (abi.SOLVER_LEHE, (1, 1, 0), None, (128, 64, 64))]


@pytest.mark.parametrize("solver,boundary,pml,grid", PML_CASES,
                         ids=[f"{'yee' if s == abi.SOLVER_YEE else 'lehe'}-{''.join(map(str, b))}"
                              f"{'-custom' if p else ''}-{g[0]}" for s, b, p, g in PML_CASES])
def test_pml_fields_bit_exact(solver, boundary, pml, grid):
    """Eight field steps through the stage calls with the PML on the open
    axes (psi of B advances in the first half step, E's in the E update),
    then eight through pic_step's field path (the plain kernels on the cells
    outside the layers, the PML kernels on the layer boxes), against the
    oracle."""
    st = pml_box(small_thermal(grid=grid, ppc=1), boundary=boundary, pml=pml)
    st.solver = solver
    st.lehe_dir = 1
    if solver == abi.SOLVER_LEHE:
        st.dt_si = 0.99 * st.cell_si[1] / setups.SPEED_OF_LIGHT_SI / 1.8
    print(f"case pml fields solver={solver} boundary={boundary} custom={pml is not None} grid={grid}", flush=True)
    sim, ora, _ = H.make_pair(st, e_amp=1.0, b_amp=1.0)
    J = H.random_fields(st, 5, 0.1, 0.1)[0]
    sim.load_field(abi.FIELD_J, J)
    ora.set_field(abi.FIELD_J, J)
    for it in range(8):
        sim.update_before_current()
        ora.update_before_current()
        sim.add_current()
        ora.add_current()
        sim.update_after_current()
        ora.update_after_current()
    for it in range(8):
        sim.fdtd_step()
        ora.update_before_current()
        ora.add_current()
        ora.update_after_current()
    for f in (abi.FIELD_E, abi.FIELD_B):
        H.assert_field_equal(sim.read_field(f), ora.get_field(f), f"field {f}")
    for f, which in ((abi.FIELD_PML_E, 0), (abi.FIELD_PML_B, 1)):
        got, want = sim.read_field(f), ora.get_pml(which)
        assert np.any(want != 0.0)
        np.testing.assert_array_equal(got.astype(np.float64), want, err_msg=f"psi {f}")
    sim.close()


# launch forms of the fused field step: 64^3 runs 8x8x4 tiles; the others
# z-march 32x8 xy tiles over chunks of 4 planes with partial xy tiles
# (200x104x96), 2 planes (128x128x68), 8 planes with a partial last chunk
# (256x256x68) and 16 planes (256x256x128)
FUSED_FIELD_CASES = [((64, 64, 64), abi.SOLVER_YEE), ((200, 104, 96), abi.SOLVER_YEE),
                     ((128, 128, 68), abi.SOLVER_YEE), ((256, 256, 68), abi.SOLVER_YEE),
                     ((256, 256, 128), abi.SOLVER_YEE), ((64, 64, 64), abi.SOLVER_LEHE),
                     ((200, 104, 96), abi.SOLVER_LEHE)]
# the same launch forms with open borders on all axes and the absorber
# (the last two: the binomial current filter)
FUSED_FIELD_CASES = [(g, s, o) for g, s in FUSED_FIELD_CASES for o in ("",)] + [
    ((64, 64, 64), abi.SOLVER_YEE, "open"), ((200, 104, 96), abi.SOLVER_YEE, "open"),
    ((256, 256, 68), abi.SOLVER_YEE, "open"), ((200, 104, 96), abi.SOLVER_LEHE, "open"),
    ((64, 64, 64), abi.SOLVER_YEE, "binomial"), ((256, 256, 68), abi.SOLVER_YEE, "binomial"),
    ((256, 256, 68), abi.SOLVER_LEHE, "dir0"), ((200, 104, 96), abi.SOLVER_LEHE, "dir2"),
    ((256, 256, 68), abi.SOLVER_LEHE, "open")]


@pytest.mark.parametrize("grid,solver,variant", FUSED_FIELD_CASES,
                         ids=[f"{'yee' if s == abi.SOLVER_YEE else 'lehe'}-{'x'.join(map(str, g))}{'-' + o if o else ''}"
                              for g, s, o in FUSED_FIELD_CASES])
def test_fused_field_step_equals_stage_calls(grid, solver, variant):
    """pic_fdtd_step (two LDS-tiled kernels) against update_beforeCurrent,
    addCurrent and update_afterCurrent called one by one, which
    test_fields_bit_exact pins to the oracle; twice, so the second call runs
    on the swapped E buffers."""
    st = setups.thermal(grid=grid, ppc=1)
    if variant == "open":
        open_box(st, boundary=(1, 1, 1), cells=((12, 7), (5, 9), (3, 8)))
    if variant == "binomial":
        st.current_interpolation = abi.CURRENT_BINOMIAL
    st.solver = solver
    st.lehe_dir = {"dir0": 0, "dir2": 2}.get(variant, 1)
    if solver == abi.SOLVER_LEHE:
        st.dt_si = 0.99 * st.cell_si[1] / setups.SPEED_OF_LIGHT_SI / 1.8
    print(f"case fused field step grid={grid} solver={solver} variant={variant}", flush=True)
    E, B = H.random_fields(st, 3, 1.0, 1.0)
    J = H.random_fields(st, 4, 0.1, 0.1)[0]
    fused, staged = abi.Simulation(st.config()), abi.Simulation(st.config())
    for s in (fused, staged):
        s.load_field(abi.FIELD_E, E)
        s.load_field(abi.FIELD_B, B)
        s.load_field(abi.FIELD_J, J)
    for it in range(2):
        fused.fdtd_step()
        staged.update_before_current()
        staged.add_current()
        staged.update_after_current()
        for f in (abi.FIELD_E, abi.FIELD_B):
            H.assert_field_equal(fused.read_field(f), staged.read_field(f), f"call {it} field {f}")
    fused.close()
    staged.close()


def small_lehe():
    st = small_thermal(abi.SHAPE_TSC, abi.PUSHER_VAY)
    st.solver = abi.SOLVER_LEHE
    st.lehe_dir = 1
    st.dt_si = 0.99 * st.cell_si[1] / setups.SPEED_OF_LIGHT_SI / 1.8
    return st


STEP_CASES = [
    ("thermal-cic-boris", lambda: small_thermal(abi.SHAPE_CIC, abi.PUSHER_BORIS)),
    ("thermal-tsc-vay-lehe", small_lehe),
    ("thermal-pqs-vay", lambda: small_thermal(abi.SHAPE_PQS, abi.PUSHER_VAY)),
    ("thermal-tsc-hc", lambda: small_thermal(abi.SHAPE_TSC, abi.PUSHER_HIGUERA_CARY, u=0.3)),
    ("kh-tsc-2species", lambda: small_two_species(abi.SHAPE_TSC)),
    ("open-xz-cic-boris", lambda: open_box(small_thermal(abi.SHAPE_CIC, abi.PUSHER_BORIS, u=0.3))),
    ("open-xyz-tsc-vay-lehe", lambda: open_box(small_lehe(), boundary=(1, 1, 1))),
    ("open-y-kh-tsc", lambda: open_box(small_two_species(abi.SHAPE_TSC), boundary=(0, 1, 0))),
    ("binomial-kh-tsc", lambda: replace(small_two_species(abi.SHAPE_TSC), current_interpolation=abi.CURRENT_BINOMIAL)),
    ("emz-kh-tsc-electrons", lambda: emz(small_two_species(abi.SHAPE_TSC), species=(0,))),
    ("emz-open-cic", lambda: emz(open_box(small_thermal(abi.SHAPE_CIC, abi.PUSHER_BORIS, u=0.3)))),
    ("pml-xz-cic-boris", lambda: pml_box(small_thermal(abi.SHAPE_CIC, abi.PUSHER_BORIS, u=0.3))),
    ("pml-xyz-tsc-vay-lehe", lambda: pml_box(small_lehe(), boundary=(1, 1, 1))),
]


@pytest.mark.parametrize("schedule", [abi.PUSH_SCHEDULE_PHASES, abi.PUSH_SCHEDULE_SINGLE], ids=["phases", "single"])
@pytest.mark.parametrize("name,make", STEP_CASES, ids=[c[0] for c in STEP_CASES])
def test_steps_bit_exact(name, make, schedule):
    """Ten full steps against the oracle, the fused push launched per
    checkerboard phase or in one launch with the gather flush (the oracle
    has one order for both)."""
    st = make()
    nsteps = 10
    print(f"case steps {name} grid={st.grid} steps={nsteps} schedule={schedule}", flush=True)
    sim, ora, _ = H.make_pair(st, e_amp=1e-3, b_amp=1e-3)
    sim.set_push_schedule(schedule)
    supp = {abi.SHAPE_CIC: 2, abi.SHAPE_TSC: 3, abi.SHAPE_PQS: 4}[st.species[0].shape]
    assert sim.push_schedule(supp) == schedule
    n0 = sum(int(sim.read_frames(sp)["counts"].sum()) for sp in range(len(st.species)))
    sim.step(nsteps)
    ora.step(nsteps)
    for sp in range(len(st.species)):
        H.assert_frames_equal(sim.read_frames(sp), ora.get_frames(sp), f"{name} species {sp}")
    if any(st.boundary):  # the case must absorb particles at its open borders
        assert sum(int(sim.read_frames(sp)["counts"].sum()) for sp in range(len(st.species))) < n0
    for f in (abi.FIELD_E, abi.FIELD_B, abi.FIELD_J):
        H.assert_field_equal(sim.read_field(f), ora.get_field(f), f"{name} field {f}")
    sim.close()


def test_load_read_roundtrip():
    st = small_two_species()
    print("case load/read roundtrip", flush=True)
    sim = abi.Simulation(st.config())
    cells, pos, mom, w = setups.host_particles(st, 0)
    sim.load_particles(0, cells, pos, mom, w)
    f = setups.bin_frames(st, 0, cells, pos, mom, w)
    H.assert_frames_equal(sim.read_frames(0), {k: v for k, v in f.items()}, "load")
    c2, p2, m2, w2 = sim.read_particles(0)
    assert c2.shape[0] == cells.shape[0]
    sim.close()


def test_graph_replay_matches_eager():
    """pic_step_graph replays a captured step; every replay, with eager work
    and synchronisation in between, must equal eager stepping bit for bit."""
    st = small_two_species(abi.SHAPE_TSC)
    print("case graph replay", flush=True)
    a, _, frames = H.make_pair(st)
    b = abi.Simulation(st.config())
    for sp, f in enumerate(frames):
        b.write_frames(sp, f)
    E, B = H.random_fields(st, 11, 1e-3, 1e-3)
    b.load_field(abi.FIELD_E, E)
    b.load_field(abi.FIELD_B, B)
    for it in range(4):
        a.step(1)
        b.step_graph(1)
        b.field_energy()  # eager work between replays
        for sp in range(len(st.species)):
            H.assert_frames_equal(b.read_frames(sp), a.read_frames(sp), f"graph replay {it} species {sp}")
        for f in (abi.FIELD_E, abi.FIELD_B, abi.FIELD_J):
            H.assert_field_equal(b.read_field(f), a.read_field(f).astype(np.float64), f"graph replay {it} field {f}")
    a.close()
    b.close()


def reflecting(st, species_axes, boundary=(1, 1, 1)):
    """Open borders with the absorber; species s reflects on the axes species_axes[s]."""
    open_box(st, boundary=boundary)
    for s, axes in species_axes.items():
        st.species[s].particle_boundary = tuple(abi.PARTICLE_REFLECTING if d in axes else abi.PARTICLE_ABSORBING
                                                for d in range(3))
    return st


def pml_reflecting(st, species_axes, boundary=(1, 1, 1)):
    """Open borders with the PML (4 to 6 cells); species s reflects on the axes
    species_axes[s] and is absorbing, so damped in the layers, on the others."""
    pml_box(st, boundary=boundary)
    for s, axes in species_axes.items():
        st.species[s].particle_boundary = tuple(abi.PARTICLE_REFLECTING if d in axes else abi.PARTICLE_ABSORBING
                                                for d in range(3))
    return st


REFLECT_CASES = [
    ("reflect-xyz-tsc-boris", lambda: reflecting(small_thermal(abi.SHAPE_TSC, abi.PUSHER_BORIS, u=0.3),
                                                 {0: (0, 1, 2)})),
    ("reflect-xy-kh-electrons-absorbing-ions", lambda: reflecting(small_two_species(abi.SHAPE_TSC), {0: (0, 1)},
                                                                  boundary=(1, 1, 0))),
    ("reflect-y-pqs-vay-lehe", lambda: reflecting(small_lehe(), {0: (1,)}, boundary=(0, 1, 0))),
    ("reflect-xz-cic-hc-emz", lambda: emz(reflecting(small_thermal(abi.SHAPE_CIC, abi.PUSHER_HIGUERA_CARY, u=0.3),
                                                     {0: (0, 2)}, boundary=(1, 0, 1)))),
    # with the PML: the species reflecting on x is damped only in the y layers
    ("pml-reflect-x-tsc-boris", lambda: pml_reflecting(small_thermal(abi.SHAPE_TSC, abi.PUSHER_BORIS, u=0.3),
                                                       {0: (0,)}, boundary=(1, 1, 0))),
    ("pml-reflect-none-cic-hc-emz", lambda: emz(pml_reflecting(small_thermal(abi.SHAPE_CIC, abi.PUSHER_HIGUERA_CARY,
                                                                             u=0.3), {0: ()}, boundary=(1, 1, 1)))),
    ("pml-reflect-y-kh-electrons", lambda: pml_reflecting(small_two_species(abi.SHAPE_TSC), {0: (1,)},
                                                          boundary=(1, 1, 0))),
]


@pytest.mark.parametrize("name,make", REFLECT_CASES, ids=[c[0] for c in REFLECT_CASES])
def test_reflecting_borders_bit_exact(name, make):
    """Species reflecting at open borders (boundary/Reflecting.hpp): 10 steps
    bit-exact against the oracle; a species that reflects on every open axis
    keeps its particles."""
    st = make()
    nsteps = 10
    print(f"case reflecting {name} grid={st.grid} steps={nsteps}", flush=True)
    sim, ora, _ = H.make_pair(st, e_amp=1e-3, b_amp=1e-3)
    n0 = [int(sim.read_frames(sp)["counts"].sum()) for sp in range(len(st.species))]
    sim.step(nsteps)
    ora.step(nsteps)
    for sp in range(len(st.species)):
        f = sim.read_frames(sp)
        H.assert_frames_equal(f, ora.get_frames(sp), f"{name} species {sp}")
        reflects_all_open = all(st.species[sp].particle_boundary[d] == abi.PARTICLE_REFLECTING
                                for d in range(3) if st.boundary[d])
        if reflects_all_open:
            assert int(f["counts"].sum()) == n0[sp]
    for fld in (abi.FIELD_E, abi.FIELD_B, abi.FIELD_J):
        H.assert_field_equal(sim.read_field(fld), ora.get_field(fld), f"{name} field {fld}")
    sim.close()
