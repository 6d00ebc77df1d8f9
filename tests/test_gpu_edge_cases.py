"""Edge cases of the frame layout, the push and the shift, bit-exact against
the oracle: empty and ragged supercells (counts at and around frame
boundaries), a supercell larger than the push kernel's leaver lists reach
(the shift then collects its leavers itself), positions on cell borders,
ultra-relativistic momenta for every pusher, and supercells that overflow
their frames in the shift: they grow, and no particle is lost."""
import copy

import numpy as np
import pytest

from picongpu_amd import abi, setups
from tests import helpers as H
from oracle import oracle as O

pytestmark = pytest.mark.gpu

GRID = (16, 16, 12)  # 2 x 2 x 3 supercells of 8 x 8 x 4


def crafted(st, counts, seed, u=0.3, positions=None):
    """Flat particles with counts[s] particles in supercell s at random cells
    of it; positions drawn uniformly unless `positions` (values to draw each
    in-cell coordinate from) is given; momenta mass * u * N(0, 1)."""
    rng = np.random.default_rng(seed)
    nscx, nscy = GRID[0] // 8, GRID[1] // 8
    cells, pos = [], []
    for s, n in enumerate(counts):
        sx, sy, sz = s % nscx, (s // nscx) % nscy, s // (nscx * nscy)
        lc = rng.integers(0, 256, n)
        c = np.stack([sx * 8 + lc % 8, sy * 8 + (lc // 8) % 8, sz * 4 + lc // 64], axis=1)
        cells.append(c)
        if positions is None:
            pos.append(rng.random((n, 3), dtype=np.float32))
        else:
            pos.append(rng.choice(np.asarray(positions, np.float32), (n, 3)))
    cells = np.concatenate(cells).astype(np.int32)
    pos = np.concatenate(pos).astype(np.float32)
    s = st.species[0]
    vol = st.cell_si[0] * st.cell_si[1] * st.cell_si[2]
    weighting = np.float32(s.density_si * vol / s.ppc)
    mass = float(np.float32(setups.abi_free_constants(st)["base_mass"] * np.float32(s.mass_ratio)) * weighting)
    n = cells.shape[0]
    mom = (rng.standard_normal((n, 3)) * u * mass).astype(np.float32)
    w = np.full(n, weighting, np.float32)
    return cells, pos, mom, w


def run_pair(st, cells, pos, mom, w, frames, nsteps):
    f = setups.bin_frames(st, 0, cells, pos, mom, w, frames_per_sc=frames)
    sim = abi.Simulation(st.config())
    sim.write_frames(0, f)
    ora = O.Oracle(st, 32, 0, jexp=sim.j_scale_exp())
    ora.set_frames(0, f)
    E, B = H.random_fields(st, 17, 1e-3, 1e-3)
    sim.load_field(abi.FIELD_E, E)
    sim.load_field(abi.FIELD_B, B)
    ora.set_field(abi.FIELD_E, E)
    ora.set_field(abi.FIELD_B, B)
    sim.step(nsteps)
    ora.step(nsteps)
    H.assert_frames_equal(sim.read_frames(0), ora.get_frames(0), "frames")
    for fld in (abi.FIELD_E, abi.FIELD_B, abi.FIELD_J):
        H.assert_field_equal(sim.read_field(fld), ora.get_field(fld), f"field {fld}")
    counts = sim.read_frames(0)["counts"]
    sim.close()
    ora.close()
    return counts


def thermal(shape=abi.SHAPE_TSC, pusher=abi.PUSHER_BORIS, frames=6):
    st = setups.thermal(grid=GRID, ppc=4, shape=shape, pusher=pusher)
    st.species[0].u_th = (0.1, 0.1, 0.1)
    st.species[0].frames_per_sc = frames
    return st


@pytest.mark.parametrize("shape", [abi.SHAPE_CIC, abi.SHAPE_TSC, abi.SHAPE_PQS])
def test_empty_and_ragged_supercells(shape):
    st = thermal(shape=shape, frames=6)  # 1536 slots per supercell
    counts = [0, 1, 63, 64, 255, 256, 257, 511, 512, 700, 1000, 0]
    print(f"case ragged supercells shape={shape} counts={counts}", flush=True)
    cells, pos, mom, w = crafted(st, counts, seed=3)
    out = run_pair(st, cells, pos, mom, w, frames=6, nsteps=4)
    assert int(out.sum()) == sum(counts)  # periodic box: nothing leaves


def test_supercell_beyond_the_leaver_lists():
    """17000 particles in one supercell (more than 512 threads x 32 iterations
    of the push's leaver lists): the shift collects that species itself."""
    st = thermal(frames=70)  # 17920 slots
    counts = [17000, 300, 0, 17, 256, 1, 2, 3, 511, 0, 64, 5]
    print(f"case large supercell counts={counts}", flush=True)
    cells, pos, mom, w = crafted(st, counts, seed=5)
    out = run_pair(st, cells, pos, mom, w, frames=70, nsteps=3)
    assert int(out.sum()) == sum(counts)


def test_cell_border_positions():
    """In-cell positions exactly on the cell borders (0, -0, the largest
    float below 1) and at 1/2, moved across the borders in both directions."""
    st = thermal(shape=abi.SHAPE_TSC, frames=4)
    borders = [0.0, -0.0, np.nextafter(np.float32(1), np.float32(0)), 0.5, np.float32(2.0 ** -24)]
    counts = [200] * 12
    print("case cell-border positions", flush=True)
    cells, pos, mom, w = crafted(st, counts, seed=7, positions=borders)
    run_pair(st, cells, pos, mom, w, frames=4, nsteps=3)


@pytest.mark.parametrize("pusher", [abi.PUSHER_BORIS, abi.PUSHER_VAY, abi.PUSHER_HIGUERA_CARY])
def test_ultra_relativistic_momenta(pusher):
    """gamma ~ 100 .. 1000: the pushers' FP64 sections and velocities near c
    (every particle moves almost a full cell per step along some axis)."""
    st = thermal(shape=abi.SHAPE_CIC, pusher=pusher, frames=6)
    counts = [300] * 12
    print(f"case ultra-relativistic pusher={pusher}", flush=True)
    cells, pos, mom, w = crafted(st, counts, seed=11, u=300.0)
    run_pair(st, cells, pos, mom, w, frames=6, nsteps=3)


def converging(st, counts, seed, u):
    """crafted() particles whose momenta point towards the box centre, so the
    supercells there fill up over the steps."""
    cells, pos, mom, w = crafted(st, counts, seed=seed, u=0.0)
    centre = np.asarray(GRID, np.float64) / 2.0
    x = cells.astype(np.float64) + pos.astype(np.float64)
    d = centre - x
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    s = st.species[0]
    vol = st.cell_si[0] * st.cell_si[1] * st.cell_si[2]
    weighting = np.float32(s.density_si * vol / s.ppc)
    mass = float(np.float32(setups.abi_free_constants(st)["base_mass"] * np.float32(s.mass_ratio)) * weighting)
    mom = (d * u * mass).astype(np.float32)
    return cells, pos, mom, w


@pytest.mark.parametrize("graph", [False, True])
def test_full_supercells_grow_without_loss(graph):
    """Supercells filled to their capacity receive more particles than fit:
    the supercells grow, the records waiting in the outboxes land where a
    larger capacity would have put them and the skipped steps run, so the
    result is bit-exact with the oracle run at a capacity that never
    overflows (ParticlesBox.hpp:96-122 takes frames on demand instead)."""
    st = thermal(shape=abi.SHAPE_CIC, frames=1)  # 256 slots
    counts = [256] * 12
    print(f"case capacity growth graph={graph}", flush=True)
    cells, pos, mom, w = converging(st, counts, seed=13, u=1.5)
    f = setups.bin_frames(st, 0, cells, pos, mom, w, frames_per_sc=1)
    sim = abi.Simulation(st.config())
    sim.write_frames(0, f)
    big = copy.deepcopy(st)
    big.species[0].frames_per_sc = 12  # holds every particle of the box in one supercell
    fb = setups.bin_frames(big, 0, cells, pos, mom, w, frames_per_sc=12)
    ora = O.Oracle(big, 32, 0, jexp=sim.j_scale_exp())
    ora.set_frames(0, fb)
    for k in range(3):
        if graph:
            sim.step_graph(2)
        else:
            sim.step(2)
        ora.step(2)
        H.assert_frames_equal(sim.read_frames(0), ora.get_frames(0), f"frames after {2 * k + 2} steps")
        for fld in (abi.FIELD_E, abi.FIELD_B, abi.FIELD_J):
            H.assert_field_equal(sim.read_field(fld), ora.get_field(fld), f"field {fld} after {2 * k + 2} steps")
    assert sim.slots(0) > 256, "the supercells did not grow"
    caps = sim.supercell_slots(0)
    assert np.all(caps % 256 == 0) and caps.max() == sim.slots(0)
    assert sim.error_flags() & abi.FLAG_CAPACITY_GREW
    assert sim.error_flags() == 0  # read and cleared
    assert sim.particle_count(0) == sum(counts)
    sim.close()
    ora.close()


def test_only_the_supercells_that_fill_up_grow():
    """The four full supercells of the middle z layer between sparse ones,
    particles converging on the box centre (which lies in that layer): the
    middle supercells overflow and grow by whole frames, the sparse ones,
    which only lose particles, keep their one frame, and the result is
    bit-exact with the oracle run at a capacity that never overflows."""
    st = thermal(shape=abi.SHAPE_TSC, frames=1)
    counts = [48] * 4 + [256] * 4 + [48] * 4
    print("case per-supercell growth", flush=True)
    cells, pos, mom, w = converging(st, counts, seed=17, u=1.5)
    f = setups.bin_frames(st, 0, cells, pos, mom, w, frames_per_sc=1)
    sim = abi.Simulation(st.config())
    sim.write_frames(0, f)
    big = copy.deepcopy(st)
    big.species[0].frames_per_sc = 12
    fb = setups.bin_frames(big, 0, cells, pos, mom, w, frames_per_sc=12)
    ora = O.Oracle(big, 32, 0, jexp=sim.j_scale_exp())
    ora.set_frames(0, fb)
    sim.step_graph(6)
    ora.step(6)
    H.assert_frames_equal(sim.read_frames(0), ora.get_frames(0), "frames after 6 steps")
    for fld in (abi.FIELD_E, abi.FIELD_B, abi.FIELD_J):
        H.assert_field_equal(sim.read_field(fld), ora.get_field(fld), f"field {fld} after 6 steps")
    caps = sim.supercell_slots(0)
    print(f"capacities after 6 steps: {caps.tolist()}", flush=True)
    assert np.all(caps[4:8] > 256) and np.all(caps[:4] == 256) and np.all(caps[8:] == 256)
    assert np.all(caps % 256 == 0)
    assert sim.particle_count(0) == sum(counts)
    sim.close()
    ora.close()


@pytest.mark.parametrize("field,index,value,fragment", [
    ("cell", None, 300, "localCellIdx"), ("mask", None, 0, "multiMask"), ("pos", 1, 1.0, "position"),
    ("pos", 0, -1e-7, "position"), ("mom", 2, np.nan, "momentum"), ("w", None, np.inf, "weighting")])
def test_frames_that_kernels_cannot_take_are_rejected(field, index, value, fragment):
    """pic_write_frames checks the occupied slots (local cell, multiMask,
    position, finite momentum and weighting) before anything reaches a kernel."""
    st = thermal(frames=2)
    print(f"case rejected frames {field}={value}", flush=True)
    cells, pos, mom, w = crafted(st, [10] * 12, seed=19)
    f = setups.bin_frames(st, 0, cells, pos, mom, w, frames_per_sc=2)
    slot = 3 * 512 + 4  # supercell 3, slot 4: occupied
    arr = f[field]
    if index is None:
        arr[slot] = value
    else:
        arr[index, slot] = value
    sim = abi.Simulation(st.config())
    with pytest.raises(abi.PicError, match=fragment):
        sim.write_frames(0, f)
    sim.close()
