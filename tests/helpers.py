"""Shared helpers of the parity tests: identical seeded inputs for the HIP
library and the CPU oracle, and bit-level comparison of their states."""
from __future__ import annotations

import numpy as np

from oracle import oracle as O
from picongpu_amd import abi, setups


def bits(a) -> np.ndarray:
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def random_fields(setup, seed, e_amp, b_amp):
    rng = np.random.default_rng(seed)
    n = setup.ncells
    E = (rng.standard_normal((3, n)) * e_amp).astype(np.float32)
    B = (rng.standard_normal((3, n)) * b_amp).astype(np.float32)
    return E, B


def make_pair(setup, e_amp=1e-3, b_amp=1e-3, seed=11, order=0, with_sim=True):
    """Simulation (GPU) and Oracle loaded with the same particles and fields."""
    frames = []
    for sp in range(len(setup.species)):
        cells, pos, mom, w = setups.host_particles(setup, sp)
        frames.append(setups.bin_frames(setup, sp, cells, pos, mom, w))
    E, B = random_fields(setup, seed, e_amp, b_amp)
    sim = None
    jexp = 40
    if with_sim:
        sim = abi.Simulation(setup.config())
        for sp, f in enumerate(frames):
            sim.write_frames(sp, f)
        sim.load_field(abi.FIELD_E, E)
        sim.load_field(abi.FIELD_B, B)
        jexp = sim.j_scale_exp()
    ora = O.Oracle(setup, 32, order, jexp=jexp)
    for sp, f in enumerate(frames):
        ora.set_frames(sp, f)
    ora.set_field(abi.FIELD_E, E)
    ora.set_field(abi.FIELD_B, B)
    return sim, ora, frames


def occupied(frames) -> np.ndarray:
    """Mask of the occupied slots [0, count) of every supercell of a frame pool
    (its slots per supercell follow from the array sizes)."""
    cap = frames["w"].shape[0] // frames["counts"].shape[0]
    return (np.arange(cap)[None, :] < frames["counts"][:, None].astype(np.int64)).ravel()


def assert_frames_equal(gpu, ora, what=""):
    """Slot-exact equality of two frame pools (float compared bitwise). The two
    pools may have different slots per supercell (a grown pool against one
    sized large enough from the start): the occupied slots of each supercell
    are compared in slot order."""
    np.testing.assert_array_equal(gpu["counts"], ora["counts"], err_msg=f"{what}: supercell counts")
    vg, vo = occupied(gpu), occupied(ora)
    for key in ("pos", "mom"):
        g = bits(gpu[key][:, vg])
        o = bits(np.asarray(ora[key][:, vo], np.float32))
        bad = np.nonzero(g != o)
        assert bad[0].size == 0, f"{what}: {key} differs at {bad[0].size} entries, first {bad[1][:5]}"
    np.testing.assert_array_equal(bits(gpu["w"][vg]), bits(np.asarray(ora["w"][vo], np.float32)))
    np.testing.assert_array_equal(gpu["cell"][vg], ora["cell"][vo], err_msg=f"{what}: localCellIdx")
    np.testing.assert_array_equal(gpu["mask"][vg], ora["mask"][vo], err_msg=f"{what}: multiMask")


def assert_field_equal(gpu, ora, what=""):
    g = bits(gpu)
    o = bits(np.asarray(ora, np.float32))
    bad = np.nonzero(g != o)
    assert bad[0].size == 0, (
        f"{what}: {bad[0].size} of {g.size} values differ; max abs diff "
        f"{np.max(np.abs(gpu.astype(np.float64) - ora)) if bad[0].size else 0}")
