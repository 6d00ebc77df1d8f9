"""Batched ensembles: every member advanced by the batched launches must equal
the same simulation advanced on its own with pic_step, bit for bit (frames,
E, B, J), whatever the schedule, eager or replayed from graphs, and the
oracle run on the member's initial state."""
import copy

import numpy as np
import pytest

from oracle import oracle as O
from picongpu_amd import abi, setups
from tests import helpers as H

pytestmark = pytest.mark.gpu


def members():
    # at least 2x2x3 supercells: the TSC fused checkerboard needs stride 2, 2, 3
    return [setups.two_stream(grid=(16, 16, 12), ppc=2, beta=0.3, seed=11),
            setups.two_stream(grid=(24, 16, 16), ppc=3, beta=0.2, seed=12),
            setups.two_stream(grid=(16, 24, 12), ppc=4, beta=0.35, seed=13)]


def init(sim, st):
    setups.init_device(sim, st)


def compare(ens_member, single, what):
    H.assert_frames_equal(ens_member.read_frames(0), single.read_frames(0), f"{what} frames")
    for f in (abi.FIELD_E, abi.FIELD_B, abi.FIELD_J):
        H.assert_field_equal(ens_member.read_field(f), single.read_field(f).astype(np.float64), f"{what} field {f}")


def singles(sts, nsteps):
    out = []
    for st in sts:
        s = abi.Simulation(st.config())
        init(s, st)
        s.step(nsteps)
        out.append(s)
    return out


@pytest.mark.parametrize("mode", ["eager", "graph"])
def test_ensemble_equals_single_runs(mode):
    sts = members()
    nsteps = 6
    print(f"case ensemble {mode} members={[s.grid for s in sts]} steps={nsteps}", flush=True)
    ens = abi.Ensemble([st.config() for st in sts])
    for m, st in zip(ens.members, sts):
        init(m, st)
    if mode == "eager":
        ens.step(nsteps)
    else:
        ens.step_graph(nsteps)
    ref = singles(sts, nsteps)
    for i, (m, s) in enumerate(zip(ens.members, ref)):
        compare(m, s, f"member {i}")
        assert m.error_flags() == 0
        s.close()
    ens.close()


def test_ensemble_schedules_and_autotune_do_not_change_results():
    sts = members()
    print("case ensemble schedules + autotune", flush=True)
    ens = abi.Ensemble([st.config() for st in sts])
    for m, st in zip(ens.members, sts):
        init(m, st)
    taken = 0
    for order in (abi.ENSEMBLE_ORDER_SIM_MAJOR, abi.ENSEMBLE_ORDER_INTERLEAVED, abi.ENSEMBLE_ORDER_HEAVY_FIRST):
        for bpc in (0, 1):
            ens.set_schedule(order, bpc)
            ens.step(1)
            taken += 1
    ms, n = ens.autotune(trial_steps=1)
    taken += n
    assert n == 12 * 2 and ms > 0.0
    ens.step_graph(1)
    taken += 1
    ref = singles(sts, taken)
    for i, (m, s) in enumerate(zip(ens.members, ref)):
        compare(m, s, f"member {i} after {taken} steps")
        s.close()
    ens.close()


def test_member_stepped_alone_in_pairs():
    """A member may be stepped on its own between ensemble steps in even
    numbers; an odd number is refused."""
    sts = members()[:2]
    print("case ensemble member stepped alone", flush=True)
    ens = abi.Ensemble([st.config() for st in sts])
    for m, st in zip(ens.members, sts):
        init(m, st)
    ens.step(1)
    ens.members[0].step(2)
    ens.step_graph(1)
    ens.members[1].step(1)
    with pytest.raises(abi.PicError):
        ens.step(1)
    ens.members[1].step(1)
    ens.step(1)
    ref = singles(sts, 0)
    ref[0].step(5)
    ref[1].step(5)
    for i, (m, s) in enumerate(zip(ens.members, ref)):
        compare(m, s, f"member {i}")
        s.close()
    ens.close()


def oracle_of(member, st, frames=None):
    """The oracle loaded with the member's current frames and fields (frames:
    slots per supercell of the oracle's pool, default the member's)."""
    big = copy.deepcopy(st)
    if frames is not None:
        big.species[0].frames_per_sc = frames
    ora = O.Oracle(big, 32, 0, jexp=member.j_scale_exp())
    f = member.read_frames(0)
    if frames is not None:
        v = H.occupied(f)
        counts = f["counts"]
        g = {"counts": counts}
        cap = frames * 256
        nsc = counts.shape[0]
        slot = (np.repeat(np.arange(nsc, dtype=np.int64) * cap, counts.astype(np.int64)) +
                np.concatenate([np.arange(c, dtype=np.int64) for c in counts]))
        for key, shape, dt in (("pos", 3, np.float32), ("mom", 3, np.float32), ("w", 0, np.float32),
                               ("cell", 0, np.uint16), ("mask", 0, np.uint8)):
            if shape:
                a = np.zeros((shape, nsc * cap), dt)
                a[:, slot] = f[key][:, v]
            else:
                a = np.zeros(nsc * cap, dt)
                a[slot] = f[key][v]
            g[key] = a
        f = g
    ora.set_frames(0, f)
    for fld in (abi.FIELD_E, abi.FIELD_B, abi.FIELD_J):
        ora.set_field(fld, member.read_field(fld))
    return ora


def compare_oracle(member, ora, what):
    H.assert_frames_equal(member.read_frames(0), ora.get_frames(0), f"{what} frames")
    for f in (abi.FIELD_E, abi.FIELD_B, abi.FIELD_J):
        H.assert_field_equal(member.read_field(f), ora.get_field(f), f"{what} field {f}")


@pytest.mark.parametrize("mode", ["eager", "graph"])
def test_ensemble_members_match_oracle(mode):
    """Each member's batched steps against the oracle started from the
    member's initial state (generated on the device), bit for bit."""
    sts = members()
    nsteps = 5
    print(f"case ensemble vs oracle {mode}", flush=True)
    ens = abi.Ensemble([st.config() for st in sts])
    oras = []
    for m, st in zip(ens.members, sts):
        init(m, st)
        oras.append(oracle_of(m, st))
    if mode == "eager":
        ens.step(nsteps)
    else:
        ens.step_graph(nsteps)
    for i, (m, o) in enumerate(zip(ens.members, oras)):
        o.step(nsteps)
        compare_oracle(m, o, f"member {i}")
        o.close()
    ens.close()


def test_ensemble_member_grows_without_loss():
    """A member whose supercells start at exactly their particle count (one
    frame per 256 particles) overflows in the first shift: it grows, runs the
    skipped steps on its own, and still matches the oracle run at a capacity
    that never overflows; the other members are untouched."""
    sts = members()
    sts[1].species[0].frames_per_sc = 3  # 3 ppc x 256 cells: exactly full
    print("case ensemble member growth", flush=True)
    ens = abi.Ensemble([st.config() for st in sts])
    oras = []
    for m, st in zip(ens.members, sts):
        init(m, st)
        oras.append(oracle_of(m, st, frames=24))
    ens.step_graph(3)
    ens.step(3)
    assert ens.members[1].slots(0) > 3 * 256
    assert ens.members[1].error_flags() & abi.FLAG_CAPACITY_GREW
    for i, (m, o) in enumerate(zip(ens.members, oras)):
        o.step(6)
        compare_oracle(m, o, f"member {i}")
        o.close()
    ens.close()
