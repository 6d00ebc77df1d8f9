"""The oracle's EmZ current deposition (currentDeposition/EmZ/EmZ.hpp:61-137,
EmZ/DepositCurrent.hpp:46-112), pinned two ways: discrete continuity
rho1 - rho0 + dt div J = 0 with FP64 B-spline charge densities, and the
decomposition that defines EmZ: its current equals the Esirkepov (2001)
closed form (FP64) summed over the one or two straight segments the relay
point (relayPoint.hpp:47-64) splits the move into."""
import math

import numpy as np
import pytest

from oracle import oracle as O
from picongpu_amd import abi, setups
from tests.test_oracle_pinning import div_j, esirkepov_closed_form, rho_nodes

SUPP = {abi.SHAPE_CIC: 2, abi.SHAPE_TSC: 3, abi.SHAPE_PQS: 4}


def moves(shape, n):
    rng = np.random.default_rng(11 + shape)
    for _ in range(n):
        pos = rng.random(3).astype(np.float32)
        v = rng.uniform(-1, 1, 3)
        v = (v / max(np.linalg.norm(v), 1e-9) * rng.uniform(0, 0.95)).astype(np.float32)
        yield pos, v


def relay(shape, p0, p1):
    """Relay point per axis in cell units (relayPoint.hpp:47-64), FP64."""
    out = np.empty(3)
    split = False
    for d in range(3):
        if SUPP[shape] % 2 == 0:
            i0, i1 = math.floor(p0[d]), math.floor(p1[d])
            out[d] = p1[d] if i0 == i1 else max(i0, i1)
        else:
            i0, i1 = math.floor(p0[d] + 0.5), math.floor(p1[d] + 0.5)
            out[d] = p1[d] if i0 == i1 else (i0 + i1) / 2.0
        split |= i0 != i1
    return out, split


@pytest.mark.parametrize("shape", [abi.SHAPE_CIC, abi.SHAPE_TSC, abi.SHAPE_PQS])
def test_emz_continuity(shape):
    st = setups.thermal(grid=(24, 24, 24))
    k = O.constants(st)
    dt = float(k.dt)
    cell = np.array(k.cell, np.float64)
    c = np.array([12, 12, 12], np.int32)
    q = np.float32(-1.0)
    for trial, (pos, v) in enumerate(moves(shape, 40)):
        J = O.deposit_single(st, shape, abi.CURRENT_SOLVER_EMZ, c, pos, v, q)
        d = (v * np.float32(dt)) / cell.astype(np.float32)
        p0 = (pos - d).astype(np.float32)
        win = ((6, 19), (6, 19), (6, 19))
        r0 = rho_nodes(st, shape, c + p0.astype(np.float64), float(q), win)
        r1 = rho_nodes(st, shape, c + pos.astype(np.float64), float(q), win)
        resid = r1 - r0 + dt * div_j(st, J, win)
        assert np.abs(resid).max() <= 2e-6 * np.abs(r1).max(), trial


@pytest.mark.parametrize("shape", [abi.SHAPE_CIC, abi.SHAPE_TSC, abi.SHAPE_PQS])
def test_emz_is_esirkepov_over_the_relay_segments(shape):
    st = setups.thermal(grid=(24, 24, 24))
    k = O.constants(st)
    dt = float(k.dt)
    cell = np.array(k.cell, np.float64)
    c = np.array([12, 12, 12], np.int32)
    q = -1.0
    splits = 0
    for pos, v in moves(shape, 30):
        J = O.deposit_single(st, shape, abi.CURRENT_SOLVER_EMZ, c, pos, v, np.float32(q))
        d = (v * np.float32(dt)) / cell.astype(np.float32)
        p0 = (pos - d).astype(np.float32).astype(np.float64)
        p1 = pos.astype(np.float64)
        r, split = relay(shape, p0, p1)
        splits += split
        want = esirkepov_closed_form(shape, c, r, r - p0, q, cell, dt, st.grid)
        if split:
            want = want + esirkepov_closed_form(shape, c, p1, p1 - r, q, cell, dt, st.grid)
        # FP32 differences of nearly equal shape values bound the agreement
        scale = np.abs(want).max()
        np.testing.assert_allclose(J, want, rtol=0, atol=1e-4 * scale)
    assert splits > 5  # the sample crosses cells


@pytest.mark.parametrize("shape", [abi.SHAPE_CIC, abi.SHAPE_TSC, abi.SHAPE_PQS])
def test_emz_equals_esirkepov_inside_one_cell(shape):
    st = setups.thermal(grid=(24, 24, 24))
    c = np.array([12, 12, 12], np.int32)
    k = O.constants(st)
    # a slow particle inside its assignment cell ([0,1) for even support,
    # [-0.5,0.5) around the cell origin for odd) does not cross
    pos = np.full(3, 0.5 if SUPP[shape] % 2 == 0 else 0.25, np.float32)
    v = np.array([0.05, -0.03, 0.02], np.float32)
    a = O.deposit_single(st, shape, abi.CURRENT_SOLVER_EMZ, c, pos, v, np.float32(-1.0))
    b = O.esirkepov_single(st, shape, c, pos, v, np.float32(-1.0))
    np.testing.assert_allclose(a, b, rtol=0, atol=1e-6 * np.abs(b).max())
    assert np.count_nonzero(a) == np.count_nonzero(b)
