"""Bounds of the fused push/deposit J tile, checked on the CPU.

The fused kernel deposits a particle's current in the frame of its cell
after the move (as the reference's deposit stage does, Esirkepov.hpp:61-242,
EmZ.hpp:61-137) but indexes the tile by the supercell the particle started
in. The tile extends the supercell by the reference's deposit margins,
lowerMargin = SUPP/2 + 1 - (SUPP+1)%2 and upperMargin = (SUPP+1)/2 + 1
(Esirkepov.hpp:42-43), and no further: a particle moves at most one cell per
axis, and the start of its path lies in its old cell, so every cell the
deposit touches lies within those margins of the OLD cell.

This test replays, in float32 with the kernel's operation order, the move
(x0 + dp, moveParticle's -0.5/+0.5 arithmetic, MoveParticle.hpp:47-160), the
reflecting border (Reflecting.hpp:38-66: -x - epsilon above, 2 - x below,
dp negated) and the deposit geometry (start p0 = p1 - dp in the new frame,
grid shift gs = min(iS, iE), one extra row when the shifted cell changes),
and checks the touched cells (Esirkepov rows gs+begin .. gs+end-1+leave, EmZ
rows s+begin .. s+end-1 for s in {iS, iE}) against the margins, for moves
below one cell (|dp| < 1: the host requires float(float(c*dt)/cell) < 1 on
every axis, and dp = float(float(v*dt)/cell) with |v| < c cannot exceed it).
A move of exactly one cell is excluded: x0 = 1 - 2^-24 plus dp = 1 rounds to
2, outside moveParticle's range, and the deposit would reach one cell
further. Beyond the margins LDS atomics would leave the tile.
"""
import numpy as np
import pytest

F = np.float32
EPS = F(1.1920929e-7)

SHAPES = {2: (0, 2), 3: (-1, 2), 4: (-1, 3)}  # SUPP: (begin, end) of ChargeAssignment


def margins(supp):
    return supp // 2 + 1 - (supp + 1) % 2, (supp + 1) // 2 + 1


def move_axis(pos):
    """moveParticle on one axis: pos in [-1, 2) -> [0, 1), returns (pos, step)."""
    p = (pos - F(0.5)).astype(F)
    step = np.where(p < F(-0.5), -1, np.where(p >= F(0.5), 1, 0))
    p = (p - step.astype(F)).astype(F)
    return (p + F(0.5)).astype(F), step


def touched(supp, x0, dp, reflect=None):
    """Touched cell range, relative to the old cell, of the fused deposit."""
    begin, end = SHAPES[supp]
    p1, d = move_axis((x0 + dp).astype(F))
    if reflect == "upper":  # the old cell is the last one; the move left the box upwards
        out = d > 0
        p1 = np.where(out, (-p1 - EPS).astype(F), p1)
        dp = np.where(out, -dp, dp).astype(F)
        p1, d2 = move_axis(p1)
        d = d + np.where(out, d2, 0)
    elif reflect == "lower":
        out = d < 0
        p1 = np.where(out, (F(2.0) - p1).astype(F), p1)
        dp = np.where(out, -dp, dp).astype(F)
        p1, d2 = move_axis(p1)
        d = d + np.where(out, d2, 0)
    p0 = (p1 - dp).astype(F)
    if supp % 2 == 0:
        iS, iE = np.floor(p0).astype(int), np.floor(p1).astype(int)
    else:
        iS, iE = np.floor((p0 + F(0.5)).astype(F)).astype(int), np.floor((p1 + F(0.5)).astype(F)).astype(int)
    gs = np.minimum(iS, iE)
    leave = (iS != iE).astype(int)
    lo = d + gs + begin
    hi = d + np.maximum(gs + end - 1 + leave, np.maximum(iS, iE) + end - 1)
    return lo, hi, d, np.abs(iS - iE)


def touched_own_axis(supp, x0, dp, reflect=None):
    """Upper end of the touched range along a current component's own axis
    (the prefix-sum loop: Esirkepov rows up to gs+end-2+leave, EmZ up to
    s+end-2), relative to the old cell."""
    begin, end = SHAPES[supp]
    lo, hi, d, _ = touched(supp, x0, dp, reflect)
    return hi - 1


def samples(rng, n):
    ulp = np.finfo(F).eps
    x_edge = np.array([0.0, 2.0**-24, 0.5 - ulp / 2, 0.5, 0.5 + ulp, 1.0 - ulp / 2, 0.25, 0.75], F)
    d_edge = np.array([0.0, 2.0**-30, 1.0 - ulp / 2, 0.5, 0.5 - ulp / 2, 2.0**-24, 0.999], F)
    d_edge = np.concatenate([d_edge, -d_edge])
    xe, de = np.meshgrid(x_edge, d_edge)
    x = np.concatenate([xe.ravel(), rng.random(n, dtype=F) * F(1.0 - 2.0**-24)]).astype(F)
    d = np.concatenate([de.ravel(), (rng.random(n, dtype=F) * F(2.0) - F(1.0))]).astype(F)
    x = np.minimum(x, F(1.0 - 2.0**-24))
    return x, np.clip(d, F(-1.0 + 2.0**-24), F(1.0 - 2.0**-24))


@pytest.mark.parametrize("supp", [2, 3, 4])
@pytest.mark.parametrize("reflect", [None, "upper", "lower"])
def test_fused_deposit_stays_inside_margins(supp, reflect):
    rng = np.random.default_rng(1234 + supp)
    x, dp = samples(rng, 400_000)
    lo, hi, d, jump = touched(supp, x, dp, reflect)
    lower, upper = margins(supp)
    assert np.all(np.abs(d) <= 1), "a particle moved more than one cell"
    assert lo.min() >= -lower, f"touched {lo.min()} below the old cell, margin {lower}"
    assert hi.max() <= upper, f"touched {hi.max()} above the old cell, margin {upper}"
    # the margins are tight: both ends are reached
    assert lo.min() == -lower and hi.max() == upper
    # the start and end shifted cells differ by at most one
    assert jump.max() <= 1
    # along a component's own axis the tile reaches one cell less above
    assert touched_own_axis(supp, x, dp, reflect).max() == upper - 1


def test_a_whole_cell_move_leaves_the_margins():
    # why the host rejects float(c*dt)/cell >= 1: x0 = 1 - 2^-24 moved by one
    # cell rounds to 2.0, moveParticle leaves the position at 1.0 and the
    # CIC deposit touches a cell beyond the upper margin
    lo, hi, d, jump = touched(2, np.array([1.0 - 2.0**-24], F), np.array([1.0], F))
    assert hi[0] == margins(2)[1] + 1


def test_margins_are_the_reference_ones():
    # Esirkepov.hpp:42-43 for CIC, TSC, PQS
    assert [margins(s) for s in (2, 3, 4)] == [(1, 2), (2, 3), (2, 3)]
