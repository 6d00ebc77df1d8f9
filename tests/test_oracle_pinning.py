"""Pins the CPU oracle against the reference's own tests and against what the
mathematics fixes, independently of the HIP library.

* moveParticle: the seven exact cases of share/picongpu/unit/MoveParticle.cpp:98-202
* shapes: partition of unity within one epsilon on 1024 uniform positions
  (share/picongpu/unit/shape.cpp:112-207) and the textbook B-splines
* Esirkepov: discrete continuity rho1 - rho0 + dt div J = 0 for single particles
  with FP64 B-spline charge densities, and the closed-form W of Esirkepov (2001)
  on the CurrentDeposition setup (share/picongpu/tests/CurrentDeposition: cell
  (8,8,8), offset (0.9,0.8,0.7), beta 0.999 along (1,0,0),(1,1,0),(1,1,1),
  24^3, dt = dx/(2c); absolute tolerance 1e-5 of the reference, applied
  relative to max |J| here)
* Boris/Vay: gyration in uniform B (share/picongpu/tests/Pusher criteria:
  radius drift per turn < 1e-5 from momentum, < 5e-5 from position, phase
  error < 0.16 rad per turn at 50 steps per turn) and phase lag ~ dt^2
  (tests/PusherScaling: exponent 2 +- 0.1)
* Yee/Lehe: standing-wave frequency against the discrete dispersion relations
* whole cycle: charge conservation to round-off, FP32 vs FP64 twin
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from picongpu_amd import abi, setups

EPS32 = float(np.finfo(np.float32).eps)


# ---------------------------------------------------------------------------
# moveParticle
# ---------------------------------------------------------------------------
def test_move_particle_reference_cases():
    # unchanged position
    p, lc, m, r = O.move_particle([0.0, 0.0, 0.0], 0, 1)
    assert (lc, m, r) == (0, 1, 1) and np.all(p == 0)
    # trivial move inside the cell
    for i in range(3):
        q = [0.0, 0.0, 0.0]
        q[i] = 0.42
        p, lc, m, r = O.move_particle(q, 0, 1)
        assert lc == 0 and m == 1 and r < 2
        assert abs(p[i] - 0.42) <= EPS32
    # out of the cell in +direction: localCellIdx 1, 8, 64
    for i, expect in enumerate([1, 8, 64]):
        q = [0.0, 0.0, 0.0]
        q[i] = 1.1
        p, lc, m, r = O.move_particle(q, 0, 1)
        assert lc == expect and m == 1 and r < 2
        assert abs(p[i] - 0.1) <= 4 * EPS32
    # out of the cell in -direction from the last cell: 254, 247, 191
    for i, expect in enumerate([254, 255 - 8, 191]):
        q = [0.0, 0.0, 0.0]
        q[i] = -0.3
        p, lc, m, r = O.move_particle(q, 255, 1)
        assert lc == expect and m == 1 and r < 2
        assert abs(p[i] - 0.7) <= 4 * EPS32
    # diagonal
    p, lc, m, r = O.move_particle([1.4, 1.6, 0.0], 0, 1)
    assert lc == 9 and m == 1
    assert abs(p[0] - 0.4) <= 4 * EPS32 and abs(p[1] - 0.6) <= 4 * EPS32
    # leaving the supercell towards -x: multiMask 3, localCellIdx 7
    p, lc, m, r = O.move_particle([-0.9, 0.0, 0.0], 0, 1)
    assert lc == 7 and m == 3 and r == 3
    assert abs(p[0] - 0.1) <= 4 * EPS32
    # rounding near zero: -eps/4 stays in the cell at 0
    p, lc, m, r = O.move_particle([-EPS32 / 4, 0.0, 0.0], 0, 1)
    assert lc == 0 and m == 1 and p[0] == 0.0


# ---------------------------------------------------------------------------
# shapes
# ---------------------------------------------------------------------------
def bspline(shape, x):
    a = np.abs(np.asarray(x, np.float64))
    if shape == abi.SHAPE_CIC:
        return np.where(a < 1, 1 - a, 0.0)
    if shape == abi.SHAPE_TSC:
        return np.where(a < 0.5, 0.75 - a * a, np.where(a < 1.5, 0.5 * (1.5 - a) ** 2, 0.0))
    return np.where(a < 1, (4 - 6 * a * a + 3 * a ** 3) / 6, np.where(a < 2, (2 - a) ** 3 / 6, 0.0))


SUPP = {abi.SHAPE_CIC: 2, abi.SHAPE_TSC: 3, abi.SHAPE_PQS: 4}
BEGIN = {abi.SHAPE_CIC: 0, abi.SHAPE_TSC: -1, abi.SHAPE_PQS: -1}


def position_shift(pos, even):
    v = np.float32(pos) - np.float32(0.5)
    s = (0 if v >= -0.5 else -1) if even else (1 if v >= 0.0 else 0)
    return np.float32(v - np.float32(s) + np.float32(0.5))


@pytest.mark.parametrize("shape", [abi.SHAPE_CIC, abi.SHAPE_TSC, abi.SHAPE_PQS])
def test_shape_partition_of_unity(shape):
    rng = np.random.default_rng(42)
    pos = rng.random(1024).astype(np.float32)
    supp = SUPP[shape]
    even = supp % 2 == 0
    # Catch::Approx(1).margin(eps): absolute eps or Catch's default relative 100*eps
    tol = max(EPS32, 100 * EPS32)
    for x in pos:
        # ChargeAssignment over [begin, end] at the unshifted position, summed in float
        tot = np.float32(0.0)
        for g in range(BEGIN[shape], BEGIN[shape] + supp + 1):
            tot = np.float32(tot + O.shape(shape, 0, np.float32(g) - x)[0])
        assert abs(float(tot) - 1.0) <= tol, (shape, x, tot)
        # on-support array at the shifted position
        xs = position_shift(x, even)
        arr = O.shape(shape, 1, xs)[:supp]
        assert abs(float(np.sum(arr, dtype=np.float32)) - 1.0) <= tol
        # values against the textbook B-spline (grid points begin.. relative to the shifted cell)
        b = -(supp // 2) + (supp + 1) % 2
        ref = bspline(shape, np.arange(b, b + supp) - np.float64(xs))
        np.testing.assert_allclose(arr, ref, rtol=0, atol=4 * EPS32)


@pytest.mark.parametrize("shape", [abi.SHAPE_CIC, abi.SHAPE_TSC, abi.SHAPE_PQS])
def test_cached_shape_array_out_of_range(shape):
    supp = SUPP[shape]
    b = BEGIN[shape]
    for x in np.linspace(0.0, 0.99, 23).astype(np.float32):
        base = x if supp % 2 == 0 else np.float32(x - 0.5)
        a_in = O.shape(shape, 2, base, 0)[: supp + 1]
        a_out = O.shape(shape, 2, np.float32(base + 1.0), 1)[: supp + 1]
        grid = np.arange(b, b + supp + 1)
        np.testing.assert_allclose(a_in, bspline(shape, grid - np.float64(base)), atol=4 * EPS32)
        np.testing.assert_allclose(a_out, bspline(shape, grid - (np.float64(base) + 1.0)), atol=4 * EPS32)


# ---------------------------------------------------------------------------
# Esirkepov
# ---------------------------------------------------------------------------
def rho_nodes(setup, shape, gpos, q, window):
    """FP64 charge density q/V * prod S(node - x) on the window of nodes."""
    k = O.constants(setup)
    vol = float(k.cell[0]) * float(k.cell[1]) * float(k.cell[2])
    (x0, x1), (y0, y1), (z0, z1) = window
    X, Y, Z = np.meshgrid(np.arange(x0, x1), np.arange(y0, y1), np.arange(z0, z1), indexing="ij")
    return q / vol * bspline(shape, X - gpos[0]) * bspline(shape, Y - gpos[1]) * bspline(shape, Z - gpos[2])


def div_j(setup, J, window):
    k = O.constants(setup)
    nx, ny, nz = setup.grid
    Jg = J.reshape(3, nz, ny, nx)
    (x0, x1), (y0, y1), (z0, z1) = window
    out = np.zeros((x1 - x0, y1 - y0, z1 - z0))
    for ix, x in enumerate(range(x0, x1)):
        for iy, y in enumerate(range(y0, y1)):
            for iz, z in enumerate(range(z0, z1)):
                d = (Jg[0, z, y, x] - Jg[0, z, y, x - 1]) / float(k.cell[0])
                d += (Jg[1, z, y, x] - Jg[1, z, y - 1, x]) / float(k.cell[1])
                d += (Jg[2, z, y, x] - Jg[2, z - 1, y, x]) / float(k.cell[2])
                out[ix, iy, iz] = d
    return out


@pytest.mark.parametrize("shape", [abi.SHAPE_CIC, abi.SHAPE_TSC, abi.SHAPE_PQS])
def test_esirkepov_single_particle_continuity(shape):
    st = setups.thermal(grid=(24, 24, 24))
    k = O.constants(st)
    rng = np.random.default_rng(7 + shape)
    dt = float(k.dt)
    cell = np.array(k.cell, np.float64)
    for trial in range(40):
        c = np.array([12, 12, 12], np.int32)
        pos = rng.random(3).astype(np.float32)
        # velocities up to 0.9 c per component, total below c
        v = rng.uniform(-1, 1, 3)
        v = (v / max(np.linalg.norm(v), 1e-9) * rng.uniform(0, 0.95)).astype(np.float32)
        q = np.float32(-1.0)
        J = O.esirkepov_single(st, shape, c, pos, v, q)
        # start position as the deposit reconstructs it (pos - v dt / cell), FP32
        d = (v.astype(np.float32) * np.float32(dt)) / cell.astype(np.float32)
        p0 = (pos - d).astype(np.float32)
        g1 = c + pos.astype(np.float64)
        g0 = c + p0.astype(np.float64)
        win = ((6, 19), (6, 19), (6, 19))
        r0 = rho_nodes(st, shape, g0, float(q), win)
        r1 = rho_nodes(st, shape, g1, float(q), win)
        dj = div_j(st, J, win)
        resid = r1 - r0 + dt * dj
        scale = np.abs(r1).max()
        assert np.abs(resid).max() <= 2e-6 * scale, (trial, np.abs(resid).max() / scale)


def esirkepov_closed_form(shape, c, p1, d, q, cell, dt, grid):
    """Esirkepov (2001) W formula in FP64 for one straight move from
    x0 = c + p1 - d to x1 = c + p1 (cell units); J_d(n) located at
    n + e_d/2 as the Yee cell of the reference."""
    x1 = np.asarray(c, np.float64) + np.asarray(p1, np.float64)
    x0 = x1 - np.asarray(d, np.float64)
    vol = cell[0] * cell[1] * cell[2]
    J = np.zeros((3,) + tuple(grid[::-1]))
    lo = np.floor(np.minimum(x0, x1)).astype(int) - 3
    hi = np.floor(np.maximum(x0, x1)).astype(int) + 4
    nodes = [np.arange(lo[a], hi[a]) for a in range(3)]
    S0 = [bspline(shape, nodes[a] - x0[a]) for a in range(3)]
    S1 = [bspline(shape, nodes[a] - x1[a]) for a in range(3)]
    DS = [S1[a] - S0[a] for a in range(3)]
    for comp in range(3):
        a, b = [x for x in range(3) if x != comp]
        W = np.einsum("i,j->ij", S0[a], S0[b]) + 0.5 * np.einsum("i,j->ij", DS[a], S0[b]) \
            + 0.5 * np.einsum("i,j->ij", S0[a], DS[b]) + np.einsum("i,j->ij", DS[a], DS[b]) / 3.0
        pref = -q / (vol * dt) * cell[comp]
        cum = np.cumsum(DS[comp])
        # W(n_comp, n_a, n_b) = DS(comp) prefix * W_ab ; J at node n located at n + 1/2 along comp
        for ia, na in enumerate(nodes[a]):
            for ib, nb in enumerate(nodes[b]):
                for ic, ncomp in enumerate(nodes[comp]):
                    idx = [0, 0, 0]
                    idx[comp] = ncomp
                    idx[a] = na
                    idx[b] = nb
                    J[comp, idx[2] % grid[2], idx[1] % grid[1], idx[0] % grid[0]] += pref * W[ia, ib] * cum[ic]
    return J.reshape(3, -1)


@pytest.mark.parametrize("shape", [abi.SHAPE_CIC, abi.SHAPE_TSC, abi.SHAPE_PQS])
@pytest.mark.parametrize("direction", [(1, 0, 0), (1, 1, 0), (1, 1, 1)])
def test_esirkepov_current_deposition_setup(shape, direction):
    # the reference's CurrentDeposition test: 24^3, dt = dx/(2c), one electron
    # at cell (8,8,8) offset (0.9,0.8,0.7) moving at beta = 0.999
    cell_si = (1e-7, 1e-7, 1e-7)
    st = setups.thermal(grid=(24, 24, 24))
    st.cell_si = cell_si
    st.dt_si = 0.5 * cell_si[0] / setups.SPEED_OF_LIGHT_SI
    k = O.constants(st)
    cell = np.array(k.cell, np.float64)
    dt = float(k.dt)
    u = np.array(direction, np.float64)
    v = (0.999 * u / np.linalg.norm(u)).astype(np.float32)
    start = np.array([8.9, 8.8, 8.7], np.float32)
    d = (v * np.float32(dt)) / cell.astype(np.float32)
    endg = start.astype(np.float32) + d  # position after the free push
    c = np.floor(endg).astype(np.int32)
    p1 = (endg - c).astype(np.float32)
    q = np.float32(-1.0)
    J = O.esirkepov_single(st, shape, c, p1, v, q)
    Jref = esirkepov_closed_form(shape, c, p1, d.astype(np.float64), float(q), cell, dt, st.grid)
    scale = np.abs(Jref).max()
    assert np.abs(J - Jref).max() <= 1e-5 * scale
    assert np.abs(J).sum() > 0


def test_esirkepov_stationary_particle_no_current():
    st = setups.thermal(grid=(24, 24, 24))
    for shape in (abi.SHAPE_CIC, abi.SHAPE_TSC, abi.SHAPE_PQS):
        J = O.esirkepov_single(st, shape, np.array([5, 6, 7], np.int32), np.float32([0.3, 0.6, 0.1]),
                               np.zeros(3, np.float32), np.float32(-1.0))
        assert not np.any(J)


# ---------------------------------------------------------------------------
# pushers
# ---------------------------------------------------------------------------
def single_particle_oracle(pusher, beta=0.5, steps_per_turn=50, precision=64):
    st = setups.thermal(grid=(16, 16, 12), ppc=1, pusher=pusher)
    ora = O.Oracle(st, precision, 1)
    k = O.constants(st)
    wtyp = k.typical_weighting
    w = np.float32(wtyp)
    mass = float(np.float32(k.species_mass[0]) * w)
    charge = float(np.float32(k.species_charge[0]) * w)
    gamma = 1.0 / math.sqrt(1 - beta * beta)
    p = np.float32(gamma * beta * mass)
    omega = 2 * math.pi / steps_per_turn  # per step, dt = 1
    b0 = omega * gamma * mass / abs(charge)
    cells = np.array([[8, 8, 6]], np.int32)
    pos = np.array([[0.5, 0.5, 0.5]], np.float32)
    mom = np.array([[0.0, p, 0.0]], np.float32)
    f = setups.bin_frames(st, 0, cells, pos, mom, np.array([w], np.float32))
    ora.set_frames(0, f)
    B = np.zeros((3, st.ncells), np.float32)
    B[2, :] = b0
    ora.set_field(abi.FIELD_B, B)
    return st, ora, k, b0, charge, mass, gamma


def track(st, ora, k):
    f = ora.get_frames(0)
    cap = ora.cap(0)
    sc = int(np.nonzero(f["counts"])[0][0])
    slot = sc * cap
    nscx, nscy = st.grid[0] // 8, st.grid[1] // 8
    sx, sy, sz = sc % nscx, (sc // nscx) % nscy, sc // (nscx * nscy)
    lc = int(f["cell"][slot])
    g = np.array([sx * 8 + lc % 8, sy * 8 + (lc // 8) % 8, sz * 4 + lc // 64], np.float64) + f["pos"][:, slot]
    return g * np.array(k.cell, np.float64), f["mom"][:, slot].copy()


@pytest.mark.parametrize("pusher", [abi.PUSHER_BORIS, abi.PUSHER_VAY, abi.PUSHER_HIGUERA_CARY])
@pytest.mark.parametrize("precision", [32, 64])
def test_pusher_gyration(pusher, precision):
    st, ora, k, b0, charge, mass, gamma = single_particle_oracle(pusher, precision=precision)
    n = 50
    L = np.array(st.grid, np.float64) * np.array(k.cell, np.float64)
    x0, p0 = track(st, ora, k)
    pts, moms = [x0], [p0]
    for _ in range(3 * n):
        ora.push(-1, False)
        ora.shift(-1)
        x, p = track(st, ora, k)
        pts.append(x)
        moms.append(p)
    pts = np.array(pts)
    # unwrap the periodic box
    for i in range(1, len(pts)):
        pts[i] = pts[i - 1] + ((pts[i] - pts[i - 1] + L / 2) % L - L / 2)
    moms = np.array(moms)
    r_mom = np.linalg.norm(moms[:, :2], axis=1) / (abs(charge) * b0)
    # radius from momentum: relative change per turn below 1e-5
    for t in range(1, 4):
        assert abs(r_mom[t * n] - r_mom[(t - 1) * n]) / r_mom[0] < 1e-5
    # radius from positions about the orbit centre, per turn below 5e-5
    centre = pts[:n].mean(axis=0)
    r_pos = np.linalg.norm((pts - centre)[:, :2], axis=1)
    for t in range(1, 3):
        a = r_pos[(t - 1) * n:t * n].mean()
        b = r_pos[t * n:(t + 1) * n].mean()
        assert abs(b - a) / a < 5e-5
    # phase error per turn against the exact cyclotron rotation
    ang = np.unwrap(np.arctan2(moms[:, 1], moms[:, 0]))
    per_turn = abs(abs(ang[n] - ang[0]) - 2 * math.pi)
    assert per_turn < 0.16


@pytest.mark.parametrize("precision", [32, 64])
def test_relativistic_exb_drift(precision):
    """Crossed E (y) and B (z) with |E| = 0.6 c |B|: a particle moving at the
    drift velocity E x B / B^2 = 0.6 c x with its Lorentz factor feels no net
    force. Vay and Higuera-Cary keep that momentum (both are built to give the
    relativistic E x B drift exactly); Boris does not (its half-step velocity
    from the averaged momentum misses the drift at relativistic speed), which
    shows the check tells the pushers apart."""
    err = {}
    for pusher in (abi.PUSHER_BORIS, abi.PUSHER_VAY, abi.PUSHER_HIGUERA_CARY):
        st, ora, k, b0, charge, mass, gamma = single_particle_oracle(pusher, precision=precision)
        bd = 0.6
        gd = 1.0 / math.sqrt(1.0 - bd * bd)
        E = np.zeros((3, st.ncells), np.float32)
        E[1, :] = bd * k.c * b0
        ora.set_field(abi.FIELD_E, E)
        f = ora.get_frames(0)
        cap = ora.cap(0)
        sc = int(np.nonzero(f["counts"])[0][0])
        f["mom"][:, sc * cap] = [gd * mass * bd * k.c, 0.0, 0.0]
        ora.set_frames(0, {key: (v.astype(np.float32) if v.dtype == np.float64 else v) for key, v in f.items()})
        _, p0 = track(st, ora, k)
        worst = 0.0
        for _ in range(200):
            ora.push(-1, False)
            ora.shift(-1)
            _, p = track(st, ora, k)
            worst = max(worst, float(np.abs(p - p0).max() / np.abs(p0).max()))
        err[pusher] = worst
        ora.close()
    tol = 1e-6 if precision == 64 else 5e-6
    assert err[abi.PUSHER_VAY] < tol and err[abi.PUSHER_HIGUERA_CARY] < tol, err
    assert err[abi.PUSHER_BORIS] > 1e-3, err


def test_boris_phase_lag_scales_with_dt_squared():
    lags, dts = [], []
    for spt in (160, 80, 40, 20, 10):
        st, ora, k, b0, charge, mass, gamma = single_particle_oracle(abi.PUSHER_BORIS, steps_per_turn=spt)
        _, p0 = track(st, ora, k)
        for _ in range(spt):
            ora.push(-1, False)
            ora.shift(-1)
        _, p1 = track(st, ora, k)
        a0 = math.atan2(p0[1], p0[0])
        a1 = math.atan2(p1[1], p1[0])
        lag = abs(((a1 - a0 + math.pi) % (2 * math.pi)) - math.pi)
        lags.append(lag)
        dts.append(1.0 / spt)
    slope = np.polyfit(np.log(dts), np.log(lags), 1)[0]
    assert abs(slope - 2.0) < 0.1, slope


# ---------------------------------------------------------------------------
# field solvers
# ---------------------------------------------------------------------------
def standing_wave(solver, axis, mode=3, steps=40, precision=64):
    grid = [8, 8, 4]
    grid[axis] = 64 if axis < 2 else 64
    st = setups.thermal(grid=tuple(grid), ppc=1)
    st.species = []
    st.solver = solver
    st.lehe_dir = 1
    if solver == abi.SOLVER_LEHE:
        st.cell_si = (1e-7, 5e-8, 1e-7)
        st.dt_si = 0.99 * st.cell_si[1] / setups.SPEED_OF_LIGHT_SI
    ora = O.Oracle(st, precision, 1)
    k = O.constants(st)
    n = st.ncells
    idx = np.arange(n)
    coord = [idx % grid[0], (idx // grid[0]) % grid[1], idx // (grid[0] * grid[1])]
    kvec = 2 * math.pi * mode / grid[axis]
    comp = (axis + 1) % 3  # transverse E component, its position along `axis` is on the node
    E = np.zeros((3, n), np.float32)
    E[comp] = np.cos(kvec * coord[axis])
    ora.set_field(abi.FIELD_E, E)
    amps = []
    basis = np.cos(kvec * coord[axis])
    for _ in range(steps):
        ora.update_before_current()
        ora.add_current()
        ora.update_after_current()
        e = ora.get_field(abi.FIELD_E)[comp]
        amps.append(float(e @ basis / (basis @ basis)))
    return st, k, np.array(amps), kvec


def test_yee_standing_wave_dispersion():
    for axis in range(3):
        st, k, amps, kvec = standing_wave(abi.SOLVER_YEE, axis)
        h = st.cell_si[axis] / (st.dt_si * setups.SPEED_OF_LIGHT_SI)  # FP64 twin cell size
        cdt = float(k.c) * float(k.dt)
        s2 = (cdt / h) ** 2 * math.sin(kvec / 2) ** 2
        omega = math.acos(1 - 2 * s2)
        ref = np.cos(omega * np.arange(1, len(amps) + 1))
        np.testing.assert_allclose(amps, ref, atol=1e-7)  # initial field enters through FP32


def test_lehe_standing_wave_dispersion():
    # wave along the Cherenkov-free axis y: curl E uses the Lehe stencil
    st, k, amps, kvec = standing_wave(abi.SOLVER_LEHE, 1)
    cell = [c / (st.dt_si * setups.SPEED_OF_LIGHT_SI) for c in st.cell_si]  # FP64 twin cell sizes
    cdt = float(k.c) * float(k.dt)
    # FP64 coefficients as the FP64 twin computes them
    h = cell[1]
    sr = h / cdt
    co = sr * math.sin(math.pi / 2 * cdt / h)
    delta = 0.25 * (1 - co * co)
    b1 = 0.125 * (h / cell[2]) ** 2
    b2 = 0.125 * (h / cell[0]) ** 2
    alpha = 1 - 2 * b1 - 2 * b2 - 3 * delta
    fac = (alpha + 2 * b1 + 2 * b2) * math.sin(kvec / 2) + delta * math.sin(3 * kvec / 2)
    s2 = (cdt / h) ** 2 * math.sin(kvec / 2) * fac
    omega = math.acos(1 - 2 * s2)
    ref = np.cos(omega * np.arange(1, len(amps) + 1))
    np.testing.assert_allclose(amps, ref, atol=1e-7)  # initial field enters through FP32
    # and it differs from the plain Yee relation
    s2y = (cdt / h) ** 2 * math.sin(kvec / 2) ** 2
    assert abs(math.acos(1 - 2 * s2y) - omega) > 1e-6


def test_lehe_coefficients_closed_form():
    st = setups.laser_plasma(grid=(16, 16, 16))
    k = O.constants(st)
    cell = [float(np.float32(c)) for c in k.cell]
    cdt = float(np.float32(k.c) * np.float32(k.dt))
    h = cell[1]
    sr = float(np.float32(np.float32(k.cell[1]) / (np.float32(k.c) * np.float32(k.dt))))
    co = sr * math.sin(math.pi / 2 * cdt / h)
    delta = np.float32(0.25 * (1 - co * co))
    b1 = 0.125 * float(np.float32(np.float32(k.cell[1]) / np.float32(k.cell[2]))) ** 2
    b2 = 0.125 * float(np.float32(np.float32(k.cell[1]) / np.float32(k.cell[0]))) ** 2
    alpha = np.float32(1 - 2 * b1 - 2 * b2 - 3 * float(delta))
    assert np.float32(k.lehe_delta) == delta
    assert np.float32(k.lehe_alpha) == alpha
    assert -0.25 < delta <= 0  # sin(pi r/2)/r >= 1 for r = c dt/dy <= 1


# ---------------------------------------------------------------------------
# unit system
# ---------------------------------------------------------------------------
def test_constants_closed_form():
    st = setups.thermal()
    k = O.constants(st)
    c = 2.99792458e8
    unit_len = st.dt_si * c
    assert k.cell[0] == np.float32(st.cell_si[0] / unit_len)
    assert k.dt == np.float32(1.0) and k.c == np.float32(1.0)
    wtyp = st.density_si * (st.cell_si[0] * st.cell_si[1] * st.cell_si[2]) / st.typical_ppc
    assert k.typical_weighting == wtyp
    # electron mass and charge per unit weighting
    assert k.species_mass[0] == np.float32(1.0 / wtyp)
    assert k.species_charge[0] == np.float32(-1.0 / wtyp)
    # eps0 in PIC units: mue0 = MUE0_SI / L / M * Q^2
    me, qe, mue0 = 9.1093837139e-31, 1.602176634e-19, 1.25663706127e-6
    m_u = me * wtyp
    q_u = qe * wtyp
    mue0_pic = np.float32(mue0 / unit_len / m_u * q_u * q_u)
    assert k.mue0 == mue0_pic
    assert k.eps0 == np.float32(1.0 / float(mue0_pic))
    assert k.add_current_coeff == np.float32(-(np.float32(1.0) / np.float32(k.eps0)) * np.float32(1.0))
    # Courant number of the thermal setup is 0.99 of the 3-D Yee limit
    assert abs(float(k.cell[0]) - math.sqrt(3) / 0.99) < 1e-5


# ---------------------------------------------------------------------------
# whole cycle
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("shape", [abi.SHAPE_CIC, abi.SHAPE_TSC, abi.SHAPE_PQS])
@pytest.mark.parametrize("order,precision", [(0, 32), (1, 32), (1, 64)])
def test_charge_conservation_over_steps(shape, order, precision):
    st = setups.thermal(grid=(16, 16, 12), ppc=3, shape=shape)
    st.species[0].u_th = (0.05, 0.05, 0.05)
    cells, pos, mom, w = setups.host_particles(st, 0)
    ora = O.Oracle(st, precision, order, jexp=40)
    ora.set_frames(0, setups.bin_frames(st, 0, cells, pos, mom, w))
    k = O.constants(st)
    eps0 = float(np.float32(k.eps0)) if precision == 32 else None
    if precision == 64:
        # FP64 twin: eps0 from the FP64 unit system
        me, qe, mue0, c = 9.1093837139e-31, 1.602176634e-19, 1.25663706127e-6, 2.99792458e8
        wt = k.typical_weighting
        mu = mue0 / (st.dt_si * c) / (me * wt) * (qe * wt) ** 2
        eps0 = 1.0 / mu
    r0 = eps0 * ora.div_e() - ora.charge_density()
    ora.step(8)
    r1 = eps0 * ora.div_e() - ora.charge_density()
    scale = np.abs(ora.charge_density()).max()
    tol = 2e-6 if precision == 32 else 1e-12
    assert np.abs(r1 - r0).max() <= tol * scale
    assert ora.get_frames(0)["counts"].sum() == cells.shape[0]


def test_fp32_against_fp64_twin_100_steps():
    """FP32 state against the FP64 twin after 100 steps (1e-4, max-norm relative)."""
    st = setups.thermal(grid=(16, 16, 12), ppc=4, shape=abi.SHAPE_CIC)
    cells, pos, mom, w = setups.host_particles(st, 0)
    f = setups.bin_frames(st, 0, cells, pos, mom, w)
    o32 = O.Oracle(st, 32, 1)
    o64 = O.Oracle(st, 64, 1)
    for o in (o32, o64):
        o.set_frames(0, f)
    o32.step(100)
    o64.step(100)
    k = O.constants(st)
    res = {}
    gp = {}
    for name, o in (("f32", o32), ("f64", o64)):
        fr = o.get_frames(0)
        ids = o.get_ids(0)
        cap = o.cap(0)
        valid = (np.arange(cap)[None, :] < fr["counts"][:, None].astype(np.int64)).ravel()
        sc = np.repeat(np.arange(o.nsc), cap)[valid]
        nscx, nscy = st.grid[0] // 8, st.grid[1] // 8
        lc = fr["cell"][valid].astype(np.int64)
        g = np.stack([(sc % nscx) * 8 + lc % 8, ((sc // nscx) % nscy) * 8 + (lc // 8) % 8,
                      (sc // (nscx * nscy)) * 4 + lc // 64]).astype(np.float64) + fr["pos"][:, valid]
        order = np.argsort(ids[valid])
        gp[name] = (g[:, order], fr["mom"][:, valid][:, order], lc[order])
    L = np.array(st.grid, np.float64)[:, None]
    dpos = (gp["f32"][0] - gp["f64"][0] + L / 2) % L - L / 2
    res["pos_cells_max"] = np.abs(dpos).max()
    res["mom_rel"] = np.abs(gp["f32"][1] - gp["f64"][1]).max() / np.abs(gp["f64"][1]).max()
    for f_id, nm in ((abi.FIELD_E, "E"), (abi.FIELD_B, "B"), (abi.FIELD_J, "J")):
        a, b = o32.get_field(f_id), o64.get_field(f_id)
        res[nm] = np.abs(a - b).max() / np.abs(b).max()
    flips = np.count_nonzero(gp["f32"][2] != gp["f64"][2])
    print("FP32 vs FP64 twin after 100 steps:", res, "cell-index flips", flips)
    assert res["pos_cells_max"] < 1e-4
    assert res["mom_rel"] < 1e-4
    for nm in ("E", "B", "J"):
        assert res[nm] < 1e-4
