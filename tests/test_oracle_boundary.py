"""Open borders and the exponential field absorber in the oracle, pinned to
closed forms and to independent numpy statements.

Open axis (the reference with the axis not periodic and the default
absorbing particle boundary, offset 0): fields read zero beyond the border,
current deposited there is dropped, particles crossing it are removed.
Exponential absorber (absorber/exponential/Exponential.hpp:52-95,
Exponential.kernel:63-112): at distance d from an open border, fields are
multiplied by exp(-strength * (T - 1 - d)) for d < T - 1, thickness T and
strength read at [axis][exchange % 2]: exchange 1 is RIGHT, the +x side, and
2 is LEFT, -x (pmacc/type/Exchange.hpp:48-49), so the negative side uses
entry 0 and the positive side entry 1, as fieldAbsorber.param orders them."""
import math

import numpy as np
import pytest

from oracle import oracle as O
from picongpu_amd import abi, setups


def field_only(grid, boundary=(1, 0, 0), absorber=0, cells=((0, 0), (0, 0), (0, 0)),
               strength=((0.0, 0.0), (0.0, 0.0), (0.0, 0.0))):
    st = setups.thermal(grid=grid, ppc=1)
    st.species = []
    st.boundary = boundary
    st.absorber = absorber
    st.absorber_cells = cells
    st.absorber_strength = strength
    return st


def test_absorber_factors_closed_form():
    nx = 64
    st = field_only((nx, 8, 4), absorber=abi.ABSORBER_EXPONENTIAL, cells=((6, 10), (0, 0), (0, 0)),
                    strength=((0.05, 0.02), (0.0, 0.0), (0.0, 0.0)))
    ora = O.Oracle(st, 32, 1)
    E = np.ones((3, st.ncells), np.float32)
    ora.set_field(abi.FIELD_E, E)
    ora.update_after_current()  # absorber on E, B half step, absorber on B
    got = ora.get_field(abi.FIELD_E).reshape(3, 4, 8, nx)
    want = np.ones(nx, np.float32)
    # LEFT (-x) is exchange 2: entry [x][0] = 6 cells, strength 0.05
    for x in range(0, 5):
        want[x] = np.exp(np.float32(-0.05) * np.float32(6 - 1 - x))
    # RIGHT (+x) is exchange 1: entry [x][1] = 10 cells, strength 0.02
    for d in range(0, 9):
        want[nx - 1 - d] = np.exp(np.float32(-0.02) * np.float32(10 - 1 - d))
    for c in range(3):
        np.testing.assert_allclose(got[c], np.broadcast_to(want, (4, 8, nx)), rtol=2e-7, atol=0)
    assert np.all(got[:, :, :, 5:nx - 9] == 1.0)
    ora.close()


def test_open_border_curl_reads_zero_beyond():
    """update_afterCurrent's B half step against a numpy Yee curl with zero
    padding along the open axis and periodic wrap on the others."""
    nx, ny, nz = 32, 16, 8
    st = field_only((nx, ny, nz))
    ora = O.Oracle(st, 32, 1)
    k = O.constants(st)
    rng = np.random.default_rng(5)
    E = rng.standard_normal((3, st.ncells)).astype(np.float32)
    B = rng.standard_normal((3, st.ncells)).astype(np.float32)
    ora.set_field(abi.FIELD_E, E)
    ora.set_field(abi.FIELD_B, B)
    ora.update_after_current()
    e = E.reshape(3, nz, ny, nx)
    b = B.reshape(3, nz, ny, nx)

    def fwd(f, axis):  # (f[i+1] - f[i]) / dx, zero beyond the open x border
        cell = np.float32(k.cell[axis])
        if axis == 0:
            nxt = np.concatenate([f[:, :, 1:], np.zeros_like(f[:, :, :1])], axis=2)
        elif axis == 1:
            nxt = np.roll(f, -1, axis=1)
        else:
            nxt = np.roll(f, -1, axis=0)
        return (nxt - f) / cell

    cx = fwd(e[2], 1) - fwd(e[1], 2)
    cy = fwd(e[0], 2) - fwd(e[2], 0)
    cz = fwd(e[1], 0) - fwd(e[0], 1)
    half = np.float32(0.5) * np.float32(k.dt)
    want = np.stack([b[0] - cx * half, b[1] - cy * half, b[2] - cz * half]).reshape(3, -1)
    np.testing.assert_array_equal(ora.get_field(abi.FIELD_B).astype(np.float32), want)
    ora.close()


def test_absorber_removes_a_pulse_an_open_border_reflects_it():
    def run(absorber):
        st = field_only((256, 8, 4), absorber=absorber, cells=((32, 32), (0, 0), (0, 0)),
                        strength=((0.1, 0.1), (0.0, 0.0), (0.0, 0.0)))
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
        for _ in range(400):  # the pulse travels about 230 cells: into the border and back
            ora.update_before_current()
            ora.add_current()
            ora.update_after_current()
        e1 = ora.field_energy()[0]
        ora.close()
        return e1 / e0
    assert run(abi.ABSORBER_NONE) > 0.99
    assert run(abi.ABSORBER_EXPONENTIAL) < 0.05


@pytest.mark.parametrize("boundary", [0, 1])
def test_particle_leaving_an_open_border_is_removed(boundary):
    st = setups.thermal(grid=(32, 8, 4), ppc=1, pusher=abi.PUSHER_BORIS)
    st.boundary = (boundary, 0, 0)
    ora = O.Oracle(st, 32, 1)
    k = O.constants(st)
    w = np.float32(k.typical_weighting)
    mass = float(np.float32(k.species_mass[0]) * w)
    beta = 0.5
    p = np.float32(beta / math.sqrt(1 - beta * beta) * mass * k.c)
    cells = np.array([[29, 4, 2]], np.int32)
    f = setups.bin_frames(st, 0, cells, np.array([[0.5, 0.5, 0.5]], np.float32), np.array([[p, 0.0, 0.0]], np.float32),
                          np.array([w], np.float32))
    ora.set_frames(0, f)
    for _ in range(20):  # about 5.7 cells: past the +x border at x = 32
        ora.step(1)
    n = int(ora.get_frames(0)["counts"].sum())
    assert n == (1 if boundary == 0 else 0)
    ora.close()


def test_binomial_filter_weights_and_conservation():
    """Binomial.hpp:40-104: a unit impulse spreads with weights 8 (centre),
    4 (faces), 2 (edges), 1 (corners) over 64; the filter keeps the total."""
    grid = (16, 16, 12)
    st = setups.thermal(grid=grid, ppc=1)
    st.species = []
    st.current_interpolation = abi.CURRENT_BINOMIAL
    ora = O.Oracle(st, 32, 1)
    k = O.constants(st)
    J = np.zeros((3, st.ncells), np.float32)
    nx, ny, nz = grid
    c0 = (5, 7, 3)
    J[1, c0[0] + nx * (c0[1] + ny * c0[2])] = 64.0
    ora.set_field(abi.FIELD_J, J)
    ora.add_current()
    E = ora.get_field(abi.FIELD_E).reshape(3, nz, ny, nx)
    coeff = np.float32(k.add_current_coeff)
    for dz in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                w = 8 >> (abs(dx) + abs(dy) + abs(dz))
                assert E[1, c0[2] + dz, c0[1] + dy, c0[0] + dx] == np.float32(coeff * np.float32(w))
    assert np.count_nonzero(E[1]) == 27 and not E[0].any() and not E[2].any()
    ora.close()
    rng = np.random.default_rng(2)
    J = rng.standard_normal((3, st.ncells)).astype(np.float32)
    ora = O.Oracle(st, 64, 1)
    ora.set_field(abi.FIELD_J, J)
    ora.add_current()
    E = ora.get_field(abi.FIELD_E)
    np.testing.assert_allclose(E.sum(axis=1), float(k.add_current_coeff) * J.astype(np.float64).sum(axis=1),
                               rtol=1e-9, atol=1e-9 * abs(float(k.add_current_coeff)) * st.ncells)
    ora.close()


# ---- reflecting particle boundary (boundary/Reflecting.hpp:38-66) ----------
def reflecting_box(axes, grid=(16, 16, 12)):
    st = setups.thermal(grid=grid, ppc=1, pusher=abi.PUSHER_BORIS)
    st.boundary = (1, 1, 1)
    st.species[0].particle_boundary = tuple(abi.PARTICLE_REFLECTING if d in axes else abi.PARTICLE_ABSORBING
                                            for d in range(3))
    return st


def one_particle(st, cell, pos, beta_dir):
    """One electron at `cell`/`pos` moving with velocity beta_dir (units of c)
    in zero fields; returns the oracle, its momentum and the cells it moves
    per step along each axis (float64)."""
    ora = O.Oracle(st, 32, 0)
    k = O.constants(st)
    w = np.float32(k.typical_weighting)
    mass = float(np.float32(k.species_mass[0]) * w)
    beta = np.asarray(beta_dir, np.float64)
    g = 1.0 / math.sqrt(1.0 - float(beta @ beta))
    p = (g * beta * mass * float(k.c)).astype(np.float32)
    f = setups.bin_frames(st, 0, np.array([cell], np.int32), np.array([pos], np.float32), p[None, :],
                          np.array([w], np.float32))
    ora.set_frames(0, f)
    p64 = p.astype(np.float64)
    v = p64 / np.sqrt(mass * mass + (p64 @ p64) / float(k.c) ** 2)  # Velocity.hpp
    per_step = v * float(k.dt) / np.array([k.cell[0], k.cell[1], k.cell[2]], np.float64)
    return ora, p, per_step


def particle_state(ora, st):
    f = ora.get_frames(0)
    sc = int(np.nonzero(f["counts"])[0][0])
    cap = f["w"].shape[0] // f["counts"].shape[0]
    slot = sc * cap
    nscx, nscy = st.grid[0] // 8, st.grid[1] // 8
    lc = int(f["cell"][slot])
    cell = np.array([(sc % nscx) * 8 + lc % 8, ((sc // nscx) % nscy) * 8 + (lc // 8) % 8,
                     (sc // (nscx * nscy)) * 4 + lc // 64])
    return cell, f["pos"][:, slot].astype(np.float64), f["mom"][:, slot].astype(np.float64), int(f["counts"].sum())


@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("side", [1, -1])
def test_reflecting_border_mirrors_the_particle(axis, side):
    """Crossing a reflecting border the particle comes back mirrored: global
    coordinate 2 X_b - x of where the unreflected move ends (X_b the border),
    to FP32 round-off plus the epsilon the upper border subtracts; the momentum
    component is negated, the other components and coordinates unchanged."""
    st = reflecting_box(axes=(axis,))
    n = st.grid[axis]
    cell = [5, 5, 5]
    pos = [0.5, 0.5, 0.5]
    cell[axis] = n - 1 if side > 0 else 0
    pos[axis] = 0.9 if side > 0 else 0.1
    beta = [0.0, 0.0, 0.0]
    beta[axis] = 0.5 * side
    ora, p, per_step = one_particle(st, cell, pos, beta)
    x0 = cell[axis] + pos[axis]
    x_unreflected = x0 + per_step[axis]
    assert (x_unreflected > n) if side > 0 else (x_unreflected < 0)  # the move crosses the border
    ora.step(1)
    c1, pos1, mom1, count = particle_state(ora, st)
    assert count == 1
    border = n if side > 0 else 0
    x1 = c1[axis] + pos1[axis]
    assert x1 == pytest.approx(2 * border - x_unreflected, abs=4e-7)
    assert c1[axis] == (n - 1 if side > 0 else 0)
    assert mom1[axis] == -float(p[axis])
    for d in range(3):
        if d != axis:
            assert c1[d] == 5 and pos1[d] == 0.5 and mom1[d] == float(p[d])
    ora.close()


def test_reflecting_corner_mirrors_both_axes():
    st = reflecting_box(axes=(0, 1))
    n = st.grid
    ora, p, per_step = one_particle(st, [n[0] - 1, 0, 5], [0.95, 0.05, 0.5], [0.4, -0.4, 0.0])
    x_unref = np.array([n[0] - 1 + 0.95, 0.05, 5.5]) + per_step
    ora.step(1)
    c1, pos1, mom1, count = particle_state(ora, st)
    assert count == 1
    assert c1[0] == n[0] - 1 and c1[1] == 0 and c1[2] == 5
    assert c1[0] + pos1[0] == pytest.approx(2 * n[0] - x_unref[0], abs=4e-7)
    assert c1[1] + pos1[1] == pytest.approx(-x_unref[1], abs=4e-7)
    assert mom1[0] == -float(p[0]) and mom1[1] == -float(p[1]) and mom1[2] == float(p[2])
    ora.close()


def test_reflecting_box_keeps_every_particle():
    """A thermal plasma in a box whose borders all reflect keeps its particles,
    every one inside the box with in-cell positions in [0, 1]; the same box with
    absorbing borders loses particles."""
    counts = {}
    for kind in (abi.PARTICLE_REFLECTING, abi.PARTICLE_ABSORBING):
        st = setups.thermal(grid=(16, 16, 12), ppc=2, pusher=abi.PUSHER_BORIS)
        st.boundary = (1, 1, 1)
        st.species[0].u_th = (0.3, 0.3, 0.3)
        st.species[0].particle_boundary = (kind, kind, kind)
        ora = O.Oracle(st, 32, 0)
        f0 = setups.bin_frames(st, 0, *setups.host_particles(st, 0))
        ora.set_frames(0, f0)
        ora.step(12)
        f = ora.get_frames(0)
        counts[kind] = int(f["counts"].sum())
        cap = f["w"].shape[0] // f["counts"].shape[0]
        valid = (np.arange(cap)[None, :] < f["counts"][:, None]).ravel()
        assert np.all((f["pos"][:, valid] >= 0.0) & (f["pos"][:, valid] <= 1.0))
        n0 = int(f0["counts"].sum())
        ora.close()
    assert counts[abi.PARTICLE_REFLECTING] == n0
    assert counts[abi.PARTICLE_ABSORBING] < n0
