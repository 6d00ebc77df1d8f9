"""Oracle pinning: incident field through a Huygens surface
(fields/incidentField: Solver.hpp, Solver.kernel, Functors.hpp,
profiles/GaussianPulse.hpp, profiles/PlaneWave.hpp, calculatePhaseVelocity.hpp).

The reference's own tests hold no vectors for this path, so each check is a
property the mathematics fixes:
  - the phase velocity against the closed-form roots of the Yee and Lehe
    discrete dispersion relations along an axis and a diagonal;
  - the GaussianPulse functor against Gaussian-beam identities: amplitude A at
    the focus at the pulse centre, exp(-r^2/w0^2) across the focal plane,
    A/2 on axis at +-z_R with the Gouy phase +-pi/4 (a sign flip swaps the
    two), the circular polarisation vectors, the Laguerre-Gauss modes;
  - the PlaneWave plateau A sin(w (t - mue));
  - TF/SF injection by the oracle's steps: the field inside the total-field
    region follows the analytic wave moving at the numerical phase velocity,
    and the scattered-field region stays near zero (a wrong sign, shift or
    time of the correction leaks the whole wave there).
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from picongpu_amd import abi, setups

C_SI = 2.99792458e8


def vacuum(grid, cell, dt, incident, solver=abi.SOLVER_YEE, lehe_dir=1, boundary=(0, 1, 0), position=None, **kw):
    return setups.Setup("vacuum", tuple(grid), tuple(cell), dt, 1e25, 1, [], solver=solver, lehe_dir=lehe_dir,
                        boundary=boundary, incident=incident,
                        incident_position=position or ((16, -16), (16, -16), (16, -16)), **kw)


def units(st):
    """FP64 constants as the FP64 oracle holds them: cell sizes and time step
    in PIC units without the FP32 rounding, and the E-field unit."""
    k = O.constants(st)
    unit_e = 1.0 / (k.unit_time ** 2 / k.unit_mass / k.unit_length * k.unit_charge)
    cell = np.array(st.cell_si, dtype=np.float64) / k.unit_length
    dt = st.dt_si / k.unit_time
    return k, unit_e, cell, dt


# ---- phase velocity -------------------------------------------------------
def test_phase_velocity_yee_axis_and_diagonal():
    cell = (0.05e-6, 0.05e-6, 0.05e-6)
    dt = setups.yee_dt(cell, 0.9)
    wl = 0.8e-6
    for direction in [(0, 1, 0), (1 / math.sqrt(2), 1 / math.sqrt(2), 0)]:
        prof = dict(kind=abi.INCIDENT_PLANE_WAVE, wavelength_si=wl, amplitude_si=1.0, pulse_duration_si=1e-15,
                    direction=direction, polarisation_direction=(0, 0, 1))
        st = vacuum((16, 64, 8), cell, dt, {(1, 0): prof})
        k, _, cellp, dtp = units(st)
        vph, _, _ = O.incident_info(st, 1, 0)
        d = cellp[0]
        omega = 2 * math.pi / (wl / k.unit_length)
        n_axes = sum(1 for c in direction if abs(c) > 0)
        # sum over the n equal axes of sin^2(k d/(2 sqrt n))/d^2 = sin^2(omega dt/2)/dt^2
        s = d / math.sqrt(n_axes) / dtp * math.sin(omega * dtp / 2)
        absk = 2 * math.sqrt(n_axes) / d * math.asin(s)
        assert vph == pytest.approx(omega / absk, rel=1e-12)
        assert vph < 1.0


def test_phase_velocity_lehe_cherenkov_free_axis():
    cell = (1e-7, 5e-8, 1e-7)
    dt = 0.995 * cell[1] / C_SI
    wl = 0.8e-6
    prof = dict(kind=abi.INCIDENT_PLANE_WAVE, wavelength_si=wl, amplitude_si=1.0, pulse_duration_si=1e-15,
                direction=(0, 1, 0), polarisation_direction=(1, 0, 0))
    st = vacuum((16, 64, 8), cell, dt, {(1, 0): prof}, solver=abi.SOLVER_LEHE, lehe_dir=1)
    k, _, cellp, dtp = units(st)
    vph, _, _ = O.incident_info(st, 1, 0)
    dy = cellp[1]
    omega = 2 * math.pi / (wl / k.unit_length)
    ratio = dy / dtp
    co = ratio * math.sin(math.pi / 2 * dtp / dy)
    delta = 0.25 * (1 - co * co)
    L2 = (math.sin(omega * dtp / 2) / dtp) ** 2
    # along the axis: u/dy^2 - 4 delta u^2/dy^2 = L2 with u = sin^2(k dy / 2)
    u = (1 - math.sqrt(1 - 16 * delta * L2 * dy * dy)) / (8 * delta)
    absk = 2 / dy * math.asin(math.sqrt(u))
    assert vph == pytest.approx(omega / absk, rel=1e-12)


# ---- GaussianPulse functor ------------------------------------------------
def gaussian_setup(pol=abi.POLARISATION_LINEAR, modes=(1.0,), phases=None, w0=1.5e-6):
    cell = (0.1e-6, 0.05e-6, 0.1e-6)
    dt = 0.995 * cell[1] / C_SI
    grid = (64, 256, 64)
    prof = dict(kind=abi.INCIDENT_GAUSSIAN_PULSE, wavelength_si=0.8e-6, amplitude_si=4e12, pulse_duration_si=5e-15,
                direction=(0, 1, 0), polarisation_direction=(1, 0, 0), focus_position_si=(0, 6e-6, 0),
                focus_origin=(abi.ORIGIN_CENTER, abi.ORIGIN_ZERO, abi.ORIGIN_CENTER), w0_si=w0, pulse_init=15.0,
                polarisation=pol, laguerre_modes=modes, laguerre_phases=phases)
    return vacuum(grid, cell, dt, {(1, 0): prof}, solver=abi.SOLVER_LEHE, lehe_dir=1, boundary=(1, 1, 0),
                  absorber=abi.ABSORBER_EXPONENTIAL, absorber_cells=((12, 12), (12, 12), (12, 12)),
                  absorber_strength=((1e-3, 1e-3),) * 3)


def pulse_centre_step(st, y_cells):
    """Source step at which the pulse centre (envelope maximum, zero carrier
    phase) crosses total position y (cells) on the axis."""
    k, _, cellp, dtp = units(st)
    vph, origin, _ = O.incident_info(st, 1, 0)
    prof = st.incident[(1, 0)]
    pd = prof["pulse_duration_si"] / k.unit_time
    time_shift = -0.5 * prof["pulse_init"] * pd
    distance = y_cells * cellp[1] - origin[1]
    return (distance / vph - time_shift) / dtp


def test_gaussian_focus_amplitude_waist_and_polarisation():
    st = gaussian_setup()
    k, unit_e, cellp, dtp = units(st)
    A = 4e12 / unit_e
    _, _, focus = O.incident_info(st, 1, 0)
    cell = cellp
    fc = focus / cell  # focus in cells
    assert fc[0] == pytest.approx(32.0) and fc[2] == pytest.approx(32.0)
    step = pulse_centre_step(st, fc[1])
    w0 = 1.5e-6 / k.unit_length
    r = np.linspace(-2 * w0, 2 * w0, 9)
    pos = np.stack([fc[0] + r / cell[0], np.full_like(r, fc[1]), np.full_like(r, fc[2])], axis=1)
    E = O.incident_field(st, 1, 0, 0, step, pos)
    np.testing.assert_allclose(E[:, 0], A * np.exp(-(r / w0) ** 2), rtol=1e-9, atol=1e-12 * A)
    assert np.all(E[:, 1:] == 0.0)
    posz = np.stack([np.full_like(r, fc[0]), np.full_like(r, fc[1]), fc[2] + r / cell[2]], axis=1)
    Ez = O.incident_field(st, 1, 0, 0, step, posz)
    np.testing.assert_allclose(Ez[:, 0], A * np.exp(-(r / w0) ** 2), rtol=1e-9, atol=1e-12 * A)
    # ApproximateIncidentB: B = y x E / c = -E_x z
    B = O.incident_field(st, 1, 0, 1, step, pos)
    np.testing.assert_allclose(B[:, 2], -E[:, 0], rtol=1e-12)
    assert np.all(B[:, :2] == 0.0)
    # before the source time origin the profile is zero
    assert np.all(O.incident_field(st, 1, 0, 0, -1.0, pos) == 0.0)


def test_gaussian_gouy_phase_at_rayleigh_length():
    st = gaussian_setup()
    k, unit_e, cellp, dtp = units(st)
    A = 4e12 / unit_e
    _, _, focus = O.incident_info(st, 1, 0)
    dy = cellp[1]
    wl = 0.8e-6 / k.unit_length
    w0 = 1.5e-6 / k.unit_length
    zr = math.pi * w0 * w0 / wl
    pd = 5e-15 / k.unit_time
    period = wl
    for sign in (+1, -1):
        y = focus[1] / dy + sign * zr / dy
        pos = np.array([[focus[0] / cellp[0], y, focus[2] / cellp[2]]])
        step = pulse_centre_step(st, y)
        # pulse centre: carrier phase w(time - focusPos/c) + Gouy = +-pi/4, width w0 sqrt 2
        E0 = O.incident_field(st, 1, 0, 0, step, pos)[0, 0]
        assert E0 == pytest.approx(A / 2, rel=1e-6)
        # an eighth of a period later the carrier has advanced by pi/4: cos(pi/4 + pi/4) = 0 at +z_R,
        # cos(-pi/4 + pi/4) = 1 at -z_R; the envelope there is exp(-(T/8 / (2 pd))^2)
        dt8 = period / 8 / dtp
        env = math.exp(-((period / 8) / (2 * pd)) ** 2)
        E1 = O.incident_field(st, 1, 0, 0, step + dt8, pos)[0, 0]
        expect = 0.0 if sign > 0 else A / math.sqrt(2) * env
        assert E1 == pytest.approx(expect, abs=2e-5 * A)


def test_gaussian_circular_polarisation_vectors():
    st = gaussian_setup(pol=abi.POLARISATION_CIRCULAR)
    k, unit_e, cellp, dtp = units(st)
    A = 4e12 / unit_e
    _, _, focus = O.incident_info(st, 1, 0)
    cell = cellp
    fc = focus / cell
    step = pulse_centre_step(st, fc[1])
    E = O.incident_field(st, 1, 0, 0, step, fc[None, :])[0]
    # cp1 (x / sqrt 2) carries cos(pi/2) = 0, cp2 = y x cp1 = -z / sqrt 2 carries cos(0) = 1
    assert E[0] == pytest.approx(0.0, abs=1e-7 * A)
    assert E[1] == 0.0
    assert E[2] == pytest.approx(-A / math.sqrt(2), rel=1e-9)
    # a quarter period later the field has turned onto the x axis
    quarter = 0.8e-6 / k.unit_length / 4 / dtp
    E = O.incident_field(st, 1, 0, 0, step + quarter, fc[None, :])[0]
    env = math.exp(-((quarter * dtp) / (2 * 5e-15 / k.unit_time)) ** 2)
    assert abs(E[0]) == pytest.approx(A / math.sqrt(2) * env, rel=1e-5)
    assert E[2] == pytest.approx(0.0, abs=1e-5 * A)


def test_gaussian_laguerre_modes():
    modes = (1.0, 0.5, -0.25)
    st = gaussian_setup(modes=modes)
    k, unit_e, cellp, dtp = units(st)
    A = 4e12 / unit_e
    _, _, focus = O.incident_info(st, 1, 0)
    cell = cellp
    fc = focus / cell
    step = pulse_centre_step(st, fc[1])
    w0 = 1.5e-6 / k.unit_length
    r = np.linspace(0, 2 * w0, 7)
    pos = np.stack([fc[0] + r / cell[0], np.full_like(r, fc[1]), np.full_like(r, fc[2])], axis=1)
    E = O.incident_field(st, 1, 0, 0, step, pos)[:, 0]
    x = 2 * (r / w0) ** 2
    lag = [np.ones_like(x), 1 - x, (x * x - 4 * x + 2) / 2]  # L_0, L_1, L_2
    # at the focus the mode phases (2m+1) xi vanish
    expect = A * np.exp(-(r / w0) ** 2) * sum(m * l for m, l in zip(modes, lag)) / sum(modes)
    np.testing.assert_allclose(E, expect, rtol=1e-8, atol=1e-10 * A)


def test_plane_wave_plateau():
    cell = (0.05e-6, 0.05e-6, 0.05e-6)
    dt = setups.yee_dt(cell, 0.9)
    prof = dict(kind=abi.INCIDENT_PLANE_WAVE, wavelength_si=0.8e-6, amplitude_si=2e12, pulse_duration_si=2e-15,
                direction=(0, 1, 0), polarisation_direction=(0, 0, 1), ramp_init=6.0,
                laser_nofocus_constant_si=20e-15)
    st = vacuum((16, 64, 8), cell, dt, {(1, 0): prof})
    k, unit_e, cellp, dtp = units(st)
    A = 2e12 / unit_e
    vph, origin, _ = O.incident_info(st, 1, 0)
    pd = 2e-15 / k.unit_time
    mue = 0.5 * 6.0 * pd
    w = 2 * math.pi / (0.8e-6 / k.unit_length)
    y = 20.0
    delay = (y * cellp[1] - origin[1]) / vph
    for frac in (0.1, 0.37, 0.8):
        t = mue + frac * 20e-15 / k.unit_time
        E = O.incident_field(st, 1, 0, 0, (t + delay) / dtp, np.array([[3.0, y, 2.0]]))[0]
        assert E[2] == pytest.approx(A * math.sin(w * (t - mue)), abs=1e-9 * A)
        assert E[0] == 0.0 and E[1] == 0.0


# ---- injection through the Huygens surface ---------------------------------
def plane_wave_run(solver, side, d=0.02e-6, pd=4e-15, prec=64):
    cell = (d, d, d)
    lehe = solver == abi.SOLVER_LEHE
    dt = 0.95 * d / C_SI if lehe else setups.yee_dt(cell, 0.95)
    grid = (8, 512, 4)
    direction = (0, 1, 0) if side == 0 else (0, -1, 0)
    prof = dict(kind=abi.INCIDENT_PLANE_WAVE, wavelength_si=0.8e-6, amplitude_si=1e12, pulse_duration_si=pd,
                direction=direction, polarisation_direction=(1, 0, 0), ramp_init=8.0)
    st = vacuum(grid, cell, dt, {(1, side): prof}, solver=solver, lehe_dir=1)
    k, unit_e, cellp, dtp = units(st)
    A = 1e12 / unit_e
    vph, origin, _ = O.incident_info(st, 1, side)
    dy = cellp[1]
    mue = 0.5 * 8.0 * pd / k.unit_time
    n = int((mue + 150 * dy / vph) / dtp)
    o = O.Oracle(st, precision=prec)
    o.step(n)
    assert o.get_step() == n
    E = o.get_field(abi.FIELD_E).reshape(3, grid[2], grid[1], grid[0])
    B = o.get_field(abi.FIELD_B).reshape(3, grid[2], grid[1], grid[0])
    o.close()
    # the closed form of the PlaneWave functor, moving at the numerical phase velocity
    wl = 0.8e-6 / k.unit_length
    w = 2 * math.pi / wl
    tau = pd / k.unit_time * math.sqrt(2)
    y = np.arange(grid[1]) * dy
    dist = (y - origin[1]) * direction[1]
    t = n * dtp - dist / vph
    env = A * np.exp(-0.5 * ((t - mue) / tau) ** 2)
    ana = np.where(t >= 0, (np.sin(w * (t - mue)) + np.cos(w * (t - mue)) * (t - mue) / (w * tau * tau)) * env, 0)
    return E, B, ana, A, grid


@pytest.mark.parametrize("solver", [abi.SOLVER_YEE, abi.SOLVER_LEHE])
@pytest.mark.parametrize("side", [0, 1])
def test_plane_wave_injection(solver, side):
    E, B, ana, A, grid = plane_wave_run(solver, side)
    ex = E[0, 0, :, 0]
    ny = grid[1]
    if side == 0:
        total, scattered = slice(17, ny), slice(0, 16)
    else:
        total, scattered = slice(0, ny - 17), slice(ny - 16, ny)
    assert np.abs(ex[total] - ana[total]).max() < 1e-2 * A
    assert np.abs(ex[scattered]).max() < 5e-4 * A
    assert np.abs(ex).max() > 0.95 * A
    # transversally uniform, only E_x and B_z
    assert np.abs(E[0] - E[0][:1, :, :1]).max() == 0.0
    assert np.abs(E[1]).max() == 0.0 and np.abs(E[2]).max() == 0.0
    assert np.abs(B[0]).max() == 0.0 and np.abs(B[1]).max() == 0.0


def test_plane_wave_injection_fp32_tracks_fp64():
    E32, _, _, A, _ = plane_wave_run(abi.SOLVER_LEHE, 0, prec=32)
    E64, _, _, _, _ = plane_wave_run(abi.SOLVER_LEHE, 0, prec=64)
    assert np.abs(E32 - E64).max() < 1e-5 * A


def test_gaussian_injection_lehe():
    """3-D Gaussian pulse from YMin (Lehe along y, x/y open with the absorber):
    the total-field region follows the functor, the scattered region stays dark.
    The surface's transverse extent is about 4 w0 wide so that its edges clip
    little of the beam; what remains is the paraxial profile not being an
    exact solution of the discrete Maxwell equations (a few percent)."""
    cell = (0.1e-6, 0.05e-6, 0.1e-6)
    dt = 0.995 * cell[1] / C_SI
    grid = (96, 96, 96)
    prof = dict(kind=abi.INCIDENT_GAUSSIAN_PULSE, wavelength_si=0.8e-6, amplitude_si=4e12, pulse_duration_si=3e-15,
                direction=(0, 1, 0), polarisation_direction=(1, 0, 0), focus_position_si=(0, 3e-6, 0),
                focus_origin=(abi.ORIGIN_CENTER, abi.ORIGIN_ZERO, abi.ORIGIN_CENTER), w0_si=1.2e-6, pulse_init=6.0)
    st = vacuum(grid, cell, dt, {(1, 0): prof}, solver=abi.SOLVER_LEHE, lehe_dir=1, boundary=(1, 1, 0),
                absorber=abi.ABSORBER_EXPONENTIAL, absorber_cells=((8, 8), (8, 8), (8, 8)),
                absorber_strength=((1e-3, 1e-3),) * 3)
    k, unit_e, cellp, dtp = units(st)
    A = 4e12 / unit_e
    yc = 3e-6 / cell[1]  # focus plane in cells
    n = int(round(pulse_centre_step(st, yc)))
    o = O.Oracle(st, precision=32)
    o.step(n)
    E = o.get_field(abi.FIELD_E).reshape(3, grid[2], grid[1], grid[0])
    o.close()
    # E_x at its Yee position (x + 0.5, y, z) in the total-field region around the axis
    xs = np.arange(40, 57)
    zs = np.arange(40, 57)
    ys = np.arange(24, 78)
    X, Y, Z = np.meshgrid(xs, ys, zs, indexing="ij")
    pos = np.stack([X + 0.5, Y, Z], axis=-1).reshape(-1, 3).astype(np.float64)
    ref = O.incident_field(st, 1, 0, 0, float(n), pos)[:, 0].reshape(X.shape)
    sim = E[0][np.ix_(zs, ys, xs)].transpose(2, 1, 0)
    assert np.abs(ref).max() > 0.9 * A
    assert np.abs(sim - ref).max() < 0.06 * A
    # scattered-field region (before the surface at y = 16)
    assert np.abs(E[:, :, :14, :]).max() < 0.02 * A
