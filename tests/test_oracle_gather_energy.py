"""Pinning of three oracle functions by what the mathematics fixes:

- the field gather (FieldToParticleInterpolation.hpp:43-131 with
  ShiftCoordinateSystem.hpp:36-80 and the Yee staggering of YeeCell.hpp:69-123):
  B-splines of every order reproduce linear functions, so a field whose grid
  values are a + b.x at each component's own Yee position must gather to
  a + b.X at the particle, for CIC, TSC and PQS, all six components and
  in-cell positions on both sides of the shift's branch points;
- the Lehe curl E off the Cherenkov-free axis (Lehe/Derivative.hpp:173-228):
  an oblique standing wave (k along x and y, Cherenkov-free axis y) oscillates
  at the frequency of the scheme's discrete dispersion relation, derived here
  from the Fourier symbols of the implemented stencils;
- the field energy (plugins/EnergyFields.x.cpp:183-247): a sinusoidal field
  over whole periods has a sum of squares of exactly half the cell count, so
  the energy in joules is 0.5*eps0*E^2*V/2 per cell (E) and 0.5/mu0*B^2*V/2 (B)
  in SI units.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from picongpu_amd import abi, setups

E_OFF = [(0.5, 0.0, 0.0), (0.0, 0.5, 0.0), (0.0, 0.0, 0.5)]
B_OFF = [(0.0, 0.5, 0.5), (0.5, 0.0, 0.5), (0.5, 0.5, 0.0)]
GRID = (16, 16, 12)


def linear_fields(a, b):
    """E and B planes whose component c holds a[c] + b . (node + offset_c) at
    every node (x fastest)."""
    nx, ny, nz = GRID
    idx = np.arange(nx * ny * nz)
    node = np.stack([idx % nx, (idx // nx) % ny, idx // (nx * ny)]).astype(np.float64)
    E = np.zeros((3, idx.size), np.float32)
    B = np.zeros((3, idx.size), np.float32)
    for c in range(3):
        E[c] = a[c] + sum(b[d] * (node[d] + E_OFF[c][d]) for d in range(3))
        B[c] = a[3 + c] + sum(b[d] * (node[d] + B_OFF[c][d]) for d in range(3))
    return E, B


def gather_positions(seed):
    """In-cell positions on the branch points of the shift (0, 1/2 and the
    largest float below 1 on every axis) plus random ones."""
    special = np.array([0.0, 0.5, np.nextafter(np.float32(1), np.float32(0)), 0.25, 0.75], np.float32)
    grid = np.stack(np.meshgrid(special, special, special, indexing="ij"), -1).reshape(-1, 3)
    rng = np.random.default_rng(seed)
    return np.concatenate([grid, rng.random((64, 3), dtype=np.float32)])


@pytest.mark.parametrize("precision", [32, 64])
@pytest.mark.parametrize("shape", [abi.SHAPE_CIC, abi.SHAPE_TSC, abi.SHAPE_PQS])
def test_gather_reproduces_linear_fields(shape, precision):
    st = setups.thermal(grid=GRID, ppc=1, shape=shape)
    ora = O.Oracle(st, precision, 0)
    a = np.array([1.0, -0.5, 0.25, 0.75, -1.25, 0.5])
    # distinct per axis (a transposed weight is visible) and dyadic, so every
    # grid value a + b.(node + offset) is exact in FP32
    b = np.array([1 / 64, 1 / 32, 5 / 128])
    E, B = linear_fields(a, b)
    ora.set_field(abi.FIELD_E, E)
    ora.set_field(abi.FIELD_B, B)
    rng = np.random.default_rng(3)
    worst = 0.0
    for pos in gather_positions(5):
        # cells away from the periodic seam, so the linear function has no jump in the stencil
        cell = np.array([rng.integers(4, 12), rng.integers(4, 12), rng.integers(4, 8)], np.int32)
        got = ora.gather(shape, cell, pos)
        X = cell.astype(np.float64) + pos.astype(np.float64)
        want = a + np.concatenate([np.full(3, b @ X), np.full(3, b @ X)])
        worst = max(worst, float(np.abs(got - want).max()))
    # FP32: values ~1.5 summed over up to 64 nodes; FP64: round-off only
    assert worst < (2e-6 if precision == 32 else 1e-13), worst
    ora.close()


def test_gather_catches_a_wrong_stagger():
    """The check above fails when the field is placed at the wrong Yee
    position: E_x sampled at the node instead of (1/2, 0, 0) is off by b_x/2."""
    st = setups.thermal(grid=GRID, ppc=1, shape=abi.SHAPE_TSC)
    ora = O.Oracle(st, 64, 0)
    a = np.zeros(6)
    b = np.array([1 / 64, 1 / 32, 5 / 128])
    E, B = linear_fields(a, b)
    nx, ny, nz = GRID
    idx = np.arange(nx * ny * nz)
    E[0] = b[0] * (idx % nx) + b[1] * ((idx // nx) % ny) + b[2] * (idx // (nx * ny))  # no x offset
    ora.set_field(abi.FIELD_E, E)
    ora.set_field(abi.FIELD_B, B)
    cell = np.array([6, 6, 5], np.int32)
    pos = np.array([0.3, 0.6, 0.2], np.float32)
    got = ora.gather(abi.SHAPE_TSC, cell, pos)
    X = cell + pos.astype(np.float64)
    assert abs(got[0] - b @ X) == pytest.approx(b[0] / 2, rel=1e-6)  # positions pass through FP32
    assert abs(got[1] - b @ X) < 1e-13
    ora.close()


# ---------------------------------------------------------------------------
# Lehe curl E off the Cherenkov-free axis
# ---------------------------------------------------------------------------
def lehe_oblique_omega(cell, cdt, kx, ky, beta_off):
    """Frequency (per step) of the mode E_z ~ cos(kx i) cos(ky j) (kx, ky the
    phase per cell) of the Yee E update with the Lehe B update, Cherenkov-free
    axis y. With s_d = sin(k_d / 2), the forward difference has the symbol 2i s/h, the
    off-axis Lehe derivative along x multiplies it by alpha' + 2 beta' cos(ky)
    = 1 - 4 beta' s_y^2, and the on-axis one along y by
    alpha + 2 beta_1 cos(kz) + 2 beta_2 cos(kx) plus delta sin(3 ky / 2)/s_y,
    i.e. 1 - 4 beta_1 s_z^2 - 4 beta_2 s_x^2 - 4 delta s_y^2. Leapfrog:
    sin^2(omega dt / 2) / (c dt)^2 = sum of the products over x and y."""
    hx, hy, hz = cell
    sx2 = math.sin(kx / 2) ** 2  # kx, ky: phase per cell
    sy2 = math.sin(ky / 2) ** 2
    co = hy / cdt * math.sin(math.pi / 2 * cdt / hy)
    delta = 0.25 * (1 - co * co)
    beta2 = 0.125 * (hy / hx) ** 2  # the on-axis stencil's average along x
    rhs = sx2 / hx ** 2 * (1 - 4 * beta_off * sy2) + sy2 / hy ** 2 * (1 - 4 * beta2 * sx2 - 4 * delta * sy2)
    return 2 * math.asin(math.sqrt(rhs) * cdt)


def test_lehe_off_axis_oblique_standing_wave():
    nx, ny, nz = 16, 32, 4
    st = setups.thermal(grid=(nx, ny, nz), ppc=1)
    st.species = []
    st.solver = abi.SOLVER_LEHE
    st.lehe_dir = 1
    st.cell_si = (1e-7, 5e-8, 1e-7)  # unequal cells: the off-axis terms do not cancel
    st.dt_si = 0.99 * st.cell_si[1] / setups.SPEED_OF_LIGHT_SI
    ora = O.Oracle(st, 64, 1)
    k = O.constants(st)
    cell = [c / (st.dt_si * setups.SPEED_OF_LIGHT_SI) for c in st.cell_si]  # FP64 twin cell sizes
    cdt = float(k.c) * float(k.dt)
    kx = 2 * math.pi * 2 / nx
    ky = 2 * math.pi * 5 / ny
    idx = np.arange(nx * ny * nz)
    x, y = idx % nx, (idx // nx) % ny
    basis = np.cos(kx * x) * np.cos(ky * y)  # E_z sits on the xy nodes
    E = np.zeros((3, idx.size), np.float32)
    E[2] = basis
    ora.set_field(abi.FIELD_E, E)
    amps = []
    for _ in range(40):
        ora.update_before_current()
        ora.add_current()
        ora.update_after_current()
        amps.append(float(ora.get_field(abi.FIELD_E)[2] @ basis / (basis @ basis)))
    n = np.arange(1, 41)
    omega = lehe_oblique_omega(cell, cdt, kx, ky, beta_off=0.125)
    np.testing.assert_allclose(amps, np.cos(omega * n), atol=1e-7)  # initial field enters through FP32
    # the off-axis average is what makes the difference: without it (beta' = 0)
    # the same mode drifts visibly within the 40 steps
    omega0 = lehe_oblique_omega(cell, cdt, kx, ky, beta_off=0.0)
    assert np.abs(np.cos(omega0 * n) - np.cos(omega * n)).max() > 1e-3
    ora.close()


# ---------------------------------------------------------------------------
# field energy
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("precision", [32, 64])
def test_field_energy_closed_form(precision):
    nx, ny, nz = GRID
    st = setups.thermal(grid=GRID, ppc=1)
    ora = O.Oracle(st, precision, 0)
    k = O.constants(st)
    idx = np.arange(nx * ny * nz)
    x, y = idx % nx, (idx // nx) % ny
    e_amp, b_amp = 0.75, 0.25  # PIC units
    E = np.zeros((3, idx.size), np.float32)
    B = np.zeros((3, idx.size), np.float32)
    E[1] = e_amp * np.sin(2 * math.pi * 3 * x / nx)
    B[2] = b_amp * np.cos(2 * math.pi * 2 * y / ny)
    ora.set_field(abi.FIELD_E, E)
    ora.set_field(abi.FIELD_B, B)
    out = ora.field_energy()
    # SI: E and B in V/m and T, cell volume in m^3
    unit_e = k.unit_mass * k.unit_length / (k.unit_time ** 2 * k.unit_charge)
    unit_b = k.unit_mass / (k.unit_time * k.unit_charge)
    vol = st.cell_si[0] * st.cell_si[1] * st.cell_si[2]
    mu0 = 1.25663706127e-6  # param/physicalConstants.param
    eps0 = 1.0 / (mu0 * setups.SPEED_OF_LIGHT_SI ** 2)
    half = idx.size / 2  # sum of sin^2 (cos^2) over whole periods
    want_e = 0.5 * eps0 * (e_amp * unit_e) ** 2 * vol * half
    want_b = 0.5 / mu0 * (b_amp * unit_b) ** 2 * vol * half
    assert out[5] == pytest.approx(want_e, rel=2e-6)  # E_y; the constants pass through FP32 in PIC units
    assert out[3] == pytest.approx(want_b, rel=2e-6)  # B_z
    assert out[1] == out[2] == out[4] == out[6] == 0.0
    assert out[0] == pytest.approx(want_e + want_b, rel=2e-6)
    ora.close()
