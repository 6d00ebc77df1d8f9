"""GPU twins of the oracle pins in test_oracle_gather_energy.py: the HIP field
gather reproduces linear fields (seen through a Boris push in a pure E field,
where the momentum after one step is 2 * (q/2 * E * dt)), and the HIP field
energy equals the oracle's and the closed form."""
import math

import numpy as np
import pytest

from oracle import oracle as O
from picongpu_amd import abi, setups
from tests.test_oracle_gather_energy import GRID, gather_positions, linear_fields

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape", [abi.SHAPE_CIC, abi.SHAPE_TSC, abi.SHAPE_PQS])
def test_gather_of_linear_e_through_the_push(shape):
    """Particles at rest in E = a + b.x (B = 0): Boris gives p = 2 fl(fl(q/2 E) dt)
    exactly (the rotation vanishes), so E at the particle follows from p and
    must be a + b.X (FieldToParticleInterpolation.hpp:43-131, YeeCell.hpp:69-123)."""
    st = setups.thermal(grid=GRID, ppc=1, shape=shape)
    st.species[0].frames_per_sc = 2
    print(f"case linear-field gather shape={shape}", flush=True)
    scale = 2.0 ** -10  # dyadic: grid values stay exact in FP32, particles stay slow
    a = np.array([1.0, -0.5, 0.25, 0.0, 0.0, 0.0]) * scale
    b = np.array([1 / 64, 1 / 32, 5 / 128]) * scale
    E, _ = linear_fields(a, b)
    pos = gather_positions(9)
    n = pos.shape[0]
    rng = np.random.default_rng(4)
    cell = np.stack([rng.integers(4, 12, n), rng.integers(4, 12, n), rng.integers(4, 8, n)], 1).astype(np.int32)
    sim = abi.Simulation(st.config())
    k = sim.constants()
    vol = st.cell_si[0] * st.cell_si[1] * st.cell_si[2]
    w = np.full(n, np.float32(st.species[0].density_si * vol / st.species[0].ppc), np.float32)
    sim.load_particles(0, cell, pos, np.zeros((n, 3), np.float32), w)
    sim.load_field(abi.FIELD_E, E)
    sim.load_field(abi.FIELD_B, np.zeros_like(E))
    sim.push(0, fuse_deposit=False)
    _, _, mom, _ = sim.read_particles(0)
    sim.close()
    # read_particles lists supercell by supercell in load order (a stable binning)
    sc = cell[:, 0] // 8 + (GRID[0] // 8) * (cell[:, 1] // 8 + (GRID[1] // 8) * (cell[:, 2] // 4))
    order = np.argsort(sc, kind="stable")
    charge = np.float32(k.species_charge[0]) * w[0]
    hq_dt = np.float64(np.float32(0.5) * charge) * np.float64(np.float32(k.dt))
    e_at = mom.astype(np.float64) / (2.0 * hq_dt)
    X = cell[order].astype(np.float64) + pos[order].astype(np.float64)
    want = a[:3][None, :] + (X @ b)[:, None]
    err = np.abs(e_at - want).max() / scale
    print(f"linear-field gather shape={shape}: max error {err:.2e} of the field scale", flush=True)
    assert err < 4e-6


def test_field_energy_matches_oracle_and_closed_form():
    st = setups.thermal(grid=GRID, ppc=1)
    print("case field energy", flush=True)
    nx, ny, nz = GRID
    idx = np.arange(nx * ny * nz)
    x, y = idx % nx, (idx // nx) % ny
    E = np.zeros((3, idx.size), np.float32)
    B = np.zeros((3, idx.size), np.float32)
    E[1] = 0.75 * np.sin(2 * math.pi * 3 * x / nx)
    B[2] = 0.25 * np.cos(2 * math.pi * 2 * y / ny)
    rng = np.random.default_rng(8)
    E[0] = (rng.standard_normal(idx.size) * 1e-3).astype(np.float32)
    B[1] = (rng.standard_normal(idx.size) * 1e-3).astype(np.float32)
    sim = abi.Simulation(st.config())
    sim.load_field(abi.FIELD_E, E)
    sim.load_field(abi.FIELD_B, B)
    got = sim.field_energy()
    sim.close()
    ora = O.Oracle(st, 32, 0)
    ora.set_field(abi.FIELD_E, E)
    ora.set_field(abi.FIELD_B, B)
    want = ora.field_energy()
    ora.close()
    # the same FP32 squares summed in FP64 in another order
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=0.0)
    # numpy restatement: per component sum of the FP32 squares
    k = O.constants(st)
    unit_energy = k.unit_mass * k.unit_length ** 2 / k.unit_time ** 2
    vol = np.float32(np.float32(k.cell[0]) * np.float32(k.cell[1])) * np.float32(k.cell[2])
    b_fac = 0.5 / np.float64(np.float32(k.mue0)) * np.float64(vol)
    e_fac = np.float64(np.float32(np.float32(k.eps0) * vol)) * 0.5
    sb = [np.sum((B[c] * B[c]).astype(np.float64)) * b_fac * unit_energy for c in range(3)]
    se = [np.sum((E[c] * E[c]).astype(np.float64)) * e_fac * unit_energy for c in range(3)]
    np.testing.assert_allclose(got[1:4], sb, rtol=1e-12)
    np.testing.assert_allclose(got[4:7], se, rtol=1e-12)
