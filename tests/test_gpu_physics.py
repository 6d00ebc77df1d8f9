"""GPU checks beyond slot-exact parity: agreement with the reference CPU's
sequential summation order (1e-5) and with the FP64 twin (1e-4 after 100
steps, max-norm relative), fused against unfused deposition, the device initial
state against an independent restatement of its generator,
size-independent properties at the full Kelvin-Helmholtz size, and the
Kelvin-Helmholtz and filamentation growth rates against linear theory."""
import math

import numpy as np
import pytest

from oracle import oracle as O
from picongpu_amd import abi, setups
from tests import helpers as H

pytestmark = pytest.mark.gpu


def rel_max(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def small_thermal(shape=abi.SHAPE_CIC, pusher=abi.PUSHER_BORIS, grid=(16, 16, 12), ppc=4, u=0.05):
    st = setups.thermal(grid=grid, ppc=ppc, shape=shape, pusher=pusher)
    st.species[0].u_th = (u, u, u)
    return st


def global_state(st, frames, ids, cap, nsc):
    valid = (np.arange(cap)[None, :] < frames["counts"][:, None].astype(np.int64)).ravel()
    sc = np.repeat(np.arange(nsc), cap)[valid]
    nscx, nscy = st.grid[0] // 8, st.grid[1] // 8
    lc = frames["cell"][valid].astype(np.int64)
    g = np.stack([(sc % nscx) * 8 + lc % 8, ((sc // nscx) % nscy) * 8 + (lc // 8) % 8,
                  (sc // (nscx * nscy)) * 4 + lc // 64]).astype(np.float64) + np.asarray(frames["pos"], np.float64)[:, valid]
    order = np.argsort(ids[valid])
    return g[:, order], np.asarray(frames["mom"], np.float64)[:, valid][:, order], lc[order]


def test_against_sequential_order_one_step():
    """One step against the reference CPU's arithmetic and sequential FP32
    sums: the push does not see this step's J, so the particles differ only
    by the device's multiply-adds in the field gather (a few ulp: in-cell
    positions within 4 * 2^-24, momenta within 1e-6 relative, the same cells
    and slots); J within 1e-5 (max-norm)."""
    st = small_thermal(abi.SHAPE_TSC)
    print(f"case sequential-order grid={st.grid}", flush=True)
    sim, ora, frames = H.make_pair(st, order=1)
    sim.step(1)
    ora.step(1)
    g = sim.read_frames(0)
    o = ora.get_frames(0)
    np.testing.assert_array_equal(g["counts"], o["counts"])
    v = H.occupied(g)
    np.testing.assert_array_equal(g["cell"][v], o["cell"][v])
    np.testing.assert_array_equal(g["mask"][v], o["mask"][v])
    dpos = float(np.abs(g["pos"][:, v].astype(np.float64) - o["pos"][:, v]).max())
    dmom = rel_max(g["mom"][:, v], o["mom"][:, v])
    print(f"particles vs sequential order: in-cell position {dpos:.3e}, momentum {dmom:.3e}")
    assert dpos <= 4 * 2.0 ** -24 and dmom < 1e-6
    rJ = rel_max(sim.read_field(abi.FIELD_J), ora.get_field(abi.FIELD_J))
    rE = rel_max(sim.read_field(abi.FIELD_E), ora.get_field(abi.FIELD_E))
    print("J vs sequential order:", rJ, "E:", rE)
    assert rJ < 1e-5 and rE < 1e-5
    sim.close()


@pytest.mark.parametrize("shape", [abi.SHAPE_CIC, abi.SHAPE_TSC])
def test_against_fp64_twin_100_steps(shape):
    st = small_thermal(shape, u=0.01)
    print(f"case fp64 twin shape={shape} 100 steps", flush=True)
    sim, o32, frames = H.make_pair(st, e_amp=0.0, b_amp=0.0)
    o64 = O.Oracle(st, 64, 1)
    o64.set_frames(0, frames[0])
    sim.step(100)
    o32.step(100)
    o64.step(100)
    cap = o32.cap(0)
    # the GPU's frame layout equals the device-order FP32 oracle's, so its ids apply
    H.assert_frames_equal(sim.read_frames(0), o32.get_frames(0), "gpu vs fp32 oracle after 100 steps")
    ids32 = o32.get_ids(0)
    gpos, gmom, glc = global_state(st, sim.read_frames(0), ids32, cap, o32.nsc)
    dpos, dmom, dlc = global_state(st, o64.get_frames(0), o64.get_ids(0), cap, o64.nsc)
    L = np.array(st.grid, np.float64)[:, None]
    err = {"pos_cells": float(np.abs((gpos - dpos + L / 2) % L - L / 2).max()), "mom": rel_max(gmom, dmom)}
    for f, nm in ((abi.FIELD_E, "E"), (abi.FIELD_B, "B"), (abi.FIELD_J, "J")):
        err[nm] = rel_max(sim.read_field(f), o64.get_field(f))
    err["cell_flips"] = int(np.count_nonzero(glc != dlc))
    print("GPU vs FP64 twin after 100 steps:", err)
    assert err["pos_cells"] < 1e-4 and err["mom"] < 1e-4
    for nm in ("E", "B", "J"):
        assert err[nm] < 1e-4
    sim.close()


def test_fused_and_unfused_deposit_agree():
    st = small_thermal(abi.SHAPE_TSC)
    print("case fused vs unfused", flush=True)
    a, _, frames = H.make_pair(st, with_sim=True)
    b = abi.Simulation(st.config())
    b.write_frames(0, frames[0])
    E, B = H.random_fields(st, 11, 1e-3, 1e-3)
    b.load_field(abi.FIELD_E, E)
    b.load_field(abi.FIELD_B, B)
    a.current_reset()
    a.push(-1, True)
    a.shift(-1)
    b.current_reset()
    b.push(-1, False)
    b.shift(-1)
    b.deposit(-1)
    H.assert_frames_equal(a.read_frames(0), {k: v for k, v in b.read_frames(0).items()}, "fused vs unfused push")
    r = rel_max(a.read_field(abi.FIELD_J), b.read_field(abi.FIELD_J))
    print("fused vs unfused J:", r)
    assert r < 1e-6
    a.close()
    b.close()


# --- device initial state against an independent numpy restatement -------
M64 = (1 << 64) - 1


def splitmix64(x):
    x = (x + 0x9E3779B97F4A7C15) & M64
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & M64
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & M64
    return x ^ (x >> 31)


def rng_key(seed, species, cell, k, stream):
    h = splitmix64(seed ^ (species << 56))
    h = splitmix64(h ^ cell)
    return splitmix64(h ^ ((k << 16) | stream))


def uniform_f(bits):
    return np.float32(np.float32(bits >> 40) * np.float32(1.0 / 16777216.0))


def uniform_d(bits):
    return float(bits >> 11) * (1.0 / 9007199254740992.0)


def det_log(x):
    m, e = math.frexp(x)  # x = m 2^e, m in [0.5,1)
    m, e = m * 2.0, e - 1  # [1,2)
    if m > 1.4142135623730951:
        m, e = m * 0.5, e + 1
    s = (m - 1.0) / (m + 1.0)
    s2 = s * s
    term = s
    acc = s
    for k in range(3, 34, 2):
        term = term * s2
        acc = acc + term / float(k)
    return 2.0 * acc + float(e) * 0.69314718055994530942


def normal_pair(seed, species, cell, k, family):
    for attempt in range(1000):
        st = 8 + family * 2048 + attempt * 2
        u1 = 2.0 * uniform_d(rng_key(seed, species, cell, k, st)) - 1.0
        u2 = 2.0 * uniform_d(rng_key(seed, species, cell, k, st + 1)) - 1.0
        s = u1 * u1 + u2 * u2
        if 0.0 < s < 1.0:
            f = math.sqrt(-2.0 * det_log(s) / s)
            return u1 * f, u2 * f
    return 0.0, 0.0


def test_device_init_matches_generator_restatement():
    st = setups.kelvin_helmholtz(grid=(16, 16, 8), ppc=2)
    print("case device init", flush=True)
    sim = abi.Simulation(st.config())
    setups.init_device(sim, st)
    k = sim.constants()
    s = st.species[0]
    f = sim.read_frames(0)
    cap = sim.slots(0)
    vol = st.cell_si[0] * st.cell_si[1] * st.cell_si[2]
    w = np.float32(s.density_si * vol / s.ppc)
    mass = np.float32(np.float32(k.species_mass[0]) * w)
    nscx, nscy = st.grid[0] // 8, st.grid[1] // 8
    rng = np.random.default_rng(3)
    for sc in rng.choice(sim.nsc, 6, replace=False):
        assert f["counts"][sc] == s.ppc * 256
        sx, sy, sz = sc % nscx, (sc // nscx) % nscy, sc // (nscx * nscy)
        for i in rng.choice(s.ppc * 256, 20, replace=False):
            g = sc * cap + i
            kk, lc = int(i) // 256, int(i) % 256
            gx, gy, gz = int(sx) * 8 + lc % 8, int(sy) * 8 + (lc // 8) % 8, int(sz) * 4 + lc // 64
            cell = gx + st.grid[0] * (gy + st.grid[1] * gz)
            pos = [uniform_f(rng_key(s.seed, 0, cell, kk, a)) for a in range(3)]
            n0, n1 = normal_pair(s.seed, 0, cell, kk, 0)
            n2, _ = normal_pair(s.seed, 0, cell, kk, 1)
            sign = 1.0 if (st.grid[1] // 4 <= gy < 3 * st.grid[1] // 4) else -1.0
            m = float(mass)
            mom = [np.float32(m * (sign * s.u_drift[0] + s.u_th[0] * n0)),
                   np.float32(m * (sign * s.u_drift[1] + s.u_th[1] * n1)),
                   np.float32(m * (sign * s.u_drift[2] + s.u_th[2] * n2))]
            assert f["cell"][g] == lc and f["w"][g] == w
            np.testing.assert_array_equal(H.bits(f["pos"][:, g]), H.bits(np.array(pos, np.float32)))
            np.testing.assert_array_equal(H.bits(f["mom"][:, g]), H.bits(np.array(mom, np.float32)))
    sim.close()


def test_quiet_start_lattice():
    """Quiet start (startPosition/QuietImpl.hpp:47-72): particle k of every cell
    sits on lattice point ppc-1-k of a 3 x 2 x 2 in-cell lattice (x fastest),
    at q * (1/n) + (1/n) * 0.5 per axis in FP32; the momenta are those of the
    random start of the same seed."""
    quiet = (3, 2, 2)
    st = setups.thermal(grid=(16, 16, 8), ppc=12)
    print("case quiet start", flush=True)
    sims = []
    for q in (quiet, None):
        st.species[0].quiet = q
        sim = abi.Simulation(st.config())
        setups.init_device(sim, st)
        sims.append(sim)
    fq, fr = (s.read_frames(0) for s in sims)
    cap = sims[0].slots(0)
    spacing = (np.float32(1) / np.array(quiet, np.float32)).astype(np.float32)
    for sc in range(sims[0].nsc):
        n = int(fq["counts"][sc])
        assert n == 12 * 256
        k = np.arange(n) // 256
        idx = 12 - 1 - k
        q = np.stack([idx % quiet[0], (idx // quiet[0]) % quiet[1], idx // (quiet[0] * quiet[1])])
        expect = (q.astype(np.float32) * spacing[:, None] + spacing[:, None] * np.float32(0.5)).astype(np.float32)
        sl = slice(sc * cap, sc * cap + n)
        np.testing.assert_array_equal(H.bits(fq["pos"][:, sl]), H.bits(expect))
        np.testing.assert_array_equal(H.bits(fq["mom"][:, sl]), H.bits(fr["mom"][:, sl]))
        np.testing.assert_array_equal(fq["cell"][sl], fr["cell"][sl])
    for s in sims:
        s.close()


def test_gauss_residual_matches_oracle():
    st = small_thermal(abi.SHAPE_TSC)
    print("case gauss residual", flush=True)
    sim, ora, _ = H.make_pair(st)
    sim.step(3)
    ora.step(3)
    k = sim.constants()
    r_gpu = sim.gauss_residual()
    rho = ora.charge_density()
    r_ora = float(np.float32(k.eps0)) * ora.div_e() - rho
    assert np.abs(r_gpu - r_ora).max() <= 1e-12 * np.abs(rho).max()
    sim.close()


def test_full_size_kelvin_helmholtz_properties():
    """256^3, 2 x 16 ppc (536,870,912 particles): counts conserved, no capacity
    error, Gauss residual unchanged to FP32 round-off (charge conservation),
    field energy finite and growing from the seeded noise."""
    st = setups.kelvin_helmholtz()
    print("case full-size KH 256^3, 4 steps", flush=True)
    sim = abi.Simulation(st.config())
    setups.init_device(sim, st)
    n0 = [sim.particle_count(i) for i in range(2)]
    e0 = sim.field_energy()
    r0 = sim.gauss_residual()
    sim.step(4)
    n1 = [sim.particle_count(i) for i in range(2)]
    e1 = sim.field_energy()
    r1 = sim.gauss_residual()
    assert n0 == n1 == [256 ** 3 * 16] * 2
    assert sim.error_flags() == 0
    assert np.all(np.isfinite(e1)) and e1[0] > e0[0]
    # both species at the same density: |rho| per species ~ 1/V; use that scale
    k = sim.constants()
    scale = abs(float(k.species_charge[0])) * k.typical_weighting * float(k.inv_cell_volume_dt)
    drift = np.abs(r1 - r0).max() / scale
    print("Gauss residual drift / rho_species:", drift)
    assert drift < 2e-6
    sim.close()


def macro_rule(w0, ppc, f):
    """WeightMacroParticles (startPosition/detail/WeightMacroParticles.hpp:30-42)
    per cell of density factor f: floor(min(ppc, N / 10)) macro particles of
    weighting N / count, N = w0 * ppc * f real particles in FP32; where the
    count is ppc the weighting is w0 * f."""
    real = (np.float64(w0) * ppc * f).astype(np.float32)
    count = np.floor(np.minimum(np.float32(ppc), real / np.float32(10.0))).astype(np.int64)
    count = np.where(f > 0.0, count, 0)
    w_low = np.where(count > 0, real / np.maximum(count, 1).astype(np.float32), 0).astype(np.float32)
    w = np.where(count == ppc, (np.float64(w0) * f).astype(np.float32), w_low)
    return count, w


@pytest.mark.parametrize("dense", [True, False])
def test_density_ramp_init(dense):
    """Linear ramp along y: cells whose origin lies before the ramp start get
    no particles, the others WeightMacroParticles' count, never a weighting
    below MIN_WEIGHTING; the dilute case (40 real particles per macro
    particle at full density) gives the ramp's first cells 0, 1 and 2
    particles instead of 3."""
    st = setups.two_stream(grid=(16, 32, 8), ppc=3)
    dy = st.cell_si[1]
    vol = st.cell_si[0] * st.cell_si[1] * st.cell_si[2]
    if not dense:
        st.species[0].density_si = 40.0 * 3 / vol
    st.species[0].ramp = (1, 8 * dy, 16 * dy)
    print(f"case density ramp dense={dense}", flush=True)
    sim = abi.Simulation(st.config())
    setups.init_device(sim, st)
    cells, pos, mom, w = sim.read_particles(0)
    w0 = st.species[0].density_si * vol / st.species[0].ppc
    start, length = (8 * dy) / dy, (16 * dy) / dy  # the ramp in cells, as the library converts it
    fy = np.clip((np.arange(32) - start) / length, 0.0, 1.0)
    count_y, w_y = macro_rule(w0, 3, fy)
    gy = cells[:, 1].astype(np.int64)
    np.testing.assert_array_equal(H.bits(w), H.bits(w_y[gy]))
    assert np.all(w >= 10.0)
    flat = cells[:, 0] + 16 * (cells[:, 1] + 32 * cells[:, 2])
    per_cell = np.bincount(flat, minlength=16 * 32 * 8).reshape(8, 32, 16)
    assert np.all(per_cell == count_y[None, :, None])
    if not dense:
        assert list(count_y[9:13]) == [0, 1, 2, 3]
    sim.close()


def test_laser_initial_field():
    """The Gaussian pulse at the Yee positions against its formula (FP64
    evaluated on the host, rounded once to FP32)."""
    st = setups.two_stream(grid=(16, 64, 16), ppc=1)
    cell = np.array(st.cell_si)
    laser = dict(wavelength_si=0.8e-6, a0=1.5, waist_si=0.6e-6, fwhm_si=5e-15, center_si=(0.8e-6, 3.2e-6, 0.8e-6),
                 propagation_axis=1, polarisation_axis=0)
    print("case laser initial field", flush=True)
    sim = abi.Simulation(st.config())
    sim.add_laser(**laser)
    E = sim.read_field(abi.FIELD_E).reshape(3, 16, 64, 16)
    B = sim.read_field(abi.FIELD_B).reshape(3, 16, 64, 16)
    k = sim.constants()
    L = k.unit_length
    unit_e = k.unit_mass * L / (k.unit_time ** 2 * k.unit_charge)
    c_si, me, qe = setups.SPEED_OF_LIGHT_SI, 9.1093837139e-31, 1.602176634e-19
    omega = 2 * math.pi * c_si / laser["wavelength_si"]
    e0 = laser["a0"] * me * c_si * omega / qe / unit_e
    kk = 2 * math.pi / (laser["wavelength_si"] / L)
    w0 = laser["waist_si"] / L
    sigma = c_si * laser["fwhm_si"] / (2 * math.sqrt(math.log(2))) / L
    cpic = [float(np.float32(v)) for v in k.cell]
    ctr = np.array(laser["center_si"]) / L
    z, y, x = np.meshgrid(np.arange(16), np.arange(64), np.arange(16), indexing="ij")

    def pulse(off):
        px = (x + off[0]) * cpic[0] - ctr[0]
        py = (y + off[1]) * cpic[1] - ctr[1]
        pz = (z + off[2]) * cpic[2] - ctr[2]
        return e0 * np.exp(-(px * px + pz * pz) / (w0 * w0)) * np.exp(-py * py / (2 * sigma * sigma)) * np.cos(kk * py)

    ex = pulse((0.5, 0.0, 0.0)).astype(np.float32)
    bz = (-pulse((0.5, 0.5, 0.0))).astype(np.float32)
    assert np.abs(E[0] - ex).max() <= 2e-7 * np.abs(ex).max()
    assert np.abs(B[2] - bz).max() <= 2e-7 * np.abs(bz).max()
    assert np.abs(E[1]).max() == 0.0 and np.abs(E[2]).max() == 0.0 and np.abs(B[0]).max() == 0.0
    assert np.abs(ex).max() > 0.5 * e0
    sim.close()


def test_khi_growth_rate_matches_theory():
    """The reference's KHI growth-rate test (share/picongpu/tests/KHI_growthRate,
    validated by lib/python/test/setups/ESKHI/config.py): on its setup (192 x 512
    x 12 cells, 2 x 25 quiet-start particles per cell, gamma 1.021 shear flow,
    3000 steps) the growth rate of the B_x field energy f,
    Gamma(t_k) = 0.5 ln(f(t_k+1) / f(t_k-1)) / (t_k+1 - t_k-1) with f every 10
    steps and t in units of the relativistic plasma frequency
    sqrt(n e^2 / (eps0 gamma m_e)) (testsuite Math/math.py growthRate,
    Math/physics.py plasmafrequence), must reach 1 / (sqrt(8) gamma) within the
    acceptance of 20 %, judged on the maximum of Gamma (Math/deviation.py
    getTestResult)."""
    st = setups.khi_growth_rate()
    gamma = 1.021
    steps, period = 3000, 10
    print(f"case KHI growth rate grid={st.grid} steps={steps}", flush=True)
    sim = abi.Simulation(st.config())
    setups.init_device(sim, st)
    f = []
    for k in range(steps // period + 1):
        if k:
            sim.step(period)
        bx = sim.read_field(abi.FIELD_B)[0].astype(np.float64)
        f.append(float(np.sum(bx * bx)))
        if k % 50 == 0:
            print(f"  step {k * period} sum Bx^2 {f[-1]:.4e}", flush=True)
    sim.close()
    f = np.array(f)
    e, eps0, me = 1.602176634e-19, 8.8541878188e-12, 9.1093837139e-31
    omega = math.sqrt(1e25 * e * e / (eps0 * gamma * me))
    t = np.arange(len(f)) * period * st.dt_si * omega
    growth = 0.5 * np.log(f[3:] / f[1:-2]) / (t[3:] - t[1:-2])
    theory = 1.0 / (math.sqrt(8.0) * gamma)
    deviation = abs(theory - growth.max()) / theory
    print(f"  max growth rate {growth.max():.4f} (theory {theory:.4f}), deviation {100 * deviation:.1f} %", flush=True)
    assert deviation <= 0.2


def filamentation_rate(k_c_over_omega, beta):
    """Growth rate (units of omega_pe, the non-relativistic plasma frequency
    of the total density) of the purely transverse filamentation mode of two
    symmetric cold beams at +-beta c: the positive root of
    G^4 + G^2 (K^2 + 1/gamma^3) - K^2 beta^2 / gamma = 0, K = k c / omega_pe
    (cold relativistic two-fluid beams with an immobile neutralising
    background; large K gives beta / sqrt(gamma))."""
    gamma = 1.0 / math.sqrt(1.0 - beta * beta)
    K2 = k_c_over_omega ** 2
    b = K2 + gamma ** -3
    c = K2 * beta * beta / gamma
    # the positive root of x^2 + b x - c, in the form free of cancellation
    return math.sqrt(2.0 * c / (b + math.sqrt(b * b + 4.0 * c)))


def test_filamentation_growth_rate_matches_theory():
    """Filamentation (Weibel) growth of counter-streaming beams, judged in the
    form of the reference's MI test (lib/python/test/setups/MI/config.py:
    theory v / (c sqrt(gamma)) in units of the non-relativistic plasma
    frequency, acceptance 20 %). Two electron beams at +-0.3 c along z, each
    of half the 1e25 m^-3 density, cold (0.001 c spread, quiet start), on
    64 x 64 x 16 cells: the 16-cell z extent leaves the two-stream modes
    stable (they need k_z v < omega_pe / gamma^1.5 = 0.93 omega_pe; the
    longest z mode has k_z v = 1.98 omega_pe). The power of each k_z = 0
    transverse mode of (Bx, By) is fitted log-linearly over its linear phase
    (steps 350..700); every fitted rate must lie within 10 % of the cold
    dispersion relation's root at that k, and the fastest within 20 % of
    beta / sqrt(gamma)."""
    beta = 0.3
    st = setups.two_stream(grid=(64, 64, 16), ppc=8, beta=beta, seed=21)
    # the theory is the cold-beam rate: a quiet start (the beams' particles
    # paired on a 2 x 2 x 2 lattice, so the initial current noise is small)
    # and a 0.001 c thermal spread
    st.species[0].quiet = (2, 2, 2)
    st.species[0].u_th = (1e-3, 1e-3, 1e-3)
    steps, period = 700, 10
    print(f"case filamentation growth rate grid={st.grid} steps={steps}", flush=True)
    sim = abi.Simulation(st.config())
    setups.init_device(sim, st)
    modes = []
    for k in range(steps // period + 1):
        if k:
            sim.step(period)
        b = sim.read_field(abi.FIELD_B).astype(np.float64).reshape(3, st.grid[2], st.grid[1], st.grid[0])
        # power of the k_z = 0 transverse modes: x-y Fourier plane of the z mean, [ky, kx]
        modes.append(np.sum(np.abs(np.fft.fft2(b[:2].mean(axis=1))) ** 2, axis=0))
    sim.close()
    e, eps0, me, c = 1.602176634e-19, 8.8541878188e-12, 9.1093837139e-31, 2.99792458e8
    omega = math.sqrt(1e25 * e * e / (eps0 * me))
    t = np.arange(len(modes)) * period * st.dt_si * omega
    window = slice(350 // period, 700 // period + 1)
    P = np.log(np.array(modes[window]))  # B starts at zero: the log only over the window
    t = t[window]
    L = st.grid[0] * st.cell_si[0]
    fastest = 0.0
    for ny, nx in ((0, 1), (1, 0), (1, 1), (0, 2), (2, 0), (2, 1), (0, 3), (3, 0), (2, 2), (0, 4), (4, 0)):
        rate = np.polyfit(t, P[:, ny, nx], 1)[0] / 2.0
        kc = 2.0 * math.pi * math.hypot(nx, ny) / L * c / omega
        expect = filamentation_rate(kc, beta)
        print(f"  mode (kx={nx}, ky={ny}) k c/omega_pe {kc:.2f}: rate {rate:.4f}, cold theory {expect:.4f}", flush=True)
        assert abs(rate - expect) <= 0.1 * expect
        fastest = max(fastest, rate)
    gamma = 1.0 / math.sqrt(1.0 - beta * beta)
    theory = beta / math.sqrt(gamma)
    deviation = abs(theory - fastest) / theory
    print(f"  fastest mode {fastest:.4f} (beta/sqrt(gamma) {theory:.4f}), deviation {100 * deviation:.1f} %", flush=True)
    assert deviation <= 0.2


@pytest.mark.parametrize("shape", [abi.SHAPE_CIC, abi.SHAPE_TSC, abi.SHAPE_PQS])
@pytest.mark.parametrize("direction", [(1, 0, 0), (1, 1, 0), (1, 1, 1)])
def test_current_deposition_setup_closed_form(shape, direction):
    """The reference's CurrentDeposition test (share/picongpu/tests/CurrentDeposition)
    through the GPU step: 24^3 cells, dt = dx / (2c), one electron at cell
    (8,8,8) offset (0.9,0.8,0.7) moving at beta = 0.999 along the direction,
    zero fields. J after one pic_step must equal Esirkepov's FP64 closed form
    for that move within 1e-5 of its maximum (the reference's tolerance)."""
    from tests.test_oracle_pinning import esirkepov_closed_form

    st = setups.thermal(grid=(24, 24, 24), shape=shape)
    st.species[0].frames_per_sc = 1
    st.cell_si = (1e-7, 1e-7, 1e-7)
    st.dt_si = 0.5 * st.cell_si[0] / setups.SPEED_OF_LIGHT_SI
    print(f"case CurrentDeposition setup shape={shape} direction={direction}", flush=True)
    sim = abi.Simulation(st.config())
    k = sim.constants()
    vol = st.cell_si[0] ** 3
    w = np.float32(st.species[0].density_si * vol / st.species[0].ppc)
    mass = float(np.float32(k.species_mass[0]) * w)
    u = np.array(direction, np.float64)
    beta = 0.999 * u / np.linalg.norm(u)
    gamma = 1.0 / math.sqrt(1.0 - 0.999 ** 2)
    cells = np.array([[8, 8, 8]], np.int32)
    pos = np.array([[0.9, 0.8, 0.7]], np.float32)
    mom = (gamma * mass * float(k.c) * beta)[None, :].astype(np.float32)
    f = setups.bin_frames(st, 0, cells, pos, mom, np.array([w], np.float32))
    sim.write_frames(0, f)
    zero = np.zeros((3, sim.ncells), np.float32)
    sim.load_field(abi.FIELD_E, zero)
    sim.load_field(abi.FIELD_B, zero)
    sim.step(1)
    J = sim.read_field(abi.FIELD_J).astype(np.float64)
    sim.close()
    cell = np.array(k.cell, np.float64)
    dt = float(k.dt)
    d = beta * float(k.c) * dt / cell
    end = np.array([8.9, 8.8, 8.7]) + d
    c = np.floor(end).astype(np.int64)
    q = float(np.float32(k.species_charge[0]) * w)
    Jref = esirkepov_closed_form(shape, c, end - c, d, q, cell, dt, st.grid)
    scale = np.abs(Jref).max()
    err = np.abs(J - Jref).max() / scale
    print(f"  max |J - closed form| / max |J| = {err:.2e}", flush=True)
    assert err <= 1e-5


def one_particle_position(st, frames):
    """Global position (cells, FP64) of the only particle in the frames."""
    cap = len(frames["cell"]) // len(frames["counts"])
    sc = int(np.flatnonzero(frames["counts"])[0])
    g = sc * cap
    nscx, nscy = st.grid[0] // 8, st.grid[1] // 8
    lc = int(frames["cell"][g])
    cell = np.array([(sc % nscx) * 8 + lc % 8, ((sc // nscx) % nscy) * 8 + (lc // 8) % 8,
                     (sc // (nscx * nscy)) * 4 + lc // 64], np.float64)
    return cell + np.asarray(frames["pos"], np.float64)[:, g]


@pytest.mark.parametrize("pusher", [abi.PUSHER_BORIS, abi.PUSHER_VAY, abi.PUSHER_HIGUERA_CARY])
def test_single_particle_parallel_e_and_b(pusher):
    """The reference's SingleParticle setup (share/picongpu/tests/SingleParticle,
    fieldBackground.param: E = (0, 0, -10 MV/m), B = (0, 0, 50 T)): one
    electron starting with u = (0.5, 0, 0.2) gyrates about z while E
    accelerates it along z. Its weighting is 1e-6, so its own current leaves
    the uniform fields unchanged (q/m and the trajectory do not depend on it).
    After 1000 steps its position must match the FP64 solution of
    dp/dt = q (E + v x B), dx/dt = v (scipy DOP853, rtol 1e-12) within the
    FP32 bound N 2^-24 x the distance travelled (each step's momentum update
    rounds at 2^-24 of |p|); the momentum starts half a step before the
    position, as the leapfrog keeps it."""
    from scipy.integrate import solve_ivp

    st = setups.thermal(grid=(16, 16, 12), ppc=1, shape=abi.SHAPE_CIC, pusher=pusher)
    st.species[0].frames_per_sc = 1
    print(f"case single particle E || B pusher={pusher}", flush=True)
    sim = abi.Simulation(st.config())
    k = sim.constants()
    uL, uT, uM, uQ = k.unit_length, k.unit_time, k.unit_mass, k.unit_charge
    E_si, B_si = np.array([0.0, 0.0, -10.0e6]), np.array([0.0, 0.0, 50.0])
    m_si = float(k.species_mass[0]) * uM
    q_si = float(k.species_charge[0]) * uQ
    c_si = setups.SPEED_OF_LIGHT_SI
    u0 = np.array([0.5, 0.0, 0.2])
    w = 1e-6
    p_int = (u0 * m_si * w * c_si / (uM * uL / uT)).astype(np.float32)
    start = np.array([[8, 8, 6]], np.int32)
    pos0 = np.array([[0.25, 0.5, 0.75]], np.float32)
    f = setups.bin_frames(st, 0, start, pos0, p_int[None, :], np.array([w], np.float32))
    sim.write_frames(0, f)
    ones = np.ones(sim.ncells, np.float32)
    sim.load_field(abi.FIELD_E, np.stack([0 * ones, 0 * ones, ones * np.float32(E_si[2] / (uM * uL / (uT * uT * uQ)))]))
    sim.load_field(abi.FIELD_B, np.stack([0 * ones, 0 * ones, ones * np.float32(B_si[2] / (uM / (uT * uQ)))]))
    nsteps = 1000
    sim.step(nsteps)
    x_gpu = one_particle_position(st, sim.read_frames(0))
    sim.close()

    dt = float(np.float32(k.dt)) * uT
    p_start = np.asarray(p_int, np.float64) * (uM * uL / uT) / float(np.float32(w))  # at t = -dt/2, per electron

    def rhs(t, y):
        p = y[3:]
        gamma = math.sqrt(1.0 + float(p @ p) / (m_si * c_si) ** 2)
        v = p / (gamma * m_si)
        return np.concatenate([v, q_si * (E_si + np.cross(v, B_si))])

    # the fields are uniform, so p evolves on its own: bring it to t = 0 first
    p0 = solve_ivp(lambda t, p: rhs(t, np.concatenate([np.zeros(3), p]))[3:], (-dt / 2, 0.0), p_start,
                   method="DOP853", rtol=1e-12, atol=1e-40).y[:, -1]
    cell_si = np.array(st.cell_si)
    x0 = (start[0] + pos0[0].astype(np.float64)) * cell_si
    sol = solve_ivp(rhs, (0.0, nsteps * dt), np.concatenate([x0, p0]), method="DOP853", rtol=1e-12,
                    atol=np.concatenate([np.full(3, 1e-20), np.full(3, 1e-40)]))
    x_ref = sol.y[:3, -1] / cell_si
    L = np.array(st.grid, np.float64)
    diff = (x_gpu - x_ref + L / 2) % L - L / 2
    travelled = float(np.sum(np.linalg.norm(np.diff(sol.y[:3], axis=1), axis=0))) / cell_si[0]
    bound = nsteps * 2.0 ** -24 * travelled
    print(f"  travelled {travelled:.1f} cells, |x - FP64| = {np.abs(diff).max():.2e} cells (FP32 bound {bound:.1e})",
          flush=True)
    assert np.abs(diff).max() <= bound
