"""The workloads at their stated lengths and sizes, checked through properties
that do not depend on size (the oracle cannot follow them there):

- Kelvin-Helmholtz 256^3, 2 x 16 ppc, TSC, Boris, Yee, 500 steps: particle
  counts conserved, no particle lost to capacity, the Gauss residual
  (eps0 div E - rho) unchanged up to accumulated FP32 round-off, field energy
  finite.
- Laser-plasma 192x1536x192, PQS, Vay, Lehe, open x/y borders, 2000 steps:
  protons untouched, electrons only leave through the open borders, no
  particle lost to capacity (the wake compresses the plasma, supercells grow),
  the interior Gauss residual changes no more than the FP32 rounding of the
  laser field itself changes it (the same pulse in vacuum), fields finite.
- Batched ensemble (64 two-stream boxes, seed 7), 20 graph-replayed steps:
  every member conserves its particles and its Gauss residual.
- Two members of that ensemble for their 300 steps against the FP64 oracle:
  the growth rate of the magnetic field energy.
"""
import time

import numpy as np
import pytest

from oracle import oracle as O
from picongpu_amd import abi, setups

pytestmark = pytest.mark.gpu


def rho_scale(sim):
    """|rho| of one species at the typical weighting (PIC units)."""
    k = sim.constants()
    return abs(float(k.species_charge[0])) * k.typical_weighting * float(k.inv_cell_volume_dt)


def drift_bound(steps):
    """Bound on the Gauss residual's drift, relative to one species' charge
    density, that FP32 round-off allows after `steps` steps: the residual
    changes by the rounding of E += J-like terms, a few 2^-24 of |rho| per step
    at worst (linear accumulation). A lost or misplaced macro-particle changes
    it by its charge, 1/ppc of rho (1/16 here), orders of magnitude more."""
    return max(2e-6, steps * 4 * 2.0 ** -24)


def no_loss(flags):
    """Supercells may have grown; nothing else is flagged."""
    return (flags & ~abi.FLAG_CAPACITY_GREW) == 0


def test_kelvin_helmholtz_500_steps():
    st = setups.kelvin_helmholtz()
    print(f"case KH {st.grid} 2 x 16 ppc, 500 steps", flush=True)
    sim = abi.Simulation(st.config())
    setups.init_device(sim, st)
    r0 = sim.gauss_residual()
    e0 = sim.field_energy()
    for _ in range(5):
        sim.step_graph(100)
        sim.sync()
    n1 = [sim.particle_count(i) for i in range(2)]
    e1 = sim.field_energy()
    r1 = sim.gauss_residual()
    flags = sim.error_flags()
    drift = np.abs(r1 - r0).max() / rho_scale(sim)
    print(f"KH 500 steps: counts {n1}, flags {flags:#x}, field energy {e0[0]:.4e} -> {e1[0]:.4e} J, "
          f"Gauss residual drift / rho_species {drift:.3e}", flush=True)
    assert n1 == [256 ** 3 * 16] * 2
    assert no_loss(flags)
    assert np.all(np.isfinite(e1)) and e1[0] > e0[0]
    assert drift < drift_bound(500)
    sim.close()


def interior_drift(sim, st, r0, m=16):
    """Largest change of the Gauss residual, relative to one species' charge
    density, away from the open x and y borders by more than the absorber
    and the deposit's reach."""
    r1 = sim.gauss_residual().reshape(st.grid[2], st.grid[1], st.grid[0])
    return float(np.abs(r1 - r0)[:, m:-m, m:-m].max()) / rho_scale(sim)


def test_laser_plasma_2000_steps():
    """The Gauss residual's bound here is set by the laser: at a0 = 2,
    eps0 |E| (1/dx + 1/dy + 1/dz) is about 24000 rho, so each FP32 rounding
    of E moves eps0 div E by about 1.4e-3 rho, and 2000 steps of such rounding
    leave a drift of order 1e-2 rho with no particle at all. The control run
    (the same pulse, solver and borders in vacuum) measures that drift; the
    plasma run may add its own field's rounding to it, but a systematic error
    (a step's current missing or counted twice moves the residual by about
    c dt / dx = 0.5 rho in the wake) or misplaced electrons (a PQS electron
    at the box's highest density carries up to 0.04 rho to one cell) would
    show. Single-particle faithfulness is the bit-exact parity tests' job."""
    st = setups.laser_plasma()
    vac = setups.laser_plasma()
    for s in vac.species:
        s.ramp = (1, 1.0, 100e-6)  # the ramp starts 1 m away: no particles in the box
    print(f"case laser-plasma {st.grid}, 2000 steps: vacuum control", flush=True)
    sim = abi.Simulation(vac.config())
    setups.init_device(sim, vac)
    assert sim.particle_count(0) == 0 and sim.particle_count(1) == 0
    r0 = sim.gauss_residual().reshape(st.grid[2], st.grid[1], st.grid[0])
    for _ in range(4):
        sim.step_graph(500)
        sim.sync()
    drift_vacuum = interior_drift(sim, st, r0)
    assert np.all(np.isfinite(sim.field_energy()))
    sim.close()
    print(f"  vacuum: interior Gauss drift / rho {drift_vacuum:.3e}", flush=True)

    print(f"case laser-plasma {st.grid}, 2000 steps", flush=True)
    sim = abi.Simulation(st.config())
    setups.init_device(sim, st)
    n0 = [sim.particle_count(i) for i in range(2)]
    r0 = sim.gauss_residual().reshape(st.grid[2], st.grid[1], st.grid[0])
    e0 = sim.field_energy()
    t0 = time.perf_counter()
    for k in range(20):
        sim.step_graph(100)
        sim.sync()
        caps = sim.supercell_slots(0)
        print(f"  step {100 * (k + 1)}: {time.perf_counter() - t0:.1f} s, electron slots per supercell up to "
              f"{int(caps.max())}, {int(np.count_nonzero(caps > caps.min()))} supercells grown", flush=True)
    n1 = [sim.particle_count(i) for i in range(2)]
    e1 = sim.field_energy()
    flags = sim.error_flags()
    caps = sim.supercell_slots(0)
    slots = f"max {int(caps.max())}, {int(np.count_nonzero(caps > caps.min()))} of {caps.size} supercells grown"
    drift = interior_drift(sim, st, r0)
    print(f"laser-plasma 2000 steps: counts {n0} -> {n1}, electron slots per supercell {slots}, flags {flags:#x}, "
          f"field energy {e0[0]:.4e} -> {e1[0]:.4e} J, interior Gauss drift / rho {drift:.3e} "
          f"(vacuum {drift_vacuum:.3e})", flush=True)
    assert n1[1] == n0[1]  # immobile protons
    assert 0 < n1[0] <= n0[0]  # electrons leave only through the open borders
    assert no_loss(flags)
    assert np.all(np.isfinite(e1))
    assert 1e-3 < drift_vacuum < 5e-2  # the laser's rounding is what sets the scale
    assert drift < 2.0 * drift_vacuum
    sim.close()


def test_full_ensemble_20_steps():
    sts = setups.ensemble()
    print(f"case full ensemble: {len(sts)} members, 20 graph-replayed steps", flush=True)
    ens = abi.Ensemble([st.config() for st in sts])
    for m, st in zip(ens.members, sts):
        setups.init_device(m, st)
    n0 = [m.particle_count(0) for m in ens.members]
    r0 = [m.gauss_residual() for m in ens.members]
    ens.step_graph(20)
    worst = 0.0
    for i, m in enumerate(ens.members):
        assert m.particle_count(0) == n0[i], f"member {i} lost particles"
        assert no_loss(m.error_flags()), f"member {i} flags {m.error_flags():#x}"
        drift = np.abs(m.gauss_residual() - r0[i]).max() / rho_scale(m)
        worst = max(worst, drift)
        assert drift < drift_bound(20), f"member {i}: Gauss drift {drift}"
    print(f"ensemble: {sum(n0)} particles conserved, worst Gauss drift / rho {worst:.3e}", flush=True)
    ens.close()


def growth_rate(t, wb):
    """Growth rate of the magnetic field amplitude, 0.5 d ln(W_B)/dt, fitted
    over the second half of the run (t in units of 1/omega_pe)."""
    sel = t >= t[-1] / 2
    return 0.5 * np.polyfit(t[sel], np.log(wb[sel]), 1)[0]


def test_ensemble_growth_rates_against_fp64():
    """Members 62 and 53 of the 64 (40x32x60 cells, 2 and 5 ppc, beams at
    0.18 and 0.40 c), batched on the GPU for their 300 steps, against the FP64
    oracle from the same initial state: the growth rate of the magnetic energy
    (sampled every 10 steps) agrees."""
    every, steps = 10, 300
    pick = [62, 53]
    sts = [setups.ensemble()[i] for i in pick]
    print(f"case ensemble growth rates: members {pick}, {steps} steps", flush=True)
    ens = abi.Ensemble([st.config() for st in sts])
    oras = []
    for m, st in zip(ens.members, sts):
        setups.init_device(m, st)
        ora = O.Oracle(st, 64, 1)
        ora.set_frames(0, m.read_frames(0))
        oras.append(ora)
    wb_gpu = [[] for _ in sts]
    wb_f64 = [[] for _ in sts]
    for _ in range(steps // every):
        ens.step_graph(every)
        for i, (m, o) in enumerate(zip(ens.members, oras)):
            o.step(every)
            wb_gpu[i].append(m.field_energy()[1:4].sum())
            wb_f64[i].append(o.field_energy()[1:4].sum())
    for i, st in enumerate(sts):
        # electron plasma frequency of the full density (both beams), SI
        wpe = np.sqrt(st.density_si * 1.602176634e-19 ** 2 / (8.8541878188e-12 * setups.ELECTRON_MASS_SI))
        t = np.arange(1, steps // every + 1) * every * st.dt_si * wpe
        g32 = growth_rate(t, np.array(wb_gpu[i]))
        g64 = growth_rate(t, np.array(wb_f64[i]))
        err = abs(g32 - g64) / abs(g64)
        print(f"member {pick[i]} {st.grid} ppc {st.species[0].ppc} beam u_z {st.species[0].u_drift[2]:.3f} c: "
              f"growth rate FP32 GPU {g32:.6f}, FP64 oracle {g64:.6f} (1/omega_pe), relative error {err:.2e}; "
              f"W_B {wb_gpu[i][0]:.3e} -> {wb_gpu[i][-1]:.3e} J", flush=True)
        assert g64 > 0
        assert err < 1e-4
    for o in oras:
        o.close()
    ens.close()
