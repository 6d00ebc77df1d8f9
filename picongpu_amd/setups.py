"""Simulation setups in the manner of a PIConGPU setup's .param files.

Each setup fixes what param/simulation.param (DELTA_T_SI, CELL_*_SI),
param/density.param (BASE_DENSITY_SI), param/particle.param
(TYPICAL_PARTICLES_PER_CELL), param/species.param (UsedParticlePusher,
UsedParticleShape), param/speciesDefinition.param (massRatio, chargeRatio) and
param/fieldSolver.param (Solver) fix in the reference, plus the initial state.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field, replace

import numpy as np

from . import abi

SPEED_OF_LIGHT_SI = 2.99792458e8
MASS_RATIO_PROTON = 1836.152672  # param/speciesDefinition.param:75 (MassRatioIons)


@dataclass
class SpeciesSetup:
    name: str
    mass_ratio: float = 1.0
    charge_ratio: float = 1.0
    shape: int = abi.SHAPE_CIC
    pusher: int = abi.PUSHER_BORIS
    ppc: int = 8
    density_si: float = 1e25
    u_drift: tuple = (0.0, 0.0, 0.0)
    u_th: tuple = (0.0, 0.0, 0.0)
    drift_band: int = 0
    seed: int = 0
    ramp: tuple = None  # (axis, start_si, length_si) of a linear density ramp
    frames_per_sc: int = 0  # 0: derived from ppc
    current_solver: int = 0  # abi.CURRENT_SOLVER_*
    particle_boundary: tuple = (0, 0, 0)  # abi.PARTICLE_* per axis (open axes)
    quiet: tuple = None  # particles per dimension of the quiet start (None: random positions)

    def frames(self) -> int:
        if self.frames_per_sc:
            return self.frames_per_sc
        return int(math.ceil(self.ppc * 1.5)) + 2


@dataclass
class Setup:
    name: str
    grid: tuple
    cell_si: tuple
    dt_si: float
    density_si: float
    typical_ppc: int
    species: list
    solver: int = abi.SOLVER_YEE
    lehe_dir: int = 1
    steps: int = 100
    b_noise: float = 0.0
    b_noise_seed: int = 1
    laser: dict = None  # keyword arguments of Simulation.add_laser
    boundary: tuple = (0, 0, 0)  # abi.BOUNDARY_* per axis
    absorber: int = 0  # abi.ABSORBER_*
    absorber_cells: tuple = ((0, 0), (0, 0), (0, 0))  # [axis][negative, positive]
    absorber_strength: tuple = ((0.0, 0.0), (0.0, 0.0), (0.0, 0.0))
    current_interpolation: int = 0  # abi.CURRENT_*
    # incidentField.param: {(axis, side): keyword arguments of abi.fill_incident_profile}
    incident: dict = None
    incident_position: tuple = ((16, -16), (16, -16), (16, -16))
    # fieldAbsorber.param namespace pml for ABSORBER_PML (None: its defaults)
    pml: dict = None

    def config(self, device: int = 0, j_scale_exp: int = 0) -> abi.Config:
        sp = [dict(mass_ratio=s.mass_ratio, charge_ratio=s.charge_ratio, shape=s.shape, pusher=s.pusher,
                   frames_per_sc=s.frames(), current_solver=s.current_solver,
                   particle_boundary=s.particle_boundary) for s in self.species]
        return abi.make_config(self.grid, self.cell_si, self.dt_si, self.density_si, self.typical_ppc, sp,
                               solver=self.solver, lehe_dir=self.lehe_dir, j_scale_exp=j_scale_exp, device=device,
                               boundary=self.boundary, absorber=self.absorber, absorber_cells=self.absorber_cells,
                               absorber_strength=self.absorber_strength,
                               current_interpolation=self.current_interpolation, incident=self.incident,
                               incident_position=self.incident_position, pml=self.pml)

    @property
    def ncells(self) -> int:
        return int(np.prod(self.grid))

    def mobile_particles(self) -> int:
        return sum(self.ncells * s.ppc for s in self.species if s.pusher != abi.PUSHER_NONE)


def yee_dt(cell_si, courant=0.99) -> float:
    inv = sum(1.0 / (c * c) for c in cell_si)
    return courant / (SPEED_OF_LIGHT_SI * math.sqrt(inv))


def gamma_beta(beta: float) -> float:
    return beta / math.sqrt(1.0 - beta * beta)


def thermal(grid=(64, 64, 64), ppc=8, shape=abi.SHAPE_CIC, pusher=abi.PUSHER_BORIS, steps=100, seed=0) -> Setup:
    """Plain thermal plasma: electrons over an immobile neutralising
    background, 1e25 m^-3, u_th = 0.01 c, dx = 1e-7 m, dt = 0.99 Yee limit."""
    cell = (1e-7, 1e-7, 1e-7)
    e = SpeciesSetup("e", shape=shape, pusher=pusher, ppc=ppc, density_si=1e25, u_th=(0.01, 0.01, 0.01), seed=seed)
    return Setup("thermal", tuple(grid), cell, yee_dt(cell), 1e25, ppc, [e], steps=steps)


def kelvin_helmholtz(grid=(256, 256, 256), ppc=16, shape=abi.SHAPE_TSC, steps=500, seed=0) -> Setup:
    """Two-species shear flow: electrons and protons at 1e25 m^-3, the middle
    half of y drifting at +0.2 c in x and the outer halves at -0.2 c, u_th =
    0.005 c, B seeded with 1e-6 noise. Cell and time step of the reference's
    KHI test (share/picongpu/tests/KHI_growthRate/include/picongpu/param/simulation.param:52,59)."""
    cell = (8.037861e-8, 8.037861e-8, 8.037861e-8)
    dt = 1.5394e-16
    ud = (gamma_beta(0.2), 0.0, 0.0)
    uth = (0.005, 0.005, 0.005)
    e = SpeciesSetup("e", shape=shape, ppc=ppc, density_si=1e25, u_drift=ud, u_th=uth, drift_band=1, seed=seed)
    p = SpeciesSetup("p", mass_ratio=MASS_RATIO_PROTON, charge_ratio=-1.0, shape=shape, ppc=ppc, density_si=1e25,
                     u_drift=ud, u_th=uth, drift_band=1, seed=seed + 1)
    return Setup("kelvin_helmholtz", tuple(grid), cell, dt, 1e25, ppc, [e, p], steps=steps, b_noise=1e-6)


ELECTRON_MASS_SI = 9.1093837139e-31
ELECTRON_CHARGE_SI = -1.602176634e-19


def a0_amplitude_si(a0: float, wavelength_si: float) -> float:
    """E amplitude [V/m] of normalised vector potential a0, as the
    reference's UNITCONV_A0_to_Amplitude_SI (LaserWakefield incidentField.param:101-103)."""
    return a0 * (-2.0 * math.pi / wavelength_si * ELECTRON_MASS_SI * SPEED_OF_LIGHT_SI * SPEED_OF_LIGHT_SI
                 / ELECTRON_CHARGE_SI)


def khi_growth_rate(grid=(192, 512, 12), steps=3000, seed=0) -> Setup:
    """The reference's KHI growth-rate test (share/picongpu/tests/KHI_growthRate:
    -g 192 512 12, periodic, 3000 steps; param/simulation.param cell
    9.34635e-8 m * 0.86, dt 1.79e-16 s * 0.86, 1e25 m^-3; param/particle.param
    gamma 1.021 drifts along x, quiet start 5 x 5 x 1 = 25 per cell,
    electron temperature 0.5 eV; TSC, Boris, Esirkepov, Yee). Electrons and
    protons start on the same lattice (the reference derives the ions from the
    electrons); the drift sign is +1 in the middle half of y and -1 outside,
    the mirror of the reference's bands."""
    cell = (9.34635e-8 * 0.86,) * 3
    dt = 1.79e-16 * 0.86
    gamma = 1.021
    ud = (math.sqrt(gamma * gamma - 1.0), 0.0, 0.0)
    uth_e = math.sqrt(0.5 / 510998.95)  # kT / (m_e c^2) for 0.5 eV
    # the vortices of the nonlinear stage compress the plasma well beyond the
    # mean density: 80 frames (3.2 x the mean) per supercell
    e = SpeciesSetup("e", shape=abi.SHAPE_TSC, ppc=25, density_si=1e25, u_drift=ud, u_th=(uth_e,) * 3,
                     drift_band=1, seed=seed, quiet=(5, 5, 1), frames_per_sc=80)
    p = SpeciesSetup("p", mass_ratio=MASS_RATIO_PROTON, charge_ratio=-1.0, shape=abi.SHAPE_TSC, ppc=25,
                     density_si=1e25, u_drift=ud, drift_band=1, seed=seed + 1, quiet=(5, 5, 1), frames_per_sc=80)
    return Setup("khi_growth_rate", tuple(grid), cell, dt, 1e25, 25, [e, p], steps=steps)


def laser_plasma(grid=(192, 1536, 192), ppc=4, steps=2000, seed=0, laser="initial", absorber="pml") -> Setup:
    """Laser-plasma geometry: PQS shape, Vay pusher, Lehe solver Cherenkov-free
    along y, dx = dz = 0.1 um, dy = 0.05 um, cold electrons and immobile
    protons on a linear density ramp along y rising over 100 um to 3e24 m^-3
    from y = 20 um (the plateau lies beyond the 76.8 um box), and a Gaussian
    laser pulse as the initial field: 0.8 um, a0 = 2, waist 8 um, intensity
    FWHM 25 fs, polarised along x, travelling along +y, centred at y = 10 um
    on the transverse axis. Borders as the reference's LaserWakefield example
    (etc/picongpu/1.cfg: --periodic 0 0 1, no --fieldAbsorber): x and y
    open, z periodic, with the field absorber the reference then runs, the
    PML (simulation/stage/FieldAbsorber.hpp: kindName "pml") of
    param/fieldAbsorber.param (12 cells, its default grading, sigma, kappa
    and alpha). absorber="exponential" takes that file's exponential
    absorber instead (12 cells, strength 1e-3 on every side).

    laser="incident" injects the same pulse as the reference's
    LaserWakefield example does instead (its incidentField.param): a
    GaussianPulse from YMin through the Huygens surface 16 cells inside the
    borders, PULSE_DURATION_SI = FWHM / (2 sqrt(2 ln 2)), PULSE_INIT = 15,
    focused at y = 10 um on the transverse centre (FOCUS_ORIGIN Center)."""
    cell = (1e-7, 5e-8, 1e-7)
    dt = 0.995 * cell[1] / SPEED_OF_LIGHT_SI
    ramp = (1, 20e-6, 100e-6)
    e = SpeciesSetup("e", shape=abi.SHAPE_PQS, pusher=abi.PUSHER_VAY, ppc=ppc, density_si=3e24, seed=seed,
                     ramp=ramp)
    p = SpeciesSetup("p", mass_ratio=MASS_RATIO_PROTON, charge_ratio=-1.0, shape=abi.SHAPE_PQS,
                     pusher=abi.PUSHER_NONE, ppc=ppc, density_si=3e24, seed=seed + 1, ramp=ramp)
    initial = incident = None
    if laser == "initial":
        initial = dict(wavelength_si=0.8e-6, a0=2.0, waist_si=8e-6, fwhm_si=25e-15,
                       center_si=(grid[0] * cell[0] / 2, 10e-6, grid[2] * cell[2] / 2), propagation_axis=1,
                       polarisation_axis=0)
    elif laser == "incident":
        incident = {(1, 0): dict(kind=abi.INCIDENT_GAUSSIAN_PULSE, wavelength_si=0.8e-6,
                                 amplitude_si=a0_amplitude_si(2.0, 0.8e-6),
                                 pulse_duration_si=25e-15 / (2.0 * math.sqrt(2.0 * math.log(2.0))),
                                 direction=(0, 1, 0), polarisation_direction=(1, 0, 0),
                                 focus_position_si=(0.0, 10e-6, 0.0),
                                 focus_origin=(abi.ORIGIN_CENTER, abi.ORIGIN_ZERO, abi.ORIGIN_CENTER),
                                 w0_si=8e-6, pulse_init=15.0)}
    elif laser is not None:
        raise ValueError("laser must be 'initial', 'incident' or None")
    kinds = {"pml": abi.ABSORBER_PML, "exponential": abi.ABSORBER_EXPONENTIAL}
    if absorber not in kinds:
        raise ValueError("absorber must be 'pml' or 'exponential'")
    return Setup("laser_plasma", tuple(grid), cell, dt, 3e24, ppc, [e, p], solver=abi.SOLVER_LEHE, lehe_dir=1,
                 steps=steps, laser=initial, boundary=(abi.BOUNDARY_OPEN, abi.BOUNDARY_OPEN, abi.BOUNDARY_PERIODIC),
                 absorber=kinds[absorber], absorber_cells=((12, 12), (12, 12), (12, 12)),
                 absorber_strength=((1e-3, 1e-3), (1e-3, 1e-3), (1e-3, 1e-3)), incident=incident)


CONFIGS = {
    "thermal": thermal,
    "khi_growth_rate": khi_growth_rate,
    "kelvin_helmholtz": kelvin_helmholtz,
    "laser_plasma": laser_plasma,
}


def two_stream(grid=(64, 64, 64), ppc=8, beta=0.25, steps=300, seed=7) -> Setup:
    """Two counter-streaming electron beams along z (drift +-beta c, alternating
    with the particle's index in its cell), u_th = 0.01 c, 1e25 m^-3,
    dx = 1e-7 m, dt at 0.99 of the Yee limit, TSC, Boris, Yee."""
    cell = (1e-7, 1e-7, 1e-7)
    e = SpeciesSetup("e", shape=abi.SHAPE_TSC, pusher=abi.PUSHER_BORIS, ppc=ppc, density_si=1e25,
                     u_drift=(0.0, 0.0, gamma_beta(beta)), u_th=(0.01, 0.01, 0.01), drift_band=2, seed=seed)
    return Setup("two_stream", tuple(grid), cell, yee_dt(cell), 1e25, ppc, [e], steps=steps)


CONFIGS["two_stream"] = two_stream


def ensemble(n=64, lo=32, hi=96, ppc_range=(2, 16), beta_range=(0.1, 0.4), steps=300, seed=7) -> list:
    """Batched-ensemble workload: n independent two-stream boxes whose grid per
    axis is drawn from the supercell multiples in [lo, hi], ppc from
    ppc_range and beam drift from beta_range, all from one seeded generator."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        gx = int(rng.choice(np.arange(lo, hi + 1, 8)))
        gy = int(rng.choice(np.arange(lo, hi + 1, 8)))
        gz = int(rng.choice(np.arange(lo, hi + 1, 4)))
        ppc = int(rng.integers(ppc_range[0], ppc_range[1] + 1))
        beta = float(rng.uniform(*beta_range))
        st = two_stream(grid=(gx, gy, gz), ppc=ppc, beta=beta, steps=steps, seed=seed * 1000 + i)
        out.append(st)
    return out


# -- initial state --------------------------------------------------------
def differentiable(grid=(64, 64, 64), ppc=8, steps=200, seed=0) -> Setup:
    """Cold electron plasma for the differentiable form: dx = dy = dz = 1e-7 m,
    dt at 0.99 of the Yee Courant limit, 1e25 m^-3, CIC, Boris, Yee, with a
    uniform immobile neutralising background (no ion particles)."""
    cell = (1e-7, 1e-7, 1e-7)
    e = SpeciesSetup("e", shape=abi.SHAPE_CIC, pusher=abi.PUSHER_BORIS, ppc=ppc, density_si=1e25, seed=seed)
    return Setup("differentiable", tuple(grid), cell, yee_dt(cell), 1e25, ppc, [e], steps=steps)


def init_device(sim: "abi.Simulation", setup: Setup) -> None:
    """Initial particles (and B noise) generated on the device."""
    for i, s in enumerate(setup.species):
        sim.init_species(i, s.seed, s.ppc, s.density_si, s.u_drift, s.u_th, s.drift_band, ramp=s.ramp, quiet=s.quiet)
    if setup.b_noise:
        sim.add_field_noise(abi.FIELD_B, setup.b_noise_seed, setup.b_noise)
    if setup.laser:
        sim.add_laser(**setup.laser)


def host_particles(setup: Setup, sp: int, seed: int = None):
    """Flat host-side particle arrays for a species (numpy generator, for small
    tests): global cell coords, in-cell positions, momenta, weightings."""
    s = setup.species[sp]
    k = abi_free_constants(setup)
    rng = np.random.default_rng(s.seed if seed is None else seed)
    nx, ny, nz = setup.grid
    zz, yy, xx = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    cells = np.stack([xx.ravel(), yy.ravel(), zz.ravel()], axis=1).astype(np.int32)
    cells = np.repeat(cells, s.ppc, axis=0)
    n = cells.shape[0]
    pos = rng.random((n, 3), dtype=np.float32)
    vol = setup.cell_si[0] * setup.cell_si[1] * setup.cell_si[2]
    w = np.float32(s.density_si * vol / s.ppc)
    mass = np.float32(k["base_mass"] * np.float32(s.mass_ratio)) * w
    u = np.array(s.u_drift, np.float64)[None, :] * np.ones((n, 1))
    if s.drift_band == 1:
        band = (cells[:, 1] >= ny // 4) & (cells[:, 1] < 3 * ny // 4)
        u = u * np.where(band, 1.0, -1.0)[:, None]
    elif s.drift_band == 2:  # counter-streaming: sign alternates with the particle's index in its cell
        k = np.tile(np.arange(s.ppc), nx * ny * nz)
        u = u * np.where(k % 2 == 0, 1.0, -1.0)[:, None]
    u = u + rng.standard_normal((n, 3)) * np.array(s.u_th)[None, :]
    mom = (np.float64(mass) * u).astype(np.float32)
    weights = np.full(n, w, np.float32)
    return cells, pos, mom, weights


def abi_free_constants(setup: Setup) -> dict:
    """Base mass/charge in PIC units without loading the library."""
    vol = setup.cell_si[0] * setup.cell_si[1] * setup.cell_si[2]
    wtyp = setup.density_si * vol / setup.typical_ppc
    me = 9.1093837139e-31
    qe = -1.602176634e-19
    return {"base_mass": np.float32(me / (me * wtyp)), "base_charge": np.float32(qe / (-1.0 * qe * wtyp)),
            "wtyp": wtyp}


def bin_frames(setup: Setup, sp: int, cells, pos, mom, w, frames_per_sc: int = None):
    """Bin flat particles into the supercell frame layout, stable in input
    order (the layout pic_load_particles produces)."""
    nx, ny, nz = setup.grid
    nscx, nscy, nscz = nx // 8, ny // 8, nz // 4
    nsc = nscx * nscy * nscz
    cap = (frames_per_sc or setup.species[sp].frames()) * 256
    sc = cells[:, 0] // 8 + nscx * (cells[:, 1] // 8 + nscy * (cells[:, 2] // 4))
    lc = cells[:, 0] % 8 + 8 * (cells[:, 1] % 8 + 8 * (cells[:, 2] % 4))
    order = np.argsort(sc, kind="stable")
    counts = np.bincount(sc, minlength=nsc).astype(np.uint32)
    if counts.max(initial=0) > cap:
        raise ValueError("supercell capacity exceeded")
    starts = np.zeros(nsc, np.int64)
    starts[1:] = np.cumsum(counts.astype(np.int64))[:-1]
    rank = np.arange(len(order), dtype=np.int64) - starts[sc[order]]
    slot = sc[order].astype(np.int64) * cap + rank
    slots = nsc * cap
    f = {"counts": counts, "pos": np.zeros((3, slots), np.float32), "mom": np.zeros((3, slots), np.float32),
         "w": np.zeros(slots, np.float32), "cell": np.zeros(slots, np.uint16), "mask": np.zeros(slots, np.uint8)}
    f["pos"][:, slot] = pos[order].T
    f["mom"][:, slot] = mom[order].T
    f["w"][slot] = w[order]
    f["cell"][slot] = lc[order]
    f["mask"][slot] = 1
    return f


def small(setup: Setup, grid) -> Setup:
    return replace(setup, grid=tuple(grid))
