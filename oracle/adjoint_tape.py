"""FP64 tape-based adjoint of the differentiable PIC scheme (PyTorch autograd,
on the device of the tensors it is given).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the reference gradient of the
HIP adjoint. Never imported by the picongpu_amd package.

The discrete scheme is the reference's PIC step for one mobile species with
the CIC shape, the Boris pusher and the Yee solver on a periodic box
(Simulation.hpp:522-542 stage order): CIC field gather at the Yee positions
(FieldToParticleInterpolation.hpp:100-124, ShiftCoordinateSystem.hpp:53-79),
Boris push (particlePusherBoris.hpp:41-90), position update, Esirkepov
deposit from x1 - v*dt/dx to x1 (Esirkepov.hpp:61-242), B half step, E full
step, E += -(dt/eps0) J, B half step (FDTDBase.kernel:50-124). Particles keep
their identity (global positions X = cell + in-cell position); cell indices
are floor(X) and carry no derivative. Everything is FP64; autograd records
the tape.
"""
from __future__ import annotations

import math

import torch

F64 = torch.float64


def cic(x):
    """CIC assignment function 1-|x| on |x|<1 (CIC.hpp:89-107)."""
    return torch.clamp(1.0 - torch.abs(x), min=0.0)


class Grid:
    def __init__(self, shape, cell, dt, eps0, charge, mass):
        self.nx, self.ny, self.nz = shape
        self.cell = [float(c) for c in cell]
        self.dt = float(dt)
        self.eps0 = float(eps0)
        self.q = float(charge)  # macro-particle charge
        self.m = float(mass)    # macro-particle mass
        self.vol = self.cell[0] * self.cell[1] * self.cell[2]

    def flat(self, ix, iy, iz):
        return (ix % self.nx) + self.nx * ((iy % self.ny) + self.ny * (iz % self.nz))


# Yee positions of E and B components in the cell (YeeCell.hpp:69-123)
POS_E = [(0.5, 0.0, 0.0), (0.0, 0.5, 0.0), (0.0, 0.0, 0.5)]
POS_B = [(0.0, 0.5, 0.5), (0.5, 0.0, 0.5), (0.5, 0.5, 0.0)]


def gather(g: Grid, F, X, field_pos):
    """CIC interpolation of the three components of F (3, ncells) at global
    positions X (n, 3), with the reference's coordinate shift per component
    (field_pos: POS_E or POS_B)."""
    cell = torch.floor(X).detach()
    pos = X - cell
    out = []
    for c in range(3):
        fpos = field_pos[c]
        idx_base = []
        w = []
        for d in range(3):
            v = pos[:, d] - fpos[d] - 0.5
            s = torch.where(v >= -0.5, torch.zeros_like(v), -torch.ones_like(v)).detach()
            p = v - s + 0.5
            w.append((1.0 - p, p))
            idx_base.append(cell[:, d].long() + s.long())
        acc = torch.zeros(X.shape[0], dtype=F64, device=X.device)
        for a in range(2):
            for b in range(2):
                for e in range(2):
                    ii = g.flat(idx_base[0] + a, idx_base[1] + b, idx_base[2] + e)
                    acc = acc + F[c][ii] * w[0][a] * w[1][b] * w[2][e]
        out.append(acc)
    return torch.stack(out, dim=1)


def boris(g: Grid, p, E, B):
    qm = g.q / g.m
    pm = p + 0.5 * g.q * E * g.dt
    gam = torch.sqrt(1.0 + (pm * pm).sum(dim=1, keepdim=True) / (g.m * g.m))
    t = 0.5 * qm * B / gam * g.dt
    s = 2.0 * t / (1.0 + (t * t).sum(dim=1, keepdim=True))
    pp = pm + torch.cross(pm, t, dim=1)
    pl = pm + torch.cross(pp, s, dim=1)
    return pl + 0.5 * g.q * E * g.dt


def velocity(g: Grid, p):
    return p / torch.sqrt(g.m * g.m + (p * p).sum(dim=1, keepdim=True))


def deposit(g: Grid, X1, v):
    """Esirkepov CIC current of each particle moving from X1 - v dt/dx to X1."""
    cellv = torch.tensor(g.cell, dtype=F64, device=X1.device)
    X0 = X1 - v * g.dt / cellv
    gs = torch.minimum(torch.floor(X0), torch.floor(X1)).detach()
    p0 = X0 - gs
    p1 = X1 - gs
    n = X1.shape[0]
    nodes = torch.arange(3, dtype=F64, device=X1.device)
    S0 = [cic(nodes[None, :] - p0[:, d:d + 1]) for d in range(3)]  # (n,3) weights at nodes gs+0..2
    S1 = [cic(nodes[None, :] - p1[:, d:d + 1]) for d in range(3)]
    DS = [S1[d] - S0[d] for d in range(3)]
    J = torch.zeros(3, g.nx * g.ny * g.nz, dtype=F64, device=X1.device)
    base = gs.long()
    for comp in range(3):
        a, b = [x for x in range(3) if x != comp]
        csd = g.q / (g.vol * g.dt) * g.cell[comp]
        Wab = (S0[a][:, :, None] * S0[b][:, None, :] + 0.5 * DS[a][:, :, None] * S0[b][:, None, :]
               + 0.5 * S0[a][:, :, None] * DS[b][:, None, :] + DS[a][:, :, None] * DS[b][:, None, :] / 3.0)
        cum = torch.cumsum(DS[comp], dim=1)  # prefix over the current's axis
        contrib = -csd * Wab[:, :, :, None] * cum[:, None, None, :]  # (n, 3a, 3b, 3comp)
        for ia in range(3):
            for ib in range(3):
                for ic in range(2):  # J at the last node of the axis is zero
                    off = [0, 0, 0]
                    off[a], off[b], off[comp] = ia, ib, ic
                    ii = g.flat(base[:, 0] + off[0], base[:, 1] + off[1], base[:, 2] + off[2])
                    J[comp] = J[comp].index_add(0, ii, contrib[:, ia, ib, ic])
    return J


def curl_fwd(g: Grid, F):
    """Curl with forward differences (Yee curl of E)."""
    s = (g.nz, g.ny, g.nx)
    f = [F[c].reshape(s) for c in range(3)]

    def d(x, axis):
        ax = {0: 2, 1: 1, 2: 0}[axis]
        return (torch.roll(x, -1, dims=ax) - x) / g.cell[axis]

    cx = d(f[2], 1) - d(f[1], 2)
    cy = d(f[0], 2) - d(f[2], 0)
    cz = d(f[1], 0) - d(f[0], 1)
    return torch.stack([cx.reshape(-1), cy.reshape(-1), cz.reshape(-1)])


def curl_bwd(g: Grid, F):
    """Curl with backward differences (Yee curl of B)."""
    s = (g.nz, g.ny, g.nx)
    f = [F[c].reshape(s) for c in range(3)]

    def d(x, axis):
        ax = {0: 2, 1: 1, 2: 0}[axis]
        return (x - torch.roll(x, 1, dims=ax)) / g.cell[axis]

    cx = d(f[2], 1) - d(f[1], 2)
    cy = d(f[0], 2) - d(f[2], 0)
    cz = d(f[1], 0) - d(f[0], 1)
    return torch.stack([cx.reshape(-1), cy.reshape(-1), cz.reshape(-1)])


def step(g: Grid, X, p, E, B):
    Ep = gather(g, E, X, POS_E)
    Bp = gather(g, B, X, POS_B)
    p1 = boris(g, p, Ep, Bp)
    v = velocity(g, p1)
    cellv = torch.tensor(g.cell, dtype=F64, device=X.device)
    X1 = X + v * g.dt / cellv
    J = deposit(g, X1, v)
    B = B - 0.5 * g.dt * curl_fwd(g, E)
    E = E + g.dt * curl_bwd(g, B) - (g.dt / g.eps0) * J
    B = B - 0.5 * g.dt * curl_fwd(g, E)
    return X1, p1, E, B


def probe_energy(g: Grid, E, B, lo, hi):
    s = (g.nz, g.ny, g.nx)
    tot = 0.0
    for F in (E, B):
        for c in range(3):
            f = F[c].reshape(s)[lo[2]:hi[2], lo[1]:hi[1], lo[0]:hi[0]]
            tot = tot + 0.5 * (f * f).sum()
    return tot


def run(g: Grid, X, p, E, B, nsteps, lo, hi):
    for _ in range(nsteps):
        X, p, E, B = step(g, X, p, E, B)
    return probe_energy(g, E, B, lo, hi)


def gradient_all(g: Grid, X, p, E0, B0, nsteps, lo, hi):
    """Objective and its gradient with respect to the initial E, B and the
    initial particle positions X (cells) and momenta p."""
    E0 = E0.clone().detach().requires_grad_(True)
    B0 = B0.clone().detach().requires_grad_(True)
    X = X.clone().detach().requires_grad_(True)
    p = p.clone().detach().requires_grad_(True)
    f = run(g, X, p, E0, B0, nsteps, lo, hi)
    f.backward()
    return float(f.detach()), E0.grad.detach(), B0.grad.detach(), X.grad.detach(), p.grad.detach()


def grid_from_constants(shape, k, w):
    """Tape grid with the library's FP32 constants (pic_constants) and the
    macro-particle weighting w: charge and mass are the FP32 products the
    kernels use (frame charge * w, frame mass * w)."""
    import numpy as np
    q = float(np.float32(np.float32(k.species_charge[0]) * np.float32(w)))
    m = float(np.float32(np.float32(k.species_mass[0]) * np.float32(w)))
    return Grid(shape, [float(np.float32(c)) for c in k.cell], float(np.float32(k.dt)), float(np.float32(k.eps0)),
                q, m)


def grid_fp64(setup, w, species=0):
    """Tape grid with the unitless constants evaluated in FP64 throughout, as
    the FP64 twin derives them (simulation.unitless:31-500 without the FP32
    casts; speciesDefinition.unitless:45-66)."""
    c_si, mue0_si, me, qe = 2.99792458e8, 1.25663706127e-6, 9.1093837139e-31, -1.602176634e-19
    unit_time = setup.dt_si
    unit_length = unit_time * c_si
    vol = setup.cell_si[0] * setup.cell_si[1] * setup.cell_si[2]
    wtyp = setup.density_si * vol / setup.typical_ppc
    unit_mass = me * wtyp
    unit_charge = -1.0 * qe * wtyp
    mue0 = mue0_si / unit_length / unit_mass * unit_charge * unit_charge
    eps0 = 1.0 / mue0
    s = setup.species[species]
    mass = me / unit_mass * s.mass_ratio
    charge = qe / unit_charge * s.charge_ratio
    cell = [c / unit_length for c in setup.cell_si]
    return Grid(setup.grid, cell, setup.dt_si / unit_time, eps0, charge * float(w), mass * float(w))


def gradient(g: Grid, X, p, E0, B0, nsteps, lo, hi):
    """Objective and its gradient with respect to the initial E and B."""
    E0 = E0.clone().detach().requires_grad_(True)
    B0 = B0.clone().detach().requires_grad_(True)
    f = run(g, X.clone(), p.clone(), E0, B0, nsteps, lo, hi)
    f.backward()
    return float(f), E0.grad.detach(), B0.grad.detach()


def laser_fields(shape, cell, theta, lam_cells=8.0, w0_cells=4.0, tau_cells=6.0, amp=1.0):
    """Gaussian pulse along +x as an initial field, differentiable in
    theta = (a0, focus offset in cells, linear chirp): E_y and B_z = E_y
    (PIC units, c = 1), sampled at their Yee positions."""
    a0, focus, chirp = theta[0], theta[1], theta[2]
    nx, ny, nz = shape
    xs = torch.arange(nx, dtype=F64)
    ys = torch.arange(ny, dtype=F64)
    zs = torch.arange(nz, dtype=F64)
    xc, yc, zc = nx / 2.0, ny / 2.0, nz / 2.0
    k = 2 * math.pi / lam_cells
    zr = math.pi * w0_cells ** 2 / lam_cells

    def pulse(x, y, z):
        xi = x - xc
        zf = (xi - focus) / zr
        w = w0_cells * torch.sqrt(1.0 + zf * zf)
        r2 = (y - yc) ** 2 + (z - zc) ** 2
        env = torch.exp(-(xi / tau_cells) ** 2) * (w0_cells / w) * torch.exp(-r2 / (w * w))
        phase = k * xi + chirp * k * xi * xi / tau_cells - torch.atan(zf)
        return amp * a0 * env * torch.cos(phase)

    Z, Y, Xg = torch.meshgrid(zs, ys, xs, indexing="ij")
    n = nx * ny * nz
    E = torch.zeros(3, n, dtype=F64)
    B = torch.zeros(3, n, dtype=F64)
    # E_y at (i, j+1/2, k), B_z at (i+1/2, j+1/2, k)
    E1 = pulse(Xg, Y + 0.5, Z).reshape(-1)
    B2 = pulse(Xg + 0.5, Y + 0.5, Z).reshape(-1)
    E = torch.stack([torch.zeros(n, dtype=F64), E1, torch.zeros(n, dtype=F64)])
    B = torch.stack([torch.zeros(n, dtype=F64), torch.zeros(n, dtype=F64), B2])
    return E, B
