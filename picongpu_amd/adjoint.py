"""torch.autograd harness around the differentiable form of the PIC step.

The forward runs the HIP kernels of the differentiable form (one CIC/Boris
species kept in identity order, the reference's stage order
Simulation.hpp:522-542, Yee solver) from initial fields E0, B0 for a fixed
number of steps, storing a checkpoint every ``checkpoint_every`` steps. The
backward runs the hand-written FP64 adjoint kernels: it recomputes each
checkpoint segment forward and reverses it step by step (transposed curl
stencils, the Esirkepov transpose as a gather of the current's adjoint,
reverse Boris, reverse field gather).

PyTorch is used for device memory only: every kernel runs inside the HIP
library through its C ABI (pic_adjoint_forward / pic_adjoint_backward).
Fields are (3, nx*ny*nz) tensors on the library's device, x fastest.
"""
from __future__ import annotations

import math

import numpy as np
import torch

from . import abi, setups


class DifferentiablePIC:
    """A simulation context in differentiable mode with its initial particles."""

    def __init__(self, setup: setups.Setup, cell, pos, mom, w, nsteps: int, checkpoint_every: int = 20,
                 species: int = 0):
        if not torch.cuda.is_available():
            raise RuntimeError("the differentiable form needs a HIP device")
        self.setup = setup
        self.sim = abi.Simulation(setup.config())
        self.sim.adjoint_load_particles(species, cell, pos, mom, w)
        self.nsteps = int(nsteps)
        self.every = int(checkpoint_every)
        self.ncells = setup.ncells
        self.device = torch.device("cuda", 0)
        # the context keeps the checkpoints of its last forward only: every
        # forward gets a new generation, and a backward of an older one is refused
        self.generation = 0

    def close(self):
        self.sim.close()

    def fields(self, E0: torch.Tensor, B0: torch.Tensor):
        """Final (E, B) after nsteps, differentiable with respect to E0 and B0."""
        return _PICAdjoint.apply(E0, B0, self)

    def forward_raw(self, E0: torch.Tensor, B0: torch.Tensor):
        """Forward only (no autograd record); returns float32 (E, B)."""
        e0 = E0.detach().to(self.device, torch.float32).contiguous()
        b0 = B0.detach().to(self.device, torch.float32).contiguous()
        E = torch.empty_like(e0)
        B = torch.empty_like(b0)
        torch.cuda.synchronize(self.device)
        self.generation += 1
        self.sim.adjoint_forward(e0.data_ptr(), b0.data_ptr(), self.nsteps, self.every, E.data_ptr(), B.data_ptr())
        return E, B

    def backward_raw(self, gE: torch.Tensor, gB: torch.Tensor, particles: bool = False):
        """Adjoint of the last forward: (dE0, dB0[, dX0, dP0]) in float64."""
        ge = gE.detach().to(self.device, torch.float64).contiguous()
        gb = gB.detach().to(self.device, torch.float64).contiguous()
        dE0 = torch.empty_like(ge)
        dB0 = torch.empty_like(gb)
        dX0 = dP0 = None
        if particles:
            n = self.sim.adj_n
            dX0 = torch.empty((3, n), dtype=torch.float64, device=self.device)
            dP0 = torch.empty((3, n), dtype=torch.float64, device=self.device)
        torch.cuda.synchronize(self.device)
        self.sim.adjoint_backward(ge.data_ptr(), gb.data_ptr(), dE0.data_ptr(), dB0.data_ptr(),
                                  dX0.data_ptr() if particles else None, dP0.data_ptr() if particles else None)
        if particles:
            return dE0, dB0, dX0, dP0
        return dE0, dB0


class _PICAdjoint(torch.autograd.Function):
    @staticmethod
    def forward(ctx, E0, B0, pic: DifferentiablePIC):
        ctx.pic = pic
        ctx.dtypes = (E0.dtype, B0.dtype)
        E, B = pic.forward_raw(E0, B0)
        ctx.generation = pic.generation
        return E.to(torch.float64), B.to(torch.float64)

    @staticmethod
    def backward(ctx, gE, gB):
        pic = ctx.pic
        if ctx.generation != pic.generation:
            raise RuntimeError("the differentiable context ran another forward after this one; its checkpoints "
                               "belong to that forward, so this backward would return wrong gradients: "
                               "call fields() again before backward()")
        if gE is None:
            gE = torch.zeros((3, pic.ncells), dtype=torch.float64, device=pic.device)
        if gB is None:
            gB = torch.zeros((3, pic.ncells), dtype=torch.float64, device=pic.device)
        dE0, dB0 = pic.backward_raw(gE, gB)
        return dE0.to(ctx.dtypes[0]), dB0.to(ctx.dtypes[1]), None


# ---------------------------------------------------------------------------
# objective and initial fields of the differentiable configuration
# ---------------------------------------------------------------------------
def probe_energy(setup: setups.Setup, E: torch.Tensor, B: torch.Tensor, lo, hi) -> torch.Tensor:
    """Electromagnetic energy density sum 0.5*(E^2 + B^2) over the probe box
    [lo, hi) in cells (PIC units, c = 1)."""
    nx, ny, nz = setup.grid
    tot = 0.0
    for F in (E, B):
        f = F.reshape(3, nz, ny, nx)[:, lo[2]:hi[2], lo[1]:hi[1], lo[0]:hi[0]]
        tot = tot + 0.5 * (f * f).sum()
    return tot


def gaussian_pulse(setup: setups.Setup, theta: torch.Tensor, lam_cells=8.0, w0_cells=4.0, tau_cells=6.0,
                   amp=1.0):
    """Gaussian pulse travelling along +x as an initial field, differentiable in
    theta = (a0, focus offset in cells, linear chirp): E_y and B_z = E_y
    (c = 1) sampled at their Yee positions (YeeCell.hpp:69-123)."""
    nx, ny, nz = setup.grid
    dev = theta.device
    xs = torch.arange(nx, dtype=torch.float64, device=dev)
    ys = torch.arange(ny, dtype=torch.float64, device=dev)
    zs = torch.arange(nz, dtype=torch.float64, device=dev)
    xc, yc, zc = nx / 2.0, ny / 2.0, nz / 2.0
    k = 2 * math.pi / lam_cells
    zr = math.pi * w0_cells ** 2 / lam_cells
    a0, focus, chirp = theta[0], theta[1], theta[2]

    def pulse(x, y, z):
        xi = x - xc
        zf = (xi - focus) / zr
        w = w0_cells * torch.sqrt(1.0 + zf * zf)
        r2 = (y - yc) ** 2 + (z - zc) ** 2
        env = torch.exp(-(xi / tau_cells) ** 2) * (w0_cells / w) * torch.exp(-r2 / (w * w))
        phase = k * xi + chirp * k * xi * xi / tau_cells - torch.atan(zf)
        return amp * a0 * env * torch.cos(phase)

    Z, Y, X = torch.meshgrid(zs, ys, xs, indexing="ij")
    n = nx * ny * nz
    zero = torch.zeros(n, dtype=torch.float64, device=dev)
    E = torch.stack([zero, pulse(X, Y + 0.5, Z).reshape(-1), zero])
    B = torch.stack([zero, zero, pulse(X + 0.5, Y + 0.5, Z).reshape(-1)])
    return E, B


def cold_plasma(setup: setups.Setup, sp: int = 0, seed: int = None, tile=(8, 4, 1)):
    """Uniformly placed cold particles: ppc per cell at seeded uniform in-cell
    positions, zero momentum. Ordered tile-major (tiles of 8x4x1 cells), then
    by slot (the first particle of every cell of the tile, then the second,
    ...), then by cell within the tile (x fastest), so the particles a thread
    block handles sit in a compact box of cells and their current and field
    adjoint updates accumulate in the block's LDS window. Returns
    (cell (3,n) int32, pos (3,n) float32, mom (3,n) float32, w (n,) float32)."""
    s = setup.species[sp]
    nx, ny, nz = setup.grid
    tx, ty, tz = tile
    if nx % tx or ny % ty or nz % tz:
        raise ValueError("grid must be a multiple of the tile")
    rng = np.random.default_rng(seed if seed is not None else s.seed)
    ntx, nty, ntz = nx // tx, ny // ty, nz // tz
    t = np.arange(ntx * nty * ntz)
    loc = np.arange(tx * ty * tz)
    shape = (t.size, s.ppc, loc.size)
    gx = np.broadcast_to(((t % ntx) * tx)[:, None, None] + (loc % tx)[None, None, :], shape)
    gy = np.broadcast_to((((t // ntx) % nty) * ty)[:, None, None] + ((loc // tx) % ty)[None, None, :], shape)
    gz = np.broadcast_to(((t // (ntx * nty)) * tz)[:, None, None] + (loc // (tx * ty))[None, None, :], shape)
    cell = np.stack([gx.ravel(), gy.ravel(), gz.ravel()]).astype(np.int32)
    n = cell.shape[1]
    pos = rng.random((3, n), dtype=np.float32)
    mom = np.zeros((3, n), np.float32)
    vol = setup.cell_si[0] * setup.cell_si[1] * setup.cell_si[2]
    w = np.full(n, np.float32(s.density_si * vol / s.ppc), np.float32)
    return cell, pos, mom, w


def a0_field_unit(setup: setups.Setup, wavelength_si: float = 0.8e-6) -> float:
    """PIC-unit field amplitude of a0 = 1: m_e c omega / e in V/m divided by the
    unit of the electric field (UNIT_MASS UNIT_LENGTH / (UNIT_TIME^2 UNIT_CHARGE))."""
    k = abi.constants_from_config(setup.config())
    m_e, e = setups.ELECTRON_MASS_SI, 1.602176634e-19
    omega = 2 * math.pi * setups.SPEED_OF_LIGHT_SI / wavelength_si
    e_si = m_e * setups.SPEED_OF_LIGHT_SI * omega / e
    unit_e = k.unit_mass * k.unit_length / (k.unit_time ** 2 * k.unit_charge)
    return e_si / unit_e
