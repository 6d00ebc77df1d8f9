"""Pinning of the FP64 tape reference used to check the HIP adjoint.

The tape (oracle/adjoint_tape.py) restates the CIC/Boris/Yee step in torch
FP64. It is pinned here against the C++ FP64 oracle (same discrete scheme,
itself pinned against the reference's golden cases), against exact
identities the adjoint relies on (the transpose of the forward-difference
curl is the backward-difference curl; the Esirkepov current satisfies the
discrete continuity equation), and its autograd gradient against FP64
central finite differences.
"""
import numpy as np
import pytest
import torch

from oracle import adjoint_tape as T
from oracle import oracle as O
from picongpu_amd import abi, setups

F64 = torch.float64


def tape_setup(grid=(16, 16, 8), ppc=2, u=0.05):
    st = setups.thermal(grid=grid, ppc=ppc)
    st.species[0].u_th = (u, u, u)
    return st


def test_curl_transpose_identity():
    """<curl_fwd(a), b> = <a, curl_bwd(b)> on the periodic grid: the adjoint
    of the Yee E-curl is the B-curl, which the field adjoint kernels use."""
    g = T.Grid((8, 6, 4), (1.3, 0.7, 1.9), 1.0, 1.0, 1.0, 1.0)
    rng = np.random.default_rng(1)
    a = torch.tensor(rng.standard_normal((3, 8 * 6 * 4)), dtype=F64)
    b = torch.tensor(rng.standard_normal((3, 8 * 6 * 4)), dtype=F64)
    lhs = float((T.curl_fwd(g, a) * b).sum())
    rhs = float((a * T.curl_bwd(g, b)).sum())
    assert abs(lhs - rhs) <= 1e-12 * max(abs(lhs), 1.0)
    # and a mistake that transposes with the wrong sign or the same stencil fails it
    wrong = float((a * T.curl_fwd(g, b)).sum())
    assert abs(lhs - wrong) > 1e-3


def test_deposit_continuity():
    """Esirkepov current of random moves satisfies (rho1 - rho0)/dt + div J = 0
    with CIC charge density at the cell origins (charge conservation)."""
    shape = (8, 8, 8)
    cell = (1.2, 0.9, 1.5)
    g = T.Grid(shape, cell, 1.0, 1.0, 0.7, 1.0)
    rng = np.random.default_rng(2)
    n = 50
    X1 = torch.tensor(rng.random((n, 3)) * np.array(shape), dtype=F64)
    v = torch.tensor((rng.random((n, 3)) - 0.5) * 0.8, dtype=F64) * torch.tensor(cell, dtype=F64)
    J = T.deposit(g, X1, v)
    X0 = X1 - v * g.dt / torch.tensor(cell, dtype=F64)

    def rho(X):
        r = torch.zeros(g.nx * g.ny * g.nz, dtype=F64)
        base = torch.floor(X).long()
        p = X - torch.floor(X)
        for a in range(2):
            for b in range(2):
                for c in range(2):
                    w = ((1 - p[:, 0]) if a == 0 else p[:, 0]) * ((1 - p[:, 1]) if b == 0 else p[:, 1]) * \
                        ((1 - p[:, 2]) if c == 0 else p[:, 2])
                    r = r.index_add(0, g.flat(base[:, 0] + a, base[:, 1] + b, base[:, 2] + c), w)
        return r * g.q / g.vol

    s = (g.nz, g.ny, g.nx)
    div = torch.zeros(s, dtype=F64)
    for d, ax in ((0, 2), (1, 1), (2, 0)):
        j = J[d].reshape(s)
        div = div + (j - torch.roll(j, 1, dims=ax)) / cell[d]
    lhs = (rho(X1) - rho(X0)).reshape(s) / g.dt + div
    assert float(lhs.abs().max()) < 1e-12 * float(rho(X1).abs().max())


def test_tape_forward_matches_fp64_oracle():
    """Five FP64 steps of the tape equal the C++ FP64 oracle (sequential
    order) on the same particles and fields: E, B, J and particle state to
    FP64 round-off."""
    st = tape_setup()
    cells, pos, mom, w = setups.host_particles(st, 0)
    rng = np.random.default_rng(4)
    E0 = (rng.standard_normal((3, st.ncells)) * 1e-3).astype(np.float32)
    B0 = (rng.standard_normal((3, st.ncells)) * 1e-3).astype(np.float32)
    o = O.Oracle(st, 64, 1)
    o.set_frames(0, setups.bin_frames(st, 0, cells, pos, mom, w))
    o.set_field(abi.FIELD_E, E0)
    o.set_field(abi.FIELD_B, B0)
    nsteps = 5
    o.step(nsteps)
    g = T.grid_fp64(st, w[0])
    X = torch.tensor(cells.astype(np.float64) + pos.astype(np.float64))
    p = torch.tensor(mom.astype(np.float64))
    E = torch.tensor(E0.astype(np.float64))
    B = torch.tensor(B0.astype(np.float64))
    with torch.no_grad():
        for _ in range(nsteps):
            X, p, E, B = T.step(g, X, p, E, B)
    for nm, a, b in (("E", E.numpy(), o.get_field(abi.FIELD_E)), ("B", B.numpy(), o.get_field(abi.FIELD_B))):
        err = np.abs(a - b).max() / np.abs(b).max()
        assert err < 1e-9, (nm, err)
    # particles, matched through the oracle's ids (assigned in initial slot order)
    f = o.get_frames(0)
    ids = o.get_ids(0)
    cap = o.cap(0)
    valid = (np.arange(cap)[None, :] < f["counts"][:, None].astype(np.int64)).ravel()
    sc = np.repeat(np.arange(o.nsc), cap)[valid]
    lc = f["cell"][valid].astype(np.int64)
    nscx, nscy = st.grid[0] // 8, st.grid[1] // 8
    gpos = np.stack([(sc % nscx) * 8 + lc % 8, ((sc // nscx) % nscy) * 8 + (lc // 8) % 8,
                     (sc // (nscx * nscy)) * 4 + lc // 64]).astype(np.float64) + f["pos"][:, valid]
    gmom = f["mom"][:, valid]
    # input index of each initial slot: bin_frames orders stably by supercell
    sc_in = cells[:, 0] // 8 + nscx * (cells[:, 1] // 8 + nscy * (cells[:, 2] // 4))
    input_of_id = np.argsort(sc_in, kind="stable")
    src = input_of_id[ids[valid]]
    L = np.array(st.grid, np.float64)[:, None]
    dx = (gpos - X.numpy()[src].T + L / 2) % L - L / 2
    assert np.abs(dx).max() < 1e-10
    assert np.abs(gmom - p.numpy()[src].T).max() <= 1e-10 * np.abs(gmom).max()
    o.close()


def test_tape_gradient_matches_central_differences():
    """Autograd gradient of the probe energy with respect to the pulse
    parameters equals FP64 central differences (h = 1e-4)."""
    st = tape_setup(grid=(16, 16, 8), ppc=1, u=0.0)
    k = abi.constants_from_config(st.config())
    cells, pos, mom, w = setups.host_particles(st, 0)
    g = T.grid_from_constants(st.grid, k, w[0])
    X = torch.tensor(cells.astype(np.float64) + pos.astype(np.float64))
    p = torch.tensor(mom.astype(np.float64))
    lo, hi = (4, 4, 2), (12, 12, 6)
    nsteps = 4

    def objective(theta):
        E0, B0 = T.laser_fields(st.grid, g.cell, theta, amp=0.3)
        return T.run(g, X.clone(), p.clone(), E0, B0, nsteps, lo, hi)

    theta = torch.tensor([1.0, 0.5, 0.05], dtype=F64, requires_grad=True)
    f = objective(theta)
    f.backward()
    grad = theta.grad.numpy()
    h = 1e-4
    fd = np.zeros(3)
    with torch.no_grad():
        for i in range(3):
            tp = theta.detach().clone()
            tm = theta.detach().clone()
            tp[i] += h
            tm[i] -= h
            fd[i] = (float(objective(tp)) - float(objective(tm))) / (2 * h)
    err = np.abs(grad - fd).max() / np.abs(fd).max()
    assert err < 1e-4, (grad, fd)
