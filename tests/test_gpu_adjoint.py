"""The HIP differentiable form and its hand-written adjoint against the FP64
tape (oracle/adjoint_tape.py, pinned in test_adjoint_tape.py) and FP64
central finite differences.

Bars: forward fields and particles after the run within 1e-5 (max-norm
relative) of the FP64 tape; gradients with respect to the initial fields
and particles within 1e-4 of the tape's autograd gradients; gradient with
respect to the pulse parameters within 1e-3 of FP64 central differences
(h = 1e-4) of the tape's forward over 12 steps, and within 5e-3 (max-norm
relative) over the configuration's 200 steps with checkpoints every 20 (32^3
cells); gradients
independent of the checkpoint spacing to 1e-12.
"""
import numpy as np
import pytest
import torch

from oracle import adjoint_tape as T
from picongpu_amd import adjoint, setups

pytestmark = pytest.mark.gpu

F64 = torch.float64
GRID = (16, 16, 16)
LO, HI = (4, 4, 4), (12, 12, 12)


def rel_max(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def make(nsteps, every, ppc=2, grid=GRID):
    st = setups.differentiable(grid=grid, ppc=ppc, steps=nsteps)
    cell, pos, mom, w = adjoint.cold_plasma(st)
    pic = adjoint.DifferentiablePIC(st, cell, pos, mom, w, nsteps, every)
    k = pic.sim.constants()
    g = T.grid_from_constants(st.grid, k, w[0])
    X = torch.tensor((cell.astype(np.float64) + pos.astype(np.float64)).T.copy())
    p = torch.tensor(mom.astype(np.float64).T.copy())
    return st, pic, g, X, p


def pulse_amp(st):
    # a0 = 1 in PIC units, scaled so the probe energy changes by the plasma
    # response within a few steps without particles crossing many cells
    return 0.3 * adjoint.a0_field_unit(st)


def test_forward_matches_fp64_tape():
    nsteps = 12
    print(f"case adjoint forward grid={GRID} steps={nsteps}", flush=True)
    st, pic, g, X, p = make(nsteps, 5)
    theta = torch.tensor([1.0, 0.5, 0.05], dtype=F64)
    E0, B0 = adjoint.gaussian_pulse(st, theta.cuda(), amp=pulse_amp(st))
    E, B = pic.forward_raw(E0, B0)
    cell, pos, mom = pic.sim.adjoint_read_particles()
    e0 = torch.tensor(E0.float().cpu().numpy().astype(np.float64))
    b0 = torch.tensor(B0.float().cpu().numpy().astype(np.float64))
    with torch.no_grad():
        Xt, pt, Et, Bt = X.clone(), p.clone(), e0, b0
        for _ in range(nsteps):
            Xt, pt, Et, Bt = T.step(g, Xt, pt, Et, Bt)
    eE = rel_max(E.cpu().numpy(), Et.numpy())
    eB = rel_max(B.cpu().numpy(), Bt.numpy())
    Xg = cell.astype(np.float64) + pos.astype(np.float64)
    L = np.array(st.grid, np.float64)[:, None]
    dX = float(np.abs((Xg - Xt.numpy().T + L / 2) % L - L / 2).max())
    eP = rel_max(mom, pt.numpy().T)
    print("forward vs FP64 tape: E", eE, "B", eB, "X (cells)", dX, "p", eP)
    assert eE < 1e-5 and eB < 1e-5 and dX < 1e-5 and eP < 1e-5
    pic.close()


def test_adjoint_matches_fp64_tape():
    nsteps, every = 12, 5
    print(f"case adjoint gradients grid={GRID} steps={nsteps} every={every}", flush=True)
    st, pic, g, X, p = make(nsteps, every)
    theta = torch.tensor([1.0, 0.5, 0.05], dtype=F64)
    E0, B0 = adjoint.gaussian_pulse(st, theta.cuda(), amp=pulse_amp(st))
    E, B = pic.forward_raw(E0, B0)
    Ed = E.double().requires_grad_(True)
    Bd = B.double().requires_grad_(True)
    adjoint.probe_energy(st, Ed, Bd, LO, HI).backward()
    dE0, dB0, dX0, dP0 = pic.backward_raw(Ed.grad, Bd.grad, particles=True)
    e0 = torch.tensor(E0.float().cpu().numpy().astype(np.float64))
    b0 = torch.tensor(B0.float().cpu().numpy().astype(np.float64))
    f, tE, tB, tX, tP = T.gradient_all(g, X, p, e0, b0, nsteps, LO, HI)
    err = {"dE0": rel_max(dE0.cpu().numpy(), tE.numpy()), "dB0": rel_max(dB0.cpu().numpy(), tB.numpy()),
           "dX0": rel_max(dX0.cpu().numpy(), tX.numpy().T), "dP0": rel_max(dP0.cpu().numpy(), tP.numpy().T)}
    print("adjoint vs FP64 tape:", err, "|dX0|max", float(tX.abs().max()), "|dP0|max", float(tP.abs().max()))
    for nm, e in err.items():
        assert e < 1e-4, (nm, e)
    pic.close()


def test_parameter_gradient_matches_central_differences():
    """dF/dtheta through the torch.autograd.Function (pulse built in torch on
    the device, HIP forward and adjoint) against FP64 central differences of
    the tape's forward, h = 1e-4."""
    nsteps, every = 12, 5
    print(f"case parameter gradient grid={GRID} steps={nsteps}", flush=True)
    st, pic, g, X, p = make(nsteps, every)
    amp = pulse_amp(st)
    theta = torch.tensor([1.0, 0.5, 0.05], dtype=F64, device="cuda", requires_grad=True)
    E0, B0 = adjoint.gaussian_pulse(st, theta, amp=amp)
    E, B = pic.fields(E0, B0)
    F = adjoint.probe_energy(st, E, B, LO, HI)
    F.backward()
    grad = theta.grad.cpu().numpy()

    def objective(th):
        e0, b0 = T.laser_fields(st.grid, g.cell, th, amp=amp)
        return float(T.run(g, X.clone(), p.clone(), e0, b0, nsteps, LO, HI))

    h = 1e-4
    fd = np.zeros(3)
    with torch.no_grad():
        base = torch.tensor([1.0, 0.5, 0.05], dtype=F64)
        for i in range(3):
            tp, tm = base.clone(), base.clone()
            tp[i] += h
            tm[i] -= h
            fd[i] = (objective(tp) - objective(tm)) / (2 * h)
    err = np.abs(grad - fd) / np.abs(fd)
    print("dF/dtheta HIP adjoint:", grad, "FP64 central differences:", fd, "rel err:", err, "F:", float(F.detach()))
    assert np.all(err < 1e-3)
    pic.close()


def test_parameter_gradient_200_steps_matches_central_differences():
    """The configuration's length: 200 steps with a checkpoint every 20 on
    32^3 cells x 2 ppc. dF/dtheta through the torch.autograd.Function
    against FP64 central differences (h = 1e-4) of the tape's forward, which
    runs on the device in FP64 here (on the CPU it would take minutes)."""
    nsteps, every, grid = 200, 20, (32, 32, 32)
    lo, hi = (8, 8, 8), (24, 24, 24)
    print(f"case parameter gradient grid={grid} steps={nsteps} every={every}", flush=True)
    st, pic, g, X, p = make(nsteps, every, grid=grid)
    amp = pulse_amp(st)
    theta = torch.tensor([1.0, 0.5, 0.05], dtype=F64, device="cuda", requires_grad=True)
    E0, B0 = adjoint.gaussian_pulse(st, theta, amp=amp)
    E, B = pic.fields(E0, B0)
    F = adjoint.probe_energy(st, E, B, lo, hi)
    F.backward()
    grad = theta.grad.cpu().numpy()
    Xd, pd = X.cuda(), p.cuda()

    def objective(th):
        e0, b0 = T.laser_fields(st.grid, g.cell, th, amp=amp)
        return float(T.run(g, Xd.clone(), pd.clone(), e0.cuda(), b0.cuda(), nsteps, lo, hi))

    h = 1e-4
    fd = np.zeros(3)
    with torch.no_grad():
        base = torch.tensor([1.0, 0.5, 0.05], dtype=F64)
        f64 = objective(base)
        for i in range(3):
            tp, tm = base.clone(), base.clone()
            tp[i] += h
            tm[i] -= h
            fd[i] = (objective(tp) - objective(tm)) / (2 * h)
            print(f"  parameter {i}: FP64 central difference {fd[i]:.9e}", flush=True)
    err = np.abs(grad - fd) / np.abs(fd)
    err_max = float(np.abs(grad - fd).max() / np.abs(fd).max())
    print(f"200 steps: dF/dtheta HIP adjoint {grad}, FP64 central differences {fd}, relative error per "
          f"parameter {err}, max-norm relative {err_max:.3e}; F HIP {float(F.detach()):.9e}, FP64 {f64:.9e}",
          flush=True)
    # over 200 steps the FP32 forward's trajectories drift from the FP64 ones
    # (F agrees to 3e-8): measured 1.7e-3 max-norm relative, per parameter
    # 3.2e-4, 3.2e-3 and 7.8e-3 (the second derivative is 130x smaller than
    # the first)
    assert err_max < 5e-3
    pic.close()


def test_checkpoint_spacing_does_not_change_gradient():
    nsteps = 9
    print(f"case checkpoint spacing steps={nsteps}", flush=True)
    out = []
    for every in (1, 4, nsteps):
        st, pic, g, X, p = make(nsteps, every, ppc=1)
        theta = torch.tensor([1.0, 0.5, 0.05], dtype=F64)
        E0, B0 = adjoint.gaussian_pulse(st, theta.cuda(), amp=pulse_amp(st))
        E, B = pic.forward_raw(E0, B0)
        Ed = E.double().requires_grad_(True)
        Bd = B.double().requires_grad_(True)
        adjoint.probe_energy(st, Ed, Bd, LO, HI).backward()
        dE0, dB0 = pic.backward_raw(Ed.grad, Bd.grad)
        out.append((dE0.cpu().numpy(), dB0.cpu().numpy()))
        pic.close()
    for a, b in out[1:]:
        assert rel_max(a, out[0][0]) < 1e-12 and rel_max(b, out[0][1]) < 1e-12


def test_backward_of_an_older_forward_is_refused():
    """The context keeps the checkpoints of its last forward only: a backward
    through the autograd function after another forward must raise instead of
    returning the gradients of the newer forward."""
    nsteps = 4
    print("case adjoint generation check", flush=True)
    st, pic, g, X, p = make(nsteps, 2)
    theta = torch.tensor([1.0, 0.0, 0.0], dtype=F64)
    E0, B0 = adjoint.gaussian_pulse(st, theta.cuda(), amp=pulse_amp(st))
    E0 = E0.clone().requires_grad_(True)
    E, B = pic.fields(E0, B0)
    obj = adjoint.probe_energy(st, E, B, LO, HI)
    pic.forward_raw(E0.detach() * 0.5, B0)  # a second forward on the same context
    with pytest.raises(RuntimeError, match="another forward"):
        obj.backward()
    E, B = pic.fields(E0, B0)  # forward again: the backward of the newest forward works
    adjoint.probe_energy(st, E, B, LO, HI).backward()
    assert E0.grad is not None and torch.isfinite(E0.grad).all()
    pic.close()
