"""The oracle's PML weight damping of an absorbing particle boundary
(boundary/Absorbing.hpp:95-200 dampWeightsInPml, after Lehe et al. 2022),
pinned to closed forms and to numerical quadrature.

After the push, each exchange of an open axis on which a species is
absorbing scales the weighting and the momentum of the particles in the
supercells its PML layer reaches by exp(-I), I the integral of the
normalised conductivity sigma / eps0 = SIGMA_MAX / eps0 * depth^m over the
last step's straight path (Pml.kernel:326-386), the weighting kept at
MIN_WEIGHTING = 10 or above (particle.param, DampedWeighting.hpp:67-90).
The sigma maximum here is restated from fieldAbsorber.param (SIGMA_OPT_SI =
0.8 (m + 1) / (Z0 dx)) and fieldAbsorber.unitless (/ eps0 * unit time)."""
import math

import numpy as np
import pytest
from scipy import integrate

from oracle import oracle as O
from picongpu_amd import abi, setups

MU0_SI = 1.25663706127e-6  # physicalConstants.param MUE0_SI
NEG, POS = 8, 12           # layer cells at the -x and +x borders


def pml_box(grid=(32, 8, 4), reflecting=False, pml=None):
    st = setups.thermal(grid=grid, ppc=1, pusher=abi.PUSHER_BORIS)
    st.boundary = (abi.BOUNDARY_OPEN, abi.BOUNDARY_PERIODIC, abi.BOUNDARY_PERIODIC)
    st.absorber = abi.ABSORBER_PML
    st.absorber_cells = ((NEG, POS), (0, 0), (0, 0))
    st.pml = pml
    if reflecting:
        st.species[0].particle_boundary = (abi.PARTICLE_REFLECTING, abi.PARTICLE_ABSORBING, abi.PARTICLE_ABSORBING)
    return st


def sigma_max_normalised(st, k, grading=4.0, sigma_max_si=None):
    c = setups.SPEED_OF_LIGHT_SI
    z0 = MU0_SI * c
    eps0 = 1.0 / MU0_SI / c / c
    smax = 0.8 * (grading + 1.0) / (z0 * st.cell_si[0]) if sigma_max_si is None else sigma_max_si
    return smax / eps0 * k.unit_time


def sigma(x, smax, nx, grading=4.0):
    """normalised sigma at x (cells) along the open axis"""
    if x < NEG:
        return smax * ((NEG - x) / NEG) ** grading
    if x > nx - POS:
        return smax * ((x - (nx - POS)) / POS) ** grading
    return 0.0


def run_one(st, cell, pos, beta_x, w=None):
    """Push one electron (zero fields) and return (w0, w1, p0, p1, x0, x1) with
    x the global coordinate along the open axis."""
    ora = O.Oracle(st, 32, 0)
    k = O.constants(st)
    w0 = np.float32(k.typical_weighting if w is None else w)
    mass = float(np.float32(k.species_mass[0]) * w0)
    g = 1.0 / math.sqrt(1.0 - beta_x * beta_x)
    p = np.array([g * beta_x * mass * float(k.c), 0.0, 0.0], np.float32)
    f = setups.bin_frames(st, 0, np.array([cell], np.int32), np.array([pos], np.float32), p[None, :],
                          np.array([w0], np.float32))
    ora.set_frames(0, f)
    ora.push(0, fused=False)
    ora.shift(0)  # into the supercell of its new cell
    out = ora.get_frames(0)
    sc = int(np.nonzero(out["counts"])[0][0])
    cap = out["w"].shape[0] // out["counts"].shape[0]
    slot = sc * cap
    lc = int(out["cell"][slot])
    x1 = (sc % (st.grid[0] // 8)) * 8 + lc % 8 + float(out["pos"][0, slot])
    res = (float(w0), float(out["w"][slot]), p.astype(np.float64), out["mom"][:, slot].astype(np.float64),
           cell[0] + pos[0], x1)
    ora.close()
    return res, k


def path_integral(x_prev, x_cur, dt, smax, nx, grading=4.0):
    """integral of sigma(x(t)) dt along the straight path, by quadrature"""
    v = (x_cur - x_prev) / dt
    val, _ = integrate.quad(lambda x: sigma(x, smax, nx, grading), min(x_prev, x_cur), max(x_prev, x_cur),
                            points=[NEG, nx - POS], epsabs=0.0, epsrel=1e-12, limit=200)
    return val / abs(v)


def test_particle_at_rest_in_the_layer():
    st = pml_box()
    (w0, w1, p0, p1, x0, x1), k = run_one(st, [3, 4, 2], [0.25, 0.5, 0.5], 0.0)
    smax = sigma_max_normalised(st, k)
    # a particle at rest: the trapezoidal rule over the step, dt * sigma(x)
    want = w0 * math.exp(-float(k.dt) * sigma(x0, smax, st.grid[0]))
    assert want < w0 * 0.999  # the damping is visible
    assert w1 == pytest.approx(want, rel=2e-6)
    assert x1 == x0 and np.all(p1 == 0.0)


# a linear grading with a large conductivity, so that a short path near the
# interface still damps the weighting visibly
LINEAR = dict(sigma_kappa_grading_order=1.0, sigma_max_si=(2.0e6, 0.0, 0.0), kappa_max=(1.0, 1.0, 1.0),
              alpha_grading_order=1.0, alpha_max_si=(0.2, 0.2, 0.2))


@pytest.mark.parametrize("cell,pos,beta,pml", [([4, 4, 2], [0.5, 0.5, 0.5], -0.3, None),   # inside the -x layer
                                                ([25, 4, 2], [0.5, 0.5, 0.5], 0.4, None),  # inside the +x layer
                                                ([20, 4, 2], [0.1, 0.5, 0.5], -0.6, LINEAR),  # leaves the +x layer
                                                ([8, 4, 2], [0.2, 0.5, 0.5], -0.6, LINEAR),  # enters the -x layer
                                                ([19, 4, 2], [0.8, 0.5, 0.5], 0.6, LINEAR),  # enters the +x layer
                                                ([7, 4, 2], [0.3, 0.5, 0.5], 0.5, LINEAR)])
def test_moving_particle_path_integral(cell, pos, beta, pml):
    st = pml_box(pml=pml)
    (w0, w1, p0, p1, x0, x1), k = run_one(st, cell, pos, beta)
    grading = 4.0 if pml is None else pml["sigma_kappa_grading_order"]
    smax = sigma_max_normalised(st, k, grading, None if pml is None else pml["sigma_max_si"][0])
    dt = float(k.dt)
    # the path the damping integrates: back from the new position by v dt
    # (zero fields: the push keeps the momentum)
    mass = float(np.float32(k.species_mass[0]) * np.float32(w0))
    v = p0[0] / math.sqrt(mass * mass + p0[0] ** 2 / float(k.c) ** 2)
    x_prev = x1 - v * dt / float(k.cell[0])
    integral = path_integral(x_prev, x1, dt, smax, st.grid[0], grading)
    assert integral > 1e-6
    # FP32 evaluation of the closed form: a few ulps of its terms
    assert math.log(w0 / w1) == pytest.approx(integral, rel=1e-5, abs=2e-7)
    # the momentum scales with the weighting: the velocity stays
    assert p1[0] / w1 == pytest.approx(p0[0] / w0, rel=2e-7)


@pytest.mark.parametrize("cell,pos,beta", [([18, 4, 2], [0.5, 0.5, 0.5], 0.0),   # +x area (cells >= 16), no layer
                                            ([10, 4, 2], [0.5, 0.5, 0.5], -0.5),  # outside the -x area (cells < 8)
                                            ([9, 4, 2], [0.5, 0.5, 0.5], 0.5),
                                            # leaves the -x layer into supercell 1: the -x exchange's area is
                                            # the supercells the layer reaches (cells 0..7), so the part of the
                                            # path inside the layer is not damped (Absorbing.hpp:112-126)
                                            ([7, 4, 2], [0.9, 0.5, 0.5], 0.6)])
def test_no_damping_outside_the_areas(cell, pos, beta):
    st = pml_box()
    (w0, w1, p0, p1, x0, x1), k = run_one(st, cell, pos, beta)
    assert w1 == w0 and p1[0] == p0[0]


def test_reflecting_species_is_not_damped():
    st = pml_box(reflecting=True)
    (w0, w1, p0, p1, x0, x1), k = run_one(st, [3, 4, 2], [0.25, 0.5, 0.5], 0.3)
    assert w1 == w0 and p1[0] == p0[0]


def test_weighting_floor():
    """A strong layer damps the weighting to MIN_WEIGHTING, not below, and the
    momentum keeps the velocity."""
    pml = dict(sigma_kappa_grading_order=4.0, sigma_max_si=(1e15, 0.0, 0.0), kappa_max=(1.0, 1.0, 1.0),
               alpha_grading_order=1.0, alpha_max_si=(0.2, 0.2, 0.2))
    st = pml_box(pml=pml)
    (w0, w1, p0, p1, x0, x1), k = run_one(st, [1, 4, 2], [0.5, 0.5, 0.5], -0.2, w=50.0)
    assert w1 == 10.0
    assert p1[0] == np.float32(np.float32(p0[0]) * np.float32(np.float32(10.0) / np.float32(50.0)))


def test_custom_grading_order():
    pml = dict(sigma_kappa_grading_order=2.5, sigma_max_si=(3.0e5, 0.0, 0.0), kappa_max=(1.0, 1.0, 1.0),
               alpha_grading_order=1.0, alpha_max_si=(0.2, 0.2, 0.2))
    st = pml_box(pml=pml)
    (w0, w1, p0, p1, x0, x1), k = run_one(st, [2, 4, 2], [0.7, 0.5, 0.5], 0.0)
    smax = sigma_max_normalised(st, k, sigma_max_si=3.0e5)
    want = w0 * math.exp(-float(k.dt) * sigma(x0, smax, st.grid[0], grading=2.5))
    assert want < w0 * 0.999
    assert w1 == pytest.approx(want, rel=2e-6)


def test_fp64_exp_and_log_routines():
    """The damping's FP64 e^x and ln x (the device runs the same operations)
    against the C library: a few units in the last place of FP64, so their
    FP32 roundings are the correctly rounded FP32 values but for ties."""
    rng = np.random.default_rng(3)
    lib = O.lib()
    xs = np.concatenate([-rng.random(2000) * 700.0, -rng.random(2000) * 2.0, [0.0, -1e-300, -0.5 * math.log(2.0)],
                         rng.random(200) * 3.0])
    for x in xs:
        got, want = lib.oracle_exp_f64(float(x)), math.exp(float(x))
        assert got == pytest.approx(want, rel=4e-16, abs=0.0), x
    ds = np.concatenate([rng.random(2000), 10.0 ** -rng.uniform(0, 30, 500), [1.0, 0.5, math.sqrt(0.5), 0.75],
                         1.0 + rng.random(100)])
    for d in ds:
        got, want = lib.oracle_log_f64(float(d)), math.log(float(d))
        assert got == pytest.approx(want, rel=4e-16, abs=2e-16), d
    # depth^4 as the damping takes it, rounded to FP32, against numpy's power
    depth = rng.random(5000)
    got = np.array([np.float32(lib.oracle_exp_f64(4.0 * lib.oracle_log_f64(float(v)))) for v in depth])
    want = np.power(depth, 4.0).astype(np.float32)
    assert np.mean(got == want) > 0.999
    assert np.all(np.abs(got.astype(np.float64) - want) <= np.spacing(want))
