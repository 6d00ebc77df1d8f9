"""ctypes wrapper of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg, never by the picongpu_amd package.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")


class OracleSpecies(C.Structure):
    _fields_ = [("mass_ratio", C.c_double), ("charge_ratio", C.c_double), ("shape", C.c_int32),
                ("pusher", C.c_int32), ("frames_per_sc", C.c_int32), ("current_solver", C.c_int32),
                ("particle_boundary", C.c_int32 * 3)]


class OracleIncidentProfile(C.Structure):
    _fields_ = [("kind", C.c_int32), ("polarisation", C.c_int32), ("wavelength_si", C.c_double),
                ("amplitude_si", C.c_double), ("pulse_duration_si", C.c_double), ("time_delay_si", C.c_double),
                ("laser_phase", C.c_float), ("direction", C.c_double * 3), ("polarisation_direction", C.c_double * 3),
                ("focus_position_si", C.c_double * 3), ("focus_origin", C.c_int32 * 3), ("w0_si", C.c_double),
                ("pulse_init", C.c_double), ("num_modes", C.c_int32), ("laguerre_modes", C.c_float * 16),
                ("laguerre_phases", C.c_float * 16), ("laser_nofocus_constant_si", C.c_double),
                ("ramp_init", C.c_double)]


# parameter defaults of profiles::BaseParam / GaussianPulseParam / PlaneWaveParam
# as the setups spell them (keyword names of the incident profile dicts)
_INCIDENT_DEFAULTS = dict(focus_position_si=(0.0, 0.0, 0.0), focus_origin=(0, 0, 0), polarisation=0,
                          time_delay_si=0.0, laser_phase=0.0, w0_si=0.0, pulse_init=20.0, laguerre_modes=(1.0,),
                          laguerre_phases=None, laser_nofocus_constant_si=0.0, ramp_init=20.6146)


def _fill_incident(dst, prof):
    p = dict(_INCIDENT_DEFAULTS)
    p.update(prof)
    modes = list(p["laguerre_modes"])
    phases = list(p["laguerre_phases"]) if p["laguerre_phases"] is not None else [0.0] * len(modes)
    dst.kind = int(p["kind"])
    dst.polarisation = int(p["polarisation"])
    for name in ("wavelength_si", "amplitude_si", "pulse_duration_si", "time_delay_si", "w0_si", "pulse_init",
                 "laser_nofocus_constant_si", "ramp_init"):
        setattr(dst, name, float(p[name]))
    dst.laser_phase = float(p["laser_phase"])
    for d in range(3):
        dst.direction[d] = float(p["direction"][d])
        dst.polarisation_direction[d] = float(p["polarisation_direction"][d])
        dst.focus_position_si[d] = float(p["focus_position_si"][d])
        dst.focus_origin[d] = int(p["focus_origin"][d])
    dst.num_modes = len(modes) - 1
    for m in range(len(modes)):
        dst.laguerre_modes[m] = float(modes[m])
        dst.laguerre_phases[m] = float(phases[m])


class OraclePml(C.Structure):
    _fields_ = [("custom", C.c_int32), ("sigma_kappa_grading_order", C.c_double), ("sigma_max_si", C.c_double * 3),
                ("kappa_max", C.c_double * 3), ("alpha_grading_order", C.c_double), ("alpha_max_si", C.c_double * 3)]


class OracleConfig(C.Structure):
    _fields_ = [("grid", C.c_int32 * 3), ("cell_si", C.c_double * 3), ("dt_si", C.c_double),
                ("base_density_si", C.c_double), ("typical_ppc", C.c_int32), ("solver", C.c_int32),
                ("lehe_dir", C.c_int32), ("n_species", C.c_int32), ("species", OracleSpecies * 4),
                ("j_scale_exp", C.c_int32), ("device", C.c_int32), ("boundary", C.c_int32 * 3),
                ("absorber", C.c_int32), ("absorber_cells", (C.c_int32 * 2) * 3),
                ("absorber_strength", (C.c_float * 2) * 3), ("current_interpolation", C.c_int32),
                ("incident", (OracleIncidentProfile * 2) * 3), ("incident_position", (C.c_int32 * 2) * 3),
                ("pml", OraclePml)]


class OracleConstants(C.Structure):
    _fields_ = [("cell", C.c_float * 3), ("dt", C.c_float), ("c", C.c_float), ("eps0", C.c_float),
                ("mue0", C.c_float), ("mue0_eps0", C.c_float), ("add_current_coeff", C.c_float),
                ("inv_cell_volume_dt", C.c_float), ("lehe_alpha", C.c_float), ("lehe_delta", C.c_float),
                ("species_mass", C.c_float * 4), ("species_charge", C.c_float * 4), ("unit_length", C.c_double),
                ("unit_time", C.c_double), ("unit_mass", C.c_double), ("unit_charge", C.c_double),
                ("typical_weighting", C.c_double)]


_lib = None
_P = C.c_void_p
_F = C.POINTER(C.c_float)
_D = C.POINTER(C.c_double)
_U32 = C.POINTER(C.c_uint32)
_U16 = C.POINTER(C.c_uint16)
_U8 = C.POINTER(C.c_uint8)
_I32 = C.POINTER(C.c_int32)


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        sig = {
            "oracle_create": (_P, [C.POINTER(OracleConfig), C.c_int32, C.c_int32]),
            "oracle_destroy": (None, [_P]),
            "oracle_get_constants": (C.c_int32, [C.POINTER(OracleConfig), C.POINTER(OracleConstants)]),
            "oracle_set_jexp": (None, [_P, C.c_int32]),
            "oracle_set_threads": (C.c_int32, [C.c_int32]),
            "oracle_set_frames": (C.c_int32, [_P, C.c_int32, _U32, _F, _F, _F, _U16, _U8]),
            "oracle_get_frames": (C.c_int32, [_P, C.c_int32, _U32, _D, _D, _D, _U16, _U8]),
            "oracle_get_ids": (C.c_int32, [_P, C.c_int32, _U32]),
            "oracle_set_field": (C.c_int32, [_P, C.c_int32, _F]),
            "oracle_get_field": (C.c_int32, [_P, C.c_int32, _D]),
            "oracle_get_pml": (C.c_int32, [_P, C.c_int32, _D]),
            "oracle_exp_f64": (C.c_double, [C.c_double]),
            "oracle_log_f64": (C.c_double, [C.c_double]),
            "oracle_current_reset": (None, [_P]),
            "oracle_push": (None, [_P, C.c_int32, C.c_int32]),
            "oracle_deposit": (None, [_P, C.c_int32]),
            "oracle_gather": (None, [_P, C.c_int32, _I32, _F, _D]),
            "oracle_shift": (C.c_int32, [_P, C.c_int32]),
            "oracle_fdtd_before_current": (None, [_P]),
            "oracle_add_current": (None, [_P]),
            "oracle_fdtd_after_current": (None, [_P]),
            "oracle_step": (C.c_int32, [_P, C.c_int32]),
            "oracle_set_step": (None, [_P, C.c_int64]),
            "oracle_incident_field": (None, [C.POINTER(OracleConfig), C.c_int32, C.c_int32, C.c_int32, C.c_double,
                                             _D, C.c_int64, _D]),
            "oracle_incident_info": (None, [C.POINTER(OracleConfig), C.c_int32, C.c_int32, _D]),
            "oracle_get_step": (C.c_int64, [_P]),
            "oracle_field_energy": (None, [_P, _D]),
            "oracle_charge_density": (None, [_P, _D]),
            "oracle_div_e": (None, [_P, _D]),
            "oracle_esirkepov_single": (None, [C.POINTER(OracleConfig), C.c_int32, _I32, _F, _F, C.c_float, _D]),
            "oracle_deposit_single": (None, [C.POINTER(OracleConfig), C.c_int32, C.c_int32, _I32, _F, _F, C.c_float,
                                             _D]),
            "oracle_move_particle": (C.c_int32, [_F, _U16, _U8]),
            "oracle_shape": (None, [C.c_int32, C.c_int32, C.c_float, C.c_int32, _F]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def config_from(setup) -> OracleConfig:
    """OracleConfig from a picongpu_amd.setups.Setup (plain attribute reads)."""
    c = OracleConfig()
    for d in range(3):
        c.grid[d] = int(setup.grid[d])
        c.cell_si[d] = float(setup.cell_si[d])
    c.dt_si = float(setup.dt_si)
    c.base_density_si = float(setup.density_si)
    c.typical_ppc = int(setup.typical_ppc)
    c.solver = int(setup.solver)
    c.lehe_dir = int(setup.lehe_dir)
    c.n_species = len(setup.species)
    for i, s in enumerate(setup.species):
        c.species[i].mass_ratio = s.mass_ratio
        c.species[i].charge_ratio = s.charge_ratio
        c.species[i].shape = s.shape
        c.species[i].pusher = s.pusher
        c.species[i].frames_per_sc = s.frames()
        c.species[i].current_solver = s.current_solver
        for d in range(3):
            c.species[i].particle_boundary[d] = int(s.particle_boundary[d])
    for d in range(3):
        c.boundary[d] = int(setup.boundary[d])
        for side in range(2):
            c.absorber_cells[d][side] = int(setup.absorber_cells[d][side])
            c.absorber_strength[d][side] = float(setup.absorber_strength[d][side])
    c.absorber = int(setup.absorber)
    c.current_interpolation = int(setup.current_interpolation)
    for d in range(3):
        for side in range(2):
            c.incident_position[d][side] = int(setup.incident_position[d][side])
    for (axis, side), prof in (setup.incident or {}).items():
        _fill_incident(c.incident[axis][side], prof)
    pml = getattr(setup, "pml", None)
    if pml is not None:
        c.pml.custom = 1
        c.pml.sigma_kappa_grading_order = float(pml["sigma_kappa_grading_order"])
        c.pml.alpha_grading_order = float(pml["alpha_grading_order"])
        for d in range(3):
            c.pml.sigma_max_si[d] = float(pml["sigma_max_si"][d])
            c.pml.kappa_max[d] = float(pml["kappa_max"][d])
            c.pml.alpha_max_si[d] = float(pml["alpha_max_si"][d])
    return c


def constants(setup) -> OracleConstants:
    k = OracleConstants()
    cfg = config_from(setup)
    lib().oracle_get_constants(C.byref(cfg), C.byref(k))
    return k


class Oracle:
    """order 0: device summation order (fixed-point tiles, fused deposit);
    order 1: the reference CPU's sequential order. precision 32 or 64."""

    def __init__(self, setup, precision=32, order=0, jexp=40):
        self.setup = setup
        self.cfg = config_from(setup)
        self.h = lib().oracle_create(C.byref(self.cfg), precision, order)
        self.grid = tuple(setup.grid)
        self.ncells = int(np.prod(self.grid))
        self.nsc = (self.grid[0] // 8) * (self.grid[1] // 8) * (self.grid[2] // 4)
        lib().oracle_set_jexp(self.h, jexp)

    def close(self):
        if self.h:
            lib().oracle_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def cap(self, sp):
        return self.setup.species[sp].frames() * 256

    def set_jexp(self, e):
        lib().oracle_set_jexp(self.h, int(e))

    def set_frames(self, sp, f):
        slots = self.nsc * self.cap(sp)
        if len(f["counts"]) != self.nsc or len(f["cell"]) != slots or len(f["mask"]) != slots \
                or np.shape(f["pos"]) != (3, slots) or np.shape(f["mom"]) != (3, slots) or len(f["w"]) != slots:
            raise ValueError(f"frames of species {sp} do not match the layout ({self.nsc} supercells x {self.cap(sp)} slots)")
        if np.asarray(f["counts"]).max(initial=0) > self.cap(sp):
            raise ValueError(f"a supercell count of species {sp} exceeds its {self.cap(sp)} slots")
        lib().oracle_set_frames(self.h, sp, _p(f["counts"], C.c_uint32), _p(np.ascontiguousarray(f["pos"], np.float32), C.c_float),
                                _p(np.ascontiguousarray(f["mom"], np.float32), C.c_float),
                                _p(np.ascontiguousarray(f["w"], np.float32), C.c_float),
                                _p(f["cell"], C.c_uint16), _p(f["mask"], C.c_uint8))

    def get_frames(self, sp):
        slots = self.nsc * self.cap(sp)
        f = {"counts": np.zeros(self.nsc, np.uint32), "pos": np.zeros((3, slots)), "mom": np.zeros((3, slots)),
             "w": np.zeros(slots), "cell": np.zeros(slots, np.uint16), "mask": np.zeros(slots, np.uint8)}
        lib().oracle_get_frames(self.h, sp, _p(f["counts"], C.c_uint32), _p(f["pos"], C.c_double),
                                _p(f["mom"], C.c_double), _p(f["w"], C.c_double), _p(f["cell"], C.c_uint16),
                                _p(f["mask"], C.c_uint8))
        return f

    def get_ids(self, sp):
        ids = np.zeros(self.nsc * self.cap(sp), np.uint32)
        lib().oracle_get_ids(self.h, sp, _p(ids, C.c_uint32))
        return ids

    def set_field(self, which, data):
        data = np.ascontiguousarray(data, np.float32).reshape(3, -1)
        lib().oracle_set_field(self.h, which, _p(data, C.c_float))

    def get_field(self, which):
        out = np.zeros((3, self.ncells))
        lib().oracle_get_field(self.h, which, _p(out, C.c_double))
        return out

    def get_pml(self, which):
        """psi of E (which 0) or B (1): (6, ncells) in NodeValues order xy, xz, yx, yz, zx, zy."""
        out = np.zeros((6, self.ncells))
        if lib().oracle_get_pml(self.h, which, _p(out, C.c_double)) != 0:
            raise ValueError("the state has no PML")
        return out

    def current_reset(self):
        lib().oracle_current_reset(self.h)

    def push(self, species=-1, fused=True):
        lib().oracle_push(self.h, species, 1 if fused else 0)

    def deposit(self, species=-1):
        lib().oracle_deposit(self.h, species)

    def shift(self, species=-1):
        r = lib().oracle_shift(self.h, species)
        if r != 0:
            raise RuntimeError("oracle: supercell capacity exceeded")

    def update_before_current(self):
        lib().oracle_fdtd_before_current(self.h)

    def add_current(self):
        lib().oracle_add_current(self.h)

    def update_after_current(self):
        lib().oracle_fdtd_after_current(self.h)

    def set_step(self, step):
        lib().oracle_set_step(self.h, int(step))

    def get_step(self) -> int:
        return int(lib().oracle_get_step(self.h))

    def step(self, n=1):
        r = lib().oracle_step(self.h, int(n))
        if r != 0:
            raise RuntimeError("oracle: supercell capacity exceeded")

    def gather(self, shape, cell, pos):
        """E and B of the current fields at one particle (global cell, in-cell
        position), interpolated with the shape: (Ex, Ey, Ez, Bx, By, Bz)."""
        c = np.ascontiguousarray(cell, np.int32)
        q = np.ascontiguousarray(pos, np.float32)
        out = np.zeros(6)
        lib().oracle_gather(self.h, int(shape), _p(c, C.c_int32), _p(q, C.c_float), _p(out, C.c_double))
        return out

    def field_energy(self):
        """EnergyFields output in joules: total, B x, y, z, E x, y, z."""
        out = np.zeros(7)
        lib().oracle_field_energy(self.h, _p(out, C.c_double))
        return out

    def charge_density(self):
        out = np.zeros(self.ncells)
        lib().oracle_charge_density(self.h, _p(out, C.c_double))
        return out

    def div_e(self):
        out = np.zeros(self.ncells)
        lib().oracle_div_e(self.h, _p(out, C.c_double))
        return out


def set_threads(n: int) -> int:
    return lib().oracle_set_threads(int(n))


def esirkepov_single(setup, shape, cell, pos, vel, charge):
    cfg = config_from(setup)
    n = int(np.prod(setup.grid))
    J = np.zeros((3, n))
    cell = np.ascontiguousarray(cell, np.int32)
    pos = np.ascontiguousarray(pos, np.float32)
    vel = np.ascontiguousarray(vel, np.float32)
    lib().oracle_esirkepov_single(C.byref(cfg), shape, _p(cell, C.c_int32), _p(pos, C.c_float), _p(vel, C.c_float),
                                  float(charge), _p(J, C.c_double))
    return J


def deposit_single(setup, shape, scheme, cell, pos, vel, charge):
    """One particle's current with the Esirkepov (scheme 0) or EmZ (1) solver."""
    cfg = config_from(setup)
    n = int(np.prod(setup.grid))
    J = np.zeros((3, n))
    cell = np.ascontiguousarray(cell, np.int32)
    pos = np.ascontiguousarray(pos, np.float32)
    vel = np.ascontiguousarray(vel, np.float32)
    lib().oracle_deposit_single(C.byref(cfg), shape, scheme, _p(cell, C.c_int32), _p(pos, C.c_float),
                                _p(vel, C.c_float), float(charge), _p(J, C.c_double))
    return J


def move_particle(pos, local_cell=0, mask=1):
    p = np.ascontiguousarray(pos, np.float32).copy()
    lc = C.c_uint16(local_cell)
    m = C.c_uint8(mask)
    r = lib().oracle_move_particle(_p(p, C.c_float), C.byref(lc), C.byref(m))
    return p, lc.value, m.value, r


def shape(shape_id, kind, x, flag=0):
    out = np.zeros(5, np.float32)
    lib().oracle_shape(shape_id, kind, float(x), int(flag), _p(out, C.c_float))
    return out


def incident_field(setup, axis, side, which, current_step, pos):
    """FP64 incident field (which 0: E, 1: B) of face (axis, side) at total
    cell positions pos (n, 3) and source step current_step; returns (n, 3)."""
    cfg = config_from(setup)
    pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
    out = np.zeros_like(pos)
    lib().oracle_incident_field(C.byref(cfg), axis, side, which, float(current_step), _p(pos, C.c_double),
                                pos.shape[0], _p(out, C.c_double))
    return out


def incident_info(setup, axis, side):
    """(phase velocity / c, origin[3], focus[3]) of face (axis, side), FP64, PIC units."""
    cfg = config_from(setup)
    out = np.zeros(7)
    lib().oracle_incident_info(C.byref(cfg), axis, side, _p(out, C.c_double))
    return out[0], out[1:4], out[4:7]
