"""ctypes binding of the C ABI (include/picongpu_amd.h).

`Simulation` owns one pic_ctx and exposes the stages of one PIC step with the
names of the reference's stage functors (simulation/stage/*.hpp,
simulation/control/Simulation.hpp:522-542). The library must be built
(picongpu_amd.build); there is no fallback path: a missing library raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import build as _build

MAX_SPECIES = 4

SHAPE_CIC, SHAPE_TSC, SHAPE_PQS = 1, 2, 3
PUSHER_NONE, PUSHER_BORIS, PUSHER_VAY, PUSHER_HIGUERA_CARY = 0, 1, 2, 3
BOUNDARY_PERIODIC, BOUNDARY_OPEN = 0, 1
PARTICLE_ABSORBING, PARTICLE_REFLECTING = 0, 1
ABSORBER_NONE, ABSORBER_EXPONENTIAL, ABSORBER_PML = 0, 1, 2
CURRENT_NONE, CURRENT_BINOMIAL = 0, 1
CURRENT_SOLVER_ESIRKEPOV, CURRENT_SOLVER_EMZ = 0, 1
SOLVER_YEE, SOLVER_LEHE = 0, 1
INCIDENT_NONE, INCIDENT_GAUSSIAN_PULSE, INCIDENT_PLANE_WAVE = 0, 1, 2
POLARISATION_LINEAR, POLARISATION_CIRCULAR = 0, 1
ORIGIN_ZERO, ORIGIN_CENTER = 0, 1
INCIDENT_MAX_MODES = 16
FIELD_E, FIELD_B, FIELD_J = 0, 1, 2
# the PML's convolution fields: six planes in NodeValues order xy, xz, yx, yz, zx, zy
FIELD_PML_E, FIELD_PML_B = 3, 4
PML_COMPONENTS = ("xy", "xz", "yx", "yz", "zx", "zy")
TIMER_PUSH, TIMER_SHIFT, TIMER_FIELD = 0, 1, 2
PUSH_SCHEDULE_AUTO, PUSH_SCHEDULE_PHASES, PUSH_SCHEDULE_SINGLE = 0, 1, 2

ERR = {0: "ok", -1: "invalid", -2: "hip", -3: "capacity", -4: "state"}
FLAG_CAPACITY_GREW = 1  # PIC_FLAG_CAPACITY_GREW: a supercell overflowed and the supercells grew


class SpeciesDesc(C.Structure):
    _fields_ = [
        ("mass_ratio", C.c_double),
        ("charge_ratio", C.c_double),
        ("shape", C.c_int32),
        ("pusher", C.c_int32),
        ("frames_per_sc", C.c_int32),
        ("current_solver", C.c_int32),
        ("particle_boundary", C.c_int32 * 3),
    ]


class IncidentProfile(C.Structure):
    """pic_incident_profile: one face's profile of incidentField.param, in SI."""
    _fields_ = [
        ("kind", C.c_int32),
        ("polarisation", C.c_int32),
        ("wavelength_si", C.c_double),
        ("amplitude_si", C.c_double),
        ("pulse_duration_si", C.c_double),
        ("time_delay_si", C.c_double),
        ("laser_phase", C.c_float),
        ("direction", C.c_double * 3),
        ("polarisation_direction", C.c_double * 3),
        ("focus_position_si", C.c_double * 3),
        ("focus_origin", C.c_int32 * 3),
        ("w0_si", C.c_double),
        ("pulse_init", C.c_double),
        ("num_modes", C.c_int32),
        ("laguerre_modes", C.c_float * INCIDENT_MAX_MODES),
        ("laguerre_phases", C.c_float * INCIDENT_MAX_MODES),
        ("laser_nofocus_constant_si", C.c_double),
        ("ramp_init", C.c_double),
    ]


def fill_incident_profile(dst, kind, wavelength_si, amplitude_si, pulse_duration_si, direction,
                          polarisation_direction, focus_position_si=(0.0, 0.0, 0.0), focus_origin=(0, 0, 0),
                          polarisation=POLARISATION_LINEAR, time_delay_si=0.0, laser_phase=0.0, w0_si=0.0,
                          pulse_init=20.0, laguerre_modes=(1.0,), laguerre_phases=None,
                          laser_nofocus_constant_si=0.0, ramp_init=20.6146):
    """Fill a pic_incident_profile-shaped ctypes struct from the parameter
    names of profiles::BaseParam / GaussianPulseParam / PlaneWaveParam (SI)."""
    modes = list(laguerre_modes)
    phases = list(laguerre_phases) if laguerre_phases is not None else [0.0] * len(modes)
    if not 1 <= len(modes) <= INCIDENT_MAX_MODES or len(phases) != len(modes):
        raise ValueError("laguerre_modes / laguerre_phases: 1..16 entries of equal length")
    dst.kind = int(kind)
    dst.polarisation = int(polarisation)
    dst.wavelength_si = float(wavelength_si)
    dst.amplitude_si = float(amplitude_si)
    dst.pulse_duration_si = float(pulse_duration_si)
    dst.time_delay_si = float(time_delay_si)
    dst.laser_phase = float(laser_phase)
    for d in range(3):
        dst.direction[d] = float(direction[d])
        dst.polarisation_direction[d] = float(polarisation_direction[d])
        dst.focus_position_si[d] = float(focus_position_si[d])
        dst.focus_origin[d] = int(focus_origin[d])
    dst.w0_si = float(w0_si)
    dst.pulse_init = float(pulse_init)
    dst.num_modes = len(modes) - 1
    for m in range(len(modes)):
        dst.laguerre_modes[m] = float(modes[m])
        dst.laguerre_phases[m] = float(phases[m])
    dst.laser_nofocus_constant_si = float(laser_nofocus_constant_si)
    dst.ramp_init = float(ramp_init)
    return dst


class PmlParams(C.Structure):
    _fields_ = [
        ("custom", C.c_int32),
        ("sigma_kappa_grading_order", C.c_double),
        ("sigma_max_si", C.c_double * 3),
        ("kappa_max", C.c_double * 3),
        ("alpha_grading_order", C.c_double),
        ("alpha_max_si", C.c_double * 3),
    ]


def fill_pml(dst: PmlParams, pml) -> None:
    """pml: None (the fieldAbsorber.param defaults) or a dict with
    sigma_kappa_grading_order, sigma_max_si (3), kappa_max (3),
    alpha_grading_order, alpha_max_si (3)."""
    if pml is None:
        dst.custom = 0
        return
    dst.custom = 1
    dst.sigma_kappa_grading_order = float(pml["sigma_kappa_grading_order"])
    dst.alpha_grading_order = float(pml["alpha_grading_order"])
    for d in range(3):
        dst.sigma_max_si[d] = float(pml["sigma_max_si"][d])
        dst.kappa_max[d] = float(pml["kappa_max"][d])
        dst.alpha_max_si[d] = float(pml["alpha_max_si"][d])


class Config(C.Structure):
    _fields_ = [
        ("grid", C.c_int32 * 3),
        ("cell_si", C.c_double * 3),
        ("dt_si", C.c_double),
        ("base_density_si", C.c_double),
        ("typical_ppc", C.c_int32),
        ("solver", C.c_int32),
        ("lehe_dir", C.c_int32),
        ("n_species", C.c_int32),
        ("species", SpeciesDesc * MAX_SPECIES),
        ("j_scale_exp", C.c_int32),
        ("device", C.c_int32),
        ("boundary", C.c_int32 * 3),
        ("absorber", C.c_int32),
        ("absorber_cells", (C.c_int32 * 2) * 3),
        ("absorber_strength", (C.c_float * 2) * 3),
        ("current_interpolation", C.c_int32),
        ("incident", (IncidentProfile * 2) * 3),
        ("incident_position", (C.c_int32 * 2) * 3),
        ("pml", PmlParams),
    ]


class Constants(C.Structure):
    _fields_ = [
        ("cell", C.c_float * 3),
        ("dt", C.c_float),
        ("c", C.c_float),
        ("eps0", C.c_float),
        ("mue0", C.c_float),
        ("mue0_eps0", C.c_float),
        ("add_current_coeff", C.c_float),
        ("inv_cell_volume_dt", C.c_float),
        ("lehe_alpha", C.c_float),
        ("lehe_delta", C.c_float),
        ("species_mass", C.c_float * MAX_SPECIES),
        ("species_charge", C.c_float * MAX_SPECIES),
        ("unit_length", C.c_double),
        ("unit_time", C.c_double),
        ("unit_mass", C.c_double),
        ("unit_charge", C.c_double),
        ("typical_weighting", C.c_double),
    ]


class InitDesc(C.Structure):
    _fields_ = [
        ("seed", C.c_uint64),
        ("ppc", C.c_int32),
        ("drift_band", C.c_int32),
        ("density_si", C.c_double),
        ("u_drift", C.c_double * 3),
        ("u_th", C.c_double * 3),
        ("profile", C.c_int32),
        ("profile_axis", C.c_int32),
        ("ramp_start_si", C.c_double),
        ("ramp_length_si", C.c_double),
        ("quiet", C.c_int32 * 3),
    ]


class LaserDesc(C.Structure):
    _fields_ = [
        ("wavelength_si", C.c_double),
        ("a0", C.c_double),
        ("waist_si", C.c_double),
        ("fwhm_si", C.c_double),
        ("center_si", C.c_double * 3),
        ("propagation_axis", C.c_int32),
        ("polarisation_axis", C.c_int32),
    ]


_P = C.c_void_p
_F = C.POINTER(C.c_float)
_I32 = C.POINTER(C.c_int32)
_U32 = C.POINTER(C.c_uint32)
_U16 = C.POINTER(C.c_uint16)
_U8 = C.POINTER(C.c_uint8)
_D = C.POINTER(C.c_double)

# name -> (restype, argtypes); the exact set of symbols include/picongpu_amd.h declares
SIGNATURES = {
    "pic_last_error": (C.c_char_p, []),
    "pic_abi_version": (C.c_int32, []),
    "pic_constants_from_config": (C.c_int, [C.POINTER(Config), C.POINTER(Constants)]),
    "pic_create": (C.c_int, [C.POINTER(Config), C.POINTER(_P)]),
    "pic_destroy": (C.c_int, [_P]),
    "pic_get_constants": (C.c_int, [_P, C.POINTER(Constants)]),
    "pic_sync": (C.c_int, [_P]),
    "pic_load_particles": (C.c_int, [_P, C.c_int32, C.c_int64, _I32, _F, _F, _F]),
    "pic_read_particles": (C.c_int, [_P, C.c_int32, C.c_int64, C.POINTER(C.c_int64), _I32, _F, _F, _F]),
    "pic_particle_count": (C.c_int64, [_P, C.c_int32]),
    "pic_num_supercells": (C.c_int32, [_P]),
    "pic_slots_per_supercell": (C.c_int32, [_P, C.c_int32]),
    "pic_supercell_slots": (C.c_int, [_P, C.c_int32, _U32]),
    "pic_read_frames": (C.c_int, [_P, C.c_int32, _U32, _F, _F, _F, _U16, _U8]),
    "pic_write_frames": (C.c_int, [_P, C.c_int32, _U32, _F, _F, _F, _U16, _U8]),
    "pic_init_species": (C.c_int, [_P, C.c_int32, C.POINTER(InitDesc)]),
    "pic_load_field": (C.c_int, [_P, C.c_int32, _F]),
    "pic_read_field": (C.c_int, [_P, C.c_int32, _F]),
    "pic_add_field_noise": (C.c_int, [_P, C.c_int32, C.c_uint64, C.c_float]),
    "pic_add_laser": (C.c_int, [_P, C.POINTER(LaserDesc)]),
    "pic_current_reset": (C.c_int, [_P]),
    "pic_push": (C.c_int, [_P, C.c_int32, C.c_int32]),
    "pic_shift": (C.c_int, [_P, C.c_int32]),
    "pic_deposit": (C.c_int, [_P, C.c_int32]),
    "pic_fdtd_before_current": (C.c_int, [_P]),
    "pic_fold_current": (C.c_int, [_P]),
    "pic_add_current": (C.c_int, [_P]),
    "pic_fdtd_after_current": (C.c_int, [_P]),
    "pic_fdtd_step": (C.c_int, [_P]),
    "pic_set_step": (C.c_int, [_P, C.c_int64]),
    "pic_get_step": (C.c_int, [_P, C.POINTER(C.c_int64)]),
    "pic_step": (C.c_int, [_P, C.c_int32]),
    "pic_step_graph": (C.c_int, [_P, C.c_int32]),
    "pic_field_energy": (C.c_int, [_P, _D]),
    "pic_gauss_residual": (C.c_int, [_P, _D]),
    "pic_gauss_residual_noalloc": (C.c_int, [_P, _D]),
    "pic_error_flags": (C.c_int, [_P, _U32]),
    "pic_timing_enable": (C.c_int, [_P, C.c_int32]),
    "pic_kernel_timer": (C.c_int, [_P, C.c_int32, _D, C.POINTER(C.c_int64)]),
    "pic_j_scale_exp": (C.c_int32, [_P]),
    "pic_num_push_phases": (C.c_int32, [_P, C.c_int32]),
    "pic_set_push_schedule": (C.c_int, [_P, C.c_int32]),
    "pic_push_schedule": (C.c_int32, [_P, C.c_int32]),
    "pic_adjoint_load_particles": (C.c_int, [_P, C.c_int32, C.c_int64, _I32, _F, _F, _F]),
    "pic_adjoint_read_particles": (C.c_int, [_P, _I32, _F, _F]),
    "pic_adjoint_forward": (C.c_int, [_P, _P, _P, C.c_int32, C.c_int32, _P, _P]),
    "pic_adjoint_backward": (C.c_int, [_P, _P, _P, _P, _P, _P, _P]),
    "pic_ensemble_create": (C.c_int, [C.POINTER(Config), C.c_int32, C.POINTER(_P)]),
    "pic_ensemble_destroy": (C.c_int, [_P]),
    "pic_ensemble_size": (C.c_int32, [_P]),
    "pic_ensemble_member": (_P, [_P, C.c_int32]),
    "pic_ensemble_step": (C.c_int, [_P, C.c_int32]),
    "pic_ensemble_step_graph": (C.c_int, [_P, C.c_int32]),
    "pic_ensemble_set_schedule": (C.c_int, [_P, C.c_int32, C.c_int32]),
    "pic_ensemble_get_schedule": (C.c_int, [_P, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "pic_ensemble_autotune": (C.c_int, [_P, C.c_int32, C.POINTER(C.c_double), C.POINTER(C.c_int32)]),
}

_lib = None


def library_path() -> str:
    return _build.LIB_PATH


def load_library() -> C.CDLL:
    """Load the HIP library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    path = library_path()
    if not os.path.exists(path):
        raise RuntimeError(f"picongpu_amd HIP library missing at {path}; run picongpu_amd.build.build_library()")
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class PicError(RuntimeError):
    def __init__(self, code: int, where: str, msg: str):
        super().__init__(f"{where} failed ({ERR.get(code, code)}): {msg}")
        self.code = code


def check(code: int, where: str) -> None:
    if code != 0:
        raise PicError(code, where, load_library().pic_last_error().decode())


def ptr(a: np.ndarray, ctype):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(ctype))


def make_config(grid, cell_si, dt_si, density_si, ppc, species, solver=SOLVER_YEE, lehe_dir=1, j_scale_exp=0,
                device=0, boundary=(0, 0, 0), absorber=0, absorber_cells=((0, 0), (0, 0), (0, 0)),
                absorber_strength=((0.0, 0.0), (0.0, 0.0), (0.0, 0.0)), current_interpolation=0, incident=None,
                incident_position=((16, -16), (16, -16), (16, -16)), pml=None) -> Config:
    """species: list of dicts with mass_ratio, charge_ratio, shape, pusher, frames_per_sc;
    boundary: BOUNDARY_* per axis; absorber: ABSORBER_*, with thickness and
    strength per [axis][negative, positive] side as in fieldAbsorber.param,
    pml the PML parameters (fill_pml; None: the reference's defaults);
    incident: {(axis, side): keyword arguments of fill_incident_profile} for
    the faces of incidentField.param that carry a profile, incident_position
    its POSITION table."""
    cfg = Config()
    for d in range(3):
        cfg.grid[d] = int(grid[d])
        cfg.cell_si[d] = float(cell_si[d])
    cfg.dt_si = float(dt_si)
    cfg.base_density_si = float(density_si)
    cfg.typical_ppc = int(ppc)
    cfg.solver = int(solver)
    cfg.lehe_dir = int(lehe_dir)
    cfg.n_species = len(species)
    for i, s in enumerate(species):
        cfg.species[i].mass_ratio = float(s["mass_ratio"])
        cfg.species[i].charge_ratio = float(s["charge_ratio"])
        cfg.species[i].shape = int(s["shape"])
        cfg.species[i].pusher = int(s["pusher"])
        cfg.species[i].frames_per_sc = int(s["frames_per_sc"])
        cfg.species[i].current_solver = int(s.get("current_solver", 0))
        for d, kind in enumerate(s.get("particle_boundary", (0, 0, 0))):
            cfg.species[i].particle_boundary[d] = int(kind)
    cfg.j_scale_exp = int(j_scale_exp)
    cfg.device = int(device)
    for d in range(3):
        cfg.boundary[d] = int(boundary[d])
        for side in range(2):
            cfg.absorber_cells[d][side] = int(absorber_cells[d][side])
            cfg.absorber_strength[d][side] = float(absorber_strength[d][side])
    cfg.absorber = int(absorber)
    cfg.current_interpolation = int(current_interpolation)
    for d in range(3):
        for side in range(2):
            cfg.incident_position[d][side] = int(incident_position[d][side])
    for (axis, side), prof in (incident or {}).items():
        fill_incident_profile(cfg.incident[axis][side], **prof)
    fill_pml(cfg.pml, pml)
    return cfg


def constants_from_config(cfg: Config) -> Constants:
    k = Constants()
    check(load_library().pic_constants_from_config(C.byref(cfg), C.byref(k)), "pic_constants_from_config")
    return k


class Simulation:
    """One periodic simulation box on the GPU (a pic_ctx)."""

    def __init__(self, cfg: Config, handle=None):
        self.lib = load_library()
        self.cfg = cfg
        self.owned = handle is None
        if handle is None:
            h = C.c_void_p()
            check(self.lib.pic_create(C.byref(cfg), C.byref(h)), "pic_create")
        else:
            h = C.c_void_p(handle)
        self.h = h
        self.grid = tuple(cfg.grid)
        self.ncells = int(np.prod(self.grid))
        self.nsc = self.lib.pic_num_supercells(h)

    def close(self):
        if self.h:
            if self.owned:
                self.lib.pic_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- state -------------------------------------------------------------
    def constants(self) -> Constants:
        k = Constants()
        check(self.lib.pic_get_constants(self.h, C.byref(k)), "pic_get_constants")
        return k

    def slots(self, sp: int) -> int:
        return self.lib.pic_slots_per_supercell(self.h, sp)

    def supercell_slots(self, sp: int) -> np.ndarray:
        """Capacity (slots) of every supercell of the species."""
        out = np.zeros(self.nsc, np.uint32)
        check(self.lib.pic_supercell_slots(self.h, sp, ptr(out, C.c_uint32)), "pic_supercell_slots")
        return out

    def load_particles(self, sp, cell, pos, mom, w):
        cell = np.ascontiguousarray(cell, dtype=np.int32)
        pos = np.ascontiguousarray(pos, dtype=np.float32)
        mom = np.ascontiguousarray(mom, dtype=np.float32)
        w = np.ascontiguousarray(w, dtype=np.float32)
        n = w.shape[0]
        check(self.lib.pic_load_particles(self.h, sp, n, ptr(cell, C.c_int32), ptr(pos, C.c_float),
                                          ptr(mom, C.c_float), ptr(w, C.c_float)), "pic_load_particles")

    def read_particles(self, sp):
        n = C.c_int64()
        check(self.lib.pic_read_particles(self.h, sp, 0, C.byref(n), None, None, None, None), "pic_read_particles")
        n = n.value
        cell = np.zeros((n, 3), np.int32)
        pos = np.zeros((n, 3), np.float32)
        mom = np.zeros((n, 3), np.float32)
        w = np.zeros(n, np.float32)
        m = C.c_int64()
        check(self.lib.pic_read_particles(self.h, sp, n, C.byref(m), ptr(cell, C.c_int32), ptr(pos, C.c_float),
                                          ptr(mom, C.c_float), ptr(w, C.c_float)), "pic_read_particles")
        return cell, pos, mom, w

    def particle_count(self, sp) -> int:
        n = self.lib.pic_particle_count(self.h, sp)
        if n < 0:
            raise PicError(-2, "pic_particle_count", self.lib.pic_last_error().decode())
        return int(n)

    def read_frames(self, sp):
        cap = self.slots(sp)
        slots = self.nsc * cap
        f = {
            "counts": np.zeros(self.nsc, np.uint32),
            "pos": np.zeros((3, slots), np.float32),
            "mom": np.zeros((3, slots), np.float32),
            "w": np.zeros(slots, np.float32),
            "cell": np.zeros(slots, np.uint16),
            "mask": np.zeros(slots, np.uint8),
        }
        check(self.lib.pic_read_frames(self.h, sp, ptr(f["counts"], C.c_uint32), ptr(f["pos"], C.c_float),
                                       ptr(f["mom"], C.c_float), ptr(f["w"], C.c_float), ptr(f["cell"], C.c_uint16),
                                       ptr(f["mask"], C.c_uint8)), "pic_read_frames")
        return f

    def write_frames(self, sp, f):
        check(self.lib.pic_write_frames(self.h, sp, ptr(f["counts"], C.c_uint32), ptr(f["pos"], C.c_float),
                                        ptr(f["mom"], C.c_float), ptr(f["w"], C.c_float), ptr(f["cell"], C.c_uint16),
                                        ptr(f["mask"], C.c_uint8)), "pic_write_frames")

    def init_species(self, sp, seed, ppc, density_si, u_drift=(0, 0, 0), u_th=(0, 0, 0), drift_band=0, ramp=None,
                     quiet=None):
        """ramp: None (uniform) or (axis, start_si, length_si) of a linear density ramp;
        quiet: None (random positions) or particles per dimension of the quiet start"""
        d = InitDesc()
        d.seed = int(seed)
        d.ppc = int(ppc)
        d.drift_band = int(drift_band)
        d.density_si = float(density_si)
        for i in range(3):
            d.u_drift[i] = float(u_drift[i])
            d.u_th[i] = float(u_th[i])
        if ramp is not None:
            d.profile = 1
            d.profile_axis = int(ramp[0])
            d.ramp_start_si = float(ramp[1])
            d.ramp_length_si = float(ramp[2])
        if quiet is not None:
            for i in range(3):
                d.quiet[i] = int(quiet[i])
        check(self.lib.pic_init_species(self.h, sp, C.byref(d)), "pic_init_species")

    def add_laser(self, wavelength_si, a0, waist_si, fwhm_si, center_si, propagation_axis, polarisation_axis):
        d = LaserDesc()
        d.wavelength_si = float(wavelength_si)
        d.a0 = float(a0)
        d.waist_si = float(waist_si)
        d.fwhm_si = float(fwhm_si)
        for i in range(3):
            d.center_si[i] = float(center_si[i])
        d.propagation_axis = int(propagation_axis)
        d.polarisation_axis = int(polarisation_axis)
        check(self.lib.pic_add_laser(self.h, C.byref(d)), "pic_add_laser")

    @staticmethod
    def field_planes(which):
        return 6 if which in (FIELD_PML_E, FIELD_PML_B) else 3

    def load_field(self, which, data):
        data = np.ascontiguousarray(data, dtype=np.float32).reshape(self.field_planes(which), -1)
        assert data.shape[1] == self.ncells
        check(self.lib.pic_load_field(self.h, which, ptr(data, C.c_float)), "pic_load_field")

    def read_field(self, which):
        out = np.zeros((self.field_planes(which), self.ncells), np.float32)
        check(self.lib.pic_read_field(self.h, which, ptr(out, C.c_float)), "pic_read_field")
        return out

    def add_field_noise(self, which, seed, amp):
        check(self.lib.pic_add_field_noise(self.h, which, int(seed), float(amp)), "pic_add_field_noise")

    # -- stages (Simulation::runOneStep order) -------------------------------
    def current_reset(self):
        check(self.lib.pic_current_reset(self.h), "pic_current_reset")

    def push(self, species=-1, fuse_deposit=True):
        check(self.lib.pic_push(self.h, species, 1 if fuse_deposit else 0), "pic_push")

    def shift(self, species=-1):
        check(self.lib.pic_shift(self.h, species), "pic_shift")

    def deposit(self, species=-1):
        check(self.lib.pic_deposit(self.h, species), "pic_deposit")

    def update_before_current(self):
        check(self.lib.pic_fdtd_before_current(self.h), "pic_fdtd_before_current")

    def fold_current(self):
        """FieldJ::asyncCommunication: the fold is part of the deposit's flush; this only checks the context."""
        check(self.lib.pic_fold_current(self.h), "pic_fold_current")

    def add_current(self):
        check(self.lib.pic_add_current(self.h), "pic_add_current")

    def update_after_current(self):
        check(self.lib.pic_fdtd_after_current(self.h), "pic_fdtd_after_current")

    def fdtd_step(self):
        """update_beforeCurrent + addCurrent + update_afterCurrent in the fused kernels."""
        check(self.lib.pic_fdtd_step(self.h), "pic_fdtd_step")

    def set_step(self, step):
        """Set the field step counter (the solver's currentStep, which times the incident field)."""
        check(self.lib.pic_set_step(self.h, int(step)), "pic_set_step")

    def get_step(self) -> int:
        v = C.c_int64()
        check(self.lib.pic_get_step(self.h, C.byref(v)), "pic_get_step")
        return int(v.value)

    def step(self, n=1):
        check(self.lib.pic_step(self.h, int(n)), "pic_step")

    def step_graph(self, n=1):
        check(self.lib.pic_step_graph(self.h, int(n)), "pic_step_graph")

    def sync(self):
        check(self.lib.pic_sync(self.h), "pic_sync")

    # -- diagnostics ---------------------------------------------------------
    def field_energy(self):
        """EnergyFields output in joules: total, B x, y, z, E x, y, z."""
        out = np.zeros(7, np.float64)
        check(self.lib.pic_field_energy(self.h, ptr(out, C.c_double)), "pic_field_energy")
        return out

    def gauss_residual(self):
        out = np.zeros(self.ncells, np.float64)
        check(self.lib.pic_gauss_residual(self.h, ptr(out, C.c_double)), "pic_gauss_residual")
        return out

    def gauss_residual_noalloc(self):
        """pic_gauss_residual_noalloc: the same residual with no device
        allocation (grids up to 2^21 cells)."""
        out = np.zeros(self.ncells, np.float64)
        check(self.lib.pic_gauss_residual_noalloc(self.h, ptr(out, C.c_double)), "pic_gauss_residual_noalloc")
        return out

    def error_flags(self) -> int:
        f = C.c_uint32()
        check(self.lib.pic_error_flags(self.h, C.byref(f)), "pic_error_flags")
        return f.value

    def timing(self, enable: bool):
        check(self.lib.pic_timing_enable(self.h, 1 if enable else 0), "pic_timing_enable")

    def kernel_timer(self, timer: int):
        ms = C.c_double()
        n = C.c_int64()
        check(self.lib.pic_kernel_timer(self.h, timer, C.byref(ms), C.byref(n)), "pic_kernel_timer")
        return ms.value, n.value

    def j_scale_exp(self) -> int:
        return self.lib.pic_j_scale_exp(self.h)

    def num_push_phases(self, fused=True) -> int:
        return self.lib.pic_num_push_phases(self.h, 1 if fused else 0)

    def set_push_schedule(self, schedule: int) -> None:
        """PUSH_SCHEDULE_AUTO, _PHASES or _SINGLE for the fused push."""
        check(self.lib.pic_set_push_schedule(self.h, int(schedule)), "pic_set_push_schedule")

    def push_schedule(self, support: int) -> int:
        return self.lib.pic_push_schedule(self.h, int(support))

    # --- differentiable form (device buffers are passed as integer addresses) ---
    def adjoint_load_particles(self, sp, cell, pos, mom, w):
        """cell: (3, n) int32 global cell index; pos, mom: (3, n) float32; w: (n,) float32"""
        cell = np.ascontiguousarray(cell, np.int32)
        pos = np.ascontiguousarray(pos, np.float32)
        mom = np.ascontiguousarray(mom, np.float32)
        w = np.ascontiguousarray(w, np.float32)
        n = w.shape[0]
        assert cell.shape == (3, n) and pos.shape == (3, n) and mom.shape == (3, n)
        check(self.lib.pic_adjoint_load_particles(self.h, sp, n, ptr(cell, C.c_int32), ptr(pos, C.c_float),
                                                  ptr(mom, C.c_float), ptr(w, C.c_float)),
              "pic_adjoint_load_particles")
        self.adj_n = n

    def adjoint_read_particles(self):
        n = self.adj_n
        cell = np.zeros((3, n), np.int32)
        pos = np.zeros((3, n), np.float32)
        mom = np.zeros((3, n), np.float32)
        check(self.lib.pic_adjoint_read_particles(self.h, ptr(cell, C.c_int32), ptr(pos, C.c_float),
                                                  ptr(mom, C.c_float)), "pic_adjoint_read_particles")
        return cell, pos, mom

    def adjoint_forward(self, E0_addr, B0_addr, nsteps, every, E_out_addr=None, B_out_addr=None):
        check(self.lib.pic_adjoint_forward(self.h, E0_addr, B0_addr, int(nsteps), int(every), E_out_addr,
                                           B_out_addr), "pic_adjoint_forward")

    def adjoint_backward(self, dE_addr, dB_addr, dE0_addr, dB0_addr, dX0_addr=None, dP0_addr=None):
        check(self.lib.pic_adjoint_backward(self.h, dE_addr, dB_addr, dE0_addr, dB0_addr, dX0_addr, dP0_addr),
              "pic_adjoint_backward")


ENSEMBLE_ORDER_SIM_MAJOR, ENSEMBLE_ORDER_INTERLEAVED, ENSEMBLE_ORDER_HEAVY_FIRST = 0, 1, 2


class Ensemble:
    """Independent simulations advanced by batched launches (a pic_ensemble)."""

    def __init__(self, cfgs):
        self.lib = load_library()
        n = len(cfgs)
        arr = (Config * n)(*cfgs)
        h = C.c_void_p()
        check(self.lib.pic_ensemble_create(arr, n, C.byref(h)), "pic_ensemble_create")
        self.h = h
        self.members = []
        for i in range(n):
            m = self.lib.pic_ensemble_member(h, i)
            if not m:
                raise PicError(-1, "pic_ensemble_member", self.lib.pic_last_error().decode())
            self.members.append(Simulation(cfgs[i], handle=m))

    def __len__(self):
        return len(self.members)

    def close(self):
        if self.h:
            for m in self.members:
                m.h = None
            self.lib.pic_ensemble_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, n=1):
        check(self.lib.pic_ensemble_step(self.h, int(n)), "pic_ensemble_step")

    def step_graph(self, n=1):
        check(self.lib.pic_ensemble_step_graph(self.h, int(n)), "pic_ensemble_step_graph")

    def set_schedule(self, order, blocks_per_cu):
        check(self.lib.pic_ensemble_set_schedule(self.h, int(order), int(blocks_per_cu)), "pic_ensemble_set_schedule")

    def schedule(self):
        o, g = C.c_int32(), C.c_int32()
        check(self.lib.pic_ensemble_get_schedule(self.h, C.byref(o), C.byref(g)), "pic_ensemble_get_schedule")
        return o.value, g.value

    def autotune(self, trial_steps=2):
        ms, n = C.c_double(), C.c_int32()
        check(self.lib.pic_ensemble_autotune(self.h, int(trial_steps), C.byref(ms), C.byref(n)),
              "pic_ensemble_autotune")
        return ms.value, n.value
