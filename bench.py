"""Throughput of the PIC cycle (push + deposit + FDTD + shift) on one MI355X.

Workload (default): the two-species Kelvin-Helmholtz shear flow, 256^3 cells,
16 macro-particles per cell per species (536,870,912 particles), TSC shape,
Boris pusher, Yee solver, Esirkepov deposit, FP32, initial state generated on
the device. One step is one full PIC cycle. K steps are timed twice with
every input resident in HBM: eagerly with HIP events on the simulation's
stream around each launch (the dominant kernel's HBM roofline and the
push/shift/field breakdown come from this run), then replayed from the
captured step graphs; `value` is the faster of the two wall-clock rates.
Further objects on the same line: the differentiable form's adjoint/forward
time ratio and the batched ensemble against one-at-a-time stepping.

Metric: particle-steps/s = mobile particles * steps / wall time.
Algorithmic bytes: 55 B per particle-step (read x,p,w,cellIdx,mask = 31 B,
write x,p = 24 B) and 96 B per cell-step for the whole step; the fused
push/deposit kernel alone moves 55 B per particle plus 36 B per cell
(E, B read; J written).

The CPU baseline is the FP32 oracle in the reference CPU's order (C++/OpenMP,
the same per-particle arithmetic, the omp2b deposit order) timed on this host
on a bounded sample of the same physics (same species, shapes, ppc and
cell/time step on a smaller periodic box), 10 steps after 2 warm-ups.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

HBM_PEAK_GBS = 8000.0

# HBM bytes per launch of the fused push/deposit kernel measured with
# rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE (KiB per dispatch) on this
# workload, scaled by what the counters read for a 1 GiB device copy
# (FETCH_SIZE reads 0.5 GiB for it: factor 2.0; WRITE_SIZE 1 GiB: factor 1.0).
# The kernel's epilogue also stages the supercell shift (leaver records to
# the outbox, compaction of the supercell), and the TSC kernels' spilled
# registers go to scratch, so these bytes include both.
# Command: rocprofv3 --pmc FETCH_SIZE (then WRITE_SIZE) -- python3 bench.py
# --steps 1 --warmup 1 --no-cpu-baseline --no-adjoint --no-ensemble --calibrate-copy
MEASURED_TRAFFIC = {
    ("kelvin_helmholtz", (256, 256, 256)): {"fetch_kib": 896118.6, "write_kib": 1340060.5, "fetch_scale": 2.0,
                                            "write_scale": 1.0},
}
# Instruction counters of the fused push/deposit kernel on this workload,
# means per dispatch: rocprofv3 --pmc SQ_WAVES SQ_INSTS_VALU SQ_INSTS_LDS
# SQ_LDS_BANK_CONFLICT SQ_ACTIVE_INST_VALU SQ_BUSY_CYCLES SQ_WAIT_INST_LDS
# SQ_ACTIVE_INST_LDS GRBM_GUI_ACTIVE -- python3 bench.py --steps 1 --warmup 1
# --no-cpu-baseline --no-adjoint --no-ensemble. SQ counters are summed over
# the chip; GRBM_GUI_ACTIVE is summed over the 8 XCDs (cycles of one XCD x 8).
# Keyed by workload, grid and field absorber: with the PML the laser box runs
# the push kernel that also damps the weightings in the layers
# (kPushDepositPml).
MEASURED_INSTS = {
    ("kelvin_helmholtz", (256, 256, 256), "none"): {"valu": 826.64e6, "lds": 90.875e6, "lds_bank_conflict": 488.81e6,
                                                    "grbm_gui_active": 31.881e6},
    ("laser_plasma", (192, 1536, 192), "exponential"): {"valu": 495.28e6, "lds": 52.231e6,
                                                        "lds_bank_conflict": 197.88e6, "grbm_gui_active": 20.089e6},
    ("laser_plasma", (192, 1536, 192), "pml"): {"valu": 517.90e6, "lds": 52.208e6, "lds_bank_conflict": 197.81e6,
                                                "grbm_gui_active": 22.487e6},
}
# VALU issue peak: a wave64 VALU instruction occupies one of the 1,024 SIMDs
# (256 CUs x 4) for 4 cycles at the 2.4 GHz peak engine clock
VALU_SIMDS = 1024
VALU_CYCLES_PER_INST = 4
CLOCK_HZ = 2.4e9
CUS = 256
XCDS = 8
BYTES_PER_PARTICLE = 55.0
BYTES_PER_CELL_STEP = 96.0
BYTES_PER_CELL_PUSH = 36.0


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(setup_fn, grid, steps=10, warmup=2):
    """The oracle in the reference CPU's order (alpaka omp2b semantics: push and
    shift, then per supercell a float cache in reverse frame order flushed
    under a stride-2 checkerboard), FP32, OpenMP over this host's cores, on a
    bounded sample of the same physics (same species, shape, ppc, cell and
    time step on a smaller periodic box)."""
    from oracle import oracle as O
    from picongpu_amd import setups

    st = setup_fn(grid=grid)
    threads = O.set_threads(int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)))
    ora = O.Oracle(st, 32, 1)
    for sp in range(len(st.species)):
        cells, pos, mom, w = setups.host_particles(st, sp)
        ora.set_frames(sp, setups.bin_frames(st, sp, cells, pos, mom, w))
    ora.step(warmup)
    t0 = time.perf_counter()
    ora.step(steps)
    dt = time.perf_counter() - t0
    n = st.mobile_particles()
    return {"value": n * steps / dt, "unit": "particle-steps/s", "cores": threads, "kind": "port",
            "sample": f"{st.name} {grid[0]}x{grid[1]}x{grid[2]} cells, {n} particles, {steps} steps after "
                      f"{warmup} warm-up, FP32 oracle in the reference CPU order (omp2b), {threads} OpenMP threads "
                      f"on {cpu_model()}, {dt:.2f} s"}


def adjoint_timing(steps=200, every=20, grid=(64, 64, 64), ppc=8):
    """Differentiable form (cold plasma, CIC, Boris, Yee, Gaussian pulse as the
    initial field): wall time of the checkpointed forward and of the adjoint
    (segment recompute + reverse), objective = field energy in a 16^3 probe."""
    import torch

    from picongpu_amd import adjoint, setups

    st = setups.differentiable(grid=grid, ppc=ppc, steps=steps)
    cell, pos, mom, w = adjoint.cold_plasma(st)
    pic = adjoint.DifferentiablePIC(st, cell, pos, mom, w, steps, every)
    theta = torch.tensor([1.0, 0.0, 0.0], dtype=torch.float64, device="cuda")
    E0, B0 = adjoint.gaussian_pulse(st, theta, lam_cells=8.0, w0_cells=12.0, tau_cells=12.0,
                                    amp=adjoint.a0_field_unit(st))
    lo, hi = (40, 24, 24), (56, 40, 40)

    def grads(E, B):
        Ed = E.double().requires_grad_(True)
        Bd = B.double().requires_grad_(True)
        adjoint.probe_energy(st, Ed, Bd, lo, hi).backward()
        return Ed.grad, Bd.grad

    E, B = pic.forward_raw(E0, B0)  # warm-up (allocates checkpoints)
    gE, gB = grads(E, B)
    pic.backward_raw(gE, gB)  # warm-up (allocates the segment tape)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    E, B = pic.forward_raw(E0, B0)
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0
    gE, gB = grads(E, B)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    pic.backward_raw(gE, gB)
    torch.cuda.synchronize()
    t_bwd = time.perf_counter() - t0
    pic.close()
    n = len(w)
    return {"workload": f"differentiable {grid[0]}x{grid[1]}x{grid[2]} cold plasma {ppc} ppc, CIC, Boris, Yee, "
                        f"{steps} steps, checkpoints every {every}", "particles": n,
            "forward_ms": t_fwd * 1e3, "backward_ms": t_bwd * 1e3, "ratio": t_bwd / t_fwd,
            "forward_particle_steps_per_s": n * steps / t_fwd}


def ensemble_timing(steps=10, warmup=2):
    """Batched ensemble (64 two-stream boxes of 32..96 cells per axis, 2..16
    ppc, TSC, Boris, Yee, seed 7): particle-steps/s of the autotuned batched
    step replayed from graphs, against the same members advanced one at a time
    with their own graph-replayed pic_step (automatic push schedule), and one
    at a time with one push launch per checkerboard phase."""
    from picongpu_amd import abi, setups

    sts = setups.ensemble()
    ens = abi.Ensemble([st.config() for st in sts])
    for m, st in zip(ens.members, sts):
        setups.init_device(m, st)
    particles = sum(m.particle_count(0) for m in ens.members)
    tuned_ms, taken = ens.autotune(trial_steps=2)
    ens.step_graph(warmup)
    t0 = time.perf_counter()
    ens.step_graph(steps)
    t_batch = time.perf_counter() - t0
    for m in ens.members:
        m.step_graph(warmup)
    t0 = time.perf_counter()
    for m in ens.members:
        m.step_graph(steps)
    t_seq = time.perf_counter() - t0
    order, bpc = ens.schedule()
    ens.close()
    # the same members one at a time with one push launch per checkerboard
    # phase (the reference's per-colour launches; the automatic schedule
    # above runs small members' pushes in one launch)
    sims = []
    for st in sts:
        sim = abi.Simulation(st.config())
        sim.set_push_schedule(abi.PUSH_SCHEDULE_PHASES)
        setups.init_device(sim, st)
        sims.append(sim)
    for sim in sims:
        sim.step_graph(warmup)
    t0 = time.perf_counter()
    for sim in sims:
        sim.step_graph(steps)
    t_phases = time.perf_counter() - t0
    for sim in sims:
        sim.close()
    names = {0: "sim-major", 1: "interleaved", 2: "heavy-first"}
    return {"workload": f"{len(sts)} two-stream boxes (32..96 cells per axis, 2..16 ppc, TSC, Boris, Yee, seed 7)",
            "particles": particles, "cells": int(sum(int(np.prod(st.grid)) for st in sts)), "steps": steps,
            "batched_ms_per_step": t_batch * 1e3 / steps, "one_at_a_time_ms_per_step": t_seq * 1e3 / steps,
            "batched_particle_steps_per_s": particles * steps / t_batch,
            "one_at_a_time_particle_steps_per_s": particles * steps / t_seq, "speedup": t_seq / t_batch,
            "one_at_a_time_per_phase_ms_per_step": t_phases * 1e3 / steps,
            "speedup_vs_per_phase": t_phases / t_batch,
            "schedule": {"order": names.get(order, order), "blocks_per_cu": bpc, "autotune_ms_per_step": tuned_ms,
                         "autotune_steps": taken}}


def copy_calibration(nbytes=1 << 30):
    """One device-to-device copy of nbytes (a copy kernel of known size) so
    rocprofv3's FETCH_SIZE/WRITE_SIZE can be scaled to bytes."""
    import ctypes as C

    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(nbytes)) == 0
    assert hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="kelvin_helmholtz")
    ap.add_argument("--grid", default="", help="override grid, e.g. 128,128,128")
    ap.add_argument("--laser", default="initial", choices=["initial", "incident"],
                    help="laser_plasma only: the pulse as the initial field or injected through the Huygens surface")
    ap.add_argument("--absorber", default="pml", choices=["pml", "exponential"],
                    help="laser_plasma only: the field absorber on the open x and y borders")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-grid", default="128,64,64")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint/forward timing")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the batched ensemble timing")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run a 1 GiB device copy first (scale for rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported (the path shards by "
              f"spatial domain decomposition, which this build does not run)", file=sys.stderr)
        sys.exit(2)

    # torch first: its HIP device discovery must run before the library
    # initialises the runtime in this process (the adjoint harness uses torch)
    import torch

    torch.cuda.is_available()
    from picongpu_amd import abi, setups

    if args.calibrate_copy:
        copy_calibration()
    setup_fn = setups.CONFIGS[args.config]
    kw = {}
    if args.grid:
        kw["grid"] = tuple(int(v) for v in args.grid.split(","))
    if args.config == "laser_plasma":
        kw["laser"] = args.laser
        kw["absorber"] = args.absorber
    st = setup_fn(**kw)
    n_particles = st.mobile_particles()
    ncells = st.ncells

    sim = abi.Simulation(st.config())
    setups.init_device(sim, st)
    sim.sync()
    n_particles = sum(sim.particle_count(i) for i, s in enumerate(st.species) if s.pusher != abi.PUSHER_NONE)
    for _ in range(args.warmup):
        sim.step(1)
    sim.sync()

    sim.timing(True)
    sim.sync()
    t0 = time.perf_counter()
    sim.step(args.steps)
    sim.sync()
    wall = time.perf_counter() - t0
    push_ms, push_launches = sim.kernel_timer(abi.TIMER_PUSH)
    shift_ms, _ = sim.kernel_timer(abi.TIMER_SHIFT)
    field_ms, _ = sim.kernel_timer(abi.TIMER_FIELD)
    sim.timing(False)
    # the same steps replayed from the captured graphs (one per E buffer parity)
    sim.step_graph(2)
    sim.sync()
    t0 = time.perf_counter()
    sim.step_graph(args.steps)
    sim.sync()
    wall_graph = time.perf_counter() - t0
    counts = [sim.particle_count(i) for i in range(len(st.species))]
    sim.close()

    eager_ms = wall * 1e3 / args.steps
    graph_ms = wall_graph * 1e3 / args.steps
    path = "graph" if wall_graph < wall else "eager"
    best = min(wall, wall_graph)
    ms_per_step = best * 1e3 / args.steps
    value = n_particles * args.steps / best
    push_bytes = (n_particles * BYTES_PER_PARTICLE + ncells * BYTES_PER_CELL_PUSH) * args.steps
    achieved = push_bytes / (push_ms * 1e-3) / 1e9
    step_bytes = (n_particles * BYTES_PER_PARTICLE + ncells * BYTES_PER_CELL_STEP) * args.steps
    step_gbs = step_bytes / wall / 1e9

    adj = None
    if not args.no_adjoint:
        adj = adjoint_timing()
    ens = None
    if not args.no_ensemble:
        ens = ensemble_timing()

    baseline = None
    if not args.no_cpu_baseline:
        try:
            baseline = cpu_baseline(setup_fn, tuple(int(v) for v in args.cpu_grid.split(",")))
        except Exception as exc:  # the baseline is reported, never required
            baseline = {"error": str(exc)}

    traffic = None
    t = MEASURED_TRAFFIC.get((st.name, tuple(st.grid)))
    if t is not None:
        traffic = (t["fetch_kib"] * t["fetch_scale"] + t["write_kib"] * t["write_scale"]) * 1024.0
    avg_launch_ms = push_ms / max(push_launches, 1)
    compute = {"valu_frac": None}
    absorber_name = ({0: "none", 1: "exponential", 2: "pml"}.get(st.absorber)
                     if any(b == abi.BOUNDARY_OPEN for b in st.boundary) else "none")
    mi = MEASURED_INSTS.get((st.name, tuple(st.grid), absorber_name))
    if mi is not None:
        particles_per_launch = n_particles * args.steps / max(push_launches, 1)
        compute = {
            # measured VALU instructions per launch at the VALU issue peak,
            # over the launch time measured now
            "valu_frac": mi["valu"] * VALU_CYCLES_PER_INST / (VALU_SIMDS * CLOCK_HZ * avg_launch_ms * 1e-3),
            "valu_per_64_particles": mi["valu"] / (particles_per_launch / 64.0),
            "lds_per_64_particles": mi["lds"] / (particles_per_launch / 64.0),
            # cycles with an LDS bank conflict per CU over the profiled launch's cycles
            "lds_conflict_frac": mi["lds_bank_conflict"] / CUS / (mi["grbm_gui_active"] / XCDS),
            "valu_basis": "SQ_INSTS_VALU x 4 cycles / (1024 SIMDs x 2.4 GHz x launch time)",
        }

    out = {
        "metric": "particle-steps/s (push+deposit+FDTD) on 1 GPU; adjoint/forward time ratio",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (device-seeded initial state)",
        "config": {"workload": st.name + ("+incident" if st.incident else ""), "grid": list(st.grid), "species": [s.name for s in st.species],
                   "ppc": [s.ppc for s in st.species], "shape": "TSC" if st.species[0].shape == 2 else
                   ("CIC" if st.species[0].shape == 1 else "PQS"),
                   "pusher": {1: "Boris", 2: "Vay", 3: "HigueraCary"}.get(st.species[0].pusher, "none"),
                   "solver": "Yee" if st.solver == 0 else "Lehe", "particles": n_particles,
                   "cells": ncells,
                   "field_absorber": absorber_name},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "kPushDeposit (fused gather+push+move+Esirkepov)",
                     "avg_launch_ms": avg_launch_ms, "launches": push_launches, **compute},
        "step_roofline": {"achieved": step_gbs, "frac": step_gbs / HBM_PEAK_GBS,
                          "push_ms": push_ms / args.steps, "shift_ms": shift_ms / args.steps,
                          "field_ms": field_ms / args.steps},
        "step_path": {"timed": path, "eager_ms_per_step": eager_ms, "graph_ms_per_step": graph_ms},
        "particles_after": counts,
        "adjoint": adj,
        "ensemble": ens,
        "cpu_baseline": baseline,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
