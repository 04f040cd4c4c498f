"""Benchmark: particle-steps/s of the per-step SPH pipeline on one MI355X.

Workload: the 3D Sedov blast, 64^3 = 262,144 particles in the periodic unit
box (Wendland C6, 100 neighbours, FP32 pair sums with FP64 integrator, keys
and tree), one full driver iteration per step: leapfrog kick/drift, periodic
copies, range/keys/sort, octree, neighbour tiles, density/EOS/signal speed,
acceleration, du/dt, smoothing length and time step.  Inputs are resident in
HBM before the timed region; WALL copies are not counted in N.

Prints one JSON line (see the keys below).  --gpus N > 1 is not supported:
this build runs one GPU.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, AMD published

# Algorithmic HBM bytes per particle for the single-kernel phases, FP32
# derived fields (4 B) and FP64 state (8 B): each field the kernel must read
# or write, once.  Candidate gathers beyond the particle's own record,
# tiles and masks are implementation traffic, not counted.
KERNEL_BYTES = {
    # x,y,z,h,m,u read (48) ; rho,P,c written (12)
    "density": 60,
    # x,y,z,vx,vy,vz read (48); c, c_old read (8); vsig written (4)
    "vsig": 60,
    # x,y,z,vhx,vhy,vhz,h,m read (64); rho,P,c read (12); a, g, phi (28)
    "accel": 104,
    # x,y,z,vh,v,h,m,u read (88); rho,P,c read (12); dudt,P,c written (12)
    "dudt": 112,
    # x,y,z,h read (32); candidate ids written (4 B per candidate ~ not
    # counted); node boxes are tree traffic
    "traverse": 32,
    # x,y,z,h read (32); mask bits written (not counted)
    "masks": 32,
}
# Compulsory particle-state traffic of one step (bytes per particle), the
# B_hydro / B_grav of the benchmark definition.
B_HYDRO = 864
B_GRAV = 968


def workload(name):
    """(particles, params, label, step bytes) of a named workload."""
    from flecsph_amd import generators
    if name == "sedov64":
        parts, prm = generators.sedov(n_side=64)
        return parts, prm, "sedov3d_64^3_periodic", B_HYDRO
    if name == "polytrope1m":
        parts, prm = generators.polytrope(n=1_000_000)
        return parts, prm, "polytrope_n1_1M_fmm_theta0.5_quad_leaf32", B_GRAV
    raise SystemExit(f"unknown workload {name}")


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="sedov64",
                    choices=["sedov64", "polytrope1m"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--cpu-steps", type=int, default=3)
    return ap.parse_args()


def cpu_baseline(name, threads, steps):
    """The FP64 CPU oracle (a port of the reference path, OpenMP over
    groups and FMM sinks) on the same workload: one initial iteration and
    one warm step untimed, then `steps` timed steps."""
    from flecsph_amd import _abi
    from oracle import oracle
    parts, prm, label, _ = workload(name)
    prm["precision"] = _abi.PREC_FP64
    s = oracle.OracleSystem(_abi.default_params(**prm), threads=threads)
    s.upload(parts)
    s.step()
    s.step()
    t0 = time.perf_counter()
    for _ in range(steps):
        s.step()
    dt = time.perf_counter() - t0
    n_real = s.count()[1]
    return dict(value=n_real * steps / dt, unit="particle-steps/s",
                cores=threads, kind="port",
                sample=f"{label}, {steps} timed steps after 2 untimed, FP64 "
                       f"oracle, {threads} OpenMP threads, {dt:.1f} s")


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; "
                          "--gpus must be 1"}))
        sys.exit(2)
    from flecsph_amd import BodySystem, _abi

    parts, prm, label, step_bytes = workload(args.workload)
    g = BodySystem(_abi.default_params(**prm), device=0)
    g.upload(parts)
    for _ in range(args.warmup):
        g.step()
    g.enable_timing(True)
    n_tot, n_real = g.count()
    phase = {k: 0.0 for k in _abi.TIMERS}
    n_tot_acc = 0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        g.step()  # returns after the step's work has completed
        for k, v in g.phase_times().items():
            phase[k] += v
        n_tot_acc += g.count()[0]
    elapsed = time.perf_counter() - t0
    g.enable_timing(False)
    K = args.steps
    value = n_real * K / elapsed
    ms_step = elapsed / K * 1e3
    avg_ntot = n_tot_acc / K
    per_phase = {k: v / K for k, v in phase.items()}
    single = {k: per_phase[k] for k in KERNEL_BYTES}
    dom = max(single, key=single.get)
    achieved = KERNEL_BYTES[dom] * avg_ntot / (single[dom] * 1e-3) / 1e9
    it, dt, tt = g.time()
    out = {
        "metric": "particle-steps/s",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": label,
            "n_particles": n_real,
            "n_with_periodic_copies": round(avg_ntot),
            "kernel": "wendland_c6", "n_ngb": 100,
            "arithmetic": "FP32 pair sums (compensated), FP64 keys/tree/"
                          "integrator/reductions",
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dom,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
        },
        "step_hbm_frac": step_bytes * value / (HBM_PEAK_GBS * 1e9),
        "phase_ms": {k: round(v, 4) for k, v in per_phase.items()},
        "sim_state": {"iteration": it, "dt": dt, "time": tt},
    }
    if not args.no_cpu_baseline:
        thr = args.cpu_threads or min(16, os.cpu_count() or 1)
        out["cpu_baseline"] = cpu_baseline(args.workload, thr, args.cpu_steps)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
