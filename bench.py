"""Benchmark: particle-steps/s of the per-step SPH pipeline on one MI355X.

Headline workload (default): the self-gravitating n=1 polytrope, 1,000,000
particles (R = 1, M = 1, G = 1, K = 1, gamma = 2, Wendland C6 with 100
neighbours), FMM self-gravity with opening angle 0.5, quadrupole sources and
buckets of 32, FP32 pair and multipole arithmetic with FP64 positions,
integrator, keys and tree; one full driver iteration per step (leapfrog,
range/keys/sort, octree, neighbour tiles, density/EOS/signal speed,
acceleration, gravity, du/dt, smoothing length, time step).  Inputs are
resident in HBM before the timed region.  Before the warm-up the library's
autotuner picks the pair kernels' tile width and LDS chunk (untimed; it
times every candidate from the same saved state and restores it; neighbour
lists do not depend on the choice), reported in config.tuning.

The roofline object is that of the step's dominant single kernel (for the
star: the P2P/local-expansion kernel of the FMM), from its HIP-event time in
further steps: algorithmic bytes per particle x particles per launch over
the launch time, against 8 TB/s; `valu` gives the same kernel against the
FP32 vector peak.

Side measurements in the same JSON line:
  "sedov64": the periodic 3D Sedov blast, 64^3 = 262,144 particles;
  "learned_h": Newton-h against the learned smoothing length on the cold
  collapse of an FCC-lattice sphere with self-gravity (default 4M particles,
  the first --lh-learn steps train the predictor; times are averaged over
  the steps after training): h-phase and whole-step speedups and the
  neighbour-count error of both;
  "batched": 64 instances with log-uniform particle counts in [8k, 200k],
  each a periodic Sedov blast or an open Sod tube by a seeded coin, in one
  batched context (autotuned tile width and LDS chunk, graph replay of the
  step tail): aggregate particle-steps/s and p50/p99 step latency;
  "cpu_baseline": the FP64 CPU oracle on the headline workload with the
  GPU box's CPU share of threads, plus a single-threaded leg on a bounded
  sample, the host's CPU model and core counts.

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
# tiles, masks, FMM cells and interaction lists are implementation traffic,
# not counted.
KERNEL_BYTES = {
    # x,y,z,h,m,u read (48) ; rho,P,c written (12)
    "density": 60,
    # x,y,z,vx,vy,vz read (48); c, c_old read (8); vsig written (4)
    "vsig": 60,
    # x,y,z,vhx,vhy,vhz,h,m read (64); rho,P,c read (12); a, g, phi (28);
    # with the signal speed folded in: v (24), c_old (4) read, vsig (4)
    "accel": 136,
    # x,y,z,vh,v,h,m,u read (88); rho,P,c read (12); dudt,P,c written (12)
    "dudt": 112,
    # x,y,z,h read (32); candidate ids written (4 B per candidate ~ not
    # counted); node boxes are tree traffic
    "traverse": 32,
    # x,y,z,h read (32); mask bits written (not counted)
    "masks": 32,
    # FMM local-expansion evaluation + direct sums: x,y,z,m read (32);
    # gx,gy,gz,phi written (16)
    "p2p": 48,
}
# Compulsory particle-state traffic of one step (bytes per particle), the
# B_hydro / B_grav of the benchmark definition.
B_HYDRO = 864
B_GRAV = 968
# the star does not evolve u (polytropic EOS, evolve_internal_energy = no):
# no du/dt pass (100 B) and no u kick (20 B)
B_GRAV_NO_U = B_GRAV - 120

# FP32 operations per neighbour pair of the pair kernels and per
# particle-particle interaction of P2P, counted from flecsph_amd/csrc/
# k_sph.hip and k_grav.hip (Wendland C6; FMA = 2, rcp/rsqrt = 1; position
# differences and the compensated sums included; the driver's acceleration
# pass includes the signal speed).
PAIR_FLOPS = {"density": 32, "vsig": 20, "accel": 89, "dudt": 69, "p2p": 20}
VALU_FP32_PEAK_TFLOPS = 157.3  # MI355X FP32 vector, AMD published

# HBM bytes per launch, (workload, kernel) -> ((tile_rows, chunk), bytes),
# from rocprofv3 FETCH_SIZE and WRITE_SIZE runs (KiB per dispatch), each
# scaled by what the counter reads for a plain 256 MiB device copy
# (FETCH_SIZE x 2.00, WRITE_SIZE x 1.00); the profiles/ summaries have the
# commands.  A run with another tuning reports no traffic (None: the kernel
# does not depend on the tuning).
TRAFFIC = {
    ("sedov64", "accel"): ((64, 64), 637_743_869),
    ("sedov64", "masks"): ((64, 64), 223_822_856),
    ("polytrope1m", "p2p"): (None, 389_688_786),
}

# Side workloads measured after the headline one.
SIDE_WORKLOADS = ("sedov64",)
# Threads of the CPU baseline: the GPU box's CPU share per GPU.
BOX_CPU_SHARE = 16
# Single-threaded CPU leg: a bounded sample of the headline workload.
CPU_1T_SAMPLE = {"polytrope1m": "polytrope125k", "sedov64": "sedov32"}


def workload(name):
    """(particles, params, label, step bytes) of a named workload."""
    from flecsph_amd import generators
    if name == "sedov64":
        parts, prm = generators.sedov(n_side=64)
        return parts, prm, "sedov3d_64^3_periodic", B_HYDRO
    if name == "sedov32":
        parts, prm = generators.sedov(n_side=32)
        return parts, prm, "sedov3d_32^3_periodic", B_HYDRO
    if name == "polytrope1m":
        parts, prm = generators.polytrope(n=1_000_000)
        return (parts, prm, "polytrope_n1_1M_fmm_theta0.5_quad_leaf32",
                B_GRAV_NO_U)
    if name == "polytrope125k":
        parts, prm = generators.polytrope(n=125_000)
        return (parts, prm, "polytrope_n1_125k_fmm_theta0.5_quad_leaf32",
                B_GRAV_NO_U)
    raise SystemExit(f"unknown workload {name}")


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="polytrope1m",
                    choices=["sedov64", "polytrope1m", "collapse4m_newton",
                             "collapse4m_learned", "sedov_batch64"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--copy-probe", action="store_true",
                    help="only three 256 MiB device-to-device copies (the "
                    "reference kernel for scaling HBM counters)")
    ap.add_argument("--no-side", action="store_true",
                    help="skip the side workloads (Sedov 64^3)")
    ap.add_argument("--no-learned-h", action="store_true")
    ap.add_argument("--lh-n", type=int, default=4_000_000)
    ap.add_argument("--lh-steps", type=int, default=200)
    ap.add_argument("--lh-learn", type=int, default=50)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep tile_rows 64 and chunk 128 instead of "
                         "autotuning them before the warm-up")
    ap.add_argument("--tune-steps", type=int, default=3)
    ap.add_argument("--tuning", default="",
                    help="TILE_ROWS,CHUNK: use these instead of autotuning "
                         "(profiling runs use the bench line's choice)")
    ap.add_argument("--batch-inst", type=int, default=64)
    ap.add_argument("--batch-steps", type=int, default=10)
    return ap.parse_args()


def log(msg):
    print(msg, file=sys.stderr, flush=True)


# Tuning (tile rows, chunk) of both runs of the learned-h comparison (the
# autotuner is not used there: the predictor would keep learning during its
# steps); the fastest of the autotuner's candidates on this workload.  The
# grouping stays the reference's (group_max 128).
LH_TUNING = (64, 64)


def learned_h_comparison(n, steps, learn):
    """Newton-h and learned-h runs of the same collapse; per-step h-phase
    (HIP events) and wall times averaged over the steps after training."""
    import numpy as np
    from flecsph_amd import BodySystem, _abi, generators, set_tuning
    parts, prm = generators.collapse(n=n)
    n_real = len(parts["x"])
    out = {}
    for name, mode in (("newton", _abi.H_NEWTON), ("learned", _abi.H_LEARNED)):
        g = BodySystem(_abi.default_params(
            **dict(prm, h_mode=mode, learn_steps=learn)))
        set_tuning(g, *LH_TUNING)
        g.upload(parts)
        g.enable_timing(True)
        hms, sms = [], []
        for k in range(steps):
            t0 = time.perf_counter()
            g.step()
            t1 = time.perf_counter()
            if k >= learn:
                hms.append(g.phase_times()["h"])
                sms.append((t1 - t0) * 1e3)
            if k % 20 == 19:
                log(f"learned-h comparison: {name} step {k + 1}/{steps}")
        ngb = g.get("hngb")[:]
        ids = g.get("id")
        real = ids < n_real
        it = g.get("hiter")[real]
        rec = {"h_ms": float(np.mean(hms)), "step_ms": float(np.mean(sms)),
               "ngb_abs_err": float(np.mean(np.abs(ngb[real] - prm["n_target"]))),
               "iterations": float(it.mean())}
        if mode == _abi.H_LEARNED:
            st = g.h_stats()
            rec.update(fallback_last=st["fallback"], clamped_last=st["clamped"],
                       train_loss=st["last_loss"])
        out[name] = rec
        g.close()
    nw, le = out["newton"], out["learned"]
    return {
        "workload": f"collapse_fcc_sphere_{n_real}_fmm_theta0.5_quad_leaf32",
        "steps": steps, "learn_steps": learn,
        "timed_steps": steps - learn,
        "tuning": {"group_max": _abi.default_params(**prm).group_max,
                   "tile_rows": LH_TUNING[0], "chunk": LH_TUNING[1]},
        "newton_h_ms": nw["h_ms"], "learned_h_ms": le["h_ms"],
        "h_phase_speedup": nw["h_ms"] / le["h_ms"],
        "newton_step_ms": nw["step_ms"], "learned_step_ms": le["step_ms"],
        "step_speedup": nw["step_ms"] / le["step_ms"],
        "ngb_abs_err_newton": nw["ngb_abs_err"],
        "ngb_abs_err_learned": le["ngb_abs_err"],
        "iterations_newton": nw["iterations"],
        "iterations_learned": le["iterations"],
        "fallback_last_step": le["fallback_last"],
        "train_loss": le["train_loss"],
    }


def batched_run(n_inst, steps):
    """Batched Sedov workload: graph replay, autotune (of the replayed
    steps), then `steps` timed steps."""
    import numpy as np
    from flecsph_amd import BatchSystem, _abi, autotune, generators
    sets, prm, periodic = generators.mixed_batch(n_inst)
    N = sum(len(p["x"]) for p in sets)
    b = BatchSystem(_abi.default_params(**prm), n_inst)
    b.upload(sets)
    for i, per in enumerate(periodic):
        b.instance_set_periodic(i, per)
    b.step()
    b.enable_graph(True)  # the tuner times the replayed steps it picks for
    gm, ch = autotune(b, steps=2)
    log(f"batched: autotune chose tile_rows {gm}, chunk {ch}")
    b.step()
    lat = []
    for _ in range(steps):
        t0 = time.perf_counter()
        b.step()
        lat.append((time.perf_counter() - t0) * 1e3)
    b.close()
    lat = np.array(lat)
    return {
        "workload": f"sedov_sod_x{n_inst}_loguniform_8k_200k",
        "instances": n_inst, "blasts": int(sum(periodic)),
        "tubes": int(n_inst - sum(periodic)), "particles": N, "steps": steps,
        "value": N / (lat.mean() * 1e-3), "unit": "particle-steps/s",
        "ms_per_step": float(lat.mean()),
        "p50_step_ms": float(np.percentile(lat, 50)),
        "p99_step_ms": float(np.percentile(lat, 99)),
        "group_max": _abi.default_params(**prm).group_max, "tile_rows": gm,
        "chunk": ch, "graph": True,
    }


def cpu_baseline(name, threads, steps):
    """The FP64 CPU oracle (a port of the reference path, OpenMP over
    groups and FMM sinks) on a workload: the initial iteration untimed, then
    `steps` timed steps."""
    from flecsph_amd import _abi
    from oracle import oracle
    parts, prm, label, _ = workload(name)
    prm["precision"] = _abi.PREC_FP64
    s = oracle.OracleSystem(_abi.default_params(**prm), threads=threads)
    s.upload(parts)
    s.step()
    dt = 0.0
    for k in range(steps):
        t0 = time.perf_counter()
        s.step()
        dt += time.perf_counter() - t0
        log(f"CPU baseline {label}: step {k + 1}/{steps} "
            f"({time.perf_counter() - t0:.1f} s)")
    n_real = s.count()[1]
    s.close()
    return dict(value=n_real * steps / dt, unit="particle-steps/s",
                cores=threads, kind="port",
                sample=f"{label}, {steps} timed steps after the initial "
                       f"iteration, FP64 oracle, {threads} OpenMP "
                       f"thread{'s' if threads > 1 else ''}, {dt:.1f} s")


def side_workload(args):
    """The learned-h, Newton-h and batched workloads on their own: warmup
    steps, then `steps` timed steps; one JSON line (used for profiling)."""
    from flecsph_amd import BatchSystem, BodySystem, _abi, generators
    name = args.workload
    if name == "sedov_batch64":
        sets, prm, periodic = generators.mixed_batch(64)
        s = BatchSystem(_abi.default_params(**prm), 64)
        s.upload(sets)
        for i, per in enumerate(periodic):
            s.instance_set_periodic(i, per)
        n = sum(len(p["x"]) for p in sets)
        label = "sedov_sod_x64_loguniform_8k_200k_graph"
    else:
        parts, prm = generators.collapse(n=4_000_000)
        mode = _abi.H_LEARNED if name.endswith("learned") else _abi.H_NEWTON
        s = BodySystem(_abi.default_params(
            **dict(prm, h_mode=mode, learn_steps=args.lh_learn)))
        s.upload(parts)
        n = len(parts["x"])
        label = f"collapse_fcc_sphere_{n}_{name.split('_')[1]}_h"
    if args.tuning:
        from flecsph_amd import set_tuning
        set_tuning(s, *(int(v) for v in args.tuning.split(",")))
    for k in range(args.warmup):
        s.step()
        if name == "sedov_batch64" and k == 0:
            s.enable_graph(True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        s.step()
    el = time.perf_counter() - t0
    print(json.dumps({"metric": "particle-steps/s", "value": n * args.steps / el,
                      "unit": "particle-steps/s", "n_gpus": 1,
                      "steps": args.steps, "warmup": args.warmup,
                      "ms_per_step": el / args.steps * 1e3,
                      "higher_is_better": True, "data": "synthetic",
                      "config": {"workload": label}}))


def copy_probe():
    """Three plain 256 MiB device-to-device copies by the HIP runtime's copy
    kernel: the known-size reference for FETCH_SIZE / WRITE_SIZE."""
    import ctypes as C
    lib = C.CDLL("libamdhip64.so")
    nbytes = 256 << 20
    a, b = C.c_void_p(), C.c_void_p()
    for p in (a, b):
        if lib.hipMalloc(C.byref(p), C.c_size_t(nbytes)) != 0:
            raise SystemExit("hipMalloc failed")
    if lib.hipMemset(a, 1, C.c_size_t(nbytes)) != 0:
        raise SystemExit("hipMemset failed")
    for _ in range(3):
        if lib.hipMemcpy(b, a, C.c_size_t(nbytes), 3) != 0:  # device->device
            raise SystemExit("hipMemcpy failed")
    lib.hipDeviceSynchronize()
    for p in (a, b):
        lib.hipFree(p)
    print(json.dumps({"copy_bytes": nbytes, "copies": 3}))


def host_info():
    """CPU model and core counts of this host (from /proc/cpuinfo)."""
    model, cores = "unknown", set()
    phys = core = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                k, _, v = line.partition(":")
                k, v = k.strip(), v.strip()
                if k == "model name" and model == "unknown":
                    model = v
                elif k == "physical id":
                    phys = v
                elif k == "core id":
                    core = v
                elif not k and phys is not None and core is not None:
                    cores.add((phys, core))
                    phys = core = None
        if phys is not None and core is not None:
            cores.add((phys, core))
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return {"cpu_model": model, "logical_cpus": os.cpu_count(),
            "physical_cores": len(cores) or None, "usable_cpus": usable}


def run_single(name, args, tuning_arg="", autotune_on=True):
    """One single-simulation workload: optional autotune (untimed), warm-up,
    `args.steps` timed whole driver iterations, then per-kernel HIP-event
    times from further steps.  Returns the measurement record."""
    from flecsph_amd import BodySystem, _abi, autotune

    parts, prm, label, step_bytes = workload(name)
    g = BodySystem(_abi.default_params(**prm), device=0)
    g.upload(parts)
    tuning = {"group_max": prm.get("group_max", 128), "tile_rows": 64,
              "chunk": 128}
    if tuning_arg:
        from flecsph_amd import set_tuning
        tr, ch = (int(v) for v in tuning_arg.split(","))
        set_tuning(g, tr, ch)
        tuning.update(tile_rows=tr, chunk=ch, given=True)
    elif autotune_on:
        # untimed: each (tile_rows, chunk) candidate is timed over
        # tune_steps steps from the state after the initial iteration, which
        # the autotuner restores afterwards (the warm-up then continues from
        # there)
        tr, ch = autotune(g, steps=args.tune_steps)
        tuning.update(tile_rows=tr, chunk=ch,
                      tune_steps_per_candidate=args.tune_steps)
        log(f"{name}: autotune chose tile_rows {tr}, chunk {ch}")
    for _ in range(args.warmup):
        g.step()
    n_tot, n_real = g.count()
    n_tot_acc = 0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        g.step()  # returns after the step's work has completed
        n_tot_acc += g.count()[0]
    elapsed = time.perf_counter() - t0
    K = args.steps
    value = n_real * K / elapsed
    avg_ntot = n_tot_acc / K
    log(f"{name}: {K} timed steps, {elapsed / K * 1e3:.3f} ms/step")
    # phase times (HIP events on the stream) from further steps, so the
    # events do not sit in the timed region
    g.enable_timing(True)
    phase = {k: 0.0 for k in _abi.TIMERS}
    n_phase = max(3, min(10, K))
    gstats = None
    for _ in range(n_phase):
        g.step()
        for k, v in g.phase_times().items():
            phase[k] += v
    g.enable_timing(False)
    if prm.get("gravity"):
        gstats = g.gravity_stats()
    per_phase = {k: v / n_phase for k, v in phase.items()}
    single = {k: per_phase[k] for k in KERNEL_BYTES if per_phase[k] > 0}
    dom = max(single, key=single.get)
    achieved = KERNEL_BYTES[dom] * avg_ntot / (single[dom] * 1e-3) / 1e9
    traffic = None
    meas = TRAFFIC.get((name, dom))
    # the FMM kernels do not depend on the pair kernels' tuning
    if meas and (meas[0] is None or
                 (tuning["tile_rows"], tuning["chunk"]) == meas[0]):
        traffic = meas[1]
    # work of the dominant kernel: neighbour pairs (one mask bit per (a, b))
    # for the SPH passes, particle-particle interactions for P2P
    if dom == "p2p":
        work = float(gstats["p2p_interactions"])
    else:
        work = float(g.get("nngb").sum())
    valu = None
    if dom in PAIR_FLOPS:
        tf = PAIR_FLOPS[dom] * work / (single[dom] * 1e-3) / 1e12
        valu = {"kernel": dom, "flops_per_pair": PAIR_FLOPS[dom],
                "pairs": work, "achieved": tf,
                "peak": VALU_FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": tf / VALU_FP32_PEAK_TFLOPS}
    it, dt, tt = g.time()
    g.close()
    rec = {
        "value": value,
        "ms_per_step": elapsed / K * 1e3,
        "config": {
            "workload": label,
            "n_particles": n_real,
            "kernel": "wendland_c6", "n_ngb": 100,
            "arithmetic": "FP32 pair sums (compensated), FP64 keys/tree/"
                          "integrator/reductions",
            "tuning": tuning,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dom,
            "bytes_per_particle": KERNEL_BYTES[dom],
            "particles_per_launch": round(avg_ntot),
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            # measured HBM bytes per launch over the same duration
            "traffic_gbs": (traffic / (single[dom] * 1e-3) / 1e9
                            if traffic else None),
            "kernel_ms": single[dom],
        },
        "valu": valu,
        "step_hbm_frac": step_bytes * value / (HBM_PEAK_GBS * 1e9),
        "phase_ms": {k: round(v, 4) for k, v in per_phase.items()},
        "sim_state": {"iteration": it, "dt": dt, "time": tt},
    }
    if n_tot != n_real:
        rec["config"]["n_with_periodic_copies"] = round(avg_ntot)
    if prm.get("gravity"):
        rec["config"].update(fmm={"theta": prm["theta"],
                                  "leaf_size": prm["leaf_size"],
                                  "order": "quadrupole"})
        rec["gravity_stats"] = gstats
        # M2L runs on FP32 VALU inside the list walk: each pair's operator
        # depends on its own offset (centre-of-mass expansions on an
        # adaptive octree), so a batch of pairs is a set of matrix-vector
        # products with no operand reuse, which matrix cores cannot speed up
        rec["m2l"] = {"engine": "fp32_valu", "pairs": gstats["m2l_pairs"],
                      "mfma_utilisation": 0.0}
    return rec


def cpu_legs(name, threads):
    """The CPU baseline: the FP64 oracle on the headline workload with the
    box's CPU share, and single-threaded on a bounded sample."""
    legs = {}
    log(f"CPU baseline: {name}, {threads} threads ...")
    legs["threads"] = cpu_baseline(name, threads, 3)
    sample = CPU_1T_SAMPLE.get(name)
    if sample:
        log(f"CPU baseline: {sample}, 1 thread ...")
        legs["single"] = cpu_baseline(sample, 1, 2)
    return legs


def main():
    args = parse()
    if args.copy_probe:
        copy_probe()
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; "
                          "--gpus must be 1"}))
        sys.exit(2)
    if args.workload in ("collapse4m_newton", "collapse4m_learned",
                         "sedov_batch64"):
        side_workload(args)
        return
    head = run_single(args.workload, args, args.tuning, not args.no_autotune)
    out = {
        "metric": "particle-steps/s",
        "value": head["value"],
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
    }
    out.update({k: v for k, v in head.items()
                if k not in ("value", "ms_per_step")})
    if not args.no_side:
        for side in SIDE_WORKLOADS:
            if side == args.workload:
                continue
            log(f"side workload {side} ...")
            r = run_single(side, args, "", not args.no_autotune)
            out[side] = {"value": r["value"], "unit": "particle-steps/s",
                         "ms_per_step": r["ms_per_step"],
                         "config": r["config"], "roofline": r["roofline"],
                         "valu": r["valu"], "phase_ms": r["phase_ms"]}
    if not args.no_learned_h:
        log("learned-h comparison ...")
        out["learned_h"] = learned_h_comparison(args.lh_n, args.lh_steps,
                                                args.lh_learn)
    if not args.no_batch:
        log("batched run ...")
        out["batched"] = batched_run(args.batch_inst, args.batch_steps)
    if not args.no_cpu_baseline:
        host = host_info()
        thr = args.cpu_threads or min(BOX_CPU_SHARE, host["usable_cpus"])
        legs = cpu_legs(args.workload, thr)
        base = legs["threads"]
        base["host"] = host
        if "single" in legs:
            base["single_thread"] = legs["single"]
        if host["physical_cores"]:
            # linear scaling of the measured per-thread rate to every
            # physical core of the host: an upper bound that favours the CPU
            est = base["value"] / thr * host["physical_cores"]
            base["all_physical_cores_linear_estimate"] = est
            base["gpu_over_all_cores_estimate"] = out["value"] / est
        base["gpu_over_cpu"] = out["value"] / base["value"]
        out["cpu_baseline"] = base
    print(json.dumps(out))


if __name__ == "__main__":
    main()
