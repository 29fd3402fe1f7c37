"""Benchmark: bead updates/s on the HEG N=256, M=128 FP64 workload.

A step is one sweep of every replica: N*M/2^4 = 2048 bisection attempts (15 bead
updates each) and N = 256 displace attempts (M = 128 bead updates each), then
the adjoint dS/dr of every bead and the device estimators (thermal energy,
S(k), g(r)).  `value` is the whole-job throughput of R independent replicas
advanced together (batched+adjoint), inputs resident in HBM.  The plain method
(one chain, one attempt at a time through the same kernels) is timed beside it,
and the CPU oracle on a bounded sample of the same workload is the CPU baseline.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "bead updates/s, HEG N=256 M=128 FP64, batched+adjoint vs plain GPU kernel"
HBM_PEAK_GBPS = 8000.0
MEASURED_RHO_ALG_BYTES = 1786919139.56  # algorithmic bytes per k_rho_loop launch, default workload
MEASURED_RHO_TRAFFIC_BYTES = 1841.70e6  # HBM bytes per k_rho_loop launch, same workload (rocprofv3 --pmc)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--replicas", type=int, default=1024)
    ap.add_argument("--graph", action="store_true", help="replay captured hipGraphs of 16 rounds (about 2 %% "
                    "faster here; rocprofv3 kernel traces of graph replays are not taken, so the default is eager)")
    ap.add_argument("--plain-attempts", type=int, default=1024)
    ap.add_argument("--cpu-sweeps", type=int, default=1,
                    help="whole sweeps of the CPU oracle's moves on all host cores (the estimators run once)")
    ap.add_argument("--cpu-one-thread-seconds", type=float, default=4.0)
    ap.add_argument("--cpu-only", action="store_true", help="only the CPU baseline (no GPU work)")
    ap.add_argument("--no-autotune", dest="autotune", action="store_false",
                    help="keep the engine's default launch shape instead of autotuning it before the warm-up")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-timing", action="store_true", help="no kernel clocks or events (for counter-collection runs)")
    ap.add_argument("--kernel-events", action="store_true",
                    help="time every launch with dispatch events instead of device clocks (slows the run; "
                         "gives the per-kernel breakdown)")
    ap.add_argument("--no-graph-line", dest="graph_line", action="store_false",
                    help="skip the second timing of the same sweeps replayed from captured hipGraphs")
    ap.add_argument("--perm", action="store_true",
                    help="time the permuting bisection (PermBisectIterative) on the same workload with Bose "
                         "statistics instead of the headline sweep; prints its own JSON line")
    ap.add_argument("--uneven", action="store_true",
                    help="time sweeps of the 64-replica uneven batch (N 16-128, M 16-64, r_s 1-4, each with its own "
                         "box and k-set), eager and replayed from captured hipGraphs, and print that line instead")
    ap.add_argument("--no-extra", dest="extra", action="store_false",
                    help="skip the uneven-batch (64 replicas) and large-case (8 x N=1024, M=256) sub-lines of the "
                         "default run")
    ap.add_argument("--c5-rounds", type=int, default=0,
                    help="with --c5: time this many bisect rounds (after as many untimed ones) instead of whole "
                         "sweeps, for hardware-counter passes over the rho_k kernel")
    ap.add_argument("--c5", action="store_true",
                    help="time only the large case (8 replicas of N=1024, M=256, 4012 k-vectors) and print that line")
    ap.add_argument("--copy-mib", type=int, default=4096,
                    help="size of the device-to-device copy timed first (the practical HBM ceiling for a "
                         "read-one-write-one stream, reported as hbm_copy_GBps; 0 skips it)")
    return ap.parse_args()


def copy_ceiling(mib: int) -> float:
    """GB/s (read + write bytes) of a device-to-device hipMemcpy of `mib` MiB, best
    of three, timed with HIP events: the practical HBM rate of a stream that
    reads one array and writes another, as the rho_k update does."""
    import ctypes as C

    hip = C.CDLL("libamdhip64.so")
    n = mib << 20
    a, b = C.c_void_p(), C.c_void_p()
    e0, e1 = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(n)) == 0 and hip.hipMalloc(C.byref(b), C.c_size_t(n)) == 0
    assert hip.hipEventCreate(C.byref(e0)) == 0 and hip.hipEventCreate(C.byref(e1)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(n)) == 0
    best = float("inf")
    for _ in range(3):
        assert hip.hipEventRecord(e0, None) == 0
        assert hip.hipMemcpyAsync(b, a, C.c_size_t(n), 3, None) == 0  # hipMemcpyDeviceToDevice
        assert hip.hipEventRecord(e1, None) == 0
        assert hip.hipEventSynchronize(e1) == 0
        ms = C.c_float()
        assert hip.hipEventElapsedTime(C.byref(ms), e0, e1) == 0
        best = min(best, ms.value)
    hip.hipEventDestroy(e0)
    hip.hipEventDestroy(e1)
    hip.hipFree(a)
    hip.hipFree(b)
    gbps = 2.0 * n / (best * 1e-3) / 1e9
    print(f"device copy of {mib} MiB: {gbps:.0f} GB/s (read + write)", file=sys.stderr, flush=True)
    return gbps


def _cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(sweeps: int, one_thread_s: float):
    """The CPU oracle built with the reference's release flags (-O3 -ffast-math
    -fopenmp, reference CMakeLists.txt:64-66) on one replica of the same
    workload: `sweeps` whole sweeps of moves (2048 bisect + 256 displace
    attempts each) on every host core, then the sweep's estimators timed apart
    (virial energy with the reference's finite-difference GetActionGradient,
    S(k), g(r) with 200 bins), and a bounded sample of the moves on one thread.
    `value` is a whole sweep (moves and estimators) per its time, as on the GPU."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_ffi
    from simpimc_amd import abi, system

    oracle_ffi.use_timing_build()
    cores = len(os.sched_getaffinity(0))
    s = system.config_c2()
    kc = 0.95 * float(system.kmag(s.L, s.k_n).max())

    def fresh(threads):
        oracle_ffi.set_threads(threads)
        o = oracle_ffi.Oracle(s, seed=17)
        o.upload(system.random_path(s, seed=17, spread=0.05))
        o.obs_configure(energy=abi.ENERGY_VIRIAL_COMPAT, sk_k_cut=kc, gr_n_r=200)
        o.bisect(64)  # warm caches and OpenMP pools
        o.displace(8)
        return o

    o = fresh(cores)
    per_sweep = 0
    t_moves = 0.0
    for k in range(sweeps):
        u0 = o.bead_updates()
        t0 = time.perf_counter()
        o.sweep(1, 0)
        t_moves += time.perf_counter() - t0
        per_sweep = o.bead_updates() - u0
        print(f"cpu oracle: sweep {k + 1}/{sweeps} moves done ({t_moves:.1f} s)", file=sys.stderr, flush=True)
    est = {}
    for which, name in ((abi.OBS_ENERGY, "virial_energy_fd_gradient"), (abi.OBS_SK, "sk"), (abi.OBS_GR, "gr")):
        t0 = time.perf_counter()
        o.obs_accumulate(which)
        est[name] = time.perf_counter() - t0
    print(f"cpu oracle: estimators {est}", file=sys.stderr, flush=True)
    sweep_moves_s = t_moves / sweeps
    sweep_s = sweep_moves_s + sum(est.values())

    o1 = fresh(1)
    u0 = o1.bead_updates()
    t0 = time.perf_counter()
    n_b = n_d = 0
    while time.perf_counter() - t0 < one_thread_s:
        o1.bisect(64)
        o1.displace(8)
        n_b += 64
        n_d += 8
    dt1 = time.perf_counter() - t0
    du1 = o1.bead_updates() - u0
    print(f"cpu oracle, one thread: {du1 / dt1:.4g} bead updates/s over the moves", file=sys.stderr, flush=True)
    return {
        "value": per_sweep / sweep_s,
        "unit": "bead updates/s",
        "cores": cores,
        "kind": "port",
        "build": "-O3 -ffast-math -fopenmp (the oracle's timing build)",
        "cpu_model": _cpu_model(),
        "nproc": os.cpu_count(),
        "sample": (f"1 replica of the same workload: {sweeps} whole sweep(s) of moves on {cores} threads "
                   f"({per_sweep} bead updates each), then one sweep's estimators"),
        "moves_only": per_sweep / sweep_moves_s,
        "sweep_moves_s": sweep_moves_s,
        "estimators_s": est,
        "one_thread": {"moves_only": du1 / dt1,
                       "sample": f"{n_b} bisect + {n_d} displace attempts ({du1} bead updates) in {dt1:.1f} s"},
    }


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py: this engine runs on one GPU; --gpus {a.gpus} is not supported (replicas would "
              f"shard across GPUs with no traffic between them)", file=sys.stderr)
        sys.exit(2)
    from simpimc_amd import abi, system

    if a.perm:
        return perm_bench(a, abi, system)
    if a.uneven:
        print(json.dumps(uneven_bench(a, abi, system)))
        return
    if a.c5:
        print(json.dumps(c5_bench(a, abi, system, copy_ceiling(a.copy_mib) if a.copy_mib else None,
                                  rounds=a.c5_rounds)))
        return
    if a.cpu_only:
        print(json.dumps({"cpu_baseline": cpu_baseline(a.cpu_sweeps, a.cpu_one_thread_seconds)}))
        return
    copy_gbps = copy_ceiling(a.copy_mib) if a.copy_mib else None
    base = system.config_c2()
    R = a.replicas
    systems = [base] * R
    flags = abi.SWEEP_ADJOINT | abi.SWEEP_ENERGY | abi.SWEEP_SK | abi.SWEEP_GR
    if a.graph:
        flags |= abi.SWEEP_GRAPH
    eng = abi.Engine(systems, seed=2026)
    for r in range(R):
        eng.upload(system.random_path(base, seed=1000 + r, spread=0.05), r=r)
    print(f"{R} replicas uploaded", file=sys.stderr, flush=True)
    # launch shape: tried on short bisection stages (ordinary, untimed moves)
    shape = eng.autotune(32) if a.autotune else None
    if shape:
        print(f"autotuned launch shape: {shape}", file=sys.stderr, flush=True)
    kc = 0.95 * float(system.kmag(base.L, base.k_n).max())
    eng.obs_configure(energy=abi.ENERGY_VIRIAL_ADJOINT, sk_k_cut=kc, gr_n_r=200)
    # the timing mode is set before the warm-up, so the graphs it captures are the timed ones
    if a.no_timing:
        eng.timing(0)
    elif a.kernel_events:
        eng.timing(abi.TIMING_EVENTS)  # events cannot be captured: these sweeps run eagerly
    else:
        eng.timing(abi.TIMING_CLOCK)
    for i in range(a.warmup):
        eng.sweep(1, flags)
        print(f"warmup sweep {i + 1}/{a.warmup} done", file=sys.stderr, flush=True)
    eng.timing_reset()
    u0 = eng.bead_updates()
    t0 = time.perf_counter()
    for i in range(a.steps):
        eng.sweep(1, flags)  # returns after the device work has finished
        print(f"timed sweep {i + 1}/{a.steps} done at {time.perf_counter() - t0:.2f} s", file=sys.stderr, flush=True)
    dt = time.perf_counter() - t0
    du = eng.bead_updates() - u0
    value = du / dt
    kernels = {}
    names = ["propose", "pair", "rho", "decide", "adjoint", "estimators"]
    for i, n in enumerate(names):
        ms, cnt, by = eng.kernel_timing(i)
        if cnt:
            kernels[n] = {"ms": ms, "launches": cnt, "avg_us": 1e3 * ms / cnt}
    if "pair" in kernels:
        _, cnt, links = eng.kernel_timing(abi.KERNEL_PAIR)
        kernels["pair"]["link_evals_per_s"] = links / (kernels["pair"]["ms"] * 1e-3)
    rho_ms, rho_n, rho_bytes = eng.kernel_timing(abi.KERNEL_RHO)
    roofline = None
    if rho_n:
        ach = rho_bytes / (rho_ms * 1e-3) / 1e9
        # HBM bytes per k_rho_loop launch of this workload, measured with rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE
        # (reads x2.000, writes x1.000, the factors a 1 GiB device copy gives), for the run whose algorithmic
        # bytes per launch were 1786.9 MB; reported only when this run's algorithmic bytes are the same
        same = abs(rho_bytes / rho_n - MEASURED_RHO_ALG_BYTES) < 1e-3 * MEASURED_RHO_ALG_BYTES
        traffic = MEASURED_RHO_TRAFFIC_BYTES if same else None
        roofline = {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBPS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBPS,
                    "traffic": traffic, "traffic_unit": "bytes per launch",
                    "traffic_from": ("profiles/k_rho_loop_k_pair_lds_heg256_m128_1024rep_counters.md" if same else None),
                    "kernel": "k_rho (incremental rho_k update + long-range action)",
                    "bytes_per_launch": rho_bytes / rho_n, "avg_launch_us": 1e3 * rho_ms / rho_n,
                    "duration_from": ("dispatch events" if a.kernel_events else
                                      "device wall clock, first block start to last block end, every launch "
                                      "of the timed region"),
                    "hbm_copy_GBps": copy_gbps, "frac_of_copy": (ach / copy_gbps) if copy_gbps else None}
    c0 = eng.counters(0)
    # the same sweeps replayed from captured hipGraphs (16 rounds per graph, the
    # fork and join included); one untimed sweep captures the graphs
    graph_line = None
    if a.graph_line and not a.graph and not a.kernel_events:
        eng.sweep(1, flags | abi.SWEEP_GRAPH)
        u0g = eng.bead_updates()
        t0g = time.perf_counter()
        for i in range(a.steps):
            eng.sweep(1, flags | abi.SWEEP_GRAPH)
        dtg = time.perf_counter() - t0g
        graph_line = {"value": (eng.bead_updates() - u0g) / dtg, "ms_per_step": 1e3 * dtg / a.steps, "steps": a.steps}
        print(f"graph replay: {graph_line['value']:.4g} bead updates/s", file=sys.stderr, flush=True)
    eng.close()

    plain = None
    if not a.no_plain:
        pe = abi.Engine(base, seed=7)
        pe.upload(system.random_path(base, seed=7, spread=0.05))
        pe.bisect(64)
        pe.displace(8)
        n_b = a.plain_attempts
        n_d = max(1, n_b // 8)
        u0p = pe.bead_updates()
        t1 = time.perf_counter()
        pe.bisect(n_b)
        pe.displace(n_d)
        dtp = time.perf_counter() - t1
        dup = pe.bead_updates() - u0p
        plain = {"value": dup / dtp, "attempts": {"bisect": n_b, "displace": n_d}, "seconds": dtp}
        print(f"plain single chain: {dup / dtp:.4g} bead updates/s", file=sys.stderr, flush=True)
        pe.close()

    extra_uneven = extra_c5 = None
    if a.extra:
        extra_uneven = uneven_bench(a, abi, system, steps=5)
        extra_c5 = c5_bench(a, abi, system, copy_gbps, steps=2)
    cpu = None if a.no_cpu else cpu_baseline(a.cpu_sweeps, a.cpu_one_thread_seconds)
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "bead updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1e3 * dt / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": (f"HEG N=256 M=128 r_s=1 tau=0.02 FP64, David pair action n_order=2 (2000-point table), "
                         f"Ewald with {base.n_k} stored k-vectors, bisect level 4 + displace; "
                         f"{R} replicas batched, adjoint dS/dr and estimators every sweep"),
            "replicas": R,
            "launch_shape": shape,
            "graph": bool(a.graph) and not a.kernel_events,
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
        "plain_single_chain": plain,
        "batched_vs_plain": (value / plain["value"]) if plain else None,
        "graph_replay": graph_line,
        "rho_k_update_GBps": roofline["achieved"] if roofline else None,
        "hbm_copy_GBps": copy_gbps,
        "kernels": kernels,
        "acceptance_replica0": {"bisect": float(c0[1]) / max(1, c0[0]), "displace": float(c0[3]) / max(1, c0[2])},
        "uneven": extra_uneven,
        "c5": extra_c5,
    }
    print(json.dumps(out))


def perm_bench(a, abi, system):
    """Permuting bisection on the headline workload's system: a step is
    N M / 2^l PermBisectIterative attempts on every replica.  Reported: bead
    updates/s (n_chain (2^l - 1) per attempt that reached the bisection), the
    attempts' cycle lengths and acceptance, and the time per kernel class."""
    base = system.config_c2()
    R = a.replicas
    eng = abi.Engine([base] * R, seed=2027)
    for r in range(R):
        eng.upload(system.random_path(base, seed=3000 + r, spread=0.05), r=r)
    n_att = (base.n_part * base.n_bead) >> base.bisect_n_level
    eng.timing(abi.TIMING_EVENTS)
    for i in range(a.warmup):
        eng.perm_bisect(n_att)
        eng.bisect(n_att)
        print(f"warmup step {i + 1}/{a.warmup} done", file=sys.stderr, flush=True)
    eng.timing_reset()
    u0 = eng.bead_updates()
    t0 = time.perf_counter()
    for i in range(a.steps):
        eng.perm_bisect(n_att)
        print(f"timed step {i + 1}/{a.steps} done at {time.perf_counter() - t0:.2f} s", file=sys.stderr, flush=True)
    dt = time.perf_counter() - t0
    value = (eng.bead_updates() - u0) / dt
    kernels = {}
    for i, n in enumerate(["propose", "pair", "rho", "decide"]):
        ms, cnt, _ = eng.kernel_timing(i)
        if cnt:
            kernels[n] = {"ms": ms, "launches": cnt, "avg_us": 1e3 * ms / cnt}
    att, acc, pa, pc = eng.perm_counters(0)
    eng.close()
    print(json.dumps({
        "metric": "bead updates/s, permuting bisection, HEG N=256 M=128 FP64",
        "value": value, "unit": "bead updates/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": 1e3 * dt / a.steps, "higher_is_better": True, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"HEG N=256 M=128 bosons, David n_order=2, Ewald {base.n_k} k-vectors, "
                               f"PermBisectIterative level {base.bisect_n_level}, {n_att} attempts per step on each "
                               f"of {R} replicas", "replicas": R},
        "kernels": kernels,
        "replica0": {"attempts": att, "accepts": acc, "perm_attempt": pa[:8].tolist(), "perm_accept": pc[:8].tolist()},
    }))


def uneven_bench(a, abi, system, steps=None):
    """Sweeps of the uneven batch: 64 replicas of N in {16, 32, 64, 128},
    M in {16, 32, 64} and r_s in [1, 4], each with the C1 interaction on its
    own box and k-set, in one context.  A sweep is every replica's own
    N M / 2^l bisect and N displace attempts, the adjoint and the estimators.
    Two launch designs are timed: the autotuned rounds of launches, and the
    persistent stage kernel (one launch per stage, a workgroup per replica);
    `value` is the faster one's eager rate, and each is also replayed from
    captured hipGraphs."""
    steps = steps or a.steps
    systems = system.config_c4()
    R = len(systems)
    kc = min(0.95 * float(system.kmag(s.L, s.k_n).max()) for s in systems)
    flags = abi.SWEEP_ADJOINT | abi.SWEEP_ENERGY | abi.SWEEP_SK | abi.SWEEP_GR
    designs = {}
    for design in ("rounds", "persistent"):
        os.environ["SPIMC_PERSISTENT"] = "512" if design == "persistent" else "0"
        eng = abi.Engine(systems, seed=2028)
        for r, s in enumerate(systems):
            eng.upload(system.random_path(s, seed=4000 + r, spread=0.05), r=r)
        shape = eng.autotune(32) if (a.autotune and design == "rounds") else None
        eng.obs_configure(energy=abi.ENERGY_VIRIAL_ADJOINT, sk_k_cut=kc, gr_n_r=200)
        lines = {}
        for mode, f in (("eager", flags), ("graph", flags | abi.SWEEP_GRAPH)):
            for i in range(max(1, a.warmup)):
                eng.sweep(1, f)  # the first graph sweep captures the graphs
            u0 = eng.bead_updates()
            t0 = time.perf_counter()
            for i in range(steps):
                eng.sweep(1, f)
            dt = time.perf_counter() - t0
            lines[mode] = {"value": (eng.bead_updates() - u0) / dt, "ms_per_step": 1e3 * dt / steps}
            print(f"uneven {design} {mode}: {lines[mode]['value']:.4g} bead updates/s", file=sys.stderr, flush=True)
        eng.close()
        designs[design] = {"eager": lines["eager"], "graph_replay": lines["graph"], "launch_shape": shape}
    os.environ.pop("SPIMC_PERSISTENT", None)
    best = max(designs, key=lambda d: designs[d]["eager"]["value"])
    return {
        "metric": "bead updates/s, uneven batch of 64 replicas FP64",
        "value": designs[best]["eager"]["value"], "unit": "bead updates/s", "n_gpus": 1, "steps": steps,
        "warmup": a.warmup, "ms_per_step": designs[best]["eager"]["ms_per_step"], "higher_is_better": True,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": "64 replicas, N in {16,32,64,128}, M in {16,32,64}, r_s in [1,4], Bare Ewald pair "
                               "action on each replica's own box and k-set, bisect level 3 + displace, adjoint and "
                               "estimators every sweep", "replicas": R,
                   "beads": int(sum(s.n_part * s.n_bead for s in systems)), "design": best},
        "designs": designs,
    }


C5_TRAFFIC_RATIO = 1.048  # 16.14 MB of HBM traffic per 15.41 MB (algorithmic) rho_k launch
C5_ROOFLINE_ROUNDS = 1024  # bisect rounds that time the rho_k kernel of the combined case


def c5_bench(a, abi, system, copy_gbps, steps=None, rounds=0):
    """The combined large case: 8 replicas of HEG N=1024, M=256, r_s=1, 4012
    stored k-vectors, David n_order=2, batched with the adjoint, the virial
    energy from it, S(k) on all k-vectors and g(r) on 200 bins every sweep.
    Reported: bead updates/s of the timed sweeps, and the rho_k kernel's
    algorithmic GB/s from the device clocks (32 B per (slice, k) update) over
    bisect rounds run after them."""
    steps = steps or a.steps
    base = system.config_c5()
    R = 8
    eng = abi.Engine([base] * R, seed=2029)
    for r in range(R):
        eng.upload(system.random_path(base, seed=5000 + r, spread=0.05), r=r)
    shape = eng.autotune(16) if a.autotune else None
    eng.obs_configure(energy=abi.ENERGY_VIRIAL_ADJOINT, sk_k_cut=1e300, gr_n_r=200)
    flags = abi.SWEEP_ADJOINT | abi.SWEEP_ENERGY | abi.SWEEP_SK | abi.SWEEP_GR
    if rounds:
        # bisect rounds only: every rho_k launch updates the same number of slices
        eng.timing(abi.TIMING_CLOCK)
        eng.bisect(rounds)
        print(f"c5 {rounds} untimed bisect rounds done", file=sys.stderr, flush=True)
        eng.timing_reset()
        u0 = eng.bead_updates()
        t0 = time.perf_counter()
        eng.bisect(rounds)
        steps = 1
        print(f"c5 {rounds} timed bisect rounds done", file=sys.stderr, flush=True)
    else:
        for i in range(max(1, a.warmup)):
            eng.sweep(1, flags)
            print(f"c5 warmup sweep {i + 1} done", file=sys.stderr, flush=True)
        u0 = eng.bead_updates()
        t0 = time.perf_counter()
        for i in range(steps):
            eng.sweep(1, flags)
            print(f"c5 timed sweep {i + 1}/{steps} done at {time.perf_counter() - t0:.2f} s", file=sys.stderr,
                  flush=True)
    dt = time.perf_counter() - t0
    value = (eng.bead_updates() - u0) / dt
    if not rounds:
        # the rho_k kernel's own launches, timed on the device clocks over bisect
        # rounds on a second context of the same replicas with the default launch
        # shape (rho_k after the pair kernel, a workgroup per item: the shape of the
        # counter passes); with the clocks on, bisection rounds run as three
        # launches, so every rho_k update is one timed k_rho_loop item
        eng.close()
        eng = abi.Engine([base] * R, seed=2030)
        for r in range(R):
            eng.upload(system.random_path(base, seed=5000 + r, spread=0.05), r=r)
        eng.timing(abi.TIMING_CLOCK)
        eng.bisect(C5_ROOFLINE_ROUNDS)
        eng.timing_reset()
        eng.bisect(C5_ROOFLINE_ROUNDS)
        print(f"c5 {C5_ROOFLINE_ROUNDS} bisect rounds on the device clocks done", file=sys.stderr, flush=True)
    rho_ms, rho_n, rho_bytes = eng.kernel_timing(abi.KERNEL_RHO)
    pair_ms, pair_n, links = eng.kernel_timing(abi.KERNEL_PAIR)
    eng.close()
    ach = rho_bytes / (rho_ms * 1e-3) / 1e9 if rho_n else None
    return {
        "metric": "bead updates/s, HEG N=1024 M=256 FP64, 8 replicas batched+adjoint",
        "value": value, "unit": "bead updates/s", "n_gpus": 1, "steps": steps, "warmup": a.warmup,
        "ms_per_step": 1e3 * dt / steps, "higher_is_better": True, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"HEG N=1024 M=256 r_s=1 tau=0.02, David n_order=2, Ewald {base.n_k} stored k-vectors, "
                               f"bisect level 4 + displace; {R} replicas, adjoint, virial energy, S(k) on all "
                               f"k-vectors and g(r) on 200 bins every sweep"
                               + (f" (timed: {rounds} bisect rounds, not whole sweeps)" if rounds else ""),
                   "replicas": R, "launch_shape": shape},
        "rho_k_update_GBps": ach,
        "roofline": ({"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                      "frac": ach / HBM_PEAK_GBPS, "kernel": "k_rho (incremental rho_k update + long-range action)",
                      "bytes_per_launch": rho_bytes / rho_n, "avg_launch_us": 1e3 * rho_ms / rho_n,
                      "duration_from": "device wall clock, first block start to last block end, over "
                                       + (f"the {rounds} timed bisect rounds" if rounds else
                                          f"{C5_ROOFLINE_ROUNDS} bisect rounds of the same replicas with the "
                                          "default launch shape (rho_k after the pair kernel, a workgroup per "
                                          "item), after the timed sweeps"),
                      # FETCH_SIZE / WRITE_SIZE passes over bisect rounds of this case, scaled by a copy kernel
                      "traffic_over_algorithmic": C5_TRAFFIC_RATIO,
                      "traffic_from": "profiles/k_rho_loop_heg1024_m256_8rep_counters.md",
                      "hbm_copy_GBps": copy_gbps, "frac_of_copy": (ach / copy_gbps) if copy_gbps else None}
                     if rho_n else None),
        "pair": ({"avg_us": 1e3 * pair_ms / pair_n, "link_evals_per_s": links / (pair_ms * 1e-3)} if pair_n else None),
    }


if __name__ == "__main__":
    main()
