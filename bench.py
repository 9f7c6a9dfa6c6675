"""Throughput of the sparse statevector engine on one MI355X.

Workload (one step): the 48-qubit state of 2^20 distinct uniform keys with
normalised complex-normal amplitudes, taken through the seeded 2000-gate circuit
(60% diagonal RZ/PHASE/CZ, 35% permutation X/CX/CCX, 5% branching H/RY on 4
fixed qubits; nonzeros grow from 2^20 toward 2^24), FP64 amplitudes, prune
threshold 1e-14.  The state is resident in HBM before timing; each step restores
the device-side snapshot of the initial state and applies the circuit.

    python bench.py [--steps K] [--warmup W] [--prec f64|f32] [--fuse-window N]

Prints one JSON line: value = nonzero-amplitude gate updates per second
(sum over gates of the nonzeros entering each gate / wall time), with the
roofline of the dominant kernel (the fused local sweep, timed with HIP events on
the engine's stream) and the CPU oracle's rate on a bounded sample of the same
workload.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12          # MI355X HBM3E bytes/s
VALU_PEAK = 1024 * 2.4e9 / 4   # wave64 VALU instructions/s: 1024 SIMDs x 2.4 GHz, 4 cycles each
KEY_B = 8
# bytes per update of an unfused per-gate engine on this gate mix (0.6 DIAG at
# K+2A, 0.35 PERM at 2(K+A), 0.05 BRANCH at (1+2)(K+A) per entering nonzero)
PER_UPDATE_BYTES = {"f64": 0.6 * 40 + 0.35 * 48 + 0.05 * 72, "f32": 0.6 * 24 + 0.35 * 32 + 0.05 * 48}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--prec", default="f64", choices=["f64", "f32"])
    ap.add_argument("--fuse-window", type=int, default=0)
    ap.add_argument("--bucket-bits", type=int, default=-1)
    ap.add_argument("--n-init-log2", type=int, default=20)
    ap.add_argument("--gates", type=int, default=2000)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--fused-blocks", type=int, nargs="?", const=6, default=0,
                    help="FP32 fused-block variant: windows of at least N (default 6) bucket-qubit "
                         "gates as dense 16x16 MFMA blocks")
    ap.add_argument("--circuit", default="mix", choices=["mix", "layers"],
                    help="mix: the 60/35/5 diagonal/permutation/branching circuit; layers: layered "
                         "single-qubit gates on the 4 branching qubits (long fusable windows)")
    ap.add_argument("--layer-qubits", type=int, default=4,
                    help="layers circuit: branching qubits the layers act on (5 with --bucket-bits 5 and "
                         "--fused-blocks: 32x32 complex blocks)")
    ap.add_argument("--frames", default="on", choices=["on", "off"],
                    help="Pauli gates on bucket qubits through the per-bucket frame (on) or the amplitude "
                         "registers (off)")
    ap.add_argument("--renorm-mode", type=int, default=1, choices=[0, 1],
                    help="FP32 renormalisation: 0 one launch per period, 1 one pass per run of gates (checked)")
    ap.add_argument("--tail-gates", type=int, default=200,
                    help="also run the last N gates one gate per launch (0: skip)")
    ap.add_argument("--cpu-chunk", type=int, default=100,
                    help="gates per CPU-baseline chunk (chunks start at gates 0, 500, 1000, 1500)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the FP32, batched and serving measurements of the default line")
    ap.add_argument("--mode", default="large", choices=["large", "batch", "serve", "c1", "wide"],
                    help="large: the 48-qubit 2^20-key circuit (default, the headline metric); "
                         "batch: 256 independent circuits in one batch; serve: Poisson arrivals; "
                         "c1: the 20-qubit 200-gate circuit from |0...0> (CPU baseline: the FC engine); "
                         "wide: the 48-qubit circuit branching on --branch qubits (cross-bucket gates)")
    ap.add_argument("--branch", type=int, default=6, help="wide mode: branching qubits")
    ap.add_argument("--circuits", type=int, default=256, help="batch mode: circuits")
    ap.add_argument("--eager", action="store_true",
                    help="batch mode: launch every step directly instead of replaying a captured hipGraph")
    ap.add_argument("--requests", type=int, default=600,
                    help="requests per rate (serve mode, and the serving detail of the default line)")
    ap.add_argument("--rates", default="200,1000,5000", help="serve mode: arrival rates (1/s)")
    ap.add_argument("--step-gates", type=int, default=64, help="serve mode: gates per batched step")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device copy 5 times (scale factor for the "
                         "FETCH_SIZE / WRITE_SIZE counters)")
    return ap.parse_args()


def copy_calibration():
    """A device-to-device copy kernel of known size: 5 x 1 GiB read + 1 GiB written."""
    import torch
    n = 1 << 27                      # 2^27 float64 = 1 GiB
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes_per_launch": n * 8, "launches": 5}), flush=True)


def stream_copy():
    """Device-to-device copy bandwidth (hipMemcpy of 1 GiB, 5 times, timed with
    HIP events): the practical HBM ceiling beside the 8 TB/s datasheet peak."""
    import ctypes
    try:
        hip = ctypes.CDLL("libamdhip64.so")
    except OSError:
        return None
    n = 1 << 30
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    e0, e1 = ctypes.c_void_p(), ctypes.c_void_p()
    if hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(n)) or hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(n)):
        return None
    try:
        hip.hipMemset(a, 1, ctypes.c_size_t(n))
        hip.hipEventCreate(ctypes.byref(e0))
        hip.hipEventCreate(ctypes.byref(e1))
        hip.hipMemcpy(b, a, ctypes.c_size_t(n), 3)                 # warm-up
        hip.hipEventRecord(e0, None)
        for _ in range(5):
            hip.hipMemcpyAsync(b, a, ctypes.c_size_t(n), 3, None)
        hip.hipEventRecord(e1, None)
        hip.hipEventSynchronize(e1)
        ms = ctypes.c_float()
        hip.hipEventElapsedTime(ctypes.byref(ms), e0, e1)
        hip.hipEventDestroy(e0)
        hip.hipEventDestroy(e1)
    finally:
        hip.hipFree(a)
        hip.hipFree(b)
    if ms.value <= 0:
        return None
    return {"bytes_per_copy": 2 * n, "copies": 5, "GBps": 5 * 2 * n / (ms.value * 1e-3) / 1e9,
            "frac_of_peak": 5 * 2 * n / (ms.value * 1e-3) / HBM_PEAK}


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(w, chunk=100, points=(0, 500, 1000, 1500)):
    """The CPU oracle (FP64 sparse engine, OpenMP) timed on a sample spread
    across the circuit's nonzero growth: `chunk` gates from each of the gate
    positions in `points`, each started from the exact state the circuit has
    there (taken from the GPU engine, which matches the oracle bit for bit on
    keys and within 1e-12 on amplitudes).  The value is the sampled updates
    over the sampled time, on up to 16 host cores; one core on the first two
    chunks.  Sampling, not the whole circuit, keeps the bench within minutes
    (the whole circuit takes the oracle about a minute on 16 cores)."""
    from oracle import oracle as O
    from sparqc_amd import engine as E
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    cores = min(16, avail)
    gl = w["gates"]
    pts = [p for p in points if p < len(gl)]
    starts = []
    with E.Engine(w["nq"], prec="f64") as e:
        e.load(w["keys"], w["re"], w["im"])
        e.prepare(gl)
        done = 0
        for p in pts:
            e.apply(gl[done:p])
            done = p
            starts.append(e.export(sort=False))

    def timed(threads, which):
        O.set_threads(threads)
        upd, secs, nnz = 0, 0.0, []
        for i in which:
            k, r, im = starts[i]
            s = O.Sparse(w["nq"])
            s.load(k, r, im)
            nnz.append(int(k.size))
            t0 = time.perf_counter()
            s.apply(gl[pts[i]:pts[i] + chunk])
            secs += time.perf_counter() - t0
            upd += s.updates()
        return upd, secs, nnz

    upd, secs, nnz = timed(cores, range(len(pts)))
    upd1, secs1, _ = timed(1, range(min(2, len(pts))))
    return {"value": upd / secs, "unit": "updates/s", "cores": cores, "kind": "port",
            "sample": "CPU oracle (FP64 sparse engine) on %d-gate chunks at gates %s of the same 2000-gate "
                      "workload, each from the exact state there (%s nonzeros): %d updates in %.1f s"
                      % (chunk, ",".join(str(p) for p in pts), "/".join(str(n) for n in nnz), upd, secs),
            "one_core": {"value": upd1 / secs1 if secs1 > 0 else None, "chunks": min(2, len(pts)),
                         "seconds": secs1},
            "cpu_model": cpu_model(), "cores_available": avail}


def unfused_tail(a, w, eng):
    """The last gates of the circuit one gate per launch, from the state the
    fused run reaches before them (buckets full): the HBM-bound regime of the
    same sweep kernel."""
    from sparqc_amd import engine as E
    n = min(a.tail_gates, len(w["gates"]))
    head, tail = w["gates"][:len(w["gates"]) - n], w["gates"][len(w["gates"]) - n:]
    amp_b = 16 if a.prec == "f64" else 8
    with E.Engine(48, prec=a.prec, timing=True, fuse_window=1, bucket_bits=a.bucket_bits,
                  renorm_every=0) as e:
        e.load(w["keys"], w["re"], w["im"])
        e.prepare(w["gates"])
        e.apply(head)
        e.sync()
        e.snapshot_save()
        e.apply(tail)                    # warm-up
        e.sync()
        e.snapshot_restore()
        s0 = e.stats()
        e.apply(tail)
        e.sync()
        s1 = e.stats()
    k = s1["sweeps"] - s0["sweeps"]
    ms = (s1["sweep_ms"] - s0["sweep_ms"]) / max(k, 1)
    mv = (s1["moved"] - s0["moved"]) / max(k, 1)
    model = mv * (KEY_B + amp_b)
    # what one gate-per-launch sweep must move in this layout: every bucket's
    # base key and its 2^MB amplitude slots, read once and written once
    alg = 2 * s1["buckets"] * (KEY_B + (1 << s1["bucket_bits"]) * amp_b)
    traffic = traffic_from_profiles(a.prec, 1)
    return {"gates": n, "launches": k, "avg_launch_ms": ms, "algorithmic_bytes_per_launch": alg,
            "model_bytes_per_launch": model,
            "achieved_GBps": alg / (ms * 1e-3) / 1e9 if ms > 0 else 0.0,
            "frac": alg / (ms * 1e-3) / HBM_PEAK if ms > 0 else 0.0,
            "traffic": traffic,
            "traffic_GBps": traffic / (ms * 1e-3) / 1e9 if (traffic and ms > 0) else None,
            "updates_per_s": (s1["updates"] - s0["updates"]) / ((s1["sweep_ms"] - s0["sweep_ms"]) * 1e-3)}


def precision_report(got32, want64):
    """FP32 final state against the FP64 engine's on the same workload:
    infidelity 1 - |<psi64|psi32>|^2 / (|psi64|^2 |psi32|^2) and max |a32 - a64|
    over the key union (a key missing on one side counts with amplitude 0)."""
    k32, r32, i32 = got32
    k64, r64, i64 = want64
    keys = np.union1d(k32, k64)
    a32 = np.zeros(keys.size, dtype=np.complex128)
    a64 = np.zeros(keys.size, dtype=np.complex128)
    a32[np.searchsorted(keys, k32)] = r32 + 1j * i32
    a64[np.searchsorted(keys, k64)] = r64 + 1j * i64
    ov = np.vdot(a64, a32)
    infid = 1.0 - abs(ov) ** 2 / (np.vdot(a64, a64).real * np.vdot(a32, a32).real)
    return {"infidelity": float(infid), "max_abs_diff": float(np.max(np.abs(a32 - a64))),
            "keys_fp32": int(k32.size), "keys_fp64": int(k64.size),
            "keys_only_one_side": int(keys.size * 2 - k32.size - k64.size)}


def run_timed(w, prec, steps, warmup, timing=True, **kw):
    """Steps of the workload (restore the initial snapshot, apply the circuit,
    sync) on one engine; returns (updates per step, seconds per step, engine).
    timing: HIP events around every launch (per-class kernel times)."""
    from sparqc_amd import engine as E
    eng = E.Engine(w["nq"], prec=prec, timing=timing, renorm_every=100 if prec == "f32" else 0, **kw)
    eng.load(w["keys"], w["re"], w["im"])
    eng.prepare(w["gates"])
    eng.snapshot_save()
    for _ in range(warmup):
        eng.snapshot_restore()
        eng.apply(w["gates"])
        eng.sync()
    s0 = eng.stats()
    t0 = time.perf_counter()
    for _ in range(steps):
        eng.snapshot_restore()
        eng.apply(w["gates"])
        eng.sync()
    dt = (time.perf_counter() - t0) / steps
    s1 = eng.stats()
    return (s1["updates"] - s0["updates"]) / steps, dt, eng, s0, s1


def fp32_detail(w, fp64_value, fp64_state, steps=10, warmup=2):
    """The FP32 form on the same workload: updates/s, its ratio to the FP64
    line's value, and its precision against the FP64 final state."""
    upd, dt, eng, s0, s1 = run_timed(w, "f32", steps, warmup)
    got = eng.export()
    eng.close()
    out = {"updates_per_s": upd / dt, "ms_per_step": dt * 1e3, "ratio_to_fp64": upd / dt / fp64_value,
           "renormalise_every": 100, "renormalisation": "one pass with FP64 norm checkpoints (checked)",
           "sweep_launches_per_step": (s1["sweeps"] - s0["sweeps"]) / steps,
           "reruns": s1["reruns"] - s0["reruns"], "steps": steps}
    out.update(precision_report(got, fp64_state))
    return out


def blocks_detail(seed, steps=3, warmup=1):
    """The fused-block variant of the FP32 form on the layered circuits (long
    windows of single-qubit gates on the branching qubits): 16x16 blocks with
    4 branching qubits, 32x32 blocks with 5 (bucket_bits 5), each against the
    FP64 engine on the same circuit (gate by gate, its default bucket qubits)
    and the FP32 form gate by gate.  Updates are counted by a gate-by-gate FP32
    run (a block does not see the nonzeros between its gates)."""
    from sparqc_amd import engine as E
    from sparqc_amd import gen
    out = {}
    for nb, bits, label in ((4, -1, "16x16_blocks_4_qubits"), (5, 5, "32x32_blocks_5_qubits")):
        w = gen.circuit_layers(seed, nq=48, n_init=1 << 20, n_gates=2000, n_branch=nb)
        upd64, dt64, e64, _, _ = run_timed(w, "f64", steps, warmup)
        want = e64.export()
        e64.close()
        upd32, dt32, e32, _, _ = run_timed(w, "f32", steps, warmup, bucket_bits=bits)
        e32.close()
        _, dtb, eb, s0, s1 = run_timed(w, "f32", steps, warmup, bucket_bits=bits, fuse_blocks=6)
        got = eb.export()
        eb.close()
        row = {"workload": "48 qubits, 2^20 initial keys, 2000 gates in layers on %d branching qubits" % nb,
               "fp64_updates_per_s": upd64 / dt64, "fp32_gate_by_gate_updates_per_s": upd32 / dt32,
               "fp32_blocks_updates_per_s": upd32 / dtb, "blocks_ratio_to_fp64": (upd32 / dtb) / (upd64 / dt64),
               "blocks_per_step": (s1["blocks"] - s0["blocks"]) / steps,
               "block_gates_per_step": (s1["block_gates"] - s0["block_gates"]) / steps,
               "fp32_bucket_bits": s1["bucket_bits"], "ms_per_step_fp64": dt64 * 1e3, "ms_per_step_blocks": dtb * 1e3}
        row.update(precision_report(got, want))
        out[label] = row
    return out


def traffic_from_profiles(prec, window):
    """HBM bytes per launch measured by rocprofv3 counters, when committed."""
    p = os.path.join(ROOT, "profiles", "sweep_traffic.json")
    try:
        with open(p) as f:
            d = json.load(f)
        return d.get("%s_w%d" % (prec, window))
    except (OSError, ValueError):
        return None


def counters_source():
    """Which build the committed counters (profiles/sweep_traffic.json) were taken on."""
    try:
        with open(os.path.join(ROOT, "profiles", "sweep_traffic.json")) as f:
            src = json.load(f).get("source", {})
        return {"file": "profiles/sweep_traffic.json", "commit": src.get("commit")}
    except (OSError, ValueError):
        return None


def valu_from_profiles(a, sweep_s_per_step):
    """VALU issue of the fused sweep: wave instructions per step counted by
    rocprofv3 (SQ_INSTS_VALU, committed summary) over the live kernel time,
    against the issue peak.  Only for the default workload the counters
    were taken on."""
    if (a.circuit != "mix" or a.n_init_log2 != 20 or a.gates != 2000 or a.seed != 1 or a.fuse_window
            or a.fused_blocks or a.bucket_bits != -1):
        return None
    p = os.path.join(ROOT, "profiles", "sweep_traffic.json")
    try:
        with open(p) as f:
            v = json.load(f).get("valu_%s" % a.prec)
    except (OSError, ValueError):
        return None
    if not v or sweep_s_per_step <= 0:
        return None
    ach = v["valu_insts"] / sweep_s_per_step
    return {"bound": "valu", "achieved": ach, "peak": VALU_PEAK, "unit": "wave64 VALU instructions/s",
            "frac": ach / VALU_PEAK, "valu_per_wave_gate": v["valu_per_wave_gate"],
            "insts_per_step": v["valu_insts"]}


def run_batch(seed, n_circuits, prec, steps, warmup, eager=False):
    """n_circuits independent circuits of the batched workload advanced
    together: autotuned, then timed over whole runs.  Each run starts from
    the admitted states (a snapshot restore, a device copy of the pool, done
    and synchronised before the run's timed region: the inputs are resident
    when it starts)."""
    from sparqc_amd import engine as E
    from sparqc_amd import gen
    cs = gen.batch_circuits(seed, n_circuits=n_circuits)
    with E.Batch(prec=prec, renorm_every=100 if prec == "f32" else 0, timing=False, eager=eager) as b:
        t0 = time.perf_counter()
        ids = [b.add(c["nq"], c["keys"], c["re"], c["im"], c["gates"]) for c in cs]
        admit_s = time.perf_counter() - t0
        tune = b.autotune()
        sizes = sorted({b.info(i)["bucket_bits"] for i in ids})
        b.snapshot_save()
        launches_per_run = None
        for _ in range(warmup):
            b.snapshot_restore()
            l0 = b.tuning()["launches"]
            b.run()
            b.sync()
            launches_per_run = b.tuning()["launches"] - l0
        def timed_runs():
            run_s, t_all = 0.0, time.perf_counter()
            for _ in range(steps):
                b.snapshot_restore()
                b.sync()
                t1 = time.perf_counter()
                b.run()
                b.sync()
                run_s += time.perf_counter() - t1
            return run_s / steps, (time.perf_counter() - t_all) / steps
        dt, dt_restore = timed_runs()
        upd = sum(b.info(i)["updates"] for i in ids)
        nnz = sum(b.info(i)["nnz"] for i in ids)
        used = b.usage()["used"]
        # both launch forms at the tuned workgroup size and tiles per item,
        # outside the timed region: one sweep launch per step for the circuits
        # of 0-3 bucket qubits, or one per bucket size on graph branches
        forms = {}
        for one in (0, 1):
            b.set_tuning(one_launch=one)
            b.snapshot_restore()
            b.run()
            b.sync()
            forms["one_launch" if one else "launch_per_bucket_size"] = timed_runs()[0] * 1e3
        b.set_tuning(one_launch=tune["one_launch"])
    return {"updates_per_s": upd / dt, "ms_per_run": dt * 1e3, "ms_per_run_with_restore": dt_restore * 1e3,
            "circuits": n_circuits, "prec": prec,
            "updates_per_run": upd, "nnz_final": nnz, "initial_keys": int(sum(c["keys"].size for c in cs)),
            "gates": int(sum(len(c["gates"]) for c in cs)), "pool_bytes": used, "admission_s": admit_s,
            "autotune": tune, "circuits_per_s": n_circuits / dt,
            "launch": "eager steps" if eager else "captured hipGraph replay", "runs_timed": steps,
            "bucket_sizes": sizes, "sweep_launches_per_run": launches_per_run,
            "ms_per_run_by_launch_form": forms,
            "renormalisation": ("one pass: checkpoint fold, then a scale launch per bucket size"
                                if prec == "f32" else None)}


def batch_mode(a):
    """256 independent circuits of uneven size and depth advanced together."""
    r = run_batch(a.seed, a.circuits, a.prec, a.steps, a.warmup, eager=a.eager)
    out = {"metric": "nonzero-amp gate updates/s (batched)", "value": r["updates_per_s"], "unit": "updates/s",
           "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": r["ms_per_run"],
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": a.prec,
           "data": "synthetic (seeded splitmix64 circuits and states)",
           "config": {"workload": "%d circuits, 16-40 qubits, 2^8-2^18 initial keys, depth 50-500, "
                                  "60/35/5 diag/perm/branch mix on 3 branching qubits each" % a.circuits,
                      "circuits": a.circuits, "seed": a.seed, "prec": a.prec, "launch": r["launch"]},
           "detail": r}
    print(json.dumps(out), flush=True)


def run_serve(seed, n_requests, rates, prec, step_gates):
    """Poisson arrivals at each rate into the serving loop: p50/p99 latency."""
    from sparqc_amd import gen
    from sparqc_amd.serve import ServeLoop, pin_requests
    # requests arrive in page-locked receive buffers (admitted without a host
    # copy); preparing them is not part of the loop
    reqs = pin_requests(gen.batch_circuits(seed, n_circuits=n_requests))
    rows = []
    for rate in rates:
        arr = gen.poisson_arrivals(seed + int(rate), rate, len(reqs))
        with ServeLoop(prec=prec, step_gates=step_gates, hbm_budget_bytes=200 * 10**9) as s:
            # warm-up: the same arrivals once (kernels, device pools and the
            # page-locked export buffers reach their steady-state sizes)
            s.run(reqs, arr)
            rep = s.run(reqs, arr)
        rows.append({"rate_per_s": rate, "p50_latency_ms": rep["p50_latency_s"] * 1e3,
                     "p99_latency_ms": rep["p99_latency_s"] * 1e3, "updates_per_s": rep["updates_per_s"],
                     "served_per_s": len(reqs) / rep["wall_s"],
                     "wall_s": rep["wall_s"], "steps": rep["steps"], "peak_live": rep["peak_live"],
                     "peak_exports_in_flight": rep["peak_exports_in_flight"], "errors": len(rep["errors"]),
                     "peak_pool_bytes": rep["peak_pool_bytes"]})
    return {"requests_per_rate": n_requests, "prec": prec, "step_gates": step_gates, "hbm_budget_bytes": 200 * 10**9,
            "admission": "request states in page-locked buffers (sqe_batch_add_pinned)", "rows": rows}


def serve_mode(a):
    """Poisson arrivals at each rate, continuous batching, p50/p99 latency."""
    r = run_serve(a.seed, a.requests, [float(x) for x in a.rates.split(",")], a.prec, a.step_gates)
    out = {"metric": "serving latency p50/p99 and nonzero-amp gate updates/s", "unit": "ms / updates/s",
           "n_gpus": 1, "dtype": a.prec, "data": "synthetic (seeded splitmix64 circuits, Poisson arrivals)",
           "config": {"workload": "%d requests per rate drawn like the batched workload, HBM budget 200 GB, "
                                  "%d gates per batched step" % (a.requests, a.step_gates),
                      "rates": a.rates, "seed": a.seed, "prec": a.prec},
           "rows": r["rows"]}
    print(json.dumps(out), flush=True)


def kernel_classes(s0, s1, steps):
    """Per-step kernel time by class (engine timing) and cross-bucket traffic."""
    return {"sweep_ms_per_step": (s1["sweep_ms"] - s0["sweep_ms"]) / steps,
            "pair_ms_per_step": (s1["pair_ms"] - s0["pair_ms"]) / steps,
            "other_ms_per_step": (s1["other_ms"] - s0["other_ms"]) / steps,
            "sweeps_per_step": (s1["sweeps"] - s0["sweeps"]) / steps,
            "pairs_per_step": (s1["pairs"] - s0["pairs"]) / steps,
            "pair_moved_per_step": (s1["moved_pair"] - s0["moved_pair"]) / steps}


def c1_mode(a):
    """The 20-qubit, 200-gate circuit from |0...0> (H / RY on 10 active
    qubits, X / CX / RZ / CZ anywhere): more branching qubits than bucket bits,
    so cross-bucket launches and the hash index carry part of it.  CPU
    baseline: the FC engine (the reference's dense-slot table on the CPU)."""
    from oracle import oracle as O
    from sparqc_amd import gen
    w = gen.circuit_small(a.seed)
    steps = max(a.steps, 50)
    # the rate without per-launch events (a step is about 120 small launches,
    # so the timing events would be part of what is measured), then the
    # per-class kernel times from a second run with them
    upd, dt, eng, s0, s1 = run_timed(w, "f64", steps, a.warmup, timing=False)
    nnz1 = s1["nnz"]
    eng.close()
    _, dt_ev, eng, s0, s1 = run_timed(w, "f64", steps, a.warmup)
    eng.close()
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    cores = min(16, avail)

    def fc_rate(threads, reps):
        O.set_threads(threads)
        u, t = 0, 0.0
        for _ in range(reps):
            f = O.FC(20)
            f.load(w["keys"], w["re"], w["im"])
            t0 = time.perf_counter()
            f.apply(w["gates"])
            t += time.perf_counter() - t0
            u += f.updates()
        return u / t, t
    v, t = fc_rate(cores, 20)
    v1, t1 = fc_rate(1, 5)
    out = {"metric": "nonzero-amp gate updates/s", "value": upd / dt, "unit": "updates/s", "n_gpus": 1,
           "steps": steps, "warmup": a.warmup, "ms_per_step": dt * 1e3, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded splitmix64 circuit)",
           "config": {"workload": "sparse statevector, 20 qubits from |0...0>, 200 gates: H / RY on 10 active "
                                  "qubits, X / CX / RZ / CZ on any, prune 1e-14", "nq": 20, "gates": 200,
                      "seed": a.seed},
           "detail": dict(kernel_classes(s0, s1, steps), updates_per_step=upd, nnz_final=nnz1,
                          bucket_bits=s1["bucket_bits"], ms_per_step_with_timing_events=dt_ev * 1e3),
           "cpu_baseline": {"value": v, "unit": "updates/s", "cores": cores, "kind": "port",
                            "sample": "FC engine (dense slots addressed by key, address book, two buffers: the "
                                      "reference's table restated on the CPU) on the whole circuit, 20 runs, %.2f s"
                                      % t,
                            "one_core": {"value": v1, "runs": 5, "seconds": t1}, "cpu_model": cpu_model()}}
    print(json.dumps(out), flush=True)


def wide_mode(a):
    """The 48-qubit circuit branching on more qubits than a bucket holds:
    2^18 keys, 2000 gates of the 60/35/5 mix with branching on --branch
    qubits, so branching gates on the extra qubits run as cross-bucket launches
    (pairing through the hash index, prune and scan compaction)."""
    from sparqc_amd import gen
    w = gen.circuit_large(a.seed, nq=48, n_init=1 << 18, n_gates=2000, n_branch=a.branch)
    upd, dt, eng, s0, s1 = run_timed(w, "f64", a.steps, a.warmup)
    nnz1 = s1["nnz"]
    eng.close()
    kc = kernel_classes(s0, s1, a.steps)
    pair_bytes = kc["pair_moved_per_step"] * (KEY_B + 16)
    out = {"metric": "nonzero-amp gate updates/s", "value": upd / dt, "unit": "updates/s", "n_gpus": 1,
           "steps": a.steps, "warmup": a.warmup, "ms_per_step": dt * 1e3, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded splitmix64 circuit and state)",
           "config": {"workload": "sparse statevector, 48 qubits, 2^18 initial keys, 2000-gate mix 60%% diag / "
                                  "35%% perm / 5%% branch on %d qubits, prune 1e-14" % a.branch,
                      "nq": 48, "n_init": 1 << 18, "gates": 2000, "branch": a.branch, "seed": a.seed},
           "roofline": {"bound": "hbm", "kernel": "k_pair (cross-bucket gate)",
                        "achieved": pair_bytes / (kc["pair_ms_per_step"] * 1e-3) / 1e9 if kc["pair_ms_per_step"] else None,
                        "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                        "frac": pair_bytes / (kc["pair_ms_per_step"] * 1e-3) / HBM_PEAK if kc["pair_ms_per_step"] else None,
                        "traffic": None,
                        "note": "model bytes (nonzeros read + written) x 24 B over the summed cross-bucket launch "
                                "time; measured HBM bytes per kernel are in the rocprof summaries"},
           "detail": dict(kc, updates_per_step=upd, nnz_final=nnz1, bucket_bits=s1["bucket_bits"])}
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(w, chunk=a.cpu_chunk)
    print(json.dumps(out), flush=True)


def main():
    a = parse()
    if a.copy_calibration:
        copy_calibration()
        return
    if a.gpus != 1:
        print("bench.py measures one GPU only: this build runs on a single MI355X "
              "(multi-GPU sharding is a design note, not built); use --gpus 1", file=sys.stderr)
        sys.exit(2)
    if a.mode == "batch":
        batch_mode(a)
        return
    if a.mode == "serve":
        serve_mode(a)
        return
    if a.mode == "c1":
        c1_mode(a)
        return
    if a.mode == "wide":
        wide_mode(a)
        return
    from sparqc_amd import engine as E
    from sparqc_amd import gen

    if a.circuit == "layers":
        w = gen.circuit_layers(a.seed, nq=48, n_init=1 << a.n_init_log2, n_gates=a.gates, n_branch=a.layer_qubits)
    else:
        w = gen.circuit_large(a.seed, nq=48, n_init=1 << a.n_init_log2, n_gates=a.gates)
    if a.fused_blocks and a.prec != "f32":
        print("--fused-blocks needs --prec f32", file=sys.stderr)
        sys.exit(2)
    eng = E.Engine(48, prec=a.prec, timing=True, fuse_window=a.fuse_window, bucket_bits=a.bucket_bits,
                   renorm_every=100 if a.prec == "f32" else 0, fuse_blocks=a.fused_blocks,
                   frames=a.frames == "on", renorm_mode=a.renorm_mode)
    eng.load(w["keys"], w["re"], w["im"])
    eng.prepare(w["gates"])
    eng.snapshot_save()
    nnz0 = int(w["keys"].size)            # distinct keys, nonzero amplitudes
    for _ in range(a.warmup):
        eng.snapshot_restore()
        eng.apply(w["gates"])
        eng.sync()
    s0 = eng.stats()
    eng.sync()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        eng.snapshot_restore()
        eng.apply(w["gates"])
        eng.sync()
    t1 = time.perf_counter()
    s1 = eng.stats()
    nnz1 = s1["nnz"]                      # counted by the last sweep (no extra launch)
    dt = t1 - t0
    upd = s1["updates"] - s0["updates"]
    if a.fused_blocks:
        # a block does not see the nonzeros between its gates: take the update
        # count of the gate-by-gate FP32 run of the same circuit
        with E.Engine(48, prec="f32", renorm_every=100) as ref:
            ref.load(w["keys"], w["re"], w["im"])
            ref.apply(w["gates"])
            upd = ref.stats()["updates"] * a.steps
    amp_b = 16 if a.prec == "f64" else 8
    sweeps = s1["sweeps"] - s0["sweeps"]
    sweep_s = (s1["sweep_ms"] - s0["sweep_ms"]) / 1e3
    moved = s1["moved"] - s0["moved"]
    alg_bytes = moved * (KEY_B + amp_b)
    per_launch_bytes = alg_bytes / max(sweeps, 1)
    per_launch_s = sweep_s / max(sweeps, 1)
    achieved = per_launch_bytes / per_launch_s if per_launch_s > 0 else 0.0
    traffic = traffic_from_profiles(a.prec, a.fuse_window)
    unfused = unfused_tail(a, w, eng) if a.fuse_window == 0 and a.tail_gates > 0 else None
    out = {
        "metric": "nonzero-amp gate updates/s",
        "value": upd / dt,
        "unit": "updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": dt / a.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": a.prec,
        "data": "synthetic (seeded splitmix64 circuit and state)",
        "config": {"workload": ("sparse statevector, 48 qubits, 2^%d initial keys, %d-gate mix "
                                "60%% diag / 35%% perm / 5%% branch on 4 qubits, prune 1e-14"
                                % (a.n_init_log2, a.gates)) if a.circuit == "mix" else
                               ("sparse statevector, 48 qubits, 2^%d initial keys, %d gates in layers of "
                                "single-qubit gates on %d qubits with diagonal / CX gates between"
                                % (a.n_init_log2, a.gates, a.layer_qubits)),
                   "circuit": a.circuit,
                   "nq": 48, "n_init": 1 << a.n_init_log2, "gates": a.gates, "seed": a.seed,
                   "prec": a.prec, "fuse_window": a.fuse_window, "fused_blocks": a.fused_blocks,
                   "bucket_bits": a.bucket_bits},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK, "traffic": traffic,
                     "kernel": "k_sweep (fused local sweep)",
                     "algorithmic_bytes_per_launch": per_launch_bytes,
                     "avg_launch_ms": per_launch_s * 1e3, "launches": sweeps,
                     "note": "the fused sweep keeps each bucket in registers for the whole program, so it "
                             "is instruction-issue bound, not HBM bound ('issue'); 'unfused_tail' runs the "
                             "same kernel one gate per launch, where HBM is the bound",
                     "issue": valu_from_profiles(a, sweep_s / a.steps),
                     "counters_from": counters_source(),
                     "per_gate_model": {"bytes_per_update": PER_UPDATE_BYTES[a.prec],
                                        "equivalent_GBps": upd / dt * PER_UPDATE_BYTES[a.prec] / 1e9,
                                        "equivalent_frac": upd / dt * PER_UPDATE_BYTES[a.prec] / HBM_PEAK},
                     "unfused_tail": unfused},
        "detail": {"updates_per_step": upd / a.steps, "nnz_initial": nnz0, "nnz_final": nnz1,
                   "blocks_per_step": (s1["blocks"] - s0["blocks"]) / a.steps,
                   "block_gates_per_step": (s1["block_gates"] - s0["block_gates"]) / a.steps,
                   "buckets": s1["buckets"], "bucket_bits": s1["bucket_bits"],
                   "cross_bucket_launches": s1["pairs"] - s0["pairs"],
                   "sweep_kernel_ms_per_step": sweep_s / a.steps * 1e3,
                   "device_bytes": s1["bytes_device"]},
    }
    final = eng.export()
    eng.close()
    if a.prec == "f32":
        with E.Engine(48, prec="f64") as ref:
            ref.load(w["keys"], w["re"], w["im"])
            ref.apply(w["gates"])
            out["detail"]["precision_vs_fp64"] = precision_report(final, ref.export())
    elif not a.no_extras and a.circuit == "mix" and not a.fused_blocks:
        # the FP32 form, the batched form and the serving loop, measured in
        # the same run (their own fixed sizes, a few seconds each)
        out["detail"]["fp32"] = fp32_detail(w, upd / dt, final)
        final = None
        out["detail"]["fp32_blocks"] = blocks_detail(a.seed)
        out["detail"]["batch"] = {p: run_batch(a.seed, 256, p, steps=3, warmup=1) for p in ("f64", "f32")}
        out["detail"]["serve"] = run_serve(a.seed, a.requests, [200.0, 1000.0, 5000.0], "f64", a.step_gates)
    final = None
    out["detail"]["stream_copy"] = stream_copy()
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(w, chunk=a.cpu_chunk)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
