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
    ap.add_argument("--frames", default="on", choices=["on", "off"],
                    help="Pauli gates on bucket qubits through the per-bucket frame (on) or the amplitude "
                         "registers (off)")
    ap.add_argument("--renorm-mode", type=int, default=0, choices=[0, 1],
                    help="FP32 renormalisation: 0 one launch per period, 1 one pass per run of gates (checked)")
    ap.add_argument("--tail-gates", type=int, default=200,
                    help="also run the last N gates one gate per launch (0: skip)")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--mode", default="large", choices=["large", "batch", "serve"],
                    help="large: the 48-qubit 2^20-key circuit (default, the headline metric); "
                         "batch: 256 independent circuits in one batch; serve: Poisson arrivals")
    ap.add_argument("--circuits", type=int, default=256, help="batch mode: circuits")
    ap.add_argument("--eager", action="store_true",
                    help="batch mode: launch every step directly instead of replaying a captured hipGraph")
    ap.add_argument("--requests", type=int, default=600, help="serve mode: requests per rate")
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


def cpu_baseline(w, seconds):
    """Oracle (FP64 sparse map, OpenMP) on the first gates of the same workload,
    on up to 16 host cores, then on one core over a shorter sample."""
    from oracle import oracle as O
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    cores = min(16, avail)
    gl = w["gates"]

    def timed(threads, limit):
        O.set_threads(threads)
        s = O.Sparse(w["nq"])
        s.load(w["keys"], w["re"], w["im"])
        t0 = time.perf_counter()
        done = 0
        while done < len(gl) and time.perf_counter() - t0 < limit:
            s.apply(gl[done:done + 25])
            done = min(len(gl), done + 25)
        return s.updates() / (time.perf_counter() - t0), done, time.perf_counter() - t0

    v, done, dt = timed(cores, seconds)
    v1, done1, dt1 = timed(1, seconds / 3)
    return {"value": v, "unit": "updates/s", "cores": cores, "kind": "port",
            "sample": "CPU oracle (FP64 sparse map) on the first %d of %d gates of the same "
                      "workload from the same 2^%d-key state, %.1f s" % (
                          done, len(gl), int(round(__import__("math").log2(w["keys"].size))), dt),
            "one_core": {"value": v1, "gates": done1, "seconds": dt1},
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


def precision_report(w, eng32):
    """FP32 final state against the FP64 engine on the same workload: infidelity
    1 - |<psi64|psi32>|^2 / (|psi64|^2 |psi32|^2) and max |a32 - a64| over the
    key union (a key missing on one side counts with amplitude 0)."""
    from sparqc_amd import engine as E
    k32, r32, i32 = eng32.export()
    with E.Engine(48, prec="f64") as e:
        e.load(w["keys"], w["re"], w["im"])
        e.apply(w["gates"])
        k64, r64, i64 = e.export()
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


def traffic_from_profiles(prec, window):
    """HBM bytes per launch measured by rocprofv3 counters, when committed."""
    p = os.path.join(ROOT, "profiles", "sweep_traffic.json")
    try:
        with open(p) as f:
            d = json.load(f)
        return d.get("%s_w%d" % (prec, window))
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


def batch_mode(a):
    """256 independent circuits of uneven size and depth advanced together."""
    from sparqc_amd import engine as E
    from sparqc_amd import gen
    cs = gen.batch_circuits(a.seed, n_circuits=a.circuits)
    with E.Batch(prec=a.prec, renorm_every=100 if a.prec == "f32" else 0, timing=False, eager=a.eager) as b:
        ids = [b.add(c["nq"], c["keys"], c["re"], c["im"], c["gates"]) for c in cs]
        tune = b.autotune()
        b.snapshot_save()
        for _ in range(a.warmup):
            b.snapshot_restore()
            b.run()
            b.sync()
        t0 = time.perf_counter()
        for _ in range(a.steps):
            b.snapshot_restore()
            b.run()
            b.sync()
        dt = time.perf_counter() - t0
        upd = sum(b.info(i)["updates"] for i in ids)
        nnz = sum(b.info(i)["nnz"] for i in ids)
        used = b.usage()["used"]
    out = {"metric": "nonzero-amp gate updates/s (batched)", "value": upd / (dt / a.steps), "unit": "updates/s",
           "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": dt / a.steps * 1e3,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": a.prec,
           "data": "synthetic (seeded splitmix64 circuits and states)",
           "config": {"workload": "%d circuits, 16-40 qubits, 2^8-2^18 initial keys, depth 50-500, "
                                  "60/35/5 diag/perm/branch mix on 3 branching qubits each" % a.circuits,
                      "circuits": a.circuits, "seed": a.seed, "prec": a.prec,
                      "launch": "eager steps" if a.eager else "captured hipGraph replay"},
           "detail": {"updates_per_step": upd, "nnz_final": nnz, "initial_keys": int(sum(c["keys"].size for c in cs)),
                      "gates": int(sum(len(c["gates"]) for c in cs)), "pool_bytes": used,
                      "autotune": tune, "circuits_per_s": a.circuits / (dt / a.steps)}}
    print(json.dumps(out), flush=True)


def serve_mode(a):
    """Poisson arrivals at each rate, continuous batching, p50/p99 latency."""
    from sparqc_amd import gen
    from sparqc_amd.serve import ServeLoop
    reqs = gen.batch_circuits(a.seed, n_circuits=a.requests)
    rows = []
    for rate in [float(x) for x in a.rates.split(",")]:
        arr = gen.poisson_arrivals(a.seed + int(rate), rate, len(reqs))
        with ServeLoop(prec=a.prec, step_gates=a.step_gates, hbm_budget_bytes=200 * 10**9) as s:
            s.run(reqs[:8], np.zeros(8))                       # warm-up (kernels, pools)
            rep = s.run(reqs, arr)
        rows.append({"rate_per_s": rate, "p50_latency_ms": rep["p50_latency_s"] * 1e3,
                     "p99_latency_ms": rep["p99_latency_s"] * 1e3, "updates_per_s": rep["updates_per_s"],
                     "wall_s": rep["wall_s"], "steps": rep["steps"], "peak_live": rep["peak_live"],
                     "peak_pool_bytes": rep["peak_pool_bytes"]})
        print(json.dumps({"serve": rows[-1], "prec": a.prec}), flush=True)
    out = {"metric": "serving latency p50/p99 and nonzero-amp gate updates/s", "unit": "ms / updates/s",
           "n_gpus": 1, "dtype": a.prec, "data": "synthetic (seeded splitmix64 circuits, Poisson arrivals)",
           "config": {"workload": "%d requests per rate drawn like the batched workload, HBM budget 200 GB, "
                                  "%d gates per batched step" % (a.requests, a.step_gates),
                      "rates": a.rates, "seed": a.seed, "prec": a.prec},
           "rows": rows}
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
    from sparqc_amd import engine as E
    from sparqc_amd import gen

    if a.circuit == "layers":
        w = gen.circuit_layers(a.seed, nq=48, n_init=1 << a.n_init_log2, n_gates=a.gates)
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
                                "single-qubit gates on 4 qubits with diagonal / CX gates between"
                                % (a.n_init_log2, a.gates)),
                   "circuit": a.circuit,
                   "nq": 48, "n_init": 1 << a.n_init_log2, "gates": a.gates, "seed": a.seed,
                   "prec": a.prec, "fuse_window": a.fuse_window, "fused_blocks": a.fused_blocks},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK, "traffic": traffic,
                     "kernel": "k_sweep (fused local sweep)",
                     "algorithmic_bytes_per_launch": per_launch_bytes,
                     "avg_launch_ms": per_launch_s * 1e3, "launches": sweeps,
                     "note": "the fused sweep keeps each bucket in registers for the whole program, so it "
                             "is instruction-issue bound, not HBM bound ('issue'); 'unfused_tail' runs the "
                             "same kernel one gate per launch, where HBM is the bound",
                     "issue": valu_from_profiles(a, sweep_s / a.steps),
                     "per_gate_model": {"bytes_per_update": PER_UPDATE_BYTES[a.prec],
                                        "equivalent_GBps": upd / dt * PER_UPDATE_BYTES[a.prec] / 1e9,
                                        "equivalent_frac": upd / dt * PER_UPDATE_BYTES[a.prec] / HBM_PEAK},
                     "unfused_tail": unfused},
        "detail": {"updates_per_step": upd / a.steps, "nnz_initial": nnz0, "nnz_final": nnz1,
                   "buckets": s1["buckets"], "bucket_bits": s1["bucket_bits"],
                   "cross_bucket_launches": s1["pairs"] - s0["pairs"],
                   "sweep_kernel_ms_per_step": sweep_s / a.steps * 1e3,
                   "device_bytes": s1["bytes_device"]},
    }
    if a.prec == "f32":
        out["detail"]["precision_vs_fp64"] = precision_report(w, eng)
    out["detail"]["stream_copy"] = stream_copy()
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(w, a.cpu_seconds)
    else:
        out["cpu_baseline"] = None
    eng.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
