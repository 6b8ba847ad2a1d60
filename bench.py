"""Throughput of per-state rates + mixture-averaged transport on one MI355X.

Workload: the 53-species / 325-reaction synthetic mechanism (30 three-body, 25 Troe), 2^20 states in SoA
layout (T ~ U[600, 2800] K, P ~ U[1, 40] atm, Y ~ Dirichlet(0.3)), outputs ω_k, λ, μ, D_k. A step is one pass
of the chemistry kernel(s) over all states, inputs already in HBM. Prints one JSON line.

  python bench.py [--steps K] [--warmup W] [--form fp64|straight|mixed] [--states N] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_TFLOPS = 78.6  # MI355X FP64 (vector and matrix)
PEAK_HBM_GBS = 8000.0
MECHANISM = "gri-53"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(name, budget_s=12.0):
    """The FP64 oracle (a C restatement of the reference's per-state evaluation) on the host's cores."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from helpers import setup
    from combustion_amd import synth
    _, _, o = setup(name)
    threads = max(1, min(16, os.cpu_count() or 1))
    T, P, Y = synth.config_states(name, 256, 99)
    t = time.perf_counter()
    o.batch(T, P, Y, threads=1)
    single = 256 / (time.perf_counter() - t)
    n = int(max(1024, min(1 << 20, single * threads * budget_s)))
    T, P, Y = synth.config_states(name, n, 99)
    t = time.perf_counter()
    o.batch(T, P, Y, threads=threads)
    rate = n / (time.perf_counter() - t)
    return {"value": rate, "unit": "states/s", "cores": threads, "kind": "port",
            "sample": f"{n} states of the same workload, FP64 oracle (oracle/chem.c) on {threads} host threads; "
                      f"single-thread {single:.0f} states/s"}


def traffic_from_profiles(kernel):
    path = os.path.join(ROOT, "profiles", f"{kernel}_traffic.json")
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f).get("bytes_per_launch")
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--form", default="auto")
    ap.add_argument("--states", type=int, default=1 << 20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--block-size", type=int, default=0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU: every state is independent, so N GPUs would each run "
                                   "the 1-GPU workload on their own shard with no communication; only --gpus 1 is measured"}))
        sys.exit(2)

    import torch
    from combustion_amd import Mechanism, synth

    if not torch.cuda.is_available():
        raise SystemExit("bench.py needs a GPU")
    dev = torch.device("cuda:0")
    M = Mechanism(synth.mechanism_path(MECHANISM))
    prog = M.chemistry()
    counts = prog.op_counts()
    form = args.form
    if form == "auto":
        form = "fp64"
        try:
            mod = prog.assemble(args.block_size, form=form)
        except Exception as e:
            log("fp64 form unavailable:", e)
            form = "straight"
            mod = prog.assemble(args.block_size, form=form)
    else:
        mod = prog.assemble(args.block_size, form=form)
    N = args.states
    T, P, Y = synth.config_states(MECHANISM, N, 2024)
    inputs = [torch.from_numpy(np.ascontiguousarray(x)).to(dev) for x in [T, P, *Y]]
    outputs = [torch.empty(N, dtype=torch.float64, device=dev) for _ in range(prog.n_outputs)]
    torch.cuda.synchronize()
    for _ in range(args.warmup):
        mod(inputs, outputs)
    torch.cuda.synchronize()
    kernel_ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        mod(inputs, outputs, time_kernel=True)  # HIP events on the launch stream around the kernel(s)
        kernel_ms.append(mod.last_kernel_ms)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    assert all(torch.isfinite(o).all().item() for o in outputs[:4])
    ms_per_step = elapsed / args.steps * 1e3
    value = N * args.steps / elapsed
    avg_kernel_s = float(np.mean(kernel_ms)) / 1e3
    flop_per_state = counts["weighted"]
    achieved = flop_per_state * N / avg_kernel_s / 1e12
    hbm_bytes = 8 * N * (prog.n_inputs + prog.n_outputs)
    result = {
        "metric": "states/s rates+transport, K=53/R=325: FP64 plain vs mixed+surrogate at stated err",
        "value": value, "unit": "states/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic (seeded mechanism and states)",
        "config": {"workload": "rates+transport, K=53 species / R=325 reactions (30 three-body, 25 Troe), 2^20 states",
                   "states": N, "form": form, "mechanism_seed": synth.CONFIGS[MECHANISM][4]},
        "roofline": {"bound": "fp64", "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP64_TFLOPS, "traffic": traffic_from_profiles(f"chemistry_{form}"),
                     "flop_per_state": flop_per_state, "kernel_ms": avg_kernel_s * 1e3,
                     "hbm_algorithmic_bytes": hbm_bytes, "hbm_frac": hbm_bytes / avg_kernel_s / 1e9 / PEAK_HBM_GBS},
    }
    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(MECHANISM)
    print(json.dumps(result))


if __name__ == "__main__":
    main()
