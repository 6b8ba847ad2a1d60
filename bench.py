"""Throughput of per-state rates + mixture-averaged transport on one MI355X.

Workload: the 53-species / 325-reaction synthetic mechanism (30 three-body, 25 Troe), 2^20 states in SoA layout
(T ~ U[600, 2800] K, P ~ U[1, 40] atm, Y ~ Dirichlet(0.3)), outputs ω_k, λ, μ, D_k. A step is one pass of a
form's kernels over all states, inputs already in HBM. `value` is the plain FP64 form; the line also carries the
mixed-precision forms (MFMA contractions; FP32-core straight-line rates) with their speedups and errors against the
FP64 oracle on a sample of the same states, and — when a trained network is present — the surrogate's workload:
2^20 held-out reactor-trajectory states, FP64 against FP32-core rates + network transport. Prints one JSON line.

  python bench.py [--steps K] [--warmup W] [--states N] [--no-cpu] [--surrogate PATH]
  python bench.py --only rates|transport|fp64|mixed|fast|surrogate|copy [--steps K] [--mechanism large-160]
      (one launch kind alone, for rocprofv3 runs; surrogate runs the surrogate form on 2^20 held-out trajectory
      states; copy is a device-to-device copy of --copy-mib MiB, the reference for the FETCH_SIZE/WRITE_SIZE counters)

The line also carries the 160-species / 1800-reaction workload (FP64, mixed and fast forms; --no-large leaves it
out) and the batched reactor integrator on 2^16 H2/O2 reactors (--no-integrator leaves it out).
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
LARGE = "large-160"
SURROGATE = os.path.join(ROOT, "combustion_amd", "data", "surrogate_gri53.npz")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(prog, name, budget_s=8.0):
    """The reference's CPU path as planned for the baseline: the per-state program compiled by the host compiler
    (cb_cpu_assemble, the counterpart of ir/lib.rs:225-263) and called once per state as device.rs:35-45 does, timed
    on one core and on the box's host cores (at most 16, the GPU's share of them); the FP64 oracle for information."""
    from combustion_amd import synth
    mod = prog.assemble_cpu()
    T, P, Y = synth.config_states(name, 512, 99)
    t = time.perf_counter()
    mod([T, P, *Y], threads=1)
    probe = 512 / (time.perf_counter() - t)
    n1 = int(max(1024, min(1 << 20, probe * budget_s / 2)))
    T, P, Y = synth.config_states(name, n1, 99)
    t = time.perf_counter()
    mod([T, P, *Y], threads=1)
    single = n1 / (time.perf_counter() - t)
    threads = max(1, min(16, os.cpu_count() or 1))
    n = int(max(1024, min(1 << 20, single * threads * budget_s / 2)))
    T, P, Y = synth.config_states(name, n, 99)
    t = time.perf_counter()
    mod([T, P, *Y], threads=threads)
    rate = n / (time.perf_counter() - t)
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from helpers import setup
    _, _, o = setup(name)
    T, P, Y = synth.config_states(name, 256, 98)
    t = time.perf_counter()
    o.batch(T, P, Y, threads=1)
    oracle_rate = 256 / (time.perf_counter() - t)
    return {"value": rate, "unit": "states/s", "cores": threads, "kind": "port",
            "single_core": single, "nproc": os.cpu_count(), "cpu_model": cpu_model(),
            "sample": f"{n} states (1 core: {n1}) of the same workload (rates + transport, K=53/R=325): the per-state FP64 "
                      f"program emitted as C, compiled by host clang -O2 (cb_cpu_assemble, the counterpart of the "
                      f"reference's Cranelift JIT ir/lib.rs:225-263) and called once per state as device.rs:35-45 does, on "
                      f"{threads} host threads over contiguous state ranges; the FP64 oracle (oracle/chem.c) runs "
                      f"{oracle_rate:.0f} states/s on one core"}


def issue_from_profiles(kernel, kernel_ms):
    """VALU issue of a kernel against the SIMDs' issue capacity: its VALU instructions per launch (SQ_INSTS_VALU from the
    committed counter profile) at 4 cycles per wave64 instruction on a 16-lane SIMD (16 for the quarter-rate FP64
    transcendentals), over 1024 SIMDs at the 2.4 GHz peak clock for the kernel's time."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_counters.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        c = json.load(f)
    cycles = c["waves_per_launch"] * (4 * c["valu_instructions_per_wave"] + 12 * c.get("trans_f64_instructions_per_wave", 0))
    available = 1024 * 2.4e9 * kernel_ms * 1e-3
    return {"valu_instructions_per_wave": c["valu_instructions_per_wave"], "waves": c["waves_per_launch"],
            "issue_cycles": cycles, "frac": cycles / available}


def traffic_from_profiles(kernel):
    path = os.path.join(ROOT, "profiles", f"{kernel}_traffic.json")
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f).get("bytes_per_launch")
    return None


def time_module(mod, inputs, outputs, steps, warmup):
    """(wall seconds for `steps` calls enqueued back to back, as a solver's loop would, with a synchronization on
    each side; mean kernel ms of a separate pass of `steps` calls, each timed by HIP events on the launch stream)."""
    import torch
    for _ in range(warmup):
        mod(inputs, outputs)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        mod(inputs, outputs)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kernel_ms = []
    for _ in range(steps):
        mod(inputs, outputs, time_kernel=True)
        kernel_ms.append(mod.last_kernel_ms)
    torch.cuda.synchronize()
    return wall, float(np.mean(kernel_ms))


def errors_vs_oracle(out_dev, T, P, Y, n_sample, Ka, rates=True, exact=False, mechanism=MECHANISM):
    """Per-section max and median relative error of a form's outputs on the first n_sample states: against the
    term-scale floor of the FP64 forms (exact) or the mixed forms' state floor."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from helpers import relative_error, setup, state_floor
    _, _, o = setup(mechanism)
    ref, floor = o.batch(T[:n_sample], P[:n_sample], Y[:, :n_sample], rates=rates, threads=16)
    got = np.stack([x[:n_sample].cpu().numpy() for x in out_dev])
    f = floor if exact else state_floor(floor, Ka if rates else 0)
    err = relative_error(got, ref, f)
    k = Ka if rates else 0
    sec = {"λ": err[k], "μ": err[k + 1], "D": err[k + 2:]}
    if rates:
        sec = {"ω": err[:Ka], **sec}
    return {name: {"max": float(e.max()), "median": float(np.median(e))} for name, e in sec.items()}


def surrogate_workload(args, M, prog, dev):
    """Config 4: 2^20 held-out states of constant-pressure reactor trajectories (the surrogate's domain), the FP64
    form against the surrogate transport beside either rates kernel (FP32-core straight-line; MFMA mixed precision)
    on the same states."""
    import torch
    from combustion_amd import reactor
    from combustion_amd import surrogate as sur
    N = args.states
    t0 = time.perf_counter()
    T, P, Y = reactor.trajectories(M, 2048, N, seed=4242, max_steps=800, batch=2048)
    gen_s = time.perf_counter() - t0
    inputs = [torch.from_numpy(np.ascontiguousarray(x)).to(dev) for x in [T, P, *Y]]
    outputs = [torch.empty(N, dtype=torch.float64, device=dev) for _ in range(prog.n_outputs)]
    fp64 = prog.assemble(form="fp64")
    wall_p, ms_p = time_module(fp64, inputs, outputs, args.steps, args.warmup)
    del fp64
    out = {"states": N, "generation_s": gen_s,
           "data": "held-out states of 2048 constant-pressure reactor trajectories (seed 4242), T0 ~ U[900,1600] K, "
                   "P ~ U[1,40] atm, fuel/O2/N2",
           "fp64": {"states_per_s": N * args.steps / wall_p, "kernel_ms": ms_p}}
    trained = sur.Trained.load(args.surrogate)
    # the FP32-core rates beside the network (the faster pairing), and the MFMA mixed rates (the mixed form's own)
    for rates in ("fast", "mixed"):
        smod = sur.assemble(prog, trained, rates=rates)
        wall_s, ms_s = time_module(smod, inputs, outputs, args.steps, args.warmup)
        out[f"{rates}_surrogate"] = {"states_per_s": N * args.steps / wall_s, "kernel_ms": ms_s, "speedup_vs_fp64": wall_p / wall_s,
                                     "fallback_states": sur.fallback_count(smod),
                                     "error_vs_oracle": errors_vs_oracle(outputs, T, P, Y, args.error_sample, M.active)}
        del smod
    return out


def large_mechanism_workload(args, dev):
    """Config 3: the 160-species / 1800-reaction mechanism at 2^20 states. The FP64 form (table-driven rates kernel,
    transport_tiled on v_mfma_f64_16x16x4f64), its two kernels alone, and the mixed forms, each with its error
    against the FP64 oracle on a sample of the same states."""
    import torch
    from combustion_amd import Mechanism, synth
    M = Mechanism(synth.mechanism_path(LARGE))
    N = args.states
    T, P, Y = synth.config_states(LARGE, N, 2025)
    inputs = [torch.from_numpy(np.ascontiguousarray(x)).to(dev) for x in [T, P, *Y]]
    prog = M.chemistry()
    outputs = [torch.empty(N, dtype=torch.float64, device=dev) for _ in range(prog.n_outputs)]
    steps = max(3, args.steps // 4)
    sample = max(256, args.error_sample // 4)
    out = {"workload": f"rates+transport, K={M.K} species / R={M.R} reactions, {N} states", "steps": steps,
           "error_sample": sample}
    wall64 = None
    for form in ("fp64", "mixed", "fast"):
        mod = prog.assemble(form=form)
        wall, ms = time_module(mod, inputs, outputs, steps, 1)
        entry = {"states_per_s": N * steps / wall, "kernel_ms": ms,
                 "error_vs_oracle": errors_vs_oracle(outputs, T, P, Y, sample, M.active, exact=form == "fp64", mechanism=LARGE)}
        if form == "fp64":
            wall64 = wall
            sections = {}
            for name, flags, kernel in (("rates", 1, "rates_tiled"), ("transport", 2, "transport_tiled")):
                p = M.chemistry(flags)
                m = p.assemble(form="fp64")
                outs = [torch.empty(N, dtype=torch.float64, device=dev) for _ in range(p.n_outputs)]
                _, kms = time_module(m, inputs, outs, steps, 1)
                flop = p.op_counts()["weighted"]
                sections[name] = {"kernel": kernel, "kernel_ms": kms, "flop_per_state": flop,
                                  "achieved_tflops": flop * N / (kms * 1e-3) / 1e12}
                del m, outs
            entry["sections"] = sections
        else:
            entry["speedup_vs_fp64"] = wall64 / wall
        out[form] = entry
        del mod
    return out


def integrator_workload(args, dev, jacobian="automatic", mechanism="handwritten", method="rodas4"):
    """The batched reactor integrator (cb_integrator_*, Rosenbrock ode23s or RODAS4 at rtol 1e-8; the Jacobian
    analytic or by difference quotients, "automatic" choosing by the state size) on constant-pressure reactors:
    - "handwritten": 2^16 reactors of the hand-written H2/O2 mechanism (K=6; tests/golden/handwritten.yaml, the
      twelve initial states of tests/golden/reactor_handwritten.npz repeated with T0 scaled by U[0.95, 1.05]),
      through ignition to 1 ms;
    - "gri-53": 2^14 reactors of the 53-species mechanism (the four fuel/O2/N2 mixtures of
      tests/golden/reactor_gri53.npz repeated with T0 scaled by U[0.98, 1.02]), to 0.1 ms.
    Work per reactor step: the Jacobian (the analytic program with its own right-hand side, or difference quotients
    from the right-hand side split at the mass fractions), the LU of the (K+1)² matrix, and two right-hand-side
    evaluations and three solves (ode23s) or five and six (RODAS4)."""
    import torch
    from combustion_amd import Mechanism, synth
    from combustion_amd.native import Integrator
    golden = os.path.join(ROOT, "tests", "golden")
    if mechanism == "handwritten":
        d = np.load(os.path.join(golden, "reactor_handwritten.npz"))
        M = Mechanism(os.path.join(golden, "handwritten.yaml"))
        n, spread, t_end, warm = 1 << 16, 0.05, 1e-3, 1e-5
        what = f"constant-pressure H2/O2 reactors (hand-written mechanism, K={M.K}), 0 -> 1 ms through ignition"
    else:
        d = np.load(os.path.join(golden, "reactor_gri53.npz"))
        M = Mechanism(synth.mechanism_path(mechanism))
        n, spread, t_end, warm = 1 << 14, 0.02, 1e-4, 1e-9
        what = f"constant-pressure fuel/O2/N2 reactors ({mechanism} mechanism, K={M.K}), 0 -> 0.1 ms"
    reps = -(-n // len(d["T0"]))
    rng = np.random.default_rng(16)
    T0 = np.tile(d["T0"], reps)[:n] * rng.uniform(1 - spread, 1 + spread, n)
    Y0 = np.tile(d["Y0"], (1, reps))[:, :n]
    P = np.tile(d["P"], reps)[:n]
    integ = Integrator(M, rtol=1e-8, atol_T=1e-6, atol_Y=1e-12, jacobian=jacobian, method=method)

    def run(t_end):
        y = torch.from_numpy(np.vstack([T0, Y0])).to(dev).contiguous()
        Pd = torch.from_numpy(P).to(dev)
        t = torch.zeros(n, dtype=torch.float64, device=dev)
        h = torch.full((n,), 1e-9, dtype=torch.float64, device=dev)
        status = torch.zeros(n, dtype=torch.int32, device=dev)
        steps = torch.zeros(n, dtype=torch.int32, device=dev)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        taken = integ.advance(y, Pd, t, h, status, steps, t_end, max_steps=200000)
        torch.cuda.synchronize()
        return time.perf_counter() - t0, taken, status, steps, y

    run(warm)  # warm-up: module loads
    wall, taken, status, steps, y = run(t_end)
    m = M.K + 1
    rhs = M.reactor(lean=True)  # the shape the integrator compiles
    used = Integrator.resolved_jacobian(M, jacobian)
    if used == "differences":
        # the right-hand side split at the mass fractions: the shared part once, the varying part for the state and
        # its K copies, the whole program at the raised temperature; then a subtraction and a product per entry
        shared, varying = rhs.split(2)
        jac_flop = (shared.op_counts()["weighted"] + m * varying.op_counts()["weighted"] + rhs.op_counts()["weighted"]
                    + 2 * m * m)
    else:
        jac_flop = rhs.jacobian([0] + [2 + k for k in range(M.K)]).op_counts()["weighted"]
    stages = 3 if method == "ode23s" else 6
    flop_step = (stages - 1) * rhs.op_counts()["weighted"] + jac_flop + (2 * m ** 3) // 3 + stages * 2 * m * m
    reactor_steps = int(steps.sum().item())
    rise = (y[0] - torch.from_numpy(T0).to(dev)).cpu().numpy()
    return {"workload": f"{n} {what}, rtol 1e-8", "method": method, "jacobian": used, "reactors": n, "batch_steps": taken, "wall_ms": wall * 1e3,
            "reactor_steps": reactor_steps, "reactor_steps_per_s": reactor_steps / wall,
            "steps_per_reactor": {"min": int(steps.min().item()), "max": int(steps.max().item())},
            "reached_t_end": float((status == 1).float().mean().item()),
            "temperature_rise_K": {"min": float(rise.min()), "max": float(rise.max())},
            "flop_per_reactor_step": flop_step,
            "achieved_tflops": flop_step * reactor_steps / wall / 1e12}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--states", type=int, default=1 << 20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--surrogate", default=SURROGATE)
    ap.add_argument("--error-sample", type=int, default=8192)
    ap.add_argument("--only", choices=["rates", "transport", "fp64", "mixed", "fast", "surrogate", "copy", "integrator"])
    ap.add_argument("--copy-mib", type=int, default=1024)
    ap.add_argument("--mechanism", default=MECHANISM, help="for --only: gri-53 (default) or large-160")
    ap.add_argument("--no-large", action="store_true", help="leave out the 160-species workload")
    ap.add_argument("--no-integrator", action="store_true", help="leave out the batched reactor integrator")
    ap.add_argument("--integrator-jacobian", default="automatic", choices=["automatic", "analytic", "differences"],
                    help="for --only integrator")
    ap.add_argument("--integrator-mechanism", default="handwritten", choices=["handwritten", "gri-53"],
                    help="for --only integrator")
    ap.add_argument("--integrator-method", default="rodas4", choices=["rodas4", "ode23s"], help="for --only integrator")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU: every state is independent, so N GPUs would each run "
                                   "the 1-GPU workload on their own shard with no communication; only --gpus 1 is measured"}))
        sys.exit(2)

    import torch
    from combustion_amd import Mechanism, synth
    from combustion_amd import surrogate as sur

    if not torch.cuda.is_available():
        raise SystemExit("bench.py needs a GPU")
    dev = torch.device("cuda:0")
    name = args.mechanism if args.only else MECHANISM
    M = Mechanism(synth.mechanism_path(name))
    Ka = M.active
    N = args.states
    T, P, Y = synth.config_states(name, N, 2024)
    inputs = [torch.from_numpy(np.ascontiguousarray(x)).to(dev) for x in [T, P, *Y]]
    prog = M.chemistry()
    outputs = [torch.empty(N, dtype=torch.float64, device=dev) for _ in range(prog.n_outputs)]
    if args.only == "integrator":
        print(json.dumps({"only": "integrator", **integrator_workload(args, dev, args.integrator_jacobian,
                                                                     args.integrator_mechanism, args.integrator_method)}))
        return
    if args.only == "surrogate":
        from combustion_amd import reactor
        T, P, Y = reactor.trajectories(M, 2048, N, seed=4242, max_steps=800, batch=2048)
        inputs = [torch.from_numpy(np.ascontiguousarray(x)).to(dev) for x in [T, P, *Y]]
        mod = sur.assemble(prog, sur.Trained.load(args.surrogate))
        _, ms = time_module(mod, inputs, outputs, args.steps, 1)
        print(json.dumps({"only": "surrogate", "states": N, "steps": args.steps, "kernel_ms": ms, "fallback": sur.fallback_count(mod)}))
        return
    if args.only:
        return only(args, M, inputs, N)

    # The plain FP64 form: the whole step, then each of its two kernels alone (rates; transport) for the roofline.
    fp64 = prog.assemble(form="fp64")
    wall, step_ms = time_module(fp64, inputs, outputs, args.steps, args.warmup)
    assert all(torch.isfinite(o).all().item() for o in outputs)
    value = N * args.steps / wall
    sections = {}
    for name, flags in (("rates", 1), ("transport", 2)):
        p = M.chemistry(flags)
        mod = p.assemble(form="fp64")
        outs = [torch.empty(N, dtype=torch.float64, device=dev) for _ in range(p.n_outputs)]
        _, ms = time_module(mod, inputs, outs, max(5, args.steps // 2), 2)
        flop = p.op_counts()["weighted"]
        kernel = "rates_fp64" if flags == 1 else ("transport_fp64" if M.K <= 64 else "transport_tiled")
        sections[name] = {"kernel": kernel, "kernel_ms": ms,
                          "flop_per_state": flop, "achieved_tflops": flop * N / (ms * 1e-3) / 1e12}
        del outs
    dominant = max(sections.values(), key=lambda s: s["kernel_ms"])
    flop_total = prog.op_counts()["weighted"]
    hbm_bytes = 8 * N * (prog.n_inputs + prog.n_outputs)
    achieved = dominant["achieved_tflops"]
    result = {
        "metric": "states/s rates+transport, K=53/R=325: FP64 plain vs mixed+surrogate at stated err",
        "value": value, "unit": "states/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic (seeded mechanism and states)",
        "config": {"workload": "rates+transport, K=53 species / R=325 reactions (30 three-body, 25 Troe), 2^20 states",
                   "states": N, "form": "fp64", "mechanism_seed": synth.CONFIGS[MECHANISM][4]},
        "roofline": {"bound": "valu", "compute": "FP64 VALU issue: straight-line FP64 code without MFMA, against the FP64 "
                                                 "vector peak", "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP64_TFLOPS, "traffic": traffic_from_profiles(dominant["kernel"]),
                     "kernel": dominant["kernel"], "flop_per_state": dominant["flop_per_state"],
                     "kernel_ms": dominant["kernel_ms"],
                     "valu_issue": issue_from_profiles(dominant["kernel"], dominant["kernel_ms"]),
                     "step": {"kernel_ms": step_ms, "flop_per_state": flop_total,
                              "achieved": flop_total * N / (step_ms * 1e-3) / 1e12,
                              "frac": flop_total * N / (step_ms * 1e-3) / 1e12 / PEAK_FP64_TFLOPS,
                              "hbm_algorithmic_bytes": hbm_bytes, "hbm_frac": hbm_bytes / (step_ms * 1e-3) / 1e9 / PEAK_HBM_GBS},
                     "sections": sections},
    }
    forms = {"fp64": {"states_per_s": value, "kernel_ms": step_ms,
                      "error_vs_oracle": errors_vs_oracle(outputs, T, P, Y, args.error_sample, Ka, exact=True)}}
    # The same form with host buffers (the boundary's blocking host-pointer call: copies in, kernels, copies out).
    host_in = [np.ascontiguousarray(x) for x in [T, P, *Y]]
    host_out = [np.empty(N) for _ in range(prog.n_outputs)]
    fp64(host_in, host_out)
    t0 = time.perf_counter()
    for _ in range(3):
        fp64(host_in, host_out)
    host_s = (time.perf_counter() - t0) / 3
    result["pcie_inclusive"] = {"states_per_s": N / host_s, "ms_per_step": host_s * 1e3,
                                "note": "host numpy buffers through cb_hip_call (pageable memory, blocking); not the value"}

    # the nekRK export (cuda.rs:296-337): the four emitted files compiled by hiprtc and run as one kernel over the same
    # states (fits of length 5, the reference's export length), timed like the forms
    exported = M.export_module(5)
    wall_x, ms_x = time_module(exported, inputs, outputs, args.steps, args.warmup)
    result["nekrk_export"] = {"states_per_s": N * args.steps / wall_x, "kernel_ms": ms_x, "vs_fp64_form": wall / wall_x,
                              "note": "rates.c, conductivity.c, viscosity.c, diffusivity.c as emitted, one lane per state"}
    del exported
    for form in ("mixed", "fast"):
        mod = prog.assemble(form=form)
        wall_f, ms_f = time_module(mod, inputs, outputs, args.steps, args.warmup)
        forms[form] = {"states_per_s": N * args.steps / wall_f, "kernel_ms": ms_f, "speedup_vs_fp64": wall / wall_f,
                       "error_vs_oracle": errors_vs_oracle(outputs, T, P, Y, args.error_sample, Ka)}
        del mod
    if os.path.exists(args.surrogate):
        result["surrogate_workload"] = surrogate_workload(args, M, prog, dev)
    result["forms"] = forms
    del inputs, outputs, fp64
    torch.cuda.empty_cache()
    if not args.no_large:
        log("160-species workload")
        result["large_mechanism"] = large_mechanism_workload(args, dev)
    if not args.no_integrator:
        log("batched reactor integrator")
        result["integrator"] = integrator_workload(args, dev)
        log("batched reactor integrator, difference quotients")
        result["integrator_differences"] = integrator_workload(args, dev, "differences")
        log("batched reactor integrator, 53 species")
        result["integrator_gri53"] = integrator_workload(args, dev, mechanism="gri-53")
    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(prog, MECHANISM)
    print(json.dumps(result))


def only(args, M, inputs, N):
    """One launch kind alone, args.steps times (the command that rocprofv3 summaries in profiles/ are made of)."""
    import torch
    if args.only == "copy":
        src = torch.ones(args.copy_mib << 18, dtype=torch.float32, device=inputs[0].device)
        dst = torch.empty_like(src)
        for _ in range(args.steps):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"only": "copy", "bytes_read_per_launch": src.numel() * 4, "bytes_written_per_launch": dst.numel() * 4,
                          "steps": args.steps}))
        return
    flags = {"rates": 1, "transport": 2, "fp64": 3, "mixed": 3, "fast": 3}[args.only]
    p = M.chemistry(flags)
    mod = p.assemble(form=args.only if args.only in ("mixed", "fast") else "fp64")
    outs = [torch.empty(N, dtype=torch.float64, device=inputs[0].device) for _ in range(p.n_outputs)]
    wall, ms = time_module(mod, inputs, outs, args.steps, 1)
    print(json.dumps({"only": args.only, "mechanism": args.mechanism, "states": N, "steps": args.steps, "kernel_ms": ms,
                      "flop_per_state": p.op_counts()["weighted"],
                      "algorithmic_hbm_bytes_per_launch": 8 * N * (p.n_inputs + p.n_outputs)}))


if __name__ == "__main__":
    main()
