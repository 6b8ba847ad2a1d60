"""cb_hip_call_static, the allocation-free variant of cb_hip_call (no counterpart in the reference): the arrays'
pointer table and the surrogate form's fallback counter and state list come from the library's static device array
instead of memory the module allocates. Called through the C ABI on device arrays and checked against cb_hip_call
on the same module: two calls in a row, the second on fewer states, for the FP64 form (two kernels, run here on the
caller's stream alone) and the surrogate form (a fallback list in the static array)."""
import ctypes as C
import os

import numpy as np
import pytest

from combustion_amd import native, surrogate, synth
from helpers import setup

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TRAINED = os.path.join(ROOT, "combustion_amd", "data", "surrogate_gri53.npz")


def static_call(mod, inputs, outputs, stream=None):
    f = native.lib().cb_hip_call_static
    ins = (C.c_void_p * len(inputs))(*[x.data_ptr() for x in inputs])
    outs = (C.c_void_p * len(outputs))(*[x.data_ptr() for x in outputs])
    return f(mod._h, ins, len(inputs), len(inputs[0]), outs, len(outputs), stream)


def states(name, n, seed, gpu):
    import torch
    T, P, Y = synth.config_states(name, n, seed)
    # a few states far outside the surrogate's training box, so that its fallback list is used
    T[: n // 64] = 5000.0
    return [torch.from_numpy(np.ascontiguousarray(x)).to(gpu) for x in (T, P, *Y)]


def compare(mod, n_outputs, gpu, sizes=(8192, 3000), seed=81):
    import torch
    for i, n in enumerate(sizes):
        ins = states("gri-53", n, seed + i, gpu)
        ref = [torch.empty(n, dtype=torch.float64, device=gpu) for _ in range(n_outputs)]
        got = [torch.full((n,), float("nan"), dtype=torch.float64, device=gpu) for _ in range(n_outputs)]
        mod(ins, ref)
        torch.cuda.synchronize()
        status = static_call(mod, ins, got)
        assert status == 0, native.lib().cb_last_error()
        torch.cuda.synchronize()
        print("static call", i, "states", n, flush=True)
        assert torch.equal(torch.stack(got), torch.stack(ref)), (i, n)


@pytest.mark.gpu
def test_static_call_matches_cb_hip_call_fp64(gpu):
    M, mech, o = setup("gri-53")
    prog = M.chemistry()
    compare(prog.assemble(form="fp64"), prog.n_outputs, gpu)


@pytest.mark.gpu
def test_static_call_matches_cb_hip_call_surrogate(gpu):
    M, mech, o = setup("gri-53")
    prog = M.chemistry()
    mod = surrogate.assemble(prog, surrogate.Trained.load(TRAINED))
    compare(mod, prog.n_outputs, gpu)
    assert surrogate.fallback_count(mod) >= 3000 // 64  # the cb_hip_call of the last pair used its own list too


@pytest.mark.gpu
def test_static_call_rejects_sizes_beyond_the_scratch(gpu):
    import torch
    M, mech, o = setup("gri-53")
    prog = M.chemistry()
    mod = surrogate.assemble(prog, surrogate.Trained.load(TRAINED))
    n = (1 << 21) + 1
    x = torch.zeros(n, dtype=torch.float64, device=gpu)  # the size check comes before any array is touched
    assert static_call(mod, [x] * prog.n_inputs, [x] * prog.n_outputs) == -1
    assert b"2^21" in native.lib().cb_last_error()
