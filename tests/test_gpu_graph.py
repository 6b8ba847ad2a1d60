"""A module captured in a HIP graph on the caller's stream (torch.cuda.graph), fork and join of its second stream
included, replayed several times with eager work and synchronizations between the replays: every replay writes
the eager call's outputs exactly."""
import numpy as np
import pytest

from combustion_amd import synth
from helpers import setup


@pytest.mark.gpu
@pytest.mark.parametrize("form", ["fp64", "fast"])
def test_graph_replays_match_eager(form, gpu):
    import torch
    M, mech, o = setup("gri-53")
    n = 4096
    T, P, Y = synth.config_states("gri-53", n, 41)
    ins = [torch.from_numpy(np.ascontiguousarray(x)).to(gpu) for x in (T, P, *Y)]
    mod = M.chemistry(3).assemble(form=form)
    outs = [torch.empty(n, dtype=torch.float64, device=gpu) for _ in range(M.active + M.K + 2)]
    s = torch.cuda.Stream()
    mod(ins, outs, stream=s.cuda_stream)  # eager call: uploads the module's pointer table for these arrays
    torch.cuda.synchronize()
    eager = torch.stack(outs).clone()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        mod(ins, outs, stream=s.cuda_stream)
    torch.cuda.synchronize()
    for replay in range(4):
        for x in outs:
            x.fill_(float("nan"))
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        got = torch.stack(outs)
        assert torch.equal(got, eager), (form, replay)
        _ = (ins[0] * 2).sum().item()  # eager work of the program's own between replays


@pytest.mark.gpu
def test_graph_keeps_its_arrays_after_eager_calls_with_others(gpu):
    """A module keeps one device pointer table per set of arrays and never rewrites one: a graph captured for one set
    still reads and writes that set after eager calls of the same module on other arrays (on another stream too)."""
    import torch
    M, mech, o = setup("gri-53")
    n = 2048
    mod = M.chemistry(3).assemble(form="fp64")
    s = torch.cuda.Stream()
    t = torch.cuda.Stream()
    arrays = []
    for seed in (42, 43):
        T, P, Y = synth.config_states("gri-53", n, seed)
        ins = [torch.from_numpy(np.ascontiguousarray(x)).to(gpu) for x in (T, P, *Y)]
        arrays.append((ins, [torch.empty(n, dtype=torch.float64, device=gpu) for _ in range(M.active + M.K + 2)]))
    (ins_a, out_a), (ins_b, out_b) = arrays
    mod(ins_a, out_a, stream=s.cuda_stream)
    torch.cuda.synchronize()
    eager_a = torch.stack(out_a).clone()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        mod(ins_a, out_a, stream=s.cuda_stream)
    torch.cuda.synchronize()
    mod(ins_b, out_b, stream=t.cuda_stream)  # another set of arrays, another stream
    torch.cuda.synchronize()
    eager_b = torch.stack(out_b).clone()
    assert not torch.equal(eager_a, eager_b)
    for replay in range(3):
        for x in out_a + out_b:
            x.fill_(float("nan"))
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(torch.stack(out_a), eager_a), replay
        assert torch.isnan(torch.stack(out_b)).all(), replay  # the graph never touched the other arrays
        mod(ins_b, out_b, stream=t.cuda_stream)
        torch.cuda.synchronize()
        assert torch.equal(torch.stack(out_b), eager_b), replay
