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
