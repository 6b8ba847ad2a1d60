"""Surrogate network kernels (kernels/mlp.hip) against the CPU oracle (oracle/mlp.py, which reproduces the
kernels' BF16 rounding points): forward outputs and the gradients of the mean squared error.

The oracle sums in FP64 and the kernels in FP32, so an activation that lies within an FP32 rounding of a BF16
rounding boundary rounds the other way in one of them (about 1 in 10^5 activations of the first hidden layer,
growing through the layers); such a flip moves that activation by one BF16 unit (2^-8 relative). The bars are
therefore on the 99th percentile (1e-3 of the largest value: a systematic error such as a wrong operand layout
moves every element by O(1)) and on the maximum (2e-2, a handful of flips)."""
import numpy as np
import pytest

from oracle import mlp as omlp


def run_network(n, seed, gpu):
    import torch
    from combustion_amd.native import Network
    net = Network(55, 256, 3, 55)
    rng = np.random.default_rng(seed)
    params = np.zeros(net.n_params, np.float32)
    for w, b, o, i in net.layers:
        fan_in = 55 if w == 0 else i
        W = rng.normal(0, 1.0 / np.sqrt(fan_in), (o, i)).astype(np.float32)
        if w == 0:
            W[:, 55:] = 0
        params[w:w + o * i] = W.ravel()
        params[b:b + o] = rng.normal(0, 0.1, o)
    w, b, o, i = net.layers[-1]
    params[w:w + o * i].reshape(o, i)[55:] = 0
    params[b + 55:b + o] = 0
    x = np.zeros((n, net.in_pad), np.float32)
    x[:, :55] = rng.normal(size=(n, 55))
    t = np.zeros((n, net.out_pad), np.float32)
    t[:, :55] = rng.normal(size=(n, 55))
    p = torch.from_numpy(params).to(gpu)
    p16 = p.to(torch.bfloat16)
    xd, td = torch.from_numpy(x).to(gpu), torch.from_numpy(t).to(gpu)
    y = torch.empty((n, net.out_pad), dtype=torch.float32, device=gpu)
    ws = torch.empty(net.workspace_bytes(n), dtype=torch.uint8, device=gpu)
    grad = torch.zeros(net.n_params, dtype=torch.float32, device=gpu)
    net.forward(p, p16, xd, y, ws)
    net.backward_mse(p16, y, td, ws, grad)
    torch.cuda.synchronize()
    return net, params, x, t, y.cpu().numpy(), grad.cpu().numpy()


@pytest.mark.gpu
@pytest.mark.parametrize("n", [2000, 4096])
def test_forward_and_gradients_match_oracle(n, gpu):
    net, params, x, t, y, grad = run_network(n, 3, gpu)
    ref_grad, ref_y = omlp.backward_mse(params, net.layers, x, t, 55)  # rounds W to BF16 as the kernels' copy does
    def check(name, got, ref):
        e = np.abs(got - ref) / np.abs(ref).max()
        p99, mx = np.percentile(e, 99), e.max()
        print(f"n={n} {name}: 99th percentile {p99:.2e}, max {mx:.2e}", flush=True)
        assert p99 < 1e-3 and mx < 2e-2

    check("output", y[:, :55], ref_y[:, :55])
    assert np.all(y[:, 55:] == 0)
    for l, (w, b, o, i) in enumerate(net.layers):
        check(f"layer {l} dW", grad[w:w + o * i], ref_grad[w:w + o * i])
        check(f"layer {l} db", grad[b:b + o], ref_grad[b:b + o])


@pytest.mark.gpu
def test_adam_steps_reduce_the_loss(gpu):
    """A few hundred Adam steps of the project's kernels fit a smooth 55 -> 55 map."""
    import torch
    from combustion_amd.native import Network
    net = Network(55, 256, 3, 55)
    rng = np.random.default_rng(5)
    n = 4096
    x = np.zeros((n, net.in_pad), np.float32)
    x[:, :55] = rng.uniform(-1, 1, (n, 55))
    A = rng.normal(0, 0.3, (55, 55))
    t = np.zeros((n, net.out_pad), np.float32)
    t[:, :55] = np.tanh(x[:, :55] @ A) + 0.5 * np.sin(2 * x[:, :55])
    params = np.zeros(net.n_params, np.float32)
    for w, b, o, i in net.layers:
        params[w:w + o * i] = rng.normal(0, 1.0 / np.sqrt(i), o * i)
    gpu_t = lambda a: torch.from_numpy(a).to(gpu)
    p, xd, td = gpu_t(params), gpu_t(x), gpu_t(t)
    p16 = p.to(torch.bfloat16)
    m1, m2, grad = torch.zeros_like(p), torch.zeros_like(p), torch.zeros_like(p)
    y = torch.empty((n, net.out_pad), dtype=torch.float32, device=gpu)
    ws = torch.empty(net.workspace_bytes(n), dtype=torch.uint8, device=gpu)
    losses = []
    for step in range(1, 301):
        net.forward(p, p16, xd, y, ws)
        if step in (1, 300):
            losses.append(float(((y[:, :55] - td[:, :55]) ** 2).mean()))
        net.backward_mse(p16, y, td, ws, grad)
        net.adam(p, p16, grad, m1, m2, 1e-3, step)
    torch.cuda.synchronize()
    print("loss", losses, flush=True)
    assert losses[1] < 0.1 * losses[0]
