"""The Rosenbrock coefficients of the GPU integrator (combustion_amd/csrc/kernels/rosenbrock.hip), checked on the
CPU against what the mathematics fixes: the kernel's own rodas_a/rodas_c functions are compiled by the host compiler
and the method its header states (W = I - h γ J; k_I = W⁻¹ γ (h f(Y_I) + Σ_{j<I} c_Ij k_j); Y_I = y + Σ_{j<I} a_Ij k_j;
y_new = Y_6 + k6, error estimate k6) is run in numpy:
- global order 4 on the Van der Pol oscillator (μ = 5, exact Jacobian) against scipy's DOP853 at rtol 1e-13;
- the embedded solution Y_6 of order 3: its local error falls as h^4;
- stiff accuracy: one step of y' = λ y with h λ = -1e10 leaves |y| below 1e-8 |y0|;
and ode23s's two constants d = 1/(2+√2) and e32 = 6+√2. A mistyped digit of the tableau breaks an order condition,
which the convergence orders show."""
import math
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SOURCE = os.path.join(ROOT, "combustion_amd", "csrc", "kernels", "rosenbrock.hip")


@pytest.fixture(scope="module")
def tableau(tmp_path_factory):
    """(γ, a, c) from the kernel source: a[i][j], c[i][j] for stages i = 2..6, j < i (a_6j as the kernel uses it)."""
    text = open(SOURCE).read()
    functions = [re.search(r"constexpr double %s\(int i, int j\) \{.*?\n\}" % name, text, re.S).group(0)
                 for name in ("rodas_a", "rodas_c")]
    gamma = float(re.search(r"#define GAMMA (0\.[0-9]+)", text).group(1))
    program = "#include <cstdio>\n" + "\n".join(functions) + """
int main() {
  for (int i = 2; i <= 6; i++)
    for (int j = 1; j < i; j++) std::printf("%d %d %.17g %.17g\\n", i, j, rodas_a(i, j), rodas_c(i, j));
}
"""
    d = tmp_path_factory.mktemp("tableau")
    src, exe = d / "tableau.cpp", d / "tableau"
    src.write_text(program)
    subprocess.run(["g++", "-std=c++17", "-O0", "-o", str(exe), str(src)], check=True)
    a, c = np.zeros((7, 7)), np.zeros((7, 7))
    for line in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n"):
        if line:
            i, j, av, cv = line.split()
            a[int(i), int(j)], c[int(i), int(j)] = float(av), float(cv)
    return gamma, a, c


def step(tab, f, jac, y, h):
    """One step as the kernel's phases take it: (y_new, Y_6, error estimate k6)."""
    gamma, a, c = tab
    W = np.eye(len(y)) - h * gamma * jac(y)
    k = {1: np.linalg.solve(W, gamma * h * f(y))}
    for i in range(2, 7):
        Yi = y + sum(a[i, j] * k[j] for j in range(1, i))
        k[i] = np.linalg.solve(W, gamma * (h * f(Yi) + sum(c[i, j] * k[j] for j in range(1, i))))
    Y6 = y + sum(a[6, j] * k[j] for j in range(1, 6))
    return Y6 + k[6], Y6, k[6]


MU = 5.0


def vdp(y):
    return np.array([y[1], MU * (1 - y[0] ** 2) * y[1] - y[0]])


def vdp_jac(y):
    return np.array([[0.0, 1.0], [-2 * MU * y[0] * y[1] - 1, MU * (1 - y[0] ** 2)]])


def exact(y0, t):
    from scipy.integrate import solve_ivp
    return solve_ivp(lambda _, y: vdp(y), (0, t), y0, method="DOP853", rtol=1e-13, atol=1e-15).y[:, -1]


def test_rodas4_converges_with_order_4(tableau):
    y0, T = np.array([2.0, 0.0]), 1.0
    ref = exact(y0, T)
    errors = []
    for N in (20, 40, 80):
        y = y0.copy()
        for _ in range(N):
            y = step(tableau, vdp, vdp_jac, y, T / N)[0]
        errors.append(np.abs(y - ref).max())
    orders = [math.log2(errors[i] / errors[i + 1]) for i in range(len(errors) - 1)]
    print("global errors", errors, "orders", orders)
    assert min(orders) > 3.8


def test_embedded_solution_has_order_3(tableau):
    y0 = np.array([2.0, 0.0])
    local = []
    for h in (0.01, 0.005):
        _, Y6, k6 = step(tableau, vdp, vdp_jac, y0, h)
        local.append(np.abs(Y6 - exact(y0, h)).max())
        assert np.allclose(Y6 + k6, step(tableau, vdp, vdp_jac, y0, h)[0])
    ratio = local[0] / local[1]
    print("local errors of Y_6", local, "ratio", ratio)
    assert ratio > 12  # h^4: 16 in the limit; an order-2 embedded solution gives 8


def test_rodas4_damps_stiff_components(tableau):
    lam = -1e10
    y1 = step(tableau, lambda y: lam * y, lambda y: np.array([[lam]]), np.array([1.0]), 1.0)[0]
    assert abs(y1[0]) < 1e-8


def test_ode23s_constants():
    text = open(SOURCE).read()
    d = float(re.search(r"#define D_ROS ([0-9.]+)", text).group(1))
    e32 = float(re.search(r"#define E32 ([0-9.]+)", text).group(1))
    assert d == pytest.approx(1 / (2 + math.sqrt(2)), rel=1e-15)
    assert e32 == pytest.approx(6 + math.sqrt(2), rel=1e-15)
