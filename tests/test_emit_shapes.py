"""The shape of the emitted straight-line HIP for each form (CPU only: text checks, no compilation): which FP64
building blocks and storage each kernel uses. The numerics of those blocks are checked on the GPU
(test_gpu_arithmetic.py, test_gpu_parity.py)."""
import re

from helpers import setup


def sections(name, form="fp64"):
    M, mech, o = setup(name)
    text = M.chemistry(3).hip_source(form)
    rates, transport = text.split("// ---- transport ----")
    return rates, transport


def test_fp64_rates_kernel_building_blocks():
    rates, transport = sections("gri-53")
    assert "exp64(" in rates and " exp(" not in rates  # the inlined polynomial exponential
    rcp = rates[rates.index("double rcp64"):rates.index("}\n", rates.index("double rcp64"))]
    assert rcp.count("__builtin_fma(-b, r, 1.0)") == 1  # one Newton step after v_rcp_f64
    # 291 live doubles at K=53: the clamped exp(-G/RT) values of the K_a = 52 active species live in LDS
    m = re.search(r"__shared__ double lds_v\[(\d+)\];", rates)
    assert m and int(m.group(1)) == 52 * 256
    assert rates.count('asm volatile("" ::: "memory");') > 10
    assert "lds_v" not in transport


def test_small_mechanisms_keep_their_values_in_registers():
    rates, transport = sections("h2o2-10")
    assert "lds_v" not in rates and "lds_v" not in transport


def test_transport_kernel_is_issue_lean():
    rates, transport = sections("gri-53")
    assert "__launch_bounds__(256, 2)" in transport  # two waves per SIMD
    # one reciprocal per four diffusion pairs (1378 pairs) plus four per species (1/ṽ_k, x_k/λ_k, Wilke's row
    # quotient, D_k's quotient) and a few per state
    n = transport.count("rcp64(")
    assert 1378 // 4 <= n <= 1378 // 4 + 4 * 53 + 10, n


def test_exact_form_keeps_library_arithmetic():
    M, mech, o = setup("h2o2-10")
    text = M.chemistry(3).hip_source("straight")
    assert "exp64(" not in text and "rcp64(" not in text and "lds_v" not in text
