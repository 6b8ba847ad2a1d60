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


def test_large_mechanisms_take_the_table_driven_rates_kernel():
    """Above 64 species the FP64 and FP32-core forms take rates_tiled (tables, not straight-line code) and the MFMA
    transport kernel; the modifier selects rates_tiled at any size."""
    M, mech, o = setup("large-160")
    for form in ("fp64", "fast"):
        text = M.chemistry(3).hip_source(form)
        assert "rates_tiled(" in text and "transport_tiled(" in text and "rates_fp64" not in text
        assert "#define K 160\n" in text and "#define KA 159\n" in text
    M, mech, o = setup("gri-53")
    text = M.chemistry(3).hip_source("fp64+tiled-rates")
    assert "rates_tiled(" in text and "rates_fp64" not in text
    assert "rates_fp64" in M.chemistry(3).hip_source("fp64")


def test_table_driven_rates_kernel_refuses_what_it_has_no_tables_for():
    import pytest
    from combustion_amd import native
    from combustion_amd.synth import variant_text
    from helpers import setup_text
    for variant in ("sri", "plog"):
        M, mech, o = setup_text(variant_text(variant, "h2o2-10"))
        with pytest.raises(native.Error, match="tiled FP64 rates kernel has no"):
            M.chemistry(3).hip_source("fp64+tiled-rates")
    M, mech, o = setup("h2o2-10")
    with pytest.raises(native.Error, match="mixed form has its own rates kernel"):
        M.chemistry(3).hip_source("mixed+tiled-rates")
    with pytest.raises(ValueError):
        M.chemistry(3).hip_source("fp64+rates")
