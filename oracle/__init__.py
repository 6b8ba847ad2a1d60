"""ORACLE — test infrastructure only.

CPU restatement of the reference's per-state chemistry, against which the tests compare the HIP path (and timed,
for information, beside the CPU baseline, which is the product's own host-compiled program). Only tests/,
__graft_entry__.smoke() and bench.py (its error samples and cpu_baseline leg) may import it; the product
(combustion_amd) never does.

  mechanism.py  Cantera-YAML front end (yaml.rs:40-171, lib.rs:22-140)
  chem.c        closed-form FP64 rates, NASA-7 and mixture-averaged transport (reaction.rs, transport.rs)
  interpret.c   FP64 interpreter of a program's linear form (ast/interpret.rs:52-148)
  fits.py       kinetic-theory species properties and polynomial fits (transport.rs:14-137)
  mlp.py        FP64 forward/backward of the surrogate network with the kernels' BF16 rounding points

The tests that pin each function are in tests/test_oracle_*.py.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from . import mechanism

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
SOURCES = [os.path.join(HERE, "chem.c"), os.path.join(HERE, "interpret.c")]


def build():
    tmp = LIB + f".tmp{os.getpid()}"
    subprocess.check_call(["gcc", "-O2", "-std=c11", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math", "-Wall",
                           *SOURCES, "-lm", "-lpthread", "-o", tmp])
    os.replace(tmp, LIB)
    return LIB


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB) or any(os.path.getmtime(s) > os.path.getmtime(LIB) for s in SOURCES if os.path.exists(s)):
            build()
        _lib = C.CDLL(LIB)
        _lib.oracle_batch.restype = C.c_int
        _lib.oracle_batch.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t,
                                      C.c_void_p, C.c_void_p, C.c_int]
        _lib.oracle_reaction_rates.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_void_p, C.c_void_p, C.c_void_p]
        _lib.oracle_rhs.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                    C.c_void_p, C.c_void_p]
        _lib.interpret.restype = C.c_long
        _lib.interpret.argtypes = [C.c_void_p] * 6 + [C.c_long, C.c_int32, C.c_int32, C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p,
                                                      C.c_long, C.POINTER(C.c_long)]
    return _lib


class _Mech(C.Structure):
    _fields_ = [("K", C.c_int), ("Ka", C.c_int), ("R", C.c_int), ("D", C.c_int)] + \
               [(n, C.c_void_p) for n in ("W", "split", "nasa", "reactants", "products", "model", "rate", "k0", "troe", "eff",
                                          "fit_c", "fit_v", "fit_d")] + [("T0", C.c_double), ("sri", C.c_void_p), ("plog_offset", C.c_void_p), ("plog", C.c_void_p)]


def _ptr(a):
    return a.ctypes.data if a is not None else None


class Oracle:
    """Closed-form FP64 evaluation of a mechanism (dict from mechanism.load) with given transport fits."""

    def __init__(self, mech, fits=None, T0=1000.0):
        self.m = mech
        keep = {
            "W": np.ascontiguousarray(mech["W"], np.float64), "split": np.ascontiguousarray(mech["split"], np.float64),
            "nasa": np.ascontiguousarray(mech["nasa"], np.float64), "reactants": np.ascontiguousarray(mech["reactants"], np.int32),
            "products": np.ascontiguousarray(mech["products"], np.int32), "model": np.ascontiguousarray(mech["model"], np.int32),
            "rate": np.ascontiguousarray(mech["rate"], np.float64), "k0": np.ascontiguousarray(mech["k0"], np.float64),
            "troe": np.ascontiguousarray(mech["troe"], np.float64), "eff": np.ascontiguousarray(mech["eff"], np.float64),
            "sri": np.ascontiguousarray(mech["sri"], np.float64),
            "plog_offset": np.ascontiguousarray(np.cumsum([0] + [len(p) for p in mech["plog"]]), np.int32),
            "plog": np.ascontiguousarray(np.array([row for p in mech["plog"] for row in p] or [[0.0] * 4], np.float64)),
        }
        D = 0
        if fits is not None:
            c, v, d = fits
            D = c.shape[1]
            keep.update(fit_c=np.ascontiguousarray(c, np.float64), fit_v=np.ascontiguousarray(v, np.float64),
                        fit_d=np.ascontiguousarray(d, np.float64))
        self._keep = keep
        self.s = _Mech(mech["K"], mech["Ka"], mech["R"], D, *[_ptr(keep.get(n)) for n in (
            "W", "split", "nasa", "reactants", "products", "model", "rate", "k0", "troe", "eff", "fit_c", "fit_v", "fit_d")], T0,
            _ptr(keep["sri"]), _ptr(keep["plog_offset"]), _ptr(keep["plog"]))
        self.has_transport = fits is not None

    def batch(self, T, P, Y, rates=True, transport=True, threads=1):
        """Outputs [ω (Ka), λ, μ, D (K)] (sections as requested) and their error floors, each [n_out][N]."""
        if transport and not self.has_transport:
            raise ValueError("no transport fits")
        T, P, Y = (np.ascontiguousarray(x, np.float64) for x in (T, P, Y))
        n = len(T)
        n_out = (self.m["Ka"] if rates else 0) + ((2 + self.m["K"]) if transport else 0)
        out, scale = np.zeros((n_out, n)), np.zeros((n_out, n))
        op = (C.c_void_p * n_out)(*[out[i].ctypes.data for i in range(n_out)])
        sp = (C.c_void_p * n_out)(*[scale[i].ctypes.data for i in range(n_out)])
        status = lib().oracle_batch(C.byref(self.s), int(rates), int(transport), T.ctypes.data, P.ctypes.data, Y.ctypes.data, n,
                                    op, sp, int(threads))
        if status:
            raise RuntimeError("oracle_batch failed")
        return out, scale

    def reaction_rates(self, T, P, Y, scale=False):
        """Rates of progress q_r of one state (and their scales |k_f| (R_f + R_r) when scale is set)."""
        q, qs = np.zeros(self.m["R"]), np.zeros(self.m["R"])
        Y = np.ascontiguousarray(Y, np.float64)
        lib().oracle_reaction_rates(C.byref(self.s), float(T), float(P), Y.ctypes.data, q.ctypes.data, qs.ctypes.data)
        return (q, qs) if scale else q

    def rhs(self, T, P, Y):
        omega, dT, _ = self.rhs_full(T, P, Y)
        return omega, dT

    def rhs_full(self, T, P, Y):
        """(ω, dT/dt, extras) of one state; extras holds the term scales of ω (omega_scale), the scale of dT/dt
        (dTdt_scale), the relative volume rate (dV/dt)/V at constant pressure (dlnV) and its scale (dlnV_scale)."""
        omega, scale = np.zeros(self.m["Ka"]), np.zeros(self.m["Ka"])
        v = [C.c_double() for _ in range(4)]
        Y = np.ascontiguousarray(Y, np.float64)
        lib().oracle_rhs(C.byref(self.s), float(T), float(P), Y.ctypes.data, omega.ctypes.data, scale.ctypes.data, *[C.byref(x) for x in v])
        dT, dT_scale, dlnV, dlnV_scale = (x.value for x in v)
        return omega, dT, {"omega_scale": scale, "dTdt_scale": dT_scale, "dlnV": dlnV, "dlnV_scale": dlnV_scale}


def interpret(linear, inputs):
    """Runs a program's linear form (combustion_amd Program.linear()) on SoA inputs [n_inputs][N]."""
    inputs = np.ascontiguousarray(inputs, np.float64)
    n = inputs.shape[1]
    out = np.zeros((len(linear["outputs"]), n))
    nan_at = C.c_long(-1)
    L = {k: np.ascontiguousarray(v) for k, v in linear.items() if isinstance(v, np.ndarray)}
    status = lib().interpret(L["op"].ctypes.data, L["dst"].ctypes.data, L["a"].ctypes.data, L["b"].ctypes.data, L["c"].ctypes.data,
                             L["k"].ctypes.data, len(L["op"]), linear["n_inputs"], linear["n_registers"], L["outputs"].ctypes.data,
                             len(L["outputs"]), inputs.ctypes.data, out.ctypes.data, n, C.byref(nan_at))
    if status > 0:
        raise FloatingPointError(f"NaN in state {status - 1} at instruction {nan_at.value}")
    if status < 0:
        raise RuntimeError("interpret failed")
    return out
