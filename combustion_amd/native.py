"""ctypes binding of include/combustion_amd.h (libcombustion_amd.so, built in place by combustion_amd.build).

Mirrors the reference's host API on top of the C ABI:
  parse/new            yaml.rs:54 + lib.rs:127        -> Mechanism(path or text)
  reaction::rates      reaction.rs:356                -> Mechanism.rates()
  transport::properties transport.rs:212              -> Mechanism.transport(D, T0, viscosity, conductivity)
  device::assemble     device.rs:25,75                -> Program.assemble(block_size, form) -> Module
  (the returned Fn)    device.rs:59,98-100            -> Module(inputs) -> outputs
The library is loaded once; a missing library raises instead of falling back to anything else.
"""
from __future__ import annotations

import ctypes as C
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcombustion_amd.so")

RATES, TRANSPORT, RHS, LEAN = 1, 2, 4, 8
FORM_STRAIGHT, FORM_FP64, FORM_MIXED, FORM_FAST = 0, 1, 2, 3
FORMS = {"straight": FORM_STRAIGHT, "fp64": FORM_FP64, "mixed": FORM_MIXED, "fast": FORM_FAST}
# cb_form_flags: "fp64+tiled" sends the transport section through the MFMA kernel, "fp64+tiled-rates" the rates
# section through the table-driven kernel of large mechanisms, "fp64+tiled+tiled-rates" both.
FORM_TILED_TRANSPORT, FORM_TILED_RATES = 0x10, 0x20
MODIFIERS = {"tiled": FORM_TILED_TRANSPORT, "tiled-rates": FORM_TILED_RATES}


def form_code(form):
    """cb_hip_assemble's form argument from a name: "straight", "fp64", "mixed", "fast", with "+tiled" and/or
    "+tiled-rates" modifiers."""
    if isinstance(form, int):
        return form
    base, *mods = form.split("+")
    if base not in FORMS or any(m not in MODIFIERS for m in mods):
        raise ValueError(f"unknown form {form!r}")
    code = FORMS[base]
    for m in mods:
        code |= MODIFIERS[m]
    return code

STATUS = {0: "ok", -1: "argument", -2: "parse", -3: "unsupported", -4: "build", -5: "compile", -6: "device", -7: "size"}


class Error(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"{STATUS.get(status, status)}: {message}")
        self.status = status
        self.kind = STATUS.get(status, str(status))


_lib = None
_P = C.c_void_p
_I32P = C.POINTER(C.c_int32)
_I64P = C.POINTER(C.c_int64)
_DP = C.POINTER(C.c_double)
_DPP = C.POINTER(C.c_void_p)

# name -> (restype, argtypes)
SIGNATURES = {
    "cb_last_error": (C.c_char_p, []),
    "cb_version": (C.c_char_p, []),
    "cb_mechanism_load_yaml": (C.c_int, [C.c_char_p, C.POINTER(_P)]),
    "cb_mechanism_parse_yaml": (C.c_int, [C.c_char_p, C.POINTER(_P)]),
    "cb_mechanism_free": (None, [_P]),
    "cb_mechanism_sizes": (C.c_int, [_P, _I32P, _I32P, _I32P]),
    "cb_mechanism_json": (C.c_int, [_P, C.c_char_p, C.c_size_t, C.POINTER(C.c_size_t)]),
    "cb_ignition_state_load": (C.c_int, [_P, C.c_char_p, _DP, _DP, _DP]),
    "cb_ignition_state_parse": (C.c_int, [_P, C.c_char_p, _DP, _DP, _DP]),
    "cb_transport_fits": (C.c_int, [_P, C.c_int32, C.c_double, _DP, _DP, _DP]),
    "cb_transport_species": (C.c_int, [_P, C.c_int32, C.c_int32, C.c_int32, C.c_double, _DP]),
    "cb_transport_scales": (C.c_int, [_P, C.c_double, _DP, _DP]),
    "cb_build_rates": (C.c_int, [_P, C.POINTER(_P)]),
    "cb_build_reactions_rates": (C.c_int, [_P, C.POINTER(_P)]),
    "cb_build_transport": (C.c_int, [_P, C.c_int32, C.c_double, C.c_double, C.c_double, C.POINTER(_P)]),
    "cb_build_chemistry": (C.c_int, [_P, C.c_int32, C.c_int32, C.c_double, C.c_double, C.c_double, C.POINTER(_P)]),
    "cb_build_reactor": (C.c_int, [_P, C.c_int32, C.POINTER(_P)]),
    "cb_build_jacobian": (C.c_int, [_P, C.c_int32, _I32P, C.c_int32, _I32P, C.POINTER(_P)]),
    "cb_program_split": (C.c_int, [_P, C.c_int32, C.POINTER(_P), C.POINTER(_P)]),
    "cb_program_free": (None, [_P]),
    "cb_program_sizes": (C.c_int, [_P, _I32P, _I32P, _I64P, _I64P, _I32P]),
    "cb_program_op_counts": (C.c_int, [_P, _I64P]),
    "cb_program_linear": (C.c_int, [_P, C.POINTER(C.c_uint8), _I32P, _I32P, _I32P, _I32P, _DP, _I32P]),
    "cb_program_from_linear": (C.c_int, [C.c_int32, C.c_int64, C.POINTER(C.c_uint8), _I32P, _I32P, _I32P, _I32P, _DP, C.c_int32, _I32P,
                                         C.POINTER(_P)]),
    "cb_program_hip_source": (C.c_int, [_P, C.c_int32, C.c_uint32, C.c_char_p, C.c_size_t, C.POINTER(C.c_size_t)]),
    "cb_export_device_function": (C.c_int, [_P, C.c_int32, C.c_int32, C.c_char_p, C.c_char_p, C.c_size_t, C.POINTER(C.c_size_t)]),
    "cb_hip_assemble": (C.c_int, [_P, C.c_int32, C.c_uint32, C.POINTER(_P)]),
    "cb_hip_assemble_export": (C.c_int, [_P, C.c_int32, C.POINTER(_P)]),
    "cb_hip_compile": (C.c_int, [_P, C.c_int32, C.c_uint32]),
    "cb_hip_call": (C.c_int, [_P, _DPP, C.c_int32, C.c_size_t, _DPP, C.c_int32, C.c_int32, _P, C.POINTER(C.c_float)]),
    "cb_hip_call_static": (C.c_int, [_P, _DPP, C.c_int32, C.c_size_t, _DPP, C.c_int32, _P]),
    "cb_module_free": (None, [_P]),
    "cb_integrator_create": (C.c_int, [_P, C.c_double, C.c_double, C.c_double, C.c_int32, C.POINTER(_P)]),
    "cb_integrator_advance": (C.c_int, [_P, C.c_size_t, _P, _P, _P, _P, _P, _P, C.c_double, C.c_int32, _P, _I64P]),
    "cb_integrator_free": (None, [_P]),
    "cb_cpu_assemble": (C.c_int, [_P, C.POINTER(_P)]),
    "cb_cpu_call": (C.c_int, [_P, _DPP, C.c_int32, C.c_size_t, _DPP, C.c_int32, C.c_int32]),
    "cb_cpu_module_free": (None, [_P]),
    "cb_hip_assemble_surrogate": (C.c_int, [_P, _P, C.POINTER(_P)]),
    "cb_module_fallback_count": (C.c_int, [_P, _P, C.POINTER(C.c_uint32)]),
    "cb_mlp_create": (C.c_int, [C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.POINTER(_P)]),
    "cb_mlp_free": (None, [_P]),
    "cb_mlp_sizes": (C.c_int, [_P, _I64P, _I32P, _I32P, _I64P]),
    "cb_mlp_workspace_bytes": (C.c_int, [_P, C.c_int64, C.POINTER(C.c_size_t)]),
    "cb_mlp_forward": (C.c_int, [_P, _P, _P, _P, C.c_int64, _P, _P, _P]),
    "cb_mlp_backward_mse": (C.c_int, [_P, _P, _P, _P, C.c_int64, _P, _P, _P]),
    "cb_mlp_adam": (C.c_int, [_P, _P, _P, _P, _P, _P, C.c_float, C.c_float, C.c_float, C.c_float, C.c_int32, _P]),
}


def lib():
    global _lib
    if _lib is None:
        # One HIP runtime per process: torch ships its own libamdhip64.so.7, and whichever copy loads first serves
        # every later user of that soname. Loading torch first keeps the runtime torch was built against (device
        # tensors from torch are passed straight to the kernels); without torch the system ROCm runtime is used.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m combustion_amd.build`")
        l = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(l, name)
            f.restype = res
            f.argtypes = args
        _lib = l
    return _lib


def _check(status):
    if status != 0:
        raise Error(status, lib().cb_last_error().decode())


def _text(fn, *args):
    needed = C.c_size_t(0)
    status = fn(*args, None, 0, C.byref(needed))
    if status not in (0, -7):
        _check(status)
    buf = C.create_string_buffer(needed.value)
    _check(fn(*args, buf, needed.value, C.byref(needed)))
    return buf.value.decode()


class Mechanism:
    """yaml::parse + combustion::new (yaml.rs:54, lib.rs:127)."""

    def __init__(self, path=None, text=None):
        h = _P()
        if text is not None:
            _check(lib().cb_mechanism_parse_yaml(text.encode(), C.byref(h)))
        else:
            _check(lib().cb_mechanism_load_yaml(os.fsencode(path), C.byref(h)))
        self._h = h
        K, A, R = C.c_int32(), C.c_int32(), C.c_int32()
        _check(lib().cb_mechanism_sizes(h, C.byref(K), C.byref(A), C.byref(R)))
        self.K, self.active, self.R = K.value, A.value, R.value
        self._json = None

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.cb_mechanism_free(self._h)
            self._h = None

    @property
    def tables(self):
        """Parsed species and reaction tables in SI units (dict from the library's JSON dump)."""
        if self._json is None:
            self._json = json.loads(_text(lib().cb_mechanism_json, self._h))
        return self._json

    @property
    def species_names(self):
        return [s["name"] for s in self.tables["species"]]

    @property
    def molar_mass(self):
        return np.array([s["molar_mass"] for s in self.tables["species"]])

    def ignition_state(self, path=None, text=None):
        """cantera.rs:13-29: (T [K], P [Pa], amounts [mol] in V = 1 m³) of the first state of a .ignition file."""
        T, P, n = C.c_double(), C.c_double(), np.zeros(self.K)
        if text is not None:
            _check(lib().cb_ignition_state_parse(self._h, text.encode(), C.byref(T), C.byref(P), n.ctypes.data_as(_DP)))
        else:
            _check(lib().cb_ignition_state_load(self._h, os.fsencode(path), C.byref(T), C.byref(P), n.ctypes.data_as(_DP)))
        return T.value, P.value, n

    def transport_fits(self, D=4, T0=1000.0):
        K = self.K
        c, v, d = np.zeros((K, D)), np.zeros((K, D)), np.zeros((K, K, D))
        _check(lib().cb_transport_fits(self._h, D, T0, c.ctypes.data_as(_DP), v.ctypes.data_as(_DP), d.ctypes.data_as(_DP)))
        return c, v, d

    def transport_species(self, what, a, b, T):
        out = C.c_double()
        code = {"viscosity": 0, "conductivity": 1, "diffusivity": 2, "omega22": 3, "omega11": 4}[what]
        _check(lib().cb_transport_species(self._h, code, a, b, T, C.byref(out)))
        return out.value

    def transport_scales(self, T0=1000.0):
        v, c = C.c_double(), C.c_double()
        _check(lib().cb_transport_scales(self._h, T0, C.byref(v), C.byref(c)))
        return T0, v.value, c.value

    def rates(self):
        """reaction::rates (reaction.rs:356)."""
        h = _P()
        _check(lib().cb_build_rates(self._h, C.byref(h)))
        return Program(h, self)

    def reactions_rates(self):
        h = _P()
        _check(lib().cb_build_reactions_rates(self._h, C.byref(h)))
        return Program(h, self)

    def transport(self, D=5, T0=1000.0, viscosity=1.0, conductivity=1.0):
        """transport::properties::<D> (transport.rs:212)."""
        h = _P()
        _check(lib().cb_build_transport(self._h, D, T0, viscosity, conductivity, C.byref(h)))
        return Program(h, self)

    EXPORT_FILES = {"rates": 0, "conductivity": 1, "viscosity": 2, "diffusivity": 3}

    def device_function(self, which, prefix="nekrk_", D=5):
        """cb_export_device_function: the text of one file of the nekRK device-function export (cuda.rs:319-337);
        which is "rates", "conductivity", "viscosity" or "diffusivity"."""
        return _text(lib().cb_export_device_function, self._h, self.EXPORT_FILES[which], D, prefix.encode())

    def export_device_functions(self, directory, prefix="nekrk_", D=5):
        """Writes rates.c, conductivity.c, viscosity.c and diffusivity.c into directory (the reference writes them
        to <target>/<mechanism>/, cuda.rs:319-320); returns their paths."""
        import os
        os.makedirs(directory, exist_ok=True)
        paths = []
        for which in self.EXPORT_FILES:
            path = os.path.join(directory, which + ".c")
            with open(path, "w") as f:
                f.write(self.device_function(which, prefix, D))
            paths.append(path)
        return paths

    def export_module(self, D=5, compile_only=False):
        """cb_hip_assemble_export: the four exported files compiled by hiprtc behind one SoA kernel (inputs [T, P, Y],
        outputs [ω, λ, μ, D_k]); a Module, or None after filling the code-object cache when compile_only is set."""
        if compile_only:
            _check(lib().cb_hip_assemble_export(self._h, D, None))
            return None
        h = _P()
        _check(lib().cb_hip_assemble_export(self._h, D, C.byref(h)))
        return Module(h, _Signature(self.K + 2, self.active + 2 + self.K))

    def reactor(self, lean=False):
        """cb_build_reactor: [T, P, Y_0..Y_{K-1}] -> [dT/dt, dY_0/dt..dY_{K-1}/dt] at constant pressure."""
        h = _P()
        _check(lib().cb_build_reactor(self._h, int(lean), C.byref(h)))
        return Program(h, self)

    def chemistry(self, flags=RATES | TRANSPORT, D=4, T0=1000.0, viscosity=0.0, conductivity=0.0):
        """Per-state SoA program: inputs [T, P, Y_0..Y_{K-1}] -> [ω_0..ω_{Ka-1}, λ, μ, D_0..D_{K-1}, (dT/dt)]."""
        h = _P()
        _check(lib().cb_build_chemistry(self._h, flags, D, T0, viscosity, conductivity, C.byref(h)))
        return Program(h, self)


class Program:
    """An ast::Function (ast/lib.rs:328) built by the library."""

    def __init__(self, handle, mechanism):
        self._h = handle
        self.mechanism = mechanism
        n_in, n_out, n_reg = C.c_int32(), C.c_int32(), C.c_int32()
        n_st, n_ins = C.c_int64(), C.c_int64()
        _check(lib().cb_program_sizes(handle, C.byref(n_in), C.byref(n_out), C.byref(n_st), C.byref(n_ins), C.byref(n_reg)))
        self.n_inputs, self.n_outputs = n_in.value, n_out.value
        self.n_statements, self.n_instructions, self.n_registers = n_st.value, n_ins.value, n_reg.value

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.cb_program_free(self._h)
            self._h = None

    def op_counts(self):
        a = (C.c_int64 * 8)()
        _check(lib().cb_program_op_counts(self._h, a))
        keys = ["alu", "div", "sqrt", "exp", "ln", "neg", "select", "constants"]
        d = dict(zip(keys, list(a)))
        d["weighted"] = d["alu"] + 20 * (d["exp"] + d["ln"]) + 8 * (d["div"] + d["sqrt"])
        return d

    def linear(self):
        """Linear SSA form (op, dst, a, b, c, k, outputs) as numpy arrays."""
        n = self.n_instructions
        op = np.zeros(n, np.uint8)
        dst, a, b, c = (np.zeros(n, np.int32) for _ in range(4))
        k = np.zeros(n)
        out = np.zeros(self.n_outputs, np.int32)
        _check(lib().cb_program_linear(self._h, op.ctypes.data_as(C.POINTER(C.c_uint8)), dst.ctypes.data_as(_I32P),
                                       a.ctypes.data_as(_I32P), b.ctypes.data_as(_I32P), c.ctypes.data_as(_I32P),
                                       k.ctypes.data_as(_DP), out.ctypes.data_as(_I32P)))
        return {"op": op, "dst": dst, "a": a, "b": b, "c": c, "k": k, "outputs": out, "n_inputs": self.n_inputs,
                "n_registers": self.n_registers}

    @classmethod
    def from_linear(cls, linear):
        """cb_program_from_linear: a program from a linear form (the dict that linear() returns)."""
        op = np.ascontiguousarray(linear["op"], np.uint8)
        dst, a, b, c, out = (np.ascontiguousarray(linear[x], np.int32) for x in ("dst", "a", "b", "c", "outputs"))
        k = np.ascontiguousarray(linear["k"], np.float64)
        h = _P()
        _check(lib().cb_program_from_linear(int(linear["n_inputs"]), len(op), op.ctypes.data_as(C.POINTER(C.c_uint8)),
                                            dst.ctypes.data_as(_I32P), a.ctypes.data_as(_I32P), b.ctypes.data_as(_I32P),
                                            c.ctypes.data_as(_I32P), k.ctypes.data_as(_DP), len(out), out.ctypes.data_as(_I32P),
                                            C.byref(h)))
        return cls(h, None)

    def jacobian(self, wrt, rows=None):
        """cb_build_jacobian: outputs [own outputs, d output[rows[i]] / d input[wrt[j]] row-major]; rows default to all
        outputs."""
        wrt = np.ascontiguousarray(wrt, np.int32)
        rows = np.ascontiguousarray(range(self.n_outputs) if rows is None else rows, np.int32)
        h = _P()
        _check(lib().cb_build_jacobian(self._h, len(wrt), wrt.ctypes.data_as(_I32P), len(rows), rows.ctypes.data_as(_I32P), C.byref(h)))
        return Program(h, self.mechanism)

    def split(self, first_varying):
        """cb_program_split: (shared, varying) programs; shared takes the inputs before first_varying and returns the
        values the rest reads, varying takes this program's inputs followed by those values."""
        s, v = _P(), _P()
        _check(lib().cb_program_split(self._h, int(first_varying), C.byref(s), C.byref(v)))
        return Program(s, self.mechanism), Program(v, self.mechanism)

    def hip_source(self, form="straight", block_size=0):
        return _text(lib().cb_program_hip_source, self._h, form_code(form), block_size)

    def compile(self, form="straight", block_size=0):
        """Fills the code-object cache for a form (hiprtc for gfx950, no GPU needed)."""
        _check(lib().cb_hip_compile(self._h, form_code(form), block_size))

    def assemble_cpu(self):
        """ir::compile + ir::assemble (ir/lib.rs:225-263) with the loop of device.rs:28-48: the program as C compiled
        by the host compiler and called once per state (cb_cpu_assemble)."""
        h = _P()
        _check(lib().cb_cpu_assemble(self._h, C.byref(h)))
        return CpuModule(h, self)

    def assemble(self, block_size=0, form="straight"):
        """device::assemble (device.rs:25): emit, compile with hiprtc, load the module."""
        h = _P()
        _check(lib().cb_hip_assemble(self._h, form_code(form), block_size, C.byref(h)))
        return Module(h, self)


class _Signature:
    """Input and output counts of a module that was not built from a Program."""

    def __init__(self, n_inputs, n_outputs):
        self.n_inputs, self.n_outputs = n_inputs, n_outputs


def _pointer(x):
    if isinstance(x, np.ndarray):
        if x.dtype != np.float64 or not x.flags.c_contiguous:
            raise ValueError("arrays must be contiguous float64")
        return x.ctypes.data, False
    # torch tensor (device memory)
    if getattr(x, "is_cuda", False):
        if str(x.dtype) != "torch.float64" or not x.is_contiguous():
            raise ValueError("tensors must be contiguous float64")
        return x.data_ptr(), True
    raise TypeError("inputs must be numpy arrays or torch device tensors")


class Module:
    """The callable device::assemble returns (device.rs:59): module(inputs) -> outputs (SoA)."""

    def __init__(self, handle, program):
        self._h = handle
        self.program = program
        self.last_kernel_ms = None

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.cb_module_free(self._h)
            self._h = None

    def __call__(self, inputs, outputs=None, stream=None, time_kernel=False):
        """inputs: sequence of n_inputs arrays of N states (numpy = host, torch cuda = device).
        Returns outputs (allocated like the inputs when not given)."""
        if len(inputs) != self.program.n_inputs:
            raise ValueError(f"expected {self.program.n_inputs} inputs, got {len(inputs)}")
        n = len(inputs[0])
        ptrs = [_pointer(x) for x in inputs]
        on_device = ptrs[0][1]
        if any(d != on_device for _, d in ptrs):
            raise ValueError("inputs must all be on the host or all on the device")
        if any(len(x) != n for x in inputs):
            raise ValueError("all inputs must have the same length")
        if outputs is None:
            if on_device:
                import torch
                outputs = [torch.empty(n, dtype=torch.float64, device=inputs[0].device) for _ in range(self.program.n_outputs)]
            else:
                outputs = [np.empty(n) for _ in range(self.program.n_outputs)]
        optrs = [_pointer(x) for x in outputs]
        if any(d != on_device for _, d in optrs) or len(outputs) != self.program.n_outputs:
            raise ValueError("outputs must match the inputs' placement and the program's output count")
        in_arr = (C.c_void_p * len(ptrs))(*[p for p, _ in ptrs])
        out_arr = (C.c_void_p * len(optrs))(*[p for p, _ in optrs])
        ms = C.c_float(0)
        _check(lib().cb_hip_call(self._h, in_arr, len(ptrs), n, out_arr, len(optrs), int(on_device), stream,
                                 C.byref(ms) if time_kernel else None))
        if time_kernel:
            self.last_kernel_ms = ms.value
        return outputs


class Integrator:
    """cb_integrator_*: batched constant-pressure reactors on the GPU (Rosenbrock ode23s or RODAS4; the Jacobian
    analytic or by difference quotients).
    advance(y, P, t, h, status, steps, t_end) on contiguous torch device tensors: y float64 [K+1][n] (T, Y_k), P, t, h
    float64 [n], status int32 [n], steps int32 [n]; returns the batch's step count."""

    JACOBIAN = {"automatic": 0, "analytic": 1, "differences": 2}
    ANALYTIC_MAX_STATE = 8  # "automatic" takes the analytic program up to this state size (kAnalyticMaxState)

    @classmethod
    def resolved_jacobian(cls, mechanism, jacobian="automatic"):
        """The Jacobian an integrator of this mechanism uses: "analytic" or "differences"."""
        if jacobian != "automatic":
            return jacobian
        return "analytic" if mechanism.K + 1 <= cls.ANALYTIC_MAX_STATE else "differences"

    METHOD = {"rodas4": 0, "ode23s": 0x20}

    @classmethod
    def flags(cls, jacobian="automatic", one_per_wave=False, method="rodas4"):
        """cb_integrator_flags from a Jacobian name, the Rosenbrock packing and the method."""
        return cls.JACOBIAN[jacobian] | (0x10 if one_per_wave else 0) | cls.METHOD[method]

    def __init__(self, mechanism, rtol=1e-6, atol_T=1e-6, atol_Y=1e-12, jacobian="automatic", one_per_wave=False, method="rodas4"):
        h = _P()
        _check(lib().cb_integrator_create(mechanism._h, rtol, atol_T, atol_Y, self.flags(jacobian, one_per_wave, method),
                                          C.byref(h)))
        self._h = h
        self.mechanism = mechanism

    @classmethod
    def compile(cls, mechanism, jacobian="automatic", one_per_wave=False, method="rodas4"):
        _check(lib().cb_integrator_create(mechanism._h, 1e-6, 1e-6, 1e-12, cls.flags(jacobian, one_per_wave, method), None))

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.cb_integrator_free(self._h)
            self._h = None

    def advance(self, y, P, t, h, status, steps, t_end, max_steps=100000, stream=None):
        n = P.shape[0]
        m = self.mechanism.K + 1
        if tuple(y.shape) != (m, n):
            raise ValueError(f"y must be [{m}][{n}]")
        for x, dtype in ((y, "torch.float64"), (P, "torch.float64"), (t, "torch.float64"), (h, "torch.float64"), (status, "torch.int32"),
                         (steps, "torch.int32")):
            if not getattr(x, "is_cuda", False) or str(x.dtype) != dtype or not x.is_contiguous() or x.shape[-1] != n:
                raise ValueError("integrator arrays must be contiguous device tensors of the documented dtypes")
        taken = C.c_int64(0)
        _check(lib().cb_integrator_advance(self._h, n, y.data_ptr(), P.data_ptr(), t.data_ptr(), h.data_ptr(), status.data_ptr(),
                                           steps.data_ptr(), float(t_end), int(max_steps), stream, C.byref(taken)))
        return taken.value


class CpuModule:
    """cb_cpu_call: module(inputs, outputs=None, threads=1) over SoA host arrays (numpy float64)."""

    def __init__(self, handle, program):
        self._h = handle
        self.program = program

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.cb_cpu_module_free(self._h)
            self._h = None

    def __call__(self, inputs, outputs=None, threads=1):
        if len(inputs) != self.program.n_inputs:
            raise ValueError(f"expected {self.program.n_inputs} inputs, got {len(inputs)}")
        inputs = [np.ascontiguousarray(x, np.float64) for x in inputs]
        n = len(inputs[0])
        if any(len(x) != n for x in inputs):
            raise ValueError("all inputs must have the same length")
        if outputs is None:
            outputs = [np.empty(n) for _ in range(self.program.n_outputs)]
        for x in outputs:
            if not (isinstance(x, np.ndarray) and x.dtype == np.float64 and x.flags.c_contiguous and len(x) == n):
                raise ValueError("outputs must be contiguous float64 arrays of the inputs' length")
        in_arr = (C.c_void_p * len(inputs))(*[x.ctypes.data for x in inputs])
        out_arr = (C.c_void_p * len(outputs))(*[x.ctypes.data for x in outputs])
        _check(lib().cb_cpu_call(self._h, in_arr, len(inputs), n, out_arr, len(outputs), int(threads)))
        return outputs


class SurrogateWeights(C.Structure):
    """cb_surrogate_weights."""
    _fields_ = [("width", C.c_int32), ("depth", C.c_int32)] + [
        (name, C.POINTER(C.c_float)) for name in ("params", "feature_mean", "feature_rstd", "lower", "upper", "output_mean", "output_std")] + [
        ("n_major", C.c_int32), ("major", C.POINTER(C.c_int32)), ("rates_form", C.c_int32)]


def _tensor(x, dtype):
    if not getattr(x, "is_cuda", False) or str(x.dtype) != dtype or not x.is_contiguous():
        raise ValueError(f"expected a contiguous {dtype} device tensor")
    return x.data_ptr()


class Network:
    """The transport surrogate's MLP kernels (cb_mlp_*): n_in -> width (SiLU) x depth -> n_out, BF16 weights,
    FP32 accumulation. Parameters, activations and gradients are torch device tensors owned by the caller."""

    def __init__(self, n_in, width, depth, n_out):
        h = _P()
        _check(lib().cb_mlp_create(n_in, width, depth, n_out, C.byref(h)))
        self._h = h
        self.n_in, self.width, self.depth, self.n_out = n_in, width, depth, n_out
        n_params, in_pad, out_pad = C.c_int64(), C.c_int32(), C.c_int32()
        offsets = (C.c_int64 * (2 * (depth + 1)))()
        _check(lib().cb_mlp_sizes(h, C.byref(n_params), C.byref(in_pad), C.byref(out_pad), offsets))
        self.n_params, self.in_pad, self.out_pad = n_params.value, in_pad.value, out_pad.value
        dims = [self.in_pad] + [width] * depth + [self.out_pad]
        # (W offset, b offset, out, in) per layer
        self.layers = [(offsets[2 * l], offsets[2 * l + 1], dims[l + 1], dims[l]) for l in range(depth + 1)]

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.cb_mlp_free(self._h)
            self._h = None

    def workspace_bytes(self, n):
        b = C.c_size_t()
        _check(lib().cb_mlp_workspace_bytes(self._h, n, C.byref(b)))
        return b.value

    def forward(self, params, params16, x, y, workspace, stream=None):
        n = x.shape[0]
        _check(lib().cb_mlp_forward(self._h, _tensor(params, "torch.float32"), _tensor(params16, "torch.bfloat16"),
                                    _tensor(x, "torch.float32"), n, _tensor(y, "torch.float32"), _tensor(workspace, "torch.uint8"), stream))

    def backward_mse(self, params16, y, target, workspace, grad, stream=None):
        n = y.shape[0]
        _check(lib().cb_mlp_backward_mse(self._h, _tensor(params16, "torch.bfloat16"), _tensor(y, "torch.float32"),
                                         _tensor(target, "torch.float32"), n, _tensor(workspace, "torch.uint8"),
                                         _tensor(grad, "torch.float32"), stream))

    def adam(self, params, params16, grad, m1, m2, lr, step, beta1=0.9, beta2=0.999, eps=1e-8, stream=None):
        _check(lib().cb_mlp_adam(self._h, _tensor(params, "torch.float32"), _tensor(params16, "torch.bfloat16"),
                                 _tensor(grad, "torch.float32"), _tensor(m1, "torch.float32"), _tensor(m2, "torch.float32"),
                                 lr, beta1, beta2, eps, step, stream))
