"""MI355X (gfx950) backend for per-state chemistry: net production rates and mixture-averaged transport.

The native library (libcombustion_amd.so) holds the Cantera-YAML front end, the program builders, the HIP
emitters and the launcher behind the C ABI of include/combustion_amd.h; `native` binds it with ctypes.
"""
from .native import (FORM_FP64, FORM_MIXED, FORM_STRAIGHT, RATES, RHS, TRANSPORT, Error, Mechanism, Module,
                     Program)

__all__ = ["Mechanism", "Program", "Module", "Error", "RATES", "TRANSPORT", "RHS", "FORM_STRAIGHT", "FORM_FP64",
           "FORM_MIXED"]
