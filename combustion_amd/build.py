"""Builds the native library libcombustion_amd.so in place (host C++ against the HIP runtime and hiprtc).

The mechanism-specialised gfx950 kernels are generated and compiled at assemble time by hiprtc (or ahead of
time into the code-object cache by `precompile`). Usage: python -m combustion_amd.build
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libcombustion_amd.so")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
SOURCES = ["yaml.cpp", "mechanism.cpp", "ast.cpp", "reaction.cpp", "transport.cpp", "hip_emit.cpp", "device.cpp",
           "forms.cpp", "tiled.cpp", "capi.cpp"]


def build(verbose=True):
    objs = []
    os.makedirs(os.path.join(HERE, "_build"), exist_ok=True)
    procs = []
    for src in SOURCES:
        obj = os.path.join(HERE, "_build", src.replace(".cpp", ".o"))
        cmd = ["g++", "-O2", "-g", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function", "-D__HIP_PLATFORM_AMD__",
               "-I" + os.path.join(ROCM, "include"), "-c", os.path.join(CSRC, src), "-o", obj]
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
        objs.append(obj)
    failed = False
    for src, p in procs:
        out, _ = p.communicate()
        if out and verbose:
            sys.stdout.write(out)
        if p.returncode:
            failed = True
            sys.stdout.write(f"compile failed: {src}\n")
    if failed:
        raise RuntimeError("native build failed")
    tmp = LIB + ".tmp"
    cmd = ["g++", "-shared", "-o", tmp] + objs + ["-L" + os.path.join(ROCM, "lib"), "-Wl,-rpath," + os.path.join(ROCM, "lib"),
                                                  "-lamdhip64", "-lhiprtc", "-ldl"]
    subprocess.check_call(cmd)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build())
