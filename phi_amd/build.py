"""Build recipe of the MI355X HEOM integrator (hipcc, gfx950, in-tree).

Produces
  phi_amd/libphi_hip.so   the C-ABI library (include/phi_hip.h)
  phi_amd/phi             the `phi <paramfile> <method> [threads]` program
Object files go to phi_amd/_build/.
"""

import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libphi_hip.so")
EXE = os.path.join(HERE, "phi")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-result"]
# Packed FP32 instructions are disabled for all device code.
DEVICE = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]

LIB_SOURCES = ["phi_parameters.cc", "hierarchy.cc", "engine.cc", "engine_closure.cc",
               "engine_filter.cc", "engine_td.cc", "engine_timelocal.cc", "timelocal.cc",
               "system_hamiltonian.cc", "mlp.cc",
               "integrator.cc", "capi.cc",
               "heom_kernels.hip", "heom_rk4_f64.hip", "heom_rk4_f32.hip",
               "heom_rkf45_f64.hip", "heom_rkf45_f32.hip", "mlp_kernels.hip",
               "closure_kernels.hip", "filter_kernels.hip", "bath_kernels.hip",
               "rhs_noalloc.hip"]


def _compile(src):
    obj = os.path.join(OUT, os.path.basename(src) + ".o")
    cmd = [HIPCC] + COMMON + ["-c", os.path.join(SRC, src), "-o", obj]
    if src.endswith(".hip"):
        cmd[1:1] = DEVICE
    else:
        cmd += ["-D__HIP_PLATFORM_AMD__"]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    msgs = [l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l]
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s" % (" ".join(cmd), "\n".join(msgs)))
    return obj


def build(verbose=True):
    os.makedirs(OUT, exist_ok=True)
    with concurrent.futures.ThreadPoolExecutor(max_workers=8) as ex:
        objs = list(ex.map(_compile, LIB_SOURCES))
    tmp = LIB + ".tmp"
    subprocess.check_call([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp]
                          + objs + ["-lpthread"])
    os.replace(tmp, LIB)
    cli = _compile("phi_cli.cc")
    subprocess.check_call([HIPCC, "-o", EXE + ".tmp", cli, "-L" + HERE, "-lphi_hip",
                           "-Wl,-rpath,$ORIGIN"])
    os.replace(EXE + ".tmp", EXE)
    if verbose:
        print("built", LIB, "and", EXE)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
