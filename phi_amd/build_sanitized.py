"""Host-code sanitizer build of the C-ABI library (CPU checks only).

Compiles every host translation unit (the `.prm` reader, hierarchy
construction, integrator loop, C ABI, ...) with AddressSanitizer and
UndefinedBehaviorSanitizer, the device translation units as usual (no GPU
sanitizer), and links phi_amd/_build_asan/libphi_hip.so; the CPU oracle
(oracle/heom_oracle.cc) is built beside it with the same sanitizers.  Run
the CPU tests
against it with

    python phi_amd/build_sanitized.py --run-tests

which preloads clang's ASan runtime into the Python interpreter, points
PHI_HIP_LIBRARY and HEOM_ORACLE_LIBRARY at the sanitized libraries and runs
`pytest -m "not gpu"`.
Any sanitizer report aborts the offending test process (halt_on_error).
"""

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, HERE)
import build as B  # noqa: E402  (phi_amd/build.py: sources and flags)

OUT = os.path.join(HERE, "_build_asan")
LIB = os.path.join(OUT, "libphi_hip.so")
ORACLE_LIB = os.path.join(OUT, "libheom_oracle.so")
SAN = ["-fsanitize=address", "-fsanitize=undefined", "-fno-sanitize-recover=undefined"]


def _host_san_flags():
    # Each sanitizer flag applies to the host pass only.
    out = []
    for f in SAN:
        out += ["-Xarch_host", f]
    return out


def _compile(src):
    obj = os.path.join(OUT, os.path.basename(src) + ".o")
    cmd = [B.HIPCC] + B.COMMON + ["-g", "-fno-omit-frame-pointer", "-c",
                                   os.path.join(B.SRC, src), "-o", obj]
    if src.endswith(".hip"):
        cmd[1:1] = B.DEVICE
    else:
        cmd += ["-D__HIP_PLATFORM_AMD__"] + _host_san_flags()
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s" % (" ".join(cmd), r.stdout))
    return obj


def asan_runtime():
    r = subprocess.run([B.HIPCC, "-print-file-name=libclang_rt.asan-x86_64.so"],
                       stdout=subprocess.PIPE, text=True, check=True)
    return r.stdout.strip()


def build_oracle():
    """The CPU oracle (host code only) with the same sanitizers and the same
    runtime as the library (clang's, shared)."""
    clang = os.path.join(os.path.dirname(os.path.realpath(B.HIPCC)), "..", "llvm", "bin",
                         "clang++")
    if not os.path.exists(clang):
        clang = "/opt/rocm/llvm/bin/clang++"
    cmd = [clang, "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-ffp-contract=off",
           "-fPIC", "-shared", "-shared-libsan"] + SAN + [
           "-o", ORACLE_LIB, os.path.join(ROOT, "oracle", "heom_oracle.cc"), "-lpthread"]
    subprocess.check_call(cmd)
    return ORACLE_LIB


def build():
    import concurrent.futures
    os.makedirs(OUT, exist_ok=True)
    with concurrent.futures.ThreadPoolExecutor(max_workers=8) as ex:
        objs = list(ex.map(_compile, B.LIB_SOURCES))
    subprocess.check_call([B.HIPCC, "--offload-arch=" + B.ARCH, "-shared", "-fPIC",
                           "-shared-libsan", "-o", LIB] + objs + _host_san_flags()
                          + ["-lpthread"])
    build_oracle()
    return LIB


def run_tests(extra):
    env = dict(os.environ)
    env["PHI_HIP_LIBRARY"] = LIB
    env["HEOM_ORACLE_LIBRARY"] = ORACLE_LIB
    env["LD_PRELOAD"] = asan_runtime()
    # The interpreter itself is not instrumented: leaks of Python objects
    # are not this library's, and an impossible allocation request (the
    # oracle's reader on a fuzzed size) must fail as it does without the
    # sanitizer (NULL -> MemoryError / std::bad_alloc); every other report
    # stops the process.
    env["ASAN_OPTIONS"] = ("detect_leaks=0:allocator_may_return_null=1:halt_on_error=1:"
                           "abort_on_error=1")
    env["UBSAN_OPTIONS"] = "print_stacktrace=1:halt_on_error=1"
    cmd = [sys.executable, "-m", "pytest", os.path.join(ROOT, "tests"), "-q", "-m", "not gpu",
           "-p", "no:cacheprovider"] + extra
    return subprocess.call(cmd, env=env, cwd=ROOT)


if __name__ == "__main__":
    build()
    print("built", LIB)
    if "--run-tests" in sys.argv:
        sys.exit(run_tests([a for a in sys.argv[1:] if a != "--run-tests"]))
