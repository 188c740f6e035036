"""Build recipe of the native library (hipcc for gfx950, in-tree output).

    python -m emsolver_amd.build

compiles ``emsolver_amd/lib/libemsolver_tlm.so`` from the HIP sources under
``csrc/`` and the C++ drop-in glue under ``host/`` (TlmSolver facade and
RunGpuTlm), plus the test programs ``tests/selftest/tlm_selftest`` and
``tests/selftest/tlm_stop``.

Flags that matter:
  -ffp-contract=off  no fused multiply-add: every a*b+c rounds twice, as the
                     reference CPU path does, so FP32 results match bit for bit
  -Xclang -target-feature -Xclang -packed-fp32-ops
                     keep the compiler from emitting packed FP32 VALU ops
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB_DIR = os.path.join(HERE, "lib")
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("EMSOLVER_ARCH", "gfx950")

COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall",
          f"-I{os.path.join(ROOT, 'include')}", f"-I{os.path.join(HERE, 'csrc')}",
          f"-I{os.path.join(HERE, 'host')}"]
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]

HIP_SOURCES = ["csrc/GpuTlmHip.hip", "csrc/TlmBatch.hip", "csrc/TlmVoxel.hip",
               "csrc/TlmFarField.hip"]
HOST_SOURCES = ["host/TlmSolver.cpp", "host/GpuTlm.cpp"]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    p = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(l for l in (p.stdout + p.stderr).splitlines()
                    if "packed-fp32-ops" not in l and l.strip())
    if p.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{out}")
    if verbose and out:
        print(out)


def build(verbose: bool = False) -> str:
    os.makedirs(LIB_DIR, exist_ok=True)
    lib = os.path.join(LIB_DIR, "libemsolver_tlm.so")
    srcs = [os.path.join(HERE, s) for s in HIP_SOURCES]
    srcs += [os.path.join(HERE, s) for s in HOST_SOURCES if os.path.exists(os.path.join(HERE, s))]
    tmp = lib + ".tmp"
    _run([HIPCC, *DEVICE, *COMMON, "-shared", "-o", tmp, *srcs], verbose)
    os.replace(tmp, lib)
    # the self-test and cancellation programs (test infrastructure, live with the tests)
    for name in ("tlm_selftest", "tlm_stop"):
        st = os.path.join(ROOT, "tests", "selftest", name + ".cpp")
        if os.path.exists(st):
            exe = os.path.join(ROOT, "tests", "selftest", name)
            _run([HIPCC, *COMMON, "-o", exe + ".tmp", st, f"-L{LIB_DIR}", "-lemsolver_tlm",
                  f"-Wl,-rpath,{LIB_DIR}", "-Wl,-rpath,$ORIGIN/../../emsolver_amd/lib", "-pthread",
                  "-ldl"], verbose)
            os.replace(exe + ".tmp", exe)
    return lib


if __name__ == "__main__":
    print(build(verbose="-q" not in sys.argv))
