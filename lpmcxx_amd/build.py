"""Build recipe for the native parts of lpmcxx_amd.

Products (all in-tree so they travel with the source tree):
  lpmcxx_amd/lib/liblpm_gpu.so   HIP kernels for gfx950 + the C ABI (include/lpm_gpu.h)
  lpmcxx_amd/lib/liblpm.so       C++ host library, namespace Lpm (reference API)
  tests/cpp/bin/*                C++ test drivers written against the Lpm API
  oracle/lib/liblpm_oracle.so    CPU restatement used only by tests and the CPU baseline
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
PKG = ROOT / "lpmcxx_amd"
LIB = PKG / "lib"
OBJ = LIB / "obj"
CSRC = PKG / "csrc"
HOST = PKG / "host"
ORACLE = ROOT / "oracle"
TESTCPP = ROOT / "tests" / "cpp"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# -ffp-contract=off: FMAs only where the source writes fmaf/fma, so the FP64
# pair and box arithmetic keeps the reference's rounding.  Packed FP32 is
# switched off: only scalar FP32 VALU instructions are generated.
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-I", str(ROOT / "include"),
]
# Per-file additions.  treecode_batch.hip: MFMA accumulators and results in
# ordinary VGPRs (the planning pass converts every MFMA result to BF16 with
# VALU instructions; in accumulation registers each result first needs a copy
# out, about 36 instructions per model call).
HIP_FILE_FLAGS = {"treecode_batch.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"]}
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-Wno-sign-compare",
             "-I", str(ROOT / "include"), "-I", str(HOST)]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {cmd[0]} ... {cmd[-1]}")


def build_gpu(jobs: int = 8) -> Path:
    OBJ.mkdir(parents=True, exist_ok=True)
    srcs = sorted(CSRC.glob("*.hip"))
    objs = [OBJ / (s.stem + ".o") for s in srcs]
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        futs = [ex.submit(_run, [HIPCC, *HIP_FLAGS, *HIP_FILE_FLAGS.get(s.name, []), "-c", str(s), "-o", str(o)]) for s, o in zip(srcs, objs)]
        for f in futs:
            f.result()
    out = LIB / "liblpm_gpu.so"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *map(str, objs), "-o", str(out)])
    return out


def build_host() -> Path:
    srcs = sorted(HOST.glob("*.cpp"))
    out = LIB / "liblpm.so"
    _run(["g++", *CXX_FLAGS, "-shared", *map(str, srcs), "-o", str(out), "-L", str(LIB), "-llpm_gpu",
          "-Wl,-rpath,$ORIGIN"])
    return out


def build_oracle() -> Path:
    (ORACLE / "lib").mkdir(parents=True, exist_ok=True)
    srcs = sorted(ORACLE.glob("*.c"))
    out = ORACLE / "lib" / "liblpm_oracle.so"
    _run(["gcc", "-O2", "-std=c11", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off", "-Wall",
          *map(str, srcs), "-o", str(out), "-lm"])
    return out


def build_test_drivers() -> list[Path]:
    bindir = TESTCPP / "bin"
    bindir.mkdir(parents=True, exist_ok=True)
    outs = []
    for s in sorted(TESTCPP.glob("*.cpp")):
        o = bindir / s.stem
        _run(["g++", *CXX_FLAGS, str(s), "-o", str(o), "-L", str(LIB), "-llpm", "-llpm_gpu",
              f"-Wl,-rpath,{LIB}", "-Wl,-rpath,$ORIGIN/../../../lpmcxx_amd/lib"])
        outs.append(o)
    return outs


def build_all(verbose: bool = True) -> None:
    LIB.mkdir(parents=True, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=2) as ex:
        fo = ex.submit(build_oracle)
        g = build_gpu()
        if verbose:
            print("built", g)
        fo.result()
    h = build_host()
    drivers = build_test_drivers()
    if verbose:
        print("built", h)
        print("built", ORACLE / "lib" / "liblpm_oracle.so")
        for d in drivers:
            print("built", d)
    shutil.rmtree(OBJ, ignore_errors=True)


if __name__ == "__main__":
    build_all()
