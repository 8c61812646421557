"""Build the gfx950 library, the command-line front end and the host oracle.

    python -m astrosimgpu_cuda_amd.build [--incremental] [--no-oracle]

Outputs (all in-tree, git-ignored):
    astrosimgpu_cuda_amd/lib/libastrosim.so   HIP kernels + C ABI (include/astrosim.h)
    astrosimgpu_cuda_amd/lib/libastrosim_selftest.so  test-only: device vs host math
    astrosimgpu_cuda_amd/bin/astrosim_fused   drop-in CLI linked against the library
    oracle/build/liboracle.so, astrosim_oracle  FP64 host oracle (tests / CPU baseline)
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
OBJ = PKG / "_build"
LIB = PKG / "lib" / "libastrosim.so"
SELFTEST = PKG / "lib" / "libastrosim_selftest.so"
CLI = PKG / "bin" / "astrosim_fused"
ORACLE = ROOT / "oracle"

ROCM = Path(os.environ.get("ROCM_PATH", "/opt/rocm"))
HIPCC = str(ROCM / "bin" / "hipcc")
HOSTCXX = str(ROCM / "llvm" / "bin" / "clang++")  # host-only sources (hipcc would compile them as HIP)
ARCH = "gfx950"
# Contraction off: every a*b+c rounds twice, so the FP64 kernels evaluate
# exactly like the host oracle. Packed FP32 instructions are disabled. The
# two-operand v_fmac_f64 form is disabled too: explicit FP64 fused
# multiply-adds then select the three-operand v_fma_f64 (the same single
# rounding) instead of a register copy plus v_fmac_f64, which cut the step
# loops' VALU instructions by 10-20% and their registers.
COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-Wno-unused-result",
          "-I", str(ROOT / "include")]
DEVICE = ["-x", "hip", f"--offload-arch={ARCH}",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-Xclang", "-target-feature", "-Xclang", "-fmacf64-inst"]
# Host sources use the FMA instruction for as_dmath's fused multiply-adds
# (correctly rounded either way; the instruction avoids a libm call).
HOST = ["-D__HIP_PLATFORM_AMD__", "-mfma", "-I", str(ROCM / "include")]

LIB_SOURCES = ["as_setup.cpp", "as_engine.cpp", "as_capi.cpp", "as_plain.hip", "as_batch.hip", "as_server.cpp",
               "as_gpubuild.hip", "as_record.hip"]
HEADERS = ["as_dmath.h", "as_rng.h", "as_model.h", "as_model32.h", "as_setup.h", "as_engine.h", "as_batch.h",
           "as_gpubuild.h", "as_gpubuild_run.h"]


def _run(cmd: list[str], cwd: Path | None = None) -> None:
    proc = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    if proc.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + proc.stdout + proc.stderr)
        raise RuntimeError(f"build step failed: {cmd[0]} ... {cmd[-1]}")


def _stale(target: Path, deps: list[Path]) -> bool:
    if not target.exists():
        return True
    t = target.stat().st_mtime
    return any(d.stat().st_mtime > t for d in deps)


def _compile(src: str, incremental: bool) -> Path:
    path = CSRC / src
    obj = OBJ / (src + ".o")
    deps = [path] + [CSRC / h for h in HEADERS] + [ROOT / "include" / "astrosim.h"]
    if incremental and not _stale(obj, deps):
        return obj
    if src.endswith(".hip"):
        _run([HIPCC, *COMMON, *DEVICE, "-c", str(path), "-o", str(obj)])
    else:
        _run([HOSTCXX, *COMMON, *HOST, "-c", str(path), "-o", str(obj)])
    return obj


def build_library(incremental: bool = False, jobs: int = 8) -> None:
    OBJ.mkdir(parents=True, exist_ok=True)
    LIB.parent.mkdir(parents=True, exist_ok=True)
    CLI.parent.mkdir(parents=True, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, incremental), LIB_SOURCES + ["as_selftest.hip"]))
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", str(tmp), *map(str, objs[:-1])])
    os.replace(tmp, LIB)
    tmp = SELFTEST.with_suffix(".so.tmp")
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", str(tmp), str(objs[-1]),
          "-L", str(LIB.parent), "-lastrosim", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, SELFTEST)
    cli_obj = _compile("as_cli.cpp", incremental)
    _run([HIPCC, "-o", str(CLI), str(cli_obj), "-L", str(LIB.parent), "-lastrosim",
          "-Wl,-rpath,$ORIGIN/../lib"])
    # the vector-instruction throughput and HBM streaming probes (measurement tools, not linked
    # to the library)
    probe_obj = _compile("as_valu_probe.hip", incremental)
    _run([HIPCC, "-o", str(CLI.parent / "valu_probe"), str(probe_obj)])
    bw_obj = _compile("as_bw_probe.hip", incremental)
    _run([HIPCC, "-o", str(CLI.parent / "bw_probe"), str(bw_obj)])


def build_oracle() -> None:
    make = shutil.which("make")
    if make is None:
        raise RuntimeError("make not found")
    _run([make, "-s", "-C", str(ORACLE), "all"])


def main(argv: list[str] | None = None) -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--incremental", action="store_true", help="skip up-to-date objects")
    ap.add_argument("--no-oracle", action="store_true")
    ap.add_argument("-j", "--jobs", type=int, default=8)
    args = ap.parse_args(argv)
    build_library(incremental=args.incremental, jobs=args.jobs)
    if not args.no_oracle:
        build_oracle()
    print(f"built {LIB.relative_to(ROOT)} and {CLI.relative_to(ROOT)}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
