"""Build step for the native library (libntprop.so, gfx950).

`build_library()` compiles neutrinotelescopes_amd/csrc/*.hip with hipcc into
neutrinotelescopes_amd/libntprop.so.  Flags:
  -ffp-contract=off   no implicit fma (the reference's Julia code has none)
  -packed-fp32-ops    disabled: packed FP32 instructions are not generated
"""
from __future__ import annotations

import os
import shutil
import subprocess
from pathlib import Path

PKG_DIR = Path(__file__).resolve().parent
CSRC = PKG_DIR / "csrc"
INCLUDE = PKG_DIR.parent / "include"
LIB_NAME = "libntprop.so"
ARCH = os.environ.get("NTPROP_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    "-O3",
    "-std=c++17",
    f"--offload-arch={ARCH}",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-result",
]


def hipcc_path() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the HIP toolchain is required to build libntprop")


def sources() -> list[Path]:
    return sorted(CSRC.glob("*.hip"))


def build_library(verbose: bool = False) -> Path:
    """Compile every HIP source into the package's shared library (always):
    one object per source, compiled in parallel, then one link."""
    import tempfile
    from concurrent.futures import ThreadPoolExecutor
    out = PKG_DIR / LIB_NAME
    tmp = out.with_suffix(".so.tmp")
    hipcc = hipcc_path()
    compile_flags = [f for f in HIP_FLAGS if f != "-shared"]
    with tempfile.TemporaryDirectory(prefix="ntprop_build_") as d:
        objs = [Path(d) / (src.stem + ".o") for src in sources()]

        def compile_one(src_obj):
            src, obj = src_obj
            cmd = [hipcc, *compile_flags, f"-I{INCLUDE}", "-c", "-o", str(obj), str(src)]
            if verbose:
                print(" ".join(cmd), flush=True)
            return subprocess.run(cmd, capture_output=True, text=True)

        workers = max(1, min(len(objs), os.cpu_count() or 1, 8))
        with ThreadPoolExecutor(workers) as pool:
            results = list(pool.map(compile_one, zip(sources(), objs)))
        for res in results:
            if res.returncode != 0:
                raise RuntimeError(f"hipcc failed ({res.returncode}):\n{res.stderr[-8000:]}")
        cmd = [hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError(f"link failed ({res.returncode}):\n{res.stderr[-8000:]}")
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build_library(verbose=True))
