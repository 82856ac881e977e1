"""Device code that is also host-callable, compiled for the host and checked
on the CPU: the kernel's split Philox stream against plain Philox4x32-10."""
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not Path(HIPCC).exists(), reason="needs hipcc")
def test_philox_run_equals_philox(tmp_path):
    exe = tmp_path / "philox_run_check"
    subprocess.run([HIPCC, "-O2", "--offload-arch=gfx950", f"-I{ROOT / 'include'}",
                    f"-I{ROOT / 'neutrinotelescopes_amd' / 'csrc'}", "-o", str(exe),
                    str(ROOT / "tests" / "native" / "philox_run_check.hip")], check=True, capture_output=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    assert out.returncode == 0 and "mismatches 0" in out.stdout, out.stdout
