"""The C-ABI library loads and exports every function include/ntprop.h declares
(no device calls: runs without a GPU)."""
import ctypes as C
import re
import shutil
import subprocess
from pathlib import Path

import pytest

from neutrinotelescopes_amd import _abi

ROOT = Path(__file__).resolve().parents[1]


def declared_functions():
    names = set()
    for header in sorted((ROOT / "include").glob("*.h")):
        text = re.sub(r"/\*.*?\*/", "", header.read_text(), flags=re.S)
        names |= set(re.findall(r"^\s*(?:int32_t|void|const char\*)\s+\**(nt(?:f|pmt)?_\w+)\s*\(", text, re.M))
    return sorted(names)


def test_header_declares_entry_points():
    fns = declared_functions()
    for name in ("nt_create", "nt_destroy", "nt_set_medium", "nt_set_spectrum", "nt_set_targets",
                 "nt_set_sources", "nt_propagate", "nt_launch", "nt_eval", "ntpmt_create", "ntpmt_make_reco_pulses"):
        assert name in fns


def test_library_exports_every_declared_symbol():
    lib = _abi.load_library()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    # every declared function has a ctypes signature in the binding
    assert set(declared_functions()) <= set(_abi.SIGNATURES)


def test_abi_version_and_status_strings():
    lib = _abi.load_library()
    assert lib.nt_abi_version() == 1
    for st in range(0, 8):
        assert lib.nt_status_string(st)


def test_limits_match_header():
    text = (ROOT / "include" / "ntprop.h").read_text()
    for name in ("NT_MAX_PMTS", "NT_MAX_TARGETS", "NT_MAX_STEPS"):
        m = re.search(rf"#define {name} (\d+)", text)
        assert int(m.group(1)) == getattr(_abi, name), name
    # one photon's sphere-test count stays below 2^31 (the kernel's 32-bit per-lane counters)
    assert (_abi.NT_MAX_STEPS + 1) * _abi.NT_MAX_TARGETS <= 2 ** 31


def test_max_steps_bounded():
    import dataclasses
    import neutrinotelescopes_amd as nt
    med = nt.make_cascadia_medium_properties(0.95)
    assert dataclasses.replace(med, max_steps=_abi.NT_MAX_STEPS).to_abi().max_steps == _abi.NT_MAX_STEPS
    with pytest.raises(ValueError):
        dataclasses.replace(med, max_steps=_abi.NT_MAX_STEPS + 1)
    with pytest.raises(ValueError):
        dataclasses.replace(med, max_steps=-1)


def test_flow_constants_match_header():
    text = (ROOT / "include" / "ntflow.h").read_text()
    for name in ("NTF_N_IN", "NTF_HIDDEN", "NTF_BINS", "NTF_LAYERS", "NTF_P_PAD"):
        m = re.search(rf"#define {name} (\d+)", text)
        assert int(m.group(1)) == getattr(_abi, name), name
    assert "ntf_train_step" in declared_functions()
    assert _abi.NTF_N_PARAMS == 78 and _abi.NTF_N_WEIGHTS == 29008


def test_struct_layouts():
    assert C.sizeof(_abi.nt_source) == 64
    assert C.sizeof(_abi.nt_medium) == 48
    assert C.sizeof(_abi.nt_target) == 160
    assert _abi.nt_hits.capacity.offset == 12 * 8
    assert C.sizeof(_abi.nt_particle) == 40 and C.sizeof(_abi.nt_cascade_model) == 24
    assert C.sizeof(_abi.nt_injector) == 60
    assert _abi.nt_particle.type.offset == 32


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs a C compiler")
def test_struct_layouts_match_c_compiler(tmp_path):
    """sizeof and field offsets of every ABI struct as a C compiler lays them
    out from the header, against the ctypes mirror."""
    structs = ["nt_source", "nt_medium", "nt_target", "nt_spectrum", "nt_hits", "nt_particle", "nt_cascade_model",
               "nt_injector", "ntpmt_config", "ntpmt_pulses"]
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "ntprop.h"', '#include "ntpmt.h"',
             "int main(void) {"]
    for st in structs:
        lines.append(f'printf("{st} %zu\\n", sizeof({st}));')
        for name, *_ in getattr(_abi, st)._fields_:
            lines.append(f'printf("{st}.{name} %zu\\n", offsetof({st}, {name}));')
    lines.append("return 0; }")
    src = tmp_path / "probe.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-std=c11", "-I", str(ROOT / "include"), str(src), "-o", str(exe)], check=True)
    got = dict(l.split() for l in subprocess.run([str(exe)], check=True, capture_output=True,
                                                   text=True).stdout.splitlines())
    for st in structs:
        cls = getattr(_abi, st)
        assert int(got[st]) == C.sizeof(cls), st
        for name, *_ in cls._fields_:
            assert int(got[f"{st}.{name}"]) == getattr(cls, name).offset, (st, name)


def test_library_missing_fails_loudly(tmp_path):
    with pytest.raises(RuntimeError):
        _abi.load_library(tmp_path / "libntprop.so")


def test_null_arguments_rejected_without_device():
    lib = _abi.load_library()
    assert lib.nt_set_medium(None, None) == _abi.NT_ERR_ARG
    assert lib.nt_create(0, None) == _abi.NT_ERR_ARG
    assert lib.nt_eval(None, 0, None, 1, None, 1, 1) == _abi.NT_ERR_ARG
    inj, model = _abi.nt_injector(), _abi.nt_cascade_model()
    assert lib.nt_inject_particles(None, C.byref(inj), 1, 0, None, None) == _abi.NT_ERR_ARG
    assert lib.nt_build_cascade_sources(None, None, 1, C.byref(model), 0, None, None) == _abi.NT_ERR_ARG
    assert lib.nt_set_sources_device(None, None, 1, -1, None) == _abi.NT_ERR_ARG
