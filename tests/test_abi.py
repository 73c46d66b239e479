"""The C-ABI library loads, exports every entry point include/mmprop.h declares, and
fails loudly (MMP_ERR_NO_DEVICE) instead of computing anywhere else when no device is
present. No propagation runs here."""
import ctypes
import os
import re
import subprocess

import pytest

from gpu_util import device_available
from mm_propagation_amd import _capi

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols(header):
    text = open(os.path.join(REPO, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mmp_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported(capi_lib):
    syms = declared_symbols("mmprop.h")
    assert len(syms) >= 19
    for name in syms:
        assert hasattr(capi_lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", _capi.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (mmp_\w+)", out))
    assert set(syms) <= exported
    # the ctypes table covers the whole header
    assert {n for n, _, _ in _capi.EXPORTS} == set(syms)


def test_abi_version_and_status_strings(capi_lib):
    assert capi_lib.mmp_abi_version() == 6
    for code, name in _capi.STATUS_NAMES.items():
        assert capi_lib.mmp_status_string(code).decode() == name


def test_struct_layouts_match_the_c_compiler(tmp_path):
    """Every ctypes structure has the size and field offsets gcc gives the C declaration."""
    structs = {"mmp_params": _capi.MmpParams, "mmp_options": _capi.MmpOptions, "mmp_info": _capi.MmpInfo,
               "mmp_kernel_time": _capi.MmpKernelTime, "mmp_serve_options": _capi.MmpServeOptions,
               "mmp_serve_record": _capi.MmpServeRecord, "mmp_serve_stats": _capi.MmpServeStats}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "mmprop.h"', "int main(void) {"]
    for cname, py in structs.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(REPO, "include"), str(src), "-o", str(exe)], check=True)
    got = {}
    for line in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.splitlines():
        a, b, c = line.split()
        got[(a, b)] = int(c)
    for cname, py in structs.items():
        assert got[(cname, "size")] == ctypes.sizeof(py), cname
        for fname, _ in py._fields_:
            assert got[(cname, fname)] == getattr(py, fname).offset, (cname, fname)


@pytest.mark.skipif(device_available(), reason="a device is present")
def test_no_device_is_an_error_not_a_fallback(capi_lib):
    h = ctypes.c_void_p()
    st = capi_lib.mmp_ctx_create(0, ctypes.byref(h))
    assert st == 5  # MMP_ERR_NO_DEVICE
    assert capi_lib.mmp_last_error().decode()
    from mm_propagation_amd import configs
    from mm_propagation_amd.solver import MultimodePropagationGPU_RamanON
    with pytest.raises(_capi.MmpError):
        MultimodePropagationGPU_RamanON(configs.two_mode_gaussian(nt=64))


def test_library_matches_the_sources(capi_lib):
    """The loaded library carries the digest of the sources it was built from: the tree's."""
    from mm_propagation_amd import build
    assert capi_lib.mmp_source_hash().decode() == build.source_hash()
