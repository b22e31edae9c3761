"""The C-ABI libraries load and export every function their headers declare
(no device work)."""
import os
import re
import subprocess

import pytest

from neuronephys_clean_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(nnx_[a-z0-9_]+)\s*\(", text)))


def exported(so):
    out = subprocess.run(["nm", "-D", "--defined-only", so], capture_output=True, text=True,
                         check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if " T " in line}


@pytest.mark.parametrize("header,so,sigs", [
    ("nnx.h", _lib.LIB_PATH, 0),
    ("nnx_host.h", _lib.HOST_LIB_PATH, 1),
])
def test_every_declared_symbol_is_exported_and_bound(header, so, sigs):
    names = declared(header)
    assert len(names) > 10
    syms = exported(so)
    missing = [n for n in names if n not in syms]
    assert not missing, missing
    bound = set(_lib.exported_symbols()[sigs])
    assert not [n for n in names if n not in bound]


def test_libraries_load_without_gpu():
    assert _lib.lib().nnx_version().decode().startswith("nnx")
    p = _lib.params_squid()
    assert (p.ena, p.ek, p.el, p.v_shift) == (50.0, -77.0, -54.4, 5.0)
    r = _lib.params_reference()
    assert (r.ena, r.ek, r.el, r.out_thresh) == (55.0, -72.0, -50.0, 0.1)


def test_struct_sizes_match_header_layout():
    import ctypes as C
    assert C.sizeof(_lib.HHState) == 16 + 19 * 8
    assert C.sizeof(_lib.HHParams) == 26 * 8
    assert C.sizeof(_lib.Graph) == 16 + 7 * 8
    assert C.sizeof(_lib.Grid) == 16 + 16 + 24
    assert C.sizeof(_lib.CgOpts) == 32
    assert C.sizeof(_lib.CgResult) == 16 + 24
