"""The C-ABI libraries load and export every function their headers declare
(no device work)."""
import os
import re
import subprocess

import pytest

from neuronephys_clean_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "nnx.h")


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


STRUCTS = {"nnx_hh_state": "HHState", "nnx_hh_params": "HHParams", "nnx_graph": "Graph",
           "nnx_grid": "Grid", "nnx_cg_opts": "CgOpts", "nnx_cg_result": "CgResult",
           "nnx_step_plan": "StepPlan", "nnx_ens_plan": "EnsPlan", "nnx_ens_config": "EnsConfig",
           "nnx_regions": "Regions", "nnx_hebbian": "Hebbian",
           "nnx_diekman_state": "DiekmanState", "nnx_diekman_params": "DiekmanParams",
           "nnx_san_state": "SanState", "nnx_san_params": "SanParams", "nnx_paracrine": "Paracrine"}


def test_struct_layouts_match_the_c_compiler(tmp_path):
    """sizeof and every field offset of the ctypes mirrors equal what gcc
    computes from include/nnx.h."""
    import ctypes as C
    import shutil
    import subprocess
    if shutil.which("gcc") is None:
        pytest.skip("no host C compiler")
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', "int main(void) {"]
    for cname, pyname in STRUCTS.items():
        lines.append(f'printf("{pyname} size %zu\\n", sizeof({cname}));')
        for fname, _ in getattr(_lib, pyname)._fields_:
            lines.append(f'printf("{pyname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-o", str(exe), str(src)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    for line in filter(None, out):
        pyname, field, val = line.split()
        cls = getattr(_lib, pyname)
        got = C.sizeof(cls) if field == "size" else getattr(cls, field).offset
        assert got == int(val), (pyname, field, got, val)
