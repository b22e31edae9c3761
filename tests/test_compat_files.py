"""The reference-interface file classes (compat FileNetwork and
FileMorphology, nnx_compat.hpp) compiled by the host compiler and run on the
CPU against the Python loaders over the HDF5 fixtures."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from neuronephys_clean_amd import host

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "neuronephys_clean_amd")
H5 = os.path.join(ROOT, "tests", "golden", "h5")

SRC = r'''
#include <cstdio>
#include "nnx_compat.hpp"
using namespace nnx_compat;
int main(int argc, char** argv) {
    FileNetwork net(argv[1]);
    if (!net.Generate(std::atoi(argv[3]))) return 1;
    for (int v : net.GetCpuOffsetsIntoIndices()) std::printf("%d ", v);
    std::printf("\n");
    for (int v : net.GetCpuIndices()) std::printf("%d ", v);
    std::printf("\n");
    FileMorphology m(argv[2], true);
    if (!m.Generate(std::atoi(argv[4]))) return 2;
    for (float v : m.xyz) std::printf("%.9g ", v);
    std::printf("\n");
    for (int v : m.regions) std::printf("%d ", v);
    std::printf("\n%zu\n", m.unique_regions.size());
    FileNetwork bad("x.txt");
    return bad.Generate(3) ? 3 : 0;
}
'''


def test_compat_file_classes(tmp_path):
    if shutil.which("g++") is None:
        pytest.skip("no host C++ compiler")
    src = tmp_path / "files.cpp"
    src.write_text(SRC)
    exe = tmp_path / "files"
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(PKG, "compat"), "-o", str(exe), str(src),
                    "-L", PKG, "-lnnx", "-lnnx_host", f"-Wl,-rpath,{PKG}"], check=True)
    net, morph = os.path.join(H5, "net_earliest.h5"), os.path.join(H5, "morph_earliest.h5")
    r = subprocess.run([str(exe), net, morph, "25", "40"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.strip().split("\n")
    off, tgt, _ = host.load_hdf5_network(net, 25)
    assert [int(v) for v in lines[0].split()] == off.tolist()
    assert [int(v) for v in lines[1].split()] == tgt.tolist()
    pos, rid = host.load_hdf5_morphology(morph, 40, regions=True)
    assert np.array_equal(np.array(lines[2].split(), np.float32), pos.ravel())
    assert [int(v) for v in lines[3].split()] == rid.tolist()
    assert int(lines[4]) == len(set(rid.tolist()))
