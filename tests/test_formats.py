"""On-disk formats (libnnx_host.so): the HDF5 reader against fixtures written
by the HDF5 library (tests/golden/make_h5.c, both the earliest and the
latest file-format versions), the reference's file loaders
(FileNetwork::GenerateFromHDF, src/network_factory.cu:171-212;
FileMorphology::GenerateFromHDF, src_wholebrain/morphology.cpp:301-370) on
hand-derived examples, and the text CSR format of
Csr_matrix::input_from_connection_file (src/paracrine.cuh:49-108)."""
import os

import numpy as np
import pytest

from neuronephys_clean_amd import host

H5 = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "h5")
TAGS = ["earliest", "latest"]


def f(name):
    return os.path.join(H5, name)


@pytest.mark.parametrize("tag", TAGS)
def test_network_and_morphology_datasets(tag):
    i = np.arange(51)
    assert host.h5_read(f(f"net_{tag}.h5"), "/row_ptr", np.int32).tolist() == (3 * i).tolist()
    e = np.arange(150)
    assert host.h5_read(f(f"net_{tag}.h5"), "col_idx", np.int32).tolist() == ((7 * e + 3) % 50).tolist()
    p = np.arange(200)
    exp = {"x": np.float32(0.25) * p - np.float32(3), "y": np.float32(100) - np.float32(0.5) * p,
           "z": (p % 13).astype(np.float32) * np.float32(1.5), "region_id": (p % 5).astype(np.float32)}
    for k, v in exp.items():
        assert np.array_equal(host.h5_read(f(f"morph_{tag}.h5"), "/" + k, np.float32), v.astype(np.float32)), k


@pytest.mark.parametrize("tag", TAGS)
def test_types_layouts_and_filters(tag):
    path = f(f"types_{tag}.h5")
    assert np.array_equal(host.h5_read(path, "/grp/sub/f64be", np.float64), np.arange(64) / 8.0)
    assert host.h5_read(path, "grp/i16", np.int32).tolist() == list(range(-20, 20))
    assert host.h5_read(path, "i64", np.int64).tolist() == [k << 33 for k in range(16)]
    assert np.array_equal(host.h5_read(path, "compact_f32", np.float32),
                          np.float32(0.25) * np.arange(200) - np.float32(3))
    big = (1000003 * np.arange(10000, dtype=np.int64)) % 2 ** 31
    assert np.array_equal(host.h5_read(path, "chunked_i32", np.int64), big)      # deflate + shuffle
    assert np.array_equal(host.h5_read(path, "chunked_plain", np.int64), big)
    # conversions are C casts
    assert host.h5_read(path, "compact_f32", np.int32)[:4].tolist() == [-3, -2, -2, -2]
    assert host.h5_read(path, "grp/i16", np.float64)[0] == -20.0


def test_symbol_table_with_many_nodes():
    for k in range(40):
        assert host.h5_read(f("many_earliest.h5"), f"d{k:02d}", np.int32).tolist() == [k, k + 1, k + 2, k + 3]


def test_unsupported_and_broken_files(tmp_path):
    with pytest.raises(IOError, match="dense link storage"):
        host.h5_read(f("many_latest.h5"), "d03", np.int32)
    with pytest.raises(IOError, match="no object named"):
        host.h5_read(f("net_earliest.h5"), "/nope", np.int32)
    with pytest.raises(IOError, match="cannot open"):
        host.h5_read(str(tmp_path / "missing.h5"), "/x")
    txt = tmp_path / "plain.h5"
    txt.write_text("not hdf5 at all" * 100)
    with pytest.raises(IOError, match="not an HDF5 file"):
        host.h5_read(str(txt), "/x")
    big = (1000003 * np.arange(10000, dtype=np.int64)) % 2 ** 31
    for name in ("types_earliest.h5", "types_latest.h5"):
        data = open(f(name), "rb").read()
        errors = 0
        for cut in range(200, len(data), 997):  # truncated copies: an error or the right data, never a crash
            t = tmp_path / "cut.h5"
            t.write_bytes(data[:cut])
            try:
                assert np.array_equal(host.h5_read(str(t), "chunked_i32", np.int64), big)
            except IOError:
                errors += 1
        assert errors > 10


def test_subsample_hand_example():
    # rows: 0 -> {0,1,2}, 1 -> {3}, 2 -> {0,2}, 3 -> {1,3}; N = 2, sparsity 2:
    # neuron 0 = row 0 keeps targets 0, 2 -> 0, 1; neuron 1 = row 2 keeps 0, 2 -> 0, 1
    rp = [0, 3, 4, 6, 8]
    col = [0, 1, 2, 3, 0, 2, 1, 3]
    off, tgt, dropped = host.csr_subsample(rp, col, 2)
    assert off.tolist() == [0, 2, 4] and tgt.tolist() == [0, 1, 0, 1] and dropped == 0
    # 5 points, N = 2: sparsity 2, a target 4 maps to 2 >= N and is dropped
    off, tgt, dropped = host.csr_subsample([0, 2, 2, 3, 3, 3], [4, 2, 4], 2)
    assert off.tolist() == [0, 1, 1] and tgt.tolist() == [1] and dropped == 2
    # N = numpnts keeps everything
    off, tgt, _ = host.csr_subsample(rp, col, 4)
    assert off.tolist() == rp and tgt.tolist() == col
    with pytest.raises(ValueError):
        host.csr_subsample(rp, col, 5)


def test_hdf5_network_loader():
    off, tgt, dropped = host.load_hdf5_network(f("net_earliest.h5"), 25)
    # sparsity 2: neuron i = file row 2i (targets (7e+3) % 50 for e in 6i..6i+2), even targets halved
    exp_off, exp_tgt = [0], []
    for i in range(25):
        for e in range(6 * i, 6 * i + 3):
            c = (7 * e + 3) % 50
            if c % 2 == 0:
                exp_tgt.append(c // 2)
        exp_off.append(len(exp_tgt))
    assert off.tolist() == exp_off and tgt.tolist() == exp_tgt and dropped == 0


def test_morphology_hand_example():
    x = [0.0, 10.0, 5.0, 2.5]
    y = [-1.0, 1.0, 0.0, 0.0]
    z = [3.0, 3.0, 3.0, 4.0]
    pos = host.morphology_from_axes(x, y, z, 2)  # points 0 and 2, boxes over all four
    assert pos.tolist() == [[-1.0, -1.0, -1.0], [0.0, 0.0, -1.0]]
    with pytest.raises(ValueError):
        host.morphology_from_axes(x, y, z, 5)


def test_hdf5_morphology_loader():
    pos, rid = host.load_hdf5_morphology(f("morph_latest.h5"), 50, regions=True)
    assert pos.shape == (50, 3) and pos.min() >= -1 and pos.max() <= 1
    assert pos[0].tolist() == [-1.0, 1.0, -1.0]  # x min, y max, z min at point 0
    assert rid.tolist() == [(4 * i) % 5 for i in range(50)]


def test_connection_file(tmp_path):
    p = tmp_path / "conn.txt"
    p.write_text("3 4\n0 2 3 4\n1 2 0 1\n0.1 -2.5 3e-3 7\n")
    rp, col, val = host.read_connection_file(str(p))
    assert rp.tolist() == [0, 2, 3, 4] and col.tolist() == [1, 2, 0, 1]
    assert val.tolist() == np.array([0.1, -2.5, 3e-3, 7.0], np.float32).tolist()
    q = tmp_path / "short.txt"
    q.write_text("3 4\n0 2 3 4\n1 2 0\n0.1 -2.5 3e-3 7\n")
    with pytest.raises(IOError, match="short"):
        host.read_connection_file(str(q))
    with pytest.raises(IOError, match="cannot open"):
        host.read_connection_file(str(tmp_path / "none.txt"))
