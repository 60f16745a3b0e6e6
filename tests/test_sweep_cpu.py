"""The sweep command line (cuopo3d_amd.sweep) without a GPU: range parsing
as the awk loops of src/cuOPO3D.sh:56-86 and the .dat text format of
Files.cuh:73-94 / :180-197."""
import numpy as np

from cuopo3d_amd import sweep


def test_parse_range_matches_awk_loops():
    # awk 'BEGIN{for(i=500; i<=500; i=i+10)print i}' -> 500
    assert sweep.parse_range("500:500:10") == [500.0]
    assert sweep.parse_range("500:540:10") == [500.0, 510.0, 520.0, 530.0, 540.0]
    assert sweep.parse_range("31:45:5") == [31.0, 36.0, 41.0]
    assert sweep.parse_range("0.5:0.7:0.1") == [0.5, 0.6, 0.7]
    assert sweep.parse_range("250") == [250.0]
    assert sweep.parse_range("1,4,9") == [1.0, 4.0, 9.0]
    assert sweep.parse_range("5:4:1") == []


def test_dat_writer_format(tmp_path):
    v = np.array([[1 + 2j, 0.1 - 3j], [-0.5 + 0j, 1e-30 + 7j]], dtype=np.complex64)
    stem = str(tmp_path / "slice")
    sweep.write_dat(stem, v, rows=True)
    r = open(stem + "_r.dat").read()
    i = open(stem + "_i.dat").read()
    # one line per y row, every value followed by a tab (saveMatrixComplex_TimeSlice)
    assert r.splitlines()[0].endswith("\t") and len(r.splitlines()) == 2
    back_r = np.array([[float(x) for x in line.split("\t") if x] for line in r.splitlines()])
    back_i = np.array([[float(x) for x in line.split("\t") if x] for line in i.splitlines()])
    assert np.array_equal(back_r.astype(np.float32), v.real)
    assert np.array_equal(back_i.astype(np.float32), v.imag)
    # 20 significant digits as setprecision(20)
    assert r.split("\t")[1] == f"{np.float32(0.1):.20g}"
    stem = str(tmp_path / "trace")
    sweep.write_dat(stem, v[:, 0], rows=False)
    lines = open(stem + "_r.dat").read().splitlines()
    assert [float(x) for x in lines] == [1.0, -0.5]


def test_reference_physical_is_the_reference_cavity():
    ph = sweep.reference_physical(10)
    assert ph.lcr == 29.0e3 and ph.lcav == 4 * 29.0e3
    assert ph.resonant_s == 1 and ph.resonant_i == 1
    assert ph.rs == 0.98 and ph.ri == 0.98
    assert ph.dq == 0.0  # the crystal's own dQ = 2 d33 / pi
