"""The device line operators compiled for the host (g++, one thread; the
stand-in header tests/emu/include/hip/hip_runtime.h turns the device
qualifiers into plain C++) against an FP64 DFT evaluated term by term:
two-level and three-level time lines in the padded tile layout of the fused
pass, and x/y lines of the xy planes with the wave-uniform constant path.
FP32 arithmetic with (hi, lo) constants reaches ~1.5e-7 relative L2 on one
operator; 5e-7 fails a wrong sign, index or twiddle at once."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "emu", "line_ops.cpp")


@pytest.fixture(scope="module")
def results(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path_factory.mktemp("emu") / "line_ops")
    subprocess.run([gxx, "-O2", "-std=c++17", "-I", os.path.join(HERE, "emu", "include"), "-o", exe, SRC],
                   check=True)
    out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    return {ln.split()[0]: float(ln.split()[1]) for ln in out.strip().splitlines()}


CASES = ["t256_two_level", "t512_two_level", "t1024_two_level", "t512_three_level", "t1024_three_level",
         "t2048_three_level", "t2048_three_level_one_pixel", "x32_rows", "x64_rows_uniform",
         "x64_columns_uniform", "x96_columns", "x128_rows_uniform"]


@pytest.mark.parametrize("case", CASES)
def test_line_operator_matches_fp64_dft(results, case):
    err = results[case]
    print(case, err)
    assert err < 5e-7, (case, err)


FUSED_SRC = os.path.join(HERE, "emu", "fused_tile.cpp")
FUSED_CASES = ["f32_nt128_tx8_mid", "f32_nt256_tx16_mid", "f32_nt256_tx4_first", "f32_nt1024_tx4_mid",
               "f32_nt2048_tx2_mid", "f32_nt256_tx8_tripend_gdd", "f32_nt1024_tx2_tripend_gdd",
               "f32_nt128_tx16_tripend_nogdd", "f64_nt256_tx4_mid", "f64_nt512_tx2_mid",
               "f64_nt1024_tx2_tripend_gdd"]


@pytest.fixture(scope="module")
def fused(tmp_path_factory):
    """The fused t-pass tile (outer line-operator level in registers, straight
    from and to the fields) against the classic tile, both compiled for the
    host and run as one thread over every tile of a small grid."""
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path_factory.mktemp("emu") / "fused_tile")
    subprocess.run([gxx, "-O2", "-std=c++17", "-ffp-contract=off", "-I", os.path.join(HERE, "emu", "include"),
                    "-o", exe, FUSED_SRC], check=True)
    out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    return {ln.split()[0]: (int(ln.split()[1]), int(ln.split()[2])) for ln in out.strip().splitlines()}


@pytest.mark.parametrize("case", FUSED_CASES)
def test_fused_tile_is_bitwise_the_classic_tile(fused, case):
    diff, changed = fused[case]
    assert changed > 0, case     # the pass did act on the fields
    assert diff == 0, (case, diff)
