"""The device line operators compiled for the host (g++, one thread; the
stand-in header tests/emu/include/hip/hip_runtime.h turns the device
qualifiers into plain C++) against an FP64 DFT evaluated term by term:
two-level and three-level time lines in the padded tile layout of the fused
pass, x/y lines of the xy planes with the wave-uniform constant path, and
the row and column bands of planes beyond the LDS (256 and 512 points).
The chunked HBM <-> LDS copy is checked for coverage the same way.
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


CASES = ["t256_two_level", "t512_two_level", "t1024_two_level", "t256_three_level", "t512_three_level", "t1024_three_level",
         "t2048_three_level", "t2048_three_level_one_pixel", "t4096_three_level_one_pixel",
         "t2048_three_level_pad16", "x32_rows", "x64_rows_uniform",
         "x64_columns_uniform", "x96_columns", "x128_rows_uniform",
         "b256_rows", "b256_columns", "b512_rows_three_level", "b512_columns_three_level"]
COPY_CASES = ["copy_f32_16384_512", "copy_f64_8192_256", "copy_f32_9216_256", "copy_f64_9216_1024",
              "copy_f32_3000_256", "copy_f64_1024_256"]


@pytest.mark.parametrize("case", CASES)
def test_line_operator_matches_fp64_dft(results, case):
    err = results[case]
    print(case, err)
    assert err < 5e-7, (case, err)


@pytest.mark.parametrize("case", COPY_CASES)
def test_chunked_copy_touches_every_element_once(results, case):
    """The HBM <-> LDS copy helper of the band and plane kernels: over all
    threads of a workgroup, every element is read and written exactly once,
    no index leaves [0, TOT), and each value lands at its own index
    (uneven counts, partial last chunks and the plain-loop case)."""
    assert results[case] == 0, case

