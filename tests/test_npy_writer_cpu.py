"""opo3d::saveNpy (include/opo3d.hpp), the binary alternative to the
reference's .dat text writers: numpy reads back the shape (NT, NY, NX), the
dtype and every value."""
import os
import shutil
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def test_save_npy_round_trip(tmp_path):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path / "npy_writer")
    subprocess.run([gxx, "-O1", "-std=c++17", "-I", os.path.join(ROOT, "include"), "-o", exe,
                    os.path.join(HERE, "emu", "npy_writer.cpp")], check=True)
    f32, f64 = str(tmp_path / "a32.npy"), str(tmp_path / "a64.npy")
    subprocess.run([exe, f32, f64], check=True)
    t, y, x = np.meshgrid(np.arange(7), np.arange(3), np.arange(5), indexing="ij")
    want = (t + 0.5 * y) + 1j * (x - 0.25 * t)
    for path, dt in ((f32, np.complex64), (f64, np.complex128)):
        a = np.load(path)
        assert a.dtype == dt and a.shape == (7, 3, 5)
        assert np.array_equal(a, want.astype(dt))
