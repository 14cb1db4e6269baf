"""The Stockham passes of csrc/pv_fftk.h on the host (their load/store halves
are __host__ __device__): the mixed-radix transform used for lengths
2^a 3^b and the power-of-two one, forward and inverse, against a direct DFT
in long double.  Compiled with hipcc as host code; no GPU."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def exe(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("fftk") / "fftk_check")
    subprocess.run(["/opt/rocm/bin/hipcc", "-O2", "-ffp-contract=off", "-std=c++17", "-o", out,
                    os.path.join(HERE, "host", "fftk_check.cpp")], check=True, capture_output=True)
    return out


@pytest.mark.parametrize("M", [8, 16, 24, 48, 96, 192, 384, 768, 1536, 64, 512, 2048, 72, 216])
def test_stockham_matches_direct_dft(exe, M):
    out = subprocess.run([exe, str(M)], check=True, capture_output=True, text=True).stdout.split()
    fwd_g, fwd_p, inv_g, inv_p = float(out[1]), float(out[2]), float(out[4]), float(out[5])
    for e in (fwd_g, fwd_p, inv_g, inv_p):
        assert e < 5e-15 * max(1, M).bit_length(), (M, out)
