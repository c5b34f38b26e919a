"""Arithmetic identities the kernels rely on, proved exhaustively on the CPU."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_fma_division_by_three_is_ieee(tmp_path):
    """update_sig's pair kernel replaces the IEEE division of the squared
    distance by 3 with an FMA-corrected product; every non-negative finite
    float must give the same bits."""
    exe = str(tmp_path / "div3")
    subprocess.run(["gcc", "-O2", "-march=x86-64-v3", "-ffp-contract=off", "-fopenmp",
                    os.path.join(HERE, "arith", "div3_exhaustive.c"), "-o", exe, "-lm"], check=True)
    r = subprocess.run([exe], stdout=subprocess.PIPE, text=True, timeout=600)
    assert r.returncode == 0, r.stdout
    assert "mismatches: 0 " in r.stdout
