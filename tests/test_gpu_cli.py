"""The `fast tau` CLI end to end on the 64^3 sphere volume (radius 6, solid 0.60, seed 0):
X_POS, BiCGSTAB, FP64, tol 1e-6 -> the CSV row's tau and iteration count equal the oracle's,
and the reachable-pore mask written with --reach equals the oracle's bit for bit."""
import csv
import subprocess
from pathlib import Path

import numpy as np
import pytest

import fast_amd as F
from oracle import oracle as O

pytestmark = pytest.mark.gpu
EXE = Path(__file__).resolve().parent.parent / "fast_amd" / "fast"


def test_cli_config1_tau_and_reachable_mask(tmp_path):
    m = F.generate_spheres(64, 6.0, 0.60, 0)
    vin, csv_out, reach = tmp_path / "v.bin", tmp_path / "out.csv", tmp_path / "reach.bin"
    F.save_bin(str(vin), m)
    cmd = [str(EXE), "tau", str(vin), "-d", "x", "--solver", "BICGSTAB", "--tol", "6", "--iter", "10000",
           "-p", "double", "-o", str(csv_out), "--reach", str(reach)]
    out = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    print(out.stdout, out.stderr)
    assert out.returncode == 0
    rows = list(csv.DictReader(open(csv_out)))
    assert len(rows) == 1
    ref, _ = O.tau(m, 0, solver=0, tol=1e-6)
    # the CSV carries the reference's default 6-significant-digit formatting
    assert float(rows[0]["tau"]) == pytest.approx(ref.tau, rel=5e-6)
    assert int(rows[0]["tauIter"]) == ref.iterations
    mask_ref, _, _ = O.reachable(m, 0)
    got = F.load_bin(str(reach))
    assert got.shape == m.shape
    assert np.array_equal(got, mask_ref)
