"""The lorasp command line, the drop-in for the reference driver
(main.cpp:36-217, README's `lorasp N theta leaf eps rank sampling`): its
report lines, the 25 CSV columns plus the residual history of write_to_csv
(test_funcs.hpp:94-111), the HSS preconditioner switch (main.cpp:137-140) and
leveled ranks (solver.cpp:28-36)."""
import os
import re
import subprocess

import numpy as np
import pytest

import oracle_lib as O
from mxc_amd.solver import H2, Problem

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LORASP = os.path.join(ROOT, "mxc_amd", "lib", "lorasp")
REPORT = ["Construct Err", "H^2-Matrix Construct Time", "H^2-Matvec Time", "Dense Matvec Time",
          "H^2-Preconditioner Construct Time", "H^2-Preconditioner Construct Err", "H^2-Matrix Factorization Time",
          "H^2-Matrix Substitution Time", "H^2-Matrix Substitution Err", "GMRES Residual", "GMRES Time"]


def run(args, tmp_path):
    csv = tmp_path / "out.csv"
    p = subprocess.run([LORASP, *map(str, args), "--csv", str(csv)], capture_output=True, text=True, timeout=600)
    print(p.stdout)
    assert p.returncode == 0, p.stderr
    lines = p.stdout.splitlines()
    rep = {}
    for ln in lines:
        key, _, val = ln.partition(":")
        rep.setdefault(key.strip(), val.strip())
    rows = [r.split(",") for r in csv.read_text().splitlines()]
    return rep, lines, rows


@pytest.mark.timeout(600)
def test_lorasp_cube_20000_report_and_csv(tmp_path):
    rep, lines, rows = run([20000, 2, 256, 1e-10, 100, 2000], tmp_path)
    for k in REPORT:
        assert k in rep, k
    cerr = float(rep["Construct Err"])
    resid, iters = re.match(r"([0-9.e+-]+), Iters: (\d+)", rep["GMRES Residual"]).groups()
    iters = int(iters)
    hist = [float(re.match(r"iter (\d+): (.*)", ln).group(2)) for ln in lines if ln.startswith("iter ")]
    assert len(hist) == iters + 1 and hist[0] == 1.0
    assert cerr < 1e-10 and float(resid) < 1e-10
    assert float(rep["Dense Check Residual"]) < 1e-9
    assert len(rows) == 1
    row = rows[0]
    assert len(row) == 25 + iters + 1
    assert row[0] == "1" and row[1] == "20000" and row[2] == "2" and row[3] == "256" and row[4] == "100"
    assert float(row[5]) == 1e-10 and row[6] == "h2"
    assert float(row[7]) == pytest.approx(cerr, rel=1e-5)
    assert float(row[15]) == pytest.approx(float(rep["H^2-Preconditioner Construct Err"]), rel=1e-5)
    assert float(row[20]) == pytest.approx(float(rep["H^2-Matrix Substitution Err"]), rel=1e-5)
    assert float(row[21]) == pytest.approx(float(resid), rel=1e-5) and float(row[22]) == iters
    assert np.allclose([float(v) for v in row[25:]], hist, rtol=1e-5)


@pytest.mark.timeout(600)
def test_lorasp_hss_mode(tmp_path):
    rep, lines, rows = run([8192, 2, 256, 1e-10, 64, 2000, "--mode", "hss"], tmp_path)
    assert rows[0][6] == "hss"
    assert float(re.match(r"([0-9.e+-]+),", rep["GMRES Residual"]).group(1)) < 1e-10
    assert float(rep["Dense Check Residual"]) < 1e-9


@pytest.mark.timeout(600)
def test_leveled_rank_grows_per_level(ctx):
    """rank_func(l) = (L - l) * leveled_rank + rank caps the fixed ranks of
    level l (solver.cpp:29); the device construction keeps the oracle's ranks."""
    n, leaf, rank, lr, samp = 8192, 128, 16, 8, 2000
    pts = O.cube_points(n, 4)
    prob = Problem(ctx, pts, leaf)
    M = H2.preconditioner(prob, rank, samp, leveled_rank=lr)
    oM = O.Problem(pts, leaf).h2(rank=rank, leveled_rank=lr, orth=True, hss_far=True, sampling=samp)
    L = prob.levels
    grew = False
    for l in range(1, L + 1):
        d, r = M.dims(l)
        cap = (L - l) * lr + rank
        assert np.all(r <= cap) and np.array_equal(r, oM.dims(l)[1])
        assert np.array_equal(r, np.minimum(d, cap))
        grew = grew or r.max() > rank
    assert grew
    M.close()
    prob.close()


@pytest.mark.timeout(600)
def test_serial_sweep_log_and_csv(tmp_path):
    from mxc_amd.sweep import serial

    log, csv = serial(str(tmp_path / "log"), sizes=(2048, 4096))
    lines = open(log).read().splitlines()
    assert sum(ln.startswith("Construct Err:") for ln in lines) == 2
    assert sum(ln.startswith("GMRES Residual:") for ln in lines) == 2
    rows = [r.split(",") for r in open(csv).read().splitlines()]
    assert [r[1] for r in rows] == ["2048", "4096"]
    for r in rows:
        assert r[2] == "2" and r[3] == "256" and r[4] == "100" and float(r[5]) == 1e-10
        iters = int(float(r[22]))
        assert len(r) == 25 + iters + 1
        assert float(r[21]) < 1e-10
