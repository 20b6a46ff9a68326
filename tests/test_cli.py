"""Drop-in command-line tools: reference argv, CSV schema and exit behaviour.

Mirrors the reference drivers' mains (main-cg.cu:1682-1765, main-cg-mixed.cu:3071-3145,
main-bicg.cu:1696-1726).  Argument errors and non-square inputs end before any device call, so
those cases run without a GPU; the solves are marked gpu.
"""
import csv
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT

PKG_DIR = str(ROOT / "mille-feuille_amd")

CLIS = ["main-cg", "main-cg-mixed", "main-bicg", "main-bicg-mixed"]


def _cli(name):
    path = os.path.join(PKG_DIR, name)
    if not os.path.exists(path):
        pytest.skip(f"{name} not built")
    return path


def _write_mtx(path, n, rows, cols, vals, symmetric=False):
    with open(path, "w") as f:
        kind = "symmetric" if symmetric else "general"
        f.write(f"%%MatrixMarket matrix coordinate real {kind}\n")
        f.write(f"{n[0]} {n[1]} {len(vals)}\n")
        for r, c, v in zip(rows, cols, vals):
            f.write(f"{r + 1} {c + 1} {v!r}\n")


def _poisson_mtx(path, m):
    """Lower triangle of the 2D 5-point Laplacian on an m×m grid, as a symmetric file."""
    rows, cols, vals = [], [], []
    for i in range(m):
        for j in range(m):
            k = i * m + j
            rows.append(k); cols.append(k); vals.append(4.0)
            if j > 0:
                rows.append(k); cols.append(k - 1); vals.append(-1.0)
            if i > 0:
                rows.append(k); cols.append(k - m); vals.append(-1.0)
    _write_mtx(path, (m * m, m * m), rows, cols, vals, symmetric=True)


@pytest.mark.parametrize("name", CLIS)
def test_usage_exit_code(name):
    r = subprocess.run([_cli(name)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2
    assert "usage" in r.stderr


@pytest.mark.parametrize("name", CLIS)
def test_non_square_prints_unequal(tmp_path, name):
    p = str(tmp_path / "rect.mtx")
    _write_mtx(p, (3, 5), [0, 1, 2], [0, 3, 4], [1.0, 2.0, 3.0])
    args = [_cli(name), p] + (["10"] if "cg" in name and "bicg" not in name else [])
    r = subprocess.run(args, capture_output=True, text=True, timeout=60)
    assert r.returncode == 0
    assert "unequal" in r.stdout


def test_missing_file_fails(tmp_path):
    r = subprocess.run([_cli("main-cg"), str(tmp_path / "nope.mtx"), "10"], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 1


@pytest.mark.gpu
@pytest.mark.parametrize("name,variant", [("main-cg", "cg"), ("main-cg-mixed", "cg_mixed")])
def test_cg_cli_csv(tmp_path, orc, name, variant):
    """Fixed-iteration CG on 2D Poisson 48² (n = 2304, a multiple of 16): CSV row as the reference
    writes it, and the iteration budget honoured exactly."""
    mtx = str(tmp_path / "poisson48.mtx")
    _poisson_mtx(mtx, 48)
    out = str(tmp_path / "out.csv")
    print(f"cli {name} poisson48 max_iter=60", flush=True)
    r = subprocess.run([_cli(name), mtx, "60", out], capture_output=True, text=True, timeout=120,
                       cwd=str(tmp_path))
    assert r.returncode == 0, r.stderr
    assert f"time_{variant}_avg=" in r.stdout
    with open(out) as f:
        rows = list(csv.DictReader(f))
    assert len(rows) == 1
    row = rows[0]
    assert list(row.keys()) == ["matrix", "variant", "iterations", "time_ms", "nnz", "l2_norm", "residual"]
    assert row["variant"] == variant and int(row["iterations"]) == 60
    assert int(row["nnz"]) == 48 * 48 + 2 * 2 * 48 * 47
    # after 60 CG iterations on this system the oracle's relative residual is known; the FP64
    # true residual the tool reports must match it closely
    n = 48 * 48
    rp, ci, va = _poisson_csr(48)
    b = np.asarray([va[rp[i]:rp[i + 1]].sum() for i in range(n)])
    res = orc.cg(rp, ci, va, b, tol=0.0, maxit=60)
    assert res["iterations"] == 60
    rel = np.linalg.norm(b - orc.csr_spmv(rp, ci, va, res["x"])) / np.linalg.norm(b)
    # FP64: only the reduction order differs; mixed: binary32 products of exact int8 values
    assert float(row["l2_norm"]) == pytest.approx(rel, rel=1e-6 if variant == "cg" else 2e-2)


@pytest.mark.gpu
def test_cg_cli_tolerance_mode(tmp_path):
    mtx = str(tmp_path / "poisson40.mtx")
    _poisson_mtx(mtx, 40)
    out = str(tmp_path / "out.csv")
    env = dict(os.environ, MF_TOL="1e-10")
    r = subprocess.run([_cli("main-cg-mixed"), mtx, "5000", out], capture_output=True, text=True, timeout=120,
                       cwd=str(tmp_path), env=env)
    assert r.returncode == 0, r.stderr
    with open(out) as f:
        row = next(csv.DictReader(f))
    assert 0 < int(row["iterations"]) < 5000
    assert float(row["l2_norm"]) <= 1.2e-10


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["main-bicg", "main-bicg-mixed"])
@pytest.mark.parametrize("m", [32, 33, 120])
def test_bicg_cli_runs(tmp_path, name, m):
    """main-bicg keeps every row (33^2 = 1089 is not a multiple of 16), runs 50 iterations
    below 10^5 entries and the multi-kernel bicgstab above (120^2 Poisson: 71,520 entries
    -> bicgstab_sync; the mixed tool always runs 50), and appends
    path,iterations,time,nnz,l2,residual to data/bicg_performance.csv
    (main-bicg.cu:678-688, 1696-1726)."""
    mtx = str(tmp_path / f"poisson{m}.mtx")
    _poisson_mtx(mtx, m)
    # without data/ the row cannot be written: the reference prints "open error!"
    r0 = subprocess.run([_cli(name), mtx], capture_output=True, text=True, timeout=120, cwd=str(tmp_path))
    assert r0.returncode == 0, r0.stderr
    assert "open error!" in r0.stdout
    os.makedirs(tmp_path / "data")
    r = subprocess.run([_cli(name), mtx], capture_output=True, text=True, timeout=120, cwd=str(tmp_path))
    assert r.returncode == 0, r.stderr
    rows = (tmp_path / "data" / "bicg_performance.csv").read_text().strip().splitlines()
    assert len(rows) == 1
    f = rows[0].split(",")
    assert f[0] == mtx and int(f[1]) == 50 and int(f[3]) == 5 * m * m - 4 * m
    assert np.isfinite(float(f[4])) and float(f[2]) > 0


def _poisson_csr(m):
    n = m * m
    rp = [0]
    ci, va = [], []
    for i in range(m):
        for j in range(m):
            k = i * m + j
            ent = []
            if i > 0: ent.append((k - m, -1.0))
            if j > 0: ent.append((k - 1, -1.0))
            ent.append((k, 4.0))
            if j < m - 1: ent.append((k + 1, -1.0))
            if i < m - 1: ent.append((k + m, -1.0))
            for c, v in ent:
                ci.append(c); va.append(v)
            rp.append(len(ci))
    return np.asarray(rp, np.int32), np.asarray(ci, np.int32), np.asarray(va, np.float64)
