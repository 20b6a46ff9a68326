"""The matrix-set runner (mille-feuille_amd/matrix_set.py) against the budget, discovery and
output rules of the reference's benchmark script (test_Mille_feuille.sh:41-98): budgets from
the CSV or MAX_IT, ceil(1.25 x) for the mixed budget, <Name>.mtx directly under MATRIX_ROOT or
anywhere below it, MAX_MATS, and the CSV rows the CLIs append (on the GPU)."""
import csv
import importlib.util
import math
import os
import subprocess
import sys
from fractions import Fraction

import pytest

from conftest import ROOT

PKG_DIR = ROOT / "mille-feuille_amd"
SCRIPT = PKG_DIR / "matrix_set.py"


def _runner():
    spec = importlib.util.spec_from_file_location("mf_matrix_set", SCRIPT)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def _poisson_mtx(path, m):
    """Lower triangle of the 2D 5-point Laplacian on an m x m grid, as a symmetric file."""
    lines = []
    for i in range(m):
        for j in range(m):
            k = i * m + j
            lines.append(f"{k + 1} {k + 1} 4.0")
            if j > 0:
                lines.append(f"{k + 1} {k} -1.0")
            if i > 0:
                lines.append(f"{k + 1} {k + 1 - m} -1.0")
    with open(path, "w") as f:
        f.write("%%MatrixMarket matrix coordinate real symmetric\n")
        f.write(f"{m * m} {m * m} {len(lines)}\n")
        f.write("\n".join(lines) + "\n")


def _matrix_set(tmp_path):
    """A set of four names: one file directly under the root, one found twice in
    subdirectories, one absent, one with the budgets left empty."""
    root = tmp_path / "matrices"
    (root / "g1").mkdir(parents=True)
    (root / "g2" / "deep").mkdir(parents=True)
    _poisson_mtx(root / "alpha.mtx", 16)
    _poisson_mtx(root / "g1" / "beta.mtx", 20)
    _poisson_mtx(root / "g2" / "deep" / "beta.mtx", 24)
    _poisson_mtx(root / "g1" / "delta.mtx", 32)
    set_csv = tmp_path / "set.csv"
    set_csv.write_text("id,Group,Name,rows,cols,entries,max_iter,max_iter_mix\n"
                       "1,G,alpha,256,256,1216,40,55\n"
                       "2,G,beta,400,400,1920,33,\n"
                       "3,G,gamma,10,10,10,5,7\n"
                       "4,G,delta,1024,1024,4992,,\n")
    return root, set_csv


def test_mixed_budget_is_ceil_five_quarters():
    ms = _runner()
    for x in list(range(0, 2000)) + [9999, 10000, 123457]:
        assert ms.mixed_budget(x) == math.ceil(Fraction(5, 4) * x)


def test_duration_parsing():
    ms = _runner()
    assert ms.parse_limit("4m") == 240.0
    assert ms.parse_limit("90") == 90.0
    assert ms.parse_limit("2.5s") == 2.5
    assert ms.parse_limit("1h") == 3600.0


def test_plan_budgets_and_discovery(tmp_path):
    ms = _runner()
    root, set_csv = _matrix_set(tmp_path)
    env = {"MATRIX_ROOT": str(root)}
    runs = ms.plan(str(set_csv), env)
    got = [(p.relative_to(root).as_posix(), it, mix) for p, it, mix in runs]
    assert got == [("alpha.mtx", 40, 55),
                   ("g1/beta.mtx", 33, 42),             # ceil(1.25 * 33) = 42
                   ("g2/deep/beta.mtx", 33, 42),
                   ("g1/delta.mtx", 10000, 12500)]      # empty budgets: the defaults
    # MAX_IT overrides every row, the mixed budget follows it
    runs = ms.plan(str(set_csv), dict(env, MAX_IT="7"))
    assert [(it, mix) for _, it, mix in runs] == [(7, 9)] * 4
    # MAX_MATS counts matrix files, also inside one name's matches
    runs = ms.plan(str(set_csv), dict(env, MAX_MATS="2"))
    assert [p.name for p, _, _ in runs] == ["alpha.mtx", "beta.mtx"]
    assert ms.plan(str(set_csv), dict(env, MATRIX_ROOT=str(tmp_path / "none"))) == []


def test_dry_run_prints_the_reference_argv(tmp_path):
    root, set_csv = _matrix_set(tmp_path)
    env = dict(os.environ, MATRIX_ROOT=str(root), MAX_MATS="1", OUT_CSV_CG="a.csv", OUT_CSV_MIXED="b.csv")
    r = subprocess.run([sys.executable, str(SCRIPT), str(set_csv), "--dry-run"], capture_output=True, text=True,
                       env=env, timeout=60)
    assert r.returncode == 0, r.stderr
    lines = [ln.strip() for ln in r.stdout.splitlines()]
    assert lines[0] == f"RUN {root / 'alpha.mtx'} (max_iter=40, max_iter_mix=55)"
    assert lines[1] == f"{PKG_DIR / 'main-cg'} {root / 'alpha.mtx'} 40 a.csv"
    assert lines[2] == f"{PKG_DIR / 'main-cg-mixed'} {root / 'alpha.mtx'} 55 b.csv"
    assert lines[3].startswith("Done.")


def test_missing_set_fails(tmp_path):
    r = subprocess.run([sys.executable, str(SCRIPT), str(tmp_path / "nope.csv")], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 1 and "not found" in r.stderr


@pytest.mark.gpu
def test_matrix_set_runs_both_clis(tmp_path):
    """Four matrix files through main-cg and main-cg-mixed: one CSV row per file in each
    results file, with the budget of its row as the iteration count."""
    for b in ("main-cg", "main-cg-mixed"):
        if not os.access(PKG_DIR / b, os.X_OK):
            pytest.skip(f"{b} not built")
    root, set_csv = _matrix_set(tmp_path)
    out_cg, out_mix = tmp_path / "res" / "cg.csv", tmp_path / "res" / "mixed.csv"
    env = dict(os.environ, MATRIX_ROOT=str(root), MAX_IT="25", OUT_CSV_CG=str(out_cg), OUT_CSV_MIXED=str(out_mix),
               TIMEOUT_LIMIT="90s")
    print("matrix set: 4 files x 2 CLIs, MAX_IT=25", flush=True)
    r = subprocess.run([sys.executable, str(SCRIPT), str(set_csv)], capture_output=True, text=True, env=env,
                       timeout=600)
    assert r.returncode == 0, r.stderr
    assert "killed" not in r.stdout and "exited" not in r.stdout, r.stdout
    for path, variant, it in ((out_cg, "cg", 25), (out_mix, "cg_mixed", 32)):
        with open(path) as f:
            rows = list(csv.DictReader(f))
        assert len(rows) == 4, rows
        for row in rows:
            assert row["variant"] == variant and int(row["iterations"]) == it
            assert float(row["l2_norm"]) < 1.0
