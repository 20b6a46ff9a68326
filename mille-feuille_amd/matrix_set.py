#!/usr/bin/env python3
"""Runs the drop-in CLIs over a matrix-set CSV, the way the reference's benchmark script does
(test_Mille_feuille.sh): for every row, main-cg with the row's iteration budget and
main-cg-mixed with the mixed budget, each appending to its own results CSV.

Usage:
    python3 matrix_set.py [matrix_set.csv] [--dry-run]

Input CSV (header line, then one matrix per line; valid_matrix_set.csv's columns):
    id,Group,Name,rows,cols,entries,max_iter,max_iter_mix
Only Name (column 3) is required; max_iter and max_iter_mix may be empty.

Environment (names and defaults of test_Mille_feuille.sh:12-26):
    MATRIX_ROOT    directory searched for <Name>.mtx             (/data/matrix)
    MAX_IT         iteration budget; when set it overrides the CSV's budgets
    TIMEOUT_LIMIT  per-run limit, <number>[s|m|h]                 (4m)
    MAX_MATS       stop after this many matrices, 0 = no limit    (0)
    OUT_CSV_CG     main-cg results                                (mille_feuille_cg.csv)
    OUT_CSV_MIXED  main-cg-mixed results                          (mille_feuille_cg_mixed.csv)

Budgets (test_Mille_feuille.sh:41-45,69-78): with MAX_IT set, max_iter = MAX_IT and
max_iter_mix = ceil(1.25 MAX_IT); otherwise the row's max_iter (10000 when empty) and the
row's max_iter_mix (ceil(1.25 max_iter) when empty).  <Name>.mtx is taken directly under
MATRIX_ROOT, else every file of that name below it is run (:80-98).  A run that fails or
exceeds the limit is reported and the set continues (:47-50).  Relative output paths are
relative to the directory of the CLIs, where the script runs them (:59).

The CLIs are the ones built beside this file; the script starts them as child processes.
"""
from __future__ import annotations

import csv
import math
import os
import pathlib
import subprocess
import sys

HERE = pathlib.Path(__file__).resolve().parent
DEFAULT_MAX_IT = 10000


def mixed_budget(max_iter: int) -> int:
    """ceil(1.25 max_iter) in integers: (max_iter * 125 + 99) // 100."""
    return (max_iter * 125 + 99) // 100


def parse_limit(text: str) -> float:
    """Seconds of a coreutils-style duration: 30, 30s, 4m, 1h (fractions allowed)."""
    t = text.strip()
    scale = {"s": 1.0, "m": 60.0, "h": 3600.0, "d": 86400.0}
    if t and t[-1] in scale:
        return float(t[:-1]) * scale[t[-1]]
    return float(t)


def budgets(row: list[str], env: dict) -> tuple[int, int]:
    """(max_iter, max_iter_mix) of one CSV row under the environment's MAX_IT."""
    if env.get("MAX_IT"):
        it = int(env["MAX_IT"].strip())
        return it, mixed_budget(it)
    it_text = row[6].strip() if len(row) > 6 else ""
    it = int(it_text) if it_text else DEFAULT_MAX_IT
    mix_text = row[7].strip() if len(row) > 7 else ""
    return it, int(mix_text) if mix_text else mixed_budget(it)


def find_matrices(root: pathlib.Path, name: str) -> list[pathlib.Path]:
    direct = root / f"{name}.mtx"
    if direct.is_file():
        return [direct]
    if not root.is_dir():
        return []
    return sorted(p for p in root.rglob(f"{name}.mtx") if p.is_file())


def plan(csv_path: str, env: dict) -> list[tuple[pathlib.Path, int, int]]:
    """The (matrix, max_iter, max_iter_mix) runs of a matrix set, in order."""
    root = pathlib.Path(env.get("MATRIX_ROOT") or "/data/matrix")
    max_mats = int(env.get("MAX_MATS") or 0)
    runs: list[tuple[pathlib.Path, int, int]] = []
    with open(csv_path, newline="") as f:
        rows = csv.reader(f)
        next(rows, None)  # header
        for row in rows:
            if len(row) < 3 or not row[2].strip():
                continue
            it, it_mix = budgets(row, env)
            for m in find_matrices(root, row[2].strip()):
                if max_mats and len(runs) >= max_mats:
                    return runs
                runs.append((m, it, it_mix))
    return runs


def commands(run: tuple[pathlib.Path, int, int], out_cg: str, out_mixed: str, bindir: pathlib.Path = HERE):
    m, it, it_mix = run
    return [[str(bindir / "main-cg"), str(m), str(it), out_cg],
            [str(bindir / "main-cg-mixed"), str(m), str(it_mix), out_mixed]]


def main(argv: list[str]) -> int:
    dry = "--dry-run" in argv
    args = [a for a in argv if a != "--dry-run"]
    csv_path = args[0] if args else str(HERE / "valid_matrix_set.csv")
    if not os.path.isfile(csv_path):
        print(f"ERROR: input csv not found: {csv_path}", file=sys.stderr)
        return 1
    env = dict(os.environ)
    out_cg = env.get("OUT_CSV_CG") or "mille_feuille_cg.csv"
    out_mixed = env.get("OUT_CSV_MIXED") or "mille_feuille_cg_mixed.csv"
    limit = parse_limit(env.get("TIMEOUT_LIMIT") or "4m")
    if not dry:
        for b in ("main-cg", "main-cg-mixed"):
            if not os.access(HERE / b, os.X_OK):
                print(f"ERROR: {HERE / b} not found, run make", file=sys.stderr)
                return 1
        for out in (out_cg, out_mixed):
            parent = (HERE / out).parent
            parent.mkdir(parents=True, exist_ok=True)
    csv_path = os.path.abspath(csv_path)
    for run in plan(csv_path, env):
        print(f"RUN {run[0]} (max_iter={run[1]}, max_iter_mix={run[2]})", flush=True)
        for cmd in commands(run, out_cg, out_mixed):
            if dry:
                print("  " + " ".join(cmd), flush=True)
                continue
            try:
                r = subprocess.run(cmd, cwd=HERE, timeout=limit)
                if r.returncode != 0:
                    print(f"  {os.path.basename(cmd[0])} exited with {r.returncode}", flush=True)
            except subprocess.TimeoutExpired:
                print(f"  {os.path.basename(cmd[0])} killed after {limit:g} s", flush=True)
    print(f"Done. Results: cg -> {out_cg}, cg-mixed -> {out_mixed}")
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
