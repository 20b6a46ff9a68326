"""Per-phase timing of the variant-4 single-kernel CG, from libmf_phase.so.

    python3 mille-feuille_amd/phase_profile.py --workload poisson128 [--precision mixed]
                                               [--grid G] [--no-value-cache] [--json FILE]

Workloads (the benchmark matrices):
  poisson128  3D 7-point Poisson 128^3, b = A x*, x* ~ U[-1,1] seed 1, lossless classes
  varcoef64   variable-coefficient 3D diffusion 64^3 (k = exp(GRF), seed 4), row-scaled
              classes, b ~ U[-1,1] seed 4 (the differentiable solve's forward system)
  poisson64   2D 5-point Poisson 64 x 64, b = A x*, x* seed 0

The solve runs to 1e-10 three times (the first warms up).  Each workgroup's thread 0 reads a
100 MHz clock at the phase boundaries of every iteration; the table gives microseconds per
iteration: the mean over workgroups, the slowest workgroup and the root (logical workgroup
0, which gathers the grid reductions).  The timers add a block barrier before each grid
reduction, so the kernel time is compared with an uninstrumented run separately.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import pathlib
import sys

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parents[1]
PHASES = ["neighbour p wait", "SpMV", "local dots + block barrier", "grid reduction", "update + p publish",
          "residual replacement"]


def build_system(mf, name):
    if name == "poisson128":
        rp, ci, va = mf.workloads.poisson3d(128)
        b, _ = mf.workloads.rhs_from_solution(rp, ci, va, seed=1)
        return rp, ci, va, b, "lossless"
    if name == "varcoef64":
        rp, ci, va, _k = mf.workloads.varcoef_diffusion3d(64, seed=4)
        rng = np.random.default_rng(4)
        rng.uniform(-1, 1, rp.size - 1)  # the loss target, drawn first
        b = rng.uniform(-1, 1, rp.size - 1)
        return rp, ci, va, b, "rowscaled"
    if name == "poisson64":
        rp, ci, va = mf.workloads.poisson2d(64)
        b, _ = mf.workloads.rhs_from_solution(rp, ci, va, seed=0)
        return rp, ci, va, b, "lossless"
    raise ValueError(name)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="poisson128", choices=["poisson128", "varcoef64", "poisson64"])
    ap.add_argument("--precision", default="mixed", choices=["mixed", "fp64"])
    ap.add_argument("--grid", type=int, default=0)
    ap.add_argument("--no-value-cache", action="store_true")
    ap.add_argument("--json", default="")
    args = ap.parse_args()

    sys.path.insert(0, str(ROOT))
    import torch
    from __graft_entry__ import load_package

    mf = load_package()
    mf._lib.use_library("libmf_phase.so")
    rp, ci, va, b, policy = build_system(mf, args.workload)
    A = mf.TiledMatrix(rp, ci, va, class_policy=policy)
    bt = torch.as_tensor(b, device="cuda:0")
    kw = dict(tol=1e-10, precision=args.precision, variant=4, grid=args.grid, value_cache=not args.no_value_cache)
    for _ in range(3):
        x, info = A.cg(bt, **kw)
    G = info["grid"]
    buf = (ctypes.c_ulonglong * (8 * G))()
    st = mf._lib.lib().mf_phase_read(buf, G)
    if st != 0:
        raise mf._lib.MfError(st, "mf_phase_read")
    c = np.frombuffer(buf, dtype=np.uint64).reshape(G, 8).astype(np.float64)
    its = np.maximum(c[:, 6], 1.0)
    us = c[:, :6] / its[:, None] / 100.0  # 100 MHz ticks -> microseconds per iteration
    rec = c[:, 7] / its / 100.0
    per_iter = info["time_ms"] * 1e3 / max(1, info["iterations"])
    rows = []
    for k, name in enumerate(PHASES):
        rows.append({"phase": name, "mean_us": float(us[:, k].mean()), "max_us": float(us[:, k].max()),
                     "root_us": float(us[0, k])})
    total = us.sum(axis=1)
    out = {"workload": args.workload, "precision": args.precision, "grid": int(G), "layout": info.get("layout"),
           "variant": info["variant"], "iterations": info["iterations"], "replacements": info["replacements"],
           "kernel_ms": info["time_ms"], "kernel_us_per_iteration": per_iter, "phases": rows,
           "sum_of_phases_mean_us": float(total.mean()),
           "record_publish_us_per_iteration": {"mean": float(rec.mean()), "max": float(rec.max())}}
    print(f"{args.workload} {args.precision}: grid {G} layout {out['layout']} variant {info['variant']}, "
          f"{info['iterations']} iterations ({info['replacements']} replacements), kernel {info['time_ms']:.3f} ms "
          f"= {per_iter:.2f} us/iteration with the timers")
    print(f"{'phase':30s} {'mean us':>9s} {'max us':>9s} {'root us':>9s}")
    for r in rows:
        print(f"{r['phase']:30s} {r['mean_us']:9.3f} {r['max_us']:9.3f} {r['root_us']:9.3f}")
    print(f"{'sum':30s} {total.mean():9.3f} {total.max():9.3f} {total[0]:9.3f}")
    print(f"grid reduction entry -> own record published: mean {rec.mean():.3f} us, max {rec.max():.3f} us")
    if args.json:
        pathlib.Path(args.json).write_text(json.dumps(out, indent=1))
    A.close()


if __name__ == "__main__":
    main()
