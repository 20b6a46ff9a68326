"""Write the golden fixtures under tests/golden/ from the reference's own host code.

Runs in the development container only (it needs the reference tree).  It builds
oracle/_ref/ref_harness from the reference's csr2block.h, blockspmv_cpu.h,
encode.h and biio.h where they lie (oracle/Makefile ``ref`` target), feeds it the
inputs below and stores inputs and outputs as .npz (no pickles):

  tile_<case>.npz   Tile_create arrays + blockspmv_cpu tables for one input
  encode.npz        encode()/decode() of nibble strings of length 0..40
  mtx_add20.npz     read_Dmatrix_32 on the reference's test/add20.mtx
  mtx_sym.npz       read_Dmatrix_32 on a small symmetric .mtx written here
  mtx_cbd.npz       read_Dmatrix_32 on a small .cbd written here

Usage: python tests/golden/make_golden.py [--ref-src /root/reference/src]
"""
from __future__ import annotations

import argparse
import importlib.util
import pathlib
import subprocess
import sys
import tempfile

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(ROOT / "oracle"))
import oracle  # noqa: E402

_spec = importlib.util.spec_from_file_location("mf_workloads", ROOT / "mille-feuille_amd" / "workloads.py")
wl = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(wl)


def write_input(path, rowA, colA, rowptr, colidx, val, val_low):
    with open(path, "wb") as f:
        np.array([rowA, colA, colidx.size], np.int32).tofile(f)
        np.asarray(rowptr[: rowA + 1], np.int32).tofile(f)
        np.asarray(colidx, np.int32).tofile(f)
        np.asarray(val, np.float64).tofile(f)
        np.asarray(val_low, np.float32).tofile(f)


def run_tile(harness, name, rowA, colA, rowptr, colidx, val, tmp):
    rowptr = np.asarray(rowptr, np.int32)
    colidx = np.asarray(colidx, np.int32)
    val = np.asarray(val, np.float64)
    nnz_used = int(rowptr[rowA])
    colidx = colidx[:nnz_used]
    val = val[:nnz_used]
    vlow = val.astype(np.float32)
    fin = tmp / f"{name}.in"
    fout = tmp / f"{name}.out"
    write_input(fin, rowA, colA, rowptr, colidx, val, vlow)
    subprocess.run([str(harness), "tile", str(fin), str(fout)], check=True)
    rec = oracle.read_records(fout)
    np.savez_compressed(HERE / f"tile_{name}.npz", rowA=np.int32(rowA), colA=np.int32(colA),
                        rowptr=rowptr[: rowA + 1], colidx=colidx, val=val, val_low=vlow, **rec)
    print(f"tile_{name}: rowA={rowA} colA={colA} nnz={colidx.size} tiles={rec['scalars'][2]}")


def read_mtx_raw(path):
    """Plain coordinate reader for the fixture inputs (no mirroring)."""
    lines = [ln for ln in pathlib.Path(path).read_text().splitlines()]
    body = [ln for ln in lines[1:] if not ln.startswith("%")]
    m, n, nnz = (int(t) for t in body[0].split())
    ent = [ln.split() for ln in body[1: 1 + nnz]]
    r = np.array([int(e[0]) - 1 for e in ent])
    c = np.array([int(e[1]) - 1 for e in ent])
    v = np.array([float(e[2]) for e in ent])
    return m, n, r, c, v


def csr_file_order(m, r, c, v):
    """CSR keeping file order inside each row (how biio places general entries)."""
    order = np.argsort(r, kind="stable")
    rowptr = np.zeros(m + 1, np.int64)
    np.add.at(rowptr, r + 1, 1)
    return np.cumsum(rowptr).astype(np.int32), c[order].astype(np.int32), v[order]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--ref-src", default="/root/reference/src")
    args = ap.parse_args()
    harness = oracle.build_ref(args.ref_src)
    if harness is None:
        sys.exit("reference tree not present; fixtures are committed and need no rebuild")
    tmp = pathlib.Path(tempfile.mkdtemp())
    rng = np.random.default_rng(12345)

    # 1. add20.mtx exactly as main-cg.cu:1691-1705 feeds it: rows truncated to 2384.
    add20 = pathlib.Path(args.ref_src).parent / "test" / "add20.mtx"
    subprocess.run([str(harness), "mtx", str(add20), str(tmp / "add20.rec")], check=True)
    rec = oracle.read_records(tmp / "add20.rec")
    np.savez_compressed(HERE / "mtx_add20.npz", **rec)
    m, n, nnz = (int(v) for v in rec["scalars"])
    rowA = (m // 16) * 16
    run_tile(harness, "add20", rowA, n, rec["rowptr"], rec["colidx"], rec["val"], tmp)

    # 2. config-1 Poisson 64x64
    rp, ci, va = wl.poisson2d(64)
    run_tile(harness, "poisson64", 4096, 4096, rp, ci, va, tmp)

    # 3. ragged random matrices: rowA % 16 == 8 and == 1 (single-row last tile row)
    for nm, nrow in (("ragged1000", 1000), ("ragged1009", 1009)):
        dens = rng.random((nrow, nrow)) < 0.01
        np.fill_diagonal(dens, True)
        r, c = np.nonzero(dens)
        v = rng.standard_normal(r.size)
        rp = np.zeros(nrow + 1, np.int64)
        np.add.at(rp, r + 1, 1)
        run_tile(harness, nm, nrow, nrow, np.cumsum(rp), c, v, tmp)

    # 4. one fully dense 16x16 tile (256 nnz: u8 counters wrap) + a sparse rest
    nrow = 48
    dens = np.zeros((nrow, nrow), bool)
    dens[16:32, 16:32] = True
    dens[np.arange(nrow), np.arange(nrow)] = True
    dens[0, 47] = dens[47, 0] = True
    r, c = np.nonzero(dens)
    v = np.round(rng.standard_normal(r.size) * 8) / 4
    rp = np.zeros(nrow + 1, np.int64)
    np.add.at(rp, r + 1, 1)
    run_tile(harness, "densetile", nrow, nrow, np.cumsum(rp), c, v, tmp)

    # 5. empty rows and an empty tile row
    nrow = 64
    dens = rng.random((nrow, nrow)) < 0.05
    dens[16:32, :] = False
    dens[40, :] = False
    r, c = np.nonzero(dens)
    v = rng.standard_normal(r.size)
    rp = np.zeros(nrow + 1, np.int64)
    np.add.at(rp, r + 1, 1)
    run_tile(harness, "emptyrows", nrow, nrow, np.cumsum(rp), c, v, tmp)

    # 6. unsorted columns inside rows, non-square (rowA 80, colA 100)
    nrow, ncol = 80, 100
    dens = rng.random((nrow, ncol)) < 0.08
    r, c = np.nonzero(dens)
    v = rng.standard_normal(r.size)
    perm = np.lexsort((rng.random(r.size), r))
    r, c, v = r[perm], c[perm], v[perm]
    rp = np.zeros(nrow + 1, np.int64)
    np.add.at(rp, r + 1, 1)
    run_tile(harness, "unsorted", nrow, ncol, np.cumsum(rp), c, v, tmp)

    # 7. encode / decode of nibble strings
    enc = {}
    for L in range(0, 41):
        s = rng.integers(0, 16, L).astype(np.uint8)
        fin = tmp / f"enc{L}.in"
        with open(fin, "wb") as f:
            np.array([L], np.int32).tofile(f)
            s.tofile(f)
        subprocess.run([str(harness), "encode", str(fin), str(tmp / f"enc{L}.out")], check=True)
        rec = oracle.read_records(tmp / f"enc{L}.out")
        enc[f"in_{L}"] = s
        enc[f"enc_{L}"] = rec["encoded"]
        enc[f"dec_{L}"] = rec["decoded"]
    np.savez_compressed(HERE / "encode.npz", **enc)

    # 8. matrix reader semantics: symmetric real with a pattern-free body, and .cbd
    sym = tmp / "sym.mtx"
    ent = [(1, 1, 4.0), (3, 1, -1.5), (2, 2, 5.0), (4, 2, 2.25), (3, 3, 6.0), (4, 4, 7.0), (4, 3, -0.5)]
    with open(sym, "w") as f:
        f.write("%%MatrixMarket matrix coordinate real symmetric\n% fixture\n4 4 7\n")
        for a, b, w in ent:
            f.write(f"{a} {b} {w}\n")
    subprocess.run([str(harness), "mtx", str(sym), str(tmp / "sym.rec")], check=True)
    rec = oracle.read_records(tmp / "sym.rec")
    np.savez_compressed(HERE / "mtx_sym.npz", text=np.frombuffer(sym.read_bytes(), np.uint8), **rec)
    cbd = tmp / "small.cbd"
    rp, ci, va = wl.poisson2d(5)
    with open(cbd, "wb") as f:
        np.array([25, 25], np.int32).tofile(f)
        np.array([ci.size], np.int64).tofile(f)
        rp.astype(np.int64).tofile(f)
        ci.astype(np.int32).tofile(f)
        va.astype(np.float64).tofile(f)
    subprocess.run([str(harness), "mtx", str(cbd), str(tmp / "cbd.rec")], check=True)
    rec = oracle.read_records(tmp / "cbd.rec")
    np.savez_compressed(HERE / "mtx_cbd.npz", blob=np.frombuffer(cbd.read_bytes(), np.uint8), **rec)
    print("done")


if __name__ == "__main__":
    main()
