"""The oracle is pinned against the reference's own outputs (golden fixtures)."""
import numpy as np
import pytest

from conftest import GOLDEN, TILE_CASES, load_tile_case

FIELDS = ["tile_ptr", "tile_columnidx", "tile_nnz", "Format", "blknnz", "blknnznnz", "tilewidth",
          "csr_offset", "csrptr_offset", "Blockcsr_Val", "Blockcsr_Val_Low", "Tile_csr_Col",
          "Blockcsr_Ptr", "csr_compressedIdx"]


@pytest.mark.parametrize("case", TILE_CASES)
def test_oracle_tile_create_matches_reference(orc, case):
    g = load_tile_case(case)
    d = orc.tile_create(int(g["rowA"]), int(g["colA"]), g["rowptr"], g["colidx"], g["val"], g["val_low"])
    tilem, tilen, T, csrsize, csrptrlen, _ = (int(v) for v in g["scalars"])
    assert (d["tilem"], d["tilen"], d["tilenum"], d["csrsize"], d["csrptrlen"]) == \
        (tilem, tilen, T, csrsize, csrptrlen)
    for f in FIELDS:
        a = np.asarray(d[f])
        b = np.asarray(g[f])
        assert a.shape == b.shape, f
        assert np.array_equal(a.view(np.uint8) if a.dtype.kind == "f" else a,
                              b.view(np.uint8) if b.dtype.kind == "f" else b), f


@pytest.mark.parametrize("case", TILE_CASES)
def test_oracle_prep_tables_match_reference(orc, case):
    g = load_tile_case(case)
    d = orc.tile_create(int(g["rowA"]), int(g["colA"]), g["rowptr"], g["colidx"], g["val"], g["val_low"])
    p = orc.prep_tables(d)
    assert p["rowblkblock"] == int(g["scalars"][5])
    for f in ("ptroffset1", "ptroffset2", "rowblk_idx", "rowblk_colstart", "rowblk_colstop"):
        assert np.array_equal(p[f], g[f]), f


def test_oracle_encode_matches_reference(orc):
    with np.load(GOLDEN / "encode.npz", allow_pickle=False) as z:
        for L in range(41):
            s = z[f"in_{L}"]
            assert np.array_equal(orc.encode(s), z[f"enc_{L}"]), L
            assert np.array_equal(orc.decode(z[f"enc_{L}"], L), z[f"dec_{L}"]), L


@pytest.mark.parametrize("case", TILE_CASES)
def test_oracle_row_lists_rebuild_dense_blocks(orc, case):
    """P2 lists + Tile_csr_Col + values rebuild each tile's dense block from the CSR input."""
    g = load_tile_case(case)
    rowA, colA = int(g["rowA"]), int(g["colA"])
    d = orc.tile_create(rowA, colA, g["rowptr"], g["colidx"], g["val"], g["val_low"])
    rl = orc.row_lists(d)
    nzr, rid, rend = rl["nonzero_row_new"], rl["blockrowid_new"], rl["blockcsr_ptr_new"]
    T = d["tilenum"]
    tilem = d["tilem"]
    dense = np.zeros((tilem * 16, d["tilen"] * 16))
    rp, ci, va = g["rowptr"], g["colidx"], g["val"]
    for i in range(rowA):
        for j in range(rp[i], rp[i + 1]):
            dense[i, ci[j]] += va[j]
    rebuilt = np.zeros_like(dense)
    tile_row = np.repeat(np.arange(tilem), np.diff(d["tile_ptr"]))
    for t in range(T):
        s1, s2 = nzr[t], nzr[t + 1]
        assert rid[s1] == 0 and rend[s1] == 0  # sentinel
        off = d["tile_nnz"][t]
        nz = d["tile_nnz"][t + 1] - off
        rows_seen = rid[s1 + 1:s2]
        assert np.all(np.diff(rows_seen.astype(int)) > 0)
        for k in range(s2 - s1 - 1):
            ro = int(rid[s1 + k + 1])
            a, b = int(rend[s1 + k]), int(rend[s1 + k + 1])
            if k == s2 - s1 - 2 and nz == 256:
                b = 256  # u8 end pointer of a full tile wraps to 0
            for e in range(a, b):
                rebuilt[tile_row[t] * 16 + ro, d["tile_columnidx"][t] * 16 + d["Tile_csr_Col"][off + e]] += \
                    d["Blockcsr_Val"][off + e]
    last_rowlen = rowA - (tilem - 1) * 16
    if last_rowlen == 1:
        # the reference's unscanned start byte drops the single row of the last tile row
        dense[(tilem - 1) * 16, :] = 0.0
    assert np.array_equal(rebuilt[:rowA], dense[:rowA])


@pytest.mark.parametrize("case", TILE_CASES)
def test_oracle_tile_spmv_vs_csr(orc, case):
    g = load_tile_case(case)
    rowA, colA = int(g["rowA"]), int(g["colA"])
    d = orc.tile_create(rowA, colA, g["rowptr"], g["colidx"], g["val"], g["val_low"])
    rng = np.random.default_rng(7)
    x = rng.standard_normal(colA)
    y = orc.tile_spmv(d, x)
    import scipy.sparse as sp
    A = sp.csr_matrix((g["val"], g["colidx"], g["rowptr"]), shape=(rowA, colA))
    ref = A @ x
    if rowA - (d["tilem"] - 1) * 16 == 1:
        ref[rowA - 1] = 0.0  # the reference leaves a single-row last tile row empty
    assert np.allclose(y, ref, rtol=1e-13, atol=1e-13 * np.abs(ref).max())
