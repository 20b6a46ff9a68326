"""Product tile converter, tables, packing, classes and reader vs the golden fixtures
and the oracle (CPU only: no device is touched)."""
import numpy as np
import pytest

from conftest import GOLDEN, TILE_CASES, load_tile_case

FIELDS = ["tile_ptr", "tile_columnidx", "tile_nnz", "Format", "blknnz", "blknnznnz", "tilewidth",
          "csr_offset", "csrptr_offset", "Blockcsr_Val", "Blockcsr_Val_Low", "Tile_csr_Col",
          "Blockcsr_Ptr", "csr_compressedIdx"]


def _bits(a):
    a = np.asarray(a)
    return a.view(np.uint8) if a.dtype.kind == "f" else a


@pytest.mark.parametrize("case", TILE_CASES)
def test_tile_create_bit_exact(mf, case):
    g = load_tile_case(case)
    d = mf.tile_create(int(g["rowA"]), int(g["colA"]), g["rowptr"], g["colidx"], g["val"], g["val_low"])
    tilem, tilen, T, csrsize, csrptrlen, _ = (int(v) for v in g["scalars"])
    assert (d["tilem"], d["tilen"], d["tilenum"], d["csrsize"], d["csrptrlen"]) == (tilem, tilen, T, csrsize,
                                                                                     csrptrlen)
    for f in FIELDS:
        assert d[f].shape == g[f].shape, f
        assert np.array_equal(_bits(d[f]), _bits(g[f])), f


@pytest.mark.parametrize("case", TILE_CASES)
def test_prep_tables_bit_exact(mf, case):
    g = load_tile_case(case)
    t = mf.Tiles(int(g["rowA"]), int(g["colA"]), g["rowptr"], g["colidx"], g["val"], g["val_low"])
    p = t.prep_tables()
    assert p["rowblkblock"] == int(g["scalars"][5])
    for f in ("ptroffset1", "ptroffset2", "rowblk_idx", "rowblk_colstart", "rowblk_colstop"):
        assert np.array_equal(p[f], g[f]), f
    t.close()


@pytest.mark.parametrize("case", TILE_CASES)
def test_row_lists_match_oracle(mf, orc, case):
    g = load_tile_case(case)
    rowA, colA = int(g["rowA"]), int(g["colA"])
    t = mf.Tiles(rowA, colA, g["rowptr"], g["colidx"], g["val"], g["val_low"])
    mine = t.row_lists()
    ref = orc.row_lists(orc.tile_create(rowA, colA, g["rowptr"], g["colidx"], g["val"], g["val_low"]))
    for f in ("nonzero_row_new", "blockrowid_new", "blockcsr_ptr_new"):
        assert np.array_equal(mine[f], ref[f]), f
    t.close()


def test_encode_decode_bit_exact(mf):
    with np.load(GOLDEN / "encode.npz", allow_pickle=False) as z:
        for L in range(41):
            assert np.array_equal(mf.encode(z[f"in_{L}"]), z[f"enc_{L}"]), L
            assert np.array_equal(mf.decode(z[f"enc_{L}"], L), z[f"dec_{L}"]), L


def test_value_copies_closed_forms(mf):
    rng = np.random.default_rng(3)
    edge = np.array([0.0, -0.0, 1.0, -1.0, 0.5, 127.9, -128.9, 126.5, 65504.0, 65519.99, 65520.0, 1e6, -1e6,
                     6.103515625e-05, 5.960464477539063e-08, 2.9802322387695312e-08, 1e-30, 3.4e38, 1.0 / 3.0,
                     2.0 ** -24 * 1.5, np.nextafter(1.0, 2.0), 1.0 + 2.0 ** -11, 1.0 + 3 * 2.0 ** -11])
    v = np.concatenate([edge, rng.standard_normal(4000) * 10.0 ** rng.uniform(-9, 6, 4000)])
    f32, f16, i8 = mf.value_copies(v)
    assert np.array_equal(f32.view(np.uint32), v.astype(np.float32).view(np.uint32))
    with np.errstate(over="ignore"):
        assert np.array_equal(f16, v.astype(np.float32).astype(np.float16).view(np.uint16))
    assert np.array_equal(i8, np.clip(np.trunc(v), -128, 127).astype(np.int8))


@pytest.mark.parametrize("case", TILE_CASES)
def test_static_classes_match_oracle(mf, orc, case):
    g = load_tile_case(case)
    rowA, colA = int(g["rowA"]), int(g["colA"])
    mine = mf.tile_classes(rowA, colA, g["rowptr"], g["colidx"], g["val"], "lossless")
    # the oracle classifies on the exported value stream, where a single-row last
    # tile row is zeroed by the reference; compare on the full tile structure instead
    d = orc.tile_create(rowA, colA, g["rowptr"], g["colidx"], g["val"], g["val_low"])
    if rowA - (d["tilem"] - 1) * 16 == 1:
        pytest.skip("reference drops the values of a single-row last tile row")
    ref = orc.tile_classes(d)
    assert np.array_equal(mine, 3 - (3 - ref))
    per = np.array([orc.static_class(d["Blockcsr_Val"][d["tile_nnz"][t]:d["tile_nnz"][t + 1]])
                    for t in range(d["tilenum"])])
    assert np.array_equal(mine, per)


def test_static_class_closed_forms(orc):
    cases = [([-1.0, 4.0, 127.0, -128.0], 3), ([0.5, 1.0], 2), ([65504.0], 2), ([65505.0], 1),
             ([2.0 ** -24], 2), ([2.0 ** -25], 1), ([1.0 + 2.0 ** -10], 2), ([1.0 + 2.0 ** -11], 1),
             ([0.1], 0), ([np.float32(0.1)], 1), ([1e40], 0), ([128.0], 2), ([2049.0], 1), ([np.nan], 0)]
    for vals, want in cases:
        assert orc.static_class(np.array(vals, np.float64)) == want, vals
    rng = np.random.default_rng(9)
    v = np.concatenate([rng.integers(-200, 200, 500).astype(float), rng.standard_normal(500),
                        rng.standard_normal(500).astype(np.float32).astype(float),
                        rng.standard_normal(500).astype(np.float16).astype(float)])
    for x in v:
        assert orc.static_class(np.array([x])) == orc.static_class_np(np.array([x]))[0]


def test_dyn_class_thresholds(orc):
    table = [(2e-5, 1), (1e-5, 4), (5e-6, 4), (1e-6, 3), (5e-7, 3), (1e-7, 2), (5e-8, 2), (1e-8, 0), (0.0, 0)]
    for d, code in table:
        assert orc.dyn_code(d) == code, d
    z = np.zeros(16)
    assert orc.dyn_class(z, z) == 1
    assert orc.dyn_class(z, np.full(16, 5e-8)) == 2
    p = np.full(16, 5e-8)
    p[3] = 3e-6
    assert orc.dyn_class(z, p) == 4
    p[9] = 1.0
    assert orc.dyn_class(z, p) == 0


def test_read_matrix_matches_reference_reader(mf, tmp_path):
    with np.load(GOLDEN / "mtx_add20.npz", allow_pickle=False) as z:
        ref = {k: z[k] for k in z.files}
    # the add20 input itself is not committed; rebuild it from the fixture CSR in file order
    m, n, nnz = (int(v) for v in ref["scalars"])
    rows = np.repeat(np.arange(m), np.diff(ref["rowptr"]))
    p = tmp_path / "a.mtx"
    with open(p, "w") as f:
        f.write(f"%%MatrixMarket matrix coordinate real general\n{m} {n} {nnz}\n")
        for r, c, v in zip(rows, ref["colidx"], ref["val"]):
            f.write(f"{r + 1} {c + 1} {float(v)!r}\n")
    mm, nn, rp, ci, va, sym = mf.read_matrix(p)
    assert (mm, nn, sym) == (m, n, False)
    assert np.array_equal(rp, ref["rowptr"]) and np.array_equal(ci, ref["colidx"])
    assert np.array_equal(va, ref["val"])
    with np.load(GOLDEN / "mtx_sym.npz", allow_pickle=False) as z:
        p = tmp_path / "s.mtx"
        p.write_bytes(z["text"].tobytes())
        mm, nn, rp, ci, va, sym = mf.read_matrix(p)
        assert sym and np.array_equal(rp, z["rowptr"]) and np.array_equal(ci, z["colidx"])
        assert np.array_equal(va, z["val"])
    with np.load(GOLDEN / "mtx_cbd.npz", allow_pickle=False) as z:
        p = tmp_path / "c.cbd"
        p.write_bytes(z["blob"].tobytes())
        mm, nn, rp, ci, va, sym = mf.read_matrix(p)
        assert np.array_equal(rp, z["rowptr"]) and np.array_equal(ci, z["colidx"])
        assert np.array_equal(va, z["val"])
        q = tmp_path / "d.cbd"
        mf.write_cbd(q, mm, nn, rp, ci, va)
        assert q.read_bytes() == p.read_bytes()


def test_library_exports_every_declared_symbol(mf):
    import re
    hdr = (GOLDEN.parents[1] / "include" / "mf.h").read_text()
    declared = set(re.findall(r"\b(mf_[a-z0-9_]+)\s*\(", hdr))
    lib = mf.lib()
    missing = sorted(s for s in declared if getattr(lib, s, None) is None)
    assert not missing, missing


def test_phase_timers_only_in_the_phase_library(mf):
    """libmf.so carries the per-phase timer entry point but no timers: it refuses (no GPU call
    is made); libmf_phase.so, built from the same sources, exports the same symbols."""
    import ctypes
    buf = (ctypes.c_ulonglong * 8)()
    assert mf.lib().mf_phase_read(buf, 1) == 10  # MF_ERR_UNSUPPORTED
    phase = ctypes.CDLL(str(GOLDEN.parents[1] / "mille-feuille_amd" / "libmf_phase.so"))
    for name in ("mf_phase_read", "mf_cg_solve", "mf_matrix_create", "mf_batch_timeline"):
        assert getattr(phase, name, None) is not None, name
