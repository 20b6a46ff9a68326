"""The single-system CG kernels on the row-per-lane sliced-ELL streams (layout "sell") against
the oracle and against the tile-chunk kernels: every rows-per-thread shape, both precisions,
ragged and irregular row lengths (padding units, empty slices between non-empty ones), an
initial guess, the fixed-iteration trajectory, and the fallback to tile chunks for columns
outside the 16-bit offset range."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _poisson(mf, nx, ny=None, seed=0):
    rp, ci, va = mf.workloads.poisson2d(nx, ny)
    b, xs = mf.workloads.rhs_from_solution(rp, ci, va, seed)
    return rp, ci, va, b, xs


# grid -> rows per thread on 2D Poisson 64 x 64 (4096 rows) and 128 x 128 (16384 rows); the
# partition balances chunks, so 4 workgroups hold a few rows over 1024 each
@pytest.mark.parametrize("nx,grid,rpt", [(64, 8, 1), (64, 4, 2), (64, 1, 4), (128, 2, 8)])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_sell_every_shape_matches_oracle(mf, orc, torch, nx, grid, rpt, precision):
    rp, ci, va, b, xs = _poisson(mf, nx)
    A = mf.TiledMatrix(rp, ci, va)
    x, info = A.cg(b, tol=1e-10, precision=precision, grid=grid, layout="sell")
    ref = orc.cg(rp, ci, va, b, tol=1e-10)
    print(f"sell nx={nx} grid={grid} {precision}: it {info['iterations']} (oracle {ref['iterations']}) "
          f"rpt {info['rows_per_thread']} repl {info['replacements']}", flush=True)
    assert info["status"] == 0, info
    assert info["rows_per_thread"] == rpt
    assert abs(info["iterations"] - ref["iterations"]) <= max(1, int(0.02 * ref["iterations"]))
    assert orc.true_relres(rp, ci, va, x.cpu().numpy(), b) <= 1.2e-10
    A.close()


@pytest.mark.parametrize("variant", [2, 4])
@pytest.mark.parametrize("layout", ["sell", "chunks"])
def test_fixed_iterations_follow_the_oracle(mf, orc, torch, layout, variant):
    """Seven FP64 iterations on a ragged matrix (n = 1037) reproduce the textbook iterate."""
    rp, ci, va, b, xs = _poisson(mf, 17, 61, seed=3)
    A = mf.TiledMatrix(rp, ci, va)
    x, info = A.cg(b, max_iter=7, fixed_iterations=True, layout=layout, variant=variant, grid=5)
    ref = orc.cg(rp, ci, va, b, tol=0.0, maxit=7)
    assert info["status"] == 0 and info["iterations"] == 7
    assert np.allclose(x.cpu().numpy(), ref["x"], rtol=1e-10, atol=1e-13)
    A.close()


def test_sell_irregular_rows_and_classes(mf, orc, torch):
    """Lognormal row lengths in a band of 400: slices of very different lengths, and classes
    that are absent from some slices of a wave (empty slices between non-empty ones)."""
    rp, ci, va, b, xs = mf.workloads.irregular_spd(n=20000, band=400, logd_max=2.0)
    A = mf.TiledMatrix(rp, ci, va, class_policy="lossless")
    # two iterations: every product enters, while the ill-conditioned recurrence has not yet
    # amplified the summation-order rounding (a wrong column, value or row would show at
    # order one)
    ref = orc.cg(rp, ci, va, b, tol=0.0, maxit=2)
    x, info = A.cg(b, max_iter=2, fixed_iterations=True, precision="fp64", layout="sell")
    assert np.allclose(x.cpu().numpy(), ref["x"], rtol=1e-10, atol=1e-13 * np.abs(ref["x"]).max())
    # tile-mixed sliced-ELL and tile-mixed chunks compute the same products; only the
    # binary64 summation order differs
    xs_, _ = A.cg(b, max_iter=2, fixed_iterations=True, precision="mixed", layout="sell")
    xc_, _ = A.cg(b, max_iter=2, fixed_iterations=True, precision="mixed", layout="chunks")
    xs_, xc_ = xs_.cpu().numpy(), xc_.cpu().numpy()
    assert np.max(np.abs(xs_ - xc_)) <= 1e-10 * np.abs(xc_).max()
    x, info = A.cg(b, tol=1e-10, precision="mixed", layout="sell")
    assert info["status"] == 0 and info["layout"] in (2, 3, 6, 7), info
    assert orc.true_relres(rp, ci, va, x.cpu().numpy(), b) <= 1.2e-10
    A.close()


@pytest.mark.parametrize("variant", [2, 4])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_sell_rows_sorted_by_length(mf, orc, torch, precision, variant):
    """Four workgroups of 5,000 lognormal-length rows (8 rows per thread): each workgroup's
    rows are sorted by length into slices (layout 4/5), against the oracle's iterate after
    two iterations and to the FP64 tolerance; a variant without permuted kernels on the same
    partition falls back to the tile chunks."""
    rp, ci, va, b, xs = mf.workloads.irregular_spd(n=20000, band=400, logd_max=2.0)
    A = mf.TiledMatrix(rp, ci, va, class_policy="lossless")
    ref = orc.cg(rp, ci, va, b, tol=0.0, maxit=2)
    x, info = A.cg(b, max_iter=2, fixed_iterations=True, precision="fp64", layout="sell", grid=4,
                   variant=variant)
    assert info["layout"] in (4, 5) and info["rows_per_thread"] == 8, info
    assert np.allclose(x.cpu().numpy(), ref["x"], rtol=1e-10, atol=1e-13 * np.abs(ref["x"]).max())
    x, info = A.cg(b, tol=1e-10, precision=precision, layout="sell", grid=4, variant=variant)
    assert info["status"] == 0 and info["layout"] in (4, 5), info
    assert orc.true_relres(rp, ci, va, x.cpu().numpy(), b) <= 1.2e-10
    x, info = A.cg(b, tol=1e-10, precision=precision, layout="sell", grid=4, variant=1)
    assert info["status"] == 0 and info["layout"] == 1, info
    assert orc.true_relres(rp, ci, va, x.cpu().numpy(), b) <= 1.2e-10
    A.close()


def test_sell_initial_guess_and_zero_rhs(mf, orc, torch):
    rp, ci, va, b, xs = _poisson(mf, 31, 33, seed=2)
    A = mf.TiledMatrix(rp, ci, va)
    x0 = xs + 1e-3 * np.random.default_rng(0).standard_normal(xs.size)
    for precision in ("fp64", "mixed"):
        x, info = A.cg(b, x0=x0, tol=1e-10, precision=precision, layout="sell")
        assert info["status"] == 0
        assert orc.true_relres(rp, ci, va, x.cpu().numpy(), b) <= 1.2e-10
    x, info = A.cg(np.zeros_like(b), layout="sell")
    assert info["status"] == 0 and info["iterations"] == 0 and float(x.abs().max()) == 0.0
    A.close()


def test_sell_pipelined_variant(mf, orc, torch):
    rp, ci, va, b, xs = _poisson(mf, 64)
    A = mf.TiledMatrix(rp, ci, va)
    for precision in ("fp64", "mixed"):
        x, info = A.cg(b, tol=1e-10, precision=precision, variant=3, layout="sell")
        assert info["status"] == 0, info
        assert orc.true_relres(rp, ci, va, x.cpu().numpy(), b) <= 1.2e-10
    A.close()


def test_wide_offsets_fall_back_to_chunks(mf, orc, torch):
    """A coupling between rows 0 and n-1 (offset 39999) is outside the 16-bit offsets: the
    solve runs on the tile chunks and still matches the oracle."""
    n = 40000
    rows = [np.arange(n), np.arange(1, n), np.arange(n - 1), np.array([0, n - 1])]
    cols = [np.arange(n), np.arange(n - 1), np.arange(1, n), np.array([n - 1, 0])]
    vals = [np.full(n, 4.0), np.full(n - 1, -1.0), np.full(n - 1, -1.0), np.array([-1.0, -1.0])]
    import scipy.sparse as sp
    S = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, n))
    rp, ci, va = S.indptr.astype(np.int32), S.indices.astype(np.int32), S.data
    b, _ = mf.workloads.rhs_from_solution(rp, ci, va, 7)
    A = mf.TiledMatrix(rp, ci, va)
    ref = orc.cg(rp, ci, va, b, tol=1e-10)
    for layout in ("sell", "auto"):
        x, info = A.cg(b, tol=1e-10, precision="mixed", layout=layout)
        assert info["status"] == 0
        assert orc.true_relres(rp, ci, va, x.cpu().numpy(), b) <= 1.2e-10
        assert info["iterations"] <= 1.5 * ref["iterations"] + 5
    A.close()


def test_unit_profiles(mf, orc, torch):
    """Rows spread over two classes (convection-diffusion: binary16 and int8 tiles) take the
    narrow unit profile, single-class rows (Poisson: int8) the wide one; both reach the
    FP64 tolerance, and the layout and variant used are reported.  A small all-FP64 solve
    keeps its binary64 stream in LDS."""
    rp, ci, va = mf.workloads.convdiff3d(16)
    b, _ = mf.workloads.rhs_from_solution(rp, ci, va, 1)
    A = mf.TiledMatrix(rp, ci, va, class_policy="lossless")
    st = A.stats()["tiles_by_class"]
    assert st[2] > 0 and st[3] > 0
    x, info = A.bicgstab(b, tol=1e-10, precision="mixed")
    assert info["status"] == 0 and info["layout"] == 3, info
    assert orc.true_relres(rp, ci, va, x.cpu().numpy(), b) <= 1.2e-10
    x, info = A.bicgstab(b, tol=1e-10, precision="mixed", layout="chunks")
    assert info["status"] == 0 and info["layout"] == 1
    A.close()
    rp, ci, va = mf.workloads.poisson3d(16)
    b, _ = mf.workloads.rhs_from_solution(rp, ci, va, 1)
    A = mf.TiledMatrix(rp, ci, va, class_policy="lossless")
    x, info = A.cg(b, tol=1e-10, precision="mixed")
    assert info["status"] == 0 and info["layout"] in (2, 6) and info["variant"] == 4, info
    # an all-FP64 stream this small stays in LDS: sliced-ELL with on-chip values; without
    # the value cache the all-FP64 solve takes the tile chunks
    x, info = A.cg(b, tol=1e-10, precision="fp64")
    assert info["status"] == 0 and info["layout"] == 6 and info["variant"] == 4, info
    assert orc.true_relres(rp, ci, va, x.cpu().numpy(), b) <= 1.2e-10
    x, info = A.cg(b, tol=1e-10, precision="fp64", value_cache=False)
    assert info["status"] == 0 and info["layout"] == 1 and info["variant"] == 4, info
    A.close()


def test_layout_choice_by_bytes(mf, orc, torch):
    """Few entries per tile: the all-FP64 CG takes the sliced-ELL streams (tile-chunk headers
    would cost more per entry).  A tile-mixed request whose class streams would move as many
    bytes as the binary64 stream runs binary64 arithmetic on it, reports so, and matches the
    FP64 solve's iteration count."""
    rp, ci, va, b, xs = mf.workloads.irregular_spd(n=20000, band=8000, logd_max=2.0)
    A = mf.TiledMatrix(rp, ci, va, class_policy="lossless")
    st = A.stats()
    assert ci.size < 8 * st["tiles"]
    x, info = A.cg(b, tol=1e-10, precision="fp64")
    assert info["status"] == 0 and info["layout"] in (2, 4, 6), info
    it64 = info["iterations"]
    x, info = A.cg(b, tol=1e-10, precision="mixed")
    assert info["status"] == 0
    assert orc.true_relres(rp, ci, va, x.cpu().numpy(), b) <= 1.2e-10
    if info["precision"] == 0:
        assert info["iterations"] == it64
    # forcing the sliced-ELL layout keeps the tile classes
    x, info = A.cg(b, tol=1e-10, precision="mixed", layout="sell")
    assert info["status"] == 0 and info["precision"] == 1
    A.close()


def _poisson3d_diag(mf, nx, diag):
    """3D 7-point Poisson with the diagonal set to `diag` (SPD for diag >= 6)."""
    rp, ci, va = mf.workloads.poisson3d(nx)
    va = va.copy()
    rows = np.repeat(np.arange(rp.size - 1), np.diff(rp))
    va[rows == ci] = diag
    return rp, ci, va


@pytest.mark.parametrize("case", ["int8_wide", "fp16_int8", "fp64_sell"])
def test_value_streams_in_lds_match_memory_streams(mf, orc, torch, case):
    """On-chip value streams (the reference's LDS "shared queue", Mille-feuille-cg.h:586,
    main-cg-mixed.cu:484): the kernel copies each workgroup's value words into LDS at launch
    and every SpMV reads them there.  The words and the summation order are those of the
    memory-read kernel, so the results are bit-identical to it, and the iteration count is
    the oracle's."""
    if case == "int8_wide":
        rp, ci, va = _poisson3d_diag(mf, 32, 6.0)
        prec, layout = "mixed", "auto"
    elif case == "fp16_int8":
        rp, ci, va = _poisson3d_diag(mf, 28, 6.5)   # diagonal tiles binary16, the rest int8
        prec, layout = "mixed", "auto"
    else:
        rp, ci, va = mf.workloads.poisson2d(64)
        prec, layout = "fp64", "sell"
    b, _ = mf.workloads.rhs_from_solution(rp, ci, va, 5)
    A = mf.TiledMatrix(rp, ci, va, class_policy="lossless")
    x1, i1 = A.cg(b, tol=1e-10, precision=prec, layout=layout)
    x0, i0 = A.cg(b, tol=1e-10, precision=prec, layout=layout, value_cache=False)
    print(f"{case}: on-chip layout {i1['layout']} {i1['time_ms']:.3f} ms; memory layout {i0['layout']} "
          f"{i0['time_ms']:.3f} ms; iterations {i1['iterations']}", flush=True)
    assert i1["layout"] == i0["layout"] + 4 and i1["layout"] in (6, 7)
    assert i1["iterations"] == i0["iterations"] and i1["replacements"] == i0["replacements"]
    assert np.array_equal(x1.cpu().numpy(), x0.cpu().numpy())
    ref = orc.cg(rp, ci, va, b, tol=1e-10)
    assert abs(i1["iterations"] - ref["iterations"]) <= max(1, 0.02 * ref["iterations"])
    assert orc.true_relres(rp, ci, va, x1.cpu().numpy(), b) <= 1.2e-10
    A.close()
