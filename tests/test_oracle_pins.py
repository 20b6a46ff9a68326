"""Pins for the oracle's solver and emulation functions (CPU only).

Each check is chosen so that a plausible mistake in the restatement (a dropped
term, a wrong sign, a swapped operand, a wrong class mapping) fails it:

* ``oracle.cg`` (restating Mille-feuille-cg.h:1188-1346) against SciPy's
  independent ``scipy.sparse.linalg.cg``: iteration count within one, the
  residual history against SciPy's iterates step by step, the solution against
  ``spsolve``, and finite termination on matrices with k distinct eigenvalues.
* ``oracle.bicgstab`` (restating Mille-feuille-bicg.h:562-822, initial state
  main-bicg.cu:534-549) against ``scipy.sparse.linalg.bicgstab`` the same way,
  on nonsymmetric systems.
* ``oracle.mixed_tile_spmv`` (the tile-mixed product of main-cg-mixed.cu:1130-1214
  with binary32 products) against exact integer arithmetic, against SciPy on
  operands pre-rounded to the tile's precision (the product of two binary32
  numbers is exact in binary64, so only the product rounding remains, bounded
  by 2^-24 per term), and against ``tile_spmv`` under the FP64 policy.
* ``oracle.dyn_classes_rel`` at its thresholds 2^-13 and 2^-24.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from conftest import load_package


def _poisson2d(nx):
    return load_package().workloads.poisson2d(nx)


def _convdiff(nx):
    return load_package().workloads.convdiff3d(nx)


def _csr(rp, ci, va, n=None):
    n = rp.size - 1 if n is None else n
    return sp.csr_matrix((va, ci, rp), shape=(n, n))


def _random_spd(n, seed, density=0.02, shift=0.5):
    rng = np.random.default_rng(seed)
    B = sp.random(n, n, density=density, random_state=rng, format="csr")
    A = (B @ B.T + shift * sp.eye(n)).tocsr()
    A.sort_indices()
    return A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data.astype(np.float64)


def _random_nonsym(n, seed, density=0.03):
    """Diagonally dominant nonsymmetric matrix with a random pattern."""
    rng = np.random.default_rng(seed)
    B = sp.random(n, n, density=density, random_state=rng, format="csr")
    B.data = rng.uniform(-1.0, 1.0, B.data.size)
    rs = np.asarray(abs(B).sum(axis=1)).ravel()
    A = (B + sp.diags(rs + 1.0 + rng.uniform(0, 1, n))).tocsr()
    A.sort_indices()
    return A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data.astype(np.float64)


def _scipy_history(solver, A, b, tol, maxiter):
    """Relative true residuals of SciPy's iterates (x0 = 0), one per completed iteration."""
    nb = np.linalg.norm(b)
    hist = [1.0]
    xs = []

    def cb(xk):
        xs.append(xk.copy())
        hist.append(np.linalg.norm(b - A @ xk) / nb)

    x, info = solver(A, b, rtol=tol, atol=0.0, maxiter=maxiter, callback=cb)
    return x, info, np.array(hist), xs


SPD_CASES = {
    "poisson2d_32": lambda: _poisson2d(32),
    "poisson2d_45": lambda: _poisson2d(45),
    "random_spd_600": lambda: _random_spd(600, seed=11, shift=2.0),
}


@pytest.mark.parametrize("case", sorted(SPD_CASES))
def test_oracle_cg_matches_scipy(orc, case):
    rp, ci, va = SPD_CASES[case]()
    A = _csr(rp, ci, va)
    n = A.shape[0]
    b = A @ np.random.default_rng(5).uniform(-1, 1, n)
    tol = 1e-10
    ref = orc.cg(rp, ci, va, b, tol=tol, want_hist=True)
    x_sp, info, hist_sp, _ = _scipy_history(spla.cg, A, b, tol, 10 * n)
    assert info == 0 and ref["status"] == 0
    it_sp = len(hist_sp) - 1
    # same recurrence, different summation order of the norms: within one iteration
    assert abs(ref["iterations"] - it_sp) <= 1, (ref["iterations"], it_sp)
    # step by step: the oracle's recursive residuals follow SciPy's true residuals over
    # the first ten iterations (later, rounding alone moves individual residual norms of
    # the non-monotone CG history, so only the count is compared there)
    k = min(10, len(hist_sp), len(ref["hist"]))
    assert np.allclose(ref["hist"][:k], hist_sp[:k], rtol=1e-8, atol=0)
    # solution against a direct solve, within the conditioning allowed by 1e-10
    x_d = spla.spsolve(A.tocsc(), b)
    ev = np.linalg.eigvalsh(A.toarray()) if n <= 2500 else None
    kappa = ev[-1] / ev[0]
    err = np.linalg.norm(ref["x"] - x_d) / np.linalg.norm(x_d)
    assert err <= 2 * kappa * tol, (err, kappa)
    # the reported residual is the recursive one; the true one agrees at convergence
    assert orc.true_relres(rp, ci, va, ref["x"], b) <= 1.5 * tol


@pytest.mark.parametrize("k", [1, 2, 5, 9])
def test_oracle_cg_finite_termination(orc, k):
    """CG on a diagonal SPD matrix with k distinct eigenvalues stops in at most k iterations."""
    n = 200
    rng = np.random.default_rng(k)
    eig = np.linspace(1.0, 10.0, k) if k > 1 else np.array([3.0])
    d = eig[rng.integers(0, k, n)]
    d[:k] = eig  # every eigenvalue present
    rp = np.arange(n + 1, dtype=np.int32)
    ci = np.arange(n, dtype=np.int32)
    b = rng.uniform(-1, 1, n)
    ref = orc.cg(rp, ci, d, b, tol=1e-10)
    assert ref["status"] == 0
    assert ref["iterations"] <= k, ref["iterations"]
    assert np.allclose(ref["x"], b / d, rtol=1e-9, atol=0)
    # one eigenvalue: x = b / lambda after exactly one step (alpha = 1/lambda)
    if k == 1:
        assert ref["iterations"] == 1


def test_oracle_cg_zero_rhs_and_budget(orc):
    rp, ci, va = _poisson2d(16)
    n = rp.size - 1
    z = orc.cg(rp, ci, va, np.zeros(n), tol=1e-10)
    assert z["status"] == 0 and z["iterations"] == 0 and not z["x"].any()
    b = np.ones(n)
    capped = orc.cg(rp, ci, va, b, tol=1e-10, maxit=7)
    assert capped["status"] == 1 and capped["iterations"] == 7
    # seven steps of CG minimise the A-norm error over the 7-dimensional Krylov space:
    # the same iterate as SciPy after seven callbacks
    _, _, _, xs = _scipy_history(spla.cg, _csr(rp, ci, va), b, 1e-30, 7)
    assert np.allclose(capped["x"], xs[6], rtol=1e-10, atol=1e-12)


NONSYM_CASES = {
    "convdiff_12": lambda: _convdiff(12),
    "convdiff_16": lambda: _convdiff(16),
    "random_nonsym_500": lambda: _random_nonsym(500, seed=21),
}


@pytest.mark.parametrize("case", sorted(NONSYM_CASES))
def test_oracle_bicgstab_matches_scipy(orc, case):
    rp, ci, va = NONSYM_CASES[case]()
    A = _csr(rp, ci, va)
    n = A.shape[0]
    assert abs(A - A.T).max() > 0  # nonsymmetric
    b = A @ np.random.default_rng(6).uniform(-1, 1, n)
    tol = 1e-10
    ref = orc.bicgstab(rp, ci, va, b, tol=tol, want_hist=True)
    x_sp, info, hist_sp, xs = _scipy_history(spla.bicgstab, A, b, tol, 10 * n)
    assert info == 0 and ref["status"] == 0
    it_sp = len(hist_sp) - 1
    # SciPy may also stop half-way through an iteration (on ||s||), before its callback
    assert it_sp <= ref["iterations"] <= it_sp + 1, (ref["iterations"], it_sp)
    k = min(10, len(hist_sp), len(ref["hist"]))
    assert np.allclose(ref["hist"][:k], hist_sp[:k], rtol=1e-8, atol=0)
    x_d = spla.spsolve(A.tocsc(), b)
    err = np.linalg.norm(ref["x"] - x_d) / np.linalg.norm(x_d)
    cond = np.linalg.cond(A.toarray())
    assert err <= 2 * cond * tol, (err, cond)
    assert orc.true_relres(rp, ci, va, ref["x"], b) <= 1.5 * tol


def test_oracle_bicgstab_iterates_match_scipy(orc):
    """The first iterates agree with SciPy's to rounding (catches a wrong p or x update
    that still converges)."""
    rp, ci, va = _convdiff(10)
    A = _csr(rp, ci, va)
    b = A @ np.random.default_rng(8).uniform(-1, 1, A.shape[0])
    _, _, _, xs = _scipy_history(spla.bicgstab, A, b, 1e-30, 6)
    for m in (1, 3, 6):
        capped = orc.bicgstab(rp, ci, va, b, tol=1e-30, maxit=m)
        assert capped["iterations"] == m
        assert np.allclose(capped["x"], xs[m - 1], rtol=1e-9, atol=1e-12 * np.abs(xs[m - 1]).max())


@pytest.mark.parametrize("k", [2, 4, 7])
def test_oracle_bicgstab_finite_termination(orc, k):
    """Diagonal nonsingular matrix with k distinct eigenvalues: BiCG's polynomial
    annihilates r0 by step k, so BiCGSTAB stops in at most k iterations."""
    n = 160
    rng = np.random.default_rng(40 + k)
    eig = np.linspace(1.0, 6.0, k)
    d = eig[rng.integers(0, k, n)]
    d[:k] = eig
    rp = np.arange(n + 1, dtype=np.int32)
    ci = np.arange(n, dtype=np.int32)
    b = rng.uniform(-1, 1, n)
    ref = orc.bicgstab(rp, ci, d, b, tol=1e-10)
    assert ref["status"] == 0 and ref["iterations"] <= k
    assert np.allclose(ref["x"], b / d, rtol=1e-8, atol=0)


# ------------------------------------------------------------------------------------
# tile-mixed SpMV emulation
# ------------------------------------------------------------------------------------

def _tiles(orc, rp, ci, va, ncols=None):
    n = rp.size - 1
    return orc.tile_create(n, n if ncols is None else ncols, rp, ci, va)


def test_mixed_spmv_exact_on_integers(orc):
    """int8-exact values times small integers: every product and partial sum is exact,
    so the mixed product equals the FP64 tile product and the CSR product bit for bit."""
    rng = np.random.default_rng(3)
    n = 333
    A = sp.random(n, n, density=0.05, random_state=rng, format="csr")
    A.data = rng.integers(-128, 128, A.data.size).astype(np.float64)
    A.sort_indices()
    rp, ci, va = A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data
    d = _tiles(orc, rp, ci, va)
    cls = orc.tile_classes(d)
    assert np.all(cls == 3)
    x = rng.integers(-50, 51, n).astype(np.float64)
    y = orc.mixed_tile_spmv(d, x, cls)
    assert np.array_equal(y, A @ x)
    assert np.array_equal(y, orc.tile_spmv(d, x))


@pytest.mark.parametrize("cls_value", [1, 2, 3])
def test_mixed_spmv_rounding_bound(orc, cls_value):
    """Class c tiles: value at the class precision, x rounded to binary32, product
    rounded to binary32, sum in binary64.  Pre-rounding the operands makes each exact
    product a binary64 number, so the emulation may differ from SciPy's product of the
    rounded operands only by the binary32 product rounding, |err| <= 2^-24 sum|a x|;
    the unrounded operands differ by much more, which the test also asserts."""
    rng = np.random.default_rng(10 + cls_value)
    n = 400
    A = sp.random(n, n, density=0.04, random_state=rng, format="csr")
    if cls_value == 3:
        A.data = rng.integers(-100, 101, A.data.size).astype(np.float64)
    else:
        A.data = rng.uniform(-2.0, 2.0, A.data.size)
    A.sort_indices()
    rp, ci, va = A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data
    d = _tiles(orc, rp, ci, va)
    T = d["tilenum"]
    cls = np.full(T, cls_value, np.int8)
    x = rng.uniform(-1.0, 1.0, n)
    y = orc.mixed_tile_spmv(d, x, cls)
    if cls_value == 2:
        # the matrix's binary16 copy rounds through binary32 (main-cg-mixed.cu:2333-2350)
        a_r = va.astype(np.float32).astype(np.float16).astype(np.float64)
    else:
        a_r = va.astype(np.float32).astype(np.float64)
    x_r = x.astype(np.float32).astype(np.float64)
    Ar = sp.csr_matrix((a_r, ci, rp), shape=(n, n))
    exact = Ar @ x_r
    bound = 2.0 ** -24 * (abs(Ar) @ np.abs(x_r)) + 1e-15 * (abs(Ar) @ np.abs(x_r))
    assert np.all(np.abs(y - exact) <= bound)
    # the rounding is really there: the FP64 product of the unrounded operands differs
    assert not np.allclose(y, A @ x, rtol=1e-12, atol=0)
    # and the FP64 policy is the FP64 tile product
    y64 = orc.mixed_tile_spmv(d, x, cls, policy="fp64")
    assert np.allclose(y64, orc.tile_spmv(d, x), rtol=1e-14, atol=1e-14)
    assert np.allclose(y64, A @ x, rtol=1e-13, atol=1e-13)


def test_mixed_spmv_class0_tiles_stay_binary64(orc):
    """A class-0 tile keeps binary64 products: with every tile at class 0 the mixed
    policy is the FP64 product."""
    rp, ci, va = _random_spd(300, seed=4)
    d = _tiles(orc, rp, ci, va)
    x = np.random.default_rng(2).standard_normal(300)
    cls = np.zeros(d["tilenum"], np.int8)
    y = orc.mixed_tile_spmv(d, x, cls)
    assert np.allclose(y, _csr(rp, ci, va) @ x, rtol=1e-14, atol=1e-14)


def test_mixed_spmv_ragged_last_tile_row(orc):
    """n not a multiple of 16 (ragged last tile row), integers: exact."""
    rng = np.random.default_rng(9)
    n = 101
    A = sp.random(n, n, density=0.1, random_state=rng, format="csr")
    A.data = rng.integers(-9, 10, A.data.size).astype(np.float64)
    A.sort_indices()
    d = _tiles(orc, A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data)
    x = rng.integers(-3, 4, n).astype(np.float64)
    y = orc.mixed_tile_spmv(d, x, orc.tile_classes(d))
    assert np.array_equal(y, A @ x)


# ------------------------------------------------------------------------------------
# magnitude rule of the dynamic classes
# ------------------------------------------------------------------------------------

def test_dyn_classes_rel_thresholds(orc):
    """m = max|p| / rms per 16-entry segment: m > 2^-13 -> 0 (binary64), 2^-24 < m <= 2^-13
    -> 2 (binary16), m <= 2^-24 -> 4 (skip).  Power-of-two rms keeps every ratio exact."""
    rms = 2.0 ** 3
    t13, t24 = 2.0 ** -13, 2.0 ** -24
    mags = [t13 * 1.0000001, t13, t13 * 0.999999, t24 * 1.0000001, t24, t24 * 0.5, 0.0, 1.0]
    expect = [0, 2, 2, 2, 4, 4, 4, 0]
    p = np.zeros(16 * len(mags))
    for s, m in enumerate(mags):
        p[16 * s + 5] = -m * rms  # sign must not matter
    got = orc.dyn_classes_rel(p, rms)
    assert list(got) == expect
    # the segment maximum decides, not the first entry
    q = np.full(16, t24 * 0.5 * rms)
    q[15] = 2 * t13 * rms
    assert list(orc.dyn_classes_rel(q, rms)) == [0]
    # a ragged last segment and rms = 0
    assert list(orc.dyn_classes_rel(np.full(20, 1.0), 1.0)) == [0, 0]
    assert list(orc.dyn_classes_rel(np.full(20, 1.0), 0.0)) == [4, 4]


def test_vectorised_mixed_emulation_matches_tile_emulation(orc):
    """The per-entry vectorised emulation (used at n = 10^6) against the tile-order one:
    the same classes, products within binary64 summation-order rounding."""
    rng = np.random.default_rng(31)
    n = 500
    A = sp.random(n, n, density=0.03, random_state=rng, format="csr")
    vals = rng.uniform(-3.0, 3.0, A.data.size)
    pick = rng.random(A.data.size)
    vals[pick < 0.3] = -1.0                        # int8-exact
    vals[(pick >= 0.3) & (pick < 0.5)] = 0.375     # binary16-exact
    A.data = vals
    A.sort_indices()
    rp, ci, va = A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data
    d = _tiles(orc, rp, ci, va)
    tcls = orc.tile_classes(d)
    ecls = orc.entry_tile_classes(rp, ci, va)
    assert set(np.unique(ecls)) >= {0, 3}
    x = rng.uniform(-1, 1, n)
    y_t = orc.mixed_tile_spmv(d, x, tcls)
    y_v = orc.mixed_csr_spmv(rp, ci, va, x, ecls)
    scale = abs(A) @ np.abs(x)
    assert np.all(np.abs(y_t - y_v) <= 1e-14 * scale)
    # the classes follow the tiles: every entry of a tile carries the tile's class
    tile_of = np.repeat(np.arange(n), np.diff(rp)) // 16 * ((n + 15) // 16) + ci // 16
    for t in np.unique(tile_of)[:200]:
        assert np.unique(ecls[tile_of == t]).size == 1


@pytest.mark.parametrize("nx", [40, 37])
def test_tiled_openmp_cg_matches_oracle_cg(orc, nx):
    """The CPU baseline (tiled traversal, OpenMP) runs the oracle's CG: same iteration count
    within one (reduction order differs), same solution, on a ragged last tile row too."""
    rp, ci, va = _poisson2d(nx)
    n = rp.size - 1
    b = _csr(rp, ci, va) @ np.random.default_rng(nx).uniform(-1, 1, n)
    d = _tiles(orc, rp, ci, va)
    ref = orc.cg(rp, ci, va, b, tol=1e-10)
    got = orc.cg_tiled_omp(d, b, tol=1e-10)
    assert got["status"] == 0 and got["threads"] >= 1
    assert abs(got["iterations"] - ref["iterations"]) <= 1
    assert np.allclose(got["x"], ref["x"], rtol=0, atol=1e-8)
    assert orc.true_relres(rp, ci, va, got["x"], b) <= 1.5e-10
