"""Synthetic sparse systems for the solver's benchmarks and parity tests.

Every generator is seeded with ``numpy.random.default_rng`` and returns CSR arrays
(int32 row pointer, int32 column index, float64 values; columns ascending inside a
row) plus the right-hand side.  Grid ordering is lexicographic with x fastest:
``idx = i + nx*(j + ny*k)``.

* ``poisson2d`` / ``poisson3d`` -- 5-point / 7-point Laplacian, diagonal 4 / 6,
  off-diagonals -1, homogeneous Dirichlet boundary (missing neighbours dropped).
* ``convdiff3d`` -- 7-point convection-diffusion with first-order upwinding.
  Diffusion kappa = 1, spacing h = 1; the velocity is scaled so that the cell
  Peclet number ||v||_inf h / (2 kappa) equals ``peclet``, giving v = (1, 0.5, 0.25)
  at Pe 0.5.  Row i: diagonal 6 + |vx| + |vy| + |vz|, upstream neighbour -1 - |v_d|,
  downstream neighbour -1.
* ``irregular_spd`` -- lognormal row lengths, banded random columns, symmetrised,
  diagonally dominant, then symmetrically rescaled by D = 10^U[0,6]; a fraction of
  off-diagonal pairs is then set to exactly -1 (kept symmetric).
* ``batched_poisson`` -- independent shifted 2D/3D Poisson systems of log-uniform size.
* ``varcoef_diffusion3d`` -- 7-point diffusion with k = exp(g), g a Gaussian random
  field with squared-exponential covariance (FFT spectral synthesis), harmonic-mean
  face coefficients, Dirichlet boundary faces weighted by the cell coefficient.
"""
from __future__ import annotations

import math

import numpy as np


def _csr_from_coo(n, rows, cols, vals, ncols=None):
    rows = np.asarray(rows, np.int64)
    cols = np.asarray(cols, np.int64)
    vals = np.asarray(vals, np.float64)
    order = np.lexsort((cols, rows))
    rows, cols, vals = rows[order], cols[order], vals[order]
    rowptr = np.zeros(n + 1, np.int64)
    np.add.at(rowptr, rows + 1, 1)
    rowptr = np.cumsum(rowptr)
    return rowptr.astype(np.int32), cols.astype(np.int32), vals


def _stencil(dims, center, neigh):
    """Assemble a constant-coefficient stencil on a box grid.

    ``neigh`` is a list of (offset vector, coefficient)."""
    dims = tuple(int(d) for d in dims)
    n = int(np.prod(dims))
    grid = np.indices(dims[::-1]).reshape(len(dims), -1)[::-1]  # x fastest
    idx = np.arange(n, dtype=np.int64)
    rows = [idx]
    cols = [idx]
    vals = [np.full(n, float(center))]
    strides = [1]
    for d in dims[:-1]:
        strides.append(strides[-1] * d)
    for off, coef in neigh:
        ok = np.ones(n, bool)
        shift = 0
        for ax, o in enumerate(off):
            if o:
                ok &= (grid[ax] + o >= 0) & (grid[ax] + o < dims[ax])
                shift += o * strides[ax]
        rows.append(idx[ok])
        cols.append(idx[ok] + shift)
        vals.append(np.full(int(ok.sum()), float(coef)))
    return _csr_from_coo(n, np.concatenate(rows), np.concatenate(cols), np.concatenate(vals))


def poisson2d(nx, ny=None):
    ny = nx if ny is None else ny
    nb = [((1, 0), -1.0), ((-1, 0), -1.0), ((0, 1), -1.0), ((0, -1), -1.0)]
    return _stencil((nx, ny), 4.0, nb)


def poisson3d(nx, ny=None, nz=None):
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    nb = []
    for ax in range(3):
        for s in (1, -1):
            o = [0, 0, 0]
            o[ax] = s
            nb.append((tuple(o), -1.0))
    return _stencil((nx, ny, nz), 6.0, nb)


def convdiff3d(nx, velocity=(1.0, 0.5, 0.25), peclet=0.5, ny=None, nz=None):
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    v = np.asarray(velocity, np.float64)
    kappa, h = 1.0, 1.0
    v = v * (peclet * 2.0 * kappa / h) / np.max(np.abs(v))
    center = 6.0 * kappa / (h * h) + float(np.sum(np.abs(v))) / h
    nb = []
    for ax in range(3):
        up = -1 if v[ax] >= 0 else 1          # neighbour on the upstream side
        o_up = [0, 0, 0]
        o_up[ax] = up
        o_dn = [0, 0, 0]
        o_dn[ax] = -up
        nb.append((tuple(o_up), -kappa / (h * h) - abs(v[ax]) / h))
        nb.append((tuple(o_dn), -kappa / (h * h)))
    return _stencil((nx, ny, nz), center, nb)


def rhs_from_solution(rowptr, colidx, val, seed):
    """b = A x* with x* ~ U[-1, 1] from default_rng(seed)."""
    n = rowptr.size - 1
    rng = np.random.default_rng(seed)
    xs = rng.uniform(-1.0, 1.0, n)
    b = spmv(rowptr, colidx, val, xs)
    return b, xs


def spmv(rowptr, colidx, val, x):
    n = rowptr.size - 1
    rows = np.repeat(np.arange(n), np.diff(rowptr))
    y = np.zeros(n)
    np.add.at(y, rows, val * x[colidx])
    return y


def irregular_spd(n=1_000_000, seed=2, mu=3.0, sigma=0.8, lo=3, hi=400, band=20000,
                  logd_max=6.0, frac_minus_one=0.3):
    """Irregular SPD system (lognormal rows, banded columns, D-scaled).

    Returns (rowptr, colidx, val, b, xstar)."""
    rng = np.random.default_rng(seed)
    lens = np.clip(np.round(rng.lognormal(mu, sigma, n)), lo, hi).astype(np.int64)
    # off-diagonal entries of the strictly lower half: half of each row's length
    half = np.maximum(1, lens // 2)
    rows = np.repeat(np.arange(n, dtype=np.int64), half)
    off = rng.integers(1, band // 2 + 1, rows.size)
    cols = rows - off
    keep = cols >= 0
    rows, cols = rows[keep], cols[keep]
    vals = -rng.uniform(0.0, 1.0, rows.size)
    # merge duplicates, symmetrise
    key = rows * n + cols
    key, inv = np.unique(key, return_inverse=True)
    v = np.zeros(key.size)
    np.add.at(v, inv, vals)
    r = key // n
    c = key % n
    R = np.concatenate([r, c])
    C = np.concatenate([c, r])
    V = np.concatenate([v, v])
    rowsum = np.zeros(n)
    np.add.at(rowsum, R, np.abs(V))
    diag = rowsum + 1e-3
    d = 10.0 ** rng.uniform(0.0, logd_max, n)
    V = V * d[R] * d[C]
    diag = diag * d * d
    # a fraction of the symmetric off-diagonal pairs set to exactly -1
    npair = r.size
    pick = rng.random(npair) < frac_minus_one
    pick2 = np.concatenate([pick, pick])
    V[pick2] = -1.0
    rowsA = np.concatenate([R, np.arange(n)])
    colsA = np.concatenate([C, np.arange(n)])
    valsA = np.concatenate([V, diag])
    rowptr, colidx, val = _csr_from_coo(n, rowsA, colsA, valsA)
    b, xs = rhs_from_solution(rowptr, colidx, val, seed)
    return rowptr, colidx, val, b, xs


def batched_poisson(count=512, seed=3, nmin=1024, nmax=65536):
    """List of (rowptr, colidx, val, b) independent SPD systems.

    Sizes log-uniform in [nmin, nmax]; even systems 2D 5-point, odd systems 3D
    7-point; each diagonal shifted by U[0, 0.1]; b ~ U[-1, 1]."""
    rng = np.random.default_rng(seed)
    out = []
    for s in range(count):
        n = int(round(math.exp(rng.uniform(math.log(nmin), math.log(nmax)))))
        shift = rng.uniform(0.0, 0.1)
        if s % 2 == 0:
            nx = max(2, int(round(math.sqrt(n))))
            ny = max(2, int(round(n / nx)))
            rp, ci, va = poisson2d(nx, ny)
        else:
            nx = max(2, int(round(n ** (1.0 / 3.0))))
            nz = max(2, int(round(n / (nx * nx))))
            rp, ci, va = poisson3d(nx, nx, nz)
        m = rp.size - 1
        va = va.copy()
        diag_pos = _diag_positions(rp, ci)
        va[diag_pos] += shift
        b = rng.uniform(-1.0, 1.0, m)
        out.append((rp, ci, va, b))
    return out


def _diag_positions(rowptr, colidx):
    n = rowptr.size - 1
    rows = np.repeat(np.arange(n), np.diff(rowptr))
    pos = np.nonzero(rows == colidx)[0]
    return pos


def gaussian_random_field(shape, length_scale=8.0, variance=1.0, seed=4):
    """Stationary GRF with covariance variance*exp(-r^2 / (2 l^2)) on a periodic
    box by spectral synthesis, normalised to the requested sample variance."""
    rng = np.random.default_rng(seed)
    noise = rng.standard_normal(shape)
    freqs = np.meshgrid(*[np.fft.fftfreq(s) for s in shape], indexing="ij")
    k2 = sum((2.0 * np.pi * f) ** 2 for f in freqs)
    spec = np.exp(-0.5 * k2 * length_scale ** 2)  # SE spectral density (up to a constant)
    g = np.real(np.fft.ifftn(np.fft.fftn(noise) * np.sqrt(spec)))
    g = (g - g.mean()) / g.std() * math.sqrt(variance)
    return g


def varcoef_diffusion3d(nx=64, seed=4, length_scale=8.0):
    """Returns (rowptr, colidx, val, kcell) for -div(k grad u) on an nx^3 grid."""
    g = gaussian_random_field((nx, nx, nx), length_scale, 1.0, seed)
    k = np.exp(g).transpose(2, 1, 0).reshape(-1)  # field indexed [x,y,z] -> idx x fastest
    return _varcoef_assemble(nx, k) + (k,)


def _varcoef_assemble(nx, k):
    n = nx ** 3
    idx = np.arange(n, dtype=np.int64)
    coords = [idx % nx, (idx // nx) % nx, idx // (nx * nx)]
    strides = [1, nx, nx * nx]
    diag = np.zeros(n)
    rows, cols, vals = [], [], []
    for ax in range(3):
        for s in (1, -1):
            nb_ok = (coords[ax] + s >= 0) & (coords[ax] + s < nx)
            j = idx + s * strides[ax]
            kf = np.where(nb_ok, 2.0 * k * k[np.where(nb_ok, j, idx)] /
                          (k + k[np.where(nb_ok, j, idx)]), k)
            diag += kf
            rows.append(idx[nb_ok])
            cols.append(j[nb_ok])
            vals.append(-kf[nb_ok])
    rows.append(idx)
    cols.append(idx)
    vals.append(diag)
    return _csr_from_coo(n, np.concatenate(rows), np.concatenate(cols), np.concatenate(vals))
