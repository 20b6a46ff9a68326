"""Host-side Python mirror of the solver's call surface.

Mirrors the reference's operator interface for this path: ``Tile_create`` and the
blockspmv_cpu tables over the 16x16 tiled format (csr2block.h:267-422,
blockspmv_cpu.h:35-97), and the single-kernel CG / BiCGSTAB solves
(main-cg.cu:19,405,1110; main-bicg.cu:13,711,1336) -- all through libmf.so's C-ABI.
GPU vectors are torch tensors (device memory and streams only); every compute call
runs HIP kernels, and nothing here computes on the CPU.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib as L
from ._lib import MfError, as_f64, as_i32, check, f32p, f64p, i32p, ptr, u8p

_PREC = {"fp64": L.MF_PREC_FP64, "mixed": L.MF_PREC_MIXED}
_POLICY = {"lossless": L.MF_CLASS_LOSSLESS, "rowscaled": L.MF_CLASS_ROWSCALED, "fp64": L.MF_CLASS_ALL_FP64}


# --------------------------------------------------------------------------------------
# tiled format (host)
# --------------------------------------------------------------------------------------
class Tiles:
    """Reference-layout tile matrix produced by ``mf_tile_create``."""

    def __init__(self, rowA, colA, rowptr, colidx, val, val_low=None):
        self._keep = (as_i32(rowptr), as_i32(colidx), as_f64(val),
                      None if val_low is None else np.ascontiguousarray(val_low, np.float32))
        rp, ci, va, vl = self._keep
        self.rowA, self.colA = int(rowA), int(colA)
        self.m = L.TileMatrix()
        st = L.lib().mf_tile_create(ctypes.byref(self.m), self.rowA, self.colA, int(ci.size), ptr(rp, i32p),
                                    ptr(ci, i32p), ptr(va, f64p), None if vl is None else ptr(vl, f32p))
        check(st, "mf_tile_create")

    def close(self):
        if self.m is not None:
            L.lib().mf_tile_destroy(ctypes.byref(self.m))
            self.m = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def arrays(self):
        m = self.m
        T = m.tilenum
        cs = m.csrsize // 2 if m.csrsize % 2 == 0 else m.csrsize // 2 + 1

        def a(p, n, dt):
            if n <= 0:
                return np.zeros(0, dt)
            return np.ctypeslib.as_array(ctypes.cast(p, ctypes.POINTER(np.ctypeslib.as_ctypes_type(dt))),
                                         shape=(n,)).copy()
        return dict(
            tilem=m.tilem, tilen=m.tilen, tilenum=T, csrsize=m.csrsize, csrptrlen=m.csrptrlen,
            tile_ptr=a(m.tile_ptr, m.tilem + 1, np.int32), tile_columnidx=a(m.tile_columnidx, T, np.int32),
            tile_nnz=a(m.tile_nnz, T + 1, np.int32), Format=a(m.Format, T, np.int8),
            blknnz=a(m.blknnz, T + 1, np.int32), blknnznnz=a(m.blknnznnz, T + 1, np.uint8),
            tilewidth=a(m.tilewidth, T, np.int8), csr_offset=a(m.csr_offset, T + 1, np.int32),
            csrptr_offset=a(m.csrptr_offset, T + 1, np.int32),
            Blockcsr_Val=a(m.Blockcsr_Val, m.csrsize, np.float64),
            Blockcsr_Val_Low=a(m.Blockcsr_Val_Low, m.csrsize, np.float32),
            Tile_csr_Col=a(m.Tile_csr_Col, m.csrsize, np.uint8),
            Blockcsr_Ptr=a(m.Blockcsr_Ptr, m.csrptrlen, np.uint8),
            csr_compressedIdx=a(m.csr_compressedIdx, cs, np.uint8),
        )

    def prep_tables(self):
        T = self.m.tilenum
        p1 = np.zeros(max(T, 1), np.int32)
        p2 = np.zeros(max(T, 1), np.int32)
        n = ctypes.c_int(0)
        lb = L.lib()
        check(lb.mf_tile_prep_tables(ctypes.byref(self.m), self.rowA, ptr(p1, i32p), ptr(p2, i32p),
                                     ctypes.byref(n), None, None, None), "mf_tile_prep_tables")
        k = n.value
        idx = np.zeros(max(k, 1), np.uint32)
        cs = np.zeros(max(k, 1), np.int32)
        ce = np.zeros(max(k, 1), np.int32)
        check(lb.mf_tile_prep_tables(ctypes.byref(self.m), self.rowA, ptr(p1, i32p), ptr(p2, i32p),
                                     ctypes.byref(n), idx.ctypes.data_as(ctypes.POINTER(ctypes.c_uint)),
                                     ptr(cs, i32p), ptr(ce, i32p)), "mf_tile_prep_tables")
        return dict(ptroffset1=p1[:T], ptroffset2=p2[:T], rowblkblock=k, rowblk_idx=idx[:k],
                    rowblk_colstart=cs[:k], rowblk_colstop=ce[:k])

    def row_lists(self):
        T = self.m.tilenum
        nzr = np.zeros(T + 1, np.int32)
        tot = ctypes.c_int(0)
        lb = L.lib()
        check(lb.mf_tile_row_lists(ctypes.byref(self.m), self.rowA, ptr(nzr, i32p), None, None,
                                   ctypes.byref(tot)), "mf_tile_row_lists")
        n = tot.value
        rid = np.zeros(max(n, 1), np.uint8)
        rend = np.zeros(max(n, 1), np.uint8)
        check(lb.mf_tile_row_lists(ctypes.byref(self.m), self.rowA, ptr(nzr, i32p), ptr(rid, u8p), ptr(rend, u8p),
                                   ctypes.byref(tot)), "mf_tile_row_lists")
        return dict(nonzero_row_new=nzr, blockrowid_new=rid[:n], blockcsr_ptr_new=rend[:n])


def tile_create(rowA, colA, rowptr, colidx, val, val_low=None):
    """Tile_create-compatible arrays as a dict of NumPy arrays."""
    t = Tiles(rowA, colA, rowptr, colidx, val, val_low)
    try:
        return t.arrays()
    finally:
        t.close()


def encode(s):
    s = np.ascontiguousarray(s, np.uint8)
    n = s.size
    res = np.zeros(n // 2 + 2, np.uint8)
    L.lib().mf_encode(ptr(s if n else np.zeros(1, np.uint8), u8p), ptr(res, u8p), n, 0)
    return res[: (n + 1) // 2]


def decode(packed, n):
    packed = np.ascontiguousarray(packed, np.uint8)
    out = np.zeros(n + 1, np.uint8)
    L.lib().mf_decode(ptr(packed if packed.size else np.zeros(1, np.uint8), u8p), ptr(out, u8p), int(n))
    return out[:n]


def value_copies(val):
    """(f32, f16 bits, i8) copies as in main-cg-mixed.cu:2333-2350."""
    v = as_f64(val)
    f32 = np.zeros(v.size, np.float32)
    f16 = np.zeros(v.size, np.uint16)
    i8 = np.zeros(v.size, np.int8)
    L.lib().mf_value_copies(ptr(v, f64p), v.size, ptr(f32, f32p),
                            f16.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16)),
                            i8.ctypes.data_as(ctypes.POINTER(ctypes.c_int8)))
    return f32, f16, i8


def tile_classes(rowA, colA, rowptr, colidx, val, policy="lossless"):
    rp, ci, va = as_i32(rowptr), as_i32(colidx), as_f64(val)
    n = ctypes.c_int64(0)
    lb = L.lib()
    check(lb.mf_tile_classes(rowA, colA, ptr(rp, i32p), ptr(ci, i32p), ptr(va, f64p), _POLICY[policy], None, 0,
                             ctypes.byref(n)), "mf_tile_classes")
    out = np.zeros(max(n.value, 1), np.uint8)
    check(lb.mf_tile_classes(rowA, colA, ptr(rp, i32p), ptr(ci, i32p), ptr(va, f64p), _POLICY[policy],
                             ptr(out, u8p), out.size, ctypes.byref(n)), "mf_tile_classes")
    return out[: n.value]


def read_matrix(path):
    """Matrix Market / .cbd reader with biio semantics -> (m, n, rowptr, colidx, val, is_symmetric)."""
    m, n, z, sym = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    rp, ci, va = i32p(), i32p(), f64p()
    lb = L.lib()
    check(lb.mf_read_matrix(str(path).encode(), ctypes.byref(m), ctypes.byref(n), ctypes.byref(z),
                            ctypes.byref(rp), ctypes.byref(ci), ctypes.byref(va), ctypes.byref(sym)),
          "mf_read_matrix")
    try:
        rowptr = np.ctypeslib.as_array(rp, shape=(m.value + 1,)).copy()
        colidx = np.ctypeslib.as_array(ci, shape=(max(z.value, 1),))[: z.value].copy()
        val = np.ctypeslib.as_array(va, shape=(max(z.value, 1),))[: z.value].copy()
    finally:
        for p in (rp, ci, va):
            lb.mf_free(ctypes.cast(p, ctypes.c_void_p))
    return m.value, n.value, rowptr, colidx, val, bool(sym.value)


def write_cbd(path, m, n, rowptr, colidx, val):
    rp, ci, va = as_i32(rowptr), as_i32(colidx), as_f64(val)
    check(L.lib().mf_write_cbd(str(path).encode(), int(m), int(n), int(ci.size), ptr(rp, i32p), ptr(ci, i32p),
                               ptr(va, f64p)), "mf_write_cbd")


# --------------------------------------------------------------------------------------
# GPU matrix and solvers
# --------------------------------------------------------------------------------------
def _torch():
    import torch
    return torch


def _stream_handle():
    torch = _torch()
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _dev_vec(v, n, device):
    torch = _torch()
    if isinstance(v, torch.Tensor):
        t = v.to(device=device, dtype=torch.float64).contiguous()
    else:
        t = torch.as_tensor(np.ascontiguousarray(v, np.float64), device=device)
    if t.numel() != n:
        raise ValueError(f"vector has {t.numel()} entries, expected {n}")
    return t


def solve_options(tol=1e-10, max_iter=100000, precision="fp64", fixed_iterations=False, dynamic_precision=0,
                  replace_delta=0.0, max_replacements=64, grid=0, variant=0, use_x0=False, watchdog_ms=0.0,
                  layout="auto", value_cache=True, transpose=False):
    o = L.SolveOptions()
    L.lib().mf_solve_options_default(ctypes.byref(o))
    o.tol = float(tol)
    o.max_iter = int(max_iter)
    o.precision = _PREC[precision] if isinstance(precision, str) else int(precision)
    o.fixed_iterations = int(bool(fixed_iterations))
    o.dynamic_precision = int(dynamic_precision)
    o.replace_delta = float(replace_delta)
    o.max_replacements = int(max_replacements)
    o.grid = int(grid)
    o.variant = int(variant)
    o.use_x0 = int(bool(use_x0))
    o.watchdog_ms = float(watchdog_ms)
    o.layout = {"auto": 0, "chunks": 1, "sell": 2}[layout] if isinstance(layout, str) else int(layout)
    o.value_cache = 0 if value_cache else 1
    o.transpose = int(bool(transpose))
    return o


class TiledMatrix:
    """A CSR matrix converted to the tiled format and resident on one GPU."""

    def __init__(self, rowptr, colidx, val, n_cols=None, class_policy="lossless", device=0, keep_mixed=True):
        rp, ci, va = as_i32(rowptr), as_i32(colidx), as_f64(val)
        self.n = int(rp.size - 1)
        self.n_cols = int(n_cols if n_cols is not None else self.n)
        self.device = int(device)
        mo = L.MatrixOptions()
        L.lib().mf_matrix_options_default(ctypes.byref(mo))
        mo.class_policy = _POLICY[class_policy] if isinstance(class_policy, str) else int(class_policy)
        mo.device = self.device
        mo.keep_mixed = int(bool(keep_mixed))
        h = ctypes.c_void_p()
        torch = _torch()
        torch.cuda.set_device(self.device)
        check(L.lib().mf_matrix_create(self.n, self.n_cols, ptr(rp, i32p), ptr(ci, i32p), ptr(va, f64p),
                                       ctypes.byref(mo), _stream_handle(), ctypes.byref(h)), "mf_matrix_create")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            L.lib().mf_matrix_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def stats(self):
        s = L.MatrixStats()
        check(L.lib().mf_matrix_get_stats(self.h, ctypes.byref(s)), "mf_matrix_get_stats")
        d = {f: getattr(s, f) for f, _ in s._fields_}
        d["tiles_by_class"] = list(s.tiles_by_class)
        d["nnz_by_class"] = list(s.nnz_by_class)
        return d

    def update_values(self, val):
        va = as_f64(val)
        check(L.lib().mf_matrix_update_values(self.h, ptr(va, f64p), _stream_handle()), "mf_matrix_update_values")

    def spmv(self, x, precision="fp64", transpose=False, out=None):
        torch = _torch()
        dev = torch.device("cuda", self.device)
        n_in = self.n if transpose else self.n_cols
        n_out = self.n_cols if transpose else self.n
        xt = _dev_vec(x, n_in, dev)
        y = out if out is not None else torch.empty(n_out, dtype=torch.float64, device=dev)
        check(L.lib().mf_spmv(self.h, ctypes.c_void_p(xt.data_ptr()), ctypes.c_void_p(y.data_ptr()),
                              _PREC[precision], int(bool(transpose)), _stream_handle()), "mf_spmv")
        return y

    def pattern_outer(self, u, v, scale=1.0, out=None):
        """out[k] = scale * u[row(k)] * v[col(k)] over the stored entries, CSR order."""
        torch = _torch()
        dev = torch.device("cuda", self.device)
        ut = _dev_vec(u, self.n, dev)
        vt = _dev_vec(v, self.n_cols, dev)
        nnz = int(self.stats()["nnz"])
        o = out if out is not None else torch.empty(nnz, dtype=torch.float64, device=dev)
        check(L.lib().mf_pattern_outer(self.h, ctypes.c_void_p(ut.data_ptr()), ctypes.c_void_p(vt.data_ptr()),
                                       float(scale), ctypes.c_void_p(o.data_ptr()), _stream_handle()),
              "mf_pattern_outer")
        return o

    def _solve(self, fn_name, b, x0, opts):
        torch = _torch()
        dev = torch.device("cuda", self.device)
        bt = _dev_vec(b, self.n, dev)
        if x0 is not None:
            xt = _dev_vec(x0, self.n, dev).clone()
            opts.use_x0 = 1
        else:
            xt = torch.zeros(self.n, dtype=torch.float64, device=dev)
        info = L.SolveInfo()
        st = getattr(L.lib(), fn_name)(self.h, ctypes.c_void_p(bt.data_ptr()), ctypes.c_void_p(xt.data_ptr()),
                                       ctypes.byref(opts), ctypes.byref(info), _stream_handle())
        d = info.as_dict()
        d["status_name"] = L.STATUS_NAMES.get(st, str(st))
        if st not in (L.MF_SUCCESS, 4):  # MF_NOT_CONVERGED is reported, not raised
            raise MfError(st, fn_name)
        return xt, d

    def autotune(self, b, precision="fp64"):
        """Pick workgroups and variant of the persistent CG for this matrix; returns
        (grid, variant, ms per iteration)."""
        torch = _torch()
        dev = torch.device("cuda", self.device)
        bt = _dev_vec(b, self.n, dev)
        xt = torch.zeros(self.n, dtype=torch.float64, device=dev)
        g, v, t = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        check(L.lib().mf_cg_autotune(self.h, ctypes.c_void_p(bt.data_ptr()), ctypes.c_void_p(xt.data_ptr()),
                                     _PREC[precision], ctypes.byref(g), ctypes.byref(v), ctypes.byref(t),
                                     _stream_handle()), "mf_cg_autotune")
        return g.value, v.value, t.value

    def cg(self, b, x0=None, **kw):
        """Single-kernel CG; returns (x tensor, info dict)."""
        return self._solve("mf_cg_solve", b, x0, solve_options(**kw))

    def bicgstab(self, b, x0=None, **kw):
        """Single-kernel BiCGSTAB; returns (x tensor, info dict)."""
        return self._solve("mf_bicgstab_solve", b, x0, solve_options(**kw))

    def cg_multikernel(self, b, x0=None, **kw):
        return self._solve("mf_cg_solve_multikernel", b, x0, solve_options(**kw))

    def bicgstab_multikernel(self, b, x0=None, **kw):
        return self._solve("mf_bicgstab_solve_multikernel", b, x0, solve_options(**kw))


def dyn_classify(p_old, p_new, mode=2, rms=1.0, device=0):
    """Partial-convergence class per 16-element segment, computed on the GPU."""
    torch = _torch()
    dev = torch.device("cuda", device)
    po = torch.as_tensor(np.ascontiguousarray(p_old, np.float64), device=dev)
    pn = torch.as_tensor(np.ascontiguousarray(p_new, np.float64), device=dev)
    n = po.numel()
    out = torch.zeros((n + 15) // 16, dtype=torch.uint8, device=dev)
    check(L.lib().mf_dyn_classify(ctypes.c_void_p(po.data_ptr()), ctypes.c_void_p(pn.data_ptr()), n, int(mode),
                                  float(rms), ctypes.c_void_p(out.data_ptr()), _stream_handle()), "mf_dyn_classify")
    return out.cpu().numpy()


def device_name(device=0):
    buf = ctypes.create_string_buffer(256)
    check(L.lib().mf_device_name(int(device), buf, 256), "mf_device_name")
    return buf.value.decode()


class Batch:
    """Many independent square systems solved by one persistent launch."""

    def __init__(self, systems, class_policy="lossless", device=0, keep_mixed=True):
        """systems: list of (rowptr, colidx, val) CSR triples."""
        self._keep = [(as_i32(r), as_i32(c), as_f64(v)) for r, c, v in systems]
        cnt = len(self._keep)
        self.count = cnt
        self.device = int(device)
        rows = np.array([r.size - 1 for r, _, _ in self._keep], np.int32)
        RP = (i32p * cnt)(*[ptr(r, i32p) for r, _, _ in self._keep])
        CI = (i32p * cnt)(*[ptr(c, i32p) for _, c, _ in self._keep])
        VA = (f64p * cnt)(*[ptr(v, f64p) for _, _, v in self._keep])
        mo = L.MatrixOptions()
        L.lib().mf_matrix_options_default(ctypes.byref(mo))
        mo.class_policy = _POLICY[class_policy] if isinstance(class_policy, str) else int(class_policy)
        mo.device = self.device
        mo.keep_mixed = int(bool(keep_mixed))
        h = ctypes.c_void_p()
        _torch().cuda.set_device(self.device)
        check(L.lib().mf_batch_create(cnt, ptr(rows, i32p), RP, CI, VA, ctypes.byref(mo), _stream_handle(),
                                      ctypes.byref(h)), "mf_batch_create")
        self.h = h
        off = np.zeros(cnt + 1, np.int64)
        check(L.lib().mf_batch_row_offsets(self.h, off.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))),
              "mf_batch_row_offsets")
        self.offsets = off
        self.rows = rows

    def close(self):
        if getattr(self, "h", None):
            L.lib().mf_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_rows_per_workgroup(self, r):
        check(L.lib().mf_batch_set_groups(self.h, int(r)), "mf_batch_set_groups")

    def autotune(self, b_packed, precision="mixed", iterations=30):
        """Pick rows per workgroup and chunk groups per wave for this batch (kept for later
        solves); returns (rows_per_wg, unroll, ms of the fastest trial)."""
        r, u, t = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        check(L.lib().mf_batch_autotune(self.h, ctypes.c_void_p(b_packed.data_ptr()), _PREC[precision], int(iterations),
                                        ctypes.byref(r), ctypes.byref(u), ctypes.byref(t), _stream_handle()),
              "mf_batch_autotune")
        return r.value, u.value, t.value

    def timeline(self):
        """Per-system (start_ms, end_ms, group size) of the last batched launch."""
        st, en = np.zeros(self.count), np.zeros(self.count)
        gs = np.zeros(self.count, np.int32)
        check(L.lib().mf_batch_timeline(self.h, st.ctypes.data_as(L.f64p), en.ctypes.data_as(L.f64p),
                                        gs.ctypes.data_as(L.i32p)), "mf_batch_timeline")
        return st, en, gs

    def pack(self, vectors):
        """Concatenate per-system vectors at the padded offsets (zeros in the padding)."""
        torch = _torch()
        out = np.zeros(int(self.offsets[-1]))
        for s, v in enumerate(vectors):
            out[self.offsets[s]: self.offsets[s] + self.rows[s]] = np.asarray(v, np.float64)
        return torch.as_tensor(out, device=torch.device("cuda", self.device))

    def unpack(self, t):
        a = t.detach().cpu().numpy() if hasattr(t, "detach") else np.asarray(t)
        return [a[self.offsets[s]: self.offsets[s] + self.rows[s]].copy() for s in range(self.count)]

    def cg(self, b_packed, use_graph=True, **kw):
        """Solve every system; returns (x packed tensor, list of info dicts, launch ms)."""
        torch = _torch()
        opts = solve_options(**kw)
        x = torch.zeros_like(b_packed)
        infos = (L.SolveInfo * self.count)()
        ms = ctypes.c_double(0)
        st = L.lib().mf_cg_solve_batched(self.h, ctypes.c_void_p(b_packed.data_ptr()), ctypes.c_void_p(x.data_ptr()),
                                          ctypes.byref(opts), int(bool(use_graph)), infos, ctypes.byref(ms),
                                          _stream_handle())
        if st not in (L.MF_SUCCESS, 4):
            raise MfError(st, "mf_cg_solve_batched")
        return x, [infos[i].as_dict() for i in range(self.count)], ms.value
