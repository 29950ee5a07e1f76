"""Device-resident CSR matrices.

Mirrors the reference's ``SparseMatrix = leto_ops::CsrMatrix``
(crates/cfd-math/src/sparse/mod.rs:53), built with
``CsrMatrix::from_parts(values, col_indices, row_ptr, nrows, ncols)``, and
``SparseMatrixBuilder`` (crates/cfd-math/src/sparse/builder.rs:70-386:
duplicates summed in insertion order, rows sorted by column).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N


def _is_torch(a) -> bool:
    return hasattr(a, "data_ptr") and hasattr(a, "is_cuda")


def require_device_vectors(*tensors) -> None:
    """Every tensor handed to a *_device entry point must be a contiguous
    CUDA tensor: a host pointer or a strided view would be read or written
    by the kernels as if it were a dense device array."""
    for t in tensors:
        if not _is_torch(t):
            raise N.CfdlaError(N.ERR_INVALID_ARG, "device entry points take torch tensors")
        if not t.is_cuda:
            raise N.CfdlaError(N.ERR_INVALID_ARG, "device entry points take CUDA tensors, got a CPU tensor")
        if not t.is_contiguous():
            raise N.CfdlaError(N.ERR_INVALID_ARG, "device vectors must be contiguous")


def bind_torch_stream() -> None:
    """Issue library work on torch's current HIP stream."""
    import torch

    N.call("cfdla_set_stream", C.c_void_p(torch.cuda.current_stream().cuda_stream))


class CsrMatrix:
    """CSR matrix copied to the GPU (int32 indices, FP64 values + FP32 copy).

    The caller keeps its host arrays; ``row_ptr``/``col_indices``/``values``
    return host copies in sorted-column order.
    """

    def __init__(self, handle, nrows: int, ncols: int, nnz: int, row_ptr, col_indices, values, order=None):
        self._h = handle
        self.nrows, self.ncols, self.nnz = nrows, ncols, nnz
        self.row_ptr, self.col_indices, self._values = row_ptr, col_indices, values
        self._order = order  # sorted position -> creation position (None: rows were sorted)

    @property
    def values(self):
        """Values in sorted-column order (fetched from the device after a
        device update)."""
        if self._values is None:
            v = np.empty(self.nnz)
            N.call("cfdla_matrix_get_values", self.handle, N.ptr(v))
            self._values = v if self._order is None else v[self._order]
        return self._values

    def creation_order_values(self):
        """Values in the order the matrix was created with (the order every
        values argument of the library takes)."""
        v = self.values
        if self._order is None:
            return v
        out = np.empty_like(v)
        out[self._order] = v
        return out

    @classmethod
    def from_parts(cls, values, col_indices, row_ptr, nrows: int, ncols: int) -> "CsrMatrix":
        rp = np.ascontiguousarray(row_ptr, dtype=np.int64)
        ci = np.ascontiguousarray(col_indices, dtype=np.int64)
        v = np.ascontiguousarray(values, dtype=np.float64)
        if rp.shape != (nrows + 1,):
            raise N.CfdlaError(N.ERR_DIMENSION, f"row_ptr has {rp.size} entries, expected {nrows + 1}")
        if ci.size != v.size or ci.size != int(rp[-1]):
            raise N.CfdlaError(N.ERR_DIMENSION, "col_indices/values length differs from row_ptr[nrows]")
        h = C.c_void_p()
        N.call("cfdla_matrix_create", nrows, ncols, N.ptr(rp), N.ptr(ci), N.ptr(v), C.byref(h))
        # keep host arrays in sorted order
        order = np.concatenate([np.argsort(ci[rp[i]:rp[i + 1]], kind="stable") + rp[i] for i in range(nrows)]) \
            if nrows and not _rows_sorted(rp, ci) else None
        if order is not None:
            ci, v = ci[order], v[order]
        return cls(h, nrows, ncols, int(rp[-1]), rp, ci, v, order)

    @classmethod
    def from_scipy(cls, m) -> "CsrMatrix":
        m = m.tocsr()
        return cls.from_parts(m.data, m.indices, m.indptr, m.shape[0], m.shape[1])

    @property
    def shape(self):
        return (self.nrows, self.ncols)

    @property
    def handle(self):
        if self._h is None:
            raise ValueError("matrix has been closed")
        return self._h

    def update_values(self, values) -> None:
        """Replace the values of the same pattern (caller's original order)."""
        v = np.ascontiguousarray(values, dtype=np.float64)
        if v.size != self.nnz:
            raise N.CfdlaError(N.ERR_DIMENSION, "values length differs from nnz")
        N.call("cfdla_matrix_update_values", self.handle, N.ptr(v))
        self._values = v.copy() if self._order is None else v[self._order]

    def update_values_device(self, values) -> None:
        """Replace the values from a CUDA float64 tensor of nnz entries in the
        caller's original order; the device copies are rewritten on the GPU
        and no host copy is made."""
        import torch

        if not (_is_torch(values) and values.is_cuda and values.dtype == torch.float64 and values.is_contiguous()):
            raise N.CfdlaError(N.ERR_INVALID_ARG, "values must be a contiguous CUDA float64 tensor")
        if values.numel() != self.nnz:
            raise N.CfdlaError(N.ERR_DIMENSION, "values length differs from nnz")
        bind_torch_stream()
        N.call("cfdla_matrix_update_values_device", self.handle, values.data_ptr())
        self._values = None

    def spmv(self, x):
        """y = A x (numpy host vector or torch device tensor, FP64 or FP32)."""
        if _is_torch(x):
            import torch

            y = torch.empty(self.nrows, dtype=x.dtype, device=x.device)
            self.spmv_into(x, y)
            return y
        x = np.ascontiguousarray(x, dtype=np.float64)
        if x.shape != (self.ncols,):
            raise N.CfdlaError(N.ERR_DIMENSION, f"x has shape {x.shape}, expected ({self.ncols},)")
        y = np.empty(self.nrows)
        N.call("cfdla_spmv", self.handle, N.ptr(x), N.ptr(y))
        return y

    def spmv_into(self, x, y) -> None:
        if not _is_torch(x):
            y[:] = self.spmv(x)
            return
        import torch

        require_device_vectors(x, y)
        if x.numel() != self.ncols or y.numel() != self.nrows:
            raise N.CfdlaError(N.ERR_DIMENSION, "vector length differs from matrix shape")
        dt = {torch.float64: N.DT_F64, torch.float32: N.DT_F32}.get(x.dtype)
        if dt is None or y.dtype != x.dtype:
            raise N.CfdlaError(N.ERR_INVALID_ARG, "x and y must be float64 or float32 tensors of one dtype")
        bind_torch_stream()
        N.call("cfdla_spmv_device", self.handle, dt, x.data_ptr(), y.data_ptr())

    def storage(self) -> dict:
        """Device storage of the matrix: format ("dia", "sell64" or
        "csr_rowblock"), stored entries (diagonals x rows for dia, padded
        slots for sell64, nnz for csr) and stored diagonals / slices / row
        blocks."""
        fmt, ent, units = C.c_int32(), C.c_int64(), C.c_int64()
        N.call("cfdla_matrix_storage", self.handle, C.byref(fmt), C.byref(ent), C.byref(units))
        return {"format": {0: "csr_rowblock", 1: "sell64", 2: "dia"}[fmt.value], "stored_entries": ent.value,
                "units": units.value}

    def varcoef_refresh_device(self, nx: int, ny: int, nz: int, inv_h2: float, k, dir_mask: int, dir_val,
                               rhs=None) -> None:
        """Rewrite the values (cfdla_problem_varcoef's pattern for this grid)
        from a float64 CUDA tensor of cell coefficients k, on the device and
        bit for bit as the host generator; rhs (float64 CUDA tensor or None)
        receives the Dirichlet terms."""
        import torch

        if not (_is_torch(k) and k.is_cuda and k.dtype == torch.float64 and k.is_contiguous()
                and k.numel() == nx * ny * nz):
            raise N.CfdlaError(N.ERR_INVALID_ARG, "k must be a contiguous float64 CUDA tensor of nx*ny*nz")
        if rhs is not None and not (_is_torch(rhs) and rhs.is_cuda and rhs.dtype == torch.float64
                                    and rhs.is_contiguous() and rhs.numel() == nx * ny * nz):
            raise N.CfdlaError(N.ERR_INVALID_ARG, "rhs must be a contiguous float64 CUDA tensor of nx*ny*nz")
        dv = np.ascontiguousarray(np.broadcast_to(np.asarray(dir_val, np.float64), (6,)))
        bind_torch_stream()
        N.call("cfdla_varcoef_refresh_device", self.handle, nx, ny, nz, inv_h2, k.data_ptr(), dir_mask, N.ptr(dv),
               None if rhs is None else rhs.data_ptr())
        self._values = None  # the device copy is newer

    STORAGE_FORMATS = {"auto": -1, "csr_rowblock": 0, "sell64": 1, "dia": 2, "csr_int32": 3}

    def set_storage(self, fmt: str) -> dict:
        """Restrict the device format ("auto", "csr_rowblock", "sell64",
        "dia", or "csr_int32": row-block CSR with 32-bit columns only); a
        format the matrix cannot take falls back to row-block CSR.  Returns
        storage().  Meant for timing one product kernel."""
        if fmt not in self.STORAGE_FORMATS:
            raise ValueError(f"unknown storage format {fmt!r}")
        N.call("cfdla_matrix_set_storage", self.handle, self.STORAGE_FORMATS[fmt])
        return self.storage()

    GATHER_LAYOUTS = {"auto": -1, "quad": 0, "rows": 1, "strided": 2, "sorted": 3}

    def set_gather(self, layout: str) -> dict:
        """Gather layout of the row-block CSR product: "auto" (chosen per
        matrix and value type at upload), "quad" (x gathered for 4
        consecutive entries per thread as they are loaded), "rows" (the
        block staged in LDS, each row's lanes gathering as they walk it),
        "strided" (staged, entries dealt round robin) or "sorted" (each
        block stored in column order, products placed back in row order;
        quad where a column reaches 2^21).  Returns gather()."""
        if layout not in self.GATHER_LAYOUTS:
            raise ValueError(f"unknown gather layout {layout!r}")
        N.call("cfdla_matrix_set_gather", self.handle, self.GATHER_LAYOUTS[layout])
        return self.gather()

    def gather(self) -> dict:
        """Gather layout in use for FP64 and FP32 products ("quad" for
        matrices not stored as row-block CSR)."""
        g64, g32 = C.c_int32(), C.c_int32()
        N.call("cfdla_matrix_gather", self.handle, C.byref(g64), C.byref(g32))
        names = {0: "quad", 1: "rows", 2: "strided", 3: "sorted"}
        return {"fp64": names[g64.value], "fp32": names[g32.value]}

    def diagonal(self):
        d = np.zeros(self.nrows)
        for i in range(self.nrows):
            s, e = self.row_ptr[i], self.row_ptr[i + 1]
            hit = np.nonzero(self.col_indices[s:e] == i)[0]
            if hit.size:
                d[i] = self.values[s + hit[0]]
        return d

    def to_dense(self):
        d = np.zeros((self.nrows, self.ncols))
        rows = np.repeat(np.arange(self.nrows), np.diff(self.row_ptr))
        np.add.at(d, (rows, self.col_indices), self.values)
        return d

    def close(self) -> None:
        if self._h is not None and N._lib is not None:
            N.lib().cfdla_matrix_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _rows_sorted(rp, ci) -> bool:
    if ci.size < 2:
        return True
    d = np.diff(ci)
    row_start = np.zeros(ci.size, dtype=bool)
    row_start[rp[:-1][rp[:-1] < ci.size]] = True
    return bool(np.all((d > 0) | row_start[1:]))


class SparseMatrixBuilder:
    """Triplet accumulator; ``build()`` sums duplicates in insertion order and
    sorts each row (crates/cfd-math/src/sparse/builder.rs:303-324)."""

    def __init__(self, rows: int, cols: int):
        self.rows, self.cols = rows, cols
        self._r: list[int] = []
        self._c: list[int] = []
        self._v: list[float] = []

    def add_entry(self, row: int, col: int, value: float) -> None:
        if not (0 <= row < self.rows and 0 <= col < self.cols):
            raise N.CfdlaError(N.ERR_INVALID_ARG,
                               f"Entry ({row}, {col}) out of bounds for {self.rows}x{self.cols} matrix")
        self._r.append(row)
        self._c.append(col)
        self._v.append(float(value))

    def add_triplets(self, triplets) -> None:
        for r, c, v in triplets:
            self.add_entry(r, c, v)

    def build_arrays(self):
        if not self._r:
            return np.zeros(self.rows + 1, np.int64), np.zeros(0, np.int64), np.zeros(0)
        r = np.asarray(self._r, np.int64)
        c = np.asarray(self._c, np.int64)
        v = np.asarray(self._v, np.float64)
        order = np.lexsort((c, r))  # stable: insertion order kept among duplicates
        r, c, v = r[order], c[order], v[order]
        key_change = np.ones(r.size, dtype=bool)
        key_change[1:] = (r[1:] != r[:-1]) | (c[1:] != c[:-1])
        starts = np.nonzero(key_change)[0]
        vals = np.empty(starts.size)
        for k, s in enumerate(starts):  # sequential sum, first value as-is
            e = starts[k + 1] if k + 1 < starts.size else r.size
            acc = v[s]
            for q in range(s + 1, e):
                acc += v[q]
            vals[k] = acc
        rows, cols = r[starts], c[starts]
        rp = np.zeros(self.rows + 1, np.int64)
        np.add.at(rp, rows + 1, 1)
        return np.cumsum(rp), cols, vals

    def build(self) -> CsrMatrix:
        rp, ci, v = self.build_arrays()
        return CsrMatrix.from_parts(v, ci, rp, self.rows, self.cols)
