"""ugcore_amd: MI355X-native discretize-and-solve path for ugcore (GPU1 algebra).

Python bindings over the C ABI in include/ugcore_amd.h (the reference's
ug4py module is absent; its CMake option USE_PYBIND11 exists at
cmake/ug/pybind11.cmake:33-55). Names follow the ugcore Lua/registry API:
DomainDiscretization, GeometricMultiGrid (AssembledMultiGridCycle), CG,
BiCGStab, Jacobi, ConvCheck (StdConvCheck), SparseMatrix, Vector.

The shared library is built in-tree (ugcore_amd/libugcore_amd.so). If it is
missing or cannot be loaded the package raises; there is no CPU fallback.

HIP runtime: when torch is installed, importing this package first loads
torch's own HIP runtime (torch/lib/libamdhip64.so, the same soname as the
system runtime this library links), so that the library binds to it and one
runtime serves both in the process. Device tensors and library buffers then
share one address space and can be exchanged by pointer (ugcore_amd.adjoint,
ugcore_amd.ug4py). UGX_HIP_RUNTIME=system keeps the system runtime.
"""
import ctypes as C
import importlib.util
import os

import numpy as np


def _bind_torch_runtime():
    """load torch's HIP runtime before this library (see the module notes);
    returns its path, or None when torch or its runtime is absent"""
    if os.environ.get("UGX_HIP_RUNTIME", "").lower() == "system":
        return None
    try:
        spec = importlib.util.find_spec("torch")
    except (ImportError, ValueError):
        return None
    if spec is None or not spec.origin:
        return None
    path = os.path.join(os.path.dirname(spec.origin), "lib", "libamdhip64.so")
    if not os.path.exists(path):
        return None
    try:
        C.CDLL(path, mode=C.RTLD_GLOBAL)
    except OSError:
        return None
    return path


TORCH_HIP_RUNTIME = _bind_torch_runtime()


def hip_runtimes():
    """the HIP runtime libraries mapped into this process (one when this
    library and torch share a runtime)"""
    with open("/proc/self/maps") as f:
        return sorted({line.split()[-1] for line in f if "libamdhip64" in line})

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "libugcore_amd.so")

_lib = None

_P = C.c_void_p
_I64 = C.c_int64
_D = C.c_double
_I = C.c_int

_SIGNATURES = {
    "ugx_init": (_I, []),
    "ugx_last_error": (C.c_char_p, []),
    "ugx_stream": (_P, []),
    "ugx_synchronize": (_I, []),
    "ugx_malloc": (_P, [_I64]),
    "ugx_free": (_I, [_P]),
    "ugx_memcpy_htod": (_I, [_P, _P, _I64]),
    "ugx_memcpy_dtoh": (_I, [_P, _P, _I64]),
    "ugx_memcpy_dtod": (_I, [_P, _P, _I64]),
    "ugx_event_create": (_P, []),
    "ugx_event_destroy": (_I, [_P]),
    "ugx_event_record": (_I, [_P, _P]),
    "ugx_event_elapsed_ms": (_I, [_P, _P, _P]),
    "ugx_spmv_csr_f64": (_I, [_I64, _P, _P, _P, _D, _P, _D, _P, _P, _P]),
    "ugx_spmv_csr_f32": (_I, [_I64, _P, _P, _P, _D, _P, _D, _P, _P, _P]),
    "ugx_spmv_csr_ignore_zero_rows_f64": (_I, [_I64, _P, _P, _P, _D, _P, _P, _P]),
    "ugx_vec_scale_add2_f64": (_I, [_I64, _P, _D, _P, _D, _P, _P]),
    "ugx_vec_scale_add3_f64": (_I, [_I64, _P, _D, _P, _D, _P, _D, _P, _P]),
    "ugx_vec_set_f64": (_I, [_I64, _P, _D, _P]),
    "ugx_dot_f64": (_I, [_I64, _P, _P, _P, _P]),
    "ugx_norm_f64": (_I, [_I64, _P, _P, _P]),
    "ugx_dot_f64_noalloc": (_I, [_I64, _P, _P, _P, _P]),
    "ugx_jacobi_diag_inv_f64": (_I, [_I64, _P, _P, _P, _D, _P, _P]),
    "ugx_jacobi_apply_f64": (_I, [_I64, _P, _P, _P, _P]),
    "ugx_hierarchy_numbering": (_I, [_I, _I, _I, _P, _P]),
    "ugx_hierarchy_tets": (_I, [_I, _I, _P, _P]),
    "ugx_disc_set_velocity": (_I, [_P, _D, _D, _D]),
    "ugx_disc_set_jitter": (_I, [_P, C.c_uint64, _D]),
    "ugx_disc_adjoint_operator": (_I, [_P, _P, _P]),
    "ugx_gmg_set_transposed": (_I, [_P, _I]),
    "ugx_gmg_set_matrix_free": (_I, [_P, _I]),
    "ugx_gmg_set_rap": (_I, [_P, _I]),
    "ugx_gmg_set_graphs": (_I, [_P, _I]),
    "ugx_gmg_set_fused_smoothing": (_I, [_P, _I]),
    "ugx_solver_set_restart": (_I, [_P, _I]),
    "ugx_solver_set_min_orthogonality": (_I, [_P, _D]),
    "ugx_disc_create": (_P, [_I, _I]),
    "ugx_disc_create_ordered": (_P, [_I, _I, _I]),
    "ugx_disc_create_setup": (_P, [_I, _I, _I, _I]),
    "ugx_disc_level_tets": (_I, [_P, _I, _P, _P]),
    "ugx_disc_num_inner": (_I64, [_P, _I]),
    "ugx_disc_destroy": (None, [_P]),
    "ugx_disc_num_levels": (_I, [_P]),
    "ugx_disc_num_dofs": (_I64, [_P, _I]),
    "ugx_disc_dof_of_lattice": (_I, [_P, _I, _P]),
    "ugx_disc_dof_positions": (_I, [_P, _I, _P]),
    "ugx_cv_write_matrix": (_I, [C.c_char_p, _I64, _P, _P, _P, _P, _I]),
    "ugx_cv_write_vector": (_I, [C.c_char_p, _I64, _P, _P, _I]),
    "ugx_vector_file_write": (_I, [C.c_char_p, _I64, _P]),
    "ugx_vector_file_length": (_I64, [C.c_char_p]),
    "ugx_vector_file_read": (_I, [C.c_char_p, _I64, _P]),
    "ugx_vector_save": (_I, [_P, C.c_char_p]),
    "ugx_vector_read": (_I, [_P, C.c_char_p]),
    "ugx_disc_set_theta": (_I, [_P, _P]),
    "ugx_disc_set_theta_device": (_I, [_P, _P]),
    "ugx_disc_adjust_defect": (_I, [_P, _P]),
    "ugx_disc_adjoint_gradient": (_I, [_P, _P, _P, _P]),
    "ugx_q2_hex_element_matrices": (_I, [_I64, _P, _P, _D, _P, _P]),
    "ugx_q2_assemble_lattice": (_I, [_I, _P, _D, _P, _P]),
    "ugx_disc_set_source_const": (_I, [_P, _D]),
    "ugx_disc_set_source_dofs": (_I, [_P, _P]),
    "ugx_disc_assemble_linear": (_I, [_P, _P, _P]),
    "ugx_disc_assemble_seconds": (_D, [_P]),
    "ugx_disc_assemble_jacobian": (_I, [_P, _P, _P]),
    "ugx_disc_assemble_defect": (_I, [_P, _P, _P]),
    "ugx_elemdisc_create": (_P, [_I, _I64, _P, _I64, _P]),
    "ugx_elemdisc_destroy": (None, [_P]),
    "ugx_elemdisc_nnz": (_I64, [_P]),
    "ugx_elemdisc_num_colours": (_I, [_P]),
    "ugx_elemdisc_set_diffusion": (_I, [_P, _P]),
    "ugx_elemdisc_set_diffusion_at_points": (_I, [_P, _P]),
    "ugx_elemdisc_set_source_const": (_I, [_P, _D]),
    "ugx_elemdisc_set_source_dofs": (_I, [_P, _P]),
    "ugx_elemdisc_set_velocity": (_I, [_P, _D, _D, _D]),
    "ugx_elemdisc_set_dirichlet": (_I, [_P, _P, _I64]),
    "ugx_elemdisc_assemble_linear": (_I, [_P, _P, _P]),
    "ugx_elemdisc_assemble_jacobian": (_I, [_P, _P, _P]),
    "ugx_elemdisc_assemble_defect": (_I, [_P, _P, _P]),
    "ugx_matrix_create": (_P, []),
    "ugx_matrix_destroy": (None, [_P]),
    "ugx_matrix_num_rows": (_I64, [_P]),
    "ugx_matrix_nnz": (_I64, [_P]),
    "ugx_matrix_upload_csr": (_I, [_P, _I64, _I64, _P, _P, _P]),
    "ugx_matrix_download_csr": (_I, [_P, _P, _P, _P]),
    "ugx_matrix_device_csr": (_I, [_P, _P, _P, _P]),
    "ugx_matrix_axpy": (_I, [_P, _P, _D, _P, _D, _P]),
    "ugx_matrix_set_spmv_lanes": (_I, [_P, _I]),
    "ugx_matrix_set_spmv_layout": (_I, [_P, _I]),
    "ugx_matrix_spmv_layout": (_I, [_P]),
    "ugx_matrix_set_row_order": (_I, [_P, _P, _I64]),
    "ugx_matrix_set_slice_order": (_I, [_P, _P, _I64]),
    "ugx_matrix_has_row_order": (_I, [_P]),
    "ugx_matrix_finalize": (_I, [_P]),
    "ugx_matrix_apply_sub": (_I, [_P, _P, _P]),
    "ugx_matrix_num_cols": (_I64, [_P]),
    "ugx_matrix_device_csr_readonly": (_I, [_P, _P, _P, _P]),
    "ugx_matrix_apply_ignore_zero_rows": (_I, [_P, _P, _D, _P]),
    "ugx_matrix_resize_and_clear": (_I, [_P, _I64, _I64]),
    "ugx_matrix_resize_and_keep_values": (_I, [_P, _I64, _I64]),
    "ugx_matrix_clear_retain_structure": (_I, [_P]),
    "ugx_matrix_is_building": (_I, [_P]),
    "ugx_matrix_entry": (C.POINTER(C.c_double), [_P, _I64, _I64]),
    "ugx_matrix_get": (_I, [_P, _I64, _I64, _P]),
    "ugx_matrix_row_length": (_I64, [_P, _I64]),
    "ugx_matrix_row": (_I, [_P, _I64, _P, _P]),
    "ugx_matrix_set_as_transpose_of": (_I, [_P, _P, _D, _I]),
    "ugx_matrix_set_as_copy_of": (_I, [_P, _P, _D]),
    "ugx_matrix_scale": (_I, [_P, _D]),
    "ugx_matrix_axpy_transposed": (_I, [_P, _P, _D, _P, _D, _P]),
    "ugx_matrix_set_dirichlet_rows": (_I, [_P, _P, _I64]),
    "ugx_vector_create": (_P, [_I64]),
    "ugx_vector_destroy": (None, [_P]),
    "ugx_vector_size": (_I64, [_P]),
    "ugx_vector_device_ptr": (_P, [_P]),
    "ugx_vector_upload": (_I, [_P, _P]),
    "ugx_vector_download": (_I, [_P, _P]),
    "ugx_vector_set": (_I, [_P, _D]),
    "ugx_gmg_create": (_P, [_P, _I, _I, _I, _I, _D, _I]),
    "ugx_gmg_destroy": (None, [_P]),
    "ugx_gmg_init": (_I, [_P, _P]),
    "ugx_gmg_apply": (_I, [_P, _P, _P]),
    "ugx_gmg_smooth_level": (_I, [_P, _I, _I]),
    "ugx_gmg_surface_apply_lattice": (_I, [_P, _P, _P]),
    "ugx_gmg_surface_direction_lattice": (_I, [_P, _P, _P, _P, _P, _P, _P, _P]),
    "ugx_gmg_apply_lattice": (_I, [_P, _P, _P, _P]),
    "ugx_gmg_lattice_uses_f32_defect": (_I, [_P]),
    "ugx_gmg_lattice_halo": (_I64, [_P]),
    "ugx_gmg_level_num_diags": (_I, [_P, _I]),
    "ugx_gmg_level_dia": (_I, [_P, _I, _P]),
    "ugx_jacobi_create": (_P, [_D]),
    "ugx_jacobi_destroy": (None, [_P]),
    "ugx_solver_create": (_P, [_I]),
    "ugx_solver_destroy": (None, [_P]),
    "ugx_solver_set_preconditioner": (_I, [_P, _I, _P]),
    "ugx_q2disc_create": (_P, [_I]),
    "ugx_q2disc_destroy": (None, [_P]),
    "ugx_q2disc_num_levels": (_I, [_P]),
    "ugx_q2disc_num_dofs": (_I64, [_P, _I]),
    "ugx_q2disc_dof_positions": (_I, [_P, _I, _P]),
    "ugx_q2disc_dirichlet": (_I, [_P, _I, _P]),
    "ugx_q2disc_set_theta": (_I, [_P, _P]),
    "ugx_q2disc_set_source": (_I, [_P, _D]),
    "ugx_q2disc_assemble_linear": (_I, [_P, _I, _P, _P]),
    "ugx_q2disc_prolongation": (_I, [_P, _I, _P]),
    "ugx_q2disc_restriction": (_I, [_P, _I, _P, _P]),
    "ugx_mgmg_create": (_P, [_I, _P, _P, _P, _I, _I, _I, _I, _D]),
    "ugx_mgmg_destroy": (None, [_P]),
    "ugx_mgmg_init": (_I, [_P, _P]),
    "ugx_mgmg_apply": (_I, [_P, _P, _P]),
    "ugx_solver_set_conv_check": (_I, [_P, _I, _D, _D, _I]),
    "ugx_solver_init": (_I, [_P, _P]),
    "ugx_solver_apply": (_I, [_P, _P, _P]),
    "ugx_solver_apply_return_defect": (_I, [_P, _P, _P]),
    "ugx_solver_steps": (_I, [_P]),
    "ugx_solver_defects": (_I, [_P, _P, _I]),
    "ugx_solver_set_lattice_path": (_I, [_P, _I]),
    "ugx_solver_used_lattice_path": (_I, [_P]),
}


class UGError(RuntimeError):
    pass


def lib():
    """Load libugcore_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise UGError(f"{LIB_PATH} not found: build it with `python -m ugcore_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(rc, what):
    if rc is None or (isinstance(rc, int) and rc < 0):
        raise UGError(f"{what}: {lib().ugx_last_error().decode()}")
    return rc


def _handle(h, what):
    if not h:
        raise UGError(f"{what}: {lib().ugx_last_error().decode()}")
    return h


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def init():
    _check(lib().ugx_init(), "ugx_init")


def synchronize():
    _check(lib().ugx_synchronize(), "ugx_synchronize")


class DeviceArray:
    """A raw device buffer owned by the library (for the kernel entry points)."""

    def __init__(self, host):
        self.host = np.ascontiguousarray(host)
        self.nbytes = max(self.host.nbytes, 8)
        self.ptr = _handle(lib().ugx_malloc(self.nbytes), "ugx_malloc")
        if self.host.nbytes:
            _check(lib().ugx_memcpy_htod(self.ptr, _ptr(self.host), self.host.nbytes), "ugx_memcpy_htod")

    def numpy(self):
        out = np.empty_like(self.host)
        if out.nbytes:
            _check(lib().ugx_memcpy_dtoh(_ptr(out), self.ptr, out.nbytes), "ugx_memcpy_dtoh")
        return out

    def __del__(self):
        if getattr(self, "ptr", None) and _lib is not None:
            _lib.ugx_free(self.ptr)
            self.ptr = None


def time_launches(fn, reps, stream=None):
    """Average seconds per call of fn (which enqueues work on the library
    stream), measured with HIP events recorded on that stream."""
    L = lib()
    e0 = _handle(L.ugx_event_create(), "event")
    e1 = _handle(L.ugx_event_create(), "event")
    try:
        fn()
        synchronize()
        _check(L.ugx_event_record(e0, stream), "event_record")
        for _ in range(reps):
            fn()
        _check(L.ugx_event_record(e1, stream), "event_record")
        ms = C.c_double()
        _check(L.ugx_event_elapsed_ms(e0, e1, C.byref(ms)), "event_elapsed")
        return ms.value * 1e-3 / reps
    finally:
        L.ugx_event_destroy(e0)
        L.ugx_event_destroy(e1)


def hierarchy_numbering(kind, num_refs, level):
    """LEVEL DoF numbering (lattice point -> DoF) computed on the host; returns
    (dof_of_lattice, n_inner)."""
    dim = 2 if kind == 1 else 3
    W = (3 if kind == 2 else 1) * 2 ** level + 1
    out = np.empty(W ** dim, np.int32)
    ninner = C.c_int64(0)
    _check(lib().ugx_hierarchy_numbering(kind, num_refs, level, _ptr(out), C.byref(ninner)), "hierarchy_numbering")
    return out, ninner.value


def hierarchy_tets(num_refs, level):
    """Tetrahedra of level `level` of the kind-2 hierarchy (host): per lattice
    cell 6 packed corner codes (creation order), and per lattice vertex the
    parent-edge direction code (None on level 0)."""
    n = 3 * 2 ** level
    cells = np.empty(n ** 3 * 6, np.uint16)
    par = np.empty((n + 1) ** 3, np.uint8) if level > 0 else None
    _check(lib().ugx_hierarchy_tets(num_refs, level, _ptr(cells), _ptr(par) if par is not None else None),
           "hierarchy_tets")
    return cells.reshape(-1, 6), par


class Vector:
    """GPUVector (vector_type of the GPU1 algebra), FP64 device storage."""

    def __init__(self, n=None, values=None):
        if values is not None:
            values = np.ascontiguousarray(values, np.float64)
            n = values.size
        self.h = _handle(lib().ugx_vector_create(int(n)), "Vector")
        if values is not None:
            self.upload(values)

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.ugx_vector_destroy(self.h)
            self.h = None

    def size(self):
        return lib().ugx_vector_size(self.h)

    def upload(self, a):
        a = np.ascontiguousarray(a, np.float64)
        assert a.size == self.size()
        _check(lib().ugx_vector_upload(self.h, _ptr(a)), "Vector.upload")

    def numpy(self):
        out = np.empty(self.size(), np.float64)
        _check(lib().ugx_vector_download(self.h, _ptr(out)), "Vector.download")
        return out

    def set(self, v):
        _check(lib().ugx_vector_set(self.h, float(v)), "Vector.set")

    def device_ptr(self):
        return lib().ugx_vector_device_ptr(self.h)


class SparseMatrix:
    """GPUSparseMatrix (matrix_type of the GPU1 algebra): device CSR, int32 indices."""

    def __init__(self):
        self.h = _handle(lib().ugx_matrix_create(), "SparseMatrix")

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.ugx_matrix_destroy(self.h)
            self.h = None

    def num_rows(self):
        return lib().ugx_matrix_num_rows(self.h)

    def nnz(self):
        return lib().ugx_matrix_nnz(self.h)

    def upload_csr(self, rowptr, col, val, n_cols=None):
        rowptr = np.ascontiguousarray(rowptr, np.int32)
        col = np.ascontiguousarray(col, np.int32)
        val = np.ascontiguousarray(val, np.float64)
        rows = rowptr.size - 1
        _check(lib().ugx_matrix_upload_csr(self.h, rows, n_cols if n_cols is not None else rows, _ptr(rowptr),
                                           _ptr(col), _ptr(val)), "SparseMatrix.upload_csr")

    def download_csr(self):
        n = self.num_rows()
        nnz = self.nnz()
        rp = np.empty(n + 1, np.int32)
        col = np.empty(nnz, np.int32)
        val = np.empty(nnz, np.float64)
        _check(lib().ugx_matrix_download_csr(self.h, _ptr(rp), _ptr(col), _ptr(val)), "SparseMatrix.download")
        return rp, col, val

    def device_csr(self):
        rp, col, val = C.c_void_p(), C.c_void_p(), C.c_void_p()
        _check(lib().ugx_matrix_device_csr(self.h, C.byref(rp), C.byref(col), C.byref(val)),
               "SparseMatrix.device_csr")
        return rp.value, col.value, val.value

    # ---- the SparseMatrix surface (sparsematrix.h / sparsematrix_impl.h) ----
    def num_cols(self):
        return lib().ugx_matrix_num_cols(self.h)

    def resize_and_clear(self, rows, cols):
        """start a host-side pattern build (SparseMatrix::resize_and_clear)"""
        _check(lib().ugx_matrix_resize_and_clear(self.h, int(rows), int(cols)), "SparseMatrix.resize_and_clear")

    def resize_and_keep_values(self, rows, cols):
        _check(lib().ugx_matrix_resize_and_keep_values(self.h, int(rows), int(cols)),
               "SparseMatrix.resize_and_keep_values")

    def clear_retain_structure(self):
        _check(lib().ugx_matrix_clear_retain_structure(self.h), "SparseMatrix.clear_retain_structure")

    def building(self):
        return bool(lib().ugx_matrix_is_building(self.h))

    def _entry(self, r, c):
        p = lib().ugx_matrix_entry(self.h, int(r), int(c))
        if not p:
            raise UGError(f"SparseMatrix({r}, {c}): {lib().ugx_last_error().decode()}")
        return p

    def __setitem__(self, rc, value):
        """A[r, c] = v: operator()(r, c) = v (creates the entry; host build only)"""
        self._entry(*rc)[0] = float(value)

    def add(self, r, c, value):
        """A(r, c) += v (creates the entry with 0 first; host build only)"""
        p = self._entry(r, c)
        p[0] = p[0] + float(value)

    def __getitem__(self, rc):
        """const operator()(r, c): the stored value or 0"""
        v = C.c_double()
        _check(lib().ugx_matrix_get(self.h, int(rc[0]), int(rc[1]), C.byref(v)), "SparseMatrix.get")
        return v.value

    def has_connection(self, r, c):
        v = C.c_double()
        return _check(lib().ugx_matrix_get(self.h, int(r), int(c), C.byref(v)), "SparseMatrix.get") == 1

    def row(self, r):
        """(columns, values) of row r (begin_row/end_row)"""
        n = lib().ugx_matrix_row_length(self.h, int(r))
        if n < 0:
            raise UGError(f"SparseMatrix.row: {lib().ugx_last_error().decode()}")
        cols = np.empty(n, np.int32)
        vals = np.empty(n, np.float64)
        _check(lib().ugx_matrix_row(self.h, int(r), _ptr(cols), _ptr(vals)), "SparseMatrix.row")
        return cols, vals

    def finalize(self):
        """end a host build (the rows become the device CSR), or rebuild the
        SpMV layout after raw edits through device_csr()"""
        _check(lib().ugx_matrix_finalize(self.h), "SparseMatrix.finalize")

    def set_as_transpose_of(self, B, scale=1.0):
        """scale * B^T keeping explicit zeros (set_as_transpose_of)"""
        _check(lib().ugx_matrix_set_as_transpose_of(self.h, B.h, float(scale), 1), "SparseMatrix.set_as_transpose_of")

    def set_as_transpose_of2(self, B, scale=1.0):
        """scale * B^T without zero entries (set_as_transpose_of2)"""
        _check(lib().ugx_matrix_set_as_transpose_of(self.h, B.h, float(scale), 0), "SparseMatrix.set_as_transpose_of2")

    def set_as_copy_of(self, B, scale=1.0):
        _check(lib().ugx_matrix_set_as_copy_of(self.h, B.h, float(scale)), "SparseMatrix.set_as_copy_of")

    def scale(self, a):
        _check(lib().ugx_matrix_scale(self.h, float(a)), "SparseMatrix.scale")

    def axpy_transposed(self, dest, alpha, v1, beta, w):
        """dest = alpha*v1 + beta*A^T*w"""
        _check(lib().ugx_matrix_axpy_transposed(self.h, dest.h, float(alpha), v1.h, float(beta), w.h),
               "SparseMatrix.axpy_transposed")

    def apply_transposed(self, dest, w):
        self.axpy_transposed(dest, 0.0, dest, 1.0, w)

    def set_dirichlet_rows(self, rows):
        """SetDirichletRow for every listed row"""
        r = np.ascontiguousarray(rows, np.int32)
        _check(lib().ugx_matrix_set_dirichlet_rows(self.h, _ptr(r), r.size), "SparseMatrix.set_dirichlet_rows")

    def axpy(self, dest, alpha, v1, beta, w):
        _check(lib().ugx_matrix_axpy(self.h, dest.h, float(alpha), v1.h, float(beta), w.h), "SparseMatrix.axpy")

    def apply(self, dest, w):
        self.axpy(dest, 0.0, dest, 1.0, w)

    def apply_sub(self, d, x):
        _check(lib().ugx_matrix_apply_sub(self.h, d.h, x.h), "SparseMatrix.apply_sub")

    _LAYOUTS = {"csr": 0, "sell": 1, "dict": 2}

    def set_spmv_layout(self, layout):
        """'dict' (sliced ELL with a column dictionary per slice where the
        matrix has a slice order, else sliced ELL; default), 'sell' (sliced
        ELL) or 'csr'"""
        if layout not in self._LAYOUTS:
            raise ValueError(f"layout must be one of {sorted(self._LAYOUTS)}")
        _check(lib().ugx_matrix_set_spmv_layout(self.h, self._LAYOUTS[layout]), "set_spmv_layout")

    def spmv_layout(self):
        """the layout SpMV runs on: 'csr', 'sell' or 'dict'"""
        return {0: "csr", 1: "sell", 2: "dict"}[lib().ugx_matrix_spmv_layout(self.h)]

    def set_row_order(self, order):
        """Row permutation that fills the sliced-ELL slices (None clears it);
        SpMV results do not depend on it."""
        if order is None:
            _check(lib().ugx_matrix_set_row_order(self.h, None, 0), "set_row_order")
            return
        o = np.ascontiguousarray(order, dtype=np.int32)
        _check(lib().ugx_matrix_set_row_order(self.h, o.ctypes.data, o.size), "set_row_order")

    def set_slice_order(self, order):
        """Slices of the dictionary layout: slice t holds rows
        order[64t .. 64t+63] (-1: empty slot), every row once; None clears it.
        SpMV uses the layout when every slice touches at most 255 distinct
        columns (and no row of a slice is longer than 64)."""
        if order is None:
            _check(lib().ugx_matrix_set_slice_order(self.h, None, 0), "set_slice_order")
            return
        o = np.ascontiguousarray(order, dtype=np.int32)
        _check(lib().ugx_matrix_set_slice_order(self.h, o.ctypes.data, o.size), "set_slice_order")

    def has_row_order(self):
        return bool(lib().ugx_matrix_has_row_order(self.h))

    def set_spmv_lanes(self, lanes):
        _check(lib().ugx_matrix_set_spmv_lanes(self.h, int(lanes)), "SparseMatrix.set_spmv_lanes")


class DomainDiscretization:
    """Structured domain with device assembly. kind 0: unit cube from one
    hexahedron, Q1 FE; kind 1: unit square from two triangles, FV1; kind 2:
    unit cube as 3x3x3 hexahedra x 6 tetrahedra, FV1 convection-diffusion with
    full upwind (theta per fine lattice cell)."""

    def __init__(self, kind, num_refs, ordering="reference", setup="device"):
        self.kind = kind
        self.num_refs = num_refs
        self.dim = 2 if kind == 1 else 3
        # fine elements (kind 0/1) or fine lattice cells (kind 2) per side
        self.n_fine = (3 if kind == 2 else 1) << num_refs
        o = {"reference": 0, "lex": 1}[ordering]
        # "device": hierarchy and numbering refined on the device; "host":
        # built on the host and uploaded (the same numbering)
        st = {"device": 0, "host": 1}[setup]
        self.h = _handle(lib().ugx_disc_create_setup(int(kind), int(num_refs), o, st), "DomainDiscretization")

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.ugx_disc_destroy(self.h)
            self.h = None

    def num_levels(self):
        return lib().ugx_disc_num_levels(self.h)

    def num_dofs(self, level=None):
        if level is None:
            level = self.num_levels() - 1
        return lib().ugx_disc_num_dofs(self.h, level)

    def num_inner(self, level=None):
        """DoFs [0, num_inner) of the level are subset Inner, the rest Dirichlet"""
        level = self.num_refs if level is None else level
        return lib().ugx_disc_num_inner(self.h, level)

    def lattice_width(self, level):
        return (3 if self.kind == 2 else 1) * 2 ** level + 1

    def dof_positions(self, level=None):
        """coordinates of the level's DoFs in DoF order, shape (N, dim)"""
        level = self.num_refs if level is None else level
        n = lib().ugx_disc_num_dofs(self.h, level)
        out = np.empty((n, self.dim), np.float64)
        _check(lib().ugx_disc_dof_positions(self.h, level, _ptr(out)), "dof_positions")
        return out

    def level_tets(self, level):
        """tetrahedral domain: (cell_tets (n^3, 6) uint16, parent codes
        ((n+1)^3,) uint8 or None on level 0) of a level"""
        n = 3 * 2 ** level
        cells = np.empty((n ** 3, 6), np.uint16)
        par = np.empty((n + 1) ** 3, np.uint8) if level > 0 else None
        _check(lib().ugx_disc_level_tets(self.h, level, _ptr(cells), _ptr(par) if par is not None else None),
               "level_tets")
        return cells, par

    def dof_of_lattice(self, level):
        n = self.lattice_width(level) ** self.dim
        out = np.empty(n, np.int32)
        _check(lib().ugx_disc_dof_of_lattice(self.h, level, _ptr(out)), "dof_of_lattice")
        return out

    def set_theta(self, theta):
        theta = np.ascontiguousarray(theta, np.float64)
        if theta.size != self.n_fine ** self.dim:
            raise ValueError("theta needs one value per fine element / cell")
        _check(lib().ugx_disc_set_theta(self.h, _ptr(theta)), "set_theta")

    def set_velocity(self, v):
        _check(lib().ugx_disc_set_velocity(self.h, float(v[0]), float(v[1]), float(v[2])), "set_velocity")

    def set_jitter(self, seed, amp):
        _check(lib().ugx_disc_set_jitter(self.h, int(seed), float(amp)), "set_jitter")

    def adjoint_operator(self, A):
        """matrix of the adjoint solve: (A, Dirichlet-eliminated)^T"""
        AT = SparseMatrix()
        _check(lib().ugx_disc_adjoint_operator(self.h, A.h, AT.h), "adjoint_operator")
        return AT

    def set_source(self, f):
        if np.isscalar(f):
            _check(lib().ugx_disc_set_source_const(self.h, float(f)), "set_source")
        else:
            f = np.ascontiguousarray(f, np.float64)
            _check(lib().ugx_disc_set_source_dofs(self.h, _ptr(f)), "set_source")

    def assemble_linear(self, A, b):
        _check(lib().ugx_disc_assemble_linear(self.h, A.h, b.h), "assemble_linear")

    def assemble_jacobian(self, J, u):
        """J = dd/du at u (the discretisations are linear: the stiffness matrix)"""
        _check(lib().ugx_disc_assemble_jacobian(self.h, J.h, u.h), "assemble_jacobian")

    def assemble_defect(self, d, u):
        """d = sum_e (A_e u_e - f_e), Dirichlet entries 0 (element loop)"""
        _check(lib().ugx_disc_assemble_defect(self.h, d.h, u.h), "assemble_defect")

    def assemble_seconds(self):
        return lib().ugx_disc_assemble_seconds(self.h)

    def set_theta_device(self, ptr):
        _check(lib().ugx_disc_set_theta_device(self.h, ptr), "set_theta_device")

    def adjust_defect(self, v):
        _check(lib().ugx_disc_adjust_defect(self.h, v.h), "adjust_defect")

    def adjoint_gradient(self, u, lam, grad_ptr=None):
        """g_e = -lambda_e^T (dA_e/dtheta_e) u_e per fine element; returns a
        numpy array unless a device pointer for the result is given."""
        ne = self.n_fine ** 3
        if grad_ptr is not None:
            _check(lib().ugx_disc_adjoint_gradient(self.h, u.h, lam.h, grad_ptr), "adjoint_gradient")
            return None
        buf = DeviceArray(np.zeros(ne))
        _check(lib().ugx_disc_adjoint_gradient(self.h, u.h, lam.h, buf.ptr), "adjoint_gradient")
        return buf.numpy()


_CYCLES = {"V": 1, "W": 2, "F": -1}
_PRECISIONS = {"f64": 0, "float64": 0, 0: 0, "f32": 1, "float32": 1, 1: 1}


class GeometricMultiGrid:
    """AssembledMultiGridCycle: damped Jacobi smoother, P1 transfer, dense
    FP64 LU on the base level. cycle: 1 or "V", 2 or "W", gamma > 2, -1 or
    "F" (set_cycle_type, mg_solver.h:155-163); anything else raises
    ValueError. precision 'f64' or 'f32' for the level operators; rap=True
    builds the coarse levels as Galerkin products R A P (set_rap)."""

    def __init__(self, disc, base_level=0, nu1=2, nu2=2, cycle=1, omega=0.66, precision="f64", transposed=False,
                 matrix_free=True, rap=False, graphs=True, fused_smoothing=True):
        self.disc = disc
        if isinstance(cycle, str):
            if cycle.strip() not in _CYCLES:
                raise ValueError(f"GeometricMultiGrid: cycle '{cycle}' not supported (V, W or F)")
            cycle = _CYCLES[cycle.strip()]
        if precision not in _PRECISIONS:
            raise ValueError(f"GeometricMultiGrid: precision {precision!r} not supported (f64 or f32)")
        p = _PRECISIONS[precision]
        self.h = _handle(lib().ugx_gmg_create(disc.h, base_level, nu1, nu2, cycle, float(omega), p),
                         "GeometricMultiGrid")
        if transposed:
            _check(lib().ugx_gmg_set_transposed(self.h, 1), "GMG.set_transposed")
        if not matrix_free:
            _check(lib().ugx_gmg_set_matrix_free(self.h, 0), "GMG.set_matrix_free")
        if rap:
            _check(lib().ugx_gmg_set_rap(self.h, 1), "GMG.set_rap")
        if not graphs:
            _check(lib().ugx_gmg_set_graphs(self.h, 0), "GMG.set_graphs")
        if not fused_smoothing:
            _check(lib().ugx_gmg_set_fused_smoothing(self.h, 0), "GMG.set_fused_smoothing")

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.ugx_gmg_destroy(self.h)
            self.h = None

    def init(self, A):
        _check(lib().ugx_gmg_init(self.h, A.h), "GMG.init")

    def apply(self, c, d):
        _check(lib().ugx_gmg_apply(self.h, c.h, d.h), "GMG.apply")

    def smooth_level(self, level, nsteps=1):
        _check(lib().ugx_gmg_smooth_level(self.h, level, nsteps), "GMG.smooth_level")

    def level_dia(self, level):
        nd = lib().ugx_gmg_level_num_diags(self.h, level)
        n = self.disc.lattice_width(level) ** self.disc.dim
        out = np.empty(nd * n, np.float64)
        _check(lib().ugx_gmg_level_dia(self.h, level, _ptr(out)), "GMG.level_dia")
        return out.reshape(nd, n)


class Jacobi:
    def __init__(self, omega=1.0):
        self.h = _handle(lib().ugx_jacobi_create(float(omega)), "Jacobi")

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.ugx_jacobi_destroy(self.h)
            self.h = None


class Q2Discretization:
    """Q2 Lagrange on the unit cube (one hexahedron refined num_refs times):
    every level in the reference's LEVEL DoF order, operators by the element
    loop, StdTransfer's prolongations and restrictions (ugx_q2disc_*)."""

    def __init__(self, num_refs):
        self.num_refs = num_refs
        self.h = _handle(lib().ugx_q2disc_create(int(num_refs)), "Q2Discretization")

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.ugx_q2disc_destroy(self.h)
            self.h = None

    def num_levels(self):
        return lib().ugx_q2disc_num_levels(self.h)

    def num_dofs(self, level):
        return lib().ugx_q2disc_num_dofs(self.h, level)

    def dof_positions(self, level):
        out = np.empty((self.num_dofs(level), 3), np.float64)
        _check(lib().ugx_q2disc_dof_positions(self.h, level, _ptr(out)), "Q2Discretization.dof_positions")
        return out

    def dirichlet(self, level):
        out = np.empty(self.num_dofs(level), np.uint8)
        _check(lib().ugx_q2disc_dirichlet(self.h, level, _ptr(out)), "Q2Discretization.dirichlet")
        return out

    def set_theta(self, theta):
        if theta is None:
            _check(lib().ugx_q2disc_set_theta(self.h, None), "Q2Discretization.set_theta")
            return
        t = np.ascontiguousarray(theta, np.float64)
        if t.size != 8 ** self.num_refs:
            raise ValueError(f"theta needs {8 ** self.num_refs} values (one per top-level element)")
        _check(lib().ugx_q2disc_set_theta(self.h, _ptr(t)), "Q2Discretization.set_theta")

    def set_source(self, f):
        _check(lib().ugx_q2disc_set_source(self.h, float(f)), "Q2Discretization.set_source")

    def assemble_linear(self, level, A, b):
        _check(lib().ugx_q2disc_assemble_linear(self.h, level, A.h, b.h), "Q2Discretization.assemble_linear")

    def prolongation(self, level, P):
        _check(lib().ugx_q2disc_prolongation(self.h, level, P.h), "Q2Discretization.prolongation")

    def restriction(self, level, P, R):
        _check(lib().ugx_q2disc_restriction(self.h, level, P.h, R.h), "Q2Discretization.restriction")

    def hierarchy(self):
        """level operators, prolongations and restrictions of every level:
        (A, P, R, b_top) with P[0] = R[0] = None"""
        A, P, R = [], [None], [None]
        b = None
        for lev in range(self.num_levels()):
            a = SparseMatrix()
            bv = Vector(self.num_dofs(lev))
            self.assemble_linear(lev, a, bv)
            A.append(a)
            b = bv
            if lev > 0:
                p, r = SparseMatrix(), SparseMatrix()
                self.prolongation(lev, p)
                self.restriction(lev, p, r)
                P.append(p)
                R.append(r)
        return A, P, R, b


class MatrixMultiGrid:
    """AssembledMultiGridCycle over level operators and transfers given as
    matrices (ugx_mgmg_*): A[l] per level, P[l] and R[l] for l >= 1."""

    def __init__(self, A, P, R, base_level=0, nu1=2, nu2=2, cycle=1, omega=0.66):
        n = len(A)
        if len(P) != n or len(R) != n:
            raise ValueError("one A, P and R entry per level")
        arr = lambda ms: (C.c_void_p * n)(*[m.h if m is not None else None for m in ms])
        self._keep = (A, P, R)
        self.h = _handle(lib().ugx_mgmg_create(n, arr(A), arr(P), arr(R), int(base_level), int(nu1), int(nu2),
                                               int(cycle), float(omega)), "MatrixMultiGrid")

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.ugx_mgmg_destroy(self.h)
            self.h = None

    def init(self, A):
        _check(lib().ugx_mgmg_init(self.h, A.h), "MatrixMultiGrid.init")

    def apply(self, c, d):
        _check(lib().ugx_mgmg_apply(self.h, c.h, d.h), "MatrixMultiGrid.apply")


class _Solver:
    _type = 0

    def __init__(self, precond=None, max_steps=100, min_defect=1e-50, reduction=1e-10, verbose=False):
        self.h = _handle(lib().ugx_solver_create(self._type), type(self).__name__)
        self._precond = None
        self.set_conv_check(max_steps, min_defect, reduction, verbose)
        if precond is not None:
            self.set_preconditioner(precond)

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.ugx_solver_destroy(self.h)
            self.h = None

    def set_preconditioner(self, p):
        kind = (1 if isinstance(p, GeometricMultiGrid) else 2 if isinstance(p, Jacobi)
                else 3 if isinstance(p, MatrixMultiGrid) else 0)
        self._precond = p
        _check(lib().ugx_solver_set_preconditioner(self.h, kind, p.h if p is not None else None), "set_preconditioner")

    def set_conv_check(self, max_steps=100, min_defect=1e-50, reduction=1e-10, verbose=False):
        _check(lib().ugx_solver_set_conv_check(self.h, int(max_steps), float(min_defect), float(reduction),
                                               1 if verbose else 0), "set_conv_check")

    def init(self, A):
        self._A = A
        _check(lib().ugx_solver_init(self.h, A.h), "solver.init")

    def apply(self, x, b):
        rc = _check(lib().ugx_solver_apply(self.h, x.h, b.h), "solver.apply")
        return rc == 1

    def apply_return_defect(self, x, b):
        """solves from the initial x and leaves the final defect b - A x in b
        (linear_operator_inverse.h:188)"""
        rc = _check(lib().ugx_solver_apply_return_defect(self.h, x.h, b.h), "solver.apply_return_defect")
        return rc == 1

    def steps(self):
        return lib().ugx_solver_steps(self.h)

    def defects(self):
        n = lib().ugx_solver_defects(self.h, None, 0)
        out = np.empty(n, np.float64)
        lib().ugx_solver_defects(self.h, _ptr(out), n)
        return out


class CG(_Solver):
    _type = 0

    def set_lattice_path(self, on):
        _check(lib().ugx_solver_set_lattice_path(self.h, 1 if on else 0), "set_lattice_path")

    def used_lattice_path(self):
        return lib().ugx_solver_used_lattice_path(self.h) == 1


class BiCGStab(_Solver):
    _type = 1

    def set_restart(self, n):
        """restart every n steps (0: never), BiCGStab::set_restart"""
        _check(lib().ugx_solver_set_restart(self.h, int(n)), "set_restart")

    def set_min_orthogonality(self, m):
        """restart when |(r0, r)| <= m ||r|| ||r0||"""
        _check(lib().ugx_solver_set_min_orthogonality(self.h, float(m)), "set_min_orthogonality")


# ---- file formats ------------------------------------------------------------------------------

def _path(filename):
    return os.fsencode(os.fspath(filename))


def SaveToFile(v, filename):
    """binary vector file of a serial build (restart_bridge.cpp:146-158)"""
    _check(lib().ugx_vector_save(v.h, _path(filename)), "SaveToFile")


def ReadFromFile(v, filename):
    """reads a SaveToFile vector into v (resized to the stored length)"""
    _check(lib().ugx_vector_read(v.h, _path(filename)), "ReadFromFile")


def write_vector_file(filename, values):
    """host-array form of SaveToFile"""
    a = np.ascontiguousarray(values, np.float64)
    _check(lib().ugx_vector_file_write(_path(filename), a.size, _ptr(a)), "write_vector_file")


def read_vector_file(filename):
    """host-array form of ReadFromFile"""
    n = lib().ugx_vector_file_length(_path(filename))
    if n < 0:
        raise UGError(f"read_vector_file: {lib().ugx_last_error().decode()}")
    out = np.empty(n, np.float64)
    _check(lib().ugx_vector_file_read(_path(filename), n, _ptr(out)), "read_vector_file")
    return out


def write_connection_viewer_matrix(filename, rowptr, col, val, positions):
    """ConnectionViewer::WriteMatrix (connection_viewer_output.h:74-141) for a
    host CSR and one position per row (shape (n, dim))"""
    rowptr = np.ascontiguousarray(rowptr, np.int32)
    col = np.ascontiguousarray(col, np.int32)
    val = np.ascontiguousarray(val, np.float64)
    pos = np.ascontiguousarray(positions, np.float64)
    n = rowptr.size - 1
    if pos.shape[0] != n:
        raise UGError("write_connection_viewer_matrix: one position per row is needed")
    _check(lib().ugx_cv_write_matrix(_path(filename), n, _ptr(rowptr), _ptr(col), _ptr(val), _ptr(pos),
                                     pos.shape[1]), "write_connection_viewer_matrix")


def write_connection_viewer_vector(filename, values, positions):
    """ConnectionViewer::WriteVector (connection_viewer_output.h:402-422)"""
    a = np.ascontiguousarray(values, np.float64)
    pos = np.ascontiguousarray(positions, np.float64)
    if pos.shape[0] != a.size:
        raise UGError("write_connection_viewer_vector: one position per value is needed")
    _check(lib().ugx_cv_write_vector(_path(filename), a.size, _ptr(a), _ptr(pos), pos.shape[1]),
           "write_connection_viewer_vector")


def SaveMatrixForConnectionViewer(disc, A, filename):
    """the surface matrix with the top-level DoF positions
    (SaveMatrixForConnectionViewer, grid_function_util.h)"""
    rp, col, val = A.download_csr()
    write_connection_viewer_matrix(filename, rp, col, val, disc.dof_positions())


def SaveVectorForConnectionViewer(disc, v, filename):
    """a surface vector with the top-level DoF positions
    (SaveVectorForConnectionViewer, grid_function_util.h)"""
    write_connection_viewer_vector(filename, v.numpy(), disc.dof_positions())


class ElementDiscretization:
    """Element-loop assembly from caller-supplied connectivity
    (ugx_elemdisc_*): elem_type "triangle_fv1", "tetrahedron_fv1" or
    "hexahedron_q1"; elem_dofs [num_elems, nv] in the reference element's local
    vertex order; dof_coords [num_dofs, dim]. The pattern, the slot of every
    local entry and the element colouring are built on the device here."""

    TYPES = {"triangle_fv1": (0, 3, 2), "tetrahedron_fv1": (1, 4, 3), "hexahedron_q1": (2, 8, 3),
             "hexahedron_q2": (3, 27, 3)}

    def __init__(self, elem_type, elem_dofs, dof_coords):
        if elem_type not in self.TYPES:
            raise ValueError(f"elem_type must be one of {sorted(self.TYPES)}")
        t, nv, dim = self.TYPES[elem_type]
        E = np.ascontiguousarray(elem_dofs, np.int32)
        X = np.ascontiguousarray(dof_coords, np.float64)
        if E.ndim != 2 or E.shape[1] != nv:
            raise ValueError(f"elem_dofs must have shape [num_elems, {nv}]")
        if X.ndim != 2 or X.shape[1] != dim:
            raise ValueError(f"dof_coords must have shape [num_dofs, {dim}]")
        self.num_elems, self.num_dofs = E.shape[0], X.shape[0]
        self.h = _handle(lib().ugx_elemdisc_create(t, E.shape[0], _ptr(E), X.shape[0], _ptr(X)),
                         "ElementDiscretization")

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.ugx_elemdisc_destroy(self.h)
            self.h = None

    def nnz(self):
        return lib().ugx_elemdisc_nnz(self.h)

    def num_colours(self):
        return lib().ugx_elemdisc_num_colours(self.h)

    def set_diffusion(self, k):
        if k is None:
            _check(lib().ugx_elemdisc_set_diffusion(self.h, None), "set_diffusion")
            return
        k = np.ascontiguousarray(k, np.float64)
        if k.size != self.num_elems:
            raise ValueError(f"one diffusion value per element ({self.num_elems}) expected, got {k.size}")
        _check(lib().ugx_elemdisc_set_diffusion(self.h, _ptr(k)), "set_diffusion")

    def set_diffusion_at_points(self, kq):
        """Q2: k at the 3x3x3 Gauss points of every element [num_elems, 27] (None: per element)"""
        if kq is None:
            _check(lib().ugx_elemdisc_set_diffusion_at_points(self.h, None), "set_diffusion_at_points")
            return
        kq = np.ascontiguousarray(kq, np.float64)
        if kq.size != 27 * self.num_elems:
            raise ValueError(f"27 values per element ({27 * self.num_elems}) expected, got {kq.size}")
        _check(lib().ugx_elemdisc_set_diffusion_at_points(self.h, _ptr(kq)), "set_diffusion_at_points")

    def set_source(self, f):
        if np.isscalar(f):
            _check(lib().ugx_elemdisc_set_source_const(self.h, float(f)), "set_source")
        else:
            f = np.ascontiguousarray(f, np.float64)
            if f.size != self.num_dofs:
                raise ValueError(f"one source value per DoF ({self.num_dofs}) expected, got {f.size}")
            _check(lib().ugx_elemdisc_set_source_dofs(self.h, _ptr(f)), "set_source")

    def set_velocity(self, v):
        _check(lib().ugx_elemdisc_set_velocity(self.h, float(v[0]), float(v[1]), float(v[2])), "set_velocity")

    def set_dirichlet(self, dofs):
        d = np.ascontiguousarray(dofs, np.int32)
        _check(lib().ugx_elemdisc_set_dirichlet(self.h, _ptr(d), d.size), "set_dirichlet")

    def assemble_linear(self, A, b):
        _check(lib().ugx_elemdisc_assemble_linear(self.h, A.h, b.h), "assemble_linear")

    def assemble_jacobian(self, J, u):
        _check(lib().ugx_elemdisc_assemble_jacobian(self.h, J.h, u.h), "assemble_jacobian")

    def assemble_defect(self, d, u):
        _check(lib().ugx_elemdisc_assemble_defect(self.h, d.h, u.h), "assemble_defect")


def q2_hex_element_matrices(corners, k=None, f_const=0.0):
    """Q2 hexahedron element matrices on the device (ugx_q2_hex_element_matrices):
    corners [n, 8, 3] in the reference hexahedron order, k [n] or None (1);
    returns (Ke [n, 27, 27], fe [n, 27]) as numpy arrays"""
    X = np.ascontiguousarray(corners, dtype=np.float64)
    n = X.shape[0]
    if X.shape[1:] != (8, 3):
        raise ValueError("corners must have shape [n, 8, 3]")
    if k is not None:
        k = np.ascontiguousarray(k, dtype=np.float64)
        if k.size != n:
            raise ValueError(f"k must hold one coefficient per element ({n}), got {k.size}")
    dx = DeviceArray(X)
    dk = DeviceArray(k) if k is not None else None
    ke = DeviceArray(np.zeros((n, 27, 27)))
    fe = DeviceArray(np.zeros((n, 27)))
    _check(lib().ugx_q2_hex_element_matrices(n, dx.ptr, dk.ptr if dk is not None else None, float(f_const),
                                             ke.ptr, fe.ptr), "q2_hex_element_matrices")
    synchronize()
    return ke.numpy(), fe.numpy()


# largest n with nnz = (8n+1)^3 < 2^31 (int32 column indices)
Q2_LATTICE_MAX_N = 161


def q2_assemble_lattice(n, theta=None, f_const=1.0):
    """Q2 assembly of -div(exp(theta) grad u) = f on the unit cube of n^3
    hexahedra with u = 0 on the boundary (ugx_q2_assemble_lattice); returns
    (SparseMatrix, Vector) on the device, nodes of the (2n+1)^3 lattice
    numbered x fastest"""
    n = int(n)
    if not 1 <= n <= Q2_LATTICE_MAX_N:
        raise ValueError(f"n must be in 1..{Q2_LATTICE_MAX_N} (the pattern's nnz must fit int32)")
    if theta is not None:
        theta = np.ascontiguousarray(theta, dtype=np.float64)
        if theta.size != n ** 3:
            raise ValueError(f"theta must hold one value per element ({n ** 3}), got {theta.size}")
    A = SparseMatrix()
    b = Vector(1)
    th = DeviceArray(theta) if theta is not None else None
    _check(lib().ugx_q2_assemble_lattice(int(n), th.ptr if th is not None else None, float(f_const), A.h, b.h),
           "q2_assemble_lattice")
    return A, b
