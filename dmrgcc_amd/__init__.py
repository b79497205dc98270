"""MI355X-native DMRG hot path (dmrgcc drop-in), Python host binding.

Thin ctypes layer over the C ABI in ``include/dmrgcc_hip.h`` implemented by
``libdmrgcc_amd.so`` (HIP kernels for gfx950 + C++ host engine).  Device
buffers are passed as raw HIP pointers (torch CUDA tensors' ``data_ptr()``);
PyTorch is used only to own device memory.

The library must be built (``python dmrgcc_amd/build.py`` or
``__graft_entry__.build()``); importing the compute entry points without it
raises -- there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdmrgcc_amd.so")

FP64, FP32, BF16X3 = 0, 1, 2
PRECISION_NAMES = {FP64: "fp64", FP32: "fp32", BF16X3: "bf16x3"}

_lib = None

_c_int_p = ctypes.POINTER(ctypes.c_int)
_c_dbl_p = ctypes.POINTER(ctypes.c_double)
_c_i64_p = ctypes.POINTER(ctypes.c_int64)
_vp = ctypes.c_void_p

# name -> (restype, argtypes)
_SIGS = {
    "dmrgcc_last_error": (ctypes.c_char_p, []),
    "dmrgcc_version": (ctypes.c_int, []),
    "dmrgcc_rand_draws": (ctypes.c_int, [ctypes.c_uint, ctypes.c_int64, ctypes.c_int, _c_int_p]),
    "dmrgcc_random_matrices": (ctypes.c_int, [ctypes.c_uint, ctypes.c_int64, ctypes.c_int, _c_int_p,
                                              _c_int_p, ctypes.c_int, _c_dbl_p, _c_dbl_p]),
    "dmrgcc_quicksort": (ctypes.c_int, [ctypes.c_int, _c_dbl_p, _c_int_p, ctypes.c_int]),
    "dmrgcc_truncate_dims": (ctypes.c_int, [ctypes.c_int, _c_int_p, _c_dbl_p, ctypes.c_int,
                                            ctypes.c_double, _c_int_p]),
    "dmrgcc_sym_eig": (ctypes.c_int, [ctypes.c_int, _c_dbl_p, _c_dbl_p]),
    "dmrgcc_sym_eig_lowest": (ctypes.c_int, [ctypes.c_int, _c_dbl_p, _c_dbl_p, _c_dbl_p]),
    "dmrgcc_model_spinless_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                    ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                                    ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                                    ctypes.c_int, ctypes.POINTER(_vp)]),
    "dmrgcc_model_zspinless_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                      ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                                      ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                                      ctypes.c_int, ctypes.POINTER(_vp)]),
    "dmrgcc_model_spinful_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                    ctypes.c_int, ctypes.c_int, _c_dbl_p, _c_dbl_p,
                                                    _c_dbl_p, _c_dbl_p, _c_dbl_p, ctypes.c_int,
                                                    ctypes.c_int, ctypes.POINTER(_vp)]),
    "dmrgcc_model_info": (ctypes.c_int, [_vp, _c_int_p, _c_int_p, _c_int_p, _c_int_p]),
    "dmrgcc_input_model_create": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int,
                                                 ctypes.POINTER(_vp)]),
    "dmrgcc_input_string": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int,
                                           _c_int_p]),
    "dmrgcc_input_run_config": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, _c_int_p, _c_int_p, _c_int_p,
                                               _c_int_p, _c_int_p, _c_int_p, _c_dbl_p, _c_dbl_p, _c_int_p]),
    "dmrgcc_model_destroy": (ctypes.c_int, [_vp]),
    "dmrgcc_model_num_site": (ctypes.c_int, [_vp]),
    "dmrgcc_model_space_table": (ctypes.c_int, [_vp, ctypes.c_int, _c_int_p, _c_int_p]),
    "dmrgcc_model_mpo_site": (ctypes.c_int, [_vp, ctypes.c_int, _c_int_p, _c_int_p, _c_dbl_p]),
    "dmrgcc_model_mpo_table": (ctypes.c_int, [_vp, ctypes.c_int, _c_int_p, _c_int_p]),
    "dmrgcc_block_gemm": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                         ctypes.POINTER(_vp), _c_int_p, _c_int_p, _c_int_p,
                                         ctypes.c_int, _vp]),
    "dmrgcc_block_gemm_noalloc": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                                 ctypes.POINTER(_vp), _c_int_p, _c_int_p, _c_int_p,
                                                 ctypes.c_int, _vp]),
    "dmrgcc_heff_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, _c_dbl_p,
                                          ctypes.c_int, ctypes.c_int, _c_int_p,
                                          ctypes.c_int, _c_int_p, _c_int_p, _c_int_p, _c_int_p,
                                          ctypes.c_int, _c_int_p, _c_int_p, _c_int_p, _c_int_p,
                                          _c_int_p, _c_i64_p, _vp,
                                          ctypes.c_int, _c_int_p, _c_int_p, _c_int_p, _c_int_p,
                                          _c_int_p, _c_i64_p, _vp, _vp, ctypes.POINTER(_vp)]),
    "dmrgcc_zheff_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, _c_dbl_p,
                                           ctypes.c_int, ctypes.c_int, _c_int_p,
                                           ctypes.c_int, _c_int_p, _c_int_p, _c_int_p, _c_int_p,
                                           ctypes.c_int, _c_int_p, _c_int_p, _c_int_p, _c_int_p,
                                           _c_int_p, _c_i64_p, _vp,
                                           ctypes.c_int, _c_int_p, _c_int_p, _c_int_p, _c_int_p,
                                           _c_int_p, _c_i64_p, _vp, _vp, ctypes.POINTER(_vp)]),
    "dmrgcc_zblock_gemm": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                          ctypes.POINTER(_vp), _c_int_p, _c_int_p, _c_int_p,
                                          ctypes.c_int, _vp]),
    "dmrgcc_zenv_update": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                          _c_dbl_p, ctypes.c_int, ctypes.c_int, _c_int_p,
                                          ctypes.c_int, _c_int_p, _c_int_p, _c_int_p, _c_int_p, _vp,
                                          ctypes.c_int, _c_int_p, _c_int_p, _c_int_p, _c_int_p,
                                          _c_int_p, _c_i64_p, _vp, _vp, ctypes.POINTER(_vp)]),
    "dmrgcc_zsector_svd": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_vp), _c_int_p, _c_int_p,
                                          ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                          ctypes.POINTER(_vp), _vp]),
    "dmrgcc_heff_apply": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int]),
    "dmrgcc_heff_apply_async": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                               _vp]),
    "dmrgcc_alloc": (ctypes.c_int, [ctypes.c_int64, ctypes.POINTER(_vp)]),
    "dmrgcc_dmrg_set_autotune": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int64]),
    "dmrgcc_model_write_mpo": (ctypes.c_int, [_vp, ctypes.c_char_p]),
    "dmrgcc_model_read_mpo": (ctypes.c_int, [_vp, ctypes.c_char_p]),
    "dmrgcc_dmrg_write_env": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_char_p]),
    "dmrgcc_env_file_read": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.POINTER(_vp)]),
    "dmrgcc_input_get": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int,
                                        ctypes.c_int, ctypes.c_int, ctypes.c_double, _c_dbl_p, _c_int_p]),
    "dmrgcc_free": (ctypes.c_int, [_vp]),
    "dmrgcc_upload": (ctypes.c_int, [_vp, _vp, ctypes.c_int64, _vp]),
    "dmrgcc_download": (ctypes.c_int, [_vp, _vp, ctypes.c_int64, _vp]),
    "dmrgcc_lattice_canonical": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                _c_int_p, _c_int_p, ctypes.c_int, _c_int_p, _c_int_p,
                                                _c_int_p, _c_int_p, _vp, ctypes.c_int, ctypes.c_double, _vp,
                                                ctypes.POINTER(_vp), ctypes.POINTER(_vp), _c_int_p, _c_int_p,
                                                _c_dbl_p]),
    "dmrgcc_heff_info": (ctypes.c_int, [_vp, _c_dbl_p]),
    "dmrgcc_heff_destroy": (ctypes.c_int, [_vp]),
    "dmrgcc_lanczos": (ctypes.c_int, [_vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _c_dbl_p,
                                      _c_int_p]),
    "dmrgcc_blocks_count": (ctypes.c_int, [_vp]),
    "dmrgcc_blocks_get": (ctypes.c_int, [_vp, ctypes.c_int, _c_int_p, _c_int_p, _c_int_p, _c_int_p,
                                         _c_int_p, _c_i64_p]),
    "dmrgcc_blocks_data": (_vp, [_vp]),
    "dmrgcc_blocks_download": (ctypes.c_int, [_vp, _c_dbl_p]),
    "dmrgcc_blocks_destroy": (ctypes.c_int, [_vp]),
    "dmrgcc_env_update": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                         _c_dbl_p, ctypes.c_int, ctypes.c_int, _c_int_p,
                                         ctypes.c_int, _c_int_p, _c_int_p, _c_int_p, _c_int_p, _vp,
                                         ctypes.c_int, _c_int_p, _c_int_p, _c_int_p, _c_int_p,
                                         _c_int_p, _c_i64_p, _vp, _vp, ctypes.POINTER(_vp)]),
    "dmrgcc_sector_svd": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_vp), _c_int_p, _c_int_p,
                                         ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                         ctypes.POINTER(_vp), ctypes.c_int, _vp]),
    "dmrgcc_dmrg_create": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _c_int_p,
                                          _c_int_p, _c_int_p, _c_dbl_p, _c_dbl_p, ctypes.c_int,
                                          ctypes.c_int, ctypes.c_int, ctypes.POINTER(_vp)]),
    "dmrgcc_dmrg_create_files": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _c_int_p,
                                                _c_int_p, _c_int_p, _c_dbl_p, _c_dbl_p, ctypes.c_int,
                                                ctypes.c_int, ctypes.c_int, ctypes.c_char_p, ctypes.c_char_p,
                                                ctypes.POINTER(_vp)]),
    "dmrgcc_dmrg_write_space": (ctypes.c_int, [_vp, ctypes.c_char_p]),
    "dmrgcc_dmrg_heff2": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.POINTER(_vp)]),
    "dmrgcc_heff_batch_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                                ctypes.POINTER(_vp), ctypes.POINTER(_vp)]),
    "dmrgcc_heff_batch_apply": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int]),
    "dmrgcc_heff_batch_time": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _c_dbl_p]),
    "dmrgcc_heff_batch_info": (ctypes.c_int, [_vp, _c_dbl_p]),
    "dmrgcc_heff_batch_destroy": (ctypes.c_int, [_vp]),
    "dmrgcc_heff_state": (ctypes.c_int, [_vp, _vp]),
    "dmrgcc_dmrg_advance": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _c_int_p]),
    "dmrgcc_dmrg_set_precision": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int]),
    "dmrgcc_dmrg_heff": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.POINTER(_vp)]),
    "dmrgcc_dmrg_ket": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.POINTER(_vp), _c_i64_p]),
    "dmrgcc_dmrg_site_blocks": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _c_int_p, _c_i64_p,
                                               _c_int_p, _c_int_p, _c_int_p, _c_int_p, _c_int_p,
                                               _c_i64_p, _c_dbl_p]),
    "dmrgcc_dmrg_phys": (ctypes.c_int, [_vp, ctypes.c_int, _c_int_p, _c_int_p, _c_int_p]),
    "dmrgcc_dmrg_result": (ctypes.c_int, [_vp, _c_int_p, _c_dbl_p, _c_dbl_p, _c_dbl_p, _c_dbl_p]),
    "dmrgcc_dmrg_destroy": (ctypes.c_int, [_vp]),
    "dmrgcc_heff_time": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _c_dbl_p]),
    "dmrgcc_heff_tune": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _c_int_p, _c_int_p,
                                        _c_int_p, _c_dbl_p]),
    "dmrgcc_heff_config": (ctypes.c_int, [_vp, _c_int_p, _c_int_p]),
    "dmrgcc_gemm_config": (ctypes.c_int, [ctypes.c_int, ctypes.c_int]),
    "dmrgcc_dmrg_env_time": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _c_dbl_p]),
    "dmrgcc_vector_time": (ctypes.c_int, [ctypes.c_int64, ctypes.c_int, _c_dbl_p]),
    "dmrgcc_run": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _c_int_p, _c_int_p,
                                  _c_int_p, _c_dbl_p, _c_dbl_p, ctypes.c_int, ctypes.c_int,
                                  ctypes.c_int, _c_dbl_p, _c_dbl_p, _c_dbl_p, _c_dbl_p]),
    "dmrgcc_run_two_site": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _c_int_p, _c_int_p,
                                  _c_int_p, _c_dbl_p, _c_dbl_p, ctypes.c_int, ctypes.c_int,
                                  ctypes.c_int, _c_dbl_p, _c_dbl_p, _c_dbl_p, _c_dbl_p]),
    "dmrgcc_run_files": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _c_int_p, _c_int_p,
                                        _c_int_p, _c_dbl_p, _c_dbl_p, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_int, ctypes.c_char_p, ctypes.c_char_p, _c_dbl_p, _c_dbl_p,
                                        _c_dbl_p, _c_dbl_p]),
}


class DmrgError(RuntimeError):
    pass


def lib():
    """Load the native library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("libdmrgcc_amd.so is not built; run python dmrgcc_amd/build.py")
        # torch ships its own libamdhip64.so.7.  Loaded first, it is the one
        # HIP runtime of the process and the library binds to it (same
        # soname); loaded after the system copy it would start a second
        # runtime that cannot see the GPU.  So torch, when installed, loads
        # first, and device tensors and library calls share one runtime.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        l = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(l, name)
            f.restype = res
            f.argtypes = args
        _lib = l
    return _lib


def exported_symbols():
    return list(_SIGS)


def _check(rc):
    if rc != 0:
        msg = lib().dmrgcc_last_error().decode()
        raise DmrgError("dmrgcc status %d: %s" % (rc, msg))


def _ip(a):
    return a.ctypes.data_as(_c_int_p)


def _dp(a):
    return a.ctypes.data_as(_c_dbl_p)


def _lp(a):
    return a.ctypes.data_as(_c_i64_p)


def _i32(x):
    return np.ascontiguousarray(np.asarray(x, dtype=np.int32))


def _f64(x):
    return np.ascontiguousarray(np.asarray(x, dtype=np.float64))


def _i64(x):
    return np.ascontiguousarray(np.asarray(x, dtype=np.int64))


# --------------------------------------------------------------------------
# host-side restatements
# --------------------------------------------------------------------------
def rand_draws(n, seed=1, skip=0):
    out = np.zeros(n, dtype=np.int32)
    _check(lib().dmrgcc_rand_draws(seed, skip, n, _ip(out)))
    return out


def random_matrices(shapes, seed=1, skip=0, complex=False, w=None):
    """RandomMatrix for each (rows, cols) in draw order; returns a list of
    arrays (complex128 when complex).  w: per-matrix diagonal value source
    (diagonal := 0.1 * w), or None for the raw draws."""
    rows = _i32([r for r, _ in shapes])
    cols = _i32([c for _, c in shapes])
    cs = 2 if complex else 1
    out = np.zeros(int(cs * np.sum(rows.astype(np.int64) * cols)))
    wv = None
    if w is not None:
        wa = np.asarray(w, dtype=np.complex128 if complex else np.float64)
        wv = _f64(wa.view(np.float64) if complex else wa)
    _check(lib().dmrgcc_random_matrices(seed, skip, len(shapes), _ip(rows), _ip(cols), int(complex),
                                        _dp(wv) if wv is not None else None, _dp(out)))
    res, acc = [], 0
    for r, c in shapes:
        k = cs * r * c
        blk = out[acc:acc + k]
        res.append(blk.view(np.complex128).reshape(r, c) if complex else blk.reshape(r, c))
        acc += k
    return res


def quicksort(vals, index=None, flag=0):
    v = _f64(vals).copy()
    idx = None if index is None else _i32(index).copy()
    _check(lib().dmrgcc_quicksort(len(v), _dp(v), None if idx is None else _ip(idx), flag))
    return v, idx


def sym_eig(a):
    """RealSymMatrixDiag: (ascending eigenvalues, eigenvectors as columns)."""
    a = np.ascontiguousarray(a, dtype=np.float64).copy()
    n = a.shape[0]
    w = np.zeros(n)
    _check(lib().dmrgcc_sym_eig(n, _dp(a), _dp(w)))
    return w, a


def sym_eig_lowest(a):
    """Lowest eigenvalue and unit eigenvector of a dense symmetric matrix."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    n = a.shape[0]
    e0 = np.zeros(1)
    v0 = np.zeros(n)
    _check(lib().dmrgcc_sym_eig_lowest(n, _dp(a), _dp(e0), _dp(v0)))
    return float(e0[0]), v0


def truncate_dims(svals, max_dim, canpre):
    """ComputeTruncateDim: svals is a list of descending arrays (one per
    sector).  Returns (kept dims, svals after the sigma := 0 forcing)."""
    counts = _i32([len(s) for s in svals])
    flat = _f64(np.concatenate([np.asarray(s, dtype=np.float64) for s in svals]))
    trunc = np.zeros(len(svals), dtype=np.int32)
    _check(lib().dmrgcc_truncate_dims(len(svals), _ip(counts), _dp(flat), max_dim, canpre, _ip(trunc)))
    out, acc = [], 0
    for c in counts:
        out.append(flat[acc:acc + c].copy())
        acc += c
    return trunc, out


class _Model:
    """Model handle: MPS QN tables and the deparallelised MPO (real or
    complex), and the sweep engine on it."""

    def _init_info(self):
        L, d, nq, cz = (ctypes.c_int() for _ in range(4))
        _check(lib().dmrgcc_model_info(self._h, ctypes.byref(L), ctypes.byref(d), ctypes.byref(nq),
                                       ctypes.byref(cz)))
        self.L, self.d, self.nq, self.complex = L.value, d.value, nq.value, bool(cz.value)

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.dmrgcc_model_destroy(self._h)
            self._h = None

    def space_table(self, bond):
        """MPS bond QN table: (count,) for one QN per sector, else (count, nq)."""
        c = ctypes.c_int()
        _check(lib().dmrgcc_model_space_table(self._h, bond, ctypes.byref(c), None))
        q = np.zeros(c.value * self.nq, dtype=np.int32)
        _check(lib().dmrgcc_model_space_table(self._h, bond, ctypes.byref(c), _ip(q)))
        return q if self.nq == 1 else q.reshape(c.value, self.nq)

    def mpo_site(self, site):
        lb, rb = ctypes.c_int(), ctypes.c_int()
        _check(lib().dmrgcc_model_mpo_site(self._h, site, ctypes.byref(lb), ctypes.byref(rb), None))
        shape = (lb.value, rb.value, self.d, self.d)
        W = np.zeros(shape, dtype=np.complex128 if self.complex else np.float64)
        _check(lib().dmrgcc_model_mpo_site(self._h, site, ctypes.byref(lb), ctypes.byref(rb),
                                           _dp(W.view(np.float64))))
        return W

    def write_mpo(self, path):
        """RealTensorHamiltonian::WriteTensorHamiltonian of this model's MPO."""
        _check(lib().dmrgcc_model_write_mpo(self._h, os.fsencode(path)))

    def read_mpo(self, path):
        """Replace this model's MPO by the one in an MPO file
        (ReadTensorHamiltonian)."""
        _check(lib().dmrgcc_model_read_mpo(self._h, os.fsencode(path)))

    def mpo_table(self, bond):
        c = ctypes.c_int()
        _check(lib().dmrgcc_model_mpo_table(self._h, bond, ctypes.byref(c), None))
        q = np.zeros(c.value * self.nq, dtype=np.int32)
        _check(lib().dmrgcc_model_mpo_table(self._h, bond, ctypes.byref(c), _ip(q)))
        return q if self.nq == 1 else q.reshape(c.value, self.nq)

    def start(self, table, first_block=20, first_dim=10, precision=FP64, refine=True, verbose=False,
              resume=None, record=None):
        """A resumable run (DmrgRun) of this model over the sweep table."""
        return DmrgRun(self, table, first_block, first_dim, precision, refine, verbose, resume, record)

    def run(self, table, first_block=20, first_dim=10, precision=FP64, refine=True, verbose=False,
            resume=None, record=None, two_site=False):
        """RealTensorRundmrg::Run (ComplexTensorRundmrg::Run on a complex
        model).  table: rows (maxm, maxb, niter, canpre, noise_exponent) per
        half sweep.  resume: tensor-space file to start from instead of the
        random initial MPS; record: file the tensor space is written to after
        every sweep.  two_site: two-site sweeps instead (no expansion; no
        files).  Returns a dict of per-half-sweep energies, discarded
        weights, seconds and run statistics."""
        n, maxm, maxb, niter, canpre, noise = _table_args(table)
        e, t, s, st = np.zeros(n), np.zeros(n), np.zeros(n), np.zeros(24)
        if two_site:
            if resume is not None or record is not None:
                raise ValueError("two-site runs take no tensor-space files")
            _check(lib().dmrgcc_run_two_site(self._h, first_block, first_dim, n, _ip(maxm), _ip(maxb),
                                             _ip(niter), _dp(canpre), _dp(noise), precision, int(refine),
                                             int(verbose), _dp(e), _dp(t), _dp(s), _dp(st)))
            return _result_dict(e, t, s, st)
        _check(lib().dmrgcc_run_files(self._h, first_block, first_dim, n, _ip(maxm), _ip(maxb), _ip(niter),
                                      _dp(canpre), _dp(noise), precision, int(refine), int(verbose),
                                      _path(resume), _path(record), _dp(e), _dp(t), _dp(s), _dp(st)))
        return _result_dict(e, t, s, st)


class SpinlessModel(_Model):
    """RealSpinlessBoseSquareSpace + RealSpinlessBoseSquareHamiltonian, or with
    complex=True their complex twins (same MPO on complex storage, complex
    sweep engine)."""

    def __init__(self, nx, ny, num_boson, physics_dim, hop_x=-1.0, hop_y=-1.0, U=0.0, mu=0.0,
                 period_x=False, period_y=False, complex=False):
        self.nx, self.ny, self.N = nx, ny, num_boson
        h = _vp()
        create = lib().dmrgcc_model_zspinless_create if complex else lib().dmrgcc_model_spinless_create
        _check(create(nx, ny, num_boson, physics_dim, hop_x, hop_y, U, mu, int(period_x), int(period_y),
                      ctypes.byref(h)))
        self._h = h
        self._init_info()


class SpinfulModel(_Model):
    """ComplexSpinfulBoseSquareSpace + ComplexSpinfulBoseSquareHamiltonian:
    two boson species on an nx x ny lattice (MPS sites 2p = up, 2p+1 =
    down), y-bond phase exp(i 2 pi (x+1) flux).  Scalars apply to both
    species; U_uu, U_dd, U_ud are the intra/intra/inter interactions."""

    def __init__(self, nx, ny, n_up, n_down, physics_dim, flux=0.0, hop_x=-1.0, hop_y=-1.0,
                 U_uu=0.0, U_dd=0.0, U_ud=0.0, mu=0.0, period_x=False, period_y=False):
        self.nx, self.ny, self.N = nx, ny, (n_up, n_down)
        pair = lambda v: _f64(v if np.ndim(v) else [v, v])
        h = _vp()
        _check(lib().dmrgcc_model_spinful_create(nx, ny, n_up, n_down, physics_dim, _dp(pair(flux)),
                                                 _dp(pair(hop_x)), _dp(pair(hop_y)),
                                                 _dp(_f64([U_uu, U_dd, U_ud])), _dp(pair(mu)),
                                                 int(period_x), int(period_y), ctypes.byref(h)))
        self._h = h
        self._init_info()


class InputModel(_Model):
    """The model a reference driver builds from its input file (group
    "input"): driver is "real_spinless_bose_square",
    "complex_spinless_bose_square" or "complex_spinful_bose_square"."""

    def __init__(self, path, driver, echo=False):
        h = _vp()
        _check(lib().dmrgcc_input_model_create(os.fsencode(path), driver.encode(), int(echo), ctypes.byref(h)))
        self._h = h
        self._init_info()


def input_run_config(path, echo=False):
    """The run section of an input file: first_block, first_dim, the sweep
    table as rows (maxm, maxb, niter, canpre, noise exponent) per half sweep,
    and the cache flags."""
    n, fb, fd = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    _check(lib().dmrgcc_input_run_config(os.fsencode(path), int(echo), ctypes.byref(fb), ctypes.byref(fd),
                                         ctypes.byref(n), None, None, None, None, None, None))
    k = n.value
    maxm, maxb, niter = (np.zeros(k, dtype=np.int32) for _ in range(3))
    canpre, noise = np.zeros(k), np.zeros(k)
    flags = np.zeros(4, dtype=np.int32)
    _check(lib().dmrgcc_input_run_config(os.fsencode(path), 0, ctypes.byref(fb), ctypes.byref(fd),
                                         ctypes.byref(n), _ip(maxm), _ip(maxb), _ip(niter), _dp(canpre),
                                         _dp(noise), _ip(flags)))
    rows = [(int(maxm[i]), int(maxb[i]), int(niter[i]), float(canpre[i]), float(noise[i])) for i in range(k)]
    out = {"first_block": fb.value, "first_dim": fd.value, "table": rows,
           "disk_cache": bool(flags[0]), "cache_record": bool(flags[1]),
           "cache_resume": bool(flags[2]), "record_process": bool(flags[3])}
    for key in ("cache_name", "record_name", "resume_name", "process_name"):
        out[key] = input_string(path, key)
    return out


def input_value(path, key, kind="int", group="input", default=None, echo=False):
    """An int / double / bool value as the reference's InputGroup reads it
    (None when absent and no default is given)."""
    k = {"int": 0, "double": 1, "bool": 2}[kind]
    v, f = ctypes.c_double(), ctypes.c_int()
    _check(lib().dmrgcc_input_get(os.fsencode(path), group.encode(), key.encode(), k, int(echo),
                                  int(default is not None), float(default or 0), ctypes.byref(v),
                                  ctypes.byref(f)))
    if not f.value and default is None:
        return None
    return {"int": int, "double": float, "bool": bool}[kind](v.value)


def input_string(path, key):
    """String value of `key` in the file's "input" group (None if absent)."""
    buf = ctypes.create_string_buffer(4096)
    found = ctypes.c_int()
    _check(lib().dmrgcc_input_string(os.fsencode(path), key.encode(), buf, len(buf), ctypes.byref(found)))
    return buf.value.decode() if found.value else None


def run_input(path, driver, precision=FP64, refine=True, verbose=False, echo=False):
    """A reference model driver end to end: the model and the sweep table
    from the input file, then Run().  cache_resume starts from the file
    resume_name; cache_record writes cache_name/record_name after every
    sweep (the folder is created), as the reference's Run does."""
    m = InputModel(path, driver, echo)
    c = input_run_config(path, echo)
    resume = c["resume_name"] if c["cache_resume"] else None
    record = None
    if c["cache_record"]:
        os.makedirs(c["cache_name"], exist_ok=True)
        record = os.path.join(c["cache_name"], c["record_name"])
    return m.run(c["table"], c["first_block"], c["first_dim"], precision, refine, verbose, resume, record)


def _path(p):
    return None if p is None else os.fsencode(p)


def _table_args(table):
    rows = [tuple(r) for r in table]
    return (len(rows), _i32([r[0] for r in rows]), _i32([r[1] for r in rows]),
            _i32([r[2] for r in rows]), _f64([r[3] for r in rows]), _f64([r[4] for r in rows]))


def _result_dict(e, t, s, st):
    return {"energies": e, "trunc_err": t, "seconds": s, "total_seconds": st[0],
            "matvecs": int(st[1]), "dense_solves": int(st[2]), "matvec_seconds": st[3],
            "matvec_flops": st[4], "rand_draws": int(st[5]), "solves": int(st[6]),
            "max_bond": int(st[7]),
            "phase_seconds": {"lanczos": st[8], "expansion": st[9], "expand": st[10],
                              "canonical": st[11], "svd": st[12], "env": st[13],
                              "match_merge": st[14], "plan_build": st[19]},
            "svd_max_m": int(st[15]), "svd_max_n": int(st[16]), "svd_calls": int(st[17]),
            "svd_work": st[18], "autotune": {"tunes": int(st[20]), "tiling": (int(st[21]), int(st[22]))}}


class DmrgRun:
    """Resumable RealTensorRundmrg over a sweep table (see SpinlessModel.run)."""

    def __init__(self, model, table, first_block=20, first_dim=10, precision=FP64, refine=True,
                 verbose=False, resume=None, record=None):
        self.model = model
        self.nhalf, maxm, maxb, niter, canpre, noise = _table_args(table)
        h = _vp()
        _check(lib().dmrgcc_dmrg_create_files(model._h, first_block, first_dim, self.nhalf, _ip(maxm),
                                              _ip(maxb), _ip(niter), _dp(canpre), _dp(noise), precision,
                                              int(refine), int(verbose), _path(resume), _path(record),
                                              ctypes.byref(h)))
        self._h = h

    def write_space(self, path):
        """RealTensorSpace::WriteTensorSpace of the current state."""
        _check(lib().dmrgcc_dmrg_write_space(self._h, os.fsencode(path)))

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.dmrgcc_dmrg_destroy(self._h)
            self._h = None

    def advance(self, half=-1, site=-1):
        """Run until the Lanczos solve of `site` in half sweep `half` is next
        (True) or to the end of the table (False)."""
        r = ctypes.c_int()
        _check(lib().dmrgcc_dmrg_advance(self._h, half, site, ctypes.byref(r)))
        return bool(r.value)

    def set_precision(self, precision, refine=True):
        _check(lib().dmrgcc_dmrg_set_precision(self._h, precision, int(refine)))

    def set_autotune(self, on=True, min_dim=0):
        """Tile autotuner inside the sweep (see dmrgcc_dmrg_set_autotune);
        plans with psi_dim below ``min_dim`` (20000 when 0) keep their tiling."""
        _check(lib().dmrgcc_dmrg_set_autotune(self._h, int(on), int(min_dim)))

    def heff(self, site):
        h = _vp()
        _check(lib().dmrgcc_dmrg_heff(self._h, site, ctypes.byref(h)))
        return HeffPlan._from_handle(h, self)

    def env_time(self, site, leigh, reps=5):
        """GPU time of the environment update of `site` on the current state
        (leigh 1: left, from L_site; 2: right, from R_site): mean ms of the F
        launch and of the output launch, the bytes with every used block read
        or written once (and with the intermediate F), executed flops."""
        v = np.zeros(5)
        _check(lib().dmrgcc_dmrg_env_time(self._h, site, leigh, reps, _dp(v)))
        return {"f_ms": v[0], "out_ms": v[1], "ms": v[0] + v[1], "bytes_min": v[2], "bytes_exec": v[3],
                "flops": v[4]}

    def write_env(self, site, leigh, path):
        """WriteTensorContraction of the left (leigh 0) or right (1)
        environment of `site` on the current state."""
        _check(lib().dmrgcc_dmrg_write_env(self._h, site, leigh, os.fsencode(path)))

    def heff2(self, site):
        """Two-site H_eff of (site, site+1) on the current state (a HeffPlan)
        and its two-site state Theta = A_site B_site+1 as a device tensor."""
        import torch
        h = _vp()
        _check(lib().dmrgcc_dmrg_heff2(self._h, site, ctypes.byref(h)))
        plan = HeffPlan._from_handle(h, self)
        n = plan.info()["psi_dim"]
        theta = torch.empty(n * (2 if plan.complex else 1), dtype=torch.float64, device="cuda")
        _check(lib().dmrgcc_heff_state(h, _ptr(theta)))
        return plan, theta

    def ket(self, site):
        p = _vp()
        n = ctypes.c_int64()
        _check(lib().dmrgcc_dmrg_ket(self._h, site, ctypes.byref(p), ctypes.byref(n)))
        return p.value, n.value

    def site_blocks(self, site, which):
        """(blocks [(o, l, r, rows, cols, off)], host data) of the ket (0),
        left (1) or right (2) environment of `site` (complex128 data on a
        complex model; offsets count entries)."""
        c = ctypes.c_int()
        tot = ctypes.c_int64()
        _check(lib().dmrgcc_dmrg_site_blocks(self._h, site, which, ctypes.byref(c), ctypes.byref(tot),
                                             None, None, None, None, None, None, None))
        n = c.value
        o, l, r, rows, cols = (np.zeros(max(n, 1), dtype=np.int32) for _ in range(5))
        off = np.zeros(max(n, 1), dtype=np.int64)
        data = np.zeros(max(tot.value, 1), dtype=np.complex128 if self.model.complex else np.float64)
        _check(lib().dmrgcc_dmrg_site_blocks(self._h, site, which, ctypes.byref(c), ctypes.byref(tot),
                                             _ip(o), _ip(l), _ip(r), _ip(rows), _ip(cols), _lp(off),
                                             _dp(data.view(np.float64))))
        blocks = [(int(o[i]), int(l[i]), int(r[i]), int(rows[i]), int(cols[i]), int(off[i]))
                  for i in range(n)]
        return blocks, data[:tot.value]

    def phys(self, site):
        nl, nr = ctypes.c_int(), ctypes.c_int()
        _check(lib().dmrgcc_dmrg_phys(self._h, site, ctypes.byref(nl), ctypes.byref(nr), None))
        p = np.zeros(max(nl.value * nr.value, 1), dtype=np.int32)
        _check(lib().dmrgcc_dmrg_phys(self._h, site, ctypes.byref(nl), ctypes.byref(nr), _ip(p)))
        return p[:nl.value * nr.value].reshape(nl.value, nr.value)

    def result(self):
        nd = ctypes.c_int()
        e, t, s, st = np.zeros(self.nhalf), np.zeros(self.nhalf), np.zeros(self.nhalf), np.zeros(24)
        _check(lib().dmrgcc_dmrg_result(self._h, ctypes.byref(nd), _dp(e), _dp(t), _dp(s), _dp(st)))
        k = nd.value
        return _result_dict(e[:k], t[:k], s[:k], st)


# --------------------------------------------------------------------------
# device entry points
# --------------------------------------------------------------------------
def _ptr(t):
    return _vp(t.data_ptr()) if t is not None else _vp()


def _vec_ptr(t, doubles, cplx, what):
    """Device pointer of a vector the library reads or writes `doubles`
    float64 words of.  The C ABI takes no length, so a tensor that is not a
    contiguous CUDA float64 (complex128 for a complex plan) vector of at least
    that many words is refused here, before any library call."""
    import torch
    if not isinstance(t, torch.Tensor):
        raise DmrgError("%s: expected a torch tensor, got %s" % (what, type(t).__name__))
    ok = (torch.float64, torch.complex128) if cplx else (torch.float64,)
    if t.dtype not in ok:
        raise DmrgError("%s: dtype %s, expected %s" % (what, t.dtype, " or ".join(str(d) for d in ok)))
    if not t.is_cuda:
        raise DmrgError("%s: tensor is on %s, expected a CUDA device" % (what, t.device))
    if not t.is_contiguous():
        raise DmrgError("%s: tensor is not contiguous" % what)
    words = t.numel() * (2 if t.dtype == torch.complex128 else 1)
    if words < doubles:
        raise DmrgError("%s: %d float64 words, the plan needs %d" % (what, words, doubles))
    return _vp(t.data_ptr())


def _sync_torch():
    """Order the caller's pending torch work before a library call."""
    import sys
    torch = sys.modules.get("torch")
    if torch is not None and torch.cuda.is_initialized():
        torch.cuda.synchronize()


def block_gemm(As, Bs, Cs, precision=FP64, stream=None):
    """C_i = A_i @ B_i for lists of contiguous float64 (or complex128:
    ComplexMatrix::MultiplyToMatrix) device tensors."""
    n = len(As)
    arr = lambda ts: (_vp * max(n, 1))(*[_vp(t.data_ptr()) for t in ts])
    M = _i32([a.shape[0] for a in As])
    K = _i32([a.shape[1] for a in As])
    N = _i32([b.shape[1] for b in Bs])
    fn = lib().dmrgcc_zblock_gemm if n and As[0].is_complex() else lib().dmrgcc_block_gemm
    _sync_torch()
    _check(fn(n, arr(As), arr(Bs), arr(Cs), _ip(M), _ip(N), _ip(K), precision,
              _vp(stream) if stream else _vp()))


def _w_arg(W):
    """MPO site tensor as the ABI's double array: (real, float64) or
    (complex, interleaved float64 view of complex128)."""
    if np.iscomplexobj(W):
        return True, np.ascontiguousarray(W, dtype=np.complex128).view(np.float64)
    return False, _f64(W)


class BlockList:
    """Host description of a block-sparse matrix: per block (channel, l, r,
    rows, cols, offset)."""

    def __init__(self, o, l, r, rows, cols, off):
        self.o, self.l, self.r = _i32(o), _i32(l), _i32(r)
        self.rows, self.cols, self.off = _i32(rows), _i32(cols), _i64(off)

    def __len__(self):
        return len(self.l)


class HeffPlan:
    """H_eff matvec plan of one site (MultiplyEffectHamilton)."""

    def __init__(self, W, phys, ket_l, ket_r, ket_rows, ket_cols, Lb, L_data, Rb, R_data, stream=None):
        """A complex W (ComplexTensorContraction) makes a complex plan: L_data,
        R_data and the vectors are then complex128 device tensors."""
        lb, rb, d = W.shape[0], W.shape[1], W.shape[2]
        self.complex, W = _w_arg(W)
        phys = _i32(phys)
        nl, nr = phys.shape
        self._keep = (W, phys, _i32(ket_l), _i32(ket_r), _i32(ket_rows), _i32(ket_cols), Lb, Rb,
                      L_data, R_data)
        _, _, kl, kr, krows, kcols, _, _, _, _ = self._keep
        h = _vp()
        fn = lib().dmrgcc_zheff_create if self.complex else lib().dmrgcc_heff_create
        _sync_torch()
        _check(fn(
            d, lb, rb, _dp(W), nl, nr, _ip(phys), len(kl), _ip(kl), _ip(kr), _ip(krows), _ip(kcols),
            len(Lb), _ip(Lb.o), _ip(Lb.l), _ip(Lb.r), _ip(Lb.rows), _ip(Lb.cols), _lp(Lb.off),
            _ptr(L_data), len(Rb), _ip(Rb.o), _ip(Rb.l), _ip(Rb.r), _ip(Rb.rows), _ip(Rb.cols),
            _lp(Rb.off), _ptr(R_data), _vp(stream) if stream else _vp(), ctypes.byref(h)))
        self._h = h
        self._set_words()

    def _set_words(self):
        v = np.zeros(11)
        _check(lib().dmrgcc_heff_info(self._h, _dp(v)))
        self.words = int(v[0]) * (2 if self.complex else 1)

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.dmrgcc_heff_destroy(self._h)
            self._h = None

    def apply(self, x, y, precision=FP64):
        """y += H_eff x (contiguous CUDA float64 tensors of psi_dim entries,
        or complex128 / 2 psi_dim float64 for a complex plan)."""
        px = _vec_ptr(x, self.words, self.complex, "x")
        py = _vec_ptr(y, self.words, self.complex, "y")
        _sync_torch()
        _check(lib().dmrgcc_heff_apply(self._h, px, py, precision))

    def apply_async(self, x, y, alpha=1.0, beta=1.0, precision=FP64, stream=None):
        """y = alpha H_eff x + beta y, launched on `stream` (a torch.cuda.Stream
        or raw handle; default: torch's current stream) without waiting, so
        it can be captured in a CUDA graph."""
        import torch
        px = _vec_ptr(x, self.words, self.complex, "x")
        py = _vec_ptr(y, self.words, self.complex, "y")
        if stream is None:
            stream = torch.cuda.current_stream()
        raw = stream.cuda_stream if hasattr(stream, "cuda_stream") else stream
        _check(lib().dmrgcc_heff_apply_async(self._h, px, py, float(alpha), float(beta), precision,
                                             _vp(raw)))

    @classmethod
    def _from_handle(cls, h, owner):
        self = cls.__new__(cls)
        self._h = h
        self._keep = owner
        self.complex = bool(getattr(getattr(owner, "model", None), "complex", False))
        self._set_words()
        return self

    def info(self):
        v = np.zeros(11)
        _check(lib().dmrgcc_heff_info(self._h, _dp(v)))
        return {"psi_dim": int(v[0]), "flops_ref": v[1], "flops_exec": v[2], "bytes_min": v[3],
                "tasks": int(v[4]), "tiles": int(v[5]), "flops_stage1": v[6],
                "flops_stage2": v[7], "tiles_stage1": int(v[8]), "tiles_stage2": int(v[9]),
                "presplit": bool(v[10])}

    def time(self, precision=FP64, reps=20):
        """Mean milliseconds per stage-1 launch, stage-2 launch (with its
        split-K reduction), whole matvec, and each stage's GEMM kernel alone
        (HIP events on the plan's stream)."""
        ms = np.zeros(5)
        _check(lib().dmrgcc_heff_time(self._h, precision, reps, _dp(ms)))
        return {"stage1_ms": ms[0], "stage2_ms": ms[1], "matvec_ms": ms[2],
                "stage1_gemm_ms": ms[3], "stage2_gemm_ms": ms[4]}

    def tune(self, precision=BF16X3, reps=10):
        """Tile autotuner: time every (tile mode, split-K) candidate on this
        plan and keep the fastest.  Returns {"best": (tile_mode, split_k),
        "candidates": [(tile_mode, split_k, ms), ...]}."""
        cap = 16
        n = ctypes.c_int()
        tm = np.zeros(cap, dtype=np.int32)
        sk = np.zeros(cap, dtype=np.int32)
        ms = np.zeros(cap)
        _check(lib().dmrgcc_heff_tune(self._h, precision, reps, cap, ctypes.byref(n), _ip(tm), _ip(sk),
                                      _dp(ms)))
        cands = [(int(tm[i]), int(sk[i]), float(ms[i])) for i in range(n.value)]
        return {"best": self.config(), "candidates": cands}

    def config(self):
        """(tile_mode, split_k) of the plan's current tiling."""
        t, k = ctypes.c_int(), ctypes.c_int()
        _check(lib().dmrgcc_heff_config(self._h, ctypes.byref(t), ctypes.byref(k)))
        return t.value, k.value

    def lanczos(self, ket, num_iter, precision=FP64, refine=True):
        e = ctypes.c_double()
        it = ctypes.c_int()
        pk = _vec_ptr(ket, self.words, self.complex, "ket")
        _sync_torch()
        _check(lib().dmrgcc_lanczos(self._h, pk, num_iter, precision, int(refine),
                                    ctypes.byref(e), ctypes.byref(it)))
        return e.value, it.value


def vector_time(n, reps=50):
    """Mean ms of the Lanczos vector kernels on n-double vectors: the
    two-pass dot (16 n bytes), axpy (24 n) and the Lanczos step (32 n)."""
    ms = np.zeros(3)
    _check(lib().dmrgcc_vector_time(int(n), int(reps), _dp(ms)))
    return {"dot_ms": ms[0], "axpy_ms": ms[1], "lanczos_step_ms": ms[2],
            "bytes": {"dot": 16.0 * n, "axpy": 24.0 * n, "lanczos_step": 32.0 * n}}


def gemm_config(tile_mode=-1, split_k=-1):
    """Process-wide tiling (tile mode 32 or 64, split-K threshold) for GEMM
    batches built from now on, e.g. a tuned plan's config(); -1 restores the
    default of that field."""
    _check(lib().dmrgcc_gemm_config(int(tile_mode), int(split_k)))


class HeffBatch:
    """y_i += H_i x_i for many plans (independent instances) in one grouped
    launch per stage; x_i / y_i are device tensors bound at creation."""

    def __init__(self, plans, xs, ys):
        n = len(plans)
        if len(xs) != n or len(ys) != n:
            raise DmrgError("HeffBatch: one x and one y per plan")
        self._keep = (list(plans), list(xs), list(ys))
        arr = lambda v: (_vp * n)(*v)
        px = [_vec_ptr(x, p.words, p.complex, "x[%d]" % i) for i, (p, x) in enumerate(zip(plans, xs))]
        py = [_vec_ptr(y, p.words, p.complex, "y[%d]" % i) for i, (p, y) in enumerate(zip(plans, ys))]
        h = _vp()
        _sync_torch()
        _check(lib().dmrgcc_heff_batch_create(n, arr([p._h for p in plans]), arr(px), arr(py),
                                              ctypes.byref(h)))
        self._h = h

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.dmrgcc_heff_batch_destroy(self._h)
            self._h = None

    def apply(self, precision=FP64, graph=True):
        _sync_torch()
        _check(lib().dmrgcc_heff_batch_apply(self._h, precision, int(graph)))

    def time(self, precision=FP64, reps=20):
        """Mean ms of the whole batch as one graph replay, and of the same
        products plan by plan."""
        ms = np.zeros(2)
        _check(lib().dmrgcc_heff_batch_time(self._h, precision, reps, _dp(ms)))
        return {"batch_ms": ms[0], "each_ms": ms[1]}

    def info(self):
        v = np.zeros(3)
        _check(lib().dmrgcc_heff_batch_info(self._h, _dp(v)))
        return {"flops_ref": v[0], "tiles": int(v[1]), "plans": int(v[2])}


def env_update(leigh, W, phys, ket_l, ket_r, ket_rows, ket_cols, ket_data, Eb, E_data, stream=None):
    """Left (leigh 0) / right (leigh 1) environment update.  Returns
    (list of (o, l, r, rows, cols, offset), device data pointer holder)."""
    lb, rb, d = W.shape[0], W.shape[1], W.shape[2]
    cplx, W = _w_arg(W)
    phys = _i32(phys)
    nl, nr = phys.shape
    kl, kr, krows, kcols = _i32(ket_l), _i32(ket_r), _i32(ket_rows), _i32(ket_cols)
    h = _vp()
    fn = lib().dmrgcc_zenv_update if cplx else lib().dmrgcc_env_update
    _sync_torch()
    _check(fn(
        leigh, d, lb, rb, _dp(W), nl, nr, _ip(phys), len(kl), _ip(kl), _ip(kr), _ip(krows),
        _ip(kcols), _ptr(ket_data), len(Eb), _ip(Eb.o), _ip(Eb.l), _ip(Eb.r), _ip(Eb.rows),
        _ip(Eb.cols), _lp(Eb.off), _ptr(E_data), _vp(stream) if stream else _vp(), ctypes.byref(h)))
    return EnvResult(h, cplx)


class EnvResult:
    def __init__(self, h, cplx=False):
        self._h = h
        self.complex = cplx
        n = lib().dmrgcc_blocks_count(h)
        self.blocks = []
        o, l, r, rows, cols = (ctypes.c_int() for _ in range(5))
        off = ctypes.c_int64()
        for i in range(n):
            _check(lib().dmrgcc_blocks_get(h, i, ctypes.byref(o), ctypes.byref(l), ctypes.byref(r),
                                           ctypes.byref(rows), ctypes.byref(cols), ctypes.byref(off)))
            self.blocks.append((o.value, l.value, r.value, rows.value, cols.value, off.value))
        self.data_ptr = lib().dmrgcc_blocks_data(h)
        self.size = sum(b[3] * b[4] for b in self.blocks)

    def to_host(self):
        cs = 2 if self.complex else 1
        out = np.zeros(max(cs * self.size, 1))
        _check(lib().dmrgcc_blocks_download(self._h, _dp(out)))
        out = out[:cs * self.size]
        return out.view(np.complex128) if self.complex else out

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.dmrgcc_blocks_destroy(self._h)
            self._h = None


def env_file_read(path, nbond, complex=False):
    """ReadTensorContraction of an environment file with nbond channels:
    an EnvResult (blocks with their channel, device data)."""
    h = _vp()
    _check(lib().dmrgcc_env_file_read(os.fsencode(path), int(nbond), int(complex), ctypes.byref(h)))
    return EnvResult(h, complex)


def lattice_canonical(leigh, ldim, rdim, ket_l, ket_r, ket_rows, ket_cols, ket_data, max_dim, canpre):
    """Left (leigh 0) / right (leigh 1) canonical form of one site
    (Left/RightCanonicalTensorLattice).  Returns (ket EnvResult, canonical
    EnvResult, new sector dims, match dims, discarded weight)."""
    cplx = ket_data.is_complex() if hasattr(ket_data, "is_complex") else False
    ldim, rdim = _i32(ldim), _i32(rdim)
    kl, kr, krows, kcols = _i32(ket_l), _i32(ket_r), _i32(ket_rows), _i32(ket_cols)
    k, c = _vp(), _vp()
    ns = len(rdim) if leigh == 0 else len(ldim)
    dims = np.zeros(ns, dtype=np.int32)
    match = np.zeros(ns, dtype=np.int32)
    disc = ctypes.c_double()
    _sync_torch()
    _check(lib().dmrgcc_lattice_canonical(leigh, int(cplx), len(ldim), len(rdim), _ip(ldim), _ip(rdim), len(kl),
                                          _ip(kl), _ip(kr), _ip(krows), _ip(kcols), _ptr(ket_data), int(max_dim),
                                          float(canpre), _vp(), ctypes.byref(k), ctypes.byref(c), _ip(dims),
                                          _ip(match), ctypes.byref(disc)))
    return EnvResult(k, cplx), EnvResult(c, cplx), dims, match, disc.value


def sector_svd(mats, stream=None, precondition=0):
    """Batched thin SVD of device float64 matrices.  Returns lists (U, S, Vt)
    of device tensors with S descending.  precondition: 0 plain Jacobi; 1 QR
    first and Jacobi on R^T (long-side factor exactly orthonormal, short-side
    one only for singular values well above eps*|A|); 2 two QRs (short-side
    factor exactly orthonormal).  Complex matrices ignore precondition: sectors
    with min(M, N) >= 48 are QR-preconditioned (two QRs, columns presorted by
    norm) unless the environment sets DMRGCC_ZSVD_QR=0."""
    import torch
    n = len(mats)
    cz = n > 0 and mats[0].is_complex()
    dt = torch.complex128 if cz else torch.float64
    U = [torch.empty((m.shape[0], min(m.shape)), dtype=dt, device=m.device) for m in mats]
    S = [torch.empty((min(m.shape),), dtype=torch.float64, device=m.device) for m in mats]
    Vt = [torch.empty((min(m.shape), m.shape[1]), dtype=dt, device=m.device) for m in mats]
    arr = lambda ts: (_vp * max(n, 1))(*[_vp(t.data_ptr()) for t in ts])
    M = _i32([m.shape[0] for m in mats])
    N = _i32([m.shape[1] for m in mats])
    _sync_torch()
    if cz:
        _check(lib().dmrgcc_zsector_svd(n, arr(mats), _ip(M), _ip(N), arr(U), arr(S), arr(Vt),
                                        _vp(stream) if stream else _vp()))
        return U, S, Vt
    _check(lib().dmrgcc_sector_svd(n, arr(mats), _ip(M), _ip(N), arr(U), arr(S), arr(Vt),
                                   int(precondition), _vp(stream) if stream else _vp()))
    return U, S, Vt
