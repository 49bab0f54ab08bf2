"""hacapk_hmvm_amd -- MI355X (gfx950) H-matrix-vector multiply.

Python mirror of the C ABI in include/hacapk.h, include/hgen.h and
include/hmvm_hip.h.  All compute runs in the native library
``hacapk_hmvm_amd/libhmvm_hip.so`` (HIP kernels); this module only marshals
arguments.  Importing it fails loudly when the library has not been built:
there is no Python or CPU fallback for the HMVM.

Names follow the reference benchmark (src_hmvm/src_c/hacapk.h): ``matrix`` /
``matrix2`` structs, ``loadHmatrix`` / ``loadHmatrix2`` loaders and the
``hmvm_hip_1/1t/2/2t`` kernels with the argument order of ``hmvm_seq_*``
(v, mat, b) and the same accumulate semantics (v += A b).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libhmvm_hip.so")

c_int_p = ctypes.POINTER(ctypes.c_int)
c_double_p = ctypes.POINTER(ctypes.c_double)


class submatrix(ctypes.Structure):
    """src_c/hacapk.h:5-8"""
    _fields_ = [("ltmtx", ctypes.c_int), ("kt", ctypes.c_int), ("ndl", ctypes.c_int), ("ndt", ctypes.c_int),
                ("nstrtl", ctypes.c_int), ("nstrtt", ctypes.c_int),
                ("a1", c_double_p), ("a2", c_double_p), ("a2t", c_double_p)]


class matrix(ctypes.Structure):
    """src_c/hacapk.h:9-17"""
    _fields_ = [("nd", ctypes.c_int), ("nlf", ctypes.c_int), ("ktmax", ctypes.c_int), ("nleavs", c_int_p),
                ("submat", ctypes.POINTER(submatrix)),
                ("b1", ctypes.c_int), ("b2", ctypes.c_int), ("b3", ctypes.c_int),
                ("e1", ctypes.c_int), ("e2", ctypes.c_int), ("e3", ctypes.c_int)]


class matrix2(ctypes.Structure):
    """src_c/hacapk.h:19-28"""
    _fields_ = [("nd", ctypes.c_int), ("nlf", ctypes.c_int), ("ktmax", ctypes.c_int), ("len", ctypes.c_int),
                ("nleavs", c_int_p),
                ("ltmtx", c_int_p), ("kt", c_int_p), ("ndl", c_int_p), ("ndt", c_int_p),
                ("nstrtl", c_int_p), ("nstrtt", c_int_p), ("a1", c_int_p), ("a2", c_int_p),
                ("rowmat", c_double_p), ("rowmat_t", c_double_p)]


class matrix2_64(ctypes.Structure):
    """int64 extension of matrix2 (include/hacapk.h)."""
    _fields_ = [("nd", ctypes.c_int), ("nlf", ctypes.c_int), ("ktmax", ctypes.c_int), ("len", ctypes.c_longlong),
                ("ltmtx", c_int_p), ("kt", c_int_p), ("ndl", c_int_p), ("ndt", c_int_p),
                ("nstrtl", c_int_p), ("nstrtt", c_int_p),
                ("a1", ctypes.POINTER(ctypes.c_longlong)), ("a2", ctypes.POINTER(ctypes.c_longlong)),
                ("rowmat", c_double_p), ("rowmat_t", c_double_p)]


class hgen_params(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int), ("geometry", ctypes.c_int), ("leaf", ctypes.c_int), ("kmax", ctypes.c_int),
                ("eta", ctypes.c_double), ("eps", ctypes.c_double), ("aca_eps_factor", ctypes.c_double),
                ("jitter", ctypes.c_double), ("seed", ctypes.c_ulonglong)]


class hmvm_options(ctypes.Structure):
    _fields_ = [("storage", ctypes.c_int), ("w1", ctypes.c_int), ("w2", ctypes.c_int),
                ("block_threads", ctypes.c_int), ("kernel", ctypes.c_int), ("order", ctypes.c_int),
                ("mode", ctypes.c_int), ("t_mode", ctypes.c_int), ("max_rows", ctypes.c_int),
                ("max_rows_b", ctypes.c_int), ("w3", ctypes.c_int)]


class hmvm_panel(ctypes.Structure):
    _fields_ = [("val_off", ctypes.c_longlong), ("ld", ctypes.c_int), ("ncols", ctypes.c_int),
                ("vec_off", ctypes.c_int), ("vec_sel", ctypes.c_int), ("factor", ctypes.c_int),
                ("vec_chunks", ctypes.c_int)]


class hmvm_tile(ctypes.Structure):
    _fields_ = [("nrows", ctypes.c_int), ("panel0", ctypes.c_int), ("npanel", ctypes.c_int), ("out_off", ctypes.c_int),
                ("map_off", ctypes.c_int), ("flags", ctypes.c_int), ("ncols", ctypes.c_int), ("map_off_multi", ctypes.c_int),
                ("first", hmvm_panel)]


class hmvm_repack(ctypes.Structure):
    _fields_ = [("src_off", ctypes.c_longlong), ("dst_off", ctypes.c_longlong), ("rows", ctypes.c_int),
                ("cols", ctypes.c_int), ("src_rs", ctypes.c_int), ("src_cs", ctypes.c_int), ("dst_cs", ctypes.c_int),
                ("factor", ctypes.c_int)]


class hmvm_segsum(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int), ("nseg", ctypes.c_int), ("seg", ctypes.POINTER(ctypes.c_int)),
                ("ptr", ctypes.POINTER(ctypes.c_int)), ("list", ctypes.POINTER(ctypes.c_longlong))]


class hmvm_schedule(ctypes.Structure):
    _fields_ = [("ntiles_a", ctypes.c_int), ("ntiles_b", ctypes.c_int), ("npanels", ctypes.c_int),
                ("nrepack", ctypes.c_int), ("t_fused", ctypes.c_int),
                ("nvals", ctypes.c_longlong), ("n_t", ctypes.c_longlong), ("n_tslots", ctypes.c_longlong),
                ("n_zslots", ctypes.c_longlong),
                ("tiles_a", ctypes.POINTER(hmvm_tile)), ("tiles_b", ctypes.POINTER(hmvm_tile)),
                ("panels", ctypes.POINTER(hmvm_panel)), ("repack", ctypes.POINTER(hmvm_repack)),
                ("tsum", hmvm_segsum), ("ysum", hmvm_segsum),
                ("t_direct", ctypes.c_int), ("ntmap", ctypes.c_longlong), ("tmap", ctypes.POINTER(ctypes.c_int))]


class hmvm_plan_info(ctypes.Structure):
    _fields_ = [("nd", ctypes.c_int), ("nlf", ctypes.c_int), ("nlowrank", ctypes.c_int), ("ndense", ctypes.c_int),
                ("len", ctypes.c_longlong), ("device_elems", ctypes.c_longlong),
                ("storage", ctypes.c_int), ("elem_bytes", ctypes.c_int),
                ("ntiles_a", ctypes.c_int), ("ntiles_b", ctypes.c_int), ("npanels", ctypes.c_int),
                ("ncol_groups", ctypes.c_int), ("nrow_groups", ctypes.c_int), ("t_fused", ctypes.c_int),
                ("n_t", ctypes.c_longlong), ("n_tslots", ctypes.c_longlong), ("n_zslots", ctypes.c_longlong),
                ("elems_a", ctypes.c_longlong), ("elems_b", ctypes.c_longlong),
                ("elems_stage1", ctypes.c_longlong), ("elems_dense", ctypes.c_longlong),
                ("elems_stage2", ctypes.c_longlong),
                ("vec_reads_a", ctypes.c_longlong), ("vec_reads_b", ctypes.c_longlong),
                ("algorithmic_bytes", ctypes.c_longlong), ("device_bytes", ctypes.c_longlong),
                ("mode", ctypes.c_int), ("t_direct", ctypes.c_int)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class hmvm_timing(ctypes.Structure):
    _fields_ = [("iters", ctypes.c_int), ("ms_avg", ctypes.c_double), ("ms_min", ctypes.c_double),
                ("ms_median", ctypes.c_double), ("ms_a", ctypes.c_double), ("ms_tsum", ctypes.c_double),
                ("ms_b", ctypes.c_double), ("ms_ysum", ctypes.c_double), ("ms_stream", ctypes.c_double)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class hmvm_krylov_result(ctypes.Structure):
    _fields_ = [("iters", ctypes.c_int), ("hmvms", ctypes.c_int), ("relres", ctypes.c_double),
                ("ms_loop", ctypes.c_double), ("ms_per_iter", ctypes.c_double), ("graph_used", ctypes.c_int)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


HGEN_SPHERE, HGEN_CUBE = 0, 1
HMVM_STORAGE_F64, HMVM_STORAGE_F32, HMVM_STORAGE_BF16 = 0, 1, 2
HMVM_A2_ROWMAJOR, HMVM_A2_TRANSPOSED = 0, 1
STORAGE = {"f64": 0, "f32": 1, "bf16": 2}
HACAPK_WITH_ROWMAT_T = 1
HMVM_NFEAT = 18
HMVM_TILE_ROW_PER_LANE = 1


HMVM_ERR_ARG, HMVM_ERR_MATRIX, HMVM_ERR_HIP, HMVM_ERR_NOMEM = -101, -102, -103, -104
HMVM_ERR_NODEVICE, HMVM_ERR_UNSUPPORTED, HMVM_ERR_BREAKDOWN, HMVM_ERR_WAIT = -105, -106, -107, -108


class HMVMError(RuntimeError):
    pass


class KrylovBreakdown(HMVMError):
    """BiCGSTAB met a zero divisor or a non-finite scalar; ``u`` and ``result``
    hold the state at the stop."""

    def __init__(self, msg, u, result):
        super().__init__(msg)
        self.u = u
        self.result = result


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            "hacapk_hmvm_amd: native library %s is missing; build it with "
            "`python hacapk_hmvm_amd/build.py` (there is no CPU fallback)" % LIB_PATH)
    lib = ctypes.CDLL(LIB_PATH)
    P = ctypes.POINTER
    sig = {
        "loadHmatrix": (ctypes.c_int, [ctypes.c_char_p, P(matrix), P(matrix2)]),
        "loadHmatrix2": (ctypes.c_int, [ctypes.c_char_p, P(matrix2)]),
        "dummymatrix": (ctypes.c_int, [P(matrix), P(matrix2), ctypes.c_int, ctypes.c_int, ctypes.c_int]),
        "freeHmatrix": (None, [P(matrix), P(matrix2)]),
        "writeHmatrix1": (ctypes.c_int, [ctypes.c_char_p, P(matrix2)]),
        "writeHmatrix2": (ctypes.c_int, [ctypes.c_char_p, P(matrix2)]),
        "hacapk_strerror": (ctypes.c_char_p, [ctypes.c_int]),
        "hgen_default_params": (None, [P(hgen_params), ctypes.c_int]),
        "hgen_generate": (ctypes.c_int, [P(hgen_params), P(matrix2)]),
        "hgen_matrix_from_matrix2": (ctypes.c_int, [P(matrix2), P(matrix)]),
        "hgen_points": (ctypes.c_int, [P(hgen_params), c_double_p, c_double_p]),
        "hgen_kernel_block": (ctypes.c_int, [P(hgen_params), ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                             ctypes.c_int, c_double_p]),
        "hgen_accuracy": (ctypes.c_double, [P(hgen_params), P(matrix2), ctypes.c_int]),
        "hgen_generate64": (ctypes.c_int, [P(hgen_params), P(matrix2_64), ctypes.c_int]),
        "hgen_accuracy64": (ctypes.c_double, [P(hgen_params), P(matrix2_64), ctypes.c_int]),
        "loadHmatrix_64": (ctypes.c_int, [ctypes.c_char_p, P(matrix2_64), ctypes.c_int]),
        "writeHmatrix2_64": (ctypes.c_int, [ctypes.c_char_p, P(matrix2_64)]),
        "writeHmatrix1_64": (ctypes.c_int, [ctypes.c_char_p, P(matrix2_64)]),
        "hacapk_widen_matrix2": (ctypes.c_int, [P(matrix2), P(matrix2_64)]),
        "freeHmatrix64": (None, [P(matrix2_64)]),
        "hacapk_scan_file": (ctypes.c_int, [ctypes.c_char_p, P(matrix2_64), P(ctypes.c_longlong), c_int_p]),
        "hmvm_plan_create64": (ctypes.c_int, [P(matrix2_64), ctypes.c_int, P(hmvm_options), P(ctypes.c_void_p)]),
        "hmvm_plan_create_file": (ctypes.c_int, [ctypes.c_char_p, P(hmvm_options), P(ctypes.c_void_p)]),
        "hmvm_plan_create_batch": (ctypes.c_int, [P(P(matrix2)), ctypes.c_int, ctypes.c_int, P(hmvm_options),
                                                  P(ctypes.c_void_p), c_int_p]),
        "hmvm_hip_2_64": (None, [c_double_p, P(matrix2_64), c_double_p]),
        "hmvm_hip_2t_64": (None, [c_double_p, P(matrix2_64), c_double_p]),
        "hmvm_plan_graph_create": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                                  ctypes.c_int, P(ctypes.c_void_p)]),
        "hmvm_graph_launch": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
        "hmvm_graph_destroy": (None, [ctypes.c_void_p]),
        "hmvm_plan_time_graph": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                P(hmvm_timing)]),
        "hmvm_plan_set_phase_launch": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
        "hmvm_plan_tile_features": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                                   ctypes.POINTER(ctypes.c_float)]),
        "hmvm_plan_profile_tiles": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, c_double_p,
                                                   c_double_p, c_double_p]),
        "hmvm_plan_set_tile_order": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, c_int_p]),
        "hmvm_plan_set_tile_flags": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, c_int_p]),
        "hmvm_plan_get_tile_flags": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, c_int_p]),
        "hmvm_tile_occupancy": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, c_int_p]),
        "hmvm_schedule_export_batch": (ctypes.c_int, [P(P(matrix2)), ctypes.c_int, ctypes.c_int, P(hmvm_options),
                                                      P(hmvm_schedule)]),
        "hmvm_schedule_export64": (ctypes.c_int, [P(matrix2_64), ctypes.c_int, P(hmvm_options), P(hmvm_schedule)]),
        "hmvm_plan_reset_launch": (ctypes.c_int, [ctypes.c_void_p]),
        "hmvm_mlp_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_ulonglong, P(ctypes.c_void_p)]),
        "hmvm_mlp_destroy": (None, [ctypes.c_void_p]),
        "hmvm_mlp_num_params": (ctypes.c_int, [ctypes.c_void_p]),
        "hmvm_mlp_get_params": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float),
                                               ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)]),
        "hmvm_mlp_set_params": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float),
                                               ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)]),
        "hmvm_mlp_fit_scaler": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float), ctypes.c_int]),
        "hmvm_mlp_predict": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float), ctypes.c_int,
                                            ctypes.POINTER(ctypes.c_float)]),
        "hmvm_mlp_loss_grad": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float),
                                              ctypes.POINTER(ctypes.c_float), ctypes.c_int,
                                              ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)]),
        "hmvm_mlp_train": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float),
                                          ctypes.POINTER(ctypes.c_float), ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                          ctypes.c_float, ctypes.c_ulonglong, ctypes.POINTER(ctypes.c_float)]),
        "hmvm_hip": (None, [P(matrix), P(matrix2), c_double_p, ctypes.c_int]),
        "hmvm_hip_bench": (None, [P(matrix), P(matrix2), c_double_p]),
        "hmvm_hip_1": (None, [c_double_p, P(matrix), c_double_p]),
        "hmvm_hip_1t": (None, [c_double_p, P(matrix), c_double_p]),
        "hmvm_hip_2": (None, [c_double_p, P(matrix2), c_double_p]),
        "hmvm_hip_2t": (None, [c_double_p, P(matrix2), c_double_p]),
        "hmvm_hip_last_error": (ctypes.c_int, []),
        "hmvm_hip_last_error_message": (ctypes.c_char_p, []),
        "hmvm_hip_release": (None, [ctypes.c_void_p]),
        "hmvm_options_default": (None, [P(hmvm_options)]),
        "hmvm_plan_create": (ctypes.c_int, [P(matrix2), ctypes.c_int, P(hmvm_options), P(ctypes.c_void_p)]),
        "hmvm_plan_create_aos": (ctypes.c_int, [P(matrix), ctypes.c_int, P(hmvm_options), P(ctypes.c_void_p)]),
        "hmvm_plan_destroy": (None, [ctypes.c_void_p]),
        "hmvm_plan_get_info": (ctypes.c_int, [ctypes.c_void_p, P(hmvm_plan_info)]),
        "hmvm_plan_apply": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                           ctypes.c_void_p]),
        "hmvm_plan_apply_host": (ctypes.c_int, [ctypes.c_void_p, c_double_p, c_double_p, ctypes.c_int]),
        "hmvm_plan_time": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, P(hmvm_timing)]),
        "hmvm_plan_time_flushed": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_longlong, ctypes.c_int, P(hmvm_timing)]),
        "hmvm_plan_bench_apply": (ctypes.c_int, [ctypes.c_void_p, c_double_p, ctypes.c_int, c_double_p]),
        "hmvm_plan_reserve_multi": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
        "hmvm_plan_apply_multi": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                                 ctypes.c_int, ctypes.c_void_p]),
        "hmvm_plan_apply_multi_host": (ctypes.c_int, [ctypes.c_void_p, c_double_p, c_double_p, ctypes.c_int,
                                                      ctypes.c_int]),
        "hmvm_plan_time_multi": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                P(hmvm_timing)]),
        "hmvm_bicgstab": (ctypes.c_int, [ctypes.c_void_p, c_double_p, c_double_p, ctypes.c_int, ctypes.c_double,
                                         ctypes.c_int, ctypes.c_int, P(hmvm_krylov_result)]),
        "hmvm_bicgstab_static": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                                ctypes.c_double, ctypes.c_int, ctypes.c_void_p,
                                                P(hmvm_krylov_result)]),
        "hmvm_plan_check": (ctypes.c_int, [P(matrix2), P(hmvm_options), P(hmvm_plan_info)]),
        "hmvm_plan_set_launch": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]),
        "hmvm_plan_set_mode": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
        "hmvm_plan_fault": (ctypes.c_int, [ctypes.c_void_p]),
        "hmvm_schedule_export": (ctypes.c_int, [P(matrix2), ctypes.c_int, P(hmvm_options), P(hmvm_schedule)]),
        "hmvm_schedule_release": (None, [P(hmvm_schedule)]),
        "hmvm_bandwidth_probe": (ctypes.c_int, [ctypes.c_size_t, ctypes.c_int, c_double_p, c_double_p]),
        "hmvm_device_count": (ctypes.c_int, []),
        "hmvm_device_alloc": (ctypes.c_int, [P(ctypes.c_void_p), ctypes.c_size_t]),
        "hmvm_device_free": (ctypes.c_int, [ctypes.c_void_p]),
        "hmvm_memcpy_h2d": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]),
        "hmvm_memcpy_d2h": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]),
        "hmvm_device_synchronize": (ctypes.c_int, []),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


lib = _load()


def _dptr(a):
    return a.ctypes.data_as(c_double_p)


def _check_vec(a, n, name):
    if not isinstance(a, np.ndarray) or a.dtype != np.float64 or a.ndim != 1 or a.shape[0] < n:
        raise HMVMError("%s must be a float64 vector of length >= %d" % (name, n))
    if not a.flags["C_CONTIGUOUS"]:
        raise HMVMError("%s must be contiguous" % name)


def _raise_hmvm(what):
    code = lib.hmvm_hip_last_error()
    if code != 0:
        raise HMVMError("%s: %s (code %d)" % (what, lib.hmvm_hip_last_error_message().decode(), code))


class HMatrix:
    """Owns a ``matrix`` and/or ``matrix2`` filled by the native loaders or generator."""

    def __init__(self):
        self.mat = matrix()
        self.mat2 = matrix2()
        self.has_mat = False
        self.has_mat2 = False

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def free(self):
        if self.has_mat or self.has_mat2:
            if self.has_mat:
                lib.hmvm_hip_release(ctypes.cast(ctypes.pointer(self.mat), ctypes.c_void_p))
            if self.has_mat2:
                lib.hmvm_hip_release(ctypes.cast(ctypes.pointer(self.mat2), ctypes.c_void_p))
            lib.freeHmatrix(ctypes.byref(self.mat) if self.has_mat else None,
                            ctypes.byref(self.mat2) if self.has_mat2 else None)
            self.has_mat = self.has_mat2 = False

    @property
    def nd(self):
        return self.mat2.nd if self.has_mat2 else self.mat.nd

    @property
    def mat_ptr(self):
        return ctypes.pointer(self.mat) if self.has_mat else None

    @property
    def mat2_ptr(self):
        return ctypes.pointer(self.mat2) if self.has_mat2 else None

    def arrays(self):
        """numpy views of the matrix2 fields (no copies)."""
        m = self.mat2
        n = m.nlf
        out = {k: np.ctypeslib.as_array(getattr(m, k), (n,))
               for k in ("ltmtx", "kt", "ndl", "ndt", "nstrtl", "nstrtt", "a1", "a2")}
        out["rowmat"] = np.ctypeslib.as_array(m.rowmat, (m.len,)) if m.len > 0 else np.zeros(0)
        out["rowmat_t"] = np.ctypeslib.as_array(m.rowmat_t, (m.len,)) if m.len > 0 else np.zeros(0)
        return out

    def build_aos(self):
        """Adds the array-of-structs form (with a2t) built from matrix2."""
        if not self.has_mat2:
            raise HMVMError("no matrix2")
        if not self.has_mat:
            rc = lib.hgen_matrix_from_matrix2(ctypes.byref(self.mat2), ctypes.byref(self.mat))
            if rc:
                raise HMVMError("hgen_matrix_from_matrix2: %s" % lib.hacapk_strerror(rc).decode())
            self.has_mat = True
        return self


def _hacapk_err(rc, what):
    if rc:
        raise HMVMError("%s: %s (code %d)" % (what, lib.hacapk_strerror(rc).decode(), rc))


def loadHmatrix(fname):
    """Format 1 -> HMatrix with both matrix and matrix2 (src_c/loadmatrix.c:8)."""
    h = HMatrix()
    rc = lib.loadHmatrix(os.fsencode(fname), ctypes.byref(h.mat), ctypes.byref(h.mat2))
    _hacapk_err(rc, "loadHmatrix(%s)" % fname)
    h.has_mat = h.has_mat2 = True
    return h


def loadHmatrix2(fname):
    """Format 2 -> HMatrix with matrix2 only (src_c/loadmatrix.c:134)."""
    h = HMatrix()
    rc = lib.loadHmatrix2(os.fsencode(fname), ctypes.byref(h.mat2))
    _hacapk_err(rc, "loadHmatrix2(%s)" % fname)
    h.has_mat2 = True
    return h


def dummymatrix(ndl, ndt, n):
    """All-dense synthetic matrix (src_c/loadmatrix.c:240)."""
    h = HMatrix()
    _hacapk_err(lib.dummymatrix(ctypes.byref(h.mat), ctypes.byref(h.mat2), ndl, ndt, n), "dummymatrix")
    h.has_mat = h.has_mat2 = True
    return h


def writeHmatrix1(fname, h):
    _hacapk_err(lib.writeHmatrix1(os.fsencode(fname), ctypes.byref(h.mat2)), "writeHmatrix1")


def writeHmatrix2(fname, h):
    _hacapk_err(lib.writeHmatrix2(os.fsencode(fname), ctypes.byref(h.mat2)), "writeHmatrix2")


def gen_params(n, geometry="sphere", leaf=32, eta=2.0, eps=1e-4, seed=20261202, jitter=0.25, kmax=1000):
    p = hgen_params()
    lib.hgen_default_params(ctypes.byref(p), int(n))
    p.geometry = {"sphere": HGEN_SPHERE, "cube": HGEN_CUBE}[geometry] if isinstance(geometry, str) else geometry
    p.leaf, p.eta, p.eps, p.seed, p.jitter, p.kmax = leaf, eta, eps, seed, jitter, kmax
    return p


def generate(n, **kw):
    """Generates an H-matrix (matrix2 only; call .build_aos() for matrix)."""
    p = gen_params(n, **kw)
    h = HMatrix()
    _hacapk_err(lib.hgen_generate(ctypes.byref(p), ctypes.byref(h.mat2)), "hgen_generate")
    h.has_mat2 = True
    h.params = p
    return h


def kernel_block(params, i0, nr, j0, nc):
    out = np.empty((nr, nc))
    _hacapk_err(lib.hgen_kernel_block(ctypes.byref(params), i0, nr, j0, nc, _dptr(out)), "hgen_kernel_block")
    return out


def accuracy(h, stride=1):
    return lib.hgen_accuracy(ctypes.byref(h.params), ctypes.byref(h.mat2), stride)


class HMatrix64:
    """Owns a ``matrix2_64`` (int64 offsets) filled by the native loaders or generator."""

    def __init__(self):
        self.mat2 = matrix2_64()
        self.has_mat2 = False

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def free(self):
        if self.has_mat2:
            lib.hmvm_hip_release(ctypes.cast(ctypes.pointer(self.mat2), ctypes.c_void_p))
            lib.freeHmatrix64(ctypes.byref(self.mat2))
            self.has_mat2 = False

    @property
    def nd(self):
        return self.mat2.nd

    def arrays(self):
        m = self.mat2
        n = m.nlf
        out = {k: np.ctypeslib.as_array(getattr(m, k), (n,))
               for k in ("ltmtx", "kt", "ndl", "ndt", "nstrtl", "nstrtt", "a1", "a2")}
        out["rowmat"] = np.ctypeslib.as_array(m.rowmat, (m.len,)) if m.len > 0 and m.rowmat else np.zeros(0)
        out["rowmat_t"] = (np.ctypeslib.as_array(m.rowmat_t, (m.len,)) if m.len > 0 and m.rowmat_t
                           else None)
        return out


def leaf_subset(h, idx):
    """An HMatrix64 view of the leaves ``idx`` of ``h`` (HMatrix or HMatrix64):
    same matrix order and value buffer, only the selected leaves.  The view
    borrows ``h``'s values (keep ``h`` alive) and frees nothing itself."""
    idx = np.asarray(idx, dtype=np.int64)
    a = h.arrays()
    v = HMatrix64()
    v._keep = {k: np.ascontiguousarray(a[k][idx], dtype=np.int32)
               for k in ("ltmtx", "kt", "ndl", "ndt", "nstrtl", "nstrtt")}
    v._keep.update({k: np.ascontiguousarray(a[k][idx], dtype=np.int64) for k in ("a1", "a2")})
    v._parent = h
    m = v.mat2
    m.nd, m.nlf, m.len = h.mat2.nd, len(idx), h.mat2.len
    m.ktmax = int(v._keep["kt"].max()) if len(idx) else 0
    for k in ("ltmtx", "kt", "ndl", "ndt", "nstrtl", "nstrtt"):
        setattr(m, k, v._keep[k].ctypes.data_as(c_int_p))
    for k in ("a1", "a2"):
        setattr(m, k, v._keep[k].ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)))
    m.rowmat = h.mat2.rowmat
    m.rowmat_t = h.mat2.rowmat_t
    v.has_mat2 = False  # borrowed: never freed through this object
    return v


def loadHmatrix_64(fname, with_rowmat_t=False):
    """Format 1, 2 or 2L -> HMatrix64."""
    h = HMatrix64()
    rc = lib.loadHmatrix_64(os.fsencode(fname), ctypes.byref(h.mat2), HACAPK_WITH_ROWMAT_T if with_rowmat_t else 0)
    _hacapk_err(rc, "loadHmatrix_64(%s)" % fname)
    h.has_mat2 = True
    return h


def writeHmatrix2_64(fname, h):
    _hacapk_err(lib.writeHmatrix2_64(os.fsencode(fname), ctypes.byref(h.mat2)), "writeHmatrix2_64")


def writeHmatrix1_64(fname, h):
    _hacapk_err(lib.writeHmatrix1_64(os.fsencode(fname), ctypes.byref(h.mat2)), "writeHmatrix1_64")


def widen(h):
    """Deep copy of an HMatrix's matrix2 into the int64 form."""
    w = HMatrix64()
    _hacapk_err(lib.hacapk_widen_matrix2(ctypes.byref(h.mat2), ctypes.byref(w.mat2)), "hacapk_widen_matrix2")
    w.has_mat2 = True
    return w


def scan_file(fname):
    """Leaf descriptors of a format 1/2/2L file: (HMatrix64 without values, data_pos, format)."""
    h = HMatrix64()
    pos = ctypes.c_longlong()
    fmt = ctypes.c_int()
    _hacapk_err(lib.hacapk_scan_file(os.fsencode(fname), ctypes.byref(h.mat2), ctypes.byref(pos), ctypes.byref(fmt)),
                "hacapk_scan_file(%s)" % fname)
    h.has_mat2 = True
    return h, pos.value, fmt.value


def generate64(n, with_rowmat_t=False, **kw):
    """Generates an H-matrix in the int64 form (for more than 2^31-1 elements)."""
    p = gen_params(n, **kw)
    h = HMatrix64()
    _hacapk_err(lib.hgen_generate64(ctypes.byref(p), ctypes.byref(h.mat2),
                                    HACAPK_WITH_ROWMAT_T if with_rowmat_t else 0), "hgen_generate64")
    h.has_mat2 = True
    h.params = p
    return h


def accuracy64(h, stride=1):
    return lib.hgen_accuracy64(ctypes.byref(h.params), ctypes.byref(h.mat2), stride)


# ---- drop-in HMVM (v += A b) -------------------------------------------------------
def _variant(fn, v, h, b, use_mat2):
    _check_vec(v, h.nd, "v")
    _check_vec(b, h.nd, "b")
    if use_mat2 and not h.has_mat2:
        raise HMVMError("this variant needs matrix2")
    if not use_mat2 and not h.has_mat:
        raise HMVMError("this variant needs matrix")
    fn(_dptr(v), ctypes.byref(h.mat2) if use_mat2 else ctypes.byref(h.mat), _dptr(b))
    _raise_hmvm(fn.__name__)
    return v


def hmvm_hip_1(v, h, b):
    return _variant(lib.hmvm_hip_1, v, h, b, False)


def hmvm_hip_1t(v, h, b):
    return _variant(lib.hmvm_hip_1t, v, h, b, False)


def hmvm_hip_2(v, h, b):
    return _variant(lib.hmvm_hip_2, v, h, b, True)


def hmvm_hip_2t(v, h, b):
    return _variant(lib.hmvm_hip_2t, v, h, b, True)


def hmvm_hip_2_64(v, h, b):
    _check_vec(v, h.nd, "v")
    _check_vec(b, h.nd, "b")
    lib.hmvm_hip_2_64(_dptr(v), ctypes.byref(h.mat2), _dptr(b))
    _raise_hmvm("hmvm_hip_2_64")
    return v


def hmvm_hip_2t_64(v, h, b):
    _check_vec(v, h.nd, "v")
    _check_vec(b, h.nd, "b")
    lib.hmvm_hip_2t_64(_dptr(v), ctypes.byref(h.mat2), _dptr(b))
    _raise_hmvm("hmvm_hip_2t_64")
    return v


def hmvm_hip(h, b, dump_result=0):
    _check_vec(b, h.nd, "b")
    lib.hmvm_hip(h.mat_ptr, h.mat2_ptr, _dptr(b), int(dump_result))
    _raise_hmvm("hmvm_hip")


def hmvm_hip_bench(h, b):
    _check_vec(b, h.nd, "b")
    lib.hmvm_hip_bench(h.mat_ptr, h.mat2_ptr, _dptr(b))
    _raise_hmvm("hmvm_hip_bench")


def release(h):
    for m in (h.mat_ptr, h.mat2_ptr):
        if m is not None:
            lib.hmvm_hip_release(ctypes.cast(m, ctypes.c_void_p))


def device_count():
    return lib.hmvm_device_count()


def tile_occupancy(storage, kernel, block_threads):
    """Resident wavefronts of the tile kernel on the current device."""
    w = ctypes.c_int()
    rc = lib.hmvm_tile_occupancy(STORAGE[storage] if isinstance(storage, str) else storage, kernel, block_threads,
                                 ctypes.byref(w))
    if rc:
        raise HMVMError("hmvm_tile_occupancy: %s" % lib.hmvm_hip_last_error_message().decode())
    return w.value


# ---- device-resident plan ---------------------------------------------------------------
def bandwidth_probe(nbytes=2 << 30, iters=10):
    """Achievable (read, copy) GB/s of this device with plain streaming kernels."""
    r, c = ctypes.c_double(), ctypes.c_double()
    rc = lib.hmvm_bandwidth_probe(nbytes, iters, ctypes.byref(r), ctypes.byref(c))
    if rc:
        raise HMVMError("hmvm_bandwidth_probe: %s" % lib.hmvm_hip_last_error_message().decode())
    return r.value, c.value


def options(storage="f64", w1=0, w2=0, block_threads=0, kernel=None, order=None, mode=0, t_mode=None, max_rows=0,
            max_rows_b=0, w3=0):
    """hmvm_options; None / 0 keep the library defaults."""
    o = hmvm_options()
    lib.hmvm_options_default(ctypes.byref(o))
    o.max_rows = max_rows
    o.max_rows_b = max_rows_b
    o.w3 = w3
    if kernel is not None:
        o.kernel = kernel
    if order is not None:
        o.order = order
    if t_mode is not None:
        o.t_mode = t_mode
    o.mode = mode
    o.storage = STORAGE[storage] if isinstance(storage, str) else int(storage)
    if w1:
        o.w1 = w1
    if w2:
        o.w2 = w2
    if block_threads:
        o.block_threads = block_threads
    return o


def plan_check(h, **kw):
    """Host-only schedule check (no GPU); returns the plan info dict."""
    info = hmvm_plan_info()
    o = options(**kw)
    rc = lib.hmvm_plan_check(ctypes.byref(h.mat2), ctypes.byref(o), ctypes.byref(info))
    if rc:
        raise HMVMError("hmvm_plan_check: %s" % lib.hmvm_hip_last_error_message().decode())
    return info.as_dict()


def _struct_array(ptr, n, dtype):
    if n == 0:
        return np.zeros(0, dtype=dtype)
    raw = np.ctypeslib.as_array(ctypes.cast(ptr, ctypes.POINTER(ctypes.c_uint8)), (n * dtype.itemsize,))
    return raw.view(dtype).copy()


PANEL_DT = np.dtype([("val_off", "<i8"), ("ld", "<i4"), ("ncols", "<i4"), ("vec_off", "<i4"), ("vec_sel", "<i4"),
                     ("factor", "<i4"), ("vec_chunks", "<i4")])
TILE_DT = np.dtype([("nrows", "<i4"), ("panel0", "<i4"), ("npanel", "<i4"), ("out_off", "<i4"), ("map_off", "<i4"),
                    ("flags", "<i4"), ("ncols", "<i4"), ("map_off_multi", "<i4"), ("first", PANEL_DT)])
REPACK_DT = np.dtype([("src_off", "<i8"), ("dst_off", "<i8"), ("rows", "<i4"), ("cols", "<i4"), ("src_rs", "<i4"),
                      ("src_cs", "<i4"), ("dst_cs", "<i4"), ("factor", "<i4")])


def schedule(h, source="rowmat", **kw):
    """The tile schedule a plan would run for this matrix (or list of matrices,
    as a batch), as numpy arrays (host only)."""
    s = hmvm_schedule()
    o = options(**kw)
    if isinstance(h, (list, tuple)):
        arr = (ctypes.POINTER(matrix2) * len(h))(*[ctypes.pointer(m.mat2) for m in h])
        rc = lib.hmvm_schedule_export_batch(arr, len(h), 1 if source == "rowmat_t" else 0, ctypes.byref(o),
                                            ctypes.byref(s))
    elif isinstance(h, HMatrix64):
        rc = lib.hmvm_schedule_export64(ctypes.byref(h.mat2), 1 if source == "rowmat_t" else 0, ctypes.byref(o),
                                        ctypes.byref(s))
    else:
        rc = lib.hmvm_schedule_export(ctypes.byref(h.mat2), 1 if source == "rowmat_t" else 0, ctypes.byref(o),
                                      ctypes.byref(s))
    if rc:
        raise HMVMError("hmvm_schedule_export: %s" % lib.hmvm_hip_last_error_message().decode())
    try:
        def seg(ss):
            return {"n": ss.n, "seg": np.ctypeslib.as_array(ss.seg, (max(ss.n, 1),))[:ss.n].copy(),
                    "ptr": np.ctypeslib.as_array(ss.ptr, (ss.nseg + 1,)).copy(),
                    "list": (np.ctypeslib.as_array(ss.list, (max(ss.ptr[ss.nseg], 1),))[:ss.ptr[ss.nseg]].copy())}
        return {"t_fused": bool(s.t_fused), "nvals": s.nvals, "n_t": s.n_t, "n_tslots": s.n_tslots,
                "n_zslots": s.n_zslots,
                "tiles_a": _struct_array(s.tiles_a, s.ntiles_a, TILE_DT),
                "tiles_b": _struct_array(s.tiles_b, s.ntiles_b, TILE_DT),
                "panels": _struct_array(s.panels, s.npanels, PANEL_DT),
                "repack": _struct_array(s.repack, s.nrepack, REPACK_DT),
                "tsum": seg(s.tsum), "ysum": seg(s.ysum), "t_direct": bool(s.t_direct),
                "tmap": np.ctypeslib.as_array(s.tmap, (max(s.ntmap, 1),))[:s.ntmap].copy()}
    finally:
        lib.hmvm_schedule_release(ctypes.byref(s))


class Plan:
    """Repacked device copy of an H-matrix; y (+)= A x on device or host vectors.

    ``h`` is an HMatrix (source rowmat / rowmat_t / a2 / a2t), an HMatrix64
    (rowmat / rowmat_t), a file name (streamed, formats 1/2/2L) or a list of
    HMatrix for a block-diagonal batch (``self.row_offsets`` then gives each
    matrix's rows)."""

    def __init__(self, h, source="rowmat", **kw):
        self._p = ctypes.c_void_p()
        self.row_offsets = None
        o = options(**kw)
        if isinstance(h, (str, bytes, os.PathLike)):
            rc = lib.hmvm_plan_create_file(os.fsencode(h), ctypes.byref(o), ctypes.byref(self._p))
            if rc:
                raise HMVMError("hmvm_plan_create_file: %s (code %d)" % (lib.hmvm_hip_last_error_message().decode(),
                                                                         rc))
            self.nd = self.info()["nd"]
            return
        if isinstance(h, (list, tuple)):
            arr = (ctypes.POINTER(matrix2) * len(h))(*[ctypes.pointer(m.mat2) for m in h])
            offs = (ctypes.c_int * (len(h) + 1))()
            rc = lib.hmvm_plan_create_batch(arr, len(h), 1 if source == "rowmat_t" else 0, ctypes.byref(o),
                                            ctypes.byref(self._p), offs)
            if rc:
                raise HMVMError("hmvm_plan_create_batch: %s (code %d)" % (lib.hmvm_hip_last_error_message().decode(),
                                                                          rc))
            self.row_offsets = np.array(offs[:], dtype=np.int64)
            self.nd = int(self.row_offsets[-1])
            return
        if isinstance(h, HMatrix64):
            rc = lib.hmvm_plan_create64(ctypes.byref(h.mat2), 1 if source == "rowmat_t" else 0, ctypes.byref(o),
                                        ctypes.byref(self._p))
            if rc:
                raise HMVMError("hmvm_plan_create64: %s (code %d)" % (lib.hmvm_hip_last_error_message().decode(), rc))
            self.nd = h.nd
            return
        if source in ("rowmat", "rowmat_t"):
            rc = lib.hmvm_plan_create(ctypes.byref(h.mat2), 1 if source == "rowmat_t" else 0, ctypes.byref(o),
                                      ctypes.byref(self._p))
        elif source in ("a2", "a2t"):
            rc = lib.hmvm_plan_create_aos(ctypes.byref(h.mat), 1 if source == "a2t" else 0, ctypes.byref(o),
                                          ctypes.byref(self._p))
        else:
            raise HMVMError("unknown source %r" % source)
        if rc:
            raise HMVMError("hmvm_plan_create: %s (code %d)" % (lib.hmvm_hip_last_error_message().decode(), rc))
        self.nd = h.nd

    def __del__(self):
        self.close()

    def close(self):
        if getattr(self, "_p", None) and self._p.value:
            lib.hmvm_plan_destroy(self._p)
            self._p = ctypes.c_void_p()

    def info(self):
        i = hmvm_plan_info()
        lib.hmvm_plan_get_info(self._p, ctypes.byref(i))
        return i.as_dict()

    def apply_host(self, x, y=None, accumulate=False):
        _check_vec(x, self.nd, "x")
        if y is None:
            y = np.zeros(self.nd)
        _check_vec(y, self.nd, "y")
        rc = lib.hmvm_plan_apply_host(self._p, _dptr(x), _dptr(y), int(bool(accumulate)))
        if rc:
            raise HMVMError("hmvm_plan_apply_host: %s" % lib.hmvm_hip_last_error_message().decode())
        return y

    def apply(self, d_x, d_y, accumulate=False, stream=None):
        rc = lib.hmvm_plan_apply(self._p, ctypes.c_void_p(d_x), ctypes.c_void_p(d_y), int(bool(accumulate)),
                                 ctypes.c_void_p(stream) if stream else None)
        if rc:
            raise HMVMError("hmvm_plan_apply: %s" % lib.hmvm_hip_last_error_message().decode())

    def set_launch(self, kernel, block_threads=256):
        rc = lib.hmvm_plan_set_launch(self._p, kernel, block_threads)
        if rc:
            raise HMVMError("hmvm_plan_set_launch: %s" % lib.hmvm_hip_last_error_message().decode())

    def set_phase_launch(self, which, kernel, block_threads=256):
        rc = lib.hmvm_plan_set_phase_launch(self._p, which, kernel, block_threads)
        if rc:
            raise HMVMError("hmvm_plan_set_phase_launch: %s" % lib.hmvm_hip_last_error_message().decode())

    def tile_features(self, which, nrhs=1):
        n = self.info()["ntiles_a" if which == 0 else "ntiles_b"]
        f = np.zeros((n, HMVM_NFEAT), dtype=np.float32)
        if n:
            rc = lib.hmvm_plan_tile_features(self._p, which, nrhs, f.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
            if rc:
                raise HMVMError("hmvm_plan_tile_features: %s" % lib.hmvm_hip_last_error_message().decode())
        return f

    def profile_tiles(self, which, iters=3):
        """(per-tile duration us, per-tile start offset us, mean launch ms) of launch A (0) or B (1)."""
        n = self.info()["ntiles_a" if which == 0 else "ntiles_b"]
        d = np.zeros(max(n, 1))
        s0 = np.zeros(max(n, 1))
        lm = ctypes.c_double()
        rc = lib.hmvm_plan_profile_tiles(self._p, which, iters, _dptr(d), _dptr(s0), ctypes.byref(lm))
        if rc:
            raise HMVMError("hmvm_plan_profile_tiles: %s" % lib.hmvm_hip_last_error_message().decode())
        return d[:n], s0[:n], lm.value

    def set_tile_flags(self, which, flags):
        """Per-tile kernel path of launch A (0) or B (1), in dispatch order:
        HMVM_TILE_ROW_PER_LANE = lane-per-row path, 0 = column-split (default)."""
        flags = np.ascontiguousarray(flags, dtype=np.int32)
        n = self.info()["ntiles_a" if which == 0 else "ntiles_b"]
        if flags.shape != (n,):
            raise HMVMError("one flag per tile (%d)" % n)
        rc = lib.hmvm_plan_set_tile_flags(self._p, which, flags.ctypes.data_as(c_int_p))
        if rc:
            raise HMVMError("hmvm_plan_set_tile_flags: %s" % lib.hmvm_hip_last_error_message().decode())

    def tile_flags(self, which):
        n = self.info()["ntiles_a" if which == 0 else "ntiles_b"]
        f = np.zeros(max(n, 1), dtype=np.int32)
        rc = lib.hmvm_plan_get_tile_flags(self._p, which, f.ctypes.data_as(c_int_p))
        if rc:
            raise HMVMError("hmvm_plan_get_tile_flags: %s" % lib.hmvm_hip_last_error_message().decode())
        return f[:n]

    def set_tile_order(self, which, perm):
        perm = np.ascontiguousarray(perm, dtype=np.int32)
        rc = lib.hmvm_plan_set_tile_order(self._p, which, perm.ctypes.data_as(c_int_p))
        if rc:
            raise HMVMError("hmvm_plan_set_tile_order: %s" % lib.hmvm_hip_last_error_message().decode())

    def time_graph(self, count=1, warmup=3, iters=20):
        """Timing of hipGraph replays of ``count`` applies (per apply)."""
        t = hmvm_timing()
        rc = lib.hmvm_plan_time_graph(self._p, count, warmup, iters, ctypes.byref(t))
        if rc:
            raise HMVMError("hmvm_plan_time_graph: %s" % lib.hmvm_hip_last_error_message().decode())
        return t.as_dict()

    def graph(self, d_x, d_y, accumulate=False, count=1):
        """hipGraph of ``count`` applies on fixed device vectors; returns a Graph."""
        g = ctypes.c_void_p()
        rc = lib.hmvm_plan_graph_create(self._p, ctypes.c_void_p(d_x), ctypes.c_void_p(d_y), int(bool(accumulate)),
                                        count, ctypes.byref(g))
        if rc:
            raise HMVMError("hmvm_plan_graph_create: %s" % lib.hmvm_hip_last_error_message().decode())
        return Graph(g)

    def set_mode(self, mode):
        rc = lib.hmvm_plan_set_mode(self._p, mode)
        if rc:
            raise HMVMError("hmvm_plan_set_mode: %s" % lib.hmvm_hip_last_error_message().decode())

    def fault(self):
        """1 when a bounded wait of the persistent launch gave up."""
        return lib.hmvm_plan_fault(self._p)

    def apply_multi_host(self, X, Y=None, accumulate=False):
        """Y (+)= A X for X of shape (nd, nrhs), nrhs in {16, 32, 64}, on the matrix cores."""
        X = np.ascontiguousarray(X, dtype=np.float64)
        if X.ndim != 2 or X.shape[0] != self.nd:
            raise HMVMError("X must be (nd, nrhs)")
        nrhs = X.shape[1]
        if Y is None:
            Y = np.zeros_like(X)
        if Y.shape != X.shape or Y.dtype != np.float64 or not Y.flags["C_CONTIGUOUS"]:
            raise HMVMError("Y must be a contiguous float64 array shaped like X")
        rc = lib.hmvm_plan_apply_multi_host(self._p, _dptr(X), _dptr(Y), nrhs, int(bool(accumulate)))
        if rc:
            raise HMVMError("hmvm_plan_apply_multi_host: %s" % lib.hmvm_hip_last_error_message().decode())
        return Y

    def time_multi(self, nrhs=32, warmup=3, iters=10):
        t = hmvm_timing()
        rc = lib.hmvm_plan_time_multi(self._p, nrhs, warmup, iters, ctypes.byref(t))
        if rc:
            raise HMVMError("hmvm_plan_time_multi: %s" % lib.hmvm_hip_last_error_message().decode())
        return t.as_dict()

    def bicgstab(self, b, u0=None, max_iter=100, tol=1e-8, use_graph=True, check_every=10):
        """BiCGSTAB for A u = b on the device (m_HACApK_solve.f90:306-380); returns (u, result dict)."""
        _check_vec(b, self.nd, "b")
        u = np.zeros(self.nd) if u0 is None else np.array(u0, dtype=np.float64)
        res = hmvm_krylov_result()
        rc = lib.hmvm_bicgstab(self._p, _dptr(b), _dptr(u), int(max_iter), float(tol), int(bool(use_graph)),
                               int(check_every), ctypes.byref(res))
        if rc == HMVM_ERR_BREAKDOWN:
            raise KrylovBreakdown("hmvm_bicgstab: breakdown after %d iterations (zero divisor or non-finite "
                                  "scalar)" % res.iters, u, res.as_dict())
        if rc:
            raise HMVMError("hmvm_bicgstab failed (code %d): %s" % (rc, lib.hmvm_hip_last_error_message().decode()))
        return u, res.as_dict()

    def time(self, warmup=5, iters=20):
        t = hmvm_timing()
        rc = lib.hmvm_plan_time(self._p, warmup, iters, ctypes.byref(t))
        if rc:
            raise HMVMError("hmvm_plan_time: %s" % lib.hmvm_hip_last_error_message().decode())
        return t.as_dict()

    def time_flushed(self, iters=20, flush_mb=1024):
        """Per-apply timing with L2 and the Infinity Cache flushed before each
        apply (a kernel writes flush_mb MiB of scratch)."""
        t = hmvm_timing()
        rc = lib.hmvm_plan_time_flushed(self._p, int(flush_mb) << 20, iters, ctypes.byref(t))
        if rc:
            raise HMVMError("hmvm_plan_time_flushed: %s" % lib.hmvm_hip_last_error_message().decode())
        return t.as_dict()


class Graph:
    """A captured hipGraph of plan applies (see Plan.graph)."""

    def __init__(self, handle):
        self._g = handle

    def launch(self, stream=None):
        rc = lib.hmvm_graph_launch(self._g, ctypes.c_void_p(stream) if stream else None)
        if rc:
            raise HMVMError("hmvm_graph_launch: %s" % lib.hmvm_hip_last_error_message().decode())

    def close(self):
        if getattr(self, "_g", None) and self._g.value:
            lib.hmvm_graph_destroy(self._g)
            self._g = ctypes.c_void_p()

    def __del__(self):
        self.close()


class DeviceBuffer:
    """A device allocation through the library's helpers (for device-resident tests and benchmarks)."""

    def __init__(self, nbytes):
        self.ptr = ctypes.c_void_p()
        self.nbytes = int(nbytes)
        if lib.hmvm_device_alloc(ctypes.byref(self.ptr), self.nbytes):
            raise HMVMError("hmvm_device_alloc: %s" % lib.hmvm_hip_last_error_message().decode())

    def upload(self, a):
        a = np.ascontiguousarray(a)
        if a.nbytes > self.nbytes or lib.hmvm_memcpy_h2d(self.ptr, a.ctypes.data_as(ctypes.c_void_p), a.nbytes):
            raise HMVMError("upload failed")

    def download(self, a):
        if a.nbytes > self.nbytes or lib.hmvm_memcpy_d2h(a.ctypes.data_as(ctypes.c_void_p), self.ptr, a.nbytes):
            raise HMVMError("download failed")
        return a

    @property
    def value(self):
        return self.ptr.value

    def close(self):
        if getattr(self, "ptr", None) and self.ptr.value:
            lib.hmvm_device_free(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        self.close()


# ---- the learned cost model -----------------------------------------------------------
def _fptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


class CostMLP:
    """nin -> nh -> nh -> 1 ReLU MLP whose training and inference run as HIP kernels."""

    def __init__(self, nin=HMVM_NFEAT, nh=64, seed=1):
        self._m = ctypes.c_void_p()
        rc = lib.hmvm_mlp_create(nin, nh, seed, ctypes.byref(self._m))
        if rc:
            raise HMVMError("hmvm_mlp_create failed (code %d)" % rc)
        self.nin, self.nh = nin, nh

    def close(self):
        if getattr(self, "_m", None) and self._m.value:
            lib.hmvm_mlp_destroy(self._m)
            self._m = ctypes.c_void_p()

    def __del__(self):
        self.close()

    @property
    def nparams(self):
        return lib.hmvm_mlp_num_params(self._m)

    def get(self):
        p = np.zeros(self.nparams, dtype=np.float32)
        mu = np.zeros(self.nin, dtype=np.float32)
        sd = np.zeros(self.nin, dtype=np.float32)
        lib.hmvm_mlp_get_params(self._m, _fptr(p), _fptr(mu), _fptr(sd))
        return p, mu, sd

    def set(self, params=None, mean=None, stdev=None):
        arrs = [None if a is None else np.ascontiguousarray(a, dtype=np.float32) for a in (params, mean, stdev)]
        lib.hmvm_mlp_set_params(self._m, *[None if a is None else _fptr(a) for a in arrs])

    def fit_scaler(self, X):
        X = np.ascontiguousarray(X, dtype=np.float32)
        lib.hmvm_mlp_fit_scaler(self._m, _fptr(X), X.shape[0])

    def predict(self, X):
        X = np.ascontiguousarray(X, dtype=np.float32)
        y = np.zeros(X.shape[0], dtype=np.float32)
        rc = lib.hmvm_mlp_predict(self._m, _fptr(X), X.shape[0], _fptr(y))
        if rc:
            raise HMVMError("hmvm_mlp_predict failed (code %d)" % rc)
        return y

    def loss_grad(self, X, y):
        X = np.ascontiguousarray(X, dtype=np.float32)
        y = np.ascontiguousarray(y, dtype=np.float32)
        loss = ctypes.c_float()
        g = np.zeros(self.nparams, dtype=np.float32)
        rc = lib.hmvm_mlp_loss_grad(self._m, _fptr(X), _fptr(y), X.shape[0], ctypes.byref(loss), _fptr(g))
        if rc:
            raise HMVMError("hmvm_mlp_loss_grad failed (code %d)" % rc)
        return loss.value, g

    def train(self, X, y, epochs=100, batch=1024, lr=1e-3, seed=7):
        X = np.ascontiguousarray(X, dtype=np.float32)
        y = np.ascontiguousarray(y, dtype=np.float32)
        loss = ctypes.c_float()
        rc = lib.hmvm_mlp_train(self._m, _fptr(X), _fptr(y), X.shape[0], epochs, batch, lr, seed, ctypes.byref(loss))
        if rc:
            raise HMVMError("hmvm_mlp_train failed (code %d)" % rc)
        return loss.value
