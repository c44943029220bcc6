"""libparanumal_amd: MI355X (gfx950) build of libParanumal's C0 hex elliptic path.

The product is the in-tree shared library ``libp_amd.so`` (HIP kernels, the
C++ elliptic_t/pcg/p-multigrid host layer and its C ABI, declared in
``include/libp_amd.h``) plus the ``ellipticMain`` driver. This module only
binds that C ABI with ctypes for Python callers; it has no compute path of its
own and raises if the library is missing.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libp_amd.so")
MAIN_PATH = os.path.join(_HERE, "ellipticMain")

_lib = None

c_int, c_long, c_double, c_float, c_void_p, c_char_p = (
    ctypes.c_int, ctypes.c_long, ctypes.c_double, ctypes.c_float, ctypes.c_void_p, ctypes.c_char_p)

# name -> (restype, argtypes) for every entry point of include/libp_amd.h
SIGNATURES = {
    "libp_status_string": (c_char_p, [c_int]),
    "libp_last_error": (c_char_p, []),
    "libp_ax_hex_partial_f64": (c_int, [c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_double,
                                        c_void_p, c_void_p, c_void_p]),
    "libp_ax_hex_partial_f32": (c_int, [c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_float,
                                        c_void_p, c_void_p, c_void_p]),
    "libp_ax_hex_local_f64": (c_int, [c_int, c_int, c_void_p, c_void_p, c_void_p, c_double, c_void_p, c_void_p,
                                      c_void_p]),
    "libp_ax_hex_partial_mfma_n15": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_double,
                                             c_void_p, c_void_p, c_void_p]),
    "libp_hbm_probe_f64": (c_int, [c_int, c_long, c_void_p, c_void_p, c_void_p]),
    "libp_ax_hex_face_nodes": (c_int, [c_int]),
    "libp_ax_hex_split_blocks": (c_int, [c_int, c_int]),
    "libp_ax_hex_split_f64": (c_int, [c_int, c_int] + [c_void_p] * 5 + [c_double] + [c_void_p] * 6),
    "libp_ogs_gather_face_blocks": (c_int, [c_int]),
    "libp_ogs_gather_face_f64": (c_int, [c_int] + [c_void_p] * 6),
    "libp_pcg_gather_face_pap_f64": (c_int, [c_int] + [c_void_p] * 6 + [c_int] + [c_void_p] * 3),
    "libp_ogs_gather_f64": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "libp_ogs_gather_f32": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "libp_ogs_scatter_f64": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "libp_ogs_scatter_i32": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "libp_axpy_f64": (c_int, [c_int, c_double, c_void_p, c_double, c_void_p, c_void_p]),
    "libp_axpy_f32": (c_int, [c_int, c_float, c_void_p, c_float, c_void_p, c_void_p]),
    "libp_zaxpy_f64": (c_int, [c_int, c_double, c_void_p, c_double, c_void_p, c_void_p, c_void_p]),
    "libp_amx_f64": (c_int, [c_int, c_double, c_void_p, c_void_p, c_void_p]),
    "libp_amx_f32": (c_int, [c_int, c_float, c_void_p, c_void_p, c_void_p]),
    "libp_amxpy_f64": (c_int, [c_int, c_double, c_void_p, c_void_p, c_double, c_void_p, c_void_p]),
    "libp_amxpy_f32": (c_int, [c_int, c_float, c_void_p, c_void_p, c_float, c_void_p, c_void_p]),
    "libp_set_f64": (c_int, [c_int, c_double, c_void_p, c_void_p]),
    "libp_add_f64": (c_int, [c_int, c_double, c_void_p, c_void_p]),
    "libp_scale_f64": (c_int, [c_int, c_double, c_void_p, c_void_p]),
    "libp_convert_f64_to_f32": (c_int, [c_int, c_void_p, c_void_p, c_void_p]),
    "libp_convert_f32_to_f64": (c_int, [c_int, c_void_p, c_void_p, c_void_p]),
    "libp_reduce_ws_bytes": (c_long, []),
    "libp_inner_prod_f64": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "libp_weighted_inner_prod_f64": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "libp_norm2sq_f64": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_void_p]),
    "libp_sum_f64": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_void_p]),
    "libp_pcg_update_f64": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "libp_pcg_rdotz_f64": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_void_p]),
    "libp_pcg_update_p_f64": (c_int, [c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p]),
    "libp_scalars_to_host_f64": (c_int, [c_int, c_void_p, c_void_p, c_void_p]),
    "libp_pcg_pap_f64": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "libp_pcg_jacobi_rdotz_f64": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p,
                                          c_void_p]),
    "libp_mg_coarsen_hex_f64": (c_int, [c_int, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                        c_void_p]),
    "libp_mg_coarsen_hex_f32": (c_int, [c_int, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                        c_void_p]),
    "libp_mg_prolong_hex_f64": (c_int, [c_int, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                        c_void_p]),
    "libp_mg_prolong_hex_f32": (c_int, [c_int, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                        c_void_p]),
    "libp_cheb_start_f32": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_float, c_void_p, c_void_p, c_void_p,
                                    c_int, c_void_p]),
    "libp_cheb_step_f32": (c_int, [c_int, c_void_p, c_void_p, c_float, c_float, c_void_p, c_void_p, c_void_p,
                                   c_void_p]),
    "libp_cheb_start_f64": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_double, c_void_p, c_void_p, c_void_p,
                                    c_int, c_void_p]),
    "libp_cheb_step_f64": (c_int, [c_int, c_void_p, c_void_p, c_double, c_double, c_void_p, c_void_p, c_void_p,
                                   c_void_p]),
    "libp_geofacs_hex_f64": (c_int, [c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                     c_double, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                     c_void_p]),
    "libp_diag_hex_local_f64": (c_int, [c_int, c_int, c_void_p, c_void_p, c_void_p, c_double, c_void_p, c_int,
                                        c_double, c_void_p, c_void_p]),
    "libp_rhs_hex_f64": (c_int, [c_int, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_double,
                                 c_void_p, c_void_p, c_void_p, c_void_p]),
    "libp_addbc_hex_f64": (c_int, [c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                   c_void_p]),
    "libp_pcg_jacobi_hex1_iteration_f64": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_double, c_int,
                                                   c_void_p, c_void_p, c_void_p, c_int] + [c_void_p] * 9),
    "libp_pcg_jacobi_hex1_check_f64": (c_int, [c_int, c_void_p, c_void_p, c_void_p]),
    "libp_pcg_jacobi_hex1_cg1_f64": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_double, c_int,
                                             c_void_p, c_void_p, c_void_p, c_double, c_int, c_int]
                                    + [c_void_p] * 10),
    "libp_pcg_jacobi_csr_cg1_f64": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_double, c_int, c_int]
                                   + [c_void_p] * 9),
    "libp_pcg_jacobi_csr_cg1_check_f64": (c_int, [c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p]),
    "libp_pcg_gather_pap_f64": (c_int, [c_int] + [c_void_p] * 8),
    "libp_pcg_jacobi_update_f64": (c_int, [c_int] + [c_void_p] * 9),
    "libp_pcg_jacobi_update_noalloc_f64": (c_int, [c_int] + [c_void_p] * 8),
    "libp_pcg_gather_jacobi_update_f64": (c_int, [c_int] + [c_void_p] * 11),
    "libp_pcg_gather_jacobi_partials": (c_long, [c_int]),
    "libp_pcg_ax_hex_pap_f64": (c_int, [c_int, c_int] + [c_void_p] * 5 + [c_double] + [c_void_p] * 5),
    "libp_pcg_jacobi_update_noalloc_max_n": (c_long, []),
    "libp_pcg_relative_tol_f64": (c_int, [c_double, c_void_p, c_void_p]),
    "libp_pcg_rdotz_from_f32_f64": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p,
                                            c_void_p]),
    "libp_ogs_gather_residual_f64": (c_int, [c_int] + [c_void_p] * 6),
    "libp_ogs_gather_residual_f32": (c_int, [c_int] + [c_void_p] * 6),
    "libp_ogs_gather_add_f64": (c_int, [c_int] + [c_void_p] * 5),
    "libp_ogs_gather_add_f32": (c_int, [c_int] + [c_void_p] * 5),
    "libp_ogs_gather_cheb_start_f64": (c_int, [c_int] + [c_void_p] * 5 + [c_double] + [c_void_p] * 4),
    "libp_ogs_gather_cheb_start_f32": (c_int, [c_int] + [c_void_p] * 5 + [c_float] + [c_void_p] * 4),
    "libp_ogs_gather_cheb_step_f64": (c_int, [c_int] + [c_void_p] * 4 + [c_double, c_double] + [c_void_p] * 4),
    "libp_ogs_gather_cheb_step_f32": (c_int, [c_int] + [c_void_p] * 4 + [c_float, c_float] + [c_void_p] * 4),
    "libp_mg_coarsen_weighted_hex_f64": (c_int, [c_int, c_int, c_int] + [c_void_p] * 7),
    "libp_mg_coarsen_weighted_hex_f32": (c_int, [c_int, c_int, c_int] + [c_void_p] * 7),
    "libp_learned_assemble_add_f32": (c_int, [c_int, c_void_p, c_void_p, c_int, c_void_p, c_long] + [c_void_p] * 6),
    "libp_nbpcg_update1_f64": (c_int, [c_int] + [c_void_p] * 7),
    "libp_nbpcg_update2_f64": (c_int, [c_int, c_int] + [c_void_p] * 7),
    "libp_axpy_dev_f64": (c_int, [c_int] + [c_void_p] * 4),
    "libp_nbfpcg_update0_f64": (c_int, [c_int] + [c_void_p] * 6),
    "libp_nbfpcg_update1_f64": (c_int, [c_int] + [c_void_p] * 11),
    "libp_nbfpcg_update2_f64": (c_int, [c_int] + [c_void_p] * 10),
    "libp_mlp_linear_fwd": (c_int, [c_int, c_int, c_int, c_void_p, c_long, c_void_p, c_long, c_void_p, c_int,
                                    c_void_p, c_long, c_void_p]),
    "libp_mlp_linear_fwd_scaled": (c_int, [c_int, c_int, c_int, c_void_p, c_long, c_void_p, c_long, c_void_p,
                                           c_void_p, c_void_p, c_long, c_void_p]),
    "libp_mlp_linear_bwd_data": (c_int, [c_int, c_int, c_int, c_void_p, c_long, c_void_p, c_long, c_void_p, c_long,
                                         c_void_p, c_long, c_void_p]),
    "libp_mlp_linear_bwd_weight": (c_int, [c_int, c_int, c_int, c_void_p, c_long, c_void_p, c_long, c_void_p,
                                           c_void_p, c_void_p, c_void_p]),
    "libp_mlp_ws_bytes": (c_long, [c_int, c_int, c_int]),
    "libp_adam_step": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_float, c_float, c_float,
                               c_float, c_int, c_void_p]),
    "libp_f32_to_bf16": (c_int, [c_long, c_void_p, c_void_p, c_void_p]),
    "libp_learned_features_f32": (c_int, [c_int, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                          c_long, c_void_p, c_void_p]),
    "libp_learned_pack_weights_bf16": (c_int, [c_int, c_int, c_void_p, c_long, c_void_p, c_void_p]),
    "libp_learned_linear_layer_f32": (c_int, [c_int, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p,
                                              c_void_p, c_void_p, c_long, c_void_p, c_void_p]),
    "libp_learned_assemble_f32": (c_int, [c_int, c_void_p, c_void_p, c_int, c_void_p, c_long, c_void_p, c_void_p,
                                          c_void_p, c_void_p, c_void_p, c_void_p]),
    "libp_learned_alpha_f32": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "libp_learned_update_f32": (c_int, [c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "libp_learned_loss_f32": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                      c_void_p, c_void_p, c_void_p]),
    "libp_learned_scatter_grad_f32": (c_int, [c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_long,
                                              c_void_p]),
    "libp_learned_element_target_f32": (c_int, [c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_long,
                                                c_void_p]),
    "libp_random_uniform_f32": (c_int, [c_long, ctypes.c_ulonglong, c_void_p, c_void_p]),
    "libp_inv_row_length_f32": (c_int, [c_int, c_void_p, c_void_p, c_void_p]),
    "libp_elliptic_create": (c_void_p, [c_char_p]),
    "libp_elliptic_free": (None, [c_void_p]),
    "libp_elliptic_geti": (c_long, [c_void_p, c_char_p]),
    "libp_elliptic_getd": (c_double, [c_void_p, c_char_p]),
    "libp_elliptic_array": (c_void_p, [c_void_p, c_char_p, ctypes.POINTER(c_long), ctypes.POINTER(c_int),
                                       ctypes.POINTER(c_int)]),
    "libp_elliptic_stream": (c_void_p, [c_void_p]),
    "libp_elliptic_operator": (c_int, [c_void_p, c_void_p, c_void_p]),
    "libp_elliptic_partial_ax": (c_int, [c_void_p, c_void_p, c_void_p]),
    "libp_elliptic_precon": (c_int, [c_void_p, c_void_p, c_void_p]),
    "libp_elliptic_learned_train": (c_int, [c_void_p, c_char_p, ctypes.POINTER(c_double), c_int]),
    "libp_elliptic_learned_load": (c_int, [c_void_p, c_char_p]),
    "libp_elliptic_learned_save": (c_int, [c_void_p, c_char_p]),
    "libp_elliptic_learned_info": (c_int, [c_void_p, ctypes.POINTER(c_double), c_int]),
    "libp_elliptic_learned_status": (c_char_p, [c_void_p]),
    "libp_elliptic_learned_get_params": (c_int, [c_void_p, c_void_p, c_long]),
    "libp_elliptic_learned_set_params": (c_int, [c_void_p, c_void_p, c_long, c_int]),
    "libp_elliptic_learned_correction": (c_int, [c_void_p, c_void_p, c_void_p]),
    "libp_elliptic_learned_loss_grad": (c_int, [c_void_p, c_void_p, c_void_p, c_double, ctypes.POINTER(c_double)]),
    "libp_elliptic_learned_linear_gram": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "libp_csr_spmv_f64": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_double, c_void_p, c_double, c_void_p,
                                  c_void_p, c_int, c_void_p]),
    "libp_dense_gemv_f64": (c_int, [c_int, c_void_p, c_void_p, c_void_p, c_void_p]),
    "libp_elliptic_csr_nnz": (c_long, [c_void_p]),
    "libp_elliptic_csr": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "libp_elliptic_amg_setup": (c_int, [c_void_p]),
    "libp_elliptic_amg_vcycle": (c_int, [c_void_p, c_void_p, c_void_p]),
    "libp_elliptic_amg_info": (c_int, [c_void_p, c_void_p, c_int]),
    "libp_elliptic_amg_level_array": (c_int, [c_void_p, c_int, c_int, c_char_p, c_void_p, ctypes.POINTER(c_long),
                                              ctypes.POINTER(c_int)]),
    "libp_elliptic_amg_smooth": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_int]),
    "libp_elliptic_mg_level_ndofs": (c_long, [c_void_p, c_int, ctypes.POINTER(c_int)]),
    "libp_elliptic_mg_level_op": (c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p, c_int]),
    "libp_elliptic_mg_level_lambda": (c_int, [c_void_p, c_int, ctypes.POINTER(c_double)]),
    "libp_elliptic_run": (c_int, [c_void_p, ctypes.POINTER(c_double), c_int]),
    "libp_elliptic_solution": (c_int, [c_void_p, c_void_p]),
}


class LibpError(RuntimeError):
    pass


def lib():
    """The loaded libp_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise LibpError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                            "or `make -C libparanumal_amd`")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(status, what=""):
    if status != 0:
        L = lib()
        raise LibpError(f"{what}: {L.libp_status_string(status).decode()} ({L.libp_last_error().decode()})")


def settings_text(settings):
    """Render a dict of settings in the .rc format (test/test.py:84-92)."""
    return "".join(f"[{k}]\n{v}\n\n" for k, v in settings.items())


_DTYPES = {0: "float64", 1: "int32", 2: "int64", 3: "float32"}


class Elliptic:
    """elliptic_t built from settings, driven through the C ABI."""

    def __init__(self, settings):
        text = settings if isinstance(settings, str) else settings_text(settings)
        self._h = lib().libp_elliptic_create(text.encode())
        if not self._h:
            raise LibpError("libp_elliptic_create: " + lib().libp_last_error().decode())

    def close(self):
        if self._h:
            lib().libp_elliptic_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def geti(self, name):
        return int(lib().libp_elliptic_geti(self._h, name.encode()))

    def getd(self, name):
        return float(lib().libp_elliptic_getd(self._h, name.encode()))

    @property
    def stream(self):
        return lib().libp_elliptic_stream(self._h)

    def array_info(self, name):
        n, dt, host = c_long(), c_int(), c_int()
        p = lib().libp_elliptic_array(self._h, name.encode(), ctypes.byref(n), ctypes.byref(dt), ctypes.byref(host))
        if not p:
            raise LibpError(f"no array named {name}")
        return p, n.value, _DTYPES[dt.value], bool(host.value)

    def host_array(self, name):
        """Copy of a named array as a numpy array (device arrays are copied back)."""
        import numpy as np
        p, n, dt, on_host = self.array_info(name)
        out = np.empty(n, dtype=dt)
        if on_host:
            ctypes.memmove(out.ctypes.data, p, out.nbytes)
        else:
            hip_d2h(out.ctypes.data, p, out.nbytes, self.stream)
        return out

    def operator(self, q, Aq):
        check(lib().libp_elliptic_operator(self._h, q.data_ptr(), Aq.data_ptr()), "operator")

    def partial_ax(self, q, AqL):
        check(lib().libp_elliptic_partial_ax(self._h, q.data_ptr(), AqL.data_ptr()), "partial_ax")

    def precon(self, r, z):
        check(lib().libp_elliptic_precon(self._h, r.data_ptr(), z.data_ptr()), "precon")

    def run(self):
        out = (c_double * 6)()
        check(lib().libp_elliptic_run(self._h, out, 6), "run")
        return {"norm": out[0], "iterations": int(out[1]), "seconds": out[2], "r_final": out[3], "r0": out[4],
                "error_l2": out[5]}

    def solution(self):
        import numpy as np
        x = np.empty(self.geti("Nelements") * self.geti("Np"))
        check(lib().libp_elliptic_solution(self._h, x.ctypes.data), "solution")
        return x

    # learned fine smoother (MULTIGRID SMOOTHER LEARNED, MULTIGRID PRECISION FP32)
    def learned_train(self, **options):
        """Train the learned smoother; options as in libp_elliptic_learned_train.
        Returns the mean loss of each logged block of steps."""
        text = " ".join(f"{k}={','.join(map(str, v)) if isinstance(v, (list, tuple)) else v}"
                        for k, v in options.items())
        out = (c_double * 4096)()
        n = lib().libp_elliptic_learned_train(self._h, text.encode(), out, 4096)
        if n < 0:
            check(-n, "learned_train")
        return [out[i] for i in range(min(n, 4096))]

    def learned_load(self, path):
        """True when the weights were accepted; False (Chebyshev stays in use) otherwise."""
        return lib().libp_elliptic_learned_load(self._h, str(path).encode()) == 0

    def learned_save(self, path):
        check(lib().libp_elliptic_learned_save(self._h, str(path).encode()), "learned_save")

    def learned_info(self):
        out = (c_double * 10)()
        check(lib().libp_elliptic_learned_info(self._h, out, 10), "learned_info")
        keys = ("active", "fallback", "rejected_steps", "parameters", "alpha", "loss", "inputs", "hidden", "arch",
                "sweeps")
        d = {k: out[i] for i, k in enumerate(keys)}
        d["arch"] = "LINEAR" if d["arch"] == 1 else "MLP"
        return d

    def assembled_csr(self):
        """The fine operator assembled as CSR (numpy rowStarts, cols, vals); N <= 3."""
        import numpy as np
        nnz = lib().libp_elliptic_csr_nnz(self._h)
        if nnz < 0:
            check(-nnz, "csr_nnz")
        n = self.geti("Ngather")
        rs = np.empty(n + 1, dtype=np.int32); ci = np.empty(nnz, dtype=np.int32); v = np.empty(nnz)
        check(lib().libp_elliptic_csr(self._h, rs.ctypes.data, ci.ctypes.data, v.ctypes.data), "csr")
        return rs, ci, v

    def amg_setup(self):
        check(lib().libp_elliptic_amg_setup(self._h), "amg_setup")

    def amg_vcycle(self, b, x):
        check(lib().libp_elliptic_amg_vcycle(self._h, b.data_ptr(), x.data_ptr()), "amg_vcycle")

    def amg_info(self):
        out = (c_double * 64)()
        check(lib().libp_elliptic_amg_info(self._h, out, 64), "amg_info")
        L = int(out[0])
        return {"levels": L, "rows": [int(out[1 + 2 * l]) for l in range(L)],
                "nnz": [int(out[2 + 2 * l]) for l in range(L)]}

    def amg_level_array(self, level, name, which=0):
        """Host copy (numpy) of a named array of AMG level `level` (which: 0 the
        amg_setup AMG, 1 the p-MG base level)."""
        import numpy as np
        n, dt = c_long(), c_int()
        check(lib().libp_elliptic_amg_level_array(self._h, which, level, name.encode(), None, ctypes.byref(n),
                                                  ctypes.byref(dt)), "amg_level_array")
        out = np.empty(n.value, dtype=np.float64 if dt.value == 0 else np.int32)
        check(lib().libp_elliptic_amg_level_array(self._h, which, level, name.encode(), out.ctypes.data,
                                                  ctypes.byref(n), ctypes.byref(dt)), "amg_level_array")
        return out

    def amg_smooth(self, level, b, x, x_is_zero):
        check(lib().libp_elliptic_amg_smooth(self._h, level, b.data_ptr(), x.data_ptr(), int(x_is_zero)),
              "amg_smooth")

    def learned_linear_gram(self, r, eh):
        """Gram matrices (G = X^T X, B = T^T X; numpy FP64) of one least-squares sample of the
        linear smoother, from device FP32 gathered r and eh."""
        import numpy as np
        info = self.learned_info()
        nIn, Np = int(info["inputs"]), self.geti("Np")
        G = np.empty((nIn, nIn)); B = np.empty((Np, nIn))
        check(lib().libp_elliptic_learned_linear_gram(self._h, r.data_ptr(), eh.data_ptr(), G.ctypes.data,
                                                      B.ctypes.data), "learned_linear_gram")
        return G, B

    # p-MG level access (libp_elliptic_mg_level_*): operations on device vectors
    MG_OPERATOR, MG_COARSEN, MG_PROLONGATE, MG_SMOOTH, MG_INVDIAG = range(5)

    def mg_level_ndofs(self, level):
        bits = c_int()
        n = lib().libp_elliptic_mg_level_ndofs(self._h, level, ctypes.byref(bits))
        if n < 0:
            check(-n, "mg_level_ndofs")
        return int(n), bits.value

    def mg_level_op(self, level, op, vin, vout, flag=0):
        check(lib().libp_elliptic_mg_level_op(self._h, level, op, vin.data_ptr() if vin is not None else None,
                                              vout.data_ptr(), int(flag)), "mg_level_op")

    def mg_level_lambda(self, level):
        out = (c_double * 2)()
        check(lib().libp_elliptic_mg_level_lambda(self._h, level, out), "mg_level_lambda")
        return out[0], out[1]

    def learned_status(self):
        return lib().libp_elliptic_learned_status(self._h).decode()

    def learned_params(self):
        import numpy as np
        n = int(self.learned_info()["parameters"])
        th = np.empty(n, dtype=np.float32)
        check(lib().libp_elliptic_learned_get_params(self._h, th.ctypes.data, n), "learned_get_params")
        return th

    def learned_set_params(self, theta, activate=True):
        import numpy as np
        th = np.ascontiguousarray(theta, dtype=np.float32)
        check(lib().libp_elliptic_learned_set_params(self._h, th.ctypes.data, th.size, int(activate)),
              "learned_set_params")

    def learned_correction(self, r, c):
        check(lib().libp_elliptic_learned_correction(self._h, r.data_ptr(), c.data_ptr()), "learned_correction")

    def learned_loss_grad(self, r, eh, nrm):
        loss = c_double()
        check(lib().libp_elliptic_learned_loss_grad(self._h, r.data_ptr(), eh.data_ptr(), float(nrm),
                                                    ctypes.byref(loss)), "learned_loss_grad")
        return loss.value


def hip_d2h(dst, src, nbytes, stream=None):
    """Synchronise `stream` and copy device memory to host with the HIP runtime
    that libp_amd.so already loaded."""
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipStreamSynchronize.restype = c_int
    hip.hipStreamSynchronize.argtypes = [c_void_p]
    hip.hipMemcpy.restype = c_int
    hip.hipMemcpy.argtypes = [c_void_p, c_void_p, ctypes.c_size_t, c_int]
    if stream:
        if hip.hipStreamSynchronize(c_void_p(stream)) != 0:
            raise LibpError("hipStreamSynchronize failed")
    st = hip.hipMemcpy(c_void_p(dst), c_void_p(src), ctypes.c_size_t(nbytes), 2)  # hipMemcpyDeviceToHost
    if st != 0:
        raise LibpError(f"hipMemcpy failed ({st})")
