"""ctypes binding of libpunc_gpu.so (the C ABI in include/punc_gpu.h).

Loading never falls back to anything else: if the library has not been built,
import of the operators raises.  Build it with ``python -m puncpp_amd.build``.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpunc_gpu.so")

_lib = None

d_p = C.POINTER(C.c_double)
i32_p = C.POINTER(C.c_int32)
i64_p = C.POINTER(C.c_int64)
u64_p = C.POINTER(C.c_uint64)
vp = C.c_void_p

# name -> (restype, argtypes)
_SIGS = {
    "punc_gpu_last_error": (C.c_char_p, []),
    "punc_gpu_abi_version": (C.c_int, []),
    "punc_gpu_meshgen_box": (C.c_int, [C.c_int, d_p, d_p, C.c_double, C.c_int, C.POINTER(vp)]),
    "punc_gpu_hostmesh_sizes": (C.c_int, [vp, i64_p, i64_p, i64_p]),
    "punc_gpu_hostmesh_copy": (C.c_int, [vp, d_p, i32_p, i32_p, i32_p]),
    "punc_gpu_hostmesh_free": (None, [vp]),
    "punc_gpu_mesh_create": (C.c_int, [C.c_int64, d_p, C.c_int64, i32_p, i32_p, i32_p, C.c_int64,
                                       C.POINTER(vp)]),
    "punc_gpu_mesh_destroy": (None, [vp]),
    "punc_gpu_mesh_info": (C.c_int, [vp, i64_p, i64_p, i64_p, i64_p, d_p]),
    "punc_gpu_mesh_localizer": (C.c_int, [vp, i64_p, d_p, d_p, d_p, i32_p]),
    "punc_gpu_element_volume": (C.c_int, [vp, C.c_int, d_p]),
    "punc_gpu_exterior_facets": (C.c_int, [vp, d_p, d_p, d_p, d_p, i32_p]),
    "punc_gpu_mesh_locate": (C.c_int, [vp, C.c_int64, d_p, i64_p]),
    "punc_internal_inside_tol": (C.c_int, [vp, d_p]),
    "punc_gpu_maxwellian_flux": (C.c_int, [vp, C.c_double, d_p, d_p, d_p]),
    "punc_gpu_vdf_info": (C.c_int, [C.c_int, C.c_double, d_p, C.c_double, C.c_double, C.c_double, d_p, d_p,
                                    i32_p, i32_p]),
    "punc_gpu_vdf_flux": (C.c_int, [vp, C.c_int, C.c_double, d_p, C.c_double, C.c_double, C.c_double, C.c_uint64,
                                    C.c_int, C.c_int64, d_p, d_p]),
    "punc_gpu_pop_load_vdf": (C.c_int, [vp, C.c_int, C.c_int64, C.c_uint64, C.c_double, C.c_double, C.c_int,
                                        C.c_int, C.c_double, d_p, C.c_double, C.c_double, C.c_double]),
    "punc_gpu_injector_set_vdf": (C.c_int, [vp, C.c_int, C.c_int, C.c_double, d_p, C.c_double, C.c_double,
                                            C.c_double, C.c_double, d_p, d_p]),
    "punc_gpu_pop_create": (C.c_int, [vp, C.c_int, d_p, d_p, C.c_int64, vp, C.POINTER(vp)]),
    "punc_gpu_pop_destroy": (None, [vp]),
    "punc_gpu_pop_add_species": (C.c_int, [vp, C.c_double, C.c_double, i32_p]),
    "punc_gpu_pop_add": (C.c_int, [vp, C.c_int, C.c_int64, d_p, d_p, i64_p]),
    "punc_gpu_pop_add_located": (C.c_int, [vp, C.c_int64, d_p, d_p, i32_p, i64_p]),
    "punc_gpu_pop_load": (C.c_int, [vp, C.c_int, C.c_int64, C.c_uint64, C.c_double, C.c_double,
                                    C.c_int, C.c_double, d_p]),
    "punc_gpu_pop_size": (C.c_int, [vp, i64_p]),
    "punc_gpu_pop_counts": (C.c_int, [vp, i64_p]),
    "punc_gpu_pop_get": (C.c_int, [vp, d_p, d_p, i32_p, i64_p]),
    "punc_gpu_pop_save_file": (C.c_int, [vp, C.c_char_p, C.c_int]),
    "punc_gpu_pop_species": (C.c_int, [vp, i32_p, d_p, d_p]),
    "punc_gpu_pop_load_file": (C.c_int, [vp, C.c_char_p, C.c_int, i64_p]),
    "punc_gpu_set_field": (C.c_int, [vp, d_p]),
    "punc_gpu_host_alloc": (C.c_int, [C.c_int64, C.POINTER(vp)]),
    "punc_gpu_host_free": (C.c_int, [vp]),
    "punc_gpu_accel": (C.c_int, [vp, C.c_double, d_p]),
    "punc_gpu_boris": (C.c_int, [vp, d_p, C.c_double, d_p]),
    "punc_gpu_move": (C.c_int, [vp, C.c_double]),
    "punc_gpu_move_periodic": (C.c_int, [vp, C.c_double, d_p]),
    "punc_gpu_update": (C.c_int, [vp, i64_p]),
    "punc_gpu_distribute": (C.c_int, [vp, d_p, d_p]),
    "punc_gpu_kinetic_energy": (C.c_int, [vp, d_p]),
    "punc_gpu_potential_energy": (C.c_int, [vp, d_p, d_p]),
    "punc_gpu_density_cg1": (C.c_int, [vp, d_p, C.c_double, C.c_double, d_p, d_p]),
    "punc_gpu_density_dg0": (C.c_int, [vp, C.c_double, C.c_double, d_p, d_p]),
    "punc_gpu_statistics": (C.c_int, [vp, d_p]),
    "punc_gpu_rho_ema": (C.c_int, [vp, C.c_double, C.c_double, C.c_int, d_p]),
    "punc_gpu_density_ema": (C.c_int, [vp, C.c_double, C.c_double, C.c_int, d_p, d_p]),
    "punc_gpu_sort": (C.c_int, [vp]),
    "punc_gpu_field_create": (C.c_int, [vp, C.c_double, C.c_int, C.c_int, C.c_int, C.POINTER(vp)]),
    "punc_gpu_field_destroy": (None, [vp]),
    "punc_gpu_field_set_max_iterations": (C.c_int, [vp, C.c_int]),
    "punc_gpu_field_set_warm_start": (C.c_int, [vp, C.c_int]),
    "punc_gpu_field_set_preconditioner": (C.c_int, [vp, C.c_int]),
    "punc_gpu_field_get_preconditioner": (C.c_int, [vp, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                                    C.POINTER(C.c_int64)]),
    "punc_gpu_poisson_solve": (C.c_int, [vp, d_p, d_p, d_p, d_p, i32_p]),
    "punc_gpu_efield_solve": (C.c_int, [vp, d_p, d_p, i32_p]),
    "punc_gpu_field_solve_pop": (C.c_int, [vp, vp, d_p, d_p, d_p, d_p, i32_p]),
    "punc_gpu_object_charges": (C.c_int, [vp, d_p, d_p]),
    "punc_gpu_circuit_set": (C.c_int, [vp, C.c_double, C.c_int, i32_p, i32_p, d_p, C.c_int, i32_p, i32_p, d_p, d_p]),
    "punc_gpu_poisson_solve_circuit": (C.c_int, [vp, d_p, d_p, d_p, d_p, i32_p]),
    "punc_gpu_field_solve_pop_circuit": (C.c_int, [vp, vp, d_p, d_p, d_p, d_p, i32_p]),
    "punc_gpu_injector_set": (C.c_int, [vp, C.c_int, C.c_double, d_p, C.c_double, C.c_double]),
    "punc_gpu_species_instance": (C.c_int, [vp, C.c_int, C.c_int, C.c_uint64, C.c_int64, C.c_int64,
                                            C.c_int64, C.c_int64]),
    "punc_gpu_inject": (C.c_int, [vp, C.c_double, C.c_uint64, C.c_uint64, i64_p]),
    "punc_gpu_step_graph": (C.c_int, [vp, C.c_int, C.c_double, d_p, C.c_double, d_p, C.c_int,
                                      C.c_uint64, C.c_uint64]),
    "punc_gpu_graph_launch": (C.c_int, [vp, C.c_int]),
    "punc_gpu_step": (C.c_int, [vp, C.c_int, C.c_double, d_p, C.c_double, d_p, C.c_int, C.c_uint64,
                                C.c_uint64]),
    "punc_gpu_step_results": (C.c_int, [vp, d_p, i64_p, i64_p]),
    "punc_gpu_near_face": (C.c_int, [vp, i64_p]),
    "punc_gpu_pop_digest": (C.c_int, [vp, u64_p, i64_p]),
    "punc_gpu_vdf_flux_facets": (C.c_int, [C.c_int64, d_p, d_p, C.c_int, C.c_double, d_p, C.c_double, C.c_double,
                                           C.c_double, C.c_uint64, C.c_int, C.c_int64, d_p, d_p]),
    "punc_gpu_ema": (C.c_int, [C.c_int64, d_p, d_p, C.c_double, C.c_double]),
    "punc_gpu_vdf_flux_facets_noalloc": (C.c_int, [C.c_int64, d_p, d_p, C.c_int, C.c_double, d_p, C.c_double,
                                                   C.c_double, C.c_double, C.c_uint64, C.c_int, C.c_int64, d_p,
                                                   d_p]),
    "punc_gpu_get_rho": (C.c_int, [vp, d_p]),
    "punc_gpu_set_potential": (C.c_int, [vp, d_p]),
    "punc_gpu_step_potential_energy": (C.c_int, [vp, d_p]),
    "punc_gpu_set_launch": (C.c_int, [vp, C.c_int, C.c_int, C.c_int]),
    "punc_gpu_get_launch": (C.c_int, [vp, i32_p, i32_p, i32_p]),
    "punc_gpu_set_inject_in_step": (C.c_int, [vp, C.c_int]),
    "punc_gpu_get_inject_in_step": (C.c_int, [vp, i32_p]),
    "punc_gpu_autotune": (C.c_int, [vp, C.c_int, C.c_double, d_p, C.c_double, d_p, C.c_int, C.c_uint64,
                                    C.c_int, i32_p, i32_p, i32_p, d_p]),
    "punc_gpu_timing": (C.c_int, [vp, C.c_int]),
    "punc_gpu_timing_read": (C.c_int, [vp, d_p, i64_p]),
    "punc_gpu_synchronize": (C.c_int, [vp]),
    "punc_gpu_copy_probe": (C.c_int, [C.c_int64, C.c_int, d_p]),
    "punc_gpu_adjoint_last_error": (C.c_char_p, []),
    "punc_gpu_adjoint_collected": (C.c_int, [vp, d_p, i32_p, i32_p, i64_p]),
    "punc_gpu_adjoint_backward_q": (C.c_int, [vp, d_p, d_p, d_p, d_p, d_p]),
    "punc_gpu_efield_adjoint": (C.c_int, [vp, d_p, d_p, i32_p]),
    "punc_gpu_adjoint_create": (C.c_int, [vp, C.c_int, d_p, d_p, C.c_int64, i32_p, C.POINTER(vp)]),
    "punc_gpu_adjoint_destroy": (None, [vp]),
    "punc_gpu_adjoint_forward": (C.c_int, [vp, d_p, i64_p, d_p, d_p, d_p, C.c_double, d_p, C.c_int, C.c_int,
                                           d_p, d_p, d_p, i64_p]),
    "punc_gpu_adjoint_backward": (C.c_int, [vp, d_p, d_p, d_p, d_p]),
}


class PuncGpuError(RuntimeError):
    pass


def lib():
    """Load libpunc_gpu.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise PuncGpuError(
                f"{LIB_PATH} is missing: build it with `python -m puncpp_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        raise PuncGpuError(lib().punc_gpu_last_error().decode())


def dptr(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(d_p)


def i32ptr(a):
    if a is None:
        return None
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(i32_p)


def i64ptr(a):
    if a is None:
        return None
    assert a.dtype == np.int64 and a.flags.c_contiguous
    return a.ctypes.data_as(i64_p)


def vec3(v):
    return np.ascontiguousarray(np.asarray(v, dtype=np.float64).reshape(3))


class PinnedArray:
    """A float64 numpy array in page-locked host memory (punc_gpu_host_alloc):
    set_field / get_rho with it copy by DMA.  Keep this object alive while
    ``.array`` is in use; the memory is freed with it."""

    def __init__(self, shape):
        n = int(np.prod(shape))
        self._ptr = vp()
        check(lib().punc_gpu_host_alloc(8 * max(n, 1), C.byref(self._ptr)))
        buf = (C.c_double * max(n, 1)).from_address(self._ptr.value)
        self.array = np.ctypeslib.as_array(buf)[:n].reshape(shape)

    def __del__(self):
        p = getattr(self, "_ptr", None)
        if p is not None and p.value:
            self.array = None
            lib().punc_gpu_host_free(p)
            self._ptr = None
