"""ctypes binding of libpypat_hip.so (include/pypat_hip.h).

The library is built in-tree (``python -m pypatools_amd.build``).  Loading fails
loudly when it is missing: there is no CPU fallback for the push or the gather.
"""
import ctypes
import os
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpypat_hip.so")

PP_OK, PP_ERR_INVALID, PP_ERR_HIP, PP_ERR_NOMEM, PP_ERR_UNSUPPORTED = 0, 1, 2, 3, 4

ALG_IDS = {
    "leapfrog": 0, "boris": 1, "rk4": 2, "yoshida": 3, "vay_rel": 4, "rk4_rel": 5, "yoshida_rel": 6,
    "boris_rel": 7, "vay": 8,
}
PRECISIONS = {"fp64": 0, "fp64_fast": 1, "fp32_mixed": 2}
FIELD_ZERO, FIELD_CONST, FIELD_MAP, FIELD_ARRAY = 0, 1, 2, 3
RF_NONE, RF_COS, RF_SIN = 0, 1, 2
MAP_PERIODIC_Z, MAP_NO_FP32, MAP_NO_COEF, MAP_NO_FP64 = 1, 2, 4, 8

c_double_p = ctypes.POINTER(ctypes.c_double)
c_int64 = ctypes.c_int64
c_int32 = ctypes.c_int32
c_uint32 = ctypes.c_uint32
c_void_p = ctypes.c_void_p


class PPField(ctypes.Structure):
    _fields_ = [
        ("kind", c_int32), ("comp0", c_int32), ("map", c_void_p), ("array", c_void_p),
        ("scale1", ctypes.c_double), ("scale2", ctypes.c_double), ("offset", ctypes.c_double),
        ("has_scale2", c_int32), ("has_offset", c_int32), ("rf", c_int32), ("_pad", c_int32),
        ("value", ctypes.c_double * 3),
    ]


class PPPushArgs(ctypes.Structure):
    _fields_ = [
        ("algorithm", c_int32), ("precision", c_int32), ("n", c_int64),
        ("x", c_void_p), ("y", c_void_p), ("z", c_void_p),
        ("vx", c_void_p), ("vy", c_void_p), ("vz", c_void_p),
        ("alive", c_void_p), ("dt_pp", c_void_p), ("ids", c_void_p),
        ("dt", ctypes.c_double), ("q_over_m", ctypes.c_double),
        ("nsteps", c_int64), ("step0", c_int64),
        ("efield", PPField), ("bfield", PPField),
        ("rf_table", c_void_p), ("rf_step0", c_int64), ("rf_nsteps", c_int64),
        ("r_hist", c_void_p), ("v_hist", c_void_p), ("rec_every", c_int64), ("n_records", c_int64),
        ("kick_only", c_int32), ("block_size", c_int32), ("hist_rows", c_int64),
    ]


class PPBunch(ctypes.Structure):
    _fields_ = [
        ("offset", c_int64), ("count", c_int64), ("algorithm", c_int32), ("nsteps", c_int32),
        ("map_index", c_int32), ("flags", c_int32), ("dt", ctypes.c_double), ("q_over_m", ctypes.c_double),
        ("escale", ctypes.c_double), ("bscale", ctypes.c_double),
    ]


SPLINE_NEAREST_MODE, SPLINE_RGI = 1, 2          # pp_spline_flags
TILES_ALL, TILES_SINGLE_GATHER, TILES_RK4 = 0, 1, 2   # pp_tile_class
RK4_ALGS = ("rk4", "rk4_rel")
MAX_BATCH_MAPS = 4
BUNCH_KICK_START, BUNCH_KICK_END = 1, 2          # pp_bunch_flags
MOMENT_SUMS, MOMENT_CENTRAL = 18, 11          # include/pypat_hip.h PP_MOMENT_SUMS, PP_MOMENT_CENTRAL


class PPBatchArgs(ctypes.Structure):
    _fields_ = [
        ("precision", c_int32), ("block_size", c_int32), ("n", c_int64),
        ("x", c_void_p), ("y", c_void_p), ("z", c_void_p), ("vx", c_void_p), ("vy", c_void_p), ("vz", c_void_p),
        ("bunches", c_void_p), ("nbunch", c_int32), ("nmaps", c_int32),
        ("tile_bunch", c_void_p), ("tile_start", c_void_p), ("ntiles", c_int64),
        ("step0", c_int64), ("nsteps", c_int64), ("step0_dev", c_void_p),
        ("maps", c_void_p * MAX_BATCH_MAPS), ("tile_class", c_int32), ("_pad", c_int32),
    ]


class HipError(RuntimeError):
    """A HIP runtime failure inside libpypat_hip.so."""


class UnsupportedError(NotImplementedError):
    pass


_lib = None
_lock = threading.Lock()

# name -> (restype, argtypes); every entry is an extern "C" symbol of include/pypat_hip.h
SIGNATURES = {
    "pp_init": (c_int32, [ctypes.c_int]),
    "pp_last_error": (ctypes.c_char_p, []),
    "pp_abi_version": (c_int32, []),
    "pp_kernel_digest": (ctypes.c_char_p, []),
    "pp_get_constants": (c_int32, [c_double_p, c_int32]),
    "pp_struct_sizes": (c_int32, [ctypes.POINTER(c_int64), ctypes.POINTER(c_int64)]),
    "pp_batch_struct_sizes": (c_int32, [ctypes.POINTER(c_int64), ctypes.POINTER(c_int64)]),
    "pp_device_count": (c_int32, [ctypes.POINTER(ctypes.c_int)]),
    "pp_mem_info": (c_int32, [ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)]),
    "pp_malloc": (c_int32, [ctypes.POINTER(c_void_p), ctypes.c_size_t]),
    "pp_free": (c_int32, [c_void_p]),
    "pp_memcpy_h2d": (c_int32, [c_void_p, c_void_p, ctypes.c_size_t, c_void_p]),
    "pp_memcpy_d2h": (c_int32, [c_void_p, c_void_p, ctypes.c_size_t, c_void_p]),
    "pp_memcpy_d2d": (c_int32, [c_void_p, c_void_p, ctypes.c_size_t, c_void_p]),
    "pp_fill_f64": (c_int32, [c_void_p, ctypes.c_double, c_int64, c_void_p]),
    "pp_stream_create": (c_int32, [ctypes.POINTER(c_void_p)]),
    "pp_stream_destroy": (c_int32, [c_void_p]),
    "pp_stream_sync": (c_int32, [c_void_p]),
    "pp_device_sync": (c_int32, []),
    "pp_event_create": (c_int32, [ctypes.POINTER(c_void_p)]),
    "pp_event_destroy": (c_int32, [c_void_p]),
    "pp_event_record": (c_int32, [c_void_p, c_void_p]),
    "pp_event_elapsed_ms": (c_int32, [ctypes.POINTER(ctypes.c_float), c_void_p, c_void_p]),
    "pp_spline_create": (c_int32, [c_void_p, c_int32, c_int32, c_void_p, c_void_p, c_void_p, c_void_p, c_int32,
                                   c_int32, ctypes.c_double, ctypes.POINTER(c_void_p)]),
    "pp_spline_destroy": (c_int32, [c_void_p]),
    "pp_spline_eval": (c_int32, [c_void_p, c_void_p, c_int64, c_int32, c_int32, c_void_p, c_void_p]),
    "pp_host_alloc": (c_int32, [ctypes.POINTER(c_void_p), ctypes.c_size_t]),
    "pp_host_free": (c_int32, [c_void_p]),
    "pp_memcpy_h2d_async": (c_int32, [c_void_p, c_void_p, ctypes.c_size_t, c_void_p]),
    "pp_memcpy_d2h_async": (c_int32, [c_void_p, c_void_p, ctypes.c_size_t, c_void_p]),
    "pp_event_query": (c_int32, [c_void_p, ctypes.POINTER(c_int32)]),
    "pp_event_sync": (c_int32, [c_void_p]),
    "pp_map_create": (c_int32, [c_void_p, c_int32, c_int32, c_int32, c_int32, c_void_p, c_void_p, c_void_p,
                                ctypes.c_double, c_int32, ctypes.POINTER(c_void_p)]),
    "pp_map_destroy": (c_int32, [c_void_p]),
    "pp_map_bytes": (c_int32, [c_void_p, ctypes.POINTER(ctypes.c_size_t)]),
    "pp_field_eval": (c_int32, [ctypes.POINTER(PPField), c_int32, c_void_p, c_int64, ctypes.c_double, c_void_p,
                                c_void_p]),
    "pp_push": (c_int32, [ctypes.POINTER(PPPushArgs), c_void_p]),
    "pp_boris_push_batch": (c_int32, [c_void_p, c_void_p, c_void_p, c_int64, ctypes.c_double, ctypes.c_double,
                                      c_void_p, c_void_p]),
    "pp_vay_push_batch": (c_int32, [c_void_p, c_void_p, c_void_p, c_int64, ctypes.c_double, ctypes.c_double,
                                    c_void_p, c_void_p]),
    "pp_rk4_rel_dbetagamma_dt_batch": (c_int32, [c_void_p, c_void_p, c_void_p, c_int64, ctypes.c_double,
                                                 c_void_p, c_void_p]),
    "pp_position_update_batch": (c_int32, [c_void_p, c_void_p, c_int64, ctypes.c_double, c_void_p, c_void_p]),
    "pp_lorentz_batch": (c_int32, [c_void_p, c_void_p, c_void_p, c_int64, ctypes.c_double, c_void_p, c_void_p]),
    "pp_axpy_batch": (c_int32, [c_void_p, ctypes.c_double, c_void_p, c_void_p, c_int64, c_void_p, c_void_p]),
    "pp_rk4_combine_batch": (c_int32, [c_void_p, ctypes.c_double, c_void_p, c_void_p, c_void_p, c_void_p,
                                       c_void_p, c_int64, c_void_p, c_void_p]),
    "pp_clamp_speed_batch": (c_int32, [c_void_p, c_int64, c_void_p, c_void_p]),
    "pp_scale_offset_batch": (c_int32, [c_void_p, ctypes.c_double, c_int32, ctypes.c_double, c_int32, c_int64,
                                        c_void_p, c_void_p]),
    "pp_sort_workspace_bytes": (c_int32, [c_int64, ctypes.POINTER(ctypes.c_size_t)]),
    "pp_sort_state": (c_int32, [ctypes.POINTER(PPPushArgs), c_void_p, c_void_p, ctypes.c_size_t, c_void_p]),
    "pp_unsort_state": (c_int32, [ctypes.POINTER(PPPushArgs), c_void_p, c_void_p, ctypes.c_size_t, c_void_p]),
    "pp_iota_i32": (c_int32, [c_void_p, c_int64, c_void_p]),
    "pp_batch_tiles": (c_int32, [c_void_p, c_int32, c_int32, c_void_p, c_void_p, ctypes.POINTER(c_int64)]),
    "pp_push_batched": (c_int32, [c_void_p, c_void_p]),
    "pp_sort_state_grouped": (c_int32, [ctypes.POINTER(PPPushArgs), c_void_p, c_void_p, c_int32, c_void_p, c_void_p,
                                        ctypes.c_size_t, c_void_p]),
    "pp_sort_state_ex": (c_int32, [ctypes.POINTER(PPPushArgs), c_void_p, c_void_p, c_int32, c_int32, c_int32,
                                   c_void_p, c_void_p, ctypes.c_size_t, c_void_p]),
    "pp_graph_begin": (c_int32, [c_void_p]),
    "pp_graph_end": (c_int32, [c_void_p, ctypes.POINTER(c_void_p)]),
    "pp_graph_launch": (c_int32, [c_void_p, c_void_p]),
    "pp_graph_destroy": (c_int32, [c_void_p]),
    "pp_add_i64": (c_int32, [c_void_p, c_int64, c_void_p]),
    "pp_aos_to_soa": (c_int32, [c_void_p, c_int64, c_void_p, c_void_p, c_void_p, c_void_p]),
    "pp_gather_probe": (c_int32, [c_void_p, c_int32, c_uint32, c_int32, c_int64, c_void_p, c_void_p]),
    "pp_beam_workspace_bytes": (c_int32, [ctypes.POINTER(ctypes.c_size_t)]),
    "pp_poisson_create": (c_int32, [c_int32, c_int32, c_int32, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                    ctypes.POINTER(c_void_p)]),
    "pp_poisson_destroy": (c_int32, [c_void_p]),
    "pp_cic_deposit": (c_int32, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, ctypes.c_double, c_void_p, c_int64,
                                 c_void_p, c_void_p]),
    "pp_poisson_solve": (c_int32, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "pp_poisson_set_conductors": (c_int32, [c_void_p, c_void_p, c_int64, c_void_p, c_int64]),
    "pp_poisson_set_ray_data": (c_int32, [c_void_p, c_void_p, c_void_p]),
    "pp_poisson_conductor_data": (c_int32, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "pp_poisson_set_tolerance": (c_int32, [c_void_p, ctypes.c_double, c_int32]),
    "pp_poisson_solve_stats": (c_int32, [c_void_p, ctypes.POINTER(c_int32), ctypes.POINTER(ctypes.c_double)]),
    "pp_efield_from_potential": (c_int32, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "pp_mesh_create": (c_int32, [c_void_p, c_int64, c_void_p, c_int64, ctypes.POINTER(c_void_p)]),
    "pp_mesh_destroy": (c_int32, [c_void_p]),
    "pp_mesh_info": (c_int32, [c_void_p, ctypes.POINTER(c_int64), ctypes.POINTER(c_int64)]),
    "pp_collide_segments": (c_int32, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int64,
                                      c_void_p, c_void_p, c_void_p]),
    "pp_push_lossy": (c_int32, [ctypes.POINTER(PPPushArgs), c_void_p, c_void_p]),
    "pp_compact_state": (c_int32, [ctypes.POINTER(PPPushArgs), c_void_p, c_int32, c_int32, c_void_p,
                                   ctypes.POINTER(c_int64), c_void_p, ctypes.c_size_t, c_void_p]),
    "pp_beam_moments": (c_int32, [ctypes.POINTER(PPPushArgs), c_int32, ctypes.c_double, c_int32, c_void_p,
                                  c_void_p, c_void_p, ctypes.c_size_t, c_void_p]),
    "pp_beam_ellipse_counts": (c_int32, [ctypes.POINTER(PPPushArgs), c_int32, c_void_p, c_void_p, c_void_p,
                                         ctypes.c_size_t, c_void_p]),
    "pp_beam_moments_noalloc": (c_int32, [ctypes.POINTER(PPPushArgs), c_int32, ctypes.c_double, c_int32, c_void_p,
                                          c_void_p, c_void_p]),
    "pp_beam_noalloc_scratch_bytes": (c_int32, [ctypes.POINTER(ctypes.c_size_t)]),
    "pp_soa_to_aos": (c_int32, [c_void_p, c_void_p, c_void_p, c_int64, c_void_p, c_void_p]),
    "pp_v_to_u": (c_int32, [c_void_p, c_void_p, c_void_p, c_int64, c_void_p]),
    "pp_u_to_v": (c_int32, [c_void_p, c_void_p, c_void_p, c_int64, c_void_p]),
}


def lib():
    """Load libpypat_hip.so once; raise ImportError if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise ImportError(f"{LIB_PATH} is missing: build it with `python -m pypatools_amd.build` "
                                  "(there is no CPU fallback for the HIP push/gather path)")
            h = ctypes.CDLL(LIB_PATH, mode=ctypes.RTLD_GLOBAL)
            for name, (res, args) in SIGNATURES.items():
                fn = getattr(h, name)
                fn.restype = res
                fn.argtypes = args
            _lib = h
    return _lib


def check(rc):
    """Turn a PP_* return code into the matching Python exception."""
    if rc == PP_OK:
        return
    msg = lib().pp_last_error().decode(errors="replace")
    if rc == PP_ERR_INVALID:
        raise ValueError(msg)
    if rc == PP_ERR_UNSUPPORTED:
        raise UnsupportedError(msg)
    if rc == PP_ERR_NOMEM:
        raise MemoryError(msg)
    raise HipError(msg)


def call(name, *args):
    check(getattr(lib(), name)(*args))


def kernel_digest():
    """sha256 of the fast-path kernel sources the loaded library was built from."""
    return lib().pp_kernel_digest().decode()


def layout_ok():
    """True if the ctypes structures match the compiled C layout."""
    a, b = c_int64(), c_int64()
    check(lib().pp_struct_sizes(ctypes.byref(a), ctypes.byref(b)))
    c, d = c_int64(), c_int64()
    check(lib().pp_batch_struct_sizes(ctypes.byref(c), ctypes.byref(d)))
    return (a.value == ctypes.sizeof(PPField) and b.value == ctypes.sizeof(PPPushArgs)
            and c.value == ctypes.sizeof(PPBunch) and d.value == ctypes.sizeof(PPBatchArgs))


def constants():
    out = (ctypes.c_double * 9)()
    check(lib().pp_get_constants(out, 9))
    return list(out)
