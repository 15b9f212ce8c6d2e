"""ctypes declarations of include/ela3d.h and the library loader.

The library is the in-tree ``ela3d_amd/libela3d.so`` built by ``ela3d_amd.build``
(ELA3D_LIB names another build of it, such as the sanitized one).  There is no
CPU fallback: if the library is missing, every call raises.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("ELA3D_LIB") or os.path.join(HERE, "libela3d.so")

ABI_VERSION = 1
E_NAMES = {
    0: "OK", -1: "E_INVALID", -2: "E_OOM", -3: "E_UNSTABLE", -4: "E_HIP", -5: "E_NODEVICE",
    -6: "E_RANGE", -7: "E_STATE", -8: "E_BUDGET", -9: "E_NOTFOUND", -10: "E_TIMEOUT", -11: "E_IO",
}


class ConfigStruct(C.Structure):
    _fields_ = [
        ("struct_size", C.c_uint32), ("abi_version", C.c_uint32),
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double), ("dt", C.c_double),
        ("nstep", C.c_int32), ("npoints_pml", C.c_int32),
        ("npower", C.c_double), ("k_max_pml", C.c_double), ("alpha_max_pml", C.c_double),
        ("rcoef", C.c_double), ("vpml", C.c_double),
        ("f0", C.c_double), ("t0", C.c_double), ("factor", C.c_double), ("angle_force", C.c_double),
        ("src_kind", C.c_int32), ("wavelet", C.c_int32),
        ("isrc", C.c_int32), ("jsrc", C.c_int32), ("ksrc", C.c_int32),
        ("nrec", C.c_int32), ("rec", C.POINTER(C.c_int32)),
        ("sample_every", C.c_int32), ("energy_every", C.c_int32), ("storage", C.c_int32),
        ("compat_flags", C.c_uint32), ("dirichlet", C.c_int32), ("fp16_rescale_every", C.c_int32),
        ("device", C.c_int32), ("use_graph", C.c_int32), ("block", C.c_int32), ("zseg", C.c_int32),
        ("free_surface", C.c_int32), ("reserved", C.c_int32 * 5),
    ]


class InfoStruct(C.Structure):
    _fields_ = [
        ("cells", C.c_int64), ("device_bytes", C.c_int64),
        ("alg_bytes_per_step", C.c_double), ("alg_bytes_stress", C.c_double),
        ("alg_bytes_velocity", C.c_double), ("pml_bytes_per_step", C.c_double),
        ("nsamples", C.c_int32), ("nenergy", C.c_int32), ("steps_done", C.c_int32),
        ("block", C.c_int32), ("zseg", C.c_int32), ("zseg_slab", C.c_int32), ("vec", C.c_int32),
        ("region_lo", C.c_int32 * 3), ("region_hi", C.c_int32 * 3),
        ("block_slab", C.c_int32), ("vec_slab", C.c_int32), ("overlap", C.c_int32),
        ("merged", C.c_int32), ("dbuf", C.c_int32), ("fused", C.c_int32), ("tile", C.c_int32),
        ("strips", C.c_int32),
    ]


class KernelStats(C.Structure):
    _fields_ = [("name", C.c_char * 48), ("launches", C.c_int64), ("total_ms", C.c_double),
                ("alg_bytes_per_launch", C.c_double)]


class TuningStruct(C.Structure):
    _fields_ = [("block", C.c_int32), ("zseg_interior", C.c_int32), ("zseg_slab", C.c_int32),
                ("vec", C.c_int32), ("candidates", C.c_int32), ("ms_per_step", C.c_double),
                ("block_slab", C.c_int32), ("vec_slab", C.c_int32), ("overlap", C.c_int32),
                ("merged", C.c_int32), ("dbuf", C.c_int32), ("fused", C.c_int32), ("tile", C.c_int32),
                ("strips", C.c_int32)]


class LayeredSpec(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("h", C.c_double),
        ("nlayers", C.c_int32), ("vp_top", C.c_double), ("vp_bottom", C.c_double),
        ("noise", C.c_double), ("seed", C.c_uint64), ("bump_amp", C.c_double),
        ("bump_width", C.c_double), ("vp_vs", C.c_double), ("gardner_a", C.c_double),
        ("gardner_b", C.c_double),
    ]


class ServerOpts(C.Structure):
    _fields_ = [
        ("struct_size", C.c_uint32), ("device", C.c_int32), ("hbm_budget_bytes", C.c_int64),
        ("max_active", C.c_int32), ("steps_per_round", C.c_int32), ("use_graph", C.c_int32),
        ("plain_loop", C.c_int32), ("batch_cells", C.c_int32), ("reserved", C.c_int32 * 7),
    ]


class ResultStruct(C.Structure):
    _fields_ = [
        ("ticket", C.c_uint64), ("status", C.c_int32), ("nsamples", C.c_int32), ("nrec", C.c_int32),
        ("nenergy", C.c_int32), ("t_submit", C.c_double), ("t_admit", C.c_double),
        ("t_done", C.c_double), ("device_bytes", C.c_int64),
    ]


_lib = None


def lib() -> C.CDLL:
    """Load libela3d.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"ela3d native library not found at {LIB_PATH}: build it with "
            "`python -m ela3d_amd.build` (there is no CPU fallback)")
    L = C.CDLL(LIB_PATH)
    P = C.POINTER
    f = C.c_float
    L.ela3d_config_default.argtypes = [P(ConfigStruct)]
    L.ela3d_config_default.restype = None
    L.ela3d_validate.argtypes = [P(ConfigStruct)]
    L.ela3d_create.argtypes = [P(ConfigStruct), P(f), P(f), P(f), P(C.c_void_p)]
    L.ela3d_step.argtypes = [C.c_void_p, C.c_int32]
    L.ela3d_sync.argtypes = [C.c_void_p]
    L.ela3d_run.argtypes = [C.c_void_p]
    L.ela3d_seismograms.argtypes = [C.c_void_p, P(C.c_double), C.c_size_t]
    L.ela3d_energy.argtypes = [C.c_void_p, P(C.c_double), C.c_size_t]
    L.ela3d_snapshot.argtypes = [C.c_void_p, C.c_int32, P(C.c_double), C.c_size_t]
    L.ela3d_set_field.argtypes = [C.c_void_p, C.c_int32, P(C.c_double), C.c_size_t]
    L.ela3d_info_get.argtypes = [C.c_void_p, P(InfoStruct)]
    L.ela3d_profile.argtypes = [C.c_void_p, C.c_int32, P(KernelStats), C.c_int32, P(C.c_int32),
                                P(C.c_double)]
    L.ela3d_profile_sampled.argtypes = [C.c_void_p, C.c_int32, C.c_int32, P(KernelStats), C.c_int32,
                                        P(C.c_int32), P(C.c_double)]
    L.ela3d_run_noalloc.argtypes = [P(ConfigStruct), P(f), P(f), P(f), P(C.c_double), C.c_size_t,
                                    P(C.c_double), C.c_size_t, C.c_void_p]
    L.ela3d_create_slab.argtypes = [P(ConfigStruct), C.c_int32, C.c_int32, P(f), P(f), P(f), P(C.c_void_p),
                                    P(C.c_int32), C.c_size_t, P(C.c_int32)]
    L.ela3d_step_phase.argtypes = [C.c_void_p, C.c_int32]
    L.ela3d_halo_bytes.argtypes = [C.c_void_p, C.c_int32, C.c_int32, P(C.c_size_t)]
    L.ela3d_halo_pack.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_void_p, C.c_size_t]
    L.ela3d_halo_unpack.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_void_p, C.c_size_t]
    L.ela3d_halo_exchange.argtypes = [C.c_void_p, C.c_void_p, C.c_int32]
    L.ela3d_set_device_limit.argtypes = [C.c_int64]
    L.ela3d_destroy.argtypes = [C.c_void_p]
    L.ela3d_destroy.restype = None
    L.ela3d_plan.argtypes = [P(ConfigStruct), P(InfoStruct)]
    L.ela3d_autotune.argtypes = [P(ConfigStruct), P(f), P(f), P(f), C.c_int32, P(TuningStruct)]
    L.ela3d_tuning_set.argtypes = [P(ConfigStruct), P(TuningStruct)]
    L.ela3d_tuning_get.argtypes = [P(ConfigStruct), P(TuningStruct)]
    L.ela3d_pml_tables.argtypes = [P(ConfigStruct), C.c_int32, P(C.c_double), C.c_size_t]
    L.ela3d_strerror.argtypes = [C.c_int]
    L.ela3d_strerror.restype = C.c_char_p
    L.ela3d_last_error.argtypes = []
    L.ela3d_last_error.restype = C.c_char_p
    L.ela3d_device_count.argtypes = [P(C.c_int32)]
    L.ela3d_copy_bench.argtypes = [C.c_int32, C.c_int64, C.c_int32, C.c_int32, P(C.c_double)]
    L.ela3d_model_homogeneous.argtypes = [C.c_int32, C.c_int32, C.c_int32, f, f, f, P(f), P(f), P(f)]
    L.ela3d_layered_default.argtypes = [P(LayeredSpec)]
    L.ela3d_layered_default.restype = None
    L.ela3d_model_layered.argtypes = [P(LayeredSpec), P(f), P(f), P(f), P(C.c_double)]
    L.ela3d_source_trace.argtypes = [C.c_void_p, P(C.c_double), C.c_size_t]
    L.ela3d_checkpoint_save.argtypes = [C.c_void_p, C.c_char_p]
    L.ela3d_checkpoint_load.argtypes = [C.c_void_p, C.c_char_p]
    L.ela3d_model_anticline.argtypes = [C.c_int32, P(f), P(f), P(f)]
    L.ela3d_read_raw.argtypes = [C.c_char_p, C.c_int64, C.c_int32, P(f)]
    L.ela3d_write_raw.argtypes = [C.c_char_p, C.c_int64, C.c_int32, P(f)]
    L.ela3d_write_traces.argtypes = [C.c_char_p, P(C.c_double), C.c_int64, C.c_int64, C.c_int32,
                                     C.c_int32, C.c_int32, C.c_int32]
    if hasattr(L, "ela3d_server_create"):
        L.ela3d_server_opts_default.argtypes = [P(ServerOpts)]
        L.ela3d_server_opts_default.restype = None
        L.ela3d_server_create.argtypes = [P(ServerOpts), P(C.c_void_p)]
        L.ela3d_submit.argtypes = [C.c_void_p, P(ConfigStruct), P(f), P(f), P(f), P(C.c_uint64)]
        L.ela3d_poll.argtypes = [C.c_void_p, C.c_uint64, P(C.c_int32)]
        L.ela3d_wait.argtypes = [C.c_void_p, C.c_uint64, C.c_double, P(ResultStruct)]
        L.ela3d_result_seismograms.argtypes = [C.c_void_p, C.c_uint64, P(C.c_double), C.c_size_t]
        L.ela3d_result_energy.argtypes = [C.c_void_p, C.c_uint64, P(C.c_double), C.c_size_t]
        L.ela3d_result_release.argtypes = [C.c_void_p, C.c_uint64]
        L.ela3d_server_stats.argtypes = [C.c_void_p, P(C.c_int64)]
        L.ela3d_server_destroy.argtypes = [C.c_void_p]
        L.ela3d_server_destroy.restype = None
    _lib = L
    return L


class Ela3dError(RuntimeError):
    def __init__(self, code: int, message: str):
        self.code = code
        super().__init__(f"{E_NAMES.get(code, code)}: {message}")


def check(rc: int) -> int:
    if rc != 0:
        msg = lib().ela3d_last_error().decode(errors="replace")
        raise Ela3dError(rc, msg)
    return rc
