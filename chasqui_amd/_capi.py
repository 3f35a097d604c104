"""ctypes binding of the C ABI in include/chasqui_amd.h (libchasqui_amd.so, built in-tree).

Loading fails loudly when the native library is missing: there is no Python or CPU
fallback for any kernel.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_lib", "libchasqui_amd.so")

ARRAY_IDS = {
    "u0": 0, "u1": 1, "w": 2, "bcc": 3, "ecc": 4,
    "b1f0": 5, "b2f0": 6, "b3f0": 7, "b1f1": 8, "b2f1": 9, "b3f1": 10,
    "f1": 11, "f2": 12, "f3": 13,
    "e2x1": 14, "e3x1": 15, "e1x2": 16, "e3x2": 17, "e1x3": 18, "e2x3": 19,
    "e1": 20, "e2": 21, "e3": 22, "flag": 23,
}
CELL_ARRAYS = {"u0": 6, "u1": 6, "w": 6, "bcc": 3, "ecc": 3, "flag": 1}
FACE_MULTI = {"f1": 6, "f2": 6, "f3": 6}

RECON = {"dc": 0, "plm": 1, "ppm": 2, "wenoz": 3, "ppmx": 4}
RSOLVER = {"hlle": 0, "hlld": 1, "lhlld": 2}
INTEGRATOR = {"rk1": 1, "rk2": 2, "rk3": 3, "rk4": 4, "ssprk43": 4}
BC = {"periodic": 0, "outflow": 1, "reflect": 2}
PRECISION = {"fp64": 0, "mixed": 1}
KERNEL_CLASSES = {"flux": 0, "corner": 1, "update": 2, "bc": 3, "c2p": 4, "repair": 5, "all": 64}
CLASS_NAMES = ["flux", "corner", "update", "bc", "c2p", "repair"]


class MeshDesc(C.Structure):
    _fields_ = [("nx1", C.c_int), ("nx2", C.c_int), ("nx3", C.c_int), ("ng", C.c_int),
                ("x1min", C.c_double), ("x1max", C.c_double), ("x2min", C.c_double),
                ("x2max", C.c_double), ("x3min", C.c_double), ("x3max", C.c_double)]


class Config(C.Structure):
    _fields_ = [("recon", C.c_int), ("rsolver", C.c_int), ("nstage", C.c_int),
                ("gamma", C.c_double), ("dfloor", C.c_double), ("pfloor", C.c_double),
                ("cfl", C.c_double), ("flatten", C.c_int), ("low_mach", C.c_int),
                ("fofc", C.c_int), ("dual_energy", C.c_int), ("de_eta", C.c_double),
                ("lhlld_phi", C.c_double), ("bc", C.c_int * 3), ("precision", C.c_int),
                ("characteristic", C.c_int), ("nu", C.c_double), ("kappa", C.c_double),
                ("eta", C.c_double), ("sts", C.c_int), ("flux_limiter", C.c_int),
                ("grav_uniform", C.c_int), ("grav_point", C.c_int), ("g1", C.c_double), ("g2", C.c_double),
                ("g3", C.c_double), ("gm", C.c_double), ("gx0", C.c_double), ("gy0", C.c_double),
                ("gz0", C.c_double), ("rsoft", C.c_double), ("cooling", C.c_int), ("lambda0", C.c_double),
                ("cool_alpha", C.c_double), ("tfloor", C.c_double)]


class Status(C.Structure):
    _fields_ = [("t", C.c_double), ("dt", C.c_double), ("tlim", C.c_double), ("min_dt", C.c_double),
                ("cycle", C.c_int), ("nlim", C.c_int), ("active", C.c_int), ("nbad_last", C.c_int),
                ("fofc_total", C.c_longlong), ("sts_stages", C.c_longlong), ("limiter_total", C.c_longlong)]


_lib = None


class NativeLibraryMissing(RuntimeError):
    pass


def lib():
    """Load libchasqui_amd.so (raises NativeLibraryMissing if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise NativeLibraryMissing(
            f"{LIB_PATH} is missing: build it with `python -m chasqui_amd.build` "
            "(there is no fallback path)")
    L = C.CDLL(LIB_PATH)
    vp, dp, ip = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int)
    sig = {
        "chq_last_error": (C.c_char_p, []),
        "chq_version": (C.c_char_p, []),
        "chq_create": (C.c_int, [C.POINTER(MeshDesc), C.POINTER(Config), C.POINTER(vp)]),
        "chq_destroy": (C.c_int, [vp]),
        "chq_geometry": (C.c_int, [vp, ip, dp]),
        "chq_device_bytes": (C.c_size_t, [vp]),
        "chq_array_count": (C.c_size_t, [vp, C.c_int]),
        "chq_upload": (C.c_int, [vp, C.c_int, vp, C.c_size_t]),
        "chq_download": (C.c_int, [vp, C.c_int, vp, C.c_size_t]),
        "chq_fluxes": (C.c_int, [vp, C.c_int, C.c_int]),
        "chq_corner_emf": (C.c_int, [vp]),
        "chq_update": (C.c_int, [vp, C.c_int, C.c_double, C.c_double, C.c_double]),
        "chq_boundaries": (C.c_int, [vp, C.c_int]),
        "chq_cons_to_prim": (C.c_int, [vp, C.c_int, C.c_int, dp]),
        "chq_fofc_flag": (C.c_int, [vp, C.c_int, ip]),
        "chq_fofc_repair": (C.c_int, [vp, C.c_int, C.c_int]),
        "chq_diffusion_fluxes": (C.c_int, [vp, C.c_int, C.c_int]),
        "chq_resistive_emf": (C.c_int, [vp, C.c_int]),
        "chq_diffusion_dt": (C.c_int, [vp, C.POINTER(C.c_double)]),
        "chq_limit_fluxes": (C.c_int, [vp, C.c_int, C.c_double, C.POINTER(C.c_long)]),
        "chq_apply_gravity": (C.c_int, [vp, C.c_int, C.c_double]),
        "chq_apply_cooling": (C.c_int, [vp, C.c_int, C.c_double]),
        "chq_stage": (C.c_int, [vp, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                C.POINTER(C.c_long)]),
        "chq_init": (C.c_int, [vp, C.c_double, C.c_double, C.c_int]),
        "chq_init_resume": (C.c_int, [vp, C.c_double, C.c_int, C.c_double, C.c_int]),
        "chq_restart_write": (C.c_int, [vp, C.c_char_p, C.c_double, C.c_int]),
        "chq_restart_read": (C.c_int, [vp, C.c_char_p, dp, ip]),
        "chq_step": (C.c_int, [vp, C.c_int]),
        "chq_step_timed": (C.c_int, [vp, C.c_int, dp]),
        "chq_get_status": (C.c_int, [vp, C.POINTER(Status)]),
        "chq_sync": (C.c_int, [vp]),
        "chq_stream": (vp, [vp]),
        "chq_kernel_timing": (C.c_int, [vp, C.c_int]),
        "chq_kernel_timing_read": (C.c_int, [vp, dp, ip]),
        "chq_kernel_timing_breakdown": (C.c_int, [vp, dp, ip]),
        "chq_deck_open": (C.c_int, [C.c_char_p, C.c_char_p, C.POINTER(vp)]),
        "chq_deck_open_lazy": (C.c_int, [C.c_char_p, C.c_char_p, C.POINTER(vp)]),
        "chq_deck_prepare": (C.c_int, [vp]),
        "chq_deck_close": (C.c_int, [vp]),
        "chq_deck_last_error": (C.c_char_p, []),
        "chq_deck_mesh": (C.c_int, [vp, C.POINTER(MeshDesc)]),
        "chq_deck_config": (C.c_int, [vp, C.POINTER(Config)]),
        "chq_deck_run": (C.c_int, [vp, dp, ip, ip]),
        "chq_deck_get": (C.c_char_p, [vp, C.c_char_p, C.c_char_p]),
        "chq_deck_initial": (C.c_int, [vp, C.c_int, vp, C.c_size_t]),
        "chq_deck_print": (C.c_size_t, [vp, C.c_char_p, C.c_size_t]),
        "chq_deck_upload": (C.c_int, [vp, vp]),
        "chq_write_vtk": (C.c_int, [vp, C.POINTER(MeshDesc), C.c_char_p, C.c_double, C.c_int]),
        "chq_history": (C.c_int, [vp, C.POINTER(MeshDesc), C.c_int, dp]),
        "chq_history_noalloc": (C.c_int, [vp, C.POINTER(MeshDesc), C.c_int, dp]),
        "chq_batch_create": (C.c_int, [C.c_int, C.POINTER(vp)]),
        "chq_batch_destroy": (C.c_int, [vp]),
        "chq_batch_add": (C.c_int, [vp, vp]),
        "chq_batch_remove": (C.c_int, [vp, vp]),
        "chq_batch_size": (C.c_int, [vp]),
        "chq_batch_run": (vp, [vp, C.c_int]),
        "chq_batch_step": (C.c_int, [vp, C.c_int]),
        "chq_batch_step_timed": (C.c_int, [vp, C.c_int, dp]),
        "chq_batch_sync": (C.c_int, [vp]),
        "chq_batch_unfinished": (C.c_int, [vp, ip, ip]),
        "chq_batch_use_graph": (C.c_int, [vp, C.c_int]),
        "chq_batch_tune": (C.c_int, [vp, C.c_int, dp]),
        "chq_batch_grid": (C.c_int, [vp, C.c_int, ip, C.POINTER(C.c_char_p)]),
        "chq_batch_set_grid": (C.c_int, [vp, C.c_int, C.c_int]),
        "chq_bytes_for": (C.c_size_t, [C.POINTER(MeshDesc), C.POINTER(Config)]),
        "chq_server_create": (C.c_int, [C.c_size_t, C.c_int, C.c_int, C.POINTER(vp)]),
        "chq_server_destroy": (C.c_int, [vp]),
        "chq_server_submit": (C.c_int, [vp, C.c_char_p, C.c_char_p, C.c_int, ip]),
        "chq_server_pump": (C.c_int, [vp, C.c_int, ip, ip]),
        "chq_server_job": (C.c_int, [vp, C.c_int, dp]),
        "chq_server_job_ctx": (vp, [vp, C.c_int]),
        "chq_server_job_error": (C.c_char_p, [vp, C.c_int]),
        "chq_server_release": (C.c_int, [vp, C.c_int]),
        "chq_server_stats": (C.c_int, [vp, dp]),
        "chq_server_clock": (C.c_double, []),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def exported_symbols():
    return [
        "chq_last_error", "chq_version", "chq_create", "chq_destroy", "chq_geometry",
        "chq_device_bytes", "chq_array_count", "chq_upload", "chq_download", "chq_fluxes",
        "chq_corner_emf", "chq_update", "chq_boundaries", "chq_cons_to_prim", "chq_fofc_flag",
        "chq_fofc_repair", "chq_diffusion_fluxes", "chq_resistive_emf", "chq_diffusion_dt", "chq_limit_fluxes",
        "chq_apply_gravity", "chq_apply_cooling", "chq_stage", "chq_init", "chq_step", "chq_step_timed",
        "chq_get_status", "chq_sync", "chq_stream", "chq_kernel_timing", "chq_kernel_timing_read",
        "chq_kernel_timing_breakdown",
        "chq_deck_open", "chq_deck_open_lazy", "chq_deck_prepare", "chq_deck_close", "chq_deck_last_error", "chq_deck_mesh", "chq_deck_config",
        "chq_deck_run", "chq_deck_get", "chq_deck_initial", "chq_deck_print", "chq_deck_upload",
        "chq_write_vtk", "chq_history", "chq_history_noalloc", "chq_init_resume", "chq_restart_write", "chq_restart_read",
        "chq_batch_create", "chq_batch_destroy", "chq_batch_add", "chq_batch_remove", "chq_batch_size",
        "chq_batch_run", "chq_batch_step", "chq_batch_step_timed", "chq_batch_sync", "chq_batch_unfinished",
        "chq_batch_use_graph", "chq_batch_tune", "chq_batch_grid", "chq_batch_set_grid",
        "chq_bytes_for", "chq_server_create", "chq_server_destroy", "chq_server_submit", "chq_server_pump",
        "chq_server_job", "chq_server_job_ctx", "chq_server_job_error", "chq_server_release",
        "chq_server_stats", "chq_server_clock",
    ]


def check(rc, what=""):
    if rc != 0:
        msg = lib().chq_last_error().decode()
        raise RuntimeError(f"{what} failed ({rc}): {msg}")
