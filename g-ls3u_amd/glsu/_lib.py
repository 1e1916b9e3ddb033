"""ctypes binding of libglsu.so (include/glsu.h).

The library is the product: HIP kernels for gfx950 behind a C ABI.  There is
no fallback -- if the shared object is missing or no GPU is visible, calls
raise.
"""
import ctypes
import os
import subprocess

_PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))  # g-ls3u_amd/
LIB_PATH = os.path.join(_PKG_ROOT, "lib", "libglsu.so")
CSRC = os.path.join(_PKG_ROOT, "csrc")

OK = 0
ERR_ARG = -1
ERR_PADDING = -2
ERR_SIZE = -3
ERR_DEVICE = -4
ERR_UNSUPPORTED = -5
ERR_NOT_PD = -6

WFF = 0
WFR = 1
U8 = 0
F32 = 1
F64 = 2


class GlsuError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"glsu error {code}: {msg}")
        self.code = code


class PaddingError(GlsuError):
    """Padded FFT size out of range; the reference constructors throw int -1."""


class NotPositiveDefiniteError(GlsuError):
    """An AIA normal matrix is not positive definite; the reference prints
    that and exits (AIA/aia_cpuf.cpp:163-167)."""


class WfrResults(ctypes.Structure):
    """glsu_wfr_results: device pointers of the eight WFR maps."""
    _fields_ = [(n, ctypes.c_void_p) for n in
                ("wx", "wy", "phase", "phase_comp", "b", "r", "cxx", "cyy")]


TUNING_FIELDS = ("planx", "plany", "cols_waves", "rows_waves", "ls_tile")
LS_TILES = {1: (64, 8), 2: (64, 4), 3: (32, 8), 4: (128, 4), 5: (64, 16), 6: (32, 16)}


class Tuning(ctypes.Structure):
    """glsu_tuning: plan ids and launch shapes (-1 = library default);
    ls_tile: least-squares tile shape id (LS_TILES)."""
    _fields_ = [(f, ctypes.c_int) for f in TUNING_FIELDS]

    @staticmethod
    def make(planx=-1, plany=-1, cols_waves=-1, rows_waves=-1, ls_tile=-1):
        return Tuning(int(planx), int(plany), int(cols_waves), int(rows_waves), int(ls_tile))


class SeqDesc(ctypes.Structure):
    """glsu_seq_desc: one sequence of a batch."""
    _fields_ = [("width", ctypes.c_int), ("height", ctypes.c_int), ("rr", ctypes.c_int),
                ("phi0", ctypes.POINTER(ctypes.c_double))]


def tuning_arg(tuning=None, planx=-1, plany=-1):
    """Tuning struct from a Tuning, a dict with any of its fields, or None
    (planx / plany given separately override the dict)."""
    if isinstance(tuning, Tuning):
        t = Tuning.make(*(getattr(tuning, f) for f in TUNING_FIELDS))
    else:
        t = Tuning.make(**{k: v for k, v in (tuning or {}).items() if k in TUNING_FIELDS})
    if planx >= 0:
        t.planx = int(planx)
    if plany >= 0:
        t.plany = int(plany)
    return t


def tuning_dict(t):
    return {f: getattr(t, f) for f in TUNING_FIELDS}


class WffParams(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in
                ("sigma_x", "wxl", "wxh", "wxi", "sigma_y", "wyl", "wyh", "wyi", "thr")]


_lib = None


def build(jobs=8):
    """Compile libglsu.so for gfx950 (hipcc; runs without a GPU)."""
    subprocess.run(["make", "-s", "-C", CSRC, f"-j{jobs}"], check=True)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"libglsu.so not built ({LIB_PATH}); run __graft_entry__.build()")
    L = ctypes.CDLL(LIB_PATH)
    I, D, V = ctypes.c_int, ctypes.c_double, ctypes.c_void_p
    PI, PD = ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double)
    PV = ctypes.POINTER(ctypes.c_void_p)
    PP = ctypes.POINTER(WffParams)
    PT = ctypes.POINTER(Tuning)
    PF = ctypes.c_void_p
    U64 = ctypes.c_uint64
    sig = {
        "glsu_last_error": (ctypes.c_char_p, []),
        "glsu_version": (ctypes.c_char_p, []),
        "glsu_device_count": (I, [PI]),
        "glsu_wff_params_default": (I, [PP]),
        "glsu_wfr_params_default": (I, [PP]),
        "glsu_wff_params_dpra_reference": (I, [PP]),
        "glsu_wft2_create": (I, [PV, I, I, I, PP, V]),
        "glsu_wft2_create_tuned": (I, [PV, I, I, I, PP, V, PT]),
        "glsu_wft2_get_tuning": (I, [V, PT]),
        "glsu_wft2_compute": (I, [V, V, V, PD]),
        "glsu_wft2_compute_host": (I, [V, PF, PF, PD]),
        "glsu_wft2_info": (I, [V, PI, PI, PI, PI, PI, PI, PD]),
        "glsu_wft2_destroy": (I, [V]),
        "glsu_wft2_compute_wfr": (I, [V, V, ctypes.POINTER(WfrResults), PD]),
        "glsu_wft2_compute_wfr_host": (I, [V, V, V, PD]),
        "glsu_wft2_time_kernel": (I, [V, V, I, I, PD]),
        "glsu_wft2_compute_noalloc": (I, [I, I, PP, V, V, V]),
        "glsu_dpra_create": (I, [PV, PD, I, I, I, PP, V]),
        "glsu_dpra_create_f32": (I, [PV, PF, I, I, I, PP, V]),
        "glsu_dpra_create_tuned": (I, [PV, PD, I, I, I, PP, V, PT]),
        "glsu_dpra_get_tuning": (I, [V, PT]),
        "glsu_dpra_time_kernel": (I, [V, I, I, PD]),
        "glsu_dpra_frame": (I, [V, V, I, V, V, PD]),
        "glsu_dpra_frame_host": (I, [V, PF, I, PF, PD]),
        "glsu_dpra_update_ref": (I, [V]),
        "glsu_dpra_run": (I, [V, V, I, I, V, PD]),
        "glsu_dpra_run_graph": (I, [V, V, I, I, V, PD]),
        "glsu_dpra_profile": (I, [V, V, I, I, V, PD]),
        "glsu_dpra_profile_skips": (I, [V, V, I, I, V, PD, ctypes.POINTER(ctypes.c_ulonglong),
                                        ctypes.POINTER(ctypes.c_ulonglong)]),
        "glsu_dpra_get_state": (I, [V, PF, PF]),
        "glsu_dpra_set_state": (I, [V, PF, PF]),
        "glsu_dpra_copy_z": (I, [V, V]),
        "glsu_dpra_thr": (I, [V, PD]),
        "glsu_dpra_destroy": (I, [V]),
        "glsu_gen_frames": (I, [V, I, I, I, I, U64, D, D, D, D, V]),
        "glsu_dpra_batch_create": (I, [PV, ctypes.POINTER(SeqDesc), I, PP, V, PT]),
        "glsu_dpra_batch_run": (I, [V, ctypes.POINTER(ctypes.c_void_p), I, I,
                                    ctypes.POINTER(ctypes.c_void_p), I, PD]),
        "glsu_dpra_batch_info": (I, [V, PI, PI]),
        "glsu_dpra_batch_get_tuning": (I, [V, I, PT]),
        "glsu_dpra_batch_get_state": (I, [V, I, V, V]),
        "glsu_dpra_batch_destroy": (I, [V]),
        "glsu_unet_create": (I, [PV, I, I, ctypes.c_ulonglong, V]),
        "glsu_unet_param_count": (I, [V, ctypes.POINTER(ctypes.c_long)]),
        "glsu_unet_get_params": (I, [V, V]),
        "glsu_unet_set_params": (I, [V, V]),
        "glsu_unet_get_grads": (I, [V, V]),
        "glsu_unet_forward": (I, [V, V, V, I, I, I]),
        "glsu_unet_loss_grad": (I, [V, V, V, I, I, I, PD]),
        "glsu_unet_adam_step": (I, [V, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                    ctypes.c_float]),
        "glsu_unet_destroy": (I, [V]),
        "glsu_dpra_frame_learned": (I, [V, V, V, I, V, D, I, I, PD, PI]),
        "glsu_dpra_copy_net": (I, [V, V]),
        "glsu_dpra_learned_stats": (I, [V, PI, PI, PD, PD, I]),
        "glsu_read_matrix_bin": (I, [ctypes.c_char_p, I, PI, PI, V, ctypes.c_size_t]),
        "glsu_write_matrix_bin": (I, [ctypes.c_char_p, I, I, I, V]),
        "glsu_read_complex_text": (I, [ctypes.c_char_p, PI, PI, V, ctypes.c_size_t]),
        "glsu_write_complex_text": (I, [ctypes.c_char_p, I, I, V, I]),
        "glsu_fft_plans": (I, [PI, PI, PI, PI, I, PI]),
        "glsu_fft_plan_exchange": (I, [I]),
        "glsu_phase_to_u8": (I, [V, V, ctypes.c_size_t, V]),
        "glsu_aia_create": (I, [PV, I, I, I, V]),
        "glsu_aia_run": (I, [V, V, I, V, I, U64, I, D, I, V, PI, PD, PD]),
        "glsu_aia_run_host": (I, [V, PV, I, V, I, U64, I, D, I, V, PI, PD, PD]),
        "glsu_aia_info": (I, [V, PI, PI, PI]),
        "glsu_aia_time_iter": (I, [V, V, I, V, I, I, PD]),
        "glsu_aia_destroy": (I, [V]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def check_out(t, shape, dtype, name="out"):
    """Validate a caller-supplied output tensor before its address goes to the
    C ABI: a contiguous CUDA tensor of exactly `shape` and `dtype` (the
    library writes shape-many elements through the raw pointer).  Raises
    ValueError otherwise; returns t."""
    if not getattr(t, "is_cuda", False):
        raise ValueError(f"{name} must be a CUDA tensor")
    if t.dtype != dtype:
        raise ValueError(f"{name} must be {dtype}, got {t.dtype}")
    if tuple(t.shape) != tuple(shape):
        raise ValueError(f"{name} has shape {tuple(t.shape)}, expected {tuple(shape)}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    return t


def check(rc):
    if rc != OK:
        msg = lib().glsu_last_error().decode(errors="replace")
        if rc == ERR_PADDING:
            raise PaddingError(rc, msg)
        if rc == ERR_NOT_PD:
            raise NotPositiveDefiniteError(rc, msg)
        raise GlsuError(rc, msg)
    return rc


def exported_symbols():
    """Names declared in include/glsu.h (for the load/export test)."""
    return [
        "glsu_last_error", "glsu_version", "glsu_device_count", "glsu_wff_params_default", "glsu_wfr_params_default",
        "glsu_wff_params_dpra_reference", "glsu_wft2_create", "glsu_wft2_create_tuned",
        "glsu_wft2_compute", "glsu_wft2_compute_host", "glsu_wft2_info", "glsu_wft2_get_tuning",
        "glsu_wft2_destroy", "glsu_wft2_compute_wfr", "glsu_wft2_compute_wfr_host",
        "glsu_wft2_time_kernel", "glsu_wft2_compute_noalloc",
        "glsu_dpra_create", "glsu_dpra_create_f32", "glsu_dpra_create_tuned", "glsu_dpra_get_tuning",
        "glsu_dpra_frame",
        "glsu_dpra_frame_host", "glsu_dpra_update_ref", "glsu_dpra_run", "glsu_dpra_run_graph",
        "glsu_dpra_profile", "glsu_dpra_profile_skips", "glsu_dpra_time_kernel", "glsu_dpra_get_state", "glsu_dpra_set_state", "glsu_dpra_copy_z", "glsu_dpra_thr",
        "glsu_dpra_destroy", "glsu_gen_frames", "glsu_dpra_batch_create", "glsu_dpra_batch_run",
        "glsu_dpra_batch_info", "glsu_dpra_batch_get_tuning", "glsu_dpra_batch_get_state",
        "glsu_dpra_batch_destroy", "glsu_unet_create", "glsu_unet_param_count",
        "glsu_unet_get_params", "glsu_unet_set_params", "glsu_unet_get_grads", "glsu_unet_forward",
        "glsu_unet_loss_grad", "glsu_unet_adam_step", "glsu_unet_destroy",
        "glsu_dpra_frame_learned", "glsu_dpra_copy_net", "glsu_dpra_learned_stats",
        "glsu_read_matrix_bin", "glsu_write_matrix_bin", "glsu_read_complex_text",
        "glsu_write_complex_text", "glsu_fft_plans", "glsu_fft_plan_exchange", "glsu_phase_to_u8",
        "glsu_aia_create", "glsu_aia_run", "glsu_aia_run_host", "glsu_aia_info", "glsu_aia_time_iter", "glsu_aia_destroy",
    ]


def fft_plans():
    n = ctypes.c_int()
    ids = (ctypes.c_int * 64)()
    t = (ctypes.c_int * 64)()
    e = (ctypes.c_int * 64)()
    ln = (ctypes.c_int * 64)()
    check(lib().glsu_fft_plans(ids, t, e, ln, 64, ctypes.byref(n)))
    # exchange: 0 every cross-lane stage through LDS, 1 the first through registers
    return [dict(id=ids[i], T=t[i], E=e[i], N=ln[i], exchange=lib().glsu_fft_plan_exchange(ids[i]))
            for i in range(n.value)]
