"""Loader of the in-tree HIP library ``libReconNet4D_no_tf.so``.

The library is built in-tree (``make -C reconnet_amd``); there is no CPU
fallback: every entry point of this package fails loudly when the library is
missing or its call fails.
"""

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_NAME = "libReconNet4D_no_tf.so"
LIB_PATH = os.path.join(_HERE, LIB_NAME)

_lib = None

c_int, c_float, c_void_p, c_double = ctypes.c_int, ctypes.c_float, ctypes.c_void_p, ctypes.c_double
_P = c_void_p


def _sig(lib, name, restype, argtypes):
    fn = getattr(lib, name)
    fn.restype = restype
    fn.argtypes = argtypes


def lib():
    """Load (once) and return the ctypes handle of libReconNet4D_no_tf.so."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("%s is not built: run `make -C reconnet_amd` or __graft_entry__.build()" % LIB_NAME)
    try:
        # torch ships its own HIP runtime with the same soname; loading torch
        # first makes the library bind to that one runtime.
        import torch  # noqa: F401
    except ImportError:
        pass
    h = ctypes.CDLL(LIB_PATH)
    fan = [c_int, c_int, c_int, c_float, c_float, c_float, c_int, c_int, c_int,
           c_float, c_float, c_float, c_float, c_float, c_float]
    _sig(h, "rn_last_error", ctypes.c_char_p, [])
    _sig(h, "rn_fan_plan_create", c_void_p, fan + [_P])
    _sig(h, "rn_fan_plan_destroy", None, [_P])
    _sig(h, "rn_fan_forward", c_int, [_P, c_int, _P, _P, c_int, c_int, _P])
    _sig(h, "rn_fan_backward", c_int, [_P, c_int, _P, _P, c_int, c_int, _P])
    _sig(h, "rn_fan_filter", c_int, [_P, _P, _P, c_int, c_int, c_int, c_int, _P])
    _sig(h, "rn_sqs_create", c_void_p, [_P, c_int, c_int, c_int])
    _sig(h, "rn_sqs_destroy", None, [_P])
    _sig(h, "rn_sqs_set_data", c_int, [_P, _P, _P, _P])
    _sig(h, "rn_sqs_compute_norm", c_int, [_P, _P])
    _sig(h, "rn_sqs_set_norm", c_int, [_P, _P, _P])
    _sig(h, "rn_sqs_get_norm", c_int, [_P, _P, _P])
    _sig(h, "rn_sqs_set_image", c_int, [_P, _P, _P])
    _sig(h, "rn_sqs_get_image", c_int, [_P, _P, _P])
    _sig(h, "rn_sqs_pass", c_int, [_P, c_float, c_float, c_float, c_float, c_float, _P, _P])
    _sig(h, "rn_sqs_set_profiling", c_int, [_P, c_int])
    _sig(h, "rn_sqs_kernel_times", c_int, [_P, _P, _P, _P, _P])
    _sig(h, "rn_tv_sqs", c_int, [_P, _P, _P, c_int, c_int, c_int, c_float, _P, _P])
    _sig(h, "rn_autotune_enable", c_int, [c_int])
    _sig(h, "rn_autotune_report", c_int, [ctypes.c_char_p, c_int])
    _sig(h, "rn_autotune_reset", c_int, [])
    _sig(h, "rn_autotune_pin", c_int, [c_int, c_int])
    h.cTVSQS2D.restype = c_float
    _lib = h
    return _lib


def last_error():
    msg = lib().rn_last_error()
    return msg.decode() if msg else ""


def check(status, what):
    """Raise RuntimeError if a status-returning entry point failed."""
    if status != 0:
        raise RuntimeError("%s failed: %s" % (what, last_error()))


# kernel ids of the tuned launches (TuneKernel in csrc/rn_tune.h)
TUNED_KERNELS = {"siddon_fp": 1, "siddon_bp": 2, "dd_fp": 3, "dd_adjoint": 4, "dd_bp": 5, "sqs_update": 6,
                 "dd_fbp": 7, "siddon_fp_q": 8, "siddon_bp_q": 9, "dd_bp_q": 10, "dd_fbp_q": 11, "dd_fp_q": 12,
                 "conv_mid": 13, "cnn_forward": 14, "helical_bp": 15, "dd_cone_fp": 16, "pixel_bp": 17,
                 "siddon_cone_bp": 18, "dd_cone_adj": 19}


def autotune_pin(kernel, candidate):
    """Force launch candidate `candidate` of a tuned kernel (a TUNED_KERNELS
    name or id) for every shape; None or a negative value removes the pin."""
    kid = TUNED_KERNELS[kernel] if isinstance(kernel, str) else int(kernel)
    check(lib().rn_autotune_pin(kid, -1 if candidate is None else int(candidate)), "rn_autotune_pin")


def autotune(enable=None, reset=False):
    """Switch the launch-shape autotuner on/off, clear its cache, and return
    its current choices as {(kernel, a, b, c, d, e): (choice, [ms per candidate])}."""
    h = lib()
    if enable is not None:
        check(h.rn_autotune_enable(1 if enable else 0), "rn_autotune_enable")
    if reset:
        check(h.rn_autotune_reset(), "rn_autotune_reset")
    n = h.rn_autotune_report(None, 0)
    if n < 0:
        check(n, "rn_autotune_report")
    buf = ctypes.create_string_buffer(n + 1)
    h.rn_autotune_report(buf, n + 1)
    out = {}
    for line in buf.value.decode().splitlines():
        key, rest = line.split(" -> ")
        choice, ms = rest.split(" (")
        out[tuple(int(v) for v in key.split())] = (int(choice), [float(v) for v in ms.rstrip(" ms)").split()])
    return out

