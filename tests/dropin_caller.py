"""ctypes binding of tests/dropin/libgrid_caller.so, the grid.h-only caller
(tests/dropin/grid_caller.cpp): the drop-in C++ API driven in-process as the
reference's GUI drives it."""
import ctypes as C
import os

import numpy as np

PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "dropin", "libgrid_caller.so")
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(PATH):
            raise RuntimeError(f"{PATH} missing: run __graft_entry__.build()")
        _lib = C.CDLL(PATH)
        fp = C.POINTER(C.c_float)
        _lib.dropin_recon.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, fp, C.c_float, C.c_char_p, C.c_float,
                                      C.c_float, C.c_float, C.c_float, C.c_float, C.c_float, C.c_long, C.c_int, fp, fp,
                                      C.POINTER(C.c_long), fp]
        _lib.dropin_last_error.restype = C.c_char_p
    return _lib


def recon(form, slices, n_det, center, geom=1, angles=None, filter="shlo", C_=6.0, sampl=1.0, R=1.0,
          MaxPixSiz=1.0, X0=0.0, Y0=0.0, ltbl=512, M0=None):
    """Run the caller on slices [n, n_ang, n_det]: (images [n, M0, M0], imgsiz,
    GP->sampl after recon_init, status, last error)."""
    slices = np.ascontiguousarray(slices, np.float32)
    n, n_ang = slices.shape[:2]
    size = M0 if M0 else 1
    out = np.full((n, size, size), np.nan, np.float32)
    ang = None if angles is None else np.ascontiguousarray(angles, np.float32)
    imgsiz, smp = C.c_long(0), C.c_float(0)
    fp = C.POINTER(C.c_float)
    rc = lib().dropin_recon(form, n_ang, n_det, geom, None if ang is None else ang.ctypes.data_as(fp), center,
                            filter.encode(), C_, sampl, R, MaxPixSiz, X0, Y0, ltbl, n,
                            slices.ctypes.data_as(fp), out.ctypes.data_as(fp), C.byref(imgsiz), C.byref(smp))
    return out, imgsiz.value, smp.value, rc, lib().dropin_last_error().decode()
