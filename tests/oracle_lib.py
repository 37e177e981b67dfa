"""Loader of the CPU oracle (oracle/libsdm_oracle.so) for the tests and the CPU baseline.

The oracle exports the C ABI of include/sdm.h with the prefix ``sdmcpu_``, plus a few
oracle-only entry points bound here.
"""
import ctypes
import os
import subprocess

from sudodem_amd._native import SdmLibrary

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_LIB = os.path.join(ORACLE_DIR, "libsdm_oracle.so")

_LIB = None


def load_oracle() -> SdmLibrary:
    global _LIB
    if _LIB is None:
        if not os.path.exists(ORACLE_LIB):
            subprocess.run(["make", "-C", ORACLE_DIR], check=True, capture_output=True)
        lib = SdmLibrary(ORACLE_LIB, "sdmcpu_")
        L = lib.lib
        lib.set_threads = L.sdmcpu_set_threads
        lib.set_threads.restype = ctypes.c_int
        lib.set_threads.argtypes = [ctypes.c_void_p, ctypes.c_int32]
        lib.set_init_sort = L.sdmcpu_set_init_sort
        lib.set_init_sort.restype = ctypes.c_int
        lib.set_init_sort.argtypes = [ctypes.c_void_p, ctypes.c_int32]
        d3 = ctypes.POINTER(ctypes.c_double)
        lib.support_point = L.sdmcpu_support_point
        lib.support_point.restype = ctypes.c_int
        lib.support_point.argtypes = [ctypes.c_double] * 5 + [d3, d3, d3]
        lib.curvatures = L.sdmcpu_curvatures
        lib.curvatures.restype = ctypes.c_int
        lib.curvatures.argtypes = [ctypes.c_double] * 5 + [d3, d3, d3, d3, d3, ctypes.POINTER(ctypes.c_int32)]
        lib.hertz_mindlin = L.sdmcpu_hertz_mindlin
        lib.hertz_mindlin.restype = ctypes.c_int
        lib.hertz_mindlin.argtypes = [d3, d3, ctypes.c_double, d3]
        _LIB = lib
    return _LIB
