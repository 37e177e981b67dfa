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
# the same restatement built on the correctly rounded functions of crmath.h (glibc's values
# for all but ~0.1 % of arguments): the reference arithmetic with glibc's misroundings removed
ORACLE_CR_LIB = os.path.join(ORACLE_DIR, "libsdm_oracle_cr.so")

_LIBS = {}


def load_oracle(cr: bool = False) -> SdmLibrary:
    path = ORACLE_CR_LIB if cr else ORACLE_LIB
    if path not in _LIBS:
        if not os.path.exists(path):
            subprocess.run(["make", "-C", ORACLE_DIR], check=True, capture_output=True)
        lib = SdmLibrary(path, "sdmcpu_")
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
        _LIBS[path] = lib
    return _LIBS[path]
