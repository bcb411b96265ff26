"""The C-ABI library loads and exports every symbol that include/reconnet.h
declares (no compute call: this runs without a GPU)."""

import ctypes
import os
import re
import subprocess

from reconnet_amd import LIB_PATH

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "reconnet.h")

# the 17 extern "C" exports of the reference's libReconNet4D_no_tf.so
REFERENCE_EXPORTS = [
    "SetDevice", "cDistanceDrivenFanProjection", "cDistanceDrivenFanBackprojection",
    "cDistanceDrivenTomoProjection", "cDistanceDrivenTomoBackprojection", "cFilterFanFilter",
    "cFilterParallelFilter", "cPixelDrivenFanBackprojection", "cSiddonFanProjection", "cSiddonFanBackprojection",
    "cSiddonParallelProjection", "cSiddonParallelBackprojection", "cSiddonConeProjectionAbitrary",
    "cSiddonConeBackprojectionAbitrary", "cOSSQS", "cOSSQSConeAbitrary", "cTVSQS2D",
]


def declared_functions(header=HEADER):
    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?(?:void|int|float|double|char)\s*\**\s*\*?\s*([A-Za-z_]\w*)\s*\(",
                       text, flags=re.M)
    return sorted(set(names))


def exported_symbols(path=LIB_PATH):
    out = subprocess.check_output(["nm", "-D", "--defined-only", path]).decode()
    return {line.split()[-1] for line in out.splitlines() if " T " in line}


def test_library_is_built_in_tree():
    assert os.path.exists(LIB_PATH), "run `make -C reconnet_amd`"
    assert os.path.basename(LIB_PATH) == "libReconNet4D_no_tf.so"


def test_header_declares_reference_exports():
    decl = declared_functions()
    for name in REFERENCE_EXPORTS:
        assert name in decl


def test_every_declared_symbol_is_exported():
    syms = exported_symbols()
    missing = [n for n in declared_functions() if n not in syms]
    assert not missing, missing


def test_library_loads_with_ctypes_by_file_name():
    # the reference caller does cdll.LoadLibrary(<dir>/libReconNet4D_no_tf.so)
    lib = ctypes.cdll.LoadLibrary(LIB_PATH)
    for name in REFERENCE_EXPORTS:
        assert hasattr(lib, name)
    lib.rn_last_error.restype = ctypes.c_char_p
    # a call that needs no device succeeds and clears the thread's last error
    assert lib.rn_autotune_reset() == 0
    assert lib.rn_last_error() == b""
    # pinning a launch candidate is host state only
    assert lib.rn_autotune_pin(1, 3) == 0 and lib.rn_autotune_pin(1, -1) == 0


PFS_HEADER = os.path.join(ROOT, "include", "priorfunctionsolver.h")
PFS_PATH = os.path.join(os.path.dirname(LIB_PATH), "libPriorFunctionSolver.so")


def test_prior_function_solver_library_exports():
    # the TV-SQS entry points of the reference's libPriorFunctionSolver.so
    # (PriorFunctionSolver/TVSQS.cu:108,290) and its device selector
    assert os.path.exists(PFS_PATH), "run `make -C reconnet_amd`"
    decl = declared_functions(PFS_HEADER)
    for name in ("cSetDevice", "cTVSQS2D", "cTVSQS3D"):
        assert name in decl
    syms = exported_symbols(PFS_PATH)
    assert not [n for n in decl if n not in syms]
    lib = ctypes.cdll.LoadLibrary(PFS_PATH)
    lib.pfs_last_error.restype = ctypes.c_char_p
    assert lib.pfs_last_error() == b""
