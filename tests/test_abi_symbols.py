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


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?(?:void|int|float|double|char)\s*\**\s*\*?\s*([A-Za-z_]\w*)\s*\(",
                       text, flags=re.M)
    return sorted(set(names))


def exported_symbols():
    out = subprocess.check_output(["nm", "-D", "--defined-only", LIB_PATH]).decode()
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
    assert lib.rn_last_error() == b""
