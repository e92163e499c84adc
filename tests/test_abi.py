"""The C-ABI boundary without a GPU: the library loads and exports every function that
include/nnet_mi355x.h declares, and the constants the Python binding uses match the header."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "nnet_mi355x.h")
LIB = os.path.join(ROOT, "sph-exa-nnet_amd", "lib", "libnnet_mi355x.so")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(nnet_[a-z0-9_]+)\s*\(", text)))


def header_macros():
    out = {}
    for m in re.finditer(r"#define\s+(NNET_[A-Z0-9_]+)\s+\(?(-?\d+)\)?", open(HEADER).read()):
        out[m.group(1)] = int(m.group(2))
    return out


def test_header_declares_the_entry_points():
    names = declared_functions()
    for n in ("nnet_ctx_create", "nnet_ctx_destroy", "nnet_last_error", "nnet_ctx_load_helmholtz",
              "nnet_network_create", "nnet_network_builtin", "nnet_network_destroy", "nnet_network_reactions",
              "nnet_compute", "nnet_compute_sync", "nnet_ctx_set_max_iterations"):
        assert n in names, n


@pytest.mark.skipif(not os.path.exists(LIB), reason="libnnet_mi355x.so not built (run __graft_entry__.build())")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_constants_match_header():
    import nnet_mi355x as nm
    mac = header_macros()
    assert nm.PRECISION_FP64 == mac["NNET_PRECISION_FP64"]
    assert nm.PRECISION_MIXED == mac["NNET_PRECISION_MIXED"]
    assert (nm.DRHODT_OFF, nm.DRHODT_CPU, nm.DRHODT_INTEGRATE) == (
        mac["NNET_DRHODT_OFF"], mac["NNET_DRHODT_CPU"], mac["NNET_DRHODT_INTEGRATE"])
    for name in ("NNET_STATUS_OK", "NNET_STATUS_SKIPPED", "NNET_STATUS_STALLED", "NNET_STATUS_ZERO_DT",
                 "NNET_STATUS_MAX_ITER"):
        assert name in mac
    assert sorted(nm.EXPORTS) == sorted(declared_functions())
