"""CPU checks of the C-ABI boundary: the library loads, exports every symbol that
include/pittpack_hip.h declares, and its host-only helpers answer without a GPU.  No
compute call is made here; without a visible GPU the product path must fail loudly."""
import os
import re

import pytest

import pittpack_amd as pa

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "pittpack_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pp_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_python_export_list():
    assert declared_symbols() == sorted(pa.EXPORTED)


def test_library_exports_every_declared_symbol():
    L = pa.load_library()
    for name in declared_symbols():
        assert hasattr(L, name), name


def test_error_strings_match_reference_enum():
    # PittPackGetErrorEnum numbering (src/include/params.h:86-114)
    assert pa.error_name(0) == "SUCCESS"
    assert pa.error_name(15) == "ALLOCATION_FAIL"
    assert pa.error_name(19) == "INPUT_ARGS"
    assert pa.error_name(23) == "CUFFT_FAIL_INV_Y"
    assert pa.error_name(pa.UNSUPPORTED_BC) == "UNSUPPORTED_BC"
    assert pa.error_name(pa.NO_DEVICE) == "NO_DEVICE"


def test_supported_sizes():
    # every line length from 2 to 1024: compiled Stockham plans for even N with N/2 in the
    # radix-2/3/5/7 list, the arbitrary-length transforms for the rest
    for n in range(2, 1025):
        assert pa.size_supported(n), n
    for n in (0, 1, 1025, 1280, 2048):
        assert not pa.size_supported(n), n
    assert pa.error_name(pa.IO_FAIL) == "IO_FAIL"


def test_no_cpu_fallback_without_gpu():
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is visible: covered by the gpu tests")
    with pytest.raises(pa.PittPackError) as e:
        pa.PoissonHIP(32, 32, 16)
    assert e.value.code == pa.NO_DEVICE
