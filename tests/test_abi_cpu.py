"""C-ABI checks that need no GPU: the library loads, exports every symbol the
header declares, and the argument checks that run before any device call
return the reference's codes (cuda_replace.cu:2258-2264, 2301-2302)."""
import ctypes as C
import os
import re

import pytest

from cuda_replace_amd import REQUIRED_SYMBOLS, CudaReplaceLib
from cuda_replace_amd.wrapper import DEFAULT_LIB

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "cuda_replace.h")

# the twelve names the reference exports (cuda_replace.cu:2161-2603)
REFERENCE_ABI = {
    "cuda_replace_unified", "cuda_free_host", "cuda_replace_open", "cuda_replace_reset",
    "cuda_replace_build_index", "cuda_replace_apply", "cuda_replace_apply_seeded",
    "cuda_replace_query_prefix", "cuda_replace_apply_batch", "cuda_replace_result",
    "cuda_replace_close", "cuda_replace_version",
}


def header_symbols():
    text = open(HEADER).read()
    return set(re.findall(r"^CR_API\s+[\w\s\*]*?\b(\w+)\(", text, flags=re.M))


@pytest.fixture(scope="module")
def cpu_lib():
    if not os.path.exists(DEFAULT_LIB):
        pytest.fail("libcuda_replace.so is not built (python -m cuda_replace_amd.build)")
    return CudaReplaceLib()


def test_header_declares_reference_abi():
    syms = header_symbols()
    assert REFERENCE_ABI <= syms
    assert syms == set(REQUIRED_SYMBOLS)


def test_library_exports_every_declared_symbol(cpu_lib):
    raw = C.CDLL(DEFAULT_LIB)
    for name in header_symbols():
        assert hasattr(raw, name), name


def test_version(cpu_lib):
    assert "gfx950" in cpu_lib.version()


def test_null_arguments(cpu_lib):
    lib = cpu_lib.lib
    assert lib.cuda_replace_open(None, None, 0) == -1
    assert lib.cuda_replace_apply(None, None, 0, None, 0, None) == -1
    assert lib.cuda_replace_apply_seeded(None, None, 0, None, 0, 0, None, None) == -1
    assert lib.cuda_replace_reset(None, None, 0) == -1
    assert lib.cuda_replace_build_index(None, 0) == -1
    assert lib.cuda_replace_query_prefix(None, 0, None, None) == -1
    assert lib.cuda_replace_apply_batch(None, None, None, None, None, 0, None) == -1
    assert lib.cuda_replace_result(None, None, None) == -1
    assert lib.cr_session_length(None, None) == -1
    out = C.POINTER(C.c_uint8)()
    n = C.c_size_t()
    assert lib.cuda_replace_unified(None, 0, None, 0, None, 0, None, None) == -1
    lib.cuda_replace_close(None)
    lib.cuda_free_host(None)
    del out, n


def test_pattern_mode_classification(cpu_lib):
    # border-free patterns take the single-pass path
    for p in [b"ab", b"qarn", b"&amp;", b"x", b"NEEDLE", b"REPLACE", b"hello", b"abc" * 85 + b"d"]:
        assert cpu_lib.pattern_mode(p) == 1, p
    # self-overlapping, length <= 16: carry tables
    for p in [b"aa", b"aaa", b"aba", b"abcab", b"a" * 16, b"xx", b"abab"]:
        assert cpu_lib.pattern_mode(p) == 2, p
    # long self-overlapping, or longer than the single-pass halo
    for p in [b"a" * 17, b"ab" * 9, b"abcd" * 64 + b"e"]:
        assert cpu_lib.pattern_mode(p) == 3, p
    assert cpu_lib.pattern_mode(b"") == 0


def test_open_without_device_fails_cleanly(cpu_lib):
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    h = C.c_void_p()
    rc = cpu_lib.lib.cuda_replace_open(C.byref(h), None, 0)
    assert rc > 0 and not h.value
