"""CPU checks of the boundary: the built library loads and exports exactly the
entry points the headers declare; the ctypes binding covers all of them."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "suqa_amd", "lib", "libsuqa_hip.so")


def declared(header, prefix):
    with open(os.path.join(ROOT, "include", header)) as f:
        src = f.read()
    return set(re.findall(r"\b(" + prefix + r"[a-zA-Z0-9_]+)\s*\(", src))


def exported(lib, prefix):
    out = subprocess.run(["nm", "-D", "--defined-only", lib], check=True, capture_output=True, text=True).stdout
    return {ln.split()[-1] for ln in out.splitlines() if ln.split() and ln.split()[-1].startswith(prefix)}


@pytest.fixture(scope="module")
def built():
    if not os.path.exists(LIB):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "suqa_amd")], check=True)
    return LIB


def test_header_matches_exports(built):
    decl = declared("suqa_hip.h", "suqa_hip_")
    assert len(decl) >= 30
    assert decl == exported(built, "suqa_hip_")


def test_qms_header_matches_exports(built):
    lib = os.path.join(ROOT, "suqa_amd", "lib", "libsuqa.so")
    decl = declared("suqa_qms.h", "suqa_qms_")
    assert decl == {"suqa_qms_run", "suqa_qms_batch_run"}
    assert decl == exported(lib, "suqa_qms_")
    from suqa_amd import qms
    assert set(qms.SIGNATURES) == decl


def test_qsa_header_matches_exports(built):
    lib = os.path.join(ROOT, "suqa_amd", "lib", "libsuqa.so")
    decl = declared("suqa_qsa.h", "suqa_qsa_")
    assert decl == {"suqa_qsa_run"}
    assert decl == exported(lib, "suqa_qsa_")
    from suqa_amd import qsa
    assert set(qsa.SIGNATURES) == decl


def test_binding_covers_header(built):
    from suqa_amd import capi
    assert set(capi.SIGNATURES) == declared("suqa_hip.h", "suqa_hip_")


def test_binding_constants_match_header():
    from suqa_amd import capi
    with open(os.path.join(ROOT, "include", "suqa_hip.h")) as f:
        src = f.read()
    assert int(re.search(r"#define SUQA_HIP_OUTCOME_SLOTS (\d+)u", src).group(1)) == capi.OUTCOME_SLOTS


def test_library_loads_without_gpu(built):
    lib = ctypes.CDLL(built)
    for name in declared("suqa_hip.h", "suqa_hip_"):
        assert hasattr(lib, name)
    lib.suqa_hip_last_error.restype = ctypes.c_char_p
    assert lib.suqa_hip_last_error() == b""


def test_binding_fails_loudly_without_library(tmp_path):
    from suqa_amd import capi
    with pytest.raises(ImportError):
        capi.load_library(str(tmp_path / "missing.so"))
