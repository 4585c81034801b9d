"""The C ABI library loads on the CPU host and exports every function that
include/seistr_hip.h declares (no GPU calls)."""
import os
import re

from pyseistr_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    txt = open(os.path.join(ROOT, "include", "seistr_hip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(sh_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    names = declared()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_declares_every_header_symbol():
    assert set(declared()) <= set(_lib.EXPORTED)


def test_no_cpu_fallback():
    import torch
    if torch.cuda.is_available():
        return
    import numpy as np
    import pytest
    import pyseistr_amd as ps
    with pytest.raises(ps.SeistrError):
        ps.dip2dc(np.zeros((16, 8), np.float32))
