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


def test_patch_count_host_call():
    """sh_patch_count is host arithmetic only: it agrees with the patch
    oracle's counts and rejects a patch that fits nowhere."""
    import numpy as np
    import pytest
    from oracle import patch_oracle as P
    from pyseistr_amd import patch
    for n, l, s in [((100, 60, 1), (8, 8, 1), (4, 4, 1)), ((101, 37, 9), (16, 10, 3), (8, 3, 2)),
                    ((20, 9, 3), (3, 12, 2), (5, 4, 3)), ((3, 5, 4), (5, 5, 2), (3, 3, 1))]:
        assert patch.patch_count(n, l, s) == int(np.prod(P.counts(n, l, s)))
    with pytest.raises(ValueError):
        patch.patch_count((3, 4, 1), (5, 2, 1), (2, 1, 1))


def test_neighbours_have_no_cpu_fallback():
    import torch
    if torch.cuda.is_available():
        return
    import numpy as np
    import pytest
    import pyseistr_amd as ps
    for call in (lambda: ps.patch2d(np.zeros((16, 16))), lambda: ps.fkdip(np.zeros((16, 8)), 0.1),
                 lambda: ps.smoothc(np.zeros((16, 8)), rect=[3, 1, 1]), lambda: ps.bandpassc(np.zeros(64), 0.004)):
        with pytest.raises(ps.SeistrError):
            call()


def test_complex_input_rejected_before_any_device_call():
    import numpy as np
    import pytest
    import pyseistr_amd as ps
    z = np.zeros((16, 16), np.complex64)
    for call in (lambda: ps.patch2d(z), lambda: ps.patch2d_inv(np.zeros((9, 64), complex), 16, 16),
                 lambda: ps.fkdip(z, 0.1)):
        with pytest.raises(TypeError):
            call()


def test_batch_tile_picker():
    """The batched path's tile query and picker run on the host: tile -1
    is the library default; the picker keeps the default unless the table
    predicts at least min_gain less time for the mix."""
    from pyseistr_amd import batch
    tiles = batch.tile_shapes()
    d = batch.default_tile()
    assert 0 <= d < len(tiles)
    base = [1.0] * len(tiles)
    other = (d + 1) % len(tiles)

    def table(gain):
        ns = list(base)
        ns[other] = 1.0 - gain
        return {"tiles": [list(t) for t in tiles], "buckets": {"128x64": {"ns_per_sample": ns}}}

    assert batch.pick_tile([(128, 64)] * 4, table(0.05)) == d
    assert batch.pick_tile([(128, 64)] * 4, table(0.20)) == other
    assert batch.pick_tile([(128, 64)] * 4, table(0.05), min_gain=0.03) == other
    assert batch.pick_tile([(100, 70)], table(0.20), min_gain=0.3) == d
