"""The CPU oracle built with AddressSanitizer and UndefinedBehaviorSanitizer
(make -C oracle sanitize) runs every routine of the path on small inputs
in a child process (the ASan runtime preloaded, leak checking off since
the interpreter itself is not instrumented) without a report, and gives
the same bits as the plain build."""
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SCRIPT = r'''
import sys, numpy as np
sys.path.insert(0, ROOT)
from oracle import oracle as O
F = lambda a: np.asarray(a, np.float32).ravel(order="F")
rng = np.random.default_rng(0)
out = {}
for dt in (np.float32, np.float64):
    t = "32" if dt == np.float32 else "64"
    s3 = (20, 7, 6); s2 = (20, 9, 1)
    v = rng.standard_normal(s3).astype(np.float32)
    g = rng.standard_normal(s2[:2]).astype(np.float32)
    p = (0.6 * rng.standard_normal(s3)).astype(np.float32)
    q = (0.6 * rng.standard_normal(s3)).astype(np.float32)
    m = (rng.random(s3) > 0.4).astype(np.float32)
    out["dip3" + t] = O.dip(v, s3, niter=2, liter=5, order=2, rect=(3, 3, 2), dtype=dt)
    out["dip3m" + t] = O.dip(v, s3, niter=2, liter=5, order=1, rect=(2, 2, 2), mask=m, dtype=dt,
                             p_init=np.concatenate([F(p), F(q)]))
    out["dip2" + t] = O.dip(g, s2, niter=2, liter=5, order=3, rect=(4, 3, 1), dtype=dt)
    out["spray3" + t] = O.pwspray3d(v, p, q, s3, 2, 2, order=2, dtype=dt)
    out["mean3" + t] = O.somean3d(v, p, q, s3, 2, 1, order=2, dtype=dt)
    out["mf3" + t] = O.somf3d(v, p, q, s3, 2, 2, order=2, dtype=dt)
    out["svmf3" + t] = O.somf3d(v, p, q, s3, 2, 2, order=2, option=2, dtype=dt)
    pg = (0.6 * rng.standard_normal(s2[:2])).astype(np.float32)
    out["spray2" + t] = O.pwspray2d(F(g), F(pg), s2[:2], 3, order=2, dtype=dt)
    out["spray2a" + t] = O.pwspray2d(rng.standard_normal(20 * 9 * 7).astype(np.float32), F(pg), s2[:2], 3, order=2,
                                     adj=1, dtype=dt)
    out["mean2" + t] = O.somean2d(F(v), F(p), (20, 7, 6), 2, order=2, dtype=dt)
    out["mf2" + t] = O.somf2d(F(v), F(p), (20, 7, 6), 3, order=2, option=2, dtype=dt)
    out["soint3" + t] = O.soint3d(F(v * m), F(m), F(p), F(q), s3, order=2, niter=4, dtype=dt)
    out["soint3d" + t] = O.soint3d(F(v * m), F(m), F(3 * p), F(3 * q), s3, order=1, niter=3, drift=1, dtype=dt)
    out["soint3v" + t] = O.soint3d(F(v * m), F(m), F(p), F(q), s3, order=2, niter=3, var=0.1, dtype=dt)
    m2 = m[:, :, 0]; g2 = v[:, :, 0] * m2; p2 = p[:, :, 0]
    out["soint2" + t] = O.soint2d(F(g2), F(m2), F(p2), (20, 7), order=2, niter=4, dtype=dt)
    out["soint2p" + t] = O.soint2d(F(g2), F(m2), F(p2), (20, 7), order=2, niter=3, prec=1, dtype=dt)
    out["soint2t" + t] = O.soint2d(F(g2), F(m2), F(p2), (20, 7), order=2, niter=3, prec=1, twoplane=1,
                                   dip2=F(q[:, :, 0]), dtype=dt)
    out["sint3" + t] = O.sint3d(F(v * m), F(m), F(p), F(q), s3, niter=3, eps=0.01, ns1=2, ns2=1, order1=2,
                                order2=1, dtype=dt)
    out["sint2" + t] = O.sint2d(F(g2), F(m2), F(p2), (20, 7), niter=3, eps=0.01, ns=2, order=2, dtype=dt)
    out["smooth" + t] = O.smoothc(F(v), s3, (3, 4, 9), (1, 0, 0), (0, 1, 0), 2, 1, dtype=dt)
    out["smooth0" + t] = O.smoothc(F(v), s3, (5, 2, 2), adj=0, dtype=dt)
    out["bp" + t] = O.bandpass(F(v), s3, 0.004, 5.0, 40.0, 6, 5, 0, dtype=dt)
    out["bp1" + t] = O.bandpass(F(v), s3, 0.004, 5.0, 40.0, 3, 3, 1, dtype=dt)
    out["pwd" + t] = O.pwd(F(v), F(p), s3, 2, nw=2, der=1, dtype=dt)
    out["ap3" + t] = O.allpass3(F(v), F(p), F(q), s3, adj=0, nw=2, drift=1, dtype=dt)
    out["ap3a" + t] = O.allpass3(rng.standard_normal(2 * v.size), F(p), F(q), s3, adj=1, nw=2, dtype=dt)
    out["tri" + t] = O.trianglen(F(v), s3, (3, 2, 4), dtype=dt)
    out["p2" + t] = O.predict2(v[:, 0, 0], v[:, 1, 0], p[:, 0, 0], q[:, 0, 0], 1, 0, dtype=dt)
np.savez(OUTFILE, **out)
'''


def _run(tmp_path, sanitized):
    outfile = str(tmp_path / ("san.npz" if sanitized else "plain.npz"))
    env = dict(os.environ)
    env["ORACLE_SANITIZED"] = "1" if sanitized else "0"
    if sanitized:
        asan = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()
        env["LD_PRELOAD"] = asan
        env["ASAN_OPTIONS"] = "detect_leaks=0:abort_on_error=1"
        env["UBSAN_OPTIONS"] = "halt_on_error=1:print_stacktrace=1"
    code = SCRIPT.replace("ROOT", repr(ROOT)).replace("OUTFILE", repr(outfile))
    r = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-4000:]
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr[-4000:]
    return np.load(outfile)


def test_oracle_under_asan_ubsan(tmp_path):
    try:
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize"], check=True, timeout=300)
    except (OSError, subprocess.CalledProcessError) as e:
        pytest.skip(f"sanitizer build unavailable: {e}")
    san = _run(tmp_path, True)
    plain = _run(tmp_path, False)
    assert set(san.files) == set(plain.files) and len(san.files) > 50
    for k in plain.files:
        a, b = np.asarray(plain[k]), np.asarray(san[k])
        assert a.shape == b.shape and np.array_equal(a.view(np.uint8), b.view(np.uint8)), k
