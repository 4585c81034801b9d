"""The interpolation solvers' device-resident control flow.

soint (ps_solver + ps_cgstep), soint2dc(prec=1) (ps_solver_prec) and sint
(ps_conjgrad) keep every scalar on the device, so the iterations are queued
without host round trips; one iteration is captured into a HIP graph and
replayed (SH_SOLVER_GRAPH=0 launches the kernels directly).  Checked here:
 * graph replays and direct launches give identical bits, on repeated calls
   (the graph is captured anew per call and replayed niter times, with the
   host's own work and synchronisations in between);
 * long solves on small problems, which stop early on the reference's
   tolerance tests, equal the FP32 oracle bit for bit (a missed or late
   stop would move the model);
 * the LDS-tiled i2 <-> i3 permutation of sint3d equals the plain one."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def canon(a):
    a = np.asarray(a, np.float32).ravel(order="F")
    return np.where(a == 0, np.float32(0), a).view(np.uint32)


def same(a, b, what=""):
    assert np.array_equal(canon(a), canon(b)), what


def F(a):
    return np.asarray(a, np.float32).ravel(order="F")


@pytest.fixture(scope="module")
def ps():
    import pyseistr_amd
    return pyseistr_amd


def _case3(shape, seed):
    rng = np.random.default_rng(seed)
    d = rng.standard_normal(shape).astype(np.float32)
    m = (rng.random(shape[1:]) > 0.6).astype(np.float32)[None] * np.ones(shape, np.float32)
    p = (0.6 * rng.standard_normal(shape)).astype(np.float32)
    q = (0.6 * rng.standard_normal(shape)).astype(np.float32)
    return (d * m).astype(np.float32), m, p, q


def _calls(ps, d, m, p, q):
    d2, m2, p2 = d[:, :, 0], m[:, :, 0], p[:, :, 0]
    return {
        "soint3d": lambda: ps.soint3dc(d, m, p, q, order=2, niter=12, report=True),
        "soint3d drift": lambda: ps.soint3dc(d, m, 3 * p, 3 * q, order=1, niter=7, drift=1, report=True),
        "soint2d": lambda: ps.soint2dc(d2, m2, p2, order=2, niter=12, report=True),
        "soint2d prec": lambda: ps.soint2dc(d2, m2, p2, order=2, niter=6, prec=1, report=True),
        "sint3d": lambda: ps.sint3dc(d, m, p, q, niter=9, eps=0.01, ns1=2, ns2=2, order1=2, order2=2, report=True),
        "sint2d": lambda: ps.sint2dc(d2, m2, p2, niter=9, eps=0.01, ns=3, order=2, report=True),
    }


def test_graph_replays_equal_direct_launches(ps, monkeypatch):
    d, m, p, q = _case3((48, 20, 16), 5)
    calls = _calls(ps, d, m, p, q)
    graph = {k: f() for k, f in calls.items()}
    again = {k: f() for k, f in calls.items()}
    monkeypatch.setenv("SH_SOLVER_GRAPH", "0")
    direct = {k: f() for k, f in calls.items()}
    monkeypatch.delenv("SH_SOLVER_GRAPH")
    for k in calls:
        print(k, "iterations", graph[k][1], flush=True)
        same(graph[k][0], direct[k][0], k)
        same(again[k][0], direct[k][0], k + " (second call)")
        assert graph[k][1] == direct[k][1] == again[k][1], k


@pytest.mark.parametrize("graph", ["1", "0"])
def test_long_solves_stop_like_the_oracle(ps, monkeypatch, graph):
    from oracle import oracle as O
    monkeypatch.setenv("SH_SOLVER_GRAPH", graph)
    rng = np.random.default_rng(17)
    n1, n2 = 24, 10
    d = rng.standard_normal((n1, n2)).astype(np.float32)
    m = (rng.random((n1, n2)) > 0.5).astype(np.float32)
    d = d * m
    p = (0.4 * rng.standard_normal((n1, n2))).astype(np.float32)
    s = (n1, n2)
    got, it = ps.sint2dc(d, m, p, niter=400, eps=0.01, ns=2, order=2, report=True)
    same(got, O.sint2d(F(d), F(m), F(p), s, niter=400, eps=0.01, ns=2, order=2), "sint2d")
    print("sint2d stopped after", it, flush=True)
    got, it2 = ps.soint2dc(d, m, p, order=2, niter=400, report=True)
    same(got, O.soint2d(F(d), F(m), F(p), s, order=2, niter=400), "soint2d")
    print("soint2d stopped after", it2, flush=True)
    got, it3 = ps.soint2dc(d, m, p, order=2, niter=300, prec=1, report=True)
    same(got, O.soint2d(F(d), F(m), F(p), s, order=2, niter=300, prec=1), "soint2d prec")
    print("soint2d prec stopped after", it3, flush=True)
    assert it < 400 or it2 < 400 or it3 < 300, "no solve stopped early; the test no longer covers the stop"


def test_tiled_permutation_equals_plain(ps, monkeypatch):
    d, m, p, q = _case3((40, 37, 45), 8)
    a = ps.sint3dc(d, m, p, q, niter=4, eps=0.01, ns1=2, ns2=1, order1=2, order2=2)
    monkeypatch.setenv("SH_PERMUTE_PLAIN", "1")
    b = ps.sint3dc(d, m, p, q, niter=4, eps=0.01, ns1=2, ns2=1, order1=2, order2=2)
    monkeypatch.delenv("SH_PERMUTE_PLAIN")
    same(a, b)


def test_spray2_prefetch_distances_identical(ps, monkeypatch):
    """The 2-D chain kernels' load rings at 8 and 16 samples ahead, and their
    buffer-resource and pointer accessors (SH_SPRAY2_BUF=0), give the same
    bits (spray, adjoint spray, somean2d and sint3d's shaping CG)."""
    d, m, p, q = _case3((60, 23, 9), 11)
    d2, p2 = d[:, :, 0], p[:, :, 0]
    rng = np.random.default_rng(3)
    u = rng.standard_normal(60 * 23 * 7).astype(np.float32)

    from pyseistr_amd import filters

    def calls():
        return [filters.pwspray2d(d2, p2, 3, 2, 0.01),
                filters.pwspray2d_adjoint(u, p2, 3, 2, 0.01, (60, 23)),
                ps.somean2dc(d, p, 4, 2, 0.01),
                ps.sint3dc(d, m, p, q, niter=5, eps=0.01, ns1=3, ns2=2, order1=2, order2=2)]

    out = {}
    for pf in ("8", "16"):
        monkeypatch.setenv("SH_SPRAY2_PF", pf)
        out[pf] = calls()
    monkeypatch.delenv("SH_SPRAY2_PF")
    monkeypatch.setenv("SH_SPRAY2_BUF", "0")
    out["ptr"] = calls()
    monkeypatch.delenv("SH_SPRAY2_BUF")
    for k, (a, b, c) in enumerate(zip(out["8"], out["16"], out["ptr"])):
        same(a, b, f"call {k}")
        same(a, c, f"call {k}, pointers")


def test_spray3_level_paths_identical(ps, monkeypatch):
    """The 3-D spray levels give the same bits through buffer resources (the
    default) and through pointers (SH_SPRAY_BUF=0), at every prefetch
    distance (SH_SPRAY_PF) and with two-source slots grouped per thread
    (SH_SPRAY_GROUP): pwspray3d, somean3dc and somf3dc, PWD orders 1-3."""
    d, _, p, q = _case3((52, 17, 14), 31)

    def calls():
        return [ps.pwspray3d(d, p, q, 2, 3, 2, 0.01), ps.pwspray3d(d, p, q, 1, 2, 1, 0.05),
                ps.pwspray3d(d, p, q, 2, 1, 3, 0.01), ps.somean3dc(d, p, q, 3, 2, 0.01, 2),
                ps.somf3dc(d, p, q, 4, 4, 0.01, 2)]

    ref = calls()
    for env in ({"SH_SPRAY_BUF": "0"}, {"SH_SPRAY_PF": "4"}, {"SH_SPRAY_PF": "8"},
                {"SH_SPRAY_PF": "2", "SH_SPRAY_BUF": "0"}, {"SH_SPRAY_PF": "0"}, {"SH_SPRAY_GROUP": "2"}):
        for k, v in env.items():
            monkeypatch.setenv(k, v)
        for k, (a, b) in enumerate(zip(calls(), ref)):
            same(a, b, f"{env} call {k}")
        for k in env:
            monkeypatch.delenv(k)


def test_pwd_virtual_blocks_identical(ps, monkeypatch):
    """soint's operator kernels give the same bits with 1, 2 or 4 groups of
    samples per block (the partial-sum partition is unchanged), with and
    without drift, in 3-D and 2-D."""
    d, m, p, q = _case3((45, 19, 13), 21)
    d2, m2, p2 = d[:, :, 0], m[:, :, 0], p[:, :, 0]

    def calls():
        return [ps.soint3dc(d, m, p, q, order=2, niter=9, report=True),
                ps.soint3dc(d, m, 3 * p, 3 * q, order=1, niter=6, drift=1, report=True),
                ps.soint2dc(d2, m2, p2, order=3, niter=9, report=True)]

    out = {}
    for v in ("1", "2", "4"):
        monkeypatch.setenv("SH_PWD_V", v)
        out[v] = calls()
    monkeypatch.delenv("SH_PWD_V")
    for v in ("2", "4"):
        for k, ((a, ia), (b, ib)) in enumerate(zip(out["1"], out[v])):
            same(a, b, f"V={v} call {k}")
            assert ia == ib
