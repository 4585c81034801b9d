"""Writes golden vectors for the structural hot path from the reference's own
C extension modules (compiled from the reference sources by
oracle/build_ref.sh into oracle/_ref).  Run in the build container:

    sh oracle/build_ref.sh && python tests/golden/make_golden.py

Each fixture is an .npz of seeded inputs and the reference outputs; tests
check the CPU oracle against them bit for bit and the GPU path within its
stated tolerance.  Inputs are the project's own seeded synthetics.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import ref  # noqa: E402
from pyseistr_amd import synth  # noqa: E402


def save(name, **arrays):
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, **{k: np.asarray(v) for k, v in arrays.items()})
    print(name, sum(np.asarray(v).size for v in arrays.values()))


def main():
    if not ref.available():
        raise SystemExit("oracle/_ref not built: run sh oracle/build_ref.sh")
    # 3-D dips, somean/somf, interpolation
    s3 = (48, 12, 10)
    v, clean = synth.volume3d(s3, seed=11, n_events=3, smax=1.5, sigma=0.2, spike_frac=0.01)
    dip3 = ref.dipc(v, s3, niter=5, liter=10, order=2, rect=(5, 5, 5))
    save("dip3d", data=v, shape=s3, niter=5, liter=10, order=2, rect=(5, 5, 5), dip=dip3)
    m = (np.random.default_rng(12).random(s3) > 0.05).astype(np.float32)
    dip3m = ref.dipc(v, s3, niter=3, liter=10, order=2, rect=(3, 3, 3), mask=m)
    save("dip3d_mask", data=v, mask=m, shape=s3, niter=3, liter=10, order=2, rect=(3, 3, 3), dip=dip3m)
    n = int(np.prod(s3))
    pi, px = dip3[:n], dip3[n:]
    save("somean3d", data=v, dipi=pi, dipx=px, shape=s3, r1=2, r2=2, order=2,
         out=ref.somean3d(v, pi, px, s3, 2, 2, order=2))
    save("somf3d", data=v, dipi=pi, dipx=px, shape=s3, r1=2, r2=2, order=2,
         out=ref.somf3d(v, pi, px, s3, 2, 2, order=2),
         out_r3=ref.somf3d(v, pi, px, s3, 3, 3, order=2),
         out_o1=ref.somf3d(v, pi, px, s3, 2, 1, order=1),
         out_sv=ref.somf3d(v, pi, px, s3, 2, 2, order=2, option=2))
    obs, mask, comp = synth.interp3d(s3, seed=13, keep=0.4, n_events=2)
    dipc = ref.dipc(comp, s3, niter=3, liter=10, order=2, rect=(3, 3, 3))
    qi, qx = dipc[:n], dipc[n:]
    save("soint3d", data=obs, mask=mask, dipi=qi, dipx=qx, shape=s3, order=2, niter=15,
         out=ref.soint3d(obs, mask, qi, qx, s3, order=2, niter=15),
         out_var=ref.soint3d(obs, mask, qi, qx, s3, order=2, niter=4, var=0.05),
         out_nomask=ref.soint3d(obs, mask, qi, qx, s3, order=1, niter=10, hasmask=0))
    save("sint3d", data=obs, mask=mask, dipi=qi, dipx=qx, shape=s3, niter=8, ns1=2, ns2=2, order=2,
         out=ref.sint3d(obs, mask, qi, qx, s3, niter=8, eps=0.01, ns1=2, ns2=2, order1=2, order2=2))
    # 2-D
    g = synth.gather2d(64, 24, seed=14)
    dip2 = ref.dipc(g, (64, 24, 1), niter=5, liter=20, order=2, rect=(10, 20, 1))
    save("dip2d", data=g, shape=(64, 24, 1), niter=5, liter=20, order=2, rect=(10, 20, 1), dip=dip2)
    dip2b = ref.dipc(g, (64, 24, 1), niter=4, liter=15, order=1, rect=(4, 6, 1))
    save("dip2d_o1", data=g, shape=(64, 24, 1), niter=4, liter=15, order=1, rect=(4, 6, 1), dip=dip2b)
    d3 = v[:, :, :3].copy()
    p3 = pi.reshape(s3, order="F")[:, :, :3].copy()
    s23 = d3.shape
    save("somean2d", data=d3, dip=p3, shape=s23, ns=3, order=2,
         out=ref.somean2d(d3, p3, s23, 3, order=2), out_adj=ref.somean2d(d3, p3, s23, 3, order=2, adj=1))
    save("somf2d", data=d3, dip=p3, shape=s23, ns=3, order=2,
         out=ref.somf2d(d3, p3, s23, 3, order=2), out_sv=ref.somf2d(d3, p3, s23, 3, order=2, option=2))
    o2, m2 = obs[:, :, 0].copy(), mask[:, :, 0].copy()
    q2 = qi.reshape(s3, order="F")[:, :, 0].copy()
    save("soint2d", data=o2, mask=m2, dip=q2, shape=(48, 12), order=2, niter=15,
         out=ref.soint2d(o2, m2, q2, (48, 12), order=2, niter=15),
         out_nomask=ref.soint2d(o2, m2, q2, (48, 12), order=1, niter=10, hasmask=0))
    save("sint2d", data=o2, mask=m2, dip=q2, shape=(48, 12), niter=15, ns=3, order=2,
         out=ref.sint2d(o2, m2, q2, (48, 12), niter=15, eps=0.01, ns=3, order=2))


if __name__ == "__main__":
    main()
