"""Write pll_tridiag_ref.npz: inputs and outputs of the reference's own
pll_tridiag (pll_tridiag.c, built by `make -C oracle ref` into oracle/_ref)
on seeded diagonally dominant systems, incl. lengths 1, 2 and the shapes of
the ADI systems (a = c = -fac*coef, b = 1 + 2 fac coef).
Run in the container where the reference tree exists.
"""
import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF = os.path.join(HERE, "..", "..", "oracle", "_ref", "ref_tridiag")


def systems():
    rng = np.random.default_rng(20240601)
    out = []
    for n, ns in [(1, 3), (2, 3), (7, 5), (37, 8), (256, 4)]:
        a = -rng.uniform(0.0, 1.0, (ns, n))
        c = -rng.uniform(0.0, 1.0, (ns, n))
        b = 2.2 + rng.uniform(0.0, 1.0, (ns, n))
        r = rng.normal(size=(ns, n))
        out.append((a, b, c, r))
    # ADI-like: 1 - fac*(-2 coef) on the diagonal, Neumann-modified end rows
    n, ns = 128, 4
    coef = np.exp(rng.uniform(0, 3, (ns, n + 1))) * 100.0
    fac = 0.5 * 2e-3 * (8.0 / 15.0)
    a = -fac * coef[:, :-1]
    c = -fac * coef[:, 1:]
    b = 1.0 - fac * (-(coef[:, :-1] + coef[:, 1:]))
    b[:, 0] += a[:, 0]
    b[:, -1] += c[:, -1]
    out.append((a, b, c, rng.normal(size=(ns, n))))
    return out


def main():
    data = {}
    for t, (a, b, c, r) in enumerate(systems()):
        ns, n = a.shape
        with tempfile.TemporaryDirectory() as d:
            fi, fo = os.path.join(d, "in.bin"), os.path.join(d, "out.bin")
            with open(fi, "wb") as f:
                np.array([ns, n], dtype=np.int32).tofile(f)
                np.concatenate([a.ravel(), b.ravel(), c.ravel(), r.ravel()]).tofile(f)
            subprocess.run([REF, fi, fo], check=True)
            o = np.fromfile(fo).reshape(3, ns, n)
        for k, v in zip("abcr", (a, b, c, r)):
            data["in_%s_%d" % (k, t)] = v
        for k, v in zip("bcr", o):
            data["out_%s_%d" % (k, t)] = v
    data["ncases"] = np.array(len(systems()))
    np.savez_compressed(os.path.join(HERE, "pll_tridiag_ref.npz"), **data)


if __name__ == "__main__":
    main()
