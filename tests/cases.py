"""Seeded synthetic cases shared by the tests: a nickel (or GaN) master, the Bruker geometry of
the configurations (sample tilt 70, pattern centre (0.50, 0.50, 0.55)), patterns simulated by
the oracle at uniform random orientations, optionally with Poisson noise (test_api.py:57-70 does
the same with Gaussian noise; the configurations use Poisson at a mean of 400 counts)."""
import numpy as np

from oracle import pipeline as op

PC = (0.50, 0.50, 0.55)


def random_quaternions(rng, n):
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    q[q[:, 0] < 0] *= -1
    return q


def oracle_frame(h, w, pc=PC):
    R = op.bruker_R(70.0, 0.0)
    fr = op.frame(R, h, w)
    t = op.translation(R, pc, float(h), float(w))
    return fr, t


def simulate(sheets, side, n, h=60, w=60, seed=7, counts=0.0, pc=PC):
    """(patterns float32 [n, h, w], truths [n, 4]) from the oracle's projection of `sheets`."""
    rng = np.random.default_rng(seed)
    q = random_quaternions(rng, n)
    fr, t = oracle_frame(h, w, pc)
    pats = np.stack([op.project(sheets, side, fr, t, qq) for qq in q])
    if counts > 0:
        lo = pats.min()
        pats = rng.poisson(counts * (pats - lo) / (pats.mean() - lo)).astype(np.float64)
    return pats.astype(np.float32), q
