"""The oracle: a float64 CPU restatement of the reference's indexing path, for tests only.

TEST INFRASTRUCTURE. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import it, and only to check results (or as the timed CPU baseline): the product path
(zebsdsolve_amd/) never calls it and has no CPU fallback.

Where it is cheap the oracle takes a different route to the same numbers than the product, so
that the two cannot share a mistake: numpy's Gauss-Legendre nodes, scipy's spherical harmonics,
Wigner d from the eigenvectors of J_y (checked against the exact rational Wigner sum), and the
correlation's transform along b as d^l_{m'm}(beta_j) directly instead of the reference's
contraction at pi/2 followed by an FFT. Every function cites the reference code it follows.
"""
