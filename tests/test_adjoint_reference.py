"""The sparse FP64 reference gradient (tests/adjoint_ref.py) agrees with the
dense one (numpy solve, per-branch dense Y rebuilds) on the 200-bus grid at
the oracle's own solution.  CPU only."""
import numpy as np

from adjoint_ref import reference_gradient
from gap_amd import gridgen
from test_gpu_adjoint import dense_reference, rel_err


def test_sparse_reference_matches_dense(oracle):
    base = 100e6
    net = gridgen.small_mesh(200, seed=0)
    o = oracle.solve(net, tol=1e-12, max_it=60, j_call=-1)
    assert o["converged"]
    v = o["v"]
    gp, gq, gr, gx, L = reference_gradient(oracle, net, v, base, 1.0, 2.0)
    dp, dq, dr, dx, dL = dense_reference(oracle, net, v, base, 1.0, 2.0, None, None)
    assert abs(L - dL) <= 1e-13 * abs(dL)
    for a, b in ((gp, dp), (gq, dq), (gr, dr), (gx, dx)):
        assert rel_err(a, b) < 1e-8
