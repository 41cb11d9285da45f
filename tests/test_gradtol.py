"""CPU check of the point-adjoint tolerance rule (tests/gradtol.py).

A float32 emulation of the adjoint (q - p, its norm, g/d times the
difference, sums in float32, the arithmetic k_point_backward performs) must
pass the rule on the C4 Gaussian mixture, where |q| ~ 1e3 and neighbour
distances go down to ~1e-4; a sign flip of one neighbour's term, or a wrong
neighbour, planted in that emulation must fail it.
"""
import numpy as np
import pytest

from gradtol import check_point_grads

F32 = np.float32


def f32_adjoint(P, Q, ids, g):
    ids = ids.reshape(len(Q), -1)
    g = g.reshape(len(Q), -1).astype(F32)
    gq = np.zeros((len(Q), 3), F32)
    gp = np.zeros((len(P), 3), F32)
    for j in range(ids.shape[1]):
        p = ids[:, j]
        ok = p >= 0
        diff = (Q[ok] - P[p[ok]]).astype(F32)
        d = np.sqrt(((diff[:, 0] * diff[:, 0] + diff[:, 1] * diff[:, 1]) + diff[:, 2] * diff[:, 2]).astype(F32))
        nz = d > 0
        t = np.zeros_like(diff)
        t[nz] = (g[ok, j][nz] / d[nz])[:, None].astype(F32) * diff[nz]
        gq[ok] = (gq[ok] + t).astype(F32)
        np.add.at(gp, p[ok], -t)
    return gq, gp


@pytest.fixture(scope="module")
def mixture(oracle_mod):
    from cubql_amd import synth
    P = synth.gaussian_mixture(20_000, seed=60)
    Q = synth.gaussian_mixture(2_000, seed=61)
    tree = oracle_mod.rebin_build(oracle_mod.points_to_boxes(P), 8)
    ids, d2, _ = oracle_mod.knn(tree, P, Q, 16)
    g = np.full((len(Q), 16), 1.0 / (len(Q) * 16))
    return P, Q, ids, g


def test_float32_adjoint_passes(oracle_mod, mixture):
    P, Q, ids, g = mixture
    gq, gp = f32_adjoint(P, Q, ids, g)
    check_point_grads(oracle_mod, P, Q, ids, g, gq, gp, label="f32 emulation")


def test_sign_flip_fails(oracle_mod, mixture):
    P, Q, ids, g = mixture
    g2 = g.copy()
    g2[::7, 3] *= -1.0           # one neighbour's term flipped on every 7th query
    gq, gp = f32_adjoint(P, Q, ids, g2)
    with pytest.raises(AssertionError):
        check_point_grads(oracle_mod, P, Q, ids, g, gq, label="sign flip")


def test_wrong_neighbour_fails(oracle_mod, mixture):
    P, Q, ids, g = mixture
    ids2 = ids.copy()
    ids2[::5, 15] = ids[::5, 14]     # the k-th neighbour replaced by the (k-1)-th
    gq, gp = f32_adjoint(P, Q, ids2, g)
    with pytest.raises(AssertionError):
        check_point_grads(oracle_mod, P, Q, ids, g, gq, gp, label="wrong neighbour")
