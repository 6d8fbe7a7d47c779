"""The counter stream's random rotation of product offset vectors
(smolrng_rotation, a Marsaglia unit quaternion): every draw is a proper
rotation, and the draws follow the Haar measure on SO(3), the law of the
reference's Euler-angle draw (DirCosMD with thetarandCCD, two uniform
angles, smolreact.c:2846-2848)."""
import ctypes as C

import numpy as np

import oracle_bind as ob


def _draws(n, step=7):
    L = ob.lib()
    out = np.zeros((n, 9))
    m = (C.c_double * 9)()
    for k in range(n):
        L.osim_rng_rotation(12345, 678, k + 1, step, m)
        out[k] = np.frombuffer(m, dtype=np.float64)
    return out.reshape(n, 3, 3)


def test_rotations_are_proper():
    R = _draws(2000)
    eye = np.einsum("nij,nkj->nik", R, R)
    assert np.abs(eye - np.eye(3)).max() < 1e-12
    assert np.abs(np.linalg.det(R) - 1).max() < 1e-12


def test_haar_moments():
    """For the Haar measure on SO(3): E[R] = 0, E[R_ij^2] = 1/3, E[tr R] = 0,
    E[(tr R)^2] = 1; and R applied to a fixed vector is uniform on the sphere."""
    n = 40000
    R = _draws(n)
    se = 1 / np.sqrt(n)
    assert np.abs(R.mean(axis=0)).max() < 5 * se * np.sqrt(1 / 3)
    assert np.abs((R ** 2).mean(axis=0) - 1 / 3).max() < 5 * se * 0.3
    tr = np.trace(R, axis1=1, axis2=2)
    assert abs(tr.mean()) < 5 * se
    assert abs((tr ** 2).mean() - 1) < 5 * se * 1.5
    v = R @ np.array([0.3, -1.2, 0.5])
    z = v[:, 2] / np.linalg.norm(v, axis=1)  # uniform on [-1, 1] for a uniform direction
    x = np.sort((z + 1) / 2)
    ks = max(np.max(np.arange(1, n + 1) / n - x), np.max(x - np.arange(n) / n))
    assert ks < 1.63 / np.sqrt(n)
