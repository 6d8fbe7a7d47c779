"""3D triangle, cylinder and hemisphere panels in the oracle: closed shapes
built from them keep the molecules they reflect on the side they started on.
A tetrahedron of 4 triangles, a capsule made of a cylinder and two
hemisphere caps, and a sphere inside the capsule reflect B; over many large
steps no B crosses any of them in either direction, while A, which crosses
with probability 0.5, does.  A wrong normal, a wrong crossing test or a
reflection to the wrong side lets molecules leak."""
import numpy as np

import simspec as ss


def _in_tet(p, v):
    v = np.asarray(v)
    m = np.stack([v[1] - v[0], v[2] - v[0], v[3] - v[0]], axis=1)
    lam = np.linalg.solve(m, (p - v[0]).T).T
    return np.all(lam > 0, axis=1) & (lam.sum(axis=1) < 1)


def _in_capsule(p, a, b, r):
    ab = np.asarray(b) - np.asarray(a)
    t = np.clip(((p - a) @ ab) / (ab @ ab), 0, 1)
    return np.linalg.norm(p - (a + t[:, None] * ab), axis=1) < r


def _regions(p):
    tet = [(8.0, 8.0, 8.0), (32.0, 8.0, 8.0), (20.0, 32.0, 8.0), (20.0, 16.0, 32.0)]
    return np.stack([_in_tet(p, tet),
                     _in_capsule(p, np.array([40.0, 40, 15]), np.array([40.0, 40, 45]), 6.0),
                     np.linalg.norm(p - np.array([40.0, 40, 30]), axis=1) < 3.0], axis=1)


def _by_serial(osim, name):
    out = {}
    for ll in range(osim.nlist()):
        L = osim.get_list(ll)
        for j in range(len(L["ident"])):
            if L["ident"][j] == osim.index(name):
                out[int(L["serno"][j])] = L["pos"][j].copy()
    return out


def test_closed_shapes_keep_reflected_molecules():
    spec = ss.c3_shapes_spec(n=6000, seed=37)
    spec["reactions"] = []
    spec.pop("placement")
    spec["dt"] = 0.05  # rms step 0.45 per axis: many collisions per molecule
    osim, _ = ss.build_oracle(spec)
    b0 = _by_serial(osim, "B")
    a0 = _by_serial(osim, "A")
    keys = sorted(b0)
    r0 = _regions(np.array([b0[k] for k in keys]))
    assert r0[:, 0].sum() > 20 and r0[:, 1].sum() > 50 and r0[:, 2].sum() > 3
    for _ in range(400):
        assert osim.step() == 0
    b1 = _by_serial(osim, "B")
    assert sorted(b1) == keys
    r1 = _regions(np.array([b1[k] for k in keys]))
    assert np.array_equal(r0, r1)
    # A crosses the tetrahedron (turning into D) and the capsule
    a1 = _by_serial(osim, "A")
    ka = sorted(set(a0) & set(a1))
    ra0 = _regions(np.array([a0[k] for k in ka]))
    ra1 = _regions(np.array([a1[k] for k in ka]))
    assert (ra0[:, 1] != ra1[:, 1]).sum() > 5
    assert osim.count(osim.index("D")) > 5
