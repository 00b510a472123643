"""The blocked FP64 Gauss-Jordan inverse of the V-cycle's coarse level (mpsr_spd_inverse), checked
against numpy's inverse on SPD matrices whose sizes cover one block, whole blocks and partial last
blocks, including the coarse operator's own size class.  Also: a matrix that is not positive
definite is refused."""
import numpy as np
import pytest

import simulation_project_amd as mp

pytestmark = pytest.mark.gpu


def _spd(n, seed):
    rng = np.random.default_rng(seed)
    if n <= 300:
        m = rng.standard_normal((n, n))
        return m @ m.T + n * np.eye(n)
    # a 3-D 7-point Laplacian plus a shift, like the coarse operator's stencil, in a random order
    k = int(round(n ** (1 / 3))) + 1
    idx = np.arange(k ** 3).reshape(k, k, k)[:, :, :].ravel()[:n]
    a = np.zeros((n, n))
    pos = {int(v): i for i, v in enumerate(idx)}
    for i, v in enumerate(idx):
        x, y, z = np.unravel_index(v, (k, k, k))
        a[i, i] = 6.0 + 0.05
        for d in ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)):
            w = (x + d[0], y + d[1], z + d[2])
            if min(w) < 0 or max(w) >= k:
                continue
            j = pos.get(int(np.ravel_multi_index(w, (k, k, k))))
            if j is not None:
                a[i, j] = -1.0
    p = rng.permutation(n)
    return a[np.ix_(p, p)]


@pytest.mark.parametrize("n", [1, 5, 63, 64, 65, 130, 257, 1000, 2600])
def test_spd_inverse_matches_numpy(n):
    a = _spd(n, 7 + n)
    inv = mp.spd_inverse(a)
    ref = np.linalg.inv(a)
    err = np.abs(inv - ref).max() / np.abs(ref).max()
    cond = np.linalg.cond(a)
    assert err < 1e-13 * cond, (n, err, cond)
    # and it inverts: A X = I to the same order
    assert np.abs(a @ inv - np.eye(n)).max() < 1e-12 * cond


def test_spd_inverse_refuses_indefinite():
    a = _spd(70, 3)
    a[40, 40] = -1.0
    with pytest.raises(mp.MpsrError):
        mp.spd_inverse(a)
