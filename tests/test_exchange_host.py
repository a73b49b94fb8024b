"""tpls_exchange2d on host arrays (no array attached to the mirror, so no
device is touched): mpi.f90's exchange2d (:119-170), exchange2d_aug1
(:215-270) and exchange2d_aug2 (:294-369) for one rank whose y neighbours
are itself and whose x neighbours are MPI_PROC_NULL.

The expected values follow from the MPI datatypes: an xz-plane type is one
block of ex-sx+3 values starting at x = sx-1 per z-plane, so the received
rows change only over x = sx-1..ex+1, and the yz-plane exchanges in x go to
MPI_PROC_NULL and change nothing.
"""
import numpy as np
import pytest

from tpls_amd import _lib


@pytest.fixture(scope="module", autouse=True)
def _loaded():
    try:
        _lib.lib()
    except OSError as e:
        pytest.skip(f"libtpls_hip.so not loadable: {e}")


def field(rng, lx, ux, ly, uy, lz, uz):
    return np.asfortranarray(rng.standard_normal((ux - lx + 1, uy - ly + 1, uz - lz + 1)))


@pytest.mark.parametrize("ex,ey,ez", [(9, 6, 4), (16, 8, 0), (5, 4, 7)])
def test_exchange2d_rows(ex, ey, ez):
    rng = np.random.default_rng(ex * 100 + ey)
    sx = sy = 1
    sz = 0
    a = field(rng, sx - 1, ex + 1, sy - 1, ey + 1, sz, ez)
    want = a.copy()
    # arrays index from (sx-1, sy-1): row y sits at y - (sy-1)
    want[:, 0, :] = a[:, ey, :]          # sy-1 <- ey
    want[:, ey + 1, :] = a[:, sy, :]     # ey+1 <- sy
    _lib.exchange2d(a, ex, ey, ez, sx, sy, sz, 0)
    assert np.array_equal(a, want)


@pytest.mark.parametrize("kind", [1, 2])
def test_exchange2d_augmented_rows(kind):
    rng = np.random.default_rng(kind)
    ex, ey, ez, sx, sy, sz = 11, 8, 5, 1, 1, 0
    a = field(rng, sx - 2, ex + 2, sy - 2, ey + 2, sz, ez)
    want = a.copy()
    o = sy - 2                                   # row y sits at y - o
    xs = slice((sx - 1) - (sx - 2), (ex + 1) - (sx - 2) + 1)   # x = sx-1..ex+1
    if kind == 1:
        want[xs, (sy - 1) - o, :] = a[xs, ey - o, :]
        want[xs, (ey + 1) - o, :] = a[xs, sy - o, :]
    else:
        want[xs, (sy - 2) - o, :] = a[xs, (ey - 1) - o, :]
        want[xs, (ey + 2) - o, :] = a[xs, (sy + 1) - o, :]
    _lib.exchange2d(a, ex, ey, ez, sx, sy, sz, kind)
    assert np.array_equal(a, want)
    # the outer x columns of the received rows keep their values
    assert np.array_equal(a[0], want[0]) and np.array_equal(a[-1], want[-1])


def test_exchange2d_rejects_other_decompositions():
    a = np.zeros((12, 10, 3), order="F")
    with pytest.raises(_lib.TplsError):
        _lib.exchange2d(a, 10, 8, 2, 2, 1, 0, 0)
    with pytest.raises(_lib.TplsError):
        _lib.exchange2d(a, 10, 8, 2, 1, 1, 0, 3)
