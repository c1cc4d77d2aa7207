"""The multiply-high that replaces i / r in the select-line lookup
(LevelView::sel_magic = floor((2^64 - 1) / r) + 1, r > 1) equals the true
floor division for every rate the library chooses (1 < r <= 800, tested to
1024) over the index range of any array (i < 2^44), at and around every
multiple of r and at random points."""
import numpy as np


def _magic(r):
    return ((1 << 64) - 1) // r + 1


def test_magic_division_exact():
    rng = np.random.default_rng(5)
    for r in range(2, 1025):
        m = _magic(r)
        xs = [0, 1, r - 1, r, r + 1, (1 << 44) - 1]
        xs += [int(x) for x in rng.integers(0, 1 << 44, 200)]
        k = rng.integers(1, (1 << 44) // r, 50)
        xs += [int(q) * r + d for q in k for d in (-1, 0, 1)]
        for i in xs:
            assert (i * m) >> 64 == i // r, (r, i)
