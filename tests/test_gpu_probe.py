"""The random-line probe that prices the query kernels: it runs, reports a
positive rate, and rejects invalid arguments."""
import pytest

pytestmark = pytest.mark.gpu


def test_probe_random_lines(gpu):
    from ecl_wavelet_amd import probe_random_lines
    from ecl_wavelet_amd._lib import EclError
    for lanes in (1, 2, 4, 8):
        r = probe_random_lines(1 << 26, 100_000, steps=4, lanes=lanes)
        assert r > 0
    with pytest.raises(EclError):
        probe_random_lines(1 << 20, 1000, lanes=3)
    with pytest.raises(EclError):
        probe_random_lines(64, 1000)
