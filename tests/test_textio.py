"""Raw text files and the minimal alphabet (host functions of the C ABI):
readDataFromFile and convertDataToMinAlphabet semantics
(utils/test_benchmark_utils.cuh:588-679).  No GPU needed."""
import numpy as np
import pytest

from ecl_wavelet_amd import convert_data_to_min_alphabet, read_data_from_file
from ecl_wavelet_amd._lib import EclError


@pytest.mark.parametrize("dt", [np.uint8, np.uint16, np.uint32, np.uint64])
def test_read_prefix_and_errors(tmp_path, dt):
    rng = np.random.default_rng(np.dtype(dt).itemsize)
    data = rng.integers(0, np.iinfo(dt).max, 1000, dtype=np.uint64).astype(dt)
    p = tmp_path / "t.bin"
    data.tofile(p)
    assert np.array_equal(read_data_from_file(p, 1000, dt), data)
    assert np.array_equal(read_data_from_file(p, 17, dt), data[:17])
    assert read_data_from_file(p, 0, dt).size == 0
    with pytest.raises(EclError, match="^Data size is larger than file size$"):
        read_data_from_file(p, 1001, dt)
    with pytest.raises(EclError, match="^Failed to open file: " + str(tmp_path / "none")):
        read_data_from_file(tmp_path / "none", 1, dt)
    if np.dtype(dt).itemsize > 1:
        q = tmp_path / "odd.bin"
        q.write_bytes(data.tobytes() + b"x")
        with pytest.raises(EclError, match="^File size is not a multiple of the type size$"):
            read_data_from_file(q, 10, dt)


@pytest.mark.parametrize("dt", [np.uint8, np.uint16, np.uint32, np.uint64])
@pytest.mark.parametrize("n", [1, 1000, 3_000_001])
def test_convert_discovers_sorted_distinct(dt, n):
    """Every symbol becomes its index among the sorted distinct symbols."""
    rng = np.random.default_rng(n + np.dtype(dt).itemsize)
    top = min(np.iinfo(dt).max, 2 ** 40)
    pool = np.unique(rng.integers(0, top, 300, dtype=np.uint64).astype(dt))
    data = pool[rng.integers(0, len(pool), n)]
    exp_alpha = np.unique(data)
    exp = np.searchsorted(exp_alpha, data).astype(dt)
    work = data.copy()
    sigma, alpha = convert_data_to_min_alphabet(work, return_alphabet=True)
    assert sigma == len(exp_alpha)
    assert np.array_equal(alpha, exp_alpha)
    assert np.array_equal(work, exp)


@pytest.mark.parametrize("dt", [np.uint8, np.uint16, np.uint32, np.uint64])
def test_convert_with_alphabet_is_lower_bound(dt):
    """With a given sorted alphabet every symbol maps to the first entry not
    below it (std::lower_bound), also for symbols outside the alphabet."""
    rng = np.random.default_rng(3)
    alpha = np.unique(rng.integers(1, 200, 40)).astype(dt)
    data = rng.integers(0, 256, 5000).astype(dt)
    work = data.copy()
    assert convert_data_to_min_alphabet(work, alphabet=alpha) == len(alpha)
    assert np.array_equal(work, np.searchsorted(alpha, data, side="left").astype(dt))
