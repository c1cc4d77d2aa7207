"""Python view of the library, mirroring the reference's class and method
names (ecl::BitArray, ecl::RankSelect, ecl::WaveletTree<T>; see
include/ecl_wavelet/ for the C++ drop-in).  Every call goes through the
C ABI of libecl_wavelet_hip.so.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import call

_DTYPES = {1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.uint64}

# RankSelectQuery<T> as 16 bytes: index, then the symbol in the low bytes
QUERY_DTYPE = np.dtype([("index", "<u8"), ("symbol", "<u8")])
MIXED_DTYPE = np.dtype([("arg", "<u8"), ("symbol", "<u4"), ("kind", "<u4")])
FOREST_QUERY_DTYPE = np.dtype([("arg", "<u8"), ("symbol", "<u4"), ("kind", "<u2"),
                               ("reserved", "<u2"), ("tree", "<u4"), ("reserved2", "<u4")])


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


class DeviceBuffer:
    """Raw device allocation owned by Python (inputs and outputs of the
    device-resident entry points)."""

    def __init__(self, nbytes, device=0):
        self.device = device
        self.nbytes = int(nbytes)
        ptr = C.c_void_p()
        call("ecl_device_alloc", device, max(self.nbytes, 1), C.byref(ptr))
        self.ptr = ptr

    @classmethod
    def from_array(cls, arr, device=0):
        arr = np.ascontiguousarray(arr)
        b = cls(arr.nbytes, device)
        b.upload(arr)
        return b

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes
        call("ecl_memcpy_h2d", self.device, self.ptr, _p(arr), arr.nbytes)

    def download(self, dtype, count=None):
        dtype = np.dtype(dtype)
        if count is None:
            count = self.nbytes // dtype.itemsize
        out = np.empty(count, dtype=dtype)
        assert out.nbytes <= self.nbytes
        call("ecl_memcpy_d2h", self.device, _p(out), self.ptr, out.nbytes)
        return out

    def view(self, offset, nbytes):
        """A non-owning buffer over bytes [offset, offset + nbytes)."""
        assert 0 <= offset and offset + nbytes <= self.nbytes
        v = DeviceBuffer.__new__(DeviceBuffer)
        v.device = self.device
        v.nbytes = int(nbytes)
        v.ptr = C.c_void_p(self.ptr.value + int(offset))
        v._owner = self
        return v

    def free(self):
        if getattr(self, "_owner", None) is not None:
            self.ptr = C.c_void_p()
            return
        if self.ptr:
            call("ecl_device_free", self.device, self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:  # noqa: BLE001 - interpreter shutdown
            pass


def synchronize(device=0):
    call("ecl_device_synchronize", device)


def zipf_thresholds(sigma, s=1.1):
    """u64 inverse-CDF thresholds of Zipf(s) over sigma symbols for
    generate_text (the last threshold is 2^64 - 1)."""
    w = 1.0 / np.power(np.arange(1, sigma + 1, dtype=np.float64), s)
    cdf = np.cumsum(w) / w.sum()
    thr = np.minimum(np.floor(cdf * 2.0 ** 64), 2.0 ** 64 - 1)
    out = np.array([int(x) if x < 2.0 ** 64 else 2 ** 64 - 1 for x in thr], dtype=np.uint64)
    out[-1] = np.uint64(2 ** 64 - 1)
    return out


def generate_text_into(buf, offset, n, sigma, dtype=np.uint8, seed=1, thresholds=None,
                       device=0):
    """The same draws written into an existing device buffer at a byte offset."""
    esize = np.dtype(dtype).itemsize
    assert offset + n * esize <= buf.nbytes
    thr = None if thresholds is None else np.ascontiguousarray(thresholds, np.uint64)
    call("ecl_generate_text", device, esize, seed, n, sigma,
         None if thr is None else _p(thr), C.c_void_p(buf.ptr.value + offset))


def generate_text(n, sigma, dtype=np.uint8, seed=1, thresholds=None, device=0):
    """Synthetic text on the device: hash-uniform or inverse-CDF draws."""
    esize = np.dtype(dtype).itemsize
    buf = DeviceBuffer(n * esize, device)
    thr = None if thresholds is None else np.ascontiguousarray(thresholds, np.uint64)
    call("ecl_generate_text", device, esize, seed, n, sigma,
         None if thr is None else _p(thr), buf.ptr)
    return buf


_ESIZE_DTYPE = {1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.uint64}


def read_data_from_file(path, num_symbols, dtype=np.uint8):
    """readDataFromFile<T> (utils/test_benchmark_utils.cuh:588-619): the first
    num_symbols elements of a raw text file, with the reference's checks."""
    dtype = np.dtype(dtype)
    out = np.empty(int(num_symbols), dtype=dtype)
    call("ecl_read_text_file", str(path).encode(), dtype.itemsize, out.size, _p(out))
    return out


def convert_data_to_min_alphabet(data, alphabet=None, return_alphabet=False):
    """convertDataToMinAlphabet<T> (utils/test_benchmark_utils.cuh:628-679):
    replaces every element of the contiguous numpy array `data` in place by
    its index in the sorted distinct elements (or, by lower bound, in the
    sorted `alphabet`); returns the alphabet size (and the alphabet)."""
    assert isinstance(data, np.ndarray) and data.flags.c_contiguous
    dt = data.dtype
    a = None if alphabet is None else np.ascontiguousarray(alphabet, dtype=dt)
    cap = (1 << (8 * dt.itemsize)) if dt.itemsize <= 2 else max(data.size, 1)
    out = np.empty(cap if a is None else len(a), dtype=dt)
    sigma = C.c_size_t()
    call("ecl_convert_to_min_alphabet", dt.itemsize, _p(data), data.size,
         None if a is None else _p(a), 0 if a is None else len(a), _p(out), C.byref(sigma))
    if return_alphabet:
        return sigma.value, out[:sigma.value].copy()
    return sigma.value


class BitArray:
    """ecl::BitArray (bitarray/bit_array.cuh)."""

    def __init__(self, sizes, init_value=None, device=0):
        sizes = np.ascontiguousarray(sizes, dtype=np.uint64)
        h = C.c_void_p()
        init = -1 if init_value is None else int(bool(init_value))
        call("ecl_bit_array_create", _p(sizes), len(sizes), init, device, C.byref(h))
        self._h = h
        self._owned = True
        self.device = device

    @classmethod
    def _borrow(cls, handle, device):
        self = cls.__new__(cls)
        self._h = handle
        self._owned = False
        self.device = device
        return self

    def __del__(self):
        if getattr(self, "_owned", False) and self._h:
            try:
                call("ecl_bit_array_destroy", self._h)
            except Exception:  # noqa: BLE001
                pass
            self._h = None

    def numArrays(self):
        n = C.c_size_t()
        call("ecl_bit_array_num_arrays", self._h, C.byref(n))
        return n.value

    def size(self, i):
        n = C.c_size_t()
        call("ecl_bit_array_size", self._h, i, C.byref(n))
        return n.value

    def getOffset(self, i):
        n = C.c_size_t()
        call("ecl_bit_array_offset", self._h, i, C.byref(n))
        return n.value

    def sizeInWords(self, i):
        return (self.size(i) + 31) // 32

    def writeWords(self, i, first_word, words):
        words = np.ascontiguousarray(words, dtype=np.uint32)
        call("ecl_bit_array_write_words", self._h, i, first_word, _p(words), len(words))

    def words(self, i, first_word=0, count=None):
        if count is None:
            count = self.sizeInWords(i) - first_word
        out = np.empty(count, dtype=np.uint32)
        call("ecl_bit_array_read_words", self._h, i, first_word, _p(out), count)
        return out

    def bits(self, i):
        w = self.words(i)
        return np.unpackbits(w.view(np.uint8), bitorder="little")[: self.size(i)]

    def fillRandom(self, i, seed, threshold):
        call("ecl_bit_array_fill_random", self._h, i, seed, threshold)


class RankSelect:
    """ecl::RankSelect (bitarray/rank_select.cuh).  Consumes the BitArray."""

    def __init__(self, bit_array, device=0):
        h = C.c_void_p()
        ba = C.c_void_p(bit_array._h.value if isinstance(bit_array._h, C.c_void_p)
                        else bit_array._h)
        call("ecl_rank_select_create", C.byref(ba), device, C.byref(h))
        bit_array._owned = False  # moved into the RankSelect
        bit_array._h = None
        self._h = h
        self._owned = True
        self.device = device

    @classmethod
    def _borrow(cls, handle, device):
        self = cls.__new__(cls)
        self._h = handle
        self._owned = False
        self.device = device
        return self

    def __del__(self):
        if getattr(self, "_owned", False) and self._h:
            try:
                call("ecl_rank_select_destroy", self._h)
            except Exception:  # noqa: BLE001
                pass
            self._h = None

    @property
    def bit_array_(self):
        h = C.c_void_p()
        call("ecl_rank_select_bit_array", self._h, C.byref(h))
        return BitArray._borrow(h, self.device)

    def info(self, i):
        inf = _lib.RankSelectInfo()
        call("ecl_rank_select_info_get", self._h, i, C.byref(inf))
        return inf

    def getNumL1Blocks(self, i):
        return self.info(i).num_l1_blocks

    def getNumL2Blocks(self, i):
        return self.info(i).num_l2_blocks

    def getNumLastL2Blocks(self, i):
        return self.info(i).num_last_l2_blocks

    def getTotalNumVals(self, i, value):
        inf = self.info(i)
        return inf.total_ones if value else inf.total_zeros

    def l1Entries(self, i):
        n = self.getNumL1Blocks(i)
        out = np.empty(n, dtype=np.uint64)
        call("ecl_rank_select_get_l1", self._h, i, 0, n, _p(out))
        return out

    def l2Entries(self, i):
        n = self.getNumL2Blocks(i)
        out = np.empty(n, dtype=np.uint16)
        call("ecl_rank_select_get_l2", self._h, i, 0, n, _p(out))
        return out

    def selectSamples(self, i, value):
        inf = self.info(i)
        n = inf.num_samples_1 if value else inf.num_samples_0
        out = np.empty(n, dtype=np.uint64)
        call("ecl_rank_select_get_samples", self._h, i, int(value), 0, n, _p(out))
        return out

    def rank(self, i, index, value=1):
        index = np.ascontiguousarray(index, dtype=np.uint64)
        out = np.empty(len(index), dtype=np.uint64)
        call("ecl_rank_select_rank", self._h, i, int(value), _p(index), len(index), _p(out))
        return out

    def select(self, i, k, value=1):
        k = np.ascontiguousarray(k, dtype=np.uint64)
        out = np.empty(len(k), dtype=np.uint64)
        call("ecl_rank_select_select", self._h, i, int(value), _p(k), len(k), _p(out))
        return out

    def buildQueryLines(self):
        """Second copy of the bits in 128-byte lines headed by their rank:
        a rank then reads one line."""
        call("ecl_rank_select_build_query_lines", self._h)

    def buildSelectLines(self):
        """A 128-byte line every r ones (zeros): its position and the 960 bits
        from there; a select then reads one line."""
        call("ecl_rank_select_build_select_lines", self._h)

    def rankDevice(self, i, d_index, m, d_out, value=1, stream=None):
        call("ecl_rank_select_rank_device", self._h, i, int(value), d_index.ptr, m,
             d_out.ptr, stream)

    def selectDevice(self, i, d_k, m, d_out, value=1, stream=None):
        call("ecl_rank_select_select_device", self._h, i, int(value), d_k.ptr, m,
             d_out.ptr, stream)

    def rankGenerated(self, i, seed, m, d_out, value=1, stream=None):
        call("ecl_rank_select_rank_generated", self._h, i, int(value), seed, m,
             d_out.ptr, stream)

    def selectGenerated(self, i, seed, m, d_out, value=1, stream=None):
        call("ecl_rank_select_select_generated", self._h, i, int(value), seed, m,
             d_out.ptr, stream)


class WaveletTree:
    """ecl::WaveletTree<T> (tree/wavelet_tree.cuh).

    ``data`` is a numpy array of uint8/16/32/64 (host text) or a
    DeviceBuffer together with ``dtype`` and ``n`` (device-resident text).
    ``alphabet`` must be sorted; None or empty discovers it.
    """

    def __init__(self, data, alphabet=None, GPU_index=0, dtype=None, n=None):
        self.device = GPU_index
        h = C.c_void_p()
        if isinstance(data, DeviceBuffer):
            self.dtype = np.dtype(dtype)
            esize = self.dtype.itemsize
            a = (np.zeros(0, self.dtype) if alphabet is None
                 else np.ascontiguousarray(alphabet, dtype=self.dtype))
            call("ecl_wavelet_tree_create_device", esize, data.ptr, n, _p(a), len(a),
                 GPU_index, C.byref(h))
        else:
            data = np.ascontiguousarray(data)
            self.dtype = data.dtype
            esize = data.dtype.itemsize
            a = (np.zeros(0, data.dtype) if alphabet is None
                 else np.ascontiguousarray(alphabet, dtype=data.dtype))
            call("ecl_wavelet_tree_create", esize, _p(data), len(data), _p(a), len(a),
                 GPU_index, C.byref(h))
        self._h = h

    def __del__(self):
        if getattr(self, "_h", None):
            try:
                call("ecl_wavelet_tree_destroy", self._h)
            except Exception:  # noqa: BLE001
                pass
            self._h = None

    def buildQueryLines(self):
        """Second copy of the level bits in 128-byte lines headed by their
        rank: rank and access then read one line per level."""
        call("ecl_wavelet_tree_build_query_lines", self._h)

    def buildSelectLines(self):
        """Per level, a 128-byte line every r ones (zeros) holding its
        position and the 960 bits from there: a select then reads one line
        per level."""
        call("ecl_wavelet_tree_build_select_lines", self._h)

    # ---- persistence
    def save(self, path):
        """Writes the tree (alphabet, symbol counts, level bits) to a file."""
        call("ecl_wavelet_tree_save", self._h, str(path).encode())

    @classmethod
    def fromFile(cls, path, num_symbols, dtype=np.uint8, alphabet=None, GPU_index=0):
        """A tree over the first num_symbols elements of a raw text file
        (ecl_wavelet_tree_create_from_file; no alphabet discovers it)."""
        self = cls.__new__(cls)
        self.device = GPU_index
        self.dtype = np.dtype(dtype)
        a = (np.zeros(0, self.dtype) if alphabet is None
             else np.ascontiguousarray(alphabet, dtype=self.dtype))
        h = C.c_void_p()
        call("ecl_wavelet_tree_create_from_file", self.dtype.itemsize, str(path).encode(),
             int(num_symbols), _p(a), len(a), GPU_index, C.byref(h))
        self._h = h
        return self

    @classmethod
    def load(cls, path, GPU_index=0):
        """A tree read from a file written by save(); its directory is
        rebuilt on the device."""
        self = cls.__new__(cls)
        self.device = GPU_index
        h = C.c_void_p()
        call("ecl_wavelet_tree_load", str(path).encode(), GPU_index, C.byref(h))
        self._h = h
        self.dtype = np.dtype({1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.uint64}[
            int(self.info().element_size)])
        return self

    # ---- accessors (wavelet_tree.cuh:2165-2282)
    def info(self):
        inf = _lib.TreeInfo()
        call("ecl_wavelet_tree_info_get", self._h, C.byref(inf))
        return inf

    def getAlphabetSize(self):
        return self.info().alphabet_size

    def getNumLevels(self):
        return self.info().num_levels

    def isMinAlphabet(self):
        return bool(self.info().is_min_alphabet)

    def getCodesStart(self):
        return self.info().codes_start

    def getAlphabet(self):
        out = np.empty(self.getAlphabetSize(), dtype=self.dtype)
        call("ecl_wavelet_tree_alphabet", self._h, _p(out))
        return out

    def getCounts(self):
        out = np.empty(self.getAlphabetSize(), dtype=np.uint64)
        call("ecl_wavelet_tree_counts", self._h, _p(out))
        return out

    def getTotalAppearances(self, i):
        c = self.getCounts()
        n = self.info().data_size
        return int((n if i == len(c) - 1 else c[i + 1]) - c[i])

    def codes(self):
        inf = self.info()
        k = inf.alphabet_size - inf.codes_start if not inf.is_pow_two else 0
        codes = np.empty(k, dtype=np.uint64)
        lens = np.empty(k, dtype=np.uint8)
        call("ecl_wavelet_tree_codes", self._h, _p(codes), _p(lens))
        return codes, lens

    def nodeRanks(self):
        k = self.info().num_ranks
        r = np.empty(k, dtype=np.uint64)
        s = np.empty(k, dtype=np.uint64)
        lv = np.empty(k, dtype=np.uint8)
        call("ecl_wavelet_tree_node_ranks", self._h, _p(r), _p(s), _p(lv))
        return r, s, lv

    def numNodesAtLevel(self):
        L = self.getNumLevels()
        out = np.empty(max(L - 1, 0), dtype=np.uint64)
        call("ecl_wavelet_tree_nodes_at_level", self._h, _p(out))
        return out

    @property
    def rank_select_(self):
        h = C.c_void_p()
        call("ecl_wavelet_tree_rank_select", self._h, C.byref(h))
        return RankSelect._borrow(h, self.device)

    # ---- host batch queries
    def access(self, indices, flags=0):
        idx = np.ascontiguousarray(indices, dtype=np.uint64)
        out = np.empty(len(idx), dtype=self.dtype)
        call("ecl_wavelet_tree_access", self._h, _p(idx), len(idx), _p(out), flags)
        return out

    @staticmethod
    def make_queries(indices, symbols):
        q = np.zeros(len(indices), dtype=QUERY_DTYPE)
        q["index"] = np.asarray(indices, dtype=np.uint64)
        q["symbol"] = np.asarray(symbols, dtype=np.uint64)
        return q

    def rank(self, queries, flags=0):
        q = np.ascontiguousarray(queries, dtype=QUERY_DTYPE)
        out = np.empty(len(q), dtype=np.uint64)
        call("ecl_wavelet_tree_rank", self._h, _p(q), len(q), _p(out), flags)
        return out

    def select(self, queries, flags=0):
        q = np.ascontiguousarray(queries, dtype=QUERY_DTYPE)
        out = np.empty(len(q), dtype=np.uint64)
        call("ecl_wavelet_tree_select", self._h, _p(q), len(q), _p(out), flags)
        return out

    # ---- device-resident batch queries
    def accessDevice(self, d_idx, m, d_out, flags=0, stream=None):
        call("ecl_wavelet_tree_access_device", self._h, d_idx.ptr, m, d_out.ptr, flags, stream)

    def rankDevice(self, d_q, m, d_out, flags=0, stream=None):
        call("ecl_wavelet_tree_rank_device", self._h, d_q.ptr, m, d_out.ptr, flags, stream)

    def selectDevice(self, d_q, m, d_out, flags=0, stream=None):
        call("ecl_wavelet_tree_select_device", self._h, d_q.ptr, m, d_out.ptr, flags, stream)

    def queryBatchDevice(self, d_q, m, d_out, stream=None):
        call("ecl_wavelet_tree_query_batch_device", self._h, d_q.ptr, m, d_out.ptr, stream)

    def generateQueries(self, kind, seed, m, d_out, stream=None):
        call("ecl_wavelet_tree_generate_queries", self._h, kind, seed, m, d_out.ptr, stream)

    def buildQuadLines(self):
        """Quad lines (power-of-two alphabets with an even number of levels):
        access and rank read one memory line per two levels; results do not
        change."""
        call("ecl_wavelet_tree_build_quad_lines", self._h)

    def buildHexLines(self):
        """Hex lines (power-of-two alphabets whose level count is a multiple
        of 4; builds the quad lines first when absent): access and rank read
        one memory line per four levels; results do not change."""
        call("ecl_wavelet_tree_build_hex_lines", self._h)

    def buildQuadSelectLines(self):
        """Quad select lines (builds the quad lines first when absent): select
        reads one memory line per two levels; results do not change."""
        call("ecl_wavelet_tree_build_quad_select_lines", self._h)

    def buildHostBytes(self):
        """Bytes of the build's buffers that were placed in pinned host
        memory because the device could not hold them (0: none)."""
        b = C.c_uint64()
        call("ecl_wavelet_tree_build_host_bytes", self._h, C.byref(b))
        return b.value

    def setQueryLanes(self, lanes):
        """Lanes per access/rank query for this tree (0: library-wide)."""
        call("ecl_wavelet_tree_set_query_lanes", self._h, lanes)

    def autotune(self, m=1 << 22, seed=1):
        """Times one and two lanes per access/rank query on m generated
        queries per kind and keeps the faster; returns (lanes, [ms1, ms2])."""
        lanes = C.c_uint32()
        ms = (C.c_float * 2)()
        call("ecl_wavelet_tree_autotune", self._h, m, seed, C.byref(lanes), ms)
        return lanes.value, [ms[0], ms[1]]


class Forest:
    """Many trees over uint8 texts built together (ecl_forest_create): text
    i is d_texts[starts[i] : starts[i] + lengths[i]] (starts multiples of 16)
    over the alphabet 0 .. sigmas[i]-1."""

    def __init__(self, d_texts, starts, lengths, sigmas, device=0, tile_rows=0):
        starts = np.ascontiguousarray(starts, dtype=np.uint64)
        lengths = np.ascontiguousarray(lengths, dtype=np.uint64)
        sigmas = np.ascontiguousarray(sigmas, dtype=np.uint64)
        assert starts.size == lengths.size == sigmas.size
        h = C.c_void_p()
        call("ecl_forest_create", d_texts.ptr, _p(starts), _p(lengths), _p(sigmas), sigmas.size,
             device, tile_rows, C.byref(h))
        self._h = h
        self.device = device
        self.num_trees = int(sigmas.size)

    def __del__(self):
        try:
            if self._h:
                call("ecl_forest_destroy", self._h)
                self._h = C.c_void_p()
        except Exception:  # noqa: BLE001 - interpreter shutdown
            pass

    def buildQueryLines(self):
        """Query lines and select hints for every tree (one 128-byte line per
        level for rank and access, a hint and one block for select); later
        queries and graphs read them, with the same answers."""
        call("ecl_forest_build_query_lines", self._h)

    def buildSelectLines(self):
        """Select lines for every tree (a select reads one 128-byte line per
        level), in one launch over all levels; later queries and graphs read
        them, with the same answers.  A build graph's replay drops them from
        the trees it rebuilds."""
        call("ecl_forest_build_select_lines", self._h)

    def buildQuadLayouts(self):
        """Quad lines and quad select lines for every tree over a power-of-two
        alphabet with an even number of levels (and hex lines for those over
        16 or 256 symbols); later batches (and graphs
        captured afterwards) run a pair of lanes per query over them, and
        over the binary levels for the other trees, with the same answers.
        A build graph's replay drops them from the trees it rebuilds."""
        call("ecl_forest_build_quad_layouts", self._h)

    def treeInfo(self, i):
        n, s = C.c_uint64(), C.c_uint64()
        L = C.c_uint32()
        call("ecl_forest_tree_info", self._h, i, C.byref(n), C.byref(s), C.byref(L))
        return n.value, s.value, L.value

    def queryDevice(self, d_q, m, d_out, stream=None):
        call("ecl_forest_query_device", self._h, d_q.ptr, m, d_out.ptr, stream)

    def generateQueries(self, qoff, seed, d_out, stream=None):
        qoff = np.ascontiguousarray(qoff, dtype=np.uint64)
        call("ecl_forest_generate_queries", self._h, _p(qoff), seed, d_out.ptr, stream)

    def autotune(self, d_q, m, d_out):
        b, g = C.c_uint32(), C.c_uint32()
        call("ecl_forest_autotune", self._h, d_q.ptr, m, d_out.ptr, C.byref(b), C.byref(g))
        return b.value, g.value

    def graph(self, d_q, m, d_out):
        return Graph(self, d_q, m, d_out)

    def buildGraph(self):
        """The construction as one HIP graph: launch() rebuilds every tree
        from the current contents of the texts (ecl_forest_build_graph)."""
        return Graph(self, build=True)

    def autotuneBuild(self):
        """Times the build with each tile size, keeps the fastest for later
        build graphs; returns (tile_rows, ms)."""
        r, t = C.c_uint32(), C.c_float()
        call("ecl_forest_autotune_build", self._h, C.byref(r), C.byref(t))
        return r.value, t.value


class Graph:
    """A forest batch launch (ecl_forest_graph_create), or with build=True the
    forest's construction (ecl_forest_build_graph), captured into a HIP graph."""

    def __init__(self, forest, d_q=None, m=0, d_out=None, build=False):
        h = C.c_void_p()
        if build:
            call("ecl_forest_build_graph", forest._h, C.byref(h))
        else:
            call("ecl_forest_graph_create", forest._h, d_q.ptr, m, d_out.ptr, C.byref(h))
        self._h = h
        self._keep = (forest, d_q, d_out)

    def launch(self, stream=None):
        call("ecl_graph_launch", self._h, stream)

    def __del__(self):
        try:
            if self._h:
                call("ecl_graph_destroy", self._h)
                self._h = C.c_void_p()
        except Exception:  # noqa: BLE001 - interpreter shutdown
            pass


class _ServerConfig(C.Structure):
    _fields_ = [("hbm_budget", C.c_uint64), ("max_batch", C.c_uint64),
                ("max_inflight", C.c_uint32), ("max_trees", C.c_uint32),
                ("coalesce", C.c_uint32), ("query_lines", C.c_uint32)]


class _ServerStats(C.Structure):
    _fields_ = [("requests", C.c_uint64), ("queries", C.c_uint64), ("batches", C.c_uint64),
                ("p50_ms", C.c_double), ("p99_ms", C.c_double), ("mean_ms", C.c_double),
                ("max_ms", C.c_double), ("elapsed_s", C.c_double),
                ("queries_per_s", C.c_double), ("hbm_used", C.c_uint64),
                ("trees", C.c_uint64)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class _LoadSpec(C.Structure):
    _fields_ = [("rate", C.c_double), ("seconds", C.c_double), ("seed", C.c_uint64),
                ("min_size", C.c_uint64), ("max_size", C.c_uint64), ("pool", C.c_uint64)]


class Server:
    """Query server with continuous batching (ecl_server_*): resident trees
    admitted by an HBM budget, requests coalesced into device batches.
    query_lines: 0/False none, 1/True query lines and select hints for every
    admitted tree, 2 also select lines, 3 the quad layouts (quad lines and
    quad select lines, hex lines where the level count is a multiple of 4)
    for trees that take them and query and select lines for the others."""

    def __init__(self, device=0, hbm_budget=0, max_batch=0, max_inflight=0, max_trees=0,
                 coalesce=True, query_lines=0):
        cfg = _ServerConfig(hbm_budget, max_batch, max_inflight, max_trees,
                            1 if coalesce else 0, int(query_lines))
        h = C.c_void_p()
        call("ecl_server_create", device, C.byref(cfg), C.byref(h))
        self._h = h
        self._keep = {}

    def close(self):
        if self._h:
            call("ecl_server_destroy", self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:  # noqa: BLE001 - interpreter shutdown
            pass

    def addTree(self, text, sigma, n=None):
        """text: numpy uint8 array (host) or DeviceBuffer (then pass n)."""
        t = C.c_uint32()
        if isinstance(text, DeviceBuffer):
            call("ecl_server_add_tree", self._h, text.ptr, 1, n, sigma, C.byref(t))
        else:
            text = np.ascontiguousarray(text, dtype=np.uint8)
            call("ecl_server_add_tree", self._h, _p(text), 0, text.size, sigma, C.byref(t))
        return t.value

    def addTreeFile(self, path):
        """Admits a uint8 tree over 0..sigma-1 saved by WaveletTree.save,
        without rebuilding it."""
        t = C.c_uint32()
        call("ecl_server_add_tree_file", self._h, str(path).encode(), C.byref(t))
        return t.value

    def addTreeTextFile(self, path, num_symbols):
        """Admits a tree over the first num_symbols bytes of a raw uint8 text
        file, converted to its minimal alphabet; returns (tree, alphabet):
        requests name symbols by their index in that alphabet."""
        t, sg = C.c_uint32(), C.c_uint64()
        alpha = np.zeros(256, np.uint8)
        call("ecl_server_add_tree_text_file", self._h, str(path).encode(), int(num_symbols),
             _p(alpha), C.byref(sg), C.byref(t))
        return t.value, alpha[:sg.value].copy()

    def removeTree(self, tree):
        call("ecl_server_remove_tree", self._h, tree)

    def submit(self, tree, queries):
        """queries: MIXED_DTYPE array; returns a request id (the arrays are
        kept alive until wait())."""
        q = np.ascontiguousarray(queries, dtype=MIXED_DTYPE)
        out = np.zeros(q.size, dtype=np.uint64)
        rid = C.c_uint64()
        call("ecl_server_submit", self._h, tree, _p(q), q.size, _p(out), C.byref(rid))
        self._keep[rid.value] = (q, out)
        return rid.value

    def wait(self, rid):
        lat = C.c_double()
        call("ecl_server_wait", self._h, rid, C.byref(lat))
        _, out = self._keep.pop(rid)
        return out, lat.value

    def setMode(self, coalesce):
        call("ecl_server_set_mode", self._h, 1 if coalesce else 0)

    def stats(self):
        st = _ServerStats()
        call("ecl_server_stats_get", self._h, C.byref(st))
        return st.as_dict()

    def resetStats(self):
        call("ecl_server_stats_reset", self._h)

    def load(self, rate, seconds, seed=1, min_size=1, max_size=1_000_000, pool=0):
        spec = _LoadSpec(rate, seconds, seed, min_size, max_size, pool)
        st = _ServerStats()
        call("ecl_server_load", self._h, C.byref(spec), C.byref(st))
        return st.as_dict()


class Profiler:
    """Per-kernel HIP-event timing inside the library."""

    @staticmethod
    def enable(on=True):
        call("ecl_profile_enable", 1 if on else 0)

    @staticmethod
    def reset():
        call("ecl_profile_reset")

    @staticmethod
    def get(kernel):
        n = C.c_uint64()
        ms = C.c_double()
        call("ecl_profile_get", kernel.encode(), C.byref(n), C.byref(ms))
        return n.value, ms.value


class Timer:
    """HIP events on a stream (default stream when None)."""

    def __init__(self, stream=None):
        self.stream = stream
        self.a = C.c_void_p()
        self.b = C.c_void_p()
        call("ecl_event_create", C.byref(self.a))
        call("ecl_event_create", C.byref(self.b))

    def start(self):
        call("ecl_event_record", self.a, self.stream)

    def stop(self):
        call("ecl_event_record", self.b, self.stream)
        ms = C.c_float()
        call("ecl_event_elapsed", self.a, self.b, C.byref(ms))
        return ms.value

    def __del__(self):
        try:
            call("ecl_event_destroy", self.a)
            call("ecl_event_destroy", self.b)
        except Exception:  # noqa: BLE001
            pass


def device_name(device=0):
    buf = C.create_string_buffer(256)
    call("ecl_device_name", device, buf, 256)
    return buf.value.decode()


def set_build_span(tiles_per_workgroup=0):
    """Construction tiles each workgroup takes in turn (0: automatic); a
    tuning knob that does not change results."""
    call("ecl_set_build_span", tiles_per_workgroup)


def set_query_lanes(lanes=0):
    """Lanes per access/rank query in the tree query kernels (0: automatic,
    two over query lines and one otherwise; 2: each lane of a pair fetches
    half of every level's bits; 1: one lane per query); a tuning knob that
    does not change results."""
    call("ecl_set_query_lanes", lanes)


def probe_random_lines(nbytes, queries, steps=8, lanes=2, device=0):
    """Random, dependent 128-byte line reads per second on the device (the
    bound of the tree queries: one line per level, two for select)."""
    r = C.c_double()
    call("ecl_probe_random_lines", device, nbytes, queries, steps, lanes, C.byref(r))
    return r.value


def release_cached_memory(device=0):
    """Frees the construction scratch the library keeps per device between
    builds."""
    call("ecl_release_cached_memory", device)


def set_device_memory_cap(device, nbytes):
    """Caps the free device memory the library's admission checks assume
    (0 removes the cap)."""
    call("ecl_set_device_memory_cap", device, int(nbytes))


def tree_build_memory(dtype, n, sigma, host_text=True):
    """(resident, staging) upper bounds in bytes for a tree build: resident
    must fit the device; each staging buffer goes to pinned host memory when
    the device cannot hold it after the resident part."""
    r = C.c_uint64()
    g = C.c_uint64()
    call("ecl_wavelet_tree_build_memory", np.dtype(dtype).itemsize, n, sigma,
         1 if host_text else 0, C.byref(r), C.byref(g))
    return r.value, g.value


def mem_info(device=0):
    fr = C.c_size_t()
    tot = C.c_size_t()
    call("ecl_device_mem_info", device, C.byref(fr), C.byref(tot))
    return fr.value, tot.value
