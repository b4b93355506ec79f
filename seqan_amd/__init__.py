"""seqan_amd -- MI355X (gfx950) batched affine-gap pairwise alignment.

Python access to the C ABI of ``libseqan_amd.so`` (declared in
``include/seqan_amd.h``).  The native library does all the work; this module
only marshals numpy arrays.  It mirrors the reference's batch interface:

    global_alignment_score(set_h, set_v, score, align_config=(top, left, right, bottom))
        <- globalAlignmentScore(ExecutionPolicy<Parallel,Vectorial>, setH, setV,
                                score, AlignConfig<T,L,R,B>())
           (include/seqan/align_parallel/parallel_align_interface.h:332-343)
    local_alignment_score(set_h, set_v, score)
        <- localAlignmentScore(...) (parallel_align_interface.h:354-365)

Importing this module never falls back to anything: if the shared library is
missing it raises, and every alignment call runs on the GPU.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libseqan_amd.so")

FREE_TOP, FREE_LEFT, FREE_RIGHT, FREE_BOTTOM, LOCAL = 1, 2, 4, 8, 16
RESIDUES_ON_DEVICE, SCORES_ON_DEVICE, FORCE_INT32, PLAIN_PER_PAIR = 1, 2, 4, 8
ROWS_PACKED = 0x1000000   # rows back to back in pair order, packed on the device


def force_shape(G, R):
    """Flag bits forcing the packed kernel shape (SQMI_FORCE_SHAPE)."""
    return ((G & 0xFF) << 8) | ((R & 0xFF) << 16)

# SQMI_NOALLOC_SCRATCH_BYTES: the device scratch of sqmi_align_score_noalloc.
NOALLOC_SCRATCH_BYTES = 64 << 20

WL_DNA150, WL_DNA_LOGNORMAL, WL_PROTEIN, WL_DNA1K, WL_DNA100K = 1, 2, 3, 4, 5

DNA = "ACGT"
AMINO_ACIDS = "ABCDEFGHIJKLMNOPQRSTUVWYZX*"   # SeqAn AminoAcid ordinal order

STATUS = {0: "ok", 1: "invalid argument", 2: "sequence sets differ in length",
          3: "unsupported request", 4: "no usable gfx950 device", 5: "device error",
          6: "out of device memory"}


class SqmiError(RuntimeError):
    def __init__(self, status, message=""):
        self.status = status
        super().__init__(f"{STATUS.get(status, status)}: {message}" if message else STATUS.get(status, str(status)))


class _Sequences(ctypes.Structure):
    _fields_ = [("residues", ctypes.c_void_p), ("offsets", ctypes.POINTER(ctypes.c_int64)),
                ("count", ctypes.c_int64)]


class _Scoring(ctypes.Structure):
    _fields_ = [("alphabet_size", ctypes.c_int32), ("gap_open", ctypes.c_int32),
                ("gap_extend", ctypes.c_int32), ("table", ctypes.POINTER(ctypes.c_int32))]


class Stats(ctypes.Structure):
    _fields_ = [("pairs", ctypes.c_int64), ("cells", ctypes.c_int64), ("pairs_int16", ctypes.c_int64),
                ("pairs_int32", ctypes.c_int64), ("work_items", ctypes.c_int64),
                ("padded_cells", ctypes.c_int64), ("host_plan_ms", ctypes.c_double),
                ("device_ms", ctypes.c_double), ("total_ms", ctypes.c_double),
                ("pairs_long", ctypes.c_int64), ("pairs_long_packed", ctypes.c_int64),
                ("pairs_lane", ctypes.c_int64), ("replayed", ctypes.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


# SQMI_ABI_VERSION of include/seqan_amd.h that the structures below mirror.
ABI_VERSION = 7

# Every symbol the C ABI declares (include/seqan_amd.h).
ABI_SYMBOLS = (
    "sqmi_create", "sqmi_destroy", "sqmi_set_stream", "sqmi_align_score", "sqmi_align_score_noalloc",
    "sqmi_align_counts", "sqmi_align_rows",
    "sqmi_align_local_rows", "sqmi_align_score_banded", "sqmi_align_rows_banded",
    "sqmi_get_stats", "sqmi_plan_stats", "sqmi_blosum62", "sqmi_status_string", "sqmi_last_error", "sqmi_abi_version",
    "sqmi_workload_create", "sqmi_workload_sizes", "sqmi_workload_copy", "sqmi_workload_destroy",
)

_lib = None


def lib():
    """The loaded native library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m seqan_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        i32p = ctypes.POINTER(ctypes.c_int32)
        i64p = ctypes.POINTER(ctypes.c_int64)
        L.sqmi_create.argtypes = [ctypes.c_int, ctypes.POINTER(vp)]
        L.sqmi_destroy.argtypes = [vp]
        L.sqmi_set_stream.argtypes = [vp, vp]
        L.sqmi_align_score.argtypes = [vp, ctypes.POINTER(_Sequences), ctypes.POINTER(_Sequences),
                                       ctypes.POINTER(_Scoring), ctypes.c_uint32, vp, ctypes.c_uint32]
        L.sqmi_align_score_noalloc.argtypes = [vp, ctypes.POINTER(_Sequences), ctypes.POINTER(_Sequences),
                                       ctypes.POINTER(_Scoring), ctypes.c_uint32, vp, ctypes.c_uint32]
        L.sqmi_align_counts.argtypes = [vp, ctypes.POINTER(_Sequences), ctypes.POINTER(_Sequences),
                                        ctypes.POINTER(_Scoring), ctypes.c_uint32, vp, vp, ctypes.c_uint32]
        L.sqmi_align_rows.argtypes = [vp, ctypes.POINTER(_Sequences), ctypes.POINTER(_Sequences),
                                      ctypes.POINTER(_Scoring), ctypes.c_uint32, vp, vp, vp, vp, vp, ctypes.c_uint32]
        L.sqmi_align_score_banded.argtypes = [vp, ctypes.POINTER(_Sequences), ctypes.POINTER(_Sequences),
                                              ctypes.POINTER(_Scoring), ctypes.c_uint32, ctypes.c_int64,
                                              ctypes.c_int64, vp, ctypes.c_uint32]
        L.sqmi_align_local_rows.argtypes = [vp, ctypes.POINTER(_Sequences), ctypes.POINTER(_Sequences),
                                            ctypes.POINTER(_Scoring), vp, vp, vp, vp, vp, vp, ctypes.c_uint32]
        L.sqmi_align_rows_banded.argtypes = [vp, ctypes.POINTER(_Sequences), ctypes.POINTER(_Sequences),
                                             ctypes.POINTER(_Scoring), ctypes.c_uint32, ctypes.c_int64,
                                             ctypes.c_int64, vp, vp, vp, vp, vp, vp, ctypes.c_uint32]
        L.sqmi_get_stats.argtypes = [vp, ctypes.POINTER(Stats)]
        L.sqmi_plan_stats.argtypes = [ctypes.POINTER(_Sequences), ctypes.POINTER(_Sequences),
                                      ctypes.POINTER(_Scoring), ctypes.c_uint32, ctypes.c_uint32,
                                      ctypes.POINTER(Stats)]
        L.sqmi_blosum62.restype = i32p
        L.sqmi_status_string.restype = ctypes.c_char_p
        L.sqmi_status_string.argtypes = [ctypes.c_int]
        L.sqmi_last_error.restype = ctypes.c_char_p
        L.sqmi_last_error.argtypes = [vp]
        L.sqmi_workload_create.argtypes = [ctypes.c_int, ctypes.c_uint64, ctypes.c_int64, ctypes.c_int64,
                                           ctypes.c_int64, ctypes.POINTER(vp)]
        L.sqmi_workload_sizes.argtypes = [vp, i64p, i64p]
        L.sqmi_workload_copy.argtypes = [vp, vp, i64p, vp, i64p]
        L.sqmi_workload_destroy.argtypes = [vp]
        for fn in ABI_SYMBOLS:
            f = getattr(L, fn)
            if fn not in ("sqmi_blosum62", "sqmi_status_string", "sqmi_last_error"):
                f.restype = ctypes.c_int
        if L.sqmi_abi_version() != ABI_VERSION:
            raise ImportError(f"{LIB_PATH} has ABI version {L.sqmi_abi_version()}, this package expects "
                              f"{ABI_VERSION}: rebuild it with `python -m seqan_amd.build`")
        _lib = L
    return _lib


# ---------------------------------------------------------------------------
# Scoring
# ---------------------------------------------------------------------------
@dataclass
class Score:
    """A substitution table (row = residue of H) plus affine gap costs."""
    table: np.ndarray
    gap_extend: int
    gap_open: int

    @property
    def alphabet_size(self):
        return int(self.table.shape[0])

    def _c(self):
        t = np.ascontiguousarray(self.table, dtype=np.int32)
        self._keep = t
        return _Scoring(t.shape[0], int(self.gap_open), int(self.gap_extend),
                        t.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))


def simple_score(match, mismatch, gap_extend, gap_open, alphabet_size=4) -> Score:
    """Score<int, Simple>(match, mismatch, gapExtend, gapOpen) (score_simple.h:94-97)."""
    t = np.full((alphabet_size, alphabet_size), mismatch, dtype=np.int32)
    np.fill_diagonal(t, match)
    return Score(t, gap_extend, gap_open)


def blosum62(gap_extend=-1, gap_open=-11) -> Score:
    """Blosum62(gapExtend, gapOpen) (score/score_matrix.h:112)."""
    p = lib().sqmi_blosum62()
    t = np.ctypeslib.as_array(p, shape=(27 * 27,)).reshape(27, 27).copy()
    return Score(t, gap_extend, gap_open)


# ---------------------------------------------------------------------------
# Sequence sets
# ---------------------------------------------------------------------------
@dataclass
class SeqSet:
    """Concatenated residue ordinals plus count+1 offsets (host numpy arrays)."""
    residues: np.ndarray
    offsets: np.ndarray

    def __len__(self):
        return len(self.offsets) - 1

    def __getitem__(self, i):
        return self.residues[self.offsets[i]:self.offsets[i + 1]]

    def lengths(self):
        return np.diff(self.offsets)

    @classmethod
    def from_list(cls, seqs, alphabet=DNA):
        codes = []
        lut = {c: i for i, c in enumerate(alphabet)}
        for s in seqs:
            if isinstance(s, str):
                codes.append(np.array([lut[c] for c in s.upper()], dtype=np.uint8))
            else:
                codes.append(np.asarray(s, dtype=np.uint8))
        off = np.zeros(len(codes) + 1, dtype=np.int64)
        if codes:
            off[1:] = np.cumsum([len(c) for c in codes])
        res = np.concatenate(codes) if codes and off[-1] > 0 else np.zeros(0, np.uint8)
        return cls(np.ascontiguousarray(res, dtype=np.uint8), off)


def plan_stats(set_h: SeqSet, set_v: SeqSet, score: Score, mode=0, flags=0):
    """Host-only planner dry run (no device): routing and work-item counts."""
    ch, keep_h = _seqs(set_h)
    cv, keep_v = _seqs(set_v)
    sc = score._c()
    st = Stats()
    rc = lib().sqmi_plan_stats(ctypes.byref(ch), ctypes.byref(cv), ctypes.byref(sc), mode, flags,
                               ctypes.byref(st))
    if rc:
        raise SqmiError(rc, "sqmi_plan_stats")
    return st.as_dict()


def workload(config, seed, first, count, length_override=0):
    """Seeded synthetic pair window -> (SeqSet H, SeqSet V)."""
    L = lib()
    w = ctypes.c_void_p()
    st = L.sqmi_workload_create(config, seed, first, count, length_override, ctypes.byref(w))
    if st:
        raise SqmiError(st, "workload")
    try:
        ht, vt = ctypes.c_int64(), ctypes.c_int64()
        L.sqmi_workload_sizes(w, ctypes.byref(ht), ctypes.byref(vt))
        h = np.empty(ht.value, np.uint8)
        v = np.empty(vt.value, np.uint8)
        ho = np.empty(count + 1, np.int64)
        vo = np.empty(count + 1, np.int64)
        L.sqmi_workload_copy(w, h.ctypes.data, ho.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                             v.ctypes.data, vo.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)))
    finally:
        L.sqmi_workload_destroy(w)
    return SeqSet(h, ho), SeqSet(v, vo)


# ---------------------------------------------------------------------------
# Context
# ---------------------------------------------------------------------------
def _seqs(s: SeqSet, residues_ptr=None):
    off = np.ascontiguousarray(s.offsets, dtype=np.int64)
    res = residues_ptr if residues_ptr is not None else (
        s.residues.ctypes.data if s.residues.size else np.zeros(1, np.uint8).ctypes.data)
    c = _Sequences(res, off.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), len(off) - 1)
    return c, off


class Context:
    """One device context (owns scratch buffers and a stream)."""

    def __init__(self, device=0):
        self._L = lib()
        self._h = ctypes.c_void_p()
        st = self._L.sqmi_create(device, ctypes.byref(self._h))
        if st:
            raise SqmiError(st, f"sqmi_create(device={device})")

    def close(self):
        if self._h:
            self._L.sqmi_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, st, what):
        if st:
            msg = self._L.sqmi_last_error(self._h)
            raise SqmiError(st, f"{what}: {msg.decode() if msg else ''}")

    def set_stream(self, stream_ptr):
        self._check(self._L.sqmi_set_stream(self._h, stream_ptr), "sqmi_set_stream")

    def stats(self):
        s = Stats()
        self._L.sqmi_get_stats(self._h, ctypes.byref(s))
        return s.as_dict()

    def align_score(self, set_h: SeqSet, set_v: SeqSet, score: Score, mode=0, flags=0,
                    h_dev=None, v_dev=None, scores_dev=None):
        """Scores of all pairs.  h_dev / v_dev / scores_dev: optional device
        addresses (ints) of the residues / output (sets the device flags)."""
        ch, keep_h = _seqs(set_h, h_dev)
        cv, keep_v = _seqs(set_v, v_dev)
        if h_dev is not None or v_dev is not None:
            assert h_dev is not None and v_dev is not None
            flags |= RESIDUES_ON_DEVICE
        sc = score._c()
        if scores_dev is not None:
            flags |= SCORES_ON_DEVICE
            out = None
            ptr = scores_dev
        else:
            out = np.empty(len(set_h.offsets) - 1, np.int32)
            ptr = out.ctypes.data
        st = self._L.sqmi_align_score(self._h, ctypes.byref(ch), ctypes.byref(cv), ctypes.byref(sc),
                                      mode, ptr, flags)
        self._check(st, "sqmi_align_score")
        return out

    def align_score_noalloc(self, set_h: SeqSet, set_v: SeqSet, score: Score, mode=0, flags=0,
                            h_dev=None, v_dev=None, scores_dev=None):
        """align_score through sqmi_align_score_noalloc: no device allocation,
        every buffer beyond the inputs and outputs carved from the library's
        per-GPU scratch array (SQMI_NOALLOC_SCRATCH_BYTES)."""
        ch, keep_h = _seqs(set_h, h_dev)
        cv, keep_v = _seqs(set_v, v_dev)
        if h_dev is not None or v_dev is not None:
            assert h_dev is not None and v_dev is not None
            flags |= RESIDUES_ON_DEVICE
        sc = score._c()
        if scores_dev is not None:
            flags |= SCORES_ON_DEVICE
            out, ptr = None, scores_dev
        else:
            out = np.empty(len(set_h.offsets) - 1, np.int32)
            ptr = out.ctypes.data
        st = self._L.sqmi_align_score_noalloc(self._h, ctypes.byref(ch), ctypes.byref(cv), ctypes.byref(sc),
                                              mode, ptr, flags)
        self._check(st, "sqmi_align_score_noalloc")
        return out

    def align_score_banded(self, set_h: SeqSet, set_v: SeqSet, score: Score, lower_diag: int, upper_diag: int,
                           mode=0, flags=0, h_dev=None, v_dev=None, scores_dev=None):
        """Banded scores: cells with lower_diag <= j - i <= upper_diag (global
        modes or LOCAL).  Device addresses as in align_score."""
        ch, keep_h = _seqs(set_h, h_dev)
        cv, keep_v = _seqs(set_v, v_dev)
        if h_dev is not None or v_dev is not None:
            assert h_dev is not None and v_dev is not None
            flags |= RESIDUES_ON_DEVICE
        sc = score._c()
        if scores_dev is not None:
            flags |= SCORES_ON_DEVICE
            out, ptr = None, scores_dev
        else:
            out = np.empty(len(set_h.offsets) - 1, np.int32)
            ptr = out.ctypes.data
        st = self._L.sqmi_align_score_banded(self._h, ctypes.byref(ch), ctypes.byref(cv), ctypes.byref(sc), mode,
                                             int(lower_diag), int(upper_diag), ptr, flags)
        self._check(st, "sqmi_align_score_banded")
        return out

    def align_counts(self, set_h: SeqSet, set_v: SeqSet, score: Score, mode=0, flags=0,
                     h_dev=None, v_dev=None, scores_dev=None, counts_dev=None):
        """Scores and counts (alpha^2 substitution counts, gap opens, gap
        extensions) of all pairs.  Device addresses as in align_score; with
        scores_dev, counts_dev must be given too and nothing is returned."""
        ch, keep_h = _seqs(set_h, h_dev)
        cv, keep_v = _seqs(set_v, v_dev)
        if h_dev is not None or v_dev is not None:
            assert h_dev is not None and v_dev is not None
            flags |= RESIDUES_ON_DEVICE
        sc = score._c()
        if scores_dev is not None:
            assert counts_dev is not None
            flags |= SCORES_ON_DEVICE
            st = self._L.sqmi_align_counts(self._h, ctypes.byref(ch), ctypes.byref(cv), ctypes.byref(sc),
                                           mode, scores_dev, counts_dev, flags)
            self._check(st, "sqmi_align_counts")
            return None
        n = len(set_h.offsets) - 1
        scores = np.empty(n, np.int32)
        A = score.alphabet_size
        counts = np.empty((n, A * A + 2), np.int32)
        st = self._L.sqmi_align_counts(self._h, ctypes.byref(ch), ctypes.byref(cv), ctypes.byref(sc),
                                       mode, scores.ctypes.data, counts.ctypes.data, flags)
        self._check(st, "sqmi_align_counts")
        return scores, counts

    def align_rows(self, set_h: SeqSet, set_v: SeqSet, score: Score, mode=0, flags=0):
        """Scores, counts and alignment rows of all pairs (global alignment).
        Returns (scores, counts, op_off, ops): pair i's rows are
        ops[op_off[i]:op_off[i + 1]], one byte per column: b'M' (residues of
        H and V), b'V' (a gap in the H row), b'H' (a gap in the V row)."""
        ch, keep_h = _seqs(set_h)
        cv, keep_v = _seqs(set_v)
        sc = score._c()
        n = len(set_h.offsets) - 1
        A = score.alphabet_size
        scores = np.empty(n, np.int32)
        counts = np.empty((n, A * A + 2), np.int32)
        seg = np.zeros(n + 1, np.int64)
        seg[1:] = np.cumsum(set_h.lengths() + set_v.lengths())
        buf = np.empty(max(int(seg[-1]), 1), np.uint8)
        lens = np.zeros(n, np.int64)
        st = self._L.sqmi_align_rows(self._h, ctypes.byref(ch), ctypes.byref(cv), ctypes.byref(sc), mode,
                                     scores.ctypes.data, counts.ctypes.data, buf.ctypes.data, seg.ctypes.data,
                                     lens.ctypes.data, flags | ROWS_PACKED)
        self._check(st, "sqmi_align_rows")
        return scores, counts, *_packed_rows(buf, lens)

    def align_local_rows(self, set_h: SeqSet, set_v: SeqSet, score: Score, flags=0):
        """Local alignment: scores, counts and the rows of every local window.
        Returns (scores, counts, op_off, ops, begins) with rows as in
        align_rows and begins[i] = (first H position, first V position) of
        pair i's window."""
        ch, keep_h = _seqs(set_h)
        cv, keep_v = _seqs(set_v)
        sc = score._c()
        n = len(set_h.offsets) - 1
        A = score.alphabet_size
        scores = np.empty(n, np.int32)
        counts = np.empty((n, A * A + 2), np.int32)
        seg = np.zeros(n + 1, np.int64)
        seg[1:] = np.cumsum(set_h.lengths() + set_v.lengths())
        buf = np.empty(max(int(seg[-1]), 1), np.uint8)
        lens = np.zeros(n, np.int64)
        begins = np.zeros((n, 2), np.int64)
        st = self._L.sqmi_align_local_rows(self._h, ctypes.byref(ch), ctypes.byref(cv), ctypes.byref(sc),
                                           scores.ctypes.data, counts.ctypes.data, buf.ctypes.data,
                                           seg.ctypes.data, lens.ctypes.data, begins.ctypes.data,
                                           flags | ROWS_PACKED)
        self._check(st, "sqmi_align_local_rows")
        return (scores, counts, *_packed_rows(buf, lens), begins)

    def align_rows_banded(self, set_h: SeqSet, set_v: SeqSet, score: Score, lower_diag: int, upper_diag: int,
                          mode=0, flags=0, rows=True, h_dev=None, v_dev=None, out_dev=None):
        """Banded alignment (cells with lower_diag <= j - i <= upper_diag):
        (scores, counts, op_off, ops) as align_rows, plus begins as in
        align_local_rows with mode LOCAL; rows=False returns (scores,
        counts) only.  h_dev / v_dev: device addresses of the residues;
        out_dev: device addresses (scores, counts, ops, op_lengths[, begins])
        of the outputs, rows packed from ops (returns None: the results stay
        on the device; ops must hold sum(|H| + |V|) bytes)."""
        ch, keep_h = _seqs(set_h, h_dev)
        cv, keep_v = _seqs(set_v, v_dev)
        if h_dev is not None or v_dev is not None:
            assert h_dev is not None and v_dev is not None
            flags |= RESIDUES_ON_DEVICE
        sc = score._c()
        n = len(set_h.offsets) - 1
        A = score.alphabet_size
        if out_dev is not None:
            seg = np.zeros(n + 1, np.int64)
            seg[1:] = np.cumsum(set_h.lengths() + set_v.lengths())
            begins = out_dev[4] if len(out_dev) > 4 else None
            st = self._L.sqmi_align_rows_banded(self._h, ctypes.byref(ch), ctypes.byref(cv), ctypes.byref(sc), mode,
                                                lower_diag, upper_diag, out_dev[0], out_dev[1], out_dev[2],
                                                seg.ctypes.data, out_dev[3], begins,
                                                flags | SCORES_ON_DEVICE | ROWS_PACKED)
            self._check(st, "sqmi_align_rows_banded")
            return None
        scores = np.empty(n, np.int32)
        counts = np.empty((n, A * A + 2), np.int32)
        local = bool(mode & LOCAL)
        if not rows:
            st = self._L.sqmi_align_rows_banded(self._h, ctypes.byref(ch), ctypes.byref(cv), ctypes.byref(sc), mode,
                                                lower_diag, upper_diag, scores.ctypes.data, counts.ctypes.data,
                                                None, None, None, None, flags)
            self._check(st, "sqmi_align_rows_banded")
            return scores, counts
        seg = np.zeros(n + 1, np.int64)
        seg[1:] = np.cumsum(set_h.lengths() + set_v.lengths())
        buf = np.empty(max(int(seg[-1]), 1), np.uint8)
        lens = np.zeros(n, np.int64)
        begins = np.zeros((n, 2), np.int64)
        st = self._L.sqmi_align_rows_banded(self._h, ctypes.byref(ch), ctypes.byref(cv), ctypes.byref(sc), mode,
                                            lower_diag, upper_diag, scores.ctypes.data, counts.ctypes.data,
                                            buf.ctypes.data, seg.ctypes.data, lens.ctypes.data,
                                            begins.ctypes.data if local else None, flags | ROWS_PACKED)
        self._check(st, "sqmi_align_rows_banded")
        out = (scores, counts, *_packed_rows(buf, lens))
        return (*out, begins) if local else out


def _packed_rows(buf, lens):
    """(op_off, ops) from rows the library wrote back to back (ROWS_PACKED)."""
    op_off = np.zeros(len(lens) + 1, np.int64)
    np.cumsum(lens, out=op_off[1:])
    return op_off, buf[:op_off[-1]]


# ---------------------------------------------------------------------------
# Device memory through the HIP runtime the library itself links
# (/opt/rocm/lib/libamdhip64.so.7).  A second HIP runtime in the same process
# (for instance the one bundled with a PyTorch wheel) cannot share the GPU,
# so device-resident inputs are staged here rather than through torch.
# ---------------------------------------------------------------------------
_hip = None


def hip():
    global _hip
    if _hip is None:
        lib()   # loads the same libamdhip64 instance the engine uses
        H = ctypes.CDLL("libamdhip64.so.7")
        H.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        H.hipFree.argtypes = [ctypes.c_void_p]
        H.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        H.hipDeviceSynchronize.argtypes = []
        H.hipSetDevice.argtypes = [ctypes.c_int]
        _hip = H
    return _hip


class DeviceBuffer:
    """A plain device allocation (hipMalloc) with host <-> device copies."""

    def __init__(self, nbytes, device=0):
        H = hip()
        H.hipSetDevice(device)
        self.ptr = ctypes.c_void_p()
        self.nbytes = int(nbytes)
        err = H.hipMalloc(ctypes.byref(self.ptr), max(self.nbytes, 16))
        if err:
            raise SqmiError(6, f"hipMalloc({nbytes}) failed: {err}")

    @classmethod
    def from_array(cls, a: np.ndarray):
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        if a.nbytes:
            err = hip().hipMemcpy(b.ptr, a.ctypes.data, a.nbytes, 1)   # hipMemcpyHostToDevice
            if err:
                raise SqmiError(5, f"hipMemcpy H2D failed: {err}")
        return b

    def to_array(self, dtype, count):
        out = np.empty(count, dtype=dtype)
        if out.nbytes:
            err = hip().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, 2)   # hipMemcpyDeviceToHost
            if err:
                raise SqmiError(5, f"hipMemcpy D2H failed: {err}")
        return out

    @property
    def address(self):
        return self.ptr.value

    def free(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def synchronize():
    hip().hipDeviceSynchronize()


_default_ctx = None


def default_context():
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


def _mode_from_config(align_config):
    top, left, right, bottom = (bool(x) for x in align_config)
    return (FREE_TOP if top else 0) | (FREE_LEFT if left else 0) | (FREE_RIGHT if right else 0) | (
        FREE_BOTTOM if bottom else 0)


def _as_set(s, alphabet):
    return s if isinstance(s, SeqSet) else SeqSet.from_list(s, alphabet)


def global_alignment_score(set_h, set_v, score: Score, align_config=(False, False, False, False),
                           alphabet=DNA, ctx: Context = None, lower_diag=None, upper_diag=None):
    """globalAlignmentScore(ExecutionPolicy<Parallel,Vectorial>, setH, setV, score, AlignConfig
    [, lowerDiag, upperDiag])."""
    ctx = ctx or default_context()
    if lower_diag is not None or upper_diag is not None:
        return ctx.align_score_banded(_as_set(set_h, alphabet), _as_set(set_v, alphabet), score, lower_diag,
                                      upper_diag, _mode_from_config(align_config))
    return ctx.align_score(_as_set(set_h, alphabet), _as_set(set_v, alphabet), score,
                           _mode_from_config(align_config))


def local_alignment_score(set_h, set_v, score: Score, alphabet=AMINO_ACIDS, ctx: Context = None,
                          lower_diag=None, upper_diag=None):
    """localAlignmentScore(ExecutionPolicy<Parallel,Vectorial>, setH, setV, score
    [, lowerDiag, upperDiag])."""
    ctx = ctx or default_context()
    if lower_diag is not None or upper_diag is not None:
        return ctx.align_score_banded(_as_set(set_h, alphabet), _as_set(set_v, alphabet), score, lower_diag,
                                      upper_diag, LOCAL)
    return ctx.align_score(_as_set(set_h, alphabet), _as_set(set_v, alphabet), score, LOCAL)


def global_alignment(set_h, set_v, score: Score, align_config=(False, False, False, False), alphabet=DNA,
                     ctx: Context = None, lower_diag=None, upper_diag=None):
    """globalAlignment(ExecutionPolicy<Parallel,Vectorial>, gapsSetH, gapsSetV, score, AlignConfig
    [, lowerDiag, upperDiag]): (scores, counts, op_off, ops), rows as Context.align_rows; with a band,
    as Context.align_rows_banded."""
    ctx = ctx or default_context()
    H, V = _as_set(set_h, alphabet), _as_set(set_v, alphabet)
    if (lower_diag is None) != (upper_diag is None):
        raise ValueError("give both lower_diag and upper_diag, or neither")
    if lower_diag is not None:
        return ctx.align_rows_banded(H, V, score, lower_diag, upper_diag, mode=_mode_from_config(align_config))
    return ctx.align_rows(H, V, score, _mode_from_config(align_config))


def local_alignment(set_h, set_v, score: Score, alphabet=AMINO_ACIDS, ctx: Context = None, lower_diag=None,
                    upper_diag=None):
    """localAlignment(ExecutionPolicy<Parallel,Vectorial>, gapsSetH, gapsSetV, score
    [, lowerDiag, upperDiag]): (scores, counts, op_off, ops, begins), as Context.align_local_rows;
    with a band, as Context.align_rows_banded with mode LOCAL."""
    ctx = ctx or default_context()
    H, V = _as_set(set_h, alphabet), _as_set(set_v, alphabet)
    if (lower_diag is None) != (upper_diag is None):
        raise ValueError("give both lower_diag and upper_diag, or neither")
    if lower_diag is not None:
        return ctx.align_rows_banded(H, V, score, lower_diag, upper_diag, mode=LOCAL)
    return ctx.align_local_rows(H, V, score)
