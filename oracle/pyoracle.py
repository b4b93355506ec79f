"""TEST INFRASTRUCTURE ONLY: Python access to the oracle.

Two checkers live under ``oracle/``:

* ``liboracle.so`` -- the scalar C restatement of SeqAn 2's affine DP
  (``oracle.c``; every function cites the reference file:line it follows);
* ``_ref/ref_avx2`` / ``_ref/ref_avx512`` -- the reference library itself,
  compiled from its own headers by ``oracle/Makefile`` and driven through
  ``ref_driver.cpp`` (globalAlignmentScore / localAlignmentScore /
  globalAlignment under ExecutionPolicy<Parallel, Vectorial>).

Only tests, ``__graft_entry__.smoke()`` and bench.py's CPU-baseline leg use
this module.  The product (``seqan_amd``) never imports it.
"""
from __future__ import annotations

import ctypes
import os
import struct
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

FREE_TOP, FREE_LEFT, FREE_RIGHT, FREE_BOTTOM, LOCAL = 1, 2, 4, 8, 16


class _Scoring(ctypes.Structure):
    _fields_ = [("alpha", ctypes.c_int32), ("table", ctypes.POINTER(ctypes.c_int32)),
                ("gap_open", ctypes.c_int32), ("gap_extend", ctypes.c_int32)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            subprocess.run(["make", "-C", HERE, "liboracle.so"], check=True,
                           stdout=subprocess.DEVNULL)
        L = ctypes.CDLL(path)
        u8p = ctypes.POINTER(ctypes.c_uint8)
        i64p = ctypes.POINTER(ctypes.c_int64)
        i32p = ctypes.POINTER(ctypes.c_int32)
        sp = ctypes.POINTER(_Scoring)
        L.orc_score.restype = ctypes.c_int32
        L.orc_score.argtypes = [u8p, ctypes.c_int64, u8p, ctypes.c_int64, sp, ctypes.c_uint32]
        L.orc_score_batch.restype = None
        L.orc_score_batch.argtypes = [u8p, i64p, u8p, i64p, ctypes.c_int64, sp,
                                      ctypes.c_uint32, i32p, ctypes.c_int]
        L.orc_traceback.restype = ctypes.c_int32
        L.orc_traceback.argtypes = [u8p, ctypes.c_int64, u8p, ctypes.c_int64, sp, ctypes.c_uint32,
                                    i32p, ctypes.c_char_p, i64p, i64p]
        L.orc_traceback_batch.restype = None
        L.orc_traceback_batch.argtypes = [u8p, i64p, u8p, i64p, ctypes.c_int64, sp,
                                          ctypes.c_uint32, i32p, i32p, ctypes.c_int]
        L.orc_rows_batch.restype = None
        L.orc_rows_batch.argtypes = [u8p, i64p, u8p, i64p, ctypes.c_int64, sp, ctypes.c_uint32, i32p, i32p,
                                     u8p, i64p, i64p, i64p, ctypes.c_int]
        L.orc_rows_batch_banded.restype = None
        L.orc_rows_batch_banded.argtypes = [u8p, i64p, u8p, i64p, ctypes.c_int64, sp, ctypes.c_uint32,
                                            ctypes.c_int64, ctypes.c_int64, i32p, i32p, u8p, i64p, i64p, i64p,
                                            ctypes.c_int]
        L.orc_score_banded.restype = ctypes.c_int32
        L.orc_score_banded.argtypes = [u8p, ctypes.c_int64, u8p, ctypes.c_int64, sp, ctypes.c_uint32,
                                       ctypes.c_int64, ctypes.c_int64]
        L.orc_brute_force.restype = ctypes.c_int32
        L.orc_brute_force.argtypes = [u8p, ctypes.c_int64, u8p, ctypes.c_int64, sp, ctypes.c_uint32]
        _lib = L
    return _lib


def _ptr(a, t):
    return a.ctypes.data_as(ctypes.POINTER(t))


class Scoring:
    """Substitution table (alpha x alpha, row = H residue) + affine gap costs."""

    def __init__(self, table, gap_open, gap_extend):
        self.table = np.ascontiguousarray(np.asarray(table, dtype=np.int32))
        self.alpha = int(self.table.shape[0])
        self.gap_open = int(gap_open)
        self.gap_extend = int(gap_extend)
        self._c = _Scoring(self.alpha, _ptr(self.table, ctypes.c_int32), self.gap_open,
                           self.gap_extend)

    @classmethod
    def simple(cls, match, mismatch, gap_extend, gap_open, alpha=4):
        """Score<int, Simple>(match, mismatch, gapExtend, gapOpen) (score_simple.h:94-97)."""
        t = np.full((alpha, alpha), mismatch, dtype=np.int32)
        np.fill_diagonal(t, match)
        return cls(t, gap_open, gap_extend)


def _u8(s):
    return np.ascontiguousarray(np.frombuffer(bytes(s), dtype=np.uint8) if isinstance(s, (bytes, bytearray))
                                else np.asarray(s, dtype=np.uint8))


def score(h, v, sc: Scoring, mode: int) -> int:
    h, v = _u8(h), _u8(v)
    return int(lib().orc_score(_ptr(h, ctypes.c_uint8), len(h), _ptr(v, ctypes.c_uint8), len(v),
                               ctypes.byref(sc._c), mode))


def brute_force(h, v, sc: Scoring, mode: int) -> int:
    h, v = _u8(h), _u8(v)
    return int(lib().orc_brute_force(_ptr(h, ctypes.c_uint8), len(h), _ptr(v, ctypes.c_uint8),
                                     len(v), ctypes.byref(sc._c), mode))


def traceback(h, v, sc: Scoring, mode: int = 0, with_begins: bool = False):
    """Returns (score, counts[alpha*alpha+2], ops string), plus the local
    window's begin positions (in H, in V) with with_begins."""
    h, v = _u8(h), _u8(v)
    counts = np.zeros(sc.alpha * sc.alpha + 2, dtype=np.int32)
    buf = ctypes.create_string_buffer(len(h) + len(v) + 2)
    nops = ctypes.c_int64(0)
    begins = np.zeros(2, dtype=np.int64)
    s = lib().orc_traceback(_ptr(h, ctypes.c_uint8), len(h), _ptr(v, ctypes.c_uint8), len(v),
                            ctypes.byref(sc._c), mode, _ptr(counts, ctypes.c_int32), buf,
                            ctypes.byref(nops), _ptr(begins, ctypes.c_int64))
    out = (int(s), counts, buf.raw[:nops.value].decode())
    return out + ((int(begins[0]), int(begins[1])),) if with_begins else out


def score_banded(h, v, sc: Scoring, mode: int, lower: int, upper: int) -> int:
    h, v = _u8(h), _u8(v)
    return int(lib().orc_score_banded(_ptr(h, ctypes.c_uint8), len(h), _ptr(v, ctypes.c_uint8), len(v),
                                      ctypes.byref(sc._c), mode, lower, upper))


def score_batch(h_res, h_off, v_res, v_off, sc: Scoring, mode: int, threads: int = 0):
    h_res, v_res = _u8(h_res), _u8(v_res)
    h_off = np.ascontiguousarray(h_off, dtype=np.int64)
    v_off = np.ascontiguousarray(v_off, dtype=np.int64)
    n = len(h_off) - 1
    out = np.zeros(n, dtype=np.int32)
    lib().orc_score_batch(_ptr(h_res, ctypes.c_uint8), _ptr(h_off, ctypes.c_int64),
                          _ptr(v_res, ctypes.c_uint8), _ptr(v_off, ctypes.c_int64), n,
                          ctypes.byref(sc._c), mode, _ptr(out, ctypes.c_int32), threads)
    return out


def rows_batch(h_res, h_off, v_res, v_off, sc: Scoring, mode: int = 0, threads: int = 0,
               with_begins: bool = False, band=None):
    """Scores, counts and rows of every pair: returns (scores, counts,
    op_off, ops) with pair k's ops ('M', 'V', 'H' bytes, left to right) in
    ops[op_off[k]:op_off[k + 1]]; with_begins adds begins [count, 2], the
    positions in H and V where each local window starts.  band =
    (lower_diag, upper_diag) restricts every pair to that band."""
    h_res, v_res = _u8(h_res), _u8(v_res)
    h_off = np.ascontiguousarray(h_off, dtype=np.int64)
    v_off = np.ascontiguousarray(v_off, dtype=np.int64)
    n = len(h_off) - 1
    cap = np.zeros(n + 1, dtype=np.int64)
    cap[1:] = np.cumsum(np.diff(h_off) + np.diff(v_off) + 2)
    scores = np.zeros(n, dtype=np.int32)
    counts = np.zeros((n, sc.alpha * sc.alpha + 2), dtype=np.int32)
    buf = np.zeros(int(cap[-1]), dtype=np.uint8)
    lens = np.zeros(n, dtype=np.int64)
    begins = np.zeros((n, 2), dtype=np.int64)
    if band is None:
        lib().orc_rows_batch(_ptr(h_res, ctypes.c_uint8), _ptr(h_off, ctypes.c_int64),
                             _ptr(v_res, ctypes.c_uint8), _ptr(v_off, ctypes.c_int64), n,
                             ctypes.byref(sc._c), mode, _ptr(scores, ctypes.c_int32),
                             _ptr(counts, ctypes.c_int32), _ptr(buf, ctypes.c_uint8), _ptr(cap, ctypes.c_int64),
                             _ptr(lens, ctypes.c_int64), _ptr(begins, ctypes.c_int64), threads)
    else:
        lib().orc_rows_batch_banded(_ptr(h_res, ctypes.c_uint8), _ptr(h_off, ctypes.c_int64),
                                    _ptr(v_res, ctypes.c_uint8), _ptr(v_off, ctypes.c_int64), n,
                                    ctypes.byref(sc._c), mode, ctypes.c_int64(int(band[0])),
                                    ctypes.c_int64(int(band[1])), _ptr(scores, ctypes.c_int32),
                                    _ptr(counts, ctypes.c_int32), _ptr(buf, ctypes.c_uint8),
                                    _ptr(cap, ctypes.c_int64), _ptr(lens, ctypes.c_int64),
                                    _ptr(begins, ctypes.c_int64), threads)
    op_off = np.zeros(n + 1, dtype=np.int64)
    op_off[1:] = np.cumsum(lens)
    ops = np.concatenate([buf[cap[k]:cap[k] + lens[k]] for k in range(n)]) if n else np.zeros(0, np.uint8)
    return (scores, counts, op_off, ops, begins) if with_begins else (scores, counts, op_off, ops)


def traceback_batch(h_res, h_off, v_res, v_off, sc: Scoring, mode: int = 0, threads: int = 0):
    h_res, v_res = _u8(h_res), _u8(v_res)
    h_off = np.ascontiguousarray(h_off, dtype=np.int64)
    v_off = np.ascontiguousarray(v_off, dtype=np.int64)
    n = len(h_off) - 1
    scores = np.zeros(n, dtype=np.int32)
    counts = np.zeros((n, sc.alpha * sc.alpha + 2), dtype=np.int32)
    lib().orc_traceback_batch(_ptr(h_res, ctypes.c_uint8), _ptr(h_off, ctypes.c_int64),
                              _ptr(v_res, ctypes.c_uint8), _ptr(v_off, ctypes.c_int64), n,
                              ctypes.byref(sc._c), mode, _ptr(scores, ctypes.c_int32),
                              _ptr(counts, ctypes.c_int32), threads)
    return scores, counts


# --------------------------------------------------------------------------
# The compiled reference (oracle/_ref).
# --------------------------------------------------------------------------
REF_DIR = os.path.join(HERE, "_ref")
OP_SCORE, OP_COUNTS_SIMD, OP_COUNTS_SCALAR, OP_SCORE_BANDED, OP_ROWS_BANDED = 0, 1, 2, 4, 5


def _cpu_has_avx512():
    try:
        with open("/proc/cpuinfo") as f:
            flags = f.read()
        return all(x in flags for x in ("avx512f", "avx512bw", "avx512vl", "avx512dq", "avx512cd"))
    except OSError:
        return False


def ref_binary():
    """Path of the reference driver matching this CPU (None when not built)."""
    names = (["ref_avx512"] if _cpu_has_avx512() else []) + ["ref_avx2"]
    for nm in names:
        p = os.path.join(REF_DIR, nm)
        if os.path.exists(p):
            return p
    return None


def run_reference(op, h_res, h_off, v_res, v_off, mode, *, alphabet="dna", matrix="simple",
                  match=0, mismatch=0, gap_extend=0, gap_open=0, threads=0, repeats=1,
                  timeout=3600, lower_diag=0, upper_diag=0):
    """Runs the reference library on a batch.  Returns a dict with scores,
    counts, op_off and ops (ops 1/2; begins [count, 2] too with mode LOCAL),
    best_seconds, simd_lanes, threads."""
    exe = ref_binary()
    if exe is None:
        raise FileNotFoundError("oracle/_ref is not built")
    h_res, v_res = _u8(h_res), _u8(v_res)
    h_off = np.ascontiguousarray(h_off, dtype=np.int64)
    v_off = np.ascontiguousarray(v_off, dtype=np.int64)
    n = len(h_off) - 1
    alpha_id = {"dna": 0, "aa": 1}[alphabet]
    mat_id = {"simple": 0, "blosum62": 1}[matrix]
    with tempfile.TemporaryDirectory(prefix="sqref") as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(fin, "wb") as f:
            f.write(b"SQRI")
            f.write(struct.pack("<IIIiiiiIiiq", op, mode, alpha_id, match, mismatch, gap_extend,
                                gap_open, mat_id, threads, repeats, n))
            if op in (OP_SCORE_BANDED, OP_ROWS_BANDED):
                f.write(struct.pack("<qq", lower_diag, upper_diag))
            f.write(h_off.tobytes())
            f.write(v_off.tobytes())
            f.write(h_res.tobytes())
            f.write(v_res.tobytes())
        subprocess.run([exe, fin, fout], check=True, timeout=timeout)
        raw = open(fout, "rb").read()
    cnt, secs, lanes, thr = struct.unpack_from("<qdii", raw, 0)
    off = 8 + 8 + 4 + 4
    scores = np.frombuffer(raw, dtype=np.int32, count=cnt, offset=off).copy()
    off += 4 * cnt
    out = {"scores": scores, "best_seconds": secs, "simd_lanes": lanes, "threads": thr,
           "binary": os.path.basename(exe)}
    if op in (OP_COUNTS_SIMD, OP_COUNTS_SCALAR, OP_ROWS_BANDED):
        alpha = 4 if alphabet == "dna" else 27
        S = alpha * alpha + 2
        out["counts"] = np.frombuffer(raw, dtype=np.int32, count=cnt * S, offset=off).reshape(cnt, S).copy()
        off += 4 * cnt * S
        out["op_off"] = np.frombuffer(raw, dtype=np.int64, count=cnt + 1, offset=off).copy()
        off += 8 * (cnt + 1)
        out["ops"] = np.frombuffer(raw, dtype=np.uint8, count=int(out["op_off"][-1]), offset=off).copy()
        off += int(out["op_off"][-1])
        if mode & 16:
            out["begins"] = np.frombuffer(raw, dtype=np.int64, count=2 * cnt, offset=off).reshape(cnt, 2).copy()
    return out


def ref_blosum62():
    exe = ref_binary()
    txt = subprocess.run([exe, "--blosum62"], check=True, capture_output=True, text=True).stdout
    return np.array([[int(x) for x in line.split()] for line in txt.strip().splitlines()],
                    dtype=np.int32)
