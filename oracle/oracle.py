"""ctypes bindings for the C oracle in ``oracle.c`` (test infrastructure only; see ``oracle/__init__.py``).

Sequences are passed as ``bytes``. Costs use the lanes' terms (reference lanes.mojo:187-259): a deletion, a run
of reference letters alone, of ``k`` letters costs ``del_open + k * del_ext``.
"""

from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass
from typing import Optional, Sequence, Tuple

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

NO_BAND_LOW = -(1 << 60)
NO_BAND_HIGH = 1 << 60
OUTSIDE = -5


def build() -> None:
    """Compiles the oracle (and the CPU baseline) with the Makefile beside this file."""
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)


def _load() -> ctypes.CDLL:
    if not os.path.exists(LIB_PATH):
        build()
    lib = ctypes.CDLL(LIB_PATH)
    i64, p8, pc = ctypes.c_int64, ctypes.c_char_p, ctypes.c_char_p
    pi64 = ctypes.POINTER(ctypes.c_int64)
    lib.orc_edit_distance.argtypes = [p8, i64, p8, i64]
    lib.orc_edit_distance.restype = i64
    lib.orc_myers_distance.argtypes = [p8, i64, p8, i64]
    lib.orc_myers_distance.restype = i64
    lib.orc_costs_cost.argtypes = [p8, i64, p8, i64, ctypes.c_void_p, i64, i64]
    lib.orc_costs_cost.restype = i64
    lib.orc_costs_align.argtypes = [p8, i64, p8, i64, ctypes.c_void_p, i64, i64, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_void_p, i64]
    lib.orc_costs_align.restype = i64
    lib.orc_off_band.argtypes = [ctypes.c_void_p, i64, i64]
    lib.orc_off_band.restype = i64
    lib.orc_proof.argtypes = [ctypes.c_void_p, i64, i64, i64, i64, i64, ctypes.c_int, pi64, pi64]
    lib.orc_proof.restype = ctypes.c_int
    lib.orc_scoring_global_score.argtypes = [p8, i64, p8, i64, ctypes.c_void_p]
    lib.orc_scoring_global_score.restype = i64
    lib.orc_scoring_global_align.argtypes = [p8, i64, p8, i64, ctypes.c_void_p, p8, p8, ctypes.c_int,
                                             ctypes.c_void_p, i64]
    lib.orc_scoring_global_align.restype = i64
    lib.orc_scoring_local_end.argtypes = [p8, i64, p8, i64, ctypes.c_void_p, pi64, pi64]
    lib.orc_scoring_local_end.restype = i64
    lib.orc_scoring_reach_back.argtypes = [p8, p8, i64, i64, ctypes.c_void_p, i64, pi64, pi64]
    lib.orc_scoring_reach_back.restype = None
    lib.orc_scoring_local_align.argtypes = [p8, i64, p8, i64, ctypes.c_void_p, p8, p8, ctypes.c_int,
                                            ctypes.c_void_p, i64, pi64]
    lib.orc_scoring_local_align.restype = i64
    lib.orc_costs_distance_proven.argtypes = [p8, i64, p8, i64, ctypes.c_void_p]
    lib.orc_costs_distance_proven.restype = i64
    lib.orc_costs_align_proven.argtypes = [p8, i64, p8, i64, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                           ctypes.c_void_p, i64, pi64]
    lib.orc_costs_align_proven.restype = i64
    lib.orc_cigar_cost.argtypes = [p8, i64, p8, i64, ctypes.c_void_p, p8, i64]
    lib.orc_cigar_cost.restype = i64
    vp = ctypes.c_void_p
    lib.orc_check_alignments.argtypes = [i64, vp, vp, vp, vp, vp, vp, vp, vp, pi64]
    lib.orc_check_alignments.restype = i64
    lib.orc_ends_cost.argtypes = [p8, i64, p8, i64, vp, vp, i64, i64, pi64, pi64]
    lib.orc_ends_cost.restype = i64
    lib.orc_ends_align.argtypes = [p8, i64, p8, i64, vp, vp, i64, i64, ctypes.c_int, ctypes.c_int, vp, i64, vp]
    lib.orc_ends_align.restype = i64
    del pc
    return lib


_LIB: Optional[ctypes.CDLL] = None


def lib() -> ctypes.CDLL:
    global _LIB
    if _LIB is None:
        _LIB = _load()
    return _LIB


class _Costs(ctypes.Structure):
    _fields_ = [(name, ctypes.c_int64) for name in (
        "mismatch", "del_open", "del_ext", "ins_open", "ins_ext", "pieces",
        "del_open2", "del_ext2", "ins_open2", "ins_ext2")]


@dataclass(frozen=True)
class Costs:
    """Gap costs in the lanes' terms; ``affine(x, o, e)`` is the library's ``Costs.affine``."""

    mismatch: int
    del_open: int
    del_ext: int
    ins_open: int
    ins_ext: int
    pieces: int = 1
    del_open2: int = 0
    del_ext2: int = 0
    ins_open2: int = 0
    ins_ext2: int = 0

    @staticmethod
    def affine(mismatch: int, opening: int, extension: int) -> "Costs":
        return Costs(mismatch, opening, extension, opening, extension)

    @staticmethod
    def edit() -> "Costs":
        return Costs(1, 0, 1, 0, 1)

    @staticmethod
    def two_piece(mismatch: int, opening: int, extension: int, opening2: int, extension2: int) -> "Costs":
        return Costs(mismatch, opening, extension, opening, extension, 2, opening2, extension2, opening2,
                     extension2)

    def with_deletions(self, opening: int, extension: int) -> "Costs":
        return Costs(self.mismatch, opening, extension, self.ins_open, self.ins_ext, self.pieces,
                     self.del_open2, self.del_ext2, self.ins_open2, self.ins_ext2)

    def _c(self) -> _Costs:
        return _Costs(self.mismatch, self.del_open, self.del_ext, self.ins_open, self.ins_ext, self.pieces,
                      self.del_open2, self.del_ext2, self.ins_open2, self.ins_ext2)


def _b(text) -> bytes:
    return text.encode("ascii") if isinstance(text, str) else bytes(text)


def edit_distance(ref, qry) -> int:
    a, b = _b(ref), _b(qry)
    return lib().orc_edit_distance(a, len(a), b, len(b))


def myers_distance(ref, qry) -> int:
    a, b = _b(ref), _b(qry)
    return lib().orc_myers_distance(a, len(a), b, len(b))


def costs_cost(ref, qry, costs: Costs, band: Tuple[int, int] = (NO_BAND_LOW, NO_BAND_HIGH)) -> int:
    """The least cost inside the band (library terms), or ``OUTSIDE``."""
    a, b = _b(ref), _b(qry)
    c = costs._c()
    return lib().orc_costs_cost(a, len(a), b, len(b), ctypes.byref(c), band[0], band[1])


def costs_align(ref, qry, costs: Costs, band: Tuple[int, int] = (NO_BAND_LOW, NO_BAND_HIGH),
                right_ties: bool = False, eqx: bool = True) -> Tuple[int, Optional[str]]:
    """(cost, CIGAR) by the reference's Ties rule; (``OUTSIDE``, None) with no path inside the band."""
    a, b = _b(ref), _b(qry)
    c = costs._c()
    cap = 2 * (len(a) + len(b)) + 2
    buf = ctypes.create_string_buffer(cap)
    cost = lib().orc_costs_align(a, len(a), b, len(b), ctypes.byref(c), band[0], band[1], int(right_ties),
                                 int(eqx), buf, cap)
    if cost < 0:
        return cost, None
    return cost, buf.value.decode("ascii")


def costs_distance_proven(ref, qry, costs: Costs) -> int:
    """``costs_cost`` of a global unbanded pair in memory proportional to a band (long pairs)."""
    a, b = _b(ref), _b(qry)
    c = costs._c()
    return lib().orc_costs_distance_proven(a, len(a), b, len(b), ctypes.byref(c))


def costs_align_proven(ref, qry, costs: Costs, right_ties: bool = False, eqx: bool = True,
                       with_cells: bool = False):
    """``costs_align`` of a global unbanded pair, tracing the strict proven band alone (long pairs). With
    ``with_cells``, also the stored band's cells."""
    a, b = _b(ref), _b(qry)
    c = costs._c()
    cap = 2 * (len(a) + len(b)) + 2
    buf = ctypes.create_string_buffer(cap)
    cells = ctypes.c_int64(0)
    cost = lib().orc_costs_align_proven(a, len(a), b, len(b), ctypes.byref(c), int(right_ties), int(eqx), buf, cap,
                                        ctypes.byref(cells))
    found = (cost, buf.value.decode("ascii") if cost >= 0 else None)
    return (*found, cells.value) if with_cells else found


def cigar_cost(ref, qry, costs: Costs, cigar) -> int:
    """What `cigar` costs if it spells an alignment of the whole pair, else -1 (malformed), -2 (wrong letter
    counts) or -3 (an `=` or `X` that the letters contradict)."""
    a, b = _b(ref), _b(qry)
    text = _b(cigar)
    c = costs._c()
    return lib().orc_cigar_cost(a, len(a), b, len(b), ctypes.byref(c), text, len(text))


def check_alignments(refs: Sequence[bytes], qrys: Sequence[bytes], costs: Costs, cigars: Sequence[bytes],
                     claimed: Sequence[int]) -> Tuple[int, int]:
    """(failures, first failing index or -1): each CIGAR must spell its pair and price at its claimed cost."""
    count = len(refs)
    keep = [_b(r) for r in refs], [_b(q) for q in qrys], [_b(x) for x in cigars]
    arrays = []
    for items in keep:
        ptrs = (ctypes.c_char_p * max(count, 1))(*items)
        lens = (ctypes.c_int64 * max(count, 1))(*[len(x) for x in items])
        arrays.append((ptrs, lens))
    claim = (ctypes.c_int64 * max(count, 1))(*[int(v) for v in claimed])
    c = costs._c()
    first = ctypes.c_int64(-1)
    vp = ctypes.c_void_p
    bad = lib().orc_check_alignments(count, ctypes.cast(arrays[0][0], vp), ctypes.cast(arrays[0][1], vp),
                                     ctypes.cast(arrays[1][0], vp), ctypes.cast(arrays[1][1], vp), ctypes.byref(c),
                                     ctypes.cast(arrays[2][0], vp), ctypes.cast(arrays[2][1], vp),
                                     ctypes.cast(claim, vp), ctypes.byref(first))
    return int(bad), int(first.value)


def off_band(costs: Costs, diagonal: int, end: int) -> int:
    c = costs._c()
    return lib().orc_off_band(ctypes.byref(c), diagonal, end)


def proof(costs: Costs, rows: int, columns: int, found: int, low: int, high: int, strict: bool):
    """(proven, wide_low, wide_high) of the lanes' band certificate in lane diagonals."""
    c = costs._c()
    lo, hi = ctypes.c_int64(), ctypes.c_int64()
    ok = lib().orc_proof(ctypes.byref(c), rows, columns, found, low, high, int(strict), ctypes.byref(lo),
                         ctypes.byref(hi))
    return bool(ok), lo.value, hi.value


class _Scoring(ctypes.Structure):
    _fields_ = [("table", ctypes.c_void_p), ("letters", ctypes.c_int64), ("open", ctypes.c_int64),
                ("extend", ctypes.c_int64)]


@dataclass(frozen=True)
class Scoring:
    """A substitution table over ``alphabet`` (row = reference letter) and a gap of k letters scoring
    ``opening + k * extension``, as the library's ``Scoring``."""

    alphabet: str
    table: Tuple[int, ...]
    opening: int
    extension: int

    def codes(self, text) -> bytes:
        index = {ord(ch): i for i, ch in enumerate(self.alphabet)}
        return bytes(index[ch] for ch in _b(text))

    def _c(self):
        raw = (ctypes.c_int8 * len(self.table))(*self.table)
        s = _Scoring(ctypes.cast(raw, ctypes.c_void_p), len(self.alphabet), self.opening + self.extension,
                     self.extension)
        return s, raw


def scoring_global_score(ref, qry, scoring: Scoring) -> int:
    a, b = scoring.codes(ref), scoring.codes(qry)
    s, keep = scoring._c()
    return lib().orc_scoring_global_score(a, len(a), b, len(b), ctypes.byref(s))


def scoring_global_align(ref, qry, scoring: Scoring, eqx: bool = True) -> Tuple[int, str]:
    a, b = scoring.codes(ref), scoring.codes(qry)
    s, keep = scoring._c()
    cap = 2 * (len(a) + len(b)) + 2
    buf = ctypes.create_string_buffer(cap)
    score = lib().orc_scoring_global_align(a, len(a), b, len(b), ctypes.byref(s), _b(ref), _b(qry), int(eqx), buf,
                                           cap)
    return score, buf.value.decode("ascii")


def scoring_local_end(ref, qry, scoring: Scoring) -> Tuple[int, int, int]:
    a, b = scoring.codes(ref), scoring.codes(qry)
    s, keep = scoring._c()
    er, ec = ctypes.c_int64(), ctypes.c_int64()
    best = lib().orc_scoring_local_end(a, len(a), b, len(b), ctypes.byref(s), ctypes.byref(er), ctypes.byref(ec))
    return best, er.value, ec.value


def scoring_local_align(ref, qry, scoring: Scoring, eqx: bool = True):
    """(score, CIGAR, (ref_start, ref_end, query_start, query_end)) as the host's local path gives them."""
    a, b = scoring.codes(ref), scoring.codes(qry)
    s, keep = scoring._c()
    cap = 2 * (len(a) + len(b)) + 2
    buf = ctypes.create_string_buffer(cap)
    spans = (ctypes.c_int64 * 4)()
    score = lib().orc_scoring_local_align(a, len(a), b, len(b), ctypes.byref(s), _b(ref), _b(qry), int(eqx), buf,
                                          cap, spans)
    return score, buf.value.decode("ascii"), tuple(spans)


def uniform_scoring(match: int, mismatch: int, opening: int = -4, extension: int = -2,
                    alphabet: str = "ACGT") -> Scoring:
    size = len(alphabet)
    table = tuple(match if r == c else mismatch for r in range(size) for c in range(size))
    return Scoring(alphabet, table, opening, extension)


def rows_cost(top: str, bottom: str, costs: Costs) -> int:
    """What two gapped rows cost under ``costs``, each gap run at its cheaper piece (a rescorer sharing no
    code with the sweeps)."""
    total, k = 0, 0
    run = None
    length = 0

    def gap(kind, letters):
        if kind == "D":
            cost = costs.del_open + letters * costs.del_ext
            if costs.pieces == 2:
                cost = min(cost, costs.del_open2 + letters * costs.del_ext2)
        else:
            cost = costs.ins_open + letters * costs.ins_ext
            if costs.pieces == 2:
                cost = min(cost, costs.ins_open2 + letters * costs.ins_ext2)
        return cost

    for a, b in zip(top, bottom):
        kind = "D" if b == "-" else ("I" if a == "-" else None)
        if kind != run and run is not None:
            total += gap(run, length)
            length = 0
        run = kind
        if kind is None:
            total += 0 if a == b else costs.mismatch
        else:
            length += 1
        k += 1
    if run is not None:
        total += gap(run, length)
    return total


def _ends(ends, n, m):
    """Free letters {reference start, end, query start, end}, none past its sequence (EndsFree.of)."""
    return (ctypes.c_int64 * 4)(min(ends[0], n), min(ends[1], n), min(ends[2], m), min(ends[3], m))


def ends_cost(ref, qry, costs: Costs, ends, band: Tuple[int, int] = (NO_BAND_LOW, NO_BAND_HIGH)):
    """(cost or OUTSIDE, end row, end column) under free ends `ends` = (reference start, end, query start, end)."""
    a, b = _b(ref), _b(qry)
    c = costs._c()
    er, ec = ctypes.c_int64(0), ctypes.c_int64(0)
    cost = lib().orc_ends_cost(a, len(a), b, len(b), ctypes.byref(c), _ends(ends, len(a), len(b)), band[0], band[1],
                               ctypes.byref(er), ctypes.byref(ec))
    return cost, er.value, ec.value


def ends_align(ref, qry, costs: Costs, ends, band: Tuple[int, int] = (NO_BAND_LOW, NO_BAND_HIGH),
               right_ties: bool = False, eqx: bool = True):
    """(cost, CIGAR, (reference start, end, query start, end)) the tie rule names under free ends; cost OUTSIDE
    (and no CIGAR) when no alignment stays in the band."""
    a, b = _b(ref), _b(qry)
    c = costs._c()
    cap = 2 * (len(a) + len(b)) + 2
    buf = ctypes.create_string_buffer(cap)
    spans = (ctypes.c_int64 * 4)()
    cost = lib().orc_ends_align(a, len(a), b, len(b), ctypes.byref(c), _ends(ends, len(a), len(b)), band[0],
                                band[1], int(right_ties), int(eqx), buf, cap, spans)
    if cost < 0:
        return cost, None, None
    return cost, buf.value.decode("ascii"), tuple(spans)
