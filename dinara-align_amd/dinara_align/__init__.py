"""dinara_align on an AMD Instinct MI355X: the reference package's batch calls, run by hand-written HIP kernels.

A drop-in for the reference's ``dinara_align`` on its batched device path (reference
python/dinara_align/__init__.py): the same ``Costs``, ``Mode``, ``Band``, ``Scoring`` and ``Alignment`` types, the
same ``distance``/``align``/``score``/``distances``/``alignments`` calls and their ``ValueError`` refusals, and
the Mojo API's batch ``scores`` and device ``placement`` (reference api.mojo:843-1001,1216-1252), all reaching
``libdinara_hip.so`` through ctypes. Every call runs on the GPU; there is no CPU path, and a missing library
raises.

>>> import dinara_align as da
>>> da.distance("ACGTACGTTTGCA", "ACGTCGTTTTGCA")
2
>>> found = da.align("ACGTACGTTTGCA", "ACGTCGTTTTGCA", da.Costs.affine(4, 6, 2))
>>> found.cost, found.cigar
(12, '4=3X6=')
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass, replace
from typing import List, Optional, Sequence, Union

from . import _lib

__all__ = [
    "Alignment",
    "Anchor",
    "Band",
    "Costs",
    "Mode",
    "Placement",
    "Scoring",
    "align",
    "alignments",
    "distance",
    "distances",
    "score",
    "scores",
    "device_count",
]

UNBOUNDED = 1 << 60
"""As many free letters as any sequence has, and a band edge past every diagonal."""

Text = Union[str, bytes]

_ENDS, _EXTENSION, _LOCAL = 0, 1, 2
_INT64_MIN, _INT64_MAX = -(1 << 63), (1 << 63) - 1


def _text(sequence: Text) -> bytes:
    """`sequence` as ASCII bytes: refused unless a `str` or `bytes` (TypeError) and unless ASCII (ValueError)."""
    if isinstance(sequence, bytes):
        sequence.decode("ascii")
        return sequence
    if not isinstance(sequence, str):
        raise TypeError(f"a sequence is a str or bytes, not {type(sequence).__name__}")
    if not sequence.isascii():
        raise ValueError("a sequence holds ASCII letters alone")
    return sequence.encode("ascii")


def _ints(*values) -> tuple:
    for value in values:
        if not isinstance(value, int):
            raise TypeError(f"an integer, not {type(value).__name__}")
    return values


def _raise(code: int):
    raise ValueError(_lib.MESSAGES.get(code, f"failed with code {code}"))


@dataclass(frozen=True)
class Costs:
    """What each edit costs: a substitution `mismatch`, a gap of `k` letters `opening + k * extension`
    (reference python/dinara_align/__init__.py:91-156). Deletions may cost their own (`with_deletions`)."""

    mismatch: int = 1
    opening: int = 0
    extension: int = 1
    opening2: int = -1
    extension2: int = 0
    deletion_opening: int = 0
    deletion_extension: int = 0
    deletion_opening2: int = -1
    deletion_extension2: int = 0

    @staticmethod
    def edit() -> "Costs":
        return Costs()

    @staticmethod
    def linear(mismatch: int, gap: int) -> "Costs":
        return Costs(mismatch, 0, gap)

    @staticmethod
    def affine(mismatch: int, opening: int, extension: int) -> "Costs":
        return Costs(mismatch, opening, extension)

    @staticmethod
    def two_piece(mismatch: int, opening: int, extension: int, opening2: int, extension2: int) -> "Costs":
        if _ints(extension2)[0] <= 0 or _ints(opening2)[0] < 0:
            raise ValueError(f"second piece {opening2}, {extension2}: its extension must cost")
        return Costs(mismatch, opening, extension, opening2, extension2)

    def with_deletions(self, opening: int, extension: int, opening2: int = -1, extension2: int = 0) -> "Costs":
        if _ints(extension)[0] <= 0 or _ints(opening)[0] < 0 or (_ints(opening2)[0] >= 0 and _ints(extension2)[0] <= 0):
            raise ValueError(f"deletions {opening}, {extension}: an extension must cost")
        return replace(self, deletion_opening=opening, deletion_extension=extension, deletion_opening2=opening2,
                       deletion_extension2=extension2)

    def _c(self) -> _lib.CCosts:
        return _lib.CCosts(*_ints(self.mismatch, self.opening, self.extension, self.opening2, self.extension2,
                                  self.deletion_opening, self.deletion_extension, self.deletion_opening2,
                                  self.deletion_extension2))


class Anchor:
    START = 0
    END = 1


@dataclass(frozen=True)
class Mode:
    """Which alignments count (reference python/dinara_align/__init__.py:166-255). The device serves `GLOBAL`
    under `Costs`, and `GLOBAL` and `local()` under a `Scoring`."""

    kind: int = _ENDS
    reference_start: int = 0
    reference_end: int = 0
    query_start: int = 0
    query_end: int = 0
    match_score: int = 0
    anchor: int = Anchor.START
    zdrop: int = -1
    end_bonus: int = -1

    GLOBAL = None  # set below
    INFIX = None
    PREFIX = None
    SUFFIX = None

    @staticmethod
    def ends_free(reference_start: int = 0, reference_end: int = 0, query_start: int = 0,
                  query_end: int = 0) -> "Mode":
        return Mode(_ENDS, reference_start, reference_end, query_start, query_end)

    def with_match_score(self, match_score: int) -> "Mode":
        return replace(self, match_score=match_score)

    @staticmethod
    def extension(match_score: int, anchor: int = Anchor.START, zdrop: Optional[int] = None,
                  end_bonus: Optional[int] = None) -> "Mode":
        if zdrop is not None and _ints(zdrop)[0] < 0:
            raise ValueError("a Z-drop below zero")
        if end_bonus is not None and _ints(end_bonus)[0] < 0:
            raise ValueError("an end bonus below zero")
        return Mode(_EXTENSION, match_score=match_score, anchor=anchor, zdrop=-1 if zdrop is None else zdrop,
                    end_bonus=-1 if end_bonus is None else end_bonus)

    @staticmethod
    def local(match_score: int = 0) -> "Mode":
        return Mode(_LOCAL, match_score=match_score)

    @staticmethod
    def overlap(match_score: int) -> "Mode":
        if _ints(match_score)[0] <= 0:
            raise ValueError("an overlap needs a match that earns")
        return Mode(_ENDS, UNBOUNDED, UNBOUNDED, UNBOUNDED, UNBOUNDED, match_score)

    def _c(self) -> _lib.CMode:
        _ints(self.kind, self.reference_start, self.reference_end, self.query_start, self.query_end)
        _ints(self.match_score, self.anchor, self.zdrop, self.end_bonus)
        return _lib.CMode(self.kind, min(self.reference_start, UNBOUNDED), min(self.reference_end, UNBOUNDED),
                          min(self.query_start, UNBOUNDED), min(self.query_end, UNBOUNDED), self.match_score,
                          self.anchor, self.zdrop, self.end_bonus)


Mode.GLOBAL = Mode()
Mode.INFIX = Mode(_ENDS, UNBOUNDED, UNBOUNDED, 0, 0)
Mode.PREFIX = Mode(_ENDS, 0, UNBOUNDED, 0, 0)
Mode.SUFFIX = Mode(_ENDS, UNBOUNDED, 0, 0, 0)


@dataclass(frozen=True)
class Band:
    """The diagonals every move stays on, `low ..= high`, a cell's diagonal its reference letters less its
    query letters."""

    low: int = -UNBOUNDED
    high: int = UNBOUNDED

    @staticmethod
    def around(width: int) -> "Band":
        return Band(-width, width)


@dataclass(frozen=True)
class Scoring:
    """An alphabet's substitution table, row by row (row = the reference's letter), and a gap of `k` letters
    scoring `opening + k * extension`, both zero or less (reference python/dinara_align/__init__.py:272-309)."""

    alphabet: str
    table: tuple
    opening: int = -4
    extension: int = -2

    @staticmethod
    def dna() -> "Scoring":
        return Scoring.uniform(2, -4)

    @staticmethod
    def uniform(match: int, mismatch: int, opening: int = -4, extension: int = -2, alphabet: str = "ACGT") -> "Scoring":
        size = len(alphabet)
        table = tuple(match if row == column else mismatch for row in range(size) for column in range(size))
        return Scoring(alphabet, table, opening, extension)

    @staticmethod
    def tabulated(alphabet: str, table: Sequence[int], opening: int = -4, extension: int = -2) -> "Scoring":
        return Scoring(alphabet, tuple(table), opening, extension)

    def _c(self):
        if not isinstance(self.alphabet, str) or not self.alphabet.isascii():
            raise ValueError("an alphabet is a str of ASCII letters")
        cells = list(_ints(*self.table))
        if any(cell < -128 or cell > 127 for cell in cells):
            raise ValueError("a table's cells lie within -128 and 127")
        _ints(self.opening, self.extension)
        raw = (ctypes.c_int8 * max(len(cells), 1))(*cells)
        alphabet = self.alphabet.encode("ascii")
        c = _lib.CScoring(alphabet, len(alphabet), ctypes.cast(raw, ctypes.c_void_p), self.opening, self.extension)
        return c, (raw, alphabet)


@dataclass(frozen=True)
class Placement:
    """Where a batch runs, as the reference's Mojo `Placement.on_gpu(gpu_id, threads)` (common.mojo:130-160):
    GPU `gpu_id`, its host-side packing on `threads` threads (0: the caller's alone)."""

    gpu_id: int = 0
    threads: int = 0

    @staticmethod
    def on_gpu(gpu_id: int = 0, threads: int = 0) -> "Placement":
        return Placement(gpu_id, threads)


@dataclass(frozen=True)
class Alignment:
    """An optimal alignment of `reference[reference_start:reference_end]` against `query[query_start:query_end]`
    (reference python/dinara_align/__init__.py:312-400)."""

    cost: int
    score: int
    cigar: str
    reference_start: int
    reference_end: int
    query_start: int
    query_end: int

    def runs(self) -> list:
        out, length = [], 0
        for letter in self.cigar:
            if letter.isdigit():
                length = length * 10 + int(letter)
            else:
                out.append((length, letter))
                length = 0
        return out

    def gapped(self, reference: Text, query: Text) -> tuple:
        reference, query = _text(reference).decode(), _text(query).decode()
        top, bottom = [], []
        column, row = self.reference_start, self.query_start
        for length, operation in self.runs():
            if operation == "D":
                top.append(reference[column:column + length])
                bottom.append("-" * length)
                column += length
            elif operation == "I":
                top.append("-" * length)
                bottom.append(query[row:row + length])
                row += length
            else:
                top.append(reference[column:column + length])
                bottom.append(query[row:row + length])
                column += length
                row += length
        return "".join(top), "".join(bottom)

    def clipped_cigar(self, query_length: int, hard: bool = False) -> str:
        clip = "H" if hard else "S"
        out = f"{self.query_start}{clip}" if self.query_start > 0 else ""
        out += self.cigar
        if query_length > self.query_end:
            out += f"{query_length - self.query_end}{clip}"
        return out

    def edit_distance(self, reference: Text, query: Text) -> int:
        top, bottom = self.gapped(reference, query)
        return sum(1 for a, b in zip(top, bottom) if a != b)

    def identity(self, reference: Text, query: Text) -> float:
        top, bottom = self.gapped(reference, query)
        return sum(1 for a, b in zip(top, bottom) if a == b) / len(top) if top else 0.0


def _options(band: Optional[Band], max_cost: Optional[int], eqx: bool, ties: str) -> _lib.COptions:
    if ties not in ("left", "right"):
        raise ValueError("ties: 'left' or 'right'")
    band = band or Band()
    _ints(band.low, band.high)
    low = _INT64_MIN if band.low <= -UNBOUNDED else band.low
    high = _INT64_MAX if band.high >= UNBOUNDED else band.high
    cap = -1 if max_cost is None else _ints(max_cost)[0]
    return _lib.COptions(low, high, cap, 1 if eqx else 0, 1 if ties == "right" else 0, 0)


def _placement(placement: Optional[Placement], threads: int = 0) -> Placement:
    if placement is None:
        return Placement(0, _ints(threads)[0])
    return placement


def device_count() -> int:
    """How many GPUs the library sees."""
    return int(_lib.lib().dinara_device_count())


def _distance_codes(references, queries, costs: Costs, mode: Mode, options, placement: Placement) -> List[int]:
    refs, qrys = _lib.Texts([_text(r) for r in references]), _lib.Texts([_text(q) for q in queries])
    if len(refs.items) != len(qrys.items):
        raise ValueError("as many references as queries")
    count = len(refs.items)
    results = (ctypes.c_int64 * max(count, 1))()
    c, m = costs._c(), mode._c()
    code = _lib.lib().dinara_gpu_distances(placement.gpu_id, count, *refs.args(), *qrys.args(), ctypes.byref(c),
                                           ctypes.byref(m), ctypes.byref(options), placement.threads, results)
    if code != 0:
        _raise(code)
    return list(results[:count])


def distances(references: Sequence[Text], queries: Sequence[Text], costs: Costs = Costs(), mode: Mode = Mode.GLOBAL,
              *, band: Optional[Band] = None, max_cost: Optional[int] = None, threads: int = 0,
              placement: Optional[Placement] = None) -> list:
    """Every pair's least cost, None past `max_cost` (reference python/dinara_align/__init__.py:545-563)."""
    options = _options(band, max_cost, True, "left")
    codes = _distance_codes(references, queries, costs, mode, options, _placement(placement, threads))
    if max_cost is not None and max_cost < 0:
        # A negative cap holds no alignment: the call is still checked first (reference api.mojo:710-711).
        return [None] * len(codes)
    out = []
    for value in codes:
        if value >= 0:
            out.append(value)
        elif value == _lib.ABOVE_MAX and max_cost is not None:
            out.append(None)
        else:
            _raise(value)
    return out


def _alignments(references, queries, costs, mode, options, placement: Placement, scoring_eqx=None):
    refs, qrys = _lib.Texts([_text(r) for r in references]), _lib.Texts([_text(q) for q in queries])
    if len(refs.items) != len(qrys.items):
        raise ValueError("as many references as queries")
    count = len(refs.items)
    found = (_lib.CAlignment * max(count, 1))()
    statuses = (ctypes.c_int64 * max(count, 1))()
    handle = _lib.lib()
    m = mode._c()
    if isinstance(costs, Scoring):
        scoring, keep = costs._c()
        code = handle.dinara_gpu_scoring_alignments(placement.gpu_id, count, *refs.args(), *qrys.args(),
                                                    ctypes.byref(scoring), ctypes.byref(m), int(scoring_eqx),
                                                    placement.threads, found, statuses)
    else:
        c = costs._c()
        code = handle.dinara_gpu_alignments(placement.gpu_id, count, *refs.args(), *qrys.args(), ctypes.byref(c),
                                            ctypes.byref(m), ctypes.byref(options), placement.threads, found,
                                            statuses)
    if code != 0:
        _raise(code)
    out = []
    for i in range(count):
        status = statuses[i]
        if status != 0:
            out.append(status)
            continue
        item = found[i]
        cigar = ctypes.string_at(item.cigar, item.cigar_length).decode("ascii") if item.cigar else ""
        handle.dinara_free(item.cigar)
        out.append(Alignment(item.cost, item.score, cigar, item.reference_start, item.reference_end,
                             item.query_start, item.query_end))
    return out


def alignments(references: Sequence[Text], queries: Sequence[Text], costs: Union[Costs, Scoring] = Costs(),
               mode: Mode = Mode.GLOBAL, *, band: Optional[Band] = None, max_cost: Optional[int] = None,
               ties: str = "left", eqx: bool = True, max_memory: Optional[int] = None, threads: int = 0,
               placement: Optional[Placement] = None) -> list:
    """Every pair's optimal alignment, None past `max_cost` (reference python/dinara_align/__init__.py:566-588;
    under a Scoring, reference api.mojo:895-1001)."""
    if max_memory is not None:
        _ints(max_memory)
    options = _options(band, max_cost, eqx, ties)
    found = _alignments(references, queries, costs, mode, options, _placement(placement, threads), eqx)
    if max_cost is not None and max_cost < 0:
        return [None] * len(found)
    out = []
    for item in found:
        if isinstance(item, Alignment):
            out.append(item)
        elif item == _lib.ABOVE_MAX and max_cost is not None:
            out.append(None)
        else:
            _raise(item)
    return out


def scores(references: Sequence[Text], queries: Sequence[Text], scoring: Scoring, mode: Mode = Mode.GLOBAL, *,
           placement: Optional[Placement] = None) -> List[int]:
    """Every pair's best score under `scoring`, globally or locally (reference api.mojo:843-892)."""
    placement = _placement(placement)
    refs, qrys = _lib.Texts([_text(r) for r in references]), _lib.Texts([_text(q) for q in queries])
    if len(refs.items) != len(qrys.items):
        raise ValueError("as many references as queries")
    count = len(refs.items)
    results = (ctypes.c_int64 * max(count, 1))()
    c, keep = scoring._c()
    m = mode._c()
    code = _lib.lib().dinara_gpu_scores(placement.gpu_id, count, *refs.args(), *qrys.args(), ctypes.byref(c),
                                        ctypes.byref(m), placement.threads, results)
    if code != 0:
        _raise(code)
    return list(results[:count])


def distance(reference: Text, query: Text, costs: Costs = Costs(), mode: Mode = Mode.GLOBAL, *,
             band: Optional[Band] = None, max_cost: Optional[int] = None,
             placement: Optional[Placement] = None) -> Optional[int]:
    """The least cost of aligning `query` to `reference`, or None past `max_cost`."""
    return distances([reference], [query], costs, mode, band=band, max_cost=max_cost, placement=placement)[0]


def align(reference: Text, query: Text, costs: Union[Costs, Scoring] = Costs(), mode: Mode = Mode.GLOBAL, *,
          band: Optional[Band] = None, max_cost: Optional[int] = None, ties: str = "left", eqx: bool = True,
          max_memory: Optional[int] = None, placement: Optional[Placement] = None) -> Optional[Alignment]:
    """An optimal alignment, or None past `max_cost`."""
    return alignments([reference], [query], costs, mode, band=band, max_cost=max_cost, ties=ties, eqx=eqx,
                      max_memory=max_memory, placement=placement)[0]


def score(reference: Text, query: Text, costs: Union[Costs, Scoring] = Costs(), mode: Mode = Mode.GLOBAL, *,
          band: Optional[Band] = None, placement: Optional[Placement] = None) -> int:
    """The best score: under a Scoring its score, under Costs minus the least cost."""
    if isinstance(costs, Scoring):
        return scores([reference], [query], costs, mode, placement=placement)[0]
    return -distance(reference, query, costs, mode, band=band, placement=placement)
