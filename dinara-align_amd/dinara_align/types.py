"""The value types a caller passes and gets back: Costs, Mode, Band, Scoring, Placement and Alignment.

Their names, fields, defaults and constructors are the reference package's interface
(reference python/dinara_align/__init__.py:91-400), so code written against it builds the same values here.
How they check their arguments, what their refusals say and how an Alignment is read back are this package's
own: every check lives in one place (`_require`), each refusal names the argument at fault, and an
Alignment's helpers all walk the CIGAR through one iterator (`Alignment.columns`).
"""

from __future__ import annotations

import ctypes
import re
from dataclasses import dataclass, replace
from typing import Iterator, List, Optional, Sequence, Tuple, Union

from . import _lib

UNBOUNDED = 1 << 60
"""A count of free letters, or a band edge, that no sequence reaches: 2^60."""

Text = Union[str, bytes]

# dinara_mode's kinds (reference c/dinara.h:71-89).
KIND_ENDS_FREE, KIND_EXTENSION, KIND_LOCAL = 0, 1, 2


def _require(condition: bool, message: str) -> None:
    """Raises ValueError(message) unless `condition`."""
    if not condition:
        raise ValueError(message)


def _integers(**named) -> None:
    """TypeError for the first keyword whose value is not an int."""
    for name, value in named.items():
        if not isinstance(value, int):
            raise TypeError(f"{name} must be an int (got {value!r} of type {type(value).__name__})")


def as_bytes(sequence: Text) -> bytes:
    """A sequence as the bytes the C ABI takes. A str is encoded as ASCII; bytes pass through once they are
    known to be ASCII (positions count letters, so a letter is one byte). Anything else is a TypeError;
    a letter above 127 a ValueError."""
    if isinstance(sequence, str):
        _require(sequence.isascii(), "sequences must be ASCII: one byte per letter")
        return sequence.encode("ascii")
    if isinstance(sequence, (bytes, bytearray)):
        raw = bytes(sequence)
        _require(raw.isascii(), "sequences must be ASCII: one byte per letter")
        return raw
    raise TypeError(f"sequences are str or bytes; {type(sequence).__name__} is neither")


@dataclass(frozen=True)
class Costs:
    """Edit costs to minimize (reference python/dinara_align/__init__.py:91-156): a substitution costs
    `mismatch`; a run of k gapped letters `opening + k * extension`, or, with a second piece
    (`opening2` >= 0), the cheaper of the two lines. A negative `opening2` means one piece. Deletions (runs
    of reference letters) take the insertion prices unless `deletion_extension` is above zero."""

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
        """Levenshtein: every substitution and every gapped letter costs one."""
        return Costs()

    @staticmethod
    def linear(mismatch: int, gap: int) -> "Costs":
        """No opening cost: each gapped letter costs `gap`."""
        return Costs(mismatch=mismatch, opening=0, extension=gap)

    @staticmethod
    def affine(mismatch: int, opening: int, extension: int) -> "Costs":
        """One gap line: k letters cost `opening + k * extension`."""
        return Costs(mismatch=mismatch, opening=opening, extension=extension)

    @staticmethod
    def two_piece(mismatch: int, opening: int, extension: int, opening2: int, extension2: int) -> "Costs":
        """Two gap lines, a run paying the cheaper: the second usually opens dearer and extends cheaper."""
        _integers(opening2=opening2, extension2=extension2)
        _require(opening2 >= 0 and extension2 > 0,
                 f"two_piece: a second line needs opening2 >= 0 and extension2 > 0, not ({opening2}, {extension2})")
        return Costs(mismatch=mismatch, opening=opening, extension=extension, opening2=opening2,
                     extension2=extension2)

    def with_deletions(self, opening: int, extension: int, opening2: int = -1, extension2: int = 0) -> "Costs":
        """A copy in which runs of reference letters have prices of their own (insertions keep these)."""
        _integers(opening=opening, extension=extension, opening2=opening2, extension2=extension2)
        _require(opening >= 0 and extension > 0,
                 f"with_deletions: needs opening >= 0 and extension > 0, not ({opening}, {extension})")
        _require(opening2 < 0 or extension2 > 0,
                 f"with_deletions: a second deletion line needs extension2 > 0, not {extension2}")
        return replace(self, deletion_opening=opening, deletion_extension=extension, deletion_opening2=opening2,
                       deletion_extension2=extension2)

    def as_c(self) -> _lib.CCosts:
        """This value as a dinara_costs struct."""
        values = {name: getattr(self, name) for name, _ in _lib.CCosts._fields_}
        _integers(**values)
        return _lib.CCosts(**values)


class Anchor:
    """The end of both sequences an extension starts from."""

    START = 0
    END = 1


@dataclass(frozen=True)
class Mode:
    """Which alignments are allowed (reference python/dinara_align/__init__.py:166-255). This library's kernels
    serve, under Costs, `GLOBAL` and free ends (`ends_free`, `INFIX`, `PREFIX`, `SUFFIX`, no match score, no user
    band) and, under a Scoring, `GLOBAL` or `local()`; every other mode is refused by the library with
    DINARA_INVALID_MODE."""

    kind: int = KIND_ENDS_FREE
    reference_start: int = 0
    reference_end: int = 0
    query_start: int = 0
    query_end: int = 0
    match_score: int = 0
    anchor: int = Anchor.START
    zdrop: int = -1
    end_bonus: int = -1

    GLOBAL = None  # the presets are filled in below the class
    INFIX = None
    PREFIX = None
    SUFFIX = None

    @staticmethod
    def ends_free(reference_start: int = 0, reference_end: int = 0, query_start: int = 0,
                  query_end: int = 0) -> "Mode":
        """At most that many letters at each end of each sequence may stay unaligned at no cost."""
        return Mode(KIND_ENDS_FREE, reference_start, reference_end, query_start, query_end)

    def with_match_score(self, match_score: int) -> "Mode":
        """These ends, each match earning `match_score`."""
        return replace(self, match_score=match_score)

    @staticmethod
    def extension(match_score: int, anchor: int = Anchor.START, zdrop: Optional[int] = None,
                  end_bonus: Optional[int] = None) -> "Mode":
        """An alignment fixed at one end of both sequences and free at the other (seed extension)."""
        for name, value in (("zdrop", zdrop), ("end_bonus", end_bonus)):
            if value is not None:
                _integers(**{name: value})
                _require(value >= 0, f"extension: {name} must be zero or more, not {value}")
        return Mode(KIND_EXTENSION, match_score=match_score, anchor=anchor,
                    zdrop=-1 if zdrop is None else zdrop, end_bonus=-1 if end_bonus is None else end_bonus)

    @staticmethod
    def local(match_score: int = 0) -> "Mode":
        """Smith-Waterman: the best-scoring alignment of any substrings (under a Scoring, `local()`)."""
        return Mode(KIND_LOCAL, match_score=match_score)

    @staticmethod
    def overlap(match_score: int) -> "Mode":
        """All four end gaps free, matches rewarded: an overlap or a containment."""
        _integers(match_score=match_score)
        _require(match_score > 0, f"overlap: match_score must be above zero, not {match_score}")
        return Mode(KIND_ENDS_FREE, UNBOUNDED, UNBOUNDED, UNBOUNDED, UNBOUNDED, match_score)

    def as_c(self) -> _lib.CMode:
        """This value as a dinara_mode struct, free letters capped at UNBOUNDED."""
        values = {name: getattr(self, name) for name, _ in _lib.CMode._fields_}
        _integers(**values)
        for name in ("reference_start", "reference_end", "query_start", "query_end"):
            values[name] = min(values[name], UNBOUNDED)
        return _lib.CMode(**values)


Mode.GLOBAL = Mode()
Mode.INFIX = Mode.ends_free(reference_start=UNBOUNDED, reference_end=UNBOUNDED)
Mode.PREFIX = Mode.ends_free(reference_end=UNBOUNDED)
Mode.SUFFIX = Mode.ends_free(reference_start=UNBOUNDED)


@dataclass(frozen=True)
class Band:
    """Diagonals `low..high` (a cell's diagonal: reference letters less query letters) that every move of
    the alignment must stay on (reference python/dinara_align/__init__.py:258-269)."""

    low: int = -UNBOUNDED
    high: int = UNBOUNDED

    @staticmethod
    def around(width: int) -> "Band":
        """Within `width` diagonals of the main one, on both sides."""
        return Band(low=-width, high=width)


@dataclass(frozen=True)
class Scoring:
    """Scores to maximize (reference python/dinara_align/__init__.py:272-309): `table[a * len(alphabet) + b]`
    for reference letter `alphabet[a]` against query letter `alphabet[b]`; a gap of k letters scores
    `opening + k * extension`."""

    alphabet: str
    table: tuple
    opening: int = -4
    extension: int = -2

    @staticmethod
    def dna() -> "Scoring":
        """ACGT, +2 a match, -4 a mismatch, a gap of k letters -(4 + 2k)."""
        return Scoring.uniform(2, -4)

    @staticmethod
    def uniform(match: int, mismatch: int, opening: int = -4, extension: int = -2,
                alphabet: str = "ACGT") -> "Scoring":
        """`match` on the table's diagonal, `mismatch` everywhere else."""
        size = len(alphabet)
        cells = tuple(mismatch + (match - mismatch) * (i % (size + 1) == 0) for i in range(size * size))
        return Scoring(alphabet, cells, opening, extension)

    @staticmethod
    def tabulated(alphabet: str, table: Sequence[int], opening: int = -4, extension: int = -2) -> "Scoring":
        """A full table, row-major, one row per reference letter."""
        return Scoring(alphabet, tuple(table), opening, extension)

    def as_c(self) -> Tuple[_lib.CScoring, tuple]:
        """This value as a dinara_scoring struct, and the buffers it points into (keep them alive)."""
        _require(isinstance(self.alphabet, str) and self.alphabet.isascii(),
                 "Scoring: the alphabet must be an ASCII str")
        _integers(opening=self.opening, extension=self.extension)
        cells = list(self.table)
        for at, cell in enumerate(cells):
            _integers(**{f"table[{at}]": cell})
            _require(-128 <= cell <= 127, f"Scoring: table[{at}] = {cell} does not fit a signed byte")
        buffer = (ctypes.c_int8 * max(1, len(cells)))(*cells)
        letters = self.alphabet.encode("ascii")
        struct = _lib.CScoring(letters, len(letters), ctypes.cast(buffer, ctypes.c_void_p), self.opening,
                               self.extension)
        return struct, (buffer, letters)


@dataclass(frozen=True)
class Placement:
    """Where a batch runs, after the reference's Mojo `Placement.on_gpu(gpu_id, threads)` (reference
    common.mojo:130-160): GPU `gpu_id`, the host-side packing on `threads` threads (0: the caller's)."""

    gpu_id: int = 0
    threads: int = 0

    @staticmethod
    def on_gpu(gpu_id: int = 0, threads: int = 0) -> "Placement":
        return Placement(gpu_id, threads)


# CIGAR operations: how many reference and query letters one step of each consumes.
_STEP = {"=": (1, 1), "X": (1, 1), "M": (1, 1), "D": (1, 0), "I": (0, 1)}
_RUN = re.compile(r"(\d+)(\D)")


@dataclass(frozen=True)
class Alignment:
    """An optimal alignment of `reference[reference_start:reference_end]` with
    `query[query_start:query_end]` (reference python/dinara_align/__init__.py:312-400). The CIGAR spells it:
    `=` equal letters, `X` unequal ones (`M` either, without eqx), `D` a reference letter against nothing,
    `I` a query letter against nothing."""

    cost: int
    score: int
    cigar: str
    reference_start: int
    reference_end: int
    query_start: int
    query_end: int

    def runs(self) -> List[Tuple[int, str]]:
        """The CIGAR's runs, `(length, operation)` in order."""
        return [(int(count), op) for count, op in _RUN.findall(self.cigar)]

    def columns(self, reference: Text, query: Text) -> Iterator[Tuple[str, Optional[str], Optional[str]]]:
        """The alignment column by column: `(operation, reference letter or None, query letter or None)`."""
        ref = as_bytes(reference).decode("ascii")
        qry = as_bytes(query).decode("ascii")
        r, q = self.reference_start, self.query_start
        for count, op in self.runs():
            takes_ref, takes_qry = _STEP.get(op, (0, 0))
            for _ in range(count):
                yield op, (ref[r] if takes_ref else None), (qry[q] if takes_qry else None)
                r += takes_ref
                q += takes_qry

    def gapped(self, reference: Text, query: Text) -> Tuple[str, str]:
        """Both aligned spans written one above the other, `-` where a letter faces nothing."""
        pairs = [(a or "-", b or "-") for _, a, b in self.columns(reference, query)]
        return "".join(a for a, _ in pairs), "".join(b for _, b in pairs)

    def clipped_cigar(self, query_length: int, hard: bool = False) -> str:
        """The CIGAR a SAM record carries: the query's unaligned head and tail as soft (or hard) clips."""
        kind = "H" if hard else "S"
        head, tail = self.query_start, query_length - self.query_end
        return (f"{head}{kind}" if head > 0 else "") + self.cigar + (f"{tail}{kind}" if tail > 0 else "")

    def edit_distance(self, reference: Text, query: Text) -> int:
        """SAM's NM: columns that are not two equal letters."""
        return sum(a != b for _, a, b in self.columns(reference, query))

    def identity(self, reference: Text, query: Text) -> float:
        """Equal-letter columns over all columns (0.0 for an empty alignment)."""
        flags = [a == b for _, a, b in self.columns(reference, query)]
        return sum(flags) / len(flags) if flags else 0.0

    def mismatch_string(self, reference: Text, query: Text) -> str:
        """SAM's MD tag: counts of matching columns, separated by each mismatched reference letter and by `^`
        followed by each deleted run's letters; insertions do not appear. It starts and ends with a count."""
        tokens: List[str] = []
        matched = 0
        last = None
        for op, a, b in self.columns(reference, query):
            if b is None:  # a reference letter against nothing
                if last != "del":
                    tokens.append(f"{matched}^")
                    matched = 0
                tokens.append(a)
                last = "del"
            elif a is None:  # an insertion closes a deletion run and counts nothing
                last = "ins"
            elif a == b:
                matched += 1
                last = "match"
            else:
                tokens.append(f"{matched}{a}")
                matched = 0
                last = "sub"
        tokens.append(str(matched))
        return "".join(tokens)
