"""The data formats on either side of a batch: FASTA/FASTQ records in, SAM, PAF and tab-separated lines out.

Behaviour follows the reference's command line (reference cli/dinara_align_cli.mojo): records are named by their
header up to the first blank, sequences are upper-cased with whitespace dropped (:91-131); a SAM record carries the
soft-clipped CIGAR, NM, MD and AS tags, flag 16 on the reverse strand and flag 4 unmapped, with an @HD/@SQ/@PG
header (:463-545); a PAF line counts matches and alignment columns and gives the query's span on the forward strand
(:546-566). The command line itself is not part of this package; these are the pieces a caller of the batch
calls writes records with.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Dict, Iterable, List, Optional, Sequence, Tuple

from .types import Alignment, Text, as_bytes


@dataclass(frozen=True)
class Record:
    """A FASTA or FASTQ record: its name (the header up to the first blank) and its sequence."""

    name: str
    sequence: str


class FormatError(ValueError):
    """A file that is neither FASTA nor FASTQ, or a record cut short."""


def _name_of(header: str, path: str) -> str:
    words = header.split()
    if not words:
        raise FormatError(f"{path}: a record without a name")
    return words[0]


def parse_records(text: str, path: str = "<text>") -> List[Record]:
    """The records of FASTA or FASTQ text (the two may be mixed): sequences upper-cased, whitespace dropped; a
    FASTA sequence runs over every line up to the next `>` header, a FASTQ record is four lines."""
    lines = text.split("\n")
    out: List[Record] = []
    at = 0
    while at < len(lines):
        line = lines[at].strip()
        at += 1
        if not line:
            continue
        if line[0] == ">":
            name = _name_of(line[1:], path)
            parts = []
            while at < len(lines) and not lines[at].startswith(">"):
                parts.append("".join(lines[at].split()))
                at += 1
            out.append(Record(name, "".join(parts).upper()))
        elif line[0] == "@":
            name = _name_of(line[1:], path)
            if at + 2 >= len(lines):
                raise FormatError(f"{path}: the FASTQ record `{name}` is cut short")
            sequence, plus, quality = (lines[at + i].strip() for i in range(3))
            if not plus.startswith("+") or len(quality) != len(sequence):
                raise FormatError(f"{path}: the FASTQ record `{name}` lacks its `+` line or its qualities")
            at += 3
            out.append(Record(name, sequence.upper()))
        else:
            raise FormatError(f"{path}: neither FASTA nor FASTQ at `{line}`")
    return out


def read_records(path: str) -> List[Record]:
    """The records of a FASTA or FASTQ file."""
    with open(path, "r", encoding="ascii") as handle:
        return parse_records(handle.read(), path)


_COMPLEMENT = str.maketrans("ACGT", "TGCA")


def reverse_complement(sequence: str) -> str:
    """The other strand, read 5' to 3': A<->T, C<->G, any other letter kept as it is (reference
    cli/dinara_align_cli.mojo:211-226)."""
    return sequence.translate(_COMPLEMENT)[::-1]


@dataclass(frozen=True)
class Counts:
    """Alignment columns by kind: equal letters, unequal letters, reference letters alone, query letters alone."""

    matches: int
    mismatches: int
    deleted: int
    inserted: int

    @property
    def columns(self) -> int:
        return self.matches + self.mismatches + self.deleted + self.inserted


def counts(found: Alignment, reference: Text, query: Text) -> Counts:
    tally = {"match": 0, "mismatch": 0, "deleted": 0, "inserted": 0}
    for _, a, b in found.columns(reference, query):
        kind = "deleted" if b is None else "inserted" if a is None else "match" if a == b else "mismatch"
        tally[kind] += 1
    return Counts(tally["match"], tally["mismatch"], tally["deleted"], tally["inserted"])


def sam_header(references: Iterable[Tuple[str, int]], program: str = "dinara-align") -> List[str]:
    """@HD, one @SQ a reference name (first length kept; the same name at another length is refused, SAM cannot
    tell them apart), and @PG."""
    lines = ["@HD\tVN:1.6\tSO:unsorted"]
    lengths: Dict[str, int] = {}
    for name, length in references:
        if name in lengths:
            if lengths[name] != length:
                raise FormatError(f"two references named {name} of different lengths, which SAM cannot tell apart")
            continue
        lengths[name] = length
        lines.append(f"@SQ\tSN:{name}\tLN:{length}")
    lines.append(f"@PG\tID:{program}\tPN:{program}")
    return lines


def sam_record(query_name: str, reference_name: str, found: Optional[Alignment], reference: Text, query: Text,
               reverse: bool = False) -> str:
    """One SAM line. `query` is the sequence as aligned (the reverse complement when `reverse`). No alignment, or
    one of no letters, is an unmapped record (flag 4)."""
    q = as_bytes(query).decode("ascii")
    shown = q if q else "*"
    if found is None or not found.cigar:
        return "\t".join([query_name, "4", "*", "0", "0", "*", "*", "0", "0", shown, "*"])
    fields = [query_name, "16" if reverse else "0", reference_name, str(found.reference_start + 1), "255",
              found.clipped_cigar(len(q)), "*", "0", "0", shown, "*",
              f"NM:i:{found.edit_distance(reference, query)}", f"MD:Z:{found.mismatch_string(reference, query)}",
              f"AS:i:{found.score}"]
    return "\t".join(fields)


def paf_record(query_name: str, reference_name: str, found: Alignment, reference: Text, query: Text,
               reverse: bool = False) -> str:
    """One PAF line (no line for an unaligned pair). The query's span is on its forward strand."""
    length = len(as_bytes(query))
    tally = counts(found, reference, query)
    start, end = (length - found.query_end, length - found.query_start) if reverse else (found.query_start,
                                                                                        found.query_end)
    fields = [query_name, str(length), str(start), str(end), "-" if reverse else "+", reference_name,
              str(len(as_bytes(reference))), str(found.reference_start), str(found.reference_end), str(tally.matches),
              str(tally.columns), "255", f"NM:i:{found.edit_distance(reference, query)}", f"AS:i:{found.score}",
              f"cg:Z:{found.cigar}"]
    return "\t".join(fields)


TSV_HEADER = "query\treference\tstrand\tcost\tscore\treference_start\treference_end\tquery_start\tquery_end\tcigar"


def tsv_record(query_name: str, reference_name: str, found: Optional[Alignment], reverse: bool = False) -> str:
    if found is None:
        return "\t".join([query_name, reference_name] + ["*"] * 8)
    return "\t".join([query_name, reference_name, "-" if reverse else "+", str(found.cost), str(found.score),
                      str(found.reference_start), str(found.reference_end), str(found.query_start),
                      str(found.query_end), found.cigar])


def better(first: Alignment, second: Alignment, scored: bool) -> bool:
    """Whether `second` beats `first` when both strands are aligned: a higher score under a match score, else a
    lower cost; a tie keeps the forward strand."""
    return second.score > first.score if scored else second.cost < first.cost


def pick_strands(forward: Sequence[Optional[Alignment]], backward: Sequence[Optional[Alignment]],
                 scored: bool = False) -> List[Tuple[Optional[Alignment], bool]]:
    """Per pair, the better of its forward and reverse-complement alignments, and whether it is the reverse one."""
    out = []
    for f, r in zip(forward, backward):
        if r is not None and (f is None or better(f, r, scored)):
            out.append((r, True))
        else:
            out.append((f, False))
    return out
