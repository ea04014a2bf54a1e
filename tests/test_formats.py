"""FASTA/FASTQ records in, SAM, PAF and tab-separated lines out (dinara_align.formats), checked against the lines
the reference's command-line test expects (reference tests/test_cli.sh:14-35) for the same alignments. No GPU."""

import pytest

from dinara_align import formats
from dinara_align.types import Alignment

REFERENCE = "TTTTTTTTTTACGTACGTTTGCAGGATCCAGTTTTTTTTTT"  # reference tests/test_cli.sh:14
READS = "@r1\nACGTACGTTTGCAGGATCCAG\n+\nIIIIIIIIIIIIIIIIIIIII\n@r2\nCTGGATCCTGCAAACGTACGT\n+\nIIIIIIIIIIIIIIIIIIIII\n"


def test_records_fasta_and_fastq():
    refs = formats.parse_records(f">chr some words\n{REFERENCE[:20]}\n{REFERENCE[20:].lower()}\n\n")
    assert refs == [formats.Record("chr", REFERENCE)]
    reads = formats.parse_records(READS)
    assert [r.name for r in reads] == ["r1", "r2"]
    assert reads[1].sequence == "CTGGATCCTGCAAACGTACGT"
    # A FASTA sequence runs to the next `>` line, as the reference reads it: a FASTQ record after it is letters.
    mixed = formats.parse_records("@b x\nGG\n+\nII\n>a\nAC GT\n@c\n>d\n")
    assert mixed == [formats.Record("b", "GG"), formats.Record("a", "ACGT@C"), formats.Record("d", "")]


@pytest.mark.parametrize("text", ["@r\nACGT\n+\nIII\n", "@r\nACGT\nIIII\nIIII\n", "@r\nACGT\n", "ACGT\n", ">\nAC\n"])
def test_records_refused(text):
    with pytest.raises(formats.FormatError):
        formats.parse_records(text)


def test_reverse_complement():
    assert formats.reverse_complement("ACGTACGTTTGCAGGATCCAG") == "CTGGATCCTGCAAACGTACGT"  # r2 is r1's other strand
    assert formats.reverse_complement("ANCx") == "xGNT"


def test_tsv_line_of_a_pair_given_whole():
    found = Alignment(cost=2, score=-2, cigar="4=1D2=1I6=", reference_start=0, reference_end=13, query_start=0,
                      query_end=13)
    assert formats.tsv_record("query", "reference", found) == "query\treference\t+\t2\t-2\t0\t13\t0\t13\t4=1D2=1I6="
    assert formats.tsv_record("r2", "chr", None) == "r2\tchr\t*\t*\t*\t*\t*\t*\t*\t*"


def test_sam_record_on_the_reverse_strand():
    read = formats.reverse_complement("CTGGATCCTGCAAACGTACGT")
    found = Alignment(cost=0, score=0, cigar="21=", reference_start=10, reference_end=31, query_start=0,
                      query_end=21)
    line = formats.sam_record("r2", "chr", found, REFERENCE, read, reverse=True).split("\t")
    assert "\t".join(line[:6] + line[11:13]) == "r2\t16\tchr\t11\t255\t21=\tNM:i:0\tMD:Z:21"
    assert formats.sam_record("r3", "chr", None, REFERENCE, "").split("\t") == \
        ["r3", "4", "*", "0", "0", "*", "*", "0", "0", "*", "*"]
    assert formats.sam_header([("chr", 41), ("chr", 41)]) == ["@HD\tVN:1.6\tSO:unsorted", "@SQ\tSN:chr\tLN:41",
                                                            "@PG\tID:dinara-align\tPN:dinara-align"]
    with pytest.raises(formats.FormatError):
        formats.sam_header([("chr", 41), ("chr", 40)])


def test_sam_clips_and_tags():
    found = Alignment(cost=7, score=-7, cigar="2=1X1D2=1I", reference_start=3, reference_end=9, query_start=1,
                      query_end=7)
    line = formats.sam_record("q", "chr", found, "GGGACTTCAGG", "TACGCATGG").split("\t")
    assert line[3] == "4" and line[5] == "1S2=1X1D2=1I2S"
    assert line[11:] == ["NM:i:3", "MD:Z:2T0^T2", "AS:i:-7"]


def test_paf_record_of_a_local_alignment():
    found = Alignment(cost=0, score=42, cigar="21=", reference_start=10, reference_end=31, query_start=0,
                      query_end=21)
    line = formats.paf_record("r1", "chr", found, REFERENCE, "ACGTACGTTTGCAGGATCCAG").split("\t")
    assert "\t".join(line[:12]) == "r1\t21\t0\t21\t+\tchr\t41\t10\t31\t21\t21\t255"
    # On the reverse strand the query's span is counted on its forward strand.
    back = Alignment(cost=0, score=0, cigar="3=", reference_start=0, reference_end=3, query_start=1, query_end=4)
    assert formats.paf_record("q", "r", back, "ACG", "TACGTT", reverse=True).split("\t")[2:5] == ["2", "5", "-"]


def test_pick_strands():
    f = Alignment(3, -3, "3=", 0, 3, 0, 3)
    r = Alignment(1, -1, "3=", 0, 3, 0, 3)
    assert formats.pick_strands([f, None, f, None], [r, r, None, None]) == [(r, True), (r, True), (f, False),
                                                                            (None, False)]
    assert formats.pick_strands([f], [Alignment(3, -3, "3=", 0, 3, 0, 3)]) == [(f, False)]
