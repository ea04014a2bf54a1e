"""The C ABI library loads and exports every entry point include/dinara.h declares, and the Python package's
host-side types behave as the reference's do. No GPU is needed: nothing here computes on a device."""

import ctypes
import os
import re

import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "dinara.h")
LIBRARY = os.path.join(ROOT, "dinara-align_amd", "dinara_align", "libdinara_hip.so")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(dinara_\w+)\s*\(", text)))


def test_header_declares_the_batch_calls():
    names = declared_functions()
    for name in ("dinara_distances", "dinara_alignments", "dinara_free", "dinara_gpu_distances",
                 "dinara_gpu_alignments", "dinara_gpu_scores", "dinara_gpu_scoring_alignments", "dinara_batch_new",
                 "dinara_batch_run"):
        assert name in names


@pytest.mark.skipif(not os.path.exists(LIBRARY), reason="libdinara_hip.so is not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIBRARY)
    missing = [name for name in declared_functions() if not hasattr(lib, name)]
    assert missing == []


def test_struct_layouts_match_the_header():
    from dinara_align import _lib
    # Every field is int64_t or a pointer: reference c/dinara.h:59-138.
    assert ctypes.sizeof(_lib.CCosts) == 9 * 8
    assert ctypes.sizeof(_lib.CMode) == 9 * 8
    assert ctypes.sizeof(_lib.COptions) == 6 * 8
    assert ctypes.sizeof(_lib.CAlignment) == 8 * 8
    assert ctypes.sizeof(_lib.CScoring) == 5 * 8


def test_alignment_helpers():
    import dinara_align as da
    found = da.Alignment(1, -1, "4=1X3=", 4, 12, 4, 12)
    assert found.gapped("GGGGACGTACGTGGGG", "CCCCACGTTCGTCC") == ("ACGTACGT", "ACGTTCGT")
    assert found.clipped_cigar(14) == "4S4=1X3=2S"
    assert found.edit_distance("GGGGACGTACGTGGGG", "CCCCACGTTCGTCC") == 1
    # The reference's own literal (tests/test_python.py:90).
    assert found.mismatch_string("GGGGACGTACGTGGGG", "CCCCACGTTCGTCC") == "4A3"
    assert found.identity("GGGGACGTACGTGGGG", "CCCCACGTTCGTCC") == 7 / 8
    assert found.runs() == [(4, "="), (1, "X"), (3, "=")]
    assert found.clipped_cigar(14, hard=True) == "4H4=1X3=2H"
    assert da.Band.around(3) == da.Band(-3, 3)
    assert da.Costs.affine(4, 6, 2) == da.Costs(4, 6, 2)
    with pytest.raises(ValueError):
        da.Costs.two_piece(4, 6, 2, -5, 1)
    with pytest.raises(ValueError):
        da.Costs.affine(4, 6, 2).with_deletions(6, 0)
    with pytest.raises(TypeError):
        da.types.as_bytes(5)
    with pytest.raises(ValueError):
        da.types.as_bytes("ACGTé")


def test_sam_mismatch_string_follows_the_sam_grammar():
    """MD by the SAM specification: [0-9]+(([A-Z]|\\^[A-Z]+)[0-9]+)*, insertions left out, a deleted run's
    letters after one `^`, a count (possibly 0) between every two tokens."""
    import dinara_align as da
    ref, qry = "ACGTTTGCAGGA", "ACTTGCCAGTA"
    found = da.Alignment(0, 0, "2=2D4=1I2=1X1=", 0, 12, 0, 11)
    assert found.gapped(ref, qry) == ("ACGTTTGC-AGGA", "AC--TTGCCAGTA")
    assert found.mismatch_string(ref, qry) == "2^GT6G1"
    assert found.edit_distance(ref, qry) == 4
    # A deletion straight after a substitution, and a substitution straight after a deletion.
    tail = da.Alignment(0, 0, "1X1D1X", 0, 3, 0, 2)
    assert tail.mismatch_string("ACG", "TT") == "0A0^C0G0"
    # An insertion between two deletions closes the first run.
    split = da.Alignment(0, 0, "1D1I1D", 0, 2, 0, 1)
    assert split.mismatch_string("AC", "G") == "0^A0^C0"
    assert da.Alignment(0, 0, "", 0, 0, 0, 0).mismatch_string("", "") == "0"
    # The reference's literals (tests/test_alignment.mojo:2794-2797).
    gapped = da.Alignment(3, -3, "3M2D4M1I2M", 0, 11, 0, 10)
    assert gapped.clipped_cigar(10) == "3M2D4M1I2M"
    assert gapped.mismatch_string("ACGTTTGCAAC", "ACGTGCATAC") == "3^TT6"
    assert da.Alignment(0, 0, "", 0, 0, 0, 0).identity("", "") == 0.0


def test_synthetic_generators_are_seeded():
    from dinara_align import synthetic
    a = synthetic.affine_reads(50, seed=9)
    b = synthetic.affine_reads(50, seed=9)
    assert a == b
    assert all(len(r) == 250 for r in a[0])
    refs, qrys = synthetic.local_protein(4)
    assert all(len(r) == 3000 and len(q) == 300 for r, q in zip(refs, qrys))
