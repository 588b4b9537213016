"""Alignment ingest (plk_align_*) against the Python restatement in oracle/align_oracle.py,
on the reference's own test alignments (tests/golden: turtle.fa, turtle_aa.fasta,
example.phy from test_scripts/test_data) and on hand-made alignments with known answers."""
import pathlib
import sys

import numpy as np
import pytest

import gems_amd as G

ROOT = pathlib.Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT / "oracle"))
import align_oracle as AO  # noqa: E402

GOLD = ROOT / "tests" / "golden"


def check_same(a, ref):
    assert a.names == ref["names"]
    assert a.seq_type == ref["seq_type"]
    assert np.array_equal(a.states, np.array(ref["states"], dtype=np.uint8))
    assert np.array_equal(a.freq, np.array(ref["freq"]))
    assert np.array_equal(a.site_pattern, np.array(ref["site_pattern"], dtype=np.int32))
    assert np.array_equal(a.flags, np.array(ref["flags"], dtype=np.int32))
    assert a.gaps_only == ref["gaps_only"]


def check_invariants(a, ref):
    # the coded alignment is recovered exactly from the patterns and the site map
    rows = np.array(ref["rows"], dtype=np.uint8)
    assert np.array_equal(a.states[:, a.site_pattern], rows)
    assert a.freq.sum() == a.nsites
    assert len({tuple(c) for c in a.states.T}) == a.nptn  # distinct
    first = [int(np.argmax(a.site_pattern == k)) for k in range(a.nptn)]
    assert first == sorted(first)  # first-occurrence order


@pytest.mark.parametrize("name,seq_type,ntaxa", [("turtle.fa", AO.SEQ_DNA, 12), ("turtle_aa.fasta", AO.SEQ_PROTEIN, 16),
                                                 ("example.phy", AO.SEQ_DNA, 44)])
def test_reference_alignments(name, seq_type, ntaxa):
    path = GOLD / name
    a = G.Alignment(path)
    ref = AO.read(path.read_text())
    assert a.ntaxa == ntaxa and a.seq_type == seq_type
    check_same(a, ref)
    check_invariants(a, ref)


def test_codon_reading_and_stop_codons():
    text = ">a\nATGTAAGGC\n>b\nATGTAGNNN\n>c\nATGCCRGGC\n>d\nATGTGAGG-\n"
    a = G.Alignment(text=text, seq_type=AO.SEQ_CODON)
    ref = AO.read(text, seq_type=AO.SEQ_CODON)
    check_same(a, ref)
    su = 61
    # site 1: ATG everywhere (Met); site 2: TAA/TAG/TGA are stops -> unknown, CCR ambiguous -> unknown
    assert a.nsites == 3 and a.states[0, a.site_pattern[1]] == su and a.states[2, a.site_pattern[1]] == su
    assert a.flags[a.site_pattern[0]] & 1  # constant


def test_phylip_interleaved_and_sequential():
    inter = "4 10\nt1 ACGTA\nt2 ACGTT\nt3 ACGAA\nt4 AC-TA\n\nCCGTA\nCCGTT\nCCGTA\nNCGTA\n"
    seq = "4 10\nt1 ACGTA\nCCGTA\nt2 ACGTTCCGTT\nt3 ACG\nAACCGTA\nt4 AC-TANCGTA\n"
    a, b = G.Alignment(text=inter, fmt=1), G.Alignment(text=seq, fmt=2)
    for x, t, f in ((a, inter, 1), (b, seq, 2)):
        ref = AO.read(t, fmt=f)
        check_same(x, ref)
    assert np.array_equal(a.states[:, a.site_pattern], b.states[:, b.site_pattern])


def test_ambiguity_codes_and_flags():
    text = ">a\nARGNT\n>b\nAGGNT\n>c\nCYGNT\n>d\nCTG-T\n"
    with pytest.raises(G.PlkError):  # too few proper nucleotides to detect the type
        G.Alignment(text=text)
    a = G.Alignment(text=text, seq_type=AO.SEQ_DNA)
    ref = AO.read(text, seq_type=AO.SEQ_DNA)
    check_same(a, ref)
    assert a.states[0, a.site_pattern[1]] == 8 and a.states[2, a.site_pattern[1]] == 13  # R, Y
    assert a.gaps_only == 1  # column 4: N N N -
    flags = a.flags[a.site_pattern]
    assert list(flags & 1) == [0, 0, 1, 1, 1]  # R/G/Y/T share no state; G, gaps, T are constant
    assert list(flags & 2) == [2, 0, 0, 0, 0]  # A,A,C,C: two states twice


def test_fasta_name_shortening():
    text = ">seq1 first sample\nACGTACGTAC\n>seq2 second\nACGTACGTAA\n>seq1 other\nACGTACGTTT\n"
    a = G.Alignment(text=text)
    assert a.names == AO.read(text)["names"] == ["seq1_first", "seq2_second", "seq1_other"]


@pytest.mark.parametrize("text,msg", [
    ("3 4\na ACGT\nb ACGT\na ACGT\n", "duplicated"),
    (">a\nACGT\n>b\nACG\n>c\nACGT\n", "not enough"),
    (">a\nACGT\n>b\nACJT\n>c\nACGT\n", "invalid character"),
    ("2 4\na ACGT\nb ACGT\n", "at least 3 sequences"),
    (">a\nAC%T\n>b\nACGT\n>c\nACGT\n", "Unrecognized character"),
])
def test_errors(text, msg):
    with pytest.raises(G.PlkError) as e:
        G.Alignment(text=text, seq_type=AO.SEQ_DNA)
    assert msg in str(e.value)
    with pytest.raises(AO.AlignmentError) as e2:
        AO.read(text, seq_type=AO.SEQ_DNA)
    assert msg in str(e2.value)
