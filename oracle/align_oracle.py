"""Alignment-ingest oracle: a plain-Python restatement of IQ-TREE's alignment reading and
site-pattern compression.  TEST INFRASTRUCTURE ONLY (imported by tests/ as the checker of
libplk's plk_align_*; the product never imports it).

Follows alignment/alignment.cpp:
  processSeq 2444-2472, doReadPhylip 2474-2556, doReadPhylipSequential 2567-2633,
  doReadFasta 2708-2808 (name shortening 2746-2790), detectSequenceType 1720-1786,
  buildStateMap 1788-1846, buildPattern 2219-2330 (codons 2273-2300),
  addPatternLazy 1511-1543, computeConst 1321-1420 (constant / informative flags).
Pinned by tests/test_align.py: the coded alignment is recovered exactly from
(patterns, site -> pattern map), frequencies sum to the site count, patterns are
distinct and in first-occurrence order, and hand-made alignments with known answers.
"""

PROTEIN = "ARNDCQEGHILKMFPSTWYV"
STANDARD_CODE = "KNKNTTTTRSRSIIMIQHQHPPPPRRRRLLLLEDEDAAAAGGGGVVVV*Y*YSSSS*CWCLFLF"
SEQ_DNA, SEQ_PROTEIN, SEQ_CODON = 0, 1, 2


class AlignmentError(ValueError):
    pass


def process_seq(line, line_num):
    out = []
    k = 0
    while k < len(line):
        ch = line[k]
        if ord(ch) <= 32:
            pass
        elif ch.isalnum() or ch in "-?.*~":
            out.append(ch.upper())
        elif ch == "!":
            out.append(ch)
        elif ch in "({":
            while k < len(line) and line[k] not in ")}":
                k += 1
            if k == len(line):
                raise AlignmentError(f"Line {line_num}: No matching close-bracket ) or }} found")
            out.append("?")
        else:
            raise AlignmentError(f"Line {line_num}: Unrecognized character {ch}")
        k += 1
    return "".join(out)


def _header(line):
    parts = line.split()
    try:
        nseq, nsite = int(parts[0]), int(parts[1])
    except (IndexError, ValueError):
        raise AlignmentError("Invalid PHYLIP format. First line must contain number of sequences and sites")
    if nseq < 3:
        raise AlignmentError("There must be at least 3 sequences")
    if nsite < 1:
        raise AlignmentError("No alignment columns")
    return nseq, nsite


def _cut_name(line):
    cut = [i for i, ch in enumerate(line) if ch in " \t"]
    pos = cut[0] if cut else 10
    return line[:pos], line[pos:]


def read_phylip(text, sequential=False):
    names, seqs, nseq, nsite, sid = [], [], 0, 0, 0
    for num, line in enumerate(text.splitlines(), 1):
        if line == "":
            continue
        if nseq == 0:
            nseq, nsite = _header(line)
            names, seqs = [""] * nseq, [""] * nseq
            continue
        if sequential and sid >= nseq:
            raise AlignmentError(f"Line {num}: Too many sequences detected")
        if names[sid] == "":
            names[sid], line = _cut_name(line)
        before = len(seqs[sid])
        seqs[sid] += process_seq(line, num)
        if sequential:
            if len(seqs[sid]) > nsite:
                raise AlignmentError(f"Line {num}: Sequence {names[sid]} is too long ({len(seqs[sid])})")
            if len(seqs[sid]) == nsite:
                sid += 1
        else:
            if len(seqs[sid]) != len(seqs[0]):
                raise AlignmentError(f"Line {num}: Sequence {names[sid]} has wrong sequence length {len(seqs[sid])}")
            if len(seqs[sid]) > before:
                sid += 1
            if sid == nseq:
                sid = 0
    if nseq == 0:
        raise AlignmentError("Invalid PHYLIP format. First line must contain number of sequences and sites")
    return names, seqs, nsite


def read_fasta(text):
    names, seqs = [], []
    for num, line in enumerate(text.splitlines(), 1):
        if line == "":
            continue
        if line.startswith(">"):
            names.append(line[1:].strip(" \t"))
            seqs.append("")
            continue
        if not seqs:
            raise AlignmentError("First line must begin with '>' to define sequence name")
        seqs[-1] += process_seq(line, num)
    short, remain = [""] * len(names), list(names)
    for _ in range(4):
        seen, dup = set(), False
        for i in range(len(names)):
            if remain[i] == "":
                continue
            cut = [k for k, ch in enumerate(remain[i]) if ch in " \t"]
            if not cut:
                short[i] += remain[i]
                remain[i] = ""
            else:
                short[i] += remain[i][: cut[0]]
                remain[i] = "_" + remain[i][cut[0] + 1:]
            if not dup:
                dup = short[i] in seen
                seen.add(short[i])
        if not dup:
            break
    if not seqs:
        raise AlignmentError("No sequences found")
    return short, seqs, len(seqs[0])


def detect_type(seqs):
    text = "".join(seqs)
    alpha = sum(ch.isalpha() for ch in text)
    digit = sum(ch.isdigit() for ch in text)
    if digit or alpha < 10:
        return None
    nuc = sum(ch in "ACGTURYWSMKBHDVNX" for ch in text)
    proper = sum(ch in "ACGTU" for ch in text)
    if nuc == alpha:
        return SEQ_DNA if proper / alpha > 0.9 else None
    aa = sum(ch in PROTEIN for ch in text)
    return SEQ_PROTEIN if aa / alpha > 0.9 else None


def state_unknown(seq_type):
    return {SEQ_DNA: 18, SEQ_PROTEIN: 23, SEQ_CODON: 61}[seq_type]


def nucleotide_code(ch, su):
    masks = {"A": 1, "C": 2, "G": 4, "T": 8, "U": 8, "R": 5, "Y": 10, "W": 9, "S": 6, "M": 3, "K": 12,
             "B": 14, "H": 11, "D": 13, "V": 7}
    if ch in "?-~.!NX":
        return su
    if ch not in masks:
        return None
    m = masks[ch]
    return {1: 0, 2: 1, 4: 2, 8: 3}.get(m, m + 3)


def protein_code(ch, su):
    if ch in "?-~.!X*UO":
        return su
    if ch in PROTEIN:
        return PROTEIN.index(ch)
    return {"B": 20, "Z": 21, "J": 22}.get(ch)


def code_alignment(names, seqs, seq_type):
    """Rows of state codes per taxon (codon columns joined), raising on invalid characters."""
    su = state_unknown(seq_type)
    sense = {}
    for c, aa in enumerate(STANDARD_CODE):
        if aa != "*":
            sense[c] = len(sense)
    rows, errors = [], []
    step = 3 if seq_type == SEQ_CODON else 1
    if len(seqs[0]) % step:
        raise AlignmentError("Number of sites is not multiple of 3")
    for name, s in zip(names, seqs):
        row = []
        for site in range(0, len(s), step):
            if seq_type == SEQ_PROTEIN:
                v = protein_code(s[site], su)
            else:
                codes = [nucleotide_code(ch, su) for ch in s[site:site + step]]
                if seq_type == SEQ_DNA:
                    v = codes[0]
                elif all(x is not None and x < 4 for x in codes):
                    c = codes[0] * 16 + codes[1] * 4 + codes[2]
                    v = sense.get(c, su)
                elif any(x is None for x in codes):
                    v = None
                else:
                    v = su
            if v is None:
                errors.append(f"Sequence {name} has invalid character {s[site:site + step]} at site {site + 1}")
            row.append(v)
        rows.append(row)
    if errors:
        raise AlignmentError("\n".join(errors))
    return rows


def appearance(seq_type, n, su, s):
    if s == su:
        return (1 << n) - 1
    if s < n:
        return 1 << s
    if seq_type == SEQ_DNA:
        return (s - 3) & 15
    pair = {20: (2, 3), 21: (5, 6), 22: (9, 10)}[s]
    return (1 << pair[0]) | (1 << pair[1])


def compress(rows, seq_type):
    """Patterns in first-occurrence order: (states [taxon][pattern], freq, site_pattern,
    flags (bit 0 constant, bit 1 informative), sites made of unknowns only)."""
    n = {SEQ_DNA: 4, SEQ_PROTEIN: 20, SEQ_CODON: 61}[seq_type]
    su = state_unknown(seq_type)
    ntaxa, nsites = len(rows), len(rows[0])
    index, pats, freq, site_pattern, gaps = {}, [], [], [], 0
    for site in range(nsites):
        col = tuple(rows[q][site] for q in range(ntaxa))
        gaps += all(v == su for v in col)
        if col not in index:
            index[col] = len(pats)
            pats.append(col)
            freq.append(0.0)
        freq[index[col]] += 1.0
        site_pattern.append(index[col])
    flags = []
    for col in pats:
        app = (1 << n) - 1
        counts = {}
        for v in col:
            app &= appearance(seq_type, n, su, v)
            if v < n:
                counts[v] = counts.get(v, 0) + 1
        twice = sum(1 for c in counts.values() if c >= 2)
        flags.append((1 if app else 0) | (2 if twice >= 2 else 0))
    states = [[col[q] for col in pats] for q in range(ntaxa)]
    return states, freq, site_pattern, flags, gaps


def read(text, fmt=0, seq_type=None):
    """fmt 0 detect, 1 PHYLIP, 2 PHYLIP sequential, 3 FASTA.  Returns a dict."""
    if fmt == 0:
        first = next((ln.strip(" \t") for ln in text.splitlines() if ln.strip(" \t")), "")
        fmt = 3 if first.startswith(">") else 1
    if fmt == 3:
        names, seqs, nsite = read_fasta(text)
    else:
        names, seqs, nsite = read_phylip(text, sequential=(fmt == 2))
    seen, err = set(), []
    for i, nm in enumerate(names):
        if nm == "":
            err.append(f"Sequence number {i + 1} has no names")
        if nm in seen:
            err.append(f"The sequence name {nm} is duplicated")
        seen.add(nm)
    if err:
        raise AlignmentError("\n".join(err))
    for nm, s in zip(names, seqs):
        if len(s) != nsite:
            raise AlignmentError(f"Sequence {nm} contains {'not enough' if len(s) < nsite else 'too many'} "
                                 f"characters ({len(s)})")
    if seq_type is None:
        seq_type = detect_type(seqs)
        if seq_type is None:
            raise AlignmentError("Unknown sequence type.")
    rows = code_alignment(names, seqs, seq_type)
    states, freq, site_pattern, flags, gaps = compress(rows, seq_type)
    return {"names": names, "seq_type": seq_type, "rows": rows, "states": states, "freq": freq,
            "site_pattern": site_pattern, "flags": flags, "gaps_only": gaps}
