"""Digital sequences for the filter library: text -> uint8 residue codes.

Codes follow Easel's digital alphabets (easel/esl_alphabet.c:225-300):
amino "ACDEFGHIKLMNPQRSTVWY-BJZOUX*~" (K=20, Kp=29), DNA "ACGT-RYMKSWHBVDN*~"
(K=4, Kp=18). Readers cover FASTA and the UniProt/EMBL text format (the "SQ"
block ended by "//").
"""
import numpy as np

AMINO_SYM = "ACDEFGHIKLMNPQRSTVWY-BJZOUX*~"
DNA_SYM = "ACGT-RYMKSWHBVDN*~"
RNA_SYM = "ACGU-RYMKSWHBVDN*~"


def _inmap(sym, equiv):
    m = np.full(256, 255, np.uint8)
    for i, c in enumerate(sym):
        m[ord(c)] = i
        m[ord(c.lower())] = i
    for c, to in equiv.items():
        m[ord(c)] = m[ord(to)]
        m[ord(c.lower())] = m[ord(to)]
    return m


INMAP = {
    3: _inmap(AMINO_SYM, {"_": "-", ".": "-"}),
    2: _inmap(DNA_SYM, {"U": "T", "X": "N", "_": "-", ".": "-"}),
    1: _inmap(RNA_SYM, {"T": "U", "X": "N", "_": "-", ".": "-"}),
}


def digitize(text, abctype=3):
    b = np.frombuffer("".join(text.split()).encode(), np.uint8)
    d = INMAP[abctype][b]
    if (d == 255).any():
        bad = chr(int(b[np.argmax(d == 255)]))
        raise ValueError(f"illegal residue {bad!r}")
    return d


def read_fasta(path, abctype=3):
    """List of (name, digital residues)."""
    out, name, buf = [], None, []
    with open(path) as fh:
        for line in fh:
            if line.startswith(">"):
                if name is not None:
                    out.append((name, digitize("".join(buf), abctype)))
                name, buf = line[1:].split()[0] if line[1:].split() else "", []
            else:
                buf.append(line.strip())
    if name is not None:
        out.append((name, digitize("".join(buf), abctype)))
    return out


def read_uniprot(path, abctype=3):
    """List of (name, digital residues) from UniProt/EMBL text records."""
    out, name, buf, inseq = [], None, [], False
    with open(path) as fh:
        for line in fh:
            if line.startswith("ID "):
                name = line.split()[1]
            elif line.startswith("SQ "):
                inseq, buf = True, []
            elif line.startswith("//"):
                if name is not None:
                    out.append((name, digitize("".join(buf), abctype)))
                name, inseq = None, False
            elif inseq:
                buf.append(line.strip())
    return out


def read_uniprot_records(path, abctype=3):
    """List of (name, accession, description, digital residues) from
    UniProt/EMBL text records, with the header fields as Easel's EMBL parser
    keeps them (easel/esl_sqio_ascii.c:2610-2680): the name from ID, the first
    accession of the first AC line (up to its ';'), and the DE lines' text
    after their five-character tag, right-trimmed and joined by one space."""
    out, name, acc, desc, buf, inseq = [], None, "", [], [], False
    with open(path) as fh:
        for line in fh:
            if line.startswith("ID "):
                name, acc, desc = line.split()[1], "", []
            elif line.startswith("AC   ") and not acc:
                acc = line[5:].strip().lstrip(";").split(";")[0].strip()
            elif line.startswith("DE   "):
                desc.append(line[5:].rstrip())
            elif line.startswith("SQ "):
                inseq, buf = True, []
            elif line.startswith("//"):
                if name is not None:
                    out.append((name, acc, " ".join(desc), digitize("".join(buf), abctype)))
                name, inseq = None, False
            elif inseq:
                buf.append(line.strip())
    return out
