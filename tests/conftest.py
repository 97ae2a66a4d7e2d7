import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")
REF = os.path.join(GOLDEN, "ref")
SYNTH = os.path.join(GOLDEN, "synth")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def gpu():
    from biobench2_amd import p7gpu
    p7gpu.init(0)
    return p7gpu


def golden(*parts):
    return os.path.join(REF, *parts)


def synth_profile(name):
    return os.path.join(SYNTH, name + ".hmm")


def random_seqs(rng, lengths, p=None):
    from biobench2_amd import synth
    p = synth.AMINO_BG if p is None else p
    return [synth.iid_residues(int(L), rng, p) for L in lengths]
