"""The device pipeline's stage gates (k_msv_gate and the gate tails of the bias,
Viterbi and Forward kernels, which compare bit scores with thresholds from
p7_gpu_ScoreThreshold) against the gate formulas of p7_Pipeline
(HMMER/src/p7_pipeline.c:687-731) computed in float64 from the device's own
scores (tests/gate_props.py), on a database with planted low-complexity
sequences where the bias filter's score decides later gates; and the stages
equal the oracle's."""
import numpy as np
import pytest

import oracle
from conftest import golden
from gate_props import check_gates, lowcomplexity_db

pytestmark = pytest.mark.gpu

CASES = {
    "default": ((0.02, 1e-3, 1e-5), True),
    "loose": ((0.3, 0.05, 0.01), True),
    "f2_equals_f1": ((0.02, 0.02, 1e-5), True),
    "no_bias": ((0.3, 0.05, 0.01), False),
}


@pytest.fixture(scope="module")
def lc(gpu):
    path = golden("fn3.hmm")
    om = gpu.OProfile.from_file(path)
    t = om.tables()
    db = lowcomplexity_db(path, t["compo"], 20000, 0.10, seed=302)
    sdb = gpu.SeqDB(db.residues, db.offsets, db.lengths)
    return om, t, db, sdb


@pytest.mark.parametrize("case", sorted(CASES))
def test_gpu_gates_follow_the_gate_formulas(gpu, lc, case):
    om, t, db, sdb = lc
    F, do_bias = CASES[case]
    got = gpu.Pipeline(*F, do_biasfilter=do_bias).run(om, sdb)
    flips = check_gates(got, t, F, do_bias, db.lengths)
    if do_bias:
        assert flips > 0
    if case == "f2_equals_f1":
        assert np.all(np.isnan(got["vfsc"]))
    prof, keep = oracle.make_profile(t, oracle.bias_hmm(t), t["evparam"], F1=F[0], F2=F[1], F3=F[2],
                                     do_bias=do_bias, Lmax=int(db.lengths.max()) + 1)
    ref = oracle.pipeline_db(prof, db.residues, db.offsets, db.lengths, nthreads=8)
    assert np.array_equal(got["stage"], ref["stage"].astype(np.uint8))
    for k in ("usc", "filtersc", "vfsc"):
        assert np.array_equal(got[k], ref[k], equal_nan=True), k
