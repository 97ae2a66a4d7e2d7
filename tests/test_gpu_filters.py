"""GPU filters against the oracle, per target, through the C ABI.

MSV and Viterbi: bit-exact scores and statuses (incl. overflow -> +inf,
eslERANGE). Forward/Backward (fp32): within 1e-4 nats of the oracle's fp32
restatement of the reference; the error against the FP64 log-space oracle
is measured and bounded too. Bias filter: bit-exact.

Targets cover the kernel edge cases: lengths 1, 2, 15/16/17 (16-residue
chunks), model-length multiples, long targets, planted homologs that
overflow, degenerate residues (X, B, Z, U, O) and an all-X sequence (no
cell above the B floor, which takes the exact J-state path).
"""
import numpy as np
import pytest

import oracle
from biobench2_amd import p7gpu, seqio, synth
from conftest import golden, random_seqs, synth_profile

pytestmark = pytest.mark.gpu

MODELS = [(golden("minifam"), 0), (golden("minifam"), 1), (golden("minifam"), 2), (golden("LuxC.hmm"), 0),
          (golden("M1.hmm"), 0), (golden("XYPPX.hmm"), 0), (golden("20aa.hmm"), 0), (golden("Patched.hmm"), 0),
          (golden("SMC_N.hmm"), 0), (synth_profile("synth_M400"), 0), (synth_profile("synth_M1000"), 0)]


def target_set():
    rng = np.random.default_rng(2024)
    seqs = [s for _, s in seqio.read_fasta(golden("globins45.fa"))[:12]]
    seqs += [s for _, s in seqio.read_uniprot(golden("7LESS_DROME"))]
    seqs += [s for _, s in seqio.read_fasta(golden("rndseq400-10.fa"))]
    seqs += random_seqs(rng, [1, 2, 3, 15, 16, 17, 31, 32, 33, 63, 64, 65, 149, 150, 400, 401, 1000, 4999, 9000])
    seqs.append(seqio.digitize("X" * 50))
    seqs.append(seqio.digitize("ACDXXBZUOJ" * 7))
    hm = synth.read_hmm_probs(synth_profile("synth_M400"))
    for core in synth.emit_homologs(hm, 4, rng):
        flank = random_seqs(rng, [40])[0]
        seqs.append(np.concatenate([flank, core, flank]))
    return seqs


@pytest.fixture(scope="module")
def targets(gpu):
    seqs = target_set()
    return seqs, gpu.SeqDB.from_list(seqs)


def per_target_config(t, L, lm):
    xw = t["xw"].copy()
    xw[1, 0] = xw[2, 0] = xw[3, 0] = lm.xwmove[L]
    xf = t["xf"].copy()
    xf[1, 0] = xf[2, 0] = xf[3, 0] = lm.xfmove[L]
    xf[1, 1] = xf[2, 1] = xf[3, 1] = lm.xfloop[L]
    return xw, xf


def same(a, b):
    return a == b or (np.isinf(a) and np.isinf(b) and np.sign(a) == np.sign(b))


@pytest.mark.parametrize("path,which", MODELS)
def test_msv_and_viterbi_bit_exact(gpu, targets, path, which):
    seqs, db = targets
    om = gpu.OProfile.from_file(path, which)
    t = om.tables()
    lm = oracle.LengthModel(t, max(len(s) for s in seqs) + 1)
    msc, mst = gpu.msv_batch(om, db)
    vsc, vst = gpu.viterbi_batch(om, db)
    for i, s in enumerate(seqs):
        L = len(s)
        st, sc, _ = oracle.msv_filter(s, t, tjb=int(lm.tjb[L]))
        assert st == mst[i] and same(sc, msc[i]), ("msv", i, L, st, sc, mst[i], msc[i])
        xw, _ = per_target_config(t, L, lm)
        st, sc, _ = oracle.vit_filter(s, t, xw)
        assert st == vst[i] and same(sc, vsc[i]), ("vit", i, L, st, sc, vst[i], vsc[i])


@pytest.mark.parametrize("path,which", MODELS[:5] + MODELS[9:10])
def test_forward_backward_within_1e4_nats(gpu, targets, path, which):
    seqs, db = targets
    om = gpu.OProfile.from_file(path, which)
    t = om.tables()
    lm = oracle.LengthModel(t, max(len(s) for s in seqs) + 1)
    fsc, fst = gpu.forward_batch(om, db)
    bsc, bst = gpu.backward_batch(om, db)
    err_f = err_b = err64 = 0.0
    for i, s in enumerate(seqs):
        L = len(s)
        _, xf = per_target_config(t, L, lm)
        st, sc, xmx, _ = oracle.fwd_parser(s, t, xf)
        assert st == fst[i] == 0
        st2, sc2, _, _ = oracle.bck_parser(s, t, xmx, xf)
        assert st2 == bst[i] == 0
        err_f = max(err_f, abs(sc - fsc[i]))
        err_b = max(err_b, abs(sc2 - bsc[i]))
        om.gm.reconfig_length(L)
        tsc, rsc, xsc = om.gm.scores()
        err64 = max(err64, abs(fsc[i] - oracle.forward_log64(s, tsc, rsc, xsc)))
    assert err_f <= 1e-4 and err_b <= 1e-4, (err_f, err_b)
    # fp32 against the FP64 log-space Forward over every target (L up to 9000):
    # the oracle's fp32 restatement measures at most 9.4e-6 nats on this set
    assert err64 <= 1e-4, err64


@pytest.mark.parametrize("path,which", MODELS[:4] + MODELS[9:10])
def test_bias_filter_bit_exact(gpu, targets, path, which):
    seqs, db = targets
    om = gpu.OProfile.from_file(path, which)
    t = om.tables()
    bh = oracle.bias_hmm(t)
    got = gpu.filter_score_batch(om, db)
    for i, s in enumerate(seqs):
        L = len(s)
        p1 = np.float32(L) / np.float32(L + 1)
        ref = oracle.filter_score(s, bh["eo"], p1, np.float32(1) - p1, bh["t10"], bh["t11"], bh["pi0"], bh["pi1"], p1)
        assert ref == got[i], (i, L, ref, got[i])


def test_subset_indices_and_order(gpu, targets):
    seqs, db = targets
    om = gpu.OProfile.from_file(golden("minifam"), 2)
    full, _ = gpu.msv_batch(om, db)
    idx = np.array([5, 0, len(seqs) - 1, 3, 3], np.int64)
    sub, _ = gpu.msv_batch(om, db, idx)
    assert np.array_equal(sub, full[idx])
    vf, _ = gpu.viterbi_batch(om, db, idx)
    vfull, _ = gpu.viterbi_batch(om, db)
    assert np.array_equal(vf, vfull[idx])
