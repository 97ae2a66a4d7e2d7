"""Writes the synthetic query profiles used by the tests and the benchmark.

Each profile is drawn by biobench2_amd.synth.random_profile (seeded), written
as a HMMER3/f file, read back through the library, and calibrated the way
p7_Calibrate does (HMMER/src/evalues.c:60-466, defaults EmL=EvL=200,
EmN=EvN=EfN=200, EfL=100, Eft=0.04):
  lambda = ln2 + 1.44 / (M * H), H = mean match relative entropy in bits;
  MSV / Viterbi mu = Gumbel location fit with that lambda over i.i.d. scores;
  Forward tau from a full Gumbel fit of Forward scores, shifted to the tail.
Scores come from the oracle (CPU), so this script is test infrastructure.
The resulting STATS lines make the reference and this library read the same
E-value parameters.

    python tests/golden/make_profiles.py
"""
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from biobench2_amd import p7gpu, synth   # noqa: E402
import oracle                            # noqa: E402

OUT = os.path.join(HERE, "synth")
LOG2 = 0.69314718055994529

# name, M, seed
PROFILES = [
    ("synth_M200", 200, 42),      # configuration 1
    ("synth_M400", 400, 42),      # configuration 2 (benchmark)
    ("synth_M50", 50, 7),
    ("synth_M1000", 1000, 11),
]


def gumbel_fit(x):
    """Maximum-likelihood Gumbel (mu, lambda) (Lawless, as esl_gumbel_FitComplete)."""
    x = np.asarray(x, float)
    lam = math.pi / (math.sqrt(6.0) * x.std())
    for _ in range(100):
        e = np.exp(-lam * x)
        s0, s1, s2 = e.sum(), (x * e).sum(), (x * x * e).sum()
        f = 1.0 / lam - x.mean() + s1 / s0
        df = -1.0 / lam ** 2 - (s2 / s0 - (s1 / s0) ** 2)
        step = f / df
        lam -= step
        if abs(step) < 1e-10:
            break
    mu = -math.log(np.exp(-lam * x).mean()) / lam
    return mu, lam


def calibrate(path, seed=42):
    om = p7gpu.OProfile.from_file(path, 0, L=200)
    h = synth.read_hmm_probs(path)
    f = synth.AMINO_BG / synth.AMINO_BG.sum()
    H = float(np.mean([(p[p > 0] * np.log2(p[p > 0] / f[p > 0])).sum() for p in h.mat[1:]]))
    lam = LOG2 + 1.44 / (h.M * H)
    rng = np.random.default_rng(seed)
    t = om.tables()
    lm = oracle.LengthModel(t, 200)
    # MSV and Viterbi mu at L = 200
    xm, xv = [], []
    xw = t["xw"].copy(); xw[1, 0] = xw[2, 0] = xw[3, 0] = lm.xwmove[200]
    maxm = (255 - t["base_b"]) / t["scale_b"]
    for _ in range(200):
        s = synth.iid_residues(200, rng)
        nul = oracle.null_one(200)
        st, sc, _ = oracle.msv_filter(s, t, tjb=int(lm.tjb[200]))
        xm.append(((maxm if st == 16 else sc) - nul) / LOG2)
        st, sc, _ = oracle.vit_filter(s, t, xw)
        xv.append((sc - nul) / LOG2)
    mmu = -math.log(np.mean(np.exp(-lam * np.array(xm)))) / lam
    vmu = -math.log(np.mean(np.exp(-lam * np.array(xv)))) / lam
    # Forward tau at L = 100
    lm = oracle.LengthModel(t, 100)
    xf = t["xf"].copy()
    xf[1, 0] = xf[2, 0] = xf[3, 0] = lm.xfmove[100]
    xf[1, 1] = xf[2, 1] = xf[3, 1] = lm.xfloop[100]
    xs = []
    for _ in range(200):
        s = synth.iid_residues(100, rng)
        st, sc, _, _ = oracle.fwd_parser(s, t, xf)
        xs.append((sc - oracle.null_one(100)) / LOG2)
    gmu, glam = gumbel_fit(xs)
    tailp = 0.04
    tau = gmu - math.log(-math.log(1.0 - tailp)) / glam + math.log(tailp) / lam
    return (mmu, lam, vmu, lam, tau, lam)


def main():
    os.makedirs(OUT, exist_ok=True)
    for name, M, seed in PROFILES:
        rng = np.random.default_rng(seed)
        h = synth.random_profile(M, rng, name)
        path = os.path.join(OUT, name + ".hmm")
        synth.write_hmm(path, [h])
        h.stats = calibrate(path)
        synth.write_hmm(path, [h])
        print(name, M, ["%.4f" % v for v in h.stats])


if __name__ == "__main__":
    main()
