"""Synthetic workloads: random profile HMMs and digital target databases.

Profiles follow the recipe of the benchmark configurations: match emissions
~ Dirichlet(20 * pi_bg), insert emissions = pi_bg, transitions MM/MI/MD =
.94/.03/.03, IM/II = .6/.4, DM/DD = .7/.3, each jittered by U(0.8, 1.2) and
renormalised, with HMMER's node-M convention (MD = 0, DM = 1, DD = 0;
HMMER/src/p7_hmm.c:899-902). They are written as HMMER3/f save files, so the
reference and this library read identical numbers.

Databases: lengths ~ round(lognormal(ln 300, 0.6)) clipped to [30, 5000],
residues i.i.d. pi_bg, and a fraction of planted homologs: a local segment
emitted from the profile (match/insert/delete path) between i.i.d. flanks.
"""
import math

import numpy as np

AMINO_BG = np.array([0.0787945, 0.0151600, 0.0535222, 0.0668298, 0.0397062, 0.0695071, 0.0229198,
                     0.0590092, 0.0594422, 0.0963728, 0.0237718, 0.0414386, 0.0482904, 0.0395639,
                     0.0540978, 0.0683364, 0.0540687, 0.0673417, 0.0114135, 0.0304133])
AMINO_SYM = "ACDEFGHIKLMNPQRSTVWY"
H_MM, H_MI, H_MD, H_IM, H_II, H_DM, H_DD = range(7)


class ProfileHMM:
    """Core probability model: mat[M+1,K], ins[M+1,K], t[M+1,7] (HMMER order MM MI MD IM II DM DD)."""

    def __init__(self, name, mat, ins, t):
        self.name, self.mat, self.ins, self.t = name, mat, ins, t
        self.M = mat.shape[0] - 1
        self.stats = None          # (mmu, mlambda, vmu, vlambda, ftau, flambda)

    def occupancy(self):
        """Match occupancy (HMMER/src/p7_hmm.c:1271-1289) and insert expected use."""
        t, M = self.t, self.M
        mocc = np.zeros(M + 1)
        mocc[1] = t[0, H_MI] + t[0, H_MM]
        for k in range(2, M + 1):
            mocc[k] = mocc[k - 1] * (t[k - 1, H_MM] + t[k - 1, H_MI]) + (1.0 - mocc[k - 1]) * t[k - 1, H_DM]
        iocc = np.zeros(M + 1)
        iocc[0] = t[0, H_MI] / t[0, H_IM]
        for k in range(1, M + 1):
            iocc[k] = mocc[k] * t[k, H_MI] / t[k, H_IM]
        return mocc, iocc

    def composition(self):
        """Model composition for the bias filter (occupancy-weighted match + insert emissions)."""
        mocc, iocc = self.occupancy()
        c = (mocc[1:, None] * self.mat[1:]).sum(0) + (iocc[:, None] * self.ins).sum(0)
        return c / c.sum()


def _jitter(rng, base):
    v = np.asarray(base, float) * rng.uniform(0.8, 1.2, size=len(base))
    return v / v.sum()


def random_profile(M, rng, name=None):
    K = 20
    mat = np.zeros((M + 1, K))
    mat[1:] = rng.dirichlet(20.0 * AMINO_BG, size=M)
    mat[0, 0] = 1.0                                  # node 0 has no match state
    ins = np.tile(AMINO_BG, (M + 1, 1))
    t = np.zeros((M + 1, 7))
    for k in range(M + 1):
        mm, mi, md = _jitter(rng, [0.94, 0.03, 0.03])
        im, ii = _jitter(rng, [0.6, 0.4])
        dm, dd = _jitter(rng, [0.7, 0.3])
        t[k] = [mm, mi, md, im, ii, dm, dd]
    t[0, H_DM], t[0, H_DD] = 1.0, 0.0                # D_0 does not exist
    s = t[M, H_MM] + t[M, H_MI]                      # node M: no M->D, D->M is the exit
    t[M, H_MM], t[M, H_MI], t[M, H_MD] = t[M, H_MM] / s, t[M, H_MI] / s, 0.0
    t[M, H_DM], t[M, H_DD] = 1.0, 0.0
    return ProfileHMM(name or f"synth{M}", mat, ins, t)


def _fmt(p):
    if p <= 0.0:
        return "       *"
    v = -math.log(p)
    return f"{(v if v > 0.0 else 0.0):8.5f}"


def write_hmm(path, hmms):
    """HMMER3/f ascii save file (the format HMMER/src/p7_hmmfile.c:1294 reads)."""
    with open(path, "w") as fh:
        for h in hmms:
            M = h.M
            fh.write("HMMER3/f [3.1b2 | February 2015]\n")
            fh.write(f"NAME  {h.name}\nLENG  {M}\nALPH  amino\nRF    no\nMM    no\nCONS  yes\nCS    no\nMAP   no\n")
            if h.stats is not None:
                mmu, mlam, vmu, vlam, ftau, flam = h.stats
                fh.write(f"STATS LOCAL MSV      {mmu:8.4f}  {mlam:8.5f}\n")
                fh.write(f"STATS LOCAL VITERBI  {vmu:8.4f}  {vlam:8.5f}\n")
                fh.write(f"STATS LOCAL FORWARD  {ftau:8.4f}  {flam:8.5f}\n")
            fh.write("HMM          " + "".join(f"{c:>9s}" for c in AMINO_SYM) + "\n")
            fh.write("            m->m     m->i     m->d     i->m     i->i     d->m     d->d\n")
            comp = h.composition()
            fh.write("  COMPO   " + " ".join(_fmt(p) for p in comp) + "\n")
            fh.write("          " + " ".join(_fmt(p) for p in h.ins[0]) + "\n")
            fh.write("          " + " ".join(_fmt(p) for p in h.t[0]) + "\n")
            for k in range(1, M + 1):
                cons = AMINO_SYM[int(np.argmax(h.mat[k]))].lower()
                fh.write(f"{k:7d}   " + " ".join(_fmt(p) for p in h.mat[k]) + f" {'-':>6s} {cons} - - -\n")
                fh.write("          " + " ".join(_fmt(p) for p in h.ins[k]) + "\n")
                fh.write("          " + " ".join(_fmt(p) for p in h.t[k]) + "\n")
            fh.write("//\n")


def read_hmm_probs(path, which=0):
    """Probabilities of a HMMER3 save file as a ProfileHMM (for emitting homologs)."""
    recs, cur = [], None
    with open(path) as fh:
        lines = fh.read().split("\n")
    i = 0
    while i < len(lines):
        ln = lines[i]
        if ln.startswith("HMMER3"):
            cur = {"name": None, "M": 0, "stats": {}}
        elif cur is not None and ln.startswith("NAME"):
            cur["name"] = ln.split()[1]
        elif cur is not None and ln.startswith("LENG"):
            cur["M"] = int(ln.split()[1])
        elif cur is not None and ln.startswith("HMM "):
            M = cur["M"]
            i += 2
            toks = lines[i].split()
            if toks[0] == "COMPO":
                i += 1
                toks = lines[i].split()
            p = lambda s: 0.0 if s == "*" else math.exp(-float(s))
            mat = np.zeros((M + 1, 20)); ins = np.zeros((M + 1, 20)); t = np.zeros((M + 1, 7))
            ins[0] = [p(s) for s in toks[:20]]
            i += 1
            t[0] = [p(s) for s in lines[i].split()[:7]]
            for k in range(1, M + 1):
                i += 1
                mat[k] = [p(s) for s in lines[i].split()[1:21]]
                i += 1
                ins[k] = [p(s) for s in lines[i].split()[:20]]
                i += 1
                t[k] = [p(s) for s in lines[i].split()[:7]]
            recs.append(ProfileHMM(cur["name"], mat, ins, t))
            cur = None
        i += 1
    return recs[which]


def _cum_lut(p, bits=16):
    """Residue lookup table for u16 randoms: LUT[u] = smallest x with cdf(x) > u/2^bits."""
    c = np.cumsum(np.asarray(p, float) / np.sum(p))
    u = (np.arange(1 << bits) + 0.5) / (1 << bits)
    return np.searchsorted(c, u, side="right").clip(0, len(p) - 1).astype(np.uint8)


def iid_residues(n, rng, p=AMINO_BG, chunk=1 << 26):
    lut = _cum_lut(p)
    out = np.empty(n, np.uint8)
    for a in range(0, n, chunk):
        b = min(n, a + chunk)
        out[a:b] = lut[rng.integers(0, 1 << 16, size=b - a, dtype=np.uint16)]
    return out


def lognormal_lengths(n, rng, median=300.0, sigma=0.6, lo=30, hi=5000):
    return np.clip(np.rint(rng.lognormal(math.log(median), sigma, size=n)), lo, hi).astype(np.int32)


def _sample_rows(rng, P, rows):
    """One categorical draw per entry of rows from the distributions P[rows]."""
    c = np.cumsum(P[rows], axis=1)
    u = rng.random(len(rows))[:, None] * c[:, -1:]
    return (u > c).sum(axis=1).astype(np.uint8)


def emit_homologs(hmm, n, rng, indel_rich=False):
    """n local segments emitted from hmm along sampled M/I/D paths. Returns a list of uint8 arrays."""
    M = hmm.M
    t = hmm.t.copy()
    if indel_rich:                                   # 30% more insert/delete traffic
        t[:, H_MI] *= 3.0; t[:, H_MD] *= 3.0
        t[:, :3] /= t[:, :3].sum(1, keepdims=True)
    k0 = rng.integers(1, max(2, M // 10 + 1), size=n)
    k1 = rng.integers(max(1, M - M // 10), M + 1, size=n)
    k1 = np.maximum(k1, k0)
    state = np.zeros(n, np.int8)                     # 0 = M, 1 = D
    ev_h, ev_o, ev_x = [], [], []
    for k in range(1, M + 1):
        act = np.nonzero((k >= k0) & (k <= k1))[0]
        if act.size == 0:
            continue
        st = state[act]
        m_act = act[st == 0]
        if m_act.size:
            ev_h.append(m_act); ev_o.append(np.full(m_act.size, 2 * k, np.int64))
            ev_x.append(_sample_rows(rng, hmm.mat, np.full(m_act.size, k)))
        if k == M:
            break
        r = rng.random(act.size)
        from_m = st == 0
        pm = t[k]
        go_i = from_m & (r >= pm[H_MM]) & (r < pm[H_MM] + pm[H_MI])
        go_d = (from_m & (r >= pm[H_MM] + pm[H_MI])) | (~from_m & (r >= pm[H_DM]))
        if go_i.any():
            hi = act[go_i]
            nins = rng.geometric(1.0 - t[k, H_II], size=hi.size)
            rep = np.repeat(hi, nins)
            ev_h.append(rep); ev_o.append(np.full(rep.size, 2 * k + 1, np.int64))
            ev_x.append(_sample_rows(rng, hmm.ins, np.full(rep.size, k)))
        state[act] = np.where(go_d, 1, 0).astype(np.int8)
    if not ev_h:
        return [np.zeros(0, np.uint8) for _ in range(n)]
    h = np.concatenate(ev_h); o = np.concatenate(ev_o); x = np.concatenate(ev_x)
    order = np.lexsort((o, h))
    h, x = h[order], x[order]
    cuts = np.searchsorted(h, np.arange(n + 1))
    return [x[cuts[i]:cuts[i + 1]] for i in range(n)]


class SynthDB:
    """Concatenated digital residues + offsets + lengths, with the planted-homolog truth."""

    def __init__(self, residues, offsets, lengths, homolog):
        self.residues, self.offsets, self.lengths, self.homolog = residues, offsets, lengths, homolog

    @property
    def nseq(self):
        return int(self.lengths.size)

    @property
    def nres(self):
        return int(self.lengths.sum())

    def seq(self, i):
        o = int(self.offsets[i])
        return self.residues[o:o + int(self.lengths[i])]

    def subset(self, idx):
        idx = np.asarray(idx)
        lens = self.lengths[idx]
        offs = np.zeros(idx.size, np.int64)
        if idx.size > 1:
            offs[1:] = np.cumsum(lens[:-1])
        res = np.concatenate([self.seq(i) for i in idx]) if idx.size else np.zeros(0, np.uint8)
        return SynthDB(res, offs, lens.astype(np.int32), self.homolog[idx])


def make_db(nseq, hmms, frac_homolog=0.01, seed_db=43, seed_plant=44, median=300.0, sigma=0.6, lo=30, hi=5000,
            long_frac=0.0, long_range=(5000, 35000), indel_rich=False):
    """Database of nseq sequences; a fraction frac_homolog carries a segment emitted from one of hmms."""
    rng = np.random.default_rng(seed_db)
    lengths = lognormal_lengths(nseq, rng, median, sigma, lo, hi)
    if long_frac > 0:
        nl = int(round(long_frac * nseq))
        li = rng.choice(nseq, size=nl, replace=False)
        lengths[li] = rng.integers(long_range[0], long_range[1] + 1, size=nl)
    prng = np.random.default_rng(seed_plant)
    nh = int(round(frac_homolog * nseq))
    hidx = np.sort(prng.choice(nseq, size=nh, replace=False)) if nh else np.zeros(0, np.int64)
    homolog = np.full(nseq, -1, np.int32)
    cores = {}
    if nh:
        which = prng.integers(0, len(hmms), size=nh)
        for j, hm in enumerate(hmms):
            sel = hidx[which == j]
            segs = emit_homologs(hm, sel.size, prng, indel_rich=indel_rich)
            for i, s in zip(sel, segs):
                cores[int(i)] = s
                homolog[i] = j
                lengths[i] = max(int(lengths[i]), s.size + 2)
    offsets = np.zeros(nseq, np.int64)
    offsets[1:] = np.cumsum(lengths[:-1], dtype=np.int64)
    residues = iid_residues(int(lengths.sum()), rng)
    for i, s in cores.items():
        L = int(lengths[i])
        a = int(prng.integers(0, L - s.size + 1))
        o = int(offsets[i]) + a
        residues[o:o + s.size] = s
    return SynthDB(residues, offsets, lengths, homolog)


# ---- E-value calibration (the recipe of p7_Calibrate, HMMER/src/evalues.c:60-466) ----
LOG2 = 0.69314718055994529


def gumbel_fit(x):
    """Maximum-likelihood Gumbel (mu, lambda) of a sample (Lawless; as esl_gumbel_FitComplete)."""
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


def evparams(hmm, msv_bits, vit_bits, fwd_bits, tailp=0.04):
    """(mmu, lambda, vmu, lambda, tau, lambda) from bit scores of i.i.d. sequences:
    lambda = ln2 + 1.44 / (M * H) with H the mean match relative entropy in bits
    (evalues.c:186-191); mu by fixed-lambda Gumbel fits; tau from a Gumbel fit of
    the Forward scores shifted to the tail (evalues.c:430-466)."""
    f = AMINO_BG / AMINO_BG.sum()
    H = float(np.mean([(p[p > 0] * np.log2(p[p > 0] / f[p > 0])).sum() for p in hmm.mat[1:]]))
    lam = LOG2 + 1.44 / (hmm.M * H)
    mmu = -math.log(np.mean(np.exp(-lam * np.asarray(msv_bits)))) / lam
    vmu = -math.log(np.mean(np.exp(-lam * np.asarray(vit_bits)))) / lam
    gmu, glam = gumbel_fit(fwd_bits)
    tau = gmu - math.log(-math.log(1.0 - tailp)) / glam + math.log(tailp) / lam
    return (mmu, lam, vmu, lam, tau, lam)


def calibration_seqs(nseq=200, seed=42):
    """The i.i.d. samples of a calibration: nseq sequences of L = 200 (MSV, Viterbi)
    and nseq of L = 100 (Forward)."""
    rng = np.random.default_rng(seed)
    return [iid_residues(200, rng) for _ in range(nseq)], [iid_residues(100, rng) for _ in range(nseq)]


def calibrate_with(gpu, path, nseq=200, seed=42):
    """E-value parameters of the profile in `path` from this library's own MSV,
    Viterbi (L = 200) and Forward (L = 100) scores of i.i.d. sequences."""
    om = gpu.OProfile.from_file(path, 0)
    t = om.tables()
    hmm = read_hmm_probs(path)
    s200, s100 = calibration_seqs(nseq, seed)
    bg = gpu.Background(gpu.AMINO)
    bg.set_length(200)                                   # p7_bg_SetLength before p7_bg_NullOne
    n200 = float(bg.null_one(200))
    bg.set_length(100)
    n100 = float(bg.null_one(100))
    maxm = (255 - t["base_b"]) / t["scale_b"]           # an overflowing MSV score, as p7_MSVMu counts it
    db200, db100 = gpu.SeqDB.from_list(s200), gpu.SeqDB.from_list(s100)
    msc, _ = gpu.msv_batch(om, db200)
    vsc, _ = gpu.viterbi_batch(om, db200)
    fsc, _ = gpu.forward_batch(om, db100)
    msc = np.where(np.isinf(msc), maxm, msc.astype(np.float64))
    return evparams(hmm, (msc - n200) / LOG2, (vsc.astype(np.float64) - n200) / LOG2,
                    (fsc.astype(np.float64) - n100) / LOG2)


def query_family(gpu, dirpath, nprof=8, median_M=250, sigma=0.6, Mmin=50, Mmax=3000, seed=7):
    """Calibrated query profiles with M ~ lognormal(median, sigma) clipped to [Mmin, Mmax],
    written to dirpath; returns their paths."""
    import os
    rng = np.random.default_rng(seed)
    paths = []
    for q in range(nprof):
        M = int(np.clip(round(math.exp(math.log(median_M) + sigma * rng.standard_normal())), Mmin, Mmax))
        h = random_profile(M, rng, f"query{q}_M{M}")
        path = os.path.join(dirpath, h.name + ".hmm")
        write_hmm(path, [h])
        h.stats = calibrate_with(gpu, path, seed=seed + q)
        write_hmm(path, [h])
        paths.append(path)
    return paths


def batch_family(gpu, dirpath, nprof=64, Mmin=50, Mmax=2000, seed=42):
    """Calibrated profiles with M log-uniform in [Mmin, Mmax] (a batch of uneven
    models), written to dirpath; returns their paths."""
    import os
    rng = np.random.default_rng(seed)
    paths = []
    for q in range(nprof):
        M = int(round(math.exp(rng.uniform(math.log(Mmin), math.log(Mmax)))))
        h = random_profile(M, rng, f"batch{q}_M{M}")
        path = os.path.join(dirpath, h.name + ".hmm")
        write_hmm(path, [h])
        h.stats = calibrate_with(gpu, path, seed=seed + q)
        write_hmm(path, [h])
        paths.append(path)
    return paths


def serving_family(gpu, dirpath, nfam=1000, nquery=16, ndecoy=4, median_M=250, sigma=0.6, Mmin=50, Mmax=3000,
                   seed=11):
    """A hidden family set of nfam profiles with M ~ lognormal(median, sigma)
    clipped to [Mmin, Mmax], nquery of them drawn as queries, plus ndecoy
    profiles from the same generator that are not in the set. The queries and
    decoys are calibrated and written to dirpath. Returns (family, query
    paths, decoy paths); the family's profiles are the ones to plant."""
    import os
    rng = np.random.default_rng(seed)
    Ms = np.clip(np.rint(rng.lognormal(math.log(median_M), sigma, nfam + ndecoy)), Mmin, Mmax).astype(int)
    fam = [random_profile(int(M), rng, f"fam{i}_M{M}") for i, M in enumerate(Ms[:nfam])]
    dec = [random_profile(int(M), rng, f"decoy{i}_M{M}") for i, M in enumerate(Ms[nfam:])]
    pick = np.sort(rng.choice(nfam, size=min(nquery, nfam), replace=False))
    paths = []
    for j, h in enumerate([fam[i] for i in pick] + dec):
        path = os.path.join(dirpath, h.name + ".hmm")
        write_hmm(path, [h])
        h.stats = calibrate_with(gpu, path, seed=seed + j)
        write_hmm(path, [h])
        paths.append(path)
    return fam, paths[:len(pick)], paths[len(pick):]
