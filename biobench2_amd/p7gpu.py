"""ctypes binding of libp7gpu.so, the C ABI declared in include/p7_gpu.h.

The classes mirror the HMMER objects the C entry points stand in for
(P7_HMM, P7_BG, P7_PROFILE, P7_OPROFILE, P7_OMX, P7_PIPELINE); every call goes
through the HIP library. There is no CPU fallback: if the shared library is
missing, importing this module raises.
"""
import ctypes
import os
from ctypes import POINTER, c_double, c_float, c_int, c_int16, c_int32, c_int64, c_uint8, c_uint32, c_void_p, c_char_p

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_HERE, "libp7gpu.so")

OK, FAIL, EOF, EMEM, ENOTFOUND, EFORMAT, EINCOMPAT, EINVAL, ESYS, ERANGE, ENORESULT, EDEVICE = \
    0, 1, 3, 5, 6, 7, 10, 11, 12, 16, 19, 64
RNA, DNA, AMINO = 1, 2, 3
LOCAL, GLOCAL, UNILOCAL = 1, 2, 3
BATCH_EAGER, BATCH_GRAPH = 1, 2   # p7_gpu_PipelineBatch flags: launch directly / capture and replay a HIP graph
XE, XN, XJ, XC = 0, 1, 2, 3
MOVE, LOOP = 0, 1
BM, MM, IM, DM, MD, MI, II, DD = range(8)
KP = {AMINO: 29, DNA: 18, RNA: 18}
K = {AMINO: 20, DNA: 4, RNA: 4}


class P7GpuError(RuntimeError):
    def __init__(self, status, where):
        self.status = status
        super().__init__(f"{where}: status {status}: {last_error()}")


class Tables(ctypes.Structure):
    """P7_GPU_TABLES (include/p7_gpu.h)."""
    _fields_ = [
        ("M", c_int), ("Kp", c_int), ("abctype", c_int),
        ("rbv", POINTER(c_uint8)),
        ("tbm_b", c_uint8), ("tec_b", c_uint8), ("tjb_b", c_uint8), ("base_b", c_uint8), ("bias_b", c_uint8),
        ("scale_b", c_float),
        ("rwv", POINTER(c_int16)), ("twv", POINTER(c_int16)),
        ("xw", (c_int16 * 2) * 4), ("scale_w", c_float), ("base_w", c_int16), ("ddbound_w", c_int16),
        ("rfv", POINTER(c_float)), ("tfv", POINTER(c_float)), ("xf", (c_float * 2) * 4),
        ("L", c_int), ("mode", c_int), ("nj", c_float), ("evparam", c_float * 6), ("compo", c_float * 20),
    ]


class Hits(ctypes.Structure):
    _fields_ = [(n, POINTER(c_float)) for n in ("nullsc", "usc", "filtersc", "vfsc", "fwdsc", "bcksc")] + \
               [("stage", POINTER(c_uint8)), ("nexpected", POINTER(c_float)),
                ("nregions", POINTER(c_int32)), ("nclustered", POINTER(c_int32))]


class Stats(ctypes.Structure):
    _fields_ = [(n, c_int64) for n in ("nmodels", "nseqs", "nres", "nnodes", "n_past_msv", "n_past_bias",
                                       "n_past_vit", "n_past_fwd", "n_msv_full")] + \
               [(n, c_double) for n in ("cells_msv", "cells_bias", "cells_vit", "cells_fwd", "cells_bck",
                                        "ms_msv", "ms_bias", "ms_vit", "ms_fwd", "ms_bck", "ms_total", "ms_ssv",
                                        "ms_domains")] + \
               [(n, c_int64) for n in ("n_regions", "n_clustered")] + [("ms_null2", c_double), ("ms_multidomain", c_double)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class Report(ctypes.Structure):
    """P7_GPU_REPORT: what the hit-list text needs beyond the run's results."""
    _fields_ = [("qname", c_char_p), ("qacc", c_char_p), ("M", c_int),
                ("tname", POINTER(c_char_p)), ("tacc", POINTER(c_char_p)), ("tdesc", POINTER(c_char_p)),
                ("Z", c_double), ("domZ", c_double), ("E", c_double), ("incE", c_double), ("domE", c_double),
                ("incdomE", c_double), ("textw", c_int), ("show_accessions", c_int), ("show_alignments", c_int)]


def _load():
    if not os.path.exists(LIBPATH):
        raise ImportError(f"{LIBPATH} is missing: build it with __graft_entry__.build() "
                          "(make -C biobench2_amd/csrc); there is no CPU fallback")
    lib = ctypes.CDLL(LIBPATH)
    vp, i, f, d, i64 = c_void_p, c_int, c_float, c_double, c_int64
    pu8, pf, pi, pi64, pi32 = POINTER(c_uint8), POINTER(c_float), POINTER(c_int), POINTER(c_int64), POINTER(c_int32)
    sig = {
        "p7_gpu_Init": (i, [i]), "p7_gpu_Synchronize": (i, []), "p7_gpu_LastError": (c_char_p, []),
        "p7_gpu_hmmfile_Read": (i, [c_char_p, i, POINTER(vp)]),
        "p7_gpu_hmm_Info": (i, [vp, pi, pi, c_char_p, i]), "p7_gpu_hmm_Destroy": (None, [vp]),
        "p7_gpu_bg_Create": (vp, [i]), "p7_gpu_bg_Destroy": (None, [vp]),
        "p7_gpu_bg_SetLength": (i, [vp, i]), "p7_gpu_bg_NullOne": (i, [vp, pu8, i, pf]),
        "p7_gpu_bg_SetFilter": (i, [vp, i, pf]), "p7_gpu_bg_FilterScore": (i, [vp, pu8, i, pf]),
        "p7_gpu_ProfileConfig": (i, [vp, vp, i, i, POINTER(vp)]),
        "p7_gpu_profile_ReconfigLength": (i, [vp, i]),
        "p7_gpu_profile_GetScores": (i, [vp, pf, pf, pf]), "p7_gpu_profile_Destroy": (None, [vp]),
        "p7_gpu_oprofile_Create": (vp, [i, i]), "p7_gpu_oprofile_Convert": (i, [vp, vp]),
        "p7_gpu_oprofile_Load": (i, [vp, POINTER(Tables)]), "p7_gpu_oprofile_GetTables": (i, [vp, POINTER(Tables)]),
        "p7_gpu_oprofile_ReconfigLength": (i, [vp, i]), "p7_gpu_oprofile_ReconfigMSVLength": (i, [vp, i]),
        "p7_gpu_oprofile_ReconfigRestLength": (i, [vp, i]), "p7_gpu_oprofile_Destroy": (None, [vp]),
        "p7_gpu_omx_Create": (vp, [i, i, i]), "p7_gpu_omx_GrowTo": (i, [vp, i, i, i]),
        "p7_gpu_omx_Xmx": (pf, [vp, pi, pf, pi]), "p7_gpu_omx_Destroy": (None, [vp]),
        "p7_gpu_MSVFilter": (i, [pu8, i, vp, vp, pf]), "p7_gpu_SSVFilter": (i, [pu8, i, vp, pf]),
        "p7_gpu_ViterbiFilter": (i, [pu8, i, vp, vp, pf]), "p7_gpu_ForwardParser": (i, [pu8, i, vp, vp, pf]),
        "p7_gpu_BackwardParser": (i, [pu8, i, vp, vp, vp, pf]),
        "p7_gpu_RegionTraceEnsemble": (i, [pu8, i, i, vp, i, c_uint32, pi32, pi32, i, pi32, pf, pf, pf, pi32, i,
                                           pi32]),
        "p7_gpu_seqdb_Create": (i, [pu8, pi64, pi32, i64, POINTER(vp)]),
        "p7_gpu_seqdb_NSeq": (i64, [vp]), "p7_gpu_seqdb_NRes": (i64, [vp]), "p7_gpu_seqdb_Destroy": (None, [vp]),
        "p7_gpu_MSVFilterBatch": (i, [vp, vp, pi64, i64, pf, pi]),
        "p7_gpu_ViterbiFilterBatch": (i, [vp, vp, pi64, i64, pf, pi]),
        "p7_gpu_ForwardParserBatch": (i, [vp, vp, pi64, i64, pf, pi]),
        "p7_gpu_BackwardParserBatch": (i, [vp, vp, pi64, i64, pf, pi]),
        "p7_gpu_FilterScoreBatch": (i, [vp, vp, pi64, i64, pf]),
        "p7_gpu_MSVFilterBatchNoAlloc": (i, [vp, vp, pi64, i64, pf, pi]),
        "p7_gpu_pipeline_Create": (vp, [d, d, d, i, i]), "p7_gpu_pipeline_Destroy": (None, [vp]),
        "p7_gpu_Pipeline": (i, [vp, vp, vp, POINTER(Hits)]),
        "p7_gpu_pipeline_GetStats": (i, [vp, POINTER(Stats)]),
        "p7_gpu_pipeline_SetDomains": (i, [vp, i]), "p7_gpu_pipeline_SetNull2": (i, [vp, i]),
        "p7_gpu_pipeline_SetPoolCap": (i, [vp, i64]),
        "p7_gpu_pipeline_GetRegionNull2": (i, [vp, POINTER(c_float), POINTER(c_int32), c_int64, POINTER(c_int64)]),
        "p7_gpu_pipeline_GetDomainAlignments": (i, [vp, POINTER(c_float), POINTER(c_int32), POINTER(c_int32),
                                                    POINTER(c_int32), POINTER(c_int32), c_int64, POINTER(c_int64)]),
        "p7_gpu_pipeline_GetTargetScores": (i, [vp, POINTER(c_int64), POINTER(c_float), POINTER(c_float), POINTER(c_float),
                                                POINTER(c_double), POINTER(c_int32), POINTER(c_int32), POINTER(c_int32),
                                                c_int64, POINTER(c_int64)]),
        "p7_gpu_pipeline_GetDomainScores": (i, [vp, POINTER(c_float), POINTER(c_float), POINTER(c_double), c_int64,
                                                POINTER(c_int64)]),
        "p7_gpu_pipeline_GetHitList": (i, [vp, d, d, d, d, d, POINTER(c_int64), POINTER(c_int32), POINTER(c_int32),
                                           POINTER(c_double), c_int64, POINTER(c_int64)]),
        "p7_gpu_pipeline_GetTargetCounts": (i, [vp, POINTER(c_float), POINTER(c_int32), POINTER(c_int32),
                                                POINTER(c_int32), POINTER(c_int32), c_int64, POINTER(c_int64)]),
        "p7_gpu_pipeline_TargetsText": (i, [vp, POINTER(Report), c_char_p, ctypes.c_size_t,
                                            POINTER(ctypes.c_size_t)]),
        "p7_gpu_pipeline_DomainsText": (i, [vp, POINTER(Report), c_char_p, ctypes.c_size_t,
                                            POINTER(ctypes.c_size_t)]),
        "p7_gpu_pipeline_TabularTargets": (i, [vp, POINTER(Report), i, c_char_p, ctypes.c_size_t,
                                               POINTER(ctypes.c_size_t)]),
        "p7_gpu_pipeline_TabularDomains": (i, [vp, POINTER(Report), i, c_char_p, ctypes.c_size_t,
                                               POINTER(ctypes.c_size_t)]),
        "p7_gpu_StatisticsText": (ctypes.c_size_t, [POINTER(Stats), d, d, d, d, i, d, i, c_char_p, ctypes.c_size_t]),
        "p7_gpu_pipeline_GetRegions": (i, [vp, POINTER(c_int64), POINTER(c_int32), POINTER(c_int32), POINTER(c_int32),
                                           POINTER(c_float), c_int64, POINTER(c_int64)]),
        "p7_gpu_ScoreThreshold": (f, [i, d, d, d]),
        "p7_gpu_oprofile_ReadPressed": (i, [c_char_p, c_char_p, i, POINTER(vp)]),
        "p7_gpu_seqdb_ReadFasta": (i, [c_char_p, i, POINTER(vp)]),
        "p7_gpu_seqdb_Name": (c_char_p, [vp, i64]),
        "p7_gpu_seqdb_Save": (i, [vp, c_char_p]),
        "p7_gpu_seqdb_Load": (i, [c_char_p, POINTER(vp)]),
        "p7_gpu_server_Create": (vp, [vp, i, ctypes.c_size_t, d, d, d, i]),
        "p7_gpu_server_Slots": (i, [vp]),
        "p7_gpu_server_SetCoalesce": (i, [vp, i]),
        "p7_gpu_server_Memory": (i, [vp, POINTER(ctypes.c_size_t), POINTER(ctypes.c_size_t)]),
        "p7_gpu_oprofile_DeviceBytes": (i, [vp, vp, POINTER(ctypes.c_size_t)]),
        "p7_gpu_server_Submit": (i, [vp, vp, POINTER(Hits), POINTER(c_int64)]),
        "p7_gpu_server_Wait": (i, [vp, c_int64, POINTER(Stats), POINTER(c_double)]),
        "p7_gpu_server_Destroy": (None, [vp]),
        "p7_gpu_PipelineBatch": (i, [vp, POINTER(vp), i, vp, POINTER(Hits), POINTER(Stats), i]),
        "p7_gpu_pipeline_BatchGraphed": (i, [vp]),
        "p7_gpu_Autotune": (i, [POINTER(vp), i, vp]),
        "p7_gpu_tune_Get": (i, [i, POINTER(c_int), POINTER(c_int), pf]),
        "p7_gpu_tune_Clear": (None, []),
        "p7_gpu_tune_Save": (i, [c_char_p]),
        "p7_gpu_tune_Load": (i, [c_char_p]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


def last_error():
    s = lib.p7_gpu_LastError()
    return s.decode() if s else ""


def _check(status, where, ok=(OK,)):
    if status not in ok:
        raise P7GpuError(status, where)
    return status


def _ptr(a, ctype):
    return a.ctypes.data_as(POINTER(ctype))


def score_threshold(which, mu, lam, P):
    """Smallest float bit score whose Gumbel (which=0) or exponential (which=1) P-value is <= P."""
    return lib.p7_gpu_ScoreThreshold(which, mu, lam, P)


def init(device=0):
    _check(lib.p7_gpu_Init(device), "p7_gpu_Init")


def dsq_of(residues):
    """Digital sequence with sentinels: dsq[0] = dsq[L+1] = 255, residues at 1..L."""
    r = np.asarray(residues, dtype=np.uint8)
    out = np.empty(r.size + 2, dtype=np.uint8)
    out[0] = out[-1] = 255
    out[1:-1] = r
    return out


class HMM:
    def __init__(self, path, which=0):
        h = c_void_p()
        _check(lib.p7_gpu_hmmfile_Read(str(path).encode(), which, ctypes.byref(h)), f"read {path}")
        self.h = h
        M, t = c_int(), c_int()
        name = ctypes.create_string_buffer(256)
        lib.p7_gpu_hmm_Info(h, ctypes.byref(M), ctypes.byref(t), name, 256)
        self.M, self.abctype, self.name = M.value, t.value, name.value.decode()

    def __del__(self):
        if getattr(self, "h", None):
            lib.p7_gpu_hmm_Destroy(self.h)
            self.h = None


class Background:
    def __init__(self, abctype):
        self.h = lib.p7_gpu_bg_Create(abctype)
        if not self.h:
            raise P7GpuError(EINVAL, "p7_gpu_bg_Create")
        self.abctype = abctype

    def set_length(self, L):
        _check(lib.p7_gpu_bg_SetLength(self.h, L), "bg_SetLength")

    def null_one(self, L):
        sc = c_float()
        _check(lib.p7_gpu_bg_NullOne(self.h, None, L, ctypes.byref(sc)), "bg_NullOne")
        return sc.value

    def set_filter(self, M, compo):
        c = np.ascontiguousarray(compo, dtype=np.float32)
        _check(lib.p7_gpu_bg_SetFilter(self.h, M, _ptr(c, c_float)), "bg_SetFilter")

    def filter_score(self, residues):
        dsq = dsq_of(residues)
        sc = c_float()
        _check(lib.p7_gpu_bg_FilterScore(self.h, _ptr(dsq, c_uint8), dsq.size - 2, ctypes.byref(sc)), "bg_FilterScore")
        return sc.value

    def __del__(self):
        if getattr(self, "h", None):
            lib.p7_gpu_bg_Destroy(self.h)
            self.h = None


class Profile:
    """P7_PROFILE: float log-odds scores from p7_gpu_ProfileConfig."""

    def __init__(self, hmm, bg, L=100, mode=LOCAL):
        h = c_void_p()
        _check(lib.p7_gpu_ProfileConfig(hmm.h, bg.h, L, mode, ctypes.byref(h)), "ProfileConfig")
        self.h, self.M, self.abctype, self.mode = h, hmm.M, hmm.abctype, mode

    def reconfig_length(self, L):
        _check(lib.p7_gpu_profile_ReconfigLength(self.h, L), "profile_ReconfigLength")

    def scores(self):
        """tsc[M,8] (MM,IM,DM,BM,MD,DD,MI,II), rsc[Kp,M+1,2] (MSC,ISC), xsc[4,2] ([E,N,J,C][LOOP,MOVE])."""
        Kp = KP[self.abctype]
        tsc = np.zeros((self.M, 8), np.float32)
        rsc = np.zeros((Kp, self.M + 1, 2), np.float32)
        xsc = np.zeros((4, 2), np.float32)
        _check(lib.p7_gpu_profile_GetScores(self.h, _ptr(tsc, c_float), _ptr(rsc, c_float), _ptr(xsc, c_float)), "GetScores")
        return tsc, rsc, xsc

    def __del__(self):
        if getattr(self, "h", None):
            lib.p7_gpu_profile_Destroy(self.h)
            self.h = None


class OProfile:
    """P7_OPROFILE: quantised tables; the device copy is built on first use."""

    def __init__(self, allocM, abctype):
        self.h = lib.p7_gpu_oprofile_Create(allocM, abctype)
        if not self.h:
            raise P7GpuError(EINVAL, "oprofile_Create")
        self.abctype = abctype

    @classmethod
    def from_profile(cls, gm):
        om = cls(gm.M, gm.abctype)
        _check(lib.p7_gpu_oprofile_Convert(gm.h, om.h), "oprofile_Convert")
        return om

    @classmethod
    def from_file(cls, path, which=0, L=100):
        """hmmsearch's per-query preparation: read, configure local multihit with L, convert."""
        hmm = HMM(path, which)
        bg = Background(hmm.abctype)
        gm = Profile(hmm, bg, L, LOCAL)
        om = cls.from_profile(gm)
        om.name, om.hmm, om.gm, om.bg = hmm.name, hmm, gm, bg
        return om

    @classmethod
    def from_pressed(cls, h3f_path, h3p_path, which=0):
        """Model `which` of a pressed database (.h3f/.h3p, p7_oprofile_ReadMSV + ReadRest)."""
        om = cls.__new__(cls)
        h = c_void_p()
        _check(lib.p7_gpu_oprofile_ReadPressed(str(h3f_path).encode(), str(h3p_path).encode(), which,
                                               ctypes.byref(h)), "oprofile_ReadPressed")
        om.h = h.value
        om.abctype = om.tables()["abctype"]
        return om

    def device_bytes(self, db):
        """Device memory of this profile's tables for searches of `db` (p7_gpu_oprofile_DeviceBytes)."""
        b = ctypes.c_size_t()
        _check(lib.p7_gpu_oprofile_DeviceBytes(self.h, db.h, ctypes.byref(b)), "oprofile_DeviceBytes")
        return b.value

    def tables(self):
        t = Tables()
        _check(lib.p7_gpu_oprofile_GetTables(self.h, ctypes.byref(t)), "GetTables")
        M, Kp = t.M, t.Kp
        d = {k: getattr(t, k) for k in ("M", "Kp", "abctype", "tbm_b", "tec_b", "tjb_b", "base_b", "bias_b", "scale_b",
                                         "scale_w", "base_w", "ddbound_w", "L", "mode", "nj")}
        d["rbv"] = np.ctypeslib.as_array(t.rbv, (Kp, M)).copy()
        d["rwv"] = np.ctypeslib.as_array(t.rwv, (Kp, M)).copy()
        d["twv"] = np.ctypeslib.as_array(t.twv, (8, M)).copy()
        d["rfv"] = np.ctypeslib.as_array(t.rfv, (Kp, M)).copy()
        d["tfv"] = np.ctypeslib.as_array(t.tfv, (8, M)).copy()
        d["xw"] = np.array([[t.xw[s][m] for m in range(2)] for s in range(4)], np.int16)
        d["xf"] = np.array([[t.xf[s][m] for m in range(2)] for s in range(4)], np.float32)
        d["evparam"] = np.array(list(t.evparam), np.float32)
        d["compo"] = np.array(list(t.compo), np.float32)
        return d

    def load(self, d):
        """Import tables (same keys as tables())."""
        keep = {}
        t = Tables()
        for k in ("M", "Kp", "abctype", "tbm_b", "tec_b", "tjb_b", "base_b", "bias_b", "scale_b",
                  "scale_w", "base_w", "ddbound_w", "L", "mode", "nj"):
            setattr(t, k, d[k])
        for k, ct, dt in (("rbv", c_uint8, np.uint8), ("rwv", c_int16, np.int16), ("twv", c_int16, np.int16),
                          ("rfv", c_float, np.float32), ("tfv", c_float, np.float32)):
            keep[k] = np.ascontiguousarray(d[k], dtype=dt)
            setattr(t, k, _ptr(keep[k], ct))
        for s in range(4):
            for m in range(2):
                t.xw[s][m] = int(d["xw"][s][m])
                t.xf[s][m] = float(d["xf"][s][m])
        for z in range(6):
            t.evparam[z] = float(d["evparam"][z])
        for z in range(20):
            t.compo[z] = float(d["compo"][z])
        _check(lib.p7_gpu_oprofile_Load(self.h, ctypes.byref(t)), "oprofile_Load")

    def reconfig_length(self, L):
        _check(lib.p7_gpu_oprofile_ReconfigLength(self.h, L), "oprofile_ReconfigLength")

    def __del__(self):
        if getattr(self, "h", None):
            lib.p7_gpu_oprofile_Destroy(self.h)
            self.h = None


class OMX:
    def __init__(self, allocM=1, allocL=0, allocXL=0):
        self.h = lib.p7_gpu_omx_Create(allocM, allocL, allocXL)
        if not self.h:
            raise P7GpuError(EMEM, "omx_Create")

    def xmx(self):
        L, ts, own = c_int(), c_float(), c_int()
        p = lib.p7_gpu_omx_Xmx(self.h, ctypes.byref(L), ctypes.byref(ts), ctypes.byref(own))
        return np.ctypeslib.as_array(p, ((L.value + 1) * 6,)).reshape(L.value + 1, 6).copy(), ts.value, own.value

    def __del__(self):
        if getattr(self, "h", None):
            lib.p7_gpu_omx_Destroy(self.h)
            self.h = None


def _single(fn, residues, om, *extra):
    dsq = dsq_of(residues)
    sc = c_float()
    st = fn(_ptr(dsq, c_uint8), dsq.size - 2, om.h, *extra, ctypes.byref(sc))
    return st, sc.value


def msv_filter(residues, om):
    st, sc = _single(lib.p7_gpu_MSVFilter, residues, om, None)
    _check(st, "MSVFilter", (OK, ERANGE))
    return st, sc


def ssv_filter(residues, om):
    st, sc = _single(lib.p7_gpu_SSVFilter, residues, om)
    _check(st, "SSVFilter", (OK, ERANGE, ENORESULT))
    return st, sc


def viterbi_filter(residues, om):
    st, sc = _single(lib.p7_gpu_ViterbiFilter, residues, om, None)
    _check(st, "ViterbiFilter", (OK, ERANGE))
    return st, sc


def forward_parser(residues, om, fwd=None):
    fwd = fwd or OMX()
    st, sc = _single(lib.p7_gpu_ForwardParser, residues, om, fwd.h)
    _check(st, "ForwardParser", (OK, ERANGE))
    return st, sc, fwd


def backward_parser(residues, om, fwd, bck=None):
    bck = bck or OMX()
    st, sc = _single(lib.p7_gpu_BackwardParser, residues, om, fwd.h, bck.h)
    _check(st, "BackwardParser", (OK, ERANGE))
    return st, sc, bck


def region_trace_ensemble(residues, Ltarget, om, nsamples=200, seed=42, matrix=False, segcap=0):
    """p7_gpu_RegionTraceEnsemble over one region's residues of a target of Ltarget residues:
    dict clusters [(ci, cj)] (1-based in the region), n2sc[Lr], nseg, and with matrix=True the
    Forward matrix mx[(Lr+1), Q, 3, 4] (M, D, I per striped lane), xmx[(Lr+1), 6] and the sampled
    domains segs[nseg, 5] (trace, i, j, k, m)."""
    r = np.asarray(residues, np.uint8)
    Lr = r.size
    dsq = dsq_of(r)
    cap = 64
    ci, cj = np.zeros(cap, np.int32), np.zeros(cap, np.int32)
    n2 = np.zeros(Lr, np.float32)
    nc, ns = c_int32(), c_int32()
    Q = max(2, (om.tables()["M"] - 1) // 4 + 1)
    sc = segcap if segcap > 0 else 6144
    mx = np.zeros((Lr + 1) * 12 * Q, np.float32) if matrix else None
    xmx = np.zeros((Lr + 1) * 6, np.float32) if matrix else None
    segs = np.zeros(5 * sc, np.int32) if matrix else None
    _check(lib.p7_gpu_RegionTraceEnsemble(_ptr(dsq, c_uint8), Lr, int(Ltarget), om.h, nsamples, seed, _ptr(ci, c_int32),
                                          _ptr(cj, c_int32), cap, ctypes.byref(nc), _ptr(n2, c_float),
                                          _ptr(mx, c_float) if matrix else None, _ptr(xmx, c_float) if matrix else None,
                                          _ptr(segs, c_int32) if matrix else None, segcap, ctypes.byref(ns)),
           "RegionTraceEnsemble")
    out = {"clusters": [(int(ci[c]), int(cj[c])) for c in range(min(nc.value, cap))], "n2sc": n2, "nseg": ns.value}
    if matrix:
        out["mx"] = mx.reshape(Lr + 1, Q, 3, 4)
        out["xmx"] = xmx.reshape(Lr + 1, 6)
        out["segs"] = segs[:5 * min(ns.value, sc)].reshape(-1, 5)
    return out


class SeqDB:
    """HBM-resident digital database. residues: uint8 concatenation; offsets/lengths per sequence."""

    def __init__(self, residues, offsets, lengths):
        self.residues = np.ascontiguousarray(residues, np.uint8)
        self.offsets = np.ascontiguousarray(offsets, np.int64)
        self.lengths = np.ascontiguousarray(lengths, np.int32)
        h = c_void_p()
        _check(lib.p7_gpu_seqdb_Create(_ptr(self.residues, c_uint8), _ptr(self.offsets, c_int64),
                                       _ptr(self.lengths, c_int32), self.lengths.size, ctypes.byref(h)), "seqdb_Create")
        self.h = h
        self.nseq = int(self.lengths.size)
        self.nres = int(lib.p7_gpu_seqdb_NRes(h))

    @classmethod
    def from_list(cls, seqs):
        lengths = np.array([len(s) for s in seqs], np.int32)
        offsets = np.zeros(len(seqs), np.int64)
        if len(seqs) > 1:
            offsets[1:] = np.cumsum(lengths[:-1])
        res = np.concatenate([np.asarray(s, np.uint8) for s in seqs]) if len(seqs) else np.zeros(0, np.uint8)
        if res.size == 0:
            res = np.zeros(1, np.uint8)
        return cls(res, offsets, lengths)

    @classmethod
    def _wrap(cls, h):
        db = cls.__new__(cls)
        db.h = h
        db.nseq = int(lib.p7_gpu_seqdb_NSeq(h))
        db.nres = int(lib.p7_gpu_seqdb_NRes(h))
        db.residues = db.offsets = db.lengths = None       # resident on the device only
        return db

    @classmethod
    def from_fasta(cls, path, abctype=AMINO):
        """Native FASTA reader straight into HBM (p7_gpu_seqdb_ReadFasta)."""
        h = c_void_p()
        _check(lib.p7_gpu_seqdb_ReadFasta(str(path).encode(), abctype, ctypes.byref(h)), "seqdb_ReadFasta")
        return cls._wrap(h)

    @classmethod
    def load(cls, path):
        """A database image written by save()."""
        h = c_void_p()
        _check(lib.p7_gpu_seqdb_Load(str(path).encode(), ctypes.byref(h)), "seqdb_Load")
        return cls._wrap(h)

    def save(self, path):
        _check(lib.p7_gpu_seqdb_Save(self.h, str(path).encode()), "seqdb_Save")

    def name(self, i):
        s = lib.p7_gpu_seqdb_Name(self.h, i)
        return s.decode() if s is not None else None

    def seq(self, i):
        o, L = int(self.offsets[i]), int(self.lengths[i])
        return self.residues[o:o + L]

    def __del__(self):
        if getattr(self, "h", None):
            lib.p7_gpu_seqdb_Destroy(self.h)
            self.h = None


def _batch(fn, om, db, idx, name, with_status=True):
    if idx is None:
        n = db.nseq
        ip = None
    else:
        idx = np.ascontiguousarray(idx, np.int64)
        n = idx.size
        ip = _ptr(idx, c_int64)
    sc = np.zeros(n, np.float32)
    if with_status:
        st = np.zeros(n, np.int32)
        _check(fn(om.h, db.h, ip, n, _ptr(sc, c_float), _ptr(st, c_int)), name)
        return sc, st
    _check(fn(om.h, db.h, ip, n, _ptr(sc, c_float)), name)
    return sc


def msv_batch(om, db, idx=None):
    return _batch(lib.p7_gpu_MSVFilterBatch, om, db, idx, "MSVFilterBatch")


def msv_batch_noalloc(om, db, idx=None):
    """p7_gpu_MSVFilterBatchNoAlloc: MSVFilterBatch with the library's static
    device scratch (no allocation; the profile's device copy must exist)."""
    return _batch(lib.p7_gpu_MSVFilterBatchNoAlloc, om, db, idx, "MSVFilterBatchNoAlloc")


def viterbi_batch(om, db, idx=None):
    return _batch(lib.p7_gpu_ViterbiFilterBatch, om, db, idx, "ViterbiFilterBatch")


def forward_batch(om, db, idx=None):
    return _batch(lib.p7_gpu_ForwardParserBatch, om, db, idx, "ForwardParserBatch")


def backward_batch(om, db, idx=None):
    return _batch(lib.p7_gpu_BackwardParserBatch, om, db, idx, "BackwardParserBatch")


def filter_score_batch(om, db, idx=None):
    return _batch(lib.p7_gpu_FilterScoreBatch, om, db, idx, "FilterScoreBatch", with_status=False)


def _hits(n, scores, domains=False):
    out = {"stage": np.zeros(n, np.uint8)}
    hits = Hits()
    hits.stage = _ptr(out["stage"], c_uint8)
    if scores:
        for k in ("nullsc", "usc", "filtersc", "vfsc", "fwdsc", "bcksc"):
            out[k] = np.zeros(n, np.float32)
            setattr(hits, k, _ptr(out[k], c_float))
    if domains:
        out["nexpected"] = np.zeros(n, np.float32)
        hits.nexpected = _ptr(out["nexpected"], c_float)
        for k in ("nregions", "nclustered"):
            out[k] = np.zeros(n, np.int32)
            setattr(hits, k, _ptr(out[k], c_int32))
    return out, hits


class Server:
    """Queries against a resident database, admitted as they arrive (p7_gpu_server_*)."""

    def __init__(self, db, nslots=4, hbm_budget=0, F1=0.02, F2=1e-3, F3=1e-5, do_biasfilter=True):
        self.db = db
        self.h = lib.p7_gpu_server_Create(db.h, int(nslots), int(hbm_budget), F1, F2, F3, int(do_biasfilter))
        if not self.h:
            raise P7GpuError(EMEM, "server_Create")
        self.pending = {}

    @property
    def slots(self):
        return lib.p7_gpu_server_Slots(self.h)

    def set_coalesce(self, n):
        """Queries one slot may take as one batch (1..4)."""
        _check(lib.p7_gpu_server_SetCoalesce(self.h, int(n)), "server_SetCoalesce")

    def memory(self):
        """(bytes set aside per slot, bytes left for the profiles of queries in flight)."""
        a, b = ctypes.c_size_t(), ctypes.c_size_t()
        _check(lib.p7_gpu_server_Memory(self.h, ctypes.byref(a), ctypes.byref(b)), "server_Memory")
        return a.value, b.value

    def submit(self, om, scores=True):
        out, hits = _hits(self.db.nseq, scores)
        t = c_int64()
        _check(lib.p7_gpu_server_Submit(self.h, om.h, ctypes.byref(hits), ctypes.byref(t)), "server_Submit")
        self.pending[t.value] = (om, out, hits)      # keeps the arrays and the profile alive
        return t.value

    def wait(self, ticket):
        st, lat = Stats(), c_double()
        status = lib.p7_gpu_server_Wait(self.h, ticket, ctypes.byref(st), ctypes.byref(lat))
        om, out, hits = self.pending.pop(ticket)
        _check(status, "server_Wait")
        return out, st.as_dict(), lat.value

    def close(self):
        if getattr(self, "h", None):
            lib.p7_gpu_server_Destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


def statistics_text(st, F, Z, domZ, z_by_option=False, domz_by_option=False):
    """p7_gpu_StatisticsText for a Stats structure."""
    n = lib.p7_gpu_StatisticsText(ctypes.byref(st), F[0], F[1], F[2], float(Z), int(z_by_option), float(domZ),
                                  int(domz_by_option), None, 0)
    buf = ctypes.create_string_buffer(n + 1)
    lib.p7_gpu_StatisticsText(ctypes.byref(st), F[0], F[1], F[2], float(Z), int(z_by_option), float(domZ),
                              int(domz_by_option), buf, n + 1)
    return buf.value.decode()


class Pipeline:
    """P7_PIPELINE thresholds and counters; run() is p7_Pipeline over a whole database."""

    def __init__(self, F1=0.02, F2=1e-3, F3=1e-5, do_biasfilter=True, do_max=False, domains=False, null2=True):
        self.h = lib.p7_gpu_pipeline_Create(F1, F2, F3, int(do_biasfilter), int(do_max))
        self.domains = bool(domains)
        _check(lib.p7_gpu_pipeline_SetDomains(self.h, int(self.domains)), "pipeline_SetDomains")
        _check(lib.p7_gpu_pipeline_SetNull2(self.h, int(bool(null2))), "pipeline_SetNull2")

    def set_pool_cap(self, floats):
        """p7_gpu_pipeline_SetPoolCap: each domain-definition pool at most floats floats (0: no cap)."""
        _check(lib.p7_gpu_pipeline_SetPoolCap(self.h, int(floats)), "pipeline_SetPoolCap")

    def run(self, om, db, scores=True):
        """p7_gpu_Pipeline; with domains=True (and scores) the outputs also hold
        nexpected, nregions and nclustered of the F3 survivors; regions() has
        the regions either way."""
        out, hits = _hits(db.nseq, scores, self.domains and scores)
        _check(lib.p7_gpu_Pipeline(self.h, om.h, db.h, ctypes.byref(hits)), "Pipeline")
        return out

    def regions(self):
        """[(sequence index, i, j, kind, envelope score)] of the last run with domain decoding, by
        sequence then start: kind 0 a single-domain region (one envelope), 1 a multidomain region,
        2 an envelope its sampled traces' clusters define (after the region, in start order)."""
        n = c_int64()
        _check(lib.p7_gpu_pipeline_GetRegions(self.h, None, None, None, None, None, 0, ctypes.byref(n)), "GetRegions")
        seq = np.zeros(n.value, np.int64)
        i, j, m = (np.zeros(n.value, np.int32) for _ in range(3))
        sc = np.zeros(n.value, np.float32)
        _check(lib.p7_gpu_pipeline_GetRegions(self.h, _ptr(seq, c_int64), _ptr(i, c_int32), _ptr(j, c_int32),
                                              _ptr(m, c_int32), _ptr(sc, c_float), n.value, ctypes.byref(n)),
               "GetRegions")
        return [(int(seq[k]), int(i[k]), int(j[k]), int(m[k]), float(sc[k])) for k in range(n.value)]

    def region_null2(self):
        """(domcorr[n] nats, status[n]) of the last run's regions, in regions() order: status 0
        computed (by expectation for kind 0, from the region's sampled traces for kind 2), 16 the
        decoding overflowed, 2 a multidomain region, whose domcorr is the sum of the sampled null2
        scores of all its residues."""
        n = c_int64()
        _check(lib.p7_gpu_pipeline_GetRegionNull2(self.h, None, None, 0, ctypes.byref(n)), "GetRegionNull2")
        dc = np.zeros(n.value, np.float32)
        st = np.zeros(n.value, np.int32)
        _check(lib.p7_gpu_pipeline_GetRegionNull2(self.h, _ptr(dc, c_float), _ptr(st, c_int32), n.value,
                                                  ctypes.byref(n)), "GetRegionNull2")
        return dc, st

    def domain_alignments(self):
        """Per region (regions() order), the optimal accuracy alignment: dict of arrays oasc (expected
        correctly aligned residues; NaN where not computed), hmmfrom, hmmto, alifrom, alito (target
        coordinates; 0 where not computed)."""
        n = c_int64()
        _check(lib.p7_gpu_pipeline_GetDomainAlignments(self.h, None, None, None, None, None, 0, ctypes.byref(n)),
               "GetDomainAlignments")
        out = {"oasc": np.zeros(n.value, np.float32)}
        for k in ("hmmfrom", "hmmto", "alifrom", "alito"):
            out[k] = np.zeros(n.value, np.int32)
        _check(lib.p7_gpu_pipeline_GetDomainAlignments(self.h, _ptr(out["oasc"], c_float), _ptr(out["hmmfrom"], c_int32),
                                                       _ptr(out["hmmto"], c_int32), _ptr(out["alifrom"], c_int32),
                                                       _ptr(out["alito"], c_int32), n.value, ctypes.byref(n)),
               "GetDomainAlignments")
        return out

    def target_scores(self):
        """Per target with regions (sequence order): dict of arrays seq, score, pre_score,
        sum_score (bits), lnP, ndom, best (index within the target's regions() entries, -1
        none) and status (0 scored; 1 no envelope, not a hit)."""
        n = c_int64()
        _check(lib.p7_gpu_pipeline_GetTargetScores(self.h, *([None] * 8), 0, ctypes.byref(n)), "GetTargetScores")
        o = {"seq": np.zeros(n.value, np.int64), "score": np.zeros(n.value, np.float32),
             "pre_score": np.zeros(n.value, np.float32), "sum_score": np.zeros(n.value, np.float32),
             "lnP": np.zeros(n.value, np.float64), "ndom": np.zeros(n.value, np.int32),
             "best": np.zeros(n.value, np.int32), "status": np.zeros(n.value, np.int32)}
        _check(lib.p7_gpu_pipeline_GetTargetScores(
            self.h, _ptr(o["seq"], c_int64), _ptr(o["score"], c_float), _ptr(o["pre_score"], c_float),
            _ptr(o["sum_score"], c_float), _ptr(o["lnP"], c_double), _ptr(o["ndom"], c_int32),
            _ptr(o["best"], c_int32), _ptr(o["status"], c_int32), n.value, ctypes.byref(n)), "GetTargetScores")
        return o

    def hit_list(self, Z, E=10.0, incE=0.01, domE=10.0, incdomE=0.01):
        """(order, flags, dflags, domZ): the scored targets ranked (indices into target_scores()
        arrays), their reported (1) / included (2) flags, per-region flags, and domZ."""
        n, dz = c_int64(), c_double()
        _check(lib.p7_gpu_pipeline_GetHitList(self.h, Z, E, incE, domE, incdomE, None, None, None, None, 0,
                                              ctypes.byref(n)), "GetHitList")
        nr = len(self.regions())
        order, flags = np.zeros(n.value, np.int64), np.zeros(n.value, np.int32)
        dflags = np.zeros(max(nr, 1), np.int32)
        _check(lib.p7_gpu_pipeline_GetHitList(self.h, Z, E, incE, domE, incdomE, _ptr(order, c_int64),
                                              _ptr(flags, c_int32), _ptr(dflags, c_int32), ctypes.byref(dz), n.value,
                                              ctypes.byref(n)), "GetHitList")
        return order, flags, dflags[:nr], dz.value

    def target_counts(self):
        """Per target (target_scores() order): dict of arrays nexpected, nregions, nclustered,
        noverlaps, nenvelopes (P7_DOMAINDEF's counters)."""
        n = c_int64()
        _check(lib.p7_gpu_pipeline_GetTargetCounts(self.h, *([None] * 5), 0, ctypes.byref(n)), "GetTargetCounts")
        o = {"nexpected": np.zeros(n.value, np.float32)}
        for k in ("nregions", "nclustered", "noverlaps", "nenvelopes"):
            o[k] = np.zeros(n.value, np.int32)
        _check(lib.p7_gpu_pipeline_GetTargetCounts(
            self.h, _ptr(o["nexpected"], c_float), _ptr(o["nregions"], c_int32), _ptr(o["nclustered"], c_int32),
            _ptr(o["noverlaps"], c_int32), _ptr(o["nenvelopes"], c_int32), n.value, ctypes.byref(n)),
            "GetTargetCounts")
        return o

    def report_text(self, kind, qname, M, names, accs=None, descs=None, qacc=None, Z=None, domZ=-1.0, E=10.0,
                    incE=0.01, domE=10.0, incdomE=0.01, textw=120, show_accessions=False, show_header=True,
                    show_alignments=False):
        """The last run's hit list as hmmsearch writes it: kind "targets" (p7_tophits_Targets),
        "domains" (p7_tophits_Domains without alignments), "tblout" or "domtblout". names (and
        accs, descs: None or lists with None / "" for none) are indexed by database sequence;
        Z defaults to the number of names."""
        enc = lambda v: None if v is None else str(v).encode()
        arr = lambda vals: None if vals is None else (c_char_p * len(vals))(*[enc(v) for v in vals])
        tn, ta, td = arr(names), arr(accs), arr(descs)
        r = Report(enc(qname), enc(qacc), int(M), ctypes.cast(tn, POINTER(c_char_p)),
                   ctypes.cast(ta, POINTER(c_char_p)) if ta is not None else None,
                   ctypes.cast(td, POINTER(c_char_p)) if td is not None else None,
                   float(len(names) if Z is None else Z), float(domZ), E, incE, domE, incdomE, int(textw),
                   int(bool(show_accessions)), int(bool(show_alignments)))
        fn = {"targets": lambda b, c, n: lib.p7_gpu_pipeline_TargetsText(self.h, ctypes.byref(r), b, c, n),
              "domains": lambda b, c, n: lib.p7_gpu_pipeline_DomainsText(self.h, ctypes.byref(r), b, c, n),
              "tblout": lambda b, c, n: lib.p7_gpu_pipeline_TabularTargets(self.h, ctypes.byref(r), int(show_header),
                                                                           b, c, n),
              "domtblout": lambda b, c, n: lib.p7_gpu_pipeline_TabularDomains(self.h, ctypes.byref(r),
                                                                              int(show_header), b, c, n)}[kind]
        n = ctypes.c_size_t()
        _check(fn(None, 0, ctypes.byref(n)), kind)
        buf = ctypes.create_string_buffer(n.value + 1)
        _check(fn(buf, n.value + 1, ctypes.byref(n)), kind)
        return buf.value.decode()

    def domain_scores(self):
        """(bitscore, dombias nats, lnP) arrays per region, in regions() order."""
        n = c_int64()
        _check(lib.p7_gpu_pipeline_GetDomainScores(self.h, None, None, None, 0, ctypes.byref(n)), "GetDomainScores")
        b, d, p = np.zeros(n.value, np.float32), np.zeros(n.value, np.float32), np.zeros(n.value, np.float64)
        _check(lib.p7_gpu_pipeline_GetDomainScores(self.h, _ptr(b, c_float), _ptr(d, c_float), _ptr(p, c_double),
                                                   n.value, ctypes.byref(n)), "GetDomainScores")
        return b, d, p

    def run_batch(self, oms, db, scores=True, graph=False):
        """p7_gpu_PipelineBatch: every profile of `oms` over the database in one
        call (one launch per kernel shape for the ungapped pass, concurrent
        per-profile chains; captured once and replayed as a HIP graph with
        graph=True). Returns one output dict and one stats dict per profile."""
        n = len(oms)
        arr = (c_void_p * n)(*[om.h for om in oms])
        outs, hits = [], (Hits * n)()
        for j in range(n):
            o, h = _hits(db.nseq, scores)
            outs.append(o)
            hits[j] = h
        st = (Stats * n)()
        _check(lib.p7_gpu_PipelineBatch(self.h, arr, n, db.h, hits, st, BATCH_GRAPH if graph else 0), "PipelineBatch")
        return outs, [st[j].as_dict() for j in range(n)]

    @property
    def batch_graphed(self):
        return bool(lib.p7_gpu_pipeline_BatchGraphed(self.h))

    def stats(self):
        st = Stats()
        _check(lib.p7_gpu_pipeline_GetStats(self.h, ctypes.byref(st)), "GetStats")
        return st.as_dict()

    def statistics_text(self, F=(0.02, 1e-3, 1e-5), Z=None, domZ=0.0, z_by_option=False, domz_by_option=False):
        """p7_pli_Statistics' summary text for this pipeline's counters (Z defaults to the
        number of targets searched)."""
        st = Stats()
        _check(lib.p7_gpu_pipeline_GetStats(self.h, ctypes.byref(st)), "GetStats")
        return statistics_text(st, F, st.nseqs if Z is None else Z, domZ, z_by_option, domz_by_option)

    def __del__(self):
        if getattr(self, "h", None):
            lib.p7_gpu_pipeline_Destroy(self.h)
            self.h = None


def autotune(oms, db):
    """p7_gpu_Autotune: measure the ungapped-pass shapes for the models' length classes."""
    arr = (c_void_p * len(oms))(*[om.h for om in oms])
    _check(lib.p7_gpu_Autotune(arr, len(oms), db.h), "Autotune")


def tune_get(M):
    """(lanes per target, queue grab, best us per shape [16, 32, 64, 8 lanes], tuned?) for models of length M."""
    lanes, grab = c_int(), c_int()
    us = np.zeros(4, np.float32)
    t = lib.p7_gpu_tune_Get(int(M), ctypes.byref(lanes), ctypes.byref(grab), _ptr(us, c_float))
    return lanes.value, grab.value, us, bool(t)


def tune_clear():
    lib.p7_gpu_tune_Clear()


def tune_save(path):
    _check(lib.p7_gpu_tune_Save(str(path).encode()), "tune_Save")


def tune_load(path):
    _check(lib.p7_gpu_tune_Load(str(path).encode()), "tune_Load")
