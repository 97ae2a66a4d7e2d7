"""Reader for the reference's pressed-profile files (.h3f / .h3p, format 3/f).

The record layout is the one HMMER/src/impl_sse/io.c:89-176 writes: the MSV
part (.h3f) and the rest of the optimized profile (.h3p), with the Farrar
striped vectors of the SSE build (16 uchars, 8 int16 or 4 floats per vector;
element z of vector q holds node k = q + 1 + z*Q). This module reads them as
plain data and removes the striping, giving tables in the same unstriped
layout as P7_GPU_TABLES, so they can be compared with the tables the GPU
library converts from the same .hmm file.
"""
import struct

import numpy as np

V3F_FMAGIC = 0xb3e6e6f3
V3F_PMAGIC = 0xb3e6f0f3
KP = {1: 18, 2: 18, 3: 29}


def _nq(M, w):
    return max(2, (M - 1) // w + 1)


class _Buf:
    def __init__(self, data):
        self.d, self.p = data, 0

    def take(self, n):
        b = self.d[self.p:self.p + n]
        if len(b) != n:
            raise ValueError("truncated pressed file")
        self.p += n
        return b

    def u32(self):
        return struct.unpack("<I", self.take(4))[0]

    def i32(self):
        return struct.unpack("<i", self.take(4))[0]

    def i16(self):
        return struct.unpack("<h", self.take(2))[0]

    def u8(self):
        return self.take(1)[0]

    def f32(self):
        return struct.unpack("<f", self.take(4))[0]

    def arr(self, dtype, n):
        a = np.frombuffer(self.take(np.dtype(dtype).itemsize * n), dtype=dtype)
        return a.copy()

    def done(self):
        return self.p >= len(self.d)


def _destripe(striped, M, Q, W):
    """striped: [..., Q, W] -> [..., M] with node k = q + 1 + z*Q."""
    out = np.zeros(striped.shape[:-2] + (M,), striped.dtype)
    for q in range(Q):
        for z in range(W):
            k = q + 1 + z * Q
            if k <= M:
                out[..., k - 1] = striped[..., q, z]
    return out


def read_pressed(h3f_path, h3p_path):
    """List of dicts, one per model, with unstriped tables and scalars."""
    f = _Buf(open(h3f_path, "rb").read())
    p = _Buf(open(h3p_path, "rb").read())
    models = []
    while not f.done():
        m = {}
        if f.u32() != V3F_FMAGIC:
            raise ValueError("bad .h3f magic")
        M = f.i32()
        abctype = f.i32()
        n = f.i32()
        m["name"] = f.take(n + 1)[:-1].decode()
        m["max_length"] = f.i32()
        m["tbm_b"], m["tec_b"], m["tjb_b"] = f.u8(), f.u8(), f.u8()
        m["scale_b"] = f.f32()
        m["base_b"], m["bias_b"] = f.u8(), f.u8()
        Kp = KP[abctype]
        Q16, Q8, Q4 = _nq(M, 16), _nq(M, 8), _nq(M, 4)
        sbv = f.arr(np.uint8, Kp * (Q16 + 17) * 16).reshape(Kp, Q16 + 17, 16)
        rbv = f.arr(np.uint8, Kp * Q16 * 16).reshape(Kp, Q16, 16)
        m["evparam"] = f.arr(np.float32, 6)
        f.arr(np.int64, 3)                                  # disk offsets
        m["compo"] = f.arr(np.float32, 20)
        if f.u32() != V3F_FMAGIC:
            raise ValueError("bad .h3f sentinel")
        m["M"], m["abctype"], m["Kp"] = M, abctype, Kp
        m["rbv"] = _destripe(rbv, M, Q16, 16)
        m["sbv_striped"] = sbv

        if p.u32() != V3F_PMAGIC:
            raise ValueError("bad .h3p magic")
        if p.i32() != M or p.i32() != abctype:
            raise ValueError("h3f/h3p records disagree")
        n = p.i32()
        p.take(n + 1)
        n = p.i32()
        if n > 0:
            p.take(n + 1)
        n = p.i32()
        if n > 0:
            p.take(n + 1)
        for _ in range(4):                                  # rf, mm, cs, consensus
            p.take(M + 2)
        twv = p.arr(np.int16, 8 * Q8 * 8).reshape(8 * Q8, 8)
        rwv = p.arr(np.int16, Kp * Q8 * 8).reshape(Kp, Q8, 8)
        m["xw"] = p.arr(np.int16, 8).reshape(4, 2)
        m["scale_w"] = p.f32()
        m["base_w"], m["ddbound_w"] = p.i16(), p.i16()
        m["ncj_roundoff"] = p.f32()
        tfv = p.arr(np.float32, 8 * Q4 * 4).reshape(8 * Q4, 4)
        rfv = p.arr(np.float32, Kp * Q4 * 4).reshape(Kp, Q4, 4)
        m["xf"] = p.arr(np.float32, 8).reshape(4, 2)
        m["cutoff"] = p.arr(np.float32, 6)
        m["nj"] = p.f32()
        m["mode"] = p.i32()
        m["L"] = p.i32()
        if p.u32() != V3F_PMAGIC:
            raise ValueError("bad .h3p sentinel")
        m["rwv"] = _destripe(rwv, M, Q8, 8)
        m["rfv"] = _destripe(rfv, M, Q4, 4)
        # the striped vectors as the file holds them (for a de-striping shim to read)
        m["rbv_striped"], m["rwv_striped"], m["twv_striped"] = rbv, rwv, twv
        m["rfv_striped"], m["tfv_striped"] = rfv, tfv
        m["Q"] = (Q16, Q8, Q4)
        # transitions: per q, seven vectors BM MM IM DM MD MI II, then Q DD vectors
        tw = np.zeros((8, Q8, 8), np.int16)
        tf = np.zeros((8, Q4, 4), np.float32)
        for q in range(Q8):
            for t in range(7):
                tw[t, q] = twv[q * 7 + t]
            tw[7, q] = twv[7 * Q8 + q]
        for q in range(Q4):
            for t in range(7):
                tf[t, q] = tfv[q * 7 + t]
            tf[7, q] = tfv[7 * Q4 + q]
        m["twv"] = _destripe(tw, M, Q8, 8)
        m["tfv"] = _destripe(tf, M, Q4, 4)
        models.append(m)
    return models
