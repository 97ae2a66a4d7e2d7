// gfx950 kernels of the MSV stage's ungapped pass (the first of its two
// passes, HMMER/src/impl_sse/msvfilter.c:76-210 and ssvfilter.c:717-926).
//
// While the J state is unused, xB is a per-target constant, and with
// Z = max(M, xB) - xB the row update of the MSV recurrence becomes
// Z(i,k) = max(Z(i-1,k-1) + r, 0), r = bias - rbv. Holding Z - 32768 in int16,
// the floor at 0 is exactly the saturation of v_pk_add_i16 ... clamp, so a cell
// costs one packed add and half of a three-input packed minimum (the running
// maximum, below). The pass records max Z per target; k_msv_classify decides
// from it which targets are finished, overflow, or need the full recurrence.
//
// A target is spread over G = 8, 16, 32 or 64 lanes (8, 4, 2 or 1 targets per
// wave), lane z holding model positions z*2ND+1 .. z*2ND+2ND as ND int16 pairs
// (dword d: positions d and d + ND). Which G serves a model best is measured by
// the runtime's tuner; every shape gives the same result. Small models gain
// from 8 lanes: twice the positions per lane for the same per-row overhead.
//
// The kernel serves one or many profiles of one shape: each workgroup stages
// a profile's table in LDS, its waves take groups of targets from that
// profile's work queue, and when the queue is empty the workgroup moves on to
// the next profile with work left. A batch of profiles thus runs as one
// persistent launch, and one profile is the batch of one.
#include "p7g_devfn.h"

#ifndef SSV_BLOCK
#define SSV_BLOCK 512
#endif

// Running maximum of Z over a row, two packed pairs per instruction. A value
// Z - 32768 held as an int16 has the bit pattern 0x8000 + Z (Z < 32768), which
// read as a half-precision number is -(Z * 2^-24) while Z < 1024 (zero and
// subnormals, kept exact: the kernels run with fp16 denormals preserved) and
// keeps decreasing as Z grows. So the IEEE minimum of those halves is the
// pattern of the largest Z, and v_pk_minimum3_f16 folds two pairs into the
// accumulator at once. Values that do not order this way (Z >= 31745: NaN or
// positive patterns) only occur after Z has passed every overflow threshold on
// the same diagonal, and the minimum then still decodes to an overflowing Z.
// (Written as one instruction: the compiler would otherwise regroup the
// minimum into pairwise trees, one and a half instructions per two pairs.)
__device__ __forceinline__ int zmax3(int acc, int a, int b)
{
  int r;
  asm("v_pk_minimum3_f16 %0, %1, %2, %3" : "=v"(r) : "v"(acc), "v"(a), "v"(b));
  return r;
}
__device__ __forceinline__ int zmax2(int acc, int a) { return zmax3(acc, a, a); }


// The row shift's lane move: lane z of a target takes lane z-1's last dword;
// the first lane of each target keeps `fill`. G = 32 needs two moves: lanes
// 16 and 48 take lanes 15 and 47 by row_bcast:15, then row_shr:1 fills the
// rest of each 16-lane row, leaving lanes 0 and 32 with the fill. With G = 8
// the second target of a 16-lane row takes the first one's last position: the
// 8-lane tables leave that position past M (score -32768), so it is at the
// floor on every row, which is the fill.
template <int G>
__device__ __forceinline__ int lane_in(int fill, int v)
{
  if (G == 8 || G == 16) return __builtin_amdgcn_update_dpp(fill, v, ROW_SHR1, 0xf, 0xf, false);
  if (G == 64) return __builtin_amdgcn_update_dpp(fill, v, WAVE_SHR1, 0xf, 0xf, false);
  const int b = __builtin_amdgcn_update_dpp(fill, v, ROW_BCAST15, 0xa, 0xf, false);
  return __builtin_amdgcn_update_dpp(b, v, ROW_SHR1, 0xf, 0xf, false);
}

// BND: one chunk of a model too long for one LDS table (G = 64, one target
// per wave). The chunk's first position takes its diagonal predecessor from
// bin (bin[off + r] = Z of the previous chunk's last position after row r;
// row 0 takes the floor), and the chunk's last position after row r goes to
// bout[off + r]. Both are indexed like the residues, so each target owns the
// entries of its rows. The chunk's maximum is folded into zmax when bin is
// given.
template <int G, int ND, bool BND>
__global__ __launch_bounds__(SSV_BLOCK)
void k_ssv(const SsvJobs jobs, int njobs, int tbl_dwords,
           const uint8_t *__restrict__ res, const int64_t *__restrict__ off, const int32_t *__restrict__ len,
           const int32_t *__restrict__ slots, int64_t n, uint16_t *__restrict__ zmax, int64_t zstride,
           int32_t *queues, int grab, const uint16_t *__restrict__ bin, uint16_t *__restrict__ bout,
           const int2 *__restrict__ segs, int64_t nsegs, int64_t nseg_slots)
{
  static_assert(! BND || G == 64, "chunked models take one target per wave");
  // A row segment starting at row r0 is computed from row r0 - WARM on, from
  // the floor: every diagonal of the ungapped recurrence reaching a row >= r0
  // starts at most M - 1 <= WARM rows earlier, so the segment's own rows are
  // exact and the rows before them lower bounds, and the maximum over all
  // segments is the target's.
  constexpr int WARM = 2 * ND * G;
  constexpr int NQ   = ND / 4;                        // 16-byte reads per row
  constexpr int NT   = ND % 4;                        // 4-byte reads per row
  constexpr int NQa  = NQ ? NQ : 1, NTa = NT ? NT : 1;
  constexpr int TAIL0 = P7G_NROWS * NQ * G * 4;       // dword offset of the tail region
  constexpr int TPW  = 64 / G;
  constexpr int NACC = 4;
  extern __shared__ int4 lds4[];
  __shared__ int s_job;

  const int lane = threadIdx.x & 63;
  const int grp  = lane / G;
  const int z    = lane % G;
  const int4 *qrow = (const int4 *) __builtin_assume_aligned(lds4, 16) + z;     // + (x*NQ + q)*G
  const int  *trow = (const int *) lds4 + TAIL0 + z;                           // + (x*NT + t)*G
  const int64_t nitems = n - nseg_slots + nsegs;      // row segments of the long targets, then whole targets
  const int64_t ngroups = (nitems + TPW - 1) / TPW;

  for (int j = 0;; j++) {
    // the first profile from j on whose queue still holds work
    if (threadIdx.x == 0) {
      int jj = j;
      while (jj < njobs && (int64_t) __hip_atomic_load(&queues[jj], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= ngroups) jj++;
      s_job = jj;
    }
    __syncthreads();
    j = uni(s_job);
    if (j >= njobs) break;
    stage_table((int *) lds4, jobs.tbl[j], (size_t) tbl_dwords);
    uint16_t *zm = zmax + (size_t) j * zstride;
    int32_t *queue = queues + j;
    int64_t w = 0, wend = 0;                          // groups of TPW targets, `grab` groups per atomic
    for (;;) {
      if (w == wend) { w = (int64_t) queue_next(queue, grab); wend = w + grab; }
      if (w >= ngroups) break;
      const int64_t t = w * TPW + grp;
      w++;
      const bool valid = t < nitems;
      const bool part = valid && t < nsegs;
      int64_t slot = 0;
      int rs = 0, L = 0;                              // rows rs .. rs+L-1 of the target
      int own0 = 0;                                   // the segment's own rows start at rs + own0
      if (part) {
        const int2 sg = segs[t];
        slot = sg.x;
        rs = max(0, sg.y - WARM) & ~15;
        own0 = sg.y - rs;
        L = min(len[slot], sg.y + P7G_SEG_ROWS) - rs;
      } else if (valid) {
        const int64_t u = t - nsegs + nseg_slots;
        slot = slots ? (int64_t) slots[u] : u;
        L = len[slot];
      }
      const uint8_t *sq = res + (valid ? off[slot] + rs : 0);
      int Lw = L;
#pragma unroll
      for (int o = G; o < 64; o <<= 1) Lw = max(Lw, __shfl_xor(Lw, o, 64));
      Lw = uni(Lw);

      int zr[ND];
#pragma unroll
      for (int d = 0; d < ND; d++) zr[d] = (int) NEG_PAIR;
      int acc[NACC];
#pragma unroll
      for (int a = 0; a < NACC; a++) acc[a] = (int) NEG_PAIR;
      int nb = (int) NEG_PAIR;          // DPP destination: lanes without a left neighbour keep this fill

      const uint16_t *bi = nullptr;
      uint16_t *bo = nullptr;
      // a chunk's boundary columns, from the item's first row; a segment writes
      // only its own rows (its warm-up rows are lower bounds): the previous
      // chunk's column is exact on every row, so a warm-up of the chunk's width
      // makes the segment's own rows exact
      if (BND) {
        if (bin) bi = bin + off[slot] + rs;
        if (bout) bo = bout + off[slot] + rs;
      }
      // boundary values of rows i0 .. i0+15 (bq) and of row i0-1 (pv); the floor before row 0
      uint4 bq0 = make_uint4(0x80008000u, 0x80008000u, 0x80008000u, 0x80008000u), bq1 = bq0;
      unsigned pv = 0x8000u;
#define BQ_WORD(e) (((e) < 8) ? (((e) & 6) == 0 ? bq0.x : ((e) & 6) == 2 ? bq0.y : ((e) & 6) == 4 ? bq0.z : bq0.w) \
                              : (((e) & 6) == 0 ? bq1.x : ((e) & 6) == 2 ? bq1.y : ((e) & 6) == 4 ? bq1.z : bq1.w))
#define BQ_HALF(e) ((BQ_WORD(e) >> (((e) & 1) * 16)) & 0xffffu)
      // row i0+u's predecessor in at lane 0 before the row's shift; the row's last position out from lane 63
#define BND_IN(u) do { if (BND && (u) < 16) {                                                   \
        const unsigned v_ = ((u) == 0) ? pv : BQ_HALF((u) - 1);                                  \
        nb = (z == 0) ? (int) (v_ << 16) : nb; } } while (0)
#define BND_OUT(r) do { if (BND && bo && z == 63 && (r) >= own0) bo[r] = (uint16_t) ((unsigned) zr[ND-1] >> 16); } while (0)
#define BND_LOAD(i) do { if (BND && bi) {                                                        \
        pv = bq1.w >> 16; bq0 = *(const uint4 *) (bi + (i)); bq1 = *(const uint4 *) (bi + (i) + 8); } } while (0)

      const uint4 padchunk = make_uint4(0x1f1f1f1fu, 0x1f1f1f1fu, 0x1f1f1f1fu, 0x1f1f1f1fu);
      // one DP row with the scores in (cq, ct), while the row of residue xn is fetched into (nq, nt)
#define SSV_ROW(cq, ct, nq, nt, xn, u, r) do {                                                \
        const int4 *rq_ = qrow + (xn) * (NQ * G);                                               \
        const int  *rt_ = trow + (xn) * (NT * G);                                               \
        _Pragma("unroll") for (int q = 0; q < NQ; q++) nq[q] = rq_[q * G];                      \
        _Pragma("unroll") for (int u_ = 0; u_ < NT; u_++) nt[u_] = rt_[u_ * G];                 \
        BND_IN(u);                                                                              \
        nb = lane_in<G>(nb, zr[ND-1]);                                                          \
        const int s0_ = (int) __builtin_amdgcn_alignbit((unsigned) zr[ND-1], (unsigned) nb, 16); \
        _Pragma("unroll") for (int d = ND-1; d >= 0; d--) {                                     \
          int rv_;                                                                              \
          if (d < 4 * NQ) {                                                                     \
            const int4 c_ = cq[d / 4];                                                          \
            rv_ = (d % 4 == 0) ? c_.x : (d % 4 == 1) ? c_.y : (d % 4 == 2) ? c_.z : c_.w;       \
          } else rv_ = ct[d - 4 * NQ];                                                          \
          zr[d] = adds2(d ? zr[d-1] : s0_, rv_);                                                \
        }                                                                                       \
        _Pragma("unroll") for (int d = 0; d + 1 < ND; d += 2)                                   \
          acc[(d / 2) % NACC] = zmax3(acc[(d / 2) % NACC], zr[d], zr[d+1]);                     \
        if (ND & 1) acc[((ND - 1) / 2) % NACC] = zmax2(acc[((ND - 1) / 2) % NACC], zr[ND-1]);   \
        BND_OUT(r);                                                                             \
      } while (0)
#define RES(w, u) ((int) (((w) >> (8 * ((u) & 3))) & 0xffu))

      uint4 ch = (0 < L) ? *(const uint4 *) sq : padchunk;
      int4 qA[NQa], qB[NQa];
      int  tA[NTa], tB[NTa];
      {
        const int x0 = RES(ch.x, 0);
#pragma unroll
        for (int q = 0; q < NQ; q++) qA[q] = qrow[(x0 * NQ + q) * G];
#pragma unroll
        for (int u = 0; u < NT; u++) tA[u] = trow[(x0 * NT + u) * G];
      }
      int i0 = 0;
      for (; i0 + 16 <= Lw; i0 += 16) {
        const uint4 nch = (i0 + 16 < L) ? *(const uint4 *) (sq + i0 + 16) : padchunk;
        BND_LOAD(i0);
        SSV_ROW(qA, tA, qB, tB, RES(ch.x, 1), 0, i0);       SSV_ROW(qB, tB, qA, tA, RES(ch.x, 2), 1, i0 + 1);
        SSV_ROW(qA, tA, qB, tB, RES(ch.x, 3), 2, i0 + 2);   SSV_ROW(qB, tB, qA, tA, RES(ch.y, 0), 3, i0 + 3);
        SSV_ROW(qA, tA, qB, tB, RES(ch.y, 1), 4, i0 + 4);   SSV_ROW(qB, tB, qA, tA, RES(ch.y, 2), 5, i0 + 5);
        SSV_ROW(qA, tA, qB, tB, RES(ch.y, 3), 6, i0 + 6);   SSV_ROW(qB, tB, qA, tA, RES(ch.z, 0), 7, i0 + 7);
        SSV_ROW(qA, tA, qB, tB, RES(ch.z, 1), 8, i0 + 8);   SSV_ROW(qB, tB, qA, tA, RES(ch.z, 2), 9, i0 + 9);
        SSV_ROW(qA, tA, qB, tB, RES(ch.z, 3), 10, i0 + 10); SSV_ROW(qB, tB, qA, tA, RES(ch.w, 0), 11, i0 + 11);
        SSV_ROW(qA, tA, qB, tB, RES(ch.w, 1), 12, i0 + 12); SSV_ROW(qB, tB, qA, tA, RES(ch.w, 2), 13, i0 + 13);
        SSV_ROW(qA, tA, qB, tB, RES(ch.w, 3), 14, i0 + 14); SSV_ROW(qB, tB, qA, tA, RES(nch.x, 0), 15, i0 + 15);
        ch = nch;
      }
      if (i0 < Lw) BND_LOAD(i0);
      for (int u = 0; i0 + u < Lw; u++) {              // the last rows, fewer than 16
        const unsigned wd = (u + 1 < 4) ? ch.x : (u + 1 < 8) ? ch.y : (u + 1 < 12) ? ch.z : ch.w;
        if (BND && bi) {                               // row i0+u's predecessor, by a uniform select
          const int e = u - 1;
          const unsigned bw = (e < 2) ? bq0.x : (e < 4) ? bq0.y : (e < 6) ? bq0.z : (e < 8) ? bq0.w
                            : (e < 10) ? bq1.x : (e < 12) ? bq1.y : (e < 14) ? bq1.z : bq1.w;
          const unsigned v = (u == 0) ? pv : ((bw >> ((e & 1) * 16)) & 0xffffu);
          nb = (z == 0) ? (int) (v << 16) : nb;
        }
        SSV_ROW(qA, tA, qB, tB, RES(wd, u + 1), 16, i0 + u);
#pragma unroll
        for (int q = 0; q < NQ; q++) qA[q] = qB[q];
#pragma unroll
        for (int v = 0; v < NT; v++) tA[v] = tB[v];
      }
#undef SSV_ROW
#undef RES
#undef BND_IN
#undef BND_OUT
#undef BND_LOAD
#undef BQ_HALF
#undef BQ_WORD
      int m = -32768;
#pragma unroll
      for (int a = 0; a < NACC; a++) m = max(m, max(lo16(acc[a]), hi16(acc[a])));
#pragma unroll
      for (int o = G/2; o > 0; o >>= 1) m = max(m, __shfl_xor(m, o, 64));
      if (valid && z == 0) {
        const uint16_t v = (uint16_t) (m + 32768);
        if (part) {
          // segments of one target fold their maxima into its (zeroed) entry:
          // a compare-and-swap on the 32-bit word that holds it
          unsigned *wp = (unsigned *) ((uintptr_t) &zm[slot] & ~(uintptr_t) 3);
          const int sh = ((uintptr_t) &zm[slot] & 2) ? 16 : 0;
          unsigned old = __hip_atomic_load(wp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT), seen;
          do {
            seen = old;
            if (((seen >> sh) & 0xffffu) >= v) break;
            old = atomicCAS(wp, seen, (seen & ~(0xffffu << sh)) | ((unsigned) v << sh));
          } while (old != seen);
        } else zm[slot] = (BND && bin) ? (uint16_t) max((int) zm[slot], (int) v) : v;
      }
    }
    __syncthreads();                                  // every wave is done with this table
  }
}

/*****************************************************************
 * Launchers
 *****************************************************************/
#define SSV_ND8(X) X(8,1) X(8,2) X(8,3) X(8,4) X(8,5) X(8,6) X(8,7) X(8,8) X(8,9) X(8,10) X(8,11) X(8,12) X(8,13) \
  X(8,14) X(8,15) X(8,16) X(8,17) X(8,18) X(8,19) X(8,20) X(8,21) X(8,22) X(8,23) X(8,24) X(8,25) X(8,26) \
  X(8,27) X(8,28) X(8,29) X(8,30) X(8,31) X(8,32)
#define SSV_ND16(X) X(16,1) X(16,2) X(16,3) X(16,4) X(16,5) X(16,6) X(16,7) X(16,8) X(16,9) X(16,10) X(16,11) \
  X(16,12) X(16,13) X(16,14) X(16,15) X(16,16) X(16,17) X(16,18) X(16,19) X(16,20) X(16,21) X(16,22) X(16,23) \
  X(16,24) X(16,25) X(16,26) X(16,27) X(16,28) X(16,29) X(16,30) X(16,31) X(16,32)
#define SSV_ND32(X) X(32,3) X(32,4) X(32,5) X(32,6) X(32,7) X(32,8) X(32,9) X(32,10) X(32,11) X(32,12) X(32,13) \
  X(32,14) X(32,15) X(32,16) X(32,18) X(32,20) X(32,22) X(32,24) X(32,26) X(32,28) X(32,30) X(32,32)
#define SSV_ND64(X) X(64,8) X(64,9) X(64,10) X(64,11) X(64,12) X(64,13) X(64,14) X(64,15) X(64,16) X(64,18) \
  X(64,20) X(64,22) X(64,24) X(64,26) X(64,28) X(64,30) X(64,32) X(64,36) X(64,40)
// chunk shapes of models longer than one LDS table holds (G = 64)
#define SSV_CHUNK(X) X(64,11) X(64,12) X(64,13) X(64,14) X(64,15) X(64,16) X(64,18)
#define SSV_CASE(G_, N_) case (G_) * 256 + (N_): return (void *) k_ssv<G_, N_, false>;
#define SSV_BCASE(G_, N_) case (G_) * 256 + (N_): return (void *) k_ssv<G_, N_, true>;

static void *ssv_kernel(int G, int ND, bool bnd)
{
  if (bnd) {
    switch (G * 256 + ND) { SSV_CHUNK(SSV_BCASE) }
    return nullptr;
  }
  switch (G * 256 + ND) {
    SSV_ND8(SSV_CASE) SSV_ND16(SSV_CASE) SSV_ND32(SSV_CASE) SSV_ND64(SSV_CASE)
  }
  return nullptr;
}

int ssv_shape_built(int G, int ND) { return ssv_kernel(G, ND, false) != nullptr; }
int ssv_chunk_built(int ND) { return ssv_kernel(64, ND, true) != nullptr; }

// zmax[j * stride + i] = 0 for i < count, j < njobs: the entries the row
// segments of the long targets fold their maxima into
__global__ void k_zero_u16(uint16_t *__restrict__ z, int64_t stride, int64_t count, int njobs)
{
  const int64_t total = count * njobs;
  for (int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t) gridDim.x * blockDim.x)
    z[(i / count) * stride + i % count] = 0;
}

static hipError_t ssv_run(int G, int ND, const SsvJobs &jobs, int njobs, const DevSeqs &s, const int32_t *slots,
                          int64_t n, uint16_t *zmax, int64_t zstride, int32_t *queues, int grab,
                          const uint16_t *bin, uint16_t *bout, bool bnd, hipStream_t st)
{
  void *k = ssv_kernel(G, ND, bnd);
  if (! k || njobs < 1 || njobs > P7G_SSV_MAXJOBS || grab < 1 || (bnd && njobs != 1)) return hipErrorInvalidValue;
  // long targets in row segments: over every slot, when the warm-up (the
  // shape's positions: a chunk's, for a chunked model) is at most half a
  // segment. Their entries start from zero, except for a later chunk, whose
  // segments fold into the maxima of the chunks before it.
  const int2 *segs = nullptr;
  int64_t nsegs = 0, nseg_slots = 0;
  if (! slots && s.nseg > 0 && n == s.nslot && 2 * ND * G <= P7G_SEG_ROWS / 2) {
    segs = s.seg; nsegs = s.nseg; nseg_slots = s.nseg_slots;
    if (! bin) {
      const int64_t total = nseg_slots * njobs;
      const unsigned zg = (unsigned) std::min<int64_t>((total + 255) / 256, 4096);
      hipLaunchKernelGGL(k_zero_u16, dim3(zg), dim3(256), 0, st, zmax, zstride, nseg_slots, njobs);
      hipError_t e = hipGetLastError();
      if (e != hipSuccess) return e;
    }
  }
  const size_t lds = (size_t) P7G_NROWS * G * ND * 4;
  const int64_t nitems = n - nseg_slots + nsegs;
  const int64_t nwaves = (int64_t) njobs * ((nitems + (64 / G) - 1) / (64 / G));
  const unsigned grid = grid_for_waves_blk(k, nwaves, lds, SSV_BLOCK / 64);
  int tbl_dwords = (int) (lds / 4);
  void *args[] = { (void *) &jobs, &njobs, &tbl_dwords, (void *) &s.res, (void *) &s.off, (void *) &s.len,
                   (void *) &slots, &n, &zmax, &zstride, &queues, &grab, (void *) &bin, (void *) &bout,
                   (void *) &segs, &nsegs, &nseg_slots };
  return launch(k, grid, SSV_BLOCK, args, lds, st);
}

hipError_t launch_ssv(const DevProfile &p, int v, const DevSeqs &s, const int32_t *slots, int64_t n, uint16_t *zmax,
                      int32_t *queue, int grab, hipStream_t st)
{
  if (n <= 0) return hipSuccess;
  if (v < 0 || v >= P7G_SSV_NVAR || p.ssv[v].G == 0) return hipErrorInvalidValue;
  const SsvVariant &sv = p.ssv[v];
  if (sv.dwords != (size_t) P7G_NROWS * sv.G * sv.ND) return hipErrorInvalidValue;
  SsvJobs jobs;
  memset(&jobs, 0, sizeof(jobs));
  jobs.tbl[0] = sv.tbl;
  return ssv_run(sv.G, sv.ND, jobs, 1, s, slots, n, zmax, 0, queue, grab, nullptr, nullptr, false, st);
}

hipError_t launch_ssv_chunked(const DevProfile &p, const DevSeqs &s, const int32_t *slots, int64_t n, uint16_t *zmax,
                              int32_t *queues, int grab, uint16_t *bnd, int64_t bnd_stride, hipStream_t st)
{
  if (n <= 0) return hipSuccess;
  const int nc = p.ssv_nchunk;
  if (nc < 2 || nc > P7G_SSV_MAXCHUNK || ! bnd) return hipErrorInvalidValue;
  for (int c = 0; c < nc; c++) {
    const SsvVariant &sv = p.ssvc[c];
    if (sv.G != 64 || sv.dwords != (size_t) P7G_NROWS * 64 * sv.ND) return hipErrorInvalidValue;
    SsvJobs jobs;
    memset(&jobs, 0, sizeof(jobs));
    jobs.tbl[0] = sv.tbl;
    // boundary columns alternate between the two halves of bnd
    const uint16_t *bin = c ? bnd + (size_t) ((c - 1) & 1) * bnd_stride : nullptr;
    uint16_t *bout = (c < nc - 1) ? bnd + (size_t) (c & 1) * bnd_stride : nullptr;
    hipError_t e = ssv_run(64, sv.ND, jobs, 1, s, slots, n, zmax, 0, queues + c, grab, bin, bout, true, st);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t launch_ssv_multi(int G, int ND, const int *const *tbl, int njobs, const DevSeqs &s, int64_t n,
                            uint16_t *zmax, int64_t zstride, int32_t *queues, int grab, hipStream_t st)
{
  if (n <= 0 || njobs <= 0) return hipSuccess;
  for (int j0 = 0; j0 < njobs; j0 += P7G_SSV_MAXJOBS) {
    const int nj = std::min(njobs - j0, P7G_SSV_MAXJOBS);
    SsvJobs jobs;
    memset(&jobs, 0, sizeof(jobs));
    for (int j = 0; j < nj; j++) jobs.tbl[j] = tbl[j0 + j];
    hipError_t e = ssv_run(G, ND, jobs, nj, s, nullptr, n, zmax + (size_t) j0 * zstride, zstride, queues + j0, grab,
                           nullptr, nullptr, false, st);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}
