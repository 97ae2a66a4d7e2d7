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
// A target is spread over G = 16, 32 or 64 lanes (4, 2 or 1 targets per wave),
// lane z holding model positions z*2ND+1 .. z*2ND+2ND as ND int16 pairs (dword
// d: positions d and d + ND). Which G serves a model best is measured by the
// runtime's tuner; every shape gives the same result.
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
// rest of each 16-lane row, leaving lanes 0 and 32 with the fill.
template <int G>
__device__ __forceinline__ int lane_in(int fill, int v)
{
  if (G == 16) return __builtin_amdgcn_update_dpp(fill, v, ROW_SHR1, 0xf, 0xf, false);
  if (G == 64) return __builtin_amdgcn_update_dpp(fill, v, WAVE_SHR1, 0xf, 0xf, false);
  const int b = __builtin_amdgcn_update_dpp(fill, v, ROW_BCAST15, 0xa, 0xf, false);
  return __builtin_amdgcn_update_dpp(b, v, ROW_SHR1, 0xf, 0xf, false);
}

template <int G, int ND>
__global__ __launch_bounds__(SSV_BLOCK)
void k_ssv(const SsvJobs jobs, int njobs, int tbl_dwords,
           const uint8_t *__restrict__ res, const int64_t *__restrict__ off, const int32_t *__restrict__ len,
           const int32_t *__restrict__ slots, int64_t n, uint16_t *__restrict__ zmax, int64_t zstride,
           int32_t *queues, int grab)
{
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
  const int64_t ngroups = (n + TPW - 1) / TPW;

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
      const bool valid = t < n;
      const int64_t slot = valid ? (slots ? (int64_t) slots[t] : t) : 0;
      const int L = valid ? len[slot] : 0;
      const uint8_t *sq = res + (valid ? off[slot] : 0);
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

      const uint4 padchunk = make_uint4(0x1f1f1f1fu, 0x1f1f1f1fu, 0x1f1f1f1fu, 0x1f1f1f1fu);
      // one DP row with the scores in (cq, ct), while the row of residue xn is fetched into (nq, nt)
#define SSV_ROW(cq, ct, nq, nt, xn) do {                                                      \
        const int4 *rq_ = qrow + (xn) * (NQ * G);                                               \
        const int  *rt_ = trow + (xn) * (NT * G);                                               \
        _Pragma("unroll") for (int q = 0; q < NQ; q++) nq[q] = rq_[q * G];                      \
        _Pragma("unroll") for (int u = 0; u < NT; u++) nt[u] = rt_[u * G];                      \
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
        SSV_ROW(qA, tA, qB, tB, RES(ch.x, 1));  SSV_ROW(qB, tB, qA, tA, RES(ch.x, 2));
        SSV_ROW(qA, tA, qB, tB, RES(ch.x, 3));  SSV_ROW(qB, tB, qA, tA, RES(ch.y, 0));
        SSV_ROW(qA, tA, qB, tB, RES(ch.y, 1));  SSV_ROW(qB, tB, qA, tA, RES(ch.y, 2));
        SSV_ROW(qA, tA, qB, tB, RES(ch.y, 3));  SSV_ROW(qB, tB, qA, tA, RES(ch.z, 0));
        SSV_ROW(qA, tA, qB, tB, RES(ch.z, 1));  SSV_ROW(qB, tB, qA, tA, RES(ch.z, 2));
        SSV_ROW(qA, tA, qB, tB, RES(ch.z, 3));  SSV_ROW(qB, tB, qA, tA, RES(ch.w, 0));
        SSV_ROW(qA, tA, qB, tB, RES(ch.w, 1));  SSV_ROW(qB, tB, qA, tA, RES(ch.w, 2));
        SSV_ROW(qA, tA, qB, tB, RES(ch.w, 3));  SSV_ROW(qB, tB, qA, tA, RES(nch.x, 0));
        ch = nch;
      }
      for (int u = 0; i0 + u < Lw; u++) {              // the last rows, fewer than 16
        const unsigned wd = (u + 1 < 4) ? ch.x : (u + 1 < 8) ? ch.y : (u + 1 < 12) ? ch.z : ch.w;
        SSV_ROW(qA, tA, qB, tB, RES(wd, u + 1));
#pragma unroll
        for (int q = 0; q < NQ; q++) qA[q] = qB[q];
#pragma unroll
        for (int v = 0; v < NT; v++) tA[v] = tB[v];
      }
#undef SSV_ROW
#undef RES
      int m = -32768;
#pragma unroll
      for (int a = 0; a < NACC; a++) m = max(m, max(lo16(acc[a]), hi16(acc[a])));
#pragma unroll
      for (int o = G/2; o > 0; o >>= 1) m = max(m, __shfl_xor(m, o, 64));
      if (valid && z == 0) zm[slot] = (uint16_t) (m + 32768);
    }
    __syncthreads();                                  // every wave is done with this table
  }
}

/*****************************************************************
 * Launchers
 *****************************************************************/
#define SSV_ND16(X) X(16,1) X(16,2) X(16,3) X(16,4) X(16,5) X(16,6) X(16,7) X(16,8) X(16,9) X(16,10) X(16,11) \
  X(16,12) X(16,13) X(16,14) X(16,15) X(16,16) X(16,17) X(16,18) X(16,19) X(16,20) X(16,21) X(16,22) X(16,23) \
  X(16,24) X(16,25) X(16,26) X(16,27) X(16,28) X(16,29) X(16,30) X(16,31) X(16,32)
#define SSV_ND32(X) X(32,3) X(32,4) X(32,5) X(32,6) X(32,7) X(32,8) X(32,9) X(32,10) X(32,11) X(32,12) X(32,13) \
  X(32,14) X(32,15) X(32,16) X(32,18) X(32,20) X(32,22) X(32,24) X(32,26) X(32,28) X(32,30) X(32,32)
#define SSV_ND64(X) X(64,8) X(64,9) X(64,10) X(64,11) X(64,12) X(64,13) X(64,14) X(64,15) X(64,16) X(64,18) \
  X(64,20) X(64,22) X(64,24) X(64,26) X(64,28) X(64,30) X(64,32) X(64,36) X(64,40)
#define SSV_CASE(G_, N_) case (G_) * 256 + (N_): return (void *) k_ssv<G_, N_>;

static void *ssv_kernel(int G, int ND)
{
  switch (G * 256 + ND) {
    SSV_ND16(SSV_CASE) SSV_ND32(SSV_CASE) SSV_ND64(SSV_CASE)
  }
  return nullptr;
}

int ssv_shape_built(int G, int ND) { return ssv_kernel(G, ND) != nullptr; }

static hipError_t ssv_run(int G, int ND, const SsvJobs &jobs, int njobs, const DevSeqs &s, const int32_t *slots,
                          int64_t n, uint16_t *zmax, int64_t zstride, int32_t *queues, int grab, hipStream_t st)
{
  void *k = ssv_kernel(G, ND);
  if (! k || njobs < 1 || njobs > P7G_SSV_MAXJOBS || grab < 1) return hipErrorInvalidValue;
  const size_t lds = (size_t) P7G_NROWS * G * ND * 4;
  const int64_t nwaves = (int64_t) njobs * ((n + (64 / G) - 1) / (64 / G));
  const unsigned grid = grid_for_waves_blk(k, nwaves, lds, SSV_BLOCK / 64);
  int tbl_dwords = (int) (lds / 4);
  void *args[] = { (void *) &jobs, &njobs, &tbl_dwords, (void *) &s.res, (void *) &s.off, (void *) &s.len,
                   (void *) &slots, &n, &zmax, &zstride, &queues, &grab };
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
  return ssv_run(sv.G, sv.ND, jobs, 1, s, slots, n, zmax, 0, queue, grab, st);
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
    hipError_t e = ssv_run(G, ND, jobs, nj, s, nullptr, n, zmax + (size_t) j0 * zstride, zstride, queues + j0, grab, st);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}
