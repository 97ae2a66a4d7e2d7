// gfx950 kernels of the HMMER3 filter pipeline.
//
// Layout shared by all DP kernels: a target's model positions k = 1..M are
// split into contiguous blocks, one block per lane of a G-lane group
// (G = 16 or 64). Lane z owns positions z*B + 1 .. z*B + B, so the row-to-row
// dependency M(i,k) <- M(i-1,k-1) stays inside a lane except at block starts,
// where one DPP lane shift (row_shr / wave_shr) brings the neighbour's last
// value. The integer stages hold two positions per dword as int16 halves:
// dword d carries positions j = d (low half) and j = d + ND (high half), so a
// whole row shifts with one DPP move plus one v_alignbit on the last dword.
//
// MSV stage (HMMER/src/impl_sse/msvfilter.c:76-210): the uchar recurrence
//   sv = subs_u8(adds_u8(max(M(i-1,k-1), xB), bias), rbv[x][k])
// is evaluated exactly, in two passes:
//   1. k_ssv: while the J state is unused xB is a per-target constant, and with
//      Z = max(M, xB) - xB the row update becomes Z(i,k) = max(Z(i-1,k-1) + r, 0),
//      r = bias - rbv. Holding Z - 32768 in int16, the floor at 0 is exactly the
//      saturation of v_pk_add_i16 ... clamp, so a cell costs one packed add and
//      one packed max (the running maximum). The pass records max Z per target.
//   2. k_msv_classify decides per target: overflow (some row reaches
//      xE + bias >= 255), exact score (J never above base, xJ = max xE - tec), or
//      "needs the J state" -> k_msv_full runs the row-serial recurrence.
// Viterbi filter (vitfilter.c:84-249): int16 pairs with v_pk_add_i16 clamp =
// _mm_adds_epi16 and v_pk_max_i16 = _mm_max_epi16; the lazy-F condition
// (Dmax + ddbound > xB) is evaluated per row exactly as the reference does, and
// when it holds the D->D closure is a max-plus prefix scan across the wave.
// Forward/Backward parsers (fwdback.c:258-735): fp32, the D->D chain as an
// affine prefix scan across the wave, sparse rescaling when xE > 1e4.
#include "p7g_devfn.h"

/*****************************************************************
 * 2. MSV stage: per-target decision after the ungapped pass
 *****************************************************************/
__global__ void k_msv_classify(const DevProfile p, const int32_t *__restrict__ len, const int32_t *__restrict__ slots,
                               int64_t n, const uint16_t *__restrict__ zmax, int16_t *__restrict__ xj,
                               int32_t *__restrict__ full_list, int32_t *__restrict__ full_count, LenCfg lc)
{
  for (int64_t t = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t) gridDim.x * blockDim.x) {
    const int64_t slot = slots ? (int64_t) slots[t] : t;
    const int L    = len[slot];
    const int tjbm = ((lc.on ? lc.tjb : (int) p.tjb[L]) + p.tbm) & 0xff;
    const int xB   = max(p.base - tjbm, 0);
    const int zm   = zmax[slot];
    int v;
    if (zm == 0)                          v = -2;          // no cell above the B floor: rescore exactly
    else if (xB + zm + p.bias >= 255)     v = -1;          // some row overflows (MSV values only grow with xB)
    else if (xB + zm > p.base + p.tec)    v = -2;          // the J state would raise xB
    else                                  v = max(xB + zm - p.tec, 0);
    if (slots) {                                   // explicit slot list: append here
      list_append(v == -2, (int32_t) slot, full_list, full_count);
      if (v == -2) v = 0;
    }
    xj[slot] = (int16_t) v;                        // -2 marks "full recurrence" for the compaction
  }
}

/*****************************************************************
 * 3. MSV stage, full recurrence (one target per wave, 64 lanes)
 *****************************************************************/
// LDS: the score table is staged in LDS; otherwise (models whose table does
// not fit, ND >= 24) the rows are read from global memory.
template <int ND, bool LDS>
__global__ __launch_bounds__(512)
void k_msv_full(const DevProfile p, const int *__restrict__ tbl, int tbl_dwords,
                const uint8_t *__restrict__ res, const int64_t *__restrict__ off, const int32_t *__restrict__ len,
                const int32_t *__restrict__ list, const int32_t *__restrict__ count, int16_t *__restrict__ xj, LenCfg lc)
{
  if ((int64_t) blockIdx.x * (blockDim.x >> 6) >= *count) return;
  extern __shared__ int lds[];
  const int *rows = tbl;
  if (LDS) { stage_table(lds, tbl, (size_t) tbl_dwords); rows = lds; }
  const int z = threadIdx.x & 63;
  const int64_t n = *count;
  int32_t *queue = (int32_t *) count + P7G_QOFF;

  for (int64_t t = queue_next(queue, 1); t < n; t = queue_next(queue, 1)) {
    const int slot = uni(list[t]);
    const int L = uni(len[slot]);
    const uint8_t *sq = res + off[slot];
    const int tjbm = ((lc.on ? lc.tjb : (int) p.tjb[L]) + p.tbm) & 0xff;
    int xJ = 0;
    int xB = max(p.base - tjbm, 0);
    int mp[ND];
#pragma unroll
    for (int d = 0; d < ND; d++) mp[d] = 0;
    bool overflow = false;
    ResAsc rd;
    if (L > 0) rd.init(sq);
    for (int i = 0; i < L; i++) {
      const int x = rd.at(i);
      const int *row = rows + x * (ND * 64) + z;
      const int s0 = shift_in<64>(mp[ND-1], 0);
      const int xB2 = pair16(xB);
      int E = 0;
#pragma unroll
      for (int d = ND-1; d >= 0; d--) {
        int src = (d > 0) ? mp[d-1] : s0;
        int tv = max2(src, xB2);
        tv = adds2(tv, row[d * 64]);
        tv = max2(tv, 0);
        E  = max2(E, tv);
        mp[d] = tv;
      }
      // xJ (and so xB) only changes, and the overflow test only fires, when
      // the row's xE exceeds min(xJ + tec, 254 - bias): a ballot on the lanes'
      // maxima decides whether the wave maximum is needed at all
      const int el = max(lo16(E), hi16(E));
      if (__ballot(el > min(xJ + p.tec, 254 - p.bias)) != 0ull) {
        const int xE = uni(wave_max_i(el));
        if (xE + p.bias >= 255) { overflow = true; break; }
        xJ = max(xJ, max(xE - p.tec, 0));
        xB = max(max(p.base, xJ) - tjbm, 0);
      }
    }
    if (z == 0) xj[slot] = (int16_t) (overflow ? -1 : xJ);
  }
}

// The full MSV recurrence for large models, W waves per target: wave w owns
// dwords [w·ND/W, (w+1)·ND/W) of every lane. A slice's first dword takes the
// previous row's last dword of the slice before it (slice 0: of the last
// slice, joined across lanes as shift_in does) from LDS, double-buffered by
// row parity, so one barrier per row orders the exchange. The waves' row
// maxima travel in the same buffers, and xJ and xB are updated from them at
// the top of the next row (or after the last), in the same order as the
// one-wave kernel; results are bit-identical.
template <int ND, int W>
__global__ __launch_bounds__(64 * W)
void k_msv_full_wg(const DevProfile p, const int *__restrict__ tbl, const uint8_t *__restrict__ res,
                   const int64_t *__restrict__ off, const int32_t *__restrict__ len, const int32_t *__restrict__ list,
                   const int32_t *__restrict__ count, int16_t *__restrict__ xj, LenCfg lc)
{
  static_assert(ND % W == 0, "slices of equal width");
  constexpr int NW = ND / W;
  __shared__ int sh_end[2][W][64];       // each slice's last dword, by row parity
  __shared__ int sh_el[2][W];            // each wave's row maximum, by row parity
  __shared__ int64_t sh_t;
  if ((int64_t) blockIdx.x >= *count) return;
  const int wv = threadIdx.x >> 6, z = threadIdx.x & 63;
  const int d0 = wv * NW;
  const int64_t n = *count;
  int32_t *queue = (int32_t *) count + P7G_QOFF;

  for (;;) {
    if (threadIdx.x == 0) sh_t = (int64_t) atomicAdd(queue, 1);
    __syncthreads();
    const int64_t t = sh_t;
    if (t >= n) break;
    const int slot = list[t];
    const int L = len[slot];
    const uint8_t *sq = res + off[slot];
    const int tjbm = ((lc.on ? lc.tjb : (int) p.tjb[L]) + p.tbm) & 0xff;
    int xJ = 0;
    int xB = max(p.base - tjbm, 0);
    int mp[NW];
#pragma unroll
    for (int j = 0; j < NW; j++) mp[j] = 0;
    sh_end[1][wv][z] = 0;                              // row -1
    bool overflow = false;
    // the update of xJ and xB from the row maximum the waves left in buffer b
    auto update = [&](int b) {
      int xE = sh_el[b][0];
#pragma unroll
      for (int v = 1; v < W; v++) xE = max(xE, sh_el[b][v]);
      if (xE > min(xJ + p.tec, 254 - p.bias)) {
        if (xE + p.bias >= 255) { overflow = true; return; }
        xJ = max(xJ, max(xE - p.tec, 0));
        xB = max(max(p.base, xJ) - tjbm, 0);
      }
    };
    ResAsc rd;
    if (L > 0) rd.init(sq);
    for (int i = 0; i < L; i++) {
      const int x = rd.at(i);
      const int *row = tbl + (size_t) x * (ND * 64) + z;
      const int pb = (i + 1) & 1;                      // the previous row's buffer
      __syncthreads();                                 // the previous row's slice ends and maxima are written
      if (i > 0) { update(pb); if (overflow) break; }
      int s0;
      if (wv > 0) s0 = sh_end[pb][wv-1][z];
      else {
        const int nb = (z > 0) ? sh_end[pb][W-1][z-1] : 0;
        s0 = (int) __builtin_amdgcn_alignbit((unsigned) sh_end[pb][W-1][z], (unsigned) nb, 16);
      }
      const int xB2 = pair16(xB);
      int E = 0;
#pragma unroll
      for (int j = NW-1; j >= 0; j--) {
        int src = j ? mp[j-1] : s0;
        int tv = max2(src, xB2);
        tv = adds2(tv, row[(size_t) (d0 + j) * 64]);
        tv = max2(tv, 0);
        E  = max2(E, tv);
        mp[j] = tv;
      }
      sh_end[i & 1][wv][z] = mp[NW-1];
      const int el = wave_max_i(max(lo16(E), hi16(E)));
      if (z == 0) sh_el[i & 1][wv] = el;
    }
    if (L > 0 && ! overflow) {
      __syncthreads();
      update((L - 1) & 1);
    }
    if (threadIdx.x == 0) xj[slot] = (int16_t) (overflow ? -1 : xJ);
    __syncthreads();                                   // sh_t and the buffers are free for the next target
  }
}

// MSV scores of listed targets from their final xJ (msvfilter.c:203-207,
// with the per-target length model): ((xJ - tjb) - base) / scale_b - 3 nats
// in float, the last step in double as the reference's `-= 3.0`; +inf and
// status 16 (eslERANGE) where the recurrence overflowed.
__global__ void k_msv_scores(const DevProfile p, const int32_t *__restrict__ len, const int32_t *__restrict__ slots,
                             int64_t n, const int16_t *__restrict__ xj, float *__restrict__ sc, int *__restrict__ st)
{
#pragma clang fp contract(off)
  for (int64_t t = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t) gridDim.x * blockDim.x) {
    const int slot = slots[t];
    const int v = xj[slot];
    if (v < 0) { sc[t] = __int_as_float(0x7f800000); st[t] = 16; continue; }
    float f = (float) (v - (int) p.tjb[len[slot]]) - (float) p.base;
    f = __fdiv_rn(f, p.scale_b);
    sc[t] = (float) ((double) f - 3.0);
    st[t] = 0;
  }
}

/*****************************************************************
 * 4. MSV gate: score, null score, F1 (and F2 when the bias filter is off)
 *****************************************************************/
__global__ void k_msv_gate(const DevProfile p, const int32_t *__restrict__ len, int64_t n,
                           const int16_t *__restrict__ xj, DevResults r, DevLists l, GateThresh g)
{
#pragma clang fp contract(off)
  for (int64_t slot = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; slot < n; slot += (int64_t) gridDim.x * blockDim.x) {
    const int L = len[slot];
    const int v = xj[slot];
    const float nullsc = p.nullsc[L];
    float usc;
    if (v < 0) usc = INFINITY;
    else {
      usc = __fdiv_rn((float) (v - (int) p.tjb[L]) - (float) p.base, p.scale_b);
      usc = (float) ((double) usc - 3.0);
    }
    r.usc[slot] = usc;
    r.nullsc[slot] = nullsc;
    uint8_t stage = 0;
    bool to_bias = false, to_vit = false, to_fwd = false;
    if (L > 0) {
      const float seq = (float) __ddiv_rn((double) (usc - nullsc), LOG2_D);
      if (seq >= g.T1m) {
        stage = 1;
        if (g.do_bias) to_bias = true;
        else {
          r.filtersc[slot] = nullsc;
          if (seq >= g.T2m) { to_fwd = true; stage = 3; }
          else              { to_vit = true; stage = 2; }
        }
      }
    }
    (void) to_bias; (void) to_vit; (void) to_fwd;   // the next lists are compacted from stage[]
    r.stage[slot] = stage;
  }
}

/*****************************************************************
 * 5. Bias-composition filter: 2-state HMM Forward with per-row max
 *    normalisation (easel/esl_hmm.c:351-409, HMMER/src/p7_bg.c:473-484).
 *    Float arithmetic in the reference's order, in two passes:
 *    k_bias runs the recurrence, one target per thread, and stores each
 *    row's normaliser; k_bias_sum takes the logs of a target's normalisers
 *    on all lanes at once and adds them in row order, as the reference does.
 *    The log is off the recurrence's serial chain, which bounds the stage.
 *****************************************************************/
__device__ __forceinline__ void bias_finish(const DevProfile &p, DevResults &r, const GateThresh &g, int gating,
                                            const LenCfg &lc, int slot, int L, float logsc)
{
#pragma clang fp contract(off)
  // with a caller-supplied length model the host adds the length terms
  const float filtersc = lc.on ? logsc : (logsc + p.flen_a[L]) + p.flen_b[L];
  r.filtersc[slot] = filtersc;
  if (gating) {
    const float seq = (float) __ddiv_rn((double) (r.usc[slot] - filtersc), LOG2_D);
    if (seq < g.T1m)       r.stage[slot] = 1;
    else if (seq >= g.T2m) r.stage[slot] = 3;       // the lists of the next stages are compacted from stage[]
    else                   r.stage[slot] = 2;
  }
}

__global__ __launch_bounds__(64)
void k_bias(const DevProfile p, const uint8_t *__restrict__ res, const int64_t *__restrict__ off,
            const int32_t *__restrict__ len, const int32_t *__restrict__ list, const int32_t *__restrict__ count,
            DevResults r, GateThresh g, int gating, LenCfg lc,
            float *__restrict__ rowmax, int64_t rowmax_size, unsigned long long *cursor, int64_t *__restrict__ defer_off)
{
#pragma clang fp contract(off)
  // the two states' odds of every residue code, in LDS
  __shared__ float2 eo2[P7G_NROWS];
  for (int x = threadIdx.x; x < P7G_NROWS; x += blockDim.x) eo2[x] = make_float2(p.eo[2*x], p.eo[2*x + 1]);
  __syncthreads();
  const int64_t n = *count;
  const int lane = threadIdx.x & 63;
  for (int64_t t0 = (int64_t) blockIdx.x * blockDim.x; t0 < n; t0 += (int64_t) gridDim.x * blockDim.x) {
    const int64_t t = t0 + threadIdx.x;
    const bool valid = t < n;
    const int slot = valid ? list[t] : 0;
    const int L = valid ? len[slot] : 0;
    // room for L row normalisers and the final sum: one bump of the cursor per wave
    int64_t base = -1;
    if (rowmax) {
      // whole 16-row blocks and the closing sum, in 16-byte aligned pieces
      const int64_t need = valid ? 16 * (((int64_t) L + 15) >> 4) + 4 : 0;
      int64_t incl = need;
#pragma unroll
      for (int o = 1; o < 64; o <<= 1) { const int64_t v = __shfl_up(incl, o, 64); if (lane >= o) incl += v; }
      unsigned long long b0 = 0;
      if (lane == 63) b0 = atomicAdd(cursor, (unsigned long long) incl);
      b0 = __shfl(b0, 63, 64);
      const int64_t mine = (int64_t) b0 + incl - need;
      if (valid && mine + need <= rowmax_size) base = mine;
    }
    if (! valid) continue;
    const uint8_t *sq = res + off[slot];
    const float p1  = lc.on ? lc.p1 : __fdiv_rn((float) L, (float) (L + 1));
    const float t00 = p1, t01 = lc.on ? lc.q1 : 1.0f - p1;
    const float t10 = p.ft1_0, t11 = p.ft1_1;
    float *rm = (base >= 0) ? rowmax + base : nullptr;
    float logsc = 0.0f;
    // Row 0 runs the same arithmetic as the others with the identity for the
    // transitions and (pi0, pi1) for the previous row: ((0 + pi0*1) + pi1*0)
    // * e0 is exactly pi0 * e0, the reference's first row.
    float d0 = p.pi0, d1 = p.pi1;
    float T00 = 1.0f, T01 = 0.0f, T10 = 0.0f, T11 = 1.0f;
    float f0 = d0, f1 = d1;                      // (d0, d1) after the target's last row
    const uint4 *pc = (const uint4 *) sq;
    const int nblk = (L + 15) >> 4;
    if (rm) {
      // Straight-line blocks of 16 rows, so that the compiler can count the
      // memory operations in flight. At the top of block bk every access in
      // flight was issued a whole block earlier: the block's residue chunk,
      // and the stores of block bk-2's normalisers. The block then reads its
      // odds pairs from LDS, requests the next chunk, stores block bk-1's
      // normalisers (block 0 writes its own slot, rewritten a block later),
      // and runs its rows. Rows past the end compute values no one reads,
      // into the target's own whole-block room.
      uint4 cn = pc[0];
      float qp[16];
#pragma unroll
      for (int u = 0; u < 16; u++) qp[u] = 0.0f;
      for (int bk = 0; bk < nblk; bk++) {
        float2 eb[16];
#pragma unroll
        for (int u = 0; u < 16; u++) eb[u] = eo2[chunk_byte(cn, u)];
        cn = pc[min(bk + 1, nblk)];             // the chunk after the target is still readable
        {
          float4 *o = (float4 *) (rm + max(bk - 1, 0) * 16);
          o[0] = make_float4(qp[0], qp[1], qp[2], qp[3]);     o[1] = make_float4(qp[4], qp[5], qp[6], qp[7]);
          o[2] = make_float4(qp[8], qp[9], qp[10], qp[11]);   o[3] = make_float4(qp[12], qp[13], qp[14], qp[15]);
        }
#pragma unroll
        for (int u = 0; u < 16; u++) {
          const float2 e = eb[u];
          float n0 = 0.0f, n1 = 0.0f, mx;
          n0 += d0 * T00; n0 += d1 * T10; n0 *= e.x;
          mx = 0.0f; mx = (n0 > mx) ? n0 : mx;
          n1 += d0 * T01; n1 += d1 * T11; n1 *= e.y;
          mx = (n1 > mx) ? n1 : mx;
          d0 = __fdiv_rn(n0, mx); d1 = __fdiv_rn(n1, mx);
          qp[u] = mx;
          if (u == 0) { T00 = t00; T01 = t01; T10 = t10; T11 = t11; }
          const bool last = (bk * 16 + u == L - 1);
          f0 = last ? d0 : f0; f1 = last ? d1 : f1;
        }
      }
      if (nblk > 0) {
        float4 *o = (float4 *) (rm + (nblk - 1) * 16);
        o[0] = make_float4(qp[0], qp[1], qp[2], qp[3]);     o[1] = make_float4(qp[4], qp[5], qp[6], qp[7]);
        o[2] = make_float4(qp[8], qp[9], qp[10], qp[11]);   o[3] = make_float4(qp[12], qp[13], qp[14], qp[15]);
      }
    } else {
      // no room for the normalisers (pool exhausted): their logs are summed here
      for (int i = 0; i < L; i++) {
        const float2 e = eo2[chunk_byte(pc[i >> 4], i)];
        float n0 = 0.0f, n1 = 0.0f, mx;
        n0 += d0 * T00; n0 += d1 * T10; n0 *= e.x;
        mx = 0.0f; mx = (n0 > mx) ? n0 : mx;
        n1 += d0 * T01; n1 += d1 * T11; n1 *= e.y;
        mx = (n1 > mx) ? n1 : mx;
        d0 = __fdiv_rn(n0, mx); d1 = __fdiv_rn(n1, mx);
        logsc += (float) log((double) mx);
        T00 = t00; T01 = t01; T10 = t10; T11 = t11;
      }
      f0 = d0; f1 = d1;
    }
    d0 = f0; d1 = f1;
    // the closing sum; an empty target scores log(pi_end) = log(0) (esl_hmm.c:360-364)
    float sfin = 0.0f;
    if (L > 0) { sfin += d0 * 1.0f; sfin += d1 * 1.0f; }
    if (rm) {
      rm[L] = sfin;                            // after the block stores of the same lane
      defer_off[slot] = base;                  // k_bias_log + k_bias_sum finish it
      continue;
    }
    if (rowmax) defer_off[slot] = -1;
    logsc += (float) log((double) sfin);
    bias_finish(p, r, g, gating, lc, slot, L, logsc);
  }
}

// Every stored value replaced by float(log(double(v))), in parallel.
__global__ __launch_bounds__(256)
void k_bias_log(float *__restrict__ rowmax, const unsigned long long *__restrict__ used, int64_t cap)
{
  const int64_t n = min((int64_t) *used, cap);
  for (int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t) gridDim.x * blockDim.x)
    rowmax[i] = (float) log((double) rowmax[i]);
}

// One target per wave: its L+1 logs added in row order, as esl_hmm_Forward
// sums sc[1..L+1] (esl_hmm.c:403-405). The wave loads 64 terms at a time
// (coalesced) and the sum takes them lane by lane, so the order of the float
// additions is the reference's; a long target no longer waits on its own
// dependent loads, one per four terms.
__global__ __launch_bounds__(256)
void k_bias_sum(const DevProfile p, const int32_t *__restrict__ len, const int32_t *__restrict__ list,
                const int32_t *__restrict__ count, DevResults r, GateThresh g, int gating, LenCfg lc,
                const float *__restrict__ rowmax, const int64_t *__restrict__ defer_off)
{
#pragma clang fp contract(off)
  const int lane = threadIdx.x & 63;
  const int64_t n = *count;
  for (int64_t t = wave_id(); t < n; t += wave_cnt()) {
    const int slot = uni(list[t]);
    const int64_t o0 = defer_off[slot];
    const int64_t o = ((int64_t) (unsigned) uni((int) (o0 >> 32)) << 32) | (unsigned) uni((int) o0);
    if (o < 0) continue;                       // finished in the first pass
    const int L = uni(len[slot]);
    const int nv = L + 1;                      // L row terms and the closing sum
    float logsc = 0.0f;
    float v = (lane < nv) ? rowmax[o + lane] : 0.0f;
    for (int base = 0; base < nv; base += 64) {
      const float nxt = (base + 64 + lane < nv) ? rowmax[o + base + 64 + lane] : 0.0f;    // the next block in flight
      const int cnt = min(64, nv - base);
      int k = 0;
      for (; k + 8 <= cnt; k += 8) {
#pragma unroll
        for (int u = 0; u < 8; u++) logsc += readlanef(v, k + u);
      }
      for (; k < cnt; k++) logsc += readlanef(v, k);
      v = nxt;
    }
    if (lane == 0) bias_finish(p, r, g, gating, lc, slot, L, logsc);
  }
}

/*****************************************************************
 * 6. Viterbi filter (one target per wave, 64 lanes, int16 pairs)
 *****************************************************************/
// D->D closure of one Viterbi row: D(k) = max(D0(k), D(k-1) + tDD(k-1)).
// Every tDD <= 0 and the int16 floor at -32768 absorbs, so a saturating chain
// equals the int32 maximum over paths floored at -32768. Inside a lane both
// halves run as packed int16 chains; across lanes the lane maps
// x -> max(C, x + A) are composed by an int32 max-plus scan whose window sums
// depend only on the profile and are set up once per kernel.
template <int ND>
struct DDConst {
  int t[ND];                   // tDD pairs: position d (low half), position d+ND (high half)
  int alo, ahi;                // int32 sums of the two halves
  int k1, k2, k4, k8;          // row-scan window sums; -2^28 where the step has no source lane
  int arow;                    // sum of lane totals from the row start to this lane
  int arow31, arow47;          // the same at lanes 31 and 47 (wave-uniform)
};

template <int ND>
__device__ __forceinline__ void dd_setup(DDConst<ND> &c, const int *__restrict__ tdd, int z)
{
  const int NEG = -(1 << 28);
  int alo = 0, ahi = 0;
#pragma unroll
  for (int d = 0; d < ND; d++) { c.t[d] = tdd[d * 64]; alo += lo16(c.t[d]); ahi += hi16(c.t[d]); }
  c.alo = alo; c.ahi = ahi;
  const int r = z & 15;
  int pre = alo + ahi;                      // inclusive prefix over the lanes of this row
#pragma unroll
  for (int s = 1; s < 16; s <<= 1) { const int v = __shfl_up(pre, s, 16); if (r >= s) pre += v; }
  auto back = [&](int s) { const int v = __shfl_up(pre, s, 16); return (r >= s) ? pre - v : NEG; };
  c.k1 = back(1); c.k2 = back(2); c.k4 = back(4); c.k8 = back(8);
  c.arow = pre;
  c.arow31 = __builtin_amdgcn_readlane(pre, 31);
  c.arow47 = __builtin_amdgcn_readlane(pre, 47);
}

template <int CTRL>
__device__ __forceinline__ int dpp0(int v) { return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xf, 0xf, true); }

template <int ND>
__device__ __forceinline__ void dd_closure(int (&D)[ND], const DDConst<ND> &c, int z)
{
  const int NEG = -(1 << 28);
  // chains inside the lane, no carry in
  int P[ND];
  P[0] = D[0];
#pragma unroll
  for (int d = 1; d < ND; d++) P[d] = max2(D[d], adds2(P[d-1], c.t[d]));
  const int plo = lo16(P[ND-1]), phi = hi16(P[ND-1]);
  // the lane as x -> max(C, x + A), A = alo + ahi; scanned over 16-lane rows
  int C = max(phi, plo + c.ahi);
  C = max(C, dpp0<ROW_SHR1>(C) + c.k1);
  C = max(C, dpp0<ROW_SHR2>(C) + c.k2);
  C = max(C, dpp0<ROW_SHR4>(C) + c.k4);
  C = max(C, dpp0<ROW_SHR8>(C) + c.k8);
  {
    const int c1 = __builtin_amdgcn_readlane(C, 15);
    const int c2 = max(c1 + c.arow31, __builtin_amdgcn_readlane(C, 31));
    const int c3 = max(c2 + c.arow47, __builtin_amdgcn_readlane(C, 47));
    const int row = z >> 4;
    const int rc = (row == 0) ? NEG : (row == 1) ? c1 : (row == 2) ? c2 : c3;
    C = max(C, rc + c.arow);
  }
  const int cin = dppi<WAVE_SHR1>(NEG, C);
  // carries into the low half (from the lane before) and the high half (end of the low half)
  const int clo = max(cin, -32768);
  const int chi = max(plo, max(cin + c.alo, -32768));
  int Q = (int) (((unsigned) clo & 0xffffu) | ((unsigned) chi << 16));
#pragma unroll
  for (int d = 0; d < ND; d++) { Q = adds2(Q, c.t[d]); D[d] = max2(P[d], Q); }
}

// The same closure, found by passes that stop at the fixed point (the lazy-F
// passes of vitfilter.c:200-225, over half-lanes instead of SSE segments).
// Each half of a lane is first closed on its own; then the carries into the
// halves (the previous lane's last position into the low half, the lane's own
// low half into its high half) are applied while any of them raises a first
// position. A chain closed inside each half is raised somewhere only if it is
// raised at its first position, so that one comparison per half decides.
// Rows whose D->D paths reach past DD_LAZY_PASSES half-lane boundaries finish
// with the scan above, which gives the same closure from any partly closed row.
#define DD_LAZY_PASSES 2
template <int ND>
__device__ __forceinline__ void dd_closure_lazy(int (&D)[ND], const DDConst<ND> &c, int z)
{
#pragma unroll
  for (int d = 1; d < ND; d++) D[d] = max2(D[d], adds2(D[d-1], c.t[d]));
  for (int pass = 0; ; pass++) {
    const int nb = dppi<WAVE_SHR1>((int) NEG_PAIR, D[ND-1]);       // lane 0: no lane before, -infinity
    int Q = adds2((int) __builtin_amdgcn_alignbit((unsigned) D[ND-1], (unsigned) nb, 16), c.t[0]);
    const int m0 = max2(D[0], Q);
    if (__ballot(m0 != D[0]) == 0ull) return;
    if (pass == DD_LAZY_PASSES) { dd_closure<ND>(D, c, z); return; }
    D[0] = m0;
#pragma unroll
    for (int d = 1; d < ND; d++) { Q = adds2(Q, c.t[d]); D[d] = max2(D[d], Q); }
  }
}

// Where a list kernel finds its profile tables (template argument TM):
//   TM_LDS   the whole table in LDS, transitions pinned in registers
//   TM_PIN   the table in global memory (L2-resident), transitions pinned
//   TM_TLDS  transitions in LDS, read per row; emission rows from global memory
//   TM_GLOB  the table in global memory, transitions read per row
// Large models take the last two: B positions of every transition pinned in
// registers would not fit the register file, and the whole table not in LDS.
// The per-row reads go through an offset the compiler cannot see through, so
// that they are not hoisted out of the row loop into registers again.
enum { TM_LDS = 0, TM_PIN = 1, TM_TLDS = 2, TM_GLOB = 3 };

__device__ __forceinline__ int opaque_zero()
{
  int v = 0;
  asm volatile("" : "+v"(v));
  return v;
}

// Workgroup sizes of the list kernels: 8 waves, or 4 (one per SIMD, so that a
// wave may hold 512 registers with the accumulation registers) where the
// per-lane state of large models does not fit in 256.
#define VIT_THREADS(ND) ((ND) >= 40 ? 256 : 512)
#define FWD_THREADS(B)  ((B) >= 64 ? 256 : 512)
#define BCK_THREADS(B)  ((B) >= 28 ? 256 : 512)

template <int ND, int TM>
__global__ __launch_bounds__(VIT_THREADS(ND))
void k_vit(const DevProfile p, const uint8_t *__restrict__ res, const int64_t *__restrict__ off,
           const int32_t *__restrict__ len, const int32_t *__restrict__ list, const int32_t *__restrict__ count,
           DevResults r, DevLists l, GateThresh g, int gating, LenCfg lc)
{
#pragma clang fp contract(off)
  static_assert(TM == TM_LDS || TM == TM_TLDS, "Viterbi tables are in LDS, whole or transitions only");
  constexpr bool PIN = (TM == TM_LDS);
  if ((int64_t) blockIdx.x * (blockDim.x >> 6) >= *count) return;
  extern __shared__ int lds[];
  const size_t rowd = (size_t) 64 * ND;
  const int *tbl = p.vf_tbl;                             // emission rows
  if (TM == TM_LDS) { stage_table(lds, p.vf_tbl, p.vf_dwords); tbl = lds; }
  else stage_table(lds, p.vf_tbl + (size_t) p.nrows * rowd, (size_t) P7G_VF_NT * rowd);
  const int z = threadIdx.x & 63;
  const int *tr = ((TM == TM_LDS) ? lds + (size_t) p.nrows * rowd : lds) + z;
  int tBM[ND], tMM[ND], tIM[ND], tDM[ND], tMDS[ND], tMI[ND], tII[ND];
  if (PIN) {
#pragma unroll
    for (int d = 0; d < ND; d++) {
      tBM[d] = tr[VT_BM*rowd + d*64]; tMM[d] = tr[VT_MM*rowd + d*64]; tIM[d] = tr[VT_IM*rowd + d*64]; tDM[d] = tr[VT_DM*rowd + d*64];
      tMDS[d] = tr[VT_MDS*rowd + d*64]; tMI[d] = tr[VT_MI*rowd + d*64]; tII[d] = tr[VT_II*rowd + d*64];
    }
  }
  DDConst<ND> ddc;
  dd_setup<ND>(ddc, tr + VT_DDS*rowd, z);
  const bool loops0 = (p.xwN_loop == 0 && p.xwC_loop == 0 && p.xwJ_loop == 0);
  const int64_t n = *count;

  int32_t *queue = (int32_t *) count + P7G_QOFF;
  for (int64_t t = queue_next(queue, 1); t < n; t = queue_next(queue, 1)) {
    const int slot = uni(list[t]);
    const int L = uni(len[slot]);
    const uint8_t *sq = res + off[slot];
    const int xmove = lc.on ? lc.xmove : (int) p.xw_move[L];
    int xN = p.base_w;
    int xB = (int16_t) (xN + xmove);
    int xJ = -32768, xC = -32768;
    int M[ND], I[ND], D[ND];
#pragma unroll
    for (int d = 0; d < ND; d++) { M[d] = I[d] = D[d] = (int) NEG_PAIR; }
    int status = 0;
    // DPP destinations: lane 0 never has a source and keeps the -infinity fill
    int nbM = (int) NEG_PAIR, nbI = (int) NEG_PAIR, nbD = (int) NEG_PAIR;
    int ms0 = (int) NEG_PAIR;                     // M(i-1, k-1) for the lane's first position
    ResAsc rd;
    if (L > 0) rd.init(sq);

    for (int i = 0; i < L; i++) {
      const int x = rd.at(i);
      const int *e = tbl + (size_t) x * rowd + z;
      const int *trr = tr + (PIN ? 0 : opaque_zero());
#define VTR(A, T, d) (PIN ? A[d] : trr[(T) * rowd + (d) * 64])
      nbI = dppi<WAVE_SHR1>(nbI, I[ND-1]);
      nbD = dppi<WAVE_SHR1>(nbD, D[ND-1]);
      const int is0 = (int) __builtin_amdgcn_alignbit((unsigned) I[ND-1], (unsigned) nbI, 16);
      const int ds0 = (int) __builtin_amdgcn_alignbit((unsigned) D[ND-1], (unsigned) nbD, 16);
      const int xB2 = pair16(xB);
      int xE = (int) NEG_PAIR;
#pragma unroll
      for (int d = ND-1; d >= 0; d--) {
        const int mp = d ? M[d-1] : ms0, ip = d ? I[d-1] : is0, dp = d ? D[d-1] : ds0;
        int sv = adds2(xB2, VTR(tBM, VT_BM, d));
        sv = max2(sv, adds2(mp, VTR(tMM, VT_MM, d)));
        sv = max2(sv, adds2(ip, VTR(tIM, VT_IM, d)));
        sv = max2(sv, adds2(dp, VTR(tDM, VT_DM, d)));
        sv = adds2(sv, e[d * 64]);
        xE = max2(xE, sv);
        I[d] = max2(adds2(M[d], VTR(tMI, VT_MI, d)), adds2(I[d], VTR(tII, VT_II, d)));
        M[d] = sv;
      }
      // D(k) <- M(k-1) + tMD(k-1) on the new row; the shifted row is also the next row's M(i-1, k-1)
      nbM = dppi<WAVE_SHR1>(nbM, M[ND-1]);
      ms0 = (int) __builtin_amdgcn_alignbit((unsigned) M[ND-1], (unsigned) nbM, 16);
      int Dm = (int) NEG_PAIR;
#pragma unroll
      for (int d = 0; d < ND; d++) { D[d] = adds2(d ? M[d-1] : ms0, VTR(tMDS, VT_MDS, d)); Dm = max2(Dm, D[d]); }
#undef VTR
      // With the N, C and J loops at zero (as the conversion sets them,
      // p7_oprofile.c:888-897) the specials only change, and the overflow test
      // only fires, when the row's xE exceeds th; that holds if it holds on
      // any lane, so a ballot stands in for the wave maximum on other rows.
      const int xel = max(lo16(xE), hi16(xE));
      const int th = loops0 ? min(min(xC - p.xwE_move, xJ - p.xwE_loop), 32766) : -(1 << 30);
      if (__ballot(xel > th) != 0ull) {
        const int xEs = uni(wave_max_i(xel));
        if (xEs >= 32767) { status = 16; break; }
        xN = (int16_t) (xN + p.xwN_loop);
        xC = (int16_t) max(xC + p.xwC_loop, xEs + p.xwE_move);
        xJ = (int16_t) max(xJ + p.xwJ_loop, xEs + p.xwE_loop);
        xB = (int16_t) max(xJ + xmove, xN + xmove);
      }
      // the lazy-F test Dmax + ddbound > xB (vitfilter.c:198-199) holds if it
      // holds on any lane: a ballot instead of a wave maximum
      const int dml = max(lo16(Dm), hi16(Dm));
      if (__ballot(dml + p.ddbound_w > xB) != 0ull) dd_closure_lazy<ND>(D, ddc, z);
    }
    if (z == 0) {
      float sc;
      if (status == 16) sc = INFINITY;
      else if (xC > -32768) {
        sc = (float) xC + (float) xmove - (float) p.base_w;
        sc = __fdiv_rn(sc, p.scale_w);
        sc = (float) ((double) sc - 3.0);
      } else sc = -INFINITY;
      r.vfsc[slot] = sc;
      r.vstatus[slot] = status;
      if (gating) {
        const float seq = (float) __ddiv_rn((double) (sc - r.filtersc[slot]), LOG2_D);
        if (seq >= g.T2v) r.stage[slot] = 3;
        else r.stage[slot] = 2;
      }
    }
  }
}

// Viterbi filter for large models, W waves per target. Each lane's ND dwords
// (positions d and d+ND of the lane as int16 pairs) are split into W slices
// of ND/W; wave w owns slice w of every lane, with its transitions pinned in
// registers. A slice's first dword takes its predecessor from the previous
// slice's last dword (both halves at once), and slice 0 from slice W-1 of the
// lane before (low half) and of its own lane (high half), through LDS. The
// D->D closure composes per-slice maps x -> max(C, x + A) (int32 max-plus,
// floored at -32768 as the saturating chain is) into the lane's map, scans
// the lanes as the one-wave kernel does, then closes each slice from its
// carries. The row's xE and the lazy-F test's D maximum are workgroup maxima.
// Three barriers per row; results equal the one-wave kernel's bit for bit.
template <int ND, int W>
__global__ __launch_bounds__(64 * W)
void k_vit_wg(const DevProfile p, const uint8_t *__restrict__ res, const int64_t *__restrict__ off,
              const int32_t *__restrict__ len, const int32_t *__restrict__ list, const int32_t *__restrict__ count,
              DevResults r, DevLists l, GateThresh g, int gating, LenCfg lc)
{
#pragma clang fp contract(off)
  static_assert(ND % W == 0, "slices of equal width");
  constexpr int NW = ND / W;
  const int NEG = -(1 << 28);
  __shared__ int sh_pm[W][64], sh_pi[W][64], sh_pd[W][64];   // previous row: each slice's last dword
  __shared__ int sh_m[W][64];                                // this row's M at each slice's last dword
  __shared__ int sh_plo[W][64], sh_phi[W][64];               // this row's slice chain ends (int32, per half)
  __shared__ int sh_e[W], sh_dm[W];                          // the waves' xE and D maxima
  __shared__ int64_t sh_t;
  if ((int64_t) blockIdx.x >= *count) return;
  const int wv = threadIdx.x >> 6, z = threadIdx.x & 63;
  const int d0 = wv * NW;
  const size_t rowd = (size_t) 64 * ND;
  const int *tbl = p.vf_tbl;
  const int *tr = tbl + (size_t) p.nrows * rowd + z;
  int tBM[NW], tMM[NW], tIM[NW], tDM[NW], tMDS[NW], tMI[NW], tII[NW], tDD[NW];
#pragma unroll
  for (int j = 0; j < NW; j++) {
    const size_t c = (size_t) (d0 + j) * 64;
    tBM[j] = tr[VT_BM*rowd + c]; tMM[j] = tr[VT_MM*rowd + c]; tIM[j] = tr[VT_IM*rowd + c]; tDM[j] = tr[VT_DM*rowd + c];
    tMDS[j] = tr[VT_MDS*rowd + c]; tMI[j] = tr[VT_MI*rowd + c]; tII[j] = tr[VT_II*rowd + c]; tDD[j] = tr[VT_DDS*rowd + c];
  }
  // int32 sums of the D->D steps of every slice of this lane, per half, and
  // the lane-scan constants of the whole lane (as the one-wave kernel's)
  int alo_v[W], ahi_v[W];
#pragma unroll
  for (int v = 0; v < W; v++) {
    int a = 0, b = 0;
    for (int j = 0; j < NW; j++) { const int t = tr[VT_DDS*rowd + (size_t) (v * NW + j) * 64]; a += lo16(t); b += hi16(t); }
    alo_v[v] = a; ahi_v[v] = b;
  }
  DDConst<ND> ddc;
  dd_setup<ND>(ddc, tr + VT_DDS*rowd, z);
  const bool loops0 = (p.xwN_loop == 0 && p.xwC_loop == 0 && p.xwJ_loop == 0);
  const int64_t n = *count;
  int32_t *queue = (int32_t *) count + P7G_QOFF;

  for (;;) {
    if (threadIdx.x == 0) sh_t = (int64_t) atomicAdd(queue, 1);
    __syncthreads();
    const int64_t t = sh_t;
    if (t >= n) break;
    const int slot = list[t];
    const int L = len[slot];
    const uint8_t *sq = res + off[slot];
    const int xmove = lc.on ? lc.xmove : (int) p.xw_move[L];
    int xN = p.base_w;
    int xB = (int16_t) (xN + xmove);
    int xJ = -32768, xC = -32768;
    int M[NW], I[NW], D[NW];
#pragma unroll
    for (int j = 0; j < NW; j++) { M[j] = I[j] = D[j] = (int) NEG_PAIR; }
    sh_pm[wv][z] = (int) NEG_PAIR; sh_pi[wv][z] = (int) NEG_PAIR; sh_pd[wv][z] = (int) NEG_PAIR;
    int status = 0;
    ResAsc rd;
    if (L > 0) rd.init(sq);
    for (int i = 0; i < L; i++) {
      const int x = rd.at(i);
      const int *e = tbl + (size_t) x * rowd + z;
      __syncthreads();                                   // the previous row's slice ends are written
      int mS, iS, dS;                                    // the previous row's predecessors of the slice's first dword
      if (wv > 0) { mS = sh_pm[wv-1][z]; iS = sh_pi[wv-1][z]; dS = sh_pd[wv-1][z]; }
      else {
        const int pm = (z > 0) ? sh_pm[W-1][z-1] : (int) NEG_PAIR, pi = (z > 0) ? sh_pi[W-1][z-1] : (int) NEG_PAIR;
        const int pd = (z > 0) ? sh_pd[W-1][z-1] : (int) NEG_PAIR;
        mS = (int) __builtin_amdgcn_alignbit((unsigned) sh_pm[W-1][z], (unsigned) pm, 16);
        iS = (int) __builtin_amdgcn_alignbit((unsigned) sh_pi[W-1][z], (unsigned) pi, 16);
        dS = (int) __builtin_amdgcn_alignbit((unsigned) sh_pd[W-1][z], (unsigned) pd, 16);
      }
      const int xB2 = pair16(xB);
      int xE = (int) NEG_PAIR;
#pragma unroll
      for (int j = NW-1; j >= 0; j--) {
        const int mp = j ? M[j-1] : mS, ip = j ? I[j-1] : iS, dp = j ? D[j-1] : dS;
        int sv = adds2(xB2, tBM[j]);
        sv = max2(sv, adds2(mp, tMM[j]));
        sv = max2(sv, adds2(ip, tIM[j]));
        sv = max2(sv, adds2(dp, tDM[j]));
        sv = adds2(sv, e[(size_t) (d0 + j) * 64]);
        xE = max2(xE, sv);
        I[j] = max2(adds2(M[j], tMI[j]), adds2(I[j], tII[j]));
        M[j] = sv;
      }
      sh_m[wv][z] = M[NW-1];
      __syncthreads();                                   // this row's slice-end M values are written
      int mN;
      if (wv > 0) mN = sh_m[wv-1][z];
      else {
        const int pm = (z > 0) ? sh_m[W-1][z-1] : (int) NEG_PAIR;
        mN = (int) __builtin_amdgcn_alignbit((unsigned) sh_m[W-1][z], (unsigned) pm, 16);
      }
      int Dm = (int) NEG_PAIR;
#pragma unroll
      for (int j = 0; j < NW; j++) { D[j] = adds2(j ? M[j-1] : mN, tMDS[j]); Dm = max2(Dm, D[j]); }
      // the slice's chains without carries in
      int P[NW];
      P[0] = D[0];
#pragma unroll
      for (int j = 1; j < NW; j++) P[j] = max2(D[j], adds2(P[j-1], tDD[j]));
      sh_plo[wv][z] = lo16(P[NW-1]); sh_phi[wv][z] = hi16(P[NW-1]);
      {
        const int xel = wave_max_i(max(lo16(xE), hi16(xE)));
        const int dml = wave_max_i(max(lo16(Dm), hi16(Dm)));
        if (z == 0) { sh_e[wv] = xel; sh_dm[wv] = dml; }
      }
      __syncthreads();                                   // slice chains and the waves' maxima are written
      int xEs = sh_e[0], dmax = sh_dm[0];
#pragma unroll
      for (int v = 1; v < W; v++) { xEs = max(xEs, sh_e[v]); dmax = max(dmax, sh_dm[v]); }
      const int th = loops0 ? min(min(xC - p.xwE_move, xJ - p.xwE_loop), 32766) : -(1 << 30);
      if (xEs > th) {
        if (xEs >= 32767) { status = 16; break; }
        xN = (int16_t) (xN + p.xwN_loop);
        xC = (int16_t) max(xC + p.xwC_loop, xEs + p.xwE_move);
        xJ = (int16_t) max(xJ + p.xwJ_loop, xEs + p.xwE_loop);
        xB = (int16_t) max(xJ + xmove, xN + xmove);
      }
      if (dmax + p.ddbound_w > xB) {
        // the lane's map: the low halves of the W slices in order, then the high halves
        int Clo = NEG, Alo = 0, Chi = NEG, Ahi = 0;
#pragma unroll
        for (int v = 0; v < W; v++) {
          Clo = max(sh_plo[v][z], Clo + alo_v[v]); Alo += alo_v[v];
          Chi = max(sh_phi[v][z], Chi + ahi_v[v]); Ahi += ahi_v[v];
        }
        int C = max(Chi, Clo + Ahi);
        C = max(C, dpp0<ROW_SHR1>(C) + ddc.k1);
        C = max(C, dpp0<ROW_SHR2>(C) + ddc.k2);
        C = max(C, dpp0<ROW_SHR4>(C) + ddc.k4);
        C = max(C, dpp0<ROW_SHR8>(C) + ddc.k8);
        {
          const int c1 = __builtin_amdgcn_readlane(C, 15);
          const int c2 = max(c1 + ddc.arow31, __builtin_amdgcn_readlane(C, 31));
          const int c3 = max(c2 + ddc.arow47, __builtin_amdgcn_readlane(C, 47));
          const int row = z >> 4;
          const int rc = (row == 0) ? NEG : (row == 1) ? c1 : (row == 2) ? c2 : c3;
          C = max(C, rc + ddc.arow);
        }
        const int cin = dppi<WAVE_SHR1>(NEG, C);        // the end of the lane before
        // carries into this slice's low half (through the slices before it) and
        // high half (the lane's whole low half, then the slices before it)
        int clo = cin, chi = max(Clo, cin + Alo);
        for (int v = 0; v < wv; v++) {
          clo = max(sh_plo[v][z], clo + alo_v[v]);
          chi = max(sh_phi[v][z], chi + ahi_v[v]);
        }
        clo = max(clo, -32768); chi = max(chi, -32768);
        int Q = (int) (((unsigned) clo & 0xffffu) | ((unsigned) chi << 16));
#pragma unroll
        for (int j = 0; j < NW; j++) { Q = adds2(Q, tDD[j]); D[j] = max2(P[j], Q); }
      }
      sh_pm[wv][z] = M[NW-1]; sh_pi[wv][z] = I[NW-1]; sh_pd[wv][z] = D[NW-1];
    }
    if (threadIdx.x == 0) {
      float sc;
      if (status == 16) sc = INFINITY;
      else if (xC > -32768) {
        sc = (float) xC + (float) xmove - (float) p.base_w;
        sc = __fdiv_rn(sc, p.scale_w);
        sc = (float) ((double) sc - 3.0);
      } else sc = -INFINITY;
      r.vfsc[slot] = sc;
      r.vstatus[slot] = status;
      if (gating) {
        const float seq = (float) __ddiv_rn((double) (sc - r.filtersc[slot]), LOG2_D);
        if (seq >= g.T2v) r.stage[slot] = 3;
        else r.stage[slot] = 2;
      }
    }
    __syncthreads();                                     // sh_t and the slice ends are free for the next target
  }
}

/*****************************************************************
 * 7. Forward parser (one target per wave, fp32, B positions per lane)
 *****************************************************************/
// The D->D chain of a parser row as an affine scan: lane maps f(x) = P x + C
// composed over 16-lane rows with DPP, rows joined by readlane carries. The
// products P depend only on the profile's D->D odds: they are computed once
// per kernel (AffConst, with the same multiplications in the same order as a
// per-row computation would make), so a row costs the C recurrence alone.
struct AffConst {
  float w1, w2, w4, w8;      // window products the four row steps multiply by
  float prow;                // product over the row part the lane's window covers after the steps
  float pa, pb;              // wave-uniform row products used by the row carries
};

template <int B>
__device__ __forceinline__ void aff_fwd_setup(AffConst &k, const float (&a)[B])
{
  float P = 1.0f;
#pragma unroll
  for (int j = 0; j < B; j++) P *= a[j];
  k.w1 = P; P = P * dppf<ROW_SHR1>(1.0f, P);
  k.w2 = P; P = P * dppf<ROW_SHR2>(1.0f, P);
  k.w4 = P; P = P * dppf<ROW_SHR4>(1.0f, P);
  k.w8 = P; P = P * dppf<ROW_SHR8>(1.0f, P);
  k.prow = P; k.pa = readlanef(P, 31); k.pb = readlanef(P, 47);
}

template <int B>
__device__ __forceinline__ void aff_bck_setup(AffConst &k, const float (&a)[B])
{
  float P = 1.0f;
#pragma unroll
  for (int j = B-1; j >= 0; j--) P *= a[j];
  k.w1 = P; P = P * dppf<ROW_SHL1>(1.0f, P);
  k.w2 = P; P = P * dppf<ROW_SHL2>(1.0f, P);
  k.w4 = P; P = P * dppf<ROW_SHL4>(1.0f, P);
  k.w8 = P; P = P * dppf<ROW_SHL8>(1.0f, P);
  k.prow = P; k.pa = readlanef(P, 32); k.pb = readlanef(P, 16);
}

template <int B>
__device__ __forceinline__ void affine_fwd_scan(float (&D)[B], const float (&a)[B], const AffConst &k, int z)
{
  // D(j) = D0(j) + a(j) * D(j-1), carried across lanes in increasing k
  float C = 0.0f;
#pragma unroll
  for (int j = 0; j < B; j++) C = D[j] + a[j] * C;
  C = C + k.w1 * dppz<ROW_SHR1>(C);
  C = C + k.w2 * dppz<ROW_SHR2>(C);
  C = C + k.w4 * dppz<ROW_SHR4>(C);
  C = C + k.w8 * dppz<ROW_SHR8>(C);
  {
    const float c1 = readlanef(C, 15);
    const float c2 = k.pa * c1 + readlanef(C, 31);
    const float c3 = k.pb * c2 + readlanef(C, 47);
    const int row = z >> 4;
    const float rc = (row == 0) ? 0.0f : (row == 1) ? c1 : (row == 2) ? c2 : c3;
    C = C + k.prow * rc;
  }
  const float cin = dppz<WAVE_SHR1>(C);
  float cur = cin;
#pragma unroll
  for (int j = 0; j < B; j++) { cur = D[j] + a[j] * cur; D[j] = cur; }
}

// a(j) gives the D->D odds of the lane's position j (read twice per call, so
// that large models need not hold them in registers through the scan).
template <int B, typename A>
__device__ __forceinline__ void affine_bck_scan(float (&D)[B], A a, const AffConst &k, int z)
{
  // D(j) = c(j) + a(j) * D(j+1), carried across lanes in decreasing k
  float C = 0.0f;
#pragma unroll
  for (int j = B-1; j >= 0; j--) C = D[j] + a(j) * C;
  C = C + k.w1 * dppz<ROW_SHL1>(C);
  C = C + k.w2 * dppz<ROW_SHL2>(C);
  C = C + k.w4 * dppz<ROW_SHL4>(C);
  C = C + k.w8 * dppz<ROW_SHL8>(C);
  {
    const float c2 = readlanef(C, 48);
    const float c1 = k.pa * c2 + readlanef(C, 32);
    const float c0 = k.pb * c1 + readlanef(C, 16);
    const int row = z >> 4;
    const float rc = (row == 3) ? 0.0f : (row == 2) ? c2 : (row == 1) ? c1 : c0;
    C = C + k.prow * rc;
  }
  const float cin = dppz<WAVE_SHL1>(C);
  float cur = cin;
#pragma unroll
  for (int j = B-1; j >= 0; j--) { cur = D[j] + a(j) * cur; D[j] = cur; }
}

// The end of an envelope's null2 pass (MODE 3 of k_fwd), one wave: the
// posterior sums accM/accI (this lane's B positions) and pN/pJ/pC over the
// envelope's L rows become weights (times 1/L); null2[x] for canonical x is
// the weighted sum of the positions' match odds, plus the insert weights
// (odds 1) and the N, C, J weights (impl_sse/null2.c:80-107); a degenerate
// code's value is the average over the residues it stands for
// (esl_alphabet.c:1331-1341), and gap, '*' and '~' get 1. Lane x holds
// null2[x]. The correction is sum over the envelope's residues of
// logf(null2[x]) (p7_domaindef.c:925-931). A scale product that ended
// infinite is the decoding's range error: status 16, no correction.
template <int B>
__device__ __forceinline__ void null2_epilogue(const DevProfile &p, const float *tbl, size_t rowf, int z,
                                               float (&accM)[B], float (&accI)[B], float pN, float pJ, float pC,
                                               float sp, const uint8_t *sq, int L, float *dc_out, int32_t *st_out)
{
  if (isinf(sp)) {
    if (z == 0) { *dc_out = NAN; *st_out = 16; }
    return;
  }
  const float norm = (float) (1.0 / (double) L);
  float iw = 0.0f;
#pragma unroll
  for (int j = 0; j < B; j++) { accM[j] = __fmul_rn(accM[j], norm); iw = __fadd_rn(iw, __fmul_rn(accI[j], norm)); }
  const float xfac = __fadd_rn(__fadd_rn(__fmul_rn(pN, norm), __fmul_rn(pC, norm)), __fmul_rn(pJ, norm));
  float v = 1.0f;
  for (int x = 0; x < p.K; x++) {
    const float *e = tbl + (size_t) x * rowf + z;
    float part = iw;
#pragma unroll
    for (int j = 0; j < B; j++) part = __fadd_rn(part, __fmul_rn(accM[j], e[j * 64]));
    const float s = wave_sum_f(part);
    if (z == x) v = __fadd_rn(s, xfac);
  }
  {
    const uint32_t dg = (z < 32) ? p.degen[z] : 0u;
    float r = 0.0f;
    for (int y = 0; y < p.K; y++) {
      const float vy = __shfl(v, y);
      if ((dg >> y) & 1u) r = __fadd_rn(r, vy);
    }
    if (z > p.K && z <= p.Kp - 3) v = __fdiv_rn(r, (float) (dg >> 24));
  }
  float part = 0.0f;
  for (int b0 = 0; b0 < L; b0 += 64) {
    const int pos = b0 + z;
    const int x = (pos < L) ? sq[pos] : 0;
    const float vx = __shfl(v, x & 63);
    if (pos < L) part = __fadd_rn(part, logf(vx));
  }
  const float dc = wave_sum_f(part);
  if (z == 0) { *dc_out = dc; *st_out = 0; }
}

// p7_alidisplay_EncodePostProb (p7_alidisplay.c:520-524) as a number: '*' is 10
__device__ __forceinline__ int pp_code(float p)
{
  const double q = (double) p + 0.05;
  return (q >= 1.0) ? 10 : (int) (q * 10.0);
}

// One row of the optimal accuracy fill (p7_OptimalAccuracy, impl_sse/
// optacc.c:60-175) inside the null2 Forward pass, from the row's posteriors,
// with the traceback's choices (select_m/d/i/e/c/j/b, optacc.c:288-423) made
// here from the same values and stored for k_oa_trace: four bits per cell
// (lane-minor words), one word per row. Backward's M and I values of the row
// come in registers (pbM, pbI: PRE, loaded by the caller before the row's
// Forward part) or are loaded here from cr, eight positions at a time, so
// that the row waits on memory once per eight positions, not once per cell.
// As in the SSE code, a predecessor
// behind a transition of probability 0 contributes +0.0 to the fill's maximum
// and -infinity to the traceback's choice; D(1) is -infinity; the reference's
// striped padding cells (model positions M+1..4Q) hold 0.0 and take part in
// E's maximum and in select_e's scan, which runs in the striped order
// (k = r*Q + q + 1, q outer), an M replacing on >= and a D on >. Values are
// -infinity or >= 0, so their bit patterns order as integers. The profile is
// in unihit mode (E move 1, E loop 0; N, J, C loop floop and move fmove).
template <int B, int NW, bool PIN, bool PRE>
__device__ __forceinline__ void oa_row(const DevProfile &p, int z, int i,
                                       const float (&tBM)[B], const float (&tMM)[B], const float (&tIM)[B],
                                       const float (&tDM)[B], const float (&tMDS)[B], const float (&tMI)[B],
                                       const float (&tII)[B], const float *trr, size_t rowf,
                                       const float (&fM)[B], const float (&fI)[B], const float *cr,
                                       const float *pbM, const float *pbI, float totr,
                                       float (&accM)[B], float (&accI)[B], float ppN, float ppJ, float ppC,
                                       float floop, float fmove, float (&oaM)[B], float (&oaI)[B], float (&oaD)[B],
                                       float &oaN, float &oaJ, float &oaB, float &oaC, uint32_t *obits, uint32_t *orow,
                                       uint8_t *opp)
{
#define OTR(A, T, j) (PIN ? A[j] : trr[(T) * rowf + (j) * 64])
  const float NI = -INFINITY;
  const int M = p.M;
  const int Q = max(2, (M - 1) / 4 + 1);
  uint32_t nib[NW];
#pragma unroll
  for (int w = 0; w < NW; w++) nib[w] = 0u;
  constexpr int CH = 8;
  float cbM[CH], cbI[CH];
  uint32_t ppw = 0u;
  uint32_t *opw = (uint32_t *) opp + (size_t) (i - 1) * n2_pp_words(B) * 64 + z;
  // row i-1's values at the lane's first position's predecessor (lane z-1's last)
  const float mpL = dppf<WAVE_SHR1>(NI, oaM[B-1]), ipL = dppf<WAVE_SHR1>(NI, oaI[B-1]), dpL = dppf<WAVE_SHR1>(NI, oaD[B-1]);
  // M and I in place, last position first: position j reads row i-1 at j-1 and j
#pragma unroll
  for (int j = B - 1; j >= 0; j--) {
    const int k = z * B + j + 1;
    const bool first = (k == 1);
    const int c8 = (B - 1 - j) % CH;
    if (! PRE && c8 == 0) {
#pragma unroll
      for (int u = 0; u < CH; u++)
        if (j - u >= 0) { cbM[u] = cr[(j - u) * 64]; cbI[u] = cr[rowf + (j - u) * 64]; }
    }
    const float bm = PRE ? pbM[j] : cbM[c8], bi = PRE ? pbI[j] : cbI[c8];
    // the posteriors of row i: (f * b) * (scaleproduct * scale_i), summed for null2
    const float pm = __fmul_rn(__fmul_rn(fM[j], bm), totr);
    const float pi = __fmul_rn(__fmul_rn(fI[j], bi), totr);
    accM[j] = __fadd_rn(pm, accM[j]);
    accI[j] = __fadd_rn(pi, accI[j]);
    // the cell's posterior probability codes for the alignment display
    ppw |= (uint32_t) (pp_code(pm) | (pp_code(pi) << 4)) << ((j % 4) * 8);
    if (j % 4 == 0) { opw[(j / 4) * 64] = ppw; ppw = 0u; }
    const float mp = j ? oaM[j-1] : mpL, ip = j ? oaI[j-1] : ipL, dp = j ? oaD[j-1] : dpL;
    const bool tb = OTR(tBM, VT_BM, j) > 0.0f, tm = OTR(tMM, VT_MM, j) > 0.0f;
    const bool ti = OTR(tIM, VT_IM, j) > 0.0f, td = OTR(tDM, VT_DM, j) > 0.0f;
    // select_m: paths M, I, D, B; the first maximum wins; node 0 reads 0.0
    const float p0 = tm ? (first ? 0.0f : mp) : NI, p1 = ti ? (first ? 0.0f : ip) : NI;
    const float p2 = td ? (first ? 0.0f : dp) : NI, p3 = tb ? oaB : NI;
    int c = 0;
    float bv = p0;
    if (p1 > bv) { c = 1; bv = p1; }
    if (p2 > bv) { c = 2; bv = p2; }
    if (p3 > bv) { c = 3; }
    float v = tb ? oaB : 0.0f;
    v = fmaxf(v, tm ? mp : 0.0f);
    v = fmaxf(v, ti ? ip : 0.0f);
    v = fmaxf(v, td ? dp : 0.0f);
    // select_i: M over I on ties
    const bool tmi = OTR(tMI, VT_MI, j) > 0.0f, tii = OTR(tII, VT_II, j) > 0.0f;
    const int ci = ((tmi ? oaM[j] : NI) >= (tii ? oaI[j] : NI)) ? 0 : 1;
    oaI[j] = __fadd_rn(fmaxf(tmi ? oaM[j] : 0.0f, tii ? oaI[j] : 0.0f), pi);
    oaM[j] = __fadd_rn(v, pm);
    nib[j / 8] |= (uint32_t) (c | (ci << 2)) << ((j % 8) * 4);
  }
  // D(k) = max(tMD(k-1) ? M(k-1) : 0, tDD(k-1) ? D(k-1) : 0), D(1) = -infinity:
  // per position the map x -> max(c, p ? x : -inf), composed along the lane,
  // scanned across the lanes, then applied from the carry of the lanes before
  const float mL = dppf<WAVE_SHR1>(NI, oaM[B-1]);
  float C = NI;
  int P = 1;
#pragma unroll
  for (int j = 0; j < B; j++) {
    const int k = z * B + j + 1;
    const bool tmd = OTR(tMDS, VT_MDS, j) > 0.0f, tdd = trr[VT_DDS * rowf + j * 64] > 0.0f;
    const float A = tmd ? (j ? oaM[j-1] : mL) : 0.0f;
    const int pj = (k != 1 && tdd) ? 1 : 0;
    const float cj = (k == 1) ? NI : (tdd ? A : fmaxf(A, 0.0f));
    C = fmaxf(cj, pj ? C : NI);
    P &= pj;
  }
  // x -> max(C, x + W) with W = 0 (pass) or -infinity: an exact max-plus scan
  float W = P ? 0.0f : NI;
  wave_maxplus_scan(C, W);
  float cur = dppf<WAVE_SHR1>(NI, C);
#pragma unroll
  for (int j = 0; j < B; j++) {
    const int k = z * B + j + 1;
    const float mprev = j ? oaM[j-1] : mL;
    const bool tmd = OTR(tMDS, VT_MDS, j) > 0.0f, tdd = trr[VT_DDS * rowf + j * 64] > 0.0f;
    // select_d: M over D on ties; node 1 has neither
    const float a0 = (k == 1) ? NI : (tmd ? mprev : NI), a1 = (k == 1) ? NI : (tdd ? cur : NI);
    nib[j / 8] |= (uint32_t) ((a0 >= a1) ? 0 : 1) << ((j % 8) * 4 + 3);
    const float A = tmd ? mprev : 0.0f;
    const float cj = (k == 1) ? NI : (tdd ? A : fmaxf(A, 0.0f));
    cur = fmaxf(cj, (k != 1 && tdd) ? cur : NI);
    oaD[j] = cur;
  }
  // E(i): the row maximum over the model's cells (and the striped padding's 0.0)
  float el = NI;
#pragma unroll
  for (int j = 0; j < B; j++) if (z * B + j + 1 <= M) el = fmaxf(el, fmaxf(oaM[j], oaD[j]));
  const bool refpad = 4 * Q > M;
  float V = __int_as_float(wave_max_i(__float_as_int(el)));
  if (refpad) V = fmaxf(V, 0.0f);
  // select_e: the last M cell of value V in the striped order, else the first D cell
  auto key = [Q](int k) {                       // k - 1 < 4Q: the lane r by three compares
    const int r = (k - 1 >= Q) + (k - 1 >= 2 * Q) + (k - 1 >= 3 * Q);
    return (k - 1 - r * Q) * 4 + r; };
  int km = -1, kd = 0x7fffffff;
#pragma unroll
  for (int j = 0; j < B; j++) {
    const int k = z * B + j + 1;
    if (k <= M && oaM[j] == V) km = max(km, key(k));
    if (k <= M && oaD[j] == V) kd = min(kd, key(k));
  }
  int kmax = wave_max_i(km);
  int fail = 0, es = 0, ek = 0;
  // padding M cells hold V too, and the last of them (k = 4Q, key 4Q - 1)
  // comes after every model cell in the striped order
  if (refpad && V == 0.0f) fail = 1;
  if (kmax >= 0) { ek = (kmax & 3) * Q + (kmax >> 2) + 1; }
  else {
    const int kmin = -wave_max_i(-kd);
    if (kmin == 0x7fffffff) fail = 1;
    else { es = 1; ek = (kmin & 3) * Q + (kmin >> 2) + 1; }
  }
  // specials (unihit: E loop 0, E move 1), with select_j/c/b's choices
  const float eLoop = 0.0f, eMove = 1.0f;
  const float j0 = (floop == 0.0f) ? 0.0f : __fadd_rn(oaJ, ppJ), j1 = (eLoop == 0.0f) ? 0.0f : V;
  const int bj = (((floop == 0.0f) ? NI : __fadd_rn(oaJ, ppJ)) > ((eLoop == 0.0f) ? NI : V)) ? 1 : 0;
  const float nJ = fmaxf(j0, j1);
  const float c0 = (floop == 0.0f) ? 0.0f : __fadd_rn(oaC, ppC), c1 = (eMove == 0.0f) ? 0.0f : V;
  const int bc = (((floop == 0.0f) ? NI : __fadd_rn(oaC, ppC)) > ((eMove == 0.0f) ? NI : V)) ? 1 : 0;
  const float nC = fmaxf(c0, c1);
  const float nN = (floop == 0.0f) ? 0.0f : __fadd_rn(oaN, ppN);
  const float b0 = (fmove == 0.0f) ? 0.0f : nN, b1 = (fmove == 0.0f) ? 0.0f : nJ;
  const int bb = (((fmove == 0.0f) ? NI : nN) > ((fmove == 0.0f) ? NI : nJ)) ? 1 : 0;
  oaJ = nJ; oaC = nC; oaN = nN; oaB = fmaxf(b0, b1);
#pragma unroll
  for (int w = 0; w < NW; w++) obits[((size_t) (i - 1) * NW + w) * 64 + z] = nib[w];
  if (z == 0)
    orow[i] = (uint32_t) ek | ((uint32_t) es << OA_ESTATE) | ((uint32_t) bc << OA_CBIT) | ((uint32_t) bj << OA_JBIT) |
              ((uint32_t) bb << OA_BBIT) | ((uint32_t) fail << OA_EFAIL);
#undef OTR
}

// MODE 0: the parser. MODE 1: the rows also go to the domain decoding pool
// (a separate instance, so that the plain parser keeps its register budget).
// MODE 2: envelope scores: the list is of region copies, each scored in unihit
// mode (E move 1, E loop 0) with the N/J/C odds of its whole target's length;
// its row scales go to the scale pool when one is given.
// MODE 3: the same Forward pass again for null2: each row's posteriors
// (p7_Decoding, impl_sse/decoding.c:109-139) from this row and the stored
// Backward row, summed over the rows; then the null2 odds
// (p7_Null2_ByExpectation, impl_sse/null2.c:59-118) and the envelope's
// correction sum_pos logf(null2[x_pos]) (p7_domaindef.c:925-931).
template <int B, int TM, int MODE>
__global__ __launch_bounds__(FWD_THREADS(B))
void k_fwd(const DevProfile p, const uint8_t *__restrict__ res, const int64_t *__restrict__ off,
           const int32_t *__restrict__ len, const int32_t *__restrict__ list, const int32_t *__restrict__ count,
           DevResults r, DevLists l, GateThresh g, int gating, LenCfg lc,
           float *__restrict__ pool, int64_t pool_size, int64_t *__restrict__ scale_off, float *__restrict__ xmx_out,
           DomBufs dom)
{
  constexpr bool PIN = (TM == TM_LDS || TM == TM_PIN);
  if ((int64_t) blockIdx.x * (blockDim.x >> 6) >= *count) return;
  extern __shared__ float ldsf[];
  const size_t rowf = (size_t) 64 * B;
  const float *tbl = p.fb_tbl;                           // emission rows
  if (TM == TM_LDS) { stage_table(ldsf, p.fb_tbl, p.fb_floats); tbl = ldsf; }
  if (TM == TM_TLDS) stage_table(ldsf, p.fb_tbl + (size_t) p.nrows * rowf, (size_t) P7G_FB_NT * rowf);
  const int z = threadIdx.x & 63;
  const float *tr = ((TM == TM_TLDS) ? ldsf : tbl + (size_t) p.nrows * rowf) + z;
  float tBM[B], tMM[B], tIM[B], tDM[B], tMDS[B], tMI[B], tII[B];
  if (PIN) {
#pragma unroll
    for (int j = 0; j < B; j++) {
      tBM[j] = tr[VT_BM*rowf + j*64]; tMM[j] = tr[VT_MM*rowf + j*64]; tIM[j] = tr[VT_IM*rowf + j*64]; tDM[j] = tr[VT_DM*rowf + j*64];
      tMDS[j] = tr[VT_MDS*rowf + j*64]; tMI[j] = tr[VT_MI*rowf + j*64]; tII[j] = tr[VT_II*rowf + j*64];
    }
  }
  AffConst akc;
  {
    float tdd[B];
#pragma unroll
    for (int j = 0; j < B; j++) tdd[j] = tr[VT_DDS*rowf + j*64];
    aff_fwd_setup<B>(akc, tdd);
  }
  constexpr bool DOM = (MODE == 1);
  constexpr bool ENV = (MODE >= 2);
  const int64_t n = ENV ? min((int64_t) *count, dom.env_cap) : (int64_t) *count;
  const float eMove = ENV ? 1.0f : p.xfE_move, eLoop = ENV ? 0.0f : p.xfE_loop;

  int32_t *queue = (int32_t *) count + P7G_QOFF;
  for (int64_t t = queue_next(queue, 1); t < n; t = queue_next(queue, 1)) {
    const int slot = uni(list[t]);
    const int L = uni(len[slot]);
    const uint8_t *sq = res + off[slot];
    const float floop = ENV ? dom.env_loop[slot] : lc.on ? lc.floop : p.xf_loop[L];
    const float fmove = ENV ? dom.env_move[slot] : lc.on ? lc.fmove : p.xf_move[L];
    float *scl = nullptr;
    // null2: the region's area (Backward special rows, Backward M/I rows)
    float *area = nullptr;
    if (MODE == 3) {
      const int64_t o = dom.n2off[slot];
      area = dom.n2pool + (((int64_t) (unsigned) uni((int) (o >> 32)) << 32) | (unsigned) uni((int) o));
    }
    const float *xb = nullptr, *cells = nullptr;
    float sp = 0.0f, pN = 0.0f, pJ = 0.0f, pC = 0.0f;
    int own = 0;
    if (MODE == 3) {
      const int bs = uni(dom.n2st[slot]);
      if ((bs & 0xff) != 0 || L < 1) {            // the envelope's Backward failed
        if (z == 0) { dom.n2dc[slot] = NAN; dom.n2st[slot] = 16; }
        continue;
      }
      own = (bs & 0x100) != 0;
      xb = area;
      cells = xb + n2_xrows_floats(L) + z;
      sp = (float) (1.0 / (double) xb[1]);        // 1 / Backward's N at row 0
    }
    float accM[B], accI[B];
#pragma unroll
    for (int j = 0; j < B; j++) accM[j] = accI[j] = 0.0f;
    // optimal accuracy alignment (MODE 3): the OA rows (row 0: -infinity in
    // the cells; N 0, J -infinity, B 0, C -infinity) and where the traceback's
    // choices go
    constexpr int NW = (B + 7) / 8;
    float oaM[B], oaI[B], oaD[B];
    float oaN = 0.0f, oaJ = -INFINITY, oaB = 0.0f, oaC = -INFINITY;
    uint32_t *obits = nullptr, *orow = nullptr;
    uint8_t *opp = nullptr;
    if (MODE == 3) {
#pragma unroll
      for (int j = 0; j < B; j++) oaM[j] = oaI[j] = oaD[j] = -INFINITY;
      obits = (uint32_t *) (area + n2_oa_bits_at(L, B));
      orow  = (uint32_t *) (area + n2_oa_rows_at(L, B));
      opp   = (uint8_t *) (area + n2_pp_at(L, B));
      if (z == 0) orow[0] = 1u << OA_BBIT;        // B(0) comes from N(0)
    }
    if (pool) {
      unsigned long long a64 = 0;
      if (z == 0) a64 = atomicAdd(l.cursor, (unsigned long long) (L + 1));
      const int64_t at = ((int64_t) (unsigned) uni((int) (a64 >> 32)) << 32) | (unsigned) uni((int) a64);
      if (at + L + 1 <= pool_size) { scl = pool + at; if (z == 0) scale_off[slot] = at; }
      else { if (z == 0) { scale_off[slot] = -1; atomicExch(&l.counts[6], 1); } }
    }
    // the special-state rows go to the caller's matrix (one target) or, for
    // domain decoding, to the target's entry in the decoding pool
    float *xo = xmx_out;
    if (DOM) {
      const int64_t need = 6 * (int64_t) (L + 1);
      unsigned long long a64 = 0;
      if (z == 0) a64 = atomicAdd(l.cursor + 1, (unsigned long long) need);
      const int64_t at = ((int64_t) (unsigned) uni((int) (a64 >> 32)) << 32) | (unsigned) uni((int) a64);
      if (at + need <= dom.size) { xo = dom.pool + at; if (z == 0) dom.doff[slot] = at; }
      else { xo = nullptr; if (z == 0) { dom.doff[slot] = -1; atomicOr(&l.counts[7], 1); } }
    }
    float M[B], I[B], D[B];
#pragma unroll
    for (int j = 0; j < B; j++) M[j] = I[j] = D[j] = 0.0f;
    float xN = 1.0f, xJ = 0.0f, xC = 0.0f, xB = fmove, xE = 0.0f;
    float totscale = 0.0f;
    if (xo && z == 0) {
      xo[0] = 0.0f; xo[1] = 1.0f; xo[2] = 0.0f; xo[3] = xB; xo[4] = 0.0f; xo[5] = 1.0f;
    }
    if (scl && z == 0) scl[0] = 1.0f;

    float mL = 0.0f;                            // M(i-1, k-1) of the lane's first position
    ResAsc rd;
    if (L > 0) rd.init(sq);
    for (int i = 1; i <= L; i++) {
      const int x = rd.at(i-1);
      const float *e = tbl + (size_t) x * rowf + z;
      const float *trr = tr + (PIN ? 0 : opaque_zero());
      const float pxN = xN, pxJ = xJ, pxC = xC;          // row i-1
      // null2 (MODE 3): Backward's specials and, for up to eight positions
      // per lane, its M and I values of row i, loaded before the row's
      // Forward part so that their latency overlaps it
      constexpr bool PRE = (MODE == 3 && B <= 8);
      float bx1 = 0.0f, bx2 = 0.0f, bx4 = 0.0f, bx5 = 1.0f;
      float pbM[PRE ? B : 1], pbI[PRE ? B : 1];
      if (MODE == 3) {
        const float *xr = xb + (size_t) i * 6;
        bx1 = xr[1]; bx2 = xr[2]; bx4 = xr[4]; bx5 = xr[5];
        if (PRE) {
          const float *cr = cells + (size_t) (i-1) * 2 * rowf;
#pragma unroll
          for (int j = 0; j < (PRE ? B : 1); j++) { pbM[j] = cr[j*64]; pbI[j] = cr[rowf + j*64]; }
        }
      }

#define FTR(A, T, j) (PIN ? A[j] : trr[(T) * rowf + (j) * 64])
      const float iL = dppz<WAVE_SHR1>(I[B-1]), dL = dppz<WAVE_SHR1>(D[B-1]);
      float xEl = 0.0f;
#pragma unroll
      for (int j = B-1; j >= 0; j--) {
        const float mp = j ? M[j-1] : mL, ip = j ? I[j-1] : iL, dp = j ? D[j-1] : dL;
        const float In = M[j] * FTR(tMI, VT_MI, j) + I[j] * FTR(tII, VT_II, j);
        float sv = xB * FTR(tBM, VT_BM, j) + mp * FTR(tMM, VT_MM, j) + ip * FTR(tIM, VT_IM, j) + dp * FTR(tDM, VT_DM, j);
        sv = sv * e[j * 64];
        M[j] = sv; I[j] = In;
        xEl += sv;
      }
      const float mS = dppz<WAVE_SHR1>(M[B-1]);
#pragma unroll
      for (int j = 0; j < B; j++) D[j] = (j ? M[j-1] : mS) * FTR(tMDS, VT_MDS, j);
#undef FTR
      {
        // the D->D odds are read per row, not pinned in registers: this keeps
        // the kernel at 128 VGPRs, four waves per SIMD
        float tDDS[B];
#pragma unroll
        for (int j = 0; j < B; j++) tDDS[j] = trr[VT_DDS*rowf + j*64];
        affine_fwd_scan<B>(D, tDDS, akc, z);
      }
#pragma unroll
      for (int j = 0; j < B; j++) xEl += D[j];
      xE = unif(wave_sum_f(xEl));
      xN = xN * floop;
      xC = (xC * floop) + (xE * eMove);
      xJ = (xJ * floop) + (xE * eLoop);
      xB = (xJ * fmove) + (xN * fmove);
      float scale = 1.0f;
      if (xE > 1.0e4f) {
        xN = xN / xE; xC = xC / xE; xJ = xJ / xE; xB = xB / xE;
        const float inv = (float) (1.0 / (double) xE);
#pragma unroll
        for (int j = 0; j < B; j++) { M[j] *= inv; D[j] *= inv; I[j] *= inv; }
        scale = xE;
        totscale = (float) ((double) totscale + log((double) xE));
        xE = 1.0f;
      }
      mL = dppz<WAVE_SHR1>(M[B-1]);
      if (scl && z == 0) scl[i] = scale;
      if (xo && z == 0) {
        float *xr = xo + (size_t) i * 6;
        xr[0] = xE; xr[1] = xN; xr[2] = xJ; xr[3] = xB; xr[4] = xC; xr[5] = scale;
      }
      if (MODE == 3) {
        // posteriors of row i: (f * b) * (scaleproduct * scale_i); N, J, C
        // from Forward's row i-1 and Backward's row i
        const float totr = sp * scale;
        const float *cr = cells + (size_t) (i-1) * 2 * rowf;
        const float ppN = __fmul_rn(__fmul_rn(__fmul_rn(pxN, bx1), floop), sp);
        const float ppJ = __fmul_rn(__fmul_rn(__fmul_rn(pxJ, bx2), floop), sp);
        const float ppC = __fmul_rn(__fmul_rn(__fmul_rn(pxC, bx4), floop), sp);
        pN = __fadd_rn(pN, ppN);
        pJ = __fadd_rn(pJ, ppJ);
        pC = __fadd_rn(pC, ppC);
        if (own) sp = __fmul_rn(sp, __fdiv_rn(scale, bx5));
        // the row's posteriors (summed for null2) and its optimal accuracy row
        oa_row<B, NW, PIN, PRE>(p, z, i, tBM, tMM, tIM, tDM, tMDS, tMI, tII, tr + (PIN ? 0 : opaque_zero()), rowf,
                                M, I, cr, pbM, pbI, totr, accM, accI, ppN, ppJ, ppC, floop, fmove, oaM, oaI, oaD,
                                oaN, oaJ, oaB, oaC, obits, orow, opp);
      }
    }
    if (MODE == 3) {
      if (z == 0) dom.n2oa[slot] = oaC;          // p7_OptimalAccuracy's *ret_e = C(L)
      null2_epilogue<B>(p, tbl, rowf, z, accM, accI, pN, pJ, pC, sp, sq, L, dom.n2dc + slot, dom.n2st + slot);
      continue;
    }
    if (z == 0) {
      int status = 0;
      float sc;
      if (isnan(xC) || (L > 0 && xC == 0.0f) || isinf(xC)) { status = 16; sc = NAN; }
      else sc = (float) ((double) totscale + log((double) (xC * fmove)));
      r.fwdsc[slot] = sc;
      r.fstatus[slot] = status;
      bool pass = false;
      if (gating) {
        const float seq = (float) __ddiv_rn((double) (sc - r.filtersc[slot]), LOG2_D);
        pass = (status == 0 && seq >= g.T3f);
        r.stage[slot] = pass ? 4 : 3;
      }
      // an F3 survivor's decoding entry: the Backward rows, then btot, etot, mocc
      if (DOM) {
        int64_t bo = -1;
        if (pass && xo) {
          const int64_t need = 9 * (int64_t) (L + 1);
          const int64_t at = (int64_t) atomicAdd(l.cursor + 1, (unsigned long long) need);
          if (at + need <= dom.size) bo = at;
          else atomicOr(&l.counts[7], 1);
        }
        dom.boff[slot] = bo;
      }
    }
  }
}

// Forward parser for large models, W waves per target. Each lane's B table
// positions are split into W slices of B/W; wave w owns slice w of every lane,
// with its transitions pinned in registers (one wave alone would have to read
// them from memory every row). Along the chain (lane by lane, slice by slice)
// a slice's first position takes its predecessor from the previous slice
// through LDS. The D->D chain is composed per slice (constants of the row),
// per lane (the W slices in order), then scanned across the lanes; the row sum
// is the sum of the waves' parts. Four barriers per row order the exchanges.
template <int B>
__device__ __forceinline__ float lane_carry_fwd(float C, const AffConst &k, int z)
{
  // C: the lane's map constant (carry-in 0); returns D at the end of lane z-1
  C = C + k.w1 * dppz<ROW_SHR1>(C);
  C = C + k.w2 * dppz<ROW_SHR2>(C);
  C = C + k.w4 * dppz<ROW_SHR4>(C);
  C = C + k.w8 * dppz<ROW_SHR8>(C);
  {
    const float c1 = readlanef(C, 15);
    const float c2 = k.pa * c1 + readlanef(C, 31);
    const float c3 = k.pb * c2 + readlanef(C, 47);
    const int row = z >> 4;
    const float rc = (row == 0) ? 0.0f : (row == 1) ? c1 : (row == 2) ? c2 : c3;
    C = C + k.prow * rc;
  }
  return dppz<WAVE_SHR1>(C);
}

__device__ __forceinline__ void aff_setup_lane(AffConst &k, float P)
{
  k.w1 = P; P = P * dppf<ROW_SHR1>(1.0f, P);
  k.w2 = P; P = P * dppf<ROW_SHR2>(1.0f, P);
  k.w4 = P; P = P * dppf<ROW_SHR4>(1.0f, P);
  k.w8 = P; P = P * dppf<ROW_SHR8>(1.0f, P);
  k.prow = P; k.pa = readlanef(P, 31); k.pb = readlanef(P, 47);
}

template <int B, int W>
__global__ __launch_bounds__(64 * W)
void k_fwd_wg(const DevProfile p, const uint8_t *__restrict__ res, const int64_t *__restrict__ off,
              const int32_t *__restrict__ len, const int32_t *__restrict__ list, const int32_t *__restrict__ count,
              DevResults r, DevLists l, GateThresh g, int gating, LenCfg lc,
              float *__restrict__ pool, int64_t pool_size, int64_t *__restrict__ scale_off)
{
  static_assert(B % W == 0, "slices of equal width");
  constexpr int BW = B / W;
  __shared__ float sh_pm[W][64], sh_pi[W][64], sh_pd[W][64];  // previous row, each slice's last position
  __shared__ float sh_m[W][64];                               // this row's M at each slice's last position
  __shared__ float sh_c[W][64];                               // this row's slice D-chain constants
  __shared__ float sh_e[W];                                   // the waves' parts of the row sum
  __shared__ int64_t sh_t, sh_at;
  if ((int64_t) blockIdx.x >= *count) return;
  const int wv = threadIdx.x >> 6, z = threadIdx.x & 63;
  const int j0 = wv * BW;
  const size_t rowf = (size_t) 64 * B;
  const float *tbl = p.fb_tbl;
  const float *tr = tbl + (size_t) p.nrows * rowf + z;
  float tBM[BW], tMM[BW], tIM[BW], tDM[BW], tMDS[BW], tMI[BW], tII[BW], tDD[BW];
#pragma unroll
  for (int j = 0; j < BW; j++) {
    const size_t c = (size_t) (j0 + j) * 64;
    tBM[j] = tr[VT_BM*rowf + c]; tMM[j] = tr[VT_MM*rowf + c]; tIM[j] = tr[VT_IM*rowf + c]; tDM[j] = tr[VT_DM*rowf + c];
    tMDS[j] = tr[VT_MDS*rowf + c]; tMI[j] = tr[VT_MI*rowf + c]; tII[j] = tr[VT_II*rowf + c]; tDD[j] = tr[VT_DDS*rowf + c];
  }
  // the D->D products of every slice of this lane, and of the whole lane
  float Pv[W];
#pragma unroll
  for (int v = 0; v < W; v++) {
    float P = 1.0f;
    for (int j = 0; j < BW; j++) P *= tr[VT_DDS*rowf + (size_t) (v * BW + j) * 64];
    Pv[v] = P;
  }
  AffConst akl;
  {
    float P = 1.0f;
#pragma unroll
    for (int v = 0; v < W; v++) P *= Pv[v];
    aff_setup_lane(akl, P);
  }
  const int64_t n = *count;
  int32_t *queue = (int32_t *) count + P7G_QOFF;

  for (;;) {
    if (threadIdx.x == 0) sh_t = (int64_t) atomicAdd(queue, 1);
    __syncthreads();
    const int64_t t = sh_t;
    if (t >= n) break;
    const int slot = list[t];
    const int L = len[slot];
    const uint8_t *sq = res + off[slot];
    const float floop = lc.on ? lc.floop : p.xf_loop[L], fmove = lc.on ? lc.fmove : p.xf_move[L];
    if (threadIdx.x == 0) {
      int64_t at = -1;
      if (pool) {
        at = (int64_t) atomicAdd(l.cursor, (unsigned long long) (L + 1));
        if (at + L + 1 > pool_size) { at = -1; atomicExch(&l.counts[6], 1); }
        scale_off[slot] = at;
      }
      sh_at = at;
    }
    sh_pm[wv][z] = 0.0f; sh_pi[wv][z] = 0.0f; sh_pd[wv][z] = 0.0f;
    __syncthreads();
    float *scl = (pool && sh_at >= 0) ? pool + sh_at : nullptr;
    float M[BW], I[BW], D[BW];
#pragma unroll
    for (int j = 0; j < BW; j++) M[j] = I[j] = D[j] = 0.0f;
    float xN = 1.0f, xJ = 0.0f, xC = 0.0f, xB = fmove, xE = 0.0f;
    float totscale = 0.0f;
    if (scl && threadIdx.x == 0) scl[0] = 1.0f;
    ResAsc rd;
    if (L > 0) rd.init(sq);
    for (int i = 1; i <= L; i++) {
      const int x = rd.at(i-1);
      const float *e = tbl + (size_t) x * rowf + z;
      __syncthreads();                                  // the previous row's slice ends are written
      float mL = 0.0f, iL = 0.0f, dL = 0.0f;
      if (wv > 0)     { mL = sh_pm[wv-1][z];  iL = sh_pi[wv-1][z];  dL = sh_pd[wv-1][z]; }
      else if (z > 0) { mL = sh_pm[W-1][z-1]; iL = sh_pi[W-1][z-1]; dL = sh_pd[W-1][z-1]; }
      float xEl = 0.0f;
#pragma unroll
      for (int j = BW-1; j >= 0; j--) {
        const float mp = j ? M[j-1] : mL, ip = j ? I[j-1] : iL, dp = j ? D[j-1] : dL;
        const float In = M[j] * tMI[j] + I[j] * tII[j];
        float sv = xB * tBM[j] + mp * tMM[j] + ip * tIM[j] + dp * tDM[j];
        sv = sv * e[(size_t) (j0 + j) * 64];
        M[j] = sv; I[j] = In;
        xEl += sv;
      }
      sh_m[wv][z] = M[BW-1];
      __syncthreads();                                  // this row's slice-end M values are written
      const float mS = (wv > 0) ? sh_m[wv-1][z] : (z > 0) ? sh_m[W-1][z-1] : 0.0f;
#pragma unroll
      for (int j = 0; j < BW; j++) D[j] = (j ? M[j-1] : mS) * tMDS[j];
      {
        float C = 0.0f;
#pragma unroll
        for (int j = 0; j < BW; j++) C = D[j] + tDD[j] * C;
        sh_c[wv][z] = C;
      }
      __syncthreads();                                  // the slices' chain constants are written
      {
        float Cl = 0.0f;
#pragma unroll
        for (int v = 0; v < W; v++) Cl = sh_c[v][z] + Pv[v] * Cl;
        float c = lane_carry_fwd<B>(Cl, akl, z);
        for (int v = 0; v < wv; v++) c = sh_c[v][z] + Pv[v] * c;
        float cur = c;
#pragma unroll
        for (int j = 0; j < BW; j++) { cur = D[j] + tDD[j] * cur; D[j] = cur; }
      }
#pragma unroll
      for (int j = 0; j < BW; j++) xEl += D[j];
      {
        const float ws = wave_sum_f(xEl);
        if (z == 0) sh_e[wv] = ws;
      }
      __syncthreads();                                  // the waves' parts of the row sum are written
      xE = 0.0f;
#pragma unroll
      for (int v = 0; v < W; v++) xE += sh_e[v];
      xN = xN * floop;
      xC = (xC * floop) + (xE * p.xfE_move);
      xJ = (xJ * floop) + (xE * p.xfE_loop);
      xB = (xJ * fmove) + (xN * fmove);
      float scale = 1.0f;
      if (xE > 1.0e4f) {
        xN = xN / xE; xC = xC / xE; xJ = xJ / xE; xB = xB / xE;
        const float inv = (float) (1.0 / (double) xE);
#pragma unroll
        for (int j = 0; j < BW; j++) { M[j] *= inv; D[j] *= inv; I[j] *= inv; }
        scale = xE;
        totscale = (float) ((double) totscale + log((double) xE));
        xE = 1.0f;
      }
      if (scl && threadIdx.x == 0) scl[i] = scale;
      sh_pm[wv][z] = M[BW-1]; sh_pi[wv][z] = I[BW-1]; sh_pd[wv][z] = D[BW-1];
    }
    if (threadIdx.x == 0) {
      int status = 0;
      float sc;
      if (isnan(xC) || (L > 0 && xC == 0.0f) || isinf(xC)) { status = 16; sc = NAN; }
      else sc = (float) ((double) totscale + log((double) (xC * fmove)));
      r.fwdsc[slot] = sc;
      r.fstatus[slot] = status;
      bool pass = false;
      if (gating) {
        const float seq = (float) __ddiv_rn((double) (sc - r.filtersc[slot]), LOG2_D);
        pass = (status == 0 && seq >= g.T3f);
        r.stage[slot] = pass ? 4 : 3;
      }
    }
    __syncthreads();                                    // the target is done before the next one is taken
  }
}

/*****************************************************************
 * 8. Backward parser (one target per wave), with Forward's scale factors
 *****************************************************************/
// DOM: the rows also go to the domain decoding pool (a separate instance, so
// that the plain parser keeps its register budget)
template <int B, int TM, int MODE>
__global__ __launch_bounds__(BCK_THREADS(B))
void k_bck(const DevProfile p, const uint8_t *__restrict__ res, const int64_t *__restrict__ off,
           const int32_t *__restrict__ len, const int32_t *__restrict__ list, const int32_t *__restrict__ count,
           DevResults r, const float *__restrict__ pool, const int64_t *__restrict__ scale_off,
           const float *__restrict__ fwd_xmx, float *__restrict__ xmx_out, LenCfg lc, DomBufs dom, DevLists l)
{
  constexpr bool PIN = (TM == TM_LDS || TM == TM_PIN);
  if ((int64_t) blockIdx.x * (blockDim.x >> 6) >= *count) return;
  extern __shared__ float ldsf[];
  const size_t rowf = (size_t) 64 * B;
  const float *tbl = p.fb_tbl;                           // emission rows
  if (TM == TM_LDS) { stage_table(ldsf, p.fb_tbl, p.fb_floats); tbl = ldsf; }
  if (TM == TM_TLDS) stage_table(ldsf, p.fb_tbl + (size_t) p.nrows * rowf, (size_t) P7G_FB_NT * rowf);
  const int z = threadIdx.x & 63;
  const float *tr = ((TM == TM_TLDS) ? ldsf : tbl + (size_t) p.nrows * rowf) + z;
  float tBM[B], tMMN[B], tIMN[B], tDMN[B], tMD[B], tDD[B], tMI[B], tII[B];
  if (PIN) {
#pragma unroll
    for (int j = 0; j < B; j++) {
      tBM[j] = tr[VT_BM*rowf + j*64]; tMMN[j] = tr[FT_MMN*rowf + j*64]; tIMN[j] = tr[FT_IMN*rowf + j*64]; tDMN[j] = tr[FT_DMN*rowf + j*64];
      tMD[j] = tr[FT_MD*rowf + j*64]; tDD[j] = tr[FT_DD*rowf + j*64]; tMI[j] = tr[VT_MI*rowf + j*64]; tII[j] = tr[VT_II*rowf + j*64];
    }
  }
  AffConst akc;
  {
    float tdd[B];
#pragma unroll
    for (int j = 0; j < B; j++) tdd[j] = tr[FT_DD*rowf + j*64];
    aff_bck_setup<B>(akc, tdd);
  }
#define BTR(A, T, j) (PIN ? A[j] : trr[(T) * rowf + (j) * 64])
  constexpr bool DOM = (MODE == 1), ENV = (MODE == 2);
  const int64_t n = ENV ? min((int64_t) *count, dom.env_cap) : (int64_t) *count;
  const float eMove = ENV ? 1.0f : p.xfE_move, eLoop = ENV ? 0.0f : p.xfE_loop;

  int32_t *queue = (int32_t *) count + P7G_QOFF;
  for (int64_t t = queue_next(queue, 1); t < n; t = queue_next(queue, 1)) {
    const int slot = uni(list[t]);
    const int L = uni(len[slot]);
    const uint8_t *sq = res + off[slot];
    const float floop = ENV ? dom.env_loop[slot] : lc.on ? lc.floop : p.xf_loop[L];
    const float fmove = ENV ? dom.env_move[slot] : lc.on ? lc.fmove : p.xf_move[L];
    const float *fsc = nullptr;       // forward row scale factors, fsc[i], i = 0..L
    int fstride = 1;
    float *cells = nullptr;           // envelope Backward for null2: its M and I rows
    float *xo = xmx_out;
    if (ENV) {
      const int64_t so = dom.n2soff[slot];
      if (so < 0) {                   // the envelope-score pass found no room for its scales
        if (z == 0) { r.bcksc[slot] = NAN; r.fstatus[slot] = 5; }
        continue;
      }
      fsc = dom.n2scl + (((int64_t) (unsigned) uni((int) (so >> 32)) << 32) | (unsigned) uni((int) so));
      const int64_t o = dom.n2off[slot];
      float *area = dom.n2pool + (((int64_t) (unsigned) uni((int) (o >> 32)) << 32) | (unsigned) uni((int) o));
      xo = area;
      cells = xo + n2_xrows_floats(L) + z;
    } else if (fwd_xmx) { fsc = fwd_xmx + 5; fstride = 6; }
    else {
      const int64_t so = scale_off[slot];
      if (so < 0) {
        if (z == 0) { r.bcksc[slot] = NAN; r.fstatus[slot] = 5; }
        continue;
      }
      fsc = pool + so;
    }
    // the special-state rows: the caller's matrix (one target), or the
    // target's decoding entry that the Forward parser reserved
    if (DOM) {
      const int64_t bo = dom.boff[slot];
      const int64_t bu = ((int64_t) (unsigned) uni((int) (bo >> 32)) << 32) | (unsigned) uni((int) bo);
      xo = (bu >= 0) ? dom.pool + bu : nullptr;
    }
    float M[B], I[B], D[B];
    float xJ = 0.0f, xB = 0.0f, xN = 0.0f;
    float xC = fmove;
    float xE = xC * eMove;
#pragma unroll
    for (int j = 0; j < B; j++) { M[j] = xE; D[j] = xE; I[j] = 0.0f; }
    {
      const float *trr = tr + (PIN ? 0 : opaque_zero());
      affine_bck_scan<B>(D, [&](int j) { return BTR(tDD, FT_DD, j); }, akc, z);
      const float dS = wave_shl1f(D[0]);
#pragma unroll
      for (int j = 0; j < B; j++) M[j] = M[j] + (j < B-1 ? D[j+1] : dS) * BTR(tMD, FT_MD, j);
    }
    int own = 0;
    float totscale;
    {
      const float s = fsc[(size_t) L * fstride];
      if (s > 1.0f) {
        xE = xE / s; xN = xN / s; xC = xC / s; xJ = xJ / s; xB = xB / s;
        const float inv = (float) (1.0 / (double) s);
#pragma unroll
        for (int j = 0; j < B; j++) { M[j] *= inv; D[j] *= inv; I[j] *= inv; }
      }
      totscale = (float) log((double) s);
      if (xo && z == 0) {
        float *xr = xo + (size_t) L * 6;
        xr[0] = xE; xr[1] = xN; xr[2] = xJ; xr[3] = xB; xr[4] = xC; xr[5] = s;
      }
      if (ENV && L > 0) {
        float *cr = cells + (size_t) (L-1) * 2 * rowf;
#pragma unroll
        for (int j = 0; j < B; j++) { cr[j*64] = M[j]; cr[rowf + j*64] = I[j]; }
      }
    }
    ResDesc rd;
    if (L > 0) rd.init(sq, L);
    for (int i = L-1; i >= 1; i--) {
      const int x = rd.at(i);                         // residue i+1 (1-based)
      const float *e = tbl + (size_t) x * rowf + z;
      const float *trr = tr + (PIN ? 0 : opaque_zero());
      // mn(j) = M(j) * e(j), formed one position ahead of the update that
      // overwrites M(j); xBl sums mn(j) * tBM(j) in increasing j
      float mn = M[0] * e[0];
      float xBl = 0.0f;
      xBl += mn * BTR(tBM, VT_BM, 0);
      const float mS = wave_shl1f(mn);
#pragma unroll
      for (int j = 0; j < B; j++) {
        float ms = mS;
        if (j < B-1) { ms = M[j+1] * e[(j+1) * 64]; xBl += ms * BTR(tBM, VT_BM, j+1); }
        const float ip = I[j];
        I[j] = ip * BTR(tII, VT_II, j) + ms * BTR(tIMN, FT_IMN, j);
        D[j] = ms * BTR(tDMN, FT_DMN, j);
        M[j] = ip * BTR(tMI, VT_MI, j) + ms * BTR(tMMN, FT_MMN, j);
      }
      xB = unif(wave_sum_f(xBl));
      xC = xC * floop;
      xJ = (xB * fmove) + (xJ * floop);
      xN = (xB * fmove) + (xN * floop);
      xE = (xC * eMove) + (xJ * eLoop);
#pragma unroll
      for (int j = 0; j < B; j++) D[j] = D[j] + xE;
      affine_bck_scan<B>(D, [&](int j) { return BTR(tDD, FT_DD, j); }, akc, z);
      const float dS = wave_shl1f(D[0]);
#pragma unroll
      for (int j = 0; j < B; j++) M[j] = M[j] + xE + (j < B-1 ? D[j+1] : dS) * BTR(tMD, FT_MD, j);
      if (xB > 1.0e16f) own = 1;
      const float s = own ? ((xB > 1.0e4f) ? xB : 1.0f) : fsc[(size_t) i * fstride];
      if (s > 1.0f) {
        xE /= s; xN /= s; xJ /= s; xB /= s; xC /= s;
        const float inv = (float) (1.0 / (double) s);
#pragma unroll
        for (int j = 0; j < B; j++) { M[j] *= inv; D[j] *= inv; I[j] *= inv; }
        totscale = (float) ((double) totscale + log((double) s));
      }
      if (xo && z == 0) {
        float *xr = xo + (size_t) i * 6;
        xr[0] = xE; xr[1] = xN; xr[2] = xJ; xr[3] = xB; xr[4] = xC; xr[5] = s;
      }
      if (ENV) {
        float *cr = cells + (size_t) (i-1) * 2 * rowf;
#pragma unroll
        for (int j = 0; j < B; j++) { cr[j*64] = M[j]; cr[rowf + j*64] = I[j]; }
      }
    }
    {
      const int x = (L > 0) ? rd.at(0) : P7G_PADCODE;
      const float *e = tbl + (size_t) x * rowf + z;
      const float *trr = tr + (PIN ? 0 : opaque_zero());
      float xBl = 0.0f;
#pragma unroll
      for (int j = 0; j < B; j++) xBl += (M[j] * e[j * 64]) * BTR(tBM, VT_BM, j);
      xB = unif(wave_sum_f(xBl));
      xN = (xB * fmove) + (xN * floop);
    }
    if (z == 0) {
      if (xo) {
        xo[0] = 0.0f; xo[1] = xN; xo[2] = 0.0f; xo[3] = xB; xo[4] = 0.0f; xo[5] = 1.0f;
      }
      int status = 0;
      float sc;
      if (isnan(xN) || (L > 0 && xN == 0.0f) || isinf(xN)) { status = 16; sc = NAN; }
      else sc = (float) ((double) totscale + log((double) xN));
      r.bcksc[slot] = sc;
      r.fstatus[slot] = status | (own ? 0x100 : 0);
    }
  }
}
#undef BTR

/*****************************************************************
 * 9. Domain decoding and regions (F3 survivors, one target per wave)
 *****************************************************************/
// p7_DomainDecoding (HMMER/src/impl_sse/decoding.c:166-197) from the two
// parsers' special rows: btot[i] / etot[i], the expected B / E visits up to
// row i, and mocc[i], the probability that residue i is emitted by the core
// model. Each row's terms are the reference's float products; the running
// sums over rows are taken 64 rows at a time by a wave scan (the reference
// adds them one by one, so the sums may differ in the last bits). When the
// Backward parser used its own scale factors the scale product changes from
// row to row, and the first lane walks the rows in the reference's order.
// Then the region loop of p7_domaindef_ByPosteriorHeuristics
// (p7_domaindef.c:409-489): its three occupancy tests are evaluated for 64
// rows at once into ballots, and the loop's state machine steps through
// their bits; a region's is_multidomain_region test (:534-547) is a wave
// maximum over its split points. Thresholds rt1/rt2/rt3 of
// p7_domaindef_Create (:109-111). Regions are appended as {slot, i, j,
// multidomain} records, each with a 16-byte aligned copy of its residues and
// the unihit N/J/C odds of its target's length, for the envelope scores that
// k_fwd (MODE 2) computes next.
#define DOM_RT1 0.25f
#define DOM_RT2 0.10f
#define DOM_RT3 0.20f
__global__ __launch_bounds__(256)
void k_domains(const DevProfile p, const uint8_t *__restrict__ res, const int64_t *__restrict__ off,
               const int32_t *__restrict__ len, const int32_t *__restrict__ list,
               const int32_t *__restrict__ count, const int *__restrict__ fstatus, DevLists l, LenCfg lc,
               DomBufs dom, DomOut out)
{
#pragma clang fp contract(off)
  enum { E = 0, N = 1, J = 2, Bx = 3, C = 4, S = 5 };
  const int lane = threadIdx.x & 63;
  const int64_t n = *count;
  for (int64_t t = wave_id(); t < n; t += wave_cnt()) {
    const int slot = uni(list[t]);
    const int L = uni(len[slot]);
    const int64_t fo = dom.doff[slot], bo = dom.boff[slot];
    if (fo < 0 || bo < 0) {
      if (lane == 0) { out.nexp[slot] = NAN; out.nreg[slot] = -1; out.nclu[slot] = -1; }
      continue;
    }
    const float *fx = dom.pool + fo;
    const float *bx = dom.pool + bo;
    float *btot = dom.pool + bo + 6 * (int64_t) (L + 1), *etot = btot + (L + 1), *mocc = etot + (L + 1);
    const bool own = (fstatus[slot] & 0x100) != 0;
    const float loop = lc.on ? lc.floop : p.xf_loop[L];     // N, J and C loops are equal after ReconfigLength
    // N/J/C odds of p7_oprofile_ReconfigUnihit(om, L) (impl_sse/p7_oprofile.c:1176-1186,
    // 1113-1114: nj = 0) for the envelope scores of this target's regions
    const float umove = __fdiv_rn(2.0f + 0.0f, (float) L + 2.0f + 0.0f), uloop = 1.0f - umove;
    const float sp0 = __fdiv_rn(1.0f, bx[N]);
    if (lane == 0) { btot[0] = 0.0f; etot[0] = 0.0f; mocc[0] = 0.0f; }
    if (! own) {
      float cb = 0.0f, ce = 0.0f;                           // btot, etot at the row before the block
      for (int base = 0; base < L; base += 64) {
        const int i = base + lane + 1;
        float db = 0.0f, de = 0.0f, mo = 0.0f;
        if (i <= L) {
          const float *f0 = fx + (size_t) (i - 1) * 6, *f1 = fx + (size_t) i * 6;
          const float *b0 = bx + (size_t) (i - 1) * 6, *b1 = bx + (size_t) i * 6;
          db = f0[Bx] * b0[Bx] * f0[S] * sp0;
          de = f1[E] * b1[E] * f1[S] * sp0;
          float njcp = f0[N] * b1[N] * loop * sp0;
          njcp += f0[J] * b1[J] * loop * sp0;
          njcp += f0[C] * b1[C] * loop * sp0;
          mo = 1.0f - njcp;
        }
        const float bt = cb + wave_scan_f(db), et = ce + wave_scan_f(de);
        if (i <= L) { btot[i] = bt; etot[i] = et; mocc[i] = mo; }
        cb = readlanef(bt, 63); ce = readlanef(et, 63);
      }
    } else if (lane == 0) {
      float sp = sp0, bt = 0.0f, et = 0.0f;
      for (int i = 1; i <= L; i++) {
        const float *f0 = fx + (size_t) (i - 1) * 6, *f1 = fx + (size_t) i * 6;
        const float *b0 = bx + (size_t) (i - 1) * 6, *b1 = bx + (size_t) i * 6;
        bt = bt + (f0[Bx] * b0[Bx] * f0[S] * sp);
        sp *= __fdiv_rn(f0[S], b0[S]);
        et = et + (f1[E] * b1[E] * f1[S] * sp);
        float njcp = f0[N] * b1[N] * loop * sp;
        njcp += f0[J] * b1[J] * loop * sp;
        njcp += f0[C] * b1[C] * loop * sp;
        btot[i] = bt; etot[i] = et; mocc[i] = 1.0f - njcp;
      }
    }
    __threadfence_block();                                  // the rows above are read back by other lanes
    if (lane == 0) out.nexp[slot] = btot[L];

    int nreg = 0, nclu = 0, i0 = -1;
    bool triggered = false;
    for (int base = 0; base < L; base += 64) {
      const int j = base + lane + 1;
      bool cs = false, ct = false, ce = false;
      if (j <= L) {
        const float mo = mocc[j];
        cs = mo - (btot[j] - btot[j-1]) < DOM_RT2;
        ct = mo >= DOM_RT1;
        ce = mo - (etot[j] - etot[j-1]) < DOM_RT2;
      }
      const unsigned long long ms = __ballot(cs), mt = __ballot(ct), me = __ballot(ce);
      const int nr = min(64, L - base);
      for (int r = 0; r < nr; r++) {                       // the loop's steps, on wave-uniform values
        const int jj = base + r + 1;
        if (! triggered) {
          if      ((ms >> r) & 1ull) i0 = jj;
          else if (i0 == -1)         i0 = jj;
          if      ((mt >> r) & 1ull) triggered = true;
        } else if ((me >> r) & 1ull) {
          // is_multidomain_region(i0, jj): max over split points z of
          // min(etot[z] - etot[i0-1], btot[jj] - btot[z-1])
          const float e0 = etot[i0 - 1], bj = btot[jj];
          float mx = -1.0f;
          for (int z0 = i0; z0 <= jj; z0 += 64) {
            const int z = z0 + lane;
            float en = -1.0f;
            if (z <= jj) {
              const float a = etot[z] - e0, b = bj - btot[z - 1];
              en = (a < b) ? a : b;
            }
            mx = fmaxf(mx, en);
          }
          {
            float v = mx;
            v = fmaxf(v, __shfl_xor(v, 1, 64)); v = fmaxf(v, __shfl_xor(v, 2, 64)); v = fmaxf(v, __shfl_xor(v, 4, 64));
            v = fmaxf(v, __shfl_xor(v, 8, 64)); v = fmaxf(v, __shfl_xor(v, 16, 64)); v = fmaxf(v, __shfl_xor(v, 32, 64));
            mx = unif(v);
          }
          const int multi = (mx >= DOM_RT3) ? 1 : 0;
          // record the region, and copy its residues for the envelope score
          const int Ld = jj - i0 + 1;
          const int64_t need = (((int64_t) Ld + 15) & ~(int64_t) 15) + 16;
          int k = 0;
          unsigned long long ro64 = 0;
          if (lane == 0) {
            k = atomicAdd(&out.ecount[0], 1);
            if (multi) atomicAdd(out.nclu_total, 1);
            if (k < out.reg_cap) ro64 = atomicAdd(l.cursor + 2, (unsigned long long) need);
          }
          k = uni(k);
          const int64_t ro = ((int64_t) (unsigned) uni((int) (ro64 >> 32)) << 32) | (unsigned) uni((int) ro64);
          if (k >= out.reg_cap) { if (lane == 0) atomicOr(&l.counts[7], 2); }
          else {
            const bool fits = ro + need <= out.res_cap;
            if (fits) {
              const uint8_t *src = res + off[slot] + (i0 - 1);
              for (int64_t b = lane; b < need; b += 64) out.res[ro + b] = (b < Ld) ? src[b] : (uint8_t) P7G_PADCODE;
            }
            if (lane == 0) {
              out.regs[k] = make_int4(slot, i0, jj, multi);
              out.v_off[k] = fits ? ro : 0; out.v_len[k] = fits ? Ld : 0; out.v_list[k] = k;
              out.env_loop[k] = uloop; out.env_move[k] = umove;
              if (! fits) atomicOr(&l.counts[7], 4);
            }
          }
          nreg++; nclu += multi;
          i0 = -1; triggered = false;
        }
      }
    }
    if (lane == 0) { out.nreg[slot] = nreg; out.nclu[slot] = nclu; }
  }
}

// Backward parser for large models, W waves per target: the slices of
// k_fwd_wg, traversed the other way. A slice's last position takes its
// successor (the next slice's first position) through LDS; the D->D chain runs
// towards smaller k (slice constants, lanes composed from their last slice,
// scanned across lanes from lane 63 down).
__device__ __forceinline__ float lane_carry_bck(float C, const AffConst &k, int z)
{
  // C: the lane's map constant (carry-in 0 from the right); returns D at the
  // first position of lane z+1
  C = C + k.w1 * dppz<ROW_SHL1>(C);
  C = C + k.w2 * dppz<ROW_SHL2>(C);
  C = C + k.w4 * dppz<ROW_SHL4>(C);
  C = C + k.w8 * dppz<ROW_SHL8>(C);
  {
    const float c2 = readlanef(C, 48);
    const float c1 = k.pa * c2 + readlanef(C, 32);
    const float c0 = k.pb * c1 + readlanef(C, 16);
    const int row = z >> 4;
    const float rc = (row == 3) ? 0.0f : (row == 2) ? c2 : (row == 1) ? c1 : c0;
    C = C + k.prow * rc;
  }
  return dppz<WAVE_SHL1>(C);
}

__device__ __forceinline__ void aff_setup_lane_bck(AffConst &k, float P)
{
  k.w1 = P; P = P * dppf<ROW_SHL1>(1.0f, P);
  k.w2 = P; P = P * dppf<ROW_SHL2>(1.0f, P);
  k.w4 = P; P = P * dppf<ROW_SHL4>(1.0f, P);
  k.w8 = P; P = P * dppf<ROW_SHL8>(1.0f, P);
  k.prow = P; k.pa = readlanef(P, 32); k.pb = readlanef(P, 16);
}

template <int B, int W>
__global__ __launch_bounds__(64 * W)
void k_bck_wg(const DevProfile p, const uint8_t *__restrict__ res, const int64_t *__restrict__ off,
              const int32_t *__restrict__ len, const int32_t *__restrict__ list, const int32_t *__restrict__ count,
              DevResults r, const float *__restrict__ pool, const int64_t *__restrict__ scale_off, LenCfg lc)
{
  static_assert(B % W == 0, "slices of equal width");
  constexpr int BW = B / W;
  __shared__ float sh_mn[W][64];     // M * e at each slice's first position (row being formed)
  __shared__ float sh_c[W][64];      // the slices' D-chain constants
  __shared__ float sh_d0[W][64];     // D at each slice's first position after the chain
  __shared__ float sh_x[W];          // the waves' parts of the B sum
  __shared__ int64_t sh_t;
  if ((int64_t) blockIdx.x >= *count) return;
  const int wv = threadIdx.x >> 6, z = threadIdx.x & 63;
  const int j0 = wv * BW;
  const size_t rowf = (size_t) 64 * B;
  const float *tbl = p.fb_tbl;
  const float *tr = tbl + (size_t) p.nrows * rowf + z;
  float tBM[BW], tMMN[BW], tIMN[BW], tDMN[BW], tMD[BW], tDD[BW], tMI[BW], tII[BW];
#pragma unroll
  for (int j = 0; j < BW; j++) {
    const size_t c = (size_t) (j0 + j) * 64;
    tBM[j] = tr[VT_BM*rowf + c]; tMMN[j] = tr[FT_MMN*rowf + c]; tIMN[j] = tr[FT_IMN*rowf + c]; tDMN[j] = tr[FT_DMN*rowf + c];
    tMD[j] = tr[FT_MD*rowf + c]; tDD[j] = tr[FT_DD*rowf + c]; tMI[j] = tr[VT_MI*rowf + c]; tII[j] = tr[VT_II*rowf + c];
  }
  float Pv[W];
#pragma unroll
  for (int v = 0; v < W; v++) {
    float P = 1.0f;
    for (int j = BW-1; j >= 0; j--) P *= tr[FT_DD*rowf + (size_t) (v * BW + j) * 64];
    Pv[v] = P;
  }
  AffConst akl;
  {
    float P = 1.0f;
#pragma unroll
    for (int v = W-1; v >= 0; v--) P *= Pv[v];
    aff_setup_lane_bck(akl, P);
  }
  const int64_t n = *count;
  int32_t *queue = (int32_t *) count + P7G_QOFF;

  // the D->D chain of the row in D (towards smaller k), then M += (D of the
  // next position) * tMD; two barriers
  auto dchain_and_m = [&](float (&M)[BW], float (&D)[BW], float add) {
    {
      float C = 0.0f;
#pragma unroll
      for (int j = BW-1; j >= 0; j--) C = D[j] + tDD[j] * C;
      sh_c[wv][z] = C;
    }
    __syncthreads();
    {
      float Cl = 0.0f;
#pragma unroll
      for (int v = W-1; v >= 0; v--) Cl = sh_c[v][z] + Pv[v] * Cl;
      float c = lane_carry_bck(Cl, akl, z);
      for (int v = W-1; v > wv; v--) c = sh_c[v][z] + Pv[v] * c;
      float cur = c;
#pragma unroll
      for (int j = BW-1; j >= 0; j--) { cur = D[j] + tDD[j] * cur; D[j] = cur; }
      sh_d0[wv][z] = D[0];
    }
    __syncthreads();
    const float dS = (wv < W-1) ? sh_d0[wv+1][z] : (z < 63) ? sh_d0[0][z+1] : 0.0f;
#pragma unroll
    for (int j = 0; j < BW; j++) M[j] = M[j] + add + (j < BW-1 ? D[j+1] : dS) * tMD[j];
  };

  for (;;) {
    if (threadIdx.x == 0) sh_t = (int64_t) atomicAdd(queue, 1);
    __syncthreads();
    const int64_t t = sh_t;
    if (t >= n) break;
    const int slot = list[t];
    const int L = len[slot];
    const uint8_t *sq = res + off[slot];
    const float floop = lc.on ? lc.floop : p.xf_loop[L], fmove = lc.on ? lc.fmove : p.xf_move[L];
    const int64_t so = scale_off[slot];
    if (so < 0) {
      if (threadIdx.x == 0) { r.bcksc[slot] = NAN; r.fstatus[slot] = 5; }
      __syncthreads();
      continue;
    }
    const float *fsc = pool + so;
    float M[BW], I[BW], D[BW];
    float xJ = 0.0f, xB = 0.0f, xN = 0.0f;
    float xC = fmove;
    float xE = xC * p.xfE_move;
#pragma unroll
    for (int j = 0; j < BW; j++) { M[j] = xE; D[j] = xE; I[j] = 0.0f; }
    dchain_and_m(M, D, 0.0f);
    int own = 0;
    float totscale;
    {
      const float s = fsc[L];
      if (s > 1.0f) {
        xE = xE / s; xN = xN / s; xC = xC / s; xJ = xJ / s; xB = xB / s;
        const float inv = (float) (1.0 / (double) s);
#pragma unroll
        for (int j = 0; j < BW; j++) { M[j] *= inv; D[j] *= inv; I[j] *= inv; }
      }
      totscale = (float) log((double) s);
    }
    ResDesc rd;
    if (L > 0) rd.init(sq, L);
    for (int i = L-1; i >= 1; i--) {
      const int x = rd.at(i);                         // residue i+1 (1-based)
      const float *e = tbl + (size_t) x * rowf + z;
      // M(j) * e(j) of the slice's first position, for the previous slice's last one
      sh_mn[wv][z] = M[0] * e[(size_t) j0 * 64];
      __syncthreads();
      const float mS = (wv < W-1) ? sh_mn[wv+1][z] : (z < 63) ? sh_mn[0][z+1] : 0.0f;
      float mn = M[0] * e[(size_t) j0 * 64];
      float xBl = 0.0f;
      xBl += mn * tBM[0];
#pragma unroll
      for (int j = 0; j < BW; j++) {
        float ms = mS;
        if (j < BW-1) { ms = M[j+1] * e[(size_t) (j0 + j + 1) * 64]; xBl += ms * tBM[j+1]; }
        const float ip = I[j];
        I[j] = ip * tII[j] + ms * tIMN[j];
        D[j] = ms * tDMN[j];
        M[j] = ip * tMI[j] + ms * tMMN[j];
      }
      {
        const float ws = wave_sum_f(xBl);
        if (z == 0) sh_x[wv] = ws;
      }
      __syncthreads();
      xB = 0.0f;
#pragma unroll
      for (int v = 0; v < W; v++) xB += sh_x[v];
      xC = xC * floop;
      xJ = (xB * fmove) + (xJ * floop);
      xN = (xB * fmove) + (xN * floop);
      xE = (xC * p.xfE_move) + (xJ * p.xfE_loop);
#pragma unroll
      for (int j = 0; j < BW; j++) D[j] = D[j] + xE;
      dchain_and_m(M, D, xE);
      if (xB > 1.0e16f) own = 1;
      const float s = own ? ((xB > 1.0e4f) ? xB : 1.0f) : fsc[i];
      if (s > 1.0f) {
        xE /= s; xN /= s; xJ /= s; xB /= s; xC /= s;
        const float inv = (float) (1.0 / (double) s);
#pragma unroll
        for (int j = 0; j < BW; j++) { M[j] *= inv; D[j] *= inv; I[j] *= inv; }
        totscale = (float) ((double) totscale + log((double) s));
      }
    }
    {
      const int x = (L > 0) ? rd.at(0) : P7G_PADCODE;
      const float *e = tbl + (size_t) x * rowf + z;
      float xBl = 0.0f;
#pragma unroll
      for (int j = 0; j < BW; j++) xBl += (M[j] * e[(size_t) (j0 + j) * 64]) * tBM[j];
      const float ws = wave_sum_f(xBl);
      if (z == 0) sh_x[wv] = ws;
      __syncthreads();
      xB = 0.0f;
#pragma unroll
      for (int v = 0; v < W; v++) xB += sh_x[v];
      xN = (xB * fmove) + (xN * floop);
    }
    if (threadIdx.x == 0) {
      int status = 0;
      float sc;
      if (isnan(xN) || (L > 0 && xN == 0.0f) || isinf(xN)) { status = 16; sc = NAN; }
      else sc = (float) ((double) totscale + log((double) xN));
      r.bcksc[slot] = sc;
      r.fstatus[slot] = status | (own ? 0x100 : 0);
    }
    __syncthreads();                                    // the target is done before the next one is taken
  }
}

__global__ void k_fill32(uint32_t *__restrict__ p, uint32_t v, int64_t n)
{
  for (int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t) gridDim.x * blockDim.x) p[i] = v;
}

__global__ void k_copy32(uint32_t *__restrict__ d, const uint32_t *__restrict__ s, int64_t n)
{
  for (int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t) gridDim.x * blockDim.x) d[i] = s[i];
}

/*****************************************************************
 * Launchers
 *****************************************************************/
hipError_t launch_fill32(void *p, uint32_t v, int64_t nwords, hipStream_t st)
{
  if (nwords <= 0) return hipSuccess;
  const unsigned grid = (unsigned) std::min<int64_t>((nwords + 255) / 256, 4096);
  hipLaunchKernelGGL(k_fill32, dim3(grid), dim3(256), 0, st, (uint32_t *) p, v, nwords);
  return hipGetLastError();
}

hipError_t launch_copy32(void *dst, const void *src, int64_t nwords, hipStream_t st)
{
  if (nwords <= 0) return hipSuccess;
  const unsigned grid = (unsigned) std::min<int64_t>((nwords + 255) / 256, 4096);
  hipLaunchKernelGGL(k_copy32, dim3(grid), dim3(256), 0, st, (uint32_t *) dst, (const uint32_t *) src, nwords);
  return hipGetLastError();
}

// p7_OATrace (impl_sse/optacc.c:227-276) over the choices the null2 pass
// stored (oa_row), one wave per region, from C at the region's last residue
// back to S; keeps the first and last M states (p7_alidisplay_Create's
// hmmfrom/hmmto/sqfrom/sqto, p7_alidisplay.c:74-145), and writes the
// alignment itself for the display, from its last M state back to its first
// (p7_alidisplay_Create's z2 down to z1), one word per state: state (0 M,
// 1 I, 2 D) in bits 0-1, model position in bits 2-18, the cell's posterior
// probability code in bits 19-22, the residue code in bits 23-27; at most
// n2pcap words at n2path + n2poff[region], their number to n2pn[region]
// (0 without an alignment).
//
// The walk is one thread's, kept uniform over the wave; the wave's lanes
// speculate on runs: in M, lane l loads the choice at (i-l, k-l), and the
// leading lanes whose choice is M -> M take their steps at once, each writing
// its state's word; in C, lane l reads row i-l's C -> C bit likewise. A run's
// lanes stop wherever the step-by-step walk would fail (bounds, the word
// cap, the step guard), so that the walk itself meets every failure.
__global__ void __launch_bounds__(64) k_oa_trace(const DevProfile p, const int32_t *__restrict__ list, int64_t n,
                           const int32_t *__restrict__ vlen, const uint8_t *__restrict__ vres,
                           const int64_t *__restrict__ voff, DomBufs dom)
{
  enum { S_S = 0, S_N, S_B, S_M, S_I, S_D, S_E, S_J, S_C };
  const int64_t t = (int64_t) blockIdx.x;
  const int lane = (int) threadIdx.x;
  if (t >= n) return;
  const int slot = list[t];
  int4 out = make_int4(0, 0, 0, 0);
  if (dom.n2st[slot] != 0) {                           // no decoding, no trace
    if (lane == 0) { dom.n2ali[slot] = out; dom.n2pn[slot] = 0; }
    return;
  }
  const int L = vlen[slot], B = p.fb_B, NW = n2_oa_words(B);
  const float *area = dom.n2pool + dom.n2off[slot];
  const uint32_t *bits = (const uint32_t *) (area + n2_oa_bits_at(L, B));
  const uint32_t *rows = (const uint32_t *) (area + n2_oa_rows_at(L, B));
  const uint32_t *pp = (const uint32_t *) (area + n2_pp_at(L, B));
  const uint8_t *sq = vres + voff[slot];               // residue i at sq[i-1]
  uint32_t *path = dom.n2path + dom.n2poff[slot];
  const int pcap = L + p.M + 2;
  auto cell = [&](int i, int k) -> uint32_t {
    const int z = (k - 1) / B, j = (k - 1) % B;
    return (bits[((size_t) (i - 1) * NW + j / 8) * 64 + z] >> ((j % 8) * 4)) & 15u;
  };
  const int NP = n2_pp_words(B);
  auto ppat = [&](int i, int k) -> uint32_t {
    const int z = (k - 1) / B, j = (k - 1) % B;
    return (pp[((size_t) (i - 1) * NP + j / 4) * 64 + z] >> ((j % 4) * 8)) & 0xffu;
  };
  int i = L, k = 0, s0 = S_C, nm = 0, ok = 1, nw = 0, last = 0;
  bool rec = false;                    // s0 is a display state still to be written
  bool spec = true;                    // a run may start here (false just after a broken run)
  const int guard = 4 * (L + p.M) + 8;
  for (int step = 0; s0 != S_S; step++) {
    int s1;
    if (step > guard || i < 0) { ok = 0; break; }
    if (spec && s0 == S_M && rec) {
      // M -> M steps from (i, k): the one from (i-l, k-l) needs both cells in
      // the model and the matrix, its record below the cap, its step within
      // the guard; each writes the word of the state it leaves
      const int lim = min(min(64, pcap - nw - 1), guard - step + 1);
      const int ii = i - lane, kk = k - lane;
      bool go = false;
      uint32_t wd = 0;
      if (lane < lim && ii >= 2 && kk >= 2 && kk <= p.M) {
        const uint32_t c = cell(ii, kk), b = ppat(ii, kk), x = (uint32_t) sq[ii - 1];
        go = (c & 3u) == 0u;
        wd = ((uint32_t) kk << 2) | ((b & 15u) << 19) | (x << 23);
      }
      const unsigned long long stop = __ballot(! go);
      const int r = stop ? (int) __ffsll((long long) stop) - 1 : 64;
      if (r > 0) {
        if (lane < r) path[nw + lane] = wd;
        nw += r; last = nw; nm += r; i -= r; k -= r;
        out.x = k; out.z = i;
        spec = (r == 64);
        step += r - 1;
        continue;
      }
    }
    if (spec && s0 == S_C) {
      // C -> C steps from row i down: row i-l's C bit, the next row in the
      // matrix, the step within the guard
      const int lim = min(64, guard - step + 1);
      const int ii = i - lane;
      const bool go = lane < lim && ii >= 1 && ((rows[ii] >> OA_CBIT) & 1u);
      const unsigned long long stop = __ballot(! go);
      const int r = stop ? (int) __ffsll((long long) stop) - 1 : 64;
      if (r > 0) {
        i -= r;
        spec = (r == 64);
        step += r - 1;
        continue;
      }
    }
    spec = true;
    // the display word of s0 (entered last step) is written here, so that its
    // loads and s0's choice loads go out together
    const bool mid = (s0 == S_M || s0 == S_I || s0 == S_D) && i >= 1 && k >= 1 && k <= p.M;
    const uint32_t cw = mid ? cell(i, k) : 0u;
    if (rec) {
      const uint32_t b = (s0 == S_D) ? 0u : ppat(i, k);
      const uint32_t st = (s0 == S_M) ? 0u : (s0 == S_I) ? 1u : 2u;
      const uint32_t code = (s0 == S_M) ? (b & 15u) : (s0 == S_I) ? (b >> 4) : 0u;
      const uint32_t x = (s0 == S_D) ? 0u : (uint32_t) sq[i - 1];
      if (lane == 0) path[nw] = st | ((uint32_t) k << 2) | (code << 19) | (x << 23);
      nw++;
      if (s0 == S_M) last = nw;
      rec = false;
    }
    switch (s0) {
      case S_C: s1 = ((rows[i] >> OA_CBIT) & 1u) ? S_C : S_E; break;
      case S_E: {
        const uint32_t r = rows[i];
        if ((r >> OA_EFAIL) & 1u) { ok = 0; break; }
        s1 = ((r >> OA_ESTATE) & 1u) ? S_D : S_M;
        k = (int) (r & 0xffffu);
        break;
      }
      case S_M: {
        if (i < 1 || k < 1 || k > p.M) { ok = 0; break; }
        const int c = (int) (cw & 3u);
        s1 = (c == 0) ? S_M : (c == 1) ? S_I : (c == 2) ? S_D : S_B;
        k--; i--;
        break;
      }
      case S_D:
        if (i < 1 || k < 1 || k > p.M) { ok = 0; break; }
        s1 = ((cw >> 3) & 1u) ? S_D : S_M;
        k--;
        break;
      case S_I:
        if (i < 1 || k < 1 || k > p.M) { ok = 0; break; }
        s1 = ((cw >> 2) & 1u) ? S_I : S_M;
        i--;
        break;
      case S_B: s1 = ((rows[i] >> OA_BBIT) & 1u) ? S_N : S_J; break;
      case S_N: s1 = (i == 0) ? S_S : S_N; break;
      case S_J: s1 = ((rows[i] >> OA_JBIT) & 1u) ? S_J : S_E; break;
      default: ok = 0; s1 = S_S; break;
    }
    if (! ok) break;
    if (s1 == S_M) {                     // walking back: the first M is the alignment's last
      if (nm == 0) { out.y = k; out.w = i; }
      out.x = k; out.z = i;
      nm++;
    }
    // the display's states: from the first M walked back to (states before
    // it are the E -> D end, not shown) to the last (B precedes it)
    if ((s1 == S_M || ((s1 == S_I || s1 == S_D) && nm > 0))) {
      if (i < 1 || k < 1 || k > p.M || nw >= pcap) { ok = 0; break; }
      rec = true;                        // its word, at the next step
    }
    if ((s1 == S_N || s1 == S_J || s1 == S_C) && s1 == s0) i--;
    s0 = s1;
  }
  if (! ok || nm == 0) { out = make_int4(0, 0, 0, 0); last = 0; }
  if (lane == 0) { dom.n2ali[slot] = out; dom.n2pn[slot] = last; }
}

hipError_t launch_oa_trace(const DevProfile &p, const DevSeqs &s, const int32_t *list, int64_t n, const DomBufs &dom,
                           hipStream_t st)
{
  if (n <= 0) return hipSuccess;
  if (n > 0x7fffffff) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_oa_trace, dim3((unsigned) n), dim3(64), 0, st, p, list, n, s.len, s.res, s.off, dom);
  return hipGetLastError();
}

hipError_t launch_msv_scores(const DevProfile &p, const DevSeqs &s, const int32_t *slots, int64_t n,
                             const int16_t *xj, float *sc, int *st, hipStream_t stream)
{
  if (n <= 0) return hipSuccess;
  unsigned grid = (unsigned) ((n + 255) / 256); if (grid > 65536) grid = 65536;
  hipLaunchKernelGGL(k_msv_scores, dim3(grid), dim3(256), 0, stream, p, s.len, slots, n, xj, sc, st);
  return hipGetLastError();
}

hipError_t launch_msv_classify(const DevProfile &p, const DevSeqs &s, const int32_t *slots, int64_t n,
                               const uint16_t *zmax, int16_t *xj, int32_t *full_list, int32_t *full_count,
                               LenCfg lc, hipStream_t st)
{
  if (n <= 0) return hipSuccess;
  unsigned grid = (unsigned) ((n + 255) / 256); if (grid > 65536) grid = 65536;
  hipLaunchKernelGGL(k_msv_classify, dim3(grid), dim3(256), 0, st, p, s.len, slots, n, zmax, xj, full_list, full_count, lc);
  return hipGetLastError();
}

#define MF_CASE(N, LDS) case N: k = (void *) k_msv_full<N, LDS>; break;
hipError_t launch_msv_full(const DevProfile &p, const DevSeqs &s, const int32_t *list, const int32_t *count,
                           int64_t maxn, int16_t *xj, LenCfg lc, hipStream_t st)
{
  if (maxn <= 0) return hipSuccess;
  // large models run the recurrence with several waves per target
  void *kw = nullptr;
  int ww = 0;
  switch (p.mf_ND) {
    case 20: kw = (void *) k_msv_full_wg<20, 4>; ww = 4; break;
    case 24: kw = (void *) k_msv_full_wg<24, 4>; ww = 4; break;
    case 32: kw = (void *) k_msv_full_wg<32, 8>; ww = 8; break;
    case 40: kw = (void *) k_msv_full_wg<40, 8>; ww = 8; break;
    default: break;
  }
  if (kw) {
    const unsigned grid = grid_for_waves_blk(kw, maxn * ww, 0, ww);
    void *args[] = { (void *) &p, (void *) &p.mf_tbl, (void *) &s.res, (void *) &s.off, (void *) &s.len,
                     (void *) &list, (void *) &count, &xj, &lc };
    return launch(kw, grid, 64 * ww, args, 0, st);
  }
  void *k = nullptr;
  switch (p.mf_ND) {
    MF_CASE(1, true) MF_CASE(2, true) MF_CASE(3, true) MF_CASE(4, true) MF_CASE(5, true) MF_CASE(6, true)
    MF_CASE(8, true) MF_CASE(10, true) MF_CASE(12, true) MF_CASE(16, true) MF_CASE(20, true)
    MF_CASE(24, false) MF_CASE(32, false) MF_CASE(40, false)
    default: return hipErrorInvalidValue;
  }
  const size_t lds = (p.mf_ND <= 20) ? p.mf_dwords * 4 : 0;
  const unsigned grid = grid_for_waves_blk(k, maxn, lds, LIST_BLOCK / 64);
  int tbl_dwords = (int) p.mf_dwords;
  void *args[] = { (void *) &p, (void *) &p.mf_tbl, &tbl_dwords, (void *) &s.res, (void *) &s.off, (void *) &s.len,
                   (void *) &list, (void *) &count, &xj, &lc };
  return launch(k, grid, LIST_BLOCK, args, lds, st);
}

hipError_t launch_msv_gate(const DevProfile &p, const DevSeqs &s, int64_t n, const int16_t *xj, DevResults r,
                           DevLists l, GateThresh g, hipStream_t st)
{
  if (n <= 0) return hipSuccess;
  unsigned grid = (unsigned) ((n + 255) / 256); if (grid > 65536) grid = 65536;
  hipLaunchKernelGGL(k_msv_gate, dim3(grid), dim3(256), 0, st, p, s.len, n, xj, r, l, g);
  return hipGetLastError();
}

hipError_t launch_bias(const DevProfile &p, const DevSeqs &s, const int32_t *list, const int32_t *count, int64_t maxn,
                       DevResults r, DevLists l, GateThresh g, int gating, LenCfg lc,
                       float *rowmax, int64_t rowmax_size, int64_t *defer_off, hipStream_t st)
{
  if (maxn <= 0) return hipSuccess;
  unsigned grid = (unsigned) ((maxn + 63) / 64); if (grid > 8192) grid = 8192;
  hipLaunchKernelGGL(k_bias, dim3(grid), dim3(64), 0, st, p, s.res, s.off, s.len, list, count, r, g, gating, lc,
                     rowmax, rowmax_size, l.cursor, defer_off);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || ! rowmax) return e;
  dev_props();
  hipLaunchKernelGGL(k_bias_log, dim3(ncu() * 8), dim3(256), 0, st, rowmax, (const unsigned long long *) l.cursor, rowmax_size);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  const unsigned wgrid = (unsigned) std::min<int64_t>((maxn + 3) / 4, 8192);
  hipLaunchKernelGGL(k_bias_sum, dim3(wgrid), dim3(256), 0, st, p, s.len, list, count, r, g, gating, lc,
                     (const float *) rowmax, (const int64_t *) defer_off);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  // the scale-factor pool's cursor starts from zero again for the Forward stage
  return launch_fill32(l.cursor, 0u, 2, st);
}

// Viterbi: the whole table in LDS while it fits (ND <= 16 at 160 KB), else the
// transitions in LDS and emission rows from global memory.
// The cooperative Viterbi kernel is the default for ND >= 20 (at ND = 16 it
// measured no faster than the one-wave kernel with its whole table in LDS);
// P7G_VIT_WG=0 selects the one-wave kernel (for comparisons).
static bool vit_wg_on()
{
  static const bool on = [] { const char *v = getenv("P7G_VIT_WG"); return ! (v && v[0] == '0'); }();
  return on;
}
#define VIT_CASE(N, T) case N: k = (void *) k_vit<N, T>; lds = (T == TM_LDS) ? p.vf_dwords * 4 : (size_t) P7G_VF_NT * 64 * N * 4; \
  nthr = VIT_THREADS(N); break;
hipError_t launch_vit(const DevProfile &p, const DevSeqs &s, const int32_t *list, const int32_t *count, int64_t maxn,
                      DevResults r, DevLists l, GateThresh g, int gating, LenCfg lc, hipStream_t st)
{
  if (maxn <= 0) return hipSuccess;
  // large models run the filter with several waves per target
  void *kw = nullptr;
  int ww = 0;
  switch (p.vf_ND) {
    case 20: kw = (void *) k_vit_wg<20, 4>; ww = 4; break;
    case 24: kw = (void *) k_vit_wg<24, 4>; ww = 4; break;
    case 32: kw = (void *) k_vit_wg<32, 8>; ww = 8; break;
    case 40: kw = (void *) k_vit_wg<40, 8>; ww = 8; break;
    default: break;
  }
  if (kw && vit_wg_on()) {
    const unsigned grid = grid_for_waves_blk(kw, maxn * ww, 0, ww);
    void *args[] = { (void *) &p, (void *) &s.res, (void *) &s.off, (void *) &s.len, (void *) &list, (void *) &count,
                     &r, &l, &g, &gating, &lc };
    return launch(kw, grid, 64 * ww, args, 0, st);
  }
  void *k = nullptr;
  size_t lds = 0;
  int nthr = LIST_BLOCK;
  switch (p.vf_ND) {
    VIT_CASE(1, TM_LDS) VIT_CASE(2, TM_LDS) VIT_CASE(3, TM_LDS) VIT_CASE(4, TM_LDS) VIT_CASE(5, TM_LDS)
    VIT_CASE(6, TM_LDS) VIT_CASE(8, TM_LDS) VIT_CASE(10, TM_LDS) VIT_CASE(12, TM_LDS) VIT_CASE(16, TM_LDS)
    VIT_CASE(20, TM_TLDS) VIT_CASE(24, TM_TLDS) VIT_CASE(32, TM_TLDS) VIT_CASE(40, TM_TLDS)
    default: return hipErrorInvalidValue;
  }
  const unsigned grid = grid_for_waves_blk(k, maxn, lds, nthr / 64);
  void *args[] = { (void *) &p, (void *) &s.res, (void *) &s.off, (void *) &s.len, (void *) &list, (void *) &count,
                   &r, &l, &g, &gating, &lc };
  return launch(k, grid, nthr, args, lds, st);
}

// Forward/Backward: the whole table in LDS up to B = 12 (150 KB); B = 14 reads
// the table from global memory with transitions pinned in registers;
// B = 16..48 keep the transitions in LDS (at most 160 KB) and read them per
// row; B = 64, 80 read everything from global memory per row.
#define FB_LDS_LIMIT (150 * 1024)
static int fb_mode(const DevProfile &p)
{
  if (p.fb_floats * 4 <= FB_LDS_LIMIT && p.fb_floats * 4 <= (size_t) lds_block_limit()) return TM_LDS;
  if (p.fb_B <= 14) return TM_PIN;
  if ((size_t) P7G_FB_NT * 64 * p.fb_B * 4 <= (size_t) lds_block_limit()) return TM_TLDS;
  return TM_GLOB;
}
static size_t fb_lds(const DevProfile &p, int tm)
{
  return (tm == TM_LDS) ? p.fb_floats * 4 : (tm == TM_TLDS) ? (size_t) P7G_FB_NT * 64 * p.fb_B * 4 : 0;
}
#define FB_CASES(X) \
  X(1, TM_LDS) X(2, TM_LDS) X(3, TM_LDS) X(4, TM_LDS) X(5, TM_LDS) X(6, TM_LDS) X(7, TM_LDS) X(8, TM_LDS) \
  X(10, TM_LDS) X(12, TM_LDS) X(14, TM_PIN) X(16, TM_TLDS) X(20, TM_TLDS) X(24, TM_TLDS) X(28, TM_TLDS) \
  X(32, TM_TLDS) X(40, TM_TLDS) X(48, TM_TLDS) X(64, TM_GLOB) X(80, TM_GLOB)
#define FWD_CASE(N, T) case N * 4 + T: k = dom.env_loop ? (dom.n2mode == 2 ? (void *) k_fwd<N, T, 3> : (void *) k_fwd<N, T, 2>) \
                                          : dom.pool ? (void *) k_fwd<N, T, 1> : (void *) k_fwd<N, T, 0>; \
                                          nthr = FWD_THREADS(N); break;
#define BCK_CASE(N, T) case N * 4 + T: k = dom.env_loop ? (void *) k_bck<N, T, 2> : dom.pool ? (void *) k_bck<N, T, 1> \
                                          : (void *) k_bck<N, T, 0>; nthr = BCK_THREADS(N); break;

hipError_t launch_fwd(const DevProfile &p, const DevSeqs &s, const int32_t *list, const int32_t *count, int64_t maxn,
                      DevResults r, DevLists l, GateThresh g, int gating, LenCfg lc,
                      float *scale_pool, int64_t pool_size, int64_t *scale_off, float *xmx_out, hipStream_t st,
                      const DomBufs &dom)
{
  if (maxn <= 0) return hipSuccess;
  // large models run the parser with several waves per target (4 up to 48
  // positions per lane, 8 above)
  void *kw = nullptr;
  int ww = 0;
  if (! dom.pool && ! dom.env_loop && ! xmx_out) {
    switch (p.fb_B) {
      case 16: kw = (void *) k_fwd_wg<16, 4>; ww = 4; break;
      case 20: kw = (void *) k_fwd_wg<20, 4>; ww = 4; break;
      case 24: kw = (void *) k_fwd_wg<24, 4>; ww = 4; break;
      case 28: kw = (void *) k_fwd_wg<28, 4>; ww = 4; break;
      case 32: kw = (void *) k_fwd_wg<32, 4>; ww = 4; break;
      case 40: kw = (void *) k_fwd_wg<40, 4>; ww = 4; break;
      case 48: kw = (void *) k_fwd_wg<48, 4>; ww = 4; break;
      case 64: kw = (void *) k_fwd_wg<64, 8>; ww = 8; break;
      case 80: kw = (void *) k_fwd_wg<80, 8>; ww = 8; break;
      default: break;
    }
  }
  if (kw) {
    const unsigned grid = grid_for_waves_blk(kw, maxn * ww, 0, ww);
    void *args[] = { (void *) &p, (void *) &s.res, (void *) &s.off, (void *) &s.len, (void *) &list, (void *) &count,
                     &r, &l, &g, &gating, &lc, &scale_pool, &pool_size, &scale_off };
    return launch(kw, grid, 64 * ww, args, 0, st);
  }
  const int tm = fb_mode(p);
  void *k = nullptr;
  int nthr = LIST_BLOCK;
  switch (p.fb_B * 4 + tm) {
    FB_CASES(FWD_CASE)
    default: return hipErrorInvalidValue;
  }
  // the null2 pass in workgroups of four waves (one per SIMD), so that one of
  // them fits on a compute unit beside a multidomain region's wave
  if (dom.env_loop && dom.n2mode == 2) nthr = std::min(nthr, 256);
  const size_t lds = fb_lds(p, tm);
  const unsigned grid = grid_for_waves_blk(k, maxn, lds, nthr / 64);
  void *args[] = { (void *) &p, (void *) &s.res, (void *) &s.off, (void *) &s.len, (void *) &list, (void *) &count,
                   &r, &l, &g, &gating, &lc, &scale_pool, &pool_size, &scale_off, &xmx_out, (void *) &dom };
  return launch(k, grid, nthr, args, lds, st);
}

hipError_t launch_bck(const DevProfile &p, const DevSeqs &s, const int32_t *list, const int32_t *count, int64_t maxn,
                      DevResults r, const float *scale_pool, const int64_t *scale_off, const float *fwd_xmx,
                      float *xmx_out, LenCfg lc, hipStream_t st, const DomBufs &dom, DevLists l)
{
  if (maxn <= 0) return hipSuccess;
  // large models run the parser with several waves per target, as Forward
  void *kw = nullptr;
  int ww = 0;
  if (! fwd_xmx && ! xmx_out && ! dom.pool && ! dom.env_loop) {
    switch (p.fb_B) {
      case 16: kw = (void *) k_bck_wg<16, 4>; ww = 4; break;
      case 20: kw = (void *) k_bck_wg<20, 4>; ww = 4; break;
      case 24: kw = (void *) k_bck_wg<24, 4>; ww = 4; break;
      case 28: kw = (void *) k_bck_wg<28, 4>; ww = 4; break;
      case 32: kw = (void *) k_bck_wg<32, 4>; ww = 4; break;
      case 40: kw = (void *) k_bck_wg<40, 4>; ww = 4; break;
      case 48: kw = (void *) k_bck_wg<48, 4>; ww = 4; break;
      case 64: kw = (void *) k_bck_wg<64, 8>; ww = 8; break;
      case 80: kw = (void *) k_bck_wg<80, 8>; ww = 8; break;
      default: break;
    }
  }
  if (kw) {
    const unsigned grid = grid_for_waves_blk(kw, maxn * ww, 0, ww);
    void *args[] = { (void *) &p, (void *) &s.res, (void *) &s.off, (void *) &s.len, (void *) &list, (void *) &count,
                     &r, (void *) &scale_pool, (void *) &scale_off, &lc };
    return launch(kw, grid, 64 * ww, args, 0, st);
  }
  const int tm = fb_mode(p);
  void *k = nullptr;
  int nthr = LIST_BLOCK;
  switch (p.fb_B * 4 + tm) {
    FB_CASES(BCK_CASE)
    default: return hipErrorInvalidValue;
  }
  const size_t lds = fb_lds(p, tm);
  const unsigned grid = grid_for_waves_blk(k, maxn, lds, nthr / 64);
  void *args[] = { (void *) &p, (void *) &s.res, (void *) &s.off, (void *) &s.len, (void *) &list, (void *) &count,
                   &r, (void *) &scale_pool, (void *) &scale_off, (void *) &fwd_xmx, &xmx_out, &lc, (void *) &dom, &l };
  return launch(k, grid, nthr, args, lds, st);
}

hipError_t launch_domains(const DevProfile &p, const DevSeqs &s, const int32_t *list, const int32_t *count, int64_t maxn,
                          const DevResults &r, DevLists l, LenCfg lc, const DomBufs &dom, const DomOut &out,
                          hipStream_t st)
{
  if (maxn <= 0) return hipSuccess;
  if (! dom.pool || ! dom.doff || ! dom.boff || ! out.nexp || ! out.nreg || ! out.nclu || ! out.regs || ! out.ecount ||
      ! out.res || ! out.v_off || ! out.v_len || ! out.v_list || ! out.env_loop || ! out.env_move) return hipErrorInvalidValue;
  const unsigned grid = (unsigned) std::min<int64_t>((maxn + 3) / 4, 8192);
  hipLaunchKernelGGL(k_domains, dim3(grid), dim3(256), 0, st, p, s.res, s.off, s.len, list, count, (const int *) r.fstatus,
                     l, lc, dom, out);
  return hipGetLastError();
}

/*****************************************************************
 * Pipeline counters and caller-order outputs
 *****************************************************************/
__global__ __launch_bounds__(256)
void k_stats(const int32_t *__restrict__ len, const uint8_t *__restrict__ stage, int64_t n,
             const int32_t *__restrict__ vit, const int32_t *__restrict__ fwd, const int32_t *__restrict__ counts,
             unsigned long long *__restrict__ out)
{
  unsigned long long a[8] = { 0, 0, 0, 0, 0, 0, 0, 0 };
  const int64_t i0 = (int64_t) blockIdx.x * blockDim.x + threadIdx.x, step = (int64_t) gridDim.x * blockDim.x;
  for (int64_t slot = i0; slot < n; slot += step) {
    const int sg = stage[slot];
    const unsigned long long L = (unsigned long long) len[slot];
    a[0] += sg >= 1; a[1] += sg >= 2; a[2] += sg >= 3; a[3] += sg >= 4;
    a[4] += sg >= 1 ? L : 0ull;
    a[5] += sg >= 4 ? L : 0ull;
  }
  const int64_t nv = counts[2], nf = counts[3];
  for (int64_t t = i0; t < nv; t += step) a[6] += (unsigned long long) len[vit[t]];
  for (int64_t t = i0; t < nf; t += step) a[7] += (unsigned long long) len[fwd[t]];
  __shared__ unsigned long long red[8][4];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
  for (int k = 0; k < 8; k++) {
    unsigned long long v = a[k];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if (lane == 0) red[k][wv] = v;
  }
  __syncthreads();
  if (threadIdx.x < 8) {
    const unsigned long long v = red[threadIdx.x][0] + red[threadIdx.x][1] + red[threadIdx.x][2] + red[threadIdx.x][3];
    if (v) atomicAdd(&out[threadIdx.x], v);
  }
}

__global__ void k_scatter32(const int32_t *__restrict__ perm, const uint32_t *__restrict__ src, uint32_t *__restrict__ dst, int64_t n)
{
  for (int64_t s = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; s < n; s += (int64_t) gridDim.x * blockDim.x)
    dst[perm[s]] = src[s];
}
__global__ void k_scatter8(const int32_t *__restrict__ perm, const uint8_t *__restrict__ src, uint8_t *__restrict__ dst, int64_t n)
{
  for (int64_t s = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; s < n; s += (int64_t) gridDim.x * blockDim.x)
    dst[perm[s]] = src[s];
}

hipError_t launch_stats(const DevSeqs &s, const uint8_t *stage, int64_t n, const int32_t *vit, const int32_t *fwd,
                        const int32_t *counts, unsigned long long *out, hipStream_t st)
{
  if (n <= 0) return hipSuccess;
  unsigned grid = (unsigned) std::min<int64_t>((n + 255) / 256, (int64_t) ncu() * 2);
  hipLaunchKernelGGL(k_stats, dim3(grid), dim3(256), 0, st, s.len, stage, n, vit, fwd, counts, out);
  return hipGetLastError();
}
// each region record's caller index and target length, taken while the
// database is at hand, so that the run's results outlive it
__global__ void k_rec_gather(const int4 *__restrict__ regs, int64_t n, const int32_t *__restrict__ perm,
                             const int32_t *__restrict__ len, int32_t *__restrict__ rseq, int32_t *__restrict__ rlen)
{
  for (int64_t r = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; r < n; r += (int64_t) gridDim.x * blockDim.x) {
    const int slot = regs[r].x;
    rseq[r] = perm[slot];
    rlen[r] = len[slot];
  }
}
hipError_t launch_rec_gather(const DevSeqs &s, const int4 *regs, int64_t n, int32_t *rseq, int32_t *rlen, hipStream_t st)
{
  if (n <= 0) return hipSuccess;
  unsigned grid = (unsigned) std::min<int64_t>((n + 255) / 256, 4096);
  hipLaunchKernelGGL(k_rec_gather, dim3(grid), dim3(256), 0, st, regs, n, s.perm, s.len, rseq, rlen);
  return hipGetLastError();
}
hipError_t launch_scatter32(const DevSeqs &s, const void *src, void *dst, int64_t n, hipStream_t st)
{
  if (n <= 0) return hipSuccess;
  unsigned grid = (unsigned) std::min<int64_t>((n + 255) / 256, 65536);
  hipLaunchKernelGGL(k_scatter32, dim3(grid), dim3(256), 0, st, s.perm, (const uint32_t *) src, (uint32_t *) dst, n);
  return hipGetLastError();
}
hipError_t launch_scatter8(const DevSeqs &s, const uint8_t *src, uint8_t *dst, int64_t n, hipStream_t st)
{
  if (n <= 0) return hipSuccess;
  unsigned grid = (unsigned) std::min<int64_t>((n + 255) / 256, 65536);
  hipLaunchKernelGGL(k_scatter8, dim3(grid), dim3(256), 0, st, s.perm, src, dst, n);
  return hipGetLastError();
}

/*****************************************************************
 * Order-preserving compaction of the survivor lists
 *****************************************************************/
#define CT_TILE 4096     // slots per block: 256 threads x 16 consecutive slots

template <typename T>
__device__ __forceinline__ int tile_flags(const T *__restrict__ f, int value, int64_t base, int64_t n, unsigned &mask)
{
  mask = 0u;
#pragma unroll
  for (int u = 0; u < 16; u++) {
    const int64_t s = base + u;
    if (s < n && (int) f[s] == value) mask |= 1u << u;
  }
  return __popc(mask);
}

template <typename T>
__global__ __launch_bounds__(256)
void k_compact_count(const T *__restrict__ f, int value, int64_t n, int32_t *__restrict__ tilecnt)
{
  __shared__ int part[4];
  unsigned m;
  int c = tile_flags(f, value, (int64_t) blockIdx.x * CT_TILE + threadIdx.x * 16, n, m);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) tilecnt[blockIdx.x] = part[0] + part[1] + part[2] + part[3];
}

template <typename T>
__global__ __launch_bounds__(256)
void k_compact_scatter(const T *__restrict__ f, int value, int64_t n, const int32_t *__restrict__ tilecnt,
                       int32_t *__restrict__ list, int32_t *__restrict__ count)
{
  __shared__ int part[4];
  __shared__ int wsum[4];
  // start of this tile's output: the counts of the tiles before it
  int before = 0;
  for (int b = threadIdx.x; b < (int) blockIdx.x; b += blockDim.x) before += tilecnt[b];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) before += __shfl_xor(before, o, 64);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = before;
  // this thread's 16 slots, and an exclusive scan of the counts over the block
  unsigned m;
  const int64_t base = (int64_t) blockIdx.x * CT_TILE + threadIdx.x * 16;
  const int c = tile_flags(f, value, base, n, m);
  const int lane = threadIdx.x & 63;
  int incl = c;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) { const int v = __shfl_up(incl, o, 64); if (lane >= o) incl += v; }
  if (lane == 63) wsum[threadIdx.x >> 6] = incl;
  __syncthreads();
  const int start = part[0] + part[1] + part[2] + part[3];
  int woff = 0;
  for (int w = 0; w < (int) (threadIdx.x >> 6); w++) woff += wsum[w];
  int at = start + woff + incl - c;
#pragma unroll
  for (int u = 0; u < 16; u++) if (m & (1u << u)) list[at++] = (int32_t) (base + u);
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == blockDim.x - 1) *count = start + woff + incl;
}

template <typename T>
static hipError_t launch_compact(const T *f, int value, int64_t n, int32_t *list, int32_t *count, int32_t *tilecnt,
                                 hipStream_t st)
{
  if (n <= 0) return launch_fill32(count, 0u, 1, st);
  const unsigned tiles = (unsigned) ((n + CT_TILE - 1) / CT_TILE);
  hipLaunchKernelGGL(k_compact_count<T>, dim3(tiles), dim3(256), 0, st, f, value, n, tilecnt);
  hipLaunchKernelGGL(k_compact_scatter<T>, dim3(tiles), dim3(256), 0, st, f, value, n, (const int32_t *) tilecnt, list, count);
  return hipGetLastError();
}
hipError_t launch_compact_u8(const uint8_t *flags, int value, int64_t n, int32_t *list, int32_t *count,
                             int32_t *tilecnt, hipStream_t st)
{
  return launch_compact<uint8_t>(flags, value, n, list, count, tilecnt, st);
}
hipError_t launch_compact_i16(const int16_t *flags, int value, int64_t n, int32_t *list, int32_t *count,
                              int32_t *tilecnt, hipStream_t st)
{
  return launch_compact<int16_t>(flags, value, n, list, count, tilecnt, st);
}
