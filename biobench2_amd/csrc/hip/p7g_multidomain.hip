// Multidomain regions: region_trace_ensemble (HMMER/src/p7_domaindef.c:600-690)
// as p7_domaindef_ByPosteriorHeuristics runs it for a region that looks like
// more than one domain (:431-476). One 64-lane workgroup per region:
//
//  1. p7_Forward over the region's residues with the profile in multihit
//     mode, length-configured to the whole target, keeping every row
//     (fwdback.c:258-465 with do_full). The sampled traces below are chaotic
//     in these values: a last-bit difference moves a choice, the random
//     stream then falls out of step, and every later trace differs. So the
//     matrix is computed in the reference's own operation order: its four
//     striped lanes (model position k = r*Q + q + 1) are lanes 0..3 here for
//     the sums that run along q (the row's E sum, the D->D passes), while the
//     cells' M and I values, which depend only on the row before, are spread
//     over the 64 lanes. The file is compiled without contraction into fused
//     multiply-adds and with flushed fp32 denormals, as the reference runs
//     with FTZ/DAZ set; each operation is written out in the reference's
//     order.
//  2. nsamples stochastic traces (impl_sse/stotrace.c:76-329) with Easel's
//     fast generator reset to the seed (esl_random.c:148-160, 215-230,
//     289-310). A step's candidate paths are lanes 0..3: one product each,
//     esl_vec_FNorm's float sum and division, esl_rnd_FChoose's double sums
//     and divisions, and the first lane whose threshold exceeds the roll.
//     select_e scans row i in the striped order with a double running sum.
//     Each sampled domain (p7_trace_Index, p7_trace.c:1173-1206) gets the null2
//     odds of its own trace segment (p7_Null2_ByTrace, impl_sse/null2.c:134-
//     205: every emitting M or I state counts at the M cell of its node), and
//     every residue of the region adds, per trace, that odds or 1 outside
//     domains; n2sc = logf(sum / nsamples).
//  3. The sampled domains are clustered by single linkage (esl_cluster.c:
//     136-185 with link_spsamples, p7_spensemble.c:191-222), clusters with
//     posterior >= 0.25 get the widest endpoints with >= 2% support
//     (p7_spensemble.c:282-411), sorted by start, and clusters dominated by an
//     overlapping one of higher posterior are dropped (p7_domaindef.c:650-682).
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include <algorithm>

#include "p7g_device.h"
#include "p7g_devfn.h"

// (also stated here, so that no build of this file contracts)
#pragma clang fp contract(off)

enum { MS_M = 1, MS_D, MS_I, MS_S, MS_N, MS_B, MS_E, MS_C, MS_T, MS_J };

// Bob Jenkins' 96-bit mix, as esl_random.c disperses a seed
__device__ __forceinline__ uint32_t md_mix3(uint32_t a, uint32_t b, uint32_t c)
{
  a = (a - b - c) ^ (c >> 13);
  b = (b - c - a) ^ (a << 8);
  c = (c - a - b) ^ (b >> 13);
  a = (a - b - c) ^ (c >> 12);
  b = (b - c - a) ^ (a << 16);
  c = (c - a - b) ^ (b >> 5);
  a = (a - b - c) ^ (c >> 3);
  b = (b - c - a) ^ (a << 10);
  c = (c - a - b) ^ (b >> 15);
  return c;
}

// the device Forward table's value of row `row` at model position k (0 past M)
__device__ __forceinline__ float md_tv(const DevProfile &p, int row, int k)
{
  if (k < 1 || k > p.M) return 0.0f;
  const int B = p.fb_B, zl = (k - 1) / B, j = (k - 1) - zl * B;
  return p.fb_tbl[((size_t) row * B + j) * 64 + zl];
}

// the device table row of the i-th staged transition row
__device__ __forceinline__ int md_trow(int i)
{
  switch (i) {
  case 0: return VT_BM; case 1: return VT_MM; case 2: return VT_IM; case 3: return VT_DM;
  case 4: return VT_MDS; case 5: return VT_DDS; case 6: return VT_MI; case 7: return VT_II;
  case 8: return FT_MD; default: return FT_DD;
  }
}

// esl_vec_FNorm then esl_rnd_FChoose's thresholds for n choices of weights
// v: the float sum and divisions, then the double running sums each divided
// by their total; the first n - 1 of them go to th (the last is 1)
__device__ __forceinline__ void md_thresholds(const float *v, int n, double *th)
{
  float s = 0.0f;
  for (int a = 0; a < n; a++) s = __fadd_rn(s, v[a]);
  double c[4], tot = 0.0;
  for (int a = 0; a < n; a++) {
    const float pa = (s != 0.0f) ? __fdiv_rn(v[a], s) : (float) (1.0 / (double) (float) n);
    tot = tot + (double) pa;
    c[a] = tot;
  }
  for (int a = 0; a < n - 1; a++) th[a] = c[a] / tot;
}

// The workgroup is one wavefront, whose LDS operations complete in the order
// it issues them: between lanes that only exchange data through LDS it is
// enough that the compiler moves no memory operation across this point (a
// workgroup barrier would also wait for the outstanding global stores).
__device__ __forceinline__ void md_sync()
{
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

struct MdSeg { int idx, i, j, k, m; };

// link_spsamples: overlap >= 0.8 of the smaller segment in the sequence and in
// the model (the model overlap without the +1, as the reference has it), and
// start or end diagonals within 4
__device__ __forceinline__ bool md_linked(const MdSeg &a, const MdSeg &b)
{
  int nov = min(a.j, b.j) - max(a.i, b.i) + 1;
  int n = min(a.j - a.i + 1, b.j - b.i + 1);
  if (__fdiv_rn((float) nov, (float) n) < 0.8f) return false;
  nov = min(a.m, b.m) - max(a.k, b.k);
  n = min(a.m - a.k + 1, b.m - b.k + 1);
  if (__fdiv_rn((float) nov, (float) n) < 0.8f) return false;
  if (abs((a.i - a.k) - (b.i - b.k)) <= 4) return true;
  if (abs((a.j - a.m) - (b.j - b.m)) <= 4) return true;
  return false;
}

struct MdSig { int c, i, j, k, m; float prob; };

__global__ __launch_bounds__(64) void
k_md_region(const DevProfile p, const uint8_t *__restrict__ res, const MdJob *__restrict__ jobs, MdOut *__restrict__ out,
            float *__restrict__ pool, int nsamples, uint32_t seed, int phase, int acc_cap)
{
  extern __shared__ float lds[];
  const int z = threadIdx.x;
  const MdJob J = jobs[blockIdx.x];
  MdOut *O = out + blockIdx.x;
  const int M = p.M, Q = max(2, (M - 1) / 4 + 1), Lr = J.Lr, L = J.Ltarget;
  const int TR = p.nrows;                                  // first transition row of the table
  const float xEmove = p.xfE_move, xEloop = p.xfE_loop, floop = p.xf_loop[L], fmove = p.xf_move[L];
  const uint8_t *dsq = res + J.res_off;                    // residue i (1-based) at dsq[i-1]
  float *mat = pool + J.area;                              // rows 0..Lr, [q][s][r], s = 0 M, 1 D, 2 I
  float *xmx = mat + (size_t) (Lr + 1) * 12 * Q;           // rows 0..Lr: E N J B C SCALE
  float *acc = xmx + (size_t) 6 * (Lr + 1);                // residues 0..Lr
  int   *hist = (int *) (acc + (Lr + 1));                  // max(Lr, M) + 2 endpoint counts
  MdSeg *segs = (MdSeg *) (pool + J.area + md_seg_at(M, Lr));   // the ensemble, J.segcap entries
  int status = 0;
#define CELL(i_, q_, s_, r_) mat[(((size_t) (i_) * Q + (q_)) * 3 + (s_)) * 4 + (r_)]

  // The profile's transitions, staged in LDS when they fit (md_stage_t): rows
  // BM MM IM DM (into k), MDS DDS (into D_k), MI II, MD DD (out of k), each
  // [k = 0..M+1] with 0 outside 1..M
  const bool stage = md_stage_t(M);
  const int TW = M + 2;
  float *tl = lds;
  float *ph = lds + (stage ? md_stage_floats(M) : 0);       // the phases' own LDS
  if (stage && phase == 0)
    for (int c = z; c < 10 * TW; c += 64) tl[c] = md_tv(p, TR + md_trow(c / TW), c % TW);
  auto tv = [&](int row, int k) -> float {                 // row: 0..9 as staged
    if (stage) return tl[row * TW + k];
    return md_tv(p, TR + md_trow(row), k);
  };
  enum { T_BM = 0, T_MM, T_IM, T_DM, T_MDS, T_DDS, T_MI, T_II, T_MD, T_DD };

  // ---------------- 1. the region's Forward matrix ----------------
  if (phase == 0) {
    float *prev = ph, *cur = ph + 12 * Q, *tdd = ph + 24 * Q, *erow = ph + 28 * Q;   // erow: 4Q emissions
    __syncthreads();
    for (int c = z; c < 12 * Q; c += 64) { prev[c] = 0.0f; mat[c] = 0.0f; }
    for (int c = z; c < 4 * Q; c += 64) tdd[c] = tv(T_DD, min((c & 3) * Q + (c >> 2) + 1, M + 1));
    float xN = 1.0f, xJ = 0.0f, xB = fmove, xC = 0.0f, xE = 0.0f;
    if (z == 0) { xmx[0] = 0.0f; xmx[1] = xN; xmx[2] = xJ; xmx[3] = xB; xmx[4] = xC; xmx[5] = 1.0f; }
    __syncthreads();
    // a row's odds in striped order, eight per lane: loaded one row ahead
    // while the row before is computed (models up to M = 512; larger ones load
    // theirs at the row)
    const bool pre = 4 * Q <= 64 * 8;
    float enext[8];
    auto load_row = [&](int i1) {
      const int x = dsq[i1 - 1];
#pragma unroll
      for (int u = 0; u < 8; u++) { const int c = u * 64 + z; enext[u] = md_tv(p, x, (c & 3) * Q + (c >> 2) + 1); }
    };
    // with the row's odds in registers (pre), lane z's cells are c = 64 u + z
    // for u < 8, and their transitions are the row loop's constants: in
    // registers too
    float ecur[8], tBMr[8], tMMr[8], tIMr[8], tDMr[8], tMIr[8], tIIr[8], tMDr[8];
    if (pre) {
#pragma unroll
      for (int u = 0; u < 8; u++) {
        const int c = u * 64 + z, q = c >> 2, r = c & 3, k = r * Q + q + 1, kk = min(k, M + 1);
        const bool in = c < 4 * Q;
        tBMr[u] = in ? tv(T_BM, kk) : 0.0f; tMMr[u] = in ? tv(T_MM, kk) : 0.0f;
        tIMr[u] = in ? tv(T_IM, kk) : 0.0f; tDMr[u] = in ? tv(T_DM, kk) : 0.0f;
        tMIr[u] = in ? tv(T_MI, kk) : 0.0f; tIIr[u] = in ? tv(T_II, kk) : 0.0f;
        tMDr[u] = in ? tv(T_MD, min(k - 1, M + 1)) : 0.0f;
      }
      load_row(1);
    }
    for (int i = 1; i <= Lr; i++) {
      if (pre) {
#pragma unroll
        for (int u = 0; u < 8; u++) ecur[u] = enext[u];
        if (i < Lr) load_row(i + 1);
      } else {
        const int x = dsq[i - 1];
        for (int c0 = 0; c0 < 4 * Q; c0 += 64 * 8) {         // this residue's odds, striped order
          float e[8];
#pragma unroll
          for (int u = 0; u < 8; u++) { const int c = c0 + u * 64 + z; e[u] = md_tv(p, x, (c & 3) * Q + (c >> 2) + 1); }
#pragma unroll
          for (int u = 0; u < 8; u++) { const int c = c0 + u * 64 + z; if (c < 4 * Q) erow[c] = e[u]; }
        }
      }
      md_sync();
      if (pre) {
#pragma unroll
        for (int u = 0; u < 8; u++) {
          const int c = u * 64 + z, q = c >> 2, r = c & 3;
          if (c < 4 * Q) {
            float mp = 0.0f, dp = 0.0f, ip = 0.0f;       // row i-1 at k-1 (esl_sse_rightshift_ps at q = 0)
            if (q > 0)      { mp = prev[((q-1)*3+0)*4 + r]; dp = prev[((q-1)*3+1)*4 + r]; ip = prev[((q-1)*3+2)*4 + r]; }
            else if (r > 0) { mp = prev[((Q-1)*3+0)*4 + r-1]; dp = prev[((Q-1)*3+1)*4 + r-1]; ip = prev[((Q-1)*3+2)*4 + r-1]; }
            float sv = __fmul_rn(xB, tBMr[u]);
            sv = __fadd_rn(sv, __fmul_rn(mp, tMMr[u]));
            sv = __fadd_rn(sv, __fmul_rn(ip, tIMr[u]));
            sv = __fadd_rn(sv, __fmul_rn(dp, tDMr[u]));
            sv = __fmul_rn(sv, ecur[u]);
            cur[(q*3+0)*4 + r] = sv;
            cur[(q*3+2)*4 + r] = __fadd_rn(__fmul_rn(prev[(q*3+0)*4 + r], tMIr[u]),
                                           __fmul_rn(prev[(q*3+2)*4 + r], tIIr[u]));
          }
        }
        md_sync();
        // D before the D->D passes: cell q holds M(q-1)·tMD (the q loop's dcv); q = 0 holds 0
#pragma unroll
        for (int u = 0; u < 8; u++) {
          const int c = u * 64 + z, q = c >> 2, r = c & 3;
          if (c < 4 * Q) cur[(q*3+1)*4 + r] = q ? __fmul_rn(cur[((q-1)*3+0)*4 + r], tMDr[u]) : 0.0f;
        }
      } else {
      for (int c = z; c < 4 * Q; c += 64) {
        const int q = c >> 2, r = c & 3, k = r * Q + q + 1, kk = min(k, M + 1);
        float mp = 0.0f, dp = 0.0f, ip = 0.0f;           // row i-1 at k-1 (esl_sse_rightshift_ps at q = 0)
        if (q > 0)      { mp = prev[((q-1)*3+0)*4 + r]; dp = prev[((q-1)*3+1)*4 + r]; ip = prev[((q-1)*3+2)*4 + r]; }
        else if (r > 0) { mp = prev[((Q-1)*3+0)*4 + r-1]; dp = prev[((Q-1)*3+1)*4 + r-1]; ip = prev[((Q-1)*3+2)*4 + r-1]; }
        float sv = __fmul_rn(xB, tv(T_BM, kk));
        sv = __fadd_rn(sv, __fmul_rn(mp, tv(T_MM, kk)));
        sv = __fadd_rn(sv, __fmul_rn(ip, tv(T_IM, kk)));
        sv = __fadd_rn(sv, __fmul_rn(dp, tv(T_DM, kk)));
        sv = __fmul_rn(sv, erow[c]);
        cur[(q*3+0)*4 + r] = sv;
        cur[(q*3+2)*4 + r] = __fadd_rn(__fmul_rn(prev[(q*3+0)*4 + r], tv(T_MI, kk)),
                                       __fmul_rn(prev[(q*3+2)*4 + r], tv(T_II, kk)));
      }
      md_sync();
      // D before the D->D passes: cell q holds M(q-1)·tMD (the q loop's dcv); q = 0 holds 0
      for (int c = z; c < 4 * Q; c += 64) {
        const int q = c >> 2, r = c & 3, k = r * Q + q + 1;
        cur[(q*3+1)*4 + r] = q ? __fmul_rn(cur[((q-1)*3+0)*4 + r], tv(T_MD, min(k - 1, M + 1))) : 0.0f;
      }
      }
      md_sync();
      float xe = 0.0f;
      if (z < 4) {                                         // lane z is SSE lane r
        // The row's sums along q and its D->D passes, each a chain in q order
        // (esl_sse_hsum_ps's lanes, fwdback.c's D passes): M(q) at cm[12 q],
        // D(q) at cd[12 q], tDD(q) at tq[4 q]. Eight cells at a time, the next
        // eight's loads issued before this eight's chain. The M sum and the
        // first pass are independent chains and run together; every later
        // pass also sums its new D values onto the M sum, so that the last
        // pass leaves the row's total (its D values are the final ones).
        const int r = z;
        float *cm = cur + r, *cd = cur + 4 + r;
        const float *tq = tdd + r;
        float dcv = __fmul_rn(cur[((Q-1)*3+0)*4 + r], tv(T_MD, min(r * Q + Q, M + 1)));
        float sh = __shfl(dcv, (z + 3) & 3, 4);
        dcv = r ? sh : 0.0f;
        // Reads past the row's last q stay inside the phase's LDS (the next
        // buffers) and feed no result.
        float xm = 0.0f;
        {
          // two chunk buffers in turn: the next chunk loads into one while
          // the other's chain runs
          float am[8], ad[8], at[8], bm[8], bd[8], bt[8];
          auto ld = [&](int q0, float (&m)[8], float (&d)[8], float (&t)[8]) {
#pragma unroll
            for (int u = 0; u < 8; u++) { m[u] = cm[12*(q0+u)]; d[u] = cd[12*(q0+u)]; t[u] = tq[4*(q0+u)]; }
          };
          auto run = [&](int q0, const float (&m)[8], const float (&d)[8], const float (&t)[8]) {
            auto step = [&](int u) {
              xm = __fadd_rn(xm, m[u]);
              const float dd = __fadd_rn(dcv, d[u]);
              dcv = __fmul_rn(dd, t[u]);
              cd[12*(q0+u)] = dd;
            };
            if (q0 + 8 <= Q) {
#pragma unroll
              for (int u = 0; u < 8; u++) step(u);
            } else {
#pragma unroll
              for (int u = 0; u < 8; u++) if (q0 + u < Q) step(u);
            }
          };
          ld(0, am, ad, at);
          for (int q0 = 0; q0 < Q; q0 += 16) {
            if (q0 + 8 < Q) ld(q0 + 8, bm, bd, bt);
            run(q0, am, ad, at);
            if (q0 + 8 >= Q) break;
            if (q0 + 16 < Q) ld(q0 + 16, am, ad, at);
            run(q0 + 8, bm, bd, bt);
          }
        }
        for (int pass = 1; pass < 4; pass++) {
          sh = __shfl(dcv, (z + 3) & 3, 4);
          dcv = r ? sh : 0.0f;
          int changed = 0;
          float xd = xm;
          float ad[8], at[8], bd[8], bt[8];
          auto ld = [&](int q0, float (&d)[8], float (&t)[8]) {
#pragma unroll
            for (int u = 0; u < 8; u++) { d[u] = cd[12*(q0+u)]; t[u] = tq[4*(q0+u)]; }
          };
          auto run = [&](int q0, const float (&d)[8], const float (&t)[8]) {
            auto step = [&](int u) {
              const float dd = __fadd_rn(dcv, d[u]);
              changed |= (dd > d[u]);
              dcv = __fmul_rn(dcv, t[u]);
              xd = __fadd_rn(dd, xd);
              cd[12*(q0+u)] = dd;
            };
            if (q0 + 8 <= Q) {
#pragma unroll
              for (int u = 0; u < 8; u++) step(u);
            } else {
#pragma unroll
              for (int u = 0; u < 8; u++) if (q0 + u < Q) step(u);
            }
          };
          ld(0, ad, at);
          for (int q0 = 0; q0 < Q; q0 += 16) {
            if (q0 + 8 < Q) ld(q0 + 8, bd, bt);
            run(q0, ad, at);
            if (q0 + 8 >= Q) break;
            if (q0 + 16 < Q) ld(q0 + 16, ad, at);
            run(q0 + 8, bd, bt);
          }
          xe = xd;
          const bool any = (__ballot(changed) & 0xfull) != 0ull;
          if (M >= 100 && ! any) break;
        }
      }
      // esl_sse_hsum_ps: (e0 + e1) + (e2 + e3)
      xE = __fadd_rn(__fadd_rn(readlanef(xe, 0), readlanef(xe, 1)), __fadd_rn(readlanef(xe, 2), readlanef(xe, 3)));
      xN = __fmul_rn(xN, floop);
      xC = __fadd_rn(__fmul_rn(xC, floop), __fmul_rn(xE, xEmove));
      xJ = __fadd_rn(__fmul_rn(xJ, floop), __fmul_rn(xE, xEloop));
      xB = __fadd_rn(__fmul_rn(xJ, fmove), __fmul_rn(xN, fmove));
      float scale = 1.0f;
      md_sync();
      if (xE > 1.0e4f) {
        xN = __fdiv_rn(xN, xE); xC = __fdiv_rn(xC, xE); xJ = __fdiv_rn(xJ, xE); xB = __fdiv_rn(xB, xE);
        const float inv = (float) (1.0 / (double) xE);
        for (int c = z; c < 12 * Q; c += 64) cur[c] = __fmul_rn(cur[c], inv);
        scale = xE;
        xE = 1.0f;
        md_sync();
      }
      float *row = mat + (size_t) i * 12 * Q;
      for (int c = z; c < 12 * Q; c += 64) row[c] = cur[c];
      if (z == 0) {
        float *xr = xmx + (size_t) i * 6;
        xr[0] = xE; xr[1] = xN; xr[2] = xJ; xr[3] = xB; xr[4] = xC; xr[5] = scale;
      }
      float *t = prev; prev = cur; cur = t;
      md_sync();
    }
    __syncthreads();                                       // the matrix in global memory, for every lane
    return;
  }

  // ---------------- 2. the sampled traces ----------------
  // (their thresholds were filled by k_md_thr, md_thr_fill)
  double *thr = (double *) (pool + J.area + md_thr_at(M, Lr, J.segcap));
  double *xthr = thr + (size_t) 5 * (Lr + 1) * (M + 1);
  double *ecum = xthr + (size_t) 3 * (Lr + 1);
  // LDS: a domain's counts, its null2 odds, the trace's domains, and the
  // residues' odds sums when the region is short
  float *cnt = lds;                                        // [q][r] counts of one sampled domain
  float *n2v = lds + 4 * Q;                                // its null2 odds, 32 codes
  int   *tseg = (int *) (n2v + 32);                        // this trace's domains, last first
  const bool acc_lds = (Lr + 1) <= acc_cap;
  float *accp = acc_lds ? lds + md_acc_at(Q) : acc;        // residue sums
  uint8_t *rl = (uint8_t *) (lds + md_acc_at(Q) + acc_cap);   // and the residues, with them in LDS
  const uint8_t *rsq = acc_lds ? rl : dsq;
  auto cnt_at = [&](int k) {                               // the count cell of model position k
    const int r = (k - 1 >= Q) + (k - 1 >= 2 * Q) + (k - 1 >= 3 * Q);
    return (k - 1 - r * Q) * 4 + r; };
  const bool em_lds = md_em_lds(M, acc_cap) && p.K <= 20;
  float *em = lds + md_acc_at(Q) + acc_cap + acc_cap / 4;  // [x][q][r] odds, when em_lds
  if (em_lds)
    for (int c = z; c < p.K * 4 * Q; c += 64) {
      const int x = c / (4 * Q), rem = c - x * 4 * Q;
      em[c] = md_tv(p, x, min((rem & 3) * Q + (rem >> 2) + 1, M + 1));
    }
  for (int c = z; c < 4 * Q; c += 64) cnt[c] = 0.0f;
  for (int pos = z; pos <= Lr; pos += 64) accp[pos] = 0.0f;
  if (acc_lds) for (int pos = z; pos < Lr; pos += 64) rl[pos] = dsq[pos];
  __syncthreads();
  uint32_t rng = md_mix3(seed, 87654321u, 12345678u);
  if (rng == 0) rng = 42;
  // lane l's jump: l + 1 draws of x -> 69069 x + 1 are x -> jA x + jC
  uint32_t jA = 1u, jC = 0u;
  for (int l = 0; l <= z; l++) { jA *= 69069u; jC = jC * 69069u + 1u; }
  int nseg = 0;
  const int64_t max_steps = 64 * ((int64_t) Lr + M + 16);
  for (int t = 0; t < nsamples && ! status; t++) {
    int s0 = MS_C, i = Lr, k = 0, hi = Lr, nd = 0;
    int sqfrom = 0, sqto = 0, hmmfrom = 0, hmmto = 0, Ld = 0, kmin = 0x7fffffff, kmax = 0;
    int64_t steps = 0;
    while (s0 != MS_S) {
      int s1 = -1;
      bool back = false;                                   // s1 (I or D) steps back to M at once
      if (s0 == MS_M || s0 == MS_D || s0 == MS_I || s0 == MS_C || s0 == MS_J) {
        // A state that can repeat: M along its diagonal, D along its row, I,
        // C and J down their column. Lane l takes the step that follows l
        // repeats, with the (l+1)-th draw from here and that cell's
        // thresholds; the first lane whose step leaves the state (or fails)
        // is the step the walk takes after those repeats.
        const uint32_t xl = jA * rng + jC;
        const double roll = (double) xl / 4294967296.0;
        const int li = (s0 == MS_D) ? i : i - z, lk = (s0 == MS_M) ? k - z : (s0 == MS_D) ? k - z : k;
        int sl = -1;
        // An M run's step to I or D is most often followed by one step back
        // to M. Lane l also takes that step ahead, as if it stopped here:
        // the I or D cell is lane l+1's, whose thresholds that lane loads
        // anyway, and the draw the next one. spec: that step leads to M
        // (and its M cell is inside the matrix).
        bool spec = false;
        if (s0 == MS_C || s0 == MS_J) {
          if (li >= 1) {
            const double th = xthr[(size_t) li * 3 + (s0 == MS_J)];
            sl = (roll < th) ? s0 : (th == th) ? MS_E : -1;
          }
        } else if (s0 == MS_M) {
          const bool in = (li >= 1 && lk >= 1);
          double t0 = 0.0, t1 = 0.0, t2 = 0.0, t3 = 0.0, t4 = 0.0;
          if (in) {
            const double *th = thr + ((size_t) li * (M + 1) + lk) * 5;
            t0 = th[0]; t1 = th[1]; t2 = th[2]; t3 = th[3]; t4 = th[4];
            sl = (roll < t0) ? MS_B : (roll < t1) ? MS_M : (roll < t2) ? MS_I : (t0 == t0) ? MS_D : -1;
          }
          const double n3 = __shfl_down(t3, 1, 64), n4 = __shfl_down(t4, 1, 64);   // cell (li - 1, lk - 1)
          const bool nin = __shfl_down((int) in, 1, 64) != 0 && z < 63;
          if (nin && (sl == MS_I || sl == MS_D)) {
            const double tt = (sl == MS_I) ? n4 : n3;
            const double roll2 = (double) (69069u * xl + 1u) / 4294967296.0;
            // back to M at (li - 2, lk - 1) from I, (li - 1, lk - 2) from D
            spec = (roll2 < tt) && ((sl == MS_I) ? li - 2 >= 1 : lk - 2 >= 1);
          }
        } else if (li >= 1 && lk >= 1) {
          const double *th = thr + ((size_t) li * (M + 1) + lk) * 5;
          const double tt = th[(s0 == MS_D) ? 3 : 4];
          sl = (roll < tt) ? MS_M : (tt == tt) ? s0 : -1;
        }
        const unsigned long long stop = __ballot(sl != s0);
        const int f = stop ? (int) __ffsll((long long) stop) - 1 : 64;
        if (f > 0) {                                       // the repeats: M and I states emit
          if (s0 == MS_M) {
            if (z < f) atomicAdd(&cnt[cnt_at(k - 1 - z)], 1.0f);
            kmin = min(kmin, k - f); kmax = max(kmax, k - 1);
            Ld += f;
            if (sqto == 0) { sqto = i - 1; hmmto = k - 1; }
            sqfrom = i - f; hmmfrom = k - f;
            i -= f; k -= f;
          } else if (s0 == MS_I) {
            if (z == 0) atomicAdd(&cnt[cnt_at(k)], (float) f);
            kmin = min(kmin, k); kmax = max(kmax, k);
            Ld += f;
            i -= f;
          } else if (s0 == MS_D) k -= f;
          else i -= f;
        }
        steps += f;
        if (f == 64) {
          rng = (uint32_t) __builtin_amdgcn_readlane((int) xl, 63);
          if (steps > max_steps) { status = 1; break; }
          continue;
        }
        rng = (uint32_t) __builtin_amdgcn_readlane((int) xl, f);
        s1 = __builtin_amdgcn_readlane(sl, f);
        back = (s0 == MS_M) && __builtin_amdgcn_readlane((int) spec, f) != 0;
        if (++steps > max_steps) { status = 1; break; }
      } else {
        if (++steps > max_steps) { status = 1; break; }
        switch (s0) {
        case MS_N: s1 = MS_S; i = 0; break;                // N repeats down to row 0 without draws
        case MS_B: {
          rng = rng * 69069u + 1u;
          const double roll = (double) rng / 4294967296.0, th = xthr[(size_t) i * 3 + 2];
          s1 = (roll < th) ? MS_N : (th == th) ? MS_J : -1;
          break;
        }
        case MS_E: {                                       // one roll; the first running sum above it
          if (i < 1) break;
          rng = rng * 69069u + 1u;
          const double roll = (double) rng / 4294967296.0;
          // the row's running sums are kept at every 64th element (ecum,
          // md_ncp per row): the first checkpoint above the roll names the
          // chunk, whose sums are formed again from the checkpoint before it
          const int NCP = md_ncp(Q);
          const double *cp = ecum + (size_t) i * NCP;
          int c = -1;
          for (int c0 = 0; c0 < NCP && c < 0; c0 += 64) {
            const int cc = c0 + z;
            const double v = cp[min(cc, NCP - 1)];
            const unsigned long long m = __ballot(cc < NCP && roll < v);
            if (m) c = c0 + (int) __ffsll((long long) m) - 1;
          }
          int found = -1;
          const float nf = (float) (1.0 / (double) xmx[(size_t) i * 6 + 0]);
          if (c >= 0) {
            const int j = 64 * c + z, q = j >> 3;
            const float u = (j < 8 * Q) ? __fmul_rn(mat[((size_t) i * Q + q) * 12 + (j & 7)], nf) : 0.0f;
            double sum = c ? cp[c - 1] : 0.0;
            for (int l = 0; l < 64; l++) {
              sum = sum + (double) readlanef(u, l);
              if (roll < sum) { found = 64 * c + l; break; }
            }
          }
          if (found >= 0) {
            k = (found & 3) * Q + (found >> 3) + 1;
            s1 = ((found >> 2) & 1) ? MS_D : MS_M;
          } else {
            // the reference scans the row again, the sum going on from its total
            double sum = cp[NCP - 1];
            for (int pass = 1; pass < 4 && s1 < 0; pass++)
              for (int q0 = 0; q0 < Q && s1 < 0; q0 += 8) {
                const int q = q0 + (z >> 3), sm = (z >> 2) & 1, r = z & 3;
                const float u = (q < Q) ? __fmul_rn(CELL(i, q, sm, r), nf) : 0.0f;
                for (int l = 0; l < 64; l++) {
                  sum = sum + (double) readlanef(u, l);
                  if (roll < sum) { k = (l & 3) * Q + q0 + (l >> 3) + 1; s1 = ((l >> 2) & 1) ? MS_D : MS_M; break; }
                }
              }
          }
          break;
        }
        }
      }
      if (s1 < 0) { status = 1; break; }
      if (s0 == MS_M) { k--; i--; }
      else if (s0 == MS_D) k--;
      else if (s0 == MS_I) i--;
      // the trace element s1 at (i, k), in backward order
      if (s1 == MS_M || s1 == MS_I) {
        if (k < 1 || k > M || i < 1) { status = 1; break; }
        if (z == 0) atomicAdd(&cnt[cnt_at(k)], 1.0f);      // the node's M cell
        kmin = min(kmin, k); kmax = max(kmax, k);
        Ld++;
        if (s1 == MS_M) {
          if (sqto == 0) { sqto = i; hmmto = k; }
          sqfrom = i; hmmfrom = k;
        }
      } else if (s1 == MS_D) {
        if (k < 1 || k > M || i < 1) { status = 1; break; }
      } else if (s1 == MS_E) {
        sqfrom = sqto = hmmfrom = hmmto = 0; Ld = 0; kmin = 0x7fffffff; kmax = 0;
      } else if (s1 == MS_B) {
        // p7_Null2_ByTrace over the domain: four striped sums over q per
        // canonical residue, their hsum; degenerate codes the mean
        md_sync();
        const float nrm = (float) (1.0 / (double) (float) Ld);
        // lane z sums x = (base + z) / 4 and, for base + 64 + z, x + 16 (the
        // same lane r, so the same counts): two chains per lane, one pass
        for (int base = 0; base < p.K * 4; base += 128) {
          const int xr = base + z, x = xr >> 2, r = xr & 3, x2 = x + 16;
          const bool h1 = x < p.K, h2 = x2 < p.K;
          float sv = 0.0f, sv2 = 0.0f;
          if (h1) {
            // positions kmin..kmax in lane r: the other cells' counts are 0, and
            // adding their products (+0) leaves the striped sum unchanged
            const int q0 = max(0, kmin - 1 - r * Q), q1 = min(Q - 1, kmax - 1 - r * Q);
            for (int qa = q0; qa <= q1; qa += 8) {
              float e[8], e2[8], cq[8];
#pragma unroll
              for (int u = 0; u < 8; u++) cq[u] = cnt[min(qa + u, Q - 1) * 4 + r];
              if (em_lds) {
#pragma unroll
                for (int u = 0; u < 8; u++) {
                  e[u] = em[(x * Q + min(qa + u, Q - 1)) * 4 + r];
                  e2[u] = h2 ? em[(x2 * Q + min(qa + u, Q - 1)) * 4 + r] : 0.0f;
                }
              } else {
#pragma unroll
                for (int u = 0; u < 8; u++) {
                  e[u] = md_tv(p, x, min(r * Q + qa + u + 1, M + 1));
                  e2[u] = h2 ? md_tv(p, x2, min(r * Q + qa + u + 1, M + 1)) : 0.0f;
                }
              }
              // a zero count adds +0.0, which leaves the non-negative sums as
              // they are: no branch per cell
#pragma unroll
              for (int u = 0; u < 8; u++) {
                const float w = __fmul_rn(qa + u <= q1 ? cq[u] : 0.0f, nrm);
                sv = __fadd_rn(sv, __fmul_rn(w, e[u]));
                sv2 = __fadd_rn(sv2, __fmul_rn(w, e2[u]));
              }
            }
          }
          // lanes 4x..4x+3 hold x's lane sums
          const float s01 = __fadd_rn(sv, __shfl_down(sv, 1, 4)), s23 = __shfl_down(s01, 2, 4);
          if (h1 && r == 0) n2v[x] = __fadd_rn(__fadd_rn(s01, s23), 0.0f);
          const float t01 = __fadd_rn(sv2, __shfl_down(sv2, 1, 4)), t23 = __shfl_down(t01, 2, 4);
          if (h2 && r == 0) n2v[x2] = __fadd_rn(__fadd_rn(t01, t23), 0.0f);
        }
        md_sync();
        {                                                  // degenerate codes: one lane each
          const int xx = p.K + 1 + z;
          if (xx <= p.Kp - 3) {
            const uint32_t dm = p.degen[xx];
            float rr = 0.0f;
            for (int y = 0; y < p.K; y++) if ((dm >> y) & 1u) rr = __fadd_rn(rr, n2v[y]);
            n2v[xx] = __fdiv_rn(rr, (float) (dm >> 24));
          }
          if (z == 0) { n2v[p.K] = 1.0f; n2v[p.Kp - 2] = 1.0f; n2v[p.Kp - 1] = 1.0f; }
        }
        md_sync();
        // residues (sqto, hi] are outside domains (1), (sqfrom, sqto] take the odds
        for (int pos = z + 1 + ((sqfrom >> 6) << 6); pos <= hi; pos += 64) {
          // the lane of a residue is fixed (pos mod 64), so its updates stay in order
          if (pos > sqto)        accp[pos] = __fadd_rn(accp[pos], 1.0f);
          else if (pos > sqfrom) accp[pos] = __fadd_rn(accp[pos], n2v[rsq[pos - 1]]);
        }
        hi = sqfrom;
        if (nd >= MD_TSEG) { status = 5; break; }
        if (z == 0) { tseg[nd*4 + 0] = sqfrom; tseg[nd*4 + 1] = sqto; tseg[nd*4 + 2] = hmmfrom; tseg[nd*4 + 3] = hmmto; }
        nd++;
        for (int c = z; c < 4 * Q; c += 64) cnt[c] = 0.0f;
        kmin = 0x7fffffff; kmax = 0;
        md_sync();
      }
      // the I or D state's step back to M, taken ahead in the M run's ballot:
      // its draw, then the M element at (i - 1, k) from I, (i, k - 1) from D
      if (back) {
        rng = rng * 69069u + 1u;
        if (++steps > max_steps) { status = 1; break; }
        if (s1 == MS_I) i--; else k--;
        if (k < 1 || k > M || i < 1) { status = 1; break; }
        if (z == 0) atomicAdd(&cnt[cnt_at(k)], 1.0f);
        kmin = min(kmin, k); kmax = max(kmax, k);
        Ld++;
        if (sqto == 0) { sqto = i; hmmto = k; }
        sqfrom = i; hmmfrom = k;
        s1 = MS_M;
      }
      // (N, C and J emit on the transition to themselves: the repeats above)
      s0 = s1;
    }
    if (status) break;
    for (int pos = z + 1; pos <= hi; pos += 64) accp[pos] = __fadd_rn(accp[pos], 1.0f);
    // the trace's domains join the ensemble in sequence order
    if (nseg + nd > J.segcap) { status = 5; break; }
    md_sync();
    for (int d = z; d < nd; d += 64) {
      const int e = nd - 1 - d;
      segs[nseg + d] = MdSeg{ t, tseg[e*4 + 0], tseg[e*4 + 1], tseg[e*4 + 2], tseg[e*4 + 3] };
    }
    nseg += nd;
    md_sync();
  }
  if (status) { if (z == 0) { O->status = status; O->nc = 0; O->nseg = nseg; } return; }
  for (int pos = z + 1; pos <= Lr; pos += 64) acc[pos] = logf(__fdiv_rn(accp[pos], (float) nsamples));
  __syncthreads();

  // ---------------- 3. clustering ----------------
  // the two stacks of the linkage in LDS; each element's cluster and each
  // cluster's trace count in the region's area
  int *av = (int *) lds, *bv = av + J.segcap;
  int *asg = (int *) (pool + J.area + md_clu_at(M, Lr, J.segcap)), *ninc = asg + J.segcap;
  unsigned long long *lk = (unsigned long long *) (bv + J.segcap);   // link bits of one pass
  MdSig *sig = (MdSig *) (lk + (J.segcap + 63) / 64);
  const int n = nseg;
  for (int v = z; v < n; v += 64) { av[v] = n - v - 1; ninc[v] = 0; }
  __syncthreads();
  int na = n, nb = 0, ncl = 0;
  while (na > 0) {
    na--;
    if (z == 0) bv[nb] = av[na];
    nb++;
    __syncthreads();
    while (nb > 0) {
      nb--;
      const int v = bv[nb];
      if (z == 0) asg[v] = ncl;
      const MdSeg sv = segs[v];
      const int nw = (na + 63) / 64;
      for (int w = 0; w < nw; w++) {
        const int a = w * 64 + z;
        const bool l = (a < na) && md_linked(sv, segs[av[a]]);
        const unsigned long long m = __ballot(l);
        if (z == 0) lk[w] = m;
      }
      __syncthreads();
      // the linked slots, highest first: each leaves by taking the top's element
      for (int w = nw - 1; w >= 0; w--) {
        unsigned long long m = lk[w];
        while (m) {
          const int bit = 63 - __clzll((long long) m);
          m &= ~(1ull << bit);
          const int a = w * 64 + bit;
          if (z == 0) { const int we = av[a]; av[a] = av[na - 1]; bv[nb] = we; }
          na--; nb++;
          __syncthreads();
        }
      }
      __syncthreads();
    }
    ncl++;
  }
  __syncthreads();
  // posterior of each cluster: the traces with a member (members of one trace are adjacent)
  for (int h = z; h < n; h += 64) {
    bool first = true;
    for (int h2 = h - 1; h2 >= 0 && segs[h2].idx == segs[h].idx; h2--) if (asg[h2] == asg[h]) first = false;
    if (first) atomicAdd(&ninc[asg[h]], 1);
  }
  __syncthreads();
  int nsig = 0;
  for (int c = 0; c < ncl; c++) {
    const int ni = __hip_atomic_load(&ninc[c], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);   // counted at L2
    if (__fdiv_rn((float) ni, (float) nsamples) < 0.25f) continue;
    int lo[4], hi4[4];
    for (int e = 0; e < 4; e++) { lo[e] = 0x7fffffff; hi4[e] = -1; }
    for (int h = z; h < n; h += 64) if (asg[h] == c) {
      const int v4[4] = { segs[h].i, segs[h].j, segs[h].k, segs[h].m };
      for (int e = 0; e < 4; e++) { lo[e] = min(lo[e], v4[e]); hi4[e] = max(hi4[e], v4[e]); }
    }
    for (int e = 0; e < 4; e++) { lo[e] = -wave_max_i(-lo[e]); hi4[e] = wave_max_i(hi4[e]); }
    const int thr = (int) ceilf(__fmul_rn((float) ni, 0.02f));
    int best[4];
    for (int e = 0; e < 4; e++) {                          // i, j, k, m: leftmost start, rightmost end with support
      const int w = hi4[e] - lo[e] + 1;
      for (int b = z; b < w; b += 64) hist[b] = 0;
      __syncthreads();
      for (int h = z; h < n; h += 64) if (asg[h] == c) {
        const int v = (e == 0) ? segs[h].i : (e == 1) ? segs[h].j : (e == 2) ? segs[h].k : segs[h].m;
        atomicAdd(&hist[v - lo[e]], 1);
      }
      __syncthreads();
      // the counts were made by atomics at L2: read them there too
      auto hcount = [&](int o) { return __hip_atomic_load(&hist[o], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); };
      int b = -1;
      const bool up = (e == 0 || e == 2);
      for (int base = 0; base < w && b < 0; base += 64) {
        const int o = up ? base + z : w - 1 - base - z;
        const unsigned long long m = __ballot(o >= 0 && o < w && hcount(o) >= thr);
        if (m) b = up ? base + (int) __ffsll((long long) m) - 1 : w - 1 - base - ((int) __ffsll((long long) m) - 1);
      }
      if (b < 0) {                                         // esl_vec_IArgMax: the first maximum
        int bc = -1, bi = 0x7fffffff;
        for (int o = z; o < w; o += 64) { const int h = hcount(o); if (h > bc) { bc = h; bi = o; } }
        const int mc = wave_max_i(bc);
        b = -wave_max_i(bc == mc ? -bi : -0x7fffffff);
      }
      best[e] = lo[e] + b;
      __syncthreads();
    }
    if (best[0] > best[1] || best[2] > best[3]) continue;
    if (nsig >= MD_MAXSIG) { status = 5; break; }
    if (z == 0) sig[nsig] = MdSig{ c, best[0], best[1], best[2], best[3], __fdiv_rn((float) ni, (float) nsamples) };
    nsig++;
  }
  __syncthreads();
  if (status) { if (z == 0) { O->status = status; O->nc = 0; O->nseg = nseg; } return; }
  if (z == 0) {
    for (int a = 1; a < nsig; a++) {                       // start order; equal starts keep their order
      const MdSig v = sig[a];
      int b = a - 1;
      while (b >= 0 && sig[b].i > v.i) { sig[b + 1] = sig[b]; b--; }
      sig[b + 1] = v;
    }
    int dom[MD_MAXSIG];
    for (int d = 0; d < nsig; d++) dom[d] = 0;
    for (int d = 0; d < nsig; d++)
      for (int d2 = d + 1; d2 < nsig; d2++) {
        const int nov = min(sig[d].j, sig[d2].j) - max(sig[d].i, sig[d2].i) + 1;
        if (nov == 0) break;
        const int nn = min(sig[d].j - sig[d].i + 1, sig[d2].j - sig[d2].i + 1);
        if ((double) __fdiv_rn((float) nov, (float) nn) >= 0.8) {
          if (sig[d].prob > sig[d2].prob) dom[d2] = 1; else dom[d] = 1;
        }
      }
    int nc = 0;
    for (int d = 0; d < nsig; d++) {
      if (dom[d]) continue;
      if (nc < MD_MAXC) { O->ci[nc] = sig[d].i; O->cj[nc] = sig[d].j; }
      nc++;
    }
    O->status = (nc > MD_MAXC) ? 5 : 0;
    O->nc = min(nc, MD_MAXC);
    O->nseg = nseg;
  }
#undef CELL
}

// LDS of one region's workgroup: the Forward's two rows and D->D odds, the
// traces' counts, odds and per-trace domains, or the clustering's stacks
size_t md_lds_bytes(int M, int segcap, int acc_cap)
{
  const int64_t fwd = (md_stage_t(M) ? md_stage_floats(M) : 0) + 32 * md_q(M) + 128;
  const size_t clu = (size_t) 2 * segcap * 4 + (size_t) (segcap + 63) / 64 * 8 + (size_t) MD_MAXSIG * sizeof(MdSig);
  return std::max((size_t) std::max(fwd, md_trace_floats(M, acc_cap)) * 4, clu);
}

// The traces' cell thresholds of every region (md_thresholds over each cell
// of its finished Forward matrix): a grid of gridDim.x workgroups of
// MD_THR_NT threads per region (blockIdx.y). Every choice a trace can make
// depends only on the finished matrix, so its thresholds are computed first,
// with the same operations p7_StochasticTrace performs at the step
// (md_thresholds); a step is then one draw and one to three comparisons.
// thr[(i, k)]: M's three (B | M | I | D), D's one (M | D), I's one (M | I),
// as doubles. A pass's five-double records go out through LDS, so that the
// stores of a wave are contiguous. The rows' thresholds are k_md_rows'.
#define MD_THR_NT 256
#define MD_THR_CELLS 2                                     // cells per thread per pass
#define MD_THR_LDS_OUT ((size_t) 8 * (MD_THR_CELLS * MD_THR_NT * 5 > MD_THR_NT * 8 ? MD_THR_CELLS * MD_THR_NT * 5 : MD_THR_NT * 8))
static __host__ __device__ inline size_t md_thr_tl_bytes(int M)      // the staged transitions, when they fit
{
  const size_t b = ((size_t) 10 * (M + 2) * 4 + 7) & ~(size_t) 7;
  return (b + MD_THR_LDS_OUT <= P7G_LDS_MAX) ? b : 0;
}

__global__ __launch_bounds__(MD_THR_NT) void
k_md_thr(const DevProfile p, const MdJob *__restrict__ jobs, float *__restrict__ pool)
{
  const int t = threadIdx.x, NT = MD_THR_NT;
  const MdJob J = jobs[blockIdx.y];
  const int M = p.M, Q = max(2, (M - 1) / 4 + 1), Lr = J.Lr;
  const int TR = p.nrows;
  float *mat = pool + J.area;
  float *xmx = mat + (size_t) (Lr + 1) * 12 * Q;
  extern __shared__ double ldsd[];
  const int TW = M + 2;
  const size_t tlb = md_thr_tl_bytes(M);
  float *tl = (float *) ldsd;
  double *ob = ldsd + tlb / 8;                             // the outgoing values
  if (tlb) {
    for (int c = t; c < 10 * TW; c += NT) tl[c] = md_tv(p, TR + md_trow(c / TW), c % TW);
    __syncthreads();
  }
  auto tv = [&](int row, int k) -> float { return tlb ? tl[row * TW + k] : md_tv(p, TR + md_trow(row), k); };
  enum { T_BM = 0, T_MM, T_IM, T_DM, T_MDS, T_DDS, T_MI, T_II, T_MD, T_DD };
#define CELL(i_, q_, s_, r_) mat[(((size_t) (i_) * Q + (q_)) * 3 + (s_)) * 4 + (r_)]
  double *thr = (double *) (pool + J.area + md_thr_at(M, Lr, J.segcap));
  auto cellk = [&](int i, int k, int s) {                  // the cell of model position k (1..4Q) in row i
    const int r = (k - 1 >= Q) + (k - 1 >= 2 * Q) + (k - 1 >= 3 * Q), q = k - 1 - r * Q;
    return CELL(i, q, s, r); };
  // cells c = (i - 1) M + (k - 1): each pass MD_THR_CELLS * NT of them, their
  // matrix values loaded before the threshold arithmetic of any of them
  const int64_t LM = (int64_t) Lr * M;
  constexpr int CU = MD_THR_CELLS;
  // a cell's row and column from the pass's first cell (one 64-bit division
  // per pass) and its offset o < CU * NT: x = r0 + o < M + CU * NT, divided
  // by M through a float reciprocal and corrected by one
  const float invM = 1.0f / (float) M;
  for (int64_t c0 = (int64_t) blockIdx.x * CU * NT; c0 < LM; c0 += (int64_t) gridDim.x * CU * NT) {
    const int a0 = (int) (c0 / M), r0 = (int) (c0 - (int64_t) a0 * M);
    const int olast = (int) min((int64_t) CU * NT - 1, LM - 1 - c0);
    auto rowcol = [&](int o, int &i, int &k) {
      const int x = r0 + o;
      int qd = (int) ((float) x * invM);
      if (qd * M > x) qd--; else if ((qd + 1) * M <= x) qd++;
      i = 1 + a0 + qd; k = 1 + x - qd * M;
    };
    float in[CU][8];
#pragma unroll
    for (int u = 0; u < CU; u++) {
      int i, k;
      rowcol(min(u * NT + t, olast), i, k);
      in[u][0] = xmx[(size_t) (i - 1) * 6 + 3];
      in[u][1] = k > 1 ? cellk(i - 1, k - 1, 0) : 0.0f;
      in[u][2] = k > 1 ? cellk(i - 1, k - 1, 2) : 0.0f;
      in[u][3] = k > 1 ? cellk(i - 1, k - 1, 1) : 0.0f;
      in[u][4] = k > 1 ? cellk(i, k - 1, 0) : 0.0f;
      in[u][5] = k > 1 ? cellk(i, k - 1, 1) : 0.0f;
      in[u][6] = cellk(i - 1, k, 0);
      in[u][7] = cellk(i - 1, k, 2);
    }
#pragma unroll
    for (int u = 0; u < CU; u++) {
      int i, k;
      rowcol(min(u * NT + t, olast), i, k);
      double *o = ob + (size_t) (u * NT + t) * 5;
      float v[4];
      v[0] = __fmul_rn(in[u][0], tv(T_BM, k));
      v[1] = __fmul_rn(in[u][1], tv(T_MM, k));
      v[2] = __fmul_rn(in[u][2], tv(T_IM, k));
      v[3] = __fmul_rn(in[u][3], tv(T_DM, k));
      md_thresholds(v, 4, o);
      v[0] = (k > 1) ? __fmul_rn(in[u][4], tv(T_MDS, k)) : 0.0f;
      v[1] = (k > 1) ? __fmul_rn(in[u][5], tv(T_DDS, k)) : 0.0f;
      md_thresholds(v, 2, o + 3);
      v[0] = __fmul_rn(in[u][6], tv(T_MI, k));
      v[1] = __fmul_rn(in[u][7], tv(T_II, k));
      md_thresholds(v, 2, o + 4);
    }
    __syncthreads();
    // the pass's cells are consecutive in row-major order, each row's cells at
    // (i (M + 1) + k) 5: the stores run along them
    for (int d = t; d < CU * NT * 5; d += NT) {
      const int o = d / 5;
      if (o <= olast) {
        int i, k;
        rowcol(o, i, k);
        thr[((size_t) i * (M + 1) + k) * 5 + (d - o * 5)] = ob[d];
      }
    }
    __syncthreads();
  }
#undef CELL
}

// The rows' thresholds of every region: C's (C | E), J's (J | E) and B's
// (N | J), and the E state's running-sum checkpoints (every 64th cell and
// the last, md_ncp per row). One wave per 64 rows (blockIdx.x) of a region
// (blockIdx.y). A row's sum is a chain of 8Q double additions that its lane
// runs; the wave loads the 64 rows' M and D values eight q at a time into
// LDS, so that its loads run along the rows.
__global__ __launch_bounds__(64) void
k_md_rows(const DevProfile p, const MdJob *__restrict__ jobs, float *__restrict__ pool)
{
  const int t = threadIdx.x;
  const MdJob J = jobs[blockIdx.y];
  const int M = p.M, Q = max(2, (M - 1) / 4 + 1), Lr = J.Lr, L = J.Ltarget;
  const int r0 = blockIdx.x * 64;
  if (r0 > Lr) return;
  const float xEmove = p.xfE_move, xEloop = p.xfE_loop, floop = p.xf_loop[L], fmove = p.xf_move[L];
  const float *mat = pool + J.area;
  const float *xmx = mat + (size_t) (Lr + 1) * 12 * Q;
  double *thr = (double *) (pool + J.area + md_thr_at(M, Lr, J.segcap));
  double *xthr = thr + (size_t) 5 * (Lr + 1) * (M + 1);
  double *ecum = xthr + (size_t) 3 * (Lr + 1);
  __shared__ float4 tile[64 * 17];                         // [row][q][M, D halves], 17 float4 a row
  const int i = r0 + t;
  const bool row = (i >= 1 && i <= Lr);
  float nf = 0.0f;
  if (i <= Lr) {
    float v[2];
    if (i >= 1) {
      v[0] = __fmul_rn(xmx[(size_t) (i - 1) * 6 + 4], floop);
      v[1] = __fmul_rn(__fmul_rn(xmx[(size_t) i * 6 + 0], xEmove), xmx[(size_t) i * 6 + 5]);
      md_thresholds(v, 2, xthr + (size_t) i * 3 + 0);
      v[0] = __fmul_rn(xmx[(size_t) (i - 1) * 6 + 2], floop);
      v[1] = __fmul_rn(__fmul_rn(xmx[(size_t) i * 6 + 0], xEloop), xmx[(size_t) i * 6 + 5]);
      md_thresholds(v, 2, xthr + (size_t) i * 3 + 1);
      // select_e: norm = 1/E(i) in double, the products in float, summed in double
      nf = (float) (1.0 / (double) xmx[(size_t) i * 6 + 0]);
    }
    v[0] = __fmul_rn(xmx[(size_t) i * 6 + 1], fmove);
    v[1] = __fmul_rn(xmx[(size_t) i * 6 + 2], fmove);
    md_thresholds(v, 2, xthr + (size_t) i * 3 + 2);
  }
  const int NCP = md_ncp(Q);
  double sum = 0.0;
  for (int q0 = 0; q0 < Q; q0 += 8) {
    // lane t loads parts j = t + 64 m: row j / 16, q q0 + (j % 16) / 2, half j % 2
#pragma unroll
    for (int m = 0; m < 16; m++) {
      const int j = t + 64 * m, rw = j >> 4, part = j & 15;
      const int ii = min(r0 + rw, Lr), q = min(q0 + (part >> 1), Q - 1);   // past the ends: a valid cell, unused
      tile[rw * 17 + part] = *(const float4 *) (mat + ((size_t) ii * Q + q) * 12 + (part & 1) * 4);
    }
    __syncthreads();
    if (row) {
#pragma unroll
      for (int a = 0; a < 8; a++)
        if (q0 + a < Q) {
          const float4 x0 = tile[t * 17 + 2 * a], x1 = tile[t * 17 + 2 * a + 1];
          sum = sum + (double) __fmul_rn(x0.x, nf); sum = sum + (double) __fmul_rn(x0.y, nf);
          sum = sum + (double) __fmul_rn(x0.z, nf); sum = sum + (double) __fmul_rn(x0.w, nf);
          sum = sum + (double) __fmul_rn(x1.x, nf); sum = sum + (double) __fmul_rn(x1.y, nf);
          sum = sum + (double) __fmul_rn(x1.z, nf); sum = sum + (double) __fmul_rn(x1.w, nf);
        }
      ecum[(size_t) i * NCP + (q0 >> 3)] = sum;                 // after element min(8 (q0 + 8), 8Q) - 1
    }
    __syncthreads();
  }
}

// floats of one region's area in the pool: matrix, specials, residue sums,
// endpoint counts, the ensemble, then the traces' thresholds (doubles)
int64_t md_area_floats(int M, int Lr, int segcap)
{
  const int64_t Q = md_q(M), R = (int64_t) Lr + 1;
  const int64_t f = md_thr_at(M, Lr, segcap) + 2 * (5 * R * ((int64_t) M + 1) + 3 * R + md_ncp(Q) * R);
  return (f + 63) & ~(int64_t) 63;
}

hipError_t launch_md_regions(const DevProfile &p, const uint8_t *res, const MdJob *jobs, MdOut *out, int njobs,
                             float *pool, int segcap, int nsamples, uint32_t seed, int max_lr, hipStream_t st,
                             hipEvent_t thr_done)
{
  if (njobs <= 0) return hipSuccess;
  const int acc_cap = md_acc_cap(max_lr);
  const size_t lds = md_lds_bytes(p.M, segcap, acc_cap);
  if (lds > P7G_LDS_MAX) return hipErrorInvalidValue;
  hipError_t e = hipFuncSetAttribute((const void *) k_md_region, hipFuncAttributeMaxDynamicSharedMemorySize, (int) lds);
  if (e != hipSuccess) return e;
  // the Forward matrices (one wave per region, the Forward's LDS only), the
  // thresholds (the whole device), then the traces and the clustering
  // (128 floats past the rows: the D->D chains read up to seven cells past q = Q - 1)
  const size_t lds0 = (size_t) ((md_stage_t(p.M) ? md_stage_floats(p.M) : 0) + 32 * md_q(p.M) + 128) * 4;
  hipLaunchKernelGGL(k_md_region, dim3(njobs), dim3(64), lds0, st, p, res, jobs, out, pool, nsamples, seed, 0, acc_cap);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  for (int j0 = 0; j0 < njobs; j0 += 65535) {
    const int nj = std::min(njobs - j0, 65535);
    hipLaunchKernelGGL(k_md_thr, dim3(32, nj), dim3(MD_THR_NT), md_thr_tl_bytes(p.M) + MD_THR_LDS_OUT, st, p, jobs + j0, pool);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(k_md_rows, dim3((max_lr + 1 + 63) / 64, nj), dim3(64), 0, st, p, jobs + j0, pool);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  if (thr_done && (e = hipEventRecord(thr_done, st)) != hipSuccess) return e;
  hipLaunchKernelGGL(k_md_region, dim3(njobs), dim3(64), lds, st, p, res, jobs, out, pool, nsamples, seed, 1, acc_cap);
  return hipGetLastError();
}

// Residue copies of the envelopes the clusters define, from their regions'
// copies: one block per envelope, padded to 16 bytes plus 16 with the pad code
// as the region copies are.
__global__ void k_md_copy(uint8_t *__restrict__ res, const int64_t *__restrict__ src, const int64_t *__restrict__ dst,
                          const int32_t *__restrict__ len, int n)
{
  const int e = blockIdx.x;
  if (e >= n) return;
  const int Ld = len[e];
  const int64_t need = (((int64_t) Ld + 15) & ~(int64_t) 15) + 16;
  for (int64_t b = threadIdx.x; b < need; b += blockDim.x)
    res[dst[e] + b] = (b < Ld) ? res[src[e] + b] : (uint8_t) P7G_PADCODE;
}

// The regions' residue null2 scores (their areas' residues 1..Lr), gathered
// in job order for one copy to the host: one workgroup per region.
__global__ void k_md_n2_gather(const MdJob *__restrict__ jobs, const float *__restrict__ pool, int M,
                               float *__restrict__ dst)
{
  const MdJob J = jobs[blockIdx.x];
  const float *src = pool + J.area + md_n2sc_at(M, J.Lr) + 1;
  for (int i = threadIdx.x; i < J.Lr; i += blockDim.x) dst[(size_t) J.n2o + i] = src[i];
}

hipError_t launch_md_n2_gather(const MdJob *jobs, int njobs, const float *pool, int M, float *dst, hipStream_t st)
{
  if (njobs <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_md_n2_gather, dim3(njobs), dim3(256), 0, st, jobs, pool, M, dst);
  return hipGetLastError();
}

hipError_t launch_md_copy(uint8_t *res, const int64_t *src, const int64_t *dst, const int32_t *len, int n, hipStream_t st)
{
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_md_copy, dim3(n), dim3(256), 0, st, res, src, dst, len, n);
  return hipGetLastError();
}
