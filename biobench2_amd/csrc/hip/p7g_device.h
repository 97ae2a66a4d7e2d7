// Device-side data structures shared by the kernels and the runtime.
#pragma once
#include <stdint.h>
#include <stddef.h>

// Residue code used for rows past the end of a target and for the tail
// padding of each sequence in the HBM layout: every score against it is
// "impossible", so extra rows never change a result.
#define P7G_PADCODE 31
// Residue rows of every device score table: codes 0..Kp-1 and the pad code.
#define P7G_NROWS 32

// Number of transition rows in the device Viterbi and Forward/Backward tables.
#define P7G_VF_NT 8    // BM MM IM DM (into k), MDs (M_{k-1}->D_k), MI II (out of k), DDs (D_{k-1}->D_k)
#define P7G_FB_NT 13   // the 8 above + MMn IMn DMn (into k+1), MD DD (out of k)
enum { VT_BM = 0, VT_MM, VT_IM, VT_DM, VT_MDS, VT_MI, VT_II, VT_DDS,
       FT_MMN = 8, FT_IMN, FT_DMN, FT_MD, FT_DD };

// One shape of the ungapped MSV pass: G lanes per target (64/G targets per
// wave), ND dwords (2*ND model positions) per lane. Its table holds int16
// pairs of (bias - rbv); per residue row, ND/4 quads [q][G] of int4, then
// ND%4 dwords [t][G]: [rows][G * ND] in all.
struct SsvVariant {
  int    G, ND;          // G == 0: the model does not fit this lane count
  int   *tbl;
  size_t dwords;
};
#define P7G_SSV_NVAR 4   // G = 16, 32, 64, 8
#define P7G_LDS_MAX (160 * 1024)   // LDS of one workgroup on gfx950
#define P7G_SSV_MAXCHUNK 3

struct DevProfile {
  int M, Kp;
  int K;                 // canonical residues (20 amino, 4 DNA)
  // per code x < 32: bit y set if x stands for canonical residue y, and the
  // number of such residues in bits 24..31 (Easel's degen/ndegen)
  const uint32_t *degen;
  // Every table is lane-minor: the values one read instruction fetches for
  // consecutive lanes are consecutive in memory, so LDS reads are free of
  // bank conflicts for any per-lane block size.
  // MSV stage, ungapped pass: the variants with 16, 32 and 64 lanes per target
  SsvVariant ssv[P7G_SSV_NVAR];
  int   ssv_def;         // variant used when no tuned choice applies; -1: chunked
  // models longer than one LDS table holds (M > 2432): two or three chunks of
  // 64 lanes per target, chunk c holding positions c*128*ND+1 .. (c+1)*128*ND
  int   ssv_nchunk;      // 0, 2 or 3
  SsvVariant ssvc[P7G_SSV_MAXCHUNK];
  // MSV stage, full recurrence with the J state: same scores, 64 lanes
  int   mf_ND;
  int  *mf_tbl;      size_t mf_dwords;      // [rows][ND][64]
  // Viterbi filter: int16 pairs, 64 lanes, ND dwords per lane
  int   vf_ND;
  int  *vf_tbl;      size_t vf_dwords;      // [(rows) + 8][ND][64]
  // Forward/Backward: float, 64 lanes, B positions per lane
  int   fb_B;
  float *fb_tbl;     size_t fb_floats;      // [(rows) + 13][B][64]
  int   nrows;       // Kp rounded up to include P7G_PADCODE (32)
  // per-length constants, L = 0..Lmax
  int      Lmax;
  uint8_t *tjb;
  int16_t *xw_move;
  float   *xf_move, *xf_loop, *nullsc, *flen_a, *flen_b;
  // bias-filter HMM odds, [32 codes][2 states]
  float   *eo;
  // scalars
  int   tbm, tec, base, bias;
  float scale_b;
  int   base_w, ddbound_w, xwE_move, xwE_loop, xwN_loop, xwC_loop, xwJ_loop;
  float scale_w;
  float xfE_move, xfE_loop;
  float nj;
  float ft1_0, ft1_1;     // filter HMM state-1 transitions
  float pi0, pi1;
  void *blob;             // single device allocation holding everything above
};

// A device-resident target set: residues grouped by length (longest first),
// each sequence starting on a 16-byte boundary and padded with P7G_PADCODE.
struct DevSeqs {
  const uint8_t *res;
  const int64_t *off;    // [slot]
  const int32_t *len;    // [slot]
  const int32_t *perm;   // [slot] -> the caller's index
  int64_t nslot;
  // the ungapped MSV pass cuts targets longer than P7G_SEG_ROWS into row
  // segments (slot, first row); they cover slots 0 .. nseg_slots-1
  const int2 *seg;
  int64_t nseg, nseg_slots;
};
#define P7G_SEG_ROWS 4096

// Per-slot results written by the stage kernels.
struct DevResults {
  uint16_t *zmax;      // SSV pass: best ungapped score above the B floor
  int16_t  *xj;        // MSV: final xJ, -1 on overflow
  float    *nullsc, *usc, *filtersc, *vfsc, *fwdsc, *bcksc;
  uint8_t  *stage;
  int      *vstatus;   // Viterbi status (0 ok, 16 overflow)
  int      *fstatus;   // Forward/Backward status
};

// Work lists: counts live in device memory so that stages chain without
// host round trips.
struct DevLists {
  int32_t *full, *bias, *vit, *fwd, *bck;
  // [0] full, [1] bias, [2] vit, [3] fwd, [4] bck: list lengths; [6]
  // scale-pool overflow flag; [7] domain-decoding overflow flags (1: row
  // pool, 2: region list, 4: region residues); the work-queue cursor of list
  // k sits at [k + 8]; [13] is the cursor of the ungapped pass over all targets
  // ([14], [15]: the later launches of a chunked model). Zeroed before every call.
  int32_t *counts;
  // [0] next free float in the Forward scale-factor pool, [1] in the domain
  // decoding pool, [2] next free byte of the region residue copies
  unsigned long long *cursor;
};

// Domain decoding buffers (pool == nullptr: off). The Forward parser keeps the
// special-state rows (E N J B C SCALE) of every target it scores at doff[slot];
// the Backward parser keeps its own rows, followed by btot, etot and mocc, at
// boff[slot] (9 (L+1) floats); -1 where the pool was full.
struct DomBufs {
  float   *pool = nullptr;
  int64_t  size = 0;
  int64_t *doff = nullptr, *boff = nullptr;
  // envelope scoring (the Forward parser over region copies, unihit): the
  // N/J/C loop and move odds of each region, and how many regions there are room for
  const float *env_loop = nullptr, *env_move = nullptr;
  int64_t  env_cap = 0;
  // null2 of single-domain regions (rescore_isolated_domain): each region's
  // area at n2pool + n2off[k] holds the Backward special rows and M and I
  // rows (n2_area_floats); its Forward row scales are at n2scl + n2soff[k]
  // (written by the envelope-score pass). n2mode 2: the Forward recomputed
  // with the posterior sums, null2 and the correction n2dc[k] (nats).
  // n2st[k]: the Backward's status and own-scales flag (0x100), then 0 or 16
  float   *n2pool = nullptr;
  const int64_t *n2off = nullptr;
  const float *n2scl = nullptr;
  const int64_t *n2soff = nullptr;
  float   *n2dc = nullptr;
  int32_t *n2st = nullptr;
  int      n2mode = 0;
  // optimal accuracy alignment of each region (the n2mode 2 pass fills the
  // traceback choices in its area and the OA score; k_oa_trace walks them):
  // n2oa[k] the score, n2ali[k] = {hmmfrom, hmmto, alifrom, alito} (1-based,
  // within the region copy; 0 where no trace)
  float   *n2oa = nullptr;
  int4    *n2ali = nullptr;
  // the alignments themselves (k_oa_trace): region k's states at
  // n2path + n2poff[k], at most its length + M + 2 words, n2pn[k] of them
  uint32_t *n2path = nullptr;
  const int64_t *n2poff = nullptr;
  int32_t *n2pn = nullptr;
};

// the parts of a region's null2 area, in floats, each a multiple of 64: the
// Backward special rows, the Backward M and I rows, then the optimal accuracy
// traceback's choices (n2_oa_words per row, words [row][w][64 lanes], four
// bits per model position) and one word per row for the special states
__host__ __device__ inline int64_t n2_xrows_floats(int L) { return (6 * ((int64_t) L + 1) + 63) & ~(int64_t) 63; }
__host__ __device__ inline int n2_oa_words(int B) { return (B + 7) / 8; }
__host__ __device__ inline int64_t n2_oa_bits_at(int L, int B) { return n2_xrows_floats(L) + (int64_t) L * 128 * B; }
__host__ __device__ inline int64_t n2_oa_rows_at(int L, int B) { return n2_oa_bits_at(L, B) + (int64_t) L * 64 * n2_oa_words(B); }
// then, for the alignment display, each cell's posterior probability codes
// (p7_alidisplay_EncodePostProb: 0-9, 10 for '*'), M in the low and I in the
// high four bits of a byte, four positions per word (position j in byte j % 4
// of word j / 4), words [row][j / 4][64 lanes]
__host__ __device__ inline int n2_pp_words(int B) { return (B + 3) / 4; }
__host__ __device__ inline int64_t n2_pp_at(int L, int B) { return n2_oa_rows_at(L, B) + (((int64_t) L + 1 + 63) & ~(int64_t) 63); }
__host__ __device__ inline int64_t n2_area_floats(int L, int B)
{ return n2_pp_at(L, B) + (int64_t) L * 64 * n2_pp_words(B); }
// optimal accuracy traceback choices: per cell, bits 0-1 the state before M
// (0 M, 1 I, 2 D, 3 B), bit 2 before I (1: I), bit 3 before D (1: D); per
// row, bits 0-15 the node E leaves from, bit 16 set if from D, bit 17 C from
// C (else E), bit 18 J from J (else E), bit 19 B from N (else J), bit 20 the
// E choice fell on a striped padding cell (no trace)
enum { OA_ESTATE = 16, OA_CBIT = 17, OA_JBIT = 18, OA_BBIT = 19, OA_EFAIL = 20 };
// Per-target results of the decoding (slot order) and the region records
// {slot, i, j, multidomain} in the order they were found.
struct DomOut {
  float   *nexp = nullptr;
  int32_t *nreg = nullptr, *nclu = nullptr;
  int4    *regs = nullptr;
  int64_t  reg_cap = 0;
  int32_t *rseq = nullptr, *rlen = nullptr;   // per record: the caller's index and the target's length
  int32_t *nclu_total = nullptr;   // regions flagged multidomain (zeroed before a run)
  // [0] regions written, [8] the envelope-scoring work-queue cursor (the list
  // kernels' count/queue layout); zeroed before a run
  int32_t *ecount = nullptr;
  // per region k: its residues copied to res + v_off[k] (16-byte aligned,
  // padded), v_len[k] residues, the unihit N/J/C odds for the whole target's
  // length, v_list[k] = k, and the envelope score and status
  uint8_t *res = nullptr;
  int64_t  res_cap = 0;
  int64_t *v_off = nullptr;
  int32_t *v_len = nullptr, *v_list = nullptr, *env_st = nullptr;
  float   *env_loop = nullptr, *env_move = nullptr, *envsc = nullptr;
  // the envelope-score pass's row scales: region k's at rscale + rsoff[k]
  float   *rscale = nullptr;
  int64_t *rsoff = nullptr;
  int64_t  rscale_cap = 0;
};

// Multidomain regions (p7g_multidomain.hip): one job per region, its copy at
// res + res_off (Lr residues), the whole target's length, and its area in
// the pool (md_area_floats); segcap bounds the sampled-domain ensemble.
#define MD_TSEG   512     // sampled domains in one trace
#define MD_MAXSIG 64      // significant clusters of one region
#define MD_MAXC   32      // envelopes one region may give
#define MD_ACC_LDS 8192   // regions up to this many residues + 1 sum their odds, and keep their residues, in LDS
__host__ __device__ inline int64_t md_q(int M) { return (M - 1) / 4 + 1 < 2 ? 2 : (M - 1) / 4 + 1; }
// checkpoints per row of the E state's running sum over the 8Q M and D
// cells (one at the end of every 64 of them, and the last)
__host__ __device__ inline int md_ncp(int64_t Q) { return (int) ((8 * Q + 63) / 64); }
// LDS, in floats: the transitions staged for the Forward (10 rows of M + 2),
// the Forward's rows, and the traces' buffers (from the start of LDS): a
// domain's counts (4Q), its null2 odds (32), the trace's domains, the
// residues' sums
__host__ __device__ inline int64_t md_stage_floats(int M) { return (10 * ((int64_t) M + 2) + 3) & ~(int64_t) 3; }
__host__ __device__ inline bool md_stage_t(int M) { return (md_stage_floats(M) + 32 * md_q(M)) * 4 <= P7G_LDS_MAX; }
__host__ __device__ inline int64_t md_acc_at(int64_t Q) { return 4 * Q + 32 + MD_TSEG * 4; }
// then, when it fits, the emission odds of 20 residues in striped order
// ([x][q][r]) for the sampled domains' null2
__host__ __device__ inline int64_t md_em_floats(int M) { return 20 * 4 * md_q(M); }
// a launch's room for the residue sums (floats) and the residues (bytes):
// its longest region's length + 1, rounded up, at most MD_ACC_LDS
__host__ __device__ inline int md_acc_cap(int max_lr)
{ const int c = (max_lr + 1 + 63) & ~63; return c < MD_ACC_LDS ? c : MD_ACC_LDS; }
__host__ __device__ inline bool md_em_lds(int M, int acc_cap)
{ return (md_acc_at(md_q(M)) + acc_cap + acc_cap / 4 + md_em_floats(M)) * 4 <= P7G_LDS_MAX; }
__host__ __device__ inline int64_t md_trace_floats(int M, int acc_cap)
{ return md_acc_at(md_q(M)) + acc_cap + acc_cap / 4 + (md_em_lds(M, acc_cap) ? md_em_floats(M) : 0); }
struct MdJob {
  int64_t res_off;
  int64_t area;
  int32_t Lr, Ltarget, segcap;
  int32_t n2o;          // the region's first residue in its launch's gathered null2 scores
};
// status 0, 1 (a traceback choice failed: the reference raises an exception),
// or 5 (a capacity above was exceeded); nc envelopes ci..cj, 1-based in the
// region, in start order; nseg the ensemble's size. The region's residues'
// null2 scores are left at pool + area + md_n2sc_at(M, Lr), residues 1..Lr.
struct MdOut {
  int32_t status, nc, nseg, pad;
  int32_t ci[MD_MAXC], cj[MD_MAXC];
};
__host__ __device__ inline int64_t md_n2sc_at(int M, int Lr)
{ return (int64_t) (Lr + 1) * 12 * md_q(M) + 6 * (int64_t) (Lr + 1); }
// then the endpoint counts, the sampled domains (5 ints each), the
// clustering's element clusters and cluster counts (segcap ints each), and
// (doubles) the traces' thresholds: 5 per cell, 3 per row, md_ncp running-sum
// checkpoints per row
__host__ __device__ inline int64_t md_seg_at(int M, int Lr)
{ return md_n2sc_at(M, Lr) + (Lr + 1) + ((Lr > M ? Lr : M) + 2); }
__host__ __device__ inline int64_t md_clu_at(int M, int Lr, int segcap)
{ return md_seg_at(M, Lr) + 5 * (int64_t) segcap; }
__host__ __device__ inline int64_t md_thr_at(int M, int Lr, int segcap)
{ return (md_clu_at(M, Lr, segcap) + 2 * (int64_t) segcap + 1) & ~(int64_t) 1; }
size_t  md_lds_bytes(int M, int segcap, int acc_cap);
int64_t md_area_floats(int M, int Lr, int segcap);
// thr_done (optional): recorded on st once the traces' thresholds are filled
hipError_t launch_md_regions(const DevProfile &p, const uint8_t *res, const MdJob *jobs, MdOut *out, int njobs,
                             float *pool, int segcap, int nsamples, uint32_t seed, int max_lr, hipStream_t st,
                             hipEvent_t thr_done = nullptr);
hipError_t launch_md_copy(uint8_t *res, const int64_t *src, const int64_t *dst, const int32_t *len, int n, hipStream_t st);
// the residue null2 scores of the launch's regions, job j's Lr of them at dst + jobs[j].n2o
hipError_t launch_md_n2_gather(const MdJob *jobs, int njobs, const float *pool, int M, float *dst, hipStream_t st);

// Length model of one call. on == 0: every target uses its own length
// (hmmsearch's per-target ReconfigLength); on == 1: the caller's current
// oprofile/bg scalars are used as they are (single-target drop-ins).
struct LenCfg {
  int   on;
  int   tjb, xmove;
  float floop, fmove;
  float p1, q1;          // bias filter HMM t[0][0], t[0][1]
};

struct GateThresh {
  float T1m;   // MSV/bias gate at F1 (MSV Gumbel)
  float T2m;   // "P <= F2" at the bias stage: skip the Viterbi filter
  float T2v;   // Viterbi gate at F2 (Viterbi Gumbel)
  float T3f;   // Forward gate at F3 (exponential tail)
  int   do_bias, do_max;
};

// Tables of the profiles one multi-profile launch of the ungapped pass
// serves, passed by value (kernel arguments).
#define P7G_SSV_MAXJOBS 192
struct SsvJobs { const int *tbl[P7G_SSV_MAXJOBS]; };

#ifdef __HIPCC__
#include <hip/hip_runtime.h>
// launchers (p7g_ssv.hip)
// The ungapped pass of variant v of one profile over slots[0..n) (slots ==
// NULL: every slot); waves take `grab` groups of targets per queue atomic.
hipError_t launch_ssv(const DevProfile &p, int v, const DevSeqs &s, const int32_t *slots, int64_t n, uint16_t *zmax,
                      int32_t *queue, int grab, hipStream_t st);
// The ungapped pass of njobs profiles that share one (G, ND) shape over every
// slot: one persistent launch per P7G_SSV_MAXJOBS profiles, whose blocks move
// from profile to profile as work queues drain. Profile j reads tbl[j], writes
// zmax + j * zstride and takes work from queues[j] (zero on entry).
hipError_t launch_ssv_multi(int G, int ND, const int *const *tbl, int njobs, const DevSeqs &s, int64_t n,
                            uint16_t *zmax, int64_t zstride, int32_t *queues, int grab, hipStream_t st);
// The ungapped pass of a chunked model: one launch per chunk, each reading
// the boundary column the one before wrote. bnd holds two columns of
// bnd_stride entries (at least one per residue byte of the database);
// queues[0 .. nchunk-1] are the launches' work-queue cursors (zero on entry).
hipError_t launch_ssv_chunked(const DevProfile &p, const DevSeqs &s, const int32_t *slots, int64_t n, uint16_t *zmax,
                              int32_t *queues, int grab, uint16_t *bnd, int64_t bnd_stride, hipStream_t st);
// 1 if a kernel of this shape is built
int ssv_shape_built(int G, int ND);
int ssv_chunk_built(int ND);
// launchers (p7g_kernels.hip)
// p[0..nwords) = v, as a kernel (so that a captured graph holds no memset node)
hipError_t launch_fill32(void *p, uint32_t v, int64_t nwords, hipStream_t st);
// dst[0..nwords) = src[0..nwords), device to device, as a kernel
hipError_t launch_copy32(void *dst, const void *src, int64_t nwords, hipStream_t st);
// With slots == NULL (every slot) targets that need the full recurrence are
// marked xj = -2 for launch_compact_i16; with a slot list they are appended.
hipError_t launch_msv_classify(const DevProfile &p, const DevSeqs &s, const int32_t *slots, int64_t n,
                               const uint16_t *zmax, int16_t *xj, int32_t *full_list, int32_t *full_count,
                               LenCfg lc, hipStream_t st);
hipError_t launch_msv_full(const DevProfile &p, const DevSeqs &s, const int32_t *list, const int32_t *count,
                           int64_t maxn, int16_t *xj, LenCfg lc, hipStream_t st);
// MSV score (nats) and status (0, or 16 on overflow) of listed slots from xj,
// each with its own length's tjb.
hipError_t launch_msv_scores(const DevProfile &p, const DevSeqs &s, const int32_t *slots, int64_t n,
                             const int16_t *xj, float *sc, int *st, hipStream_t stream);
hipError_t launch_msv_gate(const DevProfile &p, const DevSeqs &s, int64_t n, const int16_t *xj, DevResults r,
                           DevLists l, GateThresh g, hipStream_t st);
// The bias filter; with `rowmax` (room for L+1 floats per listed target, taken
// from l.cursor, which is zero again on return) the logs leave the serial
// recurrence. Targets that find no room are finished in the first pass.
hipError_t launch_bias(const DevProfile &p, const DevSeqs &s, const int32_t *list, const int32_t *count, int64_t maxn,
                       DevResults r, DevLists l, GateThresh g, int gating, LenCfg lc,
                       float *rowmax, int64_t rowmax_size, int64_t *defer_off, hipStream_t st);
hipError_t launch_vit(const DevProfile &p, const DevSeqs &s, const int32_t *list, const int32_t *count, int64_t maxn,
                      DevResults r, DevLists l, GateThresh g, int gating, LenCfg lc, hipStream_t st);
hipError_t launch_fwd(const DevProfile &p, const DevSeqs &s, const int32_t *list, const int32_t *count, int64_t maxn,
                      DevResults r, DevLists l, GateThresh g, int gating, LenCfg lc,
                      float *scale_pool, int64_t pool_size, int64_t *scale_off, float *xmx_out, hipStream_t st,
                      const DomBufs &dom = DomBufs());
#define P7G_Q_SSV 13        // counts[] index of the ungapped pass's work-queue cursor
#define P7G_QOFF  8         // a list's work-queue cursor is counts[list + P7G_QOFF]

// Order-preserving compaction: list = every slot s (ascending) with
// flags[s] == value (uint8) or xj[s] == value (int16); *count = its length.
// tilecnt needs (n + 4095) / 4096 ints.
hipError_t launch_compact_u8(const uint8_t *flags, int value, int64_t n, int32_t *list, int32_t *count,
                             int32_t *tilecnt, hipStream_t st);
hipError_t launch_compact_i16(const int16_t *flags, int value, int64_t n, int32_t *list, int32_t *count,
                              int32_t *tilecnt, hipStream_t st);
// Pipeline counters: out[0..3] targets with stage >= 1..4, out[4] residues with
// stage >= 1, out[5] residues with stage 4, out[6] and out[7] residues over the
// Viterbi and Forward lists. out[] must be zero on entry.
hipError_t launch_stats(const DevSeqs &s, const uint8_t *stage, int64_t n, const int32_t *vit, const int32_t *fwd,
                        const int32_t *counts, unsigned long long *out, hipStream_t st);
// dst[perm[slot]] = src[slot] for 4-byte and 1-byte per-slot arrays.
hipError_t launch_scatter32(const DevSeqs &s, const void *src, void *dst, int64_t n, hipStream_t st);
hipError_t launch_rec_gather(const DevSeqs &s, const int4 *regs, int64_t n, int32_t *rseq, int32_t *rlen, hipStream_t st);
hipError_t launch_scatter8(const DevSeqs &s, const uint8_t *src, uint8_t *dst, int64_t n, hipStream_t st);
hipError_t launch_bck(const DevProfile &p, const DevSeqs &s, const int32_t *list, const int32_t *count, int64_t maxn,
                      DevResults r, const float *scale_pool, const int64_t *scale_off, const float *fwd_xmx,
                      float *xmx_out, LenCfg lc, hipStream_t st, const DomBufs &dom = DomBufs(),
                      DevLists l = DevLists());
// The optimal accuracy traceback of each listed region (p7_OATrace,
// impl_sse/optacc.c:227-423) from the choices the null2 pass stored: one
// wave per region; the alignment's first and last M states go to dom.n2ali.
hipError_t launch_oa_trace(const DevProfile &p, const DevSeqs &s, const int32_t *list, int64_t n, const DomBufs &dom,
                           hipStream_t st);

// Posterior decoding of the special states and the region heuristics over the
// Backward list (the F3 survivors), one target per thread.
hipError_t launch_domains(const DevProfile &p, const DevSeqs &s, const int32_t *list, const int32_t *count, int64_t maxn,
                          const DevResults &r, DevLists l, LenCfg lc, const DomBufs &dom, const DomOut &out,
                          hipStream_t st);
#endif
