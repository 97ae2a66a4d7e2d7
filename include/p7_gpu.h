/* p7_gpu.h -- C ABI of the MI355X (gfx950) HMMER3 filter pipeline.
 *
 * This is the drop-in boundary. Every entry point takes plain pointers and
 * sizes, returns an Easel status code, and keeps the argument meaning of the
 * HMMER 3.1b2 function it stands in for. Citations are to the reference tree
 * (paths relative to the biobench2 root).
 *
 * Conventions shared with HMMER:
 *   - digital sequences are uint8_t (ESL_DSQ), residues dsq[1..L]; dsq[0] and
 *     dsq[L+1] are sentinels that are never read;
 *   - scores are in nats; a filter that overflows its limited range returns
 *     p7G_ERANGE with *ret_sc = +infinity (a high-scoring hit);
 *   - the quantised tables in P7_GPU_TABLES are the P7_OPROFILE tables
 *     (HMMER/src/impl_sse/impl_sse.h:76-143) with the Farrar striping removed:
 *     node k (1..M) of residue x sits at [x*M + k-1].
 */
#ifndef P7_GPU_INCLUDED
#define P7_GPU_INCLUDED

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes: numerically identical to Easel's (HMMER/easel/easel.h:95-123). */
#define p7G_OK          0
#define p7G_FAIL        1
#define p7G_EOF         3
#define p7G_EMEM        5
#define p7G_ENOTFOUND   6
#define p7G_EFORMAT     7
#define p7G_EINCOMPAT  10
#define p7G_EINVAL     11
#define p7G_ESYS       12
#define p7G_ERANGE     16
#define p7G_ENORESULT  19
#define p7G_EDEVICE    64   /* HIP runtime failure (no Easel counterpart)     */

/* Alphabet types, as Easel numbers them (HMMER/easel/esl_alphabet.h). */
#define p7G_RNA    1
#define p7G_DNA    2
#define p7G_AMINO  3

/* Profile modes (HMMER/src/hmmer.h:60-63). The filters need a local mode. */
#define p7G_LOCAL     1
#define p7G_GLOCAL    2
#define p7G_UNILOCAL  3

/* Special-state and transition indices of P7_OPROFILE
 * (HMMER/src/impl_sse/impl_sse.h:82-86). */
enum { p7G_XE = 0, p7G_XN = 1, p7G_XJ = 2, p7G_XC = 3 };
enum { p7G_MOVE = 0, p7G_LOOP = 1 };
enum { p7G_BM = 0, p7G_MM = 1, p7G_IM = 2, p7G_DM = 3,
       p7G_MD = 4, p7G_MI = 5, p7G_II = 6, p7G_DD = 7 };

/* Parser special cells per row (HMMER/src/impl_sse/impl_sse.h:173-175). */
enum { p7G_X_E = 0, p7G_X_N = 1, p7G_X_J = 2, p7G_X_B = 3, p7G_X_C = 4, p7G_X_SCALE = 5 };
#define p7G_NXCELLS 6

/* Evalue parameters (HMMER/src/hmmer.h:72). */
enum { p7G_MMU = 0, p7G_MLAMBDA = 1, p7G_VMU = 2, p7G_VLAMBDA = 3, p7G_FTAU = 4, p7G_FLAMBDA = 5 };

typedef struct p7g_hmm_s       P7_GPU_HMM;       /* probability model from a HMMER3 save file       */
typedef struct p7g_bg_s        P7_GPU_BG;        /* null model + bias-composition filter HMM (P7_BG) */
typedef struct p7g_profile_s   P7_GPU_PROFILE;   /* float log-odds profile (P7_PROFILE)             */
typedef struct p7g_oprofile_s  P7_GPU_OPROFILE;  /* quantised tables, host + device (P7_OPROFILE)   */
typedef struct p7g_omx_s       P7_GPU_OMX;       /* DP scratch and parser specials (P7_OMX)          */
typedef struct p7g_seqdb_s     P7_GPU_SEQDB;     /* HBM-resident digital target database             */
typedef struct p7g_pipeline_s  P7_GPU_PIPELINE;  /* thresholds + counters (P7_PIPELINE)             */

/* The quantised P7_OPROFILE tables, unstriped. Used to import a profile that
 * HMMER already converted (de-striped by the caller) and to export ours. */
typedef struct p7g_tables_s {
  int       M;                 /* model length                                           */
  int       Kp;                /* alphabet size incl. degenerate/gap/missing codes       */
  int       abctype;           /* p7G_AMINO, p7G_DNA, p7G_RNA                             */
  /* MSVFilter: biased uchar costs (p7_oprofile.c:764-804)                                */
  const uint8_t *rbv;          /* [Kp][M]                                                */
  uint8_t   tbm_b, tec_b, tjb_b, base_b, bias_b;
  float     scale_b;
  /* ViterbiFilter: int16 scores (p7_oprofile.c:817-912)                                 */
  const int16_t *rwv;          /* [Kp][M]                                                */
  const int16_t *twv;          /* [8][M]: BM,MM,IM,DM into node k; MD,MI,II,DD out of k  */
  int16_t   xw[4][2];          /* [E,N,J,C][MOVE,LOOP]                                   */
  float     scale_w;
  int16_t   base_w, ddbound_w;
  /* Forward/Backward: odds ratios (p7_oprofile.c:920-985)                                */
  const float *rfv;            /* [Kp][M]                                                */
  const float *tfv;            /* [8][M], same ordering as twv                           */
  float     xf[4][2];
  /* configuration                                                                          */
  int       L;                 /* currently configured target length                     */
  int       mode;              /* p7G_LOCAL or p7G_UNILOCAL                               */
  float     nj;                /* expected J uses: 1 multihit, 0 unihit                  */
  float     evparam[6];
  float     compo[20];         /* model composition for the bias filter (K entries)      */
} P7_GPU_TABLES;

/* Per-target outputs of a batched pipeline run. Every pointer may be NULL;
 * otherwise it holds one entry per database sequence, in database order.
 * Scores that a stage did not compute are NaN. */
typedef struct p7g_hits_s {
  float   *nullsc;     /* p7_bg_NullOne                                    */
  float   *usc;        /* MSV score (nats); +inf when it overflowed          */
  float   *filtersc;   /* bias filter score (MSV survivors)                 */
  float   *vfsc;       /* Viterbi filter score (when the pipeline ran it)  */
  float   *fwdsc;      /* Forward parser score                               */
  float   *bcksc;      /* Backward parser score (F3 survivors)              */
  uint8_t *stage;      /* 0 failed MSV, 1 failed bias, 2 failed Vit,
                          3 failed Fwd, 4 passed all filters               */
  /* Domain decoding of the F3 survivors (p7_gpu_Pipeline with
   * p7_gpu_pipeline_SetDomains on; other calls leave these untouched):
   * P7_DOMAINDEF's nexpected, nregions and nclustered (NaN / 0 / 0 for
   * targets that did not pass every filter). */
  float   *nexpected;  /* posterior expected number of domains, btot[L]   */
  int32_t *nregions;   /* regions found by the posterior heuristics        */
  int32_t *nclustered; /* of which flagged for trace clustering            */
} P7_GPU_HITS;

/* Counters of P7_PIPELINE (HMMER/src/hmmer.h:1223-1237) plus device timings. */
typedef struct p7g_stats_s {
  int64_t nmodels, nseqs, nres, nnodes;
  int64_t n_past_msv, n_past_bias, n_past_vit, n_past_fwd;
  int64_t n_msv_full;          /* targets the MSV stage rescored with the J state   */
  double  cells_msv, cells_bias, cells_vit, cells_fwd, cells_bck;
  double  ms_msv, ms_bias, ms_vit, ms_fwd, ms_bck, ms_total;   /* device time, last call */
  double  ms_ssv;              /* the ungapped MSV pass alone (k_ssv), last call    */
  double  ms_domains;          /* domain decoding and regions, last call (0 if off) */
  int64_t n_regions, n_clustered;   /* regions of the last call, and multidomain ones */
  double  ms_null2;            /* rescoring of the regions (null2, alignments, multidomain regions), last call (0 if off) */
  double  ms_multidomain;      /* of which the multidomain regions' sampled traces and clusters */
} P7_GPU_STATS;

/* ---- device ------------------------------------------------------------- */
int         p7_gpu_Init(int device);               /* select device, create the stream  */
int         p7_gpu_Synchronize(void);
const char *p7_gpu_LastError(void);                /* text of the last failure           */

/* ---- models (host) ------------------------------------------------------- */
/* Replaces p7_hmmfile_Read (HMMER/src/p7_hmmfile.c:1180, ascii 3/e and 3/f at :1294). */
int   p7_gpu_hmmfile_Read(const char *path, int which, P7_GPU_HMM **ret_hmm);
int   p7_gpu_hmm_Info(const P7_GPU_HMM *hmm, int *ret_M, int *ret_abctype, char *name, int namelen);
void  p7_gpu_hmm_Destroy(P7_GPU_HMM *hmm);

/* Replaces p7_bg_Create (HMMER/src/p7_bg.c:55). */
P7_GPU_BG *p7_gpu_bg_Create(int abctype);
void       p7_gpu_bg_Destroy(P7_GPU_BG *bg);
/* Replaces p7_bg_SetLength (p7_bg.c:191). */
int   p7_gpu_bg_SetLength(P7_GPU_BG *bg, int L);
/* Replaces p7_bg_NullOne (p7_bg.c:359). */
int   p7_gpu_bg_NullOne(const P7_GPU_BG *bg, const uint8_t *dsq, int L, float *ret_sc);
/* Replaces p7_bg_SetFilter (p7_bg.c:431). */
int   p7_gpu_bg_SetFilter(P7_GPU_BG *bg, int M, const float *compo);
/* Replaces p7_bg_FilterScore (p7_bg.c:473); runs the 2-state Forward on the GPU. */
int   p7_gpu_bg_FilterScore(P7_GPU_BG *bg, const uint8_t *dsq, int L, float *ret_sc);

/* Replaces p7_ProfileConfig (HMMER/src/modelconfig.c:49). */
int   p7_gpu_ProfileConfig(const P7_GPU_HMM *hmm, const P7_GPU_BG *bg, int L, int mode, P7_GPU_PROFILE **ret_gm);
/* Replaces p7_ReconfigLength (modelconfig.c:209). */
int   p7_gpu_profile_ReconfigLength(P7_GPU_PROFILE *gm, int L);
/* Copies out the float scores: tsc[M*8] (P7_PROFILE order MM,IM,DM,BM,MD,DD,MI,II),
 * rsc[Kp*(M+1)*2] ([x][k][MSC,ISC]), xsc[4*2] ([E,N,J,C][LOOP,MOVE]). */
int   p7_gpu_profile_GetScores(const P7_GPU_PROFILE *gm, float *tsc, float *rsc, float *xsc);
void  p7_gpu_profile_Destroy(P7_GPU_PROFILE *gm);

/* Replaces p7_oprofile_Create (HMMER/src/impl_sse/p7_oprofile.c:50). */
P7_GPU_OPROFILE *p7_gpu_oprofile_Create(int allocM, int abctype);
/* Replaces p7_oprofile_Convert (p7_oprofile.c:1005). */
int   p7_gpu_oprofile_Convert(const P7_GPU_PROFILE *gm, P7_GPU_OPROFILE *om);
/* Imports tables of an already converted P7_OPROFILE (de-striped by the caller). */
int   p7_gpu_oprofile_Load(P7_GPU_OPROFILE *om, const P7_GPU_TABLES *t);
/* Reads model `which` (0-based) of a pressed profile database: the .h3f and
 * .h3p files hmmpress writes (format 3/f), as p7_oprofile_ReadMSV +
 * p7_oprofile_ReadRest do (HMMER/src/impl_sse/io.c:233-338, 493-581), with
 * the SSE striping removed. p7G_EOF when the files hold fewer models. */
int   p7_gpu_oprofile_ReadPressed(const char *h3f_path, const char *h3p_path, int which, P7_GPU_OPROFILE **ret_om);
/* Exports the tables; pointers stay valid until the next Load/Convert/Destroy. */
int   p7_gpu_oprofile_GetTables(const P7_GPU_OPROFILE *om, P7_GPU_TABLES *t);
/* Replace p7_oprofile_ReconfigLength / MSVLength / RestLength (p7_oprofile.c:1062-1127). */
int   p7_gpu_oprofile_ReconfigLength(P7_GPU_OPROFILE *om, int L);
int   p7_gpu_oprofile_ReconfigMSVLength(P7_GPU_OPROFILE *om, int L);
int   p7_gpu_oprofile_ReconfigRestLength(P7_GPU_OPROFILE *om, int L);
void  p7_gpu_oprofile_Destroy(P7_GPU_OPROFILE *om);

/* Replaces p7_omx_Create / GrowTo / Destroy (HMMER/src/impl_sse/p7_omx.c:64,140). */
P7_GPU_OMX *p7_gpu_omx_Create(int allocM, int allocL, int allocXL);
int   p7_gpu_omx_GrowTo(P7_GPU_OMX *ox, int allocM, int allocL, int allocXL);
/* The parser specials xmx[(L+1)*6] ([i*6 + p7G_X_*]), totscale and has_own_scales. */
const float *p7_gpu_omx_Xmx(const P7_GPU_OMX *ox, int *ret_L, float *ret_totscale, int *ret_has_own_scales);
void  p7_gpu_omx_Destroy(P7_GPU_OMX *ox);

/* ---- single-target drop-ins (same signatures as impl_sse.h:308-344) -------- */
/* p7_MSVFilter (msvfilter.c:76). */
int   p7_gpu_MSVFilter(const uint8_t *dsq, int L, const P7_GPU_OPROFILE *om, P7_GPU_OMX *ox, float *ret_sc);
/* p7_SSVFilter (ssvfilter.c:878): p7G_ENORESULT when the J state would be used. */
int   p7_gpu_SSVFilter(const uint8_t *dsq, int L, const P7_GPU_OPROFILE *om, float *ret_sc);
/* p7_ViterbiFilter (vitfilter.c:84). */
int   p7_gpu_ViterbiFilter(const uint8_t *dsq, int L, const P7_GPU_OPROFILE *om, P7_GPU_OMX *ox, float *ret_sc);
/* p7_ForwardParser (fwdback.c:134). Fills fwd's specials for the Backward pass. */
int   p7_gpu_ForwardParser(const uint8_t *dsq, int L, const P7_GPU_OPROFILE *om, P7_GPU_OMX *fwd, float *opt_sc);
/* p7_BackwardParser (fwdback.c:238), using fwd's row scale factors. */
int   p7_gpu_BackwardParser(const uint8_t *dsq, int L, const P7_GPU_OPROFILE *om, const P7_GPU_OMX *fwd,
                            P7_GPU_OMX *bck, float *opt_sc);

/* ---- HBM-resident database and batched stages ----------------------------- */
/* residues: concatenation; seq i is residues[offset[i] .. offset[i]+length[i]-1]
 * (digital codes 0..Kp-1, no sentinels). The library keeps its own copy in HBM,
 * grouped by length; results are always reported in the caller's order. */
int     p7_gpu_seqdb_Create(const uint8_t *residues, const int64_t *offset, const int32_t *length,
                            int64_t nseq, P7_GPU_SEQDB **ret_db);
/* Reads a FASTA file (first word of each '>' line is the name; a first line
 * starting with '#', the summary line of hmmpgmd's cache format, is skipped)
 * into the HBM database, as p7_seqcache_Open (HMMER/src/cachedb.c:41-280)
 * loads hmmpgmd's resident cache. Residues follow Easel's digital alphabets. */
int     p7_gpu_seqdb_ReadFasta(const char *path, int abctype, P7_GPU_SEQDB **ret_db);
/* Name of sequence i (file order), or NULL when the database has no names. */
const char *p7_gpu_seqdb_Name(const P7_GPU_SEQDB *db, int64_t i);
/* Writes / reads an image of the device layout (length-sorted, padded residues,
 * names), so a resident database comes back without parsing or sorting.
 * No counterpart in the reference. */
int     p7_gpu_seqdb_Save(const P7_GPU_SEQDB *db, const char *path);
int     p7_gpu_seqdb_Load(const char *path, P7_GPU_SEQDB **ret_db);
int64_t p7_gpu_seqdb_NSeq(const P7_GPU_SEQDB *db);
int64_t p7_gpu_seqdb_NRes(const P7_GPU_SEQDB *db);
void    p7_gpu_seqdb_Destroy(P7_GPU_SEQDB *db);

/* Each target is scored with the profile length-configured to its own L, as
 * hmmsearch does (hmmsearch.c:1404-1405). idx == NULL means every sequence;
 * outputs are indexed like idx (or like the database when idx is NULL). */
int   p7_gpu_MSVFilterBatch    (const P7_GPU_OPROFILE *om, const P7_GPU_SEQDB *db, const int64_t *idx, int64_t n,
                                float *ret_sc, int *ret_status);
int   p7_gpu_ViterbiFilterBatch(const P7_GPU_OPROFILE *om, const P7_GPU_SEQDB *db, const int64_t *idx, int64_t n,
                                float *ret_sc, int *ret_status);
int   p7_gpu_ForwardParserBatch(const P7_GPU_OPROFILE *om, const P7_GPU_SEQDB *db, const int64_t *idx, int64_t n,
                                float *ret_sc, int *ret_status);
int   p7_gpu_BackwardParserBatch(const P7_GPU_OPROFILE *om, const P7_GPU_SEQDB *db, const int64_t *idx, int64_t n,
                                 float *ret_sc, int *ret_status);
int   p7_gpu_FilterScoreBatch  (const P7_GPU_OPROFILE *om, const P7_GPU_SEQDB *db, const int64_t *idx, int64_t n,
                                float *ret_sc);
/* p7_gpu_MSVFilterBatch for callers that cannot allocate memory (code under
 * HIP graph capture, serving loops, harnesses that forbid hipMalloc): no
 * counterpart in the reference. Same arguments and results, bit for bit; it
 * calls no hipMalloc/hipFree and no allocating routine. Its buffers live in
 * a static device array of the library (about 20 MiB per device, held until
 * the process ends), so:
 *   - db and n are limited to 1,048,576 sequences each (P7G_NOALLOC_MAXSEQ);
 *     larger ones are p7G_EINVAL;
 *   - p7_gpu_Init must have run on the device, and the profile's device copy
 *     must already cover db's lengths (any earlier batched call or
 *     p7_gpu_Pipeline with this profile and database makes it): else p7G_EINVAL;
 *   - models longer than one LDS table (M > 2432, the chunked ungapped pass)
 *     are p7G_EINVAL;
 *   - every call on one device shares that device's array (it keeps the last
 *     call's contents; each call clears its counters first). The library
 *     serialises these calls on one lock and stream; calls that overlapped on
 *     one device would race on the array. */
int   p7_gpu_MSVFilterBatchNoAlloc(const P7_GPU_OPROFILE *om, const P7_GPU_SEQDB *db, const int64_t *idx, int64_t n,
                                   float *ret_sc, int *ret_status);

/* ---- pipeline (p7_pipeline.c:99, :666-731) -------------------------------- */
/* F1/F2/F3 are P-value thresholds (defaults 0.02, 1e-3, 1e-5; p7_pipeline.c:210-212);
 * do_biasfilter mirrors --nobias; do_max mirrors --max (all filters off). */
P7_GPU_PIPELINE *p7_gpu_pipeline_Create(double F1, double F2, double F3, int do_biasfilter, int do_max);
void  p7_gpu_pipeline_Destroy(P7_GPU_PIPELINE *pli);
/* Runs MSV -> bias -> Viterbi -> Forward -> Backward over every database
 * sequence on the device, with the stage gates of p7_Pipeline. Accumulates
 * the counters like p7_pli_NewModel/NewSeq/p7_Pipeline. */
int   p7_gpu_Pipeline(P7_GPU_PIPELINE *pli, const P7_GPU_OPROFILE *om, const P7_GPU_SEQDB *db, P7_GPU_HITS *hits);
int   p7_gpu_pipeline_GetStats(const P7_GPU_PIPELINE *pli, P7_GPU_STATS *st);
/* Domain decoding for the F3 survivors of p7_gpu_Pipeline: the part of
 * p7_domaindef_ByPosteriorHeuristics (HMMER/src/p7_domaindef.c:386-489) as
 * p7_Pipeline calls it (p7_pipeline.c:737): p7_DomainDecoding
 * (impl_sse/decoding.c:166-197) from the two parsers' special states, then
 * the region heuristics with is_multidomain_region (rt1 0.25, rt2 0.10, rt3
 * 0.20) and each region's envelope score; with p7_gpu_pipeline_SetNull2 (on
 * by default) the rest of the domain definition: each multidomain region's
 * sampled traces and clusters (region_trace_ensemble, p7_domaindef.c:600-690,
 * 200 samples, seed 42), and for every envelope (a single-domain region, or a
 * cluster of a multidomain one) rescore_isolated_domain's null2 correction
 * and optimal accuracy alignment (:812-947). Off by default; batches and the
 * server do not decode. */
int   p7_gpu_pipeline_SetDomains(P7_GPU_PIPELINE *pli, int on);
/* region_trace_ensemble (HMMER/src/p7_domaindef.c:600-690), the analysis of
 * one region that p7_domaindef_ByPosteriorHeuristics finds multidomain
 * (:431-446), on residues dsq[1..Lr] of a target of Ltarget residues: the
 * region's p7_Forward with the profile in multihit mode for the whole target
 * (fwdback.c:258-465, every row kept), nsamples stochastic traces
 * (impl_sse/stotrace.c) with Easel's fast generator reset to seed (the
 * pipeline uses 200 and 42, p7_domaindef.c:112, p7_pipeline.c:102), null2 by
 * trace (impl_sse/null2.c:134), and the clustering of the sampled domains
 * (p7_spensemble.c:282-411) without the dominated ones. Out: *ret_nc
 * envelopes ci..cj (1-based in the region, start order; up to cap written),
 * n2sc[0..Lr-1] the residues' log null2 scores (ddef->n2sc[ireg..jreg]).
 * Optional: opt_mx the Forward matrix as P7_OMX's dpf keeps it, rows 0..Lr of
 * 12*Q floats ([q][M,D,I][4 lanes], Q = max(2, (M-1)/4 + 1)); opt_xmx its
 * special rows (E N J B C SCALE); opt_segs the sampled domains (trace index,
 * i, j, k, m per domain, up to segcap) and *ret_nseg their number. segcap <=
 * 0: the pipeline's 6144. p7G_EINVAL if a traceback choice fails (the
 * reference raises an exception), p7G_EMEM past segcap sampled domains. */
int   p7_gpu_RegionTraceEnsemble(const uint8_t *dsq, int Lr, int Ltarget, const P7_GPU_OPROFILE *om, int nsamples,
                                 uint32_t seed, int32_t *ci, int32_t *cj, int cap, int32_t *ret_nc, float *n2sc,
                                 float *opt_mx, float *opt_xmx, int32_t *opt_segs, int segcap, int32_t *ret_nseg);
/* The null2 correction of decoded regions (pli->do_null2, p7_pipeline.c:208;
 * hmmsearch --nonull2 turns it off). On by default; it applies when decoding
 * is on. */
int   p7_gpu_pipeline_SetNull2(P7_GPU_PIPELINE *pli, int on);
/* No counterpart in the reference (its domain definition works one target
 * at a time in host memory): caps each of the device pools that the domain
 * definition's regions take their working areas from (the envelopes' null2
 * and alignment areas, the multidomain regions' matrices and thresholds) at
 * floats floats. 0 (the default): the smaller of what the regions need and
 * a quarter of the free device memory, at least 2 GiB. Regions whose areas
 * exceed one pool run in several batches, with the same results; a cap below
 * one region's area makes p7_gpu_Pipeline fail with p7G_EMEM. For callers
 * with a device memory budget. */
int   p7_gpu_pipeline_SetPoolCap(P7_GPU_PIPELINE *pli, int64_t floats);
/* The region records of the last p7_gpu_Pipeline call with decoding on,
 * ordered by sequence (the database's order), start, then kind: sequence
 * index, first and last residue (1-based, inclusive), the kind (multi): 0 a
 * single-domain region (one envelope), 1 a multidomain region, 2 an envelope
 * that a multidomain region's clusters define (after that region, in start
 * order; only when null2 is on); and the envelope score (nats): the Forward
 * score of residues i..j with the profile in unihit mode for the whole
 * target's length, as rescore_isolated_domain computes envsc
 * (p7_domaindef.c:812, 841). The domains of a target, in the reference's
 * order, are its records of kinds 0 and 2. Writes up to cap entries (any
 * array may be NULL); *ret_n = the number of records. Valid until the
 * pipeline's next call, while that call's database lives. */
int   p7_gpu_pipeline_GetRegions(const P7_GPU_PIPELINE *pli, int64_t *seq, int32_t *ienv, int32_t *jenv,
                                 int32_t *multi, float *envsc, int64_t cap, int64_t *ret_n);
/* The null2 correction of each record, in p7_gpu_pipeline_GetRegions' order:
 * domcorr = sum over the envelope's residues of log null2(x) (nats), as
 * rescore_isolated_domain computes it (p7_domaindef.c:841-947): for a
 * single-domain region by expectation (p7_Forward and p7_Backward over
 * residues i..j in unihit mode, p7_Decoding, p7_Null2_ByExpectation); for a
 * cluster's envelope from its region's sampled traces (p7_Null2_ByTrace
 * averaged over the samples). status: 0 computed; 16 the envelope's decoding
 * overflowed (eslERANGE, for which the reference drops the domain); 2 a
 * multidomain region itself, whose domcorr is the sum of the sampled null2
 * scores of all its residues (what the sequence bias counts for it).
 * p7G_EINVAL if the last call ran without null2. */
int   p7_gpu_pipeline_GetRegionNull2(const P7_GPU_PIPELINE *pli, float *domcorr, int32_t *status, int64_t cap,
                                     int64_t *ret_n);
/* The optimal accuracy alignment of each region, in p7_gpu_pipeline_GetRegions'
 * order, as rescore_isolated_domain computes it for a one-envelope region
 * (p7_domaindef.c:846-855: p7_OptimalAccuracy and p7_OATrace,
 * impl_sse/optacc.c:60, 227, on the envelope's unihit posterior decoding):
 * oasc, the expected number of correctly aligned residues (dom->oasc; the
 * domain table's acc column is oasc / (1 + |jenv - ienv|)), and the model and
 * target coordinates of the alignment's first and last M states
 * (p7_alidisplay_Create's hmmfrom, hmmto, sqfrom, sqto, p7_alidisplay.c:
 * 140-145). NaN / 0 for multidomain regions themselves (their clusters'
 * envelopes have theirs) and envelopes whose decoding failed. p7G_EINVAL if
 * the last call ran without null2 (the alignments are computed in the same
 * pass). */
int   p7_gpu_pipeline_GetDomainAlignments(const P7_GPU_PIPELINE *pli, float *oasc, int32_t *hmmfrom, int32_t *hmmto,
                                          int32_t *alifrom, int32_t *alito, int64_t cap, int64_t *ret_n);
/* Hit-list scores of the last p7_gpu_Pipeline call with decoding and null2,
 * as p7_Pipeline computes them after domain definition (p7_pipeline.c:743-850):
 * one entry per target with at least one region, in sequence order: the
 * null2-corrected bit score (the reconstruction score when that is higher
 * and counts a domain), the uncorrected one, the reconstruction score, the
 * natural-log P-value of the score under the Forward exponential tail
 * (esl_exp_logsurv with the profile's tau, lambda), the number of domains
 * (envelopes whose rescoring succeeded), the best domain's index among the
 * target's records (GetRegions order, -1 without domains) and a status: 0,
 * or 1 for a target without envelopes (each of its regions multidomain
 * without a significant cluster), which the reference does not make a hit
 * (p7_pipeline.c:739-740) and the hit list leaves out.
 * The sequence bias sums every residue's null2 score: the single-domain
 * envelopes' and each multidomain region's sampled ones (p7_pipeline.c:760).
 * The E-value of a score is exp(lnP) * Z for a search space of Z targets. */
int   p7_gpu_pipeline_GetTargetScores(const P7_GPU_PIPELINE *pli, int64_t *seq, float *score, float *pre_score,
                                      float *sum_score, double *lnP, int32_t *ndom, int32_t *best, int32_t *status,
                                      int64_t cap, int64_t *ret_n);
/* Per region (GetRegions order), its domain's scores (p7_pipeline.c:843-850):
 * the bit score after the null2 correction and the length correction of the
 * rest of the target, dombias (nats) = p7_FLogsum(0, log(1/256) + domcorr),
 * and the natural-log P-value of the bit score (the i-Evalue is exp(lnP) * Z,
 * the c-Evalue exp(lnP) * domZ). NaN for multidomain regions themselves and
 * envelopes whose rescoring failed. */
int   p7_gpu_pipeline_GetDomainScores(const P7_GPU_PIPELINE *pli, float *bitscore, float *dombias, double *lnP,
                                      int64_t cap, int64_t *ret_n);
/* The hit list of the scored targets (GetTargetScores status 0), ranked as
 * p7_tophits_SortBySortkey ranks them under E-value inclusion (by -lnP, ties
 * by sequence index) and flagged as p7_tophits_Threshold flags them with
 * E-value thresholds (p7_tophits.c:866-935, p7_pipeline.c:407-460): order[k]
 * is an index into GetTargetScores' arrays, flags[k] bit 0 reported
 * (exp(lnP)·Z <= E), bit 1 included (also <= incE). domZ = the number of
 * reported targets. dflags (NULL, or one entry per region in GetRegions
 * order): bit 0 reported (its target reported, exp(lnP)·domZ <= domE), bit 1
 * included (its target included, <= incdomE). hmmsearch's defaults:
 * E = 10, incE = 0.01, domE = 10, incdomE = 0.01. */
int   p7_gpu_pipeline_GetHitList(const P7_GPU_PIPELINE *pli, double Z, double E, double incE, double domE,
                                 double incdomE, int64_t *order, int32_t *flags, int32_t *dflags, double *ret_domZ,
                                 int64_t cap, int64_t *ret_n);
/* P7_DOMAINDEF's counters of each target, in GetTargetScores' order
 * (p7_domaindef.c:406-483, copied into the hit at p7_pipeline.c:816-821):
 * nexpected (btot[L]), the regions the posterior heuristics found, of which
 * multidomain, the clusters that start at or before the end of the previous
 * successfully rescored cluster of their region (noverlaps), and the
 * envelopes (single-domain regions plus clusters, rescored or not). */
int   p7_gpu_pipeline_GetTargetCounts(const P7_GPU_PIPELINE *pli, float *nexpected, int32_t *nregions,
                                      int32_t *nclustered, int32_t *noverlaps, int32_t *nenvelopes, int64_t cap,
                                      int64_t *ret_n);
/* ---- the hit list as hmmsearch writes it -----------------------------------
 * What the text needs beyond the last run's results: the query profile's
 * name, accession (NULL or "" for none) and length M; the targets' names,
 * accessions and descriptions as arrays indexed by database sequence (the
 * caller's order; tacc, tdesc or any of their entries may be NULL, and ""
 * counts as none, as an ESL_SQ's empty acc and desc do); the search spaces and
 * E-value thresholds (hmmsearch: Z = the number of targets searched unless
 * --Z, domZ < 0 for the number of reported targets unless --domZ, E = 10,
 * incE = 0.01, domE = 10, incdomE = 0.01); --textw (120; 0 for --notextw)
 * and --acc; show_alignments draws each reported domain's alignment in the
 * domain section (hmmsearch's default; 0 is --noali). */
typedef struct p7g_report_s {
  const char        *qname;
  const char        *qacc;
  int                M;
  const char *const *tname;
  const char *const *tacc;
  const char *const *tdesc;
  double             Z, domZ;
  double             E, incE, domE, incdomE;
  int                textw;
  int                show_accessions;
  int                show_alignments;   /* DomainsText: each domain's alignment (0: --noali) */
} P7_GPU_REPORT;
/* p7_tophits_Targets (p7_tophits.c:1055-1180): "Scores for complete
 * sequences", one line per reported target in rank order, the inclusion
 * threshold line, or the no-hits line. p7_tophits_Domains (:1200-1380): per
 * reported target its name and description line, its reported domains'
 * table and, with show_alignments, each reported domain's optimal accuracy
 * alignment as p7_alidisplay_Print draws it (p7_alidisplay.c:59-210,
 * 572-641: CS and RF lines when the model has them, the consensus, match,
 * target and posterior probability lines, in blocks as wide as textw
 * allows; the consensus and CS/RF lines are the annotation the profile's
 * .hmm file gave, an 'x' consensus for profiles loaded from tables). The tabular forms,
 * --tblout (p7_tophits_TabularTargets, :1519-1590) and --domtblout
 * (p7_tophits_TabularDomains, :1610-1700), with their three header lines when
 * show_header is set (hmmsearch sets it for the first query). Each writes at
 * most cap bytes of text including the terminating NUL into buf (buf may be
 * NULL with cap 0 to size it) and the text's full length to *ret_len.
 * p7G_EINVAL without a decoded run with null2 or without names. */
int   p7_gpu_pipeline_TargetsText(const P7_GPU_PIPELINE *pli, const P7_GPU_REPORT *rep, char *buf, size_t cap,
                                  size_t *ret_len);
int   p7_gpu_pipeline_DomainsText(const P7_GPU_PIPELINE *pli, const P7_GPU_REPORT *rep, char *buf, size_t cap,
                                  size_t *ret_len);
int   p7_gpu_pipeline_TabularTargets(const P7_GPU_PIPELINE *pli, const P7_GPU_REPORT *rep, int show_header,
                                     char *buf, size_t cap, size_t *ret_len);
int   p7_gpu_pipeline_TabularDomains(const P7_GPU_PIPELINE *pli, const P7_GPU_REPORT *rep, int show_header,
                                     char *buf, size_t cap, size_t *ret_len);
/* The "Internal pipeline statistics summary" text of p7_pli_Statistics
 * (p7_pipeline.c:1710-1796) for a search of models against sequences, from
 * the counters of p7_gpu_pipeline_GetStats: the query and target counts, the
 * four filter lines with their expected numbers for F1/F2/F3, and the search
 * spaces Z and domZ with their notes (z_by_option / domz_by_option: set on
 * the command line, else the number of targets / of targets reported). The
 * stopwatch lines are the caller's. Writes at most cap bytes including the
 * terminating NUL; returns the text's full length. */
size_t p7_gpu_StatisticsText(const P7_GPU_STATS *st, double F1, double F2, double F3, double Z, int z_by_option,
                             double domZ, int domz_by_option, char *buf, size_t cap);
/* ---- batches: many profiles against one resident database ----------------
 * Runs the p7_gpu_Pipeline filter chain of nprof profiles over every database
 * sequence in one call. The ungapped MSV pass of all profiles that share a
 * kernel shape runs as one persistent launch, with a device work queue per
 * profile, the shapes one after another, largest models first; each
 * profile's later stages (MSV decision, bias filter, Viterbi, Forward,
 * Backward) follow on one of up to sixteen concurrent streams as soon as its
 * shape's launch is done. With flags P7_GPU_BATCH_GRAPH each profile's stage chain is
 * captured once as a HIP graph of its own and replayed, on its stream after
 * its shape's launch, while the same profiles, database, thresholds and
 * output arrays are asked for.
 * hits and stats are arrays of nprof entries; either may be NULL, and so may
 * any pointer in hits. Each profile's results equal a p7_gpu_Pipeline call
 * with the same pipeline. Counters add to the pipeline's totals; ms_total of
 * every stats entry, and of the totals, is the device time of the whole batch.
 * No counterpart in the reference, whose hmmsearch runs one query at a time. */
#define P7_GPU_BATCH_EAGER 1   /* launch directly (the default)          */
#define P7_GPU_BATCH_GRAPH 2   /* capture as a HIP graph, replay it      */
int   p7_gpu_PipelineBatch(P7_GPU_PIPELINE *pli, const P7_GPU_OPROFILE *const *oms, int nprof,
                           const P7_GPU_SEQDB *db, P7_GPU_HITS *hits, P7_GPU_STATS *stats, int flags);
/* 1 if the last batch of this pipeline ran from the captured graph. */
int   p7_gpu_pipeline_BatchGraphed(const P7_GPU_PIPELINE *pli);

/* ---- tuner of the MSV stage's ungapped pass --------------------------------
 * For the length class of each model (M rounded up to a multiple of 32), times
 * every shape of the ungapped pass the model fits -- 8, 16, 32 or 64 lanes per
 * target, i.e. 8, 4, 2 or 1 targets per wavefront, each with its
 * positions-per-lane tile -- with work-queue grabs of 1, 2, 4 and 8 target
 * groups, on every k-th sequence of db (so the sample keeps the length mix),
 * and keeps the fastest for this device unless the built-in shape is within
 * 3% of it. Shapes that compute more than twice
 * the positions of the tightest one are not tried. Pipelines, batches and the
 * server use the kept choice from then on; no result depends on it. */
int   p7_gpu_Autotune(const P7_GPU_OPROFILE *const *oms, int nprof, const P7_GPU_SEQDB *db);
/* The choice for models of length M: lanes per target, queue grab, and the
 * best time of each shape (ret_us[4]: 16, 32, 64, 8 lanes; -1 if not tried).
 * Returns 1 if M's class was measured, 0 if the built-in default applies. */
int   p7_gpu_tune_Get(int M, int *ret_lanes, int *ret_grab, float *ret_us);
void  p7_gpu_tune_Clear(void);
/* A text file of the measured choices, for the device's architecture. */
int   p7_gpu_tune_Save(const char *path);
int   p7_gpu_tune_Load(const char *path);

/* Score threshold of a stage gate: the smallest float bit score s whose
 * P-value is <= P under the Gumbel (which = 0, easel/esl_gumbel.c:129) or
 * exponential (which = 1, easel/esl_exponential.c:127) tail, so the gate
 * "P > F: reject" (p7_pipeline.c:692, 702, 720, 729) is "s < threshold". */
float p7_gpu_ScoreThreshold(int which, double mu, double lambda, double P);

/* The device memory of a profile's tables for searches of this database
 * (its device copy is made here if it has none for the database's longest
 * target yet): what a served query holds while in flight. No counterpart in
 * the reference. */
int   p7_gpu_oprofile_DeviceBytes(const P7_GPU_OPROFILE *om, const P7_GPU_SEQDB *db, size_t *ret_bytes);
/* ---- serving: queries against a resident database, admitted as they arrive ----
 * Stands in for hmmpgmd's search loop (HMMER/src/hmmdmstr.c:701-790,
 * hmmdwrkr.c:535-644), which serves one query at a time: here up to `nslots`
 * slots (pipelines with their streams and device work areas) take queries
 * from a FIFO as they free up; with p7_gpu_server_SetCoalesce, a slot that
 * finds several queries waiting takes up to four of them as one
 * p7_gpu_PipelineBatch (their ungapped passes in one persistent launch per
 * kernel shape). The slots' work areas, for
 * batches of four, are set aside within `hbm_budget` bytes of device memory
 * (0: 90% of the free memory), which cuts the slot count; what is left admits
 * the queries in flight by their profiles' device tables (a query waits until
 * its tables fit; one alone always runs). Each query's result equals a lone
 * p7_gpu_Pipeline call with the same thresholds. Every call may come from any
 * host thread. */
typedef struct p7g_server_s P7_GPU_SERVER;
P7_GPU_SERVER *p7_gpu_server_Create(const P7_GPU_SEQDB *db, int nslots, size_t hbm_budget,
                                    double F1, double F2, double F3, int do_biasfilter);
int   p7_gpu_server_Slots(const P7_GPU_SERVER *srv);
/* Queries one slot may take as one batch: 1 (each query alone, the default)
 * to 4; turning it on makes the slots' batch work areas. */
int   p7_gpu_server_SetCoalesce(P7_GPU_SERVER *srv, int n);
/* The device memory set aside per slot, and what the budget leaves for the
 * profiles of the queries in flight. */
int   p7_gpu_server_Memory(const P7_GPU_SERVER *srv, size_t *ret_slot_bytes, size_t *ret_query_budget);
/* Queues a query and returns at once; `om` and `hits` (may be NULL) must stay
 * valid until the matching Wait returns. */
int   p7_gpu_server_Submit(P7_GPU_SERVER *srv, const P7_GPU_OPROFILE *om, P7_GPU_HITS *hits, int64_t *ret_ticket);
/* Blocks until the query is done: its pipeline counters, and the time from
 * Submit to completion in milliseconds. Returns the query's status. */
int   p7_gpu_server_Wait(P7_GPU_SERVER *srv, int64_t ticket, P7_GPU_STATS *ret_stats, double *ret_latency_ms);
/* Finishes the queued queries, then frees the slots. */
void  p7_gpu_server_Destroy(P7_GPU_SERVER *srv);

#ifdef __cplusplus
}
#endif
#endif /* P7_GPU_INCLUDED */
