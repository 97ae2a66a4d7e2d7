/* Internal host-side structures of libp7gpu. Not part of the C ABI. */
#ifndef P7G_INTERNAL_INCLUDED
#define P7G_INTERNAL_INCLUDED

#include <stdint.h>
#include <stddef.h>
#include "p7_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

#define P7G_MAXABET 20
#define P7G_MAXKP   29
#define P7G_LOG2    0.69314718055994529      /* eslCONST_LOG2  (easel.h:258) */
#define P7G_LOG2R   1.44269504088896341      /* eslCONST_LOG2R (easel.h:259) */

/* Transition order of a core HMM node (HMMER/src/hmmer.h:130-138). */
enum { H_MM = 0, H_MI = 1, H_MD = 2, H_IM = 3, H_II = 4, H_DM = 5, H_DD = 6 };
/* Transition order of a P7_PROFILE node (hmmer.h:222-231). */
enum { P_MM = 0, P_IM = 1, P_DM = 2, P_BM = 3, P_MD = 4, P_DD = 5, P_MI = 6, P_II = 7 };
#define P_NTRANS 8
/* Special states of P7_PROFILE (hmmer.h:204-218): [E,N,J,C][LOOP,MOVE]. */
enum { P_E = 0, P_N = 1, P_J = 2, P_C = 3 };
enum { P_LOOP = 0, P_MOVE = 1 };

/* Digital alphabet: the subset of ESL_ALPHABET the path uses (easel/esl_alphabet.c:225-300). */
typedef struct {
  int  type, K, Kp;
  char sym[32];
  char degen[P7G_MAXKP][P7G_MAXABET];   /* degen[x][y]: code x may be canonical y */
  int  ndegen[P7G_MAXKP];
  int8_t inmap[128];                    /* ascii -> code, -1 illegal, -2 ignored    */
} P7G_ALPHABET;

const P7G_ALPHABET *p7g_alphabet(int type);   /* NULL for unsupported types */

struct p7g_hmm_s {
  int     M, abctype, max_length;
  char    name[256];
  float  *t;          /* [(M+1)][7]   transitions, t[k*7 + H_*]         */
  float  *mat;        /* [(M+1)][K]   match emissions                   */
  float  *ins;        /* [(M+1)][K]   insert emissions                  */
  float   compo[P7G_MAXABET];
  float   evparam[6];
  int     has_compo, has_stats;
  /* per-node annotation, [0] unused, NUL-terminated at M+1: the consensus
   * residues (from the file's CONS column, else p7_hmm_SetConsensus's
   * rule), and the CS and RF lines when the file has them (else NULL) */
  char   *cons, *cs, *rf;
};

struct p7g_bg_s {
  const P7G_ALPHABET *abc;
  float  f[P7G_MAXABET];
  float  p1, omega;
  /* two-state filter HMM (esl_hmm.h): t[3][3] with the E column, eo[Kp][2], pi[3] */
  float  ft[2][3];
  float  eo[P7G_MAXKP][2];
  float  pi[3];
  int    filter_set;
};

struct p7g_profile_s {
  const P7G_ALPHABET *abc;
  int     M, L, mode, max_length;
  float   nj;
  float  *tsc;        /* [M][8]        P7_PROFILE order                   */
  float  *rsc;        /* [Kp][(M+1)][2]                                     */
  float   xsc[4][2];  /* [E,N,J,C][LOOP,MOVE]                              */
  float   evparam[6];
  float   compo[P7G_MAXABET];
  char   *cons, *cs, *rf;  /* the model's annotation, as P7_HMM holds it (or NULL) */
};

struct p7g_dev_profile_s;   /* device copy; defined in the HIP layer */

struct p7g_oprofile_s {
  const P7G_ALPHABET *abc;
  int       allocM, M, L, mode, max_length;
  float     nj;
  uint8_t  *rbv;      /* [Kp][M] */
  uint8_t   tbm_b, tec_b, tjb_b, base_b, bias_b;
  float     scale_b;
  int16_t  *rwv;      /* [Kp][M] */
  int16_t  *twv;      /* [8][M]  */
  int16_t   xw[4][2];
  float     scale_w;
  int16_t   base_w, ddbound_w;
  float     ncj_roundoff;
  float    *rfv;      /* [Kp][M] */
  float    *tfv;      /* [8][M]  */
  float     xf[4][2];
  float     evparam[6];
  float     compo[P7G_MAXABET];
  char     *cons, *cs, *rf;   /* the model's annotation for the alignment display (or NULL) */
  /* device side, built on first use, dropped when the tables change */
  struct p7g_dev_profile_s *dev;
  uint64_t  version;
};

struct p7g_omx_s {
  int     allocM, allocXL;
  int     M, L;
  float  *xmx;        /* [(allocXL+1)*6] */
  float   totscale;
  int     has_own_scales;
};

/* Per-length constants of the length model, computed on the host with the
 * same libm calls as the reference so that the device never evaluates
 * logf/roundf itself (p7_oprofile.c:1086-1121, p7_bg.c:191-199, 359-363, 473-484). */
typedef struct {
  int      Lmax;
  uint8_t *tjb_b;       /* unbiased_byteify(logf(3/(L+3)))                 */
  int16_t *xw_move;     /* wordify(logf(pmove)) for N, C, J                */
  float   *xf_move;     /* pmove                                            */
  float   *xf_loop;     /* ploop = 1 - pmove                                */
  float   *nullsc;      /* (float) L*log(p1) + log(1-p1), p1 = L/(L+1)      */
  float   *flen_a;      /* (float) L * logf(p1)                             */
  float   *flen_b;      /* logf(1 - p1)                                     */
} P7G_LENCFG;

/* host model code (host_model.c) */
int  p7g_lencfg_build(const P7_GPU_OPROFILE *om, int Lmax, P7G_LENCFG *lc);
void p7g_lencfg_free(P7G_LENCFG *lc);

/* statistics (host_evd.c) */
double p7g_gumbel_surv(double x, double mu, double lambda);
double p7g_exp_surv(double x, double mu, double lambda);
/* Smallest float s with surv(s) <= P, so "P(s) <= P" == "s >= threshold". */
float  p7g_gumbel_threshold(double mu, double lambda, double P);
float  p7g_exp_threshold(double mu, double lambda, double P);

/* hit-list scores (host_hits.c, p7_pipeline.c:743-850) */
typedef struct { float envsc, domcorr; int ienv, jenv; } P7G_DOMIN;       /* one domain: nats, 1-based coords */
typedef struct { float score, pre_score, sum_score; double lnP, pre_lnP, sum_lnP; int best; } P7G_TSCORE;
typedef struct { float bitscore, dombias; double lnP; } P7G_DSCORE;        /* bits, nats, log P */
float  p7g_flogsum(float a, float b);
double p7g_exp_logsurv(double x, double mu, double lambda);
float  p7g_null_one(int L);
/* n2parts (optional, nn2 values): the target's null2 scores summed per
 * envelope or multidomain region, in sequence order; the sequence bias is
 * their float sum. NULL: the domains' corrections (the same terms when no
 * region is multidomain). */
void   p7g_target_scores(int L, float fwdsc, float nullsc, float omega, float tau, float lambda,
                         const P7G_DOMIN *dom, int ndom, const float *n2parts, int nn2, P7G_TSCORE *ts, P7G_DSCORE *ds);

/* the hit list as the text writers take it (host_report.c): targets in rank
 * order, each with its rescored domains in sequence order */
typedef struct {
  float  bitscore, dombias;      /* bits, nats */
  double lnP;
  float  oasc;
  int    hmmfrom, hmmto, alifrom, alito, ienv, jenv;
  int    flags;                  /* bit 0 reported, bit 1 included */
  /* the alignment display's lines (p7_alidisplay_Create), N columns each,
   * or NULL without one; cs and rf NULL when the model has no such line */
  const char *model, *mline, *aseq, *ppline, *csline, *rfline;
  int    N;
} P7G_RDOM;
typedef struct {
  const char *name, *acc, *desc; /* acc, desc NULL when empty */
  int    L, flags, ndom, best, nreported, nincluded;
  float  score, pre_score, nexpected;
  double lnP;
  int    nregions, nclustered, noverlaps, nenvelopes;
  const P7G_RDOM *dom;
} P7G_RHIT;
typedef struct {
  const P7G_RHIT *hit;
  int    N;
  const char *qname, *qacc;      /* qacc NULL when empty */
  int    M;
  double Z, domZ;
  int    textw, show_accessions, show_alignments;
} P7G_RLIST;
size_t p7g_targets_text(const P7G_RLIST *l, char *buf, size_t cap);
size_t p7g_domains_text(const P7G_RLIST *l, char *buf, size_t cap);
size_t p7g_tabular_targets(const P7G_RLIST *l, int show_header, char *buf, size_t cap);
size_t p7g_tabular_domains(const P7G_RLIST *l, int show_header, char *buf, size_t cap);

/* error text */
void p7g_set_error(const char *fmt, ...);

/* device layer (hip/p7g_runtime.hip) */
void   p7g_dev_profile_free(struct p7g_dev_profile_s *d);
int    p7g_current_device(void);
int    p7g_device_free_bytes(size_t *ret_bytes);
/* Device bytes one pipeline's work area takes for a database of this size. */
size_t p7g_pipeline_work_bytes(int64_t nseq, int64_t nres);
/* Creates the pipeline's stream and allocates its work area for `db` now. */
int    p7g_pipeline_Reserve(P7_GPU_PIPELINE *pli, const P7_GPU_SEQDB *db);
int    p7g_pipeline_ReserveBatch(P7_GPU_PIPELINE *pli, const P7_GPU_SEQDB *db, int nprof);
void   p7g_pipeline_ResetStats(P7_GPU_PIPELINE *pli);

#ifdef __cplusplus
}
#endif
#endif
