/* Host-side model code of libp7gpu: digital alphabets, the HMMER3 ascii
 * reader, the null/bias model, profile configuration, conversion to the
 * quantised filter tables, and the per-length constants.
 *
 * Floating-point expressions keep the reference's operand types and order
 * (float vs double promotions, libm calls) so that the integer tables and
 * per-length constants come out identical; the file is compiled with
 * -ffp-contract=off.
 */
#include <math.h>
#include <float.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <strings.h>
#include <ctype.h>

#include "p7g_internal.h"

/*****************************************************************
 * Error text
 *****************************************************************/
static __thread char p7g_errbuf[512];

void
p7g_set_error(const char *fmt, ...)
{
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(p7g_errbuf, sizeof(p7g_errbuf), fmt, ap);
  va_end(ap);
}

const char *p7g_host_error(void) { return p7g_errbuf; }

/*****************************************************************
 * Alphabets (easel/esl_alphabet.c:225-300: symbol strings, K, Kp,
 * degeneracy sets; Kp-3 is the "any" code).
 *****************************************************************/
static P7G_ALPHABET abc_amino, abc_dna, abc_rna;
static int          abc_ready = 0;

static void
abc_init_one(P7G_ALPHABET *a, int type, const char *sym, int K, int Kp)
{
  int x, y;
  memset(a, 0, sizeof(*a));
  a->type = type; a->K = K; a->Kp = Kp;
  strcpy(a->sym, sym);
  for (x = 0; x < 128; x++) a->inmap[x] = -1;
  for (x = 0; x < Kp; x++) {
    a->inmap[(int) sym[x]] = (int8_t) x;
    a->inmap[tolower((unsigned char) sym[x])] = (int8_t) x;
  }
  for (x = 0; x < K; x++) { a->degen[x][x] = 1; a->ndegen[x] = 1; }
  for (y = 0; y < K; y++) a->degen[Kp-3][y] = 1;
  a->ndegen[Kp-3] = K;
  a->inmap['_'] = a->inmap['.'] = (int8_t) K;     /* gap synonyms */
}

static void
abc_set_degen(P7G_ALPHABET *a, char c, const char *set)
{
  int x = a->inmap[(int) c];
  const char *s;
  for (s = set; *s; s++) { a->degen[x][(int) a->inmap[(int) *s]] = 1; a->ndegen[x]++; }
}

static void
abc_equiv(P7G_ALPHABET *a, char c, char to)
{
  a->inmap[(int) c] = a->inmap[(int) to];
  a->inmap[tolower((unsigned char) c)] = a->inmap[(int) to];
}

const P7G_ALPHABET *
p7g_alphabet(int type)
{
  if (! abc_ready) {
    abc_init_one(&abc_amino, p7G_AMINO, "ACDEFGHIKLMNPQRSTVWY-BJZOUX*~", 20, 29);
    abc_set_degen(&abc_amino, 'B', "ND");
    abc_set_degen(&abc_amino, 'J', "IL");
    abc_set_degen(&abc_amino, 'Z', "QE");
    abc_set_degen(&abc_amino, 'U', "C");
    abc_set_degen(&abc_amino, 'O', "K");

    abc_init_one(&abc_dna, p7G_DNA, "ACGT-RYMKSWHBVDN*~", 4, 18);
    abc_equiv(&abc_dna, 'U', 'T');
    abc_equiv(&abc_dna, 'X', 'N');
    abc_init_one(&abc_rna, p7G_RNA, "ACGU-RYMKSWHBVDN*~", 4, 18);
    abc_equiv(&abc_rna, 'T', 'U');
    abc_equiv(&abc_rna, 'X', 'N');
    {
      P7G_ALPHABET *n[2] = { &abc_dna, &abc_rna };
      int i;
      for (i = 0; i < 2; i++) {
        const char *acgt = (i == 0) ? "ACGT" : "ACGU";
        char set[8];
        abc_set_degen(n[i], 'R', "AG");  abc_set_degen(n[i], 'Y', i ? "CU" : "CT");
        abc_set_degen(n[i], 'M', "AC");  abc_set_degen(n[i], 'K', i ? "GU" : "GT");
        abc_set_degen(n[i], 'S', "CG");  abc_set_degen(n[i], 'W', i ? "AU" : "AT");
        snprintf(set, sizeof(set), "AC%c", acgt[3]);       abc_set_degen(n[i], 'H', set);
        snprintf(set, sizeof(set), "CG%c", acgt[3]);       abc_set_degen(n[i], 'B', set);
        abc_set_degen(n[i], 'V', "ACG");
        snprintf(set, sizeof(set), "AG%c", acgt[3]);       abc_set_degen(n[i], 'D', set);
      }
    }
    abc_ready = 1;
  }
  switch (type) {
  case p7G_AMINO: return &abc_amino;
  case p7G_DNA:   return &abc_dna;
  case p7G_RNA:   return &abc_rna;
  default:        return NULL;
  }
}

/*****************************************************************
 * HMMER3/e and 3/f ascii save files (HMMER/src/p7_hmmfile.c:1294-1626).
 * Probabilities are stored as -log(p); '*' is p = 0.
 *****************************************************************/
typedef struct { char *buf; size_t n, pos; char *line; char *tokp; } LINEREADER;

static int
lr_next_line(LINEREADER *lr)        /* next non-blank line; p7G_EOF at end */
{
  for (;;) {
    char *s, *e;
    if (lr->pos >= lr->n) return p7G_EOF;
    s = lr->buf + lr->pos;
    e = memchr(s, '\n', lr->n - lr->pos);
    if (e) { *e = '\0'; lr->pos = (size_t) (e - lr->buf) + 1; }
    else   { lr->pos = lr->n; }
    lr->line = s; lr->tokp = s;
    while (*s && isspace((unsigned char) *s)) s++;
    if (*s) return p7G_OK;
  }
}

static char *
lr_token(LINEREADER *lr)
{
  char *s = lr->tokp, *t;
  while (*s && isspace((unsigned char) *s)) s++;
  if (! *s) return NULL;
  t = s;
  while (*s && ! isspace((unsigned char) *s)) s++;
  if (*s) { *s = '\0'; s++; }
  lr->tokp = s;
  return t;
}

static float
parse_prob(const char *tok)         /* expf(-1.0 * atof(tok)), p7_hmmfile.c:1564 */
{
  if (*tok == '*') return 0.0f;
  return expf((float) (-1.0 * atof(tok)));
}

void
p7_gpu_hmm_Destroy(P7_GPU_HMM *hmm)
{
  if (! hmm) return;
  free(hmm->t); free(hmm->mat); free(hmm->ins);
  free(hmm->cons); free(hmm->cs); free(hmm->rf);
  free(hmm);
}

static int
read_one_hmm(LINEREADER *lr, P7_GPU_HMM **ret)
{
  P7_GPU_HMM *hmm = calloc(1, sizeof(P7_GPU_HMM));
  int   fmt = 0;            /* 'e' or 'f' */
  int   k, x, K = 0;
  char *tag, *tok;
  const P7G_ALPHABET *abc = NULL;
  int   stats = 0;
  int   has_cons = 0, has_cs = 0, has_rf = 0;

  *ret = NULL;
  if (! hmm) return p7G_EMEM;
  hmm->abctype = 0; hmm->max_length = -1;
  for (x = 0; x < 6; x++) hmm->evparam[x] = -99999.0f;

  if (lr_next_line(lr) != p7G_OK) { free(hmm); return p7G_EOF; }
  tag = lr_token(lr);
  if      (strcmp(tag, "HMMER3/f") == 0) fmt = 'f';
  else if (strcmp(tag, "HMMER3/e") == 0) fmt = 'e';
  else { p7g_set_error("unsupported HMM save format %s", tag); free(hmm); return p7G_EFORMAT; }

  for (;;) {
    if (lr_next_line(lr) != p7G_OK) goto FORMAT;
    tag = lr_token(lr);
    if (strcmp(tag, "NAME") == 0) {
      if (! (tok = lr_token(lr))) goto FORMAT;
      snprintf(hmm->name, sizeof(hmm->name), "%s", tok);
    } else if (strcmp(tag, "CONS") == 0 || strcmp(tag, "CS") == 0 || strcmp(tag, "RF") == 0) {
      const int yes = (tok = lr_token(lr)) && strcasecmp(tok, "yes") == 0;
      if      (tag[0] == 'C' && tag[1] == 'O') has_cons = yes;
      else if (tag[0] == 'C')                  has_cs = yes;
      else                                     has_rf = yes;
    } else if (strcmp(tag, "LENG") == 0) {
      if (! (tok = lr_token(lr)) || (hmm->M = atoi(tok)) <= 0) goto FORMAT;
    } else if (strcmp(tag, "MAXL") == 0) {
      if (! (tok = lr_token(lr))) goto FORMAT;
      hmm->max_length = atoi(tok);
    } else if (strcmp(tag, "ALPH") == 0) {
      if (! (tok = lr_token(lr))) goto FORMAT;
      if      (strcasecmp(tok, "amino") == 0) hmm->abctype = p7G_AMINO;
      else if (strcasecmp(tok, "dna")   == 0) hmm->abctype = p7G_DNA;
      else if (strcasecmp(tok, "rna")   == 0) hmm->abctype = p7G_RNA;
      else { p7g_set_error("unsupported alphabet %s", tok); goto ERR; }
    } else if (strcmp(tag, "STATS") == 0) {
      char *t1 = lr_token(lr), *t2 = lr_token(lr), *t3 = lr_token(lr), *t4 = lr_token(lr);
      if (! t4 || strcasecmp(t1, "LOCAL") != 0) goto FORMAT;
      if      (strcasecmp(t2, "MSV")     == 0) { hmm->evparam[p7G_MMU]  = atof(t3); hmm->evparam[p7G_MLAMBDA] = atof(t4); stats |= 1; }
      else if (strcasecmp(t2, "VITERBI") == 0) { hmm->evparam[p7G_VMU]  = atof(t3); hmm->evparam[p7G_VLAMBDA] = atof(t4); stats |= 2; }
      else if (strcasecmp(t2, "FORWARD") == 0) { hmm->evparam[p7G_FTAU] = atof(t3); hmm->evparam[p7G_FLAMBDA] = atof(t4); stats |= 4; }
      else goto FORMAT;
    } else if (strcmp(tag, "HMM") == 0) {
      break;
    }
  }
  if (stats != 0 && stats != 7) { p7g_set_error("missing STATS lines in %s", hmm->name); goto ERR; }
  hmm->has_stats = (stats == 7);
  if (hmm->M <= 0 || hmm->abctype == 0) goto FORMAT;
  abc = p7g_alphabet(hmm->abctype);
  K   = abc->K;

  hmm->t   = calloc((size_t) (hmm->M+1) * 7, sizeof(float));
  hmm->mat = calloc((size_t) (hmm->M+1) * K, sizeof(float));
  hmm->ins = calloc((size_t) (hmm->M+1) * K, sizeof(float));
  hmm->cons = calloc((size_t) hmm->M + 2, 1);
  if (has_cs) hmm->cs = calloc((size_t) hmm->M + 2, 1);
  if (has_rf) hmm->rf = calloc((size_t) hmm->M + 2, 1);
  if (! hmm->t || ! hmm->mat || ! hmm->ins || ! hmm->cons || (has_cs && ! hmm->cs) || (has_rf && ! hmm->rf)) {
    p7_gpu_hmm_Destroy(hmm); return p7G_EMEM;
  }
  hmm->cons[0] = ' ';
  if (hmm->cs) hmm->cs[0] = ' ';
  if (hmm->rf) hmm->rf[0] = ' ';

  if (lr_next_line(lr) != p7G_OK) goto FORMAT;            /* m->m m->i ... header */
  if (lr_next_line(lr) != p7G_OK) goto FORMAT;
  tok = lr_token(lr);
  if (tok && strcmp(tok, "COMPO") == 0) {
    for (x = 0; x < K; x++) { if (! (tok = lr_token(lr))) goto FORMAT; hmm->compo[x] = parse_prob(tok); }
    hmm->has_compo = 1;
    if (lr_next_line(lr) != p7G_OK) goto FORMAT;
    tok = lr_token(lr);
  }
  if (! tok) goto FORMAT;
  hmm->ins[0] = parse_prob(tok);                            /* node 0 inserts */
  for (x = 1; x < K; x++) { if (! (tok = lr_token(lr))) goto FORMAT; hmm->ins[x] = parse_prob(tok); }
  if (lr_next_line(lr) != p7G_OK) goto FORMAT;              /* node 0 transitions */
  for (x = 0; x < 7; x++) { if (! (tok = lr_token(lr))) goto FORMAT; hmm->t[x] = parse_prob(tok); }

  for (k = 1; k <= hmm->M; k++) {
    int nfields = (fmt == 'f') ? 5 : 4;       /* MAP, CONS, RF, [MM,] CS */
    if (lr_next_line(lr) != p7G_OK) goto FORMAT;
    if (! (tok = lr_token(lr)) || atoi(tok) != k) goto FORMAT;
    for (x = 0; x < K; x++) { if (! (tok = lr_token(lr))) goto FORMAT; hmm->mat[(size_t) k*K + x] = parse_prob(tok); }
    /* MAP, CONS, RF, [MM,] CS: one token each */
    for (x = 0; x < nfields; x++) {
      if (! (tok = lr_token(lr))) goto FORMAT;
      if (x == 1 && has_cons) hmm->cons[k] = tok[0];
      if (x == 2 && hmm->rf) hmm->rf[k] = tok[0];
      if (x == nfields - 1 && hmm->cs) hmm->cs[k] = tok[0];
    }
    if (! has_cons) {                         /* p7_hmm_SetConsensus (p7_hmm.c:704-734) */
      const float mthresh = (hmm->abctype == p7G_AMINO) ? 0.5f : 0.9f;
      int best = 0;
      for (x = 1; x < K; x++) if (hmm->mat[(size_t) k*K + x] > hmm->mat[(size_t) k*K + best]) best = x;
      hmm->cons[k] = (char) ((hmm->mat[(size_t) k*K + best] >= mthresh) ? toupper(abc->sym[best]) : tolower(abc->sym[best]));
    }
    if (lr_next_line(lr) != p7G_OK) goto FORMAT;
    for (x = 0; x < K; x++) { if (! (tok = lr_token(lr))) goto FORMAT; hmm->ins[(size_t) k*K + x] = parse_prob(tok); }
    if (lr_next_line(lr) != p7G_OK) goto FORMAT;
    for (x = 0; x < 7; x++) { if (! (tok = lr_token(lr))) goto FORMAT; hmm->t[(size_t) k*7 + x] = parse_prob(tok); }
  }
  if (lr_next_line(lr) != p7G_OK) goto FORMAT;
  if (! (tok = lr_token(lr)) || strcmp(tok, "//") != 0) goto FORMAT;
  *ret = hmm;
  return p7G_OK;

 FORMAT:
  p7g_set_error("HMM file format error in model %s, line: %.60s", hmm->name, lr->line ? lr->line : "");
 ERR:
  p7_gpu_hmm_Destroy(hmm);
  return p7G_EFORMAT;
}

int
p7_gpu_hmmfile_Read(const char *path, int which, P7_GPU_HMM **ret_hmm)
{
  FILE *fp;
  long  n;
  LINEREADER lr;
  P7_GPU_HMM *hmm = NULL;
  int   i, status = p7G_OK;

  *ret_hmm = NULL;
  if (! (fp = fopen(path, "rb"))) { p7g_set_error("can't open HMM file %s", path); return p7G_ENOTFOUND; }
  fseek(fp, 0, SEEK_END); n = ftell(fp); fseek(fp, 0, SEEK_SET);
  memset(&lr, 0, sizeof(lr));
  lr.buf = malloc((size_t) n + 1);
  if (! lr.buf) { fclose(fp); return p7G_EMEM; }
  lr.n = fread(lr.buf, 1, (size_t) n, fp);
  lr.buf[lr.n] = '\0';
  fclose(fp);

  for (i = 0; i <= which; i++) {
    if (hmm) { p7_gpu_hmm_Destroy(hmm); hmm = NULL; }
    status = read_one_hmm(&lr, &hmm);
    if (status != p7G_OK) break;
  }
  free(lr.buf);
  if (status == p7G_EOF) p7g_set_error("HMM file %s holds fewer than %d models", path, which+1);
  if (status == p7G_OK) *ret_hmm = hmm;
  return status;
}

int
p7_gpu_hmm_Info(const P7_GPU_HMM *hmm, int *ret_M, int *ret_abctype, char *name, int namelen)
{
  if (! hmm) return p7G_EINVAL;
  if (ret_M)       *ret_M = hmm->M;
  if (ret_abctype) *ret_abctype = hmm->abctype;
  if (name && namelen > 0) snprintf(name, (size_t) namelen, "%s", hmm->name);
  return p7G_OK;
}

/*****************************************************************
 * Null model and bias-composition filter HMM
 * (HMMER/src/p7_bg.c:55-90, 191-199, 359-363, 431-453;
 *  HMMER/src/hmmer.c:163-187; easel/esl_hmm.c:116-153)
 *****************************************************************/
static void
amino_frequencies(float *f)
{
  static const float af[20] = {
    0.0787945f, 0.0151600f, 0.0535222f, 0.0668298f, 0.0397062f, 0.0695071f, 0.0229198f,
    0.0590092f, 0.0594422f, 0.0963728f, 0.0237718f, 0.0414386f, 0.0482904f, 0.0395639f,
    0.0540978f, 0.0683364f, 0.0540687f, 0.0673417f, 0.0114135f, 0.0304133f };
  memcpy(f, af, sizeof(af));
}

P7_GPU_BG *
p7_gpu_bg_Create(int abctype)
{
  const P7G_ALPHABET *abc = p7g_alphabet(abctype);
  P7_GPU_BG *bg;
  int x;
  if (! abc) { p7g_set_error("unsupported alphabet type %d", abctype); return NULL; }
  if (! (bg = calloc(1, sizeof(P7_GPU_BG)))) return NULL;
  bg->abc = abc;
  if (abctype == p7G_AMINO) amino_frequencies(bg->f);
  else for (x = 0; x < abc->K; x++) bg->f[x] = 1. / (float) abc->K;
  bg->p1    = 350./351.;
  bg->omega = 1./256.;
  return bg;
}

void p7_gpu_bg_Destroy(P7_GPU_BG *bg) { free(bg); }

int
p7_gpu_bg_SetLength(P7_GPU_BG *bg, int L)
{
  bg->p1 = (float) L / (float) (L+1);
  bg->ft[0][0] = bg->p1;
  bg->ft[0][1] = 1.0f - bg->p1;
  return p7G_OK;
}

int
p7_gpu_bg_NullOne(const P7_GPU_BG *bg, const uint8_t *dsq, int L, float *ret_sc)
{
  (void) dsq;
  *ret_sc = (float) L * log(bg->p1) + log(1.-bg->p1);
  return p7G_OK;
}

int
p7_gpu_bg_SetFilter(P7_GPU_BG *bg, int M, const float *compo)
{
  const P7G_ALPHABET *abc = bg->abc;
  float L0 = 400.0;
  float L1 = (float) M / 8.0;
  float e[2][P7G_MAXABET];
  int   K = abc->K, Kp = abc->Kp;
  int   k, x, y;

  bg->ft[0][0] =   L0 / (L0+1.0f);
  bg->ft[0][1] = 1.0f / (L0+1.0f);
  bg->ft[0][2] = 1.0f;
  bg->ft[1][0] = 1.0f / (L1+1.0f);
  bg->ft[1][1] =   L1 / (L1+1.0f);
  bg->ft[1][2] = 1.0f;
  for (x = 0; x < K; x++) { e[0][x] = bg->f[x]; e[1][x] = compo[x]; }
  bg->pi[0] = 0.999;
  bg->pi[1] = 0.001;
  bg->pi[2] = 0.0;

  /* esl_hmm_Configure(fhmm, bg->f): odds ratios, degenerate codes averaged */
  for (x = 0; x < K; x++)
    for (k = 0; k < 2; k++) bg->eo[x][k] = e[k][x] / bg->f[x];
  for (k = 0; k < 2; k++) { bg->eo[K][k] = 1.0; bg->eo[Kp-2][k] = 1.0; bg->eo[Kp-1][k] = 1.0; }
  for (x = K+1; x <= Kp-3; x++)
    for (k = 0; k < 2; k++) {
      float denom = 0.0f;
      bg->eo[x][k] = 0.0f;
      for (y = 0; y < K; y++)
        if (abc->degen[x][y]) { bg->eo[x][k] += e[k][y]; denom += bg->f[y]; }
      bg->eo[x][k] = ((denom > 0.0f) ? bg->eo[x][k] / denom : 0.0f);
    }
  bg->filter_set = 1;
  return p7G_OK;
}

/*****************************************************************
 * Profile configuration (HMMER/src/modelconfig.c:49-235,
 * p7_profile.c:52-103, p7_hmm.c:1271-1289).
 *****************************************************************/
void
p7_gpu_profile_Destroy(P7_GPU_PROFILE *gm)
{
  if (! gm) return;
  free(gm->tsc); free(gm->rsc);
  free(gm->cons); free(gm->cs); free(gm->rf);
  free(gm);
}

#define TSC(gm,k,s)  ((gm)->tsc[(size_t)(k) * P_NTRANS + (s)])
#define MSC(gm,k,x)  ((gm)->rsc[((size_t)(x) * ((gm)->M+1) + (k)) * 2 + 0])
#define ISC(gm,k,x)  ((gm)->rsc[((size_t)(x) * ((gm)->M+1) + (k)) * 2 + 1])

int
p7_gpu_profile_ReconfigLength(P7_GPU_PROFILE *gm, int L)
{
  float ploop, pmove;
  pmove = (2.0f + gm->nj) / ((float) L + 2.0f + gm->nj);
  ploop = 1.0f - pmove;
  gm->xsc[P_N][P_LOOP] = gm->xsc[P_C][P_LOOP] = gm->xsc[P_J][P_LOOP] = log(ploop);
  gm->xsc[P_N][P_MOVE] = gm->xsc[P_C][P_MOVE] = gm->xsc[P_J][P_MOVE] = log(pmove);
  gm->L = L;
  return p7G_OK;
}

int
p7_gpu_ProfileConfig(const P7_GPU_HMM *hmm, const P7_GPU_BG *bg, int L, int mode, P7_GPU_PROFILE **ret_gm)
{
  const P7G_ALPHABET *abc = p7g_alphabet(hmm->abctype);
  P7_GPU_PROFILE *gm;
  int   M = hmm->M, K, Kp, k, x;
  float sc[P7G_MAXKP];
  float Z;
  float *occ;

  *ret_gm = NULL;
  if (! abc || bg->abc != abc) { p7g_set_error("HMM and null model alphabets differ"); return p7G_EINCOMPAT; }
  if (mode < p7G_LOCAL || mode > p7G_UNILOCAL) { p7g_set_error("bad profile mode %d", mode); return p7G_EINVAL; }
  K = abc->K; Kp = abc->Kp;

  if (! (gm = calloc(1, sizeof(P7_GPU_PROFILE)))) return p7G_EMEM;
  gm->abc = abc; gm->M = M; gm->mode = mode; gm->max_length = hmm->max_length;
  gm->tsc = malloc(sizeof(float) * (size_t) M * P_NTRANS);
  gm->rsc = malloc(sizeof(float) * (size_t) Kp * (M+1) * 2);
  occ     = malloc(sizeof(float) * (size_t) (M+1));
  if (! gm->tsc || ! gm->rsc || ! occ) { free(occ); p7_gpu_profile_Destroy(gm); return p7G_EMEM; }

  /* edges that p7_profile_Create initialises */
  for (x = 0; x < P_NTRANS; x++) gm->tsc[x] = -INFINITY;
  if (M > 1) { TSC(gm, 1, P_DM) = -INFINITY; TSC(gm, 1, P_DD) = -INFINITY; }
  for (x = 0; x < Kp; x++) { MSC(gm, 0, x) = -INFINITY; ISC(gm, 0, x) = -INFINITY; }
  for (k = 0; k <= M; k++) { MSC(gm, k, K) = ISC(gm, k, K) = -INFINITY; MSC(gm, k, Kp-1) = ISC(gm, k, Kp-1) = -INFINITY; }
  for (x = 0; x < 6; x++) gm->evparam[x] = hmm->evparam[x];
  /* the annotation goes along (modelconfig.c:80-81) */
  if (hmm->cons) gm->cons = strdup(hmm->cons);
  if (hmm->cs)   gm->cs   = strdup(hmm->cs);
  if (hmm->rf)   gm->rf   = strdup(hmm->rf);
  for (x = 0; x < P7G_MAXABET; x++) gm->compo[x] = hmm->has_compo ? hmm->compo[x] : -1.0f;

  /* entry scores */
  if (mode == p7G_LOCAL || mode == p7G_UNILOCAL) {
    occ[0] = 0.;
    occ[1] = hmm->t[0*7 + H_MI] + hmm->t[0*7 + H_MM];
    for (k = 2; k <= M; k++)
      occ[k] = occ[k-1] * (hmm->t[(k-1)*7 + H_MM] + hmm->t[(k-1)*7 + H_MI]) +
               (1.0-occ[k-1]) * hmm->t[(k-1)*7 + H_DM];
    Z = 0.;
    for (k = 1; k <= M; k++) Z += occ[k] * (float) (M-k+1);
    for (k = 1; k <= M; k++) TSC(gm, k-1, P_BM) = log(occ[k] / Z);
  } else {
    Z = log(hmm->t[0*7 + H_MD]);
    TSC(gm, 0, P_BM) = log(1.0 - hmm->t[0*7 + H_MD]);
    for (k = 1; k < M; k++) {
      TSC(gm, k, P_BM) = Z + log(hmm->t[k*7 + H_DM]);
      Z += log(hmm->t[k*7 + H_DD]);
    }
  }
  free(occ);

  if (mode == p7G_LOCAL || mode == p7G_GLOCAL) {
    gm->xsc[P_E][P_MOVE] = -P7G_LOG2;
    gm->xsc[P_E][P_LOOP] = -P7G_LOG2;
    gm->nj = 1.0f;
  } else {
    gm->xsc[P_E][P_MOVE] = 0.0f;
    gm->xsc[P_E][P_LOOP] = -INFINITY;
    gm->nj = 0.0f;
  }

  for (k = 1; k < M; k++) {
    const float *t = hmm->t + (size_t) k*7;
    TSC(gm, k, P_MM) = log(t[H_MM]);
    TSC(gm, k, P_MI) = log(t[H_MI]);
    TSC(gm, k, P_MD) = log(t[H_MD]);
    TSC(gm, k, P_IM) = log(t[H_IM]);
    TSC(gm, k, P_II) = log(t[H_II]);
    TSC(gm, k, P_DM) = log(t[H_DM]);
    TSC(gm, k, P_DD) = log(t[H_DD]);
  }

  /* match emissions, with expected scores for degenerate codes
   * (esl_abc_FExpectScVec, easel/esl_alphabet.c:1401-1415) */
  sc[K]    = -INFINITY;
  sc[Kp-2] = -INFINITY;
  sc[Kp-1] = -INFINITY;
  for (k = 1; k <= M; k++) {
    for (x = 0; x < K; x++) sc[x] = log((double) hmm->mat[(size_t) k*K + x] / bg->f[x]);
    for (x = K+1; x <= Kp-3; x++) {
      float result = 0., denom = 0.;
      int   y;
      for (y = 0; y < K; y++)
        if (abc->degen[x][y]) { result += sc[y] * bg->f[y]; denom += bg->f[y]; }
      result /= denom;
      sc[x] = result;
    }
    for (x = 0; x < Kp; x++) MSC(gm, k, x) = sc[x];
  }

  /* insert emissions are hardwired to 0 (modelconfig.c:150-166) */
  for (x = 0; x < Kp; x++) {
    for (k = 1; k < M; k++) ISC(gm, k, x) = 0.0f;
    ISC(gm, M, x) = -INFINITY;
  }
  for (k = 1; k <= M; k++) { ISC(gm, k, K) = -INFINITY; ISC(gm, k, Kp-2) = -INFINITY; ISC(gm, k, Kp-1) = -INFINITY; }

  gm->L = 0;
  p7_gpu_profile_ReconfigLength(gm, L);
  *ret_gm = gm;
  return p7G_OK;
}

int
p7_gpu_profile_GetScores(const P7_GPU_PROFILE *gm, float *tsc, float *rsc, float *xsc)
{
  int s, t;
  if (tsc) memcpy(tsc, gm->tsc, sizeof(float) * (size_t) gm->M * P_NTRANS);
  if (rsc) memcpy(rsc, gm->rsc, sizeof(float) * (size_t) gm->abc->Kp * (gm->M+1) * 2);
  if (xsc) for (s = 0; s < 4; s++) for (t = 0; t < 2; t++) xsc[s*2+t] = gm->xsc[s][t];
  return p7G_OK;
}

/*****************************************************************
 * Conversion to the quantised filter tables (p7_oprofile.c:649-985,
 * 1005-1127). Layout is unstriped: node k of residue x at [x*M + k-1].
 *****************************************************************/

/* Cephes single-precision exp (S. L. Moshier, Cephes Math Library, expf.c):
 * range reduction x = k*ln2 + f with ln2 split in two constants, degree-5
 * minimax polynomial on [-0.5 ln2, 0.5 ln2], 2^k built from the exponent
 * field. The four-lane SSE form used by the reference (easel/esl_sse.c:182-236)
 * evaluates the same sequence of single-precision operations; this is that
 * sequence for one lane, with results below FLT_MIN flushed to zero as the
 * reference's FTZ mode does (impl_sse.h:357-377). */
static float
cephes_expf(float x)
{
  const float maxlogf =  88.3762626647949f;
  const float minlogf = -88.3762626647949f;
  float fx, tmp, z, y, pow2k;
  int   k;
  uint32_t bits;

  if (x >  maxlogf) return INFINITY;
  if (x <= minlogf) return 0.0f;
  fx  = x * (float) P7G_LOG2R;
  fx  = fx + 0.5f;
  k   = (int) fx;
  tmp = (float) k;
  if (tmp > fx) tmp = tmp - 1.0f;
  fx  = tmp;
  k   = (int) fx;
  tmp = fx * 0.693359375f;
  z   = fx * -2.12194440e-4f;
  x   = x - tmp;
  x   = x - z;
  z   = x * x;
  y   = 1.9875691500E-4f;
  y   = y * x; y = y + 1.3981999507E-3f;
  y   = y * x; y = y + 8.3334519073E-3f;
  y   = y * x; y = y + 4.1665795894E-2f;
  y   = y * x; y = y + 1.6666665459E-1f;
  y   = y * x; y = y + 5.0000001201E-1f;
  y   = y * z;
  y   = y + x;
  y   = y + 1.0f;
  bits = (uint32_t) (k + 127) << 23;
  memcpy(&pow2k, &bits, sizeof(float));
  y = y * pow2k;
  if (y < FLT_MIN) y = 0.0f;
  return y;
}

static uint8_t
biased_byteify(const P7_GPU_OPROFILE *om, float sc)
{
  uint8_t b;
  sc = -1.0f * roundf(om->scale_b * sc);
  b  = (sc > 255 - om->bias_b) ? 255 : (uint8_t) ((int) sc + om->bias_b);
  return b;
}

static uint8_t
unbiased_byteify(const P7_GPU_OPROFILE *om, float sc)
{
  uint8_t b;
  sc = -1.0f * roundf(om->scale_b * sc);
  b  = (sc > 255.) ? 255 : (uint8_t) sc;
  return b;
}

static int16_t
wordify(const P7_GPU_OPROFILE *om, float sc)
{
  sc = roundf(om->scale_w * sc);
  if      (sc >=  32767.0) return  32767;
  else if (sc <= -32768.0) return -32768;
  else return (int16_t) sc;
}

static void
om_touch(P7_GPU_OPROFILE *om)
{
  if (om->dev) { p7g_dev_profile_free(om->dev); om->dev = NULL; }
  om->version++;
}

P7_GPU_OPROFILE *
p7_gpu_oprofile_Create(int allocM, int abctype)
{
  const P7G_ALPHABET *abc = p7g_alphabet(abctype);
  P7_GPU_OPROFILE *om;
  int x;
  if (! abc || allocM <= 0) { p7g_set_error("bad oprofile allocation (M=%d, alphabet %d)", allocM, abctype); return NULL; }
  if (! (om = calloc(1, sizeof(P7_GPU_OPROFILE)))) return NULL;
  om->abc = abc; om->allocM = allocM;
  om->rbv = calloc((size_t) abc->Kp * allocM, sizeof(uint8_t));
  om->rwv = calloc((size_t) abc->Kp * allocM, sizeof(int16_t));
  om->twv = calloc((size_t) 8 * allocM, sizeof(int16_t));
  om->rfv = calloc((size_t) abc->Kp * allocM, sizeof(float));
  om->tfv = calloc((size_t) 8 * allocM, sizeof(float));
  if (! om->rbv || ! om->rwv || ! om->twv || ! om->rfv || ! om->tfv) { p7_gpu_oprofile_Destroy(om); return NULL; }
  for (x = 0; x < 6; x++) om->evparam[x] = -99999.0f;
  for (x = 0; x < P7G_MAXABET; x++) om->compo[x] = -1.0f;
  om->max_length = -1;
  return om;
}

void
p7_gpu_oprofile_Destroy(P7_GPU_OPROFILE *om)
{
  if (! om) return;
  if (om->dev) p7g_dev_profile_free(om->dev);
  free(om->rbv); free(om->rwv); free(om->twv); free(om->rfv); free(om->tfv);
  free(om->cons); free(om->cs); free(om->rf);
  free(om);
}

int
p7_gpu_oprofile_Convert(const P7_GPU_PROFILE *gm, P7_GPU_OPROFILE *om)
{
  int   M = gm->M, Kp = gm->abc->Kp, K = gm->abc->K;
  int   x, k, t, z;
  float max = 0.0f;
  int   ddtmp;

  if (gm->abc != om->abc) { p7g_set_error("alphabets of the two profiles don't match"); return p7G_EINVAL; }
  if (M > om->allocM)     { p7g_set_error("oprofile is too small"); return p7G_EINVAL; }
  om_touch(om);
  om->mode = gm->mode; om->L = gm->L; om->M = M; om->nj = gm->nj; om->max_length = gm->max_length;

  /* MSVFilter part (mf_conversion, p7_oprofile.c:764-804) */
  for (x = 0; x < K; x++)
    for (z = 0; z < (M+1)*2; z++) max = (gm->rsc[(size_t) x*(M+1)*2 + z] > max) ? gm->rsc[(size_t) x*(M+1)*2 + z] : max;
  om->scale_b = 3.0 / P7G_LOG2;
  om->base_b  = 190;
  om->bias_b  = unbiased_byteify(om, -1.0 * max);
  for (x = 0; x < Kp; x++)
    for (k = 1; k <= M; k++) om->rbv[(size_t) x*M + k-1] = biased_byteify(om, MSC(gm, k, x));
  om->tbm_b = unbiased_byteify(om, logf(2.0f / ((float) gm->M * (float) (gm->M+1))));
  om->tec_b = unbiased_byteify(om, logf(0.5f));
  om->tjb_b = unbiased_byteify(om, logf(3.0f / (float) (gm->L+3)));

  /* ViterbiFilter part (vf_conversion, p7_oprofile.c:817-912) */
  om->scale_w = 500.0 / P7G_LOG2;
  om->base_w  = 12000;
  for (x = 0; x < Kp; x++)
    for (k = 1; k <= M; k++) om->rwv[(size_t) x*M + k-1] = wordify(om, MSC(gm, k, x));
  for (k = 1; k <= M; k++) {
    for (t = p7G_BM; t <= p7G_DD; t++) {
      int     tg = 0, kb = 0;
      int16_t maxval = 0, val;
      switch (t) {
      case p7G_BM: tg = P_BM; kb = k-1; maxval =  0; break;
      case p7G_MM: tg = P_MM; kb = k-1; maxval =  0; break;
      case p7G_IM: tg = P_IM; kb = k-1; maxval =  0; break;
      case p7G_DM: tg = P_DM; kb = k-1; maxval =  0; break;
      case p7G_MD: tg = P_MD; kb = k;   maxval =  0; break;
      case p7G_MI: tg = P_MI; kb = k;   maxval =  0; break;
      case p7G_II: tg = P_II; kb = k;   maxval = -1; break;
      case p7G_DD: tg = P_DD; kb = k;   maxval = 32767; break;   /* DD: no clamp */
      }
      val = (kb < M) ? wordify(om, TSC(gm, kb, tg)) : -32768;
      om->twv[(size_t) t*M + k-1] = (val <= maxval) ? val : maxval;
    }
  }
  om->xw[p7G_XE][p7G_LOOP] = wordify(om, gm->xsc[P_E][P_LOOP]);
  om->xw[p7G_XE][p7G_MOVE] = wordify(om, gm->xsc[P_E][P_MOVE]);
  om->xw[p7G_XN][p7G_MOVE] = wordify(om, gm->xsc[P_N][P_MOVE]);
  om->xw[p7G_XN][p7G_LOOP] = 0;
  om->xw[p7G_XC][p7G_MOVE] = wordify(om, gm->xsc[P_C][P_MOVE]);
  om->xw[p7G_XC][p7G_LOOP] = 0;
  om->xw[p7G_XJ][p7G_MOVE] = wordify(om, gm->xsc[P_J][P_MOVE]);
  om->xw[p7G_XJ][p7G_LOOP] = 0;
  om->ncj_roundoff = 0.0;
  om->ddbound_w = -32768;
  for (k = 2; k < M-1; k++) {
    ddtmp  = (int) wordify(om, TSC(gm, k,   P_DD));
    ddtmp += (int) wordify(om, TSC(gm, k+1, P_DM));
    ddtmp -= (int) wordify(om, TSC(gm, k+1, P_BM));
    om->ddbound_w = (om->ddbound_w > ddtmp) ? om->ddbound_w : ddtmp;
  }

  /* Forward/Backward part (fb_conversion, p7_oprofile.c:920-985) */
  for (x = 0; x < Kp; x++)
    for (k = 1; k <= M; k++) om->rfv[(size_t) x*M + k-1] = cephes_expf(MSC(gm, k, x));
  for (k = 1; k <= M; k++) {
    for (t = p7G_BM; t <= p7G_DD; t++) {
      int tg = 0, kb = 0;
      switch (t) {
      case p7G_BM: tg = P_BM; kb = k-1; break;
      case p7G_MM: tg = P_MM; kb = k-1; break;
      case p7G_IM: tg = P_IM; kb = k-1; break;
      case p7G_DM: tg = P_DM; kb = k-1; break;
      case p7G_MD: tg = P_MD; kb = k;   break;
      case p7G_MI: tg = P_MI; kb = k;   break;
      case p7G_II: tg = P_II; kb = k;   break;
      case p7G_DD: tg = P_DD; kb = k;   break;
      }
      om->tfv[(size_t) t*M + k-1] = cephes_expf((kb < M) ? TSC(gm, kb, tg) : -INFINITY);
    }
  }
  om->xf[p7G_XE][p7G_LOOP] = expf(gm->xsc[P_E][P_LOOP]);
  om->xf[p7G_XE][p7G_MOVE] = expf(gm->xsc[P_E][P_MOVE]);
  om->xf[p7G_XN][p7G_LOOP] = expf(gm->xsc[P_N][P_LOOP]);
  om->xf[p7G_XN][p7G_MOVE] = expf(gm->xsc[P_N][P_MOVE]);
  om->xf[p7G_XC][p7G_LOOP] = expf(gm->xsc[P_C][P_LOOP]);
  om->xf[p7G_XC][p7G_MOVE] = expf(gm->xsc[P_C][P_MOVE]);
  om->xf[p7G_XJ][p7G_LOOP] = expf(gm->xsc[P_J][P_LOOP]);
  om->xf[p7G_XJ][p7G_MOVE] = expf(gm->xsc[P_J][P_MOVE]);

  for (z = 0; z < 6; z++) om->evparam[z] = gm->evparam[z];
  for (z = 0; z < P7G_MAXABET; z++) om->compo[z] = gm->compo[z];
  free(om->cons); free(om->cs); free(om->rf);
  om->cons = gm->cons ? strdup(gm->cons) : NULL;
  om->cs   = gm->cs   ? strdup(gm->cs)   : NULL;
  om->rf   = gm->rf   ? strdup(gm->rf)   : NULL;
  return p7G_OK;
}

int
p7_gpu_oprofile_Load(P7_GPU_OPROFILE *om, const P7_GPU_TABLES *t)
{
  int M = t->M, Kp;
  if (! t || t->abctype != om->abc->type || t->Kp != om->abc->Kp) { p7g_set_error("tables don't match the oprofile alphabet"); return p7G_EINCOMPAT; }
  if (M <= 0 || M > om->allocM) { p7g_set_error("tables have M=%d, oprofile holds %d", M, om->allocM); return p7G_EINVAL; }
  Kp = t->Kp;
  om_touch(om);
  om->M = M; om->L = t->L; om->mode = t->mode; om->nj = t->nj;
  memcpy(om->rbv, t->rbv, (size_t) Kp * M);
  memcpy(om->rwv, t->rwv, sizeof(int16_t) * (size_t) Kp * M);
  memcpy(om->twv, t->twv, sizeof(int16_t) * (size_t) 8 * M);
  memcpy(om->rfv, t->rfv, sizeof(float) * (size_t) Kp * M);
  memcpy(om->tfv, t->tfv, sizeof(float) * (size_t) 8 * M);
  om->tbm_b = t->tbm_b; om->tec_b = t->tec_b; om->tjb_b = t->tjb_b; om->base_b = t->base_b; om->bias_b = t->bias_b;
  om->scale_b = t->scale_b;
  memcpy(om->xw, t->xw, sizeof(om->xw));
  om->scale_w = t->scale_w; om->base_w = t->base_w; om->ddbound_w = t->ddbound_w;
  memcpy(om->xf, t->xf, sizeof(om->xf));
  memcpy(om->evparam, t->evparam, sizeof(om->evparam));
  memcpy(om->compo, t->compo, sizeof(om->compo));
  return p7G_OK;
}

int
p7_gpu_oprofile_GetTables(const P7_GPU_OPROFILE *om, P7_GPU_TABLES *t)
{
  memset(t, 0, sizeof(*t));
  t->M = om->M; t->Kp = om->abc->Kp; t->abctype = om->abc->type;
  t->rbv = om->rbv; t->tbm_b = om->tbm_b; t->tec_b = om->tec_b; t->tjb_b = om->tjb_b;
  t->base_b = om->base_b; t->bias_b = om->bias_b; t->scale_b = om->scale_b;
  t->rwv = om->rwv; t->twv = om->twv;
  memcpy(t->xw, om->xw, sizeof(t->xw));
  t->scale_w = om->scale_w; t->base_w = om->base_w; t->ddbound_w = om->ddbound_w;
  t->rfv = om->rfv; t->tfv = om->tfv;
  memcpy(t->xf, om->xf, sizeof(t->xf));
  t->L = om->L; t->mode = om->mode; t->nj = om->nj;
  memcpy(t->evparam, om->evparam, sizeof(t->evparam));
  memcpy(t->compo, om->compo, sizeof(t->compo));
  return p7G_OK;
}

int
p7_gpu_oprofile_ReconfigMSVLength(P7_GPU_OPROFILE *om, int L)
{
  om->tjb_b = unbiased_byteify(om, logf(3.0f / (float) (L+3)));
  return p7G_OK;
}

int
p7_gpu_oprofile_ReconfigRestLength(P7_GPU_OPROFILE *om, int L)
{
  float pmove, ploop;
  pmove = (2.0f + om->nj) / ((float) L + 2.0f + om->nj);
  ploop = 1.0f - pmove;
  om->xf[p7G_XN][p7G_LOOP] = om->xf[p7G_XC][p7G_LOOP] = om->xf[p7G_XJ][p7G_LOOP] = ploop;
  om->xf[p7G_XN][p7G_MOVE] = om->xf[p7G_XC][p7G_MOVE] = om->xf[p7G_XJ][p7G_MOVE] = pmove;
  om->xw[p7G_XN][p7G_MOVE] = om->xw[p7G_XC][p7G_MOVE] = om->xw[p7G_XJ][p7G_MOVE] = wordify(om, logf(pmove));
  om->L = L;
  return p7G_OK;
}

int
p7_gpu_oprofile_ReconfigLength(P7_GPU_OPROFILE *om, int L)
{
  int status;
  if ((status = p7_gpu_oprofile_ReconfigMSVLength (om, L)) != p7G_OK) return status;
  if ((status = p7_gpu_oprofile_ReconfigRestLength(om, L)) != p7G_OK) return status;
  return p7G_OK;
}

/* Per-length constants for every L in 0..Lmax. */
int
p7g_lencfg_build(const P7_GPU_OPROFILE *om, int Lmax, P7G_LENCFG *lc)
{
  int L;
  memset(lc, 0, sizeof(*lc));
  lc->Lmax    = Lmax;
  lc->tjb_b   = malloc((size_t) (Lmax+1));
  lc->xw_move = malloc(sizeof(int16_t) * (size_t) (Lmax+1));
  lc->xf_move = malloc(sizeof(float) * (size_t) (Lmax+1));
  lc->xf_loop = malloc(sizeof(float) * (size_t) (Lmax+1));
  lc->nullsc  = malloc(sizeof(float) * (size_t) (Lmax+1));
  lc->flen_a  = malloc(sizeof(float) * (size_t) (Lmax+1));
  lc->flen_b  = malloc(sizeof(float) * (size_t) (Lmax+1));
  if (! lc->tjb_b || ! lc->xw_move || ! lc->xf_move || ! lc->xf_loop || ! lc->nullsc || ! lc->flen_a || ! lc->flen_b) {
    p7g_lencfg_free(lc);
    return p7G_EMEM;
  }
  for (L = 0; L <= Lmax; L++) {
    float pmove = (2.0f + om->nj) / ((float) L + 2.0f + om->nj);
    float ploop = 1.0f - pmove;
    float p1    = (float) L / (float) (L+1);
    lc->tjb_b[L]   = unbiased_byteify(om, logf(3.0f / (float) (L+3)));
    lc->xw_move[L] = wordify(om, logf(pmove));
    lc->xf_move[L] = pmove;
    lc->xf_loop[L] = ploop;
    lc->nullsc[L]  = (float) L * log(p1) + log(1.-p1);
    lc->flen_a[L]  = (float) L * logf(p1);
    lc->flen_b[L]  = logf(1.-p1);
  }
  return p7G_OK;
}

void
p7g_lencfg_free(P7G_LENCFG *lc)
{
  free(lc->tjb_b); free(lc->xw_move); free(lc->xf_move); free(lc->xf_loop);
  free(lc->nullsc); free(lc->flen_a); free(lc->flen_b);
  memset(lc, 0, sizeof(*lc));
}

/*****************************************************************
 * Parser specials (P7_OMX, p7_omx.c:64-232)
 *****************************************************************/
P7_GPU_OMX *
p7_gpu_omx_Create(int allocM, int allocL, int allocXL)
{
  P7_GPU_OMX *ox = calloc(1, sizeof(P7_GPU_OMX));
  (void) allocL;
  if (! ox) return NULL;
  ox->allocM = allocM > 0 ? allocM : 1;
  ox->allocXL = allocXL > 0 ? allocXL : 0;
  ox->xmx = calloc((size_t) (ox->allocXL + 1) * p7G_NXCELLS, sizeof(float));
  if (! ox->xmx) { free(ox); return NULL; }
  return ox;
}

int
p7_gpu_omx_GrowTo(P7_GPU_OMX *ox, int allocM, int allocL, int allocXL)
{
  (void) allocL;
  if (allocM > ox->allocM) ox->allocM = allocM;
  if (allocXL > ox->allocXL) {
    float *p = realloc(ox->xmx, sizeof(float) * (size_t) (allocXL + 1) * p7G_NXCELLS);
    if (! p) return p7G_EMEM;
    ox->xmx = p; ox->allocXL = allocXL;
  }
  return p7G_OK;
}

const float *
p7_gpu_omx_Xmx(const P7_GPU_OMX *ox, int *ret_L, float *ret_totscale, int *ret_has_own_scales)
{
  if (ret_L) *ret_L = ox->L;
  if (ret_totscale) *ret_totscale = ox->totscale;
  if (ret_has_own_scales) *ret_has_own_scales = ox->has_own_scales;
  return ox->xmx;
}

void
p7_gpu_omx_Destroy(P7_GPU_OMX *ox)
{
  if (! ox) return;
  free(ox->xmx);
  free(ox);
}
