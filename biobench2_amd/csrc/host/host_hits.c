/* Target and domain scores of the hit list, from a target's Forward score
 * and its domains' envelope scores and null2 corrections, as p7_Pipeline
 * computes them after domain definition (HMMER/src/p7_pipeline.c:743-850),
 * with p7_FLogsum (HMMER/src/logsum.c:59-112) and esl_exp_logsurv
 * (easel/esl_exponential.c:141-145). Scores are float and P-values double,
 * as in the reference; the sequence bias sums the per-residue null2 scores
 * grouped by envelope (and multidomain region) in sequence order, where the
 * reference sums them residue by residue (they are zero elsewhere).
 */
#include <inttypes.h>
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include "p7g_internal.h"

#define LOGSUM_SCALE 1000.f
#define LOGSUM_TBL   16000
#define CONST_LOG2   0.69314718055994529      /* eslCONST_LOG2 */

static float flogsum_lookup[LOGSUM_TBL];
static pthread_once_t flogsum_once = PTHREAD_ONCE_INIT;

static void
flogsum_init(void)
{
  for (int i = 0; i < LOGSUM_TBL; i++)
    flogsum_lookup[i] = log(1. + exp((double) -i / LOGSUM_SCALE));
}

float
p7g_flogsum(float a, float b)
{
  pthread_once(&flogsum_once, flogsum_init);
  const float max = (a > b) ? a : b, min = (a > b) ? b : a;
  return (min == -INFINITY || (max - min) >= 15.7f) ? max : max + flogsum_lookup[(int) ((max - min) * LOGSUM_SCALE)];
}

double
p7g_exp_logsurv(double x, double mu, double lambda)
{
  if (x < mu) return 0.0;
  return -lambda * (x - mu);
}

/* p7_bg_NullOne with p7_bg_SetLength(bg, L) (p7_bg.c:359-363, 473-484) */
float
p7g_null_one(int L)
{
  const float p1 = (float) L / (float) (L + 1);
  return (float) L * log(p1) + log(1. - p1);
}

void
p7g_target_scores(int L, float fwdsc, float nullsc, float omega, float tau, float lambda,
                  const P7G_DOMIN *dom, int ndom, const float *n2parts, int nn2, P7G_TSCORE *ts, P7G_DSCORE *ds)
{
  float seqbias = 0.0f;
  if (n2parts) for (int a = 0; a < nn2; a++) seqbias += n2parts[a];
  else         for (int d = 0; d < ndom; d++) seqbias += dom[d].domcorr;
  seqbias = p7g_flogsum(0.0f, log(omega) + seqbias);
  float pre_score = (fwdsc - nullsc) / CONST_LOG2;
  float seq_score = (fwdsc - (nullsc + seqbias)) / CONST_LOG2;

  float sum_score = 0.0f;
  int Ld = 0;
  seqbias = 0.0f;
  for (int d = 0; d < ndom; d++)
    if (dom[d].envsc - dom[d].domcorr > 0.0) {
      sum_score += dom[d].envsc;
      Ld        += dom[d].jenv - dom[d].ienv + 1;
      seqbias   += dom[d].domcorr;
    }
  seqbias = p7g_flogsum(0.0f, log(omega) + seqbias);
  sum_score += (L - Ld) * log((float) L / (float) (L + 3));
  const float pre2_score = (sum_score - nullsc) / CONST_LOG2;
  sum_score = (sum_score - (nullsc + seqbias)) / CONST_LOG2;
  if (Ld > 0 && sum_score > seq_score) { seq_score = sum_score; pre_score = pre2_score; }

  ts->score = seq_score; ts->pre_score = pre_score; ts->sum_score = sum_score;
  ts->lnP = p7g_exp_logsurv(seq_score, tau, lambda);
  ts->pre_lnP = p7g_exp_logsurv(pre_score, tau, lambda);
  ts->sum_lnP = p7g_exp_logsurv(sum_score, tau, lambda);
  ts->best = 0;
  for (int d = 0; d < ndom; d++) {
    const int Lde = dom[d].jenv - dom[d].ienv + 1;
    float bits = dom[d].envsc + (L - Lde) * log((float) L / (float) (L + 3));
    ds[d].dombias = p7g_flogsum(0.0f, log(omega) + dom[d].domcorr);
    bits = (bits - (nullsc + ds[d].dombias)) / CONST_LOG2;
    ds[d].bitscore = bits;
    ds[d].lnP = p7g_exp_logsurv(bits, tau, lambda);
    if (ds[d].bitscore > ds[ts->best].bitscore) ts->best = d;
  }
}

/* The "Internal pipeline statistics summary" text of p7_pli_Statistics for a
 * search of models against sequences (p7_pipeline.c:1710-1796, the branch for
 * p7_SEARCH_SEQS without long targets, without the stopwatch lines), written
 * into buf (NUL-terminated, truncated at cap). Returns the full length. */
size_t
p7_gpu_StatisticsText(const P7_GPU_STATS *st, double F1, double F2, double F3, double Z, int z_by_option, double domZ,
                    int domz_by_option, char *buf, size_t cap)
{
  char tmp[2048];
  const double ntargets = (double) st->nseqs;
  int n = snprintf(tmp, sizeof(tmp),
    "Internal pipeline statistics summary:\n"
    "-------------------------------------\n"
    "Query model(s):                %15" PRId64 "  (%" PRId64 " nodes)\n"
    "Target sequences:              %15" PRId64 "  (%" PRId64 " residues searched)\n"
    "Passed MSV filter:           %15" PRId64 "  (%.6g); expected %.1f (%.6g)\n"
    "Passed bias filter:          %15" PRId64 "  (%.6g); expected %.1f (%.6g)\n"
    "Passed Vit filter:           %15" PRId64 "  (%.6g); expected %.1f (%.6g)\n"
    "Passed Fwd filter:           %15" PRId64 "  (%.6g); expected %.1f (%.6g)\n"
    "Initial search space (Z):    %15.0f  %s\n"
    "Domain search space  (domZ): %15.0f  %s\n",
    st->nmodels, st->nnodes, st->nseqs, st->nres,
    st->n_past_msv, (double) st->n_past_msv / ntargets, F1 * ntargets, F1,
    st->n_past_bias, (double) st->n_past_bias / ntargets, F1 * ntargets, F1,
    st->n_past_vit, (double) st->n_past_vit / ntargets, F2 * ntargets, F2,
    st->n_past_fwd, (double) st->n_past_fwd / ntargets, F3 * ntargets, F3,
    Z, z_by_option ? "[as set by --Z on cmdline]" : "[actual number of targets]",
    domZ, domz_by_option ? "[as set by --domZ on cmdline]" : "[number of targets reported over threshold]");
  if (n < 0) n = 0;
  if (buf && cap) {
    const size_t m = ((size_t) n < cap - 1) ? (size_t) n : cap - 1;
    memcpy(buf, tmp, m);
    buf[m] = '\0';
  }
  return (size_t) n;
}
