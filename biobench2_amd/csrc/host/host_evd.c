/* P-value tails of the pipeline's stage gates, and their inversion into
 * exact float score thresholds.
 *
 * p7_Pipeline (HMMER/src/p7_pipeline.c:690-731) rejects a target when
 * P(seq_score) > F, where seq_score is a float bit score and P is the Gumbel
 * (easel/esl_gumbel.c:129-137) or exponential (easel/esl_exponential.c:127-131)
 * survivor function in double precision. Both are non-increasing in the
 * score, so the gate equals "seq_score >= T" for T = the smallest float whose
 * P-value is <= F. T is found here by bisection over the ordered float bit
 * patterns using the same double-precision formula, so the device compares
 * floats and never evaluates exp() for a gate.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "p7g_internal.h"

#define P7G_SMALLX1 5e-9      /* eslSMALLX1 (easel.h:285) */

double
p7g_gumbel_surv(double x, double mu, double lambda)
{
  double y  = lambda*(x-mu);
  double ey = -exp(-y);
  if (fabs(ey) < P7G_SMALLX1) return -ey;
  else                        return 1 - exp(ey);
}

double
p7g_exp_surv(double x, double mu, double lambda)
{
  if (x < mu) return 1.0;
  return exp(-lambda * (x-mu));
}

static int32_t
f2ord(float f)
{
  int32_t i;
  memcpy(&i, &f, sizeof(i));
  return (i >= 0) ? i : (int32_t) (INT32_MIN - i);
}

static float
ord2f(int32_t o)
{
  int32_t i = (o >= 0) ? o : (int32_t) (INT32_MIN - o);
  float f;
  memcpy(&f, &i, sizeof(f));
  return f;
}

static float
threshold(int which, double mu, double lambda, double P)
{
  /* pass(s) := surv(s) <= P, false at -inf-like scores, true at +inf */
  int64_t lo = f2ord(-1e30f), hi = f2ord(1e30f);      /* 64-bit: the span exceeds INT32_MAX */
#define PASS(s) (((which) == 0 ? p7g_gumbel_surv((double)(s), mu, lambda) : p7g_exp_surv((double)(s), mu, lambda)) <= P)
  if (PASS(ord2f((int32_t) lo))) return -INFINITY;
  if (! PASS(ord2f((int32_t) hi))) return INFINITY;
  while (hi - lo > 1) {
    int64_t mid = lo + (hi - lo) / 2;
    if (PASS(ord2f((int32_t) mid))) hi = mid; else lo = mid;
  }
#undef PASS
  return ord2f((int32_t) hi);
}

float p7g_gumbel_threshold(double mu, double lambda, double P) { return threshold(0, mu, lambda, P); }
float p7g_exp_threshold   (double mu, double lambda, double P) { return threshold(1, mu, lambda, P); }

float
p7_gpu_ScoreThreshold(int which, double mu, double lambda, double P)
{
  return which == 0 ? p7g_gumbel_threshold(mu, lambda, P) : p7g_exp_threshold(mu, lambda, P);
}
