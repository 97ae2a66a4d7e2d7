/* The hit list as text: hmmsearch's per-sequence and per-domain tables
 * (p7_tophits_Targets and p7_tophits_Domains, HMMER/src/p7_tophits.c:
 * 1055-1380, the domains' alignments drawn as p7_alidisplay_Print draws them,
 * p7_alidisplay.c:572-641) and its
 * parsable tables (--tblout and --domtblout: p7_tophits_TabularTargets and
 * p7_tophits_TabularDomains, p7_tophits.c:1519-1700), for a search of one
 * profile against sequences (p7_SEARCH_SEQS, protein targets). The column
 * layouts and printf conversions are the reference's, so that the text is
 * byte for byte what hmmsearch writes; the numbers come from the runtime's
 * hit list (ranked and thresholded targets, their rescored domains).
 */
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include "p7g_internal.h"

#define MAXI(a, b) ((a) > (b) ? (a) : (b))

typedef struct { char *buf; size_t cap, len; } SBUF;

static void
sb_printf(SBUF *s, const char *fmt, ...)
{
  va_list ap;
  va_start(ap, fmt);
  const size_t room = (s->buf && s->len < s->cap) ? s->cap - s->len : 0;
  const int n = vsnprintf(room ? s->buf + s->len : NULL, room, fmt, ap);
  va_end(ap);
  if (n > 0) s->len += (size_t) n;
}

static size_t
sb_end(SBUF *s)
{
  if (s->buf && s->cap) s->buf[s->len < s->cap ? s->len : s->cap - 1] = '\0';
  return s->len;
}

/* widest target name (or, with show_acc, the name shown: accession when there
 * is one) and accession over every hit of the list, reported or not
 * (p7_tophits_GetMaxNameLength, GetMaxAccessionLength, GetMaxShownLength) */
static int
max_name(const P7G_RLIST *l, int shown)
{
  int m = 0;
  for (int h = 0; h < l->N; h++) {
    const char *s = (shown && l->hit[h].acc) ? l->hit[h].acc : l->hit[h].name;
    if (s) m = MAXI(m, (int) strlen(s));
  }
  return m;
}

static int
max_acc(const P7G_RLIST *l)
{
  int m = 0;
  for (int h = 0; h < l->N; h++) if (l->hit[h].acc) m = MAXI(m, (int) strlen(l->hit[h].acc));
  return m;
}

static int
n_reported(const P7G_RLIST *l)
{
  int n = 0;
  for (int h = 0; h < l->N; h++) n += (l->hit[h].flags & 1) != 0;
  return n;
}

/* acc column of the domain tables: oasc per envelope position */
static double
dom_acc(const P7G_RDOM *d)
{
  return d->oasc / (1.0 + fabs((float) (d->jenv - d->ienv)));
}

size_t
p7g_targets_text(const P7G_RLIST *l, char *buf, size_t cap)
{
  SBUF s = { buf, cap, 0 };
  const int namew = MAXI(8, max_name(l, l->show_accessions));
  const int descw = (l->textw > 0) ? MAXI(32, l->textw - namew - 61) : 0;
  int incthresh = 0;

  sb_printf(&s, "Scores for complete sequences (score includes all domains):\n");
  sb_printf(&s, "  %22s  %22s  %8s\n", " --- full sequence ---", " --- best 1 domain ---", "-#dom-");
  sb_printf(&s, "  %9s %6s %5s  %9s %6s %5s  %5s %2s  %-*s %s\n",
            "E-value", " score", " bias", "E-value", " score", " bias", "  exp", "N", namew, "Sequence", "Description");
  sb_printf(&s, "  %9s %6s %5s  %9s %6s %5s  %5s %2s  %-*s %s\n",
            "-------", "------", "-----", "-------", "------", "-----", " ----", "--", namew, "--------", "-----------");
  for (int h = 0; h < l->N; h++) {
    const P7G_RHIT *t = &l->hit[h];
    if (! (t->flags & 1)) continue;
    if (! (t->flags & 2) && ! incthresh) {
      sb_printf(&s, "  ------ inclusion threshold ------\n");
      incthresh = 1;
    }
    const char *shown = (l->show_accessions && t->acc) ? t->acc : t->name;
    const P7G_RDOM *d = t->ndom ? &t->dom[t->best] : NULL;
    sb_printf(&s, "%c %9.2g %6.1f %5.1f  %9.2g %6.1f %5.1f  %5.1f %2d  %-*s ",
              ' ', exp(t->lnP) * l->Z, t->score, t->pre_score - t->score,
              d ? exp(d->lnP) * l->Z : NAN, d ? d->bitscore : NAN, d ? P7G_LOG2R * d->dombias : NAN,
              t->nexpected, t->nreported, namew, shown);
    if (l->textw > 0) sb_printf(&s, " %-.*s\n", descw, t->desc ? t->desc : "");
    else              sb_printf(&s, " %s\n", t->desc ? t->desc : "");
  }
  if (n_reported(l) == 0) sb_printf(&s, "\n   [No hits detected that satisfy reporting thresholds]\n");
  return sb_end(&s);
}

static int
text_width(long v)
{
  char tmp[32];
  return snprintf(tmp, sizeof(tmp), "%ld", v);
}

/* p7_alidisplay_Print for one domain: blocks of aliwidth columns (at least
 * min_aliwidth when textw leaves less), each with the CS and RF lines, the
 * model line with its first and last node, the match line, the target line
 * with its first and last residue ('-' for a block of deletions only), and
 * the posterior probability line. */
static void
print_alignment(SBUF *s, const P7G_RDOM *d, const char *hmmname, const char *seqname, int textw)
{
  const int min_aliwidth = 40;
  const int namewidth = MAXI((int) strlen(hmmname), (int) strlen(seqname));
  const int coordwidth = MAXI(MAXI(text_width(d->hmmfrom), text_width(d->hmmto)),
                              MAXI(text_width(d->alifrom), text_width(d->alito)));
  int aliwidth = (textw > 0) ? textw - namewidth - 2 * coordwidth - 5 : d->N;
  if (aliwidth < d->N && aliwidth < min_aliwidth) aliwidth = min_aliwidth;
  if (aliwidth < 1) aliwidth = 1;
  long i1 = d->alifrom;
  int k1 = d->hmmfrom;
  for (int pos = 0; pos < d->N; pos += aliwidth) {
    if (pos > 0) sb_printf(s, "\n");
    const int w = MAXI(0, (d->N - pos < aliwidth) ? d->N - pos : aliwidth);
    int ni = 0, nk = 0;
    for (int z = pos; z < pos + w; z++) {
      if (d->model[z] != '.') nk++;
      if (d->aseq[z] != '-') ni++;
    }
    const int k2 = k1 + nk - 1;
    const long i2 = i1 + ni - 1;
    if (d->csline) sb_printf(s, "  %*s %.*s CS\n", namewidth + coordwidth + 1, "", w, d->csline + pos);
    if (d->rfline) sb_printf(s, "  %*s %.*s RF\n", namewidth + coordwidth + 1, "", w, d->rfline + pos);
    sb_printf(s, "  %*s %*d %.*s %-*d\n", namewidth, hmmname, coordwidth, k1, w, d->model + pos, coordwidth, k2);
    sb_printf(s, "  %*s %.*s\n", namewidth + coordwidth + 1, " ", w, d->mline + pos);
    if (ni > 0) sb_printf(s, "  %*s %*ld %.*s %-*ld\n", namewidth, seqname, coordwidth, i1, w, d->aseq + pos, coordwidth, i2);
    else        sb_printf(s, "  %*s %*s %.*s %*s\n", namewidth, seqname, coordwidth, "-", w, d->aseq + pos, coordwidth, "-");
    if (d->ppline) sb_printf(s, "  %*s %.*s PP\n", namewidth + coordwidth + 1, "", w, d->ppline + pos);
    k1 += nk;
    i1 += ni;
  }
}

size_t
p7g_domains_text(const P7G_RLIST *l, char *buf, size_t cap)
{
  SBUF s = { buf, cap, 0 };
  sb_printf(&s, "Domain annotation for each sequence%s:\n", l->show_alignments ? " (and alignments)" : "");
  for (int h = 0; h < l->N; h++) {
    const P7G_RHIT *t = &l->hit[h];
    if (! (t->flags & 1)) continue;
    const char *shown = (l->show_accessions && t->acc) ? t->acc : t->name;
    if (l->textw > 0) {
      const int descw = MAXI(32, l->textw - (int) strlen(shown) - 5);
      sb_printf(&s, ">> %s  %-.*s\n", shown, descw, t->desc ? t->desc : "");
    } else
      sb_printf(&s, ">> %s  %s\n", shown, t->desc ? t->desc : "");
    if (t->nreported == 0) {
      sb_printf(&s, "   [No individual domains that satisfy reporting thresholds (although complete target did)]\n\n");
      continue;
    }
    sb_printf(&s, " %3s   %6s %5s %9s %9s %7s %7s %2s %7s %7s %2s %7s %7s %2s %4s\n",
              "#", "score", "bias", "c-Evalue", "i-Evalue", "hmmfrom", "hmm to", "  ", "alifrom", "ali to", "  ",
              "envfrom", "env to", "  ", "acc");
    sb_printf(&s, " %3s   %6s %5s %9s %9s %7s %7s %2s %7s %7s %2s %7s %7s %2s %4s\n",
              "---", "------", "-----", "---------", "---------", "-------", "-------", "  ", "-------", "-------", "  ",
              "-------", "-------", "  ", "----");
    int nd = 0;
    for (int k = 0; k < t->ndom; k++) {
      const P7G_RDOM *d = &t->dom[k];
      if (! (d->flags & 1)) continue;
      nd++;
      sb_printf(&s, " %3d %c %6.1f %5.1f %9.2g %9.2g %7d %7d %c%c %7ld %7ld %c%c %7d %7d %c%c %4.2f\n",
                nd, (d->flags & 2) ? '!' : '?', d->bitscore, d->dombias * P7G_LOG2R,
                exp(d->lnP) * l->domZ, exp(d->lnP) * l->Z,
                d->hmmfrom, d->hmmto, d->hmmfrom == 1 ? '[' : '.', d->hmmto == l->M ? ']' : '.',
                (long) d->alifrom, (long) d->alito, d->alifrom == 1 ? '[' : '.', d->alito == t->L ? ']' : '.',
                d->ienv, d->jenv, d->ienv == 1 ? '[' : '.', d->jenv == t->L ? ']' : '.', dom_acc(d));
    }
    if (l->show_alignments) {
      sb_printf(&s, "\n  Alignments for each domain:\n");
      const char *hmmname = (l->show_accessions && l->qacc) ? l->qacc : l->qname;
      nd = 0;
      for (int k = 0; k < t->ndom; k++) {
        const P7G_RDOM *d = &t->dom[k];
        if (! (d->flags & 1)) continue;
        nd++;
        sb_printf(&s, "  == domain %d", nd);
        sb_printf(&s, "  score: %.1f bits", d->bitscore);
        sb_printf(&s, ";  conditional E-value: %.2g\n", exp(d->lnP) * l->domZ);
        if (d->model) print_alignment(&s, d, hmmname, shown, l->textw);
        sb_printf(&s, "\n");
      }
    } else
      sb_printf(&s, "\n");
  }
  if (n_reported(l) == 0) sb_printf(&s, "\n   [No targets detected that satisfy reporting thresholds]\n");
  return sb_end(&s);
}

size_t
p7g_tabular_targets(const P7G_RLIST *l, int show_header, char *buf, size_t cap)
{
  SBUF s = { buf, cap, 0 };
  const int qnamew = MAXI(20, (int) strlen(l->qname));
  const int tnamew = MAXI(20, max_name(l, 0));
  const int qaccw  = l->qacc ? MAXI(10, (int) strlen(l->qacc)) : 10;
  const int taccw  = MAXI(10, max_acc(l));
  if (show_header) {
    sb_printf(&s, "#%*s %22s %22s %33s\n", tnamew + qnamew + taccw + qaccw + 2, "",
              "--- full sequence ----", "--- best 1 domain ----", "--- domain number estimation ----");
    sb_printf(&s, "#%-*s %-*s %-*s %-*s %9s %6s %5s %9s %6s %5s %5s %3s %3s %3s %3s %3s %3s %3s %s\n",
              tnamew - 1, " target name", taccw, "accession", qnamew, "query name", qaccw, "accession",
              "  E-value", " score", " bias", "  E-value", " score", " bias",
              "exp", "reg", "clu", " ov", "env", "dom", "rep", "inc", "description of target");
    sb_printf(&s, "#%*s %*s %*s %*s %9s %6s %5s %9s %6s %5s %5s %3s %3s %3s %3s %3s %3s %3s %s\n",
              tnamew - 1, "-------------------", taccw, "----------", qnamew, "--------------------", qaccw, "----------",
              "---------", "------", "-----", "---------", "------", "-----",
              "---", "---", "---", "---", "---", "---", "---", "---", "---------------------");
  }
  for (int h = 0; h < l->N; h++) {
    const P7G_RHIT *t = &l->hit[h];
    if (! (t->flags & 1)) continue;
    const P7G_RDOM *d = t->ndom ? &t->dom[t->best] : NULL;
    sb_printf(&s, "%-*s %-*s %-*s %-*s %9.2g %6.1f %5.1f %9.2g %6.1f %5.1f %5.1f %3d %3d %3d %3d %3d %3d %3d %s\n",
              tnamew, t->name, taccw, t->acc ? t->acc : "-", qnamew, l->qname, qaccw, l->qacc ? l->qacc : "-",
              exp(t->lnP) * l->Z, t->score, t->pre_score - t->score,
              d ? exp(d->lnP) * l->Z : NAN, d ? d->bitscore : NAN, d ? d->dombias * P7G_LOG2R : NAN,
              t->nexpected, t->nregions, t->nclustered, t->noverlaps, t->nenvelopes, t->ndom, t->nreported,
              t->nincluded, t->desc ? t->desc : "-");
  }
  return sb_end(&s);
}

size_t
p7g_tabular_domains(const P7G_RLIST *l, int show_header, char *buf, size_t cap)
{
  SBUF s = { buf, cap, 0 };
  const int qnamew = MAXI(20, (int) strlen(l->qname));
  const int tnamew = MAXI(20, max_name(l, 0));
  const int qaccw  = l->qacc ? MAXI(10, (int) strlen(l->qacc)) : 10;
  const int taccw  = MAXI(10, max_acc(l));
  if (show_header) {
    sb_printf(&s, "#%*s %22s %40s %11s %11s %11s\n", tnamew + qnamew - 1 + 15 + taccw + qaccw, "",
              "--- full sequence ---", "-------------- this domain -------------", "hmm coord", "ali coord",
              "env coord");
    sb_printf(&s, "#%-*s %-*s %5s %-*s %-*s %5s %9s %6s %5s %3s %3s %9s %9s %6s %5s %5s %5s %5s %5s %5s %5s %4s %s\n",
              tnamew - 1, " target name", taccw, "accession", "tlen", qnamew, "query name", qaccw, "accession", "qlen",
              "E-value", "score", "bias", "#", "of", "c-Evalue", "i-Evalue", "score", "bias",
              "from", "to", "from", "to", "from", "to", "acc", "description of target");
    sb_printf(&s, "#%*s %*s %5s %*s %*s %5s %9s %6s %5s %3s %3s %9s %9s %6s %5s %5s %5s %5s %5s %5s %5s %4s %s\n",
              tnamew - 1, "-------------------", taccw, "----------", "-----", qnamew, "--------------------", qaccw,
              "----------", "-----", "---------", "------", "-----", "---", "---", "---------", "---------", "------",
              "-----", "-----", "-----", "-----", "-----", "-----", "-----", "----", "---------------------");
  }
  for (int h = 0; h < l->N; h++) {
    const P7G_RHIT *t = &l->hit[h];
    if (! (t->flags & 1)) continue;
    int nd = 0;
    for (int k = 0; k < t->ndom; k++) {
      const P7G_RDOM *d = &t->dom[k];
      if (! (d->flags & 1)) continue;
      nd++;
      sb_printf(&s, "%-*s %-*s %5d %-*s %-*s %5d %9.2g %6.1f %5.1f %3d %3d %9.2g %9.2g %6.1f %5.1f %5d %5d %5ld %5ld "
                    "%5d %5d %4.2f %s\n",
                tnamew, t->name, taccw, t->acc ? t->acc : "-", t->L, qnamew, l->qname, qaccw,
                l->qacc ? l->qacc : "-", l->M, exp(t->lnP) * l->Z, t->score, t->pre_score - t->score, nd,
                t->nreported, exp(d->lnP) * l->domZ, exp(d->lnP) * l->Z, d->bitscore, d->dombias * P7G_LOG2R,
                d->hmmfrom, d->hmmto, (long) d->alifrom, (long) d->alito, d->ienv, d->jenv, dom_acc(d),
                t->desc ? t->desc : "-");
    }
  }
  return sb_end(&s);
}
