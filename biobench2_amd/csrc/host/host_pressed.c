/* Pressed profiles: the .h3f (MSV part) and .h3p (the rest) files that
 * hmmpress writes, format 3/f (HMMER/src/impl_sse/io.c:89-176 writes them;
 * p7_oprofile_ReadMSV io.c:233-338 and p7_oprofile_ReadRest io.c:493-581
 * read them). A record holds the P7_OPROFILE tables striped for 128-bit SSE
 * vectors: Q = max(2, ceil(M/W)) vectors of W lanes, lane z of vector q
 * holding node k = q + 1 + z*Q (impl_sse.h:24-67). This reader removes the
 * striping and hands the tables to p7_gpu_oprofile_Load, so a pressed
 * database is served without re-converting its models. Files are read as
 * written on a little-endian host with 4-byte int and 8-byte off_t, as the
 * reference's are. */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "p7_gpu.h"
#include "p7g_internal.h"

#define V3F_FMAGIC 0xb3e6e6f3u   /* io.c:48 */
#define V3F_PMAGIC 0xb3e6f0f3u   /* io.c:49 */
#define NOFFSETS   3             /* p7_NOFFSETS (hmmer.h) */
#define NCUTOFFS   6             /* p7_NCUTOFFS (hmmer.h) */
#define SB_EXTRA   17            /* p7O_EXTRA_SB: wrap vectors of the SSV table (impl_sse.h:30) */

static int nq(int M, int w) { int q = (M - 1) / w + 1; return q < 2 ? 2 : q; }

typedef struct { FILE *fp; const char *what; } rd_t;

static int
rd(rd_t *r, void *dst, size_t n)
{
  if (n == 0) return p7G_OK;
  if (fread(dst, 1, n, r->fp) != n) { p7g_set_error("truncated %s record", r->what); return p7G_EFORMAT; }
  return p7G_OK;
}

static int
skip(rd_t *r, size_t n)
{
  char buf[4096];
  while (n > 0) { size_t k = n < sizeof(buf) ? n : sizeof(buf); if (rd(r, buf, k) != p7G_OK) return p7G_EFORMAT; n -= k; }
  return p7G_OK;
}

/* one model of each file; the tables are left striped in the buffers */
typedef struct {
  int      M, abctype, Kp;
  uint8_t  tbm_b, tec_b, tjb_b, base_b, bias_b;
  float    scale_b;
  uint8_t *rbv;                    /* [Kp][Q16][16] */
  float    evparam[6], compo[20];
  int16_t *twv, *rwv;              /* [8*Q8][8], [Kp][Q8][8] */
  int16_t  xw[4][2];
  float    scale_w, ncj_roundoff;
  int16_t  base_w, ddbound_w;
  float   *tfv, *rfv;              /* [8*Q4][4], [Kp][Q4][4] */
  float    xf[4][2];
  float    nj;
  int      mode, L;
} rec_t;

static void rec_free(rec_t *m) { free(m->rbv); free(m->twv); free(m->rwv); free(m->tfv); free(m->rfv); memset(m, 0, sizeof(*m)); }

static int
read_msv(rd_t *r, rec_t *m, int want)
{
  uint32_t magic;
  int n, maxlen, status;
  if (fread(&magic, 4, 1, r->fp) != 1) { p7g_set_error("the .h3f file holds fewer models than asked for"); return p7G_EOF; }
  if (magic != V3F_FMAGIC) { p7g_set_error("bad .h3f magic (not a HMMER 3/f pressed file)"); return p7G_EFORMAT; }
  if ((status = rd(r, &m->M, 4)) || (status = rd(r, &m->abctype, 4)) || (status = rd(r, &n, 4))) return status;
  if (m->M <= 0 || n < 0 || n > 1 << 20) { p7g_set_error("bad .h3f record header"); return p7G_EFORMAT; }
  const P7G_ALPHABET *abc = p7g_alphabet(m->abctype);
  if (! abc) { p7g_set_error("unknown alphabet type %d in .h3f", m->abctype); return p7G_EFORMAT; }
  m->Kp = abc->Kp;
  const int Q16 = nq(m->M, 16);
  if ((status = skip(r, (size_t) n + 1)) || (status = rd(r, &maxlen, 4))) return status;
  if ((status = rd(r, &m->tbm_b, 1)) || (status = rd(r, &m->tec_b, 1)) || (status = rd(r, &m->tjb_b, 1)) ||
      (status = rd(r, &m->scale_b, 4)) || (status = rd(r, &m->base_b, 1)) || (status = rd(r, &m->bias_b, 1))) return status;
  if ((status = skip(r, (size_t) m->Kp * (Q16 + SB_EXTRA) * 16)) != p7G_OK) return status;      /* SSV table: rebuilt from rbv */
  if (want) {
    if (! (m->rbv = malloc((size_t) m->Kp * Q16 * 16))) return p7G_EMEM;
    if ((status = rd(r, m->rbv, (size_t) m->Kp * Q16 * 16)) != p7G_OK) return status;
  } else if ((status = skip(r, (size_t) m->Kp * Q16 * 16)) != p7G_OK) return status;
  if ((status = rd(r, m->evparam, 24)) || (status = skip(r, 8 * NOFFSETS)) || (status = rd(r, m->compo, 80))) return status;
  if ((status = rd(r, &magic, 4)) != p7G_OK) return status;
  if (magic != V3F_FMAGIC) { p7g_set_error("bad .h3f sentinel: file corrupted?"); return p7G_EFORMAT; }
  return p7G_OK;
}

static int
read_rest(rd_t *r, rec_t *m, int want)
{
  uint32_t magic;
  int M, abctype, n, status;
  if (fread(&magic, 4, 1, r->fp) != 1) { p7g_set_error("the .h3p file holds fewer models than the .h3f"); return p7G_EFORMAT; }
  if (magic != V3F_PMAGIC) { p7g_set_error("bad .h3p magic (not a HMMER 3/f pressed file)"); return p7G_EFORMAT; }
  if ((status = rd(r, &M, 4)) || (status = rd(r, &abctype, 4)) || (status = rd(r, &n, 4))) return status;
  if (M != m->M || abctype != m->abctype) { p7g_set_error(".h3f and .h3p records disagree"); return p7G_EFORMAT; }
  if (n < 0 || n > 1 << 20 || (status = skip(r, (size_t) n + 1))) return status ? status : p7G_EFORMAT;   /* name */
  for (int a = 0; a < 2; a++) {                                                                           /* acc, desc */
    if ((status = rd(r, &n, 4)) != p7G_OK) return status;
    if (n > 0 && (status = skip(r, (size_t) n + 1)) != p7G_OK) return status;
  }
  if ((status = skip(r, 4 * ((size_t) M + 2))) != p7G_OK) return status;                                  /* rf mm cs consensus */
  const int Q8 = nq(M, 8), Q4 = nq(M, 4);
  if (want) {
    if (! (m->twv = malloc(2 * 64 * (size_t) Q8)) || ! (m->rwv = malloc(2 * 8 * (size_t) m->Kp * Q8)) ||
        ! (m->tfv = malloc(4 * 32 * (size_t) Q4)) || ! (m->rfv = malloc(4 * 4 * (size_t) m->Kp * Q4))) return p7G_EMEM;
    if ((status = rd(r, m->twv, 2 * 64 * (size_t) Q8)) || (status = rd(r, m->rwv, 2 * 8 * (size_t) m->Kp * Q8))) return status;
  } else if ((status = skip(r, 2 * 64 * (size_t) Q8 + 2 * 8 * (size_t) m->Kp * Q8)) != p7G_OK) return status;
  if ((status = rd(r, m->xw, 16)) || (status = rd(r, &m->scale_w, 4)) || (status = rd(r, &m->base_w, 2)) ||
      (status = rd(r, &m->ddbound_w, 2)) || (status = rd(r, &m->ncj_roundoff, 4))) return status;
  if (want) {
    if ((status = rd(r, m->tfv, 4 * 32 * (size_t) Q4)) || (status = rd(r, m->rfv, 4 * 4 * (size_t) m->Kp * Q4))) return status;
  } else if ((status = skip(r, 4 * 32 * (size_t) Q4 + 4 * 4 * (size_t) m->Kp * Q4)) != p7G_OK) return status;
  if ((status = rd(r, m->xf, 32)) || (status = skip(r, 4 * NCUTOFFS)) || (status = rd(r, &m->nj, 4)) ||
      (status = rd(r, &m->mode, 4)) || (status = rd(r, &m->L, 4)) || (status = rd(r, &magic, 4))) return status;
  if (magic != V3F_PMAGIC) { p7g_set_error("bad .h3p sentinel: file corrupted?"); return p7G_EFORMAT; }
  return p7G_OK;
}

/* node k (1..M) of a striped table: vector (k-1) % Q, lane (k-1) / Q */
#define STRIPED(base, Q, W, k) ((base)[((size_t) (((k) - 1) % (Q))) * (W) + ((k) - 1) / (Q)])

int
p7_gpu_oprofile_ReadPressed(const char *h3f_path, const char *h3p_path, int which, P7_GPU_OPROFILE **ret_om)
{
  rd_t rf = { NULL, ".h3f" }, rp = { NULL, ".h3p" };
  rec_t m;
  int status = p7G_OK;
  P7_GPU_OPROFILE *om = NULL;
  *ret_om = NULL;
  memset(&m, 0, sizeof(m));
  if (which < 0) { p7g_set_error("negative model index"); return p7G_EINVAL; }
  if (! (rf.fp = fopen(h3f_path, "rb"))) { p7g_set_error("can't open %s", h3f_path); return p7G_ENOTFOUND; }
  if (! (rp.fp = fopen(h3p_path, "rb"))) { fclose(rf.fp); p7g_set_error("can't open %s", h3p_path); return p7G_ENOTFOUND; }
  for (int i = 0; i <= which && status == p7G_OK; i++) {
    rec_free(&m);
    if ((status = read_msv(&rf, &m, i == which)) == p7G_OK) status = read_rest(&rp, &m, i == which);
  }
  fclose(rf.fp); fclose(rp.fp);
  if (status != p7G_OK) { rec_free(&m); return status; }

  const int M = m.M, Kp = m.Kp, Q16 = nq(M, 16), Q8 = nq(M, 8), Q4 = nq(M, 4);
  uint8_t *rbv = malloc((size_t) Kp * M);
  int16_t *rwv = malloc(2 * (size_t) Kp * M), *twv = malloc(2 * 8 * (size_t) M);
  float   *rfv = malloc(4 * (size_t) Kp * M), *tfv = malloc(4 * 8 * (size_t) M);
  if (! rbv || ! rwv || ! twv || ! rfv || ! tfv) { status = p7G_EMEM; goto DONE; }
  for (int x = 0; x < Kp; x++)
    for (int k = 1; k <= M; k++) {
      rbv[(size_t) x * M + k - 1] = STRIPED(m.rbv + (size_t) x * Q16 * 16, Q16, 16, k);
      rwv[(size_t) x * M + k - 1] = STRIPED(m.rwv + (size_t) x * Q8 * 8, Q8, 8, k);
      rfv[(size_t) x * M + k - 1] = STRIPED(m.rfv + (size_t) x * Q4 * 4, Q4, 4, k);
    }
  /* transitions: for each vector q, seven vectors BM MM IM DM MD MI II, then the Q DD vectors
   * (p7_oprofile.c:840-870, 955-985) */
  for (int k = 1; k <= M; k++) {
    const int q8 = (k - 1) % Q8, z8 = (k - 1) / Q8, q4 = (k - 1) % Q4, z4 = (k - 1) / Q4;
    for (int t = 0; t < 7; t++) {
      twv[(size_t) t * M + k - 1] = m.twv[((size_t) q8 * 7 + t) * 8 + z8];
      tfv[(size_t) t * M + k - 1] = m.tfv[((size_t) q4 * 7 + t) * 4 + z4];
    }
    twv[(size_t) 7 * M + k - 1] = m.twv[((size_t) 7 * Q8 + q8) * 8 + z8];
    tfv[(size_t) 7 * M + k - 1] = m.tfv[((size_t) 7 * Q4 + q4) * 4 + z4];
  }
  {
    P7_GPU_TABLES t;
    memset(&t, 0, sizeof(t));
    t.M = M; t.Kp = Kp; t.abctype = m.abctype;
    t.rbv = rbv; t.tbm_b = m.tbm_b; t.tec_b = m.tec_b; t.tjb_b = m.tjb_b; t.base_b = m.base_b; t.bias_b = m.bias_b;
    t.scale_b = m.scale_b;
    t.rwv = rwv; t.twv = twv; memcpy(t.xw, m.xw, sizeof(t.xw));
    t.scale_w = m.scale_w; t.base_w = m.base_w; t.ddbound_w = m.ddbound_w;
    t.rfv = rfv; t.tfv = tfv; memcpy(t.xf, m.xf, sizeof(t.xf));
    t.L = m.L; t.mode = m.mode; t.nj = m.nj;
    memcpy(t.evparam, m.evparam, sizeof(t.evparam));
    memcpy(t.compo, m.compo, sizeof(t.compo));
    if (! (om = p7_gpu_oprofile_Create(M, m.abctype))) { status = p7G_EMEM; goto DONE; }
    if ((status = p7_gpu_oprofile_Load(om, &t)) != p7G_OK) { p7_gpu_oprofile_Destroy(om); om = NULL; goto DONE; }
  }
  *ret_om = om;
DONE:
  free(rbv); free(rwv); free(twv); free(rfv); free(tfv);
  rec_free(&m);
  return status;
}
