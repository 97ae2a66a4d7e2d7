/* Runs the de-striping shim of the integration guide (p7_oprofile_ToGPU,
 * compiled from the guide's own code block against tests/abi/striped_oprofile.h)
 * on the striped vectors of one model of the reference's pressed profiles,
 * and compares the library tables it loads (p7_gpu_oprofile_Load) with the
 * library's own reading of the same pressed files
 * (p7_gpu_oprofile_ReadPressed), field by field.
 *
 *   destripe_main <dump> <h3f> <h3p> <which>
 *
 * <dump> holds the model's striped vectors and scalars as the test wrote them
 * from the pressed files: int32 M, abctype, Kp, K, L, mode; uint8 tbm_b,
 * tec_b, tjb_b, base_b, bias_b and 3 pad bytes; float scale_b, scale_w, nj;
 * int16 base_w, ddbound_w, xw[8]; float xf[8], evparam[6], compo[20]; then
 * rbv [Kp][Q16][16] uint8, rwv [Kp][Q8][8] int16, twv [8*Q8][8] int16,
 * rfv [Kp][Q4][4] float, tfv [8*Q4][4] float.
 * Prints "OK" or one "MISMATCH <field>" line per differing field. */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "striped_oprofile.h"
#include "p7_gpu.h"

P7_GPU_OPROFILE *p7_oprofile_ToGPU(const P7_OPROFILE *om);

static int rd(FILE *fp, void *p, size_t n) { return fread(p, 1, n, fp) == n; }

static void *vecs(size_t n, size_t bytes)       /* n 16-byte vectors, aligned */
{
  void *p = NULL;
  if (posix_memalign(&p, 16, n * 16)) return NULL;
  memset(p, 0, n * 16);
  (void) bytes;
  return p;
}

int main(int argc, char **argv)
{
  if (argc != 5) { fprintf(stderr, "usage: %s dump h3f h3p which\n", argv[0]); return 2; }
  FILE *fp = fopen(argv[1], "rb");
  if (! fp) { perror(argv[1]); return 2; }
  int32_t hdr[6];
  uint8_t b8[8];
  float f3[3];
  int16_t w[10];
  P7_OPROFILE om;
  STRIPED_ALPHABET abc;
  memset(&om, 0, sizeof(om));
  if (! rd(fp, hdr, sizeof(hdr)) || ! rd(fp, b8, sizeof(b8)) || ! rd(fp, f3, sizeof(f3)) || ! rd(fp, w, sizeof(w)) ||
      ! rd(fp, om.xf, sizeof(om.xf)) || ! rd(fp, om.evparam, sizeof(om.evparam)) || ! rd(fp, om.compo, sizeof(om.compo))) {
    fprintf(stderr, "short dump header\n"); return 2;
  }
  om.M = hdr[0]; abc.type = hdr[1]; abc.Kp = hdr[2]; abc.K = hdr[3]; om.L = hdr[4]; om.mode = hdr[5];
  om.abc = &abc;
  om.tbm_b = b8[0]; om.tec_b = b8[1]; om.tjb_b = b8[2]; om.base_b = b8[3]; om.bias_b = b8[4];
  om.scale_b = f3[0]; om.scale_w = f3[1]; om.nj = f3[2];
  om.base_w = w[0]; om.ddbound_w = w[1];
  memcpy(om.xw, w + 2, sizeof(om.xw));
  const int M = om.M, Kp = abc.Kp, Q16 = p7O_NQB(M), Q8 = p7O_NQW(M), Q4 = p7O_NQF(M);
  om.rbv = malloc(sizeof(__m128i *) * Kp);
  om.rwv = malloc(sizeof(__m128i *) * Kp);
  om.rfv = malloc(sizeof(__m128 *) * Kp);
  int ok = om.rbv && om.rwv && om.rfv;
  for (int x = 0; ok && x < Kp; x++) { om.rbv[x] = vecs(Q16, 16); ok = om.rbv[x] && rd(fp, om.rbv[x], 16 * (size_t) Q16); }
  for (int x = 0; ok && x < Kp; x++) { om.rwv[x] = vecs(Q8, 16); ok = om.rwv[x] && rd(fp, om.rwv[x], 16 * (size_t) Q8); }
  om.twv = vecs(8 * (size_t) Q8, 16);
  ok = ok && om.twv && rd(fp, om.twv, 16 * 8 * (size_t) Q8);
  for (int x = 0; ok && x < Kp; x++) { om.rfv[x] = vecs(Q4, 16); ok = om.rfv[x] && rd(fp, om.rfv[x], 16 * (size_t) Q4); }
  om.tfv = vecs(8 * (size_t) Q4, 16);
  ok = ok && om.tfv && rd(fp, om.tfv, 16 * 8 * (size_t) Q4);
  fclose(fp);
  if (! ok) { fprintf(stderr, "short dump\n"); return 2; }

  P7_GPU_OPROFILE *gom = p7_oprofile_ToGPU(&om);
  P7_GPU_OPROFILE *rom = NULL;
  int st = p7_gpu_oprofile_ReadPressed(argv[2], argv[3], atoi(argv[4]), &rom);
  if (! gom || st != p7G_OK) { fprintf(stderr, "load failed (%d): %s\n", st, p7_gpu_LastError()); return 2; }
  P7_GPU_TABLES a, b;
  p7_gpu_oprofile_GetTables(gom, &a);
  p7_gpu_oprofile_GetTables(rom, &b);
  int bad = 0;
#define SAME(f) do { if (memcmp(&a.f, &b.f, sizeof(a.f))) { printf("MISMATCH %s\n", #f); bad++; } } while (0)
#define SAMEA(f, n) do { if (memcmp(a.f, b.f, (n) * sizeof(*a.f))) { printf("MISMATCH %s\n", #f); bad++; } } while (0)
  SAME(M); SAME(Kp); SAME(abctype);
  if (a.M != b.M || a.Kp != b.Kp) { printf("MISMATCH shape\n"); return 1; }
  SAMEA(rbv, (size_t) Kp * M); SAMEA(rwv, (size_t) Kp * M); SAMEA(twv, (size_t) 8 * M);
  SAMEA(rfv, (size_t) Kp * M); SAMEA(tfv, (size_t) 8 * M);
  SAME(tbm_b); SAME(tec_b); SAME(tjb_b); SAME(base_b); SAME(bias_b); SAME(scale_b);
  SAME(xw); SAME(scale_w); SAME(base_w); SAME(ddbound_w); SAME(xf); SAME(L); SAME(mode); SAME(nj);
  SAME(evparam);
  if (memcmp(a.compo, b.compo, sizeof(float) * abc.K)) { printf("MISMATCH compo\n"); bad++; }
  if (! bad) printf("OK\n");
  p7_gpu_oprofile_Destroy(gom);
  p7_gpu_oprofile_Destroy(rom);
  return bad ? 1 : 0;
}
