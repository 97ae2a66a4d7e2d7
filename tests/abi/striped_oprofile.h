/* A test double of the P7_OPROFILE fields that the de-striping shim of the
 * integration guide reads (HMMER/src/impl_sse/impl_sse.h:76-143: the vector
 * types, the striped table pointers, the scalars), filled here from the
 * striped vectors of the reference's pressed profiles, so that the shim's
 * code compiles and runs as written. Only these fields; nothing else of the
 * reference's structure. */
#ifndef STRIPED_OPROFILE_INCLUDED
#define STRIPED_OPROFILE_INCLUDED
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <emmintrin.h>

/* vectors per row: 16 uchars, 8 int16, 4 floats, at least 2 (impl_sse.h:24-27) */
#define p7O_NQB(M) (((((M) - 1) / 16) + 1) > 2 ? ((((M) - 1) / 16) + 1) : 2)
#define p7O_NQW(M) (((((M) - 1) / 8) + 1) > 2 ? ((((M) - 1) / 8) + 1) : 2)
#define p7O_NQF(M) (((((M) - 1) / 4) + 1) > 2 ? ((((M) - 1) / 4) + 1) : 2)

typedef struct { int type, K, Kp; } STRIPED_ALPHABET;

typedef struct {
  __m128i **rbv;                 /* [Kp][Q16] match costs, uchar               */
  uint8_t   tbm_b, tec_b, tjb_b;
  float     scale_b;
  uint8_t   base_b, bias_b;
  __m128i **rwv;                 /* [Kp][Q8] match scores, int16               */
  __m128i  *twv;                 /* [8*Q8]: per q BM MM IM DM MD MI II, then DD */
  int16_t   xw[4][2];
  float     scale_w;
  int16_t   base_w, ddbound_w;
  __m128  **rfv;                 /* [Kp][Q4] match odds                        */
  __m128   *tfv;                 /* [8*Q4], as twv                             */
  float     xf[4][2];
  float     evparam[6];
  float     compo[20];
  const STRIPED_ALPHABET *abc;
  int       M, L, mode;
  float     nj;
} P7_OPROFILE;

#endif
