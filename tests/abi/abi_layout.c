/* Prints the size and field offsets of the structs include/p7_gpu.h passes
 * across the C ABI, as a C compiler lays them out, for comparison with the
 * ctypes mirrors in biobench2_amd/p7gpu.py: "<struct> <field> <offset>" lines
 * and one "<struct> sizeof <size>" line per struct. */
#include <stddef.h>
#include <stdio.h>
#include "p7_gpu.h"

#define F(S, f) printf("%s %s %zu\n", #S, #f, offsetof(S, f))
#define Z(S)    printf("%s sizeof %zu\n", #S, sizeof(S))

int main(void)
{
  Z(P7_GPU_TABLES);
  F(P7_GPU_TABLES, M); F(P7_GPU_TABLES, Kp); F(P7_GPU_TABLES, abctype); F(P7_GPU_TABLES, rbv);
  F(P7_GPU_TABLES, tbm_b); F(P7_GPU_TABLES, tec_b); F(P7_GPU_TABLES, tjb_b); F(P7_GPU_TABLES, base_b);
  F(P7_GPU_TABLES, bias_b); F(P7_GPU_TABLES, scale_b); F(P7_GPU_TABLES, rwv); F(P7_GPU_TABLES, twv);
  F(P7_GPU_TABLES, xw); F(P7_GPU_TABLES, scale_w); F(P7_GPU_TABLES, base_w); F(P7_GPU_TABLES, ddbound_w);
  F(P7_GPU_TABLES, rfv); F(P7_GPU_TABLES, tfv); F(P7_GPU_TABLES, xf); F(P7_GPU_TABLES, L);
  F(P7_GPU_TABLES, mode); F(P7_GPU_TABLES, nj); F(P7_GPU_TABLES, evparam); F(P7_GPU_TABLES, compo);
  Z(P7_GPU_HITS);
  F(P7_GPU_HITS, nullsc); F(P7_GPU_HITS, usc); F(P7_GPU_HITS, filtersc); F(P7_GPU_HITS, vfsc);
  F(P7_GPU_HITS, fwdsc); F(P7_GPU_HITS, bcksc); F(P7_GPU_HITS, stage); F(P7_GPU_HITS, nexpected);
  F(P7_GPU_HITS, nregions); F(P7_GPU_HITS, nclustered);
  Z(P7_GPU_STATS);
  F(P7_GPU_STATS, nmodels); F(P7_GPU_STATS, nseqs); F(P7_GPU_STATS, nres); F(P7_GPU_STATS, nnodes);
  F(P7_GPU_STATS, n_past_msv); F(P7_GPU_STATS, n_past_bias); F(P7_GPU_STATS, n_past_vit);
  F(P7_GPU_STATS, n_past_fwd); F(P7_GPU_STATS, n_msv_full); F(P7_GPU_STATS, cells_msv);
  F(P7_GPU_STATS, cells_bias); F(P7_GPU_STATS, cells_vit); F(P7_GPU_STATS, cells_fwd);
  F(P7_GPU_STATS, cells_bck); F(P7_GPU_STATS, ms_msv); F(P7_GPU_STATS, ms_bias); F(P7_GPU_STATS, ms_vit);
  F(P7_GPU_STATS, ms_fwd); F(P7_GPU_STATS, ms_bck); F(P7_GPU_STATS, ms_total); F(P7_GPU_STATS, ms_ssv);
  F(P7_GPU_STATS, ms_domains); F(P7_GPU_STATS, n_regions); F(P7_GPU_STATS, n_clustered);
  F(P7_GPU_STATS, ms_null2);
  F(P7_GPU_STATS, ms_multidomain);
  Z(P7_GPU_REPORT);
  F(P7_GPU_REPORT, qname); F(P7_GPU_REPORT, qacc); F(P7_GPU_REPORT, M); F(P7_GPU_REPORT, tname);
  F(P7_GPU_REPORT, tacc); F(P7_GPU_REPORT, tdesc); F(P7_GPU_REPORT, Z); F(P7_GPU_REPORT, domZ);
  F(P7_GPU_REPORT, E); F(P7_GPU_REPORT, incE); F(P7_GPU_REPORT, domE); F(P7_GPU_REPORT, incdomE);
  F(P7_GPU_REPORT, textw); F(P7_GPU_REPORT, show_accessions); F(P7_GPU_REPORT, show_alignments);
  return 0;
}
