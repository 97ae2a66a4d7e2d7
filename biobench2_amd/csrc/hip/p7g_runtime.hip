// C ABI of libp7gpu: device setup, device copies of profiles and target
// databases, batched filter stages, the device pipeline, and the
// single-target drop-ins. Every pipeline object owns a HIP stream and its
// device work area, so pipelines driven from different host threads overlap
// on the device; the single-target and batched calls share one library
// stream. All streams are created at the default priority.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include <mutex>
#include <map>
#include <atomic>
#include <set>
#include <chrono>
#include <algorithm>
#include <string>
#include <vector>
#include <ctype.h>

#include "p7_gpu.h"
#include "p7g_internal.h"
#include "p7g_device.h"

extern "C" const char *p7g_host_error(void);

static hipStream_t g_stream = nullptr;
static int         g_device = -1;
static std::mutex  g_lock;     // serialises calls that share the library stream and its work area
static std::mutex  g_devlock;  // guards building the device copy of a profile (innermost lock)

// Device work area of one stream: per-slot results, survivor lists, the
// Forward scale-factor pool, timing events.
struct Work {
  hipStream_t st = nullptr;
  int64_t     cap = 0;
  DevResults  r;
  DevLists    l;
  int32_t    *slots = nullptr;      // explicit slot list for batched calls
  int64_t    *scale_off = nullptr;
  void       *outbuf = nullptr;     // one per-slot array in the caller's order
  int32_t    *tilecnt = nullptr;    // per-tile counts of the list compaction
  unsigned long long *stats = nullptr;
  float      *pool = nullptr;
  int64_t     pool_size = 0;
  uint16_t   *bnd = nullptr;        // boundary column of the chunked ungapped pass
  int64_t     bnd_cap = 0;          // its length in entries (residue bytes)
  void       *mem = nullptr;
  hipEvent_t  ev[10];
  bool        ev_ready = false;
  DomBufs     dom;                  // domain decoding: row pool and per-slot offsets (made on first use)
  DomOut      dout;                 // its per-slot results and region records
  // null2 of the regions: per region its area offset, correction and status,
  // the list of regions decoded, count blocks, and the pool (grown on use)
  int64_t    *n2off = nullptr;
  float      *n2dc = nullptr;
  int32_t    *n2st = nullptr, *n2list = nullptr, *n2cnt = nullptr;
  float      *n2pool = nullptr;
  int64_t     n2size = 0;
  float      *n2oa = nullptr;           // optimal accuracy score and alignment of each region
  int4       *n2ali = nullptr;
  int64_t    *n2poff = nullptr;         // the alignments' states: per region offset and count,
  int32_t    *n2pn = nullptr;
  uint32_t   *n2path = nullptr;         // and the words (grown on use)
  int64_t     n2path_cap = 0;
  // multidomain regions: per region its job and result, the envelope copies'
  // source/destination/length, one count block (grown on use)
  MdJob      *mdjobs = nullptr;
  MdOut      *mdout = nullptr;
  int64_t    *mdsrc = nullptr, *mddst = nullptr;
  int32_t    *mdlen = nullptr, *mdcnt = nullptr;
  int64_t     md_cap = 0;
  float      *mdpool = nullptr;           // the multidomain regions' areas (grown on use)
  int64_t     mdsize = 0;
  float      *mdn2 = nullptr;             // their residue null2 scores, gathered for the host
  int64_t     mdn2_cap = 0;
  int64_t     pool_cap = 0;               // the caller's cap on each of the two pools (floats; 0 none)
  hipStream_t st2 = nullptr;              // their stream, beside st (made on first use)
  hipEvent_t  md_thr = nullptr;           // on st2: the regions' traces' thresholds are filled
  void       *dom_mem = nullptr;
  int64_t     dom_cap = 0;          // slots the per-slot arrays hold
};
static Work g_work;

// Scratch of p7_gpu_MSVFilterBatchNoAlloc, for callers that cannot allocate:
// per database slot the ungapped pass's maximum and the final xJ (2 + 2
// bytes), per listed target its slot, the J-state list, the score and the
// status (4 x 4 bytes), and a block of 16 counters. The array is a host-side
// name only: kernels and copies take the address hipGetSymbolAddress gives on
// each device (g_scratch, taken by p7_gpu_Init), never the name itself.
#define P7G_NOALLOC_MAXSEQ ((int64_t) 1 << 20)
#define NOALLOC_PART(bytes) ((((size_t) (bytes)) + 255) & ~(size_t) 255)
#define P7G_NOALLOC_BYTES (2 * NOALLOC_PART(2 * P7G_NOALLOC_MAXSEQ) + 4 * NOALLOC_PART(4 * P7G_NOALLOC_MAXSEQ) + 256)
static __device__ __align__(256) unsigned char scratch[P7G_NOALLOC_BYTES];
#define P7G_MAXDEV 64
static void *g_scratch[P7G_MAXDEV];               // scratch's address on each device
static int32_t g_na_slots[P7G_NOALLOC_MAXSEQ];   // host staging of the slot list (under g_lock)

#define HIPCHK(call) do { hipError_t e_ = (call); if (e_ != hipSuccess) { \
      p7g_set_error("HIP error %s at %s:%d (%s)", hipGetErrorString(e_), __FILE__, __LINE__, #call); \
      return p7G_EDEVICE; } } while (0)

extern "C" int
p7_gpu_Init(int device)
{
  std::lock_guard<std::mutex> lk(g_lock);
  if (g_stream && device == g_device) return p7G_OK;
  int n = 0;
  HIPCHK(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) { p7g_set_error("no HIP device %d (%d present)", device, n); return p7G_EINVAL; }
  HIPCHK(hipSetDevice(device));
  if (device < P7G_MAXDEV && ! g_scratch[device])
    HIPCHK(hipGetSymbolAddress(&g_scratch[device], HIP_SYMBOL(scratch)));
  if (g_stream) { (void) hipStreamDestroy(g_stream); g_stream = nullptr; }
  HIPCHK(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking));
  g_work.st = g_stream;
  g_device = device;
  return p7G_OK;
}

static int ensure_init(void);

static int
ensure_init(void)
{
  if (g_stream) return p7G_OK;
  return p7_gpu_Init(0);
}

extern "C" int
p7_gpu_Synchronize(void)
{
  if (! g_stream) return p7G_OK;
  HIPCHK(hipStreamSynchronize(g_stream));
  return p7G_OK;
}

extern "C" const char *p7_gpu_LastError(void) { return p7g_host_error(); }

/*****************************************************************
 * Device profile: re-laid-out tables for 16/64-lane blocks, per-length
 * constants, and the bias-filter HMM, in one allocation.
 *****************************************************************/
static std::atomic<unsigned long long> g_serial(0);   // identifies device profile copies and databases

struct p7g_dev_profile_s {
  DevProfile p;
  size_t     bytes;                  // device memory of the tables
  int        Lmax;
  unsigned long long serial;
  struct p7g_dev_profile_s *older;   // copies replaced by a longer length range, kept until the profile goes
};

extern "C" void
p7g_dev_profile_free(struct p7g_dev_profile_s *d)
{
  while (d) {
    struct p7g_dev_profile_s *o = d->older;
    if (d->p.blob) (void) hipFree(d->p.blob);
    delete d;
    d = o;
  }
}

static const int vf_nd_set[] = { 1, 2, 3, 4, 5, 6, 8, 10, 12, 16, 20, 24, 32, 40 };
static const int fb_b_set[]  = { 1, 2, 3, 4, 5, 6, 7, 8, 10, 12, 14, 16, 20, 24, 28, 32, 40, 48, 64, 80 };
static const int ssv32_set[] = { 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 18, 20, 22, 24, 26, 28, 30, 32 };
static const int ssv64_set[] = { 8, 9, 10, 11, 12, 13, 14, 15, 16, 18, 20, 22, 24, 26, 28, 30, 32, 36, 40 };
static const int ssvc_set[]  = { 11, 12, 13, 14, 15, 16, 18 };   // chunk shapes (launch_ssv_chunked)

static int pick(const int *set, int nset, int need)
{
  for (int i = 0; i < nset; i++) if (set[i] >= need) return set[i];
  return -1;
}
static int cdiv(int a, int b) { return (a + b - 1) / b; }

static inline int pack2(int lo, int hi) { return (int) (((unsigned) lo & 0xffffu) | ((unsigned) hi << 16)); }

static int
dev_profile_build(P7_GPU_OPROFILE *om, int Lmax_need, struct p7g_dev_profile_s **ret)
{
  const int M = om->M, Kp = om->abc->Kp;
  const int NR = P7G_NROWS;                          // residue rows incl. P7G_PADCODE
  auto *d = new p7g_dev_profile_s();
  d->older = nullptr;
  DevProfile &p = d->p;
  memset(&p, 0, sizeof(p));
  p.M = M; p.Kp = Kp; p.nrows = NR; p.K = om->abc->K;

  // ---- SSV / MSV-full tables: int16 (bias - rbv) ----
  auto rsv = [&](int x, int k) -> int {               // score at node k (1-based) or "impossible"
    if (x >= Kp || k < 1 || k > M) return -32768;
    return (int) om->bias_b - (int) om->rbv[(size_t) x * M + k - 1];
  };
  // the shapes of the ungapped pass this model fits: 16, 32, 64 and 8 lanes per
  // target. With 8 lanes two targets share each 16-lane DPP row, so the first
  // lane of the second one takes the last position of the first: the shape is
  // sized to leave that position past M, where every row holds the floor.
  const int ssvG[P7G_SSV_NVAR] = { 16, 32, 64, 8 };
  std::vector<int> ssv[P7G_SSV_NVAR];
  for (int v = 0; v < P7G_SSV_NVAR; v++) {
    SsvVariant &sv = p.ssv[v];
    const int G = ssvG[v];
    int ND = -1;
    if (G == 16)      ND = (M <= 1024) ? std::max(1, cdiv(M, 32)) : -1;
    else if (G == 32) ND = (M <= 2048) ? pick(ssv32_set, (int) (sizeof(ssv32_set)/sizeof(int)), cdiv(M, 64)) : -1;
    else if (G == 64) ND = pick(ssv64_set, (int) (sizeof(ssv64_set)/sizeof(int)), cdiv(M, 128));
    else              ND = (M <= 352) ? cdiv(M + 1, 16) : -1;   // where it measured faster (see ssv_def)
    if (ND > 0 && (size_t) NR * G * ND * 4 + 256 > P7G_LDS_MAX) ND = -1;   // the table (and a word) must fit in LDS
    if (ND < 0) { sv.G = 0; sv.ND = 0; sv.dwords = 0; continue; }
    sv.G = G; sv.ND = ND;
    sv.dwords = (size_t) NR * G * ND;
    ssv[v].assign(sv.dwords, (int) 0x80008000u);
    // quads: [x][q][z][4] for dwords 0..4*NQ-1; tail: [x][t][z] for dwords 4*NQ + t
    const int B = 2 * ND, NQ = ND / 4, NT = ND % 4;
    const size_t tail0 = (size_t) NR * NQ * G * 4;
    for (int x = 0; x < NR; x++)
      for (int z = 0; z < G; z++)
        for (int dd = 0; dd < ND; dd++) {
          const int val = pack2(rsv(x, z*B + dd + 1), rsv(x, z*B + dd + ND + 1));
          if (dd < 4 * NQ) ssv[v][(((size_t) x * NQ + dd / 4) * G + z) * 4 + dd % 4] = val;
          else             ssv[v][tail0 + ((size_t) x * NT + (dd - 4 * NQ)) * G + z] = val;
        }
  }
  // longer models: two or three chunks of 64 lanes, each table in LDS (profile-chunk tiling)
  std::vector<int> ssvc[P7G_SSV_MAXCHUNK];
  p.ssv_nchunk = 0;
  if (p.ssv[2].G == 0) {
    const int nset = (int) (sizeof(ssvc_set)/sizeof(int));
    int nc = 2, ND = pick(ssvc_set, nset, cdiv(M, 2 * 128));
    if (ND < 0) { nc = 3; ND = pick(ssvc_set, nset, cdiv(M, 3 * 128)); }
    if (ND < 0) { delete d; p7g_set_error("model too long for the device MSV stage (M=%d)", M); return p7G_EINVAL; }
    p.ssv_nchunk = nc;
    for (int c = 0; c < nc; c++) {
      SsvVariant &sv = p.ssvc[c];
      const int G = 64, B = 2 * ND, NQ = ND / 4, NT = ND % 4, k0 = c * G * B;
      sv.G = G; sv.ND = ND; sv.dwords = (size_t) NR * G * ND;
      ssvc[c].assign(sv.dwords, (int) 0x80008000u);
      const size_t tail0 = (size_t) NR * NQ * G * 4;
      for (int x = 0; x < NR; x++)
        for (int z = 0; z < G; z++)
          for (int dd = 0; dd < ND; dd++) {
            const int val = pack2(rsv(x, k0 + z*B + dd + 1), rsv(x, k0 + z*B + dd + ND + 1));
            if (dd < 4 * NQ) ssvc[c][(((size_t) x * NQ + dd / 4) * G + z) * 4 + dd % 4] = val;
            else             ssvc[c][tail0 + ((size_t) x * NT + (dd - 4 * NQ)) * G + z] = val;
          }
    }
  }
  // 8 lanes up to M = 352, 16 up to M = 1024, else 64; chunked beyond one LDS
  // table (on MI355X the 8-lane shape measured 1.03-1.36x faster than 16 lanes
  // for M = 40..300 in the tuner's sample runs; at M = 400 the samples were
  // even, but the full pass ran 10% slower, so larger models are not offered it)
  p.ssv_def = (p.ssv[3].G != 0 && M <= 352) ? 3 : (p.ssv[0].G != 0) ? 0 : (p.ssv[2].G != 0) ? 2 : -1;
  p.mf_ND = pick(vf_nd_set, (int) (sizeof(vf_nd_set)/sizeof(int)), std::max(1, cdiv(cdiv(M, 64), 2)));
  if (p.mf_ND < 0) { delete d; p7g_set_error("model too long for the device MSV stage (M=%d)", M); return p7G_EINVAL; }
  p.mf_dwords = (size_t) NR * 64 * p.mf_ND;
  std::vector<int> mf(p.mf_dwords);
  {
    const int ND = p.mf_ND, B = 2 * ND;
    for (int x = 0; x < NR; x++)
      for (int z = 0; z < 64; z++)
        for (int dd = 0; dd < ND; dd++)
          mf[((size_t) x * ND + dd) * 64 + z] = pack2(rsv(x, z*B + dd + 1), rsv(x, z*B + dd + ND + 1));
  }

  // ---- Viterbi tables: int16 pairs, capacity >= M+1 ----
  p.vf_ND = pick(vf_nd_set, (int) (sizeof(vf_nd_set)/sizeof(int)), cdiv(cdiv(M + 1, 64), 2));
  if (p.vf_ND < 0) { delete d; p7g_set_error("model too long for the device Viterbi stage (M=%d)", M); return p7G_EINVAL; }
  const int G = 64;
  p.vf_dwords = (size_t) (NR + P7G_VF_NT) * G * p.vf_ND;
  std::vector<int> vf(p.vf_dwords);
  {
    const int ND = p.vf_ND, B = 2 * ND;
    auto tw = [&](int t, int k) -> int { return (k >= 1 && k <= M) ? (int) om->twv[(size_t) t * M + k - 1] : -32768; };
    auto ew = [&](int x, int k) -> int { return (x < Kp && k >= 1 && k <= M) ? (int) om->rwv[(size_t) x * M + k - 1] : -32768; };
    auto tv = [&](int row, int k) -> int {
      switch (row) {
      case VT_BM: return tw(p7G_BM, k);
      case VT_MM: return tw(p7G_MM, k);
      case VT_IM: return tw(p7G_IM, k);
      case VT_DM: return tw(p7G_DM, k);
      case VT_MDS: return tw(p7G_MD, k - 1);      // M_{k-1} -> D_k
      case VT_MI: return tw(p7G_MI, k);
      case VT_II: return tw(p7G_II, k);
      case VT_DDS: return tw(p7G_DD, k - 1);      // D_{k-1} -> D_k
      }
      return -32768;
    };
    for (int r = 0; r < NR + P7G_VF_NT; r++)
      for (int z = 0; z < G; z++)
        for (int dd = 0; dd < ND; dd++) {
          const int k0 = z*B + dd + 1, k1 = z*B + dd + ND + 1;
          int a, b;
          if (r < NR) { a = ew(r, k0); b = ew(r, k1); }
          else        { a = tv(r - NR, k0); b = tv(r - NR, k1); }
          vf[((size_t) r * ND + dd) * G + z] = pack2(a, b);
        }
  }

  // ---- Forward/Backward tables: float, capacity >= M+1 ----
  p.fb_B = pick(fb_b_set, (int) (sizeof(fb_b_set)/sizeof(int)), cdiv(M + 1, 64));
  if (p.fb_B < 0) { delete d; p7g_set_error("model too long for the device Forward stage (M=%d)", M); return p7G_EINVAL; }
  p.fb_floats = (size_t) (NR + P7G_FB_NT) * 64 * p.fb_B;
  std::vector<float> fb(p.fb_floats);
  {
    const int B = p.fb_B;
    auto tf = [&](int t, int k) -> float { return (k >= 1 && k <= M) ? om->tfv[(size_t) t * M + k - 1] : 0.0f; };
    auto ef = [&](int x, int k) -> float { return (x < Kp && k >= 1 && k <= M) ? om->rfv[(size_t) x * M + k - 1] : 0.0f; };
    auto tv = [&](int row, int k) -> float {
      switch (row) {
      case VT_BM: return tf(p7G_BM, k);
      case VT_MM: return tf(p7G_MM, k);
      case VT_IM: return tf(p7G_IM, k);
      case VT_DM: return tf(p7G_DM, k);
      case VT_MDS: return tf(p7G_MD, k - 1);
      case VT_MI: return tf(p7G_MI, k);
      case VT_II: return tf(p7G_II, k);
      case VT_DDS: return tf(p7G_DD, k - 1);
      case FT_MMN: return tf(p7G_MM, k + 1);     // M_k -> M_{k+1}
      case FT_IMN: return tf(p7G_IM, k + 1);
      case FT_DMN: return tf(p7G_DM, k + 1);
      case FT_MD:  return tf(p7G_MD, k);         // M_k -> D_{k+1}
      case FT_DD:  return tf(p7G_DD, k);
      }
      return 0.0f;
    };
    for (int r = 0; r < NR + P7G_FB_NT; r++)
      for (int z = 0; z < 64; z++)
        for (int j = 0; j < B; j++) {
          const int k = z*B + j + 1;
          fb[((size_t) r * B + j) * 64 + z] = (r < NR) ? ef(r, k) : tv(r - NR, k);
        }
  }

  // ---- per-length constants ----
  P7G_LENCFG lc;
  const int Lmax = std::max(Lmax_need, 65536);
  if (p7g_lencfg_build(om, Lmax, &lc) != p7G_OK) { delete d; return p7G_EMEM; }
  p.Lmax = Lmax;

  // ---- bias filter HMM: odds ratios of the two states (p7_bg_SetFilter) ----
  std::vector<float> eo((size_t) NR * 2, 1.0f);
  {
    P7_GPU_BG *bg = p7_gpu_bg_Create(om->abc->type);
    if (! bg) { p7g_lencfg_free(&lc); delete d; return p7G_EMEM; }
    float compo[P7G_MAXABET];
    for (int x = 0; x < om->abc->K; x++) compo[x] = om->compo[x];
    p7_gpu_bg_SetFilter(bg, M, compo);
    for (int x = 0; x < Kp; x++) { eo[x*2] = bg->eo[x][0]; eo[x*2+1] = bg->eo[x][1]; }
    p.ft1_0 = bg->ft[1][0]; p.ft1_1 = bg->ft[1][1]; p.pi0 = bg->pi[0]; p.pi1 = bg->pi[1];
    p7_gpu_bg_Destroy(bg);
  }

  // ---- scalars ----
  p.tbm = om->tbm_b; p.tec = om->tec_b; p.base = om->base_b; p.bias = om->bias_b; p.scale_b = om->scale_b;
  p.base_w = om->base_w; p.ddbound_w = om->ddbound_w; p.scale_w = om->scale_w;
  p.xwE_move = om->xw[p7G_XE][p7G_MOVE]; p.xwE_loop = om->xw[p7G_XE][p7G_LOOP];
  p.xwN_loop = om->xw[p7G_XN][p7G_LOOP]; p.xwC_loop = om->xw[p7G_XC][p7G_LOOP]; p.xwJ_loop = om->xw[p7G_XJ][p7G_LOOP];
  p.xfE_move = om->xf[p7G_XE][p7G_MOVE]; p.xfE_loop = om->xf[p7G_XE][p7G_LOOP];
  p.nj = om->nj;

  // ---- degenerate codes: canonical residues of each code, and how many ----
  std::vector<uint32_t> degen(32, 0u);
  for (int x = 0; x < Kp && x < 32; x++) {
    uint32_t m = 0;
    for (int y = 0; y < om->abc->K; y++) if (om->abc->degen[x][y]) m |= 1u << y;
    degen[x] = m | ((uint32_t) om->abc->ndegen[x] << 24);
  }

  // ---- one allocation ----
  const size_t L1 = (size_t) Lmax + 1;
  size_t sz[] = { p.ssv[0].dwords * 4, p.ssv[1].dwords * 4, p.ssv[2].dwords * 4, p.mf_dwords * 4, p.vf_dwords * 4,
                  p.fb_floats * 4, L1, L1 * 2, L1 * 4, L1 * 4, L1 * 4, L1 * 4, L1 * 4, eo.size() * 4,
                  p.ssv_nchunk > 0 ? p.ssvc[0].dwords * 4 : 0, p.ssv_nchunk > 1 ? p.ssvc[1].dwords * 4 : 0,
                  p.ssv_nchunk > 2 ? p.ssvc[2].dwords * 4 : 0, p.ssv[3].dwords * 4, degen.size() * 4 };
  const void *src[] = { ssv[0].data(), ssv[1].data(), ssv[2].data(), mf.data(), vf.data(), fb.data(), lc.tjb_b,
                        lc.xw_move, lc.xf_move, lc.xf_loop, lc.nullsc, lc.flen_a, lc.flen_b, eo.data(),
                        ssvc[0].data(), ssvc[1].data(), ssvc[2].data(), ssv[3].data(), degen.data() };
  const int NB = (int) (sizeof(sz) / sizeof(sz[0]));
  size_t offs[20], total = 0;
  for (int i = 0; i < NB; i++) { offs[i] = total; total += (sz[i] + 255) & ~(size_t) 255; }
  char *blob = nullptr;
  d->bytes = total;
  hipError_t e = hipMalloc((void **) &blob, total);
  if (e == hipSuccess) {
    for (int i = 0; i < NB && e == hipSuccess; i++)
      if (sz[i]) e = hipMemcpyAsync(blob + offs[i], src[i], sz[i], hipMemcpyHostToDevice, g_stream);
    if (e == hipSuccess) e = hipStreamSynchronize(g_stream);
  }
  p7g_lencfg_free(&lc);
  if (e != hipSuccess) {
    if (blob) (void) hipFree(blob);
    delete d;
    p7g_set_error("device profile upload failed: %s", hipGetErrorString(e));
    return p7G_EDEVICE;
  }
  p.blob = blob;
  const int ssv_blob[P7G_SSV_NVAR] = { 0, 1, 2, 17 };          // blob parts of the variants' tables
  for (int v = 0; v < P7G_SSV_NVAR; v++) p.ssv[v].tbl = p.ssv[v].G ? (int *) (blob + offs[ssv_blob[v]]) : nullptr;
  p.mf_tbl = (int *) (blob + offs[3]);
  p.vf_tbl = (int *) (blob + offs[4]);  p.fb_tbl = (float *) (blob + offs[5]);
  p.tjb = (uint8_t *) (blob + offs[6]); p.xw_move = (int16_t *) (blob + offs[7]);
  p.xf_move = (float *) (blob + offs[8]); p.xf_loop = (float *) (blob + offs[9]);
  p.nullsc = (float *) (blob + offs[10]); p.flen_a = (float *) (blob + offs[11]); p.flen_b = (float *) (blob + offs[12]);
  p.eo = (float *) (blob + offs[13]);
  p.degen = (const uint32_t *) (blob + offs[18]);
  for (int c = 0; c < P7G_SSV_MAXCHUNK; c++) p.ssvc[c].tbl = (c < p.ssv_nchunk) ? (int *) (blob + offs[14 + c]) : nullptr;
  d->Lmax = Lmax;
  d->serial = ++g_serial;
  *ret = d;
  return p7G_OK;
}

static int
dev_profile(const P7_GPU_OPROFILE *om_c, int Lmax_need, DevProfile **ret)
{
  P7_GPU_OPROFILE *om = (P7_GPU_OPROFILE *) om_c;      // the device copy is a cache of the const tables
  int status;
  if (om->M <= 0) { p7g_set_error("oprofile has no model"); return p7G_EINVAL; }
  if (om->mode != p7G_LOCAL && om->mode != p7G_UNILOCAL) { p7g_set_error("the filters need a local-mode profile"); return p7G_EINVAL; }
  std::lock_guard<std::mutex> lk(g_devlock);
  if (! om->dev || om->dev->Lmax < Lmax_need) {
    // a pipeline on another stream may still read the current copy: retire it, do not free it
    struct p7g_dev_profile_s *d = nullptr;
    if ((status = dev_profile_build(om, Lmax_need, &d)) != p7G_OK) return status;
    d->older = om->dev;
    om->dev = d;
  }
  *ret = &om->dev->p;
  return p7G_OK;
}

/*****************************************************************
 * Target databases in HBM
 *****************************************************************/
struct p7g_seqdb_s {
  int64_t  nseq, nres, bytes;
  int      Lmax, maxcode;
  int32_t *h_len;          // [slot]
  int64_t *h_perm;         // slot -> caller index
  int32_t *h_inv;          // caller index -> slot
  uint8_t *d_res;
  int64_t *d_off;
  int32_t *d_len;
  int32_t *d_perm;         // slot -> caller index, on the device
  char    *names;          // sequence names in caller order (when read from a file), NUL-separated
  int64_t *name_off;       // [caller index] offset into names
  unsigned long long serial;
  int2    *d_seg;          // row segments of the targets longer than P7G_SEG_ROWS (slot, first row)
  int64_t  nseg;           // their number
  int64_t  nseg_slots;     // the leading slots they cover
};

extern "C" void
p7_gpu_seqdb_Destroy(P7_GPU_SEQDB *db)
{
  if (! db) return;
  if (db->d_res) (void) hipFree(db->d_res);
  if (db->d_off) (void) hipFree(db->d_off);
  if (db->d_len) (void) hipFree(db->d_len);
  if (db->d_perm) (void) hipFree(db->d_perm);
  if (db->d_seg) (void) hipFree(db->d_seg);
  free(db->h_len); free(db->h_perm); free(db->h_inv);
  free(db->names); free(db->name_off);
  free(db);
}

// Row segments of the long targets for the ungapped MSV pass: a target of
// more than P7G_SEG_ROWS residues is cut into pieces of P7G_SEG_ROWS rows,
// which waves take like whole targets (each starting a model length earlier,
// see k_ssv). Slots run longest first, so the long targets are the leading ones.
static hipError_t
seqdb_segments(P7_GPU_SEQDB *db)
{
  std::vector<int2> seg;
  int64_t s = 0;
  for (; s < db->nseq && db->h_len[s] > P7G_SEG_ROWS; s++)
    for (int r0 = 0; r0 < db->h_len[s]; r0 += P7G_SEG_ROWS) seg.push_back(make_int2((int) s, r0));
  db->nseg = (int64_t) seg.size();
  db->nseg_slots = s;
  if (seg.empty()) return hipSuccess;
  hipError_t e = hipMalloc((void **) &db->d_seg, sizeof(int2) * seg.size());
  if (e == hipSuccess) e = hipMemcpyAsync(db->d_seg, seg.data(), sizeof(int2) * seg.size(), hipMemcpyHostToDevice, g_stream);
  if (e == hipSuccess) e = hipStreamSynchronize(g_stream);
  return e;
}

extern "C" int
p7_gpu_seqdb_Create(const uint8_t *residues, const int64_t *offset, const int32_t *length, int64_t nseq, P7_GPU_SEQDB **ret_db)
{
  int status;
  *ret_db = nullptr;
  if ((status = ensure_init()) != p7G_OK) return status;
  if (nseq <= 0 || nseq > INT32_MAX - 1) { p7g_set_error("bad database size %lld", (long long) nseq); return p7G_EINVAL; }
  P7_GPU_SEQDB *db = (P7_GPU_SEQDB *) calloc(1, sizeof(P7_GPU_SEQDB));
  if (db) db->serial = ++g_serial;
  if (! db) return p7G_EMEM;
  db->nseq = nseq;
  db->h_len  = (int32_t *) malloc(sizeof(int32_t) * (size_t) nseq);
  db->h_perm = (int64_t *) malloc(sizeof(int64_t) * (size_t) nseq);
  db->h_inv  = (int32_t *) malloc(sizeof(int32_t) * (size_t) nseq);
  int64_t *h_off = (int64_t *) malloc(sizeof(int64_t) * (size_t) nseq);
  if (! db->h_len || ! db->h_perm || ! db->h_inv || ! h_off) { free(h_off); p7_gpu_seqdb_Destroy(db); return p7G_EMEM; }

  for (int64_t i = 0; i < nseq; i++) {
    if (length[i] < 0 || length[i] > 100000000) { free(h_off); p7_gpu_seqdb_Destroy(db); p7g_set_error("bad length %d for sequence %lld", length[i], (long long) i); return p7G_EINVAL; }
    db->h_perm[i] = i;
  }
  // longest first; equal lengths keep the caller's order
  std::stable_sort(db->h_perm, db->h_perm + nseq, [&](int64_t a, int64_t b) { return length[a] > length[b]; });
  int64_t pos = 0;
  db->Lmax = 0;
  for (int64_t s = 0; s < nseq; s++) {
    const int64_t i = db->h_perm[s];
    db->h_inv[i] = (int32_t) s;
    db->h_len[s] = length[i];
    h_off[s] = pos;
    pos += ((int64_t) length[i] + 15) & ~(int64_t) 15;
    db->nres += length[i];
    if (length[i] > db->Lmax) db->Lmax = length[i];
  }
  db->bytes = pos + 16;
  uint8_t *h_res = (uint8_t *) malloc((size_t) db->bytes);
  if (! h_res) { free(h_off); p7_gpu_seqdb_Destroy(db); return p7G_EMEM; }
  memset(h_res, P7G_PADCODE, (size_t) db->bytes);
  int maxcode = 0;
  for (int64_t s = 0; s < nseq; s++) {
    const int64_t i = db->h_perm[s];
    const uint8_t *src = residues + offset[i];
    uint8_t *dst = h_res + h_off[s];
    for (int32_t k = 0; k < length[i]; k++) { const int c = src[k]; if (c > maxcode) maxcode = c; dst[k] = (uint8_t) c; }
  }
  db->maxcode = maxcode;
  if (maxcode >= 29) { free(h_res); free(h_off); p7_gpu_seqdb_Destroy(db); p7g_set_error("residue code %d is not a digital residue", maxcode); return p7G_EINVAL; }

  std::vector<int32_t> h_perm32((size_t) nseq);
  for (int64_t s = 0; s < nseq; s++) h_perm32[(size_t) s] = (int32_t) db->h_perm[s];
  hipError_t e = hipMalloc((void **) &db->d_res, (size_t) db->bytes);
  if (e == hipSuccess) e = hipMalloc((void **) &db->d_off, sizeof(int64_t) * (size_t) nseq);
  if (e == hipSuccess) e = hipMalloc((void **) &db->d_len, sizeof(int32_t) * (size_t) nseq);
  if (e == hipSuccess) e = hipMalloc((void **) &db->d_perm, sizeof(int32_t) * (size_t) nseq);
  if (e == hipSuccess) e = hipMemcpyAsync(db->d_res, h_res, (size_t) db->bytes, hipMemcpyHostToDevice, g_stream);
  if (e == hipSuccess) e = hipMemcpyAsync(db->d_off, h_off, sizeof(int64_t) * (size_t) nseq, hipMemcpyHostToDevice, g_stream);
  if (e == hipSuccess) e = hipMemcpyAsync(db->d_len, db->h_len, sizeof(int32_t) * (size_t) nseq, hipMemcpyHostToDevice, g_stream);
  if (e == hipSuccess) e = hipMemcpyAsync(db->d_perm, h_perm32.data(), sizeof(int32_t) * (size_t) nseq, hipMemcpyHostToDevice, g_stream);
  if (e == hipSuccess) e = hipStreamSynchronize(g_stream);
  if (e == hipSuccess) e = seqdb_segments(db);
  free(h_res); free(h_off);
  if (e != hipSuccess) { p7g_set_error("database upload failed: %s", hipGetErrorString(e)); p7_gpu_seqdb_Destroy(db); return p7G_EDEVICE; }
  *ret_db = db;
  return p7G_OK;
}

/*****************************************************************
 * Resident databases from files: FASTA (as hmmpgmd's sequence cache loads
 * it, HMMER/src/cachedb.c:41-280) and a binary image of the device layout
 *****************************************************************/
extern "C" int
p7_gpu_seqdb_ReadFasta(const char *path, int abctype, P7_GPU_SEQDB **ret_db)
{
  *ret_db = nullptr;
  const P7G_ALPHABET *abc = p7g_alphabet(abctype);
  if (! abc) { p7g_set_error("unknown alphabet type %d", abctype); return p7G_EINVAL; }
  FILE *fp = fopen(path, "rb");
  if (! fp) { p7g_set_error("can't open %s", path); return p7G_ENOTFOUND; }
  std::vector<uint8_t> res;
  std::vector<int64_t> off;
  std::vector<int32_t> len;
  std::vector<char>    names;
  std::vector<int64_t> noff;
  std::vector<char> buf(1 << 20);
  int64_t lineno = 0;
  // in_header: reading the name (first word) of a '>' line; each header adds one NUL-terminated name
  bool in_header = false, at_line_start = true, first_line = true, skip_line = false;
  int status = p7G_OK;
  size_t got;
  while (status == p7G_OK && (got = fread(buf.data(), 1, buf.size(), fp)) > 0) {
    for (size_t i = 0; i < got; i++) {
      const unsigned char c = (unsigned char) buf[i];
      if (c == '\n') {
        if (in_header) { names.push_back('\0'); in_header = false; }
        lineno++; at_line_start = true; first_line = false; skip_line = false;
        continue;
      }
      if (skip_line) continue;
      if (at_line_start) {
        at_line_start = false;
        // the cache format starts with a '#' line of counts (cachedb.c:79-95): skip it
        if (first_line && c == '#') { skip_line = true; continue; }
        if (c == '>') {
          if (! len.empty()) len.back() = (int32_t) ((int64_t) res.size() - off.back());
          off.push_back((int64_t) res.size()); len.push_back(0);
          noff.push_back((int64_t) names.size());
          in_header = true;
          continue;
        }
      }
      if (in_header) {
        if (c == ' ' || c == '\t' || c == '\r') {
          if ((int64_t) names.size() > noff.back()) { names.push_back('\0'); in_header = false; skip_line = true; }   // the description is dropped
          continue;
        }
        names.push_back((char) c);
        continue;
      }
      if (c == ' ' || c == '\t' || c == '\r') continue;
      if (off.empty()) { p7g_set_error("%s line %lld: residues before the first '>' header", path, (long long) lineno + 1); status = p7G_EFORMAT; break; }
      const int x = (c < 128) ? abc->inmap[c] : -1;
      if (x < 0) { p7g_set_error("%s line %lld: illegal residue '%c'", path, (long long) lineno + 1, c); status = p7G_EFORMAT; break; }
      res.push_back((uint8_t) x);
    }
  }
  fclose(fp);
  if (status != p7G_OK) return status;
  if (in_header) names.push_back('\0');
  if (off.empty()) { p7g_set_error("%s holds no sequences", path); return p7G_EFORMAT; }
  len.back() = (int32_t) ((int64_t) res.size() - off.back());
  if (res.empty()) res.push_back(0);
  P7_GPU_SEQDB *db = nullptr;
  if ((status = p7_gpu_seqdb_Create(res.data(), off.data(), len.data(), (int64_t) off.size(), &db)) != p7G_OK) return status;
  db->names = (char *) malloc(names.size() ? names.size() : 1);
  db->name_off = (int64_t *) malloc(sizeof(int64_t) * noff.size());
  if (! db->names || ! db->name_off) { p7_gpu_seqdb_Destroy(db); return p7G_EMEM; }
  if (! names.empty()) memcpy(db->names, names.data(), names.size()); else db->names[0] = '\0';
  memcpy(db->name_off, noff.data(), sizeof(int64_t) * noff.size());
  *ret_db = db;
  return p7G_OK;
}

extern "C" const char *
p7_gpu_seqdb_Name(const P7_GPU_SEQDB *db, int64_t i)
{
  if (! db || ! db->names || i < 0 || i >= db->nseq) return nullptr;
  return db->names + db->name_off[i];
}

// Binary image: magic, nseq, nres, bytes, Lmax, maxcode, h_len[nseq], h_perm[nseq],
// the padded residues as laid out in HBM, then the names (size, bytes, offsets).
#define P7G_DBMAGIC 0x50374744u        // "DG7P"

extern "C" int
p7_gpu_seqdb_Save(const P7_GPU_SEQDB *db, const char *path)
{
  std::lock_guard<std::mutex> lk(g_lock);
  int status;
  if ((status = ensure_init()) != p7G_OK) return status;
  std::vector<uint8_t> h((size_t) db->bytes);
  HIPCHK(hipMemcpy(h.data(), db->d_res, (size_t) db->bytes, hipMemcpyDeviceToHost));
  FILE *fp = fopen(path, "wb");
  if (! fp) { p7g_set_error("can't write %s", path); return p7G_ESYS; }
  const uint32_t magic = P7G_DBMAGIC;
  int64_t namebytes = 0;                 // 0: no names
  if (db->names && db->name_off)
    for (int64_t i = 0; i < db->nseq; i++)
      namebytes = std::max<int64_t>(namebytes, db->name_off[i] + (int64_t) strlen(db->names + db->name_off[i]) + 1);
  bool ok = fwrite(&magic, 4, 1, fp) == 1 && fwrite(&db->nseq, 8, 1, fp) == 1 && fwrite(&db->nres, 8, 1, fp) == 1 &&
            fwrite(&db->bytes, 8, 1, fp) == 1 && fwrite(&db->Lmax, 4, 1, fp) == 1 && fwrite(&db->maxcode, 4, 1, fp) == 1 &&
            fwrite(db->h_len, 4, (size_t) db->nseq, fp) == (size_t) db->nseq &&
            fwrite(db->h_perm, 8, (size_t) db->nseq, fp) == (size_t) db->nseq &&
            fwrite(h.data(), 1, (size_t) db->bytes, fp) == (size_t) db->bytes &&
            fwrite(&namebytes, 8, 1, fp) == 1;
  if (ok && namebytes) ok = fwrite(db->names, 1, (size_t) namebytes, fp) == (size_t) namebytes &&
                            fwrite(db->name_off, 8, (size_t) db->nseq, fp) == (size_t) db->nseq;
  ok = ok && fwrite(&magic, 4, 1, fp) == 1;
  if (fclose(fp) != 0) ok = false;
  if (! ok) { p7g_set_error("failed writing %s", path); return p7G_ESYS; }
  return p7G_OK;
}

extern "C" int
p7_gpu_seqdb_Load(const char *path, P7_GPU_SEQDB **ret_db)
{
  int status;
  *ret_db = nullptr;
  FILE *fp = fopen(path, "rb");
  if (! fp) { p7g_set_error("can't open %s", path); return p7G_ENOTFOUND; }
  P7_GPU_SEQDB *db = (P7_GPU_SEQDB *) calloc(1, sizeof(P7_GPU_SEQDB));
  if (db) db->serial = ++g_serial;
  uint32_t magic = 0;
  int64_t namebytes = 0;
  std::vector<uint8_t> h;
  std::vector<int64_t> h_off;
  const char *why = "truncated or bad header";
  bool ok = db && fread(&magic, 4, 1, fp) == 1 && magic == P7G_DBMAGIC &&
            fread(&db->nseq, 8, 1, fp) == 1 && fread(&db->nres, 8, 1, fp) == 1 && fread(&db->bytes, 8, 1, fp) == 1 &&
            fread(&db->Lmax, 4, 1, fp) == 1 && fread(&db->maxcode, 4, 1, fp) == 1 &&
            db->nseq > 0 && db->nseq < INT32_MAX && db->bytes > 0 && db->maxcode >= 0 && db->maxcode < 29;
  if (ok) {
    db->h_len = (int32_t *) malloc(4 * (size_t) db->nseq);
    db->h_perm = (int64_t *) malloc(8 * (size_t) db->nseq);
    db->h_inv = (int32_t *) malloc(4 * (size_t) db->nseq);
    ok = db->h_len && db->h_perm && db->h_inv &&
         fread(db->h_len, 4, (size_t) db->nseq, fp) == (size_t) db->nseq &&
         fread(db->h_perm, 8, (size_t) db->nseq, fp) == (size_t) db->nseq;
    if (! ok) why = "truncated length or permutation table";
  }
  if (ok) {
    // The layout follows from the lengths: each target 16-byte aligned, padded
    // to 16, longest first. The kernels size their per-length tables from Lmax
    // and index them by each target's length, and the slot <-> caller maps must
    // be a permutation, so none of this is taken on trust.
    h_off.resize((size_t) db->nseq);
    for (int64_t i = 0; i < db->nseq; i++) db->h_inv[i] = -1;
    int64_t pos = 0, nres = 0;
    int Lmax = 0;
    for (int64_t s = 0; s < db->nseq && ok; s++) {
      const int64_t i = db->h_perm[s];
      const int L = db->h_len[s];
      if (L < 0 || L > 100000000)               { ok = false; why = "bad sequence length"; break; }
      if (s > 0 && L > db->h_len[s - 1])         { ok = false; why = "lengths not in decreasing order"; break; }
      if (i < 0 || i >= db->nseq)                { ok = false; why = "permutation entry out of range"; break; }
      if (db->h_inv[i] != -1)                    { ok = false; why = "permutation has a repeated entry"; break; }
      db->h_inv[i] = (int32_t) s;
      h_off[(size_t) s] = pos;
      pos += ((int64_t) L + 15) & ~(int64_t) 15;
      nres += L;
      if (L > Lmax) Lmax = L;
    }
    if (ok && (pos + 16 != db->bytes || nres != db->nres)) { ok = false; why = "residue count or size does not match the lengths"; }
    if (ok && Lmax != db->Lmax) { ok = false; why = "header's longest length does not match the lengths"; }
  }
  if (ok) {
    h.resize((size_t) db->bytes);
    ok = fread(h.data(), 1, (size_t) db->bytes, fp) == (size_t) db->bytes;
    if (! ok) why = "truncated residues";
  }
  if (ok) {
    // residues are digital codes of the alphabet (< 29), padding is P7G_PADCODE,
    // and the header's largest code is the largest residue present
    int maxcode = 0;
    int64_t pos = 0;
    for (int64_t s = 0; s < db->nseq && ok; s++) {
      const int L = db->h_len[s];
      const int64_t end = pos + (((int64_t) L + 15) & ~(int64_t) 15);
      for (int64_t b = pos; b < pos + L; b++) {
        const int c = h[(size_t) b];
        if (c >= 29) { ok = false; why = "residue code outside the digital alphabets"; break; }
        if (c > maxcode) maxcode = c;
      }
      for (int64_t b = pos + L; ok && b < end; b++)
        if (h[(size_t) b] != P7G_PADCODE) { ok = false; why = "bad padding byte"; }
      pos = end;
    }
    for (int64_t b = pos; ok && b < db->bytes; b++)
      if (h[(size_t) b] != P7G_PADCODE) { ok = false; why = "bad padding byte"; }
    if (ok && maxcode != db->maxcode) { ok = false; why = "header's largest residue code does not match the residues"; }
  }
  if (ok) {
    ok = fread(&namebytes, 8, 1, fp) == 1 && namebytes >= 0;
    if (! ok) why = "truncated name block";
  }
  if (ok && namebytes > 0) {
    db->names = (char *) malloc((size_t) namebytes);
    db->name_off = (int64_t *) malloc(8 * (size_t) db->nseq);
    ok = db->names && db->name_off && fread(db->names, 1, (size_t) namebytes, fp) == (size_t) namebytes &&
         fread(db->name_off, 8, (size_t) db->nseq, fp) == (size_t) db->nseq;
    for (int64_t i = 0; ok && i < db->nseq; i++) ok = db->name_off[i] >= 0 && db->name_off[i] < namebytes;
    ok = ok && db->names[namebytes - 1] == '\0';
    if (! ok) why = "bad name block";
  }
  if (ok) {
    ok = fread(&magic, 4, 1, fp) == 1 && magic == P7G_DBMAGIC;
    if (! ok) why = "bad trailer";
  }
  fclose(fp);
  if (! ok) { p7g_set_error("%s is not a valid database image: %s", path, why); p7_gpu_seqdb_Destroy(db); return p7G_EFORMAT; }
  if ((status = ensure_init()) != p7G_OK) { p7_gpu_seqdb_Destroy(db); return status; }
  std::vector<int32_t> perm32((size_t) db->nseq);
  for (int64_t s = 0; s < db->nseq; s++) perm32[(size_t) s] = (int32_t) db->h_perm[s];
  hipError_t e = hipMalloc((void **) &db->d_res, (size_t) db->bytes);
  if (e == hipSuccess) e = hipMalloc((void **) &db->d_off, sizeof(int64_t) * (size_t) db->nseq);
  if (e == hipSuccess) e = hipMalloc((void **) &db->d_len, sizeof(int32_t) * (size_t) db->nseq);
  if (e == hipSuccess) e = hipMalloc((void **) &db->d_perm, sizeof(int32_t) * (size_t) db->nseq);
  if (e == hipSuccess) e = hipMemcpyAsync(db->d_res, h.data(), (size_t) db->bytes, hipMemcpyHostToDevice, g_stream);
  if (e == hipSuccess) e = hipMemcpyAsync(db->d_off, h_off.data(), sizeof(int64_t) * (size_t) db->nseq, hipMemcpyHostToDevice, g_stream);
  if (e == hipSuccess) e = hipMemcpyAsync(db->d_len, db->h_len, sizeof(int32_t) * (size_t) db->nseq, hipMemcpyHostToDevice, g_stream);
  if (e == hipSuccess) e = hipMemcpyAsync(db->d_perm, perm32.data(), sizeof(int32_t) * (size_t) db->nseq, hipMemcpyHostToDevice, g_stream);
  if (e == hipSuccess) e = hipStreamSynchronize(g_stream);
  if (e == hipSuccess) e = seqdb_segments(db);
  if (e != hipSuccess) { p7g_set_error("database upload failed: %s", hipGetErrorString(e)); p7_gpu_seqdb_Destroy(db); return p7G_EDEVICE; }
  *ret_db = db;
  return p7G_OK;
}

extern "C" int64_t p7_gpu_seqdb_NSeq(const P7_GPU_SEQDB *db) { return db ? db->nseq : 0; }

static int check_query(const P7_GPU_OPROFILE *om, const P7_GPU_SEQDB *db);

extern "C" int
p7_gpu_oprofile_DeviceBytes(const P7_GPU_OPROFILE *om, const P7_GPU_SEQDB *db, size_t *ret_bytes)
{
  int status;
  if ((status = ensure_init()) != p7G_OK) return status;
  if ((status = check_query(om, db)) != p7G_OK) return status;
  DevProfile *p;
  if ((status = dev_profile(om, db->Lmax, &p)) != p7G_OK) return status;
  const p7g_dev_profile_s *d = (const p7g_dev_profile_s *) ((const char *) p - offsetof(p7g_dev_profile_s, p));
  *ret_bytes = d->bytes;
  return p7G_OK;
}
extern "C" int64_t p7_gpu_seqdb_NRes(const P7_GPU_SEQDB *db) { return db ? db->nres : 0; }

static DevSeqs seqs_of(const P7_GPU_SEQDB *db)
{
  DevSeqs s; s.res = db->d_res; s.off = db->d_off; s.len = db->d_len; s.perm = db->d_perm; s.nslot = db->nseq;
  s.seg = db->d_seg; s.nseg = db->nseg; s.nseg_slots = db->nseg_slots;
  return s;
}

/*****************************************************************
 * Tuned shape of the ungapped MSV pass
 *****************************************************************/
#define SSV_GRAB_DEFAULT 4          // target groups per work-queue atomic
#define TUNE_MARGIN 0.97f           // a tuned shape must beat the built-in one by 3%
// A model's length class, M rounded up to a multiple of 32, fixes the 16-,
// 32- and 64-lane shapes of the ungapped pass (and the 8-lane one to within
// one dword per lane), so the models of one class share one measurement.
static int tune_class(int M) { return cdiv(M, 32); }
struct TuneEntry { int v, grab; float us[P7G_SSV_NVAR]; };
static std::mutex g_tunelock;
static std::map<int, TuneEntry> g_tune;

// The variant and queue grab the ungapped pass of this profile runs with
// (v = -1: the chunked pass).
static void ssv_choice(const DevProfile &p, int *v, int *grab)
{
  *v = p.ssv_def; *grab = SSV_GRAB_DEFAULT;
  if (p.ssv_nchunk) return;
  std::lock_guard<std::mutex> lk(g_tunelock);
  auto it = g_tune.find(tune_class(p.M));
  if (it != g_tune.end() && it->second.v >= 0 && it->second.v < P7G_SSV_NVAR && p.ssv[it->second.v].G) {
    *v = it->second.v; *grab = it->second.grab;
  }
}

// The ungapped pass in the chosen shape; a chunked model takes one queue
// cursor per chunk from queue[] and the boundary columns of a work area
// (work_bnd, bnd_cap entries each).
static hipError_t launch_ssv_chosen(const DevProfile &p, const DevSeqs &s, const int32_t *slots, int64_t n,
                                    uint16_t *zmax, int32_t *queue, uint16_t *bnd, int64_t bnd_cap, hipStream_t st)
{
  int v, grab;
  ssv_choice(p, &v, &grab);
  if (v < 0) return launch_ssv_chunked(p, s, slots, n, zmax, queue, grab, bnd, bnd_cap, st);
  return launch_ssv(p, v, s, slots, n, zmax, queue, grab, st);
}

/*****************************************************************
 * Work buffers: per-slot results and lists, grown on demand
 *****************************************************************/

// bytes per database slot of all per-slot arrays of a work area
#define WORK_BYTES_PER_SLOT (2 + 2 + 4*6 + 1 + 4 + 4 + 4*5 + 4 + 8 + 4 + 1)
#define WORK_BYTES_FIXED    (64 * 1024)     // alignment of the arrays, counters, compaction tiles
// Forward scale-factor pool of a pipeline: one float per row of every target
// that reaches Forward. F3 is only known after Forward, so every target that
// reaches it takes L+1 floats. A pool starts at min(rows of every target,
// 2^28) floats; a run that finds it full (loose thresholds, --max) grows it
// to the rows of every target (pool_rows_all) and runs again.
static int64_t pool_rows_all(int64_t nseq, int64_t nres) { return nres + nseq; }
static int64_t pipeline_pool_floats(int64_t nseq, int64_t nres) { return std::min<int64_t>(pool_rows_all(nseq, nres), (int64_t) 1 << 28); }

static int
work_reserve(Work &w, int64_t n, int64_t pool_floats)
{
  if (! w.ev_ready) { for (int i = 0; i < 10; i++) HIPCHK(hipEventCreate(&w.ev[i])); w.ev_ready = true; }
  if (n <= w.cap && pool_floats <= w.pool_size) return p7G_OK;
  if (w.mem || w.pool) HIPCHK(hipStreamSynchronize(w.st));
  if (w.mem) { (void) hipFree(w.mem); w.mem = nullptr; w.cap = 0; }
  if (w.pool) { (void) hipFree(w.pool); w.pool = nullptr; w.pool_size = 0; }
  const int64_t cap = std::max<int64_t>(n, 1024);
  size_t total = (size_t) cap * WORK_BYTES_PER_SLOT + WORK_BYTES_FIXED;
  char *m = nullptr;
  HIPCHK(hipMalloc((void **) &m, total));
  size_t o = 0;
  auto take = [&](size_t bytes) { char *p = m + o; o += (bytes + 255) & ~(size_t) 255; return (void *) p; };
  w.r.zmax = (uint16_t *) take(2 * cap); w.r.xj = (int16_t *) take(2 * cap);
  w.r.nullsc = (float *) take(4 * cap); w.r.usc = (float *) take(4 * cap); w.r.filtersc = (float *) take(4 * cap);
  w.r.vfsc = (float *) take(4 * cap); w.r.fwdsc = (float *) take(4 * cap); w.r.bcksc = (float *) take(4 * cap);
  w.r.stage = (uint8_t *) take(cap); w.r.vstatus = (int *) take(4 * cap); w.r.fstatus = (int *) take(4 * cap);
  w.l.full = (int32_t *) take(4 * cap); w.l.bias = (int32_t *) take(4 * cap); w.l.vit = (int32_t *) take(4 * cap);
  w.l.fwd = (int32_t *) take(4 * cap); w.l.bck = (int32_t *) take(4 * cap);
  w.slots = (int32_t *) take(4 * cap);
  w.scale_off = (int64_t *) take(8 * cap);
  w.outbuf = take(4 * cap);
  w.tilecnt = (int32_t *) take(4 * (size_t) (cap / 4096 + 1));
  w.l.counts = (int32_t *) take(64);
  w.l.cursor = (unsigned long long *) take(64);
  w.stats = (unsigned long long *) take(64);
  if (o > total) { p7g_set_error("internal: work buffer overrun"); return p7G_EINVAL; }
  w.mem = m; w.cap = cap;
  const int64_t pf = std::max<int64_t>(pool_floats, 1 << 20);
  HIPCHK(hipMalloc((void **) &w.pool, sizeof(float) * (size_t) pf));
  w.pool_size = pf;
  return p7G_OK;
}

// Lists, queue cursors and the pool cursor start from zero (fill kernels, so
// that a captured graph holds no memset node).
static int
reset_counts(Work &w)
{
  HIPCHK(launch_fill32(w.l.counts, 0u, 16, w.st));
  HIPCHK(launch_fill32(w.l.cursor, 0u, 16, w.st));
  return p7G_OK;
}

// The boundary columns of the chunked ungapped pass over a residue array of
// `bytes` bytes (two columns of one uint16 per byte), made when a chunked
// model first runs.
static int
work_bnd(Work &w, const DevProfile &p, int64_t bytes)
{
  if (! p.ssv_nchunk || w.bnd_cap >= bytes) return p7G_OK;
  if (w.bnd) { HIPCHK(hipStreamSynchronize(w.st)); (void) hipFree(w.bnd); w.bnd = nullptr; w.bnd_cap = 0; }
  HIPCHK(hipMalloc((void **) &w.bnd, 2 * 2 * (size_t) bytes));
  w.bnd_cap = bytes;
  return p7G_OK;
}

// Domain decoding buffers: per slot two pool offsets and three results, the
// region records with their envelope-scoring arrays and residue copies (at
// most 2^26 bytes), and a row pool of 15 floats per row of an F3 survivor (6
// per row of a target that only reaches Forward), at most 2^28 floats.
#define DOM_REG_CAP ((int64_t) 1 << 20)
// null2 of single-domain regions: regions go in batches whose areas fit a
// pool of at most a quarter of the free device memory, and at least
// N2_POOL_MIN floats (2 GiB); each batch takes three list launches, each with
// its own count/queue block of 16 ints
#define N2_POOL_MIN ((int64_t) 1 << 29)
#define N2_MAX_BATCH 256
#define N2_CNT_INTS (N2_MAX_BATCH * 3 * 16)
#define DOM_RES_CAP ((int64_t) 1 << 26)
static int64_t dom_pool_floats(int64_t nseq, int64_t nres) { return std::min<int64_t>(16 * (nres + nseq), (int64_t) 1 << 28); }

static int
dom_reserve(Work &w, int64_t n, int64_t pool_floats)
{
  pool_floats = std::max<int64_t>(pool_floats, 1 << 20);
  if (n <= w.dom_cap && pool_floats <= w.dom.size) return p7G_OK;
  if (w.dom_mem || w.dom.pool) HIPCHK(hipStreamSynchronize(w.st));
  if (w.dom_mem) { (void) hipFree(w.dom_mem); w.dom_mem = nullptr; w.dom_cap = 0; }
  if (w.dom.pool) { (void) hipFree(w.dom.pool); w.dom.pool = nullptr; w.dom.size = 0; }
  const int64_t cap = std::max<int64_t>(n, 1024);
  const size_t total = (size_t) cap * (8 + 8 + 4 + 4 + 4) + (16 + 8 + 8 + 8 + 4 * 12) * (size_t) DOM_REG_CAP
                       + (size_t) DOM_RES_CAP + 4 * (size_t) N2_CNT_INTS
                       + 4 * (size_t) (DOM_RES_CAP + DOM_REG_CAP) + 8 * (size_t) DOM_REG_CAP
                       + (4 + 16) * (size_t) DOM_REG_CAP + 30 * 256;
  char *m = nullptr;
  HIPCHK(hipMalloc((void **) &m, total));
  size_t o = 0;
  auto take = [&](size_t bytes) { char *p = m + o; o += (bytes + 255) & ~(size_t) 255; return (void *) p; };
  w.dom.doff = (int64_t *) take(8 * cap); w.dom.boff = (int64_t *) take(8 * cap);
  w.dout.nexp = (float *) take(4 * cap); w.dout.nreg = (int32_t *) take(4 * cap); w.dout.nclu = (int32_t *) take(4 * cap);
  w.dout.regs = (int4 *) take(16 * (size_t) DOM_REG_CAP); w.dout.reg_cap = DOM_REG_CAP;
  w.dout.rseq = (int32_t *) take(4 * (size_t) DOM_REG_CAP); w.dout.rlen = (int32_t *) take(4 * (size_t) DOM_REG_CAP);
  w.dout.nclu_total = (int32_t *) take(4);
  w.dout.ecount = (int32_t *) take(64);
  w.dout.v_off = (int64_t *) take(8 * (size_t) DOM_REG_CAP);
  w.dout.v_len = (int32_t *) take(4 * (size_t) DOM_REG_CAP); w.dout.v_list = (int32_t *) take(4 * (size_t) DOM_REG_CAP);
  w.dout.env_st = (int32_t *) take(4 * (size_t) DOM_REG_CAP);
  w.dout.env_loop = (float *) take(4 * (size_t) DOM_REG_CAP); w.dout.env_move = (float *) take(4 * (size_t) DOM_REG_CAP);
  w.dout.envsc = (float *) take(4 * (size_t) DOM_REG_CAP);
  w.dout.res = (uint8_t *) take((size_t) DOM_RES_CAP); w.dout.res_cap = DOM_RES_CAP;
  w.n2off = (int64_t *) take(8 * (size_t) DOM_REG_CAP);
  w.n2dc = (float *) take(4 * (size_t) DOM_REG_CAP); w.n2st = (int32_t *) take(4 * (size_t) DOM_REG_CAP);
  w.n2list = (int32_t *) take(4 * (size_t) DOM_REG_CAP);
  w.n2cnt = (int32_t *) take(4 * (size_t) N2_CNT_INTS);
  w.dout.rscale = (float *) take(4 * (size_t) (DOM_RES_CAP + DOM_REG_CAP)); w.dout.rscale_cap = DOM_RES_CAP + DOM_REG_CAP;
  w.dout.rsoff = (int64_t *) take(8 * (size_t) DOM_REG_CAP);
  w.n2oa = (float *) take(4 * (size_t) DOM_REG_CAP);
  w.n2ali = (int4 *) take(16 * (size_t) DOM_REG_CAP);
  w.n2poff = (int64_t *) take(8 * (size_t) DOM_REG_CAP); w.n2pn = (int32_t *) take(4 * (size_t) DOM_REG_CAP);
  if (o > total) { (void) hipFree(m); p7g_set_error("internal: domain buffer overrun"); return p7G_EINVAL; }
  w.dom_mem = m; w.dom_cap = cap;
  HIPCHK(hipMalloc((void **) &w.dom.pool, sizeof(float) * (size_t) pool_floats));
  w.dom.size = pool_floats;
  return p7G_OK;
}

static void
work_release(Work &w)
{
  if (w.st) (void) hipStreamSynchronize(w.st);
  if (w.dom_mem) (void) hipFree(w.dom_mem);
  if (w.dom.pool) (void) hipFree(w.dom.pool);
  if (w.n2pool) (void) hipFree(w.n2pool);
  if (w.n2path) (void) hipFree(w.n2path);
  w.dom_mem = nullptr; w.dom = DomBufs(); w.dout = DomOut(); w.dom_cap = 0;
  w.n2off = nullptr; w.n2dc = nullptr; w.n2st = nullptr; w.n2list = nullptr; w.n2cnt = nullptr;
  w.n2oa = nullptr; w.n2ali = nullptr; w.n2poff = nullptr; w.n2pn = nullptr; w.n2path = nullptr; w.n2path_cap = 0;
  w.n2pool = nullptr; w.n2size = 0;
  if (w.st2) (void) hipStreamSynchronize(w.st2);
  if (w.mdjobs) (void) hipFree(w.mdjobs);
  w.mdjobs = nullptr; w.mdout = nullptr; w.mdsrc = w.mddst = nullptr; w.mdlen = w.mdcnt = nullptr; w.md_cap = 0;
  if (w.mdpool) (void) hipFree(w.mdpool);
  w.mdpool = nullptr; w.mdsize = 0;
  if (w.mdn2) (void) hipFree(w.mdn2);
  w.mdn2 = nullptr; w.mdn2_cap = 0;
  if (w.mem) (void) hipFree(w.mem);
  if (w.pool) (void) hipFree(w.pool);
  if (w.bnd) (void) hipFree(w.bnd);
  w.bnd = nullptr; w.bnd_cap = 0;
  if (w.ev_ready) for (int i = 0; i < 10; i++) (void) hipEventDestroy(w.ev[i]);
  w.mem = nullptr; w.pool = nullptr; w.cap = 0; w.pool_size = 0; w.ev_ready = false;
}

static int
slots_from_idx(const P7_GPU_SEQDB *db, const int64_t *idx, int64_t n, std::vector<int32_t> &slots)
{
  slots.resize((size_t) n);
  for (int64_t t = 0; t < n; t++) {
    const int64_t i = idx ? idx[t] : t;
    if (i < 0 || i >= db->nseq) { p7g_set_error("target index %lld out of range", (long long) i); return p7G_EINVAL; }
    slots[t] = db->h_inv[i];
  }
  return p7G_OK;
}

static LenCfg lencfg_per_target(void) { LenCfg lc; memset(&lc, 0, sizeof(lc)); return lc; }

static float
msv_score_from_xj(const P7_GPU_OPROFILE *om, int xj, int tjb)
{
  float sc = ((float) (xj - tjb) - (float) om->base_b);
  sc /= om->scale_b;
  sc -= 3.0;
  return sc;
}

/*****************************************************************
 * Batched stages (no gating): every listed target, length-configured to
 * its own L.
 *****************************************************************/
extern "C" int
p7_gpu_MSVFilterBatch(const P7_GPU_OPROFILE *om, const P7_GPU_SEQDB *db, const int64_t *idx, int64_t n,
                      float *ret_sc, int *ret_status)
{
  std::lock_guard<std::mutex> lk(g_lock);
  int status;
  DevProfile *p;
  if ((status = ensure_init()) != p7G_OK) return status;
  if (db->maxcode >= om->abc->Kp) { p7g_set_error("database holds residue codes outside the model alphabet"); return p7G_EINCOMPAT; }
  if ((status = dev_profile(om, db->Lmax, &p)) != p7G_OK) return status;
  if (idx == nullptr) n = db->nseq;
  if ((status = work_reserve(g_work, db->nseq, 0)) != p7G_OK) return status;
  std::vector<int32_t> slots;
  if ((status = slots_from_idx(db, idx, n, slots)) != p7G_OK) return status;
  DevSeqs s = seqs_of(db);
  Work &w = g_work;
  LenCfg lc = lencfg_per_target();
  if ((status = reset_counts(g_work)) != p7G_OK) return status;
  HIPCHK(hipMemcpyAsync(w.slots, slots.data(), sizeof(int32_t) * (size_t) n, hipMemcpyHostToDevice, g_stream));
  if ((status = work_bnd(w, *p, db->bytes)) != p7G_OK) return status;
  HIPCHK(launch_ssv_chosen(*p, s, w.slots, n, w.r.zmax, &w.l.counts[P7G_Q_SSV], w.bnd, w.bnd_cap, g_stream));
  HIPCHK(launch_msv_classify(*p, s, w.slots, n, w.r.zmax, w.r.xj, w.l.full, &w.l.counts[0], lc, g_stream));
  HIPCHK(launch_msv_full(*p, s, w.l.full, &w.l.counts[0], n, w.r.xj, lc, g_stream));
  std::vector<int16_t> xj((size_t) db->nseq);
  HIPCHK(hipMemcpyAsync(xj.data(), w.r.xj, sizeof(int16_t) * (size_t) db->nseq, hipMemcpyDeviceToHost, g_stream));
  HIPCHK(hipStreamSynchronize(g_stream));
  P7G_LENCFG lcf;
  if ((status = p7g_lencfg_build(om, db->Lmax, &lcf)) != p7G_OK) return status;
  for (int64_t t = 0; t < n; t++) {
    const int slot = slots[t];
    const int v = xj[slot];
    if (v < 0) { ret_sc[t] = INFINITY; if (ret_status) ret_status[t] = p7G_ERANGE; }
    else { ret_sc[t] = msv_score_from_xj(om, v, lcf.tjb_b[db->h_len[slot]]); if (ret_status) ret_status[t] = p7G_OK; }
  }
  p7g_lencfg_free(&lcf);
  return p7G_OK;
}

// p7_gpu_MSVFilterBatch without allocating (see p7_gpu.h): the same kernels
// over a slot list, with every buffer in this device's part of `scratch`.
// The profile's device copy must already cover the database's lengths.
extern "C" int
p7_gpu_MSVFilterBatchNoAlloc(const P7_GPU_OPROFILE *om, const P7_GPU_SEQDB *db, const int64_t *idx, int64_t n,
                             float *ret_sc, int *ret_status)
{
  std::lock_guard<std::mutex> lk(g_lock);
  if (! g_stream || g_device < 0 || g_device >= P7G_MAXDEV || ! g_scratch[g_device]) {
    p7g_set_error("p7_gpu_MSVFilterBatchNoAlloc: call p7_gpu_Init first");
    return p7G_EINVAL;
  }
  if (! om || ! db || ! ret_sc) { p7g_set_error("p7_gpu_MSVFilterBatchNoAlloc: NULL argument"); return p7G_EINVAL; }
  if (db->maxcode >= om->abc->Kp) { p7g_set_error("database holds residue codes outside the model alphabet"); return p7G_EINCOMPAT; }
  if (idx == nullptr) n = db->nseq;
  if (n < 0 || n > P7G_NOALLOC_MAXSEQ || db->nseq > P7G_NOALLOC_MAXSEQ) {
    p7g_set_error("p7_gpu_MSVFilterBatchNoAlloc: %lld targets of a %lld-sequence database exceed its scratch "
                  "(at most %lld of each)", (long long) n, (long long) db->nseq, (long long) P7G_NOALLOC_MAXSEQ);
    return p7G_EINVAL;
  }
  const DevProfile *p = nullptr;
  {
    std::lock_guard<std::mutex> dl(g_devlock);
    if (om->dev && om->dev->Lmax >= db->Lmax) p = &om->dev->p;
  }
  if (! p) {
    p7g_set_error("p7_gpu_MSVFilterBatchNoAlloc: the profile has no device copy for this database's lengths "
                  "(any other batched call or p7_gpu_Pipeline with this profile and database makes it)");
    return p7G_EINVAL;
  }
  if (p->ssv_nchunk) {
    p7g_set_error("p7_gpu_MSVFilterBatchNoAlloc: models longer than one LDS table (M = %d) need boundary columns "
                  "sized by the database; use p7_gpu_MSVFilterBatch", p->M);
    return p7G_EINVAL;
  }
  if (n == 0) return p7G_OK;
  for (int64_t t = 0; t < n; t++) {
    const int64_t i = idx ? idx[t] : t;
    if (i < 0 || i >= db->nseq) { p7g_set_error("target index %lld out of range", (long long) i); return p7G_EINVAL; }
    g_na_slots[t] = db->h_inv[i];
  }
  char *m = (char *) g_scratch[g_device];
  uint16_t *zmax = (uint16_t *) m;                m += NOALLOC_PART(2 * P7G_NOALLOC_MAXSEQ);
  int16_t  *xj   = (int16_t *) m;                 m += NOALLOC_PART(2 * P7G_NOALLOC_MAXSEQ);
  int32_t  *slots = (int32_t *) m;                m += NOALLOC_PART(4 * P7G_NOALLOC_MAXSEQ);
  int32_t  *full  = (int32_t *) m;                m += NOALLOC_PART(4 * P7G_NOALLOC_MAXSEQ);
  float    *sc    = (float *) m;                  m += NOALLOC_PART(4 * P7G_NOALLOC_MAXSEQ);
  int      *st    = (int *) m;                    m += NOALLOC_PART(4 * P7G_NOALLOC_MAXSEQ);
  int32_t  *counts = (int32_t *) m;
  DevSeqs s = seqs_of(db);
  LenCfg lc = lencfg_per_target();
  // list counts and queue cursors start at zero (the last call left its own)
  HIPCHK(launch_fill32(counts, 0u, 16, g_stream));
  HIPCHK(hipMemcpyAsync(slots, g_na_slots, sizeof(int32_t) * (size_t) n, hipMemcpyHostToDevice, g_stream));
  HIPCHK(launch_ssv_chosen(*p, s, slots, n, zmax, &counts[P7G_Q_SSV], nullptr, 0, g_stream));
  HIPCHK(launch_msv_classify(*p, s, slots, n, zmax, xj, full, &counts[0], lc, g_stream));
  HIPCHK(launch_msv_full(*p, s, full, &counts[0], n, xj, lc, g_stream));
  HIPCHK(launch_msv_scores(*p, s, slots, n, xj, sc, st, g_stream));
  HIPCHK(hipMemcpyAsync(ret_sc, sc, sizeof(float) * (size_t) n, hipMemcpyDeviceToHost, g_stream));
  if (ret_status) HIPCHK(hipMemcpyAsync(ret_status, st, sizeof(int) * (size_t) n, hipMemcpyDeviceToHost, g_stream));
  HIPCHK(hipStreamSynchronize(g_stream));
  return p7G_OK;
}

// Shared driver of the list-based stages for a batch of explicit targets.
enum { ST_BIAS = 1, ST_VIT = 2, ST_FWD = 3, ST_BCK = 4 };

static int
run_list_stage(int which, const P7_GPU_OPROFILE *om, const P7_GPU_SEQDB *db, const int64_t *idx, int64_t n,
               float *ret_sc, int *ret_status)
{
  std::lock_guard<std::mutex> lk(g_lock);
  int status;
  DevProfile *p;
  if ((status = ensure_init()) != p7G_OK) return status;
  if (db->maxcode >= om->abc->Kp) { p7g_set_error("database holds residue codes outside the model alphabet"); return p7G_EINCOMPAT; }
  if ((status = dev_profile(om, db->Lmax, &p)) != p7G_OK) return status;
  if (idx == nullptr) n = db->nseq;
  std::vector<int32_t> slots;
  if ((status = slots_from_idx(db, idx, n, slots)) != p7G_OK) return status;
  int64_t poolneed = 0;
  if (which == ST_BCK) for (int64_t t = 0; t < n; t++) poolneed += db->h_len[slots[t]] + 1;
  if ((status = work_reserve(g_work, db->nseq, poolneed)) != p7G_OK) return status;
  for (int64_t t = 0; t < n; t++) if (db->h_len[slots[t]] == 0) { p7g_set_error("zero-length target in a batched stage"); return p7G_EINVAL; }
  DevSeqs s = seqs_of(db);
  Work &w = g_work;
  LenCfg lc = lencfg_per_target();
  GateThresh g; memset(&g, 0, sizeof(g));
  if ((status = reset_counts(g_work)) != p7G_OK) return status;
  int32_t *list = (which == ST_BIAS) ? w.l.bias : (which == ST_VIT) ? w.l.vit : w.l.fwd;
  int32_t *cnt  = (which == ST_BIAS) ? &w.l.counts[1] : (which == ST_VIT) ? &w.l.counts[2] : &w.l.counts[3];
  int32_t n32 = (int32_t) n;
  HIPCHK(hipMemcpyAsync(list, slots.data(), sizeof(int32_t) * (size_t) n, hipMemcpyHostToDevice, g_stream));
  HIPCHK(hipMemcpyAsync(cnt, &n32, sizeof(int32_t), hipMemcpyHostToDevice, g_stream));
  if (which == ST_BIAS) {
    HIPCHK(launch_bias(*p, s, list, cnt, n, w.r, w.l, g, 0, lc, w.pool, w.pool_size, w.scale_off, g_stream));
  } else if (which == ST_VIT) {
    HIPCHK(launch_vit(*p, s, list, cnt, n, w.r, w.l, g, 0, lc, g_stream));
  } else {
    HIPCHK(launch_fwd(*p, s, list, cnt, n, w.r, w.l, g, 0, lc,
                      which == ST_BCK ? w.pool : nullptr, w.pool_size, w.scale_off, nullptr, g_stream));
    if (which == ST_BCK) {
      // its own count, hence its own work-queue cursor (Forward used up the one of cnt)
      HIPCHK(hipMemcpyAsync(&w.l.counts[4], &n32, sizeof(int32_t), hipMemcpyHostToDevice, g_stream));
      HIPCHK(launch_bck(*p, s, list, &w.l.counts[4], n, w.r, w.pool, w.scale_off, nullptr, nullptr, lc, g_stream));
    }
  }
  std::vector<float> sc((size_t) db->nseq);
  std::vector<int>   st((size_t) db->nseq, 0);
  float *src = (which == ST_BIAS) ? w.r.filtersc : (which == ST_VIT) ? w.r.vfsc : (which == ST_FWD) ? w.r.fwdsc : w.r.bcksc;
  HIPCHK(hipMemcpyAsync(sc.data(), src, sizeof(float) * (size_t) db->nseq, hipMemcpyDeviceToHost, g_stream));
  if (which == ST_VIT) HIPCHK(hipMemcpyAsync(st.data(), w.r.vstatus, sizeof(int) * (size_t) db->nseq, hipMemcpyDeviceToHost, g_stream));
  if (which >= ST_FWD) HIPCHK(hipMemcpyAsync(st.data(), w.r.fstatus, sizeof(int) * (size_t) db->nseq, hipMemcpyDeviceToHost, g_stream));
  int32_t counts[16];
  HIPCHK(hipMemcpyAsync(counts, w.l.counts, sizeof(counts), hipMemcpyDeviceToHost, g_stream));
  HIPCHK(hipStreamSynchronize(g_stream));
  if (which == ST_BCK && counts[6]) { p7g_set_error("Forward scale-factor pool exhausted"); return p7G_EMEM; }
  for (int64_t t = 0; t < n; t++) {
    ret_sc[t] = sc[slots[t]];
    if (ret_status) ret_status[t] = (st[slots[t]] & 0xff) == 16 ? p7G_ERANGE : ((st[slots[t]] & 0xff) ? p7G_FAIL : p7G_OK);
  }
  return p7G_OK;
}

extern "C" int p7_gpu_ViterbiFilterBatch(const P7_GPU_OPROFILE *om, const P7_GPU_SEQDB *db, const int64_t *idx, int64_t n, float *ret_sc, int *ret_status)
{ return run_list_stage(ST_VIT, om, db, idx, n, ret_sc, ret_status); }
extern "C" int p7_gpu_ForwardParserBatch(const P7_GPU_OPROFILE *om, const P7_GPU_SEQDB *db, const int64_t *idx, int64_t n, float *ret_sc, int *ret_status)
{ return run_list_stage(ST_FWD, om, db, idx, n, ret_sc, ret_status); }
extern "C" int p7_gpu_BackwardParserBatch(const P7_GPU_OPROFILE *om, const P7_GPU_SEQDB *db, const int64_t *idx, int64_t n, float *ret_sc, int *ret_status)
{ return run_list_stage(ST_BCK, om, db, idx, n, ret_sc, ret_status); }
extern "C" int p7_gpu_FilterScoreBatch(const P7_GPU_OPROFILE *om, const P7_GPU_SEQDB *db, const int64_t *idx, int64_t n, float *ret_sc)
{ return run_list_stage(ST_BIAS, om, db, idx, n, ret_sc, nullptr); }

/*****************************************************************
 * Pipeline (p7_pipeline.c:666-735) on the device
 *****************************************************************/
struct Batch;
static void batch_free(Batch *b);

struct p7g_pipeline_s {
  double F1, F2, F3;
  int    do_biasfilter, do_max;
  int    do_domains = 0;      // posterior decoding and regions of the F3 survivors (p7_gpu_Pipeline)
  int    do_null2 = 1;        // and the null2 correction of single-domain regions (p7_pipeline.c:208, do_null2)
  // the last decoding run's region records: on the device until asked for,
  // then in the caller's sequence order
  mutable std::mutex regs_mu;
  mutable std::vector<int4> regs;
  mutable std::vector<float> envsc, n2dc, oasc;
  mutable std::vector<int4> ali;
  mutable std::vector<int32_t> n2st, regs_slot, regs_L;
  mutable std::vector<int64_t> regs_didx;      // each sorted record's index on the device
  // hit-list scores of the last run, computed on first request
  mutable bool hits_ready = false;
  mutable std::vector<int64_t> h_tseq;
  mutable std::vector<P7G_TSCORE> h_ts;
  mutable std::vector<int32_t> h_tnd, h_tst;
  mutable std::vector<P7G_DSCORE> h_ds;
  mutable std::vector<float> h_nexp;           // per target: P7_DOMAINDEF's nexpected,
  mutable std::vector<int4> h_tcnt;            // nregions, nclustered, noverlaps, nenvelopes
  mutable bool regs_ready = false;
  bool   regs_valid = false;
  int64_t nregs = 0;
  const int4 *regs_dev = nullptr;
  const float *envsc_dev = nullptr;
  const float *n2dc_dev = nullptr;
  const int32_t *n2st_dev = nullptr;
  const float *oa_dev = nullptr;
  const int4 *ali_dev = nullptr;
  // for the hit-list scores: the run's Forward scores (by slot), target
  // lengths (by slot) and the profile's Forward tail parameters
  const float *fwdsc_dev = nullptr;
  const float *nexp_dev = nullptr;             // nexpected by slot
  const int32_t *rseq_dev = nullptr, *rlen_dev = nullptr;   // per record: caller index, target length
  const uint32_t *path_dev = nullptr;          // the alignments' states (k_oa_trace)
  const int64_t *poff_dev = nullptr;
  const int32_t *pn_dev = nullptr;
  // what the alignment display needs of the last run's profile, copied so
  // that the results outlive it: alphabet, length, annotation, match odds
  const P7G_ALPHABET *d_abc = nullptr;
  int         d_M = 0;
  std::string d_cons, d_cs, d_rf;
  bool        d_has_cs = false, d_has_rf = false;
  std::vector<float> d_rfv;                    // [Kp][M]
  int64_t nslots = 0;
  float ftau = 0.0f, flambda = 0.0f;
  P7_GPU_STATS st;
  Work       w;             // this pipeline's stream and device work area
  Batch     *b = nullptr;   // work areas and graph of p7_gpu_PipelineBatch, made on first use
  std::mutex mu;            // one call at a time per pipeline object
};

extern "C" P7_GPU_PIPELINE *
p7_gpu_pipeline_Create(double F1, double F2, double F3, int do_biasfilter, int do_max)
{
  P7_GPU_PIPELINE *pli = new (std::nothrow) p7g_pipeline_s();
  if (! pli) return nullptr;
  memset(&pli->st, 0, sizeof(pli->st));
  pli->F1 = F1; pli->F2 = F2; pli->F3 = F3;
  pli->do_biasfilter = do_biasfilter; pli->do_max = do_max;
  if (do_max) { pli->F1 = pli->F2 = pli->F3 = 1.0; pli->do_biasfilter = 0; }   // --max (p7_pipeline.c:214-219)
  return pli;
}

extern "C" size_t
p7g_pipeline_work_bytes(int64_t nseq, int64_t nres)
{
  const int64_t cap = std::max<int64_t>(nseq, 1024);
  return (size_t) cap * WORK_BYTES_PER_SLOT + WORK_BYTES_FIXED + sizeof(float) * (size_t) std::max<int64_t>(pipeline_pool_floats(nseq, nres), 1 << 20);
}

extern "C" int
p7g_pipeline_Reserve(P7_GPU_PIPELINE *pli, const P7_GPU_SEQDB *db)
{
  std::lock_guard<std::mutex> lk(pli->mu);
  int status;
  if ((status = ensure_init()) != p7G_OK) return status;
  if (! pli->w.st) HIPCHK(hipStreamCreateWithFlags(&pli->w.st, hipStreamNonBlocking));
  return work_reserve(pli->w, db->nseq, pipeline_pool_floats(db->nseq, db->nres));
}

extern "C" void
p7g_pipeline_ResetStats(P7_GPU_PIPELINE *pli)
{
  std::lock_guard<std::mutex> lk(pli->mu);
  memset(&pli->st, 0, sizeof(pli->st));
}

extern "C" int p7g_current_device(void) { return g_device >= 0 ? g_device : 0; }

extern "C" int
p7g_device_free_bytes(size_t *ret_bytes)
{
  size_t fr = 0, tot = 0;
  *ret_bytes = 0;
  HIPCHK(hipMemGetInfo(&fr, &tot));
  *ret_bytes = fr;
  return p7G_OK;
}

extern "C" void
p7_gpu_pipeline_Destroy(P7_GPU_PIPELINE *pli)
{
  if (! pli) return;
  batch_free(pli->b);
  work_release(pli->w);
  if (pli->w.st) (void) hipStreamDestroy(pli->w.st);
  if (pli->w.st2) (void) hipStreamDestroy(pli->w.st2);
  if (pli->w.md_thr) (void) hipEventDestroy(pli->w.md_thr);
  delete pli;
}

extern "C" int
p7_gpu_pipeline_GetStats(const P7_GPU_PIPELINE *pli, P7_GPU_STATS *st)
{
  *st = pli->st;
  return p7G_OK;
}

static int
check_query(const P7_GPU_OPROFILE *om, const P7_GPU_SEQDB *db)
{
  if (db->maxcode >= om->abc->Kp) { p7g_set_error("database holds residue codes outside the model alphabet"); return p7G_EINCOMPAT; }
  for (int z = 0; z < 6; z++)
    if (om->evparam[z] == -99999.0f) { p7g_set_error("profile has no E-value parameters (STATS)"); return p7G_EINVAL; }
  return p7G_OK;
}

static GateThresh
gates_of(const P7_GPU_PIPELINE *pli, const P7_GPU_OPROFILE *om)
{
  GateThresh g;
  g.T1m = p7g_gumbel_threshold(om->evparam[p7G_MMU], om->evparam[p7G_MLAMBDA], pli->F1);
  g.T2m = p7g_gumbel_threshold(om->evparam[p7G_MMU], om->evparam[p7G_MLAMBDA], pli->F2);
  g.T2v = p7g_gumbel_threshold(om->evparam[p7G_VMU], om->evparam[p7G_VLAMBDA], pli->F2);
  g.T3f = p7g_exp_threshold(om->evparam[p7G_FTAU], om->evparam[p7G_FLAMBDA], pli->F3);
  g.do_bias = pli->do_biasfilter; g.do_max = pli->do_max;
  return g;
}

// Counters, queue cursors and the NaN marks of scores not computed, before a
// run over n slots.
static int
chain_reset(Work &w, int64_t n)
{
  int status;
  if ((status = reset_counts(w)) != p7G_OK) return status;
  HIPCHK(launch_fill32(w.r.filtersc, 0xffffffffu, n, w.st));
  HIPCHK(launch_fill32(w.r.vfsc, 0xffffffffu, n, w.st));
  HIPCHK(launch_fill32(w.r.fwdsc, 0xffffffffu, n, w.st));
  HIPCHK(launch_fill32(w.r.bcksc, 0xffffffffu, n, w.st));
  return p7G_OK;
}

// Every stage after the ungapped pass for one profile over all n slots, on the
// work area's stream: the MSV decision and full recurrence, then the bias
// filter, Viterbi, Forward and Backward over lists compacted from the per-slot
// marks in slot order (longest target first), each list kernel taking items
// from a work queue. With `timed`, events 1..5 close the MSV, bias, Viterbi,
// Forward and Backward stages.
static int
chain_stages(Work &w, const DevProfile &p, const DevSeqs &s, int64_t n, const uint16_t *zmax, const GateThresh &g,
             bool timed, bool domains = false)
{
  const DomBufs dom = domains ? w.dom : DomBufs();
  LenCfg lc = lencfg_per_target();
  HIPCHK(launch_msv_classify(p, s, nullptr, n, zmax, w.r.xj, w.l.full, &w.l.counts[0], lc, w.st));
  HIPCHK(launch_compact_i16(w.r.xj, -2, n, w.l.full, &w.l.counts[0], w.tilecnt, w.st));
  HIPCHK(launch_msv_full(p, s, w.l.full, &w.l.counts[0], n, w.r.xj, lc, w.st));
  HIPCHK(launch_msv_gate(p, s, n, w.r.xj, w.r, w.l, g, w.st));
  if (timed) HIPCHK(hipEventRecord(w.ev[1], w.st));
  if (g.do_bias) {
    HIPCHK(launch_compact_u8(w.r.stage, 1, n, w.l.bias, &w.l.counts[1], w.tilecnt, w.st));
    HIPCHK(launch_bias(p, s, w.l.bias, &w.l.counts[1], n, w.r, w.l, g, 1, lc, w.pool, w.pool_size, w.scale_off, w.st));
  }
  if (timed) HIPCHK(hipEventRecord(w.ev[2], w.st));
  HIPCHK(launch_compact_u8(w.r.stage, 2, n, w.l.vit, &w.l.counts[2], w.tilecnt, w.st));
  HIPCHK(launch_vit(p, s, w.l.vit, &w.l.counts[2], n, w.r, w.l, g, 1, lc, w.st));
  if (timed) HIPCHK(hipEventRecord(w.ev[3], w.st));
  HIPCHK(launch_compact_u8(w.r.stage, 3, n, w.l.fwd, &w.l.counts[3], w.tilecnt, w.st));
  HIPCHK(launch_fwd(p, s, w.l.fwd, &w.l.counts[3], n, w.r, w.l, g, 1, lc, w.pool, w.pool_size, w.scale_off, nullptr, w.st,
                    dom));
  if (timed) HIPCHK(hipEventRecord(w.ev[4], w.st));
  HIPCHK(launch_compact_u8(w.r.stage, 4, n, w.l.bck, &w.l.counts[4], w.tilecnt, w.st));
  HIPCHK(launch_bck(p, s, w.l.bck, &w.l.counts[4], n, w.r, w.pool, w.scale_off, nullptr, nullptr, lc, w.st, dom, w.l));
  if (timed) HIPCHK(hipEventRecord(w.ev[5], w.st));
  if (domains) {
    HIPCHK(launch_domains(p, s, w.l.bck, &w.l.counts[4], n, w.r, w.l, lc, dom, w.dout, w.st));
    // envelope scores: the Forward parser in unihit mode over the region copies
    DevSeqs vs = s;
    vs.res = w.dout.res; vs.off = w.dout.v_off; vs.len = w.dout.v_len; vs.perm = nullptr;
    vs.seg = nullptr; vs.nseg = 0; vs.nseg_slots = 0; vs.nslot = w.dout.reg_cap;
    DevResults vr = w.r;
    vr.fwdsc = w.dout.envsc; vr.fstatus = w.dout.env_st;
    DomBufs eb;
    eb.env_loop = w.dout.env_loop; eb.env_move = w.dout.env_move; eb.env_cap = w.dout.reg_cap;
    // the row scales go to their own pool (cursor 3), for the null2 Backward
    DevLists lr = w.l;
    lr.cursor = w.l.cursor + 3;
    HIPCHK(launch_fwd(p, vs, w.dout.v_list, &w.dout.ecount[0], w.dout.reg_cap, vr, lr, g, 0, lc, w.dout.rscale,
                      w.dout.rscale_cap, w.dout.rsoff, nullptr, w.st, eb));
    if (timed) HIPCHK(hipEventRecord(w.ev[7], w.st));
  }
  return p7G_OK;
}

// Pipeline counters of one profile's run from the device sums (launch_stats)
// and list counts.
static void
stats_of_run(P7_GPU_STATS &st, const unsigned long long *sums, const int32_t *counts, const P7_GPU_OPROFILE *om,
             const P7_GPU_SEQDB *db, int do_biasfilter)
{
  memset(&st, 0, sizeof(st));
  st.nmodels = 1; st.nseqs = db->nseq; st.nres = db->nres; st.nnodes = om->M;
  st.n_past_msv = (int64_t) sums[0]; st.n_past_bias = (int64_t) sums[1];
  st.n_past_vit = (int64_t) sums[2]; st.n_past_fwd = (int64_t) sums[3];
  st.n_msv_full = counts[0];
  // MSV over every target, the bias filter over MSV survivors (when on),
  // Viterbi and Forward over their lists, Backward over F3 survivors
  st.cells_msv  = (double) db->nres * om->M;
  st.cells_bias = do_biasfilter ? (double) sums[4] : 0.0;
  st.cells_vit  = (double) sums[6] * om->M;
  st.cells_fwd  = (double) sums[7] * om->M;
  st.cells_bck  = (double) sums[5] * om->M;
}

// Adds one run's counters to a pipeline's totals (p7_pipeline_Merge-like);
// cells and times are those of the last call.
static void
stats_accumulate(P7_GPU_STATS &tot, const P7_GPU_STATS &one, bool first_of_call)
{
  tot.nmodels += one.nmodels; tot.nseqs += one.nseqs; tot.nres += one.nres; tot.nnodes += one.nnodes;
  tot.n_past_msv += one.n_past_msv; tot.n_past_bias += one.n_past_bias;
  tot.n_past_vit += one.n_past_vit; tot.n_past_fwd += one.n_past_fwd;
  if (first_of_call) {
    tot.n_msv_full = 0; tot.cells_msv = tot.cells_bias = tot.cells_vit = tot.cells_fwd = tot.cells_bck = 0.0;
  }
  tot.n_msv_full += one.n_msv_full;
  tot.cells_msv += one.cells_msv; tot.cells_bias += one.cells_bias; tot.cells_vit += one.cells_vit;
  tot.cells_fwd += one.cells_fwd; tot.cells_bck += one.cells_bck;
}

// Multidomain regions (p7_domaindef.c:431-476): each region's sampled traces
// and clusters (k_md_region), in batches whose areas fit the multidomain pool.
// Every cluster becomes a record of its own after the nregs regions (w = 2):
// its residues copied from the region's copy, its envelope score computed as
// the regions' were; its correction is the sum of the region's sampled null2
// scores over the envelope (rescore_isolated_domain with null2 done,
// p7_domaindef.c:941-943), returned in env_dc, record by record. The region's
// record keeps the sum over all its residues in reg_sum (the sequence bias
// counts every residue, p7_pipeline.c:760).
//
// md_start launches the regions on the work area's second stream when one
// batch holds them all, so that they run while the null2 pass of the
// single-domain regions runs on the first (their workgroups fit on a compute
// unit beside a region's wave); md_finish waits for them (or runs every
// batch when there are several) and makes the clusters' records.
#define MD_NSAMPLES 200        // ddef->nsamples, p7_domaindef.c:112
#define MD_SEED     42u        // the pipeline's default --seed, p7_pipeline.c:102
#define MD_SEGCAP   6144       // sampled domains kept per region
struct MdRun {
  std::vector<int64_t> mdk;            // the multidomain regions' records
  std::vector<int64_t> voff;
  std::vector<float> eloop, emove;
  unsigned long long rcur = 0;
  std::vector<MdJob> jobs;
  std::vector<std::pair<size_t, size_t>> batches;
  bool launched = false;               // the one batch is on w.st2
  int max_lr = 0;                      // the longest region
};

static int
md_start(Work &w, const DevProfile &p, const std::vector<int4> &rec, const std::vector<int32_t> &vlen, int64_t nregs,
         const int32_t *h_len, MdRun &m)
{
  for (int64_t k = 0; k < nregs; k++) if (rec[(size_t) k].w == 1 && vlen[(size_t) k] > 0) m.mdk.push_back(k);
  if (m.mdk.empty()) return p7G_OK;
  for (int64_t k : m.mdk) m.max_lr = std::max(m.max_lr, (int) vlen[(size_t) k]);
  if (md_lds_bytes(p.M, MD_SEGCAP, md_acc_cap(m.max_lr)) > P7G_LDS_MAX) { p7g_set_error("model too long for multidomain regions (M=%d)", p.M); return p7G_EINVAL; }
  m.voff.resize((size_t) nregs); m.eloop.resize((size_t) nregs); m.emove.resize((size_t) nregs);
  HIPCHK(hipMemcpyAsync(m.voff.data(), w.dout.v_off, 8 * (size_t) nregs, hipMemcpyDeviceToHost, w.st));
  HIPCHK(hipMemcpyAsync(m.eloop.data(), w.dout.env_loop, 4 * (size_t) nregs, hipMemcpyDeviceToHost, w.st));
  HIPCHK(hipMemcpyAsync(m.emove.data(), w.dout.env_move, 4 * (size_t) nregs, hipMemcpyDeviceToHost, w.st));
  HIPCHK(hipMemcpyAsync(&m.rcur, w.l.cursor + 2, 8, hipMemcpyDeviceToHost, w.st));
  HIPCHK(hipStreamSynchronize(w.st));
  const int64_t nmd = (int64_t) m.mdk.size();
  if (w.md_cap < nmd) {
    if (w.mdjobs) { (void) hipFree(w.mdjobs); w.mdjobs = nullptr; w.md_cap = 0; }
    const int64_t c = std::max<int64_t>(nmd, 256), ce = c * MD_MAXC;
    char *mem = nullptr;
    HIPCHK(hipMalloc((void **) &mem, (size_t) c * (sizeof(MdJob) + sizeof(MdOut)) + (size_t) ce * 20 + 256 * 4));
    w.mdjobs = (MdJob *) mem; w.mdout = (MdOut *) (w.mdjobs + c);
    w.mdsrc = (int64_t *) (w.mdout + c); w.mddst = w.mdsrc + ce;
    w.mdlen = (int32_t *) (w.mddst + ce); w.mdcnt = w.mdlen + ce;
    w.md_cap = c;
  }
  // batches of regions whose areas fit the pool
  size_t fr = 0, tot = 0;
  HIPCHK(hipMemGetInfo(&fr, &tot));
  int64_t need_all = 0;
  for (int64_t k : m.mdk) need_all += md_area_floats(p.M, vlen[(size_t) k], MD_SEGCAP);
  int64_t cap = std::max<int64_t>(N2_POOL_MIN, std::min<int64_t>(need_all, (int64_t) (fr / 4 / sizeof(float)) + w.mdsize));
  if (w.pool_cap > 0) cap = std::min(cap, w.pool_cap);
  int64_t used = 0, maxused = 0, nres = 0, maxres = 0;
  size_t start = 0;
  for (int64_t k : m.mdk) {
    const int Lr = vlen[(size_t) k];
    const int64_t need = md_area_floats(p.M, Lr, MD_SEGCAP);
    if (need > cap) { p7g_set_error("multidomain region of %d residues too long for the pool", Lr); return p7G_EMEM; }
    if (used + need > cap || (int64_t) (m.jobs.size() - start) >= w.md_cap || nres + Lr > INT32_MAX) {
      m.batches.push_back({ start, m.jobs.size() }); start = m.jobs.size(); used = 0; nres = 0;
    }
    MdJob J;
    J.res_off = m.voff[(size_t) k]; J.area = used; J.Lr = Lr; J.Ltarget = h_len[rec[(size_t) k].x]; J.segcap = MD_SEGCAP;
    J.n2o = (int32_t) nres;
    m.jobs.push_back(J);
    used += need; maxused = std::max(maxused, used);
    nres += Lr; maxres = std::max(maxres, nres);
  }
  if (m.jobs.size() > start) m.batches.push_back({ start, m.jobs.size() });
  if (w.mdsize < maxused) {
    if (w.mdpool) { (void) hipFree(w.mdpool); w.mdpool = nullptr; w.mdsize = 0; }
    HIPCHK(hipMalloc((void **) &w.mdpool, sizeof(float) * (size_t) maxused));
    w.mdsize = maxused;
  }
  if (w.mdn2_cap < maxres) {
    if (w.mdn2) { (void) hipFree(w.mdn2); w.mdn2 = nullptr; w.mdn2_cap = 0; }
    HIPCHK(hipMalloc((void **) &w.mdn2, sizeof(float) * (size_t) maxres));
    w.mdn2_cap = maxres;
  }
  if (m.batches.size() == 1) {
    if (! w.st2) {
      HIPCHK(hipStreamCreateWithFlags(&w.st2, hipStreamNonBlocking));
      HIPCHK(hipEventCreateWithFlags(&w.md_thr, hipEventDisableTiming));
    }
    const int nb = (int) m.jobs.size();
    HIPCHK(hipMemcpyAsync(w.mdjobs, m.jobs.data(), sizeof(MdJob) * (size_t) nb, hipMemcpyHostToDevice, w.st2));
    HIPCHK(launch_md_regions(p, w.dout.res, w.mdjobs, w.mdout, nb, w.mdpool, MD_SEGCAP, MD_NSAMPLES, MD_SEED, m.max_lr, w.st2,
                             w.md_thr));
    m.launched = true;
  }
  return p7G_OK;
}

static int
md_finish(Work &w, const DevProfile &p, const DevSeqs &s, std::vector<int4> &rec, std::vector<int32_t> &vlen,
          int64_t nregs, const GateThresh &g, std::vector<float> &env_dc,
          std::vector<std::pair<int64_t, float>> &reg_sum, MdRun &m)
{
  if (m.mdk.empty()) return p7G_OK;
  hipStream_t ms = m.launched ? w.st2 : w.st;
  std::vector<float> &eloop = m.eloop, &emove = m.emove;
  std::vector<MdOut> outs(m.jobs.size());
  std::vector<float> n2;
  std::vector<int64_t> src, dst;
  std::vector<int32_t> len;
  for (size_t bi = 0; bi < m.batches.size(); bi++) {
    const auto &b = m.batches[bi];
    const int nb = (int) (b.second - b.first);
    int64_t tot_res = 0;
    for (size_t a = b.first; a < b.second; a++) tot_res += m.jobs[a].Lr;
    n2.assign((size_t) tot_res, 0.0f);
    if (! (m.launched && bi == 0)) {
      HIPCHK(hipMemcpyAsync(w.mdjobs, m.jobs.data() + b.first, sizeof(MdJob) * (size_t) nb, hipMemcpyHostToDevice, ms));
      HIPCHK(launch_md_regions(p, w.dout.res, w.mdjobs, w.mdout, nb, w.mdpool, MD_SEGCAP, MD_NSAMPLES, MD_SEED, m.max_lr, ms));
    }
    HIPCHK(launch_md_n2_gather(w.mdjobs, nb, w.mdpool, p.M, w.mdn2, ms));
    HIPCHK(hipMemcpyAsync(outs.data() + b.first, w.mdout, sizeof(MdOut) * (size_t) nb, hipMemcpyDeviceToHost, ms));
    HIPCHK(hipMemcpyAsync(n2.data(), w.mdn2, sizeof(float) * (size_t) tot_res, hipMemcpyDeviceToHost, ms));
    HIPCHK(hipStreamSynchronize(ms));
    int64_t o = 0;
    for (size_t a = b.first; a < b.second; a++) {
      const MdOut &r = outs[a];
      const int64_t k = m.mdk[a];
      const int Lr = m.jobs[a].Lr;
      if (r.status == 1) { p7g_set_error("stochastic traceback failed in a multidomain region"); return p7G_EINVAL; }
      if (r.status != 0) { p7g_set_error("multidomain region exceeds the sampled-domain capacity"); return p7G_EMEM; }
      const float *rn = n2.data() + o;
      float rs = 0.0f;
      for (int pos = 0; pos < Lr; pos++) rs += rn[pos];
      reg_sum.push_back({ k, rs });
      for (int c = 0; c < r.nc; c++) {
        const int ci = r.ci[c], cj = r.cj[c];
        if (ci < 1 || cj > Lr || ci > cj) { p7g_set_error("internal: cluster outside its region"); return p7G_EINVAL; }
        float dc = 0.0f;
        for (int pos = ci; pos <= cj; pos++) dc += rn[pos - 1];
        const int4 &pr = rec[(size_t) k];
        rec.push_back(make_int4(pr.x, pr.y + ci - 1, pr.y + cj - 1, 2));
        vlen.push_back(cj - ci + 1);
        env_dc.push_back(dc);
        src.push_back(m.voff[(size_t) k] + ci - 1);
        len.push_back(cj - ci + 1);
        eloop.push_back(eloop[(size_t) k]); emove.push_back(emove[(size_t) k]);
      }
      o += Lr;
    }
  }
  const int64_t ne = (int64_t) src.size();
  if (ne == 0) return p7G_OK;
  if (nregs + ne > w.dout.reg_cap || ne > w.md_cap * MD_MAXC) { p7g_set_error("too many domain envelopes"); return p7G_EMEM; }
  // the envelopes' residue copies after the regions'
  int64_t cur = (int64_t) m.rcur;
  for (int64_t e = 0; e < ne; e++) {
    dst.push_back(cur);
    cur += ((((int64_t) len[(size_t) e]) + 15) & ~(int64_t) 15) + 16;
  }
  if (cur > w.dout.res_cap) { p7g_set_error("domain decoding region residue pool exhausted"); return p7G_EMEM; }
  std::vector<int32_t> vl((size_t) ne);
  std::vector<int4> nrec(rec.begin() + nregs, rec.end());
  for (int64_t e = 0; e < ne; e++) vl[(size_t) e] = (int32_t) (nregs + e);
  const unsigned long long ncur = (unsigned long long) cur;
  int32_t cnt[16] = { (int32_t) ne };
  HIPCHK(hipMemcpyAsync(w.mdsrc, src.data(), 8 * (size_t) ne, hipMemcpyHostToDevice, w.st));
  HIPCHK(hipMemcpyAsync(w.mddst, dst.data(), 8 * (size_t) ne, hipMemcpyHostToDevice, w.st));
  HIPCHK(hipMemcpyAsync(w.mdlen, len.data(), 4 * (size_t) ne, hipMemcpyHostToDevice, w.st));
  HIPCHK(hipMemcpyAsync(w.mdcnt, cnt, sizeof(cnt), hipMemcpyHostToDevice, w.st));
  HIPCHK(hipMemcpyAsync(w.dout.regs + nregs, nrec.data(), sizeof(int4) * (size_t) ne, hipMemcpyHostToDevice, w.st));
  HIPCHK(hipMemcpyAsync(w.dout.v_off + nregs, dst.data(), 8 * (size_t) ne, hipMemcpyHostToDevice, w.st));
  HIPCHK(hipMemcpyAsync(w.dout.v_len + nregs, len.data(), 4 * (size_t) ne, hipMemcpyHostToDevice, w.st));
  HIPCHK(hipMemcpyAsync(w.dout.v_list + nregs, vl.data(), 4 * (size_t) ne, hipMemcpyHostToDevice, w.st));
  HIPCHK(hipMemcpyAsync(w.dout.env_loop + nregs, eloop.data() + nregs, 4 * (size_t) ne, hipMemcpyHostToDevice, w.st));
  HIPCHK(hipMemcpyAsync(w.dout.env_move + nregs, emove.data() + nregs, 4 * (size_t) ne, hipMemcpyHostToDevice, w.st));
  HIPCHK(hipMemcpyAsync(w.l.cursor + 2, &ncur, 8, hipMemcpyHostToDevice, w.st));
  HIPCHK(launch_md_copy(w.dout.res, w.mdsrc, w.mddst, w.mdlen, (int) ne, w.st));
  // their envelope scores and row scales, as the decoding step scores regions
  DevSeqs vs = s;
  vs.res = w.dout.res; vs.off = w.dout.v_off; vs.len = w.dout.v_len; vs.perm = nullptr;
  vs.seg = nullptr; vs.nseg = 0; vs.nseg_slots = 0; vs.nslot = w.dout.reg_cap;
  DevResults vr = w.r;
  vr.fwdsc = w.dout.envsc; vr.fstatus = w.dout.env_st;
  DomBufs eb;
  eb.env_loop = w.dout.env_loop; eb.env_move = w.dout.env_move; eb.env_cap = w.dout.reg_cap;
  DevLists lr = w.l;
  lr.cursor = w.l.cursor + 3;
  const LenCfg lc = lencfg_per_target();
  HIPCHK(launch_fwd(p, vs, w.dout.v_list + nregs, w.mdcnt, ne, vr, lr, g, 0, lc, w.dout.rscale,
                    w.dout.rscale_cap, w.dout.rsoff, nullptr, w.st, eb));
  int32_t ovf = 0;
  HIPCHK(hipMemcpyAsync(&ovf, w.l.counts + 6, 4, hipMemcpyDeviceToHost, w.st));
  HIPCHK(hipStreamSynchronize(w.st));               // the host vectors above are the copies' sources
  if (ovf) { p7g_set_error("envelope row-scale pool exhausted"); return p7G_EMEM; }
  return p7G_OK;
}


// The null2 pass over records k0..k1-1 that are envelopes (w != 1): the host
// gives each an area in the null2 pool and groups them in batches whose areas
// fit the pool; per batch, three list launches over the copies: the envelope
// Backward writing its special and M/I rows, the Forward again summing the
// posteriors into null2 and the correction and filling the alignment, and the
// alignment's traceback; the first batch's Forward waits for fwd_after when
// one is given. Returns after the pass has run (the host vectors are the
// copies' sources).
static int
null2_pass(Work &w, const DevProfile &p, const DevSeqs &s, const GateThresh &g, const std::vector<int4> &rec,
           const std::vector<int32_t> &vlen, int64_t k0, int64_t k1, hipEvent_t fwd_after = nullptr)
{
  // the pool: what the envelopes need, bounded by a quarter of the free memory
  // (the pool already held counts as free)
  int64_t need_all = 0;
  for (int64_t k = k0; k < k1; k++)
    if (rec[(size_t) k].w != 1) need_all += n2_area_floats(vlen[(size_t) k], p.fb_B);
  if (need_all == 0) return p7G_OK;
  size_t fr = 0, tot = 0;
  HIPCHK(hipMemGetInfo(&fr, &tot));
  int64_t cap = std::max<int64_t>(N2_POOL_MIN, std::min<int64_t>(need_all, (int64_t) (fr / 4 / sizeof(float)) + w.n2size));
  if (w.pool_cap > 0) cap = std::min(cap, w.pool_cap);
  std::vector<int64_t> off((size_t) (k1 - k0), 0);
  std::vector<int32_t> list;
  std::vector<std::pair<size_t, size_t>> batches;
  int64_t used = 0, maxused = 0;
  size_t start = 0;
  for (int64_t k = k0; k < k1; k++) {
    if (rec[(size_t) k].w == 1) continue;
    // the kernels size an envelope by its copy's length, which is its i..j
    const int Ld = vlen[(size_t) k];
    if (Ld != rec[(size_t) k].z - rec[(size_t) k].y + 1) { p7g_set_error("internal: region copy length"); return p7G_EINVAL; }
    const int64_t need = n2_area_floats(Ld, p.fb_B);
    if (need > cap) { p7g_set_error("region of %d residues too long for the null2 pool", Ld); return p7G_EMEM; }
    if (used + need > cap) { batches.push_back({ start, list.size() }); start = list.size(); used = 0; }
    off[(size_t) (k - k0)] = used; used += need; maxused = std::max(maxused, used);
    list.push_back((int32_t) k);
  }
  if (list.size() > start) batches.push_back({ start, list.size() });
  if ((int64_t) batches.size() > N2_MAX_BATCH) { p7g_set_error("too many null2 batches (%zu)", batches.size()); return p7G_EMEM; }
  if (w.n2size < maxused) {
    HIPCHK(hipStreamSynchronize(w.st));
    if (w.n2pool) { (void) hipFree(w.n2pool); w.n2pool = nullptr; w.n2size = 0; }
    HIPCHK(hipMalloc((void **) &w.n2pool, sizeof(float) * (size_t) maxused));
    w.n2size = maxused;
  }
  std::vector<int32_t> cnt(batches.size() * 3 * 16, 0);
  for (size_t b = 0; b < batches.size(); b++)
    for (int a = 0; a < 3; a++) cnt[(b * 3 + a) * 16] = (int32_t) (batches[b].second - batches[b].first);
  HIPCHK(hipMemcpyAsync(w.n2off + k0, off.data(), sizeof(int64_t) * (size_t) (k1 - k0), hipMemcpyHostToDevice, w.st));
  HIPCHK(hipMemcpyAsync(w.n2list, list.data(), sizeof(int32_t) * list.size(), hipMemcpyHostToDevice, w.st));
  HIPCHK(hipMemcpyAsync(w.n2cnt, cnt.data(), sizeof(int32_t) * cnt.size(), hipMemcpyHostToDevice, w.st));
  DevSeqs vs = s;
  vs.res = w.dout.res; vs.off = w.dout.v_off; vs.len = w.dout.v_len; vs.perm = nullptr;
  vs.seg = nullptr; vs.nseg = 0; vs.nseg_slots = 0; vs.nslot = w.dout.reg_cap;
  DevResults vf = w.r;
  vf.fwdsc = w.dout.envsc; vf.fstatus = w.dout.env_st;
  DevResults vb = w.r;
  vb.bcksc = w.n2dc; vb.fstatus = w.n2st;
  DomBufs eb;
  eb.env_loop = w.dout.env_loop; eb.env_move = w.dout.env_move; eb.env_cap = w.dout.reg_cap;
  eb.n2pool = w.n2pool; eb.n2off = w.n2off; eb.n2dc = w.n2dc; eb.n2st = w.n2st;
  eb.n2scl = w.dout.rscale; eb.n2soff = w.dout.rsoff;
  eb.n2oa = w.n2oa; eb.n2ali = w.n2ali;
  eb.n2path = w.n2path; eb.n2poff = w.n2poff; eb.n2pn = w.n2pn;
  const LenCfg lc = lencfg_per_target();
  for (size_t b = 0; b < batches.size(); b++) {
    const int32_t *bl = w.n2list + batches[b].first;
    const int64_t nb = (int64_t) (batches[b].second - batches[b].first);
    const int32_t *c = w.n2cnt + b * 3 * 16;
    eb.n2mode = 0;
    HIPCHK(launch_bck(p, vs, bl, c + 16, nb, vb, nullptr, nullptr, nullptr, nullptr, lc, w.st, eb, w.l));
    eb.n2mode = 2;
    if (b == 0 && fwd_after) HIPCHK(hipStreamWaitEvent(w.st, fwd_after, 0));
    HIPCHK(launch_fwd(p, vs, bl, c + 32, nb, vf, w.l, g, 0, lc, nullptr, 0, nullptr, nullptr, w.st, eb));
    HIPCHK(launch_oa_trace(p, vs, bl, nb, eb, w.st));
  }
  HIPCHK(hipStreamSynchronize(w.st));
  return p7G_OK;
}

// Rescoring of the regions after a decoding run left its nregs region
// records on the device (rescore_isolated_domain, p7_domaindef.c:812-947):
// the multidomain regions' sampled traces start on the second stream; the
// single-domain regions (w = 0) get their null2 correction and optimal
// accuracy alignment meanwhile; then the multidomain regions' clusters become
// envelope records of their own (w = 2), which get theirs. A cluster's
// correction is then replaced by the sum of its region's sampled null2
// scores; a multidomain region's record keeps status 2 and, as its
// correction, the sum over all its residues. *ret_nrec: records in all.
static int
null2_regions(Work &w, const DevProfile &p, const DevSeqs &s, int64_t nregs, const GateThresh &g, const int32_t *h_len,
              int64_t *ret_nrec, double *ret_ms_md)
{
  *ret_nrec = nregs;
  *ret_ms_md = 0.0;
  if (nregs <= 0) return p7G_OK;
  std::vector<int4> rec((size_t) nregs);
  std::vector<int32_t> vlen((size_t) nregs);
  HIPCHK(hipMemcpyAsync(rec.data(), w.dout.regs, sizeof(int4) * (size_t) nregs, hipMemcpyDeviceToHost, w.st));
  HIPCHK(hipMemcpyAsync(vlen.data(), w.dout.v_len, sizeof(int32_t) * (size_t) nregs, hipMemcpyDeviceToHost, w.st));
  HIPCHK(hipStreamSynchronize(w.st));
  std::vector<float> env_dc;
  std::vector<std::pair<int64_t, float>> reg_sum;
  int status;
  const auto t0 = std::chrono::steady_clock::now();
  MdRun m;
  // an early return leaves no region launch running on st2 behind it
  struct Drain { Work &w; const MdRun &m; ~Drain() { if (m.launched && w.st2) (void) hipStreamSynchronize(w.st2); } } drain{ w, m };
  if ((status = md_start(w, p, rec, vlen, nregs, h_len, m)) != p7G_OK) return status;
  // every record's status, correction and alignment start as "not rescored"
  auto clear = [&](int64_t k0, int64_t k1) -> int {
    const int64_t n = k1 - k0;
    if (n <= 0) return p7G_OK;
    HIPCHK(launch_fill32((uint32_t *) (w.n2st + k0), 2u, n, w.st));
    HIPCHK(launch_fill32((uint32_t *) (w.n2dc + k0), 0x7fc00000u, n, w.st));
    HIPCHK(launch_fill32((uint32_t *) (w.n2oa + k0), 0x7fc00000u, n, w.st));
    HIPCHK(launch_fill32((uint32_t *) (w.n2ali + k0), 0u, 4 * n, w.st));
    HIPCHK(launch_fill32((uint32_t *) (w.n2pn + k0), 0u, n, w.st));
    return p7G_OK;
  };
  if ((status = clear(0, nregs)) != p7G_OK) return status;
  // the alignments' words: an envelope's length + M + 2 each; room for the
  // clusters' envelopes too (at most MD_MAXC per multidomain region, each
  // within its region)
  std::vector<int64_t> poff((size_t) nregs, 0);
  int64_t ptot = 0, pbound = 0;
  for (int64_t k = 0; k < nregs; k++) {
    poff[(size_t) k] = ptot;
    if (rec[(size_t) k].w != 1) ptot += (int64_t) vlen[(size_t) k] + p.M + 2;
  }
  for (int64_t k : m.mdk) pbound += (int64_t) MD_MAXC * ((int64_t) vlen[(size_t) k] + p.M + 2);
  if (w.n2path_cap < ptot + pbound) {
    HIPCHK(hipStreamSynchronize(w.st));
    if (w.n2path) { (void) hipFree(w.n2path); w.n2path = nullptr; w.n2path_cap = 0; }
    HIPCHK(hipMalloc((void **) &w.n2path, sizeof(uint32_t) * (size_t) std::max<int64_t>(ptot + pbound, 1)));
    w.n2path_cap = ptot + pbound;
  }
  HIPCHK(hipMemcpyAsync(w.n2poff, poff.data(), sizeof(int64_t) * (size_t) nregs, hipMemcpyHostToDevice, w.st));
  HIPCHK(hipStreamSynchronize(w.st));          // poff is the copy's source
  // the null2 pass's persistent workgroups would hold the compute units the
  // thresholds' grid needs: its Backward runs beside the regions' Forward
  // (one wave per region), its Forward once the thresholds are filled,
  // beside the traces
  if ((status = null2_pass(w, p, s, g, rec, vlen, 0, nregs, m.launched ? w.md_thr : nullptr)) != p7G_OK) return status;
  if ((status = md_finish(w, p, s, rec, vlen, nregs, g, env_dc, reg_sum, m)) != p7G_OK) return status;
  *ret_ms_md = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  const int64_t nrec = (int64_t) rec.size();
  if (nrec > nregs) {
    if ((status = clear(nregs, nrec)) != p7G_OK) return status;
    std::vector<int64_t> cpoff((size_t) (nrec - nregs), 0);
    for (int64_t k = nregs; k < nrec; k++) {
      cpoff[(size_t) (k - nregs)] = ptot;
      ptot += (int64_t) vlen[(size_t) k] + p.M + 2;
    }
    if (ptot > w.n2path_cap) { p7g_set_error("internal: alignment words of the clusters' envelopes"); return p7G_EINVAL; }
    HIPCHK(hipMemcpyAsync(w.n2poff + nregs, cpoff.data(), sizeof(int64_t) * (size_t) (nrec - nregs), hipMemcpyHostToDevice, w.st));
    HIPCHK(hipStreamSynchronize(w.st));        // cpoff is the copy's source
    if ((status = null2_pass(w, p, s, g, rec, vlen, nregs, nrec)) != p7G_OK) return status;
  }
  if (! env_dc.empty() || ! reg_sum.empty()) {
    // corrections from the sampled traces: clusters' over their envelopes,
    // multidomain regions' over all their residues
    std::vector<float> dc((size_t) nrec);
    HIPCHK(hipMemcpyAsync(dc.data(), w.n2dc, 4 * (size_t) nrec, hipMemcpyDeviceToHost, w.st));
    HIPCHK(hipStreamSynchronize(w.st));
    for (size_t e = 0; e < env_dc.size(); e++) dc[(size_t) nregs + e] = env_dc[e];
    for (const auto &r : reg_sum) dc[(size_t) r.first] = r.second;
    HIPCHK(hipMemcpyAsync(w.n2dc, dc.data(), 4 * (size_t) nrec, hipMemcpyHostToDevice, w.st));
    HIPCHK(hipStreamSynchronize(w.st));
  }
  HIPCHK(hipStreamSynchronize(w.st));          // the host vectors above are the copies' sources
  *ret_nrec = nrec;
  return p7G_OK;
}

extern "C" int
p7_gpu_Pipeline(P7_GPU_PIPELINE *pli, const P7_GPU_OPROFILE *om, const P7_GPU_SEQDB *db, P7_GPU_HITS *hits)
{
  std::lock_guard<std::mutex> lk(pli->mu);
  int status;
  DevProfile *p;
  if ((status = ensure_init()) != p7G_OK) return status;
  if (! pli->w.st) HIPCHK(hipStreamCreateWithFlags(&pli->w.st, hipStreamNonBlocking));
  if ((status = check_query(om, db)) != p7G_OK) return status;
  if ((status = dev_profile(om, db->Lmax, &p)) != p7G_OK) return status;
  const int64_t n = db->nseq;
  // pool for the Forward row scale factors of F3 survivors: bounded by the
  // total residues of every target that may reach the Forward stage
  Work &w = pli->w;
  if ((status = work_reserve(w, n, pipeline_pool_floats(n, db->nres))) != p7G_OK) return status;
  const bool dom = pli->do_domains != 0;
  if (dom && (status = dom_reserve(w, n, dom_pool_floats(n, db->nres))) != p7G_OK) return status;
  pli->regs_valid = false;
  DevSeqs s = seqs_of(db);
  const GateThresh g = gates_of(pli, om);

  unsigned long long sums[8];
  int32_t counts[16];
  int32_t nclu_total = 0, nregs_dev = 0;
  int64_t nrecs = 0;                      // region records, with the clusters' envelopes
  double ms_md = 0.0;                     // host time of the multidomain regions (they synchronize)
  for (int attempt = 0; ; attempt++) {
  if ((status = chain_reset(w, n)) != p7G_OK) return status;
  if (dom) {                              // targets that do not reach decoding: NaN, no regions
    HIPCHK(launch_fill32(w.dout.nexp, 0x7fc00000u, n, w.st));
    HIPCHK(launch_fill32(w.dout.nreg, 0u, n, w.st));
    HIPCHK(launch_fill32(w.dout.nclu, 0u, n, w.st));
    HIPCHK(launch_fill32(w.dout.nclu_total, 0u, 1, w.st));
    HIPCHK(launch_fill32(w.dout.ecount, 0u, 16, w.st));
  }
  HIPCHK(hipEventRecord(w.ev[0], w.st));
  if ((status = work_bnd(w, *p, db->bytes)) != p7G_OK) return status;
  HIPCHK(launch_ssv_chosen(*p, s, nullptr, n, w.r.zmax, &w.l.counts[P7G_Q_SSV], w.bnd, w.bnd_cap, w.st));
  HIPCHK(hipEventRecord(w.ev[6], w.st));
  if ((status = chain_stages(w, *p, s, n, w.r.zmax, g, true, dom)) != p7G_OK) return status;

  // counters on the device, then the requested per-target arrays in the caller's order
  HIPCHK(launch_fill32(w.stats, 0u, 16, w.st));
  HIPCHK(launch_stats(s, w.r.stage, n, w.l.vit, w.l.fwd, w.l.counts, w.stats, w.st));
  HIPCHK(hipMemcpyAsync(sums, w.stats, sizeof(sums), hipMemcpyDeviceToHost, w.st));
  HIPCHK(hipMemcpyAsync(counts, w.l.counts, sizeof(counts), hipMemcpyDeviceToHost, w.st));
  if (hits && hits->stage) {
    HIPCHK(launch_scatter8(s, w.r.stage, (uint8_t *) w.outbuf, n, w.st));
    HIPCHK(hipMemcpyAsync(hits->stage, w.outbuf, (size_t) n, hipMemcpyDeviceToHost, w.st));
  }
  struct { float *dst; const float *src; } fl[] = {
    { hits ? hits->nullsc : nullptr, w.r.nullsc }, { hits ? hits->usc : nullptr, w.r.usc },
    { hits ? hits->filtersc : nullptr, w.r.filtersc }, { hits ? hits->vfsc : nullptr, w.r.vfsc },
    { hits ? hits->fwdsc : nullptr, w.r.fwdsc }, { hits ? hits->bcksc : nullptr, w.r.bcksc } };
  for (int a = 0; a < 6; a++) if (fl[a].dst) {
    // stream order keeps the staging buffer's next use behind this copy
    HIPCHK(launch_scatter32(s, fl[a].src, w.outbuf, n, w.st));
    HIPCHK(hipMemcpyAsync(fl[a].dst, w.outbuf, sizeof(float) * (size_t) n, hipMemcpyDeviceToHost, w.st));
  }
  if (dom && hits) {
    struct { void *dst; const void *src; } dl[] = {
      { hits->nexpected, w.dout.nexp }, { hits->nregions, w.dout.nreg }, { hits->nclustered, w.dout.nclu } };
    for (int a = 0; a < 3; a++) if (dl[a].dst) {
      HIPCHK(launch_scatter32(s, dl[a].src, w.outbuf, n, w.st));
      HIPCHK(hipMemcpyAsync(dl[a].dst, w.outbuf, 4 * (size_t) n, hipMemcpyDeviceToHost, w.st));
    }
  }
  if (dom) {
    HIPCHK(hipMemcpyAsync(&nclu_total, w.dout.nclu_total, 4, hipMemcpyDeviceToHost, w.st));
    HIPCHK(hipMemcpyAsync(&nregs_dev, w.dout.ecount, 4, hipMemcpyDeviceToHost, w.st));
  }
  HIPCHK(hipStreamSynchronize(w.st));
  if (! counts[6]) break;
  // the scale-factor pool was full: grow it to the rows of every target, run again
  if (attempt > 0 || w.pool_size >= pool_rows_all(n, db->nres)) { p7g_set_error("Forward scale-factor pool exhausted"); return p7G_EMEM; }
  if ((status = work_reserve(w, n, pool_rows_all(n, db->nres))) != p7G_OK) return status;
  }
  if (dom) {
    if (counts[7]) {
      p7g_set_error("domain decoding %s exhausted", (counts[7] & 1) ? "row pool" : (counts[7] & 2) ? "region list" : "region residue pool");
      return p7G_EMEM;
    }
    if (pli->do_null2) {
      HIPCHK(hipEventRecord(w.ev[8], w.st));
      int64_t nrec = nregs_dev;
      if ((status = null2_regions(w, *p, s, nregs_dev, g, db->h_len, &nrec, &ms_md)) != p7G_OK) return status;
      nrecs = nrec;
      HIPCHK(hipEventRecord(w.ev[9], w.st));
      HIPCHK(hipEventSynchronize(w.ev[9]));
    }
    if (nrecs < nregs_dev) nrecs = nregs_dev;
    // the records' caller indices and target lengths, so that the results
    // do not read the database after the call
    HIPCHK(launch_rec_gather(s, w.dout.regs, nrecs, w.dout.rseq, w.dout.rlen, w.st));
    HIPCHK(hipStreamSynchronize(w.st));
    std::lock_guard<std::mutex> rl(pli->regs_mu);
    pli->nregs = nrecs; pli->regs_dev = w.dout.regs; pli->envsc_dev = w.dout.envsc;
    pli->rseq_dev = w.dout.rseq; pli->rlen_dev = w.dout.rlen;
    pli->n2dc_dev = pli->do_null2 ? w.n2dc : nullptr; pli->n2st_dev = w.n2st;
    pli->oa_dev = pli->do_null2 ? w.n2oa : nullptr; pli->ali_dev = pli->do_null2 ? w.n2ali : nullptr;
    pli->path_dev = pli->do_null2 ? w.n2path : nullptr; pli->poff_dev = w.n2poff; pli->pn_dev = w.n2pn;
    pli->d_abc = om->abc; pli->d_M = om->M;
    pli->d_cons = om->cons ? om->cons : "";
    pli->d_has_cs = om->cs != nullptr; pli->d_cs = om->cs ? om->cs : "";
    pli->d_has_rf = om->rf != nullptr; pli->d_rf = om->rf ? om->rf : "";
    pli->d_rfv.assign(om->rfv, om->rfv + (size_t) om->abc->Kp * om->M);
    pli->fwdsc_dev = w.r.fwdsc; pli->nexp_dev = w.dout.nexp; pli->nslots = n;
    pli->ftau = om->evparam[p7G_FTAU]; pli->flambda = om->evparam[p7G_FLAMBDA];
    pli->regs_ready = false; pli->regs.clear();
    pli->hits_ready = false;
    pli->regs_valid = true;
  }

  P7_GPU_STATS one;
  stats_of_run(one, sums, counts, om, db, pli->do_biasfilter);
  stats_accumulate(pli->st, one, true);
  P7_GPU_STATS &st = pli->st;
  float ms[5], mtot;
  for (int a = 0; a < 5; a++) HIPCHK(hipEventElapsedTime(&ms[a], w.ev[a], w.ev[a+1]));
  HIPCHK(hipEventElapsedTime(&mtot, w.ev[0], w.ev[5]));
  st.ms_total = mtot;
  HIPCHK(hipEventElapsedTime(&mtot, w.ev[0], w.ev[6]));
  st.ms_ssv = mtot;
  st.ms_msv = ms[0]; st.ms_bias = ms[1]; st.ms_vit = ms[2]; st.ms_fwd = ms[3]; st.ms_bck = ms[4];
  st.ms_domains = 0.0; st.n_regions = 0; st.n_clustered = 0;
  if (dom) {
    HIPCHK(hipEventElapsedTime(&mtot, w.ev[5], w.ev[7]));
    st.ms_domains = mtot;
    st.ms_null2 = 0.0;
    if (pli->do_null2) { HIPCHK(hipEventElapsedTime(&mtot, w.ev[8], w.ev[9])); st.ms_null2 = mtot; }
    st.ms_multidomain = ms_md;
    st.n_regions = nregs_dev;
    st.n_clustered = nclu_total;
  }
  return p7G_OK;
}

extern "C" int
p7_gpu_pipeline_SetDomains(P7_GPU_PIPELINE *pli, int on)
{
  if (! pli) return p7G_EINVAL;
  std::lock_guard<std::mutex> lk(pli->mu);
  pli->do_domains = on ? 1 : 0;
  return p7G_OK;
}

extern "C" int
p7_gpu_pipeline_SetNull2(P7_GPU_PIPELINE *pli, int on)
{
  if (! pli) return p7G_EINVAL;
  std::lock_guard<std::mutex> lk(pli->mu);
  pli->do_null2 = on ? 1 : 0;
  return p7G_OK;
}

extern "C" int
p7_gpu_pipeline_SetPoolCap(P7_GPU_PIPELINE *pli, int64_t floats)
{
  if (! pli || floats < 0) return p7G_EINVAL;
  std::lock_guard<std::mutex> lk(pli->mu);
  pli->w.pool_cap = floats;
  return p7G_OK;
}

// The last decoding run's region records on the host, in sequence then start
// order (called with regs_mu held).
static int
regs_fetch(const P7_GPU_PIPELINE *pli)
{
  if (! pli->regs_ready) {
    // slot -> the caller's sequence index, then sequence and start order
    const size_t nr = (size_t) pli->nregs;
    std::vector<int4> rec(nr);
    std::vector<float> sc(nr), dc(nr, NAN);
    std::vector<int32_t> n2(nr, -1);
    std::vector<float> oa(nr, NAN);
    std::vector<int4> al(nr, make_int4(0, 0, 0, 0));
    std::vector<int32_t> rseq(nr), rlen(nr);
    if (nr) {
      HIPCHK(hipMemcpy(rec.data(), pli->regs_dev, sizeof(int4) * nr, hipMemcpyDeviceToHost));
      HIPCHK(hipMemcpy(rseq.data(), pli->rseq_dev, sizeof(int32_t) * nr, hipMemcpyDeviceToHost));
      HIPCHK(hipMemcpy(rlen.data(), pli->rlen_dev, sizeof(int32_t) * nr, hipMemcpyDeviceToHost));
      HIPCHK(hipMemcpy(sc.data(), pli->envsc_dev, sizeof(float) * nr, hipMemcpyDeviceToHost));
      if (pli->n2dc_dev) {
        HIPCHK(hipMemcpy(dc.data(), pli->n2dc_dev, sizeof(float) * nr, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(n2.data(), pli->n2st_dev, sizeof(int32_t) * nr, hipMemcpyDeviceToHost));
      }
      if (pli->oa_dev) {
        HIPCHK(hipMemcpy(oa.data(), pli->oa_dev, sizeof(float) * nr, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(al.data(), pli->ali_dev, sizeof(int4) * nr, hipMemcpyDeviceToHost));
      }
    }
    std::vector<int32_t> slot_of(nr);
    for (size_t k = 0; k < nr; k++) { slot_of[k] = rec[k].x; rec[k].x = rseq[k]; }
    std::vector<int64_t> ord((size_t) pli->nregs);
    for (int64_t k = 0; k < pli->nregs; k++) ord[(size_t) k] = k;
    std::sort(ord.begin(), ord.end(), [&](int64_t a, int64_t b) {
      const int4 &x = rec[(size_t) a], &y = rec[(size_t) b];
      return x.x != y.x ? x.x < y.x : x.y != y.y ? x.y < y.y : x.w < y.w; });
    pli->regs.resize(nr); pli->envsc.resize(nr); pli->n2dc.resize(nr); pli->n2st.resize(nr); pli->regs_slot.resize(nr);
    pli->regs_L.resize(nr); pli->regs_didx.resize(nr);
    pli->oasc.resize(nr); pli->ali.resize(nr);
    for (size_t k = 0; k < nr; k++) {
      const size_t o = (size_t) ord[k];
      pli->regs[k] = rec[o]; pli->envsc[k] = sc[o]; pli->n2dc[k] = dc[o]; pli->n2st[k] = n2[o];
      pli->regs_slot[k] = slot_of[o]; pli->regs_L[k] = rlen[o]; pli->regs_didx[k] = (int64_t) o;
      // alignment coordinates in the target: the region copy starts at residue ienv
      const int4 a = al[o];
      pli->oasc[k] = (n2[o] == 0 && a.x > 0) ? oa[o] : NAN;
      pli->ali[k] = (n2[o] == 0 && a.x > 0) ? make_int4(a.x, a.y, a.z + rec[o].y - 1, a.w + rec[o].y - 1)
                                            : make_int4(0, 0, 0, 0);
    }
    pli->regs_ready = true;
  }
  return p7G_OK;
}

extern "C" int
p7_gpu_pipeline_GetRegions(const P7_GPU_PIPELINE *pli, int64_t *seq, int32_t *ienv, int32_t *jenv, int32_t *multi,
                           float *envsc, int64_t cap, int64_t *ret_n)
{
  if (! pli || ! pli->regs_valid) { p7g_set_error("no domain regions: the last p7_gpu_Pipeline call did not decode"); return p7G_EINVAL; }
  std::lock_guard<std::mutex> rl(pli->regs_mu);
  int status;
  if ((status = regs_fetch(pli)) != p7G_OK) return status;
  const int64_t n = (int64_t) pli->regs.size();
  if (ret_n) *ret_n = n;
  for (int64_t k = 0; k < std::min(n, cap); k++) {
    const int4 &r = pli->regs[(size_t) k];
    if (seq) seq[k] = r.x;
    if (ienv) ienv[k] = r.y;
    if (jenv) jenv[k] = r.z;
    if (multi) multi[k] = r.w;
    if (envsc) envsc[k] = pli->envsc[(size_t) k];
  }
  return p7G_OK;
}

extern "C" int
p7_gpu_pipeline_GetRegionNull2(const P7_GPU_PIPELINE *pli, float *domcorr, int32_t *status, int64_t cap, int64_t *ret_n)
{
  if (! pli || ! pli->regs_valid) { p7g_set_error("no domain regions: the last p7_gpu_Pipeline call did not decode"); return p7G_EINVAL; }
  if (! pli->n2dc_dev) { p7g_set_error("the last p7_gpu_Pipeline call ran without null2"); return p7G_EINVAL; }
  std::lock_guard<std::mutex> rl(pli->regs_mu);
  int st;
  if ((st = regs_fetch(pli)) != p7G_OK) return st;
  const int64_t n = (int64_t) pli->regs.size();
  if (ret_n) *ret_n = n;
  for (int64_t k = 0; k < std::min(n, cap); k++) {
    if (domcorr) domcorr[k] = pli->n2dc[(size_t) k];
    if (status) status[k] = pli->n2st[(size_t) k];
  }
  return p7G_OK;
}

extern "C" int
p7_gpu_pipeline_GetDomainAlignments(const P7_GPU_PIPELINE *pli, float *oasc, int32_t *hmmfrom, int32_t *hmmto,
                                    int32_t *alifrom, int32_t *alito, int64_t cap, int64_t *ret_n)
{
  if (! pli || ! pli->regs_valid) { p7g_set_error("no domain regions: the last p7_gpu_Pipeline call did not decode"); return p7G_EINVAL; }
  if (! pli->oa_dev) { p7g_set_error("the last p7_gpu_Pipeline call ran without null2 (and its alignments)"); return p7G_EINVAL; }
  std::lock_guard<std::mutex> rl(pli->regs_mu);
  int st;
  if ((st = regs_fetch(pli)) != p7G_OK) return st;
  const int64_t n = (int64_t) pli->regs.size();
  if (ret_n) *ret_n = n;
  for (int64_t k = 0; k < std::min(n, cap); k++) {
    const int4 &a = pli->ali[(size_t) k];
    if (oasc) oasc[k] = pli->oasc[(size_t) k];
    if (hmmfrom) hmmfrom[k] = a.x;
    if (hmmto) hmmto[k] = a.y;
    if (alifrom) alifrom[k] = a.z;
    if (alito) alito[k] = a.w;
  }
  return p7G_OK;
}

// Hit-list scores of the last decoding run with null2: per target with
// regions, the p7_Pipeline scores after domain definition
// (p7_pipeline.c:743-850, host_hits.c), and per region its domain's.
static int
hit_scores(const P7_GPU_PIPELINE *pli)
{
  if (pli->hits_ready) return p7G_OK;
  std::vector<int64_t> &tseq = pli->h_tseq;
  std::vector<P7G_TSCORE> &ts = pli->h_ts;
  std::vector<int32_t> &tnd = pli->h_tnd, &tst = pli->h_tst;
  std::vector<P7G_DSCORE> &ds = pli->h_ds;
  tseq.clear(); ts.clear(); tnd.clear(); tst.clear(); pli->h_nexp.clear(); pli->h_tcnt.clear();
  const size_t nr = pli->regs.size();
  std::vector<float> fwd((size_t) pli->nslots), nexp((size_t) pli->nslots);
  if (pli->nslots) {
    HIPCHK(hipMemcpy(fwd.data(), pli->fwdsc_dev, sizeof(float) * (size_t) pli->nslots, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(nexp.data(), pli->nexp_dev, sizeof(float) * (size_t) pli->nslots, hipMemcpyDeviceToHost));
  }
  const float omega = (float) (1. / 256.);                    // p7_bg.c:75
  ds.assign(nr, P7G_DSCORE{ NAN, NAN, NAN });
  for (size_t a = 0; a < nr; ) {
    size_t b = a;
    while (b < nr && pli->regs[b].x == pli->regs[a].x) b++;    // regions a..b-1: one target
    const int slot = pli->regs_slot[a];
    const int L = pli->regs_L[a];
    // domains: single-domain regions (w 0) and multidomain regions' clusters
    // (w 2) whose rescoring succeeded; the sequence bias sums the null2 scores
    // of every residue: the single-domain envelopes' and, for a multidomain
    // region (w 1), all of its residues' sampled ones
    std::vector<P7G_DOMIN> dom;
    std::vector<size_t> which;
    std::vector<float> parts;
    // P7_DOMAINDEF's counters (p7_domaindef.c:426-483): a cluster overlaps
    // when it starts at or before the end of the last cluster of its region
    // whose rescoring succeeded
    int4 cnt = make_int4(0, 0, 0, 0);
    int last_j2 = 0;
    for (size_t k = a; k < b; k++) {
      const int kind = pli->regs[k].w;
      if (kind != 2) { cnt.x++; last_j2 = 0; }
      if (kind == 1) cnt.y++;
      else {
        cnt.w++;
        if (kind == 2) {
          if (pli->regs[k].y <= last_j2) cnt.z++;
          if (pli->n2st[k] == 0) last_j2 = pli->regs[k].z;
        }
      }
    }
    for (size_t k = a; k < b; k++) {
      const int kind = pli->regs[k].w;
      if (kind == 1) { parts.push_back(pli->n2dc[k]); continue; }
      if (pli->n2st[k] != 0) continue;                      // rescoring failed: no domain (p7_domaindef.c:841-843)
      if (kind == 0) parts.push_back(pli->n2dc[k]);
      dom.push_back(P7G_DOMIN{ pli->envsc[k], pli->n2dc[k], pli->regs[k].y, pli->regs[k].z });
      which.push_back(k);
    }
    P7G_TSCORE t{ NAN, NAN, NAN, NAN, NAN, NAN, -1 };
    std::vector<P7G_DSCORE> d(dom.size());
    p7g_target_scores(L, fwd[(size_t) slot], p7g_null_one(L), omega, pli->ftau, pli->flambda, dom.data(), (int) dom.size(),
                      parts.data(), (int) parts.size(), &t, d.data());
    for (size_t q = 0; q < dom.size(); q++) ds[which[q]] = d[q];
    if (! dom.empty()) t.best = (int) (which[(size_t) t.best] - a);
    else t.best = -1;
    // no envelope (every region multidomain, no significant cluster): not a
    // hit (p7_pipeline.c:739-740)
    tseq.push_back(pli->regs[a].x); ts.push_back(t); tnd.push_back((int32_t) dom.size()); tst.push_back(cnt.w ? 0 : 1);
    pli->h_nexp.push_back(nexp[(size_t) slot]); pli->h_tcnt.push_back(cnt);
    a = b;
  }
  pli->hits_ready = true;
  return p7G_OK;
}

extern "C" int
p7_gpu_pipeline_GetTargetScores(const P7_GPU_PIPELINE *pli, int64_t *seq, float *score, float *pre_score, float *sum_score,
                                double *lnP, int32_t *ndom, int32_t *best, int32_t *status, int64_t cap, int64_t *ret_n)
{
  if (! pli || ! pli->regs_valid) { p7g_set_error("no domain regions: the last p7_gpu_Pipeline call did not decode"); return p7G_EINVAL; }
  if (! pli->n2dc_dev) { p7g_set_error("the last p7_gpu_Pipeline call ran without null2"); return p7G_EINVAL; }
  std::lock_guard<std::mutex> rl(pli->regs_mu);
  int st;
  if ((st = regs_fetch(pli)) != p7G_OK) return st;
  if ((st = hit_scores(pli)) != p7G_OK) return st;
  const std::vector<int64_t> &tseq = pli->h_tseq;
  const std::vector<P7G_TSCORE> &ts = pli->h_ts;
  const std::vector<int32_t> &tnd = pli->h_tnd, &tst = pli->h_tst;
  const int64_t n = (int64_t) tseq.size();
  if (ret_n) *ret_n = n;
  for (int64_t k = 0; k < std::min(n, cap); k++) {
    if (seq) seq[k] = tseq[(size_t) k];
    if (score) score[k] = ts[(size_t) k].score;
    if (pre_score) pre_score[k] = ts[(size_t) k].pre_score;
    if (sum_score) sum_score[k] = ts[(size_t) k].sum_score;
    if (lnP) lnP[k] = ts[(size_t) k].lnP;
    if (ndom) ndom[k] = tnd[(size_t) k];
    if (best) best[k] = ts[(size_t) k].best;
    if (status) status[k] = tst[(size_t) k];
  }
  return p7G_OK;
}

extern "C" int
p7_gpu_pipeline_GetDomainScores(const P7_GPU_PIPELINE *pli, float *bitscore, float *dombias, double *lnP, int64_t cap,
                                int64_t *ret_n)
{
  if (! pli || ! pli->regs_valid) { p7g_set_error("no domain regions: the last p7_gpu_Pipeline call did not decode"); return p7G_EINVAL; }
  if (! pli->n2dc_dev) { p7g_set_error("the last p7_gpu_Pipeline call ran without null2"); return p7G_EINVAL; }
  std::lock_guard<std::mutex> rl(pli->regs_mu);
  int st;
  if ((st = regs_fetch(pli)) != p7G_OK) return st;
  if ((st = hit_scores(pli)) != p7G_OK) return st;
  const std::vector<P7G_DSCORE> &ds = pli->h_ds;
  const int64_t n = (int64_t) ds.size();
  if (ret_n) *ret_n = n;
  for (int64_t k = 0; k < std::min(n, cap); k++) {
    if (bitscore) bitscore[k] = ds[(size_t) k].bitscore;
    if (dombias) dombias[k] = ds[(size_t) k].dombias;
    if (lnP) lnP[k] = ds[(size_t) k].lnP;
  }
  return p7G_OK;
}

// The hit list of the scored targets (status 0): p7_tophits_SortBySortkey
// with the sort key of E-value inclusion (-lnP, p7_pipeline.c:827; ties by
// sequence index), then p7_tophits_Threshold (p7_tophits.c:866-935) with
// E-value thresholds: a target is reported if exp(lnP)·Z <= E and included
// if also <= incE; domZ = the number of reported targets; a domain of a
// reported target is reported if exp(lnP_d)·domZ <= domE, and included if
// its target is and exp(lnP_d)·domZ <= incdomE.
// p7_tophits_SortBySortkey and p7_tophits_Threshold with E-value thresholds
// (p7_tophits.c:866-935): the scored targets ranked by -lnP (ties in
// sequence order), their flags (bit 0 reported, bit 1 included), the regions'
// domain flags, and domZ: domZ_set when it is >= 0 (--domZ), else the number
// of reported targets
static void
hit_list(const P7_GPU_PIPELINE *pli, double Z, double E, double incE, double domE, double incdomE, double domZ_set,
         std::vector<int64_t> &ord, std::vector<int32_t> &tf, std::vector<int32_t> &df, double &domZ)
{
  const size_t nt = pli->h_tseq.size(), nr = pli->regs.size();
  ord.clear();
  for (size_t k = 0; k < nt; k++) if (pli->h_tst[k] == 0) ord.push_back((int64_t) k);
  std::stable_sort(ord.begin(), ord.end(), [&](int64_t a, int64_t b) {
    const double ka = -pli->h_ts[(size_t) a].lnP, kb = -pli->h_ts[(size_t) b].lnP;
    return ka > kb; });
  tf.assign(nt, 0);
  int64_t nrep = 0;
  for (int64_t k : ord) {
    const double lnP = pli->h_ts[(size_t) k].lnP;
    if (exp(lnP) * Z <= E) {
      tf[(size_t) k] |= 1; nrep++;
      if (exp(lnP) * Z <= incE) tf[(size_t) k] |= 2;
    }
  }
  domZ = (domZ_set >= 0.0) ? domZ_set : (double) nrep;
  df.assign(nr, 0);
  size_t a = 0;
  for (size_t k = 0; k < nt; k++) {                   // targets and regions are both in sequence order
    while (a < nr && (int64_t) pli->regs[a].x != pli->h_tseq[k]) a++;
    if (pli->h_tst[k] != 0) continue;
    for (size_t q = a; q < nr && (int64_t) pli->regs[q].x == pli->h_tseq[k]; q++) {
      if (! (tf[k] & 1) || std::isnan(pli->h_ds[q].bitscore)) continue;
      const double e = exp(pli->h_ds[q].lnP) * domZ;
      if (e <= domE) df[q] |= 1;
      if ((tf[k] & 2) && e <= incdomE) df[q] |= 2;
    }
    // workaround_bug_h74 (p7_tophits.c:678-698): in a target with
    // overlapping clusters, of two domains with the same alignment start and
    // end only the higher bit score stays reported and included
    if (pli->h_tcnt[k].z > 0) {
      std::vector<size_t> dq;
      for (size_t q = a; q < nr && (int64_t) pli->regs[q].x == pli->h_tseq[k]; q++)
        if (! std::isnan(pli->h_ds[q].bitscore)) dq.push_back(q);
      for (size_t d1 = 0; d1 < dq.size(); d1++)
        for (size_t d2 = d1 + 1; d2 < dq.size(); d2++) {
          const int4 &x = pli->ali[dq[d1]], &y = pli->ali[dq[d2]];
          if (x.z == y.z && x.w == y.w)
            df[pli->h_ds[dq[d1]].bitscore >= pli->h_ds[dq[d2]].bitscore ? dq[d2] : dq[d1]] = 0;
        }
    }
  }
}

static int
hits_ready(const P7_GPU_PIPELINE *pli)
{
  if (! pli || ! pli->regs_valid) { p7g_set_error("no domain regions: the last p7_gpu_Pipeline call did not decode"); return p7G_EINVAL; }
  if (! pli->n2dc_dev) { p7g_set_error("the last p7_gpu_Pipeline call ran without null2"); return p7G_EINVAL; }
  int st;
  if ((st = regs_fetch(pli)) != p7G_OK) return st;
  return hit_scores(pli);
}

extern "C" int
p7_gpu_pipeline_GetHitList(const P7_GPU_PIPELINE *pli, double Z, double E, double incE, double domE, double incdomE,
                           int64_t *order, int32_t *flags, int32_t *dflags, double *ret_domZ, int64_t cap, int64_t *ret_n)
{
  if (! pli) { p7g_set_error("no pipeline"); return p7G_EINVAL; }
  std::lock_guard<std::mutex> rl(pli->regs_mu);
  int st;
  if ((st = hits_ready(pli)) != p7G_OK) return st;
  std::vector<int64_t> ord;
  std::vector<int32_t> tf, df;
  double domZ;
  hit_list(pli, Z, E, incE, domE, incdomE, -1.0, ord, tf, df, domZ);
  if (dflags) for (size_t q = 0; q < df.size(); q++) dflags[q] = df[q];
  if (ret_domZ) *ret_domZ = domZ;
  const int64_t n = (int64_t) ord.size();
  if (ret_n) *ret_n = n;
  for (int64_t k = 0; k < std::min(n, cap); k++) {
    if (order) order[k] = ord[(size_t) k];
    if (flags) flags[k] = tf[(size_t) ord[(size_t) k]];
  }
  return p7G_OK;
}

extern "C" int
p7_gpu_pipeline_GetTargetCounts(const P7_GPU_PIPELINE *pli, float *nexpected, int32_t *nregions, int32_t *nclustered,
                                int32_t *noverlaps, int32_t *nenvelopes, int64_t cap, int64_t *ret_n)
{
  if (! pli) { p7g_set_error("no pipeline"); return p7G_EINVAL; }
  std::lock_guard<std::mutex> rl(pli->regs_mu);
  int st;
  if ((st = hits_ready(pli)) != p7G_OK) return st;
  const int64_t n = (int64_t) pli->h_tseq.size();
  if (ret_n) *ret_n = n;
  for (int64_t k = 0; k < std::min(n, cap); k++) {
    const int4 &c = pli->h_tcnt[(size_t) k];
    if (nexpected) nexpected[k] = pli->h_nexp[(size_t) k];
    if (nregions) nregions[k] = c.x;
    if (nclustered) nclustered[k] = c.y;
    if (noverlaps) noverlaps[k] = c.z;
    if (nenvelopes) nenvelopes[k] = c.w;
  }
  return p7G_OK;
}

// The reported list as the text writers take it: targets in rank order with
// their domains (rescored envelopes, in sequence order) and flags
// The alignment display's lines of one domain (p7_alidisplay_Create,
// p7_alidisplay.c:147-210) from its states in alignment order: consensus
// residue or '.', match line (the residue when it is the consensus residue,
// '+' when its match odds exceed 1, else ' '), the residue (upper case for
// M, lower case for I, '-' for D), its posterior probability code ('.' for
// D), and the CS and RF annotation ('.' under I).
static void
display_lines(const P7_GPU_PIPELINE *pli, const uint32_t *w, int n, std::string &model, std::string &mline,
              std::string &aseq, std::string &pp, std::string &cs, std::string &rf)
{
  const P7G_ALPHABET *abc = pli->d_abc;
  const int M = pli->d_M;
  const bool has_cons = (int) pli->d_cons.size() > M;
  model.assign((size_t) n, ' '); mline.assign((size_t) n, ' '); aseq.assign((size_t) n, ' '); pp.assign((size_t) n, ' ');
  cs.clear(); rf.clear();
  if (pli->d_has_cs) cs.assign((size_t) n, ' ');
  if (pli->d_has_rf) rf.assign((size_t) n, ' ');
  for (int z = 0; z < n; z++) {
    const uint32_t v = w[z];
    const int st = (int) (v & 3u), k = (int) ((v >> 2) & 0x1ffffu), code = (int) ((v >> 19) & 15u), x = (int) ((v >> 23) & 31u);
    const bool kin = k >= 1 && k <= M && x < abc->Kp;
    const char cons = (has_cons && kin) ? pli->d_cons[(size_t) k] : 'x';
    if (st == 0) {
      model[z] = cons;
      const int cx = abc->inmap[(unsigned char) cons & 127];
      if (cx == x) mline[z] = cons;
      else if (kin && pli->d_rfv[(size_t) x * M + k - 1] > 1.0f) mline[z] = '+';
      else mline[z] = ' ';
      aseq[z] = (char) toupper(abc->sym[x]);
    } else if (st == 1) {
      model[z] = '.'; mline[z] = ' ';
      aseq[z] = (char) tolower(abc->sym[x]);
    } else {
      model[z] = cons; mline[z] = ' '; aseq[z] = '-';
    }
    pp[z] = (st == 2) ? '.' : (code >= 10) ? '*' : (char) ('0' + code);
    if (pli->d_has_cs) cs[z] = (st == 1 || ! kin || (int) pli->d_cs.size() <= M) ? '.' : pli->d_cs[(size_t) k];
    if (pli->d_has_rf) rf[z] = (st == 1 || ! kin || (int) pli->d_rf.size() <= M) ? '.' : pli->d_rf[(size_t) k];
  }
}

static int
report_list(const P7_GPU_PIPELINE *pli, const P7_GPU_REPORT *rep, std::vector<P7G_RHIT> &hits,
            std::vector<P7G_RDOM> &doms, P7G_RLIST &lst, std::vector<std::string> &text)
{
  if (! rep || ! rep->qname || ! rep->tname) { p7g_set_error("the report needs the query's name and the targets' names"); return p7G_EINVAL; }
  int st;
  if ((st = hits_ready(pli)) != p7G_OK) return st;
  std::vector<int64_t> ord;
  std::vector<int32_t> tf, df;
  double domZ;
  hit_list(pli, rep->Z, rep->E, rep->incE, rep->domE, rep->incdomE, rep->domZ, ord, tf, df, domZ);
  // first region of each target
  const size_t nt = pli->h_tseq.size(), nr = pli->regs.size();
  std::vector<size_t> first(nt, nr);
  for (size_t k = 0, a = 0; k < nt; k++) {
    while (a < nr && (int64_t) pli->regs[a].x != pli->h_tseq[k]) a++;
    first[k] = a;
  }
  hits.clear(); doms.clear(); text.clear();
  // the alignments' states, when drawn: every record's count and offset, then
  // the words of the reported domains
  const bool ali = rep->show_alignments && pli->path_dev && pli->d_abc;
  std::vector<int32_t> pn;
  std::vector<int64_t> poff;
  if (ali && pli->nregs > 0) {
    pn.resize((size_t) pli->nregs); poff.resize((size_t) pli->nregs);
    HIPCHK(hipMemcpy(pn.data(), pli->pn_dev, sizeof(int32_t) * pn.size(), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(poff.data(), pli->poff_dev, sizeof(int64_t) * poff.size(), hipMemcpyDeviceToHost));
  }
  std::vector<int> dtext;                  // per domain: its first line in `text`, or -1
  std::vector<size_t> dstart;
  for (int64_t k64 : ord) {
    const size_t k = (size_t) k64;
    const int64_t s = pli->h_tseq[k];
    P7G_RHIT h;
    memset(&h, 0, sizeof(h));
    h.name = rep->tname[s];
    h.acc = (rep->tacc && rep->tacc[s] && rep->tacc[s][0]) ? rep->tacc[s] : nullptr;
    h.desc = (rep->tdesc && rep->tdesc[s] && rep->tdesc[s][0]) ? rep->tdesc[s] : nullptr;
    h.L = pli->regs_L[first[k]];
    h.flags = tf[k];
    h.score = pli->h_ts[k].score; h.pre_score = pli->h_ts[k].pre_score; h.lnP = pli->h_ts[k].lnP;
    h.nexpected = pli->h_nexp[k];
    const int4 &c = pli->h_tcnt[k];
    h.nregions = c.x; h.nclustered = c.y; h.noverlaps = c.z; h.nenvelopes = c.w;
    h.best = 0;
    dstart.push_back(doms.size());
    for (size_t q = first[k]; q < nr && pli->regs[q].x == s; q++) {
      if (std::isnan(pli->h_ds[q].bitscore)) continue;
      P7G_RDOM d;
      d.bitscore = pli->h_ds[q].bitscore; d.dombias = pli->h_ds[q].dombias; d.lnP = pli->h_ds[q].lnP;
      d.oasc = pli->oasc[q];
      d.hmmfrom = pli->ali[q].x; d.hmmto = pli->ali[q].y; d.alifrom = pli->ali[q].z; d.alito = pli->ali[q].w;
      d.ienv = pli->regs[q].y; d.jenv = pli->regs[q].z;
      d.flags = df[q];
      d.model = d.mline = d.aseq = d.ppline = d.csline = d.rfline = nullptr; d.N = 0;
      int ti = -1;
      if (ali && (d.flags & 1)) {
        const int64_t o = pli->regs_didx[q];
        const int n = pn[(size_t) o];
        if (n > 0) {
          std::vector<uint32_t> w((size_t) n);
          HIPCHK(hipMemcpy(w.data(), pli->path_dev + poff[(size_t) o], sizeof(uint32_t) * (size_t) n, hipMemcpyDeviceToHost));
          std::reverse(w.begin(), w.end());                  // the kernel writes them last first
          std::string a[6];
          display_lines(pli, w.data(), n, a[0], a[1], a[2], a[3], a[4], a[5]);
          ti = (int) text.size();
          for (int e = 0; e < 6; e++) text.push_back(a[e]);
          d.N = n;
          d.hmmfrom = (int) ((w[0] >> 2) & 0x1ffffu);
          d.hmmto = (int) ((w[(size_t) n - 1] >> 2) & 0x1ffffu);
        }
      }
      dtext.push_back(ti);
      if ((int) (q - first[k]) == pli->h_ts[k].best) h.best = (int) (doms.size() - dstart.back());
      if (d.flags & 1) h.nreported++;
      if (d.flags & 2) h.nincluded++;
      doms.push_back(d);
    }
    h.ndom = (int) (doms.size() - dstart.back());
    hits.push_back(h);
  }
  for (size_t h = 0; h < hits.size(); h++) hits[h].dom = doms.data() + dstart[h];
  for (size_t e = 0; e < doms.size(); e++)            // text no longer grows: its strings stay put
    if (dtext[e] >= 0) {
      const size_t t0 = (size_t) dtext[e];
      doms[e].model = text[t0].c_str(); doms[e].mline = text[t0 + 1].c_str(); doms[e].aseq = text[t0 + 2].c_str();
      doms[e].ppline = text[t0 + 3].c_str();
      doms[e].csline = pli->d_has_cs ? text[t0 + 4].c_str() : nullptr;
      doms[e].rfline = pli->d_has_rf ? text[t0 + 5].c_str() : nullptr;
    }
  memset(&lst, 0, sizeof(lst));
  lst.hit = hits.data(); lst.N = (int) hits.size();
  lst.qname = rep->qname;
  lst.qacc = (rep->qacc && rep->qacc[0]) ? rep->qacc : nullptr;
  lst.M = rep->M;
  lst.Z = rep->Z; lst.domZ = domZ;
  lst.textw = rep->textw; lst.show_accessions = rep->show_accessions; lst.show_alignments = ali;
  return p7G_OK;
}

// the four writers share their arguments: which = 0 targets, 1 domains,
// 2 tabular targets, 3 tabular domains
static int
report_text(const P7_GPU_PIPELINE *pli, const P7_GPU_REPORT *rep, int which, int show_header, char *buf, size_t cap,
            size_t *ret_len)
{
  if (! pli) { p7g_set_error("no pipeline"); return p7G_EINVAL; }
  std::lock_guard<std::mutex> rl(pli->regs_mu);
  std::vector<P7G_RHIT> hits;
  std::vector<P7G_RDOM> doms;
  std::vector<std::string> text;
  P7G_RLIST lst;
  int st;
  if ((st = report_list(pli, rep, hits, doms, lst, text)) != p7G_OK) return st;
  size_t n = 0;
  switch (which) {
  case 0: n = p7g_targets_text(&lst, buf, cap); break;
  case 1: n = p7g_domains_text(&lst, buf, cap); break;
  case 2: n = p7g_tabular_targets(&lst, show_header, buf, cap); break;
  default: n = p7g_tabular_domains(&lst, show_header, buf, cap); break;
  }
  if (ret_len) *ret_len = n;
  return p7G_OK;
}

extern "C" int
p7_gpu_pipeline_TargetsText(const P7_GPU_PIPELINE *pli, const P7_GPU_REPORT *rep, char *buf, size_t cap, size_t *ret_len)
{ return report_text(pli, rep, 0, 0, buf, cap, ret_len); }

extern "C" int
p7_gpu_pipeline_DomainsText(const P7_GPU_PIPELINE *pli, const P7_GPU_REPORT *rep, char *buf, size_t cap, size_t *ret_len)
{ return report_text(pli, rep, 1, 0, buf, cap, ret_len); }

extern "C" int
p7_gpu_pipeline_TabularTargets(const P7_GPU_PIPELINE *pli, const P7_GPU_REPORT *rep, int show_header, char *buf,
                               size_t cap, size_t *ret_len)
{ return report_text(pli, rep, 2, show_header, buf, cap, ret_len); }

extern "C" int
p7_gpu_pipeline_TabularDomains(const P7_GPU_PIPELINE *pli, const P7_GPU_REPORT *rep, int show_header, char *buf,
                               size_t cap, size_t *ret_len)
{ return report_text(pli, rep, 3, show_header, buf, cap, ret_len); }

/*****************************************************************
 * Batches: many profiles against one database in one call
 *****************************************************************/
// Concurrent streams of a batch launched directly. Each chain takes one
// profile's stage chain at a time; with 16 chains the chains of short-target
// profiles keep the device busy while others run long-target tails, provided
// the process maps its streams onto enough hardware queues (GPU_MAX_HW_QUEUES:
// 17.2k GCUPS at 32 queues on the 64-profile batch against 7.9k at 4).
#ifndef P7G_BATCH_CHAINS
#define P7G_BATCH_CHAINS 16
#endif

struct Batch {
  Work     ch[P7G_BATCH_CHAINS];   // each chain's stream and work area
  char    *mem = nullptr;
  int      nprof_cap = 0;
  int64_t  n_cap = 0;
  unsigned mask_cap = 0;
  uint16_t *zmax = nullptr;              // [row][n] ungapped-pass results, rows in shape order
  int32_t  *queues = nullptr;            // [row] ungapped-pass work-queue cursors, then one per chunk of each row
  unsigned long long *sums = nullptr;    // [profile][8] launch_stats sums
  int32_t  *counts = nullptr;            // [profile][16] list counts at the end of the run
  uint8_t  *o_stage = nullptr;           // [profile][n] outputs in the caller's order, as asked for
  float    *o_f[6] = { nullptr, nullptr, nullptr, nullptr, nullptr, nullptr };
  std::vector<hipEvent_t> evs;           // per shape (its launch done), per chain (join), then the fork
  hipEvent_t t[2];
  bool     t_ready = false;
  // captured form: one graph per profile row (its stage chain), launched on
  // the row's chain stream after its shape's event
  std::vector<hipGraph_t>     graphs;
  std::vector<hipGraphExec_t> execs;
  std::vector<unsigned long long> key;   // what the instantiated graphs were captured for
  int      last_graph = 0;
};

struct BatchItem {                 // one profile of a batch, in shape order
  int prof;                        // index in the caller's array
  const P7_GPU_OPROFILE *om;
  DevProfile *p;
  int v, grab, shape;
  GateThresh g;
};

static void
batch_drop_graph(Batch &b)
{
  for (hipGraphExec_t e : b.execs) if (e) (void) hipGraphExecDestroy(e);
  for (hipGraph_t g : b.graphs) if (g) (void) hipGraphDestroy(g);
  b.execs.clear(); b.graphs.clear(); b.key.clear();
}

static void
batch_free(Batch *b)
{
  if (! b) return;
  batch_drop_graph(*b);
  for (int k = 0; k < P7G_BATCH_CHAINS; k++) {
    work_release(b->ch[k]);
    if (b->ch[k].st) (void) hipStreamDestroy(b->ch[k].st);
    if (b->ch[k].st2) (void) hipStreamDestroy(b->ch[k].st2);
    if (b->ch[k].md_thr) (void) hipEventDestroy(b->ch[k].md_thr);
  }
  if (b->mem) (void) hipFree(b->mem);
  for (hipEvent_t e : b->evs) (void) hipEventDestroy(e);
  if (b->t_ready) { (void) hipEventDestroy(b->t[0]); (void) hipEventDestroy(b->t[1]); }
  delete b;
}

// Output arrays a call asks for: bit 0 stage, bits 1..6 nullsc, usc,
// filtersc, vfsc, fwdsc, bcksc.
static unsigned
hits_mask(const P7_GPU_HITS *hits, int nprof)
{
  unsigned m = 0;
  for (int j = 0; hits && j < nprof; j++) {
    const P7_GPU_HITS &h = hits[j];
    if (h.stage) m |= 1u;
    const void *f[6] = { h.nullsc, h.usc, h.filtersc, h.vfsc, h.fwdsc, h.bcksc };
    for (int a = 0; a < 6; a++) if (f[a]) m |= 2u << a;
  }
  return m;
}

static int
batch_reserve(Batch &b, int nprof, int64_t n, unsigned mask)
{
  if (nprof <= b.nprof_cap && n <= b.n_cap && (mask & ~b.mask_cap) == 0) return p7G_OK;
  batch_drop_graph(b);
  if (b.mem) { (void) hipFree(b.mem); b.mem = nullptr; }
  const int64_t np = std::max(nprof, b.nprof_cap), nn = std::max(n, b.n_cap);
  mask |= b.mask_cap;
  size_t o = 0;
  auto take = [&](size_t bytes) { size_t at = o; o += (bytes + 255) & ~(size_t) 255; return at; };
  const size_t ozm = take(2 * (size_t) (np * nn)), oq = take(4 * (1 + P7G_SSV_MAXCHUNK) * (size_t) np),
               osu = take(64 * (size_t) np),
               oco = take(64 * (size_t) np), ost = (mask & 1u) ? take((size_t) (np * nn)) : 0;
  size_t of[6];
  for (int a = 0; a < 6; a++) of[a] = (mask & (2u << a)) ? take(4 * (size_t) (np * nn)) : 0;
  char *m = nullptr;
  HIPCHK(hipMalloc((void **) &m, o));
  b.mem = m; b.nprof_cap = (int) np; b.n_cap = nn; b.mask_cap = mask;
  b.zmax = (uint16_t *) (m + ozm); b.queues = (int32_t *) (m + oq);
  b.sums = (unsigned long long *) (m + osu); b.counts = (int32_t *) (m + oco);
  b.o_stage = (mask & 1u) ? (uint8_t *) (m + ost) : nullptr;
  for (int a = 0; a < 6; a++) b.o_f[a] = (mask & (2u << a)) ? (float *) (m + of[a]) : nullptr;
  return p7G_OK;
}

// Sets aside what p7_gpu_PipelineBatch needs for batches of up to nprof
// profiles over this database, with every output: the chains' streams and
// work areas and the batch's arrays (the server does this for its slots, so
// that no query waits for their allocation).
extern "C" int
p7g_pipeline_ReserveBatch(P7_GPU_PIPELINE *pli, const P7_GPU_SEQDB *db, int nprof)
{
  std::lock_guard<std::mutex> lk(pli->mu);
  int status;
  if ((status = ensure_init()) != p7G_OK) return status;
  if (! pli->b) { pli->b = new (std::nothrow) Batch(); if (! pli->b) return p7G_EMEM; }
  Batch &b = *pli->b;
  const int nchain = std::min(nprof, P7G_BATCH_CHAINS);
  for (int k = 0; k < nchain; k++) {
    Work &w = b.ch[k];
    if (! w.st) HIPCHK(hipStreamCreateWithFlags(&w.st, hipStreamNonBlocking));
    if ((status = work_reserve(w, db->nseq, pipeline_pool_floats(db->nseq, db->nres))) != p7G_OK) return status;
  }
  return batch_reserve(b, nprof, db->nseq, 0x7fu);
}

// One profile row's work after its shape's ungapped pass, on its chain's
// stream: the stage chain, its counters and the outputs asked for.
static int
batch_row(Batch &b, Work &w, const BatchItem &x, const DevSeqs &s, int64_t n, unsigned mask, int r)
{
  const int j = x.prof;
  int status;
  if ((status = chain_reset(w, n)) != p7G_OK) return status;
  if ((status = chain_stages(w, *x.p, s, n, b.zmax + (size_t) r * n, x.g, false)) != p7G_OK) return status;
  HIPCHK(launch_fill32(b.sums + 8 * (size_t) j, 0u, 16, w.st));
  HIPCHK(launch_stats(s, w.r.stage, n, w.l.vit, w.l.fwd, w.l.counts, b.sums + 8 * (size_t) j, w.st));
  HIPCHK(launch_copy32(b.counts + 16 * (size_t) j, w.l.counts, 16, w.st));
  if (mask & 1u) HIPCHK(launch_scatter8(s, w.r.stage, b.o_stage + (size_t) j * n, n, w.st));
  const float *src[6] = { w.r.nullsc, w.r.usc, w.r.filtersc, w.r.vfsc, w.r.fwdsc, w.r.bcksc };
  for (int a = 0; a < 6; a++)
    if (mask & (2u << a)) HIPCHK(launch_scatter32(s, src[a], b.o_f[a] + (size_t) j * n, n, w.st));
  return p7G_OK;
}

// Captures every row's batch_row as its own graph, on the row's chain stream
// (row r on chain r mod nchain, as batch_enqueue runs it).
static int
batch_capture(Batch &b, int nchain, const std::vector<BatchItem> &it, const DevSeqs &s, int64_t n, unsigned mask)
{
  batch_drop_graph(b);
  const int nit = (int) it.size();
  b.graphs.assign((size_t) nit, nullptr);
  b.execs.assign((size_t) nit, nullptr);
  for (int r = 0; r < nit; r++) {
    Work &w = b.ch[r % nchain];
    HIPCHK(hipStreamBeginCapture(w.st, hipStreamCaptureModeThreadLocal));
    const int status = batch_row(b, w, it[(size_t) r], s, n, mask, r);
    hipGraph_t gr = nullptr;
    const hipError_t ec = hipStreamEndCapture(w.st, &gr);
    if (status != p7G_OK) { if (gr) (void) hipGraphDestroy(gr); batch_drop_graph(b); return status; }
    if (ec != hipSuccess) { batch_drop_graph(b); HIPCHK(ec); }
    b.graphs[(size_t) r] = gr;
    const hipError_t ei = hipGraphInstantiateWithFlags(&b.execs[(size_t) r], gr, 0);
    if (ei != hipSuccess) { batch_drop_graph(b); HIPCHK(ei); }
  }
  return p7G_OK;
}

// Enqueues one batch on st0 and the chain streams: the ungapped pass of every
// profile, one persistent launch per kernel shape, one after another on st0
// (largest models first), each followed by its event; then every profile
// row's later stages on chain (row mod nchain), which waits only for its own
// shape's launch, launched directly or as the row's captured graph; then st0
// joins every chain. In sequence, each shape's launch has the whole device
// and the largest models' later stages, which hold the longest serial rows,
// start while the other shapes' passes run; launched side by side on the
// chain streams, every pass finished at about the same time and those rows
// became the batch's tail (701 against 655 ms on the 64-profile batch).
static int
batch_enqueue(Batch &b, hipStream_t st0, int nchain, const std::vector<BatchItem> &it, int nshape,
              const DevSeqs &s, int64_t n, unsigned mask, uint16_t *bnd, int64_t bnd_stride, bool graphs)
{
  const int nit = (int) it.size();
  HIPCHK(launch_fill32(b.queues, 0u, (1 + P7G_SSV_MAXCHUNK) * (int64_t) nit, st0));
  for (int r0 = 0; r0 < nit; ) {
    int r1 = r0;
    while (r1 < nit && it[r1].shape == it[r0].shape) r1++;
    hipStream_t sst = st0;
    if (it[r0].v < 0) {                  // a chunked model: its own launches (one row per shape)
      HIPCHK(launch_ssv_chunked(*it[r0].p, s, nullptr, n, b.zmax + (size_t) r0 * n,
                                b.queues + nit + P7G_SSV_MAXCHUNK * r0, it[r0].grab, bnd, bnd_stride, sst));
      HIPCHK(hipEventRecord(b.evs[it[r0].shape], sst));
      r0 = r1;
      continue;
    }
    std::vector<const int *> tbl;
    for (int r = r0; r < r1; r++) tbl.push_back(it[r].p->ssv[it[r].v].tbl);
    const SsvVariant &sv = it[r0].p->ssv[it[r0].v];
    HIPCHK(launch_ssv_multi(sv.G, sv.ND, tbl.data(), r1 - r0, s, n, b.zmax + (size_t) r0 * n, n, b.queues + r0,
                            it[r0].grab, sst));
    HIPCHK(hipEventRecord(b.evs[it[r0].shape], sst));
    r0 = r1;
  }
  int status;
  for (int r = 0; r < nit; r++) {
    Work &w = b.ch[r % nchain];
    HIPCHK(hipStreamWaitEvent(w.st, b.evs[it[r].shape], 0));
    if (graphs) HIPCHK(hipGraphLaunch(b.execs[(size_t) r], w.st));
    else if ((status = batch_row(b, w, it[(size_t) r], s, n, mask, r)) != p7G_OK) return status;
  }
  for (int k = 0; k < nchain; k++) {
    HIPCHK(hipEventRecord(b.evs[nshape + k], b.ch[k].st));
    HIPCHK(hipStreamWaitEvent(st0, b.evs[nshape + k], 0));
  }
  return p7G_OK;
}

extern "C" int
p7_gpu_PipelineBatch(P7_GPU_PIPELINE *pli, const P7_GPU_OPROFILE *const *oms, int nprof, const P7_GPU_SEQDB *db,
                     P7_GPU_HITS *hits, P7_GPU_STATS *stats, int flags)
{
  std::lock_guard<std::mutex> lk(pli->mu);
  int status;
  if (nprof <= 0 || ! oms || ! db) { p7g_set_error("empty batch"); return p7G_EINVAL; }
  if ((status = ensure_init()) != p7G_OK) return status;
  if (! pli->w.st) HIPCHK(hipStreamCreateWithFlags(&pli->w.st, hipStreamNonBlocking));
  if (! pli->b) { pli->b = new (std::nothrow) Batch(); if (! pli->b) return p7G_EMEM; }
  Batch &b = *pli->b;
  const int64_t n = db->nseq;
  hipStream_t st0 = pli->w.st;

  // profiles, their device copies and chosen shapes (all before any capture)
  std::vector<BatchItem> it((size_t) nprof);
  for (int j = 0; j < nprof; j++) {
    if (! oms[j]) { p7g_set_error("batch profile %d is NULL", j); return p7G_EINVAL; }
    if ((status = check_query(oms[j], db)) != p7G_OK) return status;
    BatchItem &x = it[j];
    x.prof = j; x.om = oms[j];
    if ((status = dev_profile(oms[j], db->Lmax, &x.p)) != p7G_OK) return status;
    ssv_choice(*x.p, &x.v, &x.grab);
    x.g = gates_of(pli, oms[j]);
  }
  // rows grouped by shape (G, ND, grab), longest shapes first, chunked models
  // (one shape each) before all; input order within a shape
  auto area = [](const BatchItem &a) -> long {
    return a.v < 0 ? (long) 1 << 40 : (long) a.p->ssv[a.v].G * a.p->ssv[a.v].ND;
  };
  auto lanes = [](const BatchItem &a) -> int { return a.v < 0 ? 64 : a.p->ssv[a.v].G; };
  std::stable_sort(it.begin(), it.end(), [&](const BatchItem &a, const BatchItem &c) {
    if (area(a) != area(c)) return area(a) > area(c);
    if (lanes(a) != lanes(c)) return lanes(a) > lanes(c);
    return a.grab > c.grab;
  });
  int nshape = 0;
  for (int r = 0; r < nprof; r++) {
    if (r > 0) {
      const bool same_shape = it[r].v >= 0 && it[r-1].v >= 0 &&
        it[r].p->ssv[it[r].v].G == it[r-1].p->ssv[it[r-1].v].G &&
        it[r].p->ssv[it[r].v].ND == it[r-1].p->ssv[it[r-1].v].ND && it[r].grab == it[r-1].grab;
      if (! same_shape) nshape++;
    }
    it[r].shape = nshape;
  }
  nshape++;
  bool any_chunked = false;
  for (const BatchItem &x : it) any_chunked |= (x.v < 0);
  const bool graph = (flags & P7_GPU_BATCH_GRAPH) != 0;
  const int nchain = std::min(nprof, P7G_BATCH_CHAINS);

  // work areas (a run that finds a scale-factor pool full grows the pools to
  // the rows of every target and runs again)
  const unsigned mask = hits_mask(hits, nprof);
  std::vector<unsigned long long> sums;
  std::vector<int32_t> counts;
  int64_t pool_floats = pipeline_pool_floats(n, db->nres);
  for (int attempt = 0; ; attempt++) {
  for (int k = 0; k < nchain; k++) {
    Work &w = b.ch[k];
    if (! w.st) HIPCHK(hipStreamCreateWithFlags(&w.st, hipStreamNonBlocking));
    if ((status = work_reserve(w, n, pool_floats)) != p7G_OK) return status;
  }
  if ((status = batch_reserve(b, nprof, n, mask)) != p7G_OK) return status;
  for (const BatchItem &x : it)
    if (x.v < 0 && (status = work_bnd(pli->w, *x.p, db->bytes)) != p7G_OK) return status;
  uint16_t *bnd = any_chunked ? pli->w.bnd : nullptr;
  while ((int) b.evs.size() < nshape + nchain + 1) {
    hipEvent_t e;
    HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    b.evs.push_back(e);
  }
  if (! b.t_ready) { HIPCHK(hipEventCreate(&b.t[0])); HIPCHK(hipEventCreate(&b.t[1])); b.t_ready = true; }
  DevSeqs s = seqs_of(db);

  // the graph is replayed while everything it was captured with is the same
  std::vector<unsigned long long> key = { (unsigned long long) (uintptr_t) b.mem, db->serial, (unsigned long long) n,
                                          (unsigned long long) nprof, (unsigned long long) nchain, mask,
                                          (unsigned long long) (uintptr_t) bnd, (unsigned long long) pli->w.bnd_cap };
  for (int k = 0; k < nchain; k++) {
    key.push_back((unsigned long long) (uintptr_t) b.ch[k].mem);
    key.push_back((unsigned long long) (uintptr_t) b.ch[k].pool);
  }
  for (const BatchItem &x : it) {
    const p7g_dev_profile_s *d = (const p7g_dev_profile_s *) ((const char *) x.p - offsetof(p7g_dev_profile_s, p));
    auto bits = [](float f) { uint32_t u; memcpy(&u, &f, 4); return (unsigned long long) u; };
    key.insert(key.end(), { (unsigned long long) x.prof, d->serial, (unsigned long long) x.v, (unsigned long long) x.grab,
                            bits(x.g.T1m), bits(x.g.T2m), bits(x.g.T2v), bits(x.g.T3f),
                            (unsigned long long) (x.g.do_bias * 2 + x.g.do_max) });
  }
  if (graph && (b.execs.size() != (size_t) nprof || key != b.key)) {
    if ((status = batch_capture(b, nchain, it, s, n, mask)) != p7G_OK) return status;
    b.key = key;
  }

  HIPCHK(hipEventRecord(b.t[0], st0));
  if ((status = batch_enqueue(b, st0, nchain, it, nshape, s, n, mask, bnd, pli->w.bnd_cap, graph)) != p7G_OK) return status;
  HIPCHK(hipEventRecord(b.t[1], st0));
  b.last_graph = graph ? 1 : 0;

  sums.assign((size_t) nprof * 8, 0);
  counts.assign((size_t) nprof * 16, 0);
  HIPCHK(hipMemcpyAsync(sums.data(), b.sums, sizeof(unsigned long long) * sums.size(), hipMemcpyDeviceToHost, st0));
  HIPCHK(hipMemcpyAsync(counts.data(), b.counts, sizeof(int32_t) * counts.size(), hipMemcpyDeviceToHost, st0));
  for (int j = 0; hits && j < nprof; j++) {
    if (hits[j].stage) HIPCHK(hipMemcpyAsync(hits[j].stage, b.o_stage + (size_t) j * n, (size_t) n, hipMemcpyDeviceToHost, st0));
    float *dst[6] = { hits[j].nullsc, hits[j].usc, hits[j].filtersc, hits[j].vfsc, hits[j].fwdsc, hits[j].bcksc };
    for (int a = 0; a < 6; a++)
      if (dst[a]) HIPCHK(hipMemcpyAsync(dst[a], b.o_f[a] + (size_t) j * n, sizeof(float) * (size_t) n, hipMemcpyDeviceToHost, st0));
  }
  HIPCHK(hipStreamSynchronize(st0));
  int full = -1;
  for (int j = 0; j < nprof && full < 0; j++) if (counts[(size_t) j * 16 + 6]) full = j;
  if (full < 0) break;
  if (attempt > 0 || pool_floats >= pool_rows_all(n, db->nres)) {
    p7g_set_error("Forward scale-factor pool exhausted (batch profile %d)", full);
    return p7G_EMEM;
  }
  pool_floats = pool_rows_all(n, db->nres);
  }
  float ms = 0.0f;
  HIPCHK(hipEventElapsedTime(&ms, b.t[0], b.t[1]));
  for (int j = 0; j < nprof; j++) {
    P7_GPU_STATS one;
    stats_of_run(one, &sums[(size_t) j * 8], &counts[(size_t) j * 16], oms[j], db, pli->do_biasfilter);
    one.ms_total = ms;
    if (stats) stats[j] = one;
    stats_accumulate(pli->st, one, j == 0);
  }
  pli->st.ms_total = ms;
  pli->st.ms_msv = pli->st.ms_bias = pli->st.ms_vit = pli->st.ms_fwd = pli->st.ms_bck = pli->st.ms_ssv = 0.0;
  return p7G_OK;
}

extern "C" int
p7_gpu_pipeline_BatchGraphed(const P7_GPU_PIPELINE *pli)
{
  return (pli && pli->b) ? pli->b->last_graph : 0;
}

/*****************************************************************
 * Tuner of the ungapped MSV pass
 *****************************************************************/
static int
arch_name(char *buf, size_t len)
{
  hipDeviceProp_t prop;
  HIPCHK(hipGetDeviceProperties(&prop, g_device >= 0 ? g_device : 0));
  snprintf(buf, len, "%s", prop.gcnArchName);
  for (char *c = buf; *c; c++) if (*c == ' ' || *c == '\n') *c = '_';
  return p7G_OK;
}

// Times the ungapped pass of variant v with queue grab `grab` over the sample;
// the best of three runs after one untimed run, in microseconds.
static int
tune_time(const DevProfile &p, int v, int grab, const DevSeqs &s, const int32_t *d_slots, int64_t ns, uint16_t *zmax,
          int32_t *queue, hipStream_t st, hipEvent_t e0, hipEvent_t e1, float *ret_us)
{
  float best = 1e30f;
  for (int rep = 0; rep < 4; rep++) {
    HIPCHK(launch_fill32(queue, 0u, 1, st));
    HIPCHK(hipEventRecord(e0, st));
    HIPCHK(launch_ssv(p, v, s, d_slots, ns, zmax, queue, grab, st));
    HIPCHK(hipEventRecord(e1, st));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0.0f;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    if (rep > 0) best = std::min(best, ms * 1000.0f);
  }
  *ret_us = best;
  return p7G_OK;
}

struct TuneScratch {
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  void *mem = nullptr;
  ~TuneScratch()
  {
    if (st) (void) hipStreamSynchronize(st);
    if (mem) (void) hipFree(mem);
    if (e0) (void) hipEventDestroy(e0);
    if (e1) (void) hipEventDestroy(e1);
    if (st) (void) hipStreamDestroy(st);
  }
};

extern "C" int
p7_gpu_Autotune(const P7_GPU_OPROFILE *const *oms, int nprof, const P7_GPU_SEQDB *db)
{
  int status;
  if (nprof < 0 || (nprof > 0 && ! oms) || ! db) { p7g_set_error("bad tuner arguments"); return p7G_EINVAL; }
  if ((status = ensure_init()) != p7G_OK) return status;
  TuneScratch ts;
  HIPCHK(hipStreamCreateWithFlags(&ts.st, hipStreamNonBlocking));
  HIPCHK(hipEventCreate(&ts.e0));
  HIPCHK(hipEventCreate(&ts.e1));
  const int64_t n = db->nseq;
  char *m = nullptr;
  const size_t zbytes = ((2 * (size_t) n + 255) & ~(size_t) 255);
  HIPCHK(hipMalloc((void **) &m, zbytes + 256 + 4 * (size_t) n));
  ts.mem = m;
  uint16_t *zmax = (uint16_t *) m;
  int32_t *queue = (int32_t *) (m + zbytes);
  int32_t *d_slots = (int32_t *) (m + zbytes + 256);
  DevSeqs s = seqs_of(db);
  static const int grabs[] = { 1, 2, 4, 8 };
  std::set<int> done;
  std::vector<int32_t> sample;
  for (int j = 0; j < nprof; j++) {
    DevProfile *p;
    if (! oms[j]) { p7g_set_error("tuner profile %d is NULL", j); return p7G_EINVAL; }
    if ((status = check_query(oms[j], db)) != p7G_OK) return status;
    if ((status = dev_profile(oms[j], db->Lmax, &p)) != p7G_OK) return status;
    const int cls = tune_class(p->M);
    if (done.count(cls) || p->ssv_nchunk) continue;     // chunked models have one shape only
    // every k-th slot (slots run longest first, so the sample keeps the length
    // mix), k chosen for about 3e10 cells per timed run; without the targets
    // longer than a row segment, which a full pass splits into segments that
    // behave like targets of segment length (a listed sample is not split, and
    // their serial rows would rank the shapes by the wrong measure)
    const int64_t k = std::max<int64_t>(1, (int64_t) ((double) db->nres * p->M / 3e10));
    sample.clear();
    for (int64_t t = 0; t < n; t += k) if (db->h_len[t] <= P7G_SEG_ROWS) sample.push_back((int32_t) t);
    if (sample.empty()) sample.push_back((int32_t) (n - 1));
    const int64_t ns = (int64_t) sample.size();
    HIPCHK(hipMemcpyAsync(d_slots, sample.data(), sizeof(int32_t) * (size_t) ns, hipMemcpyHostToDevice, ts.st));
    // shapes that compute more than twice the positions of the tightest one are not tried
    int minpos = INT32_MAX;
    for (int v = 0; v < P7G_SSV_NVAR; v++) if (p->ssv[v].G) minpos = std::min(minpos, 2 * p->ssv[v].G * p->ssv[v].ND);
    TuneEntry e;
    e.v = p->ssv_def; e.grab = SSV_GRAB_DEFAULT;
    int vgrab[P7G_SSV_NVAR];
    for (int v = 0; v < P7G_SSV_NVAR; v++) {
      e.us[v] = -1.0f; vgrab[v] = SSV_GRAB_DEFAULT;
      if (! p->ssv[v].G || 2 * p->ssv[v].G * p->ssv[v].ND > 2 * minpos) continue;
      for (int gi = 0; gi < 4; gi++) {
        float us;
        if ((status = tune_time(*p, v, grabs[gi], s, d_slots, ns, zmax, queue, ts.st, ts.e0, ts.e1, &us)) != p7G_OK) return status;
        if (e.us[v] < 0.0f || us < e.us[v]) { e.us[v] = us; vgrab[v] = grabs[gi]; }
      }
    }
    // another shape replaces the built-in one only when it is clearly faster:
    // the sample runs last about a millisecond, and a full pass has been seen
    // to rank shapes within a few percent of each other the other way round
    int bv = -1;
    for (int v = 0; v < P7G_SSV_NVAR; v++)
      if (e.us[v] > 0.0f && (bv < 0 || e.us[v] < e.us[bv])) bv = v;
    if (bv >= 0) {
      const int dv = p->ssv_def;
      const bool keep = dv >= 0 && e.us[dv] > 0.0f && e.us[bv] >= TUNE_MARGIN * e.us[dv];
      e.v = keep ? dv : bv;
      e.grab = vgrab[e.v];
    }
    {
      std::lock_guard<std::mutex> lk(g_tunelock);
      g_tune[cls] = e;
    }
    done.insert(cls);
  }
  return p7G_OK;
}

extern "C" int
p7_gpu_tune_Get(int M, int *ret_lanes, int *ret_grab, float *ret_us)
{
  const int G[P7G_SSV_NVAR] = { 16, 32, 64, 8 };
  int lanes = (M <= 352) ? 8 : (M <= 1024) ? 16 : 64, grab = SSV_GRAB_DEFAULT, tuned = 0;
  if (ret_us) for (int v = 0; v < P7G_SSV_NVAR; v++) ret_us[v] = -1.0f;
  {
    std::lock_guard<std::mutex> lk(g_tunelock);
    auto f = g_tune.find(tune_class(M));
    if (M > 0 && f != g_tune.end()) {
      lanes = G[f->second.v]; grab = f->second.grab; tuned = 1;
      if (ret_us) for (int v = 0; v < P7G_SSV_NVAR; v++) ret_us[v] = f->second.us[v];
    }
  }
  if (ret_lanes) *ret_lanes = lanes;
  if (ret_grab) *ret_grab = grab;
  return tuned;
}

extern "C" void
p7_gpu_tune_Clear(void)
{
  std::lock_guard<std::mutex> lk(g_tunelock);
  g_tune.clear();
}

extern "C" int
p7_gpu_tune_Save(const char *path)
{
  int status;
  char arch[256];
  if ((status = ensure_init()) != p7G_OK) return status;
  if ((status = arch_name(arch, sizeof(arch))) != p7G_OK) return status;
  FILE *fp = fopen(path, "w");
  if (! fp) { p7g_set_error("cannot write %s", path); return p7G_EFORMAT; }
  fprintf(fp, "p7gpu-tune 2 %s\n", arch);
  {
    std::lock_guard<std::mutex> lk(g_tunelock);
    for (const auto &kv : g_tune)
      fprintf(fp, "%d %d %d %.3f %.3f %.3f %.3f\n", kv.first, kv.second.v, kv.second.grab,
              kv.second.us[0], kv.second.us[1], kv.second.us[2], kv.second.us[3]);
  }
  if (fclose(fp) != 0) { p7g_set_error("cannot write %s", path); return p7G_EFORMAT; }
  return p7G_OK;
}

extern "C" int
p7_gpu_tune_Load(const char *path)
{
  int status;
  char arch[256], farch[256];
  if ((status = ensure_init()) != p7G_OK) return status;
  if ((status = arch_name(arch, sizeof(arch))) != p7G_OK) return status;
  FILE *fp = fopen(path, "r");
  if (! fp) { p7g_set_error("cannot open %s", path); return p7G_ENOTFOUND; }
  int ver = 0;
  // version 1 files hold the times of three shapes (16, 32, 64 lanes), version 2 of four (and 8)
  if (fscanf(fp, "p7gpu-tune %d %255s", &ver, farch) != 2 || (ver != 1 && ver != 2)) {
    fclose(fp); p7g_set_error("%s is not a tuning file", path); return p7G_EFORMAT;
  }
  if (strcmp(arch, farch) != 0) {
    fclose(fp); p7g_set_error("%s was measured on %s, this device is %s", path, farch, arch); return p7G_EINCOMPAT;
  }
  std::map<int, TuneEntry> got;
  int cls, v, grab;
  float u[4] = { -1.0f, -1.0f, -1.0f, -1.0f };
  const int nfields = (ver == 1) ? 6 : 7;
  int nf;
  while ((nf = (ver == 1) ? fscanf(fp, "%d %d %d %f %f %f", &cls, &v, &grab, &u[0], &u[1], &u[2])
                          : fscanf(fp, "%d %d %d %f %f %f %f", &cls, &v, &grab, &u[0], &u[1], &u[2], &u[3])) == nfields) {
    if (cls < 1 || v < 0 || v >= nfields - 3 || (grab != 1 && grab != 2 && grab != 4 && grab != 8)) { nf = -2; break; }
    TuneEntry e; e.v = v; e.grab = grab;
    for (int i = 0; i < P7G_SSV_NVAR; i++) e.us[i] = u[i];
    got[cls] = e;
  }
  fclose(fp);
  if (nf != EOF) { p7g_set_error("%s: bad tuning line", path); return p7G_EFORMAT; }
  std::lock_guard<std::mutex> lk(g_tunelock);
  for (const auto &kv : got) g_tune[kv.first] = kv.second;
  return p7G_OK;
}

/*****************************************************************
 * Single-target drop-ins: one target staged in a scratch database
 *****************************************************************/
struct Single {
  uint8_t *res = nullptr; int64_t cap = 0;
  int64_t *off = nullptr; int32_t *len = nullptr;
  float   *xmx = nullptr; int64_t xcap = 0;     // two (L+1)*6 buffers: forward, backward
};
static Single g_one;

static int
single_stage(const uint8_t *dsq, int L, DevSeqs *s)
{
  Single &o = g_one;
  const int64_t need = (((int64_t) L + 15) & ~(int64_t) 15) + 16;
  if (need > o.cap) {
    if (o.res) (void) hipFree(o.res);
    o.res = nullptr;
    HIPCHK(hipMalloc((void **) &o.res, (size_t) need));
    o.cap = need;
  }
  if (! o.off) { HIPCHK(hipMalloc((void **) &o.off, 64)); HIPCHK(hipMalloc((void **) &o.len, 64)); }
  std::vector<uint8_t> h((size_t) need, P7G_PADCODE);
  for (int i = 0; i < L; i++) h[i] = dsq[i + 1];
  int64_t zero = 0;
  HIPCHK(hipMemcpyAsync(o.res, h.data(), (size_t) need, hipMemcpyHostToDevice, g_stream));
  HIPCHK(hipMemcpyAsync(o.off, &zero, sizeof(zero), hipMemcpyHostToDevice, g_stream));
  HIPCHK(hipMemcpyAsync(o.len, &L, sizeof(int32_t), hipMemcpyHostToDevice, g_stream));
  // the work buffers' stream-ordered use keeps `h` alive until the copies complete
  HIPCHK(hipStreamSynchronize(g_stream));
  s->res = o.res; s->off = o.off; s->len = o.len; s->nslot = 1;
  s->perm = nullptr; s->seg = nullptr; s->nseg = 0; s->nseg_slots = 0;
  return p7G_OK;
}

static int
check_dsq(const uint8_t *dsq, int L, const P7_GPU_OPROFILE *om)
{
  if (L < 0) { p7g_set_error("negative sequence length"); return p7G_EINVAL; }
  for (int i = 1; i <= L; i++) if (dsq[i] >= om->abc->Kp) { p7g_set_error("residue %d at position %d is not a digital residue", dsq[i], i); return p7G_EINVAL; }
  return p7G_OK;
}

static LenCfg
lencfg_of(const P7_GPU_OPROFILE *om)
{
  LenCfg lc;
  lc.on = 1;
  lc.tjb = om->tjb_b;
  lc.xmove = om->xw[p7G_XC][p7G_MOVE];
  lc.floop = om->xf[p7G_XC][p7G_LOOP];
  lc.fmove = om->xf[p7G_XC][p7G_MOVE];
  lc.p1 = 0.0f; lc.q1 = 0.0f;
  return lc;
}

static int
single_msv(const uint8_t *dsq, int L, const P7_GPU_OPROFILE *om, int *ret_xj)
{
  int status;
  DevProfile *p;
  if ((status = dev_profile(om, L, &p)) != p7G_OK) return status;
  if ((status = work_reserve(g_work, 1, 0)) != p7G_OK) return status;
  DevSeqs s;
  if ((status = single_stage(dsq, L, &s)) != p7G_OK) return status;
  Work &w = g_work;
  LenCfg lc = lencfg_of(om);
  if ((status = reset_counts(g_work)) != p7G_OK) return status;
  if ((status = work_bnd(w, *p, g_one.cap)) != p7G_OK) return status;
  HIPCHK(launch_ssv_chosen(*p, s, nullptr, 1, w.r.zmax, &w.l.counts[P7G_Q_SSV], w.bnd, w.bnd_cap, g_stream));
  HIPCHK(launch_msv_classify(*p, s, nullptr, 1, w.r.zmax, w.r.xj, w.l.full, &w.l.counts[0], lc, g_stream));
  HIPCHK(launch_compact_i16(w.r.xj, -2, 1, w.l.full, &w.l.counts[0], w.tilecnt, g_stream));
  HIPCHK(launch_msv_full(*p, s, w.l.full, &w.l.counts[0], 1, w.r.xj, lc, g_stream));
  int16_t xj; int32_t nfull;
  HIPCHK(hipMemcpyAsync(&xj, w.r.xj, sizeof(xj), hipMemcpyDeviceToHost, g_stream));
  HIPCHK(hipMemcpyAsync(&nfull, &w.l.counts[0], sizeof(nfull), hipMemcpyDeviceToHost, g_stream));
  HIPCHK(hipStreamSynchronize(g_stream));
  *ret_xj = (xj < 0) ? -1 : (nfull ? 1000 + xj : xj);
  return p7G_OK;
}

extern "C" int
p7_gpu_MSVFilter(const uint8_t *dsq, int L, const P7_GPU_OPROFILE *om, P7_GPU_OMX *ox, float *ret_sc)
{
  std::lock_guard<std::mutex> lk(g_lock);
  int status, v;
  if ((status = ensure_init()) != p7G_OK) return status;
  if ((status = check_dsq(dsq, L, om)) != p7G_OK) return status;
  if (ox) ox->M = om->M;
  if (L == 0) { *ret_sc = msv_score_from_xj(om, 0, om->tjb_b); return p7G_OK; }
  if ((status = single_msv(dsq, L, om, &v)) != p7G_OK) return status;
  if (v < 0) { *ret_sc = INFINITY; return p7G_ERANGE; }
  *ret_sc = msv_score_from_xj(om, v >= 1000 ? v - 1000 : v, om->tjb_b);
  return p7G_OK;
}

extern "C" int
p7_gpu_SSVFilter(const uint8_t *dsq, int L, const P7_GPU_OPROFILE *om, float *ret_sc)
{
  std::lock_guard<std::mutex> lk(g_lock);
  int status, v;
  if ((status = ensure_init()) != p7G_OK) return status;
  if ((status = check_dsq(dsq, L, om)) != p7G_OK) return status;
  if (L == 0) { *ret_sc = msv_score_from_xj(om, 0, om->tjb_b); return p7G_OK; }
  if ((status = single_msv(dsq, L, om, &v)) != p7G_OK) return status;
  if (v < 0) { *ret_sc = INFINITY; return p7G_ERANGE; }
  if (v >= 1000) return p7G_ENORESULT;      // the J state is needed: caller runs the full MSV filter
  *ret_sc = msv_score_from_xj(om, v, om->tjb_b);
  return p7G_OK;
}

extern "C" int
p7_gpu_ViterbiFilter(const uint8_t *dsq, int L, const P7_GPU_OPROFILE *om, P7_GPU_OMX *ox, float *ret_sc)
{
  std::lock_guard<std::mutex> lk(g_lock);
  int status;
  DevProfile *p;
  if ((status = ensure_init()) != p7G_OK) return status;
  if (om->mode != p7G_LOCAL && om->mode != p7G_UNILOCAL) { p7g_set_error("Fast filter only works for local alignment"); return p7G_EINVAL; }
  if ((status = check_dsq(dsq, L, om)) != p7G_OK) return status;
  if (ox) ox->M = om->M;
  if (L == 0) {                                // no rows: xC stays -infinity
    *ret_sc = -INFINITY;
    return p7G_OK;
  }
  if ((status = dev_profile(om, L, &p)) != p7G_OK) return status;
  if ((status = work_reserve(g_work, 1, 0)) != p7G_OK) return status;
  DevSeqs s;
  if ((status = single_stage(dsq, L, &s)) != p7G_OK) return status;
  Work &w = g_work;
  GateThresh g; memset(&g, 0, sizeof(g));
  if ((status = reset_counts(g_work)) != p7G_OK) return status;
  int32_t zero = 0, one = 1;
  HIPCHK(hipMemcpyAsync(w.l.vit, &zero, sizeof(zero), hipMemcpyHostToDevice, g_stream));
  HIPCHK(hipMemcpyAsync(&w.l.counts[2], &one, sizeof(one), hipMemcpyHostToDevice, g_stream));
  HIPCHK(launch_vit(*p, s, w.l.vit, &w.l.counts[2], 1, w.r, w.l, g, 0, lencfg_of(om), g_stream));
  float sc; int vst;
  HIPCHK(hipMemcpyAsync(&sc, w.r.vfsc, sizeof(sc), hipMemcpyDeviceToHost, g_stream));
  HIPCHK(hipMemcpyAsync(&vst, w.r.vstatus, sizeof(vst), hipMemcpyDeviceToHost, g_stream));
  HIPCHK(hipStreamSynchronize(g_stream));
  *ret_sc = sc;
  return vst == 16 ? p7G_ERANGE : p7G_OK;
}

static int
single_xmx(int L)
{
  Single &o = g_one;
  const int64_t need = (int64_t) (L + 1) * 6 * 2;
  if (need > o.xcap) {
    if (o.xmx) (void) hipFree(o.xmx);
    o.xmx = nullptr;
    HIPCHK(hipMalloc((void **) &o.xmx, sizeof(float) * (size_t) need));
    o.xcap = need;
  }
  return p7G_OK;
}

extern "C" int
p7_gpu_ForwardParser(const uint8_t *dsq, int L, const P7_GPU_OPROFILE *om, P7_GPU_OMX *fwd, float *opt_sc)
{
  std::lock_guard<std::mutex> lk(g_lock);
  int status;
  DevProfile *p;
  if ((status = ensure_init()) != p7G_OK) return status;
  if ((status = check_dsq(dsq, L, om)) != p7G_OK) return status;
  if (! fwd) { p7g_set_error("ForwardParser needs a matrix"); return p7G_EINVAL; }
  if ((status = p7_gpu_omx_GrowTo(fwd, om->M, 0, L)) != p7G_OK) return status;
  fwd->M = om->M; fwd->L = L; fwd->has_own_scales = 1; fwd->totscale = 0.0f;
  if (L == 0) {
    float *x = fwd->xmx;
    x[0] = 0.0f; x[1] = 1.0f; x[2] = 0.0f; x[3] = om->xf[p7G_XN][p7G_MOVE]; x[4] = 0.0f; x[5] = 1.0f;
    if (opt_sc) *opt_sc = (float) ((double) 0.0f + log((double) (0.0f * om->xf[p7G_XC][p7G_MOVE])));
    return p7G_OK;
  }
  if ((status = dev_profile(om, L, &p)) != p7G_OK) return status;
  if ((status = work_reserve(g_work, 1, 0)) != p7G_OK) return status;
  if ((status = single_xmx(L)) != p7G_OK) return status;
  DevSeqs s;
  if ((status = single_stage(dsq, L, &s)) != p7G_OK) return status;
  Work &w = g_work;
  GateThresh g; memset(&g, 0, sizeof(g));
  if ((status = reset_counts(g_work)) != p7G_OK) return status;
  int32_t zero = 0, one = 1;
  HIPCHK(hipMemcpyAsync(w.l.fwd, &zero, sizeof(zero), hipMemcpyHostToDevice, g_stream));
  HIPCHK(hipMemcpyAsync(&w.l.counts[3], &one, sizeof(one), hipMemcpyHostToDevice, g_stream));
  HIPCHK(launch_fwd(*p, s, w.l.fwd, &w.l.counts[3], 1, w.r, w.l, g, 0, lencfg_of(om), nullptr, 0, nullptr, g_one.xmx, g_stream));
  float sc; int fst;
  HIPCHK(hipMemcpyAsync(&sc, w.r.fwdsc, sizeof(sc), hipMemcpyDeviceToHost, g_stream));
  HIPCHK(hipMemcpyAsync(&fst, w.r.fstatus, sizeof(fst), hipMemcpyDeviceToHost, g_stream));
  HIPCHK(hipMemcpyAsync(fwd->xmx, g_one.xmx, sizeof(float) * (size_t) (L + 1) * 6, hipMemcpyDeviceToHost, g_stream));
  HIPCHK(hipStreamSynchronize(g_stream));
  double ts = 0.0;
  for (int i = 1; i <= L; i++) { float sf = fwd->xmx[i*6 + 5]; if (sf != 1.0f) ts = (float) (ts + log((double) sf)); }
  fwd->totscale = (float) ts;
  if (fst == 16) { p7g_set_error("forward score is NaN, zero or infinite"); return p7G_ERANGE; }
  if (opt_sc) *opt_sc = sc;
  return p7G_OK;
}

extern "C" int
p7_gpu_BackwardParser(const uint8_t *dsq, int L, const P7_GPU_OPROFILE *om, const P7_GPU_OMX *fwd, P7_GPU_OMX *bck, float *opt_sc)
{
  std::lock_guard<std::mutex> lk(g_lock);
  int status;
  DevProfile *p;
  if ((status = ensure_init()) != p7G_OK) return status;
  if ((status = check_dsq(dsq, L, om)) != p7G_OK) return status;
  if (! fwd || ! bck) { p7g_set_error("BackwardParser needs two matrices"); return p7G_EINVAL; }
  if (fwd->L != L) { p7g_set_error("fwd matrix size doesn't agree with length L"); return p7G_EINVAL; }
  if ((status = p7_gpu_omx_GrowTo(bck, om->M, 0, L)) != p7G_OK) return status;
  bck->M = om->M; bck->L = L; bck->has_own_scales = 0;
  if (L == 0) { p7g_set_error("backward score underflow on an empty sequence"); return p7G_ERANGE; }
  if ((status = dev_profile(om, L, &p)) != p7G_OK) return status;
  if ((status = work_reserve(g_work, 1, 0)) != p7G_OK) return status;
  if ((status = single_xmx(L)) != p7G_OK) return status;
  DevSeqs s;
  if ((status = single_stage(dsq, L, &s)) != p7G_OK) return status;
  Work &w = g_work;
  if ((status = reset_counts(g_work)) != p7G_OK) return status;
  int32_t zero = 0, one = 1;
  float *d_fx = g_one.xmx, *d_bx = g_one.xmx + (size_t) (L + 1) * 6;
  HIPCHK(hipMemcpyAsync(d_fx, fwd->xmx, sizeof(float) * (size_t) (L + 1) * 6, hipMemcpyHostToDevice, g_stream));
  HIPCHK(hipMemcpyAsync(w.l.bck, &zero, sizeof(zero), hipMemcpyHostToDevice, g_stream));
  HIPCHK(hipMemcpyAsync(&w.l.counts[4], &one, sizeof(one), hipMemcpyHostToDevice, g_stream));
  HIPCHK(launch_bck(*p, s, w.l.bck, &w.l.counts[4], 1, w.r, nullptr, nullptr, d_fx, d_bx, lencfg_of(om), g_stream));
  float sc; int fst;
  HIPCHK(hipMemcpyAsync(&sc, w.r.bcksc, sizeof(sc), hipMemcpyDeviceToHost, g_stream));
  HIPCHK(hipMemcpyAsync(&fst, w.r.fstatus, sizeof(fst), hipMemcpyDeviceToHost, g_stream));
  HIPCHK(hipMemcpyAsync(bck->xmx, d_bx, sizeof(float) * (size_t) (L + 1) * 6, hipMemcpyDeviceToHost, g_stream));
  HIPCHK(hipStreamSynchronize(g_stream));
  bck->has_own_scales = (fst & 0x100) ? 1 : 0;
  {
    double ts = log((double) bck->xmx[(size_t) L * 6 + 5]);
    float tsf = (float) ts;
    for (int i = L - 1; i >= 1; i--) { float sf = bck->xmx[(size_t) i * 6 + 5]; if (sf > 1.0f) tsf = (float) ((double) tsf + log((double) sf)); }
    bck->totscale = tsf;
  }
  if ((fst & 0xff) == 16) { p7g_set_error("backward score is NaN, zero or infinite"); return p7G_ERANGE; }
  if (opt_sc) *opt_sc = sc;
  return p7G_OK;
}

/*****************************************************************
 * Bias filter drop-in (p7_bg_FilterScore): the 2-state Forward on the GPU,
 * length terms on the host with the bg's own p1.
 *****************************************************************/
extern "C" int
p7_gpu_bg_FilterScore(P7_GPU_BG *bg, const uint8_t *dsq, int L, float *ret_sc)
{
  std::lock_guard<std::mutex> lk(g_lock);
  int status;
  if ((status = ensure_init()) != p7G_OK) return status;
  if (! bg->filter_set) { p7g_set_error("p7_gpu_bg_SetFilter() must be called first"); return p7G_EINVAL; }
  for (int i = 1; i <= L; i++) if (dsq[i] >= bg->abc->Kp) { p7g_set_error("bad residue code"); return p7G_EINVAL; }
  float logsc;
  if (L == 0) logsc = (float) log((double) bg->pi[2]);
  else {
    // a throwaway device profile carries only the filter HMM
    DevProfile p; memset(&p, 0, sizeof(p));
    std::vector<float> eo(64, 1.0f);
    for (int x = 0; x < bg->abc->Kp; x++) { eo[x*2] = bg->eo[x][0]; eo[x*2+1] = bg->eo[x][1]; }
    float *d_eo = nullptr;
    HIPCHK(hipMalloc((void **) &d_eo, sizeof(float) * 64));
    HIPCHK(hipMemcpyAsync(d_eo, eo.data(), sizeof(float) * 64, hipMemcpyHostToDevice, g_stream));
    p.eo = d_eo; p.ft1_0 = bg->ft[1][0]; p.ft1_1 = bg->ft[1][1]; p.pi0 = bg->pi[0]; p.pi1 = bg->pi[1];
    if ((status = work_reserve(g_work, 1, 0)) != p7G_OK) { (void) hipFree(d_eo); return status; }
    DevSeqs s;
    if ((status = single_stage(dsq, L, &s)) != p7G_OK) { (void) hipFree(d_eo); return status; }
    Work &w = g_work;
    GateThresh g; memset(&g, 0, sizeof(g));
    LenCfg lc; memset(&lc, 0, sizeof(lc)); lc.on = 1; lc.p1 = bg->ft[0][0]; lc.q1 = bg->ft[0][1];
    int32_t zero = 0, one = 1;
    HIPCHK(hipMemcpyAsync(w.l.bias, &zero, sizeof(zero), hipMemcpyHostToDevice, g_stream));
    HIPCHK(hipMemcpyAsync(&w.l.counts[1], &one, sizeof(one), hipMemcpyHostToDevice, g_stream));
    HIPCHK(launch_bias(p, s, w.l.bias, &w.l.counts[1], 1, w.r, w.l, g, 0, lc, nullptr, 0, nullptr, g_stream));
    HIPCHK(hipMemcpyAsync(&logsc, w.r.filtersc, sizeof(float), hipMemcpyDeviceToHost, g_stream));
    HIPCHK(hipStreamSynchronize(g_stream));
    (void) hipFree(d_eo);
  }
  *ret_sc = logsc + (float) L * logf(bg->p1) + logf(1.-bg->p1);
  return p7G_OK;
}

// region_trace_ensemble for one region (p7_domaindef.c:600-690, as
// p7_domaindef_ByPosteriorHeuristics calls it at :440-446): k_md_region over
// residues dsq[1..Lr] of a target of Ltarget residues.
extern "C" int
p7_gpu_RegionTraceEnsemble(const uint8_t *dsq, int Lr, int Ltarget, const P7_GPU_OPROFILE *om, int nsamples,
                           uint32_t seed, int32_t *ci, int32_t *cj, int cap, int32_t *ret_nc, float *n2sc,
                           float *opt_mx, float *opt_xmx, int32_t *opt_segs, int segcap, int32_t *ret_nseg)
{
  std::lock_guard<std::mutex> lk(g_lock);
  int status;
  DevProfile *p;
  if ((status = ensure_init()) != p7G_OK) return status;
  if (Lr < 1 || Ltarget < Lr || nsamples < 1 || ! ret_nc || ! n2sc) { p7g_set_error("RegionTraceEnsemble: bad arguments"); return p7G_EINVAL; }
  if ((status = check_dsq(dsq, Lr, om)) != p7G_OK) return status;
  if (segcap <= 0) segcap = MD_SEGCAP;
  if ((status = dev_profile(om, Ltarget, &p)) != p7G_OK) return status;
  if (md_lds_bytes(p->M, segcap, md_acc_cap(Lr)) > P7G_LDS_MAX) { p7g_set_error("model too long for multidomain regions (M=%d)", p->M); return p7G_EINVAL; }
  const int64_t Q = std::max(2, (p->M - 1) / 4 + 1);
  const int64_t area = md_area_floats(p->M, Lr, segcap);
  const size_t resb = ((size_t) Lr + 31) & ~(size_t) 15;
  char *m = nullptr;
  HIPCHK(hipMalloc((void **) &m, sizeof(float) * (size_t) area + resb + sizeof(MdJob) + sizeof(MdOut) + 512));
  float *pool = (float *) m;
  uint8_t *res = (uint8_t *) (pool + area);
  MdJob *job = (MdJob *) (res + resb);
  MdOut *out = (MdOut *) (job + 1);
  MdJob J;
  J.res_off = 0; J.area = 0; J.Lr = Lr; J.Ltarget = Ltarget; J.segcap = segcap; J.n2o = 0;
  MdOut O;
  hipError_t e = hipSuccess;
  if (e == hipSuccess) e = hipMemcpyAsync(res, dsq + 1, (size_t) Lr, hipMemcpyHostToDevice, g_stream);
  if (e == hipSuccess) e = hipMemcpyAsync(job, &J, sizeof(J), hipMemcpyHostToDevice, g_stream);
  if (e == hipSuccess) e = launch_md_regions(*p, res, job, out, 1, pool, segcap, nsamples, seed, Lr, g_stream);
  if (e == hipSuccess) e = hipMemcpyAsync(&O, out, sizeof(O), hipMemcpyDeviceToHost, g_stream);
  if (e == hipSuccess) e = hipMemcpyAsync(n2sc, pool + md_n2sc_at(p->M, Lr) + 1, sizeof(float) * (size_t) Lr, hipMemcpyDeviceToHost, g_stream);
  if (e == hipSuccess && opt_mx) e = hipMemcpyAsync(opt_mx, pool, sizeof(float) * (size_t) (Lr + 1) * 12 * Q, hipMemcpyDeviceToHost, g_stream);
  if (e == hipSuccess && opt_xmx) e = hipMemcpyAsync(opt_xmx, pool + (size_t) (Lr + 1) * 12 * Q, sizeof(float) * 6 * (size_t) (Lr + 1), hipMemcpyDeviceToHost, g_stream);
  if (e == hipSuccess) e = hipStreamSynchronize(g_stream);
  if (e == hipSuccess && opt_segs && O.nseg > 0) {
    e = hipMemcpy(opt_segs, pool + md_seg_at(p->M, Lr), sizeof(int32_t) * 5 * (size_t) std::min(O.nseg, segcap), hipMemcpyDeviceToHost);
  }
  (void) hipFree(m);
  if (e != hipSuccess) { p7g_set_error("HIP error %d (%s)", (int) e, hipGetErrorString(e)); return p7G_EDEVICE; }
  *ret_nc = O.nc;
  if (ret_nseg) *ret_nseg = O.nseg;
  for (int c = 0; c < std::min(O.nc, cap); c++) { if (ci) ci[c] = O.ci[c]; if (cj) cj[c] = O.cj[c]; }
  if (O.status == 1) { p7g_set_error("stochastic traceback failed"); return p7G_EINVAL; }
  if (O.status != 0) { p7g_set_error("region exceeds the sampled-domain capacity"); return p7G_EMEM; }
  return p7G_OK;
}
