// Device helpers shared by the kernel files: packed int16 arithmetic, DPP
// lane moves and wave reductions, work queues, wave-uniform residue readers,
// and the host-side launch and occupancy helpers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <math.h>
#include <string.h>
#include <algorithm>
#include "p7g_device.h"

typedef short s2v __attribute__((ext_vector_type(2)));

#define NEG_PAIR 0x80008000u     // (-32768, -32768)
#define LOG2_D   0.69314718055994529
// Workgroup size of the one-target-per-wave list stages: 8 waves share one LDS
// copy of the profile.
#define LIST_BLOCK 512
// 16-byte loads the bias filter's ordered sum keeps in flight.

__device__ __forceinline__ int   s2i(s2v v) { return __builtin_bit_cast(int, v); }
__device__ __forceinline__ s2v   i2s(int v) { return __builtin_bit_cast(s2v, v); }
__device__ __forceinline__ int   adds2(int a, int b) { return s2i(__builtin_elementwise_add_sat(i2s(a), i2s(b))); }
__device__ __forceinline__ int   max2(int a, int b)  { return s2i(__builtin_elementwise_max(i2s(a), i2s(b))); }
__device__ __forceinline__ int   lo16(int v) { return (int)(short)(v & 0xffff); }
__device__ __forceinline__ int   hi16(int v) { return (int)(short)((unsigned)v >> 16); }
__device__ __forceinline__ int   pair16(int v) { return (int)(((unsigned)v & 0xffffu) | ((unsigned)v << 16)); }

// DPP lane moves. Lanes without a source receive `fill`.
__device__ __forceinline__ int wave_shr1(int v, int fill) { return __builtin_amdgcn_update_dpp(fill, v, 0x138, 0xf, 0xf, false); }
__device__ __forceinline__ int wave_shl1(int v, int fill) { return __builtin_amdgcn_update_dpp(fill, v, 0x130, 0xf, 0xf, false); }
__device__ __forceinline__ int row_shr1(int v, int fill)  { return __builtin_amdgcn_update_dpp(fill, v, 0x111, 0xf, 0xf, false); }
__device__ __forceinline__ float wave_shr1f(float v) { return __builtin_bit_cast(float, wave_shr1(__builtin_bit_cast(int, v), 0)); }
__device__ __forceinline__ float wave_shl1f(float v) { return __builtin_bit_cast(float, wave_shl1(__builtin_bit_cast(int, v), 0)); }

// Row shift of a packed block: the dword that becomes the new position pair 0.
template <int G>
__device__ __forceinline__ int shift_in(int last, int fill)
{
  int nb = (G == 16) ? row_shr1(last, fill) : wave_shr1(last, fill);
  return (int) __builtin_amdgcn_alignbit((unsigned) last, (unsigned) nb, 16);
}

// DPP lane move with an identity for lanes that have no source lane.
template <int CTRL, int RM = 0xf>
__device__ __forceinline__ int dppi(int old, int v) { return __builtin_amdgcn_update_dpp(old, v, CTRL, RM, 0xf, false); }
template <int CTRL, int RM = 0xf>
__device__ __forceinline__ float dppf(float old, float v)
{
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, old), __builtin_bit_cast(int, v), CTRL, RM, 0xf, false));
}
// The same move for an additive identity: lanes without a source lane read 0
// (bound_ctrl), which lets the compiler fold the move into the consuming
// add or multiply-add as its DPP operand.
template <int CTRL, int RM = 0xf>
__device__ __forceinline__ float dppz(float v)
{
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, RM, 0xf, true));
}
__device__ __forceinline__ float readlanef(float v, int l) { return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), l)); }
enum { ROW_SHR1 = 0x111, ROW_SHR2 = 0x112, ROW_SHR4 = 0x114, ROW_SHR8 = 0x118,
       ROW_SHL1 = 0x101, ROW_SHL2 = 0x102, ROW_SHL4 = 0x104, ROW_SHL8 = 0x108,
       WAVE_SHL1 = 0x130, WAVE_SHR1 = 0x138, ROW_BCAST15 = 0x142, ROW_BCAST31 = 0x143 };
#define INT_NEG ((int) 0x80000000)

// Wave reductions: inclusive scans within each 16-lane row, then row_bcast:15
// (rows 1 and 3 take lane 15 of the row before) and row_bcast:31 (rows 2 and 3
// take lane 31), so lane 63 holds the result. Lanes a move does not write keep
// the identity, which lets the compiler fuse each move into its ALU op.
__device__ __forceinline__ int wave_max_i(int v)
{
  v = max(v, dppi<ROW_SHR1>(INT_NEG, v));
  v = max(v, dppi<ROW_SHR2>(INT_NEG, v));
  v = max(v, dppi<ROW_SHR4>(INT_NEG, v));
  v = max(v, dppi<ROW_SHR8>(INT_NEG, v));
  v = max(v, dppi<ROW_BCAST15, 0xa>(INT_NEG, v));
  v = max(v, dppi<ROW_BCAST31, 0xc>(INT_NEG, v));
  return __builtin_amdgcn_readlane(v, 63);
}
__device__ __forceinline__ float wave_sum_f(float v)
{
  v += dppz<ROW_SHR1>(v);
  v += dppz<ROW_SHR2>(v);
  v += dppz<ROW_SHR4>(v);
  v += dppz<ROW_SHR8>(v);
  v += dppz<ROW_BCAST15, 0xa>(v);
  v += dppz<ROW_BCAST31, 0xc>(v);
  return readlanef(v, 63);
}
// The same moves leave every lane with the inclusive prefix sum over lanes 0..lane.
__device__ __forceinline__ float wave_scan_f(float v)
{
  v += dppz<ROW_SHR1>(v);
  v += dppz<ROW_SHR2>(v);
  v += dppz<ROW_SHR4>(v);
  v += dppz<ROW_SHR8>(v);
  v += dppz<ROW_BCAST15, 0xa>(v);
  v += dppz<ROW_BCAST31, 0xc>(v);
  return v;
}
// Inclusive scan over the lanes of the maps x -> max(C, x + W), composed in
// lane order (each lane's map applied after those of the lanes before), for
// W in {0, -infinity}: the sums are exact. Identity: (-infinity, 0).
__device__ __forceinline__ void wave_maxplus_scan(float &C, float &W)
{
#define P7G_MPS(CTRL, RM) { const float c1 = dppf<CTRL, RM>(-INFINITY, C), w1 = dppf<CTRL, RM>(0.0f, W); \
                            C = fmaxf(C, W + c1); W = W + w1; }
  P7G_MPS(ROW_SHR1, 0xf) P7G_MPS(ROW_SHR2, 0xf) P7G_MPS(ROW_SHR4, 0xf) P7G_MPS(ROW_SHR8, 0xf)
  P7G_MPS(ROW_BCAST15, 0xa) P7G_MPS(ROW_BCAST31, 0xc)
#undef P7G_MPS
}

__device__ __forceinline__ int   uni(int v)   { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ float unif(float v){ return __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, v))); }

__device__ __forceinline__ int64_t wave_id()  { return (int64_t) blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); }
__device__ __forceinline__ int64_t wave_cnt() { return (int64_t) gridDim.x * (blockDim.x >> 6); }

template <typename T>
__device__ __forceinline__ void stage_table(T *lds, const T *__restrict__ g, size_t n)
{
  for (size_t i = threadIdx.x; i < n; i += blockDim.x) lds[i] = g[i];
  __syncthreads();
}

// Appends `val` of every active lane with `pred` set to list[], with one
// atomic per wave; the order inside the list is unspecified.
__device__ __forceinline__ void list_append(bool pred, int32_t val, int32_t *__restrict__ list, int32_t *count)
{
  const unsigned long long m = __ballot(pred);
  if (m == 0ull) return;
  const int lane = __lane_id();
  const int leader = __ffsll((long long) m) - 1;
  int base = 0;
  if (lane == leader) base = atomicAdd(count, __popcll(m));
  base = __shfl(base, leader, 64);
  if (pred) list[base + __popcll(m & ((1ull << lane) - 1ull))] = val;
}

// Next item of a device work queue shared by the waves of one launch
// (wave-uniform). Lists are in slot order, longest target first, so waves
// taking items in queue order schedule the longest work first.
__device__ __forceinline__ int queue_next(int32_t *q, int step)
{
  int v = 0;
  if (__lane_id() == 0) v = atomicAdd(q, step);
  return uni(__shfl(v, 0, 64));
}

// Wave-uniform residue readers. A target's residues start on a 16-byte boundary
// and every target is followed by at least 16 readable bytes (its padding or
// the next target, and 16 spare bytes after the last one), so the chunk after
// the current one can always be fetched ahead. Residue i is 0-based.
__device__ __forceinline__ int chunk_byte(const uint4 &c, int i)
{
  const int q = (i >> 2) & 3;
  const unsigned w = (q == 0) ? c.x : (q == 1) ? c.y : (q == 2) ? c.z : c.w;
  return (int) ((w >> ((i & 3) * 8)) & 0xffu);
}
// The chunk in use is moved to scalar registers; only the one in flight
// occupies vector registers.
__device__ __forceinline__ uint4 uni4(const uint4 &v)
{
  return make_uint4((unsigned) uni((int) v.x), (unsigned) uni((int) v.y), (unsigned) uni((int) v.z), (unsigned) uni((int) v.w));
}
struct ResAsc {                       // i = 0, 1, 2, ...
  const uint4 *p; uint4 cur, nxt; int c;
  __device__ __forceinline__ void init(const uint8_t *sq) { p = (const uint4 *) sq; c = 0; cur = uni4(p[0]); nxt = p[1]; }
  __device__ __forceinline__ int at(int i)
  {
    if ((i >> 4) != c) { c = i >> 4; cur = uni4(nxt); nxt = p[c + 1]; }
    return chunk_byte(cur, i);
  }
};
struct ResDesc {                      // i = L-1, L-2, ..., 0
  const uint4 *p; uint4 cur, nxt; int c;
  __device__ __forceinline__ void init(const uint8_t *sq, int L)
  {
    p = (const uint4 *) sq; c = (L - 1) >> 4; cur = uni4(p[c]); nxt = p[c > 0 ? c - 1 : 0];
  }
  __device__ __forceinline__ int at(int i)
  {
    if ((i >> 4) != c) { c = i >> 4; cur = uni4(nxt); nxt = p[c > 0 ? c - 1 : 0]; }
    return chunk_byte(cur, i);
  }
};

/*****************************************************************
 * Host-side launch helpers
 *****************************************************************/
static int g_ncu = 0, g_lds_block = 0, g_lds_cu = 0;
static inline void dev_props()
{
  if (g_ncu) return;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) dev = 0;
  if (hipDeviceGetAttribute(&g_ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || g_ncu <= 0) g_ncu = 256;
  if (hipDeviceGetAttribute(&g_lds_block, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || g_lds_block <= 0)
    g_lds_block = 64 * 1024;
  if (hipDeviceGetAttribute(&g_lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev) != hipSuccess || g_lds_cu <= 0)
    g_lds_cu = g_lds_block;
}
static inline int ncu() { dev_props(); return g_ncu; }
static inline int lds_block_limit() { dev_props(); return g_lds_block; }

// Launch with dynamic LDS; raises the per-kernel limit when a table needs more than 64 KB.
static inline hipError_t launch(void *k, unsigned grid, unsigned block, void **args, size_t lds, hipStream_t st)
{
  if (lds > (size_t) lds_block_limit()) return hipErrorInvalidValue;
  if (lds > 64 * 1024) (void) hipFuncSetAttribute((const void *) k, hipFuncAttributeMaxDynamicSharedMemorySize, (int) lds);
  return hipLaunchKernel(k, dim3(grid), dim3(block), args, lds, st);
}

// Persistent grid: as many blocks of `wpb` waves as can be resident at once
// (registers, LDS and wave slots, from the occupancy calculator), never more
// than the work needs. A block that could not be resident would only start
// when another finishes and run its share alone at the end.
static inline unsigned grid_for_waves_blk(const void *kernel, int64_t nwaves, size_t lds_bytes, int wpb)
{
  dev_props();
  int per_cu = 0;
  if (lds_bytes > 64 * 1024) (void) hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int) lds_bytes);
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, wpb * 64, lds_bytes) != hipSuccess || per_cu < 1) {
    per_cu = 32 / wpb;
    if (lds_bytes > 0) { int byl = (int) ((size_t) g_lds_cu / lds_bytes); if (byl < per_cu) per_cu = byl; }
    if (per_cu < 1) per_cu = 1;
  }
  int64_t g = (int64_t) ncu() * per_cu;
  int64_t need = (nwaves + wpb - 1) / wpb;
  if (need < g) g = need;
  if (g < 1) g = 1;
  return (unsigned) g;
}
