// Device-side definitions shared by the POY DO alignment kernels (gfx950).
//
// Sequences are 5-bit nucleotide bitsets (A=1 C=2 G=4 T=8 gap=16) with the
// leading gap at index 0, as POY stores them (src/fasta.ml:148).
//
// Work decomposition used by every 2D fill kernel: one 64-lane wavefront per
// pair.  A pair's DP rows are cut into stripes of 64 rows; lane t owns row
// i0+t of the stripe and the wavefront sweeps the stripe's band columns as a
// systolic pipeline: at step s lane t evaluates column jlo+s-t, so the cell
// above (row i-1, same column) was produced by lane t-1 one step earlier and
// arrives through a single DPP wave_shr:1 move; the up-left cell is the
// value received on the previous step.  Lane 0 takes its "above" row from a
// boundary row written by lane 63 of the previous stripe (or generates POY's
// row 0).  Direction information is stored per stripe as
// [step/pack][lane] words, so every store of the wavefront is one coalesced
// 256-byte transaction.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace poyhip {

constexpr int kHigh = 1000000;  // HIGH_NUM, src/algn.c:37
constexpr int kGap = 16;
constexpr int kWave = 64;

// 2D flags (src/matrices.h:22-27)
constexpr int kAlign = 1, kInsert = 2, kDelete = 4;

// Cost model resident in device memory; staged into LDS by each kernel.
struct CmDev {
  int cost[1024];     // cost[(a<<5)+b] for 5-bit codes a, b
  int worst[1024];
  int prepend[32];
  int tail[32];
  uint8_t median[1024];
  int gap_open;
  int delta;          // smallest positive entry of the full 2*32*32 table
  int cost_model;     // 0 linear, 1 affine
  int combinations;
};

// Descriptor arrays of a batch (structure of arrays, device memory).  The
// DP rows always come from the shorter sequence: `swapped` says that the
// caller's second sequence is the shorter one.  `prio` is the swap flag of
// the reference's backtraces (insert/delete and horizontal/vertical tie
// order); batches set it equal to `swapped`, the per-pair drop-ins take it
// from the caller as the reference stubs do.
struct PairTable {
  const uint8_t* blob;    // all sequences
  const int64_t* off1;    // caller's first sequence
  const int64_t* off2;    // caller's second sequence
  const int* len_a;       // rows (shorter)
  const int* len_b;       // columns
  const int* swapped;
  const int* prio;
  const int* deltawh;     // linear: Align.cost_2 band hint
  const int64_t* dm_off;  // byte offset of the pair's direction store
  const int64_t* stab_off;// offset of the pair's stripe table (int64 units)
  const int64_t* out_off; // byte offset of the pair's output block
  const int* order;       // pairs of this launch, largest first
  int n;                  // entries in `order`
  int npairs;             // pairs in the tables (backtrace grid)
};

__device__ __forceinline__ int64_t rows_off(const PairTable& t, int pid) {
  return t.swapped[pid] ? t.off2[pid] : t.off1[pid];
}
__device__ __forceinline__ int64_t cols_off(const PairTable& t, int pid) {
  return t.swapped[pid] ? t.off1[pid] : t.off2[pid];
}

struct PairResult {
  int cost;
  int gap_num;
  int newk;       // band half-width of the final round (-1: no DP rows)
  int rounds;
  int full;       // linear: 1 when the full plane was filled; affine: 1 when
                  // the directions are gap-free nibbles (else bytes)
  int pad;
  long long cells;  // cells evaluated over all rounds
  long long slots;  // lane-steps issued by the fill (cells + idle lanes)
};

// Wavefront shift right by one lane; lane 0 receives `lane0`.
__device__ __forceinline__ int wave_shr1(int v, int lane0) {
  return __builtin_amdgcn_update_dpp(lane0, v, 0x138, 0xF, 0xF, false);
}

typedef unsigned long long lmask;  // one bit per lane of the wavefront

__device__ __forceinline__ lmask ballot(bool b) { return __builtin_amdgcn_ballot_w64(b); }

// acc * 2 + (this lane's bit of m): one add with the lane mask as carry-in.
__device__ __forceinline__ int shift_in(int acc, lmask m) {
  int r;
  asm("v_addc_co_u32_e64 %0, vcc, %1, %1, %2" : "=v"(r) : "v"(acc), "s"(m) : "vcc");
  return r;
}

// Next work item of a persistent wavefront.  Every lane issues the atomic
// (lane 0 adds 1, the others 0) so the grab sits in uniform control flow and
// the loop around it keeps a single header: a lane-0-only branch before the
// broadcast let the compiler split the loop so that lanes 1..63 never
// returned to the atomic.  Lane 0's old value is the index.
__device__ __forceinline__ int wave_grab(int* counter) {
  const int old = atomicAdd(counter, (threadIdx.x & 63) == 0 ? 1 : 0);
  return __builtin_amdgcn_readfirstlane(old);
}

__device__ __forceinline__ int imax(int a, int b) { return a > b ? a : b; }
__device__ __forceinline__ int imin(int a, int b) { return a < b ? a : b; }

// HAS_GAP_OPENING (src/algn.c:1240-1253) with gap_startNO = 0.
__device__ __forceinline__ int has_gap_opening(int idx, int prev, int curr, int go) {
  if (idx == 1 && (kGap & curr)) return 0;
  if (idx > 1 && !(kGap & prev) && (kGap & curr)) return 0;
  return go;
}

// choose_dir (src/algn.c:1594-1624): 0 vertical, 1 horizontal, 2 diagonal,
// 3 align.
// Flags are 0/1 ints; written with selects and bit operations only, so the
// uniform swap flag does not turn into a branch.
__device__ __forceinline__ int choose_dir(int has_diag, int has_vert, int has_hrzn, int sw) {
  const int base = has_vert ? 0 : (has_hrzn ? 1 : (has_diag ? 2 : 3));
  return base | (has_vert & has_hrzn & sw);
}

// This lane's direction word in global block q of a pair's store (one word
// per lane and block: 32-bit, or 16-bit with nib); 0 outside the store.
__device__ __forceinline__ uint32_t dm_word(const uint8_t* dm, long long cap, bool nib,
                                            long long q) {
  const int wlog = nib ? 1 : 2;
  const long long at = (q << (6 + wlog)) + ((long long)threadIdx.x << wlog);
  if (q < 0 || at + (1 << wlog) > cap) return 0u;
  return nib ? (uint32_t)*reinterpret_cast<const uint16_t*>(dm + at)
             : *reinterpret_cast<const uint32_t*>(dm + at);
}

// Exclusive prefix sum over the 64 lanes (Hillis-Steele with shuffles) and
// the total.
__device__ __forceinline__ int wave_excl_sum(int x, int& total) {
  int v = x;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int y = __shfl_up(v, d);
    if ((int)threadIdx.x >= d) v += y;
  }
  total = __builtin_amdgcn_readlane(v, 63);
  return v - x;
}

}  // namespace poyhip
