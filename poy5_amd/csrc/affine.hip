// Banded affine-gap DO alignment (POY algn_CAML_align_affine_3,
// src/algn.c:2359) on gfx950: one wavefront per pair, systolic 64-row
// stripes, Ukkonen T-doubling inside the kernel, 7-bit directions packed in
// one byte per cell, followed by a per-pair backtrace that reproduces
// backtrace_aff (src/algn.c:1715) and its median lookups.
#include <type_traits>

#include "poyhip_device.h"
#include "poyhip_layout.h"

namespace poyhip {

// Direction byte of one affine cell, decided at fill time for the pair's
// swap flag:
//   bits 0-1  move taken from the "todo" mode   (choose_dir on DO_* flags)
//   bits 2-3  move after an align column        (choose_dir on ALIGN_TO_*)
//   bit 4     END_VERTICAL | HORIZONTAL_EQ_VERTICAL  (vertical run stops)
//   bit 5     END_HORIZONTAL | HORIZONTAL_EQ_VERTICAL (horizontal run stops)
//   bit 6     END_BLOCK                          (gap-gap block stops)
// Move codes: 0 vertical, 1 horizontal, 2 diagonal block, 3 align.

struct AffFillArgs {
  PairTable pt;
  const CmDev* cm;
  uint8_t* dm;
  int* ws;            // one workspace slot of ws_stride ints per wavefront
  long long ws_stride;
  long long* stab;    // stripe tables
  PairResult* res;
  int* counter;
  const uint8_t* cls;  // per-pair class (1: gap-free fast path), or null
  const int* ngen;     // number of general-class pairs, or null
  // output blocks and lengths (as AffBtArgs): when set, the gap-free
  // kernel's wavefronts run each pair's backtrace right after its fill
  uint8_t* out = nullptr;
  int* outlen = nullptr;
  int rotating = 1;  // 0: gap-free pairs also run the plain stripe loop
};

__device__ __forceinline__ void aff_backtrace_pair(PairTable pt, const uint8_t* dm_all,
                                                   const long long* stab_all, uint8_t* out,
                                                   int* outlen, const uint8_t* s_med, int pid,
                                                   int newk, bool nib);

// The backtrace of pair pid right after its fill, by the same wavefront; the
// median table goes into s_med (shared memory idle between pairs).
__device__ __forceinline__ void aff_fused_backtrace(PairTable pt, const CmDev* cm,
                                                    const uint8_t* dm_all,
                                                    const long long* stab_all, uint8_t* out,
                                                    int* outlen, uint8_t* s_med, int pid,
                                                    int newk, bool nib) {
  // the direction store's words come from this wavefront's lanes
  __threadfence_block();
  for (int k = threadIdx.x; k < 512; k += 64) s_med[k] = cm->median[k];
  __syncthreads();
  aff_backtrace_pair(pt, dm_all, stab_all, out, outlen, s_med, pid, newk, nib);
  __syncthreads();
}

// Reads `p[j]` for j in [0, n); other j read p[0] (the value is unused).
__device__ __forceinline__ int ldc(const uint8_t* p, int j, int n) {
  return p[(unsigned)j < (unsigned)n ? j : 0];
}

// Constants of one DP row (one position of the shorter sequence).
struct AffRow {
  int gap;       // row symbol carries the gap bit
  int isng;      // row symbol has no gap bit
  int ip_nogap;  // previous row symbol has no gap bit
  int ge;        // gap extension cost(ic, gap)
  int go;        // HAS_GAP_OPENING of the row
  int vext;      // vertical extension cost
  int cost_row;  // (ic & 15) << 5
};

struct AffCell {
  int V, C, H, E, G, f, gmax;
  uint32_t code;
};

// One cell of algn_newkk_fill_a_row_aff (src/algn.c:2113-2172): the four
// state values, ASSIGN_MINIMUM (1936), the u16 gap counts (algn_fill_gapnum,
// 127) and the direction byte.  Branch-free: every tie set is the set of
// sources equal to the minimum, which is what the reference's sequence of
// comparisons leaves.
__device__ __forceinline__ void aff_cell(const AffRow& r, int go, int sw, int j, int jc,
                                         int jprev, bool at_left, bool at_right,
                                         const int* s_cost, const int* s_prep, int uV, int uCB,
                                         int uG, int dV, int dCB, int dH, int dED, int dG,
                                         int lH, int lCB, int lG, AffCell& o) {
  const int gr = s_prep[jc];
  const int sj_go = has_gap_opening(j, jprev, jc, go);
  const bool jgap = (jc & kGap) != 0;
  const int shx = (j == 1) ? gr : (((jprev & kGap) && !jgap) ? sj_go + gr : gr);
  // horizontal (FILL_EXTEND_HORIZONTAL)
  const int extH = lH + shx, opnH = lCB + sj_go + gr;
  const int H = at_left ? kHigh : imin(extH, opnH);
  const int endH = at_left ? 0 : (extH >= opnH);
  // vertical (FILL_EXTEND_VERTICAL)
  const int extV = uV + r.vext, opnV = uCB + r.go + r.ge;
  const int V = at_right ? kHigh : imin(extV, opnV);
  const int endV = at_right ? 0 : (extV >= opnV);
  // gap-gap block (FILL_EXTEND_BLOCK_DIAGONAL)
  const bool flag = r.gap && jgap;
  const int jp_ed = at_left ? jc : jprev;
  const bool flag2 = r.ip_nogap && !(jp_ed & kGap);
  const int dg = flag ? 0 : kHigh;
  const int od = flag ? (flag2 ? 0 : 2 * go) : kHigh;
  const int extE = dED + dg, opnE = dCB + od;
  // align (FILL_CLOSE_BLOCK_DIAGONAL)
  const int d = s_cost[r.cost_row + (jc & 15)];
  const int fa = dCB + d;
  const int fv = dV + d + (r.isng ? 0 : sj_go);
  const int fh = dH + d + (jgap ? r.go : 0);
  const int fd = dED + d + imax(sj_go, r.go);
  const int cmin = imin(imin(fa, fv), imin(fh, fd));
  const bool j0 = (j == 0);
  const int E = j0 ? kHigh : imin(extE, opnE);
  const int endB = j0 ? 0 : (extE >= opnE);
  const int C = j0 ? kHigh : cmin;
  const int a2v = !j0 && fv == cmin, a2h = !j0 && fh == cmin, a2d = !j0 && fd == cmin;
  // ASSIGN_MINIMUM
  const int f = imin(imin(H, V), imin(E, C));
  const int dh = H == f, dv = V == f, dd = E == f, da = C == f;
  const int heqv = dh && (H == V);
  // algn_fill_gapnum: each set flag contributes its source, the row keeps
  // the maximum (u16 rows: operands < 2^16, results stored mod 2^16)
  const int pg1 = uG & 0xFFFF, pg2 = (int)((unsigned)uG >> 16);
  const int pg1l = dG & 0xFFFF, pg2l = (int)((unsigned)dG >> 16);
  const int g1l = lG & 0xFFFF, g2l = (int)((unsigned)lG >> 16);
  const int al = dd | da;
  const int g1 = imax(imax(al ? pg1l : 0, dh ? g1l + 1 : 0), dv ? pg1 : 0);
  const int g2 = imax(imax(al ? pg2l : 0, dh ? g2l : 0), dv ? pg2 + 1 : 0);
  o.G = (g1 & 0xFFFF) | ((g2 & 0xFFFF) << 16);
  o.gmax = imax(g1 & 0xFFFF, g2 & 0xFFFF);
  const int todo = choose_dir(dd, dv, dh, sw);
  const int amode = choose_dir(a2d, a2v, a2h, sw);
  o.code = (uint32_t)(todo | (amode << 2) | ((endV | heqv) << 4) | ((endH | heqv) << 5) |
                      (endB << 6));
  o.V = V;
  o.C = C;
  o.H = H;
  o.E = E;
  o.f = f;
}

// aff_cell for a pair whose symbols past the leading gap carry no gap bit:
// then HAS_GAP_OPENING is gap_open for every row and column past 0, the
// horizontal/vertical extensions are the plain gap costs, the gap-gap block
// state stays >= HIGH_NUM so it never reaches a minimum or a tie (the walk
// never enters it, so its END_BLOCK flag is not stored).  What remains is
// Gotoh's recurrence with POY's tie flags.
// The column's table entries come in: gr = prepend[jc] (horizontal gap
// cost of the column symbol) and d = cost(row symbol, jc).
__device__ __forceinline__ void aff_cell_fast_core(int ge, int go, int sw, int j, int gr, int d,
                                                   bool at_left, bool at_right, int uV, int uCB,
                                                   int uG, int dV, int dCB, int dH, int dG, int lH,
                                                   int lCB, int lG, AffCell& o) {
  const int hb = lCB + go;
  const int H = at_left ? kHigh : gr + imin(lH, hb);
  const int endH = at_left ? 0 : (lH >= hb);
  const int vb = uCB + go;
  const int V = at_right ? kHigh : ge + imin(uV, vb);
  const int endV = at_right ? 0 : (uV >= vb);
  const int m3 = imin(imin(dCB, dV), dH);
  const bool j0 = (j == 0);
  const int C = j0 ? kHigh : m3 + d;
  const int f = imin(imin(H, V), C);
  const int dh = H == f, dv = V == f, da = C == f;
  const int heqv = dh && (H == V);
  const int pg1 = uG & 0xFFFF, pg2 = (int)((unsigned)uG >> 16);
  const int pg1l = dG & 0xFFFF, pg2l = (int)((unsigned)dG >> 16);
  const int g1l = lG & 0xFFFF, g2l = (int)((unsigned)lG >> 16);
  const int g1 = imax(imax(da ? pg1l : 0, dh ? g1l + 1 : 0), dv ? pg1 : 0);
  const int g2 = imax(imax(da ? pg2l : 0, dh ? g2l : 0), dv ? pg2 + 1 : 0);
  o.G = (g1 & 0xFFFF) | ((g2 & 0xFFFF) << 16);
  o.gmax = imax(g1 & 0xFFFF, g2 & 0xFFFF);
  // gap-free direction nibble: todo move, vertical end, horizontal end (the
  // align-mode move equals the diagonal predecessor's todo move)
  const int todo = choose_dir(0, dv, dh, sw);
  o.code = (uint32_t)(todo | ((endV | heqv) << 2) | ((endH | heqv) << 3));
  o.V = V;
  o.C = C;
  o.H = H;
  o.E = kHigh;
  o.f = f;
}


__device__ __forceinline__ void aff_cell_fast(const AffRow& r, int go, int sw, int j, int jc,
                                              bool at_left, bool at_right, const int* s_cost,
                                              const int* s_prep, int uV, int uCB, int uG, int dV,
                                              int dCB, int dH, int dG, int lH, int lCB, int lG,
                                              AffCell& o) {
  aff_cell_fast_core(r.ge, go, sw, j, s_prep[jc], s_cost[r.cost_row + (jc & 15)], at_left,
                     at_right, uV, uCB, uG, dV, dCB, dH, dG, lH, lCB, lG, o);
}

// 1 when a pair can take aff_cell_fast: no symbol past the leading gap of
// either sequence carries the gap bit.
// One wavefront per pair, consecutive bytes across the lanes.  ngen (may be
// null) counts the pairs of the general class.
__global__ __launch_bounds__(64) void k_affine_classify(PairTable pt, uint8_t* cls, int* ngen) {
  const int pid = blockIdx.x;
  if (pid >= pt.npairs) return;
  const uint8_t* sa = pt.blob + rows_off(pt, pid);
  const uint8_t* sb = pt.blob + cols_off(pt, pid);
  const int la = pt.len_a[pid], lb = pt.len_b[pid];
  int any = 0;
#pragma unroll 4
  for (int k = 1 + (int)threadIdx.x; k < la; k += 64) any |= sa[k];
#pragma unroll 4
  for (int k = 1 + (int)threadIdx.x; k < lb; k += 64) any |= sb[k];
  const bool gap = __builtin_amdgcn_ballot_w64((any & kGap) != 0) != 0;
  if (threadIdx.x == 0) {
    cls[pid] = gap ? 0 : 1;
    if (gap && ngen != nullptr) atomicAdd(ngen, 1);
  }
}

// ---------------------------------------------------------------------------
// Rotating stripe schedule (gap-free pairs)
//
// With one wavefront per 64-row stripe, lane t works on row i0+t from step
// 2t to 2t+W-1 of a stripe that lasts W+126 steps (W = band width), so a
// wavefront idles on almost half its lane-steps for the band widths of
// typical pairs.  Here every lane moves on to its row of the next stripe as
// soon as the schedule allows: stripe b starts at global step S_b and lane t
// evaluates column j of its row at step S_b + (j - jlo_b) + t.  S_{b+1} is
// the smallest start that keeps (1) lane 0 of stripe b+1 behind lane 63 of
// stripe b on every column and (2) every lane's row of stripe b+1 after its
// row of stripe b.  Lanes 1..63 still take the row above from lane t-1
// through DPP; lane 0 takes it from a FIFO in LDS that lane 63 fills one
// stripe earlier (indexed by column mod kRotFifo).  Every lane loads its own
// column symbols: lanes of different stripes work on unrelated columns, so
// the symbol cannot travel down the wavefront as in the plain loop.
//
// The FIFO holds lag + 1 entries, lag ~ W - 123, so a band wider than about
// 560 columns outgrows it.  Such a round spills the FIFO to the wavefront's
// global workspace (one 16-byte entry per column): lane 63 still writes its
// four values per step to shared memory, lanes 0-3 copy the block's four
// entries to global memory at the end of the block, and lanes 0-4 load the
// entries lane 0 needs two blocks ahead into registers and park them in a
// 16-entry shared ring one block ahead, so the loads' latency hides behind a
// block of cells.  A round with more stripes than kRotMaxStripes runs the
// plain stripe loop instead.
//
// Direction words are laid out [global step / 4][lane] for the whole round
// and stab[b] = S_b, so the backtrace finds cell (row i0_b + t, column j) of
// stripe b at global step S_b + (j - jlo_b) + t, exactly as for the plain
// loop (whose S_b are the running sums of its stripe lengths).
constexpr int kRotFifo = 512;        // entries (power of two)
constexpr int kRotMaxStripes = 128;

struct RotShared {
  // V, CB + gap opening, the best state f and the horizontal gap count of
  // row i0_b - 1, one array each (lane 0 reads each component into its own
  // register)
  int fV[kRotFifo], fC[kRotFifo], fF[kRotFifo], fG[kRotFifo];
  int S[kRotMaxStripes + 1];         // stripe start steps
  int lag;                           // largest FIFO lag of the round
  int2 fin;                          // cost and gap counts of the last cell
};

// Band of row i in the round: [lo, hi], right edge rb (src/algn.c:2113-2172).
__device__ __forceinline__ int band_lo(int i, int newk) { return imax(i - newk, 0); }
__device__ __forceinline__ int band_hi(int i, int D, int newk, int lenj) {
  const int rb = i + D + newk;
  return rb <= lenj - 1 ? rb : lenj;
}

// Plans the round: S_b for every stripe.  Returns the total number of
// steps, or -1 when the round must use the plain stripe loop.
//   C1: S_{b+1} >= S_b + 71 + (jlo_{b+1} - jlo_b).  Lane t enters stripe b
//       at S_b + (lo(i0_b + t) - jlo_b) + t <= S_b + (jlo_{b+1} - jlo_b) + 63,
//       so every lane has entered stripe b at least 8 steps before S_{b+1}:
//       the entries into stripe b+1 are planned at the start of the 4-step
//       block before the one holding S_{b+1} (one block ahead, for the column
//       symbol loads), and lane 0 reads FIFO entries up front per block, at
//       least 4 steps after lane 63 wrote them.
//   C2: lane t's first cell of stripe b+1 comes at least 4 steps after its
//       last cell of stripe b, so it can take its new row at the 4-step
//       block boundary before that first cell.  Lane t+1 reads lane t's
//       outputs through DPP one step later, so it still sees lane t's last
//       cell of stripe b; at its own last cell (the band's right edge) it
//       ignores the row above.
// The shared FIFO is safe while lag = S_{b+1} - S_b - (jlo_{b+1} - jlo_b)
// - 63 stays within kRotFifo - 72 (no entry is overwritten before it is
// read); a larger lag spills it (lag_out).  The spilled FIFO needs lag >= 24:
// an entry reaches global memory at most 3 steps after lane 63 computes it
// and is loaded at most 12 steps before lane 0 reads it (C1 gives lag >= 8,
// the spill starts above 400).
// The gap counts are POY's u16 rows (DIRECTION_MATRIX); the rotating fill
// keeps the horizontal one in 32 bits, equal while lenj stays below 2^16.
__device__ int rot_plan(RotShared& sh, int leni, int lenj, int D, int newk, int nst,
                        long long dm_steps_cap, int& lag_out) {
  const int lane = threadIdx.x;
  if (nst > kRotMaxStripes || lenj > 65535) return -1;
  for (int b = lane; b < nst; b += 64) {
    const int i0 = 1 + (b << 6);
    const int jlo = band_lo(i0, newk);
    int d, lag = 0;
    if (b + 1 < nst) {
      const int jlon = band_lo(i0 + 64, newk);
      int c2 = 0;
      for (int t = 0; t < 64 && i0 + t + 64 <= leni; t++)
        c2 = imax(c2, band_hi(i0 + t, D, newk, lenj) - band_lo(i0 + t + 64, newk));
      c2 += (jlon - jlo) + 4;
      d = imax(71 + (jlon - jlo), c2);
      lag = d - (jlon - jlo) - 63;
    } else {
      d = 0;
      for (int t = 0; t < 64 && i0 + t <= leni; t++)
        d = imax(d, band_hi(i0 + t, D, newk, lenj) - jlo + t + 1);
    }
    sh.S[b + 1] = d;
    sh.fV[b] = lag;  // scratch: the FIFO is idle between rounds
  }
  __syncthreads();
  if (lane == 0) {
    int acc = 0, mlag = 0;
    sh.S[0] = 0;
    for (int b = 0; b < nst; b++) {
      acc += sh.S[b + 1];
      sh.S[b + 1] = acc;
      mlag = imax(mlag, sh.fV[b]);
    }
    sh.lag = mlag;
  }
  __syncthreads();
  const int total = __builtin_amdgcn_readfirstlane(sh.S[nst]);
  const int lag = __builtin_amdgcn_readfirstlane(sh.lag);
  __syncthreads();
  lag_out = lag;
  if ((long long)((total + 3) >> 2) * 4 > dm_steps_cap || total > (1 << 24)) return -1;
  return total;
}

// Buffer resources over a pair's column sequence and its direction store.
// Byte loads past the end of the sequence (negative offsets included, as
// unsigned) return 0 instead of faulting, so a lane outside its band loads
// without a clamp (the cells it evaluates there are never read).
// The pointer and length are wave-uniform; the broadcasts keep the resource
// in scalar registers.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t seq_rsrc(const uint8_t* p, int n) {
  const unsigned long long a = reinterpret_cast<unsigned long long>(p);
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)a);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(a >> 32));
  void* q = reinterpret_cast<void*>(((unsigned long long)hi << 32) | lo);
  return __builtin_amdgcn_make_buffer_rsrc(q, (short)0, __builtin_amdgcn_readfirstlane(n),
                                           0x00020000);
}
__device__ __forceinline__ int ld_sym(__amdgpu_buffer_rsrc_t r, int off) {
  return __builtin_amdgcn_raw_buffer_load_b8(r, off, 0, 0);
}

typedef unsigned short rot_u16x2 __attribute__((ext_vector_type(2)));
typedef unsigned rot_u32x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ int max_u16x2(int a, int b) {
  return __builtin_bit_cast(int, __builtin_elementwise_max(__builtin_bit_cast(rot_u16x2, a),
                                                            __builtin_bit_cast(rot_u16x2, b)));
}


// One round of the gap-free fill on the rotating schedule.  Same outputs as
// the plain stripe loop of k_affine_fill: cost and gap count of the last
// cell, cells and lane-steps, stripe table and direction nibbles.
//
// Every lane keeps its current row in step coordinates (column st + cbc,
// right edge at step rs, column 0 at step zs) and the next row as a pending
// record, planned for all lanes at once when lane 0 is about to enter a
// stripe.  A lane takes its pending row at the 4-step block boundary at or
// before the row's first cell (step fst): its previous row ended at least 4
// steps before fst (C2), so whole blocks see one row per lane and the
// block's table lookups and column loads need no per-step selection.  Cells
// a lane evaluates before fst lie left of its band and are never read.
//
// Direction nibbles: bits 0-1 the todo move (choose_dir), bit 2
// END_VERTICAL | HORIZONTAL_EQ_VERTICAL, bit 3 END_HORIZONTAL |
// HORIZONTAL_EQ_VERTICAL; the align-mode move of a cell is the todo move of
// its diagonal predecessor, where the backtrace reads it.  Layout: per 4-step
// block one 16-bit word per lane, step u of the block in bits 15-4u..12-4u.
//
// kG: the FIFO spilled to gfifo, the four components of columns -64..lenj
// at entry column + 64 (lane 63 runs from column -63 before its first row
// and lane 0 reads column -1 at most, so every offset stays non-negative;
// columns past lenj fall outside the resource).  Shared-memory slots then:
// 0-15 the ring of lane 0's entries by step, 16 lane 0's up-left entry at
// its next stripe, 32-35 lane 63's entries of the current block.
template <bool kSw, bool kG>
__device__ void rot_round_fast(RotShared& sh, int total, const uint8_t* __restrict__ sa,
                               const uint8_t* __restrict__ sb, int leni, int lenj, int D, int newk,
                               int nst, int go, int HA0, const int* s_cost, const int* s_ge,
                               const int* s_prep, uint8_t* dm, long long* stab, int* gfifo,
                               int& rc_out, int& rgn_out, long long& cells, long long& slots) {
  const int lane = threadIdx.x;
  const int kNever = 0x7fffffff;
  for (int b = lane; b <= nst; b += 64) stab[b] = sh.S[b];
  const __amdgpu_buffer_rsrc_t rsym = seq_rsrc(sb, lenj + 1);
  const __amdgpu_buffer_rsrc_t rdm = seq_rsrc(dm, ((total + 3) >> 2) << 7);
  // stores past column lenj are dropped, loads there return 0
  const __amdgpu_buffer_rsrc_t rgf =
      seq_rsrc(reinterpret_cast<const uint8_t*>(gfifo), (lenj + 65) << 4);
  rot_u32x4 pf = {0, 0, 0, 0};  // kG, lanes 0-4: entries loaded a block ago
  const int jlo_t = band_lo(1, newk);
  // ---- current row: stripe 0 -----------------------------------------------
  int symA = sa[imin(1 + lane, leni)];  // row symbols of even / odd stripes,
  int symB = nst > 1 ? sa[imin(65 + lane, leni)] : 0;  // one stripe ahead
  int cbc = jlo_t - lane, rs = kNever, zs = kNever, fst = kNever;
  int ge = s_ge[symA], crow = (symA & 15) << 4;
  {
    const int i = 1 + lane;
    if (i <= leni) {
      const int lo = band_lo(i, newk);
      fst = lo - cbc;
      rs = i + D + newk - cbc;
      zs = -cbc;
      cells += band_hi(i, D, newk, lenj) - lo + 1;
    }
  }
  // ---- pending row (none yet) ----------------------------------------------
  int nsw = kNever, ncbc = 0, nrs = kNever, nzs = kNever, nge = 0, ncrow = 0;
  // ---- lane 0's stripe (uniform) --------------------------------------------
  int b0 = 0, S0 = 0, jlo0 = jlo_t, planned = 0;
  int next0 = nst > 1 ? __builtin_amdgcn_readfirstlane(sh.S[1]) : kNever;
  int jlon0 = band_lo(65, newk);
  // lane 63's column offset and pending switch (uniform copies)
  int c63 = jlo_t - 63, n63sw = kNever, n63c = 0;
  const int blast = nst - 1;
  const int s_fin = __builtin_amdgcn_readfirstlane(sh.S[blast]) +
                    (lenj - band_lo(1 + (blast << 6), newk)) + ((leni - 1) & 63);
  const int src = (leni - 1) & 63;  // lane of the last row
  // up-left cell of lane 0's first step: row 0 at column jlo0 - 1
  // (its best state: CB 0, H HA0, V HIGH_NUM)
  int dF = jlo0 >= 1 ? imin(0, HA0) : kHigh, dG = 0;
  int oV = kHigh, oCBg = kHigh + go, oH = kHigh, oF = kHigh, oG = 0;
  int cb0run = 0;
  int c0prev = sb[imax(jlo0 - 1, 0)];
  slots += (long long)total * 64;
  const int total4 = (total + 3) & ~3;
  // lane 0's column at step st (at most one stripe change in 8 steps)
  auto col0 = [&](int st) { return st >= next0 ? jlon0 + (st - next0) : jlo0 + (st - S0); };
  // column symbols, loaded one 4-step block ahead: cn for lane 0's column
  // (row-0 generator of stripe 0), cl for the lane's own column (lanes of
  // different stripes work on unrelated columns)
  int cn[4], cl[4];
#pragma unroll
  for (int u = 0; u < 4; u++) {
    cn[u] = ld_sym(rsym, col0(u));
    cl[u] = ld_sym(rsym, u + cbc);
  }
  for (int s = 0; s < total4; s += 4) {
    // lane 63's FIFO writes of the previous block precede these reads in
    // program order; the barrier keeps the compiler from moving them
    __builtin_amdgcn_wave_barrier();
    // lane 0 enters stripe b0 + 1 within this block or the next: plan every
    // lane's entry into it (all lanes are in stripe b0 by now, C1)
    if (next0 < s + 8 && planned <= b0) {
      planned = b0 + 1;
      const int i = 1 + (planned << 6) + lane;
      const int sym = (planned & 1) ? symB : symA;
      nge = s_ge[sym];
      ncrow = (sym & 15) << 4;
      ncbc = jlon0 - next0 - lane;
      nsw = kNever;
      if (i <= leni) {
        const int lo = band_lo(i, newk);
        fst = lo - ncbc;
        nsw = fst & ~3;
        nrs = i + D + newk - ncbc;
        nzs = -ncbc;
        cells += band_hi(i, D, newk, lenj) - lo + 1;
      }
      n63sw = __builtin_amdgcn_readlane(nsw, 63);
      n63c = __builtin_amdgcn_readlane(ncbc, 63);
      // stripe planned + 1's row symbols replace those of stripe b0, whose
      // row constants every lane already holds
      if (planned + 1 < nst) {
        const int nsym = sa[imin(1 + ((planned + 1) << 6) + lane, leni)];
        if ((planned + 1) & 1) symB = nsym;
        else symA = nsym;
      }
    }
    // the lane takes its pending row at this block boundary
    if (s == nsw) {
      cbc = ncbc; rs = nrs; zs = nzs; ge = nge; crow = ncrow;
    }
    if (s == n63sw) c63 = n63c;
    int c0b[4], grb[4], cdb[4];
    int fvx[4], fvy[4], fvz[4], fvw[4];
    // table entries of the block's columns, looked up together.  Column
    // symbols are 1..15 past column 0 and 0 outside the sequence; column 0's
    // leading gap reads past its row (at most into the padding after row
    // 15), a value its cell never uses (its diagonal is HIGH_NUM).
#pragma unroll
    for (int u = 0; u < 4; u++) {
      c0b[u] = cn[u];
      grb[u] = s_prep[cl[u]];
      cdb[u] = s_cost[crow + cl[u]];
    }
    if (b0 == 0) {
#pragma unroll
      for (int u = 0; u < 4; u++) cn[u] = ld_sym(rsym, col0(s + 4 + u));
    }
    {
      const int cbn = (s + 4 >= nsw) ? ncbc : cbc;
#pragma unroll
      for (int u = 0; u < 4; u++) cl[u] = ld_sym(rsym, s + 4 + u + cbn);
    }
    int pack = 0;
    // One step.  The slow form also handles the uniform events of a block:
    // lane 0 entering a stripe, the round's last cell and lane 0's row-0
    // generator in stripe 0, and column-0 cells; quiet blocks have none.
    auto step = [&](const int u, auto slow_t) {
      constexpr bool kSlow = decltype(slow_t)::value;
      const int st = s + u;
      if constexpr (kSlow) {
        // lane 0 enters its next stripe
        if (st == next0) {
          b0++;
          S0 = next0;
          jlo0 = jlon0;
          next0 = b0 + 1 < nst ? __builtin_amdgcn_readfirstlane(sh.S[b0 + 1]) : kNever;
          jlon0 = band_lo(1 + ((b0 + 1) << 6), newk);
          // lane 0's up-left cell: row i0 - 1 at column jlo0 - 1; the other
          // lanes keep theirs
          int2 v = make_int2(kHigh, 0);
          if (jlo0 >= 1) {
            const int x = kG ? 16 : (jlo0 - 1) & (kRotFifo - 1);
            v = make_int2(sh.fF[x], sh.fG[x]);
          }
          if (lane == 0) {
            dF = v.x;
            dG = v.y;
          }
        }
      }
      // lane 0's source row: POY's row 0 in stripe 0, else the FIFO
      int sV = fvx[u], sCBg = fvy[u], sF = fvz[u], sG = fvw[u];
      if constexpr (kSlow) {
        const int c0 = c0b[u];
        if (b0 == 0) {
          const int jl0 = jlo0 + (st - S0);
          const int gr0 = s_prep[c0];
          const int gop0 = has_gap_opening(jl0, c0prev, c0, go);
          const int add0 = ((c0prev & kGap) && !(c0 & kGap)) ? gop0 + gr0 : gr0;
          cb0run = (jl0 == 0) ? 0 : cb0run + add0;
          // V HIGH_NUM, CB cb0run, H HA0 at column 0 and cb0run after it
          sV = kHigh;
          sCBg = cb0run + go;
          sF = (jl0 == 0) ? imin(0, HA0) : cb0run;
          sG = jl0;
        }
        c0prev = c0;
      }
      const int uV = wave_shr1(oV, sV);
      const int uCBg = wave_shr1(oCBg, sCBg);
      const int uF = wave_shr1(oF, sF);
      const int uG = wave_shr1(oG, sG);
      // the cell (aff_cell_fast_core); tie and end flags are lane masks,
      // combined on the scalar unit.  Flags the backtrace never reads are
      // left unmasked: column 0 (the walk stops there), END_VERTICAL at the
      // right edge and END_HORIZONTAL at the left edge (a run reaching them
      // would extend a HIGH_NUM state).
      //   Each lane passes down CB + gap opening (both uses add it) and the
      // cell's best state f, the diagonal's min(CB, V, H) one step later.
      const int gr = grb[u], d = cdb[u];
      const bool at_left = st == fst, at_right = st == rs;
      const bool j0 = kSlow && st == zs;  // quiet blocks hold no column-0 cell
      const int H = at_left ? kHigh : gr + imin(oH, oCBg);
      const int V = at_right ? kHigh : ge + imin(uV, uCBg);
      const int C = j0 ? kHigh : dF + d;
      const int CBg = C + go;
      const int f = imin(imin(H, V), C);
      const bool dh = H == f, dv = V == f, da = C == f;
      const lmask mdh = ballot(dh), mdv = ballot(dv);
      const lmask mheqv = mdh & ballot(H == V);
      // algn_fill_gapnum: horizontal moves over the tied predecessors.  Every
      // path into (i, j) has j - i more horizontal than vertical moves, so
      // the largest horizontal count also has the largest vertical one.
      const int G = imax(imax(da ? dG : 0, dh ? oG + 1 : 0), dv ? uG : 0);
      // direction nibble, high bit first: horizontal end, vertical end,
      // then choose_dir of the todo mode (bit 1: neither vertical nor
      // horizontal; bit 0: not vertical, or both with the swap flag)
      pack = shift_in(pack, ballot(oH >= oCBg) | mheqv);
      pack = shift_in(pack, ballot(uV >= uCBg) | mheqv);
      pack = shift_in(pack, ~(mdv | mdh));
      pack = shift_in(pack, kSw ? ~mdv | mdh : ~mdv);
      // lane 63 hands its row to lane 0 of the next stripe (columns outside
      // its band land in slots that are rewritten before they are read)
      if (lane == 63) {
        const int x = kG ? 32 + u : (st + c63) & (kRotFifo - 1);
        sh.fV[x] = V;
        sh.fC[x] = CBg;
        sh.fF[x] = f;
        sh.fG[x] = G;
      }
      // the last cell of the round: cost and horizontal gap count
      if constexpr (kSlow) {
        if (st == s_fin) {
          if (lane == src) sh.fin = make_int2(f, G);
        }
      }
      oV = V; oCBg = CBg; oH = H; oF = f; oG = G;
      dF = uF; dG = uG;
    };
    // lane 0's FIFO entries of the block, read at the start of each form:
    // one register per component, each the DPP destination of its up value
    // (lane 0 keeps it)
    auto read_fifo = [&]() {
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const int x = kG ? (s + u) & 15 : col0(s + u) & (kRotFifo - 1);
        fvx[u] = sh.fV[x];
        fvy[u] = sh.fC[x];
        fvz[u] = sh.fF[x];
        fvw[u] = sh.fG[x];
      }
    };
    // quiet: no uniform event and no lane at column 0 in this block
    const bool quiet = next0 >= s + 4 && (unsigned)(s_fin - s) >= 4u && b0 != 0 &&
                       ballot((unsigned)(zs - s) < 4u) == 0;
    if (quiet) {
      read_fifo();
      step(0, std::false_type{}); step(1, std::false_type{});
      step(2, std::false_type{}); step(3, std::false_type{});
    } else {
      read_fifo();
      step(0, std::true_type{}); step(1, std::true_type{});
      step(2, std::true_type{}); step(3, std::true_type{});
    }
    __builtin_amdgcn_raw_buffer_store_b16((unsigned short)pack, rdm, lane << 1, (s >> 2) << 7, 0);
    if constexpr (kG) {
      // lane 63's writes of this block precede the copies below
      __builtin_amdgcn_wave_barrier();
      if (lane < 5) {
        // lanes 0-3: this block's entries of lane 63 (column s + u + c63) to
        // global memory
        if (lane < 4) {
          const rot_u32x4 v = {(unsigned)sh.fV[32 + lane], (unsigned)sh.fC[32 + lane],
                               (unsigned)sh.fF[32 + lane], (unsigned)sh.fG[32 + lane]};
          __builtin_amdgcn_raw_buffer_store_b128(v, rgf, (s + lane + c63 + 64) << 4, 0, 0);
        }
        // the entries loaded at the end of the previous block: lane 0's
        // steps s + 4..s + 7, and lane 4 its up-left entry at the next stripe
        const int x = lane < 4 ? (s + 4 + lane) & 15 : 16;
        sh.fV[x] = (int)pf.x;
        sh.fC[x] = (int)pf.y;
        sh.fF[x] = (int)pf.z;
        sh.fG[x] = (int)pf.w;
        // load steps s + 8..s + 11 (at most one stripe change in 12 steps,
        // so col0 holds), device scope: lane 63's stores went to L2
        const int col = lane < 4 ? col0(s + 8 + lane) : jlon0 - 1;
        pf = __builtin_amdgcn_raw_buffer_load_b128(rgf, (col + 64) << 4, 0, 16);
      }
    }
  }
  __builtin_amdgcn_wave_barrier();
  const int2 fin = sh.fin;
  rc_out = fin.x;
  // the last cell (leni, lenj) has lenj - leni >= 0 more horizontal moves
  rgn_out = fin.y;
}

// At most 128 VGPRs: four wavefronts per SIMD (shared memory allows as many)
template <bool kDm, bool kFast>
__global__ __launch_bounds__(64, 4) void k_affine_fill(AffFillArgs A) {
  // gap-free pairs only index rows and columns 0..15 of the cost table,
  // kept 16 wide (the wavefront's shared memory stays within a sixteenth
  // of the CU's, for four wavefronts per SIMD), plus 16 padding entries
  constexpr int kCostN = kFast ? 256 : 1024;
  constexpr int kRowShift = kFast ? 4 : 5;
  __shared__ int s_cost[kCostN + 16];
  __shared__ int s_ge[32];
  __shared__ int s_prep[32];
  __shared__ RotShared s_rot;
  const int lane = threadIdx.x;
  // the general kernel has nothing to do when no pair carries a gap bit
  if (!kFast && A.ngen != nullptr && __builtin_amdgcn_readfirstlane(*A.ngen) == 0) return;
  for (int k = lane; k < kCostN; k += 64)
    s_cost[k] = A.cm->cost[((k >> kRowShift) << 5) + (k & ((1 << kRowShift) - 1))];
  if (lane < 16) s_cost[kCostN + lane] = 0;
  if (lane < 32) s_ge[lane] = A.cm->cost[(lane << 5) + kGap];
  if (lane < 32) s_prep[lane] = A.cm->prepend[lane];
  __syncthreads();
  const int go = A.cm->gap_open;
  const int delta = A.cm->delta;

  for (;;) {
    const int q = wave_grab(A.counter);
    if (q >= A.pt.n) break;
    const int pid = A.pt.order[q];
    if (A.cls != nullptr && A.cls[pid] != (kFast ? 1 : 0)) continue;
    const uint8_t* __restrict__ sa = A.pt.blob + rows_off(A.pt, pid);
    const uint8_t* __restrict__ sb = A.pt.blob + cols_off(A.pt, pid);
    const int La = A.pt.len_a[pid], Lb = A.pt.len_b[pid];
    const int sw = A.pt.prio[pid];
    const int leni = La - 1, lenj = Lb - 1, D = lenj - leni;
    const int W1 = lenj + 1;
    int* ws = A.ws + (long long)blockIdx.x * A.ws_stride;
    uint8_t* dm = A.dm + (kDm ? A.pt.dm_off[pid] : 0);
    int* EDA = ws;
    int* bnd = ws + aff_bnd_offset(Lb);
    long long* stab = A.stab + A.pt.stab_off[pid];

    if (leni == 0) {
      // No DP rows: the cost is final_cost_matrix[lenj] as set by
      // initialize_matrices_affine (src/algn.c:1866-1900).
      int sum = 0;
      for (int j = 1 + lane; j <= lenj; j += 64) sum += s_prep[sb[j]];
      for (int o = 32; o > 0; o >>= 1) sum += __shfl_xor(sum, o);
      int T = (D + 1) * delta, rounds = 0;
      for (;;) {
        const int p = (T - D) / 2;
        rounds++;
        const int newp = (2 * T - D) / 2;
        if (65535 < p || newp - lenj + 1 >= 0) break;
        T *= 2;
      }
      if (lane == 0) {
        PairResult r;
        r.cost = lenj == 0 ? 0 : go + sum;
        r.gap_num = 65535;
        r.newk = -1;
        r.rounds = rounds;
        r.full = 0;
        r.pad = 0;
        r.cells = 0;
        r.slots = 0;
        A.res[pid] = r;
      }
      if constexpr (kDm && kFast) {
        if (A.out != nullptr)
          aff_fused_backtrace(A.pt, A.cm, A.dm, A.stab, A.out, A.outlen,
                              reinterpret_cast<uint8_t*>(s_rot.fV), pid, -1, false);
      }
      continue;
    }

    for (int j = lane; j <= lenj; j += 64) EDA[j] = kHigh;
    int HA0 = go;
    __threadfence_block();

    int T = (D + 1) * delta;
    int rounds = 0, newk = 0, cost = 0, gap_num = 0;
    long long cells = 0, slots = 0;
    for (;;) {
      const int p = (T - D) / 2;
      newk = p >= leni ? leni - 1 : p;
      // A round whose stop test cannot pass leaves only its row-0 state
      // change: the last cell's gap count is at least D (every path makes D
      // more horizontal than vertical moves, and the count keeps the most
      // horizontal moves over tied paths, exact below 2^16 columns), so
      // gap_num < p needs p > D.  Gap-free pairs only: the general kernel's
      // rounds also leave its row-0 gap-gap buffer behind.
      if constexpr (kFast) {
        const int newp = (2 * T - D) / 2;
        if (D >= p && newp - lenj + 1 < 0 && lenj <= 65535) {
          rounds++;
          if (imin(leni, newk) >= 2) HA0 = kHigh;
          T *= 2;
          continue;
        }
      }
      const int nstripes = (leni + 63) >> 6;
      long long dmpos = 0;
      int rc = 0, rgn = 0;
      int rot_total = -1;
      if constexpr (kFast && kDm) {
        int lag = 0;
        if (A.rotating)
          rot_total =
              rot_plan(s_rot, leni, lenj, D, newk, nstripes, aff_dm_bytes(La, Lb) >> 5, lag);
        if (rot_total > 0) {
          // the global FIFO (Lb + 64 entries of 4 ints) reuses the stripe
          // loop's row buffers (16 * Lb + 16 ints; the spill needs lenj > 400)
          if (lag > kRotFifo - 72) {
            if (sw)
              rot_round_fast<true, true>(s_rot, rot_total, sa, sb, leni, lenj, D, newk, nstripes,
                                         go, HA0, s_cost, s_ge, s_prep, dm, stab, bnd, rc, rgn,
                                         cells, slots);
            else
              rot_round_fast<false, true>(s_rot, rot_total, sa, sb, leni, lenj, D, newk, nstripes,
                                          go, HA0, s_cost, s_ge, s_prep, dm, stab, bnd, rc, rgn,
                                          cells, slots);
          } else {
            if (sw)
              rot_round_fast<true, false>(s_rot, rot_total, sa, sb, leni, lenj, D, newk, nstripes,
                                          go, HA0, s_cost, s_ge, s_prep, dm, stab, bnd, rc, rgn,
                                          cells, slots);
            else
              rot_round_fast<false, false>(s_rot, rot_total, sa, sb, leni, lenj, D, newk,
                                           nstripes, go, HA0, s_cost, s_ge, s_prep, dm, stab, bnd,
                                           rc, rgn, cells, slots);
          }
        }
      }
      for (int b = 0; rot_total <= 0 && b < nstripes; b++) {
        const int i0 = 1 + (b << 6);
        const int ilast = imin(i0 + 63, leni);
        const int i = i0 + lane;
        const bool act = i <= leni;
        const int jlo = imax(i0 - newk, 0);
        const int jhi = (ilast + D + newk <= lenj - 1) ? ilast + D + newk : lenj;
        const int nsteps = (jhi - jlo + 1) + (ilast - i0);
        const int nsteps4 = (nsteps + 3) & ~3;
        // stripe start in steps (gap-free pairs: 32 direction bytes a step)
        if (kDm && lane == 0) stab[b] = dmpos >> (kFast ? 5 : 6);
        slots += (long long)nsteps * 64;

        // row constants of lane t (row i); rows past leni get the last row's
        // values and an empty band
        const int ir = imin(i, leni);
        const int ic = sa[ir], ip = sa[ir - 1];
        AffRow row;
        row.gap = ic & kGap;
        row.isng = (ic & kGap) == 0;
        row.ip_nogap = (ip & kGap) == 0;
        row.ge = s_ge[ic];
        row.go = has_gap_opening(ir, ip, ic, go);
        row.vext = (ir > 1 && (ip & kGap) && !(ic & kGap)) ? row.go + row.ge : row.ge;
        row.cost_row = (ic & 15) << kRowShift;
        const int my_lo = act ? imax(i - newk, 0) : (1 << 30);
        const int my_rb = i + D + newk;
        const int my_hi = act ? ((my_rb <= lenj - 1) ? my_rb : lenj) : -1;
        if (act) cells += my_hi - my_lo + 1;
        const int4* rb4 = reinterpret_cast<const int4*>(bnd + ((b & 1) ^ 1) * 8 * W1);
        int* wb = bnd + (b & 1) * 8 * W1;

        // Up-left cell of lane 0's first step: row i0-1 at column jlo-1.
        // Lane-0 sources are functions of lane 0's column jlo+s only, so the
        // whole wavefront evaluates them with uniform addresses.
        int dV = kHigh, dCB = kHigh, dH = kHigh, dED = kHigh, dG = 0;
        if (jlo >= 1) {
          if (b == 0) {
            dCB = 0; dH = HA0; dED = EDA[0];
          } else {
            const int4 v = rb4[2 * (jlo - 1)];
            dV = v.x; dCB = v.y; dH = v.z; dED = v.w;
            dG = reinterpret_cast<const int*>(rb4)[8 * (jlo - 1) + 4];
          }
        }
        int lH = kHigh, lCB = kHigh, lG = 0;
        int oV = kHigh, oCB = kHigh, oH = kHigh, oED = kHigh, oG = 0;
        int cb0run = 0;                               // row-0 prefix at lane 0's column
        int c0prev = sb[imax(jlo - 1, 0)];            // lane 0's previous column symbol
        int jprev = sb[imax(jlo - lane - 1, 0)];      // own previous column symbol
        for (int s = 0; s < nsteps4; s += 4) {
          uint32_t pack = 0;
#pragma unroll
          for (int u = 0; u < 4; u++) {
            const int jl0 = jlo + s + u;              // lane 0's column (uniform)
            const int jcl = imin(jl0, lenj);
            const int c0 = sb[jcl];
            int sV = kHigh, sCB, sH, sED, sG;
            if (b == 0) {
              // POY's row 0 (algn_newkk_test_aff, src/algn.c:2228-2247)
              const int gr0 = s_prep[c0];
              const int gop0 = has_gap_opening(jl0, c0prev, c0, go);
              const int add0 = ((c0prev & kGap) && !(c0 & kGap)) ? gop0 + gr0 : gr0;
              cb0run = (jl0 == 0) ? 0 : cb0run + add0;
              sCB = cb0run;
              sH = (jl0 == 0) ? HA0 : cb0run;
              sED = kFast ? kHigh : EDA[jcl];
              sG = jl0 & 0xFFFF;
            } else {
              const int4 v = rb4[2 * jcl];
              sV = v.x; sCB = v.y; sH = v.z; sED = v.w;
              sG = reinterpret_cast<const int*>(rb4)[8 * jcl + 4];
            }
            c0prev = c0;
            const int jc = wave_shr1(jprev, c0);
            const int uV = wave_shr1(oV, sV);
            const int uCB = wave_shr1(oCB, sCB);
            const int uH = wave_shr1(oH, sH);
            const int uED = kFast ? kHigh : wave_shr1(oED, sED);
            const int uG = wave_shr1(oG, sG);
            const int j = jl0 - lane;
            AffCell o;
            if constexpr (kFast)
              aff_cell_fast(row, go, sw, j, jc, j == my_lo, j == my_rb, s_cost, s_prep, uV, uCB, uG,
                            dV, dCB, dH, dG, lH, lCB, lG, o);
            else
              aff_cell(row, go, sw, j, jc, jprev, j == my_lo, j == my_rb, s_cost, s_prep, uV, uCB,
                       uG, dV, dCB, dH, dED, dG, lH, lCB, lG, o);
            const bool inb = j >= my_lo && j <= my_hi;
            if (inb) {
              if (i == leni && j == lenj) { rc = o.f; rgn = o.gmax; }
              // row-0 buffer persistence across rounds: the last even row of
              // this round that covers column j leaves its ED value there.
              if (!kFast && !(i & 1) && i + 2 > imin(leni, j + newk)) EDA[j] = o.E;
              if (lane == 63) {
                reinterpret_cast<int4*>(wb)[2 * j] = make_int4(o.V, o.C, o.H, o.E);
                wb[8 * j + 4] = o.G;
              }
            }
            oV = o.V; oCB = o.C; oH = o.H; oED = o.E; oG = o.G;
            lH = o.H; lCB = o.C; lG = o.G;
            dV = uV; dCB = uCB; dH = uH; dED = uED; dG = uG;
            jprev = jc;
            if constexpr (kFast)
              pack = (pack << 4) | o.code;  // nibbles, step 0 highest
            else
              pack |= o.code << (8 * u);
          }
          if constexpr (kDm) {
            if constexpr (kFast)
              reinterpret_cast<uint16_t*>(dm + dmpos)[((s >> 2) << 6) + lane] = (uint16_t)pack;
            else
              reinterpret_cast<uint32_t*>(dm + dmpos)[((s >> 2) << 6) + lane] = pack;
          }
        }
        dmpos += (long long)(nsteps4 >> 2) * (kFast ? 128 : 256);
        __threadfence_block();
      }
      const int src = (leni - 1) & 63;
      // uniform by construction; the broadcast keeps the round loop scalar
      cost = __builtin_amdgcn_readfirstlane(__shfl(rc, src));
      gap_num = __builtin_amdgcn_readfirstlane(__shfl(rgn, src));
      rounds++;
      const int newp = (2 * T - D) / 2;
      if (gap_num < p || newp - lenj + 1 >= 0) break;
      if (imin(leni, newk) >= 2) HA0 = kHigh;
      T *= 2;
    }
    for (int o = 32; o > 0; o >>= 1) cells += __shfl_xor(cells, o);
    if (lane == 0) {
      PairResult r;
      r.cost = cost;
      r.gap_num = gap_num;
      r.newk = newk;
      r.rounds = rounds;
      r.full = kFast ? 1 : 0;  // direction format: 1 = gap-free nibbles
      r.pad = 0;
      r.cells = cells;
      r.slots = slots;
      A.res[pid] = r;
    }
    if constexpr (kDm && kFast) {
      // its median table in the rotating schedule's FIFO
      if (A.out != nullptr)
        aff_fused_backtrace(A.pt, A.cm, A.dm, A.stab, A.out, A.outlen,
                            reinterpret_cast<uint8_t*>(s_rot.fV), pid, newk, kFast);
    }
  }
}

struct AffBtArgs {
  PairTable pt;
  const CmDev* cm;
  const uint8_t* dm;
  const long long* stab;
  const PairResult* res;
  uint8_t* out;
  int* outlen;  // 4 per pair: aligned length, aligned length, median, median-with-gaps
  // when set: only the general-class pairs (cls 0; the gap-free kernel ran
  // the others' backtraces), none when the count ngen is 0
  const uint8_t* cls = nullptr;
  const int* ngen = nullptr;
};

// backtrace_aff (src/algn.c:1715-1819) with follow_*_affine (1531-1592),
// one wavefront per pair, in two passes.
//
// Walk: uniform; the mode machine needs only the direction codes, so the
// symbols stay out of it.  The wavefront keeps the direction words of 4
// consecutive 4-step blocks in registers (block q: lane t holds the word of
// row t), the next 4 loading behind them.  The walk visits cells at strictly
// decreasing global steps (every move lowers the step within a stripe, and
// the previous stripe's cells of a column come earlier), so the windows only
// move down and the walk waits for memory about once per window.  Each step
// records its move (0 vertical, 1 horizontal, 2 gap-gap diagonal, 3 align)
// at output position cap-1-step of the first region, 64 steps per store.
//
// Expansion: 64 steps per pass over the lanes.  The row and column of step s
// are the start minus the moves before it (a prefix sum); the four outputs
// of a step follow from its move and the two symbols exactly as the walk's
// puts in the reference do (resi, resj and medianwg take one byte per step,
// the median only where it keeps a column, compacted by a second prefix
// sum).  Outputs are written right to left into four regions of capacity
// La+Lb+2 (resi, resj, median, medianwg).
// Pair pid of the table; newk and the direction format (nib: gap-free
// nibbles) as its fill left them; s_med: median table rows 0..15.
__device__ __forceinline__ void aff_backtrace_pair(PairTable pt, const uint8_t* dm_all,
                                                   const long long* stab_all, uint8_t* out,
                                                   int* outlen, const uint8_t* s_med, int pid,
                                                   int newk, bool nib) {
  const int lane = threadIdx.x;
  const uint8_t* sa = pt.blob + rows_off(pt, pid);
  const uint8_t* sb = pt.blob + cols_off(pt, pid);
  const int La = pt.len_a[pid], Lb = pt.len_b[pid];
  const int geo = pt.swapped[pid];
  const int cap = La + Lb + 2;
  uint8_t* blk = out + pt.out_off[pid];
  // regions in the caller's order: first sequence, second sequence
  uint8_t* ri = geo ? blk + cap : blk;
  uint8_t* rj = geo ? blk : blk + cap;
  uint8_t* md = blk + 2 * cap;
  uint8_t* mw = blk + 3 * cap;
  const uint8_t* dm = dm_all + pt.dm_off[pid];
  const long long* stab = stab_all + pt.stab_off[pid];
  const long long dm_cap = aff_dm_bytes(La, Lb);
  // direction format: bytes (see the top of this file) or, for gap-free
  // pairs, nibbles (todo move, vertical end, horizontal end) whose align-mode
  // move is the todo move of the diagonal predecessor
  const int vend = nib ? 4 : 16, hend = nib ? 8 : 32;

  // ---- walk ----------------------------------------------------------------
  uint32_t wa0 = 0, wa1 = 0, wa2 = 0, wa3 = 0, wb0 = 0, wb1 = 0, wb2 = 0, wb3 = 0;
  long long qa = -(1ll << 60), qb = -(1ll << 60);  // top blocks of the windows
  uint32_t mv = 0;  // move of step 64m + lane of the current chunk m
  int n = 0;        // steps recorded
  auto record = [&](int m) __attribute__((always_inline)) {
    if (lane == (n & 63)) mv = (uint32_t)m;
    n++;
    if ((n & 63) == 0) ri[cap - 1 - (n - 64) - lane] = (uint8_t)mv;
  };
  int i = La - 1, j = Lb - 1;
  int mode = 4;  // 0 vert, 1 horz, 2 diag, 3 align, 4 todo
  int cb = -1, jlo = 0;
  long long base = 0;
  while (i != 0 && j != 0) {
    const int bb = (i - 1) >> 6;
    if (bb != cb) {
      cb = bb;
      jlo = imax(1 + (bb << 6) - newk, 0);
      base = stab[bb];
    }
    const int t = (i - 1) & 63;
    const int s = j - jlo + t;
    // stab[b] = global step of the stripe's first step; cells on the path are
    // inside the band (a step before the stripe reads code 0)
    int code = 0;
    if (s >= 0) {
      const long long g = base + s;
      const long long q = g >> 2;
      if (q > qa || q < qa - 3) {
        if (q <= qb && q >= qb - 3) {
          wa0 = wb0; wa1 = wb1; wa2 = wb2; wa3 = wb3;
          qa = qb;
        } else {
          wa0 = dm_word(dm, dm_cap, nib, q);
          wa1 = dm_word(dm, dm_cap, nib, q - 1);
          wa2 = dm_word(dm, dm_cap, nib, q - 2);
          wa3 = dm_word(dm, dm_cap, nib, q - 3);
          qa = q;
        }
        // the next window loads while this one is walked
        wb0 = dm_word(dm, dm_cap, nib, qa - 4);
        wb1 = dm_word(dm, dm_cap, nib, qa - 5);
        wb2 = dm_word(dm, dm_cap, nib, qa - 6);
        wb3 = dm_word(dm, dm_cap, nib, qa - 7);
        qb = qa - 4;
      }
      const int k = (int)(qa - q);
      const uint32_t v = k < 2 ? (k == 0 ? wa0 : wa1) : (k == 2 ? wa2 : wa3);
      const uint32_t w = __builtin_amdgcn_readlane(v, t);
      code = nib ? (int)(w >> (12 - 4 * (int)(g & 3))) & 15 : (int)(w >> (8 * (int)(g & 3))) & 255;
    }
    // a todo read decides the move of this same cell
    if (mode == 4) mode = code & 3;
    record(mode);
    if (mode == 0) {
      if (code & vend) mode = 4;
      i--;
    } else if (mode == 1) {
      if (code & hend) mode = 4;
      j--;
    } else if (mode == 2) {
      if (code & 64) mode = 4;
      i--; j--;
    } else {
      mode = nib ? 4 : (code >> 2) & 3;
      i--; j--;
    }
  }
  // the reference's trailing loops: vertical to row 0, then horizontal
  while (i != 0) { record(0); i--; }
  while (j != 0) { record(1); j--; }
  if ((n & 63) != 0 && lane < (n & 63)) ri[cap - 1 - (n & ~63) - lane] = (uint8_t)mv;
  __threadfence_block();

  // ---- expansion -----------------------------------------------------------
  int ci = 0, cj = 0, cm = 0;  // rows, columns and median bytes before the chunk
  int mlast = -1;              // last median byte written
  for (int c = 0; c < n; c += 64) {
    const int s = c + lane;
    const bool ok = s < n;
    const int m = ok ? ri[cap - 1 - s] : 1;
    // moves before this one, packed: rows in bits 0-7, columns in 8-15
    const int di = ok && m != 1, dj = ok && m != 0;
    int tot;
    const int ex = wave_excl_sum(di | (dj << 8), tot);
    const int ii = La - 1 - ci - (ex & 255), jj = Lb - 1 - cj - (ex >> 8);
    const int ic = ok && m != 1 ? sa[ii] : 0;
    const int jc = ok && m != 0 ? sb[jj] : 0;
    int vri = kGap, vrj = kGap, vmw = kGap, vmd = -1;
    if (m == 0) {
      vri = ic;
      vmw = (ic & kGap) ? kGap : ic | kGap;
      if (!(ic & kGap)) vmd = ic | kGap;
    } else if (m == 1) {
      vrj = jc;
      vmw = (jc & kGap) ? kGap : jc | kGap;
      if (!(jc & kGap)) vmd = jc | kGap;
    } else if (m == 2) {
      vri = ic;
      vrj = jc;
    } else {
      vri = ic;
      vrj = jc;
      vmw = vmd = s_med[((ic & 15) << 5) + (jc & 15)];
    }
    const bool emd = ok && vmd >= 0;
    int mtot;
    const int mex = wave_excl_sum(emd ? 1 : 0, mtot);
    if (ok) {
      ri[cap - 1 - s] = (uint8_t)vri;
      rj[cap - 1 - s] = (uint8_t)vrj;
      mw[cap - 1 - s] = (uint8_t)vmw;
    }
    if (emd) md[cap - 1 - (cm + mex)] = (uint8_t)vmd;
    if (mtot > 0) {
      const int lastl = 63 - __builtin_clzll(__builtin_amdgcn_ballot_w64(emd));
      mlast = __builtin_amdgcn_readlane(vmd, lastl);
    }
    ci += tot & 255;
    cj += tot >> 8;
    cm += mtot;
  }
  // the leading gap column; the median takes it unless it already ends so
  if (lane == 0) {
    ri[cap - 1 - n] = kGap;
    rj[cap - 1 - n] = kGap;
    mw[cap - 1 - n] = kGap;
    if (cm == 0 || mlast != kGap) md[cap - 1 - cm] = kGap;
  }
  if (cm == 0 || mlast != kGap) cm++;
  if (lane == 0) {
    outlen[4 * pid + (geo ? 1 : 0)] = n + 1;
    outlen[4 * pid + (geo ? 0 : 1)] = n + 1;
    outlen[4 * pid + 2] = cm;
    outlen[4 * pid + 3] = n + 1;
  }
}

__global__ __launch_bounds__(64) void k_affine_backtrace(AffBtArgs A) {
  // median table rows 0..15 in shared memory
  __shared__ uint8_t s_med[512];
  const int lane = threadIdx.x;
  if (A.ngen != nullptr && __builtin_amdgcn_readfirstlane(*A.ngen) == 0) return;
  const int pid = __builtin_amdgcn_readfirstlane(A.pt.order[blockIdx.x]);
  if (A.cls != nullptr && A.cls[pid] != 0) return;
  for (int k = lane; k < 512; k += 64) s_med[k] = A.cm->median[k];
  __syncthreads();
  const PairResult R = A.res[pid];
  aff_backtrace_pair(A.pt, A.dm, A.stab, A.out, A.outlen, s_med, pid, R.newk, R.full != 0);
}

// algn_CAML_cost_affine_3 (src/algn.c:2457) -> algn_fill_plane_3_aff_nobt
// (1987): full matrix, no directions, either order (rows = shorter).
struct AffCostArgs {
  PairTable pt;
  const CmDev* cm;
  int* ws;
  long long ws_stride;
  PairResult* res;
  int* counter;
};

__global__ __launch_bounds__(64) void k_affine_cost_full(AffCostArgs A) {
  __shared__ int s_cost[1024];
  __shared__ int s_prep[32];
  const int lane = threadIdx.x;
  for (int k = lane; k < 1024; k += 64) s_cost[k] = A.cm->cost[k];
  if (lane < 32) s_prep[lane] = A.cm->prepend[lane];
  __syncthreads();
  const int go = A.cm->gap_open;
  for (;;) {
    const int q = wave_grab(A.counter);
    if (q >= A.pt.n) break;
    const int pid = A.pt.order[q];
    const uint8_t* __restrict__ sa = A.pt.blob + rows_off(A.pt, pid);
    const uint8_t* __restrict__ sb = A.pt.blob + cols_off(A.pt, pid);
    const int La = A.pt.len_a[pid], Lb = A.pt.len_b[pid];
    const int leni = La - 1, lenj = Lb - 1;
    const int W1 = lenj + 1;
    int* bnd = A.ws + (long long)blockIdx.x * A.ws_stride;
    int rc = 0;
    if (leni == 0) {
      // row 0 of initialize_matrices_affine_nobt: H = go + prefix sums
      int sum = 0;
      for (int j = 1 + lane; j <= lenj; j += 64) sum += s_prep[sb[j]];
      for (int o = 32; o > 0; o >>= 1) sum += __shfl_xor(sum, o);
      if (lane == 0) {
        PairResult r{};
        r.cost = lenj == 0 ? 0 : go + sum;  // min(H, V=go|HIGH, ED, CB) at lenj
        if (lenj == 0) r.cost = 0;
        r.newk = -1;
        r.rounds = 1;
        A.res[pid] = r;
      }
      continue;
    }
    const int nstripes = (leni + 63) >> 6;
    for (int b = 0; b < nstripes; b++) {
      const int i0 = 1 + (b << 6);
      const int ilast = imin(i0 + 63, leni);
      const int i = i0 + lane;
      const bool act = i <= leni;
      const int nsteps = W1 + (ilast - i0);
      int ic = 0, ip = 0, si_ge = 0, si_go = 0, si_ng = 0, si_vext = 0;
      if (act) {
        ic = sa[i];
        ip = sa[i - 1];
        si_ge = s_cost[(ic << 5) + kGap];
        si_go = has_gap_opening(i, ip, ic, go);
        si_ng = ic & 15;
        si_vext = (i > 1 && (ip & kGap) && !(ic & kGap)) ? si_go + si_ge : si_ge;
      }
      const int* rb = bnd + ((b & 1) ^ 1) * 4 * W1;
      int* wb = bnd + (b & 1) * 4 * W1;
      int dV = kHigh, dCB = kHigh, dH = kHigh, dED = kHigh;
      int lH = kHigh, lCB = kHigh;
      int oV = kHigh, oCB = kHigh, oH = kHigh, oED = kHigh;
      int h0run = go;
      int jprev = ldc(sb, -lane - 1, Lb);
      int jc_next = ldc(sb, -lane, Lb);
      for (int s = 0; s < nsteps; s++) {
        const int j = s - lane;
        const int jc = jc_next;
        jc_next = ldc(sb, j + 1, Lb);
        int sV = kHigh, sCB = kHigh, sH = kHigh, sED = kHigh;
        if (lane == 0) {
          if (b == 0) {
            // row 0 (initialize_matrices_affine_nobt, src/algn.c:1822-1864)
            if (j == 0) { sV = go; sCB = 0; sH = go; sED = kHigh; h0run = go; }
            else { h0run += s_prep[jc]; sV = kHigh; sCB = kHigh; sH = h0run; sED = kHigh; }
          } else if (j <= lenj) {
            sV = rb[j]; sCB = rb[W1 + j]; sH = rb[2 * W1 + j]; sED = rb[3 * W1 + j];
          }
        }
        const int uV = wave_shr1(oV, sV);
        const int uCB = wave_shr1(oCB, sCB);
        const int uH = wave_shr1(oH, sH);
        const int uED = wave_shr1(oED, sED);
        if (act && j >= 0 && j <= lenj) {
          int H, V, E, C;
          if (j == 0) {
            const int r = uV + si_vext;
            H = kHigh; E = kHigh; V = r; C = kHigh;
          } else {
            const int gr = s_prep[jc];
            const int sj_go = has_gap_opening(j, jprev, jc, go);
            const int shx = (j == 1) ? gr : (((jprev & kGap) && !(jc & kGap)) ? sj_go + gr : gr);
            int ext = lH + shx, opn = lCB + sj_go + gr;
            H = ext < opn ? ext : opn;
            // algn_CAML_cost_affine_3 carves its state arrays 2*largest ints
            // apart but fills them with a row offset of largest + 1
            // (src/algn.c:2488-2491 with algn_fill_plane_3_aff_nobt's
            // offset, 2006): an even row's start-of-row writes of H[0] and
            // E[0] overwrite the odd row's V and C at the last column, so
            // V at (even i, lenj) has only HIGH sources
            const bool lost = (j == lenj) && ((i & 1) == 0);
            ext = (lost ? kHigh : uV) + si_vext;
            opn = (lost ? kHigh : uCB) + si_go + si_ge;
            V = ext < opn ? ext : opn;
            const bool flag = (kGap & ic) && (kGap & jc);
            const bool flag2 = !(kGap & ip) && !(kGap & jprev);
            const int dg = flag ? 0 : kHigh;
            const int od = flag ? (flag2 ? 0 : 2 * go) : kHigh;
            ext = dED + dg; opn = dCB + od;
            E = ext < opn ? ext : opn;
            const int d = s_cost[(si_ng << 5) + (jc & 15)];
            const int extra = (sj_go < si_go) ? si_go : sj_go;
            // with lenj = 1 the same overlap makes an odd row's own V and
            // E of column 1 land on the H and V of column 0 of the row
            // above before the aligned state reads them
            const bool self1 = (lenj == 1) && ((i & 1) == 1);
            int cb = dCB + d;
            const int fv = (self1 ? E : dV) + d + ((ic == si_ng) ? 0 : sj_go);
            const int fh = (self1 ? V : dH) + d + ((jc == (jc & 15)) ? 0 : si_go);
            const int fd = dED + d + extra;
            if (cb > fv) cb = fv;
            if (cb > fh) cb = fh;
            if (cb > fd) cb = fd;
            C = cb;
          }
          oV = V; oCB = C; oH = H; oED = E;
          lH = H; lCB = C;
          if (i == leni && j == lenj) rc = imin(imin(H, V), imin(E, C));
          if (lane == 63) { wb[j] = V; wb[W1 + j] = C; wb[2 * W1 + j] = H; wb[3 * W1 + j] = E; }
        }
        dV = uV; dCB = uCB; dH = uH; dED = uED;
        jprev = jc;
      }
      __threadfence_block();
    }
    const int cost = __builtin_amdgcn_readfirstlane(__shfl(rc, (leni - 1) & 63));
    if (lane == 0) {
      PairResult r{};
      r.cost = cost;
      r.newk = leni - 1;
      r.rounds = 1;
      r.cells = (long long)leni * (lenj + 1);
      A.res[pid] = r;
    }
  }
}

template __global__ void k_affine_fill<true, false>(AffFillArgs);
template __global__ void k_affine_fill<true, true>(AffFillArgs);
template __global__ void k_affine_fill<false, false>(AffFillArgs);

}  // namespace poyhip
