// Non-affine DO alignment (POY algn_nw, src/algn.c:2972) on gfx950: full
// plane or Ukkonen band with T-doubling (algn_fill_plane_2, 1134), one
// wavefront per pair in systolic 64-row stripes.  Each cell keeps the move
// backtrace_2d (3278) would take under the pair's swap flag (2 bits) plus
// the number of deletions on the path backtrace_2d_gaps (978) would take, so
// the band stop test needs no second backtrace.
#include "poyhip_device.h"
#include "poyhip_layout.h"

namespace poyhip {

struct LinFillArgs {
  PairTable pt;
  const CmDev* cm;
  uint8_t* dm;
  int* ws;            // one workspace slot of ws_stride ints per wavefront
  long long ws_stride;
  long long* stab;
  PairResult* res;
  int* counter;
  // 0: every doubling round, directions kept from the last one
  // 1: every doubling round, no directions; the final threshold T is left
  //    in res[pid].pad for phase 2
  // 2: one round at the threshold res[pid].pad, directions kept; the
  //    statistics of phase 1 are kept
  int phase;
};

__device__ __forceinline__ int ldc8(const uint8_t* p, int j, int n) {
  return p[(unsigned)j < (unsigned)n ? j : 0];
}

// algn_fill_plane_2's choice between the full plane and the band
// (src/algn.c:1134-1177) for width 50 and height 50 + (L1 - L2)
// (algn_nw_limit, 2954).
__device__ __forceinline__ bool lin_full_plane(int L1, int L2, int deltawh) {
  int height = 50 + (L1 - L2) + deltawh;
  if (height > L1) height = L1;
  if ((float)L1 >= 1.5f * (float)L2) return true;
  if (2 * height < L1) return false;
  return 8 >= (L1 - height);
}

template <bool kDm>
__global__ __launch_bounds__(64) void k_linear_fill(LinFillArgs A) {
  __shared__ int s_cost[1024];
  __shared__ int s_prep[32];
  __shared__ int s_tail[32];
  const int lane = threadIdx.x;
  for (int k = lane; k < 1024; k += 64) s_cost[k] = A.cm->cost[k];
  if (lane < 32) {
    s_prep[lane] = A.cm->prepend[lane];
    s_tail[lane] = A.cm->tail[lane];
  }
  __syncthreads();
  const int delta = A.cm->delta;

  for (;;) {
    const int q = wave_grab(A.counter);
    if (q >= A.pt.n) break;
    const int pid = A.pt.order[q];
    const uint8_t* __restrict__ sa = A.pt.blob + rows_off(A.pt, pid);
    const uint8_t* __restrict__ sb = A.pt.blob + cols_off(A.pt, pid);
    const int L1 = A.pt.len_a[pid], L2 = A.pt.len_b[pid];
    const int sw = A.pt.prio[pid];
    const bool full = lin_full_plane(L1, L2, A.pt.deltawh[pid]);
    int* bnd = A.ws + (long long)blockIdx.x * A.ws_stride;
    uint8_t* dm = A.dm + (kDm ? A.pt.dm_off[pid] : 0);
    long long* stab = A.stab + A.pt.stab_off[pid];
    const int Dd = L2 - L1;

    int T = (Dd + 1) * delta;
    int rounds = 0, newk = 0, cost = 0, gap_num = 0;
    long long cells = 0, slots = 0;
    long long cells1 = 0;
    if (A.phase == 2) {
      const PairResult r1 = A.res[pid];
      T = r1.pad;
      rounds = r1.rounds - 1;
      cells1 = r1.cells;
      slots = lane == 0 ? r1.slots : 0;
    }
    int Tfin = T;
    for (;;) {
      Tfin = T;
      const int p = (T - Dd) / 2;
      // A band round whose stop test cannot pass is skipped: its gap count
      // is the deletion count plus Dd, so (gap_num + 1) < p needs p > Dd + 1,
      // and rounds leave no state behind.
      if (!full && A.phase != 2 && p <= Dd + 1 && (2 * T - Dd) / 2 - L2 + 1 < 0) {
        rounds++;
        T *= 2;
        continue;
      }
      newk = full ? L1 - 1 : (p >= L1 ? L1 - 1 : p);
      const int lfr = full ? L2 : imin(L2 - L1 + 1 + p, L2);
      if (lane == 0) cells += lfr;
      int rc = 0, rdc = 0;
      if (L1 == 1) {
        int sum = 0;
        for (int j = 1 + lane; j < L2; j += 64) sum += s_prep[sb[j]];
        for (int o = 32; o > 0; o >>= 1) sum += __shfl_xor(sum, o);
        rc = sum;
        rdc = 0;
      }
      const int nstripes = (L1 - 1 + 63) >> 6;
      long long dmpos = 0;
      // final cell (L1-1, L2-1): last stripe, lane (L1-2) & 63
      const int blast = nstripes - 1;
      const int jlo_last = full ? 0 : imax(1 + (blast << 6) - newk, 0);
      const int s_fin = (L2 - 1) - jlo_last + ((L1 - 2) & 63);
      for (int b = 0; b < nstripes; b++) {
        const int i0 = 1 + (b << 6);
        const int ilast = imin(i0 + 63, L1 - 1);
        const int i = i0 + lane;
        const bool act = i <= L1 - 1;
        const int jlo = full ? 0 : imax(i0 - newk, 0);
        const int jhi = full ? L2 - 1 : imin(ilast + Dd + newk, L2 - 1);
        const int nsteps = (jhi - jlo + 1) + (ilast - i0);
        const int nsteps16 = (nsteps + 15) & ~15;
        if (kDm && lane == 0) stab[b] = dmpos;
        slots += (long long)nsteps * 64;
        // row constants; lanes past L1-1 get an empty band
        const int ir = imin(i, L1 - 1);
        const int s1i = sa[ir];
        const int c = s_cost[(s1i << 5) + kGap];
        const int tl = s_tail[s1i];
        const int col0 = full ? c : tl;
        const int arow = s1i << 5;
        int lo = 1 << 30, hi = -1;
        bool hasL = false, hasR = false;
        if (act) {
          if (full) {
            lo = 0;
            hi = L2 - 1;
          } else {
            hasL = (i - newk) > 0;
            lo = hasL ? i - newk : 0;
            hasR = (i + Dd + newk) <= L2 - 1;
            hi = hasR ? i + Dd + newk : L2 - 1;
          }
          cells += hi - lo + 1;
        }
        const int2* rb = reinterpret_cast<const int2*>(bnd + ((b & 1) ^ 1) * 2 * L2);
        int2* wb = reinterpret_cast<int2*>(bnd + (b & 1) * 2 * L2);
        // up-left cell of lane 0's first step: row i0-1 at column jlo-1
        int dmv = 0, ddc = 0;
        if (jlo >= 1 && b > 0) {
          const int2 v = rb[jlo - 1];
          dmv = v.x;
          ddc = v.y;
        }
        int lmv = 0, ldc = 0;   // left
        int omv = 0, odc = 0;   // handed to the next lane
        int r0run = 0;          // POY's row 0 at lane 0's column (stripe 0)
        int jprev = ldc8(sb, jlo - lane - 1, L2);
        // lane 0's sources of a 16-step block (column symbol, row above),
        // spread over lanes 0..15 and loaded one block ahead; step u reads
        // lane u's copy with readlane
        auto load_src = [&](int s0, int& sym, int& mv, int& dcv) {
          const int jc0 = imin(jlo + s0 + (lane & 15), L2 - 1);
          sym = sb[jc0];
          mv = 0;
          dcv = 0;
          if (b > 0) {
            const int2 v = rb[jc0];
            mv = v.x;
            dcv = v.y;
          }
        };
        int nsym, nmv, ndc;
        load_src(0, nsym, nmv, ndc);
        for (int s = 0; s < nsteps16; s += 16) {
          const int csym = nsym, cmv = nmv, cdc = ndc;
          load_src(s + 16, nsym, nmv, ndc);
          uint32_t pack = 0;
#pragma unroll
          for (int u = 0; u < 16; u++) {
            const int st = s + u;
            const int j = jlo + st - lane;
            const int c0 = __builtin_amdgcn_readlane(csym, u);
            int smv = __builtin_amdgcn_readlane(cmv, u), sdc = __builtin_amdgcn_readlane(cdc, u);
            if (b == 0) {
              // row 0: prefix sums of prepend costs (algn_fill_first_row)
              if (jlo + st > 0) r0run += s_prep[c0];
              smv = r0run;
              sdc = 0;
            }
            const int jc = wave_shr1(jprev, c0);
            const int umv = wave_shr1(omv, smv);
            const int udc = wave_shr1(odc, sdc);
            // candidates of algn_fill_row / the band edge cells
            // (src/algn.c:458-533, 543-652); a disabled move is priced out
            constexpr int kOff = 1 << 29;
            const int alg = s_cost[arow + jc];
            const int gapj = s_cost[(kGap << 5) + jc];
            const bool edge_l = hasL && j == lo;
            const bool edge_r = hasR && j == hi;
            const int del = edge_r ? kOff : umv + c;
            const int ins = edge_l ? kOff : lmv + gapj;
            const int aln = dmv + alg;
            int val = imin(imin(del, ins), aln);
            int fl = (aln == val ? kAlign : 0) | (ins == val ? kInsert : 0) |
                     (del == val ? kDelete : 0);
            // algn_fill_last_column (src/algn.c:654)
            const int cst = (!hasR && j == L2 - 1) ? tl + umv : kOff;
            fl = cst < val ? kDelete : (fl | (cst == val ? kDelete : 0));
            val = imin(val, cst);
            if (j == 0) {
              val = umv + col0;
              fl = kDelete;
            }
            // backtrace_2d_gaps order: align, insert, delete
            const int dc = (fl & kAlign) ? ddc : ((fl & kInsert) ? ldc : udc + 1);
            // backtrace_2d order: align, then delete/insert by swap flag
            const uint32_t mvv = (fl & kAlign) ? 0u
                                 : (sw ? ((fl & kInsert) ? 2u : 1u) : ((fl & kDelete) ? 1u : 2u));
            const bool inb = act && j >= lo && j <= hi;
            if (inb) {
              omv = val;
              odc = dc;
              lmv = val;
              ldc = dc;
              if (lane == 63) wb[j] = make_int2(val, dc);
            }
            if (b == blast && st == s_fin) {
              rc = val;
              rdc = dc;
            }
            pack |= (inb ? mvv : 0u) << (2 * u);
            dmv = umv;
            ddc = udc;
            jprev = jc;
          }
          if (kDm) reinterpret_cast<uint32_t*>(dm + dmpos)[((s >> 4) << 6) + lane] = pack;
        }
        dmpos += (long long)(nsteps16 >> 4) * 256;
        __threadfence_block();
      }
      if (L1 > 1) {
        const int src = (L1 - 2) & 63;
        // uniform by construction; the broadcast keeps the round loop scalar
        cost = __builtin_amdgcn_readfirstlane(__shfl(rc, src));
        gap_num = __builtin_amdgcn_readfirstlane(__shfl(rdc, src)) + Dd;
      } else {
        cost = __builtin_amdgcn_readfirstlane(rc);
        gap_num = L2 - 1;
      }
      rounds++;
      if (full || A.phase == 2) break;
      const int newp = (2 * T - Dd) / 2;
      if ((gap_num + 1) < p || newp - L2 + 1 >= 0) break;
      T *= 2;
    }
    // phase 2 repeats phase 1's last round: report phase 1's cells
    if (A.phase == 2) cells = lane == 0 ? cells1 : 0;
    for (int o = 32; o > 0; o >>= 1) cells += __shfl_xor(cells, o);
    if (lane == 0) {
      PairResult r;
      r.cost = cost;
      r.gap_num = gap_num;
      r.newk = newk;
      r.rounds = rounds;
      r.full = full ? 1 : 0;
      r.pad = Tfin;
      r.cells = cells;
      r.slots = slots;
      A.res[pid] = r;
    }
  }
}

struct LinBtArgs {
  PairTable pt;
  const CmDev* cm;
  const uint8_t* dm;
  const long long* stab;
  const PairResult* res;
  uint8_t* out;
  int* outlen;  // 4 per pair: r1, r2, ancestor, median with gaps
  // limits of algn_CAML_backtrace_2d_limit (one pair; lim = 0: none): the
  // walk starts at row a1 - 1, column st2 + a2 - 1, reads the rows' symbols
  // from st1 + a1 - 1 down and the columns' from st2 + a2 - 1 down, and
  // ends when it reaches row 0 left of column st2
  int lim = 0, lim_st1 = 0, lim_st2 = 0, lim_a1 = 0, lim_a2 = 0;
};

// backtrace_2d (src/algn.c:3278) over the packed moves, emitting the aligned
// pair together with algn_ancestor_2 (3603) for the linear model and
// seq_get_median_2d_with_gaps (src/seq.c:241).  r1 belongs to the DP rows'
// sequence; medians are looked up in the caller's argument order.
//
// One wavefront per pair, in two passes.  Walk: uniform; the move of a cell
// is its 2-bit code (0 diagonal, 1 up, 2 left), read from a register window
// of 4 consecutive 16-step blocks (block q: lane t holds the word of row t)
// while the next 4 load behind it; every move lowers the cell's global step
// (within a stripe and into the stripe above), so the windows only move
// down.  Each step records its move at output position cap-1-step of r1, 64
// steps per store; the walk ends with the (0, 0) cell, which emits the two
// leading gaps.  Expansion: 64 steps per pass over the lanes; the row and
// column of step s are the start minus the moves before it (a prefix sum),
// the ancestor keeps the steps whose median is not a gap (a second prefix
// sum).
__global__ __launch_bounds__(64) void k_linear_backtrace(LinBtArgs A) {
  __shared__ uint8_t s_med[1024];  // median table, 32 x 32
  const int lane = threadIdx.x;
  for (int k = lane; k < 1024; k += 64) s_med[k] = A.cm->median[k];
  __syncthreads();
  const int pid = __builtin_amdgcn_readfirstlane(A.pt.order[blockIdx.x]);
  const uint8_t* sa = A.pt.blob + rows_off(A.pt, pid);
  const uint8_t* sb = A.pt.blob + cols_off(A.pt, pid);
  const int L1 = A.pt.len_a[pid], L2 = A.pt.len_b[pid];
  const int sw = A.pt.swapped[pid];  // caller order of the median lookup
  const int cap = (int)lin_out_cap(L1, L2);
  uint8_t* blk = A.out + A.pt.out_off[pid];
  uint8_t* r1 = sw ? blk + cap : blk;
  uint8_t* r2 = sw ? blk : blk + cap;
  uint8_t* an = blk + 2 * cap;
  uint8_t* mw = blk + 3 * cap;
  const uint8_t* dm = A.dm + A.pt.dm_off[pid];
  const long long* stab = A.stab + A.pt.stab_off[pid];
  const PairResult R = A.res[pid];
  const bool full = R.full != 0;
  const int newk = R.newk;
  const long long dm_cap = lin_dm_bytes(L1, L2);

  // ---- walk ----------------------------------------------------------------
  uint32_t wa0 = 0, wa1 = 0, wa2 = 0, wa3 = 0, wb0 = 0, wb1 = 0, wb2 = 0, wb3 = 0;
  long long qa = -(1ll << 60), qb = -(1ll << 60);  // top blocks of the windows
  uint32_t mv = 0;  // move of step 64m + lane of the current chunk m
  int n = 0;        // steps recorded
  auto record = [&](int m) __attribute__((always_inline)) {
    if (lane == (n & 63)) mv = (uint32_t)m;
    n++;
    if ((n & 63) == 0) r1[cap - 1 - (n - 64) - lane] = (uint8_t)mv;
  };
  int i = A.lim ? A.lim_a1 - 1 : L1 - 1, j = A.lim ? A.lim_st2 + A.lim_a2 - 1 : L2 - 1;
  const int stop_col = A.lim ? A.lim_st2 : 0;
  const int top_a = A.lim ? A.lim_st1 + A.lim_a1 - 1 : L1 - 1;  // last row symbol read
  const int top_b = A.lim ? A.lim_st2 + A.lim_a2 - 1 : L2 - 1;
  int cb = -1, jlo = 0;
  long long base = 0;  // first 16-step block of the stripe
  while (i > 0 && j > 0) {
    const int bb = (i - 1) >> 6;
    if (bb != cb) {
      cb = bb;
      jlo = full ? 0 : imax(1 + (bb << 6) - newk, 0);
      base = stab[bb] >> 8;  // 64 lanes x 4 bytes per block
    }
    const int t = (i - 1) & 63;
    const int s = j - jlo + t;
    // path cells are inside the band (a step before the stripe reads the
    // diagonal code 0)
    int code = 0;
    if (s >= 0) {
      const long long q = base + (s >> 4);
      if (q > qa || q < qa - 3) {
        if (q <= qb && q >= qb - 3) {
          wa0 = wb0; wa1 = wb1; wa2 = wb2; wa3 = wb3;
          qa = qb;
        } else {
          wa0 = dm_word(dm, dm_cap, false, q);
          wa1 = dm_word(dm, dm_cap, false, q - 1);
          wa2 = dm_word(dm, dm_cap, false, q - 2);
          wa3 = dm_word(dm, dm_cap, false, q - 3);
          qa = q;
        }
        // the next window loads while this one is walked
        wb0 = dm_word(dm, dm_cap, false, qa - 4);
        wb1 = dm_word(dm, dm_cap, false, qa - 5);
        wb2 = dm_word(dm, dm_cap, false, qa - 6);
        wb3 = dm_word(dm, dm_cap, false, qa - 7);
        qb = qa - 4;
      }
      const int k = (int)(qa - q);
      const uint32_t v = k < 2 ? (k == 0 ? wa0 : wa1) : (k == 2 ? wa2 : wa3);
      const uint32_t w = __builtin_amdgcn_readlane(v, t);
      code = (int)(w >> ((s & 15) << 1)) & 3;
    }
    record(code);
    if (code == 0) {
      i--;
      j--;
    } else if (code == 1) {
      i--;
    } else {
      j--;
    }
  }
  // the reference's border moves: up along column 0, left along row 0 (to
  // column stop_col), and the (0, 0) cell; a walk that starts above row 0
  // (a limit of a1 = 0) makes no move
  if (i >= 0) {
    while (i > 0) { record(1); i--; }
    while (j > 0 && j >= stop_col) { record(2); j--; }
    if (j == 0 && stop_col == 0) record(0);
  }
  if ((n & 63) != 0 && lane < (n & 63)) r1[cap - 1 - (n & ~63) - lane] = (uint8_t)mv;
  __threadfence_block();

  // ---- expansion -----------------------------------------------------------
  int ci = 0, cj = 0, ca = 0;  // rows, columns and ancestor bytes before the chunk
  for (int c = 0; c < n; c += 64) {
    const int s = c + lane;
    const bool ok = s < n;
    const int m = ok ? r1[cap - 1 - s] : 3;
    const bool di = ok && m < 2, dj = ok && m != 1;
    // moves before this one, packed: rows in bits 0-7, columns in 8-15
    int tot;
    const int ex = wave_excl_sum((di ? 1 : 0) | (dj ? 256 : 0), tot);
    const int x = di ? sa[top_a - ci - (ex & 255)] : kGap;
    const int y = dj ? sb[top_b - cj - (ex >> 8)] : kGap;
    const int md = sw ? s_med[(y << 5) + x] : s_med[(x << 5) + y];
    const bool ea = ok && md != kGap;
    int atot;
    const int aex = wave_excl_sum(ea ? 1 : 0, atot);
    if (ok) {
      r1[cap - 1 - s] = (uint8_t)x;
      r2[cap - 1 - s] = (uint8_t)y;
      mw[cap - 1 - s] = (uint8_t)md;
    }
    if (ea) an[cap - 1 - (ca + aex)] = (uint8_t)md;
    ci += tot & 255;
    cj += tot >> 8;
    ca += atot;
  }
  if (lane == 0) {
    an[cap - 1 - ca] = kGap;
    A.outlen[4 * pid + (sw ? 1 : 0)] = n;
    A.outlen[4 * pid + (sw ? 0 : 1)] = n;
    A.outlen[4 * pid + 2] = ca + 1;
    A.outlen[4 * pid + 3] = n;
  }
}

// Align.count_gaps (seq_CAML_count, src/seq.c:644) for both sequences and
// Align.cost_2's deltawh (src/sequence.ml:868-936).  has_deltaw = 0 is the
// ?deltaw=None call made by align_2; 1 is DOS.distance's ~deltaw.  One
// wavefront per pair (kDwPairsPerBlock a block): lanes count consecutive
// bytes, a wavefront sum gives the counts.
constexpr int kDwPairsPerBlock = 4;

__global__ __launch_bounds__(64 * kDwPairsPerBlock) void k_linear_deltawh(PairTable pt,
                                                                       int* deltawh,
                                                                       int has_deltaw,
                                                                       const int* deltaw) {
  const int pid = blockIdx.x * kDwPairsPerBlock + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (pid >= pt.npairs) return;
  const uint8_t* sa = pt.blob + rows_off(pt, pid);
  const uint8_t* sb = pt.blob + cols_off(pt, pid);
  const int L1 = pt.len_a[pid], L2 = pt.len_b[pid];
  // a symbol counts when it carries the gap bit
  int g1 = 0, g2 = 0;
#pragma unroll 4
  for (int k = lane; k < L1; k += 64) g1 += (sa[k] >> 4) & 1;
#pragma unroll 4
  for (int k = lane; k < L2; k += 64) g2 += (sb[k] >> 4) & 1;
  for (int o = 32; o > 0; o >>= 1) {
    g1 += __shfl_xor(g1, o);
    g2 += __shfl_xor(g2, o);
  }
  if (lane != 0) return;
  const int dif = L1 - L2;
  const int lower = L1 / 10;
  // has_deltaw: 0 = ?deltaw=None for every pair, 1 = ~deltaw for every
  // pair, 2 = per pair (a negative deltaw[pid] means None)
  int dw;
  if (!has_deltaw || (has_deltaw == 2 && deltaw[pid] < 0)) dw = dif < lower ? lower / 2 : 2;
  else dw = dif < lower ? lower : deltaw[pid];
  deltawh[pid] = imax(g1, g2) + dw;
}

template __global__ void k_linear_fill<true>(LinFillArgs);
template __global__ void k_linear_fill<false>(LinFillArgs);

}  // namespace poyhip
