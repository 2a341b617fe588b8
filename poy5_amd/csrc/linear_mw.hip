// Non-affine DO alignment (POY algn_nw, src/algn.c:2972) with one workgroup
// of NW wavefronts per pair.
//
// The DP rows are cut into 64-row stripes as in k_linear_fill; stripe b runs
// on wavefront b mod NW, so NW stripes of one pair are in flight at once.
// Lane t of a stripe owns row i0+t and evaluates column jlo+s-t at step s;
// the row above comes from lane t-1 through DPP, and lane 0 reads the last
// row of the previous stripe from a boundary row in global memory that lane
// 63 of the previous stripe's wavefront wrote.
//
// Schedule: steps are grouped in blocks of 16, and the whole workgroup
// passes a barrier after every block.  Stripe b runs during blocks
// [P*b, P*b + nblk(b)) with a fixed period P (blocks) chosen per round:
// P >= kMinLag keeps stripe b far enough behind stripe b-1 that every value
// its lane 0 loads (one block ahead) was written before an earlier barrier,
// and P >= ceil(max nblk / NW) lets a wavefront finish stripe b before its
// next stripe b+NW starts.  Wavefront w's boundary row is overwritten by
// stripe b+NW only after stripe b+1 has read the same columns (stripe b+NW
// trails stripe b+1 by (NW-1)*P blocks, and in a band its columns are at
// most 64*(NW-1) ahead), so one row per wavefront suffices.
//
// Per cell the recurrence, tie sets and band-edge rules are those of
// k_linear_fill (algn_fill_row src/algn.c:458-533, edge cells 543-652,
// algn_fill_last_column 654): blocks where every lane is strictly inside its
// band, away from column 0 and the last column, run a branch-free cell on a
// packed cost table (substitution in the low half, insertion in the high
// half); the other blocks run the general cell.  Moves are stored in
// k_linear_fill's layout (2 bits per step, step 0 in the low bits).
#include "poyhip_device.h"
#include "poyhip_layout.h"

namespace poyhip {

constexpr int kMwMinLag = 10;  // blocks between consecutive stripes, at least

// Boundary rows of one workgroup: NW rows of (cost, delete count), each
// with 16 entries before column 0 and 112 after the last column (loads one
// block ahead run past the band).
POYHIP_HD inline int64_t lin_mw_row(int Lb) { return (int64_t)Lb + 128; }
POYHIP_HD inline int64_t lin_mw_ws_ints(int Lb, int NW) { return (int64_t)NW * 2 * lin_mw_row(Lb); }

typedef unsigned long long mw_mask;

// acc * 2 + (this lane's bit of m): one add with the lane mask as carry-in.
__device__ __forceinline__ unsigned mw_shift_in(unsigned acc, mw_mask m) {
  unsigned r;
  asm("v_addc_co_u32_e64 %0, vcc, %1, %1, %2" : "=v"(r) : "v"(acc), "s"(m) : "vcc");
  return r;
}

// Steps of stripe b (rows i0..ilast) of a round with half-width newk.
__device__ __forceinline__ void mw_stripe_geom(int b, int L1, int L2, int newk, bool full,
                                               int& i0, int& ilast, int& jlo, int& jhi,
                                               int& nsteps) {
  const int Dd = L2 - L1;
  i0 = 1 + (b << 6);
  ilast = imin(i0 + 63, L1 - 1);
  jlo = full ? 0 : imax(i0 - newk, 0);
  jhi = full ? L2 - 1 : imin(ilast + Dd + newk, L2 - 1);
  nsteps = (jhi - jlo + 1) + (ilast - i0);
}

// Row a of the packed table starts at 64 a + 8 (ctz(a) & 3): the rows of
// A, C, G and T are rotated by 0, 8, 16 and 24 banks, so the 16 lookups of
// two plain symbols fall in 16 different banks (with rows of 32 entries a
// wavefront's lanes, which hold different row and column symbols, read the
// same 4 banks).
__device__ __forceinline__ int mw_ag_row(int a) {
  return (a << 6) + (a ? ((__builtin_ctz(a) & 3) << 3) : 0);
}

// Shared memory of one workgroup.
template <int NW>
struct MwShared {
  int cost[1024];
  int ag[2048];              // cost(a,b) | cost(gap,b) << 16 at mw_ag_row(a) + b
  int prep[32];
  int tail[32];
  int2 src[NW][2][16];       // lane 0's row above, per block (double buffer)
  int2 out[NW][16];          // lane 63's results of a block
  uint8_t sym[NW][2][96];    // column symbols of a block window
  int pair[4];
  int fin[2];
  long long acc[NW][2];
  int scan[64 * NW];
};

// One pair on the workgroup (all threads call it with the same pid); kSw is
// the pair's swap flag, the tie order of the stored moves.
template <bool kDm, bool kDc, bool kSw, int NW>
__device__ __forceinline__ void lin_mw_pair(MwShared<NW>& S, const LinFillArgs& A,
                                                      int pid, int2* ring0) {
  int* const s_cost = S.cost;
  int* const s_ag = S.ag;
  int* const s_prep = S.prep;
  int* const s_tail = S.tail;
  auto& s_src = S.src;
  auto& s_out = S.out;
  auto& s_sym = S.sym;
  int* const s_fin = S.fin;
  auto& s_acc = S.acc;
  int* const s_scan = S.scan;
  const int tid = threadIdx.x;
  const int w = tid >> 6, lane = tid & 63;
  const int delta = A.cm->delta;
  {
    const uint8_t* __restrict__ sa = A.pt.blob + rows_off(A.pt, pid);
    const uint8_t* __restrict__ sb = A.pt.blob + cols_off(A.pt, pid);
    const int L1 = A.pt.len_a[pid], L2 = A.pt.len_b[pid];
    constexpr bool sw = kSw;
    const bool full = lin_full_plane(L1, L2, A.pt.deltawh[pid]);
    const int Dd = L2 - L1;
    const int rstride = (int)lin_mw_row(L2);  // int2 entries per boundary row
    int2* const rin = ring0 + (long long)((w + NW - 1) % NW) * rstride + 16;
    int2* const rout = ring0 + (long long)w * rstride + 16;
    uint8_t* dm = A.dm + (kDm ? A.pt.dm_off[pid] : 0);
    long long* stab = A.stab + A.pt.stab_off[pid];

    int T = (Dd + 1) * delta;
    int rounds = 0, newk = 0, cost = 0, gap_num = 0;
    long long cells = 0, slots = 0, cells1 = 0, slots1 = 0;
    if (A.phase == 2) {
      const PairResult r1 = A.res[pid];
      T = r1.pad;
      rounds = r1.rounds - 1;
      cells1 = r1.cells;
      slots1 = r1.slots;
    }
    int Tfin = T;
    for (;;) {
      Tfin = T;
      const int p = (T - Dd) / 2;
      // band rounds whose stop test cannot pass are skipped (see
      // k_linear_fill)
      if (!full && A.phase != 2 && p <= Dd + 1 && (2 * T - Dd) / 2 - L2 + 1 < 0) {
        rounds++;
        T *= 2;
        continue;
      }
      newk = full ? L1 - 1 : (p >= L1 ? L1 - 1 : p);
      const int lfr = full ? L2 : imin(L2 - L1 + 1 + p, L2);
      if (tid == 0) cells += lfr;
      if (L1 == 1) {
        // no DP row: the cost is row 0's last cell
        int sum = 0;
        for (int j = 1 + tid; j < L2; j += 64 * NW) sum += s_prep[sb[j]];
        for (int o = 32; o > 0; o >>= 1) sum += __shfl_xor(sum, o);
        if (lane == 0) s_acc[w][0] = sum;
        __syncthreads();
        long long tot = 0;
        for (int k = 0; k < NW; k++) tot += s_acc[k][0];
        __syncthreads();
        cost = (int)tot;
        gap_num = L2 - 1;
      } else {
      // ---- row 0 into wavefront NW-1's boundary row (stripe -1) ------------
      {
        int2* const r0 = ring0 + (long long)(NW - 1) * rstride + 16;
        const int per = (L2 + 64 * NW - 1) / (64 * NW);
        const int a0 = tid * per, a1 = imin(a0 + per, L2);
        int sum = 0;
        for (int j = a0; j < a1; j++) sum += j > 0 ? s_prep[sb[j]] : 0;
        s_scan[tid] = sum;
        __syncthreads();
        if (tid < 64) {
          // exclusive scan of the 64*NW partial sums by wavefront 0
          int run = 0;
          for (int k = 0; k < NW; k++) {
            const int v = s_scan[k * 64 + lane];
            int x = v;
            for (int o = 1; o < 64; o <<= 1) {
              const int y = __shfl_up(x, o);
              if (lane >= o) x += y;
            }
            s_scan[k * 64 + lane] = run + x - v;
            run += __shfl(x, 63);
          }
        }
        __syncthreads();
        int acc = s_scan[tid];
        for (int j = a0; j < a1; j++) {
          acc += j > 0 ? s_prep[sb[j]] : 0;
          r0[j] = make_int2(acc, 0);
        }
      }
      const int nstripes = (L1 - 1 + 63) >> 6;
      const int blast = nstripes - 1;
      const int jlo_last = full ? 0 : imax(1 + (blast << 6) - newk, 0);
      const int s_fin_step = (L2 - 1) - jlo_last + ((L1 - 2) & 63);
      // period: at least kMwMinLag, and long enough that a wavefront is done
      // with a stripe when its next one starts
      int nblk_max = 0;
      for (int b = 0; b < imin(nstripes, 3); b++) {
        int i0_, il_, jl_, jh_, ns_;
        mw_stripe_geom(b, L1, L2, newk, full, i0_, il_, jl_, jh_, ns_);
        nblk_max = imax(nblk_max, (ns_ + 15) >> 4);
      }
      {
        // widest stripe: stripe 0 starts at column 0 in a band, so the
        // widest ones are in the middle; the bound below covers every stripe
        const int wmax = full ? L2 : imin(Dd + 2 * newk + 64, L2);
        nblk_max = imax(nblk_max, (wmax + 63 + 15) >> 4);
      }
      const int P = imax(kMwMinLag, (nblk_max + NW - 1) / NW);
      const int Q = P * (nstripes - 1) + nblk_max;
      __syncthreads();  // row 0 written

      // ---- this wavefront's stripes ----------------------------------------
      int b = w;
      long long dmpos = 0;  // direction bytes of stripes before b
      for (int k = 0; k < w && k < nstripes; k++) {
        int i0_, il_, jl_, jh_, ns_;
        mw_stripe_geom(k, L1, L2, newk, full, i0_, il_, jl_, jh_, ns_);
        dmpos += (long long)((ns_ + 15) >> 4) * 256;
      }
      // per-stripe state (valid while b < nstripes)
      int i0 = 0, ilast = 0, jlo = 0, jhi = 0, nsteps = 0, nblk = 0;
      int c = 0, tl = 0, col0 = 0, arow = 0, agrow = 0, lo = 0, hi = 0;
      bool act = false, hasL = false, hasR = false;
      int fast_lo = 0, fast_hi = -1;  // block starts (steps) of the fast path
      int dmv = 0, ddc = 0, lmv = 0, ldc = 0, omv = 0, odc = 0;
      int rc = 0, rdc = 0;
      bool have_fin = false;
      auto enter = [&]() {
        mw_stripe_geom(b, L1, L2, newk, full, i0, ilast, jlo, jhi, nsteps);
        nblk = (nsteps + 15) >> 4;
        const int i = i0 + lane;
        act = i <= L1 - 1;
        const int ir = imin(i, L1 - 1);
        const int s1i = sa[ir];
        c = s_cost[(s1i << 5) + kGap];
        tl = s_tail[s1i];
        col0 = full ? c : tl;
        arow = s1i << 5;
        agrow = mw_ag_row(s1i);
        lo = 1 << 30;
        hi = -1;
        hasL = hasR = false;
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
        if (lane == 0) slots += (long long)nsteps * 64;
        // fast blocks: every active lane strictly inside its band, past
        // column 0 and before the last column, for all 16 steps
        const int flo = hasL ? lo + 1 : imax(lo, 1);
        const int fhi = hasR ? hi - 1 : imin(hi, L2 - 2);
        int a = act ? flo - jlo + lane : -(1 << 30);
        int z = act ? fhi - jlo + lane - 15 : (1 << 30);
        for (int o = 32; o > 0; o >>= 1) {
          a = imax(a, __shfl_xor(a, o));
          z = imin(z, __shfl_xor(z, o));
        }
        fast_lo = __builtin_amdgcn_readfirstlane(a);
        fast_hi = __builtin_amdgcn_readfirstlane(z);
        if (kDm && lane == 0) stab[b] = dmpos;
        // lane 0's diagonal for its first step: row i0-1 at column jlo-1
        dmv = 0;
        ddc = 0;
        if (jlo >= 1) {
          const int2 v = rin[jlo - 1];
          dmv = v.x;
          ddc = v.y;
        }
        lmv = ldc = omv = odc = 0;
        // sources and symbols of block 0
        if (lane < 16) s_src[w][0][lane] = rin[jlo + lane];
        for (int k = lane; k < 80; k += 64) {
          const int j = jlo - 63 + k;
          s_sym[w][0][k] = (unsigned)j < (unsigned)L2 ? sb[j] : 0;
        }
      };
      // a stripe is entered at its start block, not earlier: its first
      // sources are written by the previous stripe by then
      bool in_stripe = false;
      for (int qb = 0; qb < Q; qb++) {
        if (!in_stripe && b < nstripes && qb == P * b) {
          enter();
          in_stripe = true;
          __builtin_amdgcn_wave_barrier();
        }
        const int m = qb - P * b;  // local block of the current stripe
        if (in_stripe) {
          const int s0 = m << 4;
          const int buf = m & 1;
          // next block's sources: issued now, parked in LDS at the end
          int2 nsrc = make_int2(0, 0);
          uint8_t nsym0 = 0, nsym1 = 0;
          const bool more = m + 1 < nblk;
          if (more) {
            if (lane < 16) nsrc = rin[jlo + s0 + 16 + lane];
            const int j0 = jlo + s0 + 16 - 63 + lane;
            nsym0 = (unsigned)j0 < (unsigned)L2 ? sb[j0] : 0;
            const int j1 = j0 + 64;
            nsym1 = (lane < 16 && (unsigned)j1 < (unsigned)L2) ? sb[j1] : 0;
          }
          unsigned pack = 0;
          const uint8_t* symw = &s_sym[w][buf][63 - lane];
          const bool fast = s0 >= fast_lo && s0 <= fast_hi;
          const int4* const src4 = reinterpret_cast<const int4*>(s_src[w][buf]);
          if (fast) {
            // two halves of 8 steps: lane 0's sources (LDS broadcasts) and
            // the packed table entries of a half are read before its cells
#pragma unroll
            for (int h = 0; h < 2; h++) {
              int ag[8];
              int2 src[8];
#pragma unroll
              for (int u = 0; u < 4; u++) {
                const int4 v = src4[4 * h + u];
                src[2 * u] = make_int2(v.x, v.y);
                src[2 * u + 1] = make_int2(v.z, v.w);
              }
#pragma unroll
              for (int u = 0; u < 8; u++) ag[u] = s_ag[agrow + symw[8 * h + u]];
#pragma unroll
              for (int u = 0; u < 8; u++) {
                const int2 sv = src[u];
                const int umv = __builtin_amdgcn_update_dpp(sv.x, omv, 0x138, 0xF, 0xF, false);
                const int del = umv + c;
                const int ins = lmv + (ag[u] >> 16);
                const int aln = dmv + (ag[u] & 0xFFFF);
                const int val = imin(imin(del, ins), aln);
                const bool ea = aln == val;
                if (kDc) {
                  const int udc = __builtin_amdgcn_update_dpp(sv.y, odc, 0x138, 0xF, 0xF, false);
                  // backtrace_2d_gaps order: align, insert, delete
                  const int dc = ea ? ddc : (ins == val ? ldc : udc + 1);
                  ddc = udc;
                  odc = dc;
                  ldc = dc;
                }
                if (kDm) {
                  // 2-bit move under the swap priority: 0 align, 1 up, 2 left
                  const mw_mask E = __builtin_amdgcn_ballot_w64(ea);
                  const mw_mask X = __builtin_amdgcn_ballot_w64(sw ? ins == val : del == val);
                  const mw_mask up = sw ? (~E & ~X) : (~E & X);
                  const mw_mask left = sw ? (~E & X) : (~E & ~X);
                  pack = mw_shift_in(pack, up);
                  pack = mw_shift_in(pack, left);
                }
                if (lane == 63) s_out[w][8 * h + u] = make_int2(val, kDc ? odc : 0);
                dmv = umv;
                omv = val;
                lmv = val;
              }
            }
          } else {
#pragma unroll
            for (int u = 0; u < 16; u++) {
              const int st = s0 + u;
              const int j = jlo + st - lane;
              const int2 sv = s_src[w][buf][u];
              const int umv = __builtin_amdgcn_update_dpp(sv.x, omv, 0x138, 0xF, 0xF, false);
              const int udc = __builtin_amdgcn_update_dpp(sv.y, odc, 0x138, 0xF, 0xF, false);
              const int jc = symw[u];
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
              const int dc = (fl & kAlign) ? ddc : ((fl & kInsert) ? ldc : udc + 1);
              const bool inb = act && j >= lo && j <= hi;
              if (kDm) {
                const bool mup = inb && !(fl & kAlign) &&
                                 (sw ? !(fl & kInsert) : (fl & kDelete) != 0);
                const bool mleft = inb && !(fl & kAlign) &&
                                   (sw ? (fl & kInsert) != 0 : !(fl & kDelete));
                pack = mw_shift_in(pack, __builtin_amdgcn_ballot_w64(mup));
                pack = mw_shift_in(pack, __builtin_amdgcn_ballot_w64(mleft));
              }
              if (inb) {
                omv = val;
                odc = dc;
                lmv = val;
                ldc = dc;
              }
              if (lane == 63) s_out[w][u] = make_int2(omv, odc);
              if (b == blast && st == s_fin_step && lane == ((L1 - 2) & 63)) {
                rc = val;
                rdc = dc;
                have_fin = true;
              }
              dmv = umv;
              ddc = udc;
            }
          }
          if (kDm) {
            // the shifts put step 0 in the top bits: reverse to k_linear_fill's
            // layout (step u in bits 2u, 2u+1; 1 = up, 2 = left)
            const unsigned word = __builtin_bitreverse32(pack);
            reinterpret_cast<unsigned*>(dm + dmpos)[(m << 6) + lane] = word;
          }
          // lane 63's 16 results to this wavefront's boundary row
          __builtin_amdgcn_wave_barrier();
          if (lane < 16) {
            const int jc = jlo + s0 + lane - 63;
            if (jc >= 0 && jc < L2) rout[jc] = s_out[w][lane];
          }
          if (more) {
            if (lane < 16) s_src[w][buf ^ 1][lane] = nsrc;
            s_sym[w][buf ^ 1][lane] = nsym0;
            if (lane < 16) s_sym[w][buf ^ 1][64 + lane] = nsym1;
          }
          if (m + 1 == nblk) {
            // stripe done: move to b + NW
            for (int k = 0; k < NW; k++) {
              int i0_, il_, jl_, jh_, ns_;
              if (b + k < nstripes) {
                mw_stripe_geom(b + k, L1, L2, newk, full, i0_, il_, jl_, jh_, ns_);
                dmpos += (long long)((ns_ + 15) >> 4) * 256;
              }
            }
            b += NW;
            in_stripe = false;
          }
        }
        __syncthreads();
      }
      // final cell
      if (have_fin) {
        s_fin[0] = rc;
        s_fin[1] = rdc;
      }
      __syncthreads();
      cost = s_fin[0];
      gap_num = s_fin[1] + Dd;
      __syncthreads();
      }
      rounds++;
      if (full || A.phase == 2) break;
      const int newp = (2 * T - Dd) / 2;
      if ((gap_num + 1) < p || newp - L2 + 1 >= 0) break;
      T *= 2;
    }
    // statistics of the pair: sums over the workgroup
    for (int o = 32; o > 0; o >>= 1) {
      cells += __shfl_xor(cells, o);
      slots += __shfl_xor(slots, o);
    }
    if (lane == 0) {
      s_acc[w][0] = cells;
      s_acc[w][1] = slots;
    }
    __syncthreads();
    if (tid == 0) {
      long long tc = 0, ts = 0;
      for (int k = 0; k < NW; k++) {
        tc += s_acc[k][0];
        ts += s_acc[k][1];
      }
      PairResult r;
      r.cost = cost;
      r.gap_num = gap_num;
      r.newk = newk;
      r.rounds = rounds;
      r.full = full ? 1 : 0;
      r.pad = Tfin;
      // phase 2 repeats phase 1's last round: report phase 1's figures
      r.cells = A.phase == 2 ? cells1 : tc;
      r.slots = A.phase == 2 ? slots1 + ts : ts;
      A.res[pid] = r;
    }
    __syncthreads();
  }
}

template <bool kDm, bool kDc, int NW>
__global__ __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(4))) void k_linear_mw(LinFillArgs A) {
  __shared__ MwShared<NW> S;
  const int tid = threadIdx.x;
  for (int k = tid; k < 1024; k += 64 * NW) {
    const int c = A.cm->cost[k];
    S.cost[k] = c;
    S.ag[mw_ag_row(k >> 5) + (k & 31)] = (c & 0xFFFF) | (A.cm->cost[(kGap << 5) + (k & 31)] << 16);
  }
  if (tid < 32) {
    S.prep[tid] = A.cm->prepend[tid];
    S.tail[tid] = A.cm->tail[tid];
  }
  __syncthreads();
  int2* const ring0 = reinterpret_cast<int2*>(A.ws + (long long)blockIdx.x * A.ws_stride);
  for (;;) {
    if (tid == 0) S.pair[0] = atomicAdd(A.counter, 1);
    __syncthreads();
    const int q = S.pair[0];
    __syncthreads();  // the next grab rewrites S.pair
    if (q >= A.pt.n) break;
    const int pid = A.pt.order[q];
    if (A.pt.prio[pid])
      lin_mw_pair<kDm, kDc, true, NW>(S, A, pid, ring0);
    else
      lin_mw_pair<kDm, kDc, false, NW>(S, A, pid, ring0);
  }
}

template __global__ void k_linear_mw<true, true, 8>(LinFillArgs);
template __global__ void k_linear_mw<false, true, 8>(LinFillArgs);
template __global__ void k_linear_mw<true, false, 8>(LinFillArgs);

}  // namespace poyhip
