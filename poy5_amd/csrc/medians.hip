// Column-wise median lookups on aligned pairs (cm-style table lookups on the
// device) and the alignment cost recomputation behind Align.max_cost_2.
#include <climits>

#include "poyhip.h"
#include "poyhip_device.h"

namespace poyhip {

enum MedianOp {
  kAncestor = 0,
  kMedianNoGaps = 1,
  kMedianWithGaps = 2,
  kWorst = 3,
  kClosest = 4,    // Align.closest's column map over an aligned pair
  kClosestEq = 5   // the same over two equal sequences (gap bits masked)
};

// Cost_matrix.Two_D.get_closest (src/cost_matrix.ml:1387-1424) for 5-bit
// sets with combinations: the state of b (its gap bit dropped unless a or b
// is the gap itself, or the gap when both carry one) cheapest against a,
// the lowest state winning ties; -1 for an empty set (the reference's
// "No bits on?" failure)
__device__ __forceinline__ int closest_state(const int* cost, int a, int b) {
  int bb;
  if (a == kGap || b == kGap) bb = b;
  else if ((a & kGap) && (b & kGap)) bb = kGap;
  else bb = b & ~kGap;
  if (bb == 0 || a == 0) return -1;
  int best = a, bc = INT_MAX;
#pragma unroll
  for (int bit = 0; bit < 5; bit++) {
    const int x = 1 << bit;
    if (bb & x) {
      const int c = cost[(a << 5) + x];
      if (c < bc) {
        bc = c;
        best = x;
      }
    }
  }
  return best;
}

// one column of Align.closest (src/sequence.ml:1199-1228): the two
// sequences as the column map sees them (kClosestEq masks the gap bit of
// every column but the first, as the branch for equal sequences does)
__device__ __forceinline__ int closest_column(const int* cost, int op, int x, int y, int i) {
  if (op == kClosestEq && i > 0) {
    x &= ~kGap;
    y &= ~kGap;
  }
  return closest_state(cost, x, y);
}

// One thread walks one aligned pair.  out receives the result right-aligned
// in a region of `cap` bytes; res[0] = length (or cost for kWorst), res[1] =
// error flag (1: a median entry was 0, algn_ancestor_2's failwith).
__device__ void pair_median_serial(const CmDev* cm, const uint8_t* a, const uint8_t* b, int len,
                                   int op, uint8_t* out, int cap, int* res) {
  const uint8_t* med = cm->median;
  int n = 0, err = 0;
#define PUT(v) out[cap - 1 - n++] = (uint8_t)(v)
  if (op == kWorst) {
    // algn_calculate_from_2_aligned (src/algn.c:3003-3089) on `worst`,
    // uselevel 0 with combinations
    int r = 0, state = 0;
    const int go = cm->gap_open;
    int i0 = (len > 0 && (kGap & a[0]) && (kGap & b[0])) ? 1 : 0;
    for (int i = i0; i < len; i++) {
      const int x = a[i], y = b[i];
      if (state == 0) {
        if ((x & kGap) && !(y & kGap)) { r += go; state = 1; }
        else if ((y & kGap) && !(x & kGap)) { r += go; state = 2; }
      } else if (state == 1) {
        if (!(x & kGap)) {
          if ((y & kGap) && !(x & kGap)) { r += go; state = 2; }
          else state = 0;
        }
      } else {
        if (!(y & kGap)) {
          if (x & kGap) { r += go; state = 1; }
          else state = 0;
        }
      }
      r += cm->worst[(x << 5) + y];
    }
    res[0] = r;
    res[1] = 0;
    return;
  }
  if (op == kMedianWithGaps) {
    // seq_get_median_2d_with_gaps (src/seq.c:241)
    for (int i = len - 1; i >= 0; i--) PUT(med[(a[i] << 5) + b[i]]);
    res[0] = n;
    res[1] = 0;
    return;
  }
  if (op == kClosest || op == kClosestEq) {
    // the column map, then remove_gaps2 (src/sequence.ml:209): symbols
    // equal to the gap dropped and one leading gap prepended
    for (int i = len - 1; i >= 0; i--) {
      const int v = closest_column(cm->cost, op, a[i], b[i], i);
      if (v < 0) { err = 1; break; }
      if (v != kGap) PUT(v);
    }
    if (err) { res[0] = 0; res[1] = 1; return; }
    PUT(kGap);
    res[0] = n;
    res[1] = 0;
    return;
  }
  if (op == kMedianNoGaps) {
    // seq_get_median_2d_no_gaps (src/seq.c:259)
    for (int i = len - 1; i >= 0; i--) {
      const int m = med[(a[i] << 5) + b[i]];
      if (m != kGap) PUT(m);
    }
    PUT(kGap);
    res[0] = n;
    res[1] = 0;
    return;
  }
  // algn_ancestor_2 (src/algn.c:3603-3626)
  const bool keep_gaps = cm->combinations && cm->cost_model == 1;
  for (int i = len - 1; i >= 0; i--) {
    const int m = med[(a[i] << 5) + b[i]];
    if (m == 0) { err = 1; break; }
    if (!keep_gaps) { if (m != kGap) PUT(m); }
    else PUT(m);
  }
  if (err) { res[0] = 0; res[1] = 1; return; }
  if (n == 0) {
    // the reference reads sm[0] past the end here; both branches give [gap]
    PUT(kGap);
  } else if (!cm->combinations || (cm->cost_model != 1 && out[cap - n] != kGap)) {
    PUT(kGap);
  } else if (cm->combinations) {
    // algn_correct_blocks_affine (src/algn.c:3561) then algn_remove_gaps
    // (3539); sm[k] = out[cap - n + k]
    int extending = 0, inside = 0, prev_block = 0;
    uint8_t* sm = out + cap - n;
    for (int i = 0; i < n; i++) {
      const int ab = a[i], bb = b[i];
      int sb = sm[i];
      if (!inside && (!(ab & kGap) || !(bb & kGap))) inside = 0;
      else if (inside && ((!(ab & kGap)) || (!(bb & kGap)))) inside = 0;
      else if (((ab & kGap) || (bb & kGap)) && ((ab != kGap) || (bb != kGap))) inside = 1;
      else inside = 0;
      if (((kGap & ab) || (kGap & bb)) && (!(sb & kGap)) && !extending) {
        prev_block = inside;
        extending = 1;
      } else if ((kGap & ab) && (kGap & bb) && (sb & kGap) && (sb != kGap) && extending &&
                 inside && !prev_block) {
        sb = (~kGap) & sb;
        prev_block = 0;
      } else if ((kGap & ab) && (kGap & bb) && (1 == extending)) {
        prev_block = inside;
        extending = 0;
      }
      sm[i] = (uint8_t)sb;
    }
    // remove gaps, keeping order, right-aligned, then the leading gap
    int w = cap - 1;
    for (int i = n - 1; i >= 0; i--) {
      const uint8_t v = sm[i];
      if (v != kGap) out[w--] = v;
    }
    out[w--] = kGap;
    n = cap - 1 - w;
  }
#undef PUT
  res[0] = n;
  res[1] = 0;
}

__global__ void k_pair_median(const CmDev* cm, const uint8_t* a, const uint8_t* b, int len,
                              int op, uint8_t* out, int cap, int* res) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  pair_median_serial(cm, a, b, len, op, out, cap, res);
}

// Many column maps at once (the medians and closest sequences of a batch of
// aligned pairs): one wavefront per job, 64 columns a step from the right
// end; the symbols a map keeps are compacted with a ballot, so each step
// reads and writes 64 consecutive bytes.  Inputs and outputs share the byte
// blob `seq`; the result lies right-aligned in the job's `cap` bytes.
__device__ __forceinline__ bool post_in(int64_t off, int64_t n, int64_t size) {
  return off >= 0 && n >= 0 && off <= size && n <= size - off;
}

__global__ __launch_bounds__(256) void k_pairs_post(const CmDev* cm, int op,
                                                    const poyhip_post_job* jobs, int n,
                                                    uint8_t* seq, int64_t seq_bytes, int* out_len,
                                                    int* status) {
  __shared__ int s_cost[1024];
  __shared__ uint8_t s_med[1024];
  for (int q = threadIdx.x; q < 1024; q += 256) {
    s_cost[q] = cm->cost[q];
    s_med[q] = cm->median[q];
  }
  __syncthreads();
  const int lane = threadIdx.x & 63;
  const int k = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (k >= n) return;
  const poyhip_post_job J = jobs[k];
  const bool gaps_out = op != kMedianWithGaps;  // a leading gap is prepended
  if (!(J.len >= 0 && J.cap > J.len && post_in(J.a, J.len, seq_bytes) &&
        post_in(J.b, J.len, seq_bytes) && post_in(J.out, J.cap, seq_bytes))) {
    if (lane == 0) {
      out_len[k] = 0;
      status[k] = POYHIP_EINVAL;
    }
    return;
  }
  const uint8_t* a = seq + J.a;
  const uint8_t* b = seq + J.b;
  // the w-th symbol from the right goes to out[last - w]: indices stay
  // non-negative (a negated index here was once widened as unsigned)
  uint8_t* out = seq + J.out;
  const int last = J.cap - 1;
  const lmask below = (1ull << lane) - 1;  // lanes holding later columns
  int w = 0;
  bool bad = false;
  for (int base = J.len - 1; base >= 0; base -= 64) {
    const int i = base - lane;
    int v = kGap;
    bool keep = false;
    if (i >= 0) {
      const int x = a[i], y = b[i];
      if (op == kClosest || op == kClosestEq) {
        v = closest_column(s_cost, op, x, y, i);
        bad |= v < 0;
        keep = v > 0 && v != kGap;
      } else {
        v = s_med[(x << 5) + y];
        keep = op == kMedianWithGaps || v != kGap;
      }
    }
    const lmask m = ballot(keep);
    if (keep) out[last - (w + __popcll(m & below))] = (uint8_t)v;
    w += __popcll(m);
  }
  const bool failed = ballot(bad) != 0;
  if (gaps_out && lane == 0) out[last - w] = kGap;
  if (gaps_out) w++;
  if (lane == 0) {
    out_len[k] = failed ? 0 : w;
    status[k] = failed ? POYHIP_EINVAL : POYHIP_OK;
  }
}

}  // namespace poyhip
