// Column-wise median lookups on aligned pairs (cm-style table lookups on the
// device) and the alignment cost recomputation behind Align.max_cost_2.
#include "poyhip_device.h"

namespace poyhip {

enum MedianOp { kAncestor = 0, kMedianNoGaps = 1, kMedianWithGaps = 2, kWorst = 3 };

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

}  // namespace poyhip
