// Union merge of POY's implied-alignment unions (src/union.c: union_merge
// 170, behind union_CAML_make 318) on gfx950.
//
// A merge walks the two children's aligned sequences and their unions from
// the right, prepending the new union's symbols and writing its offset and
// child-offset arrays; the steps depend on each other through five cursors,
// so one lane runs one merge and a batch runs many merges side by side (the
// unions of one tree level, say).  A union is a sequence right-aligned in
// `cap` bytes of the byte blob plus three int32 arrays of `cap` entries in
// the int blob (offset, union_c1, union_c2), their live part at the right
// end; the child unions need only their sequence and offsets.
#include "poyhip.h"
#include "poyhip_device.h"

namespace poyhip {

struct UnCursor {
  const uint8_t* s;  // sequence buffer (cap bytes)
  int cap, len;
  int end, position;  // offset-array and sequence cursors
};

struct UnOut {
  uint8_t* s;
  int32_t *off, *ca, *cb;
  int cap, len;
  int end, pca, pcb;
};

__device__ __forceinline__ int un_get(const UnCursor& u, int pos) {
  return u.s[u.cap - u.len + pos];
}

// seq_prepend (src/seq.c:155); false when the capacity is exhausted
__device__ __forceinline__ bool un_prepend(UnOut& c, int v) {
  if (c.len >= c.cap) return false;
  c.len++;
  c.s[c.cap - c.len] = (uint8_t)v;
  return true;
}

// union_prepend_item (src/union.c:79)
__device__ __forceinline__ bool un_item(UnCursor& src, UnOut& c, int v) {
  if (src.position >= 0) {
    v |= un_get(src, src.position);
    src.position--;
    src.end--;
  }
  return un_prepend(c, v);
}

// union_prepend_item with the symbol at src's position already loaded (sym)
__device__ __forceinline__ bool un_item_at(UnCursor& src, UnOut& c, int v, int sym) {
  if (src.position >= 0) {
    v |= sym;
    src.position--;
    src.end--;
  }
  return un_prepend(c, v);
}

// union_prepend_and_fill_items (src/union.c:136): the counts prep..1 of the
// items prepended since the last fill, their child offsets 0, then the
// column's symbol with the children's positions (the new union is cleared
// before the walk, so writes of 0 only move a cursor)
__device__ __forceinline__ bool un_fill(UnOut& c, int interm, int prep, int apos, int bpos) {
  if (c.end - prep < 0 || c.pca - prep < 0 || c.pcb - prep < 0) return false;
  for (int i = prep; i > 0; i--) c.off[c.end--] = i;
  c.pca -= prep;
  c.pcb -= prep;
  if (apos != 0) c.ca[c.pca] = apos;
  if (bpos != 0) c.cb[c.pcb] = bpos;
  c.pca--;
  c.pcb--;
  const bool ok = un_prepend(c, interm);
  c.end--;  // offset 0
  return ok;
}

// union_merge (src/union.c:170-274).  Returns false where the reference
// fails (the aligned sequences end while a union still has items) or the
// new union's capacity is exhausted.
__device__ __forceinline__ bool union_merge_one(const uint8_t* a, const uint8_t* b, const uint8_t* median,
                                int lena, UnCursor& au, const int32_t* aoff, UnCursor& bu,
                                const int32_t* boff, UnOut& c) {
  const int gap = kGap;
  int items = 0;
  int i = lena - 1;
  // begin cursors stay put; a union holds items while begin <= end
  const int abeg = au.len == 0 ? au.cap - 1 : au.cap - au.len;
  const int bbeg = bu.len == 0 ? bu.cap - 1 : bu.cap - bu.len;
  // the column's symbols do not depend on the cursors: each iteration loads
  // the next column's while it works on its own
  int cm = 0, ca = 0, cb = 0;
  if (i >= 0) {
    cm = median[i];
    ca = a[i];
    cb = b[i];
  }
  while (abeg <= au.end || bbeg <= bu.end) {
    if (c.end < 0 || c.pca < 0 || c.pcb < 0) return false;
    // the counts at both cursors and the symbols at both positions load
    // together; a symbol is loaded again when items moved its cursor
    const bool pa = au.position > -1, pb = bu.position > -1;
    const int ka = pa ? aoff[au.end] : 0;
    const int kb = pb ? boff[bu.end] : 0;
    int sa = pa ? un_get(au, au.position) : 0;
    int sb = pb ? un_get(bu, bu.position) : 0;
    // union_copy_non_homologous (src/union.c:118): the count at each
    // child's cursor is the number of its items copied across as they are
    items += ka;
    if (ka > 0) {
      for (int k = ka; k > 0; k--)
        if (!un_item(au, c, gap)) return false;
      if (au.position >= 0) sa = un_get(au, au.position);
    }
    items += kb;
    if (kb > 0) {
      for (int k = kb; k > 0; k--)
        if (!un_item(bu, c, gap)) return false;
      if (bu.position >= 0) sb = un_get(bu, bu.position);
    }
    if (i < 0) return false;  // "infinit loop" in the reference
    int interm = cm;
    const int ai = ca, bi = cb;
    if (i > 0) {
      cm = median[i - 1];
      ca = a[i - 1];
      cb = b[i - 1];
    }
    if (interm == gap) {
      if (i == 0) {
        if (!un_fill(c, gap, items, 0, 0)) return false;
        au.end--;
        bu.end--;
        return true;
      }
      bool ok;
      if (ai == gap) ok = un_item_at(bu, c, gap, sb);
      else if (bi == gap) ok = un_item_at(au, c, gap, sa);
      else {
        // union_prepend_pair (src/union.c:38)
        int v = gap;
        if (au.position >= 0) {
          v |= sa;
          au.position--;
          au.end--;
        }
        if (bu.position >= 0) {
          v |= sb;
          bu.position--;
          bu.end--;
        }
        ok = un_prepend(c, v);
      }
      if (!ok) return false;
      items += 1;
    } else {
      int apos, bpos;
      // union_move_left (src/union.c:161) as the three cases call it; in
      // the last case each guard tests the other union's cursor, as written
      if (ai == gap) {
        apos = 0;
        bpos = bu.position;
        if (bpos >= 0) {
          interm |= sb;
          bu.position--;
          bu.end--;
        }
      } else if (bi == gap) {
        bpos = 0;
        apos = au.position;
        if (apos >= 0) {
          interm |= sa;
          au.position--;
          au.end--;
        }
      } else {
        bpos = bu.position;
        apos = au.position;
        // a guard that passes with the union's own cursor below 0 reads
        // the byte before its live part, as the reference does
        if (bpos >= 0) {
          interm |= apos >= 0 ? sa : un_get(au, au.position);
          au.position--;
          au.end--;
        }
        if (apos >= 0) {
          interm |= bpos >= 0 ? sb : un_get(bu, bu.position);
          bu.position--;
          bu.end--;
        }
      }
      if (!un_fill(c, interm, items, apos, bpos)) return false;
      items = 0;
    }
    i--;
  }
  if (c.len == 0 || c.s[c.cap - c.len] != gap) {
    if (c.end < 0 || c.pca < 0 || c.pcb < 0 || !un_prepend(c, gap)) return false;
  }
  return true;
}

// every range the job names lies inside its blob
__device__ __forceinline__ bool un_in(int64_t off, int64_t n, int64_t size) {
  return off >= 0 && n >= 0 && off <= size && n <= size - off;
}

__device__ __forceinline__ bool un_job_fits(const poyhip_union_job& J, int64_t seq_bytes,
                                            int64_t n_ints) {
  return J.lena >= 0 && J.ua_len >= 0 && J.ub_len >= 0 && J.ua_cap >= 1 && J.ub_cap >= 1 &&
         J.c_cap >= 1 && J.ua_len <= J.ua_cap && J.ub_len <= J.ub_cap &&
         un_in(J.a, J.lena, seq_bytes) && un_in(J.b, J.lena, seq_bytes) &&
         un_in(J.median, J.lena, seq_bytes) && un_in(J.ua_s, J.ua_cap, seq_bytes) &&
         un_in(J.ub_s, J.ub_cap, seq_bytes) && un_in(J.c_s, J.c_cap, seq_bytes) &&
         un_in(J.ua_off, J.ua_cap, n_ints) && un_in(J.ub_off, J.ub_cap, n_ints) &&
         un_in(J.c_off, J.c_cap, n_ints) && un_in(J.c_c1, J.c_cap, n_ints) &&
         un_in(J.c_c2, J.c_cap, n_ints);
}

// One wavefront (one 64-thread block) takes `lanes` consecutive jobs, one
// lane per merge; fewer jobs per wavefront spread a level over more SIMDs,
// which hides more of each serial walk's load latency.  The whole wavefront
// first clears the new unions of its jobs with coalesced stores (entries the
// merge does not write are 0; the reference leaves them uninitialised), then
// each job's lane walks it.
__global__ __launch_bounds__(64) void k_union_merge(const poyhip_union_job* jobs, int n,
                                                    int lanes, uint8_t* seq, int64_t seq_bytes,
                                                    int32_t* ints, int64_t n_ints, int* c_len,
                                                    int* status) {
  const int lane = threadIdx.x;
  const int base = blockIdx.x * lanes;
  const int k = base + lane;
  const bool mine = lane < lanes && k < n;
  poyhip_union_job J;
  bool fits = false;
  if (mine) {
    J = jobs[k];
    fits = un_job_fits(J, seq_bytes, n_ints);
  }
  const lmask fit = ballot(fits);
  const int njobs = n - base < lanes ? n - base : lanes;
  for (int j = 0; j < njobs; j++) {
    if (!((fit >> j) & 1)) continue;
    const poyhip_union_job& Q = jobs[base + j];
    const int cap = Q.c_cap;
    uint8_t* cs = seq + Q.c_s;
    int32_t* co = ints + Q.c_off;
    int32_t* c1 = ints + Q.c_c1;
    int32_t* c2 = ints + Q.c_c2;
    for (int q = lane; q < cap; q += 64) {
      cs[q] = 0;
      co[q] = 0;
      c1[q] = 0;
      c2[q] = 0;
    }
  }
  // the clearing stores complete before any lane writes its union
  __syncthreads();
  if (!mine) return;
  if (!fits) {
    c_len[k] = 0;
    status[k] = POYHIP_EINVAL;
    return;
  }
  UnOut c;
  c.s = seq + J.c_s;
  c.off = ints + J.c_off;
  c.ca = ints + J.c_c1;
  c.cb = ints + J.c_c2;
  c.cap = J.c_cap;
  c.len = 0;
  c.end = c.pca = c.pcb = c.cap - 1;
  UnCursor au, bu;
  au.s = seq + J.ua_s;
  au.cap = J.ua_cap;
  au.len = J.ua_len;
  au.end = au.cap - 1;
  au.position = au.len - 1;
  bu.s = seq + J.ub_s;
  bu.cap = J.ub_cap;
  bu.len = J.ub_len;
  bu.end = bu.cap - 1;
  bu.position = bu.len - 1;
  const bool ok = union_merge_one(seq + J.a, seq + J.b, seq + J.median, J.lena, au,
                                  ints + J.ua_off, bu, ints + J.ub_off, c);
  c_len[k] = c.len;
  status[k] = ok ? POYHIP_OK : POYHIP_EINVAL;
}

}  // namespace poyhip
