// Three-sequence cube of POY's algn_nw_3d (src/algn.c:2977) for gfx950.
//
// algn_fill_cube (src/algn.c:2593-2763) fills planes over s1, rows over s2
// and cells over s3.  Cell (i, j, k) keeps ONE direction: the candidates are
// tried in the fixed order P3, P1, P2 (fill_parallel, 2552), then S3, S1, S2
// (fill_moved, 2524, k >= 1), then SS along the row (k >= 1), each replacing
// the running value only when strictly smaller.
//
// Row pointers.  The reference advances its "diag", "upper" and "prev" row
// pointers s2_len-1 times per plane while the plane advances s2_len rows, so
// for plane i >= 1 the rows it reads lag their geometric neighbours.  With
// R = i*L2 + j the global row number, the rows read are
//   plane 0, row j >= 1 : prev = R - 1
//   plane i >= 1, row 0 : diag = (i-1)*(L2-1)
//   plane i >= 1, row j : diag = (i-1)*(L2-1) + j-1, upper = diag + 1,
//                         prev = i*(L2-1) + j-1
// All of them are at most L1+L2-1 rows behind R, so a wavefront keeps the
// last L1+L2 rows in a ring in global memory instead of the whole cube.
//
// Mapping: one wavefront per triple; a row is cut into segments of 1024
// cells and lane l owns cells 16l..16l+15 of a segment, for every row.  A
// lane only ever reads ring cells it wrote itself; the k-1 neighbours of its
// first cell come from lane l-1 through DPP.  The SS recurrence
// mm[k] = min(x[k], mm[k-1] + ggs3[k]) is a min-plus scan: each lane folds
// its 16 cells into a map y -> min(a, y + b), a Kogge-Stone scan over the
// wavefront composes the maps, and a second pass applies the incoming value.
//
// Directions: one nibble per cell, nibble order = cell order; a row of one
// segment keeps the lane blocks below L3 (8 bytes each), a longer row 512
// bytes per segment.  The codes number the moves in the order they are tried
// (1 P3, 2 P1, 3 P2, 4 S3, 5 S1, 6 S2, 7 SS), so that among equal values the
// first move tried has the smallest code.
#pragma once
#include "poyhip_device.h"
#include "poyhip_layout.h"

namespace poyhip {

// struct cm_3d (src/cm.h:253-280) cost and median tables addressed by
// 5-bit codes: index ((a << 5) + b) << 5 + c.
struct Cm3Dev {
  int cost[32768];
  uint8_t median[32768];
};

enum : int { kD3P3 = 1, kD3P1 = 2, kD3P2 = 3, kD3S3 = 4, kD3S1 = 5, kD3S2 = 6, kD3SS = 7 };

constexpr int kCubeCpl = 16;                  // cells per lane per segment
constexpr int kCubeSeg = 64 * kCubeCpl;       // cells per segment
constexpr int kCubeInf = 1 << 28;

struct TripleTable {
  const uint8_t* blob;
  const int64_t* off1;
  const int64_t* off2;
  const int64_t* off3;
  const int* len1;
  const int* len2;
  const int* len3;
  const int64_t* dm_off;   // byte offset of the triple's direction store
  const int64_t* out_off;  // byte offset of the triple's output block
  const int* order;        // triples of this launch, largest first
  int n;
  int ntriples;
};

struct CubeFillArgs {
  TripleTable tt;
  const Cm3Dev* cm3;
  uint8_t* dm;             // null for cost-only fills
  int* ring;               // one ring per resident wavefront
  long long ring_stride;   // ints per wavefront
  int* cost;               // per triple
  int* counter;
  int seg_lo = 1, seg_hi = 1 << 30;  // the launch takes triples of seg_lo..seg_hi segments
};

__device__ __forceinline__ int cube_c3(const Cm3Dev* c, int a, int b, int x) {
  return c->cost[(((a << 5) + b) << 5) + x];
}

__device__ __forceinline__ void load16(const int* p, int v[kCubeCpl]) {
  const int4* q = reinterpret_cast<const int4*>(p);
#pragma unroll
  for (int u = 0; u < 4; u++) {
    const int4 t = q[u];
    v[4 * u] = t.x;
    v[4 * u + 1] = t.y;
    v[4 * u + 2] = t.z;
    v[4 * u + 3] = t.w;
  }
}

__device__ __forceinline__ void store16(int* p, const int v[kCubeCpl]) {
  int4* q = reinterpret_cast<int4*>(p);
#pragma unroll
  for (int u = 0; u < 4; u++) q[u] = make_int4(v[4 * u], v[4 * u + 1], v[4 * u + 2], v[4 * u + 3]);
}

// One step of the min-plus map scan: the maps (pa, pb) of the lanes before
// are applied first, then the lane's own (fa, fb):
//   y -> min(fa, min(pa, y + pb) + fb) = min(min(fa, pa + fb), y + pb + fb).
// Lanes without a source keep the identity map (kCubeInf, 0).
template <int kCtrl, int kRowMask>
__device__ __forceinline__ void cube_scan_step(int& fa, int& fb) {
  const int pa = __builtin_amdgcn_update_dpp(kCubeInf, fa, kCtrl, kRowMask, 0xF, false);
  const int pb = __builtin_amdgcn_update_dpp(0, fb, kCtrl, kRowMask, 0xF, false);
  fa = imin(fa, pa + fb);
  fb = imin(pb + fb, kCubeInf);
}

// Inclusive scan of the maps over the 64 lanes.
__device__ __forceinline__ void cube_scan_minplus(int& fa, int& fb) {
  cube_scan_step<0x111, 0xF>(fa, fb);  // row_shr:1
  cube_scan_step<0x112, 0xF>(fa, fb);  // row_shr:2
  cube_scan_step<0x114, 0xF>(fa, fb);  // row_shr:4
  cube_scan_step<0x118, 0xF>(fa, fb);  // row_shr:8
  cube_scan_step<0x142, 0xA>(fa, fb);  // row_bcast:15 into rows 1 and 3
  cube_scan_step<0x143, 0xC>(fa, fb);  // row_bcast:31 into rows 2 and 3
}

// Row kinds of algn_fill_cube.
enum : int { kRowFirst = 0, kRowPlane0 = 1, kRowPlaneStart = 2, kRowInner = 3 };

// Cells of one row segment of algn_fill_cube: the six plane/line moves in
// the reference's order, then the SS run along the row as a min-plus scan.
// cU/cP/cD are the k-1 inputs of lane 0's first cell (previous segment),
// cY the final value before it.  Returns the row values in X, the direction
// nibbles in lo/hi and the lane's last value in y.
__device__ __forceinline__ void cube_row(int kind, int k0, const int cc[kCubeCpl],
                                         const int GG[kCubeCpl], const int U[kCubeCpl],
                                         const int P[kCubeCpl], const int D[kCubeCpl], int cU,
                                         int cP, int cD, int cY, int s1g0, int g0, int s120,
                                         const int* s_s1g, const int* s_gs2, const int* s_s12,
                                         int X[kCubeCpl], uint32_t& lo, uint32_t& hi, int& y) {
  const int lane = threadIdx.x;
  int Dir[kCubeCpl];
  const int nU = wave_shr1(U[kCubeCpl - 1], cU);
  const int nP = wave_shr1(P[kCubeCpl - 1], cP);
  const int nD = wave_shr1(D[kCubeCpl - 1], cD);
#pragma unroll
  for (int m = 0; m < kCubeCpl; m++) {
    const int k = k0 + m;
    const int c = cc[m];
    const int um1 = m == 0 ? nU : U[m - 1];
    const int pm1 = m == 0 ? nP : P[m - 1];
    const int dm1 = m == 0 ? nD : D[m - 1];
    int v, d;
    if (kind == kRowFirst) {
      v = k == 0 ? 0 : kCubeInf;
      d = k == 0 ? kD3S2 : kD3SS;
    } else if (kind == kRowPlane0) {
      v = P[m] + g0;
      d = kD3P1;
      const int t1 = pm1 + s_gs2[c];
      if (k >= 1 && t1 < v) { v = t1; d = kD3S1; }
    } else if (kind == kRowPlaneStart) {
      v = D[m] + s1g0;
      d = kD3P3;
      const int t1 = dm1 + s_s1g[c];
      if (k >= 1 && t1 < v) { v = t1; d = kD3S3; }
    } else {
      v = U[m] + s1g0;
      d = kD3P3;
      int t1 = P[m] + g0;
      if (t1 < v) { v = t1; d = kD3P1; }
      t1 = D[m] + s120;
      if (t1 < v) { v = t1; d = kD3P2; }
      if (k >= 1) {
        t1 = um1 + s_s1g[c];
        if (t1 < v) { v = t1; d = kD3S3; }
        t1 = pm1 + s_gs2[c];
        if (t1 < v) { v = t1; d = kD3S1; }
        t1 = dm1 + s_s12[c];
        if (t1 < v) { v = t1; d = kD3S2; }
      }
    }
    X[m] = v;
    Dir[m] = d;
  }
  // SS scan, pass 1: the lane's map y -> min(fa, y + fb)
  const bool has_k0 = (k0 == 0);
  int fa = kCubeInf, fb = 0;
#pragma unroll
  for (int m = 0; m < kCubeCpl; m++) {
    if (has_k0 && m == 0) {
      fa = X[0];
      fb = kCubeInf;  // the chain starts at k = 0: y is ignored
    } else {
      fa = imin(X[m], fa + GG[m]);
      fb = imin(fb + GG[m], kCubeInf);
    }
  }
  // inclusive scan of the maps over the wavefront (DPP: row_shr inside rows
  // of 16 lanes, then row_bcast:15 / row_bcast:31)
  cube_scan_minplus(fa, fb);
  const int ia = wave_shr1(fa, kCubeInf);
  const int ib = wave_shr1(fb, 0);
  y = lane == 0 ? cY : imin(ia, cY + ib);
  // pass 2: apply the incoming value
  lo = 0;
  hi = 0;
#pragma unroll
  for (int m = 0; m < kCubeCpl; m++) {
    int v = X[m], d = Dir[m];
    if (!(has_k0 && m == 0)) {
      const int t1 = y + GG[m];
      if (t1 < v) { v = t1; d = kD3SS; }
    }
    X[m] = v;
    y = v;
    if (m < 8) lo |= (uint32_t)d << (4 * m);
    else hi |= (uint32_t)d << (4 * (m - 8));
  }
}

// Triples whose rows fit one segment (L3 <= 1024).  Besides the ring, the
// row pointers give registers most inputs: a plane-0 row reads the row just
// computed, and an inner row's diag row is the previous row's upper row (the
// plane-start row's input for row 1).  The two rows an inner row still
// loads, and its precalc entries, are fetched one row ahead: row R+1 of
// plane i >= 1 reads rows at least i+1 behind it, all stored before row R.
template <bool kDm>
__device__ int cube_fast(const CubeFillArgs& A, const uint8_t* s1, const uint8_t* s2, int L1,
                         int L2, int L3, int g3, int* ring, uint8_t* dm, const int cc[kCubeCpl],
                         const int GG[kCubeCpl], int* s_s1g, int* s_gs2, int* s_s12) {
  const int lane = threadIdx.x;
  const int k0 = lane * kCubeCpl;
  const int NR = L1 + L2;
  auto row_ptr = [&](int r) { return ring + (long long)(r % NR) * kCubeSeg + k0; };
  int U[kCubeCpl], P[kCubeCpl], D[kCubeCpl], nU[kCubeCpl], nP[kCubeCpl];
#pragma unroll
  for (int m = 0; m < kCubeCpl; m++) U[m] = P[m] = D[m] = nU[m] = nP[m] = 0;
  // precalc entries of the next row (lanes 0..31 hold entry `lane`)
  int tG = 0, tA = 0, nG0 = 0, nS120 = 0;
  auto fetch_tables = [&](int i, int j) {
    const int a = s1[i] & 31, b = s2[j] & 31;
    if (j >= 1) {
      tG = cube_c3(A.cm3, kGap, b, lane & 31);
      nG0 = cube_c3(A.cm3, kGap, b, g3);
      if (i >= 1) {
        tA = cube_c3(A.cm3, a, b, lane & 31);
        nS120 = cube_c3(A.cm3, a, b, g3);
      }
    }
  };
  fetch_tables(0, L2 > 1 ? 1 : 0);
  int s1g0 = 0, last = 0;
  for (int i = 0; i < L1; i++) {
    if (i >= 1) {
      const int a = s1[i] & 31;
      __builtin_amdgcn_wave_barrier();
      if (lane < 32) s_s1g[lane] = cube_c3(A.cm3, a, kGap, lane);
      s1g0 = cube_c3(A.cm3, a, kGap, g3);
    }
    for (int j = 0; j < L2; j++) {
      const int R = i * L2 + j;
      const int kind = i == 0 ? (j == 0 ? kRowFirst : kRowPlane0)
                              : (j == 0 ? kRowPlaneStart : kRowInner);
      int g0 = 0, s120 = 0;
      if (j >= 1) {
        // this row's precalc entries, fetched during the previous row
        __builtin_amdgcn_wave_barrier();
        if (lane < 32) {
          s_gs2[lane] = tG;
          if (i >= 1) s_s12[lane] = tA;
        }
        g0 = nG0;
        s120 = nS120;
      }
      __builtin_amdgcn_wave_barrier();
      if (kind == kRowPlaneStart) {
        load16(row_ptr((i - 1) * (L2 - 1)), D);
      } else if (kind == kRowInner) {
#pragma unroll
        for (int m = 0; m < kCubeCpl; m++) {
          U[m] = nU[m];
          P[m] = nP[m];
        }
      }
      // fetch the next row's inputs and precalc entries
      const int Rn = R + 1;
      if (Rn < L1 * L2) {
        const int in = Rn / L2, jn = Rn - in * L2;
        if (in >= 1 && jn >= 1) {
          load16(row_ptr((in - 1) * (L2 - 1) + jn), nU);
          load16(row_ptr(in * (L2 - 1) + jn - 1), nP);
        }
        if (jn >= 1) fetch_tables(in, jn);
      }
      int X[kCubeCpl];
      uint32_t lo, hi;
      int y;
      cube_row(kind, k0, cc, GG, U, P, D, 0, 0, 0, kCubeInf, s1g0, g0, s120, s_s1g, s_gs2, s_s12,
               X, lo, hi, y);
      store16(row_ptr(R), X);
      if (kDm && k0 < L3)
        reinterpret_cast<uint2*>(dm + (long long)R * cube_dm_row_bytes(L3))[lane] = make_uint2(lo, hi);
      // register chains for the next row
      if (kind == kRowInner) {
#pragma unroll
        for (int m = 0; m < kCubeCpl; m++) D[m] = U[m];
      }
      if (i == 0) {
#pragma unroll
        for (int m = 0; m < kCubeCpl; m++) P[m] = X[m];
      }
      if (R == L1 * L2 - 1) {
        // the last cell, k = L3 - 1, belongs to lane (L3 - 1) / 16
        const int mm = (L3 - 1) % kCubeCpl;
        int sel = 0;
#pragma unroll
        for (int m = 0; m < kCubeCpl; m++) sel = (m == mm) ? X[m] : sel;
        last = __shfl(sel, (L3 - 1) / kCubeCpl);
      }
    }
  }
  return last;
}

// kOneSeg: the launch takes only triples with L3 <= 1024 (cube_fast); the
// other instantiation takes only the longer ones.  Each instantiation gets
// the registers of its own path.
template <bool kDm, bool kOneSeg>
__global__ __launch_bounds__(64) void k_cube_fill(CubeFillArgs A) {
  // precalc rows of the current plane / row indexed by the s3 code
  // (cm_precalc_4algn_3d, src/cm.c:1390-1412)
  __shared__ int s_gg[32];   // (gap, gap, c)
  __shared__ int s_s1g[32];  // (s1[i], gap, c)
  __shared__ int s_gs2[32];  // (gap, s2[j], c)
  __shared__ int s_s12[32];  // (s1[i], s2[j], c)
  const int lane = threadIdx.x;
  if (lane < 32) s_gg[lane] = cube_c3(A.cm3, kGap, kGap, lane);
  __syncthreads();

  for (;;) {
    const int q = wave_grab(A.counter);
    if (q >= A.tt.n) break;
    const int t = A.tt.order[q];
    const uint8_t* s1 = A.tt.blob + A.tt.off1[t];
    const uint8_t* s2 = A.tt.blob + A.tt.off2[t];
    const uint8_t* s3 = A.tt.blob + A.tt.off3[t];
    const int L1 = A.tt.len1[t], L2 = A.tt.len2[t], L3 = A.tt.len3[t];
    const int nseg = cube_nseg(L3);
    if ((nseg == 1) != kOneSeg || nseg < A.seg_lo || nseg > A.seg_hi) continue;
    const int row_ints = nseg * kCubeSeg;
    const int NR = L1 + L2;
    int* ring = A.ring + (long long)blockIdx.x * A.ring_stride;
    uint8_t* dm = kDm ? A.dm + A.tt.dm_off[t] : nullptr;
    // codes are 5-bit sets; the mask keeps table reads inside the tables
    const int g3 = s3[0] & 31;  // s3[0] prices the k = 0 entries (the leading gap)

    // s3 codes of the lane's cells (cells past L3 are priced as gaps) and
    // their (gap, gap, s3[k]) entries; reloaded per segment only when a row
    // has several segments
    int cc[kCubeCpl], GG[kCubeCpl];
    auto load_codes = [&](int sg) {
#pragma unroll
      for (int m = 0; m < kCubeCpl; m++) {
        const int k = sg * kCubeSeg + lane * kCubeCpl + m;
        cc[m] = k < L3 ? (s3[k] & 31) : kGap;
        GG[m] = s_gg[cc[m]];
      }
    };
    if constexpr (kOneSeg) {
      load_codes(0);
      const int lv =
          cube_fast<kDm>(A, s1, s2, L1, L2, L3, g3, ring, dm, cc, GG, s_s1g, s_gs2, s_s12);
      if (lane == 0) A.cost[t] = lv;
      continue;
    }
    int last_val = 0;
    for (int i = 0; i < L1; i++) {
      const int a = s1[i] & 31;
      if (i >= 1) {
        __syncthreads();
        if (lane < 32) s_s1g[lane] = cube_c3(A.cm3, a, kGap, lane);
      }
      for (int j = 0; j < L2; j++) {
        const int R = i * L2 + j;
        const int kind = i == 0 ? (j == 0 ? kRowFirst : kRowPlane0)
                                : (j == 0 ? kRowPlaneStart : kRowInner);
        const int b = s2[j] & 31;
        __syncthreads();
        if (lane < 32 && j >= 1) {
          s_gs2[lane] = cube_c3(A.cm3, kGap, b, lane);
          if (i >= 1) s_s12[lane] = cube_c3(A.cm3, a, b, lane);
        }
        __syncthreads();
        // k = 0 entries of the precalc rows
        const int s1g0 = i >= 1 ? cube_c3(A.cm3, a, kGap, g3) : 0;
        const int g0 = j >= 1 ? cube_c3(A.cm3, kGap, b, g3) : 0;
        const int s120 = (i >= 1 && j >= 1) ? cube_c3(A.cm3, a, b, g3) : 0;
        // rows read (global row numbers)
        int rU = 0, rP = 0, rD = 0;
        if (kind == kRowPlane0) {
          rP = R - 1;
        } else if (kind == kRowPlaneStart) {
          rD = (i - 1) * (L2 - 1);
        } else if (kind == kRowInner) {
          rD = (i - 1) * (L2 - 1) + j - 1;
          rU = rD + 1;
          rP = i * (L2 - 1) + j - 1;
        }
        const int* pU = ring + (long long)(rU % NR) * row_ints;
        const int* pP = ring + (long long)(rP % NR) * row_ints;
        const int* pD = ring + (long long)(rD % NR) * row_ints;
        int* pC = ring + (long long)(R % NR) * row_ints;
        // carries from the previous segment of this row: the k-1 inputs of
        // lane 0's first cell and the SS scan value
        int cU = 0, cP = 0, cD = 0, cY = kCubeInf;
        for (int sg = 0; sg < nseg; sg++) {
          const int k0 = sg * kCubeSeg + lane * kCubeCpl;
          if (nseg > 1) load_codes(sg);
          int U[kCubeCpl] = {0}, P[kCubeCpl] = {0}, D[kCubeCpl] = {0};
          if (kind == kRowInner) {
            load16(pU + k0, U);
            load16(pP + k0, P);
            load16(pD + k0, D);
          } else if (kind == kRowPlane0) {
            load16(pP + k0, P);
          } else if (kind == kRowPlaneStart) {
            load16(pD + k0, D);
          }
          int X[kCubeCpl];
          uint32_t lo, hi;
          int y;
          cube_row(kind, k0, cc, GG, U, P, D, cU, cP, cD, cY, s1g0, g0, s120, s_s1g, s_gs2, s_s12,
                   X, lo, hi, y);
          store16(pC + k0, X);
          if (kDm) {
            uint2* w = reinterpret_cast<uint2*>(dm + (long long)R * (nseg * 512) + sg * 512);
            w[lane] = make_uint2(lo, hi);
          }
          // carries into the next segment come from lane 63
          cU = __shfl(U[kCubeCpl - 1], 63);
          cP = __shfl(P[kCubeCpl - 1], 63);
          cD = __shfl(D[kCubeCpl - 1], 63);
          cY = __shfl(y, 63);
          if (R == L1 * L2 - 1) {
            const int kl = L3 - 1;
            if (kl >= sg * kCubeSeg && kl < (sg + 1) * kCubeSeg) {
              const int owner = (kl - sg * kCubeSeg) / kCubeCpl;
              const int mm = kl % kCubeCpl;
              int sel = 0;
#pragma unroll
              for (int m = 0; m < kCubeCpl; m++) sel = (m == mm) ? X[m] : sel;
              last_val = __shfl(sel, owner);
            }
          }
        }
        __threadfence_block();
      }
    }
    if (lane == 0) A.cost[t] = last_val;
  }
}

struct CubeBtArgs {
  TripleTable tt;
  const Cm3Dev* cm3;
  const uint8_t* dm;
  uint8_t* out;
  int* outlen;  // 4 per triple: r1, r2, r3, median_3 lengths (-1: bad walk)
};

// backtrack_3d (src/algn.c:3345-3421) followed by algn_get_median_3d
// (3641-3654) on the aligned sequences.  Outputs are written right to left
// into four regions of capacity L1+L2+L3.
__global__ __launch_bounds__(64) void k_cube_backtrack(CubeBtArgs A) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= A.tt.n) return;
  const int t = A.tt.order[q];
  const uint8_t* s1 = A.tt.blob + A.tt.off1[t];
  const uint8_t* s2 = A.tt.blob + A.tt.off2[t];
  const uint8_t* s3 = A.tt.blob + A.tt.off3[t];
  const int L1 = A.tt.len1[t], L2 = A.tt.len2[t], L3 = A.tt.len3[t];
  const int cap = L1 + L2 + L3;
  const long long row_bytes = cube_dm_row_bytes(L3);
  const uint8_t* dm = A.dm + A.tt.dm_off[t];
  uint8_t* r1 = A.out + A.tt.out_off[t];
  uint8_t* r2 = r1 + cap;
  uint8_t* r3 = r2 + cap;
  uint8_t* md = r3 + cap;
  int i = L1 - 1, j = L2 - 1, k = L3 - 1, n = 0, bad = 0;
  while ((i | j | k) != 0) {
    if (n >= cap) { bad = 1; break; }
    const int byte = dm[(long long)(i * L2 + j) * row_bytes + (k / kCubeSeg) * 512 +
                        (k % kCubeSeg) / 2];
    const int d = (k & 1) ? (byte >> 4) : (byte & 15);
    const int di = (d == kD3S2 || d == kD3S3 || d == kD3P3 || d == kD3P2);
    const int dj = (d == kD3S2 || d == kD3S1 || d == kD3P1 || d == kD3P2);
    const int dk = (d == kD3S2 || d == kD3S3 || d == kD3S1 || d == kD3SS);
    if (d < kD3P3 || d > kD3SS || i - di < 0 || j - dj < 0 || k - dk < 0) { bad = 1; break; }
    const int p = cap - 1 - n;
    r1[p] = di ? s1[i] : kGap;
    r2[p] = dj ? s2[j] : kGap;
    r3[p] = dk ? s3[k] : kGap;
    n++;
    i -= di;
    j -= dj;
    k -= dk;
  }
  int* ol = A.outlen + 4 * (long long)t;
  if (bad) {
    ol[0] = ol[1] = ol[2] = ol[3] = -1;
    return;
  }
  // algn_get_median_3d reads the last column on every iteration (its end
  // pointers never move), so every position gets that column's median.
  if (n > 0) {
    const int c = cap - 1;
    const uint8_t m = A.cm3->median[((((int)r1[c] & 31) << 5) + (r2[c] & 31)) * 32 + (r3[c] & 31)];
    for (int u = 0; u < n; u++) md[cap - 1 - u] = m;
  }
  ol[0] = ol[1] = ol[2] = ol[3] = n;
}

// algn_get_median_3d on caller-aligned sequences (the drop-in median_3).
__global__ void k_median_3(const Cm3Dev* cm3, const uint8_t* a, const uint8_t* b,
                           const uint8_t* c, int len, uint8_t* out) {
  if (len <= 0) return;
  const uint8_t m =
      cm3->median[((((int)a[len - 1] & 31) << 5) + (b[len - 1] & 31)) * 32 + (c[len - 1] & 31)];
  for (int u = blockIdx.x * blockDim.x + threadIdx.x; u < len; u += gridDim.x * blockDim.x)
    out[u] = m;
}

}  // namespace poyhip
