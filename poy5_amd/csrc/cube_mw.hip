// Multi-wavefront fill of algn_fill_cube (src/algn.c:2593-2763) for triples
// whose rows fit one segment (L3 <= 1024).
//
// The NW wavefronts of a workgroup share one triple.  With R = i*L2 + j the
// global row number, a row of plane i >= 1 reads
//   prev = R - i - 1, upper = R - L2 - i + 1, diag = upper - 1
// (the plane-start row j = 0 reads diag = R - L2 - i + 1 only), and a row of
// plane 0 reads prev = R - 1.  Every row of plane i >= 1 therefore reads rows
// at least i+1 behind it, so any min(NW, i+1) consecutive rows of the plane
// are independent of one another.  The workgroup walks plane i in groups of
// that many rows, wavefront w taking row j0 + w of the group, with one
// workgroup barrier between groups; plane 0 runs one row per group.  Cells
// are computed exactly as in the one-wavefront fill: the same candidate
// order (P3, P1, P2, S3, S1, S2, then SS along the row) and the same strict
// comparisons, so values and direction nibbles are identical.
//
// Ring: the deepest read lags L1+L2-1 rows and a group writes up to NW rows
// while the others of the group are read, so the workgroup keeps the last
// L1+L2+NW rows.  A ring row holds the lane blocks that carry cells below
// L3: 16 * ceil(L3/16) ints.  Lanes past L3 load a valid block (the last
// real one) and store nothing; their cells never feed a cell below L3
// (values move to larger k only).
//
// Keys.  A cell is carried as key = 8*value + code, code the direction
// nibble of its move (cube.hip numbers the moves in the order they are
// tried).  The minimum key holds the minimum value and, among equal values,
// the move tried first: exactly the reference's "replace only when strictly
// smaller" chain.  Ring rows hold 8*value, and the tables hold 8*cost + the
// code of the move they price, so a candidate key is one add and the choice
// is a min3 tree.  The SS move (code 7) of pass 2 becomes
// min(key, (prev | 7) + 8*gg): a tie keeps the earlier move.  kMwInf = 2^30
// stands for a missing neighbour; values stay below 2^27, so no sum of two
// keys, or of kMwInf and a key, wraps.
//
// Tables (cm_precalc_4algn_3d, src/cm.c:1390-1412) live in LDS as pairs so
// that one 8-byte read serves two lookups:
//   pair[b][c] = {8*(gap, b, c) + S1, 8*(s1[i], b, c) + S2}   per plane
//   pl[c]      = {8*(s1[i], gap, c) + S3, 8*(gap, gap, c)}
// Directions use the one-wavefront layout (one nibble per cell, 512 bytes
// per row) so the backtrack is shared.
#pragma once
#include "cube.hip"

namespace poyhip {

constexpr int kCubeMwWaves = 8;
constexpr uint32_t kMwInf = 1u << 30;

// 32-bit words of a ring row: the lane blocks that carry cells below L3
POYHIP_HD inline int cube_mw_row_words(int L3, bool u16) {
  return (u16 ? 8 : 16) * ((L3 + 15) / 16);
}
POYHIP_HD inline int64_t cube_mw_ring_ints(int L1, int L2, int L3, int NW) {
  return (int64_t)(L1 + L2 + NW) * cube_mw_row_words(L3, false);
}
// 16-bit ring rows need 8*value < 2^16 for every cell: a value is at most
// (L1 + L2 + L3 + 15) times the largest table entry (the cells padding the
// last lane block included)
POYHIP_HD inline bool cube_mw_u16(int64_t len_sum, int cm3_max) {
  return cm3_max >= 0 && 8 * (len_sum + 16) * (int64_t)cm3_max < 65536;
}

__device__ __forceinline__ uint32_t umin(uint32_t a, uint32_t b) { return a < b ? a : b; }
__device__ __forceinline__ uint32_t umin3(uint32_t a, uint32_t b, uint32_t c) {
  return umin(umin(a, b), c);
}
__device__ __forceinline__ uint32_t ushr1(uint32_t v, uint32_t lane0) {
  return (uint32_t)wave_shr1((int)v, (int)lane0);
}

template <int kCtrl, int kRowMask>
__device__ __forceinline__ void mw_scan_step(uint32_t& fa, uint32_t& fb) {
  const uint32_t pa = (uint32_t)__builtin_amdgcn_update_dpp((int)kMwInf, (int)fa, kCtrl, kRowMask, 0xF, false);
  const uint32_t pb = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)fb, kCtrl, kRowMask, 0xF, false);
  fa = umin(fa, pa + fb);
  fb = umin(pb + fb, kMwInf);
}

// Inclusive scan over the 64 lanes of the maps y -> min(fa, y + fb) in the
// key domain (identity (kMwInf, 0)); the same DPP steps as cube_scan_minplus.
__device__ __forceinline__ void mw_scan_minplus(uint32_t& fa, uint32_t& fb) {
  mw_scan_step<0x111, 0xF>(fa, fb);
  mw_scan_step<0x112, 0xF>(fa, fb);
  mw_scan_step<0x114, 0xF>(fa, fb);
  mw_scan_step<0x118, 0xF>(fa, fb);
  mw_scan_step<0x142, 0xA>(fa, fb);
  mw_scan_step<0x143, 0xC>(fa, fb);
}

__device__ __forceinline__ uint2 lds_pair(const char* base, int byte_off) {
  return *reinterpret_cast<const uint2*>(base + byte_off);
}

__device__ __forceinline__ void load16u(const uint32_t* p, uint32_t v[kCubeCpl]) {
  load16(reinterpret_cast<const int*>(p), reinterpret_cast<int*>(v));
}

// One row of one wavefront from its input rows U (upper), P (prev) and
// D (diag), already in registers (the kinds read: first none, plane 0 P,
// plane start D, inner all three).  Byte m&3 of c8in[m>>2] is 8 * code of
// cell k0+m; b256 is 256 * s2[j] (the row of the pair table); kP3/kP1/kP2
// are the k = 0 table entries as keys of the plane moves.  X returns the
// keys, lo/hi the direction nibbles of cells 0-7 and 8-15.
template <int kKind, bool kDm>
__device__ __forceinline__ void cube_mw_cells(const uint32_t U[kCubeCpl], const uint32_t P[kCubeCpl],
                                              const uint32_t D[kCubeCpl],
                                              const uint32_t c8in[kCubeCpl / 4], int b256,
                                              uint32_t kP3, uint32_t kP1, uint32_t kP2,
                                              const char* s_pl, const char* s_pair,
                                              uint32_t X[kCubeCpl], uint32_t& lo_out,
                                              uint32_t& hi_out) {
  const int lane = threadIdx.x & 63;
  // Register pressure: the empty asm statements make values opaque to the
  // optimiser.  Without them the per-cell table offsets are hoisted out of
  // the row loop (16 live registers) and the nibble ORs are reassociated
  // into a tree that keeps every cell's code live.
  uint32_t c8p[kCubeCpl / 4];
#pragma unroll
  for (int u = 0; u < kCubeCpl / 4; u++) {
    c8p[u] = c8in[u];
    asm volatile("" : "+v"(c8p[u]));
  }
  // k-1 neighbours of the lane's first cell; lane 0's first cell is k = 0,
  // which has no such moves
  uint32_t nU = kMwInf, nP = kMwInf, nD = kMwInf;
  if (kKind == kRowInner) nU = ushr1(U[kCubeCpl - 1], kMwInf);
  if (kKind == kRowInner || kKind == kRowPlane0) nP = ushr1(P[kCubeCpl - 1], kMwInf);
  if (kKind == kRowInner || kKind == kRowPlaneStart) nD = ushr1(D[kCubeCpl - 1], kMwInf);
  uint32_t GG[kCubeCpl];
  // pass 1: the six plane and line moves, then the lane's map
  // y -> min(fa, y + fb) of the SS run
  uint32_t fa = kMwInf, fb = 0;
#pragma unroll
  for (int m = 0; m < kCubeCpl; m++) {
    const bool k0cell = (m == 0) && (lane == 0);
    const int c8 = (c8p[m >> 2] >> (8 * (m & 3))) & 255;
    const uint2 pl = lds_pair(s_pl, c8);
    GG[m] = pl.y;
    uint32_t key;
    if (kKind == kRowFirst) {
      key = k0cell ? (uint32_t)kD3S2 : kMwInf + kD3SS;
    } else if (kKind == kRowPlane0) {
      const uint2 pr = lds_pair(s_pair, b256 + c8);
      key = umin(P[m] + kP1, (m == 0 ? nP : P[m - 1]) + pr.x);
    } else if (kKind == kRowPlaneStart) {
      key = umin(D[m] + kP3, (m == 0 ? nD : D[m - 1]) + pl.x);
    } else {
      const uint2 pr = lds_pair(s_pair, b256 + c8);
      const uint32_t a = umin3(U[m] + kP3, P[m] + kP1, D[m] + kP2);
      const uint32_t b = umin3((m == 0 ? nU : U[m - 1]) + pl.x, (m == 0 ? nP : P[m - 1]) + pr.x,
                               (m == 0 ? nD : D[m - 1]) + pr.y);
      key = umin(a, b);
    }
    X[m] = key;
    if (k0cell) {
      fa = key;
      fb = kMwInf;  // the chain starts at k = 0: y is ignored
    } else {
      fa = umin(key, fa + pl.y);
      fb = umin(fb + pl.y, kMwInf);
    }
  }
  mw_scan_minplus(fa, fb);
  const uint32_t ia = ushr1(fa, kMwInf);
  const uint32_t ib = ushr1(fb, 0);
  uint32_t y = lane == 0 ? kMwInf : umin(ia, kMwInf + ib);
  // pass 2: the SS move
  uint32_t lo = 0, hi = 0;
#pragma unroll
  for (int m = 0; m < kCubeCpl; m++) {
    uint32_t key = X[m];
    if (!(m == 0 && lane == 0)) key = umin(key, (y | 7u) + GG[m]);
    X[m] = key;
    y = key;
    if (kDm) {
      if (m < 8) lo |= (key & 7u) << (4 * m);
      else hi |= (key & 7u) << (4 * (m - 8));
      asm volatile("" : "+v"(lo), "+v"(hi));
    }
  }
  lo_out = lo;
  hi_out = hi;
}

// Ring rows hold 8*value per cell, as 32-bit words or, when every value of
// the cube stays below 2^13 (kU16), as 16-bit halves: a lane's 16 cells are
// 16 words or 8 words (cell 2q in the low half of word q).
template <bool kU16>
__device__ __forceinline__ void cube_mw_load(const uint32_t* p, uint32_t v[kCubeCpl]) {
  if (kU16) {
    const uint4 a = reinterpret_cast<const uint4*>(p)[0];
    const uint4 b = reinterpret_cast<const uint4*>(p)[1];
    const uint32_t w[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
#pragma unroll
    for (int q = 0; q < 8; q++) {
      v[2 * q] = w[q] & 0xFFFFu;
      v[2 * q + 1] = w[q] >> 16;
    }
  } else {
    load16u(p, v);
  }
}

// Stores a computed row: the ring row holds 8*value, the direction row the
// nibbles.  Lanes past L3 store nothing.
template <bool kDm, bool kU16>
__device__ __forceinline__ void cube_mw_store(uint32_t* pX, uint2* pDm, bool act,
                                              const uint32_t X[kCubeCpl], uint32_t lo, uint32_t hi) {
  if (act) {
    if (kU16) {
      uint32_t w[8];
#pragma unroll
      for (int q = 0; q < 8; q++) w[q] = (X[2 * q] & 0xFFF8u) | ((X[2 * q + 1] & 0xFFF8u) << 16);
      reinterpret_cast<uint4*>(pX)[0] = make_uint4(w[0], w[1], w[2], w[3]);
      reinterpret_cast<uint4*>(pX)[1] = make_uint4(w[4], w[5], w[6], w[7]);
    } else {
      uint32_t V[kCubeCpl];
#pragma unroll
      for (int m = 0; m < kCubeCpl; m++) V[m] = X[m] & ~7u;
      store16(reinterpret_cast<int*>(pX), reinterpret_cast<const int*>(V));
    }
    if (kDm) *pDm = make_uint2(lo, hi);
  }
}

template <bool kDm, bool kU16, int NW>
__global__ __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(4))) void k_cube_mw(
    CubeFillArgs A) {
  __shared__ uint2 s_pair[32 * 32];
  __shared__ uint2 s_pl[32];
  __shared__ int s_q;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  for (int e = tid; e < 1024; e += 64 * NW)
    s_pair[e].x = 8u * cube_c3(A.cm3, kGap, e >> 5, e & 31) + kD3S1;
  if (tid < 32) s_pl[tid].y = 8u * cube_c3(A.cm3, kGap, kGap, tid);
  const char* pl_base = reinterpret_cast<const char*>(s_pl);
  const char* pair_base = reinterpret_cast<const char*>(s_pair);

  for (;;) {
    if (tid == 0) s_q = atomicAdd(A.counter, 1);
    __syncthreads();
    const int q = s_q;
    __syncthreads();  // the next grab rewrites s_q
    if (q >= A.tt.n) break;
    const int t = A.tt.order[q];
    const int L1 = A.tt.len1[t], L2 = A.tt.len2[t], L3 = A.tt.len3[t];
    if (cube_nseg(L3) != 1) continue;
    const uint8_t* s1 = A.tt.blob + A.tt.off1[t];
    const uint8_t* s2 = A.tt.blob + A.tt.off2[t];
    const uint8_t* s3 = A.tt.blob + A.tt.off3[t];
    const int RS = cube_mw_row_words(L3, kU16);
    const int NR = L1 + L2 + NW;
    const int k0 = lane * kCubeCpl;
    const bool act = k0 < L3;
    // lanes past L3 read the last real block (in bounds, never stored)
    const int bw = kU16 ? kCubeCpl / 2 : kCubeCpl;  // words per lane block
    const int kr = imin(lane * bw, RS - bw);
    uint32_t* ring = reinterpret_cast<uint32_t*>(A.ring) + (long long)blockIdx.x * A.ring_stride;
    uint8_t* dm = kDm ? A.dm + A.tt.dm_off[t] : nullptr;
    const int g3 = s3[0] & 31;  // s3[0] prices the k = 0 entries (the leading gap)
    uint32_t c8[kCubeCpl / 4] = {0, 0, 0, 0};
#pragma unroll
    for (int m = 0; m < kCubeCpl; m++) {
      const int k = k0 + m;
      c8[m >> 2] |= (uint32_t)(8 * (k < L3 ? (s3[k] & 31) : kGap)) << (8 * (m & 3));
    }
    auto rowp = [&](int r) { return ring + (long long)(r % NR) * RS + kr; };
    int i = 0, j0 = 0, g = 1;
    uint32_t kP3 = 0;
    while (i < L1) {
      __syncthreads();  // the rows of the previous group are in the ring
      if (j0 == 0 && i >= 1) {
        const int a = s1[i] & 31;
        for (int e = tid; e < 1024; e += 64 * NW)
          s_pair[e].y = 8u * cube_c3(A.cm3, a, e >> 5, e & 31) + kD3S2;
        if (tid < 32) s_pl[tid].x = 8u * cube_c3(A.cm3, a, kGap, tid) + kD3S3;
        __syncthreads();
        kP3 = s_pl[g3].x - kD3S3 + kD3P3;
      }
      const int G0 = i * L2 + j0;
      const int j = j0 + w;
      const bool busy = w < g && j < L2;
      uint32_t U[kCubeCpl], P[kCubeCpl], D[kCubeCpl];
      int kind = kRowInner;
      if (busy) {
        const int R = G0 + w;
        kind = i == 0 ? (j == 0 ? kRowFirst : kRowPlane0) : (j == 0 ? kRowPlaneStart : kRowInner);
        if (kind == kRowPlane0) {
          cube_mw_load<kU16>(rowp(R - 1), P);
        } else if (kind == kRowPlaneStart) {
          cube_mw_load<kU16>(rowp((i - 1) * (L2 - 1)), D);
        } else if (kind == kRowInner) {
          const int rD = (i - 1) * (L2 - 1) + j - 1;
          cube_mw_load<kU16>(rowp(rD + 1), U);
          cube_mw_load<kU16>(rowp(i * (L2 - 1) + j - 1), P);
          cube_mw_load<kU16>(rowp(rD), D);
        }
      }
      if (busy) {
        const int R = G0 + w;
        const int b256 = j >= 1 ? 256 * (s2[j] & 31) : 0;
        const uint32_t kP1 = j >= 1 ? s_pair[(b256 >> 3) + g3].x - kD3S1 + kD3P1 : 0;
        const uint32_t kP2 = (i >= 1 && j >= 1) ? s_pair[(b256 >> 3) + g3].y - kD3S2 + kD3P2 : 0;
        uint32_t X[kCubeCpl];
        uint32_t lo, hi;
        if (kind == kRowInner)
          cube_mw_cells<kRowInner, kDm>(U, P, D, c8, b256, kP3, kP1, kP2, pl_base, pair_base, X,
                                        lo, hi);
        else if (kind == kRowPlaneStart)
          cube_mw_cells<kRowPlaneStart, kDm>(U, P, D, c8, b256, kP3, kP1, kP2, pl_base, pair_base,
                                             X, lo, hi);
        else if (kind == kRowPlane0)
          cube_mw_cells<kRowPlane0, kDm>(U, P, D, c8, b256, kP3, kP1, kP2, pl_base, pair_base, X,
                                         lo, hi);
        else
          cube_mw_cells<kRowFirst, kDm>(U, P, D, c8, b256, kP3, kP1, kP2, pl_base, pair_base, X,
                                        lo, hi);
        uint32_t* pX = rowp(R);
        uint2* pDm = kDm ? reinterpret_cast<uint2*>(dm + (long long)R * 512) + lane : nullptr;
        cube_mw_store<kDm, kU16>(pX, pDm, act, X, lo, hi);
        if (R == L1 * L2 - 1) {
          // the last cell, k = L3 - 1, belongs to lane (L3 - 1) / 16
          const int mm = (L3 - 1) % kCubeCpl;
          uint32_t sel = 0;
#pragma unroll
          for (int m = 0; m < kCubeCpl; m++) sel = (m == mm) ? X[m] : sel;
          const int lv = __shfl((int)(sel >> 3), (L3 - 1) / kCubeCpl);
          if (lane == 0) A.cost[t] = lv;
        }
      }
      j0 += g;
      if (j0 >= L2) {
        i++;
        j0 = 0;
        g = imin(NW, i + 1);
      }
    }
  }
}

template __global__ void k_cube_mw<true, false, kCubeMwWaves>(CubeFillArgs);
template __global__ void k_cube_mw<false, false, kCubeMwWaves>(CubeFillArgs);
template __global__ void k_cube_mw<true, true, kCubeMwWaves>(CubeFillArgs);
template __global__ void k_cube_mw<false, true, kCubeMwWaves>(CubeFillArgs);

}  // namespace poyhip
