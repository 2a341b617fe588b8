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
// Directions use the one-wavefront layout (one nibble per cell, the lane
// blocks below L3 of each row) so the backtrack is shared.
#pragma once
#include "cube.hip"

namespace poyhip {

constexpr int kCubeMwWaves = 8;
// longest rows the workgroup fill takes: their codes (one byte per cell)
// fit the staging buffers, 2 * (2*8+1) * 2 KB
constexpr int kCubeMwMaxSeg = 68;
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
                                              const char* s_pl, const char* s_pair, bool seg0,
                                              uint32_t cU, uint32_t cP, uint32_t cD, uint32_t cY,
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
  // k-1 neighbours of the lane's first cell: lane 0 takes the carries from
  // the previous segment (kMwInf in the first, whose k = 0 has no such move)
  uint32_t nU = kMwInf, nP = kMwInf, nD = kMwInf;
  if (kKind == kRowInner) nU = ushr1(U[kCubeCpl - 1], cU);
  if (kKind == kRowInner || kKind == kRowPlane0) nP = ushr1(P[kCubeCpl - 1], cP);
  if (kKind == kRowInner || kKind == kRowPlaneStart) nD = ushr1(D[kCubeCpl - 1], cD);
  uint32_t GG[kCubeCpl];
  // pass 1: the six plane and line moves, then the lane's map
  // y -> min(fa, y + fb) of the SS run
  uint32_t fa = kMwInf, fb = 0;
#pragma unroll
  for (int m = 0; m < kCubeCpl; m++) {
    const bool k0cell = seg0 && (m == 0) && (lane == 0);
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
  uint32_t y = lane == 0 ? cY : umin(ia, cY + ib);
  // pass 2: the SS move
  uint32_t lo = 0, hi = 0;
#pragma unroll
  for (int m = 0; m < kCubeCpl; m++) {
    uint32_t key = X[m];
    if (!(seg0 && m == 0 && lane == 0)) key = umin(key, (y | 7u) + GG[m]);
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

// Staging of the rows a group reads, for 16-bit ring rows.  While group s
// is computed, the rows of group s+1 are copied from the ring into LDS with
// global_load_lds (no registers, completion counted like any vector load),
// whenever they all lie before group s: they are then complete since group
// s's barrier.  A buffer has 2*NW+1 slots: slot u <= NW holds diag row b+u
// (wavefront w reads D from slot w and U from slot w+1), slot NW+1+w the
// prev row p+w.  A slot holds a row's 128 chunks of 16 bytes with chunk
// 2L+h (lane L's half h) at position 64h+L, so each of a lane's two
// ds_read_b128 reads 16 consecutive bytes per lane across the wavefront.
// Two buffers alternate: a buffer is refilled one step after it was read,
// and every wavefront reads its rows before it reaches the barrier between.
using lds_ptr_t = __attribute__((address_space(3))) void*;

// s_waitcnt immediates (vmcnt in bits 3:0 and 15:14, expcnt 6:4 and
// lgkmcnt 11:8 left at their maxima): wait until at most n vector memory
// operations are outstanding.
constexpr int mw_vmcnt(int n) { return (n & 15) | (7 << 4) | (15 << 8) | ((n >> 4) << 14); }
template <bool kDm>
constexpr int kMwWaitStores = mw_vmcnt(kDm ? 3 : 2);
constexpr int kMwWaitNone = mw_vmcnt(0);

__device__ __forceinline__ void mw_stage_row(const uint32_t* row, int nchunk, uint4* slot, int lane) {
#pragma unroll
  for (int h = 0; h < 2; h++) {
    const int c = imin(2 * lane + h, nchunk - 1);
    __builtin_amdgcn_global_load_lds(row + 4 * c, (lds_ptr_t)(slot + 64 * h), 16, 0, 0);
  }
}

__device__ __forceinline__ void mw_read_slot(const uint4* slot, int lane, uint32_t v[kCubeCpl]) {
  const uint4 a = slot[lane];
  const uint4 b = slot[64 + lane];
  const uint32_t wd[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
#pragma unroll
  for (int q = 0; q < 8; q++) {
    v[2 * q] = wd[q] & 0xFFFFu;
    v[2 * q + 1] = wd[q] >> 16;
  }
}

template <bool kDm, bool kU16, int NW>
__global__ __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(4))) void k_cube_mw(
    CubeFillArgs A) {
  __shared__ uint2 s_pair[32 * 32];
  __shared__ uint2 s_pl[32];
  __shared__ int s_q;
  constexpr int kSlots = 2 * NW + 1;
  // staging buffers of one-segment triples; a longer triple keeps its
  // pre-scaled codes (8 * s3[k], one byte per cell) here instead
  __shared__ uint4 s_rows[2][kSlots][128];
  static_assert(sizeof(s_rows) >= (size_t)kCubeMwMaxSeg * kCubeSeg, "codes fit the buffers");
  uint8_t* s_codes = reinterpret_cast<uint8_t*>(&s_rows[0][0][0]);
  const int tid = threadIdx.x, lane = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave-uniform
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
    const int nseg = cube_nseg(L3);
    if (nseg < A.seg_lo || nseg > A.seg_hi) continue;
    const uint8_t* s1 = A.tt.blob + A.tt.off1[t];
    const uint8_t* s2 = A.tt.blob + A.tt.off2[t];
    const uint8_t* s3 = A.tt.blob + A.tt.off3[t];
    const int RS = cube_mw_row_words(L3, kU16);
    const int NR = L1 + L2 + NW;
    const int bw = kU16 ? kCubeCpl / 2 : kCubeCpl;  // words per lane block
    uint32_t* ring = reinterpret_cast<uint32_t*>(A.ring) + (long long)blockIdx.x * A.ring_stride;
    uint8_t* dm = kDm ? A.dm + A.tt.dm_off[t] : nullptr;
    const long long dmrow = cube_dm_row_bytes(L3);
    const int g3 = s3[0] & 31;  // s3[0] prices the k = 0 entries (the leading gap)
    // codes of a one-segment row stay in registers
    uint32_t c8[kCubeCpl / 4] = {0, 0, 0, 0};
    if (nseg == 1) {
#pragma unroll
      for (int m = 0; m < kCubeCpl; m++) {
        const int k = lane * kCubeCpl + m;
        c8[m >> 2] |= (uint32_t)(8 * (k < L3 ? (s3[k] & 31) : kGap)) << (8 * (m & 3));
      }
    } else {
      for (int k = tid; k < nseg * kCubeSeg; k += 64 * NW)
        s_codes[k] = (uint8_t)(8 * (k < L3 ? (s3[k] & 31) : kGap));
    }
    // lane block of segment sg (lanes past L3 read the last real block and
    // store nothing)
    auto blk = [&](int sg) { return imin((sg * 64 + lane) * bw, RS - bw); };
    auto rowp = [&](int r, int sg) { return ring + (long long)(r % NR) * RS + blk(sg); };
    auto row0 = [&](int r) { return ring + (long long)(r % NR) * RS; };
    int i = 0, j0 = 0, g = 1, cur = 0;
    bool have = false;  // this group's inner rows are staged in s_rows[cur]
    // s2 code of the wavefront's row, loaded during the step before (the
    // first group of a triple has only row j = 0, which needs none)
    int s2c = 0;
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
      // the next group; stage its rows when its newest input, the prev row
      // of its last inner row, lies before this group
      int i2 = i, j02 = j0 + g, g2 = g;
      if (j02 >= L2) {
        i2 = i + 1;
        j02 = 0;
        g2 = imin(NW, i2 + 1);
      }
      bool stage = false;
      if (kU16 && nseg == 1 && i2 >= 2 && i2 < L1 && L2 >= 2) {
        const int jl = imin(j02 + g2, L2) - 1;
        stage = jl >= 1 && i2 * L2 + jl - i2 - 1 < G0;
      }
      const int R = G0 + w;
      const int kind =
          i == 0 ? (j == 0 ? kRowFirst : kRowPlane0) : (j == 0 ? kRowPlaneStart : kRowInner);
      int b256 = 0;
      uint32_t kP1 = 0, kP2 = 0;
      if (busy) {
        b256 = j >= 1 ? 256 * (s2c & 31) : 0;
        kP1 = j >= 1 ? s_pair[(b256 >> 3) + g3].x - kD3S1 + kD3P1 : 0;
        kP2 = (i >= 1 && j >= 1) ? s_pair[(b256 >> 3) + g3].y - kD3S2 + kD3P2 : 0;
      }
      // the next group's s2 code, consumed a step later
      {
        const int jn = j02 + w;
        s2c = (i2 < L1 && w < g2 && jn >= 1 && jn < L2) ? (int)s2[jn] : 0;
      }
      // rows read: upper, prev, diag
      const int rU = kind == kRowInner ? (i - 1) * (L2 - 1) + j : 0;
      const int rP = kind == kRowInner ? i * (L2 - 1) + j - 1 : R - 1;
      const int rD = kind == kRowInner ? rU - 1 : (i - 1) * (L2 - 1);
      // carries from the previous segment
      uint32_t cU = kMwInf, cP = kMwInf, cD = kMwInf, cY = kMwInf;
      for (int sg = 0; sg < nseg; sg++) {
        uint32_t U[kCubeCpl], P[kCubeCpl], D[kCubeCpl];
        if (busy) {
          if (kind == kRowPlane0) {
            cube_mw_load<kU16>(rowp(rP, sg), P);
          } else if (kind == kRowPlaneStart) {
            cube_mw_load<kU16>(rowp(rD, sg), D);
          } else if (kind == kRowInner && kU16 && have) {
            mw_read_slot(&s_rows[cur][w][0], lane, D);
            mw_read_slot(&s_rows[cur][w + 1][0], lane, U);
            mw_read_slot(&s_rows[cur][NW + 1 + w][0], lane, P);
          } else if (kind == kRowInner) {
            cube_mw_load<kU16>(rowp(rU, sg), U);
            cube_mw_load<kU16>(rowp(rP, sg), P);
            cube_mw_load<kU16>(rowp(rD, sg), D);
          }
        }
        // staging is issued after this step's own rows are read (so the
        // compiler's wait before those LDS reads does not cover it) and
        // before the stores, so that waiting for it leaves the stores in
        // flight
        if (stage) {
          const int bn = (i2 - 1) * (L2 - 1) + j02 - 1;
          const int pn = i2 * (L2 - 1) + j02 - 1;
          const int nu = imin(g2, L2 - j02);  // diag slots 0..nu
          for (int sl = w; sl < kSlots; sl += NW) {
            if (sl <= NW) {
              if (sl <= nu) mw_stage_row(row0(bn + sl), RS / 4, &s_rows[cur ^ 1][sl][0], lane);
            } else if (sl - NW - 1 < nu) {
              mw_stage_row(row0(pn + sl - NW - 1), RS / 4, &s_rows[cur ^ 1][sl][0], lane);
            }
          }
        }
        if (busy) {
          uint32_t cc[kCubeCpl / 4];
          if (nseg == 1) {
#pragma unroll
            for (int u = 0; u < kCubeCpl / 4; u++) cc[u] = c8[u];
          } else {
            const uint4 v = reinterpret_cast<const uint4*>(s_codes + sg * kCubeSeg)[lane];
            cc[0] = v.x;
            cc[1] = v.y;
            cc[2] = v.z;
            cc[3] = v.w;
          }
          const bool seg0 = sg == 0;
          uint32_t X[kCubeCpl];
          uint32_t lo, hi;
          if (kind == kRowInner)
            cube_mw_cells<kRowInner, kDm>(U, P, D, cc, b256, kP3, kP1, kP2, pl_base, pair_base,
                                          seg0, cU, cP, cD, cY, X, lo, hi);
          else if (kind == kRowPlaneStart)
            cube_mw_cells<kRowPlaneStart, kDm>(U, P, D, cc, b256, kP3, kP1, kP2, pl_base,
                                               pair_base, seg0, cU, cP, cD, cY, X, lo, hi);
          else if (kind == kRowPlane0)
            cube_mw_cells<kRowPlane0, kDm>(U, P, D, cc, b256, kP3, kP1, kP2, pl_base, pair_base,
                                           seg0, cU, cP, cD, cY, X, lo, hi);
          else
            cube_mw_cells<kRowFirst, kDm>(U, P, D, cc, b256, kP3, kP1, kP2, pl_base, pair_base,
                                          seg0, cU, cP, cD, cY, X, lo, hi);
          const int k0 = sg * kCubeSeg + lane * kCubeCpl;
          uint2* pDm = kDm ? reinterpret_cast<uint2*>(dm + R * dmrow + sg * 512) + lane : nullptr;
          cube_mw_store<kDm, kU16>(rowp(R, sg), pDm, k0 < L3, X, lo, hi);
          if (R == L1 * L2 - 1 && (L3 - 1) / kCubeSeg == sg) {
            // the last cell, k = L3 - 1, belongs to lane ((L3 - 1) % 1024) / 16
            const int mm = (L3 - 1) % kCubeCpl;
            uint32_t sel = 0;
#pragma unroll
            for (int m = 0; m < kCubeCpl; m++) sel = (m == mm) ? X[m] : sel;
            const int lv = __shfl((int)(sel >> 3), ((L3 - 1) % kCubeSeg) / kCubeCpl);
            if (lane == 0) A.cost[t] = lv;
          }
          if (sg + 1 < nseg) {
            cU = (uint32_t)__builtin_amdgcn_readlane((int)U[kCubeCpl - 1], 63);
            cP = (uint32_t)__builtin_amdgcn_readlane((int)P[kCubeCpl - 1], 63);
            cD = (uint32_t)__builtin_amdgcn_readlane((int)D[kCubeCpl - 1], 63);
            cY = (uint32_t)__builtin_amdgcn_readlane((int)X[kCubeCpl - 1], 63);
          }
        }
      }
      // The staged rows land before the barrier.  Vector memory operations
      // complete in issue order for the counter, so waiting until only this
      // step's stores are outstanding (two ring stores and the direction
      // store of a busy wavefront) covers the staging loads.
      if (stage) {
        if (busy)
          __builtin_amdgcn_s_waitcnt(kMwWaitStores<kDm>);
        else
          __builtin_amdgcn_s_waitcnt(kMwWaitNone);
      }
      have = stage;
      cur ^= 1;
      i = i2;
      j0 = j02;
      g = g2;
    }
  }
}

template __global__ void k_cube_mw<true, false, kCubeMwWaves>(CubeFillArgs);
template __global__ void k_cube_mw<false, false, kCubeMwWaves>(CubeFillArgs);
template __global__ void k_cube_mw<true, true, kCubeMwWaves>(CubeFillArgs);
template __global__ void k_cube_mw<false, true, kCubeMwWaves>(CubeFillArgs);

}  // namespace poyhip
