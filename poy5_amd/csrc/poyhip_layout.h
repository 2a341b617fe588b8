// Sizes of the per-pair device regions, shared by host planning code and the
// kernels.  La/Lb are sequence lengths including the leading gap; La <= Lb.
#pragma once
#include <stdint.h>

#ifdef __HIPCC__
#define POYHIP_HD __host__ __device__
#else
#define POYHIP_HD
#endif

namespace poyhip {

POYHIP_HD inline int n_stripes(int La) { return (La - 1 + 63) >> 6; }

// ---- banded affine ------------------------------------------------------
// per-wavefront workspace ints: row-0 block-diagonal buffer (Lb, padded to
// 16 bytes) and two boundary rows of 8-int records (V, CB, H, ED, gap
// counts, padding) so a record moves as one 16-byte and one 4-byte access
POYHIP_HD inline int64_t aff_bnd_offset(int Lb) { return ((int64_t)Lb + 3) & ~(int64_t)3; }
POYHIP_HD inline int64_t aff_ws_ints(int Lb) { return aff_bnd_offset(Lb) + (int64_t)16 * Lb + 16; }
// direction bytes: one byte per (step, lane), at most Lb+63 steps a stripe
POYHIP_HD inline int64_t aff_dm_bytes(int La, int Lb) {
  return (int64_t)n_stripes(La) * (((int64_t)(Lb - 1) + 64 + 3) >> 2) * 256;
}
// outputs: resi, resj, median, medianwg, each La+Lb+2 bytes
POYHIP_HD inline int64_t aff_out_bytes(int La, int Lb) { return 4 * (int64_t)(La + Lb + 2); }

// ---- full affine cost (no directions) -----------------------------------
POYHIP_HD inline int64_t affcost_ws_ints(int Lb) { return (int64_t)8 * Lb + 16; }

// ---- non-affine ---------------------------------------------------------
// per-wavefront workspace: two boundary rows of (cost, delete count)
POYHIP_HD inline int64_t lin_ws_ints(int Lb) { return (int64_t)4 * Lb + 16; }
// 2-bit moves: 16 steps per 32-bit word per lane
POYHIP_HD inline int64_t lin_dm_bytes(int La, int Lb) {
  return (int64_t)n_stripes(La) * (((int64_t)Lb + 63 + 15) >> 4) * 256;
}
// outputs: r1, r2 (La+Lb each), ancestor median (La+Lb+1), median with gaps
POYHIP_HD inline int64_t lin_out_cap(int La, int Lb) { return (int64_t)La + Lb + 1; }
POYHIP_HD inline int64_t lin_out_bytes(int La, int Lb) { return 4 * lin_out_cap(La, Lb); }

// ---- diagonal-store aligner (affine newkkonen) -----------------------------
// per-wavefront workspace: two boundary rows of (P, CD, Q, ED) records and
// gap counts, 16-byte aligned
POYHIP_HD inline int64_t nk_ws_ints(int Lb) { return 10 * (((int64_t)Lb + 3) & ~(int64_t)3) + 64; }
// one direction byte per (step, lane), steps padded to 16; a stripe of the
// widest band has at most Lb + 63 steps
POYHIP_HD inline int64_t nk_stripe_bytes(int nsteps) { return (int64_t)((nsteps + 15) & ~15) * 64; }
POYHIP_HD inline int64_t nk_dm_bytes(int La, int Lb) {
  return (int64_t)n_stripes(La) * nk_stripe_bytes(Lb + 63);
}
// direction bytes of the round with band half-width newk
POYHIP_HD inline int64_t nk_dm_exact(int La, int Lb, int newk) {
  const int D = Lb - La;
  int64_t tot = 0;
  for (int i0 = 1; i0 <= La - 1; i0 += 64) {
    const int ilast = i0 + 63 < La - 1 ? i0 + 63 : La - 1;
    const int jlo = i0 - newk > 0 ? i0 - newk : 0;
    const int jhi = ilast + D + newk < Lb - 1 ? ilast + D + newk : Lb - 1;
    tot += nk_stripe_bytes((jhi - jlo + 1) + (ilast - i0));
  }
  return tot;
}
// outputs: s1', s2' (caller order), ancestor, median with gaps
POYHIP_HD inline int64_t nk_out_cap(int La, int Lb) { return (int64_t)La + Lb + 1; }
POYHIP_HD inline int64_t nk_out_bytes(int La, int Lb) { return 4 * nk_out_cap(La, Lb); }

// ---- 3D cube --------------------------------------------------------------
// rows are cut into segments of 1024 cells (16 per lane)
POYHIP_HD inline int cube_nseg(int L3) { return (L3 + 1023) / 1024; }
POYHIP_HD inline int64_t cube_row_ints(int L3) { return (int64_t)cube_nseg(L3) * 1024; }
// ring of the last L1+L2 rows, per resident wavefront
POYHIP_HD inline int64_t cube_ring_ints(int L1, int L2, int L3) {
  return (int64_t)(L1 + L2) * cube_row_ints(L3);
}
// one nibble per cell: a one-segment row keeps the 8-byte lane blocks that
// carry cells below L3, a longer row 512 bytes per segment
POYHIP_HD inline int64_t cube_dm_row_bytes(int L3) {
  return cube_nseg(L3) == 1 ? 8 * (int64_t)((L3 + 15) / 16) : (int64_t)cube_nseg(L3) * 512;
}
POYHIP_HD inline int64_t cube_dm_bytes(int L1, int L2, int L3) {
  return (int64_t)L1 * L2 * cube_dm_row_bytes(L3);
}
// outputs: r1, r2, r3, median_3, each L1+L2+L3 bytes
POYHIP_HD inline int64_t cube_out_cap(int L1, int L2, int L3) { return (int64_t)L1 + L2 + L3; }
POYHIP_HD inline int64_t cube_out_bytes(int L1, int L2, int L3) {
  return 4 * cube_out_cap(L1, L2, L3);
}

}  // namespace poyhip
