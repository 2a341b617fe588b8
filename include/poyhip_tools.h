/*
 * poyhip_tools.h -- benchmark and test hooks of libpoyhip.
 *
 * None of these replaces a POY stub and none belongs to the drop-in
 * interface of poyhip.h: they generate the synthetic benchmark inputs and
 * force failure paths from the host in tests.  Same conventions as
 * poyhip.h (return codes, host pointers).
 */
#ifndef POYHIP_TOOLS_H
#define POYHIP_TOOLS_H

#include "poyhip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- workloads (synthetic inputs of the benchmark configurations) ------- */
/* SplitMix64-seeded pair generators; see the package documentation for the
 * mutation process.  Two-pass: call with blob == NULL to get the total
 * bytes, then again to fill.  Sequences are written back to back; off/len
 * arrays have 2n entries (first, second sequence of each pair). */
int poyhip_gen_pairs(int config, int n, uint64_t seed, uint8_t *blob, int64_t *off,
                     int *len, int64_t *total_bytes);
/* Triples of configuration 5 (three mutations of one root); off/len arrays
 * have 3n entries. */
int poyhip_gen_triples(int config, int n, uint64_t seed, uint8_t *blob, int64_t *off,
                       int *len, int64_t *total_bytes);

/* ---- failure injection --------------------------------------------------- */
/* The server's next admission fails as a device launch error would: its
 * requests end with POYHIP_EDEVICE, the medians they were to keep are
 * marked as never coming, and the server refuses later submissions.  No
 * kernel is launched for the failed admission. */
int poyhip_server_fail_next_launch(poyhip_server *s);

#ifdef __cplusplus
}
#endif
#endif /* POYHIP_TOOLS_H */
