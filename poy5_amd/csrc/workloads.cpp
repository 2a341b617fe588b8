// Synthetic pair workloads of the benchmark configurations (host code).
//
// Each pair is two independent mutations of a random root sequence.  The
// mutation walks the root left to right; at every root position an indel
// event happens with probability `indel` (half deletions that skip a run of
// root symbols, half insertions of a run of random bases before the
// position, run lengths geometric with mean 3), then the surviving symbol is
// substituted with probability `sub` (uniformly among the three other bases)
// and replaced by one of the 11 multi-base IUPAC codes with probability
// `iupac`.  Every output sequence starts with POY's leading gap (16).
// Randomness: SplitMix64, one independent stream per pair derived from the
// seed, so any pair can be regenerated alone.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "poyhip_tools.h"

namespace {

struct SplitMix64 {
  uint64_t x;
  explicit SplitMix64(uint64_t s) : x(s) {}
  uint64_t next() {
    uint64_t z = (x += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
  int below(int n) { return (int)(uniform() * n); }
  // geometric on {1, 2, ...} with success probability p
  int geom(double p) {
    double u = uniform();
    if (u <= 0.0) u = 1e-300;
    return 1 + (int)std::floor(std::log(u) / std::log(1.0 - p));
  }
};

const uint8_t kBases[4] = {1, 2, 4, 8};
const uint8_t kIupac[11] = {3, 5, 9, 6, 10, 12, 7, 11, 13, 14, 15};

struct Params {
  double sub, indel, iupac;
};

void mutate(SplitMix64& r, const std::vector<uint8_t>& root, const Params& pr,
            std::vector<uint8_t>& out) {
  out.clear();
  out.push_back(16);
  size_t k = 0;
  while (k < root.size()) {
    if (r.uniform() < pr.indel) {
      const int len = r.geom(1.0 / 3.0);
      if (r.uniform() < 0.5) {
        k += (size_t)len;
        continue;
      }
      for (int t = 0; t < len; t++) out.push_back(kBases[r.below(4)]);
    }
    uint8_t ch = root[k++];
    if (r.uniform() < pr.sub) {
      int b = 0;
      while (kBases[b] != ch) b++;
      b = (b + 1 + r.below(3)) & 3;
      ch = kBases[b];
    }
    if (r.uniform() < pr.iupac) ch = kIupac[r.below(11)];
    out.push_back(ch);
  }
}

// Generates pair k of configuration `config`.
void gen_pair(int config, uint64_t seed, int k, std::vector<uint8_t>& a, std::vector<uint8_t>& b) {
  SplitMix64 r(seed ^ (0xD1B54A32D192ED03ULL * (uint64_t)(k + 1)));
  r.next();
  int L = 0;
  Params pr{0.10, 0.03, 0.05};
  switch (config) {
    case 1:  // root U[300,700], 10% substitutions, 3% indels, 5% IUPAC
      L = 300 + r.below(401);
      break;
    case 2: {  // root log-uniform [50,5000], divergence U[2%,25%] at 10:3
      const double lo = std::log(50.0), hi = std::log(5000.0);
      L = (int)std::floor(std::exp(lo + (hi - lo) * r.uniform()));
      const double d = 0.02 + 0.23 * r.uniform();
      pr = {d * 10.0 / 13.0, d * 3.0 / 13.0, 0.05};
      break;
    }
    case 4: {  // root U[16000,100000], divergence U[1%,8%], 10% IUPAC
      L = 16000 + r.below(84001);
      const double d = 0.01 + 0.07 * r.uniform();
      pr = {d * 10.0 / 13.0, d * 3.0 / 13.0, 0.10};
      break;
    }
    default:
      L = 0;
  }
  std::vector<uint8_t> root(L);
  for (int t = 0; t < L; t++) root[t] = kBases[r.below(4)];
  mutate(r, root, pr, a);
  mutate(r, root, pr, b);
}

// Generates triple k of configuration 5: root U[200,1000], three branches
// each with its own divergence U[3%,15%] at the 10:3 substitution:indel
// ratio, no IUPAC codes.
void gen_triple(uint64_t seed, int k, std::vector<uint8_t> out[3]) {
  SplitMix64 r(seed ^ (0xA24BAED4963EE407ULL * (uint64_t)(k + 1)));
  r.next();
  const int L = 200 + r.below(801);
  std::vector<uint8_t> root(L);
  for (int t = 0; t < L; t++) root[t] = kBases[r.below(4)];
  for (int u = 0; u < 3; u++) {
    const double d = 0.03 + 0.12 * r.uniform();
    const Params pr{d * 10.0 / 13.0, d * 3.0 / 13.0, 0.0};
    mutate(r, root, pr, out[u]);
  }
}

}  // namespace

extern "C" int poyhip_gen_triples(int config, int n, uint64_t seed, uint8_t* blob, int64_t* off,
                                  int* len, int64_t* total_bytes) {
  if (n <= 0 || !total_bytes || config != 5) return POYHIP_EINVAL;
  const int nt = std::max(1, std::min(16, (int)std::thread::hardware_concurrency()));
  std::vector<int> ln(3 * (size_t)n);
  auto lengths = [&](int t) {
    std::vector<uint8_t> s[3];
    for (int k = t; k < n; k += nt) {
      gen_triple(seed, k, s);
      for (int u = 0; u < 3; u++) ln[3 * (size_t)k + u] = (int)s[u].size();
    }
  };
  {
    std::vector<std::thread> th;
    for (int t = 0; t < nt; t++) th.emplace_back(lengths, t);
    for (auto& x : th) x.join();
  }
  int64_t tot = 0;
  std::vector<int64_t> o(3 * (size_t)n);
  for (size_t q = 0; q < 3 * (size_t)n; q++) {
    o[q] = tot;
    tot += ln[q];
  }
  *total_bytes = tot;
  if (off) memcpy(off, o.data(), sizeof(int64_t) * 3 * (size_t)n);
  if (len) memcpy(len, ln.data(), sizeof(int) * 3 * (size_t)n);
  if (!blob) return POYHIP_OK;
  auto fill = [&](int t) {
    std::vector<uint8_t> s[3];
    for (int k = t; k < n; k += nt) {
      gen_triple(seed, k, s);
      for (int u = 0; u < 3; u++) memcpy(blob + o[3 * (size_t)k + u], s[u].data(), s[u].size());
    }
  };
  std::vector<std::thread> th;
  for (int t = 0; t < nt; t++) th.emplace_back(fill, t);
  for (auto& x : th) x.join();
  return POYHIP_OK;
}

extern "C" int poyhip_gen_pairs(int config, int n, uint64_t seed, uint8_t* blob, int64_t* off,
                                int* len, int64_t* total_bytes) {
  if (n <= 0 || !total_bytes || !(config == 1 || config == 2 || config == 4)) return POYHIP_EINVAL;
  const int nt = std::max(1, std::min(16, (int)std::thread::hardware_concurrency()));
  std::vector<int> la(n), lb(n);
  auto lengths = [&](int t) {
    std::vector<uint8_t> a, b;
    for (int k = t; k < n; k += nt) {
      gen_pair(config, seed, k, a, b);
      la[k] = (int)a.size();
      lb[k] = (int)b.size();
    }
  };
  {
    std::vector<std::thread> th;
    for (int t = 0; t < nt; t++) th.emplace_back(lengths, t);
    for (auto& x : th) x.join();
  }
  int64_t tot = 0;
  std::vector<int64_t> o(2 * (size_t)n);
  for (int k = 0; k < n; k++) {
    o[2 * k] = tot;
    tot += la[k];
    o[2 * k + 1] = tot;
    tot += lb[k];
  }
  *total_bytes = tot;
  if (off) memcpy(off, o.data(), sizeof(int64_t) * 2 * (size_t)n);
  if (len)
    for (int k = 0; k < n; k++) {
      len[2 * k] = la[k];
      len[2 * k + 1] = lb[k];
    }
  if (!blob) return POYHIP_OK;
  auto fill = [&](int t) {
    std::vector<uint8_t> a, b;
    for (int k = t; k < n; k += nt) {
      gen_pair(config, seed, k, a, b);
      memcpy(blob + o[2 * k], a.data(), a.size());
      memcpy(blob + o[2 * k + 1], b.data(), b.size());
    }
  };
  std::vector<std::thread> th;
  for (int t = 0; t < nt; t++) th.emplace_back(fill, t);
  for (auto& x : th) x.join();
  return POYHIP_OK;
}
