// Host side of libpoyhip: the C ABI declared in include/poyhip.h.
//
// Single translation unit: the kernel sources are included here so every
// launch sees its kernel's definition.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <climits>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "poyhip_tools.h"
#include "poyhip_device.h"
#include "poyhip_layout.h"

#include "affine.hip"
#include "cube.hip"
#include "cube_mw.hip"
#include "linear.hip"
#include "linear_mw.hip"
#include "medians.hip"
#include "newkk.hip"
#include "unions.hip"

using namespace poyhip;

namespace {

__global__ void k_zero_ints(int* p, int n) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) p[k] = 0;
}

__global__ void k_extract_cost(const PairResult* res, int* cost, int n) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) cost[k] = res[k].cost;
}

struct DevBuf {
  void* p = nullptr;
  size_t cap = 0;
};

// device blocks that destroyed batches of a context leave for its next
// batches.  Batches of one context run in order on its stream and their
// destroy waits for that stream, so a block handed on is idle.  A block
// serves a request of at least half its size; kSpareBytes bound what is kept
struct SpareBlocks {
  static constexpr size_t kSpareBytes = size_t(4) << 30;
  static constexpr size_t kMaxBlocks = 64;
  std::mutex m;
  std::vector<std::pair<void*, size_t>> idle, live;
  size_t idle_bytes = 0;
};

}  // namespace

struct poyhip_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  char err[512] = {0};
  bool has_cm = false;
  poyhip_cm cm;
  CmDev* d_cm = nullptr;
  int num_cus = 0;
  // scratch of the per-pair drop-ins (grow-only, like Matrix.default)
  DevBuf stage, dm, ws, stab, outblk;
  std::vector<uint8_t> h_stage;
  std::vector<uint8_t> h_out;
  // state of the last poyhip_cost_2, read by poyhip_backtrace_2d
  bool last2d = false;
  int last_deltawh = 0;
  std::vector<uint8_t> last_s1, last_s2;
  // three-sequence model and the state of the last poyhip_cost_3
  Cm3Dev* d_cm3 = nullptr;
  bool has_cm3 = false;
  bool last3d = false;
  std::vector<uint8_t> last3[3];
  // the linear fill may run one workgroup of wavefronts per pair when every
  // 5-bit cost entry fits the packed 16-bit table (set by poyhip_set_cm);
  // POYHIP_LINEAR_FILL=wave forces the wavefront-per-pair kernel
  bool lin_mw = false;
  // one-segment triples (L3 <= 1024) fill with one workgroup of wavefronts
  // per triple; POYHIP_CUBE_FILL=wave forces the wavefront-per-triple kernel
  bool cube_mw = true;
  // tuning knobs (poyhip_set_tuning): resident wavefronts per CU of the
  // pair batches, the rotating schedule of gap-free affine pairs, and the
  // workgroup linear fill (used when the costs also fit, lin_fits)
  int waves_per_cu = 16;
  bool aff_rot = true;
  // 2: a workgroup per pair unless the batch alone gives every SIMD four
  // wavefronts of its own pairs; 1: a workgroup per pair; 0: a wavefront
  int lin_wg = 2;
  bool lin_fits = false;
  // largest entry of the 3D cost table, -1 when an entry is negative (the
  // workgroup fill carries 8*value + code keys and needs values below 2^27)
  int cm3_max = -1;
  // largest entry of the 5-bit cost table, -1 when an entry is negative
  // (the diagonal-store aligner bounds its sums below its infinity with it)
  int cm_max = -1;
  // the diagonal-store fill carries both 16-bit gap counts when a longer
  // sequence reaches 2^16; POYHIP_NEWKK_COUNTS=both makes it carry both
  // always (results are the same)
  bool nk_wide_always = false;
  // union merges per wavefront; 0 picks by level size, POYHIP_UNION_LANES
  // (a power of two up to 64) fixes it
  int union_lanes = 0;
  // device address of this device's copy of `scratch` (poyhip_create looks
  // it up; null when the lookup failed, with the reason in scratch_err)
  uint8_t* scratch = nullptr;
  char scratch_err[160] = {0};
  SpareBlocks spare;
};

namespace {

// a batch's device block: an idle one of the context when one fits, else
// hipMalloc
hipError_t ctx_malloc(poyhip_ctx* c, void** p, size_t n) {
  SpareBlocks& sp = c->spare;
  std::lock_guard<std::mutex> g(sp.m);
  int best = -1;
  for (size_t k = 0; k < sp.idle.size(); k++) {
    const size_t sz = sp.idle[k].second;
    if (sz >= n && sz / 2 <= n && (best < 0 || sz < sp.idle[best].second)) best = (int)k;
  }
  if (best >= 0) {
    *p = sp.idle[best].first;
    sp.live.push_back(sp.idle[best]);
    sp.idle_bytes -= sp.idle[best].second;
    sp.idle.erase(sp.idle.begin() + best);
    return hipSuccess;
  }
  const hipError_t e = hipMalloc(p, n);
  if (e == hipSuccess && *p) sp.live.push_back({*p, n});
  return e;
}

// gives a batch's block back to the context (the caller has waited for the
// context's stream)
void ctx_release(poyhip_ctx* c, void* p) {
  if (!p) return;
  SpareBlocks& sp = c->spare;
  std::lock_guard<std::mutex> g(sp.m);
  size_t sz = 0;
  for (size_t k = 0; k < sp.live.size(); k++)
    if (sp.live[k].first == p) {
      sz = sp.live[k].second;
      sp.live.erase(sp.live.begin() + k);
      break;
    }
  if (sz && sz <= SpareBlocks::kSpareBytes - sp.idle_bytes &&
      sp.idle.size() < SpareBlocks::kMaxBlocks) {
    sp.idle.push_back({p, sz});
    sp.idle_bytes += sz;
    return;
  }
  (void)hipFree(p);
}

void ctx_free_idle(poyhip_ctx* c) {
  SpareBlocks& sp = c->spare;
  std::lock_guard<std::mutex> g(sp.m);
  for (auto& blk : sp.idle) (void)hipFree(blk.first);
  sp.idle.clear();
  sp.idle_bytes = 0;
}

int set_err(poyhip_ctx* c, int code, const char* fmt, ...) {
  if (c) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(c->err, sizeof c->err, fmt, ap);
    va_end(ap);
  }
  return code;
}

#define HIPCHK(ctx, x)                                                                 \
  do {                                                                                 \
    hipError_t e_ = (x);                                                               \
    if (e_ != hipSuccess)                                                              \
      return set_err((ctx), POYHIP_EDEVICE, "%s failed: %s", #x, hipGetErrorString(e_)); \
  } while (0)

int grow(poyhip_ctx* c, DevBuf& b, size_t n) {
  if (n == 0) n = 64;
  if (n <= b.cap) return POYHIP_OK;
  if (b.p) (void)hipFree(b.p);
  b.p = nullptr;
  b.cap = 0;
  if (hipMalloc(&b.p, n) != hipSuccess) {
    b.p = nullptr;
    return set_err(c, POYHIP_ENOMEM, "hipMalloc of %zu bytes failed", n);
  }
  b.cap = n;
  return POYHIP_OK;
}

// A drop-in's sequence: present, not empty, and every byte a 5-bit code (the
// kernels index 32 x 32 tables with the raw bytes).
bool valid_seq(const uint8_t* s, int l) {
  if (s == nullptr || l < 1) return false;
  uint8_t any = 0;
  for (int k = 0; k < l; k++) any |= s[k];
  return any < 32;
}

// Offsets of the one-pair descriptor block used by the drop-in calls.
struct OneDesc {
  int64_t off1, off2, dm_off, stab_off, out_off;
  int len_a, len_b, swapped, prio, deltawh, order, counter, pad;
};
constexpr size_t kDescBytes = 128;
static_assert(sizeof(OneDesc) <= kDescBytes, "descriptor block");

PairTable one_table(void* dstage, const uint8_t* dblob) {
  char* base = static_cast<char*>(dstage);
  OneDesc* d = reinterpret_cast<OneDesc*>(base);
  PairTable t;
  t.blob = dblob;
  t.off1 = &d->off1;
  t.off2 = &d->off2;
  t.len_a = &d->len_a;
  t.len_b = &d->len_b;
  t.swapped = &d->swapped;
  t.prio = &d->prio;
  t.deltawh = &d->deltawh;
  t.dm_off = &d->dm_off;
  t.stab_off = &d->stab_off;
  t.out_off = &d->out_off;
  t.order = &d->order;
  t.n = 1;
  t.npairs = 1;
  return t;
}

// Copies the descriptor and the two sequences to the device in one transfer.
// The rows sequence is `a` (la), the columns sequence `b` (lb).
int stage_pair(poyhip_ctx* c, const uint8_t* a, int la, const uint8_t* b, int lb, int prio,
               int deltawh, PairTable* t, int** counter) {
  const size_t need = kDescBytes + (size_t)la + (size_t)lb;
  int rc = grow(c, c->stage, need);
  if (rc) return rc;
  c->h_stage.resize(need);
  OneDesc d;
  memset(&d, 0, sizeof d);
  d.off1 = 0;
  d.off2 = la;
  d.len_a = la;
  d.len_b = lb;
  d.swapped = 0;
  d.prio = prio;
  d.deltawh = deltawh;
  memcpy(c->h_stage.data(), &d, sizeof d);
  memcpy(c->h_stage.data() + kDescBytes, a, la);
  memcpy(c->h_stage.data() + kDescBytes + la, b, lb);
  HIPCHK(c, hipMemcpyAsync(c->stage.p, c->h_stage.data(), need, hipMemcpyHostToDevice, c->stream));
  *t = one_table(c->stage.p, static_cast<const uint8_t*>(c->stage.p) + kDescBytes);
  *counter = &reinterpret_cast<OneDesc*>(c->stage.p)->counter;
  return POYHIP_OK;
}

// Result block of the drop-ins: PairResult, 4 lengths, then the output.
constexpr size_t kResBytes = 64;

// ---- fixed device scratch of poyhip_align_affine_3_noalloc -----------------
// One copy per device, allocated by the code-object loader and held until the
// process ends.  Host code never uses the array's name as an address: each
// device's copy is found once with hipGetSymbolAddress (poyhip_create) and
// that pointer is what kernels and copies receive.
constexpr size_t kScratchBytes = (size_t)64 << 20;
constexpr int kScratchDevices = 64;
static __device__ __align__(256) uint8_t scratch[kScratchBytes];
std::mutex g_scratch_mu;
uint8_t* g_scratch[kScratchDevices] = {};

// Looks up the current device's scratch address (hipSetDevice(device) done by
// the caller); idempotent per device.
uint8_t* scratch_lookup(int device, char* err, size_t errn) {
  if (device < 0 || device >= kScratchDevices) {
    snprintf(err, errn, "device %d beyond the %d supported by the fixed scratch", device,
             kScratchDevices);
    return nullptr;
  }
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (!g_scratch[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      snprintf(err, errn, "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
      return nullptr;
    }
    g_scratch[device] = static_cast<uint8_t*>(p);
  }
  return g_scratch[device];
}

// Scratch layout of one banded affine pair: descriptor, the two sequences,
// direction store, workspace, stripe table, result block.  Returns the bytes
// needed; fills the offsets.
struct AffScratch {
  size_t seq, dm, ws, stab, out, total;
};
AffScratch aff_scratch_layout(int li, int lj) {
  auto up = [](size_t x) { return (x + 255) & ~(size_t)255; };
  AffScratch L;
  L.seq = kDescBytes;
  L.dm = up(L.seq + (size_t)li + (size_t)lj);
  L.ws = up(L.dm + (size_t)aff_dm_bytes(li, lj));
  L.stab = up(L.ws + (size_t)aff_ws_ints(lj) * sizeof(int));
  L.out = up(L.stab + (size_t)(n_stripes(li) + 1) * sizeof(long long));
  L.total = L.out + kResBytes + (size_t)aff_out_bytes(li, lj);
  return L;
}

// Wavefronts per workgroup of the multi-wavefront linear fill and its
// resident workgroups per CU.
constexpr int kLinMwWaves = 8;
constexpr int kLinMwPerCu = 2;

// Workspace ints per slot that either linear fill kernel needs for columns
// of up to Lb symbols.
// The workgroup cube fill takes a one-segment triple when its keys cannot
// wrap: every value of the cube is at most (L1 + L2 + L3) times the largest
// table entry.
bool cube_mw_ok(const poyhip_ctx* c, int64_t len_sum) {
  return c->cube_mw && c->cm3_max >= 0 && (len_sum + 1) * (int64_t)c->cm3_max < ((int64_t)1 << 27);
}

int64_t lin_ws_slot_ints(int Lb) {
  return std::max(lin_ws_ints(Lb), lin_mw_ws_ints(Lb, kLinMwWaves));
}

// Launches the linear fill of A.pt.n pairs: the multi-wavefront kernel (one
// workgroup per pair) when the context's table allows it, else the
// wavefront-per-pair kernel.  `slots` workspace slots of A.ws_stride ints
// are available.  Returns the kernel's launch error.
hipError_t launch_linear_fill(poyhip_ctx* c, const LinFillArgs& A, bool dm, int slots,
                              hipStream_t st) {
  const int n = A.pt.n;
  // with four of its own pairs per SIMD a batch keeps the device busy with
  // one wavefront per pair, which saves the workgroup's block barriers and
  // stripe schedule (a level of ~600-symbol pairs runs ~1.5x faster); the
  // workgroup pays off when fewer, longer pairs must fill the device
  const bool mw = c->lin_mw && (c->lin_wg == 1 || n < 16 * c->num_cus);
  if (mw) {
    const int grid = std::max(1, std::min({n, slots, c->num_cus * kLinMwPerCu}));
    const dim3 blk(64 * kLinMwWaves);
    if (A.phase == 2)
      hipLaunchKernelGGL((k_linear_mw<true, false, kLinMwWaves>), dim3(grid), blk, 0, st, A);
    else if (dm)
      hipLaunchKernelGGL((k_linear_mw<true, true, kLinMwWaves>), dim3(grid), blk, 0, st, A);
    else
      hipLaunchKernelGGL((k_linear_mw<false, true, kLinMwWaves>), dim3(grid), blk, 0, st, A);
  } else {
    const int grid = std::max(1, std::min(n, slots));
    if (dm)
      hipLaunchKernelGGL(k_linear_fill<true>, dim3(grid), dim3(64), 0, st, A);
    else
      hipLaunchKernelGGL(k_linear_fill<false>, dim3(grid), dim3(64), 0, st, A);
  }
  return hipGetLastError();
}

int run_linear_one(poyhip_ctx* c, const uint8_t* s1, int l1, const uint8_t* s2, int l2,
                   int deltawh, int prio, bool backtrace, PairResult* out_res,
                   std::vector<uint8_t>* outblk, int outlen[4], const int* limit = nullptr) {
  PairTable t;
  int* counter;
  int rc = stage_pair(c, s1, l1, s2, l2, prio, deltawh, &t, &counter);
  if (rc) return rc;
  const int64_t dmb = lin_dm_bytes(l1, l2);
  const int64_t outb = lin_out_bytes(l1, l2);
  if ((rc = grow(c, c->dm, dmb))) return rc;
  if ((rc = grow(c, c->ws, lin_ws_slot_ints(l2) * sizeof(int)))) return rc;
  if ((rc = grow(c, c->stab, (n_stripes(l1) + 1) * sizeof(long long)))) return rc;
  if ((rc = grow(c, c->outblk, kResBytes + outb))) return rc;
  PairResult* dres = static_cast<PairResult*>(c->outblk.p);
  int* doutlen = reinterpret_cast<int*>(static_cast<char*>(c->outblk.p) + sizeof(PairResult));
  uint8_t* dout = static_cast<uint8_t*>(c->outblk.p) + kResBytes;
  LinFillArgs A;
  A.pt = t;
  A.cm = c->d_cm;
  A.dm = static_cast<uint8_t*>(c->dm.p);
  A.ws = static_cast<int*>(c->ws.p);
  A.ws_stride = 0;
  A.stab = static_cast<long long*>(c->stab.p);
  A.res = dres;
  A.counter = counter;
  A.phase = 0;
  HIPCHK(c, launch_linear_fill(c, A, true, 1, c->stream));
  size_t back = sizeof(PairResult);
  if (backtrace) {
    LinBtArgs B;
    B.pt = t;
    B.cm = c->d_cm;
    B.dm = A.dm;
    B.stab = A.stab;
    B.res = dres;
    B.out = dout;
    B.outlen = doutlen;
    if (limit) {
      B.lim = 1;
      B.lim_st1 = limit[0];
      B.lim_st2 = limit[1];
      B.lim_a1 = limit[2];
      B.lim_a2 = limit[3];
    }
    hipLaunchKernelGGL(k_linear_backtrace, dim3(1), dim3(64), 0, c->stream, B);
    HIPCHK(c, hipGetLastError());
    back = kResBytes + outb;
  }
  c->h_out.resize(back);
  HIPCHK(c, hipMemcpyAsync(c->h_out.data(), c->outblk.p, back, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  memcpy(out_res, c->h_out.data(), sizeof(PairResult));
  if (backtrace) {
    memcpy(outlen, c->h_out.data() + sizeof(PairResult), 4 * sizeof(int));
    outblk->assign(c->h_out.begin() + kResBytes, c->h_out.begin() + kResBytes + outb);
  }
  return POYHIP_OK;
}

// ---- diagonal-store aligner (newkkonen.c, affine) -------------------------

// Direction stores above this size run the two-phase plan (rounds without
// directions, then the final round alone with a store of its exact size).
constexpr int64_t kNkDirectBytes = (int64_t)8 << 30;

// The aligner's preconditions: an affine table and sums that stay below its
// infinity (INT_MAX / 2) on every path.
int nk_check(poyhip_ctx* c, int l1, int l2) {
  if (!c->has_cm) return set_err(c, POYHIP_ESTATE, "no cost matrix installed");
  if (c->cm.cost_model != 1)
    return set_err(c, POYHIP_EINVAL,
                   "the diagonal-store aligner is built for the affine model only");
  if (l1 > l2) return set_err(c, POYHIP_EINVAL, "pass the shorter sequence first");
  const long long go = std::max(c->cm.gap_open, 0);
  if (c->cm_max < 0 || (c->cm_max + go) * ((long long)l1 + l2 + 2) >= (1ll << 29))
    return set_err(c, POYHIP_EINVAL, "cost entries too large for the aligner's infinity");
  return POYHIP_OK;
}

hipError_t launch_newkk_fill(const NkFillArgs& A, bool dm, bool wide, int grid, hipStream_t st) {
  if (dm && wide)
    hipLaunchKernelGGL((k_newkk_fill<true, true>), dim3(grid), dim3(64), 0, st, A);
  else if (dm)
    hipLaunchKernelGGL((k_newkk_fill<true, false>), dim3(grid), dim3(64), 0, st, A);
  else if (wide)
    hipLaunchKernelGGL((k_newkk_fill<false, true>), dim3(grid), dim3(64), 0, st, A);
  else
    hipLaunchKernelGGL((k_newkk_fill<false, false>), dim3(grid), dim3(64), 0, st, A);
  return hipGetLastError();
}

// One pair (s1 the rows, l1 <= l2) through the batch kernels: fill, and with
// `backtrace` the walk into the output block (regions s1', s2').
int run_newkk_one(poyhip_ctx* c, const uint8_t* s1, int l1, const uint8_t* s2, int l2, int swaped,
                  bool backtrace, PairResult* out_res, std::vector<uint8_t>* outblk,
                  int outlen[4]) {
  PairTable t;
  int* counter;
  int rc = stage_pair(c, s1, l1, s2, l2, swaped ? 1 : 0, 0, &t, &counter);
  if (rc) return rc;
  const bool wide = l2 >= 65536 || c->nk_wide_always;
  const int64_t outb = nk_out_bytes(l1, l2);
  if ((rc = grow(c, c->ws, nk_ws_ints(l2) * sizeof(int)))) return rc;
  if ((rc = grow(c, c->stab, (n_stripes(l1) + 1) * sizeof(long long)))) return rc;
  if ((rc = grow(c, c->outblk, kResBytes + outb))) return rc;
  PairResult* dres = static_cast<PairResult*>(c->outblk.p);
  int* doutlen = reinterpret_cast<int*>(static_cast<char*>(c->outblk.p) + sizeof(PairResult));
  uint8_t* dout = static_cast<uint8_t*>(c->outblk.p) + kResBytes;
  NkFillArgs A;
  A.pt = t;
  A.cm = c->d_cm;
  A.dm = nullptr;
  A.ws = static_cast<int*>(c->ws.p);
  A.ws_stride = 0;
  A.stab = static_cast<long long*>(c->stab.p);
  A.res = dres;
  A.counter = counter;
  // the walk follows the fill in the same kernel (outputs s1', s2')
  A.cm_med = c->d_cm;
  A.out = backtrace ? dout : nullptr;
  A.outlen = doutlen;
  A.medians = 0;
  const int64_t full_dm = nk_dm_bytes(l1, l2);
  if (backtrace && full_dm <= kNkDirectBytes) {
    if ((rc = grow(c, c->dm, full_dm))) return rc;
    A.dm = static_cast<uint8_t*>(c->dm.p);
    A.phase = 0;
    HIPCHK(c, launch_newkk_fill(A, true, wide, 1, c->stream));
  } else {
    A.phase = 1;
    HIPCHK(c, launch_newkk_fill(A, false, wide, 1, c->stream));
    if (backtrace) {
      PairResult r1;
      HIPCHK(c, hipMemcpyAsync(&r1, dres, sizeof r1, hipMemcpyDeviceToHost, c->stream));
      HIPCHK(c, hipStreamSynchronize(c->stream));
      // the trivial case needs no store; its walk still runs in phase 2
      if ((rc = grow(c, c->dm, r1.full == 2 ? 0 : nk_dm_exact(l1, l2, r1.newk)))) return rc;
      A.dm = static_cast<uint8_t*>(c->dm.p);
      A.phase = 2;
      HIPCHK(c, hipMemsetAsync(counter, 0, sizeof(int), c->stream));
      HIPCHK(c, launch_newkk_fill(A, true, wide, 1, c->stream));
    }
  }
  size_t back = sizeof(PairResult);
  if (backtrace) back = kResBytes + outb;
  c->h_out.resize(back);
  HIPCHK(c, hipMemcpyAsync(c->h_out.data(), c->outblk.p, back, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  memcpy(out_res, c->h_out.data(), sizeof(PairResult));
  if (backtrace) {
    memcpy(outlen, c->h_out.data() + sizeof(PairResult), 4 * sizeof(int));
    outblk->assign(c->h_out.begin() + kResBytes, c->h_out.begin() + kResBytes + outb);
  }
  return POYHIP_OK;
}

}  // namespace

extern "C" {

int poyhip_create(int device, poyhip_ctx** out) {
  if (!out) return POYHIP_EINVAL;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev)
    return POYHIP_EDEVICE;
  poyhip_ctx* c = new poyhip_ctx();
  c->device = device;
  if (hipSetDevice(device) != hipSuccess) { delete c; return POYHIP_EDEVICE; }
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return POYHIP_EDEVICE;
  }
  if (hipMalloc(&c->d_cm, sizeof(CmDev)) != hipSuccess) {
    (void)hipStreamDestroy(c->stream);
    delete c;
    return POYHIP_ENOMEM;
  }
  c->scratch = scratch_lookup(device, c->scratch_err, sizeof c->scratch_err);
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->num_cus = prop.multiProcessorCount;
  if (c->num_cus <= 0) c->num_cus = 256;
  const char* e = getenv("POYHIP_CUBE_FILL");
  c->cube_mw = !(e && strcmp(e, "wave") == 0);
  e = getenv("POYHIP_LINEAR_FILL");
  if (e && strcmp(e, "wave") == 0) c->lin_wg = 0;
  e = getenv("POYHIP_NEWKK_COUNTS");
  c->nk_wide_always = e && strcmp(e, "both") == 0;
  e = getenv("POYHIP_UNION_LANES");
  if (e) {
    const int v = atoi(e);
    if (v == 1 || v == 2 || v == 4 || v == 8 || v == 16 || v == 32 || v == 64) c->union_lanes = v;
  }
  *out = c;
  return POYHIP_OK;
}

int poyhip_flush(poyhip_ctx* c) {
  if (!c) return POYHIP_EINVAL;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  for (DevBuf* b : {&c->stage, &c->dm, &c->ws, &c->stab, &c->outblk}) {
    if (b->p) (void)hipFree(b->p);
    b->p = nullptr;
    b->cap = 0;
  }
  ctx_free_idle(c);
  c->last2d = false;
  c->last3d = false;
  return POYHIP_OK;
}

int poyhip_destroy(poyhip_ctx* c) {
  if (!c) return POYHIP_EINVAL;
  poyhip_flush(c);
  if (c->d_cm) (void)hipFree(c->d_cm);
  if (c->d_cm3) (void)hipFree(c->d_cm3);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return POYHIP_OK;
}

const char* poyhip_last_error(const poyhip_ctx* c) { return c ? c->err : "null context"; }

int poyhip_sync(poyhip_ctx* c) {
  if (!c) return POYHIP_EINVAL;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return POYHIP_OK;
}

void* poyhip_stream(poyhip_ctx* c) { return c ? (void*)c->stream : nullptr; }

int poyhip_set_cm(poyhip_ctx* c, const poyhip_cm* cm) {
  if (!c || !cm) return POYHIP_EINVAL;
  if (cm->lcm != 5 || cm->gap != 16)
    return set_err(c, POYHIP_EINVAL, "only 5-bit nucleotide tables (lcm 5, gap 16) are supported");
  (void)hipSetDevice(c->device);
  CmDev h;
  memset(&h, 0, sizeof h);
  memcpy(h.cost, cm->cost, sizeof h.cost);
  memcpy(h.worst, cm->worst, sizeof h.worst);
  memcpy(h.prepend, cm->prepend, sizeof h.prepend);
  memcpy(h.tail, cm->tail, sizeof h.tail);
  memcpy(h.median, cm->median, sizeof h.median);
  h.gap_open = cm->gap_open;
  // cm_get_min_non0_cost (src/cm.c:1064) scans all 2*32*32 entries
  int m = 0x3fffffff;
  for (int k = 0; k < 2048; k++)
    if (cm->cost[k] > 0 && cm->cost[k] < m) m = cm->cost[k];
  h.delta = m;
  h.cost_model = cm->cost_model;
  h.combinations = cm->combinations;
  HIPCHK(c, hipMemcpyAsync(c->d_cm, &h, sizeof h, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->cm = *cm;
  c->has_cm = true;
  c->last2d = false;
  bool fits = true;
  for (int k = 0; k < 1024; k++) fits &= cm->cost[k] >= 0 && cm->cost[k] < 32768;
  c->lin_fits = fits;
  c->lin_mw = fits && c->lin_wg != 0;
  int mx = 0;
  bool neg = false;
  for (int k = 0; k < 1024; k++) {
    neg |= cm->cost[k] < 0;
    mx = std::max(mx, cm->cost[k]);
  }
  c->cm_max = neg ? -1 : mx;
  return POYHIP_OK;
}

int poyhip_cost_2(poyhip_ctx* c, const uint8_t* s1, int l1, const uint8_t* s2, int l2,
                  int deltawh, int* cost) {
  if (!c || !cost || !valid_seq(s1, l1) || !valid_seq(s2, l2)) return POYHIP_EINVAL;
  if (!c->has_cm) return set_err(c, POYHIP_ESTATE, "no cost matrix installed");
  if (l1 > l2) return set_err(c, POYHIP_EINVAL, "pass the shorter sequence first");
  (void)hipSetDevice(c->device);
  PairResult r;
  int outlen[4];
  int rc = run_linear_one(c, s1, l1, s2, l2, deltawh, 0, false, &r, nullptr, outlen);
  if (rc) return rc;
  *cost = r.cost;
  c->last2d = true;
  c->last_deltawh = deltawh;
  c->last_s1.assign(s1, s1 + l1);
  c->last_s2.assign(s2, s2 + l2);
  return POYHIP_OK;
}

int poyhip_limit_2(poyhip_ctx* c, const uint8_t* s1, int l1, const uint8_t* s2, int l2, int w,
                   int h, int s1_st, int s2_st, int s1_len, int s2_len, int* cost) {
  (void)h;      // algn_CAML_limit_2 reads neither h
  (void)s1_st;  // nor the start positions
  (void)s2_st;
  // algn_nw_limit's height is 50 + (s1_len - s2_len) + w where algn_nw's is
  // 50 + (l1 - l2) + deltawh: the same decision with this deltawh
  const int64_t dw = (int64_t)w + ((int64_t)s1_len - s2_len) - ((int64_t)l1 - l2);
  if (dw < INT_MIN || dw > INT_MAX) return c ? set_err(c, POYHIP_EINVAL, "limits out of range") : POYHIP_EINVAL;
  return poyhip_cost_2(c, s1, l1, s2, l2, (int)dw, cost);
}

int poyhip_backtrace_2d_limit(poyhip_ctx* c, const uint8_t* s1, int l1, const uint8_t* s2,
                              int l2, uint8_t* r1, uint8_t* r2, int cap, int* len, int st_s1,
                              int st_s2, int algn_s1, int algn_s2, int swaped) {
  if (!c || !r1 || !r2 || !len || !valid_seq(s1, l1) || !valid_seq(s2, l2)) return POYHIP_EINVAL;
  if (st_s1 < 0 || st_s2 < 0 || algn_s1 < 0 || algn_s2 < 0 || st_s1 + algn_s1 > l1 ||
      st_s2 + algn_s2 > l2)
    return set_err(c, POYHIP_EINVAL, "limits outside the sequences");
  if (!c->last2d || (int)c->last_s1.size() != l1 || (int)c->last_s2.size() != l2 ||
      memcmp(c->last_s1.data(), s1, l1) || memcmp(c->last_s2.data(), s2, l2))
    return set_err(c, POYHIP_ESTATE, "backtrace without a matching poyhip_cost_2 / poyhip_limit_2");
  if (cap < l1 + l2) return set_err(c, POYHIP_ECAP, "capacity %d < %d", cap, l1 + l2);
  (void)hipSetDevice(c->device);
  PairResult r;
  int outlen[4];
  std::vector<uint8_t> blk;
  const int limit[4] = {st_s1, st_s2, algn_s1, algn_s2};
  int rc = run_linear_one(c, s1, l1, s2, l2, c->last_deltawh, swaped ? 1 : 0, true, &r, &blk,
                          outlen, limit);
  if (rc) return rc;
  const int cp = (int)lin_out_cap(l1, l2);
  *len = outlen[0];
  memcpy(r1, blk.data() + cp - outlen[0], outlen[0]);
  memcpy(r2, blk.data() + cp + cp - outlen[1], outlen[1]);
  return POYHIP_OK;
}

int poyhip_backtrace_2d(poyhip_ctx* c, const uint8_t* s1, int l1, const uint8_t* s2, int l2,
                        uint8_t* r1, uint8_t* r2, int cap, int* len, int swaped) {
  if (!c || !r1 || !r2 || !len || !valid_seq(s1, l1) || !valid_seq(s2, l2)) return POYHIP_EINVAL;
  if (!c->last2d || (int)c->last_s1.size() != l1 || (int)c->last_s2.size() != l2 ||
      memcmp(c->last_s1.data(), s1, l1) || memcmp(c->last_s2.data(), s2, l2))
    return set_err(c, POYHIP_ESTATE, "backtrace without a matching poyhip_cost_2");
  if (cap < l1 + l2) return set_err(c, POYHIP_ECAP, "capacity %d < %d", cap, l1 + l2);
  (void)hipSetDevice(c->device);
  // The stored moves are resolved for one tie order; refill when the
  // requested order differs (the fill is deterministic).
  PairResult r;
  int outlen[4];
  std::vector<uint8_t> blk;
  int rc = run_linear_one(c, s1, l1, s2, l2, c->last_deltawh, swaped ? 1 : 0, true, &r, &blk,
                          outlen);
  if (rc) return rc;
  const int cp = (int)lin_out_cap(l1, l2);
  *len = outlen[0];
  memcpy(r1, blk.data() + cp - outlen[0], outlen[0]);
  memcpy(r2, blk.data() + cp + cp - outlen[1], outlen[1]);
  return POYHIP_OK;
}

int poyhip_align_2d(poyhip_ctx* c, const uint8_t* s1, int l1, const uint8_t* s2, int l2,
                    int deltawh, int swaped, uint8_t* r1, uint8_t* r2, int cap, int* len,
                    int* cost) {
  if (!c || !cost || !r1 || !r2 || !len || !valid_seq(s1, l1) || !valid_seq(s2, l2))
    return POYHIP_EINVAL;
  if (!c->has_cm) return set_err(c, POYHIP_ESTATE, "no cost matrix installed");
  if (l1 > l2) return set_err(c, POYHIP_EINVAL, "pass the shorter sequence first");
  if (cap < l1 + l2) return set_err(c, POYHIP_ECAP, "capacity %d < %d", cap, l1 + l2);
  (void)hipSetDevice(c->device);
  PairResult r;
  int outlen[4];
  std::vector<uint8_t> blk;
  int rc = run_linear_one(c, s1, l1, s2, l2, deltawh, swaped ? 1 : 0, true, &r, &blk, outlen);
  if (rc) return rc;
  const int cp = (int)lin_out_cap(l1, l2);
  *cost = r.cost;
  *len = outlen[0];
  memcpy(r1, blk.data() + cp - outlen[0], outlen[0]);
  memcpy(r2, blk.data() + cp + cp - outlen[1], outlen[1]);
  c->last2d = true;
  c->last_deltawh = deltawh;
  c->last_s1.assign(s1, s1 + l1);
  c->last_s2.assign(s2, s2 + l2);
  return POYHIP_OK;
}

int poyhip_newkk_cost_2_affine(poyhip_ctx* c, const uint8_t* s1, int l1, const uint8_t* s2,
                               int l2, int swaped, int* cost) {
  if (!c || !cost || !valid_seq(s1, l1) || !valid_seq(s2, l2)) return POYHIP_EINVAL;
  int rc = nk_check(c, l1, l2);
  if (rc) return rc;
  (void)hipSetDevice(c->device);
  PairResult r;
  int outlen[4];
  rc = run_newkk_one(c, s1, l1, s2, l2, swaped, false, &r, nullptr, outlen);
  if (rc) return rc;
  *cost = r.cost;
  return POYHIP_OK;
}

int poyhip_newkk_align_2_affine(poyhip_ctx* c, const uint8_t* s1, int l1, const uint8_t* s2,
                                int l2, int swaped, int* cost, uint8_t* r1, uint8_t* r2, int cap,
                                int* len) {
  if (!c || !cost || !r1 || !r2 || !len || !valid_seq(s1, l1) || !valid_seq(s2, l2))
    return POYHIP_EINVAL;
  int rc = nk_check(c, l1, l2);
  if (rc) return rc;
  if (cap < l1 + l2) return set_err(c, POYHIP_ECAP, "output capacity %d < %d", cap, l1 + l2);
  (void)hipSetDevice(c->device);
  PairResult r;
  int outlen[4];
  std::vector<uint8_t> blk;
  rc = run_newkk_one(c, s1, l1, s2, l2, swaped, true, &r, &blk, outlen);
  if (rc) return rc;
  const int n = outlen[0];
  if (n < 0 || n > cap)
    return set_err(c, POYHIP_EDEVICE, "backtrace failed (length %d)", n);
  const int bc = (int)nk_out_cap(l1, l2);
  memcpy(r1, blk.data() + bc - n, n);
  memcpy(r2, blk.data() + 2 * bc - n, n);
  *len = n;
  *cost = r.cost;
  return POYHIP_OK;
}

int poyhip_union_merge_batch(poyhip_ctx* c, int n, const poyhip_union_job* d_jobs,
                             uint8_t* d_seq, int64_t seq_bytes, int32_t* d_ints, int64_t n_ints,
                             int* d_c_len, int* d_status, int sync) {
  if (!c || n < 0 || (n > 0 && (!d_jobs || !d_seq || !d_ints || !d_c_len || !d_status)))
    return POYHIP_EINVAL;
  if (n == 0) return POYHIP_OK;
  (void)hipSetDevice(c->device);
  // jobs per wavefront: a level's time is its longest walk, and a walk's
  // steps are quicker the fewer lanes of its wavefront still issue loads, so
  // wavefronts stay small (4 merges) unless the level has more than 16 per
  // CU at that size
  int lanes = c->union_lanes;
  if (lanes <= 0) {
    lanes = 64;
    while (lanes > 4 && (int64_t)((n + lanes - 1) / lanes) < 16 * (int64_t)c->num_cus) lanes >>= 1;
  }
  hipLaunchKernelGGL(k_union_merge, dim3((n + lanes - 1) / lanes), dim3(64), 0, c->stream, d_jobs,
                     n, lanes, d_seq, seq_bytes, d_ints, n_ints, d_c_len, d_status);
  HIPCHK(c, hipGetLastError());
  if (sync) HIPCHK(c, hipStreamSynchronize(c->stream));
  return POYHIP_OK;
}

int poyhip_union_merge(poyhip_ctx* c, const uint8_t* a, const uint8_t* b, const uint8_t* median,
                       int lena, const uint8_t* ua_s, int ua_len, int ua_cap,
                       const int32_t* ua_off, const uint8_t* ub_s, int ub_len, int ub_cap,
                       const int32_t* ub_off, uint8_t* c_s, int c_cap, int32_t* c_off,
                       int32_t* c_c1, int32_t* c_c2, int* c_len) {
  if (!c || !a || !b || !median || !ua_s || !ua_off || !ub_s || !ub_off || !c_s || !c_off ||
      !c_c1 || !c_c2 || !c_len || lena < 1 || ua_cap < 1 || ub_cap < 1 || c_cap < 1 ||
      ua_len < 0 || ub_len < 0 || ua_len > ua_cap || ub_len > ub_cap)
    return POYHIP_EINVAL;
  (void)hipSetDevice(c->device);
  // one staging block: job, sequences, then the int arrays (4-byte aligned)
  const int64_t nseq = 3 * (int64_t)lena + ua_cap + ub_cap + c_cap;
  const int64_t ints_at = (kDescBytes + nseq + 15) & ~(int64_t)15;
  const int64_t nints = (int64_t)ua_cap + ub_cap + 3 * (int64_t)c_cap;
  const int64_t res_at = ints_at + 4 * nints;
  const size_t need = (size_t)res_at + 16;
  static_assert(sizeof(poyhip_union_job) <= kDescBytes, "union job in the descriptor block");
  int rc = grow(c, c->stage, need);
  if (rc) return rc;
  c->h_stage.assign(need, 0);
  uint8_t* h = c->h_stage.data();
  poyhip_union_job J;
  memset(&J, 0, sizeof J);
  int64_t o = 0;
  auto put = [&](const uint8_t* src, int64_t len) {
    const int64_t at = o;
    memcpy(h + kDescBytes + o, src, (size_t)len);
    o += len;
    return at;
  };
  J.a = put(a, lena);
  J.b = put(b, lena);
  J.median = put(median, lena);
  J.ua_s = put(ua_s, ua_cap);
  J.ub_s = put(ub_s, ub_cap);
  J.c_s = o;
  J.lena = lena;
  J.ua_len = ua_len;
  J.ua_cap = ua_cap;
  J.ub_len = ub_len;
  J.ub_cap = ub_cap;
  J.c_cap = c_cap;
  int32_t* hi = reinterpret_cast<int32_t*>(h + ints_at);
  memcpy(hi, ua_off, 4 * (size_t)ua_cap);
  memcpy(hi + ua_cap, ub_off, 4 * (size_t)ub_cap);
  J.ua_off = 0;
  J.ub_off = ua_cap;
  J.c_off = (int64_t)ua_cap + ub_cap;
  J.c_c1 = J.c_off + c_cap;
  J.c_c2 = J.c_c1 + c_cap;
  memcpy(h, &J, sizeof J);
  uint8_t* d = static_cast<uint8_t*>(c->stage.p);
  HIPCHK(c, hipMemcpyAsync(d, h, (size_t)res_at, hipMemcpyHostToDevice, c->stream));
  int* dres = reinterpret_cast<int*>(d + res_at);
  hipLaunchKernelGGL(k_union_merge, dim3(1), dim3(64), 0, c->stream,
                     reinterpret_cast<const poyhip_union_job*>(d), 1, 1, d + kDescBytes,
                     (int64_t)nseq, reinterpret_cast<int32_t*>(d + ints_at), nints, dres, dres + 1);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipMemcpyAsync(h, d, need, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const int* hres = reinterpret_cast<const int*>(h + res_at);
  if (hres[1] != POYHIP_OK)
    return set_err(c, POYHIP_EINVAL, "union_merge fails on this input (or c_cap too small)");
  memcpy(c_s, h + kDescBytes + J.c_s, (size_t)c_cap);
  memcpy(c_off, hi + J.c_off, 4 * (size_t)c_cap);
  memcpy(c_c1, hi + J.c_c1, 4 * (size_t)c_cap);
  memcpy(c_c2, hi + J.c_c2, 4 * (size_t)c_cap);
  *c_len = hres[0];
  return POYHIP_OK;
}

int poyhip_cost_affine_3(poyhip_ctx* c, const uint8_t* si, int li, const uint8_t* sj, int lj,
                         int* cost) {
  if (!c || !cost || !valid_seq(si, li) || !valid_seq(sj, lj)) return POYHIP_EINVAL;
  if (!c->has_cm) return set_err(c, POYHIP_ESTATE, "no cost matrix installed");
  (void)hipSetDevice(c->device);
  const uint8_t *a = si, *b = sj;
  int la = li, lb = lj;
  if (li > lj) { a = sj; b = si; la = lj; lb = li; }
  PairTable t;
  int* counter;
  int rc = stage_pair(c, a, la, b, lb, 0, 0, &t, &counter);
  if (rc) return rc;
  if ((rc = grow(c, c->ws, affcost_ws_ints(lb) * sizeof(int)))) return rc;
  if ((rc = grow(c, c->outblk, kResBytes))) return rc;
  AffCostArgs A;
  A.pt = t;
  A.cm = c->d_cm;
  A.ws = static_cast<int*>(c->ws.p);
  A.ws_stride = 0;
  A.res = static_cast<PairResult*>(c->outblk.p);
  A.counter = counter;
  hipLaunchKernelGGL(k_affine_cost_full, dim3(1), dim3(64), 0, c->stream, A);
  HIPCHK(c, hipGetLastError());
  PairResult r;
  HIPCHK(c, hipMemcpyAsync(&r, c->outblk.p, sizeof r, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  *cost = r.cost;
  return POYHIP_OK;
}

int poyhip_align_affine_3(poyhip_ctx* c, const uint8_t* si, int li, const uint8_t* sj, int lj,
                          int swaped, uint8_t* median, int* median_len, uint8_t* medianwg,
                          int* medianwg_len, uint8_t* resi, uint8_t* resj, int* res_len, int cap,
                          int* cost) {
  if (!c || !cost || !median || !median_len || !medianwg || !medianwg_len || !resi || !resj ||
      !res_len || !valid_seq(si, li) || !valid_seq(sj, lj))
    return POYHIP_EINVAL;
  if (!c->has_cm) return set_err(c, POYHIP_ESTATE, "no cost matrix installed");
  if (lj < li) return set_err(c, POYHIP_EINVAL, "pass the shorter one as first");
  if (cap < li + lj + 2) return set_err(c, POYHIP_ECAP, "capacity %d < %d", cap, li + lj + 2);
  (void)hipSetDevice(c->device);
  PairTable t;
  int* counter;
  int rc = stage_pair(c, si, li, sj, lj, swaped ? 1 : 0, 0, &t, &counter);
  if (rc) return rc;
  const int64_t dmb = aff_dm_bytes(li, lj);
  const int64_t outb = aff_out_bytes(li, lj);
  if ((rc = grow(c, c->dm, dmb))) return rc;
  if ((rc = grow(c, c->ws, aff_ws_ints(lj) * sizeof(int)))) return rc;
  if ((rc = grow(c, c->stab, (n_stripes(li) + 1) * sizeof(long long)))) return rc;
  if ((rc = grow(c, c->outblk, kResBytes + outb))) return rc;
  PairResult* dres = static_cast<PairResult*>(c->outblk.p);
  int* doutlen = reinterpret_cast<int*>(static_cast<char*>(c->outblk.p) + sizeof(PairResult));
  uint8_t* dout = static_cast<uint8_t*>(c->outblk.p) + kResBytes;
  AffFillArgs A;
  A.rotating = c->aff_rot ? 1 : 0;
  A.pt = t;
  A.cm = c->d_cm;
  A.dm = static_cast<uint8_t*>(c->dm.p);
  A.ws = static_cast<int*>(c->ws.p);
  A.ws_stride = 0;
  A.stab = static_cast<long long*>(c->stab.p);
  A.res = dres;
  A.counter = counter;
  A.cls = nullptr;
  A.ngen = nullptr;
  int any = 0;
  for (int k = 1; k < li; k++) any |= si[k];
  for (int k = 1; k < lj; k++) any |= sj[k];
  if (any & kGap) {
    hipLaunchKernelGGL((k_affine_fill<true, false>), dim3(1), dim3(64), 0, c->stream, A);
    HIPCHK(c, hipGetLastError());
    AffBtArgs B;
    B.pt = t;
    B.cm = c->d_cm;
    B.dm = A.dm;
    B.stab = A.stab;
    B.res = dres;
    B.out = dout;
    B.outlen = doutlen;
    hipLaunchKernelGGL(k_affine_backtrace, dim3(1), dim3(64), 0, c->stream, B);
  } else {
    // the gap-free kernel runs the backtrace after the fill
    A.out = dout;
    A.outlen = doutlen;
    hipLaunchKernelGGL((k_affine_fill<true, true>), dim3(1), dim3(64), 0, c->stream, A);
  }
  HIPCHK(c, hipGetLastError());
  const size_t back = kResBytes + outb;
  c->h_out.resize(back);
  HIPCHK(c, hipMemcpyAsync(c->h_out.data(), c->outblk.p, back, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  PairResult r;
  int ol[4];
  memcpy(&r, c->h_out.data(), sizeof r);
  memcpy(ol, c->h_out.data() + sizeof(PairResult), sizeof ol);
  const uint8_t* blk = c->h_out.data() + kResBytes;
  const int cp = li + lj + 2;
  memcpy(resi, blk + cp - ol[0], ol[0]);
  memcpy(resj, blk + 2 * cp - ol[1], ol[1]);
  memcpy(median, blk + 3 * cp - ol[2], ol[2]);
  memcpy(medianwg, blk + 4 * cp - ol[3], ol[3]);
  *res_len = ol[0];
  *median_len = ol[2];
  *medianwg_len = ol[3];
  *cost = r.cost;
  return POYHIP_OK;
}

int poyhip_align_affine_3_noalloc(poyhip_ctx* c, const uint8_t* si, int li, const uint8_t* sj,
                                  int lj, int swaped, uint8_t* median, int* median_len,
                                  uint8_t* medianwg, int* medianwg_len, uint8_t* resi,
                                  uint8_t* resj, int* res_len, int cap, int* cost) {
  if (!c || !cost || !median || !median_len || !medianwg || !medianwg_len || !resi || !resj ||
      !res_len || !valid_seq(si, li) || !valid_seq(sj, lj))
    return POYHIP_EINVAL;
  if (!c->has_cm) return set_err(c, POYHIP_ESTATE, "no cost matrix installed");
  if (lj < li) return set_err(c, POYHIP_EINVAL, "pass the shorter one as first");
  if (cap < li + lj + 2) return set_err(c, POYHIP_ECAP, "capacity %d < %d", cap, li + lj + 2);
  if (!c->scratch) return set_err(c, POYHIP_EDEVICE, "no fixed scratch: %s", c->scratch_err);
  const AffScratch L = aff_scratch_layout(li, lj);
  if (L.total > kScratchBytes)
    return set_err(c, POYHIP_ECAP, "pair of %d x %d needs %zu bytes of scratch, over the %zu fixed",
                   li, lj, L.total, kScratchBytes);
  (void)hipSetDevice(c->device);
  uint8_t* S = c->scratch;
  // descriptor (work counter at zero) and sequences, straight from the
  // caller's buffers
  OneDesc d;
  memset(&d, 0, sizeof d);
  d.off2 = li;
  d.len_a = li;
  d.len_b = lj;
  d.prio = swaped ? 1 : 0;
  HIPCHK(c, hipMemcpyAsync(S, &d, sizeof d, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(S + L.seq, si, (size_t)li, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(S + L.seq + li, sj, (size_t)lj, hipMemcpyHostToDevice, c->stream));
  const PairTable t = one_table(S, S + L.seq);
  PairResult* dres = reinterpret_cast<PairResult*>(S + L.out);
  int* doutlen = reinterpret_cast<int*>(S + L.out + sizeof(PairResult));
  uint8_t* dout = S + L.out + kResBytes;
  AffFillArgs A;
  A.rotating = c->aff_rot ? 1 : 0;
  A.pt = t;
  A.cm = c->d_cm;
  A.dm = S + L.dm;
  A.ws = reinterpret_cast<int*>(S + L.ws);
  A.ws_stride = 0;
  A.stab = reinterpret_cast<long long*>(S + L.stab);
  A.res = dres;
  A.counter = &reinterpret_cast<OneDesc*>(S)->counter;
  A.cls = nullptr;
  A.ngen = nullptr;
  int any = 0;
  for (int k = 1; k < li; k++) any |= si[k];
  for (int k = 1; k < lj; k++) any |= sj[k];
  if (any & kGap) {
    hipLaunchKernelGGL((k_affine_fill<true, false>), dim3(1), dim3(64), 0, c->stream, A);
    HIPCHK(c, hipGetLastError());
    AffBtArgs B;
    B.pt = t;
    B.cm = c->d_cm;
    B.dm = A.dm;
    B.stab = A.stab;
    B.res = dres;
    B.out = dout;
    B.outlen = doutlen;
    hipLaunchKernelGGL(k_affine_backtrace, dim3(1), dim3(64), 0, c->stream, B);
  } else {
    A.out = dout;
    A.outlen = doutlen;
    hipLaunchKernelGGL((k_affine_fill<true, true>), dim3(1), dim3(64), 0, c->stream, A);
  }
  HIPCHK(c, hipGetLastError());
  // result header first, then each output region straight into the
  // caller's buffer
  uint8_t hdr[kResBytes];
  HIPCHK(c, hipMemcpyAsync(hdr, S + L.out, kResBytes, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  PairResult r;
  int ol[4];
  memcpy(&r, hdr, sizeof r);
  memcpy(ol, hdr + sizeof(PairResult), sizeof ol);
  const int cp = li + lj + 2;
  for (int q = 0; q < 4; q++)
    if (ol[q] < 0 || ol[q] > cp) return set_err(c, POYHIP_EDEVICE, "output length %d out of range", ol[q]);
  uint8_t* dst[4] = {resi, resj, median, medianwg};
  for (int q = 0; q < 4; q++)
    if (ol[q] > 0)
      HIPCHK(c, hipMemcpyAsync(dst[q], dout + (size_t)(q + 1) * cp - ol[q], (size_t)ol[q],
                               hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  *res_len = ol[0];
  *median_len = ol[2];
  *medianwg_len = ol[3];
  *cost = r.cost;
  return POYHIP_OK;
}

static int pair_median_op(poyhip_ctx* c, const uint8_t* a, const uint8_t* b, int len, int op,
                          uint8_t* out, int cap, int* out_len, int* value) {
  if (!c || !a || !b || len < 0) return POYHIP_EINVAL;
  if (!c->has_cm) return set_err(c, POYHIP_ESTATE, "no cost matrix installed");
  (void)hipSetDevice(c->device);
  const size_t need = kDescBytes + 2 * (size_t)len + 8;
  int rc = grow(c, c->stage, need);
  if (rc) return rc;
  const int ocap = len + 2;
  if ((rc = grow(c, c->outblk, kResBytes + ocap))) return rc;
  c->h_stage.resize(need);
  memcpy(c->h_stage.data() + kDescBytes, a, len);
  memcpy(c->h_stage.data() + kDescBytes + len, b, len);
  HIPCHK(c, hipMemcpyAsync(c->stage.p, c->h_stage.data(), need, hipMemcpyHostToDevice, c->stream));
  const uint8_t* da = static_cast<const uint8_t*>(c->stage.p) + kDescBytes;
  int* dres = static_cast<int*>(c->outblk.p);
  uint8_t* dout = static_cast<uint8_t*>(c->outblk.p) + kResBytes;
  hipLaunchKernelGGL(k_pair_median, dim3(1), dim3(64), 0, c->stream, c->d_cm, da, da + len, len,
                     op, dout, ocap, dres);
  HIPCHK(c, hipGetLastError());
  c->h_out.resize(kResBytes + ocap);
  HIPCHK(c, hipMemcpyAsync(c->h_out.data(), c->outblk.p, kResBytes + ocap, hipMemcpyDeviceToHost,
                           c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  int res[2];
  memcpy(res, c->h_out.data(), sizeof res);
  if (res[1]) return set_err(c, POYHIP_EINVAL, "median should not be 0");
  if (op == kWorst) {
    *value = res[0];
    return POYHIP_OK;
  }
  if (res[0] > cap) return set_err(c, POYHIP_ECAP, "capacity %d < %d", cap, res[0]);
  memcpy(out, c->h_out.data() + kResBytes + ocap - res[0], res[0]);
  *out_len = res[0];
  return POYHIP_OK;
}

int poyhip_ancestor_2(poyhip_ctx* c, const uint8_t* a, const uint8_t* b, int len, uint8_t* out,
                      int cap, int* out_len) {
  if (!out || !out_len) return POYHIP_EINVAL;
  return pair_median_op(c, a, b, len, kAncestor, out, cap, out_len, nullptr);
}

int poyhip_median_2(poyhip_ctx* c, const uint8_t* a, const uint8_t* b, int len, uint8_t* out,
                    int cap, int* out_len) {
  if (!out || !out_len) return POYHIP_EINVAL;
  return pair_median_op(c, a, b, len, kMedianNoGaps, out, cap, out_len, nullptr);
}

int poyhip_median_2_with_gaps(poyhip_ctx* c, const uint8_t* a, const uint8_t* b, int len,
                              uint8_t* out, int cap, int* out_len) {
  if (!out || !out_len) return POYHIP_EINVAL;
  return pair_median_op(c, a, b, len, kMedianWithGaps, out, cap, out_len, nullptr);
}

int poyhip_worst_2(poyhip_ctx* c, const uint8_t* a, const uint8_t* b, int len, int* cost) {
  if (!cost) return POYHIP_EINVAL;
  return pair_median_op(c, a, b, len, kWorst, nullptr, 0, nullptr, cost);
}

int poyhip_closest_2(poyhip_ctx* c, const uint8_t* a, const uint8_t* b, int len, int equal,
                     uint8_t* out, int cap, int* out_len) {
  if (!c || !out || !out_len) return POYHIP_EINVAL;
  if (c->has_cm && !c->cm.combinations)
    return set_err(c, POYHIP_EINVAL, "closest: the device map needs a table with combinations");
  return pair_median_op(c, a, b, len, equal ? kClosestEq : kClosest, out, cap, out_len, nullptr);
}

int poyhip_pairs_post_batch(poyhip_ctx* c, int op, int n, const poyhip_post_job* d_jobs,
                            uint8_t* d_seq, int64_t seq_bytes, int* d_out_len, int* d_status,
                            int sync) {
  if (!c || n < 0 || (n > 0 && (!d_jobs || !d_seq || !d_out_len || !d_status)))
    return POYHIP_EINVAL;
  if (op != POYHIP_POST_MEDIAN_2 && op != POYHIP_POST_MEDIAN_2_WITH_GAPS &&
      op != POYHIP_POST_CLOSEST && op != POYHIP_POST_CLOSEST_EQ)
    return set_err(c, POYHIP_EINVAL, "unknown column map %d", op);
  if (!c->has_cm) return set_err(c, POYHIP_ESTATE, "no cost matrix installed");
  if ((op == POYHIP_POST_CLOSEST || op == POYHIP_POST_CLOSEST_EQ) && !c->cm.combinations)
    return set_err(c, POYHIP_EINVAL, "closest: the device map needs a table with combinations");
  if (n == 0) return POYHIP_OK;
  static_assert(POYHIP_POST_MEDIAN_2 == kMedianNoGaps && POYHIP_POST_MEDIAN_2_WITH_GAPS ==
                    kMedianWithGaps && POYHIP_POST_CLOSEST == kClosest &&
                    POYHIP_POST_CLOSEST_EQ == kClosestEq,
                "column-map codes");
  (void)hipSetDevice(c->device);
  hipLaunchKernelGGL(k_pairs_post, dim3((n + 3) / 4), dim3(256), 0, c->stream, c->d_cm, op,
                     d_jobs, n, d_seq, seq_bytes, d_out_len, d_status);
  HIPCHK(c, hipGetLastError());
  if (sync) HIPCHK(c, hipStreamSynchronize(c->stream));
  return POYHIP_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Batches
// ---------------------------------------------------------------------------

struct poyhip_batch {
  poyhip_ctx* ctx = nullptr;
  int op = 0;
  int n = 0;
  std::vector<int> la, lb, swapped;
  std::vector<int64_t> out_off, dm_off, stab_off;
  std::vector<int> order;
  struct Chunk {
    int begin, count;
    int64_t dm_bytes;
  };
  std::vector<Chunk> chunks;
  int has_deltaw = 0;
  // device
  void* d_tab = nullptr;  // all descriptor arrays in one allocation
  int64_t *d_dm_off = nullptr, *d_stab_off = nullptr, *d_out_off = nullptr;
  int *d_la = nullptr, *d_lb = nullptr, *d_sw = nullptr, *d_dw = nullptr, *d_dwin = nullptr;
  int* d_order = nullptr;
  int* d_counters = nullptr;
  uint8_t* d_cls = nullptr;
  PairResult* d_res = nullptr;
  uint8_t* d_dm = nullptr;
  int* d_ws = nullptr;
  long long* d_stab = nullptr;
  int64_t dm_cap = 0;
  int grid = 0;
  long long ws_stride = 0;
  size_t out_bytes = 0;
  // timing
  std::vector<hipEvent_t> ev;
  float fill_ms = 0.f, total_ms = 0.f;
  // two-phase plan (long linear pairs): cost-only rounds for every pair,
  // then direction stores sized to each pair's final band
  bool two_phase = false;
  int64_t dm_budget = 0;
  // diagonal-store batches: some pair's longer sequence reaches 2^16 (both
  // 16-bit gap counts carried)
  bool wide = false;
  std::vector<PairResult> h_res;
  // graph replay
  bool use_graph = false;
  hipGraphExec_t gexec = nullptr;
  const void* g_key[6] = {nullptr};
};

namespace {

bool op_is_affine(int op) { return op == POYHIP_OP_ALIGN_AFFINE || op == POYHIP_OP_COST_AFFINE; }
bool op_is_newkk(int op) { return op == POYHIP_OP_NEWKK_ALIGN || op == POYHIP_OP_NEWKK_COST; }
bool op_is_linear(int op) { return op == POYHIP_OP_ALIGN_2D || op == POYHIP_OP_COST_2D; }
bool op_has_dm(int op) {
  return op == POYHIP_OP_ALIGN_AFFINE || op == POYHIP_OP_ALIGN_2D || op == POYHIP_OP_NEWKK_ALIGN;
}

// Direction-store budget of a batch when the caller gives none: the free
// device memory less an eighth of the device (at least 8 GiB) left to other
// allocations.  On a 288 GB MI355X a batch of 65,536 pairs of up to 5,000
// symbols then runs as one chunk, so the persistent fill balances its
// largest pairs against the whole batch instead of half of it.
size_t auto_dm_budget() {
  size_t fr = 0, tot = 0;
  if (hipMemGetInfo(&fr, &tot) != hipSuccess) return (size_t)4 << 30;
  const size_t reserve = std::max(tot / 8, (size_t)8 << 30);
  return fr > 2 * reserve ? fr - reserve : fr / 2;
}

int64_t pair_dm_bytes(int op, int la, int lb) {
  if (op == POYHIP_OP_ALIGN_AFFINE) return aff_dm_bytes(la, lb);
  if (op == POYHIP_OP_ALIGN_2D) return lin_dm_bytes(la, lb);
  if (op == POYHIP_OP_NEWKK_ALIGN) return nk_dm_bytes(la, lb);
  return 0;
}
int64_t pair_out_bytes(int op, int la, int lb) {
  if (op == POYHIP_OP_ALIGN_AFFINE) return aff_out_bytes(la, lb);
  if (op == POYHIP_OP_ALIGN_2D) return lin_out_bytes(la, lb);
  if (op == POYHIP_OP_NEWKK_ALIGN) return nk_out_bytes(la, lb);
  return 0;
}
int64_t ws_ints(int op, int lb) {
  if (op == POYHIP_OP_ALIGN_AFFINE) return aff_ws_ints(lb);
  if (op == POYHIP_OP_COST_AFFINE) return affcost_ws_ints(lb);
  if (op_is_newkk(op)) return nk_ws_ints(lb);
  return lin_ws_slot_ints(lb);
}

// Direction bytes of the final round of a linear pair with band half-width
// newk (the stripe loop of k_linear_fill).
int64_t lin_dm_exact(int L1, int L2, int newk, bool full) {
  const int Dd = L2 - L1;
  int64_t tot = 0;
  for (int i0 = 1; i0 <= L1 - 1; i0 += 64) {
    const int ilast = std::min(i0 + 63, L1 - 1);
    const int jlo = full ? 0 : std::max(i0 - newk, 0);
    const int jhi = full ? L2 - 1 : std::min(ilast + Dd + newk, L2 - 1);
    const int nsteps = (jhi - jlo + 1) + (ilast - i0);
    tot += (int64_t)((nsteps + 15) >> 4) * 256;
  }
  return tot;
}

void batch_free_device(poyhip_batch* b) {
  for (void* p : {(void*)b->d_tab, (void*)b->d_res, (void*)b->d_dm, (void*)b->d_ws,
                  (void*)b->d_stab, (void*)b->d_counters, (void*)b->d_cls})
    ctx_release(b->ctx, p);
  b->d_cls = nullptr;
  b->d_tab = nullptr;
  b->d_res = nullptr;
  b->d_dm = nullptr;
  b->d_ws = nullptr;
  b->d_stab = nullptr;
  b->d_counters = nullptr;
  for (hipEvent_t e : b->ev) (void)hipEventDestroy(e);
  b->ev.clear();
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  b->gexec = nullptr;
}

PairTable batch_table(const poyhip_batch* b, const uint8_t* blob, const int64_t* off1,
                      const int64_t* off2, int chunk) {
  PairTable t;
  t.blob = blob;
  t.off1 = off1;
  t.off2 = off2;
  t.len_a = b->d_la;
  t.len_b = b->d_lb;
  t.swapped = b->d_sw;
  t.prio = b->d_sw;
  t.deltawh = b->d_dw;
  t.dm_off = b->d_dm_off;
  t.stab_off = b->d_stab_off;
  t.out_off = b->d_out_off;
  const auto& ch = b->chunks[chunk < 0 ? 0 : chunk];
  t.order = b->d_order + (chunk < 0 ? 0 : ch.begin);
  t.n = chunk < 0 ? b->n : ch.count;
  t.npairs = b->n;
  return t;
}

// Enqueues every launch of one run on the context stream.
int batch_enqueue(poyhip_batch* b, const uint8_t* blob, const int64_t* off1, const int64_t* off2,
                  int* d_cost, uint8_t* d_out, int* d_outlen, bool timing) {
  poyhip_ctx* c = b->ctx;
  hipStream_t st = c->stream;
  const int nch = (int)b->chunks.size();
  if (timing) HIPCHK(c, hipEventRecord(b->ev[2 * nch], st));
  // counters: two per chunk (fast and general fill queues), then the number
  // of general-class pairs
  hipLaunchKernelGGL(k_zero_ints, dim3((2 * nch + 1 + 63) / 64), dim3(64), 0, st, b->d_counters,
                     2 * nch + 1);
  if (b->op == POYHIP_OP_ALIGN_AFFINE) {
    PairTable all = batch_table(b, blob, off1, off2, -1);
    hipLaunchKernelGGL(k_affine_classify, dim3(b->n), dim3(64), 0, st, all, b->d_cls,
                       b->d_counters + 2 * nch);
  }
  if (op_is_linear(b->op)) {
    PairTable all = batch_table(b, blob, off1, off2, -1);
    hipLaunchKernelGGL(k_linear_deltawh, dim3((b->n + kDwPairsPerBlock - 1) / kDwPairsPerBlock),
                       dim3(64 * kDwPairsPerBlock), 0, st, all, b->d_dw,
                       b->has_deltaw, b->d_dwin);
  }
  for (int k = 0; k < nch; k++) {
    PairTable t = batch_table(b, blob, off1, off2, k);
    if (timing) HIPCHK(c, hipEventRecord(b->ev[2 * k], st));
    switch (b->op) {
      case POYHIP_OP_ALIGN_AFFINE:
      case POYHIP_OP_COST_AFFINE: {
        if (b->op == POYHIP_OP_COST_AFFINE) {
          AffCostArgs A;
          A.pt = t;
          A.cm = c->d_cm;
          A.ws = b->d_ws;
          A.ws_stride = b->ws_stride;
          A.res = b->d_res;
          A.counter = b->d_counters + k;
          hipLaunchKernelGGL(k_affine_cost_full, dim3(b->grid), dim3(64), 0, st, A);
        } else {
          // gap-free pairs on the fast kernel, the others on the general one
          AffFillArgs A;
          A.rotating = c->aff_rot ? 1 : 0;
          A.pt = t;
          A.cm = c->d_cm;
          A.dm = b->d_dm;
          A.ws = b->d_ws;
          A.ws_stride = b->ws_stride;
          A.stab = b->d_stab;
          A.res = b->d_res;
          A.cls = b->d_cls;
          A.ngen = b->d_counters + 2 * nch;
          A.out = d_out;  // each pair's backtrace follows its fill
          A.outlen = d_outlen;
          A.counter = b->d_counters + 2 * k;
          hipLaunchKernelGGL((k_affine_fill<true, true>), dim3(b->grid), dim3(64), 0, st, A);
          A.counter = b->d_counters + 2 * k + 1;
          hipLaunchKernelGGL((k_affine_fill<true, false>), dim3(b->grid), dim3(64), 0, st, A);
        }
        break;
      }
      case POYHIP_OP_NEWKK_ALIGN:
      case POYHIP_OP_NEWKK_COST: {
        NkFillArgs A;
        A.pt = t;
        A.cm = c->d_cm;
        A.dm = b->d_dm;
        A.ws = b->d_ws;
        A.ws_stride = b->ws_stride;
        A.stab = b->d_stab;
        A.res = b->d_res;
        A.counter = b->d_counters + k;
        A.phase = 0;
        // each pair's walk, ancestor and median follow its fill
        A.cm_med = c->d_cm;
        A.out = b->op == POYHIP_OP_NEWKK_ALIGN ? d_out : nullptr;
        A.outlen = d_outlen;
        A.medians = 1;
        HIPCHK(c, launch_newkk_fill(A, b->op == POYHIP_OP_NEWKK_ALIGN, b->wide, b->grid, st));
        break;
      }
      default: {
        LinFillArgs A;
        A.pt = t;
        A.cm = c->d_cm;
        A.dm = b->d_dm;
        A.ws = b->d_ws;
        A.ws_stride = b->ws_stride;
        A.stab = b->d_stab;
        A.res = b->d_res;
        A.counter = b->d_counters + k;
        A.phase = 0;
        HIPCHK(c, launch_linear_fill(c, A, b->op == POYHIP_OP_ALIGN_2D, b->grid, st));
      }
    }
    HIPCHK(c, hipGetLastError());
    if (timing) HIPCHK(c, hipEventRecord(b->ev[2 * k + 1], st));
    if (b->op == POYHIP_OP_ALIGN_AFFINE) {
      // the general-class pairs (none in most batches: the kernel then ends
      // at once); the gap-free kernel ran the other backtraces
      AffBtArgs B;
      B.pt = t;
      B.cm = c->d_cm;
      B.dm = b->d_dm;
      B.stab = b->d_stab;
      B.res = b->d_res;
      B.out = d_out;
      B.outlen = d_outlen;
      B.cls = b->d_cls;
      B.ngen = b->d_counters + 2 * nch;
      hipLaunchKernelGGL(k_affine_backtrace, dim3(t.n), dim3(64), 0, st, B);
    } else if (b->op == POYHIP_OP_ALIGN_2D) {
      LinBtArgs B;
      B.pt = t;
      B.cm = c->d_cm;
      B.dm = b->d_dm;
      B.stab = b->d_stab;
      B.res = b->d_res;
      B.out = d_out;
      B.outlen = d_outlen;
      hipLaunchKernelGGL(k_linear_backtrace, dim3(t.n), dim3(64), 0, st, B);
    }
    HIPCHK(c, hipGetLastError());
  }
  hipLaunchKernelGGL(k_extract_cost, dim3((b->n + 255) / 256), dim3(256), 0, st, b->d_res, d_cost,
                     b->n);
  HIPCHK(c, hipGetLastError());
  if (timing) HIPCHK(c, hipEventRecord(b->ev[2 * nch + 1], st));
  return POYHIP_OK;
}

}  // namespace

extern "C" {

int poyhip_batch_create(poyhip_ctx* c, int op, int n, const int* len1, const int* len2,
                        const int* deltaw, size_t dm_budget, poyhip_batch** out) {
  if (!c || !out || n <= 0 || !len1 || !len2) return POYHIP_EINVAL;
  if (!(op >= POYHIP_OP_ALIGN_AFFINE && op <= POYHIP_OP_COST_2D) && !op_is_newkk(op))
    return set_err(c, POYHIP_EINVAL, "unknown batch op %d", op);
  if (!c->has_cm) return set_err(c, POYHIP_ESTATE, "no cost matrix installed");
  (void)hipSetDevice(c->device);
  *out = nullptr;
  poyhip_batch* b = new poyhip_batch();
  b->ctx = c;
  b->op = op;
  b->n = n;
  b->la.resize(n);
  b->lb.resize(n);
  b->swapped.resize(n);
  b->out_off.resize(n);
  b->dm_off.resize(n);
  b->stab_off.resize(n);
  int maxlb = 1;
  std::vector<double> work(n);
  for (int k = 0; k < n; k++) {
    if (len1[k] < 1 || len2[k] < 1) {
      delete b;
      return set_err(c, POYHIP_EINVAL, "pair %d has an empty sequence", k);
    }
    const bool sw = len1[k] > len2[k];
    b->swapped[k] = sw ? 1 : 0;
    b->la[k] = sw ? len2[k] : len1[k];
    b->lb[k] = sw ? len1[k] : len2[k];
    maxlb = std::max(maxlb, b->lb[k]);
    work[k] = (double)b->la[k] * (double)b->lb[k];
    if (op_is_newkk(op)) {
      const int rc = nk_check(c, b->la[k], b->lb[k]);
      if (rc) {
        delete b;
        return rc;
      }
    }
  }
  b->wide = maxlb >= 65536 || c->nk_wide_always;
  // largest first (LPT order for the persistent kernels)
  b->order.resize(n);
  std::iota(b->order.begin(), b->order.end(), 0);
  std::stable_sort(b->order.begin(), b->order.end(),
                   [&](int x, int y) { return work[x] > work[y]; });
  // output block
  int64_t o = 0;
  for (int k = 0; k < n; k++) {
    b->out_off[k] = o;
    o += (pair_out_bytes(op, b->la[k], b->lb[k]) + 15) & ~(int64_t)15;
  }
  b->out_bytes = (size_t)o;
  // stripe tables
  int64_t so = 0;
  for (int k = 0; k < n; k++) {
    b->stab_off[k] = so;
    so += n_stripes(b->la[k]) + 1;
  }
  // chunks by direction-store budget
  if (dm_budget == 0) dm_budget = auto_dm_budget();
  b->dm_budget = (int64_t)dm_budget;
  if (op == POYHIP_OP_ALIGN_2D || op == POYHIP_OP_NEWKK_ALIGN) {
    int64_t est = 0;
    for (int k = 0; k < n; k++) est += (pair_dm_bytes(op, b->la[k], b->lb[k]) + 255) & ~(int64_t)255;
    // direction stores sized for the full width would need several chunks
    // (few pairs in flight): plan per run from the pairs' final bands
    b->two_phase = est > (int64_t)dm_budget;
  }
  int64_t cur = 0;
  int begin = 0;
  for (int q = 0; q < n && !b->two_phase; q++) {
    const int k = b->order[q];
    const int64_t sz = (pair_dm_bytes(op, b->la[k], b->lb[k]) + 255) & ~(int64_t)255;
    if (cur > 0 && cur + sz > (int64_t)dm_budget) {
      b->chunks.push_back({begin, q - begin, cur});
      begin = q;
      cur = 0;
    }
    b->dm_off[k] = cur;
    cur += sz;
  }
  if (b->two_phase) {
    b->chunks.push_back({0, n, 0});
    b->dm_cap = b->dm_budget;
  } else {
    b->chunks.push_back({begin, n - begin, cur});
  }
  for (const auto& ch : b->chunks) b->dm_cap = std::max(b->dm_cap, ch.dm_bytes);
  // persistent grid: waves resident across the device
  b->grid = std::min(n, c->num_cus * c->waves_per_cu);
  if (b->grid < 1) b->grid = 1;
  b->ws_stride = (ws_ints(op, maxlb) + 63) & ~63LL;
  // device allocations
  const size_t tab_bytes = (size_t)n * (3 * 8 + 6 * 4) + 1024;
  int rc = POYHIP_OK;
  auto fail = [&](int code, const char* what) {
    batch_free_device(b);
    delete b;
    return set_err(c, code, "batch allocation failed: %s", what);
  };
  if (ctx_malloc(c, (void**)&b->d_tab, tab_bytes) != hipSuccess) return fail(POYHIP_ENOMEM, "tables");
  char* p = static_cast<char*>(b->d_tab);
  b->d_dm_off = reinterpret_cast<int64_t*>(p); p += 8 * (size_t)n;
  b->d_stab_off = reinterpret_cast<int64_t*>(p); p += 8 * (size_t)n;
  b->d_out_off = reinterpret_cast<int64_t*>(p); p += 8 * (size_t)n;
  b->d_la = reinterpret_cast<int*>(p); p += 4 * (size_t)n;
  b->d_lb = reinterpret_cast<int*>(p); p += 4 * (size_t)n;
  b->d_sw = reinterpret_cast<int*>(p); p += 4 * (size_t)n;
  b->d_dw = reinterpret_cast<int*>(p); p += 4 * (size_t)n;
  b->d_dwin = reinterpret_cast<int*>(p); p += 4 * (size_t)n;
  b->d_order = reinterpret_cast<int*>(p); p += 4 * (size_t)n;
  if (ctx_malloc(c, (void**)&b->d_res, sizeof(PairResult) * (size_t)n) != hipSuccess)
    return fail(POYHIP_ENOMEM, "results");
  const size_t ncounters = b->two_phase ? 2 * (size_t)n + 4 : 2 * b->chunks.size() + 2;
  if (ctx_malloc(c, (void**)&b->d_counters, sizeof(int) * ncounters) != hipSuccess)
    return fail(POYHIP_ENOMEM, "counters");
  if (ctx_malloc(c, (void**)&b->d_cls, (size_t)n + 64) != hipSuccess) return fail(POYHIP_ENOMEM, "classes");
  if (op_has_dm(op)) {
    if (ctx_malloc(c, (void**)&b->d_dm, (size_t)b->dm_cap + 256) != hipSuccess)
      return fail(POYHIP_ENOMEM, "direction store");
    if (ctx_malloc(c, (void**)&b->d_stab, sizeof(long long) * (size_t)(so + 1)) != hipSuccess)
      return fail(POYHIP_ENOMEM, "stripe tables");
  }
  if (ctx_malloc(c, (void**)&b->d_ws, sizeof(int) * (size_t)b->ws_stride * b->grid) != hipSuccess)
    return fail(POYHIP_ENOMEM, "workspace");
  std::vector<int> dwin(n, 0);
  if (deltaw) {
    b->has_deltaw = 1;
    for (int k = 0; k < n; k++) dwin[k] = deltaw[k];
  }
  hipStream_t st = c->stream;
  bool ok = true;
  ok &= hipMemcpyAsync(b->d_dm_off, b->dm_off.data(), 8 * (size_t)n, hipMemcpyHostToDevice, st) == hipSuccess;
  ok &= hipMemcpyAsync(b->d_stab_off, b->stab_off.data(), 8 * (size_t)n, hipMemcpyHostToDevice, st) == hipSuccess;
  ok &= hipMemcpyAsync(b->d_out_off, b->out_off.data(), 8 * (size_t)n, hipMemcpyHostToDevice, st) == hipSuccess;
  ok &= hipMemcpyAsync(b->d_la, b->la.data(), 4 * (size_t)n, hipMemcpyHostToDevice, st) == hipSuccess;
  ok &= hipMemcpyAsync(b->d_lb, b->lb.data(), 4 * (size_t)n, hipMemcpyHostToDevice, st) == hipSuccess;
  ok &= hipMemcpyAsync(b->d_sw, b->swapped.data(), 4 * (size_t)n, hipMemcpyHostToDevice, st) == hipSuccess;
  ok &= hipMemcpyAsync(b->d_dwin, dwin.data(), 4 * (size_t)n, hipMemcpyHostToDevice, st) == hipSuccess;
  ok &= hipMemcpyAsync(b->d_order, b->order.data(), 4 * (size_t)n, hipMemcpyHostToDevice, st) == hipSuccess;
  ok &= hipStreamSynchronize(st) == hipSuccess;
  if (!ok) return fail(POYHIP_EDEVICE, "upload");
  b->ev.resize(2 * b->chunks.size() + 2);
  for (auto& e : b->ev)
    if (hipEventCreate(&e) != hipSuccess) return fail(POYHIP_EDEVICE, "events");
  (void)rc;
  *out = b;
  return POYHIP_OK;
}

int poyhip_batch_destroy(poyhip_batch* b) {
  if (!b) return POYHIP_EINVAL;
  (void)hipSetDevice(b->ctx->device);
  (void)hipStreamSynchronize(b->ctx->stream);
  batch_free_device(b);
  delete b;
  return POYHIP_OK;
}

size_t poyhip_batch_out_bytes(const poyhip_batch* b) { return b ? b->out_bytes : 0; }

int64_t poyhip_batch_out_offset(const poyhip_batch* b, int k) {
  if (!b || k < 0 || k >= b->n) return -1;
  return b->out_off[k];
}

int poyhip_batch_out_cap(const poyhip_batch* b, int k) {
  if (!b || k < 0 || k >= b->n) return -1;
  if (b->op == POYHIP_OP_ALIGN_AFFINE) return b->la[k] + b->lb[k] + 2;
  if (b->op == POYHIP_OP_ALIGN_2D) return (int)lin_out_cap(b->la[k], b->lb[k]);
  if (b->op == POYHIP_OP_NEWKK_ALIGN) return (int)nk_out_cap(b->la[k], b->lb[k]);
  return 0;
}

int poyhip_batch_out_layout(const poyhip_batch* b, int64_t* off, int* cap) {
  if (!b || !off || !cap) return POYHIP_EINVAL;
  for (int k = 0; k < b->n; k++) {
    off[k] = poyhip_batch_out_offset(b, k);
    cap[k] = poyhip_batch_out_cap(b, k);
  }
  return POYHIP_OK;
}

int poyhip_set_tuning(poyhip_ctx* c, int knob, int value) {
  if (!c) return POYHIP_EINVAL;
  switch (knob) {
    case POYHIP_TUNE_WAVES_PER_CU:
      if (value < 1 || value > 32) return set_err(c, POYHIP_EINVAL, "waves per CU %d", value);
      c->waves_per_cu = value;
      return POYHIP_OK;
    case POYHIP_TUNE_AFFINE_ROTATING:
      c->aff_rot = value != 0;
      return POYHIP_OK;
    case POYHIP_TUNE_LINEAR_WORKGROUP:
      if (value < 0 || value > 2) return set_err(c, POYHIP_EINVAL, "linear workgroup %d", value);
      c->lin_wg = value;
      c->lin_mw = c->lin_fits && c->lin_wg != 0;
      return POYHIP_OK;
    case POYHIP_TUNE_CUBE_WORKGROUP:
      c->cube_mw = value != 0;
      return POYHIP_OK;
    default:
      return set_err(c, POYHIP_EINVAL, "unknown tuning knob %d", knob);
  }
}

int poyhip_get_tuning(poyhip_ctx* c, int knob) {
  if (!c) return POYHIP_EINVAL;
  switch (knob) {
    case POYHIP_TUNE_WAVES_PER_CU: return c->waves_per_cu;
    case POYHIP_TUNE_AFFINE_ROTATING: return c->aff_rot ? 1 : 0;
    case POYHIP_TUNE_LINEAR_WORKGROUP: return c->lin_wg;
    case POYHIP_TUNE_CUBE_WORKGROUP: return c->cube_mw ? 1 : 0;
    default: return POYHIP_EINVAL;
  }
}

int poyhip_batch_use_graph(poyhip_batch* b, int enable) {
  if (!b) return POYHIP_EINVAL;
  b->use_graph = enable != 0;
  if (!b->use_graph && b->gexec) {
    (void)hipGraphExecDestroy(b->gexec);
    b->gexec = nullptr;
  }
  return POYHIP_OK;
}

}  // extern "C"

namespace {

// Two-phase run of a linear align batch: phase 1 runs every doubling round
// of every pair without directions (all pairs in flight, no direction
// memory), phase 2 repeats each pair's final round with directions, in
// chunks planned from the final bands.  Synchronous.
int batch_run_two_phase(poyhip_batch* b, const uint8_t* d_blob, const int64_t* d_off1,
                        const int64_t* d_off2, int* d_cost, uint8_t* d_out, int* d_outlen) {
  poyhip_ctx* c = b->ctx;
  hipStream_t st = c->stream;
  const int n = b->n;
  std::vector<hipEvent_t> ev(4);
  for (auto& e : ev) HIPCHK(c, hipEventCreate(&e));
  struct EvGuard {
    std::vector<hipEvent_t>* v;
    ~EvGuard() { for (auto e : *v) (void)hipEventDestroy(e); }
  } guard{&ev};
  HIPCHK(c, hipEventRecord(ev[0], st));
  hipLaunchKernelGGL(k_zero_ints, dim3((2 * n + 4 + 255) / 256), dim3(256), 0, st, b->d_counters,
                     2 * n + 4);
  PairTable all = batch_table(b, d_blob, d_off1, d_off2, -1);
  const bool nk = b->op == POYHIP_OP_NEWKK_ALIGN;
  if (!nk)
    hipLaunchKernelGGL(k_linear_deltawh, dim3((n + kDwPairsPerBlock - 1) / kDwPairsPerBlock),
                       dim3(64 * kDwPairsPerBlock), 0, st, all, b->d_dw,
                       b->has_deltaw, b->d_dwin);
  LinFillArgs A;
  A.pt = all;
  A.cm = c->d_cm;
  A.dm = b->d_dm;
  A.ws = b->d_ws;
  A.ws_stride = b->ws_stride;
  A.stab = b->d_stab;
  A.res = b->d_res;
  A.counter = b->d_counters;
  A.phase = 1;
  NkFillArgs N;
  N.pt = all;
  N.cm = c->d_cm;
  N.dm = b->d_dm;
  N.ws = b->d_ws;
  N.ws_stride = b->ws_stride;
  N.stab = b->d_stab;
  N.res = b->d_res;
  N.counter = b->d_counters;
  N.phase = 1;
  N.cm_med = c->d_cm;
  N.out = d_out;
  N.outlen = d_outlen;
  N.medians = 1;
  HIPCHK(c, hipEventRecord(ev[1], st));
  if (nk) HIPCHK(c, launch_newkk_fill(N, false, b->wide, b->grid, st));
  else HIPCHK(c, launch_linear_fill(c, A, false, b->grid, st));
  HIPCHK(c, hipEventRecord(ev[2], st));
  b->h_res.resize(n);
  HIPCHK(c, hipMemcpyAsync(b->h_res.data(), b->d_res, sizeof(PairResult) * (size_t)n,
                           hipMemcpyDeviceToHost, st));
  HIPCHK(c, hipStreamSynchronize(st));
  float ms1 = 0.f;
  (void)hipEventElapsedTime(&ms1, ev[1], ev[2]);
  // chunks of the final round, largest pairs first
  b->chunks.clear();
  int64_t cur = 0;
  int begin = 0;
  for (int q = 0; q < n; q++) {
    const int k = b->order[q];
    const PairResult& r = b->h_res[k];
    const int64_t exact =
        nk ? (r.full == 2 ? 0 : nk_dm_exact(b->la[k], b->lb[k], r.newk))
           : lin_dm_exact(b->la[k], b->lb[k], r.newk < 0 ? 0 : r.newk, r.full != 0);
    const int64_t sz = (exact + 255) & ~(int64_t)255;
    if (sz > b->dm_budget)
      return set_err(c, POYHIP_ECAP, "pair %d needs %lld direction bytes", k, (long long)sz);
    if (cur > 0 && cur + sz > b->dm_budget) {
      b->chunks.push_back({begin, q - begin, cur});
      begin = q;
      cur = 0;
    }
    b->dm_off[k] = cur;
    cur += sz;
  }
  b->chunks.push_back({begin, n - begin, cur});
  HIPCHK(c, hipMemcpyAsync(b->d_dm_off, b->dm_off.data(), 8 * (size_t)n, hipMemcpyHostToDevice, st));
  float ms2 = 0.f;
  for (size_t k = 0; k < b->chunks.size(); k++) {
    PairTable t = batch_table(b, d_blob, d_off1, d_off2, (int)k);
    A.pt = t;
    A.counter = b->d_counters + 1 + k;
    A.phase = 2;
    N.pt = t;
    N.counter = b->d_counters + 1 + k;
    N.phase = 2;
    HIPCHK(c, hipEventRecord(ev[1], st));
    if (nk) HIPCHK(c, launch_newkk_fill(N, true, b->wide, b->grid, st));
    else HIPCHK(c, launch_linear_fill(c, A, true, b->grid, st));
    HIPCHK(c, hipEventRecord(ev[2], st));
    if (!nk) {
      LinBtArgs B;
      B.pt = t;
      B.cm = c->d_cm;
      B.dm = b->d_dm;
      B.stab = b->d_stab;
      B.res = b->d_res;
      B.out = d_out;
      B.outlen = d_outlen;
      hipLaunchKernelGGL(k_linear_backtrace, dim3(t.n), dim3(64), 0, st, B);
    }
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipEventSynchronize(ev[2]));
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, ev[1], ev[2]);
    ms2 += ms;
  }
  hipLaunchKernelGGL(k_extract_cost, dim3((n + 255) / 256), dim3(256), 0, st, b->d_res, d_cost, n);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipEventRecord(ev[3], st));
  HIPCHK(c, hipStreamSynchronize(st));
  b->fill_ms = ms1 + ms2;
  float all_ms = 0.f;
  if (hipEventElapsedTime(&all_ms, ev[0], ev[3]) == hipSuccess) b->total_ms = all_ms;
  return POYHIP_OK;
}

}  // namespace

extern "C" {

int poyhip_batch_run(poyhip_batch* b, const uint8_t* d_blob, const int64_t* d_off1,
                     const int64_t* d_off2, int* d_cost, uint8_t* d_out, int* d_outlen, int sync) {
  if (!b || !d_blob || !d_off1 || !d_off2 || !d_cost) return POYHIP_EINVAL;
  poyhip_ctx* c = b->ctx;
  if (op_has_dm(b->op) && (!d_out || !d_outlen))
    return set_err(c, POYHIP_EINVAL, "align batches need output buffers");
  (void)hipSetDevice(c->device);
  hipStream_t st = c->stream;
  if (b->two_phase) return batch_run_two_phase(b, d_blob, d_off1, d_off2, d_cost, d_out, d_outlen);
  if (b->use_graph) {
    const void* key[6] = {d_blob, d_off1, d_off2, d_cost, d_out, d_outlen};
    if (b->gexec && memcmp(key, b->g_key, sizeof key) != 0) {
      (void)hipGraphExecDestroy(b->gexec);
      b->gexec = nullptr;
    }
    if (!b->gexec) {
      hipGraph_t g;
      HIPCHK(c, hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      int rc = batch_enqueue(b, d_blob, d_off1, d_off2, d_cost, d_out, d_outlen, false);
      hipError_t e = hipStreamEndCapture(st, &g);
      if (rc) return rc;
      if (e != hipSuccess) return set_err(c, POYHIP_EDEVICE, "graph capture: %s", hipGetErrorString(e));
      e = hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (e != hipSuccess) {
        b->gexec = nullptr;
        return set_err(c, POYHIP_EDEVICE, "graph instantiate: %s", hipGetErrorString(e));
      }
      memcpy(b->g_key, key, sizeof key);
    }
    HIPCHK(c, hipGraphLaunch(b->gexec, st));
    b->fill_ms = -1.f;
  } else {
    int rc = batch_enqueue(b, d_blob, d_off1, d_off2, d_cost, d_out, d_outlen, true);
    if (rc) return rc;
  }
  if (sync) {
    HIPCHK(c, hipStreamSynchronize(st));
    if (!b->use_graph) {
      float tot = 0.f;
      for (size_t k = 0; k < b->chunks.size(); k++) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, b->ev[2 * k], b->ev[2 * k + 1]) == hipSuccess) tot += ms;
      }
      b->fill_ms = tot;
      float all = 0.f;
      const size_t nch = b->chunks.size();
      if (hipEventElapsedTime(&all, b->ev[2 * nch], b->ev[2 * nch + 1]) == hipSuccess)
        b->total_ms = all;
    }
  }
  return POYHIP_OK;
}

int poyhip_batch_stats(poyhip_batch* b, int64_t* cells, int64_t* slots, int* rounds, int* newk) {
  if (!b) return POYHIP_EINVAL;
  poyhip_ctx* c = b->ctx;
  (void)hipSetDevice(c->device);
  std::vector<PairResult> r(b->n);
  HIPCHK(c, hipMemcpyAsync(r.data(), b->d_res, sizeof(PairResult) * (size_t)b->n,
                           hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (int k = 0; k < b->n; k++) {
    if (cells) cells[k] = r[k].cells;
    if (slots) slots[k] = r[k].slots;
    if (rounds) rounds[k] = r[k].rounds;
    if (newk) newk[k] = r[k].newk;
  }
  return POYHIP_OK;
}

float poyhip_batch_last_fill_ms(const poyhip_batch* b) { return b ? b->fill_ms : -1.f; }
float poyhip_batch_last_total_ms(const poyhip_batch* b) { return b ? b->total_ms : -1.f; }
int poyhip_batch_chunks(const poyhip_batch* b) { return b ? (int)b->chunks.size() : POYHIP_EINVAL; }

}  // extern "C"

// ---------------------------------------------------------------------------
// Three sequences (algn_nw_3d, backtrack_3d, algn_get_median_3d)
// ---------------------------------------------------------------------------

namespace {

struct OneDesc3 {
  int64_t off1, off2, off3, dm_off, out_off;
  int l1, l2, l3, order, counter, pad;
};
static_assert(sizeof(OneDesc3) <= kDescBytes, "triple descriptor block");

TripleTable triple_table_one(void* dstage) {
  char* base = static_cast<char*>(dstage);
  OneDesc3* d = reinterpret_cast<OneDesc3*>(base);
  TripleTable t;
  t.blob = reinterpret_cast<const uint8_t*>(base + kDescBytes);
  t.off1 = &d->off1;
  t.off2 = &d->off2;
  t.off3 = &d->off3;
  t.len1 = &d->l1;
  t.len2 = &d->l2;
  t.len3 = &d->l3;
  t.dm_off = &d->dm_off;
  t.out_off = &d->out_off;
  t.order = &d->order;
  t.n = 1;
  t.ntriples = 1;
  return t;
}

// One triple on the context: fill (directions kept in c->dm) and, when
// asked, the backtrack.  outblk = [cost, 4 lengths | outputs].
int run_cube_one(poyhip_ctx* c, const uint8_t* s1, int l1, const uint8_t* s2, int l2,
                 const uint8_t* s3, int l3, bool backtrack, int* cost, int outlen[4],
                 std::vector<uint8_t>* outs) {
  const size_t need = kDescBytes + (size_t)l1 + l2 + l3;
  int rc = grow(c, c->stage, need);
  if (rc) return rc;
  const int64_t dmb = cube_dm_bytes(l1, l2, l3);
  const int64_t outb = cube_out_bytes(l1, l2, l3);
  if ((rc = grow(c, c->dm, dmb + 256))) return rc;
  const int64_t ring = std::max(cube_ring_ints(l1, l2, l3), cube_mw_ring_ints(l1, l2, l3, kCubeMwWaves));
  if ((rc = grow(c, c->ws, ring * sizeof(int)))) return rc;
  if ((rc = grow(c, c->outblk, kResBytes + outb))) return rc;
  c->h_stage.resize(need);
  OneDesc3 d;
  memset(&d, 0, sizeof d);
  d.off1 = 0;
  d.off2 = l1;
  d.off3 = (int64_t)l1 + l2;
  d.l1 = l1;
  d.l2 = l2;
  d.l3 = l3;
  memcpy(c->h_stage.data(), &d, sizeof d);
  memcpy(c->h_stage.data() + kDescBytes, s1, l1);
  memcpy(c->h_stage.data() + kDescBytes + l1, s2, l2);
  memcpy(c->h_stage.data() + kDescBytes + l1 + l2, s3, l3);
  HIPCHK(c, hipMemcpyAsync(c->stage.p, c->h_stage.data(), need, hipMemcpyHostToDevice, c->stream));
  TripleTable t = triple_table_one(c->stage.p);
  int* dcost = static_cast<int*>(c->outblk.p);
  int* doutlen = dcost + 4;
  uint8_t* dout = static_cast<uint8_t*>(c->outblk.p) + kResBytes;
  CubeFillArgs A;
  A.tt = t;
  A.cm3 = c->d_cm3;
  A.dm = static_cast<uint8_t*>(c->dm.p);
  A.ring = static_cast<int*>(c->ws.p);
  A.ring_stride = 0;
  A.cost = dcost;
  A.counter = &reinterpret_cast<OneDesc3*>(c->stage.p)->counter;
  const int64_t lsum = (int64_t)l1 + l2 + l3;
  const bool mw = cube_nseg(l3) <= kCubeMwMaxSeg && cube_mw_ok(c, lsum);
  if (mw && cube_mw_u16(lsum, c->cm3_max))
    hipLaunchKernelGGL((k_cube_mw<true, true, kCubeMwWaves>), dim3(1), dim3(64 * kCubeMwWaves), 0,
                       c->stream, A);
  else if (mw)
    hipLaunchKernelGGL((k_cube_mw<true, false, kCubeMwWaves>), dim3(1), dim3(64 * kCubeMwWaves),
                       0, c->stream, A);
  else if (cube_nseg(l3) == 1)
    hipLaunchKernelGGL((k_cube_fill<true, true>), dim3(1), dim3(64), 0, c->stream, A);
  else
    hipLaunchKernelGGL((k_cube_fill<true, false>), dim3(1), dim3(64), 0, c->stream, A);
  HIPCHK(c, hipGetLastError());
  size_t back = 4 * sizeof(int);
  if (backtrack) {
    CubeBtArgs B;
    B.tt = t;
    B.cm3 = c->d_cm3;
    B.dm = A.dm;
    B.out = dout;
    B.outlen = doutlen;
    hipLaunchKernelGGL(k_cube_backtrack, dim3(1), dim3(64), 0, c->stream, B);
    HIPCHK(c, hipGetLastError());
    back = kResBytes + outb;
  }
  c->h_out.resize(back);
  HIPCHK(c, hipMemcpyAsync(c->h_out.data(), c->outblk.p, back, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  memcpy(cost, c->h_out.data(), sizeof(int));
  if (backtrack) {
    memcpy(outlen, c->h_out.data() + 4 * sizeof(int), 4 * sizeof(int));
    outs->assign(c->h_out.begin() + kResBytes, c->h_out.begin() + kResBytes + outb);
  }
  return POYHIP_OK;
}

bool valid_triple(const uint8_t* s1, int l1, const uint8_t* s2, int l2, const uint8_t* s3,
                  int l3) {
  return valid_seq(s1, l1) && valid_seq(s2, l2) && valid_seq(s3, l3);
}

int copy_triple_outputs(poyhip_ctx* c, int l1, int l2, int l3, const int outlen[4],
                        const std::vector<uint8_t>& blk, uint8_t* r1, uint8_t* r2, uint8_t* r3,
                        int cap, int* len) {
  if (outlen[0] < 0) return set_err(c, POYHIP_EDEVICE, "3D backtrack left the cube");
  if (outlen[0] > cap) return set_err(c, POYHIP_ECAP, "capacity %d < %d", cap, outlen[0]);
  const int cp = (int)cube_out_cap(l1, l2, l3);
  const int n = outlen[0];
  memcpy(r1, blk.data() + cp - n, n);
  memcpy(r2, blk.data() + 2 * cp - n, n);
  memcpy(r3, blk.data() + 3 * cp - n, n);
  *len = n;
  return POYHIP_OK;
}

}  // namespace

extern "C" {

int poyhip_set_cm3(poyhip_ctx* c, const poyhip_cm3* cm3) {
  if (!c || !cm3) return POYHIP_EINVAL;
  (void)hipSetDevice(c->device);
  if (!c->d_cm3 && hipMalloc(&c->d_cm3, sizeof(Cm3Dev)) != hipSuccess) {
    c->d_cm3 = nullptr;
    return set_err(c, POYHIP_ENOMEM, "hipMalloc of the 3D tables failed");
  }
  static_assert(sizeof(Cm3Dev) == sizeof(poyhip_cm3), "3D table layout");
  HIPCHK(c, hipMemcpyAsync(c->d_cm3, cm3, sizeof(Cm3Dev), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->has_cm3 = true;
  c->last3d = false;
  int mx = 0;
  for (int k = 0; k < 32768 && mx >= 0; k++) mx = cm3->cost[k] < 0 ? -1 : std::max(mx, cm3->cost[k]);
  c->cm3_max = mx;
  return POYHIP_OK;
}

int poyhip_cost_3(poyhip_ctx* c, const uint8_t* s1, int l1, const uint8_t* s2, int l2,
                  const uint8_t* s3, int l3, int uk, int* cost) {
  (void)uk;  // algn_fill_cube ignores its Ukkonen barrier argument
  if (!c || !cost || !valid_triple(s1, l1, s2, l2, s3, l3)) return POYHIP_EINVAL;
  if (!c->has_cm3) return set_err(c, POYHIP_ESTATE, "no 3D cost matrix installed");
  (void)hipSetDevice(c->device);
  int outlen[4];
  int rc = run_cube_one(c, s1, l1, s2, l2, s3, l3, false, cost, outlen, nullptr);
  if (rc) return rc;
  c->last3d = true;
  c->last3[0].assign(s1, s1 + l1);
  c->last3[1].assign(s2, s2 + l2);
  c->last3[2].assign(s3, s3 + l3);
  return POYHIP_OK;
}

int poyhip_backtrack_3d(poyhip_ctx* c, const uint8_t* s1, int l1, const uint8_t* s2, int l2,
                        const uint8_t* s3, int l3, uint8_t* r1, uint8_t* r2, uint8_t* r3, int cap,
                        int* len) {
  if (!c || !r1 || !r2 || !r3 || !len || !valid_triple(s1, l1, s2, l2, s3, l3))
    return POYHIP_EINVAL;
  const uint8_t* s[3] = {s1, s2, s3};
  const int l[3] = {l1, l2, l3};
  bool same = c->last3d;
  for (int u = 0; u < 3 && same; u++)
    same = (int)c->last3[u].size() == l[u] && !memcmp(c->last3[u].data(), s[u], l[u]);
  if (!same) return set_err(c, POYHIP_ESTATE, "backtrack without a matching poyhip_cost_3");
  (void)hipSetDevice(c->device);
  int cost, outlen[4];
  std::vector<uint8_t> blk;
  int rc = run_cube_one(c, s1, l1, s2, l2, s3, l3, true, &cost, outlen, &blk);
  if (rc) return rc;
  return copy_triple_outputs(c, l1, l2, l3, outlen, blk, r1, r2, r3, cap, len);
}

int poyhip_align_3d(poyhip_ctx* c, const uint8_t* s1, int l1, const uint8_t* s2, int l2,
                    const uint8_t* s3, int l3, int uk, uint8_t* r1, uint8_t* r2, uint8_t* r3,
                    int cap, int* len, int* cost) {
  (void)uk;
  if (!c || !cost || !r1 || !r2 || !r3 || !len || !valid_triple(s1, l1, s2, l2, s3, l3))
    return POYHIP_EINVAL;
  if (!c->has_cm3) return set_err(c, POYHIP_ESTATE, "no 3D cost matrix installed");
  (void)hipSetDevice(c->device);
  int outlen[4];
  std::vector<uint8_t> blk;
  int rc = run_cube_one(c, s1, l1, s2, l2, s3, l3, true, cost, outlen, &blk);
  if (rc) return rc;
  c->last3d = true;
  c->last3[0].assign(s1, s1 + l1);
  c->last3[1].assign(s2, s2 + l2);
  c->last3[2].assign(s3, s3 + l3);
  return copy_triple_outputs(c, l1, l2, l3, outlen, blk, r1, r2, r3, cap, len);
}

int poyhip_median_3(poyhip_ctx* c, const uint8_t* a, const uint8_t* b, const uint8_t* s3,
                    int len, uint8_t* out, int cap, int* out_len) {
  if (!c || !a || !b || !s3 || !out || !out_len || len < 0) return POYHIP_EINVAL;
  if (!c->has_cm3) return set_err(c, POYHIP_ESTATE, "no 3D cost matrix installed");
  if (cap < len) return set_err(c, POYHIP_ECAP, "capacity %d < %d", cap, len);
  (void)hipSetDevice(c->device);
  *out_len = len;
  if (len == 0) return POYHIP_OK;
  const size_t need = 3 * (size_t)len;
  int rc = grow(c, c->stage, need);
  if (rc) return rc;
  if ((rc = grow(c, c->outblk, (size_t)len))) return rc;
  c->h_stage.resize(need);
  memcpy(c->h_stage.data(), a, len);
  memcpy(c->h_stage.data() + len, b, len);
  memcpy(c->h_stage.data() + 2 * (size_t)len, s3, len);
  HIPCHK(c, hipMemcpyAsync(c->stage.p, c->h_stage.data(), need, hipMemcpyHostToDevice, c->stream));
  const uint8_t* d = static_cast<const uint8_t*>(c->stage.p);
  hipLaunchKernelGGL(k_median_3, dim3(1), dim3(256), 0, c->stream, c->d_cm3, d, d + len,
                     d + 2 * (size_t)len, len, static_cast<uint8_t*>(c->outblk.p));
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipMemcpyAsync(out, c->outblk.p, len, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return POYHIP_OK;
}

}  // extern "C"

// Batches of triples

struct poyhip_batch3 {
  poyhip_ctx* ctx = nullptr;
  int op = 0;
  int n = 0;
  std::vector<int> l1, l2, l3, order;
  std::vector<int64_t> dm_off, out_off;
  struct Chunk {
    int begin, count;
    int64_t dm_bytes;
  };
  std::vector<Chunk> chunks;
  void* d_tab = nullptr;
  int64_t *d_dm_off = nullptr, *d_out_off = nullptr;
  int *d_l1 = nullptr, *d_l2 = nullptr, *d_l3 = nullptr, *d_order = nullptr;
  int* d_counters = nullptr;
  uint8_t* d_dm = nullptr;
  int* d_ring = nullptr;
  int64_t dm_cap = 0;
  long long ring_stride = 0;
  int grid = 0;
  int64_t len_sum_mw = 0;  // largest L1 + L2 + L3 among triples of <= kCubeMwMaxSeg segments
  size_t out_bytes = 0;
  std::vector<hipEvent_t> ev;
  float fill_ms = -1.f, total_ms = -1.f;
};

namespace {

void batch3_free(poyhip_batch3* b) {
  for (void* p : {(void*)b->d_tab, (void*)b->d_counters, (void*)b->d_dm, (void*)b->d_ring})
    ctx_release(b->ctx, p);
  b->d_tab = nullptr;
  b->d_counters = nullptr;
  b->d_dm = nullptr;
  b->d_ring = nullptr;
  for (hipEvent_t e : b->ev) (void)hipEventDestroy(e);
  b->ev.clear();
}

}  // namespace

extern "C" {

int poyhip_batch3_create(poyhip_ctx* c, int op, int n, const int* len1, const int* len2,
                         const int* len3, size_t dm_budget, poyhip_batch3** out) {
  if (!c || !out || n <= 0 || !len1 || !len2 || !len3) return POYHIP_EINVAL;
  if (op != POYHIP_OP_ALIGN_3D && op != POYHIP_OP_COST_3D)
    return set_err(c, POYHIP_EINVAL, "unknown triple batch op %d", op);
  if (!c->has_cm3) return set_err(c, POYHIP_ESTATE, "no 3D cost matrix installed");
  (void)hipSetDevice(c->device);
  *out = nullptr;
  poyhip_batch3* b = new poyhip_batch3();
  b->ctx = c;
  b->op = op;
  b->n = n;
  b->l1.assign(len1, len1 + n);
  b->l2.assign(len2, len2 + n);
  b->l3.assign(len3, len3 + n);
  b->dm_off.resize(n);
  b->out_off.resize(n);
  std::vector<double> work(n);
  for (int k = 0; k < n; k++) {
    if (len1[k] < 1 || len2[k] < 1 || len3[k] < 1) {
      delete b;
      return set_err(c, POYHIP_EINVAL, "triple %d has an empty sequence", k);
    }
    // fill time follows the row count (a row is one wavefront's pass over
    // its segments, whatever L3)
    work[k] = (double)len1[k] * len2[k] * cube_nseg(len3[k]);
    if (cube_nseg(len3[k]) <= kCubeMwMaxSeg)
      b->len_sum_mw = std::max<int64_t>(b->len_sum_mw, (int64_t)len1[k] + len2[k] + len3[k]);
    b->ring_stride = std::max<long long>(b->ring_stride, cube_ring_ints(len1[k], len2[k], len3[k]));
    b->ring_stride = std::max<long long>(
        b->ring_stride, cube_mw_ring_ints(len1[k], len2[k], len3[k], kCubeMwWaves));
  }
  b->ring_stride = (b->ring_stride + 63) & ~63LL;
  b->order.resize(n);
  std::iota(b->order.begin(), b->order.end(), 0);
  std::stable_sort(b->order.begin(), b->order.end(),
                   [&](int x, int y) { return work[x] > work[y]; });
  int64_t o = 0;
  for (int k = 0; k < n; k++) {
    b->out_off[k] = o;
    if (op == POYHIP_OP_ALIGN_3D) o += (cube_out_bytes(len1[k], len2[k], len3[k]) + 15) & ~(int64_t)15;
  }
  b->out_bytes = (size_t)o;
  size_t fr = 0, tot = 0;
  if (hipMemGetInfo(&fr, &tot) != hipSuccess) fr = (size_t)8 << 30;
  // resident wavefronts: up to 8 per CU, with the rings within a quarter of
  // the free memory
  long long grid = std::min<long long>(n, (long long)c->num_cus * 8);
  const long long ring_cap = (long long)(fr / 4) / (b->ring_stride * 4);
  if (grid > ring_cap) grid = ring_cap;
  if (grid < 1) {
    delete b;
    return set_err(c, POYHIP_ENOMEM, "ring of %lld ints does not fit", b->ring_stride);
  }
  b->grid = (int)grid;
  if (op == POYHIP_OP_ALIGN_3D) {
    // default: three quarters of the memory the rings leave free
    const size_t rings = sizeof(int) * (size_t)b->ring_stride * b->grid;
    if (dm_budget == 0) dm_budget = fr > rings ? (fr - rings) / 4 * 3 : 0;
    // A chunk takes no more work than its resident workgroups finish in the
    // time of its first (largest) triple, so that its work-sorted triples
    // run side by side without a tail.
    const double slots = c->cube_mw ? c->num_cus * 2 : b->grid;
    int64_t cur = 0;
    int begin = 0;
    double wsum = 0.0;
    for (int q = 0; q < n; q++) {
      const int k = b->order[q];
      const int64_t sz = (cube_dm_bytes(len1[k], len2[k], len3[k]) + 255) & ~(int64_t)255;
      if (sz > (int64_t)dm_budget) {
        delete b;
        return set_err(c, POYHIP_ECAP, "triple %d needs %lld direction bytes", k, (long long)sz);
      }
      if (cur > 0 &&
          (cur + sz > (int64_t)dm_budget || wsum + work[k] > slots * work[b->order[begin]])) {
        wsum = 0.0;
        b->chunks.push_back({begin, q - begin, cur});
        begin = q;
        cur = 0;
      }
      b->dm_off[k] = cur;
      cur += sz;
      wsum += work[k];
    }
    b->chunks.push_back({begin, n - begin, cur});
    for (const auto& ch : b->chunks) b->dm_cap = std::max(b->dm_cap, ch.dm_bytes);
  } else {
    b->chunks.push_back({0, n, 0});
  }
  auto fail = [&](int code, const char* what) {
    batch3_free(b);
    delete b;
    return set_err(c, code, "triple batch allocation failed: %s", what);
  };
  const size_t tab_bytes = (size_t)n * (2 * 8 + 4 * 4) + 1024;
  if (ctx_malloc(c, (void**)&b->d_tab, tab_bytes) != hipSuccess) return fail(POYHIP_ENOMEM, "tables");
  char* p = static_cast<char*>(b->d_tab);
  b->d_dm_off = reinterpret_cast<int64_t*>(p); p += 8 * (size_t)n;
  b->d_out_off = reinterpret_cast<int64_t*>(p); p += 8 * (size_t)n;
  b->d_l1 = reinterpret_cast<int*>(p); p += 4 * (size_t)n;
  b->d_l2 = reinterpret_cast<int*>(p); p += 4 * (size_t)n;
  b->d_l3 = reinterpret_cast<int*>(p); p += 4 * (size_t)n;
  b->d_order = reinterpret_cast<int*>(p); p += 4 * (size_t)n;
  if (ctx_malloc(c, (void**)&b->d_counters, sizeof(int) * (2 * b->chunks.size() + 2)) != hipSuccess)
    return fail(POYHIP_ENOMEM, "counters");
  if (op == POYHIP_OP_ALIGN_3D && ctx_malloc(c, (void**)&b->d_dm, (size_t)b->dm_cap + 256) != hipSuccess)
    return fail(POYHIP_ENOMEM, "direction store");
  if (ctx_malloc(c, (void**)&b->d_ring, sizeof(int) * (size_t)b->ring_stride * b->grid) != hipSuccess)
    return fail(POYHIP_ENOMEM, "rings");
  hipStream_t st = c->stream;
  bool ok = true;
  ok &= hipMemcpyAsync(b->d_dm_off, b->dm_off.data(), 8 * (size_t)n, hipMemcpyHostToDevice, st) == hipSuccess;
  ok &= hipMemcpyAsync(b->d_out_off, b->out_off.data(), 8 * (size_t)n, hipMemcpyHostToDevice, st) == hipSuccess;
  ok &= hipMemcpyAsync(b->d_l1, b->l1.data(), 4 * (size_t)n, hipMemcpyHostToDevice, st) == hipSuccess;
  ok &= hipMemcpyAsync(b->d_l2, b->l2.data(), 4 * (size_t)n, hipMemcpyHostToDevice, st) == hipSuccess;
  ok &= hipMemcpyAsync(b->d_l3, b->l3.data(), 4 * (size_t)n, hipMemcpyHostToDevice, st) == hipSuccess;
  ok &= hipMemcpyAsync(b->d_order, b->order.data(), 4 * (size_t)n, hipMemcpyHostToDevice, st) == hipSuccess;
  ok &= hipStreamSynchronize(st) == hipSuccess;
  if (!ok) return fail(POYHIP_EDEVICE, "upload");
  b->ev.resize(2 * b->chunks.size() + 2);
  for (auto& e : b->ev)
    if (hipEventCreate(&e) != hipSuccess) return fail(POYHIP_EDEVICE, "events");
  *out = b;
  return POYHIP_OK;
}

int poyhip_batch3_destroy(poyhip_batch3* b) {
  if (!b) return POYHIP_EINVAL;
  (void)hipSetDevice(b->ctx->device);
  (void)hipStreamSynchronize(b->ctx->stream);
  batch3_free(b);
  delete b;
  return POYHIP_OK;
}

size_t poyhip_batch3_out_bytes(const poyhip_batch3* b) { return b ? b->out_bytes : 0; }

int64_t poyhip_batch3_out_offset(const poyhip_batch3* b, int k) {
  if (!b || k < 0 || k >= b->n) return -1;
  return b->out_off[k];
}

int poyhip_batch3_out_cap(const poyhip_batch3* b, int k) {
  if (!b || k < 0 || k >= b->n || b->op != POYHIP_OP_ALIGN_3D) return -1;
  return (int)cube_out_cap(b->l1[k], b->l2[k], b->l3[k]);
}

int poyhip_batch3_chunks(const poyhip_batch3* b) { return b ? (int)b->chunks.size() : POYHIP_EINVAL; }

float poyhip_batch3_last_fill_ms(const poyhip_batch3* b) { return b ? b->fill_ms : -1.f; }

int poyhip_batch3_run(poyhip_batch3* b, const uint8_t* d_blob, const int64_t* d_off1,
                      const int64_t* d_off2, const int64_t* d_off3, int* d_cost, uint8_t* d_out,
                      int* d_outlen, int sync) {
  if (!b || !d_blob || !d_off1 || !d_off2 || !d_off3 || !d_cost) return POYHIP_EINVAL;
  poyhip_ctx* c = b->ctx;
  if (b->op == POYHIP_OP_ALIGN_3D && (!d_out || !d_outlen))
    return set_err(c, POYHIP_EINVAL, "align batches need output buffers");
  (void)hipSetDevice(c->device);
  hipStream_t st = c->stream;
  const int nch = (int)b->chunks.size();
  HIPCHK(c, hipEventRecord(b->ev[2 * nch], st));
  hipLaunchKernelGGL(k_zero_ints, dim3((2 * nch + 63) / 64), dim3(64), 0, st, b->d_counters, 2 * nch);
  for (int k = 0; k < nch; k++) {
    const auto& ch = b->chunks[k];
    TripleTable t;
    t.blob = d_blob;
    t.off1 = d_off1;
    t.off2 = d_off2;
    t.off3 = d_off3;
    t.len1 = b->d_l1;
    t.len2 = b->d_l2;
    t.len3 = b->d_l3;
    t.dm_off = b->d_dm_off;
    t.out_off = b->d_out_off;
    t.order = b->d_order + ch.begin;
    t.n = ch.count;
    t.ntriples = b->n;
    CubeFillArgs A;
    A.tt = t;
    A.cm3 = c->d_cm3;
    A.dm = b->d_dm;
    A.ring = b->d_ring;
    A.ring_stride = b->ring_stride;
    A.cost = d_cost;
    A.counter = b->d_counters + k;
    const int g = std::min(b->grid, ch.count);
    HIPCHK(c, hipEventRecord(b->ev[2 * k], st));
    // two launches (counters 2k and 2k+1): triples of up to kCubeMwMaxSeg
    // segments on workgroups of kCubeMwWaves wavefronts (two per CU), the
    // longer ones on single wavefronts; when the keys of the workgroup fill
    // could wrap, one-segment and longer triples on single wavefronts
    A.counter = b->d_counters + 2 * k;
    const bool dm = b->op == POYHIP_OP_ALIGN_3D;
    const bool mw = cube_mw_ok(c, b->len_sum_mw);
    if (mw) {
      const int gm = std::min(g, c->num_cus * 2);
      const dim3 blk(64 * kCubeMwWaves);
      A.seg_lo = 1;
      A.seg_hi = kCubeMwMaxSeg;
      if (cube_mw_u16(b->len_sum_mw, c->cm3_max)) {
        if (dm)
          hipLaunchKernelGGL((k_cube_mw<true, true, kCubeMwWaves>), dim3(gm), blk, 0, st, A);
        else
          hipLaunchKernelGGL((k_cube_mw<false, true, kCubeMwWaves>), dim3(gm), blk, 0, st, A);
      } else if (dm) {
        hipLaunchKernelGGL((k_cube_mw<true, false, kCubeMwWaves>), dim3(gm), blk, 0, st, A);
      } else {
        hipLaunchKernelGGL((k_cube_mw<false, false, kCubeMwWaves>), dim3(gm), blk, 0, st, A);
      }
    } else if (dm) {
      hipLaunchKernelGGL((k_cube_fill<true, true>), dim3(g), dim3(64), 0, st, A);
    } else {
      hipLaunchKernelGGL((k_cube_fill<false, true>), dim3(g), dim3(64), 0, st, A);
    }
    A.counter = b->d_counters + 2 * k + 1;
    A.seg_lo = mw ? kCubeMwMaxSeg + 1 : 2;
    A.seg_hi = 1 << 30;
    if (b->op == POYHIP_OP_ALIGN_3D)
      hipLaunchKernelGGL((k_cube_fill<true, false>), dim3(g), dim3(64), 0, st, A);
    else
      hipLaunchKernelGGL((k_cube_fill<false, false>), dim3(g), dim3(64), 0, st, A);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipEventRecord(b->ev[2 * k + 1], st));
    if (b->op == POYHIP_OP_ALIGN_3D) {
      CubeBtArgs B;
      B.tt = t;
      B.cm3 = c->d_cm3;
      B.dm = b->d_dm;
      B.out = d_out;
      B.outlen = d_outlen;
      hipLaunchKernelGGL(k_cube_backtrack, dim3((ch.count + 63) / 64), dim3(64), 0, st, B);
      HIPCHK(c, hipGetLastError());
    }
  }
  HIPCHK(c, hipEventRecord(b->ev[2 * nch + 1], st));
  if (sync) {
    HIPCHK(c, hipStreamSynchronize(st));
    float tot = 0.f;
    for (int k = 0; k < nch; k++) {
      float ms = 0.f;
      if (hipEventElapsedTime(&ms, b->ev[2 * k], b->ev[2 * k + 1]) == hipSuccess) tot += ms;
    }
    b->fill_ms = tot;
    float all = 0.f;
    if (hipEventElapsedTime(&all, b->ev[2 * nch], b->ev[2 * nch + 1]) == hipSuccess)
      b->total_ms = all;
  }
  return POYHIP_OK;
}

}  // extern "C"

#include "serve.hip"
