// Continuous-batching alignment server (no reference counterpart).
//
// A tree search emits alignment requests one at a time (SPR/TBR rejoins:
// cost_2 of a clade median against every edge's median; joins: align_2 plus
// the ancestor median).  The server turns that stream into batches:
//
//  * sequences live in a device-resident pool (append-only arena with a
//    pinned host mirror for uploads); medians that a request asks to keep
//    are written into the pool by the GPU and never travel to the host, so a
//    parent median feeds later requests directly;
//  * a host worker thread owns S slots, each with its own HIP stream and
//    preallocated buffers; whenever a slot is free and requests are pending
//    it admits requests in arrival order while their direction stores fit
//    the slot's budget (memory-based admission) and launches them as one
//    persistent-kernel batch, so batches of different slots overlap on the
//    device while new requests keep arriving;
//  * a finished slot hands its results to the requests (cost, aligned
//    sequences, pool id of the kept median) and is refilled at once.
//
// Every request runs exactly the kernels of the batch entry points, so its
// result does not depend on which other requests share its batch.
#pragma once

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>

namespace {

using Clock = std::chrono::steady_clock;

double us_since(Clock::time_point t0, Clock::time_point t1) {
  return std::chrono::duration<double, std::micro>(t1 - t0).count();
}

// Copies the ancestor median (region 2 of the linear output block) of the
// requests that keep it into the sequence pool.
__global__ void k_pool_store(const uint8_t* out, const int64_t* out_off, const int* cap,
                             const int* outlen, const int* pid_of, const int64_t* pool_off,
                             uint8_t* pool, int nkeep) {
  const int q = blockIdx.x;
  if (q >= nkeep) return;
  const int pid = pid_of[q];
  const int n = outlen[4 * pid + 2];
  const int c = cap[q];
  const uint8_t* src = out + out_off[pid] + 2 * (int64_t)c + (c - n);
  uint8_t* dst = pool + pool_off[q];
  for (int k = threadIdx.x; k < n; k += blockDim.x) dst[k] = src[k];
}

struct SrvRequest {
  int64_t id = 0;
  int op = 0;
  int64_t a = 0, b = 0;  // pool ids
  int deltaw = -1;       // -1: Align.cost_2's ?deltaw = None
  int flags = 0;
  // planning
  int la = 0, lb = 0;
  int64_t dm_bytes = 0;
  // results
  int cost = 0;
  std::vector<uint8_t> r1, r2;
  int64_t median_id = -1;
  int status = 0;  // 0 pending, 1 running, 2 done, <0 error code
  Clock::time_point t_submit, t_start, t_done;
};

struct PoolEntry {
  int64_t off;
  int len;
};

struct SrvSlot {
  hipStream_t st = nullptr;
  hipEvent_t ev_start = nullptr, ev_done = nullptr;
  bool busy = false;
  std::vector<SrvRequest*> reqs;
  // device buffers (sized at server creation)
  void* d_tab = nullptr;
  PairResult* d_res = nullptr;
  uint8_t* d_dm = nullptr;
  int* d_ws = nullptr;
  long long* d_stab = nullptr;
  uint8_t* d_out = nullptr;
  int* d_outlen = nullptr;
  int* d_cost = nullptr;
  int* d_counters = nullptr;
  uint8_t* d_cls = nullptr;
  // pinned host staging: tables up, results down
  uint8_t* h_up = nullptr;
  uint8_t* h_down = nullptr;
  size_t up_bytes = 0, down_bytes = 0;
  // per-launch layout
  int n = 0, nkeep = 0;
  int64_t out_bytes = 0;
  size_t off_cost = 0, off_outlen = 0, off_out = 0;
  std::vector<int> caps;
};

}  // namespace

struct poyhip_server {
  poyhip_ctx* ctx = nullptr;
  poyhip_server_opts opt;
  // pool
  uint8_t* d_pool = nullptr;
  uint8_t* h_pool = nullptr;  // pinned mirror of the host-provided part
  int64_t pool_used = 0;
  std::vector<PoolEntry> entries;
  // host-provided pool ranges not yet on the device, and the event after
  // the last upload (other streams wait for it before reading the pool)
  std::vector<std::pair<int64_t, int64_t>> uploads;
  hipEvent_t ev_upload = nullptr;
  bool upload_pending_event = false;
  // slots and queue
  std::vector<SrvSlot> slots;
  std::mutex m;
  std::condition_variable cv_work, cv_done;
  std::deque<SrvRequest*> pending;
  std::unordered_map<int64_t, SrvRequest*> reqs;
  int64_t next_id = 1;
  bool stop = false;
  int error = 0;
  // set by poyhip_server_fail_next_launch: the next admission fails as a
  // launch error would, without touching the device
  bool fail_next_launch = false;
  std::thread worker;
  // statistics
  int64_t batches = 0, pairs_done = 0;
  double cells_full = 0;
  int max_batch_seen = 0;
  int max_inflight = 0;
  Clock::time_point t_first;
  bool have_first = false;
  // slot limits
  int64_t slot_dm = 0, slot_out = 0;
  int slot_pairs = 0, grid = 0, max_len = 0;
  long long ws_stride = 0;
};

namespace {

int64_t srv_out_bytes(int op, int la, int lb) { return (pair_out_bytes(op, la, lb) + 15) & ~(int64_t)15; }

// Slot table layout in device memory (structure of arrays, n <= slot_pairs).
struct SrvTab {
  int64_t *off1, *off2, *dm_off, *stab_off, *out_off, *keep_pool;
  int *la, *lb, *sw, *dw, *dwin, *order, *keep_pid, *keep_cap;
};

SrvTab srv_tab(void* base, int cap) {
  char* p = static_cast<char*>(base);
  SrvTab t;
  auto take64 = [&](int64_t*& x) { x = reinterpret_cast<int64_t*>(p); p += 8 * (size_t)cap; };
  auto take32 = [&](int*& x) { x = reinterpret_cast<int*>(p); p += 4 * (size_t)cap; };
  take64(t.off1); take64(t.off2); take64(t.dm_off); take64(t.stab_off); take64(t.out_off);
  take64(t.keep_pool);
  take32(t.la); take32(t.lb); take32(t.sw); take32(t.dw); take32(t.dwin); take32(t.order);
  take32(t.keep_pid); take32(t.keep_cap);
  return t;
}
constexpr size_t kSrvTabPer = 6 * 8 + 8 * 4;

// Marks the requests of a slot as failed with `code`; a median a failed
// request was to keep is marked as never coming (len -2), so submissions that
// wait for it return instead of blocking.  Called with the lock held.
void srv_fail(poyhip_server* s, SrvSlot& sl, int code) {
  s->error = code;
  for (auto* r : sl.reqs) {
    r->status = code;
    if (r->median_id >= 0) s->entries[r->median_id].len = -2;
  }
  sl.reqs.clear();
  sl.busy = false;
}

// Finishes the requests of a completed slot.  Called with the lock held.
void srv_finish(poyhip_server* s, SrvSlot& sl) {
  const auto now = Clock::now();
  const int* cost = reinterpret_cast<const int*>(sl.h_down + sl.off_cost);
  const int* outlen = reinterpret_cast<const int*>(sl.h_down + sl.off_outlen);
  const uint8_t* out = sl.h_down + sl.off_out;
  int64_t o = 0;
  for (int k = 0; k < sl.n; k++) {
    SrvRequest* r = sl.reqs[k];
    r->cost = cost[k];
    const int64_t ob = srv_out_bytes(r->op, r->la, r->lb);
    if (op_has_dm(r->op)) {
      const int cap = (int)(ob > 0 ? pair_out_bytes(r->op, r->la, r->lb) / 4 : 0);
      const uint8_t* blk = out + o;
      const int n0 = outlen[4 * k], n1 = outlen[4 * k + 1];
      r->r1.assign(blk + cap - n0, blk + cap);
      r->r2.assign(blk + 2 * cap - n1, blk + 2 * cap);
      if (r->flags & POYHIP_REQ_KEEP_MEDIAN) {
        const int nm = outlen[4 * k + 2];
        s->entries[r->median_id].len = nm;
      }
    }
    o += ob;
    r->status = 2;
    r->t_done = now;
    s->pairs_done++;
    s->cells_full += (double)r->la * r->lb;
  }
  sl.reqs.clear();
  sl.busy = false;
  s->batches++;
}

// Plans and launches the admitted requests on slot `sl`.  Lock held on
// entry; the upload and the launches only touch this slot's buffers.
int srv_launch(poyhip_server* s, SrvSlot& sl) {
  poyhip_ctx* c = s->ctx;
  const int n = (int)sl.reqs.size();
  sl.n = n;
  const int op = sl.reqs[0]->op;
  SrvTab t = srv_tab(sl.h_up, s->slot_pairs);
  std::vector<double> work(n);
  int64_t dmo = 0, so = 0, oo = 0;
  int nkeep = 0, maxlb = 1;
  for (int k = 0; k < n; k++) {
    SrvRequest* r = sl.reqs[k];
    const PoolEntry& ea = s->entries[r->a];
    const PoolEntry& eb = s->entries[r->b];
    const bool swp = ea.len > eb.len;
    t.off1[k] = ea.off;
    t.off2[k] = eb.off;
    t.la[k] = swp ? eb.len : ea.len;
    t.lb[k] = swp ? ea.len : eb.len;
    t.sw[k] = swp ? 1 : 0;
    t.dwin[k] = r->deltaw;
    t.dw[k] = 0;
    t.dm_off[k] = dmo;
    dmo += (pair_dm_bytes(op, t.la[k], t.lb[k]) + 255) & ~(int64_t)255;
    t.stab_off[k] = so;
    so += n_stripes(t.la[k]) + 1;
    t.out_off[k] = oo;
    oo += srv_out_bytes(op, t.la[k], t.lb[k]);
    maxlb = std::max(maxlb, t.lb[k]);
    work[k] = (double)t.la[k] * t.lb[k];
    if (op == POYHIP_OP_ALIGN_2D && (r->flags & POYHIP_REQ_KEEP_MEDIAN)) {
      t.keep_pid[nkeep] = k;
      t.keep_cap[nkeep] = (int)lin_out_cap(t.la[k], t.lb[k]);
      t.keep_pool[nkeep] = s->entries[r->median_id].off;
      nkeep++;
    }
  }
  std::vector<int> order(n);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return work[x] > work[y]; });
  memcpy(t.order, order.data(), 4 * (size_t)n);
  sl.nkeep = nkeep;
  sl.out_bytes = oo;
  // results coming back: costs, lengths, outputs
  sl.off_cost = 0;
  sl.off_outlen = ((size_t)4 * n + 255) & ~(size_t)255;
  sl.off_out = sl.off_outlen + (((size_t)16 * n + 255) & ~(size_t)255);
  const size_t down = sl.off_out + (op_has_dm(op) ? (size_t)oo : 0);
  hipStream_t st = sl.st;
  HIPCHK(c, hipEventRecord(sl.ev_start, st));
  // earlier uploads may still be in flight on another slot's stream
  if (s->upload_pending_event) HIPCHK(c, hipStreamWaitEvent(st, s->ev_upload, 0));
  // sequences the host put into the pool since the last launch (ranges that
  // the GPU writes, kept medians, are never uploaded)
  if (!s->uploads.empty()) {
    for (const auto& u : s->uploads)
      HIPCHK(c, hipMemcpyAsync(s->d_pool + u.first, s->h_pool + u.first, (size_t)(u.second - u.first),
                               hipMemcpyHostToDevice, st));
    s->uploads.clear();
    HIPCHK(c, hipEventRecord(s->ev_upload, st));
    s->upload_pending_event = true;
  }
  HIPCHK(c, hipMemcpyAsync(sl.d_tab, sl.h_up, kSrvTabPer * (size_t)s->slot_pairs,
                           hipMemcpyHostToDevice, st));
  SrvTab dt = srv_tab(sl.d_tab, s->slot_pairs);
  PairTable pt;
  pt.blob = s->d_pool;
  pt.off1 = dt.off1;
  pt.off2 = dt.off2;
  pt.len_a = dt.la;
  pt.len_b = dt.lb;
  pt.swapped = dt.sw;
  pt.prio = dt.sw;
  pt.deltawh = dt.dw;
  pt.dm_off = dt.dm_off;
  pt.stab_off = dt.stab_off;
  pt.out_off = dt.out_off;
  pt.order = dt.order;
  pt.n = n;
  pt.npairs = n;
  const int grid = std::min(n, s->grid);
  const long long wss = (ws_ints(op, maxlb) + 63) & ~63LL;
  hipLaunchKernelGGL(k_zero_ints, dim3(1), dim3(64), 0, st, sl.d_counters, 2);
  if (op == POYHIP_OP_ALIGN_AFFINE)
    hipLaunchKernelGGL(k_affine_classify, dim3(n), dim3(64), 0, st, pt, sl.d_cls, nullptr);
  if (!op_is_affine(op))
    hipLaunchKernelGGL(k_linear_deltawh, dim3((n + kDwPairsPerBlock - 1) / kDwPairsPerBlock),
                       dim3(64 * kDwPairsPerBlock), 0, st, pt, dt.dw, 2,
                       dt.dwin);
  if (op == POYHIP_OP_ALIGN_AFFINE) {
    AffFillArgs A;
    A.rotating = c->aff_rot ? 1 : 0;
    A.pt = pt;
    A.cm = c->d_cm;
    A.dm = sl.d_dm;
    A.ws = sl.d_ws;
    A.ws_stride = wss;
    A.stab = sl.d_stab;
    A.res = sl.d_res;
    A.cls = sl.d_cls;
    A.ngen = nullptr;
    A.counter = sl.d_counters;
    // gap-free pairs: fill and backtrace in one kernel; the backtrace kernel
    // takes the pairs with gap-bit symbols
    A.out = sl.d_out;
    A.outlen = sl.d_outlen;
    hipLaunchKernelGGL((k_affine_fill<true, true>), dim3(grid), dim3(64), 0, st, A);
    A.counter = sl.d_counters + 1;
    hipLaunchKernelGGL((k_affine_fill<true, false>), dim3(grid), dim3(64), 0, st, A);
    AffBtArgs B;
    B.pt = pt;
    B.cm = c->d_cm;
    B.dm = sl.d_dm;
    B.stab = sl.d_stab;
    B.res = sl.d_res;
    B.out = sl.d_out;
    B.outlen = sl.d_outlen;
    B.cls = sl.d_cls;
    hipLaunchKernelGGL(k_affine_backtrace, dim3(n), dim3(64), 0, st, B);
  } else if (op == POYHIP_OP_COST_AFFINE) {
    AffCostArgs A;
    A.pt = pt;
    A.cm = c->d_cm;
    A.ws = sl.d_ws;
    A.ws_stride = wss;
    A.res = sl.d_res;
    A.counter = sl.d_counters;
    hipLaunchKernelGGL(k_affine_cost_full, dim3(grid), dim3(64), 0, st, A);
  } else {
    LinFillArgs A;
    A.pt = pt;
    A.cm = c->d_cm;
    A.dm = sl.d_dm;
    A.ws = sl.d_ws;
    A.ws_stride = wss;
    A.stab = sl.d_stab;
    A.res = sl.d_res;
    A.counter = sl.d_counters;
    A.phase = 0;
    if (op == POYHIP_OP_ALIGN_2D) {
      HIPCHK(c, launch_linear_fill(c, A, true, grid, st));
      LinBtArgs B;
      B.pt = pt;
      B.cm = c->d_cm;
      B.dm = sl.d_dm;
      B.stab = sl.d_stab;
      B.res = sl.d_res;
      B.out = sl.d_out;
      B.outlen = sl.d_outlen;
      hipLaunchKernelGGL(k_linear_backtrace, dim3(n), dim3(64), 0, st, B);
      if (nkeep > 0)
        hipLaunchKernelGGL(k_pool_store, dim3(nkeep), dim3(64), 0, st, sl.d_out, dt.out_off,
                           dt.keep_cap, sl.d_outlen, dt.keep_pid, dt.keep_pool, s->d_pool, nkeep);
    } else {
      HIPCHK(c, launch_linear_fill(c, A, false, grid, st));
    }
  }
  hipLaunchKernelGGL(k_extract_cost, dim3((n + 255) / 256), dim3(256), 0, st, sl.d_res, sl.d_cost,
                     n);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipMemcpyAsync(sl.h_down + sl.off_cost, sl.d_cost, 4 * (size_t)n,
                           hipMemcpyDeviceToHost, st));
  if (op_has_dm(op)) {
    HIPCHK(c, hipMemcpyAsync(sl.h_down + sl.off_outlen, sl.d_outlen, 16 * (size_t)n,
                             hipMemcpyDeviceToHost, st));
    HIPCHK(c, hipMemcpyAsync(sl.h_down + sl.off_out, sl.d_out, (size_t)oo, hipMemcpyDeviceToHost,
                             st));
  }
  (void)down;
  HIPCHK(c, hipEventRecord(sl.ev_done, st));
  sl.busy = true;
  return POYHIP_OK;
}

void srv_loop(poyhip_server* s) {
  (void)hipSetDevice(s->ctx->device);
  std::unique_lock<std::mutex> lk(s->m);
  while (!s->stop) {
    bool progress = false;
    int inflight = 0;
    // completions
    for (auto& sl : s->slots) {
      if (!sl.busy) continue;
      const hipError_t e = hipEventQuery(sl.ev_done);
      if (e == hipSuccess) {
        srv_finish(s, sl);
        progress = true;
      } else if (e != hipErrorNotReady) {
        srv_fail(s, sl, POYHIP_EDEVICE);
        progress = true;
      } else {
        inflight++;
      }
    }
    if (progress) s->cv_done.notify_all();
    // admissions: fill every free slot from the queue, oldest first, while
    // the direction stores and outputs fit the slot
    for (auto& sl : s->slots) {
      if (sl.busy || s->pending.empty()) continue;
      const int op = s->pending.front()->op;
      int64_t dm = 0, out = 0;
      while (!s->pending.empty() && (int)sl.reqs.size() < s->slot_pairs) {
        SrvRequest* r = s->pending.front();
        if (r->op != op) break;
        const int64_t ob = srv_out_bytes(op, r->la, r->lb);
        if (!sl.reqs.empty() && (dm + r->dm_bytes > s->slot_dm || out + ob > s->slot_out)) break;
        dm += r->dm_bytes;
        out += ob;
        r->status = 1;
        r->t_start = Clock::now();
        sl.reqs.push_back(r);
        s->pending.pop_front();
      }
      s->max_batch_seen = std::max(s->max_batch_seen, (int)sl.reqs.size());
      const bool inject = s->fail_next_launch;
      s->fail_next_launch = false;
      if (inject || srv_launch(s, sl) != POYHIP_OK) {
        srv_fail(s, sl, POYHIP_EDEVICE);
        s->cv_done.notify_all();
      } else {
        inflight++;
      }
      progress = true;
    }
    s->max_inflight = std::max(s->max_inflight, inflight);
    if (!progress) {
      if (inflight == 0)
        s->cv_work.wait_for(lk, std::chrono::milliseconds(5));
      else {
        lk.unlock();
        std::this_thread::sleep_for(std::chrono::microseconds(20));
        lk.lock();
      }
    }
  }
}

void srv_free(poyhip_server* s) {
  for (auto& sl : s->slots) {
    if (sl.st) (void)hipStreamSynchronize(sl.st);
    for (void* p : {(void*)sl.d_tab, (void*)sl.d_res, (void*)sl.d_dm, (void*)sl.d_ws,
                    (void*)sl.d_stab, (void*)sl.d_out, (void*)sl.d_outlen, (void*)sl.d_cost,
                    (void*)sl.d_counters, (void*)sl.d_cls})
      if (p) (void)hipFree(p);
    if (sl.h_up) (void)hipHostFree(sl.h_up);
    if (sl.h_down) (void)hipHostFree(sl.h_down);
    if (sl.ev_start) (void)hipEventDestroy(sl.ev_start);
    if (sl.ev_done) (void)hipEventDestroy(sl.ev_done);
    if (sl.st) (void)hipStreamDestroy(sl.st);
  }
  if (s->ev_upload) (void)hipEventDestroy(s->ev_upload);
  if (s->d_pool) (void)hipFree(s->d_pool);
  if (s->h_pool) (void)hipHostFree(s->h_pool);
  for (auto& kv : s->reqs) delete kv.second;
}

}  // namespace

extern "C" {

int poyhip_server_create(poyhip_ctx* c, const poyhip_server_opts* o, poyhip_server** out) {
  if (!c || !out) return POYHIP_EINVAL;
  if (!c->has_cm) return set_err(c, POYHIP_ESTATE, "no cost matrix installed");
  (void)hipSetDevice(c->device);
  *out = nullptr;
  poyhip_server* s = new poyhip_server();
  s->ctx = c;
  poyhip_server_opts d;
  d.n_slots = 3;
  d.max_batch_pairs = 4096;
  d.max_len = 16384;
  d.slot_dm_bytes = (size_t)4 << 30;
  d.pool_bytes = (size_t)256 << 20;
  if (o) {
    if (o->n_slots > 0) d.n_slots = o->n_slots;
    if (o->max_batch_pairs > 0) d.max_batch_pairs = o->max_batch_pairs;
    if (o->max_len > 0) d.max_len = o->max_len;
    if (o->slot_dm_bytes > 0) d.slot_dm_bytes = o->slot_dm_bytes;
    if (o->pool_bytes > 0) d.pool_bytes = o->pool_bytes;
  }
  if (d.n_slots > 8) d.n_slots = 8;
  s->opt = d;
  s->slot_pairs = d.max_batch_pairs;
  s->max_len = d.max_len;
  s->slot_dm = (int64_t)d.slot_dm_bytes;
  // outputs of a full slot: capped like the direction stores
  s->slot_out = std::max<int64_t>((int64_t)d.slot_dm_bytes / 8, aff_out_bytes(d.max_len, d.max_len) * 4);
  s->grid = std::min(d.max_batch_pairs, c->num_cus * 8);
  s->ws_stride = (std::max(aff_ws_ints(d.max_len), lin_ws_slot_ints(d.max_len)) + 63) & ~63LL;
  auto fail = [&](const char* what) {
    srv_free(s);
    delete s;
    return set_err(c, POYHIP_ENOMEM, "server allocation failed: %s", what);
  };
  if (hipMalloc(&s->d_pool, d.pool_bytes) != hipSuccess) return fail("pool");
  if (hipEventCreateWithFlags(&s->ev_upload, hipEventDisableTiming) != hipSuccess)
    return fail("events");
  if (hipHostMalloc((void**)&s->h_pool, d.pool_bytes, hipHostMallocDefault) != hipSuccess)
    return fail("pinned pool");
  const int P = s->slot_pairs;
  s->slots.resize(d.n_slots);
  for (auto& sl : s->slots) {
    if (hipStreamCreateWithFlags(&sl.st, hipStreamNonBlocking) != hipSuccess) return fail("stream");
    if (hipEventCreateWithFlags(&sl.ev_start, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&sl.ev_done, hipEventDisableTiming) != hipSuccess)
      return fail("events");
    if (hipMalloc(&sl.d_tab, kSrvTabPer * (size_t)P + 256) != hipSuccess) return fail("tables");
    if (hipMalloc(&sl.d_res, sizeof(PairResult) * (size_t)P) != hipSuccess) return fail("results");
    if (hipMalloc(&sl.d_dm, (size_t)s->slot_dm + 256) != hipSuccess) return fail("direction store");
    if (hipMalloc(&sl.d_ws, sizeof(int) * (size_t)s->ws_stride * s->grid) != hipSuccess)
      return fail("workspace");
    // stripe tables: one entry per 64 rows plus one per pair
    const int64_t stab_n = (int64_t)s->slot_dm / 256 + 2 * (int64_t)P + 64;
    if (hipMalloc(&sl.d_stab, sizeof(long long) * (size_t)stab_n) != hipSuccess)
      return fail("stripe tables");
    if (hipMalloc(&sl.d_out, (size_t)s->slot_out + 256) != hipSuccess) return fail("outputs");
    if (hipMalloc(&sl.d_outlen, 16 * (size_t)P) != hipSuccess) return fail("lengths");
    if (hipMalloc(&sl.d_cost, 4 * (size_t)P) != hipSuccess) return fail("costs");
    if (hipMalloc(&sl.d_counters, 64) != hipSuccess) return fail("counters");
    if (hipMalloc(&sl.d_cls, (size_t)P + 64) != hipSuccess) return fail("classes");
    sl.up_bytes = kSrvTabPer * (size_t)P + 256;
    sl.down_bytes = 4096 + 20 * (size_t)P + (size_t)s->slot_out;
    if (hipHostMalloc((void**)&sl.h_up, sl.up_bytes, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void**)&sl.h_down, sl.down_bytes, hipHostMallocDefault) != hipSuccess)
      return fail("pinned staging");
  }
  s->worker = std::thread(srv_loop, s);
  *out = s;
  return POYHIP_OK;
}

int poyhip_server_destroy(poyhip_server* s) {
  if (!s) return POYHIP_EINVAL;
  {
    std::lock_guard<std::mutex> g(s->m);
    s->stop = true;
  }
  s->cv_work.notify_all();
  s->cv_done.notify_all();
  if (s->worker.joinable()) s->worker.join();
  srv_free(s);
  delete s;
  return POYHIP_OK;
}

int poyhip_server_put_seq(poyhip_server* s, const uint8_t* seq, int len, int64_t* id) {
  if (!s || !seq || len < 1 || !id) return POYHIP_EINVAL;
  if (len > s->max_len) return set_err(s->ctx, POYHIP_ECAP, "sequence of %d > max_len %d", len, s->max_len);
  std::lock_guard<std::mutex> g(s->m);
  const int64_t off = (s->pool_used + 15) & ~(int64_t)15;
  if (off + len > (int64_t)s->opt.pool_bytes)
    return set_err(s->ctx, POYHIP_ENOMEM, "sequence pool full (%zu bytes)", s->opt.pool_bytes);
  memcpy(s->h_pool + off, seq, len);
  s->pool_used = off + len;
  if (!s->uploads.empty() && s->uploads.back().second == off) s->uploads.back().second = off + len;
  else s->uploads.push_back({off, off + len});
  s->entries.push_back({off, len});
  *id = (int64_t)s->entries.size() - 1;
  return POYHIP_OK;
}

int poyhip_server_get_seq(poyhip_server* s, int64_t id, uint8_t* out, int cap, int* len) {
  if (!s || !out || !len) return POYHIP_EINVAL;
  PoolEntry e;
  {
    std::lock_guard<std::mutex> g(s->m);
    if (id < 0 || id >= (int64_t)s->entries.size()) return POYHIP_EINVAL;
    e = s->entries[id];
  }
  if (e.len < 0) return set_err(s->ctx, POYHIP_ESTATE, "sequence %lld not computed yet", (long long)id);
  if (e.len > cap) return set_err(s->ctx, POYHIP_ECAP, "capacity %d < %d", cap, e.len);
  (void)hipSetDevice(s->ctx->device);
  HIPCHK(s->ctx, hipMemcpy(out, s->d_pool + e.off, e.len, hipMemcpyDeviceToHost));
  *len = e.len;
  return POYHIP_OK;
}

int poyhip_server_submit(poyhip_server* s, int op, int64_t a, int64_t b, int deltaw, int flags,
                         int64_t* req, int64_t* median_id) {
  if (!s || !req) return POYHIP_EINVAL;
  if (op < POYHIP_OP_ALIGN_AFFINE || op > POYHIP_OP_COST_2D) return POYHIP_EINVAL;
  if ((flags & POYHIP_REQ_KEEP_MEDIAN) && (op != POYHIP_OP_ALIGN_2D || !median_id))
    return set_err(s->ctx, POYHIP_EINVAL, "only align_2d requests can keep their median");
  std::unique_lock<std::mutex> lk(s->m);
  if (s->error) return s->error;
  const int64_t ne = (int64_t)s->entries.size();
  if (a < 0 || b < 0 || a >= ne || b >= ne) return set_err(s->ctx, POYHIP_EINVAL, "unknown sequence id");
  if (s->entries[a].len < 0 || s->entries[b].len < 0) {
    // a median that is still being computed: wait until it exists, its
    // producer failed (len -2) or the server stops
    s->cv_done.wait(lk, [&] {
      return s->error || s->stop || s->entries[a].len == -2 || s->entries[b].len == -2 ||
             (s->entries[a].len >= 0 && s->entries[b].len >= 0);
    });
    if (s->error) return set_err(s->ctx, s->error, "the server failed while this request waited");
    if (s->stop) return set_err(s->ctx, POYHIP_ESTATE, "the server is stopping");
    if (s->entries[a].len < 0 || s->entries[b].len < 0)
      return set_err(s->ctx, POYHIP_EDEVICE, "the request computing a named median failed");
  }
  SrvRequest* r = new SrvRequest();
  r->id = s->next_id++;
  r->op = op;
  r->a = a;
  r->b = b;
  r->deltaw = deltaw;
  r->flags = flags;
  const int l1 = s->entries[a].len, l2 = s->entries[b].len;
  r->la = std::min(l1, l2);
  r->lb = std::max(l1, l2);
  r->dm_bytes = (pair_dm_bytes(op, r->la, r->lb) + 255) & ~(int64_t)255;
  if (r->lb > s->max_len) {
    const int lb = r->lb;
    delete r;
    return set_err(s->ctx, POYHIP_ECAP, "sequence of %d > max_len %d", lb, s->max_len);
  }
  if (r->dm_bytes > s->slot_dm || srv_out_bytes(op, r->la, r->lb) > s->slot_out) {
    delete r;
    return set_err(s->ctx, POYHIP_EBUSY, "request exceeds a slot's memory budget");
  }
  if (flags & POYHIP_REQ_KEEP_MEDIAN) {
    const int cap = (int)lin_out_cap(r->la, r->lb);
    const int64_t off = (s->pool_used + 15) & ~(int64_t)15;
    if (off + cap > (int64_t)s->opt.pool_bytes) {
      delete r;
      return set_err(s->ctx, POYHIP_ENOMEM, "sequence pool full");
    }
    s->pool_used = off + cap;
    // device-written: nothing to upload for this range
    s->entries.push_back({off, -1});
    r->median_id = (int64_t)s->entries.size() - 1;
    *median_id = r->median_id;
  }
  r->t_submit = Clock::now();
  if (!s->have_first) {
    s->have_first = true;
    s->t_first = r->t_submit;
  }
  s->reqs[r->id] = r;
  s->pending.push_back(r);
  *req = r->id;
  lk.unlock();
  s->cv_work.notify_one();
  return POYHIP_OK;
}

int poyhip_server_wait(poyhip_server* s, int64_t req, int timeout_ms) {
  if (!s) return POYHIP_EINVAL;
  std::unique_lock<std::mutex> lk(s->m);
  auto it = s->reqs.find(req);
  if (it == s->reqs.end()) return POYHIP_EINVAL;
  SrvRequest* r = it->second;
  auto done = [&] { return r->status == 2 || r->status < 0 || s->stop; };
  if (timeout_ms < 0) s->cv_done.wait(lk, done);
  else if (!s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), done)) return POYHIP_EBUSY;
  if (r->status < 0) return r->status;
  return r->status == 2 ? POYHIP_OK : set_err(s->ctx, POYHIP_ESTATE, "the server is stopping");
}

int poyhip_server_result(poyhip_server* s, int64_t req, int* cost, uint8_t* r1, uint8_t* r2,
                         int cap, int* len, double* latency_us, int release) {
  if (!s) return POYHIP_EINVAL;
  std::lock_guard<std::mutex> g(s->m);
  auto it = s->reqs.find(req);
  if (it == s->reqs.end()) return POYHIP_EINVAL;
  SrvRequest* r = it->second;
  if (r->status < 0) return r->status;
  if (r->status != 2) return POYHIP_EBUSY;
  if (cost) *cost = r->cost;
  if (latency_us) *latency_us = us_since(r->t_submit, r->t_done);
  if (len) *len = (int)r->r1.size();
  if (r1 || r2) {
    if ((int)r->r1.size() > cap || (int)r->r2.size() > cap)
      return set_err(s->ctx, POYHIP_ECAP, "capacity %d < %zu", cap, r->r1.size());
    if (r1) memcpy(r1, r->r1.data(), r->r1.size());
    if (r2) memcpy(r2, r->r2.data(), r->r2.size());
  }
  if (release) {
    s->reqs.erase(it);
    delete r;
  }
  return POYHIP_OK;
}

int poyhip_server_fail_next_launch(poyhip_server* s) {
  if (!s) return POYHIP_EINVAL;
  std::lock_guard<std::mutex> g(s->m);
  s->fail_next_launch = true;
  return POYHIP_OK;
}

int poyhip_server_stats(poyhip_server* s, int64_t* batches, int64_t* pairs, double* cells_full,
                        int* max_batch, int* max_inflight) {
  if (!s) return POYHIP_EINVAL;
  std::lock_guard<std::mutex> g(s->m);
  if (batches) *batches = s->batches;
  if (pairs) *pairs = s->pairs_done;
  if (cells_full) *cells_full = s->cells_full;
  if (max_batch) *max_batch = s->max_batch_seen;
  if (max_inflight) *max_inflight = s->max_inflight;
  return POYHIP_OK;
}

// Open-loop load generator: submits request k at arrival_us[k] (offsets
// from the call) from a native thread, waits for all, and reports each
// request's latency (submission to completion).  ops/a/b/deltaw/flags are
// per request; median ids are written to median_ids (or -1).
int poyhip_server_replay(poyhip_server* s, int n, const double* arrival_us, const int* ops,
                         const int64_t* a, const int64_t* b, const int* deltaw, const int* flags,
                         int* costs, double* latency_us, int64_t* median_ids, double* wall_us) {
  if (!s || n < 0 || !arrival_us || !ops || !a || !b) return POYHIP_EINVAL;
  std::vector<int64_t> ids(n, -1);
  const auto t0 = Clock::now();
  int rc = POYHIP_OK;
  for (int k = 0; k < n && rc == POYHIP_OK; k++) {
    const auto due = t0 + std::chrono::duration_cast<Clock::duration>(
                              std::chrono::duration<double, std::micro>(arrival_us[k]));
    while (Clock::now() < due) {
      const double left = us_since(Clock::now(), due);
      if (left > 200) std::this_thread::sleep_for(std::chrono::microseconds((int)(left - 100)));
    }
    int64_t mid = -1;
    rc = poyhip_server_submit(s, ops[k], a[k], b[k], deltaw ? deltaw[k] : -1, flags ? flags[k] : 0,
                              &ids[k], &mid);
    if (median_ids) median_ids[k] = mid;
  }
  for (int k = 0; k < n; k++) {
    if (ids[k] < 0) continue;
    int w = poyhip_server_wait(s, ids[k], -1);
    if (w != POYHIP_OK && rc == POYHIP_OK) rc = w;
    int cst = 0;
    double lat = 0;
    poyhip_server_result(s, ids[k], &cst, nullptr, nullptr, 0, nullptr, &lat, 1);
    if (costs) costs[k] = cst;
    if (latency_us) latency_us[k] = lat;
  }
  if (wall_us) *wall_us = us_since(t0, Clock::now());
  return rc;
}

}  // extern "C"
