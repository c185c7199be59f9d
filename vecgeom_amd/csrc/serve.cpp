// Serving loop: continuous batching of track-transport requests onto the GPU.
//
// Producers submit requests (host arrays of track positions, directions and per-track step
// counts) from any thread. A dispatcher thread picks the next requests (pick_batch) into
// the next free lane's pinned staging buffer and issues, on that lane's HIP stream, one
// host-to-device copy, one locate-and-propagate launch and one device-to-host copy; a
// completion thread unpacks a lane's results when its event has fired and marks the
// requests done. Lanes run concurrently, so copies of one launch overlap the kernels of
// another. Admission: a request larger than a lane is rejected, and the queue holds at
// most `queue_tracks` tracks (back-pressure); lane sizes are checked against the free HBM
// reported by the runtime at creation, and every submission and every batch against the
// free HBM reported at that moment (vgmi_server_set_hbm_reserve). Queued requests can be
// cancelled; completion can be polled without blocking.
#include "geom_internal.h"
#include "vgmi_internal.h"

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

using namespace vgmi;
using Clock = std::chrono::steady_clock;

namespace {

// Bytes per track in a lane: 6 doubles in, step count in, 3 doubles + state + steps taken
// out, and the device-only "last exited" state.
constexpr size_t kTrackBytes = 6 * 8 + 4 + 3 * 8 + 4 + 4 + 4;

struct Request {
  const double *px, *py, *pz, *dx, *dy, *dz;
  const int32_t *nsteps;
  int64_t n;
  double *opx, *opy, *opz;
  uint32_t *onav;
  int32_t *otaken;
  Clock::time_point tSubmit, tDone;
  int state  = 0; // 0 queued, 1 in flight, 2 done
  int status = 0; // vgmi_status of the launch that carried it
  int reason = 0; // why a request that never ran was finished (kNot*)
};

// Reasons a request finishes without running.
constexpr int kNotDestroyed = 1, kNotCancelled = 2, kNotHbm = 3;

struct Lane {
  hipStream_t stream = nullptr;
  hipEvent_t done    = nullptr;
  char *hstage       = nullptr; // pinned
  char *dstage       = nullptr;
  uint32_t *dwork    = nullptr; // regrouping workspace
  int64_t cap        = 0;
  int maxSteps       = 0;       // largest step count of the current batch
  bool busy          = false; // owns a batch (from packing until its results are unpacked)
  bool issued        = false; // its copies, launch and event record are on the stream
  int status         = 0;
  int64_t ntracks    = 0;
  std::vector<Request *> batch;
  std::vector<int64_t> offset;
  // HIP graphs of the locate + transport launches for track-count buckets (capacity, graph)
  std::vector<std::pair<int64_t, hipGraphExec_t>> graphs;
  // SoA sections of the staging buffers for the current batch (same layout on host and
  // device): inputs packed contiguously from offset 0, outputs from outBegin.
  size_t oPx, oPy, oPz, oDx, oDy, oDz, oSteps, inEnd, oOpx, oOpy, oOpz, oNav, oTaken, outBegin, outEnd, oLast;
};

} // namespace

struct vgmi_server {
  const vgmi_geom *geom = nullptr;
  int precision         = VGMI_FP64;
  double limit          = 0;
  int64_t queueCap      = 0;
  std::vector<Lane> lanes;
  std::mutex m;
  std::condition_variable cvWork, cvDone;
  std::deque<Request *> queue;
  int64_t queuedTracks = 0;
  std::unordered_map<uint64_t, Request *> requests;
  uint64_t nextTicket = 1;
  bool stop           = false;
  int waiters         = 0; // threads inside vgmi_server_wait (destroy waits for them to leave)
  int64_t hbmReserve  = 0; // bytes of HBM that must stay free for a request or batch to be admitted
  std::thread dispatcher, completer;
  int device = 0;
  // statistics
  int64_t launches = 0, tracksDone = 0, stepsDone = 0, rejected = 0, graphLaunches = 0;
};

namespace {
// Free HBM on the server's device now (0 when the runtime cannot tell).
int64_t free_hbm(const vgmi_server *S)
{
  int old = -1;
  size_t freeB = 0, totalB = 0;
  if (hipGetDevice(&old) != hipSuccess) return 0;
  const bool other = old != S->device;
  if (other && hipSetDevice(S->device) != hipSuccess) return 0;
  const bool ok = hipMemGetInfo(&freeB, &totalB) == hipSuccess;
  if (other) (void)hipSetDevice(old);
  return ok ? int64_t(freeB) : 0;
}
} // namespace

namespace {

size_t lane_bytes(int64_t cap) { return size_t(cap) * kTrackBytes + 17 * 256; }

// Section offsets for a batch of n tracks, each section 256-byte aligned, after a 256-byte header
// whose first word is the track count (read by the kernels of a graph launch).
void layout(Lane &L, size_t n)
{
  size_t o  = 256;
  auto take = [&](size_t bytes) {
    size_t at = o;
    o += (bytes + 255) / 256 * 256;
    return at;
  };
  L.oPx = take(8 * n), L.oPy = take(8 * n), L.oPz = take(8 * n);
  L.oDx = take(8 * n), L.oDy = take(8 * n), L.oDz = take(8 * n);
  L.oSteps   = take(4 * n);
  L.inEnd    = o;
  L.outBegin = o;
  L.oOpx = take(8 * n), L.oOpy = take(8 * n), L.oOpz = take(8 * n);
  L.oNav   = take(4 * n);
  L.oTaken = take(4 * n);
  L.outEnd = o;
  L.oLast  = take(4 * n); // device only
}

// A launch's latency is that of its slowest track, and small launches run with fewer tracks per
// wavefront (much faster steps), so requests are not packed across sizes: the next launch starts
// with the smallest queued request (or with the oldest one once it has waited kMaxWait, so large
// requests are not starved); a small first request takes further small ones, smallest first, up to
// kPackTracks tracks in all; a large one runs alone. Called with the server lock held and a
// non-empty queue.
constexpr int64_t kPackTracks = 4096;
// Launches whose tracks need at most kGraphSteps steps replay a captured graph sized for the
// smallest power-of-two bucket (from kPackTracks up to the lane capacity) that holds them.
constexpr int kGraphSteps = 100;
constexpr auto kMaxWait       = std::chrono::milliseconds(20);

void take(vgmi_server *S, Lane &L, std::deque<Request *>::iterator it)
{
  Request *r = *it;
  S->queue.erase(it);
  S->queuedTracks -= r->n;
  r->state = 1;
  L.offset.push_back(L.ntracks);
  L.batch.push_back(r);
  L.ntracks += r->n;
}

void pick_batch(vgmi_server *S, Lane &L)
{
  auto oldest = S->queue.begin();
  auto first  = oldest;
  if (Clock::now() - (*oldest)->tSubmit < kMaxWait)
    for (auto it = S->queue.begin(); it != S->queue.end(); ++it)
      if ((*it)->n < (*first)->n) first = it;
  const int64_t pack = kPackTracks < L.cap ? kPackTracks : L.cap;
  const bool small   = (*first)->n <= pack;
  take(S, L, first);
  while (small) {
    auto best = S->queue.end();
    for (auto it = S->queue.begin(); it != S->queue.end(); ++it)
      if (L.ntracks + (*it)->n <= pack && (best == S->queue.end() || (*it)->n < (*best)->n)) best = it;
    if (best == S->queue.end()) break;
    take(S, L, best);
  }
}

void dispatch_loop(vgmi_server *S)
{
  (void)hipSetDevice(S->device);
  for (;;) {
    Lane *L = nullptr;
    {
      std::unique_lock<std::mutex> lk(S->m);
      auto ready_lane = [&]() -> Lane * {
        if (S->queue.empty()) return nullptr;
        for (auto &l : S->lanes)
          if (!l.busy) return &l;
        return nullptr;
      };
      S->cvWork.wait(lk, [&] { return S->stop || ready_lane() != nullptr; });
      if (S->stop) return;
      L = ready_lane();
      L->batch.clear();
      L->offset.clear();
      L->ntracks = 0;
      pick_batch(S, *L);
      // admission of the batch against the free HBM now: a batch that would leave less than the
      // reserve free is not launched, and its requests finish with VGMI_ERR_CAPACITY
      if (S->hbmReserve > 0 && free_hbm(S) < S->hbmReserve) {
        const auto now = Clock::now();
        for (Request *r : L->batch) {
          r->state  = 2;
          r->status = VGMI_ERR_CAPACITY;
          r->reason = kNotHbm;
          r->tDone  = now;
        }
        S->rejected += int64_t(L->batch.size());
        L->batch.clear();
        S->cvDone.notify_all();
        continue;
      }
      L->busy   = true;
      L->issued = false;
    }
    // a graph launch when a bucket holds the batch and its step counts are within kGraphSteps
    L->maxSteps = 0;
    for (Request const *r : L->batch)
      for (int64_t i = 0; i < r->n; ++i) L->maxSteps = r->nsteps[i] > L->maxSteps ? r->nsteps[i] : L->maxSteps;
    hipGraphExec_t graph = nullptr;
    int64_t cap          = L->ntracks;
    if (L->maxSteps <= kGraphSteps)
      for (auto const &g : L->graphs)
        if (g.first >= L->ntracks) {
          graph = g.second, cap = g.first;
          break;
        }
    layout(*L, size_t(cap));
    // pack inputs (outside the lock)
    char *h = L->hstage;
    *reinterpret_cast<uint32_t *>(h) = uint32_t(L->ntracks);
    for (size_t k = 0; k < L->batch.size(); ++k) {
      Request const *r = L->batch[k];
      const size_t o = size_t(L->offset[k]);
      std::memcpy(h + L->oPx + 8 * o, r->px, 8 * r->n);
      std::memcpy(h + L->oPy + 8 * o, r->py, 8 * r->n);
      std::memcpy(h + L->oPz + 8 * o, r->pz, 8 * r->n);
      std::memcpy(h + L->oDx + 8 * o, r->dx, 8 * r->n);
      std::memcpy(h + L->oDy + 8 * o, r->dy, 8 * r->n);
      std::memcpy(h + L->oDz + 8 * o, r->dz, 8 * r->n);
      std::memcpy(h + L->oSteps + 4 * o, r->nsteps, 4 * r->n);
    }
    char *d        = L->dstage;
    const size_t n = size_t(L->ntracks);
    auto D         = [&](size_t off) { return reinterpret_cast<double *>(d + off); };
    auto U         = [&](size_t off) { return reinterpret_cast<uint32_t *>(d + off); };
    auto I         = [&](size_t off) { return reinterpret_cast<int32_t *>(d + off); };
    TransportArgs a{D(L->oOpx), D(L->oOpy), D(L->oOpz), D(L->oDx),    D(L->oDy),       D(L->oDz), S->limit,
                    U(L->oNav), U(L->oLast), I(L->oTaken), I(L->oSteps), n,               L->maxSteps,
                    S->precision, nullptr};
    hipError_t e = hipMemcpyAsync(d, h, L->inEnd, hipMemcpyHostToDevice, L->stream);
    int st       = e == hipSuccess ? VGMI_OK : VGMI_ERR_DEVICE;
    if (st == VGMI_OK && graph) {
      if (hipGraphLaunch(graph, L->stream) != hipSuccess) st = VGMI_ERR_DEVICE;
    } else if (st == VGMI_OK) {
      st = launch_locate_init(S->geom, S->precision, D(L->oPx), D(L->oPy), D(L->oPz), a.px, a.py, a.pz, a.nav, a.last,
                              a.taken, n, L->stream);
      if (st == VGMI_OK) st = launch_transport(S->geom, a, L->dwork, L->stream);
    }
    if (st == VGMI_OK)
      e = hipMemcpyAsync(h + L->outBegin, d + L->outBegin, L->outEnd - L->outBegin, hipMemcpyDeviceToHost, L->stream);
    if (st == VGMI_OK && e == hipSuccess) e = hipEventRecord(L->done, L->stream);
    if (st == VGMI_OK && e != hipSuccess) st = VGMI_ERR_DEVICE;
    if (st != VGMI_OK) (void)hipStreamSynchronize(L->stream); // the event is not used then
    {
      std::lock_guard<std::mutex> lk(S->m);
      L->status = st;
      L->issued = true;
      S->launches++;
      if (graph) S->graphLaunches++;
    }
    S->cvWork.notify_all(); // wakes the completion thread
  }
}

void complete_loop(vgmi_server *S)
{
  (void)hipSetDevice(S->device);
  for (;;) {
    std::vector<Lane *> ready;
    {
      std::unique_lock<std::mutex> lk(S->m);
      bool anyBusy = false;
      for (auto &l : S->lanes) anyBusy = anyBusy || l.busy;
      if (S->stop && !anyBusy) return;
      if (!anyBusy) {
        S->cvWork.wait_for(lk, std::chrono::milliseconds(1));
        continue;
      }
      for (auto &l : S->lanes)
        if (l.issued && (l.status != VGMI_OK || hipEventQuery(l.done) != hipErrorNotReady)) ready.push_back(&l);
    }
    if (ready.empty()) {
      std::this_thread::sleep_for(std::chrono::microseconds(20));
      continue;
    }
    for (Lane *L : ready) {
      char const *h = L->hstage;
      int64_t steps = 0;
      if (L->status == VGMI_OK && hipEventSynchronize(L->done) != hipSuccess) L->status = VGMI_ERR_DEVICE;
      for (size_t k = 0; L->status == VGMI_OK && k < L->batch.size(); ++k) {
        Request *r     = L->batch[k];
        const size_t o = size_t(L->offset[k]);
        std::memcpy(r->opx, h + L->oOpx + 8 * o, 8 * r->n);
        std::memcpy(r->opy, h + L->oOpy + 8 * o, 8 * r->n);
        std::memcpy(r->opz, h + L->oOpz + 8 * o, 8 * r->n);
        if (r->onav) std::memcpy(r->onav, h + L->oNav + 4 * o, 4 * r->n);
        const int32_t *tk = reinterpret_cast<const int32_t *>(h + L->oTaken) + o;
        if (r->otaken) std::memcpy(r->otaken, tk, 4 * r->n);
        for (int64_t i = 0; i < r->n; ++i) steps += tk[i];
      }
      const auto now = Clock::now();
      {
        std::lock_guard<std::mutex> lk(S->m);
        for (Request *r : L->batch) {
          r->tDone  = now;
          r->state  = 2;
          r->status = L->status;
          if (L->status == VGMI_OK) S->tracksDone += r->n;
        }
        if (L->status == VGMI_OK) S->stepsDone += steps;
        L->batch.clear();
        L->busy   = false;
        L->issued = false;
      }
      S->cvDone.notify_all();
      S->cvWork.notify_all();
    }
  }
}

// Captures, on the lane's stream, the locate + transport launches of a batch of `cap` tracks laid
// out by layout(cap), reading the actual track count from the staging header, for each bucket.
hipError_t capture_graphs(vgmi_server *S, Lane &L)
{
  for (int64_t cap = kPackTracks; ; cap *= 2) {
    if (cap > L.cap) cap = L.cap;
    layout(L, size_t(cap));
    char *d        = L.dstage;
    auto D         = [&](size_t off) { return reinterpret_cast<double *>(d + off); };
    auto U         = [&](size_t off) { return reinterpret_cast<uint32_t *>(d + off); };
    auto I         = [&](size_t off) { return reinterpret_cast<int32_t *>(d + off); };
    const uint32_t *dn = reinterpret_cast<const uint32_t *>(d);
    TransportArgs a{D(L.oOpx), D(L.oOpy), D(L.oOpz), D(L.oDx), D(L.oDy), D(L.oDz), S->limit, U(L.oNav), U(L.oLast),
                    I(L.oTaken), I(L.oSteps), size_t(cap), kGraphSteps, S->precision, dn};
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamBeginCapture(L.stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return e;
    int st = launch_locate_init(S->geom, S->precision, D(L.oPx), D(L.oPy), D(L.oPz), a.px, a.py, a.pz, a.nav, a.last,
                                a.taken, size_t(cap), L.stream, dn);
    if (st == VGMI_OK) st = launch_transport(S->geom, a, L.dwork, L.stream);
    e = hipStreamEndCapture(L.stream, &g);
    if (e == hipSuccess && st != VGMI_OK) e = hipErrorLaunchFailure;
    hipGraphExec_t x = nullptr;
    if (e == hipSuccess) e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    if (g) (void)hipGraphDestroy(g);
    if (e != hipSuccess) return e;
    L.graphs.emplace_back(cap, x);
    if (cap == L.cap) return hipSuccess;
  }
}

} // namespace

extern "C" {

vgmi_status vgmi_server_create(const vgmi_geom *geom, int precision, double step_limit, int64_t lane_tracks,
                               int lanes, int64_t queue_tracks, vgmi_server **out)
{
  if (!out) return set_error(VGMI_ERR_INVALID, "vgmi_server_create: null output handle");
  *out = nullptr;
  if (!geom || !geom->device_block) return set_error(VGMI_ERR_INVALID, "vgmi_server_create: geometry not on a device");
  if (precision != VGMI_FP64 && precision != VGMI_MIXED)
    return set_error(VGMI_ERR_INVALID, "vgmi_server_create: unknown precision %d", precision);
  if (lane_tracks < 1 || lanes < 1 || lanes > 16 || queue_tracks < lane_tracks)
    return set_error(VGMI_ERR_INVALID, "vgmi_server_create: need lane_tracks >= 1, 1..16 lanes, queue >= lane");
  size_t freeB = 0, totalB = 0;
  if (hipSetDevice(geom->device) != hipSuccess || hipMemGetInfo(&freeB, &totalB) != hipSuccess) return set_error(VGMI_ERR_DEVICE, "vgmi_server_create: no device");
  const size_t need = size_t(lanes) * lane_bytes(lane_tracks);
  if (need > freeB / 2)
    return set_error(VGMI_ERR_CAPACITY, "vgmi_server_create: %zu bytes of lanes exceed half of the free HBM (%zu)", need,
                     freeB);
  auto *S      = new vgmi_server();
  S->geom      = geom;
  S->precision = precision;
  S->limit     = step_limit;
  S->queueCap  = queue_tracks;
  S->device    = geom->device;
  S->lanes.resize(lanes);
  for (auto &L : S->lanes) {
    L.cap = lane_tracks;
    const size_t bytes = lane_bytes(lane_tracks);
    hipError_t e = hipStreamCreateWithFlags(&L.stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&L.done, hipEventDisableTiming);
    if (e == hipSuccess) e = hipHostMalloc(reinterpret_cast<void **>(&L.hstage), bytes, hipHostMallocDefault);
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void **>(&L.dstage), bytes);
    const size_t words = transport_work_words(geom, size_t(lane_tracks));
    if (e == hipSuccess && words) e = hipMalloc(reinterpret_cast<void **>(&L.dwork), words * sizeof(uint32_t));
    if (e == hipSuccess) e = capture_graphs(S, L);
    if (e != hipSuccess) {
      vgmi_server_destroy(S);
      return set_error(VGMI_ERR_DEVICE, "vgmi_server_create: %s", hipGetErrorString(e));
    }
  }
  S->dispatcher = std::thread(dispatch_loop, S);
  S->completer  = std::thread(complete_loop, S);
  *out          = S;
  return VGMI_OK;
}

vgmi_status vgmi_server_submit(vgmi_server *S, const double *px, const double *py, const double *pz, const double *dx,
                               const double *dy, const double *dz, const int32_t *nsteps, int64_t n, double *px_out,
                               double *py_out, double *pz_out, uint32_t *nav_out, int32_t *steps_out,
                               uint64_t *ticket)
{
  if (!S || !ticket) return set_error(VGMI_ERR_INVALID, "vgmi_server_submit: null argument");
  if (n < 1 || !px || !py || !pz || !dx || !dy || !dz || !nsteps || !px_out || !py_out || !pz_out)
    return set_error(VGMI_ERR_INVALID, "vgmi_server_submit: empty request or null array");
  auto *r = new Request{px,     py,      pz,     dx,        dy,           dz,
                        nsteps, n,       px_out, py_out,    pz_out,       nav_out,
                        steps_out, Clock::now(), Clock::time_point(), 0, 0};
  {
    std::lock_guard<std::mutex> lk(S->m);
    if (S->stop) {
      delete r;
      return set_error(VGMI_ERR_INVALID, "vgmi_server_submit: the server is being destroyed");
    }
    if (S->hbmReserve > 0 && free_hbm(S) < S->hbmReserve) {
      S->rejected++;
      delete r;
      return set_error(VGMI_ERR_CAPACITY, "vgmi_server_submit: request of %lld tracks not admitted (free HBM below "
                       "the reserve of %lld bytes)", (long long)n, (long long)S->hbmReserve);
    }
    if (n > S->lanes[0].cap || S->queuedTracks + n > S->queueCap) {
      S->rejected++;
      delete r;
      return set_error(VGMI_ERR_CAPACITY, "vgmi_server_submit: request of %lld tracks not admitted (queue %lld/%lld)",
                       (long long)n, (long long)S->queuedTracks, (long long)S->queueCap);
    }
    *ticket              = S->nextTicket++;
    S->requests[*ticket] = r;
    S->queue.push_back(r);
    S->queuedTracks += n;
  }
  S->cvWork.notify_all();
  return VGMI_OK;
}

vgmi_status vgmi_server_wait(vgmi_server *S, uint64_t ticket, double *latency_us)
{
  if (!S) return set_error(VGMI_ERR_INVALID, "vgmi_server_wait: null server");
  std::unique_lock<std::mutex> lk(S->m);
  auto it = S->requests.find(ticket);
  if (it == S->requests.end()) return set_error(VGMI_ERR_INVALID, "vgmi_server_wait: unknown ticket %llu",
                                                (unsigned long long)ticket);
  Request *r = it->second;
  ++S->waiters;
  S->cvDone.wait(lk, [&] { return r->state == 2; });
  --S->waiters;
  if (latency_us) *latency_us = std::chrono::duration<double, std::micro>(r->tDone - r->tSubmit).count();
  const int st     = r->status;
  const int reason = r->reason;
  S->requests.erase(it);
  delete r;
  if (S->stop) S->cvDone.notify_all(); // vgmi_server_destroy waits for the last waiter to leave
  if (st == VGMI_OK) return VGMI_OK;
  if (reason == kNotDestroyed)
    return set_error(st, "vgmi_server_wait: ticket %llu was never launched (the server was destroyed)",
                     (unsigned long long)ticket);
  if (reason == kNotCancelled)
    return set_error(st, "vgmi_server_wait: ticket %llu was cancelled before its launch", (unsigned long long)ticket);
  if (reason == kNotHbm)
    return set_error(st, "vgmi_server_wait: ticket %llu was not admitted (free HBM below the reserve)",
                     (unsigned long long)ticket);
  return set_error(st, "vgmi_server_wait: the launch carrying ticket %llu failed", (unsigned long long)ticket);
}

vgmi_status vgmi_server_poll(vgmi_server *S, uint64_t ticket, int *done)
{
  if (!S || !done) return set_error(VGMI_ERR_INVALID, "vgmi_server_poll: null argument");
  std::lock_guard<std::mutex> lk(S->m);
  auto it = S->requests.find(ticket);
  if (it == S->requests.end())
    return set_error(VGMI_ERR_INVALID, "vgmi_server_poll: unknown ticket %llu", (unsigned long long)ticket);
  *done = it->second->state == 2 ? 1 : 0;
  return VGMI_OK;
}

vgmi_status vgmi_server_cancel(vgmi_server *S, uint64_t ticket)
{
  if (!S) return set_error(VGMI_ERR_INVALID, "vgmi_server_cancel: null server");
  {
    std::lock_guard<std::mutex> lk(S->m);
    auto it = S->requests.find(ticket);
    if (it == S->requests.end())
      return set_error(VGMI_ERR_INVALID, "vgmi_server_cancel: unknown ticket %llu", (unsigned long long)ticket);
    Request *r = it->second;
    if (r->state != 0)
      return set_error(VGMI_ERR_UNSUPPORTED, "vgmi_server_cancel: ticket %llu is %s", (unsigned long long)ticket,
                       r->state == 1 ? "on the GPU" : "already done");
    for (auto q = S->queue.begin(); q != S->queue.end(); ++q)
      if (*q == r) {
        S->queue.erase(q);
        break;
      }
    S->queuedTracks -= r->n;
    r->state  = 2;
    r->status = VGMI_ERR_CAPACITY;
    r->reason = kNotCancelled;
    r->tDone  = Clock::now();
  }
  S->cvDone.notify_all();
  return VGMI_OK;
}

vgmi_status vgmi_server_set_hbm_reserve(vgmi_server *S, int64_t bytes)
{
  if (!S) return set_error(VGMI_ERR_INVALID, "vgmi_server_set_hbm_reserve: null server");
  if (bytes < 0) return set_error(VGMI_ERR_INVALID, "vgmi_server_set_hbm_reserve: negative reserve");
  std::lock_guard<std::mutex> lk(S->m);
  S->hbmReserve = bytes;
  return VGMI_OK;
}

vgmi_status vgmi_server_stats(vgmi_server *S, int64_t stats[4])
{
  if (!S || !stats) return set_error(VGMI_ERR_INVALID, "vgmi_server_stats: null argument");
  std::lock_guard<std::mutex> lk(S->m);
  stats[0] = S->launches;
  stats[1] = S->tracksDone;
  stats[2] = S->stepsDone;
  stats[3] = S->rejected;
  return VGMI_OK;
}

vgmi_status vgmi_server_graph_stats(vgmi_server *S, int64_t stats[2])
{
  if (!S || !stats) return set_error(VGMI_ERR_INVALID, "vgmi_server_graph_stats: null argument");
  std::lock_guard<std::mutex> lk(S->m);
  stats[0] = S->graphLaunches;
  stats[1] = S->lanes.empty() ? 0 : int64_t(S->lanes[0].graphs.size());
  return VGMI_OK;
}

vgmi_status vgmi_server_destroy(vgmi_server *S)
{
  if (!S) return VGMI_OK;
  {
    std::lock_guard<std::mutex> lk(S->m);
    S->stop = true;
  }
  S->cvWork.notify_all();
  if (S->dispatcher.joinable()) S->dispatcher.join();
  if (S->completer.joinable()) S->completer.join();
  {
    // requests still queued are never launched: finish them with an error, wake their waiters and
    // let every waiter leave before the server's memory goes away
    std::unique_lock<std::mutex> lk(S->m);
    const auto now = Clock::now();
    for (auto &kv : S->requests)
      if (kv.second->state != 2) {
        kv.second->state  = 2;
        kv.second->status = VGMI_ERR_CAPACITY;
        kv.second->reason = kNotDestroyed;
        kv.second->tDone  = now;
      }
    S->queue.clear();
    S->cvDone.notify_all();
    S->cvDone.wait(lk, [&] { return S->waiters == 0; });
  }
  for (auto &L : S->lanes) {
    if (L.stream) (void)hipStreamSynchronize(L.stream);
    for (auto &g : L.graphs) (void)hipGraphExecDestroy(g.second);
    if (L.dstage) (void)hipFree(L.dstage);
    if (L.dwork) (void)hipFree(L.dwork);
    if (L.hstage) (void)hipHostFree(L.hstage);
    if (L.done) (void)hipEventDestroy(L.done);
    if (L.stream) (void)hipStreamDestroy(L.stream);
  }
  for (auto &kv : S->requests) delete kv.second;
  delete S;
  return VGMI_OK;
}

} // extern "C"
