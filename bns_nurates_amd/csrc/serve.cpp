// serve.cpp -- continuous-batching request server for the gray M1 opacities.
//
// Callers submit opacity requests (a meshblock of points in host SoA form,
// a quadrature order, FP64 or mixed precision, flags and parameters) from any
// thread and wait for them by id.  One scheduler thread owns the device side:
//
//   admission  each in-flight batch takes (21 + 20) doubles per point from one
//              device arena allocated once at start-up (the HBM budget); a
//              batch is admitted when the arena can hold it and one of the
//              server's streams is free, so nothing is allocated per request
//   pieces     a request larger than the chunk size is split into pieces of
//              at most that many points, queued one after another; the
//              request completes with its last piece;
//   batching   the pending piece of the smallest request (ties: the earliest)
//              opens a batch, so that small requests do not wait behind a
//              large one's pieces; a piece that has waited longer than the
//              starvation limit goes first; later pending pieces with the
//              same quadrature order, precision, flags and parameters join
//              it up to the chunk size, so that no launch holds the device
//              for longer than one chunk (points are independent, so a
//              request's results do not depend on its batch or pieces);
//   overlap    batches run on separate HIP streams (upload, kernel, download,
//              event); requests that arrive while others are on the device
//              are admitted as soon as a stream and arena space free up.
//   staging    pinned caller buffers are copied asynchronously; pageable ones
//              go through a pinned mirror of the device arena (a copy on the
//              host at each end), so that no copy blocks the scheduler's
//              stream work.  The scheduler holds its lock only to move
//              requests between queues: copies and launches run outside it,
//              and requests keep arriving (and batching) meanwhile.
//
// The quadrature tables of each order are built once and stay resident.
// Per-request latency is the time from submit to the scheduler observing the
// batch's completion event.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

#include "../../include/bns_nurates_amd.h"
#include "gray_kernel.h"
#include "gray_tables.h"

namespace {

using Clock = std::chrono::steady_clock;

constexpr size_t kAlign = 256;
constexpr int kFieldsPerPoint = BNSAMD_PT_NFIELDS + BNSAMD_OUT_NFIELDS;

size_t batch_bytes(int64_t npts) {
  size_t b = size_t(npts) * kFieldsPerPoint * sizeof(double);
  return (b + kAlign - 1) / kAlign * kAlign;
}

// First-fit sub-allocator over one device allocation.
class Arena {
 public:
  hipError_t init(size_t bytes) {
    size_ = bytes / kAlign * kAlign;
    hipError_t e = hipMalloc(&base_, size_);
    if (e == hipSuccess) free_[0] = size_;
    return e;
  }
  void release() {
    if (base_) (void)hipFree(base_);
    base_ = nullptr;
  }
  // offset of a block of `bytes`, or -1 when no free block is large enough
  int64_t alloc(size_t bytes) {
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second >= bytes) {
        size_t off = it->first, rest = it->second - bytes;
        free_.erase(it);
        if (rest) free_[off + bytes] = rest;
        used_ += bytes;
        return int64_t(off);
      }
    }
    return -1;
  }
  void free(size_t off, size_t bytes) {
    used_ -= bytes;
    auto it = free_.emplace(off, bytes).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
  char* base() const { return static_cast<char*>(base_); }
  size_t size() const { return size_; }
  size_t used() const { return used_; }

 private:
  void* base_ = nullptr;
  size_t size_ = 0, used_ = 0;
  std::map<size_t, size_t> free_;  // offset -> size
};

struct Request {
  int64_t id = 0, npts = 0;
  int64_t pieces_left = 0;
  const double* h_pts = nullptr;
  double* h_out = nullptr;
  bool pinned_in = false, pinned_out = false;  // page-locked caller buffers
  int nquad = 0, precision = 0, stim = 1;
  bnsamd_flags flags{};
  bnsamd_params params{};
  Clock::time_point t_submit, t_done;
  bool done = false;
  int status = BNSAMD_OK;
};

using Key = std::tuple<int, int, int, int, int, int, int, int, int, int, int, int, int, int, int,
                       int>;
Key key_of(const Request& r) {
  const bnsamd_flags& f = r.flags;
  const bnsamd_params& p = r.params;
  return Key{r.nquad, r.precision, r.stim, f.use_abs_em, f.use_pair, f.use_brem,
             f.use_inelastic_scatt, f.use_iso, p.use_dU, p.use_dm_eff, p.use_WM_ab,
             p.use_WM_sc, p.use_decay, p.brem_implementation, p.use_NN_medium_corr,
             p.neglect_blocking};
}

// rows [off, off + n) of a request's point and output blocks
struct Piece {
  Request* r = nullptr;
  int64_t off = 0, n = 0;
  Clock::time_point t_queued;
};

// whether host memory is page-locked (copies from it run asynchronously)
bool is_pinned(const void* p) {
  hipPointerAttribute_t a;
  if (hipPointerGetAttributes(&a, p) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return a.type == hipMemoryTypeHost;
}

struct Tables {
  bns::TableLayout L{};
  double* dev = nullptr;
};

struct Slot {
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  bool busy = false;
  std::vector<Piece> reqs;
  size_t off = 0, bytes = 0;
  int64_t npts = 0;
};

}  // namespace

struct bnsamd_server {
  bnsamd_server_config cfg{};
  Arena arena;
  char* hmirror = nullptr;  // pinned host mirror of the arena (pageable requests)
  std::vector<Slot> slots;
  std::map<int, Tables> tables;
  double gamma36 = 0.;

  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Piece> pending;
  std::map<int64_t, std::unique_ptr<Request>> requests;
  int64_t next_id = 1;
  bool stop = false;
  int64_t chunk = 0;                    // points per piece
  Clock::duration starve{};             // a piece waiting longer goes first
  std::string err;
  bnsamd_server_stats stats{};
  std::thread worker;

  // ---- scheduler thread ---------------------------------------------------
  // resident tables of order n (scheduler thread only)
  int table_for(int n, const Tables** out) {
    auto it = tables.find(n);
    if (it == tables.end()) {
      double x[bns::kMaxQuad], w[bns::kMaxQuad];
      bnsamd_gauss_legendre(n, 0., 1., x, w);
      Tables t;
      t.L = bns::gray_table_layout(n, n);
      std::vector<double> host(t.L.total);
      bns::build_gray_tables(n, x, w, n, x, w, host.data());
      if (hipMalloc(&t.dev, sizeof(double) * t.L.total) != hipSuccess) return BNSAMD_ERR_NOMEM;
      if (hipMemcpy(t.dev, host.data(), sizeof(double) * t.L.total, hipMemcpyHostToDevice) !=
          hipSuccess) {
        (void)hipFree(t.dev);
        return BNSAMD_ERR_HIP;
      }
      it = tables.emplace(n, t).first;
    }
    *out = &it->second;
    return BNSAMD_OK;
  }

  void finish(Slot& s, int status) {
    Clock::time_point now = Clock::now();
    for (const Piece& pc : s.reqs) {
      Request* r = pc.r;
      if (status != BNSAMD_OK) r->status = status;
      stats.points_completed += pc.n;
      if (--r->pieces_left == 0) {
        r->done = true;
        r->t_done = now;
        stats.requests_completed += 1;
      }
    }
    arena.free(s.off, s.bytes);
    s.reqs.clear();
    s.busy = false;
    cv_done.notify_all();
  }

  // (lock held) slots whose batch has completed, with their status
  void collect_done(std::vector<std::pair<Slot*, int>>& done) {
    for (Slot& s : slots) {
      if (!s.busy) continue;
      hipError_t e = hipEventQuery(s.done);
      if (e == hipSuccess) done.emplace_back(&s, BNSAMD_OK);
      else if (e != hipErrorNotReady) done.emplace_back(&s, BNSAMD_ERR_HIP);
    }
  }

  // (lock held) forms batches while a stream and arena space are free; the
  // returned slots are marked busy and hold their requests
  void form_batches(std::vector<Slot*>& fresh) {
    for (;;) {
      if (pending.empty()) return;
      Slot* slot = nullptr;
      for (Slot& s : slots)
        if (!s.busy) {
          slot = &s;
          break;
        }
      if (!slot) return;
      // pieces in the order they are served: a piece waiting beyond the
      // starvation limit first (oldest first), then by the size of its
      // request (smallest first), then by queue order (pieces of one request
      // keep theirs)
      const Clock::time_point now = Clock::now();
      std::vector<size_t> order(pending.size());
      for (size_t i = 0; i < order.size(); ++i) order[i] = i;
      auto starved = [&](const Piece& pc) { return now - pc.t_queued > starve; };
      std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) {
        const Piece &pa = pending[a], &pb = pending[b];
        const bool sa = starved(pa), sb = starved(pb);
        if (sa != sb) return sa;
        if (sa) return false;  // both starved: queue order
        return pa.r->npts < pb.r->npts;
      });
      // the first piece opens the batch; same-key pieces in that order join
      const Key k = key_of(*pending[order[0]].r);
      std::vector<size_t> take_idx{order[0]};
      int64_t n = pending[order[0]].n;
      for (size_t j = 1; j < order.size(); ++j) {
        const Piece& pc = pending[order[j]];
        if (n + pc.n > chunk) continue;  // a launch covers at most one chunk of points
        if (key_of(*pc.r) == k) {
          take_idx.push_back(order[j]);
          n += pc.n;
        }
      }
      size_t bytes = batch_bytes(n);
      int64_t off = arena.alloc(bytes);
      while (off < 0 && take_idx.size() > 1) {  // shrink to what fits now
        n -= pending[take_idx.back()].n;
        take_idx.pop_back();
        bytes = batch_bytes(n);
        off = arena.alloc(bytes);
      }
      if (off < 0) return;  // admission: wait for in-flight batches to free memory
      std::vector<Piece> take;
      for (size_t i : take_idx) take.push_back(pending[i]);
      std::sort(take_idx.begin(), take_idx.end());
      for (size_t j = take_idx.size(); j-- > 0;) pending.erase(pending.begin() + take_idx[j]);
      slot->busy = true;
      slot->reqs = std::move(take);
      slot->off = size_t(off);
      slot->bytes = bytes;
      slot->npts = n;
      stats.batches += 1;
      stats.max_inflight_bytes = std::max<int64_t>(stats.max_inflight_bytes, arena.used());
      int64_t busy = 0;
      for (Slot& s : slots) busy += s.busy;
      stats.max_concurrent_batches = std::max(stats.max_concurrent_batches, busy);
      fresh.push_back(slot);
    }
  }

  // (lock free) pinned mirror of the arena, allocated at the first pageable request
  int ensure_mirror() {
    if (hmirror) return BNSAMD_OK;
    void* p = nullptr;
    if (hipHostMalloc(&p, arena.size(), hipHostMallocDefault) != hipSuccess) return BNSAMD_ERR_NOMEM;
    hmirror = static_cast<char*>(p);
    return BNSAMD_OK;
  }

  // (lock free) upload, kernel and download of one batch on its stream
  int enqueue(Slot& s) {
    const Tables* tab = nullptr;
    const Request& h = *s.reqs.front().r;
    if (int rc = table_for(h.nquad, &tab)) return rc;
    const int64_t B = s.npts;
    const size_t ld = size_t(B) * sizeof(double);
    double* d_pts = reinterpret_cast<double*>(arena.base() + s.off);
    double* d_out = d_pts + size_t(B) * BNSAMD_PT_NFIELDS;
    bool staged = false;
    for (const Piece& pc : s.reqs) staged = staged || !pc.r->pinned_in || !pc.r->pinned_out;
    if (staged)
      if (int rc = ensure_mirror()) return rc;
    double* m_pts = staged ? reinterpret_cast<double*>(hmirror + s.off) : nullptr;
    double* m_out = staged ? m_pts + size_t(B) * BNSAMD_PT_NFIELDS : nullptr;
    int64_t o = 0;
    for (const Piece& pc : s.reqs) {
      const Request* r = pc.r;
      const size_t w = size_t(pc.n) * sizeof(double);
      const double* src = r->h_pts + pc.off;
      size_t spitch = size_t(r->npts) * sizeof(double);
      if (!r->pinned_in) {  // stage into the pinned mirror, then copy from there
        for (int f = 0; f < BNSAMD_PT_NFIELDS; ++f)
          std::memcpy(m_pts + size_t(f) * B + o, r->h_pts + size_t(f) * r->npts + pc.off, w);
        src = m_pts + o;
        spitch = ld;
      }
      if (hipMemcpy2DAsync(d_pts + o, ld, src, spitch, w, BNSAMD_PT_NFIELDS,
                           hipMemcpyHostToDevice, s.stream) != hipSuccess)
        return BNSAMD_ERR_HIP;
      o += pc.n;
    }
    bns::GrayKernelArgs a{};
    a.npts = B;
    a.ld = B;
    a.pts = d_pts;
    a.distr = nullptr;
    a.out = d_out;
    a.tab = tab->dev;
    a.L = tab->L;
    a.flags = bns::PointFlags{h.flags.use_abs_em, h.flags.use_pair, h.flags.use_brem,
                              h.flags.use_inelastic_scatt, h.flags.use_iso};
    a.params = bns::PointParams{h.params.use_dU, h.params.use_dm_eff, h.params.use_WM_ab,
                                h.params.use_WM_sc, h.params.use_decay,
                                h.params.brem_implementation, h.params.use_NN_medium_corr,
                                h.params.neglect_blocking};
    a.stim = h.stim;
    a.gamma36 = gamma36;
    hipError_t e = h.precision == BNSAMD_MIXED ? bns::launch_gray_m1_mixed(a, s.stream)
                                               : bns::launch_gray_m1_f64(a, s.stream);
    if (e != hipSuccess) return BNSAMD_ERR_HIP;
    o = 0;
    for (const Piece& pc : s.reqs) {
      const Request* r = pc.r;
      const size_t w = size_t(pc.n) * sizeof(double);
      double* dst = r->pinned_out ? r->h_out + pc.off : m_out + o;
      size_t dpitch = r->pinned_out ? size_t(r->npts) * sizeof(double) : ld;
      if (hipMemcpy2DAsync(dst, dpitch, d_out + o, ld, w, BNSAMD_OUT_NFIELDS,
                           hipMemcpyDeviceToHost, s.stream) != hipSuccess)
        return BNSAMD_ERR_HIP;
      o += pc.n;
    }
    return hipEventRecord(s.done, s.stream) == hipSuccess ? BNSAMD_OK : BNSAMD_ERR_HIP;
  }

  // (lock free) outputs of pageable requests from the pinned mirror
  void unstage(Slot& s) {
    const int64_t B = s.npts;
    const double* m_out =
        hmirror ? reinterpret_cast<double*>(hmirror + s.off) + size_t(B) * BNSAMD_PT_NFIELDS
                : nullptr;
    int64_t o = 0;
    for (const Piece& pc : s.reqs) {
      const Request* r = pc.r;
      if (!r->pinned_out && m_out)
        for (int f = 0; f < BNSAMD_OUT_NFIELDS; ++f)
          std::memcpy(r->h_out + size_t(f) * r->npts + pc.off, m_out + size_t(f) * B + o,
                      size_t(pc.n) * sizeof(double));
      o += pc.n;
    }
  }

  void run() {
    (void)hipSetDevice(cfg.device);
    std::vector<std::pair<Slot*, int>> done;
    std::vector<Slot*> fresh;
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
      done.clear();
      fresh.clear();
      collect_done(done);
      if (!done.empty()) {
        lk.unlock();
        for (auto& d : done)
          if (d.second == BNSAMD_OK) unstage(*d.first);
        lk.lock();
        for (auto& d : done) finish(*d.first, d.second);
      }
      form_batches(fresh);
      if (!fresh.empty()) {
        lk.unlock();
        std::vector<int> rcs;
        for (Slot* sl : fresh) {
          int rc = enqueue(*sl);
          if (rc != BNSAMD_OK) (void)hipStreamSynchronize(sl->stream);
          rcs.push_back(rc);
        }
        lk.lock();
        for (size_t i = 0; i < fresh.size(); ++i)
          if (rcs[i] != BNSAMD_OK) finish(*fresh[i], rcs[i]);
        continue;  // new arrivals may have come in meanwhile
      }
      const bool inflight =
          std::any_of(slots.begin(), slots.end(), [](const Slot& x) { return x.busy; });
      if (stop && pending.empty() && !inflight) return;
      if (!done.empty()) continue;
      if (inflight) {
        // poll the completion events; new submissions wake the loop early
        cv_work.wait_for(lk, std::chrono::microseconds(20));
      } else {
        cv_work.wait(lk, [&] { return stop || !pending.empty(); });
      }
    }
  }
};

extern "C" {

int bnsamd_server_create(const bnsamd_server_config* cfg, bnsamd_server** out) {
  if (!out || !cfg) return BNSAMD_ERR_INVALID;
  *out = nullptr;
  if (cfg->streams < 1 || cfg->streams > 16 || cfg->max_batch_points < 1 ||
      cfg->hbm_budget_bytes < batch_bytes(1) || cfg->chunk_points < 0 || !(cfg->starvation_ms >= 0.))
    return BNSAMD_ERR_INVALID;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return BNSAMD_ERR_NO_DEVICE;
  if (cfg->device < 0 || cfg->device >= ndev) return BNSAMD_ERR_INVALID;
  if (hipSetDevice(cfg->device) != hipSuccess) return BNSAMD_ERR_HIP;
  auto s = std::make_unique<bnsamd_server>();
  s->cfg = *cfg;
  s->chunk = cfg->chunk_points > 0 ? cfg->chunk_points : 32768;
  if (s->chunk > cfg->max_batch_points) s->chunk = cfg->max_batch_points;
  s->starve = std::chrono::duration_cast<Clock::duration>(std::chrono::duration<double, std::milli>(
      cfg->starvation_ms > 0. ? cfg->starvation_ms : 200.));
  s->gamma36 = bns::gamma_stirling_36();
  if (s->arena.init(cfg->hbm_budget_bytes) != hipSuccess) return BNSAMD_ERR_NOMEM;
  s->stats.arena_bytes = int64_t(s->arena.size());
  s->slots.resize(cfg->streams);
  for (Slot& sl : s->slots) {
    if (hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&sl.done, hipEventDisableTiming) != hipSuccess) {
      for (Slot& t : s->slots) {
        if (t.stream) (void)hipStreamDestroy(t.stream);
        if (t.done) (void)hipEventDestroy(t.done);
      }
      s->arena.release();
      return BNSAMD_ERR_HIP;
    }
  }
  bnsamd_server* raw = s.release();
  raw->worker = std::thread([raw] { raw->run(); });
  *out = raw;
  return BNSAMD_OK;
}

void bnsamd_server_destroy(bnsamd_server* s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> g(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  if (s->worker.joinable()) s->worker.join();
  (void)hipSetDevice(s->cfg.device);
  for (Slot& sl : s->slots) {
    (void)hipStreamSynchronize(sl.stream);
    (void)hipStreamDestroy(sl.stream);
    (void)hipEventDestroy(sl.done);
  }
  for (auto& kv : s->tables) (void)hipFree(kv.second.dev);
  if (s->hmirror) (void)hipHostFree(s->hmirror);
  s->arena.release();
  delete s;
}

const char* bnsamd_server_last_error(const bnsamd_server* s) { return s ? s->err.c_str() : ""; }

int bnsamd_server_submit(bnsamd_server* s, int64_t npts, const double* h_pts, int nquad,
                         const bnsamd_flags* flags, const bnsamd_params* params, int stim_abs,
                         int precision, double* h_out, int64_t* request_id) {
  if (!s || !request_id) return BNSAMD_ERR_INVALID;
  std::lock_guard<std::mutex> g(s->mu);
  auto bad = [&](const char* m) {
    s->err = m;
    return BNSAMD_ERR_INVALID;
  };
  if (npts < 1 || !h_pts || !h_out || !flags || !params) return bad("submit: invalid argument");
  if (nquad < 1 || nquad > bns::kMaxQuad) return bad("submit: quadrature order must be 1..32");
  if (precision != BNSAMD_FP64 && precision != BNSAMD_MIXED) return bad("submit: precision");
  if (stim_abs != 0 && stim_abs != 1) return bad("submit: stim_abs must be 0 or 1");
  if (params->brem_implementation < 0 || params->brem_implementation > 2)
    return bad("submit: unknown brem_implementation");
  if (npts > int64_t(0x7fffffff) || batch_bytes(std::min(npts, s->chunk)) > s->arena.size())
    return bad("submit: request piece larger than the server's HBM budget");
  if (s->stop) return bad("submit: server is stopping");
  auto r = std::make_unique<Request>();
  r->id = s->next_id++;
  r->npts = npts;
  r->h_pts = h_pts;
  r->h_out = h_out;
  r->pinned_in = is_pinned(h_pts);
  r->pinned_out = is_pinned(h_out);
  r->nquad = nquad;
  r->precision = precision;
  r->stim = stim_abs;
  r->flags = *flags;
  r->params = *params;
  r->t_submit = Clock::now();
  *request_id = r->id;
  for (int64_t off = 0; off < npts; off += s->chunk) {
    Piece pc;
    pc.r = r.get();
    pc.off = off;
    pc.n = std::min(s->chunk, npts - off);
    pc.t_queued = r->t_submit;
    s->pending.push_back(pc);
    r->pieces_left += 1;
  }
  s->requests.emplace(r->id, std::move(r));
  s->stats.requests_submitted += 1;
  s->cv_work.notify_one();
  return BNSAMD_OK;
}

int bnsamd_server_poll(bnsamd_server* s, int64_t request_id, int* done) {
  if (!s || !done) return BNSAMD_ERR_INVALID;
  std::lock_guard<std::mutex> g(s->mu);
  auto it = s->requests.find(request_id);
  if (it == s->requests.end()) return BNSAMD_ERR_INVALID;
  *done = it->second->done ? 1 : 0;
  return BNSAMD_OK;
}

int bnsamd_server_wait(bnsamd_server* s, int64_t request_id, double* latency_us) {
  if (!s) return BNSAMD_ERR_INVALID;
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->requests.find(request_id);
  if (it == s->requests.end()) return BNSAMD_ERR_INVALID;
  Request* r = it->second.get();
  s->cv_done.wait(lk, [&] { return r->done; });
  if (latency_us)
    *latency_us = std::chrono::duration<double, std::micro>(r->t_done - r->t_submit).count();
  int status = r->status;
  s->requests.erase(it);
  return status;
}

int bnsamd_server_get_stats(bnsamd_server* s, bnsamd_server_stats* out) {
  if (!s || !out) return BNSAMD_ERR_INVALID;
  std::lock_guard<std::mutex> g(s->mu);
  *out = s->stats;
  return BNSAMD_OK;
}

int bnsamd_host_alloc(size_t bytes, void** ptr) {
  if (!ptr) return BNSAMD_ERR_INVALID;
  *ptr = nullptr;
  return hipHostMalloc(ptr, bytes, hipHostMallocDefault) == hipSuccess ? BNSAMD_OK
                                                                       : BNSAMD_ERR_NOMEM;
}

void bnsamd_host_free(void* ptr) {
  if (ptr) (void)hipHostFree(ptr);
}

}  // extern "C"
