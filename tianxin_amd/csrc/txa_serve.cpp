// Serving loop: a worker thread that admits assembly requests by HBM footprint
// and runs them concurrently, one HIP stream per slot (continuous batching).
//
// The reference evaluates one Newmark/implicit step per call of
// AssemblyEngine::evaluate (disc-fe/src/Panzer_AssemblyEngine_impl.hpp:67-129),
// with the time-integration coefficients handed in through
// AssemblyEngineInArgs alpha / beta (Panzer_AssemblyEngine_InArgs.hpp:95-106).
// Here a request names a resident mesh and carries its state; the server maps
// Newmark (beta_N, gamma_N, dt) to the seeds of the tangent
//   J = dR/du_{n+1} = K + gamma_N / (beta_N dt) C + 1 / (beta_N dt^2) M,
// i.e. beta = 1, alpha = gamma_N / (beta_N dt), gamma2 = 1 / (beta_N dt^2)
// (first-order problems: alpha = 1 / dt), and runs txa_assemble's kernels on
// the slot's stream.
//
// Memory: one arena of the configured budget is allocated at creation; each
// admitted request takes f [ndofs] and vals [nnz] from it (first fit, 256 B
// aligned) until the caller releases the request, and the assembly workspace
// (element blocks and residuals of the two-phase elasticity kernels) until its
// kernels complete.  A request is admitted when a slot is free and its whole
// footprint fits; the queue is first come, first served.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <thread>

#include "txa_internal.hpp"

namespace txa {

using Clock = std::chrono::steady_clock;

namespace {

constexpr size_t kAlign = 256;

inline size_t round_up(size_t b) { return (b + kAlign - 1) / kAlign * kAlign; }

// First-fit sub-allocator over one device allocation.
struct Arena {
  DevBuf<char> mem;
  size_t cap = 0, used = 0, peak = 0;
  std::map<size_t, size_t> free_;   // offset -> size
  void init(size_t bytes) {
    mem.alloc(bytes);
    cap = bytes;
    free_.clear();
    free_[0] = bytes;
  }
  bool take(size_t bytes, size_t *off) {
    bytes = round_up(bytes);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < bytes) continue;
      *off = it->first;
      const size_t rest = it->second - bytes, next = it->first + bytes;
      free_.erase(it);
      if (rest) free_[next] = rest;
      used += bytes;
      peak = std::max(peak, used);
      return true;
    }
    return false;
  }
  void give(size_t off, size_t bytes) {
    bytes = round_up(bytes);
    used -= bytes;
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
};

enum { QUEUED = 0, RUNNING = 1, DONE = 2, FAILED = 3 };

struct Request {
  int64_t id = 0;
  Mesh *mesh = nullptr;
  txa_serve_request r{};
  size_t off_f = 0, off_v = 0, off_w = 0, bytes_f = 0, bytes_v = 0, bytes_w = 0;
  bool has_mem = false, has_ws = false;
  int slot = -1;
  int status = QUEUED;
  int error_code = TXA_ERR_DEVICE;
  std::string error;
  Clock::time_point t_submit, t_start, t_done;
  hipEvent_t ev = nullptr;      // completion (after the Dirichlet report is copied back)
  hipEvent_t ready = nullptr;   // the caller's producer stream at submission
};

}  // namespace

struct Server {
  int device = 0;
  std::vector<int> mesh_ids;
  std::map<int, Mesh *> meshes;
  Arena arena;
  std::vector<hipStream_t> streams;
  std::vector<int64_t> slot_req;   // request running on each slot, -1 = free
  DevBuf<int> bc_flag;             // [slots] Dirichlet-LID-outside-the-map flag of the slot's request
  int *bc_flag_h = nullptr;        // pinned copy, written back on the slot stream before completion
  std::mutex mu;
  std::condition_variable cv;      // wakes the worker (new work, stop)
  std::condition_variable done_cv; // wakes waiters (a request finished)
  std::deque<int64_t> queue;
  std::map<int64_t, std::unique_ptr<Request>> reqs;
  std::vector<double> lat_ms;
  int64_t n_failed = 0, elements_done = 0, next_id = 1;
  int inflight = 0, peak_inflight = 0;
  bool have_first = false;
  Clock::time_point t_first, t_last;
  bool stop = false;
  std::thread worker;

  ~Server() {
    {
      std::lock_guard<std::mutex> lk(mu);
      stop = true;
    }
    cv.notify_all();
    if (worker.joinable()) worker.join();
    (void)hipSetDevice(device);
    for (hipStream_t s : streams) (void)hipStreamSynchronize(s);
    for (auto &kv : reqs) {
      if (kv.second->ev) (void)hipEventDestroy(kv.second->ev);
      if (kv.second->ready) (void)hipEventDestroy(kv.second->ready);
    }
    for (hipStream_t s : streams) (void)hipStreamDestroy(s);
    if (bc_flag_h) (void)hipHostFree(bc_flag_h);
  }

  double *fptr(const Request &q) { return reinterpret_cast<double *>(arena.mem.p + q.off_f); }
  double *vptr(const Request &q) { return q.bytes_v ? reinterpret_cast<double *>(arena.mem.p + q.off_v) : nullptr; }

  void release_ws(Request &q) {
    if (!q.has_ws) return;
    arena.give(q.off_w, q.bytes_w);
    q.has_ws = false;
  }

  void release_mem(Request &q) {
    release_ws(q);
    if (!q.has_mem) return;
    arena.give(q.off_f, q.bytes_f);
    if (q.bytes_v) arena.give(q.off_v, q.bytes_v);
    q.has_mem = false;
  }

  // Launches q on slot s (lock held).  Errors mark the request failed.
  void launch(Request &q, int s) {
    const txa_serve_request &r = q.r;
    txa_assembly_args a{};
    a.physics = r.physics;
    a.precision = r.precision;
    a.coef = r.coef;
    a.nu = r.nu;
    a.rho = r.rho;
    for (int i = 0; i < 3; ++i) a.src[i] = r.src[i];
    a.rayleigh_a = r.rayleigh_a;
    a.rayleigh_b = r.rayleigh_b;
    a.x = r.u;
    a.xdot = r.v;
    a.xddot = r.a;
    a.beta = 1.0;
    if (r.physics == TXA_ELASTICITY) {
      a.alpha = r.newmark_gamma / (r.newmark_beta * r.dt);
      a.gamma2 = 1.0 / (r.newmark_beta * r.dt * r.dt);
    } else {
      a.alpha = 1.0 / r.dt;
      a.gamma2 = 0.0;
    }
    a.transient = 1;
    a.f = fptr(q);
    a.vals = r.want_jacobian ? vptr(q) : nullptr;
    a.bc_lids = r.bc_lids;
    a.bc_vals = r.bc_vals;
    a.n_bc = r.n_bc;
    q.slot = s;
    q.t_start = Clock::now();
    try {
      // after the caller's producer stream; the slot's Dirichlet flag cleared, reported back
      TXA_HIP(hipStreamWaitEvent(streams[s], q.ready, 0));
      TXA_HIP(hipMemsetAsync(bc_flag.p + s, 0, sizeof(int), streams[s]));
      assemble(*q.mesh, a, streams[s], q.has_ws ? arena.mem.p + q.off_w : nullptr, bc_flag.p + s);
      TXA_HIP(hipMemcpyAsync(bc_flag_h + s, bc_flag.p + s, sizeof(int), hipMemcpyDeviceToHost, streams[s]));
      TXA_HIP(hipEventRecord(q.ev, streams[s]));
      q.status = RUNNING;
      slot_req[s] = q.id;
      ++inflight;
      peak_inflight = std::max(peak_inflight, inflight);
    } catch (const std::exception &e) {
      q.status = FAILED;
      q.error = e.what();
      q.t_done = Clock::now();
      ++n_failed;
      release_mem(q);
      done_cv.notify_all();
    }
  }

  void run() {
    (void)hipSetDevice(device);
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
      // completions
      for (size_t s = 0; s < slot_req.size(); ++s) {
        if (slot_req[s] < 0) continue;
        Request &q = *reqs[slot_req[s]];
        const hipError_t e = hipEventQuery(q.ev);
        if (e == hipErrorNotReady) continue;
        q.t_done = Clock::now();
        release_ws(q);
        if (e == hipSuccess && bc_flag_h[s]) {
          q.status = FAILED;
          q.error_code = TXA_ERR_INVALID;
          q.error = "a Dirichlet LID of the request lies outside the DOF map (its row was not applied)";
          ++n_failed;
          release_mem(q);
        } else if (e == hipSuccess) {
          q.status = DONE;
          lat_ms.push_back(std::chrono::duration<double, std::milli>(q.t_done - q.t_submit).count());
          elements_done += q.mesh->nelem;
          t_last = q.t_done;
        } else {
          q.status = FAILED;
          q.error = std::string("device error: ") + hipGetErrorString(e);
          ++n_failed;
          release_mem(q);
        }
        slot_req[s] = -1;
        --inflight;
        done_cv.notify_all();
      }
      // admission, first come first served
      while (!queue.empty() && !stop) {
        auto fs = std::find(slot_req.begin(), slot_req.end(), int64_t(-1));
        if (fs == slot_req.end()) break;
        Request &q = *reqs[queue.front()];
        size_t of = 0, ov = 0, ow = 0;
        if (!arena.take(q.bytes_f, &of)) break;
        if (q.bytes_v && !arena.take(q.bytes_v, &ov)) {
          arena.give(of, q.bytes_f);
          break;
        }
        if (q.bytes_w && !arena.take(q.bytes_w, &ow)) {
          arena.give(of, q.bytes_f);
          if (q.bytes_v) arena.give(ov, q.bytes_v);
          break;
        }
        q.off_f = of;
        q.off_v = ov;
        q.off_w = ow;
        q.has_mem = true;
        q.has_ws = q.bytes_w != 0;
        queue.pop_front();
        launch(q, static_cast<int>(fs - slot_req.begin()));
      }
      if (stop && inflight == 0) break;
      if (inflight == 0 && (queue.empty() || stop)) {
        cv.wait(lk);
      } else {
        lk.unlock();
        std::this_thread::sleep_for(std::chrono::microseconds(20));
        lk.lock();
      }
    }
    // requests still queued at shutdown fail
    for (int64_t id : queue) {
      Request &q = *reqs[id];
      q.status = FAILED;
      q.error = "server stopped";
    }
    queue.clear();
    done_cv.notify_all();
  }
};

Server *serve_create(int device, const std::vector<Mesh *> &meshes, const std::vector<int> &ids, int64_t budget,
                     int slots) {
  TXA_REQUIRE(!meshes.empty(), TXA_ERR_INVALID, "a server needs at least one resident mesh");
  TXA_REQUIRE(slots >= 1 && slots <= 32, TXA_ERR_INVALID, "slots must be in [1, 32]");
  TXA_REQUIRE(budget >= static_cast<int64_t>(kAlign), TXA_ERR_INVALID, "HBM budget too small");
  for (Mesh *m : meshes) TXA_REQUIRE(m->has_graph, TXA_ERR_LOGIC, "served meshes need a DOF map and a graph");
  std::unique_ptr<Server> s(new Server());
  s->device = device;
  s->mesh_ids = ids;
  for (size_t i = 0; i < ids.size(); ++i) s->meshes[ids[i]] = meshes[i];
  s->arena.init(static_cast<size_t>(budget));
  s->streams.resize(slots);
  for (auto &st : s->streams) TXA_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  s->bc_flag.alloc(slots);
  TXA_HIP(hipMemset(s->bc_flag.p, 0, sizeof(int) * slots));
  TXA_HIP(hipHostMalloc(reinterpret_cast<void **>(&s->bc_flag_h), sizeof(int) * slots, hipHostMallocDefault));
  for (int i = 0; i < slots; ++i) s->bc_flag_h[i] = 0;
  s->slot_req.assign(slots, -1);
  Server *raw = s.get();
  s->worker = std::thread([raw] { raw->run(); });
  return s.release();
}

void serve_free(Server *s) { delete s; }

const std::vector<int> &serve_mesh_ids(const Server &s) { return s.mesh_ids; }

int64_t serve_submit(Server &s, const txa_serve_request &r) {
  auto it = s.meshes.find(r.mesh_id);
  TXA_REQUIRE(it != s.meshes.end(), TXA_ERR_INVALID, "mesh " + std::to_string(r.mesh_id) + " is not served here");
  Mesh &m = *it->second;
  TXA_REQUIRE(r.physics == TXA_DIFFUSION || r.physics == TXA_ELASTICITY, TXA_ERR_INVALID, "unknown physics");
  TXA_REQUIRE(m.nfields == (r.physics == TXA_ELASTICITY ? 3 : 1), TXA_ERR_INVALID, "DOF map does not fit the physics");
  TXA_REQUIRE(r.u && r.v && r.coef, TXA_ERR_INVALID, "u, v and coef are required");
  TXA_REQUIRE(r.dt > 0.0, TXA_ERR_INVALID, "dt must be positive");
  TXA_REQUIRE(r.physics != TXA_ELASTICITY || (r.newmark_beta > 0.0 && r.a), TXA_ERR_INVALID,
              "elasticity needs newmark_beta > 0 and an acceleration state");
  TXA_REQUIRE(r.n_bc == 0 || r.bc_lids, TXA_ERR_INVALID, "bc_lids missing");
  auto q = std::make_unique<Request>();
  q->mesh = &m;
  q->r = r;
  q->bytes_f = sizeof(double) * static_cast<size_t>(m.ndofs);
  q->bytes_v = r.want_jacobian ? sizeof(double) * static_cast<size_t>(m.nnz) : 0;
  q->bytes_w = assemble_workspace_bytes(m, r.physics, r.precision, true, r.want_jacobian != 0);
  TXA_REQUIRE(round_up(q->bytes_f) + round_up(q->bytes_v) + round_up(q->bytes_w) <= s.arena.cap, TXA_ERR_NOMEM,
              "request footprint exceeds the server's HBM budget");
  TXA_HIP(hipEventCreateWithFlags(&q->ev, hipEventDisableTiming));
  TXA_HIP(hipEventCreateWithFlags(&q->ready, hipEventDisableTiming));
  TXA_HIP(hipEventRecord(q->ready, static_cast<hipStream_t>(r.stream)));
  std::lock_guard<std::mutex> lk(s.mu);
  TXA_REQUIRE(!s.stop, TXA_ERR_LOGIC, "server is stopping");
  q->id = s.next_id++;
  q->t_submit = Clock::now();
  if (!s.have_first) {
    s.have_first = true;
    s.t_first = q->t_submit;
  }
  const int64_t id = q->id;
  s.reqs[id] = std::move(q);
  s.queue.push_back(id);
  s.cv.notify_all();
  return id;
}

namespace {
Request &find_req(Server &s, int64_t id) {
  auto it = s.reqs.find(id);
  TXA_REQUIRE(it != s.reqs.end(), TXA_ERR_INVALID, "unknown request id " + std::to_string(id));
  return *it->second;
}
}  // namespace

// status out; f / vals (device) valid while the request is done and unreleased
void serve_poll(Server &s, int64_t id, double timeout_s, int *status, const double **f, const double **vals) {
  std::unique_lock<std::mutex> lk(s.mu);
  Request &q = find_req(s, id);
  if (timeout_s > 0.0) {
    s.done_cv.wait_for(lk, std::chrono::duration<double>(timeout_s),
                       [&] { return q.status == DONE || q.status == FAILED; });
  }
  if (status) *status = q.status;
  if (q.status == FAILED) throw Error(q.error_code, "request " + std::to_string(id) + " failed: " + q.error);
  if (f) *f = q.status == DONE ? s.fptr(q) : nullptr;
  if (vals) *vals = q.status == DONE ? s.vptr(q) : nullptr;
}

void serve_release(Server &s, int64_t id) {
  std::lock_guard<std::mutex> lk(s.mu);
  Request &q = find_req(s, id);
  TXA_REQUIRE(q.status == DONE || q.status == FAILED, TXA_ERR_LOGIC, "request still queued or running");
  s.release_mem(q);
  if (q.ev) (void)hipEventDestroy(q.ev);
  if (q.ready) (void)hipEventDestroy(q.ready);
  s.reqs.erase(id);
  s.cv.notify_all();
}

// out: completed, failed, p50 ms, p99 ms, mean ms, max ms, elements/s, requests/s,
//      peak requests in flight, peak arena bytes, arena bytes in use
void serve_stats(Server &s, double *out) {
  std::lock_guard<std::mutex> lk(s.mu);
  std::vector<double> l = s.lat_ms;
  std::sort(l.begin(), l.end());
  auto pct = [&](double p) {
    if (l.empty()) return 0.0;
    const double pos = p * (l.size() - 1);
    const size_t i = static_cast<size_t>(pos);
    const double fr = pos - i;
    return i + 1 < l.size() ? l[i] * (1 - fr) + l[i + 1] * fr : l[i];
  };
  double mean = 0.0;
  for (double v : l) mean += v;
  mean = l.empty() ? 0.0 : mean / l.size();
  const double span = s.have_first && !l.empty() ? std::chrono::duration<double>(s.t_last - s.t_first).count() : 0.0;
  out[0] = static_cast<double>(l.size());
  out[1] = static_cast<double>(s.n_failed);
  out[2] = pct(0.50);
  out[3] = pct(0.99);
  out[4] = mean;
  out[5] = l.empty() ? 0.0 : l.back();
  out[6] = span > 0 ? s.elements_done / span : 0.0;
  out[7] = span > 0 ? l.size() / span : 0.0;
  out[8] = s.peak_inflight;
  out[9] = static_cast<double>(s.arena.peak);
  out[10] = static_cast<double>(s.arena.used);
}

}  // namespace txa
