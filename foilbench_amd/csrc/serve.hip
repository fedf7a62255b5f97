// Request loop: many Stable Fluids and lattice scenarios in flight on one GPU.
//
// Stable Fluids requests run on a pool of worker threads, each with its own context (stream and
// reduction slots); a request's handle is moved onto the worker's context for its run, so the
// CG polling of one request never stalls another. Lattice requests form one continuous batch per
// precision: every lattice step of every member is one launch, replayed from a graph in
// chunks, and lattices join and leave at chunk boundaries.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>
#include <vector>

#include "fb_common.h"
#include "lbm_batch.h"
#include "mac.h"
#include "poisson.h"

namespace fb {

static int make_ctx(int device, Ctx** out) {
  FB_HIP(hipSetDevice(device));
  auto* c = new Ctx();
  c->device = device;
  hipDeviceProp_t prop;
  FB_HIP(hipGetDeviceProperties(&prop, device));
  c->cu_count = prop.multiProcessorCount;
  FB_HIP(stream_create(&c->stream));
  if (dev_alloc(&c->partials, (size_t)kPartialKeys * kPartialBlocks * sizeof(unsigned long long)) != hipSuccess) {
    (void)stream_free(c->stream);
    delete c;
    set_error("cannot allocate reduction slots");
    return FB_ERR_NOMEM;
  }
  *out = c;
  return FB_OK;
}

static void free_ctx(Ctx* c) {
  if (!c) return;
  (void)hipStreamSynchronize(c->stream);
  (void)dev_free(c->partials);
  (void)stream_free(c->stream);
  delete c;
}

struct SfJob {
  int64_t id;
  MacGrid* M;
  fb_sf_params p;
  fb_pose pose;
  int steps;
  double submitted;
};

struct LbmJob {
  int64_t id;
  fb_lbm* L;
  fb_lbm_params p;
  int steps;
  double submitted;
  int prec;
};

// ------------------------------------------------------------------ batched pressure solves

// Stable Fluids workers of one grid shape meet here for their pressure solves: a worker hands its
// grid over and waits; the batcher runs the waiting grids' solves in one batched solve (every CG
// launch covers all of them, each grid stopping on its own), then releases them.
struct SolveJob {
  MacGrid* M = nullptr;
  PoissonReport rep{};
  int rc = FB_OK;
  bool done = false;
};

struct SolveKey {
  int nx, ny, mixed, maxit;
  double dx, dy, tol;
  bool operator==(const SolveKey& o) const {
    return nx == o.nx && ny == o.ny && mixed == o.mixed && maxit == o.maxit && dx == o.dx && dy == o.dy &&
           tol == o.tol;
  }
};

struct Server;

struct Batcher {
  SolveKey key{};
  Server* S = nullptr;
  Ctx* ctx = nullptr;
  Poisson P;
  int capacity = 1;
  std::vector<unsigned long long> slot_ver;  // operator tag loaded in each slot
  std::mutex mu;
  std::condition_variable cv;
  std::vector<SolveJob*> pending;
  int members = 0;  // workers running a request of this shape (they may all join a solve)
  bool stop = false;
  std::thread thread;
};

struct LbmTiming {
  int64_t id;
  double submitted, started;
  int prec;
};

struct Server {
  int device = 0;
  fb_server_options opt{};
  std::chrono::steady_clock::time_point t0;
  std::mutex mu;
  std::condition_variable cv;
  std::deque<SfJob> sf_queue;
  std::deque<LbmJob> lbm_queue;
  std::deque<fb_server_result> results;
  int sf_running = 0, lbm_running = 0;
  bool stop = false;
  std::vector<Ctx*> sf_ctx;
  Ctx* lbm_ctx = nullptr;
  std::vector<std::thread> threads;
  std::mutex bmu;  // guards `batchers`
  std::vector<Batcher*> batchers;

  double now() const {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
};

static void batcher_loop(Batcher* B) {
  (void)hipSetDevice(B->S->device);
  const auto gather = std::chrono::duration<double, std::micro>(B->S->opt.sf_gather_us);
  for (;;) {
    std::vector<SolveJob*> jobs;
    {
      std::unique_lock<std::mutex> lk(B->mu);
      B->cv.wait(lk, [&] { return B->stop || !B->pending.empty(); });
      if (B->pending.empty()) return;
      // give the other workers of this shape a moment to arrive
      const auto t0 = std::chrono::steady_clock::now();
      B->cv.wait_until(lk, t0 + gather, [&] {
        return B->stop || (int)B->pending.size() >= std::min(B->members, B->capacity);
      });
      const int take = std::min((int)B->pending.size(), B->capacity);
      jobs.assign(B->pending.begin(), B->pending.begin() + take);
      B->pending.erase(B->pending.begin(), B->pending.begin() + take);
    }
    // stable slots: the same grids keep the same operator slots from one solve to the next
    std::sort(jobs.begin(), jobs.end(), [](const SolveJob* a, const SolveJob* b) { return a->M < b->M; });
    const int nb = (int)jobs.size();
    int rc = FB_OK;
    std::vector<const double*> rhs(nb);
    std::vector<double*> x(nb);
    std::vector<PoissonReport> reps(nb);
    for (int k = 0; k < nb && rc == FB_OK; ++k) {
      const Poisson& src = jobs[k]->M->poisson;
      if (B->slot_ver[k] != src.version) {
        rc = B->P.load_slot(k, src);
        B->slot_ver[k] = src.version;
      }
      rhs[k] = jobs[k]->M->rhs;
      x[k] = jobs[k]->M->pres;
    }
    if (rc == FB_OK)
      rc = B->P.solve_batch(nb, rhs.data(), x.data(), B->key.tol, B->key.maxit, B->key.mixed, reps.data());
    if (rc) {
      // the slots may hold partial copies: reload them next time
      for (auto& v : B->slot_ver) v = 0;
    }
    {
      std::lock_guard<std::mutex> lk(B->mu);
      for (int k = 0; k < nb; ++k) {
        jobs[k]->rc = rc;
        jobs[k]->rep = reps[k];
        jobs[k]->done = true;
      }
    }
    B->cv.notify_all();
  }
}

static Batcher* batcher_for(Server* S, const SolveKey& key, int* rc) {
  std::lock_guard<std::mutex> lk(S->bmu);
  for (auto* b : S->batchers)
    if (b->key == key) return b;
  auto* B = new Batcher();
  B->key = key;
  B->S = S;
  B->capacity = std::min(S->opt.sf_batch, kMaxBatch);
  *rc = make_ctx(S->device, &B->ctx);
  if (*rc == FB_OK)
    *rc = B->P.setup(B->ctx->stream, key.nx, key.ny, key.dx, key.dy, 0, 0, nullptr, B->capacity);
  B->P.set_use_graphs(S->opt.use_graphs != 0);
  if (*rc) {
    free_ctx(B->ctx);
    delete B;
    return nullptr;
  }
  B->slot_ver.assign(B->capacity, 0ull);
  B->thread = std::thread(batcher_loop, B);
  S->batchers.push_back(B);
  return B;
}

// SolveHook of a served grid: hand the solve to the grid's batcher and wait for it.
static int batched_solve(void* user, MacGrid& M, double tol, int maxit, int mixed, PoissonReport* rep) {
  auto* B = static_cast<Batcher*>(user);
  if (B->key.tol != tol || B->key.maxit != maxit || B->key.mixed != mixed) {
    // not the batch's solve: the grid's own solver, completed here as the hook requires
    int rc = M.poisson.solve(M.rhs, M.pres, tol, maxit, mixed, rep);
    if (rc) return rc;
    FB_HIP(hipStreamSynchronize(M.ctx->stream));
    M.poisson.finish(rep);
    return FB_OK;
  }
  // the right-hand side was written on the worker's stream; the batcher reads it on its own
  FB_HIP(hipStreamSynchronize(M.ctx->stream));
  SolveJob job;
  job.M = &M;
  {
    std::unique_lock<std::mutex> lk(B->mu);
    B->pending.push_back(&job);
    B->cv.notify_all();
    B->cv.wait(lk, [&] { return job.done; });
  }
  *rep = job.rep;
  return job.rc;
}

static void sf_worker(Server* S, int w) {
  (void)hipSetDevice(S->device);
  Ctx* c = S->sf_ctx[w];
  for (;;) {
    SfJob job;
    {
      std::unique_lock<std::mutex> lk(S->mu);
      S->cv.wait(lk, [&] { return S->stop || !S->sf_queue.empty(); });
      if (S->sf_queue.empty()) return;  // stopping with nothing left
      job = S->sf_queue.front();
      S->sf_queue.pop_front();
      S->sf_running++;
    }
    fb_server_result r{};
    r.id = job.id;
    r.kind = FB_SERVE_SF;
    r.submitted_ms = job.submitted;
    r.started_ms = S->now();
    MacGrid& M = *job.M;
    Ctx* home = M.ctx;  // its stream was synchronised at submission, by the submitting thread
    int rc = FB_OK;
    Batcher* B = nullptr;
    if (S->opt.sf_batch > 1 && !M.dom.periodic_x && !M.dom.periodic_y) {
      const int mode = job.p.mixed_poisson | (job.p.warm_pressure ? kWarmStart : 0);
      const SolveKey key{M.dom.nx, M.dom.ny, mode, job.p.max_iterations, M.g.dx, M.g.dy,
                         job.p.tolerance};
      B = batcher_for(S, key, &rc);
    }
    if (rc == FB_OK) {
      mac_bind(M, c);
      mac_use_graphs(M, S->opt.use_graphs != 0);
      if (B) {
        std::lock_guard<std::mutex> lk(B->mu);
        B->members++;
        M.hook.fn = batched_solve;
        M.hook.user = B;
      }
      for (int k = 0; k < job.steps; ++k) {
        fb_sf_report rep;
        rc = mac_substep(M, job.p, job.pose, &rep);
        if (rc) break;
        r.pressure_iterations += rep.pressure_iterations;
        if (rep.status) {
          r.sf_status = rep.status;
          break;
        }
        r.steps_done = k + 1;
      }
      if (hipStreamSynchronize(c->stream) != hipSuccess && rc == FB_OK) rc = FB_ERR_HIP;
      if (B) {
        std::lock_guard<std::mutex> lk(B->mu);
        B->members--;
        M.hook = SolveHook{};
        B->cv.notify_all();
      }
      mac_use_graphs(M, true);
      mac_bind(M, home);
    }
    r.status = rc;
    r.finished_ms = S->now();
    {
      std::lock_guard<std::mutex> lk(S->mu);
      S->sf_running--;
      S->results.push_back(r);
    }
    S->cv.notify_all();
  }
}

static void lbm_worker(Server* S) {
  (void)hipSetDevice(S->device);
  std::vector<LbmBatch*> batches(4, nullptr);
  std::vector<LbmTiming> timing;
  for (;;) {
    std::vector<LbmJob> joining;
    bool any = false;
    for (auto* b : batches) any = any || lbm_batch_members(b) > 0;
    {
      std::unique_lock<std::mutex> lk(S->mu);
      if (!any) S->cv.wait(lk, [&] { return S->stop || !S->lbm_queue.empty(); });
      if (S->stop && S->lbm_queue.empty() && !any) break;
      while (!S->lbm_queue.empty()) {
        joining.push_back(S->lbm_queue.front());
        S->lbm_queue.pop_front();
        S->lbm_running++;
      }
    }
    std::vector<fb_server_result> out;
    for (auto& j : joining) {
      int rc = FB_OK;
      if (!batches[j.prec]) rc = lbm_batch_create(S->lbm_ctx, j.prec, S->opt.lbm_chunk, S->opt.use_graphs, &batches[j.prec]);
      if (rc == FB_OK) rc = lbm_batch_add(batches[j.prec], j.L, j.p, j.steps, j.id);
      if (rc == FB_OK) {
        timing.push_back({j.id, j.submitted, S->now(), j.prec});
      } else {
        fb_server_result r{};
        r.id = j.id;
        r.kind = FB_SERVE_LBM;
        r.status = rc;
        r.submitted_ms = j.submitted;
        r.started_ms = r.finished_ms = S->now();
        out.push_back(r);
      }
    }
    for (auto* b : batches) {
      if (lbm_batch_members(b) == 0) continue;
      std::vector<LbmDone> done;
      int rc = lbm_batch_run_chunk(b, done);
      const double t = S->now();
      for (auto& d : done) {
        fb_server_result r{};
        r.id = d.id;
        r.kind = FB_SERVE_LBM;
        r.status = d.status;
        r.steps_done = d.steps;
        r.lbm_stats = d.stats;
        r.finished_ms = t;
        for (size_t k = 0; k < timing.size(); ++k)
          if (timing[k].id == d.id) {
            r.submitted_ms = timing[k].submitted;
            r.started_ms = timing[k].started;
            timing.erase(timing.begin() + (long)k);
            break;
          }
        out.push_back(r);
      }
      if (rc) {
        // a failed chunk leaves the batch in an unknown state: fail its members
        const int prec = lbm_batch_prec(b);
        fb_server_result r{};
        r.kind = FB_SERVE_LBM;
        r.status = rc;
        r.finished_ms = t;
        std::vector<LbmTiming> rest;
        for (auto& tm : timing) {
          if (tm.prec != prec) {
            rest.push_back(tm);
            continue;
          }
          r.id = tm.id;
          r.submitted_ms = tm.submitted;
          r.started_ms = tm.started;
          out.push_back(r);
        }
        timing.swap(rest);
        lbm_batch_destroy(b);
        for (auto& bb : batches)
          if (bb == b) bb = nullptr;
      }
    }
    if (!out.empty()) {
      {
        std::lock_guard<std::mutex> lk(S->mu);
        S->lbm_running -= (int)out.size();
        for (auto& r : out) S->results.push_back(r);
      }
      S->cv.notify_all();
    }
  }
  for (auto* b : batches) lbm_batch_destroy(b);
}

}  // namespace fb

using fb::Server;

extern "C" {

int fb_mem_info(int device, int64_t* free_bytes, int64_t* total_bytes) {
  FB_REQUIRE(free_bytes && total_bytes, FB_ERR_ARG, "NULL argument");
  FB_HIP(hipSetDevice(device));
  fb::release_cached_blocks();
  size_t f = 0, t = 0;
  FB_HIP(hipMemGetInfo(&f, &t));
  *free_bytes = (int64_t)f;
  *total_bytes = (int64_t)t;
  return FB_OK;
}

int fb_poisson_time_precond(int device, int nx, int ny, double dx, double dy, int precision, int grids, int reps,
                            double* ms) {
  FB_REQUIRE(ms && nx >= 2 && ny >= 2 && grids >= 1 && grids <= fb::kMaxBatch && reps >= 1, FB_ERR_ARG,
             "bad preconditioner timing request");
  FB_REQUIRE(precision == FB_F64 || precision == FB_F32, FB_ERR_ARG, "precision must be FB_F64 or FB_F32");
  fb::Ctx* c = nullptr;
  int rc = fb::make_ctx(device, &c);
  if (rc) return rc;
  {
    fb::Poisson P;
    rc = P.setup(c->stream, nx, ny, dx, dy, 0, 0, nullptr, grids);
    if (rc == FB_OK) rc = P.time_precond(precision == FB_F32, grids, reps, ms);
    (void)hipStreamSynchronize(c->stream);
  }
  fb::free_ctx(c);
  return rc;
}

int fb_server_create(int device, const fb_server_options* options, fb_server** out) {
  FB_REQUIRE(options && out, FB_ERR_ARG, "NULL argument");
  FB_REQUIRE(options->sf_workers >= 1 && options->sf_workers <= 64, FB_ERR_ARG, "sf_workers must be in [1, 64]");
  FB_REQUIRE(options->lbm_chunk >= 2 && options->lbm_chunk % 2 == 0 && options->lbm_chunk <= 1024, FB_ERR_ARG,
             "lbm_chunk must be even and in [2, 1024]");
  FB_REQUIRE(options->sf_batch >= 0 && options->sf_batch <= fb::kMaxBatch && options->sf_gather_us >= 0.0,
             FB_ERR_ARG, "sf_batch must be in [0, %d] and sf_gather_us non-negative", fb::kMaxBatch);
  int n = 0;
  FB_HIP(hipGetDeviceCount(&n));
  FB_REQUIRE(device >= 0 && device < n, FB_ERR_ARG, "device %d out of range (%d devices)", device, n);
  auto* S = new Server();
  S->device = device;
  S->opt = *options;
  S->t0 = std::chrono::steady_clock::now();
  int rc = FB_OK;
  for (int w = 0; w < options->sf_workers && rc == FB_OK; ++w) {
    fb::Ctx* c = nullptr;
    rc = fb::make_ctx(device, &c);
    if (rc == FB_OK) S->sf_ctx.push_back(c);
  }
  if (rc == FB_OK) rc = fb::make_ctx(device, &S->lbm_ctx);
  if (rc) {
    for (auto* c : S->sf_ctx) fb::free_ctx(c);
    fb::free_ctx(S->lbm_ctx);
    delete S;
    return rc;
  }
  for (int w = 0; w < options->sf_workers; ++w) S->threads.emplace_back(fb::sf_worker, S, w);
  S->threads.emplace_back(fb::lbm_worker, S);
  *out = reinterpret_cast<fb_server*>(S);
  return FB_OK;
}

int fb_server_destroy(fb_server* srv) {
  if (!srv) return FB_OK;
  auto* S = reinterpret_cast<Server*>(srv);
  {
    std::lock_guard<std::mutex> lk(S->mu);
    S->stop = true;
  }
  S->cv.notify_all();
  for (auto& t : S->threads) t.join();
  for (auto* B : S->batchers) {
    {
      std::lock_guard<std::mutex> lk(B->mu);
      B->stop = true;
    }
    B->cv.notify_all();
    B->thread.join();
    fb::free_ctx(B->ctx);
    delete B;
  }
  for (auto* c : S->sf_ctx) fb::free_ctx(c);
  fb::free_ctx(S->lbm_ctx);
  delete S;
  return FB_OK;
}

int fb_server_submit_sf(fb_server* srv, int64_t id, fb_mac* mac, const fb_sf_params* params, const fb_pose* pose,
                        int steps) {
  FB_REQUIRE(srv && mac && params && pose, FB_ERR_ARG, "NULL argument");
  FB_REQUIRE(steps >= 1, FB_ERR_ARG, "steps must be positive");
  FB_REQUIRE(params->max_iterations >= 1, FB_ERR_ARG, "max_iterations must be positive");
  FB_REQUIRE(params->mixed_poisson >= 0 && params->mixed_poisson <= 2, FB_ERR_ARG, "unknown pressure precision %d",
             (int)params->mixed_poisson);
  FB_REQUIRE(params->warm_pressure >= 0 && params->warm_pressure <= 2 &&
                 !(params->warm_pressure && params->mixed_poisson == 1),
             FB_ERR_ARG, "warm_pressure must be 0, 1 or 2, and 0 with mixed_poisson 1");
  FB_REQUIRE(params->face_transport >= 0 && params->face_transport <= 2, FB_ERR_ARG,
             "face_transport must be 0 (cell), 1 (face advection) or 2 (skew-RK2)");
  auto* S = reinterpret_cast<Server*>(srv);
  auto* M = reinterpret_cast<fb::MacGrid*>(mac);
  FB_REQUIRE(M->ctx->device == S->device, FB_ERR_ARG, "handle lives on another device");
  // work pending on the handle's own stream completes here, in the caller's thread: the worker
  // must not wait on a stream that the caller may be capturing into meanwhile
  FB_HIP(hipSetDevice(S->device));
  FB_HIP(hipStreamSynchronize(M->ctx->stream));
  {
    std::lock_guard<std::mutex> lk(S->mu);
    FB_REQUIRE(!S->stop, FB_ERR_STATE, "server is stopping");
    S->sf_queue.push_back(fb::SfJob{id, M, *params, *pose, steps, S->now()});
  }
  S->cv.notify_all();
  return FB_OK;
}

int fb_server_submit_lbm(fb_server* srv, int64_t id, fb_lbm* lbm, const fb_lbm_params* params, int steps) {
  FB_REQUIRE(srv && lbm && params, FB_ERR_ARG, "NULL argument");
  FB_REQUIRE(steps >= 1, FB_ERR_ARG, "steps must be positive");
  auto* S = reinterpret_cast<Server*>(srv);
  const int prec = fb::lbm_handle_precision(lbm);
  FB_REQUIRE(prec >= FB_F64 && prec <= FB_F16_MIXED, FB_ERR_ARG, "invalid lattice handle");
  FB_HIP(hipSetDevice(S->device));
  FB_HIP(fb::lbm_handle_sync(lbm));
  {
    std::lock_guard<std::mutex> lk(S->mu);
    FB_REQUIRE(!S->stop, FB_ERR_STATE, "server is stopping");
    S->lbm_queue.push_back(fb::LbmJob{id, lbm, *params, steps, S->now(), prec});
  }
  S->cv.notify_all();
  return FB_OK;
}

int fb_server_poll(fb_server* srv, fb_server_result* out, int max, int* count) {
  FB_REQUIRE(srv && count && (out || max == 0), FB_ERR_ARG, "NULL argument");
  auto* S = reinterpret_cast<Server*>(srv);
  std::lock_guard<std::mutex> lk(S->mu);
  int k = 0;
  while (k < max && !S->results.empty()) {
    out[k++] = S->results.front();
    S->results.pop_front();
  }
  *count = k;
  return FB_OK;
}

int fb_server_load(fb_server* srv, int* queued, int* running) {
  FB_REQUIRE(srv && queued && running, FB_ERR_ARG, "NULL argument");
  auto* S = reinterpret_cast<Server*>(srv);
  std::lock_guard<std::mutex> lk(S->mu);
  *queued = (int)(S->sf_queue.size() + S->lbm_queue.size());
  *running = S->sf_running + S->lbm_running;
  return FB_OK;
}

int fb_server_clock(fb_server* srv, double* ms) {
  FB_REQUIRE(srv && ms, FB_ERR_ARG, "NULL argument");
  *ms = reinterpret_cast<Server*>(srv)->now();
  return FB_OK;
}

}  // extern "C"
