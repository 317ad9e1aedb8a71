// mscx_server: the serving form of the grid pipeline, with continuous
// batching.  No reference counterpart (the reference computes one grid per
// GridDataManager run); every result equals mscx_grid_run on the same grid,
// tiling and threshold.
//
// A scheduler thread admits queued requests against an HBM budget and, each
// round, launches the GPU stages of every newly admitted grid together (one
// launch per stage over all their leaf tiles).  The host phases of each grid
// (merge tree, simplification, the manifold pass) then run on a worker pool
// while the scheduler already admits and launches the next round, so work
// that arrives while earlier grids are in flight overlaps with them.  A
// computed grid stays resident under the request's key -- its tiles (owner
// images, codes) in HBM and its complexes as merged, before simplification,
// in HBM for grids whose merge tree runs on the GPU (on the host for small
// ones) -- until the budget needs the memory (least recently used first).  A request with a resident key skips the GPU stages and the merge
// up: it copies the merged complexes, simplifies at its own threshold, merges
// down and fills its manifolds from the resident owner images.
#include <cmath>
#include <condition_variable>
#include <functional>
#include <deque>
#include <list>
#include <map>
#include <thread>

#include "pipeline.hpp"

namespace mscx {

// Whether a grid's merge levels run their merges on several host threads.
// Small grids keep one thread each (many of them run at once on the
// workers); large ones, whose merges take hundreds of milliseconds, use the
// level parallelism the plain run uses.
static bool large_grid(const mscx_grid *G) { return (int64_t)G->sx * G->sy >= ((int64_t)1 << 22); }

namespace {

// A computed grid kept for repeat requests.
struct Resident {
  uint64_t key = 0;
  std::unique_ptr<mscx_grid> base;  // tiles (owned) + pieces after merge up
  size_t dev_bytes = 0;
  bool ready = false;
  bool failed = false;
  int users = 0;  // requests computing from it or waiting for it
  std::vector<uint64_t> waiters;
};

enum ReqState { RQ_QUEUED, RQ_RUNNING, RQ_DONE };

struct Request {
  uint64_t ticket = 0, key = 0;
  std::vector<float> data;
  int32_t sx = 0, sy = 0, lx = 0, ly = 0;
  double threshold = 0;
  uint32_t flags = 0;
  double t_arrive = 0, t_start = 0, t_done = 0;
  ReqState state = RQ_QUEUED;
  int status = MSCX_OK;
  std::string error;
  std::unique_ptr<mscx_grid> result;
  std::shared_ptr<Resident> res;
  size_t est_bytes = 0;
  bool repeat = false;
};

}  // namespace

}  // namespace mscx

using namespace mscx;

struct mscx_server {
  size_t budget = 0;
  // work limit of every request's simplification (complex.hpp, simplify):
  // a request whose cancellations would run far beyond its size fails with
  // MSCX_ELIMIT (or MSCX_ENOMEM) in seconds instead of holding a worker
  double work_limit = 128.0;
  int nworkers = 1;
  std::mutex mu;
  std::condition_variable cv_sched, cv_work, cv_done;
  bool stopping = false;       // no more admissions (set by destroy)
  bool workers_stop = false;   // workers leave once the task queue is empty
  uint64_t next_ticket = 1;
  std::map<uint64_t, std::unique_ptr<Request>> reqs;
  std::deque<uint64_t> queue;                        // admission order
  std::deque<std::function<void()>> work;            // host-phase tasks
  std::map<uint64_t, std::shared_ptr<Resident>> resident;
  std::list<uint64_t> lru;                           // resident keys, least recent first
  size_t reserved = 0;                               // admitted but not yet allocated estimates
  uint64_t n_done = 0, n_batches = 0, n_repeats = 0, n_evicted = 0;
  std::thread sched;
  std::vector<std::thread> workers;

  void touch(uint64_t key) {
    lru.remove(key);
    lru.push_back(key);
  }

  // Drops least recently used idle residents until `need` more bytes fit.
  // First only entries whose tiles nothing else holds (dropping them frees
  // their HBM now); then, if that is not enough, entries that results not
  // yet released still hold through their keep pointer (their HBM is freed
  // when the client releases those results, which a client waiting for this
  // request may only do once it has been admitted).  Called with mu held.
  void make_room(size_t need) {
    for (int pass = 0; pass < 2; ++pass)
      for (auto it = lru.begin(); it != lru.end() && device_pool().in_use() + reserved + need > budget;) {
        auto r = resident.find(*it);
        if (r != resident.end() && r->second->ready && r->second->users == 0 &&
            (pass == 1 || r->second.use_count() == 1)) {
          resident.erase(r);
          it = lru.erase(it);
          ++n_evicted;
        } else {
          ++it;
        }
      }
  }

  void finish(Request &R, int status, const std::string &err) {
    std::lock_guard<std::mutex> lk(mu);
    R.status = status;
    R.error = err;
    R.t_done = now_s();
    R.state = RQ_DONE;
    ++n_done;
    cv_done.notify_all();
  }

  // Host phases of one request whose merged complexes are ready in `src`
  // (pieces after merge up); tiles belong to the resident entry.
  void resimplify(Request &R, const std::shared_ptr<Resident> &res) {
    try {
      std::unique_ptr<mscx_grid> G(new mscx_grid);
      const mscx_grid &B = *res->base;
      G->sx = B.sx;
      G->sy = B.sy;
      G->leaves = B.leaves;
      G->tiles = B.tiles;
      G->owns_tiles = false;
      G->keep = res;
      G->threshold = R.threshold;
      G->work_limit = work_limit;
      G->flags = R.flags;
      copy_pieces(&B, G.get());
      grid_simplify_down(G.get(), large_grid(G.get()));
      std::vector<mscx_grid *> one{G.get()};
      grids_discs(one);
      R.result = std::move(G);
      release(res);
      finish(R, MSCX_OK, "");
    } catch (const std::exception &e) {
      release(res);
      int code = MSCX_ELOGIC;
      if (auto *er = dynamic_cast<const Error *>(&e)) code = er->code;
      finish(R, code, e.what());
    }
  }

  void release(const std::shared_ptr<Resident> &res) {
    std::lock_guard<std::mutex> lk(mu);
    res->users--;
    cv_sched.notify_all();
  }

  void post(std::function<void()> f) {
    {
      std::lock_guard<std::mutex> lk(mu);
      work.push_back(std::move(f));
    }
    cv_work.notify_one();
  }

  void worker_loop() {
    for (;;) {
      std::function<void()> f;
      {
        std::unique_lock<std::mutex> lk(mu);
        cv_work.wait(lk, [&] { return workers_stop || !work.empty(); });
        if (work.empty()) return;  // workers_stop: every posted task has been taken
        f = std::move(work.front());
        work.pop_front();
      }
      f();
    }
  }

  // One scheduling round: admit what fits, compute the new grids together,
  // hand their host phases to the workers.  Returns false when stopping.
  bool round() {
    std::vector<Request *> batch;
    {
      std::unique_lock<std::mutex> lk(mu);
      cv_sched.wait(lk, [&] { return stopping || !queue.empty(); });
      if (stopping && queue.empty()) return false;
      bool any_running = false;
      for (auto &kv : reqs)
        if (kv.second->state == RQ_RUNNING) any_running = true;
      for (auto it = queue.begin(); it != queue.end();) {
        Request &R = *reqs[*it];
        // repeat request: resident (or being computed) under its key
        auto rit = R.key ? resident.find(R.key) : resident.end();
        if (rit != resident.end() && !rit->second->failed) {
          std::shared_ptr<Resident> res = rit->second;
          res->users++;
          touch(R.key);
          R.state = RQ_RUNNING;
          R.t_start = now_s();
          R.repeat = true;
          ++n_repeats;
          if (res->ready) {
            Request *rp = &R;
            work.push_back([this, rp, res] { resimplify(*rp, res); });
            cv_work.notify_one();
          } else {
            res->waiters.push_back(R.ticket);
          }
          it = queue.erase(it);
          continue;
        }
        if (R.data.empty()) {
          R.state = RQ_DONE;
          R.status = MSCX_ESTATE;
          R.error = "the resident grid was evicted; submit the grid again";
          R.t_done = now_s();
          ++n_done;
          cv_done.notify_all();
          it = queue.erase(it);
          continue;
        }
        if (device_pool().in_use() + reserved + R.est_bytes > budget) make_room(R.est_bytes);
        const bool fits = device_pool().in_use() + reserved + R.est_bytes <= budget;
        if (!fits && (any_running || !batch.empty())) break;  // wait for memory to free up
        if (!fits && R.est_bytes > budget) {
          R.state = RQ_DONE;
          R.status = MSCX_ENOMEM;
          R.error = "grid does not fit the HBM budget";
          R.t_done = now_s();
          ++n_done;
          cv_done.notify_all();
          it = queue.erase(it);
          continue;
        }
        // admitted (when nothing else runs a grid may exceed the estimate)
        reserved += R.est_bytes;
        R.state = RQ_RUNNING;
        R.t_start = now_s();
        if (R.key) {
          auto res = std::make_shared<Resident>();
          res->key = R.key;
          res->users = 1;
          resident[R.key] = res;
          touch(R.key);
          R.res = res;
        }
        batch.push_back(&R);
        any_running = true;
        it = queue.erase(it);
      }
      if (!batch.empty()) ++n_batches;
    }
    if (batch.empty()) {
      // nothing admissible now: wait for a completion to free memory
      std::unique_lock<std::mutex> lk(mu);
      cv_sched.wait_for(lk, std::chrono::milliseconds(20));
      return true;
    }

    // prepare (upload) each grid; a bad grid fails alone
    std::vector<mscx_grid *> gs;
    std::vector<Request *> live;
    for (Request *R : batch) {
      std::unique_ptr<mscx_grid> G(new mscx_grid);
      try {
        prepare_grid(G.get(), R->data.data(), R->sx, R->sy, 0, R->lx, R->ly, R->threshold,
                     merge_policy(R->sx, R->sy, R->flags));
        G->work_limit = work_limit;
        R->result = std::move(G);
        gs.push_back(R->result.get());
        live.push_back(R);
      } catch (const std::exception &e) {
        fail_admitted(*R, e);
      }
    }
    try {
      grids_compute(gs);
    } catch (const std::exception &e) {
      for (Request *R : live) fail_admitted(*R, e);
      return true;
    }
    {
      std::lock_guard<std::mutex> lk(mu);
      for (Request *R : live) reserved -= std::min(reserved, R->est_bytes);
    }
    for (Request *R : live) {
      // the input copy is no longer needed once the tiles are computed
      std::vector<float>().swap(R->data);
      if (R->result->dev_input) {
        mscx_dev_free(R->result->dev_input);
        R->result->dev_input = nullptr;
      }
      post([this, R] { first_run(*R); });
    }
    return true;
  }

  void fail_admitted(Request &R, const std::exception &e) {
    {
      std::lock_guard<std::mutex> lk(mu);
      reserved -= std::min(reserved, R.est_bytes);
      if (R.res) {
        R.res->failed = true;
        R.res->users--;
        resident.erase(R.key);
        lru.remove(R.key);
        wake_waiters(*R.res, e.what());
      }
    }
    R.result.reset();
    int code = MSCX_ELOGIC;
    if (auto *er = dynamic_cast<const Error *>(&e)) code = er->code;
    finish(R, code, e.what());
  }

  // waiters of a resident that failed; called with mu held
  void wake_waiters(Resident &res, const std::string &err) {
    for (uint64_t t : res.waiters) {
      Request &W = *reqs[t];
      W.state = RQ_DONE;
      W.status = MSCX_ELOGIC;
      W.error = "the grid this request repeats failed: " + err;
      W.t_done = now_s();
      ++n_done;
      res.users--;
    }
    res.waiters.clear();
    cv_done.notify_all();
  }

  // Host phases of a newly computed grid; with a key, its tiles and merged
  // complexes become the resident entry and waiting repeats are released.
  void first_run(Request &R) {
    mscx_grid *G = R.result.get();
    try {
      grid_merge_up(G, large_grid(G));
      if (R.res) {
        std::unique_ptr<mscx_grid> base(new mscx_grid);
        base->sx = G->sx;
        base->sy = G->sy;
        base->leaves = G->leaves;
        base->tiles = G->tiles;  // the resident owns the tiles from now on
        copy_pieces(G, base.get());
        G->owns_tiles = false;
        G->keep = R.res;
        size_t bytes = device_tree_bytes(base.get());
        for (mscx_tile *t : G->tiles) bytes += t->fixed_bytes + (t->crit_block ? t->crit_block_bytes : 0);
        std::vector<uint64_t> waiters;
        {
          std::lock_guard<std::mutex> lk(mu);
          R.res->base = std::move(base);
          R.res->dev_bytes = bytes;
          R.res->ready = true;
          waiters.swap(R.res->waiters);
        }
        for (uint64_t t : waiters) {
          Request *W = nullptr;
          {
            std::lock_guard<std::mutex> lk(mu);
            W = reqs[t].get();
          }
          std::shared_ptr<Resident> res = R.res;
          post([this, W, res] { resimplify(*W, res); });
        }
      }
      grid_simplify_down(G, large_grid(G));
      std::vector<mscx_grid *> one{G};
      grids_discs(one);
      if (R.res) release(R.res);
      finish(R, MSCX_OK, "");
    } catch (const std::exception &e) {
      if (R.res) {
        std::lock_guard<std::mutex> lk(mu);
        if (!R.res->ready) {
          R.res->failed = true;
          resident.erase(R.key);
          lru.remove(R.key);
          wake_waiters(*R.res, e.what());
        }
        R.res->users--;
      }
      int code = MSCX_ELOGIC;
      if (auto *er = dynamic_cast<const Error *>(&e)) code = er->code;
      R.result.reset();
      finish(R, code, e.what());
    }
  }
};

extern "C" {

int mscx_server_create(size_t hbm_budget, int32_t host_threads_n, mscx_server **out) {
  return guarded([&] {
    MSCX_REQUIRE(out && hbm_budget > 0 && host_threads_n >= 0, MSCX_EINVAL, "bad argument");
    ensure_init();
    std::unique_ptr<mscx_server> S(new mscx_server);
    S->budget = hbm_budget;
    S->nworkers = host_threads_n > 0 ? host_threads_n : host_threads();
    mscx_server *s = S.get();
    for (int i = 0; i < s->nworkers; ++i) s->workers.emplace_back([s] { s->worker_loop(); });
    s->sched = std::thread([s] {
      while (s->round()) {
      }
    });
    *out = S.release();
  });
}

int mscx_server_set_work_limit(mscx_server *s, double factor) {
  return guarded([&] {
    MSCX_REQUIRE(s && factor >= 0 && !std::isnan(factor), MSCX_EINVAL, "bad argument");
    std::lock_guard<std::mutex> lk(s->mu);
    s->work_limit = factor;
  });
}

int mscx_server_submit(mscx_server *s, const float *data, int32_t sx, int32_t sy, int32_t lx, int32_t ly,
                       double threshold, uint64_t key, uint32_t flags, uint64_t *ticket) {
  return guarded([&] {
    MSCX_REQUIRE(s && ticket, MSCX_EINVAL, "null argument");
    MSCX_REQUIRE(!(flags & MSCX_RUN_DEVICE_INPUT), MSCX_EINVAL, "served grids are host input");
    MSCX_REQUIRE(threshold >= 0.0 && !std::isnan(threshold), MSCX_EINVAL, "threshold must be >= 0");
    std::unique_ptr<Request> R(new Request);
    R->key = key;
    R->sx = sx;
    R->sy = sy;
    R->lx = lx;
    R->ly = ly;
    R->threshold = threshold;
    R->flags = flags;
    R->est_bytes = grid_device_bytes(sx, sy, 0, lx, ly, true);
    // the grid is copied unless it is omitted for a key that is resident
    // now (should that entry be evicted before the request runs, the
    // request fails with MSCX_ESTATE)
    if (data) R->data.assign(data, data + (size_t)sx * sy);
    {
      std::lock_guard<std::mutex> lk(s->mu);
      if (!data) {
        auto it = key ? s->resident.find(key) : s->resident.end();
        MSCX_REQUIRE(it != s->resident.end() && !it->second->failed, MSCX_EINVAL,
                     "null input for a grid that is not resident");
      }
      R->t_arrive = now_s();
      R->ticket = s->next_ticket++;
      *ticket = R->ticket;
      s->queue.push_back(R->ticket);
      s->reqs[R->ticket] = std::move(R);
    }
    s->cv_sched.notify_all();
  });
}

int mscx_server_wait(mscx_server *s, uint64_t ticket, double timeout_s, const mscx_grid **out, int32_t *status,
                     double *latency_s) {
  return guarded([&] {
    MSCX_REQUIRE(s, MSCX_EINVAL, "null server");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    MSCX_REQUIRE(it != s->reqs.end(), MSCX_EINVAL, "unknown ticket");
    Request &R = *it->second;
    auto done = [&] { return R.state == RQ_DONE; };
    if (timeout_s < 0)
      s->cv_done.wait(lk, done);
    else
      s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), done);
    MSCX_REQUIRE(done(), MSCX_ESTATE, "request still running");
    if (out) *out = R.result.get();
    if (status) *status = R.status;
    if (latency_s) *latency_s = R.t_done - R.t_arrive;
  });
}

const char *mscx_server_error(mscx_server *s, uint64_t ticket) {
  if (!s) return "";
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->reqs.find(ticket);
  return it == s->reqs.end() ? "" : it->second->error.c_str();
}

int mscx_server_release(mscx_server *s, uint64_t ticket) {
  return guarded([&] {
    MSCX_REQUIRE(s, MSCX_EINVAL, "null server");
    std::unique_ptr<Request> R;
    {
      std::lock_guard<std::mutex> lk(s->mu);
      auto it = s->reqs.find(ticket);
      MSCX_REQUIRE(it != s->reqs.end(), MSCX_EINVAL, "unknown ticket");
      MSCX_REQUIRE(it->second->state == RQ_DONE, MSCX_ESTATE, "request still running");
      R = std::move(it->second);
      s->reqs.erase(it);
    }
    R.reset();  // frees the result; tiles go once their resident is gone too
    s->cv_sched.notify_all();
  });
}

int mscx_server_stats(mscx_server *s, uint64_t out[8]) {
  return guarded([&] {
    MSCX_REQUIRE(s && out, MSCX_EINVAL, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    out[0] = s->n_done;
    out[1] = s->n_batches;
    out[2] = s->n_repeats;
    out[3] = s->n_evicted;
    out[4] = s->resident.size();
    out[5] = device_pool().in_use();
    out[6] = s->budget;
    out[7] = s->queue.size();
  });
}

int mscx_server_destroy(mscx_server *s) {
  return guarded([&] {
    if (!s) return;
    {
      // requests still queued never run: they fail with MSCX_ESTATE
      std::lock_guard<std::mutex> lk(s->mu);
      s->stopping = true;
      for (uint64_t t : s->queue) {
        Request &R = *s->reqs[t];
        R.state = RQ_DONE;
        R.status = MSCX_ESTATE;
        R.error = "the server was destroyed before the request ran";
        R.t_done = now_s();
        ++s->n_done;
      }
      s->queue.clear();
      s->cv_done.notify_all();
    }
    s->cv_sched.notify_all();
    // the scheduler finishes its current round (posting that round's host
    // phases) and leaves; only then are the workers told to stop, so every
    // admitted request's tasks -- and the repeats they release -- still run
    if (s->sched.joinable()) s->sched.join();
    {
      std::lock_guard<std::mutex> lk(s->mu);
      s->workers_stop = true;
    }
    s->cv_work.notify_all();
    for (auto &t : s->workers) t.join();
    s->reqs.clear();
    s->resident.clear();
    delete s;
  });
}

}  // extern "C"
