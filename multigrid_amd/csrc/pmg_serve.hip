// pmg_serve.hip -- continuous-batching serving loop for Poisson-Boltzmann
// solves (no reference counterpart: the reference solves one problem per
// process, PBE_pmgc_class_test.cc).
//
// Requests (an atom list, a grid size and the iparm/rparm of the solve)
// queue up as they arrive.  One worker thread owns all device work:
//   - admission: a queued request is started as soon as a slot is free
//     (at most max_active solves in flight) and its context fits the
//     device-memory budget, first fit in arrival order; a request that has
//     been passed over for max_active later ones holds back everything
//     behind it until it fits, so a large request cannot starve behind a
//     stream of small ones (idle pooled contexts are evicted for it as the
//     running solves finish).  Contexts are
//     pooled by solve shape and reused, so a request whose shape was seen
//     before needs no allocation, no graph capture and no budget;
//   - every admitted solve runs on its context's own HIP stream: GPU input
//     assembly from the atoms, hierarchy build, then the solve session
//     (lin_begin / lin_enqueue / lin_poll), each V-cycle iteration replayed
//     from the context's captured HIP graph.  The worker polls all sessions
//     without blocking and enqueues the next iteration of whichever has
//     landed, so the work of different requests overlaps on the device;
//   - a finished solve restores the Dirichlet values and copies the
//     potential into the context's pinned buffer; a copier thread moves it
//     into the caller's array and hands the context back to the pool.
// Each request's arithmetic is exactly that of a solve on its own (same
// kernels, same launch shapes, deterministic reductions, its own buffers),
// so results do not depend on what else is in flight.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <new>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/pmg_mi355x.h"
#include "pmg_ctx.h"

using namespace pmg::drv;

namespace {

using Clock = std::chrono::steady_clock;

struct Request {
  int64_t id = 0;
  Params p{};
  int mode = PMG_MIXED;
  double h = 0.0, ionic = 0.0;
  std::vector<double> atoms;
  double *u_out = nullptr;
  Clock::time_point t_submit, t_done;
  int rc = PMG_OK, iters = 0;
  double relres = 0.0, solve_ms = 0.0, setup_ms = 0.0;
  bool done = false;
  pmg_ctx *ctx = nullptr;
  size_t bytes = 0;  // budget charged for a newly created context
  int passed = 0;    // later requests admitted while this one did not fit
  bool grouped = false;  // iterating in (or waiting for) a group
};

// Grouped iterations: the iterations of up to GMAX sessions run together
// on the group's stream (group_step), one launch per kernel for all of
// them; members join and leave between iterations.
struct Group {
  hipStream_t s = nullptr;
  hipEvent_t ev = nullptr;
  double *dn = nullptr, *hn = nullptr;  // stopping norms (device, pinned)
  std::vector<Request *> mem;
  int mode = PMG_MIXED;
  bool busy = false;
};

// Contexts are interchangeable when everything their buffers and captured
// graphs depend on matches.
bool same_shape(const Params &a, const Params &b) {
  return a.nx == b.nx && a.ny == b.ny && a.nz == b.nz && a.nlev == b.nlev &&
         a.nu1 == b.nu1 && a.nu2 == b.nu2 && a.mgkey == b.mgkey &&
         a.mgprol == b.mgprol && a.mgcoar == b.mgcoar &&
         a.mgdisc == b.mgdisc && a.mgsmoo == b.mgsmoo &&
         a.mgsolv == b.mgsolv && a.mode == b.mode;
}

struct Pooled {
  pmg_ctx *ctx;
  size_t bytes;
  int mode;  // precision it was budgeted for (PMG_MIXED adds FP32 copies)
  bool busy;
};

}  // namespace

struct pmg_server {
  int device = 0;
  size_t budget = 0, used = 0;
  int max_active = 1;
  int ngroups = 0;  // 0: every solve iterates on its own stream
  std::mutex mu;
  std::condition_variable cv_work, cv_done, cv_copy;
  std::deque<Request *> queue;
  std::unordered_map<int64_t, Request *> reqs;
  std::vector<Pooled> pool;  // every live context, busy or idle
  std::deque<Request *> copyq;
  std::vector<double> lat_ms;
  int64_t next_id = 1;
  bool stop = false, worker_done = false;
  long long completed = 0;
  int in_flight = 0, max_seen = 0;
  bool have_t0 = false;
  Clock::time_point t_first, t_last;
  std::thread worker, copier;
};

namespace {

// Under the lock: an idle pooled context of this shape, marked busy.
pmg_ctx *take_idle(pmg_server *S, const Params &p, int mode) {
  for (auto &e : S->pool)
    if (!e.busy && e.mode == mode && same_shape(e.ctx->p, p)) {
      e.busy = true;
      return e.ctx;
    }
  return nullptr;
}

void release_ctx(pmg_server *S, pmg_ctx *c) {  // under the lock
  for (auto &e : S->pool)
    if (e.ctx == c) e.busy = false;
}

// Under the lock: drop idle contexts (oldest first) until `need` more bytes
// fit; the contexts are returned for deletion outside the lock.
void evict_for(pmg_server *S, size_t need, std::vector<pmg_ctx *> &dead) {
  for (size_t q = 0; q < S->pool.size() && S->used + need > S->budget;) {
    if (!S->pool[q].busy) {
      S->used -= S->pool[q].bytes;
      dead.push_back(S->pool[q].ctx);
      S->pool.erase(S->pool.begin() + q);
    } else {
      q++;
    }
  }
}

void complete(pmg_server *S, Request *r) {  // under the lock
  r->t_done = Clock::now();
  r->done = true;
  S->in_flight--;
  if (r->rc == PMG_OK || r->rc == PMG_ITMAX) {
    S->completed++;
    S->lat_ms.push_back(
        std::chrono::duration<double, std::milli>(r->t_done - r->t_submit)
            .count());
    S->t_last = r->t_done;
  }
  S->cv_done.notify_all();
}

void copier_main(pmg_server *S) {
  std::unique_lock<std::mutex> lk(S->mu);
  for (;;) {
    // runs until the worker has finished: only then can no more arrive
    S->cv_copy.wait(lk, [&] { return S->worker_done || !S->copyq.empty(); });
    if (S->copyq.empty()) return;
    Request *r = S->copyq.front();
    S->copyq.pop_front();
    lk.unlock();
    const Params &p = r->ctx->p;
    std::memcpy(r->u_out, r->ctx->u_pin,
                sizeof(double) * (size_t)p.nx * p.ny * p.nz);
    lk.lock();
    release_ctx(S, r->ctx);
    r->ctx = nullptr;
    complete(S, r);
    S->cv_work.notify_all();
  }
}

// A group's iteration has landed: every member takes its stopping norm;
// the finished ones restore their boundary values on their own streams,
// after the iteration, and leave the group.
void group_land(Group &g) {
  std::vector<Request *> keep;
  for (size_t k = 0; k < g.mem.size(); k++) {
    Request *r = g.mem[k];
    pmg_ctx *c = r->ctx;
    c->hscal[S_RSNRM] = g.hn[k];
    c->ses.pending = false;
    if (ses_cycle_done(c)) {
      hipStreamWaitEvent(c->s, g.ev, 0);
      lin_finish(c, r->u_out ? c->u_pin : nullptr);
      r->grouped = false;
    } else {
      keep.push_back(r);
    }
  }
  g.mem.swap(keep);
  g.busy = false;
}

// Start the group's next iteration with the members it kept and the
// ready sessions, oldest first, of the same mode and smoothing counts (at
// most GMAX).  A ready session that does not match stops the admission
// (the group drains and the next empty group takes it), so that none waits
// behind a stream of later matching ones.
void group_launch(Group &g, std::deque<Request *> &ready) {
  while (!ready.empty() && (int)g.mem.size() < pmg::GMAX) {
    Request *r = ready.front();
    if (!g.mem.empty()) {
      const Params &p0 = g.mem[0]->ctx->p;
      if (r->mode != g.mode || r->ctx->p.nu1 != p0.nu1 ||
          r->ctx->p.nu2 != p0.nu2)
        break;
    }
    g.mode = r->mode;
    g.mem.push_back(r);
    ready.pop_front();
  }
  if (g.mem.empty()) return;
  std::vector<pmg_ctx *> cs;
  for (Request *r : g.mem) {
    cs.push_back(r->ctx);
    r->ctx->ses.pending = true;
  }
  group_step(cs.data(), (int)cs.size(), g.mode, g.s, g.dn);
  hipMemcpyAsync(g.hn, g.dn, sizeof(double) * cs.size(),
                 hipMemcpyDeviceToHost, g.s);
  hipEventRecord(g.ev, g.s);
  g.busy = true;
}

void worker_main(pmg_server *S) {
  hipSetDevice(S->device);
  std::vector<Group> groups(S->ngroups);
  for (Group &g : groups) {
    if (hipStreamCreateWithFlags(&g.s, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&g.ev, hipEventDisableTiming) != hipSuccess ||
        hipMalloc((void **)&g.dn, sizeof(double) * pmg::GMAX) != hipSuccess ||
        hipHostMalloc((void **)&g.hn, sizeof(double) * pmg::GMAX) !=
            hipSuccess) {
      groups.clear();  // iterate every solve on its own stream instead
      break;
    }
  }
  std::deque<Request *> ready;  // sessions waiting for a group
  std::vector<Request *> active;
  std::unique_lock<std::mutex> lk(S->mu);
  for (;;) {
    // admission: first fit in arrival order
    std::vector<Request *> admit;
    std::vector<pmg_ctx *> dead;
    Request *waiting = nullptr;  // first request of this pass that did not fit
    for (auto it = S->queue.begin();
         it != S->queue.end() &&
         (int)(active.size() + admit.size()) < S->max_active;) {
      Request *r = *it;
      pmg_ctx *c = take_idle(S, r->p, r->mode);
      if (!c) {
        const size_t need = ctx_bytes_estimate(r->p, r->mode);
        if (S->used + need > S->budget) evict_for(S, need, dead);
        if (S->used + need > S->budget) {
          if (r->passed >= S->max_active) break;  // aged: hold the rest
          if (!waiting) waiting = r;
          ++it;
          continue;
        }
        S->used += need;
        r->bytes = need;
      }
      if (waiting) waiting->passed++;
      r->ctx = c;
      admit.push_back(r);
      it = S->queue.erase(it);
    }
    if (S->stop && S->queue.empty() && active.empty() && admit.empty() &&
        S->in_flight == 0) {
      S->worker_done = true;
      S->cv_copy.notify_all();
      lk.unlock();
      for (Group &g : groups) {
        if (g.s) hipStreamSynchronize(g.s);
        if (g.ev) hipEventDestroy(g.ev);
        if (g.dn) hipFree(g.dn);
        if (g.hn) hipHostFree(g.hn);
        if (g.s) hipStreamDestroy(g.s);
      }
      lk.lock();
      break;
    }
    lk.unlock();
    for (pmg_ctx *c : dead) delete c;
    bool progressed = !admit.empty() || !dead.empty();
    // start the admitted requests
    for (Request *r : admit) {
      int rc = PMG_OK;
      bool fresh = false;
      if (!r->ctx) {
        rc = ctx_new(&r->ctx, r->p);
        fresh = rc == PMG_OK;
        // its iterations will run in groups: no per-context graph to capture
        if (fresh && !groups.empty() && group_ok(r->ctx))
          r->ctx->graphs = false;
      } else {
        r->ctx->p = r->p;  // stopping parameters of this request
      }
      if (!rc && r->u_out && !r->ctx->u_pin) {
        const size_t nf = (size_t)r->p.nx * r->p.ny * r->p.nz;
        if (hipHostMalloc((void **)&r->ctx->u_pin, nf * sizeof(double)) !=
            hipSuccess) {
          r->ctx->u_pin = nullptr;
          rc = PMG_ENOMEM;
        }
      }
      if (!rc) {
        setup_start(r->ctx);
        rc = pbe_enqueue(r->ctx, r->h, r->ionic, (int)(r->atoms.size() / 5),
                         r->atoms.data());
      }
      if (!rc) {
        setup_enqueue(r->ctx, true);
        rc = lin_begin(r->ctx, r->mode);
      }
      lk.lock();
      if (fresh) S->pool.push_back(Pooled{r->ctx, r->bytes, r->mode, true});
      if (rc) {
        if (r->ctx && fresh) {  // a context that failed to start is dropped
          hipStreamSynchronize(r->ctx->s);
          for (size_t q = 0; q < S->pool.size(); q++)
            if (S->pool[q].ctx == r->ctx) {
              S->used -= S->pool[q].bytes;
              S->pool.erase(S->pool.begin() + q);
              break;
            }
          lk.unlock();
          delete r->ctx;
          lk.lock();
        } else if (r->ctx) {
          hipStreamSynchronize(r->ctx->s);
          release_ctx(S, r->ctx);
        } else if (r->bytes) {
          S->used -= r->bytes;
        }
        r->ctx = nullptr;
        r->rc = rc;
        complete(S, r);
      } else {
        active.push_back(r);
        S->max_seen = std::max(S->max_seen, (int)active.size());
      }
      lk.unlock();
    }
    // grouped iterations that have landed
    for (Group &g : groups) {
      if (g.busy && hipEventQuery(g.ev) != hipErrorNotReady) {
        group_land(g);
        progressed = true;
      }
    }
    // advance every session by what has landed
    for (size_t q = 0; q < active.size();) {
      Request *r = active[q];
      pmg_ctx *c = r->ctx;
      if (r->grouped) {
        q++;
        continue;
      }
      if (c->ses.phase < 2) {
        if (lin_poll(c, false)) {
          lin_finish(c, r->u_out ? c->u_pin : nullptr);
          progressed = true;
        } else if (!c->ses.pending) {
          if (!groups.empty() && group_ok(c)) {
            r->grouped = true;  // its iterations run in a group from now
            ready.push_back(r);
          } else {
            lin_enqueue(c);
          }
          progressed = true;
        }
        q++;
        continue;
      }
      const int f = lin_finished(c, false);
      if (f == 0) {
        q++;
        continue;
      }
      r->rc = f < 0 ? PMG_EDEVICE : lin_end(c);
      r->iters = c->iters;
      r->relres = c->relres;
      r->solve_ms = c->solve_ms;
      r->setup_ms = c->setup_ms;
      active.erase(active.begin() + q);
      progressed = true;
      lk.lock();
      if (r->u_out && r->rc != PMG_EDEVICE) {
        S->copyq.push_back(r);
        S->cv_copy.notify_one();
      } else {
        release_ctx(S, c);
        r->ctx = nullptr;
        complete(S, r);
      }
      lk.unlock();
    }
    // the next iteration of every idle group, with the sessions ready
    for (Group &g : groups) {
      if (!g.busy && (!g.mem.empty() || !ready.empty())) {
        group_launch(g, ready);
        progressed = progressed || g.busy;
      }
    }
    lk.lock();
    if (!progressed) {
      // nothing landed: wait briefly (new submissions wake the worker)
      S->cv_work.wait_for(lk, std::chrono::microseconds(20));
    }
  }
}

}  // namespace

extern "C" {

int pmg_server_create(pmg_server **out, size_t mem_budget, int max_active) {
  return pmg_server_create_grouped(out, mem_budget, max_active, 0);
}

int pmg_server_create_grouped(pmg_server **out, size_t mem_budget,
                              int max_active, int groups) {
  if (!out || max_active < 1 || groups < 0 || groups > 16)
    return PMG_EBADPARM;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return PMG_EDEVICE;
  pmg_server *S = new (std::nothrow) pmg_server;
  if (!S) return PMG_ENOMEM;
  hipGetDevice(&S->device);
  if (mem_budget == 0) {
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
      delete S;
      return PMG_EDEVICE;
    }
    mem_budget = fr / 10 * 9;
  }
  S->budget = mem_budget;
  S->max_active = max_active;
  S->ngroups = groups;
  S->worker = std::thread(worker_main, S);
  S->copier = std::thread(copier_main, S);
  *out = S;
  return PMG_OK;
}

int pmg_server_submit(pmg_server *S, const int *iparm, const double *rparm,
                      int mode, double h, double ionic_strength, int natoms,
                      const double *atoms, double *u_out, int64_t *ticket) {
  if (!S || !iparm || !rparm || !ticket || natoms < 0 ||
      (natoms > 0 && !atoms) || (mode != PMG_FP64 && mode != PMG_MIXED) ||
      !(h > 0.0) || !(ionic_strength >= 0.0))
    return PMG_EBADPARM;
  Request *r = new (std::nothrow) Request;
  if (!r) return PMG_ENOMEM;
  int rc = decode(iparm, rparm, r->p);
  if (!rc && (r->p.nx != r->p.ny || r->p.nx != r->p.nz || r->p.mode != 0 ||
              r->p.istop > 1))
    rc = PMG_EBADPARM;  // cubic linear solves with a residual stopping test
  if (!rc && ctx_bytes_estimate(r->p, mode) > S->budget) rc = PMG_EBUSY;
  if (rc) {
    delete r;
    return rc;
  }
  r->mode = mode;
  r->h = h;
  r->ionic = ionic_strength;
  r->atoms.assign(atoms, atoms + 5 * (size_t)natoms);
  r->u_out = u_out;
  r->t_submit = Clock::now();
  std::lock_guard<std::mutex> lk(S->mu);
  if (S->stop) {
    delete r;
    return PMG_EBUSY;
  }
  r->id = S->next_id++;
  if (!S->have_t0) {
    S->have_t0 = true;
    S->t_first = r->t_submit;
  }
  S->reqs[r->id] = r;
  S->queue.push_back(r);
  S->in_flight++;
  *ticket = r->id;
  S->cv_work.notify_all();
  return PMG_OK;
}

int pmg_server_wait(pmg_server *S, int64_t ticket, int timeout_ms, int *iters,
                    double *relres, double *latency_ms) {
  if (!S) return PMG_EBADPARM;
  std::unique_lock<std::mutex> lk(S->mu);
  auto it = S->reqs.find(ticket);
  if (it == S->reqs.end()) return PMG_EBADPARM;
  Request *r = it->second;
  if (timeout_ms < 0) {
    S->cv_done.wait(lk, [&] { return r->done; });
  } else if (!S->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms),
                                  [&] { return r->done; })) {
    return PMG_EBUSY;  // not finished yet; the ticket stays valid
  }
  if (iters) *iters = r->iters;
  if (relres) *relres = r->relres;
  if (latency_ms)
    *latency_ms =
        std::chrono::duration<double, std::milli>(r->t_done - r->t_submit)
            .count();
  const int rc = r->rc;
  S->reqs.erase(it);
  delete r;
  return rc;
}

int pmg_server_stats(pmg_server *S, long long *completed, double *p50_ms,
                     double *p99_ms, double *solves_per_s, size_t *mem_used,
                     int *max_concurrent) {
  if (!S) return PMG_EBADPARM;
  std::lock_guard<std::mutex> lk(S->mu);
  std::vector<double> l = S->lat_ms;
  auto pct = [&](double q) -> double {
    if (l.empty()) return 0.0;
    const size_t k = std::min(l.size() - 1, (size_t)(q * (l.size() - 1) + 0.5));
    std::nth_element(l.begin(), l.begin() + k, l.end());
    return l[k];
  };
  if (completed) *completed = S->completed;
  if (p50_ms) *p50_ms = pct(0.50);
  if (p99_ms) *p99_ms = pct(0.99);
  if (solves_per_s) {
    const double s =
        S->completed
            ? std::chrono::duration<double>(S->t_last - S->t_first).count()
            : 0.0;
    *solves_per_s = s > 0.0 ? S->completed / s : 0.0;
  }
  if (mem_used) *mem_used = S->used;
  if (max_concurrent) *max_concurrent = S->max_seen;
  return PMG_OK;
}

int pmg_server_reset_stats(pmg_server *S) {
  if (!S) return PMG_EBADPARM;
  std::lock_guard<std::mutex> lk(S->mu);
  if (S->in_flight) return PMG_EBUSY;
  S->lat_ms.clear();
  S->completed = 0;
  S->have_t0 = false;
  S->max_seen = 0;
  return PMG_OK;
}

void pmg_server_destroy(pmg_server *S) {
  if (!S) return;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    S->stop = true;
    S->cv_work.notify_all();
  }
  S->worker.join();  // drains every submitted request first
  S->copier.join();
  for (auto &e : S->pool) delete e.ctx;
  for (auto &kv : S->reqs) delete kv.second;
  delete S;
}

}  // extern "C"
