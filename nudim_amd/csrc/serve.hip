// serve.hip -- continuous-batching job server over the job API.
//
// The reference runs one reconstruction per process (src/reconstruction.c
// main).  Here a server admits jobs of mixed size while others run: every
// job gets a slot thread and its own stream, so one job's host set-up
// (geometry and window tables) and transfers overlap other jobs' kernels,
// and small jobs' kernels overlap large ones'.  Admission follows the
// server's policy (arrival order, or least work with an aging bound) among
// the jobs that fit an HBM budget (nudim_estimate_bytes per job).  Pending
// jobs with the same parameters as the one admitted join it as one batch
// (nudim_batch_create, up to max_stack volumes, in counts that keep each
// volume's arithmetic that of its own job), so many small volumes of one
// geometry share every launch.  A job's arithmetic does not depend on what
// else runs: its result is the same alone and in any mix.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <memory>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "common.hpp"
#include "nudim.h"

namespace {

using clk = std::chrono::steady_clock;

struct ServerJob {
  long long id = 0;
  nudim_params p{};
  std::vector<double> proj;
  double *volume = nullptr, *rfactors = nullptr;
  size_t bytes = 0;
  double work = 0;  // np^2 nz passes
  double t_submit = 0, t_start = 0, t_end = 0;
  int status = NUDIM_OK;
  int stacked = 1;  // volumes of the batch the job ran in
  std::string error;
};

// the same job apart from its data: every parameter equal
bool same_params(const nudim_params &a, const nudim_params &b) {
  return a.nx == b.nx && a.nz == b.nz && a.np == b.np && a.gam_min == b.gam_min && a.gam_max == b.gam_max &&
         a.interval == b.interval && a.padangle == b.padangle && a.weight == b.weight && a.random == b.random &&
         a.seed == b.seed && a.config_height == b.config_height && a.nheight == b.nheight &&
         a.ncenter == b.ncenter && a.denoise == b.denoise && a.method == b.method && a.nhio == b.nhio &&
         a.ner == b.ner && a.precision == b.precision && a.device == b.device && a.nfft_n == b.nfft_n &&
         a.nfft_w == b.nfft_w && a.cg_form == b.cg_form && a.use_graph == b.use_graph && a.tv_full == b.tv_full &&
         a.diffusion_lambda == b.diffusion_lambda && std::memcmp(a.cu_mask, b.cu_mask, sizeof(a.cu_mask)) == 0;
}

}  // namespace

struct nudim_server {
  nudim_server_opts o{};
  clk::time_point epoch;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<ServerJob *> pending;
  std::deque<ServerJob *> done;
  std::vector<std::thread> slots;
  size_t used = 0;
  long long next_id = 1;
  int active = 0, active_small = 0;
  bool stop = false;

  double now() const { return std::chrono::duration<double>(clk::now() - epoch).count(); }

  bool small(const ServerJob *j) const { return partition && j->work <= o.small_work; }

  // the next job to admit among those that fit the free budget and a free
  // slot of their kind (caller holds mu): FIFO takes the first in arrival
  // order; SHORTEST takes the least work, except that a job waiting longer
  // than max_wait_s (oldest first) goes before it.  With the small-job
  // partition, small and other jobs have max_active slots each.  Pending
  // jobs with the same parameters (in arrival order, while they fit the
  // budget) join a small volume, up to max_stack volumes of kStackElems
  // Fourier elements in all and in a count for which batch_bit_exact holds;
  // the group takes one slot.
  //
  // Measured per-volume gain of a batch over one job at a time (FP32, TV,
  // K = 3 / 5 / 8): 2.7 / 3.9 / 5.1x at 64^3, 2.0 / 2.3 / 2.6x at 96^3,
  // 1.5x at 128^3 (np 128), 1.2x at 128^3 with np 192 and 1.0-1.1x at
  // 192^3, whose kernels fill the device alone; there a batch only makes
  // all its jobs finish with the last one.
  static constexpr double kStackVolume = 2097152.0;  // np^2 nz of a volume that may join a batch
  static constexpr double kStackElems = 8388608.0;   // np^2 nz of a whole batch
  std::vector<ServerJob *> take() {
    std::vector<ServerJob *> group;
    ServerJob *j = take_one();
    if (!j) return group;
    group.push_back(j);
    const double elems = (double)j->p.np * j->p.np * j->p.nz;
    const int cap = elems > kStackVolume ? 1 : std::min(o.max_stack, std::max(1, (int)(kStackElems / elems)));
    for (auto it = pending.begin(); it != pending.end() && (int)group.size() < cap;) {
      if (same_params((*it)->p, j->p) && used + (*it)->bytes <= (size_t)o.hbm_budget) {
        used += (*it)->bytes;
        group.push_back(*it);
        it = pending.erase(it);
      } else {
        ++it;
      }
    }
    // an even count at nz not a multiple of 64 gives its last job back, to
    // its place in arrival order
    if (!nudim::batch_bit_exact(j->p.nz, (int)group.size())) {
      ServerJob *back = group.back();
      group.pop_back();
      used -= back->bytes;
      auto at = std::find_if(pending.begin(), pending.end(), [&](const ServerJob *x) { return x->id > back->id; });
      pending.insert(at, back);
    }
    for (ServerJob *g : group) g->stacked = (int)group.size();
    return group;
  }

  ServerJob *take_one() {
    auto best = pending.end();
    const double t = now();
    const bool room_small = active_small < o.max_active, room_large = active - active_small < o.max_active;
    for (auto it = pending.begin(); it != pending.end(); ++it) {
      if (used + (*it)->bytes > (size_t)o.hbm_budget) continue;
      if (!(small(*it) ? room_small : room_large)) continue;
      if (o.policy != NUDIM_SERVE_SHORTEST || t - (*it)->t_submit > o.max_wait_s) {
        best = it;
        break;
      }
      if (best == pending.end() || (*it)->work < (*best)->work) best = it;
    }
    if (best == pending.end()) return nullptr;
    ServerJob *j = *best;
    pending.erase(best);
    used += j->bytes;
    active++;
    if (small(j)) active_small++;
    return j;
  }

  // CU masks of the small-job partition: the first small_cus CUs, and the rest
  unsigned small_mask[8] = {}, large_mask[8] = {};
  bool partition = false;

  void run(ServerJob *j) {
    if (partition) {
      bool own = false;
      for (unsigned w : j->p.cu_mask) own |= w != 0;
      if (!own) std::memcpy(j->p.cu_mask, j->work <= o.small_work ? small_mask : large_mask, sizeof(small_mask));
    }
    nudim_ctx *c = nullptr;
    auto check = [&](int rc) {
      if (rc != NUDIM_OK && j->status == NUDIM_OK) {
        j->status = rc;
        j->error = nudim_last_error();
      }
      return rc == NUDIM_OK;
    };
    if (check(nudim_create(&j->p, &c)) && check(nudim_load_projections(c, j->proj.data())) &&
        check(nudim_initial(c))) {
      const int npass = j->p.nhio + j->p.ner;
      std::vector<double> r(npass > 0 ? npass : 1);
      if (check(nudim_iterate(c, npass, r.data())) && check(nudim_get_volume(c, j->volume)) && j->rfactors)
        std::memcpy(j->rfactors, r.data(), sizeof(double) * npass);
    }
    nudim_destroy(c);
    std::vector<double>().swap(j->proj);
  }

  // a group of jobs with the same parameters as one batch context
  void run_batch(const std::vector<ServerJob *> &g) {
    ServerJob *j = g[0];
    if (partition) {
      bool own = false;
      for (unsigned w : j->p.cu_mask) own |= w != 0;
      if (!own) {
        const unsigned *m = j->work <= o.small_work ? small_mask : large_mask;
        for (ServerJob *x : g) std::memcpy(x->p.cu_mask, m, sizeof(small_mask));
      }
    }
    const int K = (int)g.size(), npass = j->p.nhio + j->p.ner;
    nudim_ctx *c = nullptr;
    int rc = nudim_batch_create(&j->p, K, &c);
    for (int k = 0; k < K && rc == NUDIM_OK; k++) rc = nudim_batch_load_projections(c, k, g[k]->proj.data());
    if (rc == NUDIM_OK) rc = nudim_initial(c);
    std::vector<double> r((size_t)std::max(npass, 1) * K);
    if (rc == NUDIM_OK) rc = nudim_iterate(c, npass, r.data());
    for (int k = 0; k < K && rc == NUDIM_OK; k++) rc = nudim_batch_get_volume(c, k, g[k]->volume);
    const std::string err = rc == NUDIM_OK ? std::string() : std::string(nudim_last_error());
    for (int k = 0; k < K; k++) {
      ServerJob *x = g[k];
      if (rc != NUDIM_OK) {
        x->status = rc;
        x->error = err;
      } else if (x->rfactors) {
        for (int i = 0; i < npass; i++) x->rfactors[i] = r[(size_t)i * K + k];
      }
      std::vector<double>().swap(x->proj);
    }
    nudim_destroy(c);
  }

  void slot_main() {
    for (;;) {
      std::vector<ServerJob *> g;
      {
        std::unique_lock<std::mutex> lk(mu);
        // a stopping server still runs what is pending
        cv_work.wait(lk, [&] { return !(g = take()).empty() || (stop && pending.empty()); });
        if (g.empty()) return;
      }
      const double t0 = now();
      for (ServerJob *x : g) x->t_start = t0;
      if (g.size() == 1)
        run(g[0]);
      else
        run_batch(g);
      const double t1 = now();
      {
        std::lock_guard<std::mutex> lk(mu);
        for (ServerJob *x : g) {
          x->t_end = t1;
          used -= x->bytes;
          done.push_back(x);
        }
        active--;
        if (small(g[0])) active_small--;
      }
      cv_done.notify_all();
      cv_work.notify_all();
    }
  }
};

extern "C" int nudim_server_create(const nudim_server_opts *o, nudim_server **out) {
  try {
    if (!o || !out || o->max_active < 1) {
      nudim::set_last_error("invalid server options");
      return NUDIM_EINVAL;
    }
    *out = nullptr;
    std::unique_ptr<nudim_server> s(new nudim_server());
    s->o = *o;
    if (s->o.max_wait_s <= 0) s->o.max_wait_s = 30.0;
    if (s->o.max_stack <= 0) s->o.max_stack = 8;
    if (s->o.policy != NUDIM_SERVE_FIFO && s->o.policy != NUDIM_SERVE_SHORTEST) {
      nudim::set_last_error("unknown server policy");
      return NUDIM_EINVAL;
    }
    if (s->o.hbm_budget <= 0) {
      if (hipSetDevice(o->device) != hipSuccess) {
        nudim::set_last_error("hipSetDevice failed");
        return NUDIM_EHIP;
      }
      size_t fr = 0, tot = 0;
      if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
        nudim::set_last_error("hipMemGetInfo failed");
        return NUDIM_EHIP;
      }
      s->o.hbm_budget = (long long)(0.9 * (double)fr);
    }
    if (s->o.small_cus > 0) {
      int ncu = 0;
      if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, o->device) != hipSuccess) {
        nudim::set_last_error("hipDeviceGetAttribute failed");
        return NUDIM_EHIP;
      }
      if (s->o.small_cus >= ncu || ncu > 256) {
        nudim::set_last_error("small_cus must be below the device's " + std::to_string(ncu) + " CUs");
        return NUDIM_EINVAL;
      }
      for (int i = 0; i < ncu; i++) (i < s->o.small_cus ? s->small_mask : s->large_mask)[i / 32] |= 1u << (i % 32);
      s->partition = true;
    }
    s->epoch = clk::now();
    const int nslots = o->max_active * (s->partition ? 2 : 1);
    try {
      for (int i = 0; i < nslots; i++) s->slots.emplace_back([p = s.get()] { p->slot_main(); });
    } catch (...) {
      // the slots already started hold the server: stop and join them first
      {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
      }
      s->cv_work.notify_all();
      for (auto &t : s->slots) t.join();
      throw;
    }
    *out = s.release();
    return NUDIM_OK;
  } catch (const std::exception &e) {
    nudim::set_last_error(e.what());
    return NUDIM_ENOMEM;
  }
}

extern "C" int nudim_server_submit(nudim_server *s, const nudim_params *p, const double *proj, double *volume,
                                   double *rfactors, long long *job_id) {
  try {
    if (!s || !p || !proj || !volume) {
      nudim::set_last_error("invalid argument");
      return NUDIM_EINVAL;
    }
    size_t bytes = 0;
    int rc = nudim_estimate_bytes(p, &bytes);
    if (rc != NUDIM_OK) return rc;
    if ((long long)bytes > s->o.hbm_budget) {
      nudim::set_last_error("job needs " + std::to_string(bytes) + " bytes of device memory, over the server budget");
      return NUDIM_EBUDGET;
    }
    // the tilt count the job reads (reconstruction.c:304)
    const int nmeas = (int)((p->gam_max - p->gam_min) / p->interval + 0.5) + 1;
    std::unique_ptr<ServerJob> j(new ServerJob());
    j->p = *p;
    // served jobs launch their passes directly: while one job's stream
    // captured a pass into a graph, other jobs' synchronous copies in their
    // own slot threads fail (the runtime refuses a legacy-stream operation
    // while a stream is capturing); replay gains nothing measurable here
    j->p.use_graph = 0;
    j->proj.assign(proj, proj + (size_t)nmeas * p->nz * p->nx);
    j->volume = volume;
    j->rfactors = rfactors;
    j->bytes = bytes;
    j->work = (double)p->np * p->np * p->nz * (double)std::max(1, p->nhio + p->ner);
    {
      std::lock_guard<std::mutex> lk(s->mu);
      if (s->stop) {
        nudim::set_last_error("server is shutting down");
        return NUDIM_ESTATE;
      }
      j->id = s->next_id++;
      j->t_submit = s->now();
      if (job_id) *job_id = j->id;
      s->pending.push_back(j.release());
    }
    s->cv_work.notify_all();
    return NUDIM_OK;
  } catch (const std::exception &e) {
    nudim::set_last_error(e.what());
    return NUDIM_ENOMEM;
  }
}

extern "C" int nudim_server_next(nudim_server *s, int timeout_ms, nudim_job_report *rep) {
  if (!s || !rep) {
    nudim::set_last_error("invalid argument");
    return NUDIM_EINVAL;
  }
  std::unique_lock<std::mutex> lk(s->mu);
  auto ready = [&] { return !s->done.empty(); };
  if (timeout_ms < 0) s->cv_done.wait(lk, ready);
  else if (!s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready)) return NUDIM_EAGAIN;
  ServerJob *j = s->done.front();
  s->done.pop_front();
  lk.unlock();
  std::memset(rep, 0, sizeof(*rep));
  rep->job_id = j->id;
  rep->status = j->status;
  rep->t_submit = j->t_submit;
  rep->t_start = j->t_start;
  rep->t_end = j->t_end;
  rep->device_bytes = (long long)j->bytes;
  rep->stacked = j->stacked;
  std::strncpy(rep->error, j->error.c_str(), sizeof(rep->error) - 1);
  delete j;
  return NUDIM_OK;
}

extern "C" int nudim_server_stats(nudim_server *s, int *pending, int *active, long long *bytes_in_use) {
  if (!s) {
    nudim::set_last_error("invalid argument");
    return NUDIM_EINVAL;
  }
  std::lock_guard<std::mutex> lk(s->mu);
  if (pending) *pending = (int)s->pending.size();
  if (active) *active = s->active;
  if (bytes_in_use) *bytes_in_use = (long long)s->used;
  return NUDIM_OK;
}

extern "C" void nudim_server_destroy(nudim_server *s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  for (auto &t : s->slots) t.join();
  for (ServerJob *j : s->pending) delete j;
  for (ServerJob *j : s->done) delete j;
  delete s;
}
