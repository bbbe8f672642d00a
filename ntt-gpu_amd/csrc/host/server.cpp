// Continuous-batching server (include/ntt/serve.hpp, C ABI include/ntt_serve.h).
//
// One worker thread owns every device call. Each loop iteration it
//   1. retires finished batches (event query; results copied out of the
//      slot's pinned staging, tickets released),
//   2. fills every idle stream slot with a new batch: the oldest pending
//      request chooses the group and compatible pending requests join it, up
//      to max_batch and as far as the device budget allows,
//   3. sleeps briefly only when it neither retired nor launched anything.
// Groups: NTT/INTT requests of one ring dimension share a launch whatever
// their level or parameter set (a per-N context over every registered Q
// prime and a tower job table); KeySwitch, ModUp and ModDown requests share
// one when they have the same parameter set and level.
// Admission counts the whole device footprint against the budget: the slot
// arenas of the batches in flight plus every resident plan (contexts with
// their twiddle tables and cached conversion plans, key-switch plans with
// their evaluation keys, per-N pools). A batch that does not fit first frees
// idle slot arenas, then evicts idle plans least recently used first (an
// evicted plan is rebuilt and its key uploaded again on next use), and
// otherwise waits for batches in flight to finish.
#include <algorithm>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <functional>
#include <map>
#include <mutex>
#include <set>
#include <stdexcept>
#include <string>
#include <thread>
#include <unordered_map>

#include "ntt/serve.hpp"
#include "ntt_hip.h"
#include "ntt_serve.h"

namespace ntt::serve {
namespace {
using Clock = std::chrono::steady_clock;

void ok(int status, const char *what) {
  if (status == NTT_HIP_OK) return;
  const std::string msg = std::string(what) + ": " + ntt_hip_last_error();
  if (status == NTT_HIP_EINVAL) throw std::invalid_argument(msg);
  throw std::runtime_error("HIP: " + msg);
}

struct Request {
  uint64_t ticket = 0;
  uint32_t params = 0;
  Op op = Op::Forward;
  size_t level = 0;
  uint64_t n = 0;
  std::vector<uint64_t> in;
  bool device = false;  // device-resident: the caller's buffers below
  const uint64_t *d_in = nullptr;
  uint64_t *d_out = nullptr, *d_out2 = nullptr;
  Clock::time_point t0;
  Result result;
  std::string error;
  bool done = false;
  bool claimed = false;  // a wait() has taken this ticket (a second one is an error)
};

// Something resident on the device on behalf of many requests.
struct Resident {
  size_t device_bytes = 0;
  uint64_t last_use = 0;
  size_t inflight = 0;  // batches in flight using it (never evicted while > 0)
  virtual ~Resident() = default;
};

// (parameter set, level): context over active Q then P, key-switch plan, key.
struct LevelPlan : Resident {
  ntt_hip_context *ctx = nullptr;
  ntt_hip_keyswitch *ks = nullptr;
  ~LevelPlan() override {
    if (ks) ntt_hip_keyswitch_destroy(ks);
    if (ctx) ntt_hip_context_destroy(ctx);
  }
};

// Ring dimension n: one context over every Q prime registered for n (the
// union over parameter sets), so NTT/INTT requests of any level share it.
struct Pool : Resident {
  ntt_hip_context *ctx = nullptr;
  std::map<uint64_t, uint32_t> tower_of;  // prime -> context tower
  ~Pool() override {
    if (ctx) ntt_hip_context_destroy(ctx);
  }
};

struct Slot {
  void *stream = nullptr, *event = nullptr;
  void *device = nullptr;  // grow-only device arena
  size_t device_bytes = 0;
  void *pinned = nullptr;  // grow-only pinned host staging
  size_t pinned_bytes = 0;
  std::vector<std::shared_ptr<Request>> batch;
  std::vector<size_t> out_at, out_words, out2_at;  // per request, words into the staged outputs
  std::vector<std::shared_ptr<Resident>> holders;  // the plan or pools the batch runs on
  bool busy = false;
};

}  // namespace

struct Server::Impl {
  Options opt;
  std::vector<std::shared_ptr<const ParamSet>> params;  // shared: the worker keeps its copy
  std::map<std::pair<uint32_t, size_t>, std::shared_ptr<LevelPlan>> plans;
  std::map<uint64_t, std::shared_ptr<Pool>> pools;  // by ring dimension
  std::map<uint64_t, size_t> pool_sets;             // parameter sets a pool of n covers
  std::vector<Slot> slots;
  mutable std::mutex mu;
  std::condition_variable cv_submit, cv_done;
  std::deque<std::shared_ptr<Request>> pending;
  std::unordered_map<uint64_t, std::shared_ptr<Request>> tickets;
  uint64_t next_ticket = 1, tick = 0;
  bool stop = false;
  Stats stats;
  size_t arena_bytes = 0;     // sum of slot arenas
  size_t resident_bytes = 0;  // sum of resident plans and pools
  std::thread worker;

  size_t in_words(const ParamSet &p, Op op, size_t L) const {
    return (op == Op::ModDown ? L + p.p.size() : L) * p.n;
  }
  size_t out_words(const ParamSet &p, Op op, size_t L) const {
    return (op == Op::ModUp ? L + p.p.size() : L) * p.n;
  }
  std::shared_ptr<const ParamSet> param_set(uint32_t id) const {
    std::lock_guard<std::mutex> lock(mu);
    return params.at(id);
  }
  size_t held() const { return arena_bytes + resident_bytes; }
  void publish_memory() {
    std::lock_guard<std::mutex> lock(mu);
    stats.device_bytes_in_flight = held();
    stats.resident_bytes = resident_bytes;
    stats.peak_device_bytes = std::max(stats.peak_device_bytes, held());
    stats.plans_resident = plans.size() + pools.size();
  }

  // ---- making room for `need` more bytes, the arena of keep_slot (which the
  // need replaces) not counted: idle arenas first, then idle residents, least
  // recently used first. keep (the plan the need is for) is never evicted.
  bool fits(size_t need, const Slot *keep_slot) const {
    return held() - (keep_slot ? keep_slot->device_bytes : 0) + need <= opt.hbm_budget_bytes;
  }
  bool make_room(size_t need, const Slot *keep_slot, const Resident *keep) {
    if (fits(need, keep_slot)) return true;
    for (Slot &o : slots)
      if (&o != keep_slot && !o.busy && o.device) {
        ok(ntt_hip_free(opt.device, o.device), "free");
        arena_bytes -= o.device_bytes;
        o.device = nullptr;
        o.device_bytes = 0;
        if (fits(need, keep_slot)) return true;
      }
    for (;;) {
      Resident *victim = nullptr;
      std::function<void()> erase;
      for (auto it = plans.begin(); it != plans.end(); ++it)
        if (it->second.get() != keep && it->second->inflight == 0 &&
            (!victim || it->second->last_use < victim->last_use)) {
          victim = it->second.get();
          erase = [this, key = it->first] { plans.erase(key); };
        }
      for (auto it = pools.begin(); it != pools.end(); ++it)
        if (it->second.get() != keep && it->second->inflight == 0 &&
            (!victim || it->second->last_use < victim->last_use)) {
          victim = it->second.get();
          erase = [this, key = it->first] {
            pools.erase(key);
            pool_sets.erase(key);
          };
        }
      if (!victim) return false;
      resident_bytes -= victim->device_bytes;
      {
        std::lock_guard<std::mutex> lock(mu);
        ++stats.plans_evicted;
      }
      erase();  // destroys the contexts and frees their device memory
      if (fits(need, keep_slot)) return true;
    }
  }

  // Device bytes of a new plan, before it exists: twiddle tables (two
  // directions of 16-byte pairs per tower) and the active key (both channels).
  static size_t plan_estimate(const ParamSet &p, size_t L) {
    const size_t T = L + p.p.size(), beta = (L + p.alpha - 1) / p.alpha;
    return T * p.n * 32 + 2 * beta * T * p.n * 8;
  }

  std::shared_ptr<LevelPlan> plan(uint32_t id, size_t L, const Slot &slot) {
    auto key = std::make_pair(id, L);
    auto it = plans.find(key);
    if (it != plans.end()) {
      it->second->last_use = ++tick;
      return it->second;
    }
    const ParamSet &p = *param_set(id);
    (void)slot;
    if (!make_room(plan_estimate(p, L), nullptr, nullptr)) return nullptr;
    auto lp = std::make_shared<LevelPlan>();
    std::vector<uint64_t> q(p.q.begin(), p.q.begin() + L), r(p.q_psi.begin(), p.q_psi.begin() + L);
    q.insert(q.end(), p.p.begin(), p.p.end());
    r.insert(r.end(), p.p_psi.begin(), p.p_psi.end());
    ok(ntt_hip_context_create(&lp->ctx, opt.device, p.n, q.data(), r.data(), q.size()), "context");
    ok(ntt_hip_keyswitch_create(&lp->ks, lp->ctx, L, p.p.size(), p.alpha, opt.max_batch), "keyswitch plan");
    ok(ntt_hip_keyswitch_upload_key(lp->ks, p.key_b.data(), p.key_a.data(), p.q.size(), p.q.data()), "key");
    size_t cb = 0, kb = 0;
    ok(ntt_hip_context_device_bytes(lp->ctx, &cb), "context bytes");
    ok(ntt_hip_keyswitch_device_bytes(lp->ks, &kb), "plan bytes");
    lp->device_bytes = cb + kb;
    lp->last_use = ++tick;
    resident_bytes += lp->device_bytes;
    plans.emplace(key, lp);
    publish_memory();
    return lp;
  }

  // The pool of ring dimension n over every Q prime registered for n (rebuilt
  // when parameter sets were added since, and only while idle).
  std::shared_ptr<Pool> pool(uint64_t n, const Slot &slot) {
    std::vector<uint64_t> q, r;
    size_t sets = 0;
    {
      std::lock_guard<std::mutex> lock(mu);
      for (const auto &ps : params) {
        if (ps->n != n) continue;
        ++sets;
        for (size_t t = 0; t < ps->q.size(); ++t)
          if (std::find(q.begin(), q.end(), ps->q[t]) == q.end()) {
            q.push_back(ps->q[t]);
            r.push_back(ps->q_psi[t]);
          }
      }
    }
    auto it = pools.find(n);
    if (it != pools.end() && (pool_sets[n] == sets || it->second->inflight > 0)) {
      it->second->last_use = ++tick;
      return it->second;
    }
    if (it != pools.end()) {  // idle and stale: rebuild over the larger union
      resident_bytes -= it->second->device_bytes;
      pools.erase(it);
    }
    (void)slot;
    if (q.empty() || !make_room(q.size() * n * 32, nullptr, nullptr)) return nullptr;
    auto pl = std::make_shared<Pool>();
    ok(ntt_hip_context_create(&pl->ctx, opt.device, n, q.data(), r.data(), q.size()), "pool context");
    for (size_t t = 0; t < q.size(); ++t) pl->tower_of[q[t]] = (uint32_t)t;
    ok(ntt_hip_context_device_bytes(pl->ctx, &pl->device_bytes), "pool bytes");
    pl->last_use = ++tick;
    resident_bytes += pl->device_bytes;
    pools[n] = pl;
    pool_sets[n] = sets;
    publish_memory();
    return pl;
  }

  size_t workspace_bytes(LevelPlan &lp, Op op, size_t b) {
    size_t bytes = 0;
    if (op == Op::KeySwitch) ok(ntt_hip_keyswitch_workspace_bytes(lp.ks, b, &bytes), "workspace");
    else if (op == Op::ModUp) ok(ntt_hip_modup_workspace_bytes(lp.ks, b, &bytes), "workspace");
    else if (op == Op::ModDown) ok(ntt_hip_moddown_workspace_bytes(lp.ks, b, &bytes), "workspace");
    return bytes;
  }

  // Arenas grow geometrically (x1.5, the device part only while the extra
  // fits the budget): each regrowth frees and allocates, which waits for the
  // device, so a stream whose batches grow reallocates a few times rather
  // than once per new batch size.
  void ensure_arena(Slot &s, size_t device_bytes, size_t pinned_bytes) {
    if (s.device_bytes < device_bytes) {
      size_t want = std::max(device_bytes, s.device_bytes + s.device_bytes / 2);
      if (want > device_bytes && held() - s.device_bytes + want > opt.hbm_budget_bytes) want = device_bytes;
      if (s.device) ok(ntt_hip_free(opt.device, s.device), "free");
      arena_bytes -= s.device_bytes;
      s.device = nullptr;
      s.device_bytes = 0;
      ok(ntt_hip_malloc(opt.device, &s.device, want), "arena");
      s.device_bytes = want;
      arena_bytes += want;
    }
    if (s.pinned_bytes < pinned_bytes) {
      const size_t want = std::max(pinned_bytes, s.pinned_bytes + s.pinned_bytes / 2);
      if (s.pinned) ok(ntt_hip_host_free(s.pinned), "host free");
      s.pinned = nullptr;
      s.pinned_bytes = 0;
      ok(ntt_hip_host_alloc(&s.pinned, want), "pinned staging");
      s.pinned_bytes = want;
    }
  }

  static bool is_transform(Op op) { return op == Op::Forward || op == Op::Inverse; }
  // Transforms of parameter sets whose ring dimension is 2^13 .. 2^16 and
  // whose moduli are below 2^61 share launches across ring dimensions
  // (ntt_hip_transform_mixed); others group by ring dimension.
  static bool mixable(const ParamSet &p) {
    if (p.n < 8192 || p.n > 65536) return false;
    for (uint64_t q : p.q)
      if (q >= (uint64_t{1} << 61)) return false;
    return true;
  }

  // Take the head request's group off the queue (at most max_batch).
  std::vector<std::shared_ptr<Request>> take_group() {
    std::lock_guard<std::mutex> lock(mu);
    std::vector<std::shared_ptr<Request>> batch;
    if (pending.empty()) return batch;
    const auto head = pending.front();
    const bool head_mixable = is_transform(head->op) && mixable(*params.at(head->params));
    for (auto it = pending.begin(); it != pending.end() && batch.size() < opt.max_batch;) {
      const Request &r = **it;
      const bool same =
          r.op == head->op && r.device == head->device &&
          (is_transform(head->op) ? r.n == head->n || (head_mixable && mixable(*params.at(r.params)))
                                  : r.params == head->params && r.level == head->level);
      if (same) {
        batch.push_back(*it);
        it = pending.erase(it);
      } else {
        ++it;
      }
    }
    return batch;
  }

  // Launch one batch on an idle slot. Returns false if nothing could be admitted.
  bool launch(Slot &slot) {
    std::vector<std::shared_ptr<Request>> batch = take_group();
    if (batch.empty()) return false;
    const Request &h = *batch.front();
    const auto hset = param_set(h.params);
    const bool pooled = is_transform(h.op);
    std::shared_ptr<LevelPlan> lp;
    // the pool of every ring dimension in the batch, each held (in flight) from
    // the moment it is fetched, so fetching or admitting another cannot evict it
    std::map<uint64_t, std::shared_ptr<Pool>> pls;
    auto unhold = [&]() {
      for (auto &kv : pls)
        if (kv.second) --kv.second->inflight;
      pls.clear();
    };
    try {
      if (pooled) {
        for (const auto &r : batch)
          if (!pls.count(r->n)) {
            auto pl = pool(r->n, slot);
            if (pl) ++pl->inflight;
            pls[r->n] = pl;
          }
      } else {
        lp = plan(h.params, h.level, slot);
      }
    } catch (const std::exception &e) {
      unhold();
      fail(batch, e.what());
      return true;
    }
    if (pooled) {
      // a busy pool built before newer parameter sets may lack their primes,
      // and a pool may not fit yet: those requests wait
      std::vector<std::shared_ptr<Request>> covered, later;
      for (const auto &r : batch) {
        const auto &pl = pls[r->n];
        const auto ps = param_set(r->params);
        bool all = pl != nullptr;
        for (size_t t = 0; t < r->level && all; ++t) all = pl->tower_of.count(ps->q[t]) != 0;
        (all ? covered : later).push_back(r);
      }
      if (covered.empty()) {
        const bool no_room = pls.size() == 1 && !pls.begin()->second;
        unhold();
        if (no_room) return wait_or_fail(batch, "the plan does not fit the device budget");
        requeue(later);
        return false;
      }
      if (!later.empty()) requeue(later);
      batch = std::move(covered);
    } else if (!lp) {  // no room for the plan itself
      return wait_or_fail(batch, "the plan does not fit the device budget");
    }
    // pools are held in flight, so make_room cannot evict them; a plan is kept explicitly
    Resident *res = pooled ? nullptr : lp.get();
    // per-request input/output extents (words)
    std::vector<std::shared_ptr<const ParamSet>> sets;
    for (const auto &r : batch) sets.push_back(param_set(r->params));
    auto in_w = [&](size_t i) { return in_words(*sets[i], batch[i]->op, batch[i]->level); };
    auto out_w = [&](size_t i) { return out_words(*sets[i], batch[i]->op, batch[i]->level); };
    // Device-resident key switches and device-resident transforms of ring
    // dimensions the mixed launch takes run on the callers' buffers through
    // pointer tables: no gather into the arena, no scatter back, and the arena
    // holds only the workspace.
    const bool device_ks = h.device && h.op == Op::KeySwitch;
    bool zero_copy = device_ks;
    if (pooled && h.device) {
      zero_copy = true;
      for (const auto &ps : sets) zero_copy = zero_copy && mixable(*ps);
    }
    auto footprint = [&](size_t b) {
      if (device_ks) {
        size_t bytes = 0;
        ok(ntt_hip_keyswitch_gather_workspace_bytes(lp->ks, b, &bytes), "workspace");
        return bytes;
      }
      if (zero_copy) {
        size_t towers = 0, table = 0;
        for (size_t i = 0; i < b; ++i) towers += batch[i]->level;
        ok(ntt_hip_transform_mixed_workspace_bytes(towers, &table), "mixed workspace");
        return (table + 255) / 256 * 256;
      }
      size_t words = 0;
      for (size_t i = 0; i < b; ++i) words += in_w(i) + out_w(i) * (h.op == Op::KeySwitch ? 2 : 1);
      size_t bytes = words * 8;
      if (pooled) {
        size_t towers = 0;
        for (size_t i = 0; i < b; ++i) towers += batch[i]->level;
        size_t table = 0;
        if (towers) ok(ntt_hip_transform_mixed_workspace_bytes(towers, &table), "mixed workspace");
        bytes += (std::max(towers * sizeof(ntt_hip_tower_job), table) + 255) / 256 * 256;
      } else {
        bytes += workspace_bytes(*lp, h.op, b);
      }
      return bytes;
    };
    // admission by the whole footprint: shrink the batch, then make room
    size_t b = batch.size();
    while (b > 1 && !fits(footprint(b), &slot)) --b;
    if (!make_room(footprint(b), &slot, res)) {
      unhold();
      return wait_or_fail(batch, "request does not fit the device budget");
    }
    if (b < batch.size()) {
      requeue(std::vector<std::shared_ptr<Request>>(batch.begin() + b, batch.end()));
      batch.resize(b);
      sets.resize(b);
    }
    // the residents the admitted batch runs on (pools no longer needed are released)
    std::vector<std::shared_ptr<Resident>> holders;
    if (pooled) {
      for (auto it = pls.begin(); it != pls.end();) {
        bool used = false;
        for (const auto &r : batch) used |= r->n == it->first;
        if (used) {
          holders.push_back(it->second);
          ++it;
        } else {
          if (it->second) --it->second->inflight;
          it = pls.erase(it);
        }
      }
    } else {
      holders.push_back(lp);
    }
    try {
      size_t iw = 0, ow = 0, ow2 = 0;
      std::vector<size_t> in_at(b), out_at(b), out2_at(b), out_len(b);
      for (size_t i = 0; i < b; ++i) {
        in_at[i] = iw;
        out_at[i] = ow;
        out_len[i] = out_w(i);
        iw += in_w(i);
        ow += out_len[i];
      }
      if (h.op == Op::KeySwitch) ow2 = ow;
      for (size_t i = 0; i < b; ++i) out2_at[i] = ow + out_at[i];
      const size_t dev = footprint(b);
      ensure_arena(slot, dev, h.device ? 0 : (iw + ow + ow2) * 8);
      uint64_t *d_in = static_cast<uint64_t *>(slot.device), *d_out = d_in + iw, *d_out2 = d_out + ow;
      void *d_ws = d_out2 + ow2;
      uint64_t *h_in = static_cast<uint64_t *>(slot.pinned);
      const bool direct = h.device && b == 1 && !pooled;  // run on the caller's buffers
      if (zero_copy) {
        d_ws = slot.device;  // the arena holds only the workspace
        d_in = const_cast<uint64_t *>(h.d_in);  // a single key switch runs on the caller's buffers
        d_out = h.d_out;
        d_out2 = h.d_out2;
      } else if (!h.device) {
        for (size_t i = 0; i < b; ++i) std::memcpy(h_in + in_at[i], batch[i]->in.data(), in_w(i) * 8);
        ok(ntt_hip_memcpy(opt.device, d_in, h_in, iw * 8, NTT_HIP_H2D, slot.stream), "h2d");
      } else if (direct) {
        d_in = const_cast<uint64_t *>(h.d_in);
        d_out = h.d_out;
        d_out2 = h.d_out2;
      } else {
        for (size_t i = 0; i < b; ++i)
          ok(ntt_hip_memcpy(opt.device, d_in + in_at[i], batch[i]->d_in, in_w(i) * 8, NTT_HIP_D2D, slot.stream),
             "gather");
      }
      if (zero_copy && pooled) {
        // one mixed launch per pass, every tower read and written in place at the caller's addresses
        std::vector<ntt_hip_mixed_job> jobs;
        for (size_t i = 0; i < b; ++i) {
          const auto &pl = pls[batch[i]->n];
          for (size_t t = 0; t < batch[i]->level; ++t) {
            const auto f = pl->tower_of.find(sets[i]->q[t]);
            if (f == pl->tower_of.end()) throw std::runtime_error("prime missing from the ring's pool");
            const size_t at = t * batch[i]->n;
            jobs.push_back({pl->ctx, batch[i]->d_out + at, batch[i]->d_in + at, f->second, 0});
          }
        }
        ok(ntt_hip_transform_mixed(jobs.data(), jobs.size(), d_ws, h.op == Op::Inverse, slot.stream), "transform");
      } else if (device_ks && b > 1) {
        std::vector<const uint64_t *> ins(b);
        std::vector<uint64_t *> obs(b), oas(b);
        for (size_t i = 0; i < b; ++i) {
          ins[i] = batch[i]->d_in;
          obs[i] = batch[i]->d_out;
          oas[i] = batch[i]->d_out2;
        }
        ok(ntt_hip_keyswitch_run_gather(lp->ks, ins.data(), b, obs.data(), oas.data(), d_ws, slot.stream),
           "keyswitch");
      } else if (pooled && pls.size() == 1) {
        // one job per tower of every request, each with its prime's pool tower
        const auto &pl = pls.begin()->second;
        std::vector<ntt_hip_tower_job> jobs;
        for (size_t i = 0; i < b; ++i)
          for (size_t t = 0; t < batch[i]->level; ++t) {
            const auto f = pl->tower_of.find(sets[i]->q[t]);
            if (f == pl->tower_of.end()) throw std::runtime_error("prime missing from the ring's pool");
            const uint64_t at = in_at[i] + t * batch[i]->n;  // the head may have been requeued
            jobs.push_back({at, at, f->second, 0});
          }
        ok(ntt_hip_transform_jobs(pl->ctx, d_out, d_in, jobs.data(), jobs.size(), d_ws, h.op == Op::Inverse,
                                  slot.stream),
           "transform");
      } else if (pooled) {
        // several ring dimensions: one mixed launch per pass over the pools
        std::vector<ntt_hip_mixed_job> jobs;
        for (size_t i = 0; i < b; ++i) {
          const auto &pl = pls[batch[i]->n];
          for (size_t t = 0; t < batch[i]->level; ++t) {
            const auto f = pl->tower_of.find(sets[i]->q[t]);
            if (f == pl->tower_of.end()) throw std::runtime_error("prime missing from the ring's pool");
            const uint64_t at = in_at[i] + t * batch[i]->n;
            jobs.push_back({pl->ctx, d_out + at, d_in + at, f->second, 0});
          }
        }
        ok(ntt_hip_transform_mixed(jobs.data(), jobs.size(), d_ws, h.op == Op::Inverse, slot.stream), "transform");
      } else {
        const size_t L = h.level;
        switch (h.op) {
          case Op::KeySwitch:
            ok(ntt_hip_keyswitch_run(lp->ks, d_in, b, d_out, d_out2, d_ws, slot.stream), "keyswitch");
            break;
          case Op::ModUp:
            ok(ntt_hip_modup_run(lp->ks, d_in, 1, b, d_out, d_ws, slot.stream), "modup");
            break;
          case Op::ModDown:
            ok(ntt_hip_moddown_run(lp->ks, d_in, b, d_out, d_ws, slot.stream), "moddown");
            break;
          default:
            (void)L;
            throw std::logic_error("transform outside the pool path");
        }
      }
      if (!h.device) {
        uint64_t *h_out = h_in + iw;
        ok(ntt_hip_memcpy(opt.device, h_out, d_out, (ow + ow2) * 8, NTT_HIP_D2H, slot.stream), "d2h");
      } else if (!direct && !zero_copy) {
        for (size_t i = 0; i < b; ++i) {
          ok(ntt_hip_memcpy(opt.device, batch[i]->d_out, d_out + out_at[i], out_len[i] * 8, NTT_HIP_D2D, slot.stream),
             "scatter");
          if (ow2)
            ok(ntt_hip_memcpy(opt.device, batch[i]->d_out2, d_out2 + out_at[i], out_len[i] * 8, NTT_HIP_D2D,
                              slot.stream),
               "scatter");
        }
      }
      ok(ntt_hip_event_record(opt.device, slot.event, slot.stream), "record");
      slot.batch = std::move(batch);
      slot.out_at = out_at;
      slot.out_words = out_len;
      slot.out2_at.assign(ow2 ? out2_at.begin() : out2_at.end(), out2_at.end());
      slot.holders = holders;
      for (const auto &r : slot.holders) {
        if (!pooled) ++r->inflight;  // pools were held since they were fetched
        r->last_use = ++tick;
      }
      slot.busy = true;
      publish_memory();
      std::lock_guard<std::mutex> lock(mu);
      ++stats.batches;
      size_t busy = 0;
      for (const Slot &s : slots) busy += s.busy;
      stats.max_in_flight = std::max(stats.max_in_flight, busy);
    } catch (const std::exception &e) {
      (void)ntt_hip_stream_synchronize(opt.device, slot.stream);
      unhold();
      fail(batch, e.what());
    }
    return true;
  }

  // A batch that cannot be admitted waits while others run, else fails.
  bool wait_or_fail(const std::vector<std::shared_ptr<Request>> &batch, const char *why) {
    const bool others_busy = std::any_of(slots.begin(), slots.end(), [](const Slot &o) { return o.busy; });
    if (!others_busy) {
      fail(batch, why);
      return true;
    }
    requeue(batch);
    return false;
  }

  bool retire(Slot &slot) {
    if (!slot.busy) return false;
    const int q = ntt_hip_event_query(opt.device, slot.event);
    if (q == NTT_HIP_NOT_READY) return false;
    const auto now = Clock::now();
    const size_t b = slot.batch.size();
    size_t iw = 0;
    for (const auto &r : slot.batch) iw += r->device ? 0 : r->in.size();
    const uint64_t *h_out = static_cast<const uint64_t *>(slot.pinned) + iw;
    {
      std::lock_guard<std::mutex> lock(mu);
      for (size_t i = 0; i < b; ++i) {
        Request &r = *slot.batch[i];
        if (q != NTT_HIP_OK) {
          r.error = std::string("HIP: batch failed: ") + ntt_hip_last_error();
        } else if (!r.device) {
          r.result.out.assign(h_out + slot.out_at[i], h_out + slot.out_at[i] + slot.out_words[i]);
          if (!slot.out2_at.empty())
            r.result.out2.assign(h_out + slot.out2_at[i], h_out + slot.out2_at[i] + slot.out_words[i]);
        }
        r.result.latency_ms = std::chrono::duration<double, std::milli>(now - r.t0).count();
        r.result.batch_size = b;
        r.done = true;
        ++stats.completed;
      }
    }
    for (const auto &r : slot.holders) --r->inflight;
    slot.holders.clear();
    slot.batch.clear();
    slot.busy = false;
    cv_done.notify_all();
    return true;
  }

  void fail(const std::vector<std::shared_ptr<Request>> &batch, const std::string &why) {
    std::lock_guard<std::mutex> lock(mu);
    for (const auto &r : batch) {
      r->error = why;
      r->done = true;
      ++stats.completed;
    }
    cv_done.notify_all();
  }

  void requeue(const std::vector<std::shared_ptr<Request>> &reqs) {
    std::lock_guard<std::mutex> lock(mu);
    for (auto it = reqs.rbegin(); it != reqs.rend(); ++it) pending.push_front(*it);
  }

  void run() {
    for (;;) {
      bool progress = false;
      for (Slot &s : slots) progress |= retire(s);
      for (Slot &s : slots)
        if (!s.busy) progress |= launch(s);
      std::unique_lock<std::mutex> lock(mu);
      const bool idle = std::none_of(slots.begin(), slots.end(), [](const Slot &s) { return s.busy; });
      if (stop && pending.empty() && idle) return;
      if (!progress) {
        if (idle) cv_submit.wait_for(lock, std::chrono::milliseconds(5));
        else cv_submit.wait_for(lock, std::chrono::microseconds(50));
      }
    }
  }
};

Server::Server(Options opt) : impl_(std::make_unique<Impl>()) {
  if (opt.streams < 1 || opt.max_batch < 1) throw std::invalid_argument("server needs >= 1 stream and batch");
  impl_->opt = opt;
  impl_->slots.resize(opt.streams);
  for (Slot &s : impl_->slots) {
    ok(ntt_hip_stream_create(opt.device, &s.stream), "stream");
    ok(ntt_hip_event_create(opt.device, &s.event, 0), "event");
  }
  impl_->worker = std::thread([this] { impl_->run(); });
}

Server::~Server() {
  {
    std::lock_guard<std::mutex> lock(impl_->mu);
    impl_->stop = true;
  }
  impl_->cv_submit.notify_all();
  if (impl_->worker.joinable()) impl_->worker.join();
  for (Slot &s : impl_->slots) {
    if (s.stream) ntt_hip_stream_synchronize(impl_->opt.device, s.stream);
    if (s.device) ntt_hip_free(impl_->opt.device, s.device);
    if (s.pinned) ntt_hip_host_free(s.pinned);
    if (s.event) ntt_hip_event_destroy(impl_->opt.device, s.event);
    if (s.stream) ntt_hip_stream_destroy(impl_->opt.device, s.stream);
  }
  impl_->plans.clear();
  impl_->pools.clear();
}

uint32_t Server::add_params(const ParamSet &p) {
  const size_t M = p.q.size(), K = p.p.size();
  if (p.n < 2 || M == 0 || K == 0 || p.q_psi.size() != M || p.p_psi.size() != K || p.alpha == 0)
    throw std::invalid_argument("parameter set needs n, M >= 1 Q towers, K >= 1 P towers and alpha >= 1");
  const size_t words = ((M + p.alpha - 1) / p.alpha) * (M + K) * p.n;
  if (p.key_b.size() != words || p.key_a.size() != words)
    throw std::invalid_argument("evaluation key must hold ceil(M/alpha) digits of (M+K)*n residues per channel");
  std::lock_guard<std::mutex> lock(impl_->mu);
  impl_->params.push_back(std::make_shared<const ParamSet>(p));
  return static_cast<uint32_t>(impl_->params.size() - 1);
}

size_t Server::output_words(uint32_t id, Op op, size_t level) const {
  std::lock_guard<std::mutex> lock(impl_->mu);
  if (id >= impl_->params.size()) throw std::invalid_argument("unknown parameter set");
  return impl_->out_words(*impl_->params[id], op, level);
}

uint64_t Server::submit(uint32_t id, Op op, size_t level, std::span<const uint64_t> input) {
  auto r = std::make_shared<Request>();
  {
    std::lock_guard<std::mutex> lock(impl_->mu);
    if (impl_->stop) throw std::invalid_argument("server stopping");
    if (id >= impl_->params.size()) throw std::invalid_argument("unknown parameter set");
    const ParamSet &p = *impl_->params[id];
    if (static_cast<int>(op) < 0 || static_cast<int>(op) > 4) throw std::invalid_argument("unknown operation");
    if (level < 1 || level > p.q.size()) throw std::invalid_argument("level outside the parameter set");
    if (input.size() != impl_->in_words(p, op, level)) throw std::invalid_argument("input length mismatch");
    const size_t towers = input.size() / p.n;
    for (size_t t = 0; t < towers; ++t) {
      const uint64_t q = t < level ? p.q[t] : p.p[t - level];
      for (size_t i = 0; i < p.n; ++i)
        if (input[t * p.n + i] >= q) throw std::invalid_argument("noncanonical input residue");
    }
    r->ticket = impl_->next_ticket++;
    r->params = id;
    r->op = op;
    r->level = level;
    r->n = p.n;
    r->in.assign(input.begin(), input.end());
    r->t0 = Clock::now();
    impl_->pending.push_back(r);
    impl_->tickets.emplace(r->ticket, r);
    ++impl_->stats.submitted;
  }
  impl_->cv_submit.notify_one();
  return r->ticket;
}

uint64_t Server::submit_device(uint32_t id, Op op, size_t level, const uint64_t *in, uint64_t *out, uint64_t *out2) {
  auto r = std::make_shared<Request>();
  {
    std::lock_guard<std::mutex> lock(impl_->mu);
    if (impl_->stop) throw std::invalid_argument("server stopping");
    if (id >= impl_->params.size()) throw std::invalid_argument("unknown parameter set");
    const ParamSet &p = *impl_->params[id];
    if (static_cast<int>(op) < 0 || static_cast<int>(op) > 4) throw std::invalid_argument("unknown operation");
    if (level < 1 || level > p.q.size()) throw std::invalid_argument("level outside the parameter set");
    if (!in || !out || (op == Op::KeySwitch && !out2)) throw std::invalid_argument("null device buffer");
    r->ticket = impl_->next_ticket++;
    r->params = id;
    r->op = op;
    r->level = level;
    r->n = p.n;
    r->device = true;
    r->d_in = in;
    r->d_out = out;
    r->d_out2 = op == Op::KeySwitch ? out2 : nullptr;
    r->t0 = Clock::now();
    impl_->pending.push_back(r);
    impl_->tickets.emplace(r->ticket, r);
    ++impl_->stats.submitted;
  }
  impl_->cv_submit.notify_one();
  return r->ticket;
}

bool Server::ready(uint64_t ticket) const {
  std::lock_guard<std::mutex> lock(impl_->mu);
  auto it = impl_->tickets.find(ticket);
  if (it == impl_->tickets.end()) throw std::invalid_argument("unknown ticket");
  return it->second->done;
}

Result Server::wait(uint64_t ticket) {
  std::unique_lock<std::mutex> lock(impl_->mu);
  auto it = impl_->tickets.find(ticket);
  if (it == impl_->tickets.end()) throw std::invalid_argument("unknown ticket");
  // Keep the request alive by its shared pointer and erase by key afterwards:
  // a submit() on another thread may rehash the map while this thread waits.
  std::shared_ptr<Request> r = it->second;
  if (r->claimed) throw std::invalid_argument("ticket already being waited on");
  r->claimed = true;
  impl_->cv_done.wait(lock, [&] { return r->done; });
  impl_->tickets.erase(ticket);
  if (!r->error.empty()) throw std::runtime_error(r->error);
  return std::move(r->result);
}

Stats Server::stats() const {
  std::lock_guard<std::mutex> lock(impl_->mu);
  return impl_->stats;
}

}  // namespace ntt::serve

// ---------------------------------------------------------------- C ABI
namespace {
thread_local std::string g_serve_error;
template <class F>
int serve_guard(F &&f) {
  try {
    return f();
  } catch (const std::invalid_argument &e) {
    g_serve_error = e.what();
    return NTT_HIP_EINVAL;
  } catch (const std::exception &e) {
    g_serve_error = e.what();
    return NTT_HIP_ERUNTIME;
  }
}
}  // namespace

struct ntt_serve {
  ntt::serve::Server server;
  explicit ntt_serve(ntt::serve::Options o) : server(o) {}
};

extern "C" {

const char *ntt_serve_last_error(void) { return g_serve_error.c_str(); }

int ntt_serve_create(ntt_serve **out, int device, size_t hbm_budget_bytes, size_t streams, size_t max_batch) {
  return serve_guard([&] {
    if (!out) throw std::invalid_argument("null server out");
    *out = new ntt_serve(ntt::serve::Options{device, hbm_budget_bytes, streams, max_batch});
    return NTT_HIP_OK;
  });
}

int ntt_serve_destroy(ntt_serve *s) {
  return serve_guard([&] {
    delete s;
    return NTT_HIP_OK;
  });
}

int ntt_serve_add_params(ntt_serve *s, uint64_t n, const uint64_t *q, const uint64_t *q_psi, size_t M,
                         const uint64_t *p, const uint64_t *p_psi, size_t K, size_t alpha, const uint64_t *key_b,
                         const uint64_t *key_a, uint32_t *id) {
  return serve_guard([&] {
    if (!s || !q || !q_psi || !p || !p_psi || !key_b || !key_a || !id) throw std::invalid_argument("null argument");
    ntt::serve::ParamSet ps;
    ps.n = n;
    ps.q.assign(q, q + M);
    ps.q_psi.assign(q_psi, q_psi + M);
    ps.p.assign(p, p + K);
    ps.p_psi.assign(p_psi, p_psi + K);
    ps.alpha = alpha;
    const size_t words = alpha ? ((M + alpha - 1) / alpha) * (M + K) * n : 0;
    ps.key_b.assign(key_b, key_b + words);
    ps.key_a.assign(key_a, key_a + words);
    *id = s->server.add_params(ps);
    return NTT_HIP_OK;
  });
}

int ntt_serve_output_words(ntt_serve *s, uint32_t params, int op, size_t level, size_t *words) {
  return serve_guard([&] {
    if (!s || !words) throw std::invalid_argument("null argument");
    *words = s->server.output_words(params, static_cast<ntt::serve::Op>(op), level);
    return NTT_HIP_OK;
  });
}

int ntt_serve_submit(ntt_serve *s, uint32_t params, int op, size_t level, const uint64_t *in, size_t words,
                     uint64_t *ticket) {
  return serve_guard([&] {
    if (!s || !in || !ticket) throw std::invalid_argument("null argument");
    *ticket = s->server.submit(params, static_cast<ntt::serve::Op>(op), level, std::span<const uint64_t>(in, words));
    return NTT_HIP_OK;
  });
}

int ntt_serve_submit_device(ntt_serve *s, uint32_t params, int op, size_t level, const uint64_t *in, uint64_t *out,
                            uint64_t *out2, uint64_t *ticket) {
  return serve_guard([&] {
    if (!s || !ticket) throw std::invalid_argument("null argument");
    *ticket = s->server.submit_device(params, static_cast<ntt::serve::Op>(op), level, in, out, out2);
    return NTT_HIP_OK;
  });
}

int ntt_serve_ready(ntt_serve *s, uint64_t ticket) {
  return serve_guard([&] {
    if (!s) throw std::invalid_argument("null server");
    return s->server.ready(ticket) ? NTT_HIP_OK : NTT_HIP_NOT_READY;
  });
}

int ntt_serve_wait(ntt_serve *s, uint64_t ticket, uint64_t *out, size_t out_words, uint64_t *out2, size_t out2_words,
                   double *latency_ms, size_t *batch_size) {
  return serve_guard([&] {
    if (!s) throw std::invalid_argument("null argument");
    auto r = s->server.wait(ticket);
    if (out) {  // host requests; a device request's results are already in its buffers
      if (r.out.size() != out_words || r.out2.size() != (out2 ? out2_words : r.out2.size()))
        throw std::invalid_argument("output buffer length mismatch");
      std::copy(r.out.begin(), r.out.end(), out);
    }
    if (out2) std::copy(r.out2.begin(), r.out2.end(), out2);
    if (latency_ms) *latency_ms = r.latency_ms;
    if (batch_size) *batch_size = r.batch_size;
    return NTT_HIP_OK;
  });
}

int ntt_serve_stats_ex(ntt_serve *s, size_t *out, size_t count) {
  return serve_guard([&] {
    if (!s || !out) throw std::invalid_argument("null argument");
    const auto st = s->server.stats();
    const size_t v[9] = {st.submitted,        st.completed,      st.batches,         st.max_in_flight, st.device_bytes_in_flight,
                         st.peak_device_bytes, st.resident_bytes, st.plans_resident, st.plans_evicted};
    std::copy(v, v + std::min<size_t>(count, 9), out);
    return NTT_HIP_OK;
  });
}

int ntt_serve_stats(ntt_serve *s, size_t *out6) {
  return serve_guard([&] {
    if (!s || !out6) throw std::invalid_argument("null argument");
    const auto st = s->server.stats();
    const size_t v[6] = {st.submitted, st.completed, st.batches, st.max_in_flight, st.device_bytes_in_flight,
                         st.peak_device_bytes};
    std::copy(v, v + 6, out6);
    return NTT_HIP_OK;
  });
}

}  // extern "C"
