// Request server: continuous batching of propagation solves on one device, written over the
// library's own C ABI (no reference counterpart; the reference solves one fiber per object).
//
// A request is a deep copy of its mmp_params and a step count. Requests that the batched
// column-stage kernels take are grouped by structure (modes, samples, stages, pair slots,
// nonzeros, Raman lines): a group is one batch solver with a fixed number of slots, created
// with the request's own parameters in every slot and all but the first slot parked; later
// requests of that structure are loaded into free slots at the group's chunk boundaries.
// Other requests get a solver of their own. Every group and solo solver has its own context
// (stream). A group advances by the fewest steps any of its live requests still needs, capped
// at max_chunk, so no request takes a step too many; a request's result is therefore the
// result of the same solve alone (bit for bit; to rounding with MMP_FLAG_AUTOTUNE, whose gather
// variant for a group's slot count may sum over the pair slots in another order). A request whose
// field stops being finite is parked by the device at that step and ends at the next chunk
// boundary as MMP_REQ_NONFINITE, so it holds its slot no longer than that.

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <new>
#include <set>
#include <string>
#include <thread>
#include <vector>

#include "../../include/mmprop.h"
#include "setup.h"

namespace {

using Clock = std::chrono::steady_clock;

// mmp_params with every array owned
struct OwnedParams {
  mmp_params p{};
  std::vector<double> beta, values, fields, raman;
  std::vector<uint32_t> offsets, klm;

  void copy_from(const mmp_params& in) {
    p = in;
    const size_t M = in.n_modes, NT = in.nt;
    beta.assign(in.dispersion_coefficients, in.dispersion_coefficients + M * in.n_beta);
    offsets.assign(in.coupling_offsets, in.coupling_offsets + M + 1);
    const size_t nnz = offsets[M];
    klm.assign(in.coupling_klm, in.coupling_klm + 3 * nnz);
    values.assign(in.coupling_values, in.coupling_values + nnz);
    fields.assign(in.initial_fields, in.initial_fields + 2 * M * NT);
    if (in.raman && in.raman_response) raman.assign(in.raman_response, in.raman_response + NT);
    p.dispersion_coefficients = beta.data();
    p.coupling_offsets = offsets.data();
    p.coupling_klm = klm.data();
    p.coupling_values = values.data();
    p.initial_fields = fields.data();
    p.raman_response = raman.empty() ? nullptr : raman.data();
  }
};

struct Request {
  uint64_t id = 0;
  OwnedParams prm;
  uint32_t n_steps = 0;
  mmp_info plan{};
  uint32_t state = MMP_REQ_QUEUED;
  bool batched = false;
  uint32_t group = 0, slot = 0;
  double t_submit = 0., t_admit = -1., t_finish = -1.;
  std::vector<double> result;
};

struct Key {
  uint32_t v[8];
  bool operator<(const Key& o) const { return std::lexicographical_compare(v, v + 8, o.v, o.v + 8); }
};

Key key_of(const Request& r) {
  const mmp_info& i = r.plan;
  return Key{{i.n_modes, i.nt, i.stages, i.n_pairs, i.nnz, i.n_conv_lines, r.prm.p.raman, 0u}};
}

struct Group {
  mmp_ctx* ctx = nullptr;
  mmp_solver* batch = nullptr;
  Key key{};
  std::vector<int64_t> slot_req;  // request id per slot, -1 when free
  std::vector<uint32_t> remaining;
  bool busy = false;
  uint32_t chunk = 0;
  size_t bytes = 0;
  bool alive = true;  // false once reclaimed (the index stays valid in records)
};

struct Solo {
  mmp_ctx* ctx = nullptr;
  mmp_solver* solver = nullptr;
  uint64_t req = 0;
  uint32_t remaining = 0;
  bool busy = false;
  uint32_t chunk = 0;
  size_t bytes = 0;
};

}  // namespace

struct mmp_server {
  mmp_serve_options opt{};
  mmp_options sopt{};
  Clock::time_point t0;
  std::vector<std::unique_ptr<Request>> reqs;  // index = id
  std::deque<uint64_t> queue;
  std::vector<std::unique_ptr<Group>> groups;
  std::map<Key, size_t> group_of;
  std::set<Key> unbatchable;
  std::vector<std::unique_ptr<Solo>> solos;
  size_t bytes_in_use = 0, peak_bytes = 0;
  uint64_t running = 0, done = 0, nonfinite = 0;
  uint32_t groups_created = 0, solos_created = 0;

  double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }
  void add_bytes(size_t b) {
    bytes_in_use += b;
    peak_bytes = std::max(peak_bytes, bytes_in_use);
  }
};

namespace {

using mmp::record_error;

#define SRV_TRY(expr)                   \
  do {                                  \
    const mmp_status st_ = (expr);      \
    if (st_ != MMP_OK) return st_;      \
  } while (0)

void close_group(mmp_server* s, Group& g) {
  if (g.batch) mmp_solver_destroy(g.batch);
  if (g.ctx) mmp_ctx_destroy(g.ctx);
  g.batch = nullptr;
  g.ctx = nullptr;
  s->bytes_in_use -= g.bytes;
  g.bytes = 0;
  g.alive = false;
  s->group_of.erase(g.key);
}

void close_solo(mmp_server* s, Solo& o) {
  if (o.solver) mmp_solver_destroy(o.solver);
  if (o.ctx) mmp_ctx_destroy(o.ctx);
  o.solver = nullptr;
  o.ctx = nullptr;
  s->bytes_in_use -= o.bytes;
}

// Frees idle groups without live requests until `need` more bytes fit the budget.
bool make_room(mmp_server* s, size_t need) {
  if ((double)(s->bytes_in_use + need) <= s->opt.hbm_budget) return true;
  for (auto& g : s->groups) {
    if (!g->alive || g->busy) continue;
    if (std::any_of(g->slot_req.begin(), g->slot_req.end(), [](int64_t r) { return r >= 0; })) continue;
    close_group(s, *g);
    if ((double)(s->bytes_in_use + need) <= s->opt.hbm_budget) return true;
  }
  return false;
}

// Reads fiber f's field into the request and closes it (finite or not).
mmp_status finish(mmp_server* s, Request& r, mmp_solver* solver, uint32_t f) {
  r.result.resize(2 * (size_t)r.prm.p.n_modes * r.prm.p.nt);
  const mmp_status st = mmp_fiber_result(solver, f, r.result.data());
  if (st == MMP_ERR_NONFINITE) {
    r.state = MMP_REQ_NONFINITE;
    r.result.clear();
    s->nonfinite++;
  } else if (st == MMP_OK) {
    r.state = MMP_REQ_DONE;
    s->done++;
  } else {
    return st;
  }
  r.t_finish = s->now();
  s->running--;
  return MMP_OK;
}

void mark_running(mmp_server* s, Request& r) {
  r.state = MMP_REQ_RUNNING;
  r.t_admit = s->now();
  s->running++;
}

// A solver of its own; false when the budget or the device has no room now.
mmp_status admit_solo(mmp_server* s, Request& r, bool& admitted) {
  admitted = false;
  if (!make_room(s, r.plan.device_bytes)) return MMP_OK;
  auto o = std::make_unique<Solo>();
  mmp_status st = mmp_ctx_create(s->opt.device, &o->ctx);
  if (st == MMP_OK) st = mmp_solver_create(o->ctx, &r.prm.p, &s->sopt, &o->solver);
  if (st == MMP_ERR_OUT_OF_MEMORY) {
    close_solo(s, *o);
    return MMP_OK;  // try again once memory frees
  }
  if (st != MMP_OK) {
    close_solo(s, *o);
    return st;
  }
  mmp_info info{};
  mmp_solver_info(o->solver, &info);
  o->bytes = info.device_bytes;
  o->req = r.id;
  o->remaining = r.n_steps;
  s->add_bytes(o->bytes);
  s->solos.push_back(std::move(o));
  s->solos_created++;
  r.batched = false;
  mark_running(s, r);
  admitted = true;
  return MMP_OK;
}

mmp_status admit(mmp_server* s, Request& r, bool& admitted) {
  admitted = false;
  const Key key = key_of(r);
  const bool batchable = r.prm.p.raman && r.plan.gather_waves > 0 && !s->unbatchable.count(key);
  if (!batchable) return admit_solo(s, r, admitted);
  auto it = s->group_of.find(key);
  if (it != s->group_of.end()) {
    Group& g = *s->groups[it->second];
    if (g.busy) return MMP_OK;  // slots are filled at chunk boundaries
    const auto free = std::find(g.slot_req.begin(), g.slot_req.end(), -1);
    if (free == g.slot_req.end()) return MMP_OK;
    const uint32_t slot = (uint32_t)(free - g.slot_req.begin());
    const mmp_status st = mmp_fiber_load(g.batch, slot, &r.prm.p);
    if (st == MMP_ERR_INVALID_ARGUMENT) return admit_solo(s, r, admitted);  // same sizes, other coupling pattern
    SRV_TRY(st);
    g.slot_req[slot] = (int64_t)r.id;
    g.remaining[slot] = r.n_steps;
    r.batched = true;
    r.group = (uint32_t)it->second;
    r.slot = slot;
    mark_running(s, r);
    admitted = true;
    return MMP_OK;
  }
  const uint32_t F = s->opt.slots;
  if (!make_room(s, r.plan.device_bytes * F)) return MMP_OK;
  auto g = std::make_unique<Group>();
  g->key = key;
  mmp_status st = mmp_ctx_create(s->opt.device, &g->ctx);
  std::vector<mmp_params> all(F, r.prm.p);  // every slot starts as this request; all but slot 0 parked
  if (st == MMP_OK) st = mmp_batch_create(g->ctx, F, all.data(), &s->sopt, &g->batch);
  if (st != MMP_OK) {
    if (g->batch) mmp_solver_destroy(g->batch);
    if (g->ctx) mmp_ctx_destroy(g->ctx);
    if (st == MMP_ERR_UNSUPPORTED) {
      s->unbatchable.insert(key);
      return admit_solo(s, r, admitted);
    }
    if (st == MMP_ERR_OUT_OF_MEMORY) return MMP_OK;
    return st;
  }
  for (uint32_t i = 1; i < F; i++) {
    const mmp_status ps = mmp_fiber_set_live(g->batch, i, 0);
    if (ps != MMP_OK) {  // the group is not registered yet: release it here
      mmp_solver_destroy(g->batch);
      mmp_ctx_destroy(g->ctx);
      return ps;
    }
  }
  mmp_info info{};
  mmp_solver_info(g->batch, &info);
  g->bytes = info.device_bytes;
  g->slot_req.assign(F, -1);
  g->remaining.assign(F, 0);
  g->slot_req[0] = (int64_t)r.id;
  g->remaining[0] = r.n_steps;
  s->add_bytes(g->bytes);
  r.batched = true;
  r.group = (uint32_t)s->groups.size();
  r.slot = 0;
  s->group_of[key] = s->groups.size();
  s->groups.push_back(std::move(g));
  s->groups_created++;
  mark_running(s, r);
  admitted = true;
  return MMP_OK;
}

mmp_status poll_impl(mmp_server* s) {
  // chunk boundaries: groups and solvers whose stream is idle
  for (auto& gp : s->groups) {
    Group& g = *gp;
    if (!g.alive || !g.busy) continue;
    const mmp_status q = mmp_ctx_query(g.ctx);
    if (q == MMP_ERR_STATE) continue;
    SRV_TRY(q);
    g.busy = false;
    for (uint32_t i = 0; i < g.slot_req.size(); i++) {
      if (g.slot_req[i] < 0) continue;
      g.remaining[i] -= g.chunk;
      if (g.remaining[i] > 0) {
        // a field that stopped being finite was parked by the device: the request ends now
        mmp_info fi{};
        SRV_TRY(mmp_fiber_info(g.batch, i, &fi));
        if (!fi.halted) continue;
        g.remaining[i] = 0;
      }
      SRV_TRY(finish(s, *s->reqs[(size_t)g.slot_req[i]], g.batch, i));
      g.slot_req[i] = -1;
      SRV_TRY(mmp_fiber_set_live(g.batch, i, 0));
    }
  }
  for (size_t k = 0; k < s->solos.size();) {
    Solo& o = *s->solos[k];
    if (o.busy) {
      const mmp_status q = mmp_ctx_query(o.ctx);
      if (q == MMP_ERR_STATE) {
        k++;
        continue;
      }
      SRV_TRY(q);
      o.busy = false;
      o.remaining -= o.chunk;
      mmp_info fi{};
      SRV_TRY(mmp_solver_info(o.solver, &fi));
      if (fi.halted) o.remaining = 0;  // parked by the device: its field is not finite
    }
    if (o.remaining == 0) {
      SRV_TRY(finish(s, *s->reqs[o.req], o.solver, 0));
      close_solo(s, o);
      s->solos.erase(s->solos.begin() + (long)k);
      continue;
    }
    k++;
  }
  // admission, oldest first; a request that does not fit now does not hold back later ones
  for (auto it = s->queue.begin(); it != s->queue.end();) {
    bool admitted = false;
    SRV_TRY(admit(s, *s->reqs[*it], admitted));
    it = admitted ? s->queue.erase(it) : it + 1;
  }
  // next chunks
  for (auto& gp : s->groups) {
    Group& g = *gp;
    if (!g.alive || g.busy) continue;
    uint32_t chunk = 0;
    for (uint32_t i = 0; i < g.slot_req.size(); i++)
      if (g.slot_req[i] >= 0) chunk = chunk ? std::min(chunk, g.remaining[i]) : g.remaining[i];
    if (!chunk) continue;
    g.chunk = std::min(chunk, s->opt.max_chunk);
    SRV_TRY(mmp_advance(g.batch, g.chunk));
    g.busy = true;
  }
  for (auto& op : s->solos) {
    Solo& o = *op;
    if (o.busy || o.remaining == 0) continue;
    o.chunk = std::min(o.remaining, s->opt.max_chunk);
    SRV_TRY(mmp_advance(o.solver, o.chunk));
    o.busy = true;
  }
  return MMP_OK;
}

}  // namespace

extern "C" {

mmp_status mmp_serve_create(const mmp_serve_options* options, mmp_server** out) {
  if (!options || !out) return record_error(MMP_ERR_INVALID_ARGUMENT, "options and out are required");
  *out = nullptr;
  if (options->slots < 1 || options->slots > 4096 || options->max_chunk < 1 || !(options->hbm_budget > 0.) ||
      options->precision > MMP_PRECISION_MIXED || options->pair_decode > MMP_PAIRS_PHYSICAL)
    return record_error(MMP_ERR_INVALID_ARGUMENT,
                        "slots in [1, 4096], max_chunk >= 1, hbm_budget > 0 and valid precision and pair decode");
  mmp_server* s = new (std::nothrow) mmp_server();
  if (!s) return record_error(MMP_ERR_OUT_OF_MEMORY, "host allocation failed");
  s->opt = *options;
  s->sopt.precision = options->precision;
  s->sopt.pair_decode = options->pair_decode;
  s->sopt.flags = options->flags;
  s->t0 = Clock::now();
  *out = s;
  return MMP_OK;
}

void mmp_serve_destroy(mmp_server* s) {
  if (!s) return;
  for (auto& g : s->groups)
    if (g->alive) close_group(s, *g);
  for (auto& o : s->solos) close_solo(s, *o);
  delete s;
}

mmp_status mmp_serve_submit(mmp_server* s, const mmp_params* params, uint32_t n_steps, uint64_t* id_out) {
  if (!s || !params) return record_error(MMP_ERR_INVALID_ARGUMENT, "server and params are required");
  if (n_steps < 1) return record_error(MMP_ERR_INVALID_ARGUMENT, "n_steps must be >= 1");
  try {
    auto r = std::make_unique<Request>();
    SRV_TRY(mmp_plan(params, &s->sopt, &r->plan));  // validates every input
    r->prm.copy_from(*params);
    r->n_steps = n_steps;
    r->id = s->reqs.size();
    r->t_submit = s->now();
    if (id_out) *id_out = r->id;
    s->queue.push_back(r->id);
    s->reqs.push_back(std::move(r));
    return MMP_OK;
  } catch (const std::bad_alloc&) {
    return record_error(MMP_ERR_OUT_OF_MEMORY, "host allocation failed");
  }
}

mmp_status mmp_serve_poll(mmp_server* s, uint32_t* active_out) {
  if (!s) return record_error(MMP_ERR_INVALID_ARGUMENT, "server is NULL");
  try {
    SRV_TRY(poll_impl(s));
  } catch (const std::bad_alloc&) {
    return record_error(MMP_ERR_OUT_OF_MEMORY, "host allocation failed");
  }
  if (active_out) *active_out = (uint32_t)(s->queue.size() + s->running);
  return MMP_OK;
}

mmp_status mmp_serve_drain(mmp_server* s) {
  if (!s) return record_error(MMP_ERR_INVALID_ARGUMENT, "server is NULL");
  for (;;) {
    uint32_t active = 0;
    SRV_TRY(mmp_serve_poll(s, &active));
    if (active == 0) return MMP_OK;
    if (s->running == 0) {
      // nothing runs and nothing was admitted: the queue head can never fit
      return record_error(MMP_ERR_STATE, "a queued request's device footprint exceeds the budget");
    }
    std::this_thread::sleep_for(std::chrono::microseconds(100));
  }
}

mmp_status mmp_serve_record_get(mmp_server* s, uint64_t id, mmp_serve_record* out) {
  if (!s || !out) return record_error(MMP_ERR_INVALID_ARGUMENT, "server and out are required");
  if (id >= s->reqs.size()) return record_error(MMP_ERR_INVALID_ARGUMENT, "unknown request id");
  const Request& r = *s->reqs[id];
  std::memset(out, 0, sizeof *out);
  out->state = r.state;
  out->batched = r.batched ? 1u : 0u;
  out->group = r.group;
  out->slot = r.slot;
  out->t_submit = r.t_submit;
  out->t_admit = r.t_admit;
  out->t_finish = r.t_finish;
  out->units = (uint64_t)r.prm.p.n_modes * r.prm.p.nt * r.n_steps;
  out->device_bytes = r.plan.device_bytes;
  return MMP_OK;
}

mmp_status mmp_serve_result(mmp_server* s, uint64_t id, double* fields_out) {
  if (!s || !fields_out) return record_error(MMP_ERR_INVALID_ARGUMENT, "server and fields_out are required");
  if (id >= s->reqs.size()) return record_error(MMP_ERR_INVALID_ARGUMENT, "unknown request id");
  const Request& r = *s->reqs[id];
  if (r.state == MMP_REQ_NONFINITE) return record_error(MMP_ERR_NONFINITE, "the request's field is not finite");
  if (r.state != MMP_REQ_DONE) return record_error(MMP_ERR_STATE, "the request is queued or running");
  std::memcpy(fields_out, r.result.data(), r.result.size() * sizeof(double));
  return MMP_OK;
}

mmp_status mmp_serve_stats_get(mmp_server* s, mmp_serve_stats* out) {
  if (!s || !out) return record_error(MMP_ERR_INVALID_ARGUMENT, "server and out are required");
  std::memset(out, 0, sizeof *out);
  out->submitted = s->reqs.size();
  out->queued = s->queue.size();
  out->running = s->running;
  out->done = s->done;
  out->nonfinite = s->nonfinite;
  out->groups = (uint32_t)std::count_if(s->groups.begin(), s->groups.end(), [](const auto& g) { return g->alive; });
  out->solos = (uint32_t)s->solos.size();
  out->groups_created = s->groups_created;
  out->solos_created = s->solos_created;
  out->bytes_in_use = s->bytes_in_use;
  out->peak_bytes = s->peak_bytes;
  return MMP_OK;
}

}  // extern "C"
