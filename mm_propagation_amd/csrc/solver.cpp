// solver.cpp -- the C ABI of include/mmprop.h: contexts, solver setup, device
// memory, step scheduling (eager or HIP-graph replay) and results.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/mmprop.h"
#include "kernels.h"
#include "setup.h"

using namespace mmp;

namespace {

thread_local std::string g_last_error;

mmp_status fail(mmp_status st, const std::string& msg) {
  g_last_error = msg;
  return st;
}

mmp_status hip_fail(hipError_t e, const char* what) {
  g_last_error = std::string(what) + ": " + hipGetErrorString(e);
  if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) return MMP_ERR_OUT_OF_MEMORY;
  if (e == hipErrorNoBinaryForGpu || e == hipErrorNoDevice || e == hipErrorInvalidDevice) return MMP_ERR_NO_DEVICE;
  return MMP_ERR_DEVICE;
}

}  // namespace

// Errors raised by other translation units of the library (the request server).
mmp_status mmp::record_error(mmp_status st, const std::string& msg) { return fail(st, msg); }

namespace {

#define HIP_TRY(expr, what)                        \
  do {                                             \
    hipError_t e_ = (expr);                        \
    if (e_ != hipSuccess) return hip_fail(e_, what); \
  } while (0)

#define MMP_TRY(expr)                 \
  do {                                \
    mmp_status s_ = (expr);           \
    if (s_ != MMP_OK) return s_;      \
  } while (0)

constexpr int kGraphSteps = 16;  // steps captured in one replayable HIP graph

}  // namespace

struct mmp_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
};

struct KernelRecord {
  int name;
  double bytes;
  hipEvent_t start, stop;
};

struct mmp_solver {
  mmp_ctx* ctx = nullptr;
  mmp_options opt{};
  Problem P;                      // fiber 0; its structure (sizes, tableau, slots) is the batch's
  uint32_t F = 1;                 // fibers in the batch
  std::vector<Problem> fib;       // per-fiber scalars (large per-fiber arrays are freed once uploaded)
  std::vector<double> time_window, fiber_length, z_host;  // per fiber
  FiberParams* d_fp = nullptr;    // [F]
  double* d_coef = nullptr;       // [F][s + 1][kHeadCoef] stage-head coefficients times each fiber's h
  Batch batch{};
  cd* etab0 = nullptr;            // [F] initial Lawson tables (restart copies them into etab)
  size_t ntab = 0;                // Lawson tables per fiber
  Geometry g{};
  Twiddles tw{};
  std::vector<void*> allocations;
  size_t device_bytes = 0;
  // no-allocation solve: every device array is carved from the static scratch instead
  unsigned char* arena = nullptr;
  size_t arena_cap = 0, arena_used = 0;
  cd* psi_hat = nullptr;
  cd* acc = nullptr;
  cd* ybuf = nullptr;
  cd* U = nullptr;
  cd* conv = nullptr;
  cd* stash = nullptr;
  cd* hspec = nullptr;
  cd* etab = nullptr;   // Lawson tables, one per distinct c_i
  cd* e1 = nullptr;     // exp(h L)
  double* imL = nullptr;
  double* hR = nullptr;
  double* energy = nullptr;
  StepState* st = nullptr;
  int n_lines = 0;
  std::vector<HeadArgs> heads;  // s + 1 entries
  GatherArgs gather{};
  PairLineArgs pairs{};
  FusedArgs fused{};
  bool use_fused = false;
  int chunk_lines = 0;       // column-stage gather: Raman lines per LDS chunk at most
  bool tuned = false;        // MMP_FLAG_AUTOTUNE chose fused.waves and chunk_lines
  int pair_columns = 1;      // column-stage pair pass: columns per workgroup and line groups
  int pair_groups = 1;
  int conv_cols = 0;         // the pair pass stores the Raman lines in column order (large grids)
  bool chain = false;        // the step close also issues the next step's stage-0 head
  double* qdense = nullptr;  // [P][M][M] (or parity-compact) dense coupling blocks for the register gathers
  // work arrays of the step in the work precision (the FP64 buffers above in FP64 mode)
  bool mixed = false;
  size_t wbytes = 16;        // bytes per complex work value
  void* ybuf_w = nullptr;
  void* U_w = nullptr;
  void* conv_w = nullptr;
  void* convt_w = nullptr;   // column-stage step: the Raman lines in column order after their row pass
  void* hspec_w = nullptr;
  const void* q_w = nullptr;
  const void* qdense_w = nullptr;
  std::vector<uint32_t> steps;  // per fiber, since its last restart
  std::vector<uint8_t> live;    // per fiber: advanced by the steps (1) or parked (0)
  std::vector<uint8_t> halted;  // per fiber: parked by the device because its state stopped being finite
  bool graphs = true;
  hipGraphExec_t graph_exec = nullptr;   // kGraphSteps steps
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph1_exec = nullptr;  // one step (the remainder of a call)
  hipGraph_t graph1 = nullptr;
  // profiling
  bool profile = false;
  std::vector<std::string> kernel_names;
  std::vector<double> kernel_bytes;
  std::vector<double> kernel_ms;
  std::vector<uint64_t> kernel_count;
  std::vector<KernelRecord> pending;
  std::vector<hipEvent_t> free_events;
};

// Solvers of different shapes on one context advanced together: one graph replay steps them all.
// Members on the group's own context are captured in order on its stream; members on other
// contexts (other streams of the same device) become parallel branches of the graph, forked from
// and joined back into the group's stream, so their kernels may run concurrently.
struct mmp_group {
  mmp_ctx* ctx = nullptr;
  std::vector<mmp_solver*> members;
  std::vector<hipStream_t> branches;   // distinct member streams other than ctx->stream
  std::vector<hipEvent_t> events;      // fork, then one join per branch, then one per branch for advance
  hipGraph_t graph = nullptr, graph1 = nullptr;
  hipGraphExec_t exec = nullptr, exec1 = nullptr;
};

namespace {

constexpr size_t kArenaAlign = 256;

template <class T>
mmp_status dalloc(mmp_solver* s, T** out, size_t count) {
  void* p = nullptr;
  const size_t bytes = sizeof(T) * (count ? count : 1);
  if (s->arena) {
    const size_t off = (s->arena_used + kArenaAlign - 1) & ~(kArenaAlign - 1);
    if (off > s->arena_cap || bytes > s->arena_cap - off)
      return fail(MMP_ERR_UNSUPPORTED, "the problem needs more device memory than the static scratch holds (" +
                                           std::to_string(s->arena_cap) + " bytes)");
    p = s->arena + off;
    s->arena_used = off + bytes;
  } else {
    HIP_TRY(hipMalloc(&p, bytes), "hipMalloc");
    s->allocations.push_back(p);
  }
  s->device_bytes += bytes;
  *out = static_cast<T*>(p);
  return MMP_OK;
}

template <class T>
mmp_status upload(mmp_solver* s, T* dst, const void* src, size_t bytes) {
  HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s->ctx->stream), "hipMemcpyAsync H2D");
  return MMP_OK;
}

int kernel_id(mmp_solver* s, const char* name) {
  for (size_t i = 0; i < s->kernel_names.size(); i++)
    if (s->kernel_names[i] == name) return (int)i;
  s->kernel_names.push_back(name);
  s->kernel_bytes.push_back(0.);
  s->kernel_ms.push_back(0.);
  s->kernel_count.push_back(0);
  return (int)s->kernel_names.size() - 1;
}

hipEvent_t take_event(mmp_solver* s) {
  if (!s->free_events.empty()) {
    hipEvent_t e = s->free_events.back();
    s->free_events.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

mmp_status drain_profile(mmp_solver* s) {
  if (s->pending.empty()) return MMP_OK;
  HIP_TRY(hipStreamSynchronize(s->ctx->stream), "hipStreamSynchronize");
  for (auto& r : s->pending) {
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, r.start, r.stop), "hipEventElapsedTime");
    s->kernel_ms[r.name] += ms;
    s->kernel_count[r.name] += 1;
    s->kernel_bytes[r.name] += r.bytes;
    s->free_events.push_back(r.start);
    s->free_events.push_back(r.stop);
  }
  s->pending.clear();
  return MMP_OK;
}

// Launch `fn` on the solver stream, bracketed by HIP events when profiling.
template <class F>
mmp_status run_kernel(mmp_solver* s, const char* name, double bytes, F&& fn) {
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (s->profile) {
    e0 = take_event(s);
    e1 = take_event(s);
    if (!e0 || !e1) return fail(MMP_ERR_DEVICE, "hipEventCreate failed");
    g_launch_events = LaunchEvents{e0, e1};  // recorded by the dispatch itself
  }
  hipError_t e = fn();
  g_launch_events = LaunchEvents{};
  if (e != hipSuccess) return hip_fail(e, name);
  if (s->profile) {
    s->pending.push_back({kernel_id(s, name), bytes, e0, e1});
    if (s->pending.size() >= 2048) MMP_TRY(drain_profile(s));
  }
  return MMP_OK;
}

// Row stage head i (i == s closes the step; with s->chain it also builds the next step's
// stage 0 and advances z).
template <class W>
mmp_status enqueue_head_w(mmp_solver* s, int i) {
  const Problem& P = s->P;
  const HeadArgs& h = s->heads[i];
  const double cb = sizeof(cpx<W>), cs = 16.0;
  const double MN = (double)P.M * P.N;
  // algorithmic bytes: every array the head reads or writes, once
  // (Lawson tables only when they are not computed from Im L, which is then read once, 8 bytes)
  const int tab = h.recompute_e ? 0 : 1;
  int state_arrays = 0, work_arrays = 0;
  if (h.has_prev) {
    work_arrays += 1;                                  // sigma
    state_arrays += 1 + tab;                           // E_{i-1}, acc or psi_hat written
    if (!h.prev_is_first) state_arrays += 1;           // acc read
    if (h.update_prev) state_arrays += 2 * tab;        // E1 read, E_{i-1} written
    if (h.stash_prev) state_arrays += 1;               // N_{i-1} stashed
  }
  if (h.prev_is_first || !h.final_stage) state_arrays += 1;  // psi_hat read
  if (!h.final_stage || h.chain_next) {
    state_arrays += tab;                               // E_i (E_0 of the next step when chained)
    work_arrays += 1;                                  // row-inverse output
    for (int e = 0; e < h.n_a; e++)
      if (h.a_idx[e] != h.prev_stage) state_arrays += 1;  // stashed N_j
  }
  const double hb = MN * (cs * state_arrays + cb * work_arrays + (h.recompute_e ? 8.0 : 0.0));
  return run_kernel(s, i == P.rk.s ? "row_close_step" : "row_stage_head", hb * s->F,
                    [&] { return launch_row_head<W>(s->g, s->tw, h, s->batch, (int)s->F, s->ctx->stream); });
}

mmp_status enqueue_head(mmp_solver* s, int i) {
  return s->mixed ? enqueue_head_w<float>(s, i) : enqueue_head_w<double>(s, i);
}

// One step. With s->chain the step's stage-0 head was issued by the previous step's close
// (or by advance_impl for the first step of a call).
template <class W>
mmp_status enqueue_step_w(mmp_solver* s) {
  using cw = cpx<W>;
  const Problem& P = s->P;
  const Geometry& g = s->g;
  hipStream_t st = s->ctx->stream;
  cw* ybuf = static_cast<cw*>(s->ybuf_w);
  cw* U = static_cast<cw*>(s->U_w);
  cw* conv = static_cast<cw*>(s->conv_w);
  cw* convt = static_cast<cw*>(s->convt_w);
  const cw* hspec = static_cast<const cw*>(s->hspec_w);
  const W* qdense = static_cast<const W*>(s->qdense_w);
  const double cb = sizeof(cw);  // bytes per complex work value
  const double cs = 16.0;        // bytes per complex FP64 state value
  const double MN = (double)P.M * P.N;
  const double LN = (double)s->n_lines * P.N;
  const double invN = 1.0 / P.N;
  for (int i = 0; i <= P.rk.s; i++) {
    if (i > 0 || !s->chain) MMP_TRY(enqueue_head_w<W>(s, i));
    if (i == P.rk.s) break;
    const int F = (int)s->F;
    if (s->use_fused) {
      MMP_TRY(run_kernel(s, "col_fields_pairs", cb * (2 * MN + LN) * F, [&] {
        return launch_colstage_fields_pairs<W>(g, s->tw, ybuf, U, conv, s->n_lines, s->pairs, invN, s->batch, F,
                                               s->pair_columns, s->pair_groups, s->conv_cols, st);
      }));
      MMP_TRY(run_kernel(s, "row_raman", cb * (2 * LN + P.N) * F, [&] {
        return launch_row_raman_cols<W>(g, s->tw, conv, convt, s->n_lines, hspec, s->batch, F, s->conv_cols, st);
      }));
      MMP_TRY(run_kernel(s, "col_raman_gather", cb * (2 * MN + LN) * F, [&] {
        return launch_colstage_raman_gather<W>(g, s->tw, U, convt, ybuf, s->gather, s->fused, qdense, invN, s->batch,
                                               F, st);
      }));
      continue;
    }
    MMP_TRY(run_kernel(s, "col_inverse_fields", 2 * cb * MN,
                       [&] { return launch_col_inverse<W>(g, s->tw, ybuf, U, P.M, invN, 1.0, st); }));
    if (P.raman) {
      MMP_TRY(run_kernel(s, "col_pairs_forward", cb * (MN + LN), [&] {
        return launch_col_pairs_forward<W>(g, s->tw, U, conv, s->n_lines, s->pairs, st);
      }));
      MMP_TRY(run_kernel(s, "row_raman", cb * (2 * LN + P.N),
                         [&] { return launch_row_raman<W>(g, s->tw, conv, s->n_lines, hspec, s->batch, 1, st); }));
      MMP_TRY(run_kernel(s, "col_inverse_raman", 2 * cb * LN, [&] {
        return launch_col_inverse<W>(g, s->tw, conv, conv, s->n_lines, invN, 1.0, st);
      }));
    }
    MMP_TRY(run_kernel(s, "gather_kerr_raman", cb * (2 * MN + LN), [&] {
      return qdense ? launch_gather_dense<W>(g, U, conv, ybuf, s->gather, qdense, st)
                    : launch_gather<W>(g, U, conv, ybuf, s->gather, st);
    }));
    MMP_TRY(run_kernel(s, "col_forward_sigma", 2 * cb * MN,
                       [&] { return launch_col_forward<W>(g, s->tw, ybuf, ybuf, P.M, st); }));
  }
  return MMP_OK;
}

mmp_status enqueue_step(mmp_solver* s) { return s->mixed ? enqueue_step_w<float>(s) : enqueue_step_w<double>(s); }

mmp_status capture_steps(mmp_solver* s, int steps, hipGraph_t* graph_out, hipGraphExec_t* exec_out) {
  hipStream_t st = s->ctx->stream;
  HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
  mmp_status status = MMP_OK;
  for (int k = 0; k < steps && status == MMP_OK; k++) status = enqueue_step(s);
  hipGraph_t graph = nullptr;
  hipError_t e = hipStreamEndCapture(st, &graph);
  if (status != MMP_OK) {
    if (graph) hipGraphDestroy(graph);
    return status;
  }
  if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
  hipGraphExec_t exec = nullptr;
  e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    hipGraphDestroy(graph);
    return hip_fail(e, "hipGraphInstantiate");
  }
  // upload once here, so that the first replay does not carry it
  e = hipGraphUpload(exec, st);
  if (e != hipSuccess) {
    hipGraphExecDestroy(exec);
    hipGraphDestroy(graph);
    return hip_fail(e, "hipGraphUpload");
  }
  *graph_out = graph;
  *exec_out = exec;
  return MMP_OK;
}

void drop_graph(mmp_solver* s) {
  if (s->graph_exec) hipGraphExecDestroy(s->graph_exec);
  if (s->graph) hipGraphDestroy(s->graph);
  if (s->graph1_exec) hipGraphExecDestroy(s->graph1_exec);
  if (s->graph1) hipGraphDestroy(s->graph1);
  s->graph_exec = nullptr;
  s->graph = nullptr;
  s->graph1_exec = nullptr;
  s->graph1 = nullptr;
}

// The replayable graphs of kGraphSteps steps and of one step. Built when the solver is created
// (and again after the autotuner changes the kernels' plan), so no call that steps pays for them.
// When capture is unavailable the solver keeps launching kernel by kernel.
mmp_status build_graph(mmp_solver* s) {
  if (!s->graphs || s->graph_exec) return MMP_OK;
  mmp_status st = capture_steps(s, kGraphSteps, &s->graph, &s->graph_exec);
  if (st == MMP_OK) st = capture_steps(s, 1, &s->graph1, &s->graph1_exec);
  if (st != MMP_OK) {
    drop_graph(s);
    s->graphs = false;
    hipGetLastError();
  }
  return MMP_OK;
}

mmp_status set_device(mmp_solver* s) {
  HIP_TRY(hipSetDevice(s->ctx->device), "hipSetDevice");
  return MMP_OK;
}

// initLawsonRK (src/MultimodePropagation.cpp:158-163): z = 0 and fresh Lawson tables.
// z = 0 and the initial Lawson tables for every fiber (zeroed on the host: no memset node).
mmp_status refresh_halted(mmp_solver* s);

mmp_status reset_z(mmp_solver* s) {
  MMP_TRY(refresh_halted(s));
  std::vector<StepState> zero(s->F, StepState{0., 0u, 0u, 0.});
  for (uint32_t f = 0; f < s->F; f++) zero[f].halted = s->halted[f];  // a non-finite field stays parked
  MMP_TRY(upload(s, s->st, zero.data(), zero.size() * sizeof(StepState)));
  if (s->etab)
    HIP_TRY(hipMemcpyAsync(s->etab, s->etab0, (size_t)s->F * s->ntab * s->P.M * s->P.N * sizeof(cd),
                           hipMemcpyDeviceToDevice, s->ctx->stream),
            "hipMemcpyAsync D2D");
  HIP_TRY(hipStreamSynchronize(s->ctx->stream), "hipStreamSynchronize");  // `zero` is a host staging copy
  std::fill(s->z_host.begin(), s->z_host.end(), 0.);
  std::fill(s->steps.begin(), s->steps.end(), 0u);
  return MMP_OK;
}

// The device parks a fiber whose state stops being finite (StepState::halted, FiberParams::live):
// bring the host's view of live, z and the step count in line with it. Waits for the stream.
mmp_status refresh_halted(mmp_solver* s) {
  std::vector<StepState> st(s->F);
  HIP_TRY(hipMemcpyAsync(st.data(), s->st, s->F * sizeof(StepState), hipMemcpyDeviceToHost, s->ctx->stream),
          "hipMemcpyAsync D2H");
  HIP_TRY(hipStreamSynchronize(s->ctx->stream), "hipStreamSynchronize");
  for (uint32_t f = 0; f < s->F; f++)
    if (st[f].halted && !s->halted[f]) {
      s->halted[f] = 1;
      s->live[f] = 0;
      s->z_host[f] = st[f].z;
      s->steps[f] = st[f].steps;
    }
  return MMP_OK;
}

// z = 0 and the initial Lawson tables for fiber f only.
mmp_status reset_fiber(mmp_solver* s, uint32_t f) {
  static const StepState zero{0., 0u, 0u, 0.};
  MMP_TRY(upload(s, s->st + f, &zero, sizeof zero));
  if (s->etab)
    HIP_TRY(hipMemcpyAsync(s->etab + f * s->batch.se, s->etab0 + f * s->batch.se, s->batch.se * sizeof(cd),
                           hipMemcpyDeviceToDevice, s->ctx->stream),
            "hipMemcpyAsync D2D");
  s->z_host[f] = 0.;
  s->steps[f] = 0;
  return MMP_OK;
}

// psi_hat of fiber f = F(psi) for a time-domain field already in s->U (its first M x NT values)
mmp_status forward_state_from_U(mmp_solver* s, uint32_t f) {
  hipStream_t st = s->ctx->stream;
  cd* ph = s->psi_hat + (size_t)f * s->batch.sf;
  HIP_TRY(launch_col_forward<double>(s->g, s->tw, s->U, ph, s->P.M, st), "col_forward");
  HIP_TRY(launch_row_forward(s->g, s->tw, ph, s->P.M, st), "row_forward");
  return MMP_OK;
}

// time-domain field of fiber f's psi_hat into the first M x NT values of s->U: with_phase ->
// exp(L z) applied and divided by the fiber's conversion factor
mmp_status inverse_state_to_U(mmp_solver* s, uint32_t f, bool with_phase) {
  hipStream_t st = s->ctx->stream;
  const size_t off = (size_t)f * s->batch.sf;
  HIP_TRY(launch_row_inverse_phase(s->g, s->tw, s->psi_hat + off, s->ybuf, s->imL + off, s->st + f,
                                   with_phase ? 1 : 0, st),
          "row_inverse_phase");
  HIP_TRY(launch_col_inverse<double>(s->g, s->tw, s->ybuf, s->U, s->P.M, 1.0 / s->P.N,
                                     with_phase ? s->fib[f].conversion_factor : 1.0, st),
          "col_inverse");
  return MMP_OK;
}

void release(mmp_solver* s) {
  if (!s) return;
  if (s->ctx) hipSetDevice(s->ctx->device);
  if (s->ctx && s->ctx->stream) hipStreamSynchronize(s->ctx->stream);
  for (auto& r : s->pending) {
    hipEventDestroy(r.start);
    hipEventDestroy(r.stop);
  }
  for (auto e : s->free_events) hipEventDestroy(e);
  drop_graph(s);
  for (void* p : s->allocations) hipFree(p);
  delete s;
}

// Lawson tables E_u (one per distinct c_u, advanced by E1 each step) exist in FP64 only: the
// mixed-precision stage heads form E_u from Im L and the fiber's z where they use it.
bool lawson_tables(const Problem& P, bool mixed) { return !(mixed && P.rk.s >= 2); }

// Device bytes create_impl allocates for F fibers of problem P (kept in step with its dalloc calls).
size_t planned_device_bytes(const Problem& P, const mmp_options& opt, size_t F) {
  auto c = [](size_t n) { return n ? n : (size_t)1; };
  const bool mixed = opt.precision == MMP_PRECISION_MIXED;
  const size_t MN = (size_t)P.M * P.N, L = P.line_a.size(), Fw = mixed ? 1 : F;
  bool any_stash = false;
  for (int j = 0; j < P.rk.s; j++) any_stash = any_stash || P.stash_needed[j];
  size_t b = 16 * MN * (2 * F + 2 * Fw) + 8 * MN * F + 16 * (P.N2 + P.N1 + P.N) + sizeof(StepState) * F +
             sizeof(FiberParams) * F + 8 * kHeadCoef * ((size_t)P.rk.s + 1) * F + 8 * c(P.M);
  if (lawson_tables(P, mixed)) b += 16 * MN * F * (2 * P.c_unique.size() + 1);
  if (any_stash) b += 16 * (size_t)P.rk.s * MN * F;
  if (P.raman) b += (mixed ? 8 : 16) * c(L) * P.N * F + 16 * (size_t)P.N * F + 8 * (size_t)P.N;
  b += 4 * c(P.coff.size()) + 4 * c(P.nnz) + 8 * c(P.nnz) + 4 * 4 * c(P.P) + 4 * 3 * c(L);
  if (mixed) {
    b += 8 * (2 * MN) * F + 8 * (P.N2 + P.N1 + P.N) + 4 * c(P.nnz);
    if (P.raman) b += 8 * (size_t)P.N * F;
  }
  if (P.fused_chunks > 0 && !(opt.flags & MMP_FLAG_SEPARATE_COLUMNS)) {
    b += 4 * (2 * (P.line_a.size() + 1) + c(P.P));  // chunk bounds sized for the finest plan, slot metadata
    if (P.raman) b += (mixed ? 8 : 16) * c(L) * P.N * F;  // the Raman lines in column order
  }
  const bool fused = P.fused_chunks > 0 && !(opt.flags & MMP_FLAG_SEPARATE_COLUMNS);
  const size_t block = (fused && P.parity) ? (size_t)P.M * ((P.M + 1) / 2) : (size_t)P.M * P.M;
  if (fused || dense_gather_chunk((int)P.M) > 0) b += (mixed ? 4 : 8) * (size_t)P.P * block * F;
  return b;
}

void fill_info(const Problem& P, const mmp_options& opt, mmp_info* o) {
  std::memset(o, 0, sizeof *o);
  o->n_modes = P.M;
  o->nt = P.N;
  o->stages = (uint32_t)P.rk.s;
  o->n_pairs = P.P;
  o->nnz = P.nnz;
  o->n_conv_lines = (uint32_t)P.line_a.size();
  o->n1 = P.N1;
  o->n2 = P.N2;
  o->default_steps = P.default_steps;
  o->t_final = P.t_final;
  o->delta_t = P.delta_t;
  o->z_final = P.z_final;
  o->delta_z = P.delta_z;
  o->conversion_factor = P.conversion_factor;
  o->sign_gamma = P.sign_gamma;
  o->ld = P.ld;
  o->max_beta2 = P.max_beta2;
  o->max_q = P.max_q;
  o->device_bytes = planned_device_bytes(P, opt, 1);
  o->n_fibers = 1;
  o->live = 1;
  const bool fused = P.fused_chunks > 0 && !(opt.flags & MMP_FLAG_SEPARATE_COLUMNS);
  o->gather_waves = fused ? (P.M <= 6 ? 8u : 4u) : 0u;  // the defaults, before any tuning
  o->chunk_lines = fused ? (uint32_t)colstage_lines_per_chunk((int)P.M) : 0u;
  if (fused) {  // a solver of one fiber (a batch may choose differently; results do not depend on it)
    int columns = 1, groups = 1;
    colstage_pairs_plan((int)P.M, (int)P.N2, (int)P.line_a.size(), 1, opt.precision == MMP_PRECISION_MIXED, columns,
                        groups);
    o->pair_columns = (uint32_t)columns;
    o->pair_groups = (uint32_t)groups;
  }
}

// Fibers of one batch must share everything the kernels take from fiber 0's problem.
bool same_structure(const Problem& a, const Problem& b) {
  return a.M == b.M && a.N == b.N && a.N1 == b.N1 && a.rk.s == b.rk.s && a.raman == b.raman && a.P == b.P &&
         a.stage_table == b.stage_table && a.stash_needed == b.stash_needed && a.c_unique == b.c_unique &&
         a.slot_a == b.slot_a && a.slot_b == b.slot_b && a.slot_line == b.slot_line && a.slot_kind == b.slot_kind &&
         a.line_a == b.line_a && a.line_b == b.line_b && a.line_self == b.line_self &&
         a.line_slot0 == b.line_slot0 && a.parity == b.parity && a.slot_parity == b.slot_parity;
}

// Dense coupling blocks Qd[s][p][k] = sum of Q / max|Q| over the nonzeros (k, slot s) of Q^(p);
// compact (parity-structured couplings): only k = 2 kk + ((p + parity_s) % 2), as [s][p][kk].
std::vector<double> dense_blocks(const Problem& P, bool compact) {
  std::vector<double> qd((size_t)P.P * P.M * P.M, 0.);
  for (uint32_t p = 0; p < P.M; p++)
    for (uint32_t e = P.off[p]; e < P.off[p + 1]; e++) qd[((size_t)P.slot[e] * P.M + p) * P.M + P.k[e]] += P.q[e];
  if (!compact) return qd;
  const uint32_t KH = (P.M + 1) / 2;
  std::vector<double> qc((size_t)P.P * P.M * KH, 0.);
  for (uint32_t s = 0; s < P.P; s++)
    for (uint32_t p = 0; p < P.M; p++)
      for (uint32_t kk = 0; kk < KH; kk++) {
        const uint32_t k = 2 * kk + ((p + P.slot_parity[s]) & 1u);
        if (k < P.M) qc[((size_t)s * P.M + p) * KH + kk] = qd[((size_t)s * P.M + p) * P.M + k];
      }
  return qc;
}

// Uploads fiber f's own data (operators, couplings, Raman spectrum, initial state, scalars),
// resets its z and Lawson tables, then frees the large host arrays of its problem.
mmp_status load_fiber(mmp_solver* s, uint32_t f, Problem& Pf) {
  hipStream_t stream = s->ctx->stream;
  const Batch& b = s->batch;
  const size_t MN = (size_t)Pf.M * Pf.N;
  MMP_TRY(upload(s, s->imL + f * b.sf, Pf.imL_fourstep.data(), MN * 8));
  if (s->etab) {
    MMP_TRY(upload(s, s->e1 + f * b.sf, Pf.e1.data(), MN * 16));
    MMP_TRY(upload(s, s->etab0 + f * b.se, Pf.etab0.data(), Pf.etab0.size() * sizeof(double)));
    MMP_TRY(upload(s, s->etab + f * b.se, Pf.etab0.data(), Pf.etab0.size() * sizeof(double)));
  }
  std::vector<double> qd;
  std::vector<float> qd32;
  if (s->qdense_w) {
    qd = dense_blocks(Pf, s->fused.parity != 0);
    if (s->mixed) {
      qd32.assign(qd.begin(), qd.end());
      MMP_TRY(upload(s, static_cast<float*>(const_cast<void*>(s->qdense_w)) + f * b.sq, qd32.data(), qd32.size() * 4));
    } else {
      MMP_TRY(upload(s, s->qdense + f * b.sq, qd.data(), qd.size() * 8));
    }
  }
  const FiberParams fp{Pf.delta_z, Pf.sign_gamma, Pf.delta_t, Pf.kv, Pf.kr, 1u, 0u};
  MMP_TRY(upload(s, s->d_fp + f, &fp, sizeof fp));
  // b_{i-1} h and a_ij h of every head, multiplied in FP64 on the host as the single solver does
  std::vector<double> coef(b.nheads * kHeadCoef, 0.);
  for (size_t i = 0; i < b.nheads; i++) {
    const HeadArgs& a = s->heads[i];
    coef[i * kHeadCoef] = a.b_prev != 0. ? a.b_prev * Pf.delta_z : 0.;
    for (int e = 0; e < a.n_a; e++) coef[i * kHeadCoef + 1 + e] = a.a_coef[e] * Pf.delta_z;
  }
  MMP_TRY(upload(s, s->d_coef + f * b.nheads * kHeadCoef, coef.data(), coef.size() * sizeof(double)));
  const StepState zero{0., 0u, 0u, 0.};
  MMP_TRY(upload(s, s->st + f, &zero, sizeof zero));
  if (Pf.raman) {
    cd* hs = s->hspec + f * b.sh;
    MMP_TRY(upload(s, s->hR, Pf.raman_response.data(), Pf.N * 8));
    HIP_TRY(launch_col_forward_real(s->g, s->tw, s->hR, hs, stream), "col_forward_real");
    HIP_TRY(launch_row_forward(s->g, s->tw, hs, 1, stream), "row_forward");
    if (s->mixed) HIP_TRY(launch_narrow(hs, static_cast<cf*>(s->hspec_w) + f * b.sh, Pf.N, stream), "narrow");
  }
  MMP_TRY(upload(s, s->U, Pf.psi0.data(), MN * 16));
  MMP_TRY(forward_state_from_U(s, f));
  HIP_TRY(hipStreamSynchronize(stream), "hipStreamSynchronize");  // host arrays above are staging copies
  s->z_host[f] = 0.;
  s->steps[f] = 0;
  s->live[f] = 1;
  s->halted[f] = 0;
  for (auto* v : {&Pf.imL_fourstep, &Pf.e1, &Pf.etab0, &Pf.psi0, &Pf.raman_response}) std::vector<double>().swap(*v);
  return MMP_OK;
}

mmp_status autotune(mmp_solver* s);

// arena (optional): carve every device array from [arena, arena + arena_cap) instead of
// allocating; such a solver launches kernel by kernel (no graph instantiation) and takes neither
// MMP_FLAG_PROFILE nor MMP_FLAG_AUTOTUNE (both allocate).
mmp_status create_impl(mmp_ctx* ctx, uint32_t n_fibers, const mmp_params* params, const mmp_options* options,
                       mmp_solver** out, unsigned char* arena = nullptr, size_t arena_cap = 0) {
  if (!ctx || !out || !params) return fail(MMP_ERR_INVALID_ARGUMENT, "ctx, params and out are required");
  *out = nullptr;
  if (n_fibers < 1 || n_fibers > 65535) return fail(MMP_ERR_INVALID_ARGUMENT, "n_fibers must be in [1, 65535]");
  mmp_options opt{};
  if (options) opt = *options;
  mmp_solver* s = new (std::nothrow) mmp_solver();
  if (!s) return fail(MMP_ERR_OUT_OF_MEMORY, "host allocation failed");
  s->ctx = ctx;
  s->opt = opt;
  s->F = n_fibers;
  s->arena = arena;
  s->arena_cap = arena_cap;
  s->fib.resize(n_fibers);
  std::string err;
  for (uint32_t f = 0; f < n_fibers; f++) {
    mmp_status stt = prepare_problem(&params[f], opt, s->fib[f], err);
    if (stt == MMP_OK && f > 0 && !same_structure(s->fib[0], s->fib[f])) {
      stt = MMP_ERR_INVALID_ARGUMENT;
      err = "fibers of one batch must share n_modes, nt, method_order, raman and the pair structure of their "
            "coupling lists";
    }
    if (stt != MMP_OK) {
      delete s;
      return fail(stt, n_fibers > 1 ? "fiber " + std::to_string(f) + ": " + err : err);
    }
    s->time_window.push_back(params[f].time_window);
    s->fiber_length.push_back(params[f].fiber_length);
  }
  s->z_host.assign(n_fibers, 0.);
  s->steps.assign(n_fibers, 0u);
  s->live.assign(n_fibers, 1);
  s->halted.assign(n_fibers, 0);
  s->P = s->fib[0];  // structure of the batch (fiber 0's large arrays are uploaded from s->fib[0])
  for (auto* v : {&s->P.imL_fourstep, &s->P.e1, &s->P.etab0, &s->P.psi0, &s->P.raman_response})
    std::vector<double>().swap(*v);
  s->mixed = opt.precision == MMP_PRECISION_MIXED;
  s->wbytes = s->mixed ? sizeof(cf) : sizeof(cd);
  s->profile = (opt.flags & MMP_FLAG_PROFILE) != 0;
  s->graphs = !s->profile && (opt.flags & MMP_FLAG_NO_GRAPH) == 0 && !s->arena;
  Problem& P = s->P;
  const uint32_t F = n_fibers;
  s->g = Geometry{(int)P.N, (int)P.N1, (int)P.N2, (int)P.M};
  s->n_lines = (int)P.line_a.size();
  s->use_fused = P.fused_chunks > 0 && !(opt.flags & MMP_FLAG_SEPARATE_COLUMNS);
  if (F > 1 && !s->use_fused) {
    delete s;
    return fail(MMP_ERR_UNSUPPORTED,
                "batched fibers need the column-stage kernels: Raman on, nt in [4096, 65536], n_modes <= 16, "
                "and no MMP_FLAG_SEPARATE_COLUMNS");
  }
  auto cleanup = [&](mmp_status st2) {
    release(s);
    return st2;
  };
  if (set_device(s) != MMP_OK) return cleanup(MMP_ERR_DEVICE);
  const size_t MN = (size_t)P.M * P.N;
  s->ntab = P.c_unique.size();
  const size_t conv_n = (size_t)(s->n_lines ? s->n_lines : 1) * P.N;
  bool any_stash = false;
  for (int j = 0; j < P.rk.s; j++) any_stash = any_stash || P.stash_needed[j];
  // dense blocks: the column-stage gather always takes them, the separate gather up to its own limit
  const bool dense = s->use_fused || dense_gather_chunk((int)P.M) > 0;
  s->batch.sf = MN;
  s->batch.sl = conv_n;
  s->batch.se = s->ntab * MN;
  s->batch.ss = (size_t)P.rk.s * MN;
  const bool compact = s->use_fused && P.parity;
  s->batch.sq = (size_t)P.P * P.M * (compact ? (P.M + 1) / 2 : P.M);
  s->fused.parity = compact ? 1 : 0;
  s->batch.sh = P.N;
  // device buffers: per-fiber arrays are [F][...]; ybuf and U are also the FP64 scratch of
  // getResult, set_state and energy (their first M x NT values)
  cd *rowtw, *coltw, *fourtw;
  mmp_status st = MMP_OK;
  auto A = [&](mmp_status x) {
    if (st == MMP_OK) st = x;
  };
  const size_t Fw = s->mixed ? 1 : F;  // FP64 ybuf/U copies: the work arrays in FP64 mode, scratch in mixed
  A(dalloc(s, &s->psi_hat, F * MN));
  A(dalloc(s, &s->acc, F * MN));
  A(dalloc(s, &s->ybuf, Fw * MN));
  A(dalloc(s, &s->U, Fw * MN));
  A(dalloc(s, &s->imL, F * MN));
  if (lawson_tables(P, s->mixed)) {
    A(dalloc(s, &s->etab, F * s->batch.se));
    A(dalloc(s, &s->etab0, F * s->batch.se));
    A(dalloc(s, &s->e1, F * MN));
  }
  A(dalloc(s, &rowtw, P.N2));
  A(dalloc(s, &coltw, P.N1));
  A(dalloc(s, &fourtw, P.N));
  A(dalloc(s, &s->st, F));
  A(dalloc(s, &s->d_fp, F));
  s->batch.nheads = (size_t)P.rk.s + 1;
  A(dalloc(s, &s->d_coef, F * s->batch.nheads * kHeadCoef));
  A(dalloc(s, &s->energy, P.M));
  if (any_stash) A(dalloc(s, &s->stash, F * s->batch.ss));
  if (P.raman) {
    A(dalloc(s, &s->hspec, F * (size_t)P.N));
    A(dalloc(s, &s->hR, P.N));
  }
  cf *rowtw32 = nullptr, *coltw32 = nullptr, *fourtw32 = nullptr;
  if (s->mixed) {
    cf *yw, *uw, *cw = nullptr, *hw = nullptr;
    A(dalloc(s, &yw, F * MN));
    A(dalloc(s, &uw, F * MN));
    if (P.raman) {
      A(dalloc(s, &cw, F * conv_n));
      A(dalloc(s, &hw, F * (size_t)P.N));
      if (s->use_fused) {
        cf* ct;
        A(dalloc(s, &ct, F * conv_n));
        s->convt_w = ct;
      }
    }
    A(dalloc(s, &rowtw32, P.N2));
    A(dalloc(s, &coltw32, P.N1));
    A(dalloc(s, &fourtw32, P.N));
    s->ybuf_w = yw;
    s->U_w = uw;
    s->conv_w = cw;
    s->hspec_w = hw;
    if (dense) {
      float* d;
      A(dalloc(s, &d, F * s->batch.sq));
      s->qdense_w = d;
    }
  } else {
    if (P.raman) A(dalloc(s, &s->conv, F * conv_n));
    if (P.raman && s->use_fused) {
      cd* ct;
      A(dalloc(s, &ct, F * conv_n));
      s->convt_w = ct;
    }
    s->ybuf_w = s->ybuf;
    s->U_w = s->U;
    s->conv_w = s->conv;
    s->hspec_w = s->hspec;
    if (dense) {
      A(dalloc(s, &s->qdense, F * s->batch.sq));
      s->qdense_w = s->qdense;
    }
  }
  uint32_t *d_coff, *d_k, *d_sa, *d_sb, *d_sl, *d_sk, *d_la, *d_lb, *d_ls;
  double* d_q;
  const size_t nnz1 = P.nnz ? P.nnz : 1, P1 = P.P ? P.P : 1, L1 = s->n_lines ? s->n_lines : 1;
  A(dalloc(s, &d_coff, P.coff.size()));
  A(dalloc(s, &d_k, nnz1));
  A(dalloc(s, &d_q, nnz1));
  A(dalloc(s, &d_sa, P1));
  A(dalloc(s, &d_sb, P1));
  A(dalloc(s, &d_sl, P1));
  A(dalloc(s, &d_sk, P1));
  A(dalloc(s, &d_la, L1));
  A(dalloc(s, &d_lb, L1));
  A(dalloc(s, &d_ls, L1));
  if (st != MMP_OK) return cleanup(st);
  // batch-wide tables (host copies stay alive until the stream synchronisation below)
  std::vector<uint32_t> kslot(P.nnz);
  std::vector<double> tr = twiddle_table(P.N2), tc = twiddle_table(P.N1), t4 = twiddle_table(P.N);
  A(upload(s, rowtw, tr.data(), tr.size() * 8));
  A(upload(s, coltw, tc.data(), tc.size() * 8));
  A(upload(s, fourtw, t4.data(), t4.size() * 8));
  // FP32 twiddles: the FP64 values rounded once
  std::vector<float> tr32(tr.begin(), tr.end()), tc32(tc.begin(), tc.end()), t432(t4.begin(), t4.end());
  if (s->mixed) {
    A(upload(s, rowtw32, tr32.data(), tr32.size() * 4));
    A(upload(s, coltw32, tc32.data(), tc32.size() * 4));
    A(upload(s, fourtw32, t432.data(), t432.size() * 4));
  }
  A(upload(s, d_coff, P.coff.data(), P.coff.size() * 4));
  if (P.nnz) {
    for (uint32_t e = 0; e < P.nnz; e++) kslot[e] = P.k[e] | (P.slot[e] << 8);
    A(upload(s, d_k, kslot.data(), P.nnz * 4));
    A(upload(s, d_q, P.q.data(), P.nnz * 8));
  }
  if (P.P) {
    A(upload(s, d_sa, P.slot_a.data(), P.P * 4));
    A(upload(s, d_sb, P.slot_b.data(), P.P * 4));
    A(upload(s, d_sl, P.slot_line.data(), P.P * 4));
    A(upload(s, d_sk, P.slot_kind.data(), P.P * 4));
  }
  if (s->n_lines) {
    A(upload(s, d_la, P.line_a.data(), s->n_lines * 4));
    A(upload(s, d_lb, P.line_b.data(), s->n_lines * 4));
    A(upload(s, d_ls, P.line_self.data(), s->n_lines * 4));
  }
  std::vector<float> q32(P.q.begin(), P.q.end());
  float* d_q32 = nullptr;
  if (s->mixed) {
    A(dalloc(s, &d_q32, nnz1));
    if (P.nnz) A(upload(s, d_q32, q32.data(), P.nnz * 4));
  }
  if (st != MMP_OK) return cleanup(st);
  s->tw = Twiddles{rowtw, coltw, fourtw, rowtw32, coltw32, fourtw32};
  s->q_w = s->mixed ? (const void*)d_q32 : (const void*)d_q;
  s->gather = GatherArgs{d_coff, d_k, s->q_w, d_sa, d_sb, d_sl, d_sk, (int)P.P, P.gather_chunk,
                         P.gather_chunks, P.gather_cap, P.raman, P.kv, P.kr};
  s->pairs = PairLineArgs{d_la, d_lb, d_ls};
  s->batch.fp = s->d_fp;
  s->batch.coef = s->d_coef;
  s->batch.st = s->st;
  std::vector<uint32_t> slot_meta(P.P);
  if (s->use_fused) {
    uint32_t *d_cl, *d_cs, *d_meta;
    const size_t max_chunks = P.line_a.size() + 1;  // room for the finest chunking the tuner may pick
    A(dalloc(s, &d_cl, max_chunks));
    A(dalloc(s, &d_cs, max_chunks));
    A(dalloc(s, &d_meta, P1));
    if (st != MMP_OK) return cleanup(st);
    for (uint32_t j = 0; j < P.P; j++)
      slot_meta[j] = P.slot_a[j] | (P.slot_b[j] << 8) | (P.slot_kind[j] << 16) | ((uint32_t)P.slot_parity[j] << 18) |
                     (P.slot_line[j] << 19);
    A(upload(s, d_cl, P.fused_line.data(), P.fused_line.size() * 4));
    A(upload(s, d_cs, P.fused_slot.data(), P.fused_slot.size() * 4));
    A(upload(s, d_meta, slot_meta.data(), P.P * 4));
    // eight gather wavefronts from 4 to 6 modes (two per SIMD hide the slot loop's LDS latency);
    // below 4 modes a column has at most 6 Raman lines and 9 slots, which four wavefronts cover
    // (measured, 29 fibers of 3 modes at NT = 16384: 71 us with eight, 38 us with four)
    s->fused = FusedArgs{d_cl,   d_cs, d_meta, P.fused_chunks, P.fused_max_lines, P.fused_max_slots, compact ? 1 : 0,
                         P.M >= 4 && P.M <= 6 ? 8 : 4};
    s->chunk_lines = colstage_lines_per_chunk((int)P.M);
    colstage_pairs_plan((int)P.M, s->g.N2, s->n_lines, (int)s->F, s->mixed, s->pair_columns, s->pair_groups);
    // grids of 8 or more workgroups per CU
    s->fused.hi_occupancy = colstage_hi_occupancy_ok((int)P.M, s->mixed, compact) && (long)s->g.N2 * s->F >= 2048;
    // the coupling sums on the matrix cores (FP32) by default at 4 to 10 and 13 to 16 modes,
    // whatever the number of fibers (a batch runs the arithmetic of its fibers alone). Measured on
    // one MI355X, gather per launch against the vector gather: 7-8 modes 1.03-1.23x (one 32-row
    // tile, three workgroups per CU); 4-6 modes 1.06-1.31x on grids of 2048 workgroups and more,
    // 0.84-0.90x on small ones (against the eight-wavefront vector gather); 9-10 modes 1.04-1.19x
    // on the grids of one fiber or a few, 0.84-0.92x on large batches, where two tiles hold the
    // matrix-core gather to two workgroups per CU against three for the vector gather (the tuner
    // picks per batch); 11-12 modes 0.89-1.05x; 13-16 modes 1.03-1.25x on every grid (three or
    // four tiles, the wavefront pairs splitting them).
    {
      const int M = (int)P.M;
      const bool want = (opt.flags & MMP_FLAG_MATRIX_CORES) || (M >= 4 && M <= 10) || M >= 13;
      s->fused.matrix_cores =
          !(opt.flags & MMP_FLAG_NO_MATRIX_CORES) && want && colstage_matrix_cores_ok(M, s->mixed, compact) ? 1 : 0;
    }
    if (s->fused.matrix_cores) {
      s->fused.waves = 4;
      s->fused.hi_occupancy = 0;
    }
    // lines in column order before the Raman row pass too: its LDS staging costs latency that
    // only large grids hide (measured: batched M = 10, NT = 16384 x 36 fibers row pass 140 ->
    // 111 us; one grin6 fiber 7.0 -> 7.6 us)
    s->conv_cols = (long)s->g.N2 * s->F >= 2048 ? 1 : 0;
  }
  // stage heads: tableau coefficients here, each fiber's h, sgn and tables in the kernel
  // the close head also builds the next step's stage 0 unless the stage-0 Lawson table is
  // advanced by that same head (no tableau here has c_i = 0 beyond stage 0)
  s->chain = P.table_last_head[P.stage_table[0]] < P.rk.s;
  for (int i = 0; i <= P.rk.s; i++) {
    HeadArgs a{};
    a.head = i;
    a.ysig = s->ybuf_w;
    a.psi_hat = s->psi_hat;
    a.acc = s->acc;
    a.stash = s->stash;
    a.yout = s->ybuf_w;
    a.e1 = s->e1;  // null without tables
    a.scale_cur = 1.0;
    a.has_prev = i > 0;
    a.prev_stage = i - 1;
    a.prev_is_first = (i == 1);
    a.final_stage = (i == P.rk.s);
    // mixed precision: Lawson factors from Im L and the step count (no tables exist); the close
    // head takes the count at the step's start from StepState::n0, which head 1 records
    if (!lawson_tables(P, s->mixed)) {
      a.recompute_e = 1;
      a.imL = s->imL;
      a.set_n0 = (i == 1);
      if (i > 0) a.c_prev = P.rk.c[i - 1];
      a.c_cur = (i < P.rk.s) ? P.rk.c[i] : P.rk.c[0];
    }
    if (i > 0) {
      const int u = P.stage_table[i - 1];
      a.e_prev = s->etab ? s->etab + (size_t)u * MN : nullptr;
      a.update_prev = (P.table_last_head[u] == i);
      a.b_prev = P.rk.b[i - 1];
      a.stash_prev = P.stash_needed[i - 1];
    }
    if (i == P.rk.s) {
      a.advance_z = 1;
      if (s->chain) {
        a.chain_next = 1;
        a.e_cur = s->etab ? s->etab + (size_t)P.stage_table[0] * MN : nullptr;
      }
    }
    if (i < P.rk.s) {
      a.e_cur = s->etab ? s->etab + (size_t)P.stage_table[i] * MN : nullptr;
      for (int j = 0; j < i; j++)
        if (P.rk.a[i][j] != 0.) {
          a.a_idx[a.n_a] = j;
          a.a_coef[a.n_a] = P.rk.a[i][j];
          a.n_a++;
        }
    }
    s->heads.push_back(a);
  }
  if (st == MMP_OK) {
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) A(hip_fail(e, "hipStreamSynchronize"));
  }
  for (uint32_t f = 0; f < F && st == MMP_OK; f++) A(load_fiber(s, f, s->fib[f]));
  if (st == MMP_OK && (opt.flags & MMP_FLAG_AUTOTUNE) && s->use_fused) A(autotune(s));
  if (st == MMP_OK) A(build_graph(s));
  if (st != MMP_OK) return cleanup(st);
  *out = s;
  return MMP_OK;
}

// The tunable shape of the column-stage step: the gather's wavefronts per workgroup, Raman lines
// per LDS chunk and three-workgroups-per-CU instance; the pair pass's neighbouring columns per
// workgroup and line groups per column block (the column DFTs' work layout); and whether the
// Raman lines travel in column order into their row pass. Only the gather variants change the
// order of a sum (over the pair slots); the other choices leave every result bit-identical.
struct ColPlan {
  int waves, lines, hi, columns, groups, conv_cols, mf;
  bool operator==(const ColPlan& o) const {
    return waves == o.waves && lines == o.lines && hi == o.hi && columns == o.columns && groups == o.groups &&
           conv_cols == o.conv_cols && mf == o.mf;
  }
};

ColPlan current_plan(const mmp_solver* s) {
  return ColPlan{s->fused.waves, s->chunk_lines, s->fused.hi_occupancy, s->pair_columns, s->pair_groups,
                 s->conv_cols, s->fused.matrix_cores};
}

bool plan_valid(const mmp_solver* s, const ColPlan& c) {
  const int M = (int)s->P.M;
  if (c.waves != 4 && !(c.waves == 8 && colstage_wide_ok(M, s->mixed))) return false;
  if (c.hi && !(c.waves == 4 && colstage_hi_occupancy_ok(M, s->mixed, s->fused.parity != 0))) return false;
  const bool mf_ok = colstage_matrix_cores_ok(M, s->mixed, s->fused.parity != 0);
  if (c.mf && (c.waves != 4 || c.hi || (s->opt.flags & MMP_FLAG_NO_MATRIX_CORES) || !mf_ok)) return false;
  if (!c.mf && mf_ok && (s->opt.flags & MMP_FLAG_MATRIX_CORES)) return false;
  if (c.columns != 1 && c.columns != 2 && c.columns != 4) return false;
  if (s->g.N2 % c.columns || colstage_pairs_lds_bytes(M, c.columns, s->mixed) > kColStagePairsLds) return false;
  if (c.groups < 1 || c.groups > s->n_lines || (c.groups & (c.groups - 1))) return false;
  return c.conv_cols == 0 || c.conv_cols == 1;
}

// Installs plan c (chunk tables uploaded); the graphs are captured again.
mmp_status apply_colstage_plan(mmp_solver* s, const ColPlan& c) {
  if (!s->use_fused) return fail(MMP_ERR_UNSUPPORTED, "no column-stage kernels in this solver");
  if (!plan_valid(s, c)) return fail(MMP_ERR_INVALID_ARGUMENT, "no such column-stage variant for this problem");
  Problem& P = s->P;
  if (!plan_colstage_chunks(P, c.lines)) return fail(MMP_ERR_INVALID_ARGUMENT, "no chunk plan for these limits");
  MMP_TRY(upload(s, const_cast<uint32_t*>(s->fused.chunk_line), P.fused_line.data(), P.fused_line.size() * 4));
  MMP_TRY(upload(s, const_cast<uint32_t*>(s->fused.chunk_slot), P.fused_slot.data(), P.fused_slot.size() * 4));
  HIP_TRY(hipStreamSynchronize(s->ctx->stream), "hipStreamSynchronize");  // host staging copies
  s->fused.n_chunks = P.fused_chunks;
  s->fused.max_chunk_lines = P.fused_max_lines;
  s->fused.max_chunk_slots = P.fused_max_slots;
  s->fused.waves = c.waves;
  s->fused.hi_occupancy = c.hi;
  s->fused.matrix_cores = c.mf;
  s->chunk_lines = c.lines;
  s->pair_columns = c.columns;
  s->pair_groups = c.groups;
  s->conv_cols = c.conv_cols;
  drop_graph(s);
  return MMP_OK;
}

mmp_status advance_impl(mmp_solver* s, uint32_t n);

// Autotuner: the column-stage plan is chosen in three passes, each keeping the best of the passes
// before: the gather variants (wavefronts, lines per chunk that give distinct chunk plans, the
// three-workgroups-per-CU instance, the coupling sums on the matrix cores), then the pair
// pass's columns per workgroup and line groups, then the line order into the Raman row pass.
// Every candidate runs one warm-up graph and two
// timed graphs of steps on this solver's own fibers (HIP events); the fastest is kept and
// remembered per structure (modes, samples, stages, slots, lines, parity, precision, fibers) for
// later solvers of the process. The fibers' spectra are saved first and restored after; z, steps,
// Lawson tables and any fiber the device parked during the runs are reset, so the solver ends as
// created.
struct TuneKey {
  uint32_t v[8];
  bool operator<(const TuneKey& o) const { return std::lexicographical_compare(v, v + 8, o.v, o.v + 8); }
};
std::mutex g_tune_mutex;
std::map<TuneKey, ColPlan> g_tune_cache;

mmp_status time_plan(mmp_solver* s, const ColPlan& c, hipEvent_t e0, hipEvent_t e1, float& ms) {
  MMP_TRY(apply_colstage_plan(s, c));
  MMP_TRY(advance_impl(s, kGraphSteps));  // warm-up, graph build
  HIP_TRY(hipEventRecord(e0, s->ctx->stream), "hipEventRecord");
  MMP_TRY(advance_impl(s, 2 * kGraphSteps));
  HIP_TRY(hipEventRecord(e1, s->ctx->stream), "hipEventRecord");
  HIP_TRY(hipEventSynchronize(e1), "hipEventSynchronize");
  HIP_TRY(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
  return MMP_OK;
}

mmp_status tune_passes(mmp_solver* s, hipEvent_t e0, hipEvent_t e1, ColPlan& best) {
  const Problem& P = s->P;
  const int M = (int)P.M;
  best = current_plan(s);
  float best_ms = 0.f;
  MMP_TRY(time_plan(s, best, e0, e1, best_ms));
  auto consider = [&](const ColPlan& c) -> mmp_status {
    if (!plan_valid(s, c) || c == best) return MMP_OK;
    float ms = 0.f;
    MMP_TRY(time_plan(s, c, e0, e1, ms));
    if (ms < best_ms) {
      best_ms = ms;
      best = c;
    }
    return MMP_OK;
  };
  // (1) gather variants (lines per chunk only where they give distinct chunk plans)
  std::vector<std::vector<uint32_t>> seen;
  std::vector<int> lines_opts;
  for (int lines : {colstage_lines_per_chunk(M), 24, 16}) {
    if (lines > colstage_lines_per_chunk(M)) continue;
    Problem probe = P;
    if (!plan_colstage_chunks(probe, lines)) continue;
    if (std::find(seen.begin(), seen.end(), probe.fused_line) != seen.end()) continue;
    seen.push_back(probe.fused_line);
    lines_opts.push_back(lines);
  }
  const ColPlan base1 = best;
  for (int waves : {4, 8})
    for (int lines : lines_opts)
      for (int hi : {0, 1})
        for (int mf : {0, 1}) {
          ColPlan c = base1;
          c.waves = waves;
          c.lines = lines;
          c.hi = hi;
          c.mf = mf;
          MMP_TRY(consider(c));
        }
  // (2) the pair pass's columns per workgroup and line groups
  const ColPlan base2 = best;
  for (int columns : {4, 2, 1})
    for (int groups = 1; groups <= 16 && (groups == 1 || groups * 4 <= s->n_lines); groups *= 2) {
      ColPlan c = base2;
      c.columns = columns;
      c.groups = groups;
      MMP_TRY(consider(c));
    }
  // (3) the Raman lines' order into their row pass
  ColPlan c = best;
  c.conv_cols = 1 - c.conv_cols;
  MMP_TRY(consider(c));
  return MMP_OK;
}

mmp_status autotune(mmp_solver* s) {
  const Problem& P = s->P;
  const TuneKey key{{P.M, P.N, (uint32_t)P.rk.s, P.P, (uint32_t)s->n_lines, (uint32_t)P.parity, s->mixed ? 1u : 0u,
                     s->F}};
  {
    std::lock_guard<std::mutex> lock(g_tune_mutex);
    auto it = g_tune_cache.find(key);
    if (it != g_tune_cache.end()) {
      MMP_TRY(apply_colstage_plan(s, it->second));
      s->tuned = true;
      return MMP_OK;
    }
  }
  const size_t state_bytes = (size_t)s->F * s->batch.sf * sizeof(cd);
  const std::vector<uint8_t> live_before = s->live;
  void* saved = nullptr;
  HIP_TRY(hipMalloc(&saved, state_bytes), "hipMalloc");
  hipEvent_t e0 = nullptr, e1 = nullptr;
  mmp_status st = MMP_OK;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) st = fail(MMP_ERR_DEVICE, "hipEventCreate");
  if (st == MMP_OK && hipMemcpyAsync(saved, s->psi_hat, state_bytes, hipMemcpyDeviceToDevice, s->ctx->stream) != hipSuccess)
    st = fail(MMP_ERR_DEVICE, "hipMemcpyAsync D2D");
  ColPlan best = current_plan(s);
  if (st == MMP_OK) st = tune_passes(s, e0, e1, best);
  if (st == MMP_OK) st = apply_colstage_plan(s, best);
  if (st == MMP_OK && hipMemcpyAsync(s->psi_hat, saved, state_bytes, hipMemcpyDeviceToDevice, s->ctx->stream) != hipSuccess)
    st = fail(MMP_ERR_DEVICE, "hipMemcpyAsync D2D");
  // fibers the device parked during the runs get their live flag back (their state is restored)
  if (st == MMP_OK) st = refresh_halted(s);
  for (uint32_t f = 0; f < s->F && st == MMP_OK; f++) {
    if (!s->halted[f]) continue;
    const uint32_t v = live_before[f];
    st = upload(s, reinterpret_cast<char*>(s->d_fp + f) + offsetof(FiberParams, live), &v, sizeof v);
    static const StepState zero{0., 0u, 0u, 0.};  // clears StepState::halted before reset_z reads it
    if (st == MMP_OK) st = upload(s, s->st + f, &zero, sizeof zero);
    s->halted[f] = 0;
    s->live[f] = live_before[f];
  }
  if (st == MMP_OK) st = reset_z(s);  // synchronises the stream before `saved` is freed
  if (st == MMP_OK) st = drain_profile(s);
  if (st == MMP_OK) {
    std::fill(s->kernel_ms.begin(), s->kernel_ms.end(), 0.);  // the tuning steps are not the caller's
    std::fill(s->kernel_count.begin(), s->kernel_count.end(), 0);
    std::fill(s->kernel_bytes.begin(), s->kernel_bytes.end(), 0.);
  }
  hipStreamSynchronize(s->ctx->stream);
  if (e0) hipEventDestroy(e0);
  if (e1) hipEventDestroy(e1);
  hipFree(saved);
  if (st != MMP_OK) return st;
  s->tuned = true;
  std::lock_guard<std::mutex> lock(g_tune_mutex);
  g_tune_cache[key] = best;
  return MMP_OK;
}

// Host bookkeeping of n steps: z += h (src/MultimodePropagation_CPU.cpp:32) and the step count of
// every live fiber.
void account_steps(mmp_solver* s, uint32_t n) {
  for (uint32_t f = 0; f < s->F; f++) {
    if (!s->live[f]) continue;
    for (uint32_t k = 0; k < n; k++) s->z_host[f] += s->fib[f].delta_z;
    s->steps[f] += n;
  }
}

mmp_status advance_impl(mmp_solver* s, uint32_t n) {
  MMP_TRY(set_device(s));
  if (n > 0 && s->chain) MMP_TRY(enqueue_head(s, 0));  // stage 0 of the first step
  uint32_t left = n;
  if (n > 0) MMP_TRY(build_graph(s));
  if (s->graphs) {
    for (; left >= (uint32_t)kGraphSteps; left -= kGraphSteps)
      HIP_TRY(hipGraphLaunch(s->graph_exec, s->ctx->stream), "hipGraphLaunch");
    for (; left > 0; left--) HIP_TRY(hipGraphLaunch(s->graph1_exec, s->ctx->stream), "hipGraphLaunch");
  }
  for (uint32_t k = 0; k < left; k++) MMP_TRY(enqueue_step(s));
  account_steps(s, n);
  return MMP_OK;
}

// One graph of `steps` steps of every member of the group, members interleaved step by step.
mmp_status capture_group(mmp_group* g, int steps, hipGraph_t* graph_out, hipGraphExec_t* exec_out) {
  hipStream_t st = g->ctx->stream;
  HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
  mmp_status status = MMP_OK;
  const size_t nb = g->branches.size();
  // fork: every branch stream joins the capture behind this event
  size_t forked = 0;
  if (nb) {
    if (hipEventRecord(g->events[0], st) != hipSuccess) status = fail(MMP_ERR_DEVICE, "hipEventRecord");
    for (size_t b = 0; b < nb && status == MMP_OK; b++) {
      if (hipStreamWaitEvent(g->branches[b], g->events[0], 0) == hipSuccess) forked++;
      else status = fail(MMP_ERR_DEVICE, "hipStreamWaitEvent");
    }
  }
  for (int k = 0; k < steps && status == MMP_OK; k++)
    for (mmp_solver* s : g->members)
      if (status == MMP_OK) status = enqueue_step(s);
  // join every branch that joined the capture back into the group's stream, also after a failure
  // (a capture ends only with its branches joined, and the branch streams leave capture mode then)
  for (size_t b = 0; b < forked; b++) {
    if (hipEventRecord(g->events[1 + b], g->branches[b]) != hipSuccess ||
        hipStreamWaitEvent(st, g->events[1 + b], 0) != hipSuccess)
      if (status == MMP_OK) status = fail(MMP_ERR_DEVICE, "hipEventRecord / hipStreamWaitEvent");
  }
  hipGraph_t graph = nullptr;
  hipError_t e = hipStreamEndCapture(st, &graph);
  if (status != MMP_OK) {
    if (graph) hipGraphDestroy(graph);
    return status;
  }
  if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
  hipGraphExec_t exec = nullptr;
  e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  if (e == hipSuccess) e = hipGraphUpload(exec, st);
  if (e != hipSuccess) {
    if (exec) hipGraphExecDestroy(exec);
    hipGraphDestroy(graph);
    return hip_fail(e, "hipGraphInstantiate");
  }
  *graph_out = graph;
  *exec_out = exec;
  return MMP_OK;
}

void group_release(mmp_group* g) {
  if (!g) return;
  if (g->ctx) {
    hipSetDevice(g->ctx->device);
    hipStreamSynchronize(g->ctx->stream);
  }
  if (g->exec) hipGraphExecDestroy(g->exec);
  if (g->graph) hipGraphDestroy(g->graph);
  if (g->exec1) hipGraphExecDestroy(g->exec1);
  if (g->graph1) hipGraphDestroy(g->graph1);
  for (hipEvent_t e : g->events) hipEventDestroy(e);
  delete g;
}

mmp_status check_fiber(mmp_solver* s, uint32_t f) {
  if (!s) return fail(MMP_ERR_INVALID_ARGUMENT, "solver is NULL");
  if (f >= s->F) return fail(MMP_ERR_INVALID_ARGUMENT, "fiber index out of range");
  return MMP_OK;
}

mmp_status read_fields(mmp_solver* s, double* out) {
  HIP_TRY(hipMemcpyAsync(out, s->U, sizeof(cd) * s->P.M * s->P.N, hipMemcpyDeviceToHost, s->ctx->stream),
          "hipMemcpyAsync D2H");
  HIP_TRY(hipStreamSynchronize(s->ctx->stream), "hipStreamSynchronize");
  return MMP_OK;
}

}  // namespace

extern "C" {

const char* mmp_status_string(mmp_status status) {
  switch (status) {
    case MMP_OK: return "MMP_OK";
    case MMP_ERR_INVALID_ARGUMENT: return "MMP_ERR_INVALID_ARGUMENT";
    case MMP_ERR_UNSUPPORTED: return "MMP_ERR_UNSUPPORTED";
    case MMP_ERR_OUT_OF_MEMORY: return "MMP_ERR_OUT_OF_MEMORY";
    case MMP_ERR_DEVICE: return "MMP_ERR_DEVICE";
    case MMP_ERR_NO_DEVICE: return "MMP_ERR_NO_DEVICE";
    case MMP_ERR_NONFINITE: return "MMP_ERR_NONFINITE";
    case MMP_ERR_STATE: return "MMP_ERR_STATE";
  }
  return "MMP_ERR_UNKNOWN";
}

const char* mmp_last_error(void) { return g_last_error.c_str(); }

int mmp_abi_version(void) { return MMPROP_ABI_VERSION; }

#ifndef MMP_SOURCE_HASH
#define MMP_SOURCE_HASH "unknown"
#endif
const char* mmp_source_hash(void) { return MMP_SOURCE_HASH; }

mmp_status mmp_ctx_create(int device, mmp_ctx** out) {
  if (!out) return fail(MMP_ERR_INVALID_ARGUMENT, "out is NULL");
  *out = nullptr;
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count == 0) return fail(MMP_ERR_NO_DEVICE, "no HIP device available");
  if (device < 0 || device >= count) return fail(MMP_ERR_INVALID_ARGUMENT, "device index out of range");
  hipDeviceProp_t prop;
  HIP_TRY(hipGetDeviceProperties(&prop, device), "hipGetDeviceProperties");
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return fail(MMP_ERR_NO_DEVICE, std::string("kernels are built for gfx950, device is ") + prop.gcnArchName);
  HIP_TRY(hipSetDevice(device), "hipSetDevice");
  // the no-allocation entry point's scratch: its address on this device, looked up once
  HIP_TRY(static_scratch_resolve(device), "hipGetSymbolAddress(scratch)");
  mmp_ctx* c = new (std::nothrow) mmp_ctx();
  if (!c) return fail(MMP_ERR_OUT_OF_MEMORY, "host allocation failed");
  c->device = device;
  e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete c;
    return hip_fail(e, "hipStreamCreate");
  }
  *out = c;
  return MMP_OK;
}

void mmp_ctx_destroy(mmp_ctx* ctx) {
  if (!ctx) return;
  hipSetDevice(ctx->device);
  if (ctx->stream) {
    hipStreamSynchronize(ctx->stream);
    hipStreamDestroy(ctx->stream);
  }
  delete ctx;
}

mmp_status mmp_ctx_synchronize(mmp_ctx* ctx) {
  if (!ctx) return fail(MMP_ERR_INVALID_ARGUMENT, "ctx is NULL");
  HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
  HIP_TRY(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
  return MMP_OK;
}

mmp_status mmp_ctx_query(mmp_ctx* ctx) {
  if (!ctx) return fail(MMP_ERR_INVALID_ARGUMENT, "ctx is NULL");
  HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
  const hipError_t e = hipStreamQuery(ctx->stream);
  if (e == hipSuccess) return MMP_OK;
  if (e == hipErrorNotReady) return MMP_ERR_STATE;
  return hip_fail(e, "hipStreamQuery");
}

mmp_status mmp_solver_create(mmp_ctx* ctx, const mmp_params* params, const mmp_options* options,
                             mmp_solver** out) {
  try {
    return create_impl(ctx, 1, params, options, out);
  } catch (const std::bad_alloc&) {
    return fail(MMP_ERR_OUT_OF_MEMORY, "host allocation failed");
  } catch (...) {
    return fail(MMP_ERR_DEVICE, "unexpected exception");
  }
}

mmp_status mmp_batch_create(mmp_ctx* ctx, uint32_t n_fibers, const mmp_params* params, const mmp_options* options,
                            mmp_solver** out) {
  try {
    return create_impl(ctx, n_fibers, params, options, out);
  } catch (const std::bad_alloc&) {
    return fail(MMP_ERR_OUT_OF_MEMORY, "host allocation failed");
  } catch (...) {
    return fail(MMP_ERR_DEVICE, "unexpected exception");
  }
}

void mmp_solver_destroy(mmp_solver* s) { release(s); }

mmp_status mmp_group_create(mmp_ctx* ctx, uint32_t n, mmp_solver* const* solvers, mmp_group** out) {
  if (!ctx || !solvers || !out || n < 1) return fail(MMP_ERR_INVALID_ARGUMENT, "ctx, solvers (n >= 1) and out are required");
  *out = nullptr;
  for (uint32_t i = 0; i < n; i++) {
    const mmp_solver* s = solvers[i];
    if (!s) return fail(MMP_ERR_INVALID_ARGUMENT, "a solver of the group is NULL");
    if (s->ctx->device != ctx->device)
      return fail(MMP_ERR_INVALID_ARGUMENT, "every solver of a group must be on the group context's device");
    if (!s->graphs) return fail(MMP_ERR_UNSUPPORTED, "a solver of the group launches kernel by kernel (profiling, "
                                                      "MMP_FLAG_NO_GRAPH or no graph support)");
    for (uint32_t j = 0; j < i; j++)
      if (solvers[j] == s) return fail(MMP_ERR_INVALID_ARGUMENT, "a solver appears twice in the group");
  }
  try {
    mmp_group* g = new mmp_group();
    g->ctx = ctx;
    g->members.assign(solvers, solvers + n);
    for (const mmp_solver* s : g->members)
      if (s->ctx->stream != ctx->stream &&
          std::find(g->branches.begin(), g->branches.end(), s->ctx->stream) == g->branches.end())
        g->branches.push_back(s->ctx->stream);
    hipError_t e = hipSetDevice(ctx->device);
    mmp_status st = e == hipSuccess ? MMP_OK : hip_fail(e, "hipSetDevice");
    // events: the fork, a join per branch (capture), and one per branch for each advance's start
    for (size_t k = 0; k < 1 + 2 * g->branches.size() && st == MMP_OK; k++) {
      hipEvent_t ev = nullptr;
      e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
      if (e != hipSuccess) st = hip_fail(e, "hipEventCreate");
      else g->events.push_back(ev);
    }
    if (st == MMP_OK) st = capture_group(g, kGraphSteps, &g->graph, &g->exec);
    if (st == MMP_OK) st = capture_group(g, 1, &g->graph1, &g->exec1);
    if (st != MMP_OK) {
      group_release(g);
      return st;
    }
    *out = g;
    return MMP_OK;
  } catch (const std::bad_alloc&) {
    return fail(MMP_ERR_OUT_OF_MEMORY, "host allocation failed");
  }
}

mmp_status mmp_group_advance(mmp_group* g, uint32_t n) {
  if (!g) return fail(MMP_ERR_INVALID_ARGUMENT, "group is NULL");
  HIP_TRY(hipSetDevice(g->ctx->device), "hipSetDevice");
  if (n == 0) return MMP_OK;
  for (mmp_solver* s : g->members)
    if (s->chain) MMP_TRY(enqueue_head(s, 0));  // stage 0 of each member's first step, on its stream
  hipStream_t st = g->ctx->stream;
  const size_t nb = g->branches.size();
  // the graph (on the group's stream) after the members' own work; the members' streams after it
  for (size_t b = 0; b < nb; b++) {
    HIP_TRY(hipEventRecord(g->events[1 + nb + b], g->branches[b]), "hipEventRecord");
    HIP_TRY(hipStreamWaitEvent(st, g->events[1 + nb + b], 0), "hipStreamWaitEvent");
  }
  uint32_t left = n;
  for (; left >= (uint32_t)kGraphSteps; left -= kGraphSteps) HIP_TRY(hipGraphLaunch(g->exec, st), "hipGraphLaunch");
  for (; left > 0; left--) HIP_TRY(hipGraphLaunch(g->exec1, st), "hipGraphLaunch");
  if (nb) {
    HIP_TRY(hipEventRecord(g->events[0], st), "hipEventRecord");
    for (size_t b = 0; b < nb; b++) HIP_TRY(hipStreamWaitEvent(g->branches[b], g->events[0], 0), "hipStreamWaitEvent");
  }
  for (mmp_solver* s : g->members) account_steps(s, n);
  return MMP_OK;
}

void mmp_group_destroy(mmp_group* g) { group_release(g); }

// One solve whose device arrays all live in the device's static scratch (no allocation). Calls are
// serialised by this lock: the scratch is one array per device, and a call ends only when its
// result has been copied to the host.
static std::mutex g_static_solve_mutex;
static_assert(MMP_STATIC_SCRATCH_BYTES == kStaticScratchBytes, "the header states the scratch size");

mmp_status mmp_solve_static(mmp_ctx* ctx, const mmp_params* params, const mmp_options* options, uint32_t nz,
                            double* fields_out) {
  if (!ctx || !params || !fields_out) return fail(MMP_ERR_INVALID_ARGUMENT, "ctx, params and fields_out are required");
  mmp_options opt{};
  if (options) opt = *options;
  if (opt.flags & (MMP_FLAG_PROFILE | MMP_FLAG_AUTOTUNE))
    return fail(MMP_ERR_INVALID_ARGUMENT, "mmp_solve_static takes neither MMP_FLAG_PROFILE nor MMP_FLAG_AUTOTUNE");
  try {
    // the footprint first, on the host, so that a problem too large is refused before any device work
    mmp_info plan{};
    MMP_TRY(mmp_plan(params, &opt, &plan));
    const size_t slack = 64 * kArenaAlign;  // alignment of the solver's arrays within the scratch
    if (plan.device_bytes + slack > kStaticScratchBytes)
      return fail(MMP_ERR_UNSUPPORTED, "mmp_solve_static: the problem needs " + std::to_string(plan.device_bytes) +
                                           " bytes of device memory; the static scratch holds " +
                                           std::to_string(kStaticScratchBytes) + " (use mmp_solver_create)");
    std::lock_guard<std::mutex> lock(g_static_solve_mutex);
    HIP_TRY(hipSetDevice(ctx->device), "hipSetDevice");
    unsigned char* base = static_cast<unsigned char*>(static_scratch(ctx->device));
    if (!base) return fail(MMP_ERR_STATE, "the static scratch of this device is not resolved (mmp_ctx_create)");
    mmp_solver* s = nullptr;
    MMP_TRY(create_impl(ctx, 1, params, &opt, &s, base, kStaticScratchBytes));
    mmp_status st = mmp_compute_lawson_rk(s, nz);
    if (st == MMP_OK) st = mmp_fiber_result(s, 0, fields_out);
    release(s);  // waits for the stream; frees no device memory
    return st;
  } catch (const std::bad_alloc&) {
    return fail(MMP_ERR_OUT_OF_MEMORY, "host allocation failed");
  } catch (...) {
    return fail(MMP_ERR_DEVICE, "unexpected exception");
  }
}

mmp_status mmp_compute_lawson_rk(mmp_solver* s, uint32_t nz) {
  if (!s) return fail(MMP_ERR_INVALID_ARGUMENT, "solver is NULL");
  MMP_TRY(set_device(s));
  MMP_TRY(reset_z(s));  // initLawsonRK: z = 0, operators rebuilt (src/MultimodePropagation.cpp:158-163)
  return advance_impl(s, nz);
}

mmp_status mmp_compute_lawson_rk_default(mmp_solver* s) {
  if (!s) return fail(MMP_ERR_INVALID_ARGUMENT, "solver is NULL");
  return mmp_compute_lawson_rk(s, s->P.default_steps);
}

mmp_status mmp_advance(mmp_solver* s, uint32_t n) {
  if (!s) return fail(MMP_ERR_INVALID_ARGUMENT, "solver is NULL");
  return advance_impl(s, n);
}

mmp_status mmp_get_result(mmp_solver* s, double* out) { return mmp_fiber_result(s, 0, out); }

mmp_status mmp_fiber_result(mmp_solver* s, uint32_t f, double* out) {
  MMP_TRY(check_fiber(s, f));
  if (!out) return fail(MMP_ERR_INVALID_ARGUMENT, "fields_out is required");
  MMP_TRY(set_device(s));
  MMP_TRY(inverse_state_to_U(s, f, true));
  MMP_TRY(read_fields(s, out));
  const size_t n = 2 * (size_t)s->P.M * s->P.N;
  for (size_t i = 0; i < n; i++)
    if (!std::isfinite(out[i])) return fail(MMP_ERR_NONFINITE, "the propagated field is not finite");
  return MMP_OK;
}

mmp_status mmp_propagate_snapshots(mmp_solver* s, uint32_t n_steps, uint32_t n_save, uint32_t mode,
                                   double* fields_out, double* z_out) {
  if (!s) return fail(MMP_ERR_INVALID_ARGUMENT, "solver is NULL");
  if (n_save == 0 || n_save > n_steps) return fail(MMP_ERR_INVALID_ARGUMENT, "n_save must be in [1, n_steps]");
  if (mode != MMP_SNAPSHOTS_REFERENCE_RESTART && mode != MMP_SNAPSHOTS_RESUMABLE)
    return fail(MMP_ERR_INVALID_ARGUMENT, "unknown snapshot mode");
  const uint32_t nz = n_steps / n_save;
  const double dZ = s->fiber_length[0] / n_steps;
  const size_t stride = 2 * (size_t)s->P.M * s->P.N;
  double Z = 0.;
  for (uint32_t i = 0; i < n_save; i++) {
    const uint32_t n = (i + 1 < n_save) ? nz : nz + n_steps % n_save;
    if (mode == MMP_SNAPSHOTS_REFERENCE_RESTART || i == 0) MMP_TRY(mmp_compute_lawson_rk(s, n));
    else MMP_TRY(mmp_advance(s, n));
    Z += n * dZ;
    if (fields_out) MMP_TRY(mmp_get_result(s, fields_out + i * stride));
    if (z_out) z_out[i] = Z;
  }
  return MMP_OK;
}

mmp_status mmp_get_state(mmp_solver* s, double* psi_out, double* z_out) { return mmp_fiber_state(s, 0, psi_out, z_out); }

mmp_status mmp_fiber_state(mmp_solver* s, uint32_t f, double* psi_out, double* z_out) {
  MMP_TRY(check_fiber(s, f));
  MMP_TRY(set_device(s));
  MMP_TRY(refresh_halted(s));
  if (z_out) *z_out = s->z_host[f];
  if (psi_out) {
    MMP_TRY(inverse_state_to_U(s, f, false));
    MMP_TRY(read_fields(s, psi_out));
  }
  return MMP_OK;
}

mmp_status mmp_set_state(mmp_solver* s, const double* psi) { return mmp_fiber_set_state(s, 0, psi); }

mmp_status mmp_fiber_set_state(mmp_solver* s, uint32_t f, const double* psi) {
  MMP_TRY(check_fiber(s, f));
  if (!psi) return fail(MMP_ERR_INVALID_ARGUMENT, "psi is required");
  const size_t n = 2 * (size_t)s->P.M * s->P.N;
  for (size_t i = 0; i < n; i++)
    if (!std::isfinite(psi[i])) return fail(MMP_ERR_NONFINITE, "psi is not finite");
  MMP_TRY(set_device(s));
  MMP_TRY(refresh_halted(s));
  MMP_TRY(upload(s, s->U, psi, n * 8));
  MMP_TRY(forward_state_from_U(s, f));
  MMP_TRY(reset_fiber(s, f));  // clears StepState::halted
  if (s->halted[f]) {  // a finite field again: the fiber is live again
    const uint32_t v = 1;
    MMP_TRY(upload(s, reinterpret_cast<char*>(s->d_fp + f) + offsetof(FiberParams, live), &v, sizeof v));
    s->halted[f] = 0;
    s->live[f] = 1;
  }
  HIP_TRY(hipStreamSynchronize(s->ctx->stream), "hipStreamSynchronize");
  return MMP_OK;
}

mmp_status mmp_fiber_load(mmp_solver* s, uint32_t f, const mmp_params* params) {
  MMP_TRY(check_fiber(s, f));
  if (!params) return fail(MMP_ERR_INVALID_ARGUMENT, "params is required");
  try {
    Problem Pf;
    std::string err;
    const mmp_status st = prepare_problem(params, s->opt, Pf, err);
    if (st != MMP_OK) return fail(st, err);
    if (!same_structure(s->P, Pf))
      return fail(MMP_ERR_INVALID_ARGUMENT, "the fiber does not share the batch's n_modes, nt, method_order, raman "
                                            "and pair structure");
    MMP_TRY(set_device(s));
    HIP_TRY(hipStreamSynchronize(s->ctx->stream), "hipStreamSynchronize");  // queued steps read the old fiber
    MMP_TRY(load_fiber(s, f, Pf));
    s->fib[f] = std::move(Pf);
    s->time_window[f] = params->time_window;
    s->fiber_length[f] = params->fiber_length;
    return MMP_OK;
  } catch (const std::bad_alloc&) {
    return fail(MMP_ERR_OUT_OF_MEMORY, "host allocation failed");
  }
}

mmp_status mmp_fiber_set_live(mmp_solver* s, uint32_t f, uint32_t live) {
  MMP_TRY(check_fiber(s, f));
  if (live > 1) return fail(MMP_ERR_INVALID_ARGUMENT, "live must be 0 or 1");
  MMP_TRY(set_device(s));
  MMP_TRY(refresh_halted(s));
  if (live && s->halted[f])
    return fail(MMP_ERR_NONFINITE, "the fiber's field is not finite (load a new problem or set a new state)");
  if (s->live[f] == live) return MMP_OK;
  // only the column-stage step is made of kernels that all honour the flag
  if (!s->use_fused) return fail(MMP_ERR_UNSUPPORTED, "parking a fiber needs the column-stage kernels");
  MMP_TRY(set_device(s));
  const uint32_t v = live;
  HIP_TRY(hipStreamSynchronize(s->ctx->stream), "hipStreamSynchronize");  // queued steps see the old flag
  MMP_TRY(upload(s, reinterpret_cast<char*>(s->d_fp + f) + offsetof(FiberParams, live), &v, sizeof v));
  HIP_TRY(hipStreamSynchronize(s->ctx->stream), "hipStreamSynchronize");
  s->live[f] = (uint8_t)live;
  return MMP_OK;
}

mmp_status mmp_energy(mmp_solver* s, double* energy_out) { return mmp_fiber_energy(s, 0, energy_out); }

mmp_status mmp_fiber_energy(mmp_solver* s, uint32_t f, double* energy_out) {
  MMP_TRY(check_fiber(s, f));
  if (!energy_out) return fail(MMP_ERR_INVALID_ARGUMENT, "energy_out is required");
  MMP_TRY(set_device(s));
  MMP_TRY(inverse_state_to_U(s, f, true));
  HIP_TRY(launch_energy(s->g, s->U, s->energy, s->time_window[f] / s->P.N, s->ctx->stream), "energy");
  HIP_TRY(hipMemcpyAsync(energy_out, s->energy, sizeof(double) * s->P.M, hipMemcpyDeviceToHost, s->ctx->stream),
          "hipMemcpyAsync D2H");
  HIP_TRY(hipStreamSynchronize(s->ctx->stream), "hipStreamSynchronize");
  return MMP_OK;
}

mmp_status mmp_photon_number(mmp_solver* s, double wavelength, double* photons_out) {
  return mmp_fiber_photon_number(s, 0, wavelength, photons_out);
}

// N_p = sum_k E_p,k / (hbar omega_k): E_p,k = (dt / NT) |F A_p|_k^2 (W ps) is bin k's share of the
// energy (Parseval, unnormalised forward DFT) and omega_k = omega0 - alpha_k (rad/ps) its absolute
// angular frequency, alpha_k = 2 pi k / T_w on the reference's grid and the carrier
// exp(-i omega0 t) (the operator i d/dt of src/MultimodePropagation.cpp:166 maps bin alpha to
// omega0 - alpha). |F A_p| = |psi_hat_p| / conversion_factor, the Lawson phase exp(L z) having
// modulus one, so the device state is reduced directly.
mmp_status mmp_fiber_photon_number(mmp_solver* s, uint32_t f, double wavelength, double* photons_out) {
  MMP_TRY(check_fiber(s, f));
  if (!photons_out) return fail(MMP_ERR_INVALID_ARGUMENT, "photons_out is required");
  if (!(wavelength > 0.) || !std::isfinite(wavelength))
    return fail(MMP_ERR_INVALID_ARGUMENT, "the carrier wavelength must be a positive number of metres");
  const double c_light = 299792458.0;          // m/s
  const double hbar = 1.054571817e-34;         // J s
  const double two_pi = 6.283185307179586476925286766559;
  const double tw = s->time_window[f];         // ps
  const double omega0 = two_pi * c_light / wavelength * 1e-12;  // rad/ps
  const double dalpha = two_pi / tw;                            // rad/ps per bin
  if (!(omega0 - dalpha * (double)(s->P.N / 2) > 0.))
    return fail(MMP_ERR_INVALID_ARGUMENT,
                "the time grid reaches negative absolute frequencies: the carrier frequency must exceed "
                "pi * nt / time_window");
  const double conv = s->fib[f].conversion_factor;
  const double dt = tw / s->P.N;  // ps
  // W ps -> J (1e-12), rad/ps -> rad/s (1e12)
  const double scale = dt / ((double)s->P.N * conv * conv) * 1e-24 / hbar;
  MMP_TRY(set_device(s));
  HIP_TRY(launch_photon_number(s->g, s->psi_hat + (size_t)f * s->batch.sf, s->energy, omega0, dalpha, scale,
                               s->ctx->stream),
          "photon_number");
  HIP_TRY(hipMemcpyAsync(photons_out, s->energy, sizeof(double) * s->P.M, hipMemcpyDeviceToHost, s->ctx->stream),
          "hipMemcpyAsync D2H");
  HIP_TRY(hipStreamSynchronize(s->ctx->stream), "hipStreamSynchronize");
  return MMP_OK;
}

mmp_status mmp_solver_info(mmp_solver* s, mmp_info* o) { return mmp_fiber_info(s, 0, o); }

mmp_status mmp_fiber_info(mmp_solver* s, uint32_t f, mmp_info* o) {
  MMP_TRY(check_fiber(s, f));
  if (!o) return fail(MMP_ERR_INVALID_ARGUMENT, "info_out is required");
  MMP_TRY(set_device(s));
  MMP_TRY(refresh_halted(s));
  fill_info(s->fib[f], s->opt, o);
  o->steps_taken = s->steps[f];
  o->z = s->z_host[f];
  o->device_bytes = s->device_bytes;
  o->n_fibers = s->F;
  o->live = s->live[f];
  o->gather_waves = s->use_fused ? (uint32_t)s->fused.waves : 0u;
  o->chunk_lines = s->use_fused ? (uint32_t)s->chunk_lines : 0u;
  o->tuned = s->tuned ? 1u : 0u;
  o->pair_columns = s->use_fused ? (uint32_t)s->pair_columns : 0u;
  o->pair_groups = s->use_fused ? (uint32_t)s->pair_groups : 0u;
  o->gather_hi_occupancy =
      s->use_fused && s->fused.waves == 4 && s->fused.hi_occupancy && !s->fused.matrix_cores ? 1u : 0u;
  o->gather_matrix_cores = s->use_fused && s->fused.waves == 4 && s->fused.matrix_cores ? 1u : 0u;
  o->gather_chunks = s->use_fused ? (uint32_t)s->fused.n_chunks : 0u;
  o->halted = s->halted[f];
  return MMP_OK;
}

mmp_status mmp_tableau(uint32_t method_order, uint32_t* stages_out, double* a_out, double* b_out, double* c_out) {
  if (!stages_out || !a_out || !b_out || !c_out) return fail(MMP_ERR_INVALID_ARGUMENT, "output pointers are required");
  const Tableau t = make_tableau(method_order);
  *stages_out = (uint32_t)t.s;
  for (int i = 0; i < 11; i++) {
    for (int j = 0; j < 11; j++) a_out[i * 11 + j] = t.a[i][j];
    b_out[i] = t.b[i];
    c_out[i] = t.c[i];
  }
  return MMP_OK;
}

mmp_status mmp_plan(const mmp_params* params, const mmp_options* options, mmp_info* o) {
  if (!o) return fail(MMP_ERR_INVALID_ARGUMENT, "info_out is NULL");
  try {
    mmp_options opt{};
    if (options) opt = *options;
    Problem P;
    std::string err;
    const mmp_status st = prepare_problem(params, opt, P, err);
    if (st != MMP_OK) return fail(st, err);
    fill_info(P, opt, o);
    return MMP_OK;
  } catch (const std::bad_alloc&) {
    return fail(MMP_ERR_OUT_OF_MEMORY, "host allocation failed");
  }
}

mmp_status mmp_plan_arrays(const mmp_params* params, const mmp_options* options, double* psi0_out, double* imL_out,
                           uint32_t* pair_a_out, uint32_t* pair_b_out, double* e1_out, double* e_out) {
  try {
    mmp_options opt{};
    if (options) opt = *options;
    Problem P;
    std::string err;
    const mmp_status st = prepare_problem(params, opt, P, err);
    if (st != MMP_OK) return fail(st, err);
    if (psi0_out) std::memcpy(psi0_out, P.psi0.data(), P.psi0.size() * sizeof(double));
    if (imL_out)
      for (uint32_t p = 0; p < P.M; p++)
        for (uint32_t j = 0; j < P.N; j++) {
          const uint32_t k1 = j % P.N1, k2 = j / P.N1;
          imL_out[(size_t)p * P.N + j] = P.imL_fourstep[(size_t)p * P.N + (size_t)k1 * P.N2 + k2];
        }
    if (pair_a_out) std::memcpy(pair_a_out, P.pair_a_ref.data(), P.P * sizeof(uint32_t));
    if (pair_b_out) std::memcpy(pair_b_out, P.pair_b_ref.data(), P.P * sizeof(uint32_t));
    const size_t MN = (size_t)P.M * P.N;
    for (uint32_t p = 0; p < P.M; p++)
      for (uint32_t j = 0; j < P.N; j++) {
        const uint32_t k1 = j % P.N1, k2 = j / P.N1;
        const size_t nat = (size_t)p * P.N + j, four = (size_t)p * P.N + (size_t)k1 * P.N2 + k2;
        if (e1_out) {
          e1_out[2 * nat] = P.e1[2 * four];
          e1_out[2 * nat + 1] = P.e1[2 * four + 1];
        }
        if (e_out)
          for (int i = 0; i < P.rk.s; i++) {
            const size_t src = (size_t)P.stage_table[i] * MN + four, dst = (size_t)i * MN + nat;
            e_out[2 * dst] = P.etab0[2 * src];
            e_out[2 * dst + 1] = P.etab0[2 * src + 1];
          }
      }
    return MMP_OK;
  } catch (const std::bad_alloc&) {
    return fail(MMP_ERR_OUT_OF_MEMORY, "host allocation failed");
  }
}

mmp_status mmp_kernel_times(mmp_solver* s, mmp_kernel_time* out, uint32_t capacity, uint32_t* count_out) {
  if (!s) return fail(MMP_ERR_INVALID_ARGUMENT, "solver is NULL");
  MMP_TRY(set_device(s));
  MMP_TRY(drain_profile(s));
  const uint32_t n = (uint32_t)s->kernel_names.size();
  if (count_out) *count_out = n;
  for (uint32_t i = 0; i < n && i < capacity && out; i++) {
    std::memset(&out[i], 0, sizeof out[i]);
    std::snprintf(out[i].name, sizeof out[i].name, "%s", s->kernel_names[i].c_str());
    out[i].total_ms = s->kernel_ms[i];
    out[i].launches = s->kernel_count[i];
    out[i].bytes_per_launch = s->kernel_count[i] ? s->kernel_bytes[i] / (double)s->kernel_count[i] : 0.;
  }
  if (out) {
    for (uint32_t i = 0; i < n; i++) {
      s->kernel_ms[i] = 0.;
      s->kernel_count[i] = 0;
      s->kernel_bytes[i] = 0.;
    }
  }
  return MMP_OK;
}

}  // extern "C"
