// host_sanitize.cpp -- the host side of libmmprop_hip.so (problem preparation, planning, the
// request server's admission and bookkeeping, the error paths) driven without a device, built
// with AddressSanitizer and UndefinedBehaviorSanitizer on the library's host code and on this
// program. Exits non-zero on a wrong status; the sanitizers abort on a memory or UB error.
//
// Faults forced from the host: non-finite and malformed inputs, an all-zero coupling tensor, a
// request larger than the server's memory budget (refused at admission), solver creation and
// server scheduling with no device present, and concurrent planning from several threads.
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <thread>
#include <vector>

#include "mmprop.h"

namespace {

int g_bad = 0;

void expect(bool ok, const char* what) {
  if (!ok) {
    std::fprintf(stderr, "FAILED: %s (last error: %s)\n", what, mmp_last_error());
    g_bad++;
  }
}

// A GRIN-like problem: every (k, l, m) with k + l - m - p even, values from a fixed sequence.
struct Problem {
  uint32_t M, nt;
  std::vector<double> beta, qv, fields, hr;
  std::vector<uint32_t> off, klm;
  mmp_params p{};

  Problem(uint32_t modes, uint32_t samples, uint32_t raman = 1) : M(modes), nt(samples) {
    for (uint32_t q = 0; q < M; q++) {
      const double b[4] = {-1.5e3 * q, 5e-4 * q, 0.02, 3e-5};
      beta.insert(beta.end(), b, b + 4);
    }
    off.push_back(0);
    uint32_t seq = 1;
    for (uint32_t q = 0; q < M; q++) {
      for (uint32_t k = 0; k < M; k++)
        for (uint32_t l = 0; l < M; l++)
          for (uint32_t m = 0; m < M; m++)
            if ((k + l + 2 * M - m - q) % 2 == 0) {
              klm.insert(klm.end(), {k, l, m});
              seq = seq * 1103515245u + 12345u;
              qv.push_back(6.4341e12 * (-0.5 + 1.5 * ((seq >> 8) & 0xffff) / 65535.0));
            }
      off.push_back((uint32_t)qv.size());
    }
    const double tw = 20.0, t0 = 0.1;
    fields.resize(2 * (size_t)M * nt);
    for (uint32_t q = 0; q < M; q++)
      for (uint32_t j = 0; j < nt; j++) {
        const double t = ((double)j - nt / 2.0) * tw / nt;
        fields[2 * ((size_t)q * nt + j)] = std::sqrt(10.0 * (q + 1)) * std::exp(-t * t / (2 * t0 * t0));
      }
    hr.resize(nt);
    for (uint32_t j = 0; j < nt; j++) {
      const double t = j * tw / nt, t1 = 0.0122, t2 = 0.032;
      hr[j] = (t1 * t1 + t2 * t2) / (t1 * t2 * t2) * std::exp(-t / t2) * std::sin(t / t1);
    }
    p.n_modes = M;
    p.fiber_length = 1.0;
    p.n_beta = 4;
    p.dispersion_coefficients = beta.data();
    p.coupling_offsets = off.data();
    p.coupling_klm = klm.data();
    p.coupling_values = qv.data();
    p.nt = nt;
    p.time_window = tw;
    p.pulse_width = t0;
    p.initial_fields = fields.data();
    p.n_steps = 1000;
    p.method_order = 4;
    p.nonlinearity_const = 1.4e-13;
    p.raman_proportion = 0.18;
    p.raman_response = hr.data();
    p.raman = raman;
  }
};

void planning() {
  const mmp_options fp64{MMP_PRECISION_FP64, MMP_PAIRS_PHYSICAL, 0, 0};
  const mmp_options mixed{MMP_PRECISION_MIXED, MMP_PAIRS_REFERENCE_CPU, 0, 0};
  const uint32_t shapes[][2] = {{1, 64}, {2, 4096}, {3, 1024}, {6, 16384}, {10, 4096}, {15, 4096}, {16, 8192}, {20, 256}};
  for (auto& sh : shapes) {
    for (uint32_t raman : {0u, 1u}) {
      Problem pr(sh[0], sh[1], raman);
      for (const mmp_options* o : {&fp64, &mixed}) {
        mmp_info info{};
        const mmp_status st = mmp_plan(&pr.p, o, &info);
        if (o == &mixed && sh[0] > 6 && st == MMP_ERR_INVALID_ARGUMENT) continue;  // counter decode reads past M
        expect(st == MMP_OK, "mmp_plan on a valid problem");
        if (st != MMP_OK) continue;
        expect(info.n_modes == sh[0] && info.nt == sh[1] && info.device_bytes > 0, "plan sizes");
        const size_t MN = (size_t)sh[0] * sh[1];
        std::vector<double> psi(2 * MN), imL(MN), e1(2 * MN), e(2 * MN * info.stages);
        std::vector<uint32_t> pa(info.n_pairs), pb(info.n_pairs);
        expect(mmp_plan_arrays(&pr.p, o, psi.data(), imL.data(), pa.data(), pb.data(), e1.data(), e.data()) ==
                   MMP_OK,
               "mmp_plan_arrays");
        for (uint32_t j = 0; j < info.n_pairs; j++) expect(pa[j] < sh[0] && pb[j] < sh[0], "decoded pair in range");
        for (size_t j = 0; j < MN; j++)
          expect(std::abs(std::hypot(e1[2 * j], e1[2 * j + 1]) - 1.0) < 1e-12, "|E1| = 1");
      }
    }
  }
  for (uint32_t order : {4u, 6u, 8u}) {
    uint32_t s = 0;
    double a[121], b[11], c[11];
    expect(mmp_tableau(order, &s, a, b, c) == MMP_OK, "mmp_tableau");
    double sb = 0;
    for (uint32_t i = 0; i < s; i++) sb += b[i];
    expect(std::abs(sb - 1.0) < 1e-14, "sum of b = 1");
  }
}

void invalid_inputs() {
  mmp_info info{};
  Problem pr(3, 1024);
  {
    Problem bad = pr;
    bad.p.initial_fields = bad.fields.data();
    bad.fields[17] = std::nan("");
    expect(mmp_plan(&bad.p, nullptr, &info) == MMP_ERR_NONFINITE, "NaN initial field refused");
  }
  {
    Problem bad(3, 1024);
    for (double& v : bad.qv) v = 0.0;
    expect(mmp_plan(&bad.p, nullptr, &info) == MMP_ERR_INVALID_ARGUMENT, "all-zero coupling refused");
  }
  {
    Problem bad(3, 1024);
    bad.off.assign(bad.off.size(), 0);  // empty Q for every mode
    bad.p.coupling_offsets = bad.off.data();
    expect(mmp_plan(&bad.p, nullptr, &info) == MMP_ERR_INVALID_ARGUMENT, "empty coupling refused");
  }
  {
    Problem bad(3, 1000);
    expect(mmp_plan(&bad.p, nullptr, &info) == MMP_ERR_UNSUPPORTED, "nt not a power of two refused");
  }
  {
    Problem bad(3, 1024);
    bad.klm[4] = 7;  // a mode index beyond M
    expect(mmp_plan(&bad.p, nullptr, &info) == MMP_ERR_INVALID_ARGUMENT, "mode index beyond M refused");
  }
  {
    Problem bad(3, 1024);
    bad.p.n_beta = 2;
    expect(mmp_plan(&bad.p, nullptr, &info) == MMP_ERR_INVALID_ARGUMENT, "fewer than 3 dispersion orders refused");
  }
  {
    Problem bad(3, 1024);
    bad.p.raman_response = nullptr;
    expect(mmp_plan(&bad.p, nullptr, &info) == MMP_ERR_INVALID_ARGUMENT, "missing Raman response refused");
  }
  expect(mmp_plan(nullptr, nullptr, &info) == MMP_ERR_INVALID_ARGUMENT, "NULL params refused");
  expect(mmp_plan(&pr.p, nullptr, nullptr) == MMP_ERR_INVALID_ARGUMENT, "NULL info refused");
  const mmp_options bad_opt{7, 0, 0, 0};
  expect(mmp_plan(&pr.p, &bad_opt, &info) == MMP_ERR_INVALID_ARGUMENT, "unknown precision refused");
  const mmp_options both_mc{MMP_PRECISION_MIXED, 0, MMP_FLAG_MATRIX_CORES | MMP_FLAG_NO_MATRIX_CORES, 0};
  expect(mmp_plan(&pr.p, &both_mc, &info) == MMP_ERR_INVALID_ARGUMENT, "contradicting matrix-core flags refused");
  // no device here: creation fails loudly instead of computing elsewhere
  mmp_ctx* ctx = nullptr;
  const mmp_status st = mmp_ctx_create(0, &ctx);
  expect(st == MMP_ERR_NO_DEVICE || st == MMP_ERR_DEVICE || st == MMP_OK, "context status");
  if (st == MMP_OK) mmp_ctx_destroy(ctx);
  expect(mmp_solver_create(nullptr, &pr.p, nullptr, nullptr) == MMP_ERR_INVALID_ARGUMENT, "NULL context refused");
}

void server() {
  mmp_serve_options o{0, 4, 64, MMP_PRECISION_MIXED, MMP_PAIRS_PHYSICAL, 0, 1e9};
  mmp_server* srv = nullptr;
  {
    mmp_serve_options bad = o;
    bad.slots = 0;
    expect(mmp_serve_create(&bad, &srv) == MMP_ERR_INVALID_ARGUMENT && !srv, "zero slots refused");
  }
  expect(mmp_serve_create(&o, &srv) == MMP_OK, "server created");
  uint64_t ids[3];
  {
    Problem a(3, 4096), b(6, 4096), c(1, 8192);
    expect(mmp_serve_submit(srv, &a.p, 100, &ids[0]) == MMP_OK, "submit a");
    expect(mmp_serve_submit(srv, &b.p, 50, &ids[1]) == MMP_OK, "submit b");
    expect(mmp_serve_submit(srv, &c.p, 10, &ids[2]) == MMP_OK, "submit c");
    // the problems go out of scope: the server keeps its own copies
  }
  Problem bad(3, 3000);
  expect(mmp_serve_submit(srv, &bad.p, 10, nullptr) == MMP_ERR_UNSUPPORTED, "invalid request refused");
  mmp_serve_record rec{};
  expect(mmp_serve_record_get(srv, ids[1], &rec) == MMP_OK && rec.state == MMP_REQ_QUEUED &&
             rec.units == 6ull * 4096 * 50,
         "queued record");
  std::vector<double> out(2 * 6 * 4096);
  expect(mmp_serve_result(srv, ids[1], out.data()) == MMP_ERR_STATE, "no result while queued");
  mmp_serve_stats stats{};
  expect(mmp_serve_stats_get(srv, &stats) == MMP_OK && stats.submitted == 3 && stats.queued == 3, "stats");
  mmp_serve_destroy(srv);

  // a budget no request fits: drain reports it instead of waiting, before any device call
  o.hbm_budget = 1e5;
  expect(mmp_serve_create(&o, &srv) == MMP_OK, "small server created");
  Problem a(3, 4096);
  expect(mmp_serve_submit(srv, &a.p, 10, nullptr) == MMP_OK, "submit to small server");
  expect(mmp_serve_drain(srv) == MMP_ERR_STATE, "request beyond the budget reported");
  mmp_serve_destroy(srv);
}

void concurrent_planning() {
  std::vector<std::thread> th;
  std::vector<int> ok(8, 0);
  for (int t = 0; t < 8; t++)
    th.emplace_back([t, &ok] {
      Problem pr(2 + t % 5, 1024u << (t % 3));
      mmp_info info{};
      int good = 1;
      for (int r = 0; r < 4; r++) good &= mmp_plan(&pr.p, nullptr, &info) == MMP_OK;
      Problem bad(3, 1000);
      good &= mmp_plan(&bad.p, nullptr, &info) == MMP_ERR_UNSUPPORTED && std::strlen(mmp_last_error()) > 0;
      ok[t] = good;
    });
  for (auto& x : th) x.join();
  for (int t = 0; t < 8; t++) expect(ok[t] == 1, "concurrent planning");
}

}  // namespace

int main() {
  planning();
  invalid_inputs();
  server();
  concurrent_planning();
  if (g_bad) {
    std::fprintf(stderr, "%d checks failed\n", g_bad);
    return 1;
  }
  std::printf("host checks passed\n");
  return 0;
}
