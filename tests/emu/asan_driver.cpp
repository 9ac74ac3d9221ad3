// TEST-ONLY: runs every stage kind of the product library (emulated on the
// CPU, see hip/hip_runtime.h) under AddressSanitizer on small synthetic
// problems.  Exit status 0 means no out-of-bounds access was found.  The
// same driver, built with threaded blocks and ThreadSanitizer (tsan_driver),
// runs a smaller set ("threads") and fails on any intra-block data race.
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/mcv_hip.h"

static int run(int dof, int n, int scheme, int precision, bool tune = true) {
  const int H = dof - 1, npv = n * H + 1, nh = npv + 2 * H;
  const size_t np = (size_t)6 * nh * nh;
  std::vector<double> sg(np), G(4 * np), IG(4 * np), A(4 * np), IA(4 * np), f(np), ps(np);
  std::vector<double> phi(np), phiG(np), u(np), v(np), cu(np), cv(np), coef(H * npv);
  std::vector<int32_t> iref(H * npv);
  for (size_t k = 0; k < np; ++k) {
    double x = 1e-3 * (double)(k % 977);
    sg[k] = 1.0 + 0.1 * std::sin(x);
    double a = 1.0 + 0.05 * std::cos(x), b = 0.01 * std::sin(2 * x);
    G[4 * k] = a; G[4 * k + 1] = b; G[4 * k + 2] = b; G[4 * k + 3] = a;
    IG[4 * k] = 1.0 / a; IG[4 * k + 1] = -b; IG[4 * k + 2] = -b; IG[4 * k + 3] = 1.0 / a;
    A[4 * k] = 1.0; A[4 * k + 1] = b; A[4 * k + 2] = 0.0; A[4 * k + 3] = 1.0;
    IA[4 * k] = 1.0; IA[4 * k + 1] = -b; IA[4 * k + 2] = 0.0; IA[4 * k + 3] = 1.0;
    f[k] = 1e-4 * std::sin(x);
    ps[k] = 0.0;
    phi[k] = 5e4 + 10.0 * std::cos(3 * x);
    phiG[k] = phi[k] * sg[k];
    u[k] = 10.0 * std::sin(x);
    v[k] = 5.0 * std::cos(x);
    cu[k] = IG[4 * k] * u[k] + IG[4 * k + 2] * v[k];
    cv[k] = IG[4 * k + 1] * u[k] + IG[4 * k + 3] * v[k];
  }
  const double dx = (M_PI / 2) / n;
  for (int k = 0; k < H; ++k)
    for (int s = 0; s < npv; ++s) {
      int c = 1 + (s / H) % n;  // any valid cell
      iref[k * npv + s] = c;
      coef[k * npv + s] = 0.3 * dx;
    }
  mcv_config cfg{MCV_ABI_VERSION, dof, n, scheme, precision, 0, 1e-6, dx};
  mcv_mesh mesh{sg.data(), G.data(), IG.data(), A.data(), IA.data(), f.data(), ps.data(),
                coef.data(), iref.data()};
  mcv_state st{phi.data(), phiG.data(), u.data(), v.data(), cu.data(), cv.data()};
  mcv_handle *h = nullptr;
  if (mcv_hip_init(&cfg, &mesh, &st, &h)) { std::printf("init: %s\n", mcv_hip_last_error()); return 1; }
  int32_t best = -1;
  double tms[MCV_TILE_VARIANTS];
  if (tune && mcv_hip_autotune(h, 1, &best, tms)) { std::printf("autotune: %s\n", mcv_hip_last_error()); return 1; }
  for (int v = 0; v < MCV_TILE_VARIANTS; ++v) {  // every tile shape through every stage kind
    if (mcv_hip_set_tile(h, v)) return 1;
    if (mcv_hip_step(h, 1)) { std::printf("step: %s\n", mcv_hip_last_error()); return 1; }
  }
  double mass, energy;
  std::vector<double> z(np), m(np), t(3 * (size_t)6 * npv * npv);
  if (mcv_hip_sync(h, &st, &mass, &energy)) { std::printf("sync: %s\n", mcv_hip_last_error()); return 1; }
  if (mcv_hip_sphere_winds(h, z.data(), m.data())) { std::printf("winds: %s\n", mcv_hip_last_error()); return 1; }
  const int trc = mcv_hip_tendency(h, &st, t.data(), t.data() + 6 * npv * npv, t.data() + 12 * npv * npv);
  if (precision == MCV_FP64 ? trc != MCV_OK : trc != MCV_EINVAL) { std::printf("tend: %d %s\n", trc, mcv_hip_last_error()); return 1; }
  // the allocation-free seam: same bits through the library's static scratch
  std::vector<double> t2(t.size());
  const int nrc = mcv_hip_tendency_noalloc(h, &st, t2.data(), t2.data() + 6 * npv * npv,
                                           t2.data() + 12 * npv * npv);
  if (precision == MCV_FP64 ? (nrc != MCV_OK || t2 != t) : nrc != MCV_EINVAL) {
    std::printf("tend noalloc: %d %s\n", nrc, mcv_hip_last_error());
    return 1;
  }
  mcv_hip_free(h);
  std::printf("dof %d n %d scheme %d precision %d ok mass %g\n", dof, n, scheme, precision, mass);
  return 0;
}

// Batch of mixed solves: graphs, autotune, admission mid-step, slot reuse.
static int run_batch() {
  struct Spec { int dof, n, scheme, precision; int64_t steps; };
  const Spec specs[] = {{3, 5, MCV_RK4, MCV_FP64, 2}, {4, 3, MCV_RK3_TVD, MCV_FP32C, 3},
                        {4, 7, MCV_RK4, MCV_FP32C, 1}, {3, 2, MCV_RK3_TVD, MCV_FP64, 2}};
  mcv_batch *b = nullptr;
  if (mcv_batch_create(0, 3, &b)) { std::printf("create: %s\n", mcv_hip_last_error()); return 1; }
  std::vector<std::vector<double>> keep;
  int32_t slots[4];
  for (int k = 0; k < 4; ++k) {
    const Spec &sp = specs[k];
    const int H = sp.dof - 1, npv = sp.n * H + 1, nh = npv + 2 * H;
    const size_t np = (size_t)6 * nh * nh;
    std::vector<double> sg(np), G(4 * np), IG(4 * np), A(4 * np), IA(4 * np), f(np), ps(np);
    std::vector<double> phi(np), phiG(np), u(np), v(np), cu(np), cv(np), coef(H * npv);
    std::vector<int32_t> iref(H * npv);
    for (size_t q = 0; q < np; ++q) {
      double x = 1e-3 * (double)(q % 977);
      sg[q] = 1.0 + 0.1 * std::sin(x);
      G[4 * q] = 1.0; G[4 * q + 1] = 0.0; G[4 * q + 2] = 0.0; G[4 * q + 3] = 1.0;
      IG[4 * q] = 1.0; IG[4 * q + 1] = 0.0; IG[4 * q + 2] = 0.0; IG[4 * q + 3] = 1.0;
      A[4 * q] = 1.0; A[4 * q + 1] = 0.0; A[4 * q + 2] = 0.0; A[4 * q + 3] = 1.0;
      IA[4 * q] = 1.0; IA[4 * q + 1] = 0.0; IA[4 * q + 2] = 0.0; IA[4 * q + 3] = 1.0;
      f[q] = 1e-4; ps[q] = 0.0;
      phi[q] = 5e4 + 10.0 * std::cos(3 * x); phiG[q] = phi[q] * sg[q];
      u[q] = 1.0; v[q] = 0.5; cu[q] = u[q]; cv[q] = v[q];
    }
    const double dx = (M_PI / 2) / sp.n;
    for (int q = 0; q < H * npv; ++q) { iref[q] = 1 + (q / H) % sp.n; coef[q] = 0.3 * dx; }
    mcv_config cfg{MCV_ABI_VERSION, sp.dof, sp.n, sp.scheme, sp.precision, 0, 1e-6, dx};
    mcv_mesh mesh{sg.data(), G.data(), IG.data(), A.data(), IA.data(), f.data(), ps.data(),
                  coef.data(), iref.data()};
    mcv_state st{phi.data(), phiG.data(), u.data(), v.data(), cu.data(), cv.data()};
    if (k == 3) {  // the batch is full until slot 0 is retired
      if (mcv_batch_add(b, &cfg, &mesh, &st, sp.steps, &slots[k]) != MCV_ESTATE) return 1;
      int64_t left[3];
      for (int it = 0; it < 50; ++it) {
        if (mcv_batch_poll(b, left, nullptr)) return 1;
        if (left[slots[0]] == 0) break;
        if (mcv_batch_run(b, 1)) return 1;
      }
      if (mcv_batch_remove(b, slots[0])) return 1;
    }
    if (mcv_batch_add(b, &cfg, &mesh, &st, sp.steps, &slots[k])) { std::printf("add: %s\n", mcv_hip_last_error()); return 1; }
    if (k == 0) {
      int32_t t;
      if (mcv_batch_autotune(b, 1, &t, nullptr)) { std::printf("autotune: %s\n", mcv_hip_last_error()); return 1; }
    }
    if (mcv_batch_run(b, k == 1 ? 13 : 2)) { std::printf("run: %s\n", mcv_hip_last_error()); return 1; }
  }
  if (mcv_batch_run(b, 30)) return 1;
  int64_t left[3], done[3];
  if (mcv_batch_poll(b, left, done)) return 1;
  for (int k = 1; k < 4; ++k) {
    const int H = specs[k].dof - 1, npv = specs[k].n * H + 1, nh = npv + 2 * H;
    std::vector<double> phi((size_t)6 * nh * nh);
    mcv_state o{phi.data(), nullptr, nullptr, nullptr, nullptr, nullptr};
    double mass;
    if (left[slots[k]] != 0 || done[slots[k]] != specs[k].steps) { std::printf("batch progress\n"); return 1; }
    if (mcv_batch_sync(b, slots[k], &o, &mass, nullptr)) { std::printf("bsync: %s\n", mcv_hip_last_error()); return 1; }
  }
  mcv_batch_free(b);
  std::printf("batch ok\n");
  return 0;
}

// Graph lifetime across the serving loop's sequence: graph replays with
// admissions and retirements between them, the work list growing and
// shrinking (re-capture), tile-shape and graph on/off switches, a new kind of
// solve joining (a different compiled kernel), autotuning, and teardown with
// solves still resident.
static int run_batch_graphs() {
  struct Spec { int dof, n, scheme, precision; int64_t steps; };
  const Spec small = {3, 3, MCV_RK3_TVD, MCV_FP32C, 60}, big = {3, 9, MCV_RK3_TVD, MCV_FP32C, 8},
             other = {4, 4, MCV_RK4, MCV_FP64, 6};
  std::vector<std::vector<double>> keep;
  auto add = [&](mcv_batch *b, const Spec &sp, int32_t *slot) -> int {
    const int H = sp.dof - 1, npv = sp.n * H + 1, nh = npv + 2 * H;
    const size_t np = (size_t)6 * nh * nh;
    std::vector<double> sg(np, 1.0), G(4 * np, 0.0), A(4 * np, 0.0), f(np, 1e-4), ps(np, 0.0);
    std::vector<double> phi(np), phiG(np), u(np, 1.0), v(np, 0.5), coef(H * npv);
    std::vector<int32_t> iref(H * npv);
    for (size_t q = 0; q < np; ++q) {
      G[4 * q] = G[4 * q + 3] = A[4 * q] = A[4 * q + 3] = 1.0;
      phi[q] = 5e4 + 10.0 * std::cos(3e-3 * (double)(q % 977));
      phiG[q] = phi[q];
    }
    const double dx = (M_PI / 2) / sp.n;
    for (int q = 0; q < H * npv; ++q) { iref[q] = 1 + (q / H) % sp.n; coef[q] = 0.3 * dx; }
    mcv_config cfg{MCV_ABI_VERSION, sp.dof, sp.n, sp.scheme, sp.precision, 0, 1e-6, dx};
    mcv_mesh mesh{sg.data(), G.data(), G.data(), A.data(), A.data(), f.data(), ps.data(),
                  coef.data(), iref.data()};
    mcv_state st{phi.data(), phiG.data(), u.data(), v.data(), u.data(), v.data()};
    return mcv_batch_add(b, &cfg, &mesh, &st, sp.steps, slot);
  };
  mcv_batch *b = nullptr;
  if (mcv_batch_create(0, 4, &b)) return 1;
  int32_t s0, s1, s2;
#define MCV_OK_OR_FAIL(x)                                                   \
  do {                                                                      \
    if ((x) != MCV_OK) {                                                    \
      std::printf("%s: %s\n", #x, mcv_hip_last_error());                   \
      return 1;                                                             \
    }                                                                       \
  } while (0)
  MCV_OK_OR_FAIL(mcv_batch_set_graphs(b, 1));
  MCV_OK_OR_FAIL(add(b, small, &s0));
  MCV_OK_OR_FAIL(mcv_batch_run(b, 12));  // captured for the small list
  MCV_OK_OR_FAIL(add(b, big, &s1));
  MCV_OK_OR_FAIL(mcv_batch_run(b, 24));  // grown: re-captured
  MCV_OK_OR_FAIL(mcv_batch_set_tile(b, 2));
  MCV_OK_OR_FAIL(mcv_batch_run(b, 12));  // another tile shape: re-captured
  int64_t left[4], done[4];
  for (int it = 0; it < 20; ++it) {
    MCV_OK_OR_FAIL(mcv_batch_poll(b, left, done));
    if (left[s1] == 0) break;
    MCV_OK_OR_FAIL(mcv_batch_run(b, 12));
  }
  MCV_OK_OR_FAIL(mcv_batch_remove(b, s1));
  MCV_OK_OR_FAIL(mcv_batch_run(b, 12));  // shrunk: re-captured
  MCV_OK_OR_FAIL(mcv_batch_set_graphs(b, 0));
  MCV_OK_OR_FAIL(mcv_batch_run(b, 5));
  MCV_OK_OR_FAIL(mcv_batch_set_graphs(b, 1));
  MCV_OK_OR_FAIL(mcv_batch_run(b, 1));   // back to a step boundary of the RK3 solve
  MCV_OK_OR_FAIL(add(b, other, &s2));     // another kind: another kernel, re-captured
  MCV_OK_OR_FAIL(mcv_batch_run(b, 24));
  MCV_OK_OR_FAIL(mcv_batch_poll(b, left, done));
  mcv_batch_free(b);                      // solves still resident
#undef MCV_OK_OR_FAIL
  std::printf("batch graphs ok\n");
  return 0;
}

int main(int argc, char **argv) {
  setvbuf(stdout, nullptr, _IONBF, 0);
  int rc = 0;
  // "threads": the smaller set the threaded (ThreadSanitizer) build runs
  if (argc > 1 && std::string(argv[1]) == "threads") {
    // every stage kind (RK4 in FP64, RK3 in FP32C), each order once, every
    // tile shape on a partial tile, the boundary pass, diagnostics, a batch
    rc |= run(3, 4, MCV_RK4, MCV_FP64, false);
    rc |= run(4, 4, MCV_RK3_TVD, MCV_FP32C, false);
    return rc | run_batch();
  }
  for (int dof = 3; dof <= 4; ++dof)
    for (int n : {2, 5, 7, 13})
      for (int scheme : {MCV_RK4, MCV_RK3_TVD})
        for (int precision : {MCV_FP64, MCV_FP32C}) rc |= run(dof, n, scheme, precision);
  rc |= run_batch();
  rc |= run_batch_graphs();
  return rc;
}
