// TEST-ONLY: runs every stage kind of the product library (emulated on the
// CPU, see hip/hip_runtime.h) under AddressSanitizer on small synthetic
// problems.  Exit status 0 means no out-of-bounds access was found.
#include <cmath>
#include <cstdio>
#include <vector>

#include "../../include/mcv_hip.h"

static int run(int dof, int n, int scheme, int precision) {
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
  if (mcv_hip_autotune(h, 1, &best, tms)) { std::printf("autotune: %s\n", mcv_hip_last_error()); return 1; }
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
  mcv_hip_free(h);
  std::printf("dof %d n %d scheme %d precision %d ok mass %g\n", dof, n, scheme, precision, mass);
  return 0;
}

int main() {
  setvbuf(stdout, nullptr, _IONBF, 0);
  int rc = 0;
  for (int dof = 3; dof <= 4; ++dof)
    for (int n : {2, 5, 7, 13})
      for (int scheme : {MCV_RK4, MCV_RK3_TVD})
        for (int precision : {MCV_FP64, MCV_FP32C}) rc |= run(dof, n, scheme, precision);
  return rc;
}
