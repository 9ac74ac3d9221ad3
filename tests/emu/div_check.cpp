// TEST-ONLY: the product's exact division by a known divisor (div_by in
// mcv_sw_amd/csrc/mcv_device.hpp), compiled for the CPU against the shim,
// against IEEE division a / b, bit for bit.  Prints the number of
// mismatches and exits nonzero if any.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>

#include "../../mcv_sw_amd/csrc/mcv_device.hpp"

static uint64_t bits(double x) {
  uint64_t u;
  std::memcpy(&u, &x, sizeof u);
  return u;
}

int main() {
  std::mt19937_64 rng(12345);
  std::uniform_real_distribution<double> mant(1.0, 2.0);
  std::uniform_int_distribution<int> ex(-300, 300), small(-40, 40);
  long long n = 0, bad = 0;
  auto check = [&](double a, double b) {
    const mcv::Recip d = mcv::make_recip(b);
    const double got = mcv::div_by(a, d), want = a / b;
    ++n;
    if (bits(got) != bits(want)) {
      if (bad < 10) std::printf("mismatch: %a / %a = %a, div_by %a\n", a, b, want, got);
      ++bad;
    }
  };
  // the scheme's divisors: dx (pi/2/n), 2dx, dx^2 for n = 2..1024, and 27, 6, 3
  std::vector<double> divs = {27.0, 6.0, 3.0, -3.0, 1.0, 0.1};
  for (int nn = 2; nn <= 1024; nn = nn * 3 / 2 + 1) {
    const double dx = (std::atan(1.0) * 2.0) / nn;
    divs.push_back(dx);
    divs.push_back(2.0 * dx);
    divs.push_back(dx * dx);
  }
  for (int k = 0; k < 200; ++k) divs.push_back(std::ldexp(mant(rng), small(rng)));  // sqrtG-like
  for (double b : divs) {
    for (double a : {0.0, -0.0, 1.0, -1.0, b, -b, 3.0 * b, -7.0 * b}) check(a, b);
    for (int k = 0; k < 2000; ++k) {
      const double a = std::ldexp(mant(rng), ex(rng)) * ((k & 1) ? -1.0 : 1.0);
      check(a, b);
      check(a, -b);
      check(std::nearbyint(a), b);  // integer-valued numerators: many exact quotients
    }
  }
  std::printf("%lld divisions, %lld mismatches\n", n, bad);
  return bad ? 1 : 0;
}
