// Host check of the nodal updates in akx_math.hpp: the reference order
// (reference_predict / reference_gather, the plain FP64 path) against the
// FP32 form's half-step velocity (half_step_predict / half_step_gather), on
// a chain of springs with lumped masses, an external force and two blocked
// components, over one call of 2,000 steps.  The half-step run is repeated
// with the same steps split over calls (37, 1, 2, 1, ... steps).
// Prints: the largest relative differences of u, v and a between the two
// forms, and 1 if the split run has the bits of the single call (else 0).
#include "../../akantu2_amd/csrc/akx_math.hpp"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

namespace {

constexpr int N = 64;        // DOFs of the chain
constexpr int STEPS = 2000;

struct Chain {
  std::vector<double> M, k, fext;
  std::vector<char> blocked;
  double dt;
  // f_int = -K u, summed in index order
  void fint(const std::vector<double> &u, std::vector<double> &f) const {
    for (int i = 0; i < N; ++i) {
      double s = 0.;
      if (i > 0)
        s += k[i] * (u[i - 1] - u[i]);
      if (i + 1 < N)
        s += k[i + 1] * (u[i + 1] - u[i]);
      f[i] = s;
    }
  }
};

struct State {
  std::vector<double> u, v, a;
  std::vector<float> cu, cv;
};

// one call of n steps, reference order
void run_reference(const Chain &c, State &s, int n) {
  std::vector<double> f(N);
  for (int i = 0; i < N; ++i)
    if (!c.blocked[i])
      akx::reference_predict(s.a[i], c.dt, s.v[i], s.u[i]);
  for (int k = 0; k < n; ++k) {
    c.fint(s.u, f);
    for (int i = 0; i < N; ++i) {
      if (c.blocked[i])
        continue;
      if (k + 1 < n)
        akx::reference_gather<true>(c.fext[i], f[i], c.M[i], c.dt, s.a[i], s.v[i], s.u[i]);
      else
        akx::reference_gather<false>(c.fext[i], f[i], c.M[i], c.dt, s.a[i], s.v[i], s.u[i]);
    }
  }
}

// one call of n steps, half-step velocity
void run_half_step(const Chain &c, State &s, int n) {
  std::vector<double> f(N);
  for (int i = 0; i < N; ++i)
    if (!c.blocked[i])
      akx::half_step_predict(s.a[i], c.dt, s.v[i], s.cv[i], s.u[i], s.cu[i]);
  for (int k = 0; k < n; ++k) {
    c.fint(s.u, f);
    for (int i = 0; i < N; ++i) {
      if (c.blocked[i])
        continue;
      const double ftot = c.fext[i] + f[i];
      if (k + 1 < n)
        akx::half_step_gather<true>(ftot, c.M[i], c.dt, s.v[i], s.cv[i], s.u[i], s.cu[i],
                                    s.a[i]);
      else
        akx::half_step_gather<false>(ftot, c.M[i], c.dt, s.v[i], s.cv[i], s.u[i], s.cu[i],
                                     s.a[i]);
    }
  }
}

double relerr(const std::vector<double> &x, const std::vector<double> &ref) {
  double num = 0., den = 0.;
  for (int i = 0; i < N; ++i) {
    num = std::fmax(num, std::fabs(x[i] - ref[i]));
    den = std::fmax(den, std::fabs(ref[i]));
  }
  return den > 0. ? num / den : num;
}

bool same_bits(const std::vector<double> &x, const std::vector<double> &y) {
  return std::memcmp(x.data(), y.data(), N * sizeof(double)) == 0;
}

} // namespace

int main() {
  std::mt19937_64 rng(11);
  std::uniform_real_distribution<double> U(0., 1.);
  Chain c;
  c.M.resize(N);
  c.k.resize(N + 1);
  c.fext.resize(N);
  c.blocked.assign(N, 0);
  for (int i = 0; i < N; ++i) {
    c.M[i] = 0.5 + 1.5 * U(rng);
    c.fext[i] = 1e-3 * (U(rng) - 0.5);
  }
  for (auto &x : c.k)
    x = 0.5 + U(rng);
  c.blocked[0] = c.blocked[N / 2] = 1;
  // central difference is stable for dt < 2 / omega_max, omega_max^2 <=
  // max over i of (k_i + k_{i+1}) / M_i * 2
  double w2 = 0.;
  for (int i = 0; i < N; ++i)
    w2 = std::fmax(w2, 2. * (c.k[i] + c.k[i + 1]) / c.M[i]);
  c.dt = 0.5 * 2. / std::sqrt(w2);
  State s0;
  s0.u.assign(N, 0.);
  s0.v.resize(N);
  s0.a.assign(N, 0.);
  s0.cu.assign(N, 0.f);
  s0.cv.assign(N, 0.f);
  for (int i = 0; i < N; ++i)
    s0.v[i] = c.blocked[i] ? 0. : 1e-2 * std::sin(0.3 * i) * (0.5 + U(rng));

  State ref = s0, one = s0, split = s0;
  run_reference(c, ref, STEPS);
  run_half_step(c, one, STEPS);
  const int parts[] = {37, 1, 2, 1, 500, 3, 1, 455};
  int done = 0;
  for (int p : parts) {
    run_half_step(c, split, p);
    done += p;
  }
  run_half_step(c, split, STEPS - done);
  const bool bits = same_bits(one.u, split.u) && same_bits(one.v, split.v) &&
                    same_bits(one.a, split.a);
  std::printf("%.3e %.3e %.3e %d %.3e\n", relerr(one.u, ref.u), relerr(one.v, ref.v),
              relerr(one.a, ref.a), bits ? 1 : 0, std::fabs(ref.u[N - 1]));
  return 0;
}
