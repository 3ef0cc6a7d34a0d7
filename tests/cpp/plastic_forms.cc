// Host check of the two linear-isotropic-hardening updates in akx_math.hpp:
// the reference's incremental form (plastic_lih) and the total-strain form
// with the compact state (plastic_lih_total), driven along the same seeded
// strain path through loading, yielding and unloading.  Prints the largest
// relative difference of the stress and of the plastic strain, and the
// number of yielding steps.
#include "../../akantu2_amd/csrc/akx_math.hpp"

#include <cmath>
#include <cstdio>
#include <random>

int main() {
  using namespace akx;
  std::mt19937_64 rng(7);
  std::uniform_real_distribution<double> U(-1., 1.);
  double lambda, mu;
  lame(1.3, 0.27, lambda, mu);
  const double sy = 2e-3, h = 0.05;
  double dir[3][3];
  for (auto &r : dir)
    for (auto &x : r)
      x = U(rng);
  double pg[3][3] = {}, ps[3][3] = {}, pe[3][3] = {}, piso = 0.;
  double pe6[6] = {}, iso6 = 0.;
  double max_ds = 0, max_de = 0, smax = 0, emax = 0;
  int yielding = 0;
  for (int k = 0; k < 400; ++k) {
    // load to 6e-3 along dir with noise, unload, reload in another direction
    const double amp = k < 150 ? 4e-5 * k : (k < 250 ? 4e-5 * (300 - k) : 4e-5 * (k - 200));
    double g[3][3];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        g[i][j] = amp * dir[i][j] * (k < 250 ? 1. : (i == j ? -1. : 1.)) + 1e-6 * U(rng);
    double S[3][3], eps[3][3], iso;
    plastic_lih(lambda, mu, sy, h, g, pg, ps, piso, pe, S, iso, eps);
    double S2[3][3], pn[6], iso2;
    plastic_lih_total(lambda, mu, sy, h, g, pe6, iso6, S2, pn, iso2);
    yielding += iso > piso;
    const int vi[6] = {0, 1, 2, 1, 0, 0}, vj[6] = {0, 1, 2, 2, 2, 1};
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        max_ds = std::fmax(max_ds, std::fabs(S[i][j] - S2[i][j]));
        smax = std::fmax(smax, std::fabs(S[i][j]));
      }
    for (int c = 0; c < 6; ++c) {
      max_de = std::fmax(max_de, std::fabs(eps[vi[c]][vj[c]] - pn[c]));
      emax = std::fmax(emax, std::fabs(eps[vi[c]][vj[c]]));
    }
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        pg[i][j] = g[i][j];
        ps[i][j] = S[i][j];
        pe[i][j] = eps[i][j];
      }
    piso = iso;
    for (int c = 0; c < 6; ++c)
      pe6[c] = pn[c];
    iso6 = iso2;
  }
  std::printf("%.3e %.3e %d %.3e\n", max_ds / smax, max_de / (emax > 0 ? emax : 1.), yielding,
              std::fabs(piso - iso6) / (piso > 0 ? piso : 1.));
  return 0;
}
