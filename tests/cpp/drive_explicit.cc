// Drives the explicit path through the C++ host API (include/akantu_mi355x.hpp)
// the way an Akantu driver does (examples/c++/solid_mechanics_model/explicit/
// explicit_dynamic.cc): build the model, initFull, stable time step, solveStep
// loop, read the nodal arrays back.  Inputs and outputs are raw binary files
// written and read by tests/test_gpu_cpp_api.py.
//
//   drive_explicit <dir> <etype 0|1> <law name> <finite_deformation 0|1> <steps> [batch|held]
//
// <law name> "local_elastic" is a ConstitutiveLaw written here (the
// reference's Material subclass overriding computeStress), the small-strain
// elastic law in the reference's operation order; "device_elastic" is the
// same material with the law given as device code (Material::setDeviceStress),
// whose host computeStress must then never run.
//
// With "batch" the same simulation runs as one instance of an FP64 Batch
// (include/akantu_mi355x.hpp), beside a second instance (the same mesh at
// rest) that must not change its result.
//
// With "held" the driver takes `auto & u = model.getDisplacement()` before the
// loop, runs solveStep("explicit_lumped") and, after steps / 2 steps, adds
// (2e-5, -1e-5, 3e-5) to u(7) through the held reference, as a reference
// driver writing through its Array would.
//
// <dir>/nodes.f64 [nn][3], conn.i32 [nel][nnpe], blocked.u8 [nn][3],
// v0.f64 [nn][3], params.f64 {rho, E, nu, sigma_y, h, dt_factor}
// -> <dir>/u.f64, v.f64, fint.f64, stress.f64, dt.f64
#include "akantu_mi355x.hpp"

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

using namespace akantu_mi355x;

// sigma = lambda tr(grad u) I + 2 mu sym(grad u) (material_elastic_inline_impl.hh:33-49)
class LocalElastic : public ConstitutiveLaw {
public:
  LocalElastic(double E, double nu)
      : lambda(nu * E / ((1. + nu) * (1. - 2. * nu))), mu(E / (2. * (1. + nu))) {}
  void computeStress(const Array<double> &grad_u, Array<double> &stress) override {
    ++calls;
    for (std::int64_t q = 0; q < grad_u.size(); ++q) {
      auto g = [&](int i, int j) { return grad_u(q, i + 3 * j); };
      const double tr = g(0, 0) + g(1, 1) + g(2, 2), two_mu = 2. * mu;
      for (int j = 0; j < 3; ++j)
        for (int i = 0; i < 3; ++i) {
          const double eps = .5 * (g(j, i) + g(i, j));
          stress(q, i + 3 * j) = i == j ? two_mu * eps + lambda * tr : two_mu * eps;
        }
    }
  }
  double lambda, mu;
  long calls = 0;
};

// the same law for one Gauss point, as device code (params: lambda, mu)
const char *kDeviceElastic = R"(
__device__ void akx_compute_stress(const double *g, double *s, const double *p) {
  const double tr = g[0] + g[4] + g[8], two_mu = 2. * p[1];
  for (int j = 0; j < 3; ++j)
    for (int i = 0; i < 3; ++i) {
      const double eps = .5 * (g[j + 3 * i] + g[i + 3 * j]);
      s[i + 3 * j] = i == j ? two_mu * eps + p[0] * tr : two_mu * eps;
    }
}
)";

template <class T> static std::vector<T> load(const std::string &path) {
  std::ifstream f(path, std::ios::binary | std::ios::ate);
  if (!f)
    throw std::runtime_error("cannot open " + path);
  std::streamsize n = f.tellg();
  f.seekg(0);
  std::vector<T> v(n / sizeof(T));
  f.read(reinterpret_cast<char *>(v.data()), n);
  return v;
}

template <class T> static void save(const std::string &path, const T *p, size_t n) {
  std::ofstream f(path, std::ios::binary);
  f.write(reinterpret_cast<const char *>(p), n * sizeof(T));
}

int main(int argc, char **argv) {
  if (argc != 6 && argc != 7) {
    std::cerr << "usage: drive_explicit <dir> <etype> <law> <finite_deformation> <steps> "
                 "[batch|held]\n";
    return 2;
  }
  const bool batch = argc == 7 && std::string(argv[6]) == "batch";
  const bool held = argc == 7 && std::string(argv[6]) == "held";
  const std::string dir = argv[1];
  const int etype = std::atoi(argv[2]);
  const std::string law = argv[3];
  const bool fd = std::atoi(argv[4]) != 0;
  const long steps = std::atol(argv[5]);
  try {
    const int nnpe = etype == 1 ? 8 : 4;
    auto X = load<double>(dir + "/nodes.f64");
    auto C = load<std::int32_t>(dir + "/conn.i32");
    auto B = load<std::uint8_t>(dir + "/blocked.u8");
    auto V = load<double>(dir + "/v0.f64");
    auto P = load<double>(dir + "/params.f64");
    const std::int64_t nn = X.size() / 3, nel = C.size() / nnpe;

    Mesh mesh{etype == 1 ? _hexahedron_8 : _tetrahedron_4, Array<double>(nn, 3),
              Array<std::int32_t>(nel, nnpe)};
    std::copy(X.begin(), X.end(), mesh.nodes.data());
    std::copy(C.begin(), C.end(), mesh.connectivity.data());

    if (batch) {
      Array<std::uint8_t> blocked(nn, 3);
      Array<double> velocity(nn, 3);
      std::copy(B.begin(), B.end(), blocked.data());
      std::copy(V.begin(), V.end(), velocity.data());
      const std::int64_t tets = etype == 0 ? 2 * nel : 0, hexes = etype == 1 ? 2 * nel : 0;
      Batch b(Batch::fp64, 2 * nn + 512, tets, hexes);
      const int other = b.addInstance(mesh, law, {P[0], P[1], P[2], P[3], P[4]}, &blocked,
                                      nullptr, P[5], fd);
      const int id = b.addInstance(mesh, law, {P[0], P[1], P[2], P[3], P[4]}, &blocked,
                                   &velocity, P[5], fd);
      b.solveStep(steps / 2);
      b.solveStep(steps - steps / 2);
      save(dir + "/u.f64", b.getDisplacement(id).data(), 3 * nn);
      save(dir + "/v.f64", b.getVelocity(id).data(), 3 * nn);
      save(dir + "/fint.f64", b.getInternalForce(id).data(), 3 * nn);
      const double dt = b.getTimeStep(id);
      save(dir + "/dt.f64", &dt, 1);
      auto rest = b.getDisplacement(other);
      for (std::int64_t i = 0; i < 3 * nn; ++i)
        if (rest.data()[i] != 0.)
          throw std::runtime_error("the instance at rest moved");
      return 0;
    }
    SolidMechanicsModel model(mesh);
    LocalElastic local(P[1], P[2]);
    if (law == "local_elastic") {
      model.addMaterial(local, {P[0], P[1], P[2]}, fd);
    } else if (law == "device_elastic") {
      Material &mat = model.addMaterial(local, {P[0], P[1], P[2]}, fd);
      mat.setDeviceStress(kDeviceElastic, {local.lambda, local.mu});
    } else
      model.addMaterial(law, {P[0], P[1], P[2], P[3], P[4]}, fd);
    model.initFull(_explicit_lumped_mass);
    auto &blocked = model.getBlockedDOFs();
    auto &velocity = model.getVelocity();
    for (std::int64_t i = 0; i < nn; ++i)
      for (int c = 0; c < 3; ++c) {
        blocked(i, c) = B[3 * i + c];
        velocity(i, c) = V[3 * i + c];
      }
    const double dt = P[5] * model.getStableTimeStep();
    model.setTimeStep(dt);
    if (held) {
      auto &u = model.getDisplacement();
      for (long s = 0; s < steps; ++s) {
        if (s == steps / 2) {
          const double bump[3] = {2e-5, -1e-5, 3e-5};
          for (int c = 0; c < 3; ++c)
            u(7, c) += bump[c];
        }
        model.solveStep("explicit_lumped");
      }
    } else {
      for (long s = 0; s < steps; ++s)
        model.solveStep();
    }
    save(dir + "/u.f64", model.getDisplacement().data(), 3 * nn);
    save(dir + "/v.f64", model.getVelocity().data(), 3 * nn);
    save(dir + "/fint.f64", model.getInternalForce().data(), 3 * nn);
    auto stress = model.getMaterial(0).getStress();
    save(dir + "/stress.f64", stress.data(), (size_t)stress.size() * 9);
    save(dir + "/dt.f64", &dt, 1);
    if (law == "device_elastic" && local.calls != 0)
      throw std::runtime_error("the host computeStress ran beside the device law");
  } catch (const JacobianException &e) {
    std::cerr << "jacobian: " << e.what() << "\n";
    return 3;
  } catch (const std::exception &e) {
    std::cerr << "error: " << e.what() << "\n";
    return 1;
  }
  return 0;
}
