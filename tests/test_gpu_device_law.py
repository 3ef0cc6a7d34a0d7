"""User laws run on the device (akx_model_set_device_stress).

The reference's computeStress plugin runs on the host; here a user material
may instead give its law as HIP C++ for one Gauss point, compiled at run
time (hiprtc, no floating-point contraction) and run by the element pass's
stress phase without a host round trip.

- A device linear-elastic law in the built-in operation order gives the
  built-in trajectory bit for bit (Tet4, Hex8), and its Python computeStress
  is never called; a device Saint-Venant-Kirchhoff law (PK2) the built-in
  finite-strain elastic one.
- A nonlinear law with only + - * / gives the same bits on the device as the
  same law through the host callback (numpy, same order).
- A source that does not compile raises with the compiler's log; the
  material keeps its host law.  source None returns to the host law.
"""
import numpy as np
import pytest

import akantu2_amd as aka
from akantu2_amd import _native as N
from tests.helpers import make_gpu, workload
from tests.test_gpu_user_material import LocalElastic, _user_model

pytestmark = pytest.mark.gpu

ELASTIC_SRC = r"""
// sigma = lambda tr(eps) I + 2 mu eps (material_elastic_inline_impl.hh:33-49)
__device__ void akx_compute_stress(const double *g, double *s, const double *p) {
  const double lambda = p[0], two_mu = 2.0 * p[1];
  const double tr = (g[0] + g[4]) + g[8];
  for (int j = 0; j < 3; ++j)
    for (int i = 0; i < 3; ++i) {
      const double eps = 0.5 * (g[i + 3 * j] + g[j + 3 * i]);
      s[i + 3 * j] = i == j ? two_mu * eps + lambda * tr : two_mu * eps;
    }
}
"""

SVK_SRC = r"""
// S = lambda tr(E) I + 2 mu E, E = (g + g^T + g^T g) / 2
__device__ void akx_compute_stress(const double *gt, double *S, const double *p) {
  const double lambda = p[0], two_mu = 2.0 * p[1];
  double g[3][3], E[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      g[i][j] = gt[i + 3 * j];  // du_i/dX_j
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double q = g[0][i] * g[0][j];
      q = q + g[1][i] * g[1][j];
      q = q + g[2][i] * g[2][j];
      E[i][j] = (q + g[j][i] + g[i][j]) / 2.0;
    }
  const double tr = (E[0][0] + E[1][1]) + E[2][2];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      const double eps = 0.5 * (E[j][i] + E[i][j]);
      S[j + 3 * i] = i == j ? two_mu * eps + lambda * tr : two_mu * eps;
    }
}
"""

# a nonlinear law: sigma = lambda tr I + 2 mu eps + c (eps : eps) eps
CUBIC_SRC = r"""
__device__ void akx_compute_stress(const double *g, double *s, const double *p) {
  const double lambda = p[0], two_mu = 2.0 * p[1], c = p[2];
  double eps[9];
  for (int j = 0; j < 3; ++j)
    for (int i = 0; i < 3; ++i)
      eps[i + 3 * j] = 0.5 * (g[i + 3 * j] + g[j + 3 * i]);
  double ee = 0.;
  for (int k = 0; k < 9; ++k)
    ee = ee + eps[k] * eps[k];
  const double tr = (eps[0] + eps[4]) + eps[8];
  for (int k = 0; k < 9; ++k) {
    double v = two_mu * eps[k] + (c * ee) * eps[k];
    if (k == 0 || k == 4 || k == 8)
      v = v + lambda * tr;
    s[k] = v;
  }
}
"""


class DeviceElastic(LocalElastic):
    device_stress_source = ELASTIC_SRC

    def deviceStressParams(self):
        return (self.lame_lambda, self.lame_mu)

    def computeStress(self, el_type, ghost_type):
        raise AssertionError("the host law must not run")


class DeviceSVK(DeviceElastic):
    device_stress_source = SVK_SRC


class HostCubic(LocalElastic):
    """The cubic law in numpy, in the device source's order."""
    C = 40.0

    def computeStress(self, el_type, ghost_type):
        self.calls += 1
        g = self.getGradU(el_type, ghost_type)  # [q][9], i + 3 j
        s = self.getStress(el_type, ghost_type)
        idx = np.arange(9)
        i, j = idx % 3, idx // 3
        eps = 0.5 * (g[:, i + 3 * j] + g[:, j + 3 * i])
        ee = np.zeros(g.shape[0])
        for k in range(9):
            ee = ee + eps[:, k] * eps[:, k]
        tr = (eps[:, 0] + eps[:, 4]) + eps[:, 8]
        two_mu = 2.0 * self.lame_mu
        for k in range(9):
            v = two_mu * eps[:, k] + (self.C * ee) * eps[:, k]
            if k in (0, 4, 8):
                v = v + self.lame_lambda * tr
            s[:, k] = v


class DeviceCubic(HostCubic):
    device_stress_source = CUBIC_SRC

    def deviceStressParams(self):
        return (self.lame_lambda, self.lame_mu, self.C)

    def computeStress(self, el_type, ghost_type):
        raise AssertionError("the host law must not run")


FACTORY = aka.MaterialFactory.getInstance()
for name, cls in (("device_elastic", DeviceElastic), ("device_svk", DeviceSVK),
                  ("host_cubic", HostCubic), ("device_cubic", DeviceCubic)):
    FACTORY.registerAllocator(name, lambda dim, unused, model, _id, cls=cls: cls(model, _id))


@pytest.mark.parametrize("etype,shape", [("_tetrahedron_4", (8, 3, 3)),
                                         ("_hexahedron_8", (6, 3, 2))])
def test_device_elastic_equals_builtin(etype, shape):
    mesh, blocked, v0 = workload(etype, shape, jitter=0.1, seed=71)
    mat = dict(rho=1.0, E=1.3, nu=0.27)
    ref, dt = make_gpu(mesh, blocked, v0, "elastic", **mat)
    dev, law = _user_model(mesh, blocked, v0, "device_elastic", dt, **mat)
    for n in (1, 199):
        ref.solveSteps(n)
        dev.solveSteps(n)
        for f in ("getDisplacement", "getVelocity", "getInternalForce"):
            assert np.array_equal(getattr(dev, f)(), getattr(ref, f)()), (f, n)
    assert law.calls == 0
    # the stress the device law wrote, read through getStress
    assert np.array_equal(law.getStress(), ref.getMaterial(0).getStress())


def test_device_svk_equals_builtin_finite_strain():
    mesh, blocked, v0 = workload("_hexahedron_8", (5, 3, 3), jitter=0.1, seed=72, amp=2e-2)
    mat = dict(rho=1.0, E=1.1, nu=0.3)
    ref, dt = make_gpu(mesh, blocked, v0, "elastic_fd", **mat)
    dev, law = _user_model(mesh, blocked, v0, "device_svk", dt, finite_deformation=True, **mat)
    ref.solveSteps(150)
    dev.solveSteps(150)
    for f in ("getDisplacement", "getVelocity", "getInternalForce"):
        assert np.array_equal(getattr(dev, f)(), getattr(ref, f)()), f
    assert law.calls == 0


@pytest.mark.parametrize("etype,shape", [("_tetrahedron_4", (8, 3, 3)),
                                         ("_hexahedron_8", (6, 3, 2))])
def test_device_law_equals_host_callback(etype, shape):
    mesh, blocked, v0 = workload(etype, shape, jitter=0.1, seed=73, amp=3e-2)
    mat = dict(rho=1.0, E=1.0, nu=0.3)
    ref, dt = make_gpu(mesh, blocked, v0, "elastic", **mat)
    host, hlaw = _user_model(mesh, blocked, v0, "host_cubic", dt, **mat)
    dev, dlaw = _user_model(mesh, blocked, v0, "device_cubic", dt, **mat)
    for n in (1, 120):
        host.solveSteps(n)
        dev.solveSteps(n)
        for f in ("getDisplacement", "getVelocity", "getInternalForce"):
            assert np.array_equal(getattr(dev, f)(), getattr(host, f)()), (f, n)
    assert hlaw.calls == 121 and dlaw.calls == 0
    # the cubic term is felt: the trajectory is not the linear one
    ref.solveSteps(121)
    assert not np.array_equal(ref.getDisplacement(), dev.getDisplacement())


def test_device_law_compile_error_and_switch_back():
    mesh, blocked, v0 = workload("_tetrahedron_4", (6, 3, 3), jitter=0.1, seed=74)
    mat = dict(rho=1.0, E=1.3, nu=0.27)
    ref, dt = make_gpu(mesh, blocked, v0, "elastic", **mat)
    usr, law = _user_model(mesh, blocked, v0, "local_elastic", dt, **mat)
    with pytest.raises(N.AkantuError, match="did not compile"):
        law.setDeviceStress("__device__ void akx_compute_stress(const double *g, double *s, "
                            "const double *p) { s[0] = undefined_name; }", ())
    usr.solveSteps(10)  # still the host law
    assert law.calls == 10
    law.setDeviceStress(ELASTIC_SRC, (law.lame_lambda, law.lame_mu))
    usr.solveSteps(10)
    assert law.calls == 10
    law.setDeviceStress(None)
    usr.solveSteps(10)
    assert law.calls == 20
    ref.solveSteps(30)
    assert np.array_equal(usr.getDisplacement(), ref.getDisplacement())
    with pytest.raises(N.AkantuError, match="user material"):
        ref.getMaterial(0).setDeviceStress(ELASTIC_SRC, (1.0, 1.0))
