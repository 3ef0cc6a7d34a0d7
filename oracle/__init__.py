"""ctypes binding of the CPU FP64 oracle (oracle/akantu_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py``, to compare results against.  The
product package ``akantu2_amd`` never imports it.

``OracleModel`` owns numpy arrays in the reference's array-of-structures
layouts and drives the C restatement of ``SolidMechanicsModel::solveStep``.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

TET4, HEX8 = 0, 1
ELASTIC, ELASTIC_FD, NEOHOOKEAN, PLASTIC_LIH = 0, 1, 2, 3
LAWS = {"elastic": ELASTIC, "elastic_fd": ELASTIC_FD, "neohookean": NEOHOOKEAN,
        "plastic_linear_isotropic_hardening": PLASTIC_LIH}

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_bp = C.POINTER(C.c_ubyte)


class _Model(C.Structure):
    _fields_ = [
        ("type", C.c_int), ("law", C.c_int), ("nn", C.c_long), ("nel", C.c_long),
        ("X", _dp), ("conn", _ip),
        ("rho", C.c_double), ("E", C.c_double), ("nu", C.c_double),
        ("sigma_y", C.c_double), ("h", C.c_double),
        ("E_el", _dp), ("nu_el", _dp), ("sigma_y_el", _dp), ("h_el", _dp),
        ("B", _dp), ("wJ", _dp),
        ("u", _dp), ("v", _dp), ("a", _dp), ("f_int", _dp), ("f_ext", _dp),
        ("mass", _dp), ("u_prev", _dp), ("blocked", _bp),
        ("grad_u", _dp), ("stress", _dp),
        ("prev_grad_u", _dp), ("prev_stress", _dp), ("eps_p", _dp),
        ("prev_eps_p", _dp), ("iso", _dp), ("prev_iso", _dp),
        ("plastic_energy", _dp),
        ("dt", C.c_double), ("step", C.c_long),
        ("eigen_grad_u", _dp), ("alpha", C.c_double), ("delta_T", _dp),
        ("sigma_th", _dp), ("prev_sigma_th", _dp),
    ]


RESIDUAL_CB = C.CFUNCTYPE(None, C.c_void_p, _dp)

_lib = None


def build() -> str:
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE, "liboracle.so"], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        M = C.POINTER(_Model)
        L.orc_precompute.restype = C.c_long
        L.orc_precompute.argtypes = [M]
        for fn in ("orc_lumped_mass", "orc_compute_all_stresses",
                   "orc_assemble_internal_forces", "orc_internal_forces",
                   "orc_solve_step"):
            getattr(L, fn).restype = None
            getattr(L, fn).argtypes = [M]
        L.orc_solve_steps.restype = None
        L.orc_solve_steps.argtypes = [M, C.c_long]
        for fn in ("orc_kinetic_energy", "orc_external_work", "orc_potential_energy",
                   "orc_plastic_energy"):
            getattr(L, fn).restype = C.c_double
            getattr(L, fn).argtypes = [M]
        L.orc_set_threads.restype = None
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_get_threads.restype = C.c_int
        L.orc_get_threads.argtypes = []
        L.orc_apply_traction.restype = None
        L.orc_apply_traction.argtypes = [M, C.c_int, C.c_long, _ip, _dp]
        L.orc_math_reduce.restype = C.c_double
        L.orc_math_reduce.argtypes = [_dp, C.c_long]
        L.orc_stable_dt.restype = C.c_double
        L.orc_stable_dt.argtypes = [M]
        L.orc_element_inradius.restype = C.c_double
        L.orc_element_inradius.argtypes = [C.c_int, _dp]
        L.orc_gauss_points.argtypes = [C.c_int, _dp, _dp]
        L.orc_shapes.argtypes = [C.c_int, _dp, _dp]
        L.orc_dnds.argtypes = [C.c_int, _dp, _dp]
        L.orc_lame.argtypes = [C.c_double, C.c_double, _dp, _dp]
        L.orc_law_stress.restype = None
        L.orc_law_stress.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double,
                                     C.c_double, _dp, _dp, _dp, C.c_double, _dp,
                                     _dp, _dp, _dp]
        L.orc_law_stress_th.restype = None
        L.orc_law_stress_th.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double,
                                        C.c_double, _dp, _dp, _dp, C.c_double, _dp,
                                        C.c_double, C.c_double, _dp, _dp, _dp]
        L.orc_lumped_step_cb.restype = None
        L.orc_lumped_step_cb.argtypes = [C.c_long, C.c_double, _dp, _dp, _dp, _bp,
                                         _dp, RESIDUAL_CB, C.c_void_p]
        _lib = L
    return _lib


def _ptr(a, t=_dp):
    return None if a is None else a.ctypes.data_as(t)


def nnpe_of(etype: int) -> int:
    return 4 if etype == TET4 else 8


def nqp_of(etype: int) -> int:
    return 1 if etype == TET4 else 8


class OracleModel:
    """One mesh, one material, the reference's explicit lumped-mass model."""

    def __init__(self, etype, X, conn, law="elastic", rho=1.0, E=1.0, nu=0.3,
                 sigma_y=0.0, h=0.0, E_el=None, nu_el=None, sigma_y_el=None,
                 h_el=None, store_B=True):
        self.etype = int(etype)
        self.law = LAWS[law] if isinstance(law, str) else int(law)
        self.X = np.ascontiguousarray(X, dtype=np.float64).reshape(-1, 3)
        self.conn = np.ascontiguousarray(conn, dtype=np.int32).reshape(
            -1, nnpe_of(self.etype))
        nn, nel = self.X.shape[0], self.conn.shape[0]
        nqp, nnpe = nqp_of(self.etype), nnpe_of(self.etype)
        self.nn, self.nel, self.nqp = nn, nel, nqp
        z3 = lambda: np.zeros((nn, 3))
        self.u, self.v, self.a = z3(), z3(), z3()
        self.f_int, self.f_ext, self.mass, self.u_prev = z3(), z3(), z3(), z3()
        self.blocked = np.zeros((nn, 3), dtype=np.uint8)
        self.B = np.zeros((nel * nqp, 3 * nnpe)) if store_B else None
        self.wJ = np.zeros(nel * nqp)
        self.grad_u = np.zeros((nel * nqp, 9))
        self.stress = np.zeros((nel * nqp, 9))
        plastic = self.law == PLASTIC_LIH
        self.prev_grad_u = np.zeros((nel * nqp, 9)) if plastic else None
        self.prev_stress = np.zeros((nel * nqp, 9)) if plastic else None
        self.eps_p = np.zeros((nel * nqp, 9)) if plastic else None
        self.prev_eps_p = np.zeros((nel * nqp, 9)) if plastic else None
        self.iso = np.zeros(nel * nqp) if plastic else None
        self.prev_iso = np.zeros(nel * nqp) if plastic else None
        self.plastic_energy = np.zeros(nel * nqp) if plastic else None
        cvt = lambda a: None if a is None else np.ascontiguousarray(a, np.float64)
        self.E_el, self.nu_el = cvt(E_el), cvt(nu_el)
        self.sigma_y_el, self.h_el = cvt(sigma_y_el), cvt(h_el)
        s = _Model()
        s.type, s.law, s.nn, s.nel = self.etype, self.law, nn, nel
        s.X, s.conn = _ptr(self.X), _ptr(self.conn, _ip)
        s.rho, s.E, s.nu, s.sigma_y, s.h = rho, E, nu, sigma_y, h
        s.E_el, s.nu_el = _ptr(self.E_el), _ptr(self.nu_el)
        s.sigma_y_el, s.h_el = _ptr(self.sigma_y_el), _ptr(self.h_el)
        s.B, s.wJ = _ptr(self.B), _ptr(self.wJ)
        s.u, s.v, s.a = _ptr(self.u), _ptr(self.v), _ptr(self.a)
        s.f_int, s.f_ext = _ptr(self.f_int), _ptr(self.f_ext)
        s.mass, s.u_prev = _ptr(self.mass), _ptr(self.u_prev)
        s.blocked = _ptr(self.blocked, _bp)
        s.grad_u, s.stress = _ptr(self.grad_u), _ptr(self.stress)
        s.prev_grad_u, s.prev_stress = _ptr(self.prev_grad_u), _ptr(self.prev_stress)
        s.eps_p, s.prev_eps_p = _ptr(self.eps_p), _ptr(self.prev_eps_p)
        s.iso, s.prev_iso = _ptr(self.iso), _ptr(self.prev_iso)
        s.plastic_energy = _ptr(self.plastic_energy)
        s.dt, s.step = 0.0, 0
        self.eigen_grad_u = self.delta_T = self.sigma_th = self.prev_sigma_th = None
        s.alpha = 0.0
        self._s = s

    def set_eigen_grad_u(self, eig):
        """Material::eigengradu: a 3x3 matrix for every quadrature point
        (the "eigen_grad_u" parameter, material.cc:80-85) or an array
        [nel*nqp][3][3]; None clears it."""
        if eig is None:
            self.eigen_grad_u = None
        else:
            e = np.asarray(eig, np.float64)
            if e.shape == (3, 3):
                e = np.broadcast_to(e, (self.nel * self.nqp, 3, 3))
            # column-major 3x3 per point, as grad_u
            self.eigen_grad_u = np.ascontiguousarray(
                np.transpose(e.reshape(-1, 3, 3), (0, 2, 1)).reshape(-1, 9))
        self._s.eigen_grad_u = _ptr(self.eigen_grad_u)

    def set_thermal(self, alpha, delta_T):
        """MaterialThermal's alpha and delta_T (a value for every quadrature
        point, or an array [nel*nqp])."""
        nq = self.nel * self.nqp
        self.delta_T = np.ascontiguousarray(np.broadcast_to(
            np.asarray(delta_T, np.float64), (nq,)))
        if self.sigma_th is None:
            self.sigma_th, self.prev_sigma_th = np.zeros(nq), np.zeros(nq)
        self._s.alpha = float(alpha)
        self._s.delta_T = _ptr(self.delta_T)
        self._s.sigma_th, self._s.prev_sigma_th = _ptr(self.sigma_th), _ptr(self.prev_sigma_th)

    # initFull(_explicit_lumped_mass): precompute + lumped mass
    def init_full(self):
        bad = lib().orc_precompute(C.byref(self._s))
        if bad:
            raise ValueError(f"negative jacobian at quadrature point {bad - 1}")
        lib().orc_lumped_mass(C.byref(self._s))

    def stable_dt(self) -> float:
        return lib().orc_stable_dt(C.byref(self._s))

    def set_time_step(self, dt: float):
        self._s.dt = float(dt)

    @property
    def dt(self) -> float:
        return self._s.dt

    def internal_forces(self):
        lib().orc_internal_forces(C.byref(self._s))

    def solve_steps(self, n: int = 1):
        lib().orc_solve_steps(C.byref(self._s), int(n))

    @property
    def step(self) -> int:
        return self._s.step

    def apply_traction(self, facet_type, facets, traction):
        """BC::Neumann::FromTraction on facets [n][3] (Tri3, type 2) or
        [n][4] (Quad4, type 3)."""
        f = np.ascontiguousarray(facets, dtype=np.int32)
        t = np.ascontiguousarray(traction, dtype=np.float64)
        lib().orc_apply_traction(C.byref(self._s), int(facet_type), f.shape[0],
                                 _ptr(f, _ip), _ptr(t))

    def energy(self, energy_id: str) -> float:
        """SolidMechanicsModel::getEnergy(energy_id) for "kinetic",
        "external work", "potential" and "plastic" (0 for other ids, as the
        reference's Material::getEnergy)."""
        L, s = lib(), C.byref(self._s)
        if energy_id == "kinetic":
            return L.orc_kinetic_energy(s)
        if energy_id == "external work":
            return L.orc_external_work(s)
        if energy_id == "potential":
            return L.orc_potential_energy(s)
        if energy_id == "plastic":
            return L.orc_plastic_energy(s)
        return 0.0


def set_threads(n: int):
    """Threads of the oracle's element and DOF loops (results unchanged)."""
    lib().orc_set_threads(int(n))


def math_reduce(values) -> float:
    """The reference's Math::reduce (sort, pairwise sums) of a copy."""
    a = np.array(values, dtype=np.float64)
    return lib().orc_math_reduce(_ptr(a), a.size)


def lumped_step_cb(u, v, a, blocked, mass, dt, residual):
    """One lumped central-difference step on flat DOF arrays; ``residual(r)``
    fills the f_ext + f_int part of the residual after the predictor."""
    ndof = u.size

    def _cb(_ctx, rp):
        r = np.ctypeslib.as_array(rp, shape=(ndof,))
        residual(r)

    cb = RESIDUAL_CB(_cb)
    lib().orc_lumped_step_cb(ndof, dt, _ptr(u), _ptr(v), _ptr(a),
                             _ptr(blocked, _bp), _ptr(mass), cb, None)


def gauss_points(etype):
    nqp = nqp_of(etype)
    xi, w = np.zeros((nqp, 3)), np.zeros(nqp)
    lib().orc_gauss_points(etype, _ptr(xi), _ptr(w))
    return xi, w


def shapes(etype, xi):
    N = np.zeros(nnpe_of(etype))
    lib().orc_shapes(etype, _ptr(np.ascontiguousarray(xi, np.float64)), _ptr(N))
    return N


def dnds(etype, xi):
    d = np.zeros((3, nnpe_of(etype)))
    lib().orc_dnds(etype, _ptr(np.ascontiguousarray(xi, np.float64)), _ptr(d))
    return d


def inradius(etype, Xe):
    Xe = np.ascontiguousarray(Xe, np.float64)
    return lib().orc_element_inradius(etype, _ptr(Xe))


def lame(E, nu):
    lam, mu = C.c_double(), C.c_double()
    lib().orc_lame(E, nu, C.byref(lam), C.byref(mu))
    return lam.value, mu.value


def law_stress(law, lam, mu, grad_u, sigma_y=0.0, h=0.0, prev_grad_u=None,
               prev_sigma=None, prev_iso=0.0, prev_eps_p=None, sigma_th=None,
               delta_sigma_th=0.0):
    """Per-quadrature-point constitutive law; tensors are 3x3 numpy arrays.
    sigma_th / delta_sigma_th: the thermal terms (None: the default -0)."""
    law = LAWS[law] if isinstance(law, str) else int(law)
    col = lambda m: np.ascontiguousarray(np.asarray(m, np.float64).T).reshape(9)
    g = col(grad_u)
    pg = col(prev_grad_u if prev_grad_u is not None else np.zeros((3, 3)))
    ps = col(prev_sigma if prev_sigma is not None else np.zeros((3, 3)))
    pe = col(prev_eps_p if prev_eps_p is not None else np.zeros((3, 3)))
    sig, eps = np.zeros(9), np.zeros(9)
    iso = C.c_double(0.0)
    if sigma_th is None:
        lib().orc_law_stress(law, lam, mu, sigma_y, h, _ptr(g), _ptr(pg), _ptr(ps),
                             prev_iso, _ptr(pe), _ptr(sig), C.cast(C.byref(iso), _dp),
                             _ptr(eps))
    else:
        lib().orc_law_stress_th(law, lam, mu, sigma_y, h, _ptr(g), _ptr(pg), _ptr(ps),
                                prev_iso, _ptr(pe), float(sigma_th), float(delta_sigma_th),
                                _ptr(sig), C.cast(C.byref(iso), _dp), _ptr(eps))
    unc = lambda v: v.reshape(3, 3).T.copy()
    return unc(sig), iso.value, unc(eps)
