"""Python front end of the MI355X explicit path, shaped like the reference's
Python API (python/py_solid_mechanics_model.cc:57-173, py_model.cc:131-139,
py_material.cc:45-80): ``SolidMechanicsModel(mesh)``, ``initFull``,
``getStableTimeStep``, ``setTimeStep``, ``solveStep``, ``getDisplacement`` ...

Nodal arrays are numpy arrays in the reference's [nb_nodes][3] layout.  A
getter returns a view of a host copy of the device-resident field, and the
copy behaves like the reference's live ``Array`` (solid_mechanics_model.hh
:440-468): while any array referring to it is alive, every call that changes
the device field refreshes the copy in place, and edits made through it are
uploaded before the next call that reads the field (the reference lets users
write through ``getDisplacement()``).  A copy nothing refers to any more is
dropped, so a ``solveStep`` with no array held costs O(1) on the host; every
array held across steps costs one device-to-host copy per step.

User constitutive laws subclass :class:`Material` and register with
:class:`MaterialFactory`, as in the reference; ``getFEEngine()`` and
``getDOFManager()`` give the FE-engine gradient / BtD / integrate and the
elemental assembly on caller fields.
"""
from __future__ import annotations

import ctypes as C
import sys
import weakref

import numpy as np

from . import _native as N
from .mesh import BoxMesh

_explicit_lumped_mass = "explicit_lumped_mass"
_not_ghost, _ghost = "_not_ghost", "_ghost"
# parameter flags of Material.registerParamReal (accepted, not interpreted)
_pat_parsable, _pat_readable, _pat_writable, _pat_modifiable = 1, 2, 4, 6

_FIELDS = {
    "displacement": N.FIELD_DISPLACEMENT,
    "velocity": N.FIELD_VELOCITY,
    "acceleration": N.FIELD_ACCELERATION,
    "internal_force": N.FIELD_INTERNAL_FORCE,
    "external_force": N.FIELD_EXTERNAL_FORCE,
    "mass": N.FIELD_MASS,
    "blocked_dofs": N.FIELD_BLOCKED_DOFS,
}
_WRITABLE = {"displacement", "velocity", "acceleration", "external_force",
             "blocked_dofs"}
_LAWS = {"elastic": N.LAW_ELASTIC, "elastic_fd": N.LAW_ELASTIC_FD,
         "neohookean": N.LAW_NEOHOOKEAN,
         "plastic_linear_isotropic_hardening": N.LAW_PLASTIC_LIH}


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


_libc = C.CDLL(None)
_libc.memcmp.restype = C.c_int
_libc.memcmp.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t]
_SOLVER_IDS = ("", "explicit_lumped")


def _check_solver_id(solver_id):
    """ModelSolver::getSolver (model_solver.cc:201-209): "" is the default
    solver, which initFull(_explicit_lumped_mass) names "explicit_lumped"
    (solid_mechanics_model.cc:165-169)."""
    if not isinstance(solver_id, str):
        raise TypeError(f"solver_id must be a str, not {type(solver_id).__name__} "
                        "(solveSteps(n) runs n steps)")
    if solver_id not in _SOLVER_IDS:
        raise RuntimeError(f"The non linear solver dof_manager:tss:{solver_id} does not "
                           "exists in dof_manager")


class _Mirror:
    """Host copy of one device field.  Every array a getter hands out is a
    view of ``base`` (numpy points a view's ``.base`` at the owner of the
    memory, also for views of views), so ``base`` has references beyond this
    object exactly while some handed-out array is alive."""

    __slots__ = ("base", "snap", "view", "writable")

    def __init__(self, base, writable):
        self.base = base
        self.writable = writable
        self.snap = base.copy() if writable else None
        self.view = None  # weak reference to the last view handed out

    def held(self):
        # this object's reference + getrefcount's argument
        return sys.getrefcount(self.base) > 2

    def edited(self):
        return self.writable and _libc.memcmp(self.base.ctypes.data, self.snap.ctypes.data,
                                              self.base.nbytes) != 0


class MaterialFactory:
    """The reference's MaterialFactory (material.hh:610-621; from Python
    py_material.cc:237-240): ``registerAllocator(name, allocator)`` makes
    ``model.addMaterial(name, ...)`` create ``allocator(dim, unused, model,
    id)``, a subclass of :class:`Material` whose ``computeStress`` the
    model calls at every stress evaluation."""

    _instance = None

    def __init__(self):
        self._allocators = {}

    @staticmethod
    def getInstance():
        if MaterialFactory._instance is None:
            MaterialFactory._instance = MaterialFactory()
        return MaterialFactory._instance

    def registerAllocator(self, name, allocator):
        if name in _LAWS:
            raise ValueError(f"{name!r} is a built-in law")
        self._allocators[name] = allocator

    def allocator(self, name):
        return self._allocators.get(name)


class Material:
    """A material of the model (Material's accessors).

    Subclass it to write a constitutive law, as in the reference
    (py_material.cc:45-80): override ``computeStress(el_type, ghost_type)``
    and fill ``self.getStress(el_type, ghost_type)`` ([nb_qp][9], sigma, or
    the second Piola-Kirchhoff stress for finite deformation) from
    ``self.getGradU(el_type, ghost_type)`` ([nb_qp][9], column-major
    du_i/dX_j at i + 3 j).  Register the subclass with
    ``MaterialFactory.getInstance().registerAllocator``.  The model runs such
    a law on the plain FP64 path; its push-wave speed (stable time step) is
    the elastic one of its rho, E, nu."""

    def __init__(self, model, mid, law=None, params=None, nb_elements=None):
        self._model, self.id, self.law = model, mid, law
        self.params = dict(params or {})
        self.nb_elements = nb_elements
        self._live = None  # (grad_u, stress) views while computeStress runs
        self._param_doc = {}

    def getParam(self, name):
        return self.params[name]

    get = getParam

    # parameters of a user law (Material::registerParam / getReal)
    def registerParamReal(self, name, flags=0, description=""):
        self._param_doc[name] = description

    def getReal(self, name):
        if name == "alpha":
            return self._alpha
        return float(self.params[name])

    # MaterialThermal's parameters (material_thermal.cc:36-43): "alpha" and
    # "delta_T", the uniform temperature field; the plain FP64 path applies
    # sigma_th = -E / (1 - 2 nu) alpha delta_T (akx_model_set_thermal)
    _alpha = 0.0
    _delta_T = None

    def setReal(self, name, value):
        """ParameterRegistry::set (py_parser.cc:68-69) for the parameters a
        running model can change here: "alpha" and "delta_T"."""
        if name == "alpha":
            self._alpha = float(value)
        elif name == "delta_T":
            self._delta_T = np.full(self._nq(), float(value))
        else:
            raise NotImplementedError(f"parameter {name!r} cannot be changed on this path")
        self._push_thermal()

    def setDeltaT(self, delta_T):
        """The "delta_T" internal field, one value per Gauss point
        ([nb_elements * nqp], the material's element order)."""
        d = np.ascontiguousarray(delta_T, np.float64).reshape(-1)
        if d.size != self._nq():
            raise ValueError("delta_T needs one value per Gauss point")
        self._delta_T = d
        self._push_thermal()

    def _nq(self):
        return self.nb_elements * self._model.nqp

    def _push_thermal(self):
        if self._delta_T is None:
            return
        N.check(N.lib().akx_model_set_thermal(self._model._h, self.id, self._alpha,
                                              _p(self._delta_T)))
        self._model._stresses_fresh = False

    def applyEigenGradU(self, prescribed_eigen_grad_u, ghost_type=_not_ghost):
        """Material::applyEigenGradU (material.cc:811-824): the same 3x3
        eigen grad u at every Gauss point of the material."""
        g = np.asarray(prescribed_eigen_grad_u, np.float64).reshape(3, 3)
        self.setEigenGradU(np.broadcast_to(g, (self._nq(), 3, 3)))

    def setEigenGradU(self, eigen_grad_u):
        """The "eigen_grad_u" internal field: [nb_elements * nqp][3][3] (or
        [..][9] column-major, du_i/dX_j at i + 3 j); None clears it."""
        if eigen_grad_u is None:
            N.check(N.lib().akx_model_set_eigen_grad_u(self._model._h, self.id, None))
        else:
            e = np.asarray(eigen_grad_u, np.float64)
            if e.ndim == 3:  # matrices -> column-major rows
                e = np.transpose(e, (0, 2, 1)).reshape(-1, 9)
            e = np.ascontiguousarray(e.reshape(-1, 9))
            if e.shape[0] != self._nq():
                raise ValueError("eigen_grad_u needs one matrix per Gauss point")
            N.check(N.lib().akx_model_set_eigen_grad_u(self._model._h, self.id, _p(e)))
        self._model._stresses_fresh = False

    def initMaterial(self):
        """Called once at initFull (Material::initMaterial)."""

    # A law run on the device (akx_model_set_device_stress; the reference's
    # plugin runs on the host).  A subclass that sets device_stress_source --
    # HIP C++ defining
    #   __device__ void akx_compute_stress(const double *grad_u,
    #                                      double *stress, const double *params)
    # for one Gauss point (3x3 column-major, as getGradU / getStress) -- has
    # it compiled at initFull, after initMaterial, with the values
    # deviceStressParams() returns; its computeStress is then not called.
    device_stress_source = None

    def deviceStressParams(self):
        return ()

    def setDeviceStress(self, source, params=()):
        """Compile and use a device law for this material (None: back to
        computeStress)."""
        p = np.ascontiguousarray(params, np.float64).reshape(-1)
        N.check(N.lib().akx_model_set_device_stress(
            self._model._h, self.id, None if source is None else source.encode(),
            _p(p) if p.size else None, int(p.size)))
        self._model._stresses_fresh = False

    def computeStress(self, el_type, ghost_type=_not_ghost):
        raise NotImplementedError("computeStress of a user material")

    def _internal(self, which, comps=9):
        m = self._model
        m._push()
        if self._needs_compute():
            m._run(N.lib().akx_model_compute_all_stresses)
            m._stresses_fresh = True
        out = np.empty((self.nb_elements * m.nqp, comps))
        N.check(N.lib().akx_model_get_internal(m._h, self.id, which, _p(out)))
        return out

    def _needs_compute(self):
        m = self._model
        if self.law == N.LAW_PLASTIC_LIH:
            return False  # the history buffers hold the current state
        return not m._stresses_fresh

    def getGradU(self, el_type=None, ghost_type=_not_ghost):
        if self._live is not None:
            return self._live[0]
        return self._internal(N.INTERNAL_GRAD_U)

    def getStress(self, el_type=None, ghost_type=_not_ghost):
        """sigma for small strain laws, the second Piola-Kirchhoff stress for
        finite deformation (as Material::getStress / piola_kirchhoff_2).
        Inside computeStress: the array to fill."""
        if self._live is not None:
            return self._live[1]
        return self._internal(N.INTERNAL_STRESS)

    def getInelasticStrain(self):
        return self._internal(N.INTERNAL_INELASTIC_STRAIN)

    def getIsoHardening(self):
        return self._internal(N.INTERNAL_ISO_HARDENING, 1)[:, 0]

    def getEnergy(self, energy_id: str) -> float:
        """Material::getEnergy: "potential" or "plastic" (0 otherwise)."""
        return self._model._energy(energy_id, self.id)


class FEEngine:
    """The model's FEEngine calls on caller fields (fe_engine.hh:85-158),
    computed on the device in the reference's operation order.  Per-point
    arrays have the points of the i-th filtered element at rows i * nqp ..,
    matrices per point column-major; ``filter`` lists elements (None: all)."""

    def __init__(self, model):
        self._m = model

    def _filter(self, filter_elements):
        if filter_elements is None:
            return None, self._m.mesh.nb_elements
        f = np.ascontiguousarray(filter_elements, np.int32)
        return f, f.size

    def gradientOnIntegrationPoints(self, u, nb_degree_of_freedom=None, el_type=None,
                                    ghost_type=_not_ghost, filter_elements=None):
        """[nb_nodes][d] -> [n * nqp][d * 3], du_k/dX_j at k + d j."""
        u = np.ascontiguousarray(u, np.float64).reshape(self._m.nb_nodes, -1)
        d = u.shape[1] if nb_degree_of_freedom is None else int(nb_degree_of_freedom)
        f, n = self._filter(filter_elements)
        out = np.empty((n * self._m.nqp, 3 * d))
        N.check(N.lib().akx_fe_gradient_on_integration_points(self._m._h, _p(u), d, _p(f),
                                                              n, _p(out)))
        return out

    def computeBtD(self, Ds, el_type=None, ghost_type=_not_ghost, filter_elements=None):
        """[n * nqp][3 d] (d x 3 per point) -> [n * nqp][d * nnpe], D B."""
        f, n = self._filter(filter_elements)
        Ds = np.ascontiguousarray(Ds, np.float64).reshape(n * self._m.nqp, -1)
        nnpe = self._m.mesh.conn.shape[1]
        out = np.empty((Ds.shape[0], Ds.shape[1] // 3 * nnpe))
        N.check(N.lib().akx_fe_compute_btd(self._m._h, _p(Ds), Ds.shape[1], _p(f), n,
                                           _p(out)))
        return out

    def integrate(self, f_q, nb_degree_of_freedom=None, el_type=None, ghost_type=_not_ghost,
                  filter_elements=None):
        """[n * nqp][d] -> [n][d], sum over the Gauss points of f w detJ."""
        f, n = self._filter(filter_elements)
        f_q = np.ascontiguousarray(f_q, np.float64).reshape(n * self._m.nqp, -1)
        d = f_q.shape[1] if nb_degree_of_freedom is None else int(nb_degree_of_freedom)
        out = np.empty((n, d))
        N.check(N.lib().akx_fe_integrate(self._m._h, _p(f_q), d, _p(f), n, _p(out)))
        return out


class DOFManager:
    """DOFManager::assembleElementalArrayLocalArray (dof_manager.cc:65-120)
    on the device: ``out`` [nb_nodes][d] += scale * elementary [n][d * nnpe],
    each node's contributions in filter order."""

    def __init__(self, model):
        self._m = model

    def assembleElementalArrayLocalArray(self, elementary, out, el_type=None,
                                         ghost_type=_not_ghost, scale_factor=1.0,
                                         filter_elements=None):
        if not (isinstance(out, np.ndarray) and out.dtype == np.float64
                and out.flags.c_contiguous):
            raise TypeError("out must be a C-contiguous float64 array (updated in place)")
        f = None if filter_elements is None else np.ascontiguousarray(filter_elements,
                                                                      np.int32)
        n = self._m.mesh.nb_elements if f is None else f.size
        elementary = np.ascontiguousarray(elementary, np.float64).reshape(n, -1)
        d = out.size // self._m.nb_nodes
        N.check(N.lib().akx_dof_assemble_elemental_array(
            self._m._h, _p(elementary), d, _p(f), n, float(scale_factor), _p(out)))
        return out


class SolidMechanicsModel:
    """One mesh of one element type on one gfx950 device."""

    def __init__(self, mesh: BoxMesh, device: int = 0):
        if N.device_count() == 0:
            raise RuntimeError("no gfx950 device visible: the MI355X path has no "
                               "CPU fallback")
        h = C.c_void_p()
        N.check(N.lib().akx_model_create(device, C.byref(h)))
        self._h = h
        self.mesh = mesh
        self.nqp = mesh.nqp
        nodes = np.ascontiguousarray(mesh.nodes, np.float64)
        conn = np.ascontiguousarray(mesh.conn, np.int32)
        N.check(N.lib().akx_model_set_mesh(h, mesh.etype, nodes.shape[0], _p(nodes),
                                           conn.shape[0], _p(conn)))
        self.nb_nodes = nodes.shape[0]
        self.materials = []
        self._mirrors = {}
        self._initialized = False
        self._stresses_fresh = False
        self.time_step = 0.0
        self._cb = None
        self._cb_error = None

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            try:
                N.lib().akx_model_destroy(h)
            except Exception:
                pass
            self._h = None

    # ------------------------------------------------------------ materials
    def addMaterial(self, law="elastic", rho=1.0, E=1.0, nu=0.3, sigma_y=0.0, h=0.0,
                    finite_deformation=False, elements=None, E_el=None, nu_el=None,
                    sigma_y_el=None, h_el=None, **user_params):
        """Register a material (the ``material <law> [...]`` section of the
        reference's input file).  ``elements`` is the element filter in
        assembly order (default: all elements).  A name registered with
        ``MaterialFactory`` creates that user material (``user_params`` are
        its further parameters, read with ``getReal``)."""
        alloc = MaterialFactory.getInstance().allocator(law) if isinstance(law, str) else None
        if alloc is not None:
            return self._add_user_material(alloc, rho, E, nu, finite_deformation,
                                           elements, user_params)
        if user_params:
            raise TypeError(f"unknown parameters {sorted(user_params)} for {law!r}")
        if law == "elastic_fd":
            law, finite_deformation = "elastic", True
        if isinstance(law, str):
            # the factory lookup of the C ABI: unknown names and the finite-
            # deformation plastic branch are refused there
            lid_c = C.c_int(-1)
            N.check(N.lib().akx_material_law_by_name(law.encode(), int(bool(finite_deformation)),
                                                     C.byref(lid_c)))
            lid = lid_c.value
        else:
            lid = int(law)
        params = np.array([rho, E, nu, sigma_y, h], np.float64)
        mid = C.c_int(-1)
        filt = None if elements is None else np.ascontiguousarray(elements, np.int32)
        n = self.mesh.nb_elements if filt is None else filt.size
        N.check(N.lib().akx_model_add_material(
            self._h, lid, params.ctypes.data_as(C.POINTER(C.c_double)), 5,
            _p(filt), 0 if filt is None else filt.size, C.byref(mid)))
        if any(x is not None for x in (E_el, nu_el, sigma_y_el, h_el)):
            cv = lambda a: None if a is None else np.ascontiguousarray(a, np.float64)
            arrs = [cv(E_el), cv(nu_el), cv(sigma_y_el), cv(h_el)]
            for a in arrs:
                if a is not None and a.size != n:
                    raise ValueError("per-element parameters must match the filter")
            N.check(N.lib().akx_model_set_element_parameters(
                self._h, mid.value, *[_p(a) for a in arrs]))
        mat = Material(self, mid.value, lid,
                       dict(rho=rho, E=E, nu=nu, sigma_y=sigma_y, h=h), n)
        mat.element_ids = None if filt is None else filt.copy()
        self.materials.append(mat)
        return mat

    def _add_user_material(self, alloc, rho, E, nu, finite_deformation, elements, extra):
        lid = N.LAW_USER_FD if finite_deformation else N.LAW_USER
        params = np.array([rho, E, nu], np.float64)
        mid = C.c_int(-1)
        filt = None if elements is None else np.ascontiguousarray(elements, np.int32)
        n = self.mesh.nb_elements if filt is None else filt.size
        N.check(N.lib().akx_model_add_material(
            self._h, lid, params.ctypes.data_as(C.POINTER(C.c_double)), 3,
            _p(filt), 0 if filt is None else filt.size, C.byref(mid)))
        mat = alloc(3, None, self, mid.value)
        if not isinstance(mat, Material):
            raise TypeError("a material allocator must return an akantu2_amd.Material")
        mat._model, mat.id, mat.law, mat.nb_elements = self, mid.value, lid, n
        mat.params.update(dict(rho=rho, E=E, nu=nu, finite_deformation=finite_deformation),
                          **extra)
        mat.element_ids = None if filt is None else filt.copy()
        self.materials.append(mat)
        self._install_stress_callback()
        return mat

    def _install_stress_callback(self):
        if self._cb is not None:
            return
        ref = weakref.ref(self)
        el_type = "_tetrahedron_4" if self.mesh.etype == 0 else "_hexahedron_8"

        def compute_stress(user, mid, nq, grad_p, stress_p):
            model = ref()
            if model is None:
                return 1
            mat = model.materials[mid]
            try:
                grad = np.ctypeslib.as_array(grad_p, shape=(nq, 9))
                stress = np.ctypeslib.as_array(stress_p, shape=(nq, 9))
                mat._live = (grad, stress)
                mat.computeStress(el_type, _not_ghost)
                return 0
            except BaseException as e:  # re-raised by the call that ran the law
                model._cb_error = e
                return 1
            finally:
                mat._live = None

        self._cb = N.COMPUTE_STRESS_FN(compute_stress)
        N.check(N.lib().akx_model_set_stress_callback(self._h, self._cb, None))

    def _run(self, fn, *args):
        """A library call that may run user laws: an exception raised in a
        computeStress propagates from here."""
        self._cb_error = None
        status = fn(self._h, *args)
        if self._cb_error is not None:
            e, self._cb_error = self._cb_error, None
            raise e
        N.check(status)

    def getMaterial(self, i):
        return self.materials[i]

    def getFEEngine(self, name=None):
        return FEEngine(self)

    def getDOFManager(self):
        return DOFManager(self)

    def setPrecision(self, precision):
        p = {"fp64": N.PRECISION_FP64, "fp32": N.PRECISION_FP32,
             "fp64_fast": N.PRECISION_FP64_FAST}[precision]
        N.check(N.lib().akx_model_set_precision(self._h, p))

    # ---------------------------------------------------------------- init
    def initFull(self, _analysis_method=_explicit_lumped_mass, **material):
        if _analysis_method != _explicit_lumped_mass:
            raise NotImplementedError("only _explicit_lumped_mass is provided")
        if not self.materials:
            self.addMaterial(**(material or {}))
        N.check(N.lib().akx_model_init_full(self._h))
        for mat in self.materials:
            if mat.law in (N.LAW_USER, N.LAW_USER_FD):
                mat.initMaterial()
                if mat.device_stress_source is not None:
                    mat.setDeviceStress(mat.device_stress_source, mat.deviceStressParams())
        self._initialized = True

    def getStableTimeStep(self) -> float:
        self._push()
        dt = C.c_double()
        N.check(N.lib().akx_model_get_stable_time_step(self._h, C.byref(dt)))
        return dt.value

    def setTimeStep(self, dt: float, solver_id: str = ""):
        """ModelSolver::setTimeStep(time_step, solver_id) (model_solver.cc:323)."""
        _check_solver_id(solver_id)
        N.check(N.lib().akx_model_set_time_step(self._h, float(dt)))
        self.time_step = float(dt)

    def getTimeStep(self, solver_id: str = "") -> float:
        _check_solver_id(solver_id)
        return self.time_step

    # ---------------------------------------------------------------- steps
    def solveStep(self, solver_id: str = ""):
        """ModelSolver::solveStep(solver_id="") (model_solver.cc:267-280,
        py_model.cc:130-135): one explicit step of the default solver, which
        ``"explicit_lumped"`` also names."""
        _check_solver_id(solver_id)
        self._push()
        self._run(N.lib().akx_model_solve_step, solver_id.encode())
        self._refresh()

    def solveSteps(self, nb_steps: int):
        """``nb_steps`` consecutive ``solveStep()`` calls in one library call
        (no counterpart in the reference): the same results bit for bit, with
        the next step's predictor fused into each step's nodal gather."""
        self._push()
        self._run(N.lib().akx_model_solve_steps, int(nb_steps))
        self._refresh()

    def assembleInternalForces(self):
        self._push()
        self._run(N.lib().akx_model_assemble_internal_forces)
        self._refresh()

    def synchronize(self):
        N.check(N.lib().akx_model_synchronize(self._h))

    # --------------------------------------------------------------- fields
    def _push(self):
        """Upload the edits made through the arrays handed out, and forget
        the copies nothing refers to any more.  A copy is current with the
        device whenever an edit is made through it (``_refresh`` runs after
        every call that changes a field), so uploading it whole is exact."""
        for name in list(self._mirrors):
            mir = self._mirrors[name]
            if mir.edited():
                N.check(N.lib().akx_model_set_field(self._h, _FIELDS[name], _p(mir.base)))
                np.copyto(mir.snap, mir.base)
                self._stresses_fresh = False
            if not mir.held():
                del self._mirrors[name]

    def _refresh(self, names=None):
        """After a call that changed device fields: bring the held copies up
        to date in place (all of them, or those in ``names``)."""
        self._stresses_fresh = False
        for name in list(self._mirrors):
            if names is not None and name not in names:
                continue
            mir = self._mirrors[name]
            if not mir.held():
                del self._mirrors[name]
                continue
            N.check(N.lib().akx_model_get_field(self._h, _FIELDS[name], _p(mir.base)))
            if mir.writable:
                np.copyto(mir.snap, mir.base)

    def _get(self, name):
        mir = self._mirrors.get(name)
        if mir is None:
            dtype = np.uint8 if name == "blocked_dofs" else np.float64
            base = np.empty((self.nb_nodes, 3), dtype)
            N.check(N.lib().akx_model_get_field(self._h, _FIELDS[name], _p(base)))
            mir = _Mirror(base, name in _WRITABLE)
            self._mirrors[name] = mir
        view = mir.view() if mir.view is not None else None
        if view is None:
            view = mir.base.view()
            if not mir.writable:
                view.flags.writeable = False
            mir.view = weakref.ref(view)
        return view

    def getDisplacement(self):
        return self._get("displacement")

    def getVelocity(self):
        return self._get("velocity")

    def getAcceleration(self):
        return self._get("acceleration")

    def getInternalForce(self):
        return self._get("internal_force")

    def getExternalForce(self):
        return self._get("external_force")

    def getMass(self):
        return self._get("mass")

    def getBlockedDOFs(self):
        return self._get("blocked_dofs")

    # ---------------------------------------------------------------- dumps
    def setBaseName(self, name, directory="paraview"):
        """Model::setBaseName: ParaView output under ``directory``."""
        from .dumper import DumperParaview
        fields = self._dumper.fields if getattr(self, "_dumper", None) else []
        self._dumper = DumperParaview(self, name, directory)
        for f in fields:
            self._dumper.addDumpField(f)

    def addDumpField(self, name):
        if getattr(self, "_dumper", None) is None:
            self.setBaseName("solid_mechanics_model")
        self._dumper.addDumpField(name)

    def dump(self, time=None, step=None):
        """Model::dump: write the registered fields; returns the file."""
        if getattr(self, "_dumper", None) is None:
            self.setBaseName("solid_mechanics_model")
        return self._dumper.dump(time, step)

    # -------------------------------------------------- boundary conditions
    def applyBC(self, bc, group):
        """BoundaryCondition::applyBC(functor, group): ``group`` is a group
        name of the mesh or a group dict {"nodes", "facet_type", "facets"}."""
        from . import bc as BC
        if isinstance(group, str):
            group = self.mesh.meta["groups"][group]
        if isinstance(bc, BC.Neumann.FromTraction):
            f = np.ascontiguousarray(group["facets"], np.int32)
            t = np.ascontiguousarray(bc.traction, np.float64)
            self._push()
            N.check(N.lib().akx_model_apply_traction(
                self._h, int(group["facet_type"]), f.shape[0], _p(f),
                t.ctypes.data_as(C.POINTER(C.c_double))))
            self._refresh({"external_force"})
            return
        nodes = np.asarray(group["nodes"])
        bc.apply(self.getBlockedDOFs(), self.getDisplacement(), nodes)

    # ------------------------------------------------------------- energies
    def _energy(self, energy_id, material_id):
        self._push()
        out = C.c_double()
        self._run(N.lib().akx_model_get_energy, energy_id.encode(), int(material_id),
                  C.byref(out))
        if energy_id == "potential":
            self._stresses_fresh = True
        return out.value

    def getEnergy(self, energy_id: str) -> float:
        """SolidMechanicsModel::getEnergy: "kinetic", "external work",
        "potential", "plastic" (0 for other ids)."""
        return self._energy(energy_id, -1)

    def setPipeline(self, parts: int) -> bool:
        """Pipelined launch schedule (see akx_model_set_pipeline); returns
        whether it applies to this model."""
        active = C.c_int(0)
        N.check(N.lib().akx_model_set_pipeline(self._h, int(parts), C.byref(active)))
        return bool(active.value)

    _ASSEMBLIES = {"per_corner": 0, "default": 1, "block_partials": 2}

    def setBlockAssembly(self, enable):
        """Element-force assembly (see akx_model_set_block_assembly), before
        initFull: True / "default" (Tet4 FP64: prefix partials, same bits;
        FP32 Tet4 and Hex8: block partials), False / "per_corner", or
        "block_partials" (also for Tet4 in FP64, other rounding)."""
        mode = self._ASSEMBLIES[enable] if isinstance(enable, str) else int(bool(enable))
        N.check(N.lib().akx_model_set_block_assembly(self._h, mode))

    def assembly(self) -> str:
        """The assembly the initialised model uses: "per_corner",
        "prefix_partials" or "block_partials"."""
        active = C.c_int(0)
        N.check(N.lib().akx_model_get_block_assembly(self._h, C.byref(active)))
        return ("per_corner", "prefix_partials", "block_partials")[active.value]

    def blockAssembly(self) -> bool:
        """Whether the initialised model sums element forces in blocks
        (prefix or block partials)."""
        return self.assembly() != "per_corner"

    # ----------------------------------------------------------- measurement
    def enableKernelTiming(self, enable=True):
        N.check(N.lib().akx_model_enable_kernel_timing(self._h, int(enable)))

    def kernelTiming(self, kernel_class):
        ms, n = C.c_double(), C.c_int64()
        N.check(N.lib().akx_model_get_kernel_timing(self._h, kernel_class,
                                                    C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def kernelBytes(self, kernel_class):
        b = C.c_double()
        N.check(N.lib().akx_model_get_kernel_bytes(self._h, kernel_class, C.byref(b)))
        return b.value
