// akantu_mi355x.hpp -- C++17 host side above the C ABI (akantu_mi355x.h),
// shaped like the reference's SolidMechanicsModel / Material API so that
// driver code written against Akantu reads the same:
//
//   akantu_mi355x::SolidMechanicsModel model(mesh);
//   model.addMaterial("neohookean", {1.0, 1.0, 0.3});
//   model.initFull(akantu_mi355x::_explicit_lumped_mass);
//   model.setTimeStep(0.5 * model.getStableTimeStep());
//   for (int s = 0; s < 1000; ++s) model.solveStep();
//   const auto & u = model.getDisplacement();   // [nb_nodes][3]
//
// The nodal Arrays the getters return are host copies of the device fields
// in the reference's layout (Array<Real>: one tuple of nb_component values per
// row) that behave like the reference's live Arrays (solid_mechanics_model.hh
// :440-468) when held across steps: a call that changes the device fields only
// marks the copies stale (O(1)); the next element access of a stale copy
// downloads it, and a non-const access marks it edited, so that the next
// call that reads the field uploads it (the reference lets users write through
// getDisplacement()).  A raw pointer from data() sees the values of that
// data() call: call data() again after a step.  Errors throw
// akantu_mi355x::Exception (the reference throws debug::Exception; a negative
// jacobian throws JacobianException like IntegratorGaussException).
#pragma once

#include "akantu_mi355x.h"

#include <array>
#include <cstdint>
#include <deque>
#include <exception>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

namespace akantu_mi355x {

struct Exception : std::runtime_error {
  int code;
  Exception(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};
struct JacobianException : Exception {
  using Exception::Exception;
};

inline void check(int status) {
  if (status == AKX_OK)
    return;
  std::string msg = akx_last_error();
  if (status == AKX_ERR_JACOBIAN)
    throw JacobianException(status, msg);
  throw Exception(status, msg);
}

enum AnalysisMethod { _explicit_lumped_mass };
enum ElementType { _tetrahedron_4 = AKX_TET4, _hexahedron_8 = AKX_HEX8 };

// Keeps a model's host copies of its device fields current: bit f of
// `stale` says copy f is older than the device field, bit f of `dirty` that
// it was accessed for writing since it was last uploaded.
class FieldSync {
public:
  std::uint32_t stale = 0, dirty = 0;
  virtual void pull(int field) = 0; // download copy `field`, clear its stale bit

protected:
  ~FieldSync() = default;
};

// Minimal Array<T>: contiguous rows of nb_component values (aka_array.hh).
// An Array a model hands out is bound to the model's FieldSync; a copy of it
// is a plain Array of the values it holds at the copy.
template <class T> class Array {
public:
  Array() = default;
  Array(std::int64_t size, int nb_component, T value = T())
      : nb_comp(nb_component), data_(size * nb_component, value) {}
  Array(const Array &o) : nb_comp(o.nb_comp), data_((o.touch(), o.data_)) {}
  Array(Array &&o) : nb_comp(o.nb_comp), data_((o.touch(), std::move(o.data_))) {}
  Array &operator=(const Array &o) {
    if (this != &o) {
      o.touch();
      touch_write();
      nb_comp = o.nb_comp;
      data_ = o.data_;
    }
    return *this;
  }
  std::int64_t size() const { return (std::int64_t)data_.size() / nb_comp; }
  int getNbComponent() const { return nb_comp; }
  T &operator()(std::int64_t i, int c = 0) {
    touch_write();
    return data_[i * nb_comp + c];
  }
  const T &operator()(std::int64_t i, int c = 0) const {
    touch();
    return data_[i * nb_comp + c];
  }
  T *data() {
    touch_write();
    return data_.data();
  }
  const T *data() const {
    touch();
    return data_.data();
  }

private:
  friend class SolidMechanicsModel;
  void touch() const {
    if (sync && (sync->stale >> field & 1u))
      sync->pull(field);
  }
  void touch_write() {
    if (sync) {
      touch();
      sync->dirty |= 1u << field;
    }
  }
  T *raw() { return data_.data(); } // no synchronisation (the owner's own I/O)
  void bind(FieldSync *s, int f) {
    sync = s;
    field = f;
  }

  int nb_comp = 1;
  std::vector<T> data_;
  FieldSync *sync = nullptr;
  int field = -1;
};

// Mesh given as arrays (Mesh::getNodes / Mesh::getConnectivity).
struct Mesh {
  ElementType type;
  Array<double> nodes;          // [nb_nodes][3]
  Array<std::int32_t> connectivity; // [nb_elements][nnpe]
};

struct MaterialParameters {
  double rho, E, nu, sigma_y = 0., h = 0.;
};

class SolidMechanicsModel;

// A constitutive law written by the caller: the reference's Material
// subclass overriding computeStress (material.hh:78-81).  grad_u and stress
// are the material's Gauss-point arrays [nb_qp][9] (column-major, du_i/dX_j
// at i + 3 j), as Material::getGradU / getStress; stress is sigma, or the
// second Piola-Kirchhoff stress for a finite-deformation law.  It runs on
// the host at every stress evaluation, beside the device passes.
class ConstitutiveLaw {
public:
  virtual ~ConstitutiveLaw() = default;
  virtual void computeStress(const Array<double> &grad_u, Array<double> &stress) = 0;
};

// Handle on one material (Material's accessors).
class Material {
public:
  Material(SolidMechanicsModel &m, int id, std::int64_t n) : model(&m), id(id), n(n) {}
  int getID() const { return id; }
  // Material::getGradU / getStress (or PK2 for finite deformation)
  Array<double> getGradU() const;
  Array<double> getStress() const;
  // Material::applyEigenGradU (material.cc:811-824): the same eigen grad u
  // (column-major 3x3, du_i/dX_j at i + 3 j) at every Gauss point, or one
  // per Gauss point ([n * nqp][9]); subtracted from grad u before the law
  void applyEigenGradU(const double (&eigen_grad_u)[9]);
  void setEigenGradU(const Array<double> &eigen_grad_u);
  // MaterialThermal's "alpha" and "delta_T" (uniform, or one value per
  // Gauss point): sigma_th = -E / (1 - 2 nu) alpha delta_T
  void setThermal(double alpha, double delta_T);
  void setThermal(double alpha, const Array<double> &delta_T);
  // a user material's computeStress run on the device (no counterpart in
  // the reference): HIP C++ defining
  //   __device__ void akx_compute_stress(const double *grad_u,
  //                                      double *stress, const double *params)
  // compiled at this call; the ConstitutiveLaw is then not called
  void setDeviceStress(const std::string &source, const std::vector<double> &params = {});

private:
  SolidMechanicsModel *model;
  int id;
  std::int64_t n;
};

class SolidMechanicsModel : private FieldSync {
public:
  explicit SolidMechanicsModel(const Mesh &mesh, int device = 0)
      : nb_nodes(mesh.nodes.size()),
        nqp(mesh.type == _hexahedron_8 ? 8 : 1),
        nb_elements(mesh.connectivity.size()) {
    check(akx_model_create(device, &h));
    check(akx_model_set_mesh(h, mesh.type, nb_nodes, mesh.nodes.data(), nb_elements,
                             mesh.connectivity.data()));
    for (int f = 0; f < NF; ++f) {
      fields[f] = Array<double>(nb_nodes, 3);
      fields[f].bind(this, f);
    }
    blocked = Array<std::uint8_t>(nb_nodes, 3);
    blocked.bind(this, AKX_FIELD_BLOCKED_DOFS);
    stale = ALL;
  }
  ~SolidMechanicsModel() { akx_model_destroy(h); }
  SolidMechanicsModel(const SolidMechanicsModel &) = delete;
  SolidMechanicsModel &operator=(const SolidMechanicsModel &) = delete;

  // material <name> [ rho = .. E = .. nu = .. ] ; finite_deformation selects
  // the total-Lagrangian branch of "elastic" (with
  // "plastic_linear_isotropic_hardening" it throws: not provided)
  Material &addMaterial(const std::string &name, const MaterialParameters &p,
                        bool finite_deformation = false,
                        const std::vector<std::int32_t> &filter = {}) {
    int law = 0, id = -1;
    check(akx_material_law_by_name(name.c_str(), finite_deformation, &law));
    double params[5] = {p.rho, p.E, p.nu, p.sigma_y, p.h};
    check(akx_model_add_material(h, law, params, 5,
                                 filter.empty() ? nullptr : filter.data(),
                                 (std::int64_t)filter.size(), &id));
    materials.emplace_back(*this, id,
                           filter.empty() ? nb_elements : (std::int64_t)filter.size());
    return materials.back();
  }
  // a material whose computeStress is `law` (kept by reference; the caller
  // keeps it alive); rho, E, nu give the lumped mass and the push-wave speed
  Material &addMaterial(ConstitutiveLaw &law, const MaterialParameters &p,
                        bool finite_deformation = false,
                        const std::vector<std::int32_t> &filter = {}) {
    int id = -1;
    double params[3] = {p.rho, p.E, p.nu};
    check(akx_model_add_material(h, finite_deformation ? AKX_LAW_USER_FD : AKX_LAW_USER,
                                 params, 3, filter.empty() ? nullptr : filter.data(),
                                 (std::int64_t)filter.size(), &id));
    user_laws[id] = &law;
    check(akx_model_set_stress_callback(h, &SolidMechanicsModel::stress_trampoline, this));
    materials.emplace_back(*this, id,
                           filter.empty() ? nb_elements : (std::int64_t)filter.size());
    return materials.back();
  }
  Material &getMaterial(int i) { return materials.at(i); }

  void initFull(AnalysisMethod = _explicit_lumped_mass) {
    push();
    check(akx_model_init_full(h));
    stale = ALL;
  }
  double getStableTimeStep() {
    push();
    double dt = 0.;
    check(akx_model_get_stable_time_step(h, &dt));
    return dt;
  }
  // ModelSolver::setTimeStep / getTimeStep (model_solver.cc:316-326)
  void setTimeStep(double dt, const std::string &solver_id = "") {
    check_solver(solver_id);
    check(akx_model_set_time_step(h, dt));
    time_step = dt;
  }
  double getTimeStep(const std::string &solver_id = "") const {
    check_solver(solver_id);
    return time_step;
  }
  // ModelSolver::solveStep(const ID & solver_id = "") (model_solver.hh:78):
  // "" and "explicit_lumped" name the solver initFull(_explicit_lumped_mass)
  // registers (solid_mechanics_model.cc:165-169)
  void solveStep(const std::string &solver_id = "") {
    push();
    run(akx_model_solve_step(h, solver_id.c_str()));
    stale = ALL;
  }
  // nb_steps solveStep() calls in one library call (no counterpart in the
  // reference; the same results bit for bit)
  void solveSteps(std::int64_t nb_steps) {
    push();
    run(akx_model_solve_steps(h, nb_steps));
    stale = ALL;
  }
  void assembleInternalForces() {
    push();
    run(akx_model_assemble_internal_forces(h));
    stale = ALL;
  }

  // nodal Arrays (solid_mechanics_model.hh:440-468)
  Array<double> &getDisplacement() { return fields[AKX_FIELD_DISPLACEMENT]; }
  Array<double> &getVelocity() { return fields[AKX_FIELD_VELOCITY]; }
  Array<double> &getAcceleration() { return fields[AKX_FIELD_ACCELERATION]; }
  Array<double> &getExternalForce() { return fields[AKX_FIELD_EXTERNAL_FORCE]; }
  const Array<double> &getInternalForce() const { return fields[AKX_FIELD_INTERNAL_FORCE]; }
  const Array<double> &getMass() const { return fields[AKX_FIELD_MASS]; }
  Array<std::uint8_t> &getBlockedDOFs() { return blocked; }

  akx_model *handle() { return h; }
  int nbQuadraturePoints() const { return nqp; }

private:
  friend class Material;
  static constexpr int NF = 6;
  static constexpr std::uint32_t ALL = ((1u << NF) - 1) | (1u << AKX_FIELD_BLOCKED_DOFS);
  static void check_solver(const std::string &id) {
    if (!id.empty() && id != "explicit_lumped")
      throw Exception(AKX_ERR_INVALID, "The non linear solver dof_manager:tss:" + id +
                                           " does not exists in dof_manager");
  }
  void pull(int f) override {
    void *dst = f == AKX_FIELD_BLOCKED_DOFS ? (void *)blocked.raw() : (void *)fields[f].raw();
    check(akx_model_get_field(h, f, dst));
    stale &= ~(1u << f);
  }
  // upload the copies accessed for writing (current with the device when the
  // access was made, so the upload is exact)
  void push() {
    for (int f : {AKX_FIELD_DISPLACEMENT, AKX_FIELD_VELOCITY, AKX_FIELD_ACCELERATION,
                  AKX_FIELD_EXTERNAL_FORCE, AKX_FIELD_BLOCKED_DOFS})
      if (dirty >> f & 1u) {
        const void *src =
            f == AKX_FIELD_BLOCKED_DOFS ? (void *)blocked.raw() : (void *)fields[f].raw();
        check(akx_model_set_field(h, f, src));
        dirty &= ~(1u << f);
      }
  }

  // computeStress of the user materials; an exception a law throws is
  // rethrown from the call that ran it
  static int stress_trampoline(void *self, int material_id, std::int64_t nb_qp,
                               const double *grad_u, double *stress) {
    auto *m = static_cast<SolidMechanicsModel *>(self);
    try {
      Array<double> g(nb_qp, 9), s(nb_qp, 9);
      std::copy(grad_u, grad_u + 9 * nb_qp, g.data());
      m->user_laws.at(material_id)->computeStress(g, s);
      std::copy(s.data(), s.data() + 9 * nb_qp, stress);
      return 0;
    } catch (...) {
      m->law_error = std::current_exception();
      return 1;
    }
  }
  void run(int status) {
    if (law_error) {
      std::exception_ptr e = law_error;
      law_error = nullptr;
      std::rethrow_exception(e);
    }
    check(status);
  }

  akx_model *h = nullptr;
  std::map<int, ConstitutiveLaw *> user_laws;
  std::exception_ptr law_error;
  std::int64_t nb_nodes;
  int nqp;
  std::int64_t nb_elements;
  double time_step = 0.;
  std::array<Array<double>, NF> fields;
  Array<std::uint8_t> blocked;
  std::deque<Material> materials; // stable references
};

// Many independent models stepped together (akx_batch_*): each instance is
// one SolidMechanicsModel with one material and zero external force.
class Batch {
public:
  enum Precision { fp64 = AKX_PRECISION_FP64, fp32 = AKX_PRECISION_FP32 };
  Batch(Precision p, std::int64_t node_capacity, std::int64_t tet_capacity,
        std::int64_t hex_capacity, bool plasticity = false, int device = 0) {
    check(akx_batch_create(device, p, node_capacity, tet_capacity, hex_capacity,
                           plasticity, &h));
  }
  ~Batch() { akx_batch_destroy(h); }
  Batch(const Batch &) = delete;
  Batch &operator=(const Batch &) = delete;

  // SolidMechanicsModel(mesh) + addMaterial + initFull + setTimeStep(
  // dt_factor * getStableTimeStep()); returns the instance id
  int addInstance(const Mesh &mesh, const std::string &law, const MaterialParameters &p,
                  const Array<std::uint8_t> *blocked = nullptr,
                  const Array<double> *velocity = nullptr, double dt_factor = 0.5,
                  bool finite_deformation = false) {
    int lid = 0, id = -1;
    check(akx_material_law_by_name(law.c_str(), finite_deformation, &lid));
    double params[5] = {p.rho, p.E, p.nu, p.sigma_y, p.h};
    check(akx_batch_add_instance(h, mesh.type, lid, params, mesh.nodes.size(),
                                 mesh.nodes.data(), mesh.connectivity.size(),
                                 mesh.connectivity.data(),
                                 blocked ? blocked->data() : nullptr,
                                 velocity ? velocity->data() : nullptr, dt_factor, &id));
    if ((int)nb_nodes.size() <= id)
      nb_nodes.resize(id + 1);
    nb_nodes[id] = mesh.nodes.size();
    return id;
  }
  void removeInstance(int id) { check(akx_batch_remove_instance(h, id)); }
  // nb_steps solveStep() calls of every live instance
  void solveStep(std::int64_t nb_steps = 1) { check(akx_batch_solve_step(h, nb_steps)); }
  double getTimeStep(int id) {
    double dt = 0.;
    check(akx_batch_get_time_step(h, id, &dt, nullptr));
    return dt;
  }
  Array<double> getDisplacement(int id) { return field(id, AKX_FIELD_DISPLACEMENT); }
  Array<double> getVelocity(int id) { return field(id, AKX_FIELD_VELOCITY); }
  Array<double> getInternalForce(int id) { return field(id, AKX_FIELD_INTERNAL_FORCE); }
  // launch shape chosen by timing the element pass of the live instances
  void autotune(int repeats = 3) {
    check(akx_batch_autotune(h, repeats, nullptr, nullptr, nullptr, nullptr));
  }
  akx_batch *handle() { return h; }

private:
  Array<double> field(int id, int f) {
    Array<double> out(nb_nodes.at(id), 3);
    check(akx_batch_get_field(h, id, f, out.data()));
    return out;
  }
  akx_batch *h = nullptr;
  std::vector<std::int64_t> nb_nodes;
};

inline Array<double> Material::getGradU() const {
  model->push();
  model->run(akx_model_compute_all_stresses(model->h));
  Array<double> out(n * model->nqp, 9);
  check(akx_model_get_internal(model->h, id, AKX_INTERNAL_GRAD_U, out.data()));
  return out;
}

inline Array<double> Material::getStress() const {
  model->push();
  model->run(akx_model_compute_all_stresses(model->h));
  Array<double> out(n * model->nqp, 9);
  check(akx_model_get_internal(model->h, id, AKX_INTERNAL_STRESS, out.data()));
  return out;
}

inline void Material::applyEigenGradU(const double (&eigen_grad_u)[9]) {
  Array<double> e(n * model->nqp, 9);
  for (std::int64_t q = 0; q < n * model->nqp; ++q)
    for (int k = 0; k < 9; ++k)
      e(q, k) = eigen_grad_u[k];
  setEigenGradU(e);
}

inline void Material::setEigenGradU(const Array<double> &eigen_grad_u) {
  if (eigen_grad_u.size() != n * model->nqp || eigen_grad_u.getNbComponent() != 9)
    throw Exception(AKX_ERR_INVALID, "eigen_grad_u needs [nb_qp][9]");
  check(akx_model_set_eigen_grad_u(model->h, id, eigen_grad_u.data()));
}

inline void Material::setThermal(double alpha, double delta_T) {
  setThermal(alpha, Array<double>(n * model->nqp, 1, delta_T));
}

inline void Material::setThermal(double alpha, const Array<double> &delta_T) {
  if (delta_T.size() != n * model->nqp)
    throw Exception(AKX_ERR_INVALID, "delta_T needs one value per Gauss point");
  check(akx_model_set_thermal(model->h, id, alpha, delta_T.data()));
}

inline void Material::setDeviceStress(const std::string &source,
                                      const std::vector<double> &params) {
  check(akx_model_set_device_stress(model->h, id, source.c_str(),
                                    params.empty() ? nullptr : params.data(),
                                    (int)params.size()));
}

} // namespace akantu_mi355x
