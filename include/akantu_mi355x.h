/*
 * akantu_mi355x.h -- C ABI of the MI355X (gfx950) explicit-dynamics
 * internal-force path.
 *
 * One `akx_model` is one SolidMechanicsModel: a mesh of one element type,
 * one or more materials (each with an element filter), device-resident
 * nodal state, one HIP stream.  Every function returns AKX_OK (0) or a
 * negative status; akx_last_error() gives the message of the last failure on
 * the calling thread.  Nothing throws across this boundary.
 *
 * Host arrays passed in or out use the reference's Array<T> layout
 * (array-of-structures, one tuple per row: nodal fields [nb_nodes][3],
 * connectivity [nb_elements][nnpe], quadrature-point tensors
 * [nb_elements*nqp][9] column-major).  The library converts to its own
 * structure-of-arrays device layout.  The caller owns host memory; the
 * library owns device memory.
 *
 * Reference interfaces replaced (paths relative to the reference root):
 *   SolidMechanicsModel(mesh) + initFull(_analysis_method =
 *     _explicit_lumped_mass)      src/model/solid_mechanics/solid_mechanics_model.cc:52-78,186-232
 *   ModelSolver::solveStep       src/model/common/model_solver.cc:267-280
 *   getStableTimeStep            solid_mechanics_model.cc:423-473
 *   setTimeStep                  src/model/common/model_solver.cc (setTimeStep)
 *   getDisplacement/getVelocity/getAcceleration/getInternalForce/
 *   getExternalForce/getBlockedDOFs/getMass
 *                                solid_mechanics_model.hh:440-468 (the nodal Arrays)
 *   Material::computeAllStresses + Material::computeStress (plugin)
 *                                src/model/solid_mechanics/material.cc:128-154, material.hh:78-81
 *   MaterialFactory / instantiateMaterial<Mat>("name")
 *                                material.hh:41-42,610-621
 *   SolidMechanicsModel::assembleInternalForces
 *                                solid_mechanics_model.cc:333-370
 */
#ifndef AKANTU_MI355X_H
#define AKANTU_MI355X_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AKX_OK 0
#define AKX_ERR_INVALID (-1)    /* bad argument or call order */
#define AKX_ERR_HIP (-2)        /* HIP runtime failure */
#define AKX_ERR_JACOBIAN (-3)   /* negative jacobian (IntegratorGaussException) */
#define AKX_ERR_MEMORY (-4)     /* device allocation failed */
#define AKX_ERR_UNSUPPORTED (-5)

/* ElementType subset (src/common/aka_element_classes_info.hh) */
#define AKX_TET4 0  /* _tetrahedron_4 */
#define AKX_HEX8 1  /* _hexahedron_8 */
#define AKX_TRI3 2  /* _triangle_3, boundary facet of Tet4 */
#define AKX_QUAD4 3 /* _quadrangle_4, boundary facet of Hex8 */

/* Constitutive laws registered in the material factory under the
 * reference's names: "elastic" (material_elastic.cc:216-217), "neohookean"
 * (material_neohookean.cc:161-162), "plastic_linear_isotropic_hardening". */
#define AKX_LAW_ELASTIC 0     /* finite_deformation = false */
#define AKX_LAW_ELASTIC_FD 1  /* "elastic" with finite_deformation = true */
#define AKX_LAW_NEOHOOKEAN 2
#define AKX_LAW_PLASTIC_LIH 3
/* A law supplied by the caller: the Material::computeStress plugin
 * (material.hh:78-81, a subclass registered with
 * MaterialFactory::registerAllocator, material.hh:610-621; from Python
 * py_material.cc:45-80,237-240).  params = {rho, E, nu}: E and nu give the
 * push-wave speed of the stable time step and rho the lumped mass.  The plain
 * FP64 path runs it (AKX_PRECISION_FP64). */
#define AKX_LAW_USER 4     /* small strain: the callback writes sigma */
#define AKX_LAW_USER_FD 5  /* finite deformation: the callback writes PK2 */

/* Arithmetic of the element pass */
#define AKX_PRECISION_FP64 0 /* plain path (default): FP64 in the reference's
                                operation order, no FMA; the CPU oracle is
                                reproduced bit for bit except the last bit
                                of log() */
#define AKX_PRECISION_FP32 1 /* grad u, stress, B^T sigma in FP32; nodal
                                update in FP64 with compensated summation.
                                Between the steps of one solve call the
                                velocity field holds v + dt/2 a and the
                                acceleration field is not updated; both are
                                v and a again when the call returns */
#define AKX_PRECISION_FP64_FAST 2 /* FP64 with FMA and the regrouped
                                contraction r_e = sum_q Q_q dN/dxi_q; equal to
                                the plain path to rounding per evaluation */

/* Nodal fields, [nb_nodes][3] doubles unless stated */
#define AKX_FIELD_DISPLACEMENT 0
#define AKX_FIELD_VELOCITY 1
#define AKX_FIELD_ACCELERATION 2
#define AKX_FIELD_INTERNAL_FORCE 3
#define AKX_FIELD_EXTERNAL_FORCE 4
#define AKX_FIELD_MASS 5          /* [nb_nodes][3], read only */
#define AKX_FIELD_BLOCKED_DOFS 6  /* [nb_nodes][3] uint8 (bool) */
#define AKX_FIELD_NODES 7         /* [nb_nodes][3], read only */

/* Quadrature-point internals, [nb_filter_elements * nqp][9] column-major */
#define AKX_INTERNAL_GRAD_U 0
#define AKX_INTERNAL_STRESS 1          /* sigma (small strain) or PK2 */
#define AKX_INTERNAL_INELASTIC_STRAIN 2 /* plastic only */
#define AKX_INTERNAL_ISO_HARDENING 3    /* plastic only, [n*nqp][1] */

typedef struct akx_model akx_model;

/* Library ---------------------------------------------------------------- */
const char *akx_last_error(void);
int akx_version(void);
/* Number of visible gfx950 devices (0 on a machine without GPU). */
int akx_device_count(int *count);

/* Model life cycle ---------------------------------------------------------
 * Replaces `SolidMechanicsModel model(mesh)`: the mesh is given as arrays. */
int akx_model_create(int device, akx_model **out);
int akx_model_destroy(akx_model *m);
int akx_model_set_mesh(akx_model *m, int element_type, int64_t nb_nodes,
                       const double *nodes, int64_t nb_elements,
                       const int32_t *connectivity);

/* Material -----------------------------------------------------------------
 * Replaces the parsed `material <name> [ rho = .. E = .. nu = .. ]` section
 * and MaterialFactory::allocate.  params = {rho, E, nu, sigma_y, h} (the last
 * two only for the plastic law).  `filter` lists the elements of the
 * material in assembly order (NULL: all elements in ascending order).
 * Materials are assembled in creation order, like for_each_constitutive_law
 * (constitutive_laws_handler.hh:116-126). */
int akx_model_add_material(akx_model *m, int law, const double *params,
                           int nparams, const int32_t *filter,
                           int64_t nb_filter, int *material_id);
/* Per-element E, nu (and sigma_y, h for plasticity) over the material's
 * filter, for the per-element-parameter material.  Any pointer may be NULL
 * (keeps the scalar). */
int akx_model_set_element_parameters(akx_model *m, int material_id,
                                     const double *E, const double *nu,
                                     const double *sigma_y, const double *h);
/* Look a law up by the reference's material name ("elastic", "neohookean",
 * "plastic_linear_isotropic_hardening"); finite_deformation selects the total
 * Lagrangian branch of "elastic".  "plastic_linear_isotropic_hardening" with
 * finite_deformation != 0 (material_linear_isotropic_hardening.cc:36-52) is
 * not provided and returns AKX_ERR_UNSUPPORTED; unknown names too. */
int akx_material_law_by_name(const char *name, int finite_deformation,
                             int *law);

/* Material::eigengradu (material.cc:35-36,80-85; applyEigenGradU
 * material.cc:811-824): subtracted from grad u at every Gauss point of the
 * material before computeStress (gradu_vect -= eigengradu, material.cc:147),
 * so the law, the finite-deformation force (B_NL) and the plastic history all
 * see grad u - eigen_grad_u.  eigen_grad_u [nb_filter * nqp][9]
 * (column-major, points in the material's element order); NULL: zero (the
 * default).  Plain FP64 path only (AKX_ERR_UNSUPPORTED otherwise). */
int akx_model_set_eigen_grad_u(akx_model *m, int material_id, const double *eigen_grad_u);
/* MaterialThermal (material_thermal.cc:30-44, .hh:91-95): alpha and the
 * temperature change delta_T [nb_filter * nqp] give sigma_th = -E / (1 -
 * 2 nu) * alpha * delta_T at every Gauss point (the material's E, nu, per
 * element when set); the elastic laws add it to the diagonal of the stress
 * (material_elastic_inline_impl.hh:33-49; of S for finite deformation), the
 * plastic law adds previous sigma_th - sigma_th to its trial stress
 * (material_linear_isotropic_hardening_inline_impl.hh:33-47), the
 * neo-Hookean law does not read it.  delta_T NULL: no thermal stress (the
 * default, alpha = 0).  Plain FP64 path only. */
int akx_model_set_thermal(akx_model *m, int material_id, double alpha, const double *delta_T);

int akx_model_set_precision(akx_model *m, int precision);

/* computeStress of the AKX_LAW_USER / AKX_LAW_USER_FD materials.  Called on
 * the calling thread, once per user material and per stress evaluation
 * (every solveStep, assembleInternalForces, computeAllStresses), with the
 * material's Gauss-point gradients grad_u [nb_qp][9] (column-major,
 * du_i/dX_j at i + 3 j) and its stress [nb_qp][9] to fill (sigma, or PK2
 * for AKX_LAW_USER_FD), as Material::getGradU / getStress give them to
 * computeStress.  Gauss points are in the material's element order, nqp
 * per element.  A non-zero return aborts the call with AKX_ERR_INVALID.
 * The arrays are valid during the call only. */
typedef int (*akx_compute_stress_fn)(void *user, int material_id, int64_t nb_qp,
                                     const double *grad_u, double *stress);
int akx_model_set_stress_callback(akx_model *m, akx_compute_stress_fn fn, void *user);

/* computeStress of one user material run on the device: no counterpart in
 * the reference, whose plugin runs on the host (the callback above, which
 * costs a device-host round trip of the material's grad u and stress per
 * evaluation).  source is HIP C++ defining
 *   __device__ void akx_compute_stress(const double *grad_u, double *stress,
 *                                      const double *params);
 * for one Gauss point: grad_u and stress are 3x3, column-major as for the
 * callback (stress is zero on entry), params the nparams values given here
 * (copied).  At this call the source is compiled into the material's
 * element pass for the model's GPU (hiprtc, the library's own kernel
 * headers, -O3, no floating-point contraction: every multiply and add
 * rounds on its own, as on the plain path): the law runs per Gauss point
 * inside the pass, between grad u and the force terms.  A compile error
 * returns AKX_ERR_INVALID with the compiler's log in akx_last_error.  From
 * then on the callback is not called for the material.  Calling again
 * replaces the law; source NULL returns the material to the callback.  The
 * pipelined schedule does not apply to a model with a device law. */
int akx_model_set_device_stress(akx_model *m, int material_id, const char *source,
                                const double *params, int nparams);

/* FEEngine / DOFManager calls on caller fields, run on the device in the
 * reference's operation order (the one of the element passes).  filter
 * lists elements in order (NULL: all, ascending).  Per-point arrays are
 * [nb_filter * nqp][comps], points of the i-th filtered element at rows
 * i * nqp ..; a matrix per point is column-major.
 *
 * FEEngine::gradientOnIntegrationPoints(u, nablauq, nb_dof, type, ghost,
 * filter) (fe_engine.hh:136-139, shape_functions_inline_impl.hh:263-315):
 * field [nb_nodes][nb_dof] -> out [nb_filter * nqp][nb_dof * 3], the
 * nb_dof x 3 gradient du_k/dX_j at k + nb_dof j. */
int akx_fe_gradient_on_integration_points(akx_model *m, const double *field, int nb_dof,
                                          const int32_t *filter, int64_t nb_filter,
                                          double *out);
/* FEEngine::computeBtD(Ds, BtDs, type, ghost, filter) (fe_engine.hh:155-158,
 * shape_lagrange_inline_impl.hh:373-410): Ds [nb_filter * nqp][3 d] (a d x 3
 * matrix per point) -> BtDs [nb_filter * nqp][d * nnpe], D B at k + d n. */
int akx_fe_compute_btd(akx_model *m, const double *Ds, int nb_comp, const int32_t *filter,
                       int64_t nb_filter, double *BtDs);
/* FEEngine::integrate(f, intf, nb_dof, type, ghost, filter) (fe_engine.hh:85-94,
 * integrator_gauss_inline_impl.hh:401-441): f [nb_filter * nqp][nb_dof] ->
 * intf [nb_filter][nb_dof], sum over the Gauss points of f w detJ. */
int akx_fe_integrate(akx_model *m, const double *f, int nb_dof, const int32_t *filter,
                     int64_t nb_filter, double *intf);
/* DOFManager::assembleElementalArrayLocalArray(elementary, assembled, type,
 * ghost, scale, filter) (dof_manager.hh:85, dof_manager.cc:65-120):
 * assembled [nb_nodes][nb_dof] += scale * elementary [nb_filter][nb_dof *
 * nnpe] (node-major, dof k of local node n at k + nb_dof n), each node's
 * contributions added in filter order. */
int akx_dof_assemble_elemental_array(akx_model *m, const double *elementary, int nb_dof,
                                     const int32_t *filter, int64_t nb_filter,
                                     double scale, double *assembled);

/* initFull(_analysis_method = _explicit_lumped_mass): checks the jacobians
 * (AKX_ERR_JACOBIAN, like IntegratorGauss::checkJacobians), builds the
 * node->element gather table and assembles the lumped mass. */
int akx_model_init_full(akx_model *m);

/* Time stepping -------------------------------------------------------------*/
int akx_model_get_stable_time_step(akx_model *m, double *dt);
int akx_model_set_time_step(akx_model *m, double dt);
/* ModelSolver::solveStep(const ID & solver_id = "") (model_solver.hh:78,
 * model_solver.cc:267-280): one explicit step.  solver_id NULL or "" is the
 * default solver, "explicit_lumped" names it (the id initFull(
 * _explicit_lumped_mass) registers, solid_mechanics_model.cc:165-169); any
 * other id fails with AKX_ERR_INVALID, as DOFManager::getTimeStepSolver
 * throws for a solver that does not exist (dof_manager.cc:585-596).  The
 * step is queued on the model's stream; a following field read waits for it. */
int akx_model_solve_step(akx_model *m, const char *solver_id);
/* nb_steps consecutive solveStep() calls in one call, no counterpart in the
 * reference: bit-identical to nb_steps calls of akx_model_solve_step, with the
 * next step's predictor fused into each step's nodal gather. */
int akx_model_solve_steps(akx_model *m, int64_t nb_steps);
/* SolidMechanicsModel::assembleInternalForces on the current displacement
 * (f_int only, no time integration). */
int akx_model_assemble_internal_forces(akx_model *m);

/* Fields --------------------------------------------------------------------*/
int akx_model_get_nb_nodes(akx_model *m, int64_t *nb_nodes);
int akx_model_get_field(akx_model *m, int field, void *host);
int akx_model_set_field(akx_model *m, int field, const void *host);
/* BC::Neumann::FromTraction applied on a group of boundary facets
 * (boundary_condition_tmpl.hh:100-160, boundary_condition_functor_inline_impl
 * .hh:109-113): f_ext += the integral of N^T t over the facets, added in
 * facet order.  facet_type AKX_TRI3 ([nb][3]) or AKX_QUAD4 ([nb][4]);
 * traction[3] constant. */
int akx_model_apply_traction(akx_model *m, int facet_type, int64_t nb_facets,
                             const int32_t *facet_connectivity,
                             const double *traction);
/* Material::computeAllStresses on the current displacement, then copy one
 * internal of one material to the host. */
int akx_model_compute_all_stresses(akx_model *m);
int akx_model_get_internal(akx_model *m, int material_id, int internal,
                           double *host);

/* SolidMechanicsModel::getEnergy(energy_id) (material_id = -1) or
 * Material::getEnergy(energy_id) of one material, on the device state:
 * "kinetic", "external work" (solid_mechanics_model.cc:476-606; model
 * energies only), "potential" (material.cc:604-619 and the laws' densities),
 * "plastic" (material_plastic.cc:47-66: the plastic work accumulated after
 * every step, material_plastic.cc:86-99).  Other ids give 0, as in the
 * reference.  Sums are deterministic tree sums. */
int akx_model_get_energy(akx_model *m, const char *energy_id, int material_id,
                         double *energy);

/* No-allocation variant of akx_model_get_energy, with no counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * ids, same values (the same tree sum, its first level fused with the
 * per-node / per-element values); it calls no hipMalloc, hipFree or
 * allocating library routine.  Its partial sums live in one library-owned
 * device array per GPU, AKX_NOALLOC_MAX_ITEMS / 2048 + 256 doubles (about
 * 0.5 MB), held until the process ends: the model's node count ("kinetic",
 * "external work") or each selected material's element count ("potential",
 * "plastic") may be at most AKX_NOALLOC_MAX_ITEMS, else AKX_ERR_INVALID.
 * "potential" needs the stress arrays an earlier akx_model_get_energy(
 * "potential") or akx_model_compute_all_stresses allocated (AKX_ERR_INVALID
 * otherwise) and a model without user materials (AKX_ERR_UNSUPPORTED).
 * Calls on one GPU share its array: calls that overlap in time on one GPU
 * (two threads, or models on different streams) race and must be
 * serialised by the caller. */
#define AKX_NOALLOC_MAX_ITEMS ((int64_t)1 << 27) /* 134,217,728 */
int akx_model_get_energy_noalloc(akx_model *m, const char *energy_id, int material_id,
                                 double *energy);

/* Launch schedule of solve_step, no counterpart in the reference: with
 * parts > 1 the elements are cut in `parts` consecutive parts and the nodal
 * gather of the nodes a part completes runs on a second stream while the
 * next part's element pass runs (same values and summation orders as the
 * plain schedule).  Available for one contiguous non-plastic material whose
 * node numbering follows the elements (structured meshes); *active reports
 * whether it applies.  parts = 0 or 1: the plain schedule (default). */
int akx_model_set_pipeline(akx_model *m, int parts, int *active);

/* Assembly of the element passes, no counterpart in the reference (call
 * before akx_model_init_full).  It applies to models whose materials are all
 * contiguous element ranges: Tet4 (in FP64 with built-in laws) and, in FP32,
 * Hex8.  A node's corners, in assembly order (materials in order, elements
 * ascending), come in runs: its corners in one block of elements (256 Tet4,
 * 192 Hex8).
 * - AKX_ASSEMBLY_PER_CORNER: one table entry per element corner, added by
 *   the nodal gather in assembly order.
 * - AKX_ASSEMBLY_DEFAULT: in FP64, prefix partials: the block of a node's
 *   first run sums it, in that order, into one entry and every later corner
 *   keeps its own; the gather's sums keep their bits (same f_int as per
 *   corner, fewer entries).  In FP32, block partials (below).
 * - AKX_ASSEMBLY_BLOCK_PARTIALS: every run is summed in its block, one entry
 *   per (block, node), added in block order: bit-reproducible, but rounded
 *   otherwise than the per-corner order (FP64: opt-in; FP32: the default).
 * get_block_assembly reports the assembly the initialised model uses:
 * AKX_ASSEMBLY_PER_CORNER, AKX_ASSEMBLY_PREFIX_PARTIALS or
 * AKX_ASSEMBLY_BLOCK_PARTIALS. */
#define AKX_ASSEMBLY_PER_CORNER 0
#define AKX_ASSEMBLY_DEFAULT 1
#define AKX_ASSEMBLY_PREFIX_PARTIALS 1
#define AKX_ASSEMBLY_BLOCK_PARTIALS 2
int akx_model_set_block_assembly(akx_model *m, int enable);
int akx_model_get_block_assembly(akx_model *m, int *active);

/* Synchronise the model's stream (all queued steps finished). */
int akx_model_synchronize(akx_model *m);

/* Kernel timing -------------------------------------------------------------
 * When enabled, every launch of the element pass and of the nodal gather is
 * bracketed by HIP events on the stream it runs on.  Kernel classes:
 * 0 = element pass, 1 = nodal gather + update, 2 = predictor. */
int akx_model_enable_kernel_timing(akx_model *m, int enable);
int akx_model_get_kernel_timing(akx_model *m, int kernel_class,
                                double *total_ms, int64_t *launches);
/* Bytes each kernel class moves per launch in the unique-bytes model (each
 * array read or written once, node data once per node); for class 1 the
 * gather of a step followed by another, which does not store f_int. */
int akx_model_get_kernel_bytes(akx_model *m, int kernel_class, double *bytes);
/* Streaming device copy of `bytes` (read + write), repeated; average time
 * per copy in ms.  No counterpart in the reference: it calibrates the HBM
 * byte counters and reports the measured copy bandwidth. */
int akx_calibration_copy(int device, int64_t bytes, int repeats, double *ms);

/* Batches ------------------------------------------------------------------
 * An `akx_batch` holds many independent instances (each one what the
 * reference runs as its own SolidMechanicsModel with one material, explicit
 * lumped mass and zero external force) in device arenas sized at creation,
 * and advances all live instances together: per step one element launch per
 * (element type, law) group present (one tile of elements per workgroup)
 * and one nodal gather + update launch (one 256-node slice per workgroup),
 * captured as a hipGraph of two steps.  FP32 batches assemble by block
 * partials (one force-table entry per (element block, node), added in block
 * order).  The
 * reference has no batch object; per instance the calls stand for
 *   SolidMechanicsModel(mesh) + addMaterial + initFull   -> add_instance
 *   getStableTimeStep / setTimeStep                      -> get/set_time_step
 *   solveStep (nb_steps times, every live instance)      -> solve_step
 *   getDisplacement / getVelocity / getAcceleration /
 *   getInternalForce                                     -> get_field
 * with the same arithmetic as akx_model (FP64: the reference's operation
 * order; FP32: the precision form).  An instance's results do not depend on
 * which other instances share the batch (per-instance arithmetic, per-node
 * gather in fixed order). */
typedef struct akx_batch akx_batch;

/* precision: AKX_PRECISION_FP64 or AKX_PRECISION_FP32.  Capacities: nodes
 * (each instance takes its node count rounded up to 256), Tet4 and Hex8
 * elements; with_plasticity reserves the plastic history (ping-pong). */
int akx_batch_create(int device, int precision, int64_t node_capacity,
                     int64_t tet_capacity, int64_t hex_capacity,
                     int with_plasticity, akx_batch **out);
int akx_batch_destroy(akx_batch *b);
/* Device bytes one instance takes (admission by HBM footprint). */
int akx_batch_footprint(int precision, int element_type, int law,
                        int64_t nb_nodes, int64_t nb_elements, int64_t *bytes);
/* Add an instance: mesh arrays as in akx_model_set_mesh, params = {rho, E,
 * nu, sigma_y, h}, blocked_dofs [nb_nodes][3] uint8 (NULL: none), initial
 * velocity [nb_nodes][3] (NULL: zero), dt = dt_factor * stable time step
 * (computed here).  AKX_ERR_MEMORY when an arena has no room (nothing is
 * kept), AKX_ERR_JACOBIAN for an inverted element.  Instance ids of removed
 * instances are reused.  Calls between solve_step calls only. */
int akx_batch_add_instance(akx_batch *b, int element_type, int law,
                           const double *params, int64_t nb_nodes,
                           const double *nodes, int64_t nb_elements,
                           const int32_t *connectivity,
                           const uint8_t *blocked_dofs, const double *velocity,
                           double dt_factor, int *instance_id);
int akx_batch_remove_instance(akx_batch *b, int instance_id);
int akx_batch_get_time_step(akx_batch *b, int instance_id, double *dt,
                            double *dt_crit);
int akx_batch_set_time_step(akx_batch *b, int instance_id, double dt);
/* Steps the instance has done since it was added. */
int akx_batch_get_steps(akx_batch *b, int instance_id, int64_t *steps);
/* nb_steps steps of every live instance (queued on the batch's stream). */
int akx_batch_solve_step(akx_batch *b, int64_t nb_steps);
int akx_batch_synchronize(akx_batch *b);
/* Nodal field of one instance, [nb_nodes][3] doubles: displacement,
 * velocity, acceleration, internal force or nodes. */
int akx_batch_get_field(akx_batch *b, int instance_id, int field, double *host);
/* Launch shape: threads per block (64, 128, 256; Hex8 uses 8 lanes per
 * element), elements per tile, hipGraph on/off.  Every workgroup does one
 * tile; blocks_per_cu (0..32) is kept for compatibility and has no effect. */
int akx_batch_set_launch(akx_batch *b, int block, int tile, int blocks_per_cu,
                         int use_graph);
int akx_batch_get_launch(akx_batch *b, int *block, int *tile,
                         int *blocks_per_cu, int *use_graph);
/* Time the element pass of the live instances over the launch-shape
 * candidates (repeats launches each) and keep the fastest. */
int akx_batch_autotune(akx_batch *b, int repeats, int *block, int *tile,
                       int *blocks_per_cu, double *best_ms);
int akx_batch_get_stats(akx_batch *b, int64_t *live, int64_t *elements,
                        int64_t *quadrature_points, int64_t *nodes);
/* Largest free range of each arena (nodes, Tet4, Hex8, force-table entries). */
int akx_batch_capacity_left(akx_batch *b, int64_t *nodes, int64_t *tets,
                            int64_t *hexes, int64_t *entries);
/* Per-launch HIP-event timing (disables the graph while on); classes
 * 0 = element pass (all groups of one step), 1 = gather + update. */
int akx_batch_enable_kernel_timing(akx_batch *b, int enable);
int akx_batch_get_kernel_timing(akx_batch *b, int kernel_class,
                                double *total_ms, int64_t *launches);
/* Unique bytes one launch of the class moves over the live instances. */
int akx_batch_get_kernel_bytes(akx_batch *b, int kernel_class, double *bytes);

#ifdef __cplusplus
}
#endif
#endif
