/*
 * akantu_oracle.h -- CPU FP64 oracle for the explicit central-difference /
 * lumped-mass step of Akantu's SolidMechanicsModel.
 *
 * TEST INFRASTRUCTURE ONLY.  This file and akantu_oracle.c restate the
 * reference algorithm in plain C so that the HIP product path can be checked
 * against it.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  The product library never links it.
 *
 * Layouts follow the reference's Array<T> (array-of-structures, one tuple per
 * row): nodal fields are [nn][3], connectivity is [nel][nnpe], quadrature
 * point tensors are [nel*nqp][9] stored column-major (entry (i,j) at i+3j),
 * shape derivatives are [nel*nqp][3*nnpe] with dN_n/dX_j at j+3n
 * (shape_lagrange_inline_impl.hh:190-245), elemental vectors are
 * [nel][3*nnpe] with component i of node n at i+3n.
 */
#ifndef AKANTU_ORACLE_H
#define AKANTU_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_TET4 = 0, ORC_HEX8 = 1 };

enum {
  ORC_ELASTIC = 0,     /* material_elastic, finite_deformation = false */
  ORC_ELASTIC_FD = 1,  /* material_elastic, finite_deformation = true  */
  ORC_NEOHOOKEAN = 2,  /* material_neohookean (always finite deformation) */
  ORC_PLASTIC_LIH = 3  /* plastic_linear_isotropic_hardening, small strain */
};

typedef struct orc_model {
  int type; /* ORC_TET4 / ORC_HEX8 */
  int law;  /* ORC_ELASTIC ... */
  long nn;  /* number of nodes */
  long nel; /* number of elements */
  const double *X; /* [nn][3] reference (initial) coordinates */
  const int *conn; /* [nel][nnpe] */

  /* material parameters: scalars, or per-element arrays when non-NULL */
  double rho, E, nu, sigma_y, h;
  const double *E_el, *nu_el, *sigma_y_el, *h_el;

  /* FEEngine precompute (caller-allocated; B may be NULL -> recomputed on the
   * fly, which gives the same bits since the arithmetic is identical) */
  double *B;  /* [nel*nqp][3*nnpe] */
  double *wJ; /* [nel*nqp] */

  /* nodal state [nn][3] */
  double *u, *v, *a, *f_int, *f_ext, *mass, *u_prev;
  unsigned char *blocked; /* [nn][3] */

  /* quadrature-point internals [nel*nqp][9] (stress holds sigma for small
   * strain laws and the second Piola-Kirchhoff stress for finite
   * deformation) */
  double *grad_u, *stress;
  /* plasticity history (NULL for the other laws) */
  double *prev_grad_u, *prev_stress, *eps_p, *prev_eps_p; /* [nel*nqp][9] */
  double *iso, *prev_iso;                                 /* [nel*nqp] */
  /* plastic energy accumulated by afterSolveStep (material_plastic.cc
   * :86-99), [nel*nqp]; NULL: not tracked */
  double *plastic_energy;

  double dt;
  long step;

  /* Material::eigengradu (material.cc:35-36,80-85): subtracted from grad u at
   * every quadrature point before computeStress (material.cc:147),
   * [nel*nqp][9]; NULL: zero */
  const double *eigen_grad_u;
  /* MaterialThermal (material_thermal.cc:30-44, .hh:91-95): alpha and the
   * per-quadrature-point delta_T [nel*nqp]; NULL: sigma_th = -0 (alpha = 0
   * by default).  sigma_th and its previous value (history,
   * material_thermal.cc:34) [nel*nqp], written when delta_T is set. */
  double alpha;
  const double *delta_T;
  double *sigma_th, *prev_sigma_th;
} orc_model;

/* Threads of the element and DOF loops (default 1); results are
 * bit-identical for every count. */
void orc_set_threads(int n);
int orc_get_threads(void);

int orc_nnpe(int type);
int orc_nqp(int type);

/* Shape functions / natural derivatives / Gauss tables of the reference. */
void orc_gauss_points(int type, double *xi /* [nqp][3] */, double *w /* [nqp] */);
void orc_shapes(int type, const double *xi, double *N /* [nnpe] */);
void orc_dnds(int type, const double *xi, double *dnds /* [3][nnpe] row-major d/dxi_i */);

/* a13: B and wJ.  Returns 0, or (1 + global qp index) of the first negative
 * jacobian (the reference throws IntegratorGaussException there). */
long orc_precompute(orc_model *m);

/* a14: lumped mass M_a = sum_e sum_q rho N_a(q) wJ_q, same value in the three
 * components. */
void orc_lumped_mass(orc_model *m);

/* a15: stable time step min_e h_e / c_p,e on the current positions X + u. */
double orc_stable_dt(const orc_model *m);
double orc_element_inradius(int type, const double *Xe /* [nnpe][3] */);

/* a6+a7: grad_u and the constitutive law on all quadrature points. */
void orc_compute_all_stresses(orc_model *m);
/* a8+a9: f_int = 0 then assemble -int B^T sigma. */
void orc_assemble_internal_forces(orc_model *m);
/* a5: both of the above */
void orc_internal_forces(orc_model *m);

/* a1..a12: one solveStep() with _explicit_lumped_mass / central difference. */
void orc_solve_step(orc_model *m);
void orc_solve_steps(orc_model *m, long nsteps);

/* The lumped central-difference step on an arbitrary set of DOFs, with the
 * residual (f_ext + f_int part) supplied by a callback that is called after
 * the predictor.  Pins the integrator independently of the FE engine. */
typedef void (*orc_residual_cb)(void *ctx, double *r /* [ndof], zeroed */);
void orc_lumped_step_cb(long ndof, double dt, double *u, double *v, double *a,
                        const unsigned char *blocked, const double *mass,
                        orc_residual_cb cb, void *ctx);

/* Energies (solid_mechanics_model.cc:476-606, material.cc:604-649,
 * material_plastic.cc:47-99): kinetic and external work on the current
 * nodal state; potential on the current grad_u / stress (plastic: and
 * inelastic strain); plastic from the accumulated plastic_energy. */
double orc_kinetic_energy(const orc_model *m);
double orc_external_work(const orc_model *m);
double orc_potential_energy(const orc_model *m);
double orc_plastic_energy(const orc_model *m);
/* BC::Neumann::FromTraction on a set of boundary facets
 * (boundary_condition_tmpl.hh:100-160, boundary_condition_functor_inline_impl
 * .hh:109-113): f_ext += assemble(integrate(N^T t)) over the facets, in facet
 * order.  facet_type 2 = _triangle_3 (1 Gauss point), 3 = _quadrangle_4
 * (2 x 2); conn [nb_facets][3 or 4]. */
void orc_apply_traction(orc_model *m, int facet_type, long nb_facets, const int *conn,
                        const double *traction);

/* Math::reduce (aka_math.cc:79-107): sort ascending, then pairwise sums. */
double orc_math_reduce(double *values, long n);

/* Small helpers exposed for closed-form tests. */
void orc_law_stress(int law, double lambda, double mu, double sigma_y,
                    double h, const double *grad_u, const double *prev_grad_u,
                    const double *prev_sigma, double prev_iso,
                    const double *prev_eps_p, double *sigma, double *iso,
                    double *eps_p);
void orc_lame(double E, double nu, double *lambda, double *mu);
/* The same with the thermal terms: sigma_th of the elastic laws, and
 * delta_sigma_th = previous sigma_th - sigma_th of the plastic trial stress
 * (material_linear_isotropic_hardening_inline_impl.hh:33-47). */
void orc_law_stress_th(int law, double lambda, double mu, double sigma_y,
                       double h, const double *grad_u, const double *prev_grad_u,
                       const double *prev_sigma, double prev_iso,
                       const double *prev_eps_p, double sigma_th,
                       double delta_sigma_th, double *sigma, double *iso,
                       double *eps_p);

#ifdef __cplusplus
}
#endif
#endif
