/*
 * beta_oracle.h -- CPU restatement of BETA's finite-element hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the
 * HIP path (libbeta_hip.so).  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  Nothing in betanoox_amd/ may
 * link, import or call it.
 *
 * Every function restates one routine of the reference (file:line cited at
 * each definition in beta_oracle.c) in plain serial C, keeping the
 * reference's loop orders so that its FP64 summation order is reproduced.
 */
#ifndef BETA_ORACLE_H
#define BETA_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (mirror include/beta_hip.h) ---- */
#define ORC_OK              0
#define ORC_ERR_ARG        -1
#define ORC_NOT_CONVERGED  -4
#define ORC_ZERO_DIAG      -5
#define ORC_NEG_DIAG       -6
#define ORC_NEG_DETJ       -7

/* ---- quadrature ---- */
void orc_gauss(int k, double *coords, double *weights);
int  orc_hex_quad_points(int order, double *xi, double *eta, double *zeta, double *w);
int  orc_quad_quad_points(int order, double *xi, double *eta, double *w);

/* ---- shape functions ---- */
void orc_shape_hex8(double *S, double xi, double eta, double zeta);
void orc_deriv_hex8(double *Sxi, double *Seta, double *Szeta, double xi, double eta, double zeta);
void orc_shape_quad4(double *S, double xi, double eta);
void orc_deriv_quad4(double *Sxi, double *Seta, double xi, double eta);

/* ---- Jacobian: returns detJ, writes global derivatives ---- */
double orc_jacobian3d(int n, const double *dxi, const double *deta, const double *dzeta,
                      const double *x, const double *y, const double *z,
                      double *dx, double *dy, double *dz);
double orc_jacobian2d(int n, const double *dxi, const double *deta,
                      const double *x, const double *y, double *dx, double *dy);

/* ---- dense LU (Numerical Recipes, as matrix.cpp) ---- */
int  orc_ludcmp(int n, double *a, int *indx, double *d);
void orc_lubksb(int n, const double *a, const int *indx, double *b);
int  orc_invert(int n, double *a);

/* ---- materials ---- */
int  orc_elastic_cmat(const double *moduli9, double *C36);      /* E11 E22 E33 nu12 nu23 nu13 G12 G23 G13 */
int  orc_elastic_iso_cmat(double E, double nu, double *C36);
void orc_plane_stress_cmat(double e11, double e22, double nu12, double G12, double *C9);

/* ---- element matrices (row-major nd x nd, upper triangle written) ---- */
int orc_ke_hex8_elastic(const double *xyz24, const double *C36, int order, double *Ke);
int orc_ke_quad4_elastic(const double *xy8, const double *C9, int order, double factor, double *Ke);
int orc_ke_hex8_heat(const double *xyz24, const double *k9, double capacity, int order,
                     double *Ke, double *Me);
int orc_fe_hex8_elastic(const double *xyz24, const double *C36, int order,
                        const double *ue24, double *Fe);

/* ---- Equations + SymmetricSparseMatrix restatement ---- */
typedef struct orc_eq orc_eq;
orc_eq   *orc_eq_create(int n_dof);
void      orc_eq_destroy(orc_eq *eq);
void      orc_eq_set_restraint(orc_eq *eq, int eqn, double value);
void      orc_eq_specify(orc_eq *eq, const int *map, int size);
void      orc_eq_allocate(orc_eq *eq);
long long orc_eq_nnz(const orc_eq *eq);
void      orc_eq_get_pattern(const orc_eq *eq, int *rowptr, int *colind);
void      orc_eq_get_values(const orc_eq *eq, double *val);
void      orc_eq_set_values(orc_eq *eq, const double *val);
void      orc_eq_zero(orc_eq *eq);
int       orc_eq_add_matrix(orc_eq *eq, const double *Ke, const int *map, int size,
                            const double *force);
int       orc_eq_add_load(orc_eq *eq, int eqn, double load);
void      orc_eq_build_rhs(const orc_eq *eq, double *b);
void      orc_eq_get_pseudo(const orc_eq *eq, double *f);
void      orc_eq_set_replace_zero_diag(orc_eq *eq, double value);
void      orc_eq_spmv(const orc_eq *eq, const double *x, double *y);
int       orc_eq_check_fix_diag(orc_eq *eq);
int       orc_eq_pcg(orc_eq *eq, double *b, const double *x0, double eps, int maxit,
                     int stop_rule, int *iters, double *final_norm);
int       orc_eq_skyline_solve(const orc_eq *eq, double *b);
/* ParCGMatrix::solve (ParCG.cpp:93-347): two OpenMP threads, ||r||/||x|| stop rule */
int       orc_eq_pcg_parcg(orc_eq *eq, double *b, const double *x0, double eps, int maxit,
                           int *iters, double *final_norm);

/* ---- element DOF map (BasicElement::calculateDofList) ---- */
int orc_dof_list(int nnpe, const int *conn, const int *dof_first, const int *ndof, int *map);

/* ---- serial assembler drivers (SerialAssembler order) ---- */
void orc_specify_mesh(orc_eq *eq, int n_el, int nnpe, const int *conn,
                      const int *dof_first, const int *ndof);
int  orc_assemble_elastic3d(orc_eq *eq, int n_el, const int *conn, const double *xyz,
                            const int *dof_first, const int *ndof,
                            const double *cmats, const int *elem_group, int order);
int  orc_assemble_elastic2d(orc_eq *eq, int n_el, const int *conn, const double *xy,
                            const int *dof_first, const int *ndof,
                            const double *cmats, const int *elem_group, int order,
                            double thickness);
int  orc_assemble_heat_transient(orc_eq *eq, int n_el, const int *conn, const double *xyz,
                                 const int *dof_first, const int *ndof,
                                 const double *kcap, const int *elem_group, int order,
                                 double alpha, double dt, const double *u_s);

/* ---- J2 / Hill plasticity (plastic.cpp) ---- */
typedef struct {
    double E, nu, sigma0, H;       /* linear isotropic hardening: m_E=H, m_n=1 */
} orc_j2_params;
int orc_j2_stress_increment(const orc_j2_params *p, const double *stress, double eps_bar,
                            const double *dstrain, double *dstress, double *deps_bar,
                            double *Cep36, int *plastic);
/* Same update; Dalg36 = exact derivative d(dstress)/d(dstrain) of the loop as executed. */
int orc_j2_stress_increment_alg(const orc_j2_params *p, const double *stress, double eps_bar,
                                const double *dstrain, double *dstress, double *deps_bar,
                                double *Cep36, double *Dalg36, int *plastic);
#define ORC_TANGENT_CONTINUUM   0   /* getTangentStiffness (plastic.cpp:583-610) */
#define ORC_TANGENT_ALGORITHMIC 1   /* symmetric part of the return map's derivative */

/* Plastic hex8 element (plasticElement.cpp:39-250).  Per Gauss point record
 * of ORC_PISV doubles: stress[6], eps_bar, dstress[6], deps_bar, plastic,
 * return-map iterations, Cep[36]. */
#define ORC_PISV 52
int  orc_plastic_update(const double *xyz24, const orc_j2_params *p, int order, const double *ue24,
                        double *isv);                                            /* "V" */
int  orc_plastic_update_tangent(const double *xyz24, const orc_j2_params *p, int order, const double *ue24,
                                double *isv, int tangent);
int  orc_plastic_ke_fe(const double *xyz24, const orc_j2_params *p, int order, const double *isv,
                       double *Ke, double *Fe);                                  /* "KI" + "F" */
void orc_plastic_commit(double *isv, int nip);                                   /* "U" */

#ifdef __cplusplus
}
#endif
#endif
