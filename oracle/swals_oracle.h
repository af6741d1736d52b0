/*
 * CPU restatement of the SWALS time-stepping hot path (FP64, serial, plain C).
 *
 * TEST INFRASTRUCTURE ONLY: this is the parity checker. Only tests/, the smoke()
 * entry point and bench.py's cpu_baseline leg load it; the product (ptha_amd/)
 * never links or calls it.
 *
 * Arrays follow the reference's Fortran layout: U(nx, ny, 4) column-major with the
 * i (x) index fastest and planes [STG, UH, VH, ELV]
 * (propagation/SWALS/src/shallow_water/domain_mod.f90:79).
 */
#ifndef SWALS_ORACLE_H
#define SWALS_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* timestepping_method ids: index into the reference's timestepping_metadata table
 * (timestepping_metadata_mod.f90:69-140). */
enum { OR_RK2 = 1, OR_RK2N = 2, OR_MIDPOINT = 3, OR_EULER = 4, OR_LINEAR = 5, OR_LLPNF = 6, OR_LEAPFROG_NONLINEAR = 8 };

typedef struct or_domain or_domain;

or_domain *or_domain_create(int nx, int ny, const double lw[2], const double ll[2],
                            int method, int spherical, int coriolis, int truely_linear,
                            double msl_linear, double linear_friction_coeff,
                            int friction_type, double cfl, double theta);
void or_domain_destroy(or_domain *d);
void or_domain_set_U(or_domain *d, const double *U);
void or_domain_get_U(const or_domain *d, double *U);
void or_domain_set_manning_squared(or_domain *d, const double *m);
void or_domain_get_metrics(const or_domain *d, double *x, double *y, double *dbe, double *dle,
                           double *area, double *cor, double *cor_be, double *tanlat);
double or_domain_evolve(or_domain *d, double dt, int nsteps);
void or_domain_get_scalars(const or_domain *d, double *out);
void or_domain_set_time(or_domain *d, double t);
double or_domain_volume_interior(const or_domain *d);
double or_domain_stationary_timestep_max(const or_domain *d);
void or_domain_get_max_U(const or_domain *d, int k, double *out);
void or_domain_set_stage_bc(or_domain *d, double amp, double period, double msl, const int sides[4]);
/* Flather boundary (boundary_mod.f90:502-720) on sides [N, E, S, W]; use_fun: outside
 * state from the sine stage msl + amp*sin(2*pi*t/period) instead of the sea at rest. */
void or_domain_set_flather_bc(or_domain *d, int use_fun, double amp, double period, double msl, const int sides[4]);
/* Other boundary subroutines of boundary_mod.f90 on sides [N, E, S, W]: kind 3
 * transmissive_boundary, 4 boundary_stage_radiation_momentum, 5
 * boundary_stage_transmissive_momentum (use_fun: the sine stage function, else
 * msl_linear; 5 needs it). */
void or_domain_set_bc(or_domain *d, int kind, int use_fun, double amp, double period, double msl, const int sides[4]);
void or_domain_set_record_max_U(or_domain *d, int on);
void or_domain_get_gauges(const or_domain *d, int n, const int *ij, double *out);

/* Dispersion on staggered grids (linear_dispersive_solver_mod.f90, default form):
 * setup (td1, td2, tridiagonal_inner_iter), store_last_U, the extrapolation level
 * append (U(UH:VH) at the domain time), and solve_dispersive_terms. */
void or_domain_dispersion_setup(or_domain *d, double td1, double td2, int niter);
void or_domain_dispersion_store(or_domain *d);
void or_domain_dispersion_append(or_domain *d);
void or_domain_dispersion_solve(or_domain *d, int rhs_up, int estimate, double ftime);

/* Individual pieces, exposed for unit tests of the restatement itself. */
void or_mc_limited_gradient(const double *u_local, const double *u_lower, const double *u_upper,
                            const double *theta, double *grad, int n);
double or_area_lonlat_rectangle(double lon1, double lat1, double dlon, double dlat, int flat);

#ifdef __cplusplus
}
#endif
#endif
