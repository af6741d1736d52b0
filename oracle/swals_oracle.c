/*
 * CPU restatement of the SWALS time-stepping hot path (FP64, serial).
 *
 * TEST INFRASTRUCTURE ONLY (the parity checker): loaded by tests/, smoke() and
 * bench.py's cpu_baseline leg; never by the product.
 *
 * Each routine cites the reference file:line whose semantics it restates, with
 * paths relative to propagation/SWALS/src/shallow_water/. Arithmetic is written in
 * the same evaluation order as the reference expressions, and the build uses
 * -ffp-contract=off, so on the same inputs the restatement reproduces the compiled
 * reference to the last bit or a few ulps (libm differences only).
 */
#include "swals_oracle.h"
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* global_mod.f90:37-44 */
static const double GRAVITY = 9.8;
static const double RADIUS_EARTH = 6371000.0;
static const double MIN_DEPTH = 1.0e-05;
static const double MAXIMUM_TIMESTEP = 1.0e+20;
/* spherical_mod.f90:15 */
static const double EARTH_ANGULAR_FREQ = 7.292115e-05;
/* domain_mod.f90:111-117 */
static const double LIMITER_COEF1 = 0.1, LIMITER_COEF2 = 0.3, LIMITER_COEF3 = 100.0;

enum { STG = 0, UH = 1, VH = 2, ELV = 3 };
enum { MX_STAGE = 0, MX_SPEED, MX_FLUX, MX_ARRIVAL, MX_TOMS, MX_MIN_STAGE, N_MAXQ };

struct or_domain {
    int nx, ny, method, spherical, coriolis_on, truely_linear, friction_type;
    int xL, xU, yL, yU, n_ext;
    int boundary_exterior[4];
    double lw[2], ll[2], dx[2];
    double msl_linear, linear_friction_coeff, cfl, theta, maximum_timestep;
    double *x, *y, *dbe, *dle, *area, *cor, *cor_be, *tanlat;
    double *U, *manning_sq, *friction_work; /* friction_work: 2 planes (UH, VH) */
    int friction_work_is_setup;
    /* finite-volume work arrays */
    double *depth, *vel, *flux_NS, *flux_EW, *es, *es_vh_jm1, *backup_U;
    /* scalars */
    double time, max_dt, evolve_step_dt, dt_last_update;
    double boundary_flux_store[4], boundary_flux_evolve_integral, boundary_flux_time_integral;
    double negative_depth_fix_counter;
    int nsteps_advanced;
    /* max quantities */
    int record_max_U;
    double *max_U;
    double arrival_stage_threshold;
    /* boundary: bc_on 1 = transmissive stage boundary, 2 = Flather; bc_fun: the Flather
     * outside state comes from the sine stage function (else msl_linear at rest) */
    int bc_on, bc_fun;
    double bc_amp, bc_period, bc_msl;
    /* dispersion (staggered grids): RHS (UH, VH), last_U (UH, VH, ELV), Ax (UH, VH),
     * the three extrapolation levels (yn2, yn1, y0; UH and VH each) and their times */
    int disp_on, disp_iter;
    double td1, td2;
    double *dRHS, *dlast, *dAx, *qy[3], qt[3];
};

#define IX(d, i, j) ((size_t)((j) - 1) * (size_t)(d)->nx + (size_t)((i) - 1))
#define P(d, k) ((d)->U + (size_t)(k) * (size_t)(d)->nx * (size_t)(d)->ny)
#define A(d, k, i, j) (P(d, k)[IX(d, i, j)])

static double dmax(double a, double b) { return a > b ? a : b; }
static double dmin(double a, double b) { return a < b ? a : b; }

/* spherical_mod.f90:23-64 */
double or_area_lonlat_rectangle(double lon1, double lat1, double dlon, double dlat, int flat)
{
    const double deg2rad = acos(-1.0) / 180.0;
    if (flat) {
        return RADIUS_EARTH * RADIUS_EARTH * cos((lat1 + 0.5 * dlat) * deg2rad) * fabs(dlat) * deg2rad *
               fmod(fabs(dlon), 360.0) * deg2rad;
    }
    double s1 = sin(lat1 * deg2rad), s2 = sin((lat1 + dlat) * deg2rad);
    return RADIUS_EARTH * RADIUS_EARTH * fabs(s1 - s2) * fmod(fabs(dlon), 360.0) * deg2rad;
}

/* Metric arrays: domain_mod.f90:918-1014 (x/y centres, cos/tan lat, Coriolis, edge
 * lengths, flat spherical cell areas). */
static void build_metrics(or_domain *d)
{
    const int nx = d->nx, ny = d->ny;
    const double deg2rad = acos(-1.0) / 180.0;
    d->dx[0] = d->lw[0] / (nx * 1.0);
    d->dx[1] = d->lw[1] / (ny * 1.0);
    for (int i = 1; i <= nx; i++) d->x[i - 1] = d->ll[0] + ((i - 0.5) / (nx * 1.0)) * d->lw[0];
    for (int i = 1; i <= ny; i++) d->y[i - 1] = d->ll[1] + ((i - 0.5) / (ny * 1.0)) * d->lw[1];
    if (d->spherical) {
        double *coslat_be = calloc((size_t)ny + 1, sizeof(double));
        for (int j = 0; j < ny; j++) coslat_be[j] = cos((d->y[j] - 0.5 * d->dx[1]) * deg2rad);
        coslat_be[ny] = cos((d->y[ny - 1] + 0.5 * d->dx[1]) * deg2rad);
        for (int j = 0; j < ny; j++) d->tanlat[j] = tan(d->y[j] * deg2rad) / RADIUS_EARTH;
        if (d->coriolis_on) {
            for (int j = 0; j < ny; j++) d->cor[j] = 2.0 * sin(d->y[j] * deg2rad) * EARTH_ANGULAR_FREQ;
            for (int j = 0; j < ny; j++)
                d->cor_be[j] = 2.0 * EARTH_ANGULAR_FREQ * sin((d->y[j] - 0.5 * d->dx[1]) * deg2rad);
            d->cor_be[ny] = 2.0 * EARTH_ANGULAR_FREQ * sin((d->y[ny - 1] + 0.5 * d->dx[1]) * deg2rad);
        }
        for (int i = 0; i <= nx; i++) d->dle[i] = d->dx[1] * deg2rad * RADIUS_EARTH;
        for (int j = 0; j <= ny; j++) d->dbe[j] = d->dx[0] * deg2rad * RADIUS_EARTH * coslat_be[j];
        for (int j = 0; j < ny; j++)
            d->area[j] = or_area_lonlat_rectangle(d->x[0] - d->dx[0] * 0.5, d->y[j] - d->dx[1] * 0.5, d->dx[0],
                                                  d->dx[1], 1);
        free(coslat_be);
    } else {
        for (int i = 0; i <= nx; i++) d->dle[i] = d->dx[1];
        for (int j = 0; j <= ny; j++) d->dbe[j] = d->dx[0];
        for (int j = 0; j < ny; j++) d->area[j] = d->dx[0] * d->dx[1];
    }
}

or_domain *or_domain_create(int nx, int ny, const double lw[2], const double ll[2], int method, int spherical,
                            int coriolis, int truely_linear, double msl_linear, double linear_friction_coeff,
                            int friction_type, double cfl, double theta)
{
    or_domain *d = calloc(1, sizeof(or_domain));
    size_t n = (size_t)nx * ny;
    d->nx = nx; d->ny = ny; d->method = method;
    d->spherical = spherical; d->coriolis_on = spherical && coriolis;
    d->truely_linear = truely_linear; d->msl_linear = msl_linear;
    d->linear_friction_coeff = linear_friction_coeff; d->friction_type = friction_type;
    d->lw[0] = lw[0]; d->lw[1] = lw[1]; d->ll[0] = ll[0]; d->ll[1] = ll[1];
    /* Default CFL / theta from timestepping_metadata_mod.f90:73-121 */
    const int fv2 = method == OR_RK2 || method == OR_RK2N || method == OR_MIDPOINT;
    double def_cfl = fv2 ? 0.99 : (method == OR_EULER ? 0.9 : 0.7);
    double def_theta = fv2 ? 1.6 : (method == OR_EULER ? 0.9 : -HUGE_VAL);
    d->cfl = cfl > 0 ? cfl : def_cfl;
    d->theta = theta > 0 ? theta : def_theta;
    d->maximum_timestep = MAXIMUM_TIMESTEP;
    d->xL = 1; d->xU = nx; d->yL = 1; d->yU = ny; d->n_ext = 2;
    for (int k = 0; k < 4; k++) d->boundary_exterior[k] = 1;
    d->x = calloc(nx, sizeof(double)); d->y = calloc(ny, sizeof(double));
    d->dbe = calloc(ny + 1, sizeof(double)); d->dle = calloc(nx + 1, sizeof(double));
    d->area = calloc(ny, sizeof(double)); d->cor = calloc(ny, sizeof(double));
    d->cor_be = calloc(ny + 1, sizeof(double)); d->tanlat = calloc(ny, sizeof(double));
    d->U = calloc(4 * n, sizeof(double));
    d->manning_sq = calloc(n, sizeof(double));
    d->friction_work = calloc(2 * n, sizeof(double));
    d->depth = calloc(n, sizeof(double)); d->vel = calloc(2 * n, sizeof(double));
    d->flux_NS = calloc(3 * (size_t)nx * (ny + 1), sizeof(double));
    d->flux_EW = calloc(3 * (size_t)(nx + 1) * ny, sizeof(double));
    d->es = calloc(2 * n, sizeof(double));
    d->es_vh_jm1 = calloc((size_t)nx * (ny + 1), sizeof(double));
    d->backup_U = calloc(3 * n, sizeof(double));
    d->max_U = calloc(N_MAXQ * n, sizeof(double));
    d->record_max_U = 1;
    d->arrival_stage_threshold = 0.01; /* domain_mod.f90:434 default */
    /* max_U initial values: domain_mod.f90:1119-1136 (computed in float32) */
    float big = 0.99f * FLT_MAX;
    for (size_t k = 0; k < N_MAXQ * n; k++) d->max_U[k] = -(double)big;
    for (size_t k = 0; k < n; k++) d->max_U[MX_MIN_STAGE * n + k] = (double)big;
    build_metrics(d);
    return d;
}

void or_domain_destroy(or_domain *d)
{
    if (!d) return;
    free(d->x); free(d->y); free(d->dbe); free(d->dle); free(d->area); free(d->cor); free(d->cor_be);
    free(d->tanlat); free(d->U); free(d->manning_sq); free(d->friction_work); free(d->depth); free(d->vel);
    free(d->flux_NS); free(d->flux_EW); free(d->es); free(d->es_vh_jm1); free(d->backup_U); free(d->max_U);
    free(d);
}

void or_domain_set_U(or_domain *d, const double *U) { memcpy(d->U, U, 4 * sizeof(double) * d->nx * d->ny); }
void or_domain_get_U(const or_domain *d, double *U) { memcpy(U, d->U, 4 * sizeof(double) * d->nx * d->ny); }
void or_domain_set_manning_squared(or_domain *d, const double *m)
{
    memcpy(d->manning_sq, m, sizeof(double) * d->nx * d->ny);
    d->friction_work_is_setup = 0;
}
void or_domain_set_time(or_domain *d, double t) { d->time = t; }
void or_domain_set_record_max_U(or_domain *d, int on) { d->record_max_U = on; }

void or_domain_get_metrics(const or_domain *d, double *x, double *y, double *dbe, double *dle, double *area,
                           double *cor, double *cor_be, double *tanlat)
{
    memcpy(x, d->x, sizeof(double) * d->nx);
    memcpy(y, d->y, sizeof(double) * d->ny);
    memcpy(dbe, d->dbe, sizeof(double) * (d->ny + 1));
    memcpy(dle, d->dle, sizeof(double) * (d->nx + 1));
    memcpy(area, d->area, sizeof(double) * d->ny);
    memcpy(cor, d->cor, sizeof(double) * d->ny);
    memcpy(cor_be, d->cor_be, sizeof(double) * (d->ny + 1));
    if (d->spherical) memcpy(tanlat, d->tanlat, sizeof(double) * d->ny);
    else memset(tanlat, 0, sizeof(double) * d->ny);
}

/* boundary_mod.f90:78-164 (boundary_stage_transmissive_normal_momentum) with the
 * stage function msl + amp*sin(2*pi*t/period) and elevation left unchanged. */
/* Outside state of the Flather boundary at cell (i, j) (boundary_mod.f90:528-545):
 * without a boundary function, the sea at rest at max(msl_linear, elevation); with
 * one, its stage and momenta (here the sine stage, zero momenta, the cell's
 * elevation), velocities zero where the outside is dry. */
static void flather_outside(const or_domain *d, int i, int j, double *stage_o, double *u_o, double *v_o,
                            double *depth_o)
{
    if (!d->bc_fun) {
        *stage_o = dmax(d->msl_linear, A(d, ELV, i, j));
        *u_o = 0.0;
        *v_o = 0.0;
        *depth_o = *stage_o - A(d, ELV, i, j);
    } else {
        const double two_pi = 2.0 * acos(-1.0);
        const double st = d->bc_msl + d->bc_amp * sin(two_pi * d->time / d->bc_period);
        const double uh_o = 0.0, vh_o = 0.0;
        *stage_o = st;
        *depth_o = st - A(d, ELV, i, j);
        if (*depth_o <= MIN_DEPTH) {
            *u_o = 0.0;
            *v_o = 0.0;
        } else {
            *u_o = uh_o / *depth_o;
            *v_o = vh_o / *depth_o;
        }
    }
}

/* One Flather edge cell (boundary_mod.f90:547-569 and its three mirror images).
 * (b, ib, jb): boundary cell; (ii, ji): the cell inside it; sgn = -1 on the west and
 * south edges (the characteristic entering from outside moves in +x / +y), +1 on
 * the east and north; normal = UH on the west/east edges, VH on the south/north. */
static void flather_cell(or_domain *d, int ib, int jb, int ii, int ji, double sgn, int normal)
{
    const int tang = normal == UH ? VH : UH;
    double stage_o, u_o, v_o, depth_o;
    flather_outside(d, ib, jb, &stage_o, &u_o, &v_o, &depth_o);
    const double vel_o = normal == UH ? u_o : v_o;
    const double depth_inside = A(d, STG, ii, ji) - A(d, ELV, ii, ji);
    if ((depth_inside > MIN_DEPTH) && (depth_o > MIN_DEPTH)) {
        const double depth_inside_inv = 1.0 / depth_inside;
        const double sqrt_g_on_di = sqrt(GRAVITY * depth_inside_inv);
        const double qn = A(d, normal, ii, ji);
        double w1, w2, w3, stage;
        if (sgn < 0) {
            w1 = vel_o + sqrt_g_on_di * stage_o;
            w2 = (qn < 0.0) ? A(d, tang, ii, ji) * depth_inside_inv : 0.0;
            w3 = qn * depth_inside_inv - sqrt_g_on_di * A(d, STG, ii, ji);
            stage = (-w3 + w1) / (2.0 * sqrt_g_on_di);
        } else {
            w1 = vel_o - sqrt_g_on_di * stage_o;
            w2 = (qn > 0.0) ? A(d, tang, ii, ji) * depth_inside_inv : 0.0;
            w3 = qn * depth_inside_inv + sqrt_g_on_di * A(d, STG, ii, ji);
            stage = (w3 - w1) / (2.0 * sqrt_g_on_di);
        }
        A(d, STG, ib, jb) = stage;
        A(d, normal, ib, jb) = (w3 + w1) * 0.5 * depth_inside;
        A(d, tang, ib, jb) = w2 * depth_inside;
    } else {
        A(d, STG, ib, jb) = stage_o;
        A(d, UH, ib, jb) = 0.0;
        A(d, VH, ib, jb) = 0.0;
    }
}

/* boundary_mod.f90:502-720 (flather_boundary): west and east columns for every row,
 * then south and north rows for every column (so the corners end with the
 * south/north treatment, reading the west/east results next to them). */
static void flather_boundary(or_domain *d)
{
    const int nx = d->nx, ny = d->ny;
    for (int j = 1; j <= ny; j++) {
        if (d->boundary_exterior[3]) flather_cell(d, 1, j, 2, j, -1.0, UH);
        if (d->boundary_exterior[1]) flather_cell(d, nx, j, nx - 1, j, 1.0, UH);
    }
    for (int i = 1; i <= nx; i++) {
        if (d->boundary_exterior[2]) flather_cell(d, i, 1, i, 2, -1.0, VH);
        if (d->boundary_exterior[0]) flather_cell(d, i, ny, i, ny - 1, 1.0, VH);
    }
}

/* boundary_mod.f90:23-75 (transmissive_boundary): every variable, elevation included,
 * copied from the cell inside; west/east columns, then south/north rows */
static void transmissive_boundary(or_domain *d)
{
    const int nx = d->nx, ny = d->ny;
    for (int k = 0; k < 4; k++)
        for (int j = 1; j <= ny; j++) {
            if (d->boundary_exterior[3]) A(d, k, 1, j) = A(d, k, 2, j);
            if (d->boundary_exterior[1]) A(d, k, nx, j) = A(d, k, nx - 1, j);
        }
    for (int k = 0; k < 4; k++)
        for (int i = 1; i <= nx; i++) {
            if (d->boundary_exterior[2]) A(d, k, i, 1) = A(d, k, i, 2);
            if (d->boundary_exterior[0]) A(d, k, i, ny) = A(d, k, i, ny - 1);
        }
}

/* boundary values [stage, elevation] at (i, j): the sine stage function with the cell's
 * elevation, or (radiation without a function) msl_linear and the cell's elevation */
static void bc_values(const or_domain *d, int i, int j, double *stage, double *elev)
{
    if (d->bc_fun) {
        const double two_pi = 2.0 * acos(-1.0);
        *stage = d->bc_msl + d->bc_amp * sin(two_pi * d->time / d->bc_period);
    } else {
        *stage = d->msl_linear;
    }
    *elev = A(d, ELV, i, j);
}

/* boundary_mod.f90:166-318 (boundary_stage_radiation_momentum, sgn = -1 west/south) and
 * :321-405 (boundary_stage_transmissive_momentum); staggered grids set stage and
 * elevation only */
static void stage_momentum_cell(or_domain *d, int ib, int jb, int ii, int ji, double sgn, int normal, int radiation,
                                int stag)
{
    const int tang = normal == UH ? VH : UH;
    double st, el;
    bc_values(d, ib, jb, &st, &el);
    A(d, STG, ib, jb) = st;
    A(d, ELV, ib, jb) = el;
    if (stag) return;
    if (radiation) {
        const double local_h = dmax(st - el, 0.0);
        const double q = (sqrt(GRAVITY * dmax(A(d, STG, ii, ji) - el, 0.0)) - sqrt(GRAVITY * local_h)) * local_h;
        A(d, normal, ib, jb) = sgn < 0 ? -q : q;
        A(d, tang, ib, jb) = A(d, tang, ii, ji);
    } else {
        A(d, normal, ib, jb) = A(d, normal, ii, ji);
        A(d, tang, ib, jb) = A(d, tang, ii, ji);
    }
}

static void stage_momentum_boundary(or_domain *d, int radiation)
{
    const int nx = d->nx, ny = d->ny;
    const int stag = (d->method == OR_LINEAR || d->method == OR_LLPNF || d->method == OR_LEAPFROG_NONLINEAR);
    for (int j = 1; j <= ny; j++) {
        if (d->boundary_exterior[3]) stage_momentum_cell(d, 1, j, 2, j, -1.0, UH, radiation, stag);
        if (d->boundary_exterior[1]) stage_momentum_cell(d, nx, j, nx - 1, j, 1.0, UH, radiation, stag);
    }
    for (int i = 1; i <= nx; i++) {
        if (d->boundary_exterior[2]) stage_momentum_cell(d, i, 1, i, 2, -1.0, VH, radiation, stag);
        if (d->boundary_exterior[0]) stage_momentum_cell(d, i, ny, i, ny - 1, 1.0, VH, radiation, stag);
    }
}

static void update_boundary(or_domain *d)
{
    if (!d->bc_on) return;
    if (d->bc_on == 2) {
        flather_boundary(d);
        return;
    }
    if (d->bc_on == 3) { transmissive_boundary(d); return; }
    if (d->bc_on == 4) { stage_momentum_boundary(d, 1); return; }
    if (d->bc_on == 5) { stage_momentum_boundary(d, 0); return; }
    const int nx = d->nx, ny = d->ny;
    const int stag = (d->method == OR_LINEAR || d->method == OR_LLPNF || d->method == OR_LEAPFROG_NONLINEAR);
    const double two_pi = 2.0 * acos(-1.0);
    double stage = d->bc_msl + d->bc_amp * sin(two_pi * d->time / d->bc_period);
    for (int j = 1; j <= ny; j++) {
        if (d->boundary_exterior[3]) {
            int i = 1;
            A(d, STG, i, j) = stage;
            if (!stag) { A(d, UH, i, j) = A(d, UH, i + 1, j); A(d, VH, i, j) = 0.0; }
        }
        if (d->boundary_exterior[1]) {
            int i = nx;
            A(d, STG, i, j) = stage;
            if (!stag) { A(d, UH, i, j) = A(d, UH, i - 1, j); A(d, VH, i, j) = 0.0; }
        }
    }
    for (int i = 1; i <= nx; i++) {
        if (d->boundary_exterior[2]) {
            int j = 1;
            A(d, STG, i, j) = stage;
            if (!stag) { A(d, VH, i, j) = A(d, VH, i, j + 1); A(d, UH, i, j) = 0.0; }
        }
        if (d->boundary_exterior[0]) {
            int j = ny;
            A(d, STG, i, j) = stage;
            if (!stag) { A(d, VH, i, j) = A(d, VH, i, j - 1); A(d, UH, i, j) = 0.0; }
        }
    }
}

void or_domain_set_stage_bc(or_domain *d, double amp, double period, double msl, const int sides[4])
{
    d->bc_on = 1; d->bc_fun = 1; d->bc_amp = amp; d->bc_period = period; d->bc_msl = msl;
    for (int k = 0; k < 4; k++) d->boundary_exterior[k] = sides[k] != 0;
}

void or_domain_set_bc(or_domain *d, int kind, int use_fun, double amp, double period, double msl, const int sides[4])
{
    d->bc_on = kind; d->bc_fun = use_fun != 0; d->bc_amp = amp; d->bc_period = period; d->bc_msl = msl;
    for (int k = 0; k < 4; k++) d->boundary_exterior[k] = sides[k] != 0;
}

void or_domain_set_flather_bc(or_domain *d, int use_fun, double amp, double period, double msl, const int sides[4])
{
    d->bc_on = 2; d->bc_fun = use_fun != 0; d->bc_amp = amp; d->bc_period = period; d->bc_msl = msl;
    for (int k = 0; k < 4; k++) d->boundary_exterior[k] = sides[k] != 0;
}

/* domain_mod_leapfrog_solver_friction_include.f90:21-85 (precompute_friction_work) */
static void precompute_friction_work(or_domain *d, int truely_linear)
{
    const int nx = d->nx, ny = d->ny;
    const double power = (d->friction_type == 1) ? -2.0 : -7.0 / 3.0;
    double *fU = d->friction_work, *fV = d->friction_work + (size_t)nx * ny;
    const double msl = d->msl_linear;
    for (int j = 1; j <= ny; j++) {
        for (int i = 1; i <= nx; i++) {
            int ip1 = i + 1 < nx ? i + 1 : nx, jp1 = j + 1 < ny ? j + 1 : ny;
            double depth_iph, depth_jph;
            if (truely_linear) {
                depth_iph = 0.5 * (msl - A(d, ELV, i, j) + msl - A(d, ELV, ip1, j));
                depth_jph = 0.5 * (msl - A(d, ELV, i, j) + msl - A(d, ELV, i, jp1));
            } else {
                depth_iph = 0.5 * (A(d, STG, i, j) - A(d, ELV, i, j) + A(d, STG, ip1, j) - A(d, ELV, ip1, j));
                depth_jph = 0.5 * (A(d, STG, i, j) - A(d, ELV, i, j) + A(d, STG, i, jp1) - A(d, ELV, i, jp1));
            }
            depth_iph = dmax(depth_iph, MIN_DEPTH);
            depth_jph = dmax(depth_jph, MIN_DEPTH);
            double nsq_iph = 0.5 * (d->manning_sq[IX(d, i, j)] + d->manning_sq[IX(d, ip1, j)]);
            double nsq_jph = 0.5 * (d->manning_sq[IX(d, i, j)] + d->manning_sq[IX(d, i, jp1)]);
            fU[IX(d, i, j)] = GRAVITY * nsq_iph * pow(depth_iph, power);
            fV[IX(d, i, j)] = GRAVITY * nsq_jph * pow(depth_jph, power);
        }
    }
    if (truely_linear) d->friction_work_is_setup = 1;
}

/* Staggered-grid leapfrog step: domain_mod_linear_solver_include.f90:59-571, with the
 * four variants of domain_mod_timestepping_alternatives_include.f90:339-434. */
static void lf_stage_boundary(or_domain *d, double dt)
{
    const int nx = d->nx, ny = d->ny;
    const int xL = d->xL, xU = d->xU, yL = d->yL, yU = d->yU;
    /* Stage half-step: include.f90:92-113 */
    for (int j = yL + 1; j <= yU - 1; j++) {
        double inv_cell_area_dt = dt / d->area[j - 1];
        for (int i = xL + 1; i <= xU - 1; i++) {
            A(d, STG, i, j) = A(d, STG, i, j) -
                              inv_cell_area_dt * ((A(d, UH, i, j) - A(d, UH, i - 1, j)) * d->dle[i - 1] +
                                                  (A(d, VH, i, j) * d->dbe[j] - A(d, VH, i, j - 1) * d->dbe[j - 1]));
        }
    }
    d->time = d->time + dt * 0.5;
    update_boundary(d);

    /* Boundary flux integration: include.f90:126-205 (single domain => exterior == store) */
    {
        const int ne = d->n_ext;
        double s;
        s = 0.0; for (int i = ne + 1; i <= nx - ne; i++) s += A(d, VH, i, ny - ne);
        d->boundary_flux_store[0] = s * d->dbe[ny - ne];
        s = 0.0; for (int j = 1 + ne; j <= ny - ne; j++) s += A(d, UH, nx - ne, j);
        d->boundary_flux_store[1] = s * d->dle[nx - ne];
        s = 0.0; for (int i = ne + 1; i <= nx - ne; i++) s += A(d, VH, i, ne);
        d->boundary_flux_store[2] = -s * d->dbe[ne];
        s = 0.0; for (int j = 1 + ne; j <= ny - ne; j++) s += A(d, UH, ne, j);
        d->boundary_flux_store[3] = -s * d->dle[ne];
        d->boundary_flux_evolve_integral =
            (((d->boundary_flux_store[0] + d->boundary_flux_store[1]) + d->boundary_flux_store[2]) +
             d->boundary_flux_store[3]) * dt;
    }
}

static void leapfrog_step(or_domain *d, double dt)
{
    const int nx = d->nx, ny = d->ny;
    const int xL = d->xL, xU = d->xU, yL = d->yL, yU = d->yU;
    const int friction = (d->method == OR_LLPNF);
    const int cor_or_fric = d->coriolis_on || friction;
    const double msl = d->msl_linear;
    d->dt_last_update = dt;
    lf_stage_boundary(d, dt);

    if (friction && !d->friction_work_is_setup) precompute_friction_work(d, d->truely_linear);
    const double *fwU = d->friction_work, *fwV = d->friction_work + (size_t)nx * ny;

    /* Momentum half-step: include.f90:211-571, serial loop order yL..yU-1. */
    double *h_jph = calloc(nx + 1, sizeof(double)), *h_iph = calloc(nx + 1, sizeof(double));
    double *wet_jph = calloc(nx + 1, sizeof(double)), *wet_iph = calloc(nx + 1, sizeof(double));
    double *dw_j = calloc(nx + 1, sizeof(double));
    double *vh_iph_jmh = calloc(nx + 1, sizeof(double)), *vh_iph_jph = calloc(nx + 1, sizeof(double));
    double *uh_i_j = calloc(nx + 1, sizeof(double)), *uh_i_jp1 = calloc(nx + 1, sizeof(double));
    double *fm_UH = calloc(nx + 1, sizeof(double)), *fm_VH = calloc(nx + 1, sizeof(double));
    const int yl = yL, yu = yU - 1;
    if (cor_or_fric) {
        /* include.f90:295-331: 'old' values bounding the j-loop */
        int jm = (yl == 1) ? 1 : yl - 1;
        for (int i = xL; i <= xU - 1; i++) vh_iph_jmh[i] = 0.5 * (A(d, VH, i, jm) + A(d, VH, i + 1, jm));
        for (int i = xL + 1; i <= xU - 1; i++) uh_i_j[i] = 0.5 * (A(d, UH, i, yl) + A(d, UH, i - 1, yl));
        uh_i_j[xL] = 0.5 * (A(d, UH, xL, yl) + A(d, UH, xL - 1 > 1 ? xL - 1 : 1, yl));
    }
    for (int j = yl; j <= yu; j++) {
        double inv_cell_area_dt_g = GRAVITY * dt / d->area[j - 1];
        double inv_cell_area_dt_vh_g = GRAVITY * dt / (0.5 * (d->area[j - 1] + d->area[j]));
        for (int i = xL; i <= xU - 1; i++) dw_j[i] = A(d, STG, i, j + 1) - A(d, STG, i, j);
        if (d->truely_linear) {
            for (int i = xL; i <= xU; i++) {
                int c = (A(d, ELV, i, j + 1) < -MIN_DEPTH + msl) && (A(d, ELV, i, j) < -MIN_DEPTH + msl);
                h_jph[i] = c ? msl - 0.5 * (A(d, ELV, i, j + 1) + A(d, ELV, i, j)) : 0.0;
                wet_jph[i] = 1.0;
            }
            for (int i = xL; i <= xU - 1; i++) {
                int c = (A(d, ELV, i, j) < -MIN_DEPTH + msl) && (A(d, ELV, i + 1, j) < -MIN_DEPTH + msl);
                h_iph[i] = c ? msl - 0.5 * (A(d, ELV, i + 1, j) + A(d, ELV, i, j)) : 0.0;
                wet_iph[i] = 1.0;
            }
        } else {
            for (int i = xL; i <= xU; i++) {
                int c = (A(d, STG, i, j + 1) - A(d, ELV, i, j + 1) > MIN_DEPTH) &&
                        (A(d, STG, i, j) - A(d, ELV, i, j) > MIN_DEPTH);
                h_jph[i] = c ? 0.5 * ((A(d, STG, i, j + 1) + A(d, STG, i, j)) -
                                      (A(d, ELV, i, j + 1) + A(d, ELV, i, j)))
                             : 0.0;
                wet_jph[i] = h_jph[i] > 0.0 ? 1.0 : 0.0;
            }
            for (int i = xL; i <= xU - 1; i++) {
                int c = (A(d, STG, i, j) - A(d, ELV, i, j) > MIN_DEPTH) &&
                        (A(d, STG, i + 1, j) - A(d, ELV, i + 1, j) > MIN_DEPTH);
                h_iph[i] = c ? 0.5 * ((A(d, STG, i + 1, j) + A(d, STG, i, j)) -
                                      (A(d, ELV, i + 1, j) + A(d, ELV, i, j)))
                             : 0.0;
                wet_iph[i] = h_iph[i] > 0.0 ? 1.0 : 0.0;
            }
        }
        if (cor_or_fric) {
            /* include.f90:408-430 */
            for (int i = xL; i <= xU - 1; i++) vh_iph_jph[i] = 0.5 * (A(d, VH, i, j) + A(d, VH, i + 1, j));
            for (int i = xL + 1; i <= xU - 1; i++) uh_i_jp1[i] = 0.5 * (A(d, UH, i, j + 1) + A(d, UH, i - 1, j + 1));
            uh_i_jp1[xL] = 0.5 * (A(d, UH, xL, j + 1) + A(d, UH, xL - 1 > 1 ? xL - 1 : 1, j + 1));
        }
        if (friction) {
            /* include.f90:486-503 (no ambient flux) */
            for (int i = xL; i <= xU - 1; i++) {
                double uh = A(d, UH, i, j), vh = A(d, VH, i, j);
                double vbar = 0.5 * (vh_iph_jmh[i] + vh_iph_jph[i]);
                double ubar = 0.5 * (uh_i_j[i] + uh_i_jp1[i]);
                fm_UH[i] = 1.0 / (1.0 + dt * d->linear_friction_coeff +
                                  dt * fwU[IX(d, i, j)] * (sqrt(uh * uh + vbar * vbar)));
                fm_VH[i] = 1.0 / (1.0 + dt * d->linear_friction_coeff +
                                  dt * fwV[IX(d, i, j)] * (sqrt(vh * vh + ubar * ubar)));
            }
        }
        for (int i = xL; i <= xU - 1; i++) {
            double uh_new = A(d, UH, i, j) * wet_iph[i] - inv_cell_area_dt_g * h_iph[i] *
                                                                (A(d, STG, i + 1, j) - A(d, STG, i, j)) * d->dle[i];
            double vh_new = A(d, VH, i, j) * wet_jph[i] - inv_cell_area_dt_vh_g * h_jph[i] * dw_j[i] * d->dbe[j];
            if (d->coriolis_on) {
                /* include.f90:432-439, 523-540 */
                double dhc = dt * d->cor[j - 1] * 0.5 * (h_iph[i] > 0.0 ? 1.0 : 0.0);
                double dhc_jph = dt * d->cor_be[j] * 0.5 * (h_jph[i] > 0.0 ? 1.0 : 0.0);
                uh_new = uh_new + dhc * (vh_iph_jmh[i] + vh_iph_jph[i]);
                vh_new = vh_new - dhc_jph * (uh_i_j[i] + uh_i_jp1[i]);
            }
            if (friction) {
                uh_new = fm_UH[i] * uh_new;
                vh_new = fm_VH[i] * vh_new;
            }
            A(d, UH, i, j) = uh_new;
            A(d, VH, i, j) = vh_new;
        }
        if (!friction && d->linear_friction_coeff != 0.0) {
            /* include.f90:550-557 */
            double f = 1.0 / (1.0 + dt * d->linear_friction_coeff);
            for (int i = xL; i <= xU - 1; i++) {
                A(d, UH, i, j) = f * A(d, UH, i, j);
                A(d, VH, i, j) = f * A(d, VH, i, j);
            }
        }
        if (cor_or_fric) {
            for (int i = xL; i <= xU - 1; i++) { vh_iph_jmh[i] = vh_iph_jph[i]; uh_i_j[i] = uh_i_jp1[i]; }
        }
    }
    d->time = d->time + 0.5 * dt;
    free(h_jph); free(h_iph); free(wet_jph); free(wet_iph); free(dw_j); free(vh_iph_jmh); free(vh_iph_jph);
    free(uh_i_j); free(uh_i_jp1); free(fm_UH); free(fm_VH);
}

/* domain_mod.f90:1360-1419 */
static void compute_depth_and_velocity(or_domain *d)
{
    const size_t n = (size_t)d->nx * d->ny;
    double *stg = P(d, STG), *uh = P(d, UH), *vh = P(d, VH), *elv = P(d, ELV);
    for (size_t k = 0; k < n; k++) {
        d->depth[k] = stg[k] - elv[k];
        if (d->depth[k] > MIN_DEPTH) {
            double depth_inv = 1.0 / d->depth[k];
            d->vel[k] = uh[k] * depth_inv;
            d->vel[n + k] = vh[k] * depth_inv;
        } else {
            if (d->depth[k] < 0.0) {
                d->depth[k] = 0.0;
                stg[k] = elv[k];
                d->negative_depth_fix_counter += 1.0;
            }
            d->vel[k] = 0.0; d->vel[n + k] = 0.0;
            uh[k] = 0.0; vh[k] = 0.0;
        }
    }
}

/* extrapolation_limiting_mod.f90:44-79 (limiter_type = 'MC') */
void or_mc_limited_gradient(const double *u_local, const double *u_lower, const double *u_upper,
                            const double *theta, double *grad, int n)
{
    for (int i = 0; i < n; i++) {
        double a = u_upper[i] - u_local[i];
        double b = u_local[i] - u_lower[i];
        double th = theta[i];
        double sa = copysign(1.0, a), sb = copysign(1.0, b);
        double half_sasb = 0.5 * (sa + sb);
        double dd = dmin(fabs(a), fabs(b)) * half_sasb * th;
        double e = 0.5 * (a + b);
        double c = (dd == 0.0) ? 0.0 : e;
        grad[i] = (dd > 0.0) ? dmin(c, dd) : dmax(c, dd);
    }
}

/* Wet-dry theta taper: domain_mod_compute_fluxes_DE1_inner_include.f90:751-757 / 816-822 */
static double theta_wd(const or_domain *d, double d0, double d1, double d2)
{
    double mindep = dmin(dmin(d0, d1), d2) - MIN_DEPTH;
    double maxdep = dmax(dmax(d0, d1), d2) + LIMITER_COEF3 * MIN_DEPTH;
    double theta_local = (1.0 / (LIMITER_COEF2 - LIMITER_COEF1)) * (mindep / maxdep - LIMITER_COEF1);
    return dmax(d->theta * dmin(1.0, theta_local), 0.0);
}

/* Row-j NS gradients: include.f90:735-795. g[0..3] = stage, depth, u, v (length nx, 0-based i-1). */
static void ns_gradients(const or_domain *d, int j, double *th, double *g[4])
{
    const int nx = d->nx, ny = d->ny;
    const size_t n = (size_t)nx * ny;
    if (j > 1 && j < ny) {
        for (int i = 1; i <= nx; i++)
            th[i - 1] = theta_wd(d, d->depth[IX(d, i, j - 1)], d->depth[IX(d, i, j)], d->depth[IX(d, i, j + 1)]);
        const double *src[4] = {P(d, STG), d->depth, d->vel, d->vel + n};
        for (int q = 0; q < 4; q++)
            or_mc_limited_gradient(src[q] + IX(d, 1, j), src[q] + IX(d, 1, j - 1), src[q] + IX(d, 1, j + 1), th, g[q],
                                   nx);
    } else {
        for (int i = 0; i < nx; i++) { th[i] = 0.0; for (int q = 0; q < 4; q++) g[q][i] = 0.0; }
    }
}

/* Row-j EW gradients: include.f90:800-862 (interior i = 2..nx-1, zero at i = 1 and nx) */
static void ew_gradients(const or_domain *d, int j, double *th, double *g[4])
{
    const int nx = d->nx;
    const size_t n = (size_t)nx * d->ny;
    for (int i = 2; i <= nx - 1; i++)
        th[i - 1] = theta_wd(d, d->depth[IX(d, i - 1, j)], d->depth[IX(d, i, j)], d->depth[IX(d, i + 1, j)]);
    const double *src[4] = {P(d, STG), d->depth, d->vel, d->vel + n};
    for (int q = 0; q < 4; q++) {
        if (nx > 2)
            or_mc_limited_gradient(src[q] + IX(d, 2, j), src[q] + IX(d, 1, j), src[q] + IX(d, 3, j), th + 1,
                                   g[q] + 1, nx - 2);
        g[q][0] = 0.0;
        g[q][nx - 1] = 0.0;
    }
}

/* compute_fluxes_DE1_low_fr_diffusion: domain_mod_compute_fluxes_DE1_inner_include.f90:78-727
 * (reduced_momentum_diffusion = .true., no eddy viscosity, flux-conservative pressure)
 * followed by the boundary flux sums of domain_mod.f90:1483-1499. Returns max_dt. */
static double compute_fluxes(or_domain *d)
{
    const int nx = d->nx, ny = d->ny;
    const size_t n = (size_t)nx * ny;
    const double EPS = 1.0e-12;
    const double EPS_gs = sqrt(GRAVITY * MIN_DEPTH);
    const double half_gravity = 0.5 * GRAVITY;
    const double half_cfl = 0.5 * d->cfl;
    double *stg = P(d, STG), *uhp = P(d, UH), *vhp = P(d, VH), *elv = P(d, ELV);
    double *u = d->vel, *v = d->vel + n;
    double *fNS = d->flux_NS, *fEW = d->flux_EW;
    const size_t nNS = (size_t)nx * (ny + 1), nEW = (size_t)(nx + 1) * ny;
#define FNS(i, j, k) fNS[(size_t)(k) * nNS + (size_t)((j) - 1) * nx + ((i) - 1)]
#define FEW(i, j, k) fEW[(size_t)(k) * nEW + (size_t)((j) - 1) * (nx + 1) + ((i) - 1)]
#define ES(i, j, k) d->es[(size_t)((k) - 1) * n + IX(d, i, j)]
#define ESJ(i, j) d->es_vh_jm1[(size_t)((j) - 1) * nx + ((i) - 1)]

    compute_depth_and_velocity(d);

    double max_dt = d->maximum_timestep;
    double *work = malloc(sizeof(double) * nx);
    for (int i = 0; i < nx; i++) work[i] = 1.0 / max_dt;
    double *thL = calloc(nx, sizeof(double)), *th = calloc(nx, sizeof(double)), *thE = calloc(nx, sizeof(double));
    double *gl[4], *gc[4], *ge[4];
    for (int q = 0; q < 4; q++) {
        gl[q] = calloc(nx, sizeof(double)); gc[q] = calloc(nx, sizeof(double)); ge[q] = calloc(nx, sizeof(double));
    }
    double *es_im1 = calloc(nx, sizeof(double));
    ns_gradients(d, 1, th, gc);
    for (int j = 2; j <= ny; j++) {
        for (int i = 1; i <= nx; i++) { ES(i, j, 1) = 0.0; ES(i, j, 2) = 0.0; ESJ(i, j) = 0.0; es_im1[i - 1] = 0.0; }
        double dx_cfl_half_inv1 = 1.0 / (d->dbe[j - 1] * half_cfl);
        double dx_cfl_half_inv2 = 1.0 / (d->dle[0] * half_cfl);
        for (int q = 0; q < 4; q++) { double *t = gl[q]; gl[q] = gc[q]; gc[q] = t; }
        { double *t = thL; thL = th; th = t; }
        ns_gradients(d, j, th, gc);
        for (int i = 2; i <= nx; i++) {
            /* NS edge between (i, j-1) and (i, j): include.f90:178-432 */
            size_t km = IX(d, i, j - 1), kc = IX(d, i, j);
            double stage_neg = stg[km] + 0.5 * gl[0][i - 1];
            double stage_pos = stg[kc] - 0.5 * gc[0][i - 1];
            double depth_neg = d->depth[km] + 0.5 * gl[1][i - 1];
            double depth_pos = d->depth[kc] - 0.5 * gc[1][i - 1];
            double u_neg = u[km] + 0.5 * gl[2][i - 1], u_pos = u[kc] - 0.5 * gc[2][i - 1];
            double v_neg = v[km] + 0.5 * gl[3][i - 1], v_pos = v[kc] - 0.5 * gc[3][i - 1];
            double depth_neg_c = d->depth[km], depth_pos_c = d->depth[kc];
            double z_neg = stage_neg - depth_neg, z_pos = stage_pos - depth_pos;
            double z_half = dmax(z_neg, z_pos);
            double sns = dmax(stage_neg, z_half), dns = sns - z_half;
            if (dns == 0.0) { v_neg = 0.0; u_neg = 0.0; }
            double vd_neg = v_neg * dns, ud_neg = u_neg * dns;
            double gs_neg = sqrt(GRAVITY * dns);
            double sps = dmax(stage_pos, z_half), dps = sps - z_half;
            if (dps == 0.0) { v_pos = 0.0; u_pos = 0.0; }
            double vd_pos = v_pos * dps, ud_pos = u_pos * dps;
            double gs_pos = sqrt(GRAVITY * dps);
            double vbn = v_neg * 1.0, vbp = v_pos * 1.0;
            double s_max = dmax(dmax(vbn + gs_neg, vbp + gs_pos), 0.0);
            double s_min = dmin(dmin(vbn - gs_neg, vbp - gs_pos), 0.0);
            double denom = s_max - s_min;
            double inv_denom = (denom > EPS && (gs_pos > EPS_gs || gs_neg > EPS_gs)) ? d->dbe[j - 1] / denom : 0.0;
            double sminsmax = s_min * s_max * 1.0;
            double fr2 = 1.0 * sqrt(dmax(0.001, dmin(1.0, (v_neg * v_neg + v_pos * v_pos + u_neg * u_neg + u_pos * u_pos) /
                                                              (gs_neg * gs_neg + gs_pos * gs_pos + 1.0e-10))));
            FNS(i, j, 0) = (s_max * vd_neg - s_min * vd_pos + sminsmax * (sps - sns)) * inv_denom;
            FNS(i, j, 1) = (s_max * ud_neg * vbn - s_min * ud_pos * vbp + fr2 * sminsmax * (ud_pos - ud_neg)) * inv_denom;
            FNS(i, j, 2) = (s_max * vd_neg * vbn - s_min * vd_pos * vbp + fr2 * sminsmax * (vd_pos - vd_neg)) * inv_denom;
            double hghh = inv_denom * half_gravity * (s_max * dns * dns - s_min * dps * dps);
            double bsp_n = half_gravity * ((dns - depth_neg) * (dns + depth_neg) -
                                           (depth_neg + depth_neg_c) * (z_neg - elv[km]));
            ESJ(i, j) = bsp_n * d->dbe[j - 1] - hghh;
            double bsp_s = half_gravity * ((dps - depth_pos) * (dps + depth_pos) -
                                           (depth_pos + depth_pos_c) * (z_pos - elv[kc]));
            ES(i, j, 2) = ES(i, j, 2) - bsp_s * d->dbe[j - 1] + hghh;
            double max_speed = dmax(s_max, -s_min);
            if (max_speed > EPS) work[i - 1] = dmax(work[i - 1], max_speed * dx_cfl_half_inv2);
        }
        ew_gradients(d, j, thE, ge);
        for (int i = 2; i <= nx; i++) {
            /* EW edge between (i-1, j) and (i, j): include.f90:454-718 */
            size_t km = IX(d, i - 1, j), kc = IX(d, i, j);
            double stage_neg = stg[km] + 0.5 * ge[0][i - 2];
            double stage_pos = stg[kc] - 0.5 * ge[0][i - 1];
            double depth_neg = d->depth[km] + 0.5 * ge[1][i - 2];
            double depth_pos = d->depth[kc] - 0.5 * ge[1][i - 1];
            double u_neg = u[km] + 0.5 * ge[2][i - 2], u_pos = u[kc] - 0.5 * ge[2][i - 1];
            double v_neg = v[km] + 0.5 * ge[3][i - 2], v_pos = v[kc] - 0.5 * ge[3][i - 1];
            double depth_neg_c = d->depth[km], depth_pos_c = d->depth[kc];
            double z_neg = stage_neg - depth_neg, z_pos = stage_pos - depth_pos;
            double z_half = dmax(z_neg, z_pos);
            double sns = dmax(stage_neg, z_half), dns = sns - z_half;
            if (dns == 0.0) { u_neg = 0.0; v_neg = 0.0; }
            double ud_neg = u_neg * dns, vd_neg = v_neg * dns;
            double gs_neg = sqrt(GRAVITY * dns);
            double sps = dmax(stage_pos, z_half), dps = sps - z_half;
            if (dps == 0.0) { u_pos = 0.0; v_pos = 0.0; }
            double ud_pos = u_pos * dps, vd_pos = v_pos * dps;
            double gs_pos = sqrt(GRAVITY * dps);
            double vbn = u_neg * 1.0, vbp = u_pos * 1.0;
            double s_max = dmax(dmax(vbn + gs_neg, vbp + gs_pos), 0.0);
            double s_min = dmin(dmin(vbn - gs_neg, vbp - gs_pos), 0.0);
            double denom = s_max - s_min;
            double inv_denom = (denom > EPS && (gs_pos > EPS_gs || gs_neg > EPS_gs)) ? d->dle[i - 1] / denom : 0.0;
            double sminsmax = s_min * s_max * 1.0;
            double fr2 = 1.0 * sqrt(dmax(0.001, dmin(1.0, (v_neg * v_neg + v_pos * v_pos + u_neg * u_neg + u_pos * u_pos) /
                                                              (gs_neg * gs_neg + gs_pos * gs_pos + 1.0e-10))));
            FEW(i, j, 0) = (s_max * ud_neg - s_min * ud_pos + sminsmax * (sps - sns)) * inv_denom;
            FEW(i, j, 1) = (s_max * ud_neg * vbn - s_min * ud_pos * vbp + fr2 * sminsmax * (ud_pos - ud_neg)) * inv_denom;
            FEW(i, j, 2) = (s_max * vd_neg * vbn - s_min * vd_pos * vbp + fr2 * sminsmax * (vd_pos - vd_neg)) * inv_denom;
            double hghh = inv_denom * half_gravity * (s_max * dns * dns - s_min * dps * dps);
            double bsp_e = half_gravity * ((dns - depth_neg) * (dns + depth_neg) -
                                           (depth_neg + depth_neg_c) * (z_neg - elv[km]));
            es_im1[i - 2] = bsp_e * d->dle[i - 1] - hghh;
            double bsp_w = half_gravity * ((dps - depth_pos) * (dps + depth_pos) -
                                           (depth_pos + depth_pos_c) * (z_pos - elv[kc]));
            ES(i, j, 1) = ES(i, j, 1) - bsp_w * d->dle[i - 1] + hghh;
            if (d->spherical) {
                /* include.f90:674-710 */
                ES(i, j, 2) = ES(i, j, 2) + half_gravity * d->depth[kc] * d->depth[kc] * (d->dbe[j] - d->dbe[j - 1]);
                double cm;
                if (d->coriolis_on) {
                    cm = d->area[j - 1] * d->cor[j - 1];
                    ES(i, j, 1) = ES(i, j, 1) + cm * vhp[kc];
                    ES(i, j, 2) = ES(i, j, 2) - cm * uhp[kc];
                }
                cm = d->area[j - 1] * d->tanlat[j - 1] * u[kc];
                ES(i, j, 1) = ES(i, j, 1) + cm * vhp[kc];
                ES(i, j, 2) = ES(i, j, 2) - cm * uhp[kc];
            }
            double max_speed = dmax(s_max, -s_min);
            if (max_speed > EPS) work[i - 1] = dmax(work[i - 1], max_speed * dx_cfl_half_inv1);
        }
        for (int i = 1; i <= nx; i++) ES(i, j, 1) = ES(i, j, 1) + es_im1[i - 1];
    }
    double max_dt_inv = work[0];
    for (int i = 1; i < nx; i++) max_dt_inv = dmax(max_dt_inv, work[i]);
    max_dt = 1.0 / max_dt_inv;
    free(work); free(thL); free(th); free(thE); free(es_im1);
    for (int q = 0; q < 4; q++) { free(gl[q]); free(gc[q]); free(ge[q]); }

    /* Boundary flux sums: domain_mod.f90:1483-1499 */
    const int ne = d->n_ext;
    double s;
    s = 0.0; for (int i = 1 + ne; i <= nx - ne; i++) s += FNS(i, ny + 1 - ne, 0);
    d->boundary_flux_store[0] = s;
    s = 0.0; for (int j = 1 + ne; j <= ny - ne; j++) s += FEW(nx + 1 - ne, j, 0);
    d->boundary_flux_store[1] = s;
    s = 0.0; for (int i = 1 + ne; i <= nx - ne; i++) s += FNS(i, 1 + ne, 0);
    d->boundary_flux_store[2] = -s;
    s = 0.0; for (int j = 1 + ne; j <= ny - ne; j++) s += FEW(1 + ne, j, 0);
    d->boundary_flux_store[3] = -s;
    return max_dt;
}

/* update_U_manning / update_U_chezy: domain_mod_update_U_DE1_inner_include.f90:14-99,
 * then update_boundary (domain_mod.f90:1564-1582). */
static void update_U(or_domain *d, double dt)
{
    const int nx = d->nx, ny = d->ny;
    const size_t n = (size_t)nx * ny;
    const size_t nNS = (size_t)nx * (ny + 1), nEW = (size_t)(nx + 1) * ny;
    double *fNS = d->flux_NS, *fEW = d->flux_EW;
    const double power = (d->friction_type == 1) ? -2.0 : -7.0 / 3.0;
    double *u = d->vel, *v = d->vel + n;
    d->dt_last_update = dt;
    double dt_gravity = dt * GRAVITY;
    /* with dispersion the edge cells are not updated (update_U_DE1_inner_include.f90:19-31) */
    const int e = d->disp_on ? 1 : 0;
    for (int j = 1 + e; j <= ny - e; j++) {
        double inv_cell_area_dt = dt / d->area[j - 1];
        for (int i = 1 + e; i <= nx - e; i++)
            for (int kk = 0; kk < 3; kk++)
                A(d, kk, i, j) = A(d, kk, i, j) - inv_cell_area_dt * ((FNS(i, j + 1, kk) - FNS(i, j, kk)) +
                                                                      (FEW(i + 1, j, kk) - FEW(i, j, kk)));
        for (int i = 1 + e; i <= nx - e; i++) {
            size_t k = IX(d, i, j);
            double depth = d->depth[k];
            if (A(d, STG, i, j) <= (A(d, ELV, i, j) + MIN_DEPTH)) {
                A(d, UH, i, j) = 0.0;
                A(d, VH, i, j) = 0.0;
            } else {
                double fs = d->manning_sq[k] * sqrt(u[k] * u[k] + v[k] * v[k]) * pow(dmax(depth, MIN_DEPTH), power);
                double implicit_factor = 1.0 / (1.0 + dt_gravity * dmax(depth, 0.0) * fs + dt * d->linear_friction_coeff);
                A(d, UH, i, j) = A(d, UH, i, j) + inv_cell_area_dt * ES(i, j, 1);
                A(d, UH, i, j) = A(d, UH, i, j) * implicit_factor;
                A(d, VH, i, j) = A(d, VH, i, j) + inv_cell_area_dt * (ES(i, j, 2) + ESJ(i, j + 1));
                A(d, VH, i, j) = A(d, VH, i, j) * implicit_factor;
            }
        }
    }
    d->time = d->time + dt;
    update_boundary(d);
#undef FNS
#undef FEW
#undef ES
#undef ESJ
}

static double boundary_flux_sum(const or_domain *d)
{
    return ((d->boundary_flux_store[0] + d->boundary_flux_store[1]) + d->boundary_flux_store[2]) +
           d->boundary_flux_store[3];
}

/* one_euler_step: domain_mod_timestepping_alternatives_include.f90:8-62 */
static void euler_step(or_domain *d, double timestep)
{
    double ts = compute_fluxes(d);
    d->max_dt = ts;
    if (timestep > 0.0) ts = timestep;
    update_U(d, ts);
    d->boundary_flux_evolve_integral = d->boundary_flux_evolve_integral + ts * boundary_flux_sum(d);
}

/* one_rk2_step: domain_mod_timestepping_alternatives_include.f90:64-146 */
static void rk2_step(or_domain *d, double timestep)
{
    const size_t n = (size_t)d->nx * d->ny;
    double backup_time = d->time;
    memcpy(d->backup_U, d->U, 3 * n * sizeof(double));
    double dt_first_step;
    if (timestep > 0.0) { euler_step(d, timestep); dt_first_step = timestep; }
    else { euler_step(d, -1.0); dt_first_step = d->max_dt; }
    double max_dt_store = d->max_dt;
    euler_step(d, dt_first_step);
    for (size_t k = 0; k < 3 * n; k++) d->U[k] = 0.5 * (d->U[k] + d->backup_U[k]);
    d->time = backup_time + dt_first_step;
    d->boundary_flux_evolve_integral = 0.5 * d->boundary_flux_evolve_integral;
    d->max_dt = max_dt_store;
}

/* one_rk2n_step: domain_mod_timestepping_alternatives_include.f90:149-254, with
 * rk2n_n_value = 5 (domain_mod.f90:88-90): n-1 Euler steps of dt/(n-1), subtract 1/n of
 * the change, one more Euler step of (n-1)/n of a sub-step, add the 1/n back. */
#define RK2N_N 5
static void rk2n_step(or_domain *d, double timestep)
{
    const size_t n3 = 3 * (size_t)d->nx * d->ny;
    const double n_inverse = 1.0 / (1.0 * RK2N_N);
    double backup_time = d->time;
    memcpy(d->backup_U, d->U, n3 * sizeof(double));
    double dt_first_step;
    if (timestep > 0.0) {
        dt_first_step = timestep / (RK2N_N - 1.0);
        euler_step(d, dt_first_step);
    } else {
        euler_step(d, -1.0);
        dt_first_step = d->max_dt;
    }
    double max_dt_store = d->max_dt;
    for (int j = 2; j <= RK2N_N - 1; j++) euler_step(d, dt_first_step);
    for (size_t k = 0; k < n3; k++) d->backup_U[k] = (d->backup_U[k] - d->U[k]) * n_inverse;
    d->boundary_flux_evolve_integral = d->boundary_flux_evolve_integral * (RK2N_N - 1.0) * n_inverse;
    double reduced_dt = (1.0 * RK2N_N - 1.0) * n_inverse * dt_first_step;
    euler_step(d, reduced_dt);
    for (size_t k = 0; k < n3; k++) d->U[k] = d->U[k] + d->backup_U[k];
    d->time = backup_time + dt_first_step * (RK2N_N - 1);
    d->max_dt = max_dt_store;
}

/* one_midpoint_step: domain_mod_timestepping_alternatives_include.f90:257-336 (no
 * dispersion): half an Euler step for the fluxes, then a full step from the start
 * state with those fluxes (and the depth/velocity of the half-step state). */
static void midpoint_step(or_domain *d, double timestep)
{
    const size_t n3 = 3 * (size_t)d->nx * d->ny;
    double backup_time = d->time;
    memcpy(d->backup_U, d->U, n3 * sizeof(double));
    double dt_first_step = compute_fluxes(d);
    d->max_dt = dt_first_step;
    if (timestep > 0.0) dt_first_step = timestep;
    update_U(d, dt_first_step * 0.5);
    /* dispersion in the predictor (timestepping_alternatives_include.f90:286-297) */
    if (d->disp_on) or_domain_dispersion_solve(d, 0, 1, backup_time + dt_first_step * 0.5);
    compute_fluxes(d);
    memcpy(d->U, d->backup_U, n3 * sizeof(double));
    d->time = backup_time;
    update_U(d, dt_first_step);
    d->boundary_flux_evolve_integral = boundary_flux_sum(d) * dt_first_step;
}

/* update_max_quantities: domain_mod.f90:2243-2404, with all six max_U variables
 * (so time_of_max_stage is tracked and the 'toms > 0' branch of max_stage applies). */
static void update_max_quantities(or_domain *d)
{
    if (!d->record_max_U) return;
    const int nx = d->nx, ny = d->ny;
    const size_t n = (size_t)nx * ny;
    double *m = d->max_U;
    const int staggered = (d->method == OR_LINEAR || d->method == OR_LLPNF || d->method == OR_LEAPFROG_NONLINEAR);
    /* depth relative to msl only for the truely-linear 'linear' / llpnf solvers (:2300-2302) */
    const int truely = d->truely_linear && (d->method == OR_LINEAR || d->method == OR_LLPNF);
    for (int j = d->yL; j <= d->yU; j++) {
        for (int i = d->xL; i <= d->xU; i++) {
            size_t k = IX(d, i, j);
            double stg = A(d, STG, i, j), uh = A(d, UH, i, j), vh = A(d, VH, i, j), elv = A(d, ELV, i, j);
            /* include_dry_cells = .true. (domain_mod.f90 default) */
            if (stg > m[MX_STAGE * n + k]) { m[MX_STAGE * n + k] = stg; m[MX_TOMS * n + k] = d->time; }
            if (staggered) {
                /* depth_E = 0.5 * sum(depth(i:ip1, j)) with ip1 = min(i+1, nx): on the east
                 * column the sum has one term (half the cell's depth), likewise depth_N on
                 * the north row (domain_mod.f90:2311-2320) */
                const double h0 = truely ? (d->msl_linear - A(d, ELV, i, j)) : (A(d, STG, i, j) - A(d, ELV, i, j));
                double depth_E = h0, depth_N = h0;
                if (i < nx)
                    depth_E = h0 + (truely ? (d->msl_linear - A(d, ELV, i + 1, j)) : (A(d, STG, i + 1, j) - A(d, ELV, i + 1, j)));
                if (j < ny)
                    depth_N = h0 + (truely ? (d->msl_linear - A(d, ELV, i, j + 1)) : (A(d, STG, i, j + 1) - A(d, ELV, i, j + 1)));
                depth_E = 0.5 * depth_E;
                depth_N = 0.5 * depth_N;
                double iE = depth_E > MIN_DEPTH ? 1.0 / depth_E : 0.0;
                double iN = depth_N > MIN_DEPTH ? 1.0 / depth_N : 0.0;
                double sp = sqrt(uh * uh * (iE * iE) + vh * vh * (iN * iN));
                m[MX_SPEED * n + k] = dmax(m[MX_SPEED * n + k], sp);
            } else {
                double ld = stg - elv;
                double li = ld > MIN_DEPTH ? 1.0 / ld : 0.0;
                m[MX_SPEED * n + k] = dmax(m[MX_SPEED * n + k], sqrt((uh * uh + vh * vh) * (li * li)));
            }
            m[MX_FLUX * n + k] = dmax(m[MX_FLUX * n + k], sqrt(uh * uh + vh * vh));
            {
                double ld = stg - elv;
                double arrival_stage = d->msl_linear + d->arrival_stage_threshold;
                if (ld > MIN_DEPTH && stg > arrival_stage && m[MX_ARRIVAL * n + k] < 0.0)
                    m[MX_ARRIVAL * n + k] = d->time;
            }
            m[MX_MIN_STAGE * n + k] = dmin(m[MX_MIN_STAGE * n + k], stg);
        }
    }
}

/* Nonlinear staggered leapfrog ('leapfrog_nonlinear'):
 * domain_mod_leapfrog_nonlinear_include.f90:61-826 with FROUDE_LIMITING
 * (domain_mod_timestepping_alternatives_include.f90:439-447), serial loop order.
 * Upwind momentum advection from work arrays formed from the pre-update UH/VH and a
 * time-lagged depth, wet/dry switches, Coriolis and the spherical metric terms
 * (spherical builds), semi-implicit friction from friction_work recomputed every step
 * from the new stage, a Froude limiter, and zero outflow from dry cells. */
static void leapfrog_nonlinear_step(or_domain *d, double dt)
{
    const int nx = d->nx, ny = d->ny;
    const size_t n = (size_t)nx * ny;
    const double beta = 1.0;      /* advection_beta, global_mod.f90:38 */
    const double fr_lim = 10.0;   /* leapfrog_froude_limit, domain_mod.f90:210 */
    d->dt_last_update = dt;
    double *wk = calloc(9 * n, sizeof(double));
    double *DL = wk, *UUH = wk + n, *UVH = wk + 2 * n, *VUH = wk + 3 * n, *VVH = wk + 4 * n;
    double *fEWu = wk + 5 * n, *fEWv = wk + 6 * n, *fNSu = wk + 7 * n, *fNSv = wk + 8 * n;
#define W(a, i, j) ((a)[IX(d, i, j)])
#define WET(i, j) (A(d, STG, i, j) > A(d, ELV, i, j) + MIN_DEPTH)
    /* :63-71 depth before the stage update; work arrays and fluxes cleared */
    for (int j = 1; j <= ny; j++)
        for (int i = 1; i <= nx; i++) W(DL, i, j) = dmax(A(d, STG, i, j) - A(d, ELV, i, j), 0.0);

    lf_stage_boundary(d, dt);   /* :97-204 */
    precompute_friction_work(d, 0);   /* :218, not truely linear: every step */

    /* :265-268 time-lagged depth on rows 1..ny-1 (row ny keeps the old depth) */
    for (int j = 1; j <= ny - 1; j++)
        for (int i = 1; i <= nx; i++)
            W(DL, i, j) = 0.5 * (W(DL, i, j) + dmax(A(d, STG, i, j) - A(d, ELV, i, j), 0.0));

    /* :274-299 U*UH and U*VH at (i+1/2, j), extrapolated to the east column and row ny */
    for (int j = 1; j <= ny - 1; j++) {
        const int jm1 = j - 1 > 1 ? j - 1 : 1;
        for (int i = 1; i <= nx - 1; i++) {
            const double depth_local = 0.5 * (W(DL, i, j) + W(DL, i + 1, j));
            const double inv = depth_local > MIN_DEPTH ? 1.0 / depth_local : 0.0;
            W(UUH, i, j) = (A(d, UH, i, j) * A(d, UH, i, j)) * inv;
            if (i < nx - 1)
                W(UVH, i, j) = A(d, UH, i, j) * inv * 0.25 *
                               (A(d, VH, i, jm1) + A(d, VH, i + 1, jm1) + A(d, VH, i, j) + A(d, VH, i + 1, j));
        }
        W(UUH, nx, j) = W(UUH, nx - 1, j);
        W(UVH, nx, j) = W(UVH, nx - 2, j);
        W(UVH, nx - 1, j) = W(UVH, nx - 2, j);
    }
    for (int i = 1; i <= nx; i++) {
        W(UUH, i, ny) = W(UUH, i, ny - 1);
        W(UVH, i, ny) = W(UVH, i, ny - 1);
    }
    /* :305-328 V*VH and V*UH at (i, j+1/2) */
    for (int j = 1; j <= ny - 1; j++) {
        for (int i = 1; i <= nx; i++) {
            const double depth_local = 0.5 * (W(DL, i, j) + W(DL, i, j + 1));
            const double inv = depth_local > MIN_DEPTH ? 1.0 / depth_local : 0.0;
            W(VVH, i, j) = A(d, VH, i, j) * A(d, VH, i, j) * inv;
            if (i > 1 && j < ny - 1)
                W(VUH, i, j) = A(d, VH, i, j) * inv * 0.25 *
                               (A(d, UH, i - 1, j) + A(d, UH, i, j) + A(d, UH, i - 1, j + 1) + A(d, UH, i, j + 1));
        }
        W(VUH, 1, j) = W(VUH, 2, j);
    }
    for (int i = 1; i <= nx; i++) {
        W(VVH, i, ny) = W(VVH, i, ny - 1);
        W(VUH, i, ny) = W(VUH, i, ny - 2);
        W(VUH, i, ny - 1) = W(VUH, i, ny - 2);
    }

    /* :338-363 EW flux of UH (upwind; the mean where the neighbours cancel; dry: 0) */
    for (int j = 1; j <= ny - 1; j++) {
        for (int i = 2; i <= nx - 1; i++)
            W(fEWu, i, j) = d->dle[i - 1] * ((A(d, UH, i, j) + A(d, UH, i - 1, j) > 0.0) ? W(UUH, i - 1, j) : W(UUH, i, j));
        W(fEWu, 1, j) = W(UUH, 1, j) * d->dle[0];
        W(fEWu, nx, j) = W(UUH, nx, j) * d->dle[nx - 1];
        for (int i = 2; i <= nx - 1; i++)
            if ((A(d, UH, i, j) + A(d, UH, i - 1, j)) == 0.0)
                W(fEWu, i, j) = d->dle[i - 1] * 0.5 * (W(UUH, i - 1, j) + W(UUH, i, j));
        for (int i = 1; i <= nx; i++)
            if (!WET(i, j)) W(fEWu, i, j) = 0.0;
    }
    /* :366-409 NS flux of UH on the top edge of row j, then the bottom edge of row 1 */
    for (int j = 0; j <= ny - 1; j++) {
        const int jt = j == 0 ? 1 : j + 1;        /* the edge written */
        const int ja = j == 0 ? 1 : j;            /* the row below it (row 1 for the bottom edge) */
        const int jb = j == 0 ? 1 : j + 1;        /* the row above it */
        const double dbe = d->dbe[jt - 1];
        for (int i = 1; i <= nx - 1; i++) {
            const double vs = A(d, VH, i, ja) + A(d, VH, i + 1, ja);
            double f = dbe * (vs > 0.0 ? W(UVH, i, ja) : W(UVH, i, jb));
            if (vs == 0.0) f = dbe * 0.5 * (W(UVH, i, ja) + W(UVH, i, jb));
            if (!(WET(i, ja) && WET(i + 1, ja) && WET(i, jb) && WET(i + 1, jb))) f = 0.0;
            W(fNSu, i, jt) = f;
        }
    }
    /* :412-455 NS flux of VH on the bottom edge of rows 1..ny */
    for (int j = 1; j <= ny; j++) {
        const int jm1 = j - 1 > 1 ? j - 1 : 1;
        const double dbe = 0.5 * (d->dbe[j - 1] + d->dbe[j]);
        for (int i = 1; i <= nx - 1; i++) {
            const double vs = A(d, VH, i, j) + A(d, VH, i, jm1);
            double f = dbe * (vs > 0.0 ? W(VVH, i, jm1) : W(VVH, i, j));
            if (vs == 0.0) f = dbe * 0.5 * (W(VVH, i, jm1) + W(VVH, i, j));
            if (!WET(i, j)) f = 0.0;
            W(fNSv, i, j) = f;
        }
    }
    /* :457-481 EW flux of VH; the west column copies column 2 before its mean/dry fixes */
    for (int j = 1; j <= ny - 1; j++) {
        for (int i = 2; i <= nx; i++)
            W(fEWv, i, j) = d->dle[i - 2] *
                            ((A(d, UH, i - 1, j) + A(d, UH, i - 1, j + 1) > 0.0) ? W(VUH, i - 1, j) : W(VUH, i, j));
        W(fEWv, 1, j) = W(fEWv, 2, j);
        for (int i = 2; i <= nx; i++) {
            if ((A(d, UH, i - 1, j) + A(d, UH, i - 1, j + 1)) == 0.0)
                W(fEWv, i, j) = d->dle[i - 2] * 0.5 * (W(VUH, i - 1, j) + W(VUH, i, j));
            if (!(WET(i - 1, j) && WET(i, j) && WET(i - 1, j + 1) && WET(i, j + 1))) W(fEWv, i, j) = 0.0;
        }
    }

    /* :575-787 momentum update, rows 1..ny-1, columns 1..nx-1, from the pre-update
     * UH/VH (the row-carried Coriolis averages are read from a copy) */
    double *old = malloc(2 * n * sizeof(double));
    memcpy(old, P(d, UH), 2 * n * sizeof(double));
#define OUH(i, j) (old[IX(d, i, j)])
#define OVH(i, j) (old[n + IX(d, i, j)])
    const double *fwU = d->friction_work, *fwV = d->friction_work + n;
    for (int j = 1; j <= ny - 1; j++) {
        const double ica_dt = dt / d->area[j - 1];
        const double ica_dt_g = GRAVITY * ica_dt;
        const double ica_dt_vh = dt / (0.5 * (d->area[j - 1] + d->area[j]));
        const double ica_dt_vh_g = GRAVITY * ica_dt_vh;
        const int jr = j > 1 ? j - 1 : 1;
        for (int i = 1; i <= nx - 1; i++) {
            const double dw_j = A(d, STG, i, j + 1) - A(d, STG, i, j);
            double h_jph = 0.5 * ((A(d, STG, i, j + 1) + A(d, STG, i, j)) - (A(d, ELV, i, j + 1) + A(d, ELV, i, j)));
            if (h_jph < MIN_DEPTH) h_jph = 0.0;
            double h_iph = 0.5 * ((A(d, STG, i + 1, j) + A(d, STG, i, j)) - (A(d, ELV, i + 1, j) + A(d, ELV, i, j)));
            if (h_iph < MIN_DEPTH) h_iph = 0.0;
            const int jph_strict = WET(i, j + 1) && WET(i, j);
            const double jph_wet = (jph_strict ||
                                    (WET(i, j + 1) && A(d, STG, i, j + 1) > A(d, ELV, i, j) + MIN_DEPTH) ||
                                    (WET(i, j) && A(d, STG, i, j) > A(d, ELV, i, j + 1) + MIN_DEPTH)) ? 1.0 : 0.0;
            const int iph_strict = WET(i + 1, j) && WET(i, j);
            const double iph_wet = (iph_strict ||
                                    (WET(i + 1, j) && A(d, STG, i + 1, j) > A(d, ELV, i, j) + MIN_DEPTH) ||
                                    (WET(i, j) && A(d, STG, i, j) > A(d, ELV, i + 1, j) + MIN_DEPTH)) ? 1.0 : 0.0;
            const double vh_iph_jmh = 0.5 * (OVH(i, jr) + OVH(i + 1, jr));
            const double vh_iph_jph = 0.5 * (OVH(i, j) + OVH(i + 1, j));
            const int im1 = i > 1 ? i - 1 : 1;
            const double uh_i_j = 0.5 * (OUH(i, j) + OUH(im1, j));
            const double uh_i_jp1 = 0.5 * (OUH(i, j + 1) + OUH(im1, j + 1));
            double dt_half_cor = 0.0, dt_half_cor_jph = 0.0, t_uvh = 0.0, t_uuh = 0.0;
            if (d->coriolis_on) {
                dt_half_cor = dt * d->cor[j - 1] * 0.5 * (iph_strict ? 1.0 : 0.0);
                dt_half_cor_jph = dt * d->cor_be[j] * 0.5 * (jph_strict ? 1.0 : 0.0);
            }
            if (d->spherical) {
                double inv = 0.0;
                if (iph_strict) inv = 1.0 / h_iph;
                t_uvh = dt * d->tanlat[j - 1] * OUH(i, j) * 0.5 * (vh_iph_jmh + vh_iph_jph) * inv;
                inv = 0.0;
                if (jph_strict) inv = 1.0 / h_jph;
                t_uuh = dt * 0.5 * (d->tanlat[j - 1] + d->tanlat[j]) * 0.5 * (uh_i_jp1 + uh_i_j) * 0.5 *
                        (uh_i_jp1 + uh_i_j) * inv;
            }
            const double vbar = 0.5 * (vh_iph_jmh + vh_iph_jph);
            const double ubar = 0.5 * (uh_i_j + uh_i_jp1);
            const double fm_UH = 1.0 / (1.0 + dt * d->linear_friction_coeff +
                                        dt * fwU[IX(d, i, j)] * sqrt(OUH(i, j) * OUH(i, j) + vbar * vbar));
            const double fm_VH = 1.0 / (1.0 + dt * d->linear_friction_coeff +
                                        dt * fwV[IX(d, i, j)] * sqrt(OVH(i, j) * OVH(i, j) + ubar * ubar));
            double uh = OUH(i, j) * iph_wet -
                        ica_dt_g * h_iph * iph_wet * (A(d, STG, i + 1, j) - A(d, STG, i, j)) * d->dle[i];
            double vh = OVH(i, j) * jph_wet - ica_dt_vh_g * h_jph * jph_wet * dw_j * d->dbe[j];
            uh = uh - ica_dt * beta * iph_wet *
                          ((W(fEWu, i + 1, j) - W(fEWu, i, j)) + (W(fNSu, i, j + 1) - W(fNSu, i, j)));
            vh = vh - ica_dt_vh * beta * jph_wet *
                          ((W(fEWv, i + 1, j) - W(fEWv, i, j)) + (W(fNSv, i, j + 1) - W(fNSv, i, j)));
            if (d->coriolis_on) {
                uh = uh + dt_half_cor * (vh_iph_jmh + vh_iph_jph);
                vh = vh - dt_half_cor_jph * (uh_i_j + uh_i_jp1);
            }
            if (d->spherical) {
                uh = uh + t_uvh;
                vh = vh - t_uuh;
            }
            uh = fm_UH * uh;
            vh = fm_VH * vh;
            A(d, UH, i, j) = uh;
            A(d, VH, i, j) = vh;
        }
        /* :763-783 Froude limiter on this row */
        for (int i = 1; i <= nx - 1; i++) {
            double dl = 0.5 * (A(d, STG, i, j) - A(d, ELV, i, j) + A(d, STG, i + 1, j) - A(d, ELV, i + 1, j));
            if (dl < MIN_DEPTH) dl = 0.0;
            const double lim = fr_lim * sqrt(dl * GRAVITY) * dl * copysign(1.0, A(d, UH, i, j));
            if (A(d, UH, i, j) * A(d, UH, i, j) > lim * lim) A(d, UH, i, j) = lim;
        }
        for (int i = 1; i <= nx - 1; i++) {
            double dl = 0.5 * (A(d, STG, i, j) - A(d, ELV, i, j) + A(d, STG, i, j + 1) - A(d, ELV, i, j + 1));
            if (dl < MIN_DEPTH) dl = 0.0;
            const double lim = fr_lim * sqrt(dl * GRAVITY) * dl * copysign(1.0, A(d, VH, i, j));
            if (A(d, VH, i, j) * A(d, VH, i, j) > lim * lim) A(d, VH, i, j) = lim;
        }
    }
#undef OUH
#undef OVH
    free(old);
    /* :792-815 no outflow from dry cells */
    for (int j = 1; j <= ny - 1; j++) {
        for (int i = 1; i <= nx; i++) {
            if (A(d, STG, i, j) < A(d, ELV, i, j) + MIN_DEPTH) {
                if (A(d, UH, i, j) > 0.0) A(d, UH, i, j) = 0.0;
                if (i > 1 && A(d, UH, i - 1, j) < 0.0) A(d, UH, i - 1, j) = 0.0;
                if (A(d, VH, i, j) > 0.0) A(d, VH, i, j) = 0.0;
            }
            if (A(d, STG, i, j + 1) < A(d, ELV, i, j + 1) + MIN_DEPTH)
                if (A(d, VH, i, j) < 0.0) A(d, VH, i, j) = 0.0;
        }
    }
#undef W
#undef WET
    d->time = d->time + 0.5 * dt;
    free(wk);
}

/* ---- Dispersion on the staggered grid: linear_dispersive_solver_mod.f90 (without
 * DISPERSIVE_PEREGRINE_IN_FLUX_FORM, the default) and quadratic_extrapolation_mod.f90 ---- */

/* quadratic_extrapolation_mod.f90:17 (missing_time) */
static const double QET_MISSING = DBL_MAX * 0.987654321;

void or_domain_dispersion_setup(or_domain *d, double td1, double td2, int niter)
{
    const size_t n = (size_t)d->nx * d->ny;
    d->disp_on = 1; d->td1 = td1; d->td2 = td2; d->disp_iter = niter;
    d->dRHS = calloc(2 * n, sizeof(double));
    d->dlast = calloc(3 * n, sizeof(double));
    d->dAx = calloc(2 * n, sizeof(double));
    for (int k = 0; k < 3; k++) {
        d->qy[k] = calloc(2 * n, sizeof(double));
        d->qt[k] = QET_MISSING;
    }
}

/* store_last_U (linear_dispersive_solver_mod.f90:127-142): UH, VH, ELV */
void or_domain_dispersion_store(or_domain *d)
{
    const size_t n = (size_t)d->nx * d->ny;
    memcpy(d->dlast, P(d, UH), 3 * n * sizeof(double));
}

/* append_value_and_time_for_quadratic_extrapolation (quadratic_extrapolation_mod.f90:52-78):
 * the levels shift down one, the newest is U(UH:VH) at the domain time */
void or_domain_dispersion_append(or_domain *d)
{
    const size_t n = (size_t)d->nx * d->ny;
    double *old = d->qy[0];
    d->qy[0] = d->qy[1]; d->qt[0] = d->qt[1];
    d->qy[1] = d->qy[2]; d->qt[1] = d->qt[2];
    d->qy[2] = old;
    memcpy(d->qy[2], P(d, UH), 2 * n * sizeof(double));
    d->qt[2] = d->time;
}

/* LAPACK dgtsv for one right-hand side (the reference vendors it:
 * util/lapack_tridiagonal_solves.f:462-673): Gaussian elimination with partial
 * pivoting by rows; dl[k], du[k] are the sub/super-diagonal entries of rows k+1 / k. */
static void or_dgtsv(int n, double *dl, double *dg, double *du, double *b)
{
    double fact, temp;
    for (int i = 0; i < n - 2; i++) {
        if (fabs(dg[i]) >= fabs(dl[i])) {
            if (dg[i] != 0.0) {
                fact = dl[i] / dg[i];
                dg[i + 1] = dg[i + 1] - fact * du[i];
                b[i + 1] = b[i + 1] - fact * b[i];
            } else {
                return;
            }
            dl[i] = 0.0;
        } else {
            fact = dg[i] / dl[i];
            dg[i] = dl[i];
            temp = dg[i + 1];
            dg[i + 1] = du[i] - fact * temp;
            dl[i] = du[i + 1];
            du[i + 1] = -fact * dl[i];
            du[i] = temp;
            temp = b[i];
            b[i] = b[i + 1];
            b[i + 1] = temp - fact * b[i + 1];
        }
    }
    if (n > 1) {
        const int i = n - 2;
        if (fabs(dg[i]) >= fabs(dl[i])) {
            if (dg[i] != 0.0) {
                fact = dl[i] / dg[i];
                dg[i + 1] = dg[i + 1] - fact * du[i];
                b[i + 1] = b[i + 1] - fact * b[i];
            } else {
                return;
            }
        } else {
            fact = dg[i] / dl[i];
            dg[i] = dl[i];
            temp = dg[i + 1];
            dg[i + 1] = du[i] - fact * temp;
            du[i] = temp;
            temp = b[i];
            b[i] = b[i + 1];
            b[i + 1] = temp - fact * b[i + 1];
        }
    }
    if (dg[n - 1] == 0.0) return;
    b[n - 1] = b[n - 1] / dg[n - 1];
    if (n > 1) b[n - 2] = (b[n - 2] - du[n - 2] * b[n - 1]) / dg[n - 2];
    for (int i = n - 3; i >= 0; i--) b[i] = (b[i] - du[i] * b[i + 1] - dl[i] * b[i + 2]) / dg[i];
}

/* depth at a UH / VH point for the dispersive terms: msl-relative, zero unless both
 * cells are below msl_linear (:209-210, :266-267) */
static double disp_depth(double msl, double e1, double e2)
{
    return (e1 < msl && e2 < msl) ? 0.5 * (msl - e1 + msl - e2) : 0.0;
}

static double disp_premult(const or_domain *d, double dep, double r)
{
    return dep * dep / (3.0 * r) * dmin(1.0, dmax(0.0, (dep - d->td2) / (d->td1 - d->td2)));
}

/* linear_dispersive_staggered_matmult (:182-417) of (elev, uh, vh) into Ax */
static void disp_matmult(or_domain *d, const double *elev, const double *uh, const double *vh, double *axu,
                         double *axv)
{
    const int nx = d->nx, ny = d->ny;
    const double msl = d->msl_linear, dlon = d->dx[0], dlat = d->dx[1];
#define E2(a, i, j) ((a)[IX(d, i, j)])
    for (int j = 2; j <= ny - 1; j++) {
        const double Rdlon = 0.5 * (d->dbe[j] + d->dbe[j - 1]);
        const double Rdlat = Rdlon * dlat / dlon;
        const double cjph = d->dbe[j] * dlat / (dlon * d->dle[0]);
        const double cjmh = d->dbe[j - 1] * dlat / (dlon * d->dle[0]);
        for (int i = 2; i <= nx - 1; i++) {
            const double dep = disp_depth(msl, E2(elev, i, j), E2(elev, i + 1, j));
            const double pm = disp_premult(d, dep, Rdlon);
            E2(axu, i, j) = pm * (1.0 / Rdlon * (E2(uh, i + 1, j) - 2.0 * E2(uh, i, j) + E2(uh, i - 1, j)) +
                                  1.0 / Rdlat * ((cjph * E2(vh, i + 1, j) - cjmh * E2(vh, i + 1, j - 1)) -
                                                 (cjph * E2(vh, i, j) - cjmh * E2(vh, i, j - 1))));
        }
    }
    for (int j = 2; j <= ny - 1; j++) {
        const double r_dlat = d->dle[0];
        const double Rdlon_jp1 = 0.5 * (d->dbe[j + 1] + d->dbe[j]);
        const double Rdlon_j = 0.5 * (d->dbe[j] + d->dbe[j - 1]);
        const double Rdlat_jp1 = Rdlon_jp1 * dlat / dlon;
        const double Rdlat_j = Rdlon_j * dlat / dlon;
        const double cjp1h = d->dbe[j + 1] * dlat / (d->dle[0] * dlon);
        const double cjph = d->dbe[j] * dlat / (d->dle[0] * dlon);
        const double cjmh = d->dbe[j - 1] * dlat / (d->dle[0] * dlon);
        for (int i = 2; i <= nx - 1; i++) {
            const double dep = disp_depth(msl, E2(elev, i, j), E2(elev, i, j + 1));
            const double pm = disp_premult(d, dep, r_dlat);
            E2(axv, i, j) = pm * ((1.0 / Rdlon_jp1 * (E2(uh, i, j + 1) - E2(uh, i - 1, j + 1)) -
                                   1.0 / Rdlon_j * (E2(uh, i, j) - E2(uh, i - 1, j))) +
                                  (1.0 / Rdlat_jp1 * (cjp1h * E2(vh, i, j + 1) - cjph * E2(vh, i, j)) -
                                   1.0 / Rdlat_j * (cjph * E2(vh, i, j) - cjmh * E2(vh, i, j - 1))));
        }
    }
}

/* linear_dispersive_sweep_staggered_TRIDIAG (:663-921): every row of UH, then every
 * column of VH, each a tridiagonal solve */
static void disp_sweep(or_domain *d)
{
    const int nx = d->nx, ny = d->ny;
    const double msl = d->msl_linear, dlon = d->dx[0], dlat = d->dx[1];
    const double *elev = P(d, ELV);
    double *uh = P(d, UH), *vh = P(d, VH);
    const double *ru = d->dRHS, *rv = d->dRHS + (size_t)nx * ny;
    const int nmax = nx > ny ? nx : ny;
    double *dg = malloc(nmax * sizeof(double)), *lo = malloc(nmax * sizeof(double));
    double *up = malloc(nmax * sizeof(double)), *b = malloc(nmax * sizeof(double));
    for (int j = 2; j <= ny - 1; j++) {
        const double Rdlon = 0.5 * (d->dbe[j] + d->dbe[j - 1]);
        const double Rdlat = Rdlon * dlat / dlon;
        const double cjph = d->dbe[j] * dlat / (dlon * d->dle[0]);
        const double cjmh = d->dbe[j - 1] * dlat / (dlon * d->dle[0]);
        const int N = nx;
        dg[0] = 1.0; dg[N - 1] = 1.0;
        lo[0] = 0.0; lo[N - 1] = 0.0; up[0] = 0.0; up[N - 1] = 0.0;
        b[0] = E2(ru, 1, j); b[N - 1] = E2(ru, N, j);
        for (int i = 2; i <= N - 1; i++) {
            const double dep = disp_depth(msl, E2(elev, i, j), E2(elev, i + 1, j));
            const double pm = disp_premult(d, dep, Rdlon);
            dg[i - 1] = 1.0 + 2.0 * pm * 1.0 / Rdlon;
            lo[i - 1] = -pm * 1.0 / Rdlon;
            up[i - 1] = -pm * 1.0 / Rdlon;
            b[i - 1] = E2(ru, i, j) + pm * (1.0 / Rdlat * ((cjph * E2(vh, i + 1, j) - cjmh * E2(vh, i + 1, j - 1)) -
                                                          (cjph * E2(vh, i, j) - cjmh * E2(vh, i, j - 1))));
        }
        or_dgtsv(N, lo + 1, dg, up, b);
        for (int i = 1; i <= N; i++) E2(uh, i, j) = b[i - 1];
    }
    for (int i = 2; i <= nx - 1; i++) {
        const int N = ny;
        dg[0] = 1.0; dg[N - 1] = 1.0;
        lo[0] = 0.0; lo[N - 1] = 0.0; up[0] = 0.0; up[N - 1] = 0.0;
        b[0] = E2(rv, i, 1); b[N - 1] = E2(rv, i, N);
        for (int j = 2; j <= N - 1; j++) {
            const double r_dlat = d->dle[0];
            const double Rdlon_jp1 = 0.5 * (d->dbe[j + 1] + d->dbe[j]);
            const double Rdlon_j = 0.5 * (d->dbe[j] + d->dbe[j - 1]);
            const double Rdlat_jp1 = Rdlon_jp1 * dlat / dlon;
            const double Rdlat_j = Rdlon_j * dlat / dlon;
            const double cjp1h = d->dbe[j + 1] * dlat / (d->dle[0] * dlon);
            const double cjph = d->dbe[j] * dlat / (d->dle[0] * dlon);
            const double cjmh = d->dbe[j - 1] * dlat / (d->dle[0] * dlon);
            const double dep = disp_depth(msl, E2(elev, i, j), E2(elev, i, j + 1));
            const double pm = disp_premult(d, dep, r_dlat);
            dg[j - 1] = 1.0 + pm * (1.0 / Rdlat_jp1 * cjph + 1.0 / Rdlat_j * cjph);
            lo[j - 1] = -pm * 1.0 / Rdlat_j * cjmh;
            up[j - 1] = -pm * 1.0 / Rdlat_jp1 * cjp1h;
            b[j - 1] = E2(rv, i, j) + pm * ((1.0 / Rdlon_jp1 * (E2(uh, i, j + 1) - E2(uh, i - 1, j + 1)) -
                                            1.0 / Rdlon_j * (E2(uh, i, j) - E2(uh, i - 1, j))));
        }
        or_dgtsv(N, lo + 1, dg, up, b);
        for (int j = 1; j <= N; j++) E2(vh, i, j) = b[j - 1];
    }
    free(dg); free(lo); free(up); free(b);
}

/* cell-centred depth for the dispersive terms: msl-relative where the cell and its two
 * neighbours along the line are all below msl_linear (:451, :493) */
static double disp_depth_cc(double msl, double em, double e0, double ep)
{
    return (em < msl && e0 < msl && ep < msl) ? (msl - e0) : 0.0;
}

/* linear_dispersive_cellcentred_matmult (:419-660) */
static void disp_matmult_cc(or_domain *d, const double *elev, const double *uh, const double *vh, double *axu,
                            double *axv)
{
    const int nx = d->nx, ny = d->ny;
    const double msl = d->msl_linear, dlon = d->dx[0], dlat = d->dx[1];
    for (int j = 2; j <= ny - 1; j++) {
        const double Rdlon = 0.5 * (d->dbe[j] + d->dbe[j - 1]);
        const double Rdlat = Rdlon * dlat / dlon;
        const double cjp1 = 0.5 * (d->dbe[j + 1] + d->dbe[j]) * dlat / (dlon * d->dle[0]);
        const double cjm1 = 0.5 * (d->dbe[j - 1] + d->dbe[j - 2]) * dlat / (dlon * d->dle[0]);
        for (int i = 2; i <= nx - 1; i++) {
            const double dep = disp_depth_cc(msl, E2(elev, i - 1, j), E2(elev, i, j), E2(elev, i + 1, j));
            const double pm = disp_premult(d, dep, Rdlon);
            E2(axu, i, j) = pm * (1.0 / Rdlon * (E2(uh, i + 1, j) - 2.0 * E2(uh, i, j) + E2(uh, i - 1, j)) +
                                  1.0 / Rdlat * ((cjp1 * E2(vh, i + 1, j + 1) - cjm1 * E2(vh, i + 1, j - 1)) * 0.25 -
                                                 (cjp1 * E2(vh, i - 1, j + 1) - cjm1 * E2(vh, i - 1, j - 1)) * 0.25));
        }
    }
    for (int j = 2; j <= ny - 1; j++) {
        const double r_dlat = d->dle[0];
        const double Rdlon_jph = d->dbe[j], Rdlon_jmh = d->dbe[j - 1];
        const double Rdlat_jph = Rdlon_jph * dlat / dlon, Rdlat_jmh = Rdlon_jmh * dlat / dlon;
        const double cjp1 = 0.5 * (d->dbe[j + 1] + d->dbe[j]) * dlat / (d->dle[0] * dlon);
        const double cj = 0.5 * (d->dbe[j] + d->dbe[j - 1]) * dlat / (d->dle[0] * dlon);
        const double cjm1 = 0.5 * (d->dbe[j - 1] + d->dbe[j - 2]) * dlat / (d->dle[0] * dlon);
        for (int i = 2; i <= nx - 1; i++) {
            const double dep = disp_depth_cc(msl, E2(elev, i, j - 1), E2(elev, i, j), E2(elev, i, j + 1));
            const double pm = disp_premult(d, dep, r_dlat);
            E2(axv, i, j) =
                pm * ((1.0 / Rdlon_jph *
                           (E2(uh, i + 1, j + 1) - E2(uh, i - 1, j + 1) + E2(uh, i + 1, j) - E2(uh, i - 1, j)) * 0.25 -
                       1.0 / Rdlon_jmh *
                           (E2(uh, i + 1, j - 1) - E2(uh, i - 1, j - 1) + E2(uh, i + 1, j) - E2(uh, i - 1, j)) * 0.25) +
                      (1.0 / Rdlat_jph * (cjp1 * E2(vh, i, j + 1) - cj * E2(vh, i, j)) -
                       1.0 / Rdlat_jmh * (cj * E2(vh, i, j) - cjm1 * E2(vh, i, j - 1))));
        }
    }
}

/* linear_dispersive_sweep_cellcentred_TRIDIAG (:923-1191) */
static void disp_sweep_cc(or_domain *d)
{
    const int nx = d->nx, ny = d->ny;
    const double msl = d->msl_linear, dlon = d->dx[0], dlat = d->dx[1];
    const double *elev = P(d, ELV);
    double *uh = P(d, UH), *vh = P(d, VH);
    const double *ru = d->dRHS, *rv = d->dRHS + (size_t)nx * ny;
    const int nmax = nx > ny ? nx : ny;
    double *dg = malloc(nmax * sizeof(double)), *lo = malloc(nmax * sizeof(double));
    double *up = malloc(nmax * sizeof(double)), *b = malloc(nmax * sizeof(double));
    for (int j = 2; j <= ny - 1; j++) {
        const double Rdlon = 0.5 * (d->dbe[j] + d->dbe[j - 1]);
        const double Rdlat = Rdlon * dlat / dlon;
        const double cjp1 = 0.5 * (d->dbe[j + 1] + d->dbe[j]) * dlat / (dlon * d->dle[0]);
        const double cjm1 = 0.5 * (d->dbe[j - 1] + d->dbe[j - 2]) * dlat / (dlon * d->dle[0]);
        const int N = nx;
        dg[0] = 1.0; dg[N - 1] = 1.0;
        lo[0] = 0.0; lo[N - 1] = 0.0; up[0] = 0.0; up[N - 1] = 0.0;
        b[0] = E2(ru, 1, j); b[N - 1] = E2(ru, N, j);
        for (int i = 2; i <= N - 1; i++) {
            const double dep = disp_depth_cc(msl, E2(elev, i - 1, j), E2(elev, i, j), E2(elev, i + 1, j));
            const double pm = disp_premult(d, dep, Rdlon);
            dg[i - 1] = 1.0 + 2.0 * pm * 1.0 / Rdlon;
            lo[i - 1] = -pm * 1.0 / Rdlon;
            up[i - 1] = -pm * 1.0 / Rdlon;
            b[i - 1] = E2(ru, i, j) + pm * (1.0 / Rdlat * ((cjp1 * E2(vh, i + 1, j + 1) - cjm1 * E2(vh, i + 1, j - 1)) * 0.25 -
                                                          (cjp1 * E2(vh, i - 1, j + 1) - cjm1 * E2(vh, i - 1, j - 1)) * 0.25));
        }
        or_dgtsv(N, lo + 1, dg, up, b);
        for (int i = 1; i <= N; i++) E2(uh, i, j) = b[i - 1];
    }
    for (int i = 2; i <= nx - 1; i++) {
        const int N = ny;
        dg[0] = 1.0; dg[N - 1] = 1.0;
        lo[0] = 0.0; lo[N - 1] = 0.0; up[0] = 0.0; up[N - 1] = 0.0;
        b[0] = E2(rv, i, 1); b[N - 1] = E2(rv, i, N);
        for (int j = 2; j <= N - 1; j++) {
            const double r_dlat = d->dle[0];
            const double Rdlon_jph = d->dbe[j], Rdlon_jmh = d->dbe[j - 1];
            const double Rdlat_jph = Rdlon_jph * dlat / dlon, Rdlat_jmh = Rdlon_jmh * dlat / dlon;
            const double cjp1 = 0.5 * (d->dbe[j + 1] + d->dbe[j]) * dlat / (d->dle[0] * dlon);
            const double cj = 0.5 * (d->dbe[j] + d->dbe[j - 1]) * dlat / (d->dle[0] * dlon);
            const double cjm1 = 0.5 * (d->dbe[j - 1] + d->dbe[j - 2]) * dlat / (d->dle[0] * dlon);
            const double dep = disp_depth_cc(msl, E2(elev, i, j - 1), E2(elev, i, j), E2(elev, i, j + 1));
            const double pm = disp_premult(d, dep, r_dlat);
            dg[j - 1] = 1.0 + pm * (1.0 / Rdlat_jph * cj + 1.0 / Rdlat_jmh * cj);
            lo[j - 1] = -pm * 1.0 / Rdlat_jmh * cjm1;
            up[j - 1] = -pm * 1.0 / Rdlat_jph * cjp1;
            b[j - 1] = E2(rv, i, j) +
                       pm * ((1.0 / Rdlon_jph *
                                  (E2(uh, i + 1, j + 1) - E2(uh, i - 1, j + 1) + E2(uh, i + 1, j) - E2(uh, i - 1, j)) * 0.25 -
                              1.0 / Rdlon_jmh *
                                  (E2(uh, i + 1, j - 1) - E2(uh, i - 1, j - 1) + E2(uh, i + 1, j) - E2(uh, i - 1, j)) * 0.25));
        }
        or_dgtsv(N, lo + 1, dg, up, b);
        for (int j = 1; j <= N; j++) E2(vh, i, j) = b[j - 1];
    }
    free(dg); free(lo); free(up); free(b);
}

/* linear_dispersive_solve_TRIDIAG (:1193-1354) as solve_dispersive_terms calls it
 * (domain_mod.f90:4319-4343) */
void or_domain_dispersion_solve(or_domain *d, int rhs_up, int estimate, double ftime)
{
    const int nx = d->nx, ny = d->ny;
    const size_t n = (size_t)nx * ny;
    const int stag = (d->method == OR_LINEAR || d->method == OR_LLPNF || d->method == OR_LEAPFROG_NONLINEAR);
    if (!rhs_up) {
        if (stag)
            disp_matmult(d, d->dlast + 2 * n, d->dlast, d->dlast + n, d->dAx, d->dAx + n);
        else
            disp_matmult_cc(d, d->dlast + 2 * n, d->dlast, d->dlast + n, d->dAx, d->dAx + n);
        for (int q = 0; q < 2; q++) {
            const double *u = P(d, UH + q), *ax = d->dAx + q * n;
            double *r = d->dRHS + q * n;
            for (int j = 1; j <= ny; j++)
                for (int i = 1; i <= nx; i++) {
                    const int edge = (j == 1 || j == ny || i == 1 || i == nx);
                    E2(r, i, j) = edge ? E2(u, i, j) : E2(u, i, j) - E2(ax, i, j);
                }
        }
    }
    int niter = d->disp_iter;
    if (estimate) {
        if (d->qt[0] == QET_MISSING || d->qt[1] == QET_MISSING || d->qt[2] == QET_MISSING) {
            niter = 10 * d->disp_iter;
        } else {
            /* extrapolate_in_time (quadratic_extrapolation_mod.f90:91-193) */
            const double tn2 = d->qt[0] - d->qt[2], tn1 = d->qt[1] - d->qt[2], t1 = ftime - d->qt[2];
            const double premult = 1.0 / (tn1 * (tn2 * tn2) - tn2 * (tn1 * tn1));
            const double a = (1.0 + t1 * premult * (-(tn2 * tn2) + tn1 * tn1) + (t1 * t1) * premult * (tn2 - tn1));
            const double bb = (t1 * premult * (tn2 * tn2) + (t1 * t1) * premult * (-tn2));
            const double c = (t1 * premult * (-(tn1 * tn1)) + (t1 * t1) * premult * tn1);
            double *u = P(d, UH);
            for (size_t k = 0; k < 2 * n; k++) u[k] = d->qy[2][k] * a + d->qy[1][k] * bb + d->qy[0][k] * c;
        }
    }
    for (int it = 0; it < niter; it++) {
        if (stag) disp_sweep(d);
        else disp_sweep_cc(d);
    }
#undef E2
}

/* evolve_one_step: domain_mod.f90:1679-1813 (max_U_update_frequency = 1) */
static void evolve_one_step(or_domain *d, double dt)
{
    double time0 = d->time;
    d->boundary_flux_evolve_integral = 0.0;
    if (d->nsteps_advanced == 0) update_max_quantities(d);
    switch (d->method) {
    case OR_RK2: rk2_step(d, dt); break;
    case OR_EULER: euler_step(d, dt); break;
    case OR_RK2N: rk2n_step(d, dt); break;
    case OR_MIDPOINT: midpoint_step(d, dt); break;
    case OR_LINEAR:
    case OR_LLPNF: leapfrog_step(d, dt); break;
    case OR_LEAPFROG_NONLINEAR: leapfrog_nonlinear_step(d, dt); break;
    default: break;
    }
    update_max_quantities(d);
    d->evolve_step_dt = d->time - time0;
    d->boundary_flux_time_integral = d->boundary_flux_time_integral + d->boundary_flux_evolve_integral;
    d->nsteps_advanced += 1;
}

double or_domain_evolve(or_domain *d, double dt, int nsteps)
{
    for (int k = 0; k < nsteps; k++) evolve_one_step(d, dt);
    return d->time;
}

void or_domain_get_scalars(const or_domain *d, double *out)
{
    out[0] = d->time; out[1] = d->max_dt; out[2] = d->evolve_step_dt;
    out[3] = d->boundary_flux_time_integral; out[4] = d->boundary_flux_evolve_integral;
    for (int k = 0; k < 4; k++) out[5 + k] = d->boundary_flux_store[k];
    out[9] = d->negative_depth_fix_counter; out[10] = d->nsteps_advanced; out[11] = d->dt_last_update;
}

/* volume_interior: domain_mod.f90:1815-1849 (serial row order) */
double or_domain_volume_interior(const or_domain *d)
{
    const int ne = d->n_ext;
    double vol = 0.0;
    for (int j = 1 + ne; j <= d->ny - ne; j++) {
        double local_sum = 0.0;
        for (int i = 1 + ne; i <= d->nx - ne; i++) local_sum += A(d, STG, i, j) - A(d, ELV, i, j);
        vol = vol + d->area[j - 1] * local_sum;
    }
    return vol;
}

/* stationary_timestep_max: domain_mod.f90:1906-2022 */
double or_domain_stationary_timestep_max(const or_domain *d)
{
    double ts_max = HUGE_VAL;
    const int linear = (d->method == OR_LINEAR || d->method == OR_LLPNF) && d->truely_linear;
    for (int j = 1; j <= d->ny; j++) {
        for (int i = 1; i <= d->nx; i++) {
            double h = linear ? d->msl_linear - A(d, ELV, i, j) : A(d, STG, i, j) - A(d, ELV, i, j);
            double dist = dmin(d->dbe[j - 1] + d->dbe[j], d->dle[i - 1] + d->dle[i]);
            ts_max = dmin(ts_max, 0.5 * dist / sqrt(GRAVITY * dmax(h, MIN_DEPTH)));
        }
    }
    double ts = ts_max * d->cfl;
    if (d->method == OR_RK2 || d->method == OR_EULER || d->method == OR_MIDPOINT) ts = ts * 0.5;
    if (d->method == OR_RK2N) ts = ts * 0.5 * (RK2N_N - 1.0);
    return ts;
}

void or_domain_get_max_U(const or_domain *d, int k, double *out)
{
    const size_t n = (size_t)d->nx * d->ny;
    memcpy(out, d->max_U + (size_t)(k - 1) * n, n * sizeof(double));
}

/* update_gauge_var: point_gauge_mod.f90:208-245 ([STG, UH, VH] at 1-based (i, j)) */
void or_domain_get_gauges(const or_domain *d, int n, const int *ij, double *out)
{
    for (int g = 0; g < n; g++)
        for (int k = 0; k < 3; k++) out[(size_t)k * n + g] = A(d, k, ij[2 * g], ij[2 * g + 1]);
}
