/*
 * mdz_model.h -- host-side model set-up for the Newton-iteration path:
 * grid allocation in the reference's layout, named field access, and the
 * seeded synthetic set-ups the path is exercised and measured on.
 *
 * This plays the role of the reference's set-up plugins (SOURCE/set_*.c,
 * SOURCE/BC.c) and of the per-time-step producers that fill the grid before
 * the Newton loop (marker interpolation, SOURCE/GridRoutines.c:410
 * ComputeLithostaticPressure).  Pure host C; no GPU.
 */
#ifndef MDZ_MODEL_H
#define MDZ_MODEL_H

#include "../../include/mdoodz_hot.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mdz_model {
    grid     *mesh;
    params    model;
    mat_prop  materials;
    scale     scaling;
    Nparams   Nmodel;
    int       nph;
    int       bc_family;
    unsigned long long seed;
} mdz_model;

/* Boundary-condition families (tag patterns of SOURCE/BC.c and set_*.c). */
enum {
    MDZ_BC_PURE_SHEAR = 0,   /* free-slip box, Dirichlet normal velocity (BC.c:63 SetBCs_freeSlipBox) */
    MDZ_BC_SIMPLE_SHEAR = 1, /* periodic in x, moving no-slip walls (set_Shear_VEVP.c shear_style=1) */
    MDZ_BC_STRESS_WE = 2,    /* traction (Neumann, tag 2) on West/East walls (BC.c:104,117) */
    MDZ_BC_FREE_SURFACE = 3, /* air layer (tag 30) above a surface-pressure row (tag 31) */
    MDZ_BC_NEUMANN_TOP = 4   /* traction on the top wall: Vz tag 2 on the north row */
};

/* Grid allocation in the reference layout; every array of mdz_fields.h plus
 * BC tags/values, comp_cells, index maps, coordinates and phase fractions. */
grid  *mdz_grid_alloc(int Nx, int Nz, int nph);
void   mdz_grid_free(grid *mesh, int nph);
/* Deep copy (same arrays, new allocations). */
grid  *mdz_grid_clone(const grid *mesh, int nph);
/* Named access: double arrays by field name ("eta_n", "BCu.val",
 * "phase_perc_n:1", ...); returns NULL for unknown names. */
double *mdz_grid_dfield(grid *mesh, const char *name, int *len);
/* Tag arrays "BCu", "BCv", "BCp", "BCg" (char). */
char   *mdz_grid_tags(grid *mesh, const char *name, int *len);
/* Integer arrays "comp_cells", "kp", "lp", "kn", "ln", "kvx", "lvx", "kvz", "lvz". */
int    *mdz_grid_ifield(grid *mesh, const char *name, int *len);
/* Number of double fields in the table and their names. */
int         mdz_grid_nfields(void);
const char *mdz_grid_field_name(int i);
int         mdz_grid_field_family(int i);
int         mdz_grid_field_role(int i);

/* Seeded set-ups.  ncx, ncz = cell counts (Nx = ncx+1, Nz = ncz+1). */
/* 2-D pure-shear visco-elasto-plastic inclusion, nondimensional. */
mdz_model *mdz_model_inclusion(int ncx, int ncz, unsigned long long seed);
/* The reference's power-law pure-shear benchmark (PureShear_pwl): matrix
 * n = 30, inclusion n = 3, viscous, background pure-shear initial iterate. */
mdz_model *mdz_model_pwl_inclusion(int ncx, int ncz);
/* Three-layer lithosphere (wet quartzite / plagioclase / olivine + Peierls),
 * 1000 x 250 km, dimensional inputs scaled as LithoScale (eta 1e22, L 1e4,
 * V 1e-9, T 100).  friction/cohesion/rate < 0 take the defaults
 * (30 deg, 10 MPa, 1e-15 1/s). */
mdz_model *mdz_model_lithosphere(int ncx, int ncz, unsigned long long seed, double friction_deg,
                                 double cohesion_pa, double rate);
/* Randomised fields exercising every rheology branch, on a chosen BC family.
 * materials: 0 = nondimensional constant-viscosity phases, 1 = dimensional
 * creep laws (power law, Peierls, linear, regularised plasticity, tension
 * cut-off). */
mdz_model *mdz_model_edgecase(int ncx, int ncz, unsigned long long seed, int bc_family,
                              int materials);
void       mdz_model_free(mdz_model *m);
mdz_model *mdz_model_clone(const mdz_model *m);

/* Structure sizes, for ABI layout checks from the bindings. */
size_t mdz_sizeof(const char *type_name);
/* offsetof(type, field) for the ABI structs; -1 when unknown. */
long   mdz_offsetof(const char *type_name, const char *field);

/* Host restatement of SOURCE/GridRoutines.c:410 ComputeLithostaticPressure
 * (mode 1: rho_n); a per-time-step producer of the path's pressure input. */
void mdz_lithostatic_pressure(grid *mesh, const params *model, double *p_lith);

#ifdef __cplusplus
}
#endif
#endif
