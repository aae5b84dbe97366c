/*
 * grasph_oracle.h — FP64 CPU restatement of Grasph.jl's particle step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker of the HIP
 * path and the timed CPU baseline; nothing in the shipped library links or
 * calls it.  It follows the reference's algorithms and evaluation order
 * (colour → cell → i → j sweeps, serial ghost generation, stable sort, the
 * leapfrog step) in Float64, with OpenMP threads over the active cells of one
 * colour as Polyester `@batch` does.  Each function of grasph_oracle.cpp
 * cites the reference file:line it restates.
 *
 * The API mirrors include/grasph_hip.h (same enums, same argument meaning)
 * with an `orc_` prefix and a scene instead of a device context.
 */
#ifndef GRASPH_ORACLE_H
#define GRASPH_ORACLE_H
#include <stdint.h>
#include "../include/grasph_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_scene orc_scene;

int orc_scene_create(orc_scene **out, int ndims);
int orc_scene_destroy(orc_scene *s);
const char *orc_last_error(const orc_scene *s);
int orc_set_threads(orc_scene *s, int nthreads);

int orc_system_create(orc_scene *s, int kind, int64_t n, int ns, double mass,
                      double c, const double *source_v, double source_rho,
                      int *sys_out);
int orc_system_add_updater(orc_scene *s, int sys, int stage, int kind,
                           const double *params, int nparams);
int orc_system_upload(orc_scene *s, int sys, int field, const double *data,
                      int64_t count);
int orc_system_download(orc_scene *s, int sys, int field, double *data,
                        int64_t count);
int orc_system_upload_index(orc_scene *s, int sys, int field,
                            const int64_t *data, int64_t count);
int orc_system_download_index(orc_scene *s, int sys, int field, int64_t *data,
                              int64_t count);
int orc_system_get_n(orc_scene *s, int sys, int64_t *n_out);

int orc_virtual_create(orc_scene *s, int sys, int nzero, const int *zero_fields,
                       const double *prescribed_v);
int orc_ghost_create(orc_scene *s, int source_sys, double cutoff, int nb,
                     const double *normals, const double *points,
                     int *ghost_out);
int orc_ghost_set_copier(orc_scene *s, int ghost, int stage, int nfields,
                         const int *fields, const int *modes);

int orc_interaction_create(orc_scene *s, int kernel_kind, double h, int sys_a,
                           int partner, int wrap, const double *normal,
                           const double *point, double beta, double lj_cutoff,
                           int n_stages, const gr_pfn_desc *pfns,
                           int *inter_out);

int orc_integrator_create(orc_scene *s, int nsys, const int *systems,
                          int ninter, const int *inters, int nghost,
                          const int *ghosts, double gamma, int *integ_out);
int orc_interaction_set_vadjust(orc_scene *s, int inter, double epsilon);
int orc_integrator_set_scheme(orc_scene *s, int integ, int scheme);
int orc_integrator_dt(orc_scene *s, int integ, double cfl, double *dt_out);
int orc_step(orc_scene *s, int integ, int nsteps, double dt);

int orc_sort(orc_scene *s, int sys, double cutoff);
int orc_create_grid(orc_scene *s, int inter);
int orc_sweep(orc_scene *s, int inter, int stage);
int orc_adjust_v(orc_scene *s, int inter);
int orc_probe_create(orc_scene *s, int sys, const double *prescribed_v, int mirror, int nzero,
                     const int *zero_fields);
int orc_integrator_set_probes(orc_scene *s, int integ, int nprobes, const int *probes, int ninters,
                              const int *inters);
int orc_measure_probes(orc_scene *s, int integ);
int orc_update_state(orc_scene *s, int sys, int stage, double dt);
int orc_generate_ghosts(orc_scene *s, int ghost);
int orc_update_ghost_kinematics(orc_scene *s, int ghost);
int orc_update_ghost(orc_scene *s, int ghost, int stage);
int orc_get_grid(orc_scene *s, int inter, double *mingridx, int64_t *ngridx,
                 int64_t *ncells);
int orc_get_cell_start(orc_scene *s, int inter, int which, int64_t *out,
                       int64_t count);

/* Brute-force checker: for every particle i of sys_a, the number of
 * particles j of sys_b (j != i when sys_a == sys_b) with
 * dot(x_i - x_j, x_i - x_j) < cutoff^2, evaluated as the reference's
 * _pair_self!/_pair_coupled! do (src/Interaction.jl:546-548,721-723). */
int orc_brute_force_counts(orc_scene *s, int sys_a, int sys_b, double cutoff,
                           int64_t *counts_out);

/* Closed forms exposed for known-answer tests (src/Kernels.jl:42-60,70-81,
 * 116-130,142-153; Julia's Float64 ^ Integer). */
double orc_kernel_norm(int kernel_kind, int ndims, double h);
double orc_kernel_w(int kernel_kind, int ndims, double h, double q);
double orc_kernel_dw_dq(int kernel_kind, int ndims, double h, double q);
double orc_pow_int(double x, int64_t n);

#ifdef __cplusplus
}
#endif
#endif
