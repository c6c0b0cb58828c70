/*
 * oracle.h -- CPU restatement of the CabanaPIC per-timestep hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * implementation in cabanapic_amd/: only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it.  Nothing in the product links it.
 *
 * Every function is compiled twice from oracle_impl.inc: suffix _f for
 * real_t = float and _d for real_t = double, mirroring the reference's
 * compile-time REAL_TYPE switch (src/types.h:4-8).  A third instantiation,
 * suffix _ds (oracle_smooth.c), is the double build with a double sqrt in
 * place of the reference's sqrtf: a smooth function of its inputs, used only
 * as the finite-difference reference of the tangent-linear step.
 *
 * Storage conventions (shared with the product's C-ABI so tests can pass the
 * same host buffers to both):
 *   Ng        = (nx+2)(ny+2)(nz+2) cells, ghost width ng = 1
 *   voxel     = x + (nx+2)*(y + (ny+2)*z)           (src/types.h:195)
 *   fields    = 9 component arrays of Ng values, component-major, in the
 *               member order ex,ey,ez,cbx,cby,cbz,jfx,jfy,jfz
 *               (src/types.h:138-170)
 *   interp    = 18 component arrays of Ng values, member order of
 *               src/types.h:64-84
 *   acc       = Ng x 3 x 4 values, cell-major: acc[(cell*3 + comp)*4 + k]
 *               (src/types.h:116-120)
 *   particles = SoA x,y,z,ux,uy,uz,w (real_t) + int32 cell (src/types.h:31-54)
 *
 * Parity status per function (how each is pinned) is listed in the header
 * comment of oracle_impl.inc.
 */
#ifndef CPIC_ORACLE_H
#define CPIC_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORA_DECLARE(R, S)                                                      \
  void ora_load_interpolator##S(const R *fields, R *interp, long nx, long ny,  \
                                long nz);                                      \
  void ora_clear_accumulator##S(R *acc, long ncells);                          \
  long ora_push##S(R *x, R *y, R *z, R *ux, R *uy, R *uz, const R *w,          \
                   int32_t *cell, long np, const R *interp, R *acc,            \
                   R qdt_2mc, R cdt_dx, R cdt_dy, R cdt_dz, R qsp, long nx,    \
                   long ny, long nz, int boundary);                            \
  long ora_push_abs##S(R *x, R *y, R *z, R *ux, R *uy, R *uz, const R *w,      \
                       int32_t *cell, long np, const R *interp, R *acc,        \
                       double *abs_acc, R qdt_2mc, R cdt_dx, R cdt_dy,         \
                       R cdt_dz, R qsp, long nx, long ny, long nz,             \
                       int boundary);                                          \
  long ora_push_sum##S(R *x, R *y, R *z, R *ux, R *uy, R *uz, const R *w,      \
                       int32_t *cell, long np, const R *interp, R *acc,        \
                       double *abs_acc, double *sum_acc, R qdt_2mc, R cdt_dx,  \
                       R cdt_dy, R cdt_dz, R qsp, long nx, long ny, long nz,   \
                       int boundary);                                          \
  long ora_push_omp##S(R *x, R *y, R *z, R *ux, R *uy, R *uz, const R *w,      \
                       int32_t *cell, long np, const R *interp, R *acc,        \
                       R qdt_2mc, R cdt_dx, R cdt_dy, R cdt_dz, R qsp,         \
                       long nx, long ny, long nz, int boundary,                \
                       int nthreads);                                          \
  void ora_uncenter##S(const R *x, const R *y, const R *z, R *ux, R *uy,       \
                       R *uz, const int32_t *cell, long np, const R *interp,   \
                       R qdt_2mc, long nx, long ny, long nz);                  \
  void ora_unload_accumulator##S(R *fields, const R *acc, long nx, long ny,    \
                                 long nz, R dx, R dy, R dz, R dt);             \
  void ora_fold_ghosts##S(R *sx, R *sy, R *sz, long nx, long ny, long nz);     \
  void ora_copy_ghosts##S(R *sx, R *sy, R *sz, long nx, long ny, long nz);     \
  void ora_advance_b##S(R *fields, R px, R py, R pz, long nx, long ny,         \
                        long nz);                                              \
  void ora_advance_e##S(R *fields, R px, R py, R pz, long nx, long ny,         \
                        long nz, R dt_eps0);                                   \
  void ora_energies##S(const R *fields, long nx, long ny, long nz,             \
                       double *e64, double *b64, R *e_real, R *b_real);        \
  void ora_es_advance_e##S(R *fields, long nx, long ny, long nz, R dt_eps0);  \
  void ora_es_energy##S(const R *fields, long nx, long ny, long nz,           \
                        double *e64, R *e_real);                               \
  double ora_kinetic_energy##S(const R *ux, const R *uy, const R *uz,          \
                               const R *w, long np, R me, R c);                \
  void ora_default_particle_init##S(R *x, R *y, R *z, R *ux, R *uy, R *uz,     \
                                    R *w, int32_t *cell, long np, long nx,     \
                                    long ny, R dxp, long nppc, R we, R v0);    \
  long ora_run##S(R *fields, R *interp, R *acc, R *x, R *y, R *z, R *ux,       \
                  R *uy, R *uz, const R *w, int32_t *cell, long np, long nx,   \
                  long ny, long nz, const R *consts, long nsteps,              \
                  int nthreads, double *hist);                                 \
  void ora_step_adjoint##S(const R *fields, const R *x, const R *y,            \
                           const R *z, const R *ux, const R *uy, const R *uz,  \
                           const R *w, const int32_t *cell, long np, long nx,  \
                           long ny, long nz, const R *consts, double *lf,      \
                           const double *lp_out, double *lp_in);               \
  void ora_stage_adjoint##S(int which, double *lf, double *lacc,               \
                            const double *li, long nx, long ny, long nz,       \
                            double px, double py, double pz, double c0,        \
                            double c1, double c2);

ORA_DECLARE(float, _f)
ORA_DECLARE(double, _d)
ORA_DECLARE(double, _ds)

#undef ORA_DECLARE

/* consts[] layout for ora_run: qdt_2mc, cdt_dx, cdt_dy, cdt_dz, qsp, dx, dy,
 * dz, dt, px, py, pz, dt_eps0, me, c  (15 values). hist[] receives nsteps
 * rows of 6 doubles: e64, b64, e_real, b_real, ke64, crossings. */
#define ORA_NCONSTS 15
#define ORA_HIST_COLS 6

/* Deck-side derivations restated from src/input/deck.h and
 * example/example.cpp for the 2stream-em regression deck
 * (tests/energy_comparison/2stream-em.cxx:81-114).  out[] receives:
 * nx, ny, nz, nppc, num_steps, np, len_x, len_y, len_z, dx, dy, dz, dt, n0,
 * v0, Npe, we, dxp, then the 15 ora_run consts (33 values, as double). */
#define ORA_DECK_NOUT 33
void ora_deck_2stream_em_f(double *out);
void ora_deck_2stream_em_d(double *out);

#ifdef __cplusplus
}
#endif
#endif
