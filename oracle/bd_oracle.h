/*
 * bd_oracle.h -- FP64 single-threaded CPU restatement of the reference
 * Brownian-dynamics path.  TEST INFRASTRUCTURE ONLY: it is the checker of
 * the HIP engine and the timed CPU baseline; no product code links it.
 *
 * Every function cites the reference file:line it restates (paths relative
 * to the reference tree).  Deviations from the reference are listed at the
 * function that makes them.
 */
#ifndef BD_ORACLE_H
#define BD_ORACLE_H

#include <stdint.h>
#include "../include/bd_abi.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- rannew64 (rannew64.cu:55-126) with explicit state ---- */
typedef struct orc_rn64 {
  uint64_t ring[250];
  uint64_t lcg;
  int32_t np, nq;
} orc_rn64;

void orc_rn64_seed(orc_rn64* s, int64_t idum);
uint64_t orc_rn64_u64(orc_rn64* s);
double orc_rn64_double(orc_rn64* s);
uint64_t orc_lane_seed(uint64_t seed, uint64_t lane);

/* ---- Gaussian transform (project-defined Box-Muller) ---- */
double orc_log(double x);
void orc_sincos_turn(double u, double* s, double* c);
void orc_gauss3(const uint64_t d[4], double out[3]);

/* ---- protocol helpers ---- */
int orc_geometry(const bd_params* p, bd_geometry* g);
double orc_temp_linear(double T0, double Tf, double t0, double tf, double t);
double orc_temp_sine(double T0, double Tf, double t0, double tf, double period, double t);
double orc_temperature(const bd_params* p, int64_t ni);
int32_t orc_sampling_indexes(float minimal_decade, float run_time, float dt,
                             int32_t per_decade, int32_t* out, int32_t maxn);

/* ---- pair potentials (potentials.cu:9-76); kind and constants from p ---- */
void orc_potential(const bd_params* p, double sigma, double dist_inv, double* potential,
                   double* normal_force);

/* ---- one system ---- */
typedef struct orc_sys orc_sys;
orc_sys* orc_create(const bd_params* p);
void orc_destroy(orc_sys* s);
int orc_get_geometry(const orc_sys* s, bd_geometry* g);
int orc_set_positions(orc_sys* s, const double* xyz);
int orc_get_positions(const orc_sys* s, double* xyz);
int orc_get_raw_positions(const orc_sys* s, double* xyz);
int orc_set_forces(orc_sys* s, const double* f);
int orc_get_forces(const orc_sys* s, double* f, double* vir, double* pot);
int orc_mark_origin(orc_sys* s);
int64_t orc_get_step(const orc_sys* s);
double orc_get_temperature(const orc_sys* s);
int orc_cell_locate(orc_sys* s);
int orc_get_cells(const orc_sys* s, int32_t species, int32_t* nocup, int32_t* cell_vec);
int orc_compute_forces(orc_sys* s);
int orc_neighbours(orc_sys* s, int32_t maxn, int32_t* counts, int32_t* list);
int orc_update(orc_sys* s);
int orc_step(orc_sys* s, int32_t nsteps);
int orc_get_rng_lane(const orc_sys* s, int32_t lane, uint64_t* ring250, uint64_t* lcg,
                     int32_t* np, int32_t* nq);
int orc_gder(const orc_sys* s, int32_t nbins, double range, uint64_t* counts);
int orc_observables(const orc_sys* s, double q, bd_obs* out);

/* all-pairs forces with the minimum image (brute force, tiny N only) */
int orc_forces_brute(const bd_params* p, const double* xyz, double* f, double* vir,
                     double* pot);

/* g(r) normalisation (main.cu:849-877) of one sample's counts[4*nbins] */
void orc_gder_normalize(const bd_params* p, int32_t nbins, double range,
                        const uint64_t* counts, double* g);

/* ---- force contraction (Contracting_Vector3D.cu:57-297) ---- */
int orc_cv3d_accumulate(int32_t bins, double range, const double* xyz, const double* f,
                        int32_t n, double side, double* ata, double* atb);
int orc_cv3d_solve(int32_t bins, const double* ata, const double* atb, double* out);
int orc_jacobi_eig(int32_t n, const double* a, double* w, double* v);

#ifdef __cplusplus
}
#endif

#endif
