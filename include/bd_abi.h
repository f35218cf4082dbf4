/*
 * bd_abi.h -- C ABI of the MI355X Brownian-dynamics engine (libbd_hip.so).
 *
 * Plain C: fixed-width integers, doubles, opaque handles, int status codes.
 * No HIP, torch or C++ types cross this boundary.  Every call returns BD_OK
 * (0) or one of the BD_E_* codes; bd_last_error() returns a text message for
 * the last failing call on the calling thread.
 *
 * Reference interfaces replaced (paths relative to the reference tree):
 *   parametros (encabezados.h:40-75)           -> bd_params + bd_geometry
 *   cell geometry (main.cu:259-298)            -> bd_compute_geometry
 *   cudaMallocManaged/cudaFree (main.cu:303-321,991-1014) -> bd_create/bd_destroy
 *   set_vec_*_dev + cell_locate_dev (cell_locate.cu:110-209, main.cu:536-543)
 *                                              -> bd_cell_locate / bd_get_cells
 *   get_forces_same_dev + get_forces_diff_dev (get_forces.cu:297-558,
 *     main.cu:556-585)                         -> bd_compute_forces
 *   setup_rng_kernel (update_ermak.cu:85-89)   -> seeding inside bd_create
 *   update_ermak_dev (update_ermak.cu:91-160)  -> bd_update / bd_step
 *   calculate_temp_linear/_sine (set_temp.cu:7-23, main.cu:520-524)
 *                                              -> evaluated on the device per step
 *   get_gder_dev + host block reduce (get_gder.cu:69-127, main.cu:719-768)
 *                                              -> bd_gder
 *   fself_isotropic, MSD, energy, pressure (fself.cu:3-21, main.cu:627-638,
 *     776-806)                                 -> bd_observables
 *
 * Threading: a handle is not thread-safe; distinct handles are independent.
 * Every handle owns one HIP stream and its device memory; calls on one handle
 * are ordered on that stream.
 */
#ifndef BD_ABI_H
#define BD_ABI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (the reference exit()s or fails silently instead) ---- */
enum {
  BD_OK = 0,
  BD_E_ARG = 1,           /* invalid argument or geometry */
  BD_E_HIP = 2,           /* HIP runtime error */
  BD_E_CELL_OVERFLOW = 3, /* a cell holds >= ntags particles (main.cu:655-659) */
  BD_E_NAN = 4,           /* a non-finite position appeared (or one more than a
                             cell side outside the box) */
  BD_E_STATE = 5,         /* call out of order */
  BD_E_NOMEM = 6
};

/* pair potentials (potentials.cu:54-76); HARMONIC is the overlap-removal
 * potential of the initial-configuration generator.  WCA49_50 is the
 * pseudo-hard-sphere potential_wca49_50_hst (potentials.cu:9-23), cutoff
 * (50/49) sigma_ij; WCA49_50_AO adds the Asakura-Oosawa depletion branch of
 * potential_wca49_50_AO_hst (potentials.cu:27-52) beyond r = sigma_ij, cutoff
 * sigma_ij (1 + sigma_ij / sigma_pol), where that branch's energy and force
 * reach zero.  The reference only holds them as commented-out call sites
 * (get_forces.cu:44-49,119). */
enum {
  BD_POT_WCA = 0,
  BD_POT_WCA_MODIFIED = 1,
  BD_POT_HARMONIC = 2,
  BD_POT_WCA49_50 = 3,
  BD_POT_WCA49_50_AO = 4
};
/* temperature protocols (set_temp.cu:7-23) */
enum { BD_TEMP_LINEAR = 0, BD_TEMP_SINE = 1 };
enum { BD_SPECIES_BIG = 0, BD_SPECIES_SML = 1 };

/* One system (one replica).  FP64 re-expression of parametros
 * (encabezados.h:40-75): the derived cell fields live in bd_geometry.
 * Particles are stored big species first (main.cu:387-393). */
typedef struct bd_params {
  int32_t n_big;          /* ngrain_big */
  int32_t n_sml;          /* ngrain_sml */
  int32_t potential;      /* BD_POT_* */
  int32_t ntags_big;      /* cell capacity, big grid (1..32) */
  int32_t ntags_sml;      /* cell capacity, small grid (1..32) */
  int32_t temp_protocol;  /* BD_TEMP_* */
  int32_t rng_lanes;      /* rannew64 lanes; 0 = one lane per particle */
  int32_t reserved;
  double side;            /* box edge L */
  double sigma_big;
  double sigma_sml;
  double dt;
  double kappa;           /* stiffness of BD_POT_HARMONIC */
  double T0, Tf, t0, tf;  /* temperature protocol (main.cu:98-104) */
  double period;          /* sine protocol only */
  uint64_t seed;          /* master seed of the rannew64 lanes */
  int64_t first_step;     /* step index ni of the first update (-ntrans) */
  /* BD_POT_WCA49_50(_AO) only: U = pot_inv_t (pot_amp (s^50 - s^49) + 1),
   * s = sigma/r, for r below the cutoff (below sigma with _AO).  The
   * reference's commented values are pot_amp = 50 (50/49)^50 and
   * pot_inv_t = 0.678575 (get_forces.cu:46-48). */
  double pot_amp;
  double pot_inv_t;
  /* BD_POT_WCA49_50_AO only: polymer diameter and volume fraction of
   * U = -phi_pol (e^3 - 1.5 e^2 x + 0.5 x^3), e = 1 + sigma/sigma_pol,
   * x = r/sigma, for sigma < r < e sigma (potentials.cu:43-50) */
  double sigma_pol;
  double phi_pol;
} bd_params;

/* Cell geometry derived from bd_params (main.cu:259-298 with the cutoff of
 * the selected potential in place of sigma). */
typedef struct bd_geometry {
  int32_t ncell_big, ncell_sml;
  int32_t nrange_bb, nrange_bs, nrange_sb, nrange_ss;
  double cell_side_big, cell_side_sml;
  double rc_bb, rc_bs, rc_ss; /* pair cutoffs */
} bd_geometry;

/* Per-sample observables of one replica (fself.cu:3-21, main.cu:627-638,
 * 776-806). */
typedef struct bd_obs {
  double fself_big;   /* <sin(q|dr|)/(q|dr|)> over the big species */
  double msd_big, msd_sml;
  double epot_big, epot_sml; /* sum(0.5*pot)/N per species */
  double virial;      /* sum of per-particle virials */
  double z;           /* 1 + virial/(3 N T) */
  double temperature; /* T after the last update (pars.temp_set) */
} bd_obs;

typedef struct bd_sys bd_sys;

/* ---- geometry and lifetime ---- */
int bd_compute_geometry(const bd_params* p, bd_geometry* g);
/* Create an engine holding n_replicas independent systems (1 = the plain
 * single system).  Replica r uses params[r]. */
int bd_create(const bd_params* params, int32_t n_replicas, bd_sys** out);
int bd_destroy(bd_sys* s);
int bd_num_replicas(const bd_sys* s, int32_t* out);
int bd_get_geometry(const bd_sys* s, int32_t replica, bd_geometry* g);

/* ---- state transfer (host buffers owned by the caller; xyz is N x 3) ---- */
int bd_set_positions(bd_sys* s, int32_t replica, const double* xyz);
int bd_get_positions(bd_sys* s, int32_t replica, double* xyz);
int bd_get_raw_positions(bd_sys* s, int32_t replica, double* xyz);
int bd_get_forces(bd_sys* s, int32_t replica, double* f, double* vir, double* pot);
int bd_set_forces(bd_sys* s, int32_t replica, const double* f);
/* raw := pos, ini := pos for every replica (main.cu:497-506, at ni == 0) */
int bd_mark_origin(bd_sys* s);
int bd_get_step(const bd_sys* s, int64_t* ni);

/* ---- the hot path ---- */
/* Bin every replica into its two cell grids (cell_locate.cu:155-209). */
int bd_cell_locate(bd_sys* s);
/* Export one grid in the reference "tags" layout: nocup[ncell^3] and
 * cell_vec[ncell^3 * ntags] (species-local indices, -1 in empty slots). */
int bd_get_cells(bd_sys* s, int32_t replica, int32_t species, int32_t* nocup, int32_t* cell_vec);
/* Largest occupancy of a cell in each grid of one replica after the last
 * binning: the nocup maxima main.cu:641-650 scans on the host for its log
 * line (the occupancies only, no cell lists). */
int bd_max_occupancy(bd_sys* s, int32_t replica, int32_t* big, int32_t* sml);
/* Forces, virials and potentials from the current cells (get_forces.cu). */
int bd_compute_forces(bd_sys* s);
/* Neighbour lists in traversal order (test hook): counts[N], list[N*maxn]
 * holding particle indices (big first). */
int bd_neighbours(bd_sys* s, int32_t replica, int32_t maxn, int32_t* counts, int32_t* list);
/* One Ermak-McCammon update of every replica at the current step with the
 * protocol temperature; advances the rannew64 lanes and the step index. */
int bd_update(bd_sys* s);
/* nsteps full steps: update -> temperature -> cell locate -> forces
 * (main.cu:494-586); replayed through captured HIP graphs.
 * Errors: cell overflow (BD_E_CELL_OVERFLOW) and non-finite positions
 * (BD_E_NAN) are raised on the device and read once, when the call ends; the
 * message names the replica and the step index of the first update (or
 * binning) that raised it, and the whole call's later steps ran on that
 * state.  A replica's error stays raised (every later call fails) until its
 * state is replaced by bd_set_positions or bd_checkpoint_load. */
int bd_step(bd_sys* s, int32_t nsteps);
/* Builds (and caches) the graphs a bd_step(s, nsteps) call with the current
 * settings would replay, without running any step (no reference
 * counterpart): lets a caller keep graph instantiation out of a timed or
 * latency-critical call. */
int bd_prepare_step(bd_sys* s, int32_t nsteps);

/* Launch mode of bd_step (no reference counterpart): steps per captured
 * HIP graph G, 1..256 (default 16), or 0 to enqueue every kernel eagerly (for
 * tracing tools that do not follow graph launches).  Without the fused
 * update a call replays floor(nsteps / G) graphs of G whole steps and runs
 * the remainder eagerly.  With the fused update a call of more than G steps
 * runs its first update and last forces eagerly and replays the nsteps - 1
 * (forces + next move, binning) pairs G at a time, the remainder eagerly; a
 * call of at most G steps uses the whole-step path.  The two graphs are
 * cached side by side, so alternating call lengths rebuild neither.
 * Results are bit-identical in every mode. */
int bd_set_graph_steps(bd_sys* s, int32_t steps);
int bd_get_graph_steps(const bd_sys* s, int32_t* steps);

/* Persistent launch mode of bd_step (no reference counterpart): steps > 0
 * runs up to `steps` whole steps per cooperative launch of one kernel whose
 * co-resident workgroups sweep the update, binning and force phases with
 * grid-wide barriers between them; 0 (default) returns to graph / eager
 * launches.  Bit-identical results; reference force order only (BD_E_STATE
 * from bd_step otherwise). */
int bd_set_persistent(bd_sys* s, int32_t steps);
int bd_get_persistent(const bd_sys* s, int32_t* steps);

/* Force-kernel shape (no reference counterpart; the autotuner's knobs):
 * brick = 1..4 selects the LDS-staged kernel over brick^3 big-grid cells,
 * 0 the thread-per-particle row walk.  Every choice gives bit-identical
 * results.  The 49-50 potentials (BD_POT_WCA49_50, _AO) run the slab walk
 * at deferral depth 16 whatever the shape settings (bricks, walks, pass
 * split, fused update) say; the settings are kept and reported. */
int bd_set_brick(bd_sys* s, int32_t brick);
int bd_get_brick(const bd_sys* s, int32_t* brick);

/* Deferral depth of the row walk (no reference counterpart): 0 evaluates
 * each pair term where it is found; 16 or 32 (default 16) records the
 * partners inside the cutoff in LDS and evaluates them in the same order
 * afterwards.  Bit-identical results at every depth.  BD_E_ARG for other
 * values, or for a nonzero depth when a grid holds 2^26 slots or more. */
int bd_set_defer_depth(bd_sys* s, int32_t depth);
int bd_get_defer_depth(const bd_sys* s, int32_t* depth);

/* Walk of the thread-per-particle force kernel (no reference counterpart):
 * BD_WALK_ROWS loops over each stencil row's slot range; BD_WALK_SLABS
 * (default) lists the ranges of all rows with one di in LDS and loops over
 * them as one stream (fewer idle lanes when range lengths differ between
 * particles); BD_WALK_CHUNK2 / _CHUNK4 walk the same stream with the range
 * bounds of a slab loaded together and the candidates loaded 2 / 4 at a
 * time (fewer exposed memory latencies); BD_WALK_FLAT lists the ranges of
 * a slab for the 64 particles of a wavefront, stages all their candidates
 * in one LDS array and tests them 64 at a time whoever owns them, recording
 * the partners inside the cutoff per particle in visiting order (a device
 * buffer of 32 records per particle, allocated on first use; a particle
 * with more partners walks its stencil again), then evaluates them in that
 * order; it needs replicas below 2^22 slots and ntags <= 12 (BD_E_ARG
 * otherwise).  Same pairs, order and arithmetic: bit-identical. */
enum {
  BD_WALK_ROWS = 0,
  BD_WALK_SLABS = 1,
  BD_WALK_CHUNK2 = 2,
  BD_WALK_CHUNK4 = 3,
  BD_WALK_LIST = 4,
  BD_WALK_FLAT = 5
};
int bd_set_walk(bd_sys* s, int32_t walk);
int bd_get_walk(const bd_sys* s, int32_t* walk);

/* Records per particle that BD_WALK_FLAT keeps (1..32, default 32; no
 * reference counterpart): a particle with more partners inside the cutoff
 * walks its stencil again with direct evaluation, bit-identical.  Small
 * values exercise that path.  BD_E_ARG outside 1..32. */
int bd_set_flat_depth(bd_sys* s, int32_t depth);

/* Neighbour lists (BD_WALK_LIST, no reference counterpart): each particle
 * lists every particle within rc + skin.  A move that takes some particle of
 * a replica more than skin/2 from its position at the last build flags the
 * replica: on the next force evaluation that replica walks its cells while
 * a build kernel, on a second stream forked from and joined back into the
 * step, rebuilds its lists for the evaluations after.  Between rebuilds a
 * step tests only the listed candidates, and sorts the partners inside the
 * cutoff into the cell walk's order (get_forces.cu:21-558, cells in
 * (di, dj, dk) order, ids ascending) before adding their terms:
 * bit-identical to the cell walks.  Replicas whose grids are too small for
 * the list stencil, and particles with more than 48 candidates or 32
 * partners, walk their cells.  The update runs as its own kernel with this
 * walk, and bd_autotune does not choose it (set it explicitly).
 * bd_set_list_skin: 0 < skin <= 2 (default 0.3, in the unit of sigma);
 * BD_E_ARG otherwise. */
int bd_set_list_skin(bd_sys* s, double skin);
int bd_get_list_skin(const bd_sys* s, double* skin);

/* Stencil passes over wavefronts (no reference counterpart): 0 (default)
 * walks both passes of a particle (main.cu:576-585) in one lane; 1 walks the
 * second half of the particle's stencil slabs (in visiting order) on a second
 * wavefront at the same time, which only records its partners for the first
 * wavefront to evaluate after its own;
 * 2 walks a particle's stencil slabs on four wavefronts (the first takes
 * the same-species slabs up to the centre, the other three the remaining
 * slabs in nearly equal runs), the other three evaluating the terms of their
 * first partners into LDS for the first to add in order.  More wavefronts for the same work, for systems too small to
 * fill the GPU.  Takes effect with a slab walk (BD_WALK_SLABS, _CHUNK2,
 * _CHUNK4) at a nonzero deferral depth.  With BD_WALK_FLAT, 1 and 2 give
 * each wavefront 32 and 16 particles instead of 64 (the 64 lanes still test
 * the staged candidates together).  Bit-identical.  BD_E_ARG for values
 * other than 0, 1, 2. */
int bd_set_pass_split(bd_sys* s, int32_t split);
int bd_get_pass_split(const bd_sys* s, int32_t* split);

/* Fused update (no reference counterpart): bd_step runs the update of each
 * step after the first (update_ermak.cu:91-160) inside the force kernel of
 * the step before, one launch fewer per step; same operations on the same
 * values, bit-identical.  It can apply to the slab walks (BD_WALK_SLABS,
 * _CHUNK2, _CHUNK4, _FLAT) in reference order when every replica has one
 * particle per rannew64 lane (rng_lanes 0 or >= N).  mode 0: never;
 * 1: wherever it can; 2 (default): the flat walk, and the other slab walks
 * with a pass split at a nonzero deferral depth (where it is faster).
 * bd_get_fused_update reports whether it applies with the current settings
 * (0 or 1). */
int bd_set_fused_update(bd_sys* s, int32_t mode);
int bd_get_fused_update(const bd_sys* s, int32_t* on);

/* Autotuner (no reference counterpart): measures the force kernel for every
 * shape (row, slab and chunked slab walks at deferral depths 0, 16 and 32,
 * the slab walks also with the pass split; bricks 1..4; the flat walk where
 * it applies; all bit-identical by construction) with forces-only launches on
 * the current state, keeps the
 * fastest and returns its brick in *chosen (the walk, depth and split are
 * read with bd_get_walk / bd_get_defer_depth / bd_get_pass_split).  With
 * cache_path != NULL the choice is looked up in / appended to that text
 * file as "key brick depth walk split", keyed by (replicas, N,
 * phi, potential, order). */
int bd_autotune(bd_sys* s, int32_t trial_launches, const char* cache_path, int32_t* chosen);

/* Force summation order (no reference counterpart).
 * BD_ORDER_REFERENCE (default): each particle's pair terms are added in the
 *   reference's traversal order (get_forces.cu:360-416, 495-550), so forces
 *   are bit-identical to the FP64 restatement of the reference.
 * BD_ORDER_SPLIT4 / _SPLIT8: 4 or 8 lanes share a particle's stencil and
 *   their partial sums are combined in a fixed order: deterministic, equal to
 *   the reference order to rounding (~1e-15 relative), faster on small
 *   systems.  Needs grids of at least 2*nrange+3 cells; BD_E_ARG with the
 *   49-50 potentials, which run in the reference order only. */
enum { BD_ORDER_REFERENCE = 1, BD_ORDER_SPLIT4 = 4, BD_ORDER_SPLIT8 = 8 };
int bd_set_force_order(bd_sys* s, int32_t order);
int bd_get_force_order(const bd_sys* s, int32_t* order);

/* Bit-exact checkpoint of the whole engine state (no reference counterpart;
 * the reference's decade dumps, main.cu:809-835, keep only %f positions):
 * positions, raw and initial positions, forces, virials, potentials, the
 * rannew64 lanes, the step index and any waiting-time origins.  Load into an
 * engine created with the same parameters: the run continues exactly as the
 * saved one would have (cells are rebuilt on the next step).  BD_E_ARG for a
 * file that is unreadable, truncated or made by an engine with other
 * parameters. */
int bd_checkpoint_save(bd_sys* s, const char* path);
int bd_checkpoint_load(bd_sys* s, const char* path);

/* Measurement hook (no reference counterpart): runs nsteps steps with the
 * kernels launched one by one on the engine's stream and HIP events around
 * each; adds the elapsed milliseconds of the update, cells and forces kernels
 * into ms[0], ms[1], ms[2]. */
int bd_kernel_times(bd_sys* s, int32_t nsteps, double* ms);
/* Timing of the fused path (no reference counterpart): nsteps >= 2 steps as
 * bd_step runs them with the fused update, kernels launched one by one with
 * HIP events around each; adds the milliseconds of the nsteps cells kernels
 * into ms[0] and of the nsteps - 1 force kernels that also move the particles
 * into ms[1].  BD_E_STATE when the fused update does not apply. */
int bd_kernel_times_fused(bd_sys* s, int32_t nsteps, double* ms);
/* HBM bandwidth probe (no reference counterpart): a grid-stride double4 copy
 * between two device buffers of `bytes` each (>= 1 MiB), timed over `reps`
 * launches after a warm-up one by runtime-stamped events; *gbs = bytes read
 * plus bytes written per second, in GB/s (the achievable peak beside the
 * data-sheet figure). */
int bd_stream_bandwidth(int64_t bytes, int32_t reps, double* gbs);
/* Grid-barrier probe (no reference counterpart): one cooperative launch of
 * `workgroups` 64-thread workgroups that pass `barriers` grid-wide barriers
 * and nothing else, after a warm-up launch; *us = microseconds per barrier
 * (runtime-stamped events around the launch).  *max_workgroups (may be NULL)
 * receives how many such workgroups are co-resident on the device; more
 * than that is BD_E_ARG, since a cooperative grid must be co-resident.  The
 * cost that the persistent step kernel pays three times per step. */
int bd_grid_barrier_cost(int32_t workgroups, int32_t barriers, double* us,
                         int32_t* max_workgroups);
/* Launch-floor probe (no reference counterpart): a kernel of `workgroups`
 * workgroups of `threads` threads (64..1024, a multiple of 64) that does next
 * to nothing, launched `reps` times after a warm-up one, each timed by the
 * start and end stamps the runtime puts on the kernel itself (as
 * bd_kernel_times does); *us = mean microseconds per launch: the part of
 * every kernel time that is not the kernel's work. */
int bd_launch_floor(int32_t workgroups, int32_t threads, int32_t reps, double* us);

/* ---- rannew64 lanes ---- */
/* State of one lane of a replica: ring[250] (slot order) and the LCG word;
 * *draws = number of draws taken so far by every lane. */
int bd_get_rng_lane(bd_sys* s, int32_t replica, int32_t lane, uint64_t* ring250,
                    uint64_t* lcg, uint64_t* draws);
/* Stand-alone stream test: `lanes` independent rannew64 streams with
 * idum = bd_lane_seed(seed, lane); writes ndraws outputs of each lane,
 * lane-major (out[lane*ndraws + k]). */
int bd_rng_stream(uint64_t seed, int32_t lanes, int32_t ndraws, uint64_t* out);
uint64_t bd_lane_seed(uint64_t seed, uint64_t lane);

/* ---- initial configurations (no reference counterpart: run-fself.sh:131-138
 * runs external set_init_config / eliminate_overlap binaries that the
 * reference does not contain) ---- */
/* Uniform positions in [0, side)^3 (N x 3) from one rannew64 stream with
 * idum = bd_lane_seed(seed, 0); host only, no device work. */
int bd_random_positions(int32_t n, double side, uint64_t seed, double* xyz);

/* ---- measurements ---- */
/* Ordered-pair distance histograms bb, bs, sb, ss (get_gder.cu:69-127),
 * integer counts added into counts[4 * nbins]. */
int bd_gder(bd_sys* s, int32_t replica, int32_t nbins, double range, uint64_t* counts);
int bd_observables(bd_sys* s, int32_t replica, double q, bd_obs* out);
/* The same histograms for every replica in one launch:
 * counts[r * 4 * nbins + ...] (overwritten, not added). */
int bd_gder_all(bd_sys* s, int32_t nbins, double range, uint64_t* counts);
/* bd_gder without any device allocation (no counterpart in the reference):
 * for callers that cannot allocate device memory, such as code under HIP
 * graph capture, tight serving loops or harnesses that forbid hipMalloc.
 * The histogram is a fixed 16 KiB device array per GPU (4 x 512 bins),
 * loaded with counts[], added to and copied back: the same result as
 * bd_gder.  nbins above 512 is BD_E_ARG.  Every call on a GPU shares that
 * array, so calls that overlap on one GPU (several threads or handles) race. */
int bd_gder_noalloc(bd_sys* s, int32_t replica, int32_t nbins, double range, uint64_t* counts);

/* Waiting-time origins (no reference counterpart: the reference re-runs the
 * binary per decade, run-fself.sh:149-197).  bd_set_origin stores the
 * current raw positions of every replica in `slot`; bd_fself_origins
 * computes, in one launch, F_s(q) of the big species and the MSD of both
 * species of every replica against origins 0..nslots-1:
 * out[(r * nslots + k) * 3 + {0: F_s, 1: msd_big, 2: msd_sml}]. */
enum { BD_MAX_ORIGINS = 8 };
int bd_set_origin(bd_sys* s, int32_t slot);
int bd_fself_origins(bd_sys* s, double q, int32_t nslots, double* out);

/* ---- force contraction (Contracting_Vector3D.cuh:39-75) ---- */
typedef struct bd_cv3d bd_cv3d;
/* Contracting_Vector3D(bins, range) (Contracting_Vector3D.cu:11-55) */
int bd_cv3d_create(int32_t bins, double range, bd_cv3d** out);
int bd_cv3d_destroy(bd_cv3d* cv);
/* add_statistics(rr, ff, n, side) (Contracting_Vector3D.cu:57-164) on the
 * positions and forces of one replica of an engine; periodic != 0 applies
 * the replica's box as `side`. */
int bd_cv3d_add(bd_cv3d* cv, bd_sys* s, int32_t replica, int32_t periodic);
/* the same on host arrays (xyz, f: n x 3); side <= 0 means no images */
int bd_cv3d_add_host(bd_cv3d* cv, const double* xyz, const double* f, int32_t n, double side);
/* calculate_vector (Contracting_Vector3D.cu:166-297): BD_E_STATE with no samples */
int bd_cv3d_solve(bd_cv3d* cv);
/* get_vector (Contracting_Vector3D.cu:299-304): BD_E_STATE before a solve */
int bd_cv3d_get(const bd_cv3d* cv, double* out);
/* test hook: the accumulated normal equations, ata[3*bins*bins], atb[3*bins] */
int bd_cv3d_get_normal(bd_cv3d* cv, double* ata, double* atb);
int bd_cv3d_samples(const bd_cv3d* cv, int32_t* out);

/* ---- learned surrogate (no reference counterpart) ----
 * An ensemble of `members` MLPs n_in -> hidden -> hidden -> n_out with GELU,
 * BF16 weights/activations on the matrix cores with FP32 accumulation, FP32
 * master weights and Adam.  Parameters per member, FP32, row-major (x @ W):
 * W1[n_in][hidden] b1[hidden] W2[hidden][hidden] b2[hidden]
 * W3[hidden][n_out] b3[n_out]. */
typedef struct bd_mlp bd_mlp;
int bd_mlp_create(int32_t members, int32_t n_in, int32_t hidden, int32_t n_out, uint64_t seed,
                  bd_mlp** out);
int bd_mlp_destroy(bd_mlp* s);
int bd_mlp_num_params(const bd_mlp* s, int64_t* per_member);
int bd_mlp_get_params(bd_mlp* s, float* params);
int bd_mlp_set_params(bd_mlp* s, const float* params);
/* y[m][b][o] of every member for rows x[b][i] */
int bd_mlp_forward(bd_mlp* s, const float* x, int32_t batch, float* y);
/* mean-squared-error gradients of every member on one batch (test hook) */
int bd_mlp_gradients(bd_mlp* s, const float* x, const float* t, int32_t batch, float* grads,
                     float* loss);
/* test hook: member `member`'s intermediates of the last bd_mlp_gradients or
 * bd_mlp_train_step call on `batch` rows, as floats in this order: z1, h1,
 * z2, h2 [batch][hidden]; y, dy, dyb [batch][n_out]; dz2, dz2b, dz1, dz1b
 * [batch][hidden] (h*, dyb and dz*b are the BF16 matrix-core operands, the
 * others FP32): 8 batch hidden + 3 batch n_out values. */
int bd_mlp_trace(bd_mlp* s, int32_t member, int32_t batch, float* out);
int bd_mlp_train_step(bd_mlp* s, const float* x, const float* t, int32_t batch, float lr,
                      float* loss);
/* device-resident training set and the full Adam loop over it */
int bd_mlp_set_data(bd_mlp* s, const float* x, const float* t, int32_t n);
int bd_mlp_train(bd_mlp* s, int32_t epochs, int32_t batch, float lr, uint64_t seed,
                 float* losses);
/* ensemble mean and standard deviation over the members */
int bd_mlp_predict(bd_mlp* s, const float* x, int32_t n, float* mean, float* stdev);

/* ---- errors ---- */
const char* bd_last_error(void);
const char* bd_strerror(int code);

#ifdef __cplusplus
}
#endif

#endif /* BD_ABI_H */
