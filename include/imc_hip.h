/*
 * imc_hip.h -- C-ABI of libimc_hip.so, the MI355X (gfx950) implicit Monte Carlo
 * particle engine.  Plain pointers and sizes only; every function returns an int
 * status (0 = ok, < 0 = error, message from imc_last_error()) unless noted.
 *
 * Conventions
 *  - Host arrays passed in are copied, never retained.  Per-cell arrays are
 *    (n1, n2) C-order doubles (index i*n2 + j), the layout of the reference's
 *    NumPy material arrays (IMC/IMC2D.py:65-84).
 *  - Particles stay resident in device memory owned by the context; the census
 *    is observable through *_get_census / *_put_census.
 *  - One host thread per context; calls are ordered on the context's stream and
 *    every call that returns host data synchronises.
 *  - Random numbers: counter-based Philox4x32-10 keyed by (seed, stream tag,
 *    particle index, step, block).  The reference draws from Numba's hidden
 *    per-thread MT19937 instead (IMC/IMC2D.py:697-707,756-765).
 */
#ifndef IMC_HIP_H
#define IMC_HIP_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* geometry codes (IMC/IMC2D.py:51-52) */
#define IMC_GEOM_XY 0
#define IMC_GEOM_RZ 1
/* 1-D geometries (IMC/IMC1D.py:859 "slab" / "spherical") */
#define IMC_GEOM_SLAB 2
#define IMC_GEOM_SPHERE 3

/* arithmetic of the tracker: FP64 throughout, or FP32 geometry with FP64
 * weights, tallies and bookkeeping */
#define IMC_FP64 0
#define IMC_FP32_TRACK 1

/* volume-source kinds */
#define IMC_SRC_EMISSION 0
#define IMC_SRC_FIXED 1
#define IMC_SRC_CF 2 /* Carter-Forest emission: QV per cell, truncated-exponential times */

typedef struct imc2d_ctx imc2d_ctx;
typedef struct imc1d_ctx imc1d_ctx;
typedef struct mgimc_ctx mgimc_ctx;

/* Message of the last failed call on this thread. */
const char* imc_last_error(void);
/* Number of visible GPUs (0 if none). */
int imc_device_count(void);

/* Page-locked host memory for per-step outputs.  An output array that lies
 * in a block from imc_host_alloc receives its values from a kernel directly
 * (no staging slot, no host copy after the stream synchronises); any other
 * host pointer still works through the context's staging.  No counterpart in
 * the reference, whose outputs are NumPy arrays in pageable memory: it is
 * the host side of this library's output path.  imc_host_free must not be
 * called while a call that writes into the block is in flight. */
int imc_host_alloc(uint64_t bytes, void** out);
int imc_host_free(void* p);

/* ---- gray 2-D (xy / rz) ------------------------------------------------ */

/* Context for one mesh.  Replaces the module-level state set up by
 * IMC2D.init_simulation (IMC/IMC2D.py:1514-1588): edges, cell volumes
 * (IMC/IMC2D.py:87-106) and the particle arrays of SimulationState2D
 * (IMC/IMC2D.py:65-84). */
int imc2d_create(imc2d_ctx** out, int geometry, int nx, int ny, const double* edges1, const double* edges2,
                 int precision, uint64_t seed, int device);
void imc2d_destroy(imc2d_ctx* ctx);

/* Step number used as the Philox counter word of every draw that follows. */
int imc2d_set_step(imc2d_ctx* ctx, uint32_t step);
/* Philox key of every draw that follows (lets one context, with its device
 * buffers, run several independent realisations). */
int imc2d_set_seed(imc2d_ctx* ctx, uint64_t seed);
/* Transport launch shape: grid 0 = occupancy-derived persistent grid;
 * batch = particles a wave claims per queue visit; tally_copies = private
 * copies of the per-cell tallies (power of two, 1..64). */
int imc2d_set_launch(imc2d_ctx* ctx, int grid, int batch, int tally_copies);
/* Tracker shape (no reference counterpart: a launch parameter of this
 * library, picked by radtranbook_amd.autotune): block = 256 or 512 threads
 * per workgroup of the gray tracker (its waves share the LDS tally cache);
 * refill_min = idle lanes (1..64) a wave with active histories waits for
 * before it takes new particles from the queue (a wave with none always
 * refills).  Results do not depend on either. */
int imc2d_set_tracker(imc2d_ctx* ctx, int block, int refill_min);

/* Census access: the particle arrays of SimulationState2D
 * (weights, dir1, dir2, times, pos1, pos2, cell_i, cell_j; IMC/IMC2D.py:68-75). */
int64_t imc2d_census_size(imc2d_ctx* ctx);
int imc2d_put_census(imc2d_ctx* ctx, int64_t n, const double* w, const double* d1, const double* d2,
                     const double* t, const double* p1, const double* p2, const int64_t* ci, const int64_t* cj);
int imc2d_get_census(imc2d_ctx* ctx, int64_t n, double* w, double* d1, double* d2, double* t, double* p1,
                     double* p2, int64_t* ci, int64_t* cj);
/* Sum of census weights (the "Total Radiation Energy" of IMC/IMC2D.py:1833). */
int imc2d_census_energy(imc2d_ctx* ctx, double* out);

/* Census energy per cell (np.bincount of the weights by cell,
 * IMC/IMC2D.py:1543-1545). */
int imc2d_census_by_cell(imc2d_ctx* ctx, double* out);

/* Fleck factor and effective opacities, IMC/IMC2D.py:1646-1651:
 * beta = 4 a T^3 / c_v, f = 1/(1 + theta beta sigma c dt), sigma_s = (1-f) sigma,
 * sigma_a = f sigma.  clip=1 clips f to [0,1] (2-D); clip=0 fails if f leaves
 * [0,1] (IMC/IMC1D.py:936).  Outputs are optional. */
int imc2d_set_material(imc2d_ctx* ctx, const double* T, const double* sigma_true, const double* cv, double theta,
                       double dt, int clip, double* sigma_a_out, double* sigma_s_out);

/* Device-held material of the configs, replacing the host callables of
 * IMC2D.step when they have this form (IMC/IMC2D.py:1646-1651 sigma_a_func,
 * cv; 1800-1810 eos/inv_eos): sigma_a = sigma0 * max(T, T_floor)^-power and
 * u = c_v * T per cell.  Copies sigma0, c_v, the temperature T and the
 * internal energy density u (per cell). */
int imc2d_set_material_powerlaw(imc2d_ctx* ctx, const double* sigma0, double power, double T_floor, const double* cv,
                                const double* T, const double* u);
/* Fleck block of IMC/IMC2D.py:1646-1651 on the device-held T: sigma from the
 * power law, then f and the effective opacities as imc2d_set_material. */
int imc2d_fleck_powerlaw(imc2d_ctx* ctx, double theta, double dt, int clip);
/* Material update of IMC/IMC2D.py:1800-1810 after the step's transport (with
 * use_scalar_intensity_Tr): u += dep - E_emit / V (the emission energies of
 * the step's imc2d_sample_volume(IMC_SRC_EMISSION)), T = u / c_v,
 * T_r = (si / (a c))^(1/4).  Outputs (nullable): T, T_r, u per cell and the
 * total internal energy sum(u V). */
int imc2d_update_material(imc2d_ctx* ctx, double* T_out, double* Tr_out, double* u_out, double* internal_total_out);

/* Equilibrium census at t=0, replacing the pool: _equilibrium_sample_xy/_rz
 * (IMC/IMC2D.py:1110-1188). */
int imc2d_equilibrium_sample(imc2d_ctx* ctx, int64_t N, const double* Tr);

/* Half-Lambertian face source appended to the pool: _sample_boundary_xy/_rz
 * (IMC/IMC2D.py:157-489).  side: 0 left/rmin, 1 right/rmax, 2 bottom/zmin,
 * 3 top/zmax.  E_total: face energy a c T^4/4 A dt.  prof_cum (nullable):
 * running sum of per-face-cell emission for a position-dependent face
 * temperature (boundary_source_func, IMC/IMC2D.py:182-281). */
int imc2d_sample_boundary(imc2d_ctx* ctx, int side, int64_t n, double E_total, double dt, const double* prof_cum,
                          int nface);

/* Volume source appended to the pool: material emission
 * (IMC/IMC2D.py:1732-1764 -> _sample_source_xy 1191-1238,
 * _sample_source_rz_linear 1326-1445, _sample_source_rz_uniform 1241-1290) or
 * a fixed source (IMC/IMC2D.py:1708-1728).  Counts ceil(N E_c / E).
 * E_out (nullable): per-cell source energy; total_out: its sum; count_out:
 * particles added. */
int imc2d_sample_volume(imc2d_ctx* ctx, int kind, int64_t N, const double* QV_host, double dt, int rz_linear,
                        double* E_out, double* total_out, int64_t* count_out);

/* Transport every particle in the pool to census, escape or the weight floor:
 * _transport_particles_2d (IMC/IMC2D.py:938-1107).  reflect[4] per side;
 * weight_floor < 0 means 1e-10 * sum(w) / N (IMC/IMC2D.py:1771).  Outputs:
 * dep, si per cell (nullable), boundary loss, stats[8] = events, boundary
 * crossings, scatters, census events, captures (0), floor kills,
 * reflections, event-cap hits. */
int imc2d_transport(imc2d_ctx* ctx, double dt, const int32_t* reflect, int64_t max_events, double weight_floor,
                    double* dep_out, double* si_out, double* bl_out, int64_t* stats_out);
/* Device time of the last transport kernel in milliseconds (HIP events). */
float imc2d_last_transport_ms(imc2d_ctx* ctx);
/* imc2d_transport without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The same
 * tracker and the same results as imc2d_transport over the resident pool,
 * with every scratch buffer (tally partials, counters, pickup order,
 * weight-sum partials) in one fixed 32 MiB device array per GPU instead of
 * buffers the context grows.  dep_dev and si_dev (ncell doubles each,
 * required), bl_dev (1 double) and stats_dev (8 int64, as imc2d_transport's
 * stats) are DEVICE pointers the caller owns.  Fails with an error when the
 * scratch needed, 16 B x cells x tally copies plus 5 B per 256-particle
 * chunk and ~10 KB, exceeds 32 MiB, and on a context with the adjoint
 * replay enabled.  Calls share their GPU's array: no-allocation calls that
 * overlap on one GPU race.  Synchronises the context's stream. */
int imc2d_transport_noalloc(imc2d_ctx* ctx, double dt, const int32_t* reflect, int64_t max_events,
                            double weight_floor, double* dep_dev, double* si_dev, double* bl_dev,
                            int64_t* stats_dev);

/* ---- Carter-Forest variant (IMC/IMC2D_CarterForest.py), FP64 contexts ---- */
/* Material for the Carter-Forest step (IMC/IMC2D_CarterForest.py:948-949):
 * T (rz linear emission tilt), the true sigma_a and beta = 4 a max(T,1e-12)^3
 * / c_v per cell.  Emission then goes through imc2d_sample_volume with
 * kind IMC_SRC_CF and QV = the per-cell emission energies of
 * _sample_internal_emission_cf (876-906): the fixed-source sampler with each
 * time drawn from the truncated exponential of rate c sigma_a beta (51-56). */
int imc2d_set_material_cf(imc2d_ctx* ctx, const double* T, const double* sigma_a, const double* beta);
/* _transport_particles_2d_cf (IMC/IMC2D_CarterForest.py:346-508): analog
 * absorption with delayed re-emission or capture.  weight_floor < 0 means
 * 1e-10 sum(w)/N; fastpath_threshold, no_progress_rel_tol and
 * no_progress_max_streak as the reference's step arguments.  Outputs
 * (nullable): dep = captured energy / V and si = sum w s / (dt V) per cell,
 * boundary loss, stats[8] = events, boundary crossings, absorb-continue
 * events, census events, captures, floor kills (deposited in their cell),
 * reflections, event-cap hits. */
int imc2d_transport_cf(imc2d_ctx* ctx, double dt, const int32_t* reflect, int64_t max_events, double weight_floor,
                       double fastpath_threshold, double no_progress_rel_tol, int64_t no_progress_max_streak,
                       double* dep_out, double* si_out, double* bl_out, int64_t* stats_out);

/* One whole step of IMC/IMC2D.py:1591-1897 (step() with
 * use_scalar_intensity_Tr=True, conserve_comb_energy=False, no fixed source,
 * uniform face temperatures) on the device-held power-law material of
 * imc2d_set_material_powerlaw: the Fleck block, face sources
 * (_sample_boundary_*, Nboundary particles on each side with face_E[side] >
 * 0, face_E = a c T_b^4 / 4 x area x dt), emission (_sample_emission_*,
 * Ntarget particles), transport (_transport_particles_*), the material and
 * T_r update of IMC/IMC2D.py:1800-1810 and the comb to about Nmax particles
 * (_comb).  The same kernels and random streams as the staged entry points
 * (imc2d_fleck_powerlaw, imc2d_sample_boundary, imc2d_sample_volume,
 * imc2d_transport, imc2d_update_material, imc2d_comb) in that order, enqueued
 * as one device-resident sequence: the pool sizes the random numbers decide
 * live in device memory, so no host round trip splits the step.  use_graph
 * != 0 captures the sequence into a HIP graph on first use and replays it
 * while the parameters, the buffers and a pool bound hold (1/8 above the
 * larger of the capturing step's pool and Nmax + 3 ncell + the step's
 * sources, so steady steps replay); one graph is kept per pool-buffer
 * parity.
 * reflect[4] as imc2d_transport; rz_linear as imc2d_sample_volume.
 * Outputs (host, nullable; per cell): T, T_r, u after the update, dep
 * (deposited energy density), si (scalar intensity), E (emission energies).
 * scalars_out[6]: emission energy, boundary loss, sum(u V), pre-comb census
 * energy, post-comb census energy, transport kernel ms (device wall clock).
 * counts_out[11]: particles transported, census after the comb, emission
 * particles, then imc2d_transport's stats[8].  Synchronises the stream.
 * Refused while the adjoint replay or pinned emission counts are on. */
int imc2d_step(imc2d_ctx* ctx, double theta, double dt, int64_t Ntarget, int64_t Nboundary, const double* face_E,
               int64_t Nmax, const int32_t* reflect, int64_t max_events, int rz_linear, int use_graph, double* T_out,
               double* Tr_out, double* u_out, double* dep_out, double* si_out, double* E_out, double* scalars_out,
               int64_t* counts_out);

/* Per-cell stochastic comb to about Nmax particles: _comb (IMC/IMC2D.py:1448-1511).
 * Optional outputs: ecen_out, pre-comb census energy per cell; disc_out,
 * ecen - ecen_after per cell (the comb discrepancy of IMC/IMC2D.py:1511);
 * ecen_total_out, pre-comb census energy; total_rad_out, census energy
 * after the comb.  Times reset to 0 (IMC/IMC2D.py:1831). */
int imc2d_comb(imc2d_ctx* ctx, int64_t Nmax, double* ecen_out, double* disc_out, double* ecen_total_out,
               double* total_rad_out);

/* ---- differentiable gray 2-D step ----------------------------------------- */
/* No counterpart in the reference (its tracker is not differentiable): the
 * adjoint of the transport tallies with respect to the cell opacities, by
 * replaying every history of the transport's input pool from its Philox
 * counter.  imc2d_enable_adjoint(1) makes each imc2d_transport keep its
 * input pool; imc2d_adjoint then maps lam = dL/dE_abs (per cell, E_abs the
 * absorbed energy of the cell) to dL/dsigma_a and dL/dsigma_s per cell
 * (pathwise absorption terms, score-function terms of the scatter sampling,
 * emission weights proportional to sigma_a).  Same dt, reflect and
 * max_events as the transport; no imc2d_set_material in between. */
int imc2d_enable_adjoint(imc2d_ctx* ctx, int enable);
int imc2d_adjoint(imc2d_ctx* ctx, double dt, const int32_t* reflect, int64_t max_events, const double* lam,
                  double* gA_out, double* gS_out);
/* Emission counts per cell: out (nullable) receives those of the last
 * emission sample; fix (nullable) pins them for later samples (common random
 * numbers across perturbed runs); both null returns to the ceil rule. */
int imc2d_emission_counts(imc2d_ctx* ctx, int64_t* out, const int64_t* fix);
/* Copy the census into a device snapshot (save != 0) or back from it (an
 * error before the first save). */
int imc2d_snapshot(imc2d_ctx* ctx, int save);

/* ---- gray 1-D (slab / sphere), batched ------------------------------------ */
/* One context holds P independent problems on meshes of ncells[p] cells
 * (edges: P concatenated arrays of ncells[p]+1 ascending edges).  Particles
 * of problem p draw Philox index (p << 32) | index-within-problem, so every
 * problem reproduces a single-problem run.  Per-problem arrays below have P
 * entries, per-cell arrays sum(ncells) entries in problem order.
 * Replaces the state set up by IMC1D.init_simulation (IMC/IMC1D.py:859-891)
 * and SimulationState (IMC/IMC1D.py:54-68). */
int imc1d_create(imc1d_ctx** out, int geometry, int nprob, const int32_t* ncells, const double* edges,
                 int precision, uint64_t seed, int device);
void imc1d_destroy(imc1d_ctx* ctx);
int imc1d_set_step(imc1d_ctx* ctx, uint32_t step);
int imc1d_set_seed(imc1d_ctx* ctx, uint64_t seed);
/* Transport launch shape, as imc2d_set_launch. */
int imc1d_set_launch(imc1d_ctx* ctx, int grid, int batch, int tally_copies);
/* Tracker shape as imc2d_set_tracker (no reference counterpart): block must
 * be 256; refill_min = idle lanes (1..64) before a wave refills. */
int imc1d_set_tracker(imc1d_ctx* ctx, int block, int refill_min);

/* Census particles per problem (nullable output) and in total. */
int64_t imc1d_census_counts(imc1d_ctx* ctx, int64_t* per_problem);
/* Census arrays of SimulationState (weights, mus, times, positions,
 * cell_indices; IMC/IMC1D.py:54-68), grouped by problem: counts[p] of
 * problem p, in problem order. */
int imc1d_put_census(imc1d_ctx* ctx, const int64_t* counts, const double* w, const double* mu, const double* t,
                     const double* x, const int64_t* cell);
int imc1d_get_census(imc1d_ctx* ctx, int64_t n, double* w, double* mu, double* t, double* x, int64_t* cell,
                     int32_t* prob);

/* Census energy per cell (the np.bincount of IMC/IMC1D.py:884-886) and per
 * problem (np.sum(weights), IMC/IMC1D.py:889); outputs nullable. */
int imc1d_census_energy(imc1d_ctx* ctx, double* per_cell, double* per_problem);

/* Fleck factor (IMC/IMC1D.py:933-938), fails if f leaves [0,1]; theta and dt
 * per problem; opacity outputs nullable.  Starts a step (no source blocks). */
int imc1d_set_material(imc1d_ctx* ctx, const double* T, const double* sigma_true, const double* cv,
                       const double* theta, const double* dt, double* sigma_a_out, double* sigma_s_out);

/* Equilibrium census at t=0 replacing the pool: equilibrium_sample /
 * equilibrium_sample_spherical (IMC/IMC1D.py:220-248, 535-575).  N,
 * T_emit_floor per problem. */
int imc1d_equilibrium_sample(imc1d_ctx* ctx, const int64_t* N, const double* Tr, const double* T_emit_floor);

/* Face source block: create_boundary_particles / _spherical
 * (IMC/IMC1D.py:319-329, 744-763, placed as step() 947-975).  side 0 left /
 * inner (mu = sqrt(xi)), 1 right / outer (mu = -sqrt(xi)); n[p] particles
 * carrying E[p] in total. */
int imc1d_sample_boundary(imc1d_ctx* ctx, int side, const int64_t* n, const double* E);

/* Volume source block: emitted_particles / _spherical (IMC/IMC1D.py:251-291,
 * 629-707) for IMC_SRC_EMISSION (a c T^4 sigma_a dt V from the last
 * imc1d_set_material, zero below T_emit_floor), or sample_source /
 * _spherical (IMC/IMC1D.py:294-316, 710-741) for IMC_SRC_FIXED with QV per
 * cell.  Counts ceil(N E_c / E_p).  IMC_SRC_CF: emitted_particles /
 * _spherical of IMC/IMC1D_CarterForest.py (443-549, 890-965) with QV = the
 * Carter-Forest emission energies per cell: max(1, round(N E_c / E_p)) per
 * cell, none when E_p < 1e-30, the emission tilt of the T given to
 * imc1d_set_material_cf and times from the truncated exponential of rate
 * c sigma_a beta.  E_out per cell and Etot_out per problem nullable. */
int imc1d_sample_volume(imc1d_ctx* ctx, int kind, const int64_t* N, const double* QV_host,
                        const double* T_emit_floor, double* E_out, double* Etot_out);

/* Group census and pending source blocks into the transport pool (census,
 * then source blocks in call order, per problem: the reference's
 * concatenation order, IMC/IMC1D.py:947-1002) without tracking it; returns
 * its size (or a negative status) and fills per_problem (nullable).
 * imc1d_get_pool downloads that pool, or after imc1d_transport the
 * transported pool. */
int64_t imc1d_stage_pool(imc1d_ctx* ctx, int64_t* per_problem);
int imc1d_get_pool(imc1d_ctx* ctx, int64_t n, double* w, double* mu, double* t, double* x, int64_t* cell,
                   int32_t* prob);

/* Transport census and this step's source blocks: move_particles /
 * _move_sph_slice (IMC/IMC1D.py:133-200, 439-532).  reflect: 2 per problem;
 * weight_floor per problem (< 0: 1e-10 sum(w)/N).  Outputs (nullable): dep,
 * si per cell; leak and floor_leak per problem; stats[8] = events, boundary
 * crossings, scatters, census events, floor kills, reflections,
 * non-positive distances (the reference's AssertionError), particles
 * transported. */
int imc1d_transport(imc1d_ctx* ctx, const int32_t* reflect, const double* weight_floor, double* dep_out,
                    double* si_out, double* leak_out, double* floor_leak_out, int64_t* stats_out);
float imc1d_last_transport_ms(imc1d_ctx* ctx);

/* ---- Carter-Forest variant (IMC/IMC1D_CarterForest.py), FP64 contexts ---- */
/* Material of a Carter-Forest step (IMC/IMC1D_CarterForest.py:1177-1180): T
 * per cell (emission tilt), the true sigma_a and beta = 4 a T^3 / c_v per
 * cell, dt per problem.  Starts a new step (no source blocks). */
int imc1d_set_material_cf(imc1d_ctx* ctx, const double* T, const double* sigma_a, const double* beta,
                          const double* dt);
/* move_particles_cf / move_particles_spherical_cf
 * (IMC/IMC1D_CarterForest.py:210-332, 723-819) over the census and this
 * step's source blocks.  reflect: 2 per problem; weight_floor per problem
 * (< 0: 1e-10 sum(w)/N); max_events: event_cap_per_particle (<= 0: none);
 * fastpath_threshold as the reference's.  Outputs (nullable): dep = captured
 * energy / V and si = sum w s / (V dt) per cell, leak per problem, stats[8] =
 * events, boundary crossings, absorb-continue events, census events,
 * captures, floor kills (kept in census), reflections, event-cap hits;
 * transported_out: particles in the transport pool. */
int imc1d_transport_cf(imc1d_ctx* ctx, const int32_t* reflect, const double* weight_floor, int64_t max_events,
                       double fastpath_threshold, double* dep_out, double* si_out, double* leak_out,
                       int64_t* stats_out, int64_t* transported_out);

/* Per-problem comb to Nmax[p]: comb (IMC/IMC1D.py:770-828); times reset to 0
 * (IMC/IMC1D.py:1068).  Outputs (nullable): ecen, disc per cell; ecen_total,
 * total_rad per problem. */
int imc1d_comb(imc1d_ctx* ctx, const int64_t* Nmax, double* ecen_out, double* disc_out, double* ecen_total_out,
               double* total_rad_out);

/* ---- multigroup 2-D (xy / rz) ------------------------------------------- */
/* Per-(group, cell) arrays are (G, ncell) with cell c = i*ny + j, the layout
 * of the reference's (n_groups, nx, ny) arrays.  Replaces the state set up by
 * MG_IMC2D.init_simulation (MG_IMC/MG_IMC2D.py:1431-1543) and
 * SimulationState2DMG (MG_IMC/MG_IMC2D.py:116-136). */
int mgimc_create(mgimc_ctx** out, int geometry, int nx, int ny, const double* edges1, const double* edges2, int G,
                 const double* energy_edges, int precision, uint64_t seed, int device);
void mgimc_destroy(mgimc_ctx* ctx);
int mgimc_set_step(mgimc_ctx* ctx, uint32_t step);
int mgimc_set_seed(mgimc_ctx* ctx, uint64_t seed);
int mgimc_set_launch(mgimc_ctx* ctx, int grid, int batch, int tally_copies);
/* Tracker shape of the 2-D multigroup tracker as imc2d_set_tracker (no
 * reference counterpart): block must be 256; refill_min 1..64. */
int mgimc_set_tracker(mgimc_ctx* ctx, int block, int refill_min);
/* Copy of the particle pool on the device (no reference counterpart; lets
 * a staged step be transported several times, as the autotuner does):
 * save = 1 keeps the current pool, save = 0 puts the kept one back (an
 * error before the first save). */
int mgimc_snapshot(mgimc_ctx* ctx, int save);

/* Census: weights, dir1, dir2, times, pos1, pos2, cell_i, cell_j, groups
 * (MG_IMC2D.py:119-127). */
int64_t mgimc_census_size(mgimc_ctx* ctx);
int mgimc_put_census(mgimc_ctx* ctx, int64_t n, const double* w, const double* d1, const double* d2, const double* t,
                     const double* p1, const double* p2, const int64_t* ci, const int64_t* cj, const int32_t* g);
int mgimc_get_census(mgimc_ctx* ctx, int64_t n, double* w, double* d1, double* d2, double* t, double* p1,
                     double* p2, int64_t* ci, int64_t* cj, int32_t* g);
/* Census energy per (group, cell) (the bincounts of MG_IMC2D.py:1500-1506,
 * 2004-2011) and in total; outputs nullable. */
int mgimc_census_by_bin(mgimc_ctx* ctx, double* out, double* total_out);

/* Planck-mean Fleck factor and effective opacities (MG_IMC2D.py:1693-1719),
 * group fractions b*_g normalised over the groups or the caller's
 * emission_fractions (1866-1885), emission energies sigma_a,g b*_g a c T^4 dt V
 * with cells below T_emit_floor silent (1888-1894), and the cell CDF the
 * emission sampler draws from.  sigma: (G, ncell).  Outputs nullable:
 * sa, ss, E (G, ncell); Etot; Eplanck (per cell, Planck b*_g, no floor: the
 * particle-budget estimate of 1741-1755). */
int mgimc_set_material(mgimc_ctx* ctx, const double* T, const double* sigma, const double* cv, double theta,
                       double dt, const double* emission_fractions, double T_emit_floor, double* sa_out,
                       double* ss_out, double* E_out, double* Etot_out, double* Eplanck_out);

/* Equilibrium census at t=0 replacing the pool: _equilibrium_sample_xy_mg /
 * _rz_mg (MG_IMC2D.py:992-1095): cells by a Tr^4 V, groups from the Planck
 * spectrum at the cell's Tr. */
int mgimc_equilibrium_sample(mgimc_ctx* ctx, int64_t N, const double* Tr, double T_emit_floor);

/* Face source appended to the pool: _sample_boundary_xy / _rz
 * (MG_IMC2D.py:392-577), groups from the Planck spectrum at the face
 * temperature (Tb, or Tface[m] with the profile prof_cum of a position-
 * dependent face).  E_by_group_out (nullable, G): energy by group. */
int mgimc_sample_boundary(mgimc_ctx* ctx, int side, int64_t n, double E_total, double dt, double Tb,
                          const double* prof_cum, const double* Tface, int nface, double* E_by_group_out);

/* Fixed source appended to the pool: _sample_source_xy_mg / _rz_mg
 * (MG_IMC2D.py:1098-1234).  S: (G, ncell) energy (source x V x dt). */
int mgimc_sample_source(mgimc_ctx* ctx, int64_t N, const double* S, double dt, double* total_out);

/* Material emission appended to the pool (MG_IMC2D.py:1896-1943): N
 * particles of weight E/N from the energies of the last mgimc_set_material. */
int mgimc_sample_emission(mgimc_ctx* ctx, int64_t N, double dt, double* Etot_out);

/* Transport the pool: _transport_particles_2d_mg (MG_IMC2D.py:774-989).
 * weight_floor < 0: 1e-10 sum(w)/N (2048).  Outputs nullable: dep, si
 * (G, ncell); bl[1 + G + 4] = boundary loss total, by group, by side;
 * stats[8] as the reference's. */
int mgimc_transport(mgimc_ctx* ctx, double dt, const int32_t* reflect, int64_t max_events, double weight_floor,
                    double* dep_out, double* si_out, double* bl_out, int64_t* stats_out);
float mgimc_last_transport_ms(mgimc_ctx* ctx);

/* (group, cell) comb: _comb_mg (MG_IMC2D.py:1314-1378); times reset to 0.
 * Outputs nullable: ecen, disc, post (G, ncell); ecen_total, total_rad. */
int mgimc_comb(mgimc_ctx* ctx, int64_t Nmax, double* ecen_out, double* disc_out, double* post_out,
               double* ecen_total_out, double* total_rad_out);

/* Per-cell sums over groups of a (G, ncell) device array, accumulated from
 * group 0 upwards as numpy sums axis 0 (the group sums of MG_IMC2D.py:
 * 1978-1979,2013: deposition, emission and comb discrepancy per cell)
 * without moving the (G, ncell) array to the host.  which: 0 emission
 * energies of the last mgimc_set_material, 1 deposition of the last
 * mgimc_transport, 2 pre- minus post-comb census energy of the last
 * mgimc_comb called with disc_out or post_out.  out: ncell doubles. */
int mgimc_group_sum(mgimc_ctx* ctx, int which, double* out);

/* ---- multigroup 1-D spherical -------------------------------------------- */
/* A mgimc_ctx over I radial shells (edges: I+1 radii), the state of
 * MG_IMC1D.init_simulation (MG_IMC/MG_IMC1D.py:850-914) and
 * SimulationState1DMG (116-138).  Per-(group, cell) arrays are (G, I).  The
 * census calls above apply with d1 = mu, p1 = r, ci = cell (d2, p2, cj
 * unused); mgimc_set_material is the Planck-mean Fleck factor of
 * MG_IMC1D.py:987-1002 and the emission energies of 600-612 (pass
 * emission_fractions = NULL); mgimc_destroy, mgimc_set_step/seed/launch and
 * mgimc_last_transport_ms apply unchanged. */
int mgimc_create_sphere(mgimc_ctx** out, int I, const double* edges, int G, const double* energy_edges,
                        int precision, uint64_t seed, int device);

/* Equilibrium census at t=0 replacing the pool: _equilibrium_sample_spherical_mg
 * (MG_IMC1D.py:514-566): max(1, round(N E_i / E)) particles per emitting
 * cell, groups from the Planck spectrum at Tr_i. */
int mgimc_sphere_equilibrium(mgimc_ctx* ctx, int64_t N, const double* Tr, double T_emit_floor);

/* Half-Lambertian source on the inner (side 0, mu > 0, cell 0) or outer
 * (side 1, mu < 0, cell I-1) sphere appended to the pool:
 * _create_boundary_spherical_mg (MG_IMC1D.py:653-676) as placed by step()
 * (1061-1088).  E_out: a c T^4 / 4 dt 4 pi R^2. */
int mgimc_sphere_face(mgimc_ctx* ctx, int side, int64_t n, double T, double dt, double* E_out);

/* Volume source by (group, cell) bin appended to the pool:
 * _emitted_particles_spherical_mg (MG_IMC1D.py:569-650, kind IMC_SRC_EMISSION,
 * energies of the last mgimc_set_material, max(1, round(.)) per emitting bin)
 * or _sample_source_spherical_mg (679-743, kind IMC_SRC_FIXED, QV (G, I)
 * source energies this step, round(.) per bin).  total_out: energy added. */
int mgimc_sphere_volume(mgimc_ctx* ctx, int kind, int64_t N, const double* QV, double dt, double* total_out);

/* Transport the pool: _move_particles_spherical_mg (MG_IMC1D.py:368-507).
 * weight_floor < 0: 1e-14 sum(w)/N (1138).  Outputs nullable: dep, si
 * (G, I); bl (G) escaping weight by group; stats[8] = events, boundary,
 * scatter, census, 0, floor kills, reflections, non-positive distances. */
int mgimc_sphere_transport(mgimc_ctx* ctx, double dt, const int32_t* reflect, double weight_floor, double* dep_out,
                           double* si_out, double* bl_out, int64_t* stats_out);

/* Census after transport (MG_IMC1D.py:1147-1215 with _comb_mg_1d 746-843):
 * Nmax > 0 combs per (cell, group) bin to about Nmax, Nmax < 0 combs to
 * |Nmax| once at least |Nmax| particles live, Nmax == 0 only drops dead
 * particles; radiation != 0 rescales each bin's survivors to its pre-comb
 * energy.  Outputs nullable: disc (G, I) = pre - post comb energy per bin
 * (radiation mode: the energy of bins left without copies, which the
 * reference drops unbooked); n_alive; combed. */
int mgimc_sphere_comb(mgimc_ctx* ctx, int64_t Nmax, int radiation, double* disc_out, int64_t* n_alive_out,
                      int* combed_out);

/* Census energy and sum w mu per (group, cell) bin (MG_IMC1D.py:1262-1276,
 * before the c / V scaling of the flux) and the total energy. */
int mgimc_sphere_census(mgimc_ctx* ctx, double* E_out, double* wmu_out, double* total_out);

#ifdef __cplusplus
}
#endif
#endif
