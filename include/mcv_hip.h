/*
 * mcv_hip.h -- C-ABI of the MI355X (gfx950) MCV shallow-water time step.
 *
 * Drop-in for the reference's time-integration seam: a host program that
 * owns the namelist, mesh, initial state and output (the reference's
 * MCV_SW.F90) hands the mesh and the initial state over once, then advances
 * the state on the GPU and reads it back at output times.
 *
 * Arrays use the reference layout: column-major (i, j, panel), the halo
 * included, i and j in [1-H, nPV+H] with H = dof-1 and nPV = n*H+1, panel in
 * [1, 6] (stat_mod.F90:7-45).  2x2 matrices are (2, 2, i, j, panel)
 * column-major (mesh_mod.F90:20-23).  The library copies what it needs during
 * each call and never keeps a host pointer after the call returns.
 *
 * Every entry point returns 0 on success or a negative MCV_E* code; the
 * message of the last failure on the calling thread is mcv_hip_last_error().
 * No entry point calls exit() or abort().  A handle is used by one host
 * thread at a time.
 */
#ifndef MCV_HIP_H
#define MCV_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MCV_ABI_VERSION 1

enum {
  MCV_OK = 0,
  MCV_EINVAL = -1,     /* bad argument or configuration */
  MCV_ENOMEM = -2,     /* device or host allocation failed */
  MCV_EDEVICE = -3,    /* HIP runtime error */
  MCV_ENONFINITE = -4, /* NaN/Inf found in the FP64 mass/energy integrals */
  MCV_ESTATE = -5      /* call not valid in the handle's current state */
};

/* time_settings%integral_scheme (parameters_mod.F90:208-214) */
enum { MCV_RK3_TVD = 3, MCV_RK4 = 4 };

/* arithmetic of the step */
enum {
  MCV_FP64 = 0, /* the reference's FP64 arithmetic, bit for bit */
  MCV_FP32C = 1 /* FP32 bulk with FP64 metrics, ghost weights and integrals */
};

/* Earth radius, gravity and the rotation rate enter the step only through the
 * host's arrays (metric terms and A/IA carry the radius, f carries Omega, the
 * state is the geopotential phi = g h), so the configuration needs none of
 * them; the halo width is H = dof - 1. */
typedef struct mcv_config {
  int32_t abi_version; /* MCV_ABI_VERSION */
  int32_t dof;         /* 3 (MCV3) or 4 (MCV4): constants_mod.F90:4-10 */
  int32_t n;           /* cells per panel edge, Nx = 90/dx: parameters_mod.F90:148 */
  int32_t scheme;      /* MCV_RK4 or MCV_RK3_TVD */
  int32_t precision;   /* MCV_FP64 or MCV_FP32C */
  int32_t device;      /* HIP device ordinal */
  double dt;           /* time step, s (parameters_mod.F90:11) */
  double dx;           /* element width, radians (parameters_mod.F90:200) */
} mcv_config;

/* Host-computed geometry (mesh_mod.F90:18-40, ghost_mod.f90:11-18).
 * matrixIG must be symmetric bit for bit (projection_mod.F90:186 makes it so). */
typedef struct mcv_mesh {
  const double *sqrtG;      /* (nh, nh, 6) */
  const double *matrixG;    /* (2, 2, nh, nh, 6) */
  const double *matrixIG;   /* (2, 2, nh, nh, 6) */
  const double *matrixA;    /* (2, 2, nh, nh, 6) */
  const double *matrixIA;   /* (2, 2, nh, nh, 6) */
  const double *f;          /* Coriolis parameter, (nh, nh, 6) */
  const double *phi_s;      /* surface geopotential, (nh, nh, 6) */
  const double *ghost_coef; /* (nPV, H): offset of the ghost point in its cell */
  const int32_t *ghost_iref;/* (nPV, H): 1-based cell holding the ghost point */
} mcv_mesh;

/* Prognostic state with its diagnosed companions (stat_mod.F90:7-20).
 * phiG must equal phi*sqrtG bit for bit (test_case_mod.F90:30). */
typedef struct mcv_state {
  double *phi, *phiG, *u, *v, *contraU, *contraV; /* each (nh, nh, 6) */
} mcv_state;

typedef struct mcv_handle mcv_handle;

/* Replaces the device-side half of MCV_SW.F90:36-41 (initStat, initTend and
 * the state handoff of test_case_mod.F90:30-43): uploads the mesh and the
 * full initial state, halo included.  The first stage of the first step uses
 * the given contraU/contraV and halo, as the reference does. */
int mcv_hip_init(const mcv_config *cfg, const mcv_mesh *mesh,
                 const mcv_state *initial, mcv_handle **out);

/* Replaces `call RK4(stat(new), stat(old))` / `call RK3_TVD(...)` followed by
 * switch_stat (MCV_SW.F90:58-59,76; time_scheme_mod.F90:13-142): enqueues
 * nsteps full time steps on the handle's stream and returns. */
int mcv_hip_step(mcv_handle *h, int32_t nsteps);

/* Blocks until queued steps finish, then copies the state back in the
 * reference layout (any field pointer may be NULL) and returns the FP64
 * integrals of diag_mod.F90:10-67 (mass, energy may be NULL).  Values at the
 * four halo corners of each panel are unspecified (the reference never reads
 * them).  Returns MCV_ENONFINITE if either integral is NaN or Inf. */
int mcv_hip_sync(mcv_handle *h, mcv_state *out, double *mass, double *energy);

/* Sphere (zonal, meridional) winds of the current state on every point, as
 * the reference's history output computes them (spatial_operators_mod.F90:
 * 391-405, called at MCV_SW.F90:65).  Blocks like mcv_hip_sync. */
int mcv_hip_sphere_winds(mcv_handle *h, double *zonal, double *meridional);

/* The spatial operator alone (spatial_operators_mod.F90:12-133) applied to
 * the given full state: tendencies of phiG, u, v on (nPV, nPV, 6).  Does not
 * change the handle's state.  Blocks. */
int mcv_hip_tendency(mcv_handle *h, const mcv_state *in, double *t_phiG,
                     double *t_u, double *t_v);

/* No counterpart in the reference: mcv_hip_tendency for callers that cannot
 * allocate memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid allocation).  The same kernels and results, bit for
 * bit, but every temporary lives in a fixed scratch array in the library's
 * static device memory (32 MiB per device, held until the process ends) or in
 * buffers the handle allocated at mcv_hip_init; the call allocates no device
 * or host memory.  Meshes whose input pairs plus tendencies exceed the
 * scratch (about 800k points per field: MCV3 up to n ~ 180, MCV4 up to
 * n ~ 120) are refused with MCV_EINVAL.  All handles on one device share its
 * scratch: calls that overlap on one device (from several host threads or
 * on handles whose streams run concurrently) race.  Blocks. */
int mcv_hip_tendency_noalloc(mcv_handle *h, const mcv_state *in, double *t_phiG,
                             double *t_u, double *t_v);

/* Number of completed time steps since mcv_hip_init. */
int mcv_hip_steps_done(const mcv_handle *h, int64_t *steps);

/* Measurement aid (no counterpart in the reference, whose only timer is the
 * whole-program clock of MCV_SW.F90:33,80-82): like mcv_hip_step, but records
 * a HIP event after every kernel on the handle's stream, blocks, and returns
 * the summed durations of the stage kernels and of the boundary kernels in
 * milliseconds (either pointer may be NULL). */
int mcv_hip_step_timed(mcv_handle *h, int32_t nsteps, double *stage_ms,
                       double *boundary_ms);

/* Tile-shape autotuner (no counterpart in the reference): times each of the
 * library's MCV_TILE_VARIANTS stage-kernel tile shapes for this solve (reps
 * launches each, on scratch levels, so the trajectory is not changed), keeps
 * the fastest for later steps and returns its index and, if ms_per_stage is
 * not NULL, the mean milliseconds of every variant (MCV_TILE_VARIANTS
 * doubles).  Blocks. */
#define MCV_TILE_VARIANTS 4
int mcv_hip_autotune(mcv_handle *h, int32_t reps, int32_t *tile, double *ms_per_stage);

/* Selects a stage-kernel tile shape by index (0 is the default). */
int mcv_hip_set_tile(mcv_handle *h, int32_t tile);

/* Measurement aid (no counterpart in the reference): copies `bytes` between
 * two device buffers `reps` times with a plain streaming kernel and returns
 * the mean milliseconds per copy.  Used to calibrate HBM traffic counters and
 * as the achievable-bandwidth yardstick. */
int mcv_hip_copy_probe(int32_t device, int64_t bytes, int32_t reps, double *ms_per_copy);

/* Measurement aid (no counterpart in the reference): the streaming-copy
 * yardstick.  Times copies of `bytes` between two device buffers with
 * several 16-byte loads in flight per thread (2, 4 or 8) and non-temporal
 * loads and stores, on 2, 4 or 8 workgroups per CU, reps launches each, and
 * returns the fastest mean milliseconds per copy and the variant (1..3 for
 * 2, 4, 8 loads) and workgroups per CU that reached it (either may be NULL). */
int mcv_hip_stream_probe(int32_t device, int64_t bytes, int32_t reps, double *ms,
                         int32_t *variant, int32_t *per_cu);

/* Releases the handle and its device memory.  NULL is accepted. */
int mcv_hip_free(mcv_handle *h);

/* Device bytes a solve of this configuration allocates at most (no
 * counterpart in the reference; the admission test of the batch API), not
 * counting the history buffers that the first mcv_hip_cell_averages (3 cell
 * fields) or mcv_hip_history_* call (8 point fields more) adds. */
int mcv_hip_footprint(const mcv_config *cfg, int64_t *bytes);

/* Device bytes the handle holds. */
int mcv_hip_device_bytes(const mcv_handle *h, int64_t *bytes);

/* ---------------------------------------------------------------------------
 * Batches (no counterpart in the reference, which runs one solve per
 * process: MCV_SW.F90:48-77).  A batch holds up to `capacity` solves of any
 * order, resolution, precision, scheme and time step on one device.  One
 * mcv_batch_run "stage" advances every solve in the batch by one RK stage of
 * its own scheme with one fused stage launch and one fused boundary launch
 * (plus a one-workgroup launch that advances the solves' stage counters);
 * solves admitted at different times run interleaved, each at its own
 * stage.  Each solve advances by exactly the nsteps given at admission and
 * then idles until removed.  Results are bit-identical to the same solve run
 * alone through mcv_hip_step.  One host thread per batch.
 * ------------------------------------------------------------------------- */
typedef struct mcv_batch mcv_batch;

int mcv_batch_create(int32_t device, int32_t capacity, mcv_batch **out);

/* Admits a solve (uploads mesh and initial state exactly as mcv_hip_init
 * does) that will advance nsteps time steps; returns its slot.  Returns
 * MCV_ESTATE when every slot is taken. */
int mcv_batch_add(mcv_batch *b, const mcv_config *cfg, const mcv_mesh *mesh,
                  const mcv_state *initial, int64_t nsteps, int32_t *slot);

/* Retires a solve and releases its device memory (blocks). */
int mcv_batch_remove(mcv_batch *b, int32_t slot);

/* Enqueues nstages stage launches (replays of a captured hipGraph where
 * possible) and returns. */
int mcv_batch_run(mcv_batch *b, int32_t nstages);

/* Blocks, then writes for every slot (capacity entries, either pointer may
 * be NULL) the time steps still to run and those completed, -1 for free
 * slots. */
int mcv_batch_poll(mcv_batch *b, int64_t *steps_left, int64_t *steps_done);

/* mcv_hip_sync for one slot; MCV_ESTATE while the solve is inside a step. */
int mcv_batch_sync(mcv_batch *b, int32_t slot, mcv_state *out, double *mass,
                   double *energy);

/* The slot's solve as a handle (owned by the batch: never step or free it). */
int mcv_batch_handle(mcv_batch *b, int32_t slot, mcv_handle **h);

/* Times the MCV_TILE_VARIANTS tile shapes on the batch's current solves
 * (reps stage launches each, trajectory unchanged; every solve must be at a
 * step boundary) and keeps the fastest.  Blocks. */
int mcv_batch_autotune(mcv_batch *b, int32_t reps, int32_t *tile, double *ms_per_stage);

int mcv_batch_set_tile(mcv_batch *b, int32_t tile);

/* Graph replay on (default) or off: off launches every kernel directly. */
int mcv_batch_set_graphs(mcv_batch *b, int32_t on);

/* Measurement aid: like mcv_batch_run without graphs, with HIP events around
 * every kernel; blocks and returns the summed stage-kernel and
 * boundary-kernel milliseconds. */
int mcv_batch_run_timed(mcv_batch *b, int32_t nstages, double *stage_ms,
                        double *boundary_ms);

/* Releases the batch and every solve in it.  NULL is accepted. */
int mcv_batch_free(mcv_batch *b);

/* ---------------------------------------------------------------------------
 * Diagnostics and history output (output_mod.F90, diag_mod.F90:69-93).
 * The history file is the reference's mcv_output.nc: netCDF classic format
 * with 64-bit offsets (NF90_64BIT_OFFSET), the same dimensions, variables,
 * attributes and record layout, written without libnetcdf.
 * ------------------------------------------------------------------------- */

/* Global attributes of history_init (output_mod.F90:86-106). */
typedef struct mcv_history_info {
  int32_t dof;      /* MCV_ORDER */
  int32_t n;        /* cells per panel edge */
  int32_t xhalo;    /* halo cells (must be 1: the library's arrays carry H halo points) */
  int32_t yhalo;
  int32_t case_num;
  double dx, dy;    /* radians (written in degrees, x R2D, as the reference does) */
  double dt;        /* s */
} mcv_history_info;

/* One history record (the arguments of history_write_stat,
 * output_mod.F90:200-212).  Point fields (nh, nh, 6); cell fields
 * (n+2, n+2, 6), the halo ring of cells zero (calc_VIA). */
typedef struct mcv_history_fields {
  double *u, *v, *contraU, *contraV, *phi, *phi_t, *zonal, *meridional;
  double *uC, *vC, *phiC;
} mcv_history_fields;

/* Replaces history_init (output_mod.F90:13-166, called at MCV_SW.F90:49):
 * creates (clobbers) `path` and writes the static variables.  Coordinates
 * in radians, (n+2, n+2, 6) for cells and (nh, nh, 6) for points; areaCell
 * (n+2, n+2, 6); phi_s (nh, nh, 6).  Host only (no device needed). */
int mcv_history_init(const char *path, const mcv_history_info *info, const double *lonC,
                     const double *latC, const double *lonP, const double *latP,
                     const double *areaCell, const double *phi_s);

/* Replaces history_write_stat (output_mod.F90:168-223) for host arrays:
 * writes record `time_slot` (1-based) of an existing history file.  The
 * arrays are those of a dof, n solve (points (nh, nh, 6), cells
 * (n+2, n+2, 6)); a file whose record variables have other sizes is refused
 * with MCV_EINVAL.  Used for the first record, which the reference writes
 * from the initial state (MCV_SW.F90:50). */
int mcv_history_write(const char *path, int32_t time_slot, int32_t dof, int32_t n,
                      const mcv_history_fields *rec);

/* calc_VIA (diag_mod.F90:69-93) of the current u, v, phi on the device, FP64,
 * in the reference's summation order: (n+2, n+2, 6) each, halo ring zero.
 * Blocks. */
int mcv_hip_cell_averages(mcv_handle *h, double *uC, double *vC, double *phiC);

/* The record fields the output block of MCV_SW.F90:62-70 produces after a
 * time step (addFillValue, convert_wind_P2SP, calc_VIA x 3, phi + phi_s),
 * computed on the device; any pointer in *out may be NULL.  Halo corners:
 * u, v and phi hold FillValue, zonal/meridional and contraU/V are formed
 * from those FillValues.  MCV_ESTATE before the first step.  Blocks. */
int mcv_hip_history_fields(mcv_handle *h, const mcv_history_fields *out);

/* The output block of MCV_SW.F90:62-70 as one call: mcv_hip_history_fields
 * into host staging, then mcv_history_write of record `time_slot`. */
int mcv_hip_history_write(mcv_handle *h, const char *path, int32_t time_slot);

/* Message of the last failed call on this thread ("" if none). */
const char *mcv_hip_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
