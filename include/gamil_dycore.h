/*
 * gamil_dycore.h -- C ABI of the MI355X barotropic time step.
 *
 * Drop-in boundary for the time-stepping path of gamil-dycore.  The reference
 * has no FFI: its seam is the argument-less Fortran facade
 *   dycore_init / dycore_restart / dycore_run / dycore_final
 * (src/dycore_mod.F90:22-25, 60-157) fed through module globals (the
 * &dycore_params namelist, src/params_mod.F90:68-98; state(1) and static%ghs
 * written by the test-case IC modules, src/data_mod.F90:19-22).  The
 * functions below are what that facade binds through iso_c_binding
 * (gamil-dycore_amd/fortran/gd_capi_mod.F90); every type is a plain C type.
 *
 * Conventions
 *   - Return value: GD_OK (0) or a GD_ERR_* code; gd_last_error() gives text.
 *   - The caller owns host buffers, the library owns device buffers; no host
 *     pointer is kept after a call returns.
 *   - Host field layout: pointer to element (i=1, j=1), longitude fastest,
 *     consecutive latitude rows `ld` elements apart (ld >= num_lon).  A
 *     Fortran halo'd array x(-1:N+2,-1:M+2) is passed as x(1,1) with
 *     ld = N+4; a dense array as ld = N.  Full-lat fields (u, gd, ghs) have
 *     num_lat rows, half-lat fields (v) num_lat-1 rows.
 *   - A handle is bound to one HIP device and one stream and is not
 *     thread-safe; several handles may coexist.
 *   - A handle holds `n_members` independent runs of one configuration
 *     (ensemble members); member 0 is the single-run case.
 */
#ifndef GAMIL_DYCORE_H
#define GAMIL_DYCORE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  GD_OK = 0,
  GD_ERR_INVALID_ARG = 1,  /* bad config / size / pointer                    */
  GD_ERR_HIP = 2,          /* HIP runtime error (device, memory, launch)      */
  GD_ERR_NAN = 3,          /* total mass or energy became NaN (diag_mod.F90:79-87) */
  GD_ERR_NOT_SUPPORTED = 4 /* option of the reference not provided here       */
};

enum { GD_SPLIT_NONE = 0, GD_SPLIT_CSP1 = 1, GD_SPLIT_CSP2 = 2, GD_SPLIT_ISP = 3 };
enum { GD_ADV_CENTER_DIFF = 0, GD_ADV_UPWIND = 1, GD_ADV_WENO = 2 };
enum { GD_PASS_ALL = 0, GD_PASS_FAST = 1, GD_PASS_SLOW = 2 };
enum {
  GD_PREC_FP64 = 0,        /* all arithmetic in FP64 (reference -fdefault-real-8) */
  GD_PREC_MIXED = 1,       /* FP32 momentum stencils; FP64 state update, stored
                              tendencies, pole sums, mass flux divergence, polar
                              filter and integrals                               */
  GD_PREC_MIXED_BF16 = 2,  /* MIXED with the polar filter of the momentum rows as
                              batched matrix-core products, BF16 inputs and FP32
                              accumulation (64 <= num_lon <= 1440)               */
  GD_PREC_MIXED_MFMA = 3   /* the same with FP32 inputs                           */
};

/* &dycore_params (src/params_mod.F90:13-98) plus the device knobs. */
typedef struct gd_config {
  int32_t num_lon;                 /* num_lon                                   */
  int32_t num_lat;                 /* num_lat (full latitudes incl. poles)       */
  double time_step_size;           /* time_step_size [s]                         */
  int32_t split_scheme;            /* split_scheme: GD_SPLIT_*                   */
  int32_t subcycles;               /* subcycles                                  */
  int32_t qcon_modified;           /* qcon_modified (logical)                    */
  int32_t uv_adv_scheme;           /* uv_adv_scheme: GD_ADV_*                    */
  double uv_adv_upwind_lon_beta;   /* uv_adv_upwind_lon_beta                     */
  double uv_adv_upwind_lat_beta;   /* uv_adv_upwind_lat_beta                     */
  int32_t use_diffusion;           /* use_diffusion (logical)                    */
  int32_t diffusion_order;         /* diffusion_order                            */
  double diffusion_coef;           /* diffusion_coef [m2/s]                      */
  int32_t use_zonal_tend_filter;   /* use_zonal_tend_filter (logical)            */
  int32_t n_cutoff;                /* number of entries in cutoff[]              */
  const int32_t *cutoff;           /* zonal_tend_filter_cutoff_wavenumber(1:n),
                                      counted from each pole                      */
  int32_t precision;               /* GD_PREC_*                                  */
  int32_t device;                  /* HIP device ordinal                         */
} gd_config;

typedef struct gd_handle gd_handle;

/* dycore_init (dycore_mod.F90:60-111): mesh_init, data_init, filter_init and
 * the integrator/split/advection selection, for n_members runs.  Also applies
 * reset_cos_lat_at_poles (dycore_mod.F90:159-173), which dycore_run performs
 * before its first use of the coefficients. */
int gd_create(const gd_config *cfg, int32_t n_members, gd_handle **out);

/* dycore_final (dycore_mod.F90:144-157). */
int gd_destroy(gd_handle *h);

/* Text of the last error on this handle (or of the last failed gd_create on
 * this thread when h is NULL). */
const char *gd_last_error(const gd_handle *h);

/* IC hand-over: the writes of the test-case modules into state(1)%{u,v,gd}
 * and static%ghs (e.g. rossby_haurwitz_wave_test_mod.F90:33-85), followed by
 * iap_transform(state(old)) (dycore_mod.F90:123, types_mod.F90:399-426) on
 * the device.  u must vanish on the two pole rows (the reference never
 * updates u there, dycore_mod.F90:207,636). */
int gd_set_state(gd_handle *h, int32_t member, const double *u, const double *v,
                 const double *gd, const double *ghs, int64_t ld);

/* diag_run (diag_mod.F90:42-89) of the current state: total mass and energy
 * for every member (arrays of n_members). */
int gd_diag(gd_handle *h, double *total_mass, double *total_energy);

/* The body of the dycore_run loop (dycore_mod.F90:131-140) for nsteps steps:
 * time_integrate (654-669: split scheme + predict_correct + diffusion),
 * time_advance and diag_run.  Outputs (any may be NULL) are laid out
 * [step][member]: total mass, total energy and the beta of the last
 * integrator call of the step (the value log_step prints).  Returns
 * GD_ERR_NAN if mass or energy became NaN; *nan_step (may be NULL) then holds
 * the 1-based first such step. */
int gd_step(gd_handle *h, int32_t nsteps, double *total_mass, double *total_energy,
            double *beta, int32_t *nan_step);

/* Ensemble bookkeeping (no counterpart in the reference, which holds one
 * run): gd_step, gd_predict_correct and gd_ordinary_diffusion advance members
 * 0..n_active-1 only (default: all); the other members keep their state.
 * Their entries in gd_step's outputs are not defined.  gd_copy_member copies
 * member src's complete device state over member dst, so a run continues bit
 * for bit in another slot. */
int gd_set_active_members(gd_handle *h, int32_t n_active);
int gd_copy_member(gd_handle *h, int32_t dst, int32_t src);

/* D2H copy of the prognostic state (u, v, gd; any pointer may be NULL) and
 * of the history diagnostics div and vor (diag_mod.F90:49-77, may be NULL). */
int gd_get_state(gd_handle *h, int32_t member, double *u, double *v, double *gd,
                 int64_t ld);
int gd_get_diag_fields(gd_handle *h, int32_t member, double *div, double *vor,
                       int64_t ld);

/* ---- operator seams (the reference's internal procedures) --------------- */

/* check_antisymmetry(tend, state) (dycore_mod.F90:794-851) of the member's
 * current state, evaluated on the device: the ten inner products of the
 * operator terms with the state (u_adv_lon, u_adv_lat, fv, u_pgf with iap u
 * and full cos; v_adv_lon, v_adv_lat, fu, v_pgf with iap v and half cos;
 * mass_div_lon, mass_div_lat with gd + ghs and full cos; the operators of
 * dycore_mod.F90:367-598 before the filter) into terms[10], the same sums of
 * magnitudes into magnitudes[10], and the four sums the reference prints,
 * which vanish to round-off for the center scheme, into sums[4]: advection,
 * u_pgf + mass_div_lon, fv - fu, v_pgf + mass_div_lat.  Any pointer may be
 * NULL.  Uses the member's corrector scratch, which holds nothing between
 * steps; the state is not changed. */
int gd_check_antisymmetry(gd_handle *h, int32_t member, double *terms,
                          double *magnitudes, double *sums);

/* space_operators(state(old), tend, pass) (dycore_mod.F90:184-365) including
 * the polar filter and its inner-product rescale; du (num_lat rows, pole rows
 * 0), dv (num_lat-1 rows), dgd (num_lat rows).  The state is not changed.
 * member must be active (GD_ERR_INVALID_ARG otherwise). */
int gd_space_operators(gd_handle *h, int32_t member, int32_t pass, double *du,
                       double *dv, double *dgd, int64_t ld);

/* One integrator call predict_correct(dt, old, new, pass)
 * (dycore_mod.F90:754-792) whose result becomes the current state of the
 * active members; beta (n_active values, may be NULL) as logged at
 * dycore_mod.F90:787. */
int gd_predict_correct(gd_handle *h, double dt, int32_t pass, int32_t qcon_modified,
                       double *beta);

/* ordinary_diffusion(dt, state) (diffusion_mod.F90:74-217) applied to the
 * current state, including its closing iap_transform. */
int gd_ordinary_diffusion(gd_handle *h, double dt);

/* filter_array_at_full_lat / _half_lat (filter_mod.F90:105-167) applied to
 * every flagged row of x (half = 0: num_lat rows; half = 1: num_lat-1 rows;
 * unflagged rows are left as they are).  Runs the device filter kernel. */
int gd_filter_rows(gd_handle *h, int32_t half, double *x, int64_t ld);

/* No counterpart in the reference.  gd_filter_rows for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): x is in DEVICE memory (num_lat rows for
 * half = 0, num_lat-1 rows for half = 1, ld elements apart) and is filtered in
 * place.  The flagged rows of member 0's table are staged in a fixed 4 MiB
 * array of the library on the handle's device, so the call allocates nothing;
 * it returns GD_ERR_INVALID_ARG when flagged rows x pitch (num_lon rounded up
 * to 32) x 8 bytes exceed 4 MiB.  stream: a hipStream_t to enqueue on,
 * returning without waiting; NULL runs on the handle's stream and waits.
 * Every call on one device shares that array: calls that overlap on one
 * device (other streams, other threads, other handles) race. */
int gd_filter_rows_noalloc(gd_handle *h, int32_t half, double *x, int64_t ld, void *stream);

/* Per-row filter table as set up by filter_init (filter_mod.F90:35-103):
 * flag (0/1) and kept cutoff wavenumber c (-1 when the row has no mask) for
 * full rows (num_lat) and half rows (num_lat-1). */
int gd_filter_table(const gd_handle *h, int32_t *full_flag, int32_t *half_flag,
                    int32_t *full_cut, int32_t *half_cut);

/* ---- initial conditions (one-time host setup, not on the time step) ---- */

/* Seeded geostrophically balanced state: height perturbation from random
 * spherical-harmonic coefficients (total wavenumber 1..nmax, coefficient
 * standard deviation ~ n^slope, tapered by sin^2(lat) so the geostrophic
 * wind stays finite at the equator, rescaled to an area-weighted rms of
 * rms_height m) on a mean depth of mean_depth m; winds in geostrophic balance
 * with it.  n_mountains > 0 adds Gaussian mountains (height uniform in
 * [500, 2000] m, half-width uniform in [10, 20] deg, centre latitude uniform
 * in [-60, 60] deg) drawn from mountain_seed; ghs = g*hs and
 * gd = g*(mean_depth + h') - ghs.  When pert_rms > 0 a second independent
 * balanced field (seed pert_seed, spectrum to pert_nmax, rms pert_rms m) is
 * added: ensemble members on a common state.  u is 0 on the pole rows. */
typedef struct gd_ic_params {
  uint64_t seed;
  int32_t nmax;
  double slope;
  double rms_height;
  double mean_depth;
  int32_t n_mountains;
  uint64_t mountain_seed;
  uint64_t pert_seed;
  int32_t pert_nmax;
  double pert_rms;
} gd_ic_params;

int gd_ic_balanced(int32_t num_lon, int32_t num_lat, const gd_ic_params *p, double *u,
                   double *v, double *gd, double *ghs, int64_t ld);

/* The reference test cases with closed forms:
 * rossby_haurwitz_wave_test_mod.F90:33-85 (R, omg, gd0) and
 * mountain_zonal_flow_test_mod.F90:28-97 (smooth_mountain = .false.). */
int gd_ic_rossby_haurwitz(int32_t num_lon, int32_t num_lat, double R, double omg,
                          double gd0, double *u, double *v, double *gd, double *ghs,
                          int64_t ld);
int gd_ic_mountain_zonal_flow(int32_t num_lon, int32_t num_lat, double *u, double *v,
                              double *gd, double *ghs, int64_t ld);

/* steady_geostrophic_flow_test_mod.F90:20-62 (Williamson test 2, alpha = 0). */
int gd_ic_steady_geostrophic(int32_t num_lon, int32_t num_lat, double *u, double *v, double *gd,
                             double *ghs, int64_t ld);

/* jet_zonal_flow_test_mod.F90:29-105 (Galewsky et al. 2004): the balanced
 * jet, plus the Gaussian height bump when perturb != 0 (the reference always
 * adds it).  The balance integral is evaluated to round-off (the reference
 * uses QUADPACK qags with epsrel 1e-3). */
int gd_ic_jet_zonal_flow(int32_t num_lon, int32_t num_lat, int32_t perturb, double *u,
                         double *v, double *gd, double *ghs, int64_t ld);

/* shallow_water_waves_test_mod.F90:96-305 (Shamir & Paldor 2016 analytic
 * waves, (n, k) = (5, 10)) at time 0.  flag: 0 Rossby wave (the reference's
 * choice), 1 / -1 eastward / westward inertia-gravity wave.  phase_speed (may
 * be NULL) receives C in rad/s. */
int gd_ic_shallow_water_waves(int32_t num_lon, int32_t num_lat, int32_t flag, double *u,
                              double *v, double *gd, double *ghs, int64_t ld,
                              double *phase_speed);

/* Polar filter table whose kept wavenumber falls toward the poles as
 * cos(lat)/cos(lat0): entry k (k = 1.. from each pole, the layout of
 * zonal_tend_filter_cutoff_wavenumber, filter_mod.F90:44-99) drives full row
 * 1+k; c_k = floor(num_lon/2 * cos(lat_{1+k}) / cos(lat0)) clipped to
 * [1, num_lon/2-1], for every row poleward of lat0.  Writes at most cap
 * entries and returns the number of entries (or -1 on bad arguments). */
int gd_cutoff_from_lat0(int32_t num_lon, int32_t num_lat, double lat0_deg, int32_t *cutoff,
                        int32_t cap);

/* ---- history and restart files (host only, no device call) ------------- */

/* Dataset metadata of one output time.  Strings may be NULL (written empty). */
typedef struct gd_io_info {
  const char *desc;           /* global attribute desc (case_desc)                   */
  const char *author;         /* global attribute author                              */
  const char *time_units;     /* time:units = "<time_units> since <start_time>"      */
  const char *start_time;     /* start_time_format, e.g. "0001-01-01T00_00_00"       */
  double time_value;          /* time(1) = elapsed seconds / time_units_in_seconds    */
  double time_step_size;      /* history: global attribute time_step_size            */
  const char *time_scheme;    /* history: global attribute time_scheme               */
  const char *split_scheme;   /* history: global attribute split_scheme              */
  int32_t subcycles;          /* history: global attribute subcycles                 */
  const char *restart_time;   /* restart: global attribute restart_time              */
  double elapsed_seconds;     /* restart: global attribute elapsed_seconds           */
} gd_io_info;

/* history_write_state (history_mod.F90:93-124) into a new NetCDF classic file
 * (NF90_CLOBBER, io_mod.F90:417-495) holding one output time: coordinates
 * lon, lat, ilon, ilat [deg] and time; u and v moved to the A grid
 * (0.5 (u_i + u_{i-1}), 0.5 (v_j + v_{j-1}) with v_0 = 0) on (lon, lat); gh =
 * gd + ghs; ghs; vor on (ilon, ilat); div on (lon, lat); te and tm.  Input
 * fields are the C-grid u (num_lat rows), v (num_lat-1), gd, ghs, div
 * (num_lat) and vor (num_lat-1) in the host layout above.  Differences from
 * the reference file, which writes through the count of each declared
 * dimension: the reference declares vor on (ilon, lat) and div on (lon, ilat)
 * (history_mod.F90:48-49) and so writes vor with an undefined last row and
 * div without its north pole row; here both keep their own grids.  The A-grid
 * v has num_lat-1 defined rows (history_mod.F90:115); its last row holds the
 * NetCDF double fill value. */
int gd_history_write(const char *path, const gd_io_info *info, int32_t num_lon, int32_t num_lat,
                     const double *u, const double *v, const double *gd, const double *ghs,
                     const double *div, const double *vor, int64_t ld, double total_energy,
                     double total_mass);

/* restart_write (restart_mod.F90:57-75): u on (ilon, lat), v on (lon, ilat),
 * gd and ghs on (lon, lat), one time record, global attributes restart_time
 * and elapsed_seconds.  The interior points are written (the reference hands
 * the whole halo'd arrays to a write of num_lon x num_lat values,
 * io_mod.F90:577-621, so its files hold the first elements of the halo'd
 * storage instead). */
int gd_restart_write(const char *path, const gd_io_info *info, int32_t num_lon, int32_t num_lat,
                     const double *u, const double *v, const double *gd, const double *ghs,
                     int64_t ld);

/* restart_read (restart_mod.F90:37-55): u, v, gd, ghs of the last time record
 * (any pointer may be NULL) and the global attributes restart_time (copied
 * into restart_time, capacity cap bytes, may be NULL) and elapsed_seconds (may
 * be NULL).  Reads NetCDF classic and 64-bit-offset files whose variables are
 * double. */
int gd_restart_read(const char *path, int32_t num_lon, int32_t num_lat, double *u, double *v,
                    double *gd, double *ghs, int64_t ld, char *restart_time, int32_t cap,
                    double *elapsed_seconds);

/* Text of the last failed history/restart call on this thread. */
const char *gd_io_last_error(void);

/* ---- serving: continuous batching of independent runs ------------------ */
/* No counterpart in the reference (a single-run program): the batching and
 * serving layer over the step above.  Runs of one (resolution, time step,
 * diffusion coefficient) class share a pool of device-resident member slots
 * sized by the HBM budget; requests are admitted when they have arrived and a
 * slot of their class is free; pools of different classes run concurrently
 * on their own streams; a run gives the same bits alone and in any mix. */

typedef struct gd_serve_request {
  int32_t num_lon, num_lat;
  double time_step_size;
  double diffusion_coef;       /* used when base->use_diffusion is set          */
  int32_t nsteps;              /* run length                                    */
  double arrival_s;            /* arrival, seconds after the stream starts      */
  const double *u, *v, *gd, *ghs;  /* initial state, dense rows (ld = num_lon)  */
} gd_serve_request;

typedef struct gd_serve_result {
  double arrival_s;            /* copied from the request                        */
  double start_s;              /* admitted into a slot                           */
  double finish_s;             /* result ready (time-to-result = finish - arrival) */
  double total_mass;           /* of the final state                             */
  double total_energy;
  uint64_t state_hash;         /* FNV-1a of the bytes of the final u, v, gd      */
  int32_t status;              /* GD_OK, GD_ERR_NAN, ...                         */
  int32_t slot;
} gd_serve_result;

typedef struct gd_serve_config {
  const gd_config *base;       /* options shared by every run; num_lon, num_lat,
                                  time_step_size, diffusion_coef and the filter
                                  table are set per class                        */
  double filter_lat0;          /* per-class filter table from cos(lat)/cos(lat0)
                                  (gd_cutoff_from_lat0); 0: no filter rows       */
  int64_t hbm_budget_bytes;    /* device memory for resident runs               */
  int32_t max_slots;           /* members (slots) per pool at most              */
  int32_t chunk_steps;         /* steps between admission checks                 */
} gd_serve_config;

/* Online server: runs are submitted while others are on the GPU and polled
 * for their results.  A run's class gets a pool (one handle, runs as
 * members) when it has runs waiting and none live, with as many slots (up to
 * max_slots) as the budget left by the live pools and the free device memory
 * allow; a pool idle for 50 ms is released and its bytes go back to the
 * budget.  Waiting runs are admitted into free slots between chunks of at
 * most chunk_steps steps; a run whose mass or energy becomes NaN is evicted
 * with status GD_ERR_NAN (the other runs are not affected).
 * gd_server_submit copies the request's state (arrival_s is ignored: the run
 * arrives at the call) and returns its id; gd_server_poll returns its result
 * in *res when done (*done = 1; wait != 0 blocks until then);
 * gd_server_destroy waits for every submitted run, then frees the pools. */
typedef struct gd_server gd_server;
int gd_server_create(const gd_serve_config *cfg, gd_server **out);
int gd_server_submit(gd_server *s, const gd_serve_request *req, int64_t *id);
int gd_server_poll(gd_server *s, int64_t id, int32_t wait, gd_serve_result *res, int32_t *done);
int gd_server_destroy(gd_server *s);

/* Serves the n requests (arrival times relative to the call) through a
 * server, submitting each at its arrival time, and returns when every run
 * has finished; res[k] belongs to req[k].  wall_seconds (may be NULL)
 * receives the duration of the stream. */
int gd_serve(const gd_serve_config *cfg, const gd_serve_request *req, int32_t n,
             gd_serve_result *res, double *wall_seconds);
const char *gd_serve_last_error(void);

/* ---- measurement -------------------------------------------------------- */

/* Runs nsteps steps launching every kernel eagerly with HIP events around
 * each launch (same stream) and returns, per kernel family, the number of
 * launches and the summed milliseconds.  names receives a comma-separated
 * list of the family names (capacity name_cap bytes). */
int gd_profile_kernels(gd_handle *h, int32_t nsteps, int32_t max_families,
                       int32_t *n_families, int64_t *launches, double *ms,
                       char *names, int32_t name_cap);

/* Device bytes held by the handle (all members). */
int64_t gd_device_bytes(const gd_handle *h);

/* Average duration in microseconds of one launch of a single kernel of the
 * step, replayed `reps` times back to back on the handle's stream between
 * one pair of HIP events (inputs are whatever the state slots hold; the
 * state X is not modified).  kernel: GD_TIME_TEND (tendency kernel, fast
 * pass, second predictor), GD_TIME_FILTER (the filter of that phase),
 * GD_TIME_UPDATE (beta + corrector), GD_TIME_PHASE (that phase as one fused
 * launch, in the form gd_fused_phases reports: tendency, filter rows and
 * epilogue; GD_ERR_NOT_SUPPORTED when the handle steps with separate
 * tendency and filter launches), GD_TIME_FILTER_MM
 * (the matrix-core projection of the momentum rows of every member; the
 * mixed_mfma and mixed_bf16 modes only).  No counterpart in the reference: a
 * measurement hook. */
enum { GD_TIME_TEND = 0, GD_TIME_FILTER = 1, GD_TIME_UPDATE = 2, GD_TIME_PHASE = 3, GD_TIME_FILTER_MM = 4 };

/* How the handle's step, at its current active-member count, launches each
 * tendency phase: 0 a tendency launch then a filter launch, 1 one fused
 * launch (tendency, filter rows and epilogue), 2 one fused launch in
 * workgroups of two row segments (a filtered row's segments evaluated two at
 * a time; rows of 255 to 762 longitudes).  By default phases are fused with
 * filtered rows while active members x num_lat x ceil(num_lon/254) row
 * tiles stay within 24576 (num_lon <= 1024) or 6144 (num_lon <= 2048), in
 * form 2 for rows of 255 to 508 longitudes while its workgroups (filtered
 * rows plus far rows) fit two per compute unit; gd_autotune may choose any
 * form.  GD_FUSED=0, 1 or 2 in the environment when the handle is created
 * forces a form (2 falls back to 1 outside 255 to 762 longitudes).  Every
 * form gives the same bits.  No counterpart in the reference. */
int gd_fused_phases(const gd_handle *h);

/* Tile autotuner (no counterpart in the reference).  For the current
 * active-member count, times `steps` steps (1..512) of every launch shape
 * that cannot change a result -- the form of the phase launches (separate
 * tendency and filter launches, fused, fused with a workgroup per row; see
 * gd_fused_phases), and the workgroups per member of the corrector update --
 * from the members' current state, puts that state back, and keeps the
 * fastest shape for later gd_step calls at this active-member count (the
 * results are bit for bit those of the default shapes).  cache_path (may be
 * NULL) is a text file of earlier results keyed by the device architecture,
 * grid, member counts, precision and schemes: a cached entry is applied
 * without timing, a new result is appended.  Outputs (may be NULL): the
 * chosen form (0, 1 or 2), workgroups per member and milliseconds per
 * step. */
int gd_autotune(gd_handle *h, int32_t steps, const char *cache_path, int32_t *fused,
                int32_t *upd_wpm, double *ms_per_step);
int gd_time_kernel(gd_handle *h, int32_t kernel, int32_t reps, double *us_per_launch);

#ifdef __cplusplus
}
#endif
#endif
