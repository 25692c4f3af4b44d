/*
 * gd_oracle.h -- FP64 CPU restatement of the gamil-dycore barotropic time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This oracle is the checker for the HIP product
 * path: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it.  The product library never links or calls it.
 *
 * It follows the reference Fortran (dongli/gamil-dycore, built there with
 * -fdefault-real-8) loop by loop and keeps the reference's halo'd array
 * layout (lon halo 2, lat halo 2), its state(-1:2)/tend(-2:2) buffer roles and
 * the left-to-right operation order of every source expression.  Reductions
 * (Fortran sum/loop accumulations) run sequentially in the source loop order.
 *
 * Dense exchange layout: row-major, longitude fastest, rows j = 1..M for
 * full-lat fields (u, gd, ghs) and j = 1..M-1 for half-lat fields (v).
 */
#ifndef GD_ORACLE_H
#define GD_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct og_params {
  int num_lon;                /* N (params_mod.F90:13)                       */
  int num_lat;                /* M, full latitudes incl. both poles          */
  double dt;                  /* time_step_size [s]                          */
  int split_scheme;           /* 0 none, 1 csp1 (runs unsplit), 2 csp2, 3 isp */
  int subcycles;
  int qcon_modified;
  int uv_adv_scheme;          /* 0 center_diff, 1 upwind                     */
  double upwind_lon_beta;
  double upwind_lat_beta;
  int use_diffusion;
  int diffusion_order;
  double diffusion_coef;
  int use_zonal_tend_filter;
  int n_cutoff;               /* entries in cutoff[]                         */
  const int *cutoff;          /* zonal_tend_filter_cutoff_wavenumber(1:n)    */
} og_params;

typedef struct og_model og_model;

/* field ids for og_get_field / og_set_field */
enum { OG_U = 0, OG_V, OG_GD, OG_IAP_U, OG_IAP_V, OG_IAP_GD };
enum { OG_T_U_ADV_LON = 0, OG_T_U_ADV_LAT, OG_T_V_ADV_LON, OG_T_V_ADV_LAT,
       OG_T_FU, OG_T_FV, OG_T_U_PGF, OG_T_V_PGF, OG_T_MASS_DIV_LON,
       OG_T_MASS_DIV_LAT, OG_T_DU, OG_T_DV, OG_T_DGD };
enum { OG_ALL_PASS = 0, OG_FAST_PASS = 1, OG_SLOW_PASS = 2 };

og_model *og_create(const og_params *p);
void og_destroy(og_model *m);
void og_set_threads(int n);
int og_old(const og_model *m);
int og_new(const og_model *m);

/* mesh / coefficient access: which = 0 full_cos,1 half_cos,2 full_sin,
 * 3 half_sin,4 full_f,5 full_c,6 full_dlon,7 full_dlat,8 half_dlon,
 * 9 half_dlat,10 full_lat,11 half_lat.  out has M (full) or M-1 (half). */
void og_get_coef(const og_model *m, int which, double *out);
/* filter rows: flags/cutoff per full row (M) and half row (M-1); cutoff is
 * the largest kept wavenumber c (mask keeps indices 1..2c+2), -1 if no mask */
void og_get_filter(const og_model *m, int *full_flag, int *half_flag,
                   int *full_cut, int *half_cut);

/* state (kind 0, idx -1..2) or tend (kind 1, idx -2..2) field access */
void og_get_field(const og_model *m, int kind, int idx, int field, double *out);
void og_set_field(og_model *m, int kind, int idx, int field, const double *in);
void og_set_ghs(og_model *m, const double *ghs);

/* dycore_test IC write into state(old) + dycore_run prologue */
void og_set_initial_state(og_model *m, const double *u, const double *v,
                          const double *gd, const double *ghs);
void og_start(og_model *m, double *mass, double *energy);

/* primitives (state / tend buffer indices as in the reference) */
void og_reset_cos_lat_at_poles(og_model *m);
void og_iap_transform(og_model *m, int s);
void og_space_operators(og_model *m, int s, int t, int pass);
void og_update_state(og_model *m, double dt, int t, int s_old, int s_new);
double og_inner_product_tt(const og_model *m, int t1, int t2);
double og_inner_product_ts(const og_model *m, int t, int s);
double og_predict_correct(og_model *m, double dt, int s_old, int s_new,
                          int pass, int qcon);
void og_time_integrate(og_model *m, double *beta_last);
void og_ordinary_diffusion(og_model *m, double dt, int s);
void og_diag(og_model *m, int s, double *mass, double *energy,
             double *div, double *vor);
void og_filter_full(const og_model *m, int j, double *x);   /* x: N values */
void og_filter_half(const og_model *m, int j, double *x);
void og_check_antisymmetry(const og_model *m, int t, int s, double *out4);

/* nsteps full steps (time_integrate, time_advance, diag_run).  Returns the
 * 1-based step at which mass or energy first became NaN, else 0. */
int og_run(og_model *m, int nsteps, double *mass, double *energy, double *beta);

#ifdef __cplusplus
}
#endif
#endif
