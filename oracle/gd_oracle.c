/*
 * gd_oracle.c -- FP64 CPU restatement of the gamil-dycore barotropic step.
 *
 * TEST INFRASTRUCTURE ONLY (see gd_oracle.h): the checker for the HIP path and
 * the "port" CPU baseline.  Nothing in the product library calls it.
 *
 * Citations are to the reference tree (dongli/gamil-dycore).  Every function
 * keeps the reference loop ranges, halo conventions and the left-to-right
 * evaluation order of the Fortran expressions; Fortran `sum` and the explicit
 * accumulation loops are restated as sequential sums in source order.
 * Pointwise loops are split over latitude rows with OpenMP; each element is
 * computed by exactly one thread with the same expression, so the results do
 * not depend on the thread count.
 *
 * Parity pins (tests/test_oracle_*.py): closed-form mesh/coef values, the
 * FFTPACK 5.1 library of the reference (compiled into oracle/_ref) for the
 * filter, the four antisymmetry identities of check_antisymmetry, exact
 * energy conservation of the beta-corrected step, telescoping mass flux,
 * the spherical-Laplacian eigenvalue of the diffusion operator and the
 * Rossby-Haurwitz / resting-state special cases.
 */
#include "gd_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* params_mod.F90:5-11 (real is real(8) via -fdefault-real-8) */
static const double PI = 3.14159265358979323846264338327950288;
#define OMEGA (2.0 * PI / 86400.0)
static const double RADIUS = 6.37122e6;

#define NSTATE 4 /* state(-1:2) */
#define NTEND 5  /* tend(-2:2)  */
#define NTF 13

typedef struct {
  double *u, *v, *gd, *iu, *iv, *igd;
} og_state;

typedef struct {
  double *f[NTF];
} og_tend;

struct og_model {
  og_params p;
  int *cutoff;
  int N, M;       /* num_lon, num_lat */
  int nx;         /* N + 4 (lon halo 2 on each side) */
  int ny;         /* M + 4 rows allocated for every 2-D field (lat -1..M+2) */
  /* mesh_mod.F90:14-40, 1-based with one guard entry each side */
  double dlon, dlat;
  double *full_lat, *half_lat, *full_cos, *half_cos, *full_sin, *half_sin;
  /* types_mod.F90:25-38 / data_mod.F90:31-47 */
  double *full_f, *full_c, *full_dlon, *full_dlat;
  double *half_f, *half_c, *half_dlon, *half_dlat;
  /* filter_mod.F90:22-31 */
  int *full_flag, *half_flag;
  unsigned char *full_mask, *half_mask; /* [row][N] 0/1, FFTPACK layout */
  double *basis;                        /* [N][N] real Fourier basis */
  og_state s[NSTATE];
  og_tend t[NTEND];
  double *ghs;
  /* diffusion_mod.F90:20-26 work arrays */
  double *du_w, *dv_w, *dgd_w, *wu, *wv, *wgd;
  double *div, *vor;
  int old_idx, new_idx;
};

#define A2(a, i, j) ((a)[(size_t)((j) + 1) * m->nx + (size_t)((i) + 1)])

static double *zalloc(size_t n) { return (double *)calloc(n, sizeof(double)); }

void og_set_threads(int n) {
#ifdef _OPENMP
  omp_set_num_threads(n > 0 ? n : 1);
#else
  (void)n;
#endif
}

int og_old(const og_model *m) { return m->old_idx; }
int og_new(const og_model *m) { return m->new_idx; }

/* parallel_mod.F90:466-524: periodic longitude halo, every allocated row */
static void fill_halo(const og_model *m, double *a) {
  const int N = m->N;
  for (int j = -1; j <= m->M + 2; ++j) {
    A2(a, -1, j) = A2(a, N - 1, j);
    A2(a, 0, j) = A2(a, N, j);
    A2(a, N + 1, j) = A2(a, 1, j);
    A2(a, N + 2, j) = A2(a, 2, j);
  }
}

/* mesh_mod.F90:44-114 */
static void mesh_init(og_model *m) {
  const int N = m->N, M = m->M;
  m->dlon = 2 * PI / N;
  m->dlat = PI / (M - 1);
  for (int j = 1; j <= M - 1; ++j) {
    m->full_lat[j] = -0.5 * PI + (j - 1) * m->dlat;
    m->half_lat[j] = m->full_lat[j] + 0.5 * m->dlat;
  }
  m->full_lat[M] = 0.5 * PI;
  for (int j = 1; j <= M - 1; ++j) {
    m->half_cos[j] = cos(m->half_lat[j]);
    m->half_sin[j] = sin(m->half_lat[j]);
  }
  for (int j = 1; j <= M; ++j) {
    m->full_cos[j] = cos(m->full_lat[j]);
    m->full_sin[j] = sin(m->full_lat[j]);
  }
  m->full_cos[1] = 0.0;
  m->full_cos[M] = 0.0;
  m->full_sin[1] = -1.0;
  m->full_sin[M] = 1.0;
  /* half_cos_lat(0) and (M) are read out of bounds by v_adv_lat
   * (dycore_mod.F90:443-444) and always multiplied by the zero lat-halo v;
   * the guard entries hold 0. */
  m->half_cos[0] = 0.0;
  m->half_cos[M] = 0.0;
}

/* data_mod.F90:31-47 */
static void coef_init(og_model *m) {
  const int M = m->M;
  for (int j = 1; j <= M; ++j) {
    m->full_f[j] = 2.0 * OMEGA * m->full_sin[j];
    if (j == 1 || j == M)
      m->full_c[j] = 0.0;
    else
      m->full_c[j] = m->full_sin[j] / m->full_cos[j] / RADIUS;
    m->full_dlon[j] = RADIUS * m->dlon * m->full_cos[j];
    m->full_dlat[j] = RADIUS * m->dlat * m->full_cos[j];
  }
  for (int j = 1; j <= M - 1; ++j) {
    m->half_f[j] = 2.0 * OMEGA * m->half_sin[j];
    m->half_c[j] = m->half_sin[j] / m->half_cos[j] / RADIUS;
    m->half_dlon[j] = RADIUS * m->dlon * m->half_cos[j];
    m->half_dlat[j] = RADIUS * m->dlat * m->half_cos[j];
  }
}

/* dycore_mod.F90:159-173 */
void og_reset_cos_lat_at_poles(og_model *m) {
  const int M = m->M;
  int j = 1;
  m->full_cos[j] = m->half_cos[1] * 0.25;
  m->full_dlon[j] = RADIUS * m->dlon * m->full_cos[j];
  m->full_dlat[j] = RADIUS * m->dlat * m->full_cos[j];
  j = M;
  m->full_cos[j] = m->half_cos[M - 1] * 0.25;
  m->full_dlon[j] = RADIUS * m->dlon * m->full_cos[j];
  m->full_dlat[j] = RADIUS * m->dlat * m->full_cos[j];
}

/* filter_mod.F90:35-103.  Flags and 0/1 masks in the FFTPACK5 coefficient
 * layout [a0, a1, b1, a2, b2, ...] (rfftf1.f:87-108).  The north half-lat
 * flags sit one row poleward of their masks (filter_mod.F90:56 vs 94-95); the
 * j=1 north half flag lands on index M, outside the (M-1)-row array, and is
 * dropped here. */
static void filter_init(og_model *m) {
  const int N = m->N, M = m->M;
  const int *c = m->cutoff;
  const int nc = m->p.n_cutoff;
  if (m->p.use_zonal_tend_filter) {
    for (int j = 1; j <= nc; ++j)
      if (c[j - 1] != 0) {
        m->full_flag[1 + j] = 1;
        m->half_flag[1 + j - 1] = 1;
      }
    for (int j = 1; j <= nc; ++j)
      if (c[j - 1] != 0) {
        m->full_flag[M - j] = 1;
        int hj = M - j + 1;
        if (hj <= M - 1) m->half_flag[hj] = 1;
      }
  }
  for (int j = 1; j <= nc; ++j)
    if (c[j - 1] != 0)
      for (int i = 1; i <= c[j - 1] + 1; ++i) {
        m->full_mask[(size_t)(1 + j) * N + 2 * i - 2] = 1;
        m->full_mask[(size_t)(1 + j) * N + 2 * i - 1] = 1;
        m->half_mask[(size_t)(1 + j - 1) * N + 2 * i - 2] = 1;
        m->half_mask[(size_t)(1 + j - 1) * N + 2 * i - 1] = 1;
      }
  for (int j = 1; j <= nc; ++j)
    if (c[j - 1] != 0)
      for (int i = 1; i <= c[j - 1] + 1; ++i) {
        m->full_mask[(size_t)(M - j) * N + 2 * i - 2] = 1;
        m->full_mask[(size_t)(M - j) * N + 2 * i - 1] = 1;
        m->half_mask[(size_t)((M - 1) - j + 1) * N + 2 * i - 2] = 1;
        m->half_mask[(size_t)((M - 1) - j + 1) * N + 2 * i - 1] = 1;
      }
  /* basis[m][l]: FFTPACK5 real basis vector for coefficient slot m (0-based):
   * slot 0 -> 1, slot 2k-1 -> cos(k t_l), slot 2k -> sin(k t_l), and for even N
   * slot N-1 -> cos(N/2 t_l); t_l = 2 pi l / N.  Angles reduced mod N exactly. */
  for (int s = 0; s < N; ++s)
    for (int l = 0; l < N; ++l) {
      double v;
      if (s == 0) {
        v = 1.0;
      } else {
        int k = (s + 1) / 2;
        long r = ((long)k * l) % N;
        double ang = 2.0 * PI * (double)r / (double)N;
        v = (s % 2 == 1) ? cos(ang) : sin(ang);
      }
      m->basis[(size_t)s * N + l] = v;
    }
}

/* Orthogonal weight of FFTPACK5 slot s: forward scaling 1/N (a0, a_{N/2}) and
 * 2/N (a_k, b_k) from rfftf1.f:87-108; the backward transform (rfftb1.f)
 * rebuilds sum(coef * basis). */
static double slot_weight(int N, int s) {
  if (s == 0) return 1.0 / N;
  if (N % 2 == 0 && s == N - 1) return 1.0 / N;
  return 2.0 / N;
}

/* filter_array_at_full_lat / _half_lat (filter_mod.F90:105-167):
 * rfft1f -> multiply by mask -> rfft1b equals the projection
 * y = sum_{kept s} w_s <basis_s, x> basis_s. */
static void filter_row(const og_model *m, const unsigned char *mask, double *x) {
  const int N = m->N;
  double *coef = (double *)malloc(sizeof(double) * N);
  double *y = (double *)calloc(N, sizeof(double));
  for (int s = 0; s < N; ++s) {
    coef[s] = 0.0;
    if (!mask[s]) continue;
    const double *b = m->basis + (size_t)s * N;
    double acc = 0.0;
    for (int l = 0; l < N; ++l) acc += x[l] * b[l];
    coef[s] = slot_weight(N, s) * acc;
  }
  for (int s = 0; s < N; ++s) {
    if (!mask[s]) continue;
    const double *b = m->basis + (size_t)s * N;
    for (int l = 0; l < N; ++l) y[l] += coef[s] * b[l];
  }
  memcpy(x, y, sizeof(double) * N);
  free(coef);
  free(y);
}

void og_filter_full(const og_model *m, int j, double *x) {
  filter_row(m, m->full_mask + (size_t)j * m->N, x);
}
void og_filter_half(const og_model *m, int j, double *x) {
  filter_row(m, m->half_mask + (size_t)j * m->N, x);
}

og_model *og_create(const og_params *p) {
  og_model *m = (og_model *)calloc(1, sizeof(og_model));
  m->p = *p;
  m->cutoff = (int *)calloc(p->n_cutoff > 0 ? p->n_cutoff : 1, sizeof(int));
  for (int k = 0; k < p->n_cutoff; ++k) m->cutoff[k] = p->cutoff[k];
  m->p.cutoff = m->cutoff;
  m->N = p->num_lon;
  m->M = p->num_lat;
  m->nx = m->N + 4;
  m->ny = m->M + 4;
  const int M = m->M, N = m->N;
  double **v1[] = {&m->full_lat, &m->half_lat, &m->full_cos, &m->half_cos,
                   &m->full_sin, &m->half_sin, &m->full_f, &m->full_c,
                   &m->full_dlon, &m->full_dlat, &m->half_f, &m->half_c,
                   &m->half_dlon, &m->half_dlat};
  for (size_t k = 0; k < sizeof(v1) / sizeof(v1[0]); ++k) *v1[k] = zalloc(M + 2);
  m->full_flag = (int *)calloc(M + 2, sizeof(int));
  m->half_flag = (int *)calloc(M + 2, sizeof(int));
  m->full_mask = (unsigned char *)calloc((size_t)(M + 2) * N, 1);
  m->half_mask = (unsigned char *)calloc((size_t)(M + 2) * N, 1);
  m->basis = zalloc((size_t)N * N);
  const size_t n2 = (size_t)m->nx * m->ny;
  for (int k = 0; k < NSTATE; ++k) {
    og_state *s = &m->s[k];
    s->u = zalloc(n2); s->v = zalloc(n2); s->gd = zalloc(n2);
    s->iu = zalloc(n2); s->iv = zalloc(n2); s->igd = zalloc(n2);
  }
  for (int k = 0; k < NTEND; ++k)
    for (int f = 0; f < NTF; ++f) m->t[k].f[f] = zalloc(n2);
  m->ghs = zalloc(n2);
  m->du_w = zalloc(n2); m->dv_w = zalloc(n2); m->dgd_w = zalloc(n2);
  m->wu = zalloc(n2); m->wv = zalloc(n2); m->wgd = zalloc(n2);
  m->div = zalloc(n2); m->vor = zalloc(n2);
  mesh_init(m);
  coef_init(m);
  filter_init(m);
  m->old_idx = 1; /* time_mod.F90:68-69 */
  m->new_idx = 2;
  return m;
}

void og_destroy(og_model *m) {
  if (!m) return;
  double *v1[] = {m->full_lat, m->half_lat, m->full_cos, m->half_cos,
                  m->full_sin, m->half_sin, m->full_f, m->full_c,
                  m->full_dlon, m->full_dlat, m->half_f, m->half_c,
                  m->half_dlon, m->half_dlat};
  for (size_t k = 0; k < sizeof(v1) / sizeof(v1[0]); ++k) free(v1[k]);
  free(m->full_flag); free(m->half_flag); free(m->full_mask); free(m->half_mask);
  free(m->basis);
  for (int k = 0; k < NSTATE; ++k) {
    og_state *s = &m->s[k];
    free(s->u); free(s->v); free(s->gd); free(s->iu); free(s->iv); free(s->igd);
  }
  for (int k = 0; k < NTEND; ++k)
    for (int f = 0; f < NTF; ++f) free(m->t[k].f[f]);
  free(m->ghs);
  free(m->du_w); free(m->dv_w); free(m->dgd_w);
  free(m->wu); free(m->wv); free(m->wgd); free(m->div); free(m->vor);
  free(m->cutoff);
  free(m);
}

static og_state *S(og_model *m, int idx) { return &m->s[idx + 1]; }
static og_tend *T(og_model *m, int idx) { return &m->t[idx + 2]; }

static double *state_field(og_model *m, int idx, int field) {
  og_state *s = S(m, idx);
  switch (field) {
    case OG_U: return s->u;
    case OG_V: return s->v;
    case OG_GD: return s->gd;
    case OG_IAP_U: return s->iu;
    case OG_IAP_V: return s->iv;
    default: return s->igd;
  }
}

static int field_rows(const og_model *m, int kind, int field) {
  if (kind == 0) return (field == OG_V || field == OG_IAP_V) ? m->M - 1 : m->M;
  return (field == OG_T_V_ADV_LON || field == OG_T_V_ADV_LAT || field == OG_T_FU ||
          field == OG_T_V_PGF || field == OG_T_DV) ? m->M - 1 : m->M;
}

void og_get_field(const og_model *m, int kind, int idx, int field, double *out) {
  og_model *mm = (og_model *)m;
  double *a = kind == 0 ? state_field(mm, idx, field) : T(mm, idx)->f[field];
  int rows = field_rows(m, kind, field);
  for (int j = 1; j <= rows; ++j)
    for (int i = 1; i <= m->N; ++i) out[(size_t)(j - 1) * m->N + i - 1] = A2(a, i, j);
}

void og_set_field(og_model *m, int kind, int idx, int field, const double *in) {
  double *a = kind == 0 ? state_field(m, idx, field) : T(m, idx)->f[field];
  int rows = field_rows(m, kind, field);
  for (int j = 1; j <= rows; ++j)
    for (int i = 1; i <= m->N; ++i) A2(a, i, j) = in[(size_t)(j - 1) * m->N + i - 1];
  fill_halo(m, a);
}

void og_set_ghs(og_model *m, const double *ghs) {
  for (int j = 1; j <= m->M; ++j)
    for (int i = 1; i <= m->N; ++i) A2(m->ghs, i, j) = ghs[(size_t)(j - 1) * m->N + i - 1];
  fill_halo(m, m->ghs);
}

void og_get_coef(const og_model *m, int which, double *out) {
  const double *src[] = {m->full_cos, m->half_cos, m->full_sin, m->half_sin,
                         m->full_f, m->full_c, m->full_dlon, m->full_dlat,
                         m->half_dlon, m->half_dlat, m->full_lat, m->half_lat};
  int half = (which == 1 || which == 3 || which == 8 || which == 9 || which == 11);
  int n = half ? m->M - 1 : m->M;
  for (int j = 1; j <= n; ++j) out[j - 1] = src[which][j];
}

static int mask_cut(const og_model *m, const unsigned char *mask) {
  int last = -1;
  for (int s = 0; s < m->N; ++s)
    if (mask[s]) last = s;
  if (last < 0) return -1;
  return (last + 1) / 2 - 1; /* indices 1..2c+2 kept */
}

void og_get_filter(const og_model *m, int *full_flag, int *half_flag,
                   int *full_cut, int *half_cut) {
  for (int j = 1; j <= m->M; ++j) {
    full_flag[j - 1] = m->full_flag[j];
    full_cut[j - 1] = mask_cut(m, m->full_mask + (size_t)j * m->N);
  }
  for (int j = 1; j <= m->M - 1; ++j) {
    half_flag[j - 1] = m->half_flag[j];
    half_cut[j - 1] = mask_cut(m, m->half_mask + (size_t)j * m->N);
  }
}

/* types_mod.F90:399-426 */
void og_iap_transform(og_model *m, int sidx) {
  og_state *s = S(m, sidx);
  const int N = m->N, M = m->M;
#pragma omp parallel for
  for (int j = 1; j <= M; ++j)
    for (int i = 1; i <= N; ++i) A2(s->igd, i, j) = sqrt(A2(s->gd, i, j));
  fill_halo(m, s->igd);
#pragma omp parallel for
  for (int j = 1; j <= M; ++j)
    for (int i = 1; i <= N; ++i)
      A2(s->iu, i, j) = 0.5 * (A2(s->igd, i, j) + A2(s->igd, i + 1, j)) * A2(s->u, i, j);
  fill_halo(m, s->iu);
#pragma omp parallel for
  for (int j = 1; j <= M - 1; ++j)
    for (int i = 1; i <= N; ++i)
      A2(s->iv, i, j) = 0.5 * (A2(s->igd, i, j) + A2(s->igd, i, j + 1)) * A2(s->v, i, j);
  fill_halo(m, s->iv);
}

/* dycore_mod.F90:367-420 */
static void zonal_momentum_advection(og_model *m, og_state *s, og_tend *t) {
  const int N = m->N, M = m->M;
  double *ual = t->f[OG_T_U_ADV_LON], *val = t->f[OG_T_V_ADV_LON];
  if (m->p.uv_adv_scheme == 0) {
#pragma omp parallel for
    for (int j = 2; j <= M - 1; ++j)
      for (int i = 1; i <= N; ++i) {
        double u1 = A2(s->u, i, j) + A2(s->u, i - 1, j);
        double u2 = A2(s->u, i, j) + A2(s->u, i + 1, j);
        A2(ual, i, j) = 0.25 / m->full_dlon[j] *
                        (u2 * A2(s->iu, i + 1, j) - u1 * A2(s->iu, i - 1, j));
      }
#pragma omp parallel for
    for (int j = 1; j <= M - 1; ++j)
      for (int i = 1; i <= N; ++i) {
        double u1 = A2(s->u, i - 1, j) + A2(s->u, i - 1, j + 1);
        double u2 = A2(s->u, i, j) + A2(s->u, i, j + 1);
        A2(val, i, j) = 0.25 / m->half_dlon[j] *
                        (u2 * A2(s->iv, i + 1, j) - u1 * A2(s->iv, i - 1, j));
      }
  } else {
    const double bl = m->p.upwind_lon_beta;
#pragma omp parallel for
    for (int j = 2; j <= M - 1; ++j)
      for (int i = 1; i <= N; ++i) {
        double u1 = A2(s->u, i, j) + A2(s->u, i - 1, j);
        double u2 = A2(s->u, i, j) + A2(s->u, i + 1, j);
        double U0 = A2(s->iu, i, j), Up = A2(s->iu, i + 1, j), Um = A2(s->iu, i - 1, j);
        A2(ual, i, j) = 0.25 / m->full_dlon[j] *
                        (u2 * (U0 + Up) - bl * fabs(u2) * (Up - U0) -
                         u1 * (U0 + Um) + bl * fabs(u1) * (U0 - Um) - (u2 - u1) * U0);
      }
#pragma omp parallel for
    for (int j = 1; j <= M - 1; ++j)
      for (int i = 1; i <= N; ++i) {
        double u1 = A2(s->u, i - 1, j) + A2(s->u, i - 1, j + 1);
        double u2 = A2(s->u, i, j) + A2(s->u, i, j + 1);
        double V0 = A2(s->iv, i, j), Vp = A2(s->iv, i + 1, j), Vm = A2(s->iv, i - 1, j);
        A2(val, i, j) = 0.25 / m->half_dlon[j] *
                        (u2 * (V0 + Vp) - bl * fabs(u2) * (Vp - V0) -
                         u1 * (V0 + Vm) + bl * fabs(u1) * (V0 - Vm) - (u2 - u1) * V0);
      }
  }
}

/* dycore_mod.F90:422-475 */
static void meridional_momentum_advection(og_model *m, og_state *s, og_tend *t) {
  const int N = m->N, M = m->M;
  const double *hc = m->half_cos;
  double *ual = t->f[OG_T_U_ADV_LAT], *val = t->f[OG_T_V_ADV_LAT];
  if (m->p.uv_adv_scheme == 0) {
#pragma omp parallel for
    for (int j = 2; j <= M - 1; ++j)
      for (int i = 1; i <= N; ++i) {
        double v1 = (A2(s->v, i, j - 1) + A2(s->v, i + 1, j - 1)) * hc[j - 1];
        double v2 = (A2(s->v, i, j) + A2(s->v, i + 1, j)) * hc[j];
        A2(ual, i, j) = 0.25 / m->full_dlat[j] *
                        (v2 * A2(s->iu, i, j + 1) - v1 * A2(s->iu, i, j - 1));
      }
#pragma omp parallel for
    for (int j = 1; j <= M - 1; ++j)
      for (int i = 1; i <= N; ++i) {
        double v1 = A2(s->v, i, j) * hc[j] + A2(s->v, i, j - 1) * hc[j - 1];
        double v2 = A2(s->v, i, j) * hc[j] + A2(s->v, i, j + 1) * hc[j + 1];
        A2(val, i, j) = 0.25 / m->half_dlat[j] *
                        (v2 * A2(s->iv, i, j + 1) - v1 * A2(s->iv, i, j - 1));
      }
  } else {
    const double bl = m->p.upwind_lat_beta;
#pragma omp parallel for
    for (int j = 2; j <= M - 1; ++j)
      for (int i = 1; i <= N; ++i) {
        double v1 = (A2(s->v, i, j - 1) + A2(s->v, i + 1, j - 1)) * hc[j - 1];
        double v2 = (A2(s->v, i, j) + A2(s->v, i + 1, j)) * hc[j];
        double U0 = A2(s->iu, i, j), Up = A2(s->iu, i, j + 1), Um = A2(s->iu, i, j - 1);
        A2(ual, i, j) = 0.25 / m->full_dlat[j] *
                        (v2 * (U0 + Up) - bl * fabs(v2) * (Up - U0) -
                         v1 * (U0 + Um) + bl * fabs(v1) * (U0 - Um) - (v2 - v1) * U0);
      }
#pragma omp parallel for
    for (int j = 1; j <= M - 1; ++j)
      for (int i = 1; i <= N; ++i) {
        double v1 = A2(s->v, i, j) * hc[j] + A2(s->v, i, j - 1) * hc[j - 1];
        double v2 = A2(s->v, i, j) * hc[j] + A2(s->v, i, j + 1) * hc[j + 1];
        double V0 = A2(s->iv, i, j), Vp = A2(s->iv, i, j + 1), Vm = A2(s->iv, i, j - 1);
        A2(val, i, j) = 0.25 / m->half_dlat[j] *
                        (v2 * (V0 + Vp) - bl * fabs(v2) * (Vp - V0) -
                         v1 * (V0 + Vm) + bl * fabs(v1) * (V0 - Vm) - (v2 - v1) * V0);
      }
  }
}

/* dycore_mod.F90:477-505 */
static void coriolis(og_model *m, og_state *s, og_tend *t) {
  const int N = m->N, M = m->M;
  double *fv = t->f[OG_T_FV], *fu = t->f[OG_T_FU];
#pragma omp parallel for
  for (int j = 2; j <= M - 1; ++j) {
    double c1 = m->half_cos[j - 1] / m->full_cos[j];
    double c2 = m->half_cos[j] / m->full_cos[j];
    for (int i = 1; i <= N; ++i)
      A2(fv, i, j) = 0.25 * (m->full_f[j] + m->full_c[j] * A2(s->u, i, j)) *
                     (c1 * (A2(s->iv, i, j - 1) + A2(s->iv, i + 1, j - 1)) +
                      c2 * (A2(s->iv, i, j) + A2(s->iv, i + 1, j)));
  }
#pragma omp parallel for
  for (int j = 1; j <= M - 1; ++j)
    for (int i = 1; i <= N; ++i)
      A2(fu, i, j) = 0.25 *
          ((m->full_f[j] + m->full_c[j] * A2(s->u, i, j)) * A2(s->iu, i, j) +
           (m->full_f[j] + m->full_c[j] * A2(s->u, i - 1, j)) * A2(s->iu, i - 1, j) +
           (m->full_f[j + 1] + m->full_c[j + 1] * A2(s->u, i, j + 1)) * A2(s->iu, i, j + 1) +
           (m->full_f[j + 1] + m->full_c[j + 1] * A2(s->u, i - 1, j + 1)) * A2(s->iu, i - 1, j + 1));
}

/* dycore_mod.F90:507-537 */
static void pressure_gradient_force(og_model *m, og_state *s, og_tend *t) {
  const int N = m->N, M = m->M;
  double *up = t->f[OG_T_U_PGF], *vp = t->f[OG_T_V_PGF];
  const double *ghs = m->ghs;
#pragma omp parallel for
  for (int j = 2; j <= M - 1; ++j)
    for (int i = 1; i <= N; ++i)
      A2(up, i, j) = 0.5 * (A2(s->igd, i, j) + A2(s->igd, i + 1, j)) / m->full_dlon[j] *
                     (A2(s->gd, i + 1, j) + A2(ghs, i + 1, j) - A2(s->gd, i, j) - A2(ghs, i, j));
#pragma omp parallel for
  for (int j = 1; j <= M - 1; ++j)
    for (int i = 1; i <= N; ++i)
      A2(vp, i, j) = 0.5 * (A2(s->igd, i, j) + A2(s->igd, i, j + 1)) / m->half_dlat[j] *
                     m->half_cos[j] *
                     (A2(s->gd, i, j + 1) + A2(ghs, i, j + 1) - A2(s->gd, i, j) - A2(ghs, i, j));
}

/* dycore_mod.F90:539-598 */
static void mass_divergence(og_model *m, og_state *s, og_tend *t) {
  const int N = m->N, M = m->M;
  double *mdl = t->f[OG_T_MASS_DIV_LON], *mdp = t->f[OG_T_MASS_DIV_LAT];
  const double *hc = m->half_cos;
#pragma omp parallel for
  for (int j = 2; j <= M - 1; ++j)
    for (int i = 1; i <= N; ++i)
      A2(mdl, i, j) = ((A2(s->igd, i, j) + A2(s->igd, i + 1, j)) * A2(s->iu, i, j) -
                       (A2(s->igd, i, j) + A2(s->igd, i - 1, j)) * A2(s->iu, i - 1, j)) *
                      0.5 / m->full_dlon[j];
#pragma omp parallel for
  for (int j = 2; j <= M - 1; ++j)
    for (int i = 1; i <= N; ++i)
      A2(mdp, i, j) = ((A2(s->igd, i, j) + A2(s->igd, i, j + 1)) * A2(s->iv, i, j) * hc[j] -
                       (A2(s->igd, i, j) + A2(s->igd, i, j - 1)) * A2(s->iv, i, j - 1) * hc[j - 1]) *
                      0.5 / m->full_dlat[j];
  {
    int j = 1;
    double sp = 0.0;
    for (int i = 1; i <= N; ++i)
      sp = sp + (A2(s->igd, i, j) + A2(s->igd, i, j + 1)) * A2(s->iv, i, j);
    sp = sp * 2.0 / N / RADIUS / m->dlat;
    for (int i = 1; i <= N; ++i) A2(mdp, i, j) = sp;
  }
  {
    int j = M;
    double np = 0.0;
    for (int i = 1; i <= N; ++i)
      np = np - (A2(s->igd, i, j) + A2(s->igd, i, j - 1)) * A2(s->iv, i, j - 1);
    np = np * 2.0 / N / RADIUS / m->dlat;
    for (int i = 1; i <= N; ++i) A2(mdp, i, j) = np;
  }
}

static void zero_field(og_model *m, double *a) {
  memset(a, 0, sizeof(double) * (size_t)m->nx * m->ny);
}

/* filter + inner-product-preserving rescale of one tendency row
 * (dycore_mod.F90:212-219 and the five identical blocks after it) */
static void smooth_row(og_model *m, double *tf, const double *w1, const double *w2,
                       int j, int half) {
  const int N = m->N;
  double s1 = 0.0;
  for (int i = 1; i <= N; ++i)
    s1 += A2(tf, i, j) * (w2 ? (A2(w1, i, j) + A2(w2, i, j)) : A2(w1, i, j));
  if (fabs(s1) > 1.0e-16) {
    double *x = (double *)malloc(sizeof(double) * N);
    for (int i = 1; i <= N; ++i) x[i - 1] = A2(tf, i, j);
    if (half)
      og_filter_half(m, j, x);
    else
      og_filter_full(m, j, x);
    for (int i = 1; i <= N; ++i) A2(tf, i, j) = x[i - 1];
    free(x);
    double s2 = 0.0;
    for (int i = 1; i <= N; ++i)
      s2 += A2(tf, i, j) * (w2 ? (A2(w1, i, j) + A2(w2, i, j)) : A2(w1, i, j));
    for (int i = 1; i <= N; ++i) A2(tf, i, j) = A2(tf, i, j) * s1 / s2;
  }
}

/* dycore_mod.F90:184-365 */
void og_space_operators(og_model *m, int sidx, int tidx, int pass) {
  og_state *s = S(m, sidx);
  og_tend *t = T(m, tidx);
  const int N = m->N, M = m->M;
  double *du = t->f[OG_T_DU], *dv = t->f[OG_T_DV], *dgd = t->f[OG_T_DGD];
  double **f = t->f;
  if (pass == OG_ALL_PASS) {
    zonal_momentum_advection(m, s, t);
    meridional_momentum_advection(m, s, t);
    coriolis(m, s, t);
    pressure_gradient_force(m, s, t);
    mass_divergence(m, s, t);
#pragma omp parallel for
    for (int j = 2; j <= M - 1; ++j) {
      for (int i = 1; i <= N; ++i)
        A2(du, i, j) = -A2(f[OG_T_U_ADV_LON], i, j) - A2(f[OG_T_U_ADV_LAT], i, j) +
                       A2(f[OG_T_FV], i, j) - A2(f[OG_T_U_PGF], i, j);
      if (m->full_flag[j]) smooth_row(m, du, s->iu, NULL, j, 0);
    }
#pragma omp parallel for
    for (int j = 1; j <= M - 1; ++j) {
      for (int i = 1; i <= N; ++i)
        A2(dv, i, j) = -A2(f[OG_T_V_ADV_LON], i, j) - A2(f[OG_T_V_ADV_LAT], i, j) -
                       A2(f[OG_T_FU], i, j) - A2(f[OG_T_V_PGF], i, j);
      if (m->half_flag[j]) smooth_row(m, dv, s->iv, NULL, j, 1);
    }
#pragma omp parallel for
    for (int j = 1; j <= M; ++j) {
      for (int i = 1; i <= N; ++i)
        A2(dgd, i, j) = -A2(f[OG_T_MASS_DIV_LON], i, j) - A2(f[OG_T_MASS_DIV_LAT], i, j);
      if (m->full_flag[j]) smooth_row(m, dgd, s->gd, m->ghs, j, 0);
    }
  } else if (pass == OG_SLOW_PASS) {
    /* #ifndef NDEBUG block (dycore_mod.F90:255-262); NDEBUG is never defined */
    zero_field(m, f[OG_T_FV]);
    zero_field(m, f[OG_T_U_PGF]);
    zero_field(m, f[OG_T_FU]);
    zero_field(m, f[OG_T_V_PGF]);
    zero_field(m, f[OG_T_MASS_DIV_LON]);
    zero_field(m, f[OG_T_MASS_DIV_LAT]);
    zonal_momentum_advection(m, s, t);
    meridional_momentum_advection(m, s, t);
#pragma omp parallel for
    for (int j = 2; j <= M - 1; ++j) {
      for (int i = 1; i <= N; ++i)
        A2(du, i, j) = -A2(f[OG_T_U_ADV_LON], i, j) - A2(f[OG_T_U_ADV_LAT], i, j);
      if (m->full_flag[j]) smooth_row(m, du, s->iu, NULL, j, 0);
    }
#pragma omp parallel for
    for (int j = 1; j <= M - 1; ++j) {
      for (int i = 1; i <= N; ++i)
        A2(dv, i, j) = -A2(f[OG_T_V_ADV_LON], i, j) - A2(f[OG_T_V_ADV_LAT], i, j);
      if (m->half_flag[j]) smooth_row(m, dv, s->iv, NULL, j, 1);
    }
    zero_field(m, dgd);
  } else {
    /* #ifndef NDEBUG block (dycore_mod.F90:299-304) */
    zero_field(m, f[OG_T_U_ADV_LON]);
    zero_field(m, f[OG_T_U_ADV_LAT]);
    zero_field(m, f[OG_T_V_ADV_LON]);
    zero_field(m, f[OG_T_V_ADV_LAT]);
    coriolis(m, s, t);
    pressure_gradient_force(m, s, t);
    mass_divergence(m, s, t);
#pragma omp parallel for
    for (int j = 1; j <= M; ++j) {
      for (int i = 1; i <= N; ++i)
        A2(dgd, i, j) = -A2(f[OG_T_MASS_DIV_LON], i, j) - A2(f[OG_T_MASS_DIV_LAT], i, j);
      if (m->full_flag[j]) smooth_row(m, dgd, s->gd, m->ghs, j, 0);
    }
#pragma omp parallel for
    for (int j = 2; j <= M - 1; ++j) {
      for (int i = 1; i <= N; ++i)
        A2(du, i, j) = A2(f[OG_T_FV], i, j) - A2(f[OG_T_U_PGF], i, j);
      if (m->full_flag[j]) smooth_row(m, du, s->iu, NULL, j, 0);
    }
#pragma omp parallel for
    for (int j = 1; j <= M - 1; ++j) {
      for (int i = 1; i <= N; ++i)
        A2(dv, i, j) = -A2(f[OG_T_FU], i, j) - A2(f[OG_T_V_PGF], i, j);
      if (m->half_flag[j]) smooth_row(m, dv, s->iv, NULL, j, 1);
    }
  }
}

/* dycore_mod.F90:600-652 */
void og_update_state(og_model *m, double dt, int tidx, int so, int sn) {
  og_state *o = S(m, so), *n = S(m, sn);
  og_tend *t = T(m, tidx);
  const int N = m->N, M = m->M;
  const double *du = t->f[OG_T_DU], *dv = t->f[OG_T_DV], *dgd = t->f[OG_T_DGD];
#pragma omp parallel for
  for (int j = 1; j <= M; ++j)
    for (int i = 1; i <= N; ++i) A2(n->gd, i, j) = A2(o->gd, i, j) + dt * A2(dgd, i, j);
  fill_halo(m, n->gd);
#pragma omp parallel for
  for (int j = -1; j <= M + 2; ++j)
    for (int i = -1; i <= N + 2; ++i) A2(n->igd, i, j) = sqrt(A2(n->gd, i, j));
#pragma omp parallel for
  for (int j = 1; j <= M; ++j)
    for (int i = 1; i <= N; ++i) A2(n->iu, i, j) = A2(o->iu, i, j) + dt * A2(du, i, j);
#pragma omp parallel for
  for (int j = 1; j <= M - 1; ++j)
    for (int i = 1; i <= N; ++i) A2(n->iv, i, j) = A2(o->iv, i, j) + dt * A2(dv, i, j);
#pragma omp parallel for
  for (int j = 2; j <= M - 1; ++j)
    for (int i = 1; i <= N; ++i)
      A2(n->u, i, j) = A2(n->iu, i, j) * 2.0 / (A2(n->igd, i, j) + A2(n->igd, i + 1, j));
#pragma omp parallel for
  for (int j = 1; j <= M - 1; ++j)
    for (int i = 1; i <= N; ++i)
      A2(n->v, i, j) = A2(n->iv, i, j) * 2.0 / (A2(n->igd, i, j) + A2(n->igd, i, j + 1));
  fill_halo(m, n->iu);
  fill_halo(m, n->iv);
  fill_halo(m, n->u);
  fill_halo(m, n->v);
}

/* types_mod.F90:347-371 */
double og_inner_product_tt(const og_model *m, int t1i, int t2i) {
  og_model *mm = (og_model *)m;
  og_tend *t1 = T(mm, t1i), *t2 = T(mm, t2i);
  const int N = m->N, M = m->M;
  double res = 0.0;
  for (int j = 2; j <= M - 1; ++j)
    for (int i = 1; i <= N; ++i)
      res = res + A2(t1->f[OG_T_DU], i, j) * A2(t2->f[OG_T_DU], i, j) * m->full_cos[j];
  for (int j = 1; j <= M - 1; ++j)
    for (int i = 1; i <= N; ++i)
      res = res + A2(t1->f[OG_T_DV], i, j) * A2(t2->f[OG_T_DV], i, j) * m->half_cos[j];
  for (int j = 1; j <= M; ++j)
    for (int i = 1; i <= N; ++i)
      res = res + A2(t1->f[OG_T_DGD], i, j) * A2(t2->f[OG_T_DGD], i, j) * m->full_cos[j];
  return res;
}

/* types_mod.F90:373-397 */
double og_inner_product_ts(const og_model *m, int ti, int si) {
  og_model *mm = (og_model *)m;
  og_tend *t = T(mm, ti);
  og_state *s = S(mm, si);
  const int N = m->N, M = m->M;
  double res = 0.0;
  for (int j = 2; j <= M - 1; ++j)
    for (int i = 1; i <= N; ++i)
      res = res + A2(t->f[OG_T_DU], i, j) * A2(s->iu, i, j) * m->full_cos[j];
  for (int j = 1; j <= M - 1; ++j)
    for (int i = 1; i <= N; ++i)
      res = res + A2(t->f[OG_T_DV], i, j) * A2(s->iv, i, j) * m->half_cos[j];
  for (int j = 1; j <= M; ++j)
    for (int i = 1; i <= N; ++i)
      res = res + A2(t->f[OG_T_DGD], i, j) * A2(s->gd, i, j) * m->full_cos[j];
  return res;
}

/* dycore_mod.F90:754-792 */
double og_predict_correct(og_model *m, double time_step_size, int old, int new_,
                          int pass, int flag) {
  double dt = time_step_size * 0.5;
  og_space_operators(m, old, old, pass);
  og_update_state(m, dt, old, old, new_);
  og_space_operators(m, new_, old, pass);
  og_update_state(m, dt, old, old, new_);
  og_space_operators(m, new_, new_, pass);
  double ip1 = og_inner_product_tt(m, old, new_);
  double ip2 = og_inner_product_tt(m, new_, new_);
  double beta = (flag && ip1 != 0.0 && ip2 != 0.0) ? ip1 / ip2 : 1.0;
  dt = time_step_size * beta;
  og_update_state(m, dt, new_, old, new_);
  return beta;
}

/* types_mod.F90:229-345 tend algebra over whole halo'd arrays */
static void copy_state(og_model *m, int a, int b) {
  const size_t n2 = (size_t)m->nx * m->ny * sizeof(double);
  og_state *x = S(m, a), *y = S(m, b);
  memcpy(y->u, x->u, n2); memcpy(y->v, x->v, n2); memcpy(y->gd, x->gd, n2);
  memcpy(y->iu, x->iu, n2); memcpy(y->iv, x->iv, n2); memcpy(y->igd, x->igd, n2);
}
static void tend_axpy(og_model *m, int a, int b, double sign) { /* b = b + sign*a */
  const size_t n2 = (size_t)m->nx * m->ny;
  for (int f = 0; f < NTF; ++f) {
    double *x = T(m, a)->f[f], *y = T(m, b)->f[f];
    if (sign > 0)
      for (size_t k = 0; k < n2; ++k) y[k] = y[k] + x[k];
    else
      for (size_t k = 0; k < n2; ++k) y[k] = y[k] - x[k];
  }
}
static void tend_scale(og_model *m, double sc, int a) {
  const size_t n2 = (size_t)m->nx * m->ny;
  for (int f = 0; f < NTF; ++f) {
    double *y = T(m, a)->f[f];
    for (size_t k = 0; k < n2; ++k) y[k] = y[k] * sc;
  }
}
static void tend_zero(og_model *m, int a) {
  for (int f = 0; f < NTF; ++f) zero_field(m, T(m, a)->f[f]);
}

/* dycore_mod.F90:671-687 */
static double csp2_splitting(og_model *m) {
  const double dt = m->p.dt;
  double fast_dt = dt / m->p.subcycles;
  int t1 = 0, t2 = m->old_idx;
  og_predict_correct(m, 0.5 * dt, m->old_idx, t1, OG_SLOW_PASS, m->p.qcon_modified);
  for (int k = 1; k <= m->p.subcycles; ++k) {
    og_predict_correct(m, fast_dt, t1, t2, OG_FAST_PASS, m->p.qcon_modified);
    int tmp = t1; t1 = t2; t2 = tmp;
  }
  return og_predict_correct(m, 0.5 * dt, t1, m->new_idx, OG_SLOW_PASS, m->p.qcon_modified);
}

/* dycore_mod.F90:689-752 */
static double isp_splitting(og_model *m) {
  const double tss = m->p.dt;
  double fast_dt = tss / m->p.subcycles;
  double half_dt = tss * 0.5;
  int t1 = m->old_idx, t2 = 0;
  const int saved = -1, slow = -2, accum = -1;
  const int old = m->old_idx, new_ = m->new_idx;
  copy_state(m, old, saved);
  tend_zero(m, accum);
  og_space_operators(m, old, slow, OG_SLOW_PASS);
  for (int k = 1; k <= m->p.subcycles; ++k) {
    og_space_operators(m, t1, t1, OG_FAST_PASS);
    tend_axpy(m, slow, t1, 1.0);
    og_update_state(m, fast_dt * 0.5, t1, t1, t2);
    og_space_operators(m, t2, t1, OG_FAST_PASS);
    tend_axpy(m, slow, t1, 1.0);
    og_update_state(m, fast_dt * 0.5, t1, t1, t2);
    og_space_operators(m, t2, t2, OG_FAST_PASS);
    tend_axpy(m, t2, accum, 1.0);
    tend_axpy(m, slow, t2, 1.0);
    og_update_state(m, fast_dt, t2, t1, t2);
    int tmp = t1; t1 = t2; t2 = tmp;
  }
  tend_scale(m, 2.0 / m->p.subcycles, accum);
  og_space_operators(m, t1, t1, OG_SLOW_PASS);
  tend_axpy(m, slow, t1, -1.0);
  og_update_state(m, half_dt, t1, t1, new_);
  og_space_operators(m, new_, t1, OG_SLOW_PASS);
  tend_axpy(m, slow, t1, -1.0);
  og_update_state(m, half_dt, t1, t1, new_);
  og_space_operators(m, new_, new_, OG_SLOW_PASS);
  tend_axpy(m, slow, new_, 1.0);
  tend_axpy(m, accum, new_, 1.0);
  double ip1 = og_inner_product_ts(m, new_, saved);
  double ip2 = og_inner_product_tt(m, new_, new_);
  double beta = (m->p.qcon_modified && ip1 != 0.0 && ip2 != 0.0) ? ip1 / ip2 : 1.0;
  beta = beta * 4.0 / tss;
  half_dt = half_dt * beta;
  og_update_state(m, half_dt, new_, saved, new_);
  return beta;
}

/* diffusion_mod.F90:74-217 */
void og_ordinary_diffusion(og_model *m, double dt, int sidx) {
  og_state *s = S(m, sidx);
  const int N = m->N, M = m->M;
  const size_t n2 = (size_t)m->nx * m->ny * sizeof(double);
  double *u = m->wu, *v = m->wv, *gd = m->wgd;
  double *ud = m->du_w, *vd = m->dv_w, *gdd = m->dgd_w;
  const double *hc = m->half_cos, *fc = m->full_cos;
  memcpy(u, s->u, n2);
  memcpy(v, s->v, n2);
  memcpy(gd, s->gd, n2);
  const int norder = m->p.diffusion_order / 2;
  for (int order = 1; order <= norder; ++order) {
#pragma omp parallel for
    for (int j = 2; j <= M - 1; ++j)
      for (int i = 1; i <= N; ++i)
        A2(gdd, i, j) = (A2(gd, i + 1, j) - 2 * A2(gd, i, j) + A2(gd, i - 1, j)) /
                            (m->full_dlon[j] * m->full_dlon[j]) +
                        ((A2(gd, i, j + 1) - A2(gd, i, j)) * hc[j] -
                         (A2(gd, i, j) - A2(gd, i, j - 1)) * hc[j - 1]) /
                            (m->full_dlat[j] * m->full_dlat[j]) * fc[j];
    {
      int j = 1;
      double sp = 0.0;
      for (int i = 1; i <= N; ++i) sp = sp + A2(gd, i, j + 1) - A2(gd, i, j);
      double val = sp * hc[j] / (m->full_dlat[j] * m->full_dlat[j]) * fc[j] / N;
      for (int i = -1; i <= N + 2; ++i) A2(gdd, i, j) = val;
    }
    {
      int j = M;
      double np = 0.0;
      for (int i = 1; i <= N; ++i) np = np - (A2(gd, i, j) - A2(gd, i, j - 1));
      double val = np * hc[j - 1] / (m->full_dlat[j] * m->full_dlat[j]) * fc[j] / N;
      for (int i = -1; i <= N + 2; ++i) A2(gdd, i, j) = val;
    }
#pragma omp parallel for
    for (int j = 2; j <= M - 1; ++j)
      for (int i = 1; i <= N; ++i)
        A2(ud, i, j) = (A2(u, i + 1, j) - 2 * A2(u, i, j) + A2(u, i - 1, j)) /
                           (m->full_dlon[j] * m->full_dlon[j]) +
                       ((A2(u, i, j + 1) - A2(u, i, j)) * hc[j] -
                        (A2(u, i, j) - A2(u, i, j - 1)) * hc[j - 1]) /
                           (m->full_dlat[j] * m->full_dlat[j]) * fc[j];
#pragma omp parallel for
    for (int j = 1; j <= M - 1; ++j)
      for (int i = 1; i <= N; ++i)
        A2(vd, i, j) = (A2(v, i + 1, j) - 2 * A2(v, i, j) + A2(v, i - 1, j)) /
                       (m->half_dlon[j] * m->half_dlon[j]);
#pragma omp parallel for
    for (int j = 2; j <= M - 2; ++j)
      for (int i = 1; i <= N; ++i)
        A2(vd, i, j) = A2(vd, i, j) + ((A2(v, i, j + 1) - A2(v, i, j)) * fc[j + 1] -
                                       (A2(v, i, j) - A2(v, i, j - 1)) * fc[j]) /
                                          (m->half_dlat[j] * m->half_dlat[j]) * hc[j];
    {
      int j = 1;
      for (int i = 1; i <= N; ++i)
        A2(vd, i, j) = A2(vd, i, j) + (A2(v, i, j + 1) - A2(v, i, j)) * fc[j + 1] /
                                          (m->half_dlat[j] * m->half_dlat[j]) * hc[j];
    }
    {
      int j = M - 1;
      for (int i = 1; i <= N; ++i)
        A2(vd, i, j) = A2(vd, i, j) - (A2(v, i, j) - A2(v, i, j - 1)) * fc[j] /
                                          (m->half_dlat[j] * m->half_dlat[j]) * hc[j];
    }
    if (order != norder) {
      fill_halo(m, gdd);
      fill_halo(m, ud);
      fill_halo(m, vd);
      memcpy(gd, gdd, n2);
      memcpy(u, ud, n2);
      memcpy(v, vd, n2);
    }
  }
  /* FFT filter of the diffusion tendencies, no rescale (diffusion_mod.F90:182-193) */
#pragma omp parallel for
  for (int j = 2; j <= M - 1; ++j)
    if (m->full_flag[j]) {
      double *x = (double *)malloc(sizeof(double) * N);
      for (int i = 1; i <= N; ++i) x[i - 1] = A2(gdd, i, j);
      og_filter_full(m, j, x);
      for (int i = 1; i <= N; ++i) A2(gdd, i, j) = x[i - 1];
      for (int i = 1; i <= N; ++i) x[i - 1] = A2(ud, i, j);
      og_filter_full(m, j, x);
      for (int i = 1; i <= N; ++i) A2(ud, i, j) = x[i - 1];
      free(x);
    }
#pragma omp parallel for
  for (int j = 1; j <= M - 1; ++j)
    if (m->half_flag[j]) {
      double *x = (double *)malloc(sizeof(double) * N);
      for (int i = 1; i <= N; ++i) x[i - 1] = A2(vd, i, j);
      og_filter_half(m, j, x);
      for (int i = 1; i <= N; ++i) A2(vd, i, j) = x[i - 1];
      free(x);
    }
  int sign = ((norder + 1) % 2 == 0) ? 1 : -1; /* (-1)**(order/2+1) */
  const double nu = m->p.diffusion_coef;
#pragma omp parallel for
  for (int j = 1; j <= M; ++j) {
    for (int i = 1; i <= N; ++i)
      A2(s->gd, i, j) = A2(s->gd, i, j) + sign * dt * nu * A2(gdd, i, j);
    for (int i = 1; i <= N; ++i)
      A2(s->u, i, j) = A2(s->u, i, j) + sign * dt * nu * A2(ud, i, j);
  }
#pragma omp parallel for
  for (int j = 1; j <= M - 1; ++j)
    for (int i = 1; i <= N; ++i)
      A2(s->v, i, j) = A2(s->v, i, j) + sign * dt * nu * A2(vd, i, j);
  fill_halo(m, s->gd);
  fill_halo(m, s->u);
  fill_halo(m, s->v);
  og_iap_transform(m, sidx);
}

/* dycore_mod.F90:654-669 */
void og_time_integrate(og_model *m, double *beta_last) {
  double beta;
  switch (m->p.split_scheme) {
    case 2: beta = csp2_splitting(m); break;
    case 3: beta = isp_splitting(m); break;
    default:
      beta = og_predict_correct(m, m->p.dt, m->old_idx, m->new_idx, OG_ALL_PASS,
                                m->p.qcon_modified);
  }
  if (beta_last) *beta_last = beta;
  if (m->p.use_diffusion) og_ordinary_diffusion(m, m->p.dt, m->new_idx);
}

/* diag_mod.F90:42-121 */
void og_diag(og_model *m, int sidx, double *mass, double *energy, double *div_out,
             double *vor_out) {
  og_state *s = S(m, sidx);
  const int N = m->N, M = m->M;
  if (div_out || vor_out) {
    for (int j = 2; j <= M - 1; ++j)
      for (int i = 1; i <= N; ++i) {
        double um1 = A2(s->u, i - 1, j), up1 = A2(s->u, i, j);
        double vm1 = A2(s->v, i, j - 1) * m->half_cos[j - 1];
        double vp1 = A2(s->v, i, j) * m->half_cos[j];
        A2(m->div, i, j) = (up1 - um1) / m->full_dlon[j] + (vp1 - vm1) / m->full_dlat[j];
      }
    fill_halo(m, m->div);
    for (int j = 1; j <= M - 1; ++j)
      for (int i = 1; i <= N; ++i) {
        double um1 = A2(s->u, i, j), up1 = A2(s->u, i, j + 1);
        double vm1 = A2(s->v, i, j), vp1 = A2(s->v, i + 1, j);
        A2(m->vor, i, j) = (vp1 - vm1) / m->half_dlon[j] - (up1 - um1) / m->half_dlat[j];
      }
    fill_halo(m, m->vor);
    if (div_out)
      for (int j = 1; j <= M; ++j)
        for (int i = 1; i <= N; ++i) div_out[(size_t)(j - 1) * N + i - 1] = A2(m->div, i, j);
    if (vor_out)
      for (int j = 1; j <= M - 1; ++j)
        for (int i = 1; i <= N; ++i) vor_out[(size_t)(j - 1) * N + i - 1] = A2(m->vor, i, j);
  }
  double tm = 0.0;
  for (int j = 1; j <= M; ++j)
    for (int i = 1; i <= N; ++i)
      tm = tm + m->full_cos[j] * m->dlon * m->dlat * A2(s->gd, i, j);
  tm = tm * (RADIUS * RADIUS);
  double te = 0.0;
  for (int j = 2; j <= M - 1; ++j)
    for (int i = 1; i <= N; ++i)
      te = te + A2(s->iu, i, j) * A2(s->iu, i, j) * m->full_cos[j];
  for (int j = 1; j <= M - 1; ++j)
    for (int i = 1; i <= N; ++i)
      te = te + A2(s->iv, i, j) * A2(s->iv, i, j) * m->half_cos[j];
  for (int j = 1; j <= M; ++j)
    for (int i = 1; i <= N; ++i) {
      double h = A2(s->gd, i, j) + A2(m->ghs, i, j);
      te = te + h * h * m->full_cos[j];
    }
  if (mass) *mass = tm;
  if (energy) *energy = te;
}

/* dycore_mod.F90:794-851 (debug invariants; not called by the reference) */
void og_check_antisymmetry(const og_model *m, int ti, int si, double *out) {
  og_model *mm = (og_model *)m;
  og_tend *t = T(mm, ti);
  og_state *s = S(mm, si);
  const int N = m->N, M = m->M;
  double a[10] = {0};
  for (int j = 1; j <= M; ++j)
    for (int i = 1; i <= N; ++i) {
      double w = A2(s->iu, i, j) * m->full_cos[j];
      a[0] += A2(t->f[OG_T_U_ADV_LON], i, j) * w;
      a[1] += A2(t->f[OG_T_U_ADV_LAT], i, j) * w;
      a[2] += A2(t->f[OG_T_FV], i, j) * w;
      a[3] += A2(t->f[OG_T_U_PGF], i, j) * w;
    }
  for (int j = 1; j <= M - 1; ++j)
    for (int i = 1; i <= N; ++i) {
      double w = A2(s->iv, i, j) * m->half_cos[j];
      a[4] += A2(t->f[OG_T_V_ADV_LON], i, j) * w;
      a[5] += A2(t->f[OG_T_V_ADV_LAT], i, j) * w;
      a[6] += A2(t->f[OG_T_FU], i, j) * w;
      a[7] += A2(t->f[OG_T_V_PGF], i, j) * w;
    }
  for (int j = 1; j <= M; ++j)
    for (int i = 1; i <= N; ++i) {
      double w = (A2(s->gd, i, j) + A2(m->ghs, i, j)) * m->full_cos[j];
      a[8] += A2(t->f[OG_T_MASS_DIV_LON], i, j) * w;
      a[9] += A2(t->f[OG_T_MASS_DIV_LAT], i, j) * w;
    }
  out[0] = a[0] + a[4] + a[1] + a[5];
  out[1] = a[3] + a[8];
  out[2] = a[2] - a[6];
  out[3] = a[7] + a[9];
}

/* dycore_test.F90:29-42 (IC written into state(1) + halos) */
void og_set_initial_state(og_model *m, const double *u, const double *v,
                          const double *gd, const double *ghs) {
  int o = m->old_idx;
  og_set_field(m, 0, o, OG_U, u);
  og_set_field(m, 0, o, OG_V, v);
  og_set_field(m, 0, o, OG_GD, gd);
  og_set_ghs(m, ghs);
}

/* dycore_mod.F90:119-129 (prologue of dycore_run) */
void og_start(og_model *m, double *mass, double *energy) {
  og_reset_cos_lat_at_poles(m);
  og_iap_transform(m, m->old_idx);
  og_diag(m, m->old_idx, mass, energy, NULL, NULL);
}

/* dycore_mod.F90:131-140 + time_mod.F90:106-130 */
int og_run(og_model *m, int nsteps, double *mass, double *energy, double *beta) {
  int nan_step = 0;
  for (int k = 0; k < nsteps; ++k) {
    double b = 1.0, tm, te;
    og_time_integrate(m, &b);
    int tmp = m->old_idx; m->old_idx = m->new_idx; m->new_idx = tmp;
    og_diag(m, m->old_idx, &tm, &te, NULL, NULL);
    if (mass) mass[k] = tm;
    if (energy) energy[k] = te;
    if (beta) beta[k] = b;
    if ((isnan(tm) || isnan(te)) && !nan_step) {
      nan_step = k + 1;
      break; /* log_error -> stop 1 (diag_mod.F90:79-87) */
    }
  }
  return nan_step;
}
