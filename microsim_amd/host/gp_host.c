/*
 * gp_host.c -- host side of the MI355X drop-in for the serial microsim_gp driver (see gp_host.h).
 *
 * Input.in is read the way the reference reads it: fgets() lines, sscanf("%100s = %100[^;];") into a key and
 * a value buffer that keep their previous contents when a line does not match (blank lines therefore repeat
 * the previous key), tuples split on "{,}" with strtok semantics (empty tokens skipped, leading blanks kept).
 * GPs = Grand_potential_Finite_difference_2D_serial in the citations below.
 */
#define _GNU_SOURCE
#include "gp_host.h"

#include <ctype.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#define P GP_MAXP
#define NCM GP_MAXNC

enum {
  SEEN_DIM = 1u << 0, SEEN_MX = 1u << 1, SEEN_MY = 1u << 2, SEEN_DX = 1u << 3, SEEN_DY = 1u << 4,
  SEEN_DT = 1u << 5, SEEN_N = 1u << 6, SEEN_K = 1u << 7, SEEN_EPS = 1u << 8, SEEN_TAU = 1u << 9,
  SEEN_GAMMA = 1u << 10, SEEN_TAUAB = 1u << 11, SEEN_DAB = 1u << 12, SEEN_GABC = 1u << 13, SEEN_T = 1u << 14,
  SEEN_TEQ = 1u << 15, SEEN_ISO = 1u << 16, SEEN_TG = 1u << 17
};

static void seterr(char *err, size_t n, const char *fmt, ...) {
  if (!err || !n) return;
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(err, n, fmt, ap);
  va_end(ap);
}

/* strtok_r(s, "{,}") over a copy of the value: returns the token count, tokens in tok[] (each <= 99 chars). */
static int split_tuple(const char *val, char tok[][100], int maxtok) {
  char buf[256];
  strncpy(buf, val, sizeof(buf) - 1);
  buf[sizeof(buf) - 1] = 0;
  int n = 0;
  char *save = NULL;
  for (char *s = strtok_r(buf, "{,}", &save); s && n < maxtok; s = strtok_r(NULL, "{,}", &save)) {
    strncpy(tok[n], s, 99);
    tok[n][99] = 0;
    n++;
  }
  return n;
}

/* ------------------------------------------------------------------------------------------------------------
 * Small dense algebra with the reference's operation order (GPs/functions/matrix.h:16-218): LU with its
 * pivot rule (the running maximum keeps the signed entry, matrix.h:165-174), forward/back substitution with
 * pow(x,-1), column/row un-permutation, and the in-place restoration of the input as P^T (L U).
 * ---------------------------------------------------------------------------------------------------------- */
static void mm_mul(int n, double m1[][4], double m2[][4], double prod[][4]) {
  for (int k = 0; k < n; k++)
    for (int i = 0; i < n; i++) {
      double s = 0;
      for (int j = 0; j < n; j++) s = s + m1[k][j] * m2[j][i];
      prod[k][i] = s;
    }
}

static void lu_inverse(int n, double a[][4], double inv[][4]) {
  double factor[4][4] = {{0}}, inv1[4][4], prod[4][4];
  int tag[4];
  for (int k = 0; k < n; k++) tag[k] = k;
  for (int k = 0; k < n; k++) {
    /* pivot */
    double big = fabs(a[k][k]);
    int t1 = k;
    for (int i = k + 1; i < n; i++)
      if (fabs(a[i][k]) > big) {
        t1 = i;
        big = a[i][k];
      }
    int ts = tag[k];
    tag[k] = tag[t1];
    tag[t1] = ts;
    for (int i = 0; i < n; i++) {
      double s = a[k][i];
      a[k][i] = a[t1][i];
      a[t1][i] = s;
    }
    for (int i = 0; i < k; i++) {
      double s = factor[k][i];
      factor[k][i] = factor[t1][i];
      factor[t1][i] = s;
    }
    for (int i = k + 1; i < n; i++) {
      double fact = -a[i][k] / a[k][k];
      factor[i][k] = -fact;
      for (int j = k; j <= n - 1; j++) a[i][j] = fact * a[k][j] + a[i][j];
    }
  }
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) {
      if (i == j) factor[i][j] = 1;
      if (j > i) factor[i][j] = 0;
    }
  for (int col = 0; col < n; col++) {
    double v1[4], v[4];
    /* forward substitution with the unit lower factor on e_col */
    v1[0] = (col == 0) ? 1.0 : 0.0;
    for (int i = 1; i < n; i++) {
      double s = 0;
      for (int j = 0; j < i; j++) s = s - factor[i][j] * v1[j];
      v1[i] = ((i == col) ? 1.0 : 0.0) + s;
    }
    /* back substitution */
    v[n - 1] = v1[n - 1] * pow(a[n - 1][n - 1], -1);
    for (int i = n - 2; i >= 0; i--) {
      double s = 0;
      for (int j = i + 1; j < n; j++) s = s - a[i][j] * v[j];
      v[i] = (v1[i] + s) * pow(a[i][i], -1);
    }
    for (int j = 0; j < n; j++) inv1[j][col] = v[j];
  }
  for (int j = 0; j < n; j++)
    for (int p = 0; p < n; p++) inv[p][tag[j]] = inv1[p][j];
  mm_mul(n, factor, a, prod);
  for (int j = 0; j < n; j++)
    for (int p = 0; p < n; p++) a[tag[j]][p] = prod[j][p];
}

/* ------------------------------------------------------------------------------------------------------------
 * Parameter tables
 * ---------------------------------------------------------------------------------------------------------- */
static int fill_pairs(double *mat, const char *val, int N, char *err, size_t errlen, const char *key) {
  char tok[64][100];
  int nt = split_tuple(val, tok, 64);
  int need = N * (N - 1) / 2;
  if (nt < need) {
    seterr(err, errlen, "%s needs %d values for %d phases, got %d", key, need, N, nt);
    return GP_E_ARG;
  }
  int k = 0;
  for (int i = 0; i < N; i++)
    for (int j = i + 1; j < N; j++) {
      mat[i * P + i] = 0.0;
      mat[i * P + j] = atof(tok[k]);
      mat[j * P + i] = mat[i * P + j];
      k++;
    }
  return GP_OK;
}

static int fill_triples(double *mat, const char *val, int N, char *err, size_t errlen) {
  char tok[64][100];
  int nt = split_tuple(val, tok, 64);
  int need = N * (N - 1) * (N - 2) / 6;
  if (nt < need) {
    seterr(err, errlen, "Gamma_abc needs %d values for %d phases, got %d", need, N, nt);
    return GP_E_ARG;
  }
  int l = 0;
  for (int i = 0; i < N; i++)
    for (int j = i + 1; j < N; j++)
      for (int k = j + 1; k < N; k++) {
        double v = atof(tok[l++]);
        int ids[6][3] = {{i, j, k}, {i, k, j}, {j, i, k}, {j, k, i}, {k, i, j}, {k, j, i}};
        for (int s = 0; s < 6; s++) mat[(ids[s][0] * P + ids[s][1]) * P + ids[s][2]] = v;
      }
  return GP_OK;
}

static void rot_axis(double R[3][3], double theta, int axis) {
  double c = cos(theta * M_PI / 180.0), s = sin(theta * M_PI / 180.0);
  if (axis == 0) {
    R[0][0] = 1.0; R[0][1] = 0.0; R[0][2] = 0.0; R[1][0] = 0.0; R[2][0] = 0.0;
    R[1][1] = c; R[1][2] = -s; R[2][1] = s; R[2][2] = c;
  } else if (axis == 1) {
    R[0][0] = c; R[0][1] = 0.0; R[0][2] = s; R[1][0] = 0.0; R[2][0] = -s;
    R[1][1] = 1.0; R[1][2] = 0.0; R[2][1] = 0.0; R[2][2] = c;
  } else {
    R[0][0] = c; R[0][1] = -s; R[0][2] = 0.0; R[1][0] = s; R[2][0] = 0.0;
    R[1][1] = c; R[1][2] = 0.0; R[2][1] = 0.0; R[2][2] = 1.0;
  }
}

/* populate_rotation_matrix (GPs/functions/utility_functions.h:700-774): R = Rx Ry Rz, inverted by the LU
 * routine, which also rewrites R in place; both copied to the (b,a) slot. */
static int set_rotation(gp_input *in, const char *val, int quiet, char *err, size_t errlen) {
  char tok[16][100];
  int nt = split_tuple(val, tok, 16);
  if (nt < 5) {
    seterr(err, errlen, "Rotation_matrix needs {a, b, theta_x, theta_y, theta_z}");
    return GP_E_ARG;
  }
  long a = atol(tok[0]), b = atol(tok[1]);
  if (a < 0 || b < 0 || a >= in->p.nphases || b >= in->p.nphases) {
    seterr(err, errlen, "Rotation_matrix phase index out of range");
    return GP_E_ARG;
  }
  double Rx[3][3], Ry[3][3], Rz[3][3];
  rot_axis(Rx, atof(tok[2]), 0);
  rot_axis(Ry, atof(tok[3]), 1);
  rot_axis(Rz, atof(tok[4]), 2);
  double m1[4][4] = {{0}}, m2[4][4] = {{0}}, mult[4][4], R[4][4], Ri[4][4];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      m1[i][j] = Rx[i][j];
      m2[i][j] = Ry[i][j];
    }
  mm_mul(3, m1, m2, mult);
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) m2[i][j] = Rz[i][j];
  mm_mul(3, mult, m2, R);
  lu_inverse(3, R, Ri);
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      in->rot3[a][b][i][j] = R[i][j];
      in->rot3_inv[a][b][i][j] = Ri[i][j];
      in->rot3[b][a][i][j] = R[i][j];
      in->rot3_inv[b][a][i][j] = Ri[i][j];
    }
  in->seen_rot[a][b] = in->seen_rot[b][a] = 1;
  if (!quiet) printf("phase1=%ld, phase2=%ld\n", a, b);
  return GP_OK;
}

/* ------------------------------------------------------------------------------------------------------------
 * Boundary tables (GPs/functions/utility_functions.h:376-654)
 * ---------------------------------------------------------------------------------------------------------- */
static int64_t rowsx(const gp_input *in) { return in->p.mesh_x + 6; }
static int64_t rowsy(const gp_input *in) { return in->p.mesh_y + 6; }

static void bc_types(gp_input *in, int f, const int t[6]) {
  int64_t rx = rowsx(in), ry = rowsy(in);
  int pts[4][3] = {{2, 1, 0}, {(int)rx - 3, (int)rx - 2, (int)rx - 1}, {(int)ry - 3, (int)ry - 2, (int)ry - 1},
                   {2, 1, 0}};
  for (int s = 0; s < 4; s++)
    for (int k = 0; k < 3; k++) in->bc_points[s][f][k] = pts[s][k];
  for (int s = 0; s < 6; s++) in->bc_type[s][f] = t[s];
  if (in->bc_type[0][f] == 3 || in->bc_type[1][f] == 3) in->bc_type[0][f] = in->bc_type[1][f] = 3;
  if (in->bc_type[2][f] == 3 || in->bc_type[3][f] == 3) in->bc_type[2][f] = in->bc_type[3][f] = 3;
  if (in->bc_type[4][f] == 3 || in->bc_type[5][f] == 3) in->bc_type[4][f] = in->bc_type[5][f] = 3;
  in->bc_type[4][f] = in->bc_type[5][f] = 0; /* 2-D */
}

static void bc_proxies(gp_input *in, int f) {
  int rx = (int)rowsx(in), ry = (int)rowsy(in);
  int low[3] = {4, 5, 6};
  int hix[3] = {rx - 5, rx - 6, rx - 7}, hiy[3] = {ry - 5, ry - 6, ry - 7};
  const int *src[4][4] = {{0}};
  /* [side][type] */
  src[0][1] = low; src[0][3] = hix;   /* X- */
  src[1][1] = hix; src[1][3] = low;   /* X+ */
  src[2][1] = hiy; src[2][3] = low;   /* Y+ */
  src[3][1] = low; src[3][3] = hiy;   /* Y- */
  for (int s = 0; s < 4; s++) {
    int t = in->bc_type[s][f];
    if (t == 1 || t == 3) {
      for (int k = 0; k < 3; k++) in->bc_proxy[s][f][k] = src[s][t][k];
      in->bc_proxy_set[s][f] = 1;
    }
  }
}

static int field_index(const char *name) {
  if (strcmp(name, "phi") == 0) return 0;
  if (strcmp(name, "mu") == 0) return 1;
  if (strcmp(name, "T") == 0) return 2;
  return -1;
}

/* ------------------------------------------------------------------------------------------------------------
 * Input.in
 * ---------------------------------------------------------------------------------------------------------- */
int64_t gp_rows_x(const gp_params *p) { return p->mesh_x + 6; }
int64_t gp_rows_y(const gp_params *p) { return p->mesh_y + 6; }

static int parse_pass1(FILE *fr, gp_input *in, int quiet, char *err, size_t errlen) {
  gp_params *p = &in->p;
  char line[1000], key[128] = "", val[128] = "";
  int rc = GP_OK;
  int fn_aniso = 1; /* FUNCTION_ANISOTROPY default (GPs/functions/global_vars.h:50) */
  int fn_w = 1;
  int N = 0, K = 0;
  while (fgets(line, sizeof line, fr)) {
    sscanf(line, "%100s = %100[^;];", key, val);
    if (key[0] == '#') continue;
    const char *v = val;
#define IS(k) (strcmp(key, k) == 0)
    if (IS("DIMENSION")) { in->dimension = atoi(v); in->seen |= SEEN_DIM; }
    else if (IS("MESH_X")) { p->mesh_x = atol(v); in->seen |= SEEN_MX; }
    else if (IS("MESH_Y")) { p->mesh_y = atol(v); in->seen |= SEEN_MY; }
    else if (IS("MESH_Z")) in->mesh_z = atol(v);
    else if (IS("DELTA_X")) { p->dx = atof(v); in->seen |= SEEN_DX; }
    else if (IS("DELTA_Y")) { p->dy = atof(v); in->seen |= SEEN_DY; }
    else if (IS("DELTA_Z")) in->dz = atof(v);
    else if (IS("DELTA_t")) { p->dt = atof(v); in->seen |= SEEN_DT; }
    else if (IS("NUMPHASES")) {
      N = atoi(v);
      if (N < 2 || N > P) { seterr(err, errlen, "NUMPHASES=%d outside [2,%d]", N, P); return GP_E_ARG; }
      p->nphases = N;
      in->seen |= SEEN_N;
    } else if (IS("NUMCOMPONENTS")) {
      K = atoi(v);
      if (K < 2 || K > NCM + 1) { seterr(err, errlen, "NUMCOMPONENTS=%d outside [2,%d]", K, NCM + 1); return GP_E_ARG; }
      p->ncomp = K;
      in->seen |= SEEN_K;
    }
    else if (IS("epsilon")) { p->epsilon = atof(v); in->seen |= SEEN_EPS; }
    else if (IS("tau")) { p->tau = atof(v); in->seen |= SEEN_TAU; }
    else if (IS("NTIMESTEPS")) in->ntimesteps = atol(v);
    else if (IS("SAVET")) in->savet = atol(v);
    else if (IS("NSMOOTH")) in->nsmooth = atol(v);
    else if (IS("STARTTIME")) in->starttime = atol(v);
    else if (IS("RESTART")) in->restart = atol(v);
    else if (IS("R")) in->R = atof(v);
    else if (IS("V")) in->V = atof(v);
    else if (IS("WRITEFORMAT")) {
      if (strcmp(v, "ASCII") == 0) in->ascii = 1;
      else if (strcmp(v, "BINARY") == 0) in->ascii = 0;
    }
    else if (IS("TRACK_PROGRESS")) in->track_progress = atol(v);
    else if (IS("TEMPERATURE_SCALE")) in->temperature_scale = atof(v);
    else if (IS("Equilibrium_temperature")) { p->Teq = atof(v); in->seen |= SEEN_TEQ; }
    else if (IS("Filling_temperature")) in->Tfill = atof(v);
    else if (IS("T")) { p->T = atof(v); in->seen |= SEEN_T; }
    else if (IS("tilt_angle")) in->tilt_angle = atof(v);
    else if (IS("Rtheta")) in->rtheta = atof(v);
    else if (IS("COMPONENTS") || IS("PHASES")) {
      char tok[16][100];
      int nt = split_tuple(v, tok, 16);
      int lim = IS("PHASES") ? P : NCM + 1;
      for (int i = 0; i < nt && i < lim; i++) {
        char *dst = IS("PHASES") ? in->phases[i] : in->components[i];
        strncpy(dst, tok[i], GP_NAMELEN - 1);
        dst[GP_NAMELEN - 1] = 0;
      }
    }
    else if (IS("Function_anisotropy") && N > 0 && K - 1 > 0) fn_aniso = atoi(v);
    else if (IS("Anisotropy_type") && fn_aniso != 0) p->fold = atoi(v);
    else if (IS("Rotation_matrix") && fn_aniso != 0) {
      if (N <= 0 || K - 1 <= 0) { seterr(err, errlen, "Rotation_matrix before NUMPHASES/NUMCOMPONENTS"); return GP_E_ARG; }
      if ((rc = set_rotation(in, v, quiet, err, errlen))) return rc;
    }
    else if (IS("OBSTACLE")) { /* superseded by Function_W (initialize_functions_solverloop.h:21-32) */ }
    else if (IS("Function_W") && N > 0 && K - 1 > 0) fn_w = atoi(v);
    else if (IS("ISOTHERMAL")) {
      in->isothermal = atoi(v);
      in->seen |= SEEN_ISO;
      if (!in->isothermal) p->tempgrady = 1; /* TEMPGRADY follows ISOTHERMAL = 0 and stays set */
    }
    else if (IS("Tempgrady") && p->tempgrady) {
      /* {base_temp, DeltaT, Distance, gradient_OFFSET, velocity}, read only once TEMPGRADY is set */
      char tok[8][100];
      int nt = split_tuple(v, tok, 8);
      if (nt < 5) { seterr(err, errlen, "Tempgrady needs {base_temp, DeltaT, Distance, gradient_OFFSET, velocity}"); return GP_E_ARG; }
      p->tg_base_temp = atof(tok[0]);
      p->tg_deltaT = atof(tok[1]);
      p->tg_distance = atof(tok[2]);
      p->tg_offset = atof(tok[3]);
      p->tg_velocity = atof(tok[4]);
      in->seen |= SEEN_TG;
    }
    else if (IS("WRITEHDF5")) in->writehdf5 = atoi(v);
    else if (IS("Shift")) in->shift = atoi(v);
    else if (IS("Shiftj")) { if (in->shift) in->shiftj = atol(v); }
    else if (IS("BINARY")) { in->binary = atoi(v); in->ternary = 0; in->dilute = 0; }
    else if (IS("TERNARY")) { in->ternary = atoi(v); in->binary = 0; in->dilute = 0; }
    else if (IS("DILUTE")) { in->dilute = atoi(v); in->binary = 0; in->ternary = 0; }
    else if (IS("Writecomposition")) in->writecomposition = atoi(v);
    else if (IS("Noise_phasefield")) in->noise_phasefield = atoi(v);
    else if (IS("Amp_Noise_Phase")) { if (in->noise_phasefield) in->amp_noise_phase = atof(v); }
    else if (IS("GAMMA") && N > 0) { if ((rc = fill_pairs(p->gamma, v, N, err, errlen, "GAMMA"))) return rc; in->seen |= SEEN_GAMMA; }
    else if (IS("Tau") && N > 0) { if ((rc = fill_pairs(p->tau_ab, v, N, err, errlen, "Tau"))) return rc; in->seen |= SEEN_TAUAB; }
    else if (IS("dab") && N > 0) { if ((rc = fill_pairs(p->dab, v, N, err, errlen, "dab"))) return rc; in->seen |= SEEN_DAB; }
    else if ((IS("ec") || IS("e2") || IS("e4") || IS("beta")) && N > 0) { /* unused by FOLD 4 */ }
    else if (IS("Gamma_abc") && N > 0) { if ((rc = fill_triples(p->gamma_abc, v, N, err, errlen))) return rc; in->seen |= SEEN_GABC; }
    else if (IS("DIFFUSIVITY") && N > 0 && K - 1 > 0) {
      /* populate_diffusivity_matrix (GPs/functions/utility_functions.h:163-218) */
      char tok[40][100];
      int nt = split_tuple(v, tok, 40);
      int NC = K - 1;
      if (nt < 2 + NC) { seterr(err, errlen, "DIFFUSIVITY needs {diag, phase, D11..}"); return GP_E_ARG; }
      long ph = atol(tok[1]);
      if (ph < 0 || ph >= N) { seterr(err, errlen, "DIFFUSIVITY phase %ld out of range", ph); return GP_E_ARG; }
      double *D = &p->diffusivity[ph * NCM * NCM];
      int l = 1;
      for (int i = 0; i < NC; i++) { D[i * NCM + i] = atof(tok[1 + l]); l++; }
      if (strcmp(tok[0], "1") == 0) {
        for (int i = 0; i < NC; i++)
          for (int j = i + 1; j < NC; j++) { D[i * NCM + j] = 0.0; D[j * NCM + i] = 0.0; }
      } else {
        if (nt < 2 + NC + NC * NC) { seterr(err, errlen, "full DIFFUSIVITY needs %d values", 2 + NC + NC * NC); return GP_E_ARG; }
        for (int i = 0; i < NC; i++)
          for (int j = 0; j < NC; j++) { D[i * NCM + j] = atof(tok[1 + l]); l++; }
      }
      in->seen_D[ph] = 1;
    }
    else if ((IS("ceq") || IS("cfill") || IS("slopes")) && N > 0 && K - 1 > 0) {
      /* populate_thermodynamic_matrix (GPs/functions/utility_functions.h:259-292) */
      char tok[16][100];
      int nt = split_tuple(v, tok, 16);
      if (nt < 2 + K - 1) { seterr(err, errlen, "%s needs {a, b, c1..}", key); return GP_E_ARG; }
      int a = atoi(tok[0]), b = atoi(tok[1]);
      if (a < 0 || b < 0 || a >= N || b >= N) { seterr(err, errlen, "%s phase index out of range", key); return GP_E_ARG; }
      double (*M)[P][NCM] = IS("ceq") ? in->ceq : IS("cfill") ? in->cfill : in->slopes;
      for (int i = 0; i < K - 1; i++) M[a][b][i] = atof(tok[2 + i]);
    }
    else if (IS("Function_F") && N > 0 && K - 1 > 0) in->fn_f = atoi(v);
    else if (IS("A") && N > 0 && K - 1 > 0) {
      /* populate_A_matrix (GPs/functions/utility_functions.h:219-257) */
      char tok[32][100];
      int nt = split_tuple(v, tok, 32);
      int NC = K - 1;
      if (nt < 1 + NC + NC * (NC - 1) / 2) { seterr(err, errlen, "A needs {phase, A11..Ann, A12..}"); return GP_E_ARG; }
      long ph = atol(tok[0]);
      if (ph < 0 || ph >= N) { seterr(err, errlen, "A phase %ld out of range", ph); return GP_E_ARG; }
      double *A = &p->A[ph * NCM * NCM];
      int l = 1;
      for (int i = 0; i < NC; i++) { A[i * NCM + i] = atof(tok[l]); l++; }
      for (int i = 0; i < NC; i++)
        for (int j = i + 1; j < NC; j++) { A[i * NCM + j] = atof(tok[l]); A[j * NCM + i] = A[i * NCM + j]; l++; }
      in->seen_A[ph] = 1;
    }
    else if ((IS("EIGEN_STRAIN") || IS("VOIGT_ISOTROPIC")) && N > 0) { /* elasticity: not on this path */ }
    else if (IS("BOUNDARY") || IS("BOUNDARY_VALUE")) {
      /* second pass */
      if (!quiet) printf("Unrecongized parameter : \"%s\"\n", key);
    }
    else {
      if (!quiet) printf("Unrecongized parameter : \"%s\"\n", key);
    }
#undef IS
  }
  p->fn_aniso = fn_aniso;
  p->fn_w = fn_w;
  return GP_OK;
}

static int parse_pass2(FILE *fr, gp_input *in, char *err, size_t errlen) {
  char line[1000], key[128] = "", val[128] = "";
  while (fgets(line, sizeof line, fr)) {
    sscanf(line, "%100s = %100[^;];", key, val);
    if (key[0] == '#') continue;
    if (in->p.nphases <= 0) continue;
    if (strcmp(key, "BOUNDARY") == 0) {
      char tok[8][100];
      int nt = split_tuple(val, tok, 8);
      if (nt < 7) { seterr(err, errlen, "BOUNDARY needs {field, X+, X-, Y+, Y-, Z+, Z-}"); return GP_E_ARG; }
      int f = field_index(tok[0]);
      /* tuple order X+, X-, Y+, Y-, Z+, Z-  ->  sides X-(0), X+(1), Y+(2), Y-(3), Z+(4), Z-(5) */
      int t[6] = {atoi(tok[2]), atoi(tok[1]), atoi(tok[3]), atoi(tok[4]), atoi(tok[5]), atoi(tok[6])};
      if (f >= 0) bc_types(in, f, t);
    } else if (strcmp(key, "BOUNDARY_VALUE") == 0) {
      char tok[8][100];
      int nt = split_tuple(val, tok, 8);
      if (nt < 1) continue;
      int f = field_index(tok[0]);
      if (f >= 0) bc_proxies(in, f);
    }
  }
  return GP_OK;
}

void gp_property_matrices(gp_input *in) {
  gp_params *p = &in->p;
  const int N = p->nphases, NC = p->ncomp - 1, L = N - 1;
  double DELTA_T[P][P] = {{0}}, DELTA_C[P][NCM] = {{0}}, dcbdT[P][P][NCM] = {{{0}}};
  /* cmu = (muc)^-1, muc_ii = 2 A_ii, muc_ij = A_ij (functionF_01.h:181-192) */
  for (int a = 0; a < N; a++) {
    double m[4][4] = {{0}}, inv[4][4] = {{0}};
    for (int i = 0; i < NC; i++)
      for (int j = 0; j < NC; j++) m[i][j] = (i == j) ? 2.0 * p->A[(a * NCM + i) * NCM + j] : p->A[(a * NCM + i) * NCM + j];
    lu_inverse(NC, m, inv);
    for (int i = 0; i < NC; i++)
      for (int j = 0; j < NC; j++) p->cmu[(a * NCM + i) * NCM + j] = inv[i][j];
  }
  for (int a = 0; a < N - 1; a++) {
    DELTA_T[a][L] = 0.0;
    DELTA_T[a][a] = 0.0;
    for (int k = 0; k < NC; k++) {
      DELTA_T[a][a] += in->slopes[a][a][k] * (in->ceq[a][L][k] - in->ceq[a][a][k]);
      DELTA_T[a][L] += in->slopes[a][L][k] * (in->ceq[a][L][k] - in->ceq[a][a][k]);
      DELTA_C[a][k] = in->ceq[a][L][k] - in->ceq[a][a][k];
    }
    for (int k = 0; k < NC; k++) {
      dcbdT[a][a][k] = DELTA_C[a][k] / DELTA_T[a][a];
      dcbdT[a][L][k] = DELTA_C[a][k] / DELTA_T[a][L];
    }
    for (int k = 0; k < NC; k++) {
      double v = 2.0 * (p->A[(L * NCM + k) * NCM + k] * dcbdT[a][L][k] - p->A[(a * NCM + k) * NCM + k] * dcbdT[a][a][k]);
      for (int i = 0; i < NC; i++)
        if (k != i) v += (p->A[(L * NCM + k) * NCM + i] * dcbdT[a][L][i] - p->A[(a * NCM + k) * NCM + i] * dcbdT[a][a][i]);
      p->dBbdT[a * NCM + k] = v;
    }
  }
  for (int k = 0; k < NC; k++) p->dBbdT[L * NCM + k] = 0.0;
  /* what B(T), C(T) of a non-isothermal run need per row (gp_frame.h); dcbdT_phase stays 0 as in the serial driver
   * (initialize_variables.h:22, never assigned afterwards) */
  for (int a = 0; a < N - 1; a++) {
    for (int k = 0; k < NC; k++) {
      p->ceq_sol[a * NCM + k] = in->ceq[a][a][k];
      p->ceq_liq[a * NCM + k] = in->ceq[a][L][k];
      p->delta_c[a * NCM + k] = DELTA_C[a][k];
    }
    p->delta_T_sol[a] = DELTA_T[a][a];
    p->delta_T_liq[a] = DELTA_T[a][L];
  }
  /* function_B / function_C (functionF_01.h:62-104) at Teq and T */
  for (int pass = 0; pass < 2; pass++) {
    double Tv = pass == 0 ? p->Teq : p->T;
    for (int a = 0; a < N; a++) {
      double cl[NCM], cs[NCM];
      double Cv = 0.0;
      if (a != L) {
        for (int k = 0; k < NC; k++) {
          cl[k] = in->ceq[a][L][k] - (DELTA_C[a][k]) * (p->Teq - Tv) / (DELTA_T[a][L]);
          cs[k] = in->ceq[a][a][k] - (DELTA_C[a][k]) * (p->Teq - Tv) / (DELTA_T[a][a]);
        }
      }
      for (int i = 0; i < NC; i++) {
        double Bai = 0.0;
        if (a != L) {
          double sc = 0.0;
          for (int k = 0; k < NC; k++)
            if (k != i) sc += p->A[(L * NCM + k) * NCM + i] * cl[k] - p->A[(a * NCM + k) * NCM + i] * cs[k];
          Bai = (2.0 * (p->A[(L * NCM + i) * NCM + i] * cl[i] - p->A[(a * NCM + i) * NCM + i] * cs[i]) + sc);
        }
        if (pass == 0) p->Beq[a * NCM + i] = Bai;
        else p->B[a * NCM + i] = Bai;
      }
      if (a != L && pass == 1) {
        for (int i = 0; i < NC; i++)
          for (int j = 0; j < NC; j++)
            if (i <= j) Cv += (p->A[(a * NCM + i) * NCM + j] * cs[i] * cs[j] - p->A[(L * NCM + i) * NCM + j] * cl[i] * cl[j]);
      }
      if (pass == 1) p->C[a] = Cv;
    }
  }
}

int gp_validate_input(const gp_input *in, char *err, size_t errlen) {
  const gp_params *p = &in->p;
  const uint32_t need = SEEN_DIM | SEEN_MX | SEEN_MY | SEEN_DX | SEEN_DY | SEEN_DT | SEEN_N | SEEN_K | SEEN_EPS |
                        SEEN_GAMMA | SEEN_TAUAB | SEEN_T | SEEN_ISO;
  if ((in->seen & need) != need) {
    seterr(err, errlen, "Input.in lacks a required key (DIMENSION, MESH_X/Y, DELTA_X/Y/t, NUMPHASES, "
                        "NUMCOMPONENTS, epsilon, GAMMA, Tau, T, ISOTHERMAL)");
    return GP_E_ARG;
  }
  if (in->dimension != 2) { seterr(err, errlen, "only DIMENSION = 2 is supported"); return GP_E_ARG; }
  if (p->mesh_x < 4 || p->mesh_y < 4) { seterr(err, errlen, "MESH_X/MESH_Y must be >= 4"); return GP_E_ARG; }
  if (p->tempgrady && !(in->seen & SEEN_TG)) { seterr(err, errlen, "ISOTHERMAL = 0 needs Tempgrady after it"); return GP_E_ARG; }
  if (p->tempgrady && (p->tg_distance == 0.0 || p->dy == 0.0)) { seterr(err, errlen, "Tempgrady Distance must be nonzero"); return GP_E_ARG; }
  if (in->fn_f != 1) { seterr(err, errlen, "only Function_F = 1 (parabolic) is supported"); return GP_E_ARG; }
  if (p->fn_w != 1 && p->fn_w != 2) { seterr(err, errlen, "Function_W must be 1 or 2"); return GP_E_ARG; }
  if (p->fn_aniso != 0 && p->fn_aniso != 1) { seterr(err, errlen, "Function_anisotropy must be 0 or 1"); return GP_E_ARG; }
  if (p->fn_aniso == 1 && p->fold != 4) { seterr(err, errlen, "Function_anisotropy = 1 needs Anisotropy_type = 4"); return GP_E_ARG; }
  if (p->fn_aniso == 1 && !(in->seen & SEEN_DAB)) { seterr(err, errlen, "Function_anisotropy = 1 needs dab"); return GP_E_ARG; }
  if (in->noise_phasefield) { seterr(err, errlen, "Noise_phasefield is not supported (rand() stream)"); return GP_E_ARG; }
  if (in->shift && (in->shiftj < 1 || in->shiftj >= rowsy(in) - 1)) { seterr(err, errlen, "Shiftj must lie inside the allocated rows"); return GP_E_ARG; }
  if (p->nphases >= 3 && !(in->seen & SEEN_GABC)) { seterr(err, errlen, "NUMPHASES >= 3 needs Gamma_abc"); return GP_E_ARG; }
  if (p->solve == GP_SOLVE_BINARY && p->ncomp != 2) { seterr(err, errlen, "BINARY = 1 needs NUMCOMPONENTS = 2"); return GP_E_ARG; }
  if (p->solve == GP_SOLVE_TERNARY && p->ncomp != 3) { seterr(err, errlen, "TERNARY = 1 needs NUMCOMPONENTS = 3"); return GP_E_ARG; }
  if (p->dx <= 0 || p->dy <= 0 || p->dt <= 0 || p->epsilon <= 0) { seterr(err, errlen, "DELTA_X/Y/t and epsilon must be > 0"); return GP_E_ARG; }
  for (int a = 0; a < p->nphases; a++) {
    if (!in->seen_A[a]) { seterr(err, errlen, "missing A for phase %d", a); return GP_E_ARG; }
    if (!in->seen_D[a]) { seterr(err, errlen, "missing DIFFUSIVITY for phase %d", a); return GP_E_ARG; }
    if (p->fn_aniso == 1)
      for (int b = 0; b < p->nphases; b++)
        if (b != a && !in->seen_rot[a][b]) { seterr(err, errlen, "missing Rotation_matrix for pair (%d,%d)", a, b); return GP_E_ARG; }
  }
  for (int f = 0; f < 2; f++)
    for (int s = 0; s < 4; s++) {
      int t = in->bc_type[s][f];
      if (t != 1 && t != 3) {
        seterr(err, errlen, "BOUNDARY %s side %d: type %d is not supported (Neumann 1 / periodic 3 only)",
               f ? "mu" : "phi", s, t);
        return GP_E_ARG;
      }
      if (!in->bc_proxy_set[s][f]) {
        seterr(err, errlen, "BOUNDARY_VALUE for %s must follow its BOUNDARY line", f ? "mu" : "phi");
        return GP_E_ARG;
      }
    }
  if (p->nphases > P || p->ncomp - 1 > NCM) { seterr(err, errlen, "too many phases/components"); return GP_E_ARG; }
  return GP_OK;
}

static double host_Mu(const gp_params *p, const double *c, double T, int a, int i);

int gp_read_input(const char *input_path, gp_input *in, int quiet, char *err, size_t errlen) {
  memset(in, 0, sizeof(*in));
  in->p.jat_skip_row = -1;
  in->p.fold = 0;
  FILE *fr = fopen(input_path, "rt");
  if (!fr) {
    seterr(err, errlen, "file %s not found", input_path);
    return GP_E_IO;
  }
  int rc = parse_pass1(fr, in, quiet, err, errlen);
  if (rc) { fclose(fr); return rc; }
  rewind(fr);
  rc = parse_pass2(fr, in, err, errlen);
  fclose(fr);
  if (rc) return rc;
  gp_params *p = &in->p;
  p->solve = in->binary ? GP_SOLVE_BINARY : in->ternary ? GP_SOLVE_TERNARY : in->dilute ? GP_SOLVE_DILUTE : GP_SOLVE_LU;
  p->jat_skip_row = (int32_t)(p->mesh_y - 2);
  for (int a = 0; a < P; a++)
    for (int b = 0; b < P; b++)
      for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++) {
          p->rot[((a * P + b) * 2 + i) * 2 + j] = in->rot3[a][b][i][j];
          p->rot_inv[((a * P + b) * 2 + i) * 2 + j] = in->rot3_inv[a][b][i][j];
        }
  /* halo maps: allocated x = 0,1,2 <- X- proxies of points 0,1,2 (= x 2,1,0); x = rx-3.. <- X+ */
  for (int f = 0; f < 2; f++) {
    for (int k = 0; k < 3; k++) {
      p->halo_src_x[f][2 - k] = in->bc_proxy[0][f][k];
      p->halo_src_x[f][3 + k] = in->bc_proxy[1][f][k];
      p->halo_src_y[f][2 - k] = in->bc_proxy[3][f][k];
      p->halo_src_y[f][3 + k] = in->bc_proxy[2][f][k];
    }
  }
  /* T's boundary copies along y (boundary[.][2]); a side that copies nothing leaves its halo column as it is */
  {
    const int64_t ry = rowsy(in);
    const int32_t hy[6] = {0, 1, 2, (int32_t)ry - 3, (int32_t)ry - 2, (int32_t)ry - 1};
    for (int k = 0; k < 3; k++) {
      p->halo_src_T_y[2 - k] = in->bc_proxy_set[3][2] ? in->bc_proxy[3][2][k] : hy[2 - k];
      p->halo_src_T_y[3 + k] = in->bc_proxy_set[2][2] ? in->bc_proxy[2][2][k] : hy[3 + k];
    }
  }
  p->shift = in->shift;
  p->shiftj = in->shift ? in->shiftj : 0;
  in->max_length = (int)strlen(in->phases[0]);
  for (int a = 1; a < p->nphases; a++)
    if ((int)strlen(in->phases[a]) > in->max_length) in->max_length = (int)strlen(in->phases[a]);
  for (int k = 0; k < p->ncomp - 1; k++)
    if ((int)strlen(in->components[k]) > in->max_length) in->max_length = (int)strlen(in->components[k]);
  gp_property_matrices(in);
  /* mu of the liquid rows the moving frame adds: Mu(cfill[L][L], Tfill, L) (GPs/functions/shift.h:24-38) */
  {
    const int L = p->nphases - 1;
    for (int k = 0; k < p->ncomp - 1; k++) p->mu_fill[k] = host_Mu(p, in->cfill[L][L], in->Tfill, L, k);
  }
  return GP_OK;
}

/* ------------------------------------------------------------------------------------------------------------
 * Model functions used on the host (fill and output)
 * ---------------------------------------------------------------------------------------------------------- */
static double host_Mu(const gp_params *p, const double *c, double T, int a, int i) {
  const int NC = p->ncomp - 1;
  double s = 2.0 * p->A[(a * NCM + i) * NCM + i] * c[i] + (p->Beq[a * NCM + i] + p->dBbdT[a * NCM + i] * (T - p->Teq));
  for (int j = 0; j < NC; j++)
    if (i != j) s += p->A[(a * NCM + i) * NCM + j] * c[j];
  return s;
}

static double host_c_mu(const gp_params *p, const double *mu, double T, int a, int i) {
  const int NC = p->ncomp - 1;
  double s = 0.0;
  for (int k = 0; k < NC; k++) s += p->cmu[(a * NCM + i) * NCM + k] * (mu[k] - (p->Beq[a * NCM + k] + p->dBbdT[a * NCM + k] * (T - p->Teq)));
  return s;
}

static double host_hphi(const gp_params *p, const double *phi, int a) {
  const int N = p->nphases;
  double s1 = 0.0;
  double s = 3.0 * phi[a] * phi[a] - 2.0 * phi[a] * phi[a] * phi[a];
  for (int b = 0; b < N; b++)
    for (int c = 0; c < N; c++)
      if (b != a && c != a && b < c) s1 += phi[b] * phi[c];
  s1 *= 2.0 * phi[a];
  return s + s1;
}

static double composition_at(const gp_params *p, const double *phi_cell, const double *mu_cell, int k, double T) {
  double c = 0.0;
  for (int b = 0; b < p->nphases; b++) c += host_c_mu(p, mu_cell, T, b, k) * host_hphi(p, phi_cell, b);
  return c;
}

double gp_composition(const gp_params *p, const double *phi_cell, const double *mu_cell, int k) {
  return composition_at(p, phi_cell, mu_cell, k, p->T);
}

double gp_composition_at(const gp_params *p, const double *phi_cell, const double *mu_cell, int k, double T) {
  return composition_at(p, phi_cell, mu_cell, k, T);
}

/* T(y) = temp_bottom + y GRADIENT with temp_bottom = base_temp - BASE_POS GRADIENT, BASE_POS = gradient_OFFSET/dy
 * - shift_OFFSET (+ velocity/dy t dt after step t), GRADIENT = DeltaT dy / Distance (GPs/microsim_gp.c:77-80,
 * 108-111; apply_temperature_gradientY, GPs/functions/Temperature_gradient.h:4-29); at t = 0 the boundary copies
 * of T follow (GPs/microsim_gp.c:84) */
void gp_temperature_rows(const gp_params *p, int64_t t, int64_t shift_offset, double *T_rows) {
  const int64_t ry = gp_rows_y(p);
  if (!p->tempgrady) {
    for (int64_t y = 0; y < ry; y++) T_rows[y] = p->T;
    return;
  }
  double base;
  if (t == 0) base = (p->tg_offset / p->dy) - (double)shift_offset;
  else base = (p->tg_offset / p->dy) - (double)shift_offset + ((p->tg_velocity / p->dy) * ((double)t * p->dt));
  const double G = (p->tg_deltaT) * p->dy / (p->tg_distance);
  const double bottom = p->tg_base_temp - base * G;
  for (int64_t y = 0; y < ry; y++) T_rows[y] = bottom + (double)y * G;
  if (t == 0) {
    const int64_t hy[6] = {0, 1, 2, ry - 3, ry - 2, ry - 1};
    for (int k = 0; k < 6; k++) T_rows[hy[k]] = T_rows[p->halo_src_T_y[k]];
  }
}

void gp_composition_grid(const gp_params *p, const double *phi, const double *mu, double *c, int64_t ncells) {
  const int N = p->nphases, NC = p->ncomp - 1;
  for (int64_t i = 0; i < ncells; i++) {
    double pc[GP_MAXP], mc[GP_MAXNC];
    for (int a = 0; a < N; a++) pc[a] = phi[(int64_t)a * ncells + i];
    for (int k = 0; k < NC; k++) mc[k] = mu[(int64_t)k * ncells + i];
    for (int k = 0; k < NC; k++) c[(int64_t)k * ncells + i] = gp_composition(p, pc, mc, k);
  }
}

/* ------------------------------------------------------------------------------------------------------------
 * fill_domain (GPs/functions/fill_domain.h, GPs/functions/filling.h)
 * ---------------------------------------------------------------------------------------------------------- */
typedef struct { double *phi; int64_t rx, ry, n; int N; } grid_t;

static inline double *PH(grid_t *g, int a, int64_t x, int64_t y) { return &g->phi[(int64_t)a * g->n + x * g->ry + y]; }

static void fill_liquid_cell(grid_t *g, int64_t x, int64_t y, int quiet) {
  const int N = g->N, b = N - 1;
  double sum = 0.0;
  for (int a = 0; a < N - 1; a++) sum += *PH(g, a, x, y);
  if (sum > 1.0) {
    if (!quiet) printf("Wrong filling operation, will fill it with liquid\n");
    *PH(g, b, x, y) = 1.0;
    for (int a = 0; a < N - 1; a++)
      if (a != b) *PH(g, a, x, y) = 0.0;
  } else {
    *PH(g, b, x, y) = 1.0 - sum;
  }
}

static void fill_solid_cell(grid_t *g, int64_t x, int64_t y, int b, int inside) {
  if (inside) {
    *PH(g, b, x, y) = 1.0;
    for (int a = 0; a < g->N; a++)
      if (a != b) *PH(g, a, x, y) = 0.0;
  } else if (*PH(g, b, x, y) != 1.0) {
    *PH(g, b, x, y) = 0.0;
  }
}

/* ------------------------------------------------------------------------------------------------------------
 * The MPI driver's random and pattern fillers (GPm/functions/fill_domain.h:146-332, GPm/functions/filling.h).
 * They work on the allocated grid (rows_x x rows_y, halo included) with centres and boxes in allocated
 * coordinates, as the single-rank MPI driver does.
 * ---------------------------------------------------------------------------------------------------------- */

/* The POSIX drand48 generator: X(n+1) = (0x5DEECE66D X(n) + 0xB) mod 2^48. The reference never seeds it, so its runs
 * start from the POSIX initial state 0x1234ABCD330E and draw the same sequence every time; drand48 = X / 2^48,
 * lrand48 = X >> 17. */
typedef struct { uint64_t x; } rng48_t;
static void rng48_init(rng48_t *r) { r->x = 0x1234ABCD330EULL; }
static uint64_t rng48_next(rng48_t *r) {
  r->x = (0x5DEECE66DULL * r->x + 0xBULL) & 0xFFFFFFFFFFFFULL;
  return r->x;
}
static double rng48_drand(rng48_t *r) { return (double)rng48_next(r) / 281474976710656.0; }
static long rng48_lrand(rng48_t *r) { return (long)(rng48_next(r) >> 17); }

/* MT19937 as GSL's default generator runs it (Matsumoto and Nishimura 1998, with the Knuth seeding
 * mt[i] = 1812433253 (mt[i-1] ^ (mt[i-1] >> 30)) + i, seed 0 taken as 4357): uniform = draw / 2^32,
 * uniform_int(n) = draw / floor((2^32 - 1) / n), redrawn while >= n. The reference seeds it with time(0). */
typedef struct { uint32_t mt[624]; int mti; } mt_t;
static void mt_seed(mt_t *s, unsigned long seed) {
  if (seed == 0) seed = 4357;
  s->mt[0] = (uint32_t)(seed & 0xffffffffUL);
  for (int i = 1; i < 624; i++) s->mt[i] = (uint32_t)(1812433253UL * (s->mt[i - 1] ^ (s->mt[i - 1] >> 30)) + (unsigned)i);
  s->mti = 624;
}
static uint32_t mt_next(mt_t *s) {
  if (s->mti >= 624) {
    for (int k = 0; k < 624; k++) {
      const uint32_t y = (s->mt[k] & 0x80000000U) | (s->mt[(k + 1) % 624] & 0x7fffffffU);
      s->mt[k] = s->mt[(k + 397) % 624] ^ (y >> 1) ^ ((y & 1U) ? 0x9908b0dfU : 0U);
    }
    s->mti = 0;
  }
  uint32_t y = s->mt[s->mti++];
  y ^= y >> 11;
  y ^= (y << 7) & 0x9d2c5680U;
  y ^= (y << 15) & 0xefc60000U;
  y ^= y >> 18;
  return y;
}
static double mt_uniform(mt_t *s) { return mt_next(s) / 4294967296.0; }
static unsigned long mt_uniform_int(mt_t *s, unsigned long n) {
  const unsigned long scale = 0xffffffffUL / n;
  unsigned long k;
  do k = mt_next(s) / scale; while (k >= n);
  return k;
}

/* seed of the GSL-driven fillers: GP_FILL_SEED when set (reproducible runs), else time(0) as in the reference */
static unsigned long fill_seed(void) {
  const char *e = getenv("GP_FILL_SEED");
  return e && *e ? strtoul(e, NULL, 10) : (unsigned long)time(NULL);
}

/* fill_phase_cube (GPm/functions/filling.h:22-71) over the allocated grid for a solid b (the matrix pass is separate) */
static void fill_cube_solid(grid_t *g, int b, long x0, long x1, long y0, long y1, long z0, long z1) {
  for (int64_t x = 0; x < g->rx; x++)
    for (int64_t y = 0; y < g->ry; y++)
      fill_solid_cell(g, x, y, b, x >= x0 && x <= x1 && y >= y0 && y <= y1 && 0 >= z0 && 0 <= z1);
}

static void fill_liquid_all(grid_t *g, int quiet) {
  for (int64_t x = 0; x < g->rx; x++)
    for (int64_t y = 0; y < g->ry; y++) fill_liquid_cell(g, x, y, quiet);
}

/* the matrix (last phase) set to 1 on every allocated cell, the other phases left as they are */
static void fill_matrix_one(grid_t *g) {
  for (int64_t x = 0; x < g->rx; x++)
    for (int64_t y = 0; y < g->ry; y++) *PH(g, g->N - 1, x, y) = 1.0;
}

/* fill_phase_cylinder_random (GPm/functions/filling.h:632-698) with checkOverlap (:7-20) and
 * fill_phase_cylinder_occupancy (:579-631); z = 0 only */
static void fill_cylinders_random(grid_t *g, rng48_t *r, int64_t MX, int64_t MY, long phase, long ppt_radius,
                                  double vf, long shield, double spread, int quiet) {
  int *occ = (int *)calloc((size_t)g->n, sizeof(int));
  const double vol = (double)MX * MY * 1;
  const double vpp = (double)(ppt_radius * ppt_radius) * M_PI;
  const int nparticles = (int)ceil(vol * vf / vpp);
  if (!quiet) printf("Domain volume = %lf, ppt_radius = %le, volume_per_particle = %lf, Number of particles = %d\n", vol,
                     (double)ppt_radius, vpp, nparticles);
  long idx = 1, count = 0;
  const long limit = 10000;
  while (idx <= nparticles) {
    const long cx = (long)(MX * rng48_drand(r)), cy = (long)(MY * rng48_drand(r));
    const double mdev = spread * ppt_radius;
    const double rad = (double)ppt_radius + (1.0 * rng48_drand(r) - 0.5) * mdev;
    if (++count > limit) {
      if (!quiet) printf("Random filling attempt count limit = %ld reached. Only %ld particles could be filled.\n", limit, idx - 1);
      break;
    }
    if (((cx - rad) < 0.0) || ((cx + rad) > (MX - 1)) || ((cy - rad) < 0.0) || ((cy + rad) > (MY - 1))) continue;
    int overlap = 0;
    if (shield)
      for (int64_t x = 0; x < g->rx && !overlap; x++)
        for (int64_t y = 0; y < g->ry; y++)
          if (occ[x * g->ry + y] == 1 && (x - cx) * (x - cx) + (y - cy) * (y - cy) + 0 <= (shield * rad) * (shield * rad)) {
            overlap = 1;
            break;
          }
    if (overlap) continue;
    for (int64_t x = 0; x < g->rx; x++)
      for (int64_t y = 0; y < g->ry; y++) {
        const int in = ((x - cx) * (x - cx) + (y - cy) * (y - cy)) <= rad * rad;
        fill_solid_cell(g, x, y, (int)phase, in);
        if (in) occ[x * g->ry + y] = 1;
      }
    idx++;
  }
  fill_liquid_all(g, quiet);
  free(occ);
}

/* fill_phase_sphere_random (GPm/functions/filling.h:516-577) with fill_phase_sphere_occupancy (:463-514) on the
 * plane z = 0. The centre's z is MESH_Z drand48() = 0, so with a positive radius every attempt fails the z-extent
 * test (0 - rad < 0) and no sphere is placed on a 2-D grid; the attempts still draw four values each, up to the
 * 1e5 limit, before the matrix pass. */
static void fill_spheres_random(grid_t *g, rng48_t *r, int64_t MX, int64_t MY, long phase, long ppt_radius,
                                double vf, long shield, double spread, int quiet) {
  int *occ = (int *)calloc((size_t)g->n, sizeof(int));
  const double vol = (double)MX * MY * 1;
  const double vpp = (double)(ppt_radius * ppt_radius * ppt_radius) * (4.0 / 3.0) * M_PI;
  const int nparticles = (int)ceil(vol * vf / vpp);
  if (!quiet) printf("Domain volume = %lf, ppt_radius = %le, volume_per_particle = %lf, Number of particles = %d\n", vol,
                     (double)ppt_radius, vpp, nparticles);
  long idx = 1, count = 0;
  const long limit = 100000;
  while (idx <= nparticles) {
    const long cx = (long)(MX * rng48_drand(r)), cy = (long)(MY * rng48_drand(r)), cz = (long)(1 * rng48_drand(r));
    const double mdev = spread * ppt_radius;
    const double rad = (double)ppt_radius + (1.0 * rng48_drand(r) - 0.5) * mdev;
    if (++count > limit) {
      if (!quiet) printf("Random filling attempt count limit = %ld reached. Only %ld particles could be filled.\n", limit, idx - 1);
      break;
    }
    if (((cx - rad) < 0.0) || ((cx + rad) > (MX - 1)) || ((cy - rad) < 0.0) || ((cy + rad) > (MY - 1)) ||
        ((cz - rad) < 0.0) || ((cz + rad) > (1 - 1)))
      continue;
    int overlap = 0;
    if (shield)
      for (int64_t x = 0; x < g->rx && !overlap; x++)
        for (int64_t y = 0; y < g->ry; y++)
          if (occ[x * g->ry + y] == 1 &&
              (x - cx) * (x - cx) + (y - cy) * (y - cy) + (0 - cz) * (0 - cz) <= (shield * rad) * (shield * rad)) {
            overlap = 1;
            break;
          }
    if (overlap) continue;
    for (int64_t x = 0; x < g->rx; x++)
      for (int64_t y = 0; y < g->ry; y++) {
        const int in = ((x - cx) * (x - cx) + (y - cy) * (y - cy) + (0 - cz) * (0 - cz)) <= rad * rad;
        fill_solid_cell(g, x, y, (int)phase, in);
        if (in) occ[x * g->ry + y] = 1;
      }
    idx++;
  }
  fill_liquid_all(g, quiet);
  free(occ);
}

/* fill_phase_voronoi_2D (GPm/functions/filling.h:700-781) over [x0, x1) x [y0, y1). A point that lands inside the
 * exclusion disc of an earlier one keeps the reference's uninitialised coordinates there; here it is left out. The
 * nearest-point search keeps the reference's start (location 0, not the first point's phase). */
static int fill_voronoi(grid_t *g, rng48_t *r, long x0, long x1, long y0, long y1, long npts, double size_min) {
  const int N = g->N;
  const long lx = x1 - x0, ly = y1 - y0;
  if (npts <= 0 || lx <= 0 || ly <= 0 || x0 < 0 || y0 < 0 || x1 > g->rx || y1 > g->ry) return GP_E_ARG;
  double *px = (double *)malloc(sizeof(double) * (size_t)npts), *py = (double *)malloc(sizeof(double) * (size_t)npts);
  int *valid = (int *)calloc((size_t)npts, sizeof(int));
  long *ph = (long *)malloc(sizeof(long) * (size_t)npts);
  char *flag = (char *)calloc((size_t)g->n, 1);
  for (long k = 0; k < npts; k++) {
    const long rx = (long)(x0 + rng48_drand(r) * lx);
    const long ry = (long)(y0 + rng48_drand(r) * ly);
    if (flag[rx * g->ry + ry] != 1) {
      px[k] = (double)rx;
      py[k] = (double)ry;
      valid[k] = 1;
      for (long x = x0; x < x1; x++)
        for (long y = y0; y < y1; y++)
          if ((py[k] - y) * (py[k] - y) + (px[k] - x) * (px[k] - x) <= size_min * size_min) flag[x * g->ry + y] = 1;
    }
  }
  for (long k = 0; k < npts; k++) ph[k] = rng48_lrand(r) % (N - 1);
  for (long x = x0; x < x1; x++)
    for (long y = y0; y < y1; y++) {
      double minimum = INFINITY;
      long location = 0, first = 1;
      for (long k = 0; k < npts; k++) {
        if (!valid[k]) continue;
        const double l = (px[k] - x) * (px[k] - x) + (py[k] - y) * (py[k] - y);
        if (first) { /* l[0] starts the search with location 0 */
          minimum = l;
          first = 0;
          if (k == 0) continue;
        }
        if (l < minimum) {
          minimum = l;
          location = ph[k];
        }
      }
      for (int s = 0; s < N - 1; s++) *PH(g, s, x, y) = location == s ? 1.0 : 0.0;
    }
  free(px);
  free(py);
  free(valid);
  free(ph);
  free(flag);
  return GP_OK;
}

/* fill_cube_pattern (GPm/functions/filling.h:73-169): a lattice of boxes of size s +- sfrac s at pitch s + gap, each
 * kept with probability 1 - 1/outof, of a random variant among the first `variants` phases */
static void fill_cube_pattern(grid_t *g, mt_t *rng, int64_t MX, int64_t MY, long variants, long sx, long sy,
                              double sfrac, long gap, double gfrac, long outof, int quiet) {
  const int N = g->N;
  if (N > 1) fill_matrix_one(g);
  ldiv_t rx = ldiv((long)MX, sx + gap), ry = ldiv((long)MY, sy + gap);
  if (rx.quot == 0) rx.quot = 1;
  if (ry.quot == 0) ry.quot = 1;
  if (!quiet) printf("Filling %ld particles.\n", rx.quot * ry.quot * 1);
  long xlo = 0, xhi = 1, ylo = 0, yhi = 1;
  for (long i = 0; i < rx.quot; i++) {
    if (MX > 1) {
      const double r = mt_uniform(rng);
      const long sgn = 2 * lround(r) - 1;
      xlo = (long)(i * (gap + sx) + gap + sgn * gfrac * gap - (2 * r - 1) * sfrac * sx);
      xhi = (long)((i + 1) * (gap + sx) + sgn * gfrac * gap + (2 * r - 1) * sfrac * sx);
    }
    for (long j = 0; j < ry.quot; j++) {
      if (MY > 1) {
        const double r = mt_uniform(rng);
        const long sgn = 2 * lround(r) - 1;
        ylo = (long)(j * (gap + sy) + gap + sgn * gfrac * gap - (2 * r - 1) * sfrac * sy);
        yhi = (long)((j + 1) * (gap + sy) + sgn * gfrac * gap + (2 * r - 1) * sfrac * sy);
      }
      if (outof > 0 && !mt_uniform_int(rng, (unsigned long)outof)) continue;
      if (variants >= N) variants = N - 1;
      const long phase = (long)(mt_uniform(rng) * variants);
      if (!quiet) printf("phase now = %ld\n", phase);
      fill_cube_solid(g, (int)phase, xlo, xhi, ylo, yhi, 0, 1);
    }
  }
}

/* fill_phase_cube_random_variants (GPm/functions/filling.h:204-295) with check_overlap_sq (:171-202): boxes of size
 * s +- sfrac s at uniform positions, skipped when the shielded box meets a cell whose matrix fraction is not
 * positive, until vf of the domain is covered. The reference retries without bound; here the attempts stop after
 * 1e7 with an error. */
static int fill_cubes_random(grid_t *g, mt_t *rng, int64_t MX, int64_t MY, long variants, long sx, long sy, long sz,
                             double sfrac, double vf, long shield, int quiet) {
  const int N = g->N;
  if (N > 1) fill_matrix_one(g);
  const long vol = (long)(MX * MY * 1);
  const double vpp = (double)(sx * sy * sz);
  int nparticles = (int)ceil(vol * vf / vpp);
  if (!quiet) printf("Domain volume = %ld, sx = %ld, sy = %ld, sz = %ld, volume_per_particle = %lf, number of particles = %d\n",
                     vol, sx, sy, sz, vpp, nparticles);
  long count = 0, attempts = 0;
  const long limit = 10000;
  while (nparticles) {
    if (++attempts > 10000000L) return GP_E_ARG;
    double r = mt_uniform(rng);
    const long sxn = (long)(sx + (2 * r - 1) * sfrac * sx), syn = (long)(sy + (2 * r - 1) * sfrac * sy);
    const long szn = (long)(sz + (2 * r - 1) * sfrac * sz);
    const long xn = (long)(MX * mt_uniform(rng)), yn = (long)(MY * mt_uniform(rng));
    const long zn = (long)(1 * mt_uniform(rng));
    const long xlo = xn - sxn / 2, xhi = xn + sxn / 2, ylo = yn - syn / 2, yhi = yn + syn / 2;
    const long zlo = zn - szn / 2, zhi = zn + szn / 2;
    int overlap = 0;
    for (long x = xlo - shield; x < xhi + shield && !overlap; x++)
      for (long y = ylo - shield; y < yhi + shield && !overlap; y++)
        for (long z = zlo - shield; z < zhi + shield; z++) {
          const long index = x * (long)g->ry + z * (long)g->ry + y;
          if (0 <= index && index <= vol) {
            const double m = g->phi[(int64_t)(N - 1) * g->n + index];
            if (N > 1 ? !(m > 0) : (m > 0)) {
              overlap = 1;
              break;
            }
          }
        }
    if (overlap) continue;
    if (variants >= N) variants = N - 1;
    r = mt_uniform(rng);
    const long phase = (long)(r * variants);
    if (!quiet) printf("phase now = %ld\n", phase);
    fill_cube_solid(g, (int)phase, xlo, xhi, ylo, yhi, zlo, zhi);
    --nparticles;
    if (++count > limit) {
      if (!quiet) printf("Warning: random filling limit = %ld reached. %d particles could not be filled.\n", limit, nparticles);
      break;
    }
  }
  return GP_OK;
}

int gp_fill_domain(const gp_input *in, const char *filling_path, double *phi, double *mu, double *dphi, int quiet,
                   char *err, size_t errlen) {
  const gp_params *p = &in->p;
  const int N = p->nphases, NC = p->ncomp - 1;
  grid_t g = {phi, gp_rows_x(p), gp_rows_y(p), gp_rows_x(p) * gp_rows_y(p), N};
  memset(phi, 0, sizeof(double) * (size_t)N * g.n);
  memset(mu, 0, sizeof(double) * (size_t)NC * g.n);
  if (dphi) memset(dphi, 0, sizeof(double) * (size_t)N * g.n);
  FILE *fr = fopen(filling_path, "rt");
  if (!fr) {
    seterr(err, errlen, "file %s not found", filling_path);
    return GP_E_IO;
  }
  char line[1000], key[128] = "", val[128] = "";
  int nfills = 0;
  rng48_t r48;
  rng48_init(&r48);
  while (fgets(line, sizeof line, fr)) {
    sscanf(line, "%100s = %100[^;];", key, val);
    if (key[0] == '#') continue;
    char tok[16][100];
    int nt;
    if (strcmp(key, "FILLCUBE") == 0) {
      nt = split_tuple(val, tok, 16);
      if (nt < 7) { fclose(fr); seterr(err, errlen, "FILLCUBE needs 7 values"); return GP_E_ARG; }
      int ph = (int)atol(tok[0]);
      long x0 = atol(tok[1]) + 3, x1 = atol(tok[4]) + 3, y0 = atol(tok[2]) + 3, y1 = atol(tok[5]) + 3;
      long z0 = atol(tok[3]), z1 = atol(tok[6]);
      if (ph < 0 || ph >= N) { fclose(fr); seterr(err, errlen, "FILLCUBE phase out of range"); return GP_E_ARG; }
      for (int pass = 0; pass < 2; pass++) {
        int b = pass == 0 ? ph : N - 1;
        for (int64_t x = 0; x < g.rx; x++)
          for (int64_t y = 0; y < g.ry; y++) {
            if (dphi) dphi[(int64_t)b * g.n + x * g.ry + y] = 0.0;
            if (b < N - 1) fill_solid_cell(&g, x, y, b, x >= x0 && x <= x1 && y >= y0 && y <= y1 && 0 >= z0 && 0 <= z1);
            else fill_liquid_cell(&g, x, y, quiet);
          }
      }
      nfills++;
    } else if (strcmp(key, "FILLCYLINDER") == 0 || strcmp(key, "FILLSPHERE") == 0) {
      int cyl = strcmp(key, "FILLCYLINDER") == 0;
      if (cyl && !quiet) printf("Filling cylinder\n");
      nt = split_tuple(val, tok, 16);
      if (nt < (cyl ? 6 : 5)) { fclose(fr); seterr(err, errlen, "%s needs %d values", key, cyl ? 6 : 5); return GP_E_ARG; }
      int ph = (int)atol(tok[0]);
      long xc = atol(tok[1]) + 3, yc = atol(tok[2]) + 3;
      long z0 = atol(tok[3]), z1 = cyl ? atol(tok[4]) : 0;
      double r = atof(tok[cyl ? 5 : 4]);
      if (ph < 0 || ph >= N) { fclose(fr); seterr(err, errlen, "%s phase out of range", key); return GP_E_ARG; }
      /* The reference sweeps the whole grid twice per seed (solid pass, liquid pass). Outside the seed the
       * solid pass is a no-op (solid fractions are only ever 0 or 1 there) and the liquid pass recomputes
       * 1 - sum(solids) from unchanged values, so both passes are restricted to the seed's bounding box here;
       * one full liquid pass after the last fill line restores every other cell. Same state, O(seed area). */
      int64_t bx0 = 0, bx1 = g.rx - 1, by0 = 0, by1 = g.ry - 1;
      if (r >= 0) {
        int64_t rr = (int64_t)ceil(r) + 1;
        bx0 = xc - rr > 0 ? xc - rr : 0;
        bx1 = xc + rr < g.rx - 1 ? xc + rr : g.rx - 1;
        by0 = yc - rr > 0 ? yc - rr : 0;
        by1 = yc + rr < g.ry - 1 ? yc + rr : g.ry - 1;
      }
      for (int pass = 0; pass < 2; pass++) {
        int b = pass == 0 ? ph : N - 1;
        for (int64_t x = bx0; x <= bx1; x++)
          for (int64_t y = by0; y <= by1; y++) {
            if (b < N - 1) {
              long d2 = (x - xc) * (x - xc) + (y - yc) * (y - yc);
              int inside;
              if (cyl) inside = (d2 <= r * r) && (0 >= z0) && (0 <= z1);
              else inside = (d2 + (0 - z0) * (0 - z0) <= r * r);
              fill_solid_cell(&g, x, y, b, inside);
            } else {
              fill_liquid_cell(&g, x, y, quiet);
            }
          }
      }
      nfills++;
      if (cyl && !quiet) printf("End filling cylinder\n");
    } else if (strcmp(key, "FILLELLIPSE") == 0) {
      /* fill_phase_ellipse (GPs/functions/filling.h:54-116): indexes x*MESH_Y+y of the halo-inclusive array
       * and reads its rotation angle before assigning it; taken as 0 here. */
      nt = split_tuple(val, tok, 16);
      if (nt < 7) { fclose(fr); seterr(err, errlen, "FILLELLIPSE needs 7 values"); return GP_E_ARG; }
      int ph = (int)atol(tok[0]);
      long xc = atol(tok[1]) + 3, yc = atol(tok[2]) + 3;
      double l = (double)atol(tok[4]), e = (double)atol(tok[5]);
      if (ph < 0 || ph >= N) { fclose(fr); seterr(err, errlen, "FILLELLIPSE phase out of range"); return GP_E_ARG; }
      int64_t MX = p->mesh_x, MY = p->mesh_y;
      for (int pass = 0; pass < 2; pass++) {
        int b = pass == 0 ? ph : N - 1;
        for (int64_t x = 0; x <= MX - 1; x++)
          for (int64_t y = 0; y <= MY - 1; y++) {
            if (b < N - 1) {
              int inside = (double)(x - xc) * (x - xc) / (l * l) + ((double)(y - yc) * (y - yc)) / (e * e * l * l) <= 1.0;
              int64_t gi = x * MY + y;
              if (gi < MX * MY && gi > 0) {
                int64_t gx = gi / g.ry, gy = gi % g.ry;
                fill_solid_cell(&g, gx, gy, b, inside);
              }
            } else {
              int64_t gi = x * MY + y;
              fill_liquid_cell(&g, gi / g.ry, gi % g.ry, quiet);
            }
          }
      }
      nfills++;
    } else if (strcmp(key, "FILLCYLINDERRANDOM") == 0 || strcmp(key, "FILLSPHERERANDOM") == 0) {
      const int cyl = strcmp(key, "FILLCYLINDERRANDOM") == 0;
      if (!quiet) printf(cyl ? "Filling cylinders at random\n" : "Filling spheres at random\n");
      nt = split_tuple(val, tok, 16);
      if (nt < 5) { fclose(fr); seterr(err, errlen, "%s needs 5 values", key); return GP_E_ARG; }
      const long ph = atol(tok[0]), rad = atol(tok[1]);
      const double vf = atof(tok[2]), spread = atof(tok[4]);
      long shield = atol(tok[3]);
      if (shield > 8) shield = 8;
      else if (shield == 1) shield = 2;
      if (ph < 0 || ph >= N - 1 || rad <= 0) { fclose(fr); seterr(err, errlen, "%s: phase or radius out of range", key); return GP_E_ARG; }
      if (nfills) fill_liquid_all(&g, quiet); /* the earlier lines' liquid passes, completed */
      nfills = 0;
      if (cyl) fill_cylinders_random(&g, &r48, p->mesh_x, p->mesh_y, ph, rad, vf, shield, spread, quiet);
      else fill_spheres_random(&g, &r48, p->mesh_x, p->mesh_y, ph, rad, vf, shield, spread, quiet);
      if (!quiet) printf(cyl ? "End filling cylinders at random\n" : "End filling spheres at random\n");
    } else if (strcmp(key, "FILLVORONOI2D") == 0) {
      if (!quiet) printf("Filling Voronoi 2D\n");
      nt = split_tuple(val, tok, 16);
      if (nt < 6) { fclose(fr); seterr(err, errlen, "FILLVORONOI2D needs 6 values"); return GP_E_ARG; }
      if (nfills) fill_liquid_all(&g, quiet);
      nfills = 0;
      if (N < 2 || fill_voronoi(&g, &r48, atol(tok[0]) + 3, atol(tok[1]) + 3, atol(tok[2]) + 3, atol(tok[3]) + 3,
                                atol(tok[4]), atof(tok[5])) != GP_OK) {
        fclose(fr);
        seterr(err, errlen, "FILLVORONOI2D: region or point count out of range");
        return GP_E_ARG;
      }
      fill_liquid_all(&g, quiet);
      if (!quiet) printf("End filling Voronoi 2D");
    } else if (strcmp(key, "FILLCUBEPATTERN") == 0 || strcmp(key, "FILLCUBERANDOM") == 0) {
      const int pat = strcmp(key, "FILLCUBEPATTERN") == 0;
      if (!quiet) printf(pat ? "Filling cube pattern.\n" : "Filling random cubes.\n");
      nt = split_tuple(val, tok, 16);
      if (nt < (pat ? 8 : 7)) { fclose(fr); seterr(err, errlen, "%s needs %d values", key, pat ? 8 : 7); return GP_E_ARG; }
      const long variants = atol(tok[0]), sx = atol(tok[1]), sy = atol(tok[2]), sz = atol(tok[3]);
      const double sfrac = atof(tok[4]);
      if (sx < 1 || sy < 1 || sz < 1 || variants < 1) { fclose(fr); seterr(err, errlen, "%s: sizes and variants must be >= 1", key); return GP_E_ARG; }
      if (nfills) fill_liquid_all(&g, quiet);
      nfills = 0;
      mt_t mt;
      mt_seed(&mt, fill_seed());
      if (pat) {
        fill_cube_pattern(&g, &mt, p->mesh_x, p->mesh_y, variants, sx, sy, sfrac, atol(tok[5]), atof(tok[6]),
                          atol(tok[7]), quiet);
      } else if (fill_cubes_random(&g, &mt, p->mesh_x, p->mesh_y, variants, sx, sy, sz, sfrac, atof(tok[5]),
                                   atol(tok[6]), quiet) != GP_OK) {
        fclose(fr);
        seterr(err, errlen, "FILLCUBERANDOM: no free place found for the requested volume fraction");
        return GP_E_ARG;
      }
      if (!quiet) printf(pat ? "End filling cube pattern.\n" : "End filling random cubes.\n");
    }
  }
  fclose(fr);
  if (nfills) /* the last line's full liquid pass */
    for (int64_t x = 0; x < g.rx; x++)
      for (int64_t y = 0; y < g.ry; y++) fill_liquid_cell(&g, x, y, 1);
  if (!quiet) printf("Filling composition\n");
  /* fill_composition_cube (GPs/functions/filling.h:221-270): solid cells at ceq[b][b], liquid at
   * cfill[L][L], both through Mu(c, Teq, .) */
  for (int64_t x = 0; x < g.rx; x++)
    for (int64_t y = 0; y < g.ry; y++) {
      int64_t idx = x * g.ry + y;
      double c[NCM];
      int filled = 0;
      for (int b = 0; b < N - 1; b++) {
        if (phi[(int64_t)b * g.n + idx] == 1.0) {
          for (int k = 0; k < NC; k++) c[k] = in->ceq[b][b][k];
          for (int k = 0; k < NC; k++) mu[(int64_t)k * g.n + idx] = host_Mu(p, c, p->Teq, b, k);
          filled = 1;
          break;
        }
      }
      if (!filled) {
        for (int k = 0; k < NC; k++) c[k] = in->cfill[N - 1][N - 1][k];
        for (int k = 0; k < NC; k++) mu[(int64_t)k * g.n + idx] = host_Mu(p, c, p->Teq, N - 1, k);
      }
    }
  return GP_OK;
}

/* apply_boundary_conditions(0) (GPs/solverloop/boundary_serial.h:176-203): X sides over all y, then Y sides
 * over all x, so a corner takes the value of its x- and y-proxy. */
void gp_apply_bc_host(const gp_params *p, double *phi, double *mu) {
  const int64_t rx = gp_rows_x(p), ry = gp_rows_y(p), n = rx * ry;
  const int64_t hx[6] = {0, 1, 2, rx - 3, rx - 2, rx - 1}, hy[6] = {0, 1, 2, ry - 3, ry - 2, ry - 1};
  for (int f = 0; f < 2; f++) {
    double *base = f == 0 ? phi : mu;
    int ncomp = f == 0 ? p->nphases : p->ncomp - 1;
    for (int c = 0; c < ncomp; c++) {
      double *v = base + (int64_t)c * n;
      for (int k = 0; k < 6; k++)
        for (int64_t y = 0; y < ry; y++) v[hx[k] * ry + y] = v[(int64_t)p->halo_src_x[f][k] * ry + y];
      for (int k = 0; k < 6; k++)
        for (int64_t x = 0; x < rx; x++) v[x * ry + hy[k]] = v[x * ry + p->halo_src_y[f][k]];
    }
  }
}

/* ------------------------------------------------------------------------------------------------------------
 * VTK (GPs/solverloop/file_writer.h)
 * ---------------------------------------------------------------------------------------------------------- */
static void put_be(FILE *fp, double v) {
  unsigned char b[8], o[8];
  memcpy(b, &v, 8);
  for (int i = 0; i < 8; i++) o[i] = b[7 - i];
  fwrite(o, 1, 8, fp);
}

int gp_write_vtk(const gp_input *in, const char *path, const double *phi, const double *mu, const double *T_rows) {
  const gp_params *p = &in->p;
  const int N = p->nphases, NC = p->ncomp - 1;
  const int64_t ry = gp_rows_y(p), n = gp_rows_x(p) * ry;
  const int thermal = p->tempgrady != 0;
  if (thermal && !T_rows) return GP_E_ARG;
  FILE *fp = fopen(path, in->ascii ? "w" : "wb");
  if (!fp) return GP_E_IO;
  fprintf(fp, "# vtk DataFile Version 3.0\n");
  fprintf(fp, "Microsim_fields\n");
  fprintf(fp, in->ascii ? "ASCII\n" : "BINARY\n");
  fprintf(fp, "DATASET STRUCTURED_POINTS\n");
  fprintf(fp, "DIMENSIONS %ld %ld %ld\n", (long)p->mesh_y, (long)p->mesh_x, (long)1);
  fprintf(fp, "ORIGIN 0 0 0\n");
  fprintf(fp, "SPACING %le %le %le\n", p->dx, p->dy, in->dz);
  fprintf(fp, "POINT_DATA %ld\n", (long)p->mesh_x * (long)p->mesh_y);
  const int64_t x0 = 3, x1 = p->mesh_x + 2, y0 = 3, y1 = p->mesh_y + 2;
  const int nc = in->writecomposition ? NC : 0;
  for (int f = 0; f < N + NC + nc + thermal; f++) {
    const int temp = f == N + NC + nc;
    if (f < N) fprintf(fp, "SCALARS %s double 1\n", in->phases[f]);
    else if (f < N + NC) fprintf(fp, "SCALARS Mu_%s double 1\n", in->components[f - N]);
    else if (!temp) fprintf(fp, "SCALARS Composition_%s double 1\n", in->components[f - N - NC]);
    else fprintf(fp, "SCALARS Temperature double 1\n");
    fprintf(fp, "LOOKUP_TABLE default\n");
    for (int64_t x = x0; x <= x1; x++)
      for (int64_t y = y0; y <= y1; y++) {
        int64_t idx = x * ry + y;
        double v;
        if (f < N) v = phi[(int64_t)f * n + idx];
        else if (f < N + NC) v = mu[(int64_t)(f - N) * n + idx];
        else if (temp) v = T_rows[y];
        else {
          double pc[GP_MAXP], mc[GP_MAXNC];
          for (int a = 0; a < N; a++) pc[a] = phi[(int64_t)a * n + idx];
          for (int k = 0; k < NC; k++) mc[k] = mu[(int64_t)k * n + idx];
          v = composition_at(p, pc, mc, f - N - NC, thermal ? T_rows[y] : p->T);
        }
        if (in->ascii) fprintf(fp, f < N + NC ? "%le\n" : "%le \n", v);
        else put_be(fp, v);
      }
    /* the Temperature block is the last and ends without an empty line (file_writer.h:142-147) */
    if (in->ascii && !temp) fprintf(fp, "\n");
  }
  fclose(fp);
  return GP_OK;
}

/* fscanf(fp, "%*[^\n]\n"): skip one non-empty line, then any whitespace bytes (file_writer.h:341-353). */
static void skip_line(FILE *fp) {
  int c;
  while ((c = fgetc(fp)) != EOF && c != '\n') {
  }
  while ((c = fgetc(fp)) != EOF && isspace(c)) {
  }
  if (c != EOF) ungetc(c, fp);
}

int gp_read_vtk(const gp_input *in, const char *path, double *phi, double *mu) {
  const gp_params *p = &in->p;
  const int N = p->nphases, NC = p->ncomp - 1;
  const int64_t ry = gp_rows_y(p), n = gp_rows_x(p) * ry;
  FILE *fp = fopen(path, in->ascii ? "r" : "rb");
  if (!fp) return GP_E_IO;
  for (int i = 0; i < 8; i++) skip_line(fp);
  for (int f = 0; f < N + NC; f++) {
    skip_line(fp);
    skip_line(fp);
    double *dst = f < N ? phi + (int64_t)f * n : mu + (int64_t)(f - N) * n;
    for (int64_t x = 3; x <= p->mesh_x + 2; x++)
      for (int64_t y = 3; y <= p->mesh_y + 2; y++) {
        double v = 0.0;
        if (in->ascii) {
          if (fscanf(fp, "%le \n", &v) != 1) { fclose(fp); return GP_E_IO; }
        } else {
          unsigned char b[8], o[8];
          if (fread(b, 1, 8, fp) != 8) { fclose(fp); return GP_E_IO; }
          for (int i = 0; i < 8; i++) o[i] = b[7 - i];
          memcpy(&v, o, 8);
        }
        dst[x * ry + y] = v;
      }
  }
  fclose(fp);
  return GP_OK;
}

/* ------------------------------------------------------------------------------------------------------------
 * Echo files (PRINT_* in GPs/functions/utility_functions.h:777-922)
 * ---------------------------------------------------------------------------------------------------------- */
static void pr_matrix(FILE *fp, const char *key0, const double *m, int rows, int cols, int stride) {
  char key[1100];
  snprintf(key, sizeof key, "%s", key0);
  fprintf(fp, "%s = [", key);
  strcat(key, "= [ ");
  for (int i = 0; i < rows; i++) {
    for (int j = 0; j < cols; j++) fprintf(fp, j < cols - 1 ? "%le," : "%le", m[i * stride + j]);
    if (i < rows - 1) fprintf(fp, "\n%*s", (int)strlen(key), "");
    else fprintf(fp, "]\n");
  }
  fprintf(fp, "\n");
}

static void pr_vector(FILE *fp, const char *key, const double *v, int m) {
  fprintf(fp, "%s = [", key);
  for (int i = 0; i < m; i++) fprintf(fp, i < m - 1 ? "%le," : "%le]\n", v[i]);
  fprintf(fp, "\n");
}

static int stem_of(const char *input_path, char *stem, size_t n) {
  char buf[1000];
  snprintf(buf, sizeof buf, "%s", input_path);
  char *save = NULL, *t = strtok_r(buf, ".", &save);
  if (!t) return GP_E_ARG;
  snprintf(stem, n, "%s", t);
  return GP_OK;
}

int gp_write_echo(const gp_input *in, const char *input_path) {
  const gp_params *p = &in->p;
  const int N = p->nphases, K = p->ncomp, NC = K - 1;
  char stem[1000], fn[1100], key[1000];
  if (stem_of(input_path, stem, sizeof stem)) return GP_E_ARG;
  snprintf(fn, sizeof fn, "%s.out", stem);
  FILE *fp = fopen(fn, "w");
  if (!fp) return GP_E_IO;
#define PL(k, v) fprintf(fp, "%s = %ld\n\n", k, (long)(v))
#define PD(k, v) fprintf(fp, "%s = %le\n\n", k, (double)(v))
#define PI(k, v) fprintf(fp, "%s = %d\n\n", k, (int)(v))
  PL("MESH_X", p->mesh_x); PL("MESH_Y", p->mesh_y); PL("MESH_Z", in->mesh_z);
  PD("DELTA_X", p->dx); PD("DELTA_Y", p->dy); PD("DELTA_Z", in->dz);
  PI("NUMPHASES", N); PI("NUMCOMPONENTS", K); PI("DIMENSION", in->dimension);
  PL("NTIMESTEPS", in->ntimesteps); PL("SAVET", in->savet); PL("NSMOOTH", in->nsmooth);
  PL("STARTTIME", in->starttime); PL("RESTART", in->restart);
  PD("R", in->R); PD("V", in->V); PD("TEMPERATURE_SCALE", in->temperature_scale); PD("TEMPERATURE", p->T);
  PD("Equilibrium_temperature", p->Teq); PD("Filling_temperature", in->Tfill);
  pr_matrix(fp, "GAMMA", p->gamma, N, N, P);
  pr_matrix(fp, "dab", p->dab, N, N, P);
  pr_matrix(fp, "Tau", p->tau_ab, N, N, P);
  for (int i = 0; i < N; i++) {
    snprintf(key, sizeof key, "DIFFUSIVITY[%s]", in->phases[i]);
    pr_matrix(fp, key, &p->diffusivity[i * NCM * NCM], NC, NC, NCM);
  }
  for (int i = 0; i < N; i++) { snprintf(key, sizeof key, "Slopes[Solidus,%s]", in->phases[i]); pr_vector(fp, key, in->slopes[i][i], NC); }
  for (int i = 0; i < N; i++) { snprintf(key, sizeof key, "Slopes[Liquidus,%s]", in->phases[i]); pr_vector(fp, key, in->slopes[i][N - 1], NC); }
  for (int i = 0; i < N; i++) {
    snprintf(key, sizeof key, "A[%s]", in->phases[i]);
    pr_matrix(fp, key, &p->A[i * NCM * NCM], NC, NC, NCM);
  }
  for (int i = 0; i < N; i++)
    for (int j = 0; j < N; j++)
      if (i != j) {
        snprintf(key, sizeof key, "Rotation_matrix[%s][%s]", in->phases[i], in->phases[j]);
        pr_matrix(fp, key, &in->rot3[i][j][0][0], 3, 3, 3);
      }
  for (int i = 0; i < N; i++)
    for (int j = 0; j < N; j++)
      if (i != j) {
        snprintf(key, sizeof key, "Inv_Rotation_matrix[%s][%s]", in->phases[i], in->phases[j]);
        pr_matrix(fp, key, &in->rot3_inv[i][j][0][0], 3, 3, 3);
      }
  PI("Function_F", in->fn_f); PI("Function_W", p->fn_w); PI("Function_anisotropy", p->fn_aniso);
  PI("Anisotropy_type", p->fold);
  fprintf(fp, "COMPONENTS = [");
  for (int i = 0; i < K; i++) fprintf(fp, i < K - 1 ? "%s," : "%s]\n", in->components[i]);
  fprintf(fp, "\n");
  fprintf(fp, "PHASES = [");
  for (int i = 0; i < K; i++) fprintf(fp, i < K - 1 ? "%s," : "%s]\n", in->phases[i]);
  fprintf(fp, "\n");
  for (int i = 0; i < N; i++) { snprintf(key, sizeof key, "ceq[Solidus,%s]", in->phases[i]); pr_vector(fp, key, in->ceq[i][i], NC); }
  for (int i = 0; i < N; i++) { snprintf(key, sizeof key, "ceq[Liquidus,%s]", in->phases[i]); pr_vector(fp, key, in->ceq[i][N - 1], NC); }
  for (int i = 0; i < N; i++) { snprintf(key, sizeof key, "cfill[Liquidus,%s]", in->phases[i]); pr_vector(fp, key, in->cfill[i][N - 1], NC); }
  for (int i = 0; i < N; i++) {
    const char *c[6] = {"xx", "yy", "zz", "yz", "xz", "xy"};
    for (int s = 0; s < 6; s++) fprintf(fp, "EIGEN_STRAIN[%s].%s = %le\n", in->phases[i], c[s], 0.0);
    fprintf(fp, "\n");
  }
  for (int i = 0; i < N; i++) {
    const char *c[3] = {"C11", "C12", "C44"};
    for (int s = 0; s < 3; s++) fprintf(fp, "Stiffness_cubic[%s].%s = %le\n", in->phases[i], c[s], 0.0);
    fprintf(fp, "\n");
  }
#undef PL
#undef PD
#undef PI
  fclose(fp);
  snprintf(fn, sizeof fn, "%s.bd", stem);
  fp = fopen(fn, "w");
  if (!fp) return GP_E_IO;
  const char *scal[3] = {"PHI", "MU", "T"};
  const char *side[6] = {"LEFT", "RIGHT", "FRONT", "BACK", "TOP", "BOTTOM"};
  const char *msg = "DIRICHLET and FREE boundary conditions have not yet been implemented. Please wait for next "
                    "release. Will default to PERIODIC\n";
  for (int j = 0; j < 3; j++)
    for (int s = 0; s < 6; s++) {
      int t = in->bc_type[s][j];
      snprintf(key, sizeof key, "BOUNDARY_%s[%s]", side[s], scal[j]);
      if (t == 0) fprintf(fp, "%s = %s\n\n%s", key, "FREE", msg);
      if (t == 1) fprintf(fp, "%s = %s\n\n", key, "NEUMANN");
      if (t == 2) fprintf(fp, "%s = %s\n\n%s", key, "DIRICHLET", msg);
      if (t == 3) fprintf(fp, "%s = %s\n\n", key, "PERIODIC");
    }
  fclose(fp);
  return GP_OK;
}

size_t gp_sizeof_input(void) { return sizeof(gp_input); }
size_t gp_sizeof_params(void) { return sizeof(gp_params); }
size_t gp_sizeof_stats(void) { return sizeof(gp_stats); }
