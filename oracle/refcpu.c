/* refcpu.c — FP64 CPU restatement of the reference hot path (THE ORACLE).
 *
 * Test infrastructure only: loaded by tests/, the smoke check and bench.py's
 * cpu_baseline leg as the checker.  The product (fdtd_amd/) never links it.
 *
 * Every update keeps the reference's state (D/B fluxes, E/H derived by the
 * material step), its sweep order and its floating-point operation order, so
 * that results agree with the reference's own templates to the last bits
 * (compiled with -ffp-contract=off).  File:line citations point into
 * pederson/fdtd.
 */
#include "refcpu.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* include/FDTDConstants.hpp:394-398 */
static const double RC_PI = 3.14159265358979323846264338327950288;
static const double RC_EPS0 = 8.854e-12;
static const double RC_C0 = 2.99792458e+8;
static double rc_mu0_v(void) { return RC_PI * 4.0e-7; }
/* constexpr Newton-Raphson square root, FDTDConstants.hpp:136-156 */
static double rc_newton_sqrt(double x) {
  double curr = x, prev = 0;
  while (curr != prev) {
    double nxt = 0.5 * (curr + x / curr);
    prev = curr;
    curr = nxt;
  }
  return curr;
}
double rc_eps0(void) { return RC_EPS0; }
double rc_mu0(void) { return rc_mu0_v(); }
double rc_c0(void) { return RC_C0; }
double rc_imp0(void) { return rc_newton_sqrt(rc_mu0_v() / RC_EPS0); }

/* Levi-Civita for distinct axes (FDTDConstants.hpp:602-637) */
static int lc(int I, int J, int K) {
  if ((I == 0 && J == 1 && K == 2) || (I == 1 && J == 2 && K == 0) || (I == 2 && J == 0 && K == 1)) return 1;
  return -1;
}
/* d1 = next(I), d2 = the remaining axis (YeeUpdates.hpp:397-398) */
static const int D1[3] = {1, 2, 0};
static const int D2[3] = {2, 0, 1};
static int third(int I, int J) { return 3 - I - J; }

/* ---------------------------------------------------------------- PML */
/* NewPML::setPMLParameters, include/PML.hpp:85-102 */
void rc_pml_coeffs(double K, double S, double A, double dt, double out[4]) {
  double nu = 1.0 / RC_EPS0 * (S / K + A);
  double B = exp(-dt * nu);
  double C = (nu == 0) ? 0.0 : S / K * 1.0 / (S + K * A) * (B - 1.0);
  double u = -S / (RC_EPS0 * K * K);
  double F = (nu == 0) ? 1.0 / K : 1.0 / K - 1.0 / dt * (1.0 - B - nu * dt) * u / (nu * nu);
  double G = (nu == 0) ? 0.0 : 1.0 / dt * (1.0 - B) / nu;
  out[0] = B;
  out[1] = C;
  out[2] = F;
  out[3] = G;
}

/* PMLParameterModel::K/S/A, include/PML.hpp:378-391 */
void rc_pml_model(double x, double m, double ma, double smax, double kmax, double amax, double out[3]) {
  if (x < 0.0) {
    out[0] = 1.0;
    out[1] = 0.0;
    out[2] = 0.0;
    return;
  }
  out[0] = (x > 1.0 ? 1.0 : 1.0 + pow(x, m) * (kmax - 1.0));
  out[1] = (x > 1.0 ? 0.0 : pow(x, m) * smax);
  out[2] = (x > 1.0 ? 0.0 : pow(1.0 - x, ma) * amax);
}

static void set_row(double* row, double K, double S, double A, double dt) {
  double bcfg[4];
  rc_pml_coeffs(K, S, A, dt, bcfg);
  row[0] = K;
  row[1] = S;
  row[2] = A;
  row[3] = bcfg[0];
  row[4] = bcfg[1];
  row[5] = bcfg[2];
  row[6] = bcfg[3];
}

/* placement idiom: tests/testTM.cpp:46-71, include/Source.hpp:380-389 */
void rc_pml_profile(int L, int n, double m, double ma, double smax, double kmax, double amax, double dt,
                    double* outE, double* outH) {
  for (int i = 0; i < L; i++) {
    double d[7] = {1, 0, 0, 0, 0, 1, 0}; /* NewPML() defaults, PML.hpp:40-55 */
    memcpy(outE + 7 * i, d, sizeof d);
    memcpy(outH + 7 * i, d, sizeof d);
  }
  for (int i = 1; i < n + 2; i++) {
    double x = (double)(n - i + 1) / (double)n;
    double xm = x - 0.5 / (double)n;
    double e[3], h[3];
    rc_pml_model(x, m, ma, smax, kmax, amax, e);
    rc_pml_model(xm, m, ma, smax, kmax, amax, h);
    set_row(outE + 7 * i, e[0], e[1], e[2], dt);
    set_row(outH + 7 * i, h[0], h[1], h[2], dt);
    set_row(outE + 7 * (L - i), e[0], e[1], e[2], dt);
    set_row(outH + 7 * (L - 1 - i), h[0], h[1], h[2], dt);
  }
}

/* ---------------------------------------------------------------- Aux1D */
struct rc_aux {
  int L;
  double dx, dt, t;
  double *Dz, *Ez, *By, *Hy, *psiE, *psiH;
  double *pE, *pH; /* [L][7] */
};

rc_aux* rc_aux_create(int ncells, double dx, double dt) {
  rc_aux* a = (rc_aux*)calloc(1, sizeof(rc_aux));
  const int npml = 15; /* Source.hpp:362 */
  a->L = ncells + 2 * npml + 20; /* Source.hpp:404 */
  a->dx = dx;
  a->dt = dt;
  int L = a->L;
  a->Dz = calloc(L, sizeof(double));
  a->Ez = calloc(L, sizeof(double));
  a->By = calloc(L, sizeof(double));
  a->Hy = calloc(L, sizeof(double));
  a->psiE = calloc(L, sizeof(double));
  a->psiH = calloc(L, sizeof(double));
  a->pE = malloc(sizeof(double) * 7 * L);
  a->pH = malloc(sizeof(double) * 7 * L);
  /* PMLParameterModel(dx) defaults, PML.hpp:372-373 */
  double m = 3, smax = 0.8 * (m + 1) / (rc_imp0() * dx);
  rc_pml_profile(L, npml, m, 1, smax, 5, 0.05, dt, a->pE, a->pH);
  return a;
}

void rc_aux_destroy(rc_aux* a) {
  if (!a) return;
  free(a->Dz);
  free(a->Ez);
  free(a->By);
  free(a->Hy);
  free(a->psiE);
  free(a->psiH);
  free(a->pE);
  free(a->pH);
  free(a);
}

int rc_aux_length(const rc_aux* a) { return a->L; }

/* Aux1D::time_step, Source.hpp:413-435.  Ghost neighbours outside [0,L) hold zero. */
void rc_aux_step(rc_aux* a, double source_value) {
  const int L = a->L;
  const double dt = a->dt, dx = a->dx;
  const int npml = 15;
  a->t += dt;
  /* YeeUpdate<TEM,Dz>: Dz += dt*(F^E_x/dx*(Hy - Hy[-1]) + 0) */
  for (int n = 0; n < L; n++) {
    double hm = n > 0 ? a->Hy[n - 1] : 0.0;
    double ax = 1 * a->pE[7 * n + 5] / (dx) * (a->Hy[n] - hm);
    a->Dz[n] = a->Dz[n] + 1.0 * 1.0 * dt * (ax + 0.0);
  }
  /* UpdatePML<TEM,Electric,X> on [0,npml+2) and [L-(npml+3),L) */
  for (int pass = 0; pass < 2; pass++) {
    int b = pass == 0 ? 0 : L - (npml + 3), e = pass == 0 ? npml + 2 : L;
    for (int n = b; n < e; n++) {
      double hm = n > 0 ? a->Hy[n - 1] : 0.0;
      a->Dz[n] += 1.0 * dt * a->psiE[n] * a->pE[7 * n + 6];
      a->psiE[n] = a->pE[7 * n + 3] * a->psiE[n] + a->pE[7 * n + 4] * 1 / dx * (a->Hy[n] - hm);
    }
  }
  for (int n = 0; n < L; n++) a->Ez[n] = a->Dz[n] / (1.0 * RC_EPS0);
  a->Ez[npml + 5] += source_value; /* Source.hpp:425 */
  /* YeeUpdate<TEM,By>: By += -dt*(0 + (-F^H_x/dx)*(Ez[+1]-Ez)) */
  for (int n = 0; n < L; n++) {
    double ep = n + 1 < L ? a->Ez[n + 1] : 0.0;
    double ax = -1 * a->pH[7 * n + 5] / (dx) * (ep - a->Ez[n]);
    a->By[n] = a->By[n] + -1.0 * 1.0 * dt * (0.0 + ax);
  }
  for (int pass = 0; pass < 2; pass++) {
    int b = pass == 0 ? 0 : L - (npml + 3), e = pass == 0 ? npml + 2 : L;
    for (int n = b; n < e; n++) {
      double ep = n + 1 < L ? a->Ez[n + 1] : 0.0;
      a->By[n] += -1.0 * dt * a->psiH[n] * a->pH[7 * n + 6];
      a->psiH[n] = a->pH[7 * n + 3] * a->psiH[n] + a->pH[7 * n + 4] * -1 / dx * (ep - a->Ez[n]);
    }
  }
  const double mu0 = rc_mu0_v();
  for (int n = 0; n < L; n++) a->Hy[n] = a->By[n] / (1.0 * mu0);
}

/* Aux1D::interpolateE / interpolateH, Source.hpp:442-458 */
double rc_aux_interp_E(const rc_aux* a, double loc) {
  double nd = (double)a->L / 2.0 + loc / a->dx;
  size_t n = (size_t)nd;
  double frac = (nd - n);
  return (1.0 - frac) * a->Ez[n] + frac * a->Ez[n + 1];
}
double rc_aux_interp_H(const rc_aux* a, double loc) {
  double nd = (double)a->L / 2.0 + loc / a->dx - 0.5;
  size_t n = (size_t)nd;
  double frac = (nd - n);
  return (1.0 - frac) * a->Hy[n] + frac * a->Hy[n + 1];
}
void rc_aux_state(const rc_aux* a, double* Ez, double* Hy) {
  memcpy(Ez, a->Ez, sizeof(double) * a->L);
  memcpy(Hy, a->Hy, sizeof(double) * a->L);
}

/* ------------------------------------------------------------ grid sim */
typedef struct {
  double eps, mu;
  int np;
  int kind[RC_MAXPOLES];
  double freq[RC_MAXPOLES], gamma[RC_MAXPOLES], weight[RC_MAXPOLES];
  int mag;        /* magnetized Drude: 0 none, 1 fixed bias, 2 the cell's own B */
  double bias[3]; /* [T] */
  int fluid;      /* fluid Drude: omega_p, gamma from the cell's ne, nu_m */
  /* magnetic side: Drude (1), Lorentz (2), conductive (3) poles on B, H, M, K */
  int mnp;
  int mkind[RC_MAXPOLES];
  double mfreq[RC_MAXPOLES], mgamma[RC_MAXPOLES], mweight[RC_MAXPOLES];
} rc_mat;

typedef struct {
  int field, comp;
  size_t cell;
  double* table;
  int n;
} rc_src;

typedef struct {
  int region[6];
  int mask, nfreq, ncomp, nphase;
  size_t ncells;
  double *phE, *phH, *acc;
} rc_dft;

typedef struct {
  int field, comp;
  size_t cell;
} rc_probe;

/* one single TFSF plane with its own incident line */
typedef struct {
  int axis;
  double nrm; /* outward normal along axis: -1 MIN, +1 MAX */
  int lo[3], hi[3]; /* cells, inclusive */
  double kvec[3], pol[3], polH[3], origin[3];
  rc_aux* aux;
  double* signal;
  int nsignal;
} rc_plane;

struct rc_sim {
  int mode, nx, ny, nz;
  size_t ncell;
  double dx, dt;
  int actE[3], actH[3], axes[3];
  double *D[3], *E[3], *B[3], *H[3];
  double* psi[2][3][3];
  double* pml[2][3]; /* [ft][axis] -> L*4 = B,C,F,G */
  uint8_t* matid;
  rc_mat mat[RC_MAXMAT];
  double *P[RC_MAXPOLES][3], *J[RC_MAXPOLES][3];
  int npoles_alloc;
  double *M[RC_MAXPOLES][3], *K[RC_MAXPOLES][3]; /* magnetization and magnetic current per pole */
  int mpoles_alloc;
  rc_src* src;
  int nsrc;
  int tfsf;
  int box[6];
  double kvec[3], pol[3], polH[3], origin[3];
  rc_aux* aux;
  double* signal;
  int nsignal;
  rc_plane plane[RC_MAXPLANE];
  int nplane;
  rc_dft* dft;
  int ndft;
  rc_probe* probes;
  int nprobes;
  double* ptrace;
  int ptrace_cap;
  long step;
  /* design box: per-cell eps_r (build-defined) */
  int has_design;
  int dsg[6];
  double* dsg_eps;
  /* adjoint state: D, B and psi adjoints carried across steps; E and H
   * adjoints are per-step accumulators of the derived fields */
  double *Db[3], *Bb[3], *Eb[3], *Hb[3];
  double* Pb[2][3][3];
  double *PPb[RC_MAXPOLES][3], *JPb[RC_MAXPOLES][3]; /* pole P and J adjoints (dispersive cells) */
  int obj_dft, obj_mask;
  /* boundary condition per [axis][side] (RC_BC_*; 0: ghost layer left alone) */
  int bc[3][2];
  double *fluid_ne, *fluid_nu; /* per cell, NULL until rc_set_fluid_state */
  /* complex fields (a Bloch-periodic face): the imaginary part, a copy of the
   * configuration without sources or TFSF, made at the first step */
  double bloch[3][2][2];
  struct rc_sim* im;
};

static size_t cidx(const rc_sim* s, int i, int j, int k) {
  return (size_t)i + (size_t)s->nx * ((size_t)j + (size_t)s->ny * (size_t)k);
}

/* index of (i,j,k) in the design box (x fastest), -1 outside */
static long design_local(const rc_sim* s, int i, int j, int k) {
  if (!s->has_design) return -1;
  const int* b = s->dsg;
  if (i < b[0] || i >= b[1] || j < b[2] || j >= b[3] || k < b[4] || k >= b[5]) return -1;
  return (long)(i - b[0]) + (long)(b[1] - b[0]) * ((long)(j - b[2]) + (long)(b[3] - b[2]) * (long)(k - b[4]));
}

rc_sim* rc_create(int mode, int nx, int ny, int nz, double dx, double dt) {
  rc_sim* s = (rc_sim*)calloc(1, sizeof(rc_sim));
  s->mode = mode;
  s->nx = nx;
  s->ny = mode == RC_TEM ? 1 : ny;
  s->nz = mode == RC_3D ? nz : 1;
  s->ncell = (size_t)s->nx * s->ny * s->nz;
  s->dx = dx;
  s->dt = dt;
  int aE[3] = {0, 0, 0}, aH[3] = {0, 0, 0}, ax[3] = {1, 0, 0};
  if (mode == RC_3D) {
    aE[0] = aE[1] = aE[2] = aH[0] = aH[1] = aH[2] = 1;
    ax[1] = ax[2] = 1;
  } else if (mode == RC_TM) {
    aE[2] = aH[0] = aH[1] = 1;
    ax[1] = 1;
  } else if (mode == RC_TE) {
    aE[0] = aE[1] = aH[2] = 1;
    ax[1] = 1;
  } else {
    aE[2] = aH[1] = 1;
  }
  for (int c = 0; c < 3; c++) {
    s->actE[c] = aE[c];
    s->actH[c] = aH[c];
    s->axes[c] = ax[c];
    if (aE[c]) {
      s->D[c] = calloc(s->ncell, sizeof(double));
      s->E[c] = calloc(s->ncell, sizeof(double));
    }
    if (aH[c]) {
      s->B[c] = calloc(s->ncell, sizeof(double));
      s->H[c] = calloc(s->ncell, sizeof(double));
    }
  }
  for (int ft = 0; ft < 2; ft++)
    for (int I = 0; I < 3; I++)
      for (int J = 0; J < 3; J++)
        if ((ft == 0 ? aE[I] : aH[I]) && I != J && ax[J]) s->psi[ft][I][J] = calloc(s->ncell, sizeof(double));
  int len[3] = {s->nx, s->ny, s->nz};
  for (int ft = 0; ft < 2; ft++)
    for (int a = 0; a < 3; a++) {
      s->pml[ft][a] = malloc(sizeof(double) * 4 * len[a]);
      for (int i = 0; i < len[a]; i++) {
        s->pml[ft][a][4 * i + 0] = 0.0; /* NewPML() defaults B=0,C=0,F=1,G=0 */
        s->pml[ft][a][4 * i + 1] = 0.0;
        s->pml[ft][a][4 * i + 2] = 1.0;
        s->pml[ft][a][4 * i + 3] = 0.0;
      }
    }
  s->matid = calloc(s->ncell, 1);
  for (int m = 0; m < RC_MAXMAT; m++) {
    s->mat[m].eps = 1.0;
    s->mat[m].mu = 1.0;
  }
  return s;
}

void rc_destroy(rc_sim* s) {
  if (!s) return;
  for (int c = 0; c < 3; c++) {
    free(s->D[c]);
    free(s->E[c]);
    free(s->B[c]);
    free(s->H[c]);
    for (int p = 0; p < RC_MAXPOLES; p++) {
      free(s->P[p][c]);
      free(s->J[p][c]);
      free(s->M[p][c]);
      free(s->K[p][c]);
    }
  }
  for (int ft = 0; ft < 2; ft++)
    for (int I = 0; I < 3; I++) {
      s->pml[ft][I] ? free(s->pml[ft][I]) : (void)0;
      for (int J = 0; J < 3; J++) free(s->psi[ft][I][J]);
    }
  free(s->matid);
  for (int q = 0; q < s->nsrc; q++) free(s->src[q].table);
  free(s->src);
  rc_aux_destroy(s->aux);
  free(s->signal);
  for (int p = 0; p < s->nplane; p++) {
    rc_aux_destroy(s->plane[p].aux);
    free(s->plane[p].signal);
  }
  for (int q = 0; q < s->ndft; q++) {
    free(s->dft[q].phE);
    free(s->dft[q].phH);
    free(s->dft[q].acc);
  }
  free(s->dft);
  free(s->probes);
  free(s->ptrace);
  free(s->dsg_eps);
  for (int c = 0; c < 3; c++) {
    free(s->Db[c]);
    free(s->Bb[c]);
    free(s->Eb[c]);
    free(s->Hb[c]);
  }
  for (int ft = 0; ft < 2; ft++)
    for (int I = 0; I < 3; I++)
      for (int J = 0; J < 3; J++) free(s->Pb[ft][I][J]);
  for (int p = 0; p < RC_MAXPOLES; p++)
    for (int c = 0; c < 3; c++) {
      free(s->PPb[p][c]);
      free(s->JPb[p][c]);
    }
  free(s->fluid_ne);
  free(s->fluid_nu);
  rc_destroy(s->im);
  free(s);
}

int rc_set_pml_axis(rc_sim* s, int axis, int ft, const double* coeffs) {
  int len[3] = {s->nx, s->ny, s->nz};
  if (axis < 0 || axis > 2 || ft < 0 || ft > 1) return -1;
  memcpy(s->pml[ft][axis], coeffs, sizeof(double) * 4 * len[axis]);
  return 0;
}

int rc_set_material(rc_sim* s, int id, double eps_r, double mu_r, int npoles, const double* poles) {
  if (id < 0 || id >= RC_MAXMAT || npoles < 0 || npoles > RC_MAXPOLES) return -1;
  rc_mat* m = &s->mat[id];
  m->eps = eps_r;
  m->mu = mu_r;
  m->np = npoles;
  m->mag = 0;
  m->fluid = 0;
  for (int p = 0; p < npoles; p++) {
    m->kind[p] = (int)poles[4 * p];
    m->freq[p] = poles[4 * p + 1];
    m->gamma[p] = poles[4 * p + 2];
    m->weight[p] = poles[4 * p + 3];
  }
  if (npoles > s->npoles_alloc) {
    for (int p = s->npoles_alloc; p < npoles; p++)
      for (int c = 0; c < 3; c++)
        if (s->actE[c]) {
          s->P[p][c] = calloc(s->ncell, sizeof(double));
          s->J[p][c] = calloc(s->ncell, sizeof(double));
        }
    s->npoles_alloc = npoles;
  }
  if (s->im) rc_set_material(s->im, id, eps_r, mu_r, npoles, poles);
  return 0;
}

/* the magnetic side of material id: the reference's pole functors on
 * FieldType::Magnetic (DispersiveMaterials.hpp, the *Call structs with
 * val = mu0), which step B, H, M, K as the electric side steps D, E, P, J */
int rc_set_material_magnetic(rc_sim* s, int id, int npoles, const double* poles) {
  if (id < 0 || id >= RC_MAXMAT || npoles < 0 || npoles > RC_MAXPOLES) return -1;
  rc_mat* m = &s->mat[id];
  m->mnp = npoles;
  for (int p = 0; p < npoles; p++) {
    m->mkind[p] = (int)poles[4 * p];
    m->mfreq[p] = poles[4 * p + 1];
    m->mgamma[p] = poles[4 * p + 2];
    m->mweight[p] = poles[4 * p + 3];
  }
  if (npoles > s->mpoles_alloc) {
    for (int p = s->mpoles_alloc; p < npoles; p++)
      for (int c = 0; c < 3; c++)
        if (s->actH[c]) {
          s->M[p][c] = calloc(s->ncell, sizeof(double));
          s->K[p][c] = calloc(s->ncell, sizeof(double));
        }
    s->mpoles_alloc = npoles;
  }
  if (s->im) rc_set_material_magnetic(s->im, id, npoles, poles);
  return 0;
}

int rc_set_magnetized_drude(rc_sim* s, int id, double eps_r, double mu_r, double freq, double gamma,
                            const double* bias) {
  const double pole[4] = {1.0, freq, gamma, 1.0};
  if (rc_set_material(s, id, eps_r, mu_r, 1, pole)) return -1;
  s->mat[id].mag = bias ? 1 : 2;
  for (int c = 0; c < 3; c++) s->mat[id].bias[c] = bias ? bias[c] : 0.0;
  if (s->im) s->im->mat[id] = s->mat[id];
  return 0;
}

int rc_set_fluid_drude(rc_sim* s, int id, double mu_r, int magnetization, const double* bias) {
  const double pole[4] = {1.0, 0.0, 0.0, 1.0};
  if (magnetization < 0 || magnetization > 2 || (magnetization == 1 && !bias)) return -1;
  if (rc_set_material(s, id, 1.0, mu_r, 1, pole)) return -1;
  s->mat[id].fluid = 1;
  s->mat[id].mag = magnetization;
  for (int c = 0; c < 3; c++) s->mat[id].bias[c] = magnetization == 1 ? bias[c] : 0.0;
  if (s->im) s->im->mat[id] = s->mat[id];
  return 0;
}

int rc_set_fluid_state(rc_sim* s, const double* ne, const double* nu_m) {
  if (!s->fluid_ne) {
    s->fluid_ne = malloc(s->ncell * sizeof(double));
    s->fluid_nu = malloc(s->ncell * sizeof(double));
  }
  memcpy(s->fluid_ne, ne, s->ncell * sizeof(double));
  memcpy(s->fluid_nu, nu_m, s->ncell * sizeof(double));
  if (s->im) rc_set_fluid_state(s->im, ne, nu_m);
  return 0;
}

/* omega_p and gamma of material m at cell c: the fluid storages' getFreq =
 * Kval sqrt(ne), Kval = sqrt(q_e^2/(m_e eps0)) by the reference's constexpr
 * Newton iteration, and getGamma = nu_m (DispersiveMaterials.hpp:1160-1170) */
static void plasma_params(const rc_sim* s, const rc_mat* m, size_t c, double* freq, double* gamma) {
  if (m->fluid) {
    const double kval = rc_newton_sqrt(1.6e-19 * 1.6e-19 / (9.11e-31 * RC_EPS0));
    *freq = kval * sqrt(s->fluid_ne[c]);
    *gamma = s->fluid_nu[c];
  } else {
    *freq = m->freq[0];
    *gamma = m->gamma[0];
  }
}

/* MagnetizedDrudeUpdate3D::implicit_update (DispersiveMaterials.hpp:1283-1340),
 * in its own operation order; B is the bias (MagnetizedDrudeStorage,
 * MagnetizedFluidDrudeStorage) or the cell's own B (the self-magnetized
 * storages, 1816-1943). */
static void magnetized_drude_3d(rc_sim* s, const rc_mat* m, size_t c, double dt) {
  const double qe = 1.6e-19, me = 9.11e-31; /* drude::q_e, drude::m_e (899-900) */
  double freq, gamma;
  plasma_params(s, m, c, &freq, &gamma);
  const double Bx = m->mag == 1 ? m->bias[0] : s->B[0][c];
  const double By = m->mag == 1 ? m->bias[1] : s->B[1][c];
  const double Bz = m->mag == 1 ? m->bias[2] : s->B[2][c];
  double *Dx = &s->D[0][c], *Dy = &s->D[1][c], *Dz = &s->D[2][c];
  double *Px = &s->P[0][0][c], *Py = &s->P[0][1][c], *Pz = &s->P[0][2][c];
  double *Jx = &s->J[0][0][c], *Jy = &s->J[0][1][c], *Jz = &s->J[0][2][c];
  const double wp = dt * freq;
  const double vc = dt * gamma;
  const double vx = dt * Bx * qe / me;
  const double vy = dt * By * qe / me;
  const double vz = dt * Bz * qe / me;
  const double oneplus = (1.0 + vc + wp * wp);
  const double denom = oneplus * (oneplus * oneplus + (vx * vx + vy * vy + vz * vz));
  const double a11 = oneplus * oneplus + vx * vx, a21 = vz * oneplus + vx * vy, a31 = -vy * oneplus + vx * vz;
  const double a12 = -vz * oneplus + vx * vy, a22 = oneplus * oneplus + vy * vy, a32 = vx * oneplus + vy * vz;
  const double a13 = vy * oneplus + vx * vz, a23 = -vx * oneplus + vy * vz, a33 = oneplus * oneplus + vz * vz;
  /* J + (wp^2/dt)(D - P), written as the reference groups it */
  const double rx = *Jx - wp * wp / dt * *Px + wp * wp / dt * *Dx;
  const double ry = *Jy - wp * wp / dt * *Py + wp * wp / dt * *Dy;
  const double rz = *Jz - wp * wp / dt * *Pz + wp * wp / dt * *Dz;
  const double sx = dt * *Jx + wp * wp * *Dx - wp * wp * *Px;
  const double sy = dt * *Jy + wp * wp * *Dy - wp * wp * *Py;
  const double sz = dt * *Jz + wp * wp * *Dz - wp * wp * *Pz;
  const double Jxh = 1.0 / denom * (a11 * rx + a12 * ry + a13 * rz);
  const double Jyh = 1.0 / denom * (a21 * rx + a22 * ry + a23 * rz);
  const double Jzh = 1.0 / denom * (a31 * rx + a32 * ry + a33 * rz);
  const double Pxh = *Px + 1.0 / denom * (a11 * sx + a12 * sy + a13 * sz);
  const double Pyh = *Py + 1.0 / denom * (a21 * sx + a22 * sy + a23 * sz);
  const double Pzh = *Pz + 1.0 / denom * (a31 * sx + a32 * sy + a33 * sz);
  *Jx = Jxh;
  *Jy = Jyh;
  *Jz = Jzh;
  *Px = Pxh;
  *Py = Pyh;
  *Pz = Pzh;
  s->E[0][c] = (*Dx - *Px) / (RC_EPS0 * m->eps);
  s->E[1][c] = (*Dy - *Py) / (RC_EPS0 * m->eps);
  s->E[2][c] = (*Dz - *Pz) / (RC_EPS0 * m->eps);
}

int rc_set_matid(rc_sim* s, const uint8_t* ids) {
  memcpy(s->matid, ids, s->ncell);
  if (s->im) rc_set_matid(s->im, ids);
  return 0;
}

int rc_add_source(rc_sim* s, int field, int comp, int i, int j, int k, const double* table, int ntable) {
  if (field < 0 || field > 3 || comp < 0 || comp > 2) return -1;
  if (!(field % 2 == 0 ? s->actE[comp] : s->actH[comp])) return -1;
  s->src = realloc(s->src, sizeof(rc_src) * (s->nsrc + 1));
  rc_src* q = &s->src[s->nsrc++];
  q->field = field;
  q->comp = comp;
  q->cell = cidx(s, i, j, k);
  q->n = ntable;
  q->table = malloc(sizeof(double) * ntable);
  memcpy(q->table, table, sizeof(double) * ntable);
  return 0;
}

int rc_set_tfsf(rc_sim* s, const int box[6], const double kvec[3], const double pol[3], const double polH[3],
                const double origin[3], int aux_ncells, const double* signal, int nsignal) {
  s->tfsf = 1;
  memcpy(s->box, box, sizeof(int) * 6);
  /* TFSFPlane normalises the wave vector, Source.hpp:566-570 */
  int nd = s->mode == RC_3D ? 3 : (s->mode == RC_TEM ? 1 : 2);
  double K = 0;
  for (int d = 0; d < nd; d++) K += kvec[d] * kvec[d];
  K = sqrt(K);
  for (int d = 0; d < 3; d++) s->kvec[d] = d < nd ? kvec[d] / K : 0.0;
  memcpy(s->pol, pol, sizeof(double) * 3);
  memcpy(s->polH, polH, sizeof(double) * 3);
  memcpy(s->origin, origin, sizeof(double) * 3);
  rc_aux_destroy(s->aux);
  s->aux = rc_aux_create(aux_ncells, s->dx, s->dt);
  free(s->signal);
  s->signal = malloc(sizeof(double) * nsignal);
  memcpy(s->signal, signal, sizeof(double) * nsignal);
  s->nsignal = nsignal;
  return 0;
}

int rc_add_tfsf_plane(rc_sim* s, int axis, int side, const int cells[6], const double kvec[3], const double pol[3],
                      const double polH[3], const double origin[3], int aux_ncells, const double* signal,
                      int nsignal) {
  if (s->nplane >= RC_MAXPLANE || axis < 0 || axis > 2 || !s->axes[axis]) return -1;
  rc_plane* P = &s->plane[s->nplane++];
  P->axis = axis;
  P->nrm = side ? 1.0 : -1.0;
  for (int d = 0; d < 3; d++) {
    P->lo[d] = cells[2 * d];
    P->hi[d] = cells[2 * d + 1] - 1;
  }
  /* TFSFPlane normalises the wave vector, Source.hpp:566-570 */
  int nd = s->mode == RC_3D ? 3 : (s->mode == RC_TEM ? 1 : 2);
  double K = 0;
  for (int d = 0; d < nd; d++) K += kvec[d] * kvec[d];
  K = sqrt(K);
  for (int d = 0; d < 3; d++) P->kvec[d] = d < nd ? kvec[d] / K : 0.0;
  memcpy(P->pol, pol, sizeof(double) * 3);
  memcpy(P->polH, polH, sizeof(double) * 3);
  memcpy(P->origin, origin, sizeof(double) * 3);
  P->aux = rc_aux_create(aux_ncells, s->dx, s->dt);
  P->signal = malloc(sizeof(double) * (nsignal > 0 ? nsignal : 1));
  memcpy(P->signal, signal, sizeof(double) * nsignal);
  P->nsignal = nsignal;
  return 0;
}

int rc_add_dft(rc_sim* s, const int region[6], int comp_mask, int nfreq, const double* phaseE, const double* phaseH,
               int nphase) {
  s->dft = realloc(s->dft, sizeof(rc_dft) * (s->ndft + 1));
  rc_dft* d = &s->dft[s->ndft++];
  memcpy(d->region, region, sizeof(int) * 6);
  d->mask = comp_mask;
  d->nfreq = nfreq;
  d->nphase = nphase;
  d->ncomp = 0;
  for (int c = 0; c < 6; c++) d->ncomp += (comp_mask >> c) & 1;
  d->ncells = (size_t)(region[1] - region[0]) * (region[3] - region[2]) * (region[5] - region[4]);
  d->phE = malloc(sizeof(double) * 2 * nfreq * nphase);
  d->phH = malloc(sizeof(double) * 2 * nfreq * nphase);
  memcpy(d->phE, phaseE, sizeof(double) * 2 * nfreq * nphase);
  memcpy(d->phH, phaseH, sizeof(double) * 2 * nfreq * nphase);
  d->acc = calloc((size_t)d->ncomp * nfreq * 2 * d->ncells, sizeof(double));
  if (s->im) rc_add_dft(s->im, region, comp_mask, nfreq, phaseE, phaseH, nphase);
  return s->ndft - 1;
}

/* the spectrum of the (complex) field: F = F_re + i F_im for a complex grid */
int rc_get_dft(const rc_sim* s, int idx, double* out) {
  if (idx < 0 || idx >= s->ndft) return -1;
  const rc_dft* d = &s->dft[idx];
  memcpy(out, d->acc, sizeof(double) * (size_t)d->ncomp * d->nfreq * 2 * d->ncells);
  if (s->im) {
    const double* q = s->im->dft[idx].acc;
    for (size_t b = 0; b < (size_t)d->ncomp * d->nfreq; b++)
      for (size_t c = 0; c < d->ncells; c++) {
        out[b * 2 * d->ncells + c] -= q[b * 2 * d->ncells + d->ncells + c];
        out[b * 2 * d->ncells + d->ncells + c] += q[b * 2 * d->ncells + c];
      }
  }
  return 0;
}

int rc_add_probe(rc_sim* s, int field, int comp, int i, int j, int k) {
  s->probes = realloc(s->probes, sizeof(rc_probe) * (s->nprobes + 1));
  rc_probe* p = &s->probes[s->nprobes++];
  p->field = field;
  p->comp = comp;
  p->cell = cidx(s, i, j, k);
  return 0;
}

int rc_get_probes_imag(const rc_sim* s, double* out, int max_steps) {
  return s->im ? rc_get_probes(s->im, out, max_steps) : -1;
}

int rc_get_probes(const rc_sim* s, double* out, int max_steps) {
  long n = s->step < max_steps ? s->step : max_steps;
  if (s->ptrace) memcpy(out, s->ptrace, sizeof(double) * n * s->nprobes);
  return (int)n;
}

static double* field_ptr(const rc_sim* s, int field, int comp, int pole) {
  if (field == 6 || field == 7) return (comp >= 0 && comp < 9) ? s->psi[field - 6][comp / 3][comp % 3] : NULL;
  if (comp < 0 || comp > 2) return NULL;
  switch (field) {
    case 0: return s->D[comp];
    case 1: return s->E[comp];
    case 2: return s->B[comp];
    case 3: return s->H[comp];
    case 4: return pole < s->npoles_alloc ? s->P[pole][comp] : NULL;
    case 5: return pole < s->npoles_alloc ? s->J[pole][comp] : NULL;
    case 8: return pole < s->mpoles_alloc ? s->M[pole][comp] : NULL;
    case 9: return pole < s->mpoles_alloc ? s->K[pole][comp] : NULL;
  }
  return NULL;
}

int rc_get_field(const rc_sim* s, int field, int comp, int pole, double* out) {
  if (field & 16) { /* the imaginary part of a complex grid (zero before its first step) */
    if (!s->im) {
      if (!field_ptr(s, field & ~16, comp, pole)) return -1;
      memset(out, 0, sizeof(double) * s->ncell);
      return 0;
    }
    return rc_get_field(s->im, field & ~16, comp, pole, out);
  }
  double* p = field_ptr(s, field, comp, pole);
  if (!p) return -1;
  memcpy(out, p, sizeof(double) * s->ncell);
  return 0;
}

int rc_set_field(rc_sim* s, int field, int comp, int pole, const double* in) {
  double* p = field_ptr(s, field, comp, pole);
  if (!p) return -1;
  memcpy(p, in, sizeof(double) * s->ncell);
  return 0;
}

int rc_nthreads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
  return omp_get_max_threads();
#else
  (void)n;
  return 1;
#endif
}

/* interior index ranges: 1..N-2 on every axis of the mode (testTM.cpp:103) */
static void interior(const rc_sim* s, int lo[3], int hi[3]) {
  int len[3] = {s->nx, s->ny, s->nz};
  for (int a = 0; a < 3; a++) {
    lo[a] = s->axes[a] ? 1 : 0;
    hi[a] = s->axes[a] ? len[a] - 1 : 1;
  }
}

static size_t stride(const rc_sim* s, int a) { return a == 0 ? 1 : a == 1 ? (size_t)s->nx : (size_t)s->nx * s->ny; }

/* ---------------------------------------------------------------- boundaries
 * BoundaryUpdates.hpp: SingleUpdate::atomic_update (258-300) writes, for every
 * component of one field type, dest = decorate() * Mask * source, where dest
 * is the boundary cell's outward neighbour and source the boundary cell itself
 * (PEC 527-565, PMC 578-616, Symmetric 629-668, Antisymmetric 681-722) or its
 * periodic image (make_periodic_boundary, 800-832; mask and decorator 1).
 * The masks (component along the boundary axis / across it):
 *   PEC: E +1/-1, H -1/+1;  PMC: E -1/+1, H +1/-1;
 *   Symmetric: -1/+1 for E and H;  Antisymmetric: +1/-1 for E and H.
 * Placement, which the reference leaves to the caller, is the build's: the
 * ghost layer is index 0 (min) / L-1 (max) of the axis, the boundary cell the
 * adjacent interior layer 1 / L-2, the periodic image of ghost 0 is layer
 * L-2 and of ghost L-1 layer 1; the tangential range is the interior cells
 * (the ghost edges are never read by the stencils).  E ghosts are written
 * after the E update, H ghosts after the H update. */
double rc_bc_mask(int type, int ft, int comp, int axis) {
  const int along = comp == axis;
  switch (type) {
    case RC_BC_PEC: return ft == 0 ? (along ? 1.0 : -1.0) : (along ? -1.0 : 1.0);
    case RC_BC_PMC: return ft == 0 ? (along ? -1.0 : 1.0) : (along ? 1.0 : -1.0);
    case RC_BC_SYMMETRIC: return along ? -1.0 : 1.0;
    case RC_BC_ANTISYMMETRIC: return along ? 1.0 : -1.0;
    default: return 1.0;
  }
}

int rc_set_boundary(rc_sim* s, int axis, int side, int type) {
  if (axis < 0 || axis > 2 || side < 0 || side > 1 || type < RC_BC_NONE || type > RC_BC_PERIODIC) return -1;
  if (!s->axes[axis]) return -1;
  s->bc[axis][side] = type;
  return 0;
}

/* Bloch-periodic face (make_bloch_periodic_boundary, BoundaryUpdates.hpp:845-885):
 * ghost = factor * image for E, conj(factor) * image for H (mask 1), on
 * complex fields held as a real and an imaginary part */
int rc_set_bloch_boundary(rc_sim* s, int axis, int side, double fr, double fi) {
  if (axis < 0 || axis > 2 || side < 0 || side > 1 || !s->axes[axis] || s->step > 0) return -1;
  s->bc[axis][side] = RC_BC_BLOCH;
  s->bloch[axis][side][0] = fr;
  s->bloch[axis][side][1] = fi;
  return 0;
}

static int has_bloch(const rc_sim* s) {
  for (int a = 0; a < 3; a++)
    if (s->bc[a][0] == RC_BC_BLOCH || s->bc[a][1] == RC_BC_BLOCH) return 1;
  return 0;
}

/* the imaginary part: same grid, PML, materials, fluid state, boundaries,
 * DFT monitors and probes; no sources and no TFSF (the excitation is real) */
static rc_sim* clone_part(const rc_sim* s) {
  rc_sim* t = rc_create(s->mode, s->nx, s->ny, s->nz, s->dx, s->dt);
  const int len[3] = {s->nx, s->ny, s->nz};
  for (int ft = 0; ft < 2; ft++)
    for (int a = 0; a < 3; a++) memcpy(t->pml[ft][a], s->pml[ft][a], sizeof(double) * 4 * len[a]);
  memcpy(t->matid, s->matid, s->ncell);
  memcpy(t->mat, s->mat, sizeof s->mat);
  for (int p = 0; p < s->npoles_alloc; p++)
    for (int c = 0; c < 3; c++)
      if (t->actE[c]) {
        t->P[p][c] = calloc(t->ncell, sizeof(double));
        t->J[p][c] = calloc(t->ncell, sizeof(double));
      }
  t->npoles_alloc = s->npoles_alloc;
  for (int p = 0; p < s->mpoles_alloc; p++)
    for (int c = 0; c < 3; c++)
      if (t->actH[c]) {
        t->M[p][c] = calloc(t->ncell, sizeof(double));
        t->K[p][c] = calloc(t->ncell, sizeof(double));
      }
  t->mpoles_alloc = s->mpoles_alloc;
  memcpy(t->bc, s->bc, sizeof s->bc);
  memcpy(t->bloch, s->bloch, sizeof s->bloch);
  if (s->fluid_ne) rc_set_fluid_state(t, s->fluid_ne, s->fluid_nu);
  for (int q = 0; q < s->ndft; q++) {
    const rc_dft* d = &s->dft[q];
    rc_add_dft(t, d->region, d->mask, d->nfreq, d->phE, d->phH, d->nphase);
  }
  for (int p = 0; p < s->nprobes; p++) {
    const rc_probe* q = &s->probes[p];
    t->probes = realloc(t->probes, sizeof(rc_probe) * (t->nprobes + 1));
    t->probes[t->nprobes++] = *q;
  }
  return t;
}

/* the Bloch ghost layers of both parts from both parts' periodic images */
static void bloch_ghosts(rc_sim* s, int ft) {
  rc_sim* m = s->im;
  double** Fr = ft == 0 ? s->E : s->H;
  double** Fi = ft == 0 ? m->E : m->H;
  const int* act = ft == 0 ? s->actE : s->actH;
  int lo[3], hi[3];
  interior(s, lo, hi);
  const int len[3] = {s->nx, s->ny, s->nz};
  for (int a = 0; a < 3; a++)
    for (int side = 0; side < 2; side++) {
      if (s->bc[a][side] != RC_BC_BLOCH) continue;
      const double fr = s->bloch[a][side][0], fi = ft == 0 ? s->bloch[a][side][1] : -s->bloch[a][side][1];
      const int L = len[a], dst = side ? L - 1 : 0, src = side ? 1 : L - 2;
      int rlo[3] = {lo[0], lo[1], lo[2]}, rhi[3] = {hi[0], hi[1], hi[2]};
      rlo[a] = 0;
      rhi[a] = 1;
      for (int k = rlo[2]; k < rhi[2]; k++)
        for (int j = rlo[1]; j < rhi[1]; j++)
          for (int i = rlo[0]; i < rhi[0]; i++) {
            int cd[3] = {i, j, k}, cs[3] = {i, j, k};
            cd[a] = dst;
            cs[a] = src;
            const size_t d = cidx(s, cd[0], cd[1], cd[2]), q = cidx(s, cs[0], cs[1], cs[2]);
            for (int c = 0; c < 3; c++)
              if (act[c]) {
                /* (fr + i fi)(sr + i si), std::complex's product */
                const double sr = Fr[c][q], si = Fi[c][q];
                Fr[c][d] = fr * sr - fi * si;
                Fi[c][d] = fr * si + fi * sr;
              }
          }
    }
}

static void apply_boundaries(rc_sim* s, int ft) {
  double** F = ft == 0 ? s->E : s->H;
  const int* act = ft == 0 ? s->actE : s->actH;
  int lo[3], hi[3];
  interior(s, lo, hi);
  const int len[3] = {s->nx, s->ny, s->nz};
  for (int a = 0; a < 3; a++)
    for (int side = 0; side < 2; side++) {
      const int type = s->bc[a][side];
      if (!type || type == RC_BC_BLOCH) continue; /* Bloch faces: bloch_ghosts */
      const int L = len[a];
      const int dst = side ? L - 1 : 0;
      const int src = type == RC_BC_PERIODIC ? (side ? 1 : L - 2) : (side ? L - 2 : 1);
      int rlo[3] = {lo[0], lo[1], lo[2]}, rhi[3] = {hi[0], hi[1], hi[2]};
      rlo[a] = 0;
      rhi[a] = 1;
      for (int k = rlo[2]; k < rhi[2]; k++)
        for (int j = rlo[1]; j < rhi[1]; j++)
          for (int i = rlo[0]; i < rhi[0]; i++) {
            int cd[3] = {i, j, k}, cs[3] = {i, j, k};
            cd[a] = dst;
            cs[a] = src;
            const size_t d = cidx(s, cd[0], cd[1], cd[2]), q = cidx(s, cs[0], cs[1], cs[2]);
            for (int c = 0; c < 3; c++)
              if (act[c]) F[c][d] = 1.0 * rc_bc_mask(type, ft, c, a) * F[c][q];
          }
    }
}

void rc_apply_boundaries(rc_sim* s, int ft) { apply_boundaries(s, ft); }

/* Tangential TFSF range of a component (offset 0 or -1/2 along axis T):
 * the total-field region along T is the open interval (lo-1/2, hi). */
static void tf_range(int lo, int hi, int half, int* a, int* b) {
  if (half) {
    *a = lo + 1;
    *b = hi;
  } else {
    *a = lo;
    *b = hi - 1;
  }
}

/* one TFSF face correction pass; ft 0 = updateD (uses aux H, polH), 1 = updateB (aux E, pol).
 * The per-cell arithmetic is TFSFPlane::field_update, Source.hpp:506-553. */
static void tfsf_faces(rc_sim* s, int ft) {
  const double dt = s->dt, dx = s->dx;
  const int* box = s->box;
  for (int J = 0; J < 3; J++) {
    if (!s->axes[J]) continue;
    for (int side = 0; side < 2; side++) {
      const int layer = box[2 * J + side];
      const double nrm = side == 0 ? -1.0 : 1.0;
      for (int I = 0; I < 3; I++) {
        if (I == J) continue;
        if (!(ft == 0 ? s->actE[I] : s->actH[I])) continue;
        const int K = third(I, J);
        if (!(ft == 0 ? s->actH[K] : s->actE[K])) continue;
        /* ranges along the two tangential axes */
        int lo[3], hi[3];
        for (int T = 0; T < 3; T++) {
          if (T == J) {
            lo[T] = layer;
            hi[T] = layer;
            continue;
          }
          if (!s->axes[T]) {
            lo[T] = 0;
            hi[T] = 0;
            continue;
          }
          /* flux offsets: D_I is -1/2 on axes != I, B_I is -1/2 on axis I */
          int half = ft == 0 ? (T != I) : (T == I);
          tf_range(box[2 * T], box[2 * T + 1], half, &lo[T], &hi[T]);
        }
        double* F = ft == 0 ? s->D[I] : s->B[I];
        const double* pml = s->pml[ft][J];
        const double* polv = ft == 0 ? s->polH : s->pol;
        const int e = lc(I, J, K);
        for (int k = lo[2]; k <= hi[2]; k++)
          for (int j = lo[1]; j <= hi[1]; j++)
            for (int i = lo[0]; i <= hi[0]; i++) {
              int ci[3] = {i, j, k};
              /* curl-field location: cell centre + offset of the K component */
              double src_loc[3];
              double proj = 0;
              for (int dd = 0; dd < 3; dd++) {
                if (!s->axes[dd]) continue;
                double off;
                if (ft == 0)
                  off = (dd == K) ? -0.5 : 0.0; /* H_K offsets */
                else
                  off = (dd != K) ? -0.5 : 0.0; /* E_K offsets */
                src_loc[dd] = (ci[dd] - s->origin[dd]) * dx;
                src_loc[dd] += off * dx;
                proj += s->kvec[dd] * src_loc[dd];
              }
              double srcval = ft == 0 ? rc_aux_interp_H(s->aux, proj) : rc_aux_interp_E(s->aux, proj);
              double term = e * pml[4 * ci[J] + 2] / (dx)*nrm * srcval * polv[K];
              size_t c = cidx(s, i, j, k);
              double coeff = ft == 0 ? 1.0 : -1.0;
              F[c] += coeff * 1.0 * dt * (term + 0.0);
            }
      }
    }
  }
}

/* one single plane's correction pass (TFSFPlane::updateD / updateB,
 * Source.hpp:595-620): every cell of the plane, every flux I but the normal
 * axis J (whose two curl directions have a zero normal component), with the
 * per-cell arithmetic of TFSFPlane::field_update (506-553) */
static void plane_faces(rc_sim* s, const rc_plane* P, int ft) {
  const double dt = s->dt, dx = s->dx;
  const int J = P->axis;
  for (int I = 0; I < 3; I++) {
    if (I == J) continue;
    if (!(ft == 0 ? s->actE[I] : s->actH[I])) continue;
    const int K = third(I, J);
    if (!(ft == 0 ? s->actH[K] : s->actE[K])) continue;
    double* F = ft == 0 ? s->D[I] : s->B[I];
    const double* pml = s->pml[ft][J];
    const double* polv = ft == 0 ? P->polH : P->pol;
    const int e = lc(I, J, K);
    for (int k = P->lo[2]; k <= P->hi[2]; k++)
      for (int j = P->lo[1]; j <= P->hi[1]; j++)
        for (int i = P->lo[0]; i <= P->hi[0]; i++) {
          int ci[3] = {i, j, k};
          double proj = 0;
          for (int dd = 0; dd < 3; dd++) {
            if (!s->axes[dd]) continue;
            double off = ft == 0 ? ((dd == K) ? -0.5 : 0.0) : ((dd != K) ? -0.5 : 0.0);
            double loc = (ci[dd] - P->origin[dd]) * dx;
            loc += off * dx;
            proj += P->kvec[dd] * loc;
          }
          double srcval = ft == 0 ? rc_aux_interp_H(P->aux, proj) : rc_aux_interp_E(P->aux, proj);
          double term = e * pml[4 * ci[J] + 2] / (dx)*P->nrm * srcval * polv[K];
          size_t c = cidx(s, i, j, k);
          double coeff = ft == 0 ? 1.0 : -1.0;
          F[c] += coeff * 1.0 * dt * (term + 0.0);
        }
  }
}

/* soft sources on the derived fields (field 2: E, 3: H), added after the
 * material step (SoftSource::apply, Source.hpp:209-229, placed as Aux1D places
 * its own, 421-425): the next material step recomputes the field from D or B */
static void soft_derived(rc_sim* s, int field, long n) {
  for (int q = 0; q < s->nsrc; q++)
    if (s->src[q].field == field && n < s->src[q].n)
      (field == 2 ? s->E : s->H)[s->src[q].comp][s->src[q].cell] += s->src[q].table[n];
}

/* One step in three phases (the E half step with its boundaries, the H half
 * step with its boundaries, then DFT and probes), so that the two parts of a
 * complex grid advance in lockstep with the Bloch coupling between them. */
static void step_e(rc_sim* s) {
  int lo[3], hi[3];
  interior(s, lo, hi);
  const double dt = s->dt, dx = s->dx;
  const double mu0 = rc_mu0_v();
  const long n = s->step;
  (void)lo; (void)hi; (void)dt; (void)dx; (void)mu0; (void)n;
  /* ---- D curl, YeeUpdate<Mode,D_I> (YeeUpdates.hpp:395-405) */
  for (int I = 0; I < 3; I++) {
    if (!s->actE[I]) continue;
    const int j1 = D1[I], j2 = D2[I];
    const int k1 = third(I, j1), k2 = third(I, j2);
    const int h1 = s->actH[k1] && s->axes[j1], h2 = s->actH[k2] && s->axes[j2];
    const size_t s1 = stride(s, j1), s2 = stride(s, j2);
    double* Df = s->D[I];
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = lo[2]; k < hi[2]; k++)
      for (int j = lo[1]; j < hi[1]; j++)
        for (int i = lo[0]; i < hi[0]; i++) {
          size_t c = cidx(s, i, j, k);
          int ci[3] = {i, j, k};
          double a1 = 0.0, a2 = 0.0;
          if (h1) a1 = lc(I, j1, k1) * s->pml[0][j1][4 * ci[j1] + 2] / (dx) * (s->H[k1][c] - s->H[k1][c - s1]);
          if (h2) a2 = lc(I, j2, k2) * s->pml[0][j2][4 * ci[j2] + 2] / (dx) * (s->H[k2][c] - s->H[k2][c - s2]);
          Df[c] = Df[c] + 1.0 * 1.0 * dt * (a1 + a2);
        }
  }
  for (int q = 0; q < s->nsrc; q++)
    if (s->src[q].field == 0 && n < s->src[q].n) s->D[s->src[q].comp][s->src[q].cell] += s->src[q].table[n];
  /* ---- UpdatePML<Mode,Electric,J> (PML.hpp:315-323), J = x, y, z */
  for (int J = 0; J < 3; J++) {
    if (!s->axes[J]) continue;
    const size_t sJ = stride(s, J);
    const double* pm = s->pml[0][J];
    for (int I = 2; I >= 0; I--) {
      if (I == J || !s->actE[I]) continue;
      const int K = third(I, J);
      double* ps = s->psi[0][I][J];
      double* Df = s->D[I];
      const double* Hk = s->H[K];
      const int e = lc(I, J, K);
#pragma omp parallel for collapse(2) schedule(static)
      for (int k = lo[2]; k < hi[2]; k++)
        for (int j = lo[1]; j < hi[1]; j++)
          for (int i = lo[0]; i < hi[0]; i++) {
            size_t c = cidx(s, i, j, k);
            int ci[3] = {i, j, k};
            const double* p = pm + 4 * ci[J];
            Df[c] += 1.0 * dt * ps[c] * p[3];
            ps[c] = p[0] * ps[c] + p[1] * e / dx * (Hk[c] - Hk[c - sJ]);
          }
    }
  }
  /* ---- TFSF updateD, then the aux line advances (Source.hpp:595-609) */
  if (s->tfsf) {
    tfsf_faces(s, 0);
    rc_aux_step(s->aux, n < s->nsignal ? s->signal[n] : 0.0);
  }
  for (int p = 0; p < s->nplane; p++) {
    rc_plane* P = &s->plane[p];
    plane_faces(s, P, 0);
    rc_aux_step(P->aux, n < P->nsignal ? P->signal[n] : 0.0);
  }
  /* ---- E material (DispersiveMaterials.hpp:354-359, 710-735, 979-998) */
  {
    const int np = s->npoles_alloc;
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = lo[2]; k < hi[2]; k++)
      for (int j = lo[1]; j < hi[1]; j++)
        for (int i = lo[0]; i < hi[0]; i++) {
          size_t c = cidx(s, i, j, k);
          const rc_mat* m = &s->mat[s->matid[c]];
          const long dl = design_local(s, i, j, k);
          if (m->mag && m->np && np && dl < 0 && s->mode == RC_3D) {
            magnetized_drude_3d(s, m, c, dt);
            continue;
          }
          for (int I = 2; I >= 0; I--) {
            if (!s->actE[I]) continue;
            if (dl >= 0) { /* design cell: non-dispersive, its own eps_r */
              s->E[I][c] = s->D[I][c] / (s->dsg_eps[dl] * RC_EPS0);
              continue;
            }
            if (m->np == 0 || np == 0) {
              s->E[I][c] = s->D[I][c] / (m->eps * RC_EPS0);
              continue;
            }
            for (int p = 0; p < m->np; p++) {
              if (m->kind[p] == 3) { /* ConductiveUpdate::implicit_update (DispersiveMaterials.hpp:549-560) */
                const double factor = dt * m->freq[p] / (m->eps * RC_EPS0);
                s->P[p][I][c] = (s->P[p][I][c] + factor * s->D[I][c]) / (1.0 + factor);
              } else {
                s->P[p][I][c] += dt * s->J[p][I][c];
              }
            }
            double num = s->D[I][c];
            for (int p = 0; p < m->np; p++) num = num - s->P[p][I][c];
            double Ev = num / (m->eps * RC_EPS0);
            s->E[I][c] = Ev;
            for (int p = 0; p < m->np; p++) {
              double* Jp = &s->J[p][I][c];
              double* Pp = &s->P[p][I][c];
              const double wgt_eps0 = m->weight[p] * RC_EPS0;
              if (m->kind[p] == 3) continue; /* conductive: J untouched */
              if (m->kind[p] == 1) {
                double fq = m->freq[p], gm = m->gamma[p];
                if (m->fluid) plasma_params(s, m, c, &fq, &gm);
                double b = gm * dt * 0.5;
                double w = dt * fq;
                *Jp = *Jp * (1.0 - b) / (1.0 + b) + w * w / dt * wgt_eps0 * Ev;
              } else {
                double b = 1.0 + m->gamma[p] * dt;
                double bm = 1.0 - m->gamma[p] * dt;
                double w = m->freq[p] * dt;
                *Jp = bm / b * *Jp + m->freq[p] * w / b * wgt_eps0 * Ev - m->freq[p] * w / b * *Pp;
              }
            }
          }
        }
  }
  soft_derived(s, 2, n);
  apply_boundaries(s, 0);
}

static void step_h(rc_sim* s) {
  int lo[3], hi[3];
  interior(s, lo, hi);
  const double dt = s->dt, dx = s->dx;
  const double mu0 = rc_mu0_v();
  const long n = s->step;
  (void)lo; (void)hi; (void)dt; (void)dx; (void)mu0; (void)n;
  /* ---- B curl */
  for (int I = 0; I < 3; I++) {
    if (!s->actH[I]) continue;
    const int j1 = D1[I], j2 = D2[I];
    const int k1 = third(I, j1), k2 = third(I, j2);
    const int e1 = s->actE[k1] && s->axes[j1], e2 = s->actE[k2] && s->axes[j2];
    const size_t s1 = stride(s, j1), s2 = stride(s, j2);
    double* Bf = s->B[I];
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = lo[2]; k < hi[2]; k++)
      for (int j = lo[1]; j < hi[1]; j++)
        for (int i = lo[0]; i < hi[0]; i++) {
          size_t c = cidx(s, i, j, k);
          int ci[3] = {i, j, k};
          double a1 = 0.0, a2 = 0.0;
          if (e1) a1 = lc(I, j1, k1) * s->pml[1][j1][4 * ci[j1] + 2] / (dx) * (s->E[k1][c + s1] - s->E[k1][c]);
          if (e2) a2 = lc(I, j2, k2) * s->pml[1][j2][4 * ci[j2] + 2] / (dx) * (s->E[k2][c + s2] - s->E[k2][c]);
          Bf[c] = Bf[c] + -1.0 * 1.0 * dt * (a1 + a2);
        }
  }
  for (int q = 0; q < s->nsrc; q++)
    if (s->src[q].field == 1 && n < s->src[q].n) s->B[s->src[q].comp][s->src[q].cell] += s->src[q].table[n];
  /* ---- UpdatePML<Mode,Magnetic,J> */
  for (int J = 0; J < 3; J++) {
    if (!s->axes[J]) continue;
    const size_t sJ = stride(s, J);
    const double* pm = s->pml[1][J];
    for (int I = 2; I >= 0; I--) {
      if (I == J || !s->actH[I]) continue;
      const int K = third(I, J);
      double* ps = s->psi[1][I][J];
      double* Bf = s->B[I];
      const double* Ek = s->E[K];
      const int e = lc(I, J, K);
#pragma omp parallel for collapse(2) schedule(static)
      for (int k = lo[2]; k < hi[2]; k++)
        for (int j = lo[1]; j < hi[1]; j++)
          for (int i = lo[0]; i < hi[0]; i++) {
            size_t c = cidx(s, i, j, k);
            int ci[3] = {i, j, k};
            const double* p = pm + 4 * ci[J];
            Bf[c] += -1.0 * dt * ps[c] * p[3];
            ps[c] = p[0] * ps[c] + p[1] * e / dx * (Ek[c + sJ] - Ek[c]);
          }
    }
  }
  if (s->tfsf) tfsf_faces(s, 1);
  for (int p = 0; p < s->nplane; p++) plane_faces(s, &s->plane[p], 1);
  /* ---- H material: H = B/(mu_r mu0); magnetic poles: the electric side's
   * updates on B, H, M, K with mu0 (DispersiveMaterials.hpp:549-560, 710-735, 979-998) */
  const int mnp_alloc = s->mpoles_alloc;
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = lo[2]; k < hi[2]; k++)
    for (int j = lo[1]; j < hi[1]; j++)
      for (int i = lo[0]; i < hi[0]; i++) {
        size_t c = cidx(s, i, j, k);
        const rc_mat* m = &s->mat[s->matid[c]];
        for (int I = 2; I >= 0; I--) {
          if (!s->actH[I]) continue;
          if (m->mnp == 0 || mnp_alloc == 0) {
            s->H[I][c] = s->B[I][c] / (m->mu * mu0);
            continue;
          }
          for (int p = 0; p < m->mnp; p++) {
            if (m->mkind[p] == 3) { /* ConductiveUpdate::implicit_update */
              const double factor = dt * m->mfreq[p] / (m->mu * mu0);
              s->M[p][I][c] = (s->M[p][I][c] + factor * s->B[I][c]) / (1.0 + factor);
            } else {
              s->M[p][I][c] += dt * s->K[p][I][c];
            }
          }
          double num = s->B[I][c];
          for (int p = 0; p < m->mnp; p++) num = num - s->M[p][I][c];
          const double Hv = num / (m->mu * mu0);
          s->H[I][c] = Hv;
          for (int p = 0; p < m->mnp; p++) {
            double* Kp = &s->K[p][I][c];
            double* Mp = &s->M[p][I][c];
            const double wgt_mu0 = m->mweight[p] * mu0;
            if (m->mkind[p] == 3) continue; /* conductive: K untouched */
            if (m->mkind[p] == 1) {
              double b = m->mgamma[p] * dt * 0.5;
              double w = dt * m->mfreq[p];
              *Kp = *Kp * (1.0 - b) / (1.0 + b) + w * w / dt * wgt_mu0 * Hv;
            } else {
              double b = 1.0 + m->mgamma[p] * dt;
              double bm = 1.0 - m->mgamma[p] * dt;
              double w = m->mfreq[p] * dt;
              *Kp = bm / b * *Kp + m->mfreq[p] * w / b * wgt_mu0 * Hv - m->mfreq[p] * w / b * *Mp;
            }
          }
        }
      }
  soft_derived(s, 3, n);
  apply_boundaries(s, 1);
}

static void step_out(rc_sim* s) {
  int lo[3], hi[3];
  interior(s, lo, hi);
  const double dt = s->dt, dx = s->dx;
  const double mu0 = rc_mu0_v();
  const long n = s->step;
  (void)lo; (void)hi; (void)dt; (void)dx; (void)mu0; (void)n;
  /* ---- running DFT (build-defined): F += f(t) e^{-i w t} dt, E at (n+1)dt, H at (n+1.5)dt */
  for (int q = 0; q < s->ndft; q++) {
    rc_dft* d = &s->dft[q];
    if (n >= d->nphase) continue;
    const int* r = d->region;
    int slot = 0;
    for (int comp = 0; comp < 6; comp++) {
      if (!((d->mask >> comp) & 1)) continue;
      const double* fld = comp < 3 ? s->E[comp] : s->H[comp - 3];
      const double* ph = (comp < 3 ? d->phE : d->phH) + 2 * (size_t)d->nfreq * n;
      for (int f = 0; f < d->nfreq; f++) {
        double* re = d->acc + ((size_t)slot * d->nfreq + f) * 2 * d->ncells;
        double* im = re + d->ncells;
        size_t q2 = 0;
        for (int k = r[4]; k < r[5]; k++)
          for (int j = r[2]; j < r[3]; j++)
            for (int i = r[0]; i < r[1]; i++, q2++) {
              double v = fld ? fld[cidx(s, i, j, k)] : 0.0;
              re[q2] += v * ph[2 * f];
              im[q2] += v * ph[2 * f + 1];
            }
      }
      slot++;
    }
  }
  /* ---- probes */
  if (s->nprobes) {
    if (s->step + 1 > s->ptrace_cap) {
      int cap = s->ptrace_cap ? 2 * s->ptrace_cap : 1024;
      while (cap < s->step + 1) cap *= 2;
      s->ptrace = realloc(s->ptrace, sizeof(double) * cap * s->nprobes);
      s->ptrace_cap = cap;
    }
    for (int p = 0; p < s->nprobes; p++) {
      const double* f = field_ptr(s, s->probes[p].field, s->probes[p].comp, 0);
      s->ptrace[(size_t)s->step * s->nprobes + p] = f ? f[s->probes[p].cell] : 0.0;
    }
  }
  s->step++;
}

int rc_step(rc_sim* s, int nsteps) {
  if (has_bloch(s)) {
    /* the parts are stepped apart: only updates linear with real coefficients */
    for (int m = 0; m < RC_MAXMAT; m++)
      if (s->mat[m].np && s->mat[m].mag == 2 && s->mode == RC_3D) return -1;
    if (!s->im) s->im = clone_part(s);
  }
  for (int it = 0; it < nsteps; it++) {
    step_e(s);
    if (s->im) {
      step_e(s->im);
      bloch_ghosts(s, 0);
    }
    step_h(s);
    if (s->im) {
      step_h(s->im);
      bloch_ghosts(s, 1);
    }
    step_out(s);
    if (s->im) step_out(s->im);
  }
  return 0;
}

/* ---------------------------------------------------------------- surface quantities
 * SurfaceQuantities.hpp: UpdateSurfaceChargeDensity (174-183), UpdateSurfaceCurrentE
 * (233-244, TM 258-266) and UpdateSurfaceCurrentH (293-304, TM 319-327), in the
 * reference's operation order, on the current E and H of the cells of
 * `region` ([x0,x1)x[y0,y1)x[z0,z1), neighbours inside the grid).  out:
 * quantity-major, x fastest; 3-D: rho, Jex, Jey, Jez, Jhx, Jhy, Jhz; TM: Jez,
 * Jhx, Jhy (the TM charge-density functor does nothing). */
int rc_surface_quantities(const rc_sim* s, const int r[6], double* out) {
  const int n[3] = {s->nx, s->ny, s->nz};
  for (int a = 0; a < 3; a++) {
    const int lo = s->axes[a] ? 1 : 0, hi = s->axes[a] ? n[a] - 1 : 1;
    if (r[2 * a] < lo || r[2 * a + 1] > hi || r[2 * a] >= r[2 * a + 1]) return -1;
  }
  const size_t cells = (size_t)(r[1] - r[0]) * (r[3] - r[2]) * (r[5] - r[4]);
  const double dx = s->dx;
  const size_t sy = (size_t)s->nx, sz = (size_t)s->nx * s->ny;
  size_t q = 0;
  for (int k = r[4]; k < r[5]; k++)
    for (int j = r[2]; j < r[3]; j++)
      for (int i = r[0]; i < r[1]; i++, q++) {
        const size_t c = cidx(s, i, j, k);
        if (s->mode == 3) {
          double* const* E = s->E;
          double* const* H = s->H;
          out[0 * cells + q] = RC_EPS0 / dx * (E[0][c] + E[1][c] + E[2][c] - E[0][c - 1] - E[1][c - sy] - E[2][c - sz]);
          out[1 * cells + q] = 1.0 / dx * (H[2][c] - H[2][c - sy] - H[1][c] + H[1][c - sz]);
          out[2 * cells + q] = 1.0 / dx * (H[0][c] - H[0][c - sz] - H[2][c] + H[2][c - 1]);
          out[3 * cells + q] = 1.0 / dx * (H[1][c] - H[1][c - 1] - H[0][c] + H[0][c - sy]);
          out[4 * cells + q] = -1.0 / dx * (E[2][c + sy] - E[2][c] - E[1][c + sz] + E[1][c]);
          out[5 * cells + q] = -1.0 / dx * (E[0][c + sz] - E[0][c] - E[2][c + 1] + E[2][c]);
          out[6 * cells + q] = -1.0 / dx * (E[1][c + 1] - E[1][c] - E[0][c + sy] + E[0][c]);
        } else if (s->mode == RC_TE) { /* SurfaceQuantities.hpp:189-198, 253-263, 323-332 */
          const double* Ex = s->E[0];
          const double* Ey = s->E[1];
          const double* Hz = s->H[2];
          out[0 * cells + q] = RC_EPS0 / dx * (Ex[c] + Ey[c] - Ex[c - 1] - Ey[c - sy]);
          out[1 * cells + q] = 1.0 / dx * (Hz[c] - Hz[c - sy]);
          out[2 * cells + q] = 1.0 / dx * (-Hz[c] + Hz[c - 1]);
          out[3 * cells + q] = -1.0 / dx * (Ey[c + 1] - Ey[c] - Ex[c + sy] + Ex[c]);
        } else {
          const double* Ez = s->E[2];
          const double* Hx = s->H[0];
          const double* Hy = s->H[1];
          out[0 * cells + q] = 1.0 / dx * (Hy[c] - Hy[c - 1] - Hx[c] + Hx[c - sy]);
          out[1 * cells + q] = -1.0 / dx * (Ez[c + sy] - Ez[c]);
          out[2 * cells + q] = -1.0 / dx * (-Ez[c + 1] + Ez[c]);
        }
      }
  return 0;
}

/* ---------------------------------------------------------------- adjoint
 * Build-defined (the reference has no adjoint).  Reverse mode of rc_step, op
 * by op in the D/B form above: every forward update y += a x becomes
 * x_bar += a y_bar, run in the reverse order of the forward ops; sources and
 * TFSF terms do not depend on the state and drop out.  Each material op
 * E = D/(eps_r eps0) contributes dJ/d eps_r += E_bar (-E/eps_r); the cells of
 * pole materials (Drude, Lorentz, conductive) run the reverse of their P / J
 * recursion (adj_pole_cell), with P and J adjoints carried across steps.
 * Written independently of the device adjoint, which works on E and H
 * increments with gathered (not scattered) transposes. */

int rc_set_design(rc_sim* s, const int box[6], const double* eps_r) {
  const size_t n = (size_t)(box[1] - box[0]) * (box[3] - box[2]) * (box[5] - box[4]);
  free(s->dsg_eps);
  s->dsg_eps = malloc(sizeof(double) * n);
  memcpy(s->dsg_eps, eps_r, sizeof(double) * n);
  memcpy(s->dsg, box, sizeof(int) * 6);
  s->has_design = 1;
  return 0;
}

static double cell_eps(const rc_sim* s, int i, int j, int k) {
  const long dl = design_local(s, i, j, k);
  return dl >= 0 ? s->dsg_eps[dl] : s->mat[s->matid[cidx(s, i, j, k)]].eps;
}

void rc_adjoint_reset(rc_sim* s) {
  for (int c = 0; c < 3; c++) {
    if (s->actE[c]) {
      if (!s->Db[c]) s->Db[c] = malloc(sizeof(double) * s->ncell);
      if (!s->Eb[c]) s->Eb[c] = malloc(sizeof(double) * s->ncell);
      memset(s->Db[c], 0, sizeof(double) * s->ncell);
      memset(s->Eb[c], 0, sizeof(double) * s->ncell);
    }
    if (s->actH[c]) {
      if (!s->Bb[c]) s->Bb[c] = malloc(sizeof(double) * s->ncell);
      if (!s->Hb[c]) s->Hb[c] = malloc(sizeof(double) * s->ncell);
      memset(s->Bb[c], 0, sizeof(double) * s->ncell);
      memset(s->Hb[c], 0, sizeof(double) * s->ncell);
    }
  }
  for (int ft = 0; ft < 2; ft++)
    for (int I = 0; I < 3; I++)
      for (int J = 0; J < 3; J++)
        if (s->psi[ft][I][J]) {
          if (!s->Pb[ft][I][J]) s->Pb[ft][I][J] = malloc(sizeof(double) * s->ncell);
          memset(s->Pb[ft][I][J], 0, sizeof(double) * s->ncell);
        }
  for (int p = 0; p < s->npoles_alloc; p++)
    for (int c = 0; c < 3; c++)
      if (s->actE[c]) {
        if (!s->PPb[p][c]) s->PPb[p][c] = malloc(sizeof(double) * s->ncell);
        if (!s->JPb[p][c]) s->JPb[p][c] = malloc(sizeof(double) * s->ncell);
        memset(s->PPb[p][c], 0, sizeof(double) * s->ncell);
        memset(s->JPb[p][c], 0, sizeof(double) * s->ncell);
      }
}

/* Reverse of the pole material step of one component of one dispersive cell
 * (the forward's op order in step_e: P_p (conductive: implicit), num = D -
 * sum P, E = num/(eps_r eps0), J_p per kind), op by op from the last:
 * Eb (the E adjoint gathered this step) is consumed; D, P_p and J_p
 * adjoints carry across steps. */
static void adj_pole_cell(rc_sim* s, const rc_mat* m, int I, size_t c) {
  const double dt = s->dt;
  double eb = s->Eb[I][c];
  for (int p = m->np - 1; p >= 0; p--) {
    double* jb = &s->JPb[p][I][c];
    double* pb = &s->PPb[p][I][c];
    const double wgt_eps0 = m->weight[p] * RC_EPS0;
    if (m->kind[p] == 3) continue;
    if (m->kind[p] == 1) { /* J = J (1-b)/(1+b) + w^2/dt wgt eps0 E */
      const double b = m->gamma[p] * dt * 0.5, w = dt * m->freq[p];
      eb += w * w / dt * wgt_eps0 * *jb;
      *jb = *jb * (1.0 - b) / (1.0 + b);
    } else { /* J = bm/b J + freq w/b wgt eps0 E - freq w/b P */
      const double b = 1.0 + m->gamma[p] * dt, bm = 1.0 - m->gamma[p] * dt, w = m->freq[p] * dt;
      eb += m->freq[p] * w / b * wgt_eps0 * *jb;
      *pb += -(m->freq[p] * w / b) * *jb;
      *jb = bm / b * *jb;
    }
  }
  const double numb = eb / (m->eps * RC_EPS0); /* E = num/(eps_r eps0) */
  s->Db[I][c] += numb;                         /* num = D - sum_p P_p */
  for (int p = 0; p < m->np; p++) s->PPb[p][I][c] -= numb;
  for (int p = m->np - 1; p >= 0; p--) {
    double* pb = &s->PPb[p][I][c];
    if (m->kind[p] == 3) { /* P = (P + f D)/(1 + f) */
      const double f = dt * m->freq[p] / (m->eps * RC_EPS0);
      const double t = *pb / (1.0 + f);
      s->Db[I][c] += f * t;
      *pb = t;
    } else { /* P += dt J */
      s->JPb[p][I][c] += dt * *pb;
    }
  }
  s->Eb[I][c] = 0.0;
}

/* Reverse of forward step n.  w: objective weights 2F of monitor obj_dft
 * ([slot][nfreq][2][ncells]) or NULL; Edes: E after step n at the design
 * cells ([3][ndes]) or NULL; grad accumulates dJ/d eps_r. */
static void adj_step(rc_sim* s, long n, const double* w, const double* Edes, double* grad) {
  int lo[3], hi[3];
  interior(s, lo, hi);
  const double dt = s->dt, dx = s->dx, mu0 = rc_mu0_v();
  const long nd = s->has_design ? (long)(s->dsg[1] - s->dsg[0]) * (s->dsg[3] - s->dsg[2]) * (s->dsg[5] - s->dsg[4]) : 0;
  /* DFT of E after step n: J = sum |F|^2, dJ/dE_n = 2 Re(conj(F) phase_n) */
  if (w) {
    const rc_dft* d = &s->dft[s->obj_dft];
    if (n < d->nphase) {
      const double* ph = d->phE + 2 * (size_t)d->nfreq * n;
      int slot = 0;
      for (int comp = 0; comp < 3; comp++) {
        if (!((s->obj_mask >> comp) & 1)) continue;
        for (int f = 0; f < d->nfreq; f++) {
          const double* wr = w + ((size_t)slot * d->nfreq + f) * 2 * d->ncells;
          const double* wi = wr + d->ncells;
          size_t q2 = 0;
          const int* r = d->region;
          for (int k = r[4]; k < r[5]; k++)
            for (int j = r[2]; j < r[3]; j++)
              for (int i = r[0]; i < r[1]; i++, q2++) s->Eb[comp][cidx(s, i, j, k)] += wr[q2] * ph[2 * f] + wi[q2] * ph[2 * f + 1];
        }
        slot++;
      }
    }
  }
  /* H = B/(mu_r mu0) */
  for (int k = lo[2]; k < hi[2]; k++)
    for (int j = lo[1]; j < hi[1]; j++)
      for (int i = lo[0]; i < hi[0]; i++) {
        const size_t c = cidx(s, i, j, k);
        const double mu = s->mat[s->matid[c]].mu;
        for (int I = 0; I < 3; I++)
          if (s->actH[I]) s->Bb[I][c] += s->Hb[I][c] / (mu * mu0);
      }
  for (int I = 0; I < 3; I++)
    if (s->actH[I]) memset(s->Hb[I], 0, sizeof(double) * s->ncell);
  /* UpdatePML magnetic: B_I += -dt G psi; psi = B psi + C e/dx (E_K[c+sJ] - E_K[c]) */
  for (int J = 2; J >= 0; J--) {
    if (!s->axes[J]) continue;
    const size_t sJ = stride(s, J);
    const double* pm = s->pml[1][J];
    for (int I = 0; I < 3; I++) {
      if (I == J || !s->actH[I]) continue;
      const int K = third(I, J);
      const int e = lc(I, J, K);
      double* pb = s->Pb[1][I][J];
      const double* bb = s->Bb[I];
      double* ekb = s->Eb[K];
      for (int k = lo[2]; k < hi[2]; k++)
        for (int j = lo[1]; j < hi[1]; j++)
          for (int i = lo[0]; i < hi[0]; i++) {
            const size_t c = cidx(s, i, j, k);
            const int ci[3] = {i, j, k};
            const double* p = pm + 4 * ci[J];
            const double t = pb[c];
            if (ekb) {
              const double a = p[1] * e / dx * t;
              ekb[c + sJ] += a;
              ekb[c] -= a;
            }
            pb[c] = p[0] * t - dt * p[3] * bb[c];
          }
    }
  }
  /* B curl: B_I += -dt e F/dx (E_K[c+sJ] - E_K[c]) */
  for (int I = 2; I >= 0; I--) {
    if (!s->actH[I]) continue;
    const int js[2] = {D1[I], D2[I]};
    for (int t = 0; t < 2; t++) {
      const int Jx = js[t], K = third(I, Jx);
      if (!(s->actE[K] && s->axes[Jx])) continue;
      const size_t sJ = stride(s, Jx);
      const int e = lc(I, Jx, K);
      for (int k = lo[2]; k < hi[2]; k++)
        for (int j = lo[1]; j < hi[1]; j++)
          for (int i = lo[0]; i < hi[0]; i++) {
            const size_t c = cidx(s, i, j, k);
            const int ci[3] = {i, j, k};
            const double a = -dt * e * s->pml[1][Jx][4 * ci[Jx] + 2] / dx * s->Bb[I][c];
            s->Eb[K][c + sJ] += a;
            s->Eb[K][c] -= a;
          }
    }
  }
  /* E = D/(eps_r eps0): D_bar += E_bar/(eps_r eps0), dJ/d eps_r += E_bar (-E/eps_r);
   * dispersive cells: the reverse of their pole step */
  for (int k = lo[2]; k < hi[2]; k++)
    for (int j = lo[1]; j < hi[1]; j++)
      for (int i = lo[0]; i < hi[0]; i++) {
        const size_t c = cidx(s, i, j, k);
        const double eps = cell_eps(s, i, j, k);
        const long dl = design_local(s, i, j, k);
        const rc_mat* m = &s->mat[s->matid[c]];
        for (int I = 0; I < 3; I++) {
          if (!s->actE[I]) continue;
          if (dl < 0 && m->np && s->npoles_alloc) {
            adj_pole_cell(s, m, I, c);
            continue;
          }
          if (dl >= 0 && grad && Edes) grad[dl] += s->Eb[I][c] * (-Edes[(size_t)I * nd + dl] / eps);
          s->Db[I][c] += s->Eb[I][c] / (eps * RC_EPS0);
        }
      }
  for (int I = 0; I < 3; I++)
    if (s->actE[I]) memset(s->Eb[I], 0, sizeof(double) * s->ncell);
  /* UpdatePML electric: D_I += dt G psi; psi = B psi + C e/dx (H_K[c] - H_K[c-sJ]) */
  for (int J = 2; J >= 0; J--) {
    if (!s->axes[J]) continue;
    const size_t sJ = stride(s, J);
    const double* pm = s->pml[0][J];
    for (int I = 0; I < 3; I++) {
      if (I == J || !s->actE[I]) continue;
      const int K = third(I, J);
      const int e = lc(I, J, K);
      double* pb = s->Pb[0][I][J];
      const double* db = s->Db[I];
      double* hkb = s->Hb[K];
      for (int k = lo[2]; k < hi[2]; k++)
        for (int j = lo[1]; j < hi[1]; j++)
          for (int i = lo[0]; i < hi[0]; i++) {
            const size_t c = cidx(s, i, j, k);
            const int ci[3] = {i, j, k};
            const double* p = pm + 4 * ci[J];
            const double t = pb[c];
            if (hkb) {
              const double a = p[1] * e / dx * t;
              hkb[c] += a;
              hkb[c - sJ] -= a;
            }
            pb[c] = p[0] * t + dt * p[3] * db[c];
          }
    }
  }
  /* D curl: D_I += dt e F/dx (H_K[c] - H_K[c-sJ]) */
  for (int I = 2; I >= 0; I--) {
    if (!s->actE[I]) continue;
    const int js[2] = {D1[I], D2[I]};
    for (int t = 0; t < 2; t++) {
      const int Jx = js[t], K = third(I, Jx);
      if (!(s->actH[K] && s->axes[Jx])) continue;
      const size_t sJ = stride(s, Jx);
      const int e = lc(I, Jx, K);
      for (int k = lo[2]; k < hi[2]; k++)
        for (int j = lo[1]; j < hi[1]; j++)
          for (int i = lo[0]; i < hi[0]; i++) {
            const size_t c = cidx(s, i, j, k);
            const int ci[3] = {i, j, k};
            const double a = dt * e * s->pml[0][Jx][4 * ci[Jx] + 2] / dx * s->Db[I][c];
            s->Hb[K][c] += a;
            s->Hb[K][c - sJ] -= a;
          }
    }
  }
}

void rc_adjoint_step(rc_sim* s, long n) { adj_step(s, n, NULL, NULL, NULL); }

void rc_adjoint_finish(rc_sim* s) {
  int lo[3], hi[3];
  interior(s, lo, hi);
  const double mu0 = rc_mu0_v();
  for (int k = lo[2]; k < hi[2]; k++)
    for (int j = lo[1]; j < hi[1]; j++)
      for (int i = lo[0]; i < hi[0]; i++) {
        const size_t c = cidx(s, i, j, k);
        for (int I = 0; I < 3; I++)
          if (s->actH[I]) s->Bb[I][c] += s->Hb[I][c] / (s->mat[s->matid[c]].mu * mu0);
      }
  for (int I = 0; I < 3; I++)
    if (s->actH[I]) memset(s->Hb[I], 0, sizeof(double) * s->ncell);
}

void rc_sync_derived(rc_sim* s) {
  int lo[3], hi[3];
  interior(s, lo, hi);
  const double mu0 = rc_mu0_v();
  for (int k = lo[2]; k < hi[2]; k++)
    for (int j = lo[1]; j < hi[1]; j++)
      for (int i = lo[0]; i < hi[0]; i++) {
        const size_t c = cidx(s, i, j, k);
        for (int I = 0; I < 3; I++) {
          if (s->actE[I]) s->E[I][c] = s->D[I][c] / (cell_eps(s, i, j, k) * RC_EPS0);
          if (s->actH[I]) s->H[I][c] = s->B[I][c] / (s->mat[s->matid[c]].mu * mu0);
        }
      }
}

/* adjoint of D (field 0), B (2), psi E / H (6 / 7, comp 3 I + J), pole P / J
 * (8 / 9, comp 3 p + I) */
static double* adj_ptr(const rc_sim* s, int field, int comp) {
  if (field == 0 && comp >= 0 && comp < 3) return s->Db[comp];
  if (field == 2 && comp >= 0 && comp < 3) return s->Bb[comp];
  if ((field == 6 || field == 7) && comp >= 0 && comp < 9) return s->Pb[field - 6][comp / 3][comp % 3];
  if ((field == 8 || field == 9) && comp >= 0 && comp < 3 * RC_MAXPOLES)
    return field == 8 ? s->PPb[comp / 3][comp % 3] : s->JPb[comp / 3][comp % 3];
  return NULL;
}

int rc_get_adjoint(const rc_sim* s, int field, int comp, double* out) {
  const double* p = adj_ptr(s, field, comp);
  if (!p) return -1;
  memcpy(out, p, sizeof(double) * s->ncell);
  return 0;
}

int rc_set_adjoint(rc_sim* s, int field, int comp, const double* in) {
  double* p = adj_ptr(s, field, comp);
  if (!p) return -1;
  memcpy(p, in, sizeof(double) * s->ncell);
  return 0;
}

int rc_adjoint_eps_gradient(rc_sim* s, int nsteps, int dft, int comp_mask, double* objective, double* grad) {
  if (!s->has_design || s->step != 0 || nsteps < 1 || dft < 0 || dft >= s->ndft) return -1;
  if (comp_mask < 1 || comp_mask > 7 || (comp_mask & ~s->dft[dft].mask)) return -1;
  int lo[3], hi[3];
  interior(s, lo, hi);
  /* pole materials are transposed; magnetized and fluid models are not */
  for (int k = lo[2]; k < hi[2]; k++)
    for (int j = lo[1]; j < hi[1]; j++)
      for (int i = lo[0]; i < hi[0]; i++) {
        const rc_mat* m = &s->mat[s->matid[cidx(s, i, j, k)]];
        if (s->npoles_alloc && m->np > 0 && (m->mag || m->fluid) && design_local(s, i, j, k) < 0) return -1;
        if (s->mpoles_alloc && m->mnp > 0) return -1; /* magnetic poles: not transposed */
      }
  const long nd = (long)(s->dsg[1] - s->dsg[0]) * (s->dsg[3] - s->dsg[2]) * (s->dsg[5] - s->dsg[4]);
  double* hist = calloc((size_t)nsteps * 3 * nd, sizeof(double));
  for (int n = 0; n < nsteps; n++) {
    rc_step(s, 1);
    for (int k = s->dsg[4]; k < s->dsg[5]; k++)
      for (int j = s->dsg[2]; j < s->dsg[3]; j++)
        for (int i = s->dsg[0]; i < s->dsg[1]; i++) {
          const long dl = design_local(s, i, j, k);
          for (int I = 0; I < 3; I++)
            if (s->actE[I]) hist[((size_t)n * 3 + I) * nd + dl] = s->E[I][cidx(s, i, j, k)];
        }
  }
  const rc_dft* d = &s->dft[dft];
  const size_t per = (size_t)d->nfreq * 2 * d->ncells;
  double* w = calloc(3 * per, sizeof(double));
  double J = 0;
  int slot = 0;
  for (int c = 0; c < 3; c++) {
    if (!((comp_mask >> c) & 1)) continue;
    int sm = 0;
    for (int b = 0; b < c; b++) sm += (d->mask >> b) & 1;
    const double* a = d->acc + (size_t)sm * per;
    for (size_t x = 0; x < per; x++) {
      J += a[x] * a[x];
      w[(size_t)slot * per + x] = 2.0 * a[x];
    }
    slot++;
  }
  s->obj_dft = dft;
  s->obj_mask = comp_mask;
  rc_adjoint_reset(s);
  memset(grad, 0, sizeof(double) * nd);
  for (long n = nsteps - 1; n >= 0; n--) adj_step(s, n, w, hist + (size_t)n * 3 * nd, grad);
  free(hist);
  free(w);
  *objective = J;
  return 0;
}
