/* Plain-C restatement of one oracle.fdtd.Engine.step for real fields on a uniform grid without periodic axes
 * (TEST INFRASTRUCTURE ONLY: the multi-core CPU baseline and large-size checks; see oracle/cyee.py).
 *
 * Follows engine.py:630-663 (H and E curls, CPML psi terms in list order), the point sources' correct_E
 * (emission.py:120-127), the pure D-form groups (engine.py:722-742, 769-793) and the DFT sample
 * (engine.py:564-592), in the oracle's operation order, in float64 with no contraction of multiply-adds, so that
 * every field value equals the NumPy oracle's. Loops over the slowest axis run in parallel (OpenMP).
 *
 * Layout: E, H are (3, N0, N1, N2) in C order (k fastest). A box-shaped array (cb, CPML cp/kp/psi) is C order over
 * its box.
 */
#include <math.h>
#include <stddef.h>

typedef struct {
  int lo[3], hi[3];
} cy_box;

typedef struct {
  int c, src, d;    /* updated component, differenced component, CPML axis */
  cy_box r;         /* the term's box */
  const double* b;  /* per index along d (length hi[d] - lo[d]) */
  const double* cp; /* box-shaped */
  const double* kp; /* box-shaped or NULL (kappa = 1) */
  double* psi;      /* box-shaped */
} cy_pml;

typedef struct {
  long long n;
  const long long* idx;
  double inv_eps_inf;
  int n_poles;
  double coef[8][3];
  double* D;
  double* E_prev;
  double* P; /* [pole][2][n]: P^n then P^{n-1} */
} cy_group;

typedef struct {
  long long n;
  const long long* idx;
  const double* coef;
  const double* wave;
  long long n_wave;
} cy_point;

typedef struct {
  int n_freq;
  long long n_e, n_h;
  const double* omega;
  const long long* e_idx;
  const long long* h1;
  const long long* h2;
  double* acc_e; /* (n_freq, n_e) complex as re, im pairs */
  double* acc_h;
  int energy;    /* 1-based energy slot or 0 */
} cy_monitor;

typedef struct {
  int n[3];
  double* E;
  double* H;
  cy_box rE[3], rH[3];
  const double* cb[3];
  double cH, dt;
  int n_pml_h, n_pml_e, n_groups, n_points, n_monitors;
  const cy_pml* pml_h;
  const cy_pml* pml_e;
  cy_group* groups;
  const cy_point* points;
  cy_monitor* monitors;
  double energy[2];
  long long step;
} cy_engine;

static inline long long at(const cy_engine* e, int c, int i, int j, int k) {
  return (((long long)c * e->n[0] + i) * e->n[1] + j) * e->n[2] + k;
}

static inline long long off(const int* st, int d, int s) { return d == 0 ? s * st[0] : (d == 1 ? s * st[1] : s); }

/* F_out[c][R] -/+= coef * (a - b), forward (H) or backward (E) differences */
static void curl(cy_engine* e, int c, const cy_box* R, int forward, const double* cbox, double cs) {
  const int d1 = (c + 1) % 3, d2 = (c + 2) % 3;
  const int st[3] = {e->n[1] * e->n[2], e->n[2], 1};
  double* F = forward ? e->E : e->H; /* differenced field */
  double* G = forward ? e->H : e->E; /* updated field */
  const long long s1 = off(st, d1, 1), s2 = off(st, d2, 1);
  const int nj = R->hi[1] - R->lo[1], nk = R->hi[2] - R->lo[2];
#pragma omp parallel for schedule(static)
  for (int i = R->lo[0]; i < R->hi[0]; ++i)
    for (int j = R->lo[1]; j < R->hi[1]; ++j) {
      const long long bq = ((long long)(i - R->lo[0]) * nj + (j - R->lo[1])) * nk;
      for (int k = R->lo[2]; k < R->hi[2]; ++k) {
        const long long p2 = at(e, d2, i, j, k) , p1 = at(e, d1, i, j, k), q = at(e, c, i, j, k);
        double a, b;
        if (forward) {
          a = F[p2 + s1] - F[p2];
          b = F[p1 + s2] - F[p1];
        } else {
          a = F[p2] - F[p2 - s1];
          b = F[p1] - F[p1 - s2];
        }
        const double v = a - b;
        if (forward) {
          const double w = v * cs;
          G[q] = G[q] - w;
        } else {
          const double w = v * cbox[bq + (k - R->lo[2])];
          G[q] = G[q] + w;
        }
      }
    }
}

static void pml_term(cy_engine* e, const cy_pml* t, int forward) {
  const int st[3] = {e->n[1] * e->n[2], e->n[2], 1};
  const double* F = forward ? e->E : e->H;
  double* G = forward ? e->H : e->E;
  const long long sd = off(st, t->d, 1);
  const cy_box* R = &t->r;
  const int nj = R->hi[1] - R->lo[1], nk = R->hi[2] - R->lo[2];
#pragma omp parallel for schedule(static)
  for (int i = R->lo[0]; i < R->hi[0]; ++i)
    for (int j = R->lo[1]; j < R->hi[1]; ++j) {
      const long long bq = ((long long)(i - R->lo[0]) * nj + (j - R->lo[1])) * nk;
      for (int k = R->lo[2]; k < R->hi[2]; ++k) {
        const long long q = bq + (k - R->lo[2]);
        const long long ps = at(e, t->src, i, j, k), pc = at(e, t->c, i, j, k);
        const int u = t->d == 0 ? i - R->lo[0] : (t->d == 1 ? j - R->lo[1] : k - R->lo[2]);
        const double der = forward ? F[ps + sd] - F[ps] : F[ps] - F[ps - sd];
        double psi = t->psi[q] * t->b[u];
        if (t->kp) {
          const double w = t->kp[q] * der;
          G[pc] = G[pc] + w;
        }
        const double w = t->cp[q] * der;
        psi = psi + w;
        t->psi[q] = psi;
        G[pc] = G[pc] + psi;
      }
    }
}

static void group_step(cy_engine* e, cy_group* g) {
  const long long n = g->n;
#pragma omp parallel for schedule(static)
  for (long long q = 0; q < n; ++q) {
    const long long p = g->idx[q];
    const double e_now = e->E[p];
    const double ep = g->E_prev[q];
    g->D[q] = g->D[q] + (e_now - ep);
    double tot = 0.0;
    for (int s = 0; s < g->n_poles; ++s) {
      double* P0 = g->P + (2LL * s) * n;
      double* P1 = P0 + n;
      const double t1 = g->coef[s][0] * P0[q];
      const double t2 = g->coef[s][1] * P1[q];
      const double t3 = g->coef[s][2] * ep;
      const double nw = (t1 + t2) + t3;
      P1[q] = P0[q];
      P0[q] = nw;
      tot = s == 0 ? nw : tot + nw;
    }
    const double en = g->n_poles ? (g->D[q] - tot) * g->inv_eps_inf : g->D[q] * g->inv_eps_inf;
    e->E[p] = en;
    g->E_prev[q] = en;
  }
}

static void sample(cy_engine* e) {
  const double t_e = (double)e->step * e->dt, t_h = t_e - 0.5 * e->dt;
  double energy[2] = {0.0, 0.0};
  int has[2] = {0, 0};
  for (int m = 0; m < e->n_monitors; ++m) {
    cy_monitor* M = &e->monitors[m];
    double v = 0.0;
    for (int f = 0; f < M->n_freq; ++f) {
      const double ce = cos(M->omega[f] * t_e), se = sin(M->omega[f] * t_e);
      double* ae = M->acc_e + 2LL * f * M->n_e;
#pragma omp parallel for schedule(static)
      for (long long p = 0; p < M->n_e; ++p) {
        const double x = e->E[M->e_idx[p]];
        ae[2 * p] += ce * x;
        ae[2 * p + 1] += se * x;
      }
      if (M->n_h) {
        const double ch = cos(M->omega[f] * t_h), sh = sin(M->omega[f] * t_h);
        double* ah = M->acc_h + 2LL * f * M->n_h;
#pragma omp parallel for schedule(static)
        for (long long p = 0; p < M->n_h; ++p) {
          const double x = 0.5 * (e->H[M->h1[p]] + e->H[M->h2[p]]);
          ah[2 * p] += ch * x;
          ah[2 * p + 1] += sh * x;
        }
      }
    }
    if (M->energy && M->n_h) {
      for (long long p = 0; p < M->n_e; ++p) v += e->E[M->e_idx[p]] * e->E[M->e_idx[p]];
      for (long long p = 0; p < M->n_h; ++p) {
        const double x = 0.5 * (e->H[M->h1[p]] + e->H[M->h2[p]]);
        v += x * x;
      }
      energy[M->energy - 1] += v;
      has[M->energy - 1] = 1;
    }
  }
  for (int g = 0; g < 2; ++g)
    if (has[g]) e->energy[g] = energy[g];
}

/* n_steps steps; a DFT sample after every `stride` steps counted from step 0 when stride > 0 */
void cy_run(cy_engine* e, long long n_steps, int stride) {
  for (long long s = 0; s < n_steps; ++s) {
    for (int c = 0; c < 3; ++c) curl(e, c, &e->rH[c], 1, NULL, e->cH);
    for (int q = 0; q < e->n_pml_h; ++q) pml_term(e, &e->pml_h[q], 1);
    for (int c = 0; c < 3; ++c) curl(e, c, &e->rE[c], 0, e->cb[c], 0.0);
    for (int q = 0; q < e->n_pml_e; ++q) pml_term(e, &e->pml_e[q], 0);
    for (int q = 0; q < e->n_points; ++q) {
      const cy_point* P = &e->points[q];
      const double w = e->step < P->n_wave ? P->wave[e->step] : 0.0;
      for (long long r = 0; r < P->n; ++r) {
        const double d = P->coef[r] * w;
        e->E[P->idx[r]] = e->E[P->idx[r]] - d;
      }
    }
    for (int g = 0; g < e->n_groups; ++g) group_step(e, &e->groups[g]);
    e->step += 1;
    if (stride > 0 && e->step % stride == 0) sample(e);
  }
}

int cy_abi_version(void) { return 1; }
