/*
 * oracle_mg.c -- restatement of the serial multigrid solver (multigrid.c).
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Levels are no-ghost Nx x Ny arrays with element (i,j) at a[i*Ny+j].  The
 * hierarchy coarsens with (N+1)/2 down to 1 x 1 (multigrid.c:128-131,
 * 222-259); the smoother is the in-place lexicographic damped Gauss-Seidel
 * sweep with modular neighbours of multigrid.c:536-568 (its pi_prev copy is
 * never read, so it is not reproduced).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define SQ(x) ((x) * (x))

typedef struct mg_level {
  uint Nx, Ny, pow2; /* pow2: this (fine) level transfers 2:1 to the next coarser */
  real dx, dy;
  real *pi, *tmp, *rhs, *Hc, *Hw, *Hs, *Ow, *Os, *Osum, *rOsum, *wmm, *wmp, *wpm, *wpp;
  int own_pi;
} mg_level;

static mg_level *lv = NULL;
static int nlev = 0;
static bool g_full = false, g_wallEW = false, g_wallNS = false;
static real g_tol = 0, g_omega = 0;
static uint g_maxit = 0;

#define A2(l, a, i, j) ((l)->a[(size_t)(i) * (l)->Ny + (j)])

static uint coarsen(uint N) { return N > 1 ? (N + 1) / 2 : 1; }

/* multigrid.c:599-627 style operators are evaluated with these indices. */
static inline uint im1(uint i, uint N) { return (i + N - 1) % N; }
static inline uint ip1(uint i, uint N) { return (i + N + 1) % N; }

/* restrict_power2_MG / restrict_general_MG (multigrid.c:651-767). */
static void restrict_level(const mg_level *f, const real *pf, const mg_level *c, real *pc)
{
  if (f->pow2) {
    uint i, j;
    for (i = 0; i < f->Nx / 2; i++)
      for (j = 0; j < f->Ny / 2; j++)
        pc[(size_t)i * c->Ny + j] = 0.25 * (pf[(size_t)(2 * i) * f->Ny + 2 * j] + pf[(size_t)(2 * i + 1) * f->Ny + 2 * j]
                                            + pf[(size_t)(2 * i) * f->Ny + 2 * j + 1] + pf[(size_t)(2 * i + 1) * f->Ny + 2 * j + 1]);
  } else {
    uint ic, jc, ifn, jfn;
    memset(pc, 0, (size_t)c->Nx * c->Ny * sizeof(real));
    for (ic = 0; ic < c->Nx; ic++) {
      real x0 = ((real)ic) * f->Nx / c->Nx, x1 = ((real)(ic + 1)) * f->Nx / c->Nx;
      for (jc = 0; jc < c->Ny; jc++) {
        real y0 = ((real)jc) * f->Ny / c->Ny, y1 = ((real)(jc + 1)) * f->Ny / c->Ny, wsum = 0;
        for (ifn = (uint)floor(x0); ifn < (uint)ceil(x1); ifn++) {
          real wx = fmin(x1, (real)(ifn + 1)) - fmax(x0, (real)ifn);
          if (wx <= 0) continue;
          for (jfn = (uint)floor(y0); jfn < (uint)ceil(y1); jfn++) {
            real wy = fmin(y1, (real)(jfn + 1)) - fmax(y0, (real)jfn), w;
            if (wy <= 0) continue;
            w = wx * wy;
            pc[(size_t)ic * c->Ny + jc] += w * pf[(size_t)ifn * f->Ny + jfn];
            wsum += w;
          }
        }
        if (wsum > 0) pc[(size_t)ic * c->Ny + jc] /= wsum;
      }
    }
  }
}

/* interp_index_MG (multigrid.c:794-813). */
static uint interp_index(int idx, uint N, bool wall)
{
  if (N <= 1) return 0;
  if (wall) {
    if (idx < 0) return 0;
    if (idx >= (int)N) return N - 1;
    return (uint)idx;
  }
  return (uint)((idx + (int)N) % (int)N);
}

/* interpolate_power2_MG / interpolate_general_MG (multigrid.c:827-932). */
static void interpolate_level(const mg_level *f, real *pf, const mg_level *c, const real *pc)
{
  uint i, j;
  if (f->pow2) {
    uint hx = f->Nx / 2, hy = f->Ny / 2;
    for (i = 0; i < f->Nx; i++) {
      uint ip = (((i + 1) - (i + 1) % 2) / 2 + hx) % hx;
      uint im = (ip - 1 + hx) % hx;
      for (j = 0; j < f->Ny; j++) {
        uint jp = (((j + 1) - (j + 1) % 2) / 2 + hy) % hy;
        uint jm = (jp - 1 + hy) % hy;
        pf[(size_t)i * f->Ny + j] = A2(f, wmm, i, j) * pc[(size_t)im * c->Ny + jm] + A2(f, wpm, i, j) * pc[(size_t)ip * c->Ny + jm]
                                    + A2(f, wmp, i, j) * pc[(size_t)im * c->Ny + jp] + A2(f, wpp, i, j) * pc[(size_t)ip * c->Ny + jp];
      }
    }
    return;
  }
  for (i = 0; i < f->Nx; i++) {
    real x = (((real)i) + 0.5) * c->Nx / f->Nx - 0.5, tx;
    int imr;
    uint im, ip;
    if (g_wallEW && x <= 0) { imr = 0; tx = 0; }
    else if (g_wallEW && x >= c->Nx - 1) { imr = (int)c->Nx - 1; tx = 0; }
    else { imr = (int)floor(x); tx = x - imr; }
    im = interp_index(imr, c->Nx, g_wallEW);
    ip = interp_index(imr + 1, c->Nx, g_wallEW);
    for (j = 0; j < f->Ny; j++) {
      real y = (((real)j) + 0.5) * c->Ny / f->Ny - 0.5, ty;
      int jmr;
      uint jm, jp;
      if (g_wallNS && y <= 0) { jmr = 0; ty = 0; }
      else if (g_wallNS && y >= c->Ny - 1) { jmr = (int)c->Ny - 1; ty = 0; }
      else { jmr = (int)floor(y); ty = y - jmr; }
      jm = interp_index(jmr, c->Ny, g_wallNS);
      jp = interp_index(jmr + 1, c->Ny, g_wallNS);
      pf[(size_t)i * f->Ny + j] = (1 - tx) * (1 - ty) * pc[(size_t)im * c->Ny + jm] + tx * (1 - ty) * pc[(size_t)ip * c->Ny + jm]
                                  + (1 - tx) * ty * pc[(size_t)im * c->Ny + jp] + tx * ty * pc[(size_t)ip * c->Ny + jp];
    }
  }
}

static real *alloc_level_array(uint Nx, uint Ny) { return calloc((size_t)Nx * Ny, sizeof(real)); }

void or_mg_free(void)
{
  int m;
  if (!lv) return;
  for (m = 0; m < nlev; m++) {
    mg_level *l = &lv[m];
    if (l->own_pi) { free(l->pi); free(l->rhs); }
    free(l->tmp); free(l->Hc); free(l->Hw); free(l->Hs); free(l->Ow); free(l->Os);
    free(l->Osum); free(l->rOsum); free(l->wmm); free(l->wmp); free(l->wpm); free(l->wpp);
  }
  free(lv);
  lv = NULL;
  nlev = 0;
}

/* init_MG (multigrid.c:196-446). */
bool or_mg_init(uint Nx, uint Ny, real Lx, real Ly, const real *Hc, bool wallEW, bool wallNS,
                bool fullMG, real tol, uint maxiters, real omega)
{
  int m;
  uint nx = Nx, ny = Ny, i, j;
  or_mg_free();
  g_full = fullMG; g_tol = tol; g_maxit = maxiters; g_omega = omega;
  g_wallEW = wallEW; g_wallNS = wallNS;
  nlev = 1;
  while (nx > 1 || ny > 1) { nx = coarsen(nx); ny = coarsen(ny); nlev++; }
  lv = calloc((size_t)nlev, sizeof(mg_level));
  nx = Nx; ny = Ny;
  for (m = nlev - 1; m >= 0; m--) {
    lv[m].Nx = nx; lv[m].Ny = ny;
    lv[m].dx = Lx / nx; lv[m].dy = Ly / ny;
    nx = coarsen(nx); ny = coarsen(ny);
  }
  for (m = nlev - 1; m > 0; m--)
    lv[m].pow2 = (lv[m].Nx == 2 * lv[m - 1].Nx) && (lv[m].Ny == 2 * lv[m - 1].Ny);
  for (m = nlev - 1; m >= 0; m--) {
    mg_level *l = &lv[m];
    uint X = l->Nx, Y = l->Ny;
    l->tmp = alloc_level_array(X, Y); l->Hc = alloc_level_array(X, Y);
    l->Hw = alloc_level_array(X, Y); l->Hs = alloc_level_array(X, Y);
    l->Ow = alloc_level_array(X, Y); l->Os = alloc_level_array(X, Y);
    l->Osum = alloc_level_array(X, Y); l->rOsum = alloc_level_array(X, Y);
    l->wmm = alloc_level_array(X, Y); l->wmp = alloc_level_array(X, Y);
    l->wpm = alloc_level_array(X, Y); l->wpp = alloc_level_array(X, Y);
    if (m < nlev - 1) {
      l->pi = alloc_level_array(X, Y);
      l->rhs = alloc_level_array(X, Y);
      l->own_pi = 1;
    }
    if (m == nlev - 1) memcpy(l->Hc, Hc, (size_t)X * Y * sizeof(real));
    else restrict_level(&lv[m + 1], lv[m + 1].Hc, l, l->Hc);
    /* calcFaceThickness_MG: modular face averages (multigrid.c:157-173). */
    for (i = 0; i < X; i++)
      for (j = 0; j < Y; j++) {
        A2(l, Hw, i, j) = 0.5 * (A2(l, Hc, i, j) + A2(l, Hc, im1(i, X), j));
        A2(l, Hs, i, j) = 0.5 * (A2(l, Hc, i, j) + A2(l, Hc, i, im1(j, Y)));
      }
    for (i = 0; i < X; i++) {
      for (j = 0; j < Y; j++) A2(l, Os, i, j) = A2(l, Hs, i, j) / SQ(l->dy);
      if (wallNS) A2(l, Os, i, 0) = 0;
    }
    for (j = 0; j < Y; j++) {
      for (i = 0; i < X; i++) A2(l, Ow, i, j) = A2(l, Hw, i, j) / SQ(l->dx);
      if (wallEW) A2(l, Ow, 0, j) = 0;
    }
    for (i = 0; i < X; i++)
      for (j = 0; j < Y; j++) {
        A2(l, Osum, i, j) = A2(l, Ow, i, j) + A2(l, Ow, ip1(i, X), j) + A2(l, Os, i, j) + A2(l, Os, i, ip1(j, Y));
        A2(l, rOsum, i, j) = 1 / A2(l, Osum, i, j);
      }
    for (i = 0; i < X; i++)
      for (j = 0; j < Y; j++) {
        A2(l, wmm, i, j) = ((i % 2 == 1) ? 0.75 : 0.25) * ((j % 2 == 1) ? 0.75 : 0.25);
        A2(l, wmp, i, j) = ((i % 2 == 1) ? 0.75 : 0.25) * ((j % 2 == 1) ? 0.25 : 0.75);
        A2(l, wpm, i, j) = ((i % 2 == 1) ? 0.25 : 0.75) * ((j % 2 == 1) ? 0.75 : 0.25);
        A2(l, wpp, i, j) = ((i % 2 == 1) ? 0.25 : 0.75) * ((j % 2 == 1) ? 0.25 : 0.75);
      }
    if (wallNS)
      for (i = 0; i < X; i++) {
        A2(l, wmm, i, 0) = 0; A2(l, wpm, i, 0) = 0;
        A2(l, wmp, i, 0) = (i % 2 == 1) ? 0.75 : 0.25;
        A2(l, wpp, i, 0) = (i % 2 == 1) ? 0.25 : 0.75;
        A2(l, wmp, i, Y - 1) = 0; A2(l, wpp, i, Y - 1) = 0;
        A2(l, wmm, i, Y - 1) = (i % 2 == 1) ? 0.75 : 0.25;
        A2(l, wpm, i, Y - 1) = (i % 2 == 1) ? 0.25 : 0.75;
      }
    if (wallEW)
      for (j = 0; j < Y; j++) {
        A2(l, wmm, 0, j) = 0; A2(l, wmp, 0, j) = 0;
        A2(l, wpm, 0, j) = (j % 2 == 1) ? 0.75 : 0.25;
        A2(l, wpp, 0, j) = (j % 2 == 1) ? 0.25 : 0.75;
        A2(l, wpm, X - 1, j) = 0; A2(l, wpp, X - 1, j) = 0;
        A2(l, wmm, X - 1, j) = (j % 2 == 1) ? 0.75 : 0.25;
        A2(l, wmp, X - 1, j) = (j % 2 == 1) ? 0.25 : 0.75;
      }
    if (wallNS && wallEW) {
      A2(l, wmm, 0, 0) = 0; A2(l, wmp, 0, 0) = 0; A2(l, wpm, 0, 0) = 0; A2(l, wpp, 0, 0) = 1;
      A2(l, wmm, 0, Y - 1) = 0; A2(l, wmp, 0, Y - 1) = 0; A2(l, wpm, 0, Y - 1) = 1; A2(l, wpp, 0, Y - 1) = 0;
      A2(l, wmm, X - 1, 0) = 0; A2(l, wmp, X - 1, 0) = 1; A2(l, wpm, X - 1, 0) = 0; A2(l, wpp, X - 1, 0) = 0;
      A2(l, wmm, X - 1, Y - 1) = 1; A2(l, wmp, X - 1, Y - 1) = 0; A2(l, wpm, X - 1, Y - 1) = 0; A2(l, wpp, X - 1, Y - 1) = 0;
    }
  }
  return true;
}

/* relax_MG (multigrid.c:536-568): in-place lexicographic damped GS. */
static void relax(const mg_level *l, real *pi, const real *rhs)
{
  uint i, j, X = l->Nx, Y = l->Ny;
  for (i = 0; i < X; i++) {
    uint im = im1(i, X), ip = ip1(i, X);
    for (j = 0; j < Y; j++) {
      uint jm = im1(j, Y), jp = ip1(j, Y);
      pi[(size_t)i * Y + j] = g_omega * A2(l, rOsum, i, j)
                                  * (A2(l, Os, i, jp) * pi[(size_t)i * Y + jp] + A2(l, Os, i, j) * pi[(size_t)i * Y + jm]
                                     + A2(l, Ow, ip, j) * pi[(size_t)ip * Y + j] + A2(l, Ow, i, j) * pi[(size_t)im * Y + j]
                                     - rhs[(size_t)i * Y + j])
                              + (1 - g_omega) * pi[(size_t)i * Y + j];
    }
  }
}

/* residual_MG (multigrid.c:599-627): res = L(pi) - rhs. */
static void residual(const mg_level *l, const real *pi, const real *rhs, real *res)
{
  uint i, j, X = l->Nx, Y = l->Ny;
  for (i = 0; i < X; i++) {
    uint im = im1(i, X), ip = ip1(i, X);
    for (j = 0; j < Y; j++) {
      uint jm = im1(j, Y), jp = ip1(j, Y);
      res[(size_t)i * Y + j] = A2(l, Os, i, jp) * pi[(size_t)i * Y + jp] + A2(l, Os, i, j) * pi[(size_t)i * Y + jm]
                               + A2(l, Ow, ip, j) * pi[(size_t)ip * Y + j] + A2(l, Ow, i, j) * pi[(size_t)im * Y + j]
                               - A2(l, Osum, i, j) * pi[(size_t)i * Y + j] - rhs[(size_t)i * Y + j];
    }
  }
}

/* max_residual (multigrid.c:1123-1156). */
static real max_res(const mg_level *l, const real *pi, const real *rhs)
{
  uint i, j, X = l->Nx, Y = l->Ny;
  real mx = 0;
  for (i = 0; i < X; i++) {
    uint im = im1(i, X), ip = ip1(i, X);
    for (j = 0; j < Y; j++) {
      uint jm = im1(j, Y), jp = ip1(j, Y);
      real r = pi[(size_t)i * Y + j]
               - A2(l, rOsum, i, j) * (A2(l, Os, i, jp) * pi[(size_t)i * Y + jp] + A2(l, Os, i, j) * pi[(size_t)i * Y + jm]
                                       + A2(l, Ow, ip, j) * pi[(size_t)ip * Y + j] + A2(l, Ow, i, j) * pi[(size_t)im * Y + j]
                                       - rhs[(size_t)i * Y + j]);
      mx = fmax(mx, fabs(r));
    }
  }
  return mx;
}

/* vcycle_MG (multigrid.c:1023-1096); level 0 (1 x 1) is set to zero by
 * exactSolve_MG (multigrid.c:461-476). */
static void vcycle(int n)
{
  mg_level *l = &lv[n], *c;
  uint i, j;
  if (n == 0) {
    if (l->Nx == 1 && l->Ny == 1) l->pi[0] = 0;
    return;
  }
  c = &lv[n - 1];
  relax(l, l->pi, l->rhs);
  residual(l, l->pi, l->rhs, l->tmp);
  restrict_level(l, l->tmp, c, c->rhs);
  memset(c->pi, 0, (size_t)c->Nx * c->Ny * sizeof(real));
  vcycle(n - 1);
  interpolate_level(l, l->tmp, c, c->pi);
  for (i = 0; i < l->Nx; i++)
    for (j = 0; j < l->Ny; j++) l->pi[(size_t)i * l->Ny + j] -= l->tmp[(size_t)i * l->Ny + j];
  relax(l, l->pi, l->rhs);
}

/* iterate_MG (multigrid.c:1172-1213). */
static uint iterate(int n)
{
  uint it = 0;
  real mr = g_tol + 1;
  while (mr > g_tol && it < g_maxit) {
    vcycle(n);
    mr = max_res(&lv[n], lv[n].pi, lv[n].rhs);
    it++;
  }
  return it;
}

/* solve_MG (multigrid.c:1304-1343) and full_MG (1236-1270). */
uint or_mg_solve(real *pi, real *rhs)
{
  mg_level *f = &lv[nlev - 1];
  uint it = 0;
  int n;
  f->pi = pi;
  f->rhs = rhs;
  if (!g_full) return iterate(nlev - 1);
  for (n = nlev - 1; n > 0; n--) restrict_level(&lv[n], lv[n].rhs, &lv[n - 1], lv[n - 1].rhs);
  vcycle(0);
  for (n = 1; n < nlev; n++) {
    interpolate_level(&lv[n], lv[n].pi, &lv[n - 1], lv[n - 1].pi);
    it = iterate(n);
  }
  return it;
}

/* ---- per-function entry points for the unit tests ---- */
int or_mg_nlevels(void) { return nlev; }

void or_mg_level_size(int m, uint *Nx, uint *Ny)
{
  *Nx = lv[m].Nx;
  *Ny = lv[m].Ny;
}

void or_mg_relax(int m, real *pi, const real *rhs) { relax(&lv[m], pi, rhs); }
void or_mg_residual(int m, const real *pi, const real *rhs, real *res) { residual(&lv[m], pi, rhs, res); }
void or_mg_restrict(int m, const real *fine, real *coarse) { restrict_level(&lv[m], fine, &lv[m - 1], coarse); }
void or_mg_interpolate(int m, const real *coarse, real *fine) { interpolate_level(&lv[m], fine, &lv[m - 1], coarse); }
real or_mg_max_residual(int m, const real *pi, const real *rhs) { return max_res(&lv[m], pi, rhs); }

void or_mg_vcycle(real *pi, real *rhs)
{
  lv[nlev - 1].pi = pi;
  lv[nlev - 1].rhs = rhs;
  vcycle(nlev - 1);
}

void or_mg_operators(int m, real *Ow, real *Os, real *Osum)
{
  size_t n = (size_t)lv[m].Nx * lv[m].Ny;
  memcpy(Ow, lv[m].Ow, n * sizeof(real));
  memcpy(Os, lv[m].Os, n * sizeof(real));
  memcpy(Osum, lv[m].Osum, n * sizeof(real));
}
