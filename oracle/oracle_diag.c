/*
 * oracle_diag.c -- CPU restatement of the energy / enstrophy diagnostics.
 *
 * TEST INFRASTRUCTURE ONLY (the oracle): restates diagnostics.c:483-703 for
 * the configurations the timestep path supports (AL81 thickness scheme, no
 * active buoyancy), in the reference's loop order and FP64 association, so
 * that its output equals the reference serial build's EZdiags.dat.
 */
#include "oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define SQ(x) ((x) * (x))
#define P4(x) ((x) * (x) * (x) * (x))

void or_energy(const or_model *m, double *KE, double *PE, double *Z)
{
  const uint g = m->g, nx = m->cfg.Nx, ny = m->cfg.Ny, L = m->L;
  const real dx = m->dx, dy = m->dy;
  double local_KE = 0, local_PE = 0;
  uint i, j, k;
  real *eta = malloc(sizeof(real) * (L + 1));
  for (k = 0; k < L; k++) Z[k] = 0;
  for (k = 0; k < L; k++) {
    const or_field *H = &m->h[k], *U = &m->u[k], *V = &m->v[k];
    for (i = g; i < g + nx; i++) {
      const uint gi = i - g, ip1 = i + 1, im1 = i - 1;
      for (j = g; j < g + ny; j++) {
        const uint gj = j - g, jp1 = j + 1, jm1 = j - 1;
        real h_q, zeta, pv, KEdens, gtot;
        int kk;
        /* interface heights from the work thickness (work.c:1173-1203) */
        eta[L] = OF(&m->hhb_w, i, j);
        for (kk = (int)L - 1; kk >= 0; kk--) eta[kk] = eta[kk + 1] + OF(&m->hw[kk], i, j);
        if (m->cfg.wallNS && gj == 0) {
          h_q = 0.5 * (OF(H, im1, j) + OF(H, i, j));
          zeta = 0;
        } else if (m->cfg.wallEW && gi == 0) {
          h_q = 0.5 * (OF(H, i, jm1) + OF(H, i, j));
          zeta = 0;
        } else {
          h_q = 0.25 * (OF(H, i, j) + OF(H, im1, jm1) + OF(H, im1, j) + OF(H, i, jm1));
          zeta = (OF(U, i, jm1) - OF(U, i, j)) / dy + (OF(V, i, j) - OF(V, im1, j)) / dx;
        }
        if (h_q != 0) {
          pv = (2 * OF(&m->omz, i, j) + zeta) / h_q;
          Z[k] += dx * dy * 0.5 * h_q * SQ(pv);
        }
        KEdens = 0.125 * dx * dy
                 * ((OF(H, i, j) + OF(H, ip1, j)) * SQ(OF(U, ip1, j)) + (OF(H, im1, j) + OF(H, i, j)) * SQ(OF(U, i, j))
                    + (OF(H, i, j) + OF(H, i, jp1)) * SQ(OF(V, i, jp1)) + (OF(H, i, jm1) + OF(H, i, j)) * SQ(OF(V, i, j)));
        gtot = m->geff[k];
        local_KE += KEdens;
        local_PE += dx * dy * gtot * OF(H, i, j) * (eta[k] - 0.5 * OF(H, i, j));
        local_PE += dx * dy * gtot * P4(m->cfg.h0) / SQ(OF(H, i, j)) / 6;
      }
    }
  }
  free(eta);
  *KE = local_KE;
  *PE = local_PE;
}

bool or_write_EZ_if_due(or_model *m, const char *outdir)
{
  char path[1024];
  double KE, PE, *Z;
  uint k;
  if (m->cfg.savefreqEZ <= 0) return true;
  while (m->t >= m->t_next_EZ) {
    FILE *f;
    Z = calloc(m->L, sizeof(double));
    if (!Z) return false;
    or_energy(m, &KE, &PE, Z);
    snprintf(path, sizeof(path), "%s/EZdiags.dat", outdir);
    f = fopen(path, (m->n_EZ > 1) ? "a" : "w");
    if (!f) {
      free(Z);
      return false;
    }
    fprintf(f, "%le", m->t_next_EZ);
    fprintf(f, " %.50e %.50e %.50e ", KE, PE, KE + PE);
    for (k = 0; k < m->L; k++) fprintf(f, "%.50e ", Z[k]);
    fprintf(f, "\n");
    fclose(f);
    free(Z);
    m->n_EZ++;
    m->t_next_EZ = m->n_EZ * m->cfg.savefreqEZ;
  }
  return true;
}
