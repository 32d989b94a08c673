/*
 * oracle_diag.c -- CPU restatement of the energy / enstrophy diagnostics.
 *
 * TEST INFRASTRUCTURE ONLY (the oracle): restates diagnostics.c:483-703 for
 * the configurations the timestep path supports (every thickness scheme, no
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
        if (m->cfg.thkScheme == OR_THK_HK83) {
          /* diagnostics.c:581-642: HK83 weights u^2 and v^2 over the
           * neighbouring faces along the face, one-sided on wall rows */
          real hE = OF(H, i, j) + OF(H, ip1, j), hW = OF(H, im1, j) + OF(H, i, j);
          real hN = OF(H, i, j) + OF(H, i, jp1), hS = OF(H, i, jm1) + OF(H, i, j);
          KEdens = 0;
          if (m->cfg.wallNS && gj == 0)
            KEdens += hE * (1.25 * SQ(OF(U, ip1, j)) + 0.25 * SQ(OF(U, ip1, jp1)))
                      + hW * (1.25 * SQ(OF(U, i, j)) + 0.25 * SQ(OF(U, i, jp1)));
          else if (m->cfg.wallNS && gj == ny - 1)
            KEdens += hE * (1.25 * SQ(OF(U, ip1, j)) + 0.25 * SQ(OF(U, ip1, jm1)))
                      + hW * (1.25 * SQ(OF(U, i, j)) + 0.25 * SQ(OF(U, i, jm1)));
          else
            KEdens += hE * (SQ(OF(U, ip1, j)) + 0.25 * SQ(OF(U, ip1, jp1)) + 0.25 * SQ(OF(U, ip1, jm1)))
                      + hW * (SQ(OF(U, i, j)) + 0.25 * SQ(OF(U, i, jp1)) + 0.25 * SQ(OF(U, i, jm1)));
          if (m->cfg.wallEW && gi == 0)
            KEdens += hN * (1.25 * SQ(OF(V, i, jp1)) + 0.25 * SQ(OF(V, ip1, jp1)))
                      + hS * (1.25 * SQ(OF(V, i, j)) + 0.25 * SQ(OF(V, ip1, j)));
          else if (m->cfg.wallEW && gi == nx - 1)
            KEdens += hN * (1.25 * SQ(OF(V, i, jp1)) + 0.25 * SQ(OF(V, im1, jp1)))
                      + hS * (1.25 * SQ(OF(V, i, j)) + 0.25 * SQ(OF(V, im1, j)));
          else
            KEdens += hN * (SQ(OF(V, i, jp1)) + 0.25 * SQ(OF(V, ip1, jp1)) + 0.25 * SQ(OF(V, im1, jp1)))
                      + hS * (SQ(OF(V, i, j)) + 0.25 * SQ(OF(V, ip1, j)) + 0.25 * SQ(OF(V, im1, j)));
          KEdens *= dx * dy / 12.0;
        } else {
          /* diagnostics.c:566-579 (AL81, UP3, KT00) */
          KEdens = 0.125 * dx * dy
                   * ((OF(H, i, j) + OF(H, ip1, j)) * SQ(OF(U, ip1, j)) + (OF(H, im1, j) + OF(H, i, j)) * SQ(OF(U, i, j))
                      + (OF(H, i, j) + OF(H, i, jp1)) * SQ(OF(V, i, jp1)) + (OF(H, i, jm1) + OF(H, i, j)) * SQ(OF(V, i, j)));
        }
        gtot = (m->cfg.useTracer && m->cfg.useBuoyancy) ? m->geff[k] - OF(&m->b[k], i, j) : m->geff[k];
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
    f = fopen(path, (m->cfg.restart || m->n_EZ > 1) ? "a" : "w");
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

/* ---- online averages (diagnostics.c:187-470) ---- */

/* Field order: per layer U V H M W HU HV HUU HVV HUV, then W of layer L,
 * then P (write_average_state, diagnostics.c:376-425). */
enum { AV_U, AV_V, AV_H, AV_M, AV_W, AV_HU, AV_HV, AV_HUU, AV_HVV, AV_HUV, AV_PER_LAYER };

/* the B averages (one per layer) follow P when there is a tracer */
static uint av_fields(const or_model *m) { return AV_PER_LAYER * m->L + 2 + (m->cfg.useTracer ? m->L : 0); }

static bool averages_alloc(or_model *m)
{
  uint q, n = av_fields(m);
  if (m->avg) return true;
  m->avg = calloc(n, sizeof(or_field));
  if (!m->avg) return false;
  for (q = 0; q < n; q++)
    if (!or_field_alloc(&m->avg[q], m->cfg.Nx, m->cfg.Ny, m->g)) return false;
  return true;
}

void or_averages_free(or_model *m)
{
  uint q, n = av_fields(m);
  if (!m->avg) return;
  for (q = 0; q < n; q++) or_field_free(&m->avg[q]);
  free(m->avg);
  m->avg = NULL;
}

#define AVG(q, k) (&m->avg[(k) * AV_PER_LAYER + (q)])

/* pawsim_diagnostics_accumulate_averages (diagnostics.c:307-372); no
 * tracer.  W is the interface velocity the last tendency evaluation built
 * (the averages reload the state but do not rebuild wdia). */
bool or_accumulate_averages(or_model *m)
{
  uint i, j, k;
  const uint g = m->g, L = m->L;
  real avg_fac;
  if (m->cfg.savefreqAvg <= 0) return true;
  if (!averages_alloc(m)) return false;
  avg_fac = (m->t < m->t_next_avg) ? 1 : 1 - (m->t - m->t_next_avg) / m->cfg.dt;
  or_average_work(m);
  for (k = 0; k < L; k++) {
    const or_field *U = &m->u[k], *V = &m->v[k], *H = &m->h[k];
    for (i = g; i < g + m->cfg.Nx; i++)
      for (j = g; j < g + m->cfg.Ny; j++) {
        const uint im1 = i - 1, jm1 = j - 1;
        const real dtfac = avg_fac * m->cfg.dt;
        const real hq = 0.25 * (OF(H, i, j) + OF(H, im1, j) + OF(H, i, jm1) + OF(H, im1, jm1));
        const real uq = 0.5 * (OF(U, i, j) + OF(U, i, jm1));
        const real vq = 0.5 * (OF(V, i, j) + OF(V, im1, j));
        OF(AVG(AV_U, k), i, j) += dtfac * OF(U, i, j);
        OF(AVG(AV_V, k), i, j) += dtfac * OF(V, i, j);
        OF(AVG(AV_H, k), i, j) += dtfac * OF(H, i, j);
        OF(AVG(AV_M, k), i, j) += dtfac * (OF(&m->MM[k], i, j) + (m->cfg.useRL ? OF(&m->pi, i, j) : 0));
        OF(AVG(AV_W, k), i, j) += dtfac * OF(&m->wdia[k], i, j);
        OF(AVG(AV_HU, k), i, j) += dtfac * OF(&m->hwest[k], i, j) * OF(U, i, j);
        OF(AVG(AV_HV, k), i, j) += dtfac * OF(&m->hsouth[k], i, j) * OF(V, i, j);
        OF(AVG(AV_HUU, k), i, j) += dtfac * OF(&m->hwest[k], i, j) * OF(U, i, j) * OF(U, i, j);
        OF(AVG(AV_HVV, k), i, j) += dtfac * OF(&m->hsouth[k], i, j) * OF(V, i, j) * OF(V, i, j);
        OF(AVG(AV_HUV, k), i, j) += dtfac * hq * uq * vq;
        if (m->cfg.useTracer) OF(&m->avg[AV_PER_LAYER * L + 2 + k], i, j) += dtfac * OF(&m->b[k], i, j);
      }
  }
  for (i = g; i < g + m->cfg.Nx; i++)
    for (j = g; j < g + m->cfg.Ny; j++) {
      const real dtfac = avg_fac * m->cfg.dt;
      OF(&m->avg[AV_PER_LAYER * L], i, j) += dtfac * OF(&m->wdia[L], i, j);
      OF(&m->avg[AV_PER_LAYER * L + 1], i, j) += dtfac * OF(&m->pi, i, j);
    }
  return true;
}

static bool write_avg(const char *outdir, const char *prefix, int k, uint n, const or_field *f)
{
  char path[1024];
  if (k >= 0) snprintf(path, sizeof(path), "%s/%s%d_n=%u.dat", outdir, prefix, k, n);
  else snprintf(path, sizeof(path), "%s/%s_n=%u.dat", outdir, prefix, n);
  return or_write_field(path, f);
}

/* pawsim_diagnostics_write_averages_if_due (diagnostics.c:428-470). */
bool or_write_averages_if_due(or_model *m, const char *outdir)
{
  static const char *names[AV_PER_LAYER] = {"U_avg", "V_avg", "H_avg", "M_avg", "W_avg",
                                            "HU_avg", "HV_avg", "HUU_avg", "HVV_avg", "HUV_avg"};
  uint q, k, i, j;
  const uint L = m->L, nf = av_fields(m);
  if (m->cfg.savefreqAvg <= 0 || m->t < m->t_next_avg) return true;
  if (!averages_alloc(m)) return false;
  for (q = 0; q < nf; q++) {
    or_field *f = &m->avg[q];
    for (i = f->g; i < f->g + f->nx; i++)
      for (j = f->g; j < f->g + f->ny; j++) OF(f, i, j) *= 1 / m->cfg.savefreqAvg;
  }
  for (k = 0; k < L; k++) {
    for (q = 0; q < AV_PER_LAYER; q++)
      if (!write_avg(outdir, names[q], (int)k, m->n_avg, AVG(q, k))) return false;
    if (m->cfg.useTracer && !write_avg(outdir, "B_avg", (int)k, m->n_avg, &m->avg[AV_PER_LAYER * L + 2 + k]))
      return false;
  }
  if (!write_avg(outdir, "W_avg", (int)L, m->n_avg, &m->avg[AV_PER_LAYER * L])) return false;
  if (m->cfg.useRL && !write_avg(outdir, "P_avg", -1, m->n_avg, &m->avg[AV_PER_LAYER * L + 1])) return false;
  for (q = 0; q < nf; q++) or_field_zero(&m->avg[q]);
  m->n_avg++;
  m->t_next_avg = m->n_avg * m->cfg.savefreqAvg;
  return true;
}
