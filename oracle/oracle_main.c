/*
 * oracle_main.c -- driver of the CPU oracle: `pawsim_oracle <input> <outdir>`.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Mirrors pawsim.c main (355-528): preliminary rigid-lid correction and
 * rebuild (266-304), the n=0 output, then step -> correction -> scheduled
 * output (318-350, 471-511).  Prints the pressure lines of pawsim.c:76-113
 * (wall times are this program's own) plus one "ORACLE timing" line.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <time.h>

static double now_s(void)
{
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

static void print_solve(or_model *m, const char *phase, uint n, uint iters, double secs)
{
  if (!m->cfg.pressureTiming) return;
  printf("PAWSIM pressure solve: phase=%s n=%u solver=%s iters=%u walltime=%0.6f s residual_max=%0.6e "
         "residual_rms=%0.6e div_max=%0.6e div_rms=%0.6e\n",
         phase, n, m->cfg.solver == OR_SOLVER_MG ? "MG_GATHERED" : "SOR", iters, secs, m->res_max,
         m->res_rms, m->div_max, m->div_rms);
}

static bool write_scheduled(or_model *m, const char *outdir)
{
  real eps;
  if (m->cfg.savefrequency <= 0) return true;
  eps = 1e-9 * fabs(m->cfg.dt);
  for (;;) {
    real ts = m->n_saves * m->cfg.savefrequency;
    if (m->t + eps < ts) break;
    if (!or_write_state(m, outdir, m->n_saves)) return false;
    m->n_saves++;
  }
  return true;
}

int main(int argc, char **argv)
{
  or_model m;
  uint n, iters;
  double t0, tstep = 0, tall;
  if (argc < 3) {
    fprintf(stderr, "Usage: pawsim_oracle <input_file> <output_dir>\n");
    return 1;
  }
  if (!or_model_init(&m, argv[1], stderr)) {
    fprintf(stderr, "ERROR: Could not initialize oracle state\n");
    return 1;
  }
  /* readParameterData: a restart overwrites u, v, h from the outputs (pawsim.c:251-254) */
  if (m.cfg.restart && !or_model_restart(&m, argv[2], stderr)) {
    fprintf(stderr, "ERROR: Could not read the restart state\n");
    return 1;
  }
  printf("PAWSIM parsed %u x %u x %u AWSIM input\n", m.cfg.Nx, m.cfg.Ny, m.cfg.Nlay);
  tall = now_s();
  /* timeSteppingPreliminaries (pawsim.c:266-304). */
  or_prepare_work(&m);
  or_forcing_update(&m);
  or_tendency_core(&m);
  t0 = now_s();
  iters = or_pressure_correct(&m);
  print_solve(&m, "prelim", 0, iters, now_s() - t0);
  printf("PAWSIM pressure iterations: %u\n", iters);
  or_prepare_work(&m);
  or_forcing_update(&m);
  or_tendency_core(&m);
  if (m.cfg.writeInit && !or_write_init_diagnostics(&m, argv[2])) return 1;
  /* the first state written carries the restart index (pawsim.c:427, diagnostics.c:977-999) */
  m.n_saves = m.cfg.restart ? m.cfg.startIdx : 0;
  if (!or_write_state(&m, argv[2], m.n_saves)) return 1;
  m.n_saves++;
  if (m.cfg.savefreqAvg > 0) {  /* pawsim.c:429-433 */
    m.n_avg = (uint)round(m.cfg.tmin / m.cfg.savefreqAvg) + 1;
    m.t_next_avg = m.cfg.tmin + m.cfg.savefreqAvg;
  }
  if (m.cfg.savefreqEZ > 0) {  /* pawsim.c:434-438 */
    m.n_EZ = (uint)round(m.cfg.tmin / m.cfg.savefreqEZ) + 1;
    m.t_next_EZ = m.cfg.tmin + m.cfg.savefreqEZ;
  }
  for (n = 1; n <= m.cfg.Nt; n++) {
    double ts = now_s();
    if (!or_step(&m, n)) {
      fprintf(stderr, "ERROR: PAWSIM time step failed at n=%u\n", n);
      return 1;
    }
    t0 = now_s();
    iters = or_pressure_correct(&m);
    print_solve(&m, "step", n, iters, now_s() - t0);
    tstep += now_s() - ts;
    if (iters == m.cfg.maxiters) fprintf(stderr, "WARNING: PAWSIM pressure solve reached maxiters at n=%u\n", n);
    if (!write_scheduled(&m, argv[2])) return 1;
    /* pawsim.c:501-503 */
    if (!or_accumulate_averages(&m) || !or_write_averages_if_due(&m, argv[2])) return 1;
    if (!or_write_EZ_if_due(&m, argv[2])) return 1;
  }
  printf("ORACLE timing: steps=%u step_seconds=%0.6f solve_seconds=%0.6f correct_seconds=%0.6f total_seconds=%0.6f mean_iters=%0.2f\n",
         m.cfg.Nt, tstep, m.solve_seconds, m.correct_seconds, now_s() - tall,
         m.solves ? (double)m.iters_total / m.solves : 0.0);
  or_averages_free(&m);
  or_model_free(&m);
  return 0;
}
