/* A reference-side driver over the C ABI, written in C99 as the reference's
 * own code would call it: the reference-side glue (gpu_init, the step, the
 * chargeHop sweep fed by the reference's random stream) around the time loop
 * of S/main26.c:65-159 for a small shear run.  One ran1 stream serves
 * initChains, getNoise and chargeHop in the reference's order; a sweep is
 * given a look-ahead copy of the stream's uniforms and the stream then
 * advances by the number the sweep consumed.
 *
 * Usage: glue_run OUT_FILE.  Writes, per step t: "step t", the positions after
 * the step (N lines "x y z", %.17g), and on MC steps "mc consumed", the sweep
 * records (Ncharges lines of 6 ints) and the charge beads after the sweep. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "smdlt_hip.h"

/* the reference's run parameters (S/Parameters.h, Input.txt, CLI flags) */
static int Nb = 20, Nc = 3, N, Ncharges = 12, flowtype = 1, MStep = 2, tmax = 8;
static double L = 24.0, flowrate = 0.5, dt = 1e-3;
static char spring[100] = "Hookean";
static double epsilon = 0.1, kappab = 0.0, qmax = 2.0, lambda_b = 3.59, lambda_d = 5.0;
static double barrier = 1.0, barrier2 = 1.5;

static double *rx, *ry, *rz, *R;
static int *Charge_track, *Charge_old;
static smdlt_ctx* g_ctx;
static smdlt_ran1 g_rng; /* ran1's idum and statics, gasdev's pair cache */

static void die(int code) {
  printf("Error: %s\n", smdlt_strerror(code));
  exit(1);
}
#define CALL(x)          \
  do {                   \
    int e_ = (x);        \
    if (e_ < 0) die(e_); \
  } while (0)

static void gpu_init(void) {
  smdlt_params p;
  smdlt_default_params(&p);
  p.nb = Nb;
  p.nc = Nc;
  p.L = L;
  p.flow_type = flowtype;
  p.edot = flowrate;
  p.dt = dt;
  p.spring_fene = strcmp(spring, "FENE") == 0;
  p.epsilon = epsilon;
  p.kappab = kappab;
  p.kappas = 200.0; /* the shear program's bondforce sets kappas = 200 (S/main26.c:544) */
  p.qmax = qmax;
  p.lambda_b = lambda_b;
  p.lambda_d = lambda_d;
  p.barrier = barrier;
  p.barrier2 = barrier2;
  p.pbc_compat = 1; /* the reference's applyPBC/getNID */
  p.precision = SMDLT_PREC_FP64;
  p.mobility_tol = 1e-7;
  CALL(smdlt_create(&p, 1, 0, &g_ctx));
  CALL(smdlt_upload_positions(g_ctx, rx, ry, rz));
  CALL(smdlt_upload_charges(g_ctx, &Ncharges, Charge_track, NULL));
}

/* getNoise(); updateHI(); with the reference's noise R[3N] (S/main26.c:97-107) */
static void updateHI_gpu(unsigned long t) {
  smdlt_gasdev_fill(&g_rng, 3 * N, R);
  CALL(smdlt_step(g_ctx, t, R));
  CALL(smdlt_download_positions(g_ctx, rx, ry, rz));
}

/* chargeHop() (S/main26.c:1926-2390): the sweep reads a look-ahead copy of the
 * stream; the stream then advances by what the sweep consumed */
static int chargeHop_gpu(int* log) {
  const int n = 3 * Ncharges + 8;
  float* u = malloc(sizeof(float) * (size_t)n);
  smdlt_ran1 ahead = g_rng;
  smdlt_ran1_fill(&ahead, n, u);
  int consumed = 0;
  CALL(smdlt_mc_sweep(g_ctx, SMDLT_MC_REFERENCE, u, n, &consumed, log));
  for (int k = 0; k < consumed; ++k) (void)smdlt_ran1_next(&g_rng);
  CALL(smdlt_download_charges(g_ctx, Charge_track, Charge_old));
  free(u);
  return consumed;
}

int main(int argc, char** argv) {
  if (argc < 2) {
    fprintf(stderr, "usage: %s OUT_FILE\n", argv[0]);
    return 2;
  }
  N = Nb * Nc;
  rx = malloc(sizeof(double) * N);
  ry = malloc(sizeof(double) * N);
  rz = malloc(sizeof(double) * N);
  R = malloc(sizeof(double) * 3 * N);
  Charge_track = malloc(sizeof(int) * Ncharges);
  Charge_old = malloc(sizeof(int) * Ncharges);
  int* log = malloc(sizeof(int) * 6 * Ncharges);
  smdlt_ran1_seed(&g_rng, 830062920); /* the shear seed (S/Initialization.c:438) */
  CALL(smdlt_init_chains(Nb, Nc, L, &g_rng, rx, ry, rz));
  CALL(smdlt_init_charges(N, Ncharges, &g_rng, Charge_track));
  gpu_init();
  FILE* out = fopen(argv[1], "w");
  if (!out) return 3;
  fprintf(out, "init\n");
  for (int i = 0; i < N; ++i) fprintf(out, "%.17g %.17g %.17g\n", rx[i], ry[i], rz[i]);
  for (int k = 0; k < Ncharges; ++k) fprintf(out, "%d\n", Charge_track[k]);
  for (unsigned long t = 0; t < (unsigned long)tmax; ++t) {
    updateHI_gpu(t);
    fprintf(out, "step %lu\n", t);
    for (int i = 0; i < N; ++i) fprintf(out, "%.17g %.17g %.17g\n", rx[i], ry[i], rz[i]);
    if (t % (unsigned long)MStep == 0) {
      const int consumed = chargeHop_gpu(log);
      fprintf(out, "mc %d\n", consumed);
      for (int k = 0; k < Ncharges; ++k)
        fprintf(out, "%d %d %d %d %d %d\n", log[6 * k], log[6 * k + 1], log[6 * k + 2], log[6 * k + 3],
                log[6 * k + 4], log[6 * k + 5]);
      for (int k = 0; k < Ncharges; ++k) fprintf(out, "%d\n", Charge_track[k]);
    }
  }
  fclose(out);
  smdlt_destroy(g_ctx);
  printf("glue_run ok\n");
  return 0;
}
