/* main.c -- hypar_mi355x: runs a HyPar input deck on one MI355X.
 *
 * Same inputs and outputs as the reference driver for this path
 * (src/main.cpp:236-448, src/Simulation/Solve.cpp:37-307): reads
 * solver.inp / physics.inp / weno.inp / boundary.inp / initial.inp (and
 * simulation.inp + per-member files for ensembles) from the deck directory
 * and writes op.bin (op_<i>.bin per member) in the reference's binary format.
 *
 *   hypar_mi355x [--slots] [deck_dir]
 *
 * --slots composes every time step from the individual operator slots
 * (TimeRK -> RHSFunction -> HyperbolicFunction -> FFunction/Interpolate/
 * Upwind ...) instead of the fused RK4 step.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>

#include "hypar_host.h"

static double now(void) {
  struct timeval tv;
  gettimeofday(&tv, NULL);
  return tv.tv_sec + 1e-6 * tv.tv_usec;
}

static int output_solution(const char *dir, HyPar *sims, int nsims) {
  for (int n = 0; n < nsims; n++) {
    HyPar *s = &sims[n];
    if (!strcmp(s->op_file_format, "none")) continue;
    if (strcmp(s->op_file_format, "binary") && strcmp(s->op_file_format, "bin"))
      return hy_error("op_file_format %s is not supported (binary)", s->op_file_format);
    char root[HY_STR], fname[2 * HY_STR + 64];
    strcpy(root, s->op_fname_root);
    if (nsims > 1) {
      char idx[32];
      hy_member_index(n, nsims, idx);
      strcat(root, "_");
      strcat(root, idx);
    }
    if (!strcmp(s->op_overwrite, "no")) {
      snprintf(fname, sizeof(fname), "%s/%s_%05d.bin", dir, root, s->file_index);
      s->file_index++;
    } else {
      snprintf(fname, sizeof(fname), "%s/%s.bin", dir, root);
    }
    const long g = s->ghosts;
    size_t nh = (size_t)5 * (s->dim_global[0] + 2 * g) * (s->dim_global[1] + 2 * g) * (s->dim_global[2] + 2 * g);
    double *uh = (double *)malloc(nh * sizeof(double));
    if (!uh) return hy_error("out of host memory");
    if (hm_download(s->dev, uh, s->u, HM_LAYOUT_REFERENCE) != HM_OK) {
      free(uh);
      return hy_error("hm_download: %s", hm_last_error());
    }
    printf("Writing solution file %s.\n", fname);
    int ierr = hy_write_binary(fname, s, uh);
    free(uh);
    if (ierr) return ierr;
  }
  return 0;
}

int main(int argc, char **argv) {
  const double t_main = now();
  const char *dir = ".";
  int slots = 0;
  for (int i = 1; i < argc; i++) {
    if (!strcmp(argv[i], "--slots")) slots = 1;
    else dir = argv[i];
  }
  printf("HyPar (MI355X gfx950 path: NS3D WENO5 + RK)\n");
  int nsims = 1;
  if (hy_read_nsims(dir, &nsims)) return 1;
  if (nsims > 1) printf("-- Ensemble Simulation --\nNumber of simulation domains: %d\n", nsims);
  HyPar *sims = (HyPar *)calloc(nsims, sizeof(HyPar));
  MPIVariables *mpi = (MPIVariables *)calloc(nsims, sizeof(MPIVariables));
  if (!sims || !mpi) return hy_error("out of host memory");
  for (int n = 0; n < nsims; n++) {
    mpi[n].nproc = 1;
    for (int d = 0; d < 3; d++) mpi[n].iproc[d] = 1;
  }
  if (hy_read_solver(dir, sims, nsims)) return 1;
  for (int n = 0; n < nsims; n++) {
    HyPar *s = &sims[n];
    double *uh = NULL;
    if (hy_initialize(s)) return 1;
    if (hy_read_initial(dir, s, n, nsims, &uh)) return 1;
    hy_compute_metrics(s);
    if (hy_read_boundaries(dir, s, n, nsims)) return 1;
    if (hy_initialize_solvers(s, dir)) return 1;
    if (NavierStokes3DInitialize(s, &mpi[n], dir)) return 1;
    if (hm_upload(s->dev, s->u, uh, HM_LAYOUT_REFERENCE) != HM_OK)
      return hy_error("hm_upload: %s", hm_last_error());
    free(uh);
    printf("Domain %d: %d x %d x %d, %s WENO5, %s upwinding, %s\n", n, s->dim_global[0],
           s->dim_global[1], s->dim_global[2], s->interp_type, s->physics->upw_choice,
           s->physics->Re > 0 ? "viscous" : "inviscid");
    /* InitialSolution.c:117-135 and Initialize.c:304-307 */
    if (VolumeIntegral(s->VolumeIntegralInitial, s->u, s, &mpi[n])) return 1;
    if (nsims > 1) printf("Volume integral of the initial solution on domain %d:\n", n);
    else printf("Volume integral of the initial solution:\n");
    for (int v = 0; v < s->nvars; v++) printf("%2d:  %1.16E\n", v, s->VolumeIntegralInitial[v]);
    for (int v = 0; v < s->nvars; v++) {
      s->TotalBoundaryIntegral[v] = 0.0;
      s->ConservationError[v] = -1.0;
    }
  }
  TimeIntegration TS;
  printf("Setting up time integration.\n");
  if (TimeInitialize(&TS, sims, mpi, nsims, !slots)) return 1;
  printf("Time integration: %s %s (%s%s)\n", sims[0].time_scheme, sims[0].time_scheme_type,
         TS.fused ? (TS.ssp ? "fused SSPRK3 step" : "fused 3N RK4 step")
                  : (TS.RHSFunction == hm_slot_RHSFunction ? "stage loop around the fused right-hand side"
                                                           : "slot composition"),
         TS.ens ? ", batched over all members" : "");
  printf("Solving in time (from %d to %d iterations)\n", sims[0].restart_iter, TS.n_iter);
  double total = 0.0;
  for (TS.iter = sims[0].restart_iter; TS.iter < TS.n_iter; TS.iter++) {
    if (!TS.iter && output_solution(dir, sims, nsims)) return 1;
    double t0 = now();
    if (TimePreStep(&TS)) return 1;
    if (TS.TimeIntegrate(&TS)) return 1;
    if (TimePostStep(&TS)) return 1;
    for (int n = 0; n < nsims; n++)
      if (hm_synchronize(sims[n].dev) != HM_OK) return hy_error("%s", hm_last_error());
    double wct = now() - t0;
    total += wct;
    if ((TS.iter + 1) % sims[0].screen_op_iter == 0) {
      printf("iter=%7d  t=%1.3E  ", TS.iter + 1, TS.waqt);
      if (TS.max_cfl >= 0) printf("CFL=%1.3E  ", TS.max_cfl);
      printf("norm=%1.4E  wctime: %1.1E (s)\n", TS.norm, wct);
    }
    if ((TS.iter + 1) % sims[0].file_op_iter == 0 && (TS.iter + 1) < TS.n_iter &&
        output_solution(dir, sims, nsims))
      return 1;
  }
  printf("Completed time integration (Final time: %f), total wctime: %f (seconds).\n", TS.waqt, total);
  if (output_solution(dir, sims, nsims)) return 1;
  if (SimulationWriteErrors(dir, sims, nsims, total, now() - t_main)) return 1;
  TimeCleanup(&TS);
  for (int n = 0; n < nsims; n++) hy_cleanup(&sims[n]);
  free(sims);
  free(mpi);
  printf("Finished.\n");
  return 0;
}
