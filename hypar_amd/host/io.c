/* io.c -- the reference's input files and binary solution format.
 *
 * solver.inp keywords follow src/Simulation/ReadInputs.c:152-430,
 * simulation.inp (nsims) src/Simulation/EnsembleSimulationsDefine.cpp,
 * weno.inp src/InterpolationFunctions/WENOInitialize.c:62-96,
 * boundary.inp src/Simulation/InitializeBoundaries.c:53-120,
 * initial.inp (binary or ascii) src/IOFunctions/ReadArray.c:150-291 with
 * ghost-coordinate extrapolation ReadArray.c:59-90, and op.bin
 * src/IOFunctions/WriteBinary.c:34-100.
 */
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "hypar_host.h"

int hy_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  fprintf(stderr, "ERROR: ");
  vfprintf(stderr, fmt, ap);
  fprintf(stderr, "\n");
  va_end(ap);
  return 1;
}

static FILE *open_in(const char *dir, const char *name, const char *mode) {
  char path[2 * HY_STR];
  snprintf(path, sizeof(path), "%s/%s", dir, name);
  return fopen(path, mode);
}

void hy_member_index(int i, int nsims, char *out) {
  int width = (int)log10((double)nsims) + 1;
  snprintf(out, 32, "%0*d", width, i);
}

int hy_read_nsims(const char *dir, int *nsims) {
  *nsims = 1;
  FILE *in = open_in(dir, "simulation.inp", "r");
  if (!in) return 0;
  char word[HY_STR];
  if (fscanf(in, "%s", word) != 1 || strcmp(word, "begin")) {
    fclose(in);
    return hy_error("illegal format in simulation.inp");
  }
  while (fscanf(in, "%s", word) == 1 && strcmp(word, "end")) {
    if (!strcmp(word, "nsims")) {
      if (fscanf(in, "%d", nsims) != 1 || *nsims < 1) {
        fclose(in);
        return hy_error("bad nsims in simulation.inp");
      }
    } else {
      char useless[HY_STR];
      if (fscanf(in, "%s", useless) != 1) break;
      printf("Warning: keyword %s in file \"simulation.inp\" with value %s not recognized. Ignoring.\n",
             word, useless);
    }
  }
  fclose(in);
  return 0;
}

static void defaults(HyPar *s) {
  memset(s, 0, sizeof(*s));
  s->ndims = 1;
  s->nvars = 1;
  s->ghosts = 1;
  s->n_iter = 0;
  s->restart_iter = 0;
  s->screen_op_iter = 1;
  s->file_op_iter = 1000;
  s->dt = 0.0;
  strcpy(s->time_scheme, "euler");
  strcpy(s->time_scheme_type, " ");
  strcpy(s->spatial_scheme_hyp, "1");
  strcpy(s->interp_type, "characteristic");
  strcpy(s->spatial_type_par, "nonconservative-1stage");
  strcpy(s->spatial_scheme_par, "2");
  strcpy(s->SplitHyperbolicFlux, "no");
  strcpy(s->ConservationCheck, "no");
  strcpy(s->ip_file_type, "ascii");
  strcpy(s->input_mode, "serial");
  strcpy(s->output_mode, "serial");
  strcpy(s->op_file_format, "text");
  strcpy(s->op_overwrite, "no");
  strcpy(s->model, "none");
  strcpy(s->op_fname_root, "op");
}

int hy_read_solver(const char *dir, HyPar *sims, int nsims) {
  for (int n = 0; n < nsims; n++) defaults(&sims[n]);
  FILE *in = open_in(dir, "solver.inp", "r");
  if (!in) return hy_error("file \"solver.inp\" not found");
  HyPar *s = &sims[0];
  char word[HY_STR];
  int ferr = fscanf(in, "%s", word);
  if (ferr != 1 || strcmp(word, "begin")) {
    fclose(in);
    return hy_error("illegal format in file \"solver.inp\"");
  }
  int have_ndims = 0, n_io = 0;
  while (1) {
    if (fscanf(in, "%s", word) != 1) {
      fclose(in);
      return hy_error("unexpected end of solver.inp");
    }
    if (!strcmp(word, "end")) break;
    ferr = 1;
    if (!strcmp(word, "ndims")) {
      ferr = fscanf(in, "%d", &s->ndims);
      have_ndims = 1;
    } else if (!strcmp(word, "nvars")) {
      ferr = fscanf(in, "%d", &s->nvars);
    } else if (!strcmp(word, "size") || !strcmp(word, "size_exact")) {
      if (!have_ndims) {
        fclose(in);
        return hy_error("specify ndims before %s in solver.inp", word);
      }
      int is_size = !strcmp(word, "size");
      for (int n = 0; n < nsims && ferr == 1; n++)
        for (int i = 0; i < s->ndims && ferr == 1; i++) {
          int v;
          ferr = fscanf(in, "%d", &v);
          if (is_size && i < 3) sims[n].dim_global[i] = v;
        }
    } else if (!strcmp(word, "iproc")) {
      if (!have_ndims) {
        fclose(in);
        return hy_error("specify ndims before iproc in solver.inp");
      }
      for (int n = 0; n < nsims && ferr == 1; n++)
        for (int i = 0; i < s->ndims && ferr == 1; i++) {
          int v;
          ferr = fscanf(in, "%d", &v);
          if (ferr == 1 && v != 1) {
            fclose(in);
            return hy_error("iproc must be 1 in every dimension: one GPU holds one domain "
                            "(domain decomposition across GPUs is not built)");
          }
        }
    } else if (!strcmp(word, "ghost")) {
      ferr = fscanf(in, "%d", &s->ghosts);
    } else if (!strcmp(word, "n_iter")) {
      ferr = fscanf(in, "%d", &s->n_iter);
    } else if (!strcmp(word, "restart_iter")) {
      ferr = fscanf(in, "%d", &s->restart_iter);
    } else if (!strcmp(word, "time_scheme")) {
      ferr = fscanf(in, "%s", s->time_scheme);
    } else if (!strcmp(word, "time_scheme_type")) {
      ferr = fscanf(in, "%s", s->time_scheme_type);
    } else if (!strcmp(word, "hyp_space_scheme")) {
      ferr = fscanf(in, "%s", s->spatial_scheme_hyp);
    } else if (!strcmp(word, "hyp_flux_split")) {
      ferr = fscanf(in, "%s", s->SplitHyperbolicFlux);
    } else if (!strcmp(word, "hyp_interp_type")) {
      ferr = fscanf(in, "%s", s->interp_type);
    } else if (!strcmp(word, "par_space_type")) {
      ferr = fscanf(in, "%s", s->spatial_type_par);
    } else if (!strcmp(word, "par_space_scheme")) {
      ferr = fscanf(in, "%s", s->spatial_scheme_par);
    } else if (!strcmp(word, "dt")) {
      ferr = fscanf(in, "%lf", &s->dt);
    } else if (!strcmp(word, "conservation_check")) {
      ferr = fscanf(in, "%s", s->ConservationCheck);
    } else if (!strcmp(word, "screen_op_iter")) {
      ferr = fscanf(in, "%d", &s->screen_op_iter);
    } else if (!strcmp(word, "file_op_iter")) {
      ferr = fscanf(in, "%d", &s->file_op_iter);
    } else if (!strcmp(word, "op_file_format")) {
      ferr = fscanf(in, "%s", s->op_file_format);
    } else if (!strcmp(word, "ip_file_type")) {
      ferr = fscanf(in, "%s", s->ip_file_type);
    } else if (!strcmp(word, "input_mode")) {
      ferr = fscanf(in, "%s", s->input_mode);
      if (ferr == 1 && strcmp(s->input_mode, "serial")) ferr = fscanf(in, "%d", &n_io);
    } else if (!strcmp(word, "output_mode")) {
      ferr = fscanf(in, "%s", s->output_mode);
      if (ferr == 1 && strcmp(s->output_mode, "serial")) ferr = fscanf(in, "%d", &n_io);
    } else if (!strcmp(word, "op_overwrite")) {
      ferr = fscanf(in, "%s", s->op_overwrite);
    } else if (!strcmp(word, "model")) {
      ferr = fscanf(in, "%s", s->model);
    } else if (!strcmp(word, "plot_solution") || !strcmp(word, "immersed_body") ||
               !strcmp(word, "use_gpu") || !strcmp(word, "gpu_device_no")) {
      char value[HY_STR];
      ferr = fscanf(in, "%s", value);
      if (!strcmp(word, "immersed_body") && strcmp(value, "none")) {
        fclose(in);
        return hy_error("immersed bodies are not supported by the GPU path");
      }
    } else {
      char useless[HY_STR];
      ferr = fscanf(in, "%s", useless);
      printf("Warning: keyword %s in file \"solver.inp\" with value %s not recognized or extraneous. "
             "Ignoring.\n", word, useless);
    }
    if (ferr != 1) {
      fclose(in);
      return hy_error("read error in solver.inp near keyword %s", word);
    }
  }
  fclose(in);
  if (s->screen_op_iter <= 0) s->screen_op_iter = 1;
  if (s->file_op_iter <= 0) s->file_op_iter = s->n_iter;
  /* every member shares the scalar inputs (ReadInputs.c broadcasts them) */
  for (int n = 1; n < nsims; n++) {
    int dims[3];
    memcpy(dims, sims[n].dim_global, sizeof(dims));
    sims[n] = sims[0];
    memcpy(sims[n].dim_global, dims, sizeof(dims));
  }
  for (int n = 0; n < nsims; n++) memcpy(sims[n].dim_local, sims[n].dim_global, sizeof(int) * 3);
  return 0;
}

int hy_read_weno(const char *dir, WENOParameters *w) {
  w->mapped = w->borges = w->yc = w->no_limiting = 0;
  w->eps = 1e-6;
  w->p = 2.0;
  FILE *in = open_in(dir, "weno.inp", "r");
  if (!in) {
    printf("Warning: File weno.inp not found. Using default parameters for WENO5.\n");
    return 0;
  }
  char word[HY_STR];
  if (fscanf(in, "%s", word) != 1 || strcmp(word, "begin")) {
    fclose(in);
    return hy_error("illegal format in file \"weno.inp\"");
  }
  while (fscanf(in, "%s", word) == 1 && strcmp(word, "end")) {
    int ferr = 1;
    if (!strcmp(word, "mapped")) ferr = fscanf(in, "%d", &w->mapped);
    else if (!strcmp(word, "borges")) ferr = fscanf(in, "%d", &w->borges);
    else if (!strcmp(word, "yc")) ferr = fscanf(in, "%d", &w->yc);
    else if (!strcmp(word, "no_limiting")) ferr = fscanf(in, "%d", &w->no_limiting);
    else if (!strcmp(word, "epsilon")) ferr = fscanf(in, "%lf", &w->eps);
    else if (!strcmp(word, "p")) ferr = fscanf(in, "%lf", &w->p);
    else if (!strcmp(word, "rc") || !strcmp(word, "xi") || !strcmp(word, "tol")) {
      double d;
      ferr = fscanf(in, "%lf", &d);
    } else {
      char useless[HY_STR];
      ferr = fscanf(in, "%s", useless);
      printf("Warning: keyword %s in file \"weno.inp\" with value %s not recognized or extraneous. "
             "Ignoring.\n", word, useless);
    }
    if (ferr != 1) {
      fclose(in);
      return hy_error("read error in weno.inp near keyword %s", word);
    }
  }
  fclose(in);
  if (w->p != 2.0) printf("Warning from WENOInitialize(): \"p\" parameter is 2.0. Any other value will be ignored!\n");
  return 0;
}

int hy_read_boundaries(const char *dir, HyPar *sim, int member, int nsims) {
  char name[HY_STR] = "boundary";
  if (nsims > 1) {
    char idx[32];
    hy_member_index(member, nsims, idx);
    strcat(name, "_");
    strcat(name, idx);
  }
  strcat(name, ".inp");
  FILE *in = open_in(dir, name, "r");
  if (!in) in = open_in(dir, "boundary.inp", "r");
  if (!in) return hy_error("boundary condition file %s or boundary.inp not found", name);
  int nb;
  if (fscanf(in, "%d", &nb) != 1 || nb < 0) {
    fclose(in);
    return hy_error("bad boundary.inp");
  }
  int faces[3][2] = {{0, 0}, {0, 0}, {0, 0}};
  sim->all_periodic = 1;
  for (int b = 0; b < nb; b++) {
    char type[HY_STR];
    int d, face;
    double lim[6];
    if (fscanf(in, "%s %d %d", type, &d, &face) != 3) {
      fclose(in);
      return hy_error("bad boundary.inp entry %d", b);
    }
    for (int k = 0; k < 2 * sim->ndims; k++)
      if (fscanf(in, "%lf", &lim[k < 6 ? k : 5]) != 1) {
        fclose(in);
        return hy_error("bad boundary.inp extents, entry %d", b);
      }
    if (strcmp(type, "periodic")) sim->all_periodic = 0;
    if (d >= 0 && d < 3) faces[d][face > 0 ? 0 : 1] = 1;
  }
  fclose(in);
  sim->nBoundaryZones = nb;
  if (!sim->all_periodic)
    return hy_error("only periodic boundaries are supported by the GPU path");
  for (int d = 0; d < sim->ndims; d++)
    if (!faces[d][0] || !faces[d][1]) return hy_error("dimension %d lacks a periodic face pair", d);
  return 0;
}

/* ghost coordinates by linear extrapolation (ReadArray.c:59-90) */
static void extrapolate_x(double *X, int n, int g) {
  for (int i = 0; i < g; i++) {
    int delta = g - i;
    X[i] = X[g] + ((double)delta) * (X[g] - X[g + 1]);
  }
  for (int i = n + g; i < n + 2 * g; i++) {
    int delta = i - (n + g - 1);
    X[i] = X[n + g - 1] + ((double)delta) * (X[n + g - 1] - X[n + g - 2]);
  }
}

int hy_read_initial(const char *dir, HyPar *sim, int member, int nsims, double **u_out) {
  char name[HY_STR] = "initial";
  if (nsims > 1) {
    char idx[32];
    hy_member_index(member, nsims, idx);
    strcat(name, "_");
    strcat(name, idx);
  }
  strcat(name, ".inp");
  const int g = sim->ghosts, *n = sim->dim_global;
  const long npts = (long)n[0] * n[1] * n[2];
  const int nv = sim->nvars;
  double *xg = (double *)malloc(sizeof(double) * (n[0] + n[1] + n[2]));
  double *ug = (double *)malloc(sizeof(double) * npts * nv);
  if (!xg || !ug) {
    free(xg);
    free(ug);
    return hy_error("out of host memory reading %s", name);
  }
  int binary = !strcmp(sim->ip_file_type, "bin") || !strcmp(sim->ip_file_type, "binary");
  FILE *in = open_in(dir, name, binary ? "rb" : "r");
  if (!in) {
    free(xg);
    free(ug);
    return hy_error("initial solution file %s not found", name);
  }
  int ok = 1;
  long nx = n[0] + n[1] + n[2];
  if (binary) {
    ok = fread(xg, sizeof(double), nx, in) == (size_t)nx &&
         fread(ug, sizeof(double), npts * nv, in) == (size_t)(npts * nv);
  } else {
    for (long i = 0; i < nx && ok; i++) ok = fscanf(in, "%lf", &xg[i]) == 1;
    /* ascii: variable-major, x fastest (ReadArray.c:198-211) */
    for (int v = 0; v < nv && ok; v++)
      for (long p = 0; p < npts && ok; p++) ok = fscanf(in, "%lf", &ug[p * nv + v]) == 1;
  }
  fclose(in);
  if (!ok) {
    free(xg);
    free(ug);
    return hy_error("could not read %s", name);
  }
  /* coordinates with ghosts */
  long off = 0, offg = 0;
  for (int d = 0; d < 3; d++) {
    for (int i = 0; i < n[d]; i++) sim->x[offg + g + i] = xg[off + i];
    extrapolate_x(sim->x + offg, n[d], g);
    off += n[d];
    offg += n[d] + 2 * g;
  }
  /* solution into the reference layout (AoS with ghosts) */
  const long sx = n[0] + 2 * g, sy = n[1] + 2 * g, sz = n[2] + 2 * g;
  double *u = (double *)calloc((size_t)sx * sy * sz * nv, sizeof(double));
  if (!u) {
    free(xg);
    free(ug);
    return hy_error("out of host memory");
  }
  for (int k = 0; k < n[2]; k++)
    for (int j = 0; j < n[1]; j++)
      for (int i = 0; i < n[0]; i++) {
        long p = ((long)(k + g) * sy + (j + g)) * sx + (i + g);
        long q = ((long)k * n[1] + j) * n[0] + i;
        for (int v = 0; v < nv; v++) u[p * nv + v] = ug[q * nv + v];
      }
  free(xg);
  free(ug);
  *u_out = u;
  return 0;
}

int hy_write_binary(const char *fname, HyPar *sim, const double *u) {
  const int g = sim->ghosts, *n = sim->dim_global, nv = sim->nvars;
  FILE *out = fopen(fname, "wb");
  if (!out) return hy_error("could not open %s for writing", fname);
  int ndims = 3;
  fwrite(&ndims, sizeof(int), 1, out);
  fwrite(&nv, sizeof(int), 1, out);
  fwrite(n, sizeof(int), 3, out);
  long offg = 0;
  for (int d = 0; d < 3; d++) {
    fwrite(sim->x + offg + g, sizeof(double), n[d], out);
    offg += n[d] + 2 * g;
  }
  const long sx = n[0] + 2 * g, sy = n[1] + 2 * g;
  for (int k = 0; k < n[2]; k++)
    for (int j = 0; j < n[1]; j++) {
      long p = ((long)(k + g) * sy + (j + g)) * sx + g;
      fwrite(u + p * nv, sizeof(double), (size_t)n[0] * nv, out);
    }
  fclose(out);
  return 0;
}
