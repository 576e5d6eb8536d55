/*
 * rk_driver.c -- the reference's time loop (rk_run / rk_step, euler.c:3610-3699)
 * written in C against include/wadges.h, as a maintainer of the reference
 * driver would call the library: plain C, HIP runtime for memory, no Python.
 *
 *   rk_driver DIR NSTEPS MODE      MODE 0: vol -> surf -> update per stage
 *                                   MODE 1: wadges_stage, face traces swapped
 *
 * DIR holds meta.txt ("dim N Nq Nfq K dt gamma tau") and one raw little-endian
 * file per array (doubles; mapPq uint32).  Writes out_Q.bin, out_rhsQ.bin,
 * out_resQ.bin, out_Qf.bin (the current face traces) into DIR.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <hip/hip_runtime_api.h>

#include "wadges.h"

/* LSRK45 (Carpenter & Kennedy), the coefficients of euler.c:2779-2790 */
static const double rk4a[5] = {0.0, -567301805773.0 / 1357537059087.0,
                               -2404267990393.0 / 2016746695238.0,
                               -3550918686646.0 / 2091501179385.0,
                               -1275806237668.0 / 842570457699.0};
static const double rk4b[5] = {1432997174477.0 / 9575080441755.0,
                               5161836677717.0 / 13612068292357.0,
                               1720146321549.0 / 2090206949498.0,
                               3134564353537.0 / 4481467310338.0,
                               2277821191437.0 / 14882151754819.0};

static const char *g_dir;

static void die(const char *what) {
  fprintf(stderr, "rk_driver: %s\n", what);
  exit(1);
}

static void *upload(const char *name, size_t bytes) {
  char path[4096];
  snprintf(path, sizeof path, "%s/%s.bin", g_dir, name);
  FILE *f = fopen(path, "rb");
  if (!f) die(path);
  void *h = malloc(bytes ? bytes : 1);
  if (!h || fread(h, 1, bytes, f) != bytes) die(path);
  fclose(f);
  void *d = NULL;
  if (hipMalloc(&d, bytes ? bytes : 8) != hipSuccess) die("hipMalloc");
  if (hipMemcpy(d, h, bytes, hipMemcpyHostToDevice) != hipSuccess) die("hipMemcpy");
  free(h);
  return d;
}

static void download(const char *name, const void *d, size_t bytes) {
  char path[4096];
  snprintf(path, sizeof path, "%s/out_%s.bin", g_dir, name);
  void *h = malloc(bytes);
  if (!h || hipMemcpy(h, d, bytes, hipMemcpyDeviceToHost) != hipSuccess) die("copy back");
  FILE *f = fopen(path, "wb");
  if (!f || fwrite(h, 1, bytes, f) != bytes) die(path);
  fclose(f);
  free(h);
}

static void check(int rc, const char *what) {
  if (rc) {
    fprintf(stderr, "rk_driver: %s: %s (hip %d)\n", what, wadges_error_string(rc),
            wadges_last_hip_error());
    exit(1);
  }
}

int main(int argc, char **argv) {
  if (argc != 4) die("usage: rk_driver DIR NSTEPS MODE");
  g_dir = argv[1];
  const int nsteps = atoi(argv[2]), mode = atoi(argv[3]);
  int dim, N, Nq, Nfq, K;
  double dt, gamma, tau;
  char path[4096];
  snprintf(path, sizeof path, "%s/meta.txt", g_dir);
  FILE *mf = fopen(path, "r");
  if (!mf || fscanf(mf, "%d %d %d %d %d %lf %lf %lf", &dim, &N, &Nq, &Nfq, &K, &dt, &gamma,
                    &tau) != 8)
    die(path);
  fclose(mf);
  const int Nf = Nfq * (dim + 1), nf = dim + 2;
  const size_t D = sizeof(double);
  double *vgeo = upload("vgeo", D * K * dim * dim * Nq);
  double *vfgeo = upload("vfgeo", D * K * dim * dim * Nf);
  double *fgeo = upload("fgeo", D * K * (dim + 1) * Nf);
  double *Jq = upload("Jq", D * K * Nq);
  uint32_t *mapPq = upload("mapPq", sizeof(uint32_t) * K * Nf);
  double *Drq = upload("Drq", D * Nq * Nq), *Dsq = upload("Dsq", D * Nq * Nq);
  double *Dtq = dim == 3 ? upload("Dtq", D * Nq * Nq) : NULL;
  double *VqLq = upload("VqLq", D * Nq * Nf), *VqPq = upload("VqPq", D * Nq * Nq);
  double *VfPq = upload("VfPq", D * Nf * Nq);
  double *nrJ = upload("nrJ", D * Nf), *nsJ = upload("nsJ", D * Nf);
  double *ntJ = dim == 3 ? upload("ntJ", D * Nf) : NULL;
  double *Q = upload("Q", D * K * nf * Nq), *rhsQ = upload("rhsQ", D * K * nf * Nq);
  double *resQ = upload("resQ", D * K * nf * Nq), *Qf = upload("Qf", D * K * nf * Nf);
  double *rhsQf = NULL, *QfB = NULL;
  if (hipMalloc((void **)&rhsQf, D * K * nf * Nf) != hipSuccess) die("hipMalloc");
  if (hipMalloc((void **)&QfB, D * K * nf * Nf) != hipSuccess) die("hipMalloc");

  wadges_plan *plan = NULL;
  check(wadges_plan_create(&plan, dim, N, Nq, Nfq, gamma, tau), "wadges_plan_create");
  void *s = NULL; /* default stream */
  for (int step = 0; step < nsteps; ++step)
    for (int st = 0; st < 5; ++st) {
      if (mode == 0) {
        check(wadges_vol(plan, K, vgeo, vfgeo, nrJ, nsJ, ntJ, Drq, Dsq, Dtq, NULL, VqLq, VfPq, Q,
                         Qf, rhsQ, rhsQf, s), "wadges_vol");
        check(wadges_surf(plan, K, fgeo, mapPq, VqLq, Qf, rhsQf, rhsQ, s), "wadges_surf");
        check(wadges_update(plan, K, Jq, VqPq, VfPq, rk4a[st], rk4b[st], dt, rhsQ, resQ, Q, Qf,
                            s), "wadges_update");
      } else {
        check(wadges_stage(plan, K, vgeo, vfgeo, fgeo, mapPq, nrJ, nsJ, ntJ, Drq, Dsq, Dtq, VqLq,
                           VqPq, VfPq, Jq, rk4a[st], rk4b[st], dt, Qf, QfB, rhsQ, resQ, Q, s),
              "wadges_stage");
        double *t = Qf;
        Qf = QfB;
        QfB = t;
      }
    }
  if (hipDeviceSynchronize() != hipSuccess) die("hipDeviceSynchronize");
  download("Q", Q, D * K * nf * Nq);
  download("rhsQ", rhsQ, D * K * nf * Nq);
  download("resQ", resQ, D * K * nf * Nq);
  download("Qf", Qf, D * K * nf * Nf);
  check(wadges_plan_destroy(plan), "wadges_plan_destroy");
  printf("rk_driver: %d steps (mode %d) on K=%d elements done\n", nsteps, mode, K);
  return 0;
}
