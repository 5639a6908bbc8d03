// Timing of the REFERENCE build (oracle/_ref) on the host's cores: the CPU
// baseline of the benchmark, not part of the product.  Compiled against the
// reference's own headers and linked with its own objects (built in place from
// /root/reference/4d_Q16 by oracle/ref.mk as an MPI program; see there for
// the gauge group), plus oracle/ref_harness.c for the set-up.
//
//   mpiexec -n P susy_refbench_U4 nx ny nz nt n_op n_cg n_pole
//
// On links I + 0.1 (Gaussian) per element and a Gaussian source it times, on P
// ranks, with a barrier around each timed region:
//   * n_op calls of fermion_op (utilities.c:765),
//   * n_cg iterations of the 15-shift congrad_multi (congrad_multi.c:26) with
//     the Nroot = 2 MD shifts (shift4), capped by MaxCG,
//   * one fermion_force (update_h.c:928) over the first n_pole poles of that
//     solve (the global Norder lowered for the call, as the force loops over it),
// and prints one JSON line from rank 0 (seconds per call / iteration / pole).
#include "susy_includes.h"

int ref_init(int nx_, int ny_, int nz_, int nt_, int PBC_, int iseed_, int Nroot_,
             int Norder_, double lambda_, double kappa_u1_, double bmass_, double fmass_,
             double G_, int niter_, double error_per_site, double traj_length_, int nstep,
             int nstep_gauge);

static void gauss_matrix(double_prn *p, matrix *m, double scale, int unit) {
  int i, j;
  for (i = 0; i < NCOL; i++)
    for (j = 0; j < NCOL; j++) {
      m->e[i][j].real = scale * gaussian_rand_no(p) + (unit && i == j ? 1.0 : 0.0);
      m->e[i][j].imag = scale * gaussian_rand_no(p);
    }
}

int main(int argc, char **argv) {
  if (argc < 8) {
    fprintf(stderr, "usage: %s nx ny nz nt n_op n_cg n_pole\n", argv[0]);
    return 1;
  }
  int d[4], k, i, j, mu;
  for (k = 0; k < 4; k++) d[k] = atoi(argv[1 + k]);
  int n_op = atoi(argv[5]), n_cg = atoi(argv[6]), n_pole = atoi(argv[7]);
  // the benchmark's couplings: lambda 2, kappa_u1 0, bmass 0.4, fmass 0,
  // G 0.05, Nroot 2, Norder 15, antiperiodic temporal fermion BC
  double t0 = dclock();
  ref_init(d[0], d[1], d[2], d[3], -1, 1234, 2, 15, 2.0, 0.0, 0.4, 0.0, 0.05, n_cg, 1e-5,
           1.0, 20, 10);
  site *s;
  FORALLSITES(i, s) FORALLDIR(mu) gauss_matrix(&s->site_prn, &s->link[mu], 0.1, 1);
  compute_plaqdet();
  compute_Uinv();
  compute_DmuUmu();
  compute_Fmunu();
  Twist_Fermion *src = malloc(sizeof(Twist_Fermion) * sites_on_node);
  Twist_Fermion *dst = malloc(sizeof(Twist_Fermion) * sites_on_node);
  Twist_Fermion **psim = malloc(sizeof(Twist_Fermion *) * Norder);
  for (j = 0; j < Norder; j++) psim[j] = malloc(sizeof(Twist_Fermion) * sites_on_node);
  FORALLSITES(i, s) {
    gauss_matrix(&s->site_prn, &src[i].Fsite, 1.0, 0);
    for (mu = 0; mu < NUMLINK; mu++) gauss_matrix(&s->site_prn, &src[i].Flink[mu], 1.0, 0);
    for (mu = 0; mu < NPLAQ; mu++) gauss_matrix(&s->site_prn, &src[i].Fplaq[mu], 1.0, 0);
  }
  g_sync();
  double t_setup = dclock() - t0;

  fermion_op(src, dst, PLUS);   // first touch
  g_sync();
  t0 = dclock();
  for (k = 0; k < n_op; k++) fermion_op(src, dst, PLUS);
  g_sync();
  double t_op = (dclock() - t0) / n_op;

  for (j = 0; j < Norder; j++) shift[j] = shift4[j];
  Real size_r;
  g_sync();
  t0 = dclock();
  int it = congrad_multi(src, psim, n_cg, 1e-30, &size_r);
  g_sync();
  double t_cg = (dclock() - t0) / (it > 0 ? it : 1);

  if (n_pole < 1 || n_pole > Norder) n_pole = Norder;
  const int norder = Norder;
  Norder = n_pole;
  g_sync();
  t0 = dclock();
  fermion_force(0.01, src, psim);
  g_sync();
  double t_pole = (dclock() - t0) / n_pole;
  Norder = norder;

  if (this_node == 0)
    printf("{\"ranks\": %d, \"ncol\": %d, \"dims\": [%d, %d, %d, %d], \"setup_s\": %.4f, "
           "\"fermion_op_s\": %.5f, \"cg_iterations\": %d, \"cg_iteration_s\": %.5f, "
           "\"force_pole_s\": %.5f}\n",
           numnodes(), NCOL, d[0], d[1], d[2], d[3], t_setup, t_op, it, t_cg, t_pole);
  fflush(stdout);
  normal_exit(0);
  return 0;
}
