// Test harness around the REFERENCE build (oracle/_ref), compiled against the
// reference's own headers and linked with its own objects, which are compiled
// in place from /root/reference/4d_Q16 by oracle/ref.mk.  TEST INFRASTRUCTURE
// ONLY: tests/ use it to pin the C restatement (oracle/susy_oracle.c) and to
// write golden fixtures; the product never links it.
//
// It replaces only the stdin parsing of setup.c:16-119/307-556 and the main
// loop of control.c: parameters are set directly, then the reference's own
// setup routines run in the order of setup.c:240-260 and control.c:24-29.
// The reference header fixes NCOL (include/susy.h:16), so this build is U(2).
#define CONTROL
#include "susy_includes.h"

int ref_init(int nx_, int ny_, int nz_, int nt_, int PBC_, int iseed_,
             int Nroot_, int Norder_, double lambda_, double kappa_u1_,
             double bmass_, double fmass_, double G_, int niter_,
             double error_per_site, double traj_length_, int nstep,
             int nstep_gauge) {
  int argc = 1;
  char *argv0[] = {"ref", NULL};
  char **argv = argv0;
  initialize_machine(&argc, &argv);
  // initial_set (setup.c:67-117)
  nx = nx_; ny = ny_; nz = nz_; nt = nt_; PBC = PBC_; iseed = iseed_;
  length[0] = nx; length[1] = ny; length[2] = nz; length[3] = nt;
  Nroot = Nroot_;
  fnorm = malloc(sizeof fnorm * Nroot);
  max_ff = malloc(sizeof max_ff * Nroot);
  Norder = Norder_;
  amp = malloc(sizeof amp * Norder);
  amp4 = malloc(sizeof amp4 * Norder);
  amp8 = malloc(sizeof amp8 * Norder);
  shift = malloc(sizeof shift * Norder);
  shift4 = malloc(sizeof shift4 * Norder);
  shift8 = malloc(sizeof shift8 * Norder);
  this_node = mynode();
  number_of_nodes = numnodes();
  one_ov_N = 1.0 / (Real)NCOL;
  volume = nx * ny * nz * nt;
  total_iters = 0;
  minus1 = cmplx(-1.0, 0.0);
  // setup (setup.c:240-260)
  initialize_prn(&node_prn, iseed, volume + mynode());
  setup_layout();
  make_lattice();
  make_nn_gathers();
  setup_offset();
  setup_qclosed_offset();
  make_fields();
  // control.c:25-29
  setup_lambda();
  epsilon();
  setup_PtoP();
  setup_FQ();
  setup_rhmc();
  // readin (setup.c:433-459, 538-554)
  warms = 0; trajecs = 1;
  traj_length = traj_length_;
  nsteps[0] = nstep; nsteps[1] = nstep_gauge;
  niter = niter_;
  rsqmin = error_per_site * error_per_site;
  lambda = lambda_; kappa_u1 = kappa_u1_; bmass = bmass_; fmass = fmass_;
  G = G_;
  doG = (G > IMAG_TOL) ? 1 : 0;
  kappa = (Real)NCOL * 0.25 / lambda;
  ipiv = malloc(sizeof *ipiv * NCOL);
  store = malloc(sizeof *store * 2 * NCOL * NCOL);
  work = malloc(sizeof *work * 4 * NCOL);
  Rwork = malloc(sizeof *Rwork * (3 * NCOL - 2));
  eigs = malloc(sizeof *eigs * NCOL);
  coldlat();
  compute_plaqdet();
  compute_Uinv();
  compute_DmuUmu();
  compute_Fmunu();
  return NCOL;
}

static void recompute(void) {
  compute_plaqdet();
  compute_Uinv();
  compute_DmuUmu();
  compute_Fmunu();
}

int ref_load_config(char *path) {
  startlat_p = reload_lattice(RELOAD_SERIAL, path);
  recompute();
  return 0;
}

void ref_set_links(const double *l) {
  int i, mu;
  site *s;
  FORALLSITES(i, s) FORALLDIR(mu)
    memcpy(&s->link[mu], l + ((size_t)i * NUMLINK + mu) * 2 * DIMF, sizeof(matrix));
  recompute();
}
void ref_get_links(double *l) {
  int i, mu;
  site *s;
  FORALLSITES(i, s) FORALLDIR(mu)
    memcpy(l + ((size_t)i * NUMLINK + mu) * 2 * DIMF, &s->link[mu], sizeof(matrix));
}
void ref_set_mom(const double *m) {
  int i, mu;
  site *s;
  FORALLSITES(i, s) FORALLDIR(mu)
    memcpy(&s->mom[mu], m + ((size_t)i * NUMLINK + mu) * 2 * DIMF, sizeof(matrix));
}
void ref_get_mom(double *m) {
  int i, mu;
  site *s;
  FORALLSITES(i, s) FORALLDIR(mu)
    memcpy(m + ((size_t)i * NUMLINK + mu) * 2 * DIMF, &s->mom[mu], sizeof(matrix));
}

// Neighbour table of gather id g, read through the reference's own gather:
// gathering the field f[i] = i leaves f[neighbor[i]] behind each pointer.
void ref_get_gather(int g, int *out) {
  int i;
  site *s;
  int *f = malloc(sizeof(int) * sites_on_node);
  FORALLSITES(i, s) f[i] = i;
  msg_tag *tag = start_gather_field(f, sizeof(int), g, EVENANDODD, gen_pt[0]);
  wait_gather(tag);
  FORALLSITES(i, s) out[i] = *(int *)gen_pt[0][i];
  cleanup_gather(tag);
  free(f);
}
void ref_get_coords(int *xyzt) {
  int i;
  site *s;
  FORALLSITES(i, s) {
    xyzt[4 * i] = s->x; xyzt[4 * i + 1] = s->y;
    xyzt[4 * i + 2] = s->z; xyzt[4 * i + 3] = s->t;
  }
}
// Boundary-condition tables of one site: bc[10], bc2[10][10], bc3[10][10][10]
void ref_get_bc(int i, double *out) {
  memcpy(out, lattice[i].bc, sizeof(Real) * 10);
  memcpy(out + 10, lattice[i].bc2, sizeof(Real) * 100);
  memcpy(out + 110, lattice[i].bc3, sizeof(Real) * 1000);
}
void ref_get_tables(int *gidx, int *lookups) {
  // goffset[10], gq_offset[56], DbmP_d1/d2, DbpP_d1/d2, F1Q_d1/d2, F2Q_d1/d2
  memcpy(gidx, goffset, sizeof(int) * 10);
  memcpy(gidx + 10, gq_offset, sizeof(int) * 56);
  memcpy(gidx + 66, DbmP_d1, sizeof(int) * 30);
  memcpy(gidx + 96, DbmP_d2, sizeof(int) * 30);
  memcpy(gidx + 126, DbpP_d1, sizeof(int) * 30);
  memcpy(gidx + 156, DbpP_d2, sizeof(int) * 30);
  memcpy(gidx + 186, F1Q_d1, sizeof(int) * 30);
  memcpy(gidx + 216, F1Q_d2, sizeof(int) * 30);
  memcpy(gidx + 246, F2Q_d1, sizeof(int) * 30);
  memcpy(gidx + 276, F2Q_d2, sizeof(int) * 30);
  memcpy(lookups, DbplusPtoP_lookup, sizeof(int) * 150);
  memcpy(lookups + 150, DbminusPtoP_lookup, sizeof(int) * 150);
  memcpy(lookups + 300, FQ_lookup, sizeof(int) * 150);
  memcpy(lookups + 450, q_offset, sizeof(int) * 112);
}

void ref_fermion_op(const double *src, double *dest, int sign) {
  fermion_op((Twist_Fermion *)src, (Twist_Fermion *)dest, sign);
}
void ref_DSq(const double *src, double *dest) {
  DSq((Twist_Fermion *)src, (Twist_Fermion *)dest);
}
int ref_congrad_multi(const double *src, double *psim, int nshift,
                      const double *shifts, int MaxCG, double errormin,
                      double *size_r) {
  int j, save = Norder;
  Twist_Fermion *ps[64];
  Real sz;
  Norder = nshift;
  for (j = 0; j < nshift; j++) {
    shift[j] = shifts[j];
    ps[j] = (Twist_Fermion *)psim + (size_t)j * sites_on_node;
  }
  j = congrad_multi((Twist_Fermion *)src, ps, MaxCG, errormin, &sz);
  *size_r = sz;
  Norder = save;
  return j;
}
// Uses the reference's global amp4[] (Nroot table); sol holds Norder fields
double ref_fermion_force(double eps, const double *sol) {
  int j;
  Twist_Fermion *ps[64];
  for (j = 0; j < Norder; j++)
    ps[j] = (Twist_Fermion *)sol + (size_t)j * sites_on_node;
  return fermion_force(eps, NULL, ps);
}
double ref_gauge_force(double eps) { return gauge_force(eps); }
void ref_update_u(double eps) { update_u(eps); }
int ref_update(void) { return update(); }
void ref_plaquette(double *ss, double *st) { plaquette(ss, st); }
void ref_actions(double *out) {
  out[0] = gauge_action(NODET);
  out[1] = bmass_action();
  out[2] = det_action();
  out[3] = mom_action();
}
double ref_fermion_action(const double *src, const double *sol) {
  int j;
  Twist_Fermion *ps[64];
  for (j = 0; j < Norder; j++)
    ps[j] = (Twist_Fermion *)sol + (size_t)j * sites_on_node;
  return fermion_action((Twist_Fermion *)src, ps);
}
void ref_rand_stream(int i, int n, double *out) {
  int k;
  double_prn *p = (i < 0) ? &node_prn : &lattice[i].site_prn;
  for (k = 0; k < n; k++) out[k] = myrand(p);
}
void ref_ranmom(void) { ranmom(); }
int ref_grsource(double *src) { return grsource((Twist_Fermion *)src); }
void ref_get_fields(double *pd, double *ui, double *dm, double *fm) {
  int a, b, i, p;
  site *s;
  FORALLSITES(i, s) {
    for (a = 0; a < NUMLINK; a++) {
      for (b = 0; b < NUMLINK; b++) {
        double *d = pd + 2 * ((size_t)sites_on_node * (a * NUMLINK + b) + i);
        if (a == b) { d[0] = 0; d[1] = 0; }
        else { d[0] = plaqdet[a][b][i].real; d[1] = plaqdet[a][b][i].imag; }
      }
      memcpy(ui + ((size_t)a * sites_on_node + i) * 2 * DIMF, &Uinv[a][i], sizeof(matrix));
    }
    memcpy(dm + (size_t)i * 2 * DIMF, &DmuUmu[i], sizeof(matrix));
    for (p = 0; p < NPLAQ; p++)
      memcpy(fm + ((size_t)p * sites_on_node + i) * 2 * DIMF, &Fmunu[p][i], sizeof(matrix));
  }
}
