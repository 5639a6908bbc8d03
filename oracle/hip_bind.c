// The reference's own susy_hmc driver linked against libsusy_hip_U2.so: the
// drop-in test of the C-ABI boundary.  TEST INFRASTRUCTURE ONLY, built by
// oracle/ref.mk from the reference's objects (compiled in place from
// /root/reference/4d_Q16) into oracle/_ref/susy_hmc_U2_hip.
//
// The five symbols fermion_op, DSq, congrad_multi, fermion_force and
// gauge_force are localised in the reference's objects (objcopy -L), so the
// driver's calls reach the library.  This file adds the binding the library
// needs, without editing the reference's sources, through the linker's
// --wrap of two reference functions:
//   * readin (setup.c): after the reference has read its parameters, set up
//     the lattice and loaded the configuration, register its globals with the
//     library (susy_bind_*);
//   * compute_Fmunu (action.c), which the reference calls last wherever the
//     links change (update_u, update_o.c:31-34; the restore after a
//     rejection, update_o.c:239-242; readin, setup.c:551-554): tell the
//     library to re-read the links.
#include <stddef.h>

#include "susy_includes.h"

int susy_bind_lattice(int nx, int ny, int nz, int nt, int PBC);
int susy_bind_couplings(double lambda, double kappa_u1, double bmass, double fmass,
                        double G);
int susy_bind_sites(void *lattice, size_t site_bytes, size_t link_offset,
                    size_t mom_offset);
int susy_bind_rhmc(int *Norder, Real **shift, Real **amp4, int *total_iters);
int susy_links_changed(void);
const char *susy_last_error(void);

int __real_readin(int prompt);
void __real_compute_Fmunu(void);
static int bound = 0;

int __wrap_readin(int prompt) {
  int r = __real_readin(prompt);
  if (r != 0) return r;
  if (susy_bind_lattice(nx, ny, nz, nt, PBC) || susy_bind_couplings(lambda, kappa_u1, bmass,
                                                                     fmass, G) ||
      susy_bind_sites(lattice, sizeof(site), offsetof(site, link), offsetof(site, mom)) ||
      susy_bind_rhmc(&Norder, &shift, &amp4, &total_iters) || susy_links_changed()) {
    node0_printf("hip_bind: %s\n", susy_last_error());
    return 1;
  }
  bound = 1;
  node0_printf("hip_bind: fermion_op, DSq, congrad_multi, fermion_force and gauge_force "
               "run in libsusy_hip_U%d\n", NCOL);
  return 0;
}

void __wrap_compute_Fmunu(void) {
  __real_compute_Fmunu();
  if (bound && susy_links_changed()) {
    node0_printf("hip_bind: %s\n", susy_last_error());
    terminate(1);
  }
}
