// MI355X (gfx950) implementation of the 4d N=4 SYM RHMC hot path:
// twisted fermion operator, multi-shift CG, fermion and gauge forces.
//
// One shared library per gauge group U(NCOL), NCOL in {2, 3, 4}:
//   libsusy_hip_U<NCOL>.so.  Define NCOL to the library's value before
//   including this header; the Twist_Fermion layout depends on it.
//
// Three layers of entry points:
//  1. Drop-in symbols with the reference's exact C signatures
//     (4d_Q16/susy/susy_includes.h:73-81).  They read the caller's globals
//     through the pointers registered with susy_bind_*() (the reference keeps
//     them as globals in susy/lattice.h:71-185) and move Twist_Fermion arrays
//     between host and device on every call.
//  2. Binding calls a reference driver makes once (and after every link
//     update) so that the drop-in symbols see its state.
//  3. A device-resident API used by this project's own RHMC driver, tests and
//     benchmark: fields stay in HBM between calls.
//
// Errors: entry points returning int return 0 on success and a negative
// SUSY_E* code on failure; susy_last_error() gives a message.  A bad sign in
// fermion_op is fatal, as in the reference (utilities.c:789-792).
#ifndef SUSY_HIP_H
#define SUSY_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#ifndef NCOL
#error "define NCOL (2, 3, 4 or 5) to match the library before including susy_hip.h"
#endif

// ---------------------------------------------------------------------
// Types, layout-compatible with the reference (include/susy.h:29-44,
// susy/lattice.h:18-22) at PRECISION=2: Real = double.
typedef double Real;
typedef struct { double real, imag; } susy_complex;
typedef struct { susy_complex e[NCOL][NCOL]; } susy_matrix;
typedef struct {
  susy_matrix Fsite;
  susy_matrix Flink[5];
  susy_matrix Fplaq[10];
} Twist_Fermion;

enum {
  SUSY_OK = 0,
  SUSY_E_ARG = -1,      // invalid argument (sizes, sign, unsupported option)
  SUSY_E_STATE = -2,    // no lattice bound / not initialised
  SUSY_E_HIP = -3,      // HIP runtime error
  SUSY_E_NCOL = -4,     // library built for another NCOL
  SUSY_E_NOMEM = -5,    // device allocation failed
};

// ---------------------------------------------------------------------
// 1. Drop-in replacements (susy_includes.h:73-81)
// dest = D src (sign = +1) or D^dagger src (sign = -1)   [utilities.c:765]
void fermion_op(Twist_Fermion *src, Twist_Fermion *dest, int sign);
// dest = (D^dagger D + fmass^2) src                        [utilities.c:874]
void DSq(Twist_Fermion *src, Twist_Fermion *dest);
// (D^dagger D + shift[j]) psim[j] = src for j < Norder, zero initial guess;
// reads the bound globals Norder and shift[]; returns the iteration count,
// adds it to the bound total_iters and writes |r|^2 to *size_r
//                                                          [congrad_multi.c:26]
int congrad_multi(Twist_Fermion *src, Twist_Fermion **psim, int MaxCG,
                  Real errormin, Real *size_r);
// mom -= eps * sum_j amp4[j] dS_F/dU for the bound amp4[], Norder; src is
// unused, as in the reference; returns eps*|F|/volume      [update_h.c:928]
double fermion_force(Real eps, Twist_Fermion *src, Twist_Fermion **sol);
// mom -= kappa eps dS_B/dU (+ det_force when kappa_u1 > 0) [update_h.c:15]
double gauge_force(Real eps);

// ---------------------------------------------------------------------
// 2. Binding a reference driver (replaces direct access to its globals)
// Lattice geometry and temporal fermion BC (setup.c:67-72; PBC = +1 or -1).
int susy_bind_lattice(int nx, int ny, int nz, int nt, int PBC);
// The caller's site array: base pointer, sizeof(site), offsetof(site, link)
// and offsetof(site, mom).  Links are read from it by susy_links_changed();
// the force entry points update mom in it.
int susy_bind_sites(void *lattice, size_t site_bytes, size_t link_offset,
                    size_t mom_offset);
// Couplings as set in readin (setup.c:445-459): lambda, kappa_u1, bmass,
// fmass, G.  kappa = NCOL/(4 lambda) and doG = (G > 1e-8) follow from them.
int susy_bind_couplings(double lambda, double kappa_u1, double bmass,
                        double fmass, double G);
// Pointers to the caller's RHMC globals, read at every call:
// Norder, shift[], amp4[], total_iters (lattice.h:161-164, 110).
int susy_bind_rhmc(int *Norder, Real **shift, Real **amp4, int *total_iters);
// Re-read the links from the bound site array and recompute the derived
// fields (plaqdet, U^-1, DmuUmu, Fmunu) on the device.  Call after every
// change of the links: where the reference calls compute_plaqdet() and
// friends (update_o.c:31-34, setup.c:551-554).
int susy_links_changed(void);

// ---------------------------------------------------------------------
// 3. Device-resident API
typedef struct susy_ctx susy_ctx;

typedef struct {
  int nx, ny, nz, nt, PBC, iseed;
  int Nroot, Norder;             // Norder must be 15 (tables of setup_rhmc.c)
  double lambda, kappa_u1, bmass, fmass, G;
  int niter;                     // max CG iterations
  double error_per_site;         // CG stops at |r|^2 <= error^2 |src|^2
  double traj_length;
  int nstep, nstep_gauge;        // fermion and gauge integrator steps
} susy_params;

// Precision form of the solver inside a trajectory.
enum {
  SUSY_FORM_FP64 = 0,            // plain FP64 everywhere (parity form)
  SUSY_FORM_FP32 = 1,            // FP32 DSq + CG with FP64 reliable updates
  SUSY_FORM_SURROGATE = 4,       // flag: learned surrogate force on the inner MD level
  SUSY_FORM_PUREGAUGE = 8,       // flag: fermions off, the reference's -DPUREGAUGE build
                                 // (update_o.c:74-177, action.c:314): no pseudofermions,
                                 // solves or fermion forces; the action has no fermion part
  SUSY_FORM_SPLIT = 16,          // flag: the large-shift poles on a middle MD level
                                 // (susy_set_split)
};

int susy_ncol(void);
const char *susy_last_error(void);
// Host-only consistency check of the operator/force term tables (0 = ok).
int susy_validate_tables(void);
// CG tuning: residual window of the shifted-system updates (1, 4, 8, 16 or
// 32 iterations; 0 = the default, 32) and, for
// the FP32 form, iterations between FP64 true residuals (default 128).
// Process-wide.
int susy_set_cg_options(int window, int reliable);
// Progress lines on stderr from susy_update (one per solve) when v != 0.
int susy_set_verbose(int v);
int susy_create(const susy_params *p, susy_ctx **out);
void susy_destroy(susy_ctx *c);
int susy_volume(const susy_ctx *c);
// The context's RHMC coefficients (setup_rhmc.c:882-1031, the Nroot row):
// out[62] = ampdeg4, amp4[15], shift4[15], ampdeg8, amp8[15], shift8[15].
int susy_rhmc_coeffs(const susy_ctx *c, double *out);

// Site order is the reference's (even sites first, layout_hyper_prime.c:157).
// Gather table g in 0..73 as numbered by the reference (com_vanilla.c,
// setup_offset.c): out[i] = index of the site gathered to site i.
int susy_get_gather(susy_ctx *c, int g, int *out);

// Host <-> device transfers of whole fields in the reference's AoS layout:
// links/momenta = V x 5 matrices; Twist_Fermion fields = V x 16 matrices.
int susy_set_links(susy_ctx *c, const double *links);   // + derived fields
int susy_get_links(susy_ctx *c, double *links);
int susy_set_mom(susy_ctx *c, const double *mom);
int susy_get_mom(susy_ctx *c, double *mom);
// MILC binary gauge file (io_lat.c:826-1006), natural order; returns 1 if the
// file's checksums do not match, 0 if they do.
int susy_load_config(susy_ctx *c, const char *path);
int susy_cold_start(susy_ctx *c);
// Derived fields (AoS): plaqdet 25 x V complex ([a][b][site]), Uinv 5 x V,
// DmuUmu V, Fmunu 10 x V matrices.  Any pointer may be NULL.
int susy_get_fields(susy_ctx *c, double *plaqdet, double *uinv, double *dmu,
                    double *fmunu);

// Device Twist_Fermion vectors owned by the context (SoA in HBM).
int susy_tf_alloc(susy_ctx *c, int count);          // returns first handle
int susy_tf_upload(susy_ctx *c, int h, const Twist_Fermion *src);
int susy_tf_download(susy_ctx *c, int h, Twist_Fermion *dst);

// Resident operations on handles.
int susy_fermion_op_dev(susy_ctx *c, int h_src, int h_dst, int sign, int form);
int susy_dsq_dev(susy_ctx *c, int h_src, int h_dst, int form);
// Plain copy kernel of one field (reads and writes 16 NCOL^2 V complex
// numbers once): calibrates hardware memory-traffic counters.
int susy_stream_copy_dev(susy_ctx *c, int h_src, int h_dst);
// Multi-shift CG: psim handles h_psim .. h_psim+nshift-1.  Returns iterations
// (>= 0) or an error code.  form selects the precision form.
int susy_congrad_multi_dev(susy_ctx *c, int h_src, int h_psim, int nshift,
                           const double *shifts, int MaxCG, double errormin,
                           double *size_r, int form);
// The operator without allocation: dst = D src (sign +1) or D^dagger src
// (sign -1) in FP64, like susy_fermion_op_dev, for src and dst given as device
// pointers to Twist_Fermion fields in the context's layout (16 NCOL^2 planes
// of V complex doubles, element (i,j) of component k at site s at complex
// index (k NCOL^2 + i NCOL + j) V + s).  No counterpart in the reference: it
// is for callers that cannot allocate (code under HIP graph capture, serving
// loops).  It calls no hipMalloc/hipFree: its scratch (6 complex doubles per
// site) is a fixed 16 MiB device array of the library, so V <= 174762
// (larger lattices return SUSY_E_ARG).  All calls on one GPU share that
// array: calls that overlap on one GPU race.  Synchronises the context's
// stream before returning.
int susy_fermion_op_noalloc(susy_ctx *c, const double *src, double *dst, int sign);
// Fermion force from nsol solutions; updates the device momenta.
int susy_fermion_force_dev(susy_ctx *c, double eps, int h_sol, int nsol,
                           const double *amps, double *norm);
int susy_gauge_force_dev(susy_ctx *c, double eps, double *norm);
int susy_update_u_dev(susy_ctx *c, double eps);
// Action parts: gauge, bmass, det, mom (FP64, compensated sums).
int susy_actions_dev(susy_ctx *c, double *out4);
int susy_fermion_action_dev(susy_ctx *c, int h_src, int h_sol, double *out);
int susy_plaquette_dev(susy_ctx *c, double *ss, double *st);

// Heatbath pieces of a trajectory, with the reference's per-site generators
// (grsource.c:44-64, 200-303): fresh Gaussian momenta; a pseudofermion
// src = (D^dag D)^(1/8 Nroot) g with the solutions of its solve in
// h_psim .. h_psim+Norder-1.
int susy_ranmom(susy_ctx *c);
// Restart all generators from `seed` (as at start-up).
int susy_reseed(susy_ctx *c, int seed);
int susy_grsource_dev(susy_ctx *c, int h_src, int h_psim, int form, int *iters);

// One RHMC trajectory (update_o.c:127-282) in the given form.
// stats[0..22]: start S, end S, dS, accepted, CG iterations, gauge-force
// avg/max, fermion-force avg/max for roots 0 and 1, start action parts
// (gauge, bmass, det, fermion, mom, total), end action parts; stats[23]:
// CG iterations of the middle level (SUSY_FORM_SPLIT; included in stats[4]).
// stats holds 32 doubles.
int susy_update(susy_ctx *c, int form, double *stats);
// The same trajectory in pieces, for callers that time or interleave its
// outer MD steps; susy_update = begin, nstep x step, end.
//   susy_traj_begin: momenta, pseudofermions and their solves, start action,
//     saved links/momenta and the first fermion kick (update_o.c:127-208,
//     74-82).  Abandons an unfinished trajectory.
//   susy_traj_step: one outer step of update_step (update_o.c:80-113):
//     gauge level, solves + kick, gauge level, solves + kick.  Returns the
//     steps done (1..nstep); the step's device work is complete on return
//     (as it is after susy_traj_begin and susy_traj_end).
//   susy_traj_end: after exactly nstep steps, end action and Metropolis test
//     (update_o.c:209-282); returns 1 accepted / 0 rejected, stats as above.
//   susy_traj_substep: one sub-step of an outer step: the multi-shift solve
//     and fermion kick of one root, the first of each half step also running
//     the gauge level before it (2 Nroot sub-steps make one outer step, in
//     the order of susy_traj_step).  Returns the sub-steps done so far.
int susy_traj_begin(susy_ctx *c, int form);
int susy_traj_step(susy_ctx *c);
int susy_traj_substep(susy_ctx *c);
int susy_traj_end(susy_ctx *c, double *stats);
// CG iterations of the current (or last) trajectory so far: all, and those of
// the middle level of SUSY_FORM_SPLIT (either pointer may be NULL).
int susy_traj_iters(susy_ctx *c, int *iters, int *mid_iters);

// Timing of the last call's kernels is available through the profiler; the
// benchmark measures operator launches with these helpers (HIP events on the
// library's own stream).  Returns average milliseconds per call.
int susy_time_fermion_op(susy_ctx *c, int h_src, int h_dst, int form,
                         int reps, double *ms_per_call);
// Sampled HIP-event timing of the operator kernel D (fermion_op sign +1, the
// first half of DSq) wherever the context launches it (CG, forces,
// trajectories): every `every`-th such launch is timed (0 = off).
// susy_prof_read returns the average sampled duration in ms and the number
// of D launches for prec 0 (FP64 kernel) or 1 (FP32 kernel).
int susy_prof_enable(susy_ctx *c, int every);
int susy_prof_read(susy_ctx *c, int prec, double *avg_ms, long *launches,
                   long *sampled);
/* Block order of the operator kernel (a tuning knob; results do not depend
   on it): 0 = each XCD walks a contiguous range of sites, 1 = each XCD walks
   the same range of the even and the odd half of the site order, 2 = as 1
   with the range visited t-slice fastest (default; a lattice whose slabs
   are not whole t-slices of whole chunks runs order 1).  A value below 256 sets both
   precisions; otherwise the low byte is the FP64 order and the next byte the
   FP32 order. */
int susy_set_site_order(susy_ctx *c, int paired);

/* MILC serial gauge file of the context's links, readable by
   susy_load_config and by the reference's reload_serial (io_lat.c); stamp: the
   64-byte time stamp to record, or NULL for the current time. */
int susy_save_config(susy_ctx *c, const char *path, const char *stamp);
/* Measurements the reference prints after every trajectory (control.c:85-138):
   FLINK (link_trace.c) -- Tr(U_a U_a^dag)/N per direction, average, width, and
   Re det of the traceless part of U_a U_a^dag per direction, average, width;
   Polyakov / Wilson lines (ploop.c, ploop(dir, project, &mod)) in link
   direction a = 0..4 (GMES uses a = 3, LINES all five, LINES_POLAR all five
   with project = 1: links replaced by the unitary factor of their polar
   decomposition): average trace over the starting hyperplane and the
   average modulus;
   UUBAR_EIG / POLAR_EIG (scalar_eig.c, scalar_eig(project, ...)) -- per
   eigenvalue index (ascending) the average, width, minimum and maximum over
   all links of the eigenvalues of the traceless part of U U^dag (project 0)
   or of log (U U^dag)^(1/2) (project 1); arrays of NCOL;
   DET (determinant.c:93-119, measure_det) -- det5 = Re, Im <plaqdet>,
   <Re^2>, <Im^2>, <|plaqdet - 1|^2>, and WIDTHS (widths.c) -- w3 = widths
   of Re Tr, Re det, Im det of the plaquettes. */
int susy_link_trace_dev(susy_ctx *c, double *linktr, double *ave, double *width,
                        double *link_det, double *det_ave, double *det_width);
int susy_wilson_line_dev(susy_ctx *c, int a, int project, double *re, double *im,
                         double *mod);
int susy_scalar_eig_dev(susy_ctx *c, int project, double *ave, double *width, double *mn,
                        double *mx);
int susy_plaq_moments_dev(susy_ctx *c, double *det5, double *w3);

/* Extreme eigenvalues of DSq (D^dag D + fmass^2) at the context's links, the
   measurement the reference prints in testsuite/scalar/eig.U*.ref (there
   with PRIMME): Lanczos with full reorthogonalisation on the FP64 operator,
   stopping when the nlow smallest and nhigh largest Ritz values move by less
   than tol (relative) over 20 iterations, or after maxiter.  Distinct
   eigenvalues are resolved; an exact degeneracy shows up only as far as
   rounding seeds its second direction.  Holds maxiter + 1 fields in HBM. */
int susy_eig_dsq(susy_ctx *c, int nlow, int nhigh, int maxiter, double tol, int seed,
                 double *low, double *high, int *iters);

/* Pfaffian of the fermion operator (phase.c, printed as PFAFF in
   testsuite/scalar/phase.U*.ref): M_ab = <e_a|D|e_b> (bilinear) in the
   reference's generator basis, built column by column with the FP64
   operator; log|Pf M| and arg Pf M in [0, 2 pi); asym (may be NULL) =
   max|M + M^T| / max|M|.  For lattices with 16 N^2 V <= 8192 (2^4). */
int susy_pfaffian_dev(susy_ctx *c, double *log_mag, double *phase, double *asym);

/* The same matrix M (phase.c:36-213, matvec): n = 16 N^2 V, basis entry
   (site, generator j, slot q) at index (site * N^2 + j) * 16 + q, written
   row-major M[a][b] with (re, im) interleaved into 2 n^2 doubles. */
int susy_fermion_matrix_dev(susy_ctx *c, double *m);

/* Learned surrogate action S_theta(U) (no counterpart in the reference; it
   is trained on and checked against the plain fermion_force).  Two gauge-invariant convolution layers
   (16 channels, bfloat16 weights, FP32 accumulation) over plaquette and
   1x2-loop traces; F_theta = dS_theta/dUbar in the convention of
   fermion_force.  With SUSY_FORM_SURROGATE, susy_update adds F_theta to every
   inner (gauge) force and removes it from every outer (fermion) kick, so the
   outer level integrates S_F - S_theta and the exact action still decides
   the Metropolis test. */
int susy_sg_nparams(void);
int susy_sg_init(susy_ctx *c, int seed, double scale);
int susy_sg_set_params(susy_ctx *c, const float *theta);
int susy_sg_get_params(susy_ctx *c, float *theta);
/* S_theta and (if F_aos != NULL) F_theta at the context's links, AoS (V,5,N,N) */
int susy_sg_eval(susy_ctx *c, double *action, double *F_aos);
/* the 65 features as planes [feature][site] */
int susy_sg_features(susy_ctx *c, float *phi);
/* force-matching loss |F_theta - F|^2 / |F|^2 and its parameter gradient */
int susy_sg_loss_grad(susy_ctx *c, const double *U_aos, const double *F_aos, double *loss,
                      double *grad);
/* record (links, fermion force) at every outer kick of susy_update, keeping
   the last `capacity` pairs (0 stops); susy_sg_pairs = pairs held */
int susy_sg_record(susy_ctx *c, int capacity);
int susy_sg_pairs(susy_ctx *c);
int susy_sg_pair_loss(susy_ctx *c, int i, double *loss);
/* Adam over recorded pairs [first, first + count) */
int susy_sg_train(susy_ctx *c, int first, int count, int steps, int batch, double lr, int seed,
                  double *loss_hist);

/* Pole-split integrator (no counterpart in the reference, whose update_o.c
   integrates the whole pseudofermion action on the outer level).  With
   SUSY_FORM_SPLIT the rational action's poles j >= first_pole (the large
   shifts, whose multi-shift solve converges in few iterations) form a middle
   MD level: S_mid = sum_roots Phi^dag sum_{j >= first_pole} w_j (D^dag D +
   shift_j)^(-1) Phi, integrated with nstep_mid Omelyan steps inside each
   half outer step, each step running the gauge level (nstep_gauge steps)
   twice.  The outer level integrates S_F - S_mid (amplitudes amp4_j for
   j < first_pole, amp4_j - w_j above).  weights = NULL takes w_j = amp4_j
   (the outer level keeps exactly the small-shift poles).  The Hamiltonian
   and the Metropolis test are unchanged.  1 <= first_pole < Norder,
   nstep_mid >= 1.  Together with SUSY_FORM_SURROGATE the learned force still
   enters every gauge-level force and leaves every outer kick, so it stands
   in for the outer level's force only: train it on the pairs that
   susy_sg_record records in the split form (the outer forces). */
int susy_set_split(susy_ctx *c, int first_pole, int nstep_mid, const double *weights);

#ifdef __cplusplus
}
#endif
#endif  // SUSY_HIP_H
