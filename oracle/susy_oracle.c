// =====================================================================
// CPU ORACLE for the 4d N=4 SYM RHMC hot path.  TEST INFRASTRUCTURE ONLY.
//
// This file is a plain-C restatement of the reference algorithm
// (daschaich/susy, 4d_Q16) used as the checker for the HIP library:
// only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
// load it.  The product library never links or calls it.
//
// It is written from the mathematics of each term (owner-computes: every
// output site collects its own neighbours), not by following the reference's
// gather-pipelined loops, so that agreement with the reference build in
// oracle/_ref is a real cross-check.  Every function cites the reference
// file:line whose behaviour it restates.
//
// Parity pins (each checked by a test in tests/):
//   * gather tables, site order, lookup tables: bit-exact vs oracle/_ref
//     dumps (tests/golden) and closed forms.
//   * fermion_op / DSq / congrad_multi / fermion_force / gauge_force:
//     vs oracle/_ref on the same inputs (U(2)), adjointness <a,Db> = <D^+a,b>,
//     free-field spectrum, gauge covariance, and end-to-end vs the
//     reference's own regression outputs testsuite/scalar/hmc.U{2,3,4}.ref
//     (actions to 8 digits, exact CG iteration counts), and at U(5), U(7)
//     hmc_meas.U{5,7}.ref (start S to 12 digits; iteration counts within
//     0.1%, where LAPACK's LU order is not restated).
//   * determinant / inverse: U U^-1 = 1 and det against numpy; the LAPACK
//     pivoting convention is restated and pinned end to end by the det
//     action of hmc.U*.ref (kappa_u1 = 1 there).
//   * RNG streams (myrand, gaussian_rand_no): vs oracle/_ref stream dumps.
//
// Compile once per gauge group: gcc -O2 -DNCOL=N -shared -fPIC.
// =====================================================================
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../include/susy_rhmc15.h"

#ifndef NCOL
#error "compile with -DNCOL=N (N >= 2: the gauge group U(N))"
#endif
#define DIMF (NCOL * NCOL)
#define NUMLINK 5
#define NPLAQ 10
#define NTERMS 30
#define NQLINK 28
#define NGATHER 74
#define C2 1.0
#define IMAG_TOL 1.0e-8
#define OM_LAMBDA 0.193
#define OM_TWO_LAMBDA 0.386
#define OM_LAMBDA_MID 0.614

typedef struct { double re, im; } cx;
typedef struct { cx e[NCOL][NCOL]; } mat;
typedef struct { mat site; mat link[NUMLINK]; mat plaq[NPLAQ]; } tf_t;

// Reference generator state (include/random.h): 24-bit lagged FSR words and
// a 32-bit LCG whose state is kept in a 64-bit "unsigned long".
typedef struct {
  uint64_t r[7];
  uint64_t multiplier, addend, ic_state;
  float scale;
} prn_t;

// ---------------------------------------------------------------------
// Global state (the reference keeps the same state in globals:
// susy/lattice.h:71-185)
static int nx, ny, nz, nt, V, PBC, iseed_g;
static int *cx_, *cy_, *cz_, *ct_;
static int *gat[NGATHER];            // reference gather ids 0..73
static int *nb81;                    // neighbour at any offset in {-1,0,1}^4
static int q_offset[NQLINK][4];
static int plaq_index[NUMLINK][NUMLINK];
static double perm[NUMLINK][NUMLINK][NUMLINK][NUMLINK][NUMLINK];
static int pp_lookup[NTERMS][5];     // DbplusPtoP_lookup
static int pm_lookup[NTERMS][5];     // DbminusPtoP_lookup
static int fq_lookup[NTERMS][5];     // FQ_lookup
static mat Lambda[DIMF];

static double lambda_c, kappa, kappa_u1, bmass, fmass, G;
static int doG;
static int Nroot, Norder, niter, nsteps[2];
static double rsqmin, traj_length;
static double ampdeg4, amp4[15], shift4[15], ampdeg8, amp8[15], shift8[15];
static double shift_cur[15];
static int total_iters;

static mat *links, *mom, *old_links, *old_mom;
static cx *detU[NUMLINK];
static cx *plaqdet[NUMLINK][NUMLINK], *zwstar[NUMLINK][NUMLINK];
static mat *Uinv[NUMLINK], *Udag_inv[NUMLINK];
static mat *DmuUmu, *Fmunu[NPLAQ];
static prn_t *site_prn, node_prn;
static int g_iset;
static double g_gset;
static int quiet = 0;

// link offsets e_a (setup_offset.c:155-162): four unit vectors + (-1,-1,-1,-1)
static const int ev[NUMLINK][4] = {
  {1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0}, {0, 0, 0, 1}, {-1, -1, -1, -1}};

// ---------------------------------------------------------------------
// Complex and NCOL x NCOL matrix arithmetic (libraries/*.c)
static inline cx cmk(double r, double i) { cx c = {r, i}; return c; }
static inline cx cadd(cx a, cx b) { return cmk(a.re + b.re, a.im + b.im); }
static inline cx csub(cx a, cx b) { return cmk(a.re - b.re, a.im - b.im); }
static inline cx cmul(cx a, cx b) {
  return cmk(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
static inline cx cconj(cx a) { return cmk(a.re, -a.im); }
static inline cx cscale(double s, cx a) { return cmk(s * a.re, s * a.im); }

static inline cx mel(const mat *m, int i, int j, int adj) {
  return adj ? cconj(m->e[j][i]) : m->e[i][j];
}
static void mzero(mat *m) { memset(m, 0, sizeof(mat)); }
static void madj(const mat *a, mat *b) {  // b = a^dagger
  mat t;
  for (int i = 0; i < NCOL; i++)
    for (int j = 0; j < NCOL; j++) t.e[i][j] = cconj(a->e[j][i]);
  *b = t;
}
// c += s * op(a) * op(b), op = identity or dagger
static void madd_prod(mat *c, cx s, const mat *a, int adja, const mat *b,
                      int adjb) {
  for (int i = 0; i < NCOL; i++)
    for (int k = 0; k < NCOL; k++) {
      cx acc = cmk(0, 0);
      for (int j = 0; j < NCOL; j++)
        acc = cadd(acc, cmul(mel(a, i, j, adja), mel(b, j, k, adjb)));
      c->e[i][k] = cadd(c->e[i][k], cmul(s, acc));
    }
}
static void maddr_prod(mat *c, double s, const mat *a, int adja,
                       const mat *b, int adjb) {
  madd_prod(c, cmk(s, 0.0), a, adja, b, adjb);
}
// c += s * op(a)
static void madd_scaled(mat *c, cx s, const mat *a, int adja) {
  for (int i = 0; i < NCOL; i++)
    for (int j = 0; j < NCOL; j++)
      c->e[i][j] = cadd(c->e[i][j], cmul(s, mel(a, i, j, adja)));
}
static cx mtrace(const mat *a) {
  cx t = cmk(0, 0);
  for (int i = 0; i < NCOL; i++) t = cadd(t, a->e[i][i]);
  return t;
}
// Tr(op(a) op(b))
static cx mtrace_prod(const mat *a, int adja, const mat *b, int adjb) {
  cx t = cmk(0, 0);
  for (int i = 0; i < NCOL; i++)
    for (int j = 0; j < NCOL; j++)
      t = cadd(t, cmul(mel(a, i, j, adja), mel(b, j, i, adjb)));
  return t;
}
// ReTr(a^dagger b) (libraries/realtr.c:21-31)
static double mrealtrace(const mat *a, const mat *b) {
  double s = 0;
  for (int i = 0; i < NCOL; i++)
    for (int j = 0; j < NCOL; j++)
      s += a->e[i][j].re * b->e[i][j].re + a->e[i][j].im * b->e[i][j].im;
  return s;
}

// ---------------------------------------------------------------------
// Site order and neighbours
// node_index for one rank (generic/layout_hyper_prime.c:157-169):
// even sites first, each half in lexicographic (x fastest) order.
static int site_index(int x, int y, int z, int t) {
  int lex = x + nx * (y + ny * (z + nz * t));
  if ((x + y + z + t) % 2 == 0) return lex / 2;
  return (lex + V) / 2;
}
static int off81(const int d[4]) {
  return (d[0] + 1) + 3 * (d[1] + 1) + 9 * (d[2] + 1) + 27 * (d[3] + 1);
}
// Site at x + d with periodic wrapping (setup_offset.c:17-32)
static inline int nbr(int i, const int d[4]) { return nb81[off81(d) * V + i]; }
static inline int nbr1(int i, int a, int sgn) {
  int d[4];
  for (int k = 0; k < 4; k++) d[k] = sgn * ev[a][k];
  return nbr(i, d);
}
// Fermion boundary sign for a path whose net temporal offset is dt from a site
// at time t (setup_offset.c:67-145): antiperiodic/periodic factor PBC if the
// path leaves [0, nt-1] in t, else 1.
static inline double bcs(int i, int dt) {
  int tt = ct_[i] + dt;
  return (tt < 0 || tt > nt - 1) ? (double)PBC : 1.0;
}
static inline int tau(int a) { return ev[a][3]; }

// ---------------------------------------------------------------------
// Lookup tables (setup_lambda.c:123-297, setup_offset.c:207-259)
static double perm_sign(int i, int j, int k, int l, int m) {
  int s[5] = {i, j, k, l, m}, inv = 0;
  for (int p = 0; p < 5; p++)
    for (int q = p + 1; q < 5; q++)
      if (s[p] > s[q]) inv++;
  return (inv % 2) ? -1.0 : 1.0;
}

static void setup_tables(void) {
  for (int mu = 0; mu < NUMLINK; mu++) {
    plaq_index[mu][mu] = -1;
    for (int nu = mu + 1; nu < NUMLINK; nu++) {
      int idx = mu * (NUMLINK - 1) - mu * (mu + 1) / 2 + nu - 1;
      plaq_index[mu][nu] = plaq_index[nu][mu] = idx;
    }
  }
  memset(perm, 0, sizeof(perm));
  for (int i = 0; i < 5; i++)
    for (int j = 0; j < 5; j++)
      for (int k = 0; k < 5; k++)
        for (int l = 0; l < 5; l++)
          for (int m = 0; m < 5; m++) {
            int s[5] = {i, j, k, l, m}, ok = 1;
            for (int p = 0; p < 5; p++)
              for (int q = p + 1; q < 5; q++)
                if (s[p] == s[q]) ok = 0;
            if (ok) perm[i][j][k][l][m] = perm_sign(i, j, k, l, m);
          }
  // Each term {a,b,c,d,e} is a permutation of 0..4 with a<b, d<e;
  // the three tables differ only in loop nesting order.
  int n = 0;
  for (int a = 0; a < 5; a++)
    for (int b = a + 1; b < 5; b++)
      for (int c = 0; c < 5; c++) {
        if (c == a || c == b) continue;
        for (int d = 0; d < 5; d++) {
          if (d == a || d == b || d == c) continue;
          for (int e = d + 1; e < 5; e++) {
            if (e == a || e == b || e == c) continue;
            int r[5] = {a, b, c, d, e};
            memcpy(pp_lookup[n++], r, sizeof(r));
          }
        }
      }
  n = 0;
  for (int d = 0; d < 5; d++)
    for (int e = d + 1; e < 5; e++)
      for (int c = 0; c < 5; c++) {
        if (c == d || c == e) continue;
        for (int a = 0; a < 5; a++) {
          if (a == c || a == d || a == e) continue;
          for (int b = a + 1; b < 5; b++) {
            if (b == c || b == d || b == e) continue;
            int r[5] = {a, b, c, d, e};
            memcpy(pm_lookup[n++], r, sizeof(r));
          }
        }
      }
  n = 0;
  for (int c = 0; c < 5; c++)
    for (int d = 0; d < 5; d++) {
      if (d == c) continue;
      for (int e = d + 1; e < 5; e++) {
        if (e == c) continue;
        for (int a = 0; a < 5; a++) {
          if (a == d || a == e || a == c) continue;
          for (int b = a + 1; b < 5; b++) {
            if (b == d || b == e || b == c) continue;
            int r[5] = {a, b, c, d, e};
            memcpy(fq_lookup[n++], r, sizeof(r));
          }
        }
      }
    }
  // Q-closed offsets: all vectors with 2 or 3 nonzero entries in {-1,0,1}^4,
  // one of each +/- pair, in the order of the descending loops at
  // setup_offset.c:225-250.
  int cnt = 0, v[4];
  for (v[0] = 1; v[0] >= -1; v[0]--)
    for (v[1] = 1; v[1] >= -1; v[1]--)
      for (v[2] = 1; v[2] >= -1; v[2]--)
        for (v[3] = 1; v[3] >= -1; v[3]--) {
          int k = abs(v[0]) + abs(v[1]) + abs(v[2]) + abs(v[3]);
          if (k != 2 && k != 3) continue;
          int dup = 0;
          for (int p = 0; p < cnt; p++)
            if (v[0] == -q_offset[p][0] && v[1] == -q_offset[p][1] &&
                v[2] == -q_offset[p][2] && v[3] == -q_offset[p][3])
              dup = 1;
          if (!dup) memcpy(q_offset[cnt++], v, sizeof(v));
        }
}

// U(N) generators with Tr(T_a T_b) = -delta_ab (setup_lambda.c:10-64)
static void setup_lambda(void) {
  memset(Lambda, 0, sizeof(Lambda));
  int count = 0;
  double r2 = 1.0 / sqrt(2.0);
  for (int i = 0; i < NCOL; i++)
    for (int j = i + 1; j < NCOL; j++) {
      Lambda[count].e[i][j] = cmk(0, r2);
      Lambda[count].e[j][i] = cmk(0, r2);
      Lambda[count + 1].e[i][j] = cmk(r2, 0);
      Lambda[count + 1].e[j][i] = cmk(-r2, 0);
      count += 2;
    }
  for (int i = 0; i < NCOL - 1; i++) {
    int j = NCOL * (NCOL - 1) + i, k = i + 1;
    double f = 1.0 / sqrt(k * (k + 1.0));
    for (int l = 0; l <= k; l++) Lambda[j].e[l][l] = cmk(0, f);
    Lambda[j].e[k][k] = cmk(0, -k * f);
  }
  mzero(&Lambda[DIMF - 1]);
  for (int i = 0; i < NCOL; i++)
    Lambda[DIMF - 1].e[i][i] = cmk(0, sqrt(1.0 / (double)NCOL));
}

// ---------------------------------------------------------------------
// RNG (generic/ranstuff.c:22-63, libraries/gaussrand.c:12-32)
static void prn_init(prn_t *p, int seed, int index) {
  int32_t s = seed;
  int32_t m = 69607 + 8 * index;
  for (int k = 0; k < 7; k++) {
    s = (int32_t)((uint32_t)m * (uint32_t)s + 12345u);
    p->r[k] = (uint64_t)((s >> 8) & 0xffffff);
  }
  s = (int32_t)((uint32_t)m * (uint32_t)s + 12345u);
  p->ic_state = (uint64_t)(int64_t)s;       // int -> unsigned long
  p->multiplier = (uint64_t)(100005 + 8 * index);
  p->addend = 12345;
  p->scale = (float)(1.0 / (double)0x1000000);
}
static double prn_next(prn_t *p) {
  int32_t t = (int32_t)((((p->r[5] >> 7) | (p->r[6] << 17)) ^
                         ((p->r[4] >> 1) | (p->r[5] << 23))) & 0xffffff);
  for (int k = 6; k > 0; k--) p->r[k] = p->r[k - 1];
  p->r[0] = (uint64_t)t;
  int32_t s = (int32_t)(uint32_t)(p->ic_state * p->multiplier + p->addend);
  p->ic_state = (uint64_t)(int64_t)s;
  float f = p->scale * (float)(t ^ ((s >> 8) & 0xffffff));
  return (double)f;
}
// Polar Box-Muller, density exp(-x^2) (variance 1/2); the pair cache is
// shared by all generators as in the reference.
static double gauss(prn_t *p) {
  if (g_iset == 0) {
    double v1, v2, r;
    do {
      v1 = 2.0 * prn_next(p) - 1.0;
      v2 = 2.0 * prn_next(p) - 1.0;
      r = v1 * v1 + v2 * v2;
    } while (r >= 1.0);
    double fac = sqrt(-log(r) / r);
    g_gset = v1 * fac;
    g_iset = 1;
    return v2 * fac;
  }
  g_iset = 0;
  return g_gset;
}

// ---------------------------------------------------------------------
// Determinant and inverse of an NCOL x NCOL complex matrix by LU with
// partial pivoting, pivot = first max of |re|+|im| in the column (the
// zgetrf/izamax convention used via determinant.c:10-85).
static double cabs1(cx a) { return fabs(a.re) + fabs(a.im); }
static cx cdiv(cx a, cx b) {
  double t = b.re * b.re + b.im * b.im;
  return cmk((a.re * b.re + a.im * b.im) / t, (a.im * b.re - a.re * b.im) / t);
}
static void lu(const mat *in, mat *lu_out, int piv[NCOL]) {
  mat a = *in;
  for (int k = 0; k < NCOL; k++) {
    int p = k;
    double best = cabs1(a.e[k][k]);
    for (int i = k + 1; i < NCOL; i++)
      if (cabs1(a.e[i][k]) > best) { best = cabs1(a.e[i][k]); p = i; }
    piv[k] = p;
    if (p != k)
      for (int j = 0; j < NCOL; j++) {
        cx t = a.e[k][j]; a.e[k][j] = a.e[p][j]; a.e[p][j] = t;
      }
    cx inv = cdiv(cmk(1, 0), a.e[k][k]);
    for (int i = k + 1; i < NCOL; i++) {
      a.e[i][k] = cmul(a.e[i][k], inv);
      for (int j = k + 1; j < NCOL; j++)
        a.e[i][j] = csub(a.e[i][j], cmul(a.e[i][k], a.e[k][j]));
    }
  }
  *lu_out = a;
}
static cx find_det(const mat *m) {
  mat a; int piv[NCOL];
  lu(m, &a, piv);
  cx det = cmk(1, 0);
  for (int i = 0; i < NCOL; i++) {
    det = cmul(det, a.e[i][i]);
    if (piv[i] != i) det = cmk(-det.re, -det.im);
  }
  return det;
}
static void invert(const mat *m, mat *out) {
  mat a; int piv[NCOL];
  lu(m, &a, piv);
  // Solve A X = I column by column: P A = L U.
  mat x;
  for (int col = 0; col < NCOL; col++) {
    cx b[NCOL];
    for (int i = 0; i < NCOL; i++) b[i] = cmk(i == col, 0);
    for (int k = 0; k < NCOL; k++)
      if (piv[k] != k) { cx t = b[k]; b[k] = b[piv[k]]; b[piv[k]] = t; }
    for (int i = 1; i < NCOL; i++)
      for (int j = 0; j < i; j++) b[i] = csub(b[i], cmul(a.e[i][j], b[j]));
    for (int i = NCOL - 1; i >= 0; i--) {
      for (int j = i + 1; j < NCOL; j++)
        b[i] = csub(b[i], cmul(a.e[i][j], b[j]));
      b[i] = cdiv(b[i], a.e[i][i]);
    }
    for (int i = 0; i < NCOL; i++) x.e[i][col] = b[i];
  }
  *out = x;
}

// ---------------------------------------------------------------------
// Link-derived fields
// plaqdet_ab(x) = det U_a(x)^* det U_b(x) det U_b(x+a)^* det U_a(x+b),
// ZW*_ab = plaqdet_ab (plaqdet_ab - 1)^*          (utilities.c:17-110)
static void compute_plaqdet(void) {
  for (int a = 0; a < NUMLINK; a++)
    for (int i = 0; i < V; i++) detU[a][i] = find_det(&links[i * NUMLINK + a]);
  for (int a = 0; a < NUMLINK; a++)
    for (int b = 0; b < NUMLINK; b++) {
      if (a == b) continue;
      for (int i = 0; i < V; i++) {
        int xa = nbr1(i, a, 1), xb = nbr1(i, b, 1);
        cx p = cmul(cconj(detU[a][i]), detU[b][i]);
        p = cmul(p, cconj(detU[b][xa]));
        p = cmul(detU[a][xb], p);
        plaqdet[a][b][i] = p;
        zwstar[a][b][i] = cmul(p, cconj(cadd(p, cmk(-1, 0))));
      }
    }
}
// U^-1 and (U^-1)^dagger per link (utilities.c:117-127)
static void compute_Uinv(void) {
  for (int i = 0; i < V; i++)
    for (int a = 0; a < NUMLINK; a++) {
      invert(&links[i * NUMLINK + a], &Uinv[a][i]);
      madj(&Uinv[a][i], &Udag_inv[a][i]);
    }
}
// DmuUmu = sum_a [U_a Ubar_a(x) - Ubar_a U_a(x-a)] + G sum_{a!=b}(plaqdet-1) I
// (action.c:15-62)
static void compute_DmuUmu(void) {
  for (int i = 0; i < V; i++) {
    mat d; mzero(&d);
    for (int a = 0; a < NUMLINK; a++) {
      int xm = nbr1(i, a, -1);
      maddr_prod(&d, 1.0, &links[i * NUMLINK + a], 0, &links[i * NUMLINK + a], 1);
      maddr_prod(&d, -1.0, &links[xm * NUMLINK + a], 1, &links[xm * NUMLINK + a], 0);
    }
    if (doG) {
      for (int a = 0; a < NUMLINK; a++)
        for (int b = 0; b < NUMLINK; b++) {
          if (a == b) continue;
          cx tc = cscale(G, cadd(plaqdet[a][b][i], cmk(-1, 0)));
          for (int j = 0; j < NCOL; j++) d.e[j][j] = cadd(d.e[j][j], tc);
        }
    }
    DmuUmu[i] = d;
  }
}
// F_ab(x) = U_a(x) U_b(x+a) - U_b(x) U_a(x+b), a<b (action.c:71-98)
static void compute_Fmunu(void) {
  for (int a = 0; a < NUMLINK; a++)
    for (int b = a + 1; b < NUMLINK; b++) {
      int p = plaq_index[a][b];
      for (int i = 0; i < V; i++) {
        int xa = nbr1(i, a, 1), xb = nbr1(i, b, 1);
        mat f; mzero(&f);
        maddr_prod(&f, 1.0, &links[i * NUMLINK + a], 0, &links[xa * NUMLINK + b], 0);
        maddr_prod(&f, -1.0, &links[i * NUMLINK + b], 0, &links[xb * NUMLINK + a], 0);
        Fmunu[p][i] = f;
      }
    }
}
static void compute_fields(void) {
  compute_plaqdet();
  compute_Uinv();
  compute_DmuUmu();
  compute_Fmunu();
}

#define LNK(i, a) (&links[(i) * NUMLINK + (a)])

// ---------------------------------------------------------------------
// Twisted fermion operator D (sign=+1) or D^dagger (sign=-1)
// (utilities.c:765-865).  D^dagger v = -(D v^dagger)^dagger componentwise.
// Each output component is the sum of the terms that land on it:
//   plaquette chi_ab : Dplus (utilities.c:141-221)
//                      + DbminusPtoP (427-510) + DbplusPtoP (334-414)
//   link psi_a       : Dminus (234-321) + DbplusStoL (522-545)
//                      + detStoL (562-628)
//   site eta         : DbminusLtoS (641-682) + detLtoS (698-756)
static tf_t *op_src;           // adjoint-converted source (scratch)
static cx *op_treta, *op_T[NUMLINK];

// antisymmetric plaquette component chi_ab (stored for a<b)
static inline void plaq_get(const tf_t *f, int a, int b, mat *out) {
  if (a < b) *out = f->plaq[plaq_index[a][b]];
  else {
    const mat *m = &f->plaq[plaq_index[a][b]];
    for (int i = 0; i < NCOL; i++)
      for (int j = 0; j < NCOL; j++)
        out->e[i][j] = cmk(-m->e[i][j].re, -m->e[i][j].im);
  }
}

static void fermion_op_impl(const tf_t *in, tf_t *out, int sign) {
  tf_t *src = op_src;
  for (int i = 0; i < V; i++) {
    if (sign == 1) src[i] = in[i];
    else {
      madj(&in[i].site, &src[i].site);
      for (int a = 0; a < NUMLINK; a++) madj(&in[i].link[a], &src[i].link[a]);
      for (int p = 0; p < NPLAQ; p++) madj(&in[i].plaq[p], &src[i].plaq[p]);
    }
  }
  for (int i = 0; i < V; i++) {
    op_treta[i] = mtrace(&src[i].site);
    if (doG)
      for (int a = 0; a < NUMLINK; a++)
        op_T[a][i] = mtrace_prod(&Uinv[a][i], 0, &src[i].link[a], 0);
  }

  for (int x = 0; x < V; x++) {
    tf_t o;
    memset(&o, 0, sizeof(o));
    // ---- plaquette outputs
    for (int a = 0; a < NUMLINK; a++)
      for (int b = a + 1; b < NUMLINK; b++) {
        mat *acc = &o.plaq[plaq_index[a][b]];
        int xa = nbr1(x, a, 1), xb = nbr1(x, b, 1);
        maddr_prod(acc, bcs(x, tau(a)), LNK(x, a), 0, &src[xa].link[b], 0);
        maddr_prod(acc, -1.0, &src[x].link[b], 0, LNK(xb, a), 0);
        maddr_prod(acc, -bcs(x, tau(b)), LNK(x, b), 0, &src[xb].link[a], 0);
        maddr_prod(acc, 1.0, &src[x].link[a], 0, LNK(xa, b), 0);
      }
    for (int j = 0; j < NTERMS; j++) {
      // DbplusPtoP: dest_ab += 1/2 eps_abcde [bc3 chi_de(x+a+b+c) Ubar_c(x+a+b)
      //                                   - bc2 Ubar_c(x-c) chi_de(x+a+b)]
      int a = pp_lookup[j][0], b = pp_lookup[j][1], c = pp_lookup[j][2];
      int d = pp_lookup[j][3], e = pp_lookup[j][4];
      double tr = 0.5 * perm[a][b][c][d][e];
      int d1[4], d2[4];
      for (int k = 0; k < 4; k++) {
        d1[k] = ev[a][k] + ev[b][k];
        d2[k] = d1[k] + ev[c][k];
      }
      int x1 = nbr(x, d1), x2 = nbr(x, d2), xmc = nbr1(x, c, -1);
      mat *acc = &o.plaq[plaq_index[a][b]];
      int ide = plaq_index[d][e];
      maddr_prod(acc, tr * bcs(x, tau(a) + tau(b) + tau(c)),
                 &src[x2].plaq[ide], 0, LNK(x1, c), 1);
      maddr_prod(acc, -tr * bcs(x, tau(a) + tau(b)), LNK(xmc, c), 1,
                 &src[x1].plaq[ide], 0);
    }
    for (int j = 0; j < NTERMS; j++) {
      // DbminusPtoP: dest_de += 1/2 eps_abcde [bc2(-a,-b) chi_ab(x-a-b) Ubar_c(x-a-b-c)
      //                                   - bc3(-a,-b,-c) Ubar_c(x-c) chi_ab(x-a-b-c)]
      int a = pm_lookup[j][0], b = pm_lookup[j][1], c = pm_lookup[j][2];
      int d = pm_lookup[j][3], e = pm_lookup[j][4];
      double tr = 0.5 * perm[a][b][c][d][e];
      int d1[4], d2[4];
      for (int k = 0; k < 4; k++) {
        d2[k] = -ev[a][k] - ev[b][k];
        d1[k] = d2[k] - ev[c][k];
      }
      int x1 = nbr(x, d1), x2 = nbr(x, d2), xmc = nbr1(x, c, -1);
      mat *acc = &o.plaq[plaq_index[d][e]];
      int iab = plaq_index[a][b];
      maddr_prod(acc, tr * bcs(x, -tau(a) - tau(b)), &src[x2].plaq[iab], 0,
                 LNK(x1, c), 1);
      maddr_prod(acc, -tr * bcs(x, -tau(a) - tau(b) - tau(c)), LNK(xmc, c), 1,
                 &src[x1].plaq[iab], 0);
    }
    // ---- link outputs
    for (int b = 0; b < NUMLINK; b++) {
      mat *acc = &o.link[b];
      int xb = nbr1(x, b, 1);
      for (int a = 0; a < NUMLINK; a++) {
        if (a == b) continue;
        // Dminus: U_a(x+b) chi_ab(x) - bc(-a) chi_ab(x-a) U_a(x-a)
        int xma = nbr1(x, a, -1);
        mat chi, chim;
        plaq_get(&src[x], a, b, &chi);
        plaq_get(&src[xma], a, b, &chim);
        maddr_prod(acc, 1.0, LNK(xb, a), 0, &chi, 0);
        maddr_prod(acc, -bcs(x, -tau(a)), &chim, 0, LNK(xma, a), 0);
      }
      // DbplusStoL: 1/2 [bc(b) eta(x+b) Ubar_b(x) - Ubar_b(x) eta(x)]
      maddr_prod(acc, 0.5 * bcs(x, tau(b)), &src[xb].site, 0, LNK(x, b), 1);
      maddr_prod(acc, -0.5, LNK(x, b), 1, &src[x].site, 0);
      if (doG) {
        // detStoL: -1/2 C2 G U_b^-1(x) sum_{c!=b} [D_cb(x) + bc(-c) D_bc(x-c)],
        // D_ab = Tr(eta) plaqdet_ab
        cx sum = cmk(0, 0);
        for (int c = 0; c < NUMLINK; c++) {
          if (c == b) continue;
          int xmc = nbr1(x, c, -1);
          sum = cadd(sum, cmul(op_treta[x], plaqdet[c][b][x]));
          sum = cadd(sum, cscale(bcs(x, -tau(c)),
                                 cmul(op_treta[xmc], plaqdet[b][c][xmc])));
        }
        madd_scaled(acc, cscale(-0.5 * C2 * G, sum), &Uinv[b][x], 0);
      }
    }
    // ---- site output
    {
      mat *acc = &o.site;
      for (int a = 0; a < NUMLINK; a++) {
        int xma = nbr1(x, a, -1);
        // DbminusLtoS: 1/2 [psi_a(x) Ubar_a(x) - bc(-a) Ubar_a(x-a) psi_a(x-a)]
        maddr_prod(acc, 0.5, &src[x].link[a], 0, LNK(x, a), 1);
        maddr_prod(acc, -0.5 * bcs(x, -tau(a)), LNK(xma, a), 1,
                   &src[xma].link[a], 0);
      }
      if (doG) {
        // detLtoS: -(i G/2 C2 sqrt(N)) sum_{a!=b} plaqdet_ab(x)
        //          [T_b(x) + bc(b) T_a(x+b)] Lambda_{DIMF-1}, T_a = Tr(U_a^-1 psi_a)
        cx sum = cmk(0, 0);
        for (int a = 0; a < NUMLINK; a++)
          for (int b = 0; b < NUMLINK; b++) {
            if (a == b) continue;
            int xb = nbr1(x, b, 1);
            cx t2 = cadd(op_T[b][x], cscale(bcs(x, tau(b)), op_T[a][xb]));
            sum = cadd(sum, cmul(plaqdet[a][b][x], t2));
          }
        double lg = 0.5 * C2 * G * sqrt((double)NCOL);
        cx trd = cmk(-lg * sum.im, lg * sum.re);  // i * lg * sum
        madd_scaled(acc, cmk(-trd.re, -trd.im), &Lambda[DIMF - 1], 0);
      }
    }
    if (sign == 1) out[x] = o;
    else {  // negate and conjugate
      madj(&o.site, &out[x].site);
      for (int a = 0; a < NUMLINK; a++) madj(&o.link[a], &out[x].link[a]);
      for (int p = 0; p < NPLAQ; p++) madj(&o.plaq[p], &out[x].plaq[p]);
      double *v = (double *)&out[x];
      for (size_t k = 0; k < sizeof(tf_t) / sizeof(double); k++) v[k] = -v[k];
    }
  }
}

// ---------------------------------------------------------------------
// Twist_Fermion BLAS-1 (library_util.c:53-168), field-wide
#define TFD (sizeof(tf_t) / sizeof(double))
static double tf_magsq_site(const tf_t *a) {
  double s = mrealtrace(&a->site, &a->site);
  for (int k = 0; k < NUMLINK; k++) s += mrealtrace(&a->link[k], &a->link[k]);
  for (int k = 0; k < NPLAQ; k++) s += mrealtrace(&a->plaq[k], &a->plaq[k]);
  return s;
}
static double tf_rdot_site(const tf_t *a, const tf_t *b) {
  double s = mrealtrace(&a->site, &b->site);
  for (int k = 0; k < NUMLINK; k++) s += mrealtrace(&a->link[k], &b->link[k]);
  for (int k = 0; k < NPLAQ; k++) s += mrealtrace(&a->plaq[k], &b->plaq[k]);
  return s;
}
static void tf_axpy(tf_t *y, double s, const tf_t *x) {  // y += s x
  double *yy = (double *)y;
  const double *xx = (const double *)x;
  for (size_t k = 0; k < TFD; k++) yy[k] += s * xx[k];
}
static void tf_scale(tf_t *y, double s) {
  double *yy = (double *)y;
  for (size_t k = 0; k < TFD; k++) yy[k] *= s;
}

static tf_t *tmp_tf;
// DSq = D^dagger D + fmass^2 (utilities.c:874-885)
static void DSq_impl(const tf_t *src, tf_t *dest) {
  fermion_op_impl(src, tmp_tf, 1);
  fermion_op_impl(tmp_tf, dest, -1);
  if (fmass > IMAG_TOL) {
    double f2 = fmass * fmass;
    for (int i = 0; i < V; i++) tf_axpy(&dest[i], f2, &src[i]);
  }
}

// ---------------------------------------------------------------------
// Multi-shift CG (Jegerlehner), zero initial guess (congrad_multi.c:26-221)
static int congrad_multi_impl(const tf_t *src, tf_t **psim, int MaxCG,
                              double errormin, double *size_r) {
  int N = Norder, iteration = 0;
  tf_t *r = malloc(sizeof(tf_t) * V), *p0 = malloc(sizeof(tf_t) * V);
  tf_t *mp = malloc(sizeof(tf_t) * V);
  tf_t **p = malloc(sizeof(tf_t *) * N);
  int conv[15];
  double zi[15], zim1[15], zip1[15], bi[15], bim1[15], al[15];
  for (int j = 1; j < N; j++) p[j] = malloc(sizeof(tf_t) * V);
  for (int j = 0; j < N; j++) conv[j] = 0;
  for (int i = 0; i < V; i++) {
    r[i] = src[i];
    p0[i] = src[i];
    memset(&psim[0][i], 0, sizeof(tf_t));
    for (int j = 1; j < N; j++) {
      memset(&psim[j][i], 0, sizeof(tf_t));
      p[j][i] = src[i];
    }
  }
  double source_norm = 0;
  for (int i = 0; i < V; i++) source_norm += tf_magsq_site(&src[i]);
  double rsq = source_norm, rsqstop = errormin * source_norm;
  for (int j = 0; j < N; j++) { zim1[j] = 1; zi[j] = 1; al[j] = 0; bim1[j] = 1; }

  for (int it = 0; it < MaxCG && rsq > rsqstop; it++) {
    DSq_impl(p0, mp);
    iteration++;
    total_iters++;
    for (int i = 0; i < V; i++) tf_axpy(&mp[i], shift_cur[0], &p0[i]);
    double cd = 0;
    for (int i = 0; i < V; i++) cd += tf_rdot_site(&p0[i], &mp[i]);
    bi[0] = -rsq / cd;
    zip1[0] = 1;
    for (int j = 1; j < N; j++) {
      if (conv[j]) continue;
      zip1[j] = zi[j] * zim1[j] * bim1[0];
      double c1 = bi[0] * al[0] * (zim1[j] - zi[j]);
      double c2 = zim1[j] * bim1[0] * (1 - (shift_cur[j] - shift_cur[0]) * bi[0]);
      zip1[j] /= c1 + c2;
      bi[j] = bi[0] * zip1[j] / zi[j];
    }
    for (int i = 0; i < V; i++) {
      tf_axpy(&psim[0][i], -bi[0], &p0[i]);
      for (int j = 1; j < N; j++)
        if (!conv[j]) tf_axpy(&psim[j][i], -bi[j], &p[j][i]);
    }
    for (int i = 0; i < V; i++) tf_axpy(&r[i], bi[0], &mp[i]);
    double rsqnew = 0;
    for (int i = 0; i < V; i++) rsqnew += tf_magsq_site(&r[i]);
    al[0] = rsqnew / rsq;
    for (int j = 1; j < N; j++)
      if (!conv[j]) al[j] = al[0] * zip1[j] * bi[j] / (zi[j] * bi[0]);
    for (int i = 0; i < V; i++) {
      tf_scale(&p0[i], al[0]);
      tf_axpy(&p0[i], zip1[0], &r[i]);
      for (int j = 1; j < N; j++)
        if (!conv[j]) {
          tf_scale(&p[j][i], al[j]);
          tf_axpy(&p[j][i], zip1[j], &r[i]);
        }
    }
    rsq = rsqnew;
    for (int j = 1; j < N; j++)
      if (!conv[j] && rsq * zip1[j] * zip1[j] <= rsqstop) conv[j] = 1;
    for (int j = 0; j < N; j++)
      if (!conv[j]) { bim1[j] = bi[j]; zim1[j] = zi[j]; zi[j] = zip1[j]; }
  }
  if (rsq > rsqstop && !quiet)
    printf(" multi CONGRAD not converged\n rsq = %.4g\n", rsq);
  *size_r = rsq;
  for (int j = 1; j < N; j++) free(p[j]);
  free(p); free(r); free(p0); free(mp);
  return iteration;
}

// ---------------------------------------------------------------------
// Fermion force (update_h.c:272-978).  For each pole n the force matrix on
// link mu at x is assembled from sol = psi_n and psol = amp4_n D psi_n:
//   SV (709-746), VP (756-894), detF (477-651), F1Q (272-362), F2Q (372-462);
// the pole sum is adjointed (943-958) and mom -= eps * F (965-969).
static mat *ff_acc[NUMLINK], *ff_fu[NUMLINK];
static tf_t *ff_psol;

static void detF_add(int eta_kind, const tf_t *solf, const tf_t *psolf,
                     int psi_from_sol, int sign) {
  // eta(x), psi_a(x) taken from the site/link components of the given fields
  // (eta_kind / psi_from_sol select sol or adjointed psol), see
  // assemble_fermion_force's two detF calls (update_h.c:899-905).
  double lg = 0.5 * C2 * G * (sign == 1 ? -1.0 : 1.0);
  cx *treta = malloc(sizeof(cx) * V);
  cx *T[NUMLINK];
  mat *UpsiU[NUMLINK];
  for (int a = 0; a < NUMLINK; a++) {
    T[a] = malloc(sizeof(cx) * V);
    UpsiU[a] = malloc(sizeof(mat) * V);
  }
  for (int i = 0; i < V; i++) {
    mat eta;
    if (eta_kind == 0) eta = solf[i].site; else madj(&psolf[i].site, &eta);
    treta[i] = mtrace(&eta);
    for (int a = 0; a < NUMLINK; a++) {
      mat psi, t;
      if (psi_from_sol) psi = solf[i].link[a]; else madj(&psolf[i].link[a], &psi);
      mzero(&t);
      maddr_prod(&t, 1.0, &Uinv[a][i], 0, &psi, 0);
      T[a][i] = mtrace(&t);
      mzero(&UpsiU[a][i]);
      maddr_prod(&UpsiU[a][i], 1.0, &t, 0, &Uinv[a][i], 0);
    }
  }
#define DD(a, b, y) cmul(treta[y], plaqdet[a][b][y])
  for (int x = 0; x < V; x++)
    for (int a = 0; a < NUMLINK; a++) {
      cx plaq_term = cmk(0, 0), inv_term = cmk(0, 0), adj_term = cmk(0, 0);
      int xa = nbr1(x, a, 1);
      for (int b = 0; b < NUMLINK; b++) {
        if (a == b) continue;
        int xb = nbr1(x, b, 1), xmb = nbr1(x, b, -1);
        int xamb = nbr1(xa, b, -1);   // x + e_a - e_b, two hops
        // D_ba(x) [T_a(x) + bc(a) T_b(x+a)]
        cx t2 = cadd(T[a][x], cscale(bcs(x, tau(a)), T[b][xa]));
        plaq_term = cadd(plaq_term, cmul(DD(b, a, x), t2));
        // D_ab(x-b) [T_b(x-b) + bc(-b) T_a(x)]
        t2 = cadd(T[b][xmb], cscale(bcs(x, -tau(b)), T[a][x]));
        plaq_term = cadd(plaq_term, cmul(DD(a, b, xmb), t2));
        // D_ab(x) [T_b(x) + bc(b) T_a(x+b)]
        t2 = cadd(T[b][x], cscale(bcs(x, tau(b)), T[a][xb]));
        adj_term = cadd(adj_term, cmul(DD(a, b, x), t2));
        // D_ba(x-b) [T_a(x-b) + bc(a)(x-b) T_b(x-b+a)]
        t2 = cadd(T[a][xmb], cscale(bcs(xmb, tau(a)), T[b][xamb]));
        adj_term = cadd(adj_term, cmul(DD(b, a, xmb), t2));
        // D_ba(x) + bc(-b) D_ab(x-b)
        inv_term = cadd(inv_term, DD(b, a, x));
        inv_term = cadd(inv_term, cscale(bcs(x, -tau(b)), DD(a, b, xmb)));
      }
      mat *f = &ff_fu[a][x];
      madd_scaled(f, cscale(lg, plaq_term), &Uinv[a][x], 0);
      // (s Udag_inv)^dagger = s^* Uinv
      madd_scaled(f, cconj(cscale(lg, adj_term)), &Udag_inv[a][x], 1);
      cx s = cscale(lg, inv_term);
      madd_scaled(f, cmk(-s.re, -s.im), &UpsiU[a][x], 0);
    }
#undef DD
  free(treta);
  for (int a = 0; a < NUMLINK; a++) { free(T[a]); free(UpsiU[a]); }
}

static void assemble_force(const tf_t *sol, const tf_t *psol) {
  // s* = sol components; p* = adjoint of psol components
  for (int x = 0; x < V; x++) {
    for (int mu = 0; mu < NUMLINK; mu++) {
      mat u; mzero(&u);
      int xm = nbr1(x, mu, 1);
      double bm = bcs(x, tau(mu));
      mat peta_x, peta_xm, ppsi;
      madj(&psol[x].site, &peta_x);
      madj(&psol[xm].site, &peta_xm);
      madj(&psol[x].link[mu], &ppsi);
      // SV: [psi_mu(x) eta(x+mu) - eta(x) psi_mu(x)] terms, then 1/2 (.)^dagger
      maddr_prod(&u, bm, &sol[x].link[mu], 0, &peta_xm, 0);
      maddr_prod(&u, -1.0, &peta_x, 0, &sol[x].link[mu], 0);
      maddr_prod(&u, -bm, &ppsi, 0, &sol[xm].site, 0);
      maddr_prod(&u, 1.0, &sol[x].site, 0, &ppsi, 0);
      mat f; madj(&u, &f);
      for (int i = 0; i < NCOL; i++)
        for (int j = 0; j < NCOL; j++) f.e[i][j] = cscale(0.5, f.e[i][j]);
      // VP
      for (int nu = 0; nu < NUMLINK; nu++) {
        if (nu == mu) continue;
        int xmn = nbr1(x, nu, -1);
        mat pchi_x, pchi_m, schi_x, schi_m, t, ppsi_n;
        plaq_get(&psol[x], mu, nu, &t); madj(&t, &pchi_x);
        plaq_get(&psol[xmn], mu, nu, &t); madj(&t, &pchi_m);
        // adjoint of an antisymmetric component: (-S)^dag = -(S^dag)
        plaq_get(&sol[x], mu, nu, &schi_x);
        plaq_get(&sol[xmn], mu, nu, &schi_m);
        // part 1: -bc(mu) s_psi_nu(x+mu) p_chi_munu(x) + p_chi_munu(x-nu) s_psi_nu(x-nu)
        maddr_prod(&f, -bm, &sol[xm].link[nu], 0, &pchi_x, 0);
        maddr_prod(&f, 1.0, &pchi_m, 0, &sol[xmn].link[nu], 0);
        // part 2: bc(mu) p_psi_nu(x+mu) s_chi_munu(x) - s_chi_munu(x-nu) p_psi_nu(x-nu)
        madj(&psol[xm].link[nu], &ppsi_n);
        maddr_prod(&f, bm, &ppsi_n, 0, &schi_x, 0);
        madj(&psol[xmn].link[nu], &ppsi_n);
        maddr_prod(&f, -1.0, &schi_m, 0, &ppsi_n, 0);
      }
      ff_fu[mu][x] = f;
    }
  }
  if (doG) {
    detF_add(1, sol, psol, 1, 1);   // eta = psol site^dag, psi = sol links
    detF_add(0, sol, psol, 0, -1);  // eta = sol site, psi = psol links^dag
  }
  // Q-closed pieces, accumulated on link c
  for (int x = 0; x < V; x++)
    for (int j = 0; j < NTERMS; j++) {
      int a = fq_lookup[j][0], b = fq_lookup[j][1], c = fq_lookup[j][2];
      int d = fq_lookup[j][3], e = fq_lookup[j][4];
      int iab = plaq_index[a][b], ide = plaq_index[d][e];
      int dab[4], dabc[4];
      for (int k = 0; k < 4; k++) {
        dab[k] = ev[a][k] + ev[b][k];
        dabc[k] = dab[k] + ev[c][k];
      }
      int mdab[4] = {-dab[0], -dab[1], -dab[2], -dab[3]};
      int xc = nbr1(x, c, 1), xabc = nbr(x, dabc), xmab = nbr(x, mdab);
      double bcc = bcs(x, tau(c));
      double b3 = bcs(x, tau(a) + tau(b) + tau(c));
      double b2m = bcs(x, -tau(a) - tau(b));
      mat t, pa, pb;
      // F1Q: p_chi_de(x+a+b+c) [s_chi_ab(x+c)] , [s_chi_ab(x-a-b)] p_chi_de(x+c)
      double pm = perm[d][e][c][a][b];
      mzero(&t);
      madj(&psol[xabc].plaq[ide], &pa);
      madj(&psol[xc].plaq[ide], &pb);
      maddr_prod(&t, pm * b3 * bcc, &pa, 0, &sol[xc].plaq[iab], 0);
      maddr_prod(&t, -pm * b2m * bcc, &sol[xmab].plaq[iab], 0, &pb, 0);
      madd_scaled(&ff_fu[c][x], cmk(-0.5, 0), &t, 1);
      // F2Q: p_chi_ab(x-a-b) [s_chi_de(x+c)] , [s_chi_de(x+a+b+c)] p_chi_ab(x+c)
      pm = perm[a][b][c][d][e];
      mzero(&t);
      madj(&psol[xmab].plaq[iab], &pa);
      madj(&psol[xc].plaq[iab], &pb);
      maddr_prod(&t, pm * b2m * bcc, &pa, 0, &sol[xc].plaq[ide], 0);
      maddr_prod(&t, -pm * b3 * bcc, &sol[xabc].plaq[ide], 0, &pb, 0);
      madd_scaled(&ff_fu[c][x], cmk(-0.5, 0), &t, 1);
    }
}

static double fermion_force_impl(double eps, tf_t **sol, const double *amps,
                                 int nord) {
  for (int mu = 0; mu < NUMLINK; mu++)
    for (int x = 0; x < V; x++) mzero(&ff_acc[mu][x]);
  for (int n = 0; n < nord; n++) {
    fermion_op_impl(sol[n], ff_psol, 1);
    for (int i = 0; i < V; i++) tf_scale(&ff_psol[i], amps[n]);
    assemble_force(sol[n], ff_psol);
    for (int mu = 0; mu < NUMLINK; mu++)
      for (int x = 0; x < V; x++)
        madd_scaled(&ff_acc[mu][x], cmk(1, 0), &ff_fu[mu][x], 1);
  }
  double ret = 0;
  for (int x = 0; x < V; x++)
    for (int mu = 0; mu < NUMLINK; mu++) {
      madd_scaled(&mom[x * NUMLINK + mu], cmk(-eps, 0), &ff_acc[mu][x], 0);
      ret += mrealtrace(&ff_acc[mu][x], &ff_acc[mu][x]);
    }
  return eps * sqrt(ret) / V;
}

// ---------------------------------------------------------------------
// Determinant force (det_force.c:6-76), active when kappa_u1 > 0
static double det_force_impl(double eps) {
  double ret = 0;
  for (int x = 0; x < V; x++)
    for (int a = 0; a < NUMLINK; a++) {
      cx s = cmk(0, 0);
      for (int b = 0; b < NUMLINK; b++) {
        if (a == b) continue;
        s = cadd(s, zwstar[b][a][x]);
        s = cadd(s, zwstar[a][b][nbr1(x, b, -1)]);
      }
      mat f; mzero(&f);
      madd_scaled(&f, cscale(kappa_u1, s), &Uinv[a][x], 0);
      madd_scaled(&mom[x * NUMLINK + a], cmk(-eps, 0), &f, 1);
      ret += mrealtrace(&f, &f);
    }
  return ret;
}

// Gauge force (update_h.c:15-261)
static double gauge_force_impl(double eps) {
  double ret = 0;
  cx *trd = malloc(sizeof(cx) * V);
  for (int x = 0; x < V; x++) trd[x] = mtrace(&DmuUmu[x]);
  for (int x = 0; x < V; x++)
    for (int mu = 0; mu < NUMLINK; mu++) {
      mat f; mzero(&f);
      int xm = nbr1(x, mu, 1);
      // Ubar_mu(x) DmuUmu(x) - DmuUmu(x+mu) Ubar_mu(x)
      maddr_prod(&f, 1.0, LNK(x, mu), 1, &DmuUmu[x], 0);
      maddr_prod(&f, -1.0, &DmuUmu[xm], 0, LNK(x, mu), 1);
      if (doG) {
        // G U_mu^-1(x) sum_nu [Tr D(x) plaqdet_numu(x) + Tr D(x-nu) plaqdet_munu(x-nu)]
        cx s = cmk(0, 0);
        for (int nu = 0; nu < NUMLINK; nu++) {
          if (nu == mu) continue;
          int xmn = nbr1(x, nu, -1);
          s = cadd(s, cmul(trd[x], plaqdet[nu][mu][x]));
          s = cadd(s, cmul(trd[xmn], plaqdet[mu][nu][xmn]));
        }
        madd_scaled(&f, cscale(G, s), &Uinv[mu][x], 0);
      }
      // 2 sum_nu [U_nu(x+mu) F_munu(x)^dag - F_munu(x-nu)^dag U_nu(x-nu)]
      for (int nu = 0; nu < NUMLINK; nu++) {
        if (nu == mu) continue;
        int xmn = nbr1(x, nu, -1);
        mat Fx = Fmunu[plaq_index[mu][nu]][x], Fm = Fmunu[plaq_index[mu][nu]][xmn];
        double sg = (mu < nu) ? 1.0 : -1.0;
        maddr_prod(&f, 2.0 * sg, LNK(xm, nu), 0, &Fx, 1);
        maddr_prod(&f, -2.0 * sg, &Fm, 1, LNK(xmn, nu), 0);
      }
      if (bmass > IMAG_TOL) {
        double dmu = 2.0 * bmass * bmass / NCOL;
        double tr = mrealtrace(LNK(x, mu), LNK(x, mu)) / NCOL - 1.0;
        madd_scaled(&f, cmk(dmu * tr, 0), LNK(x, mu), 1);
      }
      madd_scaled(&mom[x * NUMLINK + mu], cmk(-kappa * eps, 0), &f, 1);
      ret += mrealtrace(&f, &f);
    }
  free(trd);
  ret *= kappa * kappa;
  if (kappa_u1 > IMAG_TOL) ret += det_force_impl(eps);
  return eps * sqrt(ret) / V;
}

// ---------------------------------------------------------------------
// Actions (action.c:106-328)
static double gauge_action_impl(void) {
  double s = 0;
  for (int x = 0; x < V; x++) {
    mat t; mzero(&t);
    maddr_prod(&t, 0.5 * C2, &DmuUmu[x], 0, &DmuUmu[x], 0);
    for (int p = 0; p < NPLAQ; p++) maddr_prod(&t, 2.0, &Fmunu[p][x], 1, &Fmunu[p][x], 0);
    s += mtrace(&t).re;
  }
  return s * kappa;
}
static double bmass_action_impl(void) {
  double s = 0;
  for (int x = 0; x < V; x++)
    for (int a = 0; a < NUMLINK; a++) {
      double tr = mrealtrace(LNK(x, a), LNK(x, a)) / NCOL - 1.0;
      s += tr * tr;
    }
  return s * kappa * bmass * bmass;
}
static double det_action_impl(void) {
  double s = 0;
  for (int x = 0; x < V; x++)
    for (int a = 0; a < NUMLINK; a++)
      for (int b = a + 1; b < NUMLINK; b++) {
        double re = plaqdet[a][b][x].re, im = plaqdet[a][b][x].im;
        s += (re - 1.0) * (re - 1.0) + im * im;
      }
  return s * kappa_u1;
}
static double fermion_action_impl(const tf_t *src, tf_t **sol) {
  double s = 0;
  for (int x = 0; x < V; x++) {
    s += ampdeg4 * tf_magsq_site(&src[x]);
    for (int j = 0; j < Norder; j++) s += amp4[j] * tf_rdot_site(&src[x], &sol[j][x]);
  }
  return s;
}
static double mom_action_impl(void) {
  double s = 0;
  for (int k = 0; k < V * NUMLINK; k++) s += mrealtrace(&mom[k], &mom[k]);
  return s;
}
// parts[0..5] = gauge, bmass, det, fermion(sum over roots), mom, total
static double action_impl(tf_t **src, tf_t ***sol, double *parts) {
  double g = gauge_action_impl(), bm = 0, dt = 0, f = 0;
  if (bmass > IMAG_TOL) bm = bmass_action_impl();
  if (kappa_u1 > IMAG_TOL) dt = det_action_impl();
  double tot = g + bm + dt;
  if (!quiet) printf("action: gauge %.8g bmass %.8g det %.8g ", g, bm, dt);
  for (int n = 0; n < Nroot; n++) {
    double fa = fermion_action_impl(src[n], sol[n]);
    if (!quiet) printf("fermion%d %.8g ", n, fa);
    f += fa;
    tot += fa;
  }
  double p = mom_action_impl();
  tot += p;
  if (!quiet) printf("mom %.8g sum %.8g\n", p, tot);
  if (parts) {
    parts[0] = g; parts[1] = bm; parts[2] = dt; parts[3] = f; parts[4] = p;
    parts[5] = tot;
  }
  return tot;
}

// ---------------------------------------------------------------------
// Heatbath (grsource.c:44-64, 200-303)
static void ranmom_impl(void) {
  for (int x = 0; x < V; x++)
    for (int mu = 0; mu < NUMLINK; mu++) {
      mat *m = &mom[x * NUMLINK + mu];
      mzero(m);
      for (int j = 0; j < DIMF; j++) {
        double gr = gauss(&site_prn[x]);
        double gi = gauss(&site_prn[x]);
        madd_scaled(m, cmk(gr, gi), &Lambda[j], 0);
      }
    }
}
static int grsource_impl(tf_t *src) {
  for (int x = 0; x < V; x++) {
    memset(&src[x], 0, sizeof(tf_t));
    for (int j = 0; j < DIMF; j++) {
      double gr = gauss(&site_prn[x]), gi = gauss(&site_prn[x]);
      madd_scaled(&src[x].site, cmk(gr, gi), &Lambda[j], 0);
      for (int mu = 0; mu < NUMLINK; mu++) {
        gr = gauss(&site_prn[x]); gi = gauss(&site_prn[x]);
        madd_scaled(&src[x].link[mu], cmk(gr, gi), &Lambda[j], 0);
      }
      for (int p = 0; p < NPLAQ; p++) {
        gr = gauss(&site_prn[x]); gi = gauss(&site_prn[x]);
        madd_scaled(&src[x].plaq[p], cmk(gr, gi), &Lambda[j], 0);
      }
    }
  }
  tf_t **psim = malloc(sizeof(tf_t *) * Norder);
  for (int j = 0; j < Norder; j++) psim[j] = malloc(sizeof(tf_t) * V);
  for (int j = 0; j < Norder; j++) shift_cur[j] = shift8[j];
  double size_r;
  int it = congrad_multi_impl(src, psim, niter, rsqmin, &size_r);
  for (int x = 0; x < V; x++) {
    tf_scale(&src[x], ampdeg8);
    for (int j = 0; j < Norder; j++) tf_axpy(&src[x], amp8[j], &psim[j][x]);
  }
  for (int j = 0; j < Norder; j++) free(psim[j]);
  free(psim);
  return it;
}

// ---------------------------------------------------------------------
// One RHMC trajectory: Omelyan two-level integrator, Metropolis test
// (update_o.c:20-282).  stats[]: 0 start S, 1 end S, 2 dS, 3 accepted,
// 4 iters, 5 gauge-force avg, 6 gauge-force max, 7.. fermion avg/max per root,
// 11..16 start action parts, 17..22 end action parts
static void update_u_impl(double eps) {
  for (int k = 0; k < V * NUMLINK; k++)
    madd_scaled(&links[k], cmk(eps, 0), &mom[k], 0);
  compute_fields();
}
static double gnorm, max_gf, fnorm[4], max_ff[4];
static double update_gauge_step_impl(double eps) {
  int n = nsteps[1];
  double norm = gauge_force_impl(eps * OM_LAMBDA);
  for (int i = 1; i <= n; i++) {
    update_u_impl(0.5 * eps);
    norm += gauge_force_impl(eps * OM_LAMBDA_MID);
    update_u_impl(0.5 * eps);
    if (i < n) norm += gauge_force_impl(eps * OM_TWO_LAMBDA);
    else norm += gauge_force_impl(eps * OM_LAMBDA);
  }
  return norm / n;
}
static void track_f(int n, double tr) {
  fnorm[n] += tr;
  if (tr > max_ff[n]) max_ff[n] = tr;
}
static int update_step_impl(tf_t **src, tf_t ***psim) {
  int iters = 0;
  double f_eps = traj_length / (double)nsteps[0];
  double g_eps = f_eps / (double)(2.0 * nsteps[1]), rs, tr;
  for (int n = 0; n < Nroot; n++)
    track_f(n, fermion_force_impl(f_eps * OM_LAMBDA, psim[n], amp4, Norder));
  for (int s = 1; s <= nsteps[0]; s++) {
    tr = update_gauge_step_impl(g_eps);
    gnorm += tr;
    if (tr > max_gf) max_gf = tr;
    for (int n = 0; n < Nroot; n++) {
      iters += congrad_multi_impl(src[n], psim[n], niter, rsqmin, &rs);
      track_f(n, fermion_force_impl(f_eps * OM_LAMBDA_MID, psim[n], amp4, Norder));
    }
    tr = update_gauge_step_impl(g_eps);
    gnorm += tr;
    if (tr > max_gf) max_gf = tr;
    for (int n = 0; n < Nroot; n++) {
      iters += congrad_multi_impl(src[n], psim[n], niter, rsqmin, &rs);
      double e = (s < nsteps[0]) ? f_eps * OM_TWO_LAMBDA : f_eps * OM_LAMBDA;
      track_f(n, fermion_force_impl(e, psim[n], amp4, Norder));
    }
  }
  return iters;
}
static int update_impl(double *stats) {
  int iters = 0;
  tf_t **gsrc = malloc(sizeof(tf_t *) * Nroot);
  tf_t ***psim = malloc(sizeof(tf_t **) * Nroot);
  for (int n = 0; n < Nroot; n++) {
    gsrc[n] = malloc(sizeof(tf_t) * V);
    psim[n] = malloc(sizeof(tf_t *) * Norder);
    for (int j = 0; j < Norder; j++) psim[n][j] = malloc(sizeof(tf_t) * V);
  }
  ranmom_impl();
  for (int n = 0; n < Nroot; n++) iters += grsource_impl(gsrc[n]);
  for (int j = 0; j < Norder; j++) shift_cur[j] = shift4[j];
  double rs, p0[6], p1[6];
  for (int n = 0; n < Nroot; n++)
    iters += congrad_multi_impl(gsrc[n], psim[n], niter, rsqmin, &rs);
  double startaction = action_impl(gsrc, psim, p0);
  gnorm = max_gf = 0;
  for (int n = 0; n < Nroot; n++) fnorm[n] = max_ff[n] = 0;
  memcpy(old_links, links, sizeof(mat) * V * NUMLINK);
  memcpy(old_mom, mom, sizeof(mat) * V * NUMLINK);
  iters += update_step_impl(gsrc, psim);
  double endaction = action_impl(gsrc, psim, p1);
  double change = endaction - startaction;
  if (fabs(change) > 1e20) {
    if (!quiet) printf("WARNING: Correcting `Apparent Overflow: Delta S = %.4g\n", change);
    change = 1.0e20;
  }
  double xr = prn_next(&node_prn);
  int accept = 1;
  if (exp(-change) < xr) {
    accept = 0;
    if (traj_length > 0.0) {
      memcpy(links, old_links, sizeof(mat) * V * NUMLINK);
      for (int k = 0; k < V * NUMLINK; k++) {
        mom[k] = old_mom[k];
        double *v = (double *)&mom[k];
        for (size_t q = 0; q < sizeof(mat) / sizeof(double); q++) v[q] = -v[q];
      }
      compute_fields();
    }
    if (!quiet)
      printf("REJECT: delta S = %.4g start S = %.12g end S = %.12g\n", change,
             startaction, endaction);
  } else if (!quiet) {
    printf("ACCEPT: delta S = %.4g start S = %.12g end S = %.12g\n", change,
           startaction, endaction);
  }
  if (!quiet) {
    printf("IT_PER_TRAJ %d\n", iters);
    printf("MONITOR_FORCE_GAUGE    %.4g %.4g\n", gnorm / (2.0 * nsteps[0]), max_gf);
    for (int n = 0; n < Nroot; n++)
      printf("MONITOR_FORCE_FERMION%d %.4g %.4g\n", n, fnorm[n] / (2.0 * nsteps[0]),
             max_ff[n]);
    fflush(stdout);
  }
  if (stats) {
    stats[0] = startaction; stats[1] = endaction; stats[2] = change;
    stats[3] = accept; stats[4] = iters;
    stats[5] = gnorm / (2.0 * nsteps[0]); stats[6] = max_gf;
    for (int n = 0; n < 2; n++) {
      stats[7 + 2 * n] = n < Nroot ? fnorm[n] / (2.0 * nsteps[0]) : 0;
      stats[8 + 2 * n] = n < Nroot ? max_ff[n] : 0;
    }
    for (int k = 0; k < 6; k++) { stats[11 + k] = p0[k]; stats[17 + k] = p1[k]; }
  }
  for (int n = 0; n < Nroot; n++) {
    free(gsrc[n]);
    for (int j = 0; j < Norder; j++) free(psim[n][j]);
    free(psim[n]);
  }
  free(gsrc); free(psim);
  return iters;
}

// ---------------------------------------------------------------------
// Plaquette (plaq.c:6-51): ReTr[Ubar_a(x+b) Ubar_b(x) U_a(x) U_b(x+a)]
static void plaquette_impl(double *ss, double *st) {
  double sss = 0, sst = 0;
  for (int a = 1; a < NUMLINK; a++)
    for (int b = 0; b < a; b++)
      for (int x = 0; x < V; x++) {
        int xa = nbr1(x, a, 1), xb = nbr1(x, b, 1);
        mat t, t2; mzero(&t); mzero(&t2);
        maddr_prod(&t, 1.0, LNK(x, b), 1, LNK(x, a), 0);
        maddr_prod(&t2, 1.0, &t, 0, LNK(xa, b), 0);
        double v = mrealtrace(LNK(xb, a), &t2);
        if (a == 3 || b == 3) sst += v; else sss += v;
      }
  *ss = sss / (6.0 * V);
  *st = sst / (4.0 * V);
}

// =====================================================================
// Exported C API (prefix orc_).  Fields cross the boundary in the
// reference's AoS layout: Twist_Fermion = 16 matrices per site
// (site, link[5], plaq[10]), matrix = NCOL x NCOL complex, row-major,
// (re, im) interleaved; links/momenta = 5 matrices per site.
// =====================================================================
static void free_all(void) {
  if (!cx_) return;
  free(cx_); free(cy_); free(cz_); free(ct_); cx_ = NULL;
  for (int g = 0; g < NGATHER; g++) free(gat[g]);
  free(nb81); free(links); free(mom); free(old_links); free(old_mom);
  for (int a = 0; a < NUMLINK; a++) {
    free(detU[a]); free(Uinv[a]); free(Udag_inv[a]); free(op_T[a]);
    free(ff_acc[a]); free(ff_fu[a]);
    for (int b = 0; b < NUMLINK; b++) { free(plaqdet[a][b]); free(zwstar[a][b]); }
  }
  free(DmuUmu);
  for (int p = 0; p < NPLAQ; p++) free(Fmunu[p]);
  free(site_prn); free(op_src); free(op_treta); free(tmp_tf); free(ff_psol);
}

int orc_ncol(void) { return NCOL; }
int orc_volume(void) { return V; }
int orc_num_gathers(void) { return NGATHER; }
void orc_set_quiet(int q) { quiet = q; }

// Returns 0 on success.  Parameters as in the reference input file
// (setup.c:16-119, 307-556).
int orc_init(int nx_, int ny_, int nz_, int nt_, int PBC_, int iseed_,
             int Nroot_, int Norder_, double lambda_, double kappa_u1_,
             double bmass_, double fmass_, double G_, int niter_,
             double error_per_site, double traj_length_, int nstep,
             int nstep_gauge) {
  free_all();
  nx = nx_; ny = ny_; nz = nz_; nt = nt_; PBC = PBC_; iseed_g = iseed_;
  V = nx * ny * nz * nt;
  if (V % 2 != 0 || (PBC != 1 && PBC != -1)) return 1;
  if (Norder_ != 15 || (Nroot_ != 1 && Nroot_ != 2)) return 2;
  Nroot = Nroot_; Norder = Norder_;
  lambda_c = lambda_; kappa_u1 = kappa_u1_; bmass = bmass_; fmass = fmass_;
  G = G_; doG = G > IMAG_TOL;
  kappa = (double)NCOL * 0.25 / lambda_c;
  niter = niter_; rsqmin = error_per_site * error_per_site;
  traj_length = traj_length_; nsteps[0] = nstep; nsteps[1] = nstep_gauge;
  int r = Nroot - 1;
  ampdeg4 = susy_rhmc15_ampdeg4[r]; ampdeg8 = susy_rhmc15_ampdeg8[r];
  for (int j = 0; j < 15; j++) {
    amp4[j] = susy_rhmc15_amp4[r][j]; shift4[j] = susy_rhmc15_shift4[r][j];
    amp8[j] = susy_rhmc15_amp8[r][j]; shift8[j] = susy_rhmc15_shift8[r][j];
    shift_cur[j] = shift4[j];
  }
  total_iters = 0;
  g_iset = 0;
  setup_tables();
  setup_lambda();

  cx_ = malloc(sizeof(int) * V); cy_ = malloc(sizeof(int) * V);
  cz_ = malloc(sizeof(int) * V); ct_ = malloc(sizeof(int) * V);
  site_prn = malloc(sizeof(prn_t) * V);
  for (int t = 0; t < nt; t++)
    for (int z = 0; z < nz; z++)
      for (int y = 0; y < ny; y++)
        for (int x = 0; x < nx; x++) {
          int i = site_index(x, y, z, t);
          cx_[i] = x; cy_[i] = y; cz_[i] = z; ct_[i] = t;
          prn_init(&site_prn[i], iseed_g, x + nx * (y + ny * (z + nz * t)));
        }
  prn_init(&node_prn, iseed_g, V);
  nb81 = malloc(sizeof(int) * 81 * V);
  int d[4];
  for (d[3] = -1; d[3] <= 1; d[3]++)
    for (d[2] = -1; d[2] <= 1; d[2]++)
      for (d[1] = -1; d[1] <= 1; d[1]++)
        for (d[0] = -1; d[0] <= 1; d[0]++)
          for (int i = 0; i < V; i++)
            nb81[off81(d) * V + i] =
                site_index((cx_[i] + nx + d[0]) % nx, (cy_[i] + ny + d[1]) % ny,
                           (cz_[i] + nz + d[2]) % nz, (ct_[i] + nt + d[3]) % nt);
  // Gather ids: 0-3 = +x,+y,+z,+t, 7-4 = -x,-y,-z,-t (com_vanilla.c:322-399);
  // 8+2a / 9+2a = x +/- e_a (setup_offset.c:175-177);
  // 18+2p / 19+2p = x +/- q_offset[p] (setup_offset.c:257-259)
  for (int g = 0; g < NGATHER; g++) gat[g] = malloc(sizeof(int) * V);
  for (int k = 0; k < 4; k++) {
    int e[4] = {0, 0, 0, 0}, m[4] = {0, 0, 0, 0};
    e[k] = 1; m[k] = -1;
    for (int i = 0; i < V; i++) { gat[k][i] = nbr(i, e); gat[7 - k][i] = nbr(i, m); }
  }
  for (int a = 0; a < NUMLINK; a++)
    for (int i = 0; i < V; i++) {
      gat[8 + 2 * a][i] = nbr1(i, a, 1);
      gat[9 + 2 * a][i] = nbr1(i, a, -1);
    }
  for (int p = 0; p < NQLINK; p++) {
    int m[4] = {-q_offset[p][0], -q_offset[p][1], -q_offset[p][2], -q_offset[p][3]};
    for (int i = 0; i < V; i++) {
      gat[18 + 2 * p][i] = nbr(i, q_offset[p]);
      gat[19 + 2 * p][i] = nbr(i, m);
    }
  }
  links = calloc(V * NUMLINK, sizeof(mat));
  mom = calloc(V * NUMLINK, sizeof(mat));
  old_links = calloc(V * NUMLINK, sizeof(mat));
  old_mom = calloc(V * NUMLINK, sizeof(mat));
  for (int a = 0; a < NUMLINK; a++) {
    detU[a] = malloc(sizeof(cx) * V);
    Uinv[a] = malloc(sizeof(mat) * V); Udag_inv[a] = malloc(sizeof(mat) * V);
    op_T[a] = malloc(sizeof(cx) * V);
    ff_acc[a] = malloc(sizeof(mat) * V); ff_fu[a] = malloc(sizeof(mat) * V);
    for (int b = 0; b < NUMLINK; b++) {
      plaqdet[a][b] = calloc(V, sizeof(cx)); zwstar[a][b] = calloc(V, sizeof(cx));
    }
  }
  DmuUmu = malloc(sizeof(mat) * V);
  for (int p = 0; p < NPLAQ; p++) Fmunu[p] = malloc(sizeof(mat) * V);
  op_src = malloc(sizeof(tf_t) * V);
  op_treta = malloc(sizeof(cx) * V);
  tmp_tf = malloc(sizeof(tf_t) * V);
  ff_psol = malloc(sizeof(tf_t) * V);
  // cold start (io_helpers.c:95-111)
  for (int k = 0; k < V * NUMLINK; k++)
    for (int j = 0; j < NCOL; j++) links[k].e[j][j] = cmk(1, 0);
  compute_fields();
  return 0;
}

void orc_get_gather(int g, int *out) { memcpy(out, gat[g], sizeof(int) * V); }
void orc_get_coords(int *xyzt) {
  for (int i = 0; i < V; i++) {
    xyzt[4 * i] = cx_[i]; xyzt[4 * i + 1] = cy_[i];
    xyzt[4 * i + 2] = cz_[i]; xyzt[4 * i + 3] = ct_[i];
  }
}
void orc_get_q_offsets(int *out) { memcpy(out, q_offset, sizeof(q_offset)); }
void orc_get_lookups(int *pp, int *pm, int *fq) {
  memcpy(pp, pp_lookup, sizeof(pp_lookup));
  memcpy(pm, pm_lookup, sizeof(pm_lookup));
  memcpy(fq, fq_lookup, sizeof(fq_lookup));
}
void orc_get_lambda(double *out) { memcpy(out, Lambda, sizeof(Lambda)); }

void orc_set_links(const double *l) {
  memcpy(links, l, sizeof(mat) * V * NUMLINK);
  compute_fields();
}
void orc_get_links(double *l) { memcpy(l, links, sizeof(mat) * V * NUMLINK); }
void orc_set_mom(const double *m) { memcpy(mom, m, sizeof(mat) * V * NUMLINK); }
void orc_get_mom(double *m) { memcpy(m, mom, sizeof(mat) * V * NUMLINK); }
// Derived fields for checks: plaqdet (5x5xV complex), Uinv (5xV mat),
// DmuUmu (V mat), Fmunu (10xV mat)
void orc_get_fields(double *pd, double *ui, double *dm, double *fm) {
  if (pd)
    for (int a = 0; a < NUMLINK; a++)
      for (int b = 0; b < NUMLINK; b++)
        memcpy(pd + 2 * (size_t)V * (a * NUMLINK + b), plaqdet[a][b], sizeof(cx) * V);
  if (ui)
    for (int a = 0; a < NUMLINK; a++)
      memcpy(ui + (size_t)a * V * 2 * DIMF, Uinv[a], sizeof(mat) * V);
  if (dm) memcpy(dm, DmuUmu, sizeof(mat) * V);
  if (fm)
    for (int p = 0; p < NPLAQ; p++)
      memcpy(fm + (size_t)p * V * 2 * DIMF, Fmunu[p], sizeof(mat) * V);
}
void orc_det_inv(const double *m, double *det, double *inv) {
  cx d = find_det((const mat *)m);
  det[0] = d.re; det[1] = d.im;
  invert((const mat *)m, (mat *)inv);
}

// MILC binary gauge file, serial natural order (io_lat.c:826-1006,
// io_lat_utils.c:519-618).  Returns 0 if the file's checksums match.
int orc_load_config(const char *path, unsigned int *sums) {
  FILE *fp = fopen(path, "rb");
  if (!fp) return -1;
  int32_t magic, dims[4], order;
  char stamp[64];
  uint32_t s29, s31;
  int swap = 0;
  if (fread(&magic, 4, 1, fp) != 1) { fclose(fp); return -2; }
  if (magic != 20103) {
    uint32_t u = (uint32_t)magic;
    u = (u >> 24) | ((u >> 8) & 0xff00) | ((u << 8) & 0xff0000) | (u << 24);
    if ((int32_t)u != 20103) { fclose(fp); return -3; }
    swap = 1;
  }
#define BSWAP(v) do { uint32_t _u = (uint32_t)(v); \
  _u = (_u >> 24) | ((_u >> 8) & 0xff00) | ((_u << 8) & 0xff0000) | (_u << 24); \
  memcpy(&(v), &_u, 4); } while (0)
  if (fread(dims, 4, 4, fp) != 4 || fread(stamp, 1, 64, fp) != 64 ||
      fread(&order, 4, 1, fp) != 1 || fread(&s29, 4, 1, fp) != 1 ||
      fread(&s31, 4, 1, fp) != 1) { fclose(fp); return -4; }
  if (swap) {
    for (int k = 0; k < 4; k++) BSWAP(dims[k]);
    BSWAP(order); BSWAP(s29); BSWAP(s31);
  }
  if (dims[0] != nx || dims[1] != ny || dims[2] != nz || dims[3] != nt || order != 0) {
    fclose(fp); return -5;
  }
  size_t nw = (size_t)NUMLINK * DIMF * 2;
  float *buf = malloc(sizeof(float) * nw);
  uint32_t c29 = 0, c31 = 0;
  int r29 = 0, r31 = 0;
  for (int lex = 0; lex < V; lex++) {
    if (fread(buf, sizeof(float), nw, fp) != nw) { free(buf); fclose(fp); return -6; }
    uint32_t *w = (uint32_t *)buf;
    for (size_t k = 0; k < nw; k++) {
      if (swap) BSWAP(w[k]);
      uint32_t v = w[k];
      c29 ^= r29 ? (v << r29 | v >> (32 - r29)) : v;
      c31 ^= r31 ? (v << r31 | v >> (32 - r31)) : v;
      if (++r29 >= 29) r29 = 0;
      if (++r31 >= 31) r31 = 0;
    }
    int x = lex % nx, y = (lex / nx) % ny, z = (lex / (nx * ny)) % nz,
        t = lex / (nx * ny * nz);
    int i = site_index(x, y, z, t);
    double *dst = (double *)&links[i * NUMLINK];
    for (size_t k = 0; k < nw; k++) dst[k] = (double)buf[k];
  }
#undef BSWAP
  free(buf);
  fclose(fp);
  if (sums) { sums[0] = c29; sums[1] = c31; sums[2] = s29; sums[3] = s31; }
  compute_fields();
  return (c29 == s29 && c31 == s31) ? 0 : 1;
}

void orc_fermion_op(const double *src, double *dest, int sign) {
  fermion_op_impl((const tf_t *)src, (tf_t *)dest, sign);
}
void orc_DSq(const double *src, double *dest) {
  DSq_impl((const tf_t *)src, (tf_t *)dest);
}
// psim: nshift consecutive TF fields.  Uses the given shifts.
int orc_congrad_multi(const double *src, double *psim, int nshift,
                      const double *shifts, int MaxCG, double errormin,
                      double *size_r) {
  int save = Norder;
  Norder = nshift;
  for (int j = 0; j < nshift; j++) shift_cur[j] = shifts[j];
  tf_t *ps[15];
  for (int j = 0; j < nshift; j++) ps[j] = (tf_t *)psim + (size_t)j * V;
  int it = congrad_multi_impl((const tf_t *)src, ps, MaxCG, errormin, size_r);
  Norder = save;
  return it;
}
double orc_fermion_force(double eps, const double *sol, int nord,
                         const double *amps) {
  tf_t *ps[15];
  for (int j = 0; j < nord; j++) ps[j] = (tf_t *)sol + (size_t)j * V;
  return fermion_force_impl(eps, ps, amps, nord);
}
double orc_gauge_force(double eps) { return gauge_force_impl(eps); }
void orc_update_u(double eps) { update_u_impl(eps); }
void orc_actions(double *out) {
  out[0] = gauge_action_impl();
  out[1] = bmass_action_impl();
  out[2] = det_action_impl();
  out[3] = mom_action_impl();
}
double orc_fermion_action(const double *src, const double *sol) {
  tf_t *ps[15];
  for (int j = 0; j < Norder; j++) ps[j] = (tf_t *)sol + (size_t)j * V;
  return fermion_action_impl((const tf_t *)src, ps);
}
void orc_plaquette(double *ss, double *st) { plaquette_impl(ss, st); }
int orc_update(double *stats) { return update_impl(stats); }
int orc_total_iters(void) { return total_iters; }
void orc_ranmom(void) { ranmom_impl(); }
int orc_grsource(double *src) { return grsource_impl((tf_t *)src); }
void orc_rhmc(double *out) {  // ampdeg4, amp4[15], shift4[15], ampdeg8, amp8, shift8
  out[0] = ampdeg4;
  for (int j = 0; j < 15; j++) { out[1 + j] = amp4[j]; out[16 + j] = shift4[j]; }
  out[31] = ampdeg8;
  for (int j = 0; j < 15; j++) { out[32 + j] = amp8[j]; out[47 + j] = shift8[j]; }
}
// Uniform stream of a site generator (or the node generator for site < 0)
void orc_rand_stream(int site, int n, double *out) {
  prn_t *p = site < 0 ? &node_prn : &site_prn[site];
  for (int k = 0; k < n; k++) out[k] = prn_next(p);
}
