// Host-side lattice geometry, gather tables, term tables, U(N) generators
// and the reference-compatible random number generator.
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "susy_internal.h"

static char g_err[512] = "";
void set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}
extern "C" const char *susy_last_error(void) { return g_err; }

// Link offsets: four unit vectors and (-1,-1,-1,-1) (setup_offset.c:155-162)
static const int EV[NUMLINK][4] = {
    {1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0}, {0, 0, 0, 1}, {-1, -1, -1, -1}};

// One-rank node_index: even sites first, lexicographic within each parity
// (layout_hyper_prime.c:157-169).
int site_index(const Geometry &g, int x, int y, int z, int t) {
  int lex = x + g.nx * (y + g.ny * (z + g.nz * t));
  return ((x + y + z + t) & 1) ? (lex + g.V) / 2 : lex / 2;
}

static int shifted(const Geometry &g, int i, const int d[4]) {
  return site_index(g, (g.x[i] + g.nx + d[0]) % g.nx, (g.y[i] + g.ny + d[1]) % g.ny,
                    (g.z[i] + g.nz + d[2]) % g.nz, (g.t[i] + g.nt + d[3]) % g.nt);
}

// Q-closed offset list: every vector of {-1,0,1}^4 with two or three nonzero
// entries, keeping the first of each +/- pair in the order of descending
// nested loops (setup_offset.c:225-250).
static void build_q_offsets(int q[NQLINK][4]) {
  int cnt = 0, v[4];
  for (v[0] = 1; v[0] >= -1; v[0]--)
    for (v[1] = 1; v[1] >= -1; v[1]--)
      for (v[2] = 1; v[2] >= -1; v[2]--)
        for (v[3] = 1; v[3] >= -1; v[3]--) {
          int k = abs(v[0]) + abs(v[1]) + abs(v[2]) + abs(v[3]);
          if (k != 2 && k != 3) continue;
          bool seen = false;
          for (int p = 0; p < cnt && !seen; p++)
            seen = v[0] == -q[p][0] && v[1] == -q[p][1] && v[2] == -q[p][2] &&
                   v[3] == -q[p][3];
          if (!seen) { memcpy(q[cnt], v, sizeof(v)); cnt++; }
        }
}

void build_geometry(Geometry &g, int nx, int ny, int nz, int nt) {
  g.nx = nx; g.ny = ny; g.nz = nz; g.nt = nt;
  g.V = nx * ny * nz * nt;
  int V = g.V;
  g.x.assign(V, 0); g.y.assign(V, 0); g.z.assign(V, 0); g.t.assign(V, 0);
  for (int t = 0; t < nt; t++)
    for (int z = 0; z < nz; z++)
      for (int y = 0; y < ny; y++)
        for (int x = 0; x < nx; x++) {
          int i = site_index(g, x, y, z, t);
          g.x[i] = x; g.y[i] = y; g.z[i] = z; g.t[i] = t;
        }
  build_q_offsets(g.q_offset);
  g.gathers.assign((size_t)NGATHER * V, 0);
  // ids 0..3: +x,+y,+z,+t; 7..4: -x,-y,-z,-t (com_vanilla.c:322-399)
  for (int k = 0; k < 4; k++) {
    int e[4] = {0, 0, 0, 0}, m[4] = {0, 0, 0, 0};
    e[k] = 1; m[k] = -1;
    for (int i = 0; i < V; i++) {
      g.gathers[(size_t)k * V + i] = shifted(g, i, e);
      g.gathers[(size_t)(7 - k) * V + i] = shifted(g, i, m);
    }
  }
  // 8+2a / 9+2a: x + e_a / x - e_a
  for (int a = 0; a < NUMLINK; a++) {
    int m[4] = {-EV[a][0], -EV[a][1], -EV[a][2], -EV[a][3]};
    for (int i = 0; i < V; i++) {
      g.gathers[(size_t)g_fwd(a) * V + i] = shifted(g, i, EV[a]);
      g.gathers[(size_t)g_bwd(a) * V + i] = shifted(g, i, m);
    }
  }
  // 18+2p / 19+2p: x + q_p / x - q_p
  for (int p = 0; p < NQLINK; p++) {
    int m[4] = {-g.q_offset[p][0], -g.q_offset[p][1], -g.q_offset[p][2],
                -g.q_offset[p][3]};
    for (int i = 0; i < V; i++) {
      g.gathers[(size_t)(18 + 2 * p) * V + i] = shifted(g, i, g.q_offset[p]);
      g.gathers[(size_t)(19 + 2 * p) * V + i] = shifted(g, i, m);
    }
  }
}

// Gather id that delivers x + d for a Q-closed offset d (check_list,
// setup_offset.c:38-61): the listed vector itself or its negative.
int gather_id(const int q[NQLINK][4], const int d[4]) {
  for (int p = 0; p < NQLINK; p++)
    if (d[0] == q[p][0] && d[1] == q[p][1] && d[2] == q[p][2] && d[3] == q[p][3])
      return 18 + 2 * p;
  for (int p = 0; p < NQLINK; p++)
    if (d[0] == -q[p][0] && d[1] == -q[p][1] && d[2] == -q[p][2] &&
        d[3] == -q[p][3])
      return 19 + 2 * p;
  return -1;
}

// Sign of the permutation (i,j,k,l,m) of 0..4 (setup_lambda.c:123-193)
static int perm_sign(int i, int j, int k, int l, int m) {
  int s[5] = {i, j, k, l, m}, inv = 0;
  for (int p = 0; p < 5; p++)
    for (int q = p + 1; q < 5; q++) inv += s[p] > s[q];
  return (inv & 1) ? -1 : 1;
}

void build_plaq_terms(const int q[NQLINK][4], PlaqTerms &pt) {
  memset(&pt, 0, sizeof(pt));
  int npp[NPLAQ] = {0}, npm[NPLAQ] = {0};
  for (int a = 0; a < NUMLINK; a++)
    for (int b = a + 1; b < NUMLINK; b++) {
      int p = plaq_idx(a, b);
      pt.a[p] = a; pt.b[p] = b;
    }
  // every permutation (a,b,c,d,e) of the five links with a<b and d<e gives
  // one DbplusPtoP term on plaquette ab and one DbminusPtoP term on de
  for (int a = 0; a < NUMLINK; a++)
    for (int b = a + 1; b < NUMLINK; b++)
      for (int c = 0; c < NUMLINK; c++) {
        if (c == a || c == b) continue;
        for (int d = 0; d < NUMLINK; d++) {
          if (d == a || d == b || d == c) continue;
          for (int e = d + 1; e < NUMLINK; e++) {
            if (e == a || e == b || e == c) continue;
            float tr = 0.5f * perm_sign(a, b, c, d, e);
            int vab[4], vabc[4];
            for (int k = 0; k < 4; k++) {
              vab[k] = EV[a][k] + EV[b][k];
              vabc[k] = vab[k] + EV[c][k];
            }
            int p = plaq_idx(a, b), n = npp[p]++;
            pt.pp_c[p][n] = c;
            pt.pp_de[p][n] = plaq_idx(d, e);
            pt.pp_g1[p][n] = gather_id(q, vab);
            pt.pp_g2[p][n] = gather_id(q, vabc);
            pt.pp_dt3[p][n] = link_tau(a) + link_tau(b) + link_tau(c);
            pt.pp_dt2[p][n] = link_tau(a) + link_tau(b);
            pt.pp_tr[p][n] = tr;
            int mab[4], mabc[4];
            for (int k = 0; k < 4; k++) { mab[k] = -vab[k]; mabc[k] = -vabc[k]; }
            p = plaq_idx(d, e); n = npm[p]++;
            pt.pm_c[p][n] = c;
            pt.pm_ab[p][n] = plaq_idx(a, b);
            pt.pm_g2[p][n] = gather_id(q, mab);
            pt.pm_g1[p][n] = gather_id(q, mabc);
            pt.pm_dt2[p][n] = -link_tau(a) - link_tau(b);
            pt.pm_dt3[p][n] = -link_tau(a) - link_tau(b) - link_tau(c);
            pt.pm_tr[p][n] = tr;
          }
        }
      }
}

void build_force_terms(const int q[NQLINK][4], ForceTerms &ft) {
  memset(&ft, 0, sizeof(ft));
  for (int c = 0; c < NUMLINK; c++) {
    int n = 0;
    for (int d = 0; d < NUMLINK; d++) {
      if (d == c) continue;
      for (int e = d + 1; e < NUMLINK; e++) {
        if (e == c) continue;
        for (int a = 0; a < NUMLINK; a++) {
          if (a == c || a == d || a == e) continue;
          for (int b = a + 1; b < NUMLINK; b++) {
            if (b == c || b == d || b == e) continue;
            int vab[4], vabc[4], mab[4];
            for (int k = 0; k < 4; k++) {
              vab[k] = EV[a][k] + EV[b][k];
              vabc[k] = vab[k] + EV[c][k];
              mab[k] = -vab[k];
            }
            ft.a[c][n] = a; ft.b[c][n] = b;
            ft.ab[c][n] = plaq_idx(a, b); ft.de[c][n] = plaq_idx(d, e);
            ft.g_abc[c][n] = gather_id(q, vabc);
            ft.g_mab[c][n] = gather_id(q, mab);
            ft.dt3[c][n] = link_tau(a) + link_tau(b) + link_tau(c);
            ft.dt2m[c][n] = -link_tau(a) - link_tau(b);
            ft.perm1[c][n] = (float)perm_sign(d, e, c, a, b);
            ft.perm2[c][n] = (float)perm_sign(a, b, c, d, e);
            n++;
          }
        }
      }
    }
  }
}

// U(N) generators normalised to Tr(T_a T_b) = -delta_ab (setup_lambda.c:10-64):
// N(N-1) off-diagonal, N-1 diagonal, and i I/sqrt(N) last.
void build_lambda(std::vector<double> &lam) {
  lam.assign((size_t)NN * NN * 2, 0.0);
  auto at = [&](int g, int i, int j) -> double * {
    return &lam[((size_t)g * NN + i * NCOL + j) * 2];
  };
  double r2 = 1.0 / sqrt(2.0);
  int count = 0;
  for (int i = 0; i < NCOL; i++)
    for (int j = i + 1; j < NCOL; j++) {
      at(count, i, j)[1] = r2; at(count, j, i)[1] = r2;
      at(count + 1, i, j)[0] = r2; at(count + 1, j, i)[0] = -r2;
      count += 2;
    }
  for (int i = 0; i < NCOL - 1; i++) {
    int g = NCOL * (NCOL - 1) + i, k = i + 1;
    double f = 1.0 / sqrt(k * (k + 1.0));
    for (int l = 0; l <= k; l++) at(g, l, l)[1] = f;
    at(g, k, k)[1] = -k * f;
  }
  for (int i = 0; i < NCOL; i++) at(NN - 1, i, i)[1] = sqrt(1.0 / (double)NCOL);
}

// Generator seeded per site with the reference's integer recurrences
// (ranstuff.c:22-43): 32-bit signed wrap-around, LCG state kept in a 64-bit
// word after sign extension.
void prn_init(Prn *p, int seed, int index) {
  uint32_t m = (uint32_t)(69607 + 8 * index), s = (uint32_t)seed;
  for (int k = 0; k < 7; k++) {
    s = m * s + 12345u;
    p->r[k] = (uint64_t)(((int32_t)s >> 8) & 0xffffff);
  }
  s = m * s + 12345u;
  p->ic_state = (uint64_t)(int64_t)(int32_t)s;
  p->multiplier = (uint64_t)(100005 + 8 * index);
  p->addend = 12345;
  p->scale = (float)(1.0 / (double)0x1000000);
}
// XOR of a lagged shift register and the LCG (ranstuff.c:48-63)
double prn_next(Prn *p) {
  int32_t t = (int32_t)((((p->r[5] >> 7) | (p->r[6] << 17)) ^
                         ((p->r[4] >> 1) | (p->r[5] << 23))) & 0xffffff);
  for (int k = 6; k > 0; k--) p->r[k] = p->r[k - 1];
  p->r[0] = (uint64_t)t;
  int32_t s = (int32_t)(uint32_t)(p->ic_state * p->multiplier + p->addend);
  p->ic_state = (uint64_t)(int64_t)s;
  return (double)(p->scale * (float)(t ^ ((s >> 8) & 0xffffff)));
}
