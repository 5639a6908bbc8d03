// Fermion force (reference: 4d_Q16/susy/update_h.c:272-978).
//
// For each pole n, with sol = psi_n and psol = D psi_n, the force matrix on
// link (x, mu) is assembled by the thread that owns (x, mu) from
//   SV    (update_h.c:709-746)   site/link terms
//   VP    (update_h.c:756-894)   link/plaquette terms
//   detF  (update_h.c:477-651)   plaquette-determinant terms (G > 0), both
//                                orientations (update_h.c:899-905)
//   F1Q, F2Q (update_h.c:272-462) Q-closed terms landing on link mu
// and accumulated as F += amp_n f^dag (update_h.c:943-958).  Every term is
// linear in psol, so psol is used unscaled and amp_n scales the sum.
// Finally mom -= eps F (update_h.c:965-969).
#include "dev_mat.h"

__constant__ ForceTerms c_ft;

int force_init_tables(const ForceTerms &ft) {
  HIPCHK(hipMemcpyToSymbol(HIP_SYMBOL(c_ft), &ft, sizeof(ForceTerms)));
  return 0;
}

#define FRB 128

__device__ __forceinline__ double bcs(int t, int dt, int nt, double pbc) {
  int tt = t + dt;
  return (tt < 0 || tt > nt - 1) ? pbc : 1.0;
}

// acc += s * op(A) * op(B) with A, B in memory (element stride V)
template <bool TA, bool TB>
__device__ __forceinline__ void fmm(Mat acc, double s, const double2 *__restrict__ A,
                                    const double2 *__restrict__ B, int V) {
  Mat b;
#pragma unroll
  for (int i = 0; i < NCOL; i++)
#pragma unroll
    for (int j = 0; j < NCOL; j++) {
      double2 v = TB ? B[(j * NCOL + i) * V] : B[(i * NCOL + j) * V];
      if (TB) v.y = -v.y;
      b[i][j] = v;
    }
#pragma unroll
  for (int i = 0; i < NCOL; i++)
#pragma unroll
    for (int j = 0; j < NCOL; j++) {
      double2 a = TA ? A[(j * NCOL + i) * V] : A[(i * NCOL + j) * V];
      double ar = s * a.x, ai = TA ? -s * a.y : s * a.y;
#pragma unroll
      for (int k = 0; k < NCOL; k++) {
        acc[i][k].x = fma(ar, b[j][k].x, acc[i][k].x);
        acc[i][k].x = fma(-ai, b[j][k].y, acc[i][k].x);
        acc[i][k].y = fma(ar, b[j][k].y, acc[i][k].y);
        acc[i][k].y = fma(ai, b[j][k].x, acc[i][k].y);
      }
    }
}

// Per-site scalars of the determinant terms:
// aux[0] Tr sol_site, aux[1] Tr psol_site^dag, aux[2+a] Tr(U_a^-1 sol_a),
// aux[7+a] Tr(U_a^-1 psol_a^dag)
__global__ void __launch_bounds__(FB)
k_force_prep(const double2 *__restrict__ sol, const double2 *__restrict__ psol,
             const double2 *__restrict__ uinv, double2 *__restrict__ aux, int V) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= V) return;
  const size_t M = (size_t)NN * V;
  double2 ts = make_double2(0, 0), tp = make_double2(0, 0);
#pragma unroll
  for (int i = 0; i < NCOL; i++) {
    double2 v = sol[(i * NCOL + i) * V + s], w = psol[(i * NCOL + i) * V + s];
    ts.x += v.x; ts.y += v.y;
    tp.x += w.x; tp.y -= w.y;
  }
  st_keep(aux + s, ts);
  st_keep(aux + V + s, tp);
  for (int a = 0; a < NUMLINK; a++) {
    const double2 *u = uinv + a * M + s;
    const double2 *ps = sol + (1 + a) * M + s, *pp = psol + (1 + a) * M + s;
    double2 t1 = make_double2(0, 0), t2 = make_double2(0, 0);
#pragma unroll
    for (int i = 0; i < NCOL; i++)
#pragma unroll
      for (int j = 0; j < NCOL; j++) {
        double2 x = u[(i * NCOL + j) * V];
        double2 y = ps[(j * NCOL + i) * V];           // sol_a[j][i]
        double2 z = cconjd(pp[(i * NCOL + j) * V]);   // (psol_a^dag)[j][i]
        t1.x += x.x * y.x - x.y * y.y; t1.y += x.x * y.y + x.y * y.x;
        t2.x += x.x * z.x - x.y * z.y; t2.y += x.x * z.y + x.y * z.x;
      }
    st_keep(aux + (2 + a) * V + s, t1);
    st_keep(aux + (7 + a) * V + s, t2);
  }
}

// F[mu](x) (+)= amp * f_U[mu](x)^dag for one pole
__global__ void __launch_bounds__(FRB)
k_force_assemble(const double2 *__restrict__ S, const double2 *__restrict__ P,
                 const double2 *__restrict__ U, const double2 *__restrict__ uinv,
                 const double2 *__restrict__ pd, const double2 *__restrict__ aux,
                 const int *__restrict__ nbr, const int *__restrict__ tco,
                 double2 *__restrict__ F, int V, int nt, double pbc, int doG, double G,
                 double amp, int first) {
  int gid = blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= V * NUMLINK) return;
  const int mu = gid / V, s = gid - mu * V;
  const size_t M = (size_t)NN * V;
  const int t = tco[s];
  const int xm = nbr[g_fwd(mu) * V + s];
  const double bm = bcs(t, link_tau(mu), nt, pbc);
  // X accumulates f_U[mu](x)^dagger: a term A B of f_U enters as B^dag A^dag.
  Mat X;
  mzero(X);
  // ---- SV: f_U += 1/2 u^dag, so X += 1/2 u with
  // u = bc s_psi(x) p_eta(x+mu) - p_eta(x) s_psi(x) - bc p_psi(x) s_eta(x+mu)
  //     + s_eta(x) p_psi(x),   p_eta = psol_site^dag, p_psi = psol_link^dag
  {
    const double2 *spsi = S + (1 + mu) * M, *ppsi = P + (1 + mu) * M;
    fmm<false, true>(X, 0.5 * bm, spsi + s, P + xm, V);
    fmm<true, false>(X, -0.5, P + s, spsi + s, V);
    fmm<true, false>(X, -0.5 * bm, ppsi + s, S + xm, V);
    fmm<false, true>(X, 0.5, S + s, ppsi + s, V);
  }
  // ---- VP: f_U += -bc s_psi_nu(x+mu) p_chi(x) + p_chi(x-nu) s_psi_nu(x-nu)
  //                 + bc p_psi_nu(x+mu) s_chi(x) - s_chi(x-nu) p_psi_nu(x-nu)
  // chi = chi_{mu nu} = sg * stored, p_chi = psol chi^dag, p_psi = psol psi^dag
  for (int nu = 0; nu < NUMLINK; nu++) {
    if (nu == mu) continue;
    const int xn = nbr[g_bwd(nu) * V + s];
    const double sg = mu < nu ? 1.0 : -1.0;
    const double2 *Sc = S + (6 + plaq_idx(mu, nu)) * M;
    const double2 *Pc = P + (6 + plaq_idx(mu, nu)) * M;
    const double2 *Sl = S + (1 + nu) * M, *Pl = P + (1 + nu) * M;
    fmm<false, true>(X, -bm * sg, Pc + s, Sl + xm, V);
    fmm<true, false>(X, sg, Sl + xn, Pc + xn, V);
    fmm<true, false>(X, bm * sg, Sc + s, Pl + xm, V);
    fmm<false, true>(X, -sg, Pl + xn, Sc + xn, V);
  }
  // ---- Q-closed F1Q / F2Q on link c = mu: f_U += -1/2 T^dag, X += -1/2 T.
  // The F2Q term of (a,b,d,e) equals the F1Q term of (d,e,a,b): x-a-b =
  // x+c+d+e, perm(a,b,c,d,e) = perm(d,e,c,a,b) and the two temporal offsets
  // trade places, so the F1Q terms are applied at double weight (checked
  // term by term when the tables are built).
  for (int n = 0; n < 6; n++) {
    const int ab = c_ft.ab[mu][n], de = c_ft.de[mu][n];
    const int xabc = nbr[c_ft.g_abc[mu][n] * V + s];
    const int xmab = nbr[c_ft.g_mab[mu][n] * V + s];
    const double b3 = bcs(t, c_ft.dt3[mu][n], nt, pbc) * bm;
    const double b2 = bcs(t, c_ft.dt2m[mu][n], nt, pbc) * bm;
    const double2 *Sab = S + (6 + ab) * M;
    const double2 *Pde = P + (6 + de) * M;
    const double p1 = c_ft.perm1[mu][n];
    // F1Q: T = p1 b3 Pde^dag(x+a+b+c) Sab(x+c) - p1 b2 Sab(x-a-b) Pde^dag(x+c)
    fmm<true, false>(X, -p1 * b3, Pde + xabc, Sab + xm, V);
    fmm<false, true>(X, p1 * b2, Sab + xmab, Pde + xm, V);
  }
  if (doG) {
    // determinant terms: X^dag += z U^-1 - U^-1 W U^-1, i.e.
    // X += conj(z) U^-dag - (U^-1 W U^-1)^dag
    const double lg1 = -0.5 * C2 * G, lg2 = 0.5 * C2 * G;
    double2 z = make_double2(0, 0);
    double2 w1 = make_double2(0, 0), w2 = make_double2(0, 0);
    for (int call = 0; call < 2; call++) {
      // call 0: eta = psol_site^dag, psi = sol links; call 1: eta = sol_site,
      // psi = psol links^dag
      const double2 *te = aux + (call == 0 ? 1 : 0) * V;
      const double2 *T = aux + (call == 0 ? 2 : 7) * V;
      const double lg = call == 0 ? lg1 : lg2;
      double2 plaq = make_double2(0, 0), adj = make_double2(0, 0), inv = make_double2(0, 0);
      const int a = mu;
      const double2 Ta = T[a * V + s];
      for (int b = 0; b < NUMLINK; b++) {
        if (b == a) continue;
        const int xb = nbr[g_fwd(b) * V + s], xmb = nbr[g_bwd(b) * V + s];
        const int xamb = nbr[g_bwd(b) * V + xm];   // x + e_a - e_b
        const int tmb = tco[xmb];
        const double2 Dba = cmul(te[s], pd[(b * NUMLINK + a) * V + s]);
        const double2 Dab = cmul(te[s], pd[(a * NUMLINK + b) * V + s]);
        const double2 Dab_m = cmul(te[xmb], pd[(a * NUMLINK + b) * V + xmb]);
        const double2 Dba_m = cmul(te[xmb], pd[(b * NUMLINK + a) * V + xmb]);
        double sa = bcs(t, link_tau(a), nt, pbc), smb = bcs(t, -link_tau(b), nt, pbc);
        double sb = bcs(t, link_tau(b), nt, pbc), sam = bcs(tmb, link_tau(a), nt, pbc);
        double2 Tb = T[b * V + s], Tb_xa = T[b * V + xm], Tb_m = T[b * V + xmb];
        double2 Ta_xb = T[a * V + xb], Ta_m = T[a * V + xmb], Tb_amb = T[b * V + xamb];
        double2 v;
        v = cmul(Dba, make_double2(Ta.x + sa * Tb_xa.x, Ta.y + sa * Tb_xa.y));
        plaq.x += v.x; plaq.y += v.y;
        v = cmul(Dab_m, make_double2(Tb_m.x + smb * Ta.x, Tb_m.y + smb * Ta.y));
        plaq.x += v.x; plaq.y += v.y;
        v = cmul(Dab, make_double2(Tb.x + sb * Ta_xb.x, Tb.y + sb * Ta_xb.y));
        adj.x += v.x; adj.y += v.y;
        v = cmul(Dba_m, make_double2(Ta_m.x + sam * Tb_amb.x, Ta_m.y + sam * Tb_amb.y));
        adj.x += v.x; adj.y += v.y;
        inv.x += Dba.x + smb * Dab_m.x;
        inv.y += Dba.y + smb * Dab_m.y;
      }
      // f_U += (lg plaq) U^-1 + conj(lg adj) U^-1 - (lg inv) U^-1 psi U^-1
      z.x += lg * plaq.x + lg * adj.x;
      z.y += lg * plaq.y - lg * adj.y;
      if (call == 0) w1 = make_double2(lg * inv.x, lg * inv.y);
      else w2 = make_double2(lg * inv.x, lg * inv.y);
    }
    // X (holding f_U^dag) += conj(z) (U^-1)^dag - (U^-1 W U^-1)^dag,
    // W = w1 sol_mu + w2 psol_mu^dag
    const double2 *ui = uinv + mu * M + s;
    const double2 *sl = S + (1 + mu) * M + s, *pl = P + (1 + mu) * M + s;
    Mat Wm;
#pragma unroll
    for (int i = 0; i < NCOL; i++)
#pragma unroll
      for (int j = 0; j < NCOL; j++) {
        double2 x = sl[(i * NCOL + j) * V], y = cconjd(pl[(j * NCOL + i) * V]);
        Wm[i][j] = make_double2(w1.x * x.x - w1.y * x.y + w2.x * y.x - w2.y * y.y,
                                w1.x * x.y + w1.y * x.x + w2.x * y.y + w2.y * y.x);
      }
#pragma unroll
    for (int i = 0; i < NCOL; i++) {
      // row i of Q = U^-1 W U^-1:  r_l = sum_k Ui[i][k] W[k][l]; Q[i][j] = sum_l r_l Ui[l][j]
      double2 rr[NCOL];
#pragma unroll
      for (int l = 0; l < NCOL; l++) rr[l] = make_double2(0, 0);
#pragma unroll
      for (int k = 0; k < NCOL; k++) {
        double2 u = ui[(i * NCOL + k) * V];
#pragma unroll
        for (int l = 0; l < NCOL; l++) {
          double2 v = cmul(u, Wm[k][l]);
          rr[l].x += v.x; rr[l].y += v.y;
        }
      }
#pragma unroll
      for (int j = 0; j < NCOL; j++) {
        double2 q = make_double2(0, 0);
#pragma unroll
        for (int l = 0; l < NCOL; l++) {
          double2 v = cmul(rr[l], ui[(l * NCOL + j) * V]);
          q.x += v.x; q.y += v.y;
        }
        double2 u = ui[(i * NCOL + j) * V];
        double2 zu = cmul(z, u);
        // f_U[i][j] += zu - q  ->  X[j][i] += conj(zu - q)
        X[j][i].x += zu.x - q.x;
        X[j][i].y -= zu.y - q.y;
      }
    }
  }
  // F (+)= amp X   (X = f_U^dag)
  double2 *fo = F + mu * M + s;
#pragma unroll
  for (int i = 0; i < NCOL; i++)
#pragma unroll
    for (int j = 0; j < NCOL; j++) {
      double2 v = make_double2(amp * X[i][j].x, amp * X[i][j].y);
      if (!first) {
        double2 o = fo[(i * NCOL + j) * V];
        v.x += o.x;
        v.y += o.y;
      }
      st_keep(fo + (i * NCOL + j) * V, v);
    }
}

// mom -= eps F; partial |F|^2
__global__ void __launch_bounds__(FB)
k_force_apply(const double2 *__restrict__ F, double2 *__restrict__ mom, double eps,
              int V, double *__restrict__ partial) {
  int gid = blockIdx.x * blockDim.x + threadIdx.x;
  double nrm = 0;
  if (gid < V * NUMLINK) {
    int mu = gid / V, s = gid - mu * V;
    const size_t M = (size_t)NN * V;
#pragma unroll
    for (int e = 0; e < NN; e++) {
      size_t k = mu * M + (size_t)e * V + s;
      double2 f = F[k], m = mom[k];
      m.x -= eps * f.x;
      m.y -= eps * f.y;
      st_keep(mom + k, m);
      nrm += f.x * f.x + f.y * f.y;
    }
  }
  block_partial(nrm, partial, 0, 0);
}

int fermion_force_dev(susy_ctx *c, double eps, double2 *const *sol, int nsol,
                      const double *amps, double *norm) {
  const int V = c->V;
  double2 *psol = c->d_tmp;
  for (int n = 0; n < nsol; n++) {
    int r = dslash64(c, sol[n], psol, 1, nullptr, 0.0, nullptr, nullptr);
    if (r) return r;
    if (c->doG)
      hipLaunchKernelGGL(k_force_prep, dim3((V + FB - 1) / FB), dim3(FB), 0, c->s, sol[n],
                         psol, c->d_uinv, c->d_aux, V);
    hipLaunchKernelGGL(k_force_assemble, dim3((V * NUMLINK + FRB - 1) / FRB), dim3(FRB), 0,
                       c->s, sol[n], psol, c->d_links, c->d_uinv, c->d_pdet, c->d_aux,
                       c->d_nbr, c->d_t, c->d_force, V, c->p.nt, (double)c->p.PBC, c->doG,
                       c->p.G, amps[n], n == 0);
    HIPCHK(hipGetLastError());
  }
  int nb = (V * NUMLINK + FB - 1) / FB;
  hipLaunchKernelGGL(k_force_apply, dim3(nb), dim3(FB), 0, c->s, c->d_force, c->d_mom, eps, V,
                     c->d_part);
  HIPCHK(hipGetLastError());
  int r = reduce_partials(c, c->d_part, nb, 0);
  if (r) return r;
  double sc;
  if ((r = read_scalars(c, &sc, 1))) return r;
  if (norm) *norm = eps * sqrt(sc) / V;
  return 0;
}
