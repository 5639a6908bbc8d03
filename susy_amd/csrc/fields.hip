// Link-derived fields, link update, gauge force, actions and plaquette.
//   compute_plaqdet / compute_Uinv  (4d_Q16/susy/utilities.c:17-127,
//                                    determinant.c:10-85)
//   compute_DmuUmu / compute_Fmunu  (action.c:15-98)
//   update_u                        (update_o.c:20-35)
//   gauge_force, det_force          (update_h.c:15-261, det_force.c:6-76)
//   gauge/bmass/det/mom actions     (action.c:106-286)
//   plaquette                       (plaq.c:6-51)
// All owner-computes: each thread writes only its own site's (or link's)
// outputs.  Sums use per-block partials reduced in a fixed order.
#include "dev_mat.h"

// ---------------------------------------------------------------------
// det and inverse of every link: LU with partial pivoting on |re|+|im|
// (the zgetrf convention), det = prod diag with a sign flip per row swap.
__global__ void __launch_bounds__(FB)
k_det_inv(const double2 *__restrict__ U, double2 *__restrict__ det,
          double2 *__restrict__ uinv, int V) {
  int gid = blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= V * NUMLINK) return;
  int a = gid / V, s = gid - a * V;
  const size_t M = (size_t)NN * V;
  Mat m;
  mload(m, U + a * M + s, V);
  int piv[NCOL];
#pragma unroll
  for (int k = 0; k < NCOL; k++) {
    int p = k;
    double best = cabs1(m[k][k]);
#pragma unroll
    for (int i = k + 1; i < NCOL; i++) {
      double v = cabs1(m[i][k]);
      if (v > best) { best = v; p = i; }
    }
    piv[k] = p;
    if (p != k) {
#pragma unroll
      for (int j = 0; j < NCOL; j++) {
        // swap rows k and p with register-friendly selects
        double2 tk = m[k][j];
#pragma unroll
        for (int i = k + 1; i < NCOL; i++)
          if (i == p) { m[k][j] = m[i][j]; m[i][j] = tk; }
      }
    }
    double2 inv = cdiv(make_double2(1, 0), m[k][k]);
#pragma unroll
    for (int i = k + 1; i < NCOL; i++) {
      m[i][k] = cmul(m[i][k], inv);
#pragma unroll
      for (int j = k + 1; j < NCOL; j++) {
        double2 t = cmul(m[i][k], m[k][j]);
        m[i][j].x -= t.x;
        m[i][j].y -= t.y;
      }
    }
  }
  double2 d = make_double2(1, 0);
#pragma unroll
  for (int i = 0; i < NCOL; i++) {
    d = cmul(d, m[i][i]);
    if (piv[i] != i) d = make_double2(-d.x, -d.y);
  }
  st_keep(det + a * V + s, d);
  // inverse: solve L U X = P^T I column by column
  Mat x;
#pragma unroll
  for (int col = 0; col < NCOL; col++) {
    double2 bv[NCOL];
#pragma unroll
    for (int i = 0; i < NCOL; i++) bv[i] = make_double2(i == col ? 1.0 : 0.0, 0);
#pragma unroll
    for (int k = 0; k < NCOL; k++) {
      double2 tk = bv[k];
#pragma unroll
      for (int i = k + 1; i < NCOL; i++)
        if (i == piv[k]) { bv[k] = bv[i]; bv[i] = tk; }
    }
#pragma unroll
    for (int i = 1; i < NCOL; i++)
#pragma unroll
      for (int j = 0; j < i; j++) {
        double2 t = cmul(m[i][j], bv[j]);
        bv[i].x -= t.x;
        bv[i].y -= t.y;
      }
#pragma unroll
    for (int i = NCOL - 1; i >= 0; i--) {
#pragma unroll
      for (int j = i + 1; j < NCOL; j++) {
        double2 t = cmul(m[i][j], bv[j]);
        bv[i].x -= t.x;
        bv[i].y -= t.y;
      }
      bv[i] = cdiv(bv[i], m[i][i]);
    }
#pragma unroll
    for (int i = 0; i < NCOL; i++) x[i][col] = bv[i];
  }
  mstore(uinv + a * M + s, x, V);
}

// plaqdet_ab(x) = det U_a(x)^* det U_b(x) det U_b(x+a)^* det U_a(x+b)
// and ZW*_ab = plaqdet_ab (plaqdet_ab - 1)^*, for all a != b
__global__ void __launch_bounds__(FB)
k_plaqdet(const double2 *__restrict__ det, const int *__restrict__ nbr,
          double2 *__restrict__ pd, double2 *__restrict__ zw, int V) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= V) return;
  for (int a = 0; a < NUMLINK; a++) {
    int xa = nbr[g_fwd(a) * V + s];
    for (int b = 0; b < NUMLINK; b++) {
      if (a == b) {
        st_keep(pd + (a * NUMLINK + b) * V + s, make_double2(0, 0));
        st_keep(zw + (a * NUMLINK + b) * V + s, make_double2(0, 0));
        continue;
      }
      int xb = nbr[g_fwd(b) * V + s];
      double2 t = cmul(cconjd(det[a * V + s]), det[b * V + s]);
      t = cmul(t, cconjd(det[b * V + xa]));
      t = cmul(det[a * V + xb], t);
      st_keep(pd + (a * NUMLINK + b) * V + s, t);
      st_keep(zw + (a * NUMLINK + b) * V + s, cmul(t, cconjd(make_double2(t.x - 1.0, t.y))));
    }
  }
}

// DmuUmu(x) = sum_a [U_a Ubar_a(x) - Ubar_a U_a(x-a)] + G sum_{a!=b}(plaqdet_ab - 1) I
__global__ void __launch_bounds__(FB)
k_dmuumu(const double2 *__restrict__ U, const double2 *__restrict__ pd,
         const int *__restrict__ nbr, double2 *__restrict__ dmu, int V, int doG,
         double G) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= V) return;
  const size_t M = (size_t)NN * V;
  Mat d, u;
  mzero(d);
  for (int a = 0; a < NUMLINK; a++) {
    mload(u, U + a * M + s, V);
    mmadd<false, true>(d, 1.0, u, u);
    int xm = nbr[g_bwd(a) * V + s];
    mload(u, U + a * M + xm, V);
    mmadd<true, false>(d, -1.0, u, u);
  }
  if (doG) {
    double2 g = make_double2(0, 0);
    for (int a = 0; a < NUMLINK; a++)
      for (int b = 0; b < NUMLINK; b++) {
        if (a == b) continue;
        double2 p = pd[(a * NUMLINK + b) * V + s];
        g.x += G * (p.x - 1.0);
        g.y += G * p.y;
      }
#pragma unroll
    for (int i = 0; i < NCOL; i++) { d[i][i].x += g.x; d[i][i].y += g.y; }
  }
  mstore(dmu + s, d, V);
}

// F_ab(x) = U_a(x) U_b(x+a) - U_b(x) U_a(x+b), a < b
__global__ void __launch_bounds__(FB)
k_fmunu(const double2 *__restrict__ U, const int *__restrict__ nbr,
        double2 *__restrict__ F, int V) {
  int gid = blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= V * NPLAQ) return;
  int p = gid / V, s = gid - p * V;
  int a = 0, b = 1;
  for (int i = 0; i < NUMLINK; i++)
    for (int j = i + 1; j < NUMLINK; j++)
      if (plaq_idx(i, j) == p) { a = i; b = j; }
  const size_t M = (size_t)NN * V;
  int xa = nbr[g_fwd(a) * V + s], xb = nbr[g_fwd(b) * V + s];
  Mat f, u1, u2;
  mzero(f);
  mload(u1, U + a * M + s, V);
  mload(u2, U + b * M + xa, V);
  mmadd<false, false>(f, 1.0, u1, u2);
  mload(u1, U + b * M + s, V);
  mload(u2, U + a * M + xb, V);
  mmadd<false, false>(f, -1.0, u1, u2);
  mstore(F + p * M + s, f, V);
}

template <typename T>
__global__ void k_demote(const double2 *__restrict__ in, T *__restrict__ out, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) st_keep(out + i, make_float2((float)in[i].x, (float)in[i].y));
}

int fields_recompute(susy_ctx *c) {
  const int V = c->V;
  hipLaunchKernelGGL(k_det_inv, dim3((V * NUMLINK + FB - 1) / FB), dim3(FB), 0, c->s,
                     c->d_links, c->d_det, c->d_uinv, V);
  hipLaunchKernelGGL(k_plaqdet, dim3((V + FB - 1) / FB), dim3(FB), 0, c->s, c->d_det,
                     c->d_nbr, c->d_pdet, c->d_zw, V);
  hipLaunchKernelGGL(k_dmuumu, dim3((V + FB - 1) / FB), dim3(FB), 0, c->s, c->d_links,
                     c->d_pdet, c->d_nbr, c->d_dmu, V, c->doG, c->p.G);
  hipLaunchKernelGGL(k_fmunu, dim3((V * NPLAQ + FB - 1) / FB), dim3(FB), 0, c->s,
                     c->d_links, c->d_nbr, c->d_fmunu, V);
  HIPCHK(hipGetLastError());
  c->f_fresh = 0;
  return 0;
}

int fields_refresh32(susy_ctx *c) {
  const size_t V = c->V;
  size_t n1 = (size_t)NUMLINK * NN * V, n2 = (size_t)NUMLINK * NUMLINK * V;
  hipLaunchKernelGGL(k_demote<float2>, dim3((n1 + FB - 1) / FB), dim3(FB), 0, c->s,
                     c->d_links, c->f_links, n1);
  hipLaunchKernelGGL(k_demote<float2>, dim3((n1 + FB - 1) / FB), dim3(FB), 0, c->s,
                     c->d_uinv, c->f_uinv, n1);
  hipLaunchKernelGGL(k_demote<float2>, dim3((n2 + FB - 1) / FB), dim3(FB), 0, c->s,
                     c->d_pdet, c->f_pdet, n2);
  HIPCHK(hipGetLastError());
  c->f_fresh = 1;
  return 0;
}

// U += eps * mom
__global__ void k_update_u(double2 *__restrict__ U, const double2 *__restrict__ mom,
                           double eps, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    double2 u = U[i], m = mom[i];
    u.x += eps * m.x;
    u.y += eps * m.y;
    st_keep(U + i, u);
  }
}

int update_u_dev(susy_ctx *c, double eps) {
  size_t n = (size_t)NUMLINK * NN * c->V;
  hipLaunchKernelGGL(k_update_u, dim3((n + FB - 1) / FB), dim3(FB), 0, c->s, c->d_links,
                     c->d_mom, eps, n);
  HIPCHK(hipGetLastError());
  return fields_recompute(c);
}

// ---------------------------------------------------------------------
// Gauge force on link (x, mu); mom -= kappa eps f^dag; partial |f|^2
__global__ void __launch_bounds__(FB)
k_gauge_force(const double2 *__restrict__ U, const double2 *__restrict__ dmu,
              const double2 *__restrict__ F, const double2 *__restrict__ uinv,
              const double2 *__restrict__ pd, const int *__restrict__ nbr,
              double2 *__restrict__ mom, int V, int doG, double G, double bmass,
              double kappa_eps, double *__restrict__ partial) {
  int gid = blockIdx.x * blockDim.x + threadIdx.x;
  double nrm = 0;
  if (gid < V * NUMLINK) {
    int mu = gid / V, s = gid - mu * V;
    const size_t M = (size_t)NN * V;
    int xm = nbr[g_fwd(mu) * V + s];
    Mat f, u, d, t;
    mzero(f);
    mload(u, U + mu * M + s, V);
    // Ubar_mu(x) DmuUmu(x) - DmuUmu(x+mu) Ubar_mu(x)
    mload(d, dmu + s, V);
    mmadd<true, false>(f, 1.0, u, d);
    mload(d, dmu + xm, V);
    mmadd<false, true>(f, -1.0, d, u);
    if (doG) {
      // G U_mu^-1 sum_nu [Tr D(x) plaqdet_numu(x) + Tr D(x-nu) plaqdet_munu(x-nu)]
      double2 trx = make_double2(0, 0);
#pragma unroll
      for (int i = 0; i < NCOL; i++) {
        double2 v = dmu[(i * NCOL + i) * V + s];
        trx.x += v.x; trx.y += v.y;
      }
      double2 sum = make_double2(0, 0);
      for (int nu = 0; nu < NUMLINK; nu++) {
        if (nu == mu) continue;
        int xn = nbr[g_bwd(nu) * V + s];
        double2 trm = make_double2(0, 0);
#pragma unroll
        for (int i = 0; i < NCOL; i++) {
          double2 v = dmu[(i * NCOL + i) * V + xn];
          trm.x += v.x; trm.y += v.y;
        }
        double2 p1 = cmul(trx, pd[(nu * NUMLINK + mu) * V + s]);
        double2 p2 = cmul(trm, pd[(mu * NUMLINK + nu) * V + xn]);
        sum.x += p1.x + p2.x;
        sum.y += p1.y + p2.y;
      }
      mload(t, uinv + mu * M + s, V);
      double sr = G * sum.x, si = G * sum.y;
#pragma unroll
      for (int i = 0; i < NCOL; i++)
#pragma unroll
        for (int k = 0; k < NCOL; k++) {
          f[i][k].x += sr * t[i][k].x - si * t[i][k].y;
          f[i][k].y += sr * t[i][k].y + si * t[i][k].x;
        }
    }
    // 2 sum_nu [U_nu(x+mu) F_munu(x)^dag - F_munu(x-nu)^dag U_nu(x-nu)]
    for (int nu = 0; nu < NUMLINK; nu++) {
      if (nu == mu) continue;
      double sg = mu < nu ? 2.0 : -2.0;
      int p = plaq_idx(mu, nu);
      int xn = nbr[g_bwd(nu) * V + s];
      mload(t, U + nu * M + xm, V);
      mload(d, F + p * M + s, V);
      mmadd<false, true>(f, sg, t, d);
      mload(t, U + nu * M + xn, V);
      mload(d, F + p * M + xn, V);
      mmadd<true, false>(f, -sg, d, t);
    }
    if (bmass > IMAG_TOL) {
      // double-trace scalar potential: 2 bmass^2/N (Tr(U Ubar)/N - 1) Ubar
      double tr = retrace_dag(u, u) / NCOL - 1.0;
      double dm = 2.0 * bmass * bmass / NCOL * tr;
#pragma unroll
      for (int i = 0; i < NCOL; i++)
#pragma unroll
        for (int k = 0; k < NCOL; k++) {
          f[i][k].x += dm * u[k][i].x;
          f[i][k].y -= dm * u[k][i].y;
        }
    }
    double2 *mp = mom + mu * M + s;
#pragma unroll
    for (int i = 0; i < NCOL; i++)
#pragma unroll
      for (int k = 0; k < NCOL; k++) {
        double2 m = mp[(i * NCOL + k) * V];
        m.x -= kappa_eps * f[k][i].x;
        m.y += kappa_eps * f[k][i].y;
        st_keep(mp + (i * NCOL + k) * V, m);
      }
    nrm = retrace_dag(f, f);
  }
  block_partial(nrm, partial, 0, 0);
}

// det_force: f = kappa_u1 U_a^-1 sum_{b!=a}[ZW*_ba(x) + ZW*_ab(x-b)]; mom -= eps f^dag
__global__ void __launch_bounds__(FB)
k_det_force(const double2 *__restrict__ uinv, const double2 *__restrict__ zw,
            const int *__restrict__ nbr, double2 *__restrict__ mom, int V,
            double kappa_u1, double eps, double *__restrict__ partial) {
  int gid = blockIdx.x * blockDim.x + threadIdx.x;
  double nrm = 0;
  if (gid < V * NUMLINK) {
    int a = gid / V, s = gid - a * V;
    const size_t M = (size_t)NN * V;
    double2 sum = make_double2(0, 0);
    for (int b = 0; b < NUMLINK; b++) {
      if (b == a) continue;
      int xm = nbr[g_bwd(b) * V + s];
      double2 z1 = zw[(b * NUMLINK + a) * V + s], z2 = zw[(a * NUMLINK + b) * V + xm];
      sum.x += z1.x + z2.x;
      sum.y += z1.y + z2.y;
    }
    sum.x *= kappa_u1;
    sum.y *= kappa_u1;
    Mat u, f;
    mload(u, uinv + a * M + s, V);
#pragma unroll
    for (int i = 0; i < NCOL; i++)
#pragma unroll
      for (int k = 0; k < NCOL; k++) f[i][k] = cmul(u[i][k], sum);
    double2 *mp = mom + a * M + s;
#pragma unroll
    for (int i = 0; i < NCOL; i++)
#pragma unroll
      for (int k = 0; k < NCOL; k++) {
        double2 m = mp[(i * NCOL + k) * V];
        m.x -= eps * f[k][i].x;
        m.y += eps * f[k][i].y;
        st_keep(mp + (i * NCOL + k) * V, m);
      }
    nrm = retrace_dag(f, f);
  }
  block_partial(nrm, partial, 0, 0);
}

int gauge_force_dev(susy_ctx *c, double eps, double *norm) {
  const int V = c->V;
  int nb = (V * NUMLINK + FB - 1) / FB;
  hipLaunchKernelGGL(k_gauge_force, dim3(nb), dim3(FB), 0, c->s, c->d_links, c->d_dmu,
                     c->d_fmunu, c->d_uinv, c->d_pdet, c->d_nbr, c->d_mom, V, c->doG,
                     c->p.G, c->p.bmass, c->kappa * eps, c->d_part);
  HIPCHK(hipGetLastError());
  int r = reduce_partials(c, c->d_part, nb, 0);
  if (r) return r;
  bool det = c->p.kappa_u1 > IMAG_TOL;
  if (det) {
    hipLaunchKernelGGL(k_det_force, dim3(nb), dim3(FB), 0, c->s, c->d_uinv, c->d_zw,
                       c->d_nbr, c->d_mom, V, c->p.kappa_u1, eps, c->d_part);
    HIPCHK(hipGetLastError());
    r = reduce_partials(c, c->d_part, nb, 1);
    if (r) return r;
  }
  double sc[2] = {0, 0};
  r = read_scalars(c, sc, det ? 2 : 1);
  if (r) return r;
  double ret = sc[0] * c->kappa * c->kappa + (det ? sc[1] : 0.0);
  if (norm) *norm = eps * sqrt(ret) / c->V;
  return 0;
}

// ---------------------------------------------------------------------
// Action densities: slot 0 gauge ReTr[C2/2 D^2 + 2 sum F^dag F], slot 1
// bmass (Tr(U Ubar)/N - 1)^2, slot 2 det |plaqdet - 1|^2 (a<b), slot 3 mom^2
__global__ void __launch_bounds__(FB)
k_actions(const double2 *__restrict__ U, const double2 *__restrict__ dmu,
          const double2 *__restrict__ F, const double2 *__restrict__ pd,
          const double2 *__restrict__ mom, int V, double *__restrict__ partial,
          int stride) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  double g = 0, bm = 0, dt = 0, pm = 0;
  if (s < V) {
    const size_t M = (size_t)NN * V;
    Mat d, t, acc;
    mzero(acc);
    mload(d, dmu + s, V);
    mmadd<false, false>(acc, 0.5 * C2, d, d);
    for (int p = 0; p < NPLAQ; p++) {
      mload(t, F + p * M + s, V);
      mmadd<true, false>(acc, 2.0, t, t);
    }
#pragma unroll
    for (int i = 0; i < NCOL; i++) g += acc[i][i].x;
    for (int a = 0; a < NUMLINK; a++) {
      mload(t, U + a * M + s, V);
      double tr = retrace_dag(t, t) / NCOL - 1.0;
      bm += tr * tr;
      mload(t, mom + a * M + s, V);
      pm += retrace_dag(t, t);
      for (int b = a + 1; b < NUMLINK; b++) {
        double2 p = pd[(a * NUMLINK + b) * V + s];
        dt += (p.x - 1.0) * (p.x - 1.0) + p.y * p.y;
      }
    }
  }
  block_partial(g, partial, stride, 0);
  block_partial(bm, partial, stride, 1);
  block_partial(dt, partial, stride, 2);
  block_partial(pm, partial, stride, 3);
}

int actions_dev(susy_ctx *c, double *out4) {
  const int V = c->V;
  int nb = (V + FB - 1) / FB;
  hipLaunchKernelGGL(k_actions, dim3(nb), dim3(FB), 0, c->s, c->d_links, c->d_dmu,
                     c->d_fmunu, c->d_pdet, c->d_mom, V, c->d_part, nb);
  HIPCHK(hipGetLastError());
  for (int k = 0; k < 4; k++) {
    int r = reduce_partials(c, c->d_part + (size_t)k * nb, nb, k);
    if (r) return r;
  }
  double sc[4];
  int r = read_scalars(c, sc, 4);
  if (r) return r;
  out4[0] = sc[0] * c->kappa;
  out4[1] = sc[1] * c->kappa * c->p.bmass * c->p.bmass;
  out4[2] = sc[2] * c->p.kappa_u1;
  out4[3] = sc[3];
  return 0;
}

// ReTr[Ubar_a(x+b) Ubar_b(x) U_a(x) U_b(x+a)] summed over a > b; planes with
// the temporal link 3 go to the second slot (plaq.c:6-51)
__global__ void __launch_bounds__(FB)
k_plaquette(const double2 *__restrict__ U, const int *__restrict__ nbr, int V,
            double *__restrict__ partial, int stride) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  double ss = 0, st = 0;
  if (s < V) {
    const size_t M = (size_t)NN * V;
    for (int a = 1; a < NUMLINK; a++)
      for (int b = 0; b < a; b++) {
        int xa = nbr[g_fwd(a) * V + s], xb = nbr[g_fwd(b) * V + s];
        Mat u1, u2, t, t2;
        mzero(t);
        mzero(t2);
        mload(u1, U + b * M + s, V);
        mload(u2, U + a * M + s, V);
        mmadd<true, false>(t, 1.0, u1, u2);
        mload(u1, U + b * M + xa, V);
        mmadd<false, false>(t2, 1.0, t, u1);
        mload(u1, U + a * M + xb, V);
        double v = retrace_dag(u1, t2);
        if (a == 3 || b == 3) st += v; else ss += v;
      }
  }
  block_partial(ss, partial, stride, 0);
  block_partial(st, partial, stride, 1);
}

int plaquette_dev(susy_ctx *c, double *ss, double *st) {
  const int V = c->V;
  int nb = (V + FB - 1) / FB;
  hipLaunchKernelGGL(k_plaquette, dim3(nb), dim3(FB), 0, c->s, c->d_links, c->d_nbr, V,
                     c->d_part, nb);
  HIPCHK(hipGetLastError());
  int r = reduce_partials(c, c->d_part, nb, 0);
  if (!r) r = reduce_partials(c, c->d_part + nb, nb, 1);
  if (r) return r;
  double sc[2];
  r = read_scalars(c, sc, 2);
  if (r) return r;
  *ss = sc[0] / (6.0 * V);
  *st = sc[1] / (4.0 * V);
  return 0;
}
