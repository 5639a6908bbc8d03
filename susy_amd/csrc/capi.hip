// C ABI of libsusy_hip_U<N>: context management, resident operations, the
// RHMC trajectory driver and the drop-in replacements of the reference's
// fermion_op / DSq / congrad_multi / fermion_force / gauge_force.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include <algorithm>
#include <chrono>
#include <thread>

#include "../../include/susy_rhmc15.h"
#include "susy_internal.h"

#define RET(x)               \
  do {                       \
    int _r = (x);            \
    if (_r < 0) return _r;   \
  } while (0)

extern "C" int susy_ncol(void) { return NCOL; }
extern "C" int susy_volume(const susy_ctx *c) { return c ? c->V : SUSY_E_STATE; }

// The context's rational-approximation coefficients (the Nroot row of the
// tables of setup_rhmc.c:882-1031): out[0] ampdeg4, out[1..15] amp4,
// out[16..30] shift4, out[31] ampdeg8, out[32..46] amp8, out[47..61] shift8.
extern "C" int susy_rhmc_coeffs(const susy_ctx *c, double *out) {
  if (!c) return SUSY_E_STATE;
  if (!out) return SUSY_E_ARG;
  out[0] = c->ampdeg4;
  out[31] = c->ampdeg8;
  for (int j = 0; j < SUSY_RHMC15_NORDER; j++) {
    out[1 + j] = c->amp4[j];
    out[16 + j] = c->shift4[j];
    out[32 + j] = c->amp8[j];
    out[47 + j] = c->shift8[j];
  }
  return 0;
}

static bool g_tables_ready = false;

// Every gather id used by the fused kernels must name one of the 74 tables.
static int validate_tables(const PlaqTerms &pt, const ForceTerms &ft) {
  auto ok = [](int g) { return g >= 0 && g < NGATHER; };
  for (int p = 0; p < NPLAQ; p++)
    for (int n = 0; n < 3; n++)
      if (!ok(pt.pp_g1[p][n]) || !ok(pt.pp_g2[p][n]) || !ok(pt.pm_g1[p][n]) ||
          !ok(pt.pm_g2[p][n]) || pt.pp_c[p][n] < 0 || pt.pp_c[p][n] >= NUMLINK ||
          pt.pm_c[p][n] < 0 || pt.pm_c[p][n] >= NUMLINK || pt.pp_de[p][n] < 0 ||
          pt.pp_de[p][n] >= NPLAQ || pt.pm_ab[p][n] < 0 || pt.pm_ab[p][n] >= NPLAQ) {
        set_error("operator term table entry (%d,%d) invalid", p, n);
        return SUSY_E_STATE;
      }
  for (int c = 0; c < NUMLINK; c++)
    for (int n = 0; n < 6; n++)
      if (!ok(ft.g_abc[c][n]) || !ok(ft.g_mab[c][n]) || ft.ab[c][n] < 0 ||
          ft.ab[c][n] >= NPLAQ || ft.de[c][n] < 0 || ft.de[c][n] >= NPLAQ) {
        set_error("force term table entry (%d,%d) invalid", c, n);
        return SUSY_E_STATE;
      }
  // The operator kernel applies the DbplusPtoP terms at double weight in
  // place of DbplusPtoP + DbminusPtoP: every DbminusPtoP term must equal the
  // DbplusPtoP term with the same output plaquette and c.
  for (int p = 0; p < NPLAQ; p++)
    for (int n = 0; n < 3; n++) {
      int hit = 0;
      for (int m = 0; m < 3; m++)
        hit += pt.pp_c[p][m] == pt.pm_c[p][n] && pt.pp_de[p][m] == pt.pm_ab[p][n] &&
               pt.pp_g1[p][m] == pt.pm_g1[p][n] && pt.pp_g2[p][m] == pt.pm_g2[p][n] &&
               pt.pp_dt3[p][m] == pt.pm_dt2[p][n] && pt.pp_dt2[p][m] == pt.pm_dt3[p][n] &&
               pt.pp_tr[p][m] == pt.pm_tr[p][n];
      if (hit != 1) {
        set_error("DbminusPtoP term (%d,%d) has no matching DbplusPtoP term", p, n);
        return SUSY_E_STATE;
      }
    }
  // Likewise the force kernel applies F1Q at double weight in place of
  // F1Q + F2Q: the F2Q term of (a,b,d,e) must equal the F1Q term of (d,e,a,b).
  for (int c = 0; c < NUMLINK; c++)
    for (int n = 0; n < 6; n++) {
      int hit = 0;
      for (int m = 0; m < 6; m++)
        hit += ft.ab[c][m] == ft.de[c][n] && ft.de[c][m] == ft.ab[c][n] &&
               ft.g_abc[c][m] == ft.g_mab[c][n] && ft.g_mab[c][m] == ft.g_abc[c][n] &&
               ft.dt3[c][m] == ft.dt2m[c][n] && ft.dt2m[c][m] == ft.dt3[c][n] &&
               ft.perm1[c][m] == ft.perm2[c][n];
      if (hit != 1) {
        set_error("F2Q term (%d,%d) has no matching F1Q term", c, n);
        return SUSY_E_STATE;
      }
    }
  return 0;
}

// Host-only check of the term tables (callable without a GPU).
extern "C" int susy_validate_tables(void) {
  Geometry g;
  build_geometry(g, 2, 2, 2, 2);
  PlaqTerms pt;
  ForceTerms ft;
  build_plaq_terms(g.q_offset, pt);
  build_force_terms(g.q_offset, ft);
  return validate_tables(pt, ft);
}

static void free_ctx(susy_ctx *c) {
  sg_free(c);
  traj_free(c);
  for (int p = 0; p < 2; p++)
    for (hipEvent_t e : c->prof_ev[p]) (void)hipEventDestroy(e);
  void *ptrs[] = {c->d_nbr, c->d_t, c->d_links, c->d_mom, c->d_old_links, c->d_old_mom,
                  c->d_uinv, c->d_det, c->d_pdet, c->d_zw, c->d_dmu, c->d_fmunu,
                  c->d_force, c->f_links, c->f_uinv, c->f_pdet, c->d_aux, c->f_aux,
                  c->d_aux2, c->f_aux2,
                  c->d_tmp, c->f_tmp, c->d_boundary, c->d_part, c->d_scal, c->cg_ws,
                  c->cg_t64, c->cg_state, c->f_in, c->f_out};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  for (double2 *p : c->tf)
    if (p) (void)hipFree(p);
  for (double2 *p : c->traj_pool)
    if (p) (void)hipFree(p);
  if (c->h_scal) (void)hipHostFree(c->h_scal);
  if (c->s) (void)hipStreamDestroy(c->s);
  delete c;
}

extern "C" void susy_destroy(susy_ctx *c) {
  if (c) {
    (void)hipStreamSynchronize(c->s);
    free_ctx(c);
  }
}

template <typename T>
static int dalloc(T **p, size_t n) {
  hipError_t e = hipMalloc((void **)p, n * sizeof(T));
  if (e != hipSuccess) {
    set_error("hipMalloc of %zu bytes failed: %s", n * sizeof(T), hipGetErrorString(e));
    return SUSY_E_NOMEM;
  }
  return 0;
}

static int ctx_setup(susy_ctx *c) {
  const susy_params &p = c->p;
  build_geometry(c->geo, p.nx, p.ny, p.nz, p.nt);
  const int V = c->V = c->geo.V;
  c->kappa = (double)NCOL * 0.25 / p.lambda;
  c->doG = p.G > IMAG_TOL;
  int r = p.Nroot - 1;
  c->ampdeg4 = susy_rhmc15_ampdeg4[r];
  c->ampdeg8 = susy_rhmc15_ampdeg8[r];
  for (int j = 0; j < 15; j++) {
    c->amp4[j] = susy_rhmc15_amp4[r][j];
    c->shift4[j] = susy_rhmc15_shift4[r][j];
    c->amp8[j] = susy_rhmc15_amp8[r][j];
    c->shift8[j] = susy_rhmc15_shift8[r][j];
  }
  build_lambda(c->lambda_gen);
  if (!g_tables_ready) {
    PlaqTerms pt;
    ForceTerms ft;
    build_plaq_terms(c->geo.q_offset, pt);
    build_force_terms(c->geo.q_offset, ft);
    RET(validate_tables(pt, ft));
    RET(dslash_init_tables(pt));
    RET(force_init_tables(ft));
    g_tables_ready = true;
  }
  HIPCHK(hipStreamCreateWithFlags(&c->s, hipStreamNonBlocking));
  RET(noalloc_init_device(c));
  const size_t mat5 = (size_t)NUMLINK * NN * V, tfn = (size_t)NTF * NN * V;
  RET(dalloc(&c->d_nbr, (size_t)NGATHER * V));
  RET(dalloc(&c->d_t, (size_t)V));
  RET(dalloc(&c->d_links, mat5));
  RET(dalloc(&c->d_mom, mat5));
  RET(dalloc(&c->d_old_links, mat5));
  RET(dalloc(&c->d_old_mom, mat5));
  RET(dalloc(&c->d_uinv, mat5));
  RET(dalloc(&c->d_force, mat5));
  RET(dalloc(&c->d_det, (size_t)NUMLINK * V));
  RET(dalloc(&c->d_pdet, (size_t)NUMLINK * NUMLINK * V));
  RET(dalloc(&c->d_zw, (size_t)NUMLINK * NUMLINK * V));
  RET(dalloc(&c->d_dmu, (size_t)NN * V));
  RET(dalloc(&c->d_fmunu, (size_t)NPLAQ * NN * V));
  RET(dalloc(&c->f_links, mat5));
  RET(dalloc(&c->f_uinv, mat5));
  RET(dalloc(&c->f_pdet, (size_t)NUMLINK * NUMLINK * V));
  RET(dalloc(&c->d_aux, (size_t)12 * V));
  RET(dalloc(&c->f_aux, (size_t)6 * V));
  RET(dalloc(&c->d_aux2, (size_t)6 * V));
  RET(dalloc(&c->f_aux2, (size_t)6 * V));
  RET(dalloc(&c->d_tmp, tfn));
  RET(dalloc(&c->d_boundary, tfn));
  const int tsl = p.nx * p.ny * p.nz;
  int ds_blocks = 0;
  for (int order = 0; order <= 2; order++)
    for (int sites : {DS_SITES64, DS_SITES32})
      ds_blocks = std::max(ds_blocks, ds_grid(V, tsl, order, sites, nullptr));
  c->npart = std::max({ds_blocks, 2048, 4 * ((V + 255) / 256), (NUMLINK * V + 255) / 256}) + 64;
  RET(dalloc(&c->d_part, (size_t)c->npart * 2));
  RET(dalloc(&c->d_scal, (size_t)64));
  HIPCHK(hipHostMalloc((void **)&c->h_scal, sizeof(double) * 64));
  HIPCHK(hipMemcpy(c->d_nbr, c->geo.gathers.data(), sizeof(int) * NGATHER * V,
                   hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->d_t, c->geo.t.data(), sizeof(int) * V, hipMemcpyHostToDevice));
  HIPCHK(hipMemset(c->d_mom, 0, sizeof(double2) * mat5));
  // generators: one per site (index = lexicographic site number), one for
  // the node (index = volume) (make_lattice.c:48, setup.c:246)
  c->site_prn.resize(V);
  for (int i = 0; i < V; i++) {
    int lex = c->geo.x[i] + p.nx * (c->geo.y[i] + p.ny * (c->geo.z[i] + p.nz * c->geo.t[i]));
    prn_init(&c->site_prn[i], p.iseed, lex);
  }
  prn_init(&c->node_prn, p.iseed, V);
  return susy_cold_start(c);
}

extern "C" int susy_create(const susy_params *p, susy_ctx **out) {
  if (!p || !out) return SUSY_E_ARG;
  *out = nullptr;
  if (p->nx < 1 || p->ny < 1 || p->nz < 1 || p->nt < 1 ||
      (p->nx * p->ny * p->nz * p->nt) % 2 != 0) {
    set_error("lattice %dx%dx%dx%d: need an even number of sites", p->nx, p->ny, p->nz, p->nt);
    return SUSY_E_ARG;
  }
  if (p->PBC != 1 && p->PBC != -1) {
    set_error("PBC must be +1 or -1 (setup_offset.c:190-199)");
    return SUSY_E_ARG;
  }
  if (p->Norder != SUSY_RHMC15_NORDER || (p->Nroot != 1 && p->Nroot != 2)) {
    set_error("only Norder 15 with Nroot 1 or 2 is tabulated");
    return SUSY_E_ARG;
  }
  if (p->lambda <= 0) {
    set_error("lambda must be positive");
    return SUSY_E_ARG;
  }
  susy_ctx *c = new susy_ctx();
  c->p = *p;
  int r = ctx_setup(c);
  if (r) {
    free_ctx(c);
    return r;
  }
  *out = c;
  return 0;
}

extern "C" int susy_get_gather(susy_ctx *c, int g, int *out) {
  if (!c) return SUSY_E_STATE;
  if (g < 0 || g >= NGATHER || !out) return SUSY_E_ARG;
  memcpy(out, &c->geo.gathers[(size_t)g * c->V], sizeof(int) * c->V);
  return 0;
}

// ---------------------------------------------------------------------
// Field transfers (AoS on the host)
static int upload_aos(susy_ctx *c, const double *h, double2 *d, int nmat) {
  size_t bytes = (size_t)nmat * NN * c->V * sizeof(double2);
  HIPCHK(hipMemcpyAsync(c->d_boundary, h, bytes, hipMemcpyHostToDevice, c->s));
  RET(aos_to_soa(c, (const double *)c->d_boundary, d, nmat));
  HIPCHK(hipStreamSynchronize(c->s));
  return 0;
}
static int download_aos(susy_ctx *c, const double2 *d, double *h, int nmat) {
  size_t bytes = (size_t)nmat * NN * c->V * sizeof(double2);
  RET(soa_to_aos(c, d, (double *)c->d_boundary, nmat));
  HIPCHK(hipMemcpyAsync(h, c->d_boundary, bytes, hipMemcpyDeviceToHost, c->s));
  HIPCHK(hipStreamSynchronize(c->s));
  return 0;
}

extern "C" int susy_set_links(susy_ctx *c, const double *links) {
  if (!c) return SUSY_E_STATE;
  RET(upload_aos(c, links, c->d_links, NUMLINK));
  RET(fields_recompute(c));
  HIPCHK(hipStreamSynchronize(c->s));
  return 0;
}
extern "C" int susy_get_links(susy_ctx *c, double *links) {
  if (!c) return SUSY_E_STATE;
  return download_aos(c, c->d_links, links, NUMLINK);
}
extern "C" int susy_set_mom(susy_ctx *c, const double *mom) {
  if (!c) return SUSY_E_STATE;
  return upload_aos(c, mom, c->d_mom, NUMLINK);
}
extern "C" int susy_get_mom(susy_ctx *c, double *mom) {
  if (!c) return SUSY_E_STATE;
  return download_aos(c, c->d_mom, mom, NUMLINK);
}

// unit links (io_helpers.c:95-111)
extern "C" int susy_cold_start(susy_ctx *c) {
  if (!c) return SUSY_E_STATE;
  std::vector<double> h((size_t)c->V * NUMLINK * NN * 2, 0.0);
  for (size_t k = 0; k < (size_t)c->V * NUMLINK; k++)
    for (int i = 0; i < NCOL; i++) h[(k * NN + i * NCOL + i) * 2] = 1.0;
  return susy_set_links(c, h.data());
}

extern "C" int susy_get_fields(susy_ctx *c, double *plaqdet, double *uinv, double *dmu,
                               double *fmunu) {
  if (!c) return SUSY_E_STATE;
  HIPCHK(hipStreamSynchronize(c->s));
  if (plaqdet)
    HIPCHK(hipMemcpy(plaqdet, c->d_pdet, sizeof(double2) * NUMLINK * NUMLINK * c->V,
                     hipMemcpyDeviceToHost));
  if (uinv) RET(download_aos(c, c->d_uinv, uinv, NUMLINK));
  if (dmu) RET(download_aos(c, c->d_dmu, dmu, 1));
  if (fmunu) {
    // Fmunu as 10 separate fields of one matrix per site
    std::vector<double> tmp((size_t)c->V * NN * 2);
    for (int p = 0; p < NPLAQ; p++) {
      RET(download_aos(c, c->d_fmunu + (size_t)p * NN * c->V, tmp.data(), 1));
      memcpy(fmunu + (size_t)p * c->V * NN * 2, tmp.data(), tmp.size() * sizeof(double));
    }
  }
  return 0;
}

// MILC binary gauge file, natural order, FP32 links (io_lat.c:826-1006,
// io_lat_utils.c:519-618), with the sum29/sum31 checksums.
static uint32_t bswap32(uint32_t u) {
  return (u >> 24) | ((u >> 8) & 0xff00) | ((u << 8) & 0xff0000) | (u << 24);
}
extern "C" int susy_load_config(susy_ctx *c, const char *path) {
  if (!c || !path) return SUSY_E_ARG;
  FILE *fp = fopen(path, "rb");
  if (!fp) {
    set_error("cannot open %s", path);
    return SUSY_E_ARG;
  }
  uint32_t hdr[5], stamp_order[17], sums[2];
  int swap = 0;
  bool ok = fread(hdr, 4, 5, fp) == 5;
  if (ok && hdr[0] != 20103) {
    swap = bswap32(hdr[0]) == 20103;
    if (!swap) ok = false;
  }
  ok = ok && fread(stamp_order, 4, 17, fp) == 17 && fread(sums, 4, 2, fp) == 2;
  if (!ok) {
    fclose(fp);
    set_error("%s: not a MILC gauge file", path);
    return SUSY_E_ARG;
  }
  if (swap) {
    for (int k = 1; k < 5; k++) hdr[k] = bswap32(hdr[k]);
    stamp_order[16] = bswap32(stamp_order[16]);
    sums[0] = bswap32(sums[0]);
    sums[1] = bswap32(sums[1]);
  }
  if ((int)hdr[1] != c->p.nx || (int)hdr[2] != c->p.ny || (int)hdr[3] != c->p.nz ||
      (int)hdr[4] != c->p.nt || stamp_order[16] != 0) {
    fclose(fp);
    set_error("%s: dimensions or site order do not match", path);
    return SUSY_E_ARG;
  }
  const int V = c->V;
  const size_t nw = (size_t)NUMLINK * NN * 2;
  std::vector<float> buf(nw);
  std::vector<double> links((size_t)V * nw);
  uint32_t c29 = 0, c31 = 0;
  int r29 = 0, r31 = 0;
  for (int lex = 0; lex < V; lex++) {
    if (fread(buf.data(), sizeof(float), nw, fp) != nw) {
      fclose(fp);
      set_error("%s: short read", path);
      return SUSY_E_ARG;
    }
    uint32_t *w = (uint32_t *)buf.data();
    for (size_t k = 0; k < nw; k++) {
      if (swap) w[k] = bswap32(w[k]);
      uint32_t v = w[k];
      c29 ^= r29 ? (v << r29 | v >> (32 - r29)) : v;
      c31 ^= r31 ? (v << r31 | v >> (32 - r31)) : v;
      if (++r29 >= 29) r29 = 0;
      if (++r31 >= 31) r31 = 0;
    }
    int x = lex % c->p.nx, y = (lex / c->p.nx) % c->p.ny;
    int z = (lex / (c->p.nx * c->p.ny)) % c->p.nz, t = lex / (c->p.nx * c->p.ny * c->p.nz);
    int i = site_index(c->geo, x, y, z, t);
    for (size_t k = 0; k < nw; k++) links[(size_t)i * nw + k] = (double)buf[k];
  }
  fclose(fp);
  RET(susy_set_links(c, links.data()));
  return (c29 == sums[0] && c31 == sums[1]) ? 0 : 1;
}

// MILC serial gauge file of the context's links (the format susy_load_config
// reads; io_lat.c): header (magic 20103, nx ny nz nt, 64-byte time stamp,
// site order 0, the two rotating-XOR checksums of the payload), then per
// site in lexicographic order the 5 links as single-precision complex.
// stamp: 64 bytes to record, or NULL for the current time.
extern "C" int susy_save_config(susy_ctx *c, const char *path, const char *stamp) {
  if (!c || !path) return SUSY_E_ARG;
  const int V = c->V;
  const size_t nw = (size_t)NUMLINK * NN * 2;
  std::vector<double> links((size_t)V * nw);
  RET(susy_get_links(c, links.data()));
  std::vector<float> pay((size_t)V * nw);
  uint32_t c29 = 0, c31 = 0;
  int r29 = 0, r31 = 0;
  for (int lex = 0; lex < V; lex++) {
    const int x = lex % c->p.nx, y = (lex / c->p.nx) % c->p.ny;
    const int z = (lex / (c->p.nx * c->p.ny)) % c->p.nz, t = lex / (c->p.nx * c->p.ny * c->p.nz);
    const int i = site_index(c->geo, x, y, z, t);
    for (size_t k = 0; k < nw; k++) {
      const float f = (float)links[(size_t)i * nw + k];
      pay[(size_t)lex * nw + k] = f;
      uint32_t v;
      memcpy(&v, &f, 4);
      c29 ^= r29 ? (v << r29 | v >> (32 - r29)) : v;
      c31 ^= r31 ? (v << r31 | v >> (32 - r31)) : v;
      if (++r29 >= 29) r29 = 0;
      if (++r31 >= 31) r31 = 0;
    }
  }
  uint32_t hdr[5] = {20103u, (uint32_t)c->p.nx, (uint32_t)c->p.ny, (uint32_t)c->p.nz,
                     (uint32_t)c->p.nt};
  char st[64];
  memset(st, 0, sizeof(st));
  if (stamp) {
    memcpy(st, stamp, 64);
  } else {
    time_t now = time(nullptr);
    strftime(st, sizeof(st), "%a %b %e %H:%M:%S %Y", localtime(&now));
  }
  const uint32_t order = 0, sums[2] = {c29, c31};
  FILE *fp = fopen(path, "wb");
  if (!fp) {
    set_error("cannot create %s", path);
    return SUSY_E_ARG;
  }
  bool ok = fwrite(hdr, 4, 5, fp) == 5 && fwrite(st, 1, 64, fp) == 64 &&
            fwrite(&order, 4, 1, fp) == 1 && fwrite(sums, 4, 2, fp) == 2 &&
            fwrite(pay.data(), sizeof(float), pay.size(), fp) == pay.size();
  ok = (fclose(fp) == 0) && ok;
  if (!ok) {
    set_error("%s: write failed", path);
    return SUSY_E_ARG;
  }
  return 0;
}

// ---------------------------------------------------------------------
// Twist_Fermion handles
extern "C" int susy_tf_alloc(susy_ctx *c, int count) {
  if (!c) return SUSY_E_STATE;
  if (count < 1) return SUSY_E_ARG;
  int first = (int)c->tf.size();
  for (int k = 0; k < count; k++) {
    double2 *p = nullptr;
    RET(dalloc(&p, tf_elems(c)));
    HIPCHK(hipMemset(p, 0, sizeof(double2) * tf_elems(c)));
    c->tf.push_back(p);
  }
  return first;
}
static double2 *tfh(susy_ctx *c, int h) {
  if (!c || h < 0 || h >= (int)c->tf.size()) {
    set_error("bad Twist_Fermion handle %d", h);
    return nullptr;
  }
  return c->tf[h];
}
extern "C" int susy_tf_upload(susy_ctx *c, int h, const Twist_Fermion *src) {
  double2 *d = tfh(c, h);
  if (!d) return SUSY_E_ARG;
  return upload_aos(c, (const double *)src, d, NTF);
}
extern "C" int susy_tf_download(susy_ctx *c, int h, Twist_Fermion *dst) {
  double2 *d = tfh(c, h);
  if (!d) return SUSY_E_ARG;
  return download_aos(c, d, (double *)dst, NTF);
}

// FP32 staging for the FP32 operator form (c->f_in, c->f_out, c->f_tmp)
__global__ void k_to32(const double2 *__restrict__ a, float2 *__restrict__ b, size_t n) {
  size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) b[i] = make_float2((float)a[i].x, (float)a[i].y);
}
__global__ void k_to64(const float2 *__restrict__ a, double2 *__restrict__ b, size_t n) {
  size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) b[i] = make_double2((double)a[i].x, (double)a[i].y);
}
static int f32_stage(susy_ctx *c) {
  size_t n = tf_elems(c);
  if (!c->f_in) RET(dalloc(&c->f_in, n));
  if (!c->f_out) RET(dalloc(&c->f_out, n));
  if (!c->f_tmp) RET(dalloc(&c->f_tmp, n));
  return 0;
}

extern "C" int susy_fermion_op_dev(susy_ctx *c, int hs, int hd, int sign, int form) {
  double2 *s = tfh(c, hs), *d = tfh(c, hd);
  if (!s || !d) return SUSY_E_ARG;
  if (form == SUSY_FORM_FP32) {
    RET(f32_stage(c));
    size_t n = tf_elems(c);
    hipLaunchKernelGGL(k_to32, dim3((n + 255) / 256), dim3(256), 0, c->s, s, c->f_in, n);
    RET(dslash32(c, c->f_in, c->f_out, sign, nullptr, 0.f, nullptr, nullptr));
    hipLaunchKernelGGL(k_to64, dim3((n + 255) / 256), dim3(256), 0, c->s, c->f_out, d, n);
  } else {
    RET(dslash64(c, s, d, sign, nullptr, 0.0, nullptr, nullptr));
  }
  HIPCHK(hipStreamSynchronize(c->s));
  return 0;
}

// dest = D^dag D src + fmass^2 src (utilities.c:874-885)
static int dsq64(susy_ctx *c, const double2 *s, double2 *d) {
  double f2 = c->p.fmass > IMAG_TOL ? c->p.fmass * c->p.fmass : 0.0;
  int nb;
  return dsq_pair64(c, s, c->d_tmp, d, f2, c->d_part, &nb);
}
extern "C" int susy_stream_copy_dev(susy_ctx *c, int hs, int hd) {
  double2 *s = tfh(c, hs), *d = tfh(c, hd);
  if (!s || !d) return SUSY_E_ARG;
  RET(tf_stream_copy(c, d, s));
  HIPCHK(hipStreamSynchronize(c->s));
  return 0;
}

extern "C" int susy_dsq_dev(susy_ctx *c, int hs, int hd, int form) {
  double2 *s = tfh(c, hs), *d = tfh(c, hd);
  if (!s || !d) return SUSY_E_ARG;
  if (form == SUSY_FORM_FP32) {
    RET(f32_stage(c));
    size_t n = tf_elems(c);
    float f2 = c->p.fmass > IMAG_TOL ? (float)(c->p.fmass * c->p.fmass) : 0.f;
    hipLaunchKernelGGL(k_to32, dim3((n + 255) / 256), dim3(256), 0, c->s, s, c->f_in, n);
    int nb;
    RET(dsq_pair32(c, c->f_in, c->f_tmp, c->f_out, f2, c->d_part, &nb));
    hipLaunchKernelGGL(k_to64, dim3((n + 255) / 256), dim3(256), 0, c->s, c->f_out, d, n);
  } else {
    RET(dsq64(c, s, d));
  }
  HIPCHK(hipStreamSynchronize(c->s));
  return 0;
}

extern "C" int susy_congrad_multi_dev(susy_ctx *c, int hs, int hp, int nshift,
                                      const double *shifts, int MaxCG, double errormin,
                                      double *size_r, int form) {
  double2 *s = tfh(c, hs);
  if (!s || nshift < 1 || hp < 0 || hp + nshift > (int)c->tf.size()) return SUSY_E_ARG;
  std::vector<double2 *> ps(nshift);
  for (int j = 0; j < nshift; j++) ps[j] = c->tf[hp + j];
  double sr = 0;
  int it = cg_multi(c, s, ps.data(), nshift, shifts, MaxCG, errormin, &sr, form);
  if (size_r) *size_r = sr;
  return it;
}

extern "C" int susy_fermion_force_dev(susy_ctx *c, double eps, int hsol, int nsol,
                                      const double *amps, double *norm) {
  if (!c || hsol < 0 || nsol < 1 || hsol + nsol > (int)c->tf.size()) return SUSY_E_ARG;
  std::vector<double2 *> ps(nsol);
  for (int j = 0; j < nsol; j++) ps[j] = c->tf[hsol + j];
  RET(fermion_force_dev(c, eps, ps.data(), nsol, amps, norm));
  return 0;
}
extern "C" int susy_gauge_force_dev(susy_ctx *c, double eps, double *norm) {
  if (!c) return SUSY_E_STATE;
  return gauge_force_dev(c, eps, norm);
}
extern "C" int susy_update_u_dev(susy_ctx *c, double eps) {
  if (!c) return SUSY_E_STATE;
  RET(update_u_dev(c, eps));
  HIPCHK(hipStreamSynchronize(c->s));
  return 0;
}
extern "C" int susy_actions_dev(susy_ctx *c, double *out4) {
  if (!c) return SUSY_E_STATE;
  return actions_dev(c, out4);
}
// sum_x [ampdeg4 |src|^2 + sum_j amp4_j Re<src, sol_j>] (action.c:243-268)
static int fermion_action_ptr(susy_ctx *c, const double2 *src, double2 *const *sol,
                              double *out) {
  double s = 0, v;
  RET(tf_dot_re(c, src, src, &v));
  s += c->ampdeg4 * v;
  for (int j = 0; j < c->p.Norder; j++) {
    RET(tf_dot_re(c, src, sol[j], &v));
    s += c->amp4[j] * v;
  }
  *out = s;
  return 0;
}
extern "C" int susy_fermion_action_dev(susy_ctx *c, int hs, int hsol, double *out) {
  double2 *s = tfh(c, hs);
  if (!s || hsol < 0 || hsol + c->p.Norder > (int)c->tf.size()) return SUSY_E_ARG;
  return fermion_action_ptr(c, s, &c->tf[hsol], out);
}
extern "C" int susy_plaquette_dev(susy_ctx *c, double *ss, double *st) {
  if (!c) return SUSY_E_STATE;
  return plaquette_dev(c, ss, st);
}

extern "C" int susy_time_fermion_op(susy_ctx *c, int hs, int hd, int form, int reps,
                                    double *ms) {
  double2 *s = tfh(c, hs), *d = tfh(c, hd);
  if (!s || !d || reps < 1) return SUSY_E_ARG;
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  if (form == SUSY_FORM_FP32) {
    RET(f32_stage(c));
    size_t n = tf_elems(c);
    hipLaunchKernelGGL(k_to32, dim3((n + 255) / 256), dim3(256), 0, c->s, s, c->f_in, n);
    RET(dslash32(c, c->f_in, c->f_out, 1, nullptr, 0.f, nullptr, nullptr));
    HIPCHK(hipEventRecord(e0, c->s));
    for (int k = 0; k < reps; k++)
      RET(dslash32(c, c->f_in, c->f_out, 1, nullptr, 0.f, nullptr, nullptr));
    HIPCHK(hipEventRecord(e1, c->s));
  } else {
    RET(dslash64(c, s, d, 1, nullptr, 0.0, nullptr, nullptr));
    HIPCHK(hipEventRecord(e0, c->s));
    for (int k = 0; k < reps; k++) RET(dslash64(c, s, d, 1, nullptr, 0.0, nullptr, nullptr));
    HIPCHK(hipEventRecord(e1, c->s));
  }
  HIPCHK(hipEventSynchronize(e1));
  float t = 0;
  HIPCHK(hipEventElapsedTime(&t, e0, e1));
  *ms = t / reps;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return 0;
}

// ---------------------------------------------------------------------
// Heatbath on the host with the reference's generators, so that momenta and
// pseudofermions are bit-identical to the reference's (grsource.c:44-64,
// 200-303).  Each site draws an even number of Gaussians, so the shared
// Box-Muller pair cache of gaussian_rand_no (gaussrand.c:12-32) is empty at
// every site boundary and sites can be generated in parallel.
struct Gauss {
  int iset = 0;
  double gset = 0;
  double next(Prn *p) {
    if (iset == 0) {
      double v1, v2, r;
      do {
        v1 = 2.0 * prn_next(p) - 1.0;
        v2 = 2.0 * prn_next(p) - 1.0;
        r = v1 * v1 + v2 * v2;
      } while (r >= 1.0);
      double fac = sqrt(-log(r) / r);
      gset = v1 * fac;
      iset = 1;
      return v2 * fac;
    }
    iset = 0;
    return gset;
  }
};

template <typename F>
static void parallel_sites(int V, F f) {
  unsigned nt = std::min(16u, std::max(1u, std::thread::hardware_concurrency()));
  if (V < 4096) nt = 1;
  std::vector<std::thread> th;
  for (unsigned k = 0; k < nt; k++)
    th.emplace_back([&, k]() {
      for (int i = (int)((size_t)V * k / nt); i < (int)((size_t)V * (k + 1) / nt); i++) f(i);
    });
  for (auto &t : th) t.join();
}

// out (AoS, nmat matrices per site) = sum_j (g1 + i g2) Lambda_j, drawn in the
// reference's order: generator j outer, matrix k inner.
static void heatbath_fill(susy_ctx *c, double *out, int nmat, bool j_outer) {
  const double *lam = c->lambda_gen.data();
  parallel_sites(c->V, [&](int i) {
    Gauss g;
    Prn *p = &c->site_prn[i];
    double *o = out + (size_t)i * nmat * NN * 2;
    memset(o, 0, sizeof(double) * nmat * NN * 2);
    auto add = [&](int k, int j) {
      double gr = g.next(p), gi = g.next(p);
      const double *L = lam + (size_t)j * NN * 2;
      double *m = o + (size_t)k * NN * 2;
      for (int e = 0; e < NN; e++) {
        m[2 * e] += gr * L[2 * e] - gi * L[2 * e + 1];
        m[2 * e + 1] += gr * L[2 * e + 1] + gi * L[2 * e];
      }
    };
    if (j_outer) {
      for (int j = 0; j < NN; j++)
        for (int k = 0; k < nmat; k++) add(k, j);
    } else {
      for (int k = 0; k < nmat; k++)
        for (int j = 0; j < NN; j++) add(k, j);
    }
  });
}

static int ranmom(susy_ctx *c) {
  std::vector<double> h((size_t)c->V * NUMLINK * NN * 2);
  heatbath_fill(c, h.data(), NUMLINK, false);
  return upload_aos(c, h.data(), c->d_mom, NUMLINK);
}

// Restart the site and node generators from a seed, as at start-up
// (make_lattice.c:48, setup.c:246).
extern "C" int susy_reseed(susy_ctx *c, int seed) {
  if (!c) return SUSY_E_STATE;
  const susy_params &p = c->p;
  for (int i = 0; i < c->V; i++) {
    int lex = c->geo.x[i] + p.nx * (c->geo.y[i] + p.ny * (c->geo.z[i] + p.nz * c->geo.t[i]));
    prn_init(&c->site_prn[i], seed, lex);
  }
  prn_init(&c->node_prn, seed, c->V);
  c->p.iseed = seed;
  return 0;
}

extern "C" int susy_ranmom(susy_ctx *c) {
  if (!c) return SUSY_E_STATE;
  return ranmom(c);
}

// src = (D^dag D)^(1/8 Nroot) g: g Gaussian, then the heatbath rational
// function through one multi-shift solve with shift8
static int grsource(susy_ctx *c, double2 *src, double2 *const *psim, int form, int *iters) {
  std::vector<double> h((size_t)c->V * NTF * NN * 2);
  heatbath_fill(c, h.data(), NTF, true);
  RET(upload_aos(c, h.data(), src, NTF));
  double sr;
  int it = cg_multi(c, src, psim, c->p.Norder, c->shift8, c->p.niter,
                    c->p.error_per_site * c->p.error_per_site, &sr, form);
  if (it < 0) return it;
  *iters += it;
  RET(tf_axpby(c, src, c->ampdeg8, psim[0], c->amp8[0]));
  for (int j = 1; j < c->p.Norder; j++) RET(tf_axpby(c, src, 1.0, psim[j], c->amp8[j]));
  return 0;
}

extern "C" int susy_grsource_dev(susy_ctx *c, int hsrc, int hpsim, int form, int *iters) {
  double2 *s = tfh(c, hsrc);
  if (!s || hpsim < 0 || hpsim + c->p.Norder > (int)c->tf.size()) return SUSY_E_ARG;
  int it = 0;
  RET(grsource(c, s, &c->tf[hpsim], form, &it));
  if (iters) *iters = it;
  return 0;
}

// ---------------------------------------------------------------------
// One trajectory (update_o.c:20-282)
// State of the trajectory in progress between susy_traj_begin, the outer
// steps and susy_traj_end.
struct Traj {
  susy_ctx *c;
  int form;             // precision of the CG (SUSY_FORM_FP64 / SUSY_FORM_FP32)
  bool sg = false;      // learned surrogate on the inner level
  bool pure = false;    // fermions switched off (-DPUREGAUGE)
  bool split = false;   // large-shift poles on a middle level (susy_set_split)
  std::vector<double2 *> gsrc;
  std::vector<std::vector<double2 *>> psim;
  std::vector<std::vector<double2 *>> psim_mid;   // middle-level solutions
  double amp_out[16] = {0};   // outer-level amplitudes
  // middle level, fixed at susy_traj_begin: first pole, steps per half outer
  // step, weights (entries >= mid_first)
  int mid_first = 0, mid_steps = 1;
  double mid_w[16] = {0};
  int mid_iters = 0;          // middle-level CG iterations
  double gnorm = 0, max_gf = 0, fnorm[2] = {0, 0}, max_ff[2] = {0, 0};
  int iters = 0;        // CG iterations so far
  int step = 0;         // outer steps begun
  int sub = 0;          // sub-steps done in the current outer step
  bool open = false;    // between begin and end
  double p0[6] = {0, 0, 0, 0, 0, 0};   // start action parts
};
void traj_free(susy_ctx *c) {
  delete c->traj;
  c->traj = nullptr;
}

static const double OM_LAMBDA = 0.193, OM_TWO_LAMBDA = 0.386, OM_LAMBDA_MID = 0.614;

static int update_gauge_step(Traj &T, double eps, double *norm) {
  susy_ctx *c = T.c;
  int n = c->p.nstep_gauge;
  double nm, acc;
  // the surrogate mode adds the learned force to every inner force
  auto inner = [&](double e, double *nrm) -> int {
    RET(gauge_force_dev(c, e, nrm));
    if (T.sg) RET(sg_force_dev(c, e));
    return 0;
  };
  RET(inner(eps * OM_LAMBDA, &acc));
  for (int i = 1; i <= n; i++) {
    RET(update_u_dev(c, 0.5 * eps));
    RET(inner(eps * OM_LAMBDA_MID, &nm));
    acc += nm;
    RET(update_u_dev(c, 0.5 * eps));
    RET(inner(eps * (i < n ? OM_TWO_LAMBDA : OM_LAMBDA), &nm));
    acc += nm;
  }
  *norm = acc / n;
  return 0;
}

static int traj_force(Traj &T, int root, double eps) {
  double nm;
  RET(fermion_force_dev(T.c, eps, T.psim[root].data(), T.c->p.Norder, T.amp_out, &nm));
  T.fnorm[root] += nm;
  if (nm > T.max_ff[root]) T.max_ff[root] = nm;
  return 0;
}

static int g_verbose = 0;
extern "C" int susy_set_site_order(susy_ctx *c, int paired) {
  // paired: FP64 order in the low byte, FP32 order in the next (a value
  // below 256 sets both)
  const int o64 = paired & 255, o32 = paired >= 256 ? (paired >> 8) & 255 : o64;
  if (!c || paired < 0 || o64 > 2 || o32 > 2) {
    set_error("susy_set_site_order: orders must be 0, 1 or 2");
    return SUSY_E_ARG;
  }
  c->site_order = o64;
  c->site_order32 = o32;
  return 0;
}

extern "C" int susy_set_verbose(int v) {
  g_verbose = v;
  return 0;
}
static double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

static int traj_cg(Traj &T, int root, int *iters) {
  susy_ctx *c = T.c;
  double sr, t0 = g_verbose ? now_s() : 0;
  int it = cg_multi(c, T.gsrc[root], T.psim[root].data(), c->p.Norder, c->shift4, c->p.niter,
                    c->p.error_per_site * c->p.error_per_site, &sr, T.form);
  if (it < 0) return it;
  *iters += it;
  if (g_verbose) {
    fprintf(stderr, "susy: cg root %d: %d iterations, %.3f s\n", root, it, now_s() - t0);
    fflush(stderr);
  }
  return 0;
}

// Middle-level kick of the split form: for every root the multi-shift solve
// over the large shifts only (shift4[j] from the first middle pole on; it
// converges in few iterations) and the force of those poles with the middle
// weights.
static int mid_kick(Traj &T, double eps) {
  susy_ctx *c = T.c;
  const int j0 = T.mid_first, nm = c->p.Norder - j0;
  for (int n = 0; n < c->p.Nroot; n++) {
    double sr, fn;
    int it = cg_multi(c, T.gsrc[n], T.psim_mid[n].data(), nm, c->shift4 + j0, c->p.niter,
                      c->p.error_per_site * c->p.error_per_site, &sr, T.form);
    if (it < 0) return it;
    T.iters += it;
    T.mid_iters += it;
    RET(fermion_force_dev(c, eps, T.psim_mid[n].data(), nm, T.mid_w + j0, &fn));
  }
  return 0;
}

// Middle level of the split form: mid_steps Omelyan steps advancing the
// links by len, the middle kicks at the step points and the whole gauge
// level (advancing len / (2 mid_steps)) between them.
static int mid_level(Traj &T, double len, double *norm) {
  susy_ctx *c = T.c;
  const int n = T.mid_steps;
  const double m = len / n, g = m / (2.0 * c->p.nstep_gauge);
  double tr, acc = 0;
  RET(mid_kick(T, m * OM_LAMBDA));
  for (int i = 1; i <= n; i++) {
    RET(update_gauge_step(T, g, &tr));
    acc += tr;
    RET(mid_kick(T, m * OM_LAMBDA_MID));
    RET(update_gauge_step(T, g, &tr));
    acc += tr;
    RET(mid_kick(T, m * (i < n ? OM_TWO_LAMBDA : OM_LAMBDA)));
  }
  *norm = acc / (2 * n);
  return 0;
}

// Outer (fermion) kick of size eps at the current links: the solves (unless
// the caller's psim already belong to these links), every root's force, and
// in the surrogate mode the removal of the learned force that the inner
// level applies, so that the outer level integrates S_F - S_theta.
static int outer_kick(Traj &T, double eps, bool solve, int *iters) {
  susy_ctx *c = T.c;
  if (T.pure) return 0;
  RET(sg_record_begin(c));
  for (int n = 0; n < c->p.Nroot; n++) {
    if (solve) RET(traj_cg(T, n, iters));
    RET(traj_force(T, n, eps));
    RET(sg_record_root(c));
  }
  RET(sg_record_end(c));
  if (T.sg) RET(sg_force_dev(c, -eps));
  return 0;
}

// Sub-steps of update_step's outer loop (update_o.c:80-113).  An outer step
// is two halves -- the gauge level, then for every root the solve and its
// fermion kick, at lambda_mid the first time and at 2 lambda (lambda on the
// last step) the second -- and each half is Nroot sub-steps: sub-step n
// solves and kicks root n, the first also runs the gauge level before it and
// the last also closes the outer kick (the surrogate's removal in that form).
// The operations and their order are those of the whole step.
static int sub_per_half(const Traj &T) { return T.pure ? 1 : T.c->p.Nroot; }

static int traj_substep(Traj &T) {
  susy_ctx *c = T.c;
  const double f_eps = c->p.traj_length / (double)c->p.nstep;
  const double g_eps = f_eps / (double)(2.0 * c->p.nstep_gauge);
  const int per = sub_per_half(T), half = T.sub / per, n = T.sub % per;
  if (T.sub == 0) T.step++;
  const double eps = f_eps * (half == 0 ? OM_LAMBDA_MID
                                        : (T.step < c->p.nstep ? OM_TWO_LAMBDA : OM_LAMBDA));
  if (n == 0) {
    double tr;
    if (T.split)
      RET(mid_level(T, 0.5 * f_eps, &tr));
    else
      RET(update_gauge_step(T, g_eps, &tr));
    T.gnorm += tr;
    if (tr > T.max_gf) T.max_gf = tr;
    if (!T.pure) RET(sg_record_begin(c));
  }
  if (!T.pure) {
    RET(traj_cg(T, n, &T.iters));
    RET(traj_force(T, n, eps));
    RET(sg_record_root(c));
    if (n == per - 1) {
      RET(sg_record_end(c));
      if (T.sg) RET(sg_force_dev(c, -eps));
    }
  }
  T.sub = (T.sub + 1) % (2 * per);
  return 0;
}

// One outer step: gauge level, solves and fermion kick at the midpoint, gauge
// level, solves and the closing kick.
static int outer_step(Traj &T) {
  for (int k = 0; k < 2 * sub_per_half(T); k++) RET(traj_substep(T));
  return 0;
}

static int traj_action(Traj &T, double *parts) {
  susy_ctx *c = T.c;
  double a4[4];
  RET(actions_dev(c, a4));
  double tot = a4[0] + (c->p.bmass > IMAG_TOL ? a4[1] : 0.0) +
               (c->p.kappa_u1 > IMAG_TOL ? a4[2] : 0.0);
  double f = 0;
  for (int n = 0; n < (T.pure ? 0 : c->p.Nroot); n++) {
    double fa;
    RET(fermion_action_ptr(c, T.gsrc[n], T.psim[n].data(), &fa));
    f += fa;
  }
  tot += f + a4[3];
  parts[0] = a4[0];
  parts[1] = c->p.bmass > IMAG_TOL ? a4[1] : 0.0;
  parts[2] = c->p.kappa_u1 > IMAG_TOL ? a4[2] : 0.0;
  parts[3] = f;
  parts[4] = a4[3];
  parts[5] = tot;
  return 0;
}

__global__ void k_neg(double2 *__restrict__ y, const double2 *__restrict__ x, size_t n) {
  size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) y[i] = make_double2(-x[i].x, -x[i].y);
}

// Start of a trajectory (update_o.c:127-208): momenta, pseudofermions and
// their solves, the start action, the saved links and momenta, and the first
// fermion kick of update_step (update_o.c:74-82).  An unfinished trajectory
// is abandoned (links and momenta stay as they are).
extern "C" int susy_traj_begin(susy_ctx *c, int form) {
  if (!c) return SUSY_E_STATE;
  if ((form & ~(SUSY_FORM_FP32 | SUSY_FORM_SURROGATE | SUSY_FORM_PUREGAUGE | SUSY_FORM_SPLIT)) !=
      0) {
    set_error("susy_traj_begin: unknown form bits 0x%x", form);
    return SUSY_E_ARG;
  }
  const bool sg = (form & SUSY_FORM_SURROGATE) != 0, pure = (form & SUSY_FORM_PUREGAUGE) != 0;
  const bool split = (form & SUSY_FORM_SPLIT) != 0;
  if ((sg || split) && pure) {
    set_error("susy_traj_begin: the surrogate and split forms need the fermions");
    return SUSY_E_ARG;
  }
  if (split && c->split_first < 1) {
    set_error("susy_traj_begin: split form without susy_set_split");
    return SUSY_E_STATE;
  }
  if (sg && !c->sg) {
    set_error("susy_traj_begin: surrogate form without susy_sg_init");
    return SUSY_E_STATE;
  }
  if (c->p.nstep < 1 || c->p.nstep_gauge < 1) {
    set_error("susy_traj_begin: nstep and nstep_gauge must be positive");
    return SUSY_E_ARG;
  }
  if (!c->traj) c->traj = new Traj;
  Traj &T = *c->traj;
  T = Traj();
  T.c = c;
  T.form = form & SUSY_FORM_FP32;
  T.sg = sg;
  T.pure = pure;
  T.split = split;
  const int Nr = c->p.Nroot, No = c->p.Norder;
  const int nmid = split ? No - c->split_first : 0;
  if (split) {
    T.mid_first = c->split_first;
    T.mid_steps = c->split_nmid;
    for (int j = 0; j < No; j++) T.mid_w[j] = c->split_w[j];
  }
  for (int j = 0; j < No; j++)
    T.amp_out[j] = split && j >= T.mid_first ? c->amp4[j] - T.mid_w[j] : c->amp4[j];
  if (!pure) {
    // per-trajectory fields are kept in the context across calls
    std::vector<double2 *> &pool = c->traj_pool;
    size_t need = (size_t)Nr * (1 + No + nmid);
    while (pool.size() < need) {
      double2 *p;
      RET(dalloc(&p, tf_elems(c)));
      pool.push_back(p);
    }
    T.gsrc.resize(Nr);
    T.psim.resize(Nr);
    for (int n = 0; n < Nr; n++) {
      T.gsrc[n] = pool[n * (1 + No)];
      T.psim[n].assign(pool.begin() + n * (1 + No) + 1, pool.begin() + (n + 1) * (1 + No));
    }
    T.psim_mid.resize(Nr);
    for (int n = 0; n < Nr; n++) {
      auto first = pool.begin() + Nr * (1 + No) + n * nmid;
      T.psim_mid[n].assign(first, first + nmid);
    }
  }
  RET(ranmom(c));
  for (int n = 0; n < (pure ? 0 : Nr); n++) {
    double t0 = g_verbose ? now_s() : 0;
    int before = T.iters;
    RET(grsource(c, T.gsrc[n], T.psim[n].data(), T.form, &T.iters));
    if (g_verbose) {
      fprintf(stderr, "susy: heatbath root %d: %d iterations, %.3f s\n", n, T.iters - before,
              now_s() - t0);
      fflush(stderr);
    }
  }
  for (int n = 0; n < (pure ? 0 : Nr); n++) RET(traj_cg(T, n, &T.iters));
  RET(traj_action(T, T.p0));
  const size_t mat5 = (size_t)NUMLINK * NN * c->V;
  HIPCHK(hipMemcpyAsync(c->d_old_links, c->d_links, sizeof(double2) * mat5,
                        hipMemcpyDeviceToDevice, c->s));
  HIPCHK(hipMemcpyAsync(c->d_old_mom, c->d_mom, sizeof(double2) * mat5,
                        hipMemcpyDeviceToDevice, c->s));
  RET(outer_kick(T, c->p.traj_length / (double)c->p.nstep * OM_LAMBDA, false, &T.iters));
  HIPCHK(hipStreamSynchronize(c->s));
  T.open = true;
  return 0;
}

// One outer MD step of the open trajectory, complete on the device when it
// returns; returns the steps done so far.
extern "C" int susy_traj_step(susy_ctx *c) {
  if (!c) return SUSY_E_STATE;
  if (!c->traj || !c->traj->open || c->traj->step >= c->p.nstep || c->traj->sub != 0) {
    set_error("susy_traj_step: no open trajectory at an outer-step boundary with steps left");
    return SUSY_E_STATE;
  }
  Traj &T = *c->traj;
  int r = outer_step(T);
  if (r < 0) {
    T.open = false;
    return r;
  }
  HIPCHK(hipStreamSynchronize(c->s));
  return T.step;
}

// One sub-step of the open trajectory (the solve and fermion kick of one
// root; see traj_substep), complete on the device when it returns; returns the
// sub-steps done in the trajectory so far (2 Nroot per outer step).
extern "C" int susy_traj_substep(susy_ctx *c) {
  if (!c) return SUSY_E_STATE;
  if (!c->traj || !c->traj->open || (c->traj->step >= c->p.nstep && c->traj->sub == 0)) {
    set_error("susy_traj_substep: no open trajectory with sub-steps left");
    return SUSY_E_STATE;
  }
  Traj &T = *c->traj;
  int r = traj_substep(T);
  if (r < 0) {
    T.open = false;
    return r;
  }
  HIPCHK(hipStreamSynchronize(c->s));
  const int per = 2 * sub_per_half(T);
  return T.sub == 0 ? T.step * per : (T.step - 1) * per + T.sub;
}

// End of the trajectory (update_o.c:209-282): the end action (the last
// solves belong to the final links), the Metropolis test with one draw of the
// node generator, and on rejection the saved links, reversed momenta and
// recomputed link fields.  Returns 1 (accepted) or 0 (rejected).
extern "C" int susy_traj_end(susy_ctx *c, double *stats) {
  if (!c) return SUSY_E_STATE;
  if (!c->traj || !c->traj->open || c->traj->step != c->p.nstep || c->traj->sub != 0) {
    set_error("susy_traj_end: the open trajectory has not made its %d steps", c->p.nstep);
    return SUSY_E_STATE;
  }
  Traj &T = *c->traj;
  T.open = false;
  const int Nr = c->p.Nroot;
  double p1[6];
  const double *p0 = T.p0;
  RET(traj_action(T, p1));
  double change = p1[5] - p0[5];
  if (fabs(change) > 1e20 || change != change) change = 1.0e20;
  double xr = prn_next(&c->node_prn);
  int accept = 1;
  if (exp(-change) < xr) {
    accept = 0;
    if (c->p.traj_length > 0.0) {
      const size_t mat5 = (size_t)NUMLINK * NN * c->V;
      HIPCHK(hipMemcpyAsync(c->d_links, c->d_old_links, sizeof(double2) * mat5,
                            hipMemcpyDeviceToDevice, c->s));
      hipLaunchKernelGGL(k_neg, dim3((mat5 + 255) / 256), dim3(256), 0, c->s, c->d_mom,
                         c->d_old_mom, mat5);
      HIPCHK(hipGetLastError());
      RET(fields_recompute(c));
    }
  }
  HIPCHK(hipStreamSynchronize(c->s));
  if (stats) {
    stats[0] = p0[5]; stats[1] = p1[5]; stats[2] = change; stats[3] = accept;
    stats[4] = T.iters;
    stats[5] = T.gnorm / (2.0 * c->p.nstep); stats[6] = T.max_gf;
    for (int n = 0; n < 2; n++) {
      stats[7 + 2 * n] = n < Nr && !T.pure ? T.fnorm[n] / (2.0 * c->p.nstep) : 0;
      stats[8 + 2 * n] = n < Nr && !T.pure ? T.max_ff[n] : 0;
    }
    for (int k = 0; k < 6; k++) { stats[11 + k] = p0[k]; stats[17 + k] = p1[k]; }
    stats[23] = T.mid_iters;
  }
  return accept;
}

extern "C" int susy_traj_iters(susy_ctx *c, int *iters, int *mid_iters) {
  if (!c || !c->traj) return SUSY_E_STATE;
  if (iters) *iters = c->traj->iters;
  if (mid_iters) *mid_iters = c->traj->mid_iters;
  return 0;
}

extern "C" int susy_set_split(susy_ctx *c, int first_pole, int nstep_mid, const double *weights) {
  if (!c) return SUSY_E_STATE;
  if (first_pole < 1 || first_pole >= c->p.Norder || nstep_mid < 1) {
    set_error("susy_set_split: need 1 <= first_pole < %d and nstep_mid >= 1", c->p.Norder);
    return SUSY_E_ARG;
  }
  c->split_first = first_pole;
  c->split_nmid = nstep_mid;
  for (int j = 0; j < c->p.Norder; j++)
    c->split_w[j] = j < first_pole ? 0.0 : weights ? weights[j - first_pole] : c->amp4[j];
  return 0;
}

extern "C" int susy_update(susy_ctx *c, int form, double *stats) {
  RET(susy_traj_begin(c, form));
  for (int s = 1; s <= c->p.nstep; s++) RET(susy_traj_step(c));
  return susy_traj_end(c, stats);
}

// =====================================================================
// Drop-in boundary: the reference's symbols, backed by one context built
// from the bound geometry and the caller's globals.
static struct {
  susy_ctx *c = nullptr;
  susy_params p;
  bool have_lattice = false, have_couplings = false;
  char *sites = nullptr;
  size_t site_bytes = 0, link_off = 0, mom_off = 0;
  int *Norder = nullptr, *total_iters = nullptr;
  Real **shift = nullptr, **amp4 = nullptr;
  std::vector<double2 *> sol;   // device fields for multi-field calls
  std::vector<double> host;     // staging for site-array gathers
} g_drop;

static int drop_ctx() {
  if (g_drop.c) return 0;
  if (!g_drop.have_lattice || !g_drop.have_couplings) {
    set_error("susy_bind_lattice and susy_bind_couplings must be called first");
    return SUSY_E_STATE;
  }
  return susy_create(&g_drop.p, &g_drop.c);
}

extern "C" int susy_bind_lattice(int nx, int ny, int nz, int nt, int PBC) {
  if (g_drop.c) {
    susy_destroy(g_drop.c);
    g_drop.c = nullptr;
  }
  susy_params &p = g_drop.p;
  memset(&p, 0, sizeof(p));
  p.nx = nx; p.ny = ny; p.nz = nz; p.nt = nt; p.PBC = PBC;
  p.iseed = 0; p.Nroot = 1; p.Norder = 15; p.lambda = 1.0;
  p.niter = 1000; p.error_per_site = 1e-5; p.traj_length = 1; p.nstep = 1;
  p.nstep_gauge = 1;
  g_drop.have_lattice = true;
  return 0;
}

extern "C" int susy_bind_couplings(double lambda, double kappa_u1, double bmass,
                                   double fmass, double G) {
  if (!g_drop.have_lattice) {
    set_error("susy_bind_lattice must be called first");
    return SUSY_E_STATE;
  }
  if (g_drop.c) {
    susy_destroy(g_drop.c);
    g_drop.c = nullptr;
  }
  g_drop.p.lambda = lambda; g_drop.p.kappa_u1 = kappa_u1; g_drop.p.bmass = bmass;
  g_drop.p.fmass = fmass; g_drop.p.G = G;
  g_drop.have_couplings = true;
  return drop_ctx();
}

extern "C" int susy_bind_sites(void *lattice, size_t site_bytes, size_t link_offset,
                               size_t mom_offset) {
  if (!lattice || site_bytes < sizeof(susy_matrix)) return SUSY_E_ARG;
  g_drop.sites = (char *)lattice;
  g_drop.site_bytes = site_bytes;
  g_drop.link_off = link_offset;
  g_drop.mom_off = mom_offset;
  return 0;
}

extern "C" int susy_bind_rhmc(int *Norder, Real **shift, Real **amp4, int *total_iters) {
  g_drop.Norder = Norder;
  g_drop.shift = shift;
  g_drop.amp4 = amp4;
  g_drop.total_iters = total_iters;
  return 0;
}

static void gather_sites(size_t off, double *out) {
  const size_t per = sizeof(susy_matrix) * NUMLINK;
  for (int i = 0; i < g_drop.c->V; i++)
    memcpy(out + (size_t)i * per / sizeof(double), g_drop.sites + i * g_drop.site_bytes + off,
           per);
}
static void scatter_sites(size_t off, const double *in) {
  const size_t per = sizeof(susy_matrix) * NUMLINK;
  for (int i = 0; i < g_drop.c->V; i++)
    memcpy(g_drop.sites + i * g_drop.site_bytes + off, in + (size_t)i * per / sizeof(double),
           per);
}

extern "C" int susy_links_changed(void) {
  RET(drop_ctx());
  if (!g_drop.sites) {
    set_error("susy_bind_sites must be called first");
    return SUSY_E_STATE;
  }
  g_drop.host.resize((size_t)g_drop.c->V * NUMLINK * NN * 2);
  gather_sites(g_drop.link_off, g_drop.host.data());
  return susy_set_links(g_drop.c, g_drop.host.data());
}

static void fatal(const char *what) {
  fprintf(stderr, "susy_hip: %s failed: %s\n", what, susy_last_error());
  fflush(stderr);
  exit(1);
}

static int drop_sol(int n) {
  while ((int)g_drop.sol.size() < n) {
    double2 *p;
    RET(dalloc(&p, tf_elems(g_drop.c)));
    g_drop.sol.push_back(p);
  }
  return 0;
}

extern "C" void fermion_op(Twist_Fermion *src, Twist_Fermion *dest, int sign) {
  if (sign != 1 && sign != -1) {
    fprintf(stderr, "Error: incorrect sign in fermion_op: %d\n", sign);
    exit(1);
  }
  susy_ctx *c;
  if (drop_ctx() || drop_sol(2)) fatal("fermion_op");
  c = g_drop.c;
  if (upload_aos(c, (const double *)src, g_drop.sol[0], NTF) ||
      dslash64(c, g_drop.sol[0], g_drop.sol[1], sign, nullptr, 0.0, nullptr, nullptr) ||
      download_aos(c, g_drop.sol[1], (double *)dest, NTF))
    fatal("fermion_op");
}

extern "C" void DSq(Twist_Fermion *src, Twist_Fermion *dest) {
  susy_ctx *c;
  if (drop_ctx() || drop_sol(2)) fatal("DSq");
  c = g_drop.c;
  if (upload_aos(c, (const double *)src, g_drop.sol[0], NTF) ||
      dsq64(c, g_drop.sol[0], g_drop.sol[1]) ||
      download_aos(c, g_drop.sol[1], (double *)dest, NTF))
    fatal("DSq");
}

extern "C" int congrad_multi(Twist_Fermion *src, Twist_Fermion **psim, int MaxCG,
                             Real errormin, Real *size_r) {
  if (drop_ctx()) fatal("congrad_multi");
  if (!g_drop.Norder || !g_drop.shift) {
    set_error("susy_bind_rhmc must be called first");
    fatal("congrad_multi");
  }
  susy_ctx *c = g_drop.c;
  int N = *g_drop.Norder;
  if (drop_sol(N + 1)) fatal("congrad_multi");
  if (upload_aos(c, (const double *)src, g_drop.sol[N], NTF)) fatal("congrad_multi");
  double sr = 0;
  int before = c->total_iters;
  int it = cg_multi(c, g_drop.sol[N], g_drop.sol.data(), N, *g_drop.shift, MaxCG, errormin,
                    &sr, SUSY_FORM_FP64);
  if (it < 0) fatal("congrad_multi");
  for (int j = 0; j < N; j++)
    if (download_aos(c, g_drop.sol[j], (double *)psim[j], NTF)) fatal("congrad_multi");
  if (g_drop.total_iters) *g_drop.total_iters += c->total_iters - before;
  *size_r = sr;
  return it;
}

extern "C" double fermion_force(Real eps, Twist_Fermion *src, Twist_Fermion **sol) {
  (void)src;
  if (drop_ctx()) fatal("fermion_force");
  if (!g_drop.Norder || !g_drop.amp4 || !g_drop.sites) {
    set_error("susy_bind_rhmc and susy_bind_sites must be called first");
    fatal("fermion_force");
  }
  susy_ctx *c = g_drop.c;
  int N = *g_drop.Norder;
  if (drop_sol(N)) fatal("fermion_force");
  for (int j = 0; j < N; j++)
    if (upload_aos(c, (const double *)sol[j], g_drop.sol[j], NTF)) fatal("fermion_force");
  g_drop.host.resize((size_t)c->V * NUMLINK * NN * 2);
  gather_sites(g_drop.mom_off, g_drop.host.data());
  if (susy_set_mom(c, g_drop.host.data())) fatal("fermion_force");
  double nm = 0;
  if (fermion_force_dev(c, eps, g_drop.sol.data(), N, *g_drop.amp4, &nm)) fatal("fermion_force");
  if (susy_get_mom(c, g_drop.host.data())) fatal("fermion_force");
  scatter_sites(g_drop.mom_off, g_drop.host.data());
  return nm;
}

extern "C" double gauge_force(Real eps) {
  if (drop_ctx()) fatal("gauge_force");
  if (!g_drop.sites) {
    set_error("susy_bind_sites must be called first");
    fatal("gauge_force");
  }
  susy_ctx *c = g_drop.c;
  g_drop.host.resize((size_t)c->V * NUMLINK * NN * 2);
  gather_sites(g_drop.mom_off, g_drop.host.data());
  if (susy_set_mom(c, g_drop.host.data())) fatal("gauge_force");
  double nm = 0;
  if (gauge_force_dev(c, eps, &nm)) fatal("gauge_force");
  if (susy_get_mom(c, g_drop.host.data())) fatal("gauge_force");
  scatter_sites(g_drop.mom_off, g_drop.host.data());
  return nm;
}
