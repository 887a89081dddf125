/*
 * oracle_c_body.h -- TEST INFRASTRUCTURE: body of the C oracle (oracle_c.c), instantiated once
 * for OC_REAL = float (suffix f32) and once for double (f64). Every expression keeps the float
 * order of oracle/fdtd_oracle.py (which follows the reference's JAX expressions); the cited
 * reference lines are listed in oracle_c.c.
 */
#define R OC_REAL
#define OC_CAT2(a, b) a##_##b
#define OC_CAT(a, b) OC_CAT2(a, b)
#define FN(name) OC_CAT(name, OC_SUF)

typedef struct FN(oc_ctx) {
  int nx, ny, nz;
  size_t N, sx, sy; /* strides of i and j */
  const R* E;
  const R* Hp; /* H before the H half-step */
  const R* H;
} FN(oc_ctx);

/* Field value with the zero halo beyond the domain (pad_fields, core/misc.py:615-641). */
static inline R FN(oc_at)(const R* F, const FN(oc_ctx) * x, int c, int i, int j, int k) {
  if (i < 0 || j < 0 || k < 0 || i >= x->nx || j >= x->ny || k >= x->nz) return (R)0;
  return F[c * x->N + i * x->sx + j * x->sy + k];
}

/* Time-centred H (H_prev + H) / 2 (update.py:1084-1102), zero beyond the domain. */
static inline R FN(oc_hc)(const FN(oc_ctx) * x, int c, int i, int j, int k) {
  if (i < 0 || j < 0 || k < 0 || i >= x->nx || j >= x->ny || k >= x->nz) return (R)0;
  size_t o = c * x->N + i * x->sx + j * x->sy + k;
  return (x->Hp[o] + x->H[o]) / (R)2;
}

/* interpolate_fields (curl.py:86-224): E and the time-centred H at the Ez point of cell (i,j,k);
 * out = (ex, ey, ez, hx, hy, hz). */
static void FN(oc_interp)(const FN(oc_ctx) * x, int i, int j, int k, R* out) {
  const R h = (R)0.5;
  const R* E = x->E;
  R ex_lo = h * (FN(oc_at)(E, x, 0, i, j, k) + FN(oc_at)(E, x, 0, i - 1, j, k));
  R ex_hi = h * (FN(oc_at)(E, x, 0, i, j, k + 1) + FN(oc_at)(E, x, 0, i - 1, j, k + 1));
  out[0] = (ex_lo + ex_hi) / (R)2;
  R ey_lo = h * (FN(oc_at)(E, x, 1, i, j, k) + FN(oc_at)(E, x, 1, i, j - 1, k));
  R ey_hi = h * (FN(oc_at)(E, x, 1, i, j, k + 1) + FN(oc_at)(E, x, 1, i, j - 1, k + 1));
  out[1] = (ey_lo + ey_hi) / (R)2;
  out[2] = FN(oc_at)(E, x, 2, i, j, k);
  out[3] = h * (FN(oc_hc)(x, 0, i, j, k) + FN(oc_hc)(x, 0, i, j - 1, k));
  out[4] = h * (FN(oc_hc)(x, 1, i, j, k) + FN(oc_hc)(x, 1, i - 1, j, k));
  R lo_x = h * (FN(oc_hc)(x, 2, i, j, k) + FN(oc_hc)(x, 2, i - 1, j, k));
  R lo_xy = h * (lo_x + h * (FN(oc_hc)(x, 2, i, j - 1, k) + FN(oc_hc)(x, 2, i - 1, j - 1, k)));
  R hi_x = h * (FN(oc_hc)(x, 2, i, j, k + 1) + FN(oc_hc)(x, 2, i - 1, j, k + 1));
  R hi_xy = h * (hi_x + h * (FN(oc_hc)(x, 2, i, j - 1, k + 1) + FN(oc_hc)(x, 2, i - 1, j - 1, k + 1)));
  out[5] = (lo_xy + hi_xy) / (R)2;
}

/* Slab-local offset of cell (i, j, k) in a slab's psi arrays (slab shape: the domain with `axis`
 * cut to [lo, hi)), or -1 when the cell lies outside the slab. */
static inline long FN(oc_slab_off)(const oc_slab* s, int nx, int ny, int nz, int i, int j, int k) {
  int L = s->hi - s->lo;
  if (s->axis == 0) {
    if (i < s->lo || i >= s->hi) return -1;
    return ((long)(i - s->lo) * ny + j) * nz + k;
  }
  if (s->axis == 1) {
    if (j < s->lo || j >= s->hi) return -1;
    return ((long)i * L + (j - s->lo)) * nz + k;
  }
  if (k < s->lo || k >= s->hi) return -1;
  return ((long)i * ny + j) * L + (k - s->lo);
}

/* apply_cpml (curl.py:284-308 / 371-394, step_cpml PML.py:138-190) for one cell: psi update and
 * the corrections scattered in slab order. which 0: E half-step (curl H with the E coefficients),
 * 1: H half-step. d holds the six one-sided differences in the order
 * (dx F_z, dx F_y, dy F_x, dy F_z, dz F_y, dz F_x), i.e. (d1, d2) of axis a at d[2a], d[2a+1]. */
static inline void FN(oc_cpml)(const oc_scene* s, R** psi, int which, int i, int j, int k, const R* d, R* K) {
  for (int p = 0; p < s->nslab; ++p) {
    const oc_slab* sl = &s->slabs[p];
    long o = FN(oc_slab_off)(sl, s->nx, s->ny, s->nz, i, j, k);
    if (o < 0) continue;
    int pos = (sl->axis == 0 ? i : sl->axis == 1 ? j : k) - sl->lo;
    const R* b = (const R*)(which == 0 ? sl->bE : sl->bH);
    const R* a = (const R*)(which == 0 ? sl->aE : sl->aH);
    const R* ik = (const R*)(which == 0 ? sl->ikE : sl->ikH);
    R* p1 = psi[4 * p + 2 * which];
    R* p2 = psi[4 * p + 2 * which + 1];
    R d1 = d[2 * sl->axis], d2 = d[2 * sl->axis + 1];
    R p1n = b[pos] * p1[o] + a[pos] * d1;
    R p2n = b[pos] * p2[o] + a[pos] * d2;
    p1[o] = p1n;
    p2[o] = p2n;
    R c1 = p1n, c2 = p2n;
    if (!sl->kappa_one) {
      R km = (R)(ik[pos] - (R)1);
      c1 = km * d1 + p1n;
      c2 = km * d2 + p2n;
    }
    int a1 = (sl->axis + 1) % 3, a2 = (sl->axis + 2) % 3;
    K[a1] = K[a1] + (-c1);
    K[a2] = K[a2] + c2;
  }
}

static void FN(oc_sources)(const oc_scene* s, R* F, int field, int n, size_t N, size_t sx, size_t sy) {
  const R c = (R)s->courant;
  for (int q = 0; q < s->nsrc; ++q) {
    const oc_source* src = &s->src[q];
    if (src->field != field) continue;
    if (src->kind == OC_DIPOLE) {
      size_t o = src->comp * N + src->cell[0] * sx + src->cell[1] * sy + src->cell[2];
      F[o] = F[o] + ((const R*)src->table)[n];
      continue;
    }
    /* the plane: C-order over the two other axes */
    int ax = src->axis;
    int n1 = ax == 0 ? s->ny : s->nx, n2 = ax == 2 ? s->ny : s->nz;
    size_t plane = (size_t)n1 * n2;
    const R* prof = (const R*)src->profile;
    const R* amp = (const R*)src->amp;
    for (int t = 0; t < src->nterms; ++t) {
      int comp = src->tcomp[t], sc = src->tsrc[t];
      R a_n = amp[(size_t)sc * s->T + n];
      R sg = (R)(src->sign * src->tsign[t]);
      const R* mult = (const R*)src->mult[comp];
      R ms = (R)src->mult_scalar;
      for (int u = 0; u < n1; ++u)
        for (int v = 0; v < n2; ++v) {
          size_t pidx = (size_t)u * n2 + v;
          int i = ax == 0 ? src->index : u;
          int j = ax == 1 ? src->index : (ax == 0 ? u : v);
          int k = ax == 2 ? src->index : v;
          size_t o = comp * N + i * sx + j * sy + k;
          R term = ((prof[sc * plane + pidx] * a_n) * c) * (mult ? mult[pidx] : ms);
          F[o] = F[o] + sg * term;
        }
    }
  }
}

int FN(oc_run)(const oc_scene* s) {
  const int nx = s->nx, ny = s->ny, nz = s->nz;
  const size_t sy = nz, sx = (size_t)ny * nz, N = (size_t)nx * ny * nz;
  R* E = (R*)s->E;
  R* H = (R*)s->H;
  const R* ie = (const R*)s->inv_eps;
  const R c = (R)s->courant;
  if (s->nthreads > 0) omp_set_num_threads(s->nthreads);
  /* psi: per slab [psiE1, psiE2, psiH1, psiH2] */
  R** psi = (R**)calloc(4 * (size_t)(s->nslab > 0 ? s->nslab : 1), sizeof(R*));
  for (int p = 0; p < s->nslab; ++p) {
    const oc_slab* sl = &s->slabs[p];
    size_t L = sl->hi - sl->lo;
    size_t cells = L * (sl->axis == 0 ? sx : sl->axis == 1 ? (size_t)nx * nz : (size_t)nx * ny);
    for (int q = 0; q < 4; ++q) psi[4 * p + q] = (R*)calloc(cells, sizeof(R));
  }
  /* ADE state per box: P^n and P^{n-1}, (P, 3, box) */
  R** Pc = (R**)calloc(s->nade > 0 ? s->nade : 1, sizeof(R*));
  R** Pp = (R**)calloc(s->nade > 0 ? s->nade : 1, sizeof(R*));
  for (int b = 0; b < s->nade; ++b) {
    const oc_ade* a = &s->ade[b];
    size_t cells = (size_t)(a->hi[0] - a->lo[0]) * (a->hi[1] - a->lo[1]) * (a->hi[2] - a->lo[2]);
    Pc[b] = (R*)calloc((size_t)a->P * 3 * cells, sizeof(R));
    Pp[b] = (R*)calloc((size_t)a->P * 3 * cells, sizeof(R));
  }
  R* Hp = (R*)malloc(3 * N * sizeof(R));
  FN(oc_ctx) ctx = {nx, ny, nz, N, sx, sy, E, Hp, H};

  for (int n = 0; n < s->T; ++n) {
    int any_det = 0;
    for (int d = 0; d < s->ndet; ++d) any_det |= s->det[d].time_to_idx[n] >= 0;
    if (any_det) memcpy(Hp, H, 3 * N * sizeof(R));

    /* ---- E half-step (update.py:256-354) */
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 0; i < nx; ++i)
      for (int j = 0; j < ny; ++j) {
        for (int k = 0; k < nz; ++k) {
          size_t o = i * sx + j * sy + k;
          R Hx = H[o], Hy = H[N + o], Hz = H[2 * N + o];
          R Hz_jm = j > 0 ? H[2 * N + o - sy] : (R)0, Hy_km = k > 0 ? H[N + o - 1] : (R)0;
          R Hx_km = k > 0 ? H[o - 1] : (R)0, Hz_im = i > 0 ? H[2 * N + o - sx] : (R)0;
          R Hy_im = i > 0 ? H[N + o - sx] : (R)0, Hx_jm = j > 0 ? H[o - sy] : (R)0;
          R dyHz = Hz - Hz_jm, dzHy = Hy - Hy_km, dzHx = Hx - Hx_km;
          R dxHz = Hz - Hz_im, dxHy = Hy - Hy_im, dyHx = Hx - Hx_jm;
          R K[3] = {dyHz - dzHy, dzHx - dxHz, dxHy - dyHx};
          if (s->nslab) {
            R dd[6] = {dxHz, dxHy, dyHx, dyHz, dzHy, dzHx};
            FN(oc_cpml)(s, psi, 0, i, j, k, dd, K);
          }
          R Eo[3], En[3], iev[3];
          for (int a = 0; a < 3; ++a) {
            Eo[a] = E[a * N + o];
            iev[a] = ie[(s->eps_comps == 3 ? a : 0) * N + o];
            En[a] = Eo[a] + (c * K[a]) * iev[a];
          }
          for (int b = 0; b < s->nade; ++b) {
            const oc_ade* A = &s->ade[b];
            if (i < A->lo[0] || i >= A->hi[0] || j < A->lo[1] || j >= A->hi[1] || k < A->lo[2] || k >= A->hi[2])
              continue;
            size_t bx = A->hi[0] - A->lo[0], by = A->hi[1] - A->lo[1], bz = A->hi[2] - A->lo[2];
            size_t cells = bx * by * bz;
            size_t lo_ = ((size_t)(i - A->lo[0]) * by + (j - A->lo[1])) * bz + (k - A->lo[2]);
            const R *c1 = (const R*)A->c1, *c2 = (const R*)A->c2, *c3 = (const R*)A->c3;
            for (int a = 0; a < 3; ++a) {
              R delta = (R)0;
              for (int p = 0; p < A->P; ++p) {
                size_t co = ((size_t)p * A->C + (A->C == 3 ? a : 0)) * cells + lo_;
                size_t po = ((size_t)p * 3 + a) * cells + lo_;
                R pc = Pc[b][po], pp = Pp[b][po];
                R ph = (c1[co] * pc + c2[co] * pp) + c3[co] * Eo[a];
                R dp = pc - ph;
                delta = p == 0 ? dp : delta + dp;
                Pp[b][po] = pc;
                Pc[b][po] = ph;
              }
              En[a] = En[a] + iev[a] * delta;
            }
          }
          for (int a = 0; a < 3; ++a) E[a * N + o] = En[a];
        }
      }
    FN(oc_sources)(s, E, 0, n, N, sx, sy);

    /* ---- H half-step (update.py:689-750) */
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 0; i < nx; ++i)
      for (int j = 0; j < ny; ++j) {
        for (int k = 0; k < nz; ++k) {
          size_t o = i * sx + j * sy + k;
          R Ex = E[o], Ey = E[N + o], Ez = E[2 * N + o];
          R Ez_jp = j < ny - 1 ? E[2 * N + o + sy] : (R)0, Ey_kp = k < nz - 1 ? E[N + o + 1] : (R)0;
          R Ex_kp = k < nz - 1 ? E[o + 1] : (R)0, Ez_ip = i < nx - 1 ? E[2 * N + o + sx] : (R)0;
          R Ey_ip = i < nx - 1 ? E[N + o + sx] : (R)0, Ex_jp = j < ny - 1 ? E[o + sy] : (R)0;
          R dyEz = Ez_jp - Ez, dzEy = Ey_kp - Ey, dzEx = Ex_kp - Ex;
          R dxEz = Ez_ip - Ez, dxEy = Ey_ip - Ey, dyEx = Ex_jp - Ex;
          R K[3] = {dyEz - dzEy, dzEx - dxEz, dxEy - dyEx};
          if (s->nslab) {
            R dd[6] = {dxEz, dxEy, dyEx, dyEz, dzEy, dzEx};
            FN(oc_cpml)(s, psi, 1, i, j, k, dd, K);
          }
          for (int a = 0; a < 3; ++a) H[a * N + o] = H[a * N + o] - (c * K[a]) * (R)1;
        }
      }
    FN(oc_sources)(s, H, 1, n, N, sx, sy);

    /* ---- detectors (update.py:1040-1137) */
    for (int d = 0; d < s->ndet; ++d) {
      const oc_det* D = &s->det[d];
      int idx = D->time_to_idx[n];
      if (idx < 0) continue;
      int bx = D->hi[0] - D->lo[0], by = D->hi[1] - D->lo[1], bz = D->hi[2] - D->lo[2];
      size_t box = (size_t)bx * by * bz;
      if (D->kind == OC_ENERGY || D->kind == OC_FLUX) {
        double acc = 0.0;
        const R w = (R)D->weight;
#pragma omp parallel for collapse(2) reduction(+ : acc) schedule(static)
        for (int u = 0; u < bx; ++u)
          for (int v = 0; v < by; ++v)
            for (int q = 0; q < bz; ++q) {
              int i = D->lo[0] + u, j = D->lo[1] + v, k = D->lo[2] + q;
              R f6[6];
              FN(oc_interp)(&ctx, i, j, k, f6);
              R val;
              if (D->kind == OC_ENERGY) {
                size_t o = i * sx + j * sy + k;
                R t[3];
                for (int a = 0; a < 3; ++a) t[a] = ((R)0.5 * ((R)1 / ie[(s->eps_comps == 3 ? a : 0) * N + o])) * (f6[a] * f6[a]);
                R eE = (t[0] + t[1]) + t[2];
                R hh[3];
                for (int a = 0; a < 3; ++a) hh[a] = (R)0.5 * (f6[3 + a] * f6[3 + a]);
                R eH = (hh[0] + hh[1]) + hh[2];
                val = (eE + eH) * w;
              } else {
                int b1 = (D->axis + 1) % 3, b2 = (D->axis + 2) % 3;
                R pf = f6[b1] * f6[3 + b2] - f6[b2] * f6[3 + b1];
                if (D->sign < 0) pf = -pf;
                val = pf * w;
              }
              acc += (double)val;
            }
        ((R*)D->state)[idx] = (R)acc;
      } else if (D->kind == OC_FIELD) {
        R* st = (R*)D->state + (size_t)idx * D->ncomp * box;
#pragma omp parallel for collapse(2) schedule(static)
        for (int u = 0; u < bx; ++u)
          for (int v = 0; v < by; ++v)
            for (int q = 0; q < bz; ++q) {
              R f6[6];
              FN(oc_interp)(&ctx, D->lo[0] + u, D->lo[1] + v, D->lo[2] + q, f6);
              size_t cell = ((size_t)u * by + v) * bz + q;
              for (int ci = 0; ci < D->ncomp; ++ci) st[ci * box + cell] = f6[D->comps[ci]];
            }
      } else if (D->kind == OC_PHASOR) {
        R* st = (R*)D->state; /* (1, nfreq, ncomp, box) complex: (re, im) pairs */
        const R* ct = (const R*)D->cos_t;
        const R* snt = (const R*)D->sin_t;
        const R scale = (R)D->scale;
        const R win = ((const R*)D->window)[n];
#pragma omp parallel for collapse(2) schedule(static)
        for (int u = 0; u < bx; ++u)
          for (int v = 0; v < by; ++v)
            for (int q = 0; q < bz; ++q) {
              R f6[6];
              FN(oc_interp)(&ctx, D->lo[0] + u, D->lo[1] + v, D->lo[2] + q, f6);
              size_t cell = ((size_t)u * by + v) * bz + q;
              for (int f = 0; f < D->nfreq; ++f) {
                R cs = ct[(size_t)f * s->T + n], sn = snt[(size_t)f * s->T + n];
                for (int ci = 0; ci < D->ncomp; ++ci) {
                  R val = f6[D->comps[ci]];
                  size_t o2 = 2 * (((size_t)f * D->ncomp + ci) * box + cell);
                  st[o2] = st[o2] + ((val * cs) * scale) * win;
                  st[o2 + 1] = st[o2 + 1] + ((val * sn) * scale) * win;
                }
              }
            }
      }
    }
  }
  for (int p = 0; p < 4 * s->nslab; ++p) free(psi[p]);
  free(psi);
  for (int b = 0; b < s->nade; ++b) {
    free(Pc[b]);
    free(Pp[b]);
  }
  free(Pc);
  free(Pp);
  free(Hp);
  return 0;
}

#undef FN
#undef OC_CAT
#undef OC_CAT2
#undef R
