/*
 * gitm_horiz_oracle.c -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of GITM's horizontal step for the parity tests of the HIP
 * implementation (advance_horizontal(iBlock), advance_horizontal.f90:58-924,
 * with calc_facevalues_lats/lons, :928-1034).  Nothing in the product library
 * links, loads or calls it.
 *
 * Each altitude 1..nAlts of each block is a 2-D problem (independent of the
 * others unless UseTopography, which reads the levels above and below): the
 * state of the slice (halo cells -1:0 and n+1:n+2 included) is advanced by
 * the reference's RK4 (:154-345) of horizontal_solver (:535-818), the halo
 * cells keeping their input values through the four stages.  The arithmetic
 * keeps the reference's operation order (Fortran evaluates left to right
 * within parentheses; no contraction), so a faithful GPU build matches bit
 * for bit.
 *
 * Layout ("block layout", as GITM's arrays with C index order): a state
 * field is [block][component][level][lat][lon] with level -1..nAlts+2,
 * lat -1..nLats+2 and lon -1..nLons+2 at offsets 0..; cMax_GDB is
 * [block][3][0..nAlts+1][0..nLats+1][0..nLons+1]; latitude arrays are
 * [block][-1..nLats+2].
 */
#include "gitm_vert_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

enum { V_RHO = 0, V_TEMP = 1, V_VEL = 2, V_NUM = 5, V_VV = 11, V_INUM = 17, V_MAX = 19 };
#define HMAX 64 /* nLons, nLats <= 60 here */

typedef struct {
  int nlo, nla, nA;
  long long slo, sla, slev; /* strides of a state field: lon 1, lat, level */
} geo;

static long long at3(const geo *g, int lon, int lat, int lev) {
  return (long long)(lon + 1) + (long long)(lat + 1) * g->sla + (long long)(lev + 1) * g->slev;
}

/* calc_facevalues_lons/lats (advance_horizontal.f90:928-1034) on one line:
 * v[i+2] = Var(i) for i = -2..n+2 (only -1..n+2 read), invfb[i+1] =
 * InvD*Dist_FB(i) for i = 0..n+2, dfb[i+1] = d*Dist_FB(i) for i = 1..n+1.
 * left[i], right[i] for faces i = 1..n+1. */
static void faces(int n, const double *v, const double *invfb, const double *dfb,
                  double beta, double *left, double *right) {
  const double F1 = 0.6250000, F2 = 0.0416667;
  double dl[HMAX + 4];
#define V(i) v[(i) + 2]
#define IF(i) invfb[(i) + 1]
  for (int i = 0; i <= n + 1; i++) {
    double up, down, h;
    if (i <= n) {
      h = IF(i + 1) * 2;
      up = h * (F1 * (V(i + 1) - V(i)) - F2 * (V(i + 2) - V(i - 1)));
    } else {
      up = (V(i + 1) - V(i)) * IF(i + 1);
    }
    if (i >= 1) {
      h = IF(i) * 2;
      down = h * (F1 * (V(i) - V(i - 1)) - F2 * (V(i + 1) - V(i - 2)));
    } else {
      down = (V(i) - V(i - 1)) * IF(i);
    }
    dl[i] = ov_limiter_mc(beta, up, down);
  }
  for (int i = 1; i <= n + 1; i++) {
    left[i] = V(i - 1) + 0.5 * dl[i - 1] * dfb[i + 1];
    right[i] = V(i) - 0.5 * dl[i] * dfb[i + 1];
  }
#undef V
#undef IF
}

/* calc_rusanov_lons/lats (:821-922): gradient from the averaged face values
 * and the Rusanov flux difference on cells 1..n; cm[i] = cMax(i), i =
 * 0..n+1; invgb[i] = InvD*Dist_GB(i), i = 1..n. */
static void rusanov(int n, const double *v, const double *invfb, const double *dfb,
                    const double *invgb, const double *cm, double beta, double *grad,
                    double *diff) {
  double L[HMAX + 4], R[HMAX + 4], flux[HMAX + 4];
  faces(n, v, invfb, dfb, beta, L, R);
  for (int f = 1; f <= n + 1; f++)
    flux[f] = 0.5 * fmax(cm[f - 1], cm[f]) * (R[f] - L[f]);
  for (int i = 1; i <= n; i++) {
    grad[i] = 0.5 * (L[i + 1] + R[i + 1] - L[i] - R[i]) * invgb[i];
    diff[i] = (flux[i + 1] - flux[i]) * invgb[i];
  }
}

typedef struct {
  double glon, dlon, glat, dlat;
} gd;

/* One altitude of one block. */
static int slice(const oh_grid *p, const geo *g, int b, int k, ov_status *st) {
  const int nlo = g->nlo, nla = g->nla, NLo = nlo + 4, NLg = nla + 4;
  const int nv = p->use_ion_advection ? V_MAX : V_INUM;
  const long long fsz = (long long)(g->nA + 4) * g->slev; /* one component */
  double *x = malloc(sizeof(double) * (size_t)V_MAX * NLo * NLg);
  double *orig = malloc(sizeof(double) * (size_t)V_MAX * nlo * nla);
  double *acc = malloc(sizeof(double) * (size_t)V_MAX * nlo * nla);
  double *nw = malloc(sizeof(double) * (size_t)V_MAX * nlo * nla);
  gd *d = malloc(sizeof(gd) * (size_t)V_MAX * nlo * nla);
  int rc = OV_OK;
#define XS(vv, la, lo) x[((size_t)(vv) * NLg + (la) + 1) * NLo + (lo) + 1]
#define IN(vv, lo, la) ((size_t)(vv) * nlo * nla + (size_t)((la) - 1) * nlo + (lo) - 1)
  /* field pointer of variable vv in block b */
  const double *src[V_MAX];
  src[V_RHO] = p->rho + (long long)b * fsz;
  src[V_TEMP] = p->temp + (long long)b * fsz;
  for (int c = 0; c < 3; c++) src[V_VEL + c] = p->vel + ((long long)b * 3 + c) * fsz;
  for (int s = 0; s < OV_NSPECIES; s++) {
    src[V_NUM + s] = p->nds + ((long long)b * OV_NSPECIES + s) * fsz;
    src[V_VV + s] = p->vvel + ((long long)b * OV_NSPECIES + s) * fsz;
  }
  if (p->use_ion_advection)
    for (int s = 0; s < OV_NIONS_ADVECT; s++)
      src[V_INUM + s] = p->ions + ((long long)b * OV_NIONS + s) * fsz;
  for (int vv = 0; vv < nv; vv++)
    for (int la = -1; la <= nla + 2; la++)
      for (int lo = -1; lo <= nlo + 2; lo++) XS(vv, la, lo) = src[vv][at3(g, lo, la, k)];
  for (int vv = 0; vv < nv; vv++)
    for (int la = 1; la <= nla; la++)
      for (int lo = 1; lo <= nlo; lo++) orig[IN(vv, lo, la)] = XS(vv, la, lo);

  const double *lat = p->latitude + (long long)b * NLg, *tanl = p->tan_lat + (long long)b * NLg,
               *cosl = p->cos_lat + (long long)b * NLg;
  const long long bo = (long long)b * fsz;
  const long long cmo = (long long)(g->nA + 2) * (nla + 2) * (nlo + 2);
  const double *cme = p->cmax + ((long long)b * 3 + 0) * cmo; /* iEast_  */
  const double *cmn = p->cmax + ((long long)b * 3 + 1) * cmo; /* iNorth_ */
#define CM(a, lo, la) a[((long long)k * (nla + 2) + (la)) * (nlo + 2) + (lo)]
#define GF(f, lo, la) p->f[bo + at3(g, lo, la, k)]
  /* UseTopography: each variable's altitude derivative dVarDAlt_C (:581-586
   * and before each variable's gradients, :591-662), fixed through the RK4
   * and taken from the state arrays, where level k+1 still holds its input
   * and level k-1 already the step's result */
  double *dz = NULL;
  if (p->use_topography) {
    dz = malloc(sizeof(double) * (size_t)V_MAX * nlo * nla);
    for (int vv = 0; vv < nv; vv++)
      for (int la = 1; la <= nla; la++)
        for (int lo = 1; lo <= nlo; lo++) {
          const double h = 0.5 / GF(dalt, lo, la);
          dz[IN(vv, lo, la)] =
              h * (src[vv][at3(g, lo, la, k + 1)] - src[vv][at3(g, lo, la, k - 1)]);
        }
  }

  for (int stage = 0; stage < 4; stage++) {
    /* horizontal_solver: Grad and Diff of every variable (:588-662) */
    for (int vv = 0; vv < nv; vv++) {
      double line[HMAX + 6], ifb[HMAX + 4], dfb[HMAX + 4], igb[HMAX + 4], cm[HMAX + 4],
          gr[HMAX + 4], df[HMAX + 4];
      for (int la = 1; la <= nla; la++) { /* longitude lines */
        for (int lo = -1; lo <= nlo + 2; lo++) line[lo + 2] = XS(vv, la, lo);
        for (int lo = 0; lo <= nlo + 2; lo++) ifb[lo + 1] = GF(inv_dlon_fb, lo, la);
        for (int lo = 1; lo <= nlo + 1; lo++) dfb[lo + 1] = GF(dlon_fb, lo, la);
        for (int lo = 1; lo <= nlo; lo++) igb[lo] = GF(inv_dlon_gb, lo, la);
        for (int lo = 0; lo <= nlo + 1; lo++) cm[lo] = CM(cme, lo, la);
        rusanov(nlo, line, ifb, dfb, igb, cm, p->beta_limiter, gr, df);
        for (int lo = 1; lo <= nlo; lo++) {
          if (dz) gr[lo] = gr[lo] - dz[IN(vv, lo, la)] * GF(dalt_dlon, lo, la); /* :919-920 */
          d[IN(vv, lo, la)].glon = gr[lo];
          d[IN(vv, lo, la)].dlon = df[lo];
        }
      }
      for (int lo = 1; lo <= nlo; lo++) { /* latitude lines */
        for (int la = -1; la <= nla + 2; la++) line[la + 2] = XS(vv, la, lo);
        for (int la = 0; la <= nla + 2; la++) ifb[la + 1] = GF(inv_dlat_fb, lo, la);
        for (int la = 1; la <= nla + 1; la++) dfb[la + 1] = GF(dlat_fb, lo, la);
        for (int la = 1; la <= nla; la++) igb[la] = GF(inv_dlat_gb, lo, la);
        for (int la = 0; la <= nla + 1; la++) cm[la] = CM(cmn, lo, la);
        rusanov(nla, line, ifb, dfb, igb, cm, p->beta_limiter, gr, df);
        for (int la = 1; la <= nla; la++) {
          if (dz) gr[la] = gr[la] - dz[IN(vv, lo, la)] * GF(dalt_dlat, lo, la); /* :870-871 */
          d[IN(vv, lo, la)].glat = gr[la];
          d[IN(vv, lo, la)].dlat = df[la];
        }
      }
    }
    /* the updates (:667-816), New starting from the stage's state */
    const double dt = p->dt, om = p->omega_body;
    for (int la = 1; la <= nla; la++) {
      const double sinl = sin(lat[la + 1]), cosv = cosl[la + 1], tl = tanl[la + 1];
      const double csin = sinl * 2 * om, ccos = cosv * 2 * om;
      const double cent = om * om * cosv * sinl;
      for (int lo = 1; lo <= nlo; lo++) {
        const double invr = GF(invr, lo, la), rd = GF(raddist, lo, la);
        const double gam = GF(gamma, lo, la);
        const double rho = XS(V_RHO, la, lo), tmp = XS(V_TEMP, la, lo);
        const double ve = XS(V_VEL + 0, la, lo), vn = XS(V_VEL + 1, la, lo),
                     vu = XS(V_VEL + 2, la, lo);
#define D(vv) d[IN(vv, lo, la)]
#define NW(vv) nw[IN(vv, lo, la)]
        for (int vv = 0; vv < nv; vv++) NW(vv) = XS(vv, la, lo);
        const double div = D(V_VEL + 1).glat + D(V_VEL + 0).glon - tl * vn * invr;
        NW(V_RHO) = NW(V_RHO) - dt * (rho * div + D(V_RHO).glat * vn + D(V_RHO).glon * ve) +
                    dt * (D(V_RHO).dlon + D(V_RHO).dlat);
        for (int s = 0; s < OV_NSPECIES; s++) {
          const int vv = V_NUM + s;
          NW(vv) = NW(vv) - dt * (XS(vv, la, lo) * div + D(vv).glat * vn + D(vv).glon * ve) +
                   dt * (D(vv).dlon + D(vv).dlat);
        }
        if (p->use_ion_advection) {
          const long long oi = ((long long)b * 3) * fsz + at3(g, lo, la, k);
          const double ive = p->ivel[oi], ivn = p->ivel[oi + fsz]; /* IVelocity E, N */
          for (int s = 0; s < OV_NIONS_ADVECT; s++) {
            const int vv = V_INUM + s;
            NW(vv) = NW(vv) - dt * (+D(vv).glat * ivn + D(vv).glon * ive) +
                     dt * (D(vv).dlon + D(vv).dlat);
          }
        }
        NW(V_VEL + 0) = NW(V_VEL + 0) -
                        dt * (vn * D(V_VEL + 0).glat + ve * D(V_VEL + 0).glon +
                              ve * (vu - tl * vn) * invr + D(V_TEMP).glon +
                              D(V_RHO).glon * tmp / rho) +
                        dt * (D(V_VEL + 0).dlon + D(V_VEL + 0).dlat);
        NW(V_VEL + 1) = NW(V_VEL + 1) -
                        dt * (vn * D(V_VEL + 1).glat + ve * D(V_VEL + 1).glon +
                              (vn * vu + tl * (ve * ve)) * invr + D(V_TEMP).glat +
                              D(V_RHO).glat * tmp / rho) +
                        dt * (D(V_VEL + 1).dlon + D(V_VEL + 1).dlat);
        NW(V_VEL + 2) = NW(V_VEL + 2) - dt * (vn * D(V_VEL + 2).glat + ve * D(V_VEL + 2).glon) +
                        dt * (D(V_VEL + 2).dlon + D(V_VEL + 2).dlat);
        for (int s = 0; s < OV_NSPECIES; s++) {
          const int vv = V_VV + s;
          NW(vv) = NW(vv) - dt * (vn * D(vv).glat + ve * D(vv).glon) +
                   dt * (D(vv).dlat + D(vv).dlon);
        }
        if (p->use_coriolis) {
          NW(V_VEL + 0) = NW(V_VEL + 0) + dt * (+csin * vn - ccos * vu);
          NW(V_VEL + 1) = NW(V_VEL + 1) - dt * (cent * rd + csin * ve);
        }
        NW(V_TEMP) = NW(V_TEMP) -
                     dt * ((gam - 1) * tmp * div + D(V_TEMP).glat * vn + D(V_TEMP).glon * ve) +
                     dt * (D(V_TEMP).dlon + D(V_TEMP).dlat);
#undef D
      }
    }
    /* the RK4 bookkeeping (:184-345) */
    for (int vv = 0; vv < nv; vv++)
      for (int la = 1; la <= nla; la++)
        for (int lo = 1; lo <= nlo; lo++) {
          const size_t c = IN(vv, lo, la);
          const double K = nw[c] - XS(vv, la, lo);
          double u;
          if (stage == 0) {
            acc[c] = K;
            u = 0.5 * K + orig[c];
          } else if (stage == 1) {
            acc[c] = acc[c] + 2.0 * K;
            u = 0.5 * K + orig[c];
          } else if (stage == 2) {
            acc[c] = acc[c] + 2.0 * K;
            u = K + orig[c];
          } else {
            acc[c] = acc[c] + K;
            u = orig[c] + (1.0 / 6.0) * acc[c];
          }
          XS(vv, la, lo) = u;
        }
  }
  /* write back (:349-429) */
  for (int la = 1; la <= nla && rc == OV_OK; la++)
    for (int lo = 1; lo <= nlo; lo++)
      for (int s = 0; s < OV_NSPECIES; s++)
        if (XS(V_NUM + s, la, lo) < 0.0 && rc == OV_OK) {
          rc = OV_ERR_NONPOSITIVE_DENSITY;
          if (st) {
            st->code = rc;
            st->col = (b * nla + (la - 1)) * nlo + (lo - 1);
            st->level = k;
          }
        }
  for (int la = 1; la <= nla; la++)
    for (int lo = 1; lo <= nlo; lo++) {
      const long long c3 = at3(g, lo, la, k), o = bo + c3;
      p->rho[o] = XS(V_RHO, la, lo);
      for (int c = 0; c < 3; c++) p->vel[((long long)b * 3 + c) * fsz + c3] = XS(V_VEL + c, la, lo);
      if (p->temp_source) p->temp_source[o] = XS(V_TEMP, la, lo) - p->temp[o];
      p->temp[o] = XS(V_TEMP, la, lo);
      for (int s = 0; s < OV_NSPECIES; s++) {
        const long long os = ((long long)b * OV_NSPECIES + s) * fsz + c3;
        p->vvel[os] = XS(V_VV + s, la, lo);
        p->nds[os] = XS(V_NUM + s, la, lo);
      }
      if (p->use_ion_advection) {
        const long long oi = (long long)b * OV_NIONS * fsz + c3;
        for (int s = 0; s < OV_NIONS_ADVECT; s++) {
          double v = XS(V_INUM + s, la, lo);
          if (v < 1.0e2) v = 1.0e2; /* :405-417 */
          p->ions[oi + (long long)s * fsz] = v;
        }
        double e = 0.0; /* :420-425 */
        for (int s = 0; s < OV_NIONS - 1; s++) e = e + p->ions[oi + (long long)s * fsz];
        p->ions[oi + (long long)(OV_NIONS - 1) * fsz] = e;
      }
    }
#undef XS
#undef IN
#undef CM
#undef GF
#undef NW
  free(x);
  free(orig);
  free(acc);
  free(nw);
  free(d);
  free(dz);
  return rc;
}

/* UseImprovedIonAdvection, after every altitude of block b
 * (advance_horizontal.f90:432-518): the divergence of the field-aligned ion
 * velocity from UAM_Gradient (gradient.f90:1-86), the ions divided by
 * 1 + Dt div, then the electrons re-summed on every cell and level. */
static void ion_divergence(const oh_grid *p, const geo *g, int b) {
  const int nlo = g->nlo, nla = g->nla, nA = g->nA, NLg = nla + 4;
  const long long fsz = (long long)(nA + 4) * g->slev;
  const double *vi = p->ivel_par + (long long)b * 3 * fsz;
  const double *inv = p->invr + (long long)b * fsz, *rd = p->raddist + (long long)b * fsz;
  const double *gl = p->grad_lon + (long long)b * 3 * nlo, *ga = p->grad_lat + (long long)b * 3 * nla;
  const double *cosl = p->cos_lat + (long long)b * NLg, *tanl = p->tan_lat + (long long)b * NLg;
  double *ions = p->ions + (long long)b * OV_NIONS * fsz;
  for (int k = 1; k <= nA; k++)
    for (int la = 1; la <= nla; la++)
      for (int lo = 1; lo <= nlo; lo++) {
        const long long c = at3(g, lo, la, k);
        const double ir = inv[c];
        double div = 0.0;
        /* UAM_Gradient leaves a direction of one cell at 0 (:24-26) */
        if (nlo > 1) { /* iDir = 1: east (gradient.f90:27-41) */
          const double *v = vi;
          const double icl = 1. / fmax(cosl[la + 1], 0.17);
          const double gr = (gl[lo - 1] * v[at3(g, lo - 1, la, k)] +
                             gl[nlo + lo - 1] * v[c] +
                             gl[2 * nlo + lo - 1] * v[at3(g, lo + 1, la, k)]) *
                            ir * icl;
          div = div + gr;
        }
        { /* iDir = 2: north (:44-56), then the tan(latitude) term (:462-477) */
          const double *v = vi + fsz;
          const double gr = nla > 1 ? (ga[la - 1] * v[at3(g, lo, la - 1, k)] +
                                       ga[nla + la - 1] * v[c] +
                                       ga[2 * nla + la - 1] * v[at3(g, lo, la + 1, k)]) *
                                          ir
                                    : 0.0;
          div = div + gr;
          double tl = tanl[la + 1];
          if (tl > 5.0) tl = 5.0;
          if (tl < -5.0) tl = -5.0;
          div = div - tl * v[c] * ir;
        }
        { /* iDir = 3: radial (:59-78), then 2 w / r (:480-493) */
          const double *v = vi + 2 * fsz;
          const double drp = rd[at3(g, lo, la, k + 1)] - rd[c];
          const double drm = rd[c] - rd[at3(g, lo, la, k - 1)];
          const double q = drm / drp;
          const double q2 = q * q;
          const double bottom = drm * (1 + q);
          const double gr = nA > 1 ? (q2 * v[at3(g, lo, la, k + 1)] -
                                      v[at3(g, lo, la, k - 1)] - (q2 - 1) * v[c]) /
                                         bottom
                                   : 0.0;
          div = div + gr;
          div = div + 2 * v[c] * ir;
        }
        div = div * cosl[la + 1]; /* :498-502 */
        const double change = p->dt * div;
        for (int s = 0; s < OV_NIONS_ADVECT; s++) { /* :505-510 */
          double *n = ions + (long long)s * fsz + c;
          *n = *n / (1 + change);
        }
      }
  /* electrons on every cell and level (:512-516) */
  for (long long c = 0; c < fsz; c++) {
    double e = 0.0;
    for (int s = 0; s < OV_NIONS - 1; s++) e = e + ions[(long long)s * fsz + c];
    ions[(long long)(OV_NIONS - 1) * fsz + c] = e;
  }
}

int oh_advance_horizontal(const oh_grid *p, ov_status *st) {
  if (!p || p->n_lons < 1 || p->n_lats < 1 || p->n_lons > 60 || p->n_lats > 60 ||
      p->n_alts < 1 || p->n_blocks < 1)
    return OV_ERR_BAD_ARGS;
  if (p->use_ion_advection && (!p->ions || !p->ivel)) return OV_ERR_BAD_ARGS;
  const int improved = p->use_ion_advection && p->use_improved_ion_advection;
  if (improved && (!p->ivel_par || !p->grad_lon || !p->grad_lat)) return OV_ERR_BAD_ARGS;
  if (p->use_topography && (!p->dalt || !p->dalt_dlon || !p->dalt_dlat)) return OV_ERR_BAD_ARGS;
  geo g;
  g.nlo = p->n_lons;
  g.nla = p->n_lats;
  g.nA = p->n_alts;
  g.slo = 1;
  g.sla = p->n_lons + 4;
  g.slev = (long long)(p->n_lons + 4) * (p->n_lats + 4);
  if (st) {
    st->code = OV_OK;
    st->col = -1;
    st->level = -99;
  }
  int first = OV_OK;
  for (int b = 0; b < p->n_blocks; b++) {
    for (int k = 1; k <= p->n_alts; k++) {
      ov_status s1;
      const int rc = slice(p, &g, b, k, &s1);
      if (rc != OV_OK && first == OV_OK) {
        first = rc;
        if (st) *st = s1;
      }
    }
    if (improved) ion_divergence(p, &g, b);
  }
  return first;
}

/* ---------------------------------------------------------------------------
 * The message exchange on one node (exchange_messages_sphere.f90:4-95).
 * ------------------------------------------------------------------------ */
int oh_exchange(const oh_exchange_args *a, ov_status *st) {
  const int nlo = a->n_lons, nla = a->n_lats, nA = a->n_alts;
  const int nbl = a->n_blocks_lon, nbt = a->n_blocks_lat, nb = nbl * nbt;
  const long long NLON = (long long)nbl * nlo, NLAT = (long long)nbt * nla;
  if (nlo < 1 || nla < 2 || nA < 1 || nbl < 1 || nbt < 1 || NLON % 2) return OV_ERR_BAD_ARGS;
  if (a->use_ion_advection && (!a->ions || !a->ivel)) return OV_ERR_BAD_ARGS;
  geo g;
  g.nlo = nlo;
  g.nla = nla;
  g.nA = nA;
  g.slo = 1;
  g.sla = nlo + 4;
  g.slev = (long long)(nlo + 4) * (nla + 4);
  const long long fsz = (long long)(nA + 4) * g.slev;
  struct { double *p; int nc; } f[9] = {{a->rho, 1}, {a->temp, 1}, {a->itemp, 1}, {a->etemp, 1},
                                        {a->vel, 3}, {a->vvel, OV_NSPECIES},
                                        {a->nds, OV_NSPECIES}, {a->ions, OV_NIONS},
                                        {a->ivel, 3}};
  const int nf = a->use_ion_advection ? 9 : 7;
  for (int b = 0; b < nb; b++) {
    const int bi = b % nbl, bj = b / nbl;
    for (int la = -1; la <= nla + 2; la++)
      for (int lo = -1; lo <= nlo + 2; lo++) {
        if (la >= 1 && la <= nla && lo >= 1 && lo <= nlo) continue; /* interior */
        long long J = (long long)bi * nlo + lo - 1, I = (long long)bj * nla + la - 1;
        if (I < 0) { /* past the south pole */
          I = -1 - I;
          J += NLON / 2;
        } else if (I >= NLAT) { /* past the north pole */
          I = 2 * NLAT - 1 - I;
          J += NLON / 2;
        }
        J = ((J % NLON) + NLON) % NLON;
        const int sb = (int)(I / nla) * nbl + (int)(J / nlo);
        const int slo = (int)(J % nlo) + 1, sla = (int)(I % nla) + 1;
        for (int k = 1; k <= nA; k++)
          for (int q = 0; q < nf; q++)
            for (int c = 0; c < f[q].nc; c++) {
              double *base = f[q].p + (long long)c * fsz;
              base[((long long)b * f[q].nc) * fsz + at3(&g, lo, la, k)] =
                  base[((long long)sb * f[q].nc) * fsz + at3(&g, slo, sla, k)];
            }
      }
  }
  /* the pole halo rows: east and north winds change sign (:30-72) */
  for (int b = 0; b < nb; b++) {
    const double *lat = a->latitude + (long long)b * (nla + 4);
    const double half_pi = 3.14159265358979323846 / 2.0;
    for (int side = 0; side < 2; side++) {
      const int flip = side == 0 ? lat[0 + 1] < -half_pi : lat[nla + 1 + 1] > half_pi;
      if (!flip) continue;
      const int r0 = side == 0 ? -1 : nla + 1;
      double *vv[2] = {a->vel, a->ivel};
      for (int w = 0; w < 2; w++) {
        if (!vv[w]) continue;
        for (int c = 0; c < 2; c++)
          for (int k = -1; k <= nA + 2; k++)
            for (int la = r0; la <= r0 + 1; la++)
              for (int lo = -1; lo <= nlo + 2; lo++) {
                double *x = vv[w] + ((long long)b * 3 + c) * fsz + at3(&g, lo, la, k);
                *x = -*x;
              }
      }
    }
  }
  /* checks on the blocks' cells (:74-84) */
  if (st) {
    st->code = OV_OK;
    st->col = -1;
    st->level = -99;
  }
  for (int b = 0; b < nb; b++)
    for (int pass = 0; pass < 2; pass++) { /* per block: Temperature, then Rho */
      const int code = pass == 0 ? OV_ERR_LOW_TEMPERATURE : OV_ERR_NONPOSITIVE_DENSITY;
      const double *x = (pass == 0 ? a->temp : a->rho) + (long long)b * fsz;
      for (int k = 1; k <= nA; k++)
        for (int la = 1; la <= nla; la++)
          for (int lo = 1; lo <= nlo; lo++) {
            const double v = x[at3(&g, lo, la, k)];
            if (pass == 0 ? v < 100.0 : v < 0.0) {
              if (st) {
                st->code = code;
                st->col = (b * nla + la - 1) * nlo + lo - 1;
                st->level = k;
              }
              return code;
            }
          }
    }
  return OV_OK;
}
