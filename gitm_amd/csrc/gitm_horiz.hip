// gitm_horiz.hip -- the horizontal step (gitm_horiz_advance of
// include/gitm_vert.h): advance_horizontal(iBlock)
// (advance_horizontal.f90:58-924) for every block, each altitude 1..nAlts
// an independent 2-D problem on the block's cells with the halo cells
// (-1:0, n+1:n+2) held at their input values (the caller's message
// exchange fills them, :20).
//
// Mapping: a workgroup takes APW altitudes of one block, a lane one cell of
// one of them.  The workgroup stages the APW slices of every advected
// variable (halo cells included) in LDS; each stage of the RK4 (:154-345)
// then runs horizontal_solver (:535-818) lane by lane: the lane forms the
// MC-limited face values of its own two faces in each direction
// (calc_facevalues_lons/lats, :928-1034) from the slopes at its cell and
// its two neighbours, the gradient and the Rusanov flux difference
// (calc_rusanov_lons/lats, :821-922), and the updates, holding the new
// values in registers until every lane has read the stage's state.  Every
// value is formed by the reference's operations in the reference's order,
// so the results equal the reference's bit for bit (sin(latitude) aside,
// which is the device's).  With UseImprovedIonAdvection a second kernel
// applies the divergence of the ion velocity afterwards.
#include "vert_gpu.h"

namespace {

constexpr int NIA = 2;  // nIonsAdvect (ModEarth.f90:39)
constexpr int V_RHO = 0, V_TEMP = 1, V_VEL = 2, V_NUM = 5, V_VV = 11, V_INUM = 17;
constexpr int NV_NEUTRAL = V_INUM, NV_IONS = V_INUM + NIA;

struct HzParams {
  double dt, omega, beta;
  int use_coriolis;
  int n_lons, n_lats, n_alts, n_blocks;
  int apw;         // altitudes per workgroup
  int alt_groups;  // workgroups per block
  int k_first;     // the launch's first altitude
};

struct HzFields {
  gitm_vert_field inv_dlon_gb, inv_dlon_fb, dlon_fb, inv_dlat_gb, inv_dlat_fb, dlat_fb,
      raddist, latitude, tan_latitude, cmax, temp_source;
  gitm_vert_field dalt, dalt_dlon, dalt_dlat;  // UseTopography
};

// Element (iLon, iLat, level, comp) of block b: the descriptor's ptr is at
// (1, 1, first_level) of block 0.
__device__ __forceinline__ double *hf(const gitm_vert_field &f, int lon, int lat, int lev,
                                      int comp, int b) {
  return f.ptr + (long long)(lon - 1) * f.s_lon + (long long)(lat - 1) * f.s_lat +
         (long long)(lev - f.first_level) * f.s_lev + (long long)comp * f.s_comp +
         (long long)b * f.s_blk;
}

__device__ __forceinline__ double limiter(double beta, double up, double dn) {
  // Limiter_mc (ModLimiter.f90:16-36)
  if (up > 0.0) {
    if (dn > 0.0) return fmin(fmin(beta * up, beta * dn), (up + dn) * 0.5);
    return 0.0;
  }
  if (dn < 0.0) return fmax(fmax(beta * up, beta * dn), (up + dn) * 0.5);
  return 0.0;
}

// The geometry of one cell i (1..n) along one line: InvD*Dist_FB at faces
// i-1..i+2, d*Dist_FB at faces i and i+1, InvD*Dist_GB(i) and cMax at cells
// i-1..i+1.
struct LineGeo {
  double ifb[4], dfb0, dfb1, igb, cm[3];
};

// One cell i (1..n) of one line: v(m) = Var(i + m), m = -3..3 (the line's
// cells -1..n+2 only).  Returns the gradient and the flux difference.
template <class VF>
__device__ __forceinline__ void rusanov_cell(int i, int n, VF v, const LineGeo &g,
                                             double beta, double &grad, double &diff) {
  auto ifb = [&](int m) { return g.ifb[m + 1]; };
  const double dfb0 = g.dfb0, dfb1 = g.dfb1, igb = g.igb, cmm = g.cm[0], cm0 = g.cm[1],
               cmp = g.cm[2];
  const double F1 = 0.6250000, F2 = 0.0416667;
  double dl[3];
#pragma unroll
  for (int t = 0; t < 3; t++) {  // slopes at cells i-1, i, i+1
    const int m = t - 1, c = i + m;
    double up, dn;
    if (c <= n) {
      const double h = ifb(m + 1) * 2;
      up = h * (F1 * (v(m + 1) - v(m)) - F2 * (v(m + 2) - v(m - 1)));
    } else {
      up = (v(m + 1) - v(m)) * ifb(m + 1);
    }
    if (c >= 1) {
      const double h = ifb(m) * 2;
      dn = h * (F1 * (v(m) - v(m - 1)) - F2 * (v(m + 1) - v(m - 2)));
    } else {
      dn = (v(m) - v(m - 1)) * ifb(m);
    }
    dl[t] = limiter(beta, up, dn);
  }
  const double l0 = v(-1) + 0.5 * dl[0] * dfb0, r0 = v(0) - 0.5 * dl[1] * dfb0;
  const double l1 = v(0) + 0.5 * dl[1] * dfb1, r1 = v(1) - 0.5 * dl[2] * dfb1;
  const double f0 = 0.5 * fmax(cmm, cm0) * (r0 - l0);
  const double f1 = 0.5 * fmax(cm0, cmp) * (r1 - l1);
  grad = 0.5 * (l1 + r1 - l0 - r0) * igb;
  diff = (f1 - f0) * igb;
}

// (256 lanes at most; two wavefronts per SIMD: the LDS of three workgroups
// per CU holds about that many anyway).  TOPO: UseTopography, launched for
// one altitude at a time (the host walks k = 1..nAlts), so that the level
// below already holds the step's result when dVarDAlt reads it.
// BLON, BLAT: the block's cells per direction when fixed at compile time
// (GITM's are, 9 x 9 by default), so that the LDS and cell indexing is
// constant; 0 takes them from the parameters.
template <int NV, bool TOPO, int BLON = 0, int BLAT = 0>
__global__ void __launch_bounds__(256, 2)
    horiz_kernel(HzParams p, gitm_vert_columns io, HzFields f, unsigned long long *err,
                 double *accg) {
  extern __shared__ double lds[];
  const int nlo = BLON ? BLON : p.n_lons, nla = BLAT ? BLAT : p.n_lats, NLo = nlo + 4,
            NLg = nla + 4;
  const int cells = nlo * nla, slab = NV * NLg * NLo;
  const int b = (int)blockIdx.x / p.alt_groups;
  const int k0 = p.k_first + ((int)blockIdx.x % p.alt_groups) * p.apw;
  const int napw = min(p.apw, p.n_alts - k0 + 1);
  const int tid = (int)threadIdx.x;
  const int a = tid / cells, cell = tid % cells;
  const bool active = a < napw;
  const int i = cell % nlo + 1, jl = cell / nlo + 1, k = k0 + a;
  // the advected fields (advance_horizontal.f90:154-167)
  auto field = [&](int v, int &comp) -> const gitm_vert_field & {
    if (v == V_RHO) { comp = 0; return io.rho; }
    if (v == V_TEMP) { comp = 0; return io.temperature; }
    if (v < V_NUM) { comp = v - V_VEL; return io.velocity; }
    if (v < V_VV) { comp = v - V_NUM; return io.ndensity_s; }
    if (v < V_INUM) { comp = v - V_VV; return io.vertical_velocity; }
    comp = v - V_INUM;
    return io.idensity_s;
  };
  // stage the slices: x[a][v][lat][lon], halo cells included; after the
  // slices, per altitude, the stage's new values, [v][cell].  The RK4 sums
  // live in device memory (accg, [workgroup][altitude][v][cell]): they are
  // touched once per stage, and the LDS they would take costs a workgroup
  // per CU.
  // Each thread takes the cells r2 = tid, tid + blockDim.x, .. of a slice
  // (two at most: a slice has (nla+4)(nlo+4) cells) and stages them for
  // every variable and altitude, so that the only integer divisions are the
  // per-cell ones, not one per element.
  {
    const int plane = NLg * NLo;
    for (int r2 = tid; r2 < plane; r2 += (int)blockDim.x) {
      const int lat = r2 / NLo - 1, lon = r2 % NLo - 1;
      for (int aa = 0; aa < napw; aa++) {
#pragma unroll
        for (int v = 0; v < NV; v++) {
          int comp;
          const gitm_vert_field &g = field(v, comp);
          lds[(aa * NV + v) * plane + r2] = *hf(g, lon, lat, k0 + aa, comp, b);
        }
      }
    }
  }
  const int aw = active ? a : 0;
  double *x = lds + (long long)aw * slab;
  double *nwb = lds + (long long)p.apw * slab + (long long)aw * NV * cells;
  double *accb = accg + ((long long)blockIdx.x * p.apw + aw) * NV * cells;
  auto X = [&](int v, int lat, int lon) -> double & {
    return x[(v * NLg + lat + 1) * NLo + lon + 1];
  };
  auto NW = [&](int v) -> double & { return nwb[v * cells + cell]; };
  auto ACC = [&](int v) -> double & { return accb[v * cells + cell]; };
  __syncthreads();
  // per-cell inputs of the updates
  double invr = 0, rd = 0, gam = 0, tl = 0, cosl = 0, sinl = 0, ive = 0, ivn = 0;
  if (active) {
    invr = *hf(io.inv_radial_distance, i, jl, k, 0, b);
    rd = *hf(f.raddist, i, jl, k, 0, b);
    gam = *hf(io.gamma, i, jl, k, 0, b);
    tl = *hf(f.tan_latitude, 1, jl, 0, 0, b);
    cosl = *hf(io.cos_latitude, 1, jl, 0, 0, b);
    sinl = sin(*hf(f.latitude, 1, jl, 0, 0, b));
    if (NV == NV_IONS) {
      ive = *hf(io.ivelocity, i, jl, k, 0, b);
      ivn = *hf(io.ivelocity, i, jl, k, 1, b);
    }
  }
  LineGeo glo{}, gla{};
  if (active) {
#pragma unroll
    for (int m = -1; m <= 2; m++) {
      glo.ifb[m + 1] = *hf(f.inv_dlon_fb, i + m, jl, k, 0, b);
      gla.ifb[m + 1] = *hf(f.inv_dlat_fb, i, jl + m, k, 0, b);
    }
    glo.dfb0 = *hf(f.dlon_fb, i, jl, k, 0, b);
    glo.dfb1 = *hf(f.dlon_fb, i + 1, jl, k, 0, b);
    gla.dfb0 = *hf(f.dlat_fb, i, jl, k, 0, b);
    gla.dfb1 = *hf(f.dlat_fb, i, jl + 1, k, 0, b);
    glo.igb = *hf(f.inv_dlon_gb, i, jl, k, 0, b);
    gla.igb = *hf(f.inv_dlat_gb, i, jl, k, 0, b);
#pragma unroll
    for (int m = -1; m <= 1; m++) {
      glo.cm[m + 1] = *hf(f.cmax, i + m, jl, k, 0, b);
      gla.cm[m + 1] = *hf(f.cmax, i, jl + m, k, 1, b);
    }
  }
  // UseTopography: 0.5 / dAlt and the slopes of the level surface
  // (advance_horizontal.f90:581-582, :870-871, :919-920)
  double hz = 0, slon = 0, slat = 0;
  if (TOPO && active) {
    hz = 0.5 / *hf(f.dalt, i, jl, k, 0, b);
    slon = *hf(f.dalt_dlon, i, jl, k, 0, b);
    slat = *hf(f.dalt_dlat, i, jl, k, 0, b);
  }
  const double dt = p.dt, om = p.omega;
  const double csin = sinl * 2 * om, ccos = cosl * 2 * om;
  const double cent = om * om * cosl * sinl;

#pragma unroll 1
  for (int stage = 0; stage < 4; stage++) {
    if (active) {
      // Grad and Diff of variable v in both directions (:588-662)
      auto gd = [&](int v, double &glon, double &dlon, double &glat, double &dlat) {
        rusanov_cell(i, nlo, [&](int m) { return X(v, jl, i + m); }, glo, p.beta, glon, dlon);
        rusanov_cell(jl, nla, [&](int m) { return X(v, jl + m, i); }, gla, p.beta, glat, dlat);
        if (TOPO) {  // dVarDAlt_C from the state arrays (:583-586, :591-662)
          int comp;
          const gitm_vert_field &g = field(v, comp);
          const double dz = hz * (*hf(g, i, jl, k + 1, comp, b) - *hf(g, i, jl, k - 1, comp, b));
          glon = glon - dz * slon;
          glat = glat - dz * slat;
        }
      };
      const double rho = X(V_RHO, jl, i), tmp = X(V_TEMP, jl, i);
      const double ve = X(V_VEL + 0, jl, i), vn = X(V_VEL + 1, jl, i), vu = X(V_VEL + 2, jl, i);
      double gE[4], gN[4], gU[4], gR[4], gT[4];  // glon, dlon, glat, dlat
      gd(V_VEL + 0, gE[0], gE[1], gE[2], gE[3]);
      gd(V_VEL + 1, gN[0], gN[1], gN[2], gN[3]);
      const double div = gN[2] + gE[0] - tl * vn * invr;
      gd(V_RHO, gR[0], gR[1], gR[2], gR[3]);
      gd(V_TEMP, gT[0], gT[1], gT[2], gT[3]);
      // the updates (:667-816), New starting from the stage's state
      NW(V_RHO) = rho - dt * (rho * div + gR[2] * vn + gR[0] * ve) + dt * (gR[1] + gR[3]);
      double we = ve -
                  dt * (vn * gE[2] + ve * gE[0] + ve * (vu - tl * vn) * invr + gT[0] +
                        gR[0] * tmp / rho) +
                  dt * (gE[1] + gE[3]);
      double wn = vn -
                  dt * (vn * gN[2] + ve * gN[0] + (vn * vu + tl * (ve * ve)) * invr + gT[2] +
                        gR[2] * tmp / rho) +
                  dt * (gN[1] + gN[3]);
      if (p.use_coriolis) {
        we = we + dt * (+csin * vn - ccos * vu);
        wn = wn - dt * (cent * rd + csin * ve);
      }
      NW(V_VEL + 0) = we;
      NW(V_VEL + 1) = wn;
      NW(V_TEMP) = tmp - dt * ((gam - 1) * tmp * div + gT[2] * vn + gT[0] * ve) +
                   dt * (gT[1] + gT[3]);
      gd(V_VEL + 2, gU[0], gU[1], gU[2], gU[3]);
      NW(V_VEL + 2) = vu - dt * (vn * gU[2] + ve * gU[0]) + dt * (gU[1] + gU[3]);
#pragma unroll 1
      for (int v = V_NUM; v < NV; v++) {
        double g0, g1, g2, g3;
        gd(v, g0, g1, g2, g3);
        const double q = X(v, jl, i);
        double w;
        if (v < V_VV)  // NDensityS
          w = q - dt * (q * div + g2 * vn + g0 * ve) + dt * (g1 + g3);
        else if (v < V_INUM)  // VerticalVelocity
          w = q - dt * (vn * g2 + ve * g0) + dt * (g3 + g1);
        else  // IDensityS(1:nIonsAdvect), carried by the ion velocity
          w = q - dt * (+g2 * ivn + g0 * ive) + dt * (g1 + g3);
        NW(v) = w;
      }
    }
    __syncthreads();  // every lane has read the stage's state
    if (active) {
      // the RK4 bookkeeping (:184-345)
#pragma unroll 1
      for (int v = 0; v < NV; v++) {
        int comp;
        const double orig = *hf(field(v, comp), i, jl, k, comp, b);
        const double K = NW(v) - X(v, jl, i);
        double u;
        if (stage == 0) {
          ACC(v) = K;
          u = 0.5 * K + orig;
        } else if (stage == 1) {
          ACC(v) = ACC(v) + 2.0 * K;
          u = 0.5 * K + orig;
        } else if (stage == 2) {
          ACC(v) = ACC(v) + 2.0 * K;
          u = K + orig;
        } else {
          const double sum = ACC(v) + K;
          u = orig + (1.0 / 6.0) * sum;
        }
        X(v, jl, i) = u;
      }
    }
    __syncthreads();
  }
  if (!active) return;
  // write back (:349-429)
  bool negative = false;
#pragma unroll
  for (int s = 0; s < gv::NSP; s++) negative = negative || X(V_NUM + s, jl, i) < 0.0;
  if (negative) {
    const int column = (b * nla + (jl - 1)) * nlo + (i - 1);
    atomicMin(err, gv::error_key(column, GITM_VERT_ENONPOSITIVE, k));
  }
  const double told = *hf(io.temperature, i, jl, k, 0, b);
  *hf(io.rho, i, jl, k, 0, b) = X(V_RHO, jl, i);
#pragma unroll
  for (int c = 0; c < 3; c++) *hf(io.velocity, i, jl, k, c, b) = X(V_VEL + c, jl, i);
  if (f.temp_source.ptr) *hf(f.temp_source, i, jl, k, 0, b) = X(V_TEMP, jl, i) - told;
  *hf(io.temperature, i, jl, k, 0, b) = X(V_TEMP, jl, i);
#pragma unroll
  for (int s = 0; s < gv::NSP; s++) {
    *hf(io.vertical_velocity, i, jl, k, s, b) = X(V_VV + s, jl, i);
    *hf(io.ndensity_s, i, jl, k, s, b) = X(V_NUM + s, jl, i);
  }
  if (NV == NV_IONS) {
#pragma unroll
    for (int s = 0; s < NIA; s++) {
      double v = X(V_INUM + s, jl, i);
      if (v < 1.0e2) v = 1.0e2;  // :405-417
      *hf(io.idensity_s, i, jl, k, s, b) = v;
    }
    double e = 0.0;  // electrons, :420-425
    for (int s = 0; s < gv::NIONS - 1; s++) e = e + *hf(io.idensity_s, i, jl, k, s, b);
    *hf(io.idensity_s, i, jl, k, gv::NIONS - 1, b) = e;
  }
}

// UseImprovedIonAdvection (advance_horizontal.f90:432-518), after the
// step: the divergence of the field-aligned ion velocity from UAM_Gradient
// (gradient.f90:1-86) divides the advected ions by 1 + Dt div on the
// interior cells of levels 1..nAlts; then the electrons are re-summed on
// every cell and level of the block.  One thread per element of the
// block's (-1:nLons+2,-1:nLats+2,-1:nAlts+2) array; each touches its own
// cell's ions only.
struct DivFields {
  gitm_vert_field ivel_par, glm, gl0, glp, gam, ga0, gap;
};

__global__ void __launch_bounds__(256)
    ion_divergence_kernel(HzParams p, gitm_vert_columns io, HzFields f, DivFields g) {
  const int nlo = p.n_lons, nla = p.n_lats, nA = p.n_alts;
  const long long NLo = nlo + 4, NLg = nla + 4, per_block = NLo * NLg * (nA + 4);
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= per_block * p.n_blocks) return;
  const int b = (int)(e / per_block);
  long long r = e % per_block;
  const int k = (int)(r / (NLo * NLg)) - 1;
  r %= NLo * NLg;
  const int lat = (int)(r / NLo) - 1, lon = (int)(r % NLo) - 1;
  if (lon >= 1 && lon <= nlo && lat >= 1 && lat <= nla && k >= 1 && k <= nA) {
    const double ir = *hf(io.inv_radial_distance, lon, lat, k, 0, b);
    auto v = [&](int comp, int lo, int la, int kk) {
      return *hf(g.ivel_par, lo, la, kk, comp, b);
    };
    double div = 0.0;
    if (nlo > 1) {  // east (gradient.f90:27-41)
      const double icl = 1. / fmax(*hf(io.cos_latitude, 1, lat, 0, 0, b), 0.17);
      const double gr = (*hf(g.glm, lon, 1, 0, 0, b) * v(0, lon - 1, lat, k) +
                         *hf(g.gl0, lon, 1, 0, 0, b) * v(0, lon, lat, k) +
                         *hf(g.glp, lon, 1, 0, 0, b) * v(0, lon + 1, lat, k)) *
                        ir * icl;
      div = div + gr;
    }
    {  // north (:44-56), then the tan(latitude) term (:462-477)
      const double gr = nla > 1 ? (*hf(g.gam, 1, lat, 0, 0, b) * v(1, lon, lat - 1, k) +
                                   *hf(g.ga0, 1, lat, 0, 0, b) * v(1, lon, lat, k) +
                                   *hf(g.gap, 1, lat, 0, 0, b) * v(1, lon, lat + 1, k)) *
                                      ir
                                : 0.0;
      div = div + gr;
      double tl = *hf(f.tan_latitude, 1, lat, 0, 0, b);
      if (tl > 5.0) tl = 5.0;
      if (tl < -5.0) tl = -5.0;
      div = div - tl * v(1, lon, lat, k) * ir;
    }
    {  // radial (:59-78), then 2 w / r (:480-493)
      const double r0 = *hf(f.raddist, lon, lat, k, 0, b);
      const double drp = *hf(f.raddist, lon, lat, k + 1, 0, b) - r0;
      const double drm = r0 - *hf(f.raddist, lon, lat, k - 1, 0, b);
      const double q = drm / drp;
      const double q2 = q * q;
      const double bottom = drm * (1 + q);
      const double gr =
          nA > 1 ? (q2 * v(2, lon, lat, k + 1) - v(2, lon, lat, k - 1) - (q2 - 1) * v(2, lon, lat, k)) /
                       bottom
                 : 0.0;
      div = div + gr;
      div = div + 2 * v(2, lon, lat, k) * ir;
    }
    div = div * *hf(io.cos_latitude, 1, lat, 0, 0, b);  // :498-502
    const double change = p.dt * div;
#pragma unroll
    for (int s = 0; s < NIA; s++) {  // :505-510
      double *n = hf(io.idensity_s, lon, lat, k, s, b);
      *n = *n / (1 + change);
    }
  }
  double el = 0.0;  // :512-516, every cell and level
  for (int s = 0; s < gv::NIONS - 1; s++) el = el + *hf(io.idensity_s, lon, lat, k, s, b);
  *hf(io.idensity_s, lon, lat, k, gv::NIONS - 1, b) = el;
}

// The message exchange on one device (exchange_messages_sphere.f90:4-95).
// One thread per (block, halo cell, level -1..nAlts+2): on levels 1..nAlts
// it copies every carried field from the global-grid cell the halo cell
// overlaps (past a pole: the mirrored row, the longitude turned by 180
// degrees); in the halo rows past a pole it then turns the east and north
// winds around on every level (:30-72).  Sources are blocks' own cells,
// which no thread writes.
struct XParams {
  int n_lons, n_lats, n_alts, nbl, nbt, ions;
};

struct XFields {
  gitm_vert_field itemp, etemp, ivel, latitude;
};

__global__ void __launch_bounds__(256)
    exchange_kernel(XParams p, gitm_vert_columns io, XFields f) {
  const int nlo = p.n_lons, nla = p.n_lats, nA = p.n_alts;
  const int NLo = nlo + 4, NLg = nla + 4, halo = NLo * NLg - nlo * nla;
  const long long per_block = (long long)halo * (nA + 4);
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= per_block * p.nbl * p.nbt) return;
  const int b = (int)(e / per_block);
  const int rem = (int)(e % per_block);
  const int k = rem / halo - 1, hc = rem % halo;
  // halo cell hc: the two bottom rows, the side columns of the block's
  // rows, the two top rows
  int lo, la;
  if (hc < 2 * NLo) {
    la = hc / NLo - 1;
    lo = hc % NLo - 1;
  } else if (hc < 2 * NLo + 4 * nla) {
    const int q = hc - 2 * NLo;
    la = q / 4 + 1;
    const int c = q % 4;
    lo = c < 2 ? c - 1 : nlo + c - 1;
  } else {
    const int q = hc - 2 * NLo - 4 * nla;
    la = nla + 1 + q / NLo;
    lo = q % NLo - 1;
  }
  const long long NLON = (long long)p.nbl * nlo, NLAT = (long long)p.nbt * nla;
  if (k >= 1 && k <= nA) {
    long long J = (long long)(b % p.nbl) * nlo + lo - 1, I = (long long)(b / p.nbl) * nla + la - 1;
    if (I < 0) {
      I = -1 - I;
      J += NLON / 2;
    } else if (I >= NLAT) {
      I = 2 * NLAT - 1 - I;
      J += NLON / 2;
    }
    J = ((J % NLON) + NLON) % NLON;
    const int sb = (int)(I / nla) * p.nbl + (int)(J / nlo);
    const int slo = (int)(J % nlo) + 1, sla = (int)(I % nla) + 1;
    auto cp = [&](const gitm_vert_field &g, int nc) {
      for (int c = 0; c < nc; c++) *hf(g, lo, la, k, c, b) = *hf(g, slo, sla, k, c, sb);
    };
    cp(io.rho, 1);
    cp(io.temperature, 1);
    cp(f.itemp, 1);
    cp(f.etemp, 1);
    cp(io.velocity, 3);
    cp(io.vertical_velocity, gv::NSP);
    cp(io.ndensity_s, gv::NSP);
    if (p.ions) {
      cp(io.idensity_s, gv::NIONS);
      cp(f.ivel, 3);
    }
  }
  const double half_pi = 3.14159265358979323846 / 2.0;
  const bool south = la <= 0 && *hf(f.latitude, 1, 0, 0, 0, b) < -half_pi;
  const bool north = la >= nla + 1 && *hf(f.latitude, 1, nla + 1, 0, 0, b) > half_pi;
  if (south || north) {
    for (int c = 0; c < 2; c++) {
      double *v = hf(io.velocity, lo, la, k, c, b);
      *v = -*v;
      if (f.ivel.ptr) {
        double *w = hf(f.ivel, lo, la, k, c, b);
        *w = -*w;
      }
    }
  }
}

// The checks after the exchange (:74-84): one thread per cell and level
// 1..nAlts of the blocks.
__global__ void __launch_bounds__(256)
    exchange_check_kernel(XParams p, gitm_vert_columns io, unsigned long long *err) {
  const int nlo = p.n_lons, nla = p.n_lats, nA = p.n_alts;
  const long long per_block = (long long)nlo * nla * nA;
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= per_block * p.nbl * p.nbt) return;
  const int b = (int)(e / per_block);
  const int rem = (int)(e % per_block);
  const int k = rem / (nlo * nla) + 1, cell = rem % (nlo * nla);
  const int lo = cell % nlo + 1, la = cell / nlo + 1;
  const int column = (b * nla + la - 1) * nlo + lo - 1;
  if (*hf(io.temperature, lo, la, k, 0, b) < 100.0)
    atomicMin(err, gv::error_key(column, GITM_VERT_ELOWTEMP, k));
  if (*hf(io.rho, lo, la, k, 0, b) < 0.0)
    atomicMin(err, gv::error_key(column, GITM_VERT_ENONPOSITIVE, k));
}

}  // namespace

extern "C" int gitm_horiz_exchange(gitm_vert_handle *h, const gitm_horiz_exchange_params *xp,
                                   const gitm_vert_columns *c,
                                   const gitm_horiz_exchange_fields *xf) {
  if (!h || !xp || !c || !xf) return GITM_VERT_EINVAL;
  const int nlo = c->n_lons, nla = c->n_lats;
  if (nlo < 1 || nla < 2 || c->n_alts < 1 || xp->n_blocks_lon < 1 || xp->n_blocks_lat < 1)
    return GITM_VERT_EINVAL;
  if ((long long)xp->n_blocks_lon * xp->n_blocks_lat != c->n_blocks) return GITM_VERT_EINVAL;
  if (((long long)xp->n_blocks_lon * nlo) % 2) return GITM_VERT_EINVAL;
  const gitm_vert_field *req[] = {&c->rho, &c->temperature, &c->velocity,
                                  &c->vertical_velocity, &c->ndensity_s, &xf->i_temperature,
                                  &xf->e_temperature, &xf->latitude};
  for (const gitm_vert_field *g : req)
    if (!g->ptr) return GITM_VERT_EINVAL;
  const bool ions = xp->use_ion_advection != 0;
  if (ions && (!c->idensity_s.ptr || !xf->ion_velocity.ptr)) return GITM_VERT_EINVAL;
  XParams p;
  p.n_lons = nlo;
  p.n_lats = nla;
  p.n_alts = c->n_alts;
  p.nbl = xp->n_blocks_lon;
  p.nbt = xp->n_blocks_lat;
  p.ions = ions;
  XFields f;
  f.itemp = xf->i_temperature;
  f.etemp = xf->e_temperature;
  f.ivel = xf->ion_velocity;
  f.latitude = xf->latitude;
  const long long halo = (long long)(nlo + 4) * (nla + 4) - (long long)nlo * nla;
  const long long n1 = halo * (c->n_alts + 4) * c->n_blocks;
  const long long n2 = (long long)nlo * nla * c->n_alts * c->n_blocks;
  if ((n1 + 255) / 256 > 0x7fffffffLL || (n2 + 255) / 256 > 0x7fffffffLL) return GITM_VERT_EINVAL;
  GV_HIP(hipSetDevice(h->device));
  hipLaunchKernelGGL(exchange_kernel, dim3((unsigned)((n1 + 255) / 256)), dim3(256), 0,
                     h->stream, p, *c, f);
  GV_HIP(hipGetLastError());
  hipLaunchKernelGGL(exchange_check_kernel, dim3((unsigned)((n2 + 255) / 256)), dim3(256), 0,
                     h->stream, p, *c, h->d_err);
  GV_HIP(hipGetLastError());
  return GITM_VERT_OK;
}

extern "C" int gitm_horiz_advance(gitm_vert_handle *h, const gitm_horiz_params *hp,
                                  const gitm_vert_columns *c, const gitm_horiz_fields *hf_) {
  if (!h || !hp || !c || !hf_) return GITM_VERT_EINVAL;
  const int nlo = c->n_lons, nla = c->n_lats;
  if (nlo < 1 || nla < 1 || c->n_blocks < 1 || c->n_alts < 1) return GITM_VERT_EINVAL;
  const int cells = nlo * nla;
  if (cells > 256) return GITM_VERT_EUNSUPPORTED;  // one lane per cell
  if (!(hp->dt > 0.0)) return GITM_VERT_EINVAL;
  const gitm_vert_field *req[] = {&c->rho, &c->temperature, &c->velocity,
                                  &c->vertical_velocity, &c->ndensity_s, &c->gamma,
                                  &c->inv_radial_distance, &c->cos_latitude,
                                  &hf_->inv_dlon_gb, &hf_->inv_dlon_fb, &hf_->dlon_fb,
                                  &hf_->inv_dlat_gb, &hf_->inv_dlat_fb, &hf_->dlat_fb,
                                  &hf_->radial_distance, &hf_->latitude, &hf_->tan_latitude,
                                  &hf_->cmax_horizontal};
  for (const gitm_vert_field *g : req)
    if (!g->ptr) return GITM_VERT_EINVAL;
  const bool ions = hp->use_ion_advection != 0;
  if (ions && (!c->idensity_s.ptr || !c->ivelocity.ptr)) return GITM_VERT_EINVAL;
  const bool improved = ions && hp->use_improved_ion_advection != 0;
  const bool topo = hp->use_topography != 0;
  if (topo && (!hf_->d_alt.ptr || !hf_->dalt_dlon.ptr || !hf_->dalt_dlat.ptr))
    return GITM_VERT_EINVAL;
  DivFields dv;
  dv.ivel_par = hf_->ivelocity_par;
  dv.glm = hf_->grad_lon_m;
  dv.gl0 = hf_->grad_lon_0;
  dv.glp = hf_->grad_lon_p;
  dv.gam = hf_->grad_lat_m;
  dv.ga0 = hf_->grad_lat_0;
  dv.gap = hf_->grad_lat_p;
  if (improved)
    for (const gitm_vert_field *g : {&dv.ivel_par, &dv.glm, &dv.gl0, &dv.glp, &dv.gam, &dv.ga0,
                                     &dv.gap})
      if (!g->ptr) return GITM_VERT_EINVAL;
  const int nv = ions ? NV_IONS : NV_NEUTRAL;
  // per altitude: the slab of every variable with halo cells, then the
  // stage's new values of the block's cells
  const size_t slab = (size_t)nv * ((nla + 4) * (nlo + 4) + cells) * sizeof(double);
  // altitudes per workgroup: as many as fit 256 lanes and 40 KB of LDS, so
  // that four workgroups share a CU (two wavefronts per SIMD, the register
  // limit); 9x9 blocks take one
  // (with topography one altitude per workgroup, whatever is asked)
  int apw = topo ? 1 : hp->altitudes_per_workgroup > 0 ? hp->altitudes_per_workgroup : 0;
  if (apw == 0) {
    apw = 1;
    while ((apw + 1) * cells <= 256 && (apw + 1) * slab <= 40 * 1024 && apw < 4) apw++;
  }
  if (apw * cells > 256 || apw * slab > 160 * 1024) return GITM_VERT_EINVAL;
  apw = std::min(apw, c->n_alts);
  HzParams p;
  p.dt = hp->dt;
  p.omega = hp->omega_body;
  p.beta = hp->beta_limiter;
  p.use_coriolis = hp->use_coriolis != 0;
  p.n_lons = nlo;
  p.n_lats = nla;
  p.n_alts = c->n_alts;
  p.n_blocks = c->n_blocks;
  p.apw = apw;
  p.alt_groups = topo ? 1 : (c->n_alts + apw - 1) / apw;
  p.k_first = 1;
  const long long nwg = (long long)p.alt_groups * c->n_blocks;
  if (nwg > 0x7fffffffLL) return GITM_VERT_EINVAL;
  HzFields f;
  f.inv_dlon_gb = hf_->inv_dlon_gb;
  f.inv_dlon_fb = hf_->inv_dlon_fb;
  f.dlon_fb = hf_->dlon_fb;
  f.inv_dlat_gb = hf_->inv_dlat_gb;
  f.inv_dlat_fb = hf_->inv_dlat_fb;
  f.dlat_fb = hf_->dlat_fb;
  f.raddist = hf_->radial_distance;
  f.latitude = hf_->latitude;
  f.tan_latitude = hf_->tan_latitude;
  f.cmax = hf_->cmax_horizontal;
  f.temp_source = hf_->temp_source;
  f.dalt = hf_->d_alt;
  f.dalt_dlon = hf_->dalt_dlon;
  f.dalt_dlat = hf_->dalt_dlat;
  GV_HIP(hipSetDevice(h->device));
  const int threads = ((apw * cells + 63) / 64) * 64;
  const size_t lds = (size_t)apw * slab;
  // the RK4 sums, in the handle's scratch (shared with the vertical step's
  // carries: both only use it within a launch on the handle's stream)
  const size_t need = (size_t)nwg * apw * nv * cells * sizeof(double);
  if (h->scratch_bytes < need) {
    if (h->scratch) {
      GV_HIP(hipStreamSynchronize(h->stream));
      (void)hipFree(h->scratch);
      h->scratch = nullptr;
      h->scratch_bytes = 0;
    }
    if (hipMalloc(&h->scratch, need) != hipSuccess) return GITM_VERT_ENOMEM;
    h->scratch_bytes = need;
  }
  // GITM's default 9 x 9 blocks take the kernels with the block shape fixed
  const bool b99 = nlo == 9 && nla == 9;
  auto launch = [&](const HzParams &q) -> int {
    const dim3 grid((unsigned)nwg), block(threads);
    if (ions && topo) {
      if (b99)
        hipLaunchKernelGGL((horiz_kernel<NV_IONS, true, 9, 9>), grid, block, lds, h->stream, q,
                           *c, f, h->d_err, h->scratch);
      else
        hipLaunchKernelGGL((horiz_kernel<NV_IONS, true>), grid, block, lds, h->stream, q, *c,
                           f, h->d_err, h->scratch);
    } else if (ions) {
      if (b99)
        hipLaunchKernelGGL((horiz_kernel<NV_IONS, false, 9, 9>), grid, block, lds, h->stream,
                           q, *c, f, h->d_err, h->scratch);
      else
        hipLaunchKernelGGL((horiz_kernel<NV_IONS, false>), grid, block, lds, h->stream, q, *c,
                           f, h->d_err, h->scratch);
    } else if (topo) {
      if (b99)
        hipLaunchKernelGGL((horiz_kernel<NV_NEUTRAL, true, 9, 9>), grid, block, lds,
                           h->stream, q, *c, f, h->d_err, h->scratch);
      else
        hipLaunchKernelGGL((horiz_kernel<NV_NEUTRAL, true>), grid, block, lds, h->stream, q,
                           *c, f, h->d_err, h->scratch);
    } else {
      if (b99)
        hipLaunchKernelGGL((horiz_kernel<NV_NEUTRAL, false, 9, 9>), grid, block, lds,
                           h->stream, q, *c, f, h->d_err, h->scratch);
      else
        hipLaunchKernelGGL((horiz_kernel<NV_NEUTRAL, false>), grid, block, lds, h->stream, q,
                           *c, f, h->d_err, h->scratch);
    }
    GV_HIP(hipGetLastError());
    return GITM_VERT_OK;
  };
  if (topo) {
    // one launch per altitude, in order on the stream (:150-430)
    for (int k = 1; k <= c->n_alts; k++) {
      p.k_first = k;
      const int rc = launch(p);
      if (rc) return rc;
    }
  } else {
    const int rc = launch(p);
    if (rc) return rc;
  }
  if (improved) {
    const long long n =
        (long long)(nlo + 4) * (nla + 4) * (c->n_alts + 4) * (long long)c->n_blocks;
    const long long nb = (n + 255) / 256;
    if (nb > 0x7fffffffLL) return GITM_VERT_EINVAL;
    hipLaunchKernelGGL(ion_divergence_kernel, dim3((unsigned)nb), dim3(256), 0, h->stream, p,
                       *c, f, dv);
    GV_HIP(hipGetLastError());
  }
  return GITM_VERT_OK;
}

// ---------------------------------------------------------------------------
// Host-array entry point (the Fortran drop-in): every field the step reads is
// staged on the device over exactly the index ranges the step touches, halo
// cells included; the written ones come back.
// ---------------------------------------------------------------------------
namespace {
constexpr size_t kHostStageBase = 64;  // staging slots after the vertical entry's
}  // namespace

extern "C" int gitm_horiz_advance_host(gitm_vert_handle *h, const gitm_horiz_params *hp,
                                       const gitm_vert_columns *c,
                                       const gitm_horiz_fields *hf_) {
  if (!h || !hp || !c || !hf_) return GITM_VERT_EINVAL;
  gitm_vert_columns d = *c;
  gitm_horiz_fields g = *hf_;
  const int n = c->n_lons, m = c->n_lats, nA = c->n_alts, nb = c->n_blocks;
  if (n < 1 || m < 1 || nA < 1 || nb < 1) return GITM_VERT_EINVAL;
  const bool ions = hp->use_ion_advection != 0;
  const bool improved = ions && hp->use_improved_ion_advection != 0;
  std::vector<HostSpan> sp;
  auto full = [&](gitm_vert_field *f, int nc, bool w) {
    sp.push_back({f, -1, n + 2, -1, m + 2, -1, nA + 2, nc, w});
  };
  auto lat1 = [&](gitm_vert_field *f) {  // (-1:nLats+2, nBlocks)
    sp.push_back({f, 1, 1, -1, m + 2, f->first_level, f->first_level, 1, false});
  };
  full(&d.rho, 1, true);
  full(&d.temperature, 1, true);
  full(&d.velocity, 3, true);
  full(&d.vertical_velocity, gv::NSP, true);
  full(&d.ndensity_s, gv::NSP, true);
  full(&d.gamma, 1, false);
  full(&d.inv_radial_distance, 1, false);
  lat1(&d.cos_latitude);
  if (ions) {
    full(&d.idensity_s, gv::NIONS, true);
    full(&d.ivelocity, 3, false);
  }
  for (gitm_vert_field *f : {&g.inv_dlon_gb, &g.inv_dlon_fb, &g.dlon_fb, &g.inv_dlat_gb,
                             &g.inv_dlat_fb, &g.dlat_fb, &g.radial_distance})
    full(f, 1, false);
  lat1(&g.latitude);
  lat1(&g.tan_latitude);
  sp.push_back({&g.cmax_horizontal, 0, n + 1, 0, m + 1, 0, nA + 1, 2, false});
  if (g.temp_source.ptr) sp.push_back({&g.temp_source, 1, n, 1, m, 1, nA, 1, true});
  if (hp->use_topography)
    for (gitm_vert_field *f : {&g.d_alt, &g.dalt_dlon, &g.dalt_dlat})
      sp.push_back({f, 1, n, 1, m, 1, nA, 1, false});
  if (improved) {
    full(&g.ivelocity_par, 3, false);
    for (gitm_vert_field *f : {&g.grad_lon_m, &g.grad_lon_0, &g.grad_lon_p})
      sp.push_back({f, 1, n, 1, 1, f->first_level, f->first_level, 1, false});
    for (gitm_vert_field *f : {&g.grad_lat_m, &g.grad_lat_0, &g.grad_lat_p})
      sp.push_back({f, 1, 1, 1, m, f->first_level, f->first_level, 1, false});
  }
  GV_HIP(hipSetDevice(h->device));
  HostStaging st;
  int rc = stage_in(h, sp, nb, kHostStageBase, st);
  if (rc) return rc;
  rc = gitm_horiz_advance(h, hp, &d, &g);
  if (rc) return rc;
  rc = stage_out(h, sp, kHostStageBase, st);
  if (rc) return rc;
  return gitm_vert_get_error(h, nullptr);
}
