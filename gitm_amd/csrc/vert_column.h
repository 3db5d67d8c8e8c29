// vert_column.h -- the column-step kernels of GITM's vertical dynamics,
// written for CDNA4 (gfx950).
//
// Mapping: one workgroup of NT = 64*W threads advances one (lon, lat) column.
// Thread j owns level k = j-1 (levels -1..nAlts+2, L = nAlts+4 <= NT) and
// keeps that level's RK4 state in registers.  Everything a level needs from
// its neighbours (the i+-2 reconstruction stencils, face states, the
// hydrostatic recurrence, ghost cells) is staged through LDS.  All RK4
// stages, the AUSM+-up / Rusanov fluxes, the boundary conditions and the
// implicit neutral-friction solve run inside one launch; the column goes to
// HBM once on entry and once on exit of n_steps vertical steps.
//
// Floating-point work is FP64 throughout and keeps the reference's operation
// order (Fortran evaluates left to right; the build disables contraction), so
// results differ from the reference only through the last-bit behaviour of
// exp/log/pow/tanh.  Face quantities that the reference computes twice (once
// as the M12 face of cell i, once as the P12 face of cell i-1,
// src/vertical_solver_ausm.f90:2096-2104) are computed once per face here.
//
// The same source compiles for the host (GV_HOST_EMULATION) where a test
// harness runs the NT threads of a workgroup as CPU threads to check the
// kernel logic against the CPU oracle bit for bit.
#pragma once

#include <stdint.h>

#include "gitm_vert.h"

#ifdef GV_HOST_EMULATION
#include <cmath>
#define GV_FN static inline
#define GV_MEM inline
#define GV_HD static inline
#else
#include <hip/hip_runtime.h>
#define GV_FN static __device__ __forceinline__
#define GV_MEM __device__ __forceinline__
#define GV_HD static __host__ __device__ inline
#endif

namespace gv {

// Forward taps: a context whose kTaps is true receives a stage's
// intermediates (the adjoint kernel records them for its reverse sweep);
// for every other context the calls compile away.
template <class C, class = void>
struct ctx_taps {
  static constexpr bool value = false;
};
template <class C>
struct ctx_taps<C, decltype((void)C::kTaps)> {
  static constexpr bool value = C::kTaps;
};

constexpr int NSP = 6;
constexpr int NIONS = 10;
constexpr double KB = 1.3807e-23;         // ModConstants.f90:15
constexpr double PI = 3.141592653589793;  // ModConstants.f90:23

// species (ModEarth.f90:12-18, 0-based)
constexpr int S_O = 0, S_N4S = 3, S_NO = 4;
GV_FN constexpr bool photochemical(int s) { return s == 3 || s == 4; }

// per-level state kept in registers
constexpr int F_LR = 0;   // LogRho
constexpr int F_T = 1;    // Temp (= Temperature*TempUnit)
constexpr int F_LN = 2;   // LogNS(1:6)
constexpr int F_W = 8;    // VertVel(1:6)
constexpr int F_VE = 14;  // Vel_GD(iEast_)
constexpr int F_VN = 15;  // Vel_GD(iNorth_)
constexpr int F_VU = 16;  // Vel_GD(iUp_)
constexpr int NF = 17;
// ion state (UseIonAdvection): LogINS of the nIons-1 = 9 ions (electrons are
// their sum); only the first nIonsAdvect = 2 are advected (ModEarth.f90:39)
constexpr int NI = 9;
constexpr int NI_ADVECT = 2;

// Diff0/1e17 and DiffExp (ModEarth.f90:237-261); both symmetric
constexpr double D0R[NSP][NSP] = {{0.000, 0.969, 0.969, 0.969, 0.715, 3.440},
                                  {0.969, 0.000, 0.715, 0.969, 0.715, 3.208},
                                  {0.969, 0.715, 0.000, 0.969, 0.527, 2.939},
                                  {0.969, 0.969, 0.969, 0.000, 0.969, 2.939},
                                  {0.715, 0.715, 0.527, 0.969, 0.000, 2.939},
                                  {3.440, 3.208, 2.939, 2.939, 2.939, 0.000}};
constexpr double DEX[NSP][NSP] = {{0.000, 0.774, 0.774, 0.774, 0.750, 0.749},
                                  {0.774, 0.000, 0.750, 0.774, 0.750, 0.710},
                                  {0.774, 0.750, 0.000, 0.774, 0.810, 0.718},
                                  {0.774, 0.774, 0.774, 0.000, 0.774, 0.718},
                                  {0.750, 0.750, 0.810, 0.774, 0.000, 0.718},
                                  {0.749, 0.710, 0.718, 0.718, 0.718, 0.000}};

// the distinct DiffExp values and where each pair of species takes its own
constexpr double DEXU[6] = {0.774, 0.750, 0.749, 0.810, 0.710, 0.718};
constexpr int DEXI[NSP][NSP] = {{0, 0, 0, 0, 1, 2},
                                {0, 0, 1, 0, 1, 4},
                                {0, 1, 0, 0, 3, 5},
                                {0, 0, 0, 0, 0, 5},
                                {1, 1, 3, 0, 0, 5},
                                {2, 4, 5, 5, 5, 0}};

// 4-point reconstruction of calc_facevalues_alts_* (vertical_solver_ausm.f90
// :1052-1053); Factor2 is the literal 0.0416667, not 1/24.
constexpr double FACT1 = 0.6250000, FACT2 = 0.0416667;

// Kernel-side run constants (built on the host from gitm_vert_params).
struct KParams {
  double dt, beta, maxvv, min_tec, vert_tau, omega;
  double mass[NSP], tdc[NSP];
  double lmass[NSP];  // log(mass), for the log-space forms of the device math
  double imass[NSP];  // 1/mass, ditto
  double log_min_kp, log_max_kp, log_min_ku, log_max_ku;  // log(1e-19), ...
  int n_alts, n_steps, solver;
  int friction, coriolis, ions, improved_ions, nighttime_ion_bcs;
  int msis;                 // UseMSIS: level-0 ghost values from the caller's MSIS
  int tides;                // bottom tides from the caller's values: 0 off, 1 GSWM, 2 WACCM
  int damping;              // UseDamping: the Rusanov step's VertTau profile
  int scr_fields;           // fields per level carried in device memory (see carry_in_lds)
  unsigned long long *err;  // first-error word, see error_key()
  double *scratch;          // RK4 carries: n_cols x scr_fields x (nAlts+4) doubles
};

// Where the RK4 carries (Y and acc, 2*NF per level) live: in LDS for columns
// of up to three wavefronts (they then cost 29 KiB at nAlts = 100 and leave
// room for two columns per CU), in device memory for the tallest columns,
// whose LDS would not fit.  Ion carries always live in device memory.
GV_HD constexpr bool carry_in_lds(int waves) { return waves <= 3; }
GV_HD constexpr int scratch_fields(bool carry_lds, bool ions) {
  return (carry_lds ? 0 : 2 * NF) + (ions ? 2 * NI : 0);
}

// Error word: the lowest (column, code, level) wins an atomic min.
GV_FN unsigned long long error_key(int col, int code, int level) {
  return ((unsigned long long)(unsigned)col << 32) |
         ((unsigned long long)(unsigned)code << 16) |
         (unsigned long long)(unsigned)(level + 1024);
}

// LDS layout of one column: one record per level (levels -3..nAlts+4, i.e.
// two guard rows on each side so that j-2..j+2 stays inside) holding the
// NSTAT static tables, the NBUF exchange buffers, the RK4 carries when they
// live in LDS and the ion rows; then a scalar area and an int buffer.  With
// the fields of a level side by side, every access of a lane to its own or a
// neighbouring level is the lane's record address plus a constant offset,
// and an odd record length keeps 64-bit accesses free of bank conflicts.
enum StatArray {
  A_DF = 0, A_IF, A_DC, A_CMF, A_KP, A_KU,  // also read at neighbouring levels
  A_GL1, A_GL2, A_GL3, A_GR1, A_GR2, A_GR3,  // gamma factors of face k|k+1
  A_HF, A_AP, A_AM, A_VOL, A_INVR, A_EDDY, A_GRAV,
  A_GLC0, A_GLC1, A_GLC2, A_GLC3, A_GLC4,
  NSTAT
};
constexpr int NBUF = 21;
enum Scalar {
  SC_ALT0 = 0, SC_MMM_M1, SC_MMM_0, SC_MG0, SC_MGT1, SC_MGT2,
  SC_CA = 6,    // bottom extrapolation coefficients for level 0 (5)
  SC_CB = 11,   // ... for level -1 (5)
  SC_BOT = 16,  // bottom BC results: 6 species x 6 values
  SC_TOP = 52,  // top BC results: 6 species x 4 values
  SC_CTA = 76,  // top one-sided weights for level nAlts+1 (5), ions
  SC_CTB = 81,  // ... for level nAlts+2 (5)
  SC_TEC = 86,  // TEC of the column (ions, nighttime BCs)
  NSCAL = 88
};

// Doubles per level record: tables, exchange buffers, carries (when in LDS),
// ions (9 LogINS and IVel(up)).  Odd in every configuration.
GV_HD constexpr int record_len(bool ions, bool carry_lds) {
  return NSTAT + NBUF + (carry_lds ? 2 * NF : 0) + (ions ? NI + 1 : 0);
}
static_assert(record_len(false, false) % 2 == 1 && record_len(false, true) % 2 == 1 &&
                  record_len(true, false) % 2 == 1 && record_len(true, true) % 2 == 1,
              "an odd record length keeps 64-bit LDS accesses conflict-free");

// LDS doubles of one column: nAlts+8 level records, the scalars and the int
// buffer.
GV_HD constexpr unsigned long lds_doubles(int n_alts, bool ions, bool carry_lds) {
  return (unsigned long)record_len(ions, carry_lds) * (unsigned long)(n_alts + 8) + NSCAL +
         (unsigned long)(n_alts + 9) / 2;
}
// LDS bytes of one column: at nAlts = 100, 69 KiB with the carries (two
// columns per 160 KiB CU), 77 KiB with ions as well.
GV_HD constexpr unsigned long lds_bytes(int n_alts, bool ions, bool carry_lds) {
  return lds_doubles(n_alts, ions, carry_lds) * sizeof(double);
}

// ---------------------------------------------------------------------------
// small numerics
// ---------------------------------------------------------------------------
// min/max/limiter/LU below are templates over the column code's number type
// (double, or the taped type of the adjoint build, gitm_amd/csrc/vert_ad.h)
template <class A, class B>
GV_FN auto dmax(A a, B b) -> decltype(a + b) {
  return a > b ? a : b;
}
template <class A, class B>
GV_FN auto dmin(A a, B b) -> decltype(a + b) {
  return a < b ? a : b;
}
#ifndef GV_HOST_EMULATION
// On the device two doubles take v_max_f64 / v_min_f64 (one instruction in
// place of a compare and two selects).  They agree with the selects above
// except for NaN operands and the sign of a zero result (min(-0, +0) is -0
// here, +0 above), which the column code does not meet: its operands are
// limited slopes, sound speeds, Mach numbers and winds that are finite, and
// a zero result is only ever added or multiplied.
GV_FN double dmax(double a, double b) { return __builtin_fmax(a, b); }
GV_FN double dmin(double a, double b) { return __builtin_fmin(a, b); }
#endif

// ln_fast: a natural logarithm in 33 FP64 instructions, against 76 for the
// device library's log (which carries extra precision); the AUSM stage takes
// about 15 per level.  Method: x = m 2^k with m in [sqrt(1/2), sqrt(2)),
// f = m - 1 (exact), s = f / (2 + f), and
//   log(1 + f) = 2 atanh(s) = f - (f^2/2 - s (f^2/2 + R(s^2))),
//   R(z) = sum_{i=1..10} 2 z^i / (2i + 1),
// the atanh series, whose remainder past z^10 is below 2^-58 of the result
// for |s| <= 0.1716; k log 2 is added as k LN2_HI (exact: LN2_HI has 21
// significant bits) plus k LN2_LO.  The quotient s takes a reciprocal seed
// refined by two Newton steps.  Error below one ulp (tests/test_ln_fast.py);
// zero, negative, infinite and NaN arguments give log's values.
#ifdef GV_HOST_EMULATION
GV_FN double frexp_mant(double x, int &e) { return std::frexp(x, &e); }
GV_FN double rcp_seed(double d) { return (double)(1.0f / (float)d); }
#else
GV_FN double frexp_mant(double x, int &e) {
  e = __builtin_amdgcn_frexp_exp(x);
  return __builtin_amdgcn_frexp_mant(x);
}
GV_FN double rcp_seed(double d) { return __builtin_amdgcn_rcp(d); }
#endif
GV_FN double ln_fast(double x) {
  constexpr double LN2_HI = 0.6931467056274414;      // 0x3fe62e4200000000
  constexpr double LN2_LO = 4.7493250390316726e-07;  // ln 2 - LN2_HI
  int e;
  double m = frexp_mant(x, e);  // x = m 2^e, m in [1/2, 1)
  const bool low = m < 0.7071067811865476;
  m = low ? m + m : m;
  e = low ? e - 1 : e;
  const double f = m - 1.0, d = 2.0 + f;
  double r = rcp_seed(d);
  r = fma(fma(-d, r, 1.0), r, r);
  r = fma(fma(-d, r, 1.0), r, r);
  const double s = f * r, z = s * s;
  double R = 2.0 / 21.0;
  R = fma(R, z, 2.0 / 19.0);
  R = fma(R, z, 2.0 / 17.0);
  R = fma(R, z, 2.0 / 15.0);
  R = fma(R, z, 2.0 / 13.0);
  R = fma(R, z, 2.0 / 11.0);
  R = fma(R, z, 2.0 / 9.0);
  R = fma(R, z, 2.0 / 7.0);
  R = fma(R, z, 2.0 / 5.0);
  R = fma(R, z, 2.0 / 3.0);
  R = R * z;
  const double hfsq = 0.5 * f * f, k = (double)e;
  double res = k * LN2_HI + (f - (hfsq - (s * (hfsq + R) + k * LN2_LO)));
  if (!(x > 0.0)) res = x == 0.0 ? -__builtin_inf() : __builtin_nan("");
  if (x == __builtin_inf()) res = x;
  return res;
}
// is_double<R>: the plain build's number type (not the taped or counting
// types of the adjoint and operation-count builds)
template <class T>
struct is_double {
  static constexpr bool value = false;
};
template <>
struct is_double<double> {
  static constexpr bool value = true;
};

// ln: the column code's logarithm.  The device takes ln_fast; the host
// build keeps the C library's log, so that it stays bit for bit with the
// oracle; the adjoint build's taped numbers take their own log.
#ifdef GV_HOST_EMULATION
GV_FN double ln(double x) { return std::log(x); }
#else
GV_FN double ln(double x) { return ln_fast(x); }
#endif
template <class R>
GV_FN R ln(const R &x) {
  return log(x);
}

// Limiter_mc, ModLimiter.f90:16-36
template <class R>
GV_FN R limiter_mc(double beta, R up, R dn) {
  if (up > 0.0) {
    if (dn > 0.0) return dmin(dmin(beta * up, beta * dn), (up + dn) * 0.5);
    return 0.0;
  }
  if (dn < 0.0) return dmax(dmax(beta * up, beta * dn), (up + dn) * 0.5);
  return 0.0;
}

// Which value Limiter_mc took: 0 (zero), 1 (beta*up), 2 (beta*dn) or 3
// ((up+dn)/2), as limiter_mc decides it (the adjoint build records it).
template <class R>
GV_FN int limiter_code(double beta, R up, R dn) {
  if (up > 0.0) {
    if (!(dn > 0.0)) return 0;
    const R a = beta * up, b = beta * dn;
    const R m = a < b ? a : b;
    return (m < (up + dn) * 0.5) ? (a < b ? 1 : 2) : 3;
  }
  if (!(dn < 0.0)) return 0;
  const R a = beta * up, b = beta * dn;
  const R m = a > b ? a : b;
  return (m > (up + dn) * 0.5) ? (a > b ? 1 : 2) : 3;
}

// Crout LU with implicit partial pivoting + substitution for the 6x6
// friction system.  Same pivot rule as src/ludcmp.f (rows scaled by
// 1/max|a_ij|, the last maximum wins, a zero pivot becomes 1e-20) and the
// leading-zero skip of src/lubksb.f.  Fully unrolled: every array index is a
// compile-time constant and row interchanges are selects, so the system stays
// in registers.  Returns false for a zero row (singular matrix).
template <class R>
GV_FN bool lu_solve6(R (&a)[6][6], R (&b)[6]) {
  R sc[6];
  bool ok = true;
#pragma unroll
  for (int i = 0; i < 6; i++) {
    R big = 0.0;
#pragma unroll
    for (int j = 0; j < 6; j++) {
      R v = fabs(a[i][j]);
      if (v > big) big = v;
    }
    if (big == 0.0) ok = false;
    sc[i] = 1. / big;
  }
  int perm[6];
#pragma unroll
  for (int j = 0; j < 6; j++) {
#pragma unroll
    for (int i = 0; i < j; i++) {
      R s = a[i][j];
#pragma unroll
      for (int q = 0; q < i; q++) s = s - a[i][q] * a[q][j];
      a[i][j] = s;
    }
    R best = 0.0;
    int piv = j;
#pragma unroll
    for (int i = j; i < 6; i++) {
      R s = a[i][j];
#pragma unroll
      for (int q = 0; q < j; q++) s = s - a[i][q] * a[q][j];
      a[i][j] = s;
      R t = sc[i] * fabs(s);
      if (t >= best) {
        piv = i;
        best = t;
      }
    }
#pragma unroll
    for (int r = j + 1; r < 6; r++) {
      if (piv == r) {
#pragma unroll
        for (int q = 0; q < 6; q++) {
          R t = a[r][q];
          a[r][q] = a[j][q];
          a[j][q] = t;
        }
        sc[r] = sc[j];
      }
    }
    perm[j] = piv;
    if (a[j][j] == 0.) a[j][j] = 1.0e-20;
    if (j != 5) {
      R r = 1. / a[j][j];
#pragma unroll
      for (int i = j + 1; i < 6; i++) a[i][j] = a[i][j] * r;
    }
  }
  int first = -1;
#pragma unroll
  for (int i = 0; i < 6; i++) {
    R s = b[i];
#pragma unroll
    for (int r = i + 1; r < 6; r++)
      if (perm[i] == r) {
        s = b[r];
        b[r] = b[i];
      }
    if (first >= 0) {
#pragma unroll
      for (int q = 0; q < i; q++)
        if (q >= first) s = s - a[i][q] * b[q];
    } else if (s != 0.) {
      first = i;
    }
    b[i] = s;
  }
#pragma unroll
  for (int i = 5; i >= 0; i--) {
    R s = b[i];
#pragma unroll
    for (int q = i + 1; q < 6; q++) s = s - a[i][q] * b[q];
    b[i] = s / a[i][i];
  }
  return ok;
}

// calc_neutral_friction, calc_neutral_friction.f90:47-106, one level.
// The reference form takes tp[u] = Temp**DEXU[u].  The device form
// (DEVICE) takes tp[u] = Temp**-DEXU[u] and uses that the total density
// cancels: C_iq = kT/m_i n_q/N / (1e-4 Diff0 T^e / (N 1e-6))
//              = kT/m_i n_q (1e-2/Diff0) T^-e,
// which needs no division (52 per level in the reference form).
template <bool DEVICE, class R>
GV_FN bool neutral_friction(const KParams &P, R dtin, R (&vel)[NSP], R nt,
                            const R (&ns)[NSP], const R (&glc)[NSP], R eddy, R temp,
                            const R (&tp)[6]) {
  R m[NSP][NSP], ec[NSP];
  if (DEVICE) {
    const R kT = KB * temp;
#pragma unroll
    for (int i = 0; i < NSP; i++) {
      const R kTOverM = kT * P.imass[i];
      R sum = 0.0;
#pragma unroll
      for (int q = 0; q < NSP; q++) {
        if (q == i) continue;
        const double g = 1.0e-19 / D0R[i][q];  // 1e-2 / Diff0, a constant
        R c = kTOverM * ns[q] * (g * tp[DEXI[i][q]]);
        sum = sum + c;
        m[i][q] = -dtin * c;
      }
      m[i][i] = 1.0 + dtin * sum;
      ec[i] = -1.0 * eddy * glc[i];
    }
  } else {
    R tdij[NSP][NSP];
#pragma unroll
    for (int i = 0; i < NSP; i++)
#pragma unroll
      for (int q = i + 1; q < NSP; q++) {
        // Diff0 and DiffExp are symmetric, so TempDij(i,j) == TempDij(j,i)
        const double d0 = 1.0e17 * D0R[i][q];
        R v = (1.0e-04) * (d0 * tp[DEXI[i][q]]) / (nt * (1.0e-06));
        tdij[i][q] = v;
        tdij[q][i] = v;
      }
    const R denscale = 1.0 / nt;
#pragma unroll
    for (int i = 0; i < NSP; i++) {
      const R kTOverM = KB * temp / P.mass[i];
      R sum = 0.0;
#pragma unroll
      for (int q = 0; q < NSP; q++) {
        if (q == i) continue;
        R c = kTOverM * denscale * ns[q] / tdij[i][q];
        sum = sum + c;
        m[i][q] = -dtin * c;
      }
      m[i][i] = 1.0 + dtin * sum;
      ec[i] = -1.0 * eddy * glc[i];
    }
  }
  bool ok = lu_solve6(m, vel);
#pragma unroll
  for (int i = 0; i < NSP; i++) vel[i] = vel[i] + ec[i];
  return ok;
}

// one-sided 5-point extrapolation coefficients, set_vertical_bcs.Earth.f90
// :228-248 (h1..h4 are the spacings above the point being filled)
template <class R>
GV_FN void bc_fwd_coefs(R h1, R h2, R h3, R h4, R *C) {
  R H1 = h1, H2 = h1 + h2, H3 = h1 + h2 + h3, H4 = h1 + h2 + h3 + h4;
  C[0] = -1.0 * (H2 * H3 * H4 + H1 * H3 * H4 + H1 * H2 * H4 + H1 * H2 * H3) /
         (H1 * H2 * H3 * H4);
  C[1] = 1.0 * (H2 * H3 * H4) / (h1 * h2 * (h2 + h3) * (h2 + h3 + h4));
  C[2] = -1.0 * (H1 * H3 * H4) / (H2 * h2 * h3 * (h3 + h4));
  C[3] = 1.0 * (H1 * H2 * H4) / (H3 * (h3 + h2) * h3 * h4);
  C[4] = -1.0 * (H1 * H2 * H3) / (H4 * (h2 + h3 + h4) * (h3 + h4) * h4);
}

// backward weights for the top ion slopes, set_vertical_bcs.Earth.f90
// :505-525 (h1..h4 are the spacings below the point being filled)
template <class R>
GV_FN void bc_bwd_coefs(R h1, R h2, R h3, R h4, R *C) {
  R H1 = h1, H2 = h1 + h2, H3 = h1 + h2 + h3, H4 = h1 + h2 + h3 + h4;
  C[0] = 1.0 * (H2 * H3 * H4 + H1 * H3 * H4 + H1 * H2 * H4 + H1 * H2 * H3) /
         (H1 * H2 * H3 * H4);
  C[1] = -1.0 * (H2 * H3 * H4) / (h1 * h2 * (h2 + h3) * (h2 + h3 + h4));
  C[2] = 1.0 * (H1 * H3 * H4) / (H2 * h2 * h3 * (h3 + h4));
  C[3] = -1.0 * (H1 * H2 * H4) / (H3 * (h3 + h2) * h3 * h4);
  C[4] = 1.0 * (H1 * H2 * H3) / (H4 * (h2 + h3 + h4) * (h3 + h4) * h4);
}

// Per-thread constants of the owned level.
// Per-thread constants of the owned level kept in registers; the rest live
// in the LDS tables of StatArray:
//   GL*/GR*  gamma factors 1/(g-1), g/(g-1), 2(g-1)/(g+1) of face k|k+1
//   HF       0.35*exp(-(z - z0)/45 km)                    (ausm :950)
//   AP/AM/VOL face areas r^2 and cell volume (r^3)/3       (ausm :2029-2040)
//   GLC*     5-point d ln(rho_s/rho)/dz weights            (ausm :659-676)
template <class R>
struct LevelStatic {
  R eg;   // EffectiveGravity = g + Centrifugal/InvR (ausm :582-589)
  R mgh;  // 0.5*(eg(k) + eg(k-1))                    (ausm :697)
  R gm1;  // Gamma - 1
  R tu;   // TempUnit
};

// ---------------------------------------------------------------------------
// One column.  Ctx supplies lane(), sync(), sync_or(bool), report(key),
// diff_powers(temp, sign, out[6]) = temp**(sign*DEXU[0..5]) (the friction
// coefficients' temperature dependence, calc_neutral_friction.f90:77-79) and
// device_math().  With device_math() the column uses the log-space forms of
// a few expressions: log(m n_s / rho) = log m + ln n_s - ln rho,
// log(m n_s) = log m + ln n_s, log(n_s k T) = ln n_s + log(k T), the new
// rho as the sum of m n_s instead of m exp(log n_s), and the friction
// coefficients without divisions.  They replace 18 logs
// and 6 exps per level and stage by one log, and differ from the
// reference's rounding by an ulp of the logarithm, as the device's own
// exp/log do.  The GPU always uses them; the host build of the kernels runs
// either form, so that the reference forms can be checked bit for bit.
// ---------------------------------------------------------------------------
// CARRY: the RK4 carries in LDS (true) or in device memory; by default as
// carry_in_lds of the tile.
template <int NT, class Ctx, bool IONS, bool CARRY = carry_in_lds(NT / 64)>
struct Column {
  using R = typename Ctx::real;  // double, or the adjoint build's taped type
  static constexpr bool CARRY_LDS = CARRY;
  Ctx &cx;
  const KParams &kp;
  R *lds;
  const int j, k, nA, L, P;
  const bool lv;  // owns a level
  int col;        // column number in error reports
  const R dt;  // Dt of this column's run
  LevelStatic<R> S;
  R U[NF];
  R *scr;  // this column's RK4 carries (Y, acc) in device memory
  R centrifugal, coriolis;
  bool plasma = false;  // nighttime plasmasphere column (bc :492-497)
  // UseMSIS: the caller's msis_bcs values of level 0, on the lanes that use
  // them: LogNS(0) of species j on lane j < 6; Temp(0) and the east and
  // north winds on lane 1 (level 0).  With tides the winds are the tides'
  // (set after MSIS's, :150-159), and with GSWM tides but no MSIS, msis_t
  // holds the Temp(0) of the step's first boundary pass (see bcs()).
  R msis_ln = 0.0, msis_t = 0.0, msis_ve = 0.0, msis_vn = 0.0;
  R tide_t = 0.0;  // tides: TidesTemp of level 0 on lane 1
  R vtau = 1.0;    // VertTau of the owned level (Rusanov)

  // n_alts: the column's nAlts (a batch's members have their own)
  GV_MEM Column(Ctx &c, const KParams &p, R *l, int column, R step_dt, int n_alts)
      : cx(c), kp(p), lds(l), j(c.lane()), k(c.lane() - 1), nA(n_alts), L(n_alts + 4),
        P(n_alts + 8), lv(c.lane() < n_alts + 4), col(column), dt(step_dt) {}
  GV_MEM Column(Ctx &c, const KParams &p, R *l, int column, R step_dt)
      : Column(c, p, l, column, step_dt, p.n_alts) {}

  // record layout: tables | exchange buffers | carries Y, acc | ions
  static constexpr int RS = record_len(IONS, CARRY_LDS);
  static constexpr int OFF_X = NSTAT;
  static constexpr int OFF_Y = NSTAT + NBUF;
  static constexpr int OFF_ION = OFF_Y + (CARRY_LDS ? 2 * NF : 0);
  GV_MEM R *rec(int jj) const { return lds + (jj + 2) * RS; }
  GV_MEM R &st(int a, int jj) const { return rec(jj)[a]; }
  GV_MEM R &xb(int b, int jj) const { return rec(jj)[OFF_X + b]; }
  GV_MEM R &sc(int i) const { return lds[P * RS + i]; }
  GV_MEM int &ib(int jj) const {
    return reinterpret_cast<int *>(lds + P * RS + NSCAL)[jj + 2];
  }
  // RK4 carries of the owned level: the stage-1 state Y and the running
  // combination acc.  They are read and written only at the end of a stage,
  // so they live in LDS, or for the tallest columns in device memory
  // ([2*NF][L] per column, coalesced), instead of registers.
  GV_MEM R &Yv(int f) const { return CARRY_LDS ? rec(j)[OFF_Y + f] : scr[f * L + j]; }
  GV_MEM R &Av(int f) const {
    return CARRY_LDS ? rec(j)[OFF_Y + NF + f] : scr[(NF + f) * L + j];
  }
  // Ions: LogINS of level jj-1 (field NI holds IVel(up)) in LDS, so that the
  // boundary lanes can reach the top levels; their carries follow the
  // neutral ones in device memory.
  static constexpr int ION_SCR = CARRY_LDS ? 0 : 2 * NF;  // scratch rows before them
  GV_MEM R &ion(int f, int jj) const { return rec(jj)[OFF_ION + f]; }
  GV_MEM R &Yi(int f) const { return scr[(ION_SCR + f) * L + j]; }
  GV_MEM R &Ai(int f) const { return scr[(ION_SCR + NI + f) * L + j]; }

  // advance_vertical_1d_ausm (vertical_solver_ausm.f90:167-417) for one
  // field of an interior level: stage k+1 input = Y + (New - U), and the
  // final (1/3)*(-Y + S2 + 2*S3 + S4 + (New_4 - U_4)) accumulated in the
  // reference's left-to-right order.
  GV_MEM void ausm_commit(int stg, int f, R nv) {
    if (stg < 3) {
      const R y = Yv(f);
      const R nxt = y + (nv - U[f]);
      if (stg == 0)
        Av(f) = -1.0 * y + nxt;
      else if (stg == 1)
        Av(f) = Av(f) + 2.0 * nxt;
      else
        Av(f) = Av(f) + nxt;
      U[f] = nxt;
    } else {
      U[f] = (1.0 / 3.0) * (Av(f) + (nv - U[f]));
    }
  }

  // advance_vertical_1d_rusanov (vertical_solver_rusanov.f90:120-369):
  // K = New - U; Y + K/2, Y + K/2, Y + K, then Y + (K1 + 2K2 + 2K3 + K4)/6.
  GV_MEM void rusanov_commit(int stg, int f, R nv) {
    const R K = nv - U[f];
    const R y = Yv(f);
    if (stg == 0) {
      Av(f) = K;
      U[f] = y + 0.5 * K;
    } else if (stg == 1) {
      Av(f) = Av(f) + 2.0 * K;
      U[f] = y + 0.5 * K;
    } else if (stg == 2) {
      Av(f) = Av(f) + 2.0 * K;
      U[f] = y + K;
    } else {
      U[f] = y + (1.0 / 6.0) * (Av(f) + K);
    }
  }

  // The same two combinations for an ion field held in LDS.
  GV_MEM void ausm_commit_ion(int stg, int f, R nv) {
    R &u = ion(f, j);
    if (stg < 3) {
      const R y = Yi(f);
      const R nxt = y + (nv - u);
      if (stg == 0)
        Ai(f) = -1.0 * y + nxt;
      else if (stg == 1)
        Ai(f) = Ai(f) + 2.0 * nxt;
      else
        Ai(f) = Ai(f) + nxt;
      u = nxt;
    } else {
      u = (1.0 / 3.0) * (Ai(f) + (nv - u));
    }
  }
  GV_MEM void rusanov_commit_ion(int stg, int f, R nv) {
    R &u = ion(f, j);
    const R K = nv - u;
    const R y = Yi(f);
    if (stg == 0) {
      Ai(f) = K;
      u = y + 0.5 * K;
    } else if (stg == 1) {
      Ai(f) = Ai(f) + 2.0 * K;
      u = y + 0.5 * K;
    } else if (stg == 2) {
      Ai(f) = Ai(f) + 2.0 * K;
      u = y + K;
    } else {
      u = y + (1.0 / 6.0) * (Ai(f) + K);
    }
  }

  // calc_neutral_friction of the owned level, in the device or the
  // reference form (see neutral_friction).
  GV_MEM bool friction(R dtin, R (&w)[NSP], R nt, const R (&ns)[NSP],
                       const R (&glc)[NSP], R T) {
    R tp[6];
    if (cx.device_math()) {
      cx.diff_powers(T, -1.0, tp);
      if constexpr (ctx_taps<Ctx>::value) cx.tap_powers(tp, false);
      return neutral_friction<true>(kp, dtin, w, nt, ns, glc, st(A_EDDY, j), T, tp);
    }
    cx.diff_powers(T, 1.0, tp);
    if constexpr (ctx_taps<Ctx>::value) cx.tap_powers(tp, true);
    return neutral_friction<false>(kp, dtin, w, nt, ns, glc, st(A_EDDY, j), T, tp);
  }

  // a / b in the reference form; a * (1/b) in the device form, with the
  // reciprocal ib = 1/b shared by several quotients (one division in place
  // of several)
  GV_MEM R dq(R a, R b, R ib) const {
    return cx.device_math() ? a * ib : a / b;
  }

  GV_MEM bool face() const { return k >= 0 && k <= nA; }
  GV_MEM bool cell() const { return k >= 1 && k <= nA; }
  GV_MEM void fail(int code, int level) { cx.report(error_key(col, code, level)); }

  // -------------------------------------------------------------------------
  // Column prep: geometry and per-level constants (advance_vertical.f90
  // :84-151 plus the state-independent parts of the stage functions).
  // -------------------------------------------------------------------------
  GV_MEM void setup(R alt, R grav, R invr, R dalt, R mmm,
                    R gam, R eddy, R cmax, R tu, R coslat) {
    centrifugal = (coslat * kp.omega) * (coslat * kp.omega);  // :103
    coriolis = 2 * coslat * kp.omega;                         // :104
    S.eg = grav + centrifugal / invr;
    st(A_INVR, j) = invr;
    st(A_EDDY, j) = eddy;
    S.tu = tu;
    st(A_GRAV, j) = grav;
    S.gm1 = gam - 1.0;
    // VertTau(iAlt), UseDamping: vertical_solver_rusanov.f90:581-584
    vtau = kp.damping ? 15.0 - (1.0 - exp(-1.0 * alt / 1000.0 / 40.0)) * 5.0 : R(kp.vert_tau);
    const R rd = 1.0 / invr;  // RadialDistance_C (ausm :595-597)
    xb(0, j) = alt;
    xb(1, j) = S.eg;
    xb(2, j) = gam;
    xb(3, j) = cmax;
    xb(4, j) = rd;
    if (j == 0) sc(SC_MMM_M1) = mmm;
    if (j == 1) {
      sc(SC_MMM_0) = mmm;
      sc(SC_ALT0) = alt;
    }
    cx.sync();
    // dAlt_F(k) = Alt(k) - Alt(k-1), dAlt_F(-1) = dAlt_F(0)  (:149-151)
    const R dF = (j >= 1) ? alt - xb(0, j - 1) : xb(0, 1) - alt;
    st(A_DF, j) = dF;
    st(A_IF, j) = 1.0 / dF;
    st(A_DC, j) = dalt;
    // Rusanov dissipation speed of face k-1|k: 0.5*max(cMax(k-1), cMax(k))
    st(A_CMF, j) = 0.5 * dmax(xb(3, j - 1), cmax);
    S.mgh = (j >= 1) ? 0.5 * (S.eg + xb(1, j - 1)) : 0.0;
    if (j == 1) sc(SC_MG0) = S.mgh;
    if (j == nA + 2) sc(SC_MGT1) = S.mgh;
    if (j == nA + 3) sc(SC_MGT2) = S.mgh;
    // Kp/Ku blends, vertical_solver_ausm.f90:1320-1352 (same for all species)
    if (k >= 1 && k <= nA) {
      const R t = (1.0 + tanh((alt - 160.0e+03) / 15.0e+03));
      st(A_KP, j) = exp(kp.log_min_kp + 0.5 * (kp.log_max_kp - kp.log_min_kp) * t);
      st(A_KU, j) = exp(kp.log_min_ku + 0.5 * (kp.log_max_ku - kp.log_min_ku) * t);
    } else {
      st(A_KP, j) = 0.0;
      st(A_KU, j) = 0.0;
    }
    // spherical areas and volume, vertical_solver_ausm.f90:2029-2040
    {
      const R LR = 0.5 * (rd + xb(4, j - 1));
      const R RR = 0.5 * (rd + xb(4, j + 1));
      st(A_AP, j) = RR * RR;
      st(A_AM, j) = LR * LR;
      st(A_VOL, j) = (1.0 / 3.0) * (RR * RR * RR - LR * LR * LR);
    }
    st(A_HF, j) = 0.35 * exp(-1.0 * (alt - sc(SC_ALT0)) / (45.0e+03));
    cx.sync();
    // gamma face states (calc_kt_facevalues of Gamma_1d, ausm :1535)
    R dlg = 0.0;
    if (k >= 0 && k <= nA + 1) {
      const R up = (xb(2, j + 1) - gam) * st(A_IF, j + 1), dn = (gam - xb(2, j - 1)) * st(A_IF, j);
      dlg = limiter_mc(kp.beta, up, dn);
      if constexpr (ctx_taps<Ctx>::value) cx.tap_gamma_code(limiter_code(kp.beta, up, dn));
    }
    xb(5, j) = dlg;
    // 5-point gradient weights (ausm :649-676)
    if (k >= 1 && k <= nA) {
      const R h1 = st(A_DF, j - 1), h2 = dF, h3 = st(A_DF, j + 1),
                   h4 = st(A_DF, j + 2);
      const R MeshH2 = h2 + h1;
      const R MeshH3 = h3 + h2 + h1;
      const R MeshH4 = h4 + h3 + h2 + h1;
      st(A_GLC0 + 0, j) = (h2 * h3 * (h3 + h4)) / (h1 * MeshH2 * MeshH3 * MeshH4);
      st(A_GLC0 + 1, j) = -1.0 * (MeshH2 * h3 * (h3 + h4)) / (h1 * h2 * (h2 + h3) * (h2 + h3 + h4));
      st(A_GLC0 + 3, j) = MeshH2 * h2 * (h4 + h3) / (MeshH3 * (h2 + h3) * h3 * h4);
      st(A_GLC0 + 4, j) = -1.0 * MeshH2 * h2 * h3 / (MeshH4 * (h2 + h3 + h4) * (h3 + h4) * h4);
      st(A_GLC0 + 2, j) = (h2 * h3 * (h3 + h4) + MeshH2 * h3 * (h3 + h4) -
                  MeshH2 * h2 * (h3 + h4) - MeshH2 * h2 * h3) /
                 (MeshH2 * h2 * h3 * (h3 + h4));
    } else {
      for (int q = 0; q < 5; q++) st(A_GLC0 + q, j) = 0.0;
    }
    // bottom extrapolation weights for levels 0 and -1 (bc :228-248)
    if (j == 0) {
      bc_fwd_coefs(st(A_DF, 3), st(A_DF, 4), st(A_DF, 5), st(A_DF, 6), &sc(SC_CA));
      bc_fwd_coefs(st(A_DF, 2), st(A_DF, 3), st(A_DF, 4), st(A_DF, 5), &sc(SC_CB));
    }
    if (IONS && j == 1) {
      // top ion slopes for levels nAlts+1 and nAlts+2 (bc :505-525)
      bc_bwd_coefs(st(A_DF, nA + 1), st(A_DF, nA), st(A_DF, nA - 1), st(A_DF, nA - 2),
                   &sc(SC_CTA));
      bc_bwd_coefs(st(A_DF, nA + 2), st(A_DF, nA + 1), st(A_DF, nA), st(A_DF, nA - 1),
                   &sc(SC_CTB));
    }
    cx.sync();
    if (face()) {
      const R gL = gam + 0.5 * dlg * dF;
      const R gR = xb(2, j + 1) - 0.5 * xb(5, j + 1) * st(A_DF, j + 1);
      st(A_GL1, j) = 1.0 / (gL - 1.0);
      st(A_GL2, j) = gL / (gL - 1.0);
      st(A_GL3, j) = 2.0 * ((gL - 1.0) / (gL + 1.0));
      st(A_GR1, j) = 1.0 / (gR - 1.0);
      st(A_GR2, j) = gR / (gR - 1.0);
      st(A_GR3, j) = 2.0 * ((gR - 1.0) / (gR + 1.0));
    } else {
      st(A_GL1, j) = st(A_GL2, j) = st(A_GL3, j) = st(A_GR1, j) = st(A_GR2, j) = st(A_GR3, j) = 0.0;
    }
    cx.sync();
  }

  // -------------------------------------------------------------------------
  // set_vertical_bcs, set_vertical_bcs.Earth.f90:8-652 (no BC perturbation).
  // Without MSIS and tides, ghost values depend only on the interior and on
  // Temp(0) and LogNS(0) of the non-photochemical species, which the BCs
  // never change.  `first`: the step's first pass (before the RK stages).
  // -------------------------------------------------------------------------
  GV_MEM void bcs(bool first = false) {
    // UseMSIS (set_vertical_bcs.Earth.f90:113-137): Temp(0) and the horizontal
    // winds of level 0 take the MSIS values before anything reads them; the
    // MSIS LogNS(0) replace the non-photochemical densities below, after the
    // hydrostatic fill of level -1 has used the entry values (NS, :93).
    // Tides (:150-159) then set the winds of level 0 (held in msis_ve/vn)
    // and GSWM adds TidesTemp to Temp(0), WACCM replaces it.  The drivers
    // reset the ghost levels to their values after the first pass before
    // every later pass (the stages leave them alone, vertical_solver_ausm.f90
    // :163-377, vertical_solver_rusanov.f90:120-369), so GSWM adds to that
    // Temp(0) every time; msis_t keeps it when MSIS does not set Temp(0).
    if (j == 1) {
      if (kp.msis || (kp.tides == 1 && !first)) U[F_T] = msis_t;
      if (kp.msis || kp.tides) {
        U[F_VE] = msis_ve;
        U[F_VN] = msis_vn;
      }
      if (kp.tides == 1) {
        U[F_T] = tide_t + U[F_T];
        if (first && !kp.msis) msis_t = U[F_T];
      } else if (kp.tides == 2) {
        U[F_T] = tide_t;
      }
    }
    // stage ln(6), w(6), T, vE, vN, vU
#pragma unroll
    for (int f = 0; f < 12; f++) xb(f, j) = U[F_LN + f];
    xb(12, j) = U[F_T];
    xb(13, j) = U[F_VE];
    xb(14, j) = U[F_VN];
    xb(15, j) = U[F_VU];
    cx.sync();
    if (j < NSP) {
      // bottom, one species per thread (:198-272, :277-344, :382-395)
      const int sp = j;
      R *CA = &sc(SC_CA), *CB = &sc(SC_CB);
      const R dF1 = st(A_DF, 2), dF0 = st(A_DF, 1), dFm1 = st(A_DF, 0);
      R ln0 = xb(sp, 1), ln1 = xb(sp, 2), ln2 = xb(sp, 3), ln3 = xb(sp, 4),
             ln4 = xb(sp, 5), ln5 = xb(sp, 6);
      R n0, nm1;
      if (photochemical(sp)) {
        R d = ln1 * CA[0] + ln2 * CA[1] + ln3 * CA[2] + ln4 * CA[3] + ln5 * CA[4];
        n0 = ln1 - dF1 * d;
        if (d <= 0.0) n0 = ln1;
        if (sp == S_NO && n0 < ln1 - 0.333) n0 = ln1 - 0.333;
        d = n0 * CB[0] + ln1 * CB[1] + ln2 * CB[2] + ln3 * CB[3] + ln4 * CB[4];
        nm1 = n0 - dF0 * d;
        if (d <= 0.0) nm1 = n0;
        if (sp == S_NO && nm1 < n0 - 0.333) nm1 = n0 - 0.333;
      } else {
        n0 = kp.msis ? msis_ln : ln0;  // lane sp holds species sp's value
        if (sp == S_O) {
          nm1 = n0;  // zero gradient below the boundary (:219-222)
        } else {
          const R t0 = xb(12, 1), tm1 = t0;  // Temp(-1) = Temp(0)
          const R MeanGravity = -sc(SC_MG0);
          const R MeanTemp = 0.5 * (t0 + tm1);
          const R MeanMass = 0.5 * (sc(SC_MMM_0) + sc(SC_MMM_M1));
          const R ish = MeanGravity * MeanMass / (MeanTemp * KB);
          const R ns = exp(ln0) * (t0 / tm1) * exp(1.0 * ish * dFm1);
          nm1 = ln(ns);
        }
      }
      const R w1 = xb(F_W - F_LN + sp, 2), w2 = xb(F_W - F_LN + sp, 3),
                   w3 = xb(F_W - F_LN + sp, 4), w4 = xb(F_W - F_LN + sp, 5),
                   w5 = xb(F_W - F_LN + sp, 6);
      const R w0 = w1 - dF1 * (w1 * CA[0] + w2 * CA[1] + w3 * CA[2] + w4 * CA[3] + w5 * CA[4]);
      const R wm1 = w0 - dF0 * (w0 * CB[0] + w1 * CB[1] + w2 * CB[2] + w3 * CB[3] + w4 * CB[4]);
      R w0r = w0, wm1r = wm1;
      if (photochemical(sp) && w1 > 0.0) {
        w0r = -1.0 * w1;
        wm1r = -1.0 * w1;
      }
      R *Rb = &sc(SC_BOT + 6 * sp);
      Rb[0] = n0;
      Rb[1] = nm1;
      Rb[2] = w0;
      Rb[3] = wm1;
      Rb[4] = w0r;
      Rb[5] = wm1r;
    } else if (j >= L - NSP && j < L) {
      // top, one species per thread (:429-441, :471-486), on the column's
      // last lanes: in a column of two or more wavefronts the top and the
      // bottom run side by side in different wavefronts
      const int sp = L - 1 - j;
      const int jn = nA + 1;  // offset of level nAlts
      const R tn = xb(12, jn);
      const R wn = xb(F_W - F_LN + sp, jn), wnm1 = xb(F_W - F_LN + sp, jn - 1);
      R w1 = wn, w2 = wn;
      if (photochemical(sp) && wn < 0.0) {
        w1 = -wn;
        w2 = -wnm1;
      }
      const R m = kp.mass[sp];
      R ns = exp(xb(sp, jn));
      R MeanGravity = -sc(SC_MGT1);
      R MeanTemp = 0.5 * (tn + tn);
      R ish = MeanGravity * m / (MeanTemp * KB);
      ns = ns * (tn / tn) * exp(-ish * st(A_DF, jn + 1));
      const R l1 = ln(ns);
      MeanGravity = -sc(SC_MGT2);
      ish = MeanGravity * m / (MeanTemp * KB);
      ns = ns * (tn / tn) * exp(-ish * st(A_DF, jn + 2));
      const R l2 = ln(ns);
      R *Rb = &sc(SC_TOP + 4 * sp);
      Rb[0] = l1;
      Rb[1] = l2;
      Rb[2] = w1;
      Rb[3] = w2;
    }
    if (IONS) {
      // bottom: zero-gradient ions (:277-281)
      if (j <= 1) {
#pragma unroll
        for (int f = 0; f < NI; f++) ion(f, j) = ion(f, 2);
      }
      // TEC of the column in 1e16 m^-2 (:498-503), in the reference's order
      if (j == 6 && plasma) {
        R tec = 0.0;
        for (int jj = 2; jj <= nA + 1; jj++) {
          const R v = kp.improved_ions ? ion(0, jj) : exp(ion(0, jj));
          tec = tec + st(A_DF, jj) * v;
        }
        sc(SC_TEC) = tec / 1e16;
      }
    }
    cx.sync();
    if (j <= 1) {
      // levels -1 and 0: ghost species, LogRho and bulk wind (:346-362)
      const int o = (j == 1) ? 0 : 1;
      R sr = 0.0, ns[NSP];
#pragma unroll
      for (int sp = 0; sp < NSP; sp++) {
        const R *Rb = &sc(SC_BOT + 6 * sp);
        U[F_LN + sp] = Rb[o];
        U[F_W + sp] = Rb[4 + o];
        ns[sp] = exp(Rb[o]);
        sr = sr + kp.mass[sp] * ns[sp];
      }
      U[F_LR] = ln(sr);
      const R er = exp(U[F_LR]);
      R vu = 0.0;
#pragma unroll
      for (int sp = 0; sp < NSP; sp++)
        vu = vu + ns[sp] * kp.mass[sp] * sc(SC_BOT + 6 * sp + 2 + o) / er;
      U[F_VU] = vu;
      if (!kp.msis) {
        if (!kp.tides) {  // :139-142
          U[F_VE] = 0.0;
          U[F_VN] = 0.0;
        } else if (j == 0) {  // the tides' level 0 (:150-159), Vel(-1) = Vel(0) (:192-196)
          U[F_VE] = xb(13, 1);
          U[F_VN] = xb(14, 1);
        }
      } else if (j == 0) {
        // Vel(-1) east and north extrapolated from levels 0..4 (:364-375)
        const R *CB = &sc(SC_CB);
        const R dF0 = st(A_DF, 1);
#pragma unroll
        for (int d = 0; d < 2; d++) {
          const R v0 = xb(13 + d, 1), v1 = xb(13 + d, 2), v2 = xb(13 + d, 3),
                  v3 = xb(13 + d, 4), v4 = xb(13 + d, 5);
          const R dv = v0 * CB[0] + v1 * CB[1] + v2 * CB[2] + v3 * CB[3] + v4 * CB[4];
          U[F_VE + d] = v0 - dF0 * dv;
        }
      }
      if (j == 0) U[F_T] = xb(12, 1);
    } else if (j == nA + 2 || j == nA + 3) {
      // levels nAlts+1, nAlts+2 (:420-463, :598-604)
      const int o = (j == nA + 2) ? 0 : 1;
      const int jn = nA + 1;
      R sr = 0.0;
#pragma unroll
      for (int sp = 0; sp < NSP; sp++) {
        const R *Rb = &sc(SC_TOP + 4 * sp);
        U[F_LN + sp] = Rb[o];
        U[F_W + sp] = Rb[2 + o];
        sr = sr + kp.mass[sp] * exp(Rb[o]);
      }
      U[F_LR] = ln(sr);
      U[F_VE] = xb(13, jn);
      U[F_VN] = xb(14, jn);
      const R vu = xb(15, jn);
      U[F_VU] = (vu < 0.0) ? -vu : vu;
      U[F_T] = xb(12, jn);
    }
    if (IONS && j < NI) ion_top_bc(j);
    cx.sync();
  }

  // Top ghost levels of one ion (set_vertical_bcs.Earth.f90:505-596),
  // levels nAlts+1 then nAlts+2.  With UseImprovedIonAdvection the state
  // holds densities: negative values near the top are repaired first
  // (check_for_negative_densities, :608-649), and the slope is taken in log
  // space.  The slope is limited to a decay of at least 0.1% per dAlt unless
  // the column is a thin nighttime plasmasphere.
  GV_MEM void ion_top_bc(int q) {
    const bool improved = kp.improved_ions != 0;
    const bool keep = plasma && sc(SC_TEC) < kp.min_tec;
    const R dFn = st(A_DF, nA + 1);  // dAlt_F(nAlts)
#pragma unroll 1
    for (int t = 0; t < 2; t++) {
      const int jk = nA + 2 + t;  // lane of level nAlts+1+t
      const R *C = &sc(t ? SC_CTB : SC_CTA);
      R n0, n1, n2, n3, n4, n5;
      if (improved) {
        R mn = ion(q, jk - 5);
        for (int jj = jk - 4; jj <= jk; jj++)
          if (ion(q, jj) < mn) mn = ion(q, jj);
        if (mn < 0.0) {
          if (ion(q, jk - 6) < 0.0) fail(GITM_VERT_ENEGION, jk - 1);
          for (int jj = jk - 5; jj <= jk - 1; jj++) {
            if (ion(q, jj) < 0.0) {
              if (ion(q, jj - 1) > 0.0 && ion(q, jj + 1) > 0.0)
                ion(q, jj) = (ion(q, jj - 1) + ion(q, jj + 1)) / 2.0;
              else
                ion(q, jj) = ion(q, jj - 1) * 0.95;
            }
          }
          if (ion(q, jk) < 0.0) ion(q, jk) = ion(q, jk - 1) * 0.95;
        }
        n0 = ln(ion(q, jk));
        n1 = ln(ion(q, jk - 1));
        n2 = ln(ion(q, jk - 2));
        n3 = ln(ion(q, jk - 3));
        n4 = ln(ion(q, jk - 4));
        n5 = ln(ion(q, jk - 5));
      } else {
        n0 = ion(q, jk);
        n1 = ion(q, jk - 1);
        n2 = ion(q, jk - 2);
        n3 = ion(q, jk - 3);
        n4 = ion(q, jk - 4);
        n5 = ion(q, jk - 5);
      }
      R dn = C[0] * n1 + C[1] * n2 + C[2] * n3 + C[3] * n4 + C[4] * n5;
      if (!keep) {
        if (dn > -0.001 * n0 / dFn) dn = -0.001 * n0 / dFn;
        n0 = n1 + dn * st(A_DF, jk);
        ion(q, jk) = improved ? exp(n0) : n0;
      }
    }
  }

  // -------------------------------------------------------------------------
  // 4th-order limited face values and Rusanov gradient / dissipation of up to
  // 8 fields (calc_rusanov_alts_*, vertical_solver_ausm.f90:1007-1090).
  // Uses exchange buffers 0..15.  grad/diff are valid on cell threads.
  // -------------------------------------------------------------------------
  template <int N>
  GV_MEM void rusanov_grad_diff(const R (&v)[N], R (&grad)[N],
                                R (&diff)[N]) {
    static_assert(N <= 8, "at most 8 fields per pass");
#pragma unroll
    for (int f = 0; f < N; f++) xb(f, j) = v[f];
    cx.sync();
    const R iF0 = st(A_IF, j), iF1 = st(A_IF, j + 1);
    const R dF0 = st(A_DF, j), dF1 = st(A_DF, j + 1);
    R rf[N], ln[N];
#pragma unroll
    for (int f = 0; f < N; f++) {
      const R vm2 = xb(f, j - 2), vm1 = xb(f, j - 1), v0 = v[f],
                   v1 = xb(f, j + 1), v2 = xb(f, j + 2);
      R dl = 0.0;
      if (k >= 1 && k <= nA) {
        R h = iF1 * 2.0;
        const R up = h * (FACT1 * (v1 - v0) - FACT2 * (v2 - vm1));
        h = iF0 * 2.0;
        const R dn = h * (FACT1 * (v0 - vm1) - FACT2 * (v1 - vm2));
        dl = limiter_mc(kp.beta, up, dn);
      } else if (k == 0 || k == nA + 1) {
        dl = limiter_mc(kp.beta, (v1 - v0) * iF1, (v0 - vm1) * iF0);
      }
      rf[f] = v0 - 0.5 * dl * dF0;  // right state of face k-1|k
      ln[f] = v0 + 0.5 * dl * dF1;  // left state of face k|k+1
    }
    cx.sync();
#pragma unroll
    for (int f = 0; f < N; f++) {
      xb(f, j) = rf[f];
      xb(8 + f, j) = ln[f];
    }
    cx.sync();
    const R dC = st(A_DC, j);
    const R cm0 = st(A_CMF, j), cm1 = st(A_CMF, j + 1);
#pragma unroll
    for (int f = 0; f < N; f++) {
      const R L0 = xb(8 + f, j - 1), R0 = rf[f], L1 = ln[f], R1 = xb(f, j + 1);
      grad[f] = 0.5 * (L1 + R1 - L0 - R0) / dC;
      const R df0 = cm0 * (R0 - L0), df1 = cm1 * (R1 - L1);
      diff[f] = (df1 - df0) / dC;
    }
    cx.sync();
  }

  // 5-point gradient of ln(rho_s/rho) (vertical_solver_ausm.f90:649-688)
  GV_MEM void grad_log_cons(const R (&lcs)[NSP], R (&glc)[NSP]) {
#pragma unroll
    for (int f = 0; f < NSP; f++) xb(f, j) = lcs[f];
    cx.sync();
#pragma unroll
    for (int f = 0; f < NSP; f++)
      glc[f] = st(A_GLC0 + 0, j) * xb(f, j - 2) + st(A_GLC0 + 1, j) * xb(f, j - 1) + st(A_GLC0 + 2, j) * lcs[f] +
               st(A_GLC0 + 3, j) * xb(f, j + 1) + st(A_GLC0 + 4, j) * xb(f, j + 2);
    cx.sync();
  }

  // -------------------------------------------------------------------------
  // advance_vertical_1stage_ausm, vertical_solver_ausm.f90:436-999, with
  // calc_all_fluxes_hydro (:1093-2070).  Writes New (valid on cell threads).
  // Ordered so that few per-level values stay live across the flux phase.
  // -------------------------------------------------------------------------
  // The hydrostatic densities of the AUSM stage (ausm :725-751) in the
  // device form: NS(level 1) times the running product of the per-level
  // factors a(k) b_s(k) (xb(1), xb(2+s)), split into G segments of `seg`
  // levels so that the serial chain is `seg` levels long instead of nAlts+2.
  // Lane (g, s) forms the products within segment g (from NS(level 1) for
  // g = 0, from 1 otherwise); lanes s then make the segment ends absolute
  // in order, and every other level of segments 1.. takes the end of the
  // segment below.  The products associate differently from the level-by-
  // level order of the reference (one ulp per multiplication at most, far
  // below the step's tolerance); the host build's reference forms keep that
  // order.  Reads and writes of the three phases are separated by barriers;
  // within a phase each slot has one writer and no other reader.
  GV_MEM void hydro_segmented() {
    const int G = (L / NSP) < 8 ? L / NSP : 8;  // >= 2: L >= 14 (nAlts >= 10)
    const int seg = (L - 2 + G - 1) / G;
    if (j < G * NSP) {
      const int g = j / NSP, s = j - g * NSP;
      const int lo = 2 + g * seg, hi = (lo + seg < L) ? lo + seg : L;
      R x = (g == 0) ? R(xb(8 + s, 1)) : R(1.0);
      for (int jj = lo; jj < hi; jj++) {
        x = x * xb(1, jj) * xb(2 + s, jj);
        xb(8 + s, jj) = x;
      }
      if (g == 0) xb(8 + s, 0) = xb(8 + s, 1) * xb(1, 0) * xb(2 + s, 0);
    }
    cx.sync();
    if (j < NSP) {
      const int e0 = (2 + seg < L ? 2 + seg : L) - 1;
      R x = xb(8 + j, e0);
      for (int g = 1; g < G && 2 + g * seg < L; g++) {
        const int e = (2 + (g + 1) * seg < L ? 2 + (g + 1) * seg : L) - 1;
        x = x * xb(8 + j, e);
        xb(8 + j, e) = x;
      }
    }
    cx.sync();
    if (lv && j >= 2 + seg) {
      const int g = (j - 2) / seg;
      const int below = 2 + g * seg - 1;
      const int end = (2 + (g + 1) * seg < L ? 2 + (g + 1) * seg : L) - 1;
      if (j != end) {
#pragma unroll
        for (int s = 0; s < NSP; s++) xb(8 + s, j) = xb(8 + s, below) * xb(8 + s, j);
      }
    }
  }

  GV_MEM void ausm_stage(R dtin, int stg) {
    const double *m = kp.mass;
    const R T = U[F_T];
    // --- pointwise (:591-609, :649-652, :752-773)
    R NS[NSP], rhos[NSP];
    const R rho = exp(U[F_LR]);
    R nt = 0.0;
#pragma unroll
    for (int s = 0; s < NSP; s++) {
      NS[s] = exp(U[F_LN + s]);
      nt = nt + NS[s];
    }
#pragma unroll
    for (int s = 0; s < NSP; s++) rhos[s] = m[s] * NS[s];
    const R te = nt * KB * T / S.gm1 +
                      0.5 * rho *
                          (U[F_VU] * U[F_VU] + U[F_VN] * U[F_VN] + U[F_VE] * U[F_VE]);

    // --- gradients the sources need (:611-646, :649-688): GradTemp,
    // DiffVel(E,N) and d ln(rho_s/rho)/dz
    R glc[NSP];
    {
      R lcs[NSP];
#pragma unroll
      for (int s = 0; s < NSP; s++)
        lcs[s] = cx.device_math() ? (kp.lmass[s] + U[F_LN + s]) - U[F_LR]
                                  : ln(m[s] * NS[s] / rho);
      grad_log_cons(lcs, glc);
    }
    R gr3[3], df3[3];
    {
      const R gv3[3] = {T, U[F_VE], U[F_VN]};
      rusanov_grad_diff<3>(gv3, gr3, df3);
    }
    if constexpr (ctx_taps<Ctx>::value) cx.tap_grads(glc, gr3[0], df3[1], df3[2]);
    if (IONS) ion_stage(dtin, stg, true);  // :817-829

    // --- hydrostatic background (:694-751): per-level factors in parallel,
    // the recurrence in altitude serially (one thread per species)
    xb(0, j) = T;
#pragma unroll
    for (int s = 0; s < NSP; s++) xb(8 + s, j) = NS[s];
    cx.sync();
    {
      const R tm1 = xb(0, j - 1), tp1 = xb(0, j + 1);
      const R dF = st(A_DF, j);
      if (lv && j >= 2) {
        xb(1, j) = tm1 / T;
        const R MeanTemp = 0.5 * (T + tm1);
        // the device form divides once for the six species
        const bool shared = is_double<R>::value && cx.device_math();
        const R ikt = shared ? R(1.0 / (KB * MeanTemp)) : R(0.0);
#pragma unroll
        for (int s = 0; s < NSP; s++) {
          const R ish = shared ? -1.0 * m[s] * S.mgh * ikt : -1.0 * m[s] * S.mgh / (KB * MeanTemp);
          xb(2 + s, j) = exp(-1.0 * dF * ish);
        }
      } else if (j == 0) {
        xb(1, j) = tp1 / T;
        const R MeanMass = 0.5 * (sc(SC_MMM_M1) + sc(SC_MMM_0));
        const R MeanTemp = 0.5 * (tp1 + T);
        const R ish = -1.0 * MeanMass * sc(SC_MG0) / (KB * MeanTemp);
        const R b = exp(dF * ish);
#pragma unroll
        for (int s = 0; s < NSP; s++) xb(2 + s, j) = b;
      }
    }
    cx.sync();
    if (is_double<R>::value && cx.device_math()) {
      hydro_segmented();
    } else if (j < NSP) {
      const int s = j;
      const R ns0 = xb(8 + s, 1);
      R x = ns0;
      // The product runs level by level in the reference's order; the
      // factors of the next CH levels are read while the current ones are
      // multiplied, so that the LDS latency is not paid at every level.
      constexpr int CH = 8;
      const int nch = (L - 2) / CH;
      int jj = 2;
      R a[CH], b[CH];
#pragma unroll
      for (int u = 0; u < CH; u++) {
        a[u] = xb(1, jj + u);
        b[u] = xb(2 + s, jj + u);
      }
      for (int c = 0; c < nch; c++, jj += CH) {
        R an[CH], bn[CH];
        const bool more = c + 1 < nch;
        if (more) {
#pragma unroll
          for (int u = 0; u < CH; u++) {
            an[u] = xb(1, jj + CH + u);
            bn[u] = xb(2 + s, jj + CH + u);
          }
        }
#pragma unroll
        for (int u = 0; u < CH; u++) {
          x = x * a[u] * b[u];
          xb(8 + s, jj + u) = x;
        }
        if (more) {
#pragma unroll
          for (int u = 0; u < CH; u++) {
            a[u] = an[u];
            b[u] = bn[u];
          }
        }
      }
      for (; jj < L; jj++) {
        x = x * xb(1, jj) * xb(2 + s, jj);
        xb(8 + s, jj) = x;
      }
      xb(8 + s, 0) = ns0 * xb(1, 0) * xb(2 + s, 0);
    }
    cx.sync();
    R dev[NSP];
    int sub = 0;  // bit s: SubtractHydrostatic(k, s)
    {
      R lhps[NSP], hnsv[NSP];
#pragma unroll
      for (int s = 0; s < NSP; s++) {
        const R hns = xb(8 + s, j);
        hnsv[s] = hns;
        lhps[s] = ln(hns * KB * T);
        const R hrhos = m[s] * hns;
        const R ratio = fabs(rhos[s] - hrhos) / rhos[s];
        dev[s] = rhos[s] - hrhos;
        if (!(fabs(ratio) > 1.0)) sub |= 1 << s;
      }
      if constexpr (ctx_taps<Ctx>::value) cx.tap_hydro(hnsv, sub);
      cx.sync();
      // batch 2 of the face states goes first: ln p_hydro,s and w_s
#pragma unroll
      for (int s = 0; s < NSP; s++) {
        xb(s, j) = lhps[s];
        xb(6 + s, j) = U[F_W + s];
      }
    }
    ib(j) = sub;
    cx.sync();

    // --- face states (calc_kt_facevalues, :2073-2117) of face k|k+1, one
    // face per thread, exponentiated where the flux needs them (:1356-1436)
    const int subUp = ib(j + 1);  // flags of cell k+1
    R RSL[NSP], RSR[NSP], PSL[NSP], PSR[NSP], HL[NSP], HR[NSP], VL[NSP], VR[NSP];
    R VEL, VER, VNL, VNR;
    {
      const R iF0 = st(A_IF, j), iF1 = st(A_IF, j + 1), iF2 = st(A_IF, j + 2);
      const R dF0 = st(A_DF, j), dF1 = st(A_DF, j + 1);
      // kt: the limiter choices of this cell's slopes (adjoint build only)
      unsigned long long kt = 0;
      (void)kt;
#define GV_KT(b, F, OUTL, OUTR)                                                     \
  do {                                                                              \
    const R f0 = xb(b, j), fm1 = xb(b, j - 1), f1 = xb(b, j + 1), f2 = xb(b, j + 2); \
    const R up0 = (f1 - f0) * iF1, dn0 = (f0 - fm1) * iF0;                          \
    const R d0 = limiter_mc(kp.beta, up0, dn0);                                     \
    const R d1 = limiter_mc(kp.beta, (f2 - f1) * iF2, (f1 - f0) * iF1);            \
    if constexpr (ctx_taps<Ctx>::value)                                             \
      kt |= (unsigned long long)limiter_code(kp.beta, up0, dn0) << (2 * (F));       \
    OUTL = f0 + 0.5 * d0 * dF0;                                                     \
    OUTR = f1 - 0.5 * d1 * dF1;                                                     \
  } while (0)
#pragma unroll
      for (int s = 0; s < NSP; s++) {
        R l, r;
        GV_KT(s, s, l, r);
        HL[s] = exp(l);
        HR[s] = exp(r);
        GV_KT(6 + s, 6 + s, VL[s], VR[s]);
      }
      cx.sync();
      if (cx.device_math()) {
        const R ltk = ln(T * KB);
#pragma unroll
        for (int s = 0; s < NSP; s++) {
          xb(s, j) = kp.lmass[s] + U[F_LN + s];
          xb(6 + s, j) = U[F_LN + s] + ltk;
        }
      } else {
#pragma unroll
        for (int s = 0; s < NSP; s++) {
          xb(s, j) = ln(rhos[s]);
          xb(6 + s, j) = ln(NS[s] * T * KB);
        }
      }
      xb(12, j) = U[F_VE];
      xb(13, j) = U[F_VN];
      cx.sync();
#pragma unroll
      for (int s = 0; s < NSP; s++) {
        R l, r;
        GV_KT(s, 12 + s, l, r);
        RSL[s] = exp(l);
        RSR[s] = exp(r);
        GV_KT(6 + s, 18 + s, l, r);
        PSL[s] = exp(l);
        PSR[s] = exp(r);
      }
      GV_KT(12, 24, VEL, VER);
      GV_KT(13, 25, VNL, VNR);
#undef GV_KT
      if constexpr (ctx_taps<Ctx>::value)
        cx.tap_faces(HL, HR, VL, VR, RSL, RSR, PSL, PSR, VEL, VER, VNL, VNR, kt);
    }
    cx.sync();

    // --- AUSM+-up fluxes of face k|k+1 (:1459-1970).  The P12 side belongs
    // to cell k and stays in registers; the M12 side belongs to cell k+1 (with
    // that cell's Kp, Ku, dAlt_C and SubtractHydrostatic flags) and goes to
    // LDS buffers 0..20 for the thread above.
    R FRp[NSP], FMp[NSP], NPp[NSP];
    R EFp = 0.0, MOEp = 0.0, MONp = 0.0;
#pragma unroll
    for (int s = 0; s < NSP; s++) FRp[s] = FMp[s] = NPp[s] = 0.0;
    if (face()) {
      R RL = 0.0, RR = 0.0, PL = 0.0, PR = 0.0;
#pragma unroll
      for (int s = 0; s < NSP; s++) {
        RL = RL + RSL[s];
        RR = RR + RSR[s];
      }
#pragma unroll
      for (int s = 0; s < NSP; s++) {
        PL = PL + PSL[s];
        PR = PR + PSR[s];
      }
      const R iRL = 1.0 / RL, iRR = 1.0 / RR;
      R VUL = 0.0, VUR = 0.0;  // mass-weighted face winds (:1505-1532)
#pragma unroll
      for (int s = 0; s < NSP; s++) {
        VUL = VUL + dq(RSL[s] * VL[s], RL, iRL);
        VUR = VUR + dq(RSR[s] * VR[s], RR, iRR);
      }
      // bulk energy (:1538-1562), enthalpy (:1569-1615), Liou sound speed
      // (:1619-1655)
      const R EL = st(A_GL1, j) * PL + 0.5 * RL * (VUL * VUL + VEL * VEL + VNL * VNL);
      const R ER = st(A_GR1, j) * PR + 0.5 * RR * (VUR * VUR + VER * VER + VNR * VNR);
      R CS;
      {
        const R HtL =
            0.5 * (VUL * VUL + VEL * VEL + VNL * VNL) + dq(st(A_GL2, j) * PL, RL, iRL);
        const R HtR =
            0.5 * (VUR * VUR + VER * VER + VNR * VNR) + dq(st(A_GR2, j) * PR, RR, iRR);
        R sub_cs = sqrt(st(A_GL3, j) * HtL);
        const R CSL = (sub_cs * sub_cs) / dmax(sub_cs, VUL);
        sub_cs = sqrt(st(A_GR3, j) * HtR);
        const R CSR = (sub_cs * sub_cs) / dmax(sub_cs, -1.0 * VUR);
        CS = dmin(CSL, CSR);
      }
      const R kp0 = st(A_KP, j), kp1 = st(A_KP, j + 1);
      const R ku0 = st(A_KU, j), ku1 = st(A_KU, j + 1);
      const R rdt0 = st(A_DC, j) / dtin, rdt1 = st(A_DC, j + 1) / dtin;
      R BNVp = 0.0, BNVm = 0.0;
      R nvp[NSP], nvm[NSP];  // the numerical velocities (adjoint build only)
      (void)nvp;
      (void)nvm;
      const R RLR = RL + RR;
      const R iRLR = 1.0 / RLR, icS = 1.0 / CS;
#pragma unroll
      for (int s = 0; s < NSP; s++) {
        const R MPS = 0.5 * (PSL[s] + PSR[s]);
        const R MHPS = 0.5 * (HL[s] + HR[s]);
        const R MRS = 0.5 * (RSL[s] + RSR[s]);
        const R ML = dq(VL[s], CS, icS), MR = dq(VR[s], CS, icS);  // :1689-1706
        const R M2 = 0.5 * (ML * ML + MR * MR);
        const R MZ = sqrt(dmin(1.0, dmax(M2, 1.0e-19)));
        const R FA = MZ * (2.0 - MZ);
        R MF4P, MF4N;  // :1789-1819 (16*LiouBeta = 2)
        if (fabs(ML) >= 1.0)
          MF4P = 0.5 * (ML + fabs(ML));
        else
          MF4P = 0.25 * ((ML + 1.0) * (ML + 1.0)) *
                 (1.0 - 2.0 * (-0.25 * ((ML - 1.0) * (ML - 1.0))));
        if (fabs(MR) >= 1.0)
          MF4N = 0.5 * (MR - fabs(MR));
        else
          MF4N = -0.25 * ((MR - 1.0) * (MR - 1.0)) *
                 (1.0 + 2.0 * (0.25 * ((MR + 1.0) * (MR + 1.0))));
        const R f1 = dmax((1.0 - M2), 0.0);
        const R pdh = (PSR[s] - HR[s]) - (PSL[s] - HL[s]);
        const R pd = PSR[s] - PSL[s];
        const R rvd = (RSR[s]) * VR[s] - (RSL[s]) * VL[s];
        const R mcs = MRS * CS, facs = FA * CS;
        {  // P12 side of cell k (:1822-1970)
          const bool sh = (sub >> s) & 1;
          const R MP = kp0 * f1 * (sh ? pdh : pd) / (mcs * (facs + rdt0));
          const R NV = CS * (MF4P + MF4N - MP);
          nvp[s] = NV;
          NPp[s] = (sh ? (MPS - MHPS) : MPS) - 0.5 * ku0 * CS * rvd;
          if (NV >= 0.0) {
            FRp[s] = RSL[s] * NV;
            FMp[s] = RSL[s] * VL[s] * NV;
          } else {
            FRp[s] = RSR[s] * NV;
            FMp[s] = RSR[s] * VR[s] * NV;
          }
          BNVp = BNVp + dq((RSL[s] + RSR[s]) * NV, RLR, iRLR);
        }
        {  // M12 side of cell k+1
          const bool sh = (subUp >> s) & 1;
          const R MP = kp1 * f1 * (sh ? pdh : pd) / (mcs * (facs + rdt1));
          const R NV = CS * (MF4P + MF4N - MP);
          nvm[s] = NV;
          xb(12 + s, j) = (sh ? (MPS - MHPS) : MPS) - 0.5 * ku1 * CS * rvd;
          if (NV >= 0.0) {
            xb(s, j) = RSL[s] * NV;
            xb(6 + s, j) = RSL[s] * VL[s] * NV;
          } else {
            xb(s, j) = RSR[s] * NV;
            xb(6 + s, j) = RSR[s] * VR[s] * NV;
          }
          BNVm = BNVm + dq((RSL[s] + RSR[s]) * NV, RLR, iRLR);
        }
      }
      if constexpr (ctx_taps<Ctx>::value) cx.tap_nv(nvp, nvm);
      // energy and momentum fluxes (:1973-2026)
      if (BNVp >= 0.0) {
        EFp = (EL + PL) * BNVp;
        MOEp = RL * VEL * BNVp;
        MONp = RL * VNL * BNVp;
      } else {
        EFp = (ER + PR) * BNVp;
        MOEp = RR * VER * BNVp;
        MONp = RR * VNR * BNVp;
      }
      if (BNVm >= 0.0) {
        xb(18, j) = (EL + PL) * BNVm;
        xb(19, j) = RL * VEL * BNVm;
        xb(20, j) = RL * VNL * BNVm;
      } else {
        xb(18, j) = (ER + PR) * BNVm;
        xb(19, j) = RR * VER * BNVm;
        xb(20, j) = RR * VNR * BNVm;
      }
    }
    cx.sync();
    // divergence of the face fluxes (:2042-2068); cell k takes its M12 side
    // from the face below
    R rsf[NSP], msf[NSP], ef, mfE, mfN;
    {
      const R ap = st(A_AP, j), am = st(A_AM, j), vol = st(A_VOL, j);
      const R dC = st(A_DC, j);
      const R ivol = 1.0 / vol, idC = 1.0 / dC;
#pragma unroll
      for (int s = 0; s < NSP; s++) {
        rsf[s] = dq(ap * FRp[s] - am * xb(s, j - 1), vol, ivol);
        msf[s] = dq(ap * FMp[s] - am * xb(6 + s, j - 1), vol, ivol);
        msf[s] = msf[s] + dq(NPp[s] - xb(12 + s, j - 1), dC, idC);
      }
      ef = dq(ap * EFp - am * xb(18, j - 1), vol, ivol);
      mfE = dq(ap * MOEp - am * xb(19, j - 1), vol, ivol);
      mfN = dq(ap * MONp - am * xb(20, j - 1), vol, ivol);
    }
    if constexpr (ctx_taps<Ctx>::value) cx.tap_div(rsf, msf, ef, mfE, mfN);
    cx.sync();

    // --- update of cell k (:802-997); each new value is folded into the RK4
    // carries as soon as the stage has no further use of the old one
    if (cell()) {
      R nrs[NSP], nns[NSP], w[NSP];
      R nrho = 0.0, nnt = 0.0;
#pragma unroll
      for (int s = 0; s < NSP; s++) {
        nrs[s] = rhos[s] - dtin * (rsf[s]);
        nns[s] = dq(nrs[s], m[s], kp.imass[s]);
        nrho = nrho + nrs[s];
        nnt = nnt + nns[s];
      }
      const R gradT = gr3[0];
      const R invr = st(A_INVR, j);
#pragma unroll
      for (int s = 0; s < NSP; s++) {
        const R moms = rhos[s] * U[F_W + s];
        R nm;
        if (!((sub >> s) & 1))
          nm = moms - dtin * (msf[s]) + dtin * rhos[s] * S.eg;
        else
          nm = moms - dtin * (msf[s]) + dtin * dev[s] * S.eg;
        R ww = nm / nrs[s];
        ww = ww + dtin * (U[F_VN] * U[F_VN] + U[F_VE] * U[F_VE]) * invr;
        if (kp.coriolis) ww = ww + dtin * (coriolis * U[F_VE]);
        ww = ww - dq(dtin * (kp.tdc[s] * KB * gradT), m[s], kp.imass[s]);
        w[s] = ww;
      }
      const R hf = st(A_HF, j);
      const R inrho = 1.0 / nrho;
      R nvE = dq(rho * U[F_VE] - dtin * (mfE), nrho, inrho);
      nvE = nvE + hf * dtin * df3[1];
      R nvN = dq(rho * U[F_VN] - dtin * (mfN), nrho, inrho);
      nvN = nvN + hf * dtin * df3[2];
      const R nte = te - dtin * ef + dtin * rho * U[F_VU] * S.eg;
      {
        R sr = 0.0;
#pragma unroll
        for (int s = 0; s < NSP; s++) {
          const R nl = ln(nns[s]);
          sr = sr + m[s] * (cx.device_math() ? nns[s] : exp(nl));
          ausm_commit(stg, F_LN + s, nl);
        }
        ausm_commit(stg, F_LR, ln(sr));
      }
      ausm_commit(stg, F_VE, nvE);
      ausm_commit(stg, F_VN, nvN);
      if (kp.friction && !friction(dtin, w, nnt, nns, glc, T))
        fail(GITM_VERT_ESINGULAR, k);
      if constexpr (ctx_taps<Ctx>::value) cx.tap_friction(w);
      R nvu = 0.0;
#pragma unroll
      for (int s = 0; s < NSP; s++) {
        R ww = dmax(-kp.maxvv, w[s]);
        ww = dmin(kp.maxvv, ww);
        nvu = nvu + dq(ww * nrs[s], nrho, inrho);
        ausm_commit(stg, F_W + s, ww);
      }
      const R np = (nte - 0.5 * nrho * (nvu * nvu + nvN * nvN + nvE * nvE)) * S.gm1;
      ausm_commit(stg, F_VU, nvu);
      ausm_commit(stg, F_T, np / (KB * nnt));
    }
  }

  // -------------------------------------------------------------------------
  // advance_vertical_1stage (Rusanov), vertical_solver_rusanov.f90:387-686.
  // Writes New (valid on cell threads).
  // -------------------------------------------------------------------------
  GV_MEM void rusanov_stage(int stg) {
    const double *m = kp.mass;
    const R T = U[F_T];
    R NS[NSP], lcs[NSP];
    const R rho = exp(U[F_LR]);
    R sum = 0.0;
#pragma unroll
    for (int s = 0; s < NSP; s++) {
      NS[s] = exp(U[F_LN + s]);
      sum = sum + NS[s];
    }
    const R nt = exp(ln(sum));  // NT = exp(LogNum) (:458-460)
#pragma unroll
    for (int s = 0; s < NSP; s++)
      lcs[s] = cx.device_math() ? (kp.lmass[s] + U[F_LN + s]) - U[F_LR]
                                : ln(m[s] * NS[s] / rho);
    // gradients and dissipation (:462-498)
    const R va[4] = {T, U[F_VE], U[F_VN], U[F_VU]};
    R ga[4], da[4];
    rusanov_grad_diff<4>(va, ga, da);
    R gln[NSP], dln[NSP], gw[NSP], dw[NSP];
    {
      R v6[NSP];
#pragma unroll
      for (int s = 0; s < NSP; s++) v6[s] = U[F_LN + s];
      rusanov_grad_diff<NSP>(v6, gln, dln);
#pragma unroll
      for (int s = 0; s < NSP; s++) v6[s] = U[F_W + s];
      rusanov_grad_diff<NSP>(v6, gw, dw);
    }
    R glc[NSP];
    grad_log_cons(lcs, glc);
    if (IONS) ion_stage(dt, stg, false);  // :617-630
    if (!cell()) return;
    const R divvel = ga[3] + 2 * U[F_VU] * st(A_INVR, j);  // :476
    R w[NSP], nln[NSP];
#pragma unroll
    for (int s = 0; s < NSP; s++) {  // :550-614
      const R vv = U[F_W + s];
      const R divvv = gw[s] + 2 * vv * st(A_INVR, j);
      nln[s] = U[F_LN + s] - dt * (divvv + vv * gln[s]) + dt * dln[s];
      R ww = vv -
                  dt * (vv * gw[s] -
                        (U[F_VN] * U[F_VN] + U[F_VE] * U[F_VE]) * st(A_INVR, j) +
                        T * gln[s] * KB / m[s] + ga[0] * KB / m[s] - st(A_GRAV, j)) +
                  dt * dw[s] - vv / vtau;
      if (kp.coriolis)
        ww = ww + dt * (centrifugal / st(A_INVR, j) + coriolis * U[F_VE]);
      ww = ww - dt * (kp.tdc[s] * KB * ga[0]) / m[s];
      w[s] = ww;
    }
    if (kp.friction && !friction(dt, w, nt, NS, glc, T))  // :632-641, old densities
      fail(GITM_VERT_ESINGULAR, k);
    R nvu = 0.0;
#pragma unroll
    for (int s = 0; s < NSP; s++) {  // :643-658
      R ww = dmax(-kp.maxvv, w[s]);
      ww = dmin(kp.maxvv, ww);
      w[s] = ww;
      nvu = nvu + ww * (m[s] * NS[s] / rho);
    }
    const R vu = U[F_VU];  // :660-679
    const R nvE = U[F_VE] - dt * vu * ga[1] + dt * da[1];
    const R nvN = U[F_VN] - dt * vu * ga[2] + dt * da[2];
    const R nT = T - dt * (vu * ga[0] + (S.gm1) * (T * divvel)) + dt * da[0];
    R sr = 0.0;
#pragma unroll
    for (int s = 0; s < NSP; s++) sr = sr + m[s] * exp(nln[s]);
#pragma unroll
    for (int s = 0; s < NSP; s++) {
      rusanov_commit(stg, F_LN + s, nln[s]);
      rusanov_commit(stg, F_W + s, w[s]);
    }
    rusanov_commit(stg, F_LR, ln(sr));
    rusanov_commit(stg, F_VU, nvu);
    rusanov_commit(stg, F_VE, nvE);
    rusanov_commit(stg, F_VN, nvN);
    rusanov_commit(stg, F_T, nT);
  }

  // Ion advection of one stage: the first NI_ADVECT ions move with IVel(up)
  // (Rusanov gradient and dissipation, calc_rusanov_alts_*), the others are
  // carried through the RK4 combination unchanged.  The values are final as
  // soon as they are computed, so they are committed at once.
  GV_MEM void ion_stage(R dtin, int stg, bool ausm) {
    R vi[NI_ADVECT], gi[NI_ADVECT], di[NI_ADVECT];
#pragma unroll
    for (int f = 0; f < NI_ADVECT; f++) vi[f] = ion(f, j);
    rusanov_grad_diff<NI_ADVECT>(vi, gi, di);
    if (!cell()) return;
    const R ivu = ion(NI, j);
#pragma unroll
    for (int f = 0; f < NI; f++) {
      const R nv = (f < NI_ADVECT) ? vi[f] - dtin * (ivu * gi[f]) + dtin * di[f]
                                        : ion(f, j);
      if (ausm)
        ausm_commit_ion(stg, f, nv);
      else
        rusanov_commit_ion(stg, f, nv);
    }
  }

  // Stage inputs keep the ghost ion values of the first stage's input: the
  // stage leaves ghosts unchanged, so Y + (New - U) is Y there (AUSM
  // :167-189) and so is Y + K/2 (Rusanov :139-169).  The top ion boundary
  // reads these old ghost values.
  GV_MEM void ion_ghosts_from_y() {
    if (IONS && !cell()) {
#pragma unroll
      for (int f = 0; f < NI; f++) ion(f, j) = Yi(f);
    }
  }
  GV_MEM void ion_store_y() {
    if (IONS) {
#pragma unroll
      for (int f = 0; f < NI; f++) Yi(f) = ion(f, j);
    }
  }

  // -------------------------------------------------------------------------
  // RK4 drivers
  // -------------------------------------------------------------------------
  // advance_vertical_1d_ausm, vertical_solver_ausm.f90:26-433.  Stage k+1
  // input = Y + (New_k - U_k); the final state is
  // (1/3)*(-Y + S2 + 2*S3 + S4 + (New_4 - U_4)) on 1..nAlts, accumulated in
  // the reference's left-to-right order; the boundary conditions follow every
  // stage (ghost levels never enter the combination).
  GV_MEM void ausm_step() {
    bcs(true);
    if (cell()) {
#pragma unroll
      for (int f = 0; f < NF; f++) Yv(f) = U[f];
    }
    ion_store_y();
    const R dtin[4] = {0.5 * dt, 0.5 * dt, dt, 0.5 * dt};  // :160-377
#pragma unroll 1
    for (int stg = 0; stg < 4; stg++) {
      ausm_stage(dtin[stg], stg);
      ion_ghosts_from_y();
      bcs();
    }
  }

  GV_MEM void rusanov_step() {
    bcs(true);
    if (cell()) {
#pragma unroll
      for (int f = 0; f < NF; f++) Yv(f) = U[f];
    }
    ion_store_y();
#pragma unroll 1
    for (int stg = 0; stg < 4; stg++) {
      rusanov_stage(stg);
      ion_ghosts_from_y();
      bcs();
    }
  }
};

// Field descriptors of the state adjoints (the adjoint builds): the same
// layout as the state fields of gitm_vert_columns.
struct AdjFields {
  gitm_vert_field nds, vvel, vel, temp, rho;
};

// ---------------------------------------------------------------------------
// Gather / step / scatter of one column (advance_vertical.f90:84-185)
// ---------------------------------------------------------------------------
GV_FN long long col_base(const gitm_vert_field &f, int ci, int cj, int cb) {
  return (long long)ci * f.s_lon + (long long)cj * f.s_lat + (long long)cb * f.s_blk;
}
GV_FN double *fld(const gitm_vert_field &f, long long base, int k, int comp) {
  return f.ptr + base + (long long)(k - f.first_level) * f.s_lev + (long long)comp * f.s_comp;
}

// One column of `io` (col = its index there, n_alts its nAlts): gather,
// n_steps steps of Dt = dt, scatter.  scratch: the RK4 carries of io's
// columns in device memory, scr_fields per level; err_col: the column number
// that error reports carry.
template <int NT, int SOLVER, bool IONS, bool CARRY, class Ctx>
GV_FN void run_column(Ctx &cx, const KParams &kp, const gitm_vert_columns &io, int col,
                      double *lds, double *scratch, double dt, int err_col, int n_alts,
                      int scr_fields) {
  Column<NT, Ctx, IONS, CARRY> c(cx, kp, lds, err_col, dt, n_alts);
  c.scr = scratch + (long long)col * scr_fields * (n_alts + 4);
  const int nlb = io.n_lons * io.n_lats;
  const int ci = col % io.n_lons, cj = (col / io.n_lons) % io.n_lats, cb = col / nlb;
  const int k = c.k, nA = c.nA;
  const bool lv = c.lv;
  double alt = 0.0, grav = 0.0, invr = 1.0, dalt = 1.0, mmm = 0.0, gam = 1.5,
         eddy = 0.0, cmax = 0.0, tu = 1.0;
  double nds[NSP];
#pragma unroll
  for (int f = 0; f < NF; f++) c.U[f] = 0.0;
  c.U[F_T] = 1.0;
#pragma unroll
  for (int s = 0; s < NSP; s++) nds[s] = 1.0;
  double coslat;
  {
    const gitm_vert_field &f = io.cos_latitude;
    coslat = f.ptr[col_base(f, ci, cj, cb)];
  }
  if (lv) {
#define GV_LD(F) (*fld(io.F, col_base(io.F, ci, cj, cb), k, 0))
    alt = GV_LD(altitude);
    grav = GV_LD(gravity);
    invr = GV_LD(inv_radial_distance);
    dalt = GV_LD(dalt);
    mmm = GV_LD(mean_major_mass);
    gam = GV_LD(gamma);
    tu = GV_LD(temp_unit);
    if (k >= 0 && k <= nA + 1) cmax = GV_LD(cmax);
    if (k >= 1 && k <= nA) eddy = GV_LD(eddy);
    const long long bn = col_base(io.ndensity_s, ci, cj, cb);
    const long long bw = col_base(io.vertical_velocity, ci, cj, cb);
    const long long bv = col_base(io.velocity, ci, cj, cb);
#pragma unroll
    for (int s = 0; s < NSP; s++) {
      nds[s] = *fld(io.ndensity_s, bn, k, s);
      c.U[F_W + s] = *fld(io.vertical_velocity, bw, k, s);
    }
    c.U[F_VE] = *fld(io.velocity, bv, k, 0);
    c.U[F_VN] = *fld(io.velocity, bv, k, 1);
    c.U[F_VU] = *fld(io.velocity, bv, k, 2);
    c.U[F_T] = GV_LD(temperature);
    c.U[F_LR] = GV_LD(rho);
#undef GV_LD
  }
  if (IONS) {
    // LogINS (linear densities with UseImprovedIonAdvection) and IVel(up)
    // (:120-130); SZA and MLat of the column for the nighttime BCs (:84-85)
    const long long bi = col_base(io.idensity_s, ci, cj, cb);
#pragma unroll
    for (int f = 0; f < NI; f++) {
      const double v = *fld(io.idensity_s, bi, k, f);
      c.ion(f, c.j) = kp.improved_ions ? v : ln(v);
    }
    c.ion(NI, c.j) = (k >= 1 && k <= nA)
                         ? *fld(io.ivelocity, col_base(io.ivelocity, ci, cj, cb), k, 2)
                         : 0.0;
    if (kp.nighttime_ion_bcs) {
      const double sza = io.sza.ptr[col_base(io.sza, ci, cj, cb)];
      const double mlat = io.mlat.ptr[col_base(io.mlat, ci, cj, cb)];
      c.plasma = sza > PI / 2 && ((mlat > 30.0 && mlat < 70.0) || (mlat > -60.0 && mlat < -30.0));
    }
  }
  if (kp.msis) {
    // the caller's MSIS values of the column's level 0 (msis_bcs, :113-137)
    const gitm_vert_field &f = io.msis_bc;
    const double *m = f.ptr + col_base(f, ci, cj, cb);
    if (c.j < NSP) c.msis_ln = m[(long long)c.j * f.s_comp];
    if (c.j == 1) {
      c.msis_t = m[6 * f.s_comp];
      c.msis_ve = m[7 * f.s_comp];
      c.msis_vn = m[8 * f.s_comp];
    }
  }
  if (kp.tides && c.j == 1) {
    // TidesEast/North/Temp of level 0 (:150-159); level -1's are overwritten
    // by level 0's (:192-196)
    c.msis_ve = *fld(io.tides_east, col_base(io.tides_east, ci, cj, cb), 0, 0);
    c.msis_vn = *fld(io.tides_north, col_base(io.tides_north, ci, cj, cb), 0, 0);
    c.tide_t = *fld(io.tides_temp, col_base(io.tides_temp, ci, cj, cb), 0, 0);
  }
  c.setup(alt, grav, invr, dalt, mmm, gam, eddy, cmax, tu, coslat);

  for (int step = 0; step < kp.n_steps; step++) {
    // gather (:96-116): the densities must be positive on every level
    bool bad = false;
#pragma unroll
    for (int s = 0; s < NSP; s++) bad = bad || (lv && !(nds[s] > 0.0));
    if (cx.sync_or(bad)) {
      c.ib(c.j) = bad ? k : 1 << 20;
      cx.sync();
      if (c.j == 0) {
        int lo = 1 << 20;
        for (int q = 0; q < c.L; q++)
          if (c.ib(q) < lo) lo = c.ib(q);
        c.fail(GITM_VERT_ENONPOSITIVE, lo);
      }
      return;  // the column is left as it was
    }
    c.U[F_LR] = ln(c.U[F_LR]);
    c.U[F_T] = c.U[F_T] * tu;
#pragma unroll
    for (int s = 0; s < NSP; s++) c.U[F_LN + s] = ln(nds[s]);
    if (SOLVER == GITM_VERT_AUSM)
      c.ausm_step();
    else
      c.rusanov_step();
    // scatter (:163-185)
    c.U[F_LR] = exp(c.U[F_LR]);
    c.U[F_T] = c.U[F_T] / tu;
#pragma unroll
    for (int s = 0; s < NSP; s++) nds[s] = exp(c.U[F_LN + s]);
    if (IONS) {
      // IDensityS as the scatter writes it (:187-212) and the next gather
      // reads it back: floored at 1 on 1..nAlts, or exp(LogINS)
#pragma unroll
      for (int f = 0; f < NI; f++) {
        double &u = c.ion(f, c.j);
        if (kp.improved_ions) {
          if (c.cell() && u < 1) u = 1.0;
        } else if (step + 1 < kp.n_steps) {
          u = ln(exp(u));
        }
      }
    }
  }
  if (lv) {
    const long long bn = col_base(io.ndensity_s, ci, cj, cb);
    const long long bw = col_base(io.vertical_velocity, ci, cj, cb);
    const long long bv = col_base(io.velocity, ci, cj, cb);
    double nsum = 0.0;
#pragma unroll
    for (int s = 0; s < NSP; s++) {
      *fld(io.ndensity_s, bn, k, s) = nds[s];
      *fld(io.vertical_velocity, bw, k, s) = c.U[F_W + s];
      nsum = nsum + nds[s];
    }
    if (io.log_ns.ptr) {
      const long long bl = col_base(io.log_ns, ci, cj, cb);
#pragma unroll
      for (int s = 0; s < NSP; s++) *fld(io.log_ns, bl, k, s) = c.U[F_LN + s];
    }
    if (io.ndensity.ptr) *fld(io.ndensity, col_base(io.ndensity, ci, cj, cb), k, 0) = nsum;
    *fld(io.velocity, bv, k, 0) = c.U[F_VE];
    *fld(io.velocity, bv, k, 1) = c.U[F_VN];
    *fld(io.velocity, bv, k, 2) = c.U[F_VU];
    *fld(io.temperature, col_base(io.temperature, ci, cj, cb), k, 0) = c.U[F_T];
    *fld(io.rho, col_base(io.rho, ci, cj, cb), k, 0) = c.U[F_LR];
    if (IONS) {
      // ions and electrons = their sum (:187-212)
      const long long bi = col_base(io.idensity_s, ci, cj, cb);
      double e = 0.0;
#pragma unroll
      for (int f = 0; f < NI; f++) {
        const double u = c.ion(f, c.j);
        const double v = kp.improved_ions ? u : exp(u);
        *fld(io.idensity_s, bi, k, f) = v;
        e = e + v;
      }
      *fld(io.idensity_s, bi, k, NI) = e;
    }
  }
}

// One column of a single-grid launch: the run's Dt, nAlts and tile, the
// handle's scratch.
template <int NT, int SOLVER, bool IONS, class Ctx>
GV_FN void run_column(Ctx &cx, const KParams &kp, const gitm_vert_columns &io, int col,
                      double *lds) {
  run_column<NT, SOLVER, IONS, carry_in_lds(NT / 64)>(cx, kp, io, col, lds, kp.scratch, kp.dt,
                                                      col, kp.n_alts, kp.scr_fields);
}

}  // namespace gv
