// vert_adj.h -- hand-written adjoint of the AUSM vertical step (neutrals).
//
// The adjoint of one step of src/vertical_solver_ausm.f90:26-433 (with the
// gather and scatter of src/advance_vertical.f90:84-185, the stages of
// :436-999, the fluxes of :1093-2070, the boundary conditions of
// src/set_vertical_bcs.Earth.f90 and the friction solve of
// src/calc_neutral_friction.f90:47-106) for one column, on the same mapping
// as the forward kernel: one workgroup per column, one lane per level.
//
// Forward part: the step is recomputed from its input state by the plain
// column code (vert_column.h) with a context whose taps record, per stage
// and lane, the stage's input state and the intermediates the reverse part
// needs (gradients, hydrostatic densities, face states, flux divergences,
// friction output): NREC doubles per lane and stage, in device memory laid
// out [stage][field][lane] so that a wavefront's accesses are coalesced.
//
// Reverse part: the RK4 combination of the driver is reversed stage by
// stage; each stage's reverse runs its phases backwards, recomputing the
// cheap pointwise values from the record and exchanging the adjoints that
// cross levels (stencils, face hand-offs, the hydrostatic product) through
// LDS with barriers, as the forward pass exchanges values.  The hydrostatic
// product recurrence (ausm :694-751) reverses into a suffix sum over
// levels, taken as a log-step scan.  The friction solve's reverse is one
// solve with the transposed matrix.  No operation tape is kept.
//
// The column prep inputs (TempUnit, Gamma, MeanMajorMass, cMax, the eddy
// coefficient) are inputs of the step too: their adjoints are accumulated
// and written when the caller asks for them (gfix), for the adjoint of the
// column prep (vert_prep.h) to carry back to the state.
#pragma once

#include "vert_column.h"

namespace gv {

// Fields of one stage's record (per lane).
enum AdjRecField {
  RC_U = 0,                  // the stage's input state, NF fields
  RC_GLC = RC_U + NF,        // d ln(rho_s/rho)/dz, 6
  RC_GT = RC_GLC + NSP,      // gradient of Temp
  RC_DFE, RC_DFN,            // Rusanov dissipation of the east and north winds
  RC_HNS,                    // hydrostatic densities, 6
  RC_SUB = RC_HNS + NSP,     // SubtractHydrostatic bits
  RC_HL, RC_HR = RC_HL + NSP,  // face states of face k|k+1, 6 each
  RC_VL = RC_HR + NSP, RC_VR = RC_VL + NSP,
  RC_RSL = RC_VR + NSP, RC_RSR = RC_RSL + NSP,
  RC_PSL = RC_RSR + NSP, RC_PSR = RC_PSL + NSP,
  RC_VEL = RC_PSR + NSP, RC_VER, RC_VNL, RC_VNR,
  RC_RSF, RC_MSF = RC_RSF + NSP,  // flux divergences, 6 each
  RC_EF = RC_MSF + NSP, RC_MFE, RC_MFN,
  RC_WFR,                    // species winds after the friction solve, 6
  RC_KT = RC_WFR + NSP,      // the limiter choices of the cell's face-state slopes
  RC_NVP,                    // numerical velocities of face k|k+1, P12 side, 6
  RC_NVM = RC_NVP + NSP,     // and M12 side, 6
  RC_TP = RC_NVM + NSP,      // Temp**-DiffExp of the friction, the 6 exponents
  NREC = RC_TP + 6
};
// Records per column: the four stages and the step's final state.
constexpr int ADJ_STAGES = 5;
// Work area of one column after its records (per lane): the adjoints of the
// face states of the stage being reversed.
enum AdjWork {
  WK_HL = 0, WK_HR = WK_HL + NSP, WK_VL = WK_HR + NSP, WK_VR = WK_VL + NSP,
  WK_RSL = WK_VR + NSP, WK_RSR = WK_RSL + NSP, WK_PSL = WK_RSR + NSP, WK_PSR = WK_PSL + NSP,
  WK_VEL = WK_PSR + NSP, WK_VER, WK_VNL, WK_VNR, WK_GL, WK_GR,
  NWORK
};
// The boundary conditions' results of each of the step's five passes (the
// scalars SC_BOT..SC_TOP+24 of the column's LDS), per column.
constexpr int ADJ_BC = 60;
// Doubles of one column's record, work area and boundary-condition record,
// lanes of NT.
GV_HD constexpr long adj_column_doubles(int nt) {
  return ((long)ADJ_STAGES * NREC + NWORK) * nt + (long)ADJ_STAGES * ADJ_BC;
}

// Adjoints of the column-prep inputs, per level (gfix of the kernel).
enum AdjFix { FX_TU = 0, FX_GAMMA, FX_MMM, FX_CMAX, FX_EDDY, NFIX };

// Extra LDS of the adjoint kernel beyond the forward column's: the boundary
// conditions' adjoint hand-offs.
// Layout: [0, 36) bottom Rb adjoints, [36, 60) top ones, [60, 84) scalars,
// [84, 156) bottom species contributions, [156, 180) top ones, then the
// column's cMax and Gamma by lane (nAlts+4 each), then the AX_PRE values of
// the boundary levels' state that the boundary conditions' reverse reads
// (share_boundary_state).  At nAlts = 50 the column's LDS is 40.3 KiB, four
// columns per CU.
constexpr int AX_CMAX = 180;
constexpr int AX_PRE_LEN = 87;
GV_HD constexpr int adj_lds_extra(int n_alts) {
  return AX_CMAX + 2 * (n_alts + 4) + AX_PRE_LEN;
}
GV_HD constexpr unsigned long adj_lds_bytes(int n_alts, bool carry_lds) {
  return (lds_doubles(n_alts, false, carry_lds) + adj_lds_extra(n_alts)) * sizeof(double);
}

// The forward context with taps: Base is the plain kernel context.
template <class Base>
struct TapCtx : Base {
  static constexpr bool kTaps = true;
  double *rec = nullptr;  // this column's record
  int nt = 64;            // lanes of the record's rows
  int stage = 0;
  int gamma_code = 0;     // the limiter choice of Gamma's face-value slope
  GV_MEM double &r(int f) { return rec[((long)stage * NREC + f) * nt + this->lane()]; }
  GV_MEM void tap_grads(const double (&glc)[NSP], double gt, double dfe, double dfn) {
#pragma unroll
    for (int s = 0; s < NSP; s++) r(RC_GLC + s) = glc[s];
    r(RC_GT) = gt;
    r(RC_DFE) = dfe;
    r(RC_DFN) = dfn;
  }
  GV_MEM void tap_hydro(const double (&hns)[NSP], int sub) {
#pragma unroll
    for (int s = 0; s < NSP; s++) r(RC_HNS + s) = hns[s];
    r(RC_SUB) = (double)sub;
  }
  GV_MEM void tap_gamma_code(int code) { gamma_code = code; }
  // the friction's temperature powers, recorded as Temp**-DiffExp
  // (`inverse`: the reference form's Temp**+DiffExp)
  GV_MEM void tap_powers(const double (&tp)[6], bool inverse) {
#pragma unroll
    for (int u = 0; u < 6; u++) r(RC_TP + u) = inverse ? 1.0 / tp[u] : tp[u];
  }
  GV_MEM void tap_nv(const double (&nvp)[NSP], const double (&nvm)[NSP]) {
#pragma unroll
    for (int s = 0; s < NSP; s++) {
      r(RC_NVP + s) = nvp[s];
      r(RC_NVM + s) = nvm[s];
    }
  }
  GV_MEM void tap_faces(const double (&HL)[NSP], const double (&HR)[NSP],
                        const double (&VL)[NSP], const double (&VR)[NSP],
                        const double (&RSL)[NSP], const double (&RSR)[NSP],
                        const double (&PSL)[NSP], const double (&PSR)[NSP], double VEL,
                        double VER, double VNL, double VNR, unsigned long long kt) {
    // 2 bits per field (27 fields: face states 0..25, Gamma 26)
    kt |= (unsigned long long)gamma_code << 52;
    r(RC_KT) = __builtin_bit_cast(double, kt);
#pragma unroll
    for (int s = 0; s < NSP; s++) {
      r(RC_HL + s) = HL[s];
      r(RC_HR + s) = HR[s];
      r(RC_VL + s) = VL[s];
      r(RC_VR + s) = VR[s];
      r(RC_RSL + s) = RSL[s];
      r(RC_RSR + s) = RSR[s];
      r(RC_PSL + s) = PSL[s];
      r(RC_PSR + s) = PSR[s];
    }
    r(RC_VEL) = VEL;
    r(RC_VER) = VER;
    r(RC_VNL) = VNL;
    r(RC_VNR) = VNR;
  }
  GV_MEM void tap_div(const double (&rsf)[NSP], const double (&msf)[NSP], double ef,
                      double mfE, double mfN) {
#pragma unroll
    for (int s = 0; s < NSP; s++) {
      r(RC_RSF + s) = rsf[s];
      r(RC_MSF + s) = msf[s];
    }
    r(RC_EF) = ef;
    r(RC_MFE) = mfE;
    r(RC_MFN) = mfN;
  }
  GV_MEM void tap_friction(const double (&w)[NSP]) {
#pragma unroll
    for (int s = 0; s < NSP; s++) r(RC_WFR + s) = w[s];
  }
};

// Gaussian elimination without pivoting of a 6x6 system whose matrix is
// diagonally dominant by columns, for which elimination in the natural
// order is stable: the transposed friction matrix (the friction matrix
// I + dt (diag(sum_q C_iq) - C), C >= 0, is diagonally dominant by rows).
GV_FN void solve6_dd(double (&a)[6][6], double (&b)[6]) {
  double rd[6];  // reciprocals of the pivots, for the elimination and the substitution
#pragma unroll
  for (int k = 0; k < 6; k++) {
    const double r = 1.0 / a[k][k];
    rd[k] = r;
#pragma unroll
    for (int i = k + 1; i < 6; i++) {
      const double f = a[i][k] * r;
#pragma unroll
      for (int q = k + 1; q < 6; q++) a[i][q] = a[i][q] - f * a[k][q];
      b[i] = b[i] - f * b[k];
    }
  }
#pragma unroll
  for (int i = 5; i >= 0; i--) {
    double t = b[i];
#pragma unroll
    for (int q = i + 1; q < 6; q++) t = t - a[i][q] * b[q];
    b[i] = t * rd[i];
  }
}

// Reverse of Limiter_mc (ModLimiter.f90:16-36): the adjoints of up and dn
// from that of the limited slope, on the branch the forward value took.
GV_FN void limiter_mc_rev(double beta, double up, double dn, double ld, double &lup,
                          double &ldn) {
  lup = 0.0;
  ldn = 0.0;
  if (up > 0.0) {
    if (!(dn > 0.0)) return;
    const double a = beta * up, b = beta * dn, c = (up + dn) * 0.5;
    const double m = dmin(a, b);
    if (m < c) {
      if (a < b)
        lup = beta * ld;
      else
        ldn = beta * ld;
    } else {
      lup = 0.5 * ld;
      ldn = 0.5 * ld;
    }
    return;
  }
  if (!(dn < 0.0)) return;
  const double a = beta * up, b = beta * dn, c = (up + dn) * 0.5;
  const double m = dmax(a, b);
  if (m > c) {
    if (a > b)
      lup = beta * ld;
    else
      ldn = beta * ld;
  } else {
    lup = 0.5 * ld;
    ldn = 0.5 * ld;
  }
}

// The adjoint of one column's step.  C is the forward column (its LDS
// tables, exchange buffers and state); lam[] is the adjoint of the state the
// column holds at the point the reverse has reached (all levels).
template <int NT, class Ctx>
struct AdjColumn {
  using Col = Column<NT, Ctx, false>;
  Col &c;
  Ctx &cx;
  const KParams &kp;
  double *rec;   // this column's record
  const double *bcrec;  // its boundary-condition record
  double *ax;    // the extra LDS (adj_lds_extra(nAlts) doubles)
  const int j, k, nA, L;
  double lam[NF];
  double fix[NFIX];  // adjoints of this level's column-prep inputs

  GV_MEM AdjColumn(Col &col, Ctx &ctx, double *record, double *extra)
      : c(col), cx(ctx), kp(col.kp), rec(record),
        bcrec(record + ((long)ADJ_STAGES * NREC + NWORK) * NT), ax(extra), j(col.j), k(col.k),
        nA(col.nA), L(col.L) {
#pragma unroll
    for (int f = 0; f < NF; f++) lam[f] = 0.0;
#pragma unroll
    for (int f = 0; f < NFIX; f++) fix[f] = 0.0;
  }

  GV_MEM double R(int s, int f) const { return rec[((long)s * NREC + f) * NT + j]; }
  // the work area of this lane
  GV_MEM double &W(int f) const { return rec[((long)ADJ_STAGES * NREC + f) * NT + j]; }
  // work slots of face-state field f (R4's numbering): its left state, and
  // the offset of the right state from it
  GV_MEM static int wk_left(int f) {
    return f < 6 ? WK_HL + f : f < 12 ? WK_VL + f - 6 : f < 18 ? WK_RSL + f - 12
         : f < 24 ? WK_PSL + f - 18 : f == 24 ? WK_VEL : f == 25 ? WK_VNL : WK_GL;
  }
  GV_MEM static int wk_right_offset(int f) { return f < 24 ? NSP : 1; }
  GV_MEM double &xb(int b, int jj) const { return c.xb(b, jj); }
  GV_MEM double st(int a, int jj) const { return c.st(a, jj); }
  GV_MEM bool cell() const { return k >= 1 && k <= nA; }
  GV_MEM bool face() const { return k >= 0 && k <= nA; }
  // a lane in 0..L-1 (the workgroup has exactly L lanes)
  GV_MEM bool in(int jj) const { return jj >= 0 && jj < L; }

  // -------------------------------------------------------------------------
  // set_vertical_bcs (set_vertical_bcs.Earth.f90:8-652, UseMSIS=F), reversed.
  // On entry lam holds the adjoint of the state after the boundary
  // conditions; on exit that of the state before them: the ghost levels'
  // overwritten fields are carried to the interior levels they were made
  // from, and to Temp(0) and the non-photochemical LogNS(0), which the
  // boundary conditions read and keep.  c.U must hold the state of record
  // stage s (its interior levels are the state before the boundary
  // conditions; the ghost values are recomputed here).
  // -------------------------------------------------------------------------
  // The LDS values of the boundary levels' state (AX_PRE_LEN doubles after
  // cMax and Gamma): [sp*12 + q] LogNS of species sp and [sp*12 + 6 + q] its
  // VertVel at level q = 0..5; [72] Temp(0); [73] Temp(nAlts);
  // [74 + 2 sp], [75 + 2 sp] VertVel of species sp at nAlts and nAlts-1;
  // [86] Vel(up) at nAlts.
  GV_MEM double *pre() const { return ax + AX_CMAX + 2 * L; }
  // Each lane that holds one of those levels' state (U, the state the
  // boundary conditions' reverse of the same pass reads) puts it there, so
  // that bcs_rev reads LDS instead of other lanes' records in device memory.
  // Called before a barrier that precedes the bcs_rev of the pass, and after
  // one that follows the previous bcs_rev's reads.
  GV_MEM void share_boundary_state(const double (&U)[NF]) {
    double *P = pre();
    if (j >= 1 && j <= 6) {
      const int q = j - 1;
#pragma unroll
      for (int sp = 0; sp < NSP; sp++) {
        P[sp * 12 + q] = U[F_LN + sp];
        P[sp * 12 + 6 + q] = U[F_W + sp];
      }
      if (j == 1) P[72] = U[F_T];
    } else if (j == nA + 1) {
      P[73] = U[F_T];
#pragma unroll
      for (int sp = 0; sp < NSP; sp++) P[74 + 2 * sp] = U[F_W + sp];
      P[86] = U[F_VU];
    } else if (j == nA) {
#pragma unroll
      for (int sp = 0; sp < NSP; sp++) P[75 + 2 * sp] = U[F_W + sp];
    }
  }

  GV_MEM void bcs_rev(int s) {
    const double *m = kp.mass;
    const int jn = nA + 1;  // lane of level nAlts
    // the pass's results: Rb of the bottom, 6 per species, then the top's,
    // 4 per species (the forward recorded them)
    // (stage_rev copied them from the record to the LDS scalars they came from)
    const double *bc = &c.sc(SC_BOT);
    // the values the species lanes read, from the record of state s: its
    // interior levels are the state before the boundary conditions, and
    // Temp(0) and the non-photochemical LogNS(0) are never changed by them
    // (share_boundary_state left them in LDS, P)
    const double *P = pre();
    double lnb[6], wb[6], t0 = 0.0;
    double tn = 0.0, wn = 0.0, wnm1 = 0.0;
    if (j < NSP) {
#pragma unroll
      for (int q = 0; q < 6; q++) {
        lnb[q] = P[j * 12 + q];     // LogNS of levels 0..5
        wb[q] = P[j * 12 + 6 + q];  // VertVel of levels 0..5
      }
      t0 = P[72];
    } else if (j >= L - NSP) {  // the top on the column's last lanes, as the forward
      const int sp = L - 1 - j;
      tn = P[73];
      wn = P[74 + 2 * sp];
      wnm1 = P[75 + 2 * sp];
    }
    // hand-off area: bottom [sp*6 + 0..5] adjoints of Rb (n0, nm1, w0, wm1,
    // w0r, wm1r); top [36 + sp*4 + 0..3] of (l1, l2, w1, w2); [60..63]
    // T(-1), and the top's T, VE, VN, VU copies (two levels each, 64..71)
    double *AB = ax, *AT = ax + 36, *AS = ax + 60;
    // species contributions to interior lanes: bottom [84 + sp*12 + ...],
    // top [156 + sp*4 + ...]
    double *CB = ax + 84, *CT = ax + 156;
    // 1. ghost lanes take their adjoints apart
    if (j <= 1) {
      const int o = (j == 1) ? 0 : 1;
      double ns[NSP], sr = 0.0;
#pragma unroll
      for (int sp = 0; sp < NSP; sp++) {
        ns[sp] = exp(bc[6 * sp + o]);
        sr = sr + m[sp] * ns[sp];
      }
      const double lr = ln(sr);
      const double er = exp(lr), ier = 1.0 / er;
      double vu = 0.0;
#pragma unroll
      for (int sp = 0; sp < NSP; sp++)
        vu = vu + ns[sp] * m[sp] * bc[6 * sp + 2 + o] * ier;
      const double lvu = lam[F_VU];
      const double ler = -lvu * vu * ier;
      const double lsr = (lam[F_LR] + ler * er) / sr;
#pragma unroll
      for (int sp = 0; sp < NSP; sp++) {
        const double w = bc[6 * sp + 2 + o];
        const double lns = lsr * m[sp] + lvu * m[sp] * w * ier;
        AB[sp * 6 + o] = lam[F_LN + sp] + lns * ns[sp];
        AB[sp * 6 + 2 + o] = lvu * ns[sp] * m[sp] * ier;
        AB[sp * 6 + 4 + o] = lam[F_W + sp];
      }
      if (j == 0) AS[0] = lam[F_T];
      // every field of level -1 and all but Temp(0) of level 0 are
      // overwritten; the non-photochemical LogNS(0) come back below
#pragma unroll
      for (int f = 0; f < NF; f++)
        if (!(j == 1 && f == F_T)) lam[f] = 0.0;
    } else if (j == nA + 2 || j == nA + 3) {
      const int o = (j == nA + 2) ? 0 : 1;
      double sr = 0.0, e[NSP];
#pragma unroll
      for (int sp = 0; sp < NSP; sp++) {
        e[sp] = exp(bc[36 + 4 * sp + o]);
        sr = sr + m[sp] * e[sp];
      }
      const double isr = 1.0 / sr;
#pragma unroll
      for (int sp = 0; sp < NSP; sp++) {
        AT[sp * 4 + o] = lam[F_LN + sp] + lam[F_LR] * m[sp] * e[sp] * isr;
        AT[sp * 4 + 2 + o] = lam[F_W + sp];
      }
      AS[4 + o] = lam[F_T];
      AS[6 + o] = lam[F_VE];
      AS[8 + o] = lam[F_VN];
      AS[10 + o] = lam[F_VU];
#pragma unroll
      for (int f = 0; f < NF; f++) lam[f] = 0.0;
    }
    cx.sync();
    // 2. one lane per species, bottom and top
    if (j < NSP) {
      const int sp = j;
      const double *CA = &c.sc(SC_CA), *CBw = &c.sc(SC_CB);
      const double dF1 = st(A_DF, 2), dF0 = st(A_DF, 1), dFm1 = st(A_DF, 0);
      const double *lb = &AB[sp * 6];
      double cln[6] = {0, 0, 0, 0, 0, 0}, cw[6] = {0, 0, 0, 0, 0, 0}, ct0 = 0.0;
      // winds: w0r/wm1r are the values the ghost levels keep
      double lw0 = lb[2], lwm1 = lb[3];
      if (photochemical(sp) && wb[1] > 0.0) {
        cw[1] -= lb[4] + lb[5];
      } else {
        lw0 += lb[4];
        lwm1 += lb[5];
      }
      // wm1 = w0 - dF0*(w0*CB0 + w1*CB1 + .. + w4*CB4)
      lw0 += lwm1 * (1.0 - dF0 * CBw[0]);
#pragma unroll
      for (int q = 1; q < 5; q++) cw[q] -= lwm1 * dF0 * CBw[q];
      // w0 = w1 - dF1*(w1*CA0 + w2*CA1 + .. + w5*CA4)
      cw[1] += lw0 * (1.0 - dF1 * CA[0]);
#pragma unroll
      for (int q = 1; q < 5; q++) cw[1 + q] -= lw0 * dF1 * CA[q];
      double ln0 = lb[0], lnm1 = lb[1];
      if (photochemical(sp)) {
        const double ln1 = lnb[1], ln2 = lnb[2], ln3 = lnb[3], ln4 = lnb[4], ln5 = lnb[5];
        const double d = ln1 * CA[0] + ln2 * CA[1] + ln3 * CA[2] + ln4 * CA[3] + ln5 * CA[4];
        double n0 = ln1 - dF1 * d;
        const bool n0flat = d <= 0.0;
        if (n0flat) n0 = ln1;
        const bool n0clamp = sp == S_NO && n0 < ln1 - 0.333;
        if (n0clamp) n0 = ln1 - 0.333;
        const double d2 = n0 * CBw[0] + ln1 * CBw[1] + ln2 * CBw[2] + ln3 * CBw[3] + ln4 * CBw[4];
        double nm1 = n0 - dF0 * d2;
        const bool m1flat = d2 <= 0.0;
        if (m1flat) nm1 = n0;
        const bool m1clamp = sp == S_NO && nm1 < n0 - 0.333;
        if (m1clamp || m1flat) {
          ln0 += lnm1;
        } else {
          ln0 += lnm1 * (1.0 - dF0 * CBw[0]);
#pragma unroll
          for (int q = 1; q < 5; q++) cln[q] -= lnm1 * dF0 * CBw[q];
        }
        if (n0clamp || n0flat) {
          cln[1] += ln0;
        } else {
          cln[1] += ln0 * (1.0 - dF1 * CA[0]);
#pragma unroll
          for (int q = 1; q < 5; q++) cln[1 + q] -= ln0 * dF1 * CA[q];
        }
      } else {
        cln[0] += ln0;  // n0 = LogNS(0) itself
        if (sp == S_O) {
          cln[0] += lnm1;
        } else {
          // nm1 = log(exp(ln0) * (t0/t0) * exp(ish*dFm1)), ish ~ 1/t0
          const double MeanGravity = -c.sc(SC_MG0);
          const double MeanTemp = 0.5 * (t0 + t0);
          const double MeanMass = 0.5 * (c.sc(SC_MMM_0) + c.sc(SC_MMM_M1));
          const double ish = MeanGravity * MeanMass / (MeanTemp * KB);
          cln[0] += lnm1;
          ct0 += -lnm1 * dFm1 * ish / t0;
          // MeanMass = (mmm(0) + mmm(-1))/2
          AS[12 + sp] = lnm1 * dFm1 * ish / MeanMass;
        }
      }
      double *cb = &CB[sp * 12];
#pragma unroll
      for (int q = 0; q < 6; q++) {
        cb[q] = cln[q];
        cb[6 + q] = cw[q];
      }
      AS[18 + sp] = ct0;
      if (photochemical(sp) || sp == S_O) AS[12 + sp] = 0.0;
    } else if (j >= L - NSP) {
      const int sp = L - 1 - j;
      const double *lt = &AT[sp * 4];
      const double mm = kp.mass[sp];
      const double ish1 = -c.sc(SC_MGT1) * mm / ((0.5 * (tn + tn)) * KB);
      const double ish2 = -c.sc(SC_MGT2) * mm / ((0.5 * (tn + tn)) * KB);
      const double d1 = st(A_DF, jn + 1) * ish1 / tn, d2 = st(A_DF, jn + 2) * ish2 / tn;
      double *ct = &CT[sp * 4];
      ct[0] = lt[0] + lt[1];                // LogNS(nAlts)
      ct[1] = lt[0] * d1 + lt[1] * (d1 + d2);  // Temp(nAlts)
      if (photochemical(sp) && wn < 0.0) {
        ct[2] = -lt[2];  // VertVel(nAlts)
        ct[3] = -lt[3];  // VertVel(nAlts-1)
      } else {
        ct[2] = lt[2] + lt[3];
        ct[3] = 0.0;
      }
    }
    cx.sync();
    // 3. interior lanes collect
    if (j >= 1 && j <= 6) {
      const int lev = j - 1;  // level 0..5
#pragma unroll
      for (int sp = 0; sp < NSP; sp++) {
        lam[F_LN + sp] += CB[sp * 12 + lev];
        lam[F_W + sp] += CB[sp * 12 + 6 + lev];
      }
      if (j == 1) {
        double t = AS[0];
#pragma unroll
        for (int sp = 0; sp < NSP; sp++) t += AS[18 + sp];
        lam[F_T] += t;
      }
    }
    if (j == 0 || j == 1) {
      // MeanMass of the level -1 hydrostatic fill, half from each level
      double mm = 0.0;
#pragma unroll
      for (int sp = 0; sp < NSP; sp++) mm += AS[12 + sp];
      fix[FX_MMM] += 0.5 * mm;
    }
    if (j == jn) {
      double t = AS[4] + AS[5];
#pragma unroll
      for (int sp = 0; sp < NSP; sp++) {
        lam[F_LN + sp] += CT[sp * 4 + 0];
        lam[F_W + sp] += CT[sp * 4 + 2];
        t += CT[sp * 4 + 1];
      }
      lam[F_T] += t;
      lam[F_VE] += AS[6] + AS[7];
      lam[F_VN] += AS[8] + AS[9];
      // Vel(up) of the top ghosts = |Vel(up, nAlts)|
      const double sgn = (P[86] < 0.0) ? -1.0 : 1.0;
      lam[F_VU] += sgn * (AS[10] + AS[11]);
    }
    if (j == jn - 1) {
#pragma unroll
      for (int sp = 0; sp < NSP; sp++) lam[F_W + sp] += CT[sp * 4 + 3];
    }
    cx.sync();
  }

  // -------------------------------------------------------------------------
  // Reverse of one AUSM stage (advance_vertical_1stage_ausm,
  // vertical_solver_ausm.f90:436-999) with record stage s.  On entry, -lam
  // on cell levels is the adjoint of the stage's new values (the values
  // passed to the RK4 combination: the driver has already moved the
  // combination's other terms); their contributions are added to lam, the
  // adjoint of the stage's input state (all levels).
  // -------------------------------------------------------------------------
  GV_MEM void stage_rev(int s, double dtin) {
    const double *m = kp.mass;
    double U[NF];
#pragma unroll
    for (int f = 0; f < NF; f++) U[f] = R(s, RC_U + f);
    share_boundary_state(U);  // for this pass's bcs_rev, after this stage's barriers
    // this pass's boundary results (bcs_rev reads them in the LDS scalars):
    // loaded here, stored before the hydrostatic reverse's barriers, so that
    // the load is waited for late
    const double bcv0 = (j < ADJ_BC) ? bcrec[s * ADJ_BC + j] : 0.0;
    const double bcv1 = (j + L < ADJ_BC) ? bcrec[s * ADJ_BC + j + L] : 0.0;
    const double T = U[F_T], iT = 1.0 / T;
    const double rho = exp(U[F_LR]);
    double NS[NSP], rhos[NSP], nt = 0.0;
#pragma unroll
    for (int q = 0; q < NSP; q++) {
      NS[q] = exp(U[F_LN + q]);
      nt = nt + NS[q];
      rhos[q] = m[q] * NS[q];
    }
    const double gm1 = c.S.gm1, eg = c.S.eg;
    // adjoints of pointwise intermediates, collected over the phases
    double lNS[NSP], lrho = 0.0, lnt = 0.0;
#pragma unroll
    for (int q = 0; q < NSP; q++) lNS[q] = 0.0;
    double lglc[NSP], lgt = 0.0, ldfe = 0.0, ldfn = 0.0, lhns[NSP];
    double lrsf[NSP], lmsf[NSP], lef = 0.0, lmfe = 0.0, lmfn = 0.0;
#pragma unroll
    for (int q = 0; q < NSP; q++) lglc[q] = lhns[q] = lrsf[q] = lmsf[q] = 0.0;
    const int sub = (int)R(s, RC_SUB);

    // --- R1: the update of cell k (:802-997) ---------------------------
    if (cell()) {
      double lnv[NF];
#pragma unroll
      for (int f = 0; f < NF; f++) lnv[f] = -lam[f];
      double rsf[NSP], msf[NSP], nrs[NSP], nns[NSP], hns[NSP], w[NSP], glc[NSP];
      double nrho = 0.0, nnt = 0.0;
#pragma unroll
      for (int q = 0; q < NSP; q++) {
        rsf[q] = R(s, RC_RSF + q);
        msf[q] = R(s, RC_MSF + q);
        hns[q] = R(s, RC_HNS + q);
        w[q] = R(s, RC_WFR + q);
        glc[q] = R(s, RC_GLC + q);
        nrs[q] = rhos[q] - dtin * rsf[q];
        nns[q] = nrs[q] * kp.imass[q];
        nrho = nrho + nrs[q];
        nnt = nnt + nns[q];
      }
      const double ef = R(s, RC_EF), mfE = R(s, RC_MFE), mfN = R(s, RC_MFN);
      const double dfe = R(s, RC_DFE), dfn = R(s, RC_DFN), gradT = R(s, RC_GT);
      const double invr = st(A_INVR, j), hf = st(A_HF, j);
      const double inrho = 1.0 / nrho;
      const double VE = U[F_VE], VN = U[F_VN], VU = U[F_VU];
      const double K2 = VU * VU + VN * VN + VE * VE;
      const double te = nt * KB * T / gm1 + 0.5 * rho * K2;
      double ww[NSP], nvu = 0.0, swn = 0.0;
#pragma unroll
      for (int q = 0; q < NSP; q++) {
        double x = dmax(-kp.maxvv, w[q]);
        x = dmin(kp.maxvv, x);
        ww[q] = x;
        nvu = nvu + x * nrs[q] * inrho;
        swn = swn + x * nrs[q];
      }
      const double nvE = (rho * VE - dtin * mfE) * inrho + hf * dtin * dfe;
      const double nvN = (rho * VN - dtin * mfN) * inrho + hf * dtin * dfn;
      const double nte = te - dtin * ef + dtin * rho * VU * eg;
      const double K2n = nvu * nvu + nvN * nvN + nvE * nvE;
      const double np = (nte - 0.5 * nrho * K2n) * gm1;
      const double iknt = 1.0 / (KB * nnt);
      const double nT = np * iknt;
      double sr = 0.0;
#pragma unroll
      for (int q = 0; q < NSP; q++) sr = sr + m[q] * nns[q];
      // T' = np / (KB nnt); np = (nte - nrho K2n/2) (gamma - 1)
      const double lnp = lnv[F_T] * iknt;
      double lnnt = -lnv[F_T] * nT * (KB * iknt);
      const double lnte = lnp * gm1;
      double lnrho = -0.5 * lnp * gm1 * K2n;
      const double lnvu = lnv[F_VU] - lnp * gm1 * nrho * nvu;
      const double lnvN = lnv[F_VN] - lnp * gm1 * nrho * nvN;
      const double lnvE = lnv[F_VE] - lnp * gm1 * nrho * nvE;
      fix[FX_GAMMA] += lnp * (nte - 0.5 * nrho * K2n);
      // nte = te - dtin ef + dtin rho VU eg
      const double lte = lnte;
      lef = -dtin * lnte;
      lrho += lnte * dtin * VU * eg;
      lam[F_VU] += lnte * dtin * rho * eg;
      // nvE, nvN
      double linrho = lnvE * (rho * VE - dtin * mfE) + lnvN * (rho * VN - dtin * mfN);
      lrho += (lnvE * VE + lnvN * VN) * inrho;
      lam[F_VE] += lnvE * rho * inrho;
      lam[F_VN] += lnvN * rho * inrho;
      lmfe = -lnvE * dtin * inrho;
      lmfn = -lnvN * dtin * inrho;
      ldfe = lnvE * hf * dtin;
      ldfn = lnvN * hf * dtin;
      // nvu = sum ww nrs / nrho; the clamp to +-MaximumVerticalVelocity
      double lnrs[NSP], lw[NSP], lnns[NSP];
#pragma unroll
      for (int q = 0; q < NSP; q++) {
        const double lww = lnv[F_W + q] + lnvu * nrs[q] * inrho;
        lnrs[q] = lnvu * ww[q] * inrho;
        lw[q] = (w[q] >= -kp.maxvv && w[q] <= kp.maxvv) ? lww : 0.0;
        lnns[q] = 0.0;
      }
      linrho += lnvu * swn;
      // friction: w = x + ec, M x = w_pre, ec = -eddy * glc
      double lwpre[NSP];
      if (kp.friction) {
        const double eddy = st(A_EDDY, j);
        double x[NSP], M[NSP][NSP], cc[NSP][NSP], tp[6];
#pragma unroll
        for (int u = 0; u < 6; u++) tp[u] = R(s, RC_TP + u);
        const double kT = KB * T;
#pragma unroll
        for (int i = 0; i < NSP; i++) {
          x[i] = w[i] + eddy * glc[i];
          lglc[i] += -eddy * lw[i];
          fix[FX_EDDY] -= glc[i] * lw[i];
          double sum = 0.0;
#pragma unroll
          for (int q = 0; q < NSP; q++) {
            if (q == i) {
              cc[i][q] = 0.0;
              continue;
            }
            const double g = 1.0e-19 / D0R[i][q];
            cc[i][q] = kT * kp.imass[i] * (g * tp[DEXI[i][q]]);  // C_iq / nns_q
            const double C = cc[i][q] * nns[q];
            sum = sum + C;
            // transposed: M^T[q][i] = M[i][q]
            M[q][i] = -dtin * C;
          }
          M[i][i] = 1.0 + dtin * sum;
        }
        double mu[NSP];
#pragma unroll
        for (int i = 0; i < NSP; i++) mu[i] = lw[i];
        solve6_dd(M, mu);
        double lT = 0.0;
#pragma unroll
        for (int i = 0; i < NSP; i++) {
          lwpre[i] = mu[i];
#pragma unroll
          for (int q = 0; q < NSP; q++) {
            if (q == i) continue;
            // M_iq = -dtin C_iq, M_ii = 1 + dtin sum_q C_iq
            const double lC = dtin * mu[i] * (x[q] - x[i]);
            lnns[q] += lC * cc[i][q];
            lT += lC * cc[i][q] * nns[q] * (1.0 - DEX[i][q]) * iT;
          }
        }
        lam[F_T] += lT;
      } else {
#pragma unroll
        for (int i = 0; i < NSP; i++) lwpre[i] = lw[i];
      }
      // w_pre = nm/nrs + dtin (VN^2 + VE^2) invr + dtin coriolis VE
      //         - dtin tdc KB gradT / m
      double lVN = 0.0, lVE = 0.0;
#pragma unroll
      for (int q = 0; q < NSP; q++) {
        const double l = lwpre[q];
        const double moms = rhos[q] * U[F_W + q];
        const bool sh = (sub >> q) & 1;
        const double dev = rhos[q] - m[q] * hns[q];
        const double nm = moms - dtin * msf[q] + dtin * (sh ? dev : rhos[q]) * eg;
        const double inrs = 1.0 / nrs[q];
        const double lnm = l * inrs;
        lnrs[q] -= lnm * nm * inrs;
        lVN += l * dtin * 2.0 * VN * invr;
        lVE += l * dtin * 2.0 * VE * invr;
        if (kp.coriolis) lVE += l * dtin * c.coriolis;
        lgt -= l * dtin * kp.tdc[q] * KB * kp.imass[q];
        // nm = rhos W - dtin msf + dtin (dev | rhos) eg
        double lrhos = lnm * U[F_W + q];
        lam[F_W + q] += lnm * rhos[q];
        lmsf[q] = -lnm * dtin;
        lrhos += lnm * dtin * eg;  // through rhos, or dev = rhos - m hns
        if (sh) lhns[q] -= m[q] * lnm * dtin * eg;
        lNS[q] += m[q] * lrhos;
      }
      lam[F_VN] += lVN;
      lam[F_VE] += lVE;
      // new LogNS = log(nns), new LogRho = log(sum m nns)
      const double isr = 1.0 / sr;
#pragma unroll
      for (int q = 0; q < NSP; q++) {
        lnns[q] += lnv[F_LN + q] / nns[q] + lnv[F_LR] * m[q] * isr + lnnt;
      }
      lnrho -= linrho * inrho * inrho;
#pragma unroll
      for (int q = 0; q < NSP; q++) {
        lnrs[q] += lnrho + lnns[q] * kp.imass[q];
        lNS[q] += m[q] * lnrs[q];  // nrs = rhos - dtin rsf
        lrsf[q] = -dtin * lnrs[q];
      }
      // te = nt KB T / gm1 + rho K2 / 2
      lnt += lte * KB * T / gm1;
      lam[F_T] += lte * nt * KB / gm1;
      fix[FX_GAMMA] -= lte * nt * KB * T / (gm1 * gm1);
      lrho += lte * 0.5 * K2;
      lam[F_VU] += lte * rho * VU;
      lam[F_VN] += lte * rho * VN;
      lam[F_VE] += lte * rho * VE;
    }

    // --- R2: the divergence of the face fluxes (:2042-2068) ------------
    // cell k hands the adjoints of its M12 inputs to the face below (lane
    // j-1), which computed them
    {
      const double ap = st(A_AP, j), am = st(A_AM, j), vol = st(A_VOL, j), dC = st(A_DC, j);
      const double ivol = 1.0 / vol, idC = 1.0 / dC;
      const bool cl = cell();
#pragma unroll
      for (int q = 0; q < NSP; q++) {
        xb(q, j) = cl ? -lrsf[q] * am * ivol : 0.0;
        xb(6 + q, j) = cl ? -lmsf[q] * am * ivol : 0.0;
        xb(12 + q, j) = cl ? -lmsf[q] * idC : 0.0;
      }
      xb(18, j) = cl ? -lef * am * ivol : 0.0;
      xb(19, j) = cl ? -lmfe * am * ivol : 0.0;
      xb(20, j) = cl ? -lmfn * am * ivol : 0.0;
    }
    cx.sync();

    // --- R3: AUSM+-up fluxes of face k|k+1 (:1459-1970) -----------------
    // the adjoints of the face states go to the work area (WK_*), those of
    // the exponentiated states already in log form
    if (face()) flux_rev(s, dtin, lrsf, lmsf, lef, lmfe, lmfn);
    cx.sync();

    // --- R4: face states (calc_kt_facevalues, :2073-2117) ---------------
    // 27 fields: ln p_hydro,s, w_s, ln rho_s, ln p_s, Vel(E), Vel(N) and
    // Gamma (its face values, :1535, made once per column by the setup).
    // Cell j's limited slope d(j) enters the left state of face j (+d dF/2)
    // and the right state of face j-1 (-d dF/2); d = MC(up, dn) with up =
    // (f(j+1)-f(j))/dF(j+1), dn = (f(j)-f(j-1))/dF(j), on the branch the
    // forward recorded (RC_KT).  Batches of 7 fields, 3 exchange slots each.
    double hns[NSP];
#pragma unroll
    for (int q = 0; q < NSP; q++) hns[q] = R(s, RC_HNS + q);
    {
      const unsigned long long kt = __builtin_bit_cast(unsigned long long, R(s, RC_KT));
      const bool act = j >= 1 && j <= nA + 2;
      const double iF0 = st(A_IF, j), iF1 = st(A_IF, j + 1), hdF0 = 0.5 * st(A_DF, j);
      // unrolled: every field index, and so every index into lam, the work
      // area and the limiter word, is a constant
#pragma unroll
      for (int b0 = 0; b0 < 27; b0 += 7) {
        double lo[7], lf[7];
#pragma unroll
        for (int i = 0; i < 7; i++) {
          const int f = b0 + i;
          lo[i] = 0.0;
          lf[i] = 0.0;
          double lr = 0.0;
          if (f < 27 && face()) {
            lo[i] = W(wk_left(f));
            lr = W(wk_left(f) + wk_right_offset(f));
          }
          xb(3 * i, j) = lr;
        }
        cx.sync();
#pragma unroll
        for (int i = 0; i < 7; i++) {
          const int f = b0 + i;
          double a = 0.0, b = 0.0;
          if (act && f < 27) {
            const double lrp = xb(3 * i, j - 1);  // adjoint of face j-1's right state
            const double ld = hdF0 * (lo[i] - lrp);
            lf[i] += lo[i] + lrp;
            const int code = (int)((kt >> (2 * f)) & 3);
            const double lup = code == 1 ? kp.beta * ld : (code == 3 ? 0.5 * ld : 0.0);
            const double ldn = code == 2 ? kp.beta * ld : (code == 3 ? 0.5 * ld : 0.0);
            a = lup * iF1;
            b = ldn * iF0;
            lf[i] += b - a;
          }
          xb(3 * i + 1, j) = a;
          xb(3 * i + 2, j) = b;
        }
        cx.sync();
#pragma unroll
        for (int i = 0; i < 7; i++) {
          const int f = b0 + i;
          if (f >= 27) continue;
          if (j - 1 >= 1 && j - 1 <= nA + 2) lf[i] += xb(3 * i + 1, j - 1);
          if (j + 1 <= nA + 2) lf[i] -= xb(3 * i + 2, j + 1);
          if (f < 6) {
            lhns[f] += lf[i] / hns[f];  // ln(hns KB T)
            lam[F_T] += lf[i] * iT;
          } else if (f < 12) {
            lam[F_W + f - 6] += lf[i];
          } else if (f < 18) {
            lam[F_LN + f - 12] += lf[i];  // ln(m ns)
          } else if (f < 24) {
            lam[F_LN + f - 18] += lf[i];  // ln(ns KB T)
            lam[F_T] += lf[i] * iT;
          } else if (f == 24) {
            lam[F_VE] += lf[i];
          } else if (f == 25) {
            lam[F_VN] += lf[i];
          } else {
            fix[FX_GAMMA] += lf[i];
          }
        }
      }
      cx.sync();
    }

    // --- R5: SubtractHydrostatic's deviation: done in R1 (lhns) ----------
    if (j < ADJ_BC) c.sc(SC_BOT + j) = bcv0;
    if (j + L < ADJ_BC) c.sc(SC_BOT + j + L) = bcv1;
    // --- R6: the hydrostatic background (:694-751) ----------------------
    // hns(j) = ns0 * prod_{i=2..j} a(i) b_s(i) for lanes j >= 2, hns(1) =
    // NS(level 0), hns(0) = ns0 a(0) b_s(0).  With G(j) = lhns(j) hns(j),
    // the adjoint of ln a(i) and ln b_s(i) is the suffix sum
    // S_s(i) = sum_{j >= i} G_s(j), taken as a log-step scan.
    hydro_rev(s, hns, lhns, T, NS, lNS);

    // --- R7: gradients and dissipation of Temp, Vel(E), Vel(N) (:611-646)
    grad_diff_rev(U, lgt, ldfe, ldfn);

    // --- R8: d ln(rho_s/rho)/dz (:649-688) ------------------------------
#pragma unroll
    for (int q = 0; q < NSP; q++) xb(q, j) = cell() ? lglc[q] : 0.0;
    cx.sync();
    {
      double llcs[NSP] = {0, 0, 0, 0, 0, 0};
#pragma unroll
      for (int d = -2; d <= 2; d++) {
        const int jj = j + d;
        if (jj < 2 || jj > nA + 1) continue;  // cells only
        const double wgt = st(A_GLC0 + (2 - d), jj);
#pragma unroll
        for (int q = 0; q < NSP; q++) llcs[q] += wgt * xb(q, jj);
      }
#pragma unroll
      for (int q = 0; q < NSP; q++) {
        lam[F_LN + q] += llcs[q];
        lam[F_LR] -= llcs[q];
      }
    }
    cx.sync();

    // --- R9: pointwise (:591-609) ---------------------------------------
#pragma unroll
    for (int q = 0; q < NSP; q++) lam[F_LN + q] += (lNS[q] + lnt) * NS[q];
    lam[F_LR] += lrho * rho;
  }

  // Reverse of the face's AUSM+-up flux (:1459-2026).  Inputs: the
  // adjoints of the face's outputs, on the P12 side from this lane's cell
  // (through the divergence: lrsf, lmsf, lef, lmfe, lmfn) and on the M12
  // side from the cell above (the exchange buffers of lane j+1).  Outputs, to
  // the work area: the adjoints of the face states (of the logarithms for the
  // exponentiated ones) and of the face gammas.  The face states stream from
  // the record species by species, so that few values stay live.
  GV_MEM void flux_rev(int s, double dtin, const double (&lrsf)[NSP],
                       const double (&lmsf)[NSP], double lef, double lmfe, double lmfn) {
    // the face's forward values (:1459-1655)
    double RL = 0.0, RR = 0.0, PL = 0.0, PR = 0.0;
#pragma unroll
    for (int q = 0; q < NSP; q++) {
      RL = RL + R(s, RC_RSL + q);
      RR = RR + R(s, RC_RSR + q);
    }
#pragma unroll
    for (int q = 0; q < NSP; q++) {
      PL = PL + R(s, RC_PSL + q);
      PR = PR + R(s, RC_PSR + q);
    }
    const double iRL = 1.0 / RL, iRR = 1.0 / RR;
    double VUL = 0.0, VUR = 0.0, sRV_L = 0.0, sRV_R = 0.0;
#pragma unroll
    for (int q = 0; q < NSP; q++) {
      const double a = R(s, RC_RSL + q) * R(s, RC_VL + q);
      const double b = R(s, RC_RSR + q) * R(s, RC_VR + q);
      VUL = VUL + a * iRL;
      VUR = VUR + b * iRR;
      sRV_L = sRV_L + a;
      sRV_R = sRV_R + b;
    }
    const double VEL = R(s, RC_VEL), VER = R(s, RC_VER), VNL = R(s, RC_VNL),
                 VNR = R(s, RC_VNR);
    const double GL1 = st(A_GL1, j), GL2 = st(A_GL2, j), GL3 = st(A_GL3, j);
    const double GR1 = st(A_GR1, j), GR2 = st(A_GR2, j), GR3 = st(A_GR3, j);
    const double K2L = VUL * VUL + VEL * VEL + VNL * VNL;
    const double K2R = VUR * VUR + VER * VER + VNR * VNR;
    const double EL = GL1 * PL + 0.5 * RL * K2L;
    const double ER = GR1 * PR + 0.5 * RR * K2R;
    const double HtL = 0.5 * K2L + GL2 * PL * iRL;
    const double HtR = 0.5 * K2R + GR2 * PR * iRR;
    const double sL = sqrt(GL3 * HtL), sR = sqrt(GR3 * HtR);
    const bool denL = sL > VUL;         // dmax(sL, VUL) is sL
    const bool denR = sR > -1.0 * VUR;  // dmax(sR, -VUR) is sR
    const double CSL = (sL * sL) / (denL ? sL : VUL);
    const double CSR = (sR * sR) / (denR ? sR : -1.0 * VUR);
    const bool useL = CSL < CSR;
    const double CS = useL ? CSL : CSR;
    const double kp0 = st(A_KP, j), kp1 = st(A_KP, j + 1);
    const double ku0 = st(A_KU, j), ku1 = st(A_KU, j + 1);
    const double rdt0 = st(A_DC, j) / dtin, rdt1 = st(A_DC, j + 1) / dtin;
    const double RLR = RL + RR;
    const double iRLR = 1.0 / RLR, icS = 1.0 / CS;
    const int sub = (int)R(s, RC_SUB);
    const int subUp = sub_of(s, j + 1);
    // the bulk numerical velocities from every species' NV (recorded)
    double BNVp = 0.0, BNVm = 0.0;
#pragma unroll
    for (int q = 0; q < NSP; q++) {
      const double rs = R(s, RC_RSL + q) + R(s, RC_RSR + q);
      BNVp = BNVp + rs * R(s, RC_NVP + q) * iRLR;
      BNVm = BNVm + rs * R(s, RC_NVM + q) * iRLR;
    }
    // energy and momentum fluxes (:1973-2026): P12 side to this cell, M12
    // side to the cell above
    double lEFp = 0.0, lMOEp = 0.0, lMONp = 0.0;
    if (cell()) {
      const double ap = st(A_AP, j), ivol = 1.0 / st(A_VOL, j);
      lEFp = lef * ap * ivol;
      lMOEp = lmfe * ap * ivol;
      lMONp = lmfn * ap * ivol;
    }
    const double lEFm = xb(18, j + 1), lMOEm = xb(19, j + 1), lMONm = xb(20, j + 1);
    double lBNVp, lBNVm, lEL = 0.0, lER = 0.0, lPL = 0.0, lPR = 0.0, lRL = 0.0, lRR = 0.0;
    double lVEL = 0.0, lVER = 0.0, lVNL = 0.0, lVNR = 0.0;
    if (BNVp >= 0.0) {
      lBNVp = lEFp * (EL + PL) + lMOEp * RL * VEL + lMONp * RL * VNL;
      lEL += lEFp * BNVp;
      lPL += lEFp * BNVp;
      lRL += (lMOEp * VEL + lMONp * VNL) * BNVp;
      lVEL += lMOEp * RL * BNVp;
      lVNL += lMONp * RL * BNVp;
    } else {
      lBNVp = lEFp * (ER + PR) + lMOEp * RR * VER + lMONp * RR * VNR;
      lER += lEFp * BNVp;
      lPR += lEFp * BNVp;
      lRR += (lMOEp * VER + lMONp * VNR) * BNVp;
      lVER += lMOEp * RR * BNVp;
      lVNR += lMONp * RR * BNVp;
    }
    if (BNVm >= 0.0) {
      lBNVm = lEFm * (EL + PL) + lMOEm * RL * VEL + lMONm * RL * VNL;
      lEL += lEFm * BNVm;
      lPL += lEFm * BNVm;
      lRL += (lMOEm * VEL + lMONm * VNL) * BNVm;
      lVEL += lMOEm * RL * BNVm;
      lVNL += lMONm * RL * BNVm;
    } else {
      lBNVm = lEFm * (ER + PR) + lMOEm * RR * VER + lMONm * RR * VNR;
      lER += lEFm * BNVm;
      lPR += lEFm * BNVm;
      lRR += (lMOEm * VER + lMONm * VNR) * BNVm;
      lVER += lMOEm * RR * BNVm;
      lVNR += lMONm * RR * BNVm;
    }
    // per species, both sides (:1689-1970); partial adjoints to the work area
    double lCS = 0.0, licS = 0.0, liRLR = 0.0;
    const bool cl = cell();
    const double ap = st(A_AP, j), ivol = 1.0 / st(A_VOL, j), idC = 1.0 / st(A_DC, j);
    // the species' face states stream from the record, the next species'
    // loaded while this one computes
    double nRSL = R(s, RC_RSL), nRSR = R(s, RC_RSR), nPSL = R(s, RC_PSL),
           nPSR = R(s, RC_PSR), nHL = R(s, RC_HL), nHR = R(s, RC_HR), nVL = R(s, RC_VL),
           nVR = R(s, RC_VR), nNVp = R(s, RC_NVP), nNVm = R(s, RC_NVM);
#pragma unroll 1
    for (int q = 0; q < NSP; q++) {
      const double RSL = nRSL, RSR = nRSR, PSL = nPSL, PSR = nPSR, HL = nHL, HR = nHR,
                   VL = nVL, VR = nVR, NVq[2] = {nNVp, nNVm};
      if (q + 1 < NSP) {
        nRSL = R(s, RC_RSL + q + 1);
        nRSR = R(s, RC_RSR + q + 1);
        nPSL = R(s, RC_PSL + q + 1);
        nPSR = R(s, RC_PSR + q + 1);
        nHL = R(s, RC_HL + q + 1);
        nHR = R(s, RC_HR + q + 1);
        nVL = R(s, RC_VL + q + 1);
        nVR = R(s, RC_VR + q + 1);
        nNVp = R(s, RC_NVP + q + 1);
        nNVm = R(s, RC_NVM + q + 1);
      }
      double ML, MR, M2, MZ, FA, MF4P, MF4N, f1, dMF4P, dMF4N;
      mach_terms(VL, VR, icS, ML, MR, M2, MZ, FA, MF4P, MF4N, f1, dMF4P, dMF4N);
      const double pdh = (PSR - HR) - (PSL - HL);
      const double pd = PSR - PSL;
      const double rvd = RSR * VR - RSL * VL;
      const double MRS = 0.5 * (RSL + RSR);
      const double mcs = MRS * CS, facs = FA * CS;
      double lMPS = 0.0, lMHPS = 0.0, lrvd = 0.0, lMF4P = 0.0, lMF4N = 0.0, lf1 = 0.0;
      double lpdh = 0.0, lpd = 0.0, lmcs = 0.0, lfacs = 0.0;
      double lRSLq = 0.0, lRSRq = 0.0, lVLq = 0.0, lVRq = 0.0;
#pragma unroll
      for (int side = 0; side < 2; side++) {
        // the two sides differ in Kp, Ku, dAlt_C/DtIn, the flags and where
        // their adjoints come from
        const bool sh = ((side ? subUp : sub) >> q) & 1;
        const double kpx = side ? kp1 : kp0, kux = side ? ku1 : ku0;
        const double rdt = side ? rdt1 : rdt0;
        const double D = sh ? pdh : pd;
        const double Q = mcs * (facs + rdt), iQ = 1.0 / Q;
        const double MP = kpx * f1 * D * iQ;
        const double NV = NVq[side];
        double lFR, lFM, lNP;
        if (side) {
          lFR = xb(q, j + 1);
          lFM = xb(6 + q, j + 1);
          lNP = xb(12 + q, j + 1);
        } else {
          lFR = cl ? lrsf[q] * ap * ivol : 0.0;
          lFM = cl ? lmsf[q] * ap * ivol : 0.0;
          lNP = cl ? lmsf[q] * idC : 0.0;
        }
        const double lB = side ? lBNVm : lBNVp;
        // bulk velocity term (RSL + RSR) NV / (RL + RR)
        double lNV = lB * (RSL + RSR) * iRLR;
        lRSLq += lB * NV * iRLR;
        lRSRq += lB * NV * iRLR;
        liRLR += lB * (RSL + RSR) * NV;
        // upwind species fluxes
        if (NV >= 0.0) {
          lNV += lFR * RSL + lFM * RSL * VL;
          lRSLq += (lFR + lFM * VL) * NV;
          lVLq += lFM * RSL * NV;
        } else {
          lNV += lFR * RSR + lFM * RSR * VR;
          lRSRq += (lFR + lFM * VR) * NV;
          lVRq += lFM * RSR * NV;
        }
        // numerical pressure (MPS - MHPS | MPS) - Ku CS rvd / 2
        lMPS += lNP;
        if (sh) lMHPS -= lNP;
        lCS -= lNP * 0.5 * kux * rvd;
        lrvd -= lNP * 0.5 * kux * CS;
        // NV = CS (MF4P + MF4N - MP)
        lCS += lNV * (MF4P + MF4N - MP);
        lMF4P += lNV * CS;
        lMF4N += lNV * CS;
        const double lMP = -lNV * CS;
        // MP = Kp f1 D / (mcs (facs + rdt))
        lf1 += lMP * kpx * D * iQ;
        const double lD = lMP * kpx * f1 * iQ;
        const double lQ = -lMP * MP * iQ;
        lmcs += lQ * (facs + rdt);
        lfacs += lQ * mcs;
        if (sh)
          lpdh += lD;
        else
          lpd += lD;
      }
      // rvd = RSR VR - RSL VL; mcs = MRS CS, facs = FA CS
      lRSRq += lrvd * VR;
      lVRq += lrvd * RSR;
      lRSLq -= lrvd * VL;
      lVLq -= lrvd * RSL;
      const double lMRS = lmcs * CS;
      lCS += lmcs * MRS + lfacs * FA;
      const double lFA = lfacs * CS;
      lRSLq += 0.5 * lMRS;
      lRSRq += 0.5 * lMRS;
      // f1 = max(1 - M2, 0); FA = MZ (2 - MZ); MZ = sqrt(min(1, max(M2, 1e-19)))
      double lM2 = ((1.0 - M2) > 0.0) ? -lf1 : 0.0;
      const double lMZ = lFA * (2.0 - 2.0 * MZ);
      if (M2 > 1.0e-19 && !(1.0 < M2)) lM2 += lMZ * 0.5 / MZ;
      const double lML = lM2 * ML + lMF4P * dMF4P;
      const double lMR = lM2 * MR + lMF4N * dMF4N;
      lVLq += lML * icS;
      lVRq += lMR * icS;
      licS += lML * VL + lMR * VR;
      W(WK_PSR + q) = lpdh + lpd + 0.5 * lMPS;
      W(WK_PSL + q) = -lpdh - lpd + 0.5 * lMPS;
      W(WK_HR + q) = -lpdh + 0.5 * lMHPS;
      W(WK_HL + q) = lpdh + 0.5 * lMHPS;
      W(WK_RSL + q) = lRSLq;
      W(WK_RSR + q) = lRSRq;
      W(WK_VL + q) = lVLq;
      W(WK_VR + q) = lVRq;
    }
    lCS -= licS * icS * icS;
    const double lRLR = -liRLR * iRLR * iRLR;
    lRL += lRLR;
    lRR += lRLR;
    // CS = min(CSL, CSR), Liou's interface sound speed (:1619-1655)
    double lsL = 0.0, lsR = 0.0, lVUL = 0.0, lVUR = 0.0;
    if (useL) {
      if (denL) {
        lsL += lCS;
      } else {
        lsL += lCS * 2.0 * sL / VUL;
        lVUL -= lCS * CSL / VUL;
      }
    } else {
      if (denR) {
        lsR += lCS;
      } else {
        const double den = -1.0 * VUR;
        lsR += lCS * 2.0 * sR / den;
        lVUR += lCS * CSR / den;
      }
    }
    const double lHtL = lsL * 0.5 * GL3 / sL, lHtR = lsR * 0.5 * GR3 / sR;
    // enthalpy and energy (:1538-1615)
    const double lK2L = 0.5 * lHtL + lEL * 0.5 * RL, lK2R = 0.5 * lHtR + lER * 0.5 * RR;
    lPL += lHtL * GL2 * iRL + lEL * GL1;
    lPR += lHtR * GR2 * iRR + lER * GR1;
    double liRL = lHtL * GL2 * PL, liRR = lHtR * GR2 * PR;
    lRL += lEL * 0.5 * K2L;
    lRR += lER * 0.5 * K2R;
    lVUL += 2.0 * lK2L * VUL;
    lVUR += 2.0 * lK2R * VUR;
    lVEL += 2.0 * lK2L * VEL;
    lVNL += 2.0 * lK2L * VNL;
    lVER += 2.0 * lK2R * VER;
    lVNR += 2.0 * lK2R * VNR;
    // the face's gamma factors GL1 = 1/(gL-1), GL2 = gL/(gL-1), GL3 =
    // 2(gL-1)/(gL+1) (their adjoints go to Gamma through its face values)
    {
      const double lGL1 = lEL * PL, lGL2 = lHtL * PL * iRL, lGL3 = lsL * 0.5 * HtL / sL;
      const double lGR1 = lER * PR, lGR2 = lHtR * PR * iRR, lGR3 = lsR * 0.5 * HtR / sR;
      const double gLp = 2.0 + 1.0 / GL1, gRp = 2.0 + 1.0 / GR1;  // gamma + 1
      W(WK_GL) = -GL1 * GL1 * (lGL1 + lGL2) + lGL3 * 4.0 / (gLp * gLp);
      W(WK_GR) = -GR1 * GR1 * (lGR1 + lGR2) + lGR3 * 4.0 / (gRp * gRp);
    }
    // mass-weighted face winds (:1505-1532)
    liRL += lVUL * sRV_L;
    liRR += lVUR * sRV_R;
    lRL -= liRL * iRL * iRL;
    lRR -= liRR * iRR * iRR;
    // complete the species' adjoints; the exponentiated states take theirs
    // in log form (d/d ln x = x d/dx)
#pragma unroll 2
    for (int q = 0; q < NSP; q++) {
      const double RSL = R(s, RC_RSL + q), RSR = R(s, RC_RSR + q);
      const double VL = R(s, RC_VL + q), VR = R(s, RC_VR + q);
      const double PSL = R(s, RC_PSL + q), PSR = R(s, RC_PSR + q);
      const double HL = R(s, RC_HL + q), HR = R(s, RC_HR + q);
      const double wRSL = W(WK_RSL + q), wRSR = W(WK_RSR + q), wVL = W(WK_VL + q),
                   wVR = W(WK_VR + q), wPSL = W(WK_PSL + q), wPSR = W(WK_PSR + q),
                   wHL = W(WK_HL + q), wHR = W(WK_HR + q);
      W(WK_RSL + q) = (wRSL + lVUL * VL * iRL + lRL) * RSL;
      W(WK_RSR + q) = (wRSR + lVUR * VR * iRR + lRR) * RSR;
      W(WK_VL + q) = wVL + lVUL * RSL * iRL;
      W(WK_VR + q) = wVR + lVUR * RSR * iRR;
      W(WK_PSL + q) = (wPSL + lPL) * PSL;
      W(WK_PSR + q) = (wPSR + lPR) * PSR;
      W(WK_HL + q) = wHL * HL;
      W(WK_HR + q) = wHR * HR;
    }
    W(WK_VEL) = lVEL;
    W(WK_VER) = lVER;
    W(WK_VNL) = lVNL;
    W(WK_VNR) = lVNR;
  }

  // SubtractHydrostatic bits of the cell of lane jj in record stage s.
  GV_MEM int sub_of(int s, int jj) const {
    return (int)rec[((long)s * NREC + RC_SUB) * NT + jj];
  }

  // Mach-number terms of one species at a face (:1689-1819) and the
  // derivatives of the split Mach polynomials.
  GV_MEM static void mach_terms(double VL, double VR, double icS, double &ML, double &MR,
                                double &M2, double &MZ, double &FA, double &MF4P,
                                double &MF4N, double &f1, double &dMF4P, double &dMF4N) {
    ML = VL * icS;
    MR = VR * icS;
    M2 = 0.5 * (ML * ML + MR * MR);
    MZ = sqrt(dmin(1.0, dmax(M2, 1.0e-19)));
    FA = MZ * (2.0 - MZ);
    if (fabs(ML) >= 1.0) {
      MF4P = 0.5 * (ML + fabs(ML));
      dMF4P = ML > 0.0 ? 1.0 : 0.0;
    } else {
      const double A = 0.25 * ((ML + 1.0) * (ML + 1.0));
      const double B = 1.0 - 2.0 * (-0.25 * ((ML - 1.0) * (ML - 1.0)));
      MF4P = A * B;
      dMF4P = 0.5 * (ML + 1.0) * B + A * (ML - 1.0);
    }
    if (fabs(MR) >= 1.0) {
      MF4N = 0.5 * (MR - fabs(MR));
      dMF4N = MR > 0.0 ? 0.0 : 1.0;
    } else {
      const double A = -0.25 * ((MR - 1.0) * (MR - 1.0));
      const double B = 1.0 + 2.0 * (0.25 * ((MR + 1.0) * (MR + 1.0)));
      MF4N = A * B;
      dMF4N = -0.5 * (MR - 1.0) * B + A * (MR + 1.0);
    }
    f1 = dmax((1.0 - M2), 0.0);
  }

  // Reverse of the hydrostatic background.  lhns: adjoints of the
  // hydrostatic densities of this lane; adds to lNS (NS of level 0 on lane
  // 1) and lam[F_T] (the temperature ratios and mean temperatures).
  GV_MEM void hydro_rev(int s, const double (&hns)[NSP], const double (&lhns)[NSP], double T,
                        const double (&NS)[NSP], double (&lNS)[NSP]) {
    // suffix scan over lanes of G(j) = lhns(j) hns(j) on the product lanes
    // j >= 2: S(j) = sum_{i >= j} G(i), double-buffered in exchange slots
    // 0..5 and 6..11
    int buf = 0;
#pragma unroll
    for (int q = 0; q < NSP; q++) xb(q, j) = (j >= 2) ? lhns[q] * hns[q] : 0.0;
    cx.sync();
    for (int off = 1; off < L; off <<= 1) {
      const int src = buf * 6, dst = (1 - buf) * 6;
#pragma unroll
      for (int q = 0; q < NSP; q++) {
        double v = xb(src + q, j);
        if (j + off < L) v += xb(src + q, j + off);
        xb(dst + q, j) = v;
      }
      buf = 1 - buf;
      cx.sync();
    }
    double S[NSP], S2[NSP];
#pragma unroll
    for (int q = 0; q < NSP; q++) {
      S[q] = xb(buf * 6 + q, j);
      S2[q] = xb(buf * 6 + q, 2);
    }
    cx.sync();
    // a(j) = T(j-1)/T(j), log b_s(j) = -dF ish_s with ish_s = -m_s mgh /
    // (KB MT), MT = (T(j) + T(j-1))/2 (ausm :700-712); lane 0: a(0) =
    // T(1)/T(0), log b(0) = dF(0) ish0, ish0 = -MeanMass MG0 / (KB MT0)
    // (:716-724).  cnb: the contribution to the neighbour's Temp (lane j-1,
    // or lane 1 from lane 0).
    double cown = 0.0, cnb = 0.0;
    if (j >= 2) {
      const double tm1 = tprev(s);
      const double MT = 0.5 * (T + tm1);
      const double dF = st(A_DF, j);
      double la = 0.0, lMT = 0.0;
      const double ikmt = 1.0 / (KB * MT), imt = 1.0 / MT;
#pragma unroll
      for (int q = 0; q < NSP; q++) {
        la += S[q];
        const double ish = -1.0 * kp.mass[q] * c.S.mgh * ikmt;
        const double logb = -1.0 * dF * ish;
        lMT += S[q] * (-logb * imt);
      }
      cown = -la / T + 0.5 * lMT;
      cnb = la / tm1 + 0.5 * lMT;
    } else if (j == 0) {
      const double t1 = tnext(s);
      double g = 0.0;
#pragma unroll
      for (int q = 0; q < NSP; q++) {
        const double G0 = lhns[q] * hns[q];
        g += G0;
        xb(6 + q, 0) = G0;  // hns(0) = NS(level 0) a(0) b(0): for lane 1
      }
      const double MeanMass = 0.5 * (c.sc(SC_MMM_M1) + c.sc(SC_MMM_0));
      const double MT = 0.5 * (t1 + T);
      const double ish = -1.0 * MeanMass * c.sc(SC_MG0) / (KB * MT);
      const double logb = st(A_DF, 0) * ish;
      const double lMT = g * (-logb / MT);
      cown = -g / T + 0.5 * lMT;
      cnb = g / t1 + 0.5 * lMT;
      // MeanMass = (mmm(-1) + mmm(0))/2
      const double lmm = 0.5 * g * logb / MeanMass;
      fix[FX_MMM] += lmm;
      xb(13, 0) = lmm;
    }
    lam[F_T] += cown;
    xb(12, j) = cnb;
    cx.sync();
    if (j >= 1 && j + 1 < L) lam[F_T] += xb(12, j + 1);
    if (j == 1) {
      lam[F_T] += xb(12, 0);
      fix[FX_MMM] += xb(13, 0);
#pragma unroll
      for (int q = 0; q < NSP; q++) lNS[q] += lhns[q] + (S2[q] + xb(6 + q, 0)) / NS[q];
    }
    cx.sync();
  }

  // Temp of lanes j-1 and j+1 of the current stage's input (from the record).
  GV_MEM double tprev(int s) const { return rec[((long)s * NREC + RC_U + F_T) * NT + j - 1]; }
  GV_MEM double tnext(int s) const { return rec[((long)s * NREC + RC_U + F_T) * NT + j + 1]; }

  // Reverse of rusanov_grad_diff for Temp (gradient), Vel(E) and Vel(N)
  // (dissipation), calc_rusanov_alts_ausm (:1007-1090).
  GV_MEM void grad_diff_rev(const double (&U)[NF], double lgt, double ldfe, double ldfn) {
    const double v[3] = {U[F_T], U[F_VE], U[F_VN]};
#pragma unroll
    for (int f = 0; f < 3; f++) xb(f, j) = v[f];
    cx.sync();
    const double iF0 = st(A_IF, j), iF1 = st(A_IF, j + 1);
    const double dF0 = st(A_DF, j), dF1 = st(A_DF, j + 1);
    const bool cl = cell(), edge = (k == 0 || k == nA + 1);
    double dl[3], up[3], dn[3], rf[3], ln[3];
#pragma unroll
    for (int f = 0; f < 3; f++) {
      dl[f] = up[f] = dn[f] = 0.0;
      if (cl) {
        const double vm2 = xb(f, j - 2), vm1 = xb(f, j - 1), v1 = xb(f, j + 1),
                     v2 = xb(f, j + 2);
        up[f] = (iF1 * 2.0) * (FACT1 * (v1 - v[f]) - FACT2 * (v2 - vm1));
        dn[f] = (iF0 * 2.0) * (FACT1 * (v[f] - vm1) - FACT2 * (v1 - vm2));
        dl[f] = limiter_mc(kp.beta, up[f], dn[f]);
      } else if (edge) {
        up[f] = (xb(f, j + 1) - v[f]) * iF1;
        dn[f] = (v[f] - xb(f, j - 1)) * iF0;
        dl[f] = limiter_mc(kp.beta, up[f], dn[f]);
      }
      rf[f] = v[f] - 0.5 * dl[f] * dF0;
      ln[f] = v[f] + 0.5 * dl[f] * dF1;
    }
    cx.sync();
#pragma unroll
    for (int f = 0; f < 3; f++) {
      xb(f, j) = rf[f];
      xb(3 + f, j) = ln[f];
    }
    cx.sync();
    // adjoints of (L0, R0, L1, R1) = (ln(j-1), rf(j), ln(j), rf(j+1))
    const double dC = st(A_DC, j);
    const double cm0 = st(A_CMF, j), cm1 = st(A_CMF, j + 1);
    double lL0[3] = {0, 0, 0}, lR0[3] = {0, 0, 0}, lL1[3] = {0, 0, 0}, lR1[3] = {0, 0, 0};
    double lcm0 = 0.0, lcm1 = 0.0;
    if (cl) {
      const double idC = 1.0 / dC;
      const double g = 0.5 * lgt * idC;
      lL1[0] = g;
      lR1[0] = g;
      lL0[0] = -g;
      lR0[0] = -g;
      const double ld[3] = {0.0, ldfe, ldfn};
#pragma unroll
      for (int f = 1; f < 3; f++) {
        lR1[f] = cm1 * ld[f] * idC;
        lL1[f] = -cm1 * ld[f] * idC;
        lR0[f] = -cm0 * ld[f] * idC;
        lL0[f] = cm0 * ld[f] * idC;
        const double L0 = xb(3 + f, j - 1), R1 = xb(f, j + 1);
        lcm1 += ld[f] * (R1 - ln[f]) * idC;
        lcm0 -= ld[f] * (rf[f] - L0) * idC;
      }
    }
    cx.sync();
#pragma unroll
    for (int f = 0; f < 3; f++) {
      xb(f, j) = lR1[f];  // to rf of lane j+1
      xb(3 + f, j) = lL0[f];  // to ln of lane j-1
    }
    xb(6, j) = lcm0;  // CMF(j) = max(cmax(j-1), cmax(j))/2
    xb(7, j) = lcm1;  // CMF(j+1)
    cx.sync();
    double lv[3], cM2[3], cM1[3], cP1[3], cP2[3];
    double lcmf = 0.0;  // adjoint of CMF(j)
#pragma unroll
    for (int f = 0; f < 3; f++) {
      double lrf = lR0[f], lln = lL1[f];
      if (j >= 1) lrf += xb(f, j - 1);
      if (j + 1 < L) lln += xb(3 + f, j + 1);
      lv[f] = lrf + lln;
      const double ldl = -0.5 * dF0 * lrf + 0.5 * dF1 * lln;
      cM2[f] = cM1[f] = cP1[f] = cP2[f] = 0.0;
      if (cl || edge) {
        double lup, ldn;
        limiter_mc_rev(kp.beta, up[f], dn[f], ldl, lup, ldn);
        if (cl) {
          const double h1 = iF1 * 2.0, h0 = iF0 * 2.0;
          // up = h1 (F1 (v1 - v0) - F2 (v2 - vm1)), dn = h0 (F1 (v0 - vm1) - F2 (v1 - vm2))
          cP1[f] = lup * h1 * FACT1 - ldn * h0 * FACT2;
          lv[f] += -lup * h1 * FACT1 + ldn * h0 * FACT1;
          cP2[f] = -lup * h1 * FACT2;
          cM1[f] = lup * h1 * FACT2 - ldn * h0 * FACT1;
          cM2[f] = ldn * h0 * FACT2;
        } else {
          cP1[f] = lup * iF1;
          lv[f] += -lup * iF1 + ldn * iF0;
          cM1[f] = -ldn * iF0;
        }
      }
    }
    lcmf = xb(6, j) + ((j >= 1) ? xb(7, j - 1) : 0.0);
    cx.sync();
#pragma unroll
    for (int f = 0; f < 3; f++) {
      xb(4 * f + 0, j) = cM2[f];
      xb(4 * f + 1, j) = cM1[f];
      xb(4 * f + 2, j) = cP1[f];
      xb(4 * f + 3, j) = cP2[f];
    }
    // CMF(j) = 0.5 max(cmax(j-1), cmax(j)): the adjoint goes to the larger
    // (cmax(j) on ties, as dmax returns its second argument then)
    {
      const double cmj = cmax_of(j), cmm = cmax_of(j - 1);
      const bool down = cmm > cmj;
      xb(12, j) = down ? 0.5 * lcmf : 0.0;
      fix[FX_CMAX] += down ? 0.0 : 0.5 * lcmf;
    }
    cx.sync();
#pragma unroll
    for (int f = 0; f < 3; f++) {
      double a = lv[f];
      if (j - 2 >= 0) a += xb(4 * f + 3, j - 2);
      if (j - 1 >= 0) a += xb(4 * f + 2, j - 1);
      if (j + 1 < L) a += xb(4 * f + 1, j + 1);
      if (j + 2 < L) a += xb(4 * f + 0, j + 2);
      lv[f] = a;
    }
    if (j + 1 < L) fix[FX_CMAX] += xb(12, j + 1);
    lam[F_T] += lv[0];
    lam[F_VE] += lv[1];
    lam[F_VN] += lv[2];
    cx.sync();
  }

  // cMax of lane jj (the column's input, kept in LDS by the driver).
  GV_MEM double cmax_of(int jj) const { return (jj >= 0 && jj < L) ? ax[AX_CMAX + jj] : 0.0; }
};

}  // namespace gv

namespace gv {

// Adjoints of the column-prep inputs (optional output of the adjoint).
struct AdjFixFields {
  gitm_vert_field temp_unit, gamma, mean_major_mass, cmax, eddy;
};

// The adjoint of one AUSM step of column `col` of io: gather, forward
// recompute with the record, reverse, and the input adjoints written to
// gin (and the column-prep inputs' to *gfix when given).  Ctx is a TapCtx.
// lds: the column's LDS (adj_lds_bytes); rec: the column's record and work
// area (adj_column_doubles(NT)).
template <int NT, class Ctx>
GV_FN void adj_column(Ctx &cx, const KParams &kp, const gitm_vert_columns &io,
                      const AdjFields &gout, const AdjFields &gin, const AdjFixFields *gfix,
                      int col, double *lds, double *rec, int err_col) {
  using Col = Column<NT, Ctx, false>;
  static_assert(Col::CARRY_LDS, "the adjoint keeps the RK4 carries in LDS");
  const double dt = kp.dt;
  Col c(cx, kp, lds, err_col, dt);
  c.scr = nullptr;
  double *ax = lds + lds_doubles(kp.n_alts, false, true);
  const int nlb = io.n_lons * io.n_lats;
  const int ci = col % io.n_lons, cj = (col / io.n_lons) % io.n_lats, cb = col / nlb;
  const int k = c.k, nA = c.nA, j = c.j;
  double alt, grav, invr, dalt, mmm, gam, eddy = 0.0, cmax = 0.0, tu, coslat;
  double nds[NSP], temp_in, rho_in;
  {
    const gitm_vert_field &f = io.cos_latitude;
    coslat = f.ptr[col_base(f, ci, cj, cb)];
  }
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
  {
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
    temp_in = GV_LD(temperature);
    rho_in = GV_LD(rho);
  }
#undef GV_LD
  ax[AX_CMAX + j] = cmax;
  ax[AX_CMAX + c.L + j] = gam;
  c.setup(alt, grav, invr, dalt, mmm, gam, eddy, cmax, tu, coslat);
  bool bad = false;
#pragma unroll
  for (int s = 0; s < NSP; s++) bad = bad || !(nds[s] > 0.0);
  if (cx.sync_or(bad)) {
    if (bad) c.fail(GITM_VERT_ENONPOSITIVE, k);
    return;
  }
  c.U[F_LR] = ln(rho_in);
  c.U[F_T] = temp_in * tu;
#pragma unroll
  for (int s = 0; s < NSP; s++) c.U[F_LN + s] = ln(nds[s]);

  // forward: the step (vertical_solver_ausm.f90:130-420) with its record
  cx.rec = rec;
  cx.nt = NT;
  const double dtin[4] = {0.5 * dt, 0.5 * dt, dt, 0.5 * dt};
  double *bcrec = rec + ((long)ADJ_STAGES * NREC + NWORK) * NT;
  // each pass of the boundary conditions leaves its results in the LDS
  // scalars SC_BOT..SC_TOP+24: recorded for the reverse
  auto record_bcs = [&](int pass) {
    for (int i = j; i < ADJ_BC; i += c.L) bcrec[pass * ADJ_BC + i] = c.sc(SC_BOT + i);
  };
  c.bcs();
  record_bcs(0);
  if (c.cell()) {
#pragma unroll
    for (int f = 0; f < NF; f++) c.Yv(f) = c.U[f];
  }
#pragma unroll 1
  for (int s = 0; s < 4; s++) {
    cx.stage = s;
#pragma unroll
    for (int f = 0; f < NF; f++) cx.r(RC_U + f) = c.U[f];
    c.ausm_stage(dtin[s], s);
    c.bcs();
    record_bcs(s + 1);
  }
  cx.stage = 4;
#pragma unroll
  for (int f = 0; f < NF; f++) cx.r(RC_U + f) = c.U[f];
  cx.sync();

  // the adjoints of the outputs as advance_vertical scatters them (:163-185)
  AdjColumn<NT, Ctx> a(c, cx, rec, ax);
  {
    const long long bn = col_base(gout.nds, ci, cj, cb);
    const long long bw = col_base(gout.vvel, ci, cj, cb);
    const long long bv = col_base(gout.vel, ci, cj, cb);
#pragma unroll
    for (int s = 0; s < NSP; s++) {
      a.lam[F_LN + s] = *fld(gout.nds, bn, k, s) * exp(c.U[F_LN + s]);
      a.lam[F_W + s] = *fld(gout.vvel, bw, k, s);
    }
    a.lam[F_VE] = *fld(gout.vel, bv, k, 0);
    a.lam[F_VN] = *fld(gout.vel, bv, k, 1);
    a.lam[F_VU] = *fld(gout.vel, bv, k, 2);
    const double gt = *fld(gout.temp, col_base(gout.temp, ci, cj, cb), k, 0);
    a.lam[F_T] = gt / tu;
    a.fix[FX_TU] -= gt * c.U[F_T] / (tu * tu);
    a.lam[F_LR] = *fld(gout.rho, col_base(gout.rho, ci, cj, cb), k, 0) * exp(c.U[F_LR]);
  }
  // reverse: the boundary conditions after the last stage, then the RK4
  // combination stage by stage (Stage_{k+1} = Y + New_k - U_k; Final =
  // (acc + New_4 - U_4)/3 with acc = -Y + S2 + 2 S3 + S4 on cell levels)
  a.share_boundary_state(c.U);
  cx.sync();
  a.bcs_rev(4);
  // the RK4 combination's adjoints acc and Y live where the forward kept
  // the carries (LDS)
  const bool cl = c.cell();
  if (cl) {
#pragma unroll
    for (int f = 0; f < NF; f++) {
      c.Av(f) = a.lam[f] * (1.0 / 3.0);
      c.Yv(f) = 0.0;
      a.lam[f] = -c.Av(f);  // Final = (acc + New_4 - U_4)/3
    }
  }
#pragma unroll 1
  for (int s = 3; s >= 0; s--) {
    a.stage_rev(s, dtin[s]);
    if (s == 0) {
      if (cl) {
#pragma unroll
        for (int f = 0; f < NF; f++) a.lam[f] += c.Yv(f);
      }
      a.bcs_rev(0);
      break;
    }
    a.bcs_rev(s);
    // acc gained S_s with weight 1, 2, 1 for s = 3, 2, 1 (and started from
    // -Y at s = 1); S_s = Y + New_{s-1} - U_{s-1}
    const double wacc = (s == 2) ? 2.0 : 1.0;
    if (cl) {
#pragma unroll
      for (int f = 0; f < NF; f++) {
        const double la = a.lam[f] + wacc * c.Av(f);
        double y = c.Yv(f) + la;
        if (s == 1) y -= c.Av(f);
        c.Yv(f) = y;
        a.lam[f] = -la;
      }
    }
  }
  // the gather (:105-118): LogRho = log(Rho), Temp = Temperature TempUnit,
  // LogNS = log(NDensityS)
  {
    const long long bn = col_base(gin.nds, ci, cj, cb);
    const long long bw = col_base(gin.vvel, ci, cj, cb);
    const long long bv = col_base(gin.vel, ci, cj, cb);
#pragma unroll
    for (int s = 0; s < NSP; s++) {
      *fld(gin.nds, bn, k, s) = a.lam[F_LN + s] / nds[s];
      *fld(gin.vvel, bw, k, s) = a.lam[F_W + s];
    }
    *fld(gin.vel, bv, k, 0) = a.lam[F_VE];
    *fld(gin.vel, bv, k, 1) = a.lam[F_VN];
    *fld(gin.vel, bv, k, 2) = a.lam[F_VU];
    *fld(gin.temp, col_base(gin.temp, ci, cj, cb), k, 0) = a.lam[F_T] * tu;
    *fld(gin.rho, col_base(gin.rho, ci, cj, cb), k, 0) = a.lam[F_LR] / rho_in;
    a.fix[FX_TU] += a.lam[F_T] * temp_in;
  }
  if (gfix) {
    const AdjFixFields &g = *gfix;
    *fld(g.temp_unit, col_base(g.temp_unit, ci, cj, cb), k, 0) = a.fix[FX_TU];
    *fld(g.gamma, col_base(g.gamma, ci, cj, cb), k, 0) = a.fix[FX_GAMMA];
    *fld(g.mean_major_mass, col_base(g.mean_major_mass, ci, cj, cb), k, 0) = a.fix[FX_MMM];
    *fld(g.eddy, col_base(g.eddy, ci, cj, cb), k, 0) = a.fix[FX_EDDY];
    if (k >= 0 && k <= nA + 1)
      *fld(g.cmax, col_base(g.cmax, ci, cj, cb), k, 0) = a.fix[FX_CMAX];
  }
}

}  // namespace gv
