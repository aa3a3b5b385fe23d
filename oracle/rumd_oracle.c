/*
 * rumd_oracle.c -- CPU restatement of RUMD's per-timestep hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP path and the
 * CPU baseline timed beside it; nothing in the product (rumd_amd/) links,
 * loads or calls it.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg use it (through oracle/oracle.py).
 *
 * Two kinds of function live here:
 *   - FP64 restatements of the physics (pair force/energy/virial, leapfrog with
 *     Nose-Hoover, kinetic energy, sums), used to bound the FP32 kernels' error;
 *   - FP32-exact restatements (source order, no FMA contraction: the file is
 *     compiled with -ffp-contract=off) of the arithmetic whose results must
 *     match bit for bit: the neighbour-list criterion and cell binning, and the
 *     FP32 leapfrog update.
 *
 * Pinning (what checks each function; see tests/test_oracle.py):
 *   orc_forces_brute / orc_forces_nl / orc_forces_cells: RUMD-written golden
 *     configurations (Conf/KobAndersen/KA_N1000_T0.466, KA_N256_L6.0_T0.50,
 *     Conf/KobAndersenAlt/KA_N864_L9.0_T0.50AltMass; forces, pe, vir columns),
 *     LJ_N1000_rho0.80_T0.80 at r - v dt, closed forms for SF (F(rc)=0,
 *     u(rc)=0, -du/dr = F by finite differences), Newton's third law.
 *   orc_stress_brute: per particle, sxx + syy + szz = -vir (the pinned
 *     virial); summed, the strain derivatives of the pinned total energy
 *     (dU/de for x -> x (1 + e) and for the shears x -> x + e y etc., by
 *     Richardson-extrapolated central differences on an isolated cluster).
 *   orc_forces_form_brute: each form's energy against its defining formula
 *     in the physical parameters (sigma, epsilon, ...), -du/dr = F by finite
 *     differences, special cases that reduce to the pinned 12-6 path
 *     (gLJ 12-6 = LJA, IPL_n with n = 12 = IPL 12-6 with A6 = 0), the
 *     Buckingham minimum -epsilon at rm and the smooth LJ's vanishing force
 *     at Rcut.
 *   orc_bonds, orc_angles_cossq, orc_angles_sq, orc_dihedrals_ryckaert,
 *     orc_dihedrals_periodic: closed forms for one bond (harmonic, FENE), one
 *     angle (both forms) and the trans / cis dihedral (both forms; the
 *     periodic form's force is -sin psi times the gradient force),
 *     -grad E = F by finite differences of every coordinate of a small
 *     molecule, zero total force and torque.
 *   orc_nblist_*: brute force vs cell scan, symmetry, superset of the rc sphere.
 *   orc_nvt_step: closed form for one particle, Ps = 0 => NVE, time
 *     reversibility, the r(t+dt) = r + v(t+dt/2) dt relation held by the
 *     LJ_N1000 fixture.
 *   orc_kinetic_energy, orc_thermostat_update, orc_sum_f4: closed forms,
 *     integer-valued inputs.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define NUM_PARAM 12

/* ------------------------------------------------------------------ helpers */

static inline uint32_t type_of(const float* r4) {
  uint32_t t;
  memcpy(&t, r4 + 3, sizeof t);
  return t;
}

/* FP64 minimum image (RectangularSimulationBox::calculateDistance,
 * include/rumd/SimulationBox.h:111-125, in double). */
static inline double min_image_f64(const float* a, const float* b, const double L[3], double d[3]) {
  double s = 0.0;
  for (int k = 0; k < 3; ++k) {
    double x = (double)a[k] - (double)b[k];
    x -= L[k] * rint(x / L[k]);
    d[k] = x;
    s += x * x;
  }
  return s;
}

/* FP32 minimum image exactly as the reference writes it (uncontracted). */
static inline float min_image_d2_f32(const float* a, const float* b, const float L[3], const float iL[3]) {
  float dx = a[0] - b[0];
  float dy = a[1] - b[1];
  float dz = a[2] - b[2];
  dx -= L[0] * rintf(dx * iL[0]);
  dy -= L[1] * rintf(dy * iL[1]);
  dz -= L[2] * rintf(dz * iL[2]);
  return dx * dx + dy * dy + dz * dz;
}

/* ------------------------------------------------------------------ pair term
 * Pot_IPL_12_6::ComputeInteraction (include/rumd/PairPotential.h:161-168) and
 * fij (include/rumd/PairPotential.h:1198-1244) in FP64.  param: RUMD row
 * (0 = Rcut, 1 = A12, 2 = A6).  The shifts u(rc) and s(rc) rc are formed here
 * in FP64 from the same float A12, A6, Rcut (CopyParamsToGPU,
 * src/PairPotential.cu:82-125, does this in float). */
typedef struct {
  double rc, rc2, A12, A6, ucut, src; /* src = s(rc) * rc */
} pair_par;

static pair_par make_par(const float* param) {
  pair_par p;
  p.rc = (double)param[0];
  p.rc2 = p.rc * p.rc;
  p.A12 = (double)param[1];
  p.A6 = (double)param[2];
  if (p.rc > 0.0) {
    const double i2 = 1.0 / p.rc2, i6 = i2 * i2 * i2;
    p.ucut = i6 * (p.A12 * i6 + p.A6);
    p.src = i6 * i2 * (12.0 * p.A12 * i6 + 6.0 * p.A6) * p.rc;
  } else {
    p.ucut = 0.0;
    p.src = 0.0;
  }
  return p;
}

/* s = -u'(r)/r and u of one pair (fij, include/rumd/PairPotential.h:1198-1220).
 * Returns 0 at or beyond the cut-off. */
static inline int pair_su(const pair_par* p, int cutoff, double d2, double* s_out, double* e_out) {
  if (!(d2 < p->rc2)) return 0;
  const double i2 = 1.0 / d2, i6 = i2 * i2 * i2, t = p->A12 * i6;
  double s = i6 * i2 * (12.0 * t + 6.0 * p->A6);
  double e = i6 * (t + p->A6);
  if (cutoff == 1) {        /* SP */
    e -= p->ucut;
  } else if (cutoff == 2) { /* SF */
    const double r = sqrt(d2);
    s -= p->src / r;
    e += p->src * (r - p->rc) - p->ucut;
  }
  *s_out = s;
  *e_out = e;
  return 1;
}

/* Adds the (i,j) term to fi[0..2], u, vir.  Returns 1 if inside the cut-off. */
static inline int pair_add(const pair_par* p, int cutoff, const double d[3], double d2, double* fi, double* u,
                           double* vir) {
  double s, e;
  if (!pair_su(p, cutoff, d2, &s, &e)) return 0;
  fi[0] += d[0] * s;
  fi[1] += d[1] * s;
  fi[2] += d[2] * s;
  *u += e;
  *vir += d2 * s;
  return 1;
}

static pair_par* make_table(int nt, const float* params) {
  pair_par* t = (pair_par*)malloc(sizeof(pair_par) * (size_t)nt * nt);
  for (int k = 0; k < nt * nt; ++k) t[k] = make_par(params + (size_t)k * NUM_PARAM);
  return t;
}

/* ------------------------------------------------------------------ forces
 * Output per particle: f4[4i..4i+3] = (Fx, Fy, Fz, 1/2 sum_j u_ij),
 * vir[i] = sum_j r_ij^2 s_ij (RUMD's w.w; the xyz writer stores vir/2). */

/* O(N^2) over all pairs (the nbl == 0 branch of Calcf_NBL_tp). */
void orc_forces_brute(int n, const float* pos4, const double L[3], int nt, const float* params, int cutoff,
                      double* f4, double* vir) {
  pair_par* tab = make_table(nt, params);
#pragma omp parallel for schedule(dynamic, 16)
  for (int i = 0; i < n; ++i) {
    double fi[3] = {0, 0, 0}, u = 0, w = 0, d[3];
    const uint32_t ti = type_of(pos4 + 4 * (size_t)i);
    for (int j = 0; j < n; ++j) {
      if (j == i) continue;
      const double d2 = min_image_f64(pos4 + 4 * (size_t)i, pos4 + 4 * (size_t)j, L, d);
      pair_add(&tab[ti * nt + type_of(pos4 + 4 * (size_t)j)], cutoff, d, d2, fi, &u, &w);
    }
    f4[4 * (size_t)i] = fi[0];
    f4[4 * (size_t)i + 1] = fi[1];
    f4[4 * (size_t)i + 2] = fi[2];
    f4[4 * (size_t)i + 3] = 0.5 * u;
    vir[i] = w;
  }
  free(tab);
}

/* Per-particle configurational stress over all pairs, the STR branch of fij
 * (include/rumd/PairPotential.h:1232-1240): sts6[6i..6i+5] =
 * (sxx, syy, szz, syz, sxz, sxy), each -sum_j d_a d_b s_ij.  RUMD keeps the
 * first four in sts and sxz, sxy in w.y, w.z; Sample::GetStress
 * (src/Sample.cu:393-444) sums them, subtracts 2 m v_a v_b per particle and
 * scales by 1/(2V). */
void orc_stress_brute(int n, const float* pos4, const double L[3], int nt, const float* params, int cutoff,
                      double* sts6) {
  pair_par* tab = make_table(nt, params);
#pragma omp parallel for schedule(dynamic, 16)
  for (int i = 0; i < n; ++i) {
    double a[6] = {0, 0, 0, 0, 0, 0}, d[3], s, e;
    const uint32_t ti = type_of(pos4 + 4 * (size_t)i);
    for (int j = 0; j < n; ++j) {
      if (j == i) continue;
      const double d2 = min_image_f64(pos4 + 4 * (size_t)i, pos4 + 4 * (size_t)j, L, d);
      if (!pair_su(&tab[ti * nt + type_of(pos4 + 4 * (size_t)j)], cutoff, d2, &s, &e)) continue;
      a[0] -= d[0] * d[0] * s;
      a[1] -= d[1] * d[1] * s;
      a[2] -= d[2] * d[2] * s;
      a[3] -= d[1] * d[2] * s;
      a[4] -= d[0] * d[2] * s;
      a[5] -= d[0] * d[1] * s;
    }
    memcpy(sts6 + 6 * (size_t)i, a, sizeof(a));
  }
  free(tab);
}

/* s = -u'(r)/r and u of the pair forms outside the 12-6 family, in FP64 from
 * the float parameter row q (ComputeInteraction of each class; the form
 * numbers are those of the C ABI):
 *   1 Pot_LJ_12_6_smooth_2_4 include/rumd/PairPotential.h:249-255
 *   2 Pot_gLJ_m_n            :296-303
 *   3 Pot_Gauss              :409-414
 *   4 Pot_Buckingham         :451-460
 *   5 Pot_IPL_n              :627-634
 *   6 Pot_Yukawa             :829-838
 *   7 Pot_Exp                :1046-1052
 *   8 Pot_LJ_12_6_Gauss      :367-377
 *   9, 10, 11 Pot_IPL_12, Pot_IPL_9, Pot_IPL_18  :497-504, :540-548, :585-592
 *  12 Pot_IPL6_LN            :671-679
 *  13 Pot_LJ_LN              :722-731
 *  14 Pot_Girifalco          :771-785
 *  15 Pot_Dzugutov           :878-899
 *  16 Pot_Fermi_Jagla        :944-960
 *  17 Pot_Repulsive_Shoulder :997-1011
 *  18 Pot_SAAP               :1093-1116
 *  19 Pot_ShiftedForceCoulomb :1178-1185
 * Pot_Fermi_Jagla and Pot_Repulsive_Shoulder assign the pair energy to the
 * particle's sum instead of adding it (:958, :1010); here it is added, as for
 * every other form. */
static double form_su(int form, double d2, const float* q, double* u) {
  const double p1 = q[1], p2 = q[2], p3 = q[3], p4 = q[4], p5 = q[5];
  const double p6 = q[6], p7 = q[7], p8 = q[8], p9 = q[9];
  switch (form) {
    case 1: {
      const double x2 = d2 * p1, i2 = 1.0 / x2, i6 = i2 * i2 * i2;
      *u = p2 * (i6 * (i6 - 1.0) + p3 * x2 + p4 * x2 * x2);
      return p2 * p1 * (i2 * i6 * (12.0 * i6 - 6.0) - 2.0 * p3 - 4.0 * p4 * x2);
    }
    case 2: {
      const double i2 = 1.0 / (d2 * p1);
      const double im = pow(i2, p3), in = pow(i2, p4);
      *u = p5 * (2.0 * p4 * im - 2.0 * p3 * in);
      return p1 * p2 * (im - in) * i2;
    }
    case 3: {
      const double e = exp(-d2 * p1);
      *u = p2 * e;
      return 2.0 * p1 * p2 * e;
    }
    case 4: {
      const double i2 = 1.0 / (d2 * p1), i6 = i2 * i2 * i2, i1 = sqrt(i2);
      const double e = exp(p2 * (1.0 - 1.0 / i1));
      *u = p3 * e - p4 * i6;
      return p1 * p5 * (i1 * e - i6 * i2);
    }
    case 5: {
      const double i2 = 1.0 / (d2 * p1), in = pow(i2, 0.5 * p3);
      *u = p2 * in;
      return p1 * p2 * p3 * in * i2;
    }
    case 6: {
      const double r = sqrt(d2), v = p3 * p1 / r * exp(-p2 * r);
      *u = v;
      return (1.0 / r) * (1.0 / r + p2) * v;
    }
    case 7: {
      const double r = sqrt(d2), e = exp(-r * p1);
      *u = p2 * e;
      return p1 * p2 / r * e;
    }
    case 8: { /* eps ((s/r)^12 - 2 (s/r)^6) - eps eps0 exp(-(r - r0)^2 / (2 s0^2)) */
      const double r = sqrt(d2), sr2 = 1.0 / (d2 * p1), sr6 = pow(sr2, 3.0);
      const double g = exp(-(r - p4) * (r - p4) / p3);
      *u = p6 * (sr6 * sr6 - 2.0 * sr6) - p6 * p5 * g;
      return p1 * p2 * sr2 * (sr6 * sr6 - sr6) - p2 * p5 * (r - p4) * g / (6.0 * p3 * r);
    }
    case 9:
    case 10:
    case 11: { /* eps (s/r)^n with p2 = n eps */
      const double n = form == 9 ? 12.0 : form == 10 ? 9.0 : 18.0;
      const double srn = pow(d2 * p1, -0.5 * n);
      *u = p3 * srn;
      return p2 * srn / d2;
    }
    case 12: { /* eps (s/r)^6 / ln(r/s): p1 = eps, p2 = 1/s^2, p3 = eps/s^2 */
      const double x2 = d2 * p2, lx = 0.5 * log(x2), sr6 = pow(x2, -3.0);
      *u = p1 * sr6 / lx;
      return p3 * sr6 / x2 * (6.0 * lx + 1.0) / (lx * lx);
    }
    case 13: { /* eps a ((Rm/r)^12 - (b/a) (Rm/r)^6) / ln(r/R0) */
      const double lx = 0.5 * log(d2 * p2), x2 = d2 * p1, sr6 = pow(x2, -3.0);
      *u = p3 * (sr6 * sr6 - p5 * sr6) / lx;
      return p4 * sr6 / x2 / (lx * lx) * ((12.0 * lx + 1.0) * sr6 - p5 * (6.0 * lx + 1.0));
    }
    case 14: { /* s = r/R0; -alpha(...)^-3 terms + beta(...)^-9 terms */
      const double x = sqrt(d2 * p1);
      const double am = pow(x - 1.0, -3.0), ap = pow(x + 1.0, -3.0), a0 = pow(x, -3.0);
      *u = (-p2 * (am + ap - 2.0 * a0) + p3 * (am * am * am + ap * ap * ap - 2.0 * a0 * a0 * a0)) / x;
      const double bm = pow(x - 1.0, -4.0), bp = pow(x + 1.0, -4.0);
      return p1 * a0 *
             (-p2 * ((4.0 * x - 1.0) * bm + (4.0 * x + 1.0) * bp - 8.0 * a0) +
              p3 * ((10.0 * x - 1.0) * pow(x - 1.0, -10.0) + (10.0 * x + 1.0) * pow(x + 1.0, -10.0) -
                    20.0 * pow(x, -9.0)));
    }
    case 15: { /* A (r^-16 - B) e^{c/(r-a)} [r < a] + B e^{d/(r-b)}; p1..p6 = a, b, c, d, A, B */
      const double r = sqrt(d2);
      const double eb = exp(p4 / (r - p2));
      double s = p4 * p6 * eb / ((r - p2) * (r - p2) * r);
      *u = p6 * eb;
      if (r < p1) {
        const double ea = exp(p3 / (r - p1)), v1 = p5 * ea * (pow(r, -16.0) - p6);
        s += 16.0 * p5 * ea * pow(r, -18.0) + p3 * v1 / ((r - p1) * (r - p1) * r);
        *u += v1;
      }
      return s;
    }
    case 16: { /* eps0 ((a/r)^n + A0/(1 + e^{kA (r/a - A2)}) - B0/(1 + e^{kB (r/a - B2)})) */
      const double x = sqrt(d2 * p1), xn = pow(x, -p9);
      const double ea = exp(p3 * (x - p4)), eb = exp(p6 * (x - p7));
      *u = p8 * (xn + p2 / (1.0 + ea) - p5 / (1.0 + eb));
      return p8 * p1 * (p9 * xn / (x * x) + (p2 * p3 * ea / ((1.0 + ea) * (1.0 + ea)) -
                                              p5 * p6 * eb / ((1.0 + eb) * (1.0 + eb))) / x);
    }
    case 17: { /* (s/r)^14 + eps/2 (1 - tanh(k0 (r - s1))); p2 = eps/2 */
      const double r = sqrt(d2), x2 = d2 * p1, t = tanh(p3 * (r - p4));
      *u = pow(x2, -7.0) + p2 * (1.0 - t);
      return 14.0 * p1 * pow(x2, -8.0) + p2 * p3 * (1.0 - t * t) / r;
    }
    case 18: { /* eps (a0/x e^{a1 x + a6 x^2} + a2 e^{a3 x} + a4) / (1 + a5 x^6), x = r/s;
                * the force keeps the reference's eps/(s^2 r) prefactor (:1115) */
      const double r = sqrt(d2), x = sqrt(d2 * p8);
      const double g = p1 * exp(p2 * x + p7 * x * x), h = p3 * exp(p4 * x);
      const double num = g / x + h + p5, den = 1.0 + p6 * pow(x, 6.0);
      *u = p9 * num / den;
      const double dnum = -g / (x * x) + g * (p2 + 2.0 * p7 * x) / x + p4 * h;
      const double dden = 6.0 * p6 * pow(x, 5.0);
      return -p9 * p8 / r * (dnum * den - num * dden) / (den * den);
    }
    case 19: { /* eps/r + eps/Rc^2 (r - Rc) - eps/Rc; q[0] = Rc */
      const double r = sqrt(d2);
      *u = p1 / r + p3 * (r - (double)q[0]) - p2;
      return (p1 / d2 - p3) / r;
    }
    default:
      *u = 0.0;
      return 0.0;
  }
}

/* Forces of a pair form over all pairs, with fij's cut-off handling
 * (include/rumd/PairPotential.h:1198-1220) and the shifts u(rc), s(rc) rc of
 * CopyParamsToGPU (src/PairPotential.cu:82-125) formed in FP64.  Output as
 * orc_forces_brute. */
void orc_forces_form_brute(int n, const float* pos4, const double L[3], int nt, const float* params, int form,
                           int cutoff, double* f4, double* vir) {
#pragma omp parallel for schedule(dynamic, 16)
  for (int i = 0; i < n; ++i) {
    double fi[3] = {0, 0, 0}, ui = 0, wi = 0, d[3];
    const uint32_t ti = type_of(pos4 + 4 * (size_t)i);
    for (int j = 0; j < n; ++j) {
      if (j == i) continue;
      const float* q = params + (size_t)(ti * nt + type_of(pos4 + 4 * (size_t)j)) * NUM_PARAM;
      const double rc = q[0];
      const double d2 = min_image_f64(pos4 + 4 * (size_t)i, pos4 + 4 * (size_t)j, L, d);
      if (!(d2 < rc * rc)) continue;
      double e, ec;
      double s = form_su(form, d2, q, &e);
      const double src = form_su(form, rc * rc, q, &ec) * rc;
      if (cutoff == 1) {
        e -= ec;
      } else if (cutoff == 2) {
        const double r = sqrt(d2);
        s -= src / r;
        e += src * (r - rc) - ec;
      }
      fi[0] += d[0] * s;
      fi[1] += d[1] * s;
      fi[2] += d[2] * s;
      ui += e;
      wi += d2 * s;
    }
    f4[4 * (size_t)i] = fi[0];
    f4[4 * (size_t)i + 1] = fi[1];
    f4[4 * (size_t)i + 2] = fi[2];
    f4[4 * (size_t)i + 3] = 0.5 * ui;
    vir[i] = wi;
  }
}

/* Bonds in FP64 (BondHarmonic and BondFENE, src/BondPotential.cu:281-430):
 * for each bond (a, b, type) with params (p0, k), harmonic u = k/2 (r - p0)^2
 * or (fene != 0) u = -k p0^2/2 ln(1 - (r/p0)^2); adds the forces to f4, half
 * of u to each end's f4[.,3] and F_a . (r_a - r_b) to each end's vir. */
void orc_bonds(int n, const float* pos4, const double L[3], int nbonds, const int32_t* bonds, const float* params,
               int fene, double* f4, double* vir) {
  (void)n;
  for (int q = 0; q < nbonds; ++q) {
    const int a = bonds[3 * q], b = bonds[3 * q + 1], t = bonds[3 * q + 2];
    const double l0 = params[2 * t], k = params[2 * t + 1];
    double d[3];
    const double r = sqrt(min_image_f64(pos4 + 4 * (size_t)a, pos4 + 4 * (size_t)b, L, d));
    double u, g; /* F_a = g d, d = r_a - r_b */
    if (fene) {  /* u = -k R0^2/2 ln(1 - (r/R0)^2), R0 = l0 */
      const double x = (r / l0) * (r / l0);
      u = -0.5 * k * l0 * l0 * log(1.0 - x);
      g = -k / (1.0 - x);
    } else {
      u = 0.5 * k * (r - l0) * (r - l0);
      g = -k * (r - l0) / r;
    }
    for (int c = 0; c < 3; ++c) {
      f4[4 * (size_t)a + c] += g * d[c];
      f4[4 * (size_t)b + c] -= g * d[c];
    }
    f4[4 * (size_t)a + 3] += 0.5 * u;
    f4[4 * (size_t)b + 3] += 0.5 * u;
    vir[a] += g * r * r;
    vir[b] += g * r * r;
  }
}

/* Bond angles cos^2 in FP64 (AngleCosSq, src/AnglePotential.cu:137-217), from
 * the angle theta at b between e_a = (r_a - r_b)/|.| and e_c = (r_c - r_b)/|.|:
 * u = k/2 (cos theta - cos theta0)^2, F_a = -k (cos theta - cos theta0)
 * (e_c - cos theta e_a) / |r_a - r_b|, F_c likewise, F_b = -F_a - F_c.  The whole
 * u goes to atom a's f4[.,3], as in the reference. */
void orc_angles_cossq(int n, const float* pos4, const double L[3], int nangles, const int32_t* angles,
                      const float* params, double* f4) {
  (void)n;
  for (int q = 0; q < nangles; ++q) {
    const int a = angles[4 * q], b = angles[4 * q + 1], c = angles[4 * q + 2], t = angles[4 * q + 3];
    const double th0 = params[2 * t], k = params[2 * t + 1];
    double da[3], dc[3];
    const double ra = sqrt(min_image_f64(pos4 + 4 * (size_t)a, pos4 + 4 * (size_t)b, L, da));
    const double rc = sqrt(min_image_f64(pos4 + 4 * (size_t)c, pos4 + 4 * (size_t)b, L, dc));
    double cs = 0.0;
    for (int m = 0; m < 3; ++m) {
      da[m] /= ra;
      dc[m] /= rc;
      cs += da[m] * dc[m];
    }
    const double dcos = cs - cos(th0);
    for (int m = 0; m < 3; ++m) {
      const double fa = -k * dcos * (dc[m] - cs * da[m]) / ra;
      const double fc = -k * dcos * (da[m] - cs * dc[m]) / rc;
      f4[4 * (size_t)a + m] += fa;
      f4[4 * (size_t)c + m] += fc;
      f4[4 * (size_t)b + m] -= fa + fc;
    }
    f4[4 * (size_t)a + 3] += 0.5 * k * dcos * dcos;
  }
}

/* Squared bond angles in FP64 (AngleSq, SquaredAngle, src/AnglePotential.cu:289-373):
 * theta = acos(e_a . e_c) at b, u = k/2 (theta - theta0)^2 to atom a's
 * f4[.,3]; F_a = k (theta - theta0) / (1e-6 + sin theta) (e_c - cos theta e_a)
 * / |r_a - r_b|, F_c likewise, F_b = -F_a - F_c.  The 1e-6 is the reference's
 * regularisation of the straight angle. */
void orc_angles_sq(int n, const float* pos4, const double L[3], int nangles, const int32_t* angles,
                   const float* params, double* f4) {
  (void)n;
  for (int q = 0; q < nangles; ++q) {
    const int a = angles[4 * q], b = angles[4 * q + 1], c = angles[4 * q + 2], t = angles[4 * q + 3];
    const double th0 = params[2 * t], k = params[2 * t + 1];
    double da[3], dc[3];
    const double ra = sqrt(min_image_f64(pos4 + 4 * (size_t)a, pos4 + 4 * (size_t)b, L, da));
    const double rc = sqrt(min_image_f64(pos4 + 4 * (size_t)c, pos4 + 4 * (size_t)b, L, dc));
    double cs = 0.0;
    for (int m = 0; m < 3; ++m) {
      da[m] /= ra;
      dc[m] /= rc;
      cs += da[m] * dc[m];
    }
    if (cs > 1.0) cs = 1.0;
    if (cs < -1.0) cs = -1.0;
    const double th = acos(cs), g = k * (th - th0) / (1e-6 + sin(th));
    for (int m = 0; m < 3; ++m) {
      const double fa = g * (dc[m] - cs * da[m]) / ra;
      const double fc = g * (da[m] - cs * dc[m]) / rc;
      f4[4 * (size_t)a + m] += fa;
      f4[4 * (size_t)c + m] += fc;
      f4[4 * (size_t)b + m] -= fa + fc;
    }
    f4[4 * (size_t)a + 3] += 0.5 * k * (th - th0) * (th - th0);
  }
}

static void cross3(const double* x, const double* y, double* z) {
  z[0] = x[1] * y[2] - x[2] * y[1];
  z[1] = x[2] * y[0] - x[0] * y[2];
  z[2] = x[0] * y[1] - x[1] * y[0];
}

/* Ryckaert-Bellemans dihedrals in FP64 (DihedralRyckaert, src/DihedralPotential.cu:150-260),
 * from the IUPAC torsion phi of (a, b, c, d): b1 = r_b - r_a, b2 = r_c - r_b,
 * b3 = r_d - r_c, m = b1 x b2, n = b2 x b3, phi = atan2(|b2| b1.n, m.n)
 * (0 = cis); psi = phi - pi, u = sum_k c_k cos^k(psi).  Forces from
 * dphi/dr_a = -|b2| m / |m|^2, dphi/dr_d = |b2| n / |n|^2 and the two inner
 * atoms' terms (Blondel & Karplus, J. Comput. Chem. 17, 1132 (1996)).  The
 * energy goes to atom a; nearly collinear bonds (|cos| within 1e-3 of 1 on
 * either side) contribute nothing, as in the reference.
 * periodic != 0: PeriodicDihedral (PeriodicDih, src/DihedralPotential.cu:337-458),
 * params (psi0, k1, k2, k3) of six per type, psi = acos(cos psi) in [0, pi],
 * u = 1/2 sum_m k_m (1 - cos m (psi - psi0)); the reference puts du/dpsi where
 * the Ryckaert form has du/dcos psi, so the force is du/dpsi sin(phi) dphi/dr. */
static void dihedrals(int ndih, const int32_t* dih, const float* pos4, const double L[3], const float* params,
                      int periodic, double* f4) {
  for (int q = 0; q < ndih; ++q) {
    const int a = dih[5 * q], b = dih[5 * q + 1], c = dih[5 * q + 2], d = dih[5 * q + 3], t = dih[5 * q + 4];
    const float* cf = params + 6 * t;
    double b1[3], b2[3], b3[3], m[3], nn[3];
    min_image_f64(pos4 + 4 * (size_t)b, pos4 + 4 * (size_t)a, L, b1);
    min_image_f64(pos4 + 4 * (size_t)c, pos4 + 4 * (size_t)b, L, b2);
    min_image_f64(pos4 + 4 * (size_t)d, pos4 + 4 * (size_t)c, L, b3);
    const double l1 = b1[0] * b1[0] + b1[1] * b1[1] + b1[2] * b1[2];
    const double l2 = b2[0] * b2[0] + b2[1] * b2[1] + b2[2] * b2[2];
    const double l3 = b3[0] * b3[0] + b3[1] * b3[1] + b3[2] * b3[2];
    const double d12 = b1[0] * b2[0] + b1[1] * b2[1] + b1[2] * b2[2];
    const double d23 = b2[0] * b3[0] + b2[1] * b3[1] + b2[2] * b3[2];
    if (fabs(d23 * d23 / (l2 * l3) - 1.0) < 1e-3 || fabs(d12 * d12 / (l1 * l2) - 1.0) < 1e-3) continue;
    cross3(b1, b2, m);
    cross3(b2, b3, nn);
    const double mm = m[0] * m[0] + m[1] * m[1] + m[2] * m[2];
    const double nq = nn[0] * nn[0] + nn[1] * nn[1] + nn[2] * nn[2];
    const double lb2 = sqrt(l2);
    const double y = lb2 * (b1[0] * nn[0] + b1[1] * nn[1] + b1[2] * nn[2]);
    const double x = m[0] * nn[0] + m[1] * nn[1] + m[2] * nn[2];
    const double phi = atan2(y, x);
    const double cpsi = -cos(phi); /* cos(phi - pi) */
    double u = 0.0, dudc = 0.0, pw = 1.0;
    if (periodic) {
      const double x = acos(cpsi) - cf[0];
      u = 0.5 * (cf[1] * (1.0 - cos(x)) + cf[2] * (1.0 - cos(2.0 * x)) + cf[3] * (1.0 - cos(3.0 * x)));
      dudc = 0.5 * (cf[1] * sin(x) + 2.0 * cf[2] * sin(2.0 * x) + 3.0 * cf[3] * sin(3.0 * x));
    } else {
      for (int k = 0; k < 6; ++k) {
        u += cf[k] * pw;
        if (k < 5) dudc += (k + 1) * cf[k + 1] * pw;
        pw *= cpsi;
      }
    }
    /* du/dphi = du/dcpsi * dcpsi/dphi = dudc * sin(phi) */
    const double dudphi = dudc * sin(phi);
    double ga[3], gd[3];
    for (int k = 0; k < 3; ++k) {
      ga[k] = -lb2 * m[k] / mm;
      gd[k] = lb2 * nn[k] / nq;
    }
    const double s1 = d12 / l2, s3 = d23 / l2;
    for (int k = 0; k < 3; ++k) {
      const double gb = -(1.0 + s1) * ga[k] + s3 * gd[k];
      const double gc = s1 * ga[k] - (1.0 + s3) * gd[k];
      f4[4 * (size_t)a + k] -= dudphi * ga[k];
      f4[4 * (size_t)b + k] -= dudphi * gb;
      f4[4 * (size_t)c + k] -= dudphi * gc;
      f4[4 * (size_t)d + k] -= dudphi * gd[k];
    }
    f4[4 * (size_t)a + 3] += u;
  }
}

void orc_dihedrals_ryckaert(int n, const float* pos4, const double L[3], int ndih, const int32_t* dih,
                            const float* params, double* f4) {
  (void)n;
  dihedrals(ndih, dih, pos4, L, params, 0, f4);
}

void orc_dihedrals_periodic(int n, const float* pos4, const double L[3], int ndih, const int32_t* dih,
                            const float* params, double* f4) {
  (void)n;
  dihedrals(ndih, dih, pos4, L, params, 1, f4);
}

/* Over a given neighbour list in CSR form (offsets[n+1], idx). */
void orc_forces_nl(int n, const float* pos4, const double L[3], int nt, const float* params, int cutoff,
                   const int64_t* offsets, const int32_t* idx, double* f4, double* vir) {
  pair_par* tab = make_table(nt, params);
#pragma omp parallel for schedule(dynamic, 256)
  for (int i = 0; i < n; ++i) {
    double fi[3] = {0, 0, 0}, u = 0, w = 0, d[3];
    const uint32_t ti = type_of(pos4 + 4 * (size_t)i);
    for (int64_t k = offsets[i]; k < offsets[i + 1]; ++k) {
      const int j = idx[k];
      const double d2 = min_image_f64(pos4 + 4 * (size_t)i, pos4 + 4 * (size_t)j, L, d);
      pair_add(&tab[ti * nt + type_of(pos4 + 4 * (size_t)j)], cutoff, d, d2, fi, &u, &w);
    }
    f4[4 * (size_t)i] = fi[0];
    f4[4 * (size_t)i + 1] = fi[1];
    f4[4 * (size_t)i + 2] = fi[2];
    f4[4 * (size_t)i + 3] = 0.5 * u;
    vir[i] = w;
  }
  free(tab);
}

/* ------------------------------------------------------------------ cell lists (FP64)
 * Linked-cell bins of width >= rcut_max (27-cell stencil).  Used for FP64 forces
 * at large N and by the CPU baseline; a plain acceleration of the all-pairs sum
 * (requires at least 3 cells per axis). */
typedef struct {
  int nc[3];
  int* head; /* per cell, first particle or -1 */
  int* next; /* per particle */
} cell_list;

static int cell_index_f64(const double x[3], const double L[3], const int nc[3]) {
  int c[3];
  for (int k = 0; k < 3; ++k) {
    double s = x[k] / L[k] + 0.5;
    s -= floor(s);
    int q = (int)(s * nc[k]);
    if (q >= nc[k]) q = nc[k] - 1;
    if (q < 0) q = 0;
    c[k] = q;
  }
  return c[0] + nc[0] * (c[1] + nc[1] * c[2]);
}

static int build_cells(cell_list* cl, int n, const float* pos4, const double L[3], double width) {
  for (int k = 0; k < 3; ++k) {
    cl->nc[k] = (int)floor(L[k] / width);
    if (cl->nc[k] < 3) return -1;
  }
  const size_t nc = (size_t)cl->nc[0] * cl->nc[1] * cl->nc[2];
  cl->head = (int*)malloc(sizeof(int) * nc);
  cl->next = (int*)malloc(sizeof(int) * (size_t)n);
  for (size_t c = 0; c < nc; ++c) cl->head[c] = -1;
  for (int i = n - 1; i >= 0; --i) {
    const double x[3] = {pos4[4 * (size_t)i], pos4[4 * (size_t)i + 1], pos4[4 * (size_t)i + 2]};
    const int c = cell_index_f64(x, L, cl->nc);
    cl->next[i] = cl->head[c];
    cl->head[c] = i;
  }
  return 0;
}

static void free_cells(cell_list* cl) {
  free(cl->head);
  free(cl->next);
}

/* FP64 forces through a linked-cell list; returns -1 if the box is too small. */
int orc_forces_cells(int n, const float* pos4, const double L[3], int nt, const float* params, int cutoff,
                     double* f4, double* vir) {
  pair_par* tab = make_table(nt, params);
  double rmax = 0.0;
  for (int k = 0; k < nt * nt; ++k)
    if (tab[k].rc > rmax) rmax = tab[k].rc;
  cell_list cl;
  if (build_cells(&cl, n, pos4, L, rmax) != 0) {
    free(tab);
    return -1;
  }
#pragma omp parallel for schedule(dynamic, 256)
  for (int i = 0; i < n; ++i) {
    double fi[3] = {0, 0, 0}, u = 0, w = 0, d[3];
    const float* ri = pos4 + 4 * (size_t)i;
    const uint32_t ti = type_of(ri);
    const double x[3] = {ri[0], ri[1], ri[2]};
    const int c = cell_index_f64(x, L, cl.nc);
    const int cx = c % cl.nc[0], cy = (c / cl.nc[0]) % cl.nc[1], cz = c / (cl.nc[0] * cl.nc[1]);
    for (int dz = -1; dz <= 1; ++dz)
      for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
          const int ox = (cx + dx + cl.nc[0]) % cl.nc[0];
          const int oy = (cy + dy + cl.nc[1]) % cl.nc[1];
          const int oz = (cz + dz + cl.nc[2]) % cl.nc[2];
          for (int j = cl.head[ox + cl.nc[0] * (oy + cl.nc[1] * oz)]; j >= 0; j = cl.next[j]) {
            if (j == i) continue;
            const double d2 = min_image_f64(ri, pos4 + 4 * (size_t)j, L, d);
            pair_add(&tab[ti * nt + type_of(pos4 + 4 * (size_t)j)], cutoff, d, d2, fi, &u, &w);
          }
        }
    f4[4 * (size_t)i] = fi[0];
    f4[4 * (size_t)i + 1] = fi[1];
    f4[4 * (size_t)i + 2] = fi[2];
    f4[4 * (size_t)i + 3] = 0.5 * u;
    vir[i] = w;
  }
  free_cells(&cl);
  free(tab);
  return 0;
}

/* ------------------------------------------------------------------ neighbour lists (FP32 exact)
 * box8: RUMD float[8] box.  cut_skin2[ti*nt+tj] = (rc_ij + skin)^2 formed in
 * float (src/NeighborList.cu:1239-1240).  Results are CSR over the input order:
 * counts[n] then, when idx != NULL, offsets/idx.  Entries inside a row are in
 * ascending j for the brute list and in RUMD's stencil order for the cell list. */

static void box_f32(const float box8[8], float L[3], float iL[3]) {
  const int cubic = ((unsigned)box8[0]) != 0u;
  for (int k = 0; k < 3; ++k) {
    L[k] = cubic ? box8[1] : box8[1 + k];
    iL[k] = cubic ? box8[4] : box8[4 + k];
  }
}

/* calculateNBL_N2 (src/NeighborList.cu:1001-1099): all j != i with
 * d^2 < (rc_ij + skin)^2, FP32. */
void orc_nblist_brute_f32(int n, const float* pos4, const float box8[8], int nt, const float* cut_skin2,
                          int32_t* counts, const int64_t* offsets, int32_t* idx) {
  float L[3], iL[3];
  box_f32(box8, L, iL);
#pragma omp parallel for schedule(dynamic, 16)
  for (int i = 0; i < n; ++i) {
    const uint32_t ti = type_of(pos4 + 4 * (size_t)i);
    int c = 0;
    for (int j = 0; j < n; ++j) {
      if (j == i) continue;
      const float d2 = min_image_d2_f32(pos4 + 4 * (size_t)i, pos4 + 4 * (size_t)j, L, iL);
      if (d2 < cut_skin2[ti * nt + type_of(pos4 + 4 * (size_t)j)]) {
        if (idx) idx[offsets[i] + c] = j;
        ++c;
      }
    }
    counts[i] = c;
  }
}

/* Cell grid of CalculateNBL_Sorting (src/NeighborList.cu:717-726), FP32. */
void orc_cell_grid(const float box8[8], float max_cutoff, float skin, int32_t nc[3]) {
  float L[3], iL[3];
  box_f32(box8, L, iL);
  for (int k = 0; k < 3; ++k) {
    int q = (int)floorf(2.0f * L[k] / (max_cutoff + skin));
    nc[k] = q < 5 ? 5 : q;
  }
}

/* calculateCellCoordinates (src/NeighborList.cu:1106-1133), FP32. */
static int cell_coord_f32(float x, float iL, int nc) {
  const float s = x * iL + 0.5f;
  int c = (int)((float)nc * s);
  if (c >= nc) c = nc - 1;
  if (c < 0) c = 0;
  return c;
}

void orc_cell_of(int n, const float* pos4, const float box8[8], const int32_t nc[3], int32_t* lin) {
  float L[3], iL[3];
  box_f32(box8, L, iL);
  for (int i = 0; i < n; ++i) {
    const float* r = pos4 + 4 * (size_t)i;
    const int cx = cell_coord_f32(r[0], iL[0], nc[0]);
    const int cy = cell_coord_f32(r[1], iL[1], nc[1]);
    const int cz = cell_coord_f32(r[2], iL[2], nc[2]);
    lin[i] = cx + cy * nc[0] + cz * nc[0] * nc[1];
  }
}

/* calculateNBL_CellsSorted (src/NeighborList.cu:1204-1319): bin by cell, scan
 * the 5x5x5 stencil in (dZ, dY, dX) order, particles of a cell in ascending
 * index, FP32 criterion. */
void orc_nblist_cells_f32(int n, const float* pos4, const float box8[8], int nt, const float* cut_skin2,
                          const int32_t nc[3], int32_t* counts, const int64_t* offsets, int32_t* idx) {
  float L[3], iL[3];
  box_f32(box8, L, iL);
  const int ncell = nc[0] * nc[1] * nc[2];
  int32_t* lin = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
  int32_t* start = (int32_t*)calloc((size_t)ncell + 1, sizeof(int32_t));
  int32_t* order = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
  orc_cell_of(n, pos4, box8, nc, lin);
  for (int i = 0; i < n; ++i) start[lin[i] + 1]++;
  for (int c = 0; c < ncell; ++c) start[c + 1] += start[c];
  {
    int32_t* cur = (int32_t*)malloc(sizeof(int32_t) * (size_t)ncell);
    memcpy(cur, start, sizeof(int32_t) * (size_t)ncell);
    for (int i = 0; i < n; ++i) order[cur[lin[i]]++] = i; /* stable: ascending i inside a cell */
    free(cur);
  }
#pragma omp parallel for schedule(dynamic, 64)
  for (int i = 0; i < n; ++i) {
    const float* ri = pos4 + 4 * (size_t)i;
    const uint32_t ti = type_of(ri);
    const int cx = cell_coord_f32(ri[0], iL[0], nc[0]);
    const int cy = cell_coord_f32(ri[1], iL[1], nc[1]);
    const int cz = cell_coord_f32(ri[2], iL[2], nc[2]);
    int cnt = 0;
    for (int dz = -2; dz <= 2; ++dz) {
      const int oz = (cz + dz + nc[2]) % nc[2];
      for (int dy = -2; dy <= 2; ++dy) {
        const int oy = (cy + dy + nc[1]) % nc[1];
        for (int dx = -2; dx <= 2; ++dx) {
          const int ox = (cx + dx + nc[0]) % nc[0];
          const int c = ox + oy * nc[0] + oz * nc[0] * nc[1];
          for (int k = start[c]; k < start[c + 1]; ++k) {
            const int j = order[k];
            if (j == i) continue;
            const float d2 = min_image_d2_f32(ri, pos4 + 4 * (size_t)j, L, iL);
            if (d2 < cut_skin2[ti * nt + type_of(pos4 + 4 * (size_t)j)]) {
              if (idx) idx[offsets[i] + cnt] = j;
              ++cnt;
            }
          }
        }
      }
    }
    counts[i] = cnt;
  }
  free(lin);
  free(start);
  free(order);
}

/* displacement_since_nb_build (src/NeighborList.cu:970-997): 1 if any particle
 * moved more than skin/2 since last_r (strict >, FP32). */
int orc_skin_exceeded(int n, const float* pos4, const float* last4, const float box8[8], float skin) {
  float L[3], iL[3];
  box_f32(box8, L, iL);
  const float s24 = (float)(0.25 * (double)skin * (double)skin);
  int any = 0;
  for (int i = 0; i < n; ++i)
    if (min_image_d2_f32(pos4 + 4 * (size_t)i, last4 + 4 * (size_t)i, L, iL) > s24) any = 1;
  return any;
}

/* ------------------------------------------------------------------ integrator
 * integrateNVTAlgorithm (src/IntegratorNVT.cu:364-418), FP32 in source order.
 * rlo4 != NULL: positions advanced and wrapped in FP64 from hi + lo and stored
 * back as a hi/lo float pair (the FP64 position bookkeeping variant).
 * ke2_out (may be NULL) receives sum over thermostatted particles of
 * (v.v)/v.w in FP64 (the thermostat's kinetic-energy sum). */
void orc_nvt_step(int n, float* pos4, float* vel4, const float* f4, float* im4, float* rlo4, const float box8[8],
                  float dt, double ps, uint32_t on_mask, double* ke2_out) {
  float L[3], iL[3];
  box_f32(box8, L, iL);
  const float ps_f = (float)ps;
  long double ke2 = 0.0L;
  for (int i = 0; i < n; ++i) {
    float* r = pos4 + 4 * (size_t)i;
    float* v = vel4 + 4 * (size_t)i;
    const float* f = f4 + 4 * (size_t)i;
    float* im = im4 + 4 * (size_t)i;
    const int apply = on_mask != 0u && ((on_mask >> (type_of(r) & 31u)) & 1u);
    const float lps = apply ? ps_f : 0.f;
    const float factor = 0.5f * lps * dt;
    const float plus = 1.f / (1.f + factor);
    const float minus = 1.f - factor;
    for (int k = 0; k < 3; ++k) v[k] = plus * (minus * v[k] + v[3] * f[k] * dt);
    float img[3];
    if (rlo4 == NULL) {
      for (int k = 0; k < 3; ++k) r[k] += v[k] * dt;
      for (int k = 0; k < 3; ++k) img[k] = rintf(r[k] * iL[k]);
      for (int k = 0; k < 3; ++k) r[k] -= L[k] * img[k];
    } else {
      float* lo = rlo4 + 4 * (size_t)i;
      for (int k = 0; k < 3; ++k) {
        const double Ld = (double)L[k];
        double x = (double)r[k] + (double)lo[k] + (double)v[k] * (double)dt;
        const double q = rint(x * (1.0 / Ld));
        x -= Ld * q;
        r[k] = (float)x;
        lo[k] = (float)(x - (double)r[k]);
        img[k] = (float)q;
      }
    }
    for (int k = 0; k < 3; ++k) im[k] += img[k];
    if (apply) ke2 += (long double)((v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) / v[3]);
  }
  if (ke2_out) *ke2_out = (double)ke2;
}

/* updateNVTThermostatState (src/IntegratorNVT.cu:426-437) in FP64, then the
 * target-temperature ramp (src/IntegratorNVT.cu:91). */
void orc_thermostat_update(double* ps, double* target_T, double dT_per_step, double ke2, uint32_t dof, float dt,
                           float omega2) {
  double dev = ke2 / ((double)dof * *target_T) - 1.0;
  if (dev > 10.0) dev = 10.0;
  *ps += (double)dt * (double)omega2 * dev;
  *target_T += dT_per_step;
}

/* IntegratorNVT::GetKineticEnergy (src/IntegratorNVT.cu:316-356). */
double orc_kinetic_energy(int n, const float* pos4, const float* vel4, const float* f4, float dt, double ps,
                          uint32_t on_mask) {
  long double tot = 0.0L;
  for (int i = 0; i < n; ++i) {
    const float* r = pos4 + 4 * (size_t)i;
    const float* v = vel4 + 4 * (size_t)i;
    const float* f = f4 + 4 * (size_t)i;
    const double mass = (double)(1.f / v[3]);
    const float tsim = v[3] * dt;
    const int apply = on_mask != 0u && ((on_mask >> (type_of(r) & 31u)) & 1u);
    const float factor = apply ? (float)(0.5f * dt * (on_mask ? ps : 0.0)) : 0.f;
    const float plus = 1.f / (1.f + factor);
    const float minus = 1.f - factor;
    const float vf0 = plus * (minus * v[0] + tsim * f[0]);
    const float vf1 = plus * (minus * v[1] + tsim * f[1]);
    const float vf2 = plus * (minus * v[2] + tsim * f[2]);
    const double vb = (double)(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    const double vfs = (double)(vf0 * vf0 + vf1 * vf1 + vf2 * vf2);
    tot += (long double)(mass * ((vb + vfs) / 2.0));
  }
  return (double)(0.5L * tot);
}

/* Sum of one float4 component in extended precision (Sample::GetPotentialEnergy,
 * GetVirial, src/Sample.cu:353-391). */
double orc_sum_f4(int n, const float* a4, int comp) {
  long double s = 0.0L;
  for (int i = 0; i < n; ++i) s += (long double)a4[4 * (size_t)i + comp];
  return (double)s;
}

/* ------------------------------------------------------------------ CPU baseline MD
 * The reference's step loop (Simulation.Run, Python/rumd/Simulation.py:579-597)
 * on the CPU in FP64: Verlet list with skin rebuilt when any particle moved
 * more than skin/2 (linked cells of width >= rc_max + skin), pair forces over
 * the list, Toxvaerd Nose-Hoover leapfrog, FP64 thermostat.  pos/vel are
 * 3*n doubles, types/inv_mass per particle.  Returns the number of list
 * builds; *pe_out gets the last potential energy. */
int orc_md_run(int n, double* pos, double* vel, const int32_t* types, const double* inv_mass, const double L[3],
               int nt, const float* params, int cutoff, double skin, double dt, int nsteps, int thermostat,
               double target_T, double omega2, double* ps_io, double* pe_out) {
  pair_par* tab = make_table(nt, params);
  double rmax = 0.0;
  for (int k = 0; k < nt * nt; ++k)
    if (tab[k].rc > rmax) rmax = tab[k].rc;
  const double rl = rmax + skin;
  double* cut2 = (double*)malloc(sizeof(double) * (size_t)nt * nt);
  for (int k = 0; k < nt * nt; ++k) cut2[k] = (tab[k].rc + skin) * (tab[k].rc + skin);
  float* pos4 = (float*)malloc(sizeof(float) * 4 * (size_t)n);
  double* last = (double*)malloc(sizeof(double) * 3 * (size_t)n);
  double* frc = (double*)malloc(sizeof(double) * 3 * (size_t)n);
  double* pe = (double*)malloc(sizeof(double) * (size_t)n);
  int* cnt = (int*)malloc(sizeof(int) * (size_t)n);
  int cap = (int)(1.5 * 4.0 / 3.0 * M_PI * rl * rl * rl * n / (L[0] * L[1] * L[2])) + 64;
  int32_t* nbl = (int32_t*)malloc(sizeof(int32_t) * (size_t)n * cap);
  const int dof = 3 * n - 3;
  int builds = 0, need = 1;
  double ps = *ps_io;
  for (int step = 0; step <= nsteps; ++step) {
    if (!need) {
      const double h2 = 0.25 * skin * skin;
      for (int i = 0; i < n && !need; ++i) {
        double s = 0.0;
        for (int k = 0; k < 3; ++k) {
          double x = pos[3 * (size_t)i + k] - last[3 * (size_t)i + k];
          x -= L[k] * rint(x / L[k]);
          s += x * x;
        }
        if (s > h2) need = 1;
      }
    }
    if (need) {
      for (int i = 0; i < n; ++i) {
        for (int k = 0; k < 3; ++k) pos4[4 * (size_t)i + k] = (float)pos[3 * (size_t)i + k];
        memcpy(pos4 + 4 * (size_t)i + 3, &types[i], sizeof(int32_t));
        memcpy(last + 3 * (size_t)i, pos + 3 * (size_t)i, 3 * sizeof(double));
      }
      cell_list cl;
      if (build_cells(&cl, n, pos4, L, rl) != 0) {
        free(tab); free(cut2); free(pos4); free(last); free(frc); free(pe); free(cnt); free(nbl);
        return -1;
      }
      int overflow = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(| : overflow)
      for (int i = 0; i < n; ++i) {
        const double* ri = pos + 3 * (size_t)i;
        const double x[3] = {ri[0], ri[1], ri[2]};
        const int c = cell_index_f64(x, L, cl.nc);
        const int cx = c % cl.nc[0], cy = (c / cl.nc[0]) % cl.nc[1], cz = c / (cl.nc[0] * cl.nc[1]);
        int m = 0;
        for (int dz = -1; dz <= 1; ++dz)
          for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
              const int ox = (cx + dx + cl.nc[0]) % cl.nc[0];
              const int oy = (cy + dy + cl.nc[1]) % cl.nc[1];
              const int oz = (cz + dz + cl.nc[2]) % cl.nc[2];
              for (int j = cl.head[ox + cl.nc[0] * (oy + cl.nc[1] * oz)]; j >= 0; j = cl.next[j]) {
                if (j == i) continue;
                double s = 0.0;
                for (int k = 0; k < 3; ++k) {
                  double d = ri[k] - pos[3 * (size_t)j + k];
                  d -= L[k] * rint(d / L[k]);
                  s += d * d;
                }
                if (s < cut2[types[i] * nt + types[j]]) {
                  if (m < cap) nbl[(size_t)i * cap + m] = j;
                  ++m;
                }
              }
            }
        if (m > cap) overflow = 1;
        cnt[i] = m < cap ? m : cap;
      }
      free_cells(&cl);
      if (overflow) {
        free(tab); free(cut2); free(pos4); free(last); free(frc); free(pe); free(cnt); free(nbl);
        return -2;
      }
      ++builds;
      need = 0;
    }
    /* forces */
#pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < n; ++i) {
      double fi[3] = {0, 0, 0}, u = 0, w = 0;
      const double* ri = pos + 3 * (size_t)i;
      for (int m = 0; m < cnt[i]; ++m) {
        const int j = nbl[(size_t)i * cap + m];
        double d[3], s2 = 0.0;
        for (int k = 0; k < 3; ++k) {
          double x = ri[k] - pos[3 * (size_t)j + k];
          x -= L[k] * rint(x / L[k]);
          d[k] = x;
          s2 += x * x;
        }
        pair_add(&tab[types[i] * nt + types[j]], cutoff, d, s2, fi, &u, &w);
      }
      frc[3 * (size_t)i] = fi[0];
      frc[3 * (size_t)i + 1] = fi[1];
      frc[3 * (size_t)i + 2] = fi[2];
      pe[i] = 0.5 * u;
    }
    if (step == nsteps) break;
    /* leapfrog + Nose-Hoover */
    const double factor = thermostat ? 0.5 * ps * dt : 0.0;
    const double plus = 1.0 / (1.0 + factor), minus = 1.0 - factor;
    double ke2 = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : ke2)
    for (int i = 0; i < n; ++i) {
      double vv = 0.0;
      for (int k = 0; k < 3; ++k) {
        double* v = vel + 3 * (size_t)i + k;
        *v = plus * (minus * *v + inv_mass[i] * frc[3 * (size_t)i + k] * dt);
        double* x = pos + 3 * (size_t)i + k;
        *x += *v * dt;
        *x -= L[k] * rint(*x / L[k]);
        vv += *v * *v;
      }
      ke2 += vv / inv_mass[i];
    }
    if (thermostat) {
      double dev = ke2 / ((double)dof * target_T) - 1.0;
      if (dev > 10.0) dev = 10.0;
      ps += dt * omega2 * dev;
    }
  }
  double pe_tot = 0.0;
  for (int i = 0; i < n; ++i) pe_tot += pe[i];
  *pe_out = pe_tot;
  *ps_io = ps;
  free(tab); free(cut2); free(pos4); free(last); free(frc); free(pe); free(cnt); free(nbl);
  return builds;
}

int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void orc_set_num_threads(int k) {
#ifdef _OPENMP
  if (k > 0) omp_set_num_threads(k);
#else
  (void)k;
#endif
}
