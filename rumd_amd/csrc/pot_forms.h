// Pair potential forms of RUMD's Pot_* classes outside the 12-6 family: the
// force factor s = -u'(r)/r (F_ij = s r_ij) and the energy u of one pair from
// r^2 and the parameter row, before the cut-off shift.  One definition serves
// the device kernels (T = float) and the host's shift constants and checks
// (T = double).
//
// p[k] is parameter k of RUMD's 12-float row (p[0] = Rcut is not read here);
// the parameters of each form are those its SetParams stores:
//   LJ_12_6_smooth_2_4 (include/rumd/PairPotential.h:218-256): 1/sigma^2, 4 eps, c2, c4
//   gLJ_m_n            (:264-304): 1/sigma^2, m n eps/(m-n), m/2, n/2, eps/(m-n)
//   Gauss              (:383-414): 1/sigma^2, eps
//   Buckingham         (:422-461): 1/rm^2, alpha, 6 eps/(alpha-6), alpha eps/(alpha-6),
//                                  6 alpha eps/(alpha-6)
//   IPL_n              (:600-636): 1/sigma^2, eps, n
//   Yukawa             (:795-840): sigma, 1/sigma, eps
//   Exp                (:1019-1052): 1/sigma, eps
//   LJ_12_6_Gauss      (:336-378): 1/sigma^2, 12 eps, 2 sigma0^2, r0, eps0, eps
//   IPL_12, IPL_9, IPL_18 (:469-593): 1/sigma^2, n eps, eps
//   IPL6_LN            (:644-680): eps, 1/sigma^2, eps/sigma^2
//   LJ_LN              (:689-733): 1/Rm^2, 1/R0^2, eps a, eps a/Rm^2, b/a
//   Girifalco          (:743-786): 1/R0^2, alpha, beta
//   Dzugutov           (:846-900): a, b, c, d, A, B (Rcut = b - 1e-6)
//   Fermi_Jagla        (:910-961): 1/a^2, A0, A1/A0, A2, B0, B1/B0, B2, eps0, n
//   Repulsive_Shoulder (:969-1012): 1/sigma^2, eps/2, k0, sigma1
//   SAAP               (:1060-1135): a0..a6, 1/sigma^2, eps
//   ShiftedForceCoulomb (:1143-1186): eps, eps/Rcut, eps/Rcut^2
#pragma once

#include <math.h>

#include "rumdx.h"

#ifdef __HIPCC__
#define RUMDX_HD __host__ __device__ inline
#else
#define RUMDX_HD inline
#endif

namespace rumdx {

RUMDX_HD float pf_exp(float x) { return expf(x); }
RUMDX_HD double pf_exp(double x) { return exp(x); }
RUMDX_HD float pf_log(float x) { return logf(x); }
RUMDX_HD double pf_log(double x) { return log(x); }
RUMDX_HD float pf_sqrt(float x) { return sqrtf(x); }
RUMDX_HD double pf_sqrt(double x) { return sqrt(x); }
RUMDX_HD float pf_tanh(float x) { return tanhf(x); }
RUMDX_HD double pf_tanh(double x) { return tanh(x); }

// True for the forms this header defines.
RUMDX_HD bool pot_form_known(int form) { return form >= RUMDX_POT_LJ_SMOOTH_2_4 && form <= RUMDX_POT_SF_COULOMB; }

// s and u of one pair at squared distance d2 (d2 > 0).
template <class T>
RUMDX_HD T pot_form(int form, T d2, const T* p, T& u) {
  const T one = T(1);
  switch (form) {
    case RUMDX_POT_LJ_SMOOTH_2_4: {
      // v = 4 eps ((sigma/r)^12 - (sigma/r)^6 + c2 (r/sigma)^2 + c4 (r/sigma)^4)
      const T x2 = d2 * p[1];
      const T i2 = one / x2, i6 = i2 * i2 * i2;
      u = p[2] * (i6 * (i6 - one) + p[3] * x2 + p[4] * x2 * x2);
      return p[2] * p[1] * (i2 * i6 * (T(12) * i6 - T(6)) - T(2) * p[3] - T(4) * p[4] * x2);
    }
    case RUMDX_POT_GLJ_M_N: {
      // v = eps / (m - n) (n (sigma/r)^m - m (sigma/r)^n)
      const T i2 = one / (d2 * p[1]);
      const T l = pf_log(i2);
      const T im = pf_exp(l * p[3]), in = pf_exp(l * p[4]);
      u = p[5] * (T(2) * p[4] * im - T(2) * p[3] * in);
      return p[1] * p[2] * (im - in) * i2;
    }
    case RUMDX_POT_GAUSS: {
      // v = eps exp(-(r/sigma)^2)
      const T e = pf_exp(-(d2 * p[1]));
      u = p[2] * e;
      return T(2) * p[1] * p[2] * e;
    }
    case RUMDX_POT_BUCKINGHAM: {
      // v = eps (6/(alpha-6) exp(alpha (1 - r/rm)) - alpha/(alpha-6) (rm/r)^6)
      const T i2 = one / (d2 * p[1]), i6 = i2 * i2 * i2;
      const T i1 = pf_sqrt(i2);
      const T e = pf_exp(p[2] * (one - one / i1));
      u = p[3] * e - p[4] * i6;
      return p[1] * p[5] * (i1 * e - i6 * i2);
    }
    case RUMDX_POT_IPL_N: {
      // v = eps (sigma/r)^n
      const T i2 = one / (d2 * p[1]);
      const T in = pf_exp(pf_log(i2) * p[3] * T(0.5));
      u = p[2] * in;
      return p[1] * p[2] * p[3] * in * i2;
    }
    case RUMDX_POT_YUKAWA: {
      // v = eps sigma exp(-r/sigma) / r
      const T r = pf_sqrt(d2), ir = one / r;
      const T v = p[3] * p[1] * ir * pf_exp(-p[2] * r);
      u = v;
      return ir * (ir + p[2]) * v;
    }
    case RUMDX_POT_EXP: {
      // v = eps exp(-r/sigma)
      const T x = pf_sqrt(d2) * p[1];
      const T e = pf_exp(-x);
      u = p[2] * e;
      return p[1] * p[2] * (p[1] / x) * e;
    }
    case RUMDX_POT_LJ_GAUSS: {
      // v = eps ((sigma/r)^12 - 2 (sigma/r)^6 - eps0 exp(-(r - r0)^2 / (2 sigma0^2)))
      const T i2 = one / (d2 * p[1]), i6 = i2 * i2 * i2;
      const T r = pf_sqrt(d2), x = r - p[4];
      const T e = pf_exp(-(x * x) / p[3]);
      u = p[6] * (i6 * (i6 - T(2)) - p[5] * e);
      return p[1] * p[2] * i6 * i2 * (i6 - one) - p[5] * p[2] * x * e / (T(6) * r * p[3]);
    }
    case RUMDX_POT_IPL_12:
    case RUMDX_POT_IPL_9:
    case RUMDX_POT_IPL_18: {
      // v = eps (sigma/r)^n, n = 12, 9 or 18; p[2] = n eps
      const T i2 = one / (d2 * p[1]);
      T in;
      if (form == RUMDX_POT_IPL_12) {
        const T i6 = i2 * i2 * i2;
        in = i6 * i6;
      } else if (form == RUMDX_POT_IPL_9) {
        const T i4 = i2 * i2;
        in = i4 * i4 * pf_sqrt(i2);
      } else {
        const T i6 = i2 * i2 * i2;
        in = i6 * i6 * i6;
      }
      u = p[3] * in;
      return p[1] * p[2] * in * i2;
    }
    case RUMDX_POT_IPL6_LN: {
      // v = eps (sigma/r)^6 / ln(r/sigma)
      const T x2 = d2 * p[2];
      const T l = T(0.5) * pf_log(x2), i2 = one / x2, i6 = i2 * i2 * i2;
      u = p[1] * i6 / l;
      return p[3] * i6 * i2 * (T(6) * l + one) / (l * l);
    }
    case RUMDX_POT_LJ_LN: {
      // v = eps (a (Rm/r)^12 - b (Rm/r)^6) / ln(r/R0)
      const T l = T(0.5) * pf_log(d2 * p[2]), il = one / l;
      const T i2 = one / (d2 * p[1]), i6 = i2 * i2 * i2;
      u = p[3] * i6 * (i6 - p[5]) * il;
      return p[4] * i6 * i2 * il * il * (i6 * (T(12) * l + one) - p[5] * (T(6) * l + one));
    }
    case RUMDX_POT_GIRIFALCO: {
      // x = r/R0: v = -alpha (1/(x (x-1)^3) + 1/(x (x+1)^3) - 2/x^4)
      //              + beta (1/(x (x-1)^9) + 1/(x (x+1)^9) - 2/x^10)
      const T x = pf_sqrt(d2 * p[1]);
      const T xm = x - one, xp = x + one;
      const T m2 = one / (xm * xm), q2 = one / (xp * xp);
      const T m4 = m2 * m2, q4 = q2 * q2, m3 = m4 * xm, q3 = q4 * xp;
      const T ix = one / x, i3 = ix * ix * ix;
      u = ix * (-p[2] * (m3 + q3 - T(2) * i3) + p[3] * (m3 * m3 * m3 + q3 * q3 * q3 - T(2) * i3 * i3 * i3));
      return i3 * p[1] *
             (-p[2] * ((T(4) * x - one) * m4 + (T(4) * x + one) * q4 - T(8) * i3) +
              p[3] * ((T(10) * x - one) * m4 * m3 * m3 + (T(10) * x + one) * q4 * q3 * q3 - T(20) * i3 * i3 * i3));
    }
    case RUMDX_POT_DZUGUTOV: {
      // v = A (r^-16 - B) exp(c / (r - a)) [r < a] + B exp(d / (r - b)) [r < b]
      const T r = pf_sqrt(d2), ir = one / r;
      const T ib = one / (r - p[2]);
      const T eb = pf_exp(p[4] * ib);
      T s = ir * p[4] * p[6] * ib * ib * eb;
      u = p[6] * eb;
      if (r < p[1]) {
        const T i2 = one / d2, i4 = i2 * i2, i8 = i4 * i4, i16 = i8 * i8;
        const T ia = one / (r - p[1]);
        const T ea = pf_exp(p[3] * ia);
        const T v1 = p[5] * ea * (i16 - p[6]);
        s += T(16) * p[5] * ea * i16 * i2 + ir * p[3] * ia * ia * v1;
        u += v1;
      }
      return s;
    }
    case RUMDX_POT_FERMI_JAGLA: {
      // x = r/a: v = eps0 (x^-n + A0 / (1 + exp(A1/A0 (x - A2))) - B0 / (1 + exp(B1/B0 (x - B2))))
      const T x2 = d2 * p[1], x = pf_sqrt(x2), i2 = one / x2;
      const T in = pf_exp(pf_log(i2) * p[9] * T(0.5));
      const T fa = one + pf_exp(p[3] * (x - p[4])), fb = one + pf_exp(p[6] * (x - p[7]));
      const T ia = one / fa, ib = one / fb;
      u = p[8] * (in + p[2] * ia - p[5] * ib);
      return p[8] * (p[9] * p[1] * i2 * in + p[1] * (p[2] * p[3] * ia * ia * (fa - one) - p[5] * p[6] * ib * ib * (fb - one)) / x);
    }
    case RUMDX_POT_REP_SHOULDER: {
      // v = (sigma/r)^14 + eps/2 (1 - tanh(k0 (r - sigma1)))
      const T r = pf_sqrt(d2), i2 = one / (d2 * p[1]);
      const T i4 = i2 * i2, i8 = i4 * i4;
      const T th = pf_tanh(p[3] * (r - p[4]));
      u = i8 * i4 * i2 + p[2] * (one - th);
      return T(14) * p[1] * i8 * i8 + p[2] * p[3] * (one - th * th) / r;
    }
    case RUMDX_POT_SAAP: {
      // x = r/sigma: v = eps (a0/x exp(a1 x + a6 x^2) + a2 exp(a3 x) + a4) / (1 + a5 x^6).
      // The force factor carries the reference's 1/(sigma^2 r) scale, which is
      // -v'(r)/r for sigma = 1 (include/rumd/PairPotential.h:1115).
      const T r = pf_sqrt(d2), x2 = d2 * p[8], x = pf_sqrt(x2);
      const T e1 = p[1] * pf_exp(p[2] * x + p[7] * x2);
      const T e3 = pf_exp(p[4] * x);
      const T q = one / (one + p[6] * x2 * x2 * x2);
      const T nn = e1 / x + p[3] * e3 + p[5];
      u = p[9] * nn * q;
      const T dn = p[3] * p[4] * e3 - e1 / x2 + e1 * (p[2] + T(2) * p[7] * x) / x;
      const T d6 = T(6) * p[6] * x2 * x2 * x * nn;
      return p[9] * p[8] / r * (d6 * q * q - dn * q);
    }
    case RUMDX_POT_SF_COULOMB: {
      // v = eps/r + eps/Rc^2 (r - Rc) - eps/Rc (shifted force built in); with
      // eps/Rc^2 Rc = eps/Rc the constant is -2 eps/Rc
      const T r = pf_sqrt(d2), ir = one / r;
      u = p[1] * ir + p[3] * r - T(2) * p[2];
      return (p[1] * ir * ir - p[3]) * ir;
    }
    default:
      u = T(0);
      return T(0);
  }
}

}  // namespace rumdx
