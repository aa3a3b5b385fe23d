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

// True for the forms this header defines.
RUMDX_HD bool pot_form_known(int form) { return form >= RUMDX_POT_LJ_SMOOTH_2_4 && form <= RUMDX_POT_EXP; }

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
    default:
      u = T(0);
      return T(0);
  }
}

}  // namespace rumdx
