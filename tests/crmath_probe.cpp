// Host build of crmath.h for tests/test_crmath.py: evaluates one function over arrays.
#define _GNU_SOURCE 1
#define CRM_FN inline
#include "../sudodem_amd/csrc/crmath.h"
#include <stdint.h>
extern "C" void crm_eval(int fn, const double* a, const double* b, double* out, int64_t n) {
  for (int64_t i = 0; i < n; i++) {
    switch (fn) {
      case 0: out[i] = crm::cr_pow(a[i], b[i]); break;
      case 1: out[i] = crm::cr_atan2(a[i], b[i]); break;
      case 2: out[i] = crm::cr_sin(a[i]); break;
      case 3: out[i] = crm::cr_cos(a[i]); break;
      case 4: out[i] = crm::cr_atan(a[i]); break;
      case 5: { double sn, cs; crm::cr_sincos(a[i], &sn, &cs); out[i] = sn; break; }
      case 6: { double sn, cs; crm::cr_sincos(a[i], &sn, &cs); out[i] = cs; break; }
      case 7: out[i] = crm::cr_log(a[i]); break;
      case 8: out[i] = crm::cr_log10(a[i]); break;
      case 9: out[i] = crm::cr_acos(a[i]); break;
      case 10: {  // the Hertz kn chain (4/3 M_PIl a b / c) emulated, c = 0.7 b + 1
        const double c = 0.7 * b[i] + 1.0;
        crm::DD t = crm::ext_mul_d(crm::ext_pi(), 4. / 3.);
        t = crm::ext_mul_d(crm::ext_mul_d(t, a[i]), b[i]);
        out[i] = crm::ext_to_double(crm::ext_div_d(t, c));
        break;
      }
      case 11: {  // the same in x87 long double
        const double c = 0.7 * b[i] + 1.0;
        out[i] = (double)(4. / 3. * M_PIl * a[i] * b[i] / c);
        break;
      }
      case 12: {  // the ks chain M_PIl a b s emulated, s = 1/a + 1/b
        const double sv = 1. / a[i] + 1. / b[i];
        out[i] = crm::ext_to_double(crm::ext_mul_d(crm::ext_mul_d(crm::ext_mul_d(crm::ext_pi(), a[i]), b[i]), sv));
        break;
      }
      case 14: { double p1, p2; crm::cr_pow_pair(a[i], b[i], 0.5 * b[i] + 0.25, &p1, &p2); out[i] = p1; break; }
      case 15: { double p1, p2; crm::cr_pow_pair(a[i], b[i], 0.5 * b[i] + 0.25, &p1, &p2); out[i] = p2; break; }
      case 13: {
        const double sv = 1. / a[i] + 1. / b[i];
        out[i] = (double)(M_PIl * a[i] * b[i] * sv);
        break;
      }
      default: out[i] = 0.0;
    }
  }
}
