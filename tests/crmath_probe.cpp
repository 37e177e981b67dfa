// Host build of crmath.h for tests/test_crmath.py: evaluates one function over arrays.
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
      default: out[i] = 0.0;
    }
  }
}
