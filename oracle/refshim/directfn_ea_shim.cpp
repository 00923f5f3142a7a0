// Test infrastructure only: C entry point over the reference's DIRECTFN edge-adjacent integral
// (src/src_vie/src_operators/mexcubatures/directfn_ws_ea_const.cpp), built by oracle/build_ref.sh.
#include "directfn_ws_ea_const.h"

extern "C" void vsref_directfn_ea(const double* p1, const double* p2, const double* p3, const double* q1,
                                  const double* q2, const double* q3, double ko, int np, double* out) {
  const std::complex<double> v = directfn_ws_ea_const(p1, p2, p3, q1, q2, q3, ko, np);
  out[0] = v.real();
  out[1] = v.imag();
}
