// Test infrastructure only: C entry point over the reference's DIRECTFN self-term integral
// (src/src_vie/src_operators/mexcubatures/directfn_ws_st_const.cpp, compiled where it lies by
// oracle/build_ref.sh into oracle/_ref/).  Used to pin the singular face-face integrals.
#include "directfn_ws_st_const.h"

extern "C" void vsref_directfn_st(const double* r1, const double* r2, const double* r3, double ko, int np,
                                  double* out) {
  const std::complex<double> v = directfn_ws_st_const(r1, r2, r3, ko, np);
  out[0] = v.real();
  out[1] = v.imag();
}
