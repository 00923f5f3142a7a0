// Test infrastructure only: C entry point over the reference's face-face cubature of the N operator
// (src/src_vie/src_operators/mexcubatures/cubature_nop_const.cpp), built by oracle/build_ref.sh.
#include <complex>

void cubature_nop(const int Np, const double w[], const double u[], const double v[], const double r_m[3],
                  const double r_n[3], const double R_faces[3][4][6], const double ko, const double dx,
                  std::complex<double> I_mn[6][6]);

// R_faces: 3 x 4 x 6 doubles in C order [coord][vertex][face]; out: 6 x 6 complex, row-major [obs][src]
extern "C" void vsref_cubature_nop(int np, const double* w, const double* u, const double* v, const double* r_m,
                                   const double* r_n, const double* R_faces, double ko, double dx, double* out) {
  std::complex<double> I[6][6];
  cubature_nop(np, w, u, v, r_m, r_n, reinterpret_cast<const double(*)[4][6]>(R_faces), ko, dx, I);
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) {
      out[2 * (6 * i + j)] = I[i][j].real();
      out[2 * (6 * i + j) + 1] = I[i][j].imag();
    }
}
