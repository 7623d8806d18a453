// ORACLE -- test infrastructure only (see oracle_internal.h).
// Discrete Fourier transforms with the FFTW conventions the reference uses
// (src/grid/fft.cpp:48-58, fft.h:38-60): forward e^{-2 pi i k x / n},
// unnormalised; backward e^{+2 pi i k x / n}, unnormalised (the caller
// scales).  Textbook recursive mixed-radix Cooley-Tukey for the factors
// 2, 3, 5 and a direct O(n^2) DFT for any other prime factor.
#include <cstring>

#include "oracle_internal.h"

namespace orc {

namespace {

void dft_rec(int n, const cplx* in, int istride, cplx* out, int sign) {
  if (n == 1) {
    out[0] = in[0];
    return;
  }
  int r = 0;
  for (int f : {2, 3, 5})
    if (n % f == 0) {
      r = f;
      break;
    }
  if (r == 0) {  // direct DFT
    for (int k = 0; k < n; ++k) {
      cplx s(0.0, 0.0);
      for (int x = 0; x < n; ++x) {
        double ang = sign * 2.0 * M_PI * (double)((int64_t)k * x % n) / n;
        s += in[x * istride] * cplx(std::cos(ang), std::sin(ang));
      }
      out[k] = s;
    }
    return;
  }
  int m = n / r;
  // decimation in time: sub-sequences x[r*t + q], q = 0..r-1
  std::vector<cplx> sub(n);
  for (int q = 0; q < r; ++q) dft_rec(m, in + q * istride, istride * r, sub.data() + q * m, sign);
  for (int k = 0; k < m; ++k)
    for (int p = 0; p < r; ++p) {
      int kk = k + p * m;
      cplx s(0.0, 0.0);
      for (int q = 0; q < r; ++q) {
        double ang = sign * 2.0 * M_PI * (double)((int64_t)q * kk % n) / n;
        s += sub[q * m + k] * cplx(std::cos(ang), std::sin(ang));
      }
      out[kk] = s;
    }
}

}  // namespace

void fft1d(int n, cplx* data, int stride, int sign) {
  std::vector<cplx> in(n), out(n);
  for (int i = 0; i < n; ++i) in[i] = data[(int64_t)i * stride];
  dft_rec(n, in.data(), 1, out.data(), sign);
  for (int i = 0; i < n; ++i) data[(int64_t)i * stride] = out[i];
}

// real (Nx,Ny,Nz) x fastest -> complex (nr, Nz, Ny) i fastest
void fft3d_r2c(const int cells[3], const double* in, cplx* out) {
  const int Nx = cells[0], Ny = cells[1], Nz = cells[2], nr = Nx / 2 + 1;
  std::vector<cplx> row(Nx);
  // x transforms into a (nr, Nz, Ny) buffer
  for (int z = 0; z < Nz; ++z)
    for (int y = 0; y < Ny; ++y) {
      for (int x = 0; x < Nx; ++x) row[x] = cplx(in[x + (int64_t)Nx * (y + (int64_t)Ny * z)], 0.0);
      fft1d(Nx, row.data(), 1, -1);
      for (int i = 0; i < nr; ++i) out[i + (int64_t)nr * (z + (int64_t)Nz * y)] = row[i];
    }
  // z transforms (stride nr)
  for (int y = 0; y < Ny; ++y)
    for (int i = 0; i < nr; ++i) fft1d(Nz, out + i + (int64_t)nr * Nz * y, nr, -1);
  // y transforms (stride nr*Nz)
  for (int z = 0; z < Nz; ++z)
    for (int i = 0; i < nr; ++i) fft1d(Ny, out + i + (int64_t)nr * z, nr * Nz, -1);
}

void fft3d_c2r(const int cells[3], const cplx* in, double* out) {
  const int Nx = cells[0], Ny = cells[1], Nz = cells[2], nr = Nx / 2 + 1;
  std::vector<cplx> buf(in, in + (int64_t)nr * Ny * Nz);
  for (int z = 0; z < Nz; ++z)
    for (int i = 0; i < nr; ++i) fft1d(Ny, buf.data() + i + (int64_t)nr * z, nr * Nz, +1);
  for (int y = 0; y < Ny; ++y)
    for (int i = 0; i < nr; ++i) fft1d(Nz, buf.data() + i + (int64_t)nr * Nz * y, nr, +1);
  std::vector<cplx> row(Nx);
  for (int z = 0; z < Nz; ++z)
    for (int y = 0; y < Ny; ++y) {
      const cplx* h = buf.data() + (int64_t)nr * (z + (int64_t)Nz * y);
      for (int i = 0; i < nr; ++i) row[i] = h[i];
      for (int i = nr; i < Nx; ++i) row[i] = std::conj(h[Nx - i]);
      // c2r: imaginary parts of the self-conjugate bins do not contribute
      row[0] = cplx(row[0].real(), 0.0);
      if (Nx % 2 == 0) row[Nx / 2] = cplx(row[Nx / 2].real(), 0.0);
      fft1d(Nx, row.data(), 1, +1);
      for (int x = 0; x < Nx; ++x) out[x + (int64_t)Nx * (y + (int64_t)Ny * z)] = row[x].real();
    }
}

}  // namespace orc
