// Spectral stages on CDNA4: 3D real<->complex FFTs of the 9-component
// tensor fields (LDS-staged mixed-radix Stockham passes along x, y and z),
// the fused Gamma-operator convolution + divergence residual per wavevector,
// and the streaming field operations of the solver (forward/rate fields,
// deterministic reductions, N-GMRES vector updates).  All HBM-bound.
//
// Layouts.  A real tensor field is 9 planes of N = Nx*Ny*Nz values (x
// fastest); its half spectrum is 9 planes of K = (Nx/2+1)*Ny*Nz complex
// values indexed i + nr*(y + Ny*z) (kx fastest).  Forward transforms are
// unnormalised with e^{-2 pi i k.x/N} (FFTW r2c, src/grid/fft.cpp:48-52),
// inverse transforms e^{+2 pi i ...} scaled by a caller-given factor
// (src/grid/fft.h:50-60).
#include <array>
#include <map>
#include <mutex>
#include <tuple>

#include "spectral.h"

namespace sca {

template <typename T>
struct cpx {
  T x, y;
};
template <typename T>
__device__ __forceinline__ cpx<T> cmul(cpx<T> a, cpx<T> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename T>
__device__ __forceinline__ cpx<T> cadd(cpx<T> a, cpx<T> b) {
  return {a.x + b.x, a.y + b.y};
}
template <typename T>
__device__ __forceinline__ cpx<T> csub(cpx<T> a, cpx<T> b) {
  return {a.x - b.x, a.y - b.y};
}

}  // namespace sca
#include "fft_pencil.h"
namespace sca {

// ---------------------------------------------------------------------------
// FFT passes
// ---------------------------------------------------------------------------
struct FftPassDesc {
  int n;           // transform length
  int nrad;        // number of radix passes
  int rad[16];     // radices 4, 2, 3, 5, 7, 11, 13
  int sign;        // -1 forward, +1 inverse
};

// R-point DFT in registers for R = 2, 3, 4; generic R by direct summation
// with the length-n table (exp(sign 2 pi i q k / R) = tw[q k n / R mod n]).
template <typename T, int R>
__device__ __forceinline__ void dft_r(cpx<T>* v, int sign, const cpx<T>* tw, int n) {
  if constexpr (R == 2) {
    const cpx<T> a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  } else if constexpr (R == 4) {
    const cpx<T> a = v[0], b = v[1], c = v[2], d = v[3];
    const cpx<T> s0 = cadd(a, c), s1 = csub(a, c), s2 = cadd(b, d), s3 = csub(b, d);
    // W = sign*i ; W*z = sign*(-z.y, z.x)
    const cpx<T> ws3 = {(T)sign * -s3.y, (T)sign * s3.x};
    v[0] = cadd(s0, s2);
    v[2] = csub(s0, s2);
    v[1] = cadd(s1, ws3);
    v[3] = csub(s1, ws3);
  } else if constexpr (R == 3) {
    const T h = (T)0.86602540378443864676;  // sqrt(3)/2
    const cpx<T> a = v[0], b = v[1], c = v[2];
    const cpx<T> t1 = cadd(b, c), t2 = csub(b, c);
    const cpx<T> m = {a.x - (T)0.5 * t1.x, a.y - (T)0.5 * t1.y};
    const cpx<T> it2 = {(T)sign * -h * t2.y, (T)sign * h * t2.x};
    v[0] = cadd(a, t1);
    v[1] = cadd(m, it2);
    v[2] = csub(m, it2);
  } else {
    cpx<T> o[R];
#pragma unroll
    for (int k = 0; k < R; ++k) {
      cpx<T> s = {(T)0, (T)0};
#pragma unroll
      for (int q = 0; q < R; ++q) {
        cpx<T> w = tw[((q * k) % R) * (n / R)];
        if (sign > 0) w.y = -w.y;
        s = cadd(s, cmul(v[q], w));
      }
      o[k] = s;
    }
#pragma unroll
    for (int k = 0; k < R; ++k) v[k] = o[k];
  }
}

// one Stockham radix-R pass over PB pencils (a -> b)
template <typename T, int R>
__device__ __forceinline__ void stockham_pass(const cpx<T>* a, cpx<T>* b, int PB, int ps, int n, int Ns, int sign,
                                              const cpx<T>* tw) {
  const int nb = n / R;
  const int step = n / (Ns * R);
  for (int idx = threadIdx.x; idx < PB * nb; idx += blockDim.x) {
    const int j = idx / nb, jj = idx % nb;
    const int k = jj % Ns;
    cpx<T> v[R];
    const cpx<T>* src = a + j * ps;
#pragma unroll
    for (int q = 0; q < R; ++q) {
      cpx<T> x = src[jj + q * nb];
      if (q && k) {
        cpx<T> w = tw[q * k * step];
        if (sign > 0) w.y = -w.y;
        x = cmul(x, w);
      }
      v[q] = x;
    }
    dft_r<T, R>(v, sign, tw, n);
    const int idxD = (jj / Ns) * Ns * R + k;
    cpx<T>* dst = b + j * ps;
#pragma unroll
    for (int q = 0; q < R; ++q) dst[idxD + q * Ns] = v[q];
  }
}

// Stockham autosort passes on PB pencils of length n held in LDS with
// pencil stride ps (a -> b ping-pong); returns the buffer holding the
// result.  tw: forward table tw[t] = (cos 2 pi t/n, -sin 2 pi t/n).
template <typename T>
__device__ cpx<T>* stockham(cpx<T>* a, cpx<T>* b, int PB, int ps, const FftPassDesc& d, const cpx<T>* tw) {
  const int n = d.n;
  int Ns = 1;
  for (int p = 0; p < d.nrad; ++p) {
    const int R = d.rad[p];
    switch (R) {
      case 4: stockham_pass<T, 4>(a, b, PB, ps, n, Ns, d.sign, tw); break;
      case 2: stockham_pass<T, 2>(a, b, PB, ps, n, Ns, d.sign, tw); break;
      case 3: stockham_pass<T, 3>(a, b, PB, ps, n, Ns, d.sign, tw); break;
      case 5: stockham_pass<T, 5>(a, b, PB, ps, n, Ns, d.sign, tw); break;
      case 7: stockham_pass<T, 7>(a, b, PB, ps, n, Ns, d.sign, tw); break;
      case 11: stockham_pass<T, 11>(a, b, PB, ps, n, Ns, d.sign, tw); break;
      default: stockham_pass<T, 13>(a, b, PB, ps, n, Ns, d.sign, tw); break;
    }
    __syncthreads();
    cpx<T>* t = a;
    a = b;
    b = t;
    Ns *= R;
  }
  return a;
}

// x pass, real -> half complex: rows R of Nx reals at in + R*Nx (one plane
// of the field), output rows of nr complex at out + R*nr.
template <typename T, typename TI>
__global__ void __launch_bounds__(256) fft_x_r2c(const TI* in, cpx<T>* out, int64_t nrows, FftPassDesc d,
                                                 const cpx<T>* tw, int PB) {
  extern __shared__ unsigned char smem_raw[];
  const int n = d.n, nr = n / 2 + 1;
  cpx<T>* a = reinterpret_cast<cpx<T>*>(smem_raw);
  cpx<T>* b = a + PB * n;
  const int64_t r0 = (int64_t)blockIdx.x * PB;
  for (int f = threadIdx.x; f < PB * n; f += blockDim.x) {
    const int j = f / n, e = f % n;
    const int64_t r = r0 + j;
    a[f] = {(r < nrows) ? (T)in[r * n + e] : (T)0, (T)0};
  }
  __syncthreads();
  cpx<T>* res = stockham(a, b, PB, n, d, tw);
  for (int f = threadIdx.x; f < PB * nr; f += blockDim.x) {
    const int j = f / nr, e = f % nr;
    const int64_t r = r0 + j;
    if (r < nrows) out[r * nr + e] = res[j * n + e];
  }
}

// x pass, half complex -> real (Hermitian completion), scaled
template <typename T, typename TO>
__global__ void __launch_bounds__(256) fft_x_c2r(const cpx<T>* in, TO* out, int64_t nrows, FftPassDesc d,
                                                 const cpx<T>* tw, T scale, int PB) {
  extern __shared__ unsigned char smem_raw[];
  const int n = d.n, nr = n / 2 + 1;
  cpx<T>* a = reinterpret_cast<cpx<T>*>(smem_raw);
  cpx<T>* b = a + PB * n;
  const int64_t r0 = (int64_t)blockIdx.x * PB;
  for (int f = threadIdx.x; f < PB * n; f += blockDim.x) {
    const int j = f / n, e = f % n;
    const int64_t r = r0 + j;
    cpx<T> v = {(T)0, (T)0};
    if (r < nrows) {
      if (e < nr) {
        v = in[r * nr + e];
      } else {
        v = in[r * nr + (n - e)];
        v.y = -v.y;
      }
      if (e == 0 || (2 * e == n)) v.y = (T)0;
    }
    a[f] = v;
  }
  __syncthreads();
  cpx<T>* res = stockham(a, b, PB, n, d, tw);
  for (int f = threadIdx.x; f < PB * n; f += blockDim.x) {
    const int j = f / n, e = f % n;
    const int64_t r = r0 + j;
    if (r < nrows) out[r * n + e] = (TO)(res[j * n + e].x * scale);
  }
}

// strided complex pass (y or z): pencil (o, i) starts at
// (o / d1)*s1 + (o % d1)*s2 + i and has element stride es; a block takes PB
// consecutive i of one o, so that global accesses are PB-wide contiguous.
template <typename T>
__global__ void __launch_bounds__(256) fft_strided(cpx<T>* data, int ni, int d1, int64_t s1, int64_t s2, int64_t es,
                                                   FftPassDesc d, const cpx<T>* tw, int PB) {
  extern __shared__ unsigned char smem_raw[];
  const int n = d.n;
  const int ps = n + 1;  // padded pencil stride: conflict-free LDS transposes
  cpx<T>* a = reinterpret_cast<cpx<T>*>(smem_raw);
  cpx<T>* b = a + PB * ps;
  const int i0 = blockIdx.x * PB;
  const int64_t o = blockIdx.y;
  const int64_t base = (o / d1) * s1 + (o % d1) * s2;
  for (int f = threadIdx.x; f < PB * n; f += blockDim.x) {
    const int j = f % PB, e = f / PB;
    const int i = i0 + j;
    a[j * ps + e] = (i < ni) ? data[base + i + e * es] : cpx<T>{(T)0, (T)0};
  }
  __syncthreads();
  cpx<T>* res = stockham(a, b, PB, ps, d, tw);
  for (int f = threadIdx.x; f < PB * n; f += blockDim.x) {
    const int j = f % PB, e = f / PB;
    const int i = i0 + j;
    if (i < ni) data[base + i + e * es] = res[j * ps + e];
  }
}

// ---------------------------------------------------------------------------
// Gamma convolution + divergence residual, per wavevector
// ---------------------------------------------------------------------------
struct GammaArgs {
  int Nx, Ny, Nz, nr;
  double C_ref[81];      // (l,n,m,o) at 27l+9n+3m+o
  double size[3];        // for the continuous first derivative
  double rescale[3];     // geom_size / scaled_geom_size (divergence)
  int derivative;        // 0 continuous (on the fly), else table
  const double* xi_tab;  // (3 complex, K) interleaved, derivative != 0
  const double* aim;     // 9 doubles (column-major): DC value
  int do_gamma;
};
// one RVE of a batched Gamma convolution: its spectrum (9 planes of K
// complex), arguments and divergence partial sums
struct GammaBatchEntry {
  void* data;
  GammaArgs g;
  double* div_partial;
};

// xi of wavevector (i, y, z): continuous derivative on the fly (xi1st zeroes
// each component at the Nyquist index of an even dimension, xi2nd does not),
// otherwise from the table
template <typename T>
__device__ __forceinline__ void xi_at(const GammaArgs& g, int i, int y, int z, cpx<double>* xi,
                                      bool zero_nyquist = true) {
  if (g.derivative == 0) {
    const int ks[3] = {i, (y > g.Ny / 2) ? y - g.Ny : y, (z > g.Nz / 2) ? z - g.Nz : z};
    const int idx[3] = {i, y, z};
    const int cells[3] = {g.Nx, g.Ny, g.Nz};
    const double TAU = 6.283185307179586476925286766559;
#pragma unroll
    for (int p = 0; p < 3; ++p) {
      const bool nyq = zero_nyquist && (cells[p] % 2 == 0) && (idx[p] == cells[p] / 2);
      xi[p] = {0.0, nyq ? 0.0 : TAU * ks[p] / g.size[p]};
    }
  } else {
    const int64_t kk = i + (int64_t)g.nr * (y + (int64_t)g.Ny * z);
#pragma unroll
    for (int p = 0; p < 3; ++p) xi[p] = {g.xi_tab[6 * kk + 2 * p], g.xi_tab[6 * kk + 2 * p + 1]};
  }
}

// One thread per wavevector: divergence contribution of P_hat (block
// partial sums, fixed tree) and P_hat <- Gamma_hat : P_hat (or the aim at
// k = 0).  data: 9 complex planes of K.
template <typename T>
__device__ __forceinline__ void gamma_div_block(cpx<T>* data, int64_t K, const GammaArgs& g, double* div_partial,
                                                unsigned blk, double* red) {
  const int64_t kk = (int64_t)blk * 256 + threadIdx.x;
  double contrib = 0.0;
  if (kk < K) {
    const int i = (int)(kk % g.nr);
    const int64_t yz = kk / g.nr;
    const int y = (int)(yz % g.Ny), z = (int)(yz / g.Ny);
    cpx<double> xi[3];
    xi_at<T>(g, i, y, z, xi);
    cpx<double> P[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) {
      const cpx<T> v = data[q * K + kk];
      P[q] = {(double)v.x, (double)v.y};
    }
    // divergence (mech_base.cpp:323-387): |P_hat . (conj(xi) * L/L_scaled)|^2,
    // weight 2 for 0 < i < nr-1
    {
      cpx<double> vv[3];
#pragma unroll
      for (int n = 0; n < 3; ++n) vv[n] = {xi[n].x * g.rescale[n], -xi[n].y * g.rescale[n]};
      double sre = 0.0, sim = 0.0;
#pragma unroll
      for (int m = 0; m < 3; ++m) {
        cpx<double> pr = {0.0, 0.0};
#pragma unroll
        for (int n = 0; n < 3; ++n) pr = cadd(pr, cmul(P[m + 3 * n], vv[n]));
        sre += pr.x * pr.x;
        sim += pr.y * pr.y;
      }
      const double w = (i > 0 && i < g.nr - 1) ? 2.0 : 1.0;
      contrib = w * (sre + sim);
    }
    if (g.do_gamma) {
      cpx<double> out[9];
      if (kk == 0) {
#pragma unroll
        for (int q = 0; q < 9; ++q) out[q] = {g.aim[q], 0.0};
      } else {
        // Z(l,m) = sum_no C_ref(l,n,m,o) conj(-xi_n) xi_o
        cpx<double> dy[9];
#pragma unroll
        for (int n = 0; n < 3; ++n)
#pragma unroll
          for (int o = 0; o < 3; ++o) dy[3 * n + o] = cmul(cpx<double>{-xi[n].x, xi[n].y}, xi[o]);
        cpx<double> Z[9];
#pragma unroll
        for (int l = 0; l < 3; ++l)
#pragma unroll
          for (int m = 0; m < 3; ++m) {
            cpx<double> s = {0.0, 0.0};
#pragma unroll
            for (int n = 0; n < 3; ++n)
#pragma unroll
              for (int o = 0; o < 3; ++o) {
                const double c = g.C_ref[27 * l + 9 * n + 3 * m + o];
                s.x += c * dy[3 * n + o].x;
                s.y += c * dy[3 * n + o].y;
              }
            Z[3 * l + m] = s;
          }
        const double detRe = Z[0].x * (Z[4].x * Z[8].x - Z[5].x * Z[7].x) -
                             Z[1].x * (Z[3].x * Z[8].x - Z[5].x * Z[6].x) +
                             Z[2].x * (Z[3].x * Z[7].x - Z[4].x * Z[6].x);
        if (fabs(detRe) > 1e-16) {
          // complex Cramer inverse of Z
          cpx<double> cof[9];
          cof[0] = csub(cmul(Z[4], Z[8]), cmul(Z[5], Z[7]));
          cof[3] = csub(cmul(Z[5], Z[6]), cmul(Z[3], Z[8]));
          cof[6] = csub(cmul(Z[3], Z[7]), cmul(Z[4], Z[6]));
          cof[1] = csub(cmul(Z[2], Z[7]), cmul(Z[1], Z[8]));
          cof[4] = csub(cmul(Z[0], Z[8]), cmul(Z[2], Z[6]));
          cof[7] = csub(cmul(Z[1], Z[6]), cmul(Z[0], Z[7]));
          cof[2] = csub(cmul(Z[1], Z[5]), cmul(Z[2], Z[4]));
          cof[5] = csub(cmul(Z[2], Z[3]), cmul(Z[0], Z[5]));
          cof[8] = csub(cmul(Z[0], Z[4]), cmul(Z[1], Z[3]));
          const cpx<double> det = cadd(cadd(cmul(Z[0], cof[0]), cmul(Z[1], cof[3])), cmul(Z[2], cof[6]));
          const double dd = det.x * det.x + det.y * det.y;
          const cpx<double> idet = {det.x / dd, -det.y / dd};
          // w_n = sum_o P(n,o) conj(-xi_o); u = Z^-1 w; out(l,m) = u_l xi_m
          cpx<double> w[3], u[3];
#pragma unroll
          for (int n = 0; n < 3; ++n) {
            cpx<double> s = {0.0, 0.0};
#pragma unroll
            for (int o = 0; o < 3; ++o) s = cadd(s, cmul(P[n + 3 * o], cpx<double>{-xi[o].x, xi[o].y}));
            w[n] = s;
          }
#pragma unroll
          for (int l = 0; l < 3; ++l) {
            cpx<double> s = {0.0, 0.0};
#pragma unroll
            for (int n = 0; n < 3; ++n) s = cadd(s, cmul(cmul(cof[3 * l + n], idet), w[n]));
            u[l] = s;
          }
#pragma unroll
          for (int l = 0; l < 3; ++l)
#pragma unroll
            for (int m = 0; m < 3; ++m) out[l + 3 * m] = cmul(u[l], xi[m]);
        } else {
#pragma unroll
          for (int q = 0; q < 9; ++q) out[q] = {0.0, 0.0};
        }
      }
#pragma unroll
      for (int q = 0; q < 9; ++q) data[q * K + kk] = {(T)out[q].x, (T)out[q].y};
    }
  }
  if (div_partial) {
    red[threadIdx.x] = contrib;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
      if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
      __syncthreads();
    }
    if (threadIdx.x == 0) div_partial[blk] = red[0];
  }
}

template <typename T>
__global__ void __launch_bounds__(256) gamma_div_kernel(cpx<T>* data, int64_t K, GammaArgs g, double* div_partial) {
  __shared__ double red[256];
  gamma_div_block<T>(data, K, g, div_partial, blockIdx.x, red);
}

// A batch of RVEs with the same grid: blockIdx.y is the RVE, each block does
// what its own launch does (same wavevectors, same partial sums).  The table
// is read through the constant address space (invariant loads, as for kernel
// arguments).
template <typename T>
__global__ void __launch_bounds__(256) gamma_div_batch_kernel(const GammaBatchEntry* __restrict__ tab, int64_t K) {
  __shared__ double red[256];
#if defined(__HIP_DEVICE_COMPILE__)
  typedef __attribute__((address_space(4))) const GammaBatchEntry ConstEntry;
  const GammaBatchEntry e = *(const ConstEntry*)(tab + blockIdx.y);
#else
  const GammaBatchEntry e = tab[blockIdx.y];
#endif
  gamma_div_block<T>(reinterpret_cast<cpx<T>*>(e.data), K, e.g, e.div_partial, blockIdx.x, red);
}

// ---------------------------------------------------------------------------
// Cell and node coordinates (base_update_coords, mech_base.cpp:45-154)
// ---------------------------------------------------------------------------
// Per wavevector of F_hat (9 planes): displacement fluctuation spectrum
// u_hat_l = sum_m F_hat(l,m) xi2nd_m / sum_p conj(-xi2nd_p) xi2nd_p into
// planes 0..2; the k = 0 thread keeps the 9 sums of F (its DC value).
__global__ void __launch_bounds__(256) coords_uhat_kernel(cpx<double>* data, int64_t K, GammaArgs g,
                                                          double* fsum) {
  const int64_t kk = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (kk >= K) return;
  const int i = (int)(kk % g.nr);
  const int64_t yz = kk / g.nr;
  const int y = (int)(yz % g.Ny), z = (int)(yz / g.Ny);
  cpx<double> F[9];
#pragma unroll
  for (int q = 0; q < 9; ++q) F[q] = data[q * K + kk];
  cpx<double> u[3] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
  if (kk == 0) {
#pragma unroll
    for (int q = 0; q < 9; ++q) fsum[q] = F[q].x;
  } else {
    cpx<double> xi[3];
    xi_at<double>(g, i, y, z, xi, false);
    cpx<double> den = {0.0, 0.0};
#pragma unroll
    for (int p = 0; p < 3; ++p) den = cadd(den, cmul(cpx<double>{-xi[p].x, xi[p].y}, xi[p]));
    const double dd = den.x * den.x + den.y * den.y;
#pragma unroll
    for (int l = 0; l < 3; ++l) {
      cpx<double> s = {0.0, 0.0};
#pragma unroll
      for (int m = 0; m < 3; ++m) s = cadd(s, cmul(F[l + 3 * m], xi[m]));
      u[l] = {(s.x * den.x + s.y * den.y) / dd, (s.y * den.x - s.x * den.y) / dd};
    }
  }
#pragma unroll
  for (int l = 0; l < 3; ++l) data[l * K + kk] = u[l];
}

struct CoordArgs {
  int Nx, Ny, Nz;
  double step[3];
  double wgt;
};

// nodes: F_avg x + mean of the 8 periodic neighbour cells' fluctuation
__global__ void coords_nodes_kernel(const double* u, const double* fsum, CoordArgs a, double* x_n) {
  const int64_t nn = (int64_t)(a.Nx + 1) * (a.Ny + 1) * (a.Nz + 1);
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nn) return;
  const int i = (int)(t % (a.Nx + 1));
  const int j = (int)((t / (a.Nx + 1)) % (a.Ny + 1));
  const int k = (int)(t / ((int64_t)(a.Nx + 1) * (a.Ny + 1)));
  const double pos[3] = {i * a.step[0], j * a.step[1], k * a.step[2]};
  const int64_t N = (int64_t)a.Nx * a.Ny * a.Nz;
  const int nb[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
#pragma unroll
  for (int p = 0; p < 3; ++p) {
    double s = 0.0;
#pragma unroll
    for (int q = 0; q < 3; ++q) s += (fsum[p + 3 * q] * a.wgt) * pos[q];
    for (int n = 0; n < 8; ++n) {
      const int cx = (i + nb[n][0] - 1 + a.Nx) % a.Nx;
      const int cy = (j + nb[n][1] - 1 + a.Ny) % a.Ny;
      const int cz = (k + nb[n][2] - 1 + a.Nz) % a.Nz;
      s += u[p * N + cx + (int64_t)a.Nx * (cy + (int64_t)a.Ny * cz)] * 0.125;
    }
    x_n[3 * t + p] = s;
  }
}

// cell centres: F_avg x + fluctuation
__global__ void coords_points_kernel(const double* u, const double* fsum, CoordArgs a, double* x_p) {
  const int64_t N = (int64_t)a.Nx * a.Ny * a.Nz;
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= N) return;
  const int i = (int)(c % a.Nx);
  const int j = (int)((c / a.Nx) % a.Ny);
  const int k = (int)(c / ((int64_t)a.Nx * a.Ny));
  const double pos[3] = {(i + 0.5) * a.step[0], (j + 0.5) * a.step[1], (k + 0.5) * a.step[2]};
#pragma unroll
  for (int p = 0; p < 3; ++p) {
    double s = 0.0;
#pragma unroll
    for (int q = 0; q < 3; ++q) s += (fsum[p + 3 * q] * a.wgt) * pos[q];
    x_p[3 * c + p] = u[p * N + c] + s;
  }
}

// ---------------------------------------------------------------------------
// Deterministic reductions: fixed grid, fixed trees
// ---------------------------------------------------------------------------
constexpr int RED_BLOCKS = 512;

// partial[v*RED_BLOCKS + b] = sum over this block's grid-stride share of
// op(a_v, b_v): op 0 dot(a,b), op 1 sum (a-b)^2, op 2 sum a
struct RedArgs {
  int nvec;
  int64_t n;
  unsigned char op[MAX_RED_VEC];
  const void* a[MAX_RED_VEC];
  const void* b[MAX_RED_VEC];
};

template <typename T>
__global__ void __launch_bounds__(256) partial_reduce_kernel(RedArgs r, double* partial) {
  __shared__ double red[256];
  const int v = blockIdx.y;
  const T* a = reinterpret_cast<const T*>(r.a[v]);
  const T* b = reinterpret_cast<const T*>(r.b[v]);
  const int op = r.op[v];
  // a squared norm (b == a) reads its vector once
  const bool own = (op == 2) || (a == b);
  double s = 0.0;
  constexpr int64_t STRIDE = (int64_t)RED_BLOCKS * 256;
  int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  // U grid-stride elements per trip with their loads issued first (32 bytes
  // per thread and vector in flight); the sum is accumulated in the same
  // element order as one per trip
  constexpr int U = 32 / sizeof(T) < 4 ? 4 : 32 / sizeof(T);
  for (; t + (U - 1) * STRIDE < r.n; t += U * STRIDE) {
    T x[U], y[U];
#pragma unroll
    for (int u = 0; u < U; ++u) x[u] = a[t + u * STRIDE];
#pragma unroll
    for (int u = 0; u < U; ++u) y[u] = own ? x[u] : b[t + u * STRIDE];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const double xu = (double)x[u], yu = (double)y[u];
      if (op == 0) s += xu * yu;
      else if (op == 1) {
        const double d = xu - yu;
        s += d * d;
      } else s += xu;
    }
  }
  for (; t < r.n; t += STRIDE) {
    const double x = (double)a[t];
    const double y = own ? x : (double)b[t];
    if (op == 0) s += x * y;
    else if (op == 1) {
      const double d = x - y;
      s += d * d;
    } else s += x;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[(int64_t)v * RED_BLOCKS + blockIdx.x] = red[0];
}

// Reductions of several vectors against shared operands in one pass (the
// N-GMRES inner products): MODE 0: out[2i] = <u, v_i>, out[2i+1] = <w, v_i>;
// MODE 1: out[i] = sum (v_i - u)^2.  u and w are read once per group of
// SHARED_GROUP vectors instead of once per product.  The per-thread element
// order and the block tree are those of partial_reduce_kernel, so every
// result is bitwise the one of the corresponding single reduction.
constexpr int SHARED_GROUP = 16;
template <typename T>
struct SharedRedArgs {
  int64_t n;
  const T* u;
  const T* w;
  int nv;  // vectors in this launch, 1..SHARED_GROUP
  const T* v[SHARED_GROUP];
};

// T: element type of the vectors (float vectors are summed in double, as the
// same values held in double would be).  NV: vectors in the launch (loads
// without branches, registers for exactly these); a vector equal to u (the
// N-GMRES trial residual dotted with itself) is taken from u's own load.
template <int MODE, typename T, int NV>
__global__ void __launch_bounds__(256) shared_reduce_kernel(SharedRedArgs<T> r, double* partial) {
  constexpr int NO = MODE == 0 ? 2 : 1;  // outputs per vector
  constexpr int NACC = NO * NV;
  constexpr int CH = 8;  // accumulators per block-tree pass
  // elements per trip: more when the vectors are few (bytes in flight)
  constexpr int U = NV <= 4 ? 8 : (NV <= 8 ? 4 : 2);
  __shared__ double red[CH][256];
  const int t = threadIdx.x;
  double s[NACC];
#pragma unroll
  for (int j = 0; j < NACC; ++j) s[j] = 0.0;
  bool self[NV];
#pragma unroll
  for (int i = 0; i < NV; ++i) self[i] = r.v[i] == r.u;
  constexpr int64_t STRIDE = (int64_t)RED_BLOCKS * 256;
  int64_t e = (int64_t)blockIdx.x * 256 + t;
  // U grid-stride elements per trip, loads first, the sums accumulated in
  // the same element order as one per trip
  for (; e + (U - 1) * STRIDE < r.n; e += U * STRIDE) {
    T x[U], y[U], v[NV][U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      x[u] = r.u[e + u * STRIDE];
      y[u] = MODE == 0 ? r.w[e + u * STRIDE] : (T)0;
    }
#pragma unroll
    for (int i = 0; i < NV; ++i)
      if (!self[i]) {
#pragma unroll
        for (int u = 0; u < U; ++u) v[i][u] = r.v[i][e + u * STRIDE];
      } else {
#pragma unroll
        for (int u = 0; u < U; ++u) v[i][u] = x[u];
      }
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
      for (int i = 0; i < NV; ++i) {
        const double xu = (double)x[u], vi = (double)v[i][u];
        if constexpr (MODE == 0) {
          s[2 * i] += xu * vi;
          s[2 * i + 1] += (double)y[u] * vi;
        } else {
          const double d = vi - xu;
          s[i] += d * d;
        }
      }
  }
  for (; e < r.n; e += STRIDE) {
    const double x = (double)r.u[e];
    const double y = MODE == 0 ? (double)r.w[e] : 0.0;
#pragma unroll
    for (int i = 0; i < NV; ++i) {
      const double v = self[i] ? x : (double)r.v[i][e];
      if constexpr (MODE == 0) {
        s[2 * i] += x * v;
        s[2 * i + 1] += y * v;
      } else {
        const double d = v - x;
        s[i] += d * d;
      }
    }
  }
#pragma unroll
  for (int c0 = 0; c0 < NACC; c0 += CH) {
#pragma unroll
    for (int k = 0; k < CH; ++k) red[k][t] = c0 + k < NACC ? s[c0 + k] : 0.0;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
      if (t < w) {
#pragma unroll
        for (int k = 0; k < CH; ++k) red[k][t] += red[k][t + w];
      }
      __syncthreads();
    }
    if (t < CH && c0 + t < NACC) partial[(int64_t)(c0 + t) * RED_BLOCKS + blockIdx.x] = red[t][0];
    __syncthreads();
  }
}

// out[v] = sum_b partial[v*nb + b] (fixed tree), one block per v
__global__ void __launch_bounds__(256) final_reduce_kernel(const double* partial, int nb, double* out) {
  __shared__ double red[256];
  const int v = blockIdx.x;
  double s = 0.0;
  for (int t = threadIdx.x; t < nb; t += 256) s += partial[(int64_t)v * nb + t];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[v] = red[0];
    __threadfence_system();  // out may be pinned host memory, read once the stream is done
  }
}

// final reductions of a batch: blockIdx.y is the entry, blockIdx.x the
// vector, each as final_reduce_kernel (same tree)
__global__ void __launch_bounds__(256) final_reduce_batch_kernel(const BatchReduce* __restrict__ tab) {
  __shared__ double red[256];
  const BatchReduce e = tab[blockIdx.y];
  const int v = blockIdx.x;
  if (v >= e.nvec) return;
  double s = 0.0;
  for (int t = threadIdx.x; t < e.nb; t += 256) s += e.partial[(int64_t)v * e.nb + t];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) e.out[v] = red[0];
}

// component sums from x-pass row sums: block q of an item sums its rows
// (each thread a stride-256 sequence in order, its loads issued 16 at a
// time ahead of the adds, then the block tree)
__device__ __forceinline__ void rowsum_block(const RowSumItem& e, int q, double* red) {
  const double* r = e.rowsum + (int64_t)q * e.nrow;
  double s = 0.0;
  constexpr int U = 16;
  for (int64_t t0 = threadIdx.x; t0 < e.nrow; t0 += 256 * U) {
    double x[U];
#pragma unroll
    for (int u = 0; u < U; ++u) x[u] = (t0 + 256 * u < e.nrow) ? r[t0 + 256 * u] : 0.0;
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (t0 + 256 * u < e.nrow) s += x[u];
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) e.out[q] = red[0];
}
__global__ void __launch_bounds__(256) rowsum_reduce_kernel(RowSumItem e) {
  __shared__ double red[256];
  rowsum_block(e, blockIdx.x, red);
}
__global__ void __launch_bounds__(256) rowsum_reduce_batch_kernel(const RowSumItem* __restrict__ tab) {
  __shared__ double red[256];
  rowsum_block(tab[blockIdx.y], blockIdx.x, red);
}

// copies of a batch (gathers and scatters between RVE buffers and a group's
// contiguous arena): blockIdx.y is the entry
__global__ void __launch_bounds__(256) copy_batch_kernel(const BatchCopy* __restrict__ tab) {
  const BatchCopy e = tab[blockIdx.y];
  const uint32_t* src = static_cast<const uint32_t*>(e.src);
  uint32_t* dst = static_cast<uint32_t*>(e.dst);
  if (((reinterpret_cast<uintptr_t>(src) | reinterpret_cast<uintptr_t>(dst)) & 7) == 0) {
    // both 8-byte aligned: pairs of words, then an odd last word
    const uint2* s2 = reinterpret_cast<const uint2*>(src);
    uint2* d2 = reinterpret_cast<uint2*>(dst);
    const int64_t n2 = e.words / 2;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n2; i += (int64_t)gridDim.x * 256) d2[i] = s2[i];
    if ((e.words & 1) && blockIdx.x == 0 && threadIdx.x == 0) dst[e.words - 1] = src[e.words - 1];
  } else {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < e.words; i += (int64_t)gridDim.x * 256)
      dst[i] = src[i];
  }
}

// ---------------------------------------------------------------------------
// Streaming field ops
// ---------------------------------------------------------------------------
// out = last + rate*dt  (forward_field, mech_base.cpp:217)
__global__ void fwd_field_kernel(const double* last, const double* rate, double dt, double* out, int64_t n) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) out[t] = last[t] + rate[t] * dt;
}
// field[q*N + c] -= diff[q]
__global__ void sub_plane_const_kernel(double* f, const double* diff, int64_t N) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < 9 * N) f[t] -= diff[t / N];
}
// forward_field correction: diff[q] = sums[q]*wgt - aim[q]
__global__ void fwd_diff_kernel(const double* sums, double wgt, const double* aim, double* diff) {
  const int q = threadIdx.x;
  if (q < 9) diff[q] = sums[q] * wgt - aim[q];
}
// calculate_rate (mech_base.cpp:479-497)
__global__ void rate_kernel(int het, const double* f0, const double* f, double dt, const double* av, double* out,
                            int64_t N) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < 9 * N) out[t] = het ? (f[t] - f0[t]) / dt : av[t / N];
}
// y = a*x + b*y0 (general axpby into y; a float y0 enters as its double value)
template <typename TY>
__global__ void axpby_kernel(double a, const double* x, double b, const TY* y0, double* y, int64_t n) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) y[t] = a * x[t] + b * (double)y0[t];
}
// XA = XM*(1 - alpha_total) + sum_i beta_i Xdot_i   (VecScale + VecMAXPY)
struct CombArgs {
  int l;
  double c0;
  double beta[MAX_RED_VEC];
  const double* X[MAX_RED_VEC];
};
__global__ void lincomb_kernel(const double* XM, CombArgs a, double* XA, int64_t n) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  double v = XM[t] * a.c0;
  for (int i = 0; i < a.l; ++i) v += a.beta[i] * a.X[i][t];
  XA[t] = v;
}
// XA = XM*c0 + sum_i beta_i Xdot_i (the sum in lincomb_kernel's order) and,
// in the same pass, the partials of |XA - XM|^2 and |XA - Xdot_i|^2 (the
// N-GMRES distances, shared_reduce_kernel<1> with u = XA, v = {XM, Xdot_i}):
// the history is read once instead of twice.  Element order per thread and
// the block tree are those of the shared reduction, so the distances are
// bitwise its results; l + 1 <= SHARED_GROUP.
struct CombNormArgs {
  int64_t n;
  int l;
  double c0;
  const double* XM;
  double* XA;
  double beta[SHARED_GROUP];
  const double* X[SHARED_GROUP];
};
// L: history vectors in the launch (loads without branches); U elements per
// trip, each computed as alone (the same results for any U)
template <int L>
__global__ void __launch_bounds__(256) lincomb_norms_kernel(CombNormArgs r, double* partial) {
  constexpr int CH = 8;
  constexpr int NOUT = L + 1;
  constexpr int U = L <= 4 ? 4 : 2;
  __shared__ double red[CH][256];
  const int t = threadIdx.x;
  double s[NOUT];
#pragma unroll
  for (int j = 0; j < NOUT; ++j) s[j] = 0.0;
  constexpr int64_t STRIDE = (int64_t)RED_BLOCKS * 256;
  auto one = [&](double xm, const double* h, int64_t e) {
    // the product xm c0 rounded on its own, then one fused multiply-add per
    // history vector (as in launch_lincomb, whatever L lets the compiler see)
    double v = xm * r.c0;
#pragma unroll
    for (int i = 0; i < L; ++i) v = __builtin_fma(r.beta[i], h[i], v);
    r.XA[e] = v;
    const double d0 = xm - v;
    s[0] += d0 * d0;
#pragma unroll
    for (int i = 0; i < L; ++i) {
      const double d = h[i] - v;
      s[1 + i] += d * d;
    }
  };
  int64_t e = (int64_t)blockIdx.x * 256 + t;
  for (; e + (U - 1) * STRIDE < r.n; e += U * STRIDE) {
    double xm[U], h[U][L];
#pragma unroll
    for (int u = 0; u < U; ++u) xm[u] = r.XM[e + u * STRIDE];
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
      for (int i = 0; i < L; ++i) h[u][i] = r.X[i][e + u * STRIDE];
#pragma unroll
    for (int u = 0; u < U; ++u) one(xm[u], h[u], e + u * STRIDE);
  }
  for (; e < r.n; e += STRIDE) {
    double h[L];
#pragma unroll
    for (int i = 0; i < L; ++i) h[i] = r.X[i][e];
    one(r.XM[e], h, e);
  }
#pragma unroll
  for (int c0 = 0; c0 < NOUT; c0 += CH) {
#pragma unroll
    for (int k = 0; k < CH; ++k) red[k][t] = c0 + k < NOUT ? s[c0 + k] : 0.0;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
      if (t < w) {
#pragma unroll
        for (int k = 0; k < CH; ++k) red[k][t] += red[k][t + w];
      }
      __syncthreads();
    }
    if (t < CH && c0 + t < NOUT) partial[(int64_t)(c0 + t) * RED_BLOCKS + blockIdx.x] = red[t][0];
    __syncthreads();
  }
}

// SoA (9 planes of N, component q = i+3j) <-> AoS (3,3,N) column-major
__global__ void soa_to_aos_kernel(const double* soa, double* aos, int64_t N, int ncomp) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < ncomp * N) {
    const int64_t c = t / ncomp;
    const int q = (int)(t % ncomp);
    aos[t] = soa[q * N + c];
  }
}
__global__ void aos_to_soa_kernel(const double* aos, double* soa, int64_t N, int ncomp) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < ncomp * N) {
    const int64_t c = t / ncomp;
    const int q = (int)(t % ncomp);
    soa[q * N + c] = aos[t];
  }
}
__global__ void f64_to_f32_kernel(const double* a, float* b, int64_t n) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) b[t] = (float)a[t];
}
__global__ void f32_to_f64_kernel(const float* a, double* b, int64_t n) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) b[t] = (double)a[t];
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
static inline unsigned nblk(int64_t n, int b = 256) { return (unsigned)((n + b - 1) / b); }

static FftPassDesc make_desc(int n, int sign) {
  FftPassDesc d{};
  d.n = n;
  d.sign = sign;
  int m = n, k = 0;
  while (m % 4 == 0) {
    d.rad[k++] = 4;
    m /= 4;
  }
  while (m % 2 == 0) {
    d.rad[k++] = 2;
    m /= 2;
  }
  for (int p : {3, 5, 7, 11, 13})
    while (m % p == 0) {
      d.rad[k++] = p;
      m /= p;
    }
  if (m != 1) throw std::runtime_error("FFT length " + std::to_string(n) + " has a prime factor above 13");
  d.nrad = k;
  return d;
}

// Pencil kernels for the supported lengths (compile-time N); false: use the
// generic Stockham kernels.  bs: threads per block (pencils per block =
// bs / threads per pencil); any valid value gives the same results.
#define SCA_PENCIL_LENGTHS(X)                                                                          \
  X(8) X(16) X(32) X(64) X(128) X(256) X(512) X(12) X(24) X(48) X(96) X(192) X(384) X(18) X(36) X(72) X(144) \
  X(288) X(27)

constexpr size_t MAX_BLOCK_LDS = 64 * 1024;

template <typename T, int N>
static size_t pencil_lds(int bs) {
  return (size_t)pencil::Shape<T, N>::lds(bs) * sizeof(cpx<T>);
}

// threads per pencil of a pencil length (0: not a pencil length)
template <typename T>
static int pencil_tp(int n) {
  switch (n) {
#define SCA_CASE(N) \
  case N: return pencil::Shape<T, N>::TP;
    SCA_PENCIL_LENGTHS(SCA_CASE)
#undef SCA_CASE
    default: return 0;
  }
}
template <typename T>
static size_t pencil_lds_of(int n, int bs) {
  switch (n) {
#define SCA_CASE(N) \
  case N: return pencil_lds<T, N>(bs);
    SCA_PENCIL_LENGTHS(SCA_CASE)
#undef SCA_CASE
    default: return 0;
  }
}
// block sizes a pencil pass may take: powers of two from the pencil's
// threads (and 64) to 512 within the LDS limit
template <typename T>
static std::vector<int> pencil_block_sizes(int n) {
  std::vector<int> out;
  const int tp = pencil_tp<T>(n);
  if (tp == 0) return out;
  for (int bs = 64; bs <= pencil::MAX_PENCIL_BLOCK; bs *= 2)
    if (bs >= tp && pencil_lds_of<T>(n, bs) <= MAX_BLOCK_LDS) out.push_back(bs);
  return out;
}

// rows: set_rows rows per independently paired set (rows is a multiple)
template <typename T, typename TI>
static bool pencil_x_r2c(int n, const TI* in, cpx<T>* out, int64_t rows, int64_t set_rows, const cpx<T>* tw,
                         int bs, hipStream_t st, double* rowsum) {
  switch (n) {
#define SCA_CASE(N)                                                                                          \
  case N: {                                                                                                  \
    using S = pencil::Shape<T, N>;                                                                              \
    const int64_t pencils = rows / set_rows * ((set_rows + 1) / 2), P = bs / S::TP;                          \
    hipLaunchKernelGGL((pencil::x_r2c_kernel<T, TI, N>), dim3((unsigned)((pencils + P - 1) / P)), dim3(bs),   \
                       (pencil_lds<T, N>(bs)), st, in, out, rows, set_rows, tw, rowsum);                     \
    return true;                                                                                             \
  }
    SCA_PENCIL_LENGTHS(SCA_CASE)
#undef SCA_CASE
    default: return false;
  }
}

template <typename T, typename TO>
static bool pencil_x_c2r(int n, const cpx<T>* in, TO* out, int64_t rows, int64_t set_rows, const cpx<T>* tw,
                         T scale, int bs, hipStream_t st) {
  switch (n) {
#define SCA_CASE(N)                                                                                          \
  case N: {                                                                                                  \
    using S = pencil::Shape<T, N>;                                                                              \
    const int64_t pencils = rows / set_rows * ((set_rows + 1) / 2), P = bs / S::TP;                          \
    hipLaunchKernelGGL((pencil::x_c2r_kernel<T, TO, N>), dim3((unsigned)((pencils + P - 1) / P)), dim3(bs),   \
                       (pencil_lds<T, N>(bs)), st, in, out, rows, set_rows, tw, scale);                      \
    return true;                                                                                             \
  }
    SCA_PENCIL_LENGTHS(SCA_CASE)
#undef SCA_CASE
    default: return false;
  }
}

template <typename T>
static bool pencil_strided(int n, int sign, cpx<T>* data, int ni, unsigned no, int d1, int64_t s1, int64_t s2,
                           int64_t es, const cpx<T>* tw, int bs, hipStream_t st) {
  switch (n) {
#define SCA_CASE(N)                                                                                          \
  case N: {                                                                                                  \
    using S = pencil::Shape<T, N>;                                                                              \
    const int P = bs / S::TP;                                                                                \
    const int64_t npen = (int64_t)ni * no;                                                                   \
    const dim3 grid((unsigned)((npen + P - 1) / P));                                                         \
    if (sign < 0)                                                                                            \
      hipLaunchKernelGGL((pencil::strided_kernel<T, N, -1>), grid, dim3(bs), (pencil_lds<T, N>(bs)), st, data, \
                         ni, npen, d1, s1, s2, es, tw);                                                      \
    else                                                                                                     \
      hipLaunchKernelGGL((pencil::strided_kernel<T, N, 1>), grid, dim3(bs), (pencil_lds<T, N>(bs)), st, data,  \
                         ni, npen, d1, s1, s2, es, tw);                                                      \
    return true;                                                                                             \
  }
    SCA_PENCIL_LENGTHS(SCA_CASE)
#undef SCA_CASE
    default: return false;
  }
}

template <typename T>
static int pencils_per_block(int n) {
  int PB = 16;
  while (PB > 1 && (size_t)2 * PB * (n + 1) * sizeof(cpx<T>) > MAX_BLOCK_LDS) PB /= 2;
  return PB;
}
static int x_rows_per_block(int n) { return std::max(1, std::min(256, 2048 / n)); }

// launch shape of an axis pass: the pencil block size for pencil lengths, the
// pencils (rows) per block of the generic kernels otherwise; 0 = heuristic
template <typename T>
int FftPlan<T>::shape(int axis) const {
  if (cfg[axis] > 0) return cfg[axis];
  const int n = cells[axis];
  if (pencil_tp<T>(n) > 0) {  // 256 threads, fewer where the pencils' LDS rows would not fit
    int bs = 256;
    while (bs > 64 && bs > pencil_tp<T>(n) && pencil_lds_of<T>(n, bs) > MAX_BLOCK_LDS) bs /= 2;
    return bs;
  }
  return axis == 0 ? x_rows_per_block(n) : pencils_per_block<T>(n);
}

template <typename T>
void FftPlan<T>::init(const int cells_[3]) {
  for (int p = 0; p < 3; ++p) cells[p] = cells_[p];
  nr = cells[0] / 2 + 1;
  N = (int64_t)cells[0] * cells[1] * cells[2];
  K = (int64_t)nr * cells[1] * cells[2];
  for (int p = 0; p < 3; ++p) {
    const int n = cells[p];
    std::vector<T> h(2 * n);
    for (int t = 0; t < n; ++t) {
      const double ang = 2.0 * M_PI * (double)t / (double)n;
      h[2 * t] = (T)std::cos(ang);
      h[2 * t + 1] = (T)-std::sin(ang);
    }
    tw[p] = static_cast<T*>(dev_alloc(2 * n * sizeof(T)));
    SCA_HIP(hipMemcpyAsync(tw[p], h.data(), 2 * n * sizeof(T), hipMemcpyHostToDevice, hipStreamPerThread));
    SCA_HIP(hipStreamSynchronize(hipStreamPerThread));
    make_desc(n, -1);  // validate
  }
}
template <typename T>
void FftPlan<T>::release() {
  for (int p = 0; p < 3; ++p)
    if (tw[p]) {
      dev_free(tw[p]);
      tw[p] = nullptr;
    }
}

template <typename T>
bool FftPlan<T>::row_sums_available() const {
  switch (cells[0]) {
#define SCA_CASE(N) case N:
    SCA_PENCIL_LENGTHS(SCA_CASE)
#undef SCA_CASE
    return true;
    default: return false;
  }
}

template <typename T>
void FftPlan<T>::x_pass_fwd(const void* real, bool real_is_double, cpx<T>* S, int ncomp, int shp, hipStream_t st,
                            int set_comp, double* rowsum) const {
  const int Nx = cells[0], Ny = cells[1], Nz = cells[2];
  const int64_t rows = (int64_t)ncomp * Ny * Nz;
  const int64_t set_rows = (int64_t)(set_comp > 0 ? set_comp : ncomp) * Ny * Nz;
  if (rows % set_rows != 0) throw std::runtime_error("FFT: components not a multiple of the field size");
  const cpx<T>* t0 = reinterpret_cast<const cpx<T>*>(tw[0]);
  if (rowsum && (!real_is_double || !row_sums_available()))
    throw std::runtime_error("FFT: row sums need the pencil x pass over double input");
  bool done;
  if (real_is_double)
    done = pencil_x_r2c<T, double>(Nx, static_cast<const double*>(real), S, rows, set_rows, t0, shp, st, rowsum);
  else done = pencil_x_r2c<T, T>(Nx, static_cast<const T*>(real), S, rows, set_rows, t0, shp, st, nullptr);
  if (!done) {
    FftPassDesc d = make_desc(Nx, -1);
    const size_t sh = 2 * (size_t)shp * Nx * sizeof(cpx<T>);
    if (real_is_double)
      hipLaunchKernelGGL((fft_x_r2c<T, double>), dim3(nblk(rows, shp)), dim3(256), sh, st,
                         static_cast<const double*>(real), S, rows, d, t0, shp);
    else
      hipLaunchKernelGGL((fft_x_r2c<T, T>), dim3(nblk(rows, shp)), dim3(256), sh, st, static_cast<const T*>(real), S,
                         rows, d, t0, shp);
  }
  SCA_HIP(hipGetLastError());
}

template <typename T>
template <typename TI>
void FftPlan<T>::forward(const TI* real, T* spec, int ncomp, hipStream_t st, int set_comp, double* rowsum) const {
  cpx<T>* S = reinterpret_cast<cpx<T>*>(spec);
  x_pass_fwd(real, sizeof(TI) == sizeof(double), S, ncomp, shape(0), st, set_comp, rowsum);
  strided_passes(S, ncomp, -1, st);
}

template <typename T>
void FftPlan<T>::strided_pass(cpx<T>* S, int ncomp, int sign, int axis, int shp, hipStream_t st) const {
  const int Ny = cells[1], Nz = cells[2];
  const int n = cells[axis];
  // y: pencils (c, z, i) with element stride nr; z: pencils (c, y, i), stride nr*Ny
  const unsigned no = (unsigned)(ncomp * (axis == 1 ? Nz : Ny));
  const int d1 = axis == 1 ? Nz : Ny;
  const int64_t s2 = axis == 1 ? (int64_t)nr * Ny : (int64_t)nr;
  const int64_t es = axis == 1 ? (int64_t)nr : (int64_t)nr * Ny;
  const cpx<T>* t = reinterpret_cast<const cpx<T>*>(tw[axis]);
  if (!pencil_strided<T>(n, sign, S, nr, no, d1, K, s2, es, t, shp, st)) {
    FftPassDesc d = make_desc(n, sign);
    dim3 grid((nr + shp - 1) / shp, no);
    hipLaunchKernelGGL(fft_strided<T>, grid, dim3(256), 2 * (size_t)shp * (n + 1) * sizeof(cpx<T>), st, S, nr, d1, K,
                       s2, es, d, t, shp);
  }
  SCA_HIP(hipGetLastError());
}

template <typename T>
void FftPlan<T>::strided_passes(void* Sv, int ncomp, int sign, hipStream_t st) const {
  cpx<T>* S = reinterpret_cast<cpx<T>*>(Sv);
  if (cells[1] > 1) strided_pass(S, ncomp, sign, 1, shape(1), st);
  if (cells[2] > 1) strided_pass(S, ncomp, sign, 2, shape(2), st);
}

template <typename T>
template <typename TO>
void FftPlan<T>::backward(T* spec, TO* real, int ncomp, double scale, hipStream_t st, int set_comp) const {
  const int Nx = cells[0], Ny = cells[1], Nz = cells[2];
  cpx<T>* S = reinterpret_cast<cpx<T>*>(spec);
  const int64_t set_rows = (int64_t)(set_comp > 0 ? set_comp : ncomp) * Ny * Nz;
  if (((int64_t)ncomp * Ny * Nz) % set_rows != 0) throw std::runtime_error("FFT: components not a multiple of the field size");
  strided_passes(S, ncomp, +1, st);
  const int shp = shape(0);
  if (pencil_x_c2r<T, TO>(Nx, S, real, (int64_t)ncomp * Ny * Nz, set_rows, reinterpret_cast<const cpx<T>*>(tw[0]),
                          (T)scale, shp, st)) {
    SCA_HIP(hipGetLastError());
    return;
  }
  FftPassDesc d = make_desc(Nx, +1);
  const int64_t rows = (int64_t)ncomp * Ny * Nz;
  const size_t sh = 2 * (size_t)shp * Nx * sizeof(cpx<T>);
  hipLaunchKernelGGL((fft_x_c2r<T, TO>), dim3(nblk(rows, shp)), dim3(256), sh, st, S, real, rows, d,
                     reinterpret_cast<const cpx<T>*>(tw[0]), (T)scale, shp);
  SCA_HIP(hipGetLastError());
}

// Autotuner: for every axis pass, each launch shape (pencil block sizes, or
// the generic kernels' pencils per block within the LDS limit) is timed on
// scratch arrays of the plan's size for ncomp components, and the fastest
// kept.  Shapes change no arithmetic, so a tuned plan computes bitwise what
// the heuristic one does.  Results are kept per (precision, grid, ncomp) for
// the process: later plans of the same kind take them without timing.
namespace {
std::mutex g_tune_mu;
std::map<std::tuple<int, int, int, int, int>, std::array<int, 3>> g_tuned;
}  // namespace

template <typename T>
void FftPlan<T>::tune(int ncomp, hipStream_t st) {
  const auto key = std::make_tuple((int)sizeof(T), cells[0], cells[1], cells[2], ncomp);
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tuned.find(key);
    if (it != g_tuned.end()) {
      for (int a = 0; a < 3; ++a) cfg[a] = it->second[a];
      return;
    }
  }
  double* real = nullptr;
  T* spec = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  std::array<int, 3> best = {0, 0, 0};
  try {
    real = static_cast<double*>(dev_alloc((size_t)ncomp * N * sizeof(double)));
    spec = static_cast<T*>(dev_alloc((size_t)2 * ncomp * K * sizeof(T)));
    SCA_HIP(hipMemsetAsync(real, 0, (size_t)ncomp * N * sizeof(double), st));
    SCA_HIP(hipMemsetAsync(spec, 0, (size_t)2 * ncomp * K * sizeof(T), st));
    SCA_HIP(hipEventCreate(&e0));
    SCA_HIP(hipEventCreate(&e1));
    cpx<T>* S = reinterpret_cast<cpx<T>*>(spec);
    for (int axis = 0; axis < 3; ++axis) {
      const int n = cells[axis];
      if (axis > 0 && n <= 1) continue;
      std::vector<int> cands = pencil_block_sizes<T>(n);
      if (cands.empty()) {  // generic kernels: pencils (rows) per block
        const int lim = axis == 0 ? 256 : 64;
        for (int pb = 1; pb <= lim; pb *= 2) {
          const size_t lds = 2 * (size_t)pb * (n + (axis == 0 ? 0 : 1)) * sizeof(cpx<T>);
          if (lds <= MAX_BLOCK_LDS) cands.push_back(pb);
        }
      }
      float best_ms = 1e30f;
      for (int c : cands) {
        auto run = [&] {
          if (axis == 0) x_pass_fwd(real, true, S, ncomp, c, st, 0);
          else strided_pass(S, ncomp, -1, axis, c, st);
        };
        run();  // warm-up
        float ms_min = 1e30f;
        for (int rep = 0; rep < 3; ++rep) {
          SCA_HIP(hipEventRecord(e0, st));
          run();
          SCA_HIP(hipEventRecord(e1, st));
          SCA_HIP(hipEventSynchronize(e1));
          float ms = 0.0f;
          SCA_HIP(hipEventElapsedTime(&ms, e0, e1));
          ms_min = std::min(ms_min, ms);
        }
        if (ms_min < best_ms) {
          best_ms = ms_min;
          best[axis] = c;
        }
      }
    }
  } catch (...) {
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipStreamSynchronize(st);
    dev_free(real);
    dev_free(spec);
    throw;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  dev_free(real);  // the timing runs are complete (the last event was waited for)
  dev_free(spec);
  for (int a = 0; a < 3; ++a) cfg[a] = best[a];
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tuned[key] = best;
}

template struct FftPlan<double>;
template struct FftPlan<float>;
template void FftPlan<double>::forward<double>(const double*, double*, int, hipStream_t, int, double*) const;
template void FftPlan<double>::backward<double>(double*, double*, int, double, hipStream_t, int) const;
template void FftPlan<float>::forward<double>(const double*, float*, int, hipStream_t, int, double*) const;
template void FftPlan<float>::backward<double>(float*, double*, int, double, hipStream_t, int) const;
template void FftPlan<float>::forward<float>(const float*, float*, int, hipStream_t, int, double*) const;
template void FftPlan<float>::backward<float>(float*, float*, int, double, hipStream_t, int) const;

static GammaArgs gamma_args(const GammaHost& gh, const double* d_xi_tab, const double* d_aim) {
  GammaArgs g;
  g.Nx = gh.cells[0];
  g.Ny = gh.cells[1];
  g.Nz = gh.cells[2];
  g.nr = gh.cells[0] / 2 + 1;
  for (int i = 0; i < 81; ++i) g.C_ref[i] = gh.C_ref[i];
  for (int p = 0; p < 3; ++p) {
    g.size[p] = gh.size[p];
    g.rescale[p] = gh.rescale[p];
  }
  g.derivative = gh.derivative;
  g.xi_tab = d_xi_tab;
  g.aim = d_aim;
  g.do_gamma = gh.do_gamma;
  return g;
}

template <typename T>
void launch_gamma_div(T* spec, int64_t K, const GammaHost& gh, const double* d_aim, const double* d_xi_tab,
                      double* div_partial, hipStream_t st) {
  GammaArgs g = gamma_args(gh, d_xi_tab, d_aim);
  hipLaunchKernelGGL(gamma_div_kernel<T>, dim3(gamma_blocks(K)), dim3(256), 0, st, reinterpret_cast<cpx<T>*>(spec), K,
                     g, div_partial);
  SCA_HIP(hipGetLastError());
}
template void launch_gamma_div<double>(double*, int64_t, const GammaHost&, const double*, const double*, double*,
                                       hipStream_t);
template void launch_gamma_div<float>(float*, int64_t, const GammaHost&, const double*, const double*, double*,
                                      hipStream_t);
int gamma_blocks(int64_t K) { return (int)((K + 255) / 256); }

void launch_coords_uhat(double* spec, int64_t K, const GammaHost& gh, const double* d_xi_tab, double* fsum,
                        hipStream_t st) {
  GammaArgs g = gamma_args(gh, d_xi_tab, nullptr);
  hipLaunchKernelGGL(coords_uhat_kernel, dim3(gamma_blocks(K)), dim3(256), 0, st,
                     reinterpret_cast<cpx<double>*>(spec), K, g, fsum);
  SCA_HIP(hipGetLastError());
}

static CoordArgs coord_args(const int cells[3], const double size[3]) {
  CoordArgs a;
  a.Nx = cells[0];
  a.Ny = cells[1];
  a.Nz = cells[2];
  for (int p = 0; p < 3; ++p) a.step[p] = size[p] / cells[p];
  a.wgt = 1.0 / ((double)cells[0] * cells[1] * cells[2]);
  return a;
}

void launch_coords_nodes(const double* u, const double* fsum, const int cells[3], const double size[3], double* x_n,
                         hipStream_t st) {
  const int64_t nn = (int64_t)(cells[0] + 1) * (cells[1] + 1) * (cells[2] + 1);
  hipLaunchKernelGGL(coords_nodes_kernel, dim3(nblk(nn)), dim3(256), 0, st, u, fsum, coord_args(cells, size), x_n);
  SCA_HIP(hipGetLastError());
}

void launch_coords_points(const double* u, const double* fsum, const int cells[3], const double size[3],
                          double* x_p, hipStream_t st) {
  const int64_t N = (int64_t)cells[0] * cells[1] * cells[2];
  hipLaunchKernelGGL(coords_points_kernel, dim3(nblk(N)), dim3(256), 0, st, u, fsum, coord_args(cells, size), x_p);
  SCA_HIP(hipGetLastError());
}

int red_blocks() { return RED_BLOCKS; }

template <typename T>
void launch_partial_reduce(const int* ops, int64_t n, int nvec, const T* const* a, const T* const* b, double* partial,
                           hipStream_t st) {
  if (nvec > MAX_RED_VEC) throw std::runtime_error("too many reduction vectors");
  RedArgs r{};
  r.nvec = nvec;
  r.n = n;
  for (int v = 0; v < nvec; ++v) {
    r.op[v] = (unsigned char)ops[v];
    r.a[v] = a[v];
    r.b[v] = b ? b[v] : a[v];
  }
  hipLaunchKernelGGL(partial_reduce_kernel<T>, dim3(RED_BLOCKS, nvec), dim3(256), 0, st, r, partial);
  SCA_HIP(hipGetLastError());
}
template <typename T>
void launch_shared_reduce(int mode, int64_t n, const T* u, const T* w, int nv, const T* const* v, double* partial,
                          hipStream_t st) {
  if (mode != 0 && mode != 1) throw std::runtime_error("shared reduction mode");
  const int no = mode == 0 ? 2 : 1;
  for (int g0 = 0; g0 < nv; g0 += SHARED_GROUP) {
    SharedRedArgs<T> r{};
    r.n = n;
    r.u = u;
    r.w = mode == 0 ? w : u;
    r.nv = std::min(SHARED_GROUP, nv - g0);
    for (int i = 0; i < SHARED_GROUP; ++i) r.v[i] = i < r.nv ? v[g0 + i] : u;  // unused slots stay dereferenceable
    double* out = partial + (int64_t)no * g0 * RED_BLOCKS;
    switch (mode * 100 + r.nv) {
#define SCA_SR(NVv)                                                                                              \
  case NVv:                                                                                                      \
    hipLaunchKernelGGL((shared_reduce_kernel<0, T, NVv>), dim3(RED_BLOCKS), dim3(256), 0, st, r, out);         \
    break;                                                                                                       \
  case 100 + NVv:                                                                                                \
    hipLaunchKernelGGL((shared_reduce_kernel<1, T, NVv>), dim3(RED_BLOCKS), dim3(256), 0, st, r, out);         \
    break;
      SCA_SR(1) SCA_SR(2) SCA_SR(3) SCA_SR(4) SCA_SR(5) SCA_SR(6) SCA_SR(7) SCA_SR(8)
      SCA_SR(9) SCA_SR(10) SCA_SR(11) SCA_SR(12) SCA_SR(13) SCA_SR(14) SCA_SR(15) SCA_SR(16)
#undef SCA_SR
      default: throw std::runtime_error("shared reduction size");
    }
    SCA_HIP(hipGetLastError());
  }
}
template void launch_shared_reduce<double>(int, int64_t, const double*, const double*, int, const double* const*,
                                           double*, hipStream_t);
template void launch_shared_reduce<float>(int, int64_t, const float*, const float*, int, const float* const*,
                                          double*, hipStream_t);

template void launch_partial_reduce<double>(const int*, int64_t, int, const double* const*, const double* const*,
                                            double*, hipStream_t);
template void launch_partial_reduce<float>(const int*, int64_t, int, const float* const*, const float* const*,
                                           double*, hipStream_t);
template void launch_partial_reduce<int>(const int*, int64_t, int, const int* const*, const int* const*, double*,
                                         hipStream_t);

// Cells whose tangents form C_minMaxAvg (spectral.cpp:189-207): the first
// cell of largest and of smallest |dPdF|^2, where norms within a relative
// window of the extreme count as tied (the norm is rotation invariant, so
// grains differing only in orientation tie exactly and rounding noise must
// not decide).  Pass 1: per-cell norms and block partial (max, min).
__global__ void __launch_bounds__(256) dpdf_norm_partial(const double* d, int64_t N, double* n2, double* pv) {
  __shared__ double vmx[256], vmn[256];
  double bmx = -1.0, bmn = 1.7976931348623157e308;
  for (int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x; c < N; c += (int64_t)RED_BLOCKS * 256) {
    double s = 0.0;
    for (int q = 0; q < 81; ++q) {
      const double x = d[q * N + c];
      s += x * x;
    }
    n2[c] = s;
    bmx = fmax(bmx, s);
    bmn = fmin(bmn, s);
  }
  vmx[threadIdx.x] = bmx;
  vmn[threadIdx.x] = bmn;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      vmx[threadIdx.x] = fmax(vmx[threadIdx.x], vmx[threadIdx.x + w]);
      vmn[threadIdx.x] = fmin(vmn[threadIdx.x], vmn[threadIdx.x + w]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    pv[2 * blockIdx.x] = vmx[0];
    pv[2 * blockIdx.x + 1] = vmn[0];
  }
}
// Pass 2: per block the first cell with n2 >= hi and the first with n2 <= lo (-1: none)
__global__ void __launch_bounds__(256) first_index_partial(const double* n2, int64_t N, double hi, double lo,
                                                           long long* pi) {
  __shared__ long long imx[256], imn[256];
  long long jmx = -1, jmn = -1;
  for (int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x; c < N; c += (int64_t)RED_BLOCKS * 256) {
    const double s = n2[c];
    if (jmx < 0 && s >= hi) jmx = c;
    if (jmn < 0 && s <= lo) jmn = c;
  }
  imx[threadIdx.x] = jmx;
  imn[threadIdx.x] = jmn;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      const long long a = imx[threadIdx.x], b = imx[threadIdx.x + w];
      imx[threadIdx.x] = (a < 0) ? b : ((b < 0) ? a : (a < b ? a : b));
      const long long e = imn[threadIdx.x], f = imn[threadIdx.x + w];
      imn[threadIdx.x] = (e < 0) ? f : ((f < 0) ? e : (e < f ? e : f));
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    pi[2 * blockIdx.x] = imx[0];
    pi[2 * blockIdx.x + 1] = imn[0];
  }
}
// block extremes of given per-cell norms (the tree of dpdf_norm_partial)
__global__ void __launch_bounds__(256) norm_extremes_partial(const double* n2, int64_t N, double* pv) {
  __shared__ double vmx[256], vmn[256];
  double bmx = -1.0, bmn = 1.7976931348623157e308;
  for (int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x; c < N; c += (int64_t)RED_BLOCKS * 256) {
    const double s = n2[c];
    bmx = fmax(bmx, s);
    bmn = fmin(bmn, s);
  }
  const int t = threadIdx.x;
  vmx[t] = bmx;
  vmn[t] = bmn;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) {
      vmx[t] = fmax(vmx[t], vmx[t + w]);
      vmn[t] = fmin(vmn[t], vmn[t + w]);
    }
    __syncthreads();
  }
  if (t == 0) {
    pv[2 * blockIdx.x] = vmx[0];
    pv[2 * blockIdx.x + 1] = vmn[0];
  }
}

static void select_from_extremes(const double* n2, int64_t N, double* pv, long long* pi, int64_t& imax,
                                 int64_t& imin, hipStream_t st);

void tangent_minmax_select(const double* dPdF, int64_t N, double* n2, double* pv, long long* pi, int64_t& imax,
                           int64_t& imin, hipStream_t st) {
  hipLaunchKernelGGL(dpdf_norm_partial, dim3(RED_BLOCKS), dim3(256), 0, st, dPdF, N, n2, pv);
  SCA_HIP(hipGetLastError());
  select_from_extremes(n2, N, pv, pi, imax, imin, st);
}

void tangent_select_norms(const double* n2, int64_t N, double* pv, long long* pi, int64_t& imax, int64_t& imin,
                          hipStream_t st) {
  hipLaunchKernelGGL(norm_extremes_partial, dim3(RED_BLOCKS), dim3(256), 0, st, n2, N, pv);
  SCA_HIP(hipGetLastError());
  select_from_extremes(n2, N, pv, pi, imax, imin, st);
}

// the global extremes from the block extremes in pv, then the first cell
// within TANGENT_TIE_REL of each
static void select_from_extremes(const double* n2, int64_t N, double* pv, long long* pi, int64_t& imax,
                                 int64_t& imin, hipStream_t st) {
  const int nb = RED_BLOCKS;
  std::vector<double> h(2 * nb);
  SCA_HIP(hipMemcpyAsync(h.data(), pv, 2 * nb * sizeof(double), hipMemcpyDeviceToHost, st));
  SCA_HIP(hipStreamSynchronize(st));
  double M = -1.0, m = 1.7976931348623157e308;
  for (int b = 0; b < nb; ++b) {
    M = std::max(M, h[2 * b]);
    m = std::min(m, h[2 * b + 1]);
  }
  const double hi = M * (1.0 - TANGENT_TIE_REL), lo = m * (1.0 + TANGENT_TIE_REL);
  hipLaunchKernelGGL(first_index_partial, dim3(nb), dim3(256), 0, st, n2, N, hi, lo, pi);
  SCA_HIP(hipGetLastError());
  std::vector<long long> hi_(2 * nb);
  SCA_HIP(hipMemcpyAsync(hi_.data(), pi, 2 * nb * sizeof(long long), hipMemcpyDeviceToHost, st));
  SCA_HIP(hipStreamSynchronize(st));
  imax = imin = -1;
  for (int b = 0; b < nb; ++b) {
    const long long x = hi_[2 * b], y = hi_[2 * b + 1];
    if (x >= 0 && (imax < 0 || x < imax)) imax = x;
    if (y >= 0 && (imin < 0 || y < imin)) imin = y;
  }
}
// dF_aim on the device: P_av = wgt * sums (rotated R P R^T when rotating),
// dF = S : P_av in the load frame, aim = R^T dF R in the field frame (active)
struct BcArgs {
  double S[81];  // (i,j,k,l) column-major
  double R[9];   // row-major
  double wgt;
  int rotate;
};
struct BcBatchEntry {
  const double* sums;
  BcArgs b;
  double* aim;
};
__device__ __forceinline__ void bc_aim_body(const double* sums, const BcArgs& b, double* aim) {
  double P[3][3], Pr[3][3], dF[3][3], t[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) P[i][j] = sums[i + 3 * j] * b.wgt;
  if (b.rotate) {
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) t[i][j] = b.R[3 * i] * P[0][j] + b.R[3 * i + 1] * P[1][j] + b.R[3 * i + 2] * P[2][j];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) Pr[i][j] = t[i][0] * b.R[3 * j] + t[i][1] * b.R[3 * j + 1] + t[i][2] * b.R[3 * j + 2];
  } else {
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) Pr[i][j] = P[i][j];
  }
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0.0;
      for (int n = 0; n < 3; ++n)
        for (int m = 0; m < 3; ++m) s += b.S[i + 3 * j + 9 * m + 27 * n] * Pr[m][n];
      dF[i][j] = s;
    }
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) t[i][j] = b.R[i] * dF[0][j] + b.R[3 + i] * dF[1][j] + b.R[6 + i] * dF[2][j];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      aim[i + 3 * j] = t[i][0] * b.R[j] + t[i][1] * b.R[3 + j] + t[i][2] * b.R[6 + j];
}
__global__ void bc_aim_kernel(const double* sums, BcArgs b, double* aim) {
  if (threadIdx.x != 0) return;
  bc_aim_body(sums, b, aim);
}
// one thread per RVE of a batch
__global__ void bc_aim_batch_kernel(const BcBatchEntry* __restrict__ tab, int n) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  bc_aim_body(tab[r].sums, tab[r].b, tab[r].aim);
}

void launch_bc_aim(const double* sums, const double S[81], const double R[9], double wgt, int rotate, double* aim,
                   hipStream_t st) {
  BcArgs b;
  for (int i = 0; i < 81; ++i) b.S[i] = S[i];
  for (int i = 0; i < 9; ++i) b.R[i] = R[i];
  b.wgt = wgt;
  b.rotate = rotate;
  hipLaunchKernelGGL(bc_aim_kernel, dim3(1), dim3(64), 0, st, sums, b, aim);
  SCA_HIP(hipGetLastError());
}

TableArena::~TableArena() {
  host_free(hptr);
  dev_free(dptr);
}
void TableArena::reserve(size_t bytes) {
  used = 0;
  if (bytes <= cap) return;
  bytes = std::max(bytes, 2 * cap);
  host_free(hptr);  // the batch that used them has completed
  dev_free(dptr);
  hptr = dptr = nullptr;
  cap = 0;
  hptr = static_cast<char*>(host_alloc(bytes));
  dptr = static_cast<char*>(dev_alloc(bytes));
  cap = bytes;
}

size_t batch_table_bytes_per_rve() {
  // material point (KernelArgs and P block sums, constitutive.hip) + aim +
  // Gamma + three copies + three reductions + row sums, with alignment slack
  return 1024 + sizeof(BcBatchEntry) + sizeof(GammaBatchEntry) + 3 * sizeof(BatchCopy) + 3 * sizeof(BatchReduce) +
         sizeof(RowSumItem) + 9 * 256;
}

void launch_final_reduce_batch(const BatchReduce* items, int n, TableArena& ta, hipStream_t st) {
  if (n <= 0) return;
  BatchReduce* d = nullptr;
  BatchReduce* h = ta.add<BatchReduce>(n, &d);
  int nv = 0;
  for (int i = 0; i < n; ++i) {
    h[i] = items[i];
    nv = std::max(nv, items[i].nvec);
  }
  ta.upload(h, n, st);
  hipLaunchKernelGGL(final_reduce_batch_kernel, dim3(nv, n), dim3(256), 0, st, d);
  SCA_HIP(hipGetLastError());
}

void launch_copy_batch(const BatchCopy* items, int n, TableArena& ta, hipStream_t st) {
  if (n <= 0) return;
  BatchCopy* d = nullptr;
  BatchCopy* h = ta.add<BatchCopy>(n, &d);
  int64_t mx = 0;
  for (int i = 0; i < n; ++i) {
    h[i] = items[i];
    mx = std::max(mx, items[i].words / 2 + 1);
  }
  ta.upload(h, n, st);
  const unsigned bx = (unsigned)std::min<int64_t>(std::max<int64_t>(1, (mx + 255) / 256), 1024);
  hipLaunchKernelGGL(copy_batch_kernel, dim3(bx, n), dim3(256), 0, st, d);
  SCA_HIP(hipGetLastError());
}

void launch_bc_aim_batch(const BcBatchItem* items, int n, TableArena& ta, hipStream_t st) {
  if (n <= 0) return;
  BcBatchEntry* d = nullptr;
  BcBatchEntry* h = ta.add<BcBatchEntry>(n, &d);
  for (int i = 0; i < n; ++i) {
    const BcBatchItem& it = items[i];
    h[i].sums = it.sums;
    for (int k = 0; k < 81; ++k) h[i].b.S[k] = it.S[k];
    for (int k = 0; k < 9; ++k) h[i].b.R[k] = it.R[k];
    h[i].b.wgt = it.wgt;
    h[i].b.rotate = it.rotate;
    h[i].aim = it.aim;
  }
  ta.upload(h, n, st);
  hipLaunchKernelGGL(bc_aim_batch_kernel, dim3((n + 63) / 64), dim3(64), 0, st, d, n);
  SCA_HIP(hipGetLastError());
}

template <typename T>
void launch_gamma_div_batch(const GammaBatchItem* items, int n, int64_t K, TableArena& ta, hipStream_t st) {
  if (n <= 0) return;
  GammaBatchEntry* d = nullptr;
  GammaBatchEntry* h = ta.add<GammaBatchEntry>(n, &d);
  for (int i = 0; i < n; ++i) {
    const GammaBatchItem& it = items[i];
    h[i].data = it.spec;
    h[i].g = gamma_args(it.gh, it.d_xi_tab, it.d_aim);
    h[i].div_partial = it.div_partial;
  }
  ta.upload(h, n, st);
  hipLaunchKernelGGL(gamma_div_batch_kernel<T>, dim3(gamma_blocks(K), n), dim3(256), 0, st, d, K);
  SCA_HIP(hipGetLastError());
}
template void launch_gamma_div_batch<double>(const GammaBatchItem*, int, int64_t, TableArena&, hipStream_t);
template void launch_gamma_div_batch<float>(const GammaBatchItem*, int, int64_t, TableArena&, hipStream_t);

void launch_rowsum_reduce(const RowSumItem& item, hipStream_t st) {
  hipLaunchKernelGGL(rowsum_reduce_kernel, dim3(9), dim3(256), 0, st, item);
  SCA_HIP(hipGetLastError());
}

void launch_rowsum_reduce_batch(const RowSumItem* items, int n, TableArena& ta, hipStream_t st) {
  if (n <= 0) return;
  RowSumItem* d = nullptr;
  RowSumItem* h = ta.add<RowSumItem>(n, &d);
  for (int i = 0; i < n; ++i) h[i] = items[i];
  ta.upload(h, n, st);
  hipLaunchKernelGGL(rowsum_reduce_batch_kernel, dim3(9, (unsigned)n), dim3(256), 0, st, d);
  SCA_HIP(hipGetLastError());
}

void launch_final_reduce(const double* partial, int nvec, int nb, double* out, hipStream_t st) {
  hipLaunchKernelGGL(final_reduce_kernel, dim3(nvec), dim3(256), 0, st, partial, nb, out);
  SCA_HIP(hipGetLastError());
}

void launch_fwd_field(const double* last, const double* rate, double dt, double* out, int64_t n, hipStream_t st) {
  hipLaunchKernelGGL(fwd_field_kernel, dim3(nblk(n)), dim3(256), 0, st, last, rate, dt, out, n);
  SCA_HIP(hipGetLastError());
}
void launch_sub_plane_const(double* f, const double* diff, int64_t N, hipStream_t st) {
  hipLaunchKernelGGL(sub_plane_const_kernel, dim3(nblk(9 * N)), dim3(256), 0, st, f, diff, N);
  SCA_HIP(hipGetLastError());
}
void launch_fwd_diff(const double* sums, double wgt, const double* aim, double* diff, hipStream_t st) {
  hipLaunchKernelGGL(fwd_diff_kernel, dim3(1), dim3(64), 0, st, sums, wgt, aim, diff);
  SCA_HIP(hipGetLastError());
}
void launch_rate(int het, const double* f0, const double* f, double dt, const double* av, double* out, int64_t N,
                 hipStream_t st) {
  hipLaunchKernelGGL(rate_kernel, dim3(nblk(9 * N)), dim3(256), 0, st, het, f0, f, dt, av, out, N);
  SCA_HIP(hipGetLastError());
}
void launch_axpby(double a, const double* x, double b, const double* y0, double* y, int64_t n, hipStream_t st) {
  hipLaunchKernelGGL(axpby_kernel<double>, dim3(nblk(n)), dim3(256), 0, st, a, x, b, y0, y, n);
  SCA_HIP(hipGetLastError());
}
void launch_axpby(double a, const double* x, double b, const float* y0, double* y, int64_t n, hipStream_t st) {
  hipLaunchKernelGGL(axpby_kernel<float>, dim3(nblk(n)), dim3(256), 0, st, a, x, b, y0, y, n);
  SCA_HIP(hipGetLastError());
}
void launch_lincomb(const double* XM, double c0, int l, const double* beta, const double* const* X, double* XA,
                    int64_t n, hipStream_t st) {
  if (l > MAX_RED_VEC) throw std::runtime_error("history too long");
  CombArgs a{};
  a.l = l;
  a.c0 = c0;
  for (int i = 0; i < l; ++i) {
    a.beta[i] = beta[i];
    a.X[i] = X[i];
  }
  hipLaunchKernelGGL(lincomb_kernel, dim3(nblk(n)), dim3(256), 0, st, XM, a, XA, n);
  SCA_HIP(hipGetLastError());
}
bool launch_lincomb_norms(const double* XM, double c0, int l, const double* beta, const double* const* X,
                          double* XA, int64_t n, double* partial, hipStream_t st) {
  if (l + 1 > SHARED_GROUP) return false;
  CombNormArgs a{};
  a.n = n;
  a.l = l;
  a.c0 = c0;
  a.XM = XM;
  a.XA = XA;
  for (int i = 0; i < SHARED_GROUP; ++i) {
    a.beta[i] = i < l ? beta[i] : 0.0;
    a.X[i] = i < l ? X[i] : XM;  // unused slots stay dereferenceable
  }
  switch (l) {
#define SCA_LC(Lv)                                                                                   \
  case Lv: hipLaunchKernelGGL(lincomb_norms_kernel<Lv>, dim3(RED_BLOCKS), dim3(256), 0, st, a, partial); break;
    SCA_LC(1) SCA_LC(2) SCA_LC(3) SCA_LC(4) SCA_LC(5) SCA_LC(6) SCA_LC(7) SCA_LC(8)
    SCA_LC(9) SCA_LC(10) SCA_LC(11) SCA_LC(12) SCA_LC(13) SCA_LC(14) SCA_LC(15)
#undef SCA_LC
    default: return false;
  }
  SCA_HIP(hipGetLastError());
  return true;
}
void launch_soa_to_aos(const double* soa, double* aos, int64_t N, int ncomp, hipStream_t st) {
  hipLaunchKernelGGL(soa_to_aos_kernel, dim3(nblk(ncomp * N)), dim3(256), 0, st, soa, aos, N, ncomp);
  SCA_HIP(hipGetLastError());
}
void launch_aos_to_soa(const double* aos, double* soa, int64_t N, int ncomp, hipStream_t st) {
  hipLaunchKernelGGL(aos_to_soa_kernel, dim3(nblk(ncomp * N)), dim3(256), 0, st, aos, soa, N, ncomp);
  SCA_HIP(hipGetLastError());
}
void launch_f64_to_f32(const double* a, float* b, int64_t n, hipStream_t st) {
  hipLaunchKernelGGL(f64_to_f32_kernel, dim3(nblk(n)), dim3(256), 0, st, a, b, n);
  SCA_HIP(hipGetLastError());
}
void launch_f32_to_f64(const float* a, double* b, int64_t n, hipStream_t st) {
  hipLaunchKernelGGL(f32_to_f64_kernel, dim3(nblk(n)), dim3(256), 0, st, a, b, n);
  SCA_HIP(hipGetLastError());
}

}  // namespace sca
