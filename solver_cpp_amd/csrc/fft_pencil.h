// Pencil FFTs with the transform length N fixed at compile time (included by
// spectral.hip after cpx<T> and its helpers).
//
// A pencil of N points is held by TP = N / E threads, E = 8 points each
// (E = 12 for N = 3 * 2^a; E = 16 for single-precision 2^a pencils of 32
// points or more).  Thread t holds the points t + TP * j (j < E).
// Every Stockham stage (radix R dividing E, span NS) is E / R butterflies per
// thread in registers; its butterfly b = t + TP * m reads the points
// b + q * N / R, which are exactly the thread's registers m + q * E / R.  A
// stage that is not the last writes its outputs to LDS at the Stockham
// positions and the next stage reads t + TP * j back; the last stage's
// outputs land on the thread's own positions, so the first stage reads and
// the last stage writes global memory directly, coalesced.  A block of BS
// threads (a power of two from TP to 512, chosen per plan: 256 by default,
// or by the plan's autotuner) holds P = BS / TP pencils; the block size
// changes no arithmetic.  The twiddle table (length N) is staged in LDS
// with the data.
//
// Supported: N = 2^a (8..512, E = 8), N = 3 * 2^a (12..384, E = 12),
// N = 9 * 2^a (18..288, E = 18: radix-6, 3 and 2 stages) and N = 27
// (E = 27, one thread per pencil); other lengths use the generic Stockham
// kernels.
#pragma once

namespace sca {
namespace pencil {

constexpr int odd_part(int n) {
  int m = n;
  while (m > 0 && m % 2 == 0) m /= 2;
  return m;
}
constexpr bool supported(int n) {
  if (n < 8) return false;
  const int m = odd_part(n);
  if (m == 1) return n <= 512;
  if (m == 3) return n >= 12 && n <= 384;
  if (m == 9) return n >= 18 && n <= 288;
  return n == 27;
}
constexpr int points_per_thread(int n) {
  const int m = odd_part(n);
  return m == 1 ? 8 : m == 3 ? 12 : m == 9 ? 18 : 27;
}
// single precision: 2^a pencils of 32 points or more hold 16 points per
// thread, so that a thread moves as many bytes as in double precision
template <typename T>
constexpr int points_per_thread_t(int n) {
  const int e = points_per_thread(n);
  return (sizeof(T) == 4 && e == 8 && n >= 32) ? 16 : e;
}
constexpr int pick_radix(int rem, int E) {
  for (int r : {12, 8, 6, 4, 3, 2})
    if (E % r == 0 && rem % r == 0) return r;
  return 0;
}

// cos and sin of 2 pi k / 24 (k = 0..23): every in-register twiddle of the
// radix-3, 4, 6, 8 and 12 butterflies is one of them
template <typename T>
__device__ __forceinline__ cpx<T> w24(int k) {
  constexpr double c[6] = {1.0, 0.96592582628906828675, 0.86602540378443864676, 0.70710678118654752440,
                           0.5, 0.25881904510252076235};
  // k in [0, 24): quadrant reduction
  const int q = k / 6, r = k % 6;
  const double cr = c[r], sr = (r == 0) ? 0.0 : c[6 - r];
  double x, y;
  switch (q) {
    case 0: x = cr; y = sr; break;
    case 1: x = -sr; y = cr; break;
    case 2: x = -cr; y = -sr; break;
    default: x = sr; y = -cr; break;
  }
  return {(T)x, (T)y};
}
// z * exp(SIGN 2 pi i e / M) for an exponent that is constant once the
// caller's loops are unrolled; quarter turns are exact swaps
template <typename T, int SIGN, int M>
__device__ __forceinline__ cpx<T> wmul_rt(cpx<T> z, int e) {
  const int k = ((24 / M) * e) % 24;
  if (k == 0) return z;
  if (k == 12) return {-z.x, -z.y};
  if (k == 6) return {(T)SIGN * -z.y, (T)SIGN * z.x};   // (SIGN i) z
  if (k == 18) return {(T)SIGN * z.y, (T)SIGN * -z.x};  // (-SIGN i) z
  const cpx<T> w0 = w24<T>(k);
  const cpx<T> w = {w0.x, (T)SIGN * w0.y};
  return cmul(z, w);
}

// in-register DFT of R points: X[k] = sum_q v[q] exp(SIGN 2 pi i q k / R)
template <typename T, int R, int SIGN>
__device__ __forceinline__ void dft(cpx<T>* v);

template <typename T, int SIGN>
__device__ __forceinline__ void dft2(cpx<T>& a, cpx<T>& b) {
  const cpx<T> s = cadd(a, b), d = csub(a, b);
  a = s;
  b = d;
}
template <typename T, int R, int SIGN>
__device__ __forceinline__ void dft(cpx<T>* v) {
  if constexpr (R == 2) {
    dft2<T, SIGN>(v[0], v[1]);
  } else if constexpr (R == 3) {
    const T h = (T)0.86602540378443864676;
    const cpx<T> a = v[0], b = v[1], c = v[2];
    const cpx<T> t1 = cadd(b, c), t2 = csub(b, c);
    const cpx<T> m = {a.x - (T)0.5 * t1.x, a.y - (T)0.5 * t1.y};
    const cpx<T> it2 = {(T)SIGN * -h * t2.y, (T)SIGN * h * t2.x};
    v[0] = cadd(a, t1);
    v[1] = cadd(m, it2);
    v[2] = csub(m, it2);
  } else if constexpr (R == 4) {
    const cpx<T> a = v[0], b = v[1], c = v[2], d = v[3];
    const cpx<T> s0 = cadd(a, c), s1 = csub(a, c), s2 = cadd(b, d), s3 = csub(b, d);
    const cpx<T> ws3 = {(T)SIGN * -s3.y, (T)SIGN * s3.x};  // (SIGN i) s3
    v[0] = cadd(s0, s2);
    v[2] = csub(s0, s2);
    v[1] = cadd(s1, ws3);
    v[3] = csub(s1, ws3);
  } else if constexpr (R == 6 || R == 8 || R == 12) {
    // decimation in time: R = R1 * R2 with inner DFTs of R2 points over the
    // R1 interleaved subsequences, then twiddles and DFTs of R1 points
    constexpr int R1 = (R == 12) ? 4 : 2;
    constexpr int R2 = R / R1;
    cpx<T> y[R1][R2];
#pragma unroll
    for (int n1 = 0; n1 < R1; ++n1) {
#pragma unroll
      for (int n2 = 0; n2 < R2; ++n2) y[n1][n2] = v[n1 + R1 * n2];
      dft<T, R2, SIGN>(y[n1]);
    }
    // X[k1 + R2 k2] = sum_n1 W_R1^(n1 k2) (W_R^(n1 k1) Y_n1[k1])
#pragma unroll
    for (int k1 = 0; k1 < R2; ++k1) {
      cpx<T> z[R1];
#pragma unroll
      for (int n1 = 0; n1 < R1; ++n1) z[n1] = y[n1][k1];
      if constexpr (R1 >= 2) z[1] = wmul_rt<T, SIGN, R>(z[1], 1 * k1);
      if constexpr (R1 == 4) {
        z[2] = wmul_rt<T, SIGN, R>(z[2], 2 * k1);
        z[3] = wmul_rt<T, SIGN, R>(z[3], 3 * k1);
      }
      dft<T, R1, SIGN>(z);
#pragma unroll
      for (int k2 = 0; k2 < R1; ++k2) v[k1 + R2 * k2] = z[k2];
    }
  }
}

// Stockham stages from span NS on: v holds the points t + TP * j of the
// thread's pencil; lds is the pencil's row (N points) and tw the staged
// forward table tw[i] = exp(-2 pi i i / N).
template <typename T, int N, int E, int SIGN, int NS>
__device__ __forceinline__ void stages(cpx<T>* v, int t, cpx<T>* lds, const cpx<T>* tw) {
  if constexpr (NS < N) {
    constexpr int R = pick_radix(N / NS, E);
    static_assert(R > 0, "no radix for this pencil length");
    constexpr int TP = N / E, NB = E / R;
    constexpr bool LAST = NS * R == N;
#pragma unroll
    for (int m = 0; m < NB; ++m) {
      cpx<T> x[R];
#pragma unroll
      for (int q = 0; q < R; ++q) x[q] = v[m + q * NB];
      const int b = t + TP * m;
      const int k = b % NS;
      if constexpr (NS > 1) {
#pragma unroll
        for (int q = 1; q < R; ++q) {
          cpx<T> w = tw[q * k * (N / (NS * R))];
          if constexpr (SIGN > 0) w.y = -w.y;
          x[q] = cmul(x[q], w);
        }
      }
      dft<T, R, SIGN>(x);
      if constexpr (LAST) {
#pragma unroll
        for (int q = 0; q < R; ++q) v[m + q * NB] = x[q];
      } else {
        const int base = (b / NS) * NS * R + k;
#pragma unroll
        for (int q = 0; q < R; ++q) lds[base + q * NS] = x[q];
      }
    }
    if constexpr (!LAST) {
      __syncthreads();
#pragma unroll
      for (int j = 0; j < E; ++j) v[j] = lds[t + TP * j];
      __syncthreads();  // all reads done before the next stage writes
      stages<T, N, E, SIGN, NS * R>(v, t, lds, tw);
    }
  }
}

template <typename T, int N>
struct Shape {
  static constexpr int E = points_per_thread_t<T>(N);
  static constexpr int TP = N / E;     // threads per pencil
  static constexpr int LD = N + 1;     // padded LDS row
  // complex LDS entries of a block of bs threads: rows + twiddles
  static constexpr int lds(int bs) { return bs / TP * LD + N; }
};
constexpr int MAX_PENCIL_BLOCK = 512;

template <typename T, int N>
__device__ __forceinline__ const cpx<T>* stage_tw(cpx<T>* lds_tw, const cpx<T>* tw) {
  for (int f = threadIdx.x; f < N; f += blockDim.x) lds_tw[f] = tw[f];
  return lds_tw;  // visible after the first stage's barrier
}

// strided complex pass (y or z): pencil (o, i) starts at
// (o / d1) * s1 + (o % d1) * s2 + i with element stride es; a block takes P
// consecutive i of one o, the pencil index fastest across lanes
template <typename T, int N, int SIGN>
__global__ void __launch_bounds__(MAX_PENCIL_BLOCK) strided_kernel(cpx<T>* data, int ni, int64_t npen, int d1,
                                                                   int64_t s1, int64_t s2, int64_t es,
                                                                   const cpx<T>* tw_g) {
  using S = Shape<T, N>;
  extern __shared__ unsigned char smem_raw[];
  cpx<T>* lds = reinterpret_cast<cpx<T>*>(smem_raw);
  const int P = blockDim.x / S::TP;  // a power of two
  const cpx<T>* tw = stage_tw<T, N>(lds + P * S::LD, tw_g);
  const int p = threadIdx.x & (P - 1), t = threadIdx.x >> __builtin_ctz(P);
  // pencils (o, i) numbered with i fastest across the whole launch, so that a
  // block does not hold a mostly empty row of i (ni = N/2 + 1 is odd)
  const int64_t pen = (int64_t)blockIdx.x * P + p;
  const int64_t o = pen / ni;
  const int i = (int)(pen - o * ni);
  const int64_t base = (o / d1) * s1 + (o % d1) * s2 + i;
  const bool ok = pen < npen;
  cpx<T> v[S::E];
#pragma unroll
  for (int j = 0; j < S::E; ++j) v[j] = ok ? data[base + (int64_t)(t + S::TP * j) * es] : cpx<T>{(T)0, (T)0};
  stages<T, N, S::E, SIGN, 1>(v, t, lds + p * S::LD, tw);
  if (ok) {
#pragma unroll
    for (int j = 0; j < S::E; ++j) data[base + (int64_t)(t + S::TP * j) * es] = v[j];
  }
}

// Rows of the pencil that pairs rows ra and rb of a set of set_rows rows
// (sets of rows are transformed independently: a set's last row of an odd
// count is paired with zeros, never with the next set's first row).  An
// absent row gets the index nrows.
__device__ __forceinline__ void pencil_rows(int64_t pencil, int64_t set_rows, int64_t nrows, int64_t& ra,
                                            int64_t& rb) {
  const int64_t pps = (set_rows + 1) / 2;  // pencils per set
  const int64_t set = pencil / pps, q = pencil - set * pps;
  ra = set * set_rows + 2 * q;
  rb = 2 * q + 1 < set_rows ? ra + 1 : nrows;
  if (ra >= nrows) ra = rb = nrows;
}

// x + (x of the lane DPP control CTRL selects), the halves moved separately
template <int CTRL>
__device__ __forceinline__ double dpp_add(double x) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(x), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(x), CTRL, 0xF, 0xF, false);
  return x + __hiloint2double(hi, lo);
}
// sum over each group of G consecutive lanes (G a power of two, groups
// aligned), every lane of a group ending with the same bits: DPP exchanges
// within rows of 16 lanes (quad xor 1, quad xor 2, half-row and row mirror),
// xor shuffles across rows.  All lanes of the wave take part.
template <int G>
__device__ __forceinline__ double group_sum(double x) {
  if constexpr (G >= 2) x = dpp_add<0xB1>(x);   // quad_perm [1,0,3,2]
  if constexpr (G >= 4) x = dpp_add<0x4E>(x);   // quad_perm [2,3,0,1]
  if constexpr (G >= 8) x = dpp_add<0x141>(x);  // row_half_mirror
  if constexpr (G >= 16) x = dpp_add<0x140>(x); // row_mirror
#pragma unroll
  for (int off = 16; off < G; off <<= 1) x += __shfl_xor(x, off, G);
  return x;
}

// x pass, real -> half complex, two real rows per complex pencil
// (z = x_a + i x_b): X_a[k] = (Z[k] + conj Z[N-k]) / 2,
// X_b[k] = (Z[k] - conj Z[N-k]) / (2 i).  rowsum (optional): the sum of each
// input row in double precision, in a fixed order (each thread's points in
// turn, then the pencil's TP lanes, group_sum): the same for every block
// shape, so that a field's sums can be taken from its forward transform
// instead of another pass over it.
template <typename T, typename TI, int N>
__global__ void __launch_bounds__(MAX_PENCIL_BLOCK) x_r2c_kernel(const TI* in, cpx<T>* out, int64_t nrows,
                                                                 int64_t set_rows, const cpx<T>* tw_g,
                                                                 double* rowsum) {
  using S = Shape<T, N>;
  constexpr int NR = N / 2 + 1;
  extern __shared__ unsigned char smem_raw[];
  cpx<T>* lds = reinterpret_cast<cpx<T>*>(smem_raw);
  const int P = blockDim.x / S::TP;
  const cpx<T>* tw = stage_tw<T, N>(lds + P * S::LD, tw_g);
  const int t = threadIdx.x % S::TP, p = threadIdx.x / S::TP;
  int64_t ra, rb;
  pencil_rows((int64_t)blockIdx.x * P + p, set_rows, nrows, ra, rb);
  cpx<T>* row = lds + p * S::LD;
  cpx<T> v[S::E];
  double sa = 0.0, sb = 0.0;
#pragma unroll
  for (int j = 0; j < S::E; ++j) {
    const int e = t + S::TP * j;
    const TI xa = ra < nrows ? in[ra * N + e] : (TI)0, xb = rb < nrows ? in[rb * N + e] : (TI)0;
    v[j] = {(T)xa, (T)xb};
    sa += (double)xa;
    sb += (double)xb;
  }
  if (rowsum) {
    sa = group_sum<S::TP>(sa);
    sb = group_sum<S::TP>(sb);
    if (t == 0) {
      if (ra < nrows) rowsum[ra] = sa;
      if (rb < nrows) rowsum[rb] = sb;
    }
  }
  stages<T, N, S::E, -1, 1>(v, t, row, tw);
  __syncthreads();  // the stages' last reads of the row are done
#pragma unroll
  for (int j = 0; j < S::E; ++j) row[t + S::TP * j] = v[j];
  __syncthreads();
  for (int k = t; k < NR; k += S::TP) {
    const cpx<T> z = row[k], zn = row[(N - k) % N];
    const cpx<T> xa = {(T)0.5 * (z.x + zn.x), (T)0.5 * (z.y - zn.y)};
    const cpx<T> xb = {(T)0.5 * (z.y + zn.y), (T)-0.5 * (z.x - zn.x)};
    if (ra < nrows) out[ra * NR + k] = xa;
    if (rb < nrows) out[rb * NR + k] = xb;
  }
}

// x pass, half complex -> real (Hermitian completion; the imaginary parts of
// the real modes k = 0 and k = N/2 are dropped), scaled; two rows per pencil
template <typename T, typename TO, int N>
__global__ void __launch_bounds__(MAX_PENCIL_BLOCK) x_c2r_kernel(const cpx<T>* in, TO* out, int64_t nrows,
                                                                 int64_t set_rows, const cpx<T>* tw_g, T scale) {
  using S = Shape<T, N>;
  constexpr int NR = N / 2 + 1;
  extern __shared__ unsigned char smem_raw[];
  cpx<T>* lds = reinterpret_cast<cpx<T>*>(smem_raw);
  const int P = blockDim.x / S::TP;
  const cpx<T>* tw = stage_tw<T, N>(lds + P * S::LD, tw_g);
  const int t = threadIdx.x % S::TP, p = threadIdx.x / S::TP;
  int64_t ra, rb;
  pencil_rows((int64_t)blockIdx.x * P + p, set_rows, nrows, ra, rb);
  cpx<T> v[S::E];
#pragma unroll
  for (int j = 0; j < S::E; ++j) {
    const int e = t + S::TP * j;
    const bool upper = e > N / 2;
    const int ke = upper ? N - e : e;
    cpx<T> a = ra < nrows ? in[ra * NR + ke] : cpx<T>{(T)0, (T)0};
    cpx<T> b = rb < nrows ? in[rb * NR + ke] : cpx<T>{(T)0, (T)0};
    if (e == 0 || 2 * e == N) {
      a.y = (T)0;
      b.y = (T)0;
    }
    if (upper) {
      a.y = -a.y;
      b.y = -b.y;
    }
    v[j] = {a.x - b.y, a.y + b.x};  // a + i b
  }
  stages<T, N, S::E, 1, 1>(v, t, lds + p * S::LD, tw);
#pragma unroll
  for (int j = 0; j < S::E; ++j) {
    const int e = t + S::TP * j;
    if (ra < nrows) out[ra * N + e] = (TO)(v[j].x * scale);
    if (rb < nrows) out[rb * N + e] = (TO)(v[j].y * scale);
  }
}

}  // namespace pencil
}  // namespace sca
