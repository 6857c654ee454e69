// pme_mfma.hip — the y and x transforms of a PME plane as DFT matrix
// products on the matrix cores (gfx950 MFMA), in split BF16x3.
//
// For grid edges K1, K2 in {32, 64} (planes in LDS): per kz plane of the
// half spectrum,
//   y pass    Y = X W2        (lines x, transform along y)
//   x pass    Z = W1 Y        (computed as Z^T = Y^T W1, W symmetric)
//   influence Z *= BC, energy and virial as fft_plane_conv
//   inverse x and y passes with conj(W)
// W_K[k][n] = exp(-2 pi i k n / K).  Every f32 operand is split into three
// bf16 pieces (v = h + m + l, exact to 2^-24 relative) and a real product is
// the six MFMA products h h + h m + m h + h l + l h + m m, accumulated in f32
// by the MFMA itself: f32-level products on the bf16 matrix units
// (mfma_f32_32x32x16_bf16: a 32 x 32 tile, K = 16 per instruction).
//
// Lane layout of mfma_f32_32x32x16_bf16 (checked on gfx950): A fragment of
// lane l, element j = A[l % 32][8 (l / 32) + j]; B fragment = B[8 (l / 32) +
// j][l % 32]; accumulator i = D[8 (i / 4) + 4 (l / 32) + i % 4][l % 32].
// Included by engine.hip after pme.hip (PmeParams, ConvTables, cpx).

namespace mdhip {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));

// The DFT matrices in bf16 pieces, row n of W_K (k contiguous; W is
// symmetric, so a row is a column): index ((piece * 2 + part) * K + n) * K + k,
// piece h, m, l, part re, im.
struct DftTables {
  const __bf16 *w1;   // K1
  const __bf16 *w2;   // K2
};

__device__ __forceinline__ void split3(float v, __bf16 &h, __bf16 &m, __bf16 &l) {
  h = (__bf16)v;
  float r = v - (float)h;
  m = (__bf16)r;
  r -= (float)m;
  l = (__bf16)r;
}

// The plane's f32 values (re/im, rows of stride ld) as three bf16 pieces
// each, rows of stride ldp: pieces[(piece * 2 + part) * rows * ldp + ...],
// once per pass (every task of the pass reads them as A fragments).
__device__ __forceinline__ void split_plane(const float *re, const float *im, int ld, int rows,
                                            int cols, __bf16 *pieces, int ldp) {
  const size_t ps = (size_t)rows * ldp;
  for (int w = threadIdx.x; w < rows * cols; w += blockDim.x) {
    const int r = w / cols, c = w % cols;
    __bf16 h, m, l;
    split3(re[r * ld + c], h, m, l);
    pieces[0 * ps + r * ldp + c] = h;
    pieces[2 * ps + r * ldp + c] = m;
    pieces[4 * ps + r * ldp + c] = l;
    split3(im[r * ld + c], h, m, l);
    pieces[1 * ps + r * ldp + c] = h;
    pieces[3 * ps + r * ldp + c] = m;
    pieces[5 * ps + r * ldp + c] = l;
  }
}

// 8 bf16 (16 bytes) of row `row`, columns k.., of a piece plane.
__device__ __forceinline__ bf16x8 load8(const __bf16 *p, int ldp, int row, int k) {
  return *reinterpret_cast<const bf16x8 *>(p + (size_t)row * ldp + k);
}

__device__ __forceinline__ bf16x8 negate8(bf16x8 v) {
  typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
  u32x4 u = __builtin_bit_cast(u32x4, v);
  u ^= 0x80008000u;
  return __builtin_bit_cast(bf16x8, u);
}

__device__ __forceinline__ bf16x8 b_fragment(const __bf16 *tab, int K, int piece, int part,
                                             int tn, int k0, int lane) {
  const __bf16 *src =
      tab + ((size_t)(piece * 2 + part) * K + tn + (lane & 31)) * K + k0 + 8 * (lane >> 5);
  return *reinterpret_cast<const bf16x8 *>(src);
}

__device__ __forceinline__ floatx16 mfma6(const bf16x8 &ah, const bf16x8 &am, const bf16x8 &al,
                                          const bf16x8 &bh, const bf16x8 &bm, const bf16x8 &bl,
                                          floatx16 acc) {
  acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, bh, acc, 0, 0, 0);
  acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bl, acc, 0, 0, 0);
  acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(am, bm, acc, 0, 0, 0);
  acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(am, bh, acc, 0, 0, 0);
  acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bm, acc, 0, 0, 0);
  acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bh, acc, 0, 0, 0);
  return acc;
}

// One pass: out[m][n] = sum_k in[m][k] W^(s)[k][n] for m < M, n < K, with
// W^(+1) = W and W^(-1) = conj(W); in: the pieces of split_plane (rows m,
// stride ldp); out: re/im f32 planes, row m of stride ldo, or transposed
// (row n, column m) when `transposed`.  The M x K outputs are 32 x 32 tiles,
// real and imaginary parts separately: one task per wave.
template <int K>
__device__ void dft_pass(const __bf16 *pieces, int ldp, float *out_re, float *out_im, int ldo,
                         int M, bool transposed, const __bf16 *tab, float sign) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwaves = blockDim.x >> 6;
  const int tiles_n = K / 32, ntasks = (M / 32) * tiles_n * 2;
  const size_t ps = (size_t)M * ldp;
  for (int task = wave; task < ntasks; task += nwaves) {
    const int part = task & 1, tile = task >> 1;
    const int tm = (tile / tiles_n) * 32, tn = (tile % tiles_n) * 32;
    const int row = tm + (lane & 31);
    floatx16 acc = {};
#pragma unroll
    for (int k0 = 0; k0 < K; k0 += 16) {
      // real:  Ar Br - s Ai Bi;   imaginary:  s Ar Bi + Ai Br
      const int k = k0 + 8 * (lane >> 5);
      bf16x8 ah = load8(pieces + 0 * ps, ldp, row, k);
      bf16x8 am = load8(pieces + 2 * ps, ldp, row, k);
      bf16x8 al = load8(pieces + 4 * ps, ldp, row, k);
      if (part == 1 && sign < 0.0f) {
        ah = negate8(ah);
        am = negate8(am);
        al = negate8(al);
      }
      acc = mfma6(ah, am, al, b_fragment(tab, K, 0, part, tn, k0, lane),
                  b_fragment(tab, K, 1, part, tn, k0, lane),
                  b_fragment(tab, K, 2, part, tn, k0, lane), acc);
      ah = load8(pieces + 1 * ps, ldp, row, k);
      am = load8(pieces + 3 * ps, ldp, row, k);
      al = load8(pieces + 5 * ps, ldp, row, k);
      if (part == 0 && sign > 0.0f) {
        ah = negate8(ah);
        am = negate8(am);
        al = negate8(al);
      }
      acc = mfma6(ah, am, al, b_fragment(tab, K, 0, 1 - part, tn, k0, lane),
                  b_fragment(tab, K, 1, 1 - part, tn, k0, lane),
                  b_fragment(tab, K, 2, 1 - part, tn, k0, lane), acc);
    }
    float *out = part == 0 ? out_re : out_im;
    const int col = tn + (lane & 31);
#pragma unroll
    for (int g = 0; g < 4; ++g) {
      const int r0 = tm + 8 * g + 4 * (lane >> 5);
      if (transposed) {
        *reinterpret_cast<float4 *>(out + (size_t)col * ldo + r0) =
            make_float4(acc[4 * g], acc[4 * g + 1], acc[4 * g + 2], acc[4 * g + 3]);
      } else {
#pragma unroll
        for (int i = 0; i < 4; ++i) out[(size_t)(r0 + i) * ldo + col] = acc[4 * g + i];
      }
    }
  }
}

// A kz plane: y and x DFTs, the influence product (energy, virial), the
// inverse x and y DFTs, all in LDS: the DFT tables (12 K^2 bytes each), one
// f32 plane pair (rows padded by 4 floats) that every pass writes, and the
// bf16 pieces of each pass' input (rows padded by 8).
template <int K1, int K2>
__global__ void __launch_bounds__(512) fft_plane_conv_mfma(
    cpx<float> *__restrict__ data, DftTables dt, ConvTables ct,
    double *__restrict__ energy_partial, double *__restrict__ virial_partial, size_t fs) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  __shared__ double s_e[512];
  __shared__ float s_t0[K1], s_m0[K1], s_t1[K2], s_m1[K2];
  data = frame_at(data, fs);
  energy_partial = frame_at(energy_partial, fs);
  virial_partial = frame_at(virial_partial, fs);
  constexpr int KM = K1 > K2 ? K1 : K2;
  constexpr int P = KM + 4, Q = KM + 8;        // f32 and bf16 row strides
  constexpr int NT1 = 6 * K1 * K1, NT2 = K1 == K2 ? 0 : 6 * K2 * K2;
  __bf16 *w1 = reinterpret_cast<__bf16 *>(smem);
  __bf16 *w2 = K1 == K2 ? w1 : w1 + NT1;
  float *f_re = reinterpret_cast<float *>(w1 + NT1 + NT2);
  float *f_im = f_re + KM * P;
  __bf16 *pieces = reinterpret_cast<__bf16 *>(f_im + KM * P);
  const int kz = blockIdx.x;
  cpx<float> *plane = data + (size_t)kz * K1 * K2;
  for (int w = threadIdx.x; w < NT1 / 8; w += blockDim.x)
    reinterpret_cast<bf16x8 *>(w1)[w] = reinterpret_cast<const bf16x8 *>(dt.w1)[w];
  if (K1 != K2)
    for (int w = threadIdx.x; w < NT2 / 8; w += blockDim.x)
      reinterpret_cast<bf16x8 *>(w2)[w] = reinterpret_cast<const bf16x8 *>(dt.w2)[w];
  // the plane X[x][y], rows of P
  for (int w = threadIdx.x; w < K1 * K2; w += blockDim.x) {
    const cpx<float> v = plane[w];
    const int x = w / K2, y = w % K2;
    f_re[x * P + y] = v.x;
    f_im[x * P + y] = v.y;
  }
  for (int w = threadIdx.x; w < K1; w += blockDim.x) {
    s_t0[w] = static_cast<const float *>(ct.t[0])[w];
    s_m0[w] = static_cast<const float *>(ct.m[0])[w];
  }
  for (int w = threadIdx.x; w < K2; w += blockDim.x) {
    s_t1[w] = static_cast<const float *>(ct.t[1])[w];
    s_m1[w] = static_cast<const float *>(ct.m[1])[w];
  }
  __syncthreads();
  // y: X[x][y] W2 -> Yt[ky][x]
  split_plane(f_re, f_im, P, K1, K2, pieces, Q);
  __syncthreads();
  dft_pass<K2>(pieces, Q, f_re, f_im, P, K1, true, w2, 1.0f);
  __syncthreads();
  // x: Yt[ky][x] W1 -> Zt[ky][kx]
  split_plane(f_re, f_im, P, K2, K1, pieces, Q);
  __syncthreads();
  dft_pass<K1>(pieces, Q, f_re, f_im, P, K2, false, w1, 1.0f);
  __syncthreads();
  const float t2 = static_cast<const float *>(ct.t[2])[kz];
  const float m2 = static_cast<const float *>(ct.m[2])[kz];
  const double h = (kz == 0 || (ct.K3 % 2 == 0 && kz == ct.K3 / 2)) ? 0.5 : 1.0;
  double e = 0.0;
  double vir[6] = {0, 0, 0, 0, 0, 0};
  for (int w = threadIdx.x; w < K1 * K2; w += blockDim.x) {
    const int ky = w / K1, kx = w % K1;
    const int i = ky * P + kx;
    const float vr = f_re[i], vi = f_im[i];
    float bc = 0.0f;
    if ((kx | ky | kz) != 0) {
      const float msq = s_m0[kx] * s_m0[kx] + s_m1[ky] * s_m1[ky] + m2 * m2;
      bc = s_t0[kx] * s_t1[ky] * t2 / msq;
      if (energy_partial || virial_partial) {
        const double em = h * (double)bc * ((double)vr * vr + (double)vi * vi);
        e += em;
        if (virial_partial) {
          const double c = 2.0 * (1.0 / (double)msq + (double)ct.pb2);
          const double mx = s_m0[kx], my = s_m1[ky], mz = m2;
          vir[0] += em * (1.0 - c * mx * mx);
          vir[1] += em * (1.0 - c * my * my);
          vir[2] += em * (1.0 - c * mz * mz);
          vir[3] -= em * c * mx * my;
          vir[4] -= em * c * mx * mz;
          vir[5] -= em * c * my * mz;
        }
      }
    }
    f_re[i] = vr * bc;
    f_im[i] = vi * bc;
  }
  __syncthreads();
  // inverse x: Zt[ky][kx] conj(W1) -> Y'[x][ky]
  split_plane(f_re, f_im, P, K2, K1, pieces, Q);
  __syncthreads();
  dft_pass<K1>(pieces, Q, f_re, f_im, P, K2, true, w1, -1.0f);
  __syncthreads();
  // inverse y: Y'[x][ky] conj(W2) -> X'[x][y]
  split_plane(f_re, f_im, P, K1, K2, pieces, Q);
  __syncthreads();
  dft_pass<K2>(pieces, Q, f_re, f_im, P, K1, false, w2, -1.0f);
  __syncthreads();
  for (int w = threadIdx.x; w < K1 * K2; w += blockDim.x) {
    const int x = w / K2, y = w % K2;
    plane[w] = {f_re[x * P + y], f_im[x * P + y]};
  }
  if (energy_partial) {
    s_e[threadIdx.x] = e;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
      if (threadIdx.x < s) s_e[threadIdx.x] += s_e[threadIdx.x + s];
      __syncthreads();
    }
    if (threadIdx.x == 0) energy_partial[blockIdx.x] = s_e[0];
  }
  if (virial_partial) {
    for (int c = 0; c < 6; ++c) {
      __syncthreads();
      s_e[threadIdx.x] = vir[c];
      __syncthreads();
      for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) s_e[threadIdx.x] += s_e[threadIdx.x + s];
        __syncthreads();
      }
      if (threadIdx.x == 0) virial_partial[6 * blockIdx.x + c] = s_e[0];
    }
  }
}

// ---- general axis lengths, grids whose planes do not fit LDS ------------
//
// The half spectrum of these grids is [x][y][kz] (kz fastest, as the
// Stockham column passes take it).  A block takes 32 consecutive kz of one
// set of lines: all y of one x for the y pass, all x of one y for the x
// pass.  Its 32 lines, padded with zeros to Kp (K rounded up to 32), are
// the rows of an f32 plane in LDS; their bf16 pieces are the A operand, and
// the DFT table, Kp x Kp in the same piece layout with zeros beyond K, the B
// operand read from global memory (it is shared by every block).
constexpr int kMfmaAxisMax = 224;

// out[m][n] = sum_k in[m][k] W^(s)[k][n] for the 32 rows m and n < Kp: the
// same 32 x 32 tiles as dft_pass, with Kp known at run time.
__device__ void dft_rows32(const __bf16 *pieces, int ldp, float *out_re, float *out_im, int ldo,
                           int Kp, const __bf16 *tab, float sign) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwaves = blockDim.x >> 6;
  const int ntasks = (Kp / 32) * 2;
  const size_t ps = (size_t)32 * ldp;
  for (int task = wave; task < ntasks; task += nwaves) {
    const int part = task & 1, tn = (task >> 1) * 32;
    const int row = lane & 31;
    floatx16 acc = {};
    for (int k0 = 0; k0 < Kp; k0 += 16) {
      const int k = k0 + 8 * (lane >> 5);
      bf16x8 ah = load8(pieces + 0 * ps, ldp, row, k);
      bf16x8 am = load8(pieces + 2 * ps, ldp, row, k);
      bf16x8 al = load8(pieces + 4 * ps, ldp, row, k);
      if (part == 1 && sign < 0.0f) {
        ah = negate8(ah);
        am = negate8(am);
        al = negate8(al);
      }
      acc = mfma6(ah, am, al, b_fragment(tab, Kp, 0, part, tn, k0, lane),
                  b_fragment(tab, Kp, 1, part, tn, k0, lane),
                  b_fragment(tab, Kp, 2, part, tn, k0, lane), acc);
      ah = load8(pieces + 1 * ps, ldp, row, k);
      am = load8(pieces + 3 * ps, ldp, row, k);
      al = load8(pieces + 5 * ps, ldp, row, k);
      if (part == 0 && sign > 0.0f) {
        ah = negate8(ah);
        am = negate8(am);
        al = negate8(al);
      }
      acc = mfma6(ah, am, al, b_fragment(tab, Kp, 0, 1 - part, tn, k0, lane),
                  b_fragment(tab, Kp, 1, 1 - part, tn, k0, lane),
                  b_fragment(tab, Kp, 2, 1 - part, tn, k0, lane), acc);
    }
    float *out = part == 0 ? out_re : out_im;
    const int col = tn + (lane & 31);
#pragma unroll
    for (int g = 0; g < 4; ++g) {
      const int r0 = 8 * g + 4 * (lane >> 5);
#pragma unroll
      for (int i = 0; i < 4; ++i) out[(size_t)(r0 + i) * ldo + col] = acc[4 * g + i];
    }
  }
}

// LDS bytes of dft_axis_mfma for a padded length Kp: the f32 plane of 32
// rows (re, im; stride Kp + 4) and the six bf16 piece planes (stride Kp + 8).
inline size_t axis_mfma_lds(int Kp) {
  return (size_t)2 * 32 * (Kp + 4) * sizeof(float) + (size_t)6 * 32 * (Kp + 8) * 2;
}

// AXIS 1: the y transform of every line (dft sign +1 forward, -1
// backward).  AXIS 0: the x pass of the convolution: forward x transform,
// the influence product (and its energy, one partial per block), backward
// x transform.  Grid: (lines of the other axis) x ceil(H / 32), frames in z.
template <int AXIS>
__global__ void __launch_bounds__(512) dft_axis_mfma(cpx<float> *__restrict__ data, int K1,
                                                     int K2, int H, int Kp,
                                                     const __bf16 *__restrict__ tab,
                                                     ConvTables ct, float sign,
                                                     double *__restrict__ energy_partial,
                                                     size_t fs) {
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  __shared__ double s_e[512];
  data = frame_at(data, fs);
  energy_partial = frame_at(energy_partial, fs);
  const int K = AXIS == 1 ? K2 : K1;
  const int P = Kp + 4, Q = Kp + 8;
  float *f_re = reinterpret_cast<float *>(smem);
  float *f_im = f_re + 32 * P;
  __bf16 *pieces = reinterpret_cast<__bf16 *>(f_im + 32 * P);
  const int tiles = (H + 31) / 32;
  const int line = blockIdx.x / tiles, z0 = (blockIdx.x % tiles) * 32;
  const size_t stride = AXIS == 1 ? (size_t)H : (size_t)K2 * H;
  cpx<float> *base = AXIS == 1 ? data + (size_t)line * K2 * H + z0 : data + (size_t)line * H + z0;
  // row t (kz = z0 + t), column k (the transformed axis), zero beyond
  for (int w = threadIdx.x; w < 32 * Kp; w += blockDim.x) {
    const int t = w & 31, k = w >> 5;
    cpx<float> v = {0.0f, 0.0f};
    if (k < K && z0 + t < H) v = base[(size_t)k * stride + t];
    f_re[t * P + k] = v.x;
    f_im[t * P + k] = v.y;
  }
  __syncthreads();
  split_plane(f_re, f_im, P, 32, Kp, pieces, Q);
  __syncthreads();
  dft_rows32(pieces, Q, f_re, f_im, P, Kp, tab, AXIS == 1 ? sign : 1.0f);
  __syncthreads();
  double e = 0.0;
  if (AXIS == 0) {
    const float *t0 = (const float *)ct.t[0], *t1 = (const float *)ct.t[1],
                *t2 = (const float *)ct.t[2];
    const float *m0 = (const float *)ct.m[0], *m1 = (const float *)ct.m[1],
                *m2 = (const float *)ct.m[2];
    for (int w = threadIdx.x; w < 32 * Kp; w += blockDim.x) {
      const int t = w & 31, kx = w >> 5, kz = z0 + t;
      const int i = t * P + kx;
      const float vr = f_re[i], vi = f_im[i];
      float bc = 0.0f;
      if (kx < K1 && kz < H && (kx | line | kz) != 0) {
        const float msq = m0[kx] * m0[kx] + m1[line] * m1[line] + m2[kz] * m2[kz];
        bc = t0[kx] * t1[line] * t2[kz] / msq;
        if (energy_partial) {
          const double h = (kz == 0 || (ct.K3 % 2 == 0 && kz == ct.K3 / 2)) ? 0.5 : 1.0;
          e += h * (double)bc * ((double)vr * vr + (double)vi * vi);
        }
      }
      f_re[i] = vr * bc;
      f_im[i] = vi * bc;
    }
    __syncthreads();
    split_plane(f_re, f_im, P, 32, Kp, pieces, Q);
    __syncthreads();
    dft_rows32(pieces, Q, f_re, f_im, P, Kp, tab, -1.0f);
    __syncthreads();
  }
  for (int w = threadIdx.x; w < 32 * K; w += blockDim.x) {
    const int t = w & 31, k = w >> 5;
    if (z0 + t < H) base[(size_t)k * stride + t] = {f_re[t * P + k], f_im[t * P + k]};
  }
  if (AXIS == 0 && energy_partial) {
    s_e[threadIdx.x] = e;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
      if ((int)threadIdx.x < s) s_e[threadIdx.x] += s_e[threadIdx.x + s];
      __syncthreads();
    }
    if (threadIdx.x == 0) energy_partial[blockIdx.x] = s_e[0];
  }
}

// LDS bytes of fft_plane_conv_mfma<K1, K2>.
inline size_t plane_mfma_lds(int K1, int K2) {
  const int K = K1 > K2 ? K1 : K2;
  const size_t tabs = (size_t)12 * K1 * K1 + (K1 == K2 ? 0 : (size_t)12 * K2 * K2);
  return tabs + (size_t)2 * K * (K + 4) * sizeof(float) + (size_t)6 * K * (K + 8) * 2;
}

}  // namespace mdhip
