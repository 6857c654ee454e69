// pme.hip — the smooth PME reciprocal chain for gfx950.
//
// Reference semantics (docs/paper/05-electrostatics.md:95-279):
//   places    u = K frac(x/L) in f64, first point (floor(u) - n + 1) mod K
//             (PMEWeights.cpp:8)
//   spread    Q(l) = sum_j q_j prod_a M_n(u_aj - l_a)  (PMESpread.cpp; the
//             deterministic mode adds q prod M scaled by 2^40 with 64-bit
//             integer atomics, docs/pme.md:310-331) — done that way here
//   R2C FFT   unnormalised, half spectrum along z
//   product   Q^ <- BC Q^, BC = t1 t2 t3 / m^2, C(0) = 0; energy
//             sum h BC |Q^|^2 with h = 1/2 on the k3 = 0 and Nyquist planes
//             (PMETables/PMEConvolve.cpp; the influence fused into the
//             transform as pme_callback.cu:5-21 fuses it into cuFFT's load)
//   C2R FFT   unnormalised: phi
//   gather    F_ja = -q_j (K_a/L_a) sum_l phi(l) M'(u_a - l_a) prod M
//             (PMEGather.cpp)
// The FFT is the LDS Stockham of fft.h; the forward z pass reads the 64-bit
// charge words, converts them and clears them for the next step, and the
// x pass does forward transform, influence product and backward transform
// of each x line in one LDS residency.
#include "fft.h"
#include "md_common.h"

namespace mdhip {

struct PmeParams {
  int K[3];
  int order;
  double box[3];
  double inv_box[3];
};

// B-spline weights M[t] of the points first + t, t = 0..n-1, and their
// derivatives in u, for fraction w (recursion of PMECode.h:147-186; closed
// forms of 05-electrostatics.md:290 at order 4).
template <typename T>
__device__ __forceinline__ void bspline(T w, int order, T *M, T *dM) {
  if (order == 4) {
    const T w2 = w * w, w3 = w2 * w, om = (T)1 - w;
    const T s6 = (T)1 / (T)6;
    M[0] = om * om * om * s6;
    M[1] = ((T)3 * w3 - (T)6 * w2 + (T)4) * s6;
    M[2] = ((T)-3 * w3 + (T)3 * w2 + (T)3 * w + (T)1) * s6;
    M[3] = w3 * s6;
    if (dM) {
      dM[0] = -(T)0.5 * om * om;
      dM[1] = (T)1.5 * w2 - (T)2 * w;
      dM[2] = (T)-1.5 * w2 + w + (T)0.5;
      dM[3] = (T)0.5 * w2;
    }
    return;
  }
  T A[9] = {0}, B[9] = {0};
  A[0] = w;
  A[1] = (T)1 - w;
  for (int nn = 3; nn <= order; ++nn) {
    for (int k = 0; k <= nn; ++k) B[k] = A[k];
    for (int k = 0; k < nn; ++k) {
      const T prev = k > 0 ? B[k - 1] : (T)0;
      A[k] = ((w + (T)k) * B[k] + ((T)nn - w - (T)k) * prev) / (T)(nn - 1);
    }
  }
  for (int k = 0; k < order; ++k) {
    M[order - 1 - k] = A[k];
    if (dM) dM[order - 1 - k] = B[k] - (k > 0 ? B[k - 1] : (T)0);
  }
}

__device__ __forceinline__ void pme_place(double x, double L, double invL, int K,
                                          int order, int &first, float &w) {
  const double s = x * invL;
  const double frac = s - floor(s);
  const double u = (double)K * frac;
  const double fl = floor(u);
  w = (float)(u - fl);
  int f = (int)fl - order + 1;
  f %= K;
  if (f < 0) f += K;
  first = f;
}

// ---- spread: LDS tiles of the xy columns of the cluster sort -----------
//
// The atoms of one xy column of the build lie in a contiguous range of
// places.  Their B-spline stencils cover grid points x in [gx0, gx0 + TWx),
// y in [gy0, gy0 + TWy) of that column (all z): the column's extent, the
// distance an atom may move before the validity test rebuilds the list,
// and the order.  A block accumulates one column's tile in LDS — each
// thread owns tile points (x, y) and adds the atoms in place order, so the
// f32 sums have a fixed order and need no atomics — and stores the tile.
// The first FFT pass sums the tiles that cover each grid row, in a fixed
// order (fft_rows_r2c_tiles).

struct TileParams {
  int K0;                // grid points along x (rows are K0 x K1)
  int TWx, TWy;          // tile extent in grid points (<= K)
  int ncx;               // columns along x
  int zc;                // z points per block (a chunk of K3)
  int nzc;               // chunks
  const int *gx0;        // first tile point of each column along x (mod K)
  const int *gy0;
  const int *xcov;       // per grid x: up to kMaxCover (column x, offset), -1 ends
  const int *ycov;
};
constexpr int kMaxCover = 8;
constexpr int kTileAtoms = 256;

template <int N>
__global__ void __launch_bounds__(256) pme_spread_tiles(
    const int *__restrict__ col_cl_start, const int *__restrict__ col_cl_count,
    const int *__restrict__ atom_at, const double *__restrict__ x,
    const double *__restrict__ q, PmeParams pp, TileParams tp,
    float *__restrict__ tiles, int *__restrict__ flags) {
  extern __shared__ float s_tile[];
  __shared__ int s_fx[kTileAtoms], s_fy[kTileAtoms], s_fz[kTileAtoms];
  __shared__ float s_w[kTileAtoms][3 * N];
  const int col = blockIdx.x / tp.nzc, zc = blockIdx.x % tp.nzc;
  const int K0 = pp.K[0], K1 = pp.K[1], K2 = pp.K[2];
  const int z0 = zc * tp.zc, zlen = min(tp.zc, K2 - z0);
  const int zs = zlen + 1;  // padded z stride of the LDS tile
  const int cx = col % tp.ncx, cy = col / tp.ncx;
  const int gx0 = tp.gx0[cx], gy0 = tp.gy0[cy];
  const int p0 = col_cl_start[col] * kCluster, np = col_cl_count[col] * kCluster;
  const int npts = tp.TWx * tp.TWy;
  for (int k = threadIdx.x; k < npts * zs; k += blockDim.x) s_tile[k] = 0.0f;
  for (int a0 = 0; a0 < np; a0 += kTileAtoms) {
    const int na = min(kTileAtoms, np - a0);
    __syncthreads();
    for (int t = threadIdx.x; t < na; t += blockDim.x) {
      const int atom = atom_at[p0 + a0 + t];
      int fx = -1 << 20, fy = 0, fz = 0;
      if (atom >= 0 && q[atom] != 0.0) {
        int f[3];
        float M[3][N];
        for (int d = 0; d < 3; ++d) {
          float w;
          pme_place(x[3 * atom + d], pp.box[d], pp.inv_box[d], pp.K[d], N, f[d], w);
          bspline<float>(w, N, M[d], nullptr);
        }
        fx = f[0] - gx0;
        if (fx < 0) fx += K0;
        fy = f[1] - gy0;
        if (fy < 0) fy += K1;
        fz = f[2];
        // an atom outside its column's tile would lose its charge
        if ((tp.TWx < K0 && fx > tp.TWx - N) || (tp.TWy < K1 && fy > tp.TWy - N))
          atomicOr(&flags[kFlagPmeRange], 1);
        const float qa = (float)q[atom];
#pragma unroll
        for (int k = 0; k < N; ++k) {
          s_w[t][k] = qa * M[0][k];
          s_w[t][N + k] = M[1][k];
          s_w[t][2 * N + k] = M[2][k];
        }
      }
      s_fx[t] = fx;
      s_fy[t] = fy;
      s_fz[t] = fz;
    }
    __syncthreads();
    for (int xy = threadIdx.x; xy < npts; xy += blockDim.x) {
      const int tx = xy / tp.TWy, ty = xy % tp.TWy;
      float *dst = s_tile + xy * zs;
      for (int a = 0; a < na; ++a) {
        int ox = tx - s_fx[a];
        if (tp.TWx == K0 && ox < 0) ox += K0;
        if (ox < 0 || ox >= N) continue;
        int oy = ty - s_fy[a];
        if (tp.TWy == K1 && oy < 0) oy += K1;
        if (oy < 0 || oy >= N) continue;
        float wxy = s_w[a][0], wy = s_w[a][N];
#pragma unroll
        for (int k = 1; k < N; ++k) {
          if (k == ox) wxy = s_w[a][k];
          if (k == oy) wy = s_w[a][N + k];
        }
        wxy *= wy;
        const int fz = s_fz[a];
#pragma unroll
        for (int k = 0; k < N; ++k) {
          int z = fz + k;
          if (z >= K2) z -= K2;
          const int rz = z - z0;
          if (rz >= 0 && rz < zlen) dst[rz] += wxy * s_w[a][2 * N + k];
        }
      }
    }
  }
  __syncthreads();
  float *out = tiles + (size_t)col * npts * K2 + z0;
  for (int k = threadIdx.x; k < npts * zlen; k += blockDim.x) {
    const int xy = k / zlen, z = k % zlen;
    out[(size_t)xy * K2 + z] = s_tile[xy * zs + z];
  }
}

// N lanes per atom (place order): lane t0 gathers plane x = first + t0, the
// N lanes then sum their three partial gradients.
template <int N, int LPA>
__global__ void __launch_bounds__(256) pme_gather_lanes(
    const int *__restrict__ counts, const int *__restrict__ atom_at,
    const double *__restrict__ x, const double *__restrict__ q, PmeParams pp,
    const float *__restrict__ phi, long long *__restrict__ F) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  const int place = g / LPA, t0 = g % LPA;
  const bool valid = place < counts[0];
  const int atom = valid ? atom_at[place] : -1;
  const double qa = atom >= 0 ? q[atom] : 0.0;
  float g0 = 0.0f, g1 = 0.0f, g2 = 0.0f;
  if (atom >= 0 && qa != 0.0 && t0 < N) {
    int f[3];
    float M[3][N], dM[3][N];
    for (int d = 0; d < 3; ++d) {
      float w;
      pme_place(x[3 * atom + d], pp.box[d], pp.inv_box[d], pp.K[d], N, f[d], w);
      bspline<float>(w, N, M[d], dM[d]);
    }
    float mx = M[0][0], dmx = dM[0][0];
#pragma unroll
    for (int k = 1; k < N; ++k)
      if (k == t0) {
        mx = M[0][k];
        dmx = dM[0][k];
      }
    int i0 = f[0] + t0;
    if (i0 >= pp.K[0]) i0 -= pp.K[0];
    const float *plane = phi + (size_t)i0 * pp.K[1] * pp.K[2];
#pragma unroll
    for (int t1 = 0; t1 < N; ++t1) {
      int i1 = f[1] + t1;
      if (i1 >= pp.K[1]) i1 -= pp.K[1];
      const float *row = plane + (size_t)i1 * pp.K[2];
      float s = 0.0f, ds = 0.0f;
#pragma unroll
      for (int t2 = 0; t2 < N; ++t2) {
        int i2 = f[2] + t2;
        if (i2 >= pp.K[2]) i2 -= pp.K[2];
        const float p = row[i2];
        s += p * M[2][t2];
        ds += p * dM[2][t2];
      }
      g0 += dmx * M[1][t1] * s;
      g1 += mx * dM[1][t1] * s;
      g2 += mx * M[1][t1] * ds;
    }
  }
#pragma unroll
  for (int m = 1; m < LPA; m <<= 1) {
    g0 += __shfl_xor(g0, m);
    g1 += __shfl_xor(g1, m);
    g2 += __shfl_xor(g2, m);
  }
  if (atom >= 0 && qa != 0.0 && t0 == 0) {
    const float qf = (float)qa;
    const float fx = -qf * (float)(pp.K[0] * pp.inv_box[0]) * g0;
    const float fy = -qf * (float)(pp.K[1] * pp.inv_box[1]) * g1;
    const float fz = -qf * (float)(pp.K[2] * pp.inv_box[2]) * g2;
    unsigned long long *u = reinterpret_cast<unsigned long long *>(F + 3 * place);
    atomicAdd(u + 0, (unsigned long long)to_fixed(fx));
    atomicAdd(u + 1, (unsigned long long)to_fixed(fy));
    atomicAdd(u + 2, (unsigned long long)to_fixed(fz));
  }
}

// ---- FFT passes ---------------------------------------------------------

template <typename T>
constexpr int tile_lines() { return sizeof(T) == 4 ? 16 : 8; }

// Forward along z of pairs of real rows packed as one complex line, split
// into the two half spectra.  SRC_FIXED reads (and clears) 64-bit charge
// words; otherwise a real T grid.
template <typename T, bool SRC_FIXED>
__global__ void __launch_bounds__(256) fft_rows_r2c(
    const void *__restrict__ src_v, cpx<T> *__restrict__ out, int nrows,
    FftAxis az) {
  constexpr int TL = tile_lines<T>();
  extern __shared__ unsigned char smem[];
  cpx<T> *a = reinterpret_cast<cpx<T> *>(smem);
  const int K = az.K, H = K / 2 + 1;
  cpx<T> *b = a + TL * K;
  const int row0 = blockIdx.x * 2 * TL;
  for (int w = threadIdx.x; w < 2 * TL * K; w += blockDim.x) {
    const int z = w % K, rr = w / K;  // rr: row within the tile
    const int row = row0 + rr;
    T v = (T)0;
    if (row < nrows) {
      const size_t idx = (size_t)row * K + z;
      if (SRC_FIXED) {
        unsigned long long *src = (unsigned long long *)src_v;
        v = (T)((double)(long long)src[idx] * kGridUnscale);
        src[idx] = 0ull;
      } else {
        v = ((const T *)src_v)[idx];
      }
    }
    const int t = rr >> 1;
    if (rr & 1)
      a[z * TL + t].y = v;
    else
      a[z * TL + t].x = v;
  }
  __syncthreads();
  cpx<T> *res = lds_fft<T, TL>(a, b, az, -1);
  for (int w = threadIdx.x; w < TL * H; w += blockDim.x) {
    const int k = w % H, t = w / H;
    const cpx<T> zk = res[k * TL + t];
    const cpx<T> zm = res[((K - k) % K) * TL + t];
    const int r1 = row0 + 2 * t, r2 = r1 + 1;
    if (r1 < nrows)
      out[(size_t)r1 * H + k] = {(zk.x + zm.x) * (T)0.5, (zk.y - zm.y) * (T)0.5};
    if (r2 < nrows)
      out[(size_t)r2 * H + k] = {(zk.y + zm.y) * (T)0.5, (zm.x - zk.x) * (T)0.5};
  }
}

// Forward along z of the grid rows summed from the column tiles of
// pme_spread_tiles (in a fixed order: y cover, then x cover).
template <typename T>
__global__ void __launch_bounds__(256) fft_rows_r2c_tiles(
    const float *__restrict__ tiles, TileParams tp, cpx<T> *__restrict__ out,
    int K1, FftAxis az) {
  constexpr int TL = tile_lines<T>();
  extern __shared__ unsigned char smem[];
  cpx<T> *a = reinterpret_cast<cpx<T> *>(smem);
  const int K = az.K, H = K / 2 + 1;
  cpx<T> *b = a + TL * K;
  const int row0 = blockIdx.x * 2 * TL;
  const int npts = tp.TWx * tp.TWy;
  for (int w = threadIdx.x; w < 2 * TL * K; w += blockDim.x) {
    const int z = w % K, rr = w / K;
    const int row = row0 + rr;
    const int ix = row / K1, iy = row % K1;
    float v = 0.0f;
    if (ix < tp.K0) {
      for (int cyi = 0; cyi < kMaxCover; ++cyi) {
        const int ce = tp.ycov[iy * kMaxCover + cyi];
        if (ce < 0) break;
        const int cy = ce >> 16, oy = ce & 0xffff;
        for (int cxi = 0; cxi < kMaxCover; ++cxi) {
          const int xe = tp.xcov[ix * kMaxCover + cxi];
          if (xe < 0) break;
          const int cx = xe >> 16, ox = xe & 0xffff;
          const int col = cy * tp.ncx + cx;
          v += tiles[((size_t)col * npts + ox * tp.TWy + oy) * K + z];
        }
      }
    }
    const int t = rr >> 1;
    if (rr & 1)
      a[z * TL + t].y = (T)v;
    else
      a[z * TL + t].x = (T)v;
  }
  __syncthreads();
  cpx<T> *res = lds_fft<T, TL>(a, b, az, -1);
  const int rows_total = tp.K0 * K1;
  for (int w = threadIdx.x; w < TL * H; w += blockDim.x) {
    const int k = w % H, t = w / H;
    const cpx<T> zk = res[k * TL + t];
    const cpx<T> zm = res[((K - k) % K) * TL + t];
    const int r1 = row0 + 2 * t, r2 = r1 + 1;
    if (r1 < rows_total)
      out[(size_t)r1 * H + k] = {(zk.x + zm.x) * (T)0.5, (zk.y - zm.y) * (T)0.5};
    if (r2 < rows_total)
      out[(size_t)r2 * H + k] = {(zk.y + zm.y) * (T)0.5, (zm.x - zk.x) * (T)0.5};
  }
}

// Backward along z: two half spectra as one complex line, real part to the
// first row and imaginary part to the second.  The imaginary parts of the
// zero and (even K) Nyquist terms are ignored (mdrt_fft.c:149-157).
template <typename T>
__global__ void __launch_bounds__(256) fft_rows_c2r(const cpx<T> *__restrict__ in,
                                                    T *__restrict__ dst, int nrows,
                                                    FftAxis az) {
  constexpr int TL = tile_lines<T>();
  extern __shared__ unsigned char smem[];
  cpx<T> *a = reinterpret_cast<cpx<T> *>(smem);
  const int K = az.K, H = K / 2 + 1;
  cpx<T> *b = a + TL * K;
  const int row0 = blockIdx.x * 2 * TL;
  for (int w = threadIdx.x; w < TL * K; w += blockDim.x) {
    const int k = w % K, t = w / K;
    const int r1 = row0 + 2 * t, r2 = r1 + 1;
    const int m = k < H ? k : K - k;
    cpx<T> x1 = {(T)0, (T)0}, x2 = {(T)0, (T)0};
    if (r1 < nrows) x1 = in[(size_t)r1 * H + m];
    if (r2 < nrows) x2 = in[(size_t)r2 * H + m];
    if (m == 0 || (K % 2 == 0 && m == K / 2)) {
      x1.y = (T)0;
      x2.y = (T)0;
    }
    cpx<T> zv;
    if (k < H)
      zv = {x1.x - x2.y, x1.y + x2.x};
    else
      zv = {x1.x + x2.y, -x1.y + x2.x};
    a[k * TL + t] = zv;
  }
  __syncthreads();
  cpx<T> *res = lds_fft<T, TL>(a, b, az, +1);
  for (int w = threadIdx.x; w < 2 * TL * K; w += blockDim.x) {
    const int z = w % K, rr = w / K;
    const int row = row0 + rr;
    if (row >= nrows) continue;
    const cpx<T> v = res[z * TL + (rr >> 1)];
    dst[(size_t)row * K + z] = (rr & 1) ? v.y : v.x;
  }
}

// Complex lines along y (axis 1: stride H) of the half spectrum, in place.
template <typename T>
__global__ void __launch_bounds__(256) fft_cols_y(cpx<T> *__restrict__ data, int K2,
                                                  int H, FftAxis ay, int sign) {
  constexpr int TL = tile_lines<T>();
  extern __shared__ unsigned char smem[];
  cpx<T> *a = reinterpret_cast<cpx<T> *>(smem);
  cpx<T> *b = a + TL * K2;
  const int tiles = (H + TL - 1) / TL;
  const int i0 = blockIdx.x / tiles, z0 = (blockIdx.x % tiles) * TL;
  cpx<T> *base = data + (size_t)i0 * K2 * H + z0;
  for (int w = threadIdx.x; w < TL * K2; w += blockDim.x) {
    const int t = w % TL, y = w / TL;
    a[y * TL + t] = (z0 + t < H) ? base[(size_t)y * H + t] : cpx<T>{(T)0, (T)0};
  }
  __syncthreads();
  cpx<T> *res = lds_fft<T, TL>(a, b, ay, sign);
  for (int w = threadIdx.x; w < TL * K2; w += blockDim.x) {
    const int t = w % TL, y = w / TL;
    if (z0 + t < H) base[(size_t)y * H + t] = res[y * TL + t];
  }
}

// Complex lines along x (stride K2 H): forward, influence product (and its
// energy), backward (mode 0); the product and backward only (mode 2); or a
// single direction (mode -1 forward, +1 backward) for the transform entry
// points of the runtime ABI.
struct ConvTables {
  const void *t[3];   // t_a(k) = exp(-pi^2 m_a^2 / beta^2) |b_a(k)|^2, t_0 * f/(pi V)
  const void *m[3];   // m_a(k) = signed k / L_a
  int K3;
};

template <typename T>
__global__ void __launch_bounds__(256) fft_x_conv(cpx<T> *__restrict__ data, int K1,
                                                  int K2, int H, FftAxis ax,
                                                  ConvTables ct, int mode,
                                                  double *__restrict__ energy_partial) {
  constexpr int TL = tile_lines<T>();
  extern __shared__ unsigned char smem[];
  cpx<T> *a = reinterpret_cast<cpx<T> *>(smem);
  cpx<T> *b = a + TL * K1;
  __shared__ double s_e[256];
  const int tiles = (H + TL - 1) / TL;
  const int i1 = blockIdx.x / tiles, z0 = (blockIdx.x % tiles) * TL;
  const size_t stride = (size_t)K2 * H;
  cpx<T> *base = data + (size_t)i1 * H + z0;
  for (int w = threadIdx.x; w < TL * K1; w += blockDim.x) {
    const int t = w % TL, xx = w / TL;
    a[xx * TL + t] = (z0 + t < H) ? base[xx * stride + t] : cpx<T>{(T)0, (T)0};
  }
  __syncthreads();
  if (mode == -1 || mode == 1) {  // single direction: -1 forward, +1 backward
    cpx<T> *res = lds_fft<T, TL>(a, b, ax, mode);
    for (int w = threadIdx.x; w < TL * K1; w += blockDim.x) {
      const int t = w % TL, xx = w / TL;
      if (z0 + t < H) base[xx * stride + t] = res[xx * TL + t];
    }
    return;
  }
  // mode 0: forward, product, backward; mode 2: product, backward (the
  // spectrum is already transformed along x)
  cpx<T> *res = mode == 0 ? lds_fft<T, TL>(a, b, ax, -1) : a;
  const T *t0 = (const T *)ct.t[0], *t1 = (const T *)ct.t[1], *t2 = (const T *)ct.t[2];
  const T *m0 = (const T *)ct.m[0], *m1 = (const T *)ct.m[1], *m2 = (const T *)ct.m[2];
  double e = 0.0;
  for (int w = threadIdx.x; w < TL * K1; w += blockDim.x) {
    const int t = w % TL, kx = w / TL;
    const int kz = z0 + t;
    cpx<T> v = res[kx * TL + t];
    T bc = (T)0;
    if (kz < H && (kx | i1 | kz) != 0) {
      const T msq = m0[kx] * m0[kx] + m1[i1] * m1[i1] + m2[kz] * m2[kz];
      bc = t0[kx] * t1[i1] * t2[kz] / msq;
      if (energy_partial) {
        const double h = (kz == 0 || (ct.K3 % 2 == 0 && kz == ct.K3 / 2)) ? 0.5 : 1.0;
        e += h * (double)bc * ((double)v.x * v.x + (double)v.y * v.y);
      }
    }
    res[kx * TL + t] = {v.x * bc, v.y * bc};
  }
  __syncthreads();
  cpx<T> *other = (res == a) ? b : a;
  cpx<T> *res2 = lds_fft<T, TL>(res, other, ax, +1);
  for (int w = threadIdx.x; w < TL * K1; w += blockDim.x) {
    const int t = w % TL, xx = w / TL;
    if (z0 + t < H) base[xx * stride + t] = res2[xx * TL + t];
  }
  if (energy_partial) {
    s_e[threadIdx.x] = e;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
      if (threadIdx.x < s) s_e[threadIdx.x] += s_e[threadIdx.x + s];
      __syncthreads();
    }
    if (threadIdx.x == 0) energy_partial[blockIdx.x] = s_e[0];
  }
}

template __global__ void fft_rows_r2c<float, true>(const void *, cpx<float> *, int, FftAxis);
template __global__ void fft_rows_r2c_tiles<float>(const float *, TileParams, cpx<float> *, int, FftAxis);
template __global__ void pme_spread_tiles<4>(const int *, const int *, const int *, const double *, const double *, PmeParams, TileParams, float *, int *);
template __global__ void pme_spread_tiles<6>(const int *, const int *, const int *, const double *, const double *, PmeParams, TileParams, float *, int *);
template __global__ void pme_spread_tiles<8>(const int *, const int *, const int *, const double *, const double *, PmeParams, TileParams, float *, int *);
template __global__ void pme_gather_lanes<4, 4>(const int *, const int *, const double *, const double *, PmeParams, const float *, long long *);
template __global__ void pme_gather_lanes<6, 8>(const int *, const int *, const double *, const double *, PmeParams, const float *, long long *);
template __global__ void pme_gather_lanes<8, 8>(const int *, const int *, const double *, const double *, PmeParams, const float *, long long *);
template __global__ void fft_rows_r2c<float, false>(const void *, cpx<float> *, int, FftAxis);
template __global__ void fft_rows_r2c<double, false>(const void *, cpx<double> *, int, FftAxis);
template __global__ void fft_rows_c2r<float>(const cpx<float> *, float *, int, FftAxis);
template __global__ void fft_rows_c2r<double>(const cpx<double> *, double *, int, FftAxis);
template __global__ void fft_cols_y<float>(cpx<float> *, int, int, FftAxis, int);
template __global__ void fft_cols_y<double>(cpx<double> *, int, int, FftAxis, int);
template __global__ void fft_x_conv<float>(cpx<float> *, int, int, int, FftAxis, ConvTables, int, double *);
template __global__ void fft_x_conv<double>(cpx<double> *, int, int, int, FftAxis, ConvTables, int, double *);

}  // namespace mdhip
