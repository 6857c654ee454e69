// pme.hip — the smooth PME reciprocal chain for gfx950.
//
// Reference semantics (docs/paper/05-electrostatics.md:95-279):
//   places    u = K frac(x/L) in f64, first point (floor(u) - n + 1) mod K
//             (PMEWeights.cpp:8)
//   spread    Q(l) = sum_j q_j prod_a M_n(u_aj - l_a)  (PMESpread.cpp; the
//             deterministic mode adds q prod M scaled by 2^40 with 64-bit
//             integer atomics, docs/pme.md:310-331) — done that way here,
//             in LDS, by bricks of grid rows fused with the z transform
//   R2C FFT   unnormalised, half spectrum along z
//   product   Q^ <- BC Q^, BC = t1 t2 t3 / m^2, C(0) = 0; energy
//             sum h BC |Q^|^2 with h = 1/2 on the k3 = 0 and Nyquist planes
//             (PMETables/PMEConvolve.cpp; the influence fused into the
//             transform as pme_callback.cu:5-21 fuses it into cuFFT's load)
//   C2R FFT   unnormalised: phi
//   gather    F_ja = -q_j (K_a/L_a) sum_l phi(l) M'(u_a - l_a) prod M
//             (PMEGather.cpp)
// The FFT is the LDS Stockham of fft.h; the forward z pass runs in the
// spread kernel on the brick it has just accumulated, and the x pass does
// forward transform, influence product and backward transform of each x
// line in one LDS residency.
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
  // fl is in [0, K] (K only when frac rounds up to 1), so f is in
  // [1 - order, K - order + 1] and one wrap suffices
  int f = (int)fl - order + 1;
  if (f < 0) f += K;
  first = f;
}

// ---- spread fused with the forward z transform ---------------------------
//
// A block owns a brick of BX x BY grid rows (all of z) and produces their
// half spectra along z.  pme_bin first lists, for every brick, the atoms
// whose B-spline stencil touches it (thread per atom, from the atom's own
// stencil, so no atom can be missed); each listed atom then adds the part
// of its stencil that lies in the brick as 64-bit fixed point (2^40 per e)
// by LDS integer atomics, so the sums do not depend on the order of the
// adds or of the list; the brick's rows are converted to f32 and
// transformed in place in LDS.  Every grid point belongs to one brick: no
// global atomics on the grid, no grid to clear, no real-space grid in HBM.

struct BrickParams {
  int nbx, nby;             // bricks along x and y
  int cap;                  // list capacity per brick
  int *count;               // atoms listed per brick; the gather clears it
  int *list;                // places of each brick's atoms, cap per brick
  int plane_major;          // half spectrum as [kz][x][y] instead of [x][y][kz]
};

template <int TL> struct BrickShape {
  static constexpr int R = 2 * TL;                  // real rows of a brick
  static constexpr int BY = R >= 8 ? 4 : 2;
  static constexpr int BX = R / BY;
};

// The bricks along one axis that the stencil f, f+1, .., f+N-1 (mod K)
// touches, each once.
template <int N, int B>
__device__ __forceinline__ int stencil_bricks(int f, int K, int *out) {
  int n = 0;
#pragma unroll
  for (int k = 0; k < N; ++k) {
    int p = f + k;
    if (p >= K) p -= K;
    const int b = p / B;
    bool seen = false;
    for (int j = 0; j < n; ++j) seen |= out[j] == b;
    if (!seen) out[n++] = b;
  }
  return n;
}

// Thread per place: list the place with every brick its stencil touches.
// Places run in column order, so a block touches few bricks: it counts its
// entries per brick in LDS, takes one range of each brick's list with one
// global atomic, then writes (the order within a list may vary from run to
// run; nothing downstream depends on it).
constexpr int kBinLdsBricks = 8192;
template <int N, int TL>
__global__ void __launch_bounds__(256) pme_bin(const int *__restrict__ counts,
                                               const double4 *__restrict__ xpq,
                                               PmeParams pp, BrickParams bp,
                                               int *__restrict__ flags, size_t fs) {
  constexpr int BX = BrickShape<TL>::BX, BY = BrickShape<TL>::BY;
  __shared__ int s_count[kBinLdsBricks];
  counts = frame_at(counts, fs);
  xpq = frame_at(xpq, fs);
  bp.count = frame_at(bp.count, fs);
  bp.list = frame_at(bp.list, fs);
  flags = frame_at(flags, fs);
  const int nbricks = bp.nbx * bp.nby;
  const bool lds = nbricks <= kBinLdsBricks;
  if (lds)
    for (int k = threadIdx.x; k < nbricks; k += blockDim.x) s_count[k] = 0;
  const int place = blockIdx.x * blockDim.x + threadIdx.x;
  int bx[N], by[N], nx = 0, ny = 0;
  if (place < counts[0]) {
    const double4 pq = xpq[place];
    if (pq.w != 0.0) {  // not padding, not an uncharged atom
      int f0, f1;
      float w;
      pme_place(pq.x, pp.box[0], pp.inv_box[0], pp.K[0], N, f0, w);
      pme_place(pq.y, pp.box[1], pp.inv_box[1], pp.K[1], N, f1, w);
      nx = stencil_bricks<N, BX>(f0, pp.K[0], bx);
      ny = stencil_bricks<N, BY>(f1, pp.K[1], by);
    }
  }
  int local[N * N];
  if (lds) {
    __syncthreads();
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) local[j * N + i] = atomicAdd(&s_count[by[j] * bp.nbx + bx[i]], 1);
    __syncthreads();
    for (int k = threadIdx.x; k < nbricks; k += blockDim.x) {
      const int c = s_count[k];
      if (c) s_count[k] = atomicAdd(&bp.count[k], c);
    }
    __syncthreads();
  }
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      const int brick = by[j] * bp.nbx + bx[i];
      const int slot = lds ? s_count[brick] + local[j * N + i] : atomicAdd(&bp.count[brick], 1);
      if (slot < bp.cap)
        bp.list[(size_t)brick * bp.cap + slot] = place;
      else
        atomicOr(&flags[kFlagBrickOverflow], 1);
    }
}

// The part of one atom's stencil (place-ordered position and charge pq)
// that lies in the brick at (bx0, by0), added to the brick's accumulators.
template <int N, int BX, int BY>
__device__ __forceinline__ void spread_atom(double4 pq, const PmeParams &pp, int bx0,
                                            int by0, unsigned long long *acc) {
  const int K0 = pp.K[0], K1 = pp.K[1], K2 = pp.K[2];
  const double qd = pq.w;
  int f[3];
  float w[3];
  pme_place(pq.x, pp.box[0], pp.inv_box[0], K0, N, f[0], w[0]);
  pme_place(pq.y, pp.box[1], pp.inv_box[1], K1, N, f[1], w[1]);
  pme_place(pq.z, pp.box[2], pp.inv_box[2], K2, N, f[2], w[2]);
  float M[3][N];
#pragma unroll
  for (int d = 0; d < 3; ++d) bspline<float>(w[d], N, M[d], nullptr);
  const float qf = (float)qd;
#pragma unroll
  for (int tx = 0; tx < BX; ++tx) {
    int kx = bx0 + tx - f[0];
    if (kx < 0) kx += K0;
    if (kx >= N || bx0 + tx >= K0) continue;
    float wx = M[0][0];
#pragma unroll
    for (int k = 1; k < N; ++k)
      if (k == kx) wx = M[0][k];
    wx *= qf;
#pragma unroll
    for (int ty = 0; ty < BY; ++ty) {
      int ky = by0 + ty - f[1];
      if (ky < 0) ky += K1;
      if (ky >= N || by0 + ty >= K1) continue;
      float wxy = M[1][0];
#pragma unroll
      for (int k = 1; k < N; ++k)
        if (k == ky) wxy = M[1][k];
      wxy *= wx;
      unsigned long long *row = acc + (tx * BY + ty) * K2;
#pragma unroll
      for (int k = 0; k < N; ++k) {
        int z = f[2] + k;
        if (z >= K2) z -= K2;
        const long long v = __double2ll_rn((double)(wxy * M[2][k]) * kGridScale);
        atomicAdd(row + z, (unsigned long long)v);
      }
    }
  }
}

// The brick's R accumulated real rows (64-bit words in acc) to their half
// spectra along z in `out`: rows 2t and 2t+1 as the real and imaginary
// parts of one complex line t, one LDS transform, then the two spectra
// split apart.  a, b hold TL K2 complex values each, tw the K2 twiddles.
template <int TL>
__device__ __forceinline__ void brick_forward_z(const unsigned long long *acc,
                                                cpx<float> *a, cpx<float> *b,
                                                const cpx<float> *tw, const FftAxis &az,
                                                const PmeParams &pp, const BrickParams &bp,
                                                int bx0, int by0, cpx<float> *__restrict__ out) {
  constexpr int R = BrickShape<TL>::R, BY = BrickShape<TL>::BY;
  const int K0 = pp.K[0], K1 = pp.K[1], K2 = pp.K[2], H = K2 / 2 + 1;
  for (int w = threadIdx.x; w < R * K2; w += blockDim.x) {
    const int rr = w / K2, z = w % K2;
    const float v = (float)((double)(long long)acc[w] * kGridUnscale);
    if (rr & 1)
      a[z * TL + (rr >> 1)].y = v;
    else
      a[z * TL + (rr >> 1)].x = v;
  }
  __syncthreads();
  cpx<float> *res = lds_fft<float, TL>(a, b, az, -1, tw);
  for (int w = threadIdx.x; w < TL * H; w += blockDim.x) {
    const int k = w % H, t = w / H;
    const cpx<float> zk = res[k * TL + t];
    const cpx<float> zm = res[((K2 - k) % K2) * TL + t];
    const int r1 = 2 * t, r2 = r1 + 1;
    const int ix1 = bx0 + r1 / BY, iy1 = by0 + r1 % BY;
    const int ix2 = bx0 + r2 / BY, iy2 = by0 + r2 % BY;
    const size_t rows = (size_t)K0 * K1;
    if (ix1 < K0 && iy1 < K1) {
      const size_t row = (size_t)ix1 * K1 + iy1;
      out[bp.plane_major ? k * rows + row : row * H + k] = {(zk.x + zm.x) * 0.5f,
                                                            (zk.y - zm.y) * 0.5f};
    }
    if (ix2 < K0 && iy2 < K1) {
      const size_t row = (size_t)ix2 * K1 + iy2;
      out[bp.plane_major ? k * rows + row : row * H + k] = {(zk.y + zm.y) * 0.5f,
                                                            (zm.x - zk.x) * 0.5f};
    }
  }
}

template <int N, int TL>
__global__ void __launch_bounds__(256) pme_spread_rows(
    const double4 *__restrict__ xpq, PmeParams pp, BrickParams bp,
    cpx<float> *__restrict__ out, FftAxis az, size_t fs) {
  constexpr int R = BrickShape<TL>::R, BX = BrickShape<TL>::BX, BY = BrickShape<TL>::BY;
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  xpq = frame_at(xpq, fs);
  bp.count = frame_at(bp.count, fs);
  bp.list = frame_at(bp.list, fs);
  out = frame_at(out, fs);
  const int K2 = pp.K[2];
  unsigned long long *acc = reinterpret_cast<unsigned long long *>(smem);
  cpx<float> *a = reinterpret_cast<cpx<float> *>(acc + (size_t)R * K2);
  cpx<float> *b = a + TL * K2;
  cpx<float> *tw = b + TL * K2;  // the z twiddles, read by every pass
  const int brick = blockIdx.x;
  const int bx0 = (brick % bp.nbx) * BX, by0 = (brick / bp.nbx) * BY;
  for (int w = threadIdx.x; w < R * K2; w += blockDim.x) acc[w] = 0ull;
  for (int w = threadIdx.x; w < K2; w += blockDim.x)
    tw[w] = static_cast<const cpx<float> *>(az.tw)[w];
  int nat = bp.count[brick];
  nat = nat < bp.cap ? nat : bp.cap;
  const int *lst = bp.list + (size_t)brick * bp.cap;
  __syncthreads();
  // two atoms per thread in flight: their loads are independent
  for (int i0 = threadIdx.x; i0 < nat; i0 += 2 * blockDim.x) {
    const int i1 = i0 + blockDim.x;
    const double4 a0 = xpq[lst[i0]];
    const double4 a1 = i1 < nat ? xpq[lst[i1]] : make_double4(0, 0, 0, 0);
    spread_atom<N, BX, BY>(a0, pp, bx0, by0, acc);
    if (i1 < nat) spread_atom<N, BX, BY>(a1, pp, bx0, by0, acc);
  }
  __syncthreads();
  brick_forward_z<TL>(acc, a, b, tw, az, pp, bp, bx0, by0, out);
}

// ---- backward z transform fused with the gather ---------------------------
//
// A block owns the same brick of grid rows as the spread: it transforms the
// brick's half-spectrum rows back to real potential rows in LDS, and every
// listed atom adds the part of its force that comes from the brick's points,
// as 64-bit fixed point (each part is a fixed f32 sum, so the total does not
// depend on the order of the adds).  The real potential grid never goes to
// HBM.  The block clears its brick's list count for the next step.

// One atom's force from the potential rows of the brick at (bx0, by0):
// -q (K/L) sum phi M' M M over the stencil points in the brick, added as
// 64-bit fixed point.  phi of brick row r at z is phi[z TL + r/2], .x for
// even r and .y for odd r.
template <int N, int TL, int BX, int BY>
__device__ __forceinline__ void gather_atom(double4 pq, int place, const PmeParams &pp,
                                            int bx0, int by0, const cpx<float> *phi,
                                            long long *__restrict__ F) {
  const int K0 = pp.K[0], K1 = pp.K[1], K2 = pp.K[2];
  const double qd = pq.w;
  int f[3];
  float w[3];
  pme_place(pq.x, pp.box[0], pp.inv_box[0], K0, N, f[0], w[0]);
  pme_place(pq.y, pp.box[1], pp.inv_box[1], K1, N, f[1], w[1]);
  pme_place(pq.z, pp.box[2], pp.inv_box[2], K2, N, f[2], w[2]);
  float M[3][N], dM[3][N];
#pragma unroll
  for (int d = 0; d < 3; ++d) bspline<float>(w[d], N, M[d], dM[d]);
  float g0 = 0.0f, g1 = 0.0f, g2 = 0.0f;
#pragma unroll
  for (int tx = 0; tx < BX; ++tx) {
    int kx = bx0 + tx - f[0];
    if (kx < 0) kx += K0;
    if (kx >= N || bx0 + tx >= K0) continue;
    float mx = M[0][0], dmx = dM[0][0];
#pragma unroll
    for (int k = 1; k < N; ++k)
      if (k == kx) {
        mx = M[0][k];
        dmx = dM[0][k];
      }
#pragma unroll
    for (int ty = 0; ty < BY; ++ty) {
      int ky = by0 + ty - f[1];
      if (ky < 0) ky += K1;
      if (ky >= N || by0 + ty >= K1) continue;
      float my = M[1][0], dmy = dM[1][0];
#pragma unroll
      for (int k = 1; k < N; ++k)
        if (k == ky) {
          my = M[1][k];
          dmy = dM[1][k];
        }
      const int r = tx * BY + ty;
      float sz = 0.0f, dsz = 0.0f;
#pragma unroll
      for (int k = 0; k < N; ++k) {
        int z = f[2] + k;
        if (z >= K2) z -= K2;
        const cpx<float> c = phi[z * TL + (r >> 1)];
        const float pv = (r & 1) ? c.y : c.x;
        sz += pv * M[2][k];
        dsz += pv * dM[2][k];
      }
      g0 += dmx * my * sz;
      g1 += mx * dmy * sz;
      g2 += mx * my * dsz;
    }
  }
  const float qf = (float)qd;
  unsigned long long *u = reinterpret_cast<unsigned long long *>(F + 3 * place);
  atomicAdd(u + 0, (unsigned long long)to_fixed(-qf * (float)(K0 * pp.inv_box[0]) * g0));
  atomicAdd(u + 1, (unsigned long long)to_fixed(-qf * (float)(K1 * pp.inv_box[1]) * g1));
  atomicAdd(u + 2, (unsigned long long)to_fixed(-qf * (float)(K2 * pp.inv_box[2]) * g2));
}

// The brick's half-spectrum rows of `in` back to real rows in LDS (one
// complex line t per pair of rows 2t, 2t+1): returns the buffer (a or b)
// holding the rows, phi of brick row r at z being phi[z TL + r/2], .x for
// even r and .y for odd r.
template <int TL>
__device__ __forceinline__ const cpx<float> *brick_backward_z(
    const cpx<float> *__restrict__ in, cpx<float> *a, cpx<float> *b, const cpx<float> *tw,
    const FftAxis &az, const PmeParams &pp, const BrickParams &bp, int bx0, int by0) {
  constexpr int BY = BrickShape<TL>::BY;
  const int K0 = pp.K[0], K1 = pp.K[1], K2 = pp.K[2], H = K2 / 2 + 1;
  const size_t rows = (size_t)K0 * K1;
  for (int w = threadIdx.x; w < TL * K2; w += blockDim.x) {
    const int k = w % K2, t = w / K2;
    const int r1 = 2 * t, r2 = r1 + 1;
    const int ix1 = bx0 + r1 / BY, iy1 = by0 + r1 % BY;
    const int ix2 = bx0 + r2 / BY, iy2 = by0 + r2 % BY;
    const int m = k < H ? k : K2 - k;
    cpx<float> x1 = {0.0f, 0.0f}, x2 = {0.0f, 0.0f};
    if (ix1 < K0 && iy1 < K1) {
      const size_t row = (size_t)ix1 * K1 + iy1;
      x1 = in[bp.plane_major ? m * rows + row : row * H + m];
    }
    if (ix2 < K0 && iy2 < K1) {
      const size_t row = (size_t)ix2 * K1 + iy2;
      x2 = in[bp.plane_major ? m * rows + row : row * H + m];
    }
    if (m == 0 || (K2 % 2 == 0 && m == K2 / 2)) {
      x1.y = 0.0f;
      x2.y = 0.0f;
    }
    a[k * TL + t] = k < H ? cpx<float>{x1.x - x2.y, x1.y + x2.x}
                          : cpx<float>{x1.x + x2.y, -x1.y + x2.x};
  }
  __syncthreads();
  return lds_fft<float, TL>(a, b, az, +1, tw);
}

template <int N, int TL>
__global__ void __launch_bounds__(256) pme_gather_rows(
    const double4 *__restrict__ xpq, PmeParams pp, BrickParams bp,
    const cpx<float> *__restrict__ in, FftAxis az, long long *__restrict__ F, size_t fs) {
  constexpr int BX = BrickShape<TL>::BX, BY = BrickShape<TL>::BY;
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  xpq = frame_at(xpq, fs);
  bp.count = frame_at(bp.count, fs);
  bp.list = frame_at(bp.list, fs);
  in = frame_at(in, fs);
  F = frame_at(F, fs);
  const int K2 = pp.K[2];
  cpx<float> *a = reinterpret_cast<cpx<float> *>(smem);
  cpx<float> *b = a + TL * K2;
  cpx<float> *tw = b + TL * K2;  // the z twiddles, read by every pass
  const int brick = blockIdx.x;
  const int bx0 = (brick % bp.nbx) * BX, by0 = (brick / bp.nbx) * BY;
  for (int w = threadIdx.x; w < K2; w += blockDim.x)
    tw[w] = static_cast<const cpx<float> *>(az.tw)[w];
  int nat = bp.count[brick];
  nat = nat < bp.cap ? nat : bp.cap;
  const int *lst = bp.list + (size_t)brick * bp.cap;
  const cpx<float> *phi = brick_backward_z<TL>(in, a, b, tw, az, pp, bp, bx0, by0);
  for (int i0 = threadIdx.x; i0 < nat; i0 += 2 * blockDim.x) {
    const int i1 = i0 + blockDim.x;
    const int p0 = lst[i0];
    const int p1 = i1 < nat ? lst[i1] : p0;
    const double4 a0 = xpq[p0];
    const double4 a1 = xpq[p1];
    gather_atom<N, TL, BX, BY>(a0, p0, pp, bx0, by0, phi, F);
    if (i1 < nat) gather_atom<N, TL, BX, BY>(a1, p1, pp, bx0, by0, phi, F);
  }
  __syncthreads();  // every thread has read the count
  if (threadIdx.x == 0) bp.count[brick] = 0;
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

// Backward along z: two half spectra as one complex line, real part to the
// first row and imaginary part to the second.  The imaginary parts of the
// zero and (even K) Nyquist terms are ignored (mdrt_fft.c:149-157).
template <typename T>
__global__ void __launch_bounds__(256) fft_rows_c2r(const cpx<T> *__restrict__ in,
                                                    T *__restrict__ dst, int nrows,
                                                    FftAxis az, int plane_major) {
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
    // half spectrum [row][k], or kz-plane major [k][row]
    if (r1 < nrows) x1 = in[plane_major ? (size_t)m * nrows + r1 : (size_t)r1 * H + m];
    if (r2 < nrows) x2 = in[plane_major ? (size_t)m * nrows + r2 : (size_t)r2 * H + m];
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
  float pb2;          // pi^2 / beta^2 (the reciprocal virial)
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

constexpr int kPlaneAxisMax = 256;

// The y and x passes of one kz plane, whole in LDS, for grids whose planes
// fit (K1 K2 <= 4096 points, K1, K2 <= 256): forward y, forward x,
// influence product (and its energy), backward x, backward y.  The half
// spectrum is kz-plane major ([kz][x][y]), so a plane is one contiguous read
// and write.
// CK1, CK2: the plane's edges at compile time (0: the run-time K1, K2 and
// the axes' radix lists).
template <int CK1, int CK2>
__global__ void __launch_bounds__(512) fft_plane_conv(cpx<float> *__restrict__ data, int K1,
                                                      int K2, FftAxis ax, FftAxis ay,
                                                      ConvTables ct,
                                                      double *__restrict__ energy_partial,
                                                      double *__restrict__ virial_partial,
                                                      size_t fs) {
  if (CK1 > 0) {
    K1 = CK1;
    K2 = CK2;
  }
  extern __shared__ unsigned char smem[];
  __shared__ double s_e[512];
  data = frame_at(data, fs);
  energy_partial = frame_at(energy_partial, fs);
  virial_partial = frame_at(virial_partial, fs);
  __shared__ cpx<float> s_twx[kPlaneAxisMax], s_twy[kPlaneAxisMax];
  __shared__ float s_t0[kPlaneAxisMax], s_m0[kPlaneAxisMax];
  __shared__ float s_t1[kPlaneAxisMax], s_m1[kPlaneAxisMax];
  const int np = K1 * K2;
  cpx<float> *a = reinterpret_cast<cpx<float> *>(smem);
  cpx<float> *b = a + np;
  const int kz = blockIdx.x;
  cpx<float> *plane = data + (size_t)kz * np;
  const int swz = K2 % 16 == 0 ? 15 : 0;
  // the plane (unrolled, so that the loads are issued together)
#pragma unroll 8
  for (int w = threadIdx.x; w < np; w += blockDim.x)
    a[plane_index(w / K2, w % K2, K2, swz)] = plane[w];
  // twiddles and influence tables in LDS: every pass reads them
  for (int w = threadIdx.x; w < K1; w += blockDim.x) {
    s_twx[w] = static_cast<const cpx<float> *>(ax.tw)[w];
    s_t0[w] = static_cast<const float *>(ct.t[0])[w];
    s_m0[w] = static_cast<const float *>(ct.m[0])[w];
  }
  for (int w = threadIdx.x; w < K2; w += blockDim.x) {
    s_twy[w] = static_cast<const cpx<float> *>(ay.tw)[w];
    s_t1[w] = static_cast<const float *>(ct.t[1])[w];
    s_m1[w] = static_cast<const float *>(ct.m[1])[w];
  }
  __syncthreads();
  cpx<float> *r;
  if constexpr (CK1 > 0) {
    r = lds_fft_plane_k<float, 1, CK1, CK2>(a, b, -1, swz, s_twy);
    r = lds_fft_plane_k<float, 0, CK1, CK2>(r, r == a ? b : a, -1, swz, s_twx);
  } else {
    r = lds_fft_plane<float, 1>(a, b, ay, -1, K1, K2, swz, s_twy);
    r = lds_fft_plane<float, 0>(r, r == a ? b : a, ax, -1, K1, K2, swz, s_twx);
  }
  const float t2 = static_cast<const float *>(ct.t[2])[kz];
  const float m2 = static_cast<const float *>(ct.m[2])[kz];
  const double h = (kz == 0 || (ct.K3 % 2 == 0 && kz == ct.K3 / 2)) ? 0.5 : 1.0;
  double e = 0.0;
  // the reciprocal virial sum_m E_m [delta_ab - 2 (1/m^2 + pi^2/beta^2) m_a m_b]
  // (05-electrostatics.md:262-279): xx yy zz xy xz yz
  double vir[6] = {0, 0, 0, 0, 0, 0};
  for (int w = threadIdx.x; w < np; w += blockDim.x) {
    const int kx = w / K2, ky = w % K2;
    const int i = plane_index(kx, ky, K2, swz);
    const cpx<float> v = r[i];
    float bc = 0.0f;
    if ((kx | ky | kz) != 0) {
      const float msq = s_m0[kx] * s_m0[kx] + s_m1[ky] * s_m1[ky] + m2 * m2;
      bc = s_t0[kx] * s_t1[ky] * t2 / msq;
      if (energy_partial || virial_partial) {
        const double em = h * (double)bc * ((double)v.x * v.x + (double)v.y * v.y);
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
    r[i] = {v.x * bc, v.y * bc};
  }
  __syncthreads();
  if constexpr (CK1 > 0) {
    r = lds_fft_plane_k<float, 0, CK1, CK2>(r, r == a ? b : a, +1, swz, s_twx);
    r = lds_fft_plane_k<float, 1, CK1, CK2>(r, r == a ? b : a, +1, swz, s_twy);
  } else {
    r = lds_fft_plane<float, 0>(r, r == a ? b : a, ax, +1, K1, K2, swz, s_twx);
    r = lds_fft_plane<float, 1>(r, r == a ? b : a, ay, +1, K1, K2, swz, s_twy);
  }
  for (int w = threadIdx.x; w < np; w += blockDim.x)
    plane[w] = r[plane_index(w / K2, w % K2, K2, swz)];
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

template __global__ void fft_rows_r2c<float, true>(const void *, cpx<float> *, int, FftAxis);
#define MDHIP_SPREAD(N, TL)                                                              \
  template __global__ void pme_bin<N, TL>(const int *, const double4 *, PmeParams,       \
      BrickParams, int *, size_t);                                                       \
  template __global__ void pme_spread_rows<N, TL>(const double4 *, PmeParams,            \
      BrickParams, cpx<float> *, FftAxis, size_t);                                       \
  template __global__ void pme_gather_rows<N, TL>(const double4 *, PmeParams,            \
      BrickParams, const cpx<float> *, FftAxis, long long *, size_t);
#define MDHIP_SPREAD_N(N) MDHIP_SPREAD(N, 8) MDHIP_SPREAD(N, 4) MDHIP_SPREAD(N, 2) MDHIP_SPREAD(N, 1)
MDHIP_SPREAD_N(4)
MDHIP_SPREAD_N(6)
MDHIP_SPREAD_N(8)
template __global__ void fft_rows_r2c<float, false>(const void *, cpx<float> *, int, FftAxis);
template __global__ void fft_rows_r2c<double, false>(const void *, cpx<double> *, int, FftAxis);
template __global__ void fft_rows_c2r<float>(const cpx<float> *, float *, int, FftAxis, int);
template __global__ void fft_rows_c2r<double>(const cpx<double> *, double *, int, FftAxis, int);
template __global__ void fft_cols_y<float>(cpx<float> *, int, int, FftAxis, int);
template __global__ void fft_cols_y<double>(cpx<double> *, int, int, FftAxis, int);
template __global__ void fft_x_conv<float>(cpx<float> *, int, int, int, FftAxis, ConvTables, int, double *);
template __global__ void fft_x_conv<double>(cpx<double> *, int, int, int, FftAxis, ConvTables, int, double *);

}  // namespace mdhip
