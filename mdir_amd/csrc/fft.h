// fft.h — LDS-resident mixed-radix Stockham FFT for the PME grids.
//
// Conventions of the reference's transforms (runtime/mdrt_fft.c:110-256,
// cuFFT calls of runtime/mdrt_cuda.c:2901-2991): forward exp(-2 pi i k m/K),
// backward exp(+2 pi i k m/K), unnormalised; the real grid K1 x K2 x K3 is
// row-major (z fastest) and its half spectrum K1 x K2 x (K3/2+1) complex
// numbers, interleaved.
//
// A workgroup holds a tile of TL lines of one axis in LDS (element e of line
// t at [e * TL + t], so a row of the tile is TL consecutive complex numbers
// of the grid and the global loads of a tile are coalesced) and runs every
// radix pass of the line there; only the first load and the last store touch
// HBM.  Radices 8, 4, 2 (then 3, 5, 7) in registers, twiddles from a table of
// exp(-2 pi i m / K) computed in f64 on the host.
#pragma once

#include <hip/hip_runtime.h>

namespace mdhip {

template <typename T>
struct cpx {
  T x, y;
};

template <typename T>
__device__ __forceinline__ cpx<T> cmul(cpx<T> a, cpx<T> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename T>
__device__ __forceinline__ cpx<T> cadd(cpx<T> a, cpx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
__device__ __forceinline__ cpx<T> csub(cpx<T> a, cpx<T> b) { return {a.x - b.x, a.y - b.y}; }
// a * (s i), s = +-1
template <typename T>
__device__ __forceinline__ cpx<T> cmuli(cpx<T> a, int s) {
  return s > 0 ? cpx<T>{-a.y, a.x} : cpx<T>{a.y, -a.x};
}

constexpr int kMaxRadices = 8;

struct FftAxis {
  int K;
  int nrad;
  int rad[kMaxRadices];
  const void *tw;  // cpx<T>[K]: exp(-2 pi i m / K)
};

// y[k] = sum_r v[r] exp(s 2 pi i r k / R), R known at compile time so that
// v stays in registers.
template <typename T, int R>
__device__ __forceinline__ void butterfly(cpx<T> *v, int s) {
  if constexpr (R == 2) {
    cpx<T> a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  } else if constexpr (R == 4) {
    cpx<T> a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
    cpx<T> b0 = cadd(v[1], v[3]), b1 = cmuli(csub(v[1], v[3]), s);
    v[0] = cadd(a0, b0);
    v[2] = csub(a0, b0);
    v[1] = cadd(a1, b1);
    v[3] = csub(a1, b1);
  } else if constexpr (R == 8) {
    const T h = (T)0.70710678118654752440;
    cpx<T> e[4] = {v[0], v[2], v[4], v[6]};
    cpx<T> o[4] = {v[1], v[3], v[5], v[7]};
    butterfly<T, 4>(e, s);
    butterfly<T, 4>(o, s);
    const T ss = (T)s;
    cpx<T> t1 = {h * (o[1].x - ss * o[1].y), h * (o[1].y + ss * o[1].x)};
    cpx<T> t2 = cmuli(o[2], s);
    cpx<T> t3 = {h * (-o[3].x - ss * o[3].y), h * (-o[3].y + ss * o[3].x)};
    v[0] = cadd(e[0], o[0]);
    v[4] = csub(e[0], o[0]);
    v[1] = cadd(e[1], t1);
    v[5] = csub(e[1], t1);
    v[2] = cadd(e[2], t2);
    v[6] = csub(e[2], t2);
    v[3] = cadd(e[3], t3);
    v[7] = csub(e[3], t3);
  } else {
    // radices 3, 5, 7: direct sum, constants folded at compile time
    cpx<T> in[R];
#pragma unroll
    for (int r = 0; r < R; ++r) in[r] = v[r];
#pragma unroll
    for (int k = 0; k < R; ++k) {
      cpx<T> acc = in[0];
#pragma unroll
      for (int r = 1; r < R; ++r) {
        const int m = (r * k) % R;
        const T c = (T)__builtin_cos(6.283185307179586476925286766559 * m / R);
        const T sn = (T)__builtin_sin(6.283185307179586476925286766559 * m / R);
        const T sg = (T)s * sn;
        acc.x += in[r].x * c - in[r].y * sg;
        acc.y += in[r].x * sg + in[r].y * c;
      }
      v[k] = acc;
    }
  }
}

// One Stockham pass of radix R over TL lines (Govindaraju et al., SC'08):
// after sub-transforms of length Ns, v[r] = in[j + r K/R] is twiddled by
// w^{(j mod Ns) r K/(Ns R)} and written to out[(j/Ns) Ns R + j mod Ns + r Ns].
template <typename T, int TL, int R>
__device__ __forceinline__ void stockham_pass(const cpx<T> *in, cpx<T> *out, int K,
                                              int Ns, const cpx<T> *tw, int sign) {
  const int Kr = K / R;
  const int tstep = K / (Ns * R);
  for (int w = threadIdx.x; w < TL * Kr; w += blockDim.x) {
    const int line = w % TL, j = w / TL;
    cpx<T> v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = in[(j + r * Kr) * TL + line];
    const int k = j % Ns;
    if (Ns > 1) {
#pragma unroll
      for (int r = 1; r < R; ++r) {
        cpx<T> t = tw[k * r * tstep];
        if (sign > 0) t.y = -t.y;
        v[r] = cmul(v[r], t);
      }
    }
    butterfly<T, R>(v, sign);
    const int base = (j / Ns) * Ns * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) out[(base + r * Ns) * TL + line] = v[r];
  }
}

// Runs all passes of TL lines of length ax.K held in `a` (scratch `b`),
// sign -1 forward / +1 backward.  Returns the buffer holding the result.
// tw: the twiddles to read (a copy in LDS), or null for the axis' own.
template <typename T, int TL>
__device__ cpx<T> *lds_fft(cpx<T> *a, cpx<T> *b, const FftAxis &ax, int sign,
                           const cpx<T> *tw = nullptr) {
  const int K = ax.K;
  if (!tw) tw = static_cast<const cpx<T> *>(ax.tw);
  int Ns = 1;
  cpx<T> *in = a, *out = b;
  for (int s = 0; s < ax.nrad; ++s) {
    const int R = ax.rad[s];
    switch (R) {
      case 8: stockham_pass<T, TL, 8>(in, out, K, Ns, tw, sign); break;
      case 4: stockham_pass<T, TL, 4>(in, out, K, Ns, tw, sign); break;
      case 2: stockham_pass<T, TL, 2>(in, out, K, Ns, tw, sign); break;
      case 3: stockham_pass<T, TL, 3>(in, out, K, Ns, tw, sign); break;
      case 5: stockham_pass<T, TL, 5>(in, out, K, Ns, tw, sign); break;
      default: stockham_pass<T, TL, 7>(in, out, K, Ns, tw, sign); break;
    }
    __syncthreads();
    cpx<T> *t = in;
    in = out;
    out = t;
    Ns *= R;
  }
  return in;
}

// A K1 x K2 plane of complex numbers held whole in LDS, element (x, y) at
// x K2 + (y ^ s(x)), s(x) = (x & 1) << 3 | (x >> 1) & 7 when 16 divides K2
// (swz != 0), else 0: the rows a y-pass wave touches land in all 16 LDS
// bank pairs instead of 8.
__device__ __forceinline__ int plane_index(int x, int y, int K2, int swz) {
  return x * K2 + (y ^ (swz & (((x & 1) << 3) | ((x >> 1) & 7))));
}

// One Stockham pass over the lines of a plane: along y (AXIS 1: K1 lines of
// K2) or along x (AXIS 0: K2 lines of K1).  Threads run along the elements
// of a line for y and along the lines for x, so a wave's accesses are
// consecutive up to the swizzle.
template <typename T, int R, int AXIS>
__device__ __forceinline__ void stockham_pass_plane(const cpx<T> *in, cpx<T> *out, int K1,
                                                    int K2, int swz, int Ns,
                                                    const cpx<T> *tw, int sign) {
  const int K = AXIS == 1 ? K2 : K1, nlines = AXIS == 1 ? K1 : K2;
  const int Kr = K / R;
  const int tstep = K / (Ns * R);
  for (int w = threadIdx.x; w < nlines * Kr; w += blockDim.x) {
    int line, j;
    if (AXIS == 1) {
      j = w % Kr;
      line = w / Kr;
    } else {
      line = w % nlines;
      j = w / nlines;
    }
    cpx<T> v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int e = j + r * Kr;
      v[r] = in[AXIS == 1 ? plane_index(line, e, K2, swz) : plane_index(e, line, K2, swz)];
    }
    const int k = j % Ns;
    if (Ns > 1) {
#pragma unroll
      for (int r = 1; r < R; ++r) {
        cpx<T> t = tw[k * r * tstep];
        if (sign > 0) t.y = -t.y;
        v[r] = cmul(v[r], t);
      }
    }
    butterfly<T, R>(v, sign);
    const int base = (j / Ns) * Ns * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int e = base + r * Ns;
      out[AXIS == 1 ? plane_index(line, e, K2, swz) : plane_index(e, line, K2, swz)] = v[r];
    }
  }
}

// All passes along one axis of the plane in `a` (scratch `b`); returns the
// buffer holding the result.  tw: the axis' twiddles (in LDS).
template <typename T, int AXIS>
__device__ cpx<T> *lds_fft_plane(cpx<T> *a, cpx<T> *b, const FftAxis &ax, int sign,
                                 int K1, int K2, int swz, const cpx<T> *tw) {
  int Ns = 1;
  cpx<T> *in = a, *out = b;
  for (int s = 0; s < ax.nrad; ++s) {
    const int R = ax.rad[s];
    switch (R) {
      case 8: stockham_pass_plane<T, 8, AXIS>(in, out, K1, K2, swz, Ns, tw, sign); break;
      case 4: stockham_pass_plane<T, 4, AXIS>(in, out, K1, K2, swz, Ns, tw, sign); break;
      case 2: stockham_pass_plane<T, 2, AXIS>(in, out, K1, K2, swz, Ns, tw, sign); break;
      case 3: stockham_pass_plane<T, 3, AXIS>(in, out, K1, K2, swz, Ns, tw, sign); break;
      case 5: stockham_pass_plane<T, 5, AXIS>(in, out, K1, K2, swz, Ns, tw, sign); break;
      default: stockham_pass_plane<T, 7, AXIS>(in, out, K1, K2, swz, Ns, tw, sign); break;
    }
    __syncthreads();
    cpx<T> *t = in;
    in = out;
    out = t;
    Ns *= R;
  }
  return in;
}

// The same passes with the plane's edges known at compile time (the grids
// the bench and its configs use): every index division is by a constant,
// the radix order is the host's (8s, then 4, 2, 3s, 5s, 7s), and the pass
// loop unrolls.
constexpr int next_radix(int rest) {
  return rest % 8 == 0 ? 8 : rest % 4 == 0 ? 4 : rest % 2 == 0 ? 2 : rest % 3 == 0 ? 3
                                                                 : rest % 5 == 0 ? 5 : 7;
}

template <typename T, int R, int AXIS, int K1, int K2, int NS>
__device__ __forceinline__ void stockham_pass_plane_k(const cpx<T> *in, cpx<T> *out, int swz,
                                                      const cpx<T> *tw, int sign) {
  constexpr int K = AXIS == 1 ? K2 : K1, nlines = AXIS == 1 ? K1 : K2;
  constexpr int Kr = K / R;
  constexpr int tstep = K / (NS * R);
  for (int w = threadIdx.x; w < nlines * Kr; w += blockDim.x) {
    int line, j;
    if (AXIS == 1) {
      j = w % Kr;
      line = w / Kr;
    } else {
      line = w % nlines;
      j = w / nlines;
    }
    cpx<T> v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int e = j + r * Kr;
      v[r] = in[AXIS == 1 ? plane_index(line, e, K2, swz) : plane_index(e, line, K2, swz)];
    }
    const int k = j % NS;
    if (NS > 1) {
#pragma unroll
      for (int r = 1; r < R; ++r) {
        cpx<T> t = tw[k * r * tstep];
        if (sign > 0) t.y = -t.y;
        v[r] = cmul(v[r], t);
      }
    }
    butterfly<T, R>(v, sign);
    const int base = (j / NS) * NS * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int e = base + r * NS;
      out[AXIS == 1 ? plane_index(line, e, K2, swz) : plane_index(e, line, K2, swz)] = v[r];
    }
  }
}

template <typename T, int AXIS, int K1, int K2, int NS = 1>
__device__ __forceinline__ cpx<T> *lds_fft_plane_k(cpx<T> *in, cpx<T> *out, int sign, int swz,
                                                   const cpx<T> *tw) {
  constexpr int K = AXIS == 1 ? K2 : K1;
  if constexpr (NS >= K) {
    return in;
  } else {
    constexpr int R = next_radix(K / NS);
    stockham_pass_plane_k<T, R, AXIS, K1, K2, NS>(in, out, swz, tw, sign);
    __syncthreads();
    return lds_fft_plane_k<T, AXIS, K1, K2, NS * R>(out, in, sign, swz, tw);
  }
}

}  // namespace mdhip
