// adjoint.hip — reverse products of the frame energies on gfx950.
//
// For frames n with energies U_n(x; theta), an energy cotangent g_n and a
// cotangent Gbar_n on the position gradient dU_n/dx, the product in the
// parameters is
//   p = sum_n [ g_n dU_n/dtheta + sum_i Gbar_ni . d2U_n/(dx_i dtheta) ],
// and the position Hessian product is H_n v_n = d2U_n/dx2 v_n
// (python/mdir/torch.py:139-364, the frame_vjp and frame_hvp operators;
// docs/python-frames.md, the product of force matching: the loss
// sum |F - F_ref|^2 has the cotangent Gbar = -2 (F - F_ref)).
// theta holds the per-atom charges and the per-type sigma and epsilon of
// the Lennard-Jones pairs under the Lorentz-Berthelot rules.
//
// Kernels (frame batches: blockIdx.z is the frame, as in the force path):
//   cot_prepare       cotangent to place order, scaled by a power of two c
//                     so that every component is at most 1 (force matching:
//                     from the force words and the reference forces)
//   nb_adjoint        the pair loop's traversal with the adjoint pair terms;
//                     grid plane y = 0 takes the excluded and 1-4 pairs
//   pme_spread_dipole the dipole grid D(l) = sum_i q_i c_i . grad W_i(l),
//                     fused with the forward z transform like the spread
//   pme_gather_adj    value, gradient and Hessian of the potential grid and
//                     of the convolved dipole grid at every atom
//   adj_to_atoms, frames_sum   words to f64 in atom order, sum over frames
// The products are accumulated as 64-bit fixed point (2^28 words per unit
// of the scaled cotangent) by integer atomics: bit-exact run to run.
#include "md_common.h"

namespace mdhip {

constexpr double kAdjScale = 268435456.0;  // 2^28
constexpr double kAdjUnscale = 1.0 / 268435456.0;

__device__ __forceinline__ long long to_fixed_adj(float f) {
  return __double2ll_rn((double)f * kAdjScale);
}

enum AdjKind { kAdjTheta = 0, kAdjHessian = 1, kAdjVirial = 2, kAdjPairTable = 3 };

// Doubles per (segment, group) wave slot of the per-wave partials: the
// virial (6) or the products in up to kMaxTablePairs type-pair entries
// (sigma and epsilon of each).
constexpr int kMaxTablePairs = 4;
constexpr int kPartStride = 2 * kMaxTablePairs;

// Per type pair (ti * ntypes + tj): lj = (C12, C6, 12 C12/sigma, 6 C6/sigma);
// unit = (4 sigma^12, 4 sigma^6), the coefficients per unit epsilon; sel =
// the selected entry 0 .. kMaxTablePairs-1 of the pair-table products, or -1.
struct AdjTables {
  const float4 *lj;
  const float2 *unit;
  const int *sel;
};

// Per-system constants of the adjoint pair terms (unscaled modifier
// constants, unlike NbParams's).
struct AdjParams {
  float box[3];
  float rc2, rc, rs, width;
  float beta, tbs, mb2l2e, b2tbs;     // b2tbs = 2 beta^2 tbs
  float coulomb, shift_c;
  float rc12, rc6;                    // rc^-12, rc^-6 (potential shift)
  float fa12, fb12, fc12, fa6, fb6, fc6;   // force switch: P_n = c + u^3 (a/3 + b u/4)
  int ntypes;
};

// ---- dual numbers in f32: the derivative in r^2 of the pair functions ----
struct Dual {
  float v, d;
};
__device__ __forceinline__ Dual operator+(Dual a, Dual b) { return {a.v + b.v, a.d + b.d}; }
__device__ __forceinline__ Dual operator-(Dual a, Dual b) { return {a.v - b.v, a.d - b.d}; }
__device__ __forceinline__ Dual operator*(Dual a, Dual b) {
  return {a.v * b.v, a.d * b.v + a.v * b.d};
}
__device__ __forceinline__ Dual operator+(Dual a, float b) { return {a.v + b, a.d}; }
__device__ __forceinline__ Dual operator-(Dual a, float b) { return {a.v - b, a.d}; }
__device__ __forceinline__ Dual operator*(Dual a, float b) { return {a.v * b, a.d * b}; }
__device__ __forceinline__ Dual operator*(float a, Dual b) { return {a * b.v, a * b.d}; }
__device__ __forceinline__ Dual operator+(float a, Dual b) { return {a + b.v, b.d}; }
__device__ __forceinline__ Dual operator-(float a, Dual b) { return {a - b.v, -b.d}; }

__device__ __forceinline__ float val(float x) { return x; }
__device__ __forceinline__ float val(Dual x) { return x.v; }
__device__ __forceinline__ float f_rsq(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ Dual f_rsq(Dual x) {
  const float r = __builtin_amdgcn_rsqf(x.v);
  return {r, -0.5f * r * r * r * x.d};
}
__device__ __forceinline__ float f_exp2(float x) { return __builtin_amdgcn_exp2f(x); }
__device__ __forceinline__ Dual f_exp2(Dual x) {
  const float e = __builtin_amdgcn_exp2f(x.v);
  return {e, e * 0.69314718055994531f * x.d};
}
__device__ __forceinline__ float f_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ Dual f_rcp(Dual x) {
  const float r = __builtin_amdgcn_rcpf(x.v);
  return {r, -r * r * x.d};
}
__device__ __forceinline__ float f_pos(float x) { return fmaxf(x, 0.0f); }
__device__ __forceinline__ Dual f_pos(Dual x) { return x.v > 0.0f ? x : Dual{0.0f, 0.0f}; }

// The pair functions per unit coefficient: u = qq uc + C12 u12 - C6 u6 and
// the scalar force -du/dr / r = qq hc + C12 f12 - C6 f6 (qq = f q_i q_j,
// C12 = 4 eps sigma^12, C6 = 4 eps sigma^6), with the modifier of the pair
// kernel.  erfc(x) = t exp(-x^2) P(t), t = 1/(1 + 0.475 x), P of degree 6
// fitted by weighted least squares on the relative error over [0, 3.6]
// (this project's own fit, below 1e-7 relative in exact arithmetic; the
// fit is tests/golden/erfc_fit.py), the erfc_poly of pairs.hip.
template <typename T>
struct PairParts {
  T uc, hc, u12, f12, u6, f6;
};

template <typename T>
__device__ __forceinline__ T erfc_from_exp_t(T x, T expmx2) {
  const T t = f_rcp(0.475f * x + 1.0f);
  T pol = 0.09289151430130005f * t + (-0.4289827346801758f);
  pol = pol * t + 0.6025476455688477f;
  pol = pol * t + (-0.1264616847038269f);
  pol = pol * t + 0.3417808711528778f;
  pol = pol * t + 0.24873463809490204f;
  pol = pol * t + 0.26948970556259155f;
  return t * expmx2 * pol;
}

template <int MOD, bool COUL, typename T>
__device__ __forceinline__ PairParts<T> pair_parts(T r2, const AdjParams &p) {
  PairParts<T> o;
  const T rinv = f_rsq(r2);
  const T rinv2 = rinv * rinv;
  if (COUL) {
    const T br = p.beta * (r2 * rinv);
    const T expt = f_exp2(p.mb2l2e * r2);
    const T erfcv = erfc_from_exp_t(br, expt);
    o.uc = erfcv * rinv - p.shift_c;
    o.hc = (erfcv * rinv + p.tbs * expt) * rinv2;
  } else {
    o.uc = o.hc = T{};
  }
  const T rinv6 = rinv2 * rinv2 * rinv2;
  o.u12 = rinv6 * rinv6;
  o.f12 = 12.0f * o.u12 * rinv2;
  o.u6 = rinv6;
  o.f6 = 6.0f * rinv6 * rinv2;
  if (MOD == kLJPotentialShift) {
    o.u12 = o.u12 - p.rc12;
    o.u6 = o.u6 - p.rc6;
  } else if (MOD == kLJForceSwitch) {
    const T u = f_pos(r2 * rinv - p.rs);
    const T u2 = u * u, u3 = u2 * u;
    o.f12 = o.f12 + (p.fa12 * u2 + p.fb12 * u3) * rinv;
    o.f6 = o.f6 + (p.fa6 * u2 + p.fb6 * u3) * rinv;
    o.u12 = o.u12 - (p.fc12 + u3 * (p.fa12 * (1.0f / 3.0f) + p.fb12 * 0.25f * u));
    o.u6 = o.u6 - (p.fc6 + u3 * (p.fa6 * (1.0f / 3.0f) + p.fb6 * 0.25f * u));
  } else if (MOD == kLJPotentialSwitch) {
    const T r = r2 * rinv;
    if (val(r) > p.rs) {
      const T t = (r - p.rs) * (1.0f / p.width);
      const T t2 = t * t, t3 = t2 * t;
      const T S = 1.0f + t3 * (-10.0f + t * (15.0f - 6.0f * t));
      const T dS = ((-30.0f) * t2 + 60.0f * t3 - 30.0f * t3 * t) * (1.0f / p.width);
      o.f12 = o.f12 * S - o.u12 * dS * rinv;
      o.f6 = o.f6 * S - o.u6 * dS * rinv;
      o.u12 = o.u12 * S;
      o.u6 = o.u6 * S;
    }
  }
  return o;
}

// ---- cotangent to place order -------------------------------------------
//
// One block per frame.  mode 0: the cotangent cin (atom order, 3n) and the
// energy cotangent meta[1]; mode 1 (force matching): the cotangent
// -2 w (F - F_ref) from the force words F and F_ref = cin, with the loss
// sum |F - F_ref|^2 into meta[2].  c = the least power of two >= every
// |component| (1 if all are zero) goes to meta[0], g / c to meta[1], and
// cot[place] = (cotangent / c, 0): exact scalings.
constexpr double kForceMatchWeight = -2.0;

constexpr int kCotThreads = 1024;
__global__ void __launch_bounds__(kCotThreads) cot_prepare(int mode, const int *__restrict__ counts,
                                                  const int *__restrict__ atom_at,
                                                  const long long *__restrict__ F,
                                                  const double *__restrict__ cin,
                                                  double *__restrict__ meta,
                                                  float4 *__restrict__ cot, int max_places,
                                                  size_t fs) {
  counts = frame_at(counts, fs);
  atom_at = frame_at(atom_at, fs);
  F = frame_at(F, fs);
  cin = frame_at(cin, fs);
  meta = frame_at(meta, fs);
  cot = frame_at(cot, fs);
  __shared__ double s_max[kCotThreads], s_loss[kCotThreads];
  const int np = counts[0];
  double mx = 0.0, loss = 0.0;
  for (int pl = threadIdx.x; pl < np; pl += blockDim.x) {
    const int a = atom_at[pl];
    if (a < 0) continue;
    for (int k = 0; k < 3; ++k) {
      double c;
      if (mode == 1) {
        const double r = (double)F[3 * pl + k] * kForceUnscale - cin[3 * a + k];
        loss += r * r;
        c = kForceMatchWeight * r;
      } else {
        c = cin[3 * a + k];
      }
      mx = fmax(mx, fabs(c));
    }
  }
  s_max[threadIdx.x] = mx;
  s_loss[threadIdx.x] = loss;
  __syncthreads();
  for (int s = kCotThreads / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      s_max[threadIdx.x] = fmax(s_max[threadIdx.x], s_max[threadIdx.x + s]);
      s_loss[threadIdx.x] += s_loss[threadIdx.x + s];
    }
    __syncthreads();
  }
  const double g = mode == 1 ? 0.0 : meta[1];
  const double m = fmax(s_max[0], fabs(g));
  int e = 0;
  if (m > 0.0) (void)frexp(m, &e);  // m = f 2^e, f in [0.5, 1): 2^e >= m
  const double c = ldexp(1.0, e), inv = ldexp(1.0, -e);
  __syncthreads();  // every thread has read meta[1]
  if (threadIdx.x == 0) {
    meta[0] = c;
    meta[1] = g * inv;
    meta[2] = s_loss[0];
  }
  const int cap = max_places;
  for (int pl = threadIdx.x; pl < cap; pl += blockDim.x) {
    const int a = pl < np ? atom_at[pl] : -1;
    float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (a >= 0) {
      double c3[3];
      for (int k = 0; k < 3; ++k)
        c3[k] = mode == 1
                    ? kForceMatchWeight * ((double)F[3 * pl + k] * kForceUnscale - cin[3 * a + k])
                    : cin[3 * a + k];
      v = make_float4((float)(c3[0] * inv), (float)(c3[1] * inv), (float)(c3[2] * inv), 0.0f);
    }
    cot[pl] = v;
  }
}

// ---- the excluded and 1-4 pairs, f64, thread per pair --------------------
//
// theta: only the excluded-pair Ewald correction -f q_i q_j erf(beta r)/r
// depends on theta (the 1-4 parameters are given per pair and are not
// tunables here).  Hessian: the excluded correction and the 1-4 LJ and
// Coulomb terms, u'' r^ r^T + (u'/r)(1 - r^ r^T) applied to v_i - v_j.
struct SpecialAdjArgs {
  int npairs;
  const int2 *pairs;
  const int *is14;
  const double *p14_sigma, *p14_eps, *p14_qq;
  const double *x, *q;        // x per frame (atom order)
  const int *place_of;        // per frame
  double beta, coulomb, rc, rs;
  double box[3];
  int lj_modifier;
};

// u'(r) and u''(r) of a scaled 1-4 Lennard-Jones pair under the modifier
// (the f64 forms of lj_pair_f64 in pairs.hip, differentiated once more).
__device__ void lj14_d12(double sigma, double eps, double r, const SpecialAdjArgs &p,
                         double &d1, double &d2) {
  d1 = d2 = 0.0;
  if (eps == 0.0 || sigma == 0.0) return;
  const double e4 = 4.0 * eps, rinv = 1.0 / r, sr = sigma * rinv;
  const double s3 = sr * sr * sr, s6 = s3 * s3, s12 = s6 * s6;
  const double u0 = e4 * (s12 - s6);
  double u1 = e4 * (-12.0 * s12 + 6.0 * s6) * rinv;
  double u2 = e4 * (156.0 * s12 - 42.0 * s6) * rinv * rinv;
  if (p.lj_modifier == kLJPotentialSwitch && r > p.rs) {
    const double w = p.rc - p.rs, t = (r - p.rs) / w;
    const double S = 1.0 + t * t * t * (-10.0 + t * (15.0 - 6.0 * t));
    const double dS = (-30.0 * t * t + 60.0 * t * t * t - 30.0 * t * t * t * t) / w;
    const double ddS = (-60.0 * t + 180.0 * t * t - 120.0 * t * t * t) / (w * w);
    const double s1 = u1 * S + u0 * dS;
    u2 = u2 * S + 2.0 * u1 * dS + u0 * ddS;
    u1 = s1;
  } else if (p.lj_modifier == kLJForceSwitch) {
    const double rc = p.rc, rs = p.rs, w = rc - rs;
    const double uu = r > rs ? r - rs : 0.0;
    const double g3 = sigma * sigma * sigma, g6 = g3 * g3, g12 = g6 * g6;
    const double ns[2] = {12.0, 6.0};
    double dP[2], ddP[2];
    for (int k = 0; k < 2; ++k) {
      const double nn = ns[k];
      const double an = -nn * ((nn + 4.0) * rc - (nn + 1.0) * rs) / (pow(rc, nn + 2.0) * w * w);
      const double bn = nn * ((nn + 3.0) * rc - (nn + 1.0) * rs) / (pow(rc, nn + 2.0) * w * w * w);
      dP[k] = an * uu * uu + bn * uu * uu * uu;
      ddP[k] = 2.0 * an * uu + 3.0 * bn * uu * uu;
    }
    const double r6 = pow(rinv, 6.0), r12 = r6 * r6;
    u1 = e4 * (g12 * (-12.0 * r12 * rinv - dP[0]) - g6 * (-6.0 * r6 * rinv - dP[1]));
    u2 = e4 * (g12 * (156.0 * r12 * rinv * rinv - ddP[0]) -
               g6 * (42.0 * r6 * rinv * rinv - ddP[1]));
  }
  d1 = u1;
  d2 = u2;
}

// The virial sum_pairs (-u'/r) d (x) d of the excluded and 1-4 pairs, f64:
// xx yy zz xy xz yz of this block into vspec[6 block + c] (s_v: 256
// doubles of LDS; every thread of the block takes part).
__device__ __forceinline__ void special_virial(int block, const SpecialAdjArgs &sa,
                                              double *__restrict__ vspec, double *s_v) {
  const int t = block * blockDim.x + threadIdx.x;
  double v[6] = {0, 0, 0, 0, 0, 0};
  if (t < sa.npairs) {
    const int2 ij = sa.pairs[t];
    double d[3], r2 = 0.0;
    for (int k = 0; k < 3; ++k) {
      double dd = sa.x[3 * ij.x + k] - sa.x[3 * ij.y + k];
      dd -= sa.box[k] * rint(dd / sa.box[k]);
      d[k] = dd;
      r2 += dd * dd;
    }
    const double r = sqrt(r2);
    double dudr = 0.0;
    if (sa.beta > 0.0) {
      const double qq = sa.coulomb * sa.q[ij.x] * sa.q[ij.y];
      const double ef = erf(sa.beta * r), E = exp(-sa.beta * sa.beta * r2);
      dudr += -qq * (1.1283791670955126 * sa.beta * E / r - ef / r2);
    }
    if (sa.is14 && sa.is14[t] >= 0) {
      const int k14 = sa.is14[t];
      double l1, l2;
      lj14_d12(sa.p14_sigma[k14], sa.p14_eps[k14], r, sa, l1, l2);
      dudr += l1 - sa.p14_qq[k14] / r2;
    }
    const double fsc = -dudr / r;
    v[0] = fsc * d[0] * d[0];
    v[1] = fsc * d[1] * d[1];
    v[2] = fsc * d[2] * d[2];
    v[3] = fsc * d[0] * d[1];
    v[4] = fsc * d[0] * d[2];
    v[5] = fsc * d[1] * d[2];
  }
  for (int c = 0; c < 6; ++c) {
    __syncthreads();
    s_v[threadIdx.x] = v[c];
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
      if (threadIdx.x < s) s_v[threadIdx.x] += s_v[threadIdx.x + s];
      __syncthreads();
    }
    if (threadIdx.x == 0) vspec[6 * block + c] = s_v[0];
  }
}

template <int KIND>
__device__ __forceinline__ void special_adjoint(int block, const SpecialAdjArgs &sa,
                                                const float4 *__restrict__ cot, float gs,
                                                long long *__restrict__ A) {
  const int t = block * blockDim.x + threadIdx.x;
  if (t >= sa.npairs) return;
  const int2 ij = sa.pairs[t];
  double d[3], r2 = 0.0;
  for (int k = 0; k < 3; ++k) {
    double dd = sa.x[3 * ij.x + k] - sa.x[3 * ij.y + k];
    dd -= sa.box[k] * rint(dd / sa.box[k]);
    d[k] = dd;
    r2 += dd * dd;
  }
  const double r = sqrt(r2);
  const int pi = sa.place_of[ij.x], pj = sa.place_of[ij.y];
  const float4 ci = cot[pi], cj = cot[pj];
  const double dv[3] = {(double)ci.x - cj.x, (double)ci.y - cj.y, (double)ci.z - cj.z};
  const double s = d[0] * dv[0] + d[1] * dv[1] + d[2] * dv[2];
  unsigned long long *ai = reinterpret_cast<unsigned long long *>(A + 3 * pi);
  unsigned long long *aj = reinterpret_cast<unsigned long long *>(A + 3 * pj);
  const double tbs = 1.1283791670955126 * sa.beta;
  if (KIND == kAdjTheta) {
    if (sa.beta > 0.0) {
      const double ef = erf(sa.beta * r), E = exp(-sa.beta * sa.beta * r2);
      const double phi = -ef / r;                   // per unit f q_i q_j
      const double dphi_r = (ef / r - tbs * E) / r2;  // phi'(r) / r
      const double B = (double)gs * phi + dphi_r * s;
      const double wi = sa.coulomb * sa.q[ij.y] * B, wj = sa.coulomb * sa.q[ij.x] * B;
      atomicAdd(ai, (unsigned long long)__double2ll_rn(wi * kAdjScale));
      atomicAdd(aj, (unsigned long long)__double2ll_rn(wj * kAdjScale));
    }
    return;
  }
  double u1 = 0.0, u2 = 0.0;
  if (sa.beta > 0.0) {
    const double qq = sa.coulomb * sa.q[ij.x] * sa.q[ij.y];
    const double ef = erf(sa.beta * r), E = exp(-sa.beta * sa.beta * r2);
    u1 += -qq * (tbs * E / r - ef / r2);
    u2 += -qq * (-2.0 * sa.beta * sa.beta * tbs * E - 2.0 * tbs * E / r2 + 2.0 * ef / (r2 * r));
  }
  if (sa.is14 && sa.is14[t] >= 0) {
    const int k14 = sa.is14[t];
    double l1, l2;
    lj14_d12(sa.p14_sigma[k14], sa.p14_eps[k14], r, sa, l1, l2);
    const double qq14 = sa.p14_qq[k14];
    u1 += l1 - qq14 / r2;
    u2 += l2 + 2.0 * qq14 / (r2 * r);
  }
  const double a1 = u1 / r, a2 = (u2 - u1 / r) / r2 * s;
  for (int k = 0; k < 3; ++k) {
    const long long w = __double2ll_rn((a1 * dv[k] + a2 * d[k]) * kAdjScale);
    atomicAdd(ai + k, (unsigned long long)w);
    atomicAdd(aj + k, (unsigned long long)(-w));
  }
}

// ---- the adjoint pair loop ----------------------------------------------
//
// The traversal of nb_pairs (one wave per segment of a group row, lane l
// holding i atom l & 7 and j atom l >> 3 of each 8 x 8 tile, the j sums
// over the members by DPP and through LDS, the i sums kept in registers),
// with the pair terms of the products:
//   theta: s = r_ij . (c_i - c_j);  per atom three words
//     [0] dq:   f q_other (gs uc - hc s)
//     [1] S:    12 C12/sigma P12 + 6 C6/sigma P6      (both atoms)
//     [2] E:    C12 P12 + C6 P6                        (both atoms)
//     with P12 = gs u12 - f12 s, P6 = f6 s - gs u6 (the products per unit
//     C12 and C6); dU/dsigma_a = 1/2 sum_{i of type a} S_i and
//     dU/deps_a = sum_{i of type a} E_i / (2 eps_a) under Lorentz-Berthelot
//   Hessian: with fs = -u'/r and D = (dfs/dr)/r = 2 dfs/d(r^2),
//     H (v_i - v_j) = -fs dv - D (r_ij . dv) r_ij, added to i, taken from j.
//   virial (KIND kAdjVirial): W = sum_pairs fs r (x) r, f64 per lane, one
//     partial of 6 (xx yy zz xy xz yz) per (segment, group) wave into vpart
//     and per block of the special pairs into vspec.
//   type-pair entries (KIND kAdjPairTable): for the selected table entries
//     (a, b), sum over the pairs of those types of dl/dsigma_ab =
//     12 C12/sigma P12 + 6 C6/sigma P6 and dl/deps_ab = 4 sigma^12 P12 +
//     4 sigma^6 P6, one partial of 2 kMaxTablePairs per wave into vpart.
template <int MOD, bool COUL, int KIND>
__global__ void __launch_bounds__(256) nb_adjoint(
    const float4 *__restrict__ xq, const int *__restrict__ type_at,
    const uint32_t *__restrict__ real_mask, const PairEntry *__restrict__ rows,
    const int *__restrict__ row_count, const uint64_t *__restrict__ ex_masks, int row_cap,
    int seg_len, int max_groups, const int *__restrict__ counts_in, AdjTables tab,
    AdjParams p, const float4 *__restrict__ cot, const double *__restrict__ meta,
    long long *__restrict__ A, double *__restrict__ vpart, double *__restrict__ vspec,
    SpecialAdjArgs sa, size_t fs) {
  xq = frame_at(xq, fs);
  type_at = frame_at(type_at, fs);
  real_mask = frame_at(real_mask, fs);
  rows = frame_at(rows, fs);
  row_count = frame_at(row_count, fs);
  ex_masks = frame_at(ex_masks, fs);
  counts_in = frame_at(counts_in, fs);
  cot = frame_at(cot, fs);
  meta = frame_at(meta, fs);
  A = frame_at(A, fs);
  vpart = frame_at(vpart, fs);
  vspec = frame_at(vspec, fs);
  sa.x = frame_at(sa.x, fs);
  sa.place_of = frame_at(sa.place_of, fs);
  extern __shared__ float4 s_lj4[];  // ntypes^2 x (float4, float2, int) (launch size)
  __shared__ union {
    struct {
      float aj[4][kSegmentMax * kCluster * 3];
      int cj[4][kSegmentMax];
    } pr;
    double v[256];
  } sm;
  const int nt2 = p.ntypes * p.ntypes;
  float2 *s_unit = reinterpret_cast<float2 *>(s_lj4 + nt2);
  int *s_sel = reinterpret_cast<int *>(s_unit + nt2);
  for (int k = threadIdx.x; k < nt2; k += blockDim.x) {
    s_lj4[k] = tab.lj[k];
    if (KIND == kAdjPairTable) {
      s_unit[k] = tab.unit[k];
      s_sel[k] = tab.sel[k];
    }
  }
  __syncthreads();
  const float gs = (float)meta[1];
  if (blockIdx.y == 0) {
    if ((int)blockIdx.x * (int)blockDim.x < sa.npairs) {
      if (KIND == kAdjVirial)
        special_virial(blockIdx.x, sa, vspec, sm.v);
      else if (KIND == kAdjTheta || KIND == kAdjHessian)
        special_adjoint<KIND>(blockIdx.x, sa, cot, gs, A);
    }
    return;
  }
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int gi = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + wave);
  const int seg = blockIdx.y - 1;
  const int e0 = seg * seg_len;
  const int ngroups = counts_in[1] / kGroup;
  int n = gi < ngroups ? row_count[gi] - e0 : 0;
  n = __builtin_amdgcn_readfirstlane(n < seg_len ? n : seg_len);
  const size_t slot = (size_t)seg * max_groups + gi;
  if (n <= 0) {
    if ((KIND == kAdjVirial || KIND == kAdjPairTable) && gi < max_groups && lane < kPartStride)
      vpart[kPartStride * slot + lane] = 0.0;
    return;
  }
  const int a = lane & 7, b = lane >> 3;
  const int g0 = gi * kGroup;
  double vir[6] = {0, 0, 0, 0, 0, 0};
  float tps[kMaxTablePairs] = {0, 0, 0, 0}, tpe[kMaxTablePairs] = {0, 0, 0, 0};
  float4 xi[kGroup], ci[kGroup];
  int ti[kGroup];
  bool ireal[kGroup];
  float ai[kGroup][3];
#pragma unroll
  for (int m = 0; m < kGroup; ++m) {
    xi[m] = xq[(g0 + m) * kCluster + a];
    xi[m].w *= p.coulomb;  // f q_i
    ci[m] = cot[(g0 + m) * kCluster + a];
    ti[m] = type_at[(g0 + m) * kCluster + a] * p.ntypes;
    ireal[m] = (real_mask[g0 + m] >> a) & 1u;
    ai[m][0] = ai[m][1] = ai[m][2] = 0.0f;
  }
  const PairEntry *row = rows + (size_t)gi * row_cap + e0;
  PairEntry E = row[0];
  PairEntry E1 = row[n > 1 ? 1 : 0];
  float4 xj = xq[E.cj * kCluster + b];
  float4 cj = cot[E.cj * kCluster + b];
  int tj = type_at[E.cj * kCluster + b];
  for (int e = 0; e < n; ++e) {
    const PairEntry E2 = row[e + 2 < n ? e + 2 : n - 1];
    const float4 xj1 = xq[E1.cj * kCluster + b];
    const float4 cj1 = cot[E1.cj * kCluster + b];
    const int tj1 = type_at[E1.cj * kCluster + b];
    const uint32_t info = E.info;
    float sx, sy, sz;
    shift_vector(info & 63u, p.box, sx, sy, sz);
    const float xjx = xj.x + sx, xjy = xj.y + sy, xjz = xj.z + sz;
    const bool jreal = (info >> (16 + b)) & 1u;
    const float qfj = p.coulomb * xj.w;
    float aj0 = 0.0f, aj1 = 0.0f, aj2 = 0.0f;
#pragma unroll
    for (int m = 0; m < kGroup; ++m) {
      if (!((info >> (8 + m)) & 1u)) continue;  // uniform
      const float dx = xi[m].x - xjx, dy = xi[m].y - xjy, dz = xi[m].z - xjz;
      const float r2 = dx * dx + dy * dy + dz * dz;
      bool on = E.ex >= 0 ? ((ex_masks[(size_t)E.ex * kGroup + m] >> lane) & 1ull)
                          : (ireal[m] && jreal);
      on = on && r2 < p.rc2;
      // The pair's arithmetic runs under the exec mask of the lanes that
      // are on (r2 inside rc, never zero); a member tile with no such pair
      // (about half of those of a list with a buffer) branches over it.
      if (!on) continue;
      const float4 T4 = s_lj4[ti[m] + tj];
      const float dvx = ci[m].x - cj.x, dvy = ci[m].y - cj.y, dvz = ci[m].z - cj.z;
      const float s = dx * dvx + dy * dvy + dz * dvz;
      if (KIND == kAdjPairTable) {
        const int sel = s_sel[ti[m] + tj];
        if (sel >= 0) {  // the pairs of a selected entry only (divergent)
          const PairParts<float> P = pair_parts<MOD, COUL, float>(r2, p);
          const float P12 = gs * P.u12 - P.f12 * s;
          const float P6 = P.f6 * s - gs * P.u6;
          const float2 U = s_unit[ti[m] + tj];
          const float Ss = T4.z * P12 + T4.w * P6;
          const float Eu = U.x * P12 + U.y * P6;
#pragma unroll
          for (int t = 0; t < kMaxTablePairs; ++t)
            if (sel == t) {
              tps[t] += Ss;
              tpe[t] += Eu;
            }
        }
      } else if (KIND == kAdjVirial) {
        const PairParts<float> P = pair_parts<MOD, COUL, float>(r2, p);
        float fsv = T4.x * P.f12 - T4.y * P.f6;
        if (COUL) fsv += xi[m].w * xj.w * P.hc;
        const double f = (double)fsv;
        vir[0] += f * (double)(dx * dx);
        vir[1] += f * (double)(dy * dy);
        vir[2] += f * (double)(dz * dz);
        vir[3] += f * (double)(dx * dy);
        vir[4] += f * (double)(dx * dz);
        vir[5] += f * (double)(dy * dz);
      } else if (KIND == kAdjTheta) {
        const PairParts<float> P = pair_parts<MOD, COUL, float>(r2, p);
        const float Bc = COUL ? gs * P.uc - P.hc * s : 0.0f;
        const float P12 = gs * P.u12 - P.f12 * s;
        const float P6 = P.f6 * s - gs * P.u6;
        const float Ss = T4.z * P12 + T4.w * P6;
        const float Ee = T4.x * P12 + T4.y * P6;
        const float B = Bc;
        ai[m][0] += qfj * B;
        ai[m][1] += Ss;
        ai[m][2] += Ee;
        aj0 += xi[m].w * B;
        aj1 += Ss;
        aj2 += Ee;
      } else {
        const PairParts<Dual> P = pair_parts<MOD, COUL, Dual>(Dual{r2, 1.0f}, p);
        const float qq = xi[m].w * xj.w;
        float fsv = T4.x * P.f12.v - T4.y * P.f6.v;
        float D = T4.x * P.f12.d - T4.y * P.f6.d;
        if (COUL) {
          fsv += qq * P.hc.v;
          D += qq * P.hc.d;
        }
        D *= 2.0f;
        const float hx = -fsv * dvx - D * s * dx;
        const float hy = -fsv * dvy - D * s * dy;
        const float hz = -fsv * dvz - D * s * dz;
        ai[m][0] += hx;
        ai[m][1] += hy;
        ai[m][2] += hz;
        aj0 -= hx;
        aj1 -= hy;
        aj2 -= hz;
      }
    }
    if (KIND == kAdjTheta || KIND == kAdjHessian) {
      const float g0s = sum8(aj0), g1s = sum8(aj1), g2s = sum8(aj2);
      if (a == 0) {
        float *dst = &sm.pr.aj[wave][(e * kCluster + b) * 3];
        dst[0] = g0s;
        dst[1] = g1s;
        dst[2] = g2s;
      }
      if (lane == 0) sm.pr.cj[wave][e] = E.cj;
    }
    E = E1;
    E1 = E2;
    xj = xj1;
    cj = cj1;
    tj = tj1;
  }
  if (KIND == kAdjVirial) {
#pragma unroll
    for (int c = 0; c < 6; ++c) {
#pragma unroll
      for (int s = 1; s < 64; s <<= 1) vir[c] += __shfl_xor(vir[c], s);
    }
    if (lane == 0)
      for (int c = 0; c < 6; ++c) vpart[kPartStride * slot + c] = vir[c];
    return;
  }
  if (KIND == kAdjPairTable) {
    double w[kPartStride];
#pragma unroll
    for (int t = 0; t < kMaxTablePairs; ++t) {
      w[2 * t] = tps[t];
      w[2 * t + 1] = tpe[t];
    }
#pragma unroll
    for (int c = 0; c < kPartStride; ++c) {
#pragma unroll
      for (int s = 1; s < 64; s <<= 1) w[c] += __shfl_xor(w[c], s);
    }
    if (lane == 0)
      for (int c = 0; c < kPartStride; ++c) vpart[kPartStride * slot + c] = w[c];
    return;
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  unsigned long long *u = reinterpret_cast<unsigned long long *>(A);
  for (int k = lane; k < n * kCluster * 3; k += 64) {
    const int e = k / (kCluster * 3), rem = k - e * (kCluster * 3);
    const long long w = to_fixed_adj(sm.pr.aj[wave][k]);
    if (w) atomicAdd(u + 3 * (sm.pr.cj[wave][e] * kCluster) + rem, (unsigned long long)w);
  }
#pragma unroll
  for (int m = 0; m < kGroup; ++m) {
#pragma unroll
    for (int s = 8; s < 64; s <<= 1) {
      ai[m][0] += __shfl_xor(ai[m][0], s);
      ai[m][1] += __shfl_xor(ai[m][1], s);
      ai[m][2] += __shfl_xor(ai[m][2], s);
    }
    if (b == 0 && ireal[m]) {
      unsigned long long *w = u + 3 * ((g0 + m) * kCluster + a);
      for (int k = 0; k < 3; ++k) {
        const long long v = to_fixed_adj(ai[m][k]);
        if (v) atomicAdd(w + k, (unsigned long long)v);
      }
    }
  }
}

// ---- PME: the dipole grid and the gathers ---------------------------------

// B-spline second derivatives in u at the stencil points (the second
// difference of the order n-2 weights: M_n'' = M_{n-2}(u) - 2 M_{n-2}(u-1)
// + M_{n-2}(u-2), PMECode.h's recursion run to order n-2).
template <int N>
__device__ __forceinline__ void bspline_d2(float w, float *d2) {
  float A[N] = {0};
  A[0] = w;
  A[1] = 1.0f - w;
#pragma unroll
  for (int nn = 3; nn <= N - 2; ++nn) {
    float B[N];
#pragma unroll
    for (int k = 0; k < N; ++k) B[k] = A[k];
#pragma unroll
    for (int k = 0; k < nn; ++k) {
      const float prev = k > 0 ? B[k - 1] : 0.0f;
      A[k] = ((w + (float)k) * B[k] + ((float)nn - w - (float)k) * prev) / (float)(nn - 1);
    }
  }
  // C[t'] = M_{n-2}(w + n - 3 - t') = A[n - 3 - t'], t' = 0 .. n-3
#pragma unroll
  for (int t = 0; t < N; ++t) {
    const float c0 = (t <= N - 3) ? A[N - 3 - t] : 0.0f;
    const float c1 = (t >= 1 && t - 1 <= N - 3) ? A[N - 2 - t] : 0.0f;
    const float c2 = (t >= 2) ? A[N - 1 - t] : 0.0f;
    d2[t] = c2 - 2.0f * c1 + c0;
  }
}

// The part of one atom's dipole weights q c . grad W in the brick, added to
// the brick's accumulators (c in grid units: c_a K_a / L_a).
template <int N, int BX, int BY>
__device__ __forceinline__ void spread_dipole_atom(double4 pq, float4 c, const PmeParams &pp,
                                                   int bx0, int by0, unsigned long long *acc) {
  const int K0 = pp.K[0], K1 = pp.K[1], K2 = pp.K[2];
  int f[3];
  float w[3];
  pme_place(pq.x, pp.box[0], pp.inv_box[0], K0, N, f[0], w[0]);
  pme_place(pq.y, pp.box[1], pp.inv_box[1], K1, N, f[1], w[1]);
  pme_place(pq.z, pp.box[2], pp.inv_box[2], K2, N, f[2], w[2]);
  float M[3][N], dM[3][N];
#pragma unroll
  for (int d = 0; d < 3; ++d) bspline<float>(w[d], N, M[d], dM[d]);
  const float q = (float)pq.w;
  const float cx = q * c.x * (float)(K0 * pp.inv_box[0]);
  const float cy = q * c.y * (float)(K1 * pp.inv_box[1]);
  const float cz = q * c.z * (float)(K2 * pp.inv_box[2]);
  // loops over the stencil (compile-time indices into the weights), the
  // points outside the brick skipped
#pragma unroll
  for (int i = 0; i < N; ++i) {
    int gx = f[0] + i;
    if (gx >= K0) gx -= K0;
    const int tx = gx - bx0;
    if (tx < 0 || tx >= BX) continue;
    const float mx = M[0][i], dmx = dM[0][i];
#pragma unroll
    for (int j = 0; j < N; ++j) {
      int gy = f[1] + j;
      if (gy >= K1) gy -= K1;
      const int ty = gy - by0;
      if (ty < 0 || ty >= BY) continue;
      const float my = M[1][j], dmy = dM[1][j];
      const float P = cx * dmx * my + cy * mx * dmy;
      const float Qz = cz * mx * my;
      unsigned long long *row = acc + (tx * BY + ty) * K2;
#pragma unroll
      for (int k = 0; k < N; ++k) {
        int z = f[2] + k;
        if (z >= K2) z -= K2;
        const long long v = __double2ll_rn((double)(P * M[2][k] + Qz * dM[2][k]) * kGridScale);
        atomicAdd(row + z, (unsigned long long)v);
      }
    }
  }
}

// The dipole grid of the cotangent, spread by the bricks of pme_bin's lists
// and transformed along z like pme_spread_rows.
template <int N, int TL>
__global__ void __launch_bounds__(256) pme_spread_dipole(
    const double4 *__restrict__ xpq, const float4 *__restrict__ cot, PmeParams pp,
    BrickParams bp, cpx<float> *__restrict__ out, FftAxis az, size_t fs) {
  constexpr int R = BrickShape<TL>::R, BX = BrickShape<TL>::BX, BY = BrickShape<TL>::BY;
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  xpq = frame_at(xpq, fs);
  cot = frame_at(cot, fs);
  bp.count = frame_at(bp.count, fs);
  bp.list = frame_at(bp.list, fs);
  out = frame_at(out, fs);
  const int K2 = pp.K[2];
  unsigned long long *acc = reinterpret_cast<unsigned long long *>(smem);
  cpx<float> *a = reinterpret_cast<cpx<float> *>(acc + (size_t)R * K2);
  cpx<float> *b = a + TL * K2;
  cpx<float> *tw = b + TL * K2;
  const int brick = blockIdx.x;
  const int bx0 = (brick % bp.nbx) * BX, by0 = (brick / bp.nbx) * BY;
  for (int w = threadIdx.x; w < R * K2; w += blockDim.x) acc[w] = 0ull;
  for (int w = threadIdx.x; w < K2; w += blockDim.x)
    tw[w] = static_cast<const cpx<float> *>(az.tw)[w];
  int nat = bp.count[brick];
  nat = nat < bp.cap ? nat : bp.cap;
  const int *lst = bp.list + (size_t)brick * bp.cap;
  __syncthreads();
  for (int i = threadIdx.x; i < nat; i += blockDim.x) {
    const int pl = lst[i];
    spread_dipole_atom<N, BX, BY>(xpq[pl], cot[pl], pp, bx0, by0, acc);
  }
  __syncthreads();
  brick_forward_z<TL>(acc, a, b, tw, az, pp, bp, bx0, by0, out);
}

// Gathers of the potential phi (the forward's convolved grid) and psi (the
// convolved dipole grid) at one atom, with q, the scaled cotangent c and
// gs = g / c:
//   theta:   dq += c . grad phi + gs phi + psi        (word 0)
//   Hessian: (H v)_a += q [sum_b (hess phi)_ab c_b + (grad psi)_a]
// (gradients in x: grid-unit derivatives times K_a / L_a).
template <int N, int TL, int BX, int BY, int KIND>
__device__ __forceinline__ void gather_adj_atom(double4 pq, float4 c, float gs, int place,
                                                const PmeParams &pp, int bx0, int by0,
                                                const cpx<float> *phi, const cpx<float> *psi,
                                                long long *__restrict__ A) {
  const int K0 = pp.K[0], K1 = pp.K[1], K2 = pp.K[2];
  int f[3];
  float w[3];
  pme_place(pq.x, pp.box[0], pp.inv_box[0], K0, N, f[0], w[0]);
  pme_place(pq.y, pp.box[1], pp.inv_box[1], K1, N, f[1], w[1]);
  pme_place(pq.z, pp.box[2], pp.inv_box[2], K2, N, f[2], w[2]);
  float M[3][N], dM[3][N], d2M[3][N];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    bspline<float>(w[d], N, M[d], dM[d]);
    if (KIND == kAdjHessian) bspline_d2<N>(w[d], d2M[d]);
  }
  // theta: phi value, grad; psi value.  Hessian: phi hessian, psi grad.
  float v0 = 0.0f, g0 = 0.0f, g1 = 0.0f, g2 = 0.0f, p0 = 0.0f;
  float hxx = 0.0f, hyy = 0.0f, hzz = 0.0f, hxy = 0.0f, hxz = 0.0f, hyz = 0.0f;
#pragma unroll
  for (int i = 0; i < N; ++i) {
    int gx = f[0] + i;
    if (gx >= K0) gx -= K0;
    const int tx = gx - bx0;
    if (tx < 0 || tx >= BX) continue;
    const float mx = M[0][i], dmx = dM[0][i];
    const float d2x = KIND == kAdjHessian ? d2M[0][i] : 0.0f;
#pragma unroll
    for (int j = 0; j < N; ++j) {
      int gy = f[1] + j;
      if (gy >= K1) gy -= K1;
      const int ty = gy - by0;
      if (ty < 0 || ty >= BY) continue;
      const float my = M[1][j], dmy = dM[1][j];
      const float d2y = KIND == kAdjHessian ? d2M[1][j] : 0.0f;
      const int r = tx * BY + ty;
      float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f, q0 = 0.0f, q1 = 0.0f;
#pragma unroll
      for (int k = 0; k < N; ++k) {
        int z = f[2] + k;
        if (z >= K2) z -= K2;
        const cpx<float> cp = phi[z * TL + (r >> 1)];
        const cpx<float> cs = psi[z * TL + (r >> 1)];
        const float pv = (r & 1) ? cp.y : cp.x;
        const float sv = (r & 1) ? cs.y : cs.x;
        s0 += pv * M[2][k];
        s1 += pv * dM[2][k];
        q0 += sv * M[2][k];
        if (KIND == kAdjHessian) {
          s2 += pv * d2M[2][k];
          q1 += sv * dM[2][k];
        }
      }
      if (KIND == kAdjTheta) {
        v0 += mx * my * s0;
        g0 += dmx * my * s0;
        g1 += mx * dmy * s0;
        g2 += mx * my * s1;
        p0 += mx * my * q0;
      } else {
        hxx += d2x * my * s0;
        hyy += mx * d2y * s0;
        hzz += mx * my * s2;
        hxy += dmx * dmy * s0;
        hxz += dmx * my * s1;
        hyz += mx * dmy * s1;
        g0 += dmx * my * q0;
        g1 += mx * dmy * q0;
        g2 += mx * my * q1;
      }
    }
  }
  const float sx = (float)(K0 * pp.inv_box[0]), sy = (float)(K1 * pp.inv_box[1]),
              sz = (float)(K2 * pp.inv_box[2]);
  unsigned long long *u = reinterpret_cast<unsigned long long *>(A + 3 * place);
  if (KIND == kAdjTheta) {
    const float dq = c.x * sx * g0 + c.y * sy * g1 + c.z * sz * g2 + gs * v0 + p0;
    atomicAdd(u, (unsigned long long)to_fixed_adj(dq));
  } else {
    const float q = (float)pq.w;
    const float cx = c.x * sx, cy = c.y * sy, cz = c.z * sz;
    const float ox = q * sx * (hxx * cx + hxy * cy + hxz * cz + g0);
    const float oy = q * sy * (hxy * cx + hyy * cy + hyz * cz + g1);
    const float oz = q * sz * (hxz * cx + hyz * cy + hzz * cz + g2);
    atomicAdd(u + 0, (unsigned long long)to_fixed_adj(ox));
    atomicAdd(u + 1, (unsigned long long)to_fixed_adj(oy));
    atomicAdd(u + 2, (unsigned long long)to_fixed_adj(oz));
  }
}

template <int N, int TL, int KIND>
__global__ void __launch_bounds__(256) pme_gather_adj(
    const double4 *__restrict__ xpq, const float4 *__restrict__ cot,
    const double *__restrict__ meta, PmeParams pp, BrickParams bp,
    const cpx<float> *__restrict__ in_phi, const cpx<float> *__restrict__ in_psi, FftAxis az,
    long long *__restrict__ A, size_t fs) {
  constexpr int BX = BrickShape<TL>::BX, BY = BrickShape<TL>::BY;
  extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
  xpq = frame_at(xpq, fs);
  cot = frame_at(cot, fs);
  meta = frame_at(meta, fs);
  bp.count = frame_at(bp.count, fs);
  bp.list = frame_at(bp.list, fs);
  in_phi = frame_at(in_phi, fs);
  in_psi = frame_at(in_psi, fs);
  A = frame_at(A, fs);
  const int K2 = pp.K[2];
  cpx<float> *a1 = reinterpret_cast<cpx<float> *>(smem);
  cpx<float> *b1 = a1 + TL * K2;
  cpx<float> *a2 = b1 + TL * K2;
  cpx<float> *b2 = a2 + TL * K2;
  cpx<float> *tw = b2 + TL * K2;
  const int brick = blockIdx.x;
  const int bx0 = (brick % bp.nbx) * BX, by0 = (brick / bp.nbx) * BY;
  for (int w = threadIdx.x; w < K2; w += blockDim.x)
    tw[w] = static_cast<const cpx<float> *>(az.tw)[w];
  int nat = bp.count[brick];
  nat = nat < bp.cap ? nat : bp.cap;
  const int *lst = bp.list + (size_t)brick * bp.cap;
  const float gs = (float)meta[1];
  const cpx<float> *phi = brick_backward_z<TL>(in_phi, a1, b1, tw, az, pp, bp, bx0, by0);
  __syncthreads();
  const cpx<float> *psi = brick_backward_z<TL>(in_psi, a2, b2, tw, az, pp, bp, bx0, by0);
  __syncthreads();
  for (int i = threadIdx.x; i < nat; i += blockDim.x) {
    const int pl = lst[i];
    gather_adj_atom<N, TL, BX, BY, KIND>(xpq[pl], cot[pl], gs, pl, pp, bx0, by0, phi, psi, A);
  }
  __syncthreads();  // every thread has read the count
  if (threadIdx.x == 0) bp.count[brick] = 0;
}

// ---- frame arrays in and out ---------------------------------------------

// Between a contiguous array of `count` doubles per frame and the per-frame
// buffer (fs bytes apart): to_frames != 0 copies contiguous -> frames.
__global__ void frames_copy(const double *__restrict__ src, double *__restrict__ dst, int count,
                            int to_frames, size_t fs) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= count) return;
  const size_t f = blockIdx.z;
  if (to_frames)
    frame_at(dst, fs)[j] = src[f * count + j];
  else
    dst[f * count + j] = frame_at(src, fs)[j];
}

// The small results of a batch gathered for one read by the host:
// loss[f] = meta_f[2] (the loss of frame f) and flag_or[k] = the OR over the
// frames of flag k (as a double).
// One block of 1024 threads: thread (f, k) reads flag k of frame f.
__global__ void __launch_bounds__(1024) frames_collect(int frames, const double *__restrict__ meta,
                                                       const int *__restrict__ flags,
                                                       double *__restrict__ loss,
                                                       double *__restrict__ flag_or, size_t fs) {
  __shared__ int s_or[kNumFlags];
  const int t = threadIdx.x;
  if (t < kNumFlags) s_or[t] = 0;
  __syncthreads();
  const char *mb = reinterpret_cast<const char *>(meta);
  const char *fb = reinterpret_cast<const char *>(flags);
  for (int f = t; f < frames; f += blockDim.x)
    loss[f] = reinterpret_cast<const double *>(mb + f * fs)[2];
  for (int w = t; w < frames * kNumFlags; w += blockDim.x) {
    const int f = w / kNumFlags, k = w % kNumFlags;
    const int v = reinterpret_cast<const int *>(fb + f * fs)[k];
    if (v) atomicOr(&s_or[k], v);
  }
  __syncthreads();
  if (t < kNumFlags) flag_or[t] = (double)s_or[t];
}

// ---- read-out -------------------------------------------------------------

// Words of the products to f64 in atom order, times the frame's scale c.
__global__ void adj_to_atoms(int n, const int *__restrict__ place_of,
                             const long long *__restrict__ A, const double *__restrict__ meta,
                             double *__restrict__ out, size_t fs) {
  place_of = frame_at(place_of, fs);
  A = frame_at(A, fs);
  meta = frame_at(meta, fs);
  out = frame_at(out, fs);
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double c = meta[0] * kAdjUnscale;
  const int pl = place_of[i];
  for (int k = 0; k < 3; ++k) out[3 * i + k] = (double)A[3 * pl + k] * c;
}

// sum[j] = sum over the frames, in frame order, of the per-frame arrays
// (count values each, fs bytes apart).
__global__ void frames_sum(int frames, int count, const double *__restrict__ in,
                           double *__restrict__ sum, size_t fs) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= count) return;
  double s = 0.0;
  const char *base = reinterpret_cast<const char *>(in);
  // the loads of 8 frames issued together, the adds in frame order
#pragma unroll 8
  for (int f = 0; f < frames; ++f) s += reinterpret_cast<const double *>(base + f * fs)[j];
  sum[j] = s;
}

}  // namespace mdhip
