// integrate.hip — velocity Verlet with SETTLE waters, in f64.
//
// Reference semantics: dyn.kick / dyn.drift (lib/Driver/DynamicsIR.cpp:102-112)
// fused as the integration run of MDExecKernels.cpp:1121 (08-gpu.md:112-121):
//   half kick, drift, SETTLE positions (ConstraintIR.cpp:473-593, the method
//   of Miyamoto & Kollman, J. Comput. Chem. 13, 952 (1992)), v += dx_c/dt,
//   the validity test |x - x_ref|^2 > (skin/2)^2 (ExposeValidity.cpp:23);
//   then after the forces: half kick and the velocity constraint
//   (ConstraintIR.cpp:392-471).
// A unit is a SETTLE water (3 consecutive atoms), a constraint cluster (an
// atom and up to four atoms bonded to it at fixed lengths, consecutive,
// M-SHAKE as ConstraintIR.cpp:595-747 takes it) or a free atom; one thread
// owns a unit, reads its force words and clears them for the next step.
// Blocks of 64 threads: a few thousand units still spread over every CU.
#include "md_common.h"

namespace mdhip {

// The integrator's f64 arithmetic is rounded operation by operation as
// written (no contraction into fused multiply-adds): a step then gives the
// same bits whether its two halves run in one launch (kBetween) or two,
// and follows the oracle's integration, which rounds each operation.
#pragma clang fp contract(off)

struct IntParams {
  double dt;
  double skin_half2;     // (outer list buffer / 2)^2: rebuild beyond it
  double prune_half2;    // (pruned list buffer / 2)^2: prune again beyond it
  const double *x_prune; // positions at the last prune (atom order), or null
  double half_box[3];
  double m_o, m_h;
  double ra, rb, rc;   // SETTLE canonical triangle
  double d_oh, d_hh;
  // Langevin dynamics (langevin != 0): the g-BAOAB step with the O step
  // v <- decay v + spread / sqrt(m) R, decay = exp(-gamma dt) and
  // spread = sqrt((1 - decay^2) kB T); R from langevin_normals(seed, step,
  // atom).  step[0] is the step number the next first half takes (the
  // launch's last block advances it), step[1] counts finished blocks.
  int langevin;
  double decay, spread;
  unsigned long long seed;
  long long *step;
  // constraint clusters: each member's bond length to its centre, by atom,
  // and the inverse masses (the descent step of a minimisation)
  const double *clen;
  const double *inv_mass;
};

// A unit's code (unit_size[u]): its atom count in bits 0-7, kUnitCluster
// for a constraint cluster, the cluster centre's offset from the unit's
// first atom in bits 12-15.
constexpr int kUnitCluster = 0x100;
constexpr int kUnitCentreShift = 12;

__device__ __forceinline__ void cross(const double *a, const double *b, double *c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}
__device__ __forceinline__ double dot(const double *a, const double *b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
__device__ __forceinline__ void unit(double *a) {
  const double s = rsqrt(dot(a, a));
  a[0] *= s;
  a[1] *= s;
  a[2] *= s;
}

// SETTLE of one water: old positions A0 B0 C0 (satisfying the constraints),
// new unconstrained A1 B1 C1 -> constrained, in place.
__device__ void settle_positions(const double *A0, const double *B0, const double *C0,
                                 double *A1, double *B1, double *C1,
                                 const IntParams &p) {
  const double total = p.m_o + 2.0 * p.m_h;
  double b0[3], c0[3], b1r[3], c1r[3];
  for (int k = 0; k < 3; ++k) {
    b0[k] = B0[k] - A0[k];
    c0[k] = C0[k] - A0[k];
    b1r[k] = B1[k] - A1[k];
    c1r[k] = C1[k] - A1[k];
  }
  double com[3], a1[3], b1[3], c1[3];
  for (int k = 0; k < 3; ++k) {
    com[k] = p.m_h / total * (b1r[k] + c1r[k]);
    a1[k] = -com[k];
    b1[k] = b1r[k] - com[k];
    c1[k] = c1r[k] - com[k];
  }
  double ez[3], ex[3], ey[3];
  cross(b0, c0, ez);
  unit(ez);
  cross(a1, ez, ex);
  unit(ex);
  cross(ez, ex, ey);
  const double xb0 = dot(b0, ex), yb0 = dot(b0, ey);
  const double xc0 = dot(c0, ex), yc0 = dot(c0, ey);
  const double za1 = dot(a1, ez);
  const double xb1 = dot(b1, ex), yb1 = dot(b1, ey), zb1 = dot(b1, ez);
  const double xc1 = dot(c1, ex), yc1 = dot(c1, ey), zc1 = dot(c1, ez);
  const double sinphi = za1 / p.ra;
  const double cosphi = sqrt(1.0 - sinphi * sinphi);
  const double sinpsi = (zb1 - zc1) / (2.0 * p.rc * cosphi);
  const double cospsi = sqrt(1.0 - sinpsi * sinpsi);
  const double ya2 = p.ra * cosphi;
  const double xb2 = -p.rc * cospsi;
  const double t1 = -p.rb * cosphi, t2 = p.rc * sinpsi * sinphi;
  const double yb2 = t1 - t2, yc2 = t1 + t2;
  const double alpha = xb2 * (xb0 - xc0) + yb0 * yb2 + yc0 * yc2;
  const double beta = xb2 * (yc0 - yb0) + xb0 * yb2 + xc0 * yc2;
  const double gamma = (xb0 * yb1 - xb1 * yb0) + (xc0 * yc1 - xc1 * yc0);
  const double both = alpha * alpha + beta * beta;
  const double sintheta = (alpha * gamma - beta * sqrt(both - gamma * gamma)) / both;
  const double costheta = sqrt(1.0 - sintheta * sintheta);
  const double xa3 = -ya2 * sintheta, ya3 = ya2 * costheta;
  const double xb3 = xb2 * costheta - yb2 * sintheta;
  const double yb3 = xb2 * sintheta + yb2 * costheta;
  const double xc3 = -xb2 * costheta - yc2 * sintheta;
  const double yc3 = -xb2 * sintheta + yc2 * costheta;
  double base[3] = {A1[0], A1[1], A1[2]};
  for (int k = 0; k < 3; ++k) {
    A1[k] = base[k] + xa3 * ex[k] + ya3 * ey[k] + za1 * ez[k] + com[k];
    B1[k] = base[k] + xb3 * ex[k] + yb3 * ey[k] + zb1 * ez[k] + com[k];
    C1[k] = base[k] + xc3 * ex[k] + yc3 * ey[k] + zc1 * ez[k] + com[k];
  }
}

__device__ __forceinline__ double det3(double m00, double m01, double m02, double m10,
                                       double m11, double m12, double m20, double m21,
                                       double m22) {
  return m00 * (m11 * m22 - m12 * m21) - m01 * (m10 * m22 - m12 * m20) +
         m02 * (m10 * m21 - m11 * m20);
}

// Impulses along the three bonds that zero the bond-relative velocities.
__device__ void settle_velocities(const double *xa, const double *xb, const double *xc,
                                  double *va, double *vb, double *vc,
                                  const IntParams &p) {
  const double io = 1.0 / p.m_o, ih = 1.0 / p.m_h;
  double e0[3], e1[3], e2[3], dvb[3], dvc[3], dvcb[3];
  for (int k = 0; k < 3; ++k) {
    e0[k] = xb[k] - xa[k];
    e1[k] = xc[k] - xa[k];
    e2[k] = xc[k] - xb[k];
    dvb[k] = vb[k] - va[k];
    dvc[k] = vc[k] - va[k];
    dvcb[k] = vc[k] - vb[k];
  }
  unit(e0);
  unit(e1);
  unit(e2);
  const double r0 = -dot(e0, dvb), r1 = -dot(e1, dvc), r2 = -dot(e2, dvcb);
  const double a00 = io + ih, a11 = io + ih, a22 = 2.0 * ih;
  const double a01 = dot(e0, e1) * io;
  const double a02 = -dot(e0, e2) * ih;
  const double a12 = dot(e1, e2) * ih;
  const double det = det3(a00, a01, a02, a01, a11, a12, a02, a12, a22);
  const double s0 = det3(r0, a01, a02, r1, a11, a12, r2, a12, a22) / det;
  const double s1 = det3(a00, r0, a02, a01, r1, a12, a02, r2, a22) / det;
  const double s2 = det3(a00, a01, r0, a01, a11, r1, a02, a12, r2) / det;
  for (int k = 0; k < 3; ++k) {
    const double i0 = s0 * e0[k], i1 = s1 * e1[k], i2 = s2 * e2[k];
    va[k] += -io * (i0 + i1);
    vb[k] += ih * (i0 - i2);
    vc[k] += ih * (i1 + i2);
  }
}

// The state of one unit of S atoms (a SETTLE water, S = 3, a constraint
// cluster, S = 2..5, with its centre at index 0, or a free atom, S = 1) in
// registers: S is a compile-time constant so that the arrays stay in
// registers.  at[m] is the atom of index m.
template <int S>
struct UnitState {
  int at[S];
  int pl[S];
  double im[S], x[S][3], v[S][3], f[S][3];
  double len[S];  // clusters: index m's bond length to the centre
};

// Linear equations A y = b of at most 4 unknowns by elimination (the
// cluster equations are diagonally dominant: no pivoting).
template <int M>
__device__ __forceinline__ void solve_small(double (&A)[M][M], double (&b)[M]) {
#pragma unroll
  for (int i = 0; i < M; ++i)
#pragma unroll
    for (int r = i + 1; r < M; ++r) {
      const double f = A[r][i] / A[i][i];
#pragma unroll
      for (int c = i; c < M; ++c) A[r][c] -= f * A[i][c];
      b[r] -= f * b[i];
    }
#pragma unroll
  for (int i = M - 1; i >= 0; --i) {
    double sum = b[i];
#pragma unroll
    for (int c = i + 1; c < M; ++c) sum -= A[i][c] * b[c];
    b[i] = sum / A[i][i];
  }
}

// M-SHAKE of a cluster (centre at index 0): six Newton sweeps of the
// linearized bond equations sum_l 2 (r_k . s_l) w_kl lambda_l = d_k^2 -
// r_k . r_k, the members moving by +lambda_l s_l / m_l along the old bonds
// s_l and the centre by -sum_l lambda_l s_l / m_0 (w_kk = 1/m_0 + 1/m_k,
// w_kl = 1/m_0), the sweeps of the reference (ConstraintIR.cpp:595-747).
template <int S>
__device__ __forceinline__ void mshake_positions(const double (&x0)[S][3], double (&xn)[S][3],
                                                 const double (&im)[S],
                                                 const double (&len)[S]) {
  constexpr int M = S - 1;
  double sv[M][3], d2[M];
#pragma unroll
  for (int k = 0; k < M; ++k) {
    d2[k] = len[k + 1] * len[k + 1];
#pragma unroll
    for (int j = 0; j < 3; ++j) sv[k][j] = x0[k + 1][j] - x0[0][j];
  }
  for (int sweep = 0; sweep < 6; ++sweep) {
    double r[M][3], A[M][M], lam[M];
#pragma unroll
    for (int k = 0; k < M; ++k) {
#pragma unroll
      for (int j = 0; j < 3; ++j) r[k][j] = xn[k + 1][j] - xn[0][j];
      lam[k] = d2[k] - dot(r[k], r[k]);
    }
#pragma unroll
    for (int k = 0; k < M; ++k)
#pragma unroll
      for (int l = 0; l < M; ++l) A[k][l] = 2.0 * dot(r[k], sv[l]) * ((k == l) ? im[0] + im[k + 1] : im[0]);
    solve_small<M>(A, lam);
#pragma unroll
    for (int l = 0; l < M; ++l)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        xn[0][j] -= lam[l] * sv[l][j] * im[0];
        xn[l + 1][j] += lam[l] * sv[l][j] * im[l + 1];
      }
  }
}

// The velocity constraint of a cluster: impulses mu_l along the bonds r_l
// with sum_l (r_k . r_l) w_kl mu_l = -r_k . (v_k - v_0), one solve.
template <int S>
__device__ __forceinline__ void rattle_velocities(const double (&x)[S][3], double (&v)[S][3],
                                                  const double (&im)[S]) {
  constexpr int M = S - 1;
  double r[M][3], A[M][M], mu[M];
#pragma unroll
  for (int k = 0; k < M; ++k) {
    double dv[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      r[k][j] = x[k + 1][j] - x[0][j];
      dv[j] = v[k + 1][j] - v[0][j];
    }
    mu[k] = -dot(r[k], dv);
  }
#pragma unroll
  for (int k = 0; k < M; ++k)
#pragma unroll
    for (int l = 0; l < M; ++l) A[k][l] = dot(r[k], r[l]) * ((k == l) ? im[0] + im[k + 1] : im[0]);
  solve_small<M>(A, mu);
#pragma unroll
  for (int l = 0; l < M; ++l)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      v[0][j] -= mu[l] * r[l][j] * im[0];
      v[l + 1][j] += mu[l] * r[l][j] * im[l + 1];
    }
}

// Loads the unit at `first`; with clear, the force words are zeroed for the
// next evaluation once read.  Fr (or null): reciprocal force words by atom,
// added to the place-ordered words as integers before the conversion, so
// the force is the one a single word would hold.
template <int S, bool CL>
__device__ __forceinline__ void load_unit(int first, int code, const double *__restrict__ x,
                                          const double *__restrict__ v,
                                          const double *__restrict__ inv_mass,
                                          long long *__restrict__ F,
                                          long long *__restrict__ Fr,
                                          const int *__restrict__ place_of, bool clear,
                                          const IntParams &p, UnitState<S> &s) {
  const int centre = CL ? (code >> kUnitCentreShift) & 15 : 0;
#pragma unroll
  for (int m = 0; m < S; ++m) {
    // clusters: the centre first, then the others in their order
    const int at = CL ? first + (m == 0 ? centre : (m - 1 < centre ? m - 1 : m)) : first + m;
    s.at[m] = at;
    s.len[m] = CL && m > 0 ? p.clen[at] : 0.0;
    s.pl[m] = place_of[at];
    s.im[m] = inv_mass[at];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      long long w = F[3 * s.pl[m] + k];
      if (Fr) {
        w += Fr[3 * at + k];
        if (clear) Fr[3 * at + k] = 0;
      }
      s.f[m][k] = (double)w * kForceUnscale;
      if (clear) F[3 * s.pl[m] + k] = 0;
      s.v[m][k] = v[3 * at + k];
      s.x[m][k] = x[3 * at + k];
    }
  }
}

template <int S>
__device__ __forceinline__ void half_kick(UnitState<S> &s, double dt) {
#pragma unroll
  for (int m = 0; m < S; ++m)
#pragma unroll
    for (int k = 0; k < 3; ++k) s.v[m][k] = s.v[m][k] + 0.5 * dt * s.f[m][k] * s.im[m];
}

template <int S, bool CL>
__device__ __forceinline__ void constrain_velocities(UnitState<S> &s, const double (&x)[S][3],
                                                     const IntParams &p) {
  if constexpr (CL) rattle_velocities<S>(x, s.v, s.im);
  else if constexpr (S == 3) settle_velocities(x[0], x[1], x[2], s.v[0], s.v[1], s.v[2], p);
}

// A drift of length h from positions x0 (constrained): xn = x0 + h v,
// SETTLE or M-SHAKE from x0, and v += (xn_constrained - xn) / h.
template <int S, bool CL>
__device__ __forceinline__ void drift_settle(const UnitState<S> &s, const double (&x0)[S][3],
                                             double (&v)[S][3], double (&xn)[S][3], double h,
                                             const IntParams &p) {
#pragma unroll
  for (int m = 0; m < S; ++m)
#pragma unroll
    for (int k = 0; k < 3; ++k) xn[m][k] = x0[m][k] + h * v[m][k];
  if constexpr (CL || S == 3) {
    double xu[S][3];
#pragma unroll
    for (int m = 0; m < S; ++m)
#pragma unroll
      for (int k = 0; k < 3; ++k) xu[m][k] = xn[m][k];
    if constexpr (CL) mshake_positions<S>(x0, xn, s.im, s.len);
    else settle_positions(x0[0], x0[1], x0[2], xn[0], xn[1], xn[2], p);
    const double ih = 1.0 / h;
#pragma unroll
    for (int m = 0; m < S; ++m)
#pragma unroll
      for (int k = 0; k < 3; ++k) v[m][k] += (xn[m][k] - xu[m][k]) * ih;
  }
}

// The rest of a g-BAOAB step after its first half kick (B): the geodesic
// integrator of Leimkuhler and Matthews (Proc. R. Soc. A 472, 20160138,
// 2016), as the reference's Langevin step with constraints takes it
// (lib/Driver/DynamicsIR.cpp, the geodesic branch): project v onto the
// constraints; A: half drift, SETTLE, v += dx_c / (dt/2); O: v <- decay v +
// spread / sqrt(m) R, projected again at the new positions; A: half drift
// from there, SETTLE, v += dx_c / (dt/2).  The closing B (kick and
// projection) is the second half of the step, as in velocity Verlet.
template <int S, bool CL>
__device__ __forceinline__ void baoab_rest(UnitState<S> &s, double (&xn)[S][3], long long step,
                                           const IntParams &p) {
  const double h = 0.5 * p.dt;
  constrain_velocities<S, CL>(s, s.x, p);
  double xh[S][3];
  drift_settle<S, CL>(s, s.x, s.v, xh, h, p);
#pragma unroll
  for (int m = 0; m < S; ++m) {
    if (!(s.im[m] > 0.0)) continue;  // a particle without mass keeps its velocity
    double r[3];
    langevin_normals(p.seed, step, (uint32_t)s.at[m], r);
    const double width = p.spread * sqrt(s.im[m]);
#pragma unroll
    for (int k = 0; k < 3; ++k) s.v[m][k] = p.decay * s.v[m][k] + width * r[k];
  }
  constrain_velocities<S, CL>(s, xh, p);
  drift_settle<S, CL>(s, xh, s.v, xn, h, p);
}

// Stores x, v and the f32 positions of the pair loop after a drift to xn;
// returns bit 0 when an atom moved beyond half the list buffer since the
// build, bit 1 beyond half the pruned list's buffer since the last prune.
template <int S>
__device__ __forceinline__ int store_drifted(UnitState<S> &s, const double (&xn)[S][3],
                                             double *__restrict__ x, double *__restrict__ v,
                                             const double *__restrict__ shift,
                                             const double *__restrict__ xref,
                                             const double *__restrict__ q,
                                             float4 *__restrict__ xq,
                                             const float4 *__restrict__ gorig,
                                             double4 *__restrict__ xpq,
                                             double4 *__restrict__ xpa, const IntParams &p) {
  int moved = 0;  // bit 0: beyond the list buffer, bit 1: beyond the pruned one
#pragma unroll
  for (int m = 0; m < S; ++m) {
    const int at = s.at[m];
    double d2 = 0.0, d2p = 0.0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      x[3 * at + k] = xn[m][k];
      v[3 * at + k] = s.v[m][k];
      const double d = xn[m][k] - xref[3 * at + k];
      d2 += d * d;
      if (p.x_prune) {
        const double dp = xn[m][k] - p.x_prune[3 * at + k];
        d2p += dp * dp;
      }
    }
    if (d2 > p.skin_half2) moved |= 1;
    if (d2p > p.prune_half2) moved |= 2;
    const float4 o = gorig ? gorig[s.pl[m] / (kCluster * kGroup)]
                           : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    xq[s.pl[m]] = make_float4(
        (float)(xn[m][0] - shift[3 * at + 0] - p.half_box[0] - (double)o.x),
        (float)(xn[m][1] - shift[3 * at + 1] - p.half_box[1] - (double)o.y),
        (float)(xn[m][2] - shift[3 * at + 2] - p.half_box[2] - (double)o.z), (float)q[at]);
    if (xpq) xpq[s.pl[m]] = make_double4(xn[m][0], xn[m][1], xn[m][2], q[at]);
    if (xpa) xpa[at] = make_double4(xn[m][0], xn[m][1], xn[m][2], q[at]);
  }
  return moved;
}

// Which halves of velocity Verlet a launch runs on each unit.
enum StepPart { kFirstHalf = 1, kSecondHalf = 2, kBetween = 3 };

template <int PART, int S, bool CL>
__device__ __forceinline__ void step_unit(int first, int code, long long step,
                                          double *__restrict__ x,
                                          double *__restrict__ v,
                                          const double *__restrict__ inv_mass,
                                          long long *__restrict__ F,
                                          long long *__restrict__ Fr,
                                          const int *__restrict__ place_of,
                                          const double *__restrict__ shift,
                                          const double *__restrict__ xref,
                                          const double *__restrict__ q,
                                          float4 *__restrict__ xq,
                                          const float4 *__restrict__ gorig,
                                          double4 *__restrict__ xpq,
                                          double4 *__restrict__ xpa, const IntParams &p,
                                          int *__restrict__ flags) {
  UnitState<S> s;
  // the force words are cleared by the launch that reads them last
  load_unit<S, CL>(first, code, x, v, inv_mass, F, Fr, place_of, PART != kSecondHalf, p, s);
  if (PART & kSecondHalf) {  // end of a step: half kick, velocity constraint
    half_kick<S>(s, p.dt);
    constrain_velocities<S, CL>(s, s.x, p);
  }
  if (PART & kFirstHalf) {   // start of a step: half kick, drift, constraints
    half_kick<S>(s, p.dt);
    double xn[S][3];
    if (p.langevin) baoab_rest<S, CL>(s, xn, step, p);
    else drift_settle<S, CL>(s, s.x, s.v, xn, p.dt, p);
    const int moved = store_drifted<S>(s, xn, x, v, shift, xref, q, xq, gorig, xpq, xpa, p);
    if (moved & 1) atomicOr(&flags[kFlagRebuild], 1);
    if (moved & 2) atomicOr(&flags[kFlagPrune], 1);
  } else {
#pragma unroll
    for (int m = 0; m < S; ++m)
#pragma unroll
      for (int k = 0; k < 3; ++k) v[3 * s.at[m] + k] = s.v[m][k];
  }
}

// PART kFirstHalf: half kick with the force words (cleared), drift, SETTLE,
// v += dx_c/dt, f32 positions of the pair loop, validity flag.
// kSecondHalf: half kick with the new forces (not cleared: the next step's
// first half reads them) and the velocity constraint.  kBetween: both, the
// second half of one step and the first half of the next, in one pass over
// the state with the same operations in the same order.
template <int PART>
__global__ void __launch_bounds__(64) integrate_part(
    int nunits, const int *__restrict__ unit_first, const int *__restrict__ unit_size,
    double *__restrict__ x, double *__restrict__ v, const double *__restrict__ inv_mass,
    long long *__restrict__ F, long long *__restrict__ Fr, const int *__restrict__ place_of,
    const double *__restrict__ shift, const double *__restrict__ xref,
    const double *__restrict__ q, float4 *__restrict__ xq, const float4 *__restrict__ gorig,
    double4 *__restrict__ xpq, double4 *__restrict__ xpa, IntParams p,
    int *__restrict__ flags) {
  const int u = blockIdx.x * blockDim.x + threadIdx.x;
  const bool noise = (PART & kFirstHalf) && p.langevin;
  const long long step = noise ? p.step[0] : 0;
  if (u < nunits) {
    const int first = unit_first[u], code = unit_size[u];
    if (code == 3)
      step_unit<PART, 3, false>(first, code, step, x, v, inv_mass, F, Fr, place_of, shift,
                                xref, q, xq, gorig, xpq, xpa, p, flags);
    else
      step_unit<PART, 1, false>(first, code, step, x, v, inv_mass, F, Fr, place_of, shift,
                                xref, q, xq, gorig, xpq, xpa, p, flags);
  }
  if (noise && threadIdx.x == 0) {
    // every block has read the step number before it counts itself done
    // (the fence completes this block's loads); the last one advances it
    __threadfence();
    unsigned long long *count = reinterpret_cast<unsigned long long *>(p.step + 1);
    if (atomicAdd(count, 1ull) + 1ull == (unsigned long long)gridDim.x) {
      p.step[0] = step + 1;
      p.step[1] = 0;
    }
  }
}

// The same parts for the constraint clusters, a launch of their own (their
// solves would raise the water kernel's registers): launched before
// integrate_part on the same stream, it reads the step number and leaves
// its advance to integrate_part.
template <int PART>
__global__ void __launch_bounds__(64) integrate_clusters(
    int nclusters, const int *__restrict__ cluster_first, const int *__restrict__ cluster_code,
    double *__restrict__ x, double *__restrict__ v, const double *__restrict__ inv_mass,
    long long *__restrict__ F, long long *__restrict__ Fr, const int *__restrict__ place_of,
    const double *__restrict__ shift, const double *__restrict__ xref,
    const double *__restrict__ q, float4 *__restrict__ xq, const float4 *__restrict__ gorig,
    double4 *__restrict__ xpq, double4 *__restrict__ xpa, IntParams p,
    int *__restrict__ flags) {
  const int u = blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= nclusters) return;
  const long long step = (PART & kFirstHalf) && p.langevin ? p.step[0] : 0;
  const int first = cluster_first[u], code = cluster_code[u];
#define STEP_CLUSTER(S)                                                                     \
  step_unit<PART, S, true>(first, code, step, x, v, inv_mass, F, Fr, place_of, shift, xref, q, \
                           xq, gorig, xpq, xpa, p, flags)
  switch (code & 0xff) {
    case 2: STEP_CLUSTER(2); break;
    case 3: STEP_CLUSTER(3); break;
    case 4: STEP_CLUSTER(4); break;
    default: STEP_CLUSTER(5); break;
  }
#undef STEP_CLUSTER
}

// The isotropic or axis-wise scaling of a Monte Carlo volume move: each
// unit (a rigid water or a free atom) moves with its centre of mass,
// X -> s X, and keeps its shape: x_i += (s - 1) X (per axis).  Masses come
// from the inverse masses (a massless atom has weight 0; a unit without
// mass moves with the mean of its positions).
__global__ void scale_units(int nunits, const int *__restrict__ unit_first,
                            const int *__restrict__ unit_size, const double *__restrict__ inv_mass,
                            double *__restrict__ x, double s0, double s1, double s2) {
  const int u = blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= nunits) return;
  const int first = unit_first[u], size = unit_size[u] & 0xff;
  double com[3] = {0.0, 0.0, 0.0}, total = 0.0;
  for (int m = 0; m < size; ++m) {
    const double im = inv_mass[first + m];
    const double w = im > 0.0 ? 1.0 / im : 0.0;
    total += w;
    for (int k = 0; k < 3; ++k) com[k] += w * x[3 * (first + m) + k];
  }
  if (total > 0.0) {
    for (int k = 0; k < 3; ++k) com[k] /= total;
  } else {
    for (int k = 0; k < 3; ++k) com[k] = 0.0;
    for (int m = 0; m < size; ++m)
      for (int k = 0; k < 3; ++k) com[k] += x[3 * (first + m) + k] / size;
  }
  const double s[3] = {s0, s1, s2};
  for (int m = 0; m < size; ++m)
    for (int k = 0; k < 3; ++k) x[3 * (first + m) + k] += (s[k] - 1.0) * com[k];
}

// Positions and charges by atom for PME in atom order (evaluations outside
// the step sequence, where no integration writes them).
__global__ void pme_pack_atoms(int n, const double *__restrict__ x, const double *__restrict__ q,
                               double4 *__restrict__ xpa) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) xpa[i] = make_double4(x[3 * i], x[3 * i + 1], x[3 * i + 2], q[i]);
}

// The pair loop's f32 positions (and the f64 ones of PME in place order) of
// every atom at its place in the current list, from x: the values a drift
// stores (store_drifted, the same expressions without contraction), for
// positions set outside the steps while the list stays.
__global__ void refresh_places(int n, const double *__restrict__ x,
                               const double *__restrict__ shift, const double *__restrict__ q,
                               const int *__restrict__ place_of,
                               const float4 *__restrict__ gorig, double hx, double hy,
                               double hz, float4 *__restrict__ xq,
                               double4 *__restrict__ xpq) {
  const int at = blockIdx.x * blockDim.x + threadIdx.x;
  if (at >= n) return;
  const int pl = place_of[at];
  const float4 o = gorig ? gorig[pl / (kCluster * kGroup)] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  const double x0 = x[3 * at], x1 = x[3 * at + 1], x2 = x[3 * at + 2];
  xq[pl] = make_float4((float)(x0 - shift[3 * at + 0] - hx - (double)o.x),
                       (float)(x1 - shift[3 * at + 1] - hy - (double)o.y),
                       (float)(x2 - shift[3 * at + 2] - hz - (double)o.z), (float)q[at]);
  if (xpq) xpq[pl] = make_double4(x0, x1, x2, q[at]);
}

// Reciprocal force words by atom added into the place-ordered words (and
// cleared): integer sums, so the result is the word a single buffer holds.
__global__ void pme_fold_atoms(int n, const int *__restrict__ place_of,
                               long long *__restrict__ Fr, long long *__restrict__ F) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int pl = place_of[i];
  for (int k = 0; k < 3; ++k) {
    F[3 * pl + k] += Fr[3 * i + k];
    Fr[3 * i + k] = 0;
  }
}

// Force words (place order) -> f64 forces (atom order).
__global__ void forces_to_atoms(int n, const int *__restrict__ place_of,
                                const long long *__restrict__ F, double *__restrict__ out,
                                size_t fs) {
  place_of = frame_at(place_of, fs);
  F = frame_at(F, fs);
  out = frame_at(out, fs);
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int pl = place_of[i];
  for (int k = 0; k < 3; ++k) out[3 * i + k] = (double)F[3 * pl + k] * kForceUnscale;
}

__global__ void clear_words(long long *__restrict__ w, size_t count, size_t fs) {
  w = frame_at(w, fs);
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < count; i += stride) w[i] = 0;
}

// Steepest descent of rigid waters (setup only): x_new = x + h F / |F|max,
// constrained from the old positions; the step is kept by the host when the
// energy falls.
// A descent step of a constraint cluster: each atom along its force, then
// M-SHAKE against the cluster's positions before the step.
template <int S>
__device__ void descend_cluster(int first, int code, const double *__restrict__ x0,
                                double *__restrict__ x, const long long *__restrict__ F,
                                const int *__restrict__ place_of, double step,
                                const IntParams &p) {
  const int centre = (code >> kUnitCentreShift) & 15;
  double xo[S][3], xn[S][3], im[S], len[S];
  int at[S];
#pragma unroll
  for (int m = 0; m < S; ++m) {
    at[m] = first + (m == 0 ? centre : (m - 1 < centre ? m - 1 : m));
    im[m] = p.inv_mass[at[m]];
    len[m] = m > 0 ? p.clen[at[m]] : 0.0;
    const int pl = place_of[at[m]];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      xo[m][k] = x0[3 * at[m] + k];
      xn[m][k] = xo[m][k] + step * (double)F[3 * pl + k] * kForceUnscale;
    }
  }
  mshake_positions<S>(xo, xn, im, len);
#pragma unroll
  for (int m = 0; m < S; ++m)
#pragma unroll
    for (int k = 0; k < 3; ++k) x[3 * at[m] + k] = xn[m][k];
}

__global__ void descend(int nunits, const int *__restrict__ unit_first,
                        const int *__restrict__ unit_size, const double *__restrict__ x0,
                        double *__restrict__ x, const long long *__restrict__ F,
                        const int *__restrict__ place_of, double step, IntParams p) {
  const int u = blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= nunits) return;
  const int first = unit_first[u], code = unit_size[u];
  if (code & kUnitCluster) {
    switch (code & 0xff) {
      case 2: descend_cluster<2>(first, code, x0, x, F, place_of, step, p); break;
      case 3: descend_cluster<3>(first, code, x0, x, F, place_of, step, p); break;
      case 4: descend_cluster<4>(first, code, x0, x, F, place_of, step, p); break;
      default: descend_cluster<5>(first, code, x0, x, F, place_of, step, p); break;
    }
    return;
  }
  const int size = code & 0xff;
  double xo[3][3], xn[3][3];
  for (int m = 0; m < size; ++m) {
    const int at = first + m;
    const int pl = place_of[at];
    for (int k = 0; k < 3; ++k) {
      xo[m][k] = x0[3 * at + k];
      xn[m][k] = xo[m][k] + step * (double)F[3 * pl + k] * kForceUnscale;
    }
  }
  if (size == 3) settle_positions(xo[0], xo[1], xo[2], xn[0], xn[1], xn[2], p);
  for (int m = 0; m < size; ++m)
    for (int k = 0; k < 3; ++k) x[3 * (first + m) + k] = xn[m][k];
}

// Back to the default contraction for the kernels that follow.
#pragma clang fp contract(fast)

}  // namespace mdhip
