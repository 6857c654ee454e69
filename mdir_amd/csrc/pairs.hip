// pairs.hip — the nonbonded pair loop and the excluded / 1-4 pair tuples.
//
// Reference semantics: the primary sum LJ + f qq [erfc(beta r)/r - s] over
// the non-excluded pairs within rc (lib/Driver/TermNonbonded.cpp:1423-1556,
// :456-467; LJ modifiers TermLennardJones.cpp:26-168 and the potential
// switch of ExpandTruncation.cpp:19-45), the excluded-pair correction
// -f qq erf(beta r)/r (TermNonbonded.cpp:1615-1671) and scaled 1-4 pairs
// (TermPairs14.cpp:11-17).
//
// Pair kernel: one wave per segment of a group row (kGroup i-clusters against
// up to a segment's j-clusters).  Lane l holds i atom a = l & 7 and j atom
// b = l >> 3 of each 8 x 8 tile, so a tile is one instruction stream with one
// pair per lane.  The force on j is accumulated over the members an entry
// pairs with, summed over the 8 lanes of a column by three DPP row operations
// and added once as 64-bit fixed point; the forces on the members stay in
// registers across the segment and are summed once at the end, over the
// wave's 8 columns through LDS.
#include "md_common.h"

namespace mdhip {

constexpr int kMaxTypes = 32;

// A DPP read of a lane of the same row.  Every control used here names a
// lane inside the row, so bound_ctrl never applies; with it set and the old
// value zero the compiler folds the move into the add that consumes it
// (v_add_f32_dpp).
template <int CTRL>
__device__ __forceinline__ float dpp(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, true));
}

// Sum over the 8 consecutive lanes of a half row (result in all 8 lanes):
// quad_perm [1,0,3,2], quad_perm [2,3,0,1], row_half_mirror.
__device__ __forceinline__ float sum8(float v) {
  v += dpp<0xB1>(v);
  v += dpp<0x4E>(v);
  v += dpp<0x141>(v);
  return v;
}

// erfc(x) = t exp(-x^2) P(t), t = 1/(1 + 0.475 x): a degree-6 polynomial
// fitted (weighted least squares on the relative error, this project's own
// fit, tests/golden/erfc_fit.py) on [0, 3.6]; below 1e-7 relative in exact
// arithmetic (9.8e-8 at worst), ~1e-6 in f32, where
// the rounding of exp(-x^2) itself dominates as it does for any f32 erfc.
// It reuses the exp(-x^2) that the force needs anyway.  t comes from the
// hardware reciprocal (1 ulp).  erfc_poly(t) is P(t).
__device__ __forceinline__ float erfc_poly(float t) {
  float pol = 0.09289151430130005f;
  pol = fmaf(pol, t, -0.4289827346801758f);
  pol = fmaf(pol, t, 0.6025476455688477f);
  pol = fmaf(pol, t, -0.1264616847038269f);
  pol = fmaf(pol, t, 0.3417808711528778f);
  pol = fmaf(pol, t, 0.24873463809490204f);
  pol = fmaf(pol, t, 0.26948970556259155f);
  return pol;
}

// Scalar force -dU/dr / r and energy of one pair in f32.  lj holds
// (6 c6, 12 c12), the force's own scaling; the energy divides it back out
// and the modifier constants carry the 1/6 and 1/12.  b475 = 0.475 beta
// (the t of erfc_poly is 1 / (1 + b475 r)).  Both parts are formed as
// (-dU/dr) r first and share the final 1/r^2:
//   Coulomb  qq exp(-beta^2 r^2) (t P(t) / r + 2 beta / sqrt(pi))
//   LJ       (12 c12 / r^6 - 6 c6) / r^6
// The hardware rsq, rcp and exp2 are used without range fix-ups: the
// caller runs this only on the lanes whose pair is on (r2 within rc and
// never zero or denormal: padding, excluded and i == j pairs are off, and
// overlapping atoms are not a state).
template <int MOD, bool COUL, bool ENERGY>
__device__ __forceinline__ float pair_force(float r2, float qq, float2 lj,
                                            const NbParams &p, float b475, float &e_lj,
                                            float &e_c) {
  const float rinv = __builtin_amdgcn_rsqf(r2);
  const float rinv2 = rinv * rinv;
  float g = 0.0f;  // Coulomb (-dU/dr) r
  if (COUL) {
    const float r = r2 * rinv;
    const float expt = __builtin_amdgcn_exp2f(p.mb2l2e * r2);  // exp(-beta^2 r^2)
    const float t = __builtin_amdgcn_rcpf(fmaf(b475, r, 1.0f));
    const float tp = t * erfc_poly(t);                          // erfc(beta r) exp(beta^2 r^2)
    g = qq * expt * fmaf(tp, rinv, p.tbs);
    if (ENERGY) e_c += qq * (tp * expt * rinv - p.shift_c);
  }
  const float c6x = lj.x, c12x = lj.y;
  // zero parameters give zero: no branch
  const float rinv6 = rinv2 * rinv2 * rinv2;
  const float l = fmaf(c12x, rinv6, -c6x);
  float ulj = 0.0f;
  if (ENERGY || MOD == kLJPotentialSwitch)
    ulj = fmaf(c12x * (1.0f / 12.0f), rinv6, -c6x * (1.0f / 6.0f)) * rinv6;
  float fscal;
  if (MOD == kLJForceSwitch) {
    const float r = r2 * rinv;
    const float u = fmaxf(r - p.rs, 0.0f);
    const float u2 = u * u, u3 = u2 * u;
    const float dp12 = p.sw_a12 * u2 + p.sw_b12 * u3;
    const float dp6 = p.sw_a6 * u2 + p.sw_b6 * u3;
    fscal = fmaf(l, rinv6, g) * rinv2 + (c12x * dp12 - c6x * dp6) * rinv;
    if (ENERGY) {
      const float p12 = p.sw_c12 + u3 * (p.sw_a12 * (1.0f / 3.0f) + p.sw_b12 * 0.25f * u);
      const float p6 = p.sw_c6 + u3 * (p.sw_a6 * (1.0f / 3.0f) + p.sw_b6 * 0.25f * u);
      ulj -= c12x * p12 - c6x * p6;
    }
  } else if (MOD == kLJPotentialSwitch) {
    float flj = l * rinv6 * rinv2;
    const float r = r2 * rinv;
    if (r > p.rs) {
      const float width = p.rc - p.rs;
      const float t = (r - p.rs) / width;
      const float t2 = t * t, t3 = t2 * t;
      const float S = 1.0f + t3 * (-10.0f + t * (15.0f - 6.0f * t));
      const float dS = (-30.0f * t2 + 60.0f * t3 - 30.0f * t3 * t) / width;
      flj = flj * S - ulj * dS * rinv;
      ulj *= S;
    }
    fscal = fmaf(g, rinv2, flj);
  } else {  // none, potential shift: the force is the plain one
    if (MOD == kLJPotentialShift && ENERGY) ulj -= c12x * p.shift12 - c6x * p.shift6;
    fscal = fmaf(l, rinv6, g) * rinv2;
  }
  if (ENERGY) e_lj += ulj;
  return fscal;
}

// Excluded pairs (every excluded pair, no cutoff) and scaled 1-4 pairs, in
// f64: thread per pair.  pairs14 params may be null when there are none.
// They run as the first plane of blocks of the pair kernel's launch.
struct SpecialParams {
  double box[3];
  double beta;
  double coulomb;
  double rc, rs;
  int lj_modifier;
};

__device__ void lj_pair_f64(double sigma, double eps, double r,
                            const SpecialParams &p, double &u, double &dudr) {
  u = 0.0;
  dudr = 0.0;
  if (eps == 0.0 || sigma == 0.0) return;
  const double e4 = 4.0 * eps, sr = sigma / r;
  const double s3 = sr * sr * sr, s6 = s3 * s3, s12 = s6 * s6;
  double raw = e4 * (s12 - s6), draw = e4 * (-12.0 * s12 + 6.0 * s6) / r;
  if (p.lj_modifier == kLJPotentialShift) {
    const double g3 = sigma * sigma * sigma, g6 = g3 * g3;
    raw -= e4 * (g6 * g6 * pow(p.rc, -12.0) - g6 * pow(p.rc, -6.0));
  } else if (p.lj_modifier == kLJPotentialSwitch && r > p.rs) {
    const double width = p.rc - p.rs, t = (r - p.rs) / width;
    const double S = 1.0 + t * t * t * (-10.0 + t * (15.0 - 6.0 * t));
    const double dS = (-30.0 * t * t + 60.0 * t * t * t - 30.0 * t * t * t * t) / width;
    draw = draw * S + raw * dS;
    raw *= S;
  } else if (p.lj_modifier == kLJForceSwitch) {
    const double rc = p.rc, rs = p.rs, width = rc - rs;
    const double uu = r > rs ? r - rs : 0.0;
    const double g3 = sigma * sigma * sigma, g6 = g3 * g3, g12 = g6 * g6;
    double P[2], dP[2];
    const double ns[2] = {12.0, 6.0};
    for (int k = 0; k < 2; ++k) {
      const double nn = ns[k];
      const double an = -nn * ((nn + 4.0) * rc - (nn + 1.0) * rs) / (pow(rc, nn + 2.0) * width * width);
      const double bn = nn * ((nn + 3.0) * rc - (nn + 1.0) * rs) / (pow(rc, nn + 2.0) * width * width * width);
      const double cn = pow(rc, -nn) - an / 3.0 * pow(width, 3.0) - bn / 4.0 * pow(width, 4.0);
      P[k] = cn + uu * uu * uu * (an / 3.0 + bn / 4.0 * uu);
      dP[k] = an * uu * uu + bn * uu * uu * uu;
    }
    const double rinv = 1.0 / r, r6 = pow(rinv, 6.0), r12 = r6 * r6;
    raw = e4 * (g12 * (r12 - P[0]) - g6 * (r6 - P[1]));
    draw = e4 * (g12 * (-12.0 * r12 * rinv - dP[0]) - g6 * (-6.0 * r6 * rinv - dP[1]));
  }
  u = raw;
  dudr = draw;
}

// The excluded and 1-4 pairs: thread per pair, blocks of 256 threads.
struct SpecialArgs {
  int npairs;
  const int2 *pairs;
  const int *is14;           // index into the 1-4 parameters, or -1
  const double *p14_sigma, *p14_eps, *p14_qq;
  const double *x, *q;
  const int *place_of;
  double *energy_partial;    // 3 per block, or null
  SpecialParams p;
};

__device__ __forceinline__ void special_block(int block, const SpecialArgs &sa,
                                              long long *__restrict__ F,
                                              double (*s_e)[256]) {
  const int npairs = sa.npairs;
  const int2 *__restrict__ pairs = sa.pairs;
  const int *__restrict__ is14 = sa.is14;
  const double *__restrict__ p14_sigma = sa.p14_sigma, *__restrict__ p14_eps = sa.p14_eps,
               *__restrict__ p14_qq = sa.p14_qq, *__restrict__ x = sa.x, *__restrict__ q = sa.q;
  const int *__restrict__ place_of = sa.place_of;
  double *__restrict__ energy_partial = sa.energy_partial;
  const SpecialParams &p = sa.p;
  const int t = block * blockDim.x + threadIdx.x;
  double e_ex = 0.0, e_lj14 = 0.0, e_c14 = 0.0;
  if (t < npairs) {
    const int2 ij = pairs[t];
    double d[3], r2 = 0.0;
    for (int k = 0; k < 3; ++k) {
      double dd = x[3 * ij.x + k] - x[3 * ij.y + k];
      dd -= p.box[k] * rint(dd / p.box[k]);
      d[k] = dd;
      r2 += dd * dd;
    }
    const double r = sqrt(r2);
    double dudr = 0.0;
    if (p.beta > 0.0) {
      const double qq = p.coulomb * q[ij.x] * q[ij.y];
      const double ef = erf(p.beta * r);
      e_ex = -qq * ef / r;
      dudr += -qq * (1.1283791670955126 * p.beta * exp(-p.beta * p.beta * r2) / r - ef / r2);
    }
    if (is14 && is14[t] >= 0) {
      const int k14 = is14[t];
      double u, du;
      lj_pair_f64(p14_sigma[k14], p14_eps[k14], r, p, u, du);
      const double uc = p14_qq[k14] / r;
      e_lj14 = u;
      e_c14 = uc;
      dudr += du - uc / r;
    }
    const double fs = -dudr / r;
    const long long wx = __double2ll_rn(fs * d[0] * kForceScale);
    const long long wy = __double2ll_rn(fs * d[1] * kForceScale);
    const long long wz = __double2ll_rn(fs * d[2] * kForceScale);
    unsigned long long *fi = reinterpret_cast<unsigned long long *>(F + 3 * place_of[ij.x]);
    unsigned long long *fj = reinterpret_cast<unsigned long long *>(F + 3 * place_of[ij.y]);
    atomicAdd(fi + 0, (unsigned long long)wx);
    atomicAdd(fi + 1, (unsigned long long)wy);
    atomicAdd(fi + 2, (unsigned long long)wz);
    atomicAdd(fj + 0, (unsigned long long)(-wx));
    atomicAdd(fj + 1, (unsigned long long)(-wy));
    atomicAdd(fj + 2, (unsigned long long)(-wz));
  }
  if (energy_partial) {
    s_e[0][threadIdx.x] = e_ex;
    s_e[1][threadIdx.x] = e_lj14;
    s_e[2][threadIdx.x] = e_c14;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
      if (threadIdx.x < s)
        for (int k = 0; k < 3; ++k) s_e[k][threadIdx.x] += s_e[k][threadIdx.x + s];
      __syncthreads();
    }
    if (threadIdx.x == 0)
      for (int k = 0; k < 3; ++k) energy_partial[3 * block + k] = s_e[k][0];
  }
}

// Grid: x over groups (4 waves per block), y = 0 the excluded and 1-4
// pairs, y = 1 + segment the pair segments.
template <int MOD, bool COUL, bool ENERGY, bool ORG>
__device__ __forceinline__ void nb_pairs_body(
    const float4 *__restrict__ xq, const float4 *__restrict__ gorig,
    const int *__restrict__ type_at,
    const uint32_t *__restrict__ real_mask, const PairEntry *__restrict__ rows,
    const int *__restrict__ row_count, const uint64_t *__restrict__ ex_masks,
    int row_cap, int seg_len, int max_groups, const int *__restrict__ counts_in,
    const float2 *__restrict__ ljtab, const NbParams &p, long long *__restrict__ F,
    double *__restrict__ energy_partial, int *__restrict__ flags, const SpecialArgs &sa) {
  extern __shared__ float2 s_lj[];                   // ntypes^2 (launch size)
  __shared__ union __align__(16) {
    struct {
      float fj[4][kSegmentMax * kCluster * 3];       // j forces of the entries
      int cj[4][kSegmentMax];
    } pr;
    double e[3][256];                                // special-pair energies
  } sm;
  const int nt2 = p.ntypes * p.ntypes;
  for (int k = threadIdx.x; k < nt2; k += blockDim.x) s_lj[k] = ljtab[k];
  // The build kernels of this step have finished: clear the request.
  if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) {
    flags[kFlagRebuild] = 0;
    flags[kFlagPrune] = 0;
  }
  __syncthreads();
  if (blockIdx.y == 0) {  // the excluded and 1-4 pairs
    if ((int)blockIdx.x * (int)blockDim.x < sa.npairs) special_block(blockIdx.x, sa, F, sm.e);
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
    if (ENERGY && gi < max_groups && lane == 0) {
      energy_partial[2 * slot] = 0.0;
      energy_partial[2 * slot + 1] = 0.0;
    }
    return;
  }
  const int a = lane & 7, b = lane >> 3;
  const int g0 = gi * kGroup;
  float4 xi[kGroup];
  int ti[kGroup];
  bool ireal[kGroup];
  float fi[kGroup][3];
#pragma unroll
  for (int m = 0; m < kGroup; ++m) {
    xi[m] = xq[(g0 + m) * kCluster + a];
    xi[m].w *= p.coulomb;
    ti[m] = type_at[(g0 + m) * kCluster + a] * p.ntypes;
    ireal[m] = (real_mask[g0 + m] >> a) & 1u;
    fi[m][0] = fi[m][1] = fi[m][2] = 0.0f;
  }
  // ORG: positions relative to group origins (large boxes), else relative
  // to the box centre with zero origins
  const float4 oi = ORG ? gorig[gi] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  const float b475 = 0.475f * p.beta;
  double e_lj = 0.0, e_c = 0.0;  // energy steps only: f64 partial sums
  const PairEntry *row = rows + (size_t)gi * row_cap + e0;
  // The segment's entries (n <= 48), one per lane in one vector load, read
  // out per entry with readlane (no scalar loads in the loop, whose waits
  // would also hold the LDS reads of the LJ parameters).  Software
  // pipeline: the j atoms one entry ahead are loaded before the current
  // entry is computed (indices clamped to the segment, so every load is in
  // bounds).
  const PairEntry mine = row[lane < n ? lane : n - 1];
  int cj = __builtin_amdgcn_readlane(mine.cj, 0);
  float4 xj = xq[cj * kCluster + b];
  float4 oj = ORG ? gorig[cj / kGroup] : oi;
  int tj = type_at[cj * kCluster + b];
  for (int e = 0; e < n; ++e) {
    const int cj1 = __builtin_amdgcn_readlane(mine.cj, e + 1 < n ? e + 1 : n - 1);
    const float4 xj1 = xq[cj1 * kCluster + b];
    const float4 oj1 = ORG ? gorig[cj1 / kGroup] : oi;
    const int tj1 = type_at[cj1 * kCluster + b];

    const uint32_t info = __builtin_amdgcn_readlane(mine.info, e);
    const int ex = __builtin_amdgcn_readlane(mine.ex, e);
    float2 ljm[kGroup];  // issued early: LDS latency overlaps the set-up
#pragma unroll
    for (int m = 0; m < kGroup; ++m) ljm[m] = s_lj[ti[m] + tj];
    // j positions in the i group's frame
    float xjx, xjy, xjz;
    if constexpr (ORG) {
      xjx = xj.x + group_offset(info, 0, oj.x, oi.x, p.box[0], p.box_lo[0]);
      xjy = xj.y + group_offset(info, 1, oj.y, oi.y, p.box[1], p.box_lo[1]);
      xjz = xj.z + group_offset(info, 2, oj.z, oi.z, p.box[2], p.box_lo[2]);
    } else {
      float sx, sy, sz;
      shift_vector(info & 63u, p.box, sx, sy, sz);
      xjx = xj.x + sx;
      xjy = xj.y + sy;
      xjz = xj.z + sz;
    }
    const bool jreal = (info >> (16 + b)) & 1u;
    // explicit masks (a subset of the real atoms' outer product), or all
    // pairs: wave-uniform words, taken as lane masks
    uint64_t exm[kGroup];
#pragma unroll
    for (int m = 0; m < kGroup; ++m) exm[m] = ~0ull;
    if (ex >= 0) {
#pragma unroll
      for (int m = 0; m < kGroup; ++m) exm[m] = ex_masks[(size_t)ex * kGroup + m];
    }
    float fjx = 0.0f, fjy = 0.0f, fjz = 0.0f;
#pragma unroll
    for (int m = 0; m < kGroup; ++m) {
      if (!((info >> (8 + m)) & 1u)) continue;  // uniform
      const float dx = xi[m].x - xjx, dy = xi[m].y - xjy, dz = xi[m].z - xjz;
      const float r2 = dx * dx + dy * dy + dz * dz;
      const bool on = __builtin_amdgcn_inverse_ballot_w64(exm[m]) && ireal[m] && jreal &&
                      r2 < p.rc2;
      // The pair's arithmetic runs under the exec mask of the lanes that are
      // on; a member tile with no pair inside rc (about half of those a
      // list with a buffer holds) branches over it.
      if (on) {
        float el = 0.0f, ec = 0.0f;
        const float fscal = pair_force<MOD, COUL, ENERGY>(r2, xi[m].w * xj.w, ljm[m], p, b475,
                                                          el, ec);
        if (ENERGY) {
          e_lj += (double)el;
          e_c += (double)ec;
        }
        const float fx = fscal * dx, fy = fscal * dy, fz = fscal * dz;
        fi[m][0] += fx;
        fi[m][1] += fy;
        fi[m][2] += fz;
        fjx -= fx;
        fjy -= fy;
        fjz -= fz;
      }
    }
    // j force: the 8 i lanes of column b, summed across the members, kept
    // in LDS; the atomics are issued after the loop, so that the loads of
    // the next entries never wait behind them
    const float gx = sum8(fjx), gy = sum8(fjy), gz = sum8(fjz);
    if (a == 0) {
      float *dst = &sm.pr.fj[wave][(e * kCluster + b) * 3];
      dst[0] = gx;
      dst[1] = gy;
      dst[2] = gz;
    }
    if (lane == 0) sm.pr.cj[wave][e] = cj;
    cj = cj1;
    xj = xj1;
    oj = oj1;
    tj = tj1;
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  // j forces of the segment: 24 words per entry, zero words skipped
  // (padding atoms and pairs beyond the cutoff give exact zeros).  A lane
  // per word, so one atomic instruction covers 64 consecutive words of a
  // few entries: a handful of cache lines.  (A lane per atom with its three
  // words takes three instructions over 64 atoms each, three times the
  // lines per word; that ran the kernel 8 µs slower.)
  unsigned long long *Fw = reinterpret_cast<unsigned long long *>(F);
  for (int k = lane; k < n * kCluster * 3; k += 64) {
    const int e = k / (kCluster * 3), rem = k - e * (kCluster * 3);
    const long long w = to_fixed(sm.pr.fj[wave][k]);
    if (w) atomicAdd(Fw + 3 * (sm.pr.cj[wave][e] * kCluster) + rem, (unsigned long long)w);
  }
  // i forces: the 8 j columns of the wave summed through LDS.  Lane (a, b)
  // stores its 12 sums (member m, coordinate c; v = 3 m + c) at
  // 64 v + 8 a + b, so the 8 columns of output o = 8 v + a are the 8 words
  // at 8 o; 96 outputs over two passes of the wave, summed in the tree
  // order ((b0 + b1) + (b2 + b3)) + ((b4 + b5) + (b6 + b7)).  Output o is
  // word 24 m + 3 a + c of the group's 96 consecutive force words, so each
  // pass's atomics cover consecutive words.
  float *red = sm.pr.fj[wave];
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  uint32_t rbits = 0;
#pragma unroll
  for (int m = 0; m < kGroup; ++m) {
    rbits |= (uint32_t)ireal[m] << m;
#pragma unroll
    for (int c = 0; c < 3; ++c) red[64 * (3 * m + c) + 8 * a + b] = fi[m][c];
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  for (int o = lane; o < kGroup * 3 * kCluster; o += 64) {
    const float4 u = *reinterpret_cast<const float4 *>(red + 8 * o);
    const float4 v = *reinterpret_cast<const float4 *>(red + 8 * o + 4);
    const float s = ((u.x + u.y) + (u.z + u.w)) + ((v.x + v.y) + (v.z + v.w));
    const int vi = o >> 3, m = vi / 3, c = vi - 3 * m;  // o & 7 == a
    const long long w = to_fixed(s);
    if (((rbits >> m) & 1u) && w)
      atomicAdd(Fw + 3 * ((size_t)(g0 + m) * kCluster + a) + c, (unsigned long long)w);
  }
  if (ENERGY) {
#pragma unroll
    for (int s = 1; s < 64; s <<= 1) {
      e_lj += __shfl_xor(e_lj, s);
      e_c += __shfl_xor(e_c, s);
    }
    if (lane == 0) {
      energy_partial[2 * slot] = e_lj;
      energy_partial[2 * slot + 1] = e_c;
    }
  }
}

// The kernels' entry: frame offsets, then the body.
template <int MOD, bool COUL, bool ENERGY, bool ORG>
__device__ __forceinline__ void nb_pairs_entry(
    const float4 *__restrict__ xq, const float4 *__restrict__ gorig,
    const int *__restrict__ type_at,
    const uint32_t *__restrict__ real_mask, const PairEntry *__restrict__ rows,
    const int *__restrict__ row_count, const uint64_t *__restrict__ ex_masks,
    int row_cap, int seg_len, int max_groups, const int *__restrict__ counts_in,
    const float2 *__restrict__ ljtab, const NbParams &p, long long *__restrict__ F,
    double *__restrict__ energy_partial, int *__restrict__ flags, SpecialArgs &sa,
    size_t fs) {
  xq = frame_at(xq, fs);
  gorig = frame_at(gorig, fs);
  type_at = frame_at(type_at, fs);
  real_mask = frame_at(real_mask, fs);
  rows = frame_at(rows, fs);
  row_count = frame_at(row_count, fs);
  ex_masks = frame_at(ex_masks, fs);
  counts_in = frame_at(counts_in, fs);
  F = frame_at(F, fs);
  energy_partial = frame_at(energy_partial, fs);
  flags = frame_at(flags, fs);
  sa.x = frame_at(sa.x, fs);
  sa.place_of = frame_at(sa.place_of, fs);
  sa.energy_partial = frame_at(sa.energy_partial, fs);
  nb_pairs_body<MOD, COUL, ENERGY, ORG>(xq, gorig, type_at, real_mask, rows, row_count, ex_masks, row_cap,
                                   seg_len, max_groups, counts_in, ljtab, p, F, energy_partial, flags,
                                   sa);
}

#define MDHIP_PAIRS_PARAMS                                                                  \
  const float4 *__restrict__ xq, const float4 *__restrict__ gorig,                          \
      const int *__restrict__ type_at, const uint32_t *__restrict__ real_mask,              \
      const PairEntry *__restrict__ rows, const int *__restrict__ row_count,                \
      const uint64_t *__restrict__ ex_masks, int row_cap, int seg_len, int max_groups,      \
      const int *__restrict__ counts_in, const float2 *__restrict__ ljtab, NbParams p,      \
      long long *__restrict__ F, double *__restrict__ energy_partial,                       \
      int *__restrict__ flags, SpecialArgs sa, size_t fs
#define MDHIP_PAIRS_ARGS                                                                    \
  xq, gorig, type_at, real_mask, rows, row_count, ex_masks, row_cap, seg_len, max_groups,   \
      counts_in, ljtab, p, F, energy_partial, flags, sa, fs

// The pair kernel, positions about the box centre (ORG false; boxes below
// 4 nm of half edge).
template <int MOD, bool COUL, bool ENERGY, bool ORG>
__global__ void __launch_bounds__(256) nb_pairs(MDHIP_PAIRS_PARAMS) {
  static_assert(!ORG, "boxes with group origins take nb_pairs_org");
  nb_pairs_entry<MOD, COUL, ENERGY, false>(MDHIP_PAIRS_ARGS);
}

// The pair kernel with positions relative to group origins: boxes of 4 nm
// of half edge or more.  There the pair loop is the step's critical path by
// far, so it runs with a 96-VGPR limit (five waves per SIMD instead of
// four) to hide more of its latency: 1.51 against 1.57 ms in place at
// 999,000 atoms.  In the smaller boxes the same limit speeds the kernel up
// too, but the PME chain beside it loses the CUs it takes and the step gets
// slower.
template <int MOD, bool COUL, bool ENERGY>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5)))
nb_pairs_org(MDHIP_PAIRS_PARAMS) {
  nb_pairs_entry<MOD, COUL, ENERGY, true>(MDHIP_PAIRS_ARGS);
}

// The pair kernel of a batch of frames (grid z over the frames), positions
// about each box's centre, with the same five-wave limit: a batch launches
// as many waves as a large box, and its PME chains beside the pair loop
// are short by comparison (64 frames of 3,150 atoms: 0.876-0.881 against
// 0.896-0.903 ms per force-matching call).
template <int MOD, bool COUL, bool ENERGY>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5)))
nb_pairs_batch(MDHIP_PAIRS_PARAMS) {
  nb_pairs_entry<MOD, COUL, ENERGY, false>(MDHIP_PAIRS_ARGS);
}

// One device-clock reading (wall_clock64), stored by lane 0 with a vector
// store: launched on a stream before and after a kernel, the two readings
// bracket it inside captured graphs, where events give no elapsed times.
__global__ void stamp_clock(unsigned long long *__restrict__ slot) {
  const unsigned long long t = (unsigned long long)wall_clock64();
  if (threadIdx.x == 0) slot[0] = t;
}

// Instantiations used by the engine.
#define MDHIP_PAIRS_ONE(MOD, C, E)                                                       \
  template __global__ void nb_pairs<MOD, C, E, false>(const float4 *, const float4 *,    \
      const int *, const uint32_t *, const PairEntry *, const int *, const uint64_t *, int, \
      int, int, const int *, const float2 *, NbParams, long long *, double *, int *,        \
      SpecialArgs, size_t);                                                              \
  template __global__ void nb_pairs_org<MOD, C, E>(const float4 *, const float4 *,       \
      const int *, const uint32_t *, const PairEntry *, const int *, const uint64_t *, int, \
      int, int, const int *, const float2 *, NbParams, long long *, double *, int *,        \
      SpecialArgs, size_t);                                                              \
  template __global__ void nb_pairs_batch<MOD, C, E>(const float4 *, const float4 *,     \
      const int *, const uint32_t *, const PairEntry *, const int *, const uint64_t *, int, \
      int, int, const int *, const float2 *, NbParams, long long *, double *, int *,        \
      SpecialArgs, size_t);
#define MDHIP_PAIRS(MOD)                                                                 \
  MDHIP_PAIRS_ONE(MOD, true, false)                                                      \
  MDHIP_PAIRS_ONE(MOD, true, true)                                                       \
  MDHIP_PAIRS_ONE(MOD, false, false)                                                     \
  MDHIP_PAIRS_ONE(MOD, false, true)
MDHIP_PAIRS(kLJNone)
MDHIP_PAIRS(kLJPotentialShift)
MDHIP_PAIRS(kLJForceSwitch)
MDHIP_PAIRS(kLJPotentialSwitch)

}  // namespace mdhip
