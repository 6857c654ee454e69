// md_common.h — shared definitions of the gfx950 force-path kernels.
//
// Layout in HBM (one system):
//   atom order   x, v, x_ref, wrap shift: double[3n]; mass, inverse mass,
//                charge: double[n]; place_of: int[n]
//   place order  xq: float4 (x, y, z, q) relative to the origin of the place's
//                group, type_at, atom_at (-1 for padding), force words:
//                int64[3 np]
//   groups       gorig: float4 origin of each group of 32 places, relative to
//                the box centre, on a grid of 1/64 nm (boxes with a half edge
//                from 4 nm; zero otherwise, and xq is about the box centre)
//   clusters     8 consecutive places; bbox lo/hi float4; column of cluster
//   pair list    rows of capacity row_cap entries per group of 4 i-clusters:
//                {cj, shift | member mask | j real mask, explicit-mask slot};
//                the pruned rows have the same layout
// Forces are accumulated as 64-bit fixed point, 2^32 words per kJ/mol/nm,
// by integer atomics: the sum does not depend on the order of the adds.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace mdhip {

// Coulomb constant f = 138.93545755023302 kJ nm mol^-1 e^-2 as the exact
// binary64 value of include/mdir/Support/PhysicalConstants.h:18-21.
constexpr double kCoulomb = 0x1.15def44ac2182p+7;

constexpr int kCluster = 8;          // atoms per cluster (8 x 8 tile = wave64)
constexpr double kForceScale = 4294967296.0;         // 2^32 words per kJ/mol/nm
constexpr double kForceUnscale = 1.0 / 4294967296.0;
constexpr double kGridScale = 1099511627776.0;       // 2^40 (PME charge grid)
constexpr double kGridUnscale = 1.0 / 1099511627776.0;

// Lennard-Jones modifiers (lib/Driver/TermLennardJones.cpp:26-168,
// ExpandTruncation.cpp:19-45).
enum LJModifier { kLJNone = 0, kLJPotentialShift = 1, kLJForceSwitch = 2,
                  kLJPotentialSwitch = 3 };

// Device flags word layout (int32 each).
enum Flag { kFlagRebuild = 0, kFlagOverflowColumn = 1, kFlagOverflowRow = 2,
            kFlagBuilds = 3, kFlagSettleFail = 4, kFlagBrickOverflow = 5,
            kFlagMaskCount = 6, kFlagPrune = 7, kFlagPrunes = 8, kNumFlags = 16 };

constexpr int kGroup = 4;            // i-clusters of a group (one row)
// Row entries one wave of the pair loop takes: a launch shape of 16, 24 or
// 48 entries (row capacities are multiples of the largest, LDS sized for it).
constexpr int kSegmentMax = 48;

// One entry of a group row: j cluster cj against the i clusters of the
// group whose bits are set in imask.  The interaction mask of a tile is
// the outer product of the real-atom masks of its two clusters, unless
// `ex` >= 0 points at four explicit 64-bit masks (bit a + 8 b: i atom a,
// j atom b interact) for tiles with exclusions or i == j.
struct __attribute__((aligned(16))) PairEntry {
  int32_t cj;
  uint32_t info;     // shift (sx+1) | (sy+1) << 2 | (sz+1) << 4; imask << 8
  int32_t ex;        // -1, or index of uint64[kGroup] masks
  int32_t pad;
};

// Per-system constants of the nonbonded and PME kernels (passed by value).
struct NbParams {
  float box[3];
  float box_lo[3];     // box edge minus its f32 value
  float half_box[3];
  float rc2;
  float rc;
  float rs;            // switch start
  float beta;          // Ewald splitting, 0 = no Coulomb
  float coulomb;       // f
  float shift_c;       // erfc(beta rc)/rc under POTENTIAL_SHIFT, else 0
  float mb2l2e;        // -beta^2 log2(e): exp(-beta^2 r^2) = exp2(mb2l2e r^2)
  float tbs;           // 2 beta / sqrt(pi)
  // modifier constants, with the 1/12 and 1/6 of the (6 c6, 12 c12) table
  float sw_a12, sw_b12, sw_c12, sw_a6, sw_b6, sw_c6;   // force switch
  float shift12, shift6;                               // potential shift
  int ntypes;
  int lj_modifier;
};

__device__ __forceinline__ long long to_fixed(float f) {
  return __double2ll_rn((double)f * kForceScale);
}

// Frame batches: a launch over K frames of one system runs frame blockIdx.z,
// whose per-frame buffers lie fs bytes after those of the frame before
// (fs = 0 for a single system).  Null pointers stay null.  Byte-pointer
// arithmetic (not an integer round trip) keeps the pointer's global address
// space visible to the compiler, so that loads stay global_load.
template <class T>
__device__ __forceinline__ T *frame_at(T *p, size_t fs) {
  using Byte = typename std::conditional<std::is_const<T>::value, const char, char>::type;
  return reinterpret_cast<T *>(reinterpret_cast<Byte *>(p) +
                               (p ? (size_t)blockIdx.z * fs : (size_t)0));
}

__device__ __forceinline__ void shift_vector(uint32_t code, const float *box,
                                             float &sx, float &sy, float &sz) {
  sx = (float)((int)(code & 3u) - 1) * box[0];
  sy = (float)((int)((code >> 2) & 3u) - 1) * box[1];
  sz = (float)((int)((code >> 4) & 3u) - 1) * box[2];
}

// The offset added to a j group's positions to put them in the i group's
// frame: the difference of the two group origins plus the periodic shift
// k L (k in -1, 0, 1), with the box edge L split as its f32 value box plus
// box_lo.  The origins are multiples of 1/64 nm and box a multiple of its
// own ulp u <= 1/64 nm (edges below 2^18 nm), so ogj - ogi (exact) and
// (ogj - ogi) + k box are multiples of u.  The latter, the distance of the
// two origins across the boundary, is exact while it stays below L, which
// holds in dense systems, where a group spans little more than a cluster
// (the pair's origins lie within the cut-off, the list buffer and the two
// groups' extents); a group stretched over a sparse column rounds it once,
// as positions about the box centre would.  Adding k box_lo rounds once, on
// an offset of the order of the cut-off.
__device__ __forceinline__ float group_offset(uint32_t code, int d, float ogj, float ogi,
                                              float box, float box_lo) {
  const float k = (float)((int)((code >> (2 * d)) & 3u) - 1);
  return ((ogj - ogi) + k * box) + k * box_lo;
}

// Philox 4x32 with 10 rounds (Salmon, Moraes, Dror and Shaw, "Parallel
// random numbers: as easy as 1, 2, 3", SC11): a counter-based generator,
// so a particle's numbers at a step are a function of (seed, step,
// particle) alone. One round maps the counter (c0, c1, c2, c3) under the
// key (k0, k1) to (hi(B c2) ^ c1 ^ k0, lo(B c2), hi(A c0) ^ c3 ^ k1,
// lo(A c0)) with the multipliers A = 0xD2511F53, B = 0xCD9E8D57; the key
// is bumped by the Weyl constants (0x9E3779B9, 0xBB67AE85) between rounds.
__host__ __device__ __forceinline__ void philox4x32_10(uint32_t c[4], uint32_t k0,
                                                       uint32_t k1) {
#pragma unroll
  for (int round = 0; round < 10; ++round) {
    const uint64_t pa = (uint64_t)0xD2511F53u * c[0], pb = (uint64_t)0xCD9E8D57u * c[2];
    const uint32_t hi_a = (uint32_t)(pa >> 32), lo_a = (uint32_t)pa;
    const uint32_t hi_b = (uint32_t)(pb >> 32), lo_b = (uint32_t)pb;
    c[0] = hi_b ^ c[1] ^ k0;
    c[1] = lo_b;
    c[2] = hi_a ^ c[3] ^ k1;
    c[3] = lo_a;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}

// The three standard normal numbers of particle `id` at `step` for the
// Langevin O step (the reference's dyn.noise, docs/dialects/dyn.md:441-460,
// lib/Driver/NoiseIR.cpp): one Philox block under the seed's two words,
// with the counter (step low word, step high word, id, 2 << 24), stream 2
// in the high byte of the last word; each word w gives the uniform
// (w + 1/2) 2^-32 in (0, 1); Box-Muller on (u0, u1) gives the first two
// numbers (cosine, sine), on (u2, u3) the third (cosine).  The angle 2 pi u
// is taken as sincospi / cospi of 2u (2u exact): one argument reduction,
// exact for multiples of pi, shared by the cosine and sine of the first
// pair.  cos(2 pi u) of a rounded 2 pi u differs from it by a few ulp.
__device__ __forceinline__ void langevin_normals(unsigned long long seed, long long step,
                                                 uint32_t id, double r[3]) {
  uint32_t c[4] = {(uint32_t)(unsigned long long)step,
                   (uint32_t)((unsigned long long)step >> 32), id, 2u << 24};
  philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
  double u[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) u[k] = ((double)c[k] + 0.5) * 0x1.0p-32;
  const double rad0 = sqrt(-2.0 * log(u[0]));
  const double rad1 = sqrt(-2.0 * log(u[2]));
  double s0, c0;
  sincospi(2.0 * u[1], &s0, &c0);
  r[0] = rad0 * c0;
  r[1] = rad0 * s0;
  r[2] = rad1 * cospi(2.0 * u[3]);
}

}  // namespace mdhip
