// nbnxm.hip — cell-sorted cluster build for gfx950.
//
// Replaces the reference's neighbor builds (lib/Runtime/Templates/
// NeighborsMatrix.mlir:193-460 for the pair set it defines,
// lib/Conversion/MDExecKernels/GroupsBuild.cpp:2515 for the GPU structure)
// with a layout made for wave64: atoms are binned into xy columns, each
// column is sorted along z, and every 8 consecutive places form a cluster
// with a bounding box; columns hold whole groups of 4 clusters.  A wave
// searches the clusters near one group and writes a row of {j-cluster,
// periodic shift, the members it pairs with}: an 8 x 8 tile is one pair per
// lane of a wave64.
//
// Three kernels, each a no-op unless the rebuild flag is set, so a captured
// step graph carries them every step and they cost one empty launch when the
// list is still valid (validity test: 04-neighbors.md:54-133).
#include "md_common.h"

namespace mdhip {

struct GridParams {
  double box[3];
  double inv_box[3];
  double half_box[3];
  double wx, wy;          // column widths
  int ncx, ncy, ncol;
  int col_cap;            // bucket capacity per column (atoms)
  int row_cap;            // entries per i-cluster row
  int max_clusters;
  int rx, ry;             // column search ranges
  float rlist2;           // squared reach of the bbox search
  int max_excl;           // partners per atom kept in the exclusion table
};

constexpr int kSortCapMax = 1024;   // max atoms per column the sort handles

// Sets the rebuild flag (a kernel, not a memset, so that graphs may hold it).
__global__ void nb_request_build(int *flags, size_t fs) {
  flags = frame_at(flags, fs);
  if (threadIdx.x == 0) flags[kFlagRebuild] = 1;
}

// Thread per atom: wrap into the box, bin into a column bucket, remember
// the wrap shift and the reference positions of the validity test.  A block
// first counts its atoms per column in LDS, takes one range of each bucket
// with one global atomic per column it touches, then places its atoms: the
// fill order may vary from run to run, the sort of nb_sort makes it not
// matter.
constexpr int kBinLdsColumns = 8192;
__global__ void __launch_bounds__(256) nb_bin(int n, const double *__restrict__ x,
                                             double *__restrict__ shift,
                                             double *__restrict__ xref,
                                             int *__restrict__ col_count,
                                             int *__restrict__ bucket,
                                             int *__restrict__ flags, GridParams g,
                                             size_t fs) {
  x = frame_at(x, fs);
  shift = frame_at(shift, fs);
  xref = frame_at(xref, fs);
  col_count = frame_at(col_count, fs);
  bucket = frame_at(bucket, fs);
  flags = frame_at(flags, fs);
  if (flags[kFlagRebuild] == 0) return;
  __shared__ int s_hist[kBinLdsColumns];
  const bool lds = g.ncol <= kBinLdsColumns;
  if (lds) {
    for (int c = threadIdx.x; c < g.ncol; c += blockDim.x) s_hist[c] = 0;
    __syncthreads();
  }
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  int col = -1, local = 0;
  if (i < n) {
    double w[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const double xv = x[3 * i + k];
      const double sh = floor(xv * g.inv_box[k]) * g.box[k];
      shift[3 * i + k] = sh;
      xref[3 * i + k] = xv;
      w[k] = xv - sh;
    }
    int cx = (int)(w[0] / g.wx);
    int cy = (int)(w[1] / g.wy);
    cx = cx < 0 ? 0 : (cx >= g.ncx ? g.ncx - 1 : cx);
    cy = cy < 0 ? 0 : (cy >= g.ncy ? g.ncy - 1 : cy);
    col = cy * g.ncx + cx;
    if (lds) local = atomicAdd(&s_hist[col], 1);
  }
  if (lds) {
    __syncthreads();
    for (int c = threadIdx.x; c < g.ncol; c += blockDim.x) {
      const int cnt = s_hist[c];
      if (cnt) s_hist[c] = atomicAdd(&col_count[c], cnt);
    }
    __syncthreads();
  }
  if (col < 0) return;
  const int slot = lds ? s_hist[col] + local : atomicAdd(&col_count[col], 1);
  if (slot < g.col_cap)
    bucket[col * g.col_cap + slot] = i;
  else
    atomicOr(&flags[kFlagOverflowColumn], 1);
}

// Block per column: place offset by a prefix over the preceding columns,
// rank sort of the column's atoms by (z, atom index) — a total order, so
// the result does not depend on the order the bucket was filled in — then
// places, padding, cluster boxes (about the box centre), group origins and
// the f32 positions relative to them (origins zero unless `origins`).
__global__ void __launch_bounds__(256) nb_sort(
    const double *__restrict__ x, const double *__restrict__ shift,
    const double *__restrict__ q, const int *__restrict__ type,
    const int *__restrict__ col_count, const int *__restrict__ bucket,
    int *__restrict__ atom_at, int *__restrict__ place_of,
    float4 *__restrict__ xq, float4 *__restrict__ gorig, double4 *__restrict__ xpq,
    int *__restrict__ type_at,
    float4 *__restrict__ bb_lo, float4 *__restrict__ bb_hi,
    int *__restrict__ cluster_col, uint32_t *__restrict__ real_mask,
    int *__restrict__ col_cl_start, int *__restrict__ col_cl_count,
    int *__restrict__ counts_out, int *__restrict__ flags, GridParams g, int origins,
    size_t fs) {
  x = frame_at(x, fs);
  shift = frame_at(shift, fs);
  col_count = frame_at(col_count, fs);
  bucket = frame_at(bucket, fs);
  atom_at = frame_at(atom_at, fs);
  place_of = frame_at(place_of, fs);
  xq = frame_at(xq, fs);
  gorig = frame_at(gorig, fs);
  xpq = frame_at(xpq, fs);
  type_at = frame_at(type_at, fs);
  bb_lo = frame_at(bb_lo, fs);
  bb_hi = frame_at(bb_hi, fs);
  cluster_col = frame_at(cluster_col, fs);
  real_mask = frame_at(real_mask, fs);
  col_cl_start = frame_at(col_cl_start, fs);
  col_cl_count = frame_at(col_cl_count, fs);
  counts_out = frame_at(counts_out, fs);
  flags = frame_at(flags, fs);
  if (flags[kFlagRebuild] == 0) return;
  __shared__ int s_red[256];
  __shared__ int s_atom[kSortCapMax];
  __shared__ double s_z[kSortCapMax];
  __shared__ float4 s_pos[kSortCapMax];
  __shared__ int s_ratom[kSortCapMax];                       // atom by rank
  __shared__ float4 s_gor[kSortCapMax / (kCluster * kGroup)];  // group origins
  const int col = blockIdx.x, tid = threadIdx.x;
  const int cap = g.col_cap < kSortCapMax ? g.col_cap : kSortCapMax;

  // every column takes whole groups of clusters, so that no group of the
  // pair list straddles two columns
  constexpr int kPad = kCluster * kGroup;
  int sum = 0;
  for (int c = tid; c < col; c += blockDim.x) {
    int cnt = col_count[c];
    cnt = cnt < cap ? cnt : cap;
    sum += (cnt + kPad - 1) & ~(kPad - 1);
  }
  s_red[tid] = sum;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (tid < s) s_red[tid] += s_red[tid + s];
    __syncthreads();
  }
  const int start = s_red[0];
  int count = col_count[col];
  count = count < cap ? count : cap;
  const int padded = (count + kPad - 1) & ~(kPad - 1);

  for (int t = tid; t < count; t += blockDim.x) {
    int a = bucket[col * g.col_cap + t];
    s_atom[t] = a;
    s_z[t] = x[3 * a + 2] - shift[3 * a + 2];
  }
  __syncthreads();
  for (int t = tid; t < count; t += blockDim.x) {
    double za = s_z[t];
    int ia = s_atom[t];
    int rank = 0, u = 0;
    // four independent LDS reads per pass, so that their latencies overlap
    for (; u + 4 <= count; u += 4) {
      const double z0 = s_z[u], z1 = s_z[u + 1], z2 = s_z[u + 2], z3 = s_z[u + 3];
      const int i0 = s_atom[u], i1 = s_atom[u + 1], i2 = s_atom[u + 2], i3 = s_atom[u + 3];
      rank += ((z0 < za) || (z0 == za && i0 < ia)) + ((z1 < za) || (z1 == za && i1 < ia)) +
              ((z2 < za) || (z2 == za && i2 < ia)) + ((z3 < za) || (z3 == za && i3 < ia));
    }
    for (; u < count; ++u) {
      double zb = s_z[u];
      int ib = s_atom[u];
      rank += (zb < za) || (zb == za && ib < ia);
    }
    int place = start + rank;
    atom_at[place] = ia;
    place_of[ia] = place;
    float4 p;
    p.x = (float)(x[3 * ia + 0] - shift[3 * ia + 0] - g.half_box[0]);
    p.y = (float)(x[3 * ia + 1] - shift[3 * ia + 1] - g.half_box[1]);
    p.z = (float)(x[3 * ia + 2] - shift[3 * ia + 2] - g.half_box[2]);
    p.w = (float)q[ia];
    if (!origins) xq[place] = p;  // about the box centre
    xpq[place] = make_double4(x[3 * ia + 0], x[3 * ia + 1], x[3 * ia + 2], q[ia]);
    type_at[place] = type[ia];
    s_pos[rank] = p;
    s_ratom[rank] = ia;
  }
  for (int t = count + tid; t < padded; t += blockDim.x) {
    int place = start + t;
    atom_at[place] = -1;
    // far from everything and from each other; masked out of every tile
    xq[place] = make_float4(1.0e5f, 1.0e5f, 1.0e5f + (float)(place & 1023), 0.0f);
    xpq[place] = make_double4(0.0, 0.0, 0.0, 0.0);  // zero charge: no PME work
    type_at[place] = 0;
  }
  __syncthreads();
  const int cl_start = start / kCluster, ncl = padded / kCluster;
  if (tid == 0) {
    col_cl_start[col] = cl_start;
    col_cl_count[col] = (count + kCluster - 1) / kCluster;  // clusters holding atoms
    if (col == 0) {
      atomicAdd(&flags[kFlagBuilds], 1);
      flags[kFlagMaskCount] = 0;  // explicit-mask slots of the coming search
    }
    if (col == g.ncol - 1) {
      counts_out[0] = start + padded;         // places
      counts_out[1] = (start + padded) / kCluster;  // clusters
    }
  }
  for (int c = tid; c < ncl; c += blockDim.x) {
    float4 lo = make_float4(3.0e38f, 3.0e38f, 3.0e38f, 0.0f);
    float4 hi = make_float4(-3.0e38f, -3.0e38f, -3.0e38f, 0.0f);
    uint32_t m = 0;
    for (int a = 0; a < kCluster; ++a) {
      int r = c * kCluster + a;
      if (r < count) {
        float4 p = s_pos[r];
        lo.x = fminf(lo.x, p.x); lo.y = fminf(lo.y, p.y); lo.z = fminf(lo.z, p.z);
        hi.x = fmaxf(hi.x, p.x); hi.y = fmaxf(hi.y, p.y); hi.z = fmaxf(hi.z, p.z);
        m |= 1u << a;
      }
    }
    bb_lo[cl_start + c] = lo;
    bb_hi[cl_start + c] = hi;
    cluster_col[cl_start + c] = col;
    real_mask[cl_start + c] = m;
  }
  // Group origins (large boxes): the lower corner of each group's atoms on
  // a grid of 1/64 nm (exact in f32, as are their differences), and the
  // pair loop's f32 positions relative to their group's origin, so that
  // their spacing does not grow with the box.  The pair loop adds the exact
  // difference of two origins and the periodic shift to the j atoms.
  // Without origins they are zero: positions about the box centre.
  const int ngr = ncl / kGroup;
  if (!origins) {
    for (int gl = tid; gl < ngr; gl += blockDim.x)
      gorig[cl_start / kGroup + gl] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    return;
  }
  for (int gl = tid; gl < ngr; gl += blockDim.x) {
    const int r0 = gl * kPad, r1 = r0 + kPad < count ? r0 + kPad : count;
    float lx = 3.0e38f, ly = 3.0e38f, lz = 3.0e38f;
    for (int r = r0; r < r1; ++r) {
      lx = fminf(lx, s_pos[r].x);
      ly = fminf(ly, s_pos[r].y);
      lz = fminf(lz, s_pos[r].z);
    }
    const float4 o = r1 > r0 ? make_float4(floorf(lx * 64.0f) * 0.015625f,
                                           floorf(ly * 64.0f) * 0.015625f,
                                           floorf(lz * 64.0f) * 0.015625f, 0.0f)
                             : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    s_gor[gl] = o;
    gorig[cl_start / kGroup + gl] = o;
  }
  __syncthreads();
  for (int r = tid; r < count; r += blockDim.x) {
    const int ia = s_ratom[r];
    const float4 o = s_gor[r / kPad];
    xq[start + r] = make_float4(
        (float)(x[3 * ia + 0] - shift[3 * ia + 0] - g.half_box[0] - (double)o.x),
        (float)(x[3 * ia + 1] - shift[3 * ia + 1] - g.half_box[1] - (double)o.y),
        (float)(x[3 * ia + 2] - shift[3 * ia + 2] - g.half_box[2] - (double)o.z), (float)q[ia]);
  }
}

__device__ __forceinline__ float axis_gap(float lo_i, float hi_i, float lo_j,
                                          float hi_j) {
  float g1 = lo_j - hi_i, g2 = lo_i - hi_j;
  float g = fmaxf(g1, g2);
  return g > 0.0f ? g : 0.0f;
}

// Block per group of kGroup consecutive i-clusters (one column's).  The
// candidates are the clusters of the columns within the reach of the group's
// box, at z shifts -1, 0, +1, whose z range meets the group's z range
// widened by the reach (a contiguous range of each column's z-sorted
// clusters, found by binary search).  Wave 1 tests 64 column slots at a
// time and lays out their candidates with a wave prefix sum, while wave 0
// builds the group's exclusion table; the four waves then test 64
// candidates each per round (wave w takes the round's iteration w), and a
// block prefix over the waves places their entries, so the row holds the
// candidates in their order, as one wave walking them would write it.  Bit
// m of an entry's imask is set when j is within the reach of member m and
// the tile (member, j, shift) is ordered, so that each unordered cluster
// pair occurs once.  Tiles of a cluster with itself, and tiles holding
// excluded pairs, get four explicit 64-bit masks (MDExecKernels.cpp:1748-1823
// removes the same partners from a matrix row).
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6))) nb_search(
    const float4 *__restrict__ bb_lo, const float4 *__restrict__ bb_hi,
    const int *__restrict__ cluster_col, const uint32_t *__restrict__ real_mask,
    const int *__restrict__ col_cl_start, const int *__restrict__ col_cl_count,
    const int *__restrict__ atom_at, const int *__restrict__ place_of,
    const int *__restrict__ excl_start, const int *__restrict__ excl_atoms,
    const int *__restrict__ counts_in, int *__restrict__ col_count,
    PairEntry *__restrict__ rows, int *__restrict__ row_count,
    uint64_t *__restrict__ ex_masks, int ex_cap, int *__restrict__ flags,
    GridParams g, size_t fs) {
  bb_lo = frame_at(bb_lo, fs);
  bb_hi = frame_at(bb_hi, fs);
  cluster_col = frame_at(cluster_col, fs);
  real_mask = frame_at(real_mask, fs);
  col_cl_start = frame_at(col_cl_start, fs);
  col_cl_count = frame_at(col_cl_count, fs);
  atom_at = frame_at(atom_at, fs);
  place_of = frame_at(place_of, fs);
  counts_in = frame_at(counts_in, fs);
  col_count = frame_at(col_count, fs);
  rows = frame_at(rows, fs);
  row_count = frame_at(row_count, fs);
  ex_masks = frame_at(ex_masks, fs);
  flags = frame_at(flags, fs);
  if (flags[kFlagRebuild] == 0) return;
  // The next build starts from empty buckets (nb_sort has read them).
  if (blockIdx.x == 0)
    for (int c = threadIdx.x; c < g.ncol; c += blockDim.x) col_count[c] = 0;
  constexpr int kWaves = 4;
  constexpr int kExMax = kGroup * kCluster * 16;  // 16 partners per atom at most
  __shared__ int s_ex_cluster[kExMax];
  __shared__ int s_ex_bit[kExMax];                // member << 8 | mask bit
  __shared__ int s_ex_n;
  // partner clusters of the group: open-addressing set, key -1 = empty,
  // value = the members that have a partner there
  constexpr int kPartnerSlots = kExMax;  // a power of two
  static_assert((kPartnerSlots & (kPartnerSlots - 1)) == 0, "partner set size");
  __shared__ int s_pkey[kPartnerSlots];
  __shared__ int s_pmem[kPartnerSlots];
  __shared__ int s_cstart[64], s_cbase[64], s_ccode[64];
  __shared__ int s_k0[3][64], s_n[2][64];  // per z shift: first cluster, count
  __shared__ int s_total;                  // candidates of the column chunk
  // the z bounds of the chunk's columns' clusters, staged so that the
  // binary searches read LDS instead of chains of dependent global loads
  // (chunks of more clusters read them from global memory).  1,024 keeps
  // the block at 19 KB of LDS and, with 63 VGPRs, eight blocks fit a CU:
  // a batch of frames launches thousands of these blocks.
  constexpr int kZStage = 1024;
  __shared__ float s_zlo[kZStage], s_zhi[kZStage];
  __shared__ int s_zoff[64], s_zn[64], s_kmin[64];
  __shared__ int s_ztotal;
  __shared__ int s_wcount[kWaves];         // entries per wave in a round
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int gi = blockIdx.x;
  const int ngroups = counts_in[1] / kGroup;
  if (gi >= ngroups) return;  // the whole block: no barrier skipped
  const int g0 = gi * kGroup;
  if (threadIdx.x == 0) s_ex_n = 0;
  for (int k = threadIdx.x; k < kPartnerSlots; k += blockDim.x) {
    s_pkey[k] = -1;
    s_pmem[k] = 0;
  }

  uint32_t mreal[kGroup];
  float4 mlo[kGroup], mhi[kGroup];
  float4 glo = make_float4(3.0e38f, 3.0e38f, 3.0e38f, 0.0f);
  float4 ghi = make_float4(-3.0e38f, -3.0e38f, -3.0e38f, 0.0f);
#pragma unroll
  for (int m = 0; m < kGroup; ++m) {
    mreal[m] = real_mask[g0 + m];
    mlo[m] = bb_lo[g0 + m];
    mhi[m] = bb_hi[g0 + m];
    if (mreal[m]) {
      glo.x = fminf(glo.x, mlo[m].x); glo.y = fminf(glo.y, mlo[m].y);
      glo.z = fminf(glo.z, mlo[m].z);
      ghi.x = fmaxf(ghi.x, mhi[m].x); ghi.y = fmaxf(ghi.y, mhi[m].y);
      ghi.z = fmaxf(ghi.z, mhi[m].z);
    }
  }
  uint32_t group_real = 0u;
#pragma unroll
  for (int m = 0; m < kGroup; ++m) group_real |= mreal[m];
  const bool any_real = group_real != 0u;
  const int icol = cluster_col[g0];
  const int cx = icol % g.ncx, cy = icol / g.ncx;
  const float Lx = (float)g.box[0], Ly = (float)g.box[1], Lz = (float)g.box[2];
  const float wxf = (float)g.wx, wyf = (float)g.wy;
  const float zreach = sqrtf(g.rlist2) * 1.0001f + 1.0e-5f;
  const int ncolx = 2 * g.rx + 1, ncols = any_real ? ncolx * (2 * g.ry + 1) : 0;
  __syncthreads();

  // exclusion table of the group: (partner cluster, member << 8 | mask bit)
  if (wave == 0 && lane < kGroup * kCluster) {
    const int m = lane >> 3, a = lane & 7;
    const int atom = atom_at[g0 * kCluster + lane];
    if (atom >= 0) {
      const int b0 = excl_start[atom], b1 = excl_start[atom + 1];
      for (int k = b0; k < b1; ++k) {
        const int pp = place_of[excl_atoms[k]];
        const int slot = atomicAdd(&s_ex_n, 1);
        if (slot < kExMax) {
          const int pc = pp / kCluster;
          s_ex_cluster[slot] = pc;
          s_ex_bit[slot] = (m << 8) | (a + kCluster * (pp % kCluster));
          // at most kExMax distinct keys: the probe always finds a slot
          for (int h = pc & (kPartnerSlots - 1);; h = (h + 1) & (kPartnerSlots - 1)) {
            const int prev = atomicCAS(&s_pkey[h], -1, pc);
            if (prev == -1 || prev == pc) {
              atomicOr(&s_pmem[h], 1 << m);
              break;
            }
          }
        } else {
          atomicOr(&flags[kFlagOverflowRow], 2);
        }
      }
    }
  }

  int count = 0;  // entries of the row so far (the same in every thread)
  PairEntry *row = rows + (size_t)gi * g.row_cap;
  for (int cbase = 0; cbase < ncols; cbase += 64) {
    // column slots of the chunk (wave 1): the column within reach of the
    // group's box, its periodic code and its clusters
    if (wave == 1) {
      const int c = cbase + lane;
      int cl0 = 0, code = 0, nreal = 0, kmin = 0;
      if (c < ncols) {
        const int dy = c / ncolx - g.ry, dx = c % ncolx - g.rx;
        int jy = cy + dy, sy = 0, jx = cx + dx, sx = 0;
        while (jy < 0) { jy += g.ncy; --sy; }
        while (jy >= g.ncy) { jy -= g.ncy; ++sy; }
        while (jx < 0) { jx += g.ncx; --sx; }
        while (jx >= g.ncx) { jx -= g.ncx; ++sx; }
        const float ylo = jy * wyf - 0.5f * Ly + sy * Ly, yhi = ylo + wyf;
        const float xlo = jx * wxf - 0.5f * Lx + sx * Lx, xhi = xlo + wxf;
        const float gy = axis_gap(glo.y, ghi.y, ylo, yhi);
        const float gx = axis_gap(glo.x, ghi.x, xlo, xhi);
        if (gx * gx + gy * gy < g.rlist2 && sx >= -1 && sx <= 1 && sy >= -1 && sy <= 1) {
          const int jcol = jy * g.ncx + jx;
          cl0 = col_cl_start[jcol];
          nreal = col_cl_count[jcol];
          // a tile is ordered only when cj >= ci: clusters before the group
          // never pair with it
          kmin = g0 - cl0 > 0 ? g0 - cl0 : 0;
          code = (sx + 1) | ((sy + 1) << 2);
        }
      }
      int zscan = nreal;
#pragma unroll
      for (int off = 1; off < 64; off <<= 1) {
        const int up = __shfl_up(zscan, off);
        if (lane >= off) zscan += up;
      }
      s_cstart[lane] = cl0;
      s_ccode[lane] = code;
      s_zn[lane] = nreal;
      s_zoff[lane] = zscan - nreal;
      s_kmin[lane] = kmin;
      if (lane == 63) s_ztotal = zscan;
    }
    __syncthreads();
    const int ztotal = s_ztotal;
    const bool staged = ztotal <= kZStage;
    if (staged) {
      for (int i = threadIdx.x; i < ztotal; i += blockDim.x) {
        // the slot holding i: the last slot whose offset is <= i (slots
        // without clusters share their offset with the next one)
        int lo = 0, hi = 63;
        while (lo < hi) {
          const int mid = (lo + hi + 1) >> 1;
          if (s_zoff[mid] <= i) lo = mid; else hi = mid - 1;
        }
        const int cl = s_cstart[lo] + (i - s_zoff[lo]);
        s_zlo[i] = bb_lo[cl].z;
        s_zhi[i] = bb_hi[cl].z;
      }
    }
    __syncthreads();
    if (wave == 1) {
      int ncand = 0;
      int k0[3] = {0, 0, 0}, nz[3] = {0, 0, 0};
      const int nreal = s_zn[lane], kmin = s_kmin[lane];
      if (nreal > 0) {
        const int cl0 = s_cstart[lane], zo = s_zoff[lane];
#pragma unroll
        for (int s = 0; s < 3; ++s) {
          // the z window in j's own frame: bb_hi.z >= zlo and bb_lo.z <= zhi
          const float oz = (float)(s - 1) * Lz;
          const float zlo = glo.z - zreach - oz, zhi = ghi.z + zreach - oz;
          int lo = kmin < nreal ? kmin : nreal, hi = nreal;
          while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            const float zh = staged ? s_zhi[zo + mid] : bb_hi[cl0 + mid].z;
            if (zh < zlo) lo = mid + 1; else hi = mid;
          }
          const int first = lo;
          hi = nreal;
          while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            const float zl = staged ? s_zlo[zo + mid] : bb_lo[cl0 + mid].z;
            if (zl <= zhi) lo = mid + 1; else hi = mid;
          }
          k0[s] = first;
          nz[s] = lo - first;
        }
        ncand = nz[0] + nz[1] + nz[2];
      }
      // inclusive wave scan of ncand
      int scan = ncand;
#pragma unroll
      for (int off = 1; off < 64; off <<= 1) {
        const int up = __shfl_up(scan, off);
        if (lane >= off) scan += up;
      }
      s_cbase[lane] = scan - ncand;  // exclusive
#pragma unroll
      for (int s = 0; s < 3; ++s) s_k0[s][lane] = k0[s];
      s_n[0][lane] = nz[0];
      s_n[1][lane] = nz[1];
      if (lane == 63) s_total = scan;
    }
    __syncthreads();  // column slots and (first chunk) exclusion table
    const int total = s_total;
    const int nex = s_ex_n < kExMax ? s_ex_n : kExMax;
    const int niter = (total + 63) / 64;
    for (int round = 0; round * kWaves < niter; ++round) {
      const int t = (round * kWaves + wave) * 64 + lane;
      uint32_t imask = 0, jreal = 0, scode = 0;
      int cj = 0, ex = -1;
      if (t < total) {
        // the column slot holding candidate t: last slot with cbase <= t
        int lo = 0, hi = 63;
        while (lo < hi) {
          const int mid = (lo + hi + 1) >> 1;
          if (s_cbase[mid] <= t) lo = mid; else hi = mid - 1;
        }
        int within = t - s_cbase[lo];
        const int n0 = s_n[0][lo], n1 = s_n[1][lo];
        int sz = -1;
        if (within >= n0) {
          within -= n0;
          sz = 0;
          if (within >= n1) {
            within -= n1;
            sz = 1;
          }
        }
        cj = s_cstart[lo] + s_k0[sz + 1][lo] + within;
        const int ccode = s_ccode[lo];
        const int sx = (ccode & 3) - 1, sy = ((ccode >> 2) & 3) - 1;
        const int lex = sz != 0 ? sz : (sy != 0 ? sy : sx);
        jreal = real_mask[cj];
        scode = (uint32_t)(sx + 1) | ((uint32_t)(sy + 1) << 2) | ((uint32_t)(sz + 1) << 4);
        if (jreal) {
          const float4 jlo = bb_lo[cj], jhi = bb_hi[cj];
          const float ox = sx * Lx, oy = sy * Ly, oz = sz * Lz;
#pragma unroll
          for (int m = 0; m < kGroup; ++m) {
            const int ci = g0 + m;
            const bool ordered = cj > ci || (cj == ci && lex >= 0);
            const float ex_ = axis_gap(mlo[m].x, mhi[m].x, jlo.x + ox, jhi.x + ox);
            const float ey_ = axis_gap(mlo[m].y, mhi[m].y, jlo.y + oy, jhi.y + oy);
            const float ez_ = axis_gap(mlo[m].z, mhi[m].z, jlo.z + oz, jhi.z + oz);
            if (mreal[m] && ordered && ex_ * ex_ + ey_ * ey_ + ez_ * ez_ < g.rlist2)
              imask |= 1u << m;
          }
        }
        if (imask) {
          // a tile of a member with itself, or one that holds excluded pairs
          const int self = cj - g0;
          bool special = self >= 0 && self < kGroup && ((imask >> self) & 1u);
          if (!special && nex > 0) {
            int h = cj & (kPartnerSlots - 1);
            for (int probe = 0; probe < kPartnerSlots; ++probe, h = (h + 1) & (kPartnerSlots - 1)) {
              const int key = s_pkey[h];
              if (key == cj) {
                special = (s_pmem[h] & (int)imask) != 0;
                break;
              }
              if (key == -1) break;
            }
          }
          if (special) {
            uint64_t jcols = 0;  // byte b all ones when j atom b is real
#pragma unroll
            for (int b = 0; b < kCluster; ++b)
              if ((jreal >> b) & 1u) jcols |= 0xFFull << (kCluster * b);
            uint64_t mk[kGroup];
#pragma unroll
            for (int m = 0; m < kGroup; ++m) {
              const uint64_t irows = (uint64_t)mreal[m] * 0x0101010101010101ull;
              mk[m] = ((imask >> m) & 1u) ? (irows & jcols) : 0ull;
              if (m == self) {
                // same cluster: i < j at zero shift, i != j under a shift
                uint64_t keep = 0;
#pragma unroll
                for (int b = 0; b < kCluster; ++b)
#pragma unroll
                  for (int a = 0; a < kCluster; ++a)
                    if (lex == 0 ? (a < b) : (a != b)) keep |= 1ull << (a + kCluster * b);
                mk[m] &= keep;
              }
            }
            for (int k = 0; k < nex; ++k)
              if (s_ex_cluster[k] == cj) {
                const int mm = s_ex_bit[k] >> 8;
                const uint64_t bit = 1ull << (s_ex_bit[k] & 0xff);
#pragma unroll
                for (int m = 0; m < kGroup; ++m)
                  if (m == mm) mk[m] &= ~bit;
              }
#pragma unroll
            for (int m = 0; m < kGroup; ++m)
              if (mk[m] == 0ull) imask &= ~(1u << m);
            if (imask) {
              const int slot = atomicAdd(&flags[kFlagMaskCount], 1);
              if (slot < ex_cap) {
#pragma unroll
                for (int m = 0; m < kGroup; ++m) ex_masks[(size_t)slot * kGroup + m] = mk[m];
                ex = slot;
              } else {
                atomicOr(&flags[kFlagOverflowRow], 4);
                imask = 0;
              }
            }
          }
        }
      }
      const bool accept = imask != 0u;
      const uint64_t ballot = __ballot(accept);
      if (lane == 0) s_wcount[wave] = __popcll(ballot);
      __syncthreads();
      // the entries of the waves before this one in the round come first
      int base = count;
      for (int w = 0; w < wave; ++w) base += s_wcount[w];
      if (accept) {
        const int slot = base + __popcll(ballot & ((1ull << lane) - 1ull));
        if (slot < g.row_cap) {
          PairEntry e;
          e.cj = cj;
          e.info = scode | (imask << 8) | (jreal << 16);
          e.ex = ex;
          e.pad = 0;
          row[slot] = e;
        } else {
          atomicOr(&flags[kFlagOverflowRow], 1);
        }
      }
      for (int w = 0; w < kWaves; ++w) count += s_wcount[w];
      __syncthreads();  // s_wcount is rewritten next round
    }
    // the column arrays are rewritten by the next chunk
    __syncthreads();
  }
  if (threadIdx.x == 0) row_count[gi] = count < g.row_cap ? count : g.row_cap;
}

// ---- the pruned list ------------------------------------------------------
//
// The list above is searched with a wide buffer and rebuilt rarely; the pair
// loop runs over a pruned copy that keeps, of each entry, the members with
// at least one atom pair within rc + the pruned buffer (the dual list of the
// reference, 04-neighbors.md).  Pruning runs when the list was rebuilt this
// step or an atom moved more than half the pruned buffer since the last
// prune (the flags the integration sets); it also records the positions the
// next test measures from.  Block per group: the eight waves take eight
// contiguous parts of the row and test one entry at a time as the pair
// loop does (lane l: i atom l & 7, j atom l >> 3 of each member tile, one
// ballot per member); each wave notes its kept entries (row index and new
// member mask) in LDS, and a block prefix over the parts places them, so
// the pruned row keeps the list's order (deterministic).  Dynamic LDS: one
// int per row entry (row_cap).
__global__ void __launch_bounds__(512) nb_prune(
    const float4 *__restrict__ xq, const float4 *__restrict__ gorig,
    const PairEntry *__restrict__ rows,
    const int *__restrict__ row_count, PairEntry *__restrict__ rows_in,
    int *__restrict__ row_count_in, const int *__restrict__ counts_in,
    const int *__restrict__ atom_at, const double *__restrict__ x,
    double *__restrict__ x_prune, int row_cap, float reach2, float Lx, float Ly,
    float Lz, float Lx_lo, float Ly_lo, float Lz_lo, int *__restrict__ flags) {
  if (flags[kFlagRebuild] == 0 && flags[kFlagPrune] == 0) return;
  if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&flags[kFlagPrunes], 1);
  extern __shared__ int s_kept[];  // row_cap: (entry << kGroup) | members
  constexpr int kParts = 8;        // waves of a block, parts of a row
  __shared__ int s_n[kParts];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int gi = blockIdx.x;
  const int ngroups = counts_in[1] / kGroup;
  if (gi >= ngroups) return;  // the whole block
  const int g0 = gi * kGroup;
  if (wave == 0 && lane < kGroup * kCluster) {
    const int atom = atom_at[g0 * kCluster + lane];
    if (atom >= 0)
      for (int k = 0; k < 3; ++k) x_prune[3 * atom + k] = x[3 * atom + k];
  }
  const int n = row_count[gi];
  const int part = (n + kParts - 1) / kParts;
  const int lo = wave * part < n ? wave * part : n;
  const int hi = lo + part < n ? lo + part : n;
  const int a = lane & 7, b = lane >> 3;
  float4 xi[kGroup];
#pragma unroll
  for (int m = 0; m < kGroup; ++m) xi[m] = xq[(g0 + m) * kCluster + a];
  const float4 oi = gorig[gi];
  const PairEntry *row = rows + (size_t)gi * row_cap;
  const int4 *row4 = reinterpret_cast<const int4 *>(row);
  int kept = 0;
  // 64 entries at a time, one per lane (cj, info, ex, pad), taken by
  // readlane; the j atoms of the next entry load while this one is tested
  for (int base = lo; base < hi; base += 64) {
    const int cnt = hi - base < 64 ? hi - base : 64;
    const int4 ent = row4[base + (lane < cnt ? lane : 0)];
    const int cj0 = __builtin_amdgcn_readlane(ent.x, 0);
    float4 xj = xq[cj0 * kCluster + b];
    float4 oj = gorig[cj0 / kGroup];
    for (int k = 0; k < cnt; ++k) {
      const int cj1 = __builtin_amdgcn_readlane(ent.x, k + 1 < cnt ? k + 1 : k);
      const float4 xj1 = xq[cj1 * kCluster + b];
      const float4 oj1 = gorig[cj1 / kGroup];
      const uint32_t info = (uint32_t)__builtin_amdgcn_readlane(ent.y, k);
      const float xjx = xj.x + group_offset(info, 0, oj.x, oi.x, Lx, Lx_lo);
      const float xjy = xj.y + group_offset(info, 1, oj.y, oi.y, Ly, Ly_lo);
      const float xjz = xj.z + group_offset(info, 2, oj.z, oi.z, Lz, Lz_lo);
      uint32_t members = 0;
#pragma unroll
      for (int m = 0; m < kGroup; ++m) {
        if (!((info >> (8 + m)) & 1u)) continue;  // uniform
        const float dx = xi[m].x - xjx, dy = xi[m].y - xjy, dz = xi[m].z - xjz;
        if (__ballot(dx * dx + dy * dy + dz * dz < reach2)) members |= 1u << m;
      }
      if (members) {
        if (lane == 0) s_kept[lo + kept] = ((base + k) << kGroup) | (int)members;
        ++kept;
      }
      xj = xj1;
      oj = oj1;
    }
  }
  if (lane == 0) s_n[wave] = kept;
  __syncthreads();
  int base = 0, total = 0;
  for (int w = 0; w < kParts; ++w) {
    if (w < wave) base += s_n[w];
    total += s_n[w];
  }
  PairEntry *out = rows_in + (size_t)gi * row_cap;
  for (int k = lane; k < kept; k += 64) {
    const int packed = s_kept[lo + k];
    PairEntry o = row[packed >> kGroup];
    o.info = (o.info & ~(((1u << kGroup) - 1u) << 8)) | ((uint32_t)(packed & ((1 << kGroup) - 1)) << 8);
    out[base + k] = o;
  }
  if (threadIdx.x == 0) row_count_in[gi] = total;
}

// ---- exports for parity tests -------------------------------------------

// For every tile lane with its mask bit set, the reference matrix test of
// NeighborsMatrix.mlir:176-183 and :394-412 on positions wrapped in f64 and
// rounded to f32 (each operation rounded on its own); pairs that pass are
// written as (min atom, max atom).  The host removes duplicates.
__global__ void __launch_bounds__(256) nb_export_pairs(
    const double *__restrict__ x, const int *__restrict__ atom_at,
    const uint32_t *__restrict__ real_mask, const PairEntry *__restrict__ rows,
    const int *__restrict__ row_count, const uint64_t *__restrict__ ex_masks,
    const int *__restrict__ counts_in, int row_cap, float Lx, float Ly, float Lz,
    double bx, double by, double bz, float limit2, int2 *__restrict__ out,
    unsigned long long *__restrict__ out_count, unsigned long long capacity) {
#pragma clang fp contract(off)
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int gi = blockIdx.x * 4 + wave;
  if (gi >= counts_in[1] / kGroup) return;
  const int a = lane & 7, b = lane >> 3;
  const float ilx = __fdiv_rn(1.0f, Lx), ily = __fdiv_rn(1.0f, Ly),
              ilz = __fdiv_rn(1.0f, Lz);
  const double L[3] = {bx, by, bz};
  const float Lf[3] = {Lx, Ly, Lz}, il[3] = {ilx, ily, ilz};
  const int n = row_count[gi];
  for (int m = 0; m < kGroup; ++m) {
    const int ci = gi * kGroup + m;
    const int ia = atom_at[ci * kCluster + a];
    const bool ireal = (real_mask[ci] >> a) & 1u;
    float wi[3] = {0, 0, 0};
    if (ia >= 0)
      for (int k = 0; k < 3; ++k) {
        const double v = x[3 * ia + k];
        wi[k] = (float)(v - floor(v * (1.0 / L[k])) * L[k]);
      }
    for (int e = 0; e < n; ++e) {
      const PairEntry E = rows[(size_t)gi * row_cap + e];
      if (!((E.info >> (8 + m)) & 1u)) continue;
      const bool on = E.ex >= 0 ? ((ex_masks[(size_t)E.ex * kGroup + m] >> lane) & 1ull)
                                : (ireal && ((E.info >> (16 + b)) & 1u));
      if (!on) continue;
      const int jb = atom_at[E.cj * kCluster + b];
      if (ia < 0 || jb < 0) continue;
      float wj[3];
      for (int k = 0; k < 3; ++k) {
        const double v = x[3 * jb + k];
        wj[k] = (float)(v - floor(v * (1.0 / L[k])) * L[k]);
      }
      float d2[3];
      for (int k = 0; k < 3; ++k) {
        const float d = __fsub_rn(wi[k], wj[k]);
        const float images = __fmul_rn(d, il[k]);
        const float nearest = rintf(images);
        const float sh = __fmul_rn(nearest, Lf[k]);
        const float dd = __fsub_rn(d, sh);
        d2[k] = __fmul_rn(dd, dd);
      }
      const float r2 = __fadd_rn(__fadd_rn(d2[0], d2[1]), d2[2]);
      if (r2 < limit2) {
        const unsigned long long slot = atomicAdd(out_count, 1ull);
        if (slot < capacity) out[slot] = make_int2(ia < jb ? ia : jb, ia < jb ? jb : ia);
      }
    }
  }
}

}  // namespace mdhip
