// nbnxm.hip — cell-sorted cluster build for gfx950.
//
// Replaces the reference's neighbor builds (lib/Runtime/Templates/
// NeighborsMatrix.mlir:193-460 for the pair set it defines,
// lib/Conversion/MDExecKernels/GroupsBuild.cpp:2515 for the GPU structure)
// with a layout made for wave64: atoms are binned into xy columns, each
// column is sorted along z, and every 8 consecutive places form a cluster
// with a bounding box.  A wave searches the clusters near one i-cluster and
// writes a row of {j-cluster, periodic shift, 64-bit interaction mask}: an
// 8 x 8 tile is one pair per lane of a wave64.
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
__global__ void nb_request_build(int *flags) {
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
                                             int *__restrict__ flags, GridParams g) {
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
// places, f32 positions about the box centre, padding, and cluster boxes.
__global__ void __launch_bounds__(256) nb_sort(
    const double *__restrict__ x, const double *__restrict__ shift,
    const double *__restrict__ q, const int *__restrict__ type,
    const int *__restrict__ col_count, const int *__restrict__ bucket,
    int *__restrict__ atom_at, int *__restrict__ place_of,
    float4 *__restrict__ xq, int *__restrict__ type_at,
    float4 *__restrict__ bb_lo, float4 *__restrict__ bb_hi,
    int *__restrict__ cluster_col, uint32_t *__restrict__ real_mask,
    int *__restrict__ col_cl_start, int *__restrict__ col_cl_count,
    int *__restrict__ counts_out, int *__restrict__ flags, GridParams g) {
  if (flags[kFlagRebuild] == 0) return;
  __shared__ int s_red[256];
  __shared__ int s_atom[kSortCapMax];
  __shared__ double s_z[kSortCapMax];
  __shared__ float4 s_pos[kSortCapMax];
  const int col = blockIdx.x, tid = threadIdx.x;
  const int cap = g.col_cap < kSortCapMax ? g.col_cap : kSortCapMax;

  int sum = 0;
  for (int c = tid; c < col; c += blockDim.x) {
    int cnt = col_count[c];
    cnt = cnt < cap ? cnt : cap;
    sum += (cnt + kCluster - 1) & ~(kCluster - 1);
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
  const int padded = (count + kCluster - 1) & ~(kCluster - 1);

  for (int t = tid; t < count; t += blockDim.x) {
    int a = bucket[col * g.col_cap + t];
    s_atom[t] = a;
    s_z[t] = x[3 * a + 2] - shift[3 * a + 2];
  }
  __syncthreads();
  for (int t = tid; t < count; t += blockDim.x) {
    double za = s_z[t];
    int ia = s_atom[t];
    int rank = 0;
    for (int u = 0; u < count; ++u) {
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
    xq[place] = p;
    type_at[place] = type[ia];
    s_pos[rank] = p;
  }
  for (int t = count + tid; t < padded; t += blockDim.x) {
    int place = start + t;
    atom_at[place] = -1;
    // far from everything and from each other; masked out of every tile
    xq[place] = make_float4(1.0e5f, 1.0e5f, 1.0e5f + (float)(place & 1023), 0.0f);
    type_at[place] = 0;
  }
  __syncthreads();
  const int cl_start = start / kCluster, ncl = padded / kCluster;
  if (tid == 0) {
    col_cl_start[col] = cl_start;
    col_cl_count[col] = ncl;
    if (col == 0) atomicAdd(&flags[kFlagBuilds], 1);
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
}

__device__ __forceinline__ float axis_gap(float lo_i, float hi_i, float lo_j,
                                          float hi_j) {
  float g1 = lo_j - hi_i, g2 = lo_i - hi_j;
  float g = fmaxf(g1, g2);
  return g > 0.0f ? g : 0.0f;
}

// Wave per i-cluster: candidates are the clusters of the columns within the
// reach in x and y, at z shifts -1, 0, +1; a candidate is kept when its box
// is within the reach of the i box (each unordered cluster pair once, from
// its lower index).  The 64-bit mask drops padding, i >= j within the same
// cluster at zero shift, and the excluded pairs (MDExecKernels.cpp:1748-1823
// removes the same partners from a matrix row).
__global__ void __launch_bounds__(256) nb_search(
    const float4 *__restrict__ bb_lo, const float4 *__restrict__ bb_hi,
    const int *__restrict__ cluster_col, const uint32_t *__restrict__ real_mask,
    const int *__restrict__ col_cl_start, const int *__restrict__ col_cl_count,
    const int *__restrict__ atom_at, const int *__restrict__ place_of,
    const int *__restrict__ excl_start, const int *__restrict__ excl_atoms,
    const int *__restrict__ counts_in, int *__restrict__ col_count,
    PairEntry *__restrict__ rows, int *__restrict__ row_count,
    int *__restrict__ flags, GridParams g) {
  if (flags[kFlagRebuild] == 0) return;
  // The next build starts from empty buckets (nb_sort has read them).
  if (blockIdx.x == 0)
    for (int c = threadIdx.x; c < g.ncol; c += blockDim.x) col_count[c] = 0;
  constexpr int kExMax = 8 * 16;
  __shared__ int s_ex_cluster[4][kExMax];
  __shared__ int s_ex_bit[4][kExMax];
  __shared__ int s_ex_n[4];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int ci = blockIdx.x * 4 + wave;
  const int nclusters = counts_in[1];
  const bool active = ci < nclusters;
  if (lane == 0) s_ex_n[wave] = 0;
  __syncthreads();

  // exclusion table of the i cluster: (partner cluster, mask bit)
  if (active && lane < kCluster) {
    int atom = atom_at[ci * kCluster + lane];
    if (atom >= 0) {
      int b = excl_start[atom], e = excl_start[atom + 1];
      for (int k = b; k < e; ++k) {
        int pp = place_of[excl_atoms[k]];
        int slot = atomicAdd(&s_ex_n[wave], 1);
        if (slot < kExMax) {
          s_ex_cluster[wave][slot] = pp / kCluster;
          s_ex_bit[wave][slot] = lane + kCluster * (pp % kCluster);
        } else {
          atomicOr(&flags[kFlagOverflowRow], 2);
        }
      }
    }
  }
  __syncthreads();
  if (!active) return;
  const int nex = s_ex_n[wave] < kExMax ? s_ex_n[wave] : kExMax;

  const float4 ilo = bb_lo[ci], ihi = bb_hi[ci];
  const uint32_t ireal = real_mask[ci];
  const int icol = cluster_col[ci];
  const int cx = icol % g.ncx, cy = icol / g.ncx;
  const float Lx = (float)g.box[0], Ly = (float)g.box[1], Lz = (float)g.box[2];
  const float wxf = (float)g.wx, wyf = (float)g.wy;
  int count = 0;
  PairEntry *row = rows + (size_t)ci * g.row_cap;
  // columns within reach: lanes test 64 column offsets at a time, a wave
  // prefix sum of their candidate counts (clusters x 3 z shifts) lays the
  // candidates out, and lanes then test 64 candidates at a time; every load
  // of an iteration is independent of the others.
  __shared__ int s_cstart[4][64], s_cbase[4][64], s_ccode[4][64];
  const int ncolx = 2 * g.rx + 1, ncols = ncolx * (2 * g.ry + 1);
  for (int cbase = 0; cbase < ncols; cbase += 64) {
    const int c = cbase + lane;
    int ncand = 0, cl0 = 0, code = 0;
    if (c < ncols) {
      const int dy = c / ncolx - g.ry, dx = c % ncolx - g.rx;
      int jy = cy + dy, sy = 0, jx = cx + dx, sx = 0;
      while (jy < 0) { jy += g.ncy; --sy; }
      while (jy >= g.ncy) { jy -= g.ncy; ++sy; }
      while (jx < 0) { jx += g.ncx; --sx; }
      while (jx >= g.ncx) { jx -= g.ncx; ++sx; }
      const float ylo = jy * wyf - 0.5f * Ly + sy * Ly, yhi = ylo + wyf;
      const float xlo = jx * wxf - 0.5f * Lx + sx * Lx, xhi = xlo + wxf;
      const float gy = axis_gap(ilo.y, ihi.y, ylo, yhi);
      const float gx = axis_gap(ilo.x, ihi.x, xlo, xhi);
      if (gx * gx + gy * gy < g.rlist2 && sx >= -1 && sx <= 1 && sy >= -1 && sy <= 1) {
        const int jcol = jy * g.ncx + jx;
        cl0 = col_cl_start[jcol];
        ncand = col_cl_count[jcol] * 3;
        code = (sx + 1) | ((sy + 1) << 2);
      }
    }
    // inclusive wave scan of ncand
    int scan = ncand;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const int up = __shfl_up(scan, off);
      if (lane >= off) scan += up;
    }
    const int total = __shfl(scan, 63);
    s_cstart[wave][lane] = cl0;
    s_cbase[wave][lane] = scan - ncand;  // exclusive
    s_ccode[wave][lane] = code;
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    for (int base = 0; base < total; base += 64) {
      const int t = base + lane;
      bool accept = false;
      int cj = 0;
      uint32_t scode = 0;
      uint64_t mask = 0;
      if (t < total) {
        // the column slot holding candidate t: last slot with cbase <= t
        int lo = 0, hi = 63;
        while (lo < hi) {
          const int mid = (lo + hi + 1) >> 1;
          if (s_cbase[wave][mid] <= t) lo = mid; else hi = mid - 1;
        }
        const int within = t - s_cbase[wave][lo];
        cj = s_cstart[wave][lo] + within / 3;
        const int sz = within % 3 - 1;
        const int ccode = s_ccode[wave][lo];
        const int sx = (ccode & 3) - 1, sy = ((ccode >> 2) & 3) - 1;
        const int lex = sz != 0 ? sz : (sy != 0 ? sy : sx);
        const bool ordered = cj > ci || (cj == ci && lex >= 0);
        if (ordered) {
          const float4 jlo = bb_lo[cj], jhi = bb_hi[cj];
          const float ox = sx * Lx, oy = sy * Ly, oz = sz * Lz;
          const float ex = axis_gap(ilo.x, ihi.x, jlo.x + ox, jhi.x + ox);
          const float ey = axis_gap(ilo.y, ihi.y, jlo.y + oy, jhi.y + oy);
          const float ez = axis_gap(ilo.z, ihi.z, jlo.z + oz, jhi.z + oz);
          if (ex * ex + ey * ey + ez * ez < g.rlist2) {
            const uint32_t jreal = real_mask[cj];
            for (int b = 0; b < kCluster; ++b)
              if (jreal & (1u << b)) mask |= (uint64_t)ireal << (kCluster * b);
            if (cj == ci) {
              uint64_t keep = 0;
              for (int b = 0; b < kCluster; ++b)
                for (int a = 0; a < kCluster; ++a)
                  if (lex == 0 ? (a < b) : (a != b)) keep |= 1ull << (a + kCluster * b);
              mask &= keep;
            }
            for (int k = 0; k < nex; ++k)
              if (s_ex_cluster[wave][k] == cj) mask &= ~(1ull << s_ex_bit[wave][k]);
            scode = (uint32_t)(sx + 1) | ((uint32_t)(sy + 1) << 2) | ((uint32_t)(sz + 1) << 4);
            accept = mask != 0;
          }
        }
      }
      const uint64_t ballot = __ballot(accept);
      const int before = __popcll(ballot & ((1ull << lane) - 1ull));
      if (accept) {
        const int slot = count + before;
        if (slot < g.row_cap) {
          PairEntry e;
          e.cj = cj;
          e.shift = scode;
          e.mask = mask;
          row[slot] = e;
        } else {
          atomicOr(&flags[kFlagOverflowRow], 1);
        }
      }
      count += __popcll(ballot);
    }
    __builtin_amdgcn_wave_barrier();
  }
  if (lane == 0) row_count[ci] = count < g.row_cap ? count : g.row_cap;
}

// ---- exports for parity tests -------------------------------------------

// For every tile lane with its mask bit set, the reference matrix test of
// NeighborsMatrix.mlir:176-183 and :394-412 on positions wrapped in f64 and
// rounded to f32 (each operation rounded on its own); pairs that pass are
// written as (min atom, max atom).  The host removes duplicates.
__global__ void __launch_bounds__(256) nb_export_pairs(
    const double *__restrict__ x, const int *__restrict__ atom_at,
    const PairEntry *__restrict__ rows, const int *__restrict__ row_count,
    const int *__restrict__ counts_in, int row_cap, float Lx, float Ly,
    float Lz, double bx, double by, double bz, float limit2,
    int2 *__restrict__ out, unsigned long long *__restrict__ out_count,
    unsigned long long capacity) {
#pragma clang fp contract(off)
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int ci = blockIdx.x * 4 + wave;
  if (ci >= counts_in[1]) return;
  const int a = lane & 7, b = lane >> 3;
  const int ia = atom_at[ci * kCluster + a];
  const float ilx = __fdiv_rn(1.0f, Lx), ily = __fdiv_rn(1.0f, Ly),
              ilz = __fdiv_rn(1.0f, Lz);
  float wi[3] = {0, 0, 0};
  if (ia >= 0) {
    double L[3] = {bx, by, bz};
    for (int k = 0; k < 3; ++k) {
      double v = x[3 * ia + k];
      wi[k] = (float)(v - floor(v * (1.0 / L[k])) * L[k]);
    }
  }
  const int n = row_count[ci];
  for (int e = 0; e < n; ++e) {
    PairEntry E = rows[(size_t)ci * row_cap + e];
    if (!((E.mask >> lane) & 1ull)) continue;
    int jb = atom_at[E.cj * kCluster + b];
    if (ia < 0 || jb < 0) continue;
    double L[3] = {bx, by, bz};
    float wj[3];
    for (int k = 0; k < 3; ++k) {
      double v = x[3 * jb + k];
      wj[k] = (float)(v - floor(v * (1.0 / L[k])) * L[k]);
    }
    float Lf[3] = {Lx, Ly, Lz}, il[3] = {ilx, ily, ilz};
    float d2[3];
    for (int k = 0; k < 3; ++k) {
      float d = __fsub_rn(wi[k], wj[k]);
      float images = __fmul_rn(d, il[k]);
      float nearest = rintf(images);
      float sh = __fmul_rn(nearest, Lf[k]);
      float dd = __fsub_rn(d, sh);
      d2[k] = __fmul_rn(dd, dd);
    }
    float r2 = __fadd_rn(__fadd_rn(d2[0], d2[1]), d2[2]);
    if (r2 < limit2) {
      unsigned long long slot = atomicAdd(out_count, 1ull);
      if (slot < capacity)
        out[slot] = make_int2(ia < jb ? ia : jb, ia < jb ? jb : ia);
    }
  }
}

}  // namespace mdhip
