/* oracle.c — FP64 CPU restatement of the reference's per-step force path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing under mdir_amd/ links, loads or calls
 * this library.  tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg use it as the checker the HIP path is compared against.
 *
 * Each function cites the reference file:line whose semantics it restates.
 * How each one is pinned (golden vectors, closed forms, brute force,
 * oracle/_ref) is stated beside it; tests/test_oracle_*.py hold the checks.
 *
 * Built with -ffp-contract=off so that the f32 neighbor search rounds each
 * operation as the reference's template does.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ======================================================================
 * Neighbor matrix — lib/Runtime/Templates/NeighborsMatrix.mlir:36-460.
 * Pinned by the golden counts of test/Runtime/neighbors-matrix.mlir:111-160
 * and neighbors-matrix-gpu.mlir:148-186 (tests/test_oracle_golden.py)
 * and by brute force over all pairs.
 * ====================================================================== */

/* NeighborsMatrix.mlir:36-56 (@mdrt.cell_count). */
int64_t orc_cell_count(double length, double width) {
  double padded = width * 1.0001;
  double ratio = length / padded;
  double fl = floor(ratio);
  double least = fmax(fl, 1.0);
  double bounded = fmin(least, 256.0);
  int64_t count = (int64_t)bounded;
  return count < 1 ? 1 : count;
}

/* NeighborsMatrix.mlir:58-82 (@mdrt.cell_range). */
int64_t orc_cell_range(double length, int64_t count, double reach) {
  double width = length / (double)count;
  double ratio = reach / width;
  double padded = ratio * 1.00005;
  double c = ceil(padded);
  double least = fmax(c, 1.0);
  double bounded = fmin(least, 256.0);
  int64_t range = (int64_t)bounded;
  return range < 1 ? 1 : range;
}

/* NeighborsMatrix.mlir:84-147 (@mdrt.cell_width). */
double orc_cell_width(int64_t count, const double box[3], double reach,
                      double least) {
  double particles = (double)count;
  double best = 1.0e30, chosen = least;
  for (int parts = 1; parts < 4; ++parts) {
    double narrow = reach / (double)parts;
    double tried = narrow > least ? narrow : least;
    int64_t n[3], s[3];
    for (int k = 0; k < 3; ++k) {
      n[k] = orc_cell_count(box[k], tried);
      int64_t r = orc_cell_range(box[k], n[k], reach);
      int64_t full = 2 * r + 1;
      s[k] = n[k] < full ? n[k] : full;
    }
    double rows = (double)(s[1] * s[2]);
    double visited = rows * (double)s[0];
    double cells = (double)(n[0] * n[1] * n[2]);
    double total = 12.0 * rows + particles * (visited / cells);
    if (total < best) {
      best = total;
      chosen = tried;
    }
  }
  return chosen;
}

/* NeighborsMatrix.mlir:152-183: wrap, cell coordinate, minimum image. */
static double wrap_coordinate(double x, double length, double inverse) {
  return x - floor(x * inverse) * length;
}
static int64_t cell_coordinate(double wrapped, double inverse, int64_t count) {
  double position = (wrapped * inverse) * (double)count;
  int64_t c = (int64_t)position; /* fptosi truncates */
  if (c > count - 1) c = count - 1;
  if (c < 0) c = 0;
  return c;
}
static float minimum_image_f32(float d, float length, float inverse) {
  volatile float images = d * inverse;
  float nearest = nearbyintf(images); /* round half to even (default mode) */
  volatile float shift = nearest * length;
  return d - shift;
}

/* NeighborsMatrix.mlir:185-460 (@mdrt.build_neighbors_matrix). */
int64_t orc_neighbors_matrix(int64_t n, const double *x, const double box[3],
                             double reach, double cell_width, int32_t *counts,
                             int32_t *index, int64_t row_width) {
  double L[3] = {box[0], box[1], box[2]};
  double inv[3] = {1.0 / L[0], 1.0 / L[1], 1.0 / L[2]};
  int64_t nc[3];
  for (int k = 0; k < 3; ++k) nc[k] = orc_cell_count(L[k], cell_width);
  int64_t cells = nc[0] * nc[1] * nc[2];
  float Lf[3], ilf[3];
  for (int k = 0; k < 3; ++k) {
    Lf[k] = (float)L[k];
    ilf[k] = 1.0f / Lf[k];
  }
  double lxyz = (L[0] + L[1]) + L[2];
  double margin = lxyz * 3.0e-6;
  double far = reach + margin;
  float limit2 = (float)(far * far);

  int64_t *key = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
  int64_t *start = (int64_t *)calloc((size_t)cells + 1, sizeof(int64_t));
  int64_t *cursor = (int64_t *)malloc(sizeof(int64_t) * (size_t)cells);
  int64_t *order = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
  float *wrapped = (float *)malloc(sizeof(float) * 3 * (size_t)(n > 0 ? n : 1));
  float *sorted = (float *)malloc(sizeof(float) * 3 * (size_t)(n > 0 ? n : 1));

  for (int64_t i = 0; i < n; ++i) {
    int64_t c[3];
    for (int k = 0; k < 3; ++k) {
      double w = wrap_coordinate(x[3 * i + k], L[k], inv[k]);
      wrapped[3 * i + k] = (float)w;
      c[k] = cell_coordinate(w, inv[k], nc[k]);
    }
    key[i] = (c[2] * nc[1] + c[1]) * nc[0] + c[0];
    start[key[i] + 1] += 1;
  }
  for (int64_t c = 0; c < cells; ++c) {
    start[c + 1] += start[c];
    cursor[c] = start[c];
  }
  for (int64_t i = 0; i < n; ++i) {
    int64_t p = cursor[key[i]]++;
    order[p] = i;
    for (int k = 0; k < 3; ++k) sorted[3 * p + k] = wrapped[3 * i + k];
  }

  int64_t range[3], span[3], first[3];
  for (int k = 0; k < 3; ++k) {
    range[k] = orc_cell_range(L[k], nc[k], reach);
    int64_t full = 2 * range[k] + 1;
    span[k] = nc[k] < full ? nc[k] : full;
    first[k] = (span[k] == full) ? nc[k] - range[k] : 0;
  }

#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t p = 0; p < n; ++p) {
    int64_t i = order[p];
    float xi = sorted[3 * p], yi = sorted[3 * p + 1], zi = sorted[3 * p + 2];
    int64_t k = key[i];
    int64_t cx = k % nc[0], rest = k / nc[0];
    int64_t cy = rest % nc[1], cz = rest / nc[1];
    int64_t count = 0;
    for (int64_t oz = 0; oz < span[2]; ++oz) {
      int64_t zc = (cz + first[2] + oz) % nc[2];
      for (int64_t oy = 0; oy < span[1]; ++oy) {
        int64_t yc = (cy + first[1] + oy) % nc[1];
        int64_t row_first = (zc * nc[1] + yc) * nc[0];
        int64_t xb = (cx + first[0]) % nc[0];
        int64_t xe = xb + span[0];
        int around = xe > nc[0];
        int64_t xstop = around ? nc[0] : xe;
        int64_t xrest = around ? xe - nc[0] : 0;
        for (int part = 0; part < 2; ++part) {
          int64_t from = part ? 0 : xb, to = part ? xrest : xstop;
          int64_t b = start[row_first + from], e = start[row_first + to];
          for (int64_t q = b; q < e; ++q) {
            float dx = minimum_image_f32(xi - sorted[3 * q], Lf[0], ilf[0]);
            float dy = minimum_image_f32(yi - sorted[3 * q + 1], Lf[1], ilf[1]);
            float dz = minimum_image_f32(zi - sorted[3 * q + 2], Lf[2], ilf[2]);
            volatile float dx2 = dx * dx, dy2 = dy * dy, dz2 = dz * dz;
            volatile float dxy2 = dx2 + dy2;
            float r2 = dxy2 + dz2;
            int neighbor = (r2 < limit2) && (p != q);
            if (neighbor && count < row_width)
              index[i * row_width + count] = (int32_t)order[q];
            if (neighbor) ++count;
          }
        }
      }
    }
    counts[i] = (int32_t)count;
  }

  int64_t largest = 0;
  for (int64_t i = 0; i < n; ++i) {
    if (counts[i] > largest) largest = counts[i];
    if (counts[i] > row_width) counts[i] = (int32_t)row_width;
  }
  free(key);
  free(start);
  free(cursor);
  free(order);
  free(wrapped);
  free(sorted);
  return largest;
}

/* lib/Conversion/MDExecKernels/MDExecKernels.cpp:1748-1823: a row that
   fits keeps its entries in order, less the excluded partners. */
void orc_exclusion_filter(int64_t n, int32_t *counts, int32_t *index,
                          int64_t row_width, const int32_t *excl_start,
                          const int32_t *excl_atoms) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    int32_t b = excl_start[i], e = excl_start[i + 1];
    if (counts[i] > row_width || b == e) continue;
    int32_t written = 0;
    for (int32_t r = 0; r < counts[i]; ++r) {
      int32_t j = index[i * row_width + r];
      int excluded = 0;
      for (int32_t k = b; k < e; ++k) excluded |= (excl_atoms[k] == j);
      if (!excluded) index[i * row_width + written++] = j;
    }
    counts[i] = written;
  }
}

/* ======================================================================
 * Pair physics.
 * ====================================================================== */

/* Lennard-Jones energy and dU/dr under the modifiers of
   lib/Driver/TermLennardJones.cpp:66-168 and the potential switch
   1 - 10t^3 + 15t^4 - 6t^5 of ExpandTruncation.cpp:19-45.
   Pinned by the argon lattice energy of test/Driver/argon.test:8-12 (switch),
   by finite differences of U (all modifiers), and by U(rc)=0 / U'(rc)=0
   where each modifier promises it. */
static void lj_pair(double sigma, double eps, double r, double rc, double rs,
                    int modifier, double *u, double *dudr) {
  if (eps == 0.0 || sigma == 0.0) {
    *u = 0.0;
    *dudr = 0.0;
    return;
  }
  double e4 = 4.0 * eps;
  double sr = sigma / r;
  double s3 = sr * sr * sr, s6 = s3 * s3, s12 = s6 * s6;
  if (modifier == ORC_LJ_NONE || modifier == ORC_LJ_POTENTIAL_SWITCH) {
    double raw = e4 * (s12 - s6);
    double draw = e4 * (-12.0 * s12 + 6.0 * s6) / r;
    if (modifier == ORC_LJ_NONE || r <= rs) {
      *u = raw;
      *dudr = draw;
      return;
    }
    double width = rc - rs;
    double t = (r - rs) / width;
    double S = 1.0 + t * t * t * (-10.0 + t * (15.0 - 6.0 * t));
    double dS = (-30.0 * t * t + 60.0 * t * t * t - 30.0 * t * t * t * t) / width;
    *u = raw * S;
    *dudr = draw * S + raw * dS;
    return;
  }
  double g3 = sigma * sigma * sigma, g6 = g3 * g3, g12 = g6 * g6;
  if (modifier == ORC_LJ_POTENTIAL_SHIFT) {
    double a = s12 - g12 * pow(rc, -12.0);
    double b = s6 - g6 * pow(rc, -6.0);
    *u = e4 * (a - b);
    *dudr = e4 * (-12.0 * s12 + 6.0 * s6) / r;
    return;
  }
  /* ORC_LJ_FORCE_SWITCH: U = 4 eps [sigma^12 (r^-12 - P12) - sigma^6 (r^-6 - P6)],
     P_n = c_n + u^3 (a_n/3 + b_n u/4), u = max(r - rs, 0). */
  double width = rc - rs;
  double uu = r > rs ? r - rs : 0.0;
  double P[2], dP[2];
  int ns[2] = {12, 6};
  for (int k = 0; k < 2; ++k) {
    double nn = ns[k];
    double an = -nn * ((nn + 4.0) * rc - (nn + 1.0) * rs) /
                (pow(rc, nn + 2.0) * width * width);
    double bn = nn * ((nn + 3.0) * rc - (nn + 1.0) * rs) /
                (pow(rc, nn + 2.0) * width * width * width);
    double cn = pow(rc, -nn) - an / 3.0 * pow(width, 3.0) -
                bn / 4.0 * pow(width, 4.0);
    P[k] = cn + uu * uu * uu * (an / 3.0 + bn / 4.0 * uu);
    dP[k] = an * uu * uu + bn * uu * uu * uu;
  }
  double rinv = 1.0 / r;
  double r6 = pow(rinv, 6.0), r12 = r6 * r6;
  *u = e4 * (g12 * (r12 - P[0]) - g6 * (r6 - P[1]));
  *dudr = e4 * (g12 * (-12.0 * r12 * rinv - dP[0]) -
                g6 * (-6.0 * r6 * rinv - dP[1]));
}

static void lj_pair_f32(float sigma, float eps, float r, float rc, float rs,
                        int modifier, float *u, float *dudr) {
  if (eps == 0.0f || sigma == 0.0f) {
    *u = 0.0f;
    *dudr = 0.0f;
    return;
  }
  /* Mixed mode of the CPU baseline: NONE / POTENTIAL_SHIFT in f32, the
     switches fall back to f64 (timing path only). */
  if (modifier == ORC_LJ_NONE || modifier == ORC_LJ_POTENTIAL_SHIFT) {
    float e4 = 4.0f * eps, sr = sigma / r;
    float s3 = sr * sr * sr, s6 = s3 * s3, s12 = s6 * s6;
    float shift = 0.0f;
    if (modifier == ORC_LJ_POTENTIAL_SHIFT) {
      float sc = sigma / rc, c3 = sc * sc * sc, c6 = c3 * c3;
      shift = c6 * c6 - c6;
    }
    *u = e4 * (s12 - s6 - shift);
    *dudr = e4 * (-12.0f * s12 + 6.0f * s6) / r;
    return;
  }
  double uu, dd;
  lj_pair(sigma, eps, r, rc, rs, modifier, &uu, &dd);
  *u = (float)uu;
  *dudr = (float)dd;
}

static inline double min_image(double d, double L) {
  return d - L * nearbyint(d / L);
}

/* The primary pair sum (TermNonbonded.cpp:1423-1556): LJ + f qq
   [erfc(beta r)/r - s] over the non-excluded pairs within rc, rows of
   the neighbor matrix summed owner-computes (MDExecToLoops.cpp:339-428).
   Pinned by the three-OPC-water direct sum of test/Driver/pme.test:20
   (-14.3572978 kcal/mol), by an all-pairs brute force, and by finite
   differences of the energy. */
int orc_nonbonded(int64_t n, const double *x, const double *q,
                  const int32_t *type, int32_t ntypes, const double *sigma,
                  const double *epsilon, const double box[3], double cutoff,
                  double switch_from, int32_t lj_modifier,
                  int32_t coulomb_shift, double beta, double coulomb,
                  const int32_t *excl_start, const int32_t *excl_atoms,
                  int32_t mixed, double *F, double *energies, double *virial) {
  int64_t width = 64;
  int32_t *counts = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
  int32_t *index = NULL;
  double cw = orc_cell_width(n, box, cutoff, 0.0);
  for (;;) {
    free(index);
    index = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n * width + 1));
    int64_t largest =
        orc_neighbors_matrix(n, x, box, cutoff, cw, counts, index, width);
    if (largest <= width) break;
    width = (int64_t)(largest * 1.25) + 1; /* mdrt_cuda.c:3361-3371 grows */
  }
  if (excl_start) orc_exclusion_filter(n, counts, index, width, excl_start, excl_atoms);

  double shift_c = coulomb_shift ? erfc(beta * cutoff) / cutoff : 0.0;
  double two_over_sqrt_pi = 2.0 / sqrt(M_PI);
  /* Per-row partial sums, added in row order afterwards, so the result does
     not depend on the number of threads (07-precision.md:61-75). */
  double *rows = (double *)calloc((size_t)(n > 0 ? n : 1) * 11, sizeof(double));
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t i = 0; i < n; ++i) {
    double fi[3] = {0, 0, 0};
    double *row = rows + 11 * i;
    for (int32_t r_ = 0; r_ < counts[i]; ++r_) {
      int64_t j = index[i * width + r_];
      double d[3];
      for (int k = 0; k < 3; ++k) d[k] = min_image(x[3 * i + k] - x[3 * j + k], box[k]);
      double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
      if (!(r2 < cutoff * cutoff)) continue;
      int32_t ti = type[i], tj = type[j];
      double sg = sigma[ti * ntypes + tj], ep = epsilon[ti * ntypes + tj];
      double fscal; /* -dU/dr / r */
      if (!mixed) {
        double r = sqrt(r2);
        double ulj, dlj;
        lj_pair(sg, ep, r, cutoff, switch_from, lj_modifier, &ulj, &dlj);
        double uc = 0.0, dc = 0.0;
        if (beta > 0.0) {
          double qq = coulomb * q[i] * q[j];
          double er = erfc(beta * r);
          uc = qq * (er / r - shift_c);
          dc = -qq * (er / r2 + beta * two_over_sqrt_pi * exp(-beta * beta * r2) / r);
        }
        row[0] += 0.5 * ulj;
        row[1] += 0.5 * uc;
        fscal = -(dlj + dc) / r;
      } else {
        float r2f = (float)r2;
        float rf = sqrtf(r2f);
        float ulj, dlj;
        lj_pair_f32((float)sg, (float)ep, rf, (float)cutoff, (float)switch_from,
                    lj_modifier, &ulj, &dlj);
        float uc = 0.0f, dc = 0.0f;
        if (beta > 0.0) {
          float qq = (float)(coulomb * q[i] * q[j]);
          float bf = (float)beta;
          float er = erfcf(bf * rf);
          uc = qq * (er / rf - (float)shift_c);
          dc = -qq * (er / r2f + bf * (float)two_over_sqrt_pi * expf(-bf * bf * r2f) / rf);
        }
        row[0] += 0.5 * ulj;
        row[1] += 0.5 * uc;
        fscal = -(double)((dlj + dc) / rf);
      }
      for (int k = 0; k < 3; ++k) fi[k] += fscal * d[k];
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) row[2 + 3 * a + b] += 0.5 * d[a] * fscal * d[b];
    }
    for (int k = 0; k < 3; ++k) F[3 * i + k] += fi[k];
  }
  for (int64_t i = 0; i < n; ++i) {
    energies[ORC_E_LJ] += rows[11 * i];
    energies[ORC_E_COUL_DIRECT] += rows[11 * i + 1];
    if (virial)
      for (int k = 0; k < 9; ++k) virial[k] += rows[11 * i + 2 + k];
  }
  free(rows);
  free(counts);
  free(index);
  return 0;
}

/* TermNonbonded.cpp:1615-1671: -f qq erf(beta r)/r over every excluded pair.
   Pinned by test/Driver/pme.test:21 (535.1260682 kcal/mol). */
void orc_excluded(int64_t npairs, const int32_t *pairs, const double *x,
                  const double *q, const double box[3], double beta,
                  double coulomb, double *F, double *energy) {
  double e = 0.0;
  double tsp = 2.0 / sqrt(M_PI);
  for (int64_t p = 0; p < npairs; ++p) {
    int32_t i = pairs[2 * p], j = pairs[2 * p + 1];
    double d[3];
    for (int k = 0; k < 3; ++k) d[k] = min_image(x[3 * i + k] - x[3 * j + k], box[k]);
    double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    double r = sqrt(r2);
    double qq = coulomb * q[i] * q[j];
    double ef = erf(beta * r);
    e += -qq * ef / r;
    double dudr = -qq * (beta * tsp * exp(-beta * beta * r2) / r - ef / r2);
    double fscal = -dudr / r;
    for (int k = 0; k < 3; ++k) {
      F[3 * i + k] += fscal * d[k];
      F[3 * j + k] -= fscal * d[k];
    }
  }
  *energy += e;
}

/* TermPairs14.cpp:11-17, :21-66: LJ(sigma, eps*s_LJ) under the modifier
   (no cutoff test: tuples have none) plus f qq s_C / r. */
void orc_pairs14(int64_t npairs, const int32_t *pairs, const double *sigma,
                 const double *eps_scaled, const double *qq_scaled,
                 const double *x, const double box[3], double cutoff,
                 double switch_from, int32_t lj_modifier, double *F,
                 double *e_lj, double *e_coul) {
  double el = 0.0, ec = 0.0;
  for (int64_t p = 0; p < npairs; ++p) {
    int32_t i = pairs[2 * p], j = pairs[2 * p + 1];
    double d[3];
    for (int k = 0; k < 3; ++k) d[k] = min_image(x[3 * i + k] - x[3 * j + k], box[k]);
    double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    double u, dudr;
    lj_pair(sigma[p], eps_scaled[p], r, cutoff, switch_from, lj_modifier, &u, &dudr);
    double uc = qq_scaled[p] / r;
    el += u;
    ec += uc;
    double fscal = -(dudr - uc / r) / r;
    for (int k = 0; k < 3; ++k) {
      F[3 * i + k] += fscal * d[k];
      F[3 * j + k] -= fscal * d[k];
    }
  }
  *e_lj += el;
  *e_coul += ec;
}

static void cross3(const double *a, const double *b, double *c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}
static double dot3(const double *a, const double *b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

/* ---- bonded tuples ------------------------------------------------------
 * The energies of lib/Driver/TermBond.cpp:11-37 (bond), TermUreyBradley.cpp
 * (the 1-3 distance), TermAngle.cpp:10-37, TermDihedral.cpp:14-58 and
 * TermImproper.cpp:10-45, over the coordinates of
 * lib/Dialect/MD/MDCoordinates.cpp:305-330 (distance; angle = acos of the
 * clamped cosine of the arms from the apex; dihedral = atan2 of the
 * Blondel-Karplus sine and cosine) and their gradients (:336-400). */
static void bonded_sub(const double *a, const double *b, const double box[3], double *d) {
  for (int k = 0; k < 3; ++k) d[k] = min_image(a[k] - b[k], box[k]);
}

void orc_bonded(int64_t ntuples, const int32_t *atoms, const int32_t *kind,
                const double *param, const double *x, const double box[3], double *F,
                double *energies) {
  for (int64_t t = 0; t < ntuples; ++t) {
    const int32_t *at = atoms + 4 * t;
    const double *p = param + 4 * t;
    const double *xa = x + 3 * at[0], *xb = x + 3 * at[1], *xc = x + 3 * at[2],
                 *xd = x + 3 * at[3];
    switch (kind[t]) {
      case 0:
      case 1: { /* distance of (a, b): dr/dx_a = (x_a - x_b) / r */
        double d[3];
        bonded_sub(xa, xb, box, d);
        double r = sqrt(dot3(d, d)), dr = r - p[1];
        energies[kind[t] == 0 ? ORC_E_BOND : ORC_E_UREY_BRADLEY] += 0.5 * p[0] * dr * dr;
        double dudr = p[0] * dr;
        for (int k = 0; k < 3; ++k) {
          double g = dudr * d[k] / r;
          F[3 * at[0] + k] -= g;
          F[3 * at[1] + k] += g;
        }
        break;
      }
      case 2: { /* angle at b between u = a - b and v = c - b */
        double u[3], v[3];
        bonded_sub(xa, xb, box, u);
        bonded_sub(xc, xb, box, v);
        double iu = 1.0 / sqrt(dot3(u, u)), iv = 1.0 / sqrt(dot3(v, v));
        double c = dot3(u, v) * iu * iv;
        double cc = c > 1.0 ? 1.0 : (c < -1.0 ? -1.0 : c);
        double theta = acos(cc), dt = theta - p[1];
        energies[ORC_E_ANGLE] += 0.5 * p[0] * dt * dt;
        double dudt = p[0] * dt;
        double sine = sqrt(1.0 - c * c);
        double slope = -1.0 / sine; /* dtheta/dc */
        for (int k = 0; k < 3; ++k) {
          double eu = u[k] * iu, ev = v[k] * iv;
          double ga = slope * iu * (ev - c * eu);
          double gc = slope * iv * (eu - c * ev);
          F[3 * at[0] + k] -= dudt * ga;
          F[3 * at[2] + k] -= dudt * gc;
          F[3 * at[1] + k] += dudt * (ga + gc);
        }
        break;
      }
      case 3:
      case 4: { /* f = a - b, g = b - c, h = d - c; p = f x g, q = h x g */
        double f[3], g[3], h[3], pv[3], qv[3], qp[3];
        bonded_sub(xa, xb, box, f);
        bonded_sub(xb, xc, box, g);
        bonded_sub(xd, xc, box, h);
        cross3(f, g, pv);
        cross3(h, g, qv);
        cross3(qv, pv, qp);
        double ng = sqrt(dot3(g, g)), pp = dot3(pv, pv), qq = dot3(qv, qv);
        double phi = atan2(dot3(qp, g) / ng, dot3(pv, qv));
        double dudp;
        if (kind[t] == 3) {
          double arg = p[1] * phi - p[2];
          energies[ORC_E_DIHEDRAL] += p[0] * (1.0 + cos(arg));
          dudp = -p[0] * p[1] * sin(arg);
        } else {
          double d = phi - p[1];
          d -= 2.0 * M_PI * floor((d + M_PI) / (2.0 * M_PI));
          energies[ORC_E_IMPROPER] += 0.5 * p[0] * d * d;
          dudp = p[0] * d;
        }
        double fg = dot3(f, g), hg = dot3(h, g);
        for (int k = 0; k < 3; ++k) {
          double ga = -ng * pv[k] / pp;
          double gd = ng * qv[k] / qq;
          double gb = ng * pv[k] / pp + fg * pv[k] / (pp * ng) - hg * qv[k] / (qq * ng);
          double gc = -ng * qv[k] / qq + hg * qv[k] / (qq * ng) - fg * pv[k] / (pp * ng);
          F[3 * at[0] + k] -= dudp * ga;
          F[3 * at[1] + k] -= dudp * gb;
          F[3 * at[2] + k] -= dudp * gc;
          F[3 * at[3] + k] -= dudp * gd;
        }
        break;
      }
      default:
        break;
    }
  }
}

/* ---- constraint clusters: M-SHAKE --------------------------------------
 * ConstraintIR.cpp:595-747 (Krautler, van Gunsteren and Hunenberger,
 * J. Comput. Chem. 22, 501 (2001)): a cluster's bonds k, from its centre c
 * to member h_k, of lengths d_k.  The members move along the old bonds
 * s_l by +lambda_l s_l / m_h and the centre by -sum_l lambda_l s_l / m_c;
 * each sweep solves the linearized equations
 *   sum_l 2 (r_k . s_l) w_kl lambda_l = d_k^2 - r_k . r_k,
 * w_kk = 1/m_c + 1/m_hk, w_kl = 1/m_c (k != l), by elimination; six sweeps,
 * as the reference takes. */
enum { ORC_CLUSTER_MAX = 5 };

static void solve_small(int m, double A[ORC_CLUSTER_MAX][ORC_CLUSTER_MAX], double *b) {
  for (int i = 0; i < m; ++i) {
    for (int r = i + 1; r < m; ++r) {
      double f = A[r][i] / A[i][i];
      for (int c = i; c < m; ++c) A[r][c] -= f * A[i][c];
      b[r] -= f * b[i];
    }
  }
  for (int i = m - 1; i >= 0; --i) {
    double sum = b[i];
    for (int c = i + 1; c < m; ++c) sum -= A[i][c] * b[c];
    b[i] = sum / A[i][i];
  }
}

void orc_shake_positions(const orc_system *s, const double *x_old, double *x_new) {
  for (int64_t u = 0; u < s->nclusters; ++u) {
    const int32_t first = s->cluster_first[u], size = s->cluster_size[u];
    const int32_t c = first + s->cluster_center[u];
    int32_t h[ORC_CLUSTER_MAX];
    int m = 0;
    for (int32_t a = first; a < first + size; ++a)
      if (a != c) h[m++] = a;
    const double ic = 1.0 / s->mass[c];
    double sv[ORC_CLUSTER_MAX][3], ih[ORC_CLUSTER_MAX], d2[ORC_CLUSTER_MAX];
    for (int k = 0; k < m; ++k) {
      ih[k] = 1.0 / s->mass[h[k]];
      d2[k] = s->cluster_length[h[k]] * s->cluster_length[h[k]];
      for (int j = 0; j < 3; ++j) sv[k][j] = x_old[3 * h[k] + j] - x_old[3 * c + j];
    }
    for (int sweep = 0; sweep < 6; ++sweep) {
      double r[ORC_CLUSTER_MAX][3], A[ORC_CLUSTER_MAX][ORC_CLUSTER_MAX], lam[ORC_CLUSTER_MAX];
      for (int k = 0; k < m; ++k) {
        for (int j = 0; j < 3; ++j) r[k][j] = x_new[3 * h[k] + j] - x_new[3 * c + j];
        lam[k] = d2[k] - dot3(r[k], r[k]);
      }
      for (int k = 0; k < m; ++k)
        for (int l = 0; l < m; ++l) {
          double w = (k == l) ? ic + ih[k] : ic;
          A[k][l] = 2.0 * dot3(r[k], sv[l]) * w;
        }
      solve_small(m, A, lam);
      for (int l = 0; l < m; ++l)
        for (int j = 0; j < 3; ++j) {
          x_new[3 * c + j] -= lam[l] * sv[l][j] * ic;
          x_new[3 * h[l] + j] += lam[l] * sv[l][j] * ih[l];
        }
    }
  }
}

/* The velocity constraint: impulses mu_l along the bonds r_l with
   sum_l (r_k . r_l) w_kl mu_l = -r_k . (v_hk - v_c). */
void orc_rattle_velocities(const orc_system *s, const double *x, double *v) {
  for (int64_t u = 0; u < s->nclusters; ++u) {
    const int32_t first = s->cluster_first[u], size = s->cluster_size[u];
    const int32_t c = first + s->cluster_center[u];
    int32_t h[ORC_CLUSTER_MAX];
    int m = 0;
    for (int32_t a = first; a < first + size; ++a)
      if (a != c) h[m++] = a;
    const double ic = 1.0 / s->mass[c];
    double r[ORC_CLUSTER_MAX][3], ih[ORC_CLUSTER_MAX], A[ORC_CLUSTER_MAX][ORC_CLUSTER_MAX],
        mu[ORC_CLUSTER_MAX];
    for (int k = 0; k < m; ++k) {
      ih[k] = 1.0 / s->mass[h[k]];
      double dv[3];
      for (int j = 0; j < 3; ++j) {
        r[k][j] = x[3 * h[k] + j] - x[3 * c + j];
        dv[j] = v[3 * h[k] + j] - v[3 * c + j];
      }
      mu[k] = -dot3(r[k], dv);
    }
    for (int k = 0; k < m; ++k)
      for (int l = 0; l < m; ++l) A[k][l] = dot3(r[k], r[l]) * ((k == l) ? ic + ih[k] : ic);
    solve_small(m, A, mu);
    for (int l = 0; l < m; ++l)
      for (int j = 0; j < 3; ++j) {
        v[3 * c + j] -= mu[l] * r[l][j] * ic;
        v[3 * h[l] + j] += mu[l] * r[l][j] * ih[l];
      }
  }
}

/* TermNonbonded.cpp:778-794: beta by bisection of erfc(beta rc) = tol.
   Pinned by test/Driver/pme.test:8 (3.47045919371208278 nm^-1 at 0.9 nm). */
double orc_ewald_beta(double cutoff, double tolerance) {
  double low = 0.0, high = 1.0;
  while (erfc(high * cutoff) > tolerance) high *= 2.0;
  for (int i = 0; i < 100; ++i) {
    double middle = 0.5 * (low + high);
    if (erfc(middle * cutoff) > tolerance)
      low = middle;
    else
      high = middle;
  }
  return 0.5 * (low + high);
}

/* TermNonbonded.cpp:756-767 (getSmoothSize). */
int64_t orc_smooth_size(int64_t least) {
  for (int64_t size = least;; ++size) {
    if (size % 2) continue;
    int64_t rest = size;
    int64_t f[4] = {2, 3, 5, 7};
    for (int k = 0; k < 4; ++k)
      while (rest % f[k] == 0) rest /= f[k];
    if (rest == 1) return size;
  }
}

/* 05-electrostatics.md:39-40 (Table 5.1): self term and the background of
   a net charge.  Pinned by test/Driver/pme.test:23 (-539.7928271 kcal/mol). */
double orc_self_energy(int64_t n, const double *q, double beta, double coulomb,
                       const double box[3]) {
  double sq = 0.0, net = 0.0;
  for (int64_t i = 0; i < n; ++i) {
    sq += q[i] * q[i];
    net += q[i];
  }
  double volume = box[0] * box[1] * box[2];
  return -coulomb * beta / sqrt(M_PI) * sq -
         coulomb * M_PI * net * net / (2.0 * volume * beta * beta);
}

/* TermNonbonded.cpp:595-636 (getTopologyDispersion): the tail of
   -C6/r^6 beyond rc with the mean C6 over non-excluded ordered pairs.
   Pinned by its closed form for a single type (test_oracle_physics.py). */
double orc_dispersion_energy(int64_t n, const int32_t *type, int32_t ntypes,
                             const double *sigma, const double *epsilon,
                             int64_t nexcl_pairs, const int32_t *excl_pairs,
                             const double box[3], double cutoff) {
  double *numbers = (double *)calloc((size_t)ntypes, sizeof(double));
  for (int64_t i = 0; i < n; ++i) numbers[type[i]] += 1.0;
  double sum = 0.0;
  for (int a = 0; a < ntypes; ++a)
    for (int b = 0; b < ntypes; ++b) {
      double c6 = 4.0 * epsilon[a * ntypes + b] * pow(sigma[a * ntypes + b], 6.0);
      sum += numbers[a] * (numbers[b] - (a == b ? 1.0 : 0.0)) * c6;
    }
  for (int64_t p = 0; p < nexcl_pairs; ++p) {
    int a = type[excl_pairs[2 * p]], b = type[excl_pairs[2 * p + 1]];
    sum -= 2.0 * 4.0 * epsilon[a * ntypes + b] * pow(sigma[a * ntypes + b], 6.0);
  }
  free(numbers);
  double nn = (double)n;
  double pairs = nn * (nn - 1.0) - 2.0 * (double)nexcl_pairs;
  double mean = pairs > 0.0 ? sum / pairs : 0.0;
  double volume = box[0] * box[1] * box[2];
  return -2.0 * M_PI / (3.0 * volume) * nn * nn * mean /
         (cutoff * cutoff * cutoff);
}

/* ======================================================================
 * FFT — conventions of runtime/mdrt_fft.c:110-256 (forward exp(-2 pi i),
 * backward exp(+2 pi i), unnormalised, half spectrum along z).
 * Pinned by a direct DFT and by oracle/_ref (pocketfft through the
 * reference's mdrt_fft.c) in tests/test_oracle_fft.py.
 * ====================================================================== */

static int64_t smallest_factor(int64_t n) {
  if (n % 4 == 0 && n > 4) return 4;
  for (int64_t p = 2; p * p <= n; ++p)
    if (n % p == 0) return p;
  return n;
}

/* Recursive decimation in time: n = p m, X[k] = sum_r w_n^{r k} Y_r[k mod m]
   from the p sub-transforms of stride p.  tw holds exp(sign 2 pi i j / N)
   of the top-level length N; a level of length n reads it at stride N/n. */
static void fft_rec(int64_t n, const double *in, int64_t stride, double *out,
                    const double *tw, int64_t N, double *scratch) {
  if (n == 1) {
    out[0] = in[0];
    out[1] = in[1];
    return;
  }
  int64_t p = smallest_factor(n);
  int64_t m = n / p;
  int64_t step = N / n;
  if (p == n) { /* prime length: direct DFT */
    for (int64_t k = 0; k < n; ++k) {
      double re = 0.0, im = 0.0;
      for (int64_t j = 0; j < n; ++j) {
        const double *w = tw + 2 * (((j * k) % n) * step);
        double xr = in[2 * j * stride], xi = in[2 * j * stride + 1];
        re += xr * w[0] - xi * w[1];
        im += xr * w[1] + xi * w[0];
      }
      out[2 * k] = re;
      out[2 * k + 1] = im;
    }
    return;
  }
  double *sub = scratch;
  double *deeper = scratch + 2 * n;
  for (int64_t r = 0; r < p; ++r)
    fft_rec(m, in + 2 * r * stride, stride * p, sub + 2 * m * r, tw, N, deeper);
  for (int64_t k = 0; k < n; ++k) {
    double re = 0.0, im = 0.0;
    for (int64_t r = 0; r < p; ++r) {
      const double *w = tw + 2 * (((r * k) % n) * step);
      double xr = sub[2 * (m * r + k % m)], xi = sub[2 * (m * r + k % m) + 1];
      re += xr * w[0] - xi * w[1];
      im += xr * w[1] + xi * w[0];
    }
    out[2 * k] = re;
    out[2 * k + 1] = im;
  }
}

static double *twiddles(int64_t n, int sign) {
  double *tw = (double *)malloc(sizeof(double) * 2 * (size_t)n);
  for (int64_t j = 0; j < n; ++j) {
    long double a = sign * 2.0L * 3.14159265358979323846264338327950288L *
                    (long double)j / (long double)n;
    tw[2 * j] = (double)cosl(a);
    tw[2 * j + 1] = (double)sinl(a);
  }
  return tw;
}

void orc_fft_line(int64_t n, double *data, int sign) {
  double *copy = (double *)malloc(sizeof(double) * 2 * (size_t)n);
  double *scratch = (double *)malloc(sizeof(double) * 8 * (size_t)n);
  double *tw = twiddles(n, sign);
  memcpy(copy, data, sizeof(double) * 2 * (size_t)n);
  fft_rec(n, copy, 1, data, tw, n, scratch);
  free(copy);
  free(scratch);
  free(tw);
}

void orc_dft_line(int64_t n, const double *in, double *out, int sign) {
  for (int64_t k = 0; k < n; ++k) {
    long double re = 0.0L, im = 0.0L;
    for (int64_t j = 0; j < n; ++j) {
      long double a = sign * 2.0L * 3.14159265358979323846264338327950288L *
                      (long double)((j * k) % n) / (long double)n;
      long double c = cosl(a), s = sinl(a);
      re += in[2 * j] * c - in[2 * j + 1] * s;
      im += in[2 * j] * s + in[2 * j + 1] * c;
    }
    out[2 * k] = (double)re;
    out[2 * k + 1] = (double)im;
  }
}

/* Complex lines along y (axis 1) or x (axis 0) of a half-spectrum grid. */
static void fft_columns(const int64_t K[3], double *c, int axis, int sign) {
  int64_t h = K[2] / 2 + 1;
  int64_t length = axis == 1 ? K[1] : K[0];
  int64_t along = axis == 1 ? h : K[1] * h;
  int64_t planes = axis == 1 ? K[0] : K[1];
  int64_t across = axis == 1 ? K[1] * h : h;
#pragma omp parallel
  {
    double *line = (double *)malloc(sizeof(double) * 2 * (size_t)length);
    double *scratch = (double *)malloc(sizeof(double) * 8 * (size_t)length);
    double *copy = (double *)malloc(sizeof(double) * 2 * (size_t)length);
    double *tw = twiddles(length, sign);
#pragma omp for collapse(2) schedule(static)
    for (int64_t pl = 0; pl < planes; ++pl)
      for (int64_t z = 0; z < h; ++z) {
        double *base = c + 2 * (pl * across + z);
        for (int64_t j = 0; j < length; ++j) {
          copy[2 * j] = base[2 * j * along];
          copy[2 * j + 1] = base[2 * j * along + 1];
        }
        fft_rec(length, copy, 1, line, tw, length, scratch);
        for (int64_t j = 0; j < length; ++j) {
          base[2 * j * along] = line[2 * j];
          base[2 * j * along + 1] = line[2 * j + 1];
        }
      }
    free(line);
    free(scratch);
    free(copy);
    free(tw);
  }
}

void orc_fft3d_r2c(const int64_t K[3], const double *real, double *cplx) {
  int64_t h = K[2] / 2 + 1, k3 = K[2];
#pragma omp parallel
  {
    double *line = (double *)malloc(sizeof(double) * 2 * (size_t)k3);
    double *out = (double *)malloc(sizeof(double) * 2 * (size_t)k3);
    double *scratch = (double *)malloc(sizeof(double) * 8 * (size_t)k3);
    double *tw = twiddles(k3, -1);
#pragma omp for schedule(static)
    for (int64_t row = 0; row < K[0] * K[1]; ++row) {
      for (int64_t z = 0; z < k3; ++z) {
        line[2 * z] = real[row * k3 + z];
        line[2 * z + 1] = 0.0;
      }
      fft_rec(k3, line, 1, out, tw, k3, scratch);
      for (int64_t z = 0; z < h; ++z) {
        cplx[2 * (row * h + z)] = out[2 * z];
        cplx[2 * (row * h + z) + 1] = out[2 * z + 1];
      }
    }
    free(line);
    free(out);
    free(scratch);
    free(tw);
  }
  fft_columns(K, cplx, 1, -1);
  fft_columns(K, cplx, 0, -1);
}

void orc_fft3d_c2r(const int64_t K[3], const double *cplx_in, double *real) {
  int64_t h = K[2] / 2 + 1, k3 = K[2];
  size_t total = (size_t)(K[0] * K[1] * h) * 2;
  double *c = (double *)malloc(sizeof(double) * total);
  memcpy(c, cplx_in, sizeof(double) * total);
  fft_columns(K, c, 1, +1);
  fft_columns(K, c, 0, +1);
#pragma omp parallel
  {
    double *line = (double *)malloc(sizeof(double) * 2 * (size_t)k3);
    double *out = (double *)malloc(sizeof(double) * 2 * (size_t)k3);
    double *scratch = (double *)malloc(sizeof(double) * 8 * (size_t)k3);
    double *tw = twiddles(k3, +1);
#pragma omp for schedule(static)
    for (int64_t row = 0; row < K[0] * K[1]; ++row) {
      const double *in = c + 2 * row * h;
      /* Hermitian completion; the imaginary parts of the zero and (even
         k3) Nyquist terms are ignored, as pocketfft's rfft_backward does
         with the r0, r1 i1, ... layout of mdrt_fft.c:149-157. */
      for (int64_t z = 0; z < h; ++z) {
        line[2 * z] = in[2 * z];
        line[2 * z + 1] = in[2 * z + 1];
      }
      line[1] = 0.0;
      if (k3 % 2 == 0) line[2 * (k3 / 2) + 1] = 0.0;
      for (int64_t z = h; z < k3; ++z) {
        line[2 * z] = in[2 * (k3 - z)];
        line[2 * z + 1] = -in[2 * (k3 - z) + 1];
      }
      fft_rec(k3, line, 1, out, tw, k3, scratch);
      for (int64_t z = 0; z < k3; ++z) real[row * k3 + z] = out[2 * z];
    }
    free(line);
    free(out);
    free(scratch);
    free(tw);
  }
  free(c);
}

/* ======================================================================
 * PME — 05-electrostatics.md:95-279, PMECode.h:99-186 (splines),
 * Influence.cpp:168-191 (moduli), HostPMESpread/Convolve/Gather.cpp.
 * Pinned by the Ewald sum (orc_ewald_reciprocal) and by the reciprocal
 * energies of test/Driver/pme.test:17-22 (4.6972413 at 64^3 order 8 to six
 * digits; 4.693822 at 35^3 order 4), by finite differences of the energy,
 * and by the order-4 closed forms of 05-electrostatics.md:290.
 * ====================================================================== */

/* A[k] = M_n(w + k), k = 0..n-1; dA[k] = M_n'(w + k), by the recursion
   M_n(x) = (x M_{n-1}(x) + (n - x) M_{n-1}(x - 1)) / (n - 1), M_2 = 1-|x-1|,
   M_n'(x) = M_{n-1}(x) - M_{n-1}(x - 1).  The weight of point first + t is
   M_n(w + n - 1 - t) = A[n - 1 - t]; orc_bspline returns it reordered:
   M[t] and dM[t] for t = 0..n-1. */
void orc_bspline(double w, int order, double *M, double *dM) {
  double A[17] = {0}, B[17] = {0};
  A[0] = w;       /* M_2(w)     */
  A[1] = 1.0 - w; /* M_2(w + 1) */
  for (int nn = 3; nn <= order; ++nn) {
    for (int k = 0; k <= nn; ++k) B[k] = A[k]; /* B = A_{nn-1}, B[nn-1] = 0 */
    for (int k = 0; k < nn; ++k) {
      double prev = k > 0 ? B[k - 1] : 0.0;
      A[k] = ((w + k) * B[k] + (nn - w - k) * prev) / (nn - 1);
    }
  }
  /* derivative from A_{order-1} */
  if (order == 2) {
    B[0] = 1.0;
    B[1] = 0.0;
    B[2] = 0.0;
    for (int k = 0; k < 2; ++k) dM[1 - k] = k == 0 ? 1.0 : -1.0;
  } else {
    for (int k = 0; k < order; ++k)
      dM[order - 1 - k] = B[k] - (k > 0 ? B[k - 1] : 0.0);
  }
  for (int k = 0; k < order; ++k) M[order - 1 - k] = A[k];
}

/* Influence.cpp:168-191: |b(k)|^2 = 1 / |sum_{j=0}^{n-2} M_n(j+1) e^{2 pi i k j/K}|^2. */
void orc_spline_moduli(int64_t count, int order, double *moduli) {
  double m[17] = {0}, next[17];
  m[1] = 1.0;
  for (int nn = 3; nn <= order; ++nn) {
    memset(next, 0, sizeof(next));
    for (int xx = 1; xx < nn; ++xx)
      next[xx] = (xx * m[xx] + (nn - xx) * m[xx - 1]) / (nn - 1);
    memcpy(m, next, sizeof(m));
  }
  for (int64_t k = 0; k < count; ++k) {
    double re = 0.0, im = 0.0;
    for (int j = 0; j <= order - 2; ++j) {
      double a = 2.0 * M_PI * (double)k * j / (double)count;
      re += m[j + 1] * cos(a);
      im += m[j + 1] * sin(a);
    }
    moduli[k] = 1.0 / (re * re + im * im);
  }
}

/* PMEWeights.cpp:8 / 05-electrostatics.md:289: u = K frac(x/L), first point
   (floor(u) - n + 1) mod K, fraction w. */
static void pme_place(double x, double L, int64_t K, int order, int64_t *first,
                      double *w) {
  double s = x * (1.0 / L);
  double frac = s - floor(s);
  double u = (double)K * frac;
  double fl = floor(u);
  *w = u - fl;
  int64_t f = (int64_t)fl - order + 1;
  f %= K;
  if (f < 0) f += K;
  *first = f;
}

void orc_pme_spread(int64_t n, const double *x, const double *q,
                    const double box[3], const int64_t K[3], int order,
                    double *grid) {
  memset(grid, 0, sizeof(double) * (size_t)(K[0] * K[1] * K[2]));
  for (int64_t i = 0; i < n; ++i) {
    if (q[i] == 0.0) continue;
    int64_t f[3];
    double M[3][16], dM[3][16];
    for (int a = 0; a < 3; ++a) {
      double w;
      pme_place(x[3 * i + a], box[a], K[a], order, &f[a], &w);
      orc_bspline(w, order, M[a], dM[a]);
    }
    for (int t0 = 0; t0 < order; ++t0) {
      int64_t i0 = (f[0] + t0) % K[0];
      for (int t1 = 0; t1 < order; ++t1) {
        int64_t i1 = (f[1] + t1) % K[1];
        double w01 = q[i] * M[0][t0] * M[1][t1];
        for (int t2 = 0; t2 < order; ++t2) {
          int64_t i2 = (f[2] + t2) % K[2];
          grid[(i0 * K[1] + i1) * K[2] + i2] += w01 * M[2][t2];
        }
      }
    }
  }
}

/* 05-electrostatics.md:262-279 (separable factors): on the half spectrum,
   BC = t1 t2 t3 / m^2, t_a = exp(-pi^2 m_a^2/beta^2) |b_a|^2, times
   f/(pi V); BC(0) = 0. */
void orc_pme_influence(const double box[3], const int64_t K[3], int order,
                       double beta, double coulomb, double *bc) {
  int64_t h = K[2] / 2 + 1;
  double *mod[3], *m[3], *t[3];
  double volume = box[0] * box[1] * box[2];
  for (int a = 0; a < 3; ++a) {
    mod[a] = (double *)malloc(sizeof(double) * (size_t)K[a]);
    m[a] = (double *)malloc(sizeof(double) * (size_t)K[a]);
    t[a] = (double *)malloc(sizeof(double) * (size_t)K[a]);
    orc_spline_moduli(K[a], order, mod[a]);
    for (int64_t k = 0; k < K[a]; ++k) {
      int64_t ks = k <= K[a] / 2 ? k : k - K[a];
      m[a][k] = (double)ks / box[a];
      t[a][k] = exp(-M_PI * M_PI * m[a][k] * m[a][k] / (beta * beta)) * mod[a][k];
    }
  }
  double pre = coulomb / (M_PI * volume);
  for (int64_t i0 = 0; i0 < K[0]; ++i0)
    for (int64_t i1 = 0; i1 < K[1]; ++i1)
      for (int64_t i2 = 0; i2 < h; ++i2) {
        double m2 = m[0][i0] * m[0][i0] + m[1][i1] * m[1][i1] + m[2][i2] * m[2][i2];
        double v = 0.0;
        if (i0 || i1 || i2) v = pre * t[0][i0] * t[1][i1] * t[2][i2] / m2;
        bc[(i0 * K[1] + i1) * h + i2] = v;
      }
  for (int a = 0; a < 3; ++a) {
    free(mod[a]);
    free(m[a]);
    free(t[a]);
  }
}

/* 05-electrostatics.md:244-251: F_ja = -q_j (K_a/L_a) sum phi M' prod M. */
void orc_pme_gather(int64_t n, const double *x, const double *q,
                    const double box[3], const int64_t K[3], int order,
                    const double *phi, double *F) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    if (q[i] == 0.0) continue;
    int64_t f[3];
    double M[3][16], dM[3][16];
    for (int a = 0; a < 3; ++a) {
      double w;
      pme_place(x[3 * i + a], box[a], K[a], order, &f[a], &w);
      orc_bspline(w, order, M[a], dM[a]);
    }
    double g[3] = {0, 0, 0};
    for (int t0 = 0; t0 < order; ++t0) {
      int64_t i0 = (f[0] + t0) % K[0];
      for (int t1 = 0; t1 < order; ++t1) {
        int64_t i1 = (f[1] + t1) % K[1];
        for (int t2 = 0; t2 < order; ++t2) {
          int64_t i2 = (f[2] + t2) % K[2];
          double p = phi[(i0 * K[1] + i1) * K[2] + i2];
          g[0] += p * dM[0][t0] * M[1][t1] * M[2][t2];
          g[1] += p * M[0][t0] * dM[1][t1] * M[2][t2];
          g[2] += p * M[0][t0] * M[1][t1] * dM[2][t2];
        }
      }
    }
    for (int a = 0; a < 3; ++a)
      F[3 * i + a] += -q[i] * ((double)K[a] / box[a]) * g[a];
  }
}

double orc_pme(int64_t n, const double *x, const double *q,
               const double box[3], const int64_t K[3], int order,
               double beta, double coulomb, double *F, double *virial) {
  int64_t h = K[2] / 2 + 1;
  size_t np = (size_t)(K[0] * K[1] * K[2]), nh = (size_t)(K[0] * K[1] * h);
  double *grid = (double *)malloc(sizeof(double) * np);
  double *cplx = (double *)malloc(sizeof(double) * 2 * nh);
  double *bc = (double *)malloc(sizeof(double) * nh);
  orc_pme_spread(n, x, q, box, K, order, grid);
  orc_fft3d_r2c(K, grid, cplx);
  orc_pme_influence(box, K, order, beta, coulomb, bc);
  double energy = 0.0, w[6] = {0};
  for (int64_t i0 = 0; i0 < K[0]; ++i0)
    for (int64_t i1 = 0; i1 < K[1]; ++i1)
      for (int64_t i2 = 0; i2 < h; ++i2) {
        size_t p = (size_t)((i0 * K[1] + i1) * h + i2);
        double hw = (i2 == 0 || (K[2] % 2 == 0 && i2 == K[2] / 2)) ? 0.5 : 1.0;
        double mag = cplx[2 * p] * cplx[2 * p] + cplx[2 * p + 1] * cplx[2 * p + 1];
        double em = hw * bc[p] * mag;
        energy += em;
        if (virial && bc[p] != 0.0) {
          int64_t k0 = i0 <= K[0] / 2 ? i0 : i0 - K[0];
          int64_t k1 = i1 <= K[1] / 2 ? i1 : i1 - K[1];
          double mm[3] = {k0 / box[0], k1 / box[1], i2 / box[2]};
          double m2 = mm[0] * mm[0] + mm[1] * mm[1] + mm[2] * mm[2];
          double fac = 2.0 * (1.0 / m2 + M_PI * M_PI / (beta * beta));
          w[0] += em * (1.0 - fac * mm[0] * mm[0]);
          w[1] += em * (1.0 - fac * mm[1] * mm[1]);
          w[2] += em * (1.0 - fac * mm[2] * mm[2]);
          w[3] += em * (-fac * mm[0] * mm[1]);
          w[4] += em * (-fac * mm[0] * mm[2]);
          w[5] += em * (-fac * mm[1] * mm[2]);
        }
        cplx[2 * p] *= bc[p];
        cplx[2 * p + 1] *= bc[p];
      }
  orc_fft3d_c2r(K, cplx, grid);
  orc_pme_gather(n, x, q, box, K, order, grid, F);
  if (virial)
    for (int k = 0; k < 6; ++k) virial[k] += w[k];
  free(grid);
  free(cplx);
  free(bc);
  return energy;
}

/* ewald_reference.py:99-117 method, with forces. */
double orc_ewald_reciprocal(int64_t n, const double *x, const double *q,
                            const double box[3], double beta, double coulomb,
                            double *F) {
  double volume = box[0] * box[1] * box[2];
  double limit = beta * sqrt(18.0 * log(10.0)) / M_PI;
  int64_t kmax[3];
  for (int k = 0; k < 3; ++k) kmax[k] = (int64_t)(limit * box[k]) + 1;
  double energy = 0.0;
  double *cs = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
  double *sn = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
  for (int64_t a = 0; a <= kmax[0]; ++a)
    for (int64_t b = -kmax[1]; b <= kmax[1]; ++b)
      for (int64_t c = -kmax[2]; c <= kmax[2]; ++c) {
        if (a == 0 && (b < 0 || (b == 0 && c <= 0))) continue;
        double m[3] = {a / box[0], b / box[1], c / box[2]};
        double m2 = m[0] * m[0] + m[1] * m[1] + m[2] * m[2];
        if (m2 > limit * limit) continue;
        double re = 0.0, im = 0.0;
        for (int64_t i = 0; i < n; ++i) {
          double ph = 2.0 * M_PI * (m[0] * x[3 * i] + m[1] * x[3 * i + 1] + m[2] * x[3 * i + 2]);
          cs[i] = cos(ph);
          sn[i] = sin(ph);
          re += q[i] * cs[i];
          im += q[i] * sn[i];
        }
        double g = exp(-M_PI * M_PI * m2 / (beta * beta)) / m2;
        double pre = 2.0 * coulomb / (2.0 * M_PI * volume) * g;
        energy += pre * (re * re + im * im);
        if (F)
          for (int64_t i = 0; i < n; ++i) {
            double s = pre * 2.0 * q[i] * 2.0 * M_PI * (im * cs[i] - re * sn[i]);
            for (int k = 0; k < 3; ++k) F[3 * i + k] -= s * m[k];
          }
      }
  free(cs);
  free(sn);
  return energy;
}

/* ======================================================================
 * Constraints — ConstraintIR.cpp:473-593 (SETTLE positions, the analytic
 * method of Miyamoto and Kollman, J. Comput. Chem. 13, 952 (1992)) and
 * :392-471 (velocities, a 3x3 solve for impulses along the bonds).
 * Pinned against a Newton-converged SHAKE of the same equations and the
 * invariants (bond lengths, momentum, zero bond-relative velocity) in
 * tests/test_oracle_constraints.py.
 * ====================================================================== */

static void unit3(double *a) {
  double s = 1.0 / sqrt(dot3(a, a));
  a[0] *= s;
  a[1] *= s;
  a[2] *= s;
}

void orc_settle_positions(int64_t nw, const int32_t *first, double m_o,
                          double m_h, double d_oh, double d_hh,
                          const double *x_old, double *x_new) {
  double total = m_o + 2.0 * m_h;
  double rc = 0.5 * d_hh;
  double height = sqrt(d_oh * d_oh - rc * rc);
  double ra = 2.0 * m_h * height / total;
  double rb = height - ra;
#pragma omp parallel for schedule(static)
  for (int64_t w = 0; w < nw; ++w) {
    int64_t o = first[w];
    const double *A0 = x_old + 3 * o, *B0 = A0 + 3, *C0 = A0 + 6;
    double *A1 = x_new + 3 * o, *B1 = A1 + 3, *C1 = A1 + 6;
    double b0[3], c0[3], b1r[3], c1r[3];
    for (int k = 0; k < 3; ++k) {
      b0[k] = B0[k] - A0[k]; /* old bonds, relative to the old oxygen */
      c0[k] = C0[k] - A0[k];
      b1r[k] = B1[k] - A1[k]; /* new, relative to the new oxygen */
      c1r[k] = C1[k] - A1[k];
    }
    double com[3], a1[3], b1[3], c1[3];
    for (int k = 0; k < 3; ++k) {
      com[k] = m_h / total * (b1r[k] + c1r[k]);
      a1[k] = -com[k];
      b1[k] = b1r[k] - com[k];
      c1[k] = c1r[k] - com[k];
    }
    double ez[3], ex[3], ey[3];
    cross3(b0, c0, ez);
    unit3(ez);
    cross3(a1, ez, ex);
    unit3(ex);
    cross3(ez, ex, ey);
    double xb0 = dot3(b0, ex), yb0 = dot3(b0, ey);
    double xc0 = dot3(c0, ex), yc0 = dot3(c0, ey);
    double za1 = dot3(a1, ez);
    double xb1 = dot3(b1, ex), yb1 = dot3(b1, ey), zb1 = dot3(b1, ez);
    double xc1 = dot3(c1, ex), yc1 = dot3(c1, ey), zc1 = dot3(c1, ez);
    double sinphi = za1 / ra;
    double cosphi = sqrt(1.0 - sinphi * sinphi);
    double sinpsi = (zb1 - zc1) / (2.0 * rc * cosphi);
    double cospsi = sqrt(1.0 - sinpsi * sinpsi);
    double ya2 = ra * cosphi;
    double xb2 = -rc * cospsi;
    double t1 = -rb * cosphi, t2 = rc * sinpsi * sinphi;
    double yb2 = t1 - t2, yc2 = t1 + t2;
    double alpha = xb2 * (xb0 - xc0) + yb0 * yb2 + yc0 * yc2;
    double beta = xb2 * (yc0 - yb0) + xb0 * yb2 + xc0 * yc2;
    double gamma = (xb0 * yb1 - xb1 * yb0) + (xc0 * yc1 - xc1 * yc0);
    double both = alpha * alpha + beta * beta;
    double sintheta = (alpha * gamma - beta * sqrt(both - gamma * gamma)) / both;
    double costheta = sqrt(1.0 - sintheta * sintheta);
    double xa3 = -ya2 * sintheta, ya3 = ya2 * costheta;
    double xb3 = xb2 * costheta - yb2 * sintheta;
    double yb3 = xb2 * sintheta + yb2 * costheta;
    double xc3 = -xb2 * costheta - yc2 * sintheta;
    double yc3 = -xb2 * sintheta + yc2 * costheta;
    double a3[3], b3[3], c3[3];
    for (int k = 0; k < 3; ++k) {
      a3[k] = xa3 * ex[k] + ya3 * ey[k] + za1 * ez[k] + com[k];
      b3[k] = xb3 * ex[k] + yb3 * ey[k] + zb1 * ez[k] + com[k];
      c3[k] = xc3 * ex[k] + yc3 * ey[k] + zc1 * ez[k] + com[k];
    }
    /* a3 etc. are relative to the new (unconstrained) oxygen A1. */
    double base[3] = {A1[0], A1[1], A1[2]};
    for (int k = 0; k < 3; ++k) {
      A1[k] = base[k] + a3[k];
      B1[k] = base[k] + b3[k];
      C1[k] = base[k] + c3[k];
    }
  }
}

static double det3(double m00, double m01, double m02, double m10, double m11,
                   double m12, double m20, double m21, double m22) {
  return m00 * (m11 * m22 - m12 * m21) - m01 * (m10 * m22 - m12 * m20) +
         m02 * (m10 * m21 - m11 * m20);
}

void orc_settle_velocities(int64_t nw, const int32_t *first, double m_o,
                           double m_h, const double *x, double *v) {
  double io = 1.0 / m_o, ih = 1.0 / m_h;
#pragma omp parallel for schedule(static)
  for (int64_t w = 0; w < nw; ++w) {
    int64_t o = first[w];
    const double *xa = x + 3 * o, *xb = xa + 3, *xc = xa + 6;
    double *va = v + 3 * o, *vb = va + 3, *vc = va + 6;
    double e0[3], e1[3], e2[3];
    for (int k = 0; k < 3; ++k) {
      e0[k] = xb[k] - xa[k];
      e1[k] = xc[k] - xa[k];
      e2[k] = xc[k] - xb[k];
    }
    unit3(e0);
    unit3(e1);
    unit3(e2);
    double dvb[3], dvc[3], dvcb[3];
    for (int k = 0; k < 3; ++k) {
      dvb[k] = vb[k] - va[k];
      dvc[k] = vc[k] - va[k];
      dvcb[k] = vc[k] - vb[k];
    }
    /* impulses t along e0 (O-H1), e1 (O-H2), e2 (H1-H2) that zero the
       relative velocities along the bonds. */
    double r0 = -dot3(e0, dvb), r1 = -dot3(e1, dvc), r2 = -dot3(e2, dvcb);
    double a00 = io + ih, a11 = io + ih, a22 = 2.0 * ih;
    double a01 = dot3(e0, e1) * io;
    double a02 = -dot3(e0, e2) * ih;
    double a12 = dot3(e1, e2) * ih;
    double det = det3(a00, a01, a02, a01, a11, a12, a02, a12, a22);
    double t0 = det3(r0, a01, a02, r1, a11, a12, r2, a12, a22) / det;
    double t1 = det3(a00, r0, a02, a01, r1, a12, a02, r2, a22) / det;
    double t2 = det3(a00, a01, r0, a01, a11, r1, a02, a12, r2) / det;
    for (int k = 0; k < 3; ++k) {
      double i0 = t0 * e0[k], i1 = t1 * e1[k], i2 = t2 * e2[k];
      va[k] += -io * (i0 + i1);
      vb[k] += ih * (i0 - i2);
      vc[k] += ih * (i1 + i2);
    }
  }
}

/* ======================================================================
 * The whole force and one NVE step — DynamicsIR.cpp:102-112, :275 and
 * 08-gpu.md:112-121: kick, drift, constrain, v += dx_c/dt, forces, kick,
 * constrain velocities.
 * ====================================================================== */

int orc_total(const orc_system *s, const double *x, double *F,
              double *energies, int32_t mixed) {
  int64_t n = s->n;
  memset(F, 0, sizeof(double) * 3 * (size_t)n);
  for (int k = 0; k < ORC_NTERMS; ++k) energies[k] = 0.0;
  orc_nonbonded(n, x, s->q, s->type, s->ntypes, s->sigma, s->epsilon, s->box,
                s->cutoff, s->switch_from, s->lj_modifier, s->coulomb_shift,
                s->beta, s->coulomb, s->excl_start, s->excl_atoms, mixed, F,
                energies, NULL);
  if (s->beta > 0.0) {
    if (s->nexcl_pairs)
      orc_excluded(s->nexcl_pairs, s->excl_pairs, x, s->q, s->box, s->beta,
                   s->coulomb, F, &energies[ORC_E_COUL_EXCL]);
    energies[ORC_E_COUL_RECIP] +=
        orc_pme(n, x, s->q, s->box, s->K, s->order, s->beta, s->coulomb, F, NULL);
    energies[ORC_E_COUL_SELF] += orc_self_energy(n, s->q, s->beta, s->coulomb, s->box);
  }
  if (s->dispersion)
    energies[ORC_E_DISP_CORR] += orc_dispersion_energy(
        n, s->type, s->ntypes, s->sigma, s->epsilon, s->nexcl_pairs,
        s->excl_pairs, s->box, s->cutoff);
  if (s->npairs14)
    orc_pairs14(s->npairs14, s->pairs14, s->p14_sigma, s->p14_eps, s->p14_qq, x, s->box,
                s->cutoff, s->switch_from, s->lj_modifier, F, &energies[ORC_E_LJ14],
                &energies[ORC_E_COUL14]);
  if (s->ntuples)
    orc_bonded(s->ntuples, s->tuple_atoms, s->tuple_kind, s->tuple_param, x, s->box, F,
               energies);
  return 0;
}

int orc_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

int orc_nve(const orc_system *s, double *x, double *v, double dt,
            int64_t steps, double *trace, int32_t mixed) {
  int64_t n = s->n;
  double *F = (double *)malloc(sizeof(double) * 3 * (size_t)n);
  double *xo = (double *)malloc(sizeof(double) * 3 * (size_t)n);
  double e[ORC_NTERMS];
  orc_total(s, x, F, e, mixed);
  for (int64_t step = 0; step < steps; ++step) {
    for (int64_t i = 0; i < n; ++i)
      for (int k = 0; k < 3; ++k) v[3 * i + k] += 0.5 * dt * F[3 * i + k] / s->mass[i];
    memcpy(xo, x, sizeof(double) * 3 * (size_t)n);
    for (int64_t i = 0; i < 3 * n; ++i) x[i] += dt * v[i];
    if (s->nwaters || s->nclusters) {
      double *xu = (double *)malloc(sizeof(double) * 3 * (size_t)n);
      memcpy(xu, x, sizeof(double) * 3 * (size_t)n);
      if (s->nwaters)
        orc_settle_positions(s->nwaters, s->water_first, s->m_o, s->m_h, s->d_oh,
                             s->d_hh, xo, x);
      if (s->nclusters) orc_shake_positions(s, xo, x);
      for (int64_t i = 0; i < 3 * n; ++i) v[i] += (x[i] - xu[i]) / dt;
      free(xu);
    }
    orc_total(s, x, F, e, mixed);
    for (int64_t i = 0; i < n; ++i)
      for (int k = 0; k < 3; ++k) v[3 * i + k] += 0.5 * dt * F[3 * i + k] / s->mass[i];
    if (s->nwaters)
      orc_settle_velocities(s->nwaters, s->water_first, s->m_o, s->m_h, x, v);
    if (s->nclusters) orc_rattle_velocities(s, x, v);
    if (trace) {
      double pot = 0.0, kin = 0.0;
      for (int k = 0; k < ORC_NTERMS; ++k) pot += e[k];
      for (int64_t i = 0; i < n; ++i)
        kin += 0.5 * s->mass[i] * dot3(v + 3 * i, v + 3 * i);
      trace[2 * step] = pot;
      trace[2 * step + 1] = kin;
    }
  }
  free(F);
  free(xo);
  return 0;
}

/* ======================================================================
 * The CPU baseline the benchmark times: the reference's OpenMP CPU path
 * for a step of velocity Verlet with SETTLE waters, restated for speed
 * rather than as a checker.  The neighbor matrix (NeighborsMatrix.mlir
 * :193-460) is built at rc + buffer and reused until an atom moved more
 * than half the buffer since the build (the validity test,
 * ExposeValidity.cpp:23, 04-neighbors.md:54-133); pair rows are owner-
 * computes (each pair in both rows, MDExecKernels.cpp:272) with the f32
 * kernels of the mixed mode; the excluded-pair correction in f64; PME with
 * the charge spread as 64-bit fixed point at 2^40 (docs/pme.md:310-331,
 * the deterministic mode), f64 line FFTs, the influence function computed
 * once for the box, and the gather; SETTLE and the integrator in f64.
 * Forces only: energies are not evaluated in the timed steps (the
 * reference evaluates them on output steps only).
 * ====================================================================== */

struct orc_md {
  const orc_system *s;
  double skin;
  int64_t width;
  int32_t *counts, *index;
  double *xref, *F, *xo, *xu;
  int64_t *words;        /* spread grid, fixed point */
  float *xf, *qf;        /* positions wrapped into the box and scaled charges, f32 */
  double *grid, *cplx, *bc;
  int64_t builds;
};

static void md_build(orc_md *c, const double *x) {
  const orc_system *s = c->s;
  int64_t n = s->n;
  double reach = s->cutoff + c->skin;
  double cw = orc_cell_width(n, s->box, reach, 0.0);
  for (;;) {
    int64_t largest =
        orc_neighbors_matrix(n, x, s->box, reach, cw, c->counts, c->index, c->width);
    if (largest <= c->width) break;
    c->width = (int64_t)(largest * 1.25) + 1;
    free(c->index);
    c->index = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n * c->width + 1));
  }
  if (s->excl_start)
    orc_exclusion_filter(n, c->counts, c->index, c->width, s->excl_start, s->excl_atoms);
  memcpy(c->xref, x, sizeof(double) * 3 * (size_t)n);
  ++c->builds;
}

orc_md *orc_md_create(const orc_system *s, double skin) {
  orc_md *c = (orc_md *)calloc(1, sizeof(orc_md));
  int64_t n = s->n;
  c->s = s;
  c->skin = skin;
  c->width = 512;
  c->counts = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
  c->index = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n * c->width + 1));
  c->xref = (double *)malloc(sizeof(double) * 3 * (size_t)n);
  c->F = (double *)malloc(sizeof(double) * 3 * (size_t)n);
  c->xo = (double *)malloc(sizeof(double) * 3 * (size_t)n);
  c->xu = (double *)malloc(sizeof(double) * 3 * (size_t)n);
  c->xf = (float *)malloc(sizeof(float) * 3 * (size_t)n);
  c->qf = (float *)malloc(sizeof(float) * (size_t)(n > 0 ? n : 1));
  if (s->beta > 0.0) {
    size_t np = (size_t)(s->K[0] * s->K[1] * s->K[2]);
    size_t nh = (size_t)(s->K[0] * s->K[1] * (s->K[2] / 2 + 1));
    c->words = (int64_t *)malloc(sizeof(int64_t) * np);
    c->grid = (double *)malloc(sizeof(double) * np);
    c->cplx = (double *)malloc(sizeof(double) * 2 * nh);
    c->bc = (double *)malloc(sizeof(double) * nh);
    orc_pme_influence(s->box, s->K, s->order, s->beta, s->coulomb, c->bc);
  }
  return c;
}

void orc_md_destroy(orc_md *c) {
  if (!c) return;
  free(c->counts);
  free(c->index);
  free(c->xref);
  free(c->F);
  free(c->xo);
  free(c->xu);
  free(c->xf);
  free(c->qf);
  free(c->words);
  free(c->grid);
  free(c->cplx);
  free(c->bc);
  free(c);
}

int64_t orc_md_builds(const orc_md *c) { return c->builds; }

/* exp(y) for y <= 0 in f32, vectorizable: y log2(e) = n + f with n an
   integer and f in [-1/2, 1/2], 2^f by its Taylor polynomial of degree 6
   (relative error below 3e-8 on that interval), 2^n by the exponent bits.
   Arguments below -87 give 0. */
static inline float md_expf(float y) {
  float t = y * 1.44269504088896341f;
  t = t < -126.0f ? -126.0f : t;
  float n = rintf(t), f = t - n;
  float p = 1.5403530393381606e-4f;
  p = p * f + 1.3333558146428443e-3f;
  p = p * f + 9.6181291076284772e-3f;
  p = p * f + 5.5504108664821580e-2f;
  p = p * f + 2.4022650695910071e-1f;
  p = p * f + 6.9314718055994531e-1f;
  p = p * f + 1.0f;
  int32_t bits = ((int32_t)n + 127) << 23;
  float scale;
  memcpy(&scale, &bits, 4);
  return p * scale;
}

/* One pair row (owner computes: every neighbor of i), f32, LJ without a
   modifier on the force (NONE or POTENTIAL_SHIFT) and erfc Ewald: the
   branch-free form a SIMD unit runs (the reference's CPU kernels are
   vectorised by its compiler).  erfc(x) = t e^{-x^2} P(t) with
   t = 1/(1 + 0.475 x) and P of degree 6 (the pair loop's approximation,
   below 1e-7 relative in exact arithmetic). */
__attribute__((target_clones("avx2", "default")))
static void md_row(int64_t i, const int32_t *restrict idx, int32_t cnt,
                   const float *restrict X, const float *restrict Q, const int32_t *restrict T,
                   const float *restrict c6t, const float *restrict c12t, int32_t ntypes,
                   const float L[3], const float iL[3], float rc2, float beta, float mb2,
                   float tbs, double *restrict Fi) {
  const float xi = X[3 * i], yi = X[3 * i + 1], zi = X[3 * i + 2], qi = Q[i];
  const float *c6r = c6t + T[i] * ntypes, *c12r = c12t + T[i] * ntypes;
  float fx = 0.0f, fy = 0.0f, fz = 0.0f;
#pragma omp simd reduction(+ : fx, fy, fz)
  for (int32_t k = 0; k < cnt; ++k) {
    const int32_t j = idx[k];
    float dx = xi - X[3 * j], dy = yi - X[3 * j + 1], dz = zi - X[3 * j + 2];
    dx -= L[0] * rintf(dx * iL[0]);
    dy -= L[1] * rintf(dy * iL[1]);
    dz -= L[2] * rintf(dz * iL[2]);
    float r2 = dx * dx + dy * dy + dz * dz;
    const int on = r2 < rc2;
    r2 = on ? r2 : 1.0f;
    const float rinv = 1.0f / sqrtf(r2), rinv2 = rinv * rinv;
    const float br = beta * r2 * rinv, ex = md_expf(mb2 * r2);
    const float t = 1.0f / (1.0f + 0.475f * br);
    float p = 0.09289151430130005f;
    p = p * t - 0.4289827346801758f;
    p = p * t + 0.6025476455688477f;
    p = p * t - 0.1264616847038269f;
    p = p * t + 0.3417808711528778f;
    p = p * t + 0.24873463809490204f;
    p = p * t + 0.26948970556259155f;
    const float erfcv = t * ex * p;
    const float fc = qi * Q[j] * (erfcv * rinv + tbs * ex) * rinv2;
    const float r6 = rinv2 * rinv2 * rinv2;
    const int32_t tj = T[j];
    const float flj = (c12r[tj] * r6 - c6r[tj]) * r6 * rinv2;
    const float fs = on ? fc + flj : 0.0f;
    fx += fs * dx;
    fy += fs * dy;
    fz += fs * dz;
  }
  Fi[0] += fx;
  Fi[1] += fy;
  Fi[2] += fz;
}

static void md_pairs(orc_md *c, const double *x, double *F) {
  const orc_system *s = c->s;
  int64_t n = s->n, width = c->width;
  if (s->lj_modifier == 2 || s->lj_modifier == 3) { /* switches: the oracle's rows */
    double e[ORC_NTERMS] = {0};
    orc_nonbonded(n, x, s->q, s->type, s->ntypes, s->sigma, s->epsilon, s->box, s->cutoff,
                  s->switch_from, s->lj_modifier, s->coulomb_shift, s->beta, s->coulomb,
                  s->excl_start, s->excl_atoms, 1, F, e, NULL);
    return;
  }
  int32_t nt = s->ntypes;
  float L[3], iL[3];
  for (int k = 0; k < 3; ++k) {
    L[k] = (float)s->box[k];
    iL[k] = 1.0f / L[k];
  }
  float *c6 = (float *)malloc(sizeof(float) * (size_t)(nt * nt));
  float *c12 = (float *)malloc(sizeof(float) * (size_t)(nt * nt));
  for (int32_t a = 0; a < nt * nt; ++a) {
    double s6 = pow(s->sigma[a], 6.0);
    c6[a] = (float)(6.0 * 4.0 * s->epsilon[a] * s6);
    c12[a] = (float)(12.0 * 4.0 * s->epsilon[a] * s6 * s6);
  }
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    for (int k = 0; k < 3; ++k) {
      double w = x[3 * i + k] - s->box[k] * floor(x[3 * i + k] / s->box[k]);
      c->xf[3 * i + k] = (float)w;
    }
    c->qf[i] = (float)(sqrt(s->coulomb) * s->q[i]);
  }
  float beta = (float)s->beta, mb2 = (float)(-s->beta * s->beta);
  float tbs = (float)(2.0 * s->beta / sqrt(M_PI));
  float rc2 = (float)(s->cutoff * s->cutoff);
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t i = 0; i < n; ++i)
    md_row(i, c->index + i * width, c->counts[i], c->xf, c->qf, s->type, c6, c12, nt, L, iL, rc2,
           beta, mb2, tbs, F + 3 * i);
  free(c6);
  free(c12);
}

static void md_pme(orc_md *c, const double *x, double *F) {
  const orc_system *s = c->s;
  const int64_t *K = s->K;
  int order = s->order;
  int64_t n = s->n, h = K[2] / 2 + 1;
  size_t np = (size_t)(K[0] * K[1] * K[2]), nh = (size_t)(K[0] * K[1] * h);
  memset(c->words, 0, sizeof(int64_t) * np);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    if (s->q[i] == 0.0) continue;
    int64_t f[3];
    double M[3][16], dM[3][16];
    for (int a = 0; a < 3; ++a) {
      double w;
      pme_place(x[3 * i + a], s->box[a], K[a], order, &f[a], &w);
      orc_bspline(w, order, M[a], dM[a]);
    }
    for (int t0 = 0; t0 < order; ++t0) {
      int64_t i0 = (f[0] + t0) % K[0];
      for (int t1 = 0; t1 < order; ++t1) {
        int64_t i1 = (f[1] + t1) % K[1];
        double w01 = s->q[i] * M[0][t0] * M[1][t1];
        for (int t2 = 0; t2 < order; ++t2) {
          int64_t i2 = (f[2] + t2) % K[2];
          int64_t word = llrint(w01 * M[2][t2] * 1099511627776.0);
#pragma omp atomic
          c->words[(i0 * K[1] + i1) * K[2] + i2] += word;
        }
      }
    }
  }
#pragma omp parallel for schedule(static)
  for (size_t p = 0; p < np; ++p) c->grid[p] = (double)c->words[p] * (1.0 / 1099511627776.0);
  orc_fft3d_r2c(K, c->grid, c->cplx);
#pragma omp parallel for schedule(static)
  for (size_t p = 0; p < nh; ++p) {
    c->cplx[2 * p] *= c->bc[p];
    c->cplx[2 * p + 1] *= c->bc[p];
  }
  orc_fft3d_c2r(K, c->cplx, c->grid);
  orc_pme_gather(n, x, s->q, s->box, K, order, c->grid, F);
}

static void md_forces(orc_md *c, const double *x) {
  const orc_system *s = c->s;
  int64_t n = s->n;
  double half2 = 0.25 * c->skin * c->skin;
  int moved = 0;
#pragma omp parallel for reduction(| : moved) schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    double d2 = 0.0;
    for (int k = 0; k < 3; ++k) {
      double d = x[3 * i + k] - c->xref[3 * i + k];
      d2 += d * d;
    }
    if (d2 > half2) moved = 1;
  }
  if (moved) md_build(c, x);
  memset(c->F, 0, sizeof(double) * 3 * (size_t)n);
  md_pairs(c, x, c->F);
  if (s->beta > 0.0) {
    double e = 0.0;
    if (s->nexcl_pairs)
      orc_excluded(s->nexcl_pairs, s->excl_pairs, x, s->q, s->box, s->beta, s->coulomb, c->F,
                   &e);
    md_pme(c, x, c->F);
  }
}

/* `steps` steps of velocity Verlet with SETTLE from (x, v), in place. */
int orc_md_run(orc_md *c, double *x, double *v, double dt, int64_t steps) {
  const orc_system *s = c->s;
  int64_t n = s->n;
  if (c->builds == 0) md_build(c, x);
  md_forces(c, x);
  for (int64_t step = 0; step < steps; ++step) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i)
      for (int k = 0; k < 3; ++k) {
        v[3 * i + k] += 0.5 * dt * c->F[3 * i + k] / s->mass[i];
        c->xo[3 * i + k] = x[3 * i + k];
        x[3 * i + k] += dt * v[3 * i + k];
        c->xu[3 * i + k] = x[3 * i + k];
      }
    if (s->nwaters) {
      orc_settle_positions(s->nwaters, s->water_first, s->m_o, s->m_h, s->d_oh, s->d_hh, c->xo,
                           x);
#pragma omp parallel for schedule(static)
      for (int64_t i = 0; i < 3 * n; ++i) v[i] += (x[i] - c->xu[i]) / dt;
    }
    md_forces(c, x);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i)
      for (int k = 0; k < 3; ++k) v[3 * i + k] += 0.5 * dt * c->F[3 * i + k] / s->mass[i];
    if (s->nwaters) orc_settle_velocities(s->nwaters, s->water_first, s->m_o, s->m_h, x, v);
  }
  return 0;
}

void orc_set_threads(int threads) {
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#else
  (void)threads;
#endif
}

/* ======================================================================
 * Langevin dynamics — the g-BAOAB step of the reference's Langevin run with
 * constraints (lib/Driver/DynamicsIR.cpp, geodesic branch; Leimkuhler and
 * Matthews, Proc. R. Soc. A 472, 20160138, 2016) and its O step, the
 * reference's dyn.noise (docs/dialects/dyn.md:441-460,
 * lib/Driver/NoiseIR.cpp): v <- c v + (sigma / sqrt m) R with
 * c = exp(-gamma dt), sigma = sqrt((1 - c^2) kT), R three normals of
 * Philox 4x32-10 (runtime/mdrt.c:289-312) at counter (step lo, step hi,
 * particle, 2 << 24) under the seed, uniforms (w + 1/2) 2^-32, Box-Muller
 * (cos, sin) of the first pair and cos of the second.
 * Philox is pinned against the reference's own mdrtPhilox4x32 compiled from
 * runtime/mdrt.c into oracle/_ref and against the published known-answer
 * vectors of Random123.
 * ====================================================================== */

void orc_philox4x32_10(const uint32_t counter[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t x0 = counter[0], x1 = counter[1], x2 = counter[2], x3 = counter[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; ++r) {
    uint64_t pa = (uint64_t)0xD2511F53u * x0;
    uint64_t pb = (uint64_t)0xCD9E8D57u * x2;
    uint32_t y0 = (uint32_t)(pb >> 32) ^ x1 ^ k0;
    uint32_t y2 = (uint32_t)(pa >> 32) ^ x3 ^ k1;
    x0 = y0;
    x1 = (uint32_t)pb;
    x2 = y2;
    x3 = (uint32_t)pa;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = x0;
  out[1] = x1;
  out[2] = x2;
  out[3] = x3;
}

void orc_langevin_normals(uint64_t seed, int64_t step, uint32_t id, double r[3]) {
  uint32_t ctr[4] = {(uint32_t)(uint64_t)step, (uint32_t)((uint64_t)step >> 32), id, 2u << 24};
  uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)}, w[4];
  orc_philox4x32_10(ctr, key, w);
  double u[4];
  for (int k = 0; k < 4; ++k) u[k] = ((double)w[k] + 0.5) * 0x1.0p-32;
  double a = sqrt(-2.0 * log(u[0])), b = sqrt(-2.0 * log(u[2]));
  r[0] = a * cos(6.283185307179586 * u[1]);
  r[1] = a * sin(6.283185307179586 * u[1]);
  r[2] = b * cos(6.283185307179586 * u[3]);
}

/* x += h v from x0, SETTLE the waters against x0, v += (x_c - x_u) / h. */
static void half_drift(const orc_system *s, double *x, double *v, double h, double *x0,
                       double *xu) {
  int64_t n = s->n;
  memcpy(x0, x, sizeof(double) * 3 * (size_t)n);
  for (int64_t i = 0; i < 3 * n; ++i) x[i] += h * v[i];
  if (!s->nwaters && !s->nclusters) return;
  memcpy(xu, x, sizeof(double) * 3 * (size_t)n);
  if (s->nwaters)
    orc_settle_positions(s->nwaters, s->water_first, s->m_o, s->m_h, s->d_oh, s->d_hh, x0, x);
  if (s->nclusters) orc_shake_positions(s, x0, x);
  for (int64_t i = 0; i < 3 * n; ++i) v[i] += (x[i] - xu[i]) / h;
}

/* The velocity constraints of waters and clusters. */
static void project_velocities(const orc_system *s, const double *x, double *v) {
  if (s->nwaters) orc_settle_velocities(s->nwaters, s->water_first, s->m_o, s->m_h, x, v);
  if (s->nclusters) orc_rattle_velocities(s, x, v);
}

int orc_langevin(const orc_system *s, double *x, double *v, double dt, int64_t steps,
                 int64_t first_step, uint64_t seed, double gamma, double kT,
                 double *trace, int32_t mixed) {
  int64_t n = s->n;
  double *F = (double *)malloc(sizeof(double) * 3 * (size_t)n);
  double *x0 = (double *)malloc(sizeof(double) * 3 * (size_t)n);
  double *xu = (double *)malloc(sizeof(double) * 3 * (size_t)n);
  double e[ORC_NTERMS];
  const double c = exp(-gamma * dt), sigma = sqrt((1.0 - c * c) * kT);
  orc_total(s, x, F, e, mixed);
  for (int64_t step = 0; step < steps; ++step) {
    for (int64_t i = 0; i < n; ++i)
      for (int k = 0; k < 3; ++k) v[3 * i + k] += 0.5 * dt * F[3 * i + k] / s->mass[i];
    project_velocities(s, x, v);
    half_drift(s, x, v, 0.5 * dt, x0, xu);
    for (int64_t i = 0; i < n; ++i) {
      if (!(s->mass[i] > 0.0)) continue;
      double r[3];
      orc_langevin_normals(seed, first_step + step, (uint32_t)i, r);
      double width = sigma / sqrt(s->mass[i]);
      for (int k = 0; k < 3; ++k) v[3 * i + k] = c * v[3 * i + k] + width * r[k];
    }
    project_velocities(s, x, v);
    half_drift(s, x, v, 0.5 * dt, x0, xu);
    orc_total(s, x, F, e, mixed);
    for (int64_t i = 0; i < n; ++i)
      for (int k = 0; k < 3; ++k) v[3 * i + k] += 0.5 * dt * F[3 * i + k] / s->mass[i];
    project_velocities(s, x, v);
    if (trace) {
      double pot = 0.0, kin = 0.0;
      for (int k = 0; k < ORC_NTERMS; ++k) pot += e[k];
      for (int64_t i = 0; i < n; ++i) kin += 0.5 * s->mass[i] * dot3(v + 3 * i, v + 3 * i);
      trace[2 * step] = pot;
      trace[2 * step + 1] = kin;
    }
  }
  free(F);
  free(x0);
  free(xu);
  return 0;
}

/* ======================================================================
 * Initial velocities — System.cpp:1298-1350 and InitialVelocityDraws.h:13-98:
 * splitmix64 seeds xoshiro256**, six uniforms in (0,1] per particle at
 * counter 6 id, Box-Muller normals, constrained, centre of mass at rest,
 * scaled to the temperature.  splitmix64 is pinned by its published first
 * output for seed 0, xoshiro256** by its published first outputs from the
 * state {1, 2, 3, 4}, the draws by their statistics.
 * ====================================================================== */

uint64_t orc_splitmix64(uint64_t *state) {
  *state += 0x9e3779b97f4a7c15ull;
  uint64_t z = *state;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

static inline uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

typedef struct {
  uint64_t s[4];
} xoshiro;

static uint64_t xoshiro_next(xoshiro *g) {
  uint64_t result = rotl64(g->s[1] * 5, 7) * 9;
  uint64_t t = g->s[1] << 17;
  g->s[2] ^= g->s[0];
  g->s[3] ^= g->s[1];
  g->s[1] ^= g->s[2];
  g->s[0] ^= g->s[3];
  g->s[2] ^= t;
  g->s[3] = rotl64(g->s[3], 45);
  return result;
}

/* One draw of xoshiro256** (Blackman and Vigna) on a caller's state. */
uint64_t orc_xoshiro256ss(uint64_t state[4]) {
  xoshiro g;
  for (int k = 0; k < 4; ++k) g.s[k] = state[k];
  const uint64_t r = xoshiro_next(&g);
  for (int k = 0; k < 4; ++k) state[k] = g.s[k];
  return r;
}

/* The pre-constraint normals: three per particle, particle i drawing the
   six uniforms at counters 6 i .. 6 i + 5 (InitialVelocityDraws.h:30-65;
   System.cpp:1298-1311).  Particles are visited in order, so the counter
   6 i is reached by drawing sequentially (the reference seeks otherwise). */
void orc_initial_normals(uint64_t seed, int64_t n, double *out) {
  xoshiro g;
  uint64_t st = seed;
  for (int k = 0; k < 4; ++k) g.s[k] = orc_splitmix64(&st);
  for (int64_t i = 0; i < 3 * n; ++i) {
    double u1 = ((double)(xoshiro_next(&g) >> 11) + 1.0) * 0x1.0p-53;
    double u2 = ((double)(xoshiro_next(&g) >> 11) + 1.0) * 0x1.0p-53;
    double radius = sqrt(-2.0 * log(u1));
    out[i] = radius * cos(6.283185307179586 * u2);
  }
}

void orc_initial_velocities(uint64_t seed, int64_t n, const double *mass,
                            double temperature, const double *x,
                            int64_t nwaters, const int32_t *water_first,
                            int64_t dof, double *v) {
  const double boltzmann = 0.0083144626181532395; /* kJ/(mol K), SI-derived */
  orc_initial_normals(seed, n, v);
  for (int64_t i = 0; i < n; ++i) {
    double width = mass[i] > 0.0 ? sqrt(boltzmann * temperature / mass[i]) : 0.0;
    for (int c = 0; c < 3; ++c) v[3 * i + c] *= width;
  }
  if (n < 2) return;
  /* Velocities that the constraints allow (System.cpp:1316-1321). */
  if (nwaters)
    orc_settle_velocities(nwaters, water_first, mass[water_first[0]],
                          mass[water_first[0] + 1], x, v);
  double total = 0.0, mom[3] = {0, 0, 0};
  for (int64_t i = 0; i < n; ++i) {
    total += mass[i];
    for (int c = 0; c < 3; ++c) mom[c] += mass[i] * v[3 * i + c];
  }
  for (int64_t i = 0; i < n; ++i)
    for (int c = 0; c < 3; ++c) v[3 * i + c] -= mom[c] / total;
  double kin = 0.0;
  for (int64_t i = 0; i < n; ++i) kin += 0.5 * mass[i] * dot3(v + 3 * i, v + 3 * i);
  double target = 0.5 * (double)dof * boltzmann * temperature;
  if (kin > 0.0) {
    double factor = sqrt(target / kin);
    for (int64_t i = 0; i < 3 * n; ++i) v[i] *= factor;
  }
}
