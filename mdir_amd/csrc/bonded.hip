// bonded.hip — bonded tuples in f64, a thread per tuple (the reference's
// tuple kernels, lib/Conversion/MDExecKernels/MDExecKernels.cpp:804-1011,
// over the terms of lib/Driver/TermBond.cpp, TermUreyBradley.cpp,
// TermAngle.cpp, TermDihedral.cpp, TermImproper.cpp):
//
//   bond, Urey-Bradley   (k/2) (r - r0)^2            r = |x_j - x_i|
//   angle                (k/2) (theta - theta0)^2    apex j
//   dihedral             k (1 + cos(n phi - phase))  phi of i-j-k-l, IUPAC
//   harmonic improper    (k/2) wrap(phi - phi0)^2    wrap to [-pi, pi)
//
// Differences use the minimum image.  The forces go to the 64-bit words of
// the atoms' places, as the excluded and 1-4 pairs' do; the kernel runs on
// the list's stream after the build of the step has placed the atoms, as a
// launch of its own (its f64 registers would lower the pair kernel's
// occupancy were it a part of that launch).  The tuples are sorted by kind,
// so a wave mostly runs one kind.
namespace mdhip {

enum TupleKind { kTupleBond = 0, kTupleUreyBradley = 1, kTupleAngle = 2, kTupleDihedral = 3,
                 kTupleImproper = 4, kTupleKinds = 5 };

struct TupleArgs {
  int ntuples;
  const int4 *atoms;      // i, j, k, l (unused members 0)
  const int *kind;        // TupleKind
  const double4 *param;   // bond (k, r0); angle (k, theta0); dihedral (k, n, phase);
                          // improper (k, phi0)
  double *energy_partial; // kTupleKinds per block, or null
};

__device__ __forceinline__ void min_image(double *d, const double *box) {
#pragma unroll
  for (int c = 0; c < 3; ++c) d[c] -= box[c] * rint(d[c] / box[c]);
}

__device__ __forceinline__ void add_force_words(long long *__restrict__ F, int place,
                                                const double *f) {
  unsigned long long *u = reinterpret_cast<unsigned long long *>(F + 3 * place);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const long long w = __double2ll_rn(f[c] * kForceScale);
    if (w) atomicAdd(u + c, (unsigned long long)w);
  }
}

// The energy of one tuple and the forces on its members (f[m], m < 4).
__device__ inline double tuple_force(int kind, const double *xi, const double *xj,
                                     const double *xk, const double *xl, double4 p,
                                     const double *box, double (*f)[3]) {
  for (int m = 0; m < 4; ++m) f[m][0] = f[m][1] = f[m][2] = 0.0;
  if (kind == kTupleBond || kind == kTupleUreyBradley) {
    double d[3] = {xj[0] - xi[0], xj[1] - xi[1], xj[2] - xi[2]};
    min_image(d, box);
    const double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    const double dr = r - p.y;
    const double g = r > 0.0 ? p.x * dr / r : 0.0;  // dU/dr / r
    for (int c = 0; c < 3; ++c) {
      f[0][c] = g * d[c];
      f[1][c] = -g * d[c];
    }
    return 0.5 * p.x * dr * dr;
  }
  if (kind == kTupleAngle) {
    double a[3] = {xi[0] - xj[0], xi[1] - xj[1], xi[2] - xj[2]};
    double b[3] = {xk[0] - xj[0], xk[1] - xj[1], xk[2] - xj[2]};
    min_image(a, box);
    min_image(b, box);
    const double aa = a[0] * a[0] + a[1] * a[1] + a[2] * a[2];
    const double bb = b[0] * b[0] + b[1] * b[1] + b[2] * b[2];
    const double ab = a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
    const double cx = a[1] * b[2] - a[2] * b[1], cy = a[2] * b[0] - a[0] * b[2],
                 cz = a[0] * b[1] - a[1] * b[0];
    const double s = sqrt(cx * cx + cy * cy + cz * cz);  // |a x b|
    const double theta = atan2(s, ab);
    const double dt = theta - p.y;
    const double dudt = p.x * dt;
    if (s > 0.0) {
      // dtheta/da = (a (a.b) - b |a|^2) / (|a|^2 s), and likewise for b
      for (int c = 0; c < 3; ++c) {
        const double ga = (a[c] * ab - b[c] * aa) / (aa * s);
        const double gb = (b[c] * ab - a[c] * bb) / (bb * s);
        f[0][c] = -dudt * ga;
        f[2][c] = -dudt * gb;
        f[1][c] = dudt * (ga + gb);
      }
    }
    return 0.5 * p.x * dt * dt;
  }
  // dihedrals: b1 = j - i, b2 = k - j, b3 = l - k; m = b1 x b2, n = b2 x b3;
  // phi = atan2(|b2| b1 . n, m . n)
  double b1[3] = {xj[0] - xi[0], xj[1] - xi[1], xj[2] - xi[2]};
  double b2[3] = {xk[0] - xj[0], xk[1] - xj[1], xk[2] - xj[2]};
  double b3[3] = {xl[0] - xk[0], xl[1] - xk[1], xl[2] - xk[2]};
  min_image(b1, box);
  min_image(b2, box);
  min_image(b3, box);
  const double m[3] = {b1[1] * b2[2] - b1[2] * b2[1], b1[2] * b2[0] - b1[0] * b2[2],
                       b1[0] * b2[1] - b1[1] * b2[0]};
  const double n[3] = {b2[1] * b3[2] - b2[2] * b3[1], b2[2] * b3[0] - b2[0] * b3[2],
                       b2[0] * b3[1] - b2[1] * b3[0]};
  const double mm = m[0] * m[0] + m[1] * m[1] + m[2] * m[2];
  const double nn = n[0] * n[0] + n[1] * n[1] + n[2] * n[2];
  const double b2b2 = b2[0] * b2[0] + b2[1] * b2[1] + b2[2] * b2[2];
  const double lb2 = sqrt(b2b2);
  const double y = lb2 * (b1[0] * n[0] + b1[1] * n[1] + b1[2] * n[2]);
  const double xx = m[0] * n[0] + m[1] * n[1] + m[2] * n[2];
  const double phi = atan2(y, xx);
  double u, dudp;
  if (kind == kTupleDihedral) {
    const double arg = p.y * phi - p.z;
    u = p.x * (1.0 + cos(arg));
    dudp = -p.x * p.y * sin(arg);
  } else {  // harmonic improper
    double d = phi - p.y;
    d -= 2.0 * M_PI * floor((d + M_PI) / (2.0 * M_PI));
    u = 0.5 * p.x * d * d;
    dudp = p.x * d;
  }
  if (mm > 0.0 && nn > 0.0 && lb2 > 0.0) {
    // dphi/dx_i = -|b2| m / |m|^2, dphi/dx_l = |b2| n / |n|^2, and with
    // s1 = b1.b2 / |b2|^2, s3 = b3.b2 / |b2|^2 the middle atoms'
    // dphi/dx_j = -(1 + s1) g_i + s3 g_l, dphi/dx_k = -(1 + s3) g_l + s1 g_i
    // (the four sum to zero)
    const double s1 = (b1[0] * b2[0] + b1[1] * b2[1] + b1[2] * b2[2]) / b2b2;
    const double s3 = (b3[0] * b2[0] + b3[1] * b2[1] + b3[2] * b2[2]) / b2b2;
    for (int c = 0; c < 3; ++c) {
      const double gi = -lb2 * m[c] / mm;
      const double gl = lb2 * n[c] / nn;
      const double gj = -(1.0 + s1) * gi + s3 * gl;
      const double gk = -(1.0 + s3) * gl + s1 * gi;
      f[0][c] = -dudp * gi;
      f[1][c] = -dudp * gj;
      f[2][c] = -dudp * gk;
      f[3][c] = -dudp * gl;
    }
  }
  return u;
}

// Grid-stride over the tuples; with energy partials, kTupleKinds sums per
// block in a fixed order.
__global__ void __launch_bounds__(256) bonded_tuples(TupleArgs ta, const double *__restrict__ x,
                                                     const int *__restrict__ place_of,
                                                     double bx, double by, double bz,
                                                     long long *__restrict__ F) {
  __shared__ double s_e[1][256];
  const double box[3] = {bx, by, bz};
  const int block = blockIdx.x, nblocks = gridDim.x;
  double e[kTupleKinds] = {0.0, 0.0, 0.0, 0.0, 0.0};
  for (int t = block * (int)blockDim.x + (int)threadIdx.x; t < ta.ntuples;
       t += nblocks * (int)blockDim.x) {
    const int4 a = ta.atoms[t];
    const int kind = ta.kind[t];
    const double4 p = ta.param[t];
    const double *xi = x + 3 * a.x, *xj = x + 3 * a.y, *xk = x + 3 * a.z, *xl = x + 3 * a.w;
    double f[4][3];
    const double u = tuple_force(kind, xi, xj, xk, xl, p, box, f);
#pragma unroll
    for (int k = 0; k < kTupleKinds; ++k) e[k] += k == kind ? u : 0.0;
    add_force_words(F, place_of[a.x], f[0]);
    add_force_words(F, place_of[a.y], f[1]);
    if (kind >= kTupleAngle) add_force_words(F, place_of[a.z], f[2]);
    if (kind >= kTupleDihedral) add_force_words(F, place_of[a.w], f[3]);
  }
  if (ta.energy_partial) {
    for (int k = 0; k < kTupleKinds; ++k) {
      s_e[0][threadIdx.x] = e[k];
      __syncthreads();
      for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) s_e[0][threadIdx.x] += s_e[0][threadIdx.x + s];
        __syncthreads();
      }
      if (threadIdx.x == 0) ta.energy_partial[kTupleKinds * block + k] = s_e[0][0];
      __syncthreads();
    }
  }
}

}  // namespace mdhip
