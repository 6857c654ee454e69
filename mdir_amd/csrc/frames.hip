// frames.hip — the frame evaluator: energies, forces and the reverse
// products of many frames of one system in one launch per kernel.
//
// Reference: the device frame batch (lib/Compiler/Simulation.cpp:3189
// evaluateFrames; the per-frame body lib/Driver/FrameBatchEvaluationIR.cpp
// :20-60) and the Python evaluator (python/mdir/_frames.py:458
// FrameEvaluator, :715 evaluate, :327 FrameEnergies.vjp;
// python/mdir/torch.py:139-364).
//
// The frames share the topology and the box; each frame owns a slab of the
// engine's per-frame buffers (positions, places, list, grids, force and
// product words), and every kernel of the build, the pair loop, the PME
// chain and the adjoint runs all frames at once (blockIdx.z).  The list of
// a frame is built from scratch at every evaluation, so its reach is the
// cutoff plus a small margin.
//
// Products (the parameters theta are the per-atom charges and the per-type
// sigma and epsilon, combined by the Lorentz-Berthelot rules):
//   vjp      sum_n [g_n dU_n/dtheta + sum_i Gbar_ni . d2U_n/dx_i dtheta]
//   hvp      H_n v_n per frame, H_n = d2U_n/dx2
//   force matching: the loss sum_n sum_i |F_ni - F_ref,ni|^2 and its
//            theta gradient (vjp with Gbar = -2 (F - F_ref), g = 0), the
//            forward and the adjoint in one captured graph.
#include "adjoint.hip"

namespace mdhip {

struct FrameBatch {
  Engine e;
  int max_frames = 0;
  int loaded = 0;        // frames whose list and potential grid are current
  // per-frame (arena) buffers
  double *d_fin = nullptr;      // 3n: cotangent, reference forces or v (atom order)
  double *d_fout = nullptr;     // 3n: forces or products (atom order)
  float4 *d_cot = nullptr;      // max_places: scaled cotangent in place order
  long long *d_A = nullptr;     // 3 max_places product words
  cpx<float> *d_cplx2 = nullptr;  // dipole grid spectrum
  double *d_meta = nullptr;     // 4: scale c, g / c, loss, -
  double *d_vpart = nullptr;    // 6 per (segment, group) wave: the pair virial
  double *d_vspec = nullptr;    // 6 per block of the excluded and 1-4 pairs
  // shared
  float4 *d_ljadj = nullptr;    // per type pair: C12, C6, 12 C12/sigma, 6 C6/sigma
  float2 *d_ljunit = nullptr;   // per type pair: 4 sigma^12, 4 sigma^6
  int *d_sel = nullptr;         // per type pair: selected table entry, or -1
  // small results read by the host at once: loss [max_frames], flags OR'ed
  // over the frames [kNumFlags], products summed over the frames [3n]
  double *d_result = nullptr;
  double *d_sum = nullptr;      // inside d_result
  std::vector<double> h_result;
  double *d_stage = nullptr;    // max_frames x 3n: frame arrays to and from the host
  AdjParams ap{};
  std::vector<double> sig_t, eps_t;   // per-type parameters
  // force-matching graph
  hipGraphExec_t fm_graph = nullptr;
  int fm_frames = 0;

  ~FrameBatch() {
    if (fm_graph) hipGraphExecDestroy(fm_graph);
    dfree(d_ljadj);
    dfree(d_ljunit);
    dfree(d_sel);
    dfree(d_result);
    dfree(d_stage);
  }

  void alloc_extra() {
    const int n = e.n;
    d_fin = dalloc<double>(3 * (size_t)n);
    d_fout = dalloc<double>(3 * (size_t)n);
    d_cot = dalloc<float4>(e.max_places);
    d_A = dalloc<long long>(3 * (size_t)e.max_places);
    d_meta = dalloc<double>(4);
    d_vpart = dalloc<double>((size_t)kPartStride * e.max_groups * e.nseg);
    d_vspec = dalloc<double>(6 * (size_t)std::max(1, blocks_for(e.nspecial)));
    e.frame_extra = {(void **)&d_fin, (void **)&d_fout, (void **)&d_cot, (void **)&d_A,
                     (void **)&d_meta, (void **)&d_vpart, (void **)&d_vspec};
    if (e.use_pme) {
      d_cplx2 = dalloc<cpx<float>>((size_t)e.pme_grid[0] * e.pme_grid[1] * (e.pme_grid[2] / 2 + 1));
      e.d_vrecip = dalloc<double>(6 * (size_t)e.fft.H());
      e.frame_extra.push_back((void **)&d_cplx2);
      e.frame_extra.push_back((void **)&e.d_vrecip);
    }
  }

  void create(const mdrtHipSystemDesc *d, int frames) {
    if (frames < 1) throw HipError("max_frames must be at least 1");
    if (d->ntuples > 0 || d->nclusters > 0)
      throw HipError("frame batches take systems without bonded tuples or constraint clusters");
    mdrtHipSystemDesc dd = *d;
    dd.skin = 0.02;       // the list is built at every evaluation
    dd.prune_skin = 0.0;
    dd.nwaters = 0;       // no integration
    e.pme_atoms = false;  // the adjoint reads PME inputs in place order
    e.seg_len = 48;       // many frames fill the GPU: long segments
    e.origins_allowed = false;  // the adjoint reads positions about the box centre
    e.create(&dd);
    max_frames = frames;
    // per-type parameters: the diagonal of the tables, which must follow
    // the Lorentz-Berthelot rules wherever a pair has Lennard-Jones
    const int nt = e.ntypes;
    sig_t.resize(nt);
    eps_t.resize(nt);
    for (int a = 0; a < nt; ++a) {
      sig_t[a] = e.h_sigma[a * nt + a];
      eps_t[a] = e.h_eps[a * nt + a];
    }
    for (int a = 0; a < nt; ++a)
      for (int b = 0; b < nt; ++b) {
        const double ee = sqrt(eps_t[a] * eps_t[b]), ss = 0.5 * (sig_t[a] + sig_t[b]);
        const double et = e.h_eps[a * nt + b], st = e.h_sigma[a * nt + b];
        if (fabs(et - ee) > 1e-12 * (1.0 + ee) || (et != 0.0 && fabs(st - ss) > 1e-12))
          throw HipError("the type-pair tables do not follow the Lorentz-Berthelot rules of "
                         "their diagonal");
      }
    set_adj_params();
    d_ljadj = dalloc<float4>((size_t)nt * nt);
    d_ljunit = dalloc<float2>((size_t)nt * nt);
    d_sel = dalloc<int>((size_t)nt * nt);
    const size_t nres = (size_t)frames + kNumFlags + 3 * (size_t)e.n;
    d_result = dalloc<double>(nres);
    d_sum = d_result + frames + kNumFlags;
    h_result.assign(nres, 0.0);
    d_stage = dalloc<double>((size_t)frames * 3 * e.n);
    upload_types();
    alloc_extra();
    e.batch(max_frames);
  }

  void set_adj_params() {
    for (int k = 0; k < 3; ++k) ap.box[k] = (float)e.box[k];
    ap.rc = (float)e.cutoff;
    ap.rc2 = (float)(e.cutoff * e.cutoff);
    ap.rs = (float)e.switch_from;
    ap.width = (float)(e.cutoff - e.switch_from);
    ap.beta = (float)e.beta;
    ap.tbs = (float)(2.0 * e.beta / sqrt(M_PI));
    ap.mb2l2e = (float)(-e.beta * e.beta / M_LN2);
    ap.b2tbs = (float)(2.0 * e.beta * e.beta * 2.0 * e.beta / sqrt(M_PI));
    ap.coulomb = (float)kCoulomb;
    ap.shift_c = e.nbp.shift_c;
    ap.rc12 = (float)pow(e.cutoff, -12.0);
    ap.rc6 = (float)pow(e.cutoff, -6.0);
    ap.ntypes = e.ntypes;
    if (e.lj_modifier == kLJForceSwitch) {
      const double rc = e.cutoff, rs = e.switch_from, w = rc - rs;
      double A[2], B[2], Cc[2];
      const double ns[2] = {12.0, 6.0};
      for (int k = 0; k < 2; ++k) {
        const double nn = ns[k];
        A[k] = -nn * ((nn + 4.0) * rc - (nn + 1.0) * rs) / (pow(rc, nn + 2.0) * w * w);
        B[k] = nn * ((nn + 3.0) * rc - (nn + 1.0) * rs) / (pow(rc, nn + 2.0) * w * w * w);
        Cc[k] = pow(rc, -nn) - A[k] / 3.0 * pow(w, 3.0) - B[k] / 4.0 * pow(w, 4.0);
      }
      ap.fa12 = (float)A[0]; ap.fb12 = (float)B[0]; ap.fc12 = (float)Cc[0];
      ap.fa6 = (float)A[1]; ap.fb6 = (float)B[1]; ap.fc6 = (float)Cc[1];
    }
  }

  // Type-pair tables from the per-type parameters; the engine's force
  // table, the product table and the host constants follow.
  void upload_types() {
    const int nt = e.ntypes;
    std::vector<float4> adj((size_t)nt * nt);
    std::vector<float2> unit((size_t)nt * nt);
    for (int a = 0; a < nt; ++a)
      for (int b = 0; b < nt; ++b) {
        const double ee = sqrt(eps_t[a] * eps_t[b]);
        const double ss = ee != 0.0 ? 0.5 * (sig_t[a] + sig_t[b]) : e.h_sigma[a * nt + b];
        e.h_eps[a * nt + b] = ee;
        e.h_sigma[a * nt + b] = ss;
        const double s6 = pow(ss, 6.0), c6 = 4.0 * ee * s6, c12 = c6 * s6;
        adj[(size_t)a * nt + b] =
            make_float4((float)c12, (float)c6, ss > 0 ? (float)(12.0 * c12 / ss) : 0.0f,
                        ss > 0 ? (float)(6.0 * c6 / ss) : 0.0f);
        unit[(size_t)a * nt + b] = make_float2((float)(4.0 * s6 * s6), (float)(4.0 * s6));
      }
    e.to_device(d_ljadj, adj.data(), adj.size() * sizeof(float4));
    e.to_device(d_ljunit, unit.data(), unit.size() * sizeof(float2));
    e.upload_lj();
    e.host_constants();
  }

  void set_parameters(const double *charges, const double *sigma, const double *epsilon) {
    if (charges) {
      e.h_q.assign(charges, charges + e.n);
      e.to_device(e.d_q, e.h_q.data(), e.n * sizeof(double));
    }
    if (sigma) sig_t.assign(sigma, sigma + e.ntypes);
    if (epsilon)
      for (int a = 0; a < e.ntypes; ++a) {
        if (epsilon[a] < 0) throw HipError("epsilon must not be negative");
        eps_t[a] = epsilon[a];
      }
    upload_types();
    loaded = 0;
  }

  void check_frames(int frames) const {
    if (frames < 1 || frames > max_frames)
      throw HipError("frames must be in [1, " + std::to_string(max_frames) + "]");
  }

  static bool device_pointer(const void *p) {
    hipPointerAttribute_t a;
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
      (void)hipGetLastError();  // pageable host memory
      return false;
    }
    return a.type == hipMemoryTypeDevice;
  }

  // The frame array `src` (frames x count doubles, host or device) into the
  // per-frame buffer `dst` of every frame: one kernel (host arrays first go
  // to a staging buffer in one copy).
  void scatter(double *dst, const double *src, size_t count, int frames) {
    if (!device_pointer(src)) {
      HIPCHK(hipMemcpyAsync(d_stage, src, (size_t)frames * count * sizeof(double),
                            hipMemcpyHostToDevice, e.stream));
      src = d_stage;
    }
    frames_copy<<<dim3(blocks_for((int)count), 1, frames), 256, 0, e.stream>>>(
        src, dst, (int)count, 1, e.fs);
  }
  void gather(double *dst, const double *src, size_t count, int frames) {
    const bool dev = device_pointer(dst);
    frames_copy<<<dim3(blocks_for((int)count), 1, frames), 256, 0, e.stream>>>(
        src, dev ? dst : d_stage, (int)count, 0, e.fs);
    if (!dev)
      HIPCHK(hipMemcpyAsync(dst, d_stage, (size_t)frames * count * sizeof(double),
                            hipMemcpyDeviceToHost, e.stream));
  }

  // Loss per frame and the flags OR'ed over the frames into d_result.
  void launch_collect(int frames) {
    frames_collect<<<1, 1024, 0, e.stream>>>(frames, d_meta, e.d_flags, d_result,
                                            d_result + max_frames, e.fs);
  }
  // d_result to the host (waits for the stream).
  void read_result() {
    HIPCHK(hipMemcpyAsync(h_result.data(), d_result, h_result.size() * sizeof(double),
                          hipMemcpyDeviceToHost, e.stream));
    HIPCHK(hipStreamSynchronize(e.stream));
  }

  // ---- forward ----
  void forward(int frames, bool energy) {
    e.nz = frames;
    e.clear_forces();
    e.request_build();
    e.build();
    e.forces(energy);
  }

  // The flags of the frames as read by read_result(); an overflow regrows
  // the buffers (all frames) and returns true.
  bool regrow_if_overflowed() {
    std::vector<int> m(kNumFlags, 0);
    for (int k = 0; k < kNumFlags; ++k) m[k] = (int)h_result[max_frames + k];
    if (!m[kFlagBrickOverflow] && !m[kFlagOverflowColumn] && !m[kFlagOverflowRow]) return false;
    if (fm_graph) hipGraphExecDestroy(fm_graph);
    fm_graph = nullptr;
    e.unbatch();
    for (void **p : e.frame_extra) dfree(*p);
    e.frame_extra.clear();
    e.grow_if_overflowed(m);
    alloc_extra();
    e.batch(max_frames);
    return true;
  }

  void evaluate(int frames, const double *positions, double *energies, double *forces,
                double *strain) {
    check_frames(frames);
    for (int attempt = 0; attempt < 8; ++attempt) {
      scatter(e.d_x, positions, 3 * (size_t)e.n, frames);
      forward(frames, energies != nullptr || strain != nullptr);
      if (strain) launch_adjoint_k<kAdjVirial>();
      launch_collect(frames);
      read_result();
      if (!regrow_if_overflowed()) {
        loaded = frames;
        if (forces) {
          forces_to_atoms<<<dim3(blocks_for(e.n), 1, frames), 256, 0, e.stream>>>(
              e.n, e.d_place_of, e.d_F, d_fout, e.fs);
          gather(forces, d_fout, 3 * (size_t)e.n, frames);
        }
        if (energies) frame_energies(frames, energies);
        if (strain) frame_strain(frames, strain);
        HIPCHK(hipStreamSynchronize(e.stream));
        return;
      }
    }
    throw HipError("the neighbor lists of the frames kept overflowing");
  }

  // Energy terms of every frame (MDRT_NUM_ENERGIES each), partial sums
  // added in a fixed order.
  void frame_energies(int frames, double *out) {
    const size_t ne = 2 * (size_t)e.max_groups * e.nseg;
    std::vector<double> ep((size_t)frames * ne);
    HIPCHK(hipMemcpy2DAsync(ep.data(), ne * sizeof(double), e.d_epair, e.fs, ne * sizeof(double),
                            frames, hipMemcpyDeviceToHost, e.stream));
    const size_t ns = e.nspecial ? 3 * (size_t)blocks_for(e.nspecial) : 0;
    std::vector<double> es((size_t)frames * ns);
    if (ns)
      HIPCHK(hipMemcpy2DAsync(es.data(), ns * sizeof(double), e.d_espec, e.fs,
                              ns * sizeof(double), frames, hipMemcpyDeviceToHost, e.stream));
    const size_t nr = e.use_pme ? (size_t)e.recip_blocks() : 0;
    std::vector<double> er((size_t)frames * nr);
    if (nr)
      HIPCHK(hipMemcpy2DAsync(er.data(), nr * sizeof(double), e.d_erecip, e.fs,
                              nr * sizeof(double), frames, hipMemcpyDeviceToHost, e.stream));
    HIPCHK(hipStreamSynchronize(e.stream));
    for (int fr = 0; fr < frames; ++fr) {
      double *E = out + (size_t)fr * MDRT_NUM_ENERGIES;
      for (int k = 0; k < MDRT_NUM_ENERGIES; ++k) E[k] = 0.0;
      for (size_t w = 0; w < ne / 2; ++w) {
        E[MDRT_E_LJ] += ep[fr * ne + 2 * w];
        E[MDRT_E_COUL_DIRECT] += ep[fr * ne + 2 * w + 1];
      }
      for (size_t b = 0; b < ns / 3; ++b) {
        E[MDRT_E_COUL_EXCL] += es[fr * ns + 3 * b];
        E[MDRT_E_LJ14] += es[fr * ns + 3 * b + 1];
        E[MDRT_E_COUL14] += es[fr * ns + 3 * b + 2];
      }
      for (size_t b = 0; b < nr; ++b) E[MDRT_E_COUL_RECIP] += er[fr * nr + b];
      E[MDRT_E_COUL_SELF] = e.e_self;
      E[MDRT_E_DISP_CORR] = e.e_disp;
    }
  }

  // dU/d(strain) of every frame (9 values, row alpha, column beta) for
  // x -> (1 + eps) x, h -> h (1 + eps)^T: minus the virial of the pairs,
  // the excluded and 1-4 pairs and the reciprocal sum (partials in fixed
  // order), minus the 1/V host terms (net-charge background, dispersion
  // correction) on the diagonal.
  void frame_strain(int frames, double *out) {
    auto read = [&](const double *src, size_t count) {
      std::vector<double> v((size_t)frames * count);
      if (count)
        HIPCHK(hipMemcpy2DAsync(v.data(), count * sizeof(double), src, e.fs,
                                count * sizeof(double), frames, hipMemcpyDeviceToHost,
                                e.stream));
      return v;
    };
    const size_t np = (size_t)kPartStride * e.max_groups * e.nseg;
    const size_t ns = e.nspecial ? 6 * (size_t)blocks_for(e.nspecial) : 0;
    const size_t nr = e.use_pme ? 6 * (size_t)e.fft.H() : 0;
    const std::vector<double> vp = read(d_vpart, np), vs = read(d_vspec, ns),
                              vr = read(e.d_vrecip, nr);
    HIPCHK(hipStreamSynchronize(e.stream));
    double net = 0;
    for (double q : e.h_q) net += q;
    const double volume = e.box[0] * e.box[1] * e.box[2];
    const double e_net =
        e.beta > 0 ? -kCoulomb * M_PI * net * net / (2.0 * volume * e.beta * e.beta) : 0.0;
    const int ij[6][2] = {{0, 0}, {1, 1}, {2, 2}, {0, 1}, {0, 2}, {1, 2}};
    for (int fr = 0; fr < frames; ++fr) {
      double w[6] = {0, 0, 0, 0, 0, 0};
      for (size_t k = 0; k < np; ++k)
        if (k % kPartStride < 6) w[k % kPartStride] += vp[fr * np + k];
      for (size_t k = 0; k < ns; ++k) w[k % 6] += vs[fr * ns + k];
      for (size_t k = 0; k < nr; ++k) w[k % 6] += vr[fr * nr + k];
      double *S = out + 9 * (size_t)fr;
      for (int c = 0; c < 6; ++c) {
        S[3 * ij[c][0] + ij[c][1]] = -w[c];
        S[3 * ij[c][1] + ij[c][0]] = -w[c];
      }
      for (int a = 0; a < 3; ++a) S[4 * a] -= e_net + e.e_disp;
    }
  }

  // ---- adjoint ----
  template <int MOD, bool COUL, int KIND>
  void launch_adjoint_t() {
    SpecialAdjArgs sa{};
    sa.npairs = e.nspecial;
    sa.pairs = e.d_special;
    sa.is14 = e.d_is14;
    sa.p14_sigma = e.d_p14_sigma;
    sa.p14_eps = e.d_p14_eps;
    sa.p14_qq = e.d_p14_qq;
    sa.x = e.d_x;
    sa.q = e.d_q;
    sa.place_of = e.d_place_of;
    sa.beta = e.beta;
    sa.coulomb = kCoulomb;
    sa.rc = e.cutoff;
    sa.rs = e.switch_from;
    for (int k = 0; k < 3; ++k) sa.box[k] = e.box[k];
    sa.lj_modifier = e.lj_modifier;
    const dim3 grid(std::max(blocks_for(e.max_groups, 4), blocks_for(e.nspecial)), e.nseg + 1,
                    e.nz);
    const AdjTables tab{d_ljadj, d_ljunit, d_sel};
    const bool parts = KIND == kAdjVirial || KIND == kAdjPairTable;
    const size_t lds = (size_t)e.ntypes * e.ntypes * (sizeof(float4) + sizeof(float2) + sizeof(int));
    nb_adjoint<MOD, COUL, KIND><<<grid, 256, lds, e.stream>>>(
        e.d_xq, e.d_type_at, e.d_real_mask, e.d_rows, e.d_row_count, e.d_ex, e.g.row_cap,
        e.seg_len, e.max_groups, e.d_counts, tab, ap, d_cot, d_meta, d_A, parts ? d_vpart : nullptr,
        KIND == kAdjVirial ? d_vspec : nullptr, sa, e.fs);
  }
  template <int KIND>
  void launch_adjoint_k() {
    const bool coul = e.beta > 0;
    switch (e.lj_modifier) {
      case kLJNone:
        coul ? launch_adjoint_t<kLJNone, true, KIND>() : launch_adjoint_t<kLJNone, false, KIND>();
        break;
      case kLJPotentialShift:
        coul ? launch_adjoint_t<kLJPotentialShift, true, KIND>()
             : launch_adjoint_t<kLJPotentialShift, false, KIND>();
        break;
      case kLJForceSwitch:
        coul ? launch_adjoint_t<kLJForceSwitch, true, KIND>()
             : launch_adjoint_t<kLJForceSwitch, false, KIND>();
        break;
      default:
        coul ? launch_adjoint_t<kLJPotentialSwitch, true, KIND>()
             : launch_adjoint_t<kLJPotentialSwitch, false, KIND>();
        break;
    }
  }

  template <int N, int TL, int KIND>
  void pme_adjoint_t() {
    constexpr int R = BrickShape<TL>::R, BX = BrickShape<TL>::BX, BY = BrickShape<TL>::BY;
    static_assert(BX * BY == R, "brick shape");
    const int blocks = ((e.pme_grid[0] + BX - 1) / BX) * ((e.pme_grid[1] + BY - 1) / BY);
    const size_t lds = (size_t)(32 * TL + 8) * e.pme_grid[2];
    pme_spread_dipole<N, TL><<<dim3(blocks, 1, e.nz), 256, lds, e.pstream>>>(
        e.d_xpq, d_cot, e.pmp, e.bp, d_cplx2, e.fft.ax[2].axis, e.fs);
    e.plane_conv(d_cplx2, nullptr, nullptr);
    pme_gather_adj<N, TL, KIND><<<dim3(blocks, 1, e.nz), 256, lds, e.pstream>>>(
        e.d_xpq, d_cot, d_meta, e.pmp, e.bp, e.d_cplx, d_cplx2, e.fft.ax[2].axis, d_A, e.fs);
  }
  template <int N, int KIND>
  void pme_adjoint_n() {
    switch (e.brick_tl) {
      case 8: pme_adjoint_t<N, 8, KIND>(); break;
      case 4: pme_adjoint_t<N, 4, KIND>(); break;
      case 2: pme_adjoint_t<N, 2, KIND>(); break;
      default: pme_adjoint_t<N, 1, KIND>(); break;
    }
  }
  template <int KIND>
  void pme_adjoint() {
    e.pme_bin_launch();  // the forward's gather cleared the brick lists
    if (e.pme_order == 4) pme_adjoint_n<4, KIND>();
    else if (e.pme_order == 6) pme_adjoint_n<6, KIND>();
    else pme_adjoint_n<8, KIND>();
  }

  // The products of the loaded frames for the cotangent in d_fin (mode 0,
  // the energy cotangents in meta[1]) or the force-matching residual
  // against the reference forces in d_fin (mode 1); the pair terms on the
  // engine's stream, the PME terms on the side stream, both adding words.
  template <int KIND>
  void adjoint(int mode) {
    cot_prepare<<<dim3(1, 1, e.nz), kCotThreads, 0, e.stream>>>(mode, e.d_counts, e.d_atom_at, e.d_F,
                                                        d_fin, d_meta, d_cot, e.max_places,
                                                        e.fs);
    clear_words<<<dim3(64, 1, e.nz), 256, 0, e.stream>>>(d_A, 3 * (size_t)e.max_places, e.fs);
    if (e.use_pme) {
      HIPCHK(hipEventRecord(e.ev_fork, e.stream));
      HIPCHK(hipStreamWaitEvent(e.side, e.ev_fork, 0));
      pme_adjoint<KIND>();
      HIPCHK(hipEventRecord(e.ev_join, e.side));
    }
    launch_adjoint_k<KIND>();
    if (e.use_pme) HIPCHK(hipStreamWaitEvent(e.stream, e.ev_join, 0));
    adj_to_atoms<<<dim3(blocks_for(e.n), 1, e.nz), 256, 0, e.stream>>>(e.n, e.d_place_of, d_A,
                                                                      d_meta, d_fout, e.fs);
    if (KIND == kAdjTheta)
      frames_sum<<<blocks_for(3 * e.n), 256, 0, e.stream>>>(e.nz, 3 * e.n, d_fout, d_sum, e.fs);
    launch_collect(e.nz);
  }

  // theta products from the per-atom sums (dq, S, E) and the energy
  // cotangents' share of the host constants.
  void theta_from_sums(const std::vector<double> &s, double gsum, double *dq, double *dsig,
                       double *deps) {
    const int n = e.n, nt = e.ntypes;
    double net = 0;
    for (double q : e.h_q) net += q;
    const double volume = e.box[0] * e.box[1] * e.box[2];
    for (int i = 0; i < n; ++i) {
      double v = s[3 * i];
      if (e.beta > 0)
        v += gsum * (-2.0 * kCoulomb * e.beta / sqrt(M_PI) * e.h_q[i] -
                     kCoulomb * M_PI * net / (volume * e.beta * e.beta));
      if (dq) dq[i] = v;
    }
    std::vector<double> ds(nt, 0.0), de(nt, 0.0);
    for (int i = 0; i < n; ++i) {
      ds[e.h_type[i]] += 0.5 * s[3 * i + 1];
      de[e.h_type[i]] += s[3 * i + 2];
    }
    for (int a = 0; a < nt; ++a) de[a] = eps_t[a] > 0 ? de[a] / (2.0 * eps_t[a]) : 0.0;
    if (e.dispersion && gsum != 0.0) {
      // E_disp = k sum_ab w_ab C6_ab: w from the type counts and the
      // excluded pairs (the host constant's own sum)
      std::vector<double> num(nt, 0.0);
      for (int t : e.h_type) num[t] += 1.0;
      std::vector<double> w((size_t)nt * nt, 0.0);
      for (int a = 0; a < nt; ++a)
        for (int b = 0; b < nt; ++b) w[a * nt + b] = num[a] * (num[b] - (a == b ? 1.0 : 0.0));
      const size_t nex = e.h_excl_pairs.size() / 2;
      for (size_t x = 0; x < nex; ++x) {
        const int a = e.h_type[e.h_excl_pairs[2 * x]], b = e.h_type[e.h_excl_pairs[2 * x + 1]];
        w[a * nt + b] -= 2.0;
      }
      const double nn = n, pairs = nn * (nn - 1.0) - 2.0 * (double)nex;
      const double k = pairs > 0 ? -2.0 * M_PI / (3.0 * volume) * nn * nn /
                                       (e.cutoff * e.cutoff * e.cutoff) / pairs
                                 : 0.0;
      for (int a = 0; a < nt; ++a)
        for (int b = 0; b < nt; ++b) {
          const double ee = e.h_eps[a * nt + b], ss = e.h_sigma[a * nt + b];
          if (ee == 0.0) continue;
          const double c6 = 4.0 * ee * pow(ss, 6.0);
          const double wk = gsum * k * w[a * nt + b];
          const double dc_ds = 24.0 * ee * pow(ss, 5.0) * 0.5;  // per sigma_a, sigma_b
          ds[a] += wk * dc_ds;
          ds[b] += wk * dc_ds;
          de[a] += wk * c6 / (2.0 * eps_t[a]);
          de[b] += wk * c6 / (2.0 * eps_t[b]);
        }
    }
    if (dsig) for (int a = 0; a < nt; ++a) dsig[a] = ds[a];
    if (deps) for (int a = 0; a < nt; ++a) deps[a] = de[a];
  }

  std::vector<double> sums() const {
    const size_t o = (size_t)max_frames + kNumFlags;
    return std::vector<double>(h_result.begin() + o, h_result.begin() + o + 3 * (size_t)e.n);
  }

  void require_loaded(int frames) const {
    if (loaded != frames)
      throw HipError("the products need the frames evaluated first (evaluate with the same "
                     "frame count)");
  }

  void vjp(int frames, const double *g, const double *cot, double *dq, double *dsig,
           double *deps) {
    check_frames(frames);
    require_loaded(frames);
    std::vector<double> gz(frames, 0.0);
    if (g) gz.assign(g, g + frames);
    double gsum = 0;
    for (double v : gz) gsum += v;
    if (cot)
      scatter(d_fin, cot, 3 * (size_t)e.n, frames);
    else
      HIPCHK(hipMemset2DAsync(d_fin, e.fs, 0, 3 * (size_t)e.n * sizeof(double), frames,
                              e.stream));
    HIPCHK(hipMemcpy2DAsync(reinterpret_cast<char *>(d_meta) + sizeof(double), e.fs, gz.data(),
                            sizeof(double), sizeof(double), frames, hipMemcpyHostToDevice,
                            e.stream));
    e.nz = frames;
    adjoint<kAdjTheta>(0);
    read_result();
    theta_from_sums(sums(), gsum, dq, dsig, deps);
  }

  // The product with up to kMaxTablePairs type-pair table entries (a, b)
  // (sigma_ab = sigma_ba and eps_ab = eps_ba each one parameter), summed
  // over the loaded frames: the pair terms, and the energy cotangents'
  // share of the dispersion correction.
  void vjp_pairs(int frames, const double *g, const double *cot, int npairs,
                 const int32_t *type_pairs, double *dsig, double *deps) {
    check_frames(frames);
    require_loaded(frames);
    const int nt = e.ntypes;
    if (npairs < 1 || npairs > kMaxTablePairs)
      throw HipError("1 to " + std::to_string(kMaxTablePairs) + " type pairs per call");
    std::vector<int> sel((size_t)nt * nt, -1);
    for (int p = 0; p < npairs; ++p) {
      const int a = type_pairs[2 * p], b = type_pairs[2 * p + 1];
      if (a < 0 || b < 0 || a >= nt || b >= nt) throw HipError("type pair out of range");
      if (sel[a * nt + b] >= 0) throw HipError("a type pair named twice");
      sel[a * nt + b] = p;
      sel[b * nt + a] = p;
    }
    e.to_device(d_sel, sel.data(), sel.size() * sizeof(int));
    std::vector<double> gz(frames, 0.0);
    if (g) gz.assign(g, g + frames);
    if (cot)
      scatter(d_fin, cot, 3 * (size_t)e.n, frames);
    else
      HIPCHK(hipMemset2DAsync(d_fin, e.fs, 0, 3 * (size_t)e.n * sizeof(double), frames,
                              e.stream));
    HIPCHK(hipMemcpy2DAsync(reinterpret_cast<char *>(d_meta) + sizeof(double), e.fs, gz.data(),
                            sizeof(double), sizeof(double), frames, hipMemcpyHostToDevice,
                            e.stream));
    e.nz = frames;
    cot_prepare<<<dim3(1, 1, e.nz), kCotThreads, 0, e.stream>>>(
        0, e.d_counts, e.d_atom_at, e.d_F, d_fin, d_meta, d_cot, e.max_places, e.fs);
    launch_adjoint_k<kAdjPairTable>();
    const size_t np = (size_t)kPartStride * e.max_groups * e.nseg;
    std::vector<double> vp((size_t)frames * np), meta((size_t)frames * 4);
    HIPCHK(hipMemcpy2DAsync(vp.data(), np * sizeof(double), d_vpart, e.fs, np * sizeof(double),
                            frames, hipMemcpyDeviceToHost, e.stream));
    HIPCHK(hipMemcpy2DAsync(meta.data(), 4 * sizeof(double), d_meta, e.fs, 4 * sizeof(double),
                            frames, hipMemcpyDeviceToHost, e.stream));
    HIPCHK(hipStreamSynchronize(e.stream));
    std::vector<double> ds(npairs, 0.0), de(npairs, 0.0);
    for (int fr = 0; fr < frames; ++fr) {
      std::vector<double> w(kPartStride, 0.0);
      for (size_t k = 0; k < np; ++k) w[k % kPartStride] += vp[fr * np + k];
      const double c = meta[4 * (size_t)fr];  // the frame's cotangent scale
      for (int p = 0; p < npairs; ++p) {
        ds[p] += c * w[2 * p];
        de[p] += c * w[2 * p + 1];
      }
    }
    double gsum = 0;
    for (double v : gz) gsum += v;
    if (e.dispersion && gsum != 0.0) {
      // E_disp = k sum_ab w_ab C6_ab over ordered type pairs
      std::vector<double> num(nt, 0.0);
      for (int t : e.h_type) num[t] += 1.0;
      const size_t nex = e.h_excl_pairs.size() / 2;
      const double volume = e.box[0] * e.box[1] * e.box[2];
      const double nn = e.n, pairs = nn * (nn - 1.0) - 2.0 * (double)nex;
      const double k = pairs > 0 ? -2.0 * M_PI / (3.0 * volume) * nn * nn /
                                       (e.cutoff * e.cutoff * e.cutoff) / pairs
                                 : 0.0;
      for (int p = 0; p < npairs; ++p) {
        const int a = type_pairs[2 * p], b = type_pairs[2 * p + 1];
        auto wab = [&](int x, int y) {
          double v = num[x] * (num[y] - (x == y ? 1.0 : 0.0));
          for (size_t q = 0; q < nex; ++q) {
            const int tx = e.h_type[e.h_excl_pairs[2 * q]], ty = e.h_type[e.h_excl_pairs[2 * q + 1]];
            if ((tx == x && ty == y) || (tx == y && ty == x)) v -= (x == y) ? 2.0 : 1.0;
          }
          return v;
        };
        const double wt = a == b ? wab(a, a) : wab(a, b) + wab(b, a);
        const double ee = e.h_eps[a * nt + b], ss = e.h_sigma[a * nt + b];
        ds[p] += gsum * k * wt * 24.0 * ee * pow(ss, 5.0);
        de[p] += gsum * k * wt * 4.0 * pow(ss, 6.0);
      }
    }
    if (dsig) std::copy(ds.begin(), ds.end(), dsig);
    if (deps) std::copy(de.begin(), de.end(), deps);
  }

  void hvp(int frames, const double *v, double *out) {
    check_frames(frames);
    require_loaded(frames);
    scatter(d_fin, v, 3 * (size_t)e.n, frames);
    std::vector<double> gz(frames, 0.0);
    HIPCHK(hipMemcpy2DAsync(reinterpret_cast<char *>(d_meta) + sizeof(double), e.fs, gz.data(),
                            sizeof(double), sizeof(double), frames, hipMemcpyHostToDevice,
                            e.stream));
    e.nz = frames;
    adjoint<kAdjHessian>(0);
    gather(out, d_fout, 3 * (size_t)e.n, frames);
    HIPCHK(hipStreamSynchronize(e.stream));
  }

  // A launch shape for every frame: the PME brick width (TL complex z
  // lines, at most), the plane transform and the pair-loop segment length.
  // The per-frame list and brick buffers live in the frame arena, so it is
  // laid out again around the change. The evaluator holds no evaluated
  // frames afterwards.
  void set_shape(int tl, int transform, int seg) {
    if (fm_graph) hipGraphExecDestroy(fm_graph);
    fm_graph = nullptr;
    e.unbatch();
    for (void **p : e.frame_extra) dfree(*p);
    e.frame_extra.clear();
    e.set_segment(seg);
    if (e.use_pme) {
      e.brick_tl_max = tl;
      e.make_bricks();
    }
    alloc_extra();
    e.batch(max_frames);
    if (e.use_pme) e.set_pme_transform(transform);
    loaded = 0;
  }

  // Launch-shape trial of the frame evaluator (the engine's autotune, for
  // the frames): times the graph-replayed force-matching call on the
  // caller's `frames` frames for every shape `what` names (bit 0: PME brick
  // widths that fit LDS, bit 1: plane transforms, bit 2: pair-loop segment
  // lengths), one untimed call (list build, capture) and `repeats` timed
  // ones each, in two passes (the faster counts), and keeps the fastest
  // unless the current shape is within the trial margin. ms_out as the
  // engine's autotune, in ms per call.
  void autotune(int what, int frames, const double *positions, const double *f_ref,
                int repeats, double *ms_out) {
    for (int k = 0; k < Engine::kTrialShapes; ++k) ms_out[k] = NAN;
    check_frames(frames);
    if (what <= 0 || repeats <= 0) return;
    if (!e.use_pme) what &= 4;
    const bool mfma_ok = e.use_pme && e.plane_fft() && Engine::mfma_edge(e.pme_grid[0]) &&
                         Engine::mfma_edge(e.pme_grid[1]);
    const int tl0 = e.use_pme ? e.brick_tl : 8, tr0 = e.pme_transform;
    const Engine::Shape cur{Engine::seg_index(e.seg_len), tr0,
                            tl0 == 8 ? 0 : (tl0 == 4 ? 1 : (tl0 == 2 ? 2 : 3))};
    std::vector<double> loss(frames), dq(e.n), ds(e.ntypes), de(e.ntypes);
    auto call = [&]() {
      force_matching(frames, positions, f_ref, loss.data(), dq.data(), ds.data(), de.data(), 1);
    };
    const Engine::Shape keep = Engine::trial(
        Engine::trial_shapes(what, e.use_pme ? e.pme_grid[2] : 0, mfma_ok, tl0, tr0, e.seg_len),
        cur,
        [&](const Engine::Shape &c) {
          set_shape(8 >> c.k, c.tr, Engine::kSegLens[c.s]);
          call();
          const auto t0 = std::chrono::steady_clock::now();
          for (int r = 0; r < repeats; ++r) call();
          const auto t1 = std::chrono::steady_clock::now();
          return std::chrono::duration<double, std::milli>(t1 - t0).count() / repeats;
        },
        ms_out);
    set_shape(8 >> keep.k, keep.tr, Engine::kSegLens[keep.s]);
  }

  // Forward, residual, adjoint and the sum over frames, as one graph per
  // frame count (captured at the first call, replayed after).
  void fm_sequence(int frames) {
    forward(frames, false);
    adjoint<kAdjTheta>(1);
  }

  void force_matching(int frames, const double *positions, const double *f_ref, double *loss,
                      double *dq, double *dsig, double *deps, int use_graph) {
    check_frames(frames);
    for (int attempt = 0; attempt < 8; ++attempt) {
      scatter(e.d_x, positions, 3 * (size_t)e.n, frames);
      scatter(d_fin, f_ref, 3 * (size_t)e.n, frames);
      if (use_graph) {
        if (fm_graph && fm_frames != frames) {
          hipGraphExecDestroy(fm_graph);
          fm_graph = nullptr;
        }
        if (!fm_graph) {
          hipGraph_t graph;
          HIPCHK(hipStreamBeginCapture(e.stream, hipStreamCaptureModeThreadLocal));
          fm_sequence(frames);
          HIPCHK(hipStreamEndCapture(e.stream, &graph));
          HIPCHK(hipGraphInstantiate(&fm_graph, graph, nullptr, nullptr, 0));
          HIPCHK(hipGraphDestroy(graph));
          fm_frames = frames;
        }
        HIPCHK(hipGraphLaunch(fm_graph, e.stream));
      } else {
        fm_sequence(frames);
      }
      read_result();
      if (!regrow_if_overflowed()) {
        loaded = frames;
        if (loss) std::copy(h_result.begin(), h_result.begin() + frames, loss);
        theta_from_sums(sums(), 0.0, dq, dsig, deps);
        return;
      }
    }
    throw HipError("the neighbor lists of the frames kept overflowing");
  }
};

}  // namespace mdhip
