// replicas.hip — many independent systems of uneven size advanced together:
// each chunk of steps of every replica is captured into ONE HIP graph, a
// branch per replica forked from and joined to a master stream, and
// replayed as one launch. The replicas share nothing but the GPU: each
// keeps its own list, PME grid (its own edge), launch shape and Langevin
// noise.  Every branch also copies its x and v aside (so that a replica
// whose list overflowed can be redone alone), stamps the device clock at
// its start and end (the replica's step latency inside the batch), and
// copies its flags into one array that a single transfer brings back.
//
// The batching workload of the north star: replicas of 500 to 8,000 TIP3P
// waters advanced in one persistent, graph-captured step; the reference has
// no counterpart (its device batch, lib/Compiler/Simulation.cpp:3189, batches
// frames of one system; its runs take one system).
namespace mdhip {

__global__ void copy_words8(const long long *__restrict__ src, long long *__restrict__ dst,
                            size_t count) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < count; i += stride) dst[i] = src[i];
}

__global__ void copy_flag_words(const int *__restrict__ src, int *__restrict__ dst) {
  if (threadIdx.x < kNumFlags) dst[threadIdx.x] = src[threadIdx.x];
}

struct ReplicaSet {
  std::vector<Engine *> reps;
  hipStream_t master = nullptr;
  hipEvent_t fork = nullptr;
  std::vector<hipEvent_t> joined;
  unsigned long long *d_stamp = nullptr;  // [2 r]: branch start, [2 r + 1]: branch end
  unsigned long long *h_stamp = nullptr;  // pinned
  int *d_flags = nullptr;                 // kNumFlags per replica
  int *h_flags = nullptr;                 // pinned
  hipGraphExec_t exec = nullptr;
  int exec_steps = 0;
  double exec_dt = 0.0;
  double clock_khz = 0.0;
  std::vector<std::vector<double>> latency;  // per replica: ms per step of each chunk
  int64_t redone = 0;                        // replica chunks redone alone (list growth)

  ~ReplicaSet() { release(); }

  void release() {
    if (exec) hipGraphExecDestroy(exec);
    exec = nullptr;
    invalidate_each();
    for (Engine *e : reps) delete e;
    reps.clear();
    for (hipEvent_t ev : joined) hipEventDestroy(ev);
    joined.clear();
    if (fork) hipEventDestroy(fork);
    fork = nullptr;
    if (master) hipStreamDestroy(master);
    master = nullptr;
    dfree(d_stamp);
    dfree(d_flags);
    if (h_stamp) hipHostFree(h_stamp);
    if (h_flags) hipHostFree(h_flags);
    h_stamp = nullptr;
    h_flags = nullptr;
  }

  void create(const mdrtHipSystemDesc *descs, int count) {
    if (count <= 0) throw HipError("a replica set needs at least one system");
    for (int r = 0; r < count; ++r) {
      Engine *e = new Engine();
      try {
        e->create(descs + r);
      } catch (...) {
        delete e;
        throw;
      }
      reps.push_back(e);
    }
    HIPCHK(hipStreamCreateWithFlags(&master, hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&fork, hipEventDisableTiming));
    joined.resize(count);
    for (auto &ev : joined) HIPCHK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    d_stamp = dalloc<unsigned long long>(2 * (size_t)count);
    d_flags = dalloc<int>((size_t)kNumFlags * count);
    HIPCHK(hipHostMalloc((void **)&h_stamp, 2 * (size_t)count * sizeof(unsigned long long), 0));
    HIPCHK(hipHostMalloc((void **)&h_flags, (size_t)kNumFlags * count * sizeof(int), 0));
    latency.resize(count);
    int dev = 0, khz = 0;
    HIPCHK(hipGetDevice(&dev));
    HIPCHK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
    if (khz <= 0) throw HipError("the device reports no wall-clock rate");
    clock_khz = (double)khz;
  }

  void invalidate() {
    if (exec) hipGraphExecDestroy(exec);
    exec = nullptr;
  }

  // One replica's branch: save x and v, clock, `steps` steps, clock, flags.
  void branch(int r, int steps, double dt) {
    Engine *e = reps[r];
    const size_t words = 3 * (size_t)e->n;
    HIPCHK(hipStreamWaitEvent(e->stream, fork, 0));
    stamp_clock<<<1, 64, 0, e->stream>>>(d_stamp + 2 * r);
    copy_words8<<<blocks_for(words, 256), 256, 0, e->stream>>>(
        reinterpret_cast<const long long *>(e->d_x), reinterpret_cast<long long *>(e->d_save_x),
        words);
    copy_words8<<<blocks_for(words, 256), 256, 0, e->stream>>>(
        reinterpret_cast<const long long *>(e->d_v), reinterpret_cast<long long *>(e->d_save_v),
        words);
    e->steps_seq(steps, dt);
    copy_flag_words<<<1, 64, 0, e->stream>>>(e->d_flags, d_flags + (size_t)kNumFlags * r);
    stamp_clock<<<1, 64, 0, e->stream>>>(d_stamp + 2 * r + 1);
    HIPCHK(hipEventRecord(joined[r], e->stream));
    HIPCHK(hipStreamWaitEvent(master, joined[r], 0));
  }

  // The generation of each replica's own captured graphs when the set's
  // graph was captured: a replica whose parameters or buffers changed since
  // (launch shape, Langevin, box, list growth) bumped its generation.
  std::vector<uint64_t> captured_generation;
  bool replicas_changed() const {
    for (size_t r = 0; r < reps.size(); ++r)
      if (reps[r]->graph_generation != captured_generation[r]) return true;
    return false;
  }

  // Launch mode: 0 = one graph of every replica's branch, launched on the
  // master stream; 1 = one graph per replica (the same branch), each
  // launched on the replica's own stream, so that the runtime's hardware
  // queues run replicas side by side.  Mode 1 is the default: over 256
  // TIP3P replicas of 500-8,000 waters the single graph ran its branches
  // one after another (21.9 ms per step of all, about the sum of their
  // own step times), the per-replica graphs 15.1 ms.
  int mode = 1;
  std::vector<hipGraphExec_t> each;
  int each_steps = 0;
  double each_dt = 0.0;
  void invalidate_each() {
    for (hipGraphExec_t g : each)
      if (g) hipGraphExecDestroy(g);
    each.clear();
  }
  void capture_each(int steps, double dt) {
    if (!each.empty() && each_steps == steps && each_dt == dt && !replicas_changed()) return;
    invalidate_each();
    each.assign(reps.size(), nullptr);
    for (size_t r = 0; r < reps.size(); ++r) {
      Engine *e = reps[r];
      hipGraph_t graph;
      HIPCHK(hipStreamBeginCapture(master, hipStreamCaptureModeThreadLocal));
      HIPCHK(hipEventRecord(fork, master));
      branch((int)r, steps, dt);
      HIPCHK(hipStreamEndCapture(master, &graph));
      const hipError_t st = hipGraphInstantiate(&each[r], graph, nullptr, nullptr, 0);
      HIPCHK(hipGraphDestroy(graph));
      HIPCHK(st);
      (void)e;
    }
    each_steps = steps;
    each_dt = dt;
    captured_generation.resize(reps.size());
    for (size_t r = 0; r < reps.size(); ++r) captured_generation[r] = reps[r]->graph_generation;
  }

  hipGraphExec_t capture(int steps, double dt) {
    if (exec && exec_steps == steps && exec_dt == dt && !replicas_changed()) return exec;
    invalidate();
    hipGraph_t graph;
    HIPCHK(hipStreamBeginCapture(master, hipStreamCaptureModeThreadLocal));
    HIPCHK(hipEventRecord(fork, master));
    for (int r = 0; r < (int)reps.size(); ++r) branch(r, steps, dt);
    HIPCHK(hipStreamEndCapture(master, &graph));
    const hipError_t st = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    HIPCHK(hipGraphDestroy(graph));
    HIPCHK(st);
    exec_steps = steps;
    exec_dt = dt;
    captured_generation.resize(reps.size());
    for (size_t r = 0; r < reps.size(); ++r) captured_generation[r] = reps[r]->graph_generation;
    return exec;
  }

  // Every replica ready for a chunk: list configured, forces of the current
  // positions in its words, its noise counter on the device.
  void prepare_replicas() {
    for (Engine *e : reps) {
      if (e->list_reach != e->cutoff + e->skin) {
        e->configure_list(e->cutoff + e->skin);
        invalidate();
      }
      if (!e->forces_valid) e->evaluate(false);
      e->put_step_count();
    }
  }

  void run(int64_t steps, double dt, int chunk) {
    if (chunk <= 0) throw HipError("replica runs take chunks of at least one step");
    prepare_replicas();
    int64_t done = 0;
    while (done < steps) {
      const int todo = (int)std::min<int64_t>(chunk, steps - done);
      const size_t count = reps.size();
      if (mode == 0) {
        HIPCHK(hipGraphLaunch(capture(todo, dt), master));
      } else {
        capture_each(todo, dt);
        for (size_t r = 0; r < count; ++r) HIPCHK(hipGraphLaunch(each[r], reps[r]->stream));
        for (size_t r = 0; r < count; ++r) {
          HIPCHK(hipEventRecord(joined[r], reps[r]->stream));
          HIPCHK(hipStreamWaitEvent(master, joined[r], 0));
        }
      }
      HIPCHK(hipMemcpyAsync(h_flags, d_flags, (size_t)kNumFlags * count * sizeof(int),
                            hipMemcpyDeviceToHost, master));
      HIPCHK(hipMemcpyAsync(h_stamp, d_stamp, 2 * count * sizeof(unsigned long long),
                            hipMemcpyDeviceToHost, master));
      HIPCHK(hipStreamSynchronize(master));
      for (size_t r = 0; r < count; ++r) {
        Engine *e = reps[r];
        std::vector<int> f(h_flags + kNumFlags * r, h_flags + kNumFlags * (r + 1));
        if (e->grow_if_overflowed(f)) {
          // the list outgrew its buffers: redo this replica's chunk alone
          // from the saved state with the grown list
          e->on_device(e->d_x, e->d_save_x, 3 * (size_t)e->n * sizeof(double));
          e->on_device(e->d_v, e->d_save_v, 3 * (size_t)e->n * sizeof(double));
          e->forces_valid = false;
          e->run(todo, dt, todo, nullptr);
          e->put_step_count();
          ++redone;
          continue;
        }
        e->step_count += todo;
        const unsigned long long t0 = h_stamp[2 * r], t1 = h_stamp[2 * r + 1];
        if (t1 > t0) latency[r].push_back((double)(t1 - t0) / clock_khz / todo);
      }
      done += todo;
    }
  }
};

}  // namespace mdhip

extern "C" {

void *mdrtHipReplicasCreate(const mdrtHipSystemDesc *descs, int32_t count) {
  mdhip::ReplicaSet *set = new mdhip::ReplicaSet();
  if (guarded([&] { set->create(descs, count); }) != 0) {
    delete set;
    return nullptr;
  }
  return set;
}

void mdrtHipReplicasDestroy(void *set) { delete static_cast<mdhip::ReplicaSet *>(set); }

void *mdrtHipReplicasSystem(void *set, int32_t index) {
  mdhip::ReplicaSet *s = static_cast<mdhip::ReplicaSet *>(set);
  if (index < 0 || index >= (int32_t)s->reps.size()) return nullptr;
  return s->reps[index];
}

int mdrtHipReplicasRun(void *set, int64_t steps, double dt, int32_t graph) {
  mdhip::ReplicaSet *s = static_cast<mdhip::ReplicaSet *>(set);
  return guarded([&] { s->run(steps, dt, graph); });
}

int mdrtHipReplicasSetLaunch(void *set, int32_t mode) {
  mdhip::ReplicaSet *s = static_cast<mdhip::ReplicaSet *>(set);
  return guarded([&] {
    if (mode != 0 && mode != 1) throw mdhip::HipError("replica launch mode must be 0 or 1");
    s->mode = mode;
  });
}

int mdrtHipReplicasLatency(void *set, int32_t index, int64_t capacity, double *ms_per_step,
                           int64_t *count) {
  mdhip::ReplicaSet *s = static_cast<mdhip::ReplicaSet *>(set);
  return guarded([&] {
    if (index < 0 || index >= (int32_t)s->reps.size())
      throw mdhip::HipError("no replica " + std::to_string(index));
    const std::vector<double> &l = s->latency[index];
    const int64_t k = std::min<int64_t>(capacity, (int64_t)l.size());
    for (int64_t i = 0; i < k; ++i) ms_per_step[i] = l[i];
    *count = (int64_t)l.size();
  });
}

int mdrtHipReplicasClearLatency(void *set) {
  mdhip::ReplicaSet *s = static_cast<mdhip::ReplicaSet *>(set);
  return guarded([&] {
    for (auto &l : s->latency) l.clear();
    s->redone = 0;
  });
}

}  // extern "C"
