// trajectory.hip — DCD trajectories of a system, written from staged
// copies on a thread of their own while the steps go on (the reference's
// driver: lib/Driver/Trajectory.cpp:48-206 for the file, lib/Driver/
// Output.cpp for the thread of the frames over the runtime's staging
// entry points, runtime/mdrt_cuda.c:3611-3845).
//
// The file is the reference's DCD: records framed by their byte length
// before and after; the CORD record with 20 numbers (frames, step of the
// first frame, period, frames x period, the time step in AKMA units of
// 48.88821 fs as f32 in number 9, the cell flag in 10, version 24 in 19);
// one 80-character title line; the particle count; then per frame the cell
// (a, cos gamma, b, cos beta, cos alpha, c: Angstrom and cosines, f64) and
// the x, y and z records, f32 Angstrom in the order of the input.  The
// header is rewritten after every frame, so a file cut short by a failure
// still counts what it holds.
//
// Units: positions nm x (1 / 0.1) rounded once to f32, cell edges
// nm / 0.1, as the reference scales them (Output.cpp:1239-1240, :1701).
#include <condition_variable>
#include <deque>
#include <thread>
#include <unistd.h>

namespace mdhip {

class DcdFile {
 public:
  ~DcdFile() { close(); }

  void create(const std::string &path, int64_t n, int64_t first, int64_t period,
              double timestep_ps, bool cell) {
    close();
    check_sizes(n, period);
    f_ = fopen(path.c_str(), "wb");
    if (!f_) throw HipError("cannot write '" + path + "'");
    set(n, first, period, timestep_ps, cell, 0);
    put_header();
    fflush(f_);
  }

  // Continues a file this writer wrote: the same particles, period and cell
  // flag; keeps its first `keep` frames and cuts what follows.
  void reopen(const std::string &path, int64_t n, int64_t keep, int64_t period,
              double timestep_ps, bool cell) {
    close();
    check_sizes(n, period);
    f_ = fopen(path.c_str(), "r+b");
    if (!f_) throw HipError("'" + path + "' cannot be opened to continue it");
    int32_t head[22];  // length, CORD, 20 numbers (+ the closing length)
    int32_t title[23], count[3];
    char tag[4];
    const bool read = fread(head, 4, 1, f_) == 1 && head[0] == 84 &&
                      fread(tag, 4, 1, f_) == 1 && memcmp(tag, "CORD", 4) == 0 &&
                      fread(head + 1, 4, 21, f_) == 21 && head[21] == 84 &&
                      fread(title, 4, 23, f_) == 23 && title[0] == 84 && title[22] == 84 &&
                      fread(count, 4, 3, f_) == 3 && count[0] == 4 && count[2] == 4;
    // head[1 + k] is number k
    if (!read || head[1 + 10] != (cell ? 1 : 0) || head[1 + 19] != 24) {
      close();
      throw HipError("'" + path + "' is not a trajectory this writer wrote for this cell");
    }
    if (count[1] != (int32_t)n) {
      close();
      throw HipError("'" + path + "' holds " + std::to_string(count[1]) +
                     " particles, and the system " + std::to_string(n));
    }
    if (head[1 + 2] != (int32_t)period) {
      close();
      throw HipError("'" + path + "' has a frame every " + std::to_string(head[1 + 2]) +
                     " steps, and the run writes one every " + std::to_string(period));
    }
    const long header = ftell(f_);
    fseek(f_, 0, SEEK_END);
    const long length = ftell(f_);
    const long held = (length - header) / frame_bytes(n, cell);
    if (keep < 0 || held < keep) {
      close();
      throw HipError("'" + path + "' holds " + std::to_string(held) + " frames, fewer than " +
                     std::to_string(keep));
    }
    const long kept = header + (long)keep * frame_bytes(n, cell);
    fflush(f_);
    if (kept != length && ftruncate(fileno(f_), kept) != 0) {
      close();
      throw HipError("'" + path + "' cannot be cut to " + std::to_string(keep) + " frames");
    }
    set(n, head[1 + 1], period, timestep_ps, cell, keep);
    rewind(f_);
    put_header();
    fseek(f_, kept, SEEK_SET);
    fflush(f_);
  }

  // One frame from positions in nm (n x 3, f64, the order of the input)
  // and the orthorhombic cell in nm.
  void append(const double *x_nm, const double *box_nm) {
    if (!f_) throw HipError("the trajectory is closed");
    const double to_angstrom = 1.0 / 0.1;
    if (cell_) {
      const double cell[6] = {box_nm[0] / 0.1, 0.0, box_nm[1] / 0.1, 0.0, 0.0, box_nm[2] / 0.1};
      put_record(cell, sizeof(cell));
    }
    row_.resize((size_t)n_);
    for (int c = 0; c < 3; ++c) {
      for (int64_t i = 0; i < n_; ++i) row_[i] = (float)(to_angstrom * x_nm[3 * i + c]);
      put_record(row_.data(), (int32_t)(n_ * sizeof(float)));
    }
    ++frames_;
    const long end = ftell(f_);
    rewind(f_);
    put_header();
    fseek(f_, end, SEEK_SET);
    if (fflush(f_) != 0 || ferror(f_)) throw HipError("a frame could not be written");
  }

  void close() {
    if (f_) fclose(f_);
    f_ = nullptr;
  }
  int64_t frames() const { return frames_; }
  int64_t particles() const { return n_; }

 private:
  static long frame_bytes(int64_t n, bool cell) {
    return (cell ? 8 + 48 : 0) + 3 * (8 + 4 * (long)n);
  }
  static void check_sizes(int64_t n, int64_t period) {
    if (n <= 0 || n > (int64_t)(INT32_MAX / 4) - 8) throw HipError("DCD: particle count out of range");
    if (period <= 0 || period > INT32_MAX) throw HipError("DCD: the period must be a positive int32");
  }
  void set(int64_t n, int64_t first, int64_t period, double dt, bool cell, int64_t frames) {
    n_ = n;
    first_ = first;
    period_ = period;
    dt_ = dt;
    cell_ = cell;
    frames_ = frames;
  }
  void put_record(const void *data, int32_t bytes) {
    fwrite(&bytes, 4, 1, f_);
    fwrite(data, 1, (size_t)bytes, f_);
    fwrite(&bytes, 4, 1, f_);
  }
  void put_header() {
    struct {
      char tag[4];
      int32_t numbers[20];
    } head;
    memcpy(head.tag, "CORD", 4);
    memset(head.numbers, 0, sizeof(head.numbers));
    head.numbers[0] = (int32_t)frames_;
    head.numbers[1] = (int32_t)first_;
    head.numbers[2] = (int32_t)period_;
    head.numbers[3] = (int32_t)(frames_ * period_);
    const float akma = (float)(dt_ / 0.04888821);  // ps per AKMA time unit
    memcpy(&head.numbers[9], &akma, 4);
    head.numbers[10] = cell_ ? 1 : 0;
    head.numbers[19] = 24;
    put_record(&head, (int32_t)sizeof(head));
    struct {
      int32_t lines;
      char text[80];
    } title;
    title.lines = 1;
    memset(title.text, ' ', sizeof(title.text));
    const char *who = "Written by mdir_amd";
    memcpy(title.text, who, strlen(who));
    put_record(&title, (int32_t)sizeof(title));
    const int32_t count = (int32_t)n_;
    put_record(&count, 4);
  }

  FILE *f_ = nullptr;
  int64_t n_ = 0, first_ = 0, period_ = 0, frames_ = 0;
  double dt_ = 0.0;
  bool cell_ = true;
  std::vector<float> row_;
};

// The thread of the frames: a ring of three staging slots.  frame() stages
// the positions on the engine's stream (a copy within the device, then the
// runtime's stream of copies takes them to pinned host memory beside the
// next steps) and queues the slot; the thread waits for each slot's copies
// in order and writes the frame.  A failed wait is kept and reported on the
// thread of the steps at its next call (mdrtStagingFail); a failed write
// drops the frames queued after it and is reported as an error there.
class Trajectory {
 public:
  Trajectory(Engine *e, const std::string &path, int64_t first, int64_t period,
             double timestep_ps, int64_t keep)
      : e_(e) {
    if (e->nz != 1) throw HipError("a trajectory is of one system, not of a frame batch");
    if (keep < 0)
      file_.create(path, e->n, first, period, timestep_ps, true);
    else
      file_.reopen(path, e->n, keep, period, timestep_ps, true);
    written_ = file_.frames();
    worker_ = std::thread([this] { work(); });
  }

  ~Trajectory() {
    {
      std::lock_guard<std::mutex> lock(m_);
      stopping_ = true;
    }
    cv_.notify_all();
    if (worker_.joinable()) worker_.join();
    file_.close();
    for (Slot &s : slots_)
      if (s.staging) mdrtStagingDestroy(s.staging);
  }

  void frame(int64_t step) {
    report();
    const int k = next_;
    next_ = (next_ + 1) % kSlots;
    Slot &s = slots_[k];
    {
      std::unique_lock<std::mutex> lock(m_);
      cv_.wait(lock, [&] { return !s.busy; });
    }
    report();
    if (!s.staging) s.staging = mdrtStagingCreate();
    void *before = mdrtUseStream(e_->stream);
    s.host = static_cast<const double *>(
        mdrtStagingCopy(s.staging, 0, e_->d_x, 3 * (uint64_t)e_->n * sizeof(double)));
    mdrtStagingEnd(s.staging);
    mdrtRestoreStream(before);
    s.step = step;
    for (int c = 0; c < 3; ++c) s.box[c] = e_->box[c];
    {
      std::lock_guard<std::mutex> lock(m_);
      s.busy = true;
      queue_.push_back(k);
      ++staged_;
    }
    cv_.notify_all();
  }

  void flush() {
    {
      std::unique_lock<std::mutex> lock(m_);
      cv_.wait(lock, [&] { return queue_.empty(); });
    }
    report();
  }

  int64_t frames() {
    std::lock_guard<std::mutex> lock(m_);
    return written_ + (int64_t)queue_.size();
  }

 private:
  static constexpr int kSlots = 3;
  struct Slot {
    void *staging = nullptr;
    const double *host = nullptr;
    int64_t step = 0;
    double box[3] = {0, 0, 0};
    bool busy = false;
  };

  // Failures the thread met, reported on the thread of the steps.
  void report() {
    int32_t failed;
    const char *call;
    std::string write_error;
    {
      std::lock_guard<std::mutex> lock(m_);
      failed = failed_;
      call = failed_call_;
      write_error = write_error_;
    }
    if (failed) mdrtStagingFail(failed, call);  // ends the process
    if (!write_error.empty()) throw HipError("trajectory: " + write_error);
  }

  void work() {
    for (;;) {
      int k;
      {
        std::unique_lock<std::mutex> lock(m_);
        cv_.wait(lock, [&] { return stopping_ || !queue_.empty(); });
        if (queue_.empty()) return;
        k = queue_.front();
      }
      Slot &s = slots_[k];
      bool dropped;
      {
        std::lock_guard<std::mutex> lock(m_);
        dropped = failed_ != 0 || !write_error_.empty();
      }
      bool written = false;
      if (!dropped) {
        const char *call = nullptr;
        const int32_t result = mdrtStagingWait(s.staging, &call);
        if (result != 0) {
          std::lock_guard<std::mutex> lock(m_);
          failed_ = result;
          failed_call_ = call;
        } else {
          try {
            file_.append(s.host, s.box);
            written = true;
          } catch (const std::exception &ex) {
            std::lock_guard<std::mutex> lock(m_);
            write_error_ = ex.what();
          }
        }
      }
      {
        std::lock_guard<std::mutex> lock(m_);
        written_ += written;
        queue_.pop_front();
        s.busy = false;
      }
      cv_.notify_all();
    }
  }

  Engine *e_;
  DcdFile file_;
  Slot slots_[kSlots];
  int next_ = 0;
  std::mutex m_;
  std::condition_variable cv_;
  std::deque<int> queue_;
  std::thread worker_;
  bool stopping_ = false;
  int64_t staged_ = 0;
  int64_t written_ = 0;  // frames in the file
  int32_t failed_ = 0;
  const char *failed_call_ = nullptr;
  std::string write_error_;
};

}  // namespace mdhip

extern "C" {

void *mdrtHipTrajectoryOpen(void *system, const char *path, int64_t first, int64_t period,
                            double timestep_ps, int64_t keep) {
  mdhip::Trajectory *t = nullptr;
  if (guarded([&] {
        if (!system || !path) throw mdhip::HipError("trajectory: no system or no path");
        t = new mdhip::Trajectory(static_cast<Engine *>(system), path, first, period,
                                  timestep_ps, keep);
      }) != 0)
    return nullptr;
  return t;
}

int mdrtHipTrajectoryFrame(void *trajectory, int64_t step) {
  return guarded([&] { static_cast<mdhip::Trajectory *>(trajectory)->frame(step); });
}

int mdrtHipTrajectoryFlush(void *trajectory) {
  return guarded([&] { static_cast<mdhip::Trajectory *>(trajectory)->flush(); });
}

int64_t mdrtHipTrajectoryFrames(void *trajectory) {
  return static_cast<mdhip::Trajectory *>(trajectory)->frames();
}

int mdrtHipTrajectoryClose(void *trajectory) {
  auto *t = static_cast<mdhip::Trajectory *>(trajectory);
  const int rc = guarded([&] { t->flush(); });
  delete t;
  return rc;
}

int mdrtHipDcdWrite(const char *path, int64_t n, int64_t frames, int64_t first, int64_t period,
                    double timestep_ps, const double *positions, const double *cells) {
  return guarded([&] {
    if (!path || (frames > 0 && !positions)) throw mdhip::HipError("DCD: no path or no positions");
    if (frames < 0) throw mdhip::HipError("DCD: a negative frame count");
    mdhip::DcdFile f;
    f.create(path, n, first, period, timestep_ps, cells != nullptr);
    const double none[3] = {0.0, 0.0, 0.0};
    for (int64_t k = 0; k < frames; ++k)
      f.append(positions + 3 * n * k, cells ? cells + 3 * k : none);
    f.close();
  });
}

}  // extern "C"
