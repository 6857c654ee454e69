// noalloc.hip — the forward transform of the PME grid for callers that
// cannot allocate device memory (code under HIP graph capture, tight
// serving loops, harnesses that forbid hipMalloc): the variant of
// _mlir_ciface_mdrtCudaFFTForward3DF32 whose only workspace, the three
// axes' twiddle tables, lives in a file-scope device array instead of a
// plan the library allocates.
//
// `scratch` is one array per GPU, held until the process ends.  Host code
// never passes its name to a kernel: the address the device gives for it
// (hipGetSymbolAddress, taken once per device with the device current) is
// what the kernels receive.  Every call on a GPU rewrites the tables it
// needs before its transform, so calls that overlap on one GPU (on
// different streams) race on the array; calls in one stream's order do not.
namespace mdhip {

// Twiddles of the three axes, K1 + K2 + K3 complex f32 numbers at most.
constexpr int kNoAllocAxisMax = 1024;
constexpr size_t kNoAllocScratchBytes = 3 * (size_t)kNoAllocAxisMax * sizeof(cpx<float>);

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

// exp(-2 pi i m / K) for m < K, evaluated in f64 (sincospi of -2m/K, whose
// argument is exact for K a power of two and rounded once otherwise) and
// rounded once to f32, as the plans' host tables are.
__global__ void fft_twiddles_f32(cpx<float> *__restrict__ tw, int K) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= K) return;
  double s, c;
  sincospi(-2.0 * (double)m / (double)K, &s, &c);
  tw[m] = {(float)c, (float)s};
}

namespace {
constexpr int kMaxDevices = 64;
void *g_scratch_at[kMaxDevices];  // the device address of `scratch`, per device
std::mutex g_scratch_mutex;

// The address of this device's `scratch`, looked up once with the device
// current.
void *scratch_address() {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) throw HipError("device ordinal beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_at[dev]) {
    void *p = nullptr;
    const hipError_t st = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (st != hipSuccess || !p)
      throw HipError(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(st));
    g_scratch_at[dev] = p;
  }
  return g_scratch_at[dev];
}

// Radices of K (8, 4, 2, then 3, 5, 7, the order of the plans' own
// factorisation, which the kernels' passes follow) into an FftAxis; false
// when K has another prime factor or too many radices.
bool factor_axis(int K, FftAxis &ax) {
  ax.K = K;
  ax.nrad = 0;
  int rest = K;
  const int order[6] = {8, 4, 2, 3, 5, 7};
  for (int r : order)
    while (rest % r == 0) {
      if (ax.nrad == kMaxRadices) return false;
      ax.rad[ax.nrad++] = r;
      rest /= r;
    }
  return rest == 1;
}
}  // namespace

}  // namespace mdhip

extern "C" int mdrtHipFFTForward3DF32NoAlloc(const float *real, void *complex, int64_t k1,
                                              int64_t k2, int64_t k3, void *stream) {
  using namespace mdhip;
  return guarded([&] {
    const int64_t K[3] = {k1, k2, k3};
    for (int a = 0; a < 3; ++a)
      if (K[a] < 2 || K[a] > kNoAllocAxisMax)
        throw HipError("no-alloc transform: grid edges must lie in [2, " +
                       std::to_string(kNoAllocAxisMax) + "]");
    if (k3 % 2) throw HipError("no-alloc transform: the last grid edge must be even");
    if (!real || !complex) throw HipError("no-alloc transform: null grid");
    FftAxis ax[3];
    for (int a = 0; a < 3; ++a)
      if (!factor_axis((int)K[a], ax[a]))
        throw HipError("no-alloc transform: grid edges need the factors 2, 3, 5 and 7 only");
    hipStream_t s = (hipStream_t)stream;
    cpx<float> *tw = static_cast<cpx<float> *>(scratch_address());
    size_t at = 0;
    for (int a = 0; a < 3; ++a) {
      fft_twiddles_f32<<<blocks_for((size_t)K[a], 256), 256, 0, s>>>(tw + at, (int)K[a]);
      ax[a].tw = tw + at;
      at += (size_t)K[a];
    }
    constexpr int TL = Fft3D<float>::TL;
    auto lds = [](int64_t k) { return (size_t)2 * TL * (size_t)k * sizeof(cpx<float>); };
    const int H = (int)(k3 / 2 + 1);
    const int nrows = (int)(k1 * k2);
    cpx<float> *out = static_cast<cpx<float> *>(complex);
    fft_rows_r2c<float, false><<<(nrows + 2 * TL - 1) / (2 * TL), 256, lds(k3), s>>>(
        real, out, nrows, ax[2]);
    const int tiles = (H + TL - 1) / TL;
    fft_cols_y<float><<<(int)k1 * tiles, 256, lds(k2), s>>>(out, (int)k2, H, ax[1], -1);
    ConvTables ct{};
    ct.K3 = (int)k3;
    fft_x_conv<float><<<(int)k2 * tiles, 256, lds(k1), s>>>(out, (int)k1, (int)k2, H, ax[0],
                                                            ct, -1, nullptr);
    HIPCHK(hipGetLastError());
  });
}
