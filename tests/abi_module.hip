// Kernels of a small gfx950 code object that the ABI tests load and launch
// through mgpuModuleLoad / mgpuModuleGetFunction / mgpuLaunchKernel, the way
// a compiled program drives the runtime.  Built by mdir_amd/build.py into
// mdir_amd/lib/abi_module.co (hipcc --genco).
#include <hip/hip_runtime.h>

// y[i] = a x[i] + y[i]
extern "C" __global__ void abi_axpy(float *y, const float *x, float a, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = a * x[i] + y[i];
}

// p[i] = value after `rounds` dependent integer steps per element (a kernel
// that takes a while, so that work issued after it without a wait would
// overtake it); `mask` is 0 at every call and keeps the steps alive.
extern "C" __global__ void abi_fill_slow(int *p, int value, int n, int rounds, unsigned mask) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  unsigned v = (unsigned)i;
  for (int r = 0; r < rounds; ++r) v = v * 1664525u + 1013904223u;
  p[i] = value + (int)(v & mask);
}

// out[i] = k a[i]
extern "C" __global__ void abi_scale(int *out, const int *a, int k, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = k * a[i];
}

// out[i] = a[i] + b[i]
extern "C" __global__ void abi_add(int *out, const int *a, const int *b, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = a[i] + b[i];
}

// *flag = value (a flag word in mapped host memory or in device memory)
extern "C" __global__ void abi_set_flag(int *flag, int value) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *flag = value;
}
