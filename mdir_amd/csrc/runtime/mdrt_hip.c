/* mdrt_hip.c — the runtime ABI of the reference's runtime/mdrt_cuda.c on HIP
 * for MI355X (gfx950).
 *
 * The compiled program of the reference issues all of its work through
 * these entry points: module load and launch (mgpu*), one shared stream
 * with two side streams joined by events, a caching device pool, flags in
 * mapped host memory, and the FFT of the PME grid.  The same names and
 * signatures are exported here; modules are gfx950 code objects, the FFT
 * entry points run this library's own LDS-resident FFT kernels (no vendor
 * FFT), and everything else maps onto the HIP runtime.
 *
 * Errors: a failed HIP call prints the call and the runtime's message and
 * aborts, after the failure hook (mdrt_cuda.c:44-71); failures a run can
 * recover from go to the stop handler if one is registered, else exit(1)
 * (mdrt_cuda.c:2559-2578).
 */
#include <hip/hip_runtime_api.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../../include/mdrt_hip.h"

/* Provided by engine.hip (the FFT plans and kernels). */
int mdhip_fft_execute(int direction, int single, void *in, void *out,
                      int64_t k1, int64_t k2, int64_t k3, const float *tables,
                      void *stream);
int mdhip_fft_retain(int64_t k1, int64_t k2, int64_t k3, int single);
void mdhip_fft_release(int handle);
int64_t mdhip_fft_plan_count(void);
int64_t mdhip_fft_plan_bytes(void);
void mdhip_fft_trim(void);

static void (*failureHook)(void) = NULL;
static void (*stopHandler)(const char *) = NULL;

static void check(hipError_t result, const char *what) {
  if (result == hipSuccess) return;
  fprintf(stderr, "mdrt: %s failed: %s\n", what, hipGetErrorString(result));
  fflush(stderr);
  if (failureHook) failureHook();
  abort();
}

static void stopWith(const char *message) {
  if (stopHandler) {
    stopHandler(message);
    return;
  }
  fprintf(stderr, "mdrt: %s\n", message);
  exit(1);
}

void mdrtSetFailureHook(void (*hook)(void)) { failureHook = hook; }
void mdrtCudaSetStopHandler(void (*handler)(const char *)) { stopHandler = handler; }

/* ---- device and streams ------------------------------------------------ */

static int32_t defaultDevice = 0;
static int entered = 0;
static hipStream_t sharedStream = NULL;
static int borrowedDepth = 0;
enum { SIDE_LANES = 2 };
static hipStream_t sideStreams[SIDE_LANES];
static hipEvent_t joinEvents[SIDE_LANES];
static int joinPending[SIDE_LANES];
static hipEvent_t forkEvent = NULL, scopeEvent = NULL;
static int onSide = 0, sideLane = -1;

static void enter(void) {
  if (!entered) {
    const char *chosen = getenv("MDRT_DEVICE");
    if (chosen) defaultDevice = (int32_t)strtol(chosen, NULL, 10);
    entered = 1;
  }
  check(hipSetDevice(defaultDevice), "hipSetDevice");
}

void mgpuSetDefaultDevice(int32_t device) {
  defaultDevice = device;
  entered = 1;
  check(hipSetDevice(device), "hipSetDevice");
}

intptr_t mdrtCooperativeGridSize(intptr_t requested) {
  static int units = -1;
  if (units < 0) {
    enter();
    int supported = 0;
    check(hipDeviceGetAttribute(&supported, hipDeviceAttributeCooperativeLaunch,
                                defaultDevice),
          "hipDeviceGetAttribute");
    units = 1;
    if (supported)
      check(hipDeviceGetAttribute(&units, hipDeviceAttributeMultiprocessorCount,
                                  defaultDevice),
            "hipDeviceGetAttribute");
  }
  return requested < units ? requested : units;
}

mgpuStream mgpuStreamCreate(void) {
  enter();
  if (!sharedStream && !borrowedDepth)
    check(hipStreamCreateWithFlags(&sharedStream, hipStreamNonBlocking),
          "hipStreamCreateWithFlags");
  return onSide ? (mgpuStream)sideStreams[sideLane] : (mgpuStream)sharedStream;
}

void mgpuStreamDestroy(mgpuStream stream) { (void)stream; }

void mgpuStreamSynchronize(mgpuStream stream) {
  check(hipStreamSynchronize((hipStream_t)stream), "hipStreamSynchronize");
}

void mgpuStreamWaitEvent(mgpuStream stream, mgpuEvent event) {
  check(hipStreamWaitEvent((hipStream_t)stream, (hipEvent_t)event, 0),
        "hipStreamWaitEvent");
}

mgpuEvent mgpuEventCreate(void) {
  enter();
  hipEvent_t event = NULL;
  check(hipEventCreateWithFlags(&event, hipEventDisableTiming), "hipEventCreate");
  return (mgpuEvent)event;
}

void mgpuEventDestroy(mgpuEvent event) {
  check(hipEventDestroy((hipEvent_t)event), "hipEventDestroy");
}

void mgpuEventSynchronize(mgpuEvent event) {
  check(hipEventSynchronize((hipEvent_t)event), "hipEventSynchronize");
}

void mgpuEventRecord(mgpuEvent event, mgpuStream stream) {
  check(hipEventRecord((hipEvent_t)event, (hipStream_t)stream), "hipEventRecord");
}

static void sideBegin(int lane) {
  mgpuStreamCreate();
  if (!forkEvent)
    check(hipEventCreateWithFlags(&forkEvent, hipEventDisableTiming), "hipEventCreate");
  if (!sideStreams[lane]) {
    check(hipStreamCreateWithFlags(&sideStreams[lane], hipStreamNonBlocking),
          "hipStreamCreateWithFlags");
    check(hipEventCreateWithFlags(&joinEvents[lane], hipEventDisableTiming),
          "hipEventCreate");
  }
  check(hipEventRecord(forkEvent, sharedStream), "hipEventRecord");
  check(hipStreamWaitEvent(sideStreams[lane], forkEvent, 0), "hipStreamWaitEvent");
  sideLane = lane;
  onSide = 1;
}

static void sideEnd(int lane) {
  check(hipEventRecord(joinEvents[lane], sideStreams[lane]), "hipEventRecord");
  sideLane = -1;
  onSide = 0;
  joinPending[lane] = 1;
}

void mdrtSideBegin(void) { sideBegin(0); }
void mdrtSideEnd(void) { sideEnd(0); }
void mdrtSide2Begin(void) { sideBegin(1); }
void mdrtSide2End(void) { sideEnd(1); }

void mdrtSideJoin(void) {
  for (int lane = 0; lane < SIDE_LANES; ++lane) {
    if (!joinPending[lane]) continue;
    check(hipStreamWaitEvent(sharedStream, joinEvents[lane], 0), "hipStreamWaitEvent");
    joinPending[lane] = 0;
  }
}

static void changeStream(hipStream_t next) {
  mdrtSideJoin();
  if (next != sharedStream) {
    if (!scopeEvent)
      check(hipEventCreateWithFlags(&scopeEvent, hipEventDisableTiming), "hipEventCreate");
    check(hipEventRecord(scopeEvent, sharedStream), "hipEventRecord");
    check(hipStreamWaitEvent(next, scopeEvent, 0), "hipStreamWaitEvent");
    sharedStream = next;
  }
}

void *mdrtUseStream(void *stream) {
  mgpuStreamCreate();
  hipStream_t previous = sharedStream;
  changeStream((hipStream_t)stream);
  ++borrowedDepth;
  return (void *)previous;
}

void mdrtRestoreStream(void *previous) {
  changeStream((hipStream_t)previous);
  --borrowedDepth;
}

static void finish(void) {
  if (sharedStream) check(hipStreamSynchronize(sharedStream), "hipStreamSynchronize");
}

/* ---- modules and launches --------------------------------------------- */

mgpuModule mgpuModuleLoad(void *data, size_t size) {
  (void)size;
  enter();
  hipModule_t module = NULL;
  hipError_t result = hipModuleLoadData(&module, data);
  if (result == hipErrorNoBinaryForGpu || result == hipErrorInvalidImage) {
    fprintf(stderr, "mdrt: the code object was not built for this device "
                    "(build it with --offload-arch=gfx950)\n");
    fflush(stderr);
  }
  check(result, "hipModuleLoadData");
  if (getenv("MDRT_TRACE")) fprintf(stderr, "TRACE module %p\n", (void *)module);
  return (mgpuModule)module;
}

mgpuModule mgpuModuleLoadJIT(void *data, int optLevel, size_t size) {
  (void)optLevel;
  return mgpuModuleLoad(data, size);
}

void mgpuModuleUnload(mgpuModule module) {
  if (sharedStream) hipStreamSynchronize(sharedStream);
  hipError_t result = hipModuleUnload((hipModule_t)module);
  if (result != hipErrorDeinitialized && result != hipErrorContextIsDestroyed)
    check(result, "hipModuleUnload");
}

mgpuFunction mgpuModuleGetFunction(mgpuModule module, const char *name) {
  hipFunction_t function = NULL;
  check(hipModuleGetFunction(&function, (hipModule_t)module, name),
        "hipModuleGetFunction");
  if (getenv("MDRT_TRACE"))
    fprintf(stderr, "TRACE function %p module %p %s\n", (void *)function,
            (void *)module, name);
  return (mgpuFunction)function;
}

void mgpuLaunchKernel(mgpuFunction function, intptr_t gridX, intptr_t gridY,
                      intptr_t gridZ, intptr_t blockX, intptr_t blockY,
                      intptr_t blockZ, int32_t smem, mgpuStream stream,
                      void **params, void **extra, size_t paramsCount) {
  (void)paramsCount;
  check(hipModuleLaunchKernel((hipFunction_t)function, (unsigned)gridX,
                              (unsigned)gridY, (unsigned)gridZ, (unsigned)blockX,
                              (unsigned)blockY, (unsigned)blockZ, (unsigned)smem,
                              (hipStream_t)stream, params, extra),
        "hipModuleLaunchKernel");
}

void mgpuLaunchKernelCooperative(mgpuFunction function, intptr_t gridX,
                                 intptr_t gridY, intptr_t gridZ,
                                 intptr_t clusterX, intptr_t clusterY,
                                 intptr_t clusterZ, intptr_t blockX,
                                 intptr_t blockY, intptr_t blockZ,
                                 int32_t smem, mgpuStream stream,
                                 void **params, void **extra) {
  (void)extra;
  if (clusterX != 1 || clusterY != 1 || clusterZ != 1) {
    stopWith("thread-block clusters are not available on this device");
    return;
  }
  check(hipModuleLaunchCooperativeKernel(
            (hipFunction_t)function, (unsigned)gridX, (unsigned)gridY,
            (unsigned)gridZ, (unsigned)blockX, (unsigned)blockY,
            (unsigned)blockZ, (unsigned)smem, (hipStream_t)stream, params),
        "hipModuleLaunchCooperativeKernel");
}

/* ---- the device pool ---------------------------------------------------
   Everything runs on one stream, so a freed block can serve the next
   request at once (mdrt_cuda.c:1788-1840).  A request takes the idle block
   of exactly its size if there is one, else the smallest idle block that
   holds it and is at most twice as large, else a new allocation. */

struct Block {
  void *pointer;
  uint64_t size;
};
static struct Block *liveBlocks = NULL, *idleBlocks = NULL;
static int64_t numLive = 0, roomLive = 0, numIdle = 0, roomIdle = 0;
static uint64_t liveBytes = 0, idleBytes = 0;
static int64_t numDriverBlocks = 0;

static void push(struct Block **list, int64_t *count, int64_t *room, struct Block b) {
  if (*count == *room) {
    *room = *room ? 2 * *room : 64;
    *list = (struct Block *)realloc(*list, (size_t)*room * sizeof(struct Block));
    if (!*list) {
      fprintf(stderr, "mdrt: out of host memory\n");
      abort();
    }
  }
  (*list)[(*count)++] = b;
}

void *mgpuMemAlloc(uint64_t size, mgpuStream stream, bool isHostShared) {
  (void)stream;
  enter();
  if (size == 0) return NULL;
  if (isHostShared) {
    void *p = NULL;
    check(hipMallocManaged(&p, size, hipMemAttachGlobal), "hipMallocManaged");
    return p;
  }
  int64_t best = -1;
  for (int64_t i = numIdle; i-- > 0;)
    if (idleBlocks[i].size == size) {
      best = i;
      break;
    }
  if (best < 0)
    for (int64_t i = numIdle; i-- > 0;) {
      uint64_t s = idleBlocks[i].size;
      if (s < size || s > 2 * size) continue;
      if (best < 0 || s < idleBlocks[best].size) best = i;
    }
  struct Block b;
  if (best >= 0) {
    b = idleBlocks[best];
    idleBlocks[best] = idleBlocks[--numIdle];
    idleBytes -= b.size;
  } else {
    b.size = size;
    b.pointer = NULL;
    hipError_t r = hipMalloc(&b.pointer, size);
    if (r == hipErrorOutOfMemory) {
      finish();
      mdrtDeviceTrim();
      r = hipMalloc(&b.pointer, size);
    }
    check(r, "hipMalloc");
    ++numDriverBlocks;
  }
  push(&liveBlocks, &numLive, &roomLive, b);
  liveBytes += b.size;
  return b.pointer;
}

void mgpuMemFree(void *pointer, mgpuStream stream) {
  (void)stream;
  if (!pointer) return;
  for (int64_t i = numLive; i-- > 0;)
    if (liveBlocks[i].pointer == pointer) {
      struct Block b = liveBlocks[i];
      liveBlocks[i] = liveBlocks[--numLive];
      liveBytes -= b.size;
      push(&idleBlocks, &numIdle, &roomIdle, b);
      idleBytes += b.size;
      return;
    }
  /* not from the pool: managed memory */
  check(hipFree(pointer), "hipFree");
}

uint64_t mdrtDeviceTrim(void) {
  if (!entered) return 0;
  finish();
  uint64_t released = idleBytes;
  for (int64_t i = 0; i < numIdle; ++i) {
    check(hipFree(idleBlocks[i].pointer), "hipFree");
    --numDriverBlocks;
  }
  numIdle = 0;
  idleBytes = 0;
  mdhip_fft_trim();
  return released;
}

void mdrtDeviceMemoryCounts(int64_t counts[8]) {
  counts[0] = numLive;
  counts[1] = (int64_t)liveBytes;
  counts[2] = numIdle;
  counts[3] = (int64_t)idleBytes;
  counts[4] = numDriverBlocks;
  counts[5] = mdhip_fft_plan_count();
  counts[6] = mdhip_fft_plan_bytes();
  size_t free = 0, total = 0;
  counts[7] = entered && hipMemGetInfo(&free, &total) == hipSuccess ? (int64_t)free : -1;
}

static int isOnDevice(const void *pointer) {
  hipPointerAttribute_t attributes;
  if (hipPointerGetAttributes(&attributes, pointer) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return attributes.type == hipMemoryTypeDevice;
}

void mgpuMemcpy(void *destination, void *source, size_t size, mgpuStream stream) {
  if (size == 0) return;
  check(hipMemcpyAsync(destination, source, size, hipMemcpyDefault,
                       (hipStream_t)stream),
        "hipMemcpyAsync");
  /* The host reads what it copied from the device and may change what it
     copied to it: a copy that touches the host is done on return. */
  if (!isOnDevice(destination) || !isOnDevice(source))
    check(hipStreamSynchronize((hipStream_t)stream), "hipStreamSynchronize");
}

void mgpuMemset32(void *destination, unsigned int value, size_t count,
                  mgpuStream stream) {
  check(hipMemsetD32Async((hipDeviceptr_t)destination, (int)value, count,
                          (hipStream_t)stream),
        "hipMemsetD32Async");
}

void mgpuMemset16(void *destination, unsigned short value, size_t count,
                  mgpuStream stream) {
  check(hipMemsetD16Async((hipDeviceptr_t)destination, value, count,
                          (hipStream_t)stream),
        "hipMemsetD16Async");
}

/* ---- flags and host results ------------------------------------------- */

static int64_t hostReadBytes = 0, hostWaits = 0, hostFlagWaits = 0;

void mdrtCountHostRead(size_t bytes, int waits, int flags) {
  hostReadBytes += (int64_t)bytes;
  hostWaits += waits;
  hostFlagWaits += flags;
}

enum { MAPPED_FLAGS = 1024, FLAG_SLOTS = 64 };
static volatile int32_t *mappedFlags = NULL;
static char *mappedDevice = NULL;
static int32_t *flagSlots = NULL;
static volatile int32_t *flagSources[FLAG_SLOTS];
static hipEvent_t flagEvents[FLAG_SLOTS];
static int64_t nextFlagSlot = 0;

void _mlir_ciface_mdrtFlagMemory(mdrtBuffer1 *result, int64_t place) {
  mgpuStreamCreate();
  if (!mappedFlags) {
    void *host = NULL;
    check(hipHostMalloc(&host, MAPPED_FLAGS * sizeof(int32_t), hipHostMallocMapped),
          "hipHostMalloc");
    memset(host, 0, MAPPED_FLAGS * sizeof(int32_t));
    void *device = NULL;
    check(hipHostGetDevicePointer(&device, host, 0), "hipHostGetDevicePointer");
    mappedFlags = (volatile int32_t *)host;
    mappedDevice = (char *)device;
  }
  if (place < 0 || place >= MAPPED_FLAGS) {
    fprintf(stderr, "mdrt: the module takes more than %d flags\n", MAPPED_FLAGS);
    exit(1);
  }
  mappedFlags[place] = 0;
  void *device = mappedDevice + place * (int64_t)sizeof(int32_t);
  result->allocated = device;
  result->aligned = device;
  result->offset = 0;
  result->size = 1;
  result->stride = 1;
}

int64_t mdrtFlagStart(void *device) {
  mgpuStreamCreate();
  if (!flagSlots) {
    check(hipHostMalloc((void **)&flagSlots, FLAG_SLOTS * sizeof(int32_t), 0),
          "hipHostMalloc");
    for (int i = 0; i < FLAG_SLOTS; ++i)
      check(hipEventCreateWithFlags(&flagEvents[i], hipEventDisableTiming),
            "hipEventCreate");
  }
  int64_t slot = nextFlagSlot++ % FLAG_SLOTS;
  char *address = (char *)device;
  if (mappedFlags && address >= mappedDevice &&
      address < mappedDevice + MAPPED_FLAGS * sizeof(int32_t)) {
    flagSources[slot] = mappedFlags + (address - mappedDevice) / sizeof(int32_t);
  } else {
    flagSources[slot] = NULL;
    mdrtCountHostRead(sizeof(int32_t), 0, 0);
    check(hipMemcpyAsync(&flagSlots[slot], device, sizeof(int32_t),
                         hipMemcpyDeviceToHost, sharedStream),
          "hipMemcpyAsync");
  }
  check(hipEventRecord(flagEvents[slot], sharedStream), "hipEventRecord");
  return slot;
}

int32_t mdrtFlagFinish(int64_t slot) {
  mdrtCountHostRead(0, 0, 1);
  check(hipEventSynchronize(flagEvents[slot]), "hipEventSynchronize");
  volatile int32_t *source = flagSources[slot];
  if (!source) return flagSlots[slot];
  int32_t value = *source;
  *source = 0;
  return value;
}

int32_t mdrtFlagTake(void *device) {
  char *address = (char *)device;
  if (mappedFlags && address >= mappedDevice &&
      address < mappedDevice + MAPPED_FLAGS * sizeof(int32_t)) {
    volatile int32_t *source = mappedFlags + (address - mappedDevice) / sizeof(int32_t);
    int32_t value = *source;
    *source = 0;
    return value;
  }
  return mdrtFlagFinish(mdrtFlagStart(device));
}

enum { HOST_RESULT_PLACES = 1024 };
static void *hostResults[HOST_RESULT_PLACES];
static int64_t hostResultBytes[HOST_RESULT_PLACES];

void _mlir_ciface_mdrtHostResults(mdrtBuffer1 *result, int64_t place, int64_t count) {
  mgpuStreamCreate();
  if (place < 0 || place >= HOST_RESULT_PLACES) {
    fprintf(stderr, "mdrt: the module takes more than %d buffers of results\n",
            HOST_RESULT_PLACES);
    exit(1);
  }
  int64_t bytes = count * (int64_t)sizeof(double);
  if (hostResultBytes[place] < bytes) {
    if (hostResults[place]) {
      finish();
      check(hipHostFree(hostResults[place]), "hipHostFree");
    }
    check(hipHostMalloc(&hostResults[place], (size_t)bytes, 0), "hipHostMalloc");
    hostResultBytes[place] = bytes;
  }
  result->allocated = hostResults[place];
  result->aligned = hostResults[place];
  result->offset = 0;
  result->size = count;
  result->stride = 1;
}

/* ---- FFT --------------------------------------------------------------- */

static void fft(int direction, int single, mdrtBuffer1 *in, mdrtBuffer1 *out,
                const float *tables, int64_t k1, int64_t k2, int64_t k3) {
  hipStream_t stream = (hipStream_t)mgpuStreamCreate();
  size_t elem = single ? sizeof(float) : sizeof(double);
  void *src = (char *)in->aligned + in->offset * (int64_t)elem;
  void *dst = (char *)out->aligned + out->offset * (int64_t)elem;
  if (mdhip_fft_execute(direction, single, src, dst, k1, k2, k3, tables, stream) != 0)
    stopWith("the FFT of the PME grid failed (grid sizes need factors 2, 3, 5, 7 "
             "and an even k3)");
}

void _mlir_ciface_mdrtCudaFFTForward3D(mdrtBuffer1 *real, mdrtBuffer1 *complex,
                                       int64_t k1, int64_t k2, int64_t k3) {
  fft(-1, 0, real, complex, NULL, k1, k2, k3);
}

void _mlir_ciface_mdrtCudaFFTBackward3D(mdrtBuffer1 *complex, mdrtBuffer1 *real,
                                        int64_t k1, int64_t k2, int64_t k3) {
  fft(+1, 0, complex, real, NULL, k1, k2, k3);
}

void _mlir_ciface_mdrtCudaFFTForward3DF32(mdrtBuffer1 *real, mdrtBuffer1 *complex,
                                          int64_t k1, int64_t k2, int64_t k3) {
  fft(-1, 1, real, complex, NULL, k1, k2, k3);
}

void _mlir_ciface_mdrtCudaFFTBackward3DF32(mdrtBuffer1 *complex, mdrtBuffer1 *real,
                                           int64_t k1, int64_t k2, int64_t k3) {
  fft(+1, 1, complex, real, NULL, k1, k2, k3);
}

void _mlir_ciface_mdrtCudaFFTGaussianBackwardF32(mdrtBuffer1 *complex,
                                                 mdrtBuffer1 *real,
                                                 mdrtBuffer1 *tables, int64_t k1,
                                                 int64_t k2, int64_t k3) {
  const float *t = (const float *)tables->aligned + tables->offset;
  fft(+1, 1, complex, real, t, k1, k2, k3);
}

int mdrtRetainFFTPlan(int64_t k1, int64_t k2, int64_t k3, int forward, int single,
                      const float *tables) {
  (void)forward;
  (void)tables;
  enter();
  return mdhip_fft_retain(k1, k2, k3, single);
}

void mdrtReleaseFFTPlan(int handle) { mdhip_fft_release(handle); }
