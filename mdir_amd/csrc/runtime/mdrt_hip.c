/* mdrt_hip.c — the device runtime that a program compiled by mdir calls,
 * for one MI355X (gfx950) under HIP.
 *
 * The entry points, their names and their signatures are those of the
 * reference's runtime/mdrt_cuda.c (each is cited in include/mdrt_hip.h);
 * the implementation is this file's own:
 *
 *   - An ordering context (`lanes`) decides which HIP stream each piece of
 *     work goes to: the trunk (the runtime's own stream, or a stream the
 *     caller lends with mdrtUseStream, kept on a checked stack), and two
 *     branches that fork from the trunk and join it again.
 *   - A device pool keeps freed blocks in a size-ordered free list and the
 *     blocks in use in an address-ordered table, so that a free and the
 *     question "is this address device memory" are binary searches.
 *   - One page-locked, device-mapped page holds the flags of the module's
 *     places and the staging words of flag reads from device memory.
 *   - Module loads take gfx950 code objects; the FFT entry points run this
 *     library's own LDS-resident transforms.
 *
 * Failure policy (the reference's, mdrt_cuda.c:44-71 and :2559-2578): a
 * failed HIP call or a misuse of the ABI prints what failed and the last
 * kernel launched, calls the failure hook and aborts; a failure a run can
 * recover from goes to the stop handler when one is registered, else the
 * process exits with status 1.
 */
#define _POSIX_C_SOURCE 200809L /* nanosleep */
#include <hip/hip_runtime_api.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

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
/* A one-thread kernel that stores `serial` into the mapped host word at
   device address `word`, with a system-scope release (a vector store). */
int mdhip_stage_mark(void *word, uint32_t serial, void *stream);

/* ======================================================================
   Failures
   ====================================================================== */

static struct {
  void (*hook)(void);
  void (*stop)(const char *);
  const char *last_kernel; /* name of the last function launched */
} failure;

static void fail_now(void) {
  fflush(stderr);
  if (failure.hook) failure.hook();
  abort();
}

/* A misuse of the ABI: the message, then the failure path. */
static void misuse(const char *format, ...) {
  va_list args;
  va_start(args, format);
  fputs("mdrt: ", stderr);
  vfprintf(stderr, format, args);
  fputc('\n', stderr);
  va_end(args);
  fail_now();
}

static void hip_ok(hipError_t status, const char *call) {
  if (status == hipSuccess) return;
  fprintf(stderr, "mdrt: %s: %s", call, hipGetErrorString(status));
  if (failure.last_kernel) fprintf(stderr, " (last kernel launched: %s)", failure.last_kernel);
  fputc('\n', stderr);
  fail_now();
}

static void stop_run(const char *message) {
  if (failure.stop) {
    failure.stop(message);
    return;
  }
  fprintf(stderr, "mdrt: %s\n", message);
  fflush(stderr);
  exit(1);
}

void mdrtSetFailureHook(void (*hook)(void)) { failure.hook = hook; }
void mdrtCudaSetStopHandler(void (*handler)(const char *)) { failure.stop = handler; }

/* ======================================================================
   Device
   ====================================================================== */

static struct {
  int ordinal;
  int known; /* ordinal read from MDRT_DEVICE or set by the caller */
  int used;  /* some HIP work was issued */
} device;

static void use_device(void) {
  if (!device.known) {
    const char *text = getenv("MDRT_DEVICE");
    device.ordinal = text ? (int)strtol(text, NULL, 10) : 0;
    device.known = 1;
  }
  device.used = 1;
  hip_ok(hipSetDevice(device.ordinal), "hipSetDevice");
}

void mgpuSetDefaultDevice(int32_t ordinal) {
  device.ordinal = ordinal;
  device.known = 1;
  device.used = 1;
  hip_ok(hipSetDevice(ordinal), "hipSetDevice");
}

intptr_t mdrtCooperativeGridSize(intptr_t requested) {
  static intptr_t resident = 0; /* workgroups that can run at once, 1 if none */
  if (!resident) {
    use_device();
    int cooperative = 0, units = 1;
    hip_ok(hipDeviceGetAttribute(&cooperative, hipDeviceAttributeCooperativeLaunch,
                                 device.ordinal), "hipDeviceGetAttribute");
    if (cooperative)
      hip_ok(hipDeviceGetAttribute(&units, hipDeviceAttributeMultiprocessorCount,
                                   device.ordinal), "hipDeviceGetAttribute");
    resident = units > 0 ? units : 1;
  }
  return requested < resident ? requested : resident;
}

/* ======================================================================
   Ordering: the trunk, borrowed trunks, branches
   ====================================================================== */

enum { BRANCHES = 2, LEND_DEPTH = 32 };

typedef struct {
  hipStream_t stream;
  hipEvent_t opened; /* the trunk's position when the branch forked */
  hipEvent_t closed; /* the end of the branch's work */
  int unjoined;      /* closed and not yet waited for by the trunk */
} Branch;

static struct {
  int made;                      /* the runtime's own stream exists */
  hipStream_t own;               /* the runtime's own stream */
  hipStream_t trunk;             /* where ordered work goes now */
  hipStream_t lent[LEND_DEPTH];  /* trunks replaced by mdrtUseStream */
  int depth;
  Branch branch[BRANCHES];
  int open;                      /* the branch work goes to, -1: the trunk */
  hipEvent_t handover;           /* orders a change of trunk */
} lanes = {.open = -1};

static hipEvent_t new_marker(void) {
  hipEvent_t e = NULL;
  hip_ok(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreateWithFlags");
  return e;
}

/* `later` runs what is issued to it next only after what `earlier` holds
   now; `marker` is recorded on `earlier` (an event wait sees the record
   made before it, so one marker serves any number of serial uses). */
static void order_after(hipStream_t later, hipStream_t earlier, hipEvent_t marker) {
  hip_ok(hipEventRecord(marker, earlier), "hipEventRecord");
  hip_ok(hipStreamWaitEvent(later, marker, 0), "hipStreamWaitEvent");
}

static void ensure_trunk(void) {
  use_device();
  if (!lanes.made && lanes.depth == 0) {
    hip_ok(hipStreamCreateWithFlags(&lanes.own, hipStreamNonBlocking),
           "hipStreamCreateWithFlags");
    lanes.trunk = lanes.own;
    lanes.made = 1;
  }
}

mgpuStream mgpuStreamCreate(void) {
  ensure_trunk();
  return (mgpuStream)(lanes.open >= 0 ? lanes.branch[lanes.open].stream : lanes.trunk);
}

void mgpuStreamDestroy(mgpuStream stream) {
  (void)stream; /* the runtime's streams live as long as the process */
}

static void branch_open(int which) {
  ensure_trunk();
  if (lanes.open >= 0)
    misuse("side stream %d begins while side stream %d is open", which + 1, lanes.open + 1);
  Branch *b = &lanes.branch[which];
  if (!b->stream) {
    hip_ok(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking),
           "hipStreamCreateWithFlags");
    b->opened = new_marker();
    b->closed = new_marker();
  }
  order_after(b->stream, lanes.trunk, b->opened);
  lanes.open = which;
}

static void branch_close(int which) {
  if (lanes.open != which)
    misuse("side stream %d ends but %s", which + 1,
           lanes.open < 0 ? "no side stream is open" : "another side stream is open");
  Branch *b = &lanes.branch[which];
  hip_ok(hipEventRecord(b->closed, b->stream), "hipEventRecord");
  b->unjoined = 1;
  lanes.open = -1;
}

void mdrtSideBegin(void) { branch_open(0); }
void mdrtSideEnd(void) { branch_close(0); }
void mdrtSide2Begin(void) { branch_open(1); }
void mdrtSide2End(void) { branch_close(1); }

void mdrtSideJoin(void) {
  for (int i = 0; i < BRANCHES; ++i) {
    Branch *b = &lanes.branch[i];
    if (!b->unjoined) continue;
    hip_ok(hipStreamWaitEvent(lanes.trunk, b->closed, 0), "hipStreamWaitEvent");
    b->unjoined = 0;
  }
}

/* Moves the trunk to `next` without a host wait: the branches join the old
   trunk first, and `next` starts after everything the old trunk holds. */
static void move_trunk(hipStream_t next) {
  if (lanes.open >= 0) misuse("the stream changes while side stream %d is open", lanes.open + 1);
  mdrtSideJoin();
  if (next == lanes.trunk) return;
  if (!lanes.handover) lanes.handover = new_marker();
  order_after(next, lanes.trunk, lanes.handover);
  lanes.trunk = next;
}

void *mdrtUseStream(void *stream) {
  ensure_trunk();
  if (lanes.depth == LEND_DEPTH) misuse("more than %d nested mdrtUseStream scopes", LEND_DEPTH);
  hipStream_t before = lanes.trunk;
  move_trunk((hipStream_t)stream);
  lanes.lent[lanes.depth++] = before;
  return (void *)before;
}

void mdrtRestoreStream(void *previous) {
  if (lanes.depth == 0) misuse("mdrtRestoreStream without a matching mdrtUseStream");
  hipStream_t expected = lanes.lent[lanes.depth - 1];
  if ((hipStream_t)previous != expected)
    misuse("mdrtRestoreStream(%p) does not restore the stream its scope replaced (%p)",
           previous, (void *)expected);
  move_trunk(expected);
  --lanes.depth;
}

/* Waits for everything issued on the runtime's streams. */
static void drain(void) {
  if (!device.used) return;
  for (int i = 0; i < BRANCHES; ++i)
    if (lanes.branch[i].stream)
      hip_ok(hipStreamSynchronize(lanes.branch[i].stream), "hipStreamSynchronize");
  if (lanes.made || lanes.depth)
    hip_ok(hipStreamSynchronize(lanes.trunk), "hipStreamSynchronize");
}

void mgpuStreamSynchronize(mgpuStream stream) {
  hip_ok(hipStreamSynchronize((hipStream_t)stream), "hipStreamSynchronize");
}

void mgpuStreamWaitEvent(mgpuStream stream, mgpuEvent event) {
  hip_ok(hipStreamWaitEvent((hipStream_t)stream, (hipEvent_t)event, 0), "hipStreamWaitEvent");
}

mgpuEvent mgpuEventCreate(void) {
  use_device();
  return (mgpuEvent)new_marker();
}

void mgpuEventDestroy(mgpuEvent event) {
  hip_ok(hipEventDestroy((hipEvent_t)event), "hipEventDestroy");
}

void mgpuEventSynchronize(mgpuEvent event) {
  hip_ok(hipEventSynchronize((hipEvent_t)event), "hipEventSynchronize");
}

void mgpuEventRecord(mgpuEvent event, mgpuStream stream) {
  hip_ok(hipEventRecord((hipEvent_t)event, (hipStream_t)stream), "hipEventRecord");
}

/* ======================================================================
   Host reads (counted for the log)
   ====================================================================== */

static struct {
  int64_t bytes, waits, flag_waits;
} host_reads;

void mdrtCountHostRead(size_t bytes, int waits, int flags) {
  host_reads.bytes += (int64_t)bytes;
  host_reads.waits += waits;
  host_reads.flag_waits += flags;
}

/* ======================================================================
   Modules and launches
   ====================================================================== */

/* Names of the functions handed out, for failure messages. */
enum { NAMED_FUNCTIONS = 512 };
static struct {
  hipFunction_t function;
  hipModule_t module;
  char name[96];
} named[NAMED_FUNCTIONS];
static int num_named = 0;

static const char *name_of(hipFunction_t function) {
  for (int i = 0; i < num_named; ++i)
    if (named[i].function == function) return named[i].name;
  return NULL;
}

mgpuModule mgpuModuleLoad(void *data, size_t size) {
  (void)size; /* the code object's ELF header gives its extent */
  use_device();
  hipModule_t module = NULL;
  hipError_t status = hipModuleLoadData(&module, data);
  if (status == hipErrorNoBinaryForGpu || status == hipErrorInvalidImage ||
      status == hipErrorInvalidKernelFile)
    fprintf(stderr, "mdrt: the module is not a code object for this device "
                    "(a gfx950 code object is needed)\n");
  hip_ok(status, "hipModuleLoadData");
  if (getenv("MDRT_TRACE")) fprintf(stderr, "mdrt: loaded module %p\n", (void *)module);
  return (mgpuModule)module;
}

mgpuModule mgpuModuleLoadJIT(void *data, int optLevel, size_t size) {
  (void)optLevel; /* code objects are already machine code */
  return mgpuModuleLoad(data, size);
}

void mgpuModuleUnload(mgpuModule module) {
  drain();
  hipError_t status = hipModuleUnload((hipModule_t)module);
  /* at process exit the runtime may already be gone */
  if (status == hipErrorDeinitialized || status == hipErrorContextIsDestroyed) return;
  hip_ok(status, "hipModuleUnload");
  for (int i = 0; i < num_named;)
    if (named[i].module == (hipModule_t)module) named[i] = named[--num_named];
    else ++i;
}

mgpuFunction mgpuModuleGetFunction(mgpuModule module, const char *name) {
  hipFunction_t function = NULL;
  hipError_t status = hipModuleGetFunction(&function, (hipModule_t)module, name);
  if (status != hipSuccess) fprintf(stderr, "mdrt: no function '%s' in the module\n", name);
  hip_ok(status, "hipModuleGetFunction");
  if (!name_of(function)) {
    int at = num_named < NAMED_FUNCTIONS ? num_named++ : NAMED_FUNCTIONS - 1;
    named[at].function = function;
    named[at].module = (hipModule_t)module;
    snprintf(named[at].name, sizeof named[at].name, "%s", name);
  }
  if (getenv("MDRT_TRACE"))
    fprintf(stderr, "mdrt: function %s = %p in module %p\n", name, (void *)function,
            (void *)module);
  return (mgpuFunction)function;
}

static unsigned extent(intptr_t value, const char *what) {
  if (value < 1 || value > 0x7fffffff) misuse("launch %s of %ld", what, (long)value);
  return (unsigned)value;
}

void mgpuLaunchKernel(mgpuFunction function, intptr_t gridX, intptr_t gridY,
                      intptr_t gridZ, intptr_t blockX, intptr_t blockY,
                      intptr_t blockZ, int32_t smem, mgpuStream stream,
                      void **params, void **extra, size_t paramsCount) {
  (void)paramsCount; /* the code object's metadata gives the argument layout */
  failure.last_kernel = name_of((hipFunction_t)function);
  hip_ok(hipModuleLaunchKernel((hipFunction_t)function, extent(gridX, "grid x"),
                               extent(gridY, "grid y"), extent(gridZ, "grid z"),
                               extent(blockX, "block x"), extent(blockY, "block y"),
                               extent(blockZ, "block z"), (unsigned)smem,
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
  if ((clusterX | clusterY | clusterZ) != 1) {
    stop_run("a launch asks for thread-block clusters, which this device does not have");
    return;
  }
  failure.last_kernel = name_of((hipFunction_t)function);
  hip_ok(hipModuleLaunchCooperativeKernel(
             (hipFunction_t)function, extent(gridX, "grid x"), extent(gridY, "grid y"),
             extent(gridZ, "grid z"), extent(blockX, "block x"), extent(blockY, "block y"),
             extent(blockZ, "block z"), (unsigned)smem, (hipStream_t)stream, params),
         "hipModuleLaunchCooperativeKernel");
}

/* ======================================================================
   The device pool
   ======================================================================
   All pool traffic is ordered by the trunk, so a freed block serves the
   next request at once, without a wait (the reference's premise,
   mdrt_cuda.c:1788-1840 and :2129-2145): allocation and free are refused
   while a side stream is open or on a stream other than the trunk.

   used[]: blocks handed out, ordered by address (free and the address
   query are binary searches).  spare[]: freed blocks ordered by size; a
   request takes the smallest spare block that holds it and is at most
   twice its size, else the driver allocates.  Sizes are rounded up to
   256 bytes.  Managed (host-shared) allocations are not pooled. */

typedef struct {
  char *base;
  uint64_t size;
  int managed;
} Block;

typedef struct {
  Block *at;
  int64_t count, room;
  uint64_t bytes;
} BlockList;

static BlockList used, spare;
static int64_t driver_blocks = 0;

static void list_insert(BlockList *list, int64_t index, Block block) {
  if (list->count == list->room) {
    int64_t room = list->room ? 2 * list->room : 128;
    Block *grown = (Block *)realloc(list->at, (size_t)room * sizeof(Block));
    if (!grown) misuse("out of host memory for the pool's tables");
    list->at = grown;
    list->room = room;
  }
  memmove(list->at + index + 1, list->at + index,
          (size_t)(list->count - index) * sizeof(Block));
  list->at[index] = block;
  ++list->count;
  list->bytes += block.size;
}

static Block list_take(BlockList *list, int64_t index) {
  Block block = list->at[index];
  memmove(list->at + index, list->at + index + 1,
          (size_t)(list->count - index - 1) * sizeof(Block));
  --list->count;
  list->bytes -= block.size;
  return block;
}

/* First index whose base is above `address`. */
static int64_t used_upper(const char *address) {
  int64_t lo = 0, hi = used.count;
  while (lo < hi) {
    int64_t mid = (lo + hi) / 2;
    if (used.at[mid].base <= address) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

/* Index of the used block containing `address`, -1 if none. */
static int64_t used_find(const void *address) {
  const char *a = (const char *)address;
  int64_t i = used_upper(a) - 1;
  if (i >= 0 && a < used.at[i].base + used.at[i].size) return i;
  return -1;
}

/* First index of a spare block of at least `size` bytes. */
static int64_t spare_lower(uint64_t size) {
  int64_t lo = 0, hi = spare.count;
  while (lo < hi) {
    int64_t mid = (lo + hi) / 2;
    if (spare.at[mid].size < size) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

static void pool_guard(mgpuStream stream, const char *what) {
  if (lanes.open >= 0)
    misuse("%s of device memory while side stream %d is open: the pool orders the "
           "reuse of its blocks by the runtime's first stream alone",
           what, lanes.open + 1);
  if (stream && (!lanes.made || (hipStream_t)stream != lanes.trunk))
    misuse("%s of device memory on a stream other than the runtime's first stream: "
           "the pool orders the reuse of its blocks by that stream alone", what);
}

static void release_spare(void) {
  while (spare.count) {
    Block b = list_take(&spare, spare.count - 1);
    hip_ok(hipFree(b.base), "hipFree");
    --driver_blocks;
  }
}

void *mgpuMemAlloc(uint64_t size, mgpuStream stream, bool isHostShared) {
  pool_guard(stream, "allocation");
  use_device();
  if (size == 0) return NULL;
  Block block = {NULL, (size + 255) & ~(uint64_t)255, 0};
  if (isHostShared) {
    hip_ok(hipMallocManaged((void **)&block.base, block.size, hipMemAttachGlobal),
           "hipMallocManaged");
    block.managed = 1;
  } else {
    int64_t i = spare_lower(block.size);
    if (i < spare.count && spare.at[i].size <= 2 * block.size) {
      block = list_take(&spare, i);
    } else {
      hipError_t status = hipMalloc((void **)&block.base, block.size);
      if (status == hipErrorOutOfMemory) {
        (void)hipGetLastError();
        drain();
        release_spare();
        mdhip_fft_trim();
        status = hipMalloc((void **)&block.base, block.size);
      }
      hip_ok(status, "hipMalloc");
      ++driver_blocks;
    }
  }
  list_insert(&used, used_upper(block.base), block);
  return block.base;
}

void mgpuMemFree(void *pointer, mgpuStream stream) {
  if (!pointer) return;
  pool_guard(stream, "free");
  int64_t i = used_find(pointer);
  if (i < 0 || used.at[i].base != (char *)pointer)
    misuse("free of %p, which the pool did not hand out", pointer);
  Block block = list_take(&used, i);
  if (block.managed) {
    drain();
    hip_ok(hipFree(block.base), "hipFree");
    return;
  }
  list_insert(&spare, spare_lower(block.size), block);
}

uint64_t mdrtDeviceTrim(void) {
  if (!device.used) return 0;
  drain();
  uint64_t released = spare.bytes;
  release_spare();
  mdhip_fft_trim();
  return released;
}

void mdrtDeviceMemoryCounts(int64_t counts[8]) {
  counts[0] = used.count;
  counts[1] = (int64_t)used.bytes;
  counts[2] = spare.count;
  counts[3] = (int64_t)spare.bytes;
  counts[4] = driver_blocks;
  counts[5] = mdhip_fft_plan_count();
  counts[6] = mdhip_fft_plan_bytes();
  size_t free_bytes = 0, total = 0;
  counts[7] = device.used && hipMemGetInfo(&free_bytes, &total) == hipSuccess
                  ? (int64_t)free_bytes : -1;
}

/* Device memory: a pool block, else whatever the HIP runtime says is
   device memory (memory of other allocators, such as torch's). */
static int is_device_memory(const void *pointer) {
  if (used_find(pointer) >= 0) return 1;
  hipPointerAttribute_t attributes;
  if (hipPointerGetAttributes(&attributes, pointer) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return attributes.type == hipMemoryTypeDevice;
}

void mgpuMemcpy(void *destination, void *source, size_t size, mgpuStream stream) {
  if (size == 0) return;
  int from_device = is_device_memory(source), to_device = is_device_memory(destination);
  if (from_device && !to_device) mdrtCountHostRead(size, 0, 0);
  hip_ok(hipMemcpyAsync(destination, source, size, hipMemcpyDefault, (hipStream_t)stream),
         "hipMemcpyAsync");
  /* The host reads what came from the device, and may change what went to
     it, as soon as this returns: a copy with a host end is finished here,
     with everything issued before it. */
  if (!from_device || !to_device) {
    mdrtCountHostRead(0, 1, 0);
    drain();
    hip_ok(hipStreamSynchronize((hipStream_t)stream), "hipStreamSynchronize");
  }
}

/* A copy that is finished on return without waiting for the side streams:
   only the trunk's own work before it is waited for. */
void mdrtCudaMemcpyScoped(uintptr_t destination, uintptr_t source, uint64_t size) {
  if (size == 0) return;
  ensure_trunk();
  if (is_device_memory((void *)source) && !is_device_memory((void *)destination))
    mdrtCountHostRead((size_t)size, 0, 0);
  hip_ok(hipMemcpyAsync((void *)destination, (void *)source, (size_t)size, hipMemcpyDefault,
                        lanes.trunk),
         "hipMemcpyAsync");
  mdrtCudaWaitScoped();
}

void mdrtCudaWaitScoped(void) {
  ensure_trunk();
  mdrtCountHostRead(0, 1, 0);
  hip_ok(hipStreamSynchronize(lanes.trunk), "hipStreamSynchronize");
}

void mgpuMemset32(void *destination, unsigned int value, size_t count, mgpuStream stream) {
  hip_ok(hipMemsetD32Async((hipDeviceptr_t)destination, (int)value, count,
                           (hipStream_t)stream),
         "hipMemsetD32Async");
}

void mgpuMemset16(void *destination, unsigned short value, size_t count, mgpuStream stream) {
  hip_ok(hipMemsetD16Async((hipDeviceptr_t)destination, value, count, (hipStream_t)stream),
         "hipMemsetD16Async");
}

/* ======================================================================
   Flags
   ======================================================================
   One page-locked page, mapped into the device's address space:

     words [0, FLAG_PLACES)                   the flag of each place of the
                                              module (a kernel stores to it
                                              across the bus);
     words [FLAG_PLACES, + FLAG_READS)        staging words of reads of
                                              flags that live in device
                                              memory (an async copy each).

   A read (mdrtFlagStart) takes the next of FLAG_READS read records in
   turn and marks the trunk with the record's event; mdrtFlagFinish waits
   for that event only, so the host can issue the work that does not
   depend on the flag first.  A mapped flag is cleared when read. */

enum { FLAG_PLACES = 1024, FLAG_READS = 64 };

static struct {
  volatile int32_t *host;   /* FLAG_PLACES + FLAG_READS words */
  char *device;             /* the device's address of host[0] */
  struct {
    hipEvent_t ready;
    volatile int32_t *word; /* where the value is when `ready` has passed */
    int clears;             /* a mapped flag: cleared by the read */
  } read[FLAG_READS];
  uint64_t reads;
} flags;

static void flags_ready(void) {
  if (flags.host) return;
  ensure_trunk();
  size_t bytes = (FLAG_PLACES + FLAG_READS) * sizeof(int32_t);
  void *page = NULL, *seen = NULL;
  hip_ok(hipHostMalloc(&page, bytes, hipHostMallocMapped), "hipHostMalloc");
  memset(page, 0, bytes);
  hip_ok(hipHostGetDevicePointer(&seen, page, 0), "hipHostGetDevicePointer");
  flags.host = (volatile int32_t *)page;
  flags.device = (char *)seen;
  for (int i = 0; i < FLAG_READS; ++i) flags.read[i].ready = new_marker();
}

/* The place of a mapped flag at device address `address`, -1 if the
   address is not one. */
static int64_t place_of(const void *address) {
  if (!flags.host) return -1;
  const char *a = (const char *)address;
  if (a < flags.device || a >= flags.device + FLAG_PLACES * sizeof(int32_t)) return -1;
  return (int64_t)((a - flags.device) / (ptrdiff_t)sizeof(int32_t));
}

void _mlir_ciface_mdrtFlagMemory(mdrtBuffer1 *result, int64_t place) {
  flags_ready();
  if (place < 0 || place >= FLAG_PLACES) {
    fprintf(stderr, "mdrt: flag place %ld is beyond the %d places of the flag page\n",
            (long)place, FLAG_PLACES);
    exit(1);
  }
  flags.host[place] = 0;
  char *address = flags.device + place * (int64_t)sizeof(int32_t);
  *result = (mdrtBuffer1){address, address, 0, 1, 1};
}

int64_t mdrtFlagStart(void *device_flag) {
  flags_ready();
  int64_t slot = (int64_t)(flags.reads++ % FLAG_READS);
  int64_t place = place_of(device_flag);
  if (place >= 0) {
    flags.read[slot].word = flags.host + place;
    flags.read[slot].clears = 1;
  } else {
    volatile int32_t *stage = flags.host + FLAG_PLACES + slot;
    mdrtCountHostRead(sizeof(int32_t), 0, 0);
    hip_ok(hipMemcpyAsync((void *)stage, device_flag, sizeof(int32_t), hipMemcpyDeviceToHost,
                          lanes.trunk),
           "hipMemcpyAsync");
    flags.read[slot].word = stage;
    flags.read[slot].clears = 0;
  }
  hip_ok(hipEventRecord(flags.read[slot].ready, lanes.trunk), "hipEventRecord");
  return slot;
}

int32_t mdrtFlagFinish(int64_t slot) {
  if (!flags.host || slot < 0 || slot >= FLAG_READS) misuse("no flag read %ld", (long)slot);
  mdrtCountHostRead(0, 0, 1);
  hip_ok(hipEventSynchronize(flags.read[slot].ready), "hipEventSynchronize");
  volatile int32_t *word = flags.read[slot].word;
  int32_t value = *word;
  if (flags.read[slot].clears) *word = 0;
  return value;
}

int32_t mdrtFlagTake(void *device_flag) {
  int64_t place = place_of(device_flag);
  if (place < 0) return mdrtFlagFinish(mdrtFlagStart(device_flag));
  int32_t value = flags.host[place];
  flags.host[place] = 0;
  return value;
}

/* ======================================================================
   Kept memory and staged frames
   ======================================================================
   Kept memory is device memory outside the pool, taken from HIP and given
   back to it directly: an embedding program holds it across runs (the
   state a failed part of a run returns to, the slots of staged frames).

   A staging slot carries up to STAGE_PARTS device buffers of one frame to
   pinned host memory while the steps go on:

     mdrtStagingCopy  copies a buffer into the slot's own device buffer on
                      the current stream, after the work that produced it
                      and before any later work overwrites it;
     mdrtStagingEnd   the runtime's stream of copies waits for the current
                      stream, copies the slot's device buffers to its host
                      buffers, records an event, and runs a one-thread
                      kernel that stores the end's serial number into the
                      slot's mapped word;
     mdrtStagingWait  returns at once when the word holds the serial of the
                      last end, else waits for the event.

   The word is stored by the device as stream work behind the copies, so it
   never claims copies that did not run; a serial number rather than a flag,
   so that a word that lands late (after its wait returned on the event and
   the slot was staged again) is not taken for the current end. */

enum { STAGE_PARTS = 4 };

typedef struct {
  hipEvent_t produced;        /* the current stream at the end */
  hipEvent_t copied;          /* the stream of copies after the end's copies */
  uint32_t ends;              /* serial number of the last end */
  volatile uint32_t *mark;    /* the mapped word the device stores into */
  void *mark_device;
  struct {
    void *device, *host;
    uint64_t room, size;
    int pending;              /* copied since the last end */
  } part[STAGE_PARTS];
} StagingSlot;

static hipStream_t copy_lane = NULL; /* the stream of copies */

void *mdrtDeviceAllocateKept(uint64_t size) {
  use_device();
  if (size == 0) return NULL;
  void *pointer = NULL;
  hipError_t status = hipMalloc(&pointer, size);
  if (status == hipErrorOutOfMemory) { /* the pool's spare blocks go first */
    (void)hipGetLastError();
    drain();
    release_spare();
    mdhip_fft_trim();
    status = hipMalloc(&pointer, size);
  }
  hip_ok(status, "hipMalloc");
  return pointer;
}

void mdrtDeviceFreeKept(void *pointer) {
  if (!pointer) return;
  drain(); /* work issued before may still use it */
  if (copy_lane) hip_ok(hipStreamSynchronize(copy_lane), "hipStreamSynchronize");
  hip_ok(hipFree(pointer), "hipFree");
}

void mdrtDeviceCopyWithin(void *destination, const void *source, uint64_t size) {
  if (size == 0) return;
  hipStream_t stream = (hipStream_t)mgpuStreamCreate();
  hip_ok(hipMemcpyAsync(destination, source, (size_t)size, hipMemcpyDeviceToDevice, stream),
         "hipMemcpyAsync");
}

void mdrtDeviceCopyToHost(void *destination, const void *source, uint64_t size) {
  if (size == 0) return;
  hipStream_t stream = (hipStream_t)mgpuStreamCreate();
  mdrtCountHostRead((size_t)size, 0, 0);
  hip_ok(hipMemcpyAsync(destination, source, (size_t)size, hipMemcpyDeviceToHost, stream),
         "hipMemcpyAsync");
  mdrtCountHostRead(0, 1, 0);
  drain();
  hip_ok(hipStreamSynchronize(stream), "hipStreamSynchronize");
}

void *mdrtStagingCreate(void) {
  ensure_trunk();
  StagingSlot *slot = calloc(1, sizeof *slot);
  if (!slot) {
    fputs("mdrt: out of host memory for a staging slot\n", stderr);
    exit(1);
  }
  if (!copy_lane)
    hip_ok(hipStreamCreateWithFlags(&copy_lane, hipStreamNonBlocking),
           "hipStreamCreateWithFlags");
  slot->produced = new_marker();
  slot->copied = new_marker();
  void *word = NULL;
  hip_ok(hipHostMalloc(&word, sizeof(uint32_t), hipHostMallocMapped | hipHostMallocCoherent),
         "hipHostMalloc");
  slot->mark = (volatile uint32_t *)word;
  *slot->mark = 0;
  hip_ok(hipHostGetDevicePointer(&slot->mark_device, word, 0), "hipHostGetDevicePointer");
  return slot;
}

void *mdrtStagingCopy(void *opaque, int32_t part, const void *source, uint64_t size) {
  StagingSlot *slot = opaque;
  if (!slot || part < 0 || part >= STAGE_PARTS) return NULL;
  if (slot->part[part].room < size) {
    /* the stream of copies may still read the old buffers */
    hip_ok(hipStreamSynchronize(copy_lane), "hipStreamSynchronize");
    mdrtDeviceFreeKept(slot->part[part].device);
    if (slot->part[part].host) hip_ok(hipHostFree(slot->part[part].host), "hipHostFree");
    slot->part[part].device = mdrtDeviceAllocateKept(size);
    hip_ok(hipHostMalloc(&slot->part[part].host, (size_t)size, hipHostMallocDefault),
           "hipHostMalloc");
    slot->part[part].room = size;
  }
  if (size)
    hip_ok(hipMemcpyAsync(slot->part[part].device, source, (size_t)size,
                          hipMemcpyDeviceToDevice, (hipStream_t)mgpuStreamCreate()),
           "hipMemcpyAsync");
  slot->part[part].size = size;
  slot->part[part].pending = 1;
  return slot->part[part].host;
}

/* MDRT_STAGING_MARK_DELAY: the stream of copies sleeps this long on the
   host between an end's copies and its word. */
static void hold_mark(void *opaque) {
  const long ms = (long)(intptr_t)opaque;
  struct timespec pause = {ms / 1000, (ms % 1000) * 1000000L};
  nanosleep(&pause, NULL);
}

void mdrtStagingEnd(void *opaque) {
  StagingSlot *slot = opaque;
  if (!slot) misuse("mdrtStagingEnd of no slot");
  hipStream_t stream = (hipStream_t)mgpuStreamCreate();
  ++slot->ends;
  hip_ok(hipEventRecord(slot->produced, stream), "hipEventRecord");
  hip_ok(hipStreamWaitEvent(copy_lane, slot->produced, 0), "hipStreamWaitEvent");
  for (int i = 0; i < STAGE_PARTS; ++i) {
    if (!slot->part[i].pending) continue;
    slot->part[i].pending = 0;
    if (!slot->part[i].size) continue;
    hip_ok(hipMemcpyAsync(slot->part[i].host, slot->part[i].device, (size_t)slot->part[i].size,
                          hipMemcpyDeviceToHost, copy_lane),
           "hipMemcpyAsync");
  }
  hip_ok(hipEventRecord(slot->copied, copy_lane), "hipEventRecord");
  const char *delay = getenv("MDRT_STAGING_MARK_DELAY");
  const long ms = delay && *delay ? strtol(delay, NULL, 10) : 0;
  if (ms > 0)
    hip_ok(hipLaunchHostFunc(copy_lane, hold_mark, (void *)(intptr_t)ms), "hipLaunchHostFunc");
  hip_ok((hipError_t)mdhip_stage_mark(slot->mark_device, slot->ends, copy_lane),
         "hipLaunchKernel (staging word)");
}

int32_t mdrtStagingWait(void *opaque, const char **call) {
  StagingSlot *slot = opaque;
  if (__atomic_load_n(slot->mark, __ATOMIC_ACQUIRE) == slot->ends) return 0;
  *call = "hipSetDevice";
  hipError_t result = hipSetDevice(device.ordinal); /* a writer's own thread */
  if (result == hipSuccess) {
    *call = "hipEventSynchronize";
    result = hipEventSynchronize(slot->copied);
  }
  /* the word may have landed while the event was waited for */
  if (result != hipSuccess && __atomic_load_n(slot->mark, __ATOMIC_ACQUIRE) == slot->ends)
    return 0;
  return (int32_t)result;
}

void mdrtStagingFail(int32_t result, const char *call) {
  char what[96];
  snprintf(what, sizeof what, "%s of a staged frame", call ? call : "the wait");
  hip_ok((hipError_t)result, what);
}

void mdrtStagingDestroy(void *opaque) {
  StagingSlot *slot = opaque;
  if (!slot) return;
  if (copy_lane) hip_ok(hipStreamSynchronize(copy_lane), "hipStreamSynchronize");
  for (int i = 0; i < STAGE_PARTS; ++i) {
    mdrtDeviceFreeKept(slot->part[i].device);
    if (slot->part[i].host) hip_ok(hipHostFree(slot->part[i].host), "hipHostFree");
  }
  hip_ok(hipHostFree((void *)slot->mark), "hipHostFree");
  hip_ok(hipEventDestroy(slot->produced), "hipEventDestroy");
  hip_ok(hipEventDestroy(slot->copied), "hipEventDestroy");
  free(slot);
}

/* ======================================================================
   Host results: page-locked f64 buffers, one per place of the module,
   kept for the process; a request beyond a buffer's room replaces it
   (after the device is done with the old one) with one of at least twice
   the room, zeroed.
   ====================================================================== */

enum { RESULT_PLACES = 1024 };
static struct {
  double *values;
  int64_t room; /* in values */
} results[RESULT_PLACES];

void _mlir_ciface_mdrtHostResults(mdrtBuffer1 *result, int64_t place, int64_t count) {
  ensure_trunk();
  if (place < 0 || place >= RESULT_PLACES) {
    fprintf(stderr, "mdrt: result place %ld is beyond the %d places of host results\n",
            (long)place, RESULT_PLACES);
    exit(1);
  }
  if (count < 0) misuse("a host result of %ld values", (long)count);
  if (results[place].room < count) {
    int64_t room = results[place].room * 2 > count ? results[place].room * 2 : count;
    if (results[place].values) {
      drain();
      hip_ok(hipHostFree(results[place].values), "hipHostFree");
    }
    void *fresh = NULL;
    hip_ok(hipHostMalloc(&fresh, (size_t)(room ? room : 1) * sizeof(double), 0),
           "hipHostMalloc");
    memset(fresh, 0, (size_t)(room ? room : 1) * sizeof(double));
    results[place].values = (double *)fresh;
    results[place].room = room;
  }
  *result = (mdrtBuffer1){results[place].values, results[place].values, 0, count, 1};
}

/* ======================================================================
   The 3D transforms of the PME grid
   ====================================================================== */

static void transform(int direction, int single, const mdrtBuffer1 *in, const mdrtBuffer1 *out,
                      const float *tables, int64_t k1, int64_t k2, int64_t k3) {
  hipStream_t stream = (hipStream_t)mgpuStreamCreate();
  size_t word = single ? sizeof(float) : sizeof(double);
  void *from = (char *)in->aligned + in->offset * (int64_t)word;
  void *to = (char *)out->aligned + out->offset * (int64_t)word;
  if (mdhip_fft_execute(direction, single, from, to, k1, k2, k3, tables, stream) != 0)
    stop_run("the PME grid transform cannot run: grid edges need the factors 2, 3, 5 and "
             "7 only, and the last edge must be even");
}

void _mlir_ciface_mdrtCudaFFTForward3D(mdrtBuffer1 *real, mdrtBuffer1 *complex,
                                       int64_t k1, int64_t k2, int64_t k3) {
  transform(-1, 0, real, complex, NULL, k1, k2, k3);
}

void _mlir_ciface_mdrtCudaFFTBackward3D(mdrtBuffer1 *complex, mdrtBuffer1 *real,
                                        int64_t k1, int64_t k2, int64_t k3) {
  transform(+1, 0, complex, real, NULL, k1, k2, k3);
}

void _mlir_ciface_mdrtCudaFFTForward3DF32(mdrtBuffer1 *real, mdrtBuffer1 *complex,
                                          int64_t k1, int64_t k2, int64_t k3) {
  transform(-1, 1, real, complex, NULL, k1, k2, k3);
}

void _mlir_ciface_mdrtCudaFFTBackward3DF32(mdrtBuffer1 *complex, mdrtBuffer1 *real,
                                           int64_t k1, int64_t k2, int64_t k3) {
  transform(+1, 1, complex, real, NULL, k1, k2, k3);
}

void _mlir_ciface_mdrtCudaFFTGaussianBackwardF32(mdrtBuffer1 *complex, mdrtBuffer1 *real,
                                                 mdrtBuffer1 *tables, int64_t k1,
                                                 int64_t k2, int64_t k3) {
  transform(+1, 1, complex, real, (const float *)tables->aligned + tables->offset, k1, k2, k3);
}

int mdrtRetainFFTPlan(int64_t k1, int64_t k2, int64_t k3, int forward, int single,
                      const float *tables) {
  (void)forward; /* one plan serves both directions */
  (void)tables;  /* the influence tables are an argument of each call */
  use_device();
  return mdhip_fft_retain(k1, k2, k3, single);
}

void mdrtReleaseFFTPlan(int handle) { mdhip_fft_release(handle); }
