// Batched encode: persistent kernels over a device-resident block work queue
// that spans any number of tasks (textures, mip levels) of mixed formats, one
// kernel per kernel class, running concurrently.
//
// The reference records one compute dispatch per task
// (GpuBlockCompressor::compress, src/block_compressor.rs:540-590).  Small
// tasks then leave most of the GPU idle: a 64x64 BC7 mip level is 256 blocks,
// four waves on a 256-CU part, and its dispatch lasts as long as one lane
// takes to search one block.  Here every BC1-BC5 and BC7 task of a batch is
// cut into work items ("chunks": consecutive blocks of one task), the chunks
// are ordered most expensive first, and per kernel class a grid sized to the
// resident wave count pulls them from an atomic counter until the class's
// queue is empty; the classes' kernels run side by side (launch_on).  Each
// 64-lane wave runs one chunk at a time, so the format branch is uniform
// across the wave.  A chunk's blocks are encoded by the same block functions
// as the per-format kernels, so the bytes are identical.
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <new>
#include <string>
#include <vector>

#include "bc15_block.hpp"
#include "bc7_block.hpp"
#include "bcn_internal.hpp"
#include "texel_load.hpp"

namespace bcn {

// A task as the kernel reads it (uniform per wave, so scalar loads).
struct alignas(16) QueueTask {
  const uint8_t* texture;  // texel (0, 0) of the region
  uint32_t* blocks;        // first block
  uint32_t pitch;          // bytes per texture row
  uint32_t bw;             // blocks per row
  uint32_t variant;        // bcn_variant: BC1..BC5 or BC7
  uint32_t vec;            // rows 16-byte aligned and blocks aligned to their size
  uint32_t src;            // bcn_source_format (an 8-bit layout)
  Bc7Params s;             // BC7 settings (unused by BC1-BC5)
};

// A work item: `count` consecutive blocks of task `task` from block `first`.
struct alignas(16) QueueChunk {
  uint32_t task, first, count, pad;
};

constexpr int kQueueWave = 64;

// Work-queue classes, one persistent kernel each.  Keeping the BC7 search for
// 3 and 4 channels and the BC1-BC5 encoders in separate functions lets each
// be register-allocated on its own (one function holding all three spills
// several hundred bytes per lane).
enum QueueClass { kQueueBc7Rgba = 0, kQueueBc7Rgb = 1, kQueueBc15 = 2, kQueueClasses = 3 };

__global__ void queue_rewind_kernel(uint32_t* next) {
  if (threadIdx.x < kQueueClasses) next[threadIdx.x] = 0u;
}

template <int CLASS>
__device__ __forceinline__ void queue_encode_chunk(const QueueTask* t, const QueueChunk& ch,
                                                   uint32_t lane, uint32_t* keys, uint32_t* sdata) {
  const uint32_t variant = t->variant, bw = t->bw, pitch = t->pitch;
  const bool vec = t->vec != 0;
  const uint8_t* tex = t->texture;
  uint32_t* dst = t->blocks;
  for (uint32_t b = 0; b < ch.count; b += kQueueWave) {
    const bool active = b + lane < ch.count;
    const uint32_t idx = ch.first + (active ? b + lane : 0u);
    uint32_t px[16];
    const uint32_t yy = idx / bw, xx = idx - yy * bw;
    load_block_rgba8(tex, pitch, xx, yy, t->src, vec, px);
    uint32_t out[4] = {0u, 0u, 0u, 0u};
    bool half = false;
    if (CLASS == kQueueBc7Rgba) {
      encode_bc7_block<4>(px, t->s, keys + lane, kQueueWave, sdata + lane, out);
    } else if (CLASS == kQueueBc7Rgb) {
      encode_bc7_block<3>(px, t->s, keys + lane, kQueueWave, sdata + lane, out);
    } else if (variant == BCN_BC1) {
      bc1_core(px, out);
      half = true;
    } else if (variant == BCN_BC2) {
      bc2_alpha(px, out);
      bc1_core(px, out + 2);
    } else if (variant == BCN_BC3) {
      bc3_alpha_block(px, 3, out);
      bc1_core(px, out + 2);
    } else if (variant == BCN_BC4) {
      bc3_alpha_block(px, 0, out);
      half = true;
    } else {
      bc3_alpha_block(px, 0, out);
      bc3_alpha_block(px, 1, out + 2);
    }
    if (active) {
      if (half) {
        uint32_t* o = dst + (size_t)idx * 2;
        if (vec) {
          *reinterpret_cast<uint2*>(o) = make_uint2(out[0], out[1]);
        } else {
          o[0] = out[0];
          o[1] = out[1];
        }
      } else {
        uint32_t* o = dst + (size_t)idx * 4;
        if (vec) {
          *reinterpret_cast<uint4*>(o) = make_uint4(out[0], out[1], out[2], out[3]);
        } else {
#pragma unroll
          for (int i = 0; i < 4; i++) o[i] = out[i];
        }
      }
    }
  }
}

// Every wave takes chunks from its class's counter until it passes nchunks,
// so each wave reaches the exit and the grid drains; a chunk's inactive lanes
// (past its last block) encode the chunk's first block again and store
// nothing.
template <int CLASS>
__global__ __launch_bounds__(kQueueWave)
__attribute__((amdgpu_waves_per_eu(CLASS == kQueueBc15 ? 8 : BCN_BC7_WAVES,
                                   CLASS == kQueueBc15 ? 8 : BCN_BC7_WAVES))) void queue_encode_kernel(
    const QueueTask* __restrict__ tasks, const QueueChunk* __restrict__ chunks, uint32_t nchunks,
    uint32_t* __restrict__ next) {
  __shared__ uint32_t keys[CLASS == kQueueBc15 ? 1 : 32 * kQueueWave];
  __shared__ uint32_t sdata[CLASS == kQueueBc15 ? 1 : 18 * kQueueWave];
  const uint32_t lane = threadIdx.x;
  for (;;) {
    uint32_t ci = 0;
    if (lane == 0) ci = atomicAdd(next, 1u);
    ci = (uint32_t)__builtin_amdgcn_readfirstlane((int)ci);
    if (ci >= nchunks) break;
    const QueueChunk ch = chunks[ci];
    queue_encode_chunk<CLASS>(tasks + ch.task, ch, lane, keys, sdata);
  }
}

static int queue_class(const QueueTask& t) {
  if (t.variant != BCN_BC7) return kQueueBc15;
  return t.s.channels == 4 ? kQueueBc7Rgba : kQueueBc7Rgb;
}

// Relative cost of one block, for ordering the queue (most expensive chunks
// first, so the cheap ones fill the tail).  BC7: a count of the PCA fits and
// index passes its settings ask for; the others are small constants.
static uint32_t block_cost(const QueueTask& t) {
  switch (t.variant) {
    case BCN_BC7: {
      const Bc7Params& s = t.s;
      uint32_t c = 8;
      if (s.mode_selection[0]) c += 48 + (s.skip_mode2 ? 0 : 128);
      if (s.mode_selection[1]) {
        uint32_t t13 = std::max(s.fast_skip_threshold_mode1, s.fast_skip_threshold_mode3);
        c += 64 + 4 * t13;
        if (s.fast_skip_threshold_mode7) c += 64 + 2 * s.fast_skip_threshold_mode7;
      }
      if (s.mode_selection[2] && s.channels > s.mode45_channel0)
        c += 24 * (s.channels - s.mode45_channel0) * (1 + s.refine_iterations[4]);
      if (s.mode_selection[3]) c += 6 * (1 + s.refine_iterations[6]);
      return c * 100;
    }
    case BCN_BC4: return 8;
    case BCN_BC5: return 12;
    default: return 30;
  }
}

}  // namespace bcn

struct bcn_batch {
  int device = 0;
  std::vector<bcn::Task> per_task;       // BC6H: launched per task after the queue
  std::vector<bcn::QueueTask> qtasks;    // in queue order (most expensive first)
  std::vector<uint32_t> qblocks;         // blocks of each queue task
  uint64_t nblocks = 0;
  bcn::QueueTask* d_tasks = nullptr;
  bcn::QueueChunk* d_chunks = nullptr;
  size_t chunk_capacity = 0;
  uint32_t* d_next = nullptr;            // one counter per queue class
  uint32_t nchunks = 0;
  uint32_t class_begin[bcn::kQueueClasses] = {}, class_count[bcn::kQueueClasses] = {};
  uint32_t grid[bcn::kQueueClasses] = {};
  uint32_t cus = 1;
  uint32_t max_resident[bcn::kQueueClasses] = {1, 1, 1};  // resident workgroups per kernel
  bcn_batch_config cfg{};
  hipStream_t capture = nullptr;
  // Fork/join of the class kernels: the launch stream runs the BC7 RGBA
  // class, side[0] the BC7 RGB class, side[1] the BC1-BC5 class and then the
  // BC6H tasks; default priority, non-blocking (they join graph captures).
  hipStream_t side[2] = {nullptr, nullptr};
  hipEvent_t fork = nullptr, join[2] = {nullptr, nullptr};
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::string last_error;
};

namespace {

int bfail(bcn_batch* b, int status, const char* what, hipError_t e = hipSuccess) {
  if (b) {
    if (e != hipSuccess) {
      char buf[256];
      snprintf(buf, sizeof(buf), "%s: %s", what, hipGetErrorString(e));
      b->last_error = buf;
    } else {
      b->last_error = what;
    }
  }
  return status;
}

void drop_graph(bcn_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
}

bool valid_config(const bcn_batch_config& c) {
  return c.chunk_blocks_bc7 >= 64 && c.chunk_blocks_bc7 % 64 == 0 && c.chunk_blocks_bc15 >= 64 &&
         c.chunk_blocks_bc15 % 64 == 0 && c.chunk_blocks_bc7 <= (1u << 24) &&
         c.chunk_blocks_bc15 <= (1u << 24) && c.serial_classes <= 1;
}

// Cuts the queue tasks into chunks under `cfg` (grouped by class, each class
// in queue order), uploads them and sizes the grids.  Synchronous.
int build_queue(bcn_batch* b, const bcn_batch_config& cfg) {
  if (!valid_config(cfg))
    return bfail(b, BCN_ERR_INVALID_ARGUMENT,
                 "chunk sizes must be positive multiples of 64 (at most 2^24), serial_classes 0 or 1");
  std::vector<bcn::QueueChunk> chunks;
  for (int cls = 0; cls < bcn::kQueueClasses; cls++) {
    b->class_begin[cls] = (uint32_t)chunks.size();
    for (uint32_t i = 0; i < (uint32_t)b->qtasks.size(); i++) {
      if (bcn::queue_class(b->qtasks[i]) != cls) continue;
      const uint32_t size = cls == bcn::kQueueBc15 ? cfg.chunk_blocks_bc15 : cfg.chunk_blocks_bc7;
      for (uint32_t first = 0; first < b->qblocks[i]; first += size) {
        bcn::QueueChunk c;
        c.task = i;
        c.first = first;
        c.count = std::min(size, b->qblocks[i] - first);
        c.pad = 0;
        chunks.push_back(c);
        if (chunks.size() >= (1ull << 31))
          return bfail(b, BCN_ERR_INVALID_ARGUMENT, "too many work items; use larger chunks");
      }
    }
    b->class_count[cls] = (uint32_t)chunks.size() - b->class_begin[cls];
  }
  drop_graph(b);
  if (chunks.size() > b->chunk_capacity) {
    if (b->d_chunks) (void)hipFree(b->d_chunks);
    b->d_chunks = nullptr;
    b->chunk_capacity = 0;
    hipError_t e = hipMalloc(&b->d_chunks, chunks.size() * sizeof(bcn::QueueChunk));
    if (e != hipSuccess) return bfail(b, BCN_ERR_HIP, "hipMalloc of the work queue failed", e);
    b->chunk_capacity = chunks.size();
  }
  if (!chunks.empty()) {
    hipError_t e = hipMemcpy(b->d_chunks, chunks.data(), chunks.size() * sizeof(bcn::QueueChunk),
                             hipMemcpyHostToDevice);
    if (e != hipSuccess) return bfail(b, BCN_ERR_HIP, "upload of the work queue failed", e);
  }
  b->nchunks = (uint32_t)chunks.size();
  for (int cls = 0; cls < bcn::kQueueClasses; cls++) {
    // waves_per_cu applies to the BC7 kernels; the BC1-BC5 kernel always
    // fills the device.
    uint32_t want = (cfg.waves_per_cu == 0 || cls == bcn::kQueueBc15) ? b->max_resident[cls]
                                                                       : cfg.waves_per_cu * b->cus;
    // More workgroups than chunks would only take a chunk index and leave.
    b->grid[cls] = std::min(want, b->class_count[cls]);
  }
  b->cfg = cfg;
  return BCN_OK;
}

template <int CLASS>
void launch_class(bcn_batch* b, hipStream_t stream) {
  if (b->class_count[CLASS] == 0) return;
  hipLaunchKernelGGL(bcn::queue_encode_kernel<CLASS>, dim3(b->grid[CLASS]), dim3(bcn::kQueueWave), 0,
                     stream, b->d_tasks, b->d_chunks + b->class_begin[CLASS], b->class_count[CLASS],
                     b->d_next + CLASS);
}

// The rewind kernel, then the three class kernels forked onto three streams
// so that they run concurrently: each is a persistent grid that drains its
// own queue, and when one class's queue runs dry its workgroups leave and
// the others' waiting workgroups take the freed slots, so one class's tail
// is filled by the others' work instead of idling until the next launch.
// The BC6H tasks follow the BC1-BC5 class on its stream.  The launch stream
// waits for both side streams, so work queued after the launch sees every
// block written.  Under stream capture the side streams join the capture
// through the fork event, and the graph gets the same three branches.
// With serial_classes every kernel goes on the launch stream, one class
// after the other.
int launch_on(bcn_batch* b, hipStream_t stream) {
  hipLaunchKernelGGL(bcn::queue_rewind_kernel, dim3(1), dim3(64), 0, stream, b->d_next);
  hipError_t e = hipGetLastError();
  if (b->cfg.serial_classes) {
    launch_class<bcn::kQueueBc7Rgba>(b, stream);
    launch_class<bcn::kQueueBc7Rgb>(b, stream);
    launch_class<bcn::kQueueBc15>(b, stream);
    e = hipGetLastError();
    if (e != hipSuccess) return bfail(b, BCN_ERR_HIP, "launch of the persistent kernels failed", e);
    for (const bcn::Task& t : b->per_task) {
      e = bcn::launch_task(t, stream);
      if (e == hipErrorNotSupported) return bfail(b, BCN_ERR_UNSUPPORTED, "format not built in this library");
      if (e != hipSuccess) return bfail(b, BCN_ERR_HIP, "kernel launch failed", e);
    }
    return BCN_OK;
  }
  if (e == hipSuccess) e = hipEventRecord(b->fork, stream);
  if (e == hipSuccess) e = hipStreamWaitEvent(b->side[0], b->fork, 0);
  if (e == hipSuccess) e = hipStreamWaitEvent(b->side[1], b->fork, 0);
  if (e != hipSuccess) return bfail(b, BCN_ERR_HIP, "fork of the class streams failed", e);
  launch_class<bcn::kQueueBc7Rgba>(b, stream);
  launch_class<bcn::kQueueBc7Rgb>(b, b->side[0]);
  launch_class<bcn::kQueueBc15>(b, b->side[1]);
  e = hipGetLastError();
  int st = BCN_OK;
  if (e != hipSuccess) st = bfail(b, BCN_ERR_HIP, "launch of the persistent kernels failed", e);
  for (size_t i = 0; st == BCN_OK && i < b->per_task.size(); i++) {
    e = bcn::launch_task(b->per_task[i], b->side[1]);
    if (e == hipErrorNotSupported) st = bfail(b, BCN_ERR_UNSUPPORTED, "format not built in this library");
    else if (e != hipSuccess) st = bfail(b, BCN_ERR_HIP, "kernel launch failed", e);
  }
  // Join even after a failed launch, so that a capture in progress can end.
  for (int i = 0; i < 2; i++) {
    hipError_t je = hipEventRecord(b->join[i], b->side[i]);
    if (je == hipSuccess) je = hipStreamWaitEvent(stream, b->join[i], 0);
    if (je != hipSuccess && st == BCN_OK) st = bfail(b, BCN_ERR_HIP, "join of the class streams failed", je);
  }
  return st;
}

template <int CLASS>
hipError_t resident(uint32_t cus, uint32_t* out) {
  int occ = 0;
  hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, bcn::queue_encode_kernel<CLASS>,
                                                             bcn::kQueueWave, 0);
  *out = (uint32_t)std::max(occ, 1) * cus;
  return e;
}

}  // namespace

extern "C" {

int bcn_batch_create(bcn_compressor* c, const bcn_batch_config* config, bcn_batch** out) {
  if (!c || !out) return BCN_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  bcn_batch_config cfg;
  cfg.chunk_blocks_bc7 = 64;
  cfg.chunk_blocks_bc15 = 1024;
  cfg.waves_per_cu = 0;
  cfg.serial_classes = 0;
  if (config) cfg = *config;
  if (!valid_config(cfg)) {
    c->last_error = "chunk sizes must be positive multiples of 64 (at most 2^24), serial_classes 0 or 1";
    return BCN_ERR_INVALID_ARGUMENT;
  }
  bcn::DeviceGuard guard(c->device);
  if (!guard.ok) {
    c->last_error = "hipSetDevice failed";
    return BCN_ERR_HIP;
  }
  bcn_batch* b = new (std::nothrow) bcn_batch();
  if (!b) return BCN_ERR_INVALID_ARGUMENT;
  b->device = c->device;
  // The compressor keeps its tasks until the batch exists: on any failure
  // below they stay queued, as if the call had not been made.
  const std::vector<bcn::Task>& tasks = c->tasks;
  // Queue tasks, then a stable sort by block cost (most expensive first).
  struct Entry {
    bcn::QueueTask q;
    uint32_t blocks, cost;
  };
  std::vector<Entry> entries;
  for (const bcn::Task& t : tasks) {
    uint32_t nb = bcn::blocks_x(t) * bcn::blocks_y(t);
    if (t.variant == BCN_BC6H || t.variant == BCN_BC6H_SF16) {
      b->per_task.push_back(t);
      continue;
    }
    if (nb == 0) continue;
    Entry en{};
    en.q.texture = t.texture;
    en.q.blocks = reinterpret_cast<uint32_t*>(t.blocks);
    en.q.pitch = t.pitch;
    en.q.bw = bcn::blocks_x(t);
    en.q.variant = (uint32_t)t.variant;
    en.q.src = (uint32_t)t.format;
    const uintptr_t align = (t.variant == BCN_BC1 || t.variant == BCN_BC4) ? 7u : 15u;
    en.q.vec = ((((uintptr_t)t.texture | (uintptr_t)t.pitch) & 15u) == 0 &&
                ((uintptr_t)t.blocks & align) == 0)
                   ? 1u
                   : 0u;
    memcpy(&en.q.s, t.settings, sizeof(en.q.s));
    en.blocks = nb;
    en.cost = bcn::block_cost(en.q);
    entries.push_back(en);
  }
  std::stable_sort(entries.begin(), entries.end(),
                   [](const Entry& a, const Entry& x) { return a.cost > x.cost; });
  for (const Entry& en : entries) {
    b->qtasks.push_back(en.q);
    b->qblocks.push_back(en.blocks);
    b->nblocks += en.blocks;
  }
  int cus = 0;
  hipError_t e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, b->device);
  b->cus = (uint32_t)std::max(cus, 1);
  if (e == hipSuccess) e = resident<bcn::kQueueBc7Rgba>(b->cus, &b->max_resident[bcn::kQueueBc7Rgba]);
  if (e == hipSuccess) e = resident<bcn::kQueueBc7Rgb>(b->cus, &b->max_resident[bcn::kQueueBc7Rgb]);
  if (e == hipSuccess) e = resident<bcn::kQueueBc15>(b->cus, &b->max_resident[bcn::kQueueBc15]);
  if (e == hipSuccess && !b->qtasks.empty())
    e = hipMalloc(&b->d_tasks, b->qtasks.size() * sizeof(bcn::QueueTask));
  if (e == hipSuccess && !b->qtasks.empty())
    e = hipMemcpy(b->d_tasks, b->qtasks.data(), b->qtasks.size() * sizeof(bcn::QueueTask),
                  hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMalloc(&b->d_next, bcn::kQueueClasses * sizeof(uint32_t));
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->capture, hipStreamNonBlocking);
  for (int i = 0; i < 2; i++) {
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->side[i], hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->join[i], hipEventDisableTiming);
  }
  if (e == hipSuccess) e = hipEventCreateWithFlags(&b->fork, hipEventDisableTiming);
  if (e != hipSuccess) {
    char buf[256];
    snprintf(buf, sizeof(buf), "batch set-up failed: %s", hipGetErrorString(e));
    c->last_error = buf;
    bcn_batch_destroy(b);
    return BCN_ERR_HIP;
  }
  int st = build_queue(b, cfg);
  if (st != BCN_OK) {
    c->last_error = b->last_error;
    bcn_batch_destroy(b);
    return st;
  }
  c->tasks.clear();
  c->last_error.clear();
  *out = b;
  return BCN_OK;
}

int bcn_batch_launch(bcn_batch* b, void* hip_stream) {
  if (!b) return BCN_ERR_INVALID_ARGUMENT;
  bcn::DeviceGuard guard(b->device);
  if (!guard.ok) return bfail(b, BCN_ERR_HIP, "hipSetDevice failed");
  int st = launch_on(b, static_cast<hipStream_t>(hip_stream));
  if (st == BCN_OK) b->last_error.clear();
  return st;
}

int bcn_batch_launch_graph(bcn_batch* b, void* hip_stream) {
  if (!b) return BCN_ERR_INVALID_ARGUMENT;
  bcn::DeviceGuard guard(b->device);
  if (!guard.ok) return bfail(b, BCN_ERR_HIP, "hipSetDevice failed");
  if (!b->exec) {
    hipError_t e = hipStreamBeginCapture(b->capture, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return bfail(b, BCN_ERR_HIP, "hipStreamBeginCapture failed", e);
    int st = launch_on(b, b->capture);
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(b->capture, &g);
    if (st != BCN_OK) {
      if (g) (void)hipGraphDestroy(g);
      return st;
    }
    if (e != hipSuccess) return bfail(b, BCN_ERR_HIP, "hipStreamEndCapture failed", e);
    b->graph = g;
    e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
      drop_graph(b);
      return bfail(b, BCN_ERR_HIP, "hipGraphInstantiate failed", e);
    }
  }
  hipError_t e = hipGraphLaunch(b->exec, static_cast<hipStream_t>(hip_stream));
  if (e != hipSuccess) return bfail(b, BCN_ERR_HIP, "hipGraphLaunch failed", e);
  b->last_error.clear();
  return BCN_OK;
}

int bcn_batch_autotune(bcn_batch* b, void* hip_stream, int reps, bcn_batch_config* chosen) {
  if (!b || reps < 1) return BCN_ERR_INVALID_ARGUMENT;
  bcn::DeviceGuard guard(b->device);
  if (!guard.ok) return bfail(b, BCN_ERR_HIP, "hipSetDevice failed");
  hipStream_t stream = static_cast<hipStream_t>(hip_stream);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  if (e != hipSuccess) {
    if (e0) (void)hipEventDestroy(e0);
    return bfail(b, BCN_ERR_HIP, "hipEventCreate failed", e);
  }
  const bcn_batch_config start = b->cfg;
  bcn_batch_config best = start;
  float best_ms = -1.0f;
  int status = BCN_OK;
  // Time one configuration: a warm-up launch, then the best of `reps`.
  auto time_config = [&](const bcn_batch_config& cfg, float* ms_out) -> int {
    int st = build_queue(b, cfg);
    if (st != BCN_OK) return st;
    st = launch_on(b, stream);
    if (st != BCN_OK) return st;
    float bestc = -1.0f;
    for (int r = 0; r < reps; r++) {
      hipError_t he = hipEventRecord(e0, stream);
      if (he == hipSuccess) st = launch_on(b, stream);
      if (he == hipSuccess && st == BCN_OK) he = hipEventRecord(e1, stream);
      if (he == hipSuccess && st == BCN_OK) he = hipEventSynchronize(e1);
      float ms = 0.0f;
      if (he == hipSuccess && st == BCN_OK) he = hipEventElapsedTime(&ms, e0, e1);
      if (st != BCN_OK) return st;
      if (he != hipSuccess) return bfail(b, BCN_ERR_HIP, "timing failed", he);
      if (bestc < 0.0f || ms < bestc) bestc = ms;
    }
    *ms_out = bestc;
    return BCN_OK;
  };
  auto consider = [&](const bcn_batch_config& cfg) {
    if (status != BCN_OK) return;
    float ms = 0.0f;
    status = time_config(cfg, &ms);
    if (status == BCN_OK && (best_ms < 0.0f || ms < best_ms)) {
      best_ms = ms;
      best = cfg;
    }
  };
  // Coordinate search: BC1-BC5 chunk size, then BC7 chunk size, then grid,
  // then the class schedule.
  const uint32_t bc15_sizes[] = {256, 1024, 4096};
  const uint32_t bc7_sizes[] = {64, 128, 256};
  for (uint32_t s : bc15_sizes) {
    bcn_batch_config cfg = best;
    cfg.chunk_blocks_bc15 = s;
    consider(cfg);
  }
  for (uint32_t s : bc7_sizes) {
    bcn_batch_config cfg = best;
    cfg.chunk_blocks_bc7 = s;
    consider(cfg);
  }
  const uint32_t resident_per_cu = std::max(1u, b->max_resident[bcn::kQueueBc7Rgb] / b->cus);
  for (uint32_t w : {resident_per_cu / 2, resident_per_cu * 3 / 4, 0u}) {
    if (w == 0 && best.waves_per_cu == 0) continue;  // already timed
    if (w != 0 && w == resident_per_cu / 2 && resident_per_cu < 2) continue;
    bcn_batch_config cfg = best;
    cfg.waves_per_cu = w;
    consider(cfg);
  }
  // Last, the class schedule: concurrent or one class after another.
  {
    bcn_batch_config cfg = best;
    cfg.serial_classes = best.serial_classes ? 0u : 1u;
    consider(cfg);
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (status != BCN_OK) {
    (void)build_queue(b, start);
    return status;
  }
  status = build_queue(b, best);
  if (status != BCN_OK) return status;
  if (chosen) *chosen = best;
  b->last_error.clear();
  return BCN_OK;
}

int bcn_batch_get_info(const bcn_batch* b, bcn_batch_info* info) {
  if (!b || !info) return BCN_ERR_INVALID_ARGUMENT;
  info->tasks = (uint32_t)(b->qtasks.size() + b->per_task.size());
  info->persistent_tasks = (uint32_t)b->qtasks.size();
  info->blocks = b->nblocks;
  info->chunks = b->nchunks;
  info->grid = std::max(b->grid[bcn::kQueueBc7Rgba], b->grid[bcn::kQueueBc7Rgb]);
  info->grid = std::max(info->grid, b->grid[bcn::kQueueBc15]);
  info->config = b->cfg;
  return BCN_OK;
}

const char* bcn_batch_last_error(const bcn_batch* b) { return b ? b->last_error.c_str() : ""; }

void bcn_batch_destroy(bcn_batch* b) {
  if (!b) return;
  bcn::DeviceGuard guard(b->device);
  drop_graph(b);
  if (b->capture) (void)hipStreamDestroy(b->capture);
  for (int i = 0; i < 2; i++) {
    if (b->side[i]) (void)hipStreamDestroy(b->side[i]);
    if (b->join[i]) (void)hipEventDestroy(b->join[i]);
  }
  if (b->fork) (void)hipEventDestroy(b->fork);
  if (b->d_tasks) (void)hipFree(b->d_tasks);
  if (b->d_chunks) (void)hipFree(b->d_chunks);
  if (b->d_next) (void)hipFree(b->d_next);
  delete b;
}

}  // extern "C"
