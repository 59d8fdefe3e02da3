// stream_controller.hip -- GpuStreamController / GpuStreamRunner /
// HostSeedsMemory / DeviceSeedsMemory of the reference
// (src/gpu_stream_controller.{h,cpp}, src/gpu_stream_runner.{h,cpp},
// src/host_seeds_memory.{h,cpp}, src/device_seeds_memory.{h,cu}) over the
// extender handles of this library.
//
// A controller owns two resources; a resource is a pair of streams plus the
// device seed arrays its runs use.  Run* takes a free resource, enqueues the
// host-to-device copies, the kernels and the copies back on that resource's
// streams and returns its id; WaitRun synchronizes the resource and
// ReleaseResource hands it back.  The gapped extender keeps per-handle ring
// workspace, so the two sides of a gapped run go on the resource's first
// stream one after the other (the reference puts them on two streams; the
// results do not depend on that).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <deque>
#include <memory>
#include <string>

#include "ghostz_hip.h"
#include "gz_common.h"
#include "gz_runtime.h"

using namespace gz;

namespace {

constexpr int kResources = 2;  // kMaxNumberOfConcurrentExecutions
constexpr int kStreams = 2;    // GpuStreamRunner::kNumberStream

template <class T>
hipError_t AllocArray(T **p, size_t n) {
  return gz::Malloc((void **)p, std::max<size_t>(n, 1) * sizeof(T));
}

}  // namespace

struct gz_host_seeds_memory_s {
  size_t size = 0;
  uint32_t *query_seed_ids = nullptr, *query_seed_starts = nullptr, *query_ids = nullptr,
           *query_positions = nullptr, *database_positions = nullptr;
  int *result_values = nullptr;
  ~gz_host_seeds_memory_s() {
    void *ps[] = {query_seed_ids, query_seed_starts, query_ids, query_positions,
                  database_positions, result_values};
    for (void *p : ps)
      if (p) (void)gz::HostFree(p);
  }
};

namespace {

struct DeviceSeeds {
  size_t size = 0;
  uint32_t *query_seed_ids = nullptr, *query_seed_starts = nullptr, *query_ids = nullptr,
           *query_positions = nullptr, *database_positions = nullptr;
  int *result_values = nullptr;
  void *temp = nullptr;  // SetQuerySeeds' scan storage
  size_t temp_bytes = 0;
  uint32_t *convert_temp = nullptr;  // 2 n: the gathers of a gapped run
  void Free() {
    void *ps[] = {query_seed_ids, query_seed_starts, query_ids, query_positions,
                  database_positions, result_values, temp, convert_temp};
    for (void *p : ps)
      if (p) (void)gz::Free(p);
    *this = DeviceSeeds();
  }
  hipError_t Init(size_t n) {
    Free();
    size = n;
    temp_bytes = gz_ungapped_extender_query_seeds_temp_bytes(n);
    hipError_t e = AllocArray(&query_seed_ids, n);
    if (e == hipSuccess) e = AllocArray(&query_seed_starts, n);
    if (e == hipSuccess) e = AllocArray(&query_ids, n);
    if (e == hipSuccess) e = AllocArray(&query_positions, n);
    if (e == hipSuccess) e = AllocArray(&database_positions, n);
    if (e == hipSuccess) e = AllocArray(&result_values, n);
    if (e == hipSuccess) e = gz::Malloc(&temp, std::max<size_t>(temp_bytes, 16));
    if (e == hipSuccess) e = AllocArray(&convert_temp, 2 * n);
    return e;
  }
};

}  // namespace

struct gz_stream_controller_s {
  gz_distance_calculator_t distance = nullptr;
  gz_ungapped_extender_t ungapped = nullptr;
  gz_gapped_extender_t gapped = nullptr;
  gz_stream_s streams[kResources][kStreams] = {};
  DeviceSeeds device[kResources];
  std::deque<int> available;
  bool initialized = false;
  // the gapped extender's ring workspace belongs to its handle: a gapped run
  // waits for the previous one (on whichever resource) to finish
  hipEvent_t gapped_done = nullptr;
  bool gapped_pending = false;
  ~gz_stream_controller_s() {
    if (gapped_done) (void)hipEventDestroy(gapped_done);
    for (int r = 0; r < kResources; ++r) {
      device[r].Free();
      for (int k = 0; k < kStreams; ++k)
        if (streams[r][k].s) (void)hipStreamDestroy(streams[r][k].s);
    }
  }
  gz_status Take(int *id) {
    GZ_REQUIRE(initialized, "stream controller: Init first");
    GZ_REQUIRE(!available.empty(), "stream controller: no free resource (ReleaseResource first)");
    *id = available.front();
    available.pop_front();
    return GZ_OK;
  }
};

extern "C" {

gz_status gz_host_seeds_memory_create(size_t size, gz_host_seeds_memory_t *out) {
  GZ_REQUIRE(out && size > 0, "gz_host_seeds_memory_create: bad argument");
  std::unique_ptr<gz_host_seeds_memory_s> m(new gz_host_seeds_memory_s());
  m->size = size;
  void **ps[] = {(void **)&m->query_seed_ids, (void **)&m->query_seed_starts,
                 (void **)&m->query_ids, (void **)&m->query_positions,
                 (void **)&m->database_positions, (void **)&m->result_values};
  for (void **p : ps) GZ_HIP_CHECK(gz::HostMalloc(p, size * 4, hipHostMallocDefault));
  *out = m.release();
  return GZ_OK;
}

gz_status gz_host_seeds_memory_destroy(gz_host_seeds_memory_t m) {
  delete m;
  return GZ_OK;
}

size_t gz_host_seeds_memory_size(gz_host_seeds_memory_t m) { return m ? m->size : 0; }
uint32_t *gz_host_seeds_memory_query_seed_ids(gz_host_seeds_memory_t m) {
  return m ? m->query_seed_ids : nullptr;
}
uint32_t *gz_host_seeds_memory_query_seed_starts(gz_host_seeds_memory_t m) {
  return m ? m->query_seed_starts : nullptr;
}
uint32_t *gz_host_seeds_memory_query_ids(gz_host_seeds_memory_t m) {
  return m ? m->query_ids : nullptr;
}
uint32_t *gz_host_seeds_memory_query_concatenated_positions(gz_host_seeds_memory_t m) {
  return m ? m->query_positions : nullptr;
}
uint32_t *gz_host_seeds_memory_database_positions(gz_host_seeds_memory_t m) {
  return m ? m->database_positions : nullptr;
}
int *gz_host_seeds_memory_result_values(gz_host_seeds_memory_t m) {
  return m ? m->result_values : nullptr;
}

gz_status gz_stream_controller_create(gz_stream_controller_t *out) {
  GZ_REQUIRE(out, "null handle pointer");
  std::unique_ptr<gz_stream_controller_s> c(new gz_stream_controller_s());
  for (int r = 0; r < kResources; ++r)
    for (int k = 0; k < kStreams; ++k)
      GZ_HIP_CHECK(hipStreamCreateWithFlags(&c->streams[r][k].s, hipStreamNonBlocking));
  *out = c.release();
  return GZ_OK;
}

gz_status gz_stream_controller_destroy(gz_stream_controller_t c) {
  if (c) (void)gz_stream_controller_wait_run_all(c);
  delete c;
  return GZ_OK;
}

gz_status gz_stream_controller_init(gz_stream_controller_t c, size_t device_buffer_size) {
  GZ_REQUIRE(c && device_buffer_size > 0, "gz_stream_controller_init: bad argument");
  GZ_REQUIRE(c->available.size() == (size_t)kResources || !c->initialized,
             "gz_stream_controller_init: a resource is still in use");
  for (int r = 0; r < kResources; ++r) {
    for (int k = 0; k < kStreams; ++k) GZ_HIP_CHECK(hipStreamSynchronize(c->streams[r][k].s));
    GZ_HIP_CHECK(c->device[r].Init(device_buffer_size));
  }
  c->available.clear();
  for (int r = 0; r < kResources; ++r) c->available.push_back(r);
  c->initialized = true;
  return GZ_OK;
}

gz_status gz_stream_controller_set_distance_calculator(gz_stream_controller_t c,
                                                       gz_distance_calculator_t d) {
  GZ_REQUIRE(c, "null handle");
  c->distance = d;
  return GZ_OK;
}
gz_status gz_stream_controller_set_ungapped_extender(gz_stream_controller_t c,
                                                     gz_ungapped_extender_t u) {
  GZ_REQUIRE(c, "null handle");
  c->ungapped = u;
  return GZ_OK;
}
gz_status gz_stream_controller_set_gapped_extender(gz_stream_controller_t c,
                                                   gz_gapped_extender_t g) {
  GZ_REQUIRE(c, "null handle");
  c->gapped = g;
  return GZ_OK;
}

gz_status gz_stream_controller_run_distance_calculation(gz_stream_controller_t c, size_t size,
                                                        gz_host_seeds_memory_t m,
                                                        int *resource_id) {
  GZ_REQUIRE(c && m && resource_id && c->distance, "distance run: bad argument");
  GZ_REQUIRE(size <= m->size && size <= c->device[0].size, "distance run: beyond the buffers");
  int id = -1;
  gz_status st = c->Take(&id);
  if (st != GZ_OK) return st;
  DeviceSeeds &d = c->device[id];
  // GpuStreamRunner::RunDistanceCalculationAsync (src/gpu_stream_runner.cpp:38-48)
  st = gz_distance_calculator_calculate_distances_async(
      c->distance, size, m->query_positions, m->database_positions, (char *)m->result_values,
      d.query_positions, d.database_positions, (char *)d.result_values, &c->streams[id][0]);
  if (st != GZ_OK) {
    c->available.push_front(id);
    return st;
  }
  *resource_id = id;
  return GZ_OK;
}

gz_status gz_stream_controller_run_ungapped_extension_with_trigger(gz_stream_controller_t c,
                                                                   size_t query_seed_size,
                                                                   size_t size,
                                                                   gz_host_seeds_memory_t m,
                                                                   int *resource_id) {
  GZ_REQUIRE(c && m && resource_id && c->ungapped, "ungapped run: bad argument");
  GZ_REQUIRE(size <= m->size && size <= c->device[0].size && query_seed_size <= m->size,
             "ungapped run: beyond the buffers");
  int id = -1;
  gz_status st = c->Take(&id);
  if (st != GZ_OK) return st;
  DeviceSeeds &d = c->device[id];
  gz_stream_t s = &c->streams[id][0];
  // GpuStreamRunner::RunUngappedExtensionWithTriggerAsync
  // (src/gpu_stream_runner.cpp:50-69): expand the seed runs, then extend
  st = gz_ungapped_extender_set_query_seeds(
      c->ungapped, query_seed_size, size, m->query_seed_ids, m->query_seed_starts, d.temp_bytes,
      d.temp, d.query_seed_ids, d.query_seed_starts, d.query_ids, d.query_positions,
      d.database_positions, s);
  if (st == GZ_OK)
    st = gz_ungapped_extender_extend_with_trigger_async(
        c->ungapped, size, nullptr, nullptr, m->database_positions, (char *)m->result_values,
        d.query_ids, d.query_positions, d.database_positions, (char *)d.result_values,
        (int *)d.query_seed_ids, s);
  if (st != GZ_OK) {
    c->available.push_front(id);
    return st;
  }
  *resource_id = id;
  return GZ_OK;
}

gz_status gz_stream_controller_run_gapped_extension(gz_stream_controller_t c,
                                                    size_t reverse_offset, size_t reverse_length,
                                                    size_t forward_offset, size_t forward_length,
                                                    gz_host_seeds_memory_t m, int *resource_id) {
  GZ_REQUIRE(c && m && resource_id && c->gapped, "gapped run: bad argument");
  GZ_REQUIRE(reverse_offset + reverse_length <= m->size &&
                 forward_offset + forward_length <= m->size &&
                 reverse_offset + reverse_length <= c->device[0].size &&
                 forward_offset + forward_length <= c->device[0].size,
             "gapped run: beyond the buffers");
  if (!c->gapped_done) GZ_HIP_CHECK(hipEventCreateWithFlags(&c->gapped_done, hipEventDisableTiming));
  int id = -1;
  gz_status st = c->Take(&id);
  if (st != GZ_OK) return st;
  DeviceSeeds &d = c->device[id];
  gz_stream_t s = &c->streams[id][0];
  if (c->gapped_pending && hipStreamWaitEvent(s->s, c->gapped_done, 0) != hipSuccess) {
    c->available.push_front(id);
    SetError("gapped run: hipStreamWaitEvent failed");
    return GZ_ERROR_HIP;
  }
  // GpuStreamRunner::RunGappedExtensionAsync (src/gpu_stream_runner.cpp:71-107):
  // each side gathers its seed positions by seed id, then extends; the
  // positions come back as the best cells and result_values as the scores.
  // Both sides share the gather scratch, so they run one after the other.
  struct Side {
    size_t off, len;
    int reverse;
  } sides[2] = {{reverse_offset, reverse_length, 1}, {forward_offset, forward_length, 0}};
  for (const Side &sd : sides) {
    if (sd.len == 0 || st != GZ_OK) continue;
    st = gz_gapped_extender_convert_to_gapped_extension_seeds_async(
        c->gapped, sd.len, sd.reverse, m->query_seed_ids + sd.off,
        m->query_positions + sd.off, m->database_positions + sd.off,
        d.query_seed_ids + sd.off, d.convert_temp, d.query_positions + sd.off,
        d.database_positions + sd.off, s);
    if (st == GZ_OK)
      st = gz_gapped_extender_extend_one_side_score_only_async(
          c->gapped, sd.len, sd.reverse, m->query_positions + sd.off,
          m->database_positions + sd.off, m->result_values + sd.off, d.query_positions + sd.off,
          d.database_positions + sd.off, d.result_values + sd.off, s);
  }
  if (st == GZ_OK && hipEventRecord(c->gapped_done, s->s) != hipSuccess) {
    SetError("gapped run: hipEventRecord failed");
    st = GZ_ERROR_HIP;
  }
  if (st != GZ_OK) {
    c->available.push_front(id);
    return st;
  }
  c->gapped_pending = true;
  *resource_id = id;
  return GZ_OK;
}

gz_status gz_stream_controller_wait_run(gz_stream_controller_t c, int resource_id) {
  GZ_REQUIRE(c && resource_id < kResources, "wait run: bad argument");
  if (resource_id < 0) return GZ_OK;
  for (int k = 0; k < kStreams; ++k)
    GZ_HIP_CHECK(hipStreamSynchronize(c->streams[resource_id][k].s));
  return GZ_OK;
}

gz_status gz_stream_controller_wait_run_all(gz_stream_controller_t c) {
  GZ_REQUIRE(c, "null handle");
  for (int r = 0; r < kResources; ++r) {
    const gz_status st = gz_stream_controller_wait_run(c, r);
    if (st != GZ_OK) return st;
  }
  return GZ_OK;
}

int gz_stream_controller_finished_run(gz_stream_controller_t c, int resource_id) {
  if (!c || resource_id < 0 || resource_id >= kResources) return 1;
  for (int k = 0; k < kStreams; ++k)
    if (hipStreamQuery(c->streams[resource_id][k].s) == hipErrorNotReady) return 0;
  return 1;
}

gz_status gz_stream_controller_release_resource(gz_stream_controller_t c, int resource_id) {
  GZ_REQUIRE(c && resource_id >= 0 && resource_id < kResources,
             "release resource: bad resource id");
  for (int r : c->available)
    GZ_REQUIRE(r != resource_id, "release resource: resource is not in use");
  c->available.push_back(resource_id);
  return GZ_OK;
}

uint32_t gz_stream_controller_available_resources(gz_stream_controller_t c) {
  return c ? (uint32_t)c->available.size() : 0;
}

}  // extern "C"
