// The update path for callers that cannot allocate device memory (code under HIP graph capture,
// tight serving loops, harnesses that forbid hipMalloc): gs_submit_edge_pairs_device's grouping
// and aggregation kernels, applied at once, with their scratch (directed entries, destinations
// grouped by vertex, work items, block histograms, counters) carved from one fixed device array
// instead of the stream's growable buffers.
//
// The array is a file-scope __device__ object: every GPU that runs this code holds its own copy
// until the process ends. Host code must never pass the array's name to a kernel (on the host it
// is only a stand-in symbol); its device address is looked up once per device with
// hipGetSymbolAddress and that address is used everywhere.
#include <cstring>

#include "gs_internal.h"

namespace gs {

static __device__ __align__(256) unsigned char scratch[GS_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// device address of this device's scratch (looked up on the device's first call)
int scratch_for_device(int device, char** out) {
  GS_CHECK(device >= 0 && device < kMaxDevices, GS_ERR_UNSUPPORTED, "device %d out of range", device);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return fail(GS_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", device,
                  hipGetErrorString(e));
    g_scratch_ptr[device] = ptr;
  }
  *out = static_cast<char*>(g_scratch_ptr[device]);
  return GS_OK;
}

}  // namespace

}  // namespace gs

using namespace gs;

extern "C" int gs_apply_edge_pairs_device_noalloc(gs_store* h, const uint32_t* d_pairs,
                                                  uint64_t nedges) {
  GS_CHECK(h && (d_pairs || nedges == 0), GS_ERR_INVALID, "null argument");
  GS_CHECK(nedges <= GS_NOALLOC_MAX_EDGES, GS_ERR_INVALID,
           "gs_apply_edge_pairs_device_noalloc: %llu edges exceed the limit of %llu per call",
           (unsigned long long)nedges, (unsigned long long)GS_NOALLOC_MAX_EDGES);
  if (nedges == 0) return GS_OK;
  Store* st = &h->impl;
  int device = -1;
  GS_HIP(hipGetDevice(&device));
  GS_CHECK(device == st->device, GS_ERR_INVALID,
           "gs_apply_edge_pairs_device_noalloc: the store lives on device %d, device %d is current",
           st->device, device);
  // no other writer of the store may have work in flight: their launches may write its rows
  // with plain read-modify-writes that would lose these updates
  {
    std::lock_guard<std::mutex> lk(st->mu);
    for (Stream* s : st->streams) {
      GS_CHECK(s->pend_n == 0, GS_ERR_INVALID,
               "gs_apply_edge_pairs_device_noalloc: a stream of the store holds pending updates "
               "(gs_sync first)");
      const hipError_t q = hipStreamQuery(s->stream);
      if (q == hipErrorNotReady)
        return fail(GS_ERR_INVALID, "gs_apply_edge_pairs_device_noalloc: a stream of the store "
                                    "has work in flight (gs_sync first)");
      GS_HIP(q);
    }
  }
  // a stream object on the host stack: the store's own query stream, default tuning, no
  // profiling (nothing here creates events or buffers)
  Stream s;
  s.store = st;
  s.stream = st->qstream;
  s.own_stream = false;
  PairScratch ws;
  const size_t need = carve_pair_scratch(&s, nedges, nullptr, ws);
  GS_CHECK(need <= (size_t)GS_NOALLOC_SCRATCH_BYTES, GS_ERR_INVALID,
           "gs_apply_edge_pairs_device_noalloc: %llu edges into %u vertices need %zu bytes of "
           "scratch, more than the %llu available",
           (unsigned long long)nedges, st->view.n, need,
           (unsigned long long)GS_NOALLOC_SCRATCH_BYTES);
  char* base = nullptr;
  int rc = scratch_for_device(device, &base);
  if (rc) return rc;
  carve_pair_scratch(&s, nedges, base, ws);
  // the counters start at zero (the scratch keeps the previous call's contents); a fill kernel,
  // not a memset, so the sequence also holds up inside a captured graph
  launch_fill_u32(ws.d_counters, kCounterWords, 0u, s.stream);
  GS_HIP(hipGetLastError());
  PairSegs sg{};
  uint32_t nblk = 0;
  add_pair_segment(sg, nblk, d_pairs, nedges);
  rc = run_pair_pipeline(&s, sg, nblk, nedges, ws, nullptr, -1);
  if (rc) return rc;
  uint32_t dropped = 0;
  GS_HIP(hipMemcpyAsync(&dropped, ws.d_counters + 3, sizeof(uint32_t), hipMemcpyDeviceToHost,
                        s.stream));
  GS_HIP(hipStreamSynchronize(s.stream));
  GS_CHECK(dropped == 0, GS_ERR_INVALID,
           "%u submitted updates named a vertex outside the store and were dropped", dropped);
  return GS_OK;
}
