// Groups of stores ingested and queried together (the batched-streams mode): many independent
// graphs of different sizes share one virtual vertex space (member i's vertex v is global vertex
// bases[i] + v), so one launch sequence of the grouping and aggregation kernels serves all of
// them, and one batched Boruvka answers connectivity for all of them. Each member keeps its own
// geometry (B, C, seed) and its own HBM store; nothing of one member is visible to another.
#include <algorithm>
#include <cstring>

#include "gs_internal.h"

namespace gs {

// host pairs of a group submission are offset into the group's vertex space through a pinned
// staging of this many edges
constexpr uint64_t kGroupStageEdges = 1ULL << 22;

static void unregister_from_members(Group& g) {
  for (Store* m : g.members) {
    std::lock_guard<std::mutex> lk(m->mu);
    auto& v = m->streams;
    v.erase(std::remove(v.begin(), v.end(), g.s), v.end());
  }
}

static void destroy_group(gs_group* h) {
  Group& g = h->impl;
  // pending updates are applied before the group's stream goes away (as gs_stream_destroy does)
  if (g.s && g.s->stream && g.s->d_pend[0]) (void)drain_stream(g.s);
  if (g.s && g.s->stream) (void)hipStreamSynchronize(g.s->stream);
  free_group_query(&g);
  if (g.s) {
    unregister_from_members(g);
    destroy_stream_resources(*g.s);
  }
  delete g.handle;
  (void)hipFree(g.d_descs);
  (void)hipFree(g.d_bases);
  (void)hipFree(g.d_vgraph);
  delete h;
}

}  // namespace gs

using namespace gs;

extern "C" {

int gs_group_create(gs_store* const* stores, uint32_t count, gs_group** out) {
  GS_CHECK(stores && out && count >= 1, GS_ERR_INVALID, "gs_group_create: bad argument");
  gs_group* h = new gs_group();
  Group& g = h->impl;
  uint64_t total = 0;
  std::vector<GraphDesc> descs(count);
  for (uint32_t i = 0; i < count; i++) {
    if (!stores[i]) {
      delete h;
      return fail(GS_ERR_INVALID, "gs_group_create: member %u is null", i);
    }
    Store* st = &stores[i]->impl;
    for (uint32_t j = 0; j < i; j++)
      if (&stores[j]->impl == st) {
        delete h;
        return fail(GS_ERR_INVALID, "gs_group_create: store given twice (members %u, %u)", j, i);
      }
    if (st->cols_per_sample != 1 || (st->flags & GS_STORE_CUBESKETCH) || st->view.cols > 64) {
      delete h;
      return fail(GS_ERR_UNSUPPORTED,
                  "gs_group_create: member %u is a wide sketch (%u columns, %u per sample%s); "
                  "groups take one-column-per-sample sketches of at most 64 columns",
                  i, st->view.cols, st->cols_per_sample,
                  (st->flags & GS_STORE_CUBESKETCH) ? ", CubeSketch rule" : "");
    }
    g.members.push_back(st);
    g.bases.push_back((uint32_t)total);
    const StoreView& v = st->view;
    descs[i] = GraphDesc{v.rows, v.det, v.n, v.bpc, v.cols, v.vstride, v.seed, (uint32_t)total,
                         st->num_samples};
    total += v.n;
    g.max_cols = std::max(g.max_cols, v.cols);
    g.max_bpc = std::max(g.max_bpc, v.bpc);
    g.min_bpc = i ? std::min(g.min_bpc, v.bpc) : v.bpc;
  }
  if (total > (1ULL << 26)) {
    delete h;
    return fail(GS_ERR_INVALID, "gs_group_create: %llu vertices in total exceed 2^26",
                (unsigned long long)total);
  }
  g.n_total = (uint32_t)total;
  g.part = make_part_params(g.n_total);
  std::vector<uint32_t> vgraph(total);
  for (uint32_t i = 0; i < count; i++)
    std::fill(vgraph.begin() + g.bases[i], vgraph.begin() + g.bases[i] + descs[i].n, i);
  int rc = GS_OK;
  if (hipMalloc(&g.d_descs, sizeof(GraphDesc) * count) != hipSuccess ||
      hipMalloc(&g.d_bases, sizeof(uint32_t) * count) != hipSuccess ||
      hipMalloc(&g.d_vgraph, sizeof(uint32_t) * total) != hipSuccess)
    rc = fail(GS_ERR_OOM, "gs_group_create: device allocation failed");
  if (!rc && (hipMemcpy(g.d_descs, descs.data(), sizeof(GraphDesc) * count,
                        hipMemcpyHostToDevice) != hipSuccess ||
              hipMemcpy(g.d_bases, g.bases.data(), sizeof(uint32_t) * count,
                        hipMemcpyHostToDevice) != hipSuccess ||
              hipMemcpy(g.d_vgraph, vgraph.data(), sizeof(uint32_t) * total,
                        hipMemcpyHostToDevice) != hipSuccess))
    rc = fail(GS_ERR_HIP, "gs_group_create: copy of the member table failed");
  if (!rc) {
    g.handle = new gs_stream();
    g.s = &g.handle->impl;
    g.s->group = &g;
    rc = init_stream_objects(g.s);
  }
  if (rc) {
    destroy_group(h);
    return rc;
  }
  for (Store* m : g.members) {
    std::lock_guard<std::mutex> lk(m->mu);
    m->streams.push_back(g.s);
  }
  *out = h;
  return GS_OK;
}

int gs_group_destroy(gs_group* h) {
  if (h) destroy_group(h);
  return GS_OK;
}

int gs_group_info(const gs_group* h, uint32_t member, uint32_t* base, uint32_t* total) {
  GS_CHECK(h, GS_ERR_INVALID, "null group");
  const Group& g = h->impl;
  GS_CHECK(member < g.members.size(), GS_ERR_INVALID, "member %u out of range", member);
  if (base) *base = g.bases[member];
  if (total) *total = g.n_total;
  return GS_OK;
}

int gs_group_stream(gs_group* h, gs_stream** out) {
  GS_CHECK(h && out, GS_ERR_INVALID, "null argument");
  *out = h->impl.handle;
  return GS_OK;
}

int gs_group_submit_pairs(gs_group* h, const uint32_t* const* pairs, const uint64_t* nedges) {
  GS_CHECK(h && pairs && nedges, GS_ERR_INVALID, "null argument");
  Group& g = h->impl;
  Stream* s = g.s;
  const uint32_t count = (uint32_t)g.members.size();
  // every member's ids are checked before anything is submitted: a member-local id past the
  // member's end would name another member's vertex once offset
  for (uint32_t i = 0; i < count; i++) {
    GS_CHECK(pairs[i] || nedges[i] == 0, GS_ERR_INVALID, "member %u: null pairs", i);
    const uint32_t n = g.members[i]->view.n;
    uint32_t bad = 0;
    for (uint64_t e = 0; e < 2 * nedges[i]; e++) bad |= pairs[i][e] >= n;
    if (bad)
      for (uint64_t e = 0; e < 2 * nedges[i]; e++)
        GS_CHECK(pairs[i][e] < n, GS_ERR_INVALID, "member %u edge %llu: vertex %u out of range", i,
                 (unsigned long long)(e / 2), pairs[i][e]);
  }
  const uint64_t chunk = kGroupStageEdges;
  uint32_t mi = 0;
  uint64_t ei = 0;  // next edge of member mi
  for (;;) {
    while (mi < count && ei >= nedges[mi]) {
      mi++;
      ei = 0;
    }
    if (mi >= count) break;
    int rc = ensure_stage(s, chunk * 2 * sizeof(uint32_t));
    if (rc) return rc;
    uint32_t* out = s->h_stage;
    uint64_t m = 0;
    while (m < chunk && mi < count) {
      if (ei >= nedges[mi]) {
        mi++;
        ei = 0;
        continue;
      }
      const uint32_t base = g.bases[mi];
      const uint64_t take = std::min(chunk - m, nedges[mi] - ei);
      const uint32_t* p = pairs[mi] + 2 * ei;
      for (uint64_t i = 0; i < take; i++) {
        out[2 * (m + i)] = p[2 * i] + base;
        out[2 * (m + i) + 1] = p[2 * i + 1] + base;
      }
      m += take;
      ei += take;
    }
    if (m == 0) break;
    // the staging is pinned: copied straight into the stream's fill buffer, read before return
    rc = submit_pairs_host(s, s->h_stage, m, 0, 0);
    if (rc) return rc;
  }
  return GS_OK;
}

int gs_group_submit_member_pairs(gs_group* h, uint32_t member, const uint32_t* pairs,
                                 uint64_t nedges) {
  GS_CHECK(h, GS_ERR_INVALID, "null group");
  Group& g = h->impl;
  GS_CHECK(member < g.members.size(), GS_ERR_INVALID, "member %u of a group of %zu", member,
           g.members.size());
  GS_CHECK(pairs || nedges == 0, GS_ERR_INVALID, "null pairs");
  MemberIds mid;
  mid.base = g.bases[member];
  mid.n = g.members[member]->view.n;
  return submit_pairs_host(g.s, pairs, nedges, 0, 0, mid);
}

int gs_group_submit_member_pairs_async(gs_group* h, uint32_t member, const uint32_t* pairs,
                                       uint64_t nedges) {
  GS_CHECK(h, GS_ERR_INVALID, "null group");
  Group& g = h->impl;
  GS_CHECK(member < g.members.size(), GS_ERR_INVALID, "member %u of a group of %zu", member,
           g.members.size());
  GS_CHECK(pairs || nedges == 0, GS_ERR_INVALID, "null pairs");
  MemberIds mid;
  mid.base = g.bases[member];
  mid.n = g.members[member]->view.n;
  return submit_pairs_host(g.s, pairs, nedges, 0, 0, mid, false);
}

int gs_group_submit_member_batches_async(gs_group* h, uint32_t count, const uint32_t* members,
                                         const uint32_t* const* pairs, const uint64_t* nedges) {
  GS_CHECK(h, GS_ERR_INVALID, "null group");
  Group& g = h->impl;
  GS_CHECK(count == 0 || (members && pairs && nedges), GS_ERR_INVALID, "null batch arrays");
  // every argument is checked before anything is enqueued
  std::vector<MemberIds> mids(count);
  for (uint32_t i = 0; i < count; i++) {
    GS_CHECK(members[i] < g.members.size(), GS_ERR_INVALID, "member %u of a group of %zu",
             members[i], g.members.size());
    GS_CHECK(pairs[i] || nedges[i] == 0, GS_ERR_INVALID, "null pairs (batch %u)", i);
    GS_CHECK(nedges[i] < (1ULL << 40), GS_ERR_INVALID, "batch %u too large", i);
    mids[i].base = g.bases[members[i]];
    mids[i].n = g.members[members[i]]->view.n;
  }
  return submit_member_batches_host(g.s, count, mids.data(), pairs, nedges);
}

int gs_stream_wait_copies(gs_stream* sh) {
  GS_CHECK(sh, GS_ERR_INVALID, "null stream");
  return wait_copies(&sh->impl);
}

int gs_group_submit_pairs_device(gs_group* h, const uint32_t* d_pairs, uint64_t nedges) {
  GS_CHECK(h, GS_ERR_INVALID, "null group");
  GS_CHECK(d_pairs || nedges == 0, GS_ERR_INVALID, "null pairs");
  return submit_pairs_device(h->impl.s, d_pairs, nedges, 0, 0);
}

int gs_group_prepare(gs_group* h) {
  GS_CHECK(h, GS_ERR_INVALID, "null group");
  Group& g = h->impl;
  int rc = prepare_stream(g.s);
  if (rc) return rc;
  if ((rc = ensure_stage(g.s, kGroupStageEdges * 2 * sizeof(uint32_t)))) return rc;
  if ((rc = prepare_group_query(&g))) return rc;
  GS_HIP(hipStreamSynchronize(g.s->stream));
  return GS_OK;
}

int gs_group_fence(gs_group* h) {
  GS_CHECK(h, GS_ERR_INVALID, "null group");
  Group& g = h->impl;
  int rc = drain_stream(g.s);  // launch what is pending
  if (rc) return rc;
  if (!g.ev_fence) GS_HIP(hipEventCreateWithFlags(&g.ev_fence, hipEventDisableTiming));
  GS_HIP(hipEventRecord(g.ev_fence, g.s->stream));
  g.fenced = true;
  return GS_OK;
}

int gs_group_query_cc(gs_group* h, const uint8_t* member_mask, uint32_t* labels,
                      uint32_t* num_cc, uint32_t* rounds, int32_t* status) {
  GS_CHECK(h && labels, GS_ERR_INVALID, "null argument");
  Group& g = h->impl;
  if (g.fenced) {
    // the group stream up to the fence (waited for on the device), every other writer of the
    // queried members now
    for (size_t i = 0; i < g.members.size(); i++) {
      if (member_mask && !member_mask[i]) continue;
      Store* m = g.members[i];
      std::vector<Stream*> ss;
      {
        std::lock_guard<std::mutex> lk(m->mu);
        ss = m->streams;
      }
      for (Stream* s : ss) {
        if (s == g.s) continue;
        int rc = drain_stream(s);
        if (rc) return rc;
        GS_HIP(hipStreamSynchronize(s->stream));
      }
      GS_HIP(hipStreamSynchronize(m->qstream));
    }
  } else {
    // every writer of every member must be done before the sketches are read
    for (Store* m : g.members) {
      int rc = sync_store(m);
      if (rc) return rc;
    }
  }
  const int rc = group_query_cc(&g, member_mask, labels, num_cc, rounds, status);
  g.fenced = false;
  return rc;
}

}  // extern "C"
