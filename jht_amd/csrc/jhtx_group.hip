// Ragged groups: several mixed-precision plans (different nside, lmax, batch; one spin)
// transformed with one kernel launch per stage.  The stage kernels live beside their
// one-plan forms (jhtx_mixed.hip recursion, jhtx_fft.hip ring FFT, jhtx_rec.hip chunk
// reduction); this file owns the group object, its workspace layout and the C ABI.
#include <algorithm>
#include <numeric>
#include <vector>

#include "jhtx_internal.h"

struct jhtx_group {
  jhtx::GroupLayout L;
  int order[jhtx::kGroupMax];  // launch slot -> index of the plan in the caller's arrays
  int device;
  void* rec_desc = nullptr;
  void* fft_desc = nullptr;
  void* fin_desc = nullptr;
  int start_synth[jhtx::kGroupMax + 1];
  int start_ring[jhtx::kGroupMax + 1];
  int start_adj[jhtx::kGroupMax + 1];
  int start_fin[jhtx::kGroupMax + 1];
  int fft_lds;
  size_t ws_bytes;
};

namespace {

size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

void release(jhtx_group* g) {
  if (!g) return;
  if (g->rec_desc) (void)hipFree(g->rec_desc);
  if (g->fft_desc) (void)hipFree(g->fft_desc);
  if (g->fin_desc) (void)hipFree(g->fin_desc);
  delete g;
}

// the call's per-plan buffers in launch order, and one stage's work-group ranges
jhtx::GroupCall make_call(const jhtx_group* g, const int* start, const void* const* in, void* const* out,
                          void* ws) {
  jhtx::GroupCall c;
  c.n = g->L.n;
  for (int i = 0; i <= jhtx::kGroupMax; ++i) c.start[i] = (i <= c.n) ? start[i] : start[c.n];
  for (int i = 0; i < jhtx::kGroupMax; ++i) {
    c.in[i] = (i < c.n) ? in[g->order[i]] : nullptr;
    c.out[i] = (i < c.n) ? out[g->order[i]] : nullptr;
  }
  c.ws = reinterpret_cast<char*>(ws);
  return c;
}

int check_group_call(const jhtx_group* g, const void* const* in, void* const* out, void* ws) {
  if (!g) {
    jhtx::set_error("null group");
    return JHTX_EINVAL;
  }
  if (!in || !out || !ws) {
    jhtx::set_error("null buffer pointer");
    return JHTX_EINVAL;
  }
  for (int i = 0; i < g->L.n; ++i)
    if (!in[i] || !out[i]) {
      jhtx::set_error("null buffer pointer for plan " + std::to_string(i));
      return JHTX_EINVAL;
    }
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev != g->device) {
    jhtx::set_error("the group's device (" + std::to_string(g->device) + ") is not current");
    return JHTX_EINVAL;
  }
  return JHTX_OK;
}

}  // namespace

extern "C" {

int jhtx_group_create(jhtx_plan* const* plans, const int* batch, int n, jhtx_group** group) {
  using namespace jhtx;
  if (!group || !plans || !batch) {
    set_error("null argument");
    return JHTX_EINVAL;
  }
  *group = nullptr;
  if (n < 1 || n > kGroupMax) {
    set_error("ragged group: " + std::to_string(n) + " plans, outside [1, " + std::to_string(kGroupMax) + "]");
    return JHTX_EINVAL;
  }
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess) {
    set_error("no current HIP device");
    return JHTX_EDEVICE;
  }
  const int spin = plans[0] ? plans[0]->spin : 0;
  for (int i = 0; i < n; ++i) {
    const jhtx_plan* p = plans[i];
    const std::string who = "ragged group: plan " + std::to_string(i);
    if (!p) {
      set_error(who + " is null");
      return JHTX_EINVAL;
    }
    if (p->ring_set) {
      set_error(who + " has no HEALPix geometry (a ring-set plan)");
      return JHTX_EINVAL;
    }
    if (p->precision != JHTX_PREC_MIXED) {
      set_error(who + " is not a mixed-precision plan");
      return JHTX_EINVAL;
    }
    if (p->spin != spin) {
      set_error(who + " has spin " + std::to_string(p->spin) + ", the group spin " + std::to_string(spin));
      return JHTX_EINVAL;
    }
    if (p->nside > kGroupNside) {
      set_error(who + ": nside " + std::to_string(p->nside) + " above " + std::to_string(kGroupNside) +
                " (the group's ring transforms are single 256-thread work-groups)");
      return JHTX_EUNSUPPORTED;
    }
    if (p->device != dev) {
      set_error(who + " lives on device " + std::to_string(p->device) + ", not the current one");
      return JHTX_EINVAL;
    }
    if (batch[i] < 1 || batch[i] > p->max_batch) {
      set_error(who + ": batch " + std::to_string(batch[i]) + " outside [1, " + std::to_string(p->max_batch) + "]");
      return JHTX_EINVAL;
    }
  }
  jhtx_group* g = new jhtx_group();
  g->device = dev;
  // launch order: decreasing lmax (the costliest work-groups first), stable
  std::vector<int> ord(n);
  std::iota(ord.begin(), ord.end(), 0);
  std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return plans[a]->lmax > plans[b]->lmax; });
  GroupLayout& L = g->L;
  L.n = n;
  L.spin = spin;
  size_t off = 0;
  for (int i = 0; i < n; ++i) {
    const jhtx_plan* p = plans[ord[i]];
    g->order[i] = ord[i];
    L.plan[i] = p;
    L.batch[i] = batch[ord[i]];
    const size_t nb = (size_t)L.batch[i] * p->ncomp;
    L.f_off[i] = (long long)off;
    off += align256(sizeof(float2) * nb * p->nrings * p->M);
    L.part_off[i] = (long long)off;
    off += align256(sizeof(float2) * (size_t)group_adj_chunks(p) * nb * p->K);
  }
  g->ws_bytes = off;
  int rc = group_rec_build(L, &g->rec_desc, g->start_synth, g->start_adj);
  if (rc == JHTX_OK) rc = group_fft_build(L, &g->fft_desc, g->start_ring, &g->fft_lds);
  if (rc == JHTX_OK) rc = group_fin_build(L, &g->fin_desc, g->start_fin);
  if (rc != JHTX_OK) {
    release(g);
    return rc;
  }
  *group = g;
  return JHTX_OK;
}

size_t jhtx_group_workspace_bytes(const jhtx_group* g) { return g ? g->ws_bytes : 0; }

int jhtx_group_synth(jhtx_group* g, const void* const* alm, void* const* map, void* ws, void* stream) {
  using namespace jhtx;
  int rc = check_group_call(g, alm, map, ws);
  if (rc) return rc;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  hipError_t e = group_rec_launch(true, g->L, g->rec_desc, make_call(g, g->start_synth, alm, map, ws), s);
  if (e == hipSuccess)
    e = group_fft_launch(true, g->L, g->fft_desc, make_call(g, g->start_ring, alm, map, ws), g->fft_lds, s);
  if (e != hipSuccess) {
    set_error(std::string("group synthesis launch failed: ") + hipGetErrorString(e));
    return JHTX_EDEVICE;
  }
  return JHTX_OK;
}

int jhtx_group_adjoint(jhtx_group* g, const void* const* map, void* const* alm, void* ws, void* stream) {
  using namespace jhtx;
  int rc = check_group_call(g, map, alm, ws);
  if (rc) return rc;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  hipError_t e = group_fft_launch(false, g->L, g->fft_desc, make_call(g, g->start_ring, map, alm, ws), g->fft_lds, s);
  if (e == hipSuccess) e = group_rec_launch(false, g->L, g->rec_desc, make_call(g, g->start_adj, map, alm, ws), s);
  if (e == hipSuccess) e = group_fin_launch(g->fin_desc, make_call(g, g->start_fin, map, alm, ws), s);
  if (e != hipSuccess) {
    set_error(std::string("group adjoint launch failed: ") + hipGetErrorString(e));
    return JHTX_EDEVICE;
  }
  return JHTX_OK;
}

long long jhtx_group_query(const jhtx_group* g, int what) {
  if (!g) return -1;
  switch (what) {
    case 0: return g->L.n;
    case 1: return g->L.spin;
    case 2: return 2;
    case 3: return 3;
    case 4: return g->start_synth[g->L.n];
    case 5: return g->start_ring[g->L.n];
    case 6: return g->start_ring[g->L.n];
    case 7: return g->start_adj[g->L.n];
    case 8: return g->start_fin[g->L.n];
    default: return -1;
  }
}

void jhtx_group_destroy(jhtx_group* g) { release(g); }

}  // extern "C"
