// spf_api.hip -- the extern "C" entry points of libspiffe_amd: context and
// run lifecycle, host<->device layout conversion, and the per-step entry
// points that replace the reference's hot-path functions.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "spf_internal.h"

using namespace spf;

namespace {

constexpr double PI_D = 3.141592653589793238462643;

// Allocation cleared on the context's stream: the run's later work on that
// (non-blocking) stream is ordered after the clear, which a null-stream
// hipMemset would not guarantee while other contexts keep the device busy.
template <typename T>
hipError_t dalloc(T **p, int64_t n, hipStream_t s) {
  if (n <= 0) n = 1;
  hipError_t e = hipMalloc((void **)p, sizeof(T) * (size_t)n);
  if (e != hipSuccess) return e;
  return hipMemsetAsync(*p, 0, sizeof(T) * (size_t)n, s);
}


int64_t ldz_for(int nz) { return ((int64_t)(nz + 1) + 31) / 32 * 32; }

void free_run(Run *R) {
  auto F = [](void *p) {
    if (p) (void)hipFree(p);
  };
  for (int k = 0; k < 6; k++)
    for (int j = 0; j < 2; j++) F(R->E2[k][j]);
  F(R->Bphi2); F(R->Br2); F(R->Bz2); F(R->Jz); F(R->Jr); F(R->Jphi); F(R->Psi); F(R->Q); for (int k = 0; k < 5; k++) F(R->PsiSpare[k]);
  F(R->EzImp); F(R->ErImp); F(R->PsiImp); F(R->BzImp); F(R->BrImp); F(R->flags); F(R->open); F(R->cut0); F(R->cut1);
  F(R->cp); F(R->cm); F(R->bzf); F(R->sor_c); F(R->sor_d); F(R->volume);
  for (int k = 0; k < 16; k++) F(R->P[k]);
  F(R->pid); F(R->gen); F(R->status);
  for (int k = 0; k < 16; k++) F(R->P2[k]);
  F(R->pid2); F(R->gen2); F(R->status2); F(R->wg_count); F(R->wg_off); F(R->counters); F(R->krange); F(R->evt); F(R->evflag);
  for (int k = 0; k < 7; k++) F(R->ev[k]);
  F(R->last_old); F(R->evlist); F(R->move_src); F(R->move_dst); F(R->srows); F(R->lrows);
  F(R->keys); F(R->skeys); F(R->vals); F(R->cz); F(R->scz); F(R->long_list); F(R->pcount); F(R->evoff); F(R->tcnt); F(R->toff); F(R->cbase); F(R->gflag); F(R->sort_tmp);
  F(R->stage); F(R->append_stage); F(R->emit_stage); F(R->Pdev); F(R->omega); F(R->red);
  F(R->red_counter); F(R->sor_state); F(R->sor_sums); F(R->sor_masks);
  R->live = false;
}

int64_t default_srow_cap(const spf_run_desc *d) {
  if (d->screen_row_capacity > 0) return d->screen_row_capacity;
  return d->n_screens > 0 ? d->n_screens * d->max_particles + 1024 : 1;
}
int64_t default_lrow_cap(const spf_run_desc *d) {
  return d->lost_row_capacity > 0 ? d->lost_row_capacity : d->max_particles + 1024;
}

}  // namespace

PhaseTimer::PhaseTimer(spf_ctx *ctx, int timer_id, int64_t launches) : c(ctx), id(timer_id), n(launches) {
  if (!c || c->timing_every <= 0 || n <= 0) return;
  c->t_launched[id] += n;
  if (c->t_calls[id]++ % c->timing_every) return;
  if (!c->ev_free.empty()) {
    e0 = c->ev_free.back();
    c->ev_free.pop_back();
  } else if (hipEventCreate(&e0) != hipSuccess) {
    e0 = nullptr;
    return;
  }
  if (hipEventRecord(e0, c->stream) != hipSuccess) {
    c->ev_free.push_back(e0);
    e0 = nullptr;
  }
}

PhaseTimer::~PhaseTimer() {
  if (!e0) return;
  hipEvent_t e1 = nullptr;
  if (!c->ev_free.empty()) {
    e1 = c->ev_free.back();
    c->ev_free.pop_back();
  } else if (hipEventCreate(&e1) != hipSuccess) {
    c->ev_free.push_back(e0);
    return;
  }
  if (hipEventRecord(e1, c->stream) != hipSuccess) {
    c->ev_free.push_back(e0);
    c->ev_free.push_back(e1);
    return;
  }
  c->ev_pending.push_back({id, n, e0, e1});
  if (c->ev_pending.size() >= 4096) resolve_timers(c, false);
}

// adds the elapsed times of completed event pairs, oldest first; with wait,
// the stream has been synchronized and every pair is complete
void spf::resolve_timers(spf_ctx *c, bool wait) {
  size_t k = 0;
  for (; k < c->ev_pending.size(); k++) {
    spf_ctx::Pending &p = c->ev_pending[k];
    if (!wait && hipEventQuery(p.e1) != hipSuccess) break;
    float ms = 0;
    if (hipEventElapsedTime(&ms, p.e0, p.e1) == hipSuccess) {
      c->t_ms[p.id] += ms;
      c->t_count[p.id] += p.n;
    }
    c->ev_free.push_back(p.e0);
    c->ev_free.push_back(p.e1);
  }
  c->ev_pending.erase(c->ev_pending.begin(), c->ev_pending.begin() + k);
}

int spf::batch_table(spf_ctx *c, size_t bytes, void **host, void **dev) {
  if (bytes > c->bt_slot) {
    HIPCALL(c, hipStreamSynchronize(c->stream));
    if (c->bt_host) HIPCALL(c, hipHostFree(c->bt_host));
    if (c->bt_dev) HIPCALL(c, hipFree(c->bt_dev));
    c->bt_host = c->bt_dev = nullptr;
    c->bt_slot = 0;
    const size_t slot = (std::max<size_t>(bytes, 65536) + 4095) / 4096 * 4096;
    HIPCALL(c, hipHostMalloc((void **)&c->bt_host, slot * spf_ctx::kTableSlots, hipHostMallocDefault));
    HIPCALL(c, hipMalloc((void **)&c->bt_dev, slot * spf_ctx::kTableSlots));
    c->bt_slot = slot;
    for (int i = 0; i < spf_ctx::kTableSlots; i++) c->bt_used[i] = false;
  }
  const int i = c->bt_next;
  c->bt_next = (i + 1) % spf_ctx::kTableSlots;
  if (!c->bt_ev[i]) HIPCALL(c, hipEventCreateWithFlags(&c->bt_ev[i], hipEventDisableTiming));
  if (c->bt_used[i]) HIPCALL(c, hipEventSynchronize(c->bt_ev[i]));   // its last copy has been made
  c->bt_cur = i;
  *host = c->bt_host + (size_t)i * c->bt_slot;
  *dev = c->bt_dev + (size_t)i * c->bt_slot;
  return SPF_OK;
}

int spf::deposit_scratch(spf_ctx *c, int64_t pairs, int64_t runs) {
  if (pairs <= c->dep_cap && runs <= c->dep_runs && c->dep_misc) return SPF_OK;
  // grown only between uses: the stream's queued work may still read the old buffers
  HIPCALL(c, hipStreamSynchronize(c->stream));
  for (void *p : {(void *)c->dep_k0, (void *)c->dep_k1, (void *)c->dep_v0, (void *)c->dep_v1, c->dep_tmp,
                  (void *)c->dep_long, (void *)c->dep_krange, (void *)c->dep_misc, (void *)c->dep_run_off})
    if (p) HIPCALL(c, hipFree(p));
  c->dep_k0 = c->dep_k1 = nullptr;
  c->dep_v0 = c->dep_v1 = nullptr;
  c->dep_tmp = nullptr;
  c->dep_long = nullptr;
  c->dep_krange = c->dep_misc = c->dep_run_off = nullptr;
  c->dep_cap = c->dep_runs = 0;
  const int64_t n = std::max<int64_t>(pairs, 1), r = std::max<int64_t>(runs, 1);
  c->dep_tmp_bytes = radix_sort_tmp_bytes(n);
  HIPCALL(c, hipMalloc((void **)&c->dep_k0, n * sizeof(uint32_t)));
  HIPCALL(c, hipMalloc((void **)&c->dep_k1, n * sizeof(uint32_t)));
  HIPCALL(c, hipMalloc((void **)&c->dep_v0, n * sizeof(double)));
  HIPCALL(c, hipMalloc((void **)&c->dep_v1, n * sizeof(double)));
  HIPCALL(c, hipMalloc(&c->dep_tmp, c->dep_tmp_bytes));
  HIPCALL(c, hipMalloc((void **)&c->dep_long, (n / 256 + 1) * sizeof(int32_t)));
  HIPCALL(c, hipMalloc((void **)&c->dep_krange, (2 * 256 + 2) * sizeof(int64_t)));
  HIPCALL(c, hipMalloc((void **)&c->dep_misc, 16 * sizeof(int64_t)));
  HIPCALL(c, hipMalloc((void **)&c->dep_run_off, (r + 1) * sizeof(int64_t)));
  c->dep_cap = n;
  c->dep_runs = r;
  return SPF_OK;
}

hipError_t spf::batch_table_upload(spf_ctx *c, const void *host, void *dev, size_t bytes) {
  hipError_t e = hipMemcpyAsync(dev, host, bytes, hipMemcpyHostToDevice, c->stream);
  if (e == hipSuccess) e = hipEventRecord(c->bt_ev[c->bt_cur], c->stream);
  if (e == hipSuccess) c->bt_used[c->bt_cur] = true;
  return e;
}

int spf::batch_flags(spf_ctx *c, size_t n) {
  if (n <= c->bw_flags_n) return SPF_OK;
  HIPCALL(c, hipStreamSynchronize(c->stream));
  if (c->bw_flags) HIPCALL(c, hipHostFree(c->bw_flags));
  c->bw_flags = nullptr;
  c->bw_flags_n = 0;
  HIPCALL(c, hipHostMalloc((void **)&c->bw_flags, std::max<size_t>(n, 256) * sizeof(int64_t), hipHostMallocDefault));
  c->bw_flags_n = std::max<size_t>(n, 256);
  return SPF_OK;
}

extern "C" {

int spf_timing_enable(spf_ctx *c, int32_t enable) {
  if (!c) return SPF_EINVAL;
  HIPCALL(c, hipStreamSynchronize(c->stream));
  resolve_timers(c, true);
  for (int k = 0; k < SPF_T_N; k++) {
    c->t_ms[k] = 0;
    c->t_count[k] = 0;
    c->t_calls[k] = 0;
    c->t_launched[k] = 0;
  }
  c->timing_every = enable > 0 ? enable : 0;
  return SPF_OK;
}

int spf_timing_read(spf_ctx *c, double *ms, int64_t *timed, int64_t *launched) {
  if (!c || !ms || !timed) return SPF_EINVAL;
  HIPCALL(c, hipStreamSynchronize(c->stream));
  resolve_timers(c, true);
  for (int k = 0; k < SPF_T_N; k++) {
    ms[k] = c->t_ms[k];
    timed[k] = c->t_count[k];
    if (launched) launched[k] = c->t_launched[k];
  }
  return SPF_OK;
}

int spf_ctx_create(int32_t device, spf_ctx **out) {
  if (!out) return SPF_EINVAL;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return SPF_EHIP;
  if (device < 0 || device >= n) return SPF_EINVAL;
  spf_ctx *c = new (std::nothrow) spf_ctx();
  if (!c) return SPF_ENOMEM;
  c->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipHostMalloc((void **)&c->pinned, 64 * sizeof(int64_t), hipHostMallocDefault) != hipSuccess ||
      hipEventCreateWithFlags(&c->stage_done, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->sor_ev[0], hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->sor_ev[1], hipEventDisableTiming) != hipSuccess ||
      noalloc_setup(device) != SPF_OK) {
    delete c;
    return SPF_EHIP;
  }
  *out = c;
  return SPF_OK;
}

int spf_ctx_destroy(spf_ctx *c) {
  if (!c) return SPF_EINVAL;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  for (Run *R : c->runs)
    if (R) {
      if (R->live) free_run(R);
      delete R;
    }
  if (c->pinned) (void)hipHostFree(c->pinned);
  if (c->bt_host) (void)hipHostFree(c->bt_host);
  if (c->bt_dev) (void)hipFree(c->bt_dev);
  for (hipEvent_t e : c->bt_ev)
    if (e) (void)hipEventDestroy(e);
  if (c->bw_flags) (void)hipHostFree(c->bw_flags);
  for (void *p : {(void *)c->dep_k0, (void *)c->dep_k1, (void *)c->dep_v0, (void *)c->dep_v1, c->dep_tmp,
                  (void *)c->dep_long, (void *)c->dep_krange, (void *)c->dep_misc, (void *)c->dep_run_off})
    if (p) (void)hipFree(p);
  if (c->hstage) (void)hipHostFree(c->hstage);
  if (c->stage_done) (void)hipEventDestroy(c->stage_done);
  for (hipEvent_t e : c->sor_ev)
    if (e) (void)hipEventDestroy(e);
  for (auto &p : c->ev_pending) {
    (void)hipEventDestroy(p.e0);
    (void)hipEventDestroy(p.e1);
  }
  for (hipEvent_t e : c->ev_free) (void)hipEventDestroy(e);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return SPF_OK;
}

const char *spf_last_error(spf_ctx *c) { return c ? c->err.c_str() : "null context"; }

int spf_synchronize(spf_ctx *c) {
  if (!c) return SPF_EINVAL;
  HIPCALL(c, hipStreamSynchronize(c->stream));
  return SPF_OK;
}

void *spf_stream(spf_ctx *c) { return c ? (void *)c->stream : nullptr; }

int64_t spf_run_footprint(const spf_run_desc *d) {
  if (!d || d->nz < 3 || d->nr < 3) return -1;
  int64_t cells = ldz_for(d->nz) * (int64_t)(d->nr + 1);
  int64_t field_bytes = cells * 8 * 33 + cells * 3;                      // 33 double grids, flags, open cells
  int64_t n = d->max_particles > 0 ? d->max_particles : 1;
  int64_t part_bytes = n * (16 * 8 + 4 + 2 + 4);                         // particle SoA
  int64_t event_bytes = n * (4 + 1 + 7 * 8 + 4 + 8) + (n / 256 + 2) * 4 +
                        (int64_t)(kMaxScreens + 1) * rows_tiles(n) * 8;   // staging, lists, rows pipeline
  // keys, contributions (both sorted), flags, sort tables: three components, four slots each when smeared
  int64_t sort_bytes = 3 * spf::deposit_slots_per_particle(d->modes) * n * (4 + 4 + 8 + 8 + 8 + 1 + 1);
  int64_t ring = default_srow_cap(d) * (int64_t)sizeof(spf_screen_row) + default_lrow_cap(d) * (int64_t)sizeof(spf_lost_row);
  // cell-sort runs: a second particle set and the reduce deposition's counts
  int64_t cellsort_bytes = d->cell_sort > 0 ? n * (16 * 8 + 4 + 2 + 4) + (n / 1024 + 2) * 12 : 0;
  return field_bytes + part_bytes + event_bytes + sort_bytes + ring + cellsort_bytes;
}

int spf_run_create(spf_ctx *c, const spf_run_desc *d, int32_t *run_out) {
  if (!c || !d || !run_out) return fail(c, SPF_EINVAL, "null argument");
  if (d->nz < 3 || d->nr < 3) return fail(c, SPF_EINVAL, "nz and nr must be >= 3");
  if (!(d->dz > 0) || !(d->dr > 0) || !(d->dt > 0)) return fail(c, SPF_EINVAL, "dz, dr and dt must be > 0");
  if (!d->metal_flags) return fail(c, SPF_EINVAL, "metal_flags is required");
  if (d->n_screens < 0 || d->n_screens > kMaxScreens) return fail(c, SPF_EINVAL, "n_screens must be in [0, 24]");
  if (d->n_screens > 0 && (!d->screen_z || !d->screen_direction)) return fail(c, SPF_EINVAL, "screen arrays missing");
  if (d->max_particles < 0 || d->max_particles > 0x20000000LL) return fail(c, SPF_EINVAL, "max_particles out of range");
  if ((int64_t)(d->nz + 1) * (d->nr + 1) >= (1LL << 29)) return fail(c, SPF_EINVAL, "grid too large");
  if (d->cell_sort < 0) return fail(c, SPF_EINVAL, "cell_sort must be >= 0");
  if (d->cell_sort > 0 && d->compat)
    return fail(c, SPF_EINVAL, "cell_sort needs compat = 0 (the reference's compaction defects depend on index order)");
  HIPCALL(c, hipSetDevice(c->device));
  Run *R = new (std::nothrow) Run();
  if (!R) return fail(c, SPF_ENOMEM, "host allocation failed");
  R->d = *d;
  R->d.metal_flags = nullptr;
  R->d.screen_z = nullptr;
  R->d.screen_direction = nullptr;
  R->screen_z.assign(d->screen_z, d->screen_z + d->n_screens);
  R->screen_dir.assign(d->screen_direction, d->screen_direction + d->n_screens);
  const int nz = d->nz, nr = d->nr;
  R->ldz = (int)ldz_for(nz);
  R->cells = (int64_t)R->ldz * (nr + 1);
  R->cap = d->max_particles;
  R->live = true;
  auto bail = [&](hipError_t e, const char *what) {
    free_run(R);
    delete R;
    return fail(c, e == hipErrorOutOfMemory ? SPF_ENOMEM : SPF_EHIP, std::string(what) + ": " + hipGetErrorString(e));
  };
#define ALLOC(ptr, n)                                        \
  do {                                                       \
    hipError_t e_ = dalloc(&(ptr), (n), c->stream);          \
    if (e_ != hipSuccess) return bail(e_, "alloc " #ptr);    \
  } while (0)
  const int64_t cells = R->cells;
  for (int k = 0; k < 6; k++)
    for (int j = 0; j < 2; j++) ALLOC(R->E2[k][j], cells);
  ALLOC(R->Bphi2, cells); ALLOC(R->Br2, cells); ALLOC(R->Bz2, cells);
  ALLOC(R->Jz, cells); ALLOC(R->Jr, cells); ALLOC(R->Jphi, cells);
  ALLOC(R->Psi, cells); ALLOC(R->Q, cells); for (int k = 0; k < 5; k++) ALLOC(R->PsiSpare[k], cells);
  ALLOC(R->EzImp, cells); ALLOC(R->ErImp, cells); ALLOC(R->PsiImp, cells);
  ALLOC(R->BzImp, cells); ALLOC(R->BrImp, cells);
  ALLOC(R->flags, cells);
  ALLOC(R->open, cells);
  ALLOC(R->stage, (int64_t)(nz + 1) * (nr + 1));
  ALLOC(R->cut0, nr + 1); ALLOC(R->cut1, nr + 1);
  ALLOC(R->cp, nr + 1); ALLOC(R->cm, nr + 1); ALLOC(R->bzf, nr + 1);
  ALLOC(R->sor_c, nr + 1); ALLOC(R->sor_d, nr + 1); ALLOC(R->volume, 4 * (nr + 1));
  const int64_t cap = R->cap > 0 ? R->cap : 1;
  for (int k = 0; k < 16; k++) ALLOC(R->P[k], cap);
  ALLOC(R->pid, cap); ALLOC(R->gen, cap); ALLOC(R->status, cap);
  R->cell_sort = d->cell_sort;
  if (R->cell_sort > 0) {
    for (int k = 0; k < 16; k++) ALLOC(R->P2[k], cap);
    ALLOC(R->pid2, cap); ALLOC(R->gen2, cap); ALLOC(R->status2, cap);
    ALLOC(R->wg_count, reduce_groups(cap) + 1); ALLOC(R->wg_off, reduce_groups(cap) + 1);
  }
  ALLOC(R->counters, 16);
  ALLOC(R->krange, 2 * 256 + 2);
  ALLOC(R->evt, cap); ALLOC(R->evflag, cap); ALLOC(R->pcount, cap / 256 + 2);
  ALLOC(R->evoff, cap / 256 + 2);
  ALLOC(R->tcnt, (int64_t)(kMaxScreens + 1) * rows_tiles(cap)); ALLOC(R->toff, (int64_t)(kMaxScreens + 1) * rows_tiles(cap));
  ALLOC(R->cbase, kMaxScreens + 4);
  for (int k = 0; k < 7; k++) ALLOC(R->ev[k], cap);
  ALLOC(R->last_old, 4); ALLOC(R->evlist, cap); ALLOC(R->move_src, cap); ALLOC(R->move_dst, cap);
  R->srow_cap = default_srow_cap(d);
  R->lrow_cap = default_lrow_cap(d);
  ALLOC(R->srows, R->srow_cap); ALLOC(R->lrows, R->lrow_cap);
  // ordered scatter-add: three components (Jz, Jr, Jphi) per particle, four
  // key slots each with area-weighted deposition
  const int64_t cap3 = 3 * deposit_slots_per_particle(d->modes) * cap;
  ALLOC(R->keys, cap3); ALLOC(R->skeys, cap3); ALLOC(R->cz, cap3); ALLOC(R->scz, cap3);
  ALLOC(R->vals, cap3); ALLOC(R->long_list, cap3 / 256 + 1); ALLOC(R->gflag, cap3);
  R->sort_tmp_bytes = radix_sort_tmp_bytes(cap3);
  {
    uint8_t *t = nullptr;
    ALLOC(t, (int64_t)R->sort_tmp_bytes);
    R->sort_tmp = t;
  }
  R->sor_shape = 0;   // the small shape needs the most tiles, mask words and partial rows
  R->red_n = std::max<int64_t>(sor_tiles(*R) * sor_batch(*R) * 4 * 2, 1024) + 16;   // two check groups of partials, or sum2 partials
  ALLOC(R->red, R->red_n);
  ALLOC(R->red_counter, 4); ALLOC(R->sor_state, 4); ALLOC(R->sor_sums, 16);
  ALLOC(R->sor_masks, sor_mask_words(*R));
  R->sor_shape = sor_shape_for((int64_t)(nz + 1) * (nr + 1));
  ALLOC(R->Pdev, 16);
#undef ALLOC
  hipStream_t s = c->stream;
  // flags to the device layout, and per-row tables computed from them
  std::vector<int16_t> hf(d->metal_flags, d->metal_flags + (size_t)(nz + 1) * (nr + 1));
  {
    hipError_t e = hipMemcpyAsync(R->stage, hf.data(), hf.size() * sizeof(int16_t), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = launch_transpose_in_i16((const int16_t *)R->stage, R->flags, nz, nr, R->ldz, s);
    if (e == hipSuccess) e = launch_open_cells(R->flags, R->ldz, nz, nr, R->open, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return bail(e, "flags upload");
  }
  // loop cut-offs: the reference breaks out of a row loop right after the
  // first point flagged FL_LAST_ACTIVE_PT (src/integrate.c:345-346 etc.)
  std::vector<int32_t> c0(nr + 1, nz - 1), c1(nr + 1, nz - 1);
  for (int ir = 0; ir <= nr; ir++) {
    for (int iz = 0; iz < nz; iz++)
      if (hf[(size_t)iz * (nr + 1) + ir] & SPF_FL_LAST_ACTIVE_PT) {
        c0[ir] = iz;
        break;
      }
    for (int iz = 1; iz < nz; iz++)
      if (hf[(size_t)iz * (nr + 1) + ir] & SPF_FL_LAST_ACTIVE_PT) {
        c1[ir] = iz;
        break;
      }
  }
  // coefficient tables with the reference's expressions, src/integrate.c:324-340, :515
  const double eps0 = d->epsilon_o, mu0 = d->mu_o, dr = d->dr, dti = d->dt;
  std::vector<double> cp(nr + 1, 0), cm(nr + 1, 0), bzf(nr + 1, 0);
  {
    double factor = 1 / (eps0 * mu0);
    for (int ir = 0; ir < nr; ir++) {
      double rp = (ir + 0.5) * dr, rm = (ir - 0.5) * dr, p;
      if (ir != 0) {
        p = 2 * factor / (rp * rp - rm * rm);
        cm[ir] = rm * p;
        cp[ir] = rp * p;
      } else {
        cp[ir] = 4 * factor / dr;
        cm[ir] = 0;
      }
    }
    for (int ir = 1; ir < nr; ir++) {
      long long a1 = (long long)(ir + 1) * (ir + 1), a0 = (long long)ir * ir;
      bzf[ir] = -2 * dti / (dr * (double)(a1 - a0));
    }
  }
  // SOR stencil rows, src/poisson.c:58-66; annulus areas, src/poisson_correction.c:141-154
  std::vector<double> sc(nr + 1, 0), sd(nr + 1, 0), vol(4 * (nr + 1), 0);
  {
    for (int ir = 0; ir <= nr; ir++) {
      if (ir != 0) {
        sc[ir] = 1 / (dr * dr) * (1 + 1. / (2 * ir));
        sd[ir] = 1 / (dr * dr) * (1 - 1. / (2 * ir));
      } else
        sc[ir] = sd[ir] = 1 / (dr * dr);
    }
    double dz = d->dz, dr2 = dr * dr, drdz = dr * dz;
    for (int ir = 0; ir <= nr; ir++) {
      double Ar_in, Ar_out, Az;
      if (ir == 0) {
        Ar_in = Ar_out = PI_D * drdz;
        Az = dr2 / 4 * PI_D;
      } else {
        Ar_in = PI_D * (2 * ir - 1) * drdz;
        Ar_out = PI_D * (2 * ir + 1) * drdz;
        long long q1 = (long long)(2 * ir + 1) * (2 * ir + 1), q0 = (long long)(2 * ir - 1) * (2 * ir - 1);
        Az = PI_D * dr2 / 4 * (double)(q1 - q0);
      }
      vol[ir] = Az * dz;
      vol[(nr + 1) + ir] = Ar_in;
      vol[2 * (nr + 1) + ir] = Ar_out;
      vol[3 * (nr + 1) + ir] = Az;
    }
  }
  {
    hipError_t e = hipSuccess;
    auto up = [&](void *dst, const void *src, size_t bytes) {
      if (e == hipSuccess) e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s);
    };
    up(R->cut0, c0.data(), c0.size() * 4);
    up(R->cut1, c1.data(), c1.size() * 4);
    up(R->cp, cp.data(), cp.size() * 8);
    up(R->cm, cm.data(), cm.size() * 8);
    up(R->bzf, bzf.data(), bzf.size() * 8);
    up(R->sor_c, sc.data(), sc.size() * 8);
    R->sor_c_host0 = sc[0];
    up(R->sor_d, sd.data(), sd.size() * 8);
    up(R->volume, vol.data(), vol.size() * 8);
    up(R->Pdev, R->P, sizeof(R->P));
    if (e == hipSuccess) e = hipStreamSynchronize(s);   // the host tables go out of scope below
    if (e != hipSuccess) return bail(e, "table upload");
  }
  *run_out = put_run(c, R);
  return SPF_OK;
}

int spf_run_destroy(spf_ctx *c, int32_t run) {
  Run *R = get_run(c, run);
  if (!R) return fail(c, SPF_EINVAL, "bad run id");
  HIPCALL(c, hipStreamSynchronize(c->stream));
  clear_run(c, run);
  free_run(R);
  delete R;
  return SPF_OK;
}

int spf_run_move(spf_ctx *from, int32_t run, spf_ctx *to, int32_t *run_out) {
  Run *R = get_run(from, run);
  if (!R || !to || !run_out) return fail(from, SPF_EINVAL, "bad run id or null argument");
  if (to->device != from->device) return fail(from, SPF_EINVAL, "a run moves only between contexts of one device");
  HIPCALL(from, hipStreamSynchronize(from->stream));   // the run's queued work is done
  clear_run(from, run);
  *run_out = put_run(to, R);
  return SPF_OK;
}

static double *field_ptr(Run *R, int32_t field) {
  switch (field) {
    case SPF_EZ: return R->E2[0][R->cur[0]];
    case SPF_ER: return R->E2[1][R->cur[1]];
    case SPF_BPHI: return R->E2[2][R->cur[2]];
    case SPF_EPHI: return R->E2[3][R->cur[3]];
    case SPF_BR: return R->E2[4][R->cur[4]];
    case SPF_BZ: return R->E2[5][R->cur[5]];
    case SPF_BPHI2: return R->Bphi2;
    case SPF_BR2: return R->Br2;
    case SPF_BZ2: return R->Bz2;
    case SPF_JZ: return R->Jz;
    case SPF_JR: return R->Jr;
    case SPF_JPHI: return R->Jphi;
    case SPF_PSI: return R->Psi;
    case SPF_Q: return R->Q;
    case SPF_EZ_IMPOSED: return R->EzImp;
    case SPF_ER_IMPOSED: return R->ErImp;
    case SPF_PSI_IMPOSED: return R->PsiImp;
    case SPF_BZ_IMPOSED: return R->BzImp;
    case SPF_BR_IMPOSED: return R->BrImp;
    default: return nullptr;
  }
}

int spf_field_upload(spf_ctx *c, int32_t run, int32_t field, const double *host) {
  Run *R = get_run(c, run);
  if (!R || !host) return fail(c, SPF_EINVAL, "bad run or null host pointer");
  double *dst = field_ptr(R, field);
  if (!dst) return fail(c, SPF_EINVAL, "unknown field id");
  size_t n = (size_t)(R->d.nz + 1) * (R->d.nr + 1);
  // every copy is ordered on the run's stream (never the null stream, which
  // the context's non-blocking stream does not wait for)
  HIPCALL(c, hipMemcpyAsync(R->stage, host, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPCALL(c, launch_transpose_in(R->stage, dst, R->d.nz, R->d.nr, R->ldz, c->stream));
  HIPCALL(c, hipStreamSynchronize(c->stream));
  if (field == SPF_EZ_IMPOSED || field == SPF_ER_IMPOSED) R->imposed_e = true;
  return SPF_OK;
}

int spf_field_download(spf_ctx *c, int32_t run, int32_t field, double *host) {
  Run *R = get_run(c, run);
  if (!R || !host) return fail(c, SPF_EINVAL, "bad run or null host pointer");
  double *src = field_ptr(R, field);
  if (!src) return fail(c, SPF_EINVAL, "unknown field id");
  size_t n = (size_t)(R->d.nz + 1) * (R->d.nr + 1);
  HIPCALL(c, launch_transpose_out(src, R->stage, R->d.nz, R->d.nr, R->ldz, c->stream));
  HIPCALL(c, hipMemcpyAsync(host, R->stage, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCALL(c, hipStreamSynchronize(c->stream));
  return SPF_OK;
}

int spf_set_step(spf_ctx *c, int32_t run, int64_t time_step) {
  Run *R = get_run(c, run);
  if (!R) return fail(c, SPF_EINVAL, "bad run id");
  R->time_step = time_step;
  return SPF_OK;
}

int spf_set_j_filter(spf_ctx *c, int32_t run, double multiplier) {
  Run *R = get_run(c, run);
  if (!R) return fail(c, SPF_EINVAL, "bad run id");
  if (!(multiplier >= 0 && multiplier < 1)) return fail(c, SPF_EINVAL, "J_filter_multiplier must be >=0 and < 1");
  R->jfm = multiplier;
  R->jm1 = 1;
  R->jm2 = 0;
  if (multiplier != 0) {   // src/integrate.c:110-116
    R->jm2 = 1 - multiplier;
    R->jm1 = multiplier / R->jm2;
  }
  return SPF_OK;
}

int spf_set_z_forces_start(spf_ctx *c, int32_t run, double z_forces_start) {
  Run *R = get_run(c, run);
  if (!R) return fail(c, SPF_EINVAL, "bad run id");
  R->z_forces_start = z_forces_start;
  return SPF_OK;
}

int spf_get_step(spf_ctx *c, int32_t run, int64_t *time_step, double *time) {
  Run *R = get_run(c, run);
  if (!R) return fail(c, SPF_EINVAL, "bad run id");
  if (time_step) *time_step = R->time_step;
  if (time) *time = R->time_step * R->d.dt + R->d.start_time;
  return SPF_OK;
}

}  // extern "C"

void spf::fdtd_args(Run *R, FdtdArgs &a) {
  const spf_run_desc &d = R->d;
  a.nz = d.nz;
  a.nr = d.nr;
  a.ldz = R->ldz;
  a.tm = (d.modes & SPF_TM_FIELDS) != 0;
  a.te = (d.modes & SPF_TE_FIELDS) != 0;
  a.left_neu = d.left_bc == SPF_NEUMANN;
  a.right_neu = d.right_bc == SPF_NEUMANN;
  a.upper_neu = d.upper_bc == SPF_NEUMANN;
  a.dti = d.dt;
  a.eps0 = d.epsilon_o;
  a.er_factor = 1 / (d.mu_o * d.epsilon_o * d.dz);
  a.ephi_factor = 1. / (d.epsilon_o * d.mu_o);
  a.top_factor = (d.rmax - d.dr / 2) / (d.rmax + d.dr / 2);
  a.dz = d.dz;
  a.dr = d.dr;
  a.Ez = R->E2[0][R->cur[0]];
  a.Er = R->E2[1][R->cur[1]];
  a.Bphi = R->E2[2][R->cur[2]];
  a.Ephi = R->E2[3][R->cur[3]];
  a.Br = R->E2[4][R->cur[4]];
  a.Bz = R->E2[5][R->cur[5]];
  a.Ezn = R->E2[0][1 - R->cur[0]];
  a.Ern = R->E2[1][1 - R->cur[1]];
  a.Bphin = R->E2[2][1 - R->cur[2]];
  a.Ephin = R->E2[3][1 - R->cur[3]];
  a.Brn = R->E2[4][1 - R->cur[4]];
  a.Bzn = R->E2[5][1 - R->cur[5]];
  a.Bphi2 = R->Bphi2;
  a.Br2 = R->Br2;
  a.Bz2 = R->Bz2;
  a.Jz = R->Jz;
  a.Jr = R->Jr;
  a.Jphi = R->Jphi;
  a.flags = R->flags;
  a.cut0 = R->cut0;
  a.cut1 = R->cut1;
  a.cp = R->cp;
  a.cm = R->cm;
  a.bzf = R->bzf;
}

void spf::fdtd_flip(Run *R) {
  if (R->d.modes & SPF_TM_FIELDS)
    for (int k = 0; k < 3; k++) R->cur[k] = 1 - R->cur[k];
  if (R->d.modes & SPF_TE_FIELDS)
    for (int k = 3; k < 6; k++) R->cur[k] = 1 - R->cur[k];
}

extern "C" {

int spf_fields_advance(spf_ctx *c, int32_t run) {
  Run *R = get_run(c, run);
  if (!R) return fail(c, SPF_EINVAL, "bad run id");
  FdtdArgs a;
  fdtd_args(R, a);
  if (!a.tm && !a.te) return SPF_OK;
  PhaseTimer pt(c, SPF_T_FDTD);
  HIPCALL(c, launch_fdtd(a, c->stream));
  fdtd_flip(R);
  return SPF_OK;
}

int spf_currents_reset(spf_ctx *c, int32_t run) {
  Run *R = get_run(c, run);
  if (!R) return fail(c, SPF_EINVAL, "bad run id");
  const bool tm = R->d.modes & SPF_TM_FIELDS, te = R->d.modes & SPF_TE_FIELDS;
  HIPCALL(c, launch_fill3(tm ? R->Jz : nullptr, tm ? R->Jr : nullptr, te ? R->Jphi : nullptr, R->cells, 0.0,
                          c->stream));
  return SPF_OK;
}

int spf_currents_scale(spf_ctx *c, int32_t run, double factor) {
  Run *R = get_run(c, run);
  if (!R) return fail(c, SPF_EINVAL, "bad run id");
  if (R->d.modes & SPF_TM_FIELDS) {
    HIPCALL(c, launch_scale(R->Jz, R->cells, factor, c->stream));
    HIPCALL(c, launch_scale(R->Jr, R->cells, factor, c->stream));
  }
  if (R->d.modes & SPF_TE_FIELDS) HIPCALL(c, launch_scale(R->Jphi, R->cells, factor, c->stream));
  return SPF_OK;
}

int spf_currents_add_line(spf_ctx *c, int32_t run, int32_t direction, int64_t iposition, int64_t istart,
                          int64_t iend, double J) {
  Run *R = get_run(c, run);
  if (!R) return fail(c, SPF_EINVAL, "bad run id");
  const int nz = R->d.nz, nr = R->d.nr;
  if (direction < 0 || direction > 2) return fail(c, SPF_EINVAL, "direction must be 0, 1 or 2");
  if (iend < istart) return SPF_OK;
  bool ok;
  if (direction == 1)
    ok = iposition >= 0 && iposition <= nz && istart >= 0 && iend <= nr;
  else
    ok = iposition >= 0 && iposition <= nr && istart >= 0 && iend <= nz;
  if (!ok) return fail(c, SPF_EINVAL, "line current indices outside the grid");
  double *Jarr = direction == 0 ? R->Jz : direction == 1 ? R->Jr : R->Jphi;
  HIPCALL(c, launch_add_line(Jarr, R->ldz, direction, iposition, istart, iend, J, c->stream));
  return SPF_OK;
}

int spf_currents_deposit(spf_ctx *c, int32_t run) {
  Run *R = get_run(c, run);
  if (!R) return fail(c, SPF_EINVAL, "bad run id");
  PhaseTimer pt(c, SPF_T_DEPOSIT);
  HIPCALL(c, deposit_currents(*R, c->stream));
  return SPF_OK;
}

int spf_particles_count(spf_ctx *c, int32_t run, int64_t *np, int64_t *np_active) {
  Run *R = get_run(c, run);
  if (!R) return fail(c, SPF_EINVAL, "bad run id");
  HIPCALL(c, hipMemcpyAsync(c->pinned, R->counters, 16 * sizeof(int64_t), hipMemcpyDeviceToHost, c->stream));
  HIPCALL(c, hipStreamSynchronize(c->stream));
  R->np_ub = c->pinned[0];
  if (np) *np = c->pinned[0];
  if (np_active) *np_active = c->pinned[1];
  return SPF_OK;
}

int spf_particles_append(spf_ctx *c, int32_t run, int64_t n, const spf_particles *h) {
  Run *R = get_run(c, run);
  if (!R || !h || n < 0) return fail(c, SPF_EINVAL, "bad run, null particles or negative count");
  if (n == 0) return SPF_OK;
  if (!h->Q || !h->z || !h->r || !h->pz || !h->pr || !h->r_pphi || !h->vz || !h->vr || !h->vphi || !h->r0 ||
      !h->z0 || !h->t0 || !h->rmax || !h->rmin || !h->zmax || !h->zmin || !h->particleID || !h->generation ||
      !h->status)
    return fail(c, SPF_EINVAL, "every particle column is required for append");
  if (R->np_ub + n > R->cap) {
    int rc = spf_particles_count(c, run, nullptr, nullptr);
    if (rc) return rc;
    if (R->np_ub + n > R->cap) return fail(c, SPF_ECAPACITY, "particle capacity exceeded");
  }
  size_t bytes = append_stage_bytes(n);
  HIPCALL(c, hipEventSynchronize(c->stage_done));
  if (c->hstage_bytes < bytes) {
    if (c->hstage) HIPCALL(c, hipHostFree(c->hstage));
    c->hstage = nullptr;
    HIPCALL(c, hipHostMalloc(&c->hstage, bytes, hipHostMallocDefault));
    c->hstage_bytes = bytes;
  }
  if (R->append_cap < bytes) {
    HIPCALL(c, hipStreamSynchronize(c->stream));
    if (R->append_stage) HIPCALL(c, hipFree(R->append_stage));
    R->append_stage = nullptr;
    HIPCALL(c, hipMalloc(&R->append_stage, bytes));
    R->append_cap = bytes;
  }
  double *dst = (double *)c->hstage;
  const double *cols[16] = {h->Q, h->z, h->r, h->pz, h->pr, h->r_pphi, h->vz, h->vr, h->vphi,
                            h->r0, h->z0, h->t0, h->rmax, h->rmin, h->zmax, h->zmin};
  for (int j = 0; j < 16; j++) memcpy(dst + (size_t)j * n, cols[j], sizeof(double) * n);
  uint8_t *b = (uint8_t *)c->hstage + 16 * 8 * (size_t)n;
  memcpy(b, h->particleID, 4 * n);
  memcpy(b + 4 * n, h->generation, 2 * n);
  size_t st_off = 16 * 8 * (size_t)n + 4 * (size_t)n + (((size_t)2 * n + 3) / 4) * 4;
  memcpy((uint8_t *)c->hstage + st_off, h->status, 4 * n);
  int64_t n_active = 0;
  for (int64_t k = 0; k < n; k++) n_active += (h->status[k] & SPF_PART_ACTIVE) ? 1 : 0;
  HIPCALL(c, hipMemcpyAsync(R->append_stage, c->hstage, bytes, hipMemcpyHostToDevice, c->stream));
  HIPCALL(c, hipEventRecord(c->stage_done, c->stream));
  HIPCALL(c, launch_append_kernels(*R, R->append_stage, n, n_active, R->Pdev, c->stream));
  R->np_ub += n;
  return SPF_OK;
}

int spf_particles_download(spf_ctx *c, int32_t run, const spf_particles *h, int64_t max, int64_t *n_out) {
  Run *R = get_run(c, run);
  if (!R || !h) return fail(c, SPF_EINVAL, "bad run or null particles");
  int64_t np = 0;
  int rc = spf_particles_count(c, run, &np, nullptr);
  if (rc) return rc;
  int64_t n = np < max ? np : max;
  if (n < 0) n = 0;
  double *cols[16] = {h->Q, h->z, h->r, h->pz, h->pr, h->r_pphi, h->vz, h->vr, h->vphi,
                      h->r0, h->z0, h->t0, h->rmax, h->rmin, h->zmax, h->zmin};
  if (n > 0) {
    hipStream_t s = c->stream;
    for (int j = 0; j < 16; j++)
      if (cols[j]) HIPCALL(c, hipMemcpyAsync(cols[j], R->P[j], sizeof(double) * n, hipMemcpyDeviceToHost, s));
    if (h->particleID) HIPCALL(c, hipMemcpyAsync(h->particleID, R->pid, 4 * n, hipMemcpyDeviceToHost, s));
    if (h->generation) HIPCALL(c, hipMemcpyAsync(h->generation, R->gen, 2 * n, hipMemcpyDeviceToHost, s));
    if (h->status) HIPCALL(c, hipMemcpyAsync(h->status, R->status, 4 * n, hipMemcpyDeviceToHost, s));
    HIPCALL(c, hipStreamSynchronize(s));
  }
  if (n_out) *n_out = n;
  return SPF_OK;
}

}  // extern "C"

void spf::push_args(Run *R, double zForcesStart, PushArgs &a, RowsArgs &w) {
  const spf_run_desc &d = R->d;
  double time = R->time_step * d.dt + d.start_time;
  a.nz = d.nz;
  a.nr = d.nr;
  a.ldz = R->ldz;
  a.tm = (d.modes & SPF_TM_FIELDS) != 0;
  a.te = (d.modes & SPF_TE_FIELDS) != 0;
  a.longit = (d.modes & SPF_LONGIT_INTERP) != 0;
  a.radial = (d.modes & SPF_RADIAL_INTERP) != 0;
  a.has_sol = d.has_solenoid != 0;
  a.open = R->open;
  a.zmin = d.zmin;
  a.zmax = d.zmax;
  a.rmax = d.rmax;
  a.dz = d.dz;
  a.dr = d.dr;
  a.dt = d.dt;
  a.time = time;
  a.c = d.c_mks;
  a.QoMC = d.QoMC;
  a.stiffness = d.stiffness;
  a.zForcesStart = zForcesStart;
  a.Ez = R->E2[0][R->cur[0]];
  a.Er = R->E2[1][R->cur[1]];
  a.Ephi = R->E2[3][R->cur[3]];
  a.Bphi2 = R->Bphi2;
  a.Br2 = R->Br2;
  a.Bz2 = R->Bz2;
  a.EzImp = R->EzImp;
  a.ErImp = R->ErImp;
  a.imp_e = R->imposed_e;
  a.BzImp = R->BzImp;
  a.BrImp = R->BrImp;
  a.flags = R->flags;
  for (int k = 0; k < 6; k++) {
    a.turnoff[k] = d.turn_off[k];
    a.cst[k] = d.constant_fields[k];
  }
  a.Q = R->P[0]; a.z = R->P[1]; a.r = R->P[2]; a.pz = R->P[3]; a.pr = R->P[4]; a.rpphi = R->P[5];
  a.vz = R->P[6]; a.vr = R->P[7]; a.vphi = R->P[8]; a.r0 = R->P[9]; a.z0 = R->P[10]; a.t0 = R->P[11];
  a.rmax_p = R->P[12]; a.rmin_p = R->P[13]; a.zmax_p = R->P[14]; a.zmin_p = R->P[15];
  a.status = R->status;
  a.np = R->counters;
  a.evt = R->evt;
  a.evflag = R->evflag;
  a.pcount = R->pcount;
  a.ev_total = (unsigned long long *)(R->counters + 5);
  a.ev_zold = R->ev[0]; a.ev_rold = R->ev[1]; a.ev_vz = R->ev[2]; a.ev_vr = R->ev[3];
  a.ev_pz0 = R->ev[4]; a.ev_pr0 = R->ev[5]; a.ev_rp0 = R->ev[6];
  a.last_old = R->last_old;
  a.new_active = (unsigned long long *)(R->counters + 7);
  a.nscreens = d.n_screens;
  for (int s = 0; s < kMaxScreens; s++) {
    a.sz[s] = 0;
    a.sdir[s] = 0;
  }
  for (int s = 0; s < d.n_screens; s++) {
    a.sz[s] = R->screen_z[s];
    a.sdir[s] = R->screen_dir[s];
  }
  w.nz = d.nz;
  w.nr = d.nr;
  w.ldz = R->ldz;
  w.compat = d.compat;
  w.nscreens = d.n_screens;
  w.zmin = d.zmin;
  w.zmax = d.zmax;
  w.dz = d.dz;
  w.dr = d.dr;
  w.dt = d.dt;
  w.time = time;
  w.step = R->time_step;
  w.flags = R->flags;
  w.evflag = R->evflag;
  w.pcount = R->pcount;
  w.ev_total = (unsigned long long *)(R->counters + 5);
  w.n_ub = R->np_ub;
  w.npb = (R->np_ub + 255) / 256;
  w.evlist = R->evlist;
  w.n_ev = R->counters + 2;
  w.np = R->counters;
  w.np_active = R->counters + 1;
  w.new_active = (unsigned long long *)(R->counters + 7);
  w.evt = R->evt;
  w.ev_zold = R->ev[0]; w.ev_rold = R->ev[1]; w.ev_vz = R->ev[2]; w.ev_vr = R->ev[3];
  w.ev_pz0 = R->ev[4]; w.ev_pr0 = R->ev[5]; w.ev_rp0 = R->ev[6];
  w.last_old = R->last_old;
  w.Q = R->P[0]; w.z = R->P[1]; w.r = R->P[2]; w.pz = R->P[3]; w.pr = R->P[4]; w.rpphi = R->P[5];
  w.vz = R->P[6]; w.vr = R->P[7]; w.vphi = R->P[8]; w.r0 = R->P[9]; w.z0 = R->P[10]; w.t0 = R->P[11];
  w.rmax_p = R->P[12]; w.rmin_p = R->P[13]; w.zmax_p = R->P[14]; w.zmin_p = R->P[15];
  w.pid = R->pid;
  w.gen = R->gen;
  w.status = R->status;
  for (int s = 0; s < kMaxScreens; s++) {
    w.sz[s] = a.sz[s];
    w.sdir[s] = a.sdir[s];
  }
  w.srows = R->srows;
  w.n_srows = R->counters + 3;
  w.srow_cap = R->srow_cap;
  w.lrows = R->lrows;
  w.n_lrows = R->counters + 4;
  w.lrow_cap = R->lrow_cap;
  w.move_src = R->move_src;
  w.move_dst = R->move_dst;
  w.tflag = R->gflag;
  w.overflow = (int32_t *)(R->counters + 8);
  w.evoff = R->evoff;
  w.tcnt = R->tcnt;
  w.toff = R->toff;
  w.ntmax = rows_tiles(R->cap > 0 ? R->cap : 1);
  w.cbase = R->cbase;
}

extern "C" {

int spf_particles_sort(spf_ctx *c, int32_t run) {
  Run *R = get_run(c, run);
  if (!R) return fail(c, SPF_EINVAL, "bad run id");
  if (R->cell_sort <= 0) return fail(c, SPF_EINVAL, "the run keeps the reference's particle order (cell_sort = 0)");
  PhaseTimer pt(c, SPF_T_SORT);
  HIPCALL(c, sort_particles(*R, c->stream));
  return SPF_OK;
}

int spf_set_sor_shape(spf_ctx *c, int32_t shape) {
  if (!c) return SPF_EINVAL;
  if (shape < -1 || shape > 1) return fail(c, SPF_EINVAL, "shape must be -1, 0 or 1");
  c->sor_shape = shape;
  return SPF_OK;
}

int spf_set_cell_sort(spf_ctx *c, int32_t run, int64_t interval) {
  Run *R = get_run(c, run);
  if (!R) return fail(c, SPF_EINVAL, "bad run id");
  if (R->cell_sort <= 0) return fail(c, SPF_EINVAL, "the run was created with cell_sort = 0");
  if (interval <= 0) return fail(c, SPF_EINVAL, "the re-sort interval must be > 0");
  R->cell_sort = interval;
  return SPF_OK;
}

int spf_particles_push(spf_ctx *c, int32_t run, double zForcesStart) {
  Run *R = get_run(c, run);
  if (!R) return fail(c, SPF_EINVAL, "bad run id");
  PushArgs a;
  RowsArgs w;
  push_args(R, zForcesStart, a, w);
  PhaseTimer pt(c, SPF_T_PUSH);
  HIPCALL(c, launch_push(a, R->np_ub, c->stream));
  HIPCALL(c, launch_rows(w, c->stream));
  return SPF_OK;
}

int spf_events_drain(spf_ctx *c, int32_t run, spf_screen_row *rows, int64_t max_rows, int64_t *n_rows,
                     spf_lost_row *lost, int64_t max_lost, int64_t *n_lost) {
  Run *R = get_run(c, run);
  if (!R) return fail(c, SPF_EINVAL, "bad run id");
  HIPCALL(c, hipMemcpyAsync(c->pinned, R->counters, 16 * sizeof(int64_t), hipMemcpyDeviceToHost, c->stream));
  HIPCALL(c, hipStreamSynchronize(c->stream));
  int64_t ns = c->pinned[3], nl = c->pinned[4];
  int32_t ovf = (int32_t)(c->pinned[8] & 0xffffffff);
  int64_t cs = ns < R->srow_cap ? ns : R->srow_cap, cl = nl < R->lrow_cap ? nl : R->lrow_cap;
  int64_t ks = rows ? (cs < max_rows ? cs : max_rows) : 0, kl = lost ? (cl < max_lost ? cl : max_lost) : 0;
  if (ks > 0) HIPCALL(c, hipMemcpyAsync(rows, R->srows, sizeof(spf_screen_row) * ks, hipMemcpyDeviceToHost, c->stream));
  if (kl > 0) HIPCALL(c, hipMemcpyAsync(lost, R->lrows, sizeof(spf_lost_row) * kl, hipMemcpyDeviceToHost, c->stream));
  if (n_rows) *n_rows = ns;
  if (n_lost) *n_lost = nl;
  c->pinned[3] = 0;
  c->pinned[4] = 0;
  c->pinned[8] = 0;
  HIPCALL(c, hipMemcpyAsync(R->counters + 3, c->pinned + 3, 2 * sizeof(int64_t), hipMemcpyHostToDevice, c->stream));
  HIPCALL(c, hipMemcpyAsync(R->counters + 8, c->pinned + 8, sizeof(int64_t), hipMemcpyHostToDevice, c->stream));
  HIPCALL(c, hipStreamSynchronize(c->stream));
  if (ovf || ks < ns || kl < nl) return fail(c, SPF_ECAPACITY, "event rows exceeded a buffer; rows were dropped");
  return SPF_OK;
}

}  // extern "C"
