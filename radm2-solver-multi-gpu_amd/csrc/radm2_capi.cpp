// C ABI of the RADM2 integrator (declared in include/radm2.h).
//
// Host side: option resolution with KPP's defaults and error codes
// (mgpu.cu:3664-3848), domain-table upload, schedule choice, launch, events,
// graph plans, and the KPP-convention front-end radm2_rosenbrock().
#include "radm2.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "radm2_kernels.h"

using radm2::DomainDesc;
using radm2::LaunchArgs;
using radm2::Ros3Opts;

struct radm2_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  int cu_count = 0;
  int schedule = RADM2_SCHED_PERSISTENT;
  int block = 0;       // 0 = tuned default
  int refill_min = 0;  // 0 = tuned default
  int waves = 0;       // register-budget variant (waves per SIMD), 0 = tuned default
  int placement = -1;  // working-array placement, -1 = tuned default
  double* d_ws = nullptr;  // explicit workspace (grows on demand)
  size_t ws_bytes = 0;
  DomainDesc* d_doms = nullptr;
  DomainDesc* h_doms = nullptr;  // pinned staging
  int doms_cap = 0;
  unsigned long long* d_counter = nullptr;  // [1]
  unsigned long long* d_stats = nullptr;    // [4]
  unsigned long long* d_slots = nullptr;    // [kSlotWords] grid block-slot bitmap
  bool stats_on = false;
  hipEvent_t ev0 = nullptr, ev1 = nullptr, ev_staged = nullptr;
  bool timed = false;
  void* scratch = nullptr;  // this device's array of radm2_integrate_noalloc
};

struct radm2_plan {
  radm2_ctx* ctx = nullptr;
  DomainDesc* d_doms = nullptr;
  unsigned long long* d_counter = nullptr;
  unsigned long long* d_slots = nullptr;
  double* d_ws = nullptr;
  size_t ws_bytes = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

namespace {

#define HIP_TRY(x)                       \
  do {                                   \
    hipError_t e_ = (x);                 \
    if (e_ != hipSuccess) {              \
      last_hip_error() = e_;             \
      return RADM2_E_HIP;                \
    }                                    \
  } while (0)

hipError_t& last_hip_error() {
  static thread_local hipError_t e = hipSuccess;
  return e;
}

constexpr int kDefaultGridBlock = 64;
// grid schedule: bitmap words of the block-slot pool (64 slots each) and
// the slots added to the resident-block count as a margin
constexpr int kSlotWords = 128;
constexpr int kSlotMargin = 8;
constexpr int kDefaultPersistentBlock = 64;
// Defaults measured on MI355X at 1M cells (the autotuner may override):
// persistent schedule, explicit workspace placement with its load-scheduled
// kernels, the 160 most accessed doubles per lane in LDS (two 64-lane
// blocks per CU) and the next 112 in registers (1 wave per SIMD register
// budget; 128 left no room for load batches of 20), refill once three
// lanes of a wave are idle (1: 527.3 ms, 2: 521.2, 3: 517.6, 4: 519.5, 5:
// 524.6, 6: 529.9 at 1M cells).
constexpr int kDefaultRefillMin = 3;
constexpr int kDefaultWaves = 1;
constexpr int kDefaultPlacement = radm2::PLACE_HOT160R112;

// The configured kernel variant; family: one that runs all of KPP's
// Rosenbrock methods (built for the default placement only).
const radm2::KernelVariant* pick_variant(const radm2_ctx* c, bool family) {
  const int w = c->waves ? c->waves : kDefaultWaves;
  const int pl = c->placement >= 0 ? c->placement : kDefaultPlacement;
  const radm2::KernelVariant* v = radm2::kernel_variants();
  for (int i = 0; i < radm2::kNumVariants; i++)
    if (v[i].waves == w && v[i].placement == pl && v[i].family == (int)family) return &v[i];
  return nullptr;
}

// Workspace for `waves` wave slots, grown (never shrunk) on demand.  The
// stream is drained first: a running kernel may still use the old buffer.
int ensure_ws(hipStream_t st, double** ws, size_t* have, int64_t waves, size_t wave_bytes) {
  const size_t need = (size_t)waves * wave_bytes;
  if (need <= *have) return 0;
  HIP_TRY(hipStreamSynchronize(st));
  if (*ws) HIP_TRY(hipFree(*ws));
  *ws = nullptr;
  *have = 0;
  HIP_TRY(hipMalloc(ws, need));
  *have = need;
  return 0;
}

// KPP option resolution (Rosenbrock(), mgpu.cu:3688-3811); INTEGRATE()'s
// call when `in` is null (mgpu.cu:3866-3873).
int resolve_opts(double dt, double rtol, double atol, const radm2_opts* in, Ros3Opts* o) {
  radm2_opts d;
  if (!in) {
    std::memset(&d, 0, sizeof d);
    d.hstart = 0.01 * dt;
    d.method = 2;
    in = &d;
  }
  const double span = std::fabs(dt);
  o->dt = dt;
  o->rtol = rtol;
  o->atol = atol;
  o->roundoff = DBL_EPSILON;  // WLAMCH('E') = 2^-52 (mgpu.cu:949-973)
  o->max_steps = in->max_steps == 0 ? 100000 : in->max_steps;
  if (o->max_steps < 0) return -1;
  const int method = in->method == 0 ? 2 : in->method;
  if (method < 1 || method > 5) return -2;
  o->hmin = in->hmin;
  if (in->hmin < 0.0) return -3;
  o->hmax = in->hmax == 0.0 ? span : std::min(std::fabs(in->hmax), span);
  if (in->hmax < 0.0) return -3;
  o->hstart = in->hstart == 0.0 ? std::max(o->hmin, 1.0e-6) : std::min(std::fabs(in->hstart), span);
  if (in->hstart < 0.0) return -3;
  o->facmin = in->facmin == 0.0 ? 0.2 : in->facmin;
  if (in->facmin < 0.0) return -4;
  o->facmax = in->facmax == 0.0 ? 6.0 : in->facmax;
  if (in->facmax < 0.0) return -4;
  o->facrej = in->facrej == 0.0 ? 0.1 : in->facrej;
  if (in->facrej < 0.0) return -4;
  o->facsafe = in->facsafe == 0.0 ? 0.9 : in->facsafe;
  if (in->facsafe < 0.0) return -4;
  if (atol <= 0.0 || rtol <= 10.0 * o->roundoff || rtol >= 1.0) return -5;
  // the reference initialises only Ros3 (mgpu.cu:3802-3811); KPP's other
  // four methods run through the generic stage loop
  o->method = method;
  if (in->sing_retries < 0) return RADM2_E_ARG;
  o->sing_retries = in->sing_retries;
  o->pad = 0;
  return 0;
}

// The reference's N2O5 hydrolysis rate formula (mgpu.cu:552) at RH = 0.
double default_rc_n2o5() { return 1.0 / (3.6e4 * std::exp(-std::pow(0.0 / 0.28, 2.8) + 300.0)); }

int fill_desc(const radm2_domain& in, int64_t first, DomainDesc* d) {
  if (in.ncells < 0) return RADM2_E_ARG;
  if (in.ncells > 0) {
    if (in.ld < in.ncells || !in.conc || !in.temp || !in.cair || !in.h2o || !in.jrate)
      return RADM2_E_ARG;
    if (!(in.dt > 0.0)) return RADM2_E_ARG;
  }
  std::memset(d, 0, sizeof *d);
  if (int rc = resolve_opts(in.dt, in.rtol, in.atol, in.opts, &d->opts)) return rc;
  d->first = first;
  d->n = in.ncells;
  d->ld = in.ld;
  d->conc = in.conc;
  d->temp = in.temp;
  d->cair = in.cair;
  d->h2o = in.h2o;
  d->jrate = in.jrate;
  d->rc = in.rc_n2o5;
  d->h0 = in.h0;
  d->nstep = in.nstep;
  d->nacc = in.nacc;
  d->nrej = in.nrej;
  d->status = in.status;
  d->hexit = in.hexit;
  d->hfirst = in.hfirst;
  d->rc_default = default_rc_n2o5();
  d->jrate_end = in.jrate_end;
  return 0;
}

int ensure_doms(radm2_ctx* c, int n) {
  if (n <= c->doms_cap) return 0;
  int cap = std::max(n, 64);
  if (c->ev_staged) HIP_TRY(hipEventSynchronize(c->ev_staged));
  if (c->d_doms) HIP_TRY(hipFree(c->d_doms));
  if (c->h_doms) HIP_TRY(hipHostFree(c->h_doms));
  c->d_doms = nullptr;
  c->h_doms = nullptr;
  HIP_TRY(hipMalloc(&c->d_doms, sizeof(DomainDesc) * cap));
  HIP_TRY(hipHostMalloc(&c->h_doms, sizeof(DomainDesc) * cap, hipHostMallocDefault));
  c->doms_cap = cap;
  return 0;
}

int pick_block(const radm2_ctx* c) {
  if (c->block) return c->block;
  return c->schedule == RADM2_SCHED_GRID ? kDefaultGridBlock : kDefaultPersistentBlock;
}

// Resident blocks per CU of a variant's kernel (a schedule's occupancy
// query); <= 0 records the HIP error and returns RADM2_E_HIP.
int resident_per_cu(int (*query)(int), int block, int* per_cu) {
  *per_cu = query(block);
  if (*per_cu > 0) return 0;
  last_hip_error() = *per_cu < 0 ? (hipError_t)(-*per_cu) : hipErrorInvalidConfiguration;
  return RADM2_E_HIP;
}

// Launch shape of a call: the variant, block size, grid blocks (persistent:
// resident blocks; grid: blocks of the grid), workspace wave slots and, for
// the grid on a workspace placement, the block slots of its pool.
struct Shape {
  const radm2::KernelVariant* v = nullptr;
  int block = 0;
  int64_t blocks = 0, waves = 0, slots = 0;
};

// sched: the schedule (the context's when negative); max_waves > 0 caps the
// persistent kernel's resident waves (a workspace of that many wave slots).
int launch_shape(const radm2_ctx* c, int64_t total, bool family, Shape* sh, int sched = -1,
                 int64_t max_waves = 0) {
  if (sched < 0) sched = c->schedule;
  // a schedule other than the context's runs at that schedule's default block
  sh->block = sched == c->schedule ? pick_block(c)
              : (sched == RADM2_SCHED_GRID ? kDefaultGridBlock : kDefaultPersistentBlock);
  sh->v = pick_variant(c, family);
  if (!sh->v) return family ? RADM2_E_UNSUPPORTED : RADM2_E_ARG;
  const radm2::KernelVariant* v = sh->v;
  const int block = sh->block;
  if (v->placement >= radm2::PLACE_HOT160 && block != 64) return RADM2_E_ARG;
  // cells spread over several lanes run on the persistent schedule only
  if (v->lanes > 1 && sched == RADM2_SCHED_GRID) return RADM2_E_UNSUPPORTED;
  const int64_t wpb = block / radm2::WAVE;
  const int64_t need = (total + block - 1) / block;
  int per_cu = 0;
  if (sched == RADM2_SCHED_GRID) {
    sh->blocks = need;
    if (v->placement >= radm2::PLACE_WS) {
      if (int rc = resident_per_cu(v->grid_blocks_per_cu, block, &per_cu)) return rc;
      sh->slots = std::min<int64_t>(need, (int64_t)per_cu * c->cu_count + kSlotMargin);
      if (sh->slots > 64 * kSlotWords) return RADM2_E_UNSUPPORTED;
    }
    sh->waves = sh->slots * wpb;
  } else {
    if (int rc = resident_per_cu(v->blocks_per_cu, block, &per_cu)) return rc;
    sh->blocks = std::max<int64_t>(1, std::min<int64_t>((int64_t)per_cu * c->cu_count, need));
    if (max_waves > 0 && v->placement >= radm2::PLACE_WS) {
      sh->blocks = std::min<int64_t>(sh->blocks, max_waves / wpb);
      if (sh->blocks < 1) return RADM2_E_UNSUPPORTED;
    }
    sh->waves = sh->blocks * wpb;
  }
  return 0;
}

// Launch the kernels for an uploaded table on stream st.  ws/ws_bytes: the
// workspace to use (grown here when the explicit placement needs more);
// grow = false for graph capture, where the plan sized it beforehand.
// slots: the grid schedule's block-slot bitmap (kSlotWords words).
// sched, max_waves: as launch_shape.
int enqueue(radm2_ctx* c, const DomainDesc* d_doms, int ndom, int64_t total,
            unsigned long long* counter, unsigned long long* slots, hipStream_t st, double** ws,
            size_t* ws_bytes, bool family, bool grow = true, int sched = -1,
            int64_t max_waves = 0) {
  if (sched < 0) sched = c->schedule;
  LaunchArgs a;
  a.doms = d_doms;
  a.ndom = ndom;
  a.total = total;
  a.counter = counter;
  a.stats = c->stats_on ? c->d_stats : nullptr;
  a.refill_min = c->refill_min ? c->refill_min : kDefaultRefillMin;
  a.slot_bits = slots;
  a.slot_words = 0;
  Shape sh;
  if (int rc = launch_shape(c, total, family, &sh, sched, max_waves)) return rc;
  const radm2::KernelVariant* v = sh.v;
  if (v->placement >= radm2::PLACE_WS) {
    if (grow) {
      if (int rc = ensure_ws(st, ws, ws_bytes, sh.waves, radm2::ws_wave_bytes(v))) return rc;
    } else if (*ws_bytes < (size_t)sh.waves * radm2::ws_wave_bytes(v)) {
      return RADM2_E_ARG;
    }
    a.ws = *ws;
    a.ws_waves = sh.waves;
  } else {
    a.ws = nullptr;
    a.ws_waves = 0;
  }
  if (sched == RADM2_SCHED_GRID) {
    if (sh.slots > 0) {
      a.slot_words = (int32_t)((sh.slots + 63) / 64);
      HIP_TRY(radm2::launch_slot_init(slots, a.slot_words, sh.slots, st));
    }
    HIP_TRY(v->grid(sh.block, a, st));
  } else {
    HIP_TRY(radm2::launch_fill_u64(counter, 1, 0ull, st));
    HIP_TRY(v->persistent(sh.block, (int)sh.blocks, a, st));
  }
  return 0;
}

// Device table of the domains.  *family: some domain uses a method other
// than Ros3, the non-autonomous forcing or singular-matrix retries, which
// only the family variant runs (RADM2_E_UNSUPPORTED when the configured
// placement has none).
int build_table(const radm2_ctx* c, const radm2_domain* doms, int ndom,
                std::vector<DomainDesc>& out, int64_t* total, bool* family) {
  out.resize(ndom);
  int64_t first = 0;
  bool other_methods = false;
  for (int i = 0; i < ndom; i++) {
    if (int rc = fill_desc(doms[i], first, &out[i])) return rc;
    other_methods |= (out[i].opts.method != 2 || out[i].opts.sing_retries > 0 ||
                      doms[i].jrate_end != nullptr) &&
                     doms[i].ncells > 0;
    first += doms[i].ncells;
  }
  *total = first;
  *family = other_methods;
  if (other_methods && !pick_variant(c, true)) return RADM2_E_UNSUPPORTED;
  return 0;
}

}  // namespace

extern "C" {

int radm2_create(radm2_ctx** ctx, int device, void* stream) {
  if (!ctx) return RADM2_E_ARG;
  *ctx = nullptr;
  radm2_ctx* c = new radm2_ctx();
  c->device = device;
  auto fail = [&](hipError_t e) {
    last_hip_error() = e;
    radm2_destroy(c);
    return RADM2_E_HIP;
  };
  hipError_t e = hipSetDevice(device);
  if (e != hipSuccess) return fail(e);
  e = hipDeviceGetAttribute(&c->cu_count, hipDeviceAttributeMultiprocessorCount, device);
  if (e != hipSuccess) return fail(e);
  // this device's copy of the allocation-free entry point's array
  if ((e = radm2::noalloc_scratch(device, &c->scratch)) != hipSuccess) return fail(e);
  if (stream) {
    c->stream = (hipStream_t)stream;
  } else {
    e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) return fail(e);
    c->own_stream = true;
  }
  if ((e = hipMalloc(&c->d_counter, sizeof(unsigned long long))) != hipSuccess) return fail(e);
  if ((e = hipMalloc(&c->d_stats, 4 * sizeof(unsigned long long))) != hipSuccess) return fail(e);
  if ((e = hipMalloc(&c->d_slots, kSlotWords * sizeof(unsigned long long))) != hipSuccess)
    return fail(e);
  if ((e = hipMemset(c->d_stats, 0, 4 * sizeof(unsigned long long))) != hipSuccess) return fail(e);
  if ((e = hipEventCreate(&c->ev0)) != hipSuccess) return fail(e);
  if ((e = hipEventCreate(&c->ev1)) != hipSuccess) return fail(e);
  if ((e = hipEventCreateWithFlags(&c->ev_staged, hipEventDisableTiming)) != hipSuccess)
    return fail(e);
  if (ensure_doms(c, 64) != 0) {
    radm2_destroy(c);
    return RADM2_E_HIP;
  }
  *ctx = c;
  return RADM2_OK;
}

void radm2_destroy(radm2_ctx* c) {
  if (!c) return;
  hipSetDevice(c->device);
  if (c->stream) hipStreamSynchronize(c->stream);
  if (c->d_doms) hipFree(c->d_doms);
  if (c->h_doms) hipHostFree(c->h_doms);
  if (c->d_counter) hipFree(c->d_counter);
  if (c->d_stats) hipFree(c->d_stats);
  if (c->d_slots) hipFree(c->d_slots);
  if (c->d_ws) hipFree(c->d_ws);
  if (c->ev0) hipEventDestroy(c->ev0);
  if (c->ev1) hipEventDestroy(c->ev1);
  if (c->ev_staged) hipEventDestroy(c->ev_staged);
  if (c->own_stream && c->stream) hipStreamDestroy(c->stream);
  delete c;
}

int radm2_set_schedule(radm2_ctx* c, int schedule, int block, int refill_min) {
  if (!c) return RADM2_E_ARG;
  if (schedule != RADM2_SCHED_GRID && schedule != RADM2_SCHED_PERSISTENT) return RADM2_E_ARG;
  if (block != 0 && block != 64 && block != 256 && block != 512 && block != 1024) return RADM2_E_ARG;
  if (schedule == RADM2_SCHED_PERSISTENT && block > 256) return RADM2_E_ARG;
  if (refill_min < 0 || refill_min > 64) return RADM2_E_ARG;
  c->schedule = schedule;
  c->block = block;
  c->refill_min = refill_min;
  return RADM2_OK;
}

int radm2_set_kernel(radm2_ctx* c, int placement, int waves_per_simd) {
  if (!c) return RADM2_E_ARG;
  if (placement < -1 || placement > radm2::PLACE_HOT_LAST) return RADM2_E_ARG;
  if (waves_per_simd != 0 && waves_per_simd != 1 && waves_per_simd != 2 && waves_per_simd != 4 &&
      waves_per_simd != 8)
    return RADM2_E_ARG;
  const int pl = placement >= 0 ? placement : kDefaultPlacement;
  const int w = waves_per_simd ? waves_per_simd : kDefaultWaves;
  const radm2::KernelVariant* v = radm2::kernel_variants();
  bool found = false;
  for (int i = 0; i < radm2::kNumVariants; i++)
    found |= (v[i].placement == pl && v[i].waves == w && v[i].family == 0);
  if (!found) return RADM2_E_UNSUPPORTED;
  c->placement = placement;
  c->waves = waves_per_simd;
  return RADM2_OK;
}

int radm2_integrate_batched(radm2_ctx* c, int ndom, const radm2_domain* doms) {
  if (!c || ndom < 0 || (ndom > 0 && !doms)) return RADM2_E_ARG;
  std::vector<DomainDesc> tab;
  int64_t total = 0;
  bool family = false;
  if (int rc = build_table(c, doms, ndom, tab, &total, &family)) return rc;
  c->timed = false;
  if (total == 0) return RADM2_OK;
  HIP_TRY(hipSetDevice(c->device));
  if (int rc = ensure_doms(c, ndom)) return rc;
  // the staging buffer may still feed the previous call's copy
  HIP_TRY(hipEventSynchronize(c->ev_staged));
  std::memcpy(c->h_doms, tab.data(), sizeof(DomainDesc) * ndom);
  HIP_TRY(hipMemcpyAsync(c->d_doms, c->h_doms, sizeof(DomainDesc) * ndom,
                         hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipEventRecord(c->ev_staged, c->stream));
  HIP_TRY(hipEventRecord(c->ev0, c->stream));
  if (int rc = enqueue(c, c->d_doms, ndom, total, c->d_counter, c->d_slots, c->stream, &c->d_ws,
                       &c->ws_bytes, family))
    return rc;
  HIP_TRY(hipEventRecord(c->ev1, c->stream));
  c->timed = true;
  return RADM2_OK;
}

int radm2_integrate(radm2_ctx* c, int64_t ncells, int64_t ld, double* conc, const double* temp,
                    const double* cair, const double* h2o, const double* jrate,
                    const double* rc_n2o5, double dt, double rtol, double atol,
                    const radm2_opts* opts, const double* h0, int32_t* nstep, int32_t* nacc,
                    int32_t* nrej, int32_t* status, double* hexit) {
  radm2_domain d;
  std::memset(&d, 0, sizeof d);
  d.ncells = ncells;
  d.ld = ld;
  d.conc = conc;
  d.temp = temp;
  d.cair = cair;
  d.h2o = h2o;
  d.jrate = jrate;
  d.rc_n2o5 = rc_n2o5;
  d.h0 = h0;
  d.nstep = nstep;
  d.nacc = nacc;
  d.nrej = nrej;
  d.status = status;
  d.hexit = hexit;
  d.dt = dt;
  d.rtol = rtol;
  d.atol = atol;
  d.opts = opts;
  return radm2_integrate_batched(c, 1, &d);
}

int radm2_integrate_noalloc(radm2_ctx* c, int64_t ncells, int64_t ld, double* conc,
                            const double* temp, const double* cair, const double* h2o,
                            const double* jrate, const double* rc_n2o5, double dt, double rtol,
                            double atol, const radm2_opts* opts, const double* h0,
                            int32_t* nstep, int32_t* nacc, int32_t* nrej, int32_t* status,
                            double* hexit) {
  if (!c || !c->scratch) return RADM2_E_ARG;
  radm2_domain d;
  std::memset(&d, 0, sizeof d);
  d.ncells = ncells;
  d.ld = ld;
  d.conc = conc;
  d.temp = temp;
  d.cair = cair;
  d.h2o = h2o;
  d.jrate = jrate;
  d.rc_n2o5 = rc_n2o5;
  d.h0 = h0;
  d.nstep = nstep;
  d.nacc = nacc;
  d.nrej = nrej;
  d.status = status;
  d.hexit = hexit;
  d.dt = dt;
  d.rtol = rtol;
  d.atol = atol;
  d.opts = opts;
  std::vector<DomainDesc> tab;
  int64_t total = 0;
  bool family = false;
  if (int rc = build_table(c, &d, 1, tab, &total, &family)) return rc;
  c->timed = false;
  if (total == 0) return RADM2_OK;
  HIP_TRY(hipSetDevice(c->device));
  // the array: counter, domain descriptor, workspace (no allocation, no
  // copy from host memory: the descriptor goes in as a kernel argument)
  unsigned char* base = static_cast<unsigned char*>(c->scratch);
  auto* counter = reinterpret_cast<unsigned long long*>(base);
  auto* d_dom = reinterpret_cast<DomainDesc*>(base + radm2::kNoallocDomOffset);
  double* ws = reinterpret_cast<double*>(base + radm2::kNoallocWsOffset);
  size_t ws_bytes = radm2::kNoallocScratchBytes - radm2::kNoallocWsOffset;
  const radm2::KernelVariant* v = pick_variant(c, family);
  if (!v) return family ? RADM2_E_UNSUPPORTED : RADM2_E_ARG;
  const int64_t max_waves =
      (int64_t)(ws_bytes / radm2::ws_wave_bytes(v));
  HIP_TRY(hipEventRecord(c->ev0, c->stream));
  HIP_TRY(radm2::launch_put_domain(tab[0], d_dom, c->stream));
  if (int rc = enqueue(c, d_dom, 1, total, counter, nullptr, c->stream, &ws, &ws_bytes, family,
                       false, RADM2_SCHED_PERSISTENT, max_waves))
    return rc;
  HIP_TRY(hipEventRecord(c->ev1, c->stream));
  c->timed = true;
  return RADM2_OK;
}

int radm2_plan_create(radm2_ctx* c, int ndom, const radm2_domain* doms, radm2_plan** out) {
  if (!c || !out || ndom <= 0 || !doms) return RADM2_E_ARG;
  *out = nullptr;
  std::vector<DomainDesc> tab;
  int64_t total = 0;
  bool family = false;
  if (int rc = build_table(c, doms, ndom, tab, &total, &family)) return rc;
  HIP_TRY(hipSetDevice(c->device));
  radm2_plan* p = new radm2_plan();
  p->ctx = c;
  auto fail = [&](hipError_t e) {
    last_hip_error() = e;
    radm2_plan_destroy(p);
    return RADM2_E_HIP;
  };
  // the launch shape is checked before anything is allocated or captured
  Shape sh;
  if (int rc = launch_shape(c, total, family, &sh)) {
    radm2_plan_destroy(p);
    return rc;
  }
  hipError_t e;
  if ((e = hipMalloc(&p->d_doms, sizeof(DomainDesc) * ndom)) != hipSuccess) return fail(e);
  if ((e = hipMalloc(&p->d_counter, sizeof(unsigned long long))) != hipSuccess) return fail(e);
  if ((e = hipMalloc(&p->d_slots, kSlotWords * sizeof(unsigned long long))) != hipSuccess)
    return fail(e);
  if ((e = hipMemcpy(p->d_doms, tab.data(), sizeof(DomainDesc) * ndom, hipMemcpyHostToDevice)) !=
      hipSuccess)
    return fail(e);
  // capture on a private stream: counter reset or slot-pool reset (fill
  // kernels, no memset node) followed by the solver launch
  hipStream_t cap;
  if ((e = hipStreamCreateWithFlags(&cap, hipStreamNonBlocking)) != hipSuccess) return fail(e);
  // size the plan's own workspace before capture (no allocation inside it)
  if (sh.v->placement >= radm2::PLACE_WS &&
      ensure_ws(cap, &p->d_ws, &p->ws_bytes, sh.waves, radm2::ws_wave_bytes(sh.v)) != 0) {
    hipStreamDestroy(cap);
    return fail(last_hip_error());
  }
  int rc = 0;
  e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
  if (e == hipSuccess) {
    rc = enqueue(c, p->d_doms, ndom, total, p->d_counter, p->d_slots, cap, &p->d_ws, &p->ws_bytes,
                 family, false);
    e = hipStreamEndCapture(cap, &p->graph);
  }
  hipStreamDestroy(cap);
  if (rc != 0) {
    // the enqueue's own error (its HIP cause, if any, is already recorded)
    radm2_plan_destroy(p);
    return rc;
  }
  if (e != hipSuccess) return fail(e);
  if ((e = hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0)) != hipSuccess)
    return fail(e);
  *out = p;
  return RADM2_OK;
}

int radm2_plan_launch(radm2_plan* p) {
  if (!p || !p->exec) return RADM2_E_ARG;
  radm2_ctx* c = p->ctx;
  HIP_TRY(hipSetDevice(c->device));
  HIP_TRY(hipEventRecord(c->ev0, c->stream));
  HIP_TRY(hipGraphLaunch(p->exec, c->stream));
  HIP_TRY(hipEventRecord(c->ev1, c->stream));
  c->timed = true;
  return RADM2_OK;
}

void radm2_plan_destroy(radm2_plan* p) {
  if (!p) return;
  if (p->ctx) {
    hipSetDevice(p->ctx->device);
    hipStreamSynchronize(p->ctx->stream);
  }
  if (p->exec) hipGraphExecDestroy(p->exec);
  if (p->graph) hipGraphDestroy(p->graph);
  if (p->d_doms) hipFree(p->d_doms);
  if (p->d_counter) hipFree(p->d_counter);
  if (p->d_slots) hipFree(p->d_slots);
  if (p->d_ws) hipFree(p->d_ws);
  delete p;
}

int radm2_sync(radm2_ctx* c) {
  if (!c) return RADM2_E_ARG;
  HIP_TRY(hipSetDevice(c->device));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return RADM2_OK;
}

int radm2_measure_fp64_peak(radm2_ctx* c, double* fp64_tflops) {
  if (!c || !fp64_tflops) return RADM2_E_ARG;
  HIP_TRY(hipSetDevice(c->device));
  HIP_TRY(hipStreamSynchronize(c->stream));
  HIP_TRY(radm2::measure_fp64_peak(c->stream, c->cu_count, fp64_tflops));
  return RADM2_OK;
}

int radm2_last_elapsed_ms(radm2_ctx* c, float* ms) {
  if (!c || !ms) return RADM2_E_ARG;
  if (!c->timed) {
    *ms = 0.0f;
    return RADM2_OK;
  }
  HIP_TRY(hipEventElapsedTime(ms, c->ev0, c->ev1));
  return RADM2_OK;
}

int radm2_enable_stats(radm2_ctx* c, int on) {
  if (!c) return RADM2_E_ARG;
  HIP_TRY(hipSetDevice(c->device));
  HIP_TRY(hipStreamSynchronize(c->stream));
  HIP_TRY(hipMemset(c->d_stats, 0, 4 * sizeof(unsigned long long)));
  c->stats_on = on != 0;
  return RADM2_OK;
}

int radm2_read_stats(radm2_ctx* c, uint64_t out[4]) {
  if (!c || !out) return RADM2_E_ARG;
  HIP_TRY(hipSetDevice(c->device));
  HIP_TRY(hipStreamSynchronize(c->stream));
  unsigned long long h[4];
  HIP_TRY(hipMemcpy(h, c->d_stats, sizeof h, hipMemcpyDeviceToHost));
  for (int i = 0; i < 4; i++) out[i] = h[i];
  return RADM2_OK;
}

int radm2_check_opts(double dt, double rtol, double atol, const radm2_opts* opts) {
  Ros3Opts o;
  return resolve_opts(dt, rtol, atol, opts, &o);
}

int radm2_rosenbrock(radm2_ctx* c, const radm2_domain* state, double Tstart, double Tend,
                     const double* AbsTol, const double* RelTol, double RPAR[20],
                     int32_t IPAR[20]) {
  if (!c || !state || !AbsTol || !RelTol || !RPAR || !IPAR) return RADM2_E_ARG;
  // KPP: Autonomous = IPAR[0] != 0; non-autonomous runs over state->jrate_end
  const bool autonomous = IPAR[0] != 0;
  const int ntol = IPAR[1] == 0 ? RADM2_NVAR : 1;
  for (int i = 1; i < ntol; i++)
    if (AbsTol[i] != AbsTol[0] || RelTol[i] != RelTol[0]) return RADM2_E_UNSUPPORTED;
  if (!(Tend > Tstart)) return RADM2_E_UNSUPPORTED;
  radm2_opts o;
  o.hmin = RPAR[0];
  o.hmax = RPAR[1];
  o.hstart = RPAR[2];
  o.facmin = RPAR[3];
  o.facmax = RPAR[4];
  o.facrej = RPAR[5];
  o.facsafe = RPAR[6];
  o.max_steps = IPAR[2];
  o.method = IPAR[3];
  o.sing_retries = 0;  // the reference's active code (mgpu.cu:3147-3163)
  // the reference rejects IPAR[3] outside 1..5 before defaulting (mgpu.cu:3721-3728)
  if (IPAR[3] < 1 || IPAR[3] > 5) return -2;
  const double dt = Tend - Tstart;
  Ros3Opts ro;
  if (int rc = resolve_opts(dt, RelTol[0], AbsTol[0], &o, &ro)) return rc;
  const int64_t n = state->ncells;
  if (n <= 0) return RADM2_E_ARG;
  // one allocation for the per-cell outputs: hexit first (8-byte aligned),
  // then the four int32 arrays
  HIP_TRY(hipSetDevice(c->device));
  void* buf = nullptr;
  HIP_TRY(hipMalloc(&buf, (size_t)n * (sizeof(double) + 4 * sizeof(int32_t))));
  double* hx = (double*)buf;
  int32_t* ns = (int32_t*)(hx + n);
  int32_t* na = ns + n;
  int32_t* nr = na + n;
  int32_t* st = nr + n;
  radm2_domain d = *state;
  d.nstep = ns;
  d.nacc = na;
  d.nrej = nr;
  d.status = st;
  d.hexit = hx;
  d.dt = dt;
  d.rtol = RelTol[0];
  d.atol = AbsTol[0];
  d.opts = &o;
  if (autonomous) d.jrate_end = nullptr;  // constant J without the family kernel
  int rc = radm2_integrate_batched(c, 1, &d);
  std::vector<int32_t> hns(n), hna(n), hnr(n), hst(n);
  std::vector<double> hhx(n);
  if (rc == 0 && hipStreamSynchronize(c->stream) == hipSuccess &&
      hipMemcpy(hns.data(), ns, n * 4, hipMemcpyDeviceToHost) == hipSuccess &&
      hipMemcpy(hna.data(), na, n * 4, hipMemcpyDeviceToHost) == hipSuccess &&
      hipMemcpy(hnr.data(), nr, n * 4, hipMemcpyDeviceToHost) == hipSuccess &&
      hipMemcpy(hst.data(), st, n * 4, hipMemcpyDeviceToHost) == hipSuccess &&
      hipMemcpy(hhx.data(), hx, n * 8, hipMemcpyDeviceToHost) == hipSuccess) {
    long long nstp = 0, nacc = 0, nrej = 0, nsng = 0;
    int worst = 1;
    double hmax_exit = 0.0;
    for (int64_t p = 0; p < n; p++) {
      nstp += hns[p];
      nacc += hna[p];
      nrej += hnr[p];
      if (hst[p] == RADM2_CELL_SINGULAR) nsng++;
      if (hst[p] != 1 && worst == 1) worst = hst[p];
      hmax_exit = std::max(hmax_exit, hhx[p]);
    }
    // KPP accounting: one Fun+Jac per time step (accepted attempt), and one
    // more Fun for dF/dT when non-autonomous; per attempt one Decomp, one
    // Solve per stage and one Fun per stage after the first that evaluates a
    // new F (Ros2 1, Ros3 1, Ros4 2, Rodas3 2, Rodas4 5)
    static const int kNewF[6] = {0, 1, 1, 2, 2, 5}, kStages[6] = {0, 2, 3, 4, 4, 6};
    IPAR[10] += (int32_t)((autonomous ? 1 : 2) * nacc + kNewF[ro.method] * nstp);
    IPAR[11] += (int32_t)nacc;
    IPAR[12] += (int32_t)nstp;
    IPAR[13] += (int32_t)nacc;
    IPAR[14] += (int32_t)nrej;
    IPAR[15] += (int32_t)nstp;
    IPAR[16] += (int32_t)(kStages[ro.method] * nstp);
    IPAR[17] += (int32_t)nsng;
    RPAR[10] = worst == 1 ? Tend : Tstart;
    RPAR[11] = hmax_exit;
    rc = worst;
  } else if (rc == 0) {
    rc = RADM2_E_HIP;
  }
  hipFree(buf);
  return rc;
}

const char* radm2_strerror(int code) {
  switch (code) {
    case 0: return "ok";
    case 1: return "integration successful";
    case -1: return "Improper value for maximal no of steps";
    case -2: return "Selected Rosenbrock method not implemented";
    case -3: return "Hmin/Hmax/Hstart must be positive";
    case -4: return "FacMin/FacMax/FacRej must be positive";
    case -5: return "Improper tolerance values";
    case -6: return "No of steps exceeds maximum bound";
    case -7: return "Step size too small (T + H/10 = T) or H < Roundoff";
    case -8: return "Matrix is repeatedly singular";
    case RADM2_E_ARG: return "invalid argument";
    case RADM2_E_HIP: return hipGetErrorString(last_hip_error());
    case RADM2_E_UNSUPPORTED: return "option not supported by this solver";
    default: return "unknown error code";
  }
}

}  // extern "C"
