// C-ABI implementation: contexts, plans and the host orchestration of the
// pencil passes and the coil stage.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/disorder.h"
#include "lanetab.hpp"
#include "kernels.hpp"
#include "transform.hpp"

using namespace dis;

// ---------------------------------------------------------------- context
struct dis_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  std::map<int, float2*> tw;
  void* ws = nullptr;
  size_t ws_size = 0;
  size_t ws_limit = 0;
  void* dcoef = nullptr;
  size_t dcoef_cap = 0;
  std::map<int, std::pair<void*, size_t>> bufs;
  // upload tape: while recording every staged upload gets its own persistent
  // buffer; while replaying the same sequence of uploads returns those
  // buffers without copies or synchronisation (for graph capture)
  std::map<std::array<int, 4>, PassTune> tune;  // tuned pass shapes per (n1, n2, n3, axis)
  int coil_schedule = 0;  // 0 auto, 1 static grid, 2 persistent work queue
  long long cache_budget = 208ll << 20;  // bytes of state volumes a transform keeps in flight (0 = all states)
  int tape_mode = 0;  // 0 off, 1 record, 2 replay
  // pinned host staging of the coefficient uploads: a ring of slots, each
  // reused once the event recorded behind its last copy has completed, so an
  // upload does not wait for the stream (the device buffer dcoef is rewritten
  // in stream order, after the kernels that read its previous contents)
  struct PinSlot {
    void* p = nullptr;
    size_t cap = 0;
    hipEvent_t ev = nullptr;
    bool used = false;
  };
  std::array<PinSlot, 4> pins;
  int pin_next = 0;
  std::vector<std::pair<void*, size_t>> tape;
  size_t tape_pos = 0;
  // profiling
  bool prof = false;
  struct Rec {
    int kclass;
    hipEvent_t a, b;
    double bytes;
    int launches;
  };
  std::vector<Rec> recs;
  std::vector<hipEvent_t> pool;
  int group_depth = 0;  // nesting of ProfGroup scopes
  int open = -1;        // record of a group whose end event is not yet recorded
};

struct dis_plan {
  dis_ctx* ctx = nullptr;
  dis_geom g{};
  const float2* S = nullptr;
  int* k1 = nullptr;
  int* k2 = nullptr;
  int* off = nullptr;
  float* We = nullptr;
  int* rowptr = nullptr;  // [S][n1+1] per-state CSR by k1
  int* csr = nullptr;     // [Ns]
  int* sord = nullptr;    // [Ns] register coil stage: sample of each slot
  int* adjtab = nullptr;  // register-path adjoint lane table: blocks of kThreads int4, or null
  int* adjoff = nullptr;  // [S+1] first table block of each state
  std::vector<int> hoff;
  int fold[2] = {0, 0};   // acquired phase-encode grid under SENSE folding
  int roi[2] = {0, 0};    // readout range [lo, hi) of the LM sums (E.nF); {0, 0} = all
  float* Fd = nullptr;    // [Ns, fdn] fractional finite-difference weights of LMsolver (E.Fd), or null
  int fdn = 0;
};

namespace {

int fail(dis_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

#define HIPCHK(ctx, expr)                                                                       \
  do {                                                                                          \
    hipError_t e_ = (expr);                                                                     \
    if (e_ != hipSuccess)                                                                       \
      return fail((ctx), DIS_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));        \
  } while (0)

}  // namespace

namespace dis {

int ctx_fail(dis_ctx* c, int code, const char* msg) { return fail(c, code, msg); }

hipStream_t ctx_stream(dis_ctx* c) { return c->stream; }

hipError_t ctx_set_device(dis_ctx* c) { return hipSetDevice(c->device); }

int ctx_twiddles(dis_ctx* c, int n, const float2** out) {
  auto it = c->tw.find(n);
  if (it != c->tw.end()) {
    *out = it->second;
    return DIS_OK;
  }
  float2* p = nullptr;
  HIPCHK(c, hipMalloc(&p, sizeof(float2) * n));
  HIPCHK(c, launch_twiddles(p, n, c->stream));
  c->tw[n] = p;
  *out = p;
  return DIS_OK;
}

int ctx_workspace(dis_ctx* c, size_t bytes, void** out) {
  if (c->ws_limit && bytes > c->ws_limit) return fail(c, DIS_E_NOMEM, "workspace request above the limit");
  if (bytes > c->ws_size) {
    if (c->ws) {
      HIPCHK(c, hipStreamSynchronize(c->stream));
      HIPCHK(c, hipFree(c->ws));
      c->ws = nullptr;
      c->ws_size = 0;
    }
    if (hipMalloc(&c->ws, bytes) != hipSuccess) {
      c->ws = nullptr;
      (void)hipGetLastError();
      return fail(c, DIS_E_NOMEM, "hipMalloc of the workspace failed");
    }
    c->ws_size = bytes;
  }
  *out = c->ws;
  return DIS_OK;
}

size_t ctx_workspace_limit(dis_ctx* c) { return c->ws_limit; }

int ctx_buffer(dis_ctx* c, int slot, size_t bytes, void** out) {
  auto& b = c->bufs[slot];
  if (bytes > b.second) {
    if (b.first) {
      HIPCHK(c, hipStreamSynchronize(c->stream));
      HIPCHK(c, hipFree(b.first));
      b.first = nullptr;
      b.second = 0;
    }
    if (hipMalloc(&b.first, bytes) != hipSuccess) {
      b.first = nullptr;
      (void)hipGetLastError();
      return fail(c, DIS_E_NOMEM, "hipMalloc of a solver buffer failed");
    }
    b.second = bytes;
  }
  *out = b.first;
  return DIS_OK;
}

// Upload the staged bytes into the device coefficient buffer (grown on
// demand).  Stream-ordered after all earlier work; synchronous on return, so
// the buffer may be reused by the next upload.
int ctx_upload_bytes(dis_ctx* c, const Staging& stg, const unsigned char** dptr) {
  const size_t nbytes = std::max<size_t>(stg.bytes.size(), 16);
  if (c->tape_mode == 1) {
    void* p = nullptr;
    HIPCHK(c, hipMalloc(&p, nbytes));
    c->tape.emplace_back(p, nbytes);
    if (!stg.bytes.empty()) HIPCHK(c, hipMemcpy(p, stg.bytes.data(), stg.bytes.size(), hipMemcpyHostToDevice));
    *dptr = (const unsigned char*)p;
    return DIS_OK;
  }
  if (c->tape_mode == 2) {
    if (c->tape_pos >= c->tape.size() || c->tape[c->tape_pos].second != nbytes)
      return fail(c, DIS_E_INVALID, "upload tape replay diverged from the recording");
    *dptr = (const unsigned char*)c->tape[c->tape_pos++].first;
    return DIS_OK;
  }
  if (nbytes > c->dcoef_cap) {
    if (c->dcoef) {
      HIPCHK(c, hipStreamSynchronize(c->stream));
      HIPCHK(c, hipFree(c->dcoef));
      c->dcoef = nullptr;
    }
    const size_t cap = std::max<size_t>(nbytes, 1 << 16);
    HIPCHK(c, hipMalloc(&c->dcoef, cap));
    c->dcoef_cap = cap;
  }
  if (!stg.bytes.empty()) {
    dis_ctx::PinSlot& ps = c->pins[c->pin_next];
    c->pin_next = (c->pin_next + 1) % (int)c->pins.size();
    if (ps.used) HIPCHK(c, hipEventSynchronize(ps.ev));  // the slot's previous copy has read it
    if (ps.cap < stg.bytes.size()) {
      if (ps.p) HIPCHK(c, hipHostFree(ps.p));
      ps.p = nullptr;
      ps.cap = 0;
      const size_t cap = std::max<size_t>(stg.bytes.size(), 1 << 16);
      HIPCHK(c, hipHostMalloc(&ps.p, cap));
      ps.cap = cap;
    }
    if (!ps.ev) HIPCHK(c, hipEventCreateWithFlags(&ps.ev, hipEventDisableTiming));
    std::memcpy(ps.p, stg.bytes.data(), stg.bytes.size());
    HIPCHK(c, hipMemcpyAsync(c->dcoef, ps.p, stg.bytes.size(), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipEventRecord(ps.ev, c->stream));
    ps.used = true;
  }
  *dptr = (const unsigned char*)c->dcoef;
  return DIS_OK;
}

bool ctx_tune_get(dis_ctx* c, const int n[3], int axis, PassTune* t) {
  auto it = c->tune.find({n[0], n[1], n[2], axis});
  if (it == c->tune.end()) return false;
  *t = it->second;
  return true;
}

void ctx_tune_set(dis_ctx* c, const int n[3], int axis, const PassTune& t) { c->tune[{n[0], n[1], n[2], axis}] = t; }

void ctx_tape_mode(dis_ctx* c, int mode) {
  c->tape_mode = mode;
  c->tape_pos = 0;
}

void ctx_tape_free(dis_ctx* c) {
  (void)hipStreamSynchronize(c->stream);
  for (auto& b : c->tape) (void)hipFree(b.first);
  c->tape.clear();
  c->tape_pos = 0;
  c->tape_mode = 0;
}

bool ctx_profiling(dis_ctx* c, int on) {
  const bool was = c->prof;
  if (on >= 0) c->prof = on != 0;
  return was;
}

int ctx_device(dis_ctx* c) { return c->device; }

int ctx_coil_schedule(dis_ctx* c) { return c->coil_schedule; }
long long ctx_cache_budget(dis_ctx* c) { return c->cache_budget; }

namespace {
void prof_close(dis_ctx* c) {
  if (c->open < 0) return;
  (void)hipEventRecord(c->recs[c->open].b, c->stream);
  c->open = -1;
}
}  // namespace

ProfGroup::ProfGroup(dis_ctx* ctx) : c(ctx) { ++c->group_depth; }

ProfGroup::~ProfGroup() {
  if (--c->group_depth == 0) prof_close(c);
}

int prof_begin(dis_ctx* c, int kclass) {
  if (!c->prof) return -1;
  if (c->open >= 0) {
    if (c->recs[c->open].kclass == kclass) {
      ++c->recs[c->open].launches;
      return c->open;
    }
    prof_close(c);
  }
  auto get = [&]() {
    hipEvent_t e = nullptr;
    if (!c->pool.empty()) {
      e = c->pool.back();
      c->pool.pop_back();
    } else if (hipEventCreate(&e) != hipSuccess) {
      e = nullptr;
    }
    return e;
  };
  dis_ctx::Rec r;
  r.kclass = kclass;
  r.a = get();
  r.b = get();
  r.bytes = 0;
  r.launches = 1;
  if (!r.a || !r.b) return -1;
  (void)hipEventRecord(r.a, c->stream);
  c->recs.push_back(r);
  return (int)c->recs.size() - 1;
}

void prof_end(dis_ctx* c, int slot, double bytes) {
  if (slot < 0 || !c->prof) return;
  c->recs[slot].bytes += bytes;
  if (c->group_depth > 0)
    c->open = slot;  // the end event waits for the group's last launch
  else
    (void)hipEventRecord(c->recs[slot].b, c->stream);
}

}  // namespace dis

// ---------------------------------------------------------------- C-ABI
extern "C" {

int dis_version(void) { return 100; }

const char* dis_status_string(int s) {
  switch (s) {
    case DIS_OK: return "ok";
    case DIS_E_INVALID: return "invalid argument";
    case DIS_E_UNSUPPORTED: return "unsupported";
    case DIS_E_NOMEM: return "out of device memory";
    case DIS_E_HIP: return "HIP runtime error";
    case DIS_E_NODEVICE: return "no device";
    default: return "unknown status";
  }
}

int dis_ctx_create(int device, dis_ctx** out) {
  if (!out) return DIS_E_INVALID;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
    (void)hipGetLastError();
    return DIS_E_NODEVICE;
  }
  if (device < 0 || device >= ndev) return DIS_E_INVALID;
  if (hipSetDevice(device) != hipSuccess) return DIS_E_HIP;
  // device address of the allocation-free Gram's static scratch (once per device)
  if (dis::coil_gram_scratch_init(device) != DIS_OK) return DIS_E_HIP;
  dis_ctx* c = new dis_ctx();
  c->device = device;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return DIS_E_HIP;
  }
  *out = c;
  return DIS_OK;
}

int dis_ctx_destroy(dis_ctx* c) {
  if (!c) return DIS_E_INVALID;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  for (auto& kv : c->tw) (void)hipFree(kv.second);
  for (auto& r : c->recs) {
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  for (auto e : c->pool) (void)hipEventDestroy(e);
  if (c->ws) (void)hipFree(c->ws);
  for (auto& kv : c->bufs)
    if (kv.second.first) (void)hipFree(kv.second.first);
  if (c->dcoef) (void)hipFree(c->dcoef);
  for (auto& ps : c->pins) {
    if (ps.p) (void)hipHostFree(ps.p);
    if (ps.ev) (void)hipEventDestroy(ps.ev);
  }
  for (auto& b : c->tape) (void)hipFree(b.first);
  (void)hipStreamDestroy(c->stream);
  delete c;
  return DIS_OK;
}

const char* dis_last_error(const dis_ctx* c) { return c ? c->err.c_str() : "null context"; }

int dis_sync(dis_ctx* c) {
  if (!c) return DIS_E_INVALID;
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return DIS_OK;
}

int dis_set_workspace_limit(dis_ctx* c, size_t bytes) {
  if (!c) return DIS_E_INVALID;
  c->ws_limit = bytes;
  return DIS_OK;
}

int dis_profile_enable(dis_ctx* c, int32_t on) {
  if (!c) return DIS_E_INVALID;
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (auto& r : c->recs) {
    c->pool.push_back(r.a);
    c->pool.push_back(r.b);
  }
  c->recs.clear();
  c->open = -1;
  c->prof = on != 0;
  return DIS_OK;
}

int dis_profile_read(dis_ctx* c, int32_t kclass, int64_t* launches, double* ms_total, double* bytes_total) {
  if (!c) return DIS_E_INVALID;
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  int64_t n = 0;
  double ms = 0, by = 0;
  for (auto& r : c->recs) {
    if (r.kclass != kclass) continue;
    float t = 0;
    HIPCHK(c, hipEventElapsedTime(&t, r.a, r.b));
    ms += t;
    by += r.bytes;
    n += r.launches;
  }
  if (launches) *launches = n;
  if (ms_total) *ms_total = ms;
  if (bytes_total) *bytes_total = by;
  return DIS_OK;
}

int dis_rigid_transform(dis_ctx* c, const int32_t n[3], const int32_t Nfull[3], const double* T, int32_t S,
                        int32_t di, int32_t sumR, int32_t x_per_state, const void* x, void* out) {
  if (!c) return DIS_E_INVALID;
  if (!n || !Nfull || !T || !x || !out || S <= 0 || (di != 0 && di != 1))
    return fail(c, DIS_E_INVALID, "dis_rigid_transform: invalid argument");
  HIPCHK(c, hipSetDevice(c->device));
  Grid gr;
  int st = make_grid(c, n, Nfull, &gr);
  if (st) return st;
  std::vector<double> Th(T, T + 6 * (size_t)S);
  if (di == 1) {
    st = rigid_forward(c, gr, Th.data(), S, S, (const float2*)x, x_per_state ? gr.nvox : 0, (float2*)out, false);
  } else {
    float2* work = (float2*)out;
    if (sumR) {
      void* ws = nullptr;
      st = ctx_workspace(c, sizeof(float2) * gr.nvox * S, &ws);
      if (st) return st;
      work = (float2*)ws;
    }
    st = rigid_inverse(c, gr, Th.data(), S, S, (const float2*)x, gr.nvox, work, (float2*)out, sumR != 0, false,
                       false);
  }
  if (st) return st;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return DIS_OK;
}

int dis_plan_create(dis_ctx* c, const dis_geom* g, const void* S_maps, const int32_t* k1, const int32_t* k2,
                    const int32_t* state_off, const float* We, dis_plan** out) {
  if (!c) return DIS_E_INVALID;
  if (!g || !S_maps || !state_off || !out || (g->nsamples > 0 && (!k1 || !k2)))
    return fail(c, DIS_E_INVALID, "dis_plan_create: null argument");
  *out = nullptr;
  if (g->ncoils <= 0 || g->nstates <= 0 || g->nmotion < 0 || g->nmotion > g->nstates || g->nsamples < 0)
    return fail(c, DIS_E_INVALID, "dis_plan_create: bad geometry");
  for (int d = 0; d < 3; ++d)
    if (g->n[d] <= 0 || g->Nfull[d] <= 0) return fail(c, DIS_E_INVALID, "dis_plan_create: bad grid size");
  if (state_off[0] != 0 || state_off[g->nstates] != g->nsamples)
    return fail(c, DIS_E_INVALID, "dis_plan_create: state offsets must start at 0 and end at nsamples");
  for (int s = 0; s < g->nstates; ++s)
    if (state_off[s + 1] < state_off[s]) return fail(c, DIS_E_INVALID, "dis_plan_create: offsets not monotone");
  for (int j = 0; j < g->nsamples; ++j)
    if (k1[j] < 0 || k1[j] >= g->n[0] || k2[j] < 0 || k2[j] >= g->n[1])
      return fail(c, DIS_E_INVALID, "dis_plan_create: sample index outside the phase-encode grid");
  Grid gr;
  int st = make_grid(c, g->n, g->Nfull, &gr);
  if (st) return st;
  HIPCHK(c, hipSetDevice(c->device));
  dis_plan* p = new dis_plan();
  p->ctx = c;
  p->g = *g;
  p->S = (const float2*)S_maps;
  p->hoff.assign(state_off, state_off + g->nstates + 1);
  const size_t ns = std::max(1, g->nsamples);
  bool ok = hipMalloc(&p->k1, ns * sizeof(int)) == hipSuccess && hipMalloc(&p->k2, ns * sizeof(int)) == hipSuccess &&
            hipMalloc(&p->off, (g->nstates + 1) * sizeof(int)) == hipSuccess;
  if (ok && g->nsamples > 0) {
    ok = hipMemcpy(p->k1, k1, g->nsamples * sizeof(int), hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(p->k2, k2, g->nsamples * sizeof(int), hipMemcpyHostToDevice) == hipSuccess;
  }
  if (ok) ok = hipMemcpy(p->off, state_off, (g->nstates + 1) * sizeof(int), hipMemcpyHostToDevice) == hipSuccess;
  // per-state CSR of the samples by phase-encode row k1 (stable in sample
  // order): rowptr[s][k1], entries (k2 << 16) | (j - off[s])
  if (ok) {
    const int n1 = g->n[0];
    std::vector<int> rp((size_t)g->nstates * (n1 + 1), 0), ent(std::max(1, g->nsamples));
    for (int s = 0; s < g->nstates; ++s) {
      const int o0 = state_off[s], o1 = state_off[s + 1];
      if (o1 - o0 > 65535) {
        ok = false;
        break;
      }
      int* r = rp.data() + (size_t)s * (n1 + 1);
      for (int j = o0; j < o1; ++j) ++r[k1[j] + 1];
      for (int k = 0; k < n1; ++k) r[k + 1] += r[k];
      std::vector<int> fill(r, r + n1);
      for (int j = o0; j < o1; ++j) ent[o0 + fill[k1[j]]++] = (k2[j] << 16) | (j - o0);
    }
    if (!ok) {
      dis_plan_destroy(p);
      return fail(c, DIS_E_UNSUPPORTED, "dis_plan_create: more than 65535 samples in one state");
    }
    ok = hipMalloc(&p->rowptr, rp.size() * sizeof(int)) == hipSuccess &&
         hipMalloc(&p->csr, ent.size() * sizeof(int)) == hipSuccess &&
         hipMemcpy(p->rowptr, rp.data(), rp.size() * sizeof(int), hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(p->csr, ent.data(), ent.size() * sizeof(int), hipMemcpyHostToDevice) == hipSuccess;
    std::vector<int> tab, toff;
    if (ok && adjoint_lane_table(g->n[0], g->n[1], g->nstates, state_off, rp, ent, &tab, &toff))
      ok = hipMalloc(&p->adjtab, tab.size() * sizeof(int)) == hipSuccess &&
           hipMalloc(&p->adjoff, toff.size() * sizeof(int)) == hipSuccess &&
           hipMemcpy(p->adjtab, tab.data(), tab.size() * sizeof(int), hipMemcpyHostToDevice) == hipSuccess &&
           hipMemcpy(p->adjoff, toff.data(), toff.size() * sizeof(int), hipMemcpyHostToDevice) == hipSuccess;
  }
  // slot -> sample table of the register coil stage: per state and chunk of
  // kCoilRegChunk samples, slot q takes a sample with k1 = q (mod 32) while
  // one is left (else one from the fullest residue), so the 16- and 32-lane
  // groups of a wave gather their phase-encode rows from distinct LDS banks
  if (ok) {
    std::vector<int> so(std::max(1, g->nsamples));
    for (int s = 0; s < g->nstates; ++s) {
      for (int c0 = state_off[s]; c0 < state_off[s + 1]; c0 += kCoilRegChunk) {
        const int m = std::min(kCoilRegChunk, state_off[s + 1] - c0);
        std::vector<int> bucket[32];
        for (int j = c0; j < c0 + m; ++j) bucket[k1[j] & 31].push_back(j);
        size_t head[32] = {0};
        for (int q = 0; q < m; ++q) {
          int r = q & 31;
          if (head[r] >= bucket[r].size()) {
            size_t best = 0;
            for (int b = 0; b < 32; ++b)
              if (bucket[b].size() - head[b] > best) {
                best = bucket[b].size() - head[b];
                r = b;
              }
          }
          so[c0 + q] = bucket[r][head[r]++];
        }
      }
    }
    ok = hipMalloc(&p->sord, so.size() * sizeof(int)) == hipSuccess &&
         hipMemcpy(p->sord, so.data(), so.size() * sizeof(int), hipMemcpyHostToDevice) == hipSuccess;
  }
  if (!ok) {
    dis_plan_destroy(p);
    return fail(c, DIS_E_NOMEM, "dis_plan_create: device allocation failed");
  }
  *out = p;
  if (We) return dis_plan_set_weights(p, We);
  return DIS_OK;
}

int dis_plan_destroy(dis_plan* p) {
  if (!p) return DIS_E_INVALID;
  if (p->ctx) (void)hipSetDevice(p->ctx->device);
  if (p->k1) (void)hipFree(p->k1);
  if (p->k2) (void)hipFree(p->k2);
  if (p->off) (void)hipFree(p->off);
  if (p->We) (void)hipFree(p->We);
  if (p->rowptr) (void)hipFree(p->rowptr);
  if (p->csr) (void)hipFree(p->csr);
  if (p->sord) (void)hipFree(p->sord);
  if (p->adjtab) (void)hipFree(p->adjtab);
  if (p->adjoff) (void)hipFree(p->adjoff);
  if (p->Fd) (void)hipFree(p->Fd);
  delete p;
  return DIS_OK;
}

int dis_plan_set_fold(dis_plan* p, int32_t ny1, int32_t ny2) {
  if (!p) return DIS_E_INVALID;
  dis_ctx* c = p->ctx;
  const dis_geom& g = p->g;
  if (ny1 == g.n[0] && ny2 == g.n[1]) {
    p->fold[0] = p->fold[1] = 0;
    return DIS_OK;
  }
  if (ny1 <= 0 || ny2 <= 0 || ny1 > g.n[0] || ny2 > g.n[1])
    return fail(c, DIS_E_INVALID, "dis_plan_set_fold: the acquired grid must be within the reconstruction grid");
  FftPlan pl;
  if (ny1 > 1024 || ny2 > 1024 || g.n[0] > 1024 || make_plan(ny1, &pl))
    return fail(c, DIS_E_UNSUPPORTED, "dis_plan_set_fold: unsupported acquired length");
  if (g.nsamples > 0) {
    HIPCHK(c, hipSetDevice(c->device));
    std::vector<int> h1(g.nsamples), h2(g.nsamples);
    HIPCHK(c, hipMemcpy(h1.data(), p->k1, sizeof(int) * g.nsamples, hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(h2.data(), p->k2, sizeof(int) * g.nsamples, hipMemcpyDeviceToHost));
    for (int j = 0; j < g.nsamples; ++j)
      if (h1[j] >= ny1 || h2[j] >= ny2)
        return fail(c, DIS_E_INVALID, "dis_plan_set_fold: sample index outside the acquired grid");
  }
  p->fold[0] = ny1;
  p->fold[1] = ny2;
  return DIS_OK;
}

int dis_plan_set_readout_roi(dis_plan* p, int32_t lo, int32_t hi) {
  if (!p) return DIS_E_INVALID;
  if (lo == 0 && (hi == 0 || hi == p->g.n[2])) {
    p->roi[0] = p->roi[1] = 0;
    return DIS_OK;
  }
  if (lo < 0 || hi <= lo || hi > p->g.n[2])
    return fail(p->ctx, DIS_E_INVALID, "dis_plan_set_readout_roi: need 0 <= lo < hi <= n3");
  p->roi[0] = lo;
  p->roi[1] = hi;
  return DIS_OK;
}

int dis_plan_set_lm_filter(dis_plan* p, const float* Fd, int32_t nF) {
  if (!p) return DIS_E_INVALID;
  dis_ctx* c = p->ctx;
  HIPCHK(c, hipSetDevice(c->device));
  if (!Fd) {
    if (p->Fd) HIPCHK(c, hipFree(p->Fd));
    p->Fd = nullptr;
    p->fdn = 0;
    return DIS_OK;
  }
  if (nF <= 0 || nF > p->g.n[2])
    return fail(c, DIS_E_INVALID, "dis_plan_set_lm_filter: need 0 < nF <= n3");
  const size_t cnt = (size_t)std::max(1, p->g.nsamples) * nF;
  if (p->Fd && p->fdn != nF) {
    HIPCHK(c, hipFree(p->Fd));
    p->Fd = nullptr;
  }
  if (!p->Fd) HIPCHK(c, hipMalloc(&p->Fd, cnt * sizeof(float)));
  p->fdn = nF;
  if (p->g.nsamples > 0)
    HIPCHK(c, hipMemcpy(p->Fd, Fd, (size_t)p->g.nsamples * nF * sizeof(float), hipMemcpyHostToDevice));
  return DIS_OK;
}

int dis_plan_set_weights(dis_plan* p, const float* We) {
  if (!p) return DIS_E_INVALID;
  dis_ctx* c = p->ctx;
  HIPCHK(c, hipSetDevice(c->device));
  if (!We) {
    if (p->We) HIPCHK(c, hipFree(p->We));
    p->We = nullptr;
    return DIS_OK;
  }
  if (!p->We) HIPCHK(c, hipMalloc(&p->We, std::max(1, p->g.nsamples) * sizeof(float)));
  if (p->g.nsamples > 0) HIPCHK(c, hipMemcpy(p->We, We, p->g.nsamples * sizeof(float), hipMemcpyHostToDevice));
  return DIS_OK;
}

}  // extern "C"

// ---------------------------------------------------------------- operators
namespace dis {

OpView plan_view(dis_plan* p) {
  OpView v;
  v.S = p->S;
  v.k1 = p->k1;
  v.k2 = p->k2;
  v.off = p->off;
  v.We = p->We;
  v.rowptr = p->rowptr;
  v.csr = p->csr;
  v.sord = p->sord;
  v.adjtab = p->adjtab;
  v.adjoff = p->adjoff;
  v.g = p->g;
  v.hoff = p->hoff.data();
  v.fold[0] = p->fold[0];
  v.fold[1] = p->fold[1];
  v.roi[0] = p->roi[0];
  v.roi[1] = p->roi[1];
  v.Fd = p->Fd;
  v.fdn = p->fdn;
  return v;
}

}  // namespace dis

extern "C" {

int dis_encode(dis_plan* p, const double* T, const void* x, void* y) {
  if (!p) return DIS_E_INVALID;
  dis_ctx* c = p->ctx;
  if (!x || !y || (p->g.nmotion > 0 && !T)) return fail(c, DIS_E_INVALID, "dis_encode: null argument");
  HIPCHK(c, hipSetDevice(c->device));
  int st = op_apply(c, plan_view(p), T, (const float2*)x, nullptr, (float2*)y, nullptr, COIL_ENCODE, nullptr);
  if (st) return st;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return DIS_OK;
}

int dis_decode(dis_plan* p, const double* T, const void* y, void* x) {
  if (!p) return DIS_E_INVALID;
  dis_ctx* c = p->ctx;
  if (!x || !y || (p->g.nmotion > 0 && !T)) return fail(c, DIS_E_INVALID, "dis_decode: null argument");
  HIPCHK(c, hipSetDevice(c->device));
  int st = op_apply(c, plan_view(p), T, nullptr, (const float2*)y, nullptr, (float2*)x, COIL_DECODE, nullptr);
  if (st) return st;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return DIS_OK;
}

int dis_encodedecode(dis_plan* p, const double* T, const void* x, void* out) {
  if (!p) return DIS_E_INVALID;
  dis_ctx* c = p->ctx;
  if (!x || !out || (p->g.nmotion > 0 && !T)) return fail(c, DIS_E_INVALID, "dis_encodedecode: null argument");
  HIPCHK(c, hipSetDevice(c->device));
  int st = op_apply(c, plan_view(p), T, (const float2*)x, nullptr, nullptr, (float2*)out, COIL_EHE, nullptr);
  if (st) return st;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return DIS_OK;
}

int dis_energy(dis_plan* p, const double* T, const void* x, const void* y, const float* Mbe, double* per_sample,
               double* total) {
  if (!p) return DIS_E_INVALID;
  dis_ctx* c = p->ctx;
  if (!x || !y || (p->g.nmotion > 0 && !T)) return fail(c, DIS_E_INVALID, "dis_energy: null argument");
  HIPCHK(c, hipSetDevice(c->device));
  EnergyOut eo;
  eo.Mbe = Mbe;
  eo.per_sample = per_sample;
  eo.total = total;
  int st = op_apply(c, plan_view(p), T, (const float2*)x, (const float2*)y, nullptr, nullptr, COIL_ENERGY, &eo);
  if (st) return st;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return DIS_OK;
}

}  // extern "C"

extern "C" int dis_cg(dis_plan* p, const double* T, const void* y, const float* Se, const float* Ma, double lambda,
                      void* x, int32_t x_init, int32_t nIt, int32_t tolType, double tol, double* En, int32_t* nEn,
                      int32_t* nn) {
  if (!p) return DIS_E_INVALID;
  dis_ctx* c = p->ctx;
  if (!x || !y || nIt < 0 || (p->g.nmotion > 0 && !T)) return fail(c, DIS_E_INVALID, "dis_cg: invalid argument");
  HIPCHK(c, hipSetDevice(c->device));
  int ne = 0, it = 0;
  int st = run_cg(c, plan_view(p), T, (const float2*)y, Se, Ma, lambda, (float2*)x, x_init != 0, nIt, tolType, tol,
                  En, &ne, &it);
  if (st) return st;
  if (nEn) *nEn = ne;
  if (nn) *nn = it;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return DIS_OK;
}

extern "C" int dis_cg_cs(dis_plan* p, const double* T, const void* y, const float* Se, const float* Ma, double lambda,
                         dis_shearlet* sh, double sh_scale, void* x, int32_t x_init, int32_t nIt, int32_t tolType,
                         double tol, int32_t* nn) {
  if (!p) return DIS_E_INVALID;
  dis_ctx* c = p->ctx;
  if (!x || !y || !sh || nIt < 0 || (p->g.nmotion > 0 && !T)) return fail(c, DIS_E_INVALID, "dis_cg_cs: invalid argument");
  HIPCHK(c, hipSetDevice(c->device));
  int it = 0;
  int st = run_cg(c, plan_view(p), T, (const float2*)y, Se, Ma, lambda, (float2*)x, x_init != 0, nIt, tolType, tol,
                  nullptr, nullptr, &it, sh, sh_scale);
  if (st) return st;
  if (nn) *nn = it;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return DIS_OK;
}

extern "C" int dis_autotune(dis_ctx* c, const int32_t n[3], const int32_t Nfull[3], int32_t S, int32_t* choice,
                            double* ms) {
  if (!c) return DIS_E_INVALID;
  if (!n || S <= 0) return fail(c, DIS_E_INVALID, "dis_autotune: invalid argument");
  HIPCHK(c, hipSetDevice(c->device));
  const int32_t* Nf = Nfull ? Nfull : n;
  return autotune_grid(c, n, Nf, S, choice, ms);
}

extern "C" int dis_set_pass_shape(dis_ctx* c, const int32_t n[3], int32_t axis, int32_t generic, int32_t P,
                                  int32_t maxradix) {
  if (!c) return DIS_E_INVALID;
  if (!n || axis < 0 || axis > 2 || P < 0 || (maxradix != 4 && maxradix != 8 && maxradix != 16))
    return fail(c, DIS_E_INVALID, "dis_set_pass_shape: invalid argument");
  if (P && (P & (P - 1))) return fail(c, DIS_E_INVALID, "dis_set_pass_shape: P must be a power of two");
  PassTune t;
  t.valid = 1;
  t.generic = generic ? 1 : 0;
  t.P = P;
  t.maxradix = maxradix;
  const int nn[3] = {n[0], n[1], n[2]};
  ctx_tune_set(c, nn, axis, t);
  return DIS_OK;
}

extern "C" int dis_set_cache_budget(dis_ctx* c, int64_t bytes) {
  if (!c) return DIS_E_INVALID;
  if (bytes < 0) return fail(c, DIS_E_INVALID, "dis_set_cache_budget: bytes must be >= 0");
  c->cache_budget = bytes;
  return DIS_OK;
}

extern "C" int dis_set_coil_schedule(dis_ctx* c, int32_t mode) {
  if (!c) return DIS_E_INVALID;
  if (mode < 0 || mode > 2) return fail(c, DIS_E_INVALID, "dis_set_coil_schedule: mode must be 0, 1 or 2");
  c->coil_schedule = mode;
  return DIS_OK;
}

extern "C" int dis_autotune_clear(dis_ctx* c) {
  if (!c) return DIS_E_INVALID;
  c->tune.clear();
  return DIS_OK;
}

extern "C" int dis_cg_batched(dis_ctx* c, int32_t nvol, dis_plan* const* plans, const double* const* T,
                              const void* const* y, const float* const* Se, void* const* x, int32_t nIt,
                              int32_t use_graph, double* res, double* ms) {
  if (!c) return DIS_E_INVALID;
  if (nvol <= 0 || !plans || !T || !y || !x || nIt < 0) return fail(c, DIS_E_INVALID, "dis_cg_batched: invalid argument");
  std::vector<OpView> views(nvol);
  for (int i = 0; i < nvol; ++i) {
    if (!plans[i] || plans[i]->ctx != c || !y[i] || !x[i] || (plans[i]->g.nmotion > 0 && !T[i]))
      return fail(c, DIS_E_INVALID, "dis_cg_batched: every plan must belong to the context and have y, x and T");
    views[i] = plan_view(plans[i]);
  }
  HIPCHK(c, hipSetDevice(c->device));
  return run_cg_batched(c, nvol, views.data(), T, (const float2* const*)y, Se, (float2* const*)x, nIt, use_graph != 0,
                        res, ms);
}

extern "C" int dis_rigid_transform_grad(dis_ctx* c, const int32_t n[3], const int32_t Nfull[3], const double* T,
                                        int32_t S, const void* x, void* Tx, void* G) {
  if (!c) return DIS_E_INVALID;
  if (!n || !Nfull || !T || !x || !G || S <= 0) return fail(c, DIS_E_INVALID, "dis_rigid_transform_grad: invalid argument");
  HIPCHK(c, hipSetDevice(c->device));
  Grid gr;
  int st = make_grid(c, n, Nfull, &gr);
  if (st) return st;
  const size_t vol = sizeof(float2) * gr.nvox;
  void* ws = nullptr;
  st = ctx_workspace(c, vol * (size_t)S * (Tx ? 5 : 6), &ws);
  if (st) return st;
  float2* X0 = (float2*)ws;
  float2* Y = X0 + 3 * gr.nvox * S;
  float2* Txp = Tx ? (float2*)Tx : Y + 2 * gr.nvox * S;
  std::vector<double> Th(T, T + 6 * (size_t)S);
  st = rigid_gradient(c, gr, Th.data(), S, S, (const float2*)x, Txp, (float2*)G, X0, Y);
  if (st) return st;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return DIS_OK;
}

extern "C" int dis_lm_partials(dis_plan* p, const double* T, const void* x, const void* y, double* sums) {
  if (!p) return DIS_E_INVALID;
  dis_ctx* c = p->ctx;
  if (!T || !x || !y || !sums) return fail(c, DIS_E_INVALID, "dis_lm_partials: null argument");
  HIPCHK(c, hipSetDevice(c->device));
  int st = lm_partials(c, plan_view(p), T, (const float2*)x, (const float2*)y, sums);
  if (st) return st;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return DIS_OK;
}

extern "C" int dis_lm(dis_plan* p, double* T, double* w, uint8_t* cT, double winit, const double* tL, const void* x,
                      const void* y, double* En, int32_t* n_trials) {
  if (!p) return DIS_E_INVALID;
  dis_ctx* c = p->ctx;
  if (!T || !w || !cT || !tL || !x || !y || !En) return fail(c, DIS_E_INVALID, "dis_lm: null argument");
  HIPCHK(c, hipSetDevice(c->device));
  int nt = 0;
  int st = run_lm(c, plan_view(p), T, w, cT, winit, tL, (const float2*)x, (const float2*)y, En, &nt);
  if (st) return st;
  if (n_trials) *n_trials = nt;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return DIS_OK;
}
