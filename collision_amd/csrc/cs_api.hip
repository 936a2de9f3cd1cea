// C ABI and host orchestration of the collision step (see
// include/collision_amd.h).  The pass structure is the reference's
// resolveCollision (dcollid.cpp:317-362): average velocity, one proximity
// pass, up to five CCD passes, impact-zone passes while collisions remain,
// domain box, final position, strain limiting, final velocity.  Every pass is
//   element boxes -> (proximity pass only: LBVH build) -> LBVH refit ->
//   traversal (compacted candidate stream) -> narrow phase (hit records) ->
//   canonical segmented sums -> Jacobi apply [-> zone rigid update]
// with all state resident in device memory; the host only reads a small
// counter block per pass to decide whether another pass runs and whether a
// buffer must grow.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/collision_amd.h"
#include "cs_internal.h"

using namespace cs;

// Staging memory of cs_zone_labels_noalloc: a fixed device array instead of
// a hipMalloc per call.  Each device holds its own copy for the life of the
// process; host code reaches it only through the address hipGetSymbolAddress
// returns for that device (the array's name is a host-side stand-in).
constexpr int NOALLOC_MAX_VERTICES = 1 << 22;  // 16 MiB of int32 labels
static __device__ __align__(256) int32_t scratch[NOALLOC_MAX_VERTICES];

namespace {

constexpr int COUNTER_INTS = CNT_INTS;
std::string g_create_error;

// per-device address of `scratch`, looked up once per device in cs_create
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
int32_t* g_scratch_ptr[kMaxDevices] = {};
hipError_t g_scratch_err[kMaxDevices] = {};
bool g_scratch_done[kMaxDevices] = {};

// with `device` current
void init_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) return;
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (g_scratch_done[device]) return;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  g_scratch_err[device] = e;
  g_scratch_ptr[device] = e == hipSuccess ? (int32_t*)p : nullptr;
  g_scratch_done[device] = true;
}

// sizes of the previous pass's candidate stream, hits and hit pairs
struct PrevPass {
  bool valid = false;  // previous pass was a CCD / zone pass of this step
  int npairs = 0, nhits = 0, nhp = 0;
  int nq = 0;  // re-queried pairs (incremental pass)
  int nzone = 0;  // zone members whose velocities the pass rewrote (zlist prefix)
  unsigned long long cand = 0;
};

// the swept-box candidate pairs the proximity pass found for the first CCD
// pass (S.pairs_seed), under the velocities before the proximity pass's hits
struct SeedPairs {
  bool valid = false;
  int npairs = 0;
  unsigned long long cand = 0;
};

}  // namespace

struct PassTimes {
  hipEvent_t e0, e1, e2, e3, t0, t1, n0, n1;
};

struct cs_solver {
  cs_params P{};
  hipStream_t st = nullptr;
  DevState S;
  std::string err;
  bool assembled = false;
  bool state_set = false;
  int has_collision = 0;
  int* h_counters = nullptr;  // pinned mirror of the counter block
  std::vector<cs_pass_info> passes;
  std::vector<PassTimes> ptimes;  // events of each pass, read after the step
  int64_t plog_rows = 0;
  std::vector<hipEvent_t> events;
  size_t ev_used = 0;
  PrevPass prev;
  SeedPairs seed;
  int64_t syncs = 0;  // host synchronisations of the current resolve
  LearnedState L;           // time-of-impact proposer (learned form)
  cs_learned_stats lstats{};  // its statistics since the last cs_learned_get_stats
  // graph_step: the captured predictor + resolve prologue and what it was
  // captured for (a change of any of these captures it again)
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  struct GraphKey {
    int64_t assemble_gen = -1;
    double dt = 0.0;
    int record = 0, stats = 0;
    void* pimp = nullptr;
    bool operator==(const GraphKey& o) const {
      return assemble_gen == o.assemble_gen && dt == o.dt && record == o.record &&
             stats == o.stats && pimp == o.pimp;
    }
  } graph_key;
  int64_t assemble_gen = 0;  // bumped by every cs_assemble / cs_load
  // autotuner: over its first steps every tuning step runs one candidate
  // launch shape per kernel; per-unit kernel times are averaged per candidate
  struct Tuner {
    bool active = false;
    int step = 0;
    int tuned = 0;  // cs_tuning.tuned
    double t_trav[4] = {}, t_narrow[3] = {}, t_query[4] = {};
    int n_trav[4] = {}, n_narrow[3] = {}, n_query[4] = {};
    std::string key;
  } tuner;
};

namespace {
constexpr int kTuneTrav[4] = {4, 8, 16, 32};
constexpr int kTuneNarrow[3] = {64, 128, 256};
constexpr int kTuneQuery[4] = {1024, 2048, 4096, 8192};
constexpr int kTuneSteps = 8;  // two rounds of the four-candidate lists

// CS_TUNE_CACHE: one line per key, "key traverse narrow query"
bool tune_cache_lookup(const std::string& key, int* tr, int* na, int* qu) {
  const char* path = std::getenv("CS_TUNE_CACHE");
  if (!path || !*path) return false;
  FILE* f = std::fopen(path, "r");
  if (!f) return false;
  char k[256];
  int a, b, c;
  bool found = false;
  while (std::fscanf(f, "%255s %d %d %d", k, &a, &b, &c) == 4)
    if (key == k) {  // the last entry of a key wins
      *tr = a;
      *na = b;
      *qu = c;
      found = true;
    }
  std::fclose(f);
  return found;
}

void tune_cache_store(const std::string& key, int tr, int na, int qu) {
  const char* path = std::getenv("CS_TUNE_CACHE");
  if (!path || !*path) return;
  FILE* f = std::fopen(path, "a");
  if (!f) return;
  std::fprintf(f, "%s %d %d %d\n", key.c_str(), tr, na, qu);
  std::fclose(f);
}
}  // namespace

namespace {

int fail(cs_solver* h, int code, const std::string& msg) {
  if (h) h->err = msg;
  else g_create_error = msg;
  return code;
}

#define HIPCHK(h, expr)                                                                    \
  do {                                                                                     \
    hipError_t _e = (expr);                                                                \
    if (_e != hipSuccess)                                                                  \
      return fail((h), CS_ERR_HIP, std::string(#expr " failed: ") + hipGetErrorString(_e)); \
  } while (0)

template <typename T>
hipError_t dalloc(T** p, size_t count) {
  *p = nullptr;
  if (count == 0) count = 1;
  return hipMalloc((void**)p, sizeof(T) * count);
}

template <typename T>
void dfree(T*& p) {
  if (p) (void)hipFree((void*)p);
  p = nullptr;
}

void free_dev(DevState& S) {
  dfree(S.x); dfree(S.x_old); dfree(S.vel); dfree(S.av);
  dfree(S.fixed); dfree(S.has); dfree(S.uf);
  dfree(S.vbody); dfree(S.bptr); dfree(S.bmem); dfree(S.bfirst); dfree(S.bvisit);
  dfree(S.bmass); dfree(S.bcom); dfree(S.bvel); dfree(S.bmrg); dfree(S.rg_imp); dfree(S.rg_num);
  dfree(S.pimp); dfree(S.pnum);
  dfree(S.ev); dfree(S.eflags); dfree(S.len0); dfree(S.ebox);
  dfree(S.bounds); dfree(S.mcode); dfree(S.mid); dfree(S.mcode_tmp); dfree(S.mid_tmp);
  dfree(S.sort_hist); dfree(S.nodes); dfree(S.parent); dfree(S.nright); dfree(S.split_right);
  dfree(S.lpre); dfree(S.kleft); dfree(S.kright); dfree(S.kparent); dfree(S.krmin);
  dfree(S.krmax); dfree(S.pre);
  dfree(S.rmin); dfree(S.rmax); dfree(S.arrive); dfree(S.pairs); dfree(S.hits);
  dfree(S.top_start); dfree(S.scan_part); dfree(S.leaf_of);
  dfree(S.pairs_prev); dfree(S.qpairs); dfree(S.hits_prev); dfree(S.hp); dfree(S.hp_prev);
  dfree(S.dirty_e); dfree(S.dlist); dfree(S.pbox); dfree(S.pairs_seed);
  dfree(S.vlist); dfree(S.vadj_ptr); dfree(S.vadj);
  dfree(S.dpre); dfree(S.rflag); dfree(S.rlist); dfree(S.tnext);
  {
    uint8_t* tq = static_cast<uint8_t*>(S.tq);
    dfree(tq);
    S.tq = nullptr;
  }
  dfree(S.hkey); dfree(S.hkey_tmp); dfree(S.hidx); dfree(S.hidx_tmp);
  dfree(S.rkey); dfree(S.rkey_tmp); dfree(S.rval); dfree(S.rval_tmp);
  dfree(S.zmark); dfree(S.zlist); dfree(S.zkey64); dfree(S.zkey64_tmp);
  dfree(S.zkey); dfree(S.zval); dfree(S.zval_tmp);
  dfree(S.zhead); dfree(S.zend);
  dfree(S.lvl_ptr); dfree(S.plog); dfree(S.counters); dfree(S.spread); dfree(S.scn); dfree(S.vscene);
  dfree(S.sslot); dfree(S.spre); dfree(S.slast);
  dfree(S.lvl_active);
  S = DevState();
}

// grow a device buffer, keeping its first `keep` entries
template <typename T>
hipError_t grow_keep(T** p, int keep, size_t new_count, hipStream_t st) {
  T* np = nullptr;
  CS_HIP_RET(dalloc(&np, new_count));
  if (keep > 0 && *p)
    CS_HIP_RET(hipMemcpyAsync(np, *p, sizeof(T) * (size_t)keep, hipMemcpyDeviceToDevice, st));
  CS_HIP_RET(hipStreamSynchronize(st));
  dfree(*p);
  *p = np;
  return hipSuccess;
}

// hit records: current buffer keeps its first `keep` records
hipError_t alloc_hits(DevState& S, int cap, int keep, hipStream_t st) {
  dfree(S.hkey); dfree(S.hkey_tmp); dfree(S.hidx); dfree(S.hidx_tmp);
  dfree(S.rkey); dfree(S.rkey_tmp); dfree(S.rval); dfree(S.rval_tmp);
  dfree(S.hits_prev);
  S.hit_cap = cap;
  CS_HIP_RET(grow_keep(&S.hits, keep, cap, st));
  CS_HIP_RET(dalloc(&S.hits_prev, cap));
  CS_HIP_RET(dalloc(&S.hkey, cap));
  CS_HIP_RET(dalloc(&S.hkey_tmp, cap));
  CS_HIP_RET(dalloc(&S.hidx, cap));
  CS_HIP_RET(dalloc(&S.hidx_tmp, cap));
  CS_HIP_RET(dalloc(&S.rkey, 4 * (size_t)cap));
  CS_HIP_RET(dalloc(&S.rkey_tmp, 4 * (size_t)cap));
  CS_HIP_RET(dalloc(&S.rval, 4 * (size_t)cap));
  CS_HIP_RET(dalloc(&S.rval_tmp, 4 * (size_t)cap));
  return hipSuccess;
}

hipError_t alloc_sort_hist(DevState& S) {
  const int nmax = std::max({S.ne, S.nv, 4 * S.hit_cap, 1});
  const size_t words = radix_hist_words(nmax);
  if (words <= S.sort_hist_words) return hipSuccess;
  dfree(S.sort_hist);
  dfree(S.scan_part);
  S.sort_hist_words = words;
  S.scan_part_words = scan_partial_words((int)words);
  CS_HIP_RET(dalloc(&S.scan_part, S.scan_part_words));
  return dalloc(&S.sort_hist, words);
}

void to_dvec4(const double* src, int n, std::vector<dvec4>& out) {
  out.resize((size_t)n);
  for (int i = 0; i < n; ++i) out[i] = dvec4{src[3 * i], src[3 * i + 1], src[3 * i + 2], 0.0};
}

void from_dvec4(const std::vector<dvec4>& in, double* dst) {
  for (size_t i = 0; i < in.size(); ++i) {
    dst[3 * i] = in[i].x;
    dst[3 * i + 1] = in[i].y;
    dst[3 * i + 2] = in[i].z;
  }
}

hipEvent_t next_event(cs_solver* h) {
  if (h->ev_used == h->events.size()) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    h->events.push_back(e);
  }
  return h->events[h->ev_used++];
}

void mark(cs_solver* h, hipEvent_t* slot) {
  *slot = next_event(h);
  if (*slot) (void)hipEventRecord(*slot, h->st);
}

// a per-pass timing point, recorded when pass_timing is at least `level`
// (1: broad- and narrow-phase kernels, 2: phase boundaries)
void mark_pass(cs_solver* h, hipEvent_t* slot, int level) {
  if (h->P.pass_timing >= level) mark(h, slot);
  else *slot = nullptr;
}

float elapsed(hipEvent_t a, hipEvent_t b) {
  float ms = 0.f;
  if (a && b && hipEventElapsedTime(&ms, a, b) == hipSuccess) return ms;
  return 0.f;
}

int read_counters(cs_solver* h) {
  HIPCHK(h, hipMemcpyAsync(h->h_counters, h->S.counters, sizeof(int) * COUNTER_INTS,
                           hipMemcpyDeviceToHost, h->st));
  HIPCHK(h, hipStreamSynchronize(h->st));
  h->syncs++;
  return CS_OK;
}

int device_error(cs_solver* h) {
  const int e = h->h_counters[CNT_ERR];
  if (e & CS_DEVERR_NONFINITE)
    return fail(h, CS_ERR_NONFINITE, "non-finite value in the collision step");
  if (e) return fail(h, CS_ERR_INTERNAL, "device error flag set");
  return CS_OK;
}

unsigned long long c64(const cs_solver* h, int slot) {
  return *(const unsigned long long*)(h->h_counters + slot);
}

int set_counter(cs_solver* h, int slot, int value) {
  HIPCHK(h, hipMemcpyAsync(h->S.counters + slot, &value, sizeof(int), hipMemcpyHostToDevice, h->st));
  HIPCHK(h, hipStreamSynchronize(h->st));  // `value` is a host stack variable
  h->syncs++;
  return CS_OK;
}

template <typename T>
void swap_ptr(T*& a, T*& b) {
  T* t = a;
  a = b;
  b = t;
}

// grow the hit and hit-pair buffers to hold nhits / nhp records, keeping
// the first keep_hits / keep_hp entries
hipError_t grow_hit_buffers(cs_solver* h, int nhits, int nhp, int keep_hits, int keep_hp) {
  DevState& S = h->S;
  if (nhits > S.hit_cap) {
    CS_HIP_RET(alloc_hits(S, nhits + nhits / 2 + 1024, keep_hits, h->st));
    CS_HIP_RET(alloc_sort_hist(S));
  }
  if (nhp > S.hp_cap) {
    const int cap = nhp + nhp / 2 + 1024;
    CS_HIP_RET(grow_keep(&S.hp, keep_hp, cap, h->st));
    dfree(S.hp_prev);
    CS_HIP_RET(dalloc(&S.hp_prev, cap));
    S.hp_cap = cap;
  }
  return hipSuccess;
}

// The narrow phase over a pair stream.  In the learned form a CCD pass first
// scores the stream's sub-tests with the network and tests its pairs in
// order of decreasing score (the results do not depend on the order: hits
// are sorted by key before use and zone merges commute).
hipError_t narrow_call(cs_solver* h, NarrowParams NP, const int2* pairs, int npairs,
                       const int* npairs_dev, int cap) {
  if (npairs <= 0) return hipSuccess;
  if (NP.ccd && h->P.learned) {
    const uint32_t* perm = nullptr;
    CS_HIP_RET(launch_learned_score(h->st, h->S, h->L, pairs, npairs, npairs_dev, cap, NP.dt,
                                    NP.thickness > 0 ? NP.thickness : 1.0, &perm));
    NP.lseed = h->L.seeds;
    NP.perm = perm;
    // A pair whose sub-tests the certified screen all rejected has no hit
    // and merges no zone (merges start at a pair's first hit), so the narrow
    // phase only runs the pairs with a scored slot, which the sort put first.
    // (Dataset collection needs every sub-test's label: then all pairs run.)
    if (NP.ds_cap == 0) {
      launch_narrow(h->st, h->S, NP, pairs, npairs, h->L.count + 1, npairs);
      return hipGetLastError();
    }
  }
  launch_narrow(h->st, h->S, NP, pairs, npairs, npairs_dev, cap);
  return hipGetLastError();
}

// One broad + narrow + response pass.  *any_hit = some pair reported contact.
//
// Full pass (proximity pass, first CCD pass of a step): element boxes, LBVH
// build (proximity) or refit, traversal of every leaf, narrow phase of the
// whole candidate stream.
// Incremental pass (later CCD and impact-zone passes): dirty elements are
// re-boxed and re-queried, clean pairs / hits / hit pairs are carried over,
// and only the queried pairs go through the narrow phase.
int run_pass(cs_solver* h, int kind, int index, int pass_id, bool* any_hit) {
  DevState& S = h->S;
  const cs_params& P = h->P;
  const int ccd = kind != CS_PASS_PROXIMITY;
  // with movable bodies every pass changes all body points (their zones are
  // moved rigidly in every apply), so passes are not incremental
  const bool full_only = P.force_full_passes || S.nbody > 0;
  const bool incremental = ccd && h->prev.valid && !full_only;
  // the first CCD pass starts from the proximity pass's swept-box pairs
  const bool seeded = kind == CS_PASS_CCD && !incremental && h->seed.valid;
  // the proximity pass walks the tree once for itself and the first CCD pass
  const bool dual = kind == CS_PASS_PROXIMITY && !full_only;
  const SeedPairs seed = h->seed;
  h->seed.valid = false;
  const bool exact_counts = P.log_pairs != 0;
  PassTimes T{};
  // the previous pass's results become "prev"; this pass fills the others
  swap_ptr(S.pairs, S.pairs_prev);
  swap_ptr(S.hits, S.hits_prev);
  swap_ptr(S.hp, S.hp_prev);
  const PrevPass prev = h->prev;
  h->prev.valid = false;
  mark_pass(h, &T.e0, 2);
  // the dirty-range marks of the previous pass are cleared first (their count
  // is zeroed below)
  if (S.ne >= 2) launch_clear_ranges(h->st, S);
  // pairs, hits, hit pairs, sub-tests, query pairs, dirty elements, the
  // 64-bit candidate counts and the dirty-range count
  // and the list of vertices this pass's hits change; the list of the
  // previous pass (the other parity) is read by this pass
  int* vcount_cur = S.counters + CNT_VLIST + (pass_id & 1);
  const int* vcount_prev = S.counters + CNT_VLIST + ((pass_id & 1) ^ 1);
  launch_zero_counters(h->st, S,
                       (1u << CNT_PAIRS) | (1u << CNT_HITS) | (1u << CNT_HIT_PAIRS) |
                           (1u << CNT_HIT_SUB) | (1u << CNT_QPAIRS) | (1u << CNT_DIRTY) |
                           (0x3fu << CNT_CAND64) | (1u << CNT_DRANGES) |
                           (1u << (CNT_VLIST + (pass_id & 1))));
  NarrowParams NP{};
  NP.dt = P.dt; NP.thickness = P.thickness; NP.round_eps = P.round_eps;
  NP.k = P.spring_k; NP.m = P.mass; NP.lambda = P.friction; NP.cr = P.restitution;
  NP.vbody = S.vbody; NP.bmass = S.bmass;
  NP.mixed = P.precision == 1;
  NP.mixed_stats = P.precision == 1 && P.precision_stats != 0;
  NP.ccd = ccd;
  NP.zone_mode = (kind == CS_PASS_ZONE);
  NP.log_pairs = P.log_pairs;
  NP.pass_id = pass_id;
  NP.spread = S.spread;
  if (ccd && h->L.ds_cap > 0) {  // dataset collection (cs_learned_collect)
    NP.dsx = h->L.dsx;
    NP.dse = h->L.dse;
    NP.dst = h->L.dst;
    NP.ds_count = h->L.ds_count;
    NP.ds_cap = h->L.ds_cap;
    NP.ds_stride = h->L.ds_stride;
    NP.ds_salt = h->L.ds_salt++;
  }
  int npairs = 0, carried_hits = 0, carried_hp = 0, carried_sub = 0, carried_pairs = 0;
  unsigned long long cand = 0;
  int nq = 0;
  int nhits = 0, nhp = 0;
  int narrow_n = 0;
  if (!incremental && !seeded) {
    if (dual) {  // static boxes for this pass, swept boxes for the first CCD pass
      launch_element_boxes(h->st, S, 0, P.dt, S.pbox);
      launch_element_boxes(h->st, S, 1, P.dt, S.ebox);
    } else {
      launch_element_boxes(h->st, S, ccd, P.dt, S.ebox);
    }
    if (kind == CS_PASS_PROXIMITY) HIPCHK(h, build_bvh(h->st, S));
    HIPCHK(h, refit_bvh(h->st, S));
    for (;;) {  // traversal; re-run with a larger stream if it overflowed
      mark_pass(h, &T.t0, 1);
      launch_traverse(h->st, S, dual ? S.pairs_seed : S.pairs, S.pair_cap, dual ? S.pbox : nullptr);
      mark_pass(h, &T.t1, 1);
      HIPCHK(h, hipGetLastError());
      int rc = read_counters(h);
      if (rc) return rc;
      npairs = h->h_counters[CNT_PAIRS];
      if (npairs <= S.pair_cap) break;
      S.pair_cap = (int)std::min<int64_t>((int64_t)npairs + npairs / 4 + 1024, INT32_MAX / 2);
      dfree(S.pairs); dfree(S.pairs_prev); dfree(S.pairs_seed);
      HIPCHK(h, dalloc(&S.pairs, S.pair_cap));
      HIPCHK(h, dalloc(&S.pairs_prev, S.pair_cap));
      HIPCHK(h, dalloc(&S.pairs_seed, S.pair_cap));
      HIPCHK(h, hipMemsetAsync(S.counters + CNT_PAIRS, 0, sizeof(int), h->st));
      HIPCHK(h, hipMemsetAsync(S.counters + CNT_CAND64, 0, 2 * sizeof(int), h->st));
      HIPCHK(h, hipMemsetAsync(S.counters + CNT_PCAND64, 0, 2 * sizeof(int), h->st));
    }
    cand = c64(h, CNT_CAND64);
    const int2* stream = S.pairs;
    const int* stream_dev = nullptr;
    if (dual) {
      // this pass's pairs are the flagged ones; their number is read on the
      // device by the narrow phase (at most npairs)
      h->seed.valid = true;
      h->seed.npairs = npairs;
      h->seed.cand = cand;
      HIPCHK(h, hipMemsetAsync(S.counters + CNT_PPAIRS, 0, sizeof(int), h->st));
      launch_split_seed(h->st, S, npairs);
      stream_dev = S.counters + CNT_PPAIRS;
    }
    mark_pass(h, &T.e1, 2);
    for (;;) {
      mark_pass(h, &T.n0, 1);
      HIPCHK(h, narrow_call(h, NP, stream, npairs, stream_dev, S.pair_cap));
      mark_pass(h, &T.n1, 1);
      HIPCHK(h, hipGetLastError());
      int rc = read_counters(h);
      if (rc) return rc;
      nhits = h->h_counters[CNT_HITS];
      nhp = h->h_counters[CNT_HIT_PAIRS];
      if (nhits <= S.hit_cap && nhp <= S.hp_cap) break;
      HIPCHK(h, grow_hit_buffers(h, nhits, nhp, 0, 0));
      if ((rc = set_counter(h, CNT_HITS, 0))) return rc;
      if ((rc = set_counter(h, CNT_HIT_PAIRS, 0))) return rc;
      if ((rc = set_counter(h, CNT_HIT_SUB, 0))) return rc;
    }
    if (dual) {  // (counted by the split, read with the narrow phase's counters)
      npairs = h->h_counters[CNT_PPAIRS];
      cand = c64(h, CNT_PCAND64);
    }
    narrow_n = npairs;
  } else if (seeded) {
    // The swept-box pairs of the proximity pass, with the elements whose
    // vertices that pass's hits changed queried again under their new boxes
    // (the incremental scheme below); no narrow-phase result carries over
    // (the proximity pass ran another test), so every pair is tested.
    launch_mark_dirty(h->st, S, vcount_prev, 0);
    mark_pass(h, &T.t0, 1);
    if (exact_counts) launch_query_dirty(h->st, S, 0);  // seed boxes: count only
    launch_grow_dirty(h->st, S, P.dt);  // new boxes; leaves exact, ancestors grown
    launch_query_dirty(h->st, S, 1);
    mark_pass(h, &T.t1, 1);
    launch_carry_pairs_from(h->st, S, S.pairs_seed, seed.npairs);
    for (int attempt = 0;; ++attempt) {
      int rc = read_counters(h);
      if (rc) return rc;
      nq = h->h_counters[CNT_QPAIRS];
      if (nq <= S.qpair_cap) break;
      if (attempt > 4) return fail(h, CS_ERR_INTERNAL, "query stream keeps overflowing");
      S.qpair_cap = nq + nq / 2 + 1024;
      dfree(S.qpairs);
      HIPCHK(h, dalloc(&S.qpairs, S.qpair_cap));
      HIPCHK(h, hipMemsetAsync(S.counters + CNT_QPAIRS, 0, sizeof(int), h->st));
      HIPCHK(h, hipMemsetAsync(S.counters + CNT_NEW64, 0, 2 * sizeof(int), h->st));
      launch_query_dirty(h->st, S, 1);
    }
    carried_pairs = h->h_counters[CNT_PAIRS];
    npairs = carried_pairs + nq;
    if (npairs > S.pair_cap) {
      const int cap = (int)std::min<int64_t>((int64_t)npairs + npairs / 4 + 1024, INT32_MAX / 2);
      HIPCHK(h, grow_keep(&S.pairs, carried_pairs, cap, h->st));
      dfree(S.pairs_prev);
      HIPCHK(h, dalloc(&S.pairs_prev, cap));
      HIPCHK(h, grow_keep(&S.pairs_seed, 0, cap, h->st));
      S.pair_cap = cap;
    }
    if (nq > 0)
      HIPCHK(h, hipMemcpyAsync(S.pairs + carried_pairs, S.qpairs, sizeof(int2) * (size_t)nq,
                               hipMemcpyDeviceToDevice, h->st));
    cand = exact_counts ? seed.cand - c64(h, CNT_OLD64) + c64(h, CNT_NEW64) : ~0ull;
    mark_pass(h, &T.e1, 2);
    narrow_n = npairs;
    for (;;) {
      mark_pass(h, &T.n0, 1);
      HIPCHK(h, narrow_call(h, NP, S.pairs, npairs, nullptr, 0));
      mark_pass(h, &T.n1, 1);
      HIPCHK(h, hipGetLastError());
      int rc = read_counters(h);
      if (rc) return rc;
      nhits = h->h_counters[CNT_HITS];
      nhp = h->h_counters[CNT_HIT_PAIRS];
      if (nhits <= S.hit_cap && nhp <= S.hp_cap) break;
      HIPCHK(h, grow_hit_buffers(h, nhits, nhp, 0, 0));
      if ((rc = set_counter(h, CNT_HITS, 0))) return rc;
      if ((rc = set_counter(h, CNT_HIT_PAIRS, 0))) return rc;
      if ((rc = set_counter(h, CNT_HIT_SUB, 0))) return rc;
    }
  } else {
    // One host synchronisation per pass: the narrow phase reads the number
    // of re-queried pairs from device memory, the counts carried over from
    // the previous pass are snapshotted on the device before the narrow
    // phase appends to them, and overflows are detected afterwards (the
    // narrow phase only appends records and performs idempotent zone
    // merges, so it can be run again).
    launch_mark_dirty(h->st, S, vcount_prev, prev.nzone);
    mark_pass(h, &T.t0, 1);
    // previous boxes: count only.  The count only feeds the exact candidate
    // total of the pass (a diagnostic), so it runs when the caller asks for
    // the parity taps; otherwise the candidate total is reported as -1.
    if (exact_counts) launch_query_dirty(h->st, S, 0);
    launch_grow_dirty(h->st, S, P.dt);               // new boxes; leaves exact, ancestors grown
    launch_query_dirty(h->st, S, 1);                 // current boxes: count + emit
    mark_pass(h, &T.t1, 1);
    // the candidate pair stream itself only feeds the exact pair count of the
    // pass; it is carried when the caller asks for exact counts
    launch_carry(h->st, S, exact_counts ? prev.npairs : 0, prev.nhits, prev.nhp);
    HIPCHK(h, hipMemcpyAsync(S.counters + CNT_SNAP, S.counters, sizeof(int) * 4,
                             hipMemcpyDeviceToDevice, h->st));
    mark_pass(h, &T.e1, 2);
    // impact-zone merges of the carried hit pairs (union order does not
    // change the partition); their count is read on the device
    if (kind == CS_PASS_ZONE) launch_hp_unions(h->st, S, prev.nhp, 1);
    // grid for a guess of the re-queried pair count; pairs beyond it are
    // handled after the count is known
    const int guess = std::min(S.qpair_cap, std::max(4096, 2 * prev.nq + 1024));
    for (int attempt = 0;; ++attempt) {
      mark_pass(h, &T.n0, 1);
      HIPCHK(h, narrow_call(h, NP, S.qpairs, guess, S.counters + CNT_QPAIRS, S.qpair_cap));
      mark_pass(h, &T.n1, 1);
      HIPCHK(h, hipGetLastError());
      int rc = read_counters(h);
      if (rc) return rc;
      nq = h->h_counters[CNT_QPAIRS];
      if (nq > guess && nq <= S.qpair_cap) {  // the rest of the stream
        HIPCHK(h, narrow_call(h, NP, S.qpairs + guess, nq - guess, nullptr, 0));
        mark_pass(h, &T.n1, 1);
        HIPCHK(h, hipGetLastError());
        if ((rc = read_counters(h))) return rc;
      }
      carried_pairs = h->h_counters[CNT_SNAP + CNT_PAIRS];
      carried_hits = h->h_counters[CNT_SNAP + CNT_HITS];
      carried_hp = h->h_counters[CNT_SNAP + CNT_HIT_PAIRS];
      carried_sub = h->h_counters[CNT_SNAP + CNT_HIT_SUB];
      nhits = h->h_counters[CNT_HITS];
      nhp = h->h_counters[CNT_HIT_PAIRS];
      const bool q_over = nq > S.qpair_cap;
      if (!q_over && nhits <= S.hit_cap && nhp <= S.hp_cap) break;
      if (attempt > 4) return fail(h, CS_ERR_INTERNAL, "narrow-phase buffers keep overflowing");
      if (q_over) {  // grow the query stream and query again
        S.qpair_cap = nq + nq / 2 + 1024;
        dfree(S.qpairs);
        HIPCHK(h, dalloc(&S.qpairs, S.qpair_cap));
        HIPCHK(h, hipMemsetAsync(S.counters + CNT_QPAIRS, 0, sizeof(int), h->st));
        HIPCHK(h, hipMemsetAsync(S.counters + CNT_NEW64, 0, 2 * sizeof(int), h->st));
        launch_query_dirty(h->st, S, 1);
      }
      HIPCHK(h, grow_hit_buffers(h, nhits, nhp, carried_hits, carried_hp));
      if ((rc = set_counter(h, CNT_HITS, carried_hits))) return rc;
      if ((rc = set_counter(h, CNT_HIT_PAIRS, carried_hp))) return rc;
      if ((rc = set_counter(h, CNT_HIT_SUB, carried_sub))) return rc;
    }
    narrow_n = nq;
    npairs = exact_counts && prev.npairs >= 0 ? carried_pairs + nq : -1;
    if (npairs > S.pair_cap) {
      const int cap = (int)std::min<int64_t>((int64_t)npairs + npairs / 4 + 1024, INT32_MAX / 2);
      HIPCHK(h, grow_keep(&S.pairs, carried_pairs, cap, h->st));
      dfree(S.pairs_prev);
      HIPCHK(h, dalloc(&S.pairs_prev, cap));
      dfree(S.pairs_seed);  // (its pairs were used by the first CCD pass)
      HIPCHK(h, dalloc(&S.pairs_seed, cap));
      S.pair_cap = cap;
    }
    if (nq > 0 && npairs >= 0)
      HIPCHK(h, hipMemcpyAsync(S.pairs + carried_pairs, S.qpairs, sizeof(int2) * (size_t)nq,
                               hipMemcpyDeviceToDevice, h->st));
    cand = exact_counts && prev.cand != ~0ull ? prev.cand - c64(h, CNT_OLD64) + c64(h, CNT_NEW64)
                                              : ~0ull;
  }
  mark_pass(h, &T.e2, 2);
  launch_scene_pass(h->st, S, kind, index, nhp, P.max_ccd_passes);  // batched scenes
  HIPCHK(h, apply_hits(h->st, S, nhits, vcount_cur, P.record_impulses != 0 && S.pimp));
  HIPCHK(h, apply_bodies(h->st, S, P.mass, P.dt));  // RG impulses, body zones (updateAverageVelocity)
  int nzone = 0;
  if (kind == CS_PASS_ZONE) {
    nzone = h->h_counters[CNT_ZLIST];
    HIPCHK(h, update_zones(h->st, S, P.mass, P.dt, nzone));  // (resets the zone count)
    HIPCHK(h, update_body_zones(h->st, S, P.mass, P.dt));    // body zones hold >= 2 points
  }
  if (P.log_pairs) launch_hp_log(h->st, S, nhp, pass_id);
  mark_pass(h, &T.e3, 2);
  cs_pass_info pi{};
  pi.kind = kind;
  pi.index = index;
  pi.candidates = cand == ~0ull ? -1 : (int64_t)cand;
  pi.narrow_pairs = npairs;
  pi.tested_pairs = narrow_n;
  pi.incremental = (incremental || seeded) ? 1 : 0;
  pi.hit_pairs = nhp;
  pi.hit_subtests = h->h_counters[CNT_HIT_SUB];
  if (P.log_pairs) {  // the pair log is read after every pass
    int rc = read_counters(h);
    if (rc) return rc;
    h->plog_rows = std::min<int64_t>(h->h_counters[CNT_LOG], S.plog_cap);
  }
  h->passes.push_back(pi);
  h->ptimes.push_back(T);  // durations are read once the step has finished
  *any_hit = nhp > 0;
  if (ccd) {
    h->prev.valid = true;
    h->prev.npairs = npairs;
    h->prev.nhits = nhits;
    h->prev.nhp = nhp;
    h->prev.nq = (incremental || seeded) ? nq : 0;
    h->prev.cand = cand;
    h->prev.nzone = nzone;
  }
  return device_error(h);
}

int check_ready(cs_solver* h) {
  if (!h) return CS_ERR_ARG;
  if (!h->assembled) return fail(h, CS_ERR_ARG, "cs_assemble has not been called");
  if (!h->state_set) return fail(h, CS_ERR_ARG, "cs_set_state has not been called");
  HIPCHK(h, hipSetDevice(h->P.device));
  return CS_OK;
}

}  // namespace

// checkpoint file: header, the parameters, then the arrays
namespace {
constexpr char kCkptMagic[8] = {'C', 'S', 'C', 'K', 'P', 'T', '0', '1'};
struct CkptHeader {
  char magic[8];
  int32_t nv, nt, nb, nbody;
  uint64_t topo;  // FNV-1a hash of the element table and the flags
  cs_params params;
};

uint64_t topo_hash(const std::vector<int4>& ev, const std::vector<uint8_t>& ef,
                   const std::vector<uint8_t>& fixed) {
  uint64_t hsh = 1469598103934665603ull;
  auto mix = [&](const void* p, size_t n) {
    const unsigned char* b = (const unsigned char*)p;
    for (size_t i = 0; i < n; ++i) hsh = (hsh ^ b[i]) * 1099511628211ull;
  };
  mix(ev.data(), sizeof(int4) * ev.size());
  mix(ef.data(), ef.size());
  mix(fixed.data(), fixed.size());
  return hsh;
}

int checkpoint_topology(cs_solver* h, uint64_t* out) {
  DevState& S = h->S;
  std::vector<int4> ev((size_t)S.ne);
  std::vector<uint8_t> ef((size_t)S.ne), fixed((size_t)S.nv);
  if (S.ne) {
    HIPCHK(h, hipMemcpy(ev.data(), S.ev, sizeof(int4) * S.ne, hipMemcpyDeviceToHost));
    HIPCHK(h, hipMemcpy(ef.data(), S.eflags, S.ne, hipMemcpyDeviceToHost));
  }
  HIPCHK(h, hipMemcpy(fixed.data(), S.fixed, S.nv, hipMemcpyDeviceToHost));
  *out = topo_hash(ev, ef, fixed);
  return CS_OK;
}

// the device arrays a checkpoint holds, in file order
struct CkptArray {
  void* dev;
  size_t bytes;
};
std::vector<CkptArray> checkpoint_arrays(DevState& S) {
  std::vector<CkptArray> a = {{S.x, sizeof(dvec4) * S.nv},     {S.x_old, sizeof(dvec4) * S.nv},
                              {S.vel, sizeof(dvec4) * S.nv},   {S.av, sizeof(dvec4) * S.nv},
                              {S.has, (size_t)S.nv}};
  if (S.nbody > 0) {
    a.push_back({S.bcom, sizeof(dvec4) * S.nbody});
    a.push_back({S.bvel, sizeof(dvec4) * S.nbody});
    a.push_back({S.bmrg, sizeof(dvec4) * S.nbody});
    a.push_back({S.rg_imp, sizeof(dvec4) * S.nbody});
    a.push_back({S.rg_num, sizeof(int32_t) * S.nbody});
  }
  return a;
}
}  // namespace

extern "C" {

void cs_default_params(cs_params* p) {
  if (!p) return;
  std::memset(p, 0, sizeof(*p));
  p->dt = 1e-3;            // DT (collid.h:19)
  p->thickness = 1e-4;     // dcollid.cpp:30
  p->round_eps = 1e-6;     // EPS (collid.h:18)
  p->spring_k = 1000.0;    // dcollid.cpp:32
  p->mass = 0.01;          // dcollid.cpp:33
  p->friction = 0.02;      // dcollid.cpp:34
  p->restitution = 0.0;    // dcollid.cpp:35
  for (int i = 0; i < 3; ++i) {
    p->dom_lo[i] = -1e30;
    p->dom_hi[i] = 1e30;
  }
  p->max_ccd_passes = 5;
  p->max_zone_passes = 1000;
  p->max_strain_sweeps = 10;
  p->log_pairs = 0;
  p->device = 0;
  p->pass_timing = 2;
}

int cs_create(cs_solver** out, const cs_params* p) {
  if (!out) return fail(nullptr, CS_ERR_ARG, "out is NULL");
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
    return fail(nullptr, CS_ERR_NO_DEVICE, "no HIP device available");
  cs_solver* h = new (std::nothrow) cs_solver();
  if (!h) return fail(nullptr, CS_ERR_NOMEM, "out of host memory");
  if (p) h->P = *p;
  else cs_default_params(&h->P);
  if (h->P.device < 0 || h->P.device >= ndev) {
    delete h;
    return fail(nullptr, CS_ERR_ARG, "device ordinal out of range");
  }
  if (hipSetDevice(h->P.device) != hipSuccess || hipStreamCreate(&h->st) != hipSuccess ||
      hipHostMalloc((void**)&h->h_counters, sizeof(int) * COUNTER_INTS) != hipSuccess) {
    delete h;
    return fail(nullptr, CS_ERR_HIP, "HIP stream / pinned buffer creation failed");
  }
  init_scratch(h->P.device);  // the staging array of cs_zone_labels_noalloc
  *out = h;
  return CS_OK;
}

int cs_destroy(cs_solver* h) {
  if (!h) return CS_OK;
  (void)hipSetDevice(h->P.device);
  if (h->st) (void)hipStreamSynchronize(h->st);
  free_dev(h->S);
  learned_free(h->L);
  if (h->graph_exec) (void)hipGraphExecDestroy(h->graph_exec);
  if (h->graph) (void)hipGraphDestroy(h->graph);
  for (auto e : h->events) (void)hipEventDestroy(e);
  if (h->h_counters) (void)hipHostFree(h->h_counters);
  if (h->st) (void)hipStreamDestroy(h->st);
  delete h;
  return CS_OK;
}

int cs_set_params(cs_solver* h, const cs_params* p) {
  if (!h || !p) return CS_ERR_ARG;
  if (p->device != h->P.device) return fail(h, CS_ERR_ARG, "device cannot change after create");
  if (!(p->dt >= 0) || !(p->thickness >= 0) || !(p->round_eps >= 0))
    return fail(h, CS_ERR_ARG, "dt, thickness and round_eps must be >= 0");
  if (p->precision < 0 || p->precision > 1 || p->learned < 0 || p->learned > 1)
    return fail(h, CS_ERR_ARG, "precision and learned must be 0 or 1");
  h->P = *p;
  return CS_OK;
}

int cs_get_params(cs_solver* h, cs_params* p) {
  if (!h || !p) return CS_ERR_ARG;
  *p = h->P;
  return CS_OK;
}

int cs_set_domain_boundary(cs_solver* h, const double* L, const double* U) {
  if (!h || !L || !U) return CS_ERR_ARG;
  for (int i = 0; i < 3; ++i) {
    h->P.dom_lo[i] = L[i];
    h->P.dom_hi[i] = U[i];
  }
  return CS_OK;
}

int cs_assemble(cs_solver* h, const cs_mesh_view* m) {
  if (!h || !m) return CS_ERR_ARG;
  if (m->nv <= 0 || m->nt < 0 || m->nb < 0 || (m->nt > 0 && !m->tris) ||
      (m->nb > 0 && !m->bonds))
    return fail(h, CS_ERR_ARG, "invalid mesh view");
  const int nv = m->nv, nt = m->nt, nb = m->nb, ne = nt + nb;
  if ((int64_t)ne >= (1ll << 28) || (int64_t)nv >= (1ll << 28))
    return fail(h, CS_ERR_UNSUPPORTED, "more than 2^28 elements or vertices");
  std::vector<int4> ev((size_t)std::max(ne, 1));
  std::vector<uint8_t> ef((size_t)std::max(ne, 1));
  std::vector<dvec4> l0((size_t)std::max(ne, 1));
  std::vector<uint8_t> fixed((size_t)nv, 0);
  if (m->is_fixed) std::memcpy(fixed.data(), m->is_fixed, (size_t)nv);
  std::vector<uint8_t> used((size_t)nv, 0);
  // movable bodies: a body is whole triangles of movable vertices of that
  // body; no bond holds a movable vertex
  const int nbody = m->nbody;
  if (nbody < 0 || (nbody > 0 && (!m->vert_body || !m->body_mass || !m->body_com)))
    return fail(h, CS_ERR_ARG, "invalid movable-body arrays");
  std::vector<int32_t> vbody((size_t)nv, -1);
  if (nbody > 0) {
    for (int v = 0; v < nv; ++v) {
      const int b = m->vert_body[v];
      if (b < -1 || b >= nbody) return fail(h, CS_ERR_ARG, "vertex body id out of range");
      if (b >= 0 && fixed[v]) return fail(h, CS_ERR_ARG, "a vertex cannot be static and movable");
      vbody[v] = b;
    }
  }
  auto movable = [&](int v) { return vbody[v] >= 0; };
  // batched scenes: the scene of an element goes to the upper half of its
  // surface field, the surface's id within its scene to the lower half
  const int nscene = m->nscene;
  if (nscene < 0 || (nscene > 1 && !m->elem_scene))
    return fail(h, CS_ERR_ARG, "invalid batched-scene arrays");
  if (nscene > 32767) return fail(h, CS_ERR_UNSUPPORTED, "more than 32767 scenes");
  if (nscene > 1 && nbody > 0)
    return fail(h, CS_ERR_UNSUPPORTED, "batched scenes with movable bodies");
  std::vector<int32_t> vscene;
  std::vector<std::vector<std::pair<int, int>>> scene_surf;
  if (nscene > 1) {
    vscene.assign((size_t)nv, -1);
    scene_surf.resize((size_t)nscene);
    for (int e = 0; e < ne; ++e) {
      const int sc = m->elem_scene[e];
      if (sc < 0 || sc >= nscene) return fail(h, CS_ERR_ARG, "element scene id out of range");
      const int np = e < nt ? 3 : 2;
      for (int j = 0; j < np; ++j) {
        const int v = e < nt ? m->tris[3 * e + j] : m->bonds[2 * (e - nt) + j];
        if (v < 0 || v >= nv) return fail(h, CS_ERR_ARG, "element vertex id out of range");
        if (vscene[v] >= 0 && vscene[v] != sc)
          return fail(h, CS_ERR_ARG, "elements of different scenes share a vertex");
        vscene[v] = sc;
      }
    }
  }
  auto surf_field = [&](int e, int user_surf) -> int {
    if (nscene <= 1) return user_surf;
    const int sc = m->elem_scene[e];
    if (e >= nt) return (sc << 16) | 0xffff;
    auto& lst = scene_surf[(size_t)sc];
    int local = -1;
    for (auto& pr : lst)
      if (pr.first == user_surf) local = pr.second;
    if (local < 0) {
      local = (int)lst.size();
      lst.push_back({user_surf, local});
    }
    return (sc << 16) | local;
  };
  for (int e = 0; e < nt; ++e) {
    int v[3];
    for (int j = 0; j < 3; ++j) {
      v[j] = m->tris[3 * e + j];
      if (v[j] < 0 || v[j] >= nv) return fail(h, CS_ERR_ARG, "triangle vertex id out of range");
      used[v[j]] = 1;
    }
    if (vbody[v[1]] != vbody[v[0]] || vbody[v[2]] != vbody[v[0]])
      return fail(h, CS_ERR_ARG, "a triangle mixes vertices of different bodies");
    ev[e] = make_int4(v[0], v[1], v[2], surf_field(e, m->tri_surf ? m->tri_surf[e] : 0));
    if (nscene > 1 && scene_surf[(size_t)m->elem_scene[e]].size() >= 0xffff)
      return fail(h, CS_ERR_UNSUPPORTED, "more than 65534 surfaces in one scene");
    // isRigidBody(hse): a static or movable point (dcollid.cpp:1091-1094)
    const bool rigid = fixed[v[0]] || fixed[v[1]] || fixed[v[2]] || movable(v[0]);
    ef[e] = EF_TRI | (rigid ? EF_RIGID : 0);
    l0[e] = dvec4{m->tri_len0 ? m->tri_len0[3 * e] : 0.0, m->tri_len0 ? m->tri_len0[3 * e + 1] : 0.0,
                  m->tri_len0 ? m->tri_len0[3 * e + 2] : 0.0, 0.0};
  }
  for (int b = 0; b < nb; ++b) {
    const int e = nt + b;
    const int v0 = m->bonds[2 * b], v1 = m->bonds[2 * b + 1];
    if (v0 < 0 || v0 >= nv || v1 < 0 || v1 >= nv)
      return fail(h, CS_ERR_ARG, "bond vertex id out of range");
    used[v0] = used[v1] = 1;
    if (movable(v0) || movable(v1)) return fail(h, CS_ERR_ARG, "a bond holds a movable vertex");
    ev[e] = make_int4(v0, v1, -1, surf_field(e, -1));
    ef[e] = (fixed[v0] || fixed[v1]) ? EF_RIGID : 0;
    l0[e] = dvec4{m->bond_len0 ? m->bond_len0[b] : 0.0, 0.0, 0.0, 0.0};
  }
  for (int v = 0; v < nv; ++v)
    if (!used[v])
      return fail(h, CS_ERR_ARG, "every vertex must belong to a triangle or bond");
  // strain-limit schedule: levels (cs_schedule.hip), then slots in level
  // order, element order within a level (the elements of one level share no
  // vertex, so any order is exact there)
  std::vector<int> level;
  int nlev = 0, sdelta = 1;
  build_strain_schedule(nv, ne, ev.data(), ef.data(), level, nlev, sdelta);
  int nsched = 0;
  for (int e = 0; e < ne; ++e)
    if (level[e] >= 0) ++nsched;
  std::vector<int> lptr((size_t)nlev + 1, 0), slot_of((size_t)std::max(ne, 1), -1);
  const size_t nslot = (size_t)std::max(nsched, 1);
  for (int e = 0; e < ne; ++e)
    if (level[e] >= 0) lptr[level[e] + 1]++;
  if (nlev >= (1 << 21)) return fail(h, CS_ERR_UNSUPPORTED, "strain schedule above 2^21 levels");
  for (int l = 0; l < nlev; ++l) {
    if (lptr[l + 1] > STRAIN_LEVEL_CAP)
      return fail(h, CS_ERR_INTERNAL, "strain schedule level above its size cap");
    lptr[l + 1] += lptr[l];
  }
  {
    std::vector<int> fill(lptr.begin(), lptr.end() - 1);
    for (int e = 0; e < ne; ++e)
      if (level[e] >= 0) slot_of[e] = fill[level[e]]++;
  }
  std::vector<StrainSlot> sslot(nslot);
  std::vector<int> slast;
  std::memset(sslot.data(), 0, sizeof(StrainSlot) * nslot);
  {
    // links: each visit of a vertex to its next visit, the last to the
    // first (element order and level order agree on the elements of one
    // vertex)
    std::vector<int> first((size_t)nv, -1), last((size_t)nv, -1);
    for (size_t s = 0; s < nslot; ++s)
      for (int i = 0; i < 3; ++i) {
        sslot[s].v[i] = -1;
        sslot[s].ns[i] = -1;
        sslot[s].nl[i] = -1;
      }
    for (int e = 0; e < ne; ++e) {
      if (level[e] < 0) continue;
      const int np = (ef[e] & EF_TRI) ? 3 : 2;
      const int vv[3] = {ev[e].x, ev[e].y, ev[e].z};
      const int s = slot_of[e];
      StrainSlot& r = sslot[s];
      r.level = level[e];
      r.L0[0] = l0[e].x; r.L0[1] = l0[e].y; r.L0[2] = l0[e].z;
      for (int i = 0; i < np; ++i) {
        const int v = vv[i];
        r.v[i] = v;
        if (last[v] >= 0) {
          StrainSlot& p = sslot[last[v] >> 2];
          p.ns[last[v] & 3] = 4 * s + i;
          p.nl[last[v] & 3] = (level[e] << 10) | (s - lptr[level[e]]);
        } else {
          first[v] = 4 * s + i;
        }
        last[v] = 4 * s + i;
      }
    }
    slast = last;
    for (int v = 0; v < nv; ++v) {
      if (last[v] < 0) continue;
      StrainSlot& p = sslot[last[v] >> 2];
      p.ns[last[v] & 3] = first[v];
      const int fs = first[v] >> 2, fl = sslot[fs].level;
      p.nl[last[v] & 3] = (fl << 10) | (fs - lptr[fl]);
    }
  }

  HIPCHK(h, hipSetDevice(h->P.device));
  HIPCHK(h, hipStreamSynchronize(h->st));
  free_dev(h->S);
  DevState& S = h->S;
  S.nv = nv; S.nt = nt; S.nb = nb; S.ne = ne;
  HIPCHK(h, dalloc(&S.x, nv)); HIPCHK(h, dalloc(&S.x_old, nv));
  HIPCHK(h, dalloc(&S.vel, nv)); HIPCHK(h, dalloc(&S.av, nv));
  HIPCHK(h, dalloc(&S.fixed, nv)); HIPCHK(h, dalloc(&S.has, nv)); HIPCHK(h, dalloc(&S.uf, nv));
  HIPCHK(h, dalloc(&S.ev, ne)); HIPCHK(h, dalloc(&S.eflags, ne)); HIPCHK(h, dalloc(&S.len0, ne));
  HIPCHK(h, dalloc(&S.ebox, ne));
  HIPCHK(h, dalloc(&S.bounds, 6));
  HIPCHK(h, dalloc(&S.mcode, ne)); HIPCHK(h, dalloc(&S.mid, ne));
  HIPCHK(h, dalloc(&S.mcode_tmp, ne)); HIPCHK(h, dalloc(&S.mid_tmp, ne));
  HIPCHK(h, dalloc(&S.nodes, 2 * ne - 1));
  HIPCHK(h, dalloc(&S.parent, 2 * ne - 1));
  HIPCHK(h, dalloc(&S.nright, 2 * ne - 1));
  HIPCHK(h, dalloc(&S.lpre, ne));
  HIPCHK(h, dalloc(&S.kleft, std::max(ne - 1, 1))); HIPCHK(h, dalloc(&S.kright, std::max(ne - 1, 1)));
  HIPCHK(h, dalloc(&S.kparent, 2 * ne - 1)); HIPCHK(h, dalloc(&S.pre, 2 * ne - 1));
  HIPCHK(h, dalloc(&S.krmin, std::max(ne - 1, 1))); HIPCHK(h, dalloc(&S.krmax, std::max(ne - 1, 1)));
  HIPCHK(h, dalloc(&S.split_right, std::max(ne - 1, 1)));
  HIPCHK(h, dalloc(&S.rmin, 2 * ne - 1)); HIPCHK(h, dalloc(&S.rmax, 2 * ne - 1));
  HIPCHK(h, dalloc(&S.arrive, 2 * ne - 1)); HIPCHK(h, dalloc(&S.top_start, std::max(ne - 1, 1)));
  S.pair_cap = (int)std::min<int64_t>(3ll * ne + 4096, INT32_MAX / 2);
  HIPCHK(h, dalloc(&S.pairs, S.pair_cap));
  HIPCHK(h, dalloc(&S.pairs_prev, S.pair_cap));
  HIPCHK(h, dalloc(&S.pairs_seed, S.pair_cap));
  HIPCHK(h, dalloc(&S.pbox, ne));
  S.qpair_cap = std::max(1 << 16, ne / 8);
  HIPCHK(h, dalloc(&S.qpairs, S.qpair_cap));
  HIPCHK(h, alloc_hits(S, std::max(1 << 16, ne / 4), 0, h->st));
  S.hp_cap = std::max(1 << 16, ne / 4);
  HIPCHK(h, dalloc(&S.hp, S.hp_cap));
  HIPCHK(h, dalloc(&S.hp_prev, S.hp_cap));
  HIPCHK(h, dalloc(&S.leaf_of, ne));
  HIPCHK(h, dalloc(&S.dirty_e, ((size_t)ne + 3) & ~(size_t)3));
  HIPCHK(h, dalloc(&S.dlist, ne));
  HIPCHK(h, dalloc(&S.vlist, nv));
  // vertex -> elements, in element order
  std::vector<int32_t> vptr((size_t)nv + 1, 0), vadj;
  for (int e = 0; e < ne; ++e) {
    vptr[(size_t)ev[e].x + 1]++;
    vptr[(size_t)ev[e].y + 1]++;
    if (ev[e].z >= 0) vptr[(size_t)ev[e].z + 1]++;
  }
  for (int v = 0; v < nv; ++v) vptr[(size_t)v + 1] += vptr[v];
  vadj.resize(std::max<size_t>(1, (size_t)vptr[nv]));
  {
    std::vector<int32_t> fillp(vptr.begin(), vptr.end() - 1);
    for (int e = 0; e < ne; ++e) {
      vadj[fillp[ev[e].x]++] = e;
      vadj[fillp[ev[e].y]++] = e;
      if (ev[e].z >= 0) vadj[fillp[ev[e].z]++] = e;
    }
  }
  HIPCHK(h, dalloc(&S.vadj_ptr, (size_t)nv + 1));
  HIPCHK(h, dalloc(&S.vadj, vadj.size()));
  HIPCHK(h, dalloc(&S.dpre, 2 * ne - 1));
  HIPCHK(h, dalloc(&S.tnext, kNumXcd + 2));
  {
    uint8_t* tq = nullptr;
    HIPCHK(h, dalloc(&tq, 2 * TRAV_QCAP * TRAV_ITEM_BYTES));
    S.tq = tq;
  }
  {
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, h->P.device) == hipSuccess &&
        cus > 0)
      S.num_cus = cus;
  }
  HIPCHK(h, dalloc(&S.rflag, (2 * ne - 1 + QW_RANGE - 1) / QW_RANGE));
  HIPCHK(h, dalloc(&S.rlist, (2 * ne - 1 + QW_RANGE - 1) / QW_RANGE));
  HIPCHK(h, dalloc(&S.zmark, nv)); HIPCHK(h, dalloc(&S.zlist, nv));
  HIPCHK(h, dalloc(&S.zkey64, nv)); HIPCHK(h, dalloc(&S.zkey64_tmp, nv));
  HIPCHK(h, dalloc(&S.zkey, nv));
  HIPCHK(h, dalloc(&S.zval, nv)); HIPCHK(h, dalloc(&S.zval_tmp, nv));
  HIPCHK(h, dalloc(&S.zhead, nv)); HIPCHK(h, dalloc(&S.zend, nv));
  S.n_levels = nlev;
  S.n_sched = nsched;
  HIPCHK(h, dalloc(&S.lvl_ptr, nlev + 1));
  S.strain_delta = sdelta;
  HIPCHK(h, dalloc(&S.sslot, nslot));
  HIPCHK(h, dalloc(&S.spre, nslot)); HIPCHK(h, dalloc(&S.slast, nv)); HIPCHK(h, dalloc(&S.lvl_active, std::max(nlev, 1)));
  S.plog_cap = 1 << 20;
  HIPCHK(h, dalloc(&S.plog, S.plog_cap));
  HIPCHK(h, dalloc(&S.counters, COUNTER_INTS));
  HIPCHK(h, dalloc(&S.spread, SPREAD_SLOTS * SPREAD_STRIDE));
  S.nscene = nscene > 1 ? nscene : 0;
  if (S.nscene > 1) {
    HIPCHK(h, dalloc(&S.scn, (size_t)SCN_FIELDS * S.nscene));
    HIPCHK(h, hipMemset(S.scn, 0, sizeof(int32_t) * SCN_FIELDS * S.nscene));
    HIPCHK(h, dalloc(&S.vscene, (size_t)nv));
    HIPCHK(h, hipMemcpy(S.vscene, vscene.data(), sizeof(int32_t) * nv, hipMemcpyHostToDevice));
  }
  HIPCHK(h, alloc_sort_hist(S));
  HIPCHK(h, hipMemcpy(S.ev, ev.data(), sizeof(int4) * ne, hipMemcpyHostToDevice));
  HIPCHK(h, hipMemcpy(S.eflags, ef.data(), ne, hipMemcpyHostToDevice));
  HIPCHK(h, hipMemcpy(S.len0, l0.data(), sizeof(dvec4) * ne, hipMemcpyHostToDevice));
  HIPCHK(h, hipMemcpy(S.fixed, fixed.data(), nv, hipMemcpyHostToDevice));
  HIPCHK(h, hipMemcpy(S.lvl_ptr, lptr.data(), sizeof(int) * (nlev + 1), hipMemcpyHostToDevice));
  HIPCHK(h, hipMemcpy(S.sslot, sslot.data(), sizeof(StrainSlot) * nslot, hipMemcpyHostToDevice));
  HIPCHK(h, hipMemset(S.spre, 0, sizeof(uint32_t) * nslot));
  HIPCHK(h, hipMemcpy(S.slast, slast.data(), sizeof(int32_t) * nv, hipMemcpyHostToDevice));
  HIPCHK(h, hipMemset(S.counters, 0, sizeof(int) * COUNTER_INTS));
  HIPCHK(h, hipMemset(S.has, 0, nv));
  HIPCHK(h, hipMemset(S.dirty_e, 0, ((size_t)ne + 3) & ~(size_t)3));
  HIPCHK(h, hipMemcpy(S.vadj_ptr, vptr.data(), sizeof(int32_t) * ((size_t)nv + 1), hipMemcpyHostToDevice));
  HIPCHK(h, hipMemcpy(S.vadj, vadj.data(), sizeof(int32_t) * vadj.size(), hipMemcpyHostToDevice));
  HIPCHK(h, hipMemset(S.dpre, 0, 2 * (size_t)ne - 1));
  HIPCHK(h, hipMemset(S.rflag, 0, sizeof(int32_t) * ((2 * (size_t)ne - 1 + QW_RANGE - 1) / QW_RANGE)));
  if (nbody > 0) {
    // members ascending by vertex id, and in order of first appearance in
    // element order (updateFinalForRG's visiting order)
    std::vector<int32_t> bptr((size_t)nbody + 1, 0), bfirst((size_t)nbody + 1, 0);
    std::vector<uint32_t> bmem;
    std::vector<int32_t> bvisit;
    for (int v = 0; v < nv; ++v)
      if (vbody[v] >= 0) bptr[(size_t)vbody[v] + 1]++;
    for (int b = 0; b < nbody; ++b) {
      if (bptr[(size_t)b + 1] < 2) return fail(h, CS_ERR_ARG, "a movable body needs two or more vertices");
      bptr[(size_t)b + 1] += bptr[b];
    }
    bmem.resize((size_t)bptr[nbody]);
    {
      std::vector<int32_t> fill(bptr.begin(), bptr.end() - 1);
      for (int v = 0; v < nv; ++v)
        if (vbody[v] >= 0) bmem[(size_t)fill[vbody[v]]++] = (uint32_t)v;
    }
    std::vector<std::vector<int32_t>> order((size_t)nbody);
    std::vector<uint8_t> seen((size_t)nv, 0);
    for (int e = 0; e < nt; ++e)
      for (int j = 0; j < 3; ++j) {
        const int v = m->tris[3 * e + j];
        if (vbody[v] >= 0 && !seen[v]) {
          seen[v] = 1;
          order[vbody[v]].push_back(v);
        }
      }
    for (int b = 0; b < nbody; ++b) {
      bfirst[(size_t)b + 1] = bfirst[b] + (int)order[b].size();
      bvisit.insert(bvisit.end(), order[b].begin(), order[b].end());
    }
    if ((int)bvisit.size() != bptr[nbody])
      return fail(h, CS_ERR_ARG, "every movable vertex must belong to a triangle of its body");
    std::vector<dvec4> com((size_t)nbody), zero((size_t)nbody, dvec4{0.0, 0.0, 0.0, 0.0});
    for (int b = 0; b < nbody; ++b)
      com[b] = dvec4{m->body_com[3 * b], m->body_com[3 * b + 1], m->body_com[3 * b + 2], 0.0};
    S.nbody = nbody;
    HIPCHK(h, dalloc(&S.vbody, nv)); HIPCHK(h, dalloc(&S.bptr, nbody + 1));
    HIPCHK(h, dalloc(&S.bmem, bmem.size())); HIPCHK(h, dalloc(&S.bfirst, nbody + 1));
    HIPCHK(h, dalloc(&S.bvisit, bvisit.size())); HIPCHK(h, dalloc(&S.bmass, nbody));
    HIPCHK(h, dalloc(&S.bcom, nbody)); HIPCHK(h, dalloc(&S.bvel, nbody));
    HIPCHK(h, dalloc(&S.bmrg, nbody)); HIPCHK(h, dalloc(&S.rg_imp, nbody));
    HIPCHK(h, dalloc(&S.rg_num, nbody));
    HIPCHK(h, hipMemcpy(S.vbody, vbody.data(), sizeof(int32_t) * nv, hipMemcpyHostToDevice));
    HIPCHK(h, hipMemcpy(S.bptr, bptr.data(), sizeof(int32_t) * (nbody + 1), hipMemcpyHostToDevice));
    HIPCHK(h, hipMemcpy(S.bmem, bmem.data(), sizeof(uint32_t) * bmem.size(), hipMemcpyHostToDevice));
    HIPCHK(h, hipMemcpy(S.bfirst, bfirst.data(), sizeof(int32_t) * (nbody + 1), hipMemcpyHostToDevice));
    HIPCHK(h, hipMemcpy(S.bvisit, bvisit.data(), sizeof(int32_t) * bvisit.size(), hipMemcpyHostToDevice));
    HIPCHK(h, hipMemcpy(S.bmass, m->body_mass, sizeof(double) * nbody, hipMemcpyHostToDevice));
    HIPCHK(h, hipMemcpy(S.bcom, com.data(), sizeof(dvec4) * nbody, hipMemcpyHostToDevice));
    // the reference reads mrg_com before its first write (an empty vector);
    // it starts at the body's center of mass here
    HIPCHK(h, hipMemcpy(S.bmrg, com.data(), sizeof(dvec4) * nbody, hipMemcpyHostToDevice));
    HIPCHK(h, hipMemcpy(S.bvel, zero.data(), sizeof(dvec4) * nbody, hipMemcpyHostToDevice));
    HIPCHK(h, hipMemcpy(S.rg_imp, zero.data(), sizeof(dvec4) * nbody, hipMemcpyHostToDevice));
    HIPCHK(h, hipMemset(S.rg_num, 0, sizeof(int32_t) * nbody));
  }
  h->assembled = true;
  h->assemble_gen++;  // buffers changed: a captured step graph is stale
  {  // launch shapes: the cache entry of this device, size and batch, or the defaults
    hipDeviceProp_t prop;
    std::string arch = "gpu";
    if (hipGetDeviceProperties(&prop, h->P.device) == hipSuccess) arch = prop.gcnArchName;
    for (char& ch : arch)
      if (ch == ' ' || ch == ':') ch = '_';
    int lg = 0;
    while (lg < 40 && (1ll << (lg + 1)) <= (long long)std::max(ne, 1)) ++lg;
    h->tuner = cs_solver::Tuner();
    h->tuner.key = arch + ":ne2^" + std::to_string(lg) + (S.nscene > 1 ? ":batch" : ":single");
    int tr = 8, na = 128, qu = 4096;
    if (tune_cache_lookup(h->tuner.key, &tr, &na, &qu)) h->tuner.tuned = 2;
    S.tune_trav = tr;
    S.tune_narrow = na;
    S.tune_query = qu;
  }
  h->state_set = false;
  return CS_OK;
}

int cs_set_state(cs_solver* h, const double* x, const double* vel, const uint8_t* has_collsn) {
  if (!h || !x || !vel) return CS_ERR_ARG;
  if (!h->assembled) return fail(h, CS_ERR_ARG, "cs_assemble has not been called");
  HIPCHK(h, hipSetDevice(h->P.device));
  DevState& S = h->S;
  std::vector<dvec4> buf;
  to_dvec4(x, S.nv, buf);
  HIPCHK(h, hipMemcpy(S.x, buf.data(), sizeof(dvec4) * S.nv, hipMemcpyHostToDevice));
  HIPCHK(h, hipMemcpy(S.x_old, buf.data(), sizeof(dvec4) * S.nv, hipMemcpyHostToDevice));
  to_dvec4(vel, S.nv, buf);
  HIPCHK(h, hipMemcpy(S.vel, buf.data(), sizeof(dvec4) * S.nv, hipMemcpyHostToDevice));
  HIPCHK(h, hipMemset(S.av, 0, sizeof(dvec4) * S.nv));
  if (has_collsn) HIPCHK(h, hipMemcpy(S.has, has_collsn, S.nv, hipMemcpyHostToDevice));
  else HIPCHK(h, hipMemset(S.has, 0, S.nv));
  if (S.nbody > 0) {  // initSurfaceState clears the RG accumulators (cdinit.cpp:192-197)
    HIPCHK(h, hipMemset(S.rg_imp, 0, sizeof(dvec4) * S.nbody));
    HIPCHK(h, hipMemset(S.rg_num, 0, sizeof(int32_t) * S.nbody));
  }
  h->state_set = true;
  return CS_OK;
}

int cs_get_state(cs_solver* h, double* x, double* vel, double* avg_vel, double* x_old,
                 uint8_t* has_collsn) {
  if (!h) return CS_ERR_ARG;
  if (!h->assembled) return fail(h, CS_ERR_ARG, "cs_assemble has not been called");
  HIPCHK(h, hipSetDevice(h->P.device));
  HIPCHK(h, hipStreamSynchronize(h->st));
  DevState& S = h->S;
  std::vector<dvec4> buf((size_t)S.nv);
  struct {
    double* dst;
    dvec4* src;
  } items[4] = {{x, S.x}, {vel, S.vel}, {avg_vel, S.av}, {x_old, S.x_old}};
  for (auto& it : items) {
    if (!it.dst) continue;
    HIPCHK(h, hipMemcpy(buf.data(), it.src, sizeof(dvec4) * S.nv, hipMemcpyDeviceToHost));
    from_dvec4(buf, it.dst);
  }
  if (has_collsn) HIPCHK(h, hipMemcpy(has_collsn, S.has, S.nv, hipMemcpyDeviceToHost));
  return CS_OK;
}

int cs_record_origin_position(cs_solver* h) {
  int rc = check_ready(h);
  if (rc) return rc;
  launch_record_origin(h->st, h->S);
  HIPCHK(h, hipGetLastError());
  return CS_OK;
}

int cs_predict(cs_solver* h) {
  int rc = check_ready(h);
  if (rc) return rc;
  launch_predict(h->st, h->S, h->P.dt);
  HIPCHK(h, hipGetLastError());
  return CS_OK;
}

namespace {

// The launches of a resolve before its first host synchronisation (kernels
// only, so that they can be captured in a graph): the dirty-range marks of
// the last pass are cleared while their count is still in the counter block,
// then the counters, the union-find (makeSet, dcollid.cpp:1015-1031), the
// batched scenes' pass state, the recorded impulses, the zone marks and
// computeAverageVelocity.
void resolve_prologue(cs_solver* h, bool stats) {
  DevState& S = h->S;
  const cs_params& P = h->P;
  if (S.ne >= 2) launch_clear_ranges(h->st, S);
  launch_fill_u32(h->st, S.counters, COUNTER_INTS, 0u);
  if (stats) launch_fill_u32(h->st, S.spread, (int64_t)SPREAD_SLOTS * SPREAD_STRIDE, 0u);
  launch_uf_init(h->st, S);
  launch_scene_reset(h->st, S);
  if (P.record_impulses && S.pimp) {  // the impulses of this resolve's last pass with a hit
    launch_fill_u32(h->st, S.pimp, (int64_t)S.nv * (int64_t)(sizeof(dvec4) / 4), 0u);
    launch_fill_u32(h->st, S.pnum, S.nv, 0u);
  }
  launch_fill_u32(h->st, S.zmark, S.nv, 0u);
  launch_average_velocity(h->st, S, P.dt);
}

// the predictor and the resolve prologue as one graph, captured again when
// what its launches depend on has changed
int ensure_step_graph(cs_solver* h, bool stats) {
  cs_solver::GraphKey key;
  key.assemble_gen = h->assemble_gen;
  key.dt = h->P.dt;
  key.record = h->P.record_impulses != 0;
  key.stats = stats;
  key.pimp = h->S.pimp;
  if (h->graph_exec && key == h->graph_key) return CS_OK;
  if (h->graph_exec) (void)hipGraphExecDestroy(h->graph_exec);
  if (h->graph) (void)hipGraphDestroy(h->graph);
  h->graph_exec = nullptr;
  h->graph = nullptr;
  HIPCHK(h, hipStreamBeginCapture(h->st, hipStreamCaptureModeThreadLocal));
  launch_predict(h->st, h->S, h->P.dt);
  resolve_prologue(h, stats);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(h->st, &g);
  if (e != hipSuccess) return fail(h, CS_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
  h->graph = g;
  HIPCHK(h, hipGraphInstantiate(&h->graph_exec, g, nullptr, nullptr, 0));
  h->graph_key = key;
  return CS_OK;
}

// with_predict_graph: the predictor runs first, inside the captured graph
int resolve_impl(cs_solver* h, cs_step_report* report, bool with_predict_graph) {
  int rc = check_ready(h);
  if (rc) return rc;
  if (h->P.learned && !h->L.ready)
    return fail(h, CS_ERR_ARG, "the learned form needs weights (cs_learned_init / _set_weights)");
  DevState& S = h->S;
  const cs_params& P = h->P;
  cs_step_report R{};
  h->passes.clear();
  h->ptimes.clear();
  h->plog_rows = 0;
  h->ev_used = 0;
  h->has_collision = 0;
  h->syncs = 0;
  h->prev = PrevPass();
  hipEvent_t t0, t_strain0, t_end;
  std::vector<hipEvent_t> strain_marks;
  const bool spread_stats = P.precision == 1 || P.learned;
  cs_solver::Tuner& tu = h->tuner;
  if (P.autotune && !tu.active && tu.tuned == 0 && tu.step == 0) tu.active = true;
  int saved_timing = -1;
  if (tu.active) {  // this step's candidates; kernel times need the timing markers
    S.tune_trav = kTuneTrav[tu.step % 4];
    S.tune_narrow = kTuneNarrow[tu.step % 3];
    S.tune_query = kTuneQuery[tu.step % 4];
    if (h->P.pass_timing < 1) {
      saved_timing = h->P.pass_timing;
      h->P.pass_timing = 1;
    }
  }
  if (P.record_impulses && !S.pimp) {
    HIPCHK(h, dalloc(&S.pimp, S.nv));
    HIPCHK(h, dalloc(&S.pnum, S.nv));
  }
  if (with_predict_graph) {
    if ((rc = ensure_step_graph(h, spread_stats))) return rc;
    mark(h, &t0);
    HIPCHK(h, hipGraphLaunch(h->graph_exec, h->st));
  } else {
    mark(h, &t0);
    resolve_prologue(h, spread_stats);
  }
  HIPCHK(h, hipGetLastError());
  int pass_id = 0;
  bool any = false;
  int status = CS_OK;
  // a non-finite average velocity ends the step before any pass, where the
  // reference calls clean_up(ERROR) (dcollid.cpp:178-184)
  if ((rc = read_counters(h))) return rc;
  if (h->h_counters[CNT_ERR]) status = device_error(h);
  // detectProximity (dcollid.cpp:390-406)
  if (!status) {
    rc = run_pass(h, CS_PASS_PROXIMITY, 0, pass_id++, &any);
    if (rc) status = rc;
  }
  // detectCollision (dcollid.cpp:430-468)
  bool is_collision = true;
  int niter = 1;
  while (!status && is_collision) {
    rc = run_pass(h, CS_PASS_CCD, niter - 1, pass_id++, &is_collision);
    if (rc) { status = rc; break; }
    if (niter == 1 && is_collision) h->has_collision = 1;
    R.n_ccd_passes++;
    if (++niter > P.max_ccd_passes) break;
  }
  // computeImpactZone (dcollid.cpp:227-265)
  if (!status && is_collision) {
    int zi = 0;
    while (is_collision) {
      if (zi >= P.max_zone_passes) { status = CS_WARN_ZONE_CAP; break; }
      rc = run_pass(h, CS_PASS_ZONE, zi, pass_id++, &is_collision);
      if (rc) { status = rc; break; }
      R.n_zone_passes++;
      zi++;
    }
  }
  mark(h, &t_strain0);
  if (status == CS_OK || status == CS_WARN_ZONE_CAP) {
    launch_domain_boundary(h->st, S, P.dom_lo, P.dom_hi, P.dt, P.friction, P.thickness);
    launch_final_position(h->st, S, P.dt);
    HIPCHK(h, hipMemsetAsync(S.counters + CNT_STRAIN_FLAG, 0, sizeof(int) * 4, h->st));
    HIPCHK(h, hipMemsetAsync(S.counters + CNT_STRAIN_LEVELS, 0, sizeof(int) * 2, h->st));
    HIPCHK(h, strain_limit(h->st, S, P.dt, P.max_strain_sweeps,
                           P.pass_timing >= 1 ? &strain_marks : nullptr));
    launch_final_velocity(h->st, S);
    launch_final_bodies(h->st, S, P.dt);  // updateFinalForRG
    HIPCHK(h, hipGetLastError());
  }
  mark(h, &t_end);
  rc = read_counters(h);
  if (rc) return rc;
  if (R.n_zone_passes) R.n_zones = h->h_counters[CNT_ZONES] + S.nbody;
  for (size_t i = 0; i < h->passes.size() && i < h->ptimes.size(); ++i) {
    cs_pass_info& pi = h->passes[i];
    const PassTimes& T = h->ptimes[i];
    pi.ms_broad = elapsed(T.e0, T.e1);
    pi.ms_narrow = elapsed(T.e1, T.e2);
    pi.ms_response = elapsed(T.e2, T.e3);
    pi.ms_traverse_kernel = elapsed(T.t0, T.t1);
    pi.ms_narrow_kernel = elapsed(T.n0, T.n1);
  }
  if (tu.active) {  // per-unit kernel times of this step's candidates
    double trav = 0.0, query = 0.0, narrow = 0.0;
    int64_t lanes = 0;
    int nfull = 0, ninc = 0;
    for (const cs_pass_info& pi : h->passes) {
      if (pi.incremental) {
        query += pi.ms_traverse_kernel;
        ++ninc;
      } else {
        trav += pi.ms_traverse_kernel;
        ++nfull;
      }
      narrow += pi.ms_narrow_kernel;
      lanes += std::max<int64_t>(pi.tested_pairs, 0);
    }
    const int a = tu.step % 4, b = tu.step % 3;
    if (nfull) { tu.t_trav[a] += trav / nfull; tu.n_trav[a]++; }
    if (ninc) { tu.t_query[a] += query / ninc; tu.n_query[a]++; }
    if (lanes) { tu.t_narrow[b] += narrow / (double)lanes; tu.n_narrow[b]++; }
    if (saved_timing >= 0) h->P.pass_timing = saved_timing;
    if (++tu.step >= kTuneSteps) {  // keep the fastest of each (defaults where unmeasured)
      auto best = [](const double* t, const int* n, int m, const int* cand, int def) {
        int bi = -1;
        for (int i = 0; i < m; ++i)
          if (n[i] && (bi < 0 || t[i] / n[i] < t[bi] / n[bi])) bi = i;
        return bi < 0 ? def : cand[bi];
      };
      S.tune_trav = best(tu.t_trav, tu.n_trav, 4, kTuneTrav, 8);
      S.tune_narrow = best(tu.t_narrow, tu.n_narrow, 3, kTuneNarrow, 128);
      S.tune_query = best(tu.t_query, tu.n_query, 4, kTuneQuery, 4096);
      tu.active = false;
      tu.tuned = 1;
      tune_cache_store(tu.key, S.tune_trav, S.tune_narrow, S.tune_query);
    }
  }
  if (status == CS_OK || status == CS_WARN_ZONE_CAP) {
    R.strain_sweeps = h->h_counters[CNT_STRAIN_SWEEPS];
    R.strain_edges = h->h_counters[CNT_STRAIN_EDGES];
    R.degenerate_edges = h->h_counters[CNT_STRAIN_DEGEN];
    R.strain_levels = h->h_counters[CNT_STRAIN_LEVELS];
    R.strain_slots = h->h_counters[CNT_STRAIN_EVALS];
    int e = device_error(h);
    if (e) status = e;
  }
  R.status = status;
  R.n_passes = (int)h->passes.size();
  R.has_collision = h->has_collision;
  for (auto& p : h->passes) {
    if (p.candidates >= 0 && R.candidates >= 0) R.candidates += p.candidates;
    else R.candidates = -1;
    if (p.narrow_pairs >= 0 && R.narrow_pairs >= 0) R.narrow_pairs += p.narrow_pairs;
    else R.narrow_pairs = -1;
    R.hit_pairs += p.hit_pairs;
    R.hit_subtests += p.hit_subtests;
    R.ms_broad += p.ms_broad;
    R.ms_narrow += p.ms_narrow;
    R.ms_response += p.ms_response;
    if (p.kind == CS_PASS_ZONE) R.ms_zone += p.ms_broad + p.ms_narrow + p.ms_response;
  }
  R.ms_total = elapsed(t0, t_end);
  R.ms_strain = elapsed(t_strain0, t_end);
  for (size_t i = 0; i + 1 < strain_marks.size(); i += 2)
    R.ms_strain_kernel += elapsed(strain_marks[i], strain_marks[i + 1]);
  for (hipEvent_t e : strain_marks) (void)hipEventDestroy(e);
  if (spread_stats) {  // the statistics' per-work-group copies: sums and maxima
    std::vector<int> sp((size_t)SPREAD_SLOTS * SPREAD_STRIDE);
    HIPCHK(h, hipMemcpyAsync(sp.data(), S.spread, sizeof(int) * sp.size(), hipMemcpyDeviceToHost,
                             h->st));
    HIPCHK(h, hipStreamSynchronize(h->st));
    h->syncs++;
    int64_t sum[SP_COUNT] = {};
    float mx[SP_COUNT] = {};
    for (int k = 0; k < SPREAD_SLOTS; ++k)
      for (int c = 0; c < SP_COUNT; ++c) {
        const int v = sp[(size_t)k * SPREAD_STRIDE + c];
        sum[c] += v;
        float f;
        std::memcpy(&f, &v, sizeof(float));
        mx[c] = std::max(mx[c], f);
      }
    R.mx_screened = sum[SP_MX_SCREENED];
    R.mx_rechecked = sum[SP_MX_RECHECK];
    R.mx_err_dist = mx[SP_MX_EDIST];
    R.mx_err_param = mx[SP_MX_EPARAM];
    R.mx_bound_ratio = mx[SP_MX_RATIO];
    if (P.learned) {
      cs_learned_stats& ls = h->lstats;
      ls.scored += sum[SP_LN_SCORED];
      ls.exists += sum[SP_LN_EXISTS];
      ls.predicted_exists += sum[SP_LN_PRED];
      ls.agree += sum[SP_LN_AGREE];
      ls.newton_roots += sum[SP_LN_ROOTS];
      ls.newton_iters += sum[SP_LN_ITERS];
      ls.newton_iters_mid += sum[SP_LN_ITERS_MID];
      ls.max_root_err = std::max(ls.max_root_err, (double)mx[SP_LN_EMAX]);
    }
  }
  R.host_syncs = h->syncs;
  if (report) *report = R;
  return status;
}

}  // namespace

int cs_resolve(cs_solver* h, cs_step_report* report) { return resolve_impl(h, report, false); }

int cs_step(cs_solver* h, cs_step_report* report) {
  if (h && h->P.graph_step) return resolve_impl(h, report, true);
  int rc = cs_predict(h);
  if (rc) return rc;
  return cs_resolve(h, report);
}

namespace {

int resolve_batch_impl(cs_solver* h, cs_step_report* report, cs_scene_report* per_scene,
                       int32_t nscene, bool with_predict_graph) {
  if (!h) return CS_ERR_ARG;
  if (!h->assembled) return fail(h, CS_ERR_ARG, "cs_assemble has not been called");
  if (h->S.nscene <= 1 || nscene != h->S.nscene || !per_scene)
    return fail(h, CS_ERR_ARG, "nscene must equal the assembled batch's scene count");
  const int rc = resolve_impl(h, report, with_predict_graph);
  if (rc != CS_OK && rc != CS_WARN_ZONE_CAP) return rc;
  std::vector<int32_t> f((size_t)SCN_FIELDS * nscene);
  HIPCHK(h, hipMemcpyAsync(f.data(), h->S.scn, sizeof(int32_t) * f.size(), hipMemcpyDeviceToHost,
                           h->st));
  HIPCHK(h, hipStreamSynchronize(h->st));
  for (int s = 0; s < nscene; ++s) {
    cs_scene_report& r = per_scene[s];
    r = cs_scene_report{};
    r.n_ccd_passes = f[(size_t)SCN_CCD * nscene + s];
    r.n_zone_passes = f[(size_t)SCN_ZONE * nscene + s];
    r.n_passes = 1 + r.n_ccd_passes + r.n_zone_passes;
    r.has_collision = f[(size_t)SCN_HASCOL * nscene + s];
    r.zone_cap = f[(size_t)SCN_PHASE * nscene + s] == 2;
    r.hit_pairs = f[(size_t)SCN_HITPAIRS * nscene + s];
  }
  return rc;
}

}  // namespace

int cs_resolve_batch(cs_solver* h, cs_step_report* report, cs_scene_report* per_scene,
                     int32_t nscene) {
  return resolve_batch_impl(h, report, per_scene, nscene, false);
}

int cs_step_batch(cs_solver* h, cs_step_report* report, cs_scene_report* per_scene,
                  int32_t nscene) {
  if (h && h->P.graph_step) return resolve_batch_impl(h, report, per_scene, nscene, true);
  int rc = cs_predict(h);
  if (rc) return rc;
  return cs_resolve_batch(h, report, per_scene, nscene);
}

int cs_has_collision(cs_solver* h) { return h ? h->has_collision : 0; }

int cs_get_bodies(cs_solver* h, double* com, double* com_vel) {
  if (!h) return CS_ERR_ARG;
  if (!h->assembled) return fail(h, CS_ERR_ARG, "cs_assemble has not been called");
  DevState& S = h->S;
  if (S.nbody == 0) return CS_OK;
  HIPCHK(h, hipSetDevice(h->P.device));
  HIPCHK(h, hipStreamSynchronize(h->st));
  std::vector<dvec4> buf((size_t)S.nbody);
  if (com) {
    HIPCHK(h, hipMemcpy(buf.data(), S.bcom, sizeof(dvec4) * S.nbody, hipMemcpyDeviceToHost));
    from_dvec4(buf, com);
  }
  if (com_vel) {
    HIPCHK(h, hipMemcpy(buf.data(), S.bvel, sizeof(dvec4) * S.nbody, hipMemcpyDeviceToHost));
    from_dvec4(buf, com_vel);
  }
  return CS_OK;
}

int cs_math_eval(int32_t device, int32_t fn, const double* x, double* y, int64_t n) {
  if (!x || !y || n < 0 || fn < 0 || fn > 3) return CS_ERR_ARG;
  if (n == 0) return CS_OK;
  if (hipSetDevice(device) != hipSuccess) return CS_ERR_NO_DEVICE;
  double *dx = nullptr, *dy = nullptr;
  hipError_t e = dalloc(&dx, (size_t)n);
  if (e == hipSuccess) e = dalloc(&dy, (size_t)n);
  if (e == hipSuccess) e = hipMemcpy(dx, x, sizeof(double) * (size_t)n, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = launch_math_eval(nullptr, fn, dx, dy, n);
  if (e == hipSuccess) e = hipMemcpy(y, dy, sizeof(double) * (size_t)n, hipMemcpyDeviceToHost);
  dfree(dx);
  dfree(dy);
  return e == hipSuccess ? CS_OK : CS_ERR_HIP;
}

int cs_synchronize(cs_solver* h) {
  if (!h) return CS_ERR_ARG;
  HIPCHK(h, hipStreamSynchronize(h->st));
  return CS_OK;
}

int32_t cs_pass_count(cs_solver* h) { return h ? (int32_t)h->passes.size() : 0; }

int cs_pass_infos(cs_solver* h, cs_pass_info* out, int32_t cap) {
  if (!h || !out) return CS_ERR_ARG;
  const int32_t n = std::min<int32_t>(cap, (int32_t)h->passes.size());
  for (int32_t i = 0; i < n; ++i) out[i] = h->passes[i];
  return CS_OK;
}

int64_t cs_pair_log_count(cs_solver* h) { return h ? h->plog_rows : 0; }

int cs_pair_log(cs_solver* h, int32_t* out4, int64_t cap) {
  if (!h || !out4) return CS_ERR_ARG;
  const int64_t n = std::min<int64_t>(cap, h->plog_rows);
  if (n <= 0) return CS_OK;
  HIPCHK(h, hipSetDevice(h->P.device));
  HIPCHK(h, hipMemcpy(out4, h->S.plog, sizeof(int4) * (size_t)n, hipMemcpyDeviceToHost));
  return CS_OK;
}

int cs_zone_labels(cs_solver* h, int32_t* out) {
  int rc = check_ready(h);
  if (rc) return rc;
  if (!out) return CS_ERR_ARG;
  int32_t* d = nullptr;
  HIPCHK(h, dalloc(&d, h->S.nv));
  launch_zone_labels(h->st, h->S, d);
  hipError_t e = hipMemcpyAsync(out, d, sizeof(int32_t) * h->S.nv, hipMemcpyDeviceToHost, h->st);
  if (e == hipSuccess) e = hipStreamSynchronize(h->st);
  (void)hipFree(d);
  if (e != hipSuccess) return fail(h, CS_ERR_HIP, hipGetErrorString(e));
  return CS_OK;
}

int cs_last_impulses(cs_solver* h, double* imp, int32_t* num) {
  if (!h) return CS_ERR_ARG;
  if (!h->assembled) return fail(h, CS_ERR_ARG, "cs_assemble has not been called");
  DevState& S = h->S;
  if (!S.pimp) return fail(h, CS_ERR_ARG, "record_impulses was not set for the last resolve");
  HIPCHK(h, hipSetDevice(h->P.device));
  HIPCHK(h, hipStreamSynchronize(h->st));
  if (imp) {
    std::vector<dvec4> buf((size_t)S.nv);
    HIPCHK(h, hipMemcpy(buf.data(), S.pimp, sizeof(dvec4) * S.nv, hipMemcpyDeviceToHost));
    from_dvec4(buf, imp);
  }
  if (num) HIPCHK(h, hipMemcpy(num, S.pnum, sizeof(int32_t) * S.nv, hipMemcpyDeviceToHost));
  return CS_OK;
}

int cs_write_vtk(cs_solver* h, const char* path) {
  if (!h || !path) return CS_ERR_ARG;
  DevState& S = h->S;
  const int nv = S.nv, ne = S.ne;
  std::vector<double> x(3 * (size_t)nv), imp(3 * (size_t)nv);
  std::vector<int32_t> num((size_t)nv);
  int rc = cs_get_state(h, x.data(), nullptr, nullptr, nullptr, nullptr);
  if (rc) return rc;
  if ((rc = cs_last_impulses(h, imp.data(), num.data()))) return rc;
  std::vector<int4> ev((size_t)std::max(ne, 1));
  HIPCHK(h, hipMemcpy(ev.data(), S.ev, sizeof(int4) * ne, hipMemcpyDeviceToHost));
  // points numbered in order of first appearance in element order
  std::vector<int> idx((size_t)nv, -1), order;
  order.reserve((size_t)nv);
  for (int e = 0; e < ne; ++e) {
    const int vv[3] = {ev[e].x, ev[e].y, ev[e].z};
    for (int j = 0; j < 3; ++j)
      if (vv[j] >= 0 && idx[vv[j]] < 0) {
        idx[vv[j]] = (int)order.size();
        order.push_back(vv[j]);
      }
  }
  std::unique_ptr<FILE, int (*)(FILE*)> f(std::fopen(path, "w"), &std::fclose);
  if (!f) return fail(h, CS_ERR_ARG, std::string("cannot open ") + path);
  FILE* o = f.get();
  std::fprintf(o, "<?xml version=\"1.0\"?>\n<VTKFile type=\"PolyData\" version=\"0.1\" "
                  "byte_order=\"LittleEndian\">\n<PolyData>\n");
  std::fprintf(o, "<Piece NumberOfPoints=\"%d\" NumberOfVerts=\"0\" NumberOfLines=\"0\" "
                  "NumberOfStrips=\"0\" NumberOfPolys=\"%d\">\n", (int)order.size(), ne);
  std::fprintf(o, "<PointData Vectors=\"CollsnImpulse\">\n<DataArray type=\"Float32\" "
                  "Name=\"CollsnImpulse\" NumberOfComponents=\"3\" format=\"ascii\">\n");
  for (int v : order)
    std::fprintf(o, "%.9g %.9g %.9g\n", (float)imp[3 * v], (float)imp[3 * v + 1], (float)imp[3 * v + 2]);
  std::fprintf(o, "</DataArray>\n</PointData>\n<CellData Scalars=\"CollsnRegion\">\n"
                  "<DataArray type=\"UInt8\" Name=\"CollsnRegion\" NumberOfComponents=\"3\" "
                  "format=\"ascii\">\n");
  for (int e = 0; e < ne; ++e) {
    const int vv[3] = {ev[e].x, ev[e].y, ev[e].z};
    bool red = false;
    for (int j = 0; j < 3; ++j)
      if (vv[j] >= 0 && num[vv[j]] > 0) red = true;
    std::fprintf(o, red ? "255 0 0\n" : "0 255 0\n");
  }
  std::fprintf(o, "</DataArray>\n</CellData>\n<Points>\n<DataArray type=\"Float32\" "
                  "Name=\"Points\" NumberOfComponents=\"3\" format=\"ascii\">\n");
  for (int v : order)
    std::fprintf(o, "%.9g %.9g %.9g\n", (float)x[3 * v], (float)x[3 * v + 1], (float)x[3 * v + 2]);
  std::fprintf(o, "</DataArray>\n</Points>\n<Polys>\n<DataArray type=\"Int32\" "
                  "Name=\"connectivity\" format=\"ascii\">\n");
  for (int e = 0; e < ne; ++e) {
    if (ev[e].z >= 0) std::fprintf(o, "%d %d %d\n", idx[ev[e].x], idx[ev[e].y], idx[ev[e].z]);
    else std::fprintf(o, "%d %d\n", idx[ev[e].x], idx[ev[e].y]);
  }
  std::fprintf(o, "</DataArray>\n<DataArray type=\"Int32\" Name=\"offsets\" format=\"ascii\">\n");
  long long off = 0;
  for (int e = 0; e < ne; ++e) {
    off += ev[e].z >= 0 ? 3 : 2;
    std::fprintf(o, "%lld\n", off);
  }
  std::fprintf(o, "</DataArray>\n</Polys>\n</Piece>\n</PolyData>\n</VTKFile>\n");
  if (std::ferror(o)) return fail(h, CS_ERR_ARG, std::string("write failed: ") + path);
  return CS_OK;
}


int cs_save(cs_solver* h, const char* path) {
  int rc = check_ready(h);
  if (rc) return rc;
  if (!path) return CS_ERR_ARG;
  HIPCHK(h, hipStreamSynchronize(h->st));
  CkptHeader hd{};
  std::memcpy(hd.magic, kCkptMagic, 8);
  hd.nv = h->S.nv; hd.nt = h->S.nt; hd.nb = h->S.nb; hd.nbody = h->S.nbody;
  if ((rc = checkpoint_topology(h, &hd.topo))) return rc;
  hd.params = h->P;
  std::unique_ptr<FILE, int (*)(FILE*)> f(std::fopen(path, "wb"), &std::fclose);
  if (!f) return fail(h, CS_ERR_ARG, std::string("cannot open ") + path);
  if (std::fwrite(&hd, sizeof(hd), 1, f.get()) != 1) return fail(h, CS_ERR_ARG, "write failed");
  std::vector<char> buf;
  for (const CkptArray& a : checkpoint_arrays(h->S)) {
    buf.resize(a.bytes);
    HIPCHK(h, hipMemcpy(buf.data(), a.dev, a.bytes, hipMemcpyDeviceToHost));
    if (a.bytes && std::fwrite(buf.data(), 1, a.bytes, f.get()) != a.bytes)
      return fail(h, CS_ERR_ARG, "write failed");
  }
  return CS_OK;
}

int cs_load(cs_solver* h, const char* path) {
  if (!h || !path) return CS_ERR_ARG;
  if (!h->assembled) return fail(h, CS_ERR_ARG, "cs_assemble has not been called");
  HIPCHK(h, hipSetDevice(h->P.device));
  HIPCHK(h, hipStreamSynchronize(h->st));
  std::unique_ptr<FILE, int (*)(FILE*)> f(std::fopen(path, "rb"), &std::fclose);
  if (!f) return fail(h, CS_ERR_ARG, std::string("cannot open ") + path);
  CkptHeader hd{};
  if (std::fread(&hd, sizeof(hd), 1, f.get()) != 1 || std::memcmp(hd.magic, kCkptMagic, 8) != 0)
    return fail(h, CS_ERR_ARG, "not a checkpoint of this library version");
  uint64_t topo = 0;
  int rc = checkpoint_topology(h, &topo);
  if (rc) return rc;
  if (hd.nv != h->S.nv || hd.nt != h->S.nt || hd.nb != h->S.nb || hd.nbody != h->S.nbody ||
      hd.topo != topo)
    return fail(h, CS_ERR_ARG, "checkpoint topology differs from the assembled mesh");
  std::vector<char> buf;
  for (const CkptArray& a : checkpoint_arrays(h->S)) {
    buf.resize(a.bytes);
    if (a.bytes && std::fread(buf.data(), 1, a.bytes, f.get()) != a.bytes)
      return fail(h, CS_ERR_ARG, "checkpoint is truncated");
    HIPCHK(h, hipMemcpy(a.dev, buf.data(), a.bytes, hipMemcpyHostToDevice));
  }
  const int device = h->P.device;
  h->P = hd.params;
  h->P.device = device;
  h->state_set = true;
  h->prev = PrevPass();
  h->seed = SeedPairs();
  return CS_OK;
}

int cs_zone_labels_noalloc(cs_solver* h, int32_t* out) {
  int rc = check_ready(h);
  if (rc) return rc;
  if (!out) return CS_ERR_ARG;
  const int dev = h->P.device;
  if (dev < 0 || dev >= kMaxDevices || !g_scratch_done[dev])
    return fail(h, CS_ERR_INTERNAL, "no staging array was set up for this device");
  if (g_scratch_err[dev] != hipSuccess || !g_scratch_ptr[dev])
    return fail(h, CS_ERR_HIP, std::string("hipGetSymbolAddress(scratch) failed: ") +
                                   hipGetErrorString(g_scratch_err[dev]));
  if (h->S.nv > NOALLOC_MAX_VERTICES)
    return fail(h, CS_ERR_UNSUPPORTED,
                "cs_zone_labels_noalloc handles at most 4194304 vertices; use cs_zone_labels");
  int32_t* d = g_scratch_ptr[dev];  // every label is written, nothing to clear
  launch_zone_labels(h->st, h->S, d);
  HIPCHK(h, hipGetLastError());
  HIPCHK(h, hipMemcpyAsync(out, d, sizeof(int32_t) * h->S.nv, hipMemcpyDeviceToHost, h->st));
  HIPCHK(h, hipStreamSynchronize(h->st));
  return CS_OK;
}

// ---- learned time-of-impact proposer ----
int cs_learned_set_weights(cs_solver* h, const float* w) {
  if (!h || !w) return CS_ERR_ARG;
  HIPCHK(h, hipSetDevice(h->P.device));
  HIPCHK(h, learned_set_weights(h->L, h->st, w));
  return CS_OK;
}

int cs_learned_get_weights(cs_solver* h, float* w) {
  if (!h || !w) return CS_ERR_ARG;
  if (!h->L.ready) return fail(h, CS_ERR_ARG, "no weights have been set");
  HIPCHK(h, hipSetDevice(h->P.device));
  HIPCHK(h, learned_get_weights(h->L, h->st, w));
  return CS_OK;
}

int cs_learned_init(cs_solver* h, uint64_t seed) {
  if (!h) return CS_ERR_ARG;
  std::vector<float> w(CS_MLP_PARAMS);
  learned_init_host(seed, w.data());
  return cs_learned_set_weights(h, w.data());
}

int cs_learned_infer(cs_solver* h, const float* x, float* out, int64_t n) {
  if (!h || n < 0 || (n > 0 && (!x || !out))) return CS_ERR_ARG;
  if (!h->L.ready) return fail(h, CS_ERR_ARG, "no weights have been set");
  if (n == 0) return CS_OK;
  HIPCHK(h, hipSetDevice(h->P.device));
  float *dx = nullptr, *dout = nullptr;
  hipError_t e = dalloc(&dx, (size_t)n * 24);
  if (e == hipSuccess) e = dalloc(&dout, (size_t)n * 2);
  if (e == hipSuccess)
    e = hipMemcpyAsync(dx, x, sizeof(float) * 24 * (size_t)n, hipMemcpyHostToDevice, h->st);
  if (e == hipSuccess) e = learned_infer(h->L, h->st, dx, n, dout);
  if (e == hipSuccess)
    e = hipMemcpyAsync(out, dout, sizeof(float) * 2 * (size_t)n, hipMemcpyDeviceToHost, h->st);
  if (e == hipSuccess) e = hipStreamSynchronize(h->st);
  dfree(dx);
  dfree(dout);
  if (e != hipSuccess) return fail(h, CS_ERR_HIP, std::string("learned inference: ") + hipGetErrorString(e));
  return CS_OK;
}

int cs_learned_grad(cs_solver* h, const float* x, const float* exists, const float* toi,
                    int64_t n, float* grad, float* loss) {
  if (!h || !x || !exists || !toi || !grad || n <= 0 || n > (1 << 24)) return CS_ERR_ARG;
  if (!h->L.ready) return fail(h, CS_ERR_ARG, "no weights have been set");
  HIPCHK(h, hipSetDevice(h->P.device));
  float *dx = nullptr, *de = nullptr, *dt = nullptr, *dl = nullptr;
  hipError_t e = dalloc(&dx, (size_t)n * 24);
  if (e == hipSuccess) e = dalloc(&de, (size_t)n);
  if (e == hipSuccess) e = dalloc(&dt, (size_t)n);
  if (e == hipSuccess) e = dalloc(&dl, 1);
  if (e == hipSuccess)
    e = hipMemcpyAsync(dx, x, sizeof(float) * 24 * (size_t)n, hipMemcpyHostToDevice, h->st);
  if (e == hipSuccess)
    e = hipMemcpyAsync(de, exists, sizeof(float) * (size_t)n, hipMemcpyHostToDevice, h->st);
  if (e == hipSuccess)
    e = hipMemcpyAsync(dt, toi, sizeof(float) * (size_t)n, hipMemcpyHostToDevice, h->st);
  if (e == hipSuccess) e = learned_grad(h->L, h->st, dx, de, dt, nullptr, (int)n, dl);
  if (e == hipSuccess)
    e = hipMemcpyAsync(grad, h->L.grad, sizeof(float) * CS_MLP_PARAMS, hipMemcpyDeviceToHost, h->st);
  float l = 0.f;
  if (e == hipSuccess) e = hipMemcpyAsync(&l, dl, sizeof(float), hipMemcpyDeviceToHost, h->st);
  if (e == hipSuccess) e = hipStreamSynchronize(h->st);
  dfree(dx);
  dfree(de);
  dfree(dt);
  dfree(dl);
  if (e != hipSuccess) return fail(h, CS_ERR_HIP, std::string("learned gradient: ") + hipGetErrorString(e));
  if (loss) *loss = l;
  return CS_OK;
}

int cs_learned_collect(cs_solver* h, int64_t capacity, int32_t stride) {
  if (!h || capacity < 0 || capacity > INT32_MAX / 2 || stride < 1) return CS_ERR_ARG;
  HIPCHK(h, hipSetDevice(h->P.device));
  HIPCHK(h, hipStreamSynchronize(h->st));
  HIPCHK(h, learned_alloc(h->L));
  if (capacity == 0) {
    h->L.ds_cap = 0;
    dfree(h->L.dsx);
    dfree(h->L.dse);
    dfree(h->L.dst);
    h->L.ds_alloc = 0;
  } else {
    HIPCHK(h, learned_ensure_dataset(h->L, capacity));
    h->L.ds_cap = capacity;
  }
  h->L.ds_stride = stride;
  h->L.ds_salt = 0;
  HIPCHK(h, hipMemset(h->L.ds_count, 0, sizeof(unsigned long long)));
  return CS_OK;
}

int64_t cs_learned_dataset_size(cs_solver* h) {
  if (!h) return -1;
  if (hipSetDevice(h->P.device) != hipSuccess) return -1;
  return learned_dataset_size(h->L, h->st);
}

int cs_learned_dataset(cs_solver* h, float* x, float* exists, float* toi, int64_t cap) {
  if (!h || cap < 0) return CS_ERR_ARG;
  HIPCHK(h, hipSetDevice(h->P.device));
  const int64_t n = std::min<int64_t>(learned_dataset_size(h->L, h->st), cap);
  if (n < 0) return fail(h, CS_ERR_HIP, "dataset size");
  if (n == 0) return CS_OK;
  if (!x || !exists || !toi) return CS_ERR_ARG;
  HIPCHK(h, hipMemcpy(x, h->L.dsx, sizeof(float) * 24 * (size_t)n, hipMemcpyDeviceToHost));
  HIPCHK(h, hipMemcpy(exists, h->L.dse, sizeof(float) * (size_t)n, hipMemcpyDeviceToHost));
  HIPCHK(h, hipMemcpy(toi, h->L.dst, sizeof(float) * (size_t)n, hipMemcpyDeviceToHost));
  return CS_OK;
}

int cs_learned_train(cs_solver* h, int32_t epochs, int32_t batch, float lr, uint64_t seed,
                     float* loss) {
  if (!h || epochs < 0 || batch < 1 || !(lr > 0)) return CS_ERR_ARG;
  if (!h->L.ready) return fail(h, CS_ERR_ARG, "no weights have been set");
  HIPCHK(h, hipSetDevice(h->P.device));
  if (learned_dataset_size(h->L, h->st) <= 0) return fail(h, CS_ERR_ARG, "the dataset is empty");
  HIPCHK(h, learned_train(h->L, h->st, epochs, batch, lr, seed, loss));
  return CS_OK;
}

int cs_learned_get_stats(cs_solver* h, cs_learned_stats* out) {
  if (!h || !out) return CS_ERR_ARG;
  *out = h->lstats;
  h->lstats = cs_learned_stats{};
  return CS_OK;
}

int cs_get_tuning(cs_solver* h, cs_tuning* out) {
  if (!h || !out) return CS_ERR_ARG;
  out->traverse_blocks_per_cu = h->S.tune_trav;
  out->narrow_threads = h->S.tune_narrow;
  out->query_grid = h->S.tune_query;
  out->tuned = h->tuner.tuned;
  return CS_OK;
}

int cs_set_tuning(cs_solver* h, const cs_tuning* in) {
  if (!h || !in) return CS_ERR_ARG;
  auto one_of = [](int v, const int* c, int n) {
    for (int i = 0; i < n; ++i)
      if (c[i] == v) return true;
    return false;
  };
  if (!one_of(in->traverse_blocks_per_cu, kTuneTrav, 4) || !one_of(in->narrow_threads, kTuneNarrow, 3) ||
      !one_of(in->query_grid, kTuneQuery, 4))
    return fail(h, CS_ERR_ARG, "launch shape outside the candidate lists");
  h->S.tune_trav = in->traverse_blocks_per_cu;
  h->S.tune_narrow = in->narrow_threads;
  h->S.tune_query = in->query_grid;
  h->tuner.active = false;
  h->tuner.tuned = 0;
  return CS_OK;
}

const char* cs_last_error(cs_solver* h) { return h ? h->err.c_str() : g_create_error.c_str(); }

const char* cs_version(void) { return "collision_amd 0.1 (gfx950)"; }

}  // extern "C"
