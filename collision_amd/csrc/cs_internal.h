// Internal declarations shared by the HIP translation units of the library.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "cs_device.h"

#define CS_HIP_RET(expr)                    \
  do {                                      \
    hipError_t _e = (expr);                 \
    if (_e != hipSuccess) return _e;        \
  } while (0)

namespace cs {

// element flag bits
enum : uint8_t { EF_TRI = 1, EF_RIGID = 2 };

// device error bits (OR-ed into DevState::err)
enum : int {
  CS_DEVERR_NONFINITE = 1,   // where the reference calls clean_up(ERROR)
  CS_DEVERR_STRAIN = 4,      // strain-limit schedule inconsistency
};

// slots of the small device counter block (ints; 64-bit counts at 16..21)
enum {
  CNT_PAIRS = 0,         // pairs in the candidate stream being built
  CNT_HITS = 1,          // hit records
  CNT_HIT_PAIRS = 2,     // pairs with at least one hit (hit-pair records)
  CNT_HIT_SUB = 3,       // hit sub-tests
  CNT_LOG = 4,           // pair-log rows
  CNT_TMP = 5,           // scratch count (LBVH top nodes)
  CNT_ZONES = 6,         // impact zones updated
  CNT_QPAIRS = 7,        // pairs found by the dirty-element query
  CNT_STRAIN_FLAG = 8,   // strain sweep state
  CNT_STRAIN_EDGES = 9,
  CNT_STRAIN_DEGEN = 10,
  CNT_STRAIN_SWEEPS = 11,
  CNT_DIRTY = 12,        // dirty elements
  CNT_STRAIN_LEVELS = 13,  // levels evaluated by the strain sweeps of the step
  CNT_STRAIN_EVALS = 14,   // slots evaluated by the strain sweeps of the step
  CNT_ERR = 15,
  CNT_CAND64 = 16,       // box-overlap candidates of a full traversal
  CNT_OLD64 = 18,        // candidates touching dirty elements, previous boxes
  CNT_NEW64 = 20,        // candidates touching dirty elements, current boxes
  CNT_ZLIST = 22,        // vertices in the zone member list (kept for the whole step)
  CNT_DRANGES = 23,      // tree ranges holding dirty leaves (kept until the next pass clears them)
  CNT_VLIST = 24,        // 24, 25: vertices whose average velocity a pass changed through
                        // hits (dirty vertex list), by pass parity
  CNT_SNAP = 28,         // 28..31: pair/hit counts carried into the current pass
  CNT_PCAND64 = 32,      // 32, 33: static-box overlaps found by the swept-box traversal
  CNT_PPAIRS = 34,       // pairs of that traversal whose static boxes overlap
  CNT_STRAIN_PROG = 35,  // strain sweep: level iteration reached (-1: done), read by its prefetcher
  CNT_INTS = 40
};

struct NarrowParams {
  double dt, thickness, round_eps, k, m, lambda, cr;
  const int32_t* vbody;  // body of a movable vertex, -1 otherwise (nullptr: no bodies)
  const double* bmass;   // total_mass of each body
  int ccd;        // 0 proximity, 1 CCD
  int mixed;      // FP32 screen of proximity sub-tests
  int mixed_stats;
  int zone_mode;  // merge impact zones (s_detImpZone)
  int log_pairs;
  int pass_id;
  int* spread;  // [SPREAD_SLOTS][SPREAD_STRIDE] statistics (SP_*)
  // learned form (CCD passes): the scoring kernel's per-sub-test seeds
  // (16 per pair, by pair position before reordering; sign bit set = root
  // predicted) and the reordering of the pairs (nullptr: none)
  const float* lseed;
  const uint32_t* perm;
  // dataset collection (ds_cap == 0: off): keep a CCD sub-test when a hash
  // of (pair, sub-test, salt) is divisible by ds_stride
  float* dsx;
  float* dse;
  float* dst;
  unsigned long long* ds_count;
  long long ds_cap;
  int ds_stride;
  uint32_t ds_salt;
};

// Statistics of the mixed-precision screens and the learned form, summed by
// many wavefronts: each work-group adds into one of SPREAD_SLOTS copies (its
// block index modulo the count, one 128-byte line each), so that no single
// address takes an atomic from every wavefront; the host adds the copies
// (sums) or takes their maximum (float bits of non-negative values).
constexpr int SPREAD_SLOTS = 64, SPREAD_STRIDE = 32;
enum {
  SP_MX_SCREENED = 0,  // proximity sub-tests the FP32 screen rejected
  SP_MX_RECHECK,       // sub-tests it passed on to FP64
  SP_MX_EDIST,         // max FP32 distance error / thickness (float bits)
  SP_MX_EPARAM,        // max FP32 barycentric / segment parameter error (float bits)
  SP_MX_RATIO,         // max FP32 error / its certified bound (float bits)
  SP_LN_SCORED,        // learned form: CCD sub-tests scored by the network
  SP_LN_EXISTS,        // of those, isCoplanar true (exact cubic)
  SP_LN_PRED,          // predicted to have a root (logit > 0)
  SP_LN_AGREE,         // prediction == exact outcome
  SP_LN_ROOTS,         // roots located by the seeded Newton solve
  SP_LN_ITERS,         // its iterations, seeded at the network's time of impact
  SP_LN_ITERS_MID,     // iterations of the same solve seeded at dt / 2
  SP_LN_EMAX,          // max |Newton root - closed-form root| / dt (float bits)
  SP_COUNT
};

// One slot of the strain-limit schedule (one non-rigid element), 256 B.
// Rest lengths, vertex ids and links are written at assembly; x_old of the
// vertices, the x_old edge lengths and the velocity copies are packed once
// per step.  Each slot keeps its own copy of its vertices' velocities: a
// visit that changes a vertex (or passes a change on) stores the new value
// into the copy held by the vertex's next visit, so an evaluated slot reads
// everything it needs from its own record.
struct alignas(16) StrainSlot {
  double x[9];     // x_old of vertices 0, 1, 2 (vertex 2 unused for a bond)
  double lold[3];  // x_old lengths of edges (0,1), (1,2), (2,0)
  double L0[3];    // rest lengths of the same edges
  int v[3];        // vertex ids; v[2] = -1 for a bond
  int ns[3];       // 4 * slot + position of the vertex's next visit; the last
                   // visit links to the first (a single visit to itself); -1: none
  int nl[3];       // that visit's level << 10 | its index within the level
  int level;
  dvec4 vel[3];    // velocity copies; w holds the bits of the slot that last wrote the vertex
};
static_assert(sizeof(StrainSlot) == 256, "StrainSlot layout");
constexpr int STRAIN_LEVEL_CAP = 1024;  // most slots in one level
constexpr int TRAV_QCAP = 65536;        // split walks per round of the full traversal
constexpr size_t TRAV_ITEM_BYTES = 48;

// All device memory of one solver handle.
struct DevState {
  int nv = 0, nt = 0, nb = 0, ne = 0;
  // vertices
  dvec4 *x = nullptr, *x_old = nullptr, *vel = nullptr, *av = nullptr;
  uint8_t *fixed = nullptr, *has = nullptr;
  int32_t* uf = nullptr;  // union-find parent
  // movable rigid bodies (nbody == 0: none, pointers null).  Every point of a
  // body holds the same collsnImpulse_RG / collsn_num_RG (an impulse spreads
  // over the whole body, dcollid.cpp:1101-1115), so they are kept per body.
  int nbody = 0;
  int32_t* vbody = nullptr;     // per vertex: body id, -1 if not movable
  int32_t* bptr = nullptr;      // [nbody + 1] offsets of each body's vertices in bmem
  uint32_t* bmem = nullptr;     // body vertices, ascending id within a body
  int32_t* bfirst = nullptr;    // [nbody + 1] offsets in bvisit
  int32_t* bvisit = nullptr;    // body vertices in order of first appearance in element order
  double* bmass = nullptr;      // total_mass
  dvec4* bcom = nullptr;        // center_of_mass
  dvec4* bvel = nullptr;        // center_of_mass_velo
  dvec4* bmrg = nullptr;        // the solver's mrg_com copy (collid.h:174)
  dvec4* rg_imp = nullptr;      // collsnImpulse_RG (never cleared by the solver)
  int32_t* rg_num = nullptr;    // collsn_num_RG
  // impulses of the last pass, kept when record_impulses is set
  dvec4* pimp = nullptr;
  int32_t* pnum = nullptr;
  // elements
  int4* ev = nullptr;     // (v0, v1, v2 or -1, surface id)
  uint8_t* eflags = nullptr;
  dvec4* len0 = nullptr;  // rest lengths of edges (j, j+1)
  dbox* ebox = nullptr;
  // LBVH
  float* bounds = nullptr;
  uint32_t *mcode = nullptr, *mid = nullptr, *mcode_tmp = nullptr, *mid_tmp = nullptr;
  uint32_t *sorted_codes = nullptr, *sorted_ids = nullptr;
  uint32_t* sort_hist = nullptr;
  size_t sort_hist_words = 0;
  bvh_node* nodes = nullptr;
  // nodes: 2n-1 records in depth-first pre-order (root 0; an internal
  // node's first child is the next record)
  int32_t* parent = nullptr;       // parent of every node, -1 at the root
  int32_t* nright = nullptr;       // second child of an internal node (-1 at leaves)
  int32_t *rmin = nullptr, *rmax = nullptr;  // Morton leaf range of every node
  int32_t* lpre = nullptr;         // node of Morton leaf j
  // construction arrays (radix-tree order) of the last build
  int32_t *kleft = nullptr, *kright = nullptr, *kparent = nullptr;
  int32_t *krmin = nullptr, *krmax = nullptr, *pre = nullptr;
  int32_t* split_right = nullptr;  // [g]: second child of the node splitting after leaf g
  uint32_t* arrive = nullptr;
  int32_t* top_start = nullptr;  // nodes spanning several refit tiles whose children do not
  int n_top_start = 0;           // (built once per LBVH build)
  uint32_t* scan_part = nullptr; // partial sums of the device scan
  size_t scan_part_words = 0;
  int32_t* leaf_of = nullptr;    // element -> pre-order index of its leaf
  // dirty leaves by pre-order index, and the QW_RANGE-node ranges holding them
  uint8_t* dpre = nullptr;
  int32_t *rflag = nullptr, *rlist = nullptr;
  int32_t* tnext = nullptr;        // persistent full traversal: next leaf per XCD share,
                                   // then the item counts of the two split rounds
  void* tq = nullptr;              // split walks of the full traversal, 2 x TRAV_QCAP items
  int num_cus = 256;               // compute units of the device
  // launch shapes (cs_set_tuning / the autotuner): persistent traversal
  // work-groups per CU, narrow-phase work-group size, dirty-query grid cap
  int tune_trav = 8, tune_narrow = 128, tune_query = 4096;
  // candidate stream (current and previous pass)
  int2* pairs = nullptr;
  int2* pairs_prev = nullptr;
  int pair_cap = 0;
  int2* qpairs = nullptr;        // pairs found by the dirty-element query
  int qpair_cap = 0;
  // hits (current and previous pass) and hit-pair records (lo, hi, mask, first sub-test)
  hit_rec* hits = nullptr;
  hit_rec* hits_prev = nullptr;
  int hit_cap = 0;
  int4* hp = nullptr;
  int4* hp_prev = nullptr;
  int hp_cap = 0;
  // vertices / elements whose average velocity changed since the last narrow phase
  uint8_t* dirty_e = nullptr;   // padded to whole 32-bit words (word atomics mark it)
  int32_t* dlist = nullptr;
  dbox* pbox = nullptr;         // static (proximity) boxes, beside the swept ones in ebox
  int2* pairs_seed = nullptr;   // swept-box candidate pairs of the proximity pass
  int32_t* vlist = nullptr;     // vertices the last apply changed, each once
  int32_t* vadj_ptr = nullptr;  // vertex -> elements holding it (CSR over all elements)
  int32_t* vadj = nullptr;
  uint64_t *hkey = nullptr, *hkey_tmp = nullptr;
  uint32_t *hidx = nullptr, *hidx_tmp = nullptr;
  uint32_t *rkey = nullptr, *rkey_tmp = nullptr, *rval = nullptr, *rval_tmp = nullptr;
  int rec_cap = 0;
  // impact zones
  // zone members: vertices that took part in a union this step (zmark set,
  // listed once in zlist); sorted by (root, vertex) for the update
  int32_t* zmark = nullptr;
  uint32_t* zlist = nullptr;
  uint64_t *zkey64 = nullptr, *zkey64_tmp = nullptr;
  uint32_t *zkey = nullptr, *zval = nullptr, *zval_tmp = nullptr;
  int32_t* zhead = nullptr;  // first sorted index of each zone
  int32_t* zend = nullptr;   // one past the last sorted index, by zone root
  // strain-limit schedule: slots in level order, slots of level L are
  // [lvl_ptr[L], lvl_ptr[L+1])
  int32_t* lvl_ptr = nullptr;
  StrainSlot* sslot = nullptr;
  uint32_t* spre = nullptr;        // per slot: byte 3 pre-check active, bytes 0-2 far marks
  int32_t* slast = nullptr;        // per vertex: 4 * slot + position of its last visit (-1: none)
  uint8_t* lvl_active = nullptr;   // level holds an active element (pre-check)
  int n_levels = 0, n_sched = 0;
  int strain_delta = 1;            // 1 + widest level span of a vertex's visits
  // logs
  int4* plog = nullptr;
  int plog_cap = 0;
  // counters / flags
  int* counters = nullptr;
  int* spread = nullptr;  // [SPREAD_SLOTS][SPREAD_STRIDE] statistics (SP_*)
  // batched scenes (nscene > 1): per-scene pass state, SCN_FIELDS x nscene
  int nscene = 0;
  int32_t* scn = nullptr;
  int32_t* vscene = nullptr;  // scene of every vertex (-1: none)
};

// per-scene pass state of a batched resolve (rows of DevState::scn)
enum {
  SCN_PHASE = 0,   // 0 before the CCD loop, 1 in it, 2 in the impact-zone loop, 3 done
  SCN_CCD,         // CCD passes the scene runs
  SCN_ZONE,        // impact-zone passes it runs
  SCN_HASCOL,      // hit in its first CCD pass
  SCN_HITPAIRS,    // hit pairs over the passes it runs
  SCN_PASS_HIT,    // this pass: hit pairs of the scene (reset after each pass)
  SCN_ACTIVE,      // this pass belongs to the scene's own pass loop
  SCN_FIELDS
};
void launch_scene_reset(hipStream_t st, DevState& S);
void launch_fill_u32(hipStream_t st, void* p, int64_t words, uint32_t v);
// after a pass's narrow phase: scene of each of its nhp hit pairs, then the
// per-scene state machine of the reference's pass loop
void launch_scene_pass(hipStream_t st, DevState& S, int kind, int index, int nhp, int max_ccd);

// The time-of-impact proposer of a handle: FP32 master weights, Adam state,
// BF16 matrix-core copies, the scoring and sorting buffers of the learned
// form and the collected dataset.
struct LearnedState {
  float *w = nullptr, *m = nullptr, *v = nullptr, *grad = nullptr;
  uint16_t* wb = nullptr;
  bool ready = false;  // weights set
  int64_t adam_t = 0;
  // learned form: per-slot seed codes [16 * cap], the list of slots to
  // score and its length, pair sort keys / positions and their partners
  float* seeds = nullptr;
  uint32_t* list = nullptr;
  int* count = nullptr;
  uint32_t *keys = nullptr, *idx = nullptr, *tkeys = nullptr, *tidx = nullptr;
  uint32_t *hist = nullptr, *part = nullptr;
  int score_cap = 0;
  // dataset: x [cap][24], exists [cap], toi [cap]; ds_count counts every
  // selected sub-test (those beyond ds_cap are dropped)
  float *dsx = nullptr, *dse = nullptr, *dst = nullptr;
  unsigned long long* ds_count = nullptr;
  int64_t ds_alloc = 0, ds_cap = 0;
  int32_t ds_stride = 1;
  uint32_t ds_salt = 0;
  // training partials
  float *gpart = nullptr, *lpart = nullptr;
  int train_tiles = 0;
};

hipError_t learned_alloc(LearnedState& L);
void learned_free(LearnedState& L);
hipError_t learned_set_weights(LearnedState& L, hipStream_t st, const float* host_w);
hipError_t learned_get_weights(LearnedState& L, hipStream_t st, float* host_w);
void learned_init_host(uint64_t seed, float* w);
hipError_t learned_infer(LearnedState& L, hipStream_t st, const float* dx, int64_t n, float* dout);
hipError_t learned_grad(LearnedState& L, hipStream_t st, const float* dx, const float* de,
                        const float* dtoi, const uint32_t* idx, int n, float* dloss);
hipError_t learned_adam_step(LearnedState& L, hipStream_t st, float lr);
hipError_t learned_ensure_dataset(LearnedState& L, int64_t capacity);
int64_t learned_dataset_size(LearnedState& L, hipStream_t st);
hipError_t learned_train(LearnedState& L, hipStream_t st, int epochs, int batch, float lr,
                         uint64_t seed, float* host_loss);

// sort and scan
size_t radix_hist_words(int n);
size_t scan_partial_words(int n);
hipError_t exclusive_scan_u32(hipStream_t st, uint32_t* data, int n, uint32_t* partials);
hipError_t radix_sort_u32(hipStream_t st, uint32_t* keys, uint32_t* vals, uint32_t* tmp_keys,
                          uint32_t* tmp_vals, uint32_t* hist, uint32_t* partials, int n, int bits,
                          bool* result_in_tmp);
hipError_t radix_sort_u64(hipStream_t st, uint64_t* keys, uint32_t* vals, uint64_t* tmp_keys,
                          uint32_t* tmp_vals, uint32_t* hist, uint32_t* partials, int n, int bits,
                          bool* result_in_tmp);

// broad phase
void launch_element_boxes(hipStream_t st, const DevState& S, int moving, double dt, dbox* out);
hipError_t build_bvh(hipStream_t st, DevState& S);  // also finds the refit top nodes
hipError_t refit_bvh(hipStream_t st, DevState& S);
// candidate pairs of the current boxes (ebox) into out; with pbox (static
// boxes, each inside its element's swept box in ebox), also the static-box
// overlaps among the rejected pairs (CNT_PCAND64)
void launch_traverse(hipStream_t st, DevState& S, int2* out, int cap, const dbox* pbox);
// the pairs of S.pairs_seed[0..n) whose static boxes overlap into S.pairs
// (count CNT_PPAIRS, also added to CNT_PCAND64)
void launch_split_seed(hipStream_t st, DevState& S, int n);
// incremental passes
// elements of the vertices the previous pass changed: *vcount hit vertices
// in vlist and nzone zone members in zlist
void launch_mark_dirty(hipStream_t st, DevState& S, const int* vcount, int nzone);
// clears the previous pass's dirty-range marks; its range count is zeroed by
// the caller afterwards (launch_zero_counters)
void launch_clear_ranges(hipStream_t st, DevState& S);
void launch_zero_counters(hipStream_t st, DevState& S, uint32_t mask);
constexpr int QW_RANGE_SHIFT = 3;
constexpr int QW_RANGE = 1 << QW_RANGE_SHIFT;  // pre-order nodes per dirty-query range
void launch_query_dirty(hipStream_t st, DevState& S, int emit);
// the dirty elements' new swept boxes, their leaves exact and ancestors grown
void launch_grow_dirty(hipStream_t st, DevState& S, double dt);
void launch_carry(hipStream_t st, DevState& S, int npairs_prev, int nhits_prev, int nhp_prev);
// pairs of src[0..n) with no dirty element into S.pairs (count CNT_PAIRS)
void launch_carry_pairs_from(hipStream_t st, DevState& S, const int2* src, int n);

// narrow phase
void launch_narrow(hipStream_t st, DevState& S, const NarrowParams& P, const int2* pairs,
                   int npairs, const int* npairs_dev, int cap);
// learned form: score a pair stream's CCD sub-tests and sort the pairs by
// decreasing score (*perm: positions in test order; the narrow phase reads
// L.seeds through it)
hipError_t launch_learned_score(hipStream_t st, DevState& S, LearnedState& L, const int2* pairs,
                                int npairs, const int* npairs_dev, int cap, double dt, double h,
                                const uint32_t** perm);
void launch_hp_unions(hipStream_t st, DevState& S, int n, int ccd);
void launch_hp_log(hipStream_t st, DevState& S, int n, int pass_id);
// y = f(x) with the step's elementary functions (0 acos, 1 cos, 2 sin, 3 pow(x, 1/3))
hipError_t launch_math_eval(hipStream_t st, int fn, const double* x, double* y, int64_t n);

// response
hipError_t apply_hits(hipStream_t st, DevState& S, int nhits, int* vcount, bool record);
// movable bodies, after apply_hits: the RG part of updateAverageVelocity
// (dcollid.cpp:728-734) and, when dt > 0, updateImpactZoneVelocityForRG
// (dcollid.cpp:267-288): every body zone moved rigidly
hipError_t apply_bodies(hipStream_t st, DevState& S, double m, double dt);
// updateImpactZoneVelocity's pass over the body zones (they hold two or more points)
hipError_t update_body_zones(hipStream_t st, DevState& S, double m, double dt);
// updateFinalForRG (dcollid.cpp:626-675)
void launch_final_bodies(hipStream_t st, DevState& S, double dt);
void launch_record_origin(hipStream_t st, DevState& S);
void launch_average_velocity(hipStream_t st, DevState& S, double dt);
void launch_predict(hipStream_t st, DevState& S, double dt);
void launch_uf_init(hipStream_t st, DevState& S);
hipError_t update_zones(hipStream_t st, DevState& S, double m, double dt, int nlist);
void launch_domain_boundary(hipStream_t st, DevState& S, const double lo[3], const double hi[3],
                            double dt, double mu, double thickness);
void launch_final_position(hipStream_t st, DevState& S, double dt);
// marks != nullptr: a pair of events around every k_strain_pipe launch is
// appended (the caller reads and destroys them)
hipError_t strain_limit(hipStream_t st, DevState& S, double dt, int max_sweeps,
                        std::vector<hipEvent_t>* marks);
// cs_schedule.hip (host)
void build_strain_schedule(int nv, int ne, const int4* ev, const uint8_t* ef,
                           std::vector<int>& level, int& n_levels, int& delta);
void launch_final_velocity(hipStream_t st, DevState& S);
void launch_zone_labels(hipStream_t st, DevState& S, int32_t* out);

// Work-group aggregated append: every lane of the block calls it; lanes with
// want=true get consecutive slots from one global atomic per block.
// s_cnt / s_base are __shared__ ints of the caller.
// Work-groups are dealt round-robin to the 8 XCDs, each with its own L2.
// Remapping block b to a contiguous share per XCD gives every XCD one
// contiguous range of the work (neighbouring leaves / pairs), so its L2
// sees one region of the tree instead of a scattered 1/8 of all of it.
// A bijection of [0, G).
constexpr int kNumXcd = 8;
__device__ inline int xcd_block(int b, int G) {
  const int q = G / kNumXcd, r = G % kNumXcd;
  const int x = b % kNumXcd, i = b / kNumXcd;
  return (x < r) ? x * (q + 1) + i : r * (q + 1) + (x - r) * q + i;
}

// slot in a list for the lanes with want set, one atomic per wavefront;
// every active lane of the wavefront calls it
__device__ inline int wave_append(bool want, int* counter) {
  const uint64_t m = __ballot(want);
  if (m == 0) return -1;
  const int lane = threadIdx.x & 63;
  const int leader = __ffsll((unsigned long long)m) - 1;
  int base = 0;
  if (lane == leader) base = atomicAdd(counter, __popcll(m));
  base = __shfl(base, leader);
  return want ? base + __popcll(m & ((1ull << lane) - 1ull)) : -1;
}

__device__ inline int block_append(bool want, int* counter, int* s_cnt, int* s_base) {
  if (threadIdx.x == 0) *s_cnt = 0;
  __syncthreads();
  int slot = want ? atomicAdd(s_cnt, 1) : -1;
  __syncthreads();
  if (threadIdx.x == 0) *s_base = *s_cnt ? atomicAdd(counter, *s_cnt) : 0;
  __syncthreads();
  return want ? *s_base + slot : -1;
}

__device__ inline unsigned long long block_sum_u64(unsigned long long v, unsigned long long* s_w) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
  if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = v;
  __syncthreads();
  unsigned long long t = 0;
  if (threadIdx.x == 0)
    for (int w = 0; w < (int)(blockDim.x + 63) / 64; ++w) t += s_w[w];
  __syncthreads();
  return t;  // valid in thread 0
}

// float atomics via CAS (monotone min/max on finite values)
__device__ inline void atomicMin_float(float* addr, float v) {
  int* ai = (int*)addr;
  int old = *ai;
  while (__int_as_float(old) > v) {
    int prev = atomicCAS(ai, old, __float_as_int(v));
    if (prev == old) break;
    old = prev;
  }
}
__device__ inline void atomicMax_float(float* addr, float v) {
  int* ai = (int*)addr;
  int old = *ai;
  while (__int_as_float(old) < v) {
    int prev = atomicCAS(ai, old, __float_as_int(v));
    if (prev == old) break;
    old = prev;
  }
}

}  // namespace cs
