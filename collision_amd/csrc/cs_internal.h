// Internal declarations shared by the HIP translation units of the library.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

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
  CS_DEVERR_STACK = 2,       // BVH traversal stack overflow
  CS_DEVERR_STRAIN = 4,      // strain-limit schedule inconsistency
};

// slots of the small device counter block
enum {
  CNT_PAIRS = 0,       // emitted narrow-phase pairs
  CNT_HITS = 1,        // hit records
  CNT_HIT_PAIRS = 2,   // pairs with at least one hit
  CNT_HIT_SUB = 3,     // hit sub-tests
  CNT_LOG = 4,         // pair-log rows
  CNT_NARROW = 5,      // pairs whose sub-tests ran
  CNT_ZONES = 6,       // impact zones updated
  CNT_ZMEM = 7,        // members of zones with >= 2 points
  CNT_STRAIN_FLAG = 8, // strain sweep found an edge over the limit
  CNT_STRAIN_EDGES = 9,
  CNT_STRAIN_DEGEN = 10,
  CNT_N = 16
};

struct NarrowParams {
  double dt, thickness, round_eps, k, m, lambda;
  int ccd;        // 0 proximity, 1 CCD
  int zone_mode;  // merge impact zones (s_detImpZone)
  int log_pairs;
  int pass_id;
};

// All device memory of one solver handle.
struct DevState {
  int nv = 0, nt = 0, nb = 0, ne = 0;
  // vertices
  dvec4 *x = nullptr, *x_old = nullptr, *vel = nullptr, *av = nullptr;
  uint8_t *fixed = nullptr, *has = nullptr;
  int32_t* uf = nullptr;  // union-find parent
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
  bvh_node *nodes = nullptr, *leaves = nullptr;
  int32_t *parent = nullptr, *lparent = nullptr, *rmax = nullptr;
  uint32_t* arrive = nullptr;
  // candidate stream
  int2* pairs = nullptr;
  int pair_cap = 0;
  // hits and contribution records
  hit_rec* hits = nullptr;
  int hit_cap = 0;
  uint64_t *hkey = nullptr, *hkey_tmp = nullptr;
  uint32_t *hidx = nullptr, *hidx_tmp = nullptr;
  uint32_t *rkey = nullptr, *rkey_tmp = nullptr, *rval = nullptr, *rval_tmp = nullptr;
  int rec_cap = 0;
  // impact zones
  uint32_t *zkey = nullptr, *zkey_tmp = nullptr, *zval = nullptr, *zval_tmp = nullptr;
  int32_t* zsize = nullptr;
  // strain-limit schedule (level-ordered element lists)
  int32_t *lvl_ptr = nullptr, *lvl_elems = nullptr;
  int n_levels = 0, n_sched = 0;
  // logs
  int4* plog = nullptr;
  int plog_cap = 0;
  // counters / flags
  int* counters = nullptr;
  unsigned long long* cand_count = nullptr;
  int* err = nullptr;
};

// sort
size_t radix_hist_words(int n);
hipError_t radix_sort_u32(hipStream_t st, uint32_t* keys, uint32_t* vals, uint32_t* tmp_keys,
                          uint32_t* tmp_vals, uint32_t* hist, int n, int bits, bool* result_in_tmp);
hipError_t radix_sort_u64(hipStream_t st, uint64_t* keys, uint32_t* vals, uint64_t* tmp_keys,
                          uint32_t* tmp_vals, uint32_t* hist, int n, int bits, bool* result_in_tmp);

// broad phase
void launch_element_boxes(hipStream_t st, const DevState& S, int moving, double dt);
hipError_t build_bvh(hipStream_t st, DevState& S);
hipError_t refit_bvh(hipStream_t st, DevState& S);
void launch_traverse(hipStream_t st, DevState& S, int cap);

// narrow phase
void launch_narrow(hipStream_t st, DevState& S, const NarrowParams& P, int npairs);

// response
hipError_t apply_hits(hipStream_t st, DevState& S, int nhits);
void launch_average_velocity(hipStream_t st, DevState& S, double dt);
void launch_predict(hipStream_t st, DevState& S, double dt);
void launch_uf_init(hipStream_t st, DevState& S);
hipError_t update_zones(hipStream_t st, DevState& S, double m, double dt);
void launch_domain_boundary(hipStream_t st, DevState& S, const double lo[3], const double hi[3],
                            double dt, double mu, double thickness);
void launch_final_position(hipStream_t st, DevState& S, double dt);
hipError_t strain_limit(hipStream_t st, DevState& S, double dt, int max_sweeps, int* sweeps_out);
void launch_final_velocity(hipStream_t st, DevState& S);
void launch_zone_labels(hipStream_t st, DevState& S, int32_t* out);

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
