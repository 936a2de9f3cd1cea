// Batched scenes: many independent scenes assembled as one mesh and run by
// the same launches (no counterpart in the reference, which runs one scene
// per solver).  The broad phase never pairs elements of different scenes,
// so each scene's pairs, hits, sums and zones are those of the scene run
// alone.  The pass loop runs while any scene still collides; a pass over a
// scene whose own loop has ended finds that scene unchanged and adds
// nothing (its hits were none in that loop's last pass, and nothing of it
// moved since), except for the rigid update of its impact zones, which is
// not bitwise idempotent: zone updates are therefore limited to the scenes
// still in their own impact-zone loop (SCN_ACTIVE).  So the batch
// reproduces every scene's own pass sequence.
// This file keeps, per scene, the pass counts and flags of that sequence:
// the reference's loop (dcollid.cpp:317-362, 430-468, 227-265) as a state
// machine fed with the scene's hit pairs of every pass.
#include "cs_internal.h"

namespace cs {

__global__ void k_scene_reset(int32_t* __restrict__ scn, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) scn[i] = 0;
}

// hit pairs of the pass per scene (the scene is the upper half of the
// element's surface field)
__global__ void k_scene_hits(const int4* __restrict__ hp, int nhp, const int4* __restrict__ ev,
                             int32_t* __restrict__ pass_hit) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nhp) return;
  atomicAdd(&pass_hit[ev[hp[i].x].w >> 16], 1);
}

// kind: CS_PASS_PROXIMITY 0, CS_PASS_CCD 1, CS_PASS_ZONE 2; index within its loop
__global__ void k_scene_update(int32_t* __restrict__ scn, int nscene, int kind, int index,
                               int max_ccd) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nscene) return;
  int32_t* f = scn + s;
  const int hits = f[SCN_PASS_HIT * nscene];
  f[SCN_PASS_HIT * nscene] = 0;
  const bool hit = hits > 0;
  int phase = f[SCN_PHASE * nscene];
  f[SCN_ACTIVE * nscene] = kind == 0 || (kind == 1 && phase == 1) || (kind == 2 && phase == 2);
  if (kind == 0) {  // detectProximity
    f[SCN_HITPAIRS * nscene] += hits;
    phase = 1;
  } else if (kind == 1) {  // detectCollision: until a pass without a hit, at most max_ccd
    if (phase == 1) {
      f[SCN_CCD * nscene] += 1;
      f[SCN_HITPAIRS * nscene] += hits;
      if (index == 0 && hit) f[SCN_HASCOL * nscene] = 1;
      if (!hit) phase = 3;
      else if (index + 1 >= max_ccd) phase = 2;  // still colliding: impact zones
    }
  } else {  // computeImpactZone: until a pass without a hit
    if (phase == 2) {
      f[SCN_ZONE * nscene] += 1;
      f[SCN_HITPAIRS * nscene] += hits;
      if (!hit) phase = 3;
    }
  }
  f[SCN_PHASE * nscene] = phase;
}

void launch_scene_reset(hipStream_t st, DevState& S) {
  if (S.nscene <= 1) return;
  const int n = SCN_FIELDS * S.nscene;
  hipLaunchKernelGGL(k_scene_reset, dim3((n + 255) / 256), dim3(256), 0, st, S.scn, n);
}

void launch_scene_pass(hipStream_t st, DevState& S, int kind, int index, int nhp, int max_ccd) {
  if (S.nscene <= 1) return;
  if (nhp > 0)
    hipLaunchKernelGGL(k_scene_hits, dim3((nhp + 255) / 256), dim3(256), 0, st, S.hp, nhp, S.ev,
                       S.scn + SCN_PASS_HIT * S.nscene);
  hipLaunchKernelGGL(k_scene_update, dim3((S.nscene + 255) / 256), dim3(256), 0, st, S.scn,
                     S.nscene, kind, index, max_ccd);
}

}  // namespace cs
