"""Writes tests/golden/scene_hashes.json: SHA-256 of the generated arrays of
a few workloads, so any change in the generators (or a host whose basic FP
arithmetic differs) is caught."""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from collision_amd import scenes as S  # noqa: E402
from tests.test_scenes import digest  # noqa: E402

out = {"config0": digest(S.config0()), "config1_n64": digest(S.config1(n=64)),
       "config4_n96": digest(S.combined_scene(n=96, nstrings=60)),
       "config2_16": digest(S.batched_scenes(count=16))}
json.dump(out, open(os.path.join(HERE, "scene_hashes.json"), "w"), indent=1)
print(out)
