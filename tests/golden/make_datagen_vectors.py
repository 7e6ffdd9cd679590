"""Writes tests/golden/datagen_vectors.json: the first 16 (u, y) values of
the seeded mass-spring-damper generator for seeds 0 and 1, plus the raw
SplitMix64 outputs they derive from. The generator is this project's own
(the reference's data files are unseeded), so the fixture pins it against
regressions; the SplitMix64 words are checked independently in the test."""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from gprat_amd import datagen  # noqa: E402

out = {}
for seed in (0, 1):
    u, y = datagen.msd_series(16, seed)
    words = [int(w) for w in datagen.splitmix64(seed, 4)]
    out[str(seed)] = {"u": u.tolist(), "y": y.tolist(), "splitmix64_first4": words}
with open(os.path.join(HERE, "datagen_vectors.json"), "w") as f:
    json.dump(out, f, indent=1)
print("wrote", os.path.join(HERE, "datagen_vectors.json"))
