"""Writes tests/golden/rng_prefix.json: the first draws of the shared engine after seed 0xBAADF00D
(std::mt19937_64 with libstdc++'s normal_distribution<double> and uniform_int_distribution<size_t>(0, 9)),
as produced by the oracle build. Run from the repository root: python tests/golden/make_rng_prefix.py"""
import json
import os
import sys

sys.path.insert(0, os.getcwd())
from oracle import oracle as O  # noqa: E402

SEED = 0xBAADF00D
O.seed(SEED)
normal = [float(v) for v in O.normal_draws(256)]
uniform = [int(v) for v in O.uniform_draws(256, 9)]
out = {"seed": SEED, "engine": "std::mt19937_64", "normal": normal, "uniform9": uniform}
json.dump(out, open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "rng_prefix.json"), "w"))
print("wrote", len(normal), "normal and", len(uniform), "uniform draws")
