"""Writes the RNG golden vectors, produced by oracle/_ref/randutils_probe, which is compiled
from the reference's own randutils.hpp (oracle/ref/Makefile):

seedseq_fe128.json: seed-sequence words and first mt19937_64 outputs for the seeds NEST uses
({rng_seed, 0xb84c9bae, vp}, random_manager.cpp:63-100) and two others.
mt19937_64_streams.json: the first 400 outputs of the per-VP engines of a few seeds, from
which the connection tests derive NEST's source, delay and weight draws by libstdc++'s
published distribution algorithms.
"""

import json
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
PROBE = os.path.join(HERE, "..", "..", "oracle", "_ref", "randutils_probe")

SEEDS = [[143202461, 0xb84c9bae, 0], [143202461, 0xb84c9bae, 3], [12345, 0xb84c9bae, 0], [1, 0xb84c9bae, 7],
         [4294967295, 0xb84c9bae, 1], [143202461, 0xc229212d], [42]]
STREAMS = [[143202461, 0xb84c9bae, 0], [143202461, 0xb84c9bae, 1], [2024, 0xb84c9bae, 0]]


def probe(seed, n_words, n_draws):
    txt = subprocess.run([PROBE, str(n_words), str(n_draws)] + [str(x) for x in seed], check=True,
                         capture_output=True, text=True).stdout.split()
    return [int(x) for x in txt[:n_words]], [int(x) for x in txt[n_words:]]


def main():
    out = []
    for s in SEEDS:
        words, draws = probe(s, 8, 5)
        out.append({"seed": s, "words": words, "mt19937_64": draws})
    with open(os.path.join(HERE, "seedseq_fe128.json"), "w") as f:
        json.dump(out, f, indent=1)
    streams = []
    for s in STREAMS:
        _, draws = probe(s, 1, 400)
        streams.append({"seed": s, "mt19937_64": draws})
    with open(os.path.join(HERE, "mt19937_64_streams.json"), "w") as f:
        json.dump(streams, f, indent=0)


if __name__ == "__main__":
    main()
