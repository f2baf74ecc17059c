"""Writes seedseq_fe128.json: seed-sequence words and first mt19937_64 outputs for the
seeds NEST uses ({rng_seed, 0xb84c9bae, vp}), produced by oracle/_ref/randutils_probe,
which is compiled from the reference's own randutils.hpp (oracle/ref/Makefile)."""

import json
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
PROBE = os.path.join(HERE, "..", "..", "oracle", "_ref", "randutils_probe")

SEEDS = [[143202461, 0xb84c9bae, 0], [143202461, 0xb84c9bae, 3], [12345, 0xb84c9bae, 0], [1, 0xb84c9bae, 7],
         [4294967295, 0xb84c9bae, 1], [143202461, 0xc229212d], [42]]


def main():
    out = []
    for s in SEEDS:
        txt = subprocess.run([PROBE, "8", "5"] + [str(x) for x in s], check=True, capture_output=True,
                             text=True).stdout.split()
        out.append({"seed": s, "words": [int(x) for x in txt[:8]], "mt19937_64": [int(x) for x in txt[8:]]})
    with open(os.path.join(HERE, "seedseq_fe128.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
