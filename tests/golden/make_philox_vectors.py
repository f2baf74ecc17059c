"""Writes the Philox_32 golden vectors:

philox4x32_kat.json: the philox4x32 known-answer vectors (7 and 10 rounds) of the Random123
distribution in the reference tree (thirdparty/Random123/tests/kat_vectors), as data.
philox_engine_draws.json: draws of r123::Engine<r123::Philox4x32> seeded as NEST seeds its
per-VP streams for rng_type "Philox_32" ({rng_seed, 0xb84c9bae, vp}, random_manager.cpp:95),
produced by oracle/_ref/philox_probe, which is compiled from the reference's own Random123 and
randutils headers (oracle/ref/Makefile): engine words, uniform_real(0, 1), uniform_int over
several ranges (libstdc++'s 32-bit Lemire path, rejections included) and normal(0, 1).
"""

import json
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
PROBE = os.path.join(HERE, "..", "..", "oracle", "_ref", "philox_probe")
KAT = "/root/reference/thirdparty/Random123/tests/kat_vectors"

SEEDS = [[20241, 0xb84c9bae, 0], [20241, 0xb84c9bae, 5], [143202461, 0xb84c9bae, 3], [7, 0xb84c9bae, 1]]
# (family, range): 0 engine words, 1 uniform_real, 2 uniform_int(range), 3 normal
FAMILIES = [(0, 0), (1, 0), (2, 77169), (2, 3), (2, 2147483649), (3, 0)]


def probe(family, n, rng, seed):
    txt = subprocess.run([PROBE, str(family), str(n), str(rng)] + [str(x) for x in seed], check=True,
                         capture_output=True, text=True).stdout.split()
    return [float(x) if family in (1, 3) else int(x) for x in txt]


def main():
    kat = []
    with open(KAT) as f:
        for line in f:
            if line.startswith("philox4x32 "):
                w = line.split()
                kat.append({"rounds": int(w[1]), "ctr": [int(x, 16) for x in w[2:6]],
                            "key": [int(x, 16) for x in w[6:8]], "out": [int(x, 16) for x in w[8:12]]})
    with open(os.path.join(HERE, "philox4x32_kat.json"), "w") as f:
        json.dump(kat, f, indent=1)
    draws = []
    for s in SEEDS:
        for fam, rng in FAMILIES:
            draws.append({"seed": s, "family": fam, "range": rng, "values": probe(fam, 200, rng, s)})
    with open(os.path.join(HERE, "philox_engine_draws.json"), "w") as f:
        json.dump(draws, f, indent=0)


if __name__ == "__main__":
    main()
