"""Write tests/golden/hash_kat.json: frozen values of the biclique hash.

The hash is defined by this project (the reference never emits (L, R) pairs):
key(side, x) = mix64((x << 1) | side), h = mix64(sum_A key(0,a) ^ rotl64(sum_B key(1,b), 31)),
mix64 = splitmix64's output function.  The values below were produced by the
oracle's implementation (oracle/oracle.c orc_hash_pair) and are checked against
the engine's (gmbe_hash_biclique) and a pure-Python restatement.

    python tests/golden/make_hash_kat.py
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.abspath(os.path.join(HERE, "..", "..")))

import oracle  # noqa: E402

CASES = [
    ([0], [0]),
    ([1], [0]),
    ([0], [1]),
    ([0, 1, 2], [5, 7]),
    ([4294967294], [4294967294]),
    (list(range(100)), list(range(50, 80))),
]


def main():
    out = [{"A": A, "B": B, "h": oracle.hash_pair(A, B)} for A, B in CASES]
    with open(os.path.join(HERE, "hash_kat.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
