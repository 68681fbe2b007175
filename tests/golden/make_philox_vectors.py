"""Writes tests/golden/philox_vectors.json from rocRAND's Philox4x32-10 (philox_rocrand.cpp).

Run from the repository root:  python tests/golden/make_philox_vectors.py
Needs hipcc and the rocRAND headers (host code only; no GPU).  The three published Random123
known-answer vectors are appended with their source noted.
"""

import json
import os
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))

# Random123 kat_vectors, philox4x32 with 10 rounds (Salmon et al., SC'11 distribution)
RANDOM123 = [
    ([0x00000000] * 4, [0x00000000] * 2, [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
    ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
    ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
     [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
]


def main() -> None:
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "philox_rocrand")
        subprocess.run(["hipcc", "-O1", "-o", exe, os.path.join(HERE, "philox_rocrand.cpp")],
                       check=True)
        lines = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.splitlines()
    rocrand = [json.loads(x) for x in lines]
    payload = {
        "generator": "Philox4x32-10",
        "rocrand": rocrand,
        "random123": [{"ctr": c, "key": k, "out": o} for c, k, o in RANDOM123],
    }
    with open(os.path.join(HERE, "philox_vectors.json"), "w") as f:
        json.dump(payload, f, indent=1)
        f.write("\n")


if __name__ == "__main__":
    main()
