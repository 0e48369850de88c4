"""Writes tests/golden/pcg_vectors.json: inputs and outputs of the 32-bit
PCG-RXS-M-XS hash ("pcg_hash", Jarzynski & Olano, JCGT 9(3) 2020) as used by
Core/include/Random.h:59-70, computed here with Python integers, independent of
the C oracle and of the kernels.  Run from the repo root:
    python tests/golden/make_pcg_vectors.py
"""
import json
import os

M32 = 0xFFFFFFFF


def pcg_hash(v: int) -> int:
    state = (v * 747796405 + 2891336453) & M32
    word = (((state >> ((state >> 28) + 4)) ^ state) * 277803737) & M32
    return ((word >> 22) ^ word) & M32


def main():
    inputs = [0, 1, 2, 3, 7, 255, 256, 65535, 65536, 747796405, 2891336453, 0x7FFFFFFF, 0x80000000,
              0xFFFFFFFE, 0xFFFFFFFF]
    x = 12345
    for _ in range(200):  # a fixed LCG walk over u32
        x = (x * 1664525 + 1013904223) & M32
        inputs.append(x)
    chains = []
    for seed in (0, 1, 1920 * 1080 - 1, 123456789):
        s, seq = seed, []
        for _ in range(8):  # PcgFloat advances the state by the hash itself
            s = pcg_hash(s)
            seq.append(s)
        chains.append({"seed": seed, "states": seq})
    out = {"inputs": inputs, "outputs": [pcg_hash(v) for v in inputs], "chains": chains}
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "pcg_vectors.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print(f"wrote {path}")


if __name__ == "__main__":
    main()
