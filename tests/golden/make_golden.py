"""Writes the golden fixtures of tests/golden/.

Nothing here is computed by this repository's code: every value is
transcribed from a published source or from the reference's own files.

chacha_kat.json
  * zero_key_keystream_32B: first 32 keystream bytes for key = 0^256,
    IV = 0^64 with 8, 12 and 20 rounds (ChaCha test vectors of
    draft-strombergson-chacha-test-vectors, TC1; the 20-round value is also
    RFC 8439 appendix A.1 test vector #1).
  * rfc8439_2_3_2: the ChaCha20 block function output of RFC 8439 section
    2.3.2 (key 00..1f, nonce 00:00:00:09:00:00:00:4a:00:00:00:00, counter 1).
  * pcg32_demo_42_54: the first six outputs of the PCG reference demo
    (pcg32_srandom_r(&rng, 42u, 54u); pcg32_random_r), the XSH-RR
    generator whose output permutation rand_core's seed_from_u64 uses.

reference_cases.json
  Circuits and asserted values from the reference repository:
  * README.md:71-75 (Bell probabilities, exact);
  * tests/test_mps_simulation.py:96-126 (measure + conditional; reset);
  * tests/test_distribution_e2e.py:86-216 (small circuits with closed-form
    marginals, tolerance 1e-6 as asserted there).
"""
import json
import math
import os

HERE = os.path.dirname(os.path.abspath(__file__))

chacha = {
    "zero_key_keystream_32B": {
        "8": "3e00ef2f895f40d67f5bb8e81f09a5a12c840ec3ce9a7f3b181be188ef711a1e",
        "12": "9bf49a6a0755f953811fce125f2683d50429c3bb49e074147e0089a52eae155f",
        "20": "76b8e0ada0f13d90405d6ae55386bd28bdd219b8a08ded1aa836efcc8b770dc7",
    },
    "rfc8439_2_3_2": ["e4e7f110", "15593bd1", "1fdd0f50", "c47120a3", "c7f4d1c7", "0368c033", "9aaa2204",
                      "4e6cd4c3", "466482d2", "09aa9f07", "05d7c214", "a2028bd9", "d19c12b5", "b94e16de",
                      "e883d0cb", "4e3c50a2"],
    "pcg32_demo_42_54": ["a15c02b7", "7b47f409", "ba1d3330", "83d2f293", "bfa4784b", "cbed606e"],
}


def h(q):
    return {"kind": "h", "qubit": q, "qubits": [q]}


def x(q):
    return {"kind": "x", "qubit": q, "qubits": [q]}


def cx(c, t):
    return {"kind": "cx", "control": c, "target": t, "qubits": [c, t], "params": []}


s3 = math.sin(math.pi / 6) ** 2
c3 = math.cos(math.pi / 6) ** 2
cases = [
    {"name": "readme_bell", "source": "README.md:71-75", "n": 2, "instructions": [h(0), cx(0, 1)],
     "probabilities_exact": {"0": 0.5000000000000001, "3": 0.5000000000000001}},
    {"name": "mps_measure_conditional", "source": "tests/test_mps_simulation.py:96-115", "n": 2, "cbits": 1,
     "instructions": [x(0), {"kind": "measure", "qubit": 0, "cbit": 0, "qubits": [0]},
                      {"kind": "conditional", "condition": {"creg_base": 0, "creg_size": 1, "creg_value": 1},
                       "op": x(1), "qubits": [1]}],
     "classical_bits": {"0": 1}, "probabilities_exact": {"3": 1.0}},
    {"name": "mps_reset", "source": "tests/test_mps_simulation.py:117-126", "n": 1,
     "instructions": [x(0), {"kind": "reset", "qubit": 0, "qubits": [0]}], "probabilities_exact": {"0": 1.0}},
    {"name": "x_cx_deterministic", "source": "tests/test_distribution_e2e.py:88-98", "n": 2,
     "instructions": [x(0), cx(0, 1)], "probabilities_approx": {"3": 1.0}},
    {"name": "bell_pair", "source": "tests/test_distribution_e2e.py:167-170", "n": 2,
     "instructions": [h(0), cx(0, 1)], "probabilities_approx": {"0": 0.5, "3": 0.5}},
    {"name": "cz_pair", "source": "tests/test_distribution_e2e.py:175-179", "n": 2,
     "instructions": [h(0), h(1), {"kind": "cz", "control": 0, "target": 1}],
     "probabilities_approx": {"0": 0.25, "1": 0.25, "2": 0.25, "3": 0.25}},
    {"name": "ry_cx", "source": "tests/test_distribution_e2e.py:180-183", "n": 2,
     "instructions": [{"kind": "ry", "qubit": 0, "theta": math.pi / 3}, cx(0, 1)],
     "probabilities_approx": {"0": c3, "3": s3}},
    {"name": "xyz_single", "source": "tests/test_distribution_e2e.py:189-193", "n": 2,
     "instructions": [x(0), {"kind": "y", "qubit": 1}, {"kind": "z", "qubit": 0}],
     "probabilities_approx": {"3": 1.0}},
    {"name": "ghz4", "source": "tests/test_distribution_e2e.py:200-205", "n": 4,
     "instructions": [h(0), cx(0, 1), cx(1, 2), cx(2, 3)], "probabilities_approx": {"0": 0.5, "15": 0.5}},
    {"name": "chain_cx4", "source": "tests/test_distribution_e2e.py:206-210", "n": 4,
     "instructions": [h(0), cx(0, 2), cx(1, 3)], "probabilities_approx": {"0": 0.5, "5": 0.5}},
    {"name": "local_4h", "source": "tests/test_distribution_e2e.py:211-213", "n": 4,
     "instructions": [h(q) for q in range(4)], "probabilities_approx": {str(j): 1 / 16 for j in range(16)}},
]

if __name__ == "__main__":
    with open(os.path.join(HERE, "chacha_kat.json"), "w") as f:
        json.dump(chacha, f, indent=1)
    with open(os.path.join(HERE, "reference_cases.json"), "w") as f:
        json.dump(cases, f, indent=1)
