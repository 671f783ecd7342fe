"""Writes reference_primes.json: the NTT-friendly prime tables p64 / sp64
that FIDESlib's own tests use (test/ParametrizedTest.cuh:121-181).

Run once in the build container (the reference tree is not available where
the tests run); the JSON is the committed fixture.
"""
import json
import os
import re

SRC = "/root/reference/test/ParametrizedTest.cuh"


def main():
    src = open(SRC).read()

    def grab(name):
        m = re.search(r"inline std::vector<FIDESlib::PrimeRecord> " + name + r"\{(.*?)\};", src, re.S)
        return [int(x) for x in re.findall(r"\.p = (\d+)", m.group(1))]

    d = {"source": "FIDESlib test/ParametrizedTest.cuh:121-181 (prime tables p64, sp64)",
         "p64": grab("p64"), "sp64": grab("sp64")}
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_primes.json")
    json.dump(d, open(out, "w"), indent=1)


if __name__ == "__main__":
    main()
