"""Writes the golden fixtures of tests/golden/ from the reference's own test data.

Run in the build container (it reads the reference checkout at /root/reference); the outputs are
committed, the GPU box never reads the reference:

  example.rsb         the reference's RSB fixture of its example scene (tests/fixtures/example.rsb,
                      8,768 bytes, written by `rt --write-example`): input data, copied as is
  fmath.json          the `fmath` vectors of tests/fixtures/tonemap.json (exp2, log2 and powf of
                      min(x, 1) on 48 inputs, written by `rt --tone-table`, rt-cli/src/tone.rs:103)
  energy_tables.json  size and FNV-1a hash of each energy-compensation table of
                      engine/crates/rt-core/src/bsdf/energy_table.rs (only the hashes are kept:
                      the oracle and the GPU regenerate the tables from the recipe of energy.rs)
"""
import json
import os
import re
import shutil

import numpy as np

REF = "/root/reference"
OUT = os.path.dirname(os.path.abspath(__file__))


def fnv1a_bytes(data: bytes) -> int:
    h = 0x811C9DC5
    for b in data:
        h = ((h ^ b) * 0x01000193) & 0xFFFFFFFF
    return h


def main():
    shutil.copyfile(os.path.join(REF, "tests/fixtures/example.rsb"), os.path.join(OUT, "example.rsb"))
    tone = json.load(open(os.path.join(REF, "tests/fixtures/tonemap.json")))
    with open(os.path.join(OUT, "fmath.json"), "w") as f:
        json.dump(tone["fmath"], f)
    src = open(os.path.join(REF, "engine/crates/rt-core/src/bsdf/energy_table.rs")).read()
    tables = {}
    for name, dt in [("ALBEDO", "<u2"), ("SCHLICK", "<u2"), ("CLEAR_IN", "<u2"), ("CLEAR_OUT", "<u2"),
                     ("ALBEDO_AVG", "<f4"), ("SCHLICK_AVG", "<f4")]:
        m = re.search(r"pub static %s: \[\w+; (\d+)\] = \[(.*?)\];" % name, src, re.S)
        vals = [v.strip() for v in m.group(2).replace("\n", " ").split(",") if v.strip()]
        arr = np.array([float(v) for v in vals], dtype=np.float64).astype(dt)
        assert arr.size == int(m.group(1))
        tables[name] = {"len": int(arr.size), "fnv1a": fnv1a_bytes(arr.tobytes())}
    with open(os.path.join(OUT, "energy_tables.json"), "w") as f:
        json.dump(tables, f, indent=1)


if __name__ == "__main__":
    main()
