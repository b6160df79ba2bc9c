"""Writes the golden fixtures in this directory from the compiled reference
(oracle/_ref/libphjref.so, built from the reference's own sources by
oracle/Makefile). Inputs and expected outputs only; no reference source.

    python tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import pyoracle as O  # noqa: E402


def main():
    assert O.ref_available(), "build oracle/_ref first (make -C oracle ref)"
    # intHash64 known answers from the reference's AVX-512 batch path.
    rng = np.random.default_rng(2024)
    x = np.concatenate([np.array([0, 1, 2, 42, 2**63, 2**64 - 1], dtype=np.uint64),
                        rng.integers(0, 2**64, 58, dtype=np.uint64)])
    y = O.ref_int_hash64(x)
    with open(os.path.join(HERE, "hash_kat.json"), "w") as f:
        json.dump({"in": [f"{int(v):016x}" for v in x], "out": [f"{int(v):016x}" for v in y]}, f, indent=0)

    # The reference's DataGen streams (seed 42) and its referenceOutput rows.
    b, p = O.ref_generate(1500, 4000, [2, 4], [0, 3], seed=42)
    np.savez_compressed(os.path.join(HERE, "ref_datagen.npz"), seed=42, b_keys=b.keys, b_c0=b.cols[0], b_c1=b.cols[1],
                        b_types=np.array(b.types), p_keys=p.keys, p_c0=p.cols[0], p_c1=p.cols[1],
                        p_types=np.array(p.types))

    # A join with 50% misses: inputs, sorted referenceOutput rows, checksum.
    s = O.make_spec(seed=8, build_rows=1200, probe_rows=3000, distribution=0, match_rate=0.5, build_types=(2, 3),
                    probe_types=(1, 4))
    b, p = O.gen(s)
    rs = 8 + 4 + 8 + 2 + 16
    rows = O.ref_reference_rows(b, p, rs, p.rows)
    np.savez_compressed(os.path.join(HERE, "ref_join_small.npz"), b_keys=b.keys, b_c0=b.cols[0], b_c1=b.cols[1],
                        b_types=np.array(b.types), p_keys=p.keys, p_c0=p.cols[0], p_c1=p.cols[1],
                        p_types=np.array(p.types), rows=rows, checksum=np.array(O.checksum_rows(rows), dtype=np.uint64))

    # Multi-match keys including the all-ones and zero keys: checksums of the
    # reference oracle and of each reference scheme.
    pool = rng.integers(0, 2**64, 300, dtype=np.uint64)
    pool[:2] = [2**64 - 1, 0]
    bk = rng.choice(pool, 2500)
    pk = np.concatenate([rng.choice(pool, 5000), rng.integers(0, 2**64, 1000, dtype=np.uint64)])
    b = O.Relation(bk, [rng.integers(0, 2**32, bk.size, dtype=np.uint32)], [2])
    p = O.Relation(pk, [rng.integers(0, 2**64, (pk.size, 2), dtype=np.uint64)], [4])
    out = dict(b_keys=bk, b_c0=b.cols[0], b_types=np.array([2]), p_keys=pk, p_c0=p.cols[0], p_types=np.array([4]))
    out["cs_reference"] = np.array(O.ref_reference_checksum(b, p), dtype=np.uint64)
    for name, scheme, pb in (("chj", 1, (8,)), ("phj_pure", 3, (4, 4)), ("phj", 2, (4, 4))):
        out[f"cs_{name}"] = np.array(O.ref_run(scheme, b, p, pb, threads=3, bep_budget_mib=1).checksum.tuple(),
                                     dtype=np.uint64)
    np.savez_compressed(os.path.join(HERE, "ref_multi.npz"), **out)
    print("golden fixtures written to", HERE)


if __name__ == "__main__":
    main()
