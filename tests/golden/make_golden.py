"""Regenerates the committed fixtures in tests/golden/ (run in the dev container).

- philox_vectors.json: Philox4x32-10 outputs from rocRAND's independent
  implementation (philox_rocrand.cpp, host path) for the three Random123
  known-answer inputs plus 64 seeded random (counter, key) pairs.
- graphs.npz: the edge lists the reference ships as data under datasets/
  (simple-5, simple-10, simple-30, small-10, wiki-vote with its weights) and
  SNAP's wiki-Vote.txt made undirected (published: 608,389 triangles).
Only data files are read from the reference; no reference code is run.
"""
import json
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF = os.environ.get("ZG_REFERENCE", "/root/reference")

KAT = [  # Random123 kat_vectors: philox4x32 10 rounds
    ((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
    ((0xffffffff,) * 4, (0xffffffff,) * 2, (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
    ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
     (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
]


def philox_vectors():
    exe = "/tmp/zg_philox_rocrand"
    subprocess.check_call(["/opt/rocm/bin/hipcc", "-O1", "-o", exe, os.path.join(HERE, "philox_rocrand.cpp")],
                          stderr=subprocess.DEVNULL)
    rng = np.random.default_rng(20241203)
    inputs = [list(c) + list(k) for c, k, _ in KAT]
    inputs += rng.integers(0, 2 ** 32, size=(64, 6), dtype=np.uint64).tolist()
    text = "\n".join(" ".join(f"{int(x):08x}" for x in row) for row in inputs) + "\n"
    out = subprocess.run([exe], input=text, capture_output=True, text=True, check=True).stdout
    vecs = []
    for line in out.strip().split("\n"):
        v = [int(x, 16) for x in line.split()]
        vecs.append({"ctr": v[0:4], "key": v[4:6], "out": v[6:10]})
    for (c, k, o), got in zip(KAT, vecs):
        assert tuple(got["out"]) == o, "rocRAND disagrees with Random123 KAT"
    with open(os.path.join(HERE, "philox_vectors.json"), "w") as f:
        json.dump({"source": "rocRAND rocrand_philox4x32_10.h ten_rounds (host)", "vectors": vecs}, f,
                  indent=0)


def read_text_edgelist(path, weighted=False):
    with open(path) as f:
        n, m = map(int, f.readline().split()[:2])
        rows = np.loadtxt(f, dtype=np.int64, ndmin=2)
    assert rows.shape[0] == m, path
    src = rows[:, 0].astype(np.uint32)
    dst = rows[:, 1].astype(np.uint32)
    w = rows[:, 2].astype(np.uint32) if weighted else None
    return n, src, dst, w


def graphs():
    d = os.path.join(REF, "datasets")
    out = {}
    for name in ("simple-5", "simple-10", "simple-30"):
        n, s, t, _ = read_text_edgelist(os.path.join(d, name, "dataset", name + ".edgelist"))
        b = np.fromfile(os.path.join(d, name, "dataset", name + ".biedgelist"), dtype="<u4").reshape(-1, 2)
        assert (b[:, 0] == s).all() and (b[:, 1] == t).all(), "biedgelist and edgelist differ"
        out[name] = (n, s, t)
    n, s, t, _ = read_text_edgelist(os.path.join(d, "small-10", "small-10.edgelist"))
    out["small-10"] = (n, s, t)
    n, s, t, _ = read_text_edgelist(os.path.join(d, "wiki-vote", "dataset", "wiki-vote.edgelist"))
    n2, s2, t2, w2 = read_text_edgelist(os.path.join(d, "wiki-vote", "dataset", "wiki-vote.weighted.edgelist"),
                                        weighted=True)
    assert n == n2 and (s == s2).all() and (t == t2).all()
    out["wiki-vote"] = (n, s, t)
    # SNAP wiki-Vote.txt, undirected, self-loops dropped, ids relabeled densely
    rows = np.loadtxt(os.path.join(d, "wiki-vote", "dataset", "wiki-Vote.txt"), dtype=np.int64, comments="#")
    ids, inv = np.unique(rows, return_inverse=True)
    inv = inv.reshape(rows.shape)
    a, b = np.minimum(inv[:, 0], inv[:, 1]), np.maximum(inv[:, 0], inv[:, 1])
    keep = a != b
    pairs = np.unique(np.stack([a[keep], b[keep]], 1), axis=0)
    out["snap-wiki-Vote"] = (len(ids), pairs[:, 0].astype(np.uint32), pairs[:, 1].astype(np.uint32))
    arrays = {}
    for k, (n, s, t) in out.items():
        arrays[k + "/n"] = np.array(n, np.uint32)
        arrays[k + "/src"] = s
        arrays[k + "/dst"] = t
    arrays["wiki-vote/w"] = w2
    np.savez_compressed(os.path.join(HERE, "graphs.npz"), **arrays)


if __name__ == "__main__":
    philox_vectors()
    graphs()
    print("ok", file=sys.stderr)
