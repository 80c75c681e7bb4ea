"""The reference's graph file formats (host-side parsing only).

<prefix>.csv         meta data; |V| is the first integer on line 2
                     (load_graph_meta_data, src/Utilities.cpp:24-35)
<prefix>.biedgelist  packed little-endian {u32 src, u32 dst[, u32 weight]}
                     (EdgeUnit, include/Type.hpp:12-26; DistGraphLoader::load_graph,
                     src/Graph.cpp:400-402; datasets/toolkits/edgelist_to_biedgelist.cpp)
*.edgelist           text: "n m" then one "src dst [weight]" line per edge
                     (naive_implementation/*.cpp readers)
"""
import os
import re

import numpy as np


def load_graph_meta_data(meta_path):
    with open(meta_path) as f:
        f.readline()
        line = f.readline()
    mt = re.search(r"[0-9]+", line)
    if not mt:
        raise ValueError(f"{meta_path}: no vertex count on line 2")
    return int(mt.group(0))


def load_biedgelist(path, weighted=False):
    unit = 12 if weighted else 8
    size = os.path.getsize(path)
    if size % unit:
        raise ValueError(f"{path}: size {size} is not a multiple of the {unit}-byte edge unit")
    raw = np.fromfile(path, dtype="<u4").reshape(-1, unit // 4)
    src = np.ascontiguousarray(raw[:, 0])
    dst = np.ascontiguousarray(raw[:, 1])
    w = np.ascontiguousarray(raw[:, 2]).astype(np.float32) if weighted else None
    return src, dst, w


def save_biedgelist(path, src, dst, weights=None):
    cols = [np.asarray(src, "<u4"), np.asarray(dst, "<u4")]
    if weights is not None:
        cols.append(np.asarray(weights, "<u4"))
    np.stack(cols, axis=1).astype("<u4").tofile(path)


def load_edgelist(path, weighted=False):
    with open(path) as f:
        header = f.readline().split()
        n = int(header[0])
        data = np.loadtxt(f, dtype=np.int64, ndmin=2)
    src = data[:, 0].astype(np.uint32)
    dst = data[:, 1].astype(np.uint32)
    w = data[:, 2].astype(np.float32) if weighted and data.shape[1] > 2 else None
    return n, src, dst, w
