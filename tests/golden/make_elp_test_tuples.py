"""Writes tests/golden/elp_test_tuples.npz: the ELP training tuples of the
graphs held out as the test split when the deployed network was trained.

    python -m zgraph.elp_train --graphs 256 --pilots 64 --out DIR   # on a GPU
    python tests/golden/make_elp_test_tuples.py DIR/elp_data.npz

The test-graph ids come from the deployed weights' training record
(sampling-graph-mining_amd/zgraph/elp_weights.npz, "test_graphs")."""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTS = os.path.join(HERE, "..", "..", "sampling-graph-mining_amd", "zgraph", "elp_weights.npz")


def main(data_path):
    with np.load(WEIGHTS, allow_pickle=False) as w:
        meta = json.loads(bytes(w["meta"]).decode())
    with np.load(data_path, allow_pickle=False) as d:
        te = np.isin(d["G"], meta["test_graphs"])
        np.savez_compressed(os.path.join(HERE, "elp_test_tuples.npz"), X=d["X"][te].astype(np.float32),
                            Y=d["Y"][te].astype(np.float32), P=d["P"][te].astype(np.int8))
    print(f"{int(te.sum())} tuples from {len(meta['test_graphs'])} held-out graphs")


if __name__ == "__main__":
    main(sys.argv[1])
