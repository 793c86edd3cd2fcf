"""Write checkpoints of the GPU library for OpenLB to resume from.

Run on a machine with the GPU and the built library (no reference needed):

    python tests/golden/gen_library_checkpoints.py <output dir> [<split-lattice output dir>]

For every fixture in tests/golden/checkpoints/, the library restores OpenLB's
checkpoint after k steps, advances `EXTRA` steps on the device, and writes its
own checkpoint (lbmh_checkpoint_save) as an OpenLB .dat file.  The files are
kept under tests/golden/checkpoints/library/<name>.npz; on the build machine,
where the reference binary exists, test_checkpoint_cpu.py has OpenLB load each
one and continue to step m, which must reach OpenLB's own populations.
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

import golden_cases as G  # noqa: E402
from openlb_amd import checkpoint as CK  # noqa: E402
from openlb_amd.lattice import BlockLattice  # noqa: E402

EXTRA = 8


def main(out_dir: str, multiblock_dir: str | None = None):
    os.makedirs(out_dir, exist_ok=True)
    if multiblock_dir is not None:
        write_multiblock(multiblock_dir)
    for name in G.checkpoint_names():
        case, meta, _, text = G.load_checkpoint(name)
        k, m = meta["steps"]
        real = meta["precision"]
        with BlockLattice.from_case(case, real) as lat:
            lat.restore(CK.decode_dat(text), real)
            lat.collide_and_stream(EXTRA)
            dat = CK.encode_dat(lat.checkpoint(real))
        info = {"fixture": name, "step": k + EXTRA, "precision": real,
                "source": "lbmh_checkpoint_save after restoring OpenLB's checkpoint and "
                          f"{EXTRA} device steps"}
        np.savez_compressed(os.path.join(out_dir, name + ".npz"), meta=np.array(json.dumps(info)),
                            dat=np.frombuffer(dat.encode("ascii"), dtype=np.uint8))
        print(f"library checkpoint {name}: step {k + EXTRA}", flush=True)


def write_multiblock(out_dir: str):
    """The same for the split-lattice fixtures of tests/golden/multiblock/ that
    hold OpenLB's checkpoint: a SuperLattice on OpenLB's cuboids restores it,
    advances EXTRA steps and writes its own (every block, in cuboid order)."""
    from openlb_amd.super_lattice import SuperLattice
    os.makedirs(out_dir, exist_ok=True)
    for name in G.multiblock_names():
        case, meta, arrays = G.load_multiblock(name)
        k, real = meta["checkpoint_step"], meta["precision"]
        if k is None:
            continue
        with SuperLattice(case, real, tuple(meta["blocks"])) as sup:
            sup.restore(CK.decode_dat(arrays["dat"].tobytes().decode("ascii")), real)
            sup.collide_and_stream(EXTRA)
            dat = CK.encode_dat(sup.checkpoint(real))
        info = {"fixture": name, "step": k + EXTRA, "precision": real, "blocks": meta["blocks"],
                "source": "SuperLattice.checkpoint after restoring OpenLB's checkpoint and "
                          f"{EXTRA} device steps"}
        np.savez_compressed(os.path.join(out_dir, name + ".npz"), meta=np.array(json.dumps(info)),
                            dat=np.frombuffer(dat.encode("ascii"), dtype=np.uint8))
        print(f"library split checkpoint {name}: step {k + EXTRA}", flush=True)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else os.path.join(HERE, "checkpoints", "library"),
         sys.argv[2] if len(sys.argv) > 2 else os.path.join(HERE, "multiblock", "library"))
