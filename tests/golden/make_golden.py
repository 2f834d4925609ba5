"""Writes the golden fixtures under tests/golden/ (run in the build container).

1. fixture_2744.npz -- the reference's own known-answer group
   code/test/{r,ac,p,ac_next} (R 343x2744, AC 2744x2744, P 2744x343, and the
   expected AC_next 343x343), converted from Matrix Market to 0-based int32
   CSR arrays + float64 values.  Data only.
2. ref_outputs_<case>.npz -- outputs of the reference library itself
   (oracle/_ref/libsp3mm_ref.so, compiled from the reference sources by
   ``make -C oracle ref``) on the fixture inputs and on the 64x64 2D Poisson
   problem:
     pair_symbnum  sp3mmRowByRowPair_0 with spmmRowByRow_SymbNum_0 (IDXMAP)
     pair_2d       sp3mmRowByRowPair_0 with spmmRowByRow2DBlocks_0 (default)
     merged        sp3mmRowByRowMerged_0
     pair_symbnum_1 / merged_1   the 1-based (_1) entry points on 1-based inputs
     ub_ra, symb_ra  spMMSizeUpperbound_0 / SpMM_Symb___0 of R*A
   so the oracle can be pinned bit for bit without the reference present.

Usage:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from sp3mm4amg_amd import mmio, problems  # noqa: E402
from oracle import ref  # noqa: E402

REF_TEST = "/root/reference/code/test"


def pack(prefix, m, out):
    out[prefix + "_shape"] = np.array([m.M, m.N], np.int64)
    out[prefix + "_irp"] = m.irp.astype(np.int64)
    out[prefix + "_ja"] = m.ja.astype(np.int32)
    out[prefix + "_vals"] = m.vals.astype(np.float64)


def main():
    fx = {}
    mats = {}
    for name in ("r", "ac", "p", "ac_next"):
        mats[name] = mmio.read(os.path.join(REF_TEST, name))
        pack(name, mats[name], fx)
    np.savez_compressed(os.path.join(HERE, "fixture_2744.npz"), **fx)

    cases = {
        "fixture": (mats["r"], mats["ac"], mats["p"]),
        "poisson2d64": tuple(problems.poisson2d_sa(64)[k] for k in ("R", "A", "P")),
    }
    for case, (R, A, P) in cases.items():
        out = {}
        pack("pair_symbnum", ref.sp3mm_pair(R, A, P, threads=4), out)
        pack("pair_2d", ref.sp3mm_pair(R, A, P, spmm="spmmRowByRow2DBlocks_", threads=4,
                                       grid=(4, 2)), out)
        pack("merged", ref.sp3mm_merged(R, A, P, threads=4), out)
        pack("pair_symbnum_1", ref.sp3mm_pair(R, A, P, base=1, threads=4), out)
        pack("merged_1", ref.sp3mm_merged(R, A, P, base=1, threads=4), out)
        out["ub_ra"] = ref.ub(R, A)
        out["symb_ra"] = ref.symb(R, A)
        np.savez_compressed(os.path.join(HERE, f"ref_outputs_{case}.npz"), **out)
        print(case, "written")


if __name__ == "__main__":
    main()
